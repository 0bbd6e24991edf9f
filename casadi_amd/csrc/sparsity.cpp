// Sparsity patterns, colourings, AMD and symbolic LDL/QR (host side).
// Each routine cites the CasADi function whose observable result it restates.
#include "sparsity.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace sxh {

ivec Sparsity::get_col() const {
  ivec col(nnz());
  for (int64_t c = 0; c < ncol; ++c)
    for (int64_t k = colind[c]; k < colind[c + 1]; ++k) col[k] = c;
  return col;
}

ivec Sparsity::compressed() const {
  ivec sp;
  sp.reserve(3 + ncol + nnz());
  sp.push_back(nrow);
  sp.push_back(ncol);
  sp.insert(sp.end(), colind.begin(), colind.end());
  sp.insert(sp.end(), row.begin(), row.end());
  return sp;
}

Sparsity Sparsity::from_compressed(const int64_t* sp) {
  int64_t nr = sp[0], nc = sp[1];
  ivec ci(sp + 2, sp + 3 + nc);
  ivec r(sp + 3 + nc, sp + 3 + nc + ci.back());
  return Sparsity(nr, nc, ci, r);
}

Sparsity Sparsity::dense(int64_t nr, int64_t nc) {
  Sparsity s(nr, nc);
  s.row.resize(nr * nc);
  for (int64_t c = 0; c < nc; ++c) {
    s.colind[c + 1] = (c + 1) * nr;
    for (int64_t r = 0; r < nr; ++r) s.row[c * nr + r] = r;
  }
  return s;
}

Sparsity Sparsity::diag(int64_t n) {
  Sparsity s(n, n);
  s.row.resize(n);
  for (int64_t c = 0; c < n; ++c) {
    s.colind[c + 1] = c + 1;
    s.row[c] = c;
  }
  return s;
}

// Sparsity::triplet (casadi/core/sparsity.cpp:1143-1293): entries sorted by
// column then row, duplicates merged.
Sparsity Sparsity::triplet(int64_t nr, int64_t nc, const ivec& r, const ivec& c, ivec* mapping) {
  if (r.size() != c.size()) throw std::runtime_error("triplet: inconsistent lengths");
  const int64_t n = static_cast<int64_t>(r.size());
  for (int64_t k = 0; k < n; ++k) {
    if (c[k] < 0 || c[k] >= nc || r[k] < 0 || r[k] >= nr)
      throw std::runtime_error("triplet: index out of bounds");
  }
  // stable counting sort by row, then by column
  ivec cnt(nr + 1, 0), by_row(n), by_col(n);
  for (int64_t k = 0; k < n; ++k) cnt[r[k] + 1]++;
  for (int64_t i = 0; i < nr; ++i) cnt[i + 1] += cnt[i];
  for (int64_t k = 0; k < n; ++k) by_row[cnt[r[k]]++] = k;
  ivec ccnt(nc + 1, 0);
  for (int64_t k : by_row) ccnt[c[k] + 1]++;
  for (int64_t i = 0; i < nc; ++i) ccnt[i + 1] += ccnt[i];
  for (int64_t k : by_row) by_col[ccnt[c[k]]++] = k;
  Sparsity s(nr, nc);
  if (mapping) mapping->assign(n, -1);
  int64_t pos = 0;
  for (int64_t col = 0; col < nc; ++col) {
    int64_t prev = -1;
    while (pos < n && c[by_col[pos]] == col) {
      int64_t el = by_col[pos++];
      if (r[el] != prev) s.row.push_back(r[el]);
      if (mapping) (*mapping)[el] = static_cast<int64_t>(s.row.size()) - 1;
      prev = r[el];
    }
    s.colind[col + 1] = static_cast<int64_t>(s.row.size());
  }
  return s;
}

Sparsity Sparsity::T(ivec* mapping) const {
  Sparsity t(ncol, nrow);
  t.row.resize(nnz());
  ivec cnt(nrow + 1, 0);
  for (int64_t k = 0; k < nnz(); ++k) cnt[row[k] + 1]++;
  for (int64_t i = 0; i < nrow; ++i) cnt[i + 1] += cnt[i];
  t.colind = cnt;
  if (mapping) mapping->resize(nnz());
  for (int64_t c = 0; c < ncol; ++c) {
    for (int64_t k = colind[c]; k < colind[c + 1]; ++k) {
      int64_t el = cnt[row[k]]++;
      t.row[el] = c;
      if (mapping) (*mapping)[el] = k;
    }
  }
  return t;
}

bool Sparsity::is_symmetric() const {
  if (!is_square()) return false;
  return *this == T();
}

Sparsity Sparsity::mtimes(const Sparsity& B) const {
  if (ncol != B.nrow) throw std::runtime_error("mtimes: dimension mismatch");
  Sparsity s(nrow, B.ncol);
  ivec mark(nrow, -1);
  for (int64_t j = 0; j < B.ncol; ++j) {
    ivec rows;
    for (int64_t kb = B.colind[j]; kb < B.colind[j + 1]; ++kb) {
      int64_t k = B.row[kb];
      for (int64_t ka = colind[k]; ka < colind[k + 1]; ++ka) {
        int64_t i = row[ka];
        if (mark[i] != j) {
          mark[i] = j;
          rows.push_back(i);
        }
      }
    }
    std::sort(rows.begin(), rows.end());
    s.row.insert(s.row.end(), rows.begin(), rows.end());
    s.colind[j + 1] = static_cast<int64_t>(s.row.size());
  }
  return s;
}

Sparsity Sparsity::intersect(const Sparsity& B) const {
  if (nrow != B.nrow || ncol != B.ncol) throw std::runtime_error("intersect: dimension mismatch");
  Sparsity s(nrow, ncol);
  for (int64_t c = 0; c < ncol; ++c) {
    // both columns are sorted by row: merge and keep the common rows
    int64_t ka = colind[c], kb = B.colind[c];
    while (ka < colind[c + 1] && kb < B.colind[c + 1]) {
      if (row[ka] < B.row[kb]) {
        ++ka;
      } else if (B.row[kb] < row[ka]) {
        ++kb;
      } else {
        s.row.push_back(row[ka]);
        ++ka;
        ++kb;
      }
    }
    s.colind[c + 1] = static_cast<int64_t>(s.row.size());
  }
  return s;
}

Sparsity Sparsity::enlarge(int64_t nr, int64_t nc, const ivec& rr, const ivec& cc) const {
  if (static_cast<int64_t>(rr.size()) != nrow || static_cast<int64_t>(cc.size()) != ncol)
    throw std::runtime_error("enlarge: index vectors do not match the pattern");
  ivec r(nnz()), c(nnz()), col = get_col();
  for (int64_t k = 0; k < nnz(); ++k) {
    r[k] = rr[row[k]];
    c[k] = cc[col[k]];
  }
  return triplet(nr, nc, r, c);
}

Sparsity Sparsity::sub(const ivec& rr, const ivec& cc, ivec* mapping) const {
  // rows of this -> list of positions in rr
  std::vector<ivec> rmap(nrow);
  for (size_t i = 0; i < rr.size(); ++i) rmap.at(rr[i]).push_back(static_cast<int64_t>(i));
  ivec r_out, c_out, src;
  for (size_t j = 0; j < cc.size(); ++j) {
    int64_t c = cc[j];
    for (int64_t k = colind[c]; k < colind[c + 1]; ++k) {
      for (int64_t i : rmap[row[k]]) {
        r_out.push_back(i);
        c_out.push_back(static_cast<int64_t>(j));
        src.push_back(k);
      }
    }
  }
  ivec m;
  Sparsity s = triplet(static_cast<int64_t>(rr.size()), static_cast<int64_t>(cc.size()), r_out, c_out, &m);
  if (mapping) {
    mapping->assign(s.nnz(), -1);
    for (size_t e = 0; e < m.size(); ++e) (*mapping)[m[e]] = src[e];
  }
  return s;
}

int64_t Sparsity::get_nz(int64_t r, int64_t c) const {
  auto b = row.begin() + colind[c], e = row.begin() + colind[c + 1];
  auto it = std::lower_bound(b, e, r);
  if (it != e && *it == r) return static_cast<int64_t>(it - row.begin());
  return -1;
}

ivec Sparsity::get_nz(const ivec& rows, const ivec& cols) const {
  ivec out(rows.size());
  for (size_t k = 0; k < rows.size(); ++k) out[k] = get_nz(rows[k], cols[k]);
  return out;
}

// SparsityInternal::pmult (casadi/core/sparsity_internal.cpp:3546-3600)
Sparsity Sparsity::pmult(const ivec& p, bool permute_rows, bool permute_columns, bool invert) const {
  ivec pinv;
  if (invert) {
    pinv.resize(p.size());
    for (size_t k = 0; k < p.size(); ++k) pinv[p[k]] = static_cast<int64_t>(k);
  }
  const ivec& pp = invert ? pinv : p;
  ivec col = get_col();
  ivec new_row(nnz()), new_col(nnz());
  for (int64_t k = 0; k < nnz(); ++k) {
    new_col[k] = permute_columns ? pp[col[k]] : col[k];
    new_row[k] = permute_rows ? pp[row[k]] : row[k];
  }
  return triplet(nrow, ncol, new_row, new_col);
}

// SparsityInternal::uni_coloring (casadi/core/sparsity_internal.cpp:2959-3046)
Sparsity Sparsity::uni_coloring(const Sparsity& AT, int64_t cutoff) const {
  ivec forbidden;
  ivec color(ncol, 0);
  for (int64_t i = 0; i < ncol; ++i) {
    for (int64_t el = colind[i]; el < colind[i + 1]; ++el) {
      int64_t c = row[el];
      for (int64_t el_prev = AT.colind[c]; el_prev < AT.colind[c + 1]; ++el_prev) {
        int64_t i_prev = AT.row[el_prev];
        if (i_prev >= i) break;
        forbidden[color[i_prev]] = i;
      }
    }
    int64_t color_i;
    for (color_i = 0; color_i < static_cast<int64_t>(forbidden.size()); ++color_i) {
      if (forbidden[color_i] != i) break;
    }
    color[i] = color_i;
    if (color_i == static_cast<int64_t>(forbidden.size())) {
      forbidden.push_back(0);
      if (static_cast<int64_t>(forbidden.size()) > cutoff) return Sparsity::null();
    }
  }
  const int64_t ncolor = static_cast<int64_t>(forbidden.size());
  ivec cols(ncol);
  for (int64_t i = 0; i < ncol; ++i) cols[i] = i;
  // rows = original columns, cols = colours; one entry per column
  return triplet(ncol, ncolor, cols, color);
}

// SparsityInternal::largest_first (casadi/core/sparsity_internal.cpp:3511-3544)
ivec Sparsity::largest_first() const {
  ivec degree(ncol);
  int64_t max_degree = 0;
  for (int64_t k = 0; k < ncol; ++k) {
    degree[k] = colind[k + 1] - colind[k];
    max_degree = std::max(max_degree, 1 + degree[k]);
  }
  ivec degree_count(max_degree + 1, 0);
  for (int64_t d : degree) degree_count.at(d + 1)++;
  for (int64_t d = 0; d < max_degree; ++d) degree_count[d + 1] += degree_count[d];
  ivec ordering(ncol);
  for (int64_t k = ncol - 1; k >= 0; --k) ordering[degree_count[degree[k]]++] = k;
  ivec reverse_ordering(ordering.rbegin(), ordering.rend());
  return reverse_ordering;
}

// SparsityInternal::star_coloring (casadi/core/sparsity_internal.cpp:3398-3509)
// (Gebremedhin, Manne, Pothen, Algorithm 4.1)
Sparsity Sparsity::star_coloring(int64_t ordering, int64_t cutoff) const {
  if (ordering != 0) {
    if (ordering != 1) throw std::runtime_error("star_coloring: ordering must be 0 or 1");
    ivec ord = largest_first();
    Sparsity sp_permuted = pmult(ord, true, true, true);
    Sparsity ret_permuted = sp_permuted.star_coloring(0, cutoff);
    if (ret_permuted.is_null()) return ret_permuted;
    return ret_permuted.pmult(ord, true, false, false);
  }
  ivec forbidden;
  ivec color(ncol, -1);
  for (int64_t i = 0; i < ncol; ++i) {
    for (int64_t w_el = colind[i]; w_el < colind[i + 1]; ++w_el) {
      int64_t w = row[w_el];
      if (color[w] != -1) forbidden[color[w]] = i;
      for (int64_t x_el = colind[w]; x_el < colind[w + 1]; ++x_el) {
        int64_t x = row[x_el];
        if (color[x] == -1) continue;
        if (color[w] == -1) {
          forbidden[color[x]] = i;
        } else {
          for (int64_t y_el = colind[x]; y_el < colind[x + 1]; ++y_el) {
            int64_t y = row[y_el];
            if (color[y] == -1 || y == w) continue;
            if (color[y] == color[w]) {
              forbidden[color[x]] = i;
              break;
            }
          }
        }
      }
    }
    bool new_color = true;
    for (int64_t ci = 0; ci < static_cast<int64_t>(forbidden.size()); ++ci) {
      if (forbidden[ci] != i) {
        color[i] = ci;
        new_color = false;
        break;
      }
    }
    if (new_color) {
      color[i] = static_cast<int64_t>(forbidden.size());
      forbidden.push_back(-1);
      if (static_cast<int64_t>(forbidden.size()) > cutoff) return Sparsity::null();
    }
  }
  ivec rng(ncol);
  std::iota(rng.begin(), rng.end(), 0);
  return triplet(ncol, static_cast<int64_t>(forbidden.size()), rng, color);
}

// --- elimination tree / postorder / LDL symbolic ---------------------------

// SparsityInternal::etree (casadi/core/sparsity_internal.cpp:35-71)
void sp_etree(const int64_t* sp, int64_t* parent, int64_t* w, bool ata) {
  int64_t nrow = sp[0], ncol = sp[1];
  const int64_t *colind = sp + 2, *row = sp + 2 + ncol + 1;
  int64_t* ancestor = w;
  int64_t* prev = nullptr;
  if (ata) {
    prev = w + ncol;
    for (int64_t r = 0; r < nrow; ++r) prev[r] = -1;
  }
  for (int64_t c = 0; c < ncol; ++c) {
    parent[c] = -1;
    ancestor[c] = -1;
    for (int64_t k = colind[c]; k < colind[c + 1]; ++k) {
      int64_t r = row[k];
      if (ata) r = prev[r];
      while (r != -1 && r < c) {
        int64_t rnext = ancestor[r];
        ancestor[r] = c;
        if (rnext == -1) parent[r] = c;
        r = rnext;
      }
      if (ata) prev[row[k]] = c;
    }
  }
}

static int64_t postorder_dfs(int64_t j, int64_t k, int64_t* head, const int64_t* next, int64_t* post,
                             int64_t* stack) {
  int64_t top = 0;
  stack[0] = j;
  while (top >= 0) {
    int64_t p = stack[top];
    int64_t i = head[p];
    if (i == -1) {
      top--;
      post[k++] = p;
    } else {
      head[p] = next[i];
      stack[++top] = i;
    }
  }
  return k;
}

// SparsityInternal::postorder (casadi/core/sparsity_internal.cpp:95-119)
void sp_postorder(const int64_t* parent, int64_t n, int64_t* post, int64_t* w) {
  int64_t *head = w, *next = w + n, *stack = w + 2 * n;
  for (int64_t j = 0; j < n; ++j) head[j] = -1;
  for (int64_t j = n - 1; j >= 0; --j) {
    if (parent[j] != -1) {
      next[j] = head[parent[j]];
      head[parent[j]] = j;
    }
  }
  int64_t k = 0;
  for (int64_t j = 0; j < n; j++)
    if (parent[j] == -1) k = postorder_dfs(j, k, head, next, post, stack);
}

ivec Sparsity::etree(bool ata) const {
  ivec sp = compressed();
  ivec parent(ncol), w(ata ? nrow + ncol : ncol);
  sp_etree(sp.data(), parent.data(), w.data(), ata);
  return parent;
}

// SparsityInternal::ldl_colind (casadi/core/sparsity_internal.cpp:379-410)
void sp_ldl_colind(const int64_t* sp, int64_t* parent, int64_t* l_colind, int64_t* w) {
  int64_t n = sp[0];
  const int64_t *colind = sp + 2, *row = sp + 2 + n + 1;
  int64_t* visited = w;
  for (int64_t c = 0; c < n; ++c) {
    parent[c] = -1;
    visited[c] = c;
    l_colind[1 + c] = 0;
    int64_t r;
    for (int64_t k = colind[c]; k < colind[c + 1] && (r = row[k]) < c; ++k) {
      while (visited[r] != c) {
        if (parent[r] == -1) parent[r] = c;
        l_colind[1 + r]++;
        visited[r] = c;
        r = parent[r];
      }
    }
  }
  l_colind[0] = 0;
  for (int64_t c = 0; c < n; ++c) l_colind[c + 1] += l_colind[c];
}

// SparsityInternal::ldl_row (casadi/core/sparsity_internal.cpp:412-448)
void sp_ldl_row(const int64_t* sp, const int64_t* parent, int64_t* l_colind, int64_t* l_row,
                int64_t* w) {
  int64_t n = sp[0];
  const int64_t *colind = sp + 2, *row = sp + n + 3;
  int64_t* visited = w;
  for (int64_t c = 0; c < n; ++c) {
    visited[c] = c;
    int64_t r;
    for (int64_t k = colind[c]; k < colind[c + 1] && (r = row[k]) < c; ++k) {
      while (visited[r] != c) {
        l_row[l_colind[r]++] = c;
        visited[r] = c;
        r = parent[r];
      }
    }
  }
  int64_t k = 0;
  for (int64_t c = 0; c < n; ++c) {
    int64_t r = l_colind[c];
    l_colind[c] = k;
    k = r;
  }
}

// Sparsity::ldl (casadi/core/sparsity.cpp:622-652)
Sparsity Sparsity::ldl(ivec& p, bool use_amd) const {
  if (!is_symmetric()) throw std::runtime_error("LDL factorization requires a symmetric matrix");
  if (use_amd) {
    p = amd();
    ivec tmp;
    Sparsity Aperm = sub(p, p);
    return Aperm.ldl(tmp, false);
  }
  int64_t n = nrow;
  p.resize(n);
  std::iota(p.begin(), p.end(), 0);
  ivec w(3 * n), parent(n), L_colind(1 + n);
  ivec sp = compressed();
  sp_ldl_colind(sp.data(), parent.data(), L_colind.data(), w.data());
  ivec L_row(L_colind.back());
  sp_ldl_row(sp.data(), parent.data(), L_colind.data(), L_row.data(), w.data());
  // ldl_row leaves rows of each column in increasing order
  return Sparsity(n, n, L_colind, L_row).T();
}

// --- QR symbolic ------------------------------------------------------------

// SparsityInternal::leaf (casadi/core/sparsity_internal.cpp:121-148)
static int64_t sp_leaf(int64_t i, int64_t j, const int64_t* first, int64_t* maxfirst, int64_t* prevleaf,
                       int64_t* ancestor, int64_t* jleaf) {
  *jleaf = 0;
  if (i <= j || first[j] <= maxfirst[i]) return -1;
  maxfirst[i] = first[j];
  int64_t jprev = prevleaf[i];
  prevleaf[i] = j;
  *jleaf = (jprev == -1) ? 1 : 2;
  if (*jleaf == 1) return i;
  int64_t q;
  for (q = jprev; q != ancestor[q]; q = ancestor[q]) {
  }
  for (int64_t s = jprev, sparent; s != q; s = sparent) {
    sparent = ancestor[s];
    ancestor[s] = q;
  }
  return q;
}

// SparsityInternal::qr_counts (casadi/core/sparsity_internal.cpp:150-226)
static int64_t sp_qr_counts(const int64_t* tr_sp, const int64_t* parent, const int64_t* post,
                            int64_t* counts, int64_t* w) {
  int64_t ncol = *tr_sp++, nrow = *tr_sp++;
  const int64_t *rowind = tr_sp, *col = tr_sp + nrow + 1;
  int64_t *ancestor = w, *maxfirst = w + ncol, *prevleaf = w + 2 * ncol, *first = w + 3 * ncol;
  int64_t *head = w + 4 * ncol, *next = w + 5 * ncol + 1;
  for (int64_t k = 0; k < ncol; ++k) first[k] = -1;
  for (int64_t k = 0; k < ncol; ++k) {
    int64_t j = post[k];
    counts[j] = (first[j] == -1) ? 1 : 0;
    for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
  }
  for (int64_t k = 0; k < ncol; ++k) ancestor[post[k]] = k;
  for (int64_t k = 0; k < ncol + 1; ++k) head[k] = -1;
  for (int64_t i = 0; i < nrow; ++i) {
    int64_t k = ncol;
    for (int64_t p = rowind[i]; p < rowind[i + 1]; ++p) k = std::min(k, ancestor[col[p]]);
    next[i] = head[k];
    head[k] = i;
  }
  for (int64_t k = 0; k < ncol; ++k) maxfirst[k] = -1;
  for (int64_t k = 0; k < ncol; ++k) prevleaf[k] = -1;
  for (int64_t i = 0; i < ncol; ++i) ancestor[i] = i;
  for (int64_t k = 0; k < ncol; ++k) {
    int64_t j = post[k];
    if (parent[j] != -1) counts[parent[j]]--;
    for (int64_t J = head[k]; J != -1; J = next[J]) {
      for (int64_t p = rowind[J]; p < rowind[J + 1]; ++p) {
        int64_t i = col[p], jleaf;
        int64_t q = sp_leaf(i, j, first, maxfirst, prevleaf, ancestor, &jleaf);
        if (jleaf >= 1) counts[j]++;
        if (jleaf == 2) counts[q]--;
      }
    }
    if (parent[j] != -1) ancestor[j] = parent[j];
  }
  for (int64_t j = 0; j < ncol; ++j)
    if (parent[j] != -1) counts[parent[j]] += counts[j];
  int64_t sum_counts = 0;
  for (int64_t j = 0; j < ncol; ++j) sum_counts += counts[j];
  return sum_counts;
}

// SparsityInternal::qr_nnz (casadi/core/sparsity_internal.cpp:227-289)
static int64_t sp_qr_nnz(const int64_t* sp, int64_t* pinv, int64_t* leftmost, const int64_t* parent,
                         int64_t* nrow_ext, int64_t* w) {
  int64_t nrow = sp[0], ncol = sp[1];
  const int64_t *colind = sp + 2, *row = sp + 2 + ncol + 1;
  int64_t *next = w, *head = w + nrow, *tail = w + nrow + ncol, *nque = w + nrow + 2 * ncol;
  for (int64_t c = 0; c < ncol; ++c) head[c] = -1;
  for (int64_t c = 0; c < ncol; ++c) tail[c] = -1;
  for (int64_t c = 0; c < ncol; ++c) nque[c] = 0;
  for (int64_t r = 0; r < nrow; ++r) leftmost[r] = -1;
  for (int64_t c = ncol - 1; c >= 0; --c)
    for (int64_t k = colind[c]; k < colind[c + 1]; ++k) leftmost[row[k]] = c;
  for (int64_t r = nrow - 1; r >= 0; --r) {
    pinv[r] = -1;
    int64_t c = leftmost[r];
    if (c == -1) continue;
    if (nque[c]++ == 0) tail[c] = r;
    next[r] = head[c];
    head[c] = r;
  }
  int64_t v_nnz = 0, nrow_new = nrow, c;
  for (c = 0; c < ncol; ++c) {
    int64_t r = head[c];
    v_nnz++;
    if (r < 0) r = nrow_new++;
    pinv[r] = c;
    if (--nque[c] <= 0) continue;
    v_nnz += nque[c];
    int64_t pa = parent[c];
    if (pa != -1) {
      if (nque[pa] == 0) tail[pa] = tail[c];
      next[tail[c]] = head[pa];
      head[pa] = next[r];
      nque[pa] += nque[c];
    }
  }
  for (int64_t r = 0; r < nrow; ++r)
    if (pinv[r] < 0) pinv[r] = c++;
  if (nrow_ext) *nrow_ext = nrow_new;
  return v_nnz;
}

// SparsityInternal::qr_sparsities (casadi/core/sparsity_internal.cpp:306-377)
static void sp_qr_sparsities(const int64_t* sp_a, int64_t nrow_ext, int64_t* sp_v, int64_t* sp_r,
                             const int64_t* leftmost, const int64_t* parent, const int64_t* pinv,
                             int64_t* iw) {
  int64_t ncol = sp_a[1];
  const int64_t *colind = sp_a + 2, *row = sp_a + 2 + ncol + 1;
  int64_t *v_colind = sp_v + 2, *v_row = sp_v + 2 + ncol + 1;
  int64_t *r_colind = sp_r + 2, *r_row = sp_r + 2 + ncol + 1;
  sp_v[0] = sp_r[0] = nrow_ext;
  sp_v[1] = sp_r[1] = ncol;
  int64_t* s = iw;
  iw += ncol;
  for (int64_t r = 0; r < nrow_ext; ++r) iw[r] = -1;
  int64_t nnz_r = 0, nnz_v = 0;
  for (int64_t c = 0; c < ncol; ++c) {
    r_colind[c] = nnz_r;
    v_colind[c] = nnz_v;
    iw[c] = c;
    v_row[nnz_v++] = c;
    int64_t top = ncol;
    for (int64_t k = colind[c]; k < colind[c + 1]; ++k) {
      int64_t r = leftmost[row[k]];
      int64_t len;
      for (len = 0; iw[r] != c; r = parent[r]) {
        s[len++] = r;
        iw[r] = c;
      }
      while (len > 0) s[--top] = s[--len];
      r = pinv[row[k]];
      if (r > c && iw[r] < c) {
        v_row[nnz_v++] = r;
        iw[r] = c;
      }
    }
    for (int64_t k = top; k < ncol; ++k) {
      int64_t r = s[k];
      r_row[nnz_r++] = r;
      if (parent[r] == c) {
        for (int64_t k2 = v_colind[r]; k2 < v_colind[r + 1]; ++k2) {
          int64_t r2 = v_row[k2];
          if (iw[r2] < c) {
            iw[r2] = c;
            v_row[nnz_v++] = r2;
          }
        }
      }
    }
    r_row[nnz_r++] = c;
  }
  r_colind[ncol] = nnz_r;
  v_colind[ncol] = nnz_v;
}

// Sparsity::qr_sparse (casadi/core/sparsity.cpp:654-693)
void Sparsity::qr_sparse(Sparsity& V, Sparsity& R, ivec& prinv, ivec& pc, bool use_amd) const {
  const int64_t size1 = nrow, size2 = ncol;
  if (use_amd) {
    pc = T().mtimes(*this).amd();
    ivec rr(size1), tmp;
    std::iota(rr.begin(), rr.end(), 0);
    Sparsity Aperm = sub(rr, pc);
    Aperm.qr_sparse(V, R, prinv, tmp, false);
    return;
  }
  pc.resize(size2);
  std::iota(pc.begin(), pc.end(), 0);
  ivec leftmost(size1), parent(size2);
  prinv.assign(size1 + size2, 0);
  ivec iw(size1 + 7 * size2 + 1);
  ivec sp = compressed(), sp_tr = T().compressed();
  int64_t nrow_ext, v_nnz, r_nnz;
  // SparsityInternal::qr_init (casadi/core/sparsity_internal.cpp:290-304)
  sp_etree(sp.data(), parent.data(), iw.data(), true);
  int64_t* post = iw.data();
  sp_postorder(parent.data(), size2, post, iw.data() + size2);
  r_nnz = sp_qr_counts(sp_tr.data(), parent.data(), post, iw.data() + size2, iw.data() + 2 * size2);
  v_nnz = sp_qr_nnz(sp.data(), prinv.data(), leftmost.data(), parent.data(), &nrow_ext, iw.data());
  ivec sp_v(2 + size2 + 1 + v_nnz), sp_r(2 + size2 + 1 + r_nnz);
  sp_qr_sparsities(sp.data(), nrow_ext, sp_v.data(), sp_r.data(), leftmost.data(), parent.data(),
                   prinv.data(), iw.data());
  prinv.resize(nrow_ext);
  V = from_compressed(sp_v.data());
  R = from_compressed(sp_r.data());
}

// --- approximate minimum degree ----------------------------------------------

static int64_t amd_wclear(int64_t mark, int64_t lemax, int64_t* w, int64_t n) {
  if (mark < 2 || (mark + lemax < 0)) {
    for (int64_t k = 0; k < n; ++k)
      if (w[k] != 0) w[k] = 1;
    mark = 2;
  }
  return mark;
}

// SparsityInternal::amd (casadi/core/sparsity_internal.cpp:661-973), the
// CSparse cs_amd minimum-degree ordering applied to a symmetric pattern with
// its diagonal dropped.
ivec Sparsity::amd() const {
  if (!is_symmetric()) throw std::runtime_error("AMD requires a symmetric matrix");
  const int64_t n = ncol;
  ivec cind = colind;
  ivec rw = row;
  int64_t nnz = 0, col_end = 0;
  for (int64_t c = 0; c < n; ++c) {
    int64_t col_begin = col_end;
    col_end = cind[c + 1];
    for (int64_t k = col_begin; k < col_end; ++k)
      if (rw[k] != c) rw[nnz++] = rw[k];
    cind[c + 1] = nnz;
  }
  int64_t dense = static_cast<int64_t>(10 * std::sqrt(static_cast<double>(n)));
  dense = std::max<int64_t>(16, dense);
  dense = std::min<int64_t>(n - 2, dense);
  ivec P(n + 1), len(n + 1), nv(n + 1), next(n + 1), head(n + 1), elen(n + 1), degree(n + 1),
      w(n + 1), hhead(n + 1);
  int64_t nel = 0, mindeg = 0, lemax = 0, d;
  uint64_t h;
  auto FLIP = [](int64_t i) { return -i - 2; };
  for (int64_t k = 0; k < n; ++k) len[k] = cind[k + 1] - cind[k];
  len[n] = 0;
  const int64_t nzmax = static_cast<int64_t>(row.size());
  // workspace beyond nzmax is never addressed while CasADi's own run stays in
  // bounds; it only keeps this restatement memory-safe on other inputs
  rw.resize(std::max<int64_t>(nzmax, 1) * 2 + 2 * n + 16);
  for (int64_t i = 0; i <= n; ++i) {
    head[i] = -1;
    P[i] = -1;
    next[i] = -1;
    hhead[i] = -1;
    nv[i] = 1;
    w[i] = 1;
    elen[i] = 0;
    degree[i] = len[i];
  }
  int64_t mark = amd_wclear(0, 0, w.data(), n);
  elen[n] = -2;
  cind[n] = -1;
  w[n] = 0;
  for (int64_t i = 0; i < n; ++i) {
    d = degree[i];
    if (d == 0) {
      elen[i] = -2;
      nel++;
      cind[i] = -1;
      w[i] = 0;
    } else if (d > dense) {
      nv[i] = 0;
      elen[i] = -1;
      nel++;
      cind[i] = FLIP(n);
      nv[n]++;
    } else {
      if (head[d] != -1) P[head[d]] = i;
      next[i] = head[d];
      head[d] = i;
    }
  }
  while (nel < n) {
    int64_t k;
    for (k = -1; mindeg < n && (k = head[mindeg]) == -1; mindeg++) {
    }
    if (next[k] != -1) P[next[k]] = -1;
    head[mindeg] = next[k];
    int64_t elenk = elen[k];
    int64_t nvk = nv[k];
    nel += nvk;
    if (elenk > 0 && nnz + mindeg >= nzmax) {
      for (int64_t j = 0; j < n; j++) {
        int64_t p = cind[j];
        if (p >= 0) {
          cind[j] = rw[p];
          rw[p] = FLIP(j);
        }
      }
      int64_t q, p;
      for (q = 0, p = 0; p < nnz;) {
        int64_t j = FLIP(rw[p++]);
        if (j >= 0) {
          rw[q] = cind[j];
          cind[j] = q++;
          for (int64_t k3 = 0; k3 < len[j] - 1; k3++) rw[q++] = rw[p++];
        }
      }
      nnz = q;
    }
    int64_t dk = 0;
    nv[k] = -nvk;
    int64_t p = cind[k];
    int64_t pk1 = (elenk == 0) ? p : nnz;
    int64_t pk2 = pk1;
    int64_t e, pj, ln;
    for (int64_t k1 = 1; k1 <= elenk + 1; k1++) {
      if (k1 > elenk) {
        e = k;
        pj = p;
        ln = len[k] - elenk;
      } else {
        e = rw[p++];
        pj = cind[e];
        ln = len[e];
      }
      for (int64_t k2 = 1; k2 <= ln; k2++) {
        int64_t i = rw[pj++];
        int64_t nvi = nv[i];
        if (nvi <= 0) continue;
        dk += nvi;
        nv[i] = -nvi;
        rw[pk2++] = i;
        if (next[i] != -1) P[next[i]] = P[i];
        if (P[i] != -1) {
          next[P[i]] = next[i];
        } else {
          head[degree[i]] = next[i];
        }
      }
      if (e != k) {
        cind[e] = FLIP(k);
        w[e] = 0;
      }
    }
    if (elenk != 0) nnz = pk2;
    degree[k] = dk;
    cind[k] = pk1;
    len[k] = pk2 - pk1;
    elen[k] = -2;
    mark = amd_wclear(mark, lemax, w.data(), n);
    for (int64_t pk = pk1; pk < pk2; pk++) {
      int64_t i = rw[pk];
      int64_t eln = elen[i];
      if (eln <= 0) continue;
      int64_t nvi = -nv[i];
      int64_t wnvi = mark - nvi;
      for (p = cind[i]; p <= cind[i] + eln - 1; p++) {
        e = rw[p];
        if (w[e] >= mark) {
          w[e] -= nvi;
        } else if (w[e] != 0) {
          w[e] = degree[e] + wnvi;
        }
      }
    }
    for (int64_t pk = pk1; pk < pk2; pk++) {
      int64_t i = rw[pk];
      int64_t p1 = cind[i];
      int64_t p2 = p1 + elen[i] - 1;
      int64_t pn = p1;
      for (h = 0, d = 0, p = p1; p <= p2; p++) {
        e = rw[p];
        if (w[e] != 0) {
          int64_t dext = w[e] - mark;
          if (dext > 0) {
            d += dext;
            rw[pn++] = e;
            h += static_cast<uint64_t>(e);
          } else {
            cind[e] = FLIP(k);
            w[e] = 0;
          }
        }
      }
      elen[i] = pn - p1 + 1;
      int64_t p3 = pn;
      int64_t p4 = p1 + len[i];
      for (p = p2 + 1; p < p4; p++) {
        int64_t j = rw[p];
        int64_t nvj = nv[j];
        if (nvj <= 0) continue;
        d += nvj;
        rw[pn++] = j;
        h += static_cast<uint64_t>(j);
      }
      if (d == 0) {
        cind[i] = FLIP(k);
        int64_t nvi = -nv[i];
        dk -= nvi;
        nvk += nvi;
        nel += nvi;
        nv[i] = 0;
        elen[i] = -1;
      } else {
        degree[i] = std::min(degree[i], d);
        rw[pn] = rw[p3];
        rw[p3] = rw[p1];
        rw[p1] = k;
        len[i] = pn - p1 + 1;
        h %= static_cast<uint64_t>(n);
        next[i] = hhead[h];
        hhead[h] = i;
        P[i] = static_cast<int64_t>(h);
      }
    }
    degree[k] = dk;
    lemax = std::max(lemax, dk);
    mark = amd_wclear(mark + lemax, lemax, w.data(), n);
    for (int64_t pk = pk1; pk < pk2; pk++) {
      int64_t i = rw[pk];
      if (nv[i] >= 0) continue;
      h = static_cast<uint64_t>(P[i]);
      i = hhead[h];
      hhead[h] = -1;
      for (; i != -1 && next[i] != -1; i = next[i], mark++) {
        ln = len[i];
        int64_t eln = elen[i];
        for (p = cind[i] + 1; p <= cind[i] + ln - 1; p++) w[rw[p]] = mark;
        int64_t jlast = i;
        for (int64_t j = next[i]; j != -1;) {
          bool ok = (len[j] == ln) && (elen[j] == eln);
          for (p = cind[j] + 1; ok && p <= cind[j] + ln - 1; p++)
            if (w[rw[p]] != mark) ok = false;
          if (ok) {
            cind[j] = FLIP(i);
            nv[i] += nv[j];
            nv[j] = 0;
            elen[j] = -1;
            j = next[j];
            next[jlast] = j;
          } else {
            jlast = j;
            j = next[j];
          }
        }
      }
    }
    int64_t pk;
    for (p = pk1, pk = pk1; pk < pk2; pk++) {
      int64_t i = rw[pk];
      int64_t nvi = -nv[i];
      if (nvi <= 0) continue;
      nv[i] = nvi;
      d = degree[i] + dk - nvi;
      d = std::min(d, n - nel - nvi);
      if (head[d] != -1) P[head[d]] = i;
      next[i] = head[d];
      P[i] = -1;
      head[d] = i;
      mindeg = std::min(mindeg, d);
      degree[i] = d;
      rw[p++] = i;
    }
    nv[k] = nvk;
    len[k] = p - pk1;
    if (len[k] == 0) {
      cind[k] = -1;
      w[k] = 0;
    }
    if (elenk != 0) nnz = p;
  }
  for (int64_t i = 0; i < n; i++) cind[i] = FLIP(cind[i]);
  for (int64_t j = 0; j <= n; j++) head[j] = -1;
  for (int64_t j = n; j >= 0; j--) {
    if (nv[j] > 0) continue;
    next[j] = head[cind[j]];
    head[cind[j]] = j;
  }
  for (int64_t e = n; e >= 0; e--) {
    if (nv[e] <= 0) continue;
    if (cind[e] != -1) {
      next[e] = head[cind[e]];
      head[cind[e]] = e;
    }
  }
  for (int64_t k = 0, i = 0; i <= n; i++) {
    if (cind[i] == -1) k = postorder_dfs(i, k, head.data(), next.data(), P.data(), w.data());
  }
  P.resize(n);
  return P;
}

}  // namespace sxh
