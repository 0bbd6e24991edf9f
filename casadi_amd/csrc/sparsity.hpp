// Compressed-column sparsity patterns and the integer graph algorithms the
// derivative and linear-solver paths need (host side).
//
// Restates casadi::Sparsity / SparsityInternal (casadi/core/sparsity.cpp,
// casadi/core/sparsity_internal.cpp): transposes, triplet construction,
// sub-patterns, the uni-directional and star colourings with largest-first
// ordering (sparsity_internal.cpp:2959-3600), approximate minimum degree
// ordering (sparsity_internal.cpp:661-973), the elimination tree and the
// symbolic LDL^T and Householder QR factorisations
// (sparsity_internal.cpp:35-448, sparsity.cpp:622-693).
#pragma once
#include <cstdint>
#include <vector>

namespace sxh {

using ivec = std::vector<int64_t>;

struct Sparsity {
  int64_t nrow = 0, ncol = 0;
  ivec colind{0};
  ivec row;

  Sparsity() = default;
  Sparsity(int64_t nr, int64_t nc) : nrow(nr), ncol(nc), colind(nc + 1, 0) {}
  Sparsity(int64_t nr, int64_t nc, ivec ci, ivec r)
      : nrow(nr), ncol(nc), colind(std::move(ci)), row(std::move(r)) {}

  int64_t nnz() const { return colind.empty() ? 0 : colind.back(); }
  int64_t numel() const { return nrow * ncol; }
  bool is_dense() const { return nnz() == numel(); }
  bool is_square() const { return nrow == ncol; }
  ivec get_col() const;
  // CasADi's compact form [nrow, ncol, colind..., row...]
  ivec compressed() const;
  static Sparsity from_compressed(const int64_t* sp);

  static Sparsity dense(int64_t nr, int64_t nc);
  static Sparsity diag(int64_t n);
  // Sparsity::triplet: sorted, duplicates removed; mapping[k] = nonzero index
  // of entry k (or -1 when duplicate entries are merged, only the last kept).
  static Sparsity triplet(int64_t nr, int64_t nc, const ivec& r, const ivec& c, ivec* mapping = nullptr);

  Sparsity T(ivec* mapping = nullptr) const;
  bool is_symmetric() const;
  bool operator==(const Sparsity& o) const {
    return nrow == o.nrow && ncol == o.ncol && colind == o.colind && row == o.row;
  }
  bool operator!=(const Sparsity& o) const { return !(*this == o); }
  // Pattern of the product this * B
  Sparsity mtimes(const Sparsity& B) const;
  // Entries present in both patterns (Sparsity::intersect, the pattern
  // operator* of casadi/core/sparsity.cpp:462-465; same dimensions required)
  Sparsity intersect(const Sparsity& B) const;
  // Insert empty rows / columns: row k of this becomes row rr[k] of an
  // nr-row pattern (enlargeRows), column k becomes column cc[k] (enlargeColumns)
  Sparsity enlarge(int64_t nr, int64_t nc, const ivec& rr, const ivec& cc) const;
  // A(rr, cc) for index vectors rr, cc (CasADi Sparsity::sub); mapping gives
  // the nonzero of this for every nonzero of the result.
  Sparsity sub(const ivec& rr, const ivec& cc, ivec* mapping = nullptr) const;
  // For every entry (nz of other) the nonzero index in this pattern or -1.
  ivec get_nz(const ivec& rows, const ivec& cols) const;
  int64_t get_nz(int64_t r, int64_t c) const;
  // pmult (casadi/core/sparsity_internal.cpp:3546-3600)
  Sparsity pmult(const ivec& p, bool permute_rows, bool permute_columns, bool invert) const;

  // Colourings (casadi/core/sparsity_internal.cpp:2959-3509).  A null result
  // (ncol == -1) means the cutoff was exceeded.
  Sparsity uni_coloring(const Sparsity& AT, int64_t cutoff) const;
  Sparsity star_coloring(int64_t ordering = 1, int64_t cutoff = INT64_MAX) const;
  ivec largest_first() const;

  // Orderings and symbolic factorisations
  ivec amd() const;
  ivec etree(bool ata = false) const;
  // Sparsity::ldl: pattern of L^T (strictly upper), permutation p
  Sparsity ldl(ivec& p, bool amd = true) const;
  // Sparsity::qr_sparse
  void qr_sparse(Sparsity& V, Sparsity& R, ivec& prinv, ivec& pc, bool amd = true) const;

  bool is_null() const { return ncol < 0; }
  static Sparsity null() {
    Sparsity s;
    s.ncol = -1;
    return s;
  }
};

// Low-level routines on CasADi compact sparsity arrays.
void sp_etree(const int64_t* sp, int64_t* parent, int64_t* w, bool ata);
void sp_postorder(const int64_t* parent, int64_t n, int64_t* post, int64_t* w);
void sp_ldl_colind(const int64_t* sp, int64_t* parent, int64_t* l_colind, int64_t* w);
void sp_ldl_row(const int64_t* sp, const int64_t* parent, int64_t* l_colind, int64_t* l_row,
                int64_t* w);

}  // namespace sxh
