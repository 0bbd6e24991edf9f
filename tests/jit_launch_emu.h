// Host emulation of a launch of a generated kernel, for the tests.
//
// The generated sources are plain C++ apart from a few qualifiers.  Compiled
// with g++ after this header, every memory access of a kernel goes through a
// checked reference:
//  - bounds: the address must lie inside the buffer the base pointer was
//    taken from (buffers are registered with their exact byte sizes, as the
//    driver allocates them);
//  - races: within one launch, an address written by one lane may not be
//    read or written by any other lane (lanes of one launch run in no fixed
//    order on the GPU), counting every (blockIdx.y, blockIdx.x, threadIdx.x).
// The launch itself loops over the same grid the driver launches, the padded
// last wavefront included, and calls the kernel entry point unchanged (its
// `if (k >= n) return` included).
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <unordered_map>
#include <vector>

struct cxj_dim3 {
  unsigned x = 0, y = 0, z = 0;
};
static cxj_dim3 blockIdx, threadIdx, blockDim, gridDim;

#define __global__
#define __device__
#define __forceinline__ inline
#define __launch_bounds__(x)

namespace cxj_emu {

struct Region {
  uintptr_t lo, hi;  // [lo, hi)
  int id;
};
static std::vector<Region> regions;
static long long lane = 0;   // global lane of the access being made
static long long errors = 0;  // violations found (the first few are printed)
static long long accesses = 0;

struct Touch {
  long long writer = -1, reader = -1;
  bool many_readers = false;
};
static std::unordered_map<uintptr_t, Touch> touched;

static void report(const char* what, uintptr_t a, int region) {
  if (errors++ < 8) std::fprintf(stderr, "[emu] %s: address %#lx region %d lane %lld\n", what, (unsigned long)a, region, lane);
}

static int region_of(const void* p) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  for (const Region& r : regions)
    if (a >= r.lo && a < r.hi) return r.id;
  // one past the end is a valid base only for an empty buffer
  return -1;
}

// Whether the access is in bounds (an out-of-bounds access is reported and
// not made: reads give 0).
static bool access(const void* base, const double* p, bool write) {
  ++accesses;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  const int rb = region_of(base);
  int rp = -1;
  for (const Region& r : regions)
    if (a >= r.lo && a + sizeof(double) <= r.hi) rp = r.id;
  if (rb < 0 || rp != rb) {
    report(write ? "out-of-bounds write" : "out-of-bounds read", a, rb);
    return false;
  }
  Touch& t = touched[a];
  if (write) {
    if ((t.writer >= 0 && t.writer != lane) || t.many_readers || (t.reader >= 0 && t.reader != lane))
      report("race (write)", a, rb);
    t.writer = lane;
  } else {
    if (t.writer >= 0 && t.writer != lane) report("race (read after another lane's write)", a, rb);
    if (t.reader >= 0 && t.reader != lane) t.many_readers = true;
    t.reader = lane;
  }
  return true;
}

// A double in device memory: reads and writes are checked.
struct Ref {
  const void* base;
  double* p;
  operator double() const { return access(base, p, false) ? *p : 0.0; }
  Ref& operator=(double v) {
    if (access(base, p, true)) *p = v;
    return *this;
  }
  Ref& operator=(const Ref& o) { return *this = static_cast<double>(o); }
};

// A pointer into device memory that remembers the buffer it came from.
template <class T>
struct Ptr {
  const void* base;
  T* p;
  Ptr(T* q) : base(q), p(q) {}
  Ptr(T* q, const void* b) : base(b), p(q) {}
  operator Ptr<const T>() const { return Ptr<const T>(p, base); }  // double* -> const double*
  Ptr operator+(long long i) const { return Ptr(p + i, base); }
  Ref operator[](long long i) const { return Ref{base, const_cast<double*>(p + i)}; }
  explicit operator bool() const { return p != nullptr; }
};

}  // namespace cxj_emu

// Memory macros of the generated sources, checked (the source's own
// definitions are removed by the harness).
#define CXJ_AT(t, i, j) ((t).a[i].p + (long long)(j) * (t).a[i].s_nz + k * (t).a[i].s_inst)
#define CXJ_REF(base, ptr) (cxj_emu::Ref{(const void*)(base), const_cast<double*>(ptr)})
#define CXJ_IN(i, j) ((A.arg.a[i].p) ? (double)CXJ_REF(A.arg.a[i].p, CXJ_AT(A.arg, i, j)) : 0.0)
#define CXJ_OUT(i, j, v) do { if (A.res.a[i].p) CXJ_REF(A.res.a[i].p, CXJ_AT(A.res, i, j)) = (v); } while (0)
#define CXJ_SEED(i, j) ((A.aseed.a[i].p) ? (double)CXJ_REF(A.aseed.a[i].p, CXJ_AT(A.aseed, i, j)) : 0.0)
#define CXJ_SENS(i, j, v) do { if (A.asens.a[i].p) CXJ_REF(A.asens.a[i].p, CXJ_AT(A.asens, i, j)) = (v); } while (0)
#define CXJ_ST_IN(r) CXJ_REF(A.st_in, A.st_in + (long long)(r) * A.lds + k)
#define CXJ_ST_OUT(r) CXJ_REF(A.st_out, A.st_out + (long long)(r) * A.lds + k)
#define CXJ_ADJ_IN(r) CXJ_REF(A.adj_in, A.adj_in + (long long)(r) * A.lds + k)
#define CXJ_ADJ_OUT(r) CXJ_REF(A.adj_out, A.adj_out + (long long)(r) * A.lds + k)
#define CXJ_SNAP(p, r) CXJ_REF(p, (p) + (long long)(r) * A.lds + k)
#define CXJ_KEEP(x) ((void)0)
#define CXJ_KEEP_PTR(x) ((void)0)
#define CXJ_AFTER(p, x) ((void)0)
#define CXJ_OPAQUE(p) ((void)0)
#define CXJ_OPAQUE_AFTER(p, x) ((void)0)
// lanes run one at a time: a wavefront-wide vote is the lane's own value
#define CXJ_ANY(b) (b)

extern "C" {
// Register a device buffer [p, p + bytes) before the launches that use it.
void emu_clear_regions() { cxj_emu::regions.clear(); }
void emu_add_region(const void* p, long long bytes, int id) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  cxj_emu::regions.push_back({a, a + static_cast<uintptr_t>(bytes), id});
}
long long emu_errors() { return cxj_emu::errors; }
long long emu_accesses() { return cxj_emu::accesses; }
void emu_reset_errors() { cxj_emu::errors = 0; cxj_emu::accesses = 0; }
}

// The launch loop: every lane of a grid (gx, gy) of blocks of bx threads, the
// access log cleared at the start of each launch.
#define CXJ_EMU_LAUNCH_LOOP(call)                                                   \
  do {                                                                              \
    cxj_emu::touched.clear();                                                       \
    gridDim.x = gx;                                                                 \
    gridDim.y = gy;                                                                 \
    blockDim.x = bx;                                                                \
    for (unsigned y = 0; y < gy; ++y)                                               \
      for (unsigned b = 0; b < gx; ++b)                                             \
        for (unsigned t = 0; t < bx; ++t) {                                         \
          blockIdx.x = b;                                                           \
          blockIdx.y = y;                                                           \
          threadIdx.x = t;                                                          \
          cxj_emu::lane = (static_cast<long long>(y) * gx + b) * bx + t;            \
          call;                                                                     \
        }                                                                           \
  } while (0)
