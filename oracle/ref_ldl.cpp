// Test infrastructure: exposes the reference's own casadi_ldl and casadi_det
// runtimes, compiled from casadi/core/runtime/casadi_ldl.hpp and
// casadi_det.hpp where they lie in the reference tree (include path set by
// oracle/Makefile), through C entry points for ctypes.
// casadi_types.hpp supplies casadi_int, exactly as in the reference build.
#include "casadi/core/casadi_types.hpp"
#include "casadi/core/runtime/casadi_ldl.hpp"
#include "casadi/core/runtime/casadi_det.hpp"

extern "C" {

void ref_ldl(const long long* sp_a, const double* a, const long long* sp_lt, double* lt, double* d,
             const long long* p, double* w) {
  casadi_ldl(sp_a, a, sp_lt, lt, d, p, w);
}

void ref_ldl_solve(double* x, long long nrhs, const long long* sp_lt, const double* lt, const double* d,
                   const long long* p, double* w) {
  casadi_ldl_solve(x, nrhs, sp_lt, lt, d, p, w);
}

double ref_det(const long long* sp_v, const double* nz_v, const long long* sp_r, const double* nz_r,
               const double* beta) {
  return casadi_det(sp_v, nz_v, sp_r, nz_r, beta);
}

}  // extern "C"
