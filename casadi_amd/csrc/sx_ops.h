// Scalar operation table shared by the host symbolic layer, the host VM and the
// gfx950 kernels.
//
// Operation codes are numbered exactly as casadi::Operation
// (casadi/core/calculus.hpp:60-218) so that an SXFunction algorithm_ stream
// produced by CasADi can be handed to this library unchanged.  The value and
// partial-derivative rules restate casadi_math<double>::fun / der
// (casadi/core/calculus.hpp:598-1018, dispatch at 1302-1439).
#pragma once

#if defined(__HIPCC_RTC__)
// Run-time compiled straight-line kernels: device math comes with hiprtc.
#define SXH_HD __device__ __forceinline__
#define SXH_HD_NOINLINE static __device__ __attribute__((noinline))
#ifndef NAN
#define NAN __builtin_nan("")
#endif
#ifndef HUGE_VAL
#define HUGE_VAL __builtin_huge_val()
#endif
#elif defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SXH_HD __host__ __device__ __forceinline__
// Out-of-line: keeps the constants and temporaries of the transcendental
// implementations out of the interpreter loops (otherwise hoisted there and
// held in registers for the whole kernel).
#define SXH_HD_NOINLINE static __host__ __device__ __attribute__((noinline))
#else
#include <cmath>
#define SXH_HD inline
#define SXH_HD_NOINLINE static __attribute__((noinline, unused))
#endif

namespace sxh {

enum Op : int {
  OP_ASSIGN = 0,
  OP_ADD, OP_SUB, OP_MUL, OP_DIV,
  OP_NEG, OP_EXP, OP_LOG, OP_POW, OP_CONSTPOW,
  OP_SQRT, OP_SQ, OP_TWICE,
  OP_SIN, OP_COS, OP_TAN,
  OP_ASIN, OP_ACOS, OP_ATAN,
  OP_LT, OP_LE, OP_EQ, OP_NE, OP_NOT, OP_AND, OP_OR,
  OP_FLOOR, OP_CEIL, OP_FMOD, OP_FABS, OP_SIGN, OP_COPYSIGN, OP_IF_ELSE_ZERO,
  OP_ERF, OP_FMIN, OP_FMAX,
  OP_INV,
  OP_SINH, OP_COSH, OP_TANH,
  OP_ASINH, OP_ACOSH, OP_ATANH,
  OP_ATAN2,            // 43
  OP_CONST,            // 44
  OP_INPUT, OP_OUTPUT, // 45, 46
  OP_PARAMETER,        // 47
  OP_CALL,             // 48
  OP_ERFINV = 86,
  OP_PRINTME = 87,
  OP_LIFT = 88,
  OP_LOG1P = 93,
  OP_EXPM1 = 94,
  OP_HYPOT = 95,
  OP_REMAINDER = 97,
  OP_NUM = 100
};

// One VM instruction: 16 bytes, the same layout as casadi::ScalarAtomic
// (casadi/core/sx_function.hpp:37-44) -- op, destination slot, then either a
// constant or the two source slots.
struct Instr {
  int op;
  int i0;
  union {
    double d;
    struct { int i1, i2; };
  };
};
static_assert(sizeof(Instr) == 16, "Instr must be 16 bytes");

// ndeps: casadi_math<T>::ndeps (casadi/core/calculus.hpp:1633-1647)
SXH_HD int op_ndeps(int op) {
  switch (op) {
    case OP_CONST: case OP_PARAMETER: case OP_INPUT: return 0;
    case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_POW: case OP_CONSTPOW:
    case OP_LT: case OP_LE: case OP_EQ: case OP_NE: case OP_AND: case OP_OR:
    case OP_COPYSIGN: case OP_FMOD: case OP_REMAINDER: case OP_FMIN: case OP_FMAX:
    case OP_ATAN2: case OP_PRINTME: case OP_LIFT: case OP_HYPOT: case OP_IF_ELSE_ZERO:
      return 2;
    case OP_CALL: return -1;
    default: return 1;
  }
}

SXH_HD bool op_is_binary(int op) { return op_ndeps(op) == 2; }

SXH_HD bool op_is_unary(int op) {
  switch (op) {
    case OP_ASSIGN: case OP_NEG: case OP_EXP: case OP_LOG: case OP_SQRT: case OP_SQ:
    case OP_TWICE: case OP_SIN: case OP_COS: case OP_TAN: case OP_ASIN: case OP_ACOS:
    case OP_ATAN: case OP_FLOOR: case OP_CEIL: case OP_NOT: case OP_ERF: case OP_FABS:
    case OP_SIGN: case OP_INV: case OP_SINH: case OP_COSH: case OP_TANH: case OP_ASINH:
    case OP_ACOSH: case OP_ATANH: case OP_ERFINV: case OP_LOG1P: case OP_EXPM1:
      return true;
    default:
      return false;
  }
}

// Whether the partial of op with respect to operand `which` (0 or 1) is the
// constant zero in casadi_math<SXElem>::der (casadi/core/calculus.hpp:
// 690-1018): floor, ceil, sign, not and the comparisons and logic ops in
// every operand, the exponent of constpow, the sign operand of copysign, the
// second operand of printme and lift, and the condition of if_else_zero.  A
// symbolic reverse or forward sweep multiplies by that zero SXElem, which
// simplifies away (x * 0 -> 0, w + 0 -> w), so no contribution is made: the
// numeric sweeps skip it too (an infinite or NaN seed times it would give
// NaN where CasADi's derivative Functions give nothing).
SXH_HD bool op_zero_partial(int op, int which) {
  switch (op) {
    case OP_FLOOR: case OP_CEIL: case OP_SIGN: case OP_NOT:
    case OP_LT: case OP_LE: case OP_EQ: case OP_NE: case OP_AND: case OP_OR:
      return true;
    case OP_IF_ELSE_ZERO:
      return which == 0;
    case OP_CONSTPOW: case OP_COPYSIGN: case OP_PRINTME: case OP_LIFT:
      return which == 1;
    default:
      return false;
  }
}

// True for the operations the numeric VM can execute (everything but the
// structural CONST/INPUT/OUTPUT/PARAMETER/CALL codes).
SXH_HD bool op_is_arith(int op) { return op_is_unary(op) || op_is_binary(op); }

SXH_HD double sxh_sign(double x) { return x < 0 ? -1 : x > 0 ? 1 : x; }

// std::fmin / std::fmax as the reference's host library computes them: a NaN
// operand gives the other one; otherwise the first operand when it is
// strictly smaller (larger), else the second -- so a tie of -0 and +0 gives
// the second operand.  IEEE 754 leaves that tie open, and the GPU's own
// min / max instructions order -0 below +0; spelling the rule out keeps the
// bits of the CPU evaluation.
SXH_HD double sxh_fmin(double x, double y) { return x != x ? y : y != y ? x : (x < y ? x : y); }
SXH_HD double sxh_fmax(double x, double y) { return x != x ? y : y != y ? x : (x > y ? x : y); }
SXH_HD double sxh_if_else_zero(double c, double y) { return c == 0 ? 0 : y; }

// Inverse error function: CasADi's own rational approximation followed by two
// Newton polishing steps (casadi/core/calculus.hpp:300-327), used by CasADi
// whenever the C library lacks erfinv (glibc does).
SXH_HD double sxh_erfinv(double x) {
  const double pi = 3.14159265358979323846;
  if (x >= 1) {
    return x == 1 ? HUGE_VAL : NAN;
  } else if (x <= -1) {
    return x == -1 ? -HUGE_VAL : NAN;
  } else if (x < -0.7) {
    double z = sqrt(-log((1.0 + x) / 2.0));
    return -(((1.641345311 * z + 3.429567803) * z - 1.624906493) * z - 1.970840454) /
           ((1.637067800 * z + 3.543889200) * z + 1.0);
  }
  double y;
  if (x < 0.7) {
    double z = x * x;
    y = x * (((-0.140543331 * z + 0.914624893) * z - 1.645349621) * z + 0.886226899) /
        ((((-0.329097515 * z + 0.012229801) * z + 1.442710462) * z - 2.118377725) * z + 1.0);
  } else {
    double z = sqrt(-log((1.0 - x) / 2.0));
    y = (((1.641345311 * z + 3.429567803) * z - 1.624906493) * z - 1.970840454) /
        ((1.637067800 * z + 3.543889200) * z + 1.0);
  }
  y = y - (erf(y) - x) / (2.0 / sqrt(pi) * exp(-y * y));
  y = y - (erf(y) - x) / (2.0 / sqrt(pi) * exp(-y * y));
  return y;
}

// x / y for 2^-900 <= |x| <= 2^900 and 2^-64 <= |y| <= 2^64, given yi = the
// correctly rounded 1 / y: the correctly rounded quotient, by two corrections
// of q = x * yi with exact FMA remainders (x - y q is representable in that
// range).  q is within 1.5 ulp of x / y; after the first correction it is
// within half an ulp plus ~2^-51 ulp (faithful), and then q + (x - y q) yi
// rounded to nearest is the correctly rounded x / y (Markstein's theorem: a
// faithful quotient and a reciprocal within half an ulp; P. Markstein,
// "Computation of elementary functions on the IBM RISC System/6000
// processor", IBM J. Res. Dev. 34(1), 1990).
SXH_HD double sxh_div_markstein(double x, double y, double yi) {
  double q = x * yi;
  q = fma(fma(-q, y, x), yi, q);
  return fma(fma(-q, y, x), yi, q);
}

// x / c for a constant c with 2^-64 <= |c| <= 2^64 and ci = RN(1 / c), on
// the device: sxh_div_markstein on the magnitudes, then v_div_fixup, which
// gives the sign and the IEEE results of zero, infinite and NaN dividends.
// That is x / c bit for bit except for finite nonzero |x| outside
// [2^-900, 2^900] (exponent from frexp outside [-899, 901]), where the
// remainders may round; such a dividend sets `wide`, and the caller then
// evaluates again with IEEE division.  Short dependent chain (multiply, four
// FMAs, fixup) and no branch: the kernels that use it are latency-bound
// single wavefronts whose scheduler needs long straight-line blocks.
#if defined(__HIPCC_RTC__) || defined(__HIP_DEVICE_COMPILE__)
SXH_HD double sxh_div_c(double x, double c, double ci, bool& wide) {
  wide |= static_cast<unsigned>(__builtin_amdgcn_frexp_exp(x) + 899) > 1800u;
  return __builtin_amdgcn_div_fixup(sxh_div_markstein(fabs(x), fabs(c), fabs(ci)), c, x);
}
#else
// The same on the host (the kernels' host emulation): frexp's exponent is 0
// for zero, infinity and NaN as v_frexp_exp_i32_f64 gives it, and the fixup
// follows v_div_fixup_f64 for a finite nonzero divisor c and a quotient q >= 0.
SXH_HD double sxh_div_c(double x, double c, double ci, bool& wide) {
  int e = 0;
  if (std::isfinite(x)) std::frexp(x, &e);
  wide |= static_cast<unsigned>(e + 899) > 1800u;
  const double q = sxh_div_markstein(std::fabs(x), std::fabs(c), std::fabs(ci));
  const bool neg = std::signbit(x) != std::signbit(c);
  if (std::isnan(x)) return x + x;
  if (std::isinf(x)) return neg ? -HUGE_VAL : HUGE_VAL;
  if (x == 0) return neg ? -0.0 : 0.0;
  return neg ? -q : q;
}
#endif

// sin(x) and cos(x) from one argument reduction on the device (the values
// equal sin(x) and cos(x) computed separately); two calls on the host.
SXH_HD void sxh_sincos(double x, double* s, double* c) {
#if defined(__HIPCC_RTC__) || defined(__HIP_DEVICE_COMPILE__)
  sincos(x, s, c);
#else
  *s = sin(x);
  *c = cos(x);
#endif
}

// The same pair returned by value from one out-of-line copy.  A long
// straight-line kernel calls it instead of inlining the reduction and
// polynomials at every site (about 1.3 KB of gfx950 code each), which keeps
// the kernel's code small enough to stay in the instruction cache.
struct SinCos {
  double s, c;
};
SXH_HD_NOINLINE SinCos sxh_sincos_call(double x) {
  SinCos r;
  sxh_sincos(x, &r.s, &r.c);
  return r;
}

SXH_HD double op_fun_full(int op, double x, double y);
SXH_HD void op_der_full(int op, double x, double y, double f, double* d);

// Out-of-line entry points of the full switches (one copy per kernel).
SXH_HD_NOINLINE double op_fun_all(int op, double x, double y) { return op_fun_full(op, x, y); }
SXH_HD_NOINLINE void op_der_all(int op, double x, double y, double f, double* d) {
  op_der_full(op, x, y, f, d);
}

// f = op(x, y).  Unary operations ignore y.  The IEEE-exact and comparison
// operations are inline; transcendental ones go through op_fun_all.
SXH_HD double op_fun(int op, double x, double y) {
  switch (op) {
    case OP_ADD: return x + y;
    case OP_SUB: return x - y;
    case OP_MUL: return x * y;
    case OP_DIV: return x / y;
    case OP_NEG: return -x;
    case OP_SQ: return x * x;
    case OP_TWICE: return 2. * x;
    case OP_SQRT: return sqrt(x);
    case OP_INV: return 1. / x;
    case OP_FABS: return fabs(x);
    case OP_LT: return x < y;
    case OP_LE: return x <= y;
    case OP_EQ: return x == y;
    case OP_NE: return x != y;
    case OP_IF_ELSE_ZERO: return sxh_if_else_zero(x, y);
    case OP_FMIN: return sxh_fmin(x, y);
    case OP_FMAX: return sxh_fmax(x, y);
    default: return op_fun_all(op, x, y);
  }
}

// Partial derivatives, split the same way.
SXH_HD void op_der(int op, double x, double y, double f, double* d) {
  switch (op) {
    case OP_ADD: d[0] = 1; d[1] = 1; return;
    case OP_SUB: d[0] = 1; d[1] = -1; return;
    case OP_MUL: d[0] = y; d[1] = x; return;
    case OP_DIV: d[0] = 1 / y; d[1] = -f / y; return;
    case OP_NEG: d[0] = -1; d[1] = 0; return;
    case OP_SQ: d[0] = x + x; d[1] = 0; return;
    case OP_TWICE: d[0] = 2; d[1] = 0; return;
    case OP_SQRT: d[0] = 1 / (f + f); d[1] = 0; return;
    case OP_INV: d[0] = -f * f; d[1] = 0; return;
    case OP_IF_ELSE_ZERO: d[0] = 0; d[1] = x; return;
    default: op_der_all(op, x, y, f, d);
  }
}

// Every operation.  Inlined with a constant op it folds to the one case.
SXH_HD double op_fun_full(int op, double x, double y) {
  switch (op) {
    case OP_ASSIGN: return x;
    case OP_ADD: return x + y;
    case OP_SUB: return x - y;
    case OP_MUL: return x * y;
    case OP_DIV: return x / y;
    case OP_NEG: return -x;
    case OP_EXP: return exp(x);
    case OP_LOG: return log(x);
    case OP_POW: return pow(x, y);
    case OP_CONSTPOW: return pow(x, y);
    case OP_SQRT: return sqrt(x);
    case OP_SQ: return x * x;
    case OP_TWICE: return 2. * x;
    case OP_SIN: return sin(x);
    case OP_COS: return cos(x);
    case OP_TAN: return tan(x);
    case OP_ASIN: return asin(x);
    case OP_ACOS: return acos(x);
    case OP_ATAN: return atan(x);
    case OP_LT: return x < y;
    case OP_LE: return x <= y;
    case OP_EQ: return x == y;
    case OP_NE: return x != y;
    case OP_NOT: return !x;
    case OP_AND: return x && y;
    case OP_OR: return x || y;
    case OP_FLOOR: return floor(x);
    case OP_CEIL: return ceil(x);
    case OP_FMOD: return fmod(x, y);
    case OP_REMAINDER: return remainder(x, y);
    case OP_FABS: return fabs(x);
    case OP_SIGN: return sxh_sign(x);
    case OP_COPYSIGN: return copysign(x, y);
    case OP_IF_ELSE_ZERO: return sxh_if_else_zero(x, y);
    case OP_ERF: return erf(x);
    case OP_FMIN: return sxh_fmin(x, y);
    case OP_FMAX: return sxh_fmax(x, y);
    case OP_INV: return 1. / x;
    case OP_SINH: return sinh(x);
    case OP_COSH: return cosh(x);
    case OP_TANH: return tanh(x);
    case OP_ASINH: return asinh(x);
    case OP_ACOSH: return acosh(x);
    case OP_ATANH: return atanh(x);
    case OP_ATAN2: return atan2(x, y);
    case OP_ERFINV: return sxh_erfinv(x);
    case OP_PRINTME: return x;
    case OP_LIFT: return x;
    case OP_LOG1P: return log1p(x);
    case OP_EXPM1: return expm1(x);
    case OP_HYPOT: return hypot(x, y);
    default: return NAN;
  }
}

// Partial derivatives d[0] = df/dx, d[1] = df/dy given f = op(x, y).
SXH_HD void op_der_full(int op, double x, double y, double f, double* d) {
  const double pi = 3.14159265358979323846;
  d[1] = 0;
  switch (op) {
    case OP_ASSIGN: d[0] = 1; break;
    case OP_ADD: d[0] = 1; d[1] = 1; break;
    case OP_SUB: d[0] = 1; d[1] = -1; break;
    case OP_MUL: d[0] = y; d[1] = x; break;
    case OP_DIV: d[0] = 1 / y; d[1] = -f / y; break;
    case OP_NEG: d[0] = -1; break;
    case OP_EXP: d[0] = f; break;
    case OP_LOG: d[0] = 1 / x; break;
    case OP_POW: d[0] = y * pow(x, y - 1); d[1] = log(x) * f; break;
    case OP_CONSTPOW: d[0] = y * pow(x, y - 1); break;
    case OP_SQRT: d[0] = 1 / (f + f); break;
    case OP_SQ: d[0] = x + x; break;
    case OP_TWICE: d[0] = 2; break;
    case OP_SIN: d[0] = cos(x); break;
    case OP_COS: d[0] = -sin(x); break;
    case OP_TAN: { double c = cos(x); d[0] = 1 / (c * c); } break;
    case OP_ASIN: d[0] = 1 / sqrt(1 - x * x); break;
    case OP_ACOS: d[0] = -1 / sqrt(1 - x * x); break;
    case OP_ATAN: d[0] = 1 / (1 + x * x); break;
    case OP_LT: case OP_LE: case OP_EQ: case OP_NE: case OP_AND: case OP_OR:
      d[0] = 0; d[1] = 0; break;
    case OP_NOT: d[0] = 0; break;
    case OP_FLOOR: case OP_CEIL: d[0] = 0; break;
    case OP_FMOD: case OP_REMAINDER: d[0] = 1; d[1] = (f - x) / y; break;
    case OP_FABS: d[0] = sxh_sign(x); break;
    case OP_SIGN: d[0] = 0; break;
    case OP_COPYSIGN: d[0] = copysign(1.0, y); break;
    case OP_IF_ELSE_ZERO: d[0] = 0; d[1] = x; break;
    case OP_ERF: d[0] = (2 / sqrt(pi)) * exp(-x * x); break;
    case OP_FMIN: {
      double a = x <= y, b = y <= x, c = a + b;
      d[0] = a / c; d[1] = b / c;
    } break;
    case OP_FMAX: {
      double a = y <= x, b = x <= y, c = a + b;
      d[0] = a / c; d[1] = b / c;
    } break;
    case OP_INV: d[0] = -f * f; break;
    case OP_SINH: d[0] = cosh(x); break;
    case OP_COSH: d[0] = sinh(x); break;
    case OP_TANH: d[0] = 1 - f * f; break;
    case OP_ASINH: d[0] = 1 / sqrt(1 + x * x); break;
    case OP_ACOSH: d[0] = 1 / sqrt(x - 1) / sqrt(x + 1); break;
    case OP_ATANH: d[0] = 1 / (1 - x * x); break;
    case OP_ATAN2: { double t = x * x + y * y; d[0] = y / t; d[1] = -x / t; } break;
    case OP_ERFINV: d[0] = (sqrt(pi) / 2) * exp(f * f); break;
    case OP_PRINTME: d[0] = 1; break;
    case OP_LIFT: d[0] = 1; break;
    case OP_LOG1P: d[0] = 1 / (1 + x); break;
    case OP_EXPM1: d[0] = exp(x); break;
    case OP_HYPOT: d[0] = x / f; d[1] = y / f; break;
    default: d[0] = NAN; d[1] = NAN; break;
  }
}

}  // namespace sxh
