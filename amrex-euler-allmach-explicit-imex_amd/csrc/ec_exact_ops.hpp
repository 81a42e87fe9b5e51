// ec_exact_ops.hpp -- IEEE-identical FP64 division and square root for gfx950
// with fewer instructions, for operands inside a stated magnitude window.
//
// hipcc lowers an FP64 a / b to
//   v_div_scale(b), v_rcp, 2 Newton steps (4 FMAs), v_div_scale(a), v_mul,
//   v_fma (residual), v_div_fmas, v_div_fixup
// (12 instructions), and sqrt(x) to a range test and ldexp pre-scaling,
// v_rsq, a Goldschmidt/Newton refinement (2 muls, 7 FMAs), ldexp post-scaling
// and a class test for +-0 and +inf (18 instructions).  The scaling steps are
// the identity unless an operand is near the ends of the exponent range
// (v_div_scale: a numerator below 2^-969, a denormal denominator or quotient,
// an exponent gap of 768 or more; sqrt: x below 2^-767).  v_div_fmas is a plain
// FMA when no scaling happened, and v_div_fixup returns the magnitude of its
// first operand with the quotient's sign unless an operand is 0, inf or NaN,
// in which case its result does not depend on the first operand.  So, inside
// the window:
//
//   recip(b)      = v_rcp + the same 2 Newton steps            (5 instructions)
//   quot(a, R)    = v_mul, v_fma, v_fma, v_div_fixup            (4 instructions)
//   quot_pos(a, R) = the same without v_div_fixup, for a numerator that is
//                   nonzero, or +0 over a positive denominator    (3 instructions)
//   sqrt_core(x)  = v_rsq + the same refinement                 (10 instructions)
//
// give the bits of a / b and sqrt(x) exactly, and one recip() serves every
// quotient with the same denominator.  Callers establish the window (the
// explicit step does so per tile, from the cell states) and fall back to the
// plain operators outside it.  Every step below is an explicit fma or a plain
// product: the file is compiled without FMA contraction.
#pragma once

#include <hip/hip_runtime.h>

namespace ec {

struct Recip {
  double b;  // the denominator
  double y;  // its refined reciprocal, as the division sequence forms it
};

__device__ __forceinline__ Recip recip(double b)
{
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  return Recip{b, y};
}

// a / r.b, bit for bit, for a and r.b inside the window (or a == 0).
__device__ __forceinline__ double quot(double a, const Recip& r)
{
  const double q = a * r.y;
  const double e = fma(-r.b, q, a);
  return __builtin_amdgcn_div_fixup(fma(e, r.y, q), r.b, a);
}

// a / r.b, bit for bit, for r.b inside the window and a numerator that is
// inside the window and nonzero, or +0 over a positive r.b.  Beyond the
// special operands (0, inf, NaN) v_div_fixup only gives the result the sign
// of a / b, and the refined quotient already carries that sign here: for
// a != 0 it is a * y corrected by a relative O(2^-52) term, and for a = +0,
// b > 0 every step yields +0 (the residual fma adds -0 and +0, which round
// to +0).  So v_div_fixup is the identity and is left out.
__device__ __forceinline__ double quot_pos(double a, const Recip& r)
{
  const double q = a * r.y;
  const double e = fma(-r.b, q, a);
  return fma(e, r.y, q);
}

// sqrt(x), bit for bit, for finite x >= 2^-767 (and NaN or negative x, which
// give NaN as sqrt does); not for +-0 or +inf.
__device__ __forceinline__ double sqrt_core(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// sqrt of a sum of squares that is either +0 or at least 2^-767.
__device__ __forceinline__ double sqrt_core0(double x) { return x > 0.0 ? sqrt_core(x) : x; }

// The window the explicit step's fast path needs, tested on the primitive
// state of every cell a tile reads: density and pressure in [2^-40, 2^40)
// (positive, finite), velocity components 0 or of magnitude in [2^-100, 2^40).
// Face states reconstructed with the minmod, MC or van Leer limiter lie between
// neighbouring cell values, so they stay in the window too, and every
// intermediate of the HLLC-P and HLLC main paths then stays between 2^-900
// and 2^300 in magnitude: no scaling step of the division or square-root
// sequence can trigger.
__device__ __forceinline__ bool pos_in_window(double x)
{
  // sign bit clear and biased exponent in [1023-40, 1023+40)
  const unsigned hi = (unsigned)(__double_as_longlong(x) >> 32);
  return hi - (983u << 20) < (80u << 20);
}

__device__ __forceinline__ bool vel_in_window(double x)
{
  const unsigned e = ((unsigned)(__double_as_longlong(x) >> 52)) & 0x7ffu;
  return (x == 0.0) | (e - 923u < 140u);
}

}  // namespace ec
