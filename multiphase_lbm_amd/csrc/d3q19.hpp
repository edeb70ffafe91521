// D3Q19 lattice tables and scalar physics shared by host and device code of
// libmplbm.  Velocity order, weights and opposite-direction rule follow the
// reference model file include/lbm_model.cuh:24,35-54 (the index order fixes
// the summation order of every moment, so it is kept).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mplbm {

constexpr int kQ = 19;

__host__ __device__ constexpr int ex(int i) {
  constexpr int t[kQ] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
  return t[i];
}
__host__ __device__ constexpr int ey(int i) {
  constexpr int t[kQ] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
  return t[i];
}
__host__ __device__ constexpr int ez(int i) {
  constexpr int t[kQ] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
  return t[i];
}
// opposite direction: i+9 for 1..9, i-9 for 10..18 (lbm_model.cuh:24, half = 9)
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : (i > 9 ? i - 9 : i + 9); }
// weight class: 0 rest (1/3), 1 axis (1/18), 2 diagonal (1/36)
__host__ __device__ constexpr int wclass(int i) {
  return i == 0 ? 0 : ((i <= 3 || (i >= 10 && i <= 12)) ? 1 : 2);
}
// weights as the reference declares them: double quotients narrowed to `real`
template <class R>
__host__ __device__ constexpr R weight(int i) {
  return wclass(i) == 0 ? (R)(1.0 / 3.0) : (wclass(i) == 1 ? (R)(1.0 / 18.0) : (R)(1.0 / 36.0));
}
__host__ __device__ constexpr float inv_weight_f(int i) {
  return wclass(i) == 0 ? 3.0f : (wclass(i) == 1 ? 18.0f : 36.0f);
}

// A double constant as a wave-uniform scalar operand at its point of use.
// In device code the empty asm makes the value opaque where it is used, so
// the compiler materialises it into a scalar register pair there instead of
// keeping a vector-register copy live across the caller's loop (the 13
// polynomial coefficients below held ~20 VGPRs of the marching kernel).
// The value is unchanged.
__host__ __device__ inline double sconst(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+s"(c));
#endif
  return c;
}

// exp in double: Cody-Waite reduction with the fdlibm ln2 split and a
// degree-13 Taylor polynomial evaluated by Horner's rule with fused
// multiply-adds.  Every operation is an IEEE-754 double operation, so the
// result is the same on the host and on gfx950.
__host__ __device__ inline double exp_cw(double x) {
  if (x != x) return x;
  if (x > 709.0) return __builtin_huge_val();
  if (x < -745.0) return 0.0;
  const double kLog2e = 1.4426950408889634;
  const double kLn2Hi = 6.93147180369123816490e-01;
  const double kLn2Lo = 1.90821492927058770002e-10;
  const double k = __builtin_rint(x * kLog2e);
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  double p = sconst(1.0 / 6227020800.0);
  p = __builtin_fma(p, r, sconst(1.0 / 479001600.0));
  p = __builtin_fma(p, r, sconst(1.0 / 39916800.0));
  p = __builtin_fma(p, r, sconst(1.0 / 3628800.0));
  p = __builtin_fma(p, r, sconst(1.0 / 362880.0));
  p = __builtin_fma(p, r, sconst(1.0 / 40320.0));
  p = __builtin_fma(p, r, sconst(1.0 / 5040.0));
  p = __builtin_fma(p, r, sconst(1.0 / 720.0));
  p = __builtin_fma(p, r, sconst(1.0 / 120.0));
  p = __builtin_fma(p, r, sconst(1.0 / 24.0));
  p = __builtin_fma(p, r, sconst(1.0 / 6.0));
  p = __builtin_fma(p, r, 0.5);
  p = __builtin_fma(p, r, 1.0);
  p = __builtin_fma(p, r, 1.0);
  return __builtin_ldexp(p, (int)k);
}

// Equation-of-state parameters in the arithmetic type R, converted once on
// the host the way the reference's float globals hold them.
template <class R>
struct Eos {
  R G;           // interaction strength
  R cs2;         // (R)(1.0/3.0)
  R half_cs2G;   // cs2*G/2 evaluated left to right in R (iteration_device.cuh:25)
  R T;           // Carnahan-Starling T = T_Tc * 0.0943 (iteration_device.cuh:13)
  int carn_star;
};

// dPress (include/iteration_device.cuh:7-27).  kCS = false compiles the
// Shan-Chen form only (for callers that know carn_star is 0).
template <bool kCS = true, class R>
__host__ __device__ inline R eos_pressure(const Eos<R>& e, R rho) {
  if (kCS && e.carn_star) {
    // a = R = 1 and b = 4 in the reference; products with 1 are exact
    const R eta = (R)((double)((R)4.0 * rho) / 4.0);
    const R eta2 = eta * eta;
    const R eta3 = eta2 * eta;
    const R rho2 = rho * rho;
    const R ome = (R)(1.0 - (double)eta);
    const R ome3 = ome * ome * ome;
    return rho * e.T * ((R)1 + eta + eta2 - eta3) / ome3 - rho2;
  }
  const R ex_ = (R)exp_cw(-(double)rho);
  const R psi = (R)(1.0 - (double)ex_);
  return e.cs2 * rho + e.half_cs2G * (psi * psi);
}

// dPsi (include/iteration_device.cuh:31-38); kCS as in eos_pressure.
template <bool kCS = true, class R>
__host__ __device__ inline R eos_psi(const Eos<R>& e, R rho) {
  if (kCS && e.carn_star) {
    const R p = eos_pressure(e, rho);
    return (R)__builtin_sqrt(2.0 * (double)(p - e.cs2 * rho) / (double)(e.cs2 * e.G));
  }
  const R ex_ = (R)exp_cw(-(double)rho);
  return (R)(1.0 - (double)ex_);
}

}  // namespace mplbm
