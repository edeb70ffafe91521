/*
 * Oracle for the D3Q19 Shan-Chen step: a plain-C, OpenMP CPU restatement of
 * the reference kernels (include/iteration_device.cuh, iteration_kernel.cuh,
 * src/iteration_host.cu:73-81) in float (the reference's precision) and in
 * double (the FP64 oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product library never
 * links or calls it.
 *
 * Exported entry points take a flat parameter vector `prm` (double[28], see
 * ORC_P_* below) so that ctypes callers need no struct definitions.
 *
 * Pins (tests/test_oracle_pins.py): D3Q19 moment identities, exact
 * permutation of streaming, single-node bounce-back, zero force for uniform
 * psi, the exact linear-psi force -G psi c_s^2 grad psi, FP64 mass and
 * momentum conservation, equilibrium moments, the Shan-Chen spinodal and
 * coexistence densities, Laplace's law, Poiseuille curvature, and orc_exp
 * against libm's exp/expf.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* D3Q19 velocity set in the reference's index order, include/lbm_model.cuh:49-54.
 * The opposite of i>=1 is i+9 (i<=9) or i-9 (i>9), lbm_model.cuh:24 with
 * iteration_kernel.cuh:113. */
static const int ORC_EX[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
static const int ORC_EY[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
static const int ORC_EZ[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
static const int ORC_OPP[19] = {0, 10, 11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8, 9};

/*
 * exp in double by Cody-Waite argument reduction and a Taylor polynomial:
 *   k = rint(x / ln2), r = (x - k*ln2_hi) - k*ln2_lo, |r| <= ln2/2,
 *   exp(x) = 2^k * sum_{j=0}^{13} r^j / j!     (Horner, explicit fma)
 * ln2_hi has its low 32 bits zero so k*ln2_hi is exact for |k| < 2^20
 * (the split published with fdlibm's e_exp.c).  The truncation error
 * |r|^14/14! < 5e-18 is below half an ulp of double for |r| <= 0.3466.
 * The GPU library evaluates the same sequence of IEEE operations, which is
 * what makes the two bit-identical; the tests bound this function against
 * libm exp (double) and expf (after narrowing).
 */
double orc_exp(double x) {
  if (x != x) return x;
  if (x > 709.0) return INFINITY;
  if (x < -745.0) return 0.0;
  const double LOG2E = 1.4426950408889634;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  double k = rint(x * LOG2E);
  double r = (x - k * LN2_HI) - k * LN2_LO;
  double p = 1.0 / 6227020800.0;              /* 1/13! */
  p = fma(p, r, 1.0 / 479001600.0);           /* 1/12! */
  p = fma(p, r, 1.0 / 39916800.0);            /* 1/11! */
  p = fma(p, r, 1.0 / 3628800.0);             /* 1/10! */
  p = fma(p, r, 1.0 / 362880.0);              /* 1/9!  */
  p = fma(p, r, 1.0 / 40320.0);               /* 1/8!  */
  p = fma(p, r, 1.0 / 5040.0);                /* 1/7!  */
  p = fma(p, r, 1.0 / 720.0);                 /* 1/6!  */
  p = fma(p, r, 1.0 / 120.0);                 /* 1/5!  */
  p = fma(p, r, 1.0 / 24.0);                  /* 1/4!  */
  p = fma(p, r, 1.0 / 6.0);                   /* 1/3!  */
  p = fma(p, r, 0.5);                         /* 1/2!  */
  p = fma(p, r, 1.0);                         /* 1/1!  */
  p = fma(p, r, 1.0);                         /* 1/0!  */
  return ldexp(p, (int)k);
}

/* flat parameter vector layout shared with oracle/oracle.py */
enum {
  ORC_P_LX = 0, ORC_P_LY, ORC_P_LZ, ORC_P_TAU, ORC_P_G, ORC_P_CARN_STAR, ORC_P_T_TC,
  ORC_P_RHOL_SP, ORC_P_RHOG_SP, ORC_P_INJ_PERIOD, ORC_P_RHO_INC, ORC_P_RHOS, ORC_P_INJ_POS,
  ORC_P_FACE_TYPE = 13, ORC_P_FACE_RHO = 19, ORC_P_BODY = 25, ORC_P_COUNT = 28
};

#define DEF_PARAMS(R, NAME)                                                      \
  typedef struct {                                                               \
    long lx, ly, lz;                                                             \
    R tau, G, T_Tc, rhol_spinodal, rhog_spinodal, rho_increment, rhos,            \
        inject_position;                                                         \
    int carn_star;                                                               \
    unsigned long rho_inject_period;                                             \
    int face_type[6];                                                            \
    R face_rho[6];                                                               \
    R body_force[3];                                                             \
  } NAME;                                                                        \
  static void NAME##_load(NAME* P, const double* p) {                            \
    P->lx = (long)p[ORC_P_LX]; P->ly = (long)p[ORC_P_LY]; P->lz = (long)p[ORC_P_LZ]; \
    P->tau = (R)p[ORC_P_TAU]; P->G = (R)p[ORC_P_G]; P->T_Tc = (R)p[ORC_P_T_TC];    \
    P->carn_star = p[ORC_P_CARN_STAR] != 0.0;                                    \
    P->rhol_spinodal = (R)p[ORC_P_RHOL_SP]; P->rhog_spinodal = (R)p[ORC_P_RHOG_SP]; \
    P->rho_inject_period = (unsigned long)p[ORC_P_INJ_PERIOD];                   \
    P->rho_increment = (R)p[ORC_P_RHO_INC]; P->rhos = (R)p[ORC_P_RHOS];          \
    P->inject_position = (R)p[ORC_P_INJ_POS];                                    \
    for (int f = 0; f < 6; ++f) {                                                \
      P->face_type[f] = (int)p[ORC_P_FACE_TYPE + f];                             \
      P->face_rho[f] = (R)p[ORC_P_FACE_RHO + f];                                 \
    }                                                                            \
    for (int a = 0; a < 3; ++a) P->body_force[a] = (R)p[ORC_P_BODY + a];         \
  }

DEF_PARAMS(float, orc_params__f32)
DEF_PARAMS(double, orc_params__f64)

#define REAL float
#define SUFFIX _f32
#include "oracle_impl.h"
#undef REAL
#undef SUFFIX

#define REAL double
#define SUFFIX _f64
#include "oracle_impl.h"
#undef REAL
#undef SUFFIX

/* ---- exported flat-parameter wrappers ---------------------------------- */
#define EXPORTS(R, S)                                                              \
  void orc##S##_init(const double* prm, const uint8_t* m, R* rho, R* psi, R* p,    \
                     R* fA, R* fB, R* F) {                                         \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    orc_init##S(&P, m, rho, psi, p, fA, fB, F);                                  \
  }                                                                                \
  void orc##S##_run(const double* prm, const uint8_t* m, R* fA, R* fB, R* rho,     \
                    R* p, R* psi, R* F, R* u, long step0, long nsteps, int reverse, \
                    unsigned inject_type) {                                        \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    orc_run##S(&P, m, fA, fB, rho, p, psi, F, u, step0, nsteps, reverse,         \
                inject_type);                                                      \
  }                                                                                \
  void orc##S##_stream(const double* prm, const uint8_t* m, const R* fA, R* fB,    \
                       const R* rho, const R* F) {                                 \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    orc_stream##S(&P, m, fA, fB, rho, F);                                        \
  }                                                                                \
  void orc##S##_density(const double* prm, const uint8_t* m, const R* fB, R* rho,  \
                        R* p, R* psi, long step, int reverse, unsigned itype) {    \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    orc_density##S(&P, m, fB, rho, p, psi, step, reverse, itype);                \
  }                                                                                \
  void orc##S##_force(const double* prm, const uint8_t* m, const R* psi, R* F) {   \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    orc_force##S(&P, m, psi, F);                                                 \
  }                                                                                \
  void orc##S##_collide(const double* prm, const uint8_t* m, R* fB, const R* rho,  \
                        R* F, R* u) {                                              \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    orc_collide##S(&P, m, fB, rho, F, u);                                        \
  }                                                                                \
  R orc##S##_psi(const double* prm, R rho) {                                       \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    return psi_##S(&P, rho);                                                      \
  }                                                                                \
  R orc##S##_press(const double* prm, R rho) {                                     \
    orc_params_##S P; orc_params_##S##_load(&P, prm);                           \
    return press_##S(&P, rho);                                                    \
  }

EXPORTS(float, _f32)
EXPORTS(double, _f64)

/* velocity tables for the test suite */
void orc_tables(int* ex, int* ey, int* ez, int* opp, double* w32, double* w64) {
  orc_params__f32 P32; orc_params__f64 P64;
  (void)P32; (void)P64;
  for (int i = 0; i < 19; ++i) {
    ex[i] = ORC_EX[i]; ey[i] = ORC_EY[i]; ez[i] = ORC_EZ[i]; opp[i] = ORC_OPP[i];
    w32[i] = (double)wt__f32(i);
    w64[i] = wt__f64(i);
  }
}
