/*
 * CPU restatement of the reference D3Q19 Shan-Chen step, generic over REAL.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is included twice by oracle.c, once
 * with REAL=float (the reference's own precision, `using real = float`,
 * include/lbm_model.cuh:15) and once with REAL=double (the FP64 oracle).  It
 * is the checker for the HIP library, never part of the product path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Layout follows the reference: every field is [x][y][z] with z fastest and
 * populations are [x][y][z][19] (AoS), include/lbm_model.cuh:18-22.
 *
 * Semantics (reference file:line in brackets):
 *   init          [include/iteration_kernel.cuh:9-29]
 *   stream        [include/iteration_kernel.cuh:51-62, iteration_device.cuh:252-271]
 *   boundary      [include/iteration_kernel.cuh:65-118, iteration_device.cuh:278-287]
 *   density       [include/iteration_device.cuh:40-63]
 *   force         [include/iteration_device.cuh:69-152]
 *   velocity/BGK  [include/iteration_device.cuh:162-206]
 *   kernel order  [src/iteration_host.cu:73-81]
 * with one deliberate difference: the reference's do_collision writes psi of
 * its own node while neighbours read psi in the same launch (a data race,
 * iteration_device.cuh:62 vs :89-147).  Here every psi of step n is computed
 * before any force is evaluated ("two-phase" semantics), which is the only
 * schedule-independent reading of that code.
 *
 * Type promotions mirror C++ rules on the reference source (for REAL=float):
 * `1. - exp(x)` and `2.0 *` are evaluated in double and narrowed, the BGK
 * update is evaluated in double (iteration_device.cuh:166-173), 1.0/rho and
 * 1.0/tau are double divisions narrowed to float.  Build with
 * -ffp-contract=off so that no multiply-add is fused.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)

/* Reference weights t (== Shan-Chen w), include/lbm_model.cuh:35-46:
 * t0 = 1.0/3.0, t1 = 1.0/18.0, t2 = 1.0/36.0 narrowed to `real`. */
static REAL FN(wt_)(int i) {
  const REAL t0 = (REAL)(1.0 / 3.0);
  const REAL t1 = (REAL)(1.0 / 18.0);
  const REAL t2 = (REAL)(1.0 / 36.0);
  static const int cls[19] = {0, 1, 1, 1, 2, 2, 2, 2, 2, 2,
                              1, 1, 1, 2, 2, 2, 2, 2, 2};
  return cls[i] == 0 ? t0 : (cls[i] == 1 ? t1 : t2);
}

/* exp(-rho) as the reference's `exp(-rhoxyz)` on a `real` argument:
 * float -> float exp for REAL=float, double exp for REAL=double. */
static REAL FN(expneg_)(REAL r) { return (REAL)orc_exp(-(double)r); }

/* dPress, include/iteration_device.cuh:7-27 */
static REAL FN(press_)(const FN(orc_params_) * P, REAL rho) {
  if (P->carn_star) {
    REAL a = (REAL)1.0, b = (REAL)4.0, R = (REAL)1.0, Tc = (REAL)0.0943;
    REAL T = P->T_Tc * Tc;
    REAL eta = (REAL)((double)(b * rho) / 4.0);
    REAL eta2 = eta * eta;
    REAL eta3 = eta2 * eta;
    REAL rho2 = rho * rho;
    REAL one_minus_eta = (REAL)(1.0 - (double)eta);
    REAL ome3 = one_minus_eta * one_minus_eta * one_minus_eta;
    return rho * R * T * ((REAL)1 + eta + eta2 - eta3) / ome3 - a * rho2;
  } else {
    REAL cs2 = (REAL)(1.0 / 3.0);
    REAL psi = (REAL)(1.0 - (double)FN(expneg_)(rho));
    REAL psi2 = psi * psi;
    return cs2 * rho + cs2 * P->G / (REAL)2 * psi2;
  }
}

/* dPsi, include/iteration_device.cuh:31-38 */
static REAL FN(psi_)(const FN(orc_params_) * P, REAL rho) {
  if (P->carn_star) {
    REAL cs2 = (REAL)(1.0 / 3.0);
    REAL p = FN(press_)(P, rho);
    return (REAL)sqrt(2.0 * (double)(p - cs2 * rho) / (double)(cs2 * P->G));
  }
  return (REAL)(1.0 - (double)FN(expneg_)(rho));
}

static inline size_t FN(cell_)(const FN(orc_params_) * P, long x, long y, long z) {
  return ((size_t)x * (size_t)P->ly + (size_t)y) * (size_t)P->lz + (size_t)z;
}

/* Upstream neighbour W(x - e_i) with the reference's periodic wrap
 * (iteration_device.cuh:256-266). */
static inline size_t FN(upstream_)(const FN(orc_params_) * P, long x, long y, long z, int i) {
  long nx = x - ORC_EX[i], ny = y - ORC_EY[i], nz = z - ORC_EZ[i];
  if (nx < 0) nx += P->lx; if (nx >= P->lx) nx -= P->lx;
  if (ny < 0) ny += P->ly; if (ny >= P->ly) ny -= P->ly;
  if (nz < 0) nz += P->lz; if (nz >= P->lz) nz -= P->lz;
  return FN(cell_)(P, nx, ny, nz);
}

/* Pressure-face test: does direction i at (x,y,z) enter through a face whose
 * boundary type is not periodic?  Returns the face index 0..5
 * (xmin,xmax,ymin,ymax,zmin,zmax) or -1.  Mirrors the face tests of
 * treat_boundary (iteration_kernel.cuh:73-110). */
static int FN(entering_face_)(const FN(orc_params_) * P, long x, long y, long z, int i) {
  /* the reference applies the face rules in x, y, z order, each overwriting
   * the last, so the last matching non-periodic face decides */
  int f = -1;
  if (x == 0 && ORC_EX[i] == 1 && P->face_type[0]) f = 0;
  if (x == P->lx - 1 && ORC_EX[i] == -1 && P->face_type[1]) f = 1;
  if (y == 0 && ORC_EY[i] == 1 && P->face_type[2]) f = 2;
  if (y == P->ly - 1 && ORC_EY[i] == -1 && P->face_type[3]) f = 3;
  if (z == 0 && ORC_EZ[i] == 1 && P->face_type[4]) f = 4;
  if (z == P->lz - 1 && ORC_EZ[i] == -1 && P->face_type[5]) f = 5;
  return f;
}

/* multiphaselbm_init, include/iteration_kernel.cuh:9-29.  rho is in/out:
 * solid nodes are overwritten with rhos.  F of fluid nodes is zeroed here
 * (the reference leaves it uninitialised; it is rewritten before first use
 * because do_collision computes it before compute_velocity). */
void FN(orc_init)(const FN(orc_params_) * P, const uint8_t* is_fluid, REAL* rho,
                  REAL* psi, REAL* pres, REAL* fA, REAL* fB, REAL* F) {
  const size_t n = (size_t)P->lx * P->ly * P->lz;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)n; ++c) {
    for (int i = 0; i < 19; ++i) {
      fA[(size_t)c * 19 + i] = FN(wt_)(i) * rho[c];
      fB[(size_t)c * 19 + i] = FN(wt_)(i) * rho[c];
    }
    if (is_fluid[c]) {
      psi[c] = FN(psi_)(P, rho[c]);
      pres[c] = FN(press_)(P, rho[c]);
    } else {
      rho[c] = P->rhos;
      psi[c] = FN(psi_)(P, P->rhos);
      pres[c] = (REAL)0;
    }
    F[3 * c + 0] = F[3 * c + 1] = F[3 * c + 2] = (REAL)0;
  }
}

/* do_streaming + treat_boundary: B from A for every fluid node.
 * Pull B_i(x) = A_i(W(x-e_i)) if that node is fluid, else halfway
 * bounce-back B_i(x) = A_opp(i)(x); B_0(x) = A_0(x).  At a non-periodic
 * face (pressure type 1) the entering direction uses the anti-bounce-back
 * rule of added_f_at_pressure_bound (iteration_device.cuh:213-243) with a
 * prescribed face density.  There is no node across a non-periodic face, so
 * the wrapped node's solid flag is not consulted there (the reference's face
 * branches are unreachable, global_var.cuh:39-63, so nothing pins this). */
void FN(orc_stream)(const FN(orc_params_) * P, const uint8_t* is_fluid,
                    const REAL* fA, REAL* fB, const REAL* rho, const REAL* F) {
#pragma omp parallel for schedule(static)
  for (long x = 0; x < P->lx; ++x)
    for (long y = 0; y < P->ly; ++y)
      for (long z = 0; z < P->lz; ++z) {
        const size_t c = FN(cell_)(P, x, y, z);
        if (!is_fluid[c]) continue;
        REAL* b = fB + c * 19;
        const REAL* a = fA + c * 19;
        b[0] = a[0];
        for (int i = 1; i < 19; ++i) {
          const size_t up = FN(upstream_)(P, x, y, z, i);
          const int face = FN(entering_face_)(P, x, y, z, i);
          if (face >= 0) {
            /* anti-bounce-back at a prescribed density rho_w:
             * B_i = -A_opp(i) + 2 t_i rho_w (1 + 4.5 (e_i.u)^2 - 1.5 u^2).
             * The reference evaluates u from A, tau*F and the stored rho
             * (iteration_device.cuh:220-239); F is never computed on a
             * non-periodic face (:79-87), so here u = sum(A e)/sum(A). */
            REAL ux = 0, uy = 0, uz = 0, ra = 0;
            for (int k = 0; k < 19; ++k) {
              ra += a[k];
              ux += a[k] * (REAL)ORC_EX[k];
              uy += a[k] * (REAL)ORC_EY[k];
              uz += a[k] * (REAL)ORC_EZ[k];
            }
            REAL inv_rho = (REAL)(1.0 / (double)ra);
            ux *= inv_rho; uy *= inv_rho; uz *= inv_rho;
            REAL usq = ux * ux + uy * uy + uz * uz;
            REAL eu = (REAL)ORC_EX[i] * ux + (REAL)ORC_EY[i] * uy + (REAL)ORC_EZ[i] * uz;
            REAL rw = P->face_rho[face];
            /* `2 * t[i] * rho_w * (1 + 4.5*eu*eu - 1.5*u_squ)` narrowed to
             * real, then added to -A_opp in real (iteration_device.cuh:242,282) */
            REAL added = (REAL)((double)((REAL)2 * FN(wt_)(i) * rw) *
                                (1.0 + 4.5 * (double)eu * (double)eu - 1.5 * (double)usq));
            b[i] = -a[ORC_OPP[i]] + added;
          } else if (!is_fluid[up]) {
            b[i] = a[ORC_OPP[i]];
          } else {
            b[i] = fA[up * 19 + i];
          }
        }
      }
}

/* compute_density for every fluid node (first phase of do_collision). */
void FN(orc_density)(const FN(orc_params_) * P, const uint8_t* is_fluid, const REAL* fB,
                     REAL* rho, REAL* pres, REAL* psi, long step, int reverse,
                     unsigned inject_type) {
  const int inject_now = P->rho_inject_period > 0 && (unsigned long)step % P->rho_inject_period == 0;
  const REAL inc = P->rho_increment * (REAL)reverse;
#pragma omp parallel for schedule(static)
  for (long x = 0; x < P->lx; ++x)
    for (long y = 0; y < P->ly; ++y)
      for (long z = 0; z < P->lz; ++z) {
        const size_t c = FN(cell_)(P, x, y, z);
        if (!is_fluid[c]) continue;
        REAL d = (REAL)0;
        for (int i = 0; i < 19; ++i) d += fB[c * 19 + i];
        /* density injection, iteration_device.cuh:47-59 (uses y, as the
         * reference does) */
        if (inject_now && (REAL)y < P->inject_position * (REAL)P->ly) {
          if (inject_type == 0) d += inc;
          else if (inject_type == 1) { if (d < P->rhol_spinodal) d += inc; }
          else if (inject_type == 2) { if (d > P->rhol_spinodal) d += inc; }
        }
        rho[c] = d;
        pres[c] = FN(press_)(P, d);
        psi[c] = FN(psi_)(P, d);
      }
}

/* compute_force, include/iteration_device.cuh:69-152: for a fluid node,
 * F_a = (-G psi(x)) * sum_{i=1..18} (w_i psi(W(x+e_i))) e_{i,a},
 * summed left to right in ascending i.  Nodes on a non-periodic face keep
 * F = 0 (the reference skips them, :79-87). */
void FN(orc_force)(const FN(orc_params_) * P, const uint8_t* is_fluid, const REAL* psi, REAL* F) {
#pragma omp parallel for schedule(static)
  for (long x = 0; x < P->lx; ++x)
    for (long y = 0; y < P->ly; ++y)
      for (long z = 0; z < P->lz; ++z) {
        const size_t c = FN(cell_)(P, x, y, z);
        if (!is_fluid[c]) continue;
        if ((P->face_type[0] || P->face_type[1]) && (x == 0 || x == P->lx - 1)) { F[3*c] = F[3*c+1] = F[3*c+2] = 0; continue; }
        if ((P->face_type[2] || P->face_type[3]) && (y == 0 || y == P->ly - 1)) { F[3*c] = F[3*c+1] = F[3*c+2] = 0; continue; }
        if ((P->face_type[4] || P->face_type[5]) && (z == 0 || z == P->lz - 1)) { F[3*c] = F[3*c+1] = F[3*c+2] = 0; continue; }
        REAL sx = 0, sy = 0, sz = 0;
        for (int i = 1; i < 19; ++i) {
          /* downstream neighbour W(x+e_i) = upstream of the opposite direction */
          const size_t nb = FN(upstream_)(P, x, y, z, ORC_OPP[i]);
          REAL wp = FN(wt_)(i) * psi[nb];
          REAL tx = wp * (REAL)ORC_EX[i], ty = wp * (REAL)ORC_EY[i], tz = wp * (REAL)ORC_EZ[i];
          if (i == 1) { sx = tx; sy = ty; sz = tz; }
          else { sx = sx + tx; sy = sy + ty; sz = sz + tz; }
        }
        REAL gpsi = -P->G * psi[c];
        F[3 * c + 0] = gpsi * sx;
        F[3 * c + 1] = gpsi * sy;
        F[3 * c + 2] = gpsi * sz;
      }
}

/* compute_velocity + post_collision, include/iteration_device.cuh:162-206.
 * BGK in place on B.  Body force g (not in the reference; zero reproduces it)
 * adds rho*g to the Shan-Chen force before the velocity shift.  If u is
 * non-NULL the shifted velocity (the reference's u, never stored there) is
 * written for fluid nodes. */
void FN(orc_collide)(const FN(orc_params_) * P, const uint8_t* is_fluid, REAL* fB,
                     const REAL* rho, REAL* F, REAL* u) {
  const REAL inv_tau = (REAL)(1.0 / (double)P->tau);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < (long)((size_t)P->lx * P->ly * P->lz); ++c) {
    if (!is_fluid[c]) continue;
    REAL* b = fB + (size_t)c * 19;
    if (P->body_force[0] != 0 || P->body_force[1] != 0 || P->body_force[2] != 0) {
      F[3 * c + 0] += rho[c] * P->body_force[0];
      F[3 * c + 1] += rho[c] * P->body_force[1];
      F[3 * c + 2] += rho[c] * P->body_force[2];
    }
    REAL ux = 0, uy = 0, uz = 0;
    for (int i = 0; i < 19; ++i) {
      ux += b[i] * (REAL)ORC_EX[i];
      uy += b[i] * (REAL)ORC_EY[i];
      uz += b[i] * (REAL)ORC_EZ[i];
    }
    ux += P->tau * F[3 * c + 0];
    uy += P->tau * F[3 * c + 1];
    uz += P->tau * F[3 * c + 2];
    REAL inv_rho = (REAL)(1.0 / (double)rho[c]);
    ux *= inv_rho; uy *= inv_rho; uz *= inv_rho;
    REAL usq = ux * ux + uy * uy + uz * uz;
    if (u) { u[3 * c + 0] = ux; u[3 * c + 1] = uy; u[3 * c + 2] = uz; }
    for (int i = 0; i < 19; ++i) {
      REAL eu = (REAL)ORC_EX[i] * ux + (REAL)ORC_EY[i] * uy + (REAL)ORC_EZ[i] * uz;
      REAL trho = FN(wt_)(i) * rho[c];
      b[i] = (REAL)((double)b[i] +
                    ((double)trho * (1.0 + 3.0 * (double)eu + 4.5 * (double)eu * (double)eu -
                                     1.5 * (double)usq) -
                     (double)b[i]) * (double)inv_tau);
    }
  }
}

/* One step of multiphaselbm_iteration (src/iteration_host.cu:73-81):
 * stream -> boundary -> density (all nodes) -> force -> velocity/BGK.
 * The caller swaps fA/fB afterwards (do_swap, iteration_kernel.cuh:44-48). */
void FN(orc_step)(const FN(orc_params_) * P, const uint8_t* is_fluid, const REAL* fA, REAL* fB,
                  REAL* rho, REAL* pres, REAL* psi, REAL* F, REAL* u, long step, int reverse,
                  unsigned inject_type) {
  FN(orc_stream)(P, is_fluid, fA, fB, rho, F);
  FN(orc_density)(P, is_fluid, fB, rho, pres, psi, step, reverse, inject_type);
  FN(orc_force)(P, is_fluid, psi, F);
  FN(orc_collide)(P, is_fluid, fB, rho, F, u);
}

/* Run nsteps steps from `step0`, swapping lattices; on return fA holds the
 * latest post-collision populations.  Returns nothing; the fields hold the
 * values of the last step. */
void FN(orc_run)(const FN(orc_params_) * P, const uint8_t* is_fluid, REAL* fA, REAL* fB,
                 REAL* rho, REAL* pres, REAL* psi, REAL* F, REAL* u, long step0, long nsteps,
                 int reverse, unsigned inject_type) {
  const size_t n19 = (size_t)P->lx * P->ly * P->lz * 19;
  for (long s = 0; s < nsteps; ++s) {
    FN(orc_step)(P, is_fluid, fA, fB, rho, pres, psi, F, (s == nsteps - 1) ? u : NULL, step0 + s,
                 reverse, inject_type);
    /* swap contents so that fA is always "collision" for the caller */
    REAL* t = fA; fA = fB; fB = t;
  }
  if (nsteps % 2 == 1) {
    /* after an odd number of steps the newest data sits in the caller's fB */
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)n19; ++k) { REAL v = fA[k]; fA[k] = fB[k]; fB[k] = v; }
  }
}

#undef FN
#undef CAT
#undef CAT2
