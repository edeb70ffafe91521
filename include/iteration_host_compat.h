/*
 * Reference-named entry points on top of include/mplbm.h, for a host driver
 * written against the reference's host<->device layer
 * (include/iteration_host.h:30-60, src/iteration_host.cu:16-81).  They act
 * on one process-global context.
 *
 * Differences a caller sees: sizes are set at run time
 * (mplbm_compat_set_domain replaces the #defines of include/lbm_model.cuh:4-6);
 * multiphaselbm_initialization takes host arrays instead of device pointers
 * to caller-allocated unified memory, and the library copies density and
 * pressure back into them on every step with step % statPeriod == 0, which
 * are the only steps on which the reference's host reads them
 * (src/multiphase_lbm.cu:595-606).  Every function returns 0 on success and
 * a mplbm_status otherwise (the reference calls exit()).
 */
#ifndef ITERATION_HOST_COMPAT_H_
#define ITERATION_HOST_COMPAT_H_

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Runtime replacement of the compile-time lx, ly, lz and `real`
 * (include/lbm_model.cuh:4-15); storage is a mplbm_storage value. */
int mplbm_compat_set_domain(int64_t lx, int64_t ly, int64_t lz, int32_t storage);

/* transfer_parameter (src/iteration_host.cu:16-41): same twelve inputs. */
int transfer_parameter(const float* tau, const float* G, const bool* carn_star, const float* T_Tc,
                       const float* rhol_spinodal, const float* rhog_spinodal, const bool* inject_density,
                       const size_t* rho_inject_period, const float* rho_increment, const float* rhos,
                       const float* inject_position, const size_t* statPeriod);

/* multiphaselbm_initialization + multiphaselbm_init
 * (src/iteration_host.cu:44-70, include/iteration_kernel.cuh:9-29): rho is
 * the initial density [x][y][z] and receives rhos on solid nodes, pressure
 * receives P(rho) (0 on solids); both stay registered for
 * multiphaselbm_iteration.  psi, the two lattices and the force live in
 * the library. */
int multiphaselbm_initialization(float* rho, float* pressure, const bool* is_fluid);

/* multiphaselbm_iteration (src/iteration_host.cu:73-81): one step. */
int multiphaselbm_iteration(long step, size_t statPeriod, int reverse, unsigned int inject_type);

/* Releases the global context (the reference frees in ~MultiphaseLBM,
 * src/multiphase_lbm.cu:1145-1160). */
void mplbm_compat_release(void);

#ifdef __cplusplus
}
#endif

#endif /* ITERATION_HOST_COMPAT_H_ */
