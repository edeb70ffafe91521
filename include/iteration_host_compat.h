/*
 * The reference's three host<->device entry points, with its arity and
 * argument meaning (include/iteration_host.h:30-60, defined in
 * src/iteration_host.cu:16-81, called from src/multiphase_lbm.cu:584-593),
 * on top of include/mplbm.h.  They act on one process-global context.
 *
 * A host driver written against the reference includes this header instead
 * of include/iteration_host.h and keeps its calls as they are:
 *
 *   transfer_parameter(&tau, &G, &carn_star, ..., &statPeriod);
 *   multiphaselbm_initialization(&drho, &dpsi, &dpressure, &dis_fluid,
 *                                &dcollision_f, &dstream_f, &dforce);
 *   for (; step < step_max; step++) multiphaselbm_iteration(step, statPeriod, reverse, inject_type);
 *
 * The seven pointers of multiphaselbm_initialization are, as in the
 * reference, addresses of the caller's array pointers (Realxyz**, ...,
 * include/lbm_model.cuh:18-22); they are taken as void* so that any of the
 * reference's fixed-size array types converts without a cast.
 *   rho_dd, pressure_dd, is_fluid_dd: read and written through.  *rho_dd is
 *     the initial density [x][y][z] (float) and receives rhos on solid
 *     nodes; *pressure_dd receives P(rho) (0 on solids); *is_fluid_dd is the
 *     bool mask.  Host, managed or device memory (copies use
 *     hipMemcpyDefault).  Density and pressure are written back after every
 *     multiphaselbm_iteration with step % statPeriod == 0, the only steps on
 *     which the reference's host reads them (src/multiphase_lbm.cu:595-606).
 *   psi_dd, collision_f_dd, stream_f_dd, force_dd: accepted and not used
 *     (may be NULL).  The library owns psi, both lattices and the force
 *     (fluid-node-only lattices in its own layout); the caller's buffers
 *     are never read or written.
 *
 * Domain size and storage are run-time values here
 * (mplbm_compat_set_domain replaces the #defines of
 * include/lbm_model.cuh:4-15; default 400^3 FP32 like the reference).
 *
 * Errors follow the reference's cudaErrorCheck (include/iteration_host.h:6-15):
 * the message is printed to stderr and the process exits with the status.
 * mplbm_compat_set_abort(0) is the equivalent of cudaAssert's abort=false:
 * the call prints and returns, and mplbm_compat_status() gives the status.
 */
#ifndef ITERATION_HOST_COMPAT_H_
#define ITERATION_HOST_COMPAT_H_

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run-time replacement of the compile-time lx, ly, lz and `real`
 * (include/lbm_model.cuh:4-15); storage is a mplbm_storage value.  Call
 * before multiphaselbm_initialization; returns 0 or a mplbm_status. */
int mplbm_compat_set_domain(int64_t lx, int64_t ly, int64_t lz, int32_t storage);

/* transfer_parameter (include/iteration_host.h:30-34, src/iteration_host.cu:16-41):
 * the same twelve inputs through pointers.  inject_density and statPeriod
 * are accepted and unused, as on the reference's device. */
void transfer_parameter(const float* tau, const float* G, const bool* carn_star, const float* T_Tc,
                        const float* rhol_spinodal, const float* rhog_spinodal, const bool* inject_density,
                        const size_t* rho_inject_period, const float* rho_increment, const float* rhos,
                        const float* inject_position, const size_t* statPeriod);

/* multiphaselbm_initialization + multiphaselbm_init (include/iteration_host.h:45-48,
 * src/iteration_host.cu:44-70, include/iteration_kernel.cuh:9-29).  Every
 * pointer is the address of the caller's array pointer (see above).  The
 * library's state is initialised before the call returns (the reference
 * launches asynchronously and its caller synchronises,
 * src/multiphase_lbm.cu:590). */
void multiphaselbm_initialization(void* rho_dd, void* psi_dd, void* pressure_dd, void* is_fluid_dd,
                                  void* collision_f_dd, void* stream_f_dd, void* force_dd);

/* multiphaselbm_iteration (include/iteration_host.h:58-59, src/iteration_host.cu:73-81):
 * one step; on a statistics step (step % statPeriod == 0) density and
 * pressure are in the caller's arrays when it returns. */
void multiphaselbm_iteration(long step, size_t statPeriod, int reverse, unsigned int inject_type);

/* 1 (default): a failing call prints and exits with its status, as
 * cudaErrorCheck does; 0: it prints and returns. */
void mplbm_compat_set_abort(int abort_on_error);

/* Status of the last compat call (0 = success). */
int mplbm_compat_status(void);

/* Releases the global context (the reference frees in ~MultiphaseLBM,
 * src/multiphase_lbm.cu:1145-1160). */
void mplbm_compat_release(void);

#ifdef __cplusplus
}
#endif

#endif /* ITERATION_HOST_COMPAT_H_ */
