/*
 * mplbm ensemble — many independent packings of uneven sizes advanced by
 * one persistent launch per step on one MI355X.
 *
 * No counterpart in the reference (it runs one compile-time domain per
 * process, include/lbm_model.cuh:4-6).  Each sample has its own size,
 * geometry and physical parameters (its own mplbm_config); all samples share
 * the storage form and the arithmetic.  A step launches one persistent
 * kernel over a flattened work list of (sample, tile, x-segment) items built
 * from per-sample offset/descriptor tables; steps without density injection
 * are replayed from a captured hipGraph (two steps per graph, one per lattice
 * parity).
 */
#ifndef MPLBM_ENSEMBLE_H_
#define MPLBM_ENSEMBLE_H_

#include "mplbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mplbm_ensemble mplbm_ensemble;

/* n configurations (sizes, physics); storage and math must agree.
 * block_threads: workgroup size of the persistent kernel (256, 512 or 1024;
 * 0 = 512).  Tiles are chosen per sample to fit that size. */
mplbm_status mplbm_ensemble_create(const mplbm_config* cfgs, int32_t n, int32_t block_threads,
                                   mplbm_ensemble** out);
void mplbm_ensemble_destroy(mplbm_ensemble* ens);
const char* mplbm_ensemble_last_error(const mplbm_ensemble* ens);

/* Compact tile of one sample (sample = -1: every sample): tile_y x tile_z
 * interior nodes, seg_x planes per work item (0 keeps the fitted value of
 * that field), as mplbm_set_tiling; tile_y * tile_z must not exceed the
 * workgroup size, which with it sets the interior nodes per thread.
 * Results do not depend on the tiles; used by the ensemble autotuner
 * (multiphase_lbm_amd/ensemble.py).  On failure the previous tiles stay. */
mplbm_status mplbm_ensemble_set_tiling(mplbm_ensemble* ens, int32_t sample, int32_t tile_y, int32_t tile_z,
                                       int32_t seg_x);

/* Per-sample set-up, as mplbm_set_geometry / mplbm_init_fields. */
mplbm_status mplbm_ensemble_set_geometry(mplbm_ensemble* ens, int32_t sample, const uint8_t* is_fluid);
mplbm_status mplbm_ensemble_init_fields(mplbm_ensemble* ens, int32_t sample, const float* rho);

/* Advances every sample by nsteps steps (numbered step, step+1, ...).  With
 * MPLBM_STEP_WRITE_MACRO the last step writes rho, p, u of every sample.
 * use_graph = 1 replays a captured hipGraph when no sample injects density. */
mplbm_status mplbm_ensemble_step(mplbm_ensemble* ens, int64_t step, int32_t nsteps, uint32_t flags,
                                 int32_t use_graph);
/* As above; returns the HIP-event time of the launched work (synchronises). */
mplbm_status mplbm_ensemble_step_timed(mplbm_ensemble* ens, int64_t step, int32_t nsteps, uint32_t flags,
                                       int32_t use_graph, double* elapsed_ms);

/* Per-sample read-back and FP64 reductions (synchronise). */
mplbm_status mplbm_ensemble_read_fields(mplbm_ensemble* ens, int32_t sample, float* rho, float* pressure,
                                        float* u);
mplbm_status mplbm_ensemble_read_populations_f64(mplbm_ensemble* ens, int32_t sample, double* f);
mplbm_status mplbm_ensemble_reduce(mplbm_ensemble* ens, int32_t sample, mplbm_stats64* out);
mplbm_status mplbm_ensemble_info(const mplbm_ensemble* ens, int32_t sample, mplbm_info* out);

/* Work-list size, fluid nodes of all samples, and the block size in use. */
mplbm_status mplbm_ensemble_summary(const mplbm_ensemble* ens, int64_t* work_items, int64_t* fluid_nodes,
                                    int32_t* block_threads);

#ifdef __cplusplus
}
#endif

#endif /* MPLBM_ENSEMBLE_H_ */
