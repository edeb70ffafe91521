/*
 * mplbm — C ABI of the MI355X (gfx950) D3Q19 Shan-Chen multiphase
 * lattice-Boltzmann step.  libmplbm.so exports exactly the functions below;
 * no C++ or torch types cross this boundary.
 *
 * The reference drives its GPU path through three host functions
 * (include/iteration_host.h:30-60, src/iteration_host.cu:16-81), all of
 * them with compile-time domain sizes and device pointers owned by the
 * caller.  This ABI replaces them with a context that owns device memory,
 * takes runtime sizes, reports errors through status codes (the reference
 * calls exit(), include/iteration_host.h:6-15) and runs on its own HIP
 * stream.  iteration_host_compat.h keeps the three reference names on top of
 * this ABI.
 *
 * Field arrays crossing the ABI use the reference layout: [x][y][z], z
 * fastest (include/lbm_model.cuh:18-22); populations [x][y][z][19] in the
 * reference's velocity order (include/lbm_model.cuh:49-54).
 *
 * Threading: a context belongs to one host thread at a time.  mplbm_step is
 * asynchronous on the context stream; every call that returns host data
 * synchronises that stream first.
 */
#ifndef MPLBM_H_
#define MPLBM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mplbm_status {
  MPLBM_OK = 0,
  MPLBM_ERR_ARG = 1,         /* invalid argument or configuration */
  MPLBM_ERR_HIP = 2,         /* a HIP runtime call failed */
  MPLBM_ERR_STATE = 3,       /* call out of order (e.g. step before init) */
  MPLBM_ERR_NOMEM = 4,       /* device allocation failed */
  MPLBM_ERR_UNSUPPORTED = 5  /* option not available for this storage type */
} mplbm_status;

/* Population storage.  FP32 and FP64 evaluate the reference arithmetic
 * (FP32: the reference's own float storage with its double-promoted
 * sub-expressions, include/iteration_device.cuh:7-206).  FP16S/BF16S store
 * the shifted, scaled deviation (f_i/w_i - rho_ref)*scale in 16 bits and do
 * all arithmetic in FP32 registers. */
typedef enum mplbm_storage {
  MPLBM_FP32 = 0,
  MPLBM_FP64 = 1,
  MPLBM_FP16S = 2,
  MPLBM_BF16S = 3
} mplbm_storage;

/* Step kernel.  COMPACT (what AUTO selects): populations of fluid nodes
 * only, one marching stream-collide launch per step that works through the
 * fluid nodes of each plane, pseudopotential halo in LDS.  FUSED: the same
 * schedule over the full lattice (every node stored, solid ones skipped).
 * TWOPASS: full lattice, a density/psi launch followed by a force/collision
 * launch (the schedule the reference's do_collision needs to be race-free).
 * All three give identical results. */
typedef enum mplbm_kernel {
  MPLBM_KERNEL_AUTO = 0,
  MPLBM_KERNEL_FUSED = 1,
  MPLBM_KERNEL_TWOPASS = 2,
  MPLBM_KERNEL_COMPACT = 3
} mplbm_kernel;

/* Arithmetic of the FP32 form.  REFERENCE evaluates the reference's
 * operation sequence, including the double-promoted sub-expressions C++ gives
 * it (iteration_device.cuh:7-206); FAST evaluates the same model in FP32 with
 * fused multiply-adds and hardware reciprocal/exp. */
typedef enum mplbm_math { MPLBM_MATH_REFERENCE = 0, MPLBM_MATH_FAST = 1 } mplbm_math;

/* Halfway bounce-back in the FUSED kernel.  WRITE: after collision a node
 * also stores f*_opp(i) into slot i of each solid upstream neighbour, so the
 * next pull needs no mask test.  PULL: the pull reads the node's own
 * opposite population where the link word marks a solid upstream node
 * (the default: fewer scattered stores).  Both give identical results. */
typedef enum mplbm_bounce_back { MPLBM_BB_WRITE = 0, MPLBM_BB_PULL = 1 } mplbm_bounce_back;

/* Outer-face types (index order xmin, xmax, ymin, ymax, zmin, zmax).  The
 * reference's face flags (include/global_var.cuh:39-63) are fixed to
 * periodic; PRESSURE prescribes face_rho[f] through the anti-bounce-back
 * rule of include/iteration_device.cuh:213-243. */
typedef enum mplbm_face { MPLBM_FACE_PERIODIC = 0, MPLBM_FACE_PRESSURE = 1 } mplbm_face;

/* Rounding of the 16-bit deviations stored by each step (FP16S/BF16S with
 * the COMPACT kernel).  NEAREST: round to nearest even, the same rounding
 * every step, so a steady population keeps the same rounding error and mass
 * drifts linearly.  STOCHASTIC: a deterministic per-(step, node, population)
 * dither of one unit in the last place before rounding to nearest, which
 * makes the stored value unbiased; runs stay reproducible. */
typedef enum mplbm_rounding { MPLBM_ROUND_NEAREST = 0, MPLBM_ROUND_STOCHASTIC = 1 } mplbm_rounding;

/* Shift of the 16-bit deviations (FP16S/BF16S with the COMPACT kernel).
 * GLOBAL: one shift rho_ref for every population, g_i = (f_i/w_i -
 * rho_ref)*scale.  NODE: each node's populations are stored as deviations
 * from that node's density (rounded to the 16-bit type and stored beside
 * them, 2 B per node), g_i = (f_i/w_i - rho_node)*scale, so the stored
 * values stay small in both phases of a two-phase run; 4 more bytes per
 * update; combines with either rounding. */
typedef enum mplbm_shift { MPLBM_SHIFT_GLOBAL = 0, MPLBM_SHIFT_NODE = 1 } mplbm_shift;

/* All fields of transfer_parameter (src/iteration_host.cu:16-41) plus the
 * domain size, storage, faces, body force and kernel tiling.  Fill with
 * mplbm_config_default() and override. */
typedef struct mplbm_config {
  int64_t lx, ly, lz;          /* domain size (compile-time 400^3 in the reference) */
  int32_t storage;             /* mplbm_storage */
  int32_t kernel;              /* mplbm_kernel */
  double tau;                  /* BGK relaxation time */
  double G;                    /* Shan-Chen interaction strength */
  int32_t carn_star;           /* 1: Carnahan-Starling EOS, 0: Shan-Chen */
  int32_t math;                /* mplbm_math (FP32 storage only; 16-bit forms are FAST) */
  double T_Tc;                 /* CS EOS T/Tc */
  double rhol_spinodal;        /* liquid threshold (injection types 1/2, stats) */
  double rhog_spinodal;        /* vapour threshold (stats) */
  int64_t rho_inject_period;   /* injection fires when step % period == 0 (0: never) */
  double rho_increment;        /* injection increment */
  double rhos;                 /* solid density (wetting) */
  double inject_position;      /* injection applies where y < inject_position*ly */
  int32_t face_type[6];        /* mplbm_face per face */
  double face_rho[6];          /* prescribed density of PRESSURE faces */
  double body_force[3];        /* acceleration g; the force density rho*g is added to
                                  the Shan-Chen force (0 reproduces the reference) */
  double rho_ref;              /* FP16S/BF16S shift; 0 = choose from the initial field */
  double dev_scale;            /* FP16S/BF16S scale; 0 = choose automatically */
  int32_t tile_y, tile_z;      /* FUSED: threads per tile in y and z incl. the ring; COMPACT:
                                  interior nodes per tile in y and z (0 = tuned default) */
  int32_t seg_x;               /* FUSED, COMPACT: planes marched per block (0 = tuned default) */
  int32_t plane_pad;           /* extra elements per population sub-plane (-1 = automatic) */
  int32_t bounce_back;         /* mplbm_bounce_back: how FUSED resolves halfway bounce-back */
  int32_t block_threads;       /* COMPACT: workgroup size 256, 512 or 1024 (0 = tuned default), at
                                  least the tile's interior node count */
  int32_t rounding;            /* mplbm_rounding of the 16-bit forms */
  int32_t shift;               /* mplbm_shift of the 16-bit forms */
} mplbm_config;

/* FP64 reductions over fluid nodes (mplbm_reduce). */
typedef struct mplbm_stats64 {
  double mass;                 /* sum of all populations of the newest lattice */
  double momentum[3];          /* sum of f_i e_i of the newest lattice */
  double fluid_nodes;
  double liquid_nodes;         /* rho > rhol_spinodal (from the last macro write) */
  double vapor_nodes;          /* rho < rhog_spinodal */
  double liquid_rho_sum, vapor_rho_sum;
  double liquid_p_sum, vapor_p_sum;
  double rho_min, rho_max, p_min, p_max;
  double nonfinite;            /* populations that are NaN or Inf */
} mplbm_stats64;

/* Statistics of MultiphaseLBM::calc_cluster (src/multiphase_lbm.cu:225-562)
 * with the extrema of write_stats (:1057-1083), from the last macro write
 * (mplbm_calc_cluster).  Same fields as the host harness's
 * mplbm_cluster_stats. */
typedef struct mplbm_clusters {
  double min_rho, max_rho, min_p, max_p;
  double avg_liq_press, avg_vap_press, saturation, avg_liq_rho, avg_vap_rho;
  int64_t liq_cluster_count, gas_cluster_count;
  double cluster_order, wet_coord_num;
  double ratio_liquid_solid_boundary, ratio_liquid_vapor_boundary;
  double cluster_radius;
  int64_t fluid_count, liquid_count, vapor_count;
} mplbm_clusters;

/* Read-only facts about a context (mplbm_info). */
typedef struct mplbm_info {
  int64_t fluid_nodes;
  int64_t cells;
  int32_t storage, kernel;
  int32_t tile_y, tile_z, seg_x;
  int32_t bytes_per_pop;       /* storage bytes of one population */
  double rho_ref, dev_scale;   /* deviation storage parameters in use */
  int64_t device_bytes;        /* device memory held by the context */
} mplbm_info;

typedef struct mplbm_ctx mplbm_ctx;

/* Step flags. */
#define MPLBM_STEP_WRITE_MACRO 1u /* the last step of the call writes rho, p, u */

/* Defaults of the reference constructor (src/multiphase_lbm.cu:35-80),
 * domain 400^3 (include/lbm_model.cuh:4-6), FP32, AUTO kernel. */
void mplbm_config_default(mplbm_config* cfg);

/* Replaces transfer_parameter (src/iteration_host.cu:16-41) and the
 * allocation half of MultiphaseLBM::resize_gpu (src/multiphase_lbm.cu:83-109):
 * validates cfg, creates the stream and allocates both lattices and fields. */
mplbm_status mplbm_create(const mplbm_config* cfg, mplbm_ctx** out);

/* Frees device memory and the stream (src/multiphase_lbm.cu:1145-1160). */
void mplbm_destroy(mplbm_ctx* ctx);

/* Text of the last error of ctx, or of the last failed mplbm_create when
 * ctx is NULL.  Valid until the next call on the same context. */
const char* mplbm_last_error(const mplbm_ctx* ctx);

/* Solid mask, host uint8 [x][y][z], nonzero = fluid (the reference's
 * is_fluid, include/global_var.cuh:11).  Builds the per-node bounce-back
 * link words on the GPU; for COMPACT also the plane ranks, the per-tile work
 * lists and the population lattices sized to the fluid count. */
mplbm_status mplbm_set_geometry(mplbm_ctx* ctx, const uint8_t* is_fluid);

/* Replaces multiphaselbm_initialization + multiphaselbm_init
 * (src/iteration_host.cu:44-70, include/iteration_kernel.cuh:9-29):
 * both lattices = t_i*rho, solid nodes take rho = rhos.  rho is host
 * [x][y][z]; the _f64 variant keeps FP64 initial data exact. */
mplbm_status mplbm_init_fields(mplbm_ctx* ctx, const float* rho);
mplbm_status mplbm_init_fields_f64(mplbm_ctx* ctx, const double* rho);

/* Replaces multiphaselbm_iteration (src/iteration_host.cu:73-81): runs
 * nsteps steps numbered step, step+1, ... asynchronously on the context
 * stream.  reverse and inject_type are the reference's injection controls.
 * With MPLBM_STEP_WRITE_MACRO the last step also writes rho, p and the
 * shifted velocity u to device fields (the reference writes rho and p
 * every step, include/iteration_device.cuh:60-61; its host reads them only
 * on statistics steps). */
mplbm_status mplbm_step(mplbm_ctx* ctx, int64_t step, int32_t reverse, uint32_t inject_type,
                        int32_t nsteps, uint32_t flags);

/* As mplbm_step, and returns the time of the launched work measured with HIP
 * events recorded on the context stream around it (synchronises). */
mplbm_status mplbm_step_timed(mplbm_ctx* ctx, int64_t step, int32_t reverse,
                              uint32_t inject_type, int32_t nsteps, uint32_t flags,
                              double* elapsed_ms);

/* Waits for the context stream. */
mplbm_status mplbm_sync(mplbm_ctx* ctx);

/* Copies the fields of the last macro write to host [x][y][z] arrays
 * (any pointer may be NULL; u is [x][y][z][3]).  Replaces the reference
 * host's unified-memory reads of rho_d/pressure_d
 * (src/multiphase_lbm.cu:243-262,1061-1069).  Solid nodes read rho = rhos,
 * p = 0, u = 0. */
mplbm_status mplbm_read_fields(mplbm_ctx* ctx, float* rho, float* pressure, float* u);
mplbm_status mplbm_read_fields_f64(mplbm_ctx* ctx, double* rho, double* pressure, double* u);

/* Newest post-collision populations as host [x][y][z][19] doubles (decoded
 * from deviation storage) and the inverse, for exact checkpoint/restart
 * (the reference restarts from a 5-digit density VTK only,
 * src/multiphase_lbm.cu:112-132). */
mplbm_status mplbm_read_populations_f64(mplbm_ctx* ctx, double* f);
mplbm_status mplbm_write_populations_f64(mplbm_ctx* ctx, const double* f);

/* Exact checkpoint (the reference can only restart from a density VTK,
 * src/multiphase_lbm.cu:112-132): the stored values of the newest
 * populations of every fluid node ([x][y][z] order, 19 each, in the
 * context's storage form, with its deviation reference and scale), the step
 * number and the domain, storage and geometry they belong to.  A loaded
 * context continues bit for bit as the saved one would have.  Loading checks
 * the domain, storage and fluid mask (set the geometry first) and returns
 * the saved step in *step (may be NULL).  Synchronise. */
mplbm_status mplbm_save_checkpoint(mplbm_ctx* ctx, const char* path, int64_t step);
mplbm_status mplbm_load_checkpoint(mplbm_ctx* ctx, const char* path, int64_t* step);

/* FP64 reductions of the newest lattice and of the last macro write
 * (synchronises). */
mplbm_status mplbm_reduce(mplbm_ctx* ctx, mplbm_stats64* out);

/* Solid numbering for the cluster statistics: host [x][y][z] array as
 * place_sphere writes it (src/multiphase_lbm.cu:893-927: 0 at fluid cells,
 * the 1-based sphere id at solid cells, the last sphere written wins), and
 * the number of solids (the wet coordination number's divisor).  Copied to
 * the device once.  NULL clears it (cluster order and wet coordination
 * number are then 0). */
mplbm_status mplbm_set_solid_numbers(mplbm_ctx* ctx, const int64_t* solid_number, int64_t solid_count);

/* calc_cluster on the device (src/multiphase_lbm.cu:225-562; the reference
 * runs it on the host at every statPeriod after reading rho and p through
 * unified memory, :243-262): phase counts, FP64 phase sums, extrema,
 * liquid-solid and liquid-vapour boundary ratios, liquid and vapour cluster
 * labelling (6-neighbour, periodic, clusters of fewer than 7 nodes dropped,
 * numbered in the order of their first cell with z fastest), cluster order,
 * wet coordination number and the single-cluster radius, from the fields of
 * the last macro write (mplbm_step with MPLBM_STEP_WRITE_MACRO).
 * liq_label / gas_label (host [x][y][z], may be NULL) receive the cluster
 * numbers of the VTK LiqClusters / GasClusters blocks.  Synchronises.
 * Needs fewer than 2^31 cells. */
mplbm_status mplbm_calc_cluster(mplbm_ctx* ctx, double rhol_spinodal, double rhog_spinodal, mplbm_clusters* out,
                                 int32_t* liq_label, int32_t* gas_label);

/* Sphere voxeliser on the device: place_sphere (src/multiphase_lbm.cu:
 * 893-927) for count spheres x, y, z, R (host float[count][4]) in list
 * order, as read_positions (:1106-1126) places the spheres of a positions
 * file.  The same float-stepped sample points, dist2 < R^2 test, nearest()
 * rounding and single periodic wrap as the reference's loop (cells beyond
 * one wrap are skipped).  is_fluid (host uint8 [x][y][z], lx*ly*lz) is
 * updated in place: covered cells become 0.  solid_number (host int64
 * [x][y][z], may be NULL) receives first_id + k for the last sphere k of
 * the list that covers the cell; other cells keep their value.  Every
 * coordinate +/- R must lie within +/-2^23 and R <= 510.  Synchronous, on
 * the calling thread's current device; errors go to mplbm_last_error(NULL). */
mplbm_status mplbm_place_spheres(int64_t lx, int64_t ly, int64_t lz, const float* spheres, int64_t count,
                                 int64_t first_id, uint8_t* is_fluid, int64_t* solid_number);

/* mplbm_place_spheres without any allocation, for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * is_fluid_dev (uint8 [x][y][z]) and solid_number_dev (int64 [x][y][z], may
 * be NULL) are device memory of the caller, updated in place exactly as
 * mplbm_place_spheres updates its host arrays; spheres is host memory as
 * there.  The per-cell sphere numbers and the staged sphere list live in one
 * static device buffer of 68 MiB per GPU (held until the process ends):
 * at most 2^18 spheres per call, and at most 2^24 cells when
 * solid_number_dev is given; larger calls are refused with MPLBM_ERR_ARG
 * before any device work.  Runs on the null stream of the current device
 * and synchronises it.  Calls that overlap on one GPU (other host threads)
 * share the buffer and race.  Errors go to mplbm_last_error(NULL). */
mplbm_status mplbm_place_spheres_noalloc(int64_t lx, int64_t ly, int64_t lz, const float* spheres, int64_t count,
                                         int64_t first_id, uint8_t* is_fluid_dev, int64_t* solid_number_dev);

/* Changes the FUSED or COMPACT kernel's tile (see mplbm_config.tile_y) and
 * planes per block; 0 keeps the tuned default for that field.  Used by the
 * autotuner (multiphase_lbm_amd/autotune.py); results do not depend on the
 * tile. */
mplbm_status mplbm_set_tiling(mplbm_ctx* ctx, int32_t tile_y, int32_t tile_z, int32_t seg_x);

/* Context facts (no synchronisation). */
mplbm_status mplbm_info_get(const mplbm_ctx* ctx, mplbm_info* out);

/* Library version string. */
const char* mplbm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MPLBM_H_ */
