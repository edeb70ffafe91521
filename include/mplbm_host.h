/*
 * mplbm_host — C ABI of the host harness around libmplbm (libmplbm_host.so):
 * the reference's input-file parser, sphere voxeliser, positions and sample
 * loaders, density initialisation, statistics and VTK/Stats writers, and the
 * driver loop, re-implemented in C++17 with runtime domain sizes.
 *
 * Field arrays are [x][y][z] (z fastest) as in the reference
 * (include/lbm_model.cuh:18-22).  Functions return 0 on success and a
 * negative value on error; mplbm_host_last_error() gives the text.
 */
#ifndef MPLBM_HOST_H_
#define MPLBM_HOST_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parsed input file (tokens of src/multiphase_lbm.cu:631-725 plus the
 * runtime-size, storage, kernel, boundary, statistics and checkpoint tokens). */
typedef struct mplbm_input {
  int64_t lx, ly, lz;
  char result_folder[256];
  int32_t num_threads;
  int64_t step, step_max;
  int64_t vtkPeriod, vtrPeriod, tecplotPeriod, statPeriod;
  int32_t save_vtk_ascii, save_vtr_binary, save_tecplot_ascii;
  double rho0, rhos, G, IniPerturbRate, tau;
  double res;
  int64_t rho_inject_period;
  double rho_increment;
  int32_t inject_density;
  uint32_t inject_type;
  double inject_position;
  int32_t carn_star;
  double T_Tc, rhol_spinodal, rhog_spinodal;
  int32_t read_densities;
  char densities_filename[256];
  int32_t hysteresis;
  double max_saturation, min_saturation;
  int32_t erosion;
  double target_density;
  /* build extensions */
  int32_t storage;          /* mplbm_storage */
  int32_t kernel;           /* mplbm_kernel */
  int32_t math;             /* mplbm_math */
  int32_t face_type[6];
  double face_rho[6];
  double body_force[3];
  int32_t unknown_tokens;   /* tokens reported as undefined */
  int32_t host_statistics;  /* 1: calc_cluster on the host after a field read
                               (the reference's way); 0: mplbm_calc_cluster */
  int32_t read_checkpoint;  /* 1: resume from checkpoint_filename (token read_checkpoint) */
  int64_t checkpointPeriod; /* exact checkpoint every N steps to <result_folder>/checkpoint.mplbm */
  char checkpoint_filename[256];
  int32_t rounding;         /* mplbm_rounding (token rounding nearest|stochastic) */
  int32_t device_voxeliser; /* 1: read_positions and generate_solids voxelise on the
                               device (mplbm_place_spheres); 0: on the host */
  int32_t shift;            /* mplbm_shift (token shift global|node) */
  int32_t saturation_cycle; /* 1 (token saturation_cycle 1): cycle injection and drainage
                               between max_saturation and min_saturation; 0: reverse stays
                               1 as in the reference, which parses both tokens but never
                               acts on them */
} mplbm_input;

/* Statistics of MultiphaseLBM::calc_cluster (src/multiphase_lbm.cu:225-562). */
typedef struct mplbm_cluster_stats {
  double min_rho, max_rho, min_p, max_p;
  double avg_liq_press, avg_vap_press, saturation, avg_liq_rho, avg_vap_rho;
  int64_t liq_cluster_count, gas_cluster_count;
  double cluster_order, wet_coord_num;
  double ratio_liquid_solid_boundary, ratio_liquid_vapor_boundary;
  double cluster_radius;
  int64_t fluid_count, liquid_count, vapor_count;
} mplbm_cluster_stats;

const char* mplbm_host_last_error(void);

/* Input file defaults (src/multiphase_lbm.cu:35-80, domain 400^3). */
void mplbm_host_input_default(mplbm_input* in);

/* Token parser only (no geometry work): fills `in` from the file.  Tokens
 * that act on the domain (resize, read_sample, place_sphere,
 * generate_solids, read_positions) are counted but not executed. */
int mplbm_host_parse(const char* path, mplbm_input* in);

/* place_sphere (src/multiphase_lbm.cu:893-927): float stepping, sample-point
 * test dist2 < R^2, nearest() rounding, one periodic wrap.  Marks is_fluid
 * = 0 and solid_number = solid_id (solid_number may be NULL). */
void mplbm_host_place_sphere(uint8_t* is_fluid, int64_t* solid_number, int64_t lx, int64_t ly, int64_t lz,
                             float x, float y, float z, float R, int64_t solid_id);

/* read_positions (src/multiphase_lbm.cu:1106-1126): x,y,z,r one per line,
 * each multiplied by res.  Returns the number of spheres placed or < 0. */
int64_t mplbm_host_read_positions(const char* path, float res, uint8_t* is_fluid, int64_t* solid_number,
                                  int64_t lx, int64_t ly, int64_t lz, int64_t first_id);

/* mplbm_host_read_positions with the spheres voxelised on the device in
 * one mplbm_place_spheres call (same mask and solid numbers). */
int64_t mplbm_host_read_positions_device(const char* path, float res, uint8_t* is_fluid, int64_t* solid_number,
                                         int64_t lx, int64_t ly, int64_t lz, int64_t first_id);

/* Density initialisation of resize_gpu (src/multiphase_lbm.cu:134-144):
 * srand(1), loops z, y, x, rho0*(1 + IniPerturbRate*(rand()/RAND_MAX - 0.5)). */
void mplbm_host_rho_init(float* rho, int64_t lx, int64_t ly, int64_t lz, float rho0, float ini_perturb_rate);

/* read_densities (src/multiphase_lbm.cu:112-132): Density block of a
 * legacy VTK written by mplbm_host_write_vtk.  Returns values read or < 0. */
int64_t mplbm_host_read_densities(const char* path, float* rho, int64_t lx, int64_t ly, int64_t lz);

/* read_sample (src/multiphase_lbm.cu:750-806): 0/1 voxel file with a
 * 3-number header, down-sampled by lx_file/lx.  Returns 0 or < 0. */
int mplbm_host_read_sample(const char* path, uint8_t* is_fluid, int64_t lx, int64_t ly, int64_t lz);

/* Seeded synthetic packing (not in the reference): spheres with radii
 * uniform in [rmin, rmax] and centres uniform in the box, placed until
 * n_max spheres or the voxel porosity target is reached.  nonoverlap = 1
 * rejects spheres overlapping earlier ones (random sequential addition,
 * periodic minimum-image distance).  Spheres are voxelised with
 * mplbm_host_place_sphere.  spheres_out receives x,y,z,r (capacity
 * spheres).  Returns the number of spheres placed or < 0. */
int64_t mplbm_host_generate_packing(uint8_t* is_fluid, int64_t lx, int64_t ly, int64_t lz, uint64_t seed,
                                    int64_t n_max, double rmin, double rmax, double porosity_target,
                                    int32_t nonoverlap, double* spheres_out, int64_t capacity);

/* Write spheres in the positions-file format read by read_positions. */
int mplbm_host_write_positions(const char* path, const double* spheres, int64_t n);

/* calc_cluster + write_stats extrema (src/multiphase_lbm.cu:225-562,
 * 1057-1083).  liq_label/gas_label (may be NULL) receive the cluster numbers
 * written to the VTK LiqClusters/GasClusters blocks. */
int mplbm_host_cluster_stats(const uint8_t* is_fluid, const float* rho, const float* pres,
                             const int64_t* solid_number, int64_t lx, int64_t ly, int64_t lz,
                             double rhol_spinodal, double rhog_spinodal, int64_t solid_count,
                             int64_t* liq_label, int64_t* gas_label, mplbm_cluster_stats* out);

/* write_vtk_ascii (src/multiphase_lbm.cu:948-1054), same format strings. */
int mplbm_host_write_vtk(const char* path, const uint8_t* is_fluid, const float* rho, const float* pres,
                         const int64_t* liq_label, const int64_t* gas_label, const int64_t* solid_number,
                         int64_t lx, int64_t ly, int64_t lz);

/* Binary counterpart of write_vtk for the vtrPeriod token (parsed by the
 * reference, :645-650, with no writer): VTK XML RectilinearGrid, raw
 * appended little-endian data with UInt64 block sizes, the same point arrays
 * (Zones, Density and Pressure as Float32 with 0 at solids, LiqClusters,
 * GasClusters, Solids as Int64) in the same x-fastest point order and the
 * same coordinates i/lx. */
int mplbm_host_write_vtr(const char* path, const uint8_t* is_fluid, const float* rho, const float* pres,
                         const int64_t* liq_label, const int64_t* gas_label, const int64_t* solid_number,
                         int64_t lx, int64_t ly, int64_t lz);

/* read_data (src/multiphase_lbm.cu:619-745): tokens, domain, density
 * initialisation and geometry (resize, read_sample, place_sphere,
 * generate_solids, read_positions), then erode_solid_touch_liquid when the
 * input has target_density (:739-740, :842-890).  Fills `in` (lx, ly, lz =
 * the domain); is_fluid and rho (each may be NULL, capacity cells) receive
 * the mask and the initial density [x][y][z]; eroded (may be NULL) the
 * number of solid nodes turned fluid.  Returns 0 or < 0. */
int mplbm_host_read_data(const char* input_path, mplbm_input* in, uint8_t* is_fluid, float* rho, int64_t capacity,
                         int64_t* eroded);

/* The whole program of src/main.cc:4-17 (read_data, init, run_gpu) on the
 * HIP library.  Returns 0 on success. */
int mplbm_host_run(const char* input_path);

#ifdef __cplusplus
}
#endif

#endif /* MPLBM_HOST_H_ */
