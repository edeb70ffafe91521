// libmplbm: MI355X (gfx950) D3Q19 Shan-Chen multiphase LBM step behind the
// C ABI of include/mplbm.h.
//
// Device layout (plane-major structure of arrays): population i of the cell
// (x, y, z) lives at lattice[(x*19 + i)*ly*lz + y*lz + z].  A wavefront
// reading 64 consecutive z of one direction issues one coalesced 256-byte
// access, and the 19 sub-planes of one x-plane lie within a 32-bit byte
// offset of a single wave-uniform base, which the marching kernel exploits.  Two
// lattices (pull scheme, A -> B, pointer swap on the host instead of the
// reference's do_swap kernel, include/iteration_kernel.cuh:44-48).  A 32-bit
// link word per cell holds bit 0 = "cell is solid" and bit i = "upstream
// node x - e_i is solid" (halfway bounce-back, iteration_kernel.cuh:112-115).
//
// Step kernels:
//  * k_fused: one launch per step.  A workgroup owns a (y,z) tile with a
//    one-cell ring and marches along x.  For plane x+1 it pulls all 19
//    populations of every tile+ring cell, forms rho and psi and writes psi to
//    a 4-slot LDS ring; after one barrier it evaluates the Shan-Chen force of
//    plane x from LDS, collides the populations of plane x it kept in
//    registers and stores them.  Each population is read from HBM once per
//    step (ring re-reads hit L2/MALL) and written once.
//  * k_density + k_collide: the two-pass schedule (psi of every node before
//    any force), the plain form of the reference's do_collision without its
//    psi race (iteration_device.cuh:62 vs :89-147).
//
// Arithmetic modes:
//  * ModeF32: float storage, the reference's float arithmetic including its
//    double-promoted sub-expressions (iteration_device.cuh:7-206).
//  * ModeF64: double storage and arithmetic.
//  * ModeH16 / ModeB16: 16-bit storage of g_i = (f_i/w_i - rho_ref)*scale,
//    all arithmetic in float.
// Built with -ffp-contract=off: no multiply-add is fused unless written.

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "d3q19.hpp"
#include "mplbm.h"

namespace mplbm {

// ----------------------------------------------------------------------------
// storage / arithmetic modes

// kRef: the reference's arithmetic (double-promoted sub-expressions, own
// exp, no fused multiply-add).  kDev: populations stored as 16-bit
// deviations.
struct ModeF32 {
  using S = float;
  using C = float;
  using M = float;  // macro-field type
  static constexpr bool kRef = true;
  static constexpr bool kDev = false;
  static constexpr int kStorage = MPLBM_FP32;
};
struct ModeF32Fast {  // FP32 storage, FP32 arithmetic with fused multiply-adds
  using S = float;
  using C = float;
  using M = float;
  static constexpr bool kRef = false;
  static constexpr bool kDev = false;
  static constexpr int kStorage = MPLBM_FP32;
};
struct ModeF64 {
  using S = double;
  using C = double;
  using M = double;
  static constexpr bool kRef = true;
  static constexpr bool kDev = false;
  static constexpr int kStorage = MPLBM_FP64;
};
struct ModeH16 {
  using S = __half;
  using C = float;
  using M = float;
  static constexpr bool kRef = false;
  static constexpr bool kDev = true;
  static constexpr int kStorage = MPLBM_FP16S;
};
struct ModeB16 {
  using S = __hip_bfloat16;
  using C = float;
  using M = float;
  static constexpr bool kRef = false;
  static constexpr bool kDev = true;
  static constexpr int kStorage = MPLBM_BF16S;
};

// Per-step physics, converted on the host to the arithmetic type C.
template <class C>
struct Phys {
  Eos<C> eos;
  C tau;            // (C)tau
  C inv_tau;        // (C)(1.0/(double)tau), iteration_device.cuh:165
  C neg_G;          // -G
  C rhos;           // solid density
  C rhol_sp;        // liquid spinodal threshold
  C inc;            // rho_increment * reverse
  C inject_ylim;    // inject_position * ly
  C g[3];           // body acceleration
  float rho_ref, scale, inv_scale;  // deviation storage
  int node_shift;                    // MPLBM_SHIFT_NODE: per-node shifts in Lattice::sin/sout
  float dec_mul[3], dec_add[3];      // per weight class: f = g * dec_mul + dec_add
  float enc_mul[3], enc_add;         // g = f * enc_mul + enc_add
  int inject_now;   // step % rho_inject_period == 0
  int inject_type;
  int has_body;
  int any_face;
  int face[6];      // face types
  C face_rho[6];
  int write_macro;
  int sr;            // stochastic rounding of the stored deviations
  uint32_t sr_seed;  // per step
};

// ----------------------------------------------------------------------------
// population encode / decode

template <class Mo>
__device__ __forceinline__ typename Mo::C decode(typename Mo::S s, int i, const Phys<typename Mo::C>& ph) {
  if constexpr (!Mo::kDev) {
    return s;
  } else {
    float g;
    if constexpr (Mo::kStorage == MPLBM_FP16S) g = __half2float(s);
    else g = __bfloat162float(s);
    // w_i (g / scale + rho_ref) folded into one multiply-add per class
    return __builtin_fmaf(g, ph.dec_mul[wclass(i)], ph.dec_add[wclass(i)]);
  }
}

template <class Mo>
__device__ __forceinline__ typename Mo::S encode(typename Mo::C v, int i, const Phys<typename Mo::C>& ph) {
  if constexpr (!Mo::kDev) {
    return v;
  } else {
    // (f / w_i - rho_ref) * scale folded into one multiply-add per class
    const float g = __builtin_fmaf(v, ph.enc_mul[wclass(i)], ph.enc_add);
    if constexpr (Mo::kStorage == MPLBM_FP16S) return __float2half_rn(g);
    else return __float2bfloat16(g);
  }
}

// Node-shifted form: w_i (g / scale + shift) and (f / w_i - shift) * scale,
// shift the 16-bit stored density of the node that wrote the population.
// kSolidShift marks a solid cell in the LDS shift region: a NaN bit pattern,
// so only a node whose stored density is already this NaN would be decoded
// with the pulling node's own shift instead of its own.
constexpr uint16_t kSolidShift = 0xFFFFu;
template <class Mo>
__device__ __forceinline__ float shift_value(typename Mo::S s) {
  if constexpr (Mo::kStorage == MPLBM_FP16S) return __half2float(s);
  else if constexpr (Mo::kStorage == MPLBM_BF16S) return __bfloat162float(s);
  else return (float)s;
}
template <class Mo>
__device__ __forceinline__ typename Mo::S shift_of(float rho) {
  if constexpr (Mo::kStorage == MPLBM_FP16S) return __float2half_rn(rho);
  else if constexpr (Mo::kStorage == MPLBM_BF16S) return __float2bfloat16(rho);
  else return (typename Mo::S)rho;
}
template <class Mo>
__device__ __forceinline__ typename Mo::C decode_at(typename Mo::S s, int i, const Phys<typename Mo::C>& ph,
                                                    float shift) {
  const float g = shift_value<Mo>(s);
  return __builtin_fmaf(g, ph.dec_mul[wclass(i)], weight<float>(i) * shift);
}
// enc_add = -shift * scale of the writing node
template <class Mo>
__device__ __forceinline__ typename Mo::S encode_at(typename Mo::C v, int i, const Phys<typename Mo::C>& ph,
                                                    float enc_add) {
  const float g = __builtin_fmaf((float)v, ph.enc_mul[wclass(i)], enc_add);
  return shift_of<Mo>(g);
}

// 32-bit integer mix (a bijection with good avalanche), for dither streams
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Stochastic rounding of the deviation: add u - 1/2 units in the last place
// of the 16-bit format (u uniform in [0, 1) from the top 24 bits of rnd),
// then round to nearest, so the stored value is v's deviation rounded up
// with probability equal to its distance from the value below.
template <class Mo>
__device__ __forceinline__ typename Mo::S encode_sr(typename Mo::C v, int i, const Phys<typename Mo::C>& ph,
                                                   uint32_t rnd, float enc_add) {
  const float g = __builtin_fmaf(v, ph.enc_mul[wclass(i)], enc_add);
  // 2^floor(log2|g|) from the exponent bits; the format's unit in the last
  // place is 2^-10 (FP16, at least 2^-24) or 2^-7 (BF16) of it
  const float p2 = __uint_as_float(__float_as_uint(g) & 0x7F800000u);
  const float u = (float)(rnd >> 8) * 0x1p-24f - 0.5f;
  if constexpr (Mo::kStorage == MPLBM_FP16S) {
    const float ulp = fmaxf(p2 * 0x1p-10f, 0x1p-24f);
    return __float2half_rn(__builtin_fmaf(u, ulp, g));
  } else {
    return __float2bfloat16(__builtin_fmaf(u, p2 * 0x1p-7f, g));
  }
}

// ----------------------------------------------------------------------------
// per-cell physics

// Density with injection (iteration_device.cuh:40-59): rho = sum_i f_i in
// ascending i, then the injection rule keyed on y.  kInject = false compiles
// the sum only (for steps on which no injection fires).
template <class C, bool kInject = true>
__device__ __forceinline__ C density(const C* f, int gy, const Phys<C>& ph) {
  C d = f[0];
#pragma unroll
  for (int i = 1; i < kQ; ++i) d = d + f[i];
  if (kInject && ph.inject_now && (C)gy < ph.inject_ylim) {
    if (ph.inject_type == 0) d = d + ph.inc;
    else if (ph.inject_type == 1) { if (d < ph.rhol_sp) d = d + ph.inc; }
    else if (ph.inject_type == 2) { if (d > ph.rhol_sp) d = d + ph.inc; }
  }
  return d;
}

// kCS = false: the Shan-Chen form only (carn_star known to be 0)
template <class Mo, bool kCS = true>
__device__ __forceinline__ typename Mo::C psi_of(const Phys<typename Mo::C>& ph, typename Mo::C rho) {
  if constexpr (Mo::kRef) {
    return eos_psi<kCS>(ph.eos, rho);
  } else {
    if (kCS && ph.eos.carn_star) return eos_psi(ph.eos, rho);
    return 1.0f - __expf(-rho);
  }
}

template <class Mo>
__device__ __forceinline__ typename Mo::C pressure_of(const Phys<typename Mo::C>& ph, typename Mo::C rho) {
  if constexpr (Mo::kRef) {
    return eos_pressure(ph.eos, rho);
  } else {
    if (ph.eos.carn_star) return eos_pressure(ph.eos, rho);
    const float psi = 1.0f - __expf(-rho);
    return ph.eos.cs2 * rho + ph.eos.half_cs2G * (psi * psi);
  }
}

// Shan-Chen force from the 18 neighbour psi values (psi_nb[i] = psi(x+e_i)),
// iteration_device.cuh:89-147: F_a = (-G psi(x)) * sum_i (w_i psi_nb[i]) e_ia
// with the non-zero terms added left to right in ascending i (the zero terms
// of the reference only add signed zeros).
template <class C>
__device__ __forceinline__ void sc_force(const C* psi_nb, C psi_c, const Phys<C>& ph, C* F) {
  C s[3];
  bool started[3] = {false, false, false};
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    const C wp = weight<C>(i) * psi_nb[i];
    const int e[3] = {ex(i), ey(i), ez(i)};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (e[a] == 0) continue;
      const C t = e[a] > 0 ? wp : -wp;
      s[a] = started[a] ? s[a] + t : t;
      started[a] = true;
    }
  }
  const C gpsi = ph.neg_G * psi_c;
  F[0] = gpsi * s[0];
  F[1] = gpsi * s[1];
  F[2] = gpsi * s[2];
}

// Velocity shift and BGK (iteration_device.cuh:162-206) in place on f.
// u is the shifted velocity (m + tau F) / rho that the reference builds.
template <class Mo>
__device__ __forceinline__ void collide(typename Mo::C* f, typename Mo::C rho, typename Mo::C* F,
                                        const Phys<typename Mo::C>& ph, typename Mo::C* u) {
  using C = typename Mo::C;
  if (ph.has_body) {
    F[0] = F[0] + rho * ph.g[0];
    F[1] = F[1] + rho * ph.g[1];
    F[2] = F[2] + rho * ph.g[2];
  }
  C m[3];
  bool started[3] = {false, false, false};
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    const int e[3] = {ex(i), ey(i), ez(i)};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (e[a] == 0) continue;
      const C t = e[a] > 0 ? f[i] : -f[i];
      m[a] = started[a] ? m[a] + t : t;
      started[a] = true;
    }
  }
  C ux = m[0] + ph.tau * F[0];
  C uy = m[1] + ph.tau * F[1];
  C uz = m[2] + ph.tau * F[2];
  if constexpr (Mo::kRef) {
    const C inv_rho = (C)(1.0 / (double)rho);
    ux = ux * inv_rho;
    uy = uy * inv_rho;
    uz = uz * inv_rho;
    const C usq = ux * ux + uy * uy + uz * uz;
    const double usq15 = 1.5 * (double)usq;
    const double itau = (double)ph.inv_tau;
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      C eu;
      {
        const int e[3] = {ex(i), ey(i), ez(i)};
        const C uu[3] = {ux, uy, uz};
        bool st = false;
        eu = (C)0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if (e[a] == 0) continue;
          const C t = e[a] > 0 ? uu[a] : -uu[a];
          eu = st ? eu + t : t;
          st = true;
        }
      }
      const C trho = weight<C>(i) * rho;
      const double deu = (double)eu;
      const double poly = 1.0 + sconst(3.0) * deu + sconst(4.5) * deu * deu - usq15;
      const double fi = (double)f[i];
      f[i] = (C)(fi + ((double)trho * poly - fi) * itau);
    }
  } else {
    // FP32 with fused multiply-adds: feq = w rho (c0 + eu (3 + 4.5 eu)),
    // c0 = 1 - 1.5 u^2; f += (feq - f) / tau
    const float inv_rho = __builtin_amdgcn_rcpf(rho);
    ux *= inv_rho;
    uy *= inv_rho;
    uz *= inv_rho;
    const float c0 = __builtin_fmaf(-1.5f, __builtin_fmaf(ux, ux, __builtin_fmaf(uy, uy, uz * uz)), 1.0f);
    const float rw[3] = {weight<float>(0) * rho, weight<float>(1) * rho, weight<float>(4) * rho};
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      float eu = 0.f;
      if (ex(i)) eu = ex(i) > 0 ? eu + ux : eu - ux;
      if (ey(i)) eu = ey(i) > 0 ? eu + uy : eu - uy;
      if (ez(i)) eu = ez(i) > 0 ? eu + uz : eu - uz;
      const float feq = rw[wclass(i)] * __builtin_fmaf(eu, __builtin_fmaf(4.5f, eu, 3.0f), c0);
      f[i] = __builtin_fmaf(ph.inv_tau, feq - f[i], f[i]);
    }
  }
  u[0] = ux;
  u[1] = uy;
  u[2] = uz;
}

// Anti-bounce-back at a PRESSURE face for direction i entering the domain:
// B_i = -A_opp(i) + 2 t_i rho_w (1 + 4.5 (e_i.u)^2 - 1.5 u^2) with
// u = sum(A e)/sum(A) of the face node (iteration_device.cuh:213-243,282-286).
template <class Mo>
__device__ __forceinline__ typename Mo::C abb_value(const typename Mo::C* a, int i, typename Mo::C rho_w) {
  using C = typename Mo::C;
  C ra = (C)0, mx = (C)0, my = (C)0, mz = (C)0;
#pragma unroll
  for (int k = 0; k < kQ; ++k) {
    ra = ra + a[k];
    mx = mx + a[k] * (C)ex(k);
    my = my + a[k] * (C)ey(k);
    mz = mz + a[k] * (C)ez(k);
  }
  const C inv = (C)(1.0 / (double)ra);
  const C ux = mx * inv, uy = my * inv, uz = mz * inv;
  const C usq = ux * ux + uy * uy + uz * uz;
  const C eu = (C)ex(i) * ux + (C)ey(i) * uy + (C)ez(i) * uz;
  const C added = (C)((double)((C)2 * weight<C>(i) * rho_w) *
                      (1.0 + 4.5 * (double)eu * (double)eu - 1.5 * (double)usq));
  return -a[opp(i)] + added;
}

// Which non-periodic face (0..5) direction i enters through at (x,y,z), or -1.
// Later faces win, as the reference applies x, y, z face rules in turn.
template <class C>
__device__ __forceinline__ int entering_face(const Phys<C>& ph, int i, int x, int y, int z, int lx, int ly,
                                             int lz) {
  int f = -1;
  if (ph.face[0] && x == 0 && ex(i) == 1) f = 0;
  if (ph.face[1] && x == lx - 1 && ex(i) == -1) f = 1;
  if (ph.face[2] && y == 0 && ey(i) == 1) f = 2;
  if (ph.face[3] && y == ly - 1 && ey(i) == -1) f = 3;
  if (ph.face[4] && z == 0 && ez(i) == 1) f = 4;
  if (ph.face[5] && z == lz - 1 && ez(i) == -1) f = 5;
  return f;
}

// The force is not evaluated on a node lying on the boundary planes of a
// non-periodic axis (iteration_device.cuh:79-87); F = 0 there.
template <class C>
__device__ __forceinline__ bool force_skipped(const Phys<C>& ph, int x, int y, int z, int lx, int ly, int lz) {
  if ((ph.face[0] | ph.face[1]) && (x == 0 || x == lx - 1)) return true;
  if ((ph.face[2] | ph.face[3]) && (y == 0 || y == ly - 1)) return true;
  if ((ph.face[4] | ph.face[5]) && (z == 0 || z == lz - 1)) return true;
  return false;
}

// ----------------------------------------------------------------------------
// kernel arguments

template <class Mo>
struct Lattice {
  const typename Mo::S* fin;
  typename Mo::S* fout;
  const uint32_t* link;
  typename Mo::M* rho;   // macro fields (written when ph.write_macro)
  typename Mo::M* pres;
  typename Mo::M* u;     // 3 planes of ncell
  typename Mo::C* psi;   // two-pass psi field
  // node shifts (MPLBM_SHIFT_NODE, compact layout): one per fluid node (see
  // sidx), read from sin (written with fin), written to sout (with fout);
  // null in the global-shift form
  const typename Mo::S* sin;
  typename Mo::S* sout;
  // element index of population i of cell c (plane-major layout)
  __device__ __forceinline__ uint64_t pix(int i, uint32_t c) const {
    const uint32_t x = c / plane;
    return ((uint64_t)x * kQ + i) * ps + (c - x * plane);
  }
  int lx, ly, lz;
  uint32_t plane;        // ly*lz (cells of an x-plane; field arrays use it)
  uint32_t ps;           // population sub-plane stride in elements (>= plane)
  uint32_t ncell;
  // compact layout (fluid nodes only): see cpix
  const uint2* rw;       // [x][y][w][group]: see rank_word
  uint32_t W;            // words per (x, y) row
  int compact;
};

// Compact layout.  Only fluid nodes are stored.  The fluid nodes of an
// x-plane are numbered (their plane rank r) 32-wide z block by block, and
// within a block by blocks of rows; a block of rows numbers first the
// columns 0-1 of its rows (row by row), then columns 2-29 (row by row, z
// within a row), then columns 30-31, so the cells next to the neighbouring
// z blocks form short contiguous ranges.  Each (x, y, z block) has three
// rank words, one per column group: {fluid bits of the row's columns in the
// group, rank of the group's first node in the row}.  Population i of the
// node of rank r in plane x lives at element ((x+1)*19 + i)*ps + r, ps >= the
// largest plane count.  Memory planes 0 and lx+1 are copies of planes lx-1
// and 0 (kept current by the step kernel), so the three source planes of a
// pull are always adjacent in memory.
constexpr uint32_t kRankGroups = 3;
__host__ __device__ __forceinline__ uint32_t rank_group(uint32_t c) {  // c = z mod 32
  return c < 2 ? 1u : (c >= 30 ? 2u : 0u);
}
// index of the rank word of cell (y, z) within its plane's words
__host__ __device__ __forceinline__ uint32_t rank_word(uint32_t W, uint32_t y, uint32_t z) {
  return (y * W + (z >> 5)) * kRankGroups + rank_group(z & 31u);
}
__device__ __forceinline__ uint32_t plane_rank(const uint2* rw, uint32_t W, int ly, int x, int y, int z) {
  const uint2 w = rw[(uint64_t)x * ly * W * kRankGroups + rank_word(W, y, z)];
  return w.y + __popc(w.x & ((1u << (z & 31)) - 1u));
}
__device__ __forceinline__ bool rw_fluid(const uint2* rw, uint32_t W, int ly, int x, int y, int z) {
  return (rw[(uint64_t)x * ly * W * kRankGroups + rank_word(W, y, z)].x >> (z & 31)) & 1u;
}
template <class Mo>
__device__ __forceinline__ uint64_t cpix(const Lattice<Mo>& L, int i, int mplane, uint32_t rank) {
  return ((uint64_t)mplane * kQ + i) * L.ps + rank;
}
// element of the node shift of the fluid node of rank `rank` in plane x
// (the populations' plane-rank numbering, no ghost planes)
template <class Mo>
__device__ __forceinline__ uint64_t sidx(const Lattice<Mo>& L, int x, uint32_t rank) {
  return (uint64_t)x * L.ps + rank;
}

struct Tiling {
  int TY, TZ;            // threads per tile in y and z (ring included unless wrapped)
  int tiles_y, tiles_z;
  int seg, nseg;         // planes per block, segments along x
  int wrap_y, wrap_z;    // 1: the whole axis lies in one tile, neighbours wrap in LDS
  int nblocks;           // real blocks (grid may be padded to a multiple of 8)
  int xcd_remap;
  int pull_bb;           // 1: bounce-back read on the pull side, no ghost stores
};

__device__ __forceinline__ int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

template <class C>
__device__ __forceinline__ bool on_face(const Phys<C>& ph, int x, int y, int z, int lx, int ly, int lz) {
  return (ph.face[0] && x == 0) || (ph.face[1] && x == lx - 1) || (ph.face[2] && y == 0) ||
         (ph.face[3] && y == ly - 1) || (ph.face[4] && z == 0) || (ph.face[5] && z == lz - 1);
}

// Raw buffer access: a wave-uniform buffer resource (one x-plane of
// populations), a per-lane byte offset (voffset) and a wave-uniform byte
// offset (soffset, the population's sub-plane).  No per-lane address
// arithmetic.  Word 3 0x00020000 is the gfx9-family resource format.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc make_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)0xFFFFFFFF, 0x00020000);
}
template <class S>
__device__ __forceinline__ S bload(Rsrc r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(S) == 4) {
    return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
  } else if constexpr (sizeof(S) == 8) {
    return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
  } else {
    return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0));
  }
}
// Cache policy of a population store.  The compact kernel's 4- and 8-byte
// stores carry the non-temporal hint (bit 1 on gfx950): the written lattice
// is read only in the next step, and keeping these lines out of L2's working
// set leaves room for the ring lines that neighbouring tiles re-read (FP32
// 512^3: reads 10.47 -> 9.39 GB, writes 4.76 -> 5.17 GB per step, 1-4%
// faster depending on the device).  16-bit stores keep the default policy
// (their partial lines cost more written bytes than the reads save: FP16 5%
// slower with the hint), and so does the full-lattice kernel (3% slower).
constexpr int kStoreDefault = 0;
constexpr int kStoreNT = 2;
template <class S>
constexpr int compact_store_policy() {
  return sizeof(S) >= 4 ? kStoreNT : kStoreDefault;
}
template <class S, int kPolicy = kStoreDefault>
__device__ __forceinline__ void bstore(S v, Rsrc r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(S) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)voff, (int)soff, kPolicy);
  } else if constexpr (sizeof(S) == 8) {
    using V2 = unsigned int __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, (int)voff, (int)soff, kPolicy);
  } else {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, (int)voff, (int)soff, kPolicy);
  }
}

// Global loads/stores at a 32-bit byte offset from a base pointer.
template <class S>
__device__ __forceinline__ S ld(const S* base, uint32_t off_bytes) {
  return *reinterpret_cast<const S*>(reinterpret_cast<const char*>(base) + off_bytes);
}
template <class S>
__device__ __forceinline__ void st(S* base, uint32_t off_bytes, S v) {
  *reinterpret_cast<S*>(reinterpret_cast<char*>(base) + off_bytes) = v;
}

// Face rule after the pull on a face cell: each direction entering through a
// PRESSURE face takes the anti-bounce-back value built from the cell's own
// populations of the previous step.
template <class Mo>
__device__ __forceinline__ void apply_faces(const Lattice<Mo>& L, const Phys<typename Mo::C>& ph, uint32_t self,
                                            int x, int y, int z, typename Mo::C* f) {
  using C = typename Mo::C;
  C a[kQ];
#pragma unroll
  for (int k = 0; k < kQ; ++k) a[k] = decode<Mo>(L.fin[L.pix(k, self)], k, ph);
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    const int fc = entering_face(ph, i, x, y, z, L.lx, L.ly, L.lz);
    if (fc >= 0) f[i] = abb_value<Mo>(a, i, ph.face_rho[fc]);
  }
}

// Halfway bounce-back is resolved on the write side.  After a fluid cell x
// collides it stores f*_i(x) in its own slots and, for every direction i
// whose upstream node x - e_i is solid, also stores f*_opp(i)(x) in slot i
// of that solid node.  The next step's pull of direction i at x reads slot i
// of x - e_i unconditionally and so receives exactly the bounce-back value
// B_i(x) = A_opp(i)(x) of iteration_kernel.cuh:112-115.  A solid slot (i, s)
// is written only by the unique fluid node s + e_i, and solid nodes are never
// updated otherwise, so the extra stores do not race.
template <class P, class Q>
struct PtrOff {
  P first;
  Q second;
};

template <class Mo, class OwnFn, class UpFn>
__device__ __forceinline__ void store_cell(uint32_t lk, OwnFn own, UpFn up, const typename Mo::C* f,
                                           const Phys<typename Mo::C>& ph) {
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const auto o = own(i);
    st(o.first, o.second, encode<Mo>(f[i], i, ph));
  }
  if (lk & ~1u) {
#pragma unroll
    for (int i = 1; i < kQ; ++i)
      if ((lk >> i) & 1u) {
        const auto o = up(i);
        st(o.first, o.second, encode<Mo>(f[opp(i)], i, ph));
      }
  }
}


// ----------------------------------------------------------------------------
// fused marching stream-collide kernel

// One workgroup's share of a step: tile (tyi, tzi) of the (y,z) plane,
// marched over the x-planes of segment `seg`.  Every thread of the workgroup
// must call it (it contains barriers); `lds` holds 4 * TY * TZ values of C.
template <class Mo, bool kFaces>
__device__ __forceinline__ void march_tile(const Lattice<Mo>& L, const Phys<typename Mo::C>& ph, const Tiling& T,
                                           int seg, int tyi, int tzi, typename Mo::C* lds) {
  using C = typename Mo::C;
  using S = typename Mo::S;
  const int TY = T.TY, TZ = T.TZ, TT = TY * TZ;
  const int tid = threadIdx.x;
  const bool thr = tid < TT;
  const int ty = thr ? tid / TZ : 0;
  const int tz = thr ? tid - ty * TZ : 0;

  int gy, gz, rowm, rowp, colm, colp;
  bool iny, inz;
  if (T.wrap_y) {
    gy = ty;
    iny = thr;
    rowm = ty == 0 ? L.ly - 1 : ty - 1;
    rowp = ty + 1 == L.ly ? 0 : ty + 1;
  } else {
    const int yy = tyi * (TY - 2) - 1 + ty;
    gy = wrapi(yy, L.ly);
    iny = thr && ty >= 1 && ty <= TY - 2 && yy < L.ly;
    rowm = ty - 1;
    rowp = ty + 1;
  }
  if (T.wrap_z) {
    gz = tz;
    inz = thr;
    colm = tz == 0 ? L.lz - 1 : tz - 1;
    colp = tz + 1 == L.lz ? 0 : tz + 1;
  } else {
    const int zz = tzi * (TZ - 2) - 1 + tz;
    gz = wrapi(zz, L.lz);
    inz = thr && tz >= 1 && tz <= TZ - 2 && zz < L.lz;
    colm = tz - 1;
    colp = tz + 1;
  }
  const bool interior = iny && inz;

  // byte offsets within an x-plane of the 9 in-plane neighbours, [dy+1][dz+1]
  uint32_t yzb[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const uint32_t row = (uint32_t)wrapi(gy + a - 1, L.ly) * (uint32_t)L.lz;
#pragma unroll
    for (int c = 0; c < 3; ++c) yzb[a][c] = (row + (uint32_t)wrapi(gz + c - 1, L.lz)) * (uint32_t)sizeof(S);
  }
  const uint32_t self_yz = yzb[1][1] / (uint32_t)sizeof(S);
  const int lrow[3] = {rowm, ty, rowp};
  const int lcol[3] = {colm, tz, colp};

  const C psi_solid = psi_of<Mo>(ph, ph.rhos);
  // block-uniform; readfirstlane keeps the march counter and the plane bases
  // in scalar registers
  const int x0 = __builtin_amdgcn_readfirstlane(seg * T.seg);
  const int x1 = __builtin_amdgcn_readfirstlane(min(x0 + T.seg, L.lx));

  // An x-plane of populations is one base pointer (wave-uniform) plus a
  // 32-bit byte offset i*psb + (y,z) offset: three bases per plane step.
  const uint32_t psb = L.ps * (uint32_t)sizeof(S);
  auto in_plane = [&](int xs) { return L.fin + (uint64_t)xs * kQ * L.ps; };
  auto out_plane = [&](int xs) { return L.fout + (uint64_t)xs * kQ * L.ps; };

  auto wrapx = [&](int v) { return v >= L.lx ? v - L.lx : (v < 0 ? v + L.lx : v); };
  auto link_of = [&](int xp) -> uint32_t { return thr ? L.link[(uint64_t)xp * L.plane + self_yz] : 1u; };

  // Issue the 19 pulls of plane xp (fluid cells only) into raw storage; the
  // values are consumed one pipeline stage later.
  auto issue_plane = [&](int xp, uint32_t lk, S* raw) {
    if (!(lk & 1u)) {
      const Rsrc bm = make_rsrc(in_plane(wrapx(xp - 1)));
      const Rsrc b0 = make_rsrc(in_plane(xp));
      const Rsrc bq = make_rsrc(in_plane(wrapx(xp + 1)));
#pragma unroll
      for (int i = 0; i < kQ; ++i) {
        const Rsrc base = ex(i) == -1 ? bq : (ex(i) == 1 ? bm : b0);
        raw[i] = bload<S>(base, yzb[1 - ey(i)][1 - ez(i)], (uint32_t)i * psb);
      }
      if (T.pull_bb && (lk & ~1u)) {
        // halfway bounce-back read directly: B_i(x) = A_opp(i)(x)
#pragma unroll
        for (int i = 1; i < kQ; ++i)
          if ((lk >> i) & 1u) raw[i] = bload<S>(b0, yzb[1][1], (uint32_t)opp(i) * psb);
      }
    }
  };
  // Decode the pulled populations, apply face rules, return rho.
  auto finish_plane = [&](int xp, uint32_t lk, const S* raw, C* f) -> C {
    C r = (C)0;
    if (!(lk & 1u)) {
#pragma unroll
      for (int i = 0; i < kQ; ++i) f[i] = decode<Mo>(raw[i], i, ph);
      if constexpr (kFaces) {
        if (on_face(ph, xp, gy, gz, L.lx, L.ly, L.lz))
          apply_faces<Mo>(L, ph, (uint32_t)((uint64_t)xp * L.plane) + self_yz, xp, gy, gz, f);
      }
      r = density(f, gy, ph);
    }
    return r;
  };
  const int myl = ty * TZ + tz;
  auto put_psi = [&](int slot, uint32_t lk, C rho) {
    if (thr) lds[slot * TT + myl] = (lk & 1u) ? psi_solid : psi_of<Mo>(ph, rho);
  };

  // Three-stage pipeline over the planes of the segment: while plane x is
  // collided and plane x+1's density/psi is formed, the loads of plane x+2
  // and the link word of plane x+3 are in flight.
  const int n = x1 - x0;
  S rawQ[kQ], rawN[kQ];
  C P[kQ], Qd[kQ];
  C rhoP;
  uint32_t lkP, lkQ, lkN;
  {
    const int xw = wrapx(x0 - 1);
    const uint32_t lkW = link_of(xw);
    issue_plane(xw, lkW, rawN);
    lkP = link_of(x0);
    issue_plane(x0, lkP, rawQ);
    const C rw = finish_plane(xw, lkW, rawN, Qd);
    put_psi(0, lkW, rw);
    rhoP = finish_plane(x0, lkP, rawQ, P);
    put_psi(1, lkP, rhoP);
    lkQ = link_of(wrapx(x0 + 1));
    issue_plane(wrapx(x0 + 1), lkQ, rawQ);
    lkN = n > 1 ? link_of(wrapx(x0 + 2)) : 1u;
  }

  for (int j = 0; j < n; ++j) {
    const int x = __builtin_amdgcn_readfirstlane(x0 + j);
    const int xn = wrapx(x + 1);
    uint32_t lkNext = 1u;
    if (j + 1 < n) {  // uniform: plane x+2 feeds the next iteration
      issue_plane(wrapx(x + 2), lkN, rawN);
      if (j + 2 < n) lkNext = link_of(wrapx(x + 3));
    }
    const C rhoQ = finish_plane(xn, lkQ, rawQ, Qd);
    put_psi((j + 2) & 3, lkQ, rhoQ);
    __syncthreads();

    if (interior && !(lkP & 1u)) {
      const uint64_t p0 = (uint64_t)x * L.plane;
      const int xm = wrapx(x - 1);
      C F[3];
      bool skip = false;
      if constexpr (kFaces) skip = force_skipped(ph, x, gy, gz, L.lx, L.ly, L.lz);
      if (skip) {
        F[0] = F[1] = F[2] = (C)0;
      } else {
        C pn[kQ];
        pn[0] = (C)0;
#pragma unroll
        for (int i = 1; i < kQ; ++i) {
          const int slot = (j + 1 + ex(i)) & 3;
          pn[i] = lds[slot * TT + lrow[1 + ey(i)] * TZ + lcol[1 + ez(i)]];
        }
        sc_force(pn, lds[((j + 1) & 3) * TT + myl], ph, F);
      }
      C u[3];
      collide<Mo>(P, rhoP, F, ph, u);
      const Rsrc om = make_rsrc(out_plane(xm));
      const Rsrc o0 = make_rsrc(out_plane(x));
      const Rsrc oq = make_rsrc(out_plane(xn));
#pragma unroll
      for (int i = 0; i < kQ; ++i) bstore<S>(encode<Mo>(P[i], i, ph), o0, yzb[1][1], (uint32_t)i * psb);
      if (!T.pull_bb && (lkP & ~1u)) {
        // write-side bounce-back into the solid upstream nodes (see store_cell)
#pragma unroll
        for (int i = 1; i < kQ; ++i)
          if ((lkP >> i) & 1u) {
            const Rsrc base = ex(i) == -1 ? oq : (ex(i) == 1 ? om : o0);
            bstore<S>(encode<Mo>(P[opp(i)], i, ph), base, yzb[1 - ey(i)][1 - ez(i)], (uint32_t)i * psb);
          }
      }
      if (ph.write_macro) {
        const uint64_t self = p0 + self_yz;
        L.rho[self] = (typename Mo::M)rhoP;
        L.pres[self] = (typename Mo::M)pressure_of<Mo>(ph, rhoP);
        L.u[self] = (typename Mo::M)u[0];
        L.u[(uint64_t)L.ncell + self] = (typename Mo::M)u[1];
        L.u[2ull * L.ncell + self] = (typename Mo::M)u[2];
      }
    }
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      P[i] = Qd[i];
      rawQ[i] = rawN[i];
    }
    rhoP = rhoQ;
    lkP = lkQ;
    lkQ = lkN;
    lkN = lkNext;
  }
}

template <class Mo, bool kFaces, int kBlock>
__global__ __launch_bounds__(kBlock) void k_fused(Lattice<Mo> L, Phys<typename Mo::C> ph, Tiling T) {
  extern __shared__ unsigned char smem_raw[];
  int b = blockIdx.x;
  if (T.xcd_remap) {
    // consecutive logical tiles share an XCD (and its L2) so the ring
    // re-reads of neighbouring tiles hit cache
    const int per = gridDim.x >> 3;
    b = (b & 7) * per + (b >> 3);
  }
  if (b >= T.nblocks) return;  // whole padding block leaves before any barrier
  const int ntile = T.tiles_y * T.tiles_z;
  const int seg = b / ntile;
  const int t = b - seg * ntile;
  const int tyi = t / T.tiles_z;
  march_tile<Mo, kFaces>(L, ph, T, seg, tyi, t - tyi * T.tiles_z, reinterpret_cast<typename Mo::C*>(smem_raw));
}

// ----------------------------------------------------------------------------
// two-pass kernels (one thread per cell)

template <class Mo>
__device__ __forceinline__ void cell_coords(const Lattice<Mo>& L, uint32_t c, int& x, int& y, int& z) {
  x = (int)(c / L.plane);
  const uint32_t r = c - (uint32_t)x * L.plane;
  y = (int)(r / (uint32_t)L.lz);
  z = (int)(r - (uint32_t)y * (uint32_t)L.lz);
}

template <class Mo>
__device__ __forceinline__ uint32_t cell_at(const Lattice<Mo>& L, int x, int y, int z) {
  return ((uint32_t)wrapi(x, L.lx) * (uint32_t)L.ly + (uint32_t)wrapi(y, L.ly)) * (uint32_t)L.lz +
         (uint32_t)wrapi(z, L.lz);
}

template <class Mo>
__device__ __forceinline__ typename Mo::C pull_cell(const Lattice<Mo>& L, const Phys<typename Mo::C>& ph,
                                                   uint32_t c, int x, int y, int z, typename Mo::C* f) {
#pragma unroll
  for (int i = 0; i < kQ; ++i)
    f[i] = decode<Mo>(L.fin[L.pix(i, cell_at(L, x - ex(i), y - ey(i), z - ez(i)))], i, ph);
  if (ph.any_face && on_face(ph, x, y, z, L.lx, L.ly, L.lz)) apply_faces<Mo>(L, ph, c, x, y, z, f);
  return density(f, y, ph);
}

template <class Mo>
__global__ __launch_bounds__(256) void k_density(Lattice<Mo> L, Phys<typename Mo::C> ph) {
  using C = typename Mo::C;
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= L.ncell) return;
  const uint32_t lk = L.link[c];
  if (lk & 1u) return;
  int x, y, z;
  cell_coords(L, c, x, y, z);
  C f[kQ];
  const C rho = pull_cell(L, ph, c, x, y, z, f);
  L.psi[c] = psi_of<Mo>(ph, rho);
  if (ph.write_macro) {
    L.rho[c] = (typename Mo::M)rho;
    L.pres[c] = (typename Mo::M)pressure_of<Mo>(ph, rho);
  }
}

template <class Mo>
__global__ __launch_bounds__(256) void k_collide(Lattice<Mo> L, Phys<typename Mo::C> ph) {
  using C = typename Mo::C;
  using S = typename Mo::S;
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= L.ncell) return;
  const uint32_t lk = L.link[c];
  if (lk & 1u) return;
  int x, y, z;
  cell_coords(L, c, x, y, z);
  C f[kQ];
  const C rho = pull_cell(L, ph, c, x, y, z, f);
  C F[3];
  if (ph.any_face && force_skipped(ph, x, y, z, L.lx, L.ly, L.lz)) {
    F[0] = F[1] = F[2] = (C)0;
  } else {
    C pn[kQ];
    pn[0] = (C)0;
#pragma unroll
    for (int i = 1; i < kQ; ++i) pn[i] = L.psi[cell_at(L, x + ex(i), y + ey(i), z + ez(i))];
    sc_force(pn, L.psi[c], ph, F);
  }
  C u[3];
  collide<Mo>(f, rho, F, ph, u);
  store_cell<Mo>(lk, [&](int i) { return PtrOff<S*, uint32_t>{L.fout + L.pix(i, c), 0u}; },
                 [&](int i) {
                   return PtrOff<S*, uint32_t>{L.fout + L.pix(i, cell_at(L, x - ex(i), y - ey(i), z - ez(i))), 0u};
                 },
                 f, ph);
  if (ph.write_macro) {
    L.u[c] = (typename Mo::M)u[0];
    L.u[(uint64_t)L.ncell + c] = (typename Mo::M)u[1];
    L.u[2ull * L.ncell + c] = (typename Mo::M)u[2];
  }
}

// Fill the bounce-back slots of solid nodes from the fluid populations of a
// lattice (after initialisation and after a population write), as the
// collision kernels do after every step.
template <class Mo>
__global__ void k_ghost(Lattice<Mo> L) {
  using S = typename Mo::S;
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= L.ncell) return;
  const uint32_t lk = L.link[c];
  if ((lk & 1u) || !(lk & ~1u)) return;
  int x, y, z;
  cell_coords(L, c, x, y, z);
  S* lat = L.fout;
#pragma unroll
  for (int i = 1; i < kQ; ++i)
    if ((lk >> i) & 1u)
      lat[L.pix(i, cell_at(L, x - ex(i), y - ey(i), z - ez(i)))] = lat[L.pix(opp(i), c)];
}

// ----------------------------------------------------------------------------
// compact marching kernel (fluid nodes only)

constexpr int kCompactSlots = 6;  // LDS slots per region (planes in flight)

struct CTiling {
  int TYi, TZi;          // interior nodes of a tile in y and z
  int PZ, RZ;            // row lengths of the psi region (TZi+2) and the rank region (TZi+4)
  int PYZ, RYZ;          // sizes of the psi region (TYi+2)(TZi+2) and the rank region (TYi+4)(TZi+4)
  uint32_t magic_pz, magic_rz;  // ceil(2^32/PZ), ceil(2^32/RZ): q/PZ = umulhi(q, magic) for q < 2^16
  int tiles_y, tiles_z;
  int seg, nseg;         // planes per workgroup, segments along x
  int nblocks;
  int xcd_remap;
  int threads;           // workgroup size (256, 512 or 1024)
  int seg0, seg_stride;  // a launch covers segments seg0 + k*seg_stride, k < nblocks / tiles
  int xoff, xlim;        // segment sg spans planes xoff + sg*seg .. min(+seg, xlim) (xlim 0: lx)
};

struct CLists {
  const uint16_t* pos;   // psi-region positions of the nodes a tile pulls in a plane, interior first
  const uint2* tab;      // [tile][x]: {first entry, n_interior | n_pull << 16}
};

// One workgroup marches a (y, z) tile of TYi x TZi nodes along the planes of
// an x segment.  Three regions of each plane are staged in LDS:
//  * rank region (tile + 2-node ring): byte offset of each fluid node within
//    its population sub-plane (plane rank * element size), ~0 for solid;
//  * psi region (tile + 1-node ring): psi of every node, solid ones included.
// Per plane x the workgroup (1) builds the rank region of plane x+3 and the
// solid part of its psi region, (2) pulls every fluid node of the psi region
// of plane x+1 (work list: the fluid nodes of that plane, not the thread
// grid) and writes its psi, keeping the populations of interior nodes in
// registers, (3) barrier, (4) evaluates the force of the fluid interior
// nodes of plane x from the psi ring, collides them and stores.  Threads take
// list entries tid, tid + kThreads, ..., so the work follows the fluid count
// of each plane; the workgroup has at least as many threads as the tile has
// interior nodes.  Six LDS slots per region (plane x' in slot (x' - x0 + 2) mod 6) keep every read at
// least one barrier away from the next write of its slot: during the pulls
// of plane x+1 the rank regions of planes x-2 .. x+3 and the psi regions of
// planes x-2 .. x+3 can be in use.
// kPlain: a step with no injection, no macro-field write and the Shan-Chen
// EOS, compiled without those paths (fewer registers: FP32 109 -> 101
// VGPRs, 16-bit 80 -> 72; 2-3% faster).
// kPre: rank words of the rank region this thread loads ahead (below).
// kTZ: the tile's interior width in z when fixed at compile time (0: T.TZi),
// so the row lengths of the rank and psi regions are constants and the
// in-plane LDS reads of a pull or a force pair up into ds_read2 instructions.
template <class Mo, bool kFaces, int kThreads, bool kSR = false, bool kNS = false, bool kPlain = false, int kPre = 2,
          int kTZ = 0>
__device__ __forceinline__ void compact_march(const Lattice<Mo>& L, const Phys<typename Mo::C>& ph, const CTiling& T,
                                              const CLists& W, int sg, int tile, unsigned char* smem_raw) {
  using C = typename Mo::C;
  using S = typename Mo::S;
  C* psl = reinterpret_cast<C*>(smem_raw);                                            // kSlots x PYZ
  uint32_t* rm = reinterpret_cast<uint32_t*>(smem_raw + kCompactSlots * T.PYZ * sizeof(C));  // kSlots x RYZ
  int* ycrd = reinterpret_cast<int*>(rm + kCompactSlots * T.RYZ);                    // TYi + 4 rows
  int* zcrd = ycrd + (T.TYi + 4);                                            // TZi + 4 columns
  uint32_t* qtab = reinterpret_cast<uint32_t*>(zcrd + (T.TZi + 4));          // RYZ: rank word << 5 | bit
  int* qpsi = reinterpret_cast<int*>(qtab + T.RYZ);                          // RYZ: psi index or -1
  int* pent = qpsi + T.RYZ;                                                  // kThreads: next list entry
  typename Mo::S* shr = reinterpret_cast<typename Mo::S*>(pent + kThreads);  // kSlots x RYZ node shifts (kNS)

  const int RZ = kTZ > 0 ? kTZ + 4 : T.RZ;  // row length of the rank region
  const int PZ = kTZ > 0 ? kTZ + 2 : T.PZ;  // row length of the psi region
  const int tyi = tile / T.tiles_z;
  const int tzi = tile - tyi * T.tiles_z;
  const int y0 = tyi * T.TYi, z0 = tzi * T.TZi;
  const int x0 = __builtin_amdgcn_readfirstlane(T.xoff + sg * T.seg);
  const int x1 = __builtin_amdgcn_readfirstlane(min(x0 + T.seg, T.xlim > 0 ? T.xlim : L.lx));
  const int n = x1 - x0;
  const int tid = threadIdx.x;

  constexpr uint32_t sz = sizeof(S);
  const uint32_t psb = L.ps * sz;  // bytes of a population sub-plane
  const C psi_solid = psi_of<Mo>(ph, ph.rhos);

  // global coordinates of the rank region's rows and columns
  for (int q = tid; q < T.TYi + 4; q += kThreads) ycrd[q] = ((y0 - 2 + q) % L.ly + L.ly) % L.ly;
  for (int q = tid; q < T.TZi + 4; q += kThreads) zcrd[q] = ((z0 - 2 + q) % L.lz + L.lz) % L.lz;
  __syncthreads();
  // per rank-region position: its rank word within a plane's words, the bit,
  // and its index in the psi region (the same for every plane)
  for (int q = tid; q < T.RYZ; q += kThreads) {
    const int r = (int)__umulhi((uint32_t)q, T.magic_rz);
    const int c = q - r * RZ;
    const uint32_t z = (uint32_t)zcrd[c];
    qtab[q] = (rank_word(L.W, (uint32_t)ycrd[r], z) << 5) | (z & 31u);
    qpsi[q] = (r >= 1 && r <= T.TYi + 2 && c >= 1 && c <= T.TZi + 2) ? (r - 1) * PZ + (c - 1) : -1;
  }
  __syncthreads();

  // planes x0-2 .. x1+1 are addressed: one wrap suffices for lx >= 2, and
  // every plane is plane 0 when lx = 1
  auto wrapx = [&](int v) {
    if (L.lx == 1) return 0;
    return v >= L.lx ? v - L.lx : (v < 0 ? v + L.lx : v);
  };
  auto slot = [&](int xp) { return (xp - x0 + 2) % kCompactSlots; };

  // (1) rank region and solid psi of plane xp.  expand_issue loads the
  // rank words of the first kPre positions of this thread (so their latency
  // overlaps the pulls that follow), expand_finish writes the region.
  static_assert(kPre >= 1 && kPre <= 2, "one or two rank words ahead");
  auto rw_word = [&](const uint2* rwp, int q) { return rwp[qtab[q] >> 5]; };
  // the node shift of a fluid cell of the rank region goes to LDS with its
  // rank, kSolidShift for a solid cell, so a decode tells a bounce-back from
  // the shift alone (kNS)
  auto put_rank = [&](int xp, int q, uint2 w) {
    const uint32_t bit = qtab[q] & 31u;
    const bool fl = (w.x >> bit) & 1u;
    const uint32_t rank = w.y + __popc(w.x & ((1u << bit) - 1u));
    rm[slot(xp) * T.RYZ + q] = fl ? rank * sz : ~0u;
    if constexpr (kNS) {
      shr[slot(xp) * T.RYZ + q] = fl ? L.sin[sidx(L, wrapx(xp), rank)] : __builtin_bit_cast(S, kSolidShift);
    }
    if (!fl) {
      const int pi = qpsi[q];
      if (pi >= 0) psl[slot(xp) * T.PYZ + pi] = psi_solid;
    }
  };
  // the thread's rank-region positions are recomputed from an opaque copy
  // of its index at each use instead of being kept as LDS addresses across
  // the march (a register the plain FP32 kernel would otherwise spill)
  auto thread_index = [&]() {
    int t = tid;
    asm volatile("" : "+v"(t));
    return t;
  };
  auto expand_issue = [&](int xp, uint2* w) {
    const uint2* rwp = L.rw + (uint64_t)wrapx(xp) * L.ly * L.W * kRankGroups;
    const int t = thread_index();
#pragma unroll
    for (int m = 0; m < kPre; ++m) {
      const int q = t + m * kThreads;
      if (q < T.RYZ) w[m] = rw_word(rwp, q);
    }
  };
  auto expand_finish = [&](int xp, const uint2* w) {
    const int t = thread_index();
#pragma unroll
    for (int m = 0; m < kPre; ++m) {
      const int q = t + m * kThreads;
      if (q < T.RYZ) put_rank(xp, q, w[m]);
    }
    const uint2* rwp = L.rw + (uint64_t)wrapx(xp) * L.ly * L.W * kRankGroups;
    for (int q = thread_index() + kPre * kThreads; q < T.RYZ; q += kThreads) put_rank(xp, q, rw_word(rwp, q));
  };
  auto expand = [&](int xp) {
    uint2 w[kPre];
    expand_issue(xp, w);
    expand_finish(xp, w);
  };

  // Pull of the fluid node at psi position p of plane xp: populations f,
  // face rules, density.  Sources in planes xp-1, xp, xp+1 lie in memory
  // planes xw, xw+1, xw+2 (xw = xp wrapped), one buffer resource.  A solid
  // source gives the halfway bounce-back B_i = A_opp(i) of the node itself
  // (iteration_kernel.cuh:112-115).
  // face rules (anti-bounce-back of the directions entering through a
  // pressure face) from the node's own populations (rank offset self)
  auto apply_faces = [&](int xp, C* f, int gy, int gz, uint32_t self, float osh) {
    if constexpr (kFaces) {
      const int xw = wrapx(xp);
      if (on_face(ph, xw, gy, gz, L.lx, L.ly, L.lz)) {
        const Rsrc R = make_rsrc(L.fin + (uint64_t)xw * kQ * L.ps);
        const uint32_t own = self + kQ * psb;  // own node, relative to memory plane xw
        C a[kQ];
#pragma unroll
        for (int k = 0; k < kQ; ++k) {
          const S v = bload<S>(R, own + (uint32_t)k * psb, 0);
          if constexpr (kNS) a[k] = decode_at<Mo>(v, k, ph, osh);
          else a[k] = decode<Mo>(v, k, ph);
        }
#pragma unroll
        for (int i = 1; i < kQ; ++i) {
          const int fc = entering_face(ph, i, xw, gy, gz, L.lx, L.ly, L.lz);
          if (fc >= 0) f[i] = abb_value<Mo>(a, i, ph.face_rho[fc]);
        }
      }
    }
  };

  // Node-shifted form: decode the 19 stored values pulled for the node at
  // rank-region position q0 of plane xp, each with the shift of the node
  // that wrote it (the source, or the node itself for a bounce-back), from
  // the shift regions of planes xp-1, xp, xp+1.  Returns the node's own shift.
  auto decode_shifted = [&](int xp, int q0, const S* raw, C* f) -> float {
    const S* s0 = shr + slot(xp) * T.RYZ + q0;
    const S* sm = shr + slot(xp - 1) * T.RYZ + q0;
    const S* sq = shr + slot(xp + 1) * T.RYZ + q0;
    const float own = shift_value<Mo>(s0[0]);
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const S* si = ex(i) == 1 ? sm : (ex(i) == -1 ? sq : s0);
      const S v = si[-ey(i) * RZ - ez(i)];
      float sh = own;
      if constexpr (sizeof(S) == 2) {
        if (__builtin_bit_cast(uint16_t, v) != kSolidShift) sh = shift_value<Mo>(v);
      }
      f[i] = decode_at<Mo>(raw[i], i, ph, sh);
    }
    return own;
  };

  // Pull of the fluid node at psi position p of plane xp: stored values raw,
  // populations f (face rules applied), density.  Sources in planes xp-1,
  // xp, xp+1 lie in memory planes xw, xw+1, xw+2 (xw = xp wrapped), one
  // buffer resource.  A solid source gives the halfway bounce-back B_i =
  // A_opp(i) of the node itself (iteration_kernel.cuh:112-115).
  auto pull = [&](int xp, int p, S* raw, C* f, int& gy, int& gz, uint32_t& self) -> C {
    const int xw = wrapx(xp);
    const int r = (int)__umulhi((uint32_t)p, T.magic_pz);
    const int c = p - r * PZ;
    const int q0 = (r + 1) * RZ + (c + 1);
    gy = ycrd[r + 1];
    gz = zcrd[c + 1];
    const uint32_t* m0 = rm + slot(xp) * T.RYZ + q0;
    const uint32_t* mm = rm + slot(xp - 1) * T.RYZ + q0;
    const uint32_t* mq = rm + slot(xp + 1) * T.RYZ + q0;
    self = m0[0];
    // psb is made opaque here so the 38 multiples of it are formed at their
    // use instead of being kept live in scalar registers across the loop
    uint32_t psb_o = psb;
    asm volatile("" : "+s"(psb_o));
    const Rsrc R = make_rsrc(L.fin + (uint64_t)xw * kQ * L.ps);
    // the wave issues its 19 gathers at raised priority, ahead of waves in
    // their collision arithmetic (512^3: FP32 3.05 -> 2.96 ms, FP16S 2.13 ->
    // 2.11 ms; keeping the priority up to the density, or also over the
    // collision's stores, was slower)
    __builtin_amdgcn_s_setprio(3);
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const uint32_t* mi = ex(i) == 1 ? mm : (ex(i) == -1 ? mq : m0);
      const uint32_t sr = mi[-ey(i) * RZ - ez(i)];
      // source in plane xp - e_x(i): sub-plane (1 - e_x)*19 + i from memory
      // plane xw; a solid source: own node, sub-plane 19 + opp(i)
      const bool fl = sr != ~0u;
      const uint32_t base = fl ? sr : self;
      const uint32_t k = fl ? (uint32_t)((1 - ex(i)) * kQ + i) : (uint32_t)(kQ + opp(i));
      raw[i] = bload<S>(R, base + __umul24(k, psb_o), 0);
    }
    __builtin_amdgcn_s_setprio(0);
    float osh = 0.f;
    if constexpr (kNS) {
      osh = decode_shifted(xp, q0, raw, f);
    } else {
#pragma unroll
      for (int i = 0; i < kQ; ++i) f[i] = decode<Mo>(raw[i], i, ph);
    }
    apply_faces(xp, f, gy, gz, self, osh);
    return density<C, !kPlain>(f, gy, ph);
  };

  // (2) psi of every fluid node of plane xp's psi region.  Interior nodes
  // come first in a list and number at most kThreads, so entry tid is this
  // thread's interior node when tid < n_interior: its populations, density
  // and position stay in registers (the *N set) until the collision of
  // plane xp one iteration later.  Entries from kThreads on are ring nodes
  // and are pulled first, while only the previous plane's set is live.
  // 16-bit storage keeps the stored values of the kept node, two per
  // register, and decodes them again (with the face rules) at the collision;
  // wider storage keeps the populations themselves, and so do the plain
  // node-shifted kernels, whose LDS (not their registers) limits them to
  // three workgroups per CU: the second decode would read the 19 source
  // shifts again
  constexpr bool kPackN = sizeof(S) == 2 && !(kNS && kPlain && !kFaces);
  constexpr int kNW = kPackN ? (kQ + 1) / 2 : kQ;
  using NT = std::conditional_t<kPackN, uint32_t, C>;
  NT fN[kNW];
  C rhoN = (C)0;
  int pN = 0;
  uint32_t ownN = 0;
  bool hasN = false;

  // Work of plane xp: the tile's list entries and this thread's first entry,
  // fetched one plane ahead of its pull so that the entry's load completes
  // with the previous plane's population loads instead of adding a memory
  // round trip before the pull's own address computation.
  struct PlaneWork {
    uint32_t first;
    int nint, npull;
  };
  auto plane_work = [&](int xp) -> PlaneWork {
    // uniform index: a scalar load through the constant address space
    using CU64 = const __attribute__((address_space(4))) unsigned long long;
    const unsigned long long tw = ((CU64*)W.tab)[(uint64_t)tile * L.lx + wrapx(xp)];
    PlaneWork w;
    w.first = __builtin_amdgcn_readfirstlane((uint32_t)tw);
    w.nint = (int)__builtin_amdgcn_readfirstlane((uint32_t)(tw >> 32) & 0xFFFFu);
    w.npull = (int)__builtin_amdgcn_readfirstlane((uint32_t)(tw >> 48));
    return w;
  };
  // this thread's list entry (entry 0 of the plane when it has none: the
  // list always holds one more entry than the last plane uses)
  auto entry_of = [&](const PlaneWork& w) -> int {
    return (int)W.pos[w.first + (tid < w.npull ? (uint32_t)tid : 0u)];
  };

  auto psi_plane = [&](int xp, bool keep, const PlaneWork& w, int pe) {
    const uint32_t first = w.first;
    const int nint = w.nint;
    const int npull = w.npull;
    C* pz = psl + slot(xp) * T.PYZ;
    for (int k = tid + kThreads; k < npull; k += kThreads) {
      const int p = W.pos[first + k];
      S raw[kQ];
      C f[kQ];
      int gy, gz;
      uint32_t self;
      const C rho = pull(xp, p, raw, f, gy, gz, self);
      pz[p] = psi_of<Mo, !kPlain>(ph, rho);
    }
    hasN = false;
    if (tid < npull) {
      if constexpr (sizeof(S) == 4 || (sizeof(S) == 2 && kPlain)) pN = pe;  // entry loaded one plane ahead
      else pN = W.pos[first + tid];
      int gy, gz;
      S raw[kQ];
      C f[kQ];
      rhoN = pull(xp, pN, raw, f, gy, gz, ownN);
      pz[pN] = psi_of<Mo, !kPlain>(ph, rhoN);
      hasN = keep && tid < nint;
      if constexpr (kPackN) {
#pragma unroll
        for (int k = 0; k < kNW; ++k) {
          const uint32_t lo = __builtin_bit_cast(uint16_t, raw[2 * k]);
          const uint32_t hi = 2 * k + 1 < kQ ? (uint32_t)__builtin_bit_cast(uint16_t, raw[2 * k + 1]) : 0u;
          fN[k] = lo | (hi << 16);
        }
      } else {
#pragma unroll
        for (int i = 0; i < kQ; ++i) fN[i] = f[i];
      }
    }
  };

  // collision of the fluid interior node at psi position p of plane x (rank
  // offset own): force, velocity shift, BGK, encode, stores (with the ghost
  // copies of the first and last planes); macro fields on macro steps
  auto collide_node = [&](int x, int p, C* f, C rho, int gy, int gz, uint32_t own) {
    const int xw = wrapx(x);
    const C* pzm = psl + slot(x - 1) * T.PYZ;
    const C* pz0 = psl + slot(x) * T.PYZ;
    const C* pzq = psl + slot(x + 1) * T.PYZ;
    C F[3];
    bool skip = false;
    if constexpr (kFaces) skip = force_skipped(ph, xw, gy, gz, L.lx, L.ly, L.lz);
    if (skip) {
      F[0] = F[1] = F[2] = (C)0;
    } else {
      C pn[kQ];
      pn[0] = (C)0;
#pragma unroll
      for (int i = 1; i < kQ; ++i) {
        const C* pp = ex(i) == 1 ? pzq : (ex(i) == -1 ? pzm : pz0);
        pn[i] = pp[p + ey(i) * PZ + ez(i)];
      }
      sc_force(pn, pz0[p], ph, F);
    }
    C u[3];
    collide<Mo>(f, rho, F, ph, u);
    S e[kQ];
    if constexpr (kSR) {
      // dither stream of this node and step: a mixed seed, then a linear
      // congruential sequence over the populations
      uint32_t rs = mix32(own ^ ((uint32_t)xw * 0x9E3779B1u) ^ ph.sr_seed);
      float enc_add = ph.enc_add;
      if constexpr (kNS) {  // node shift: deviations from the node's 16-bit density
        const S sh = shift_of<Mo>((float)rho);
        enc_add = -shift_value<Mo>(sh) * ph.scale;
        L.sout[sidx(L, xw, own / sz)] = sh;
      }
#pragma unroll
      for (int i = 0; i < kQ; ++i) {
        rs = rs * 1664525u + 1013904223u;
        e[i] = encode_sr<Mo>(f[i], i, ph, rs, enc_add);
      }
    } else if constexpr (kNS) {
      // the node's shift: its density in the 16-bit type, stored beside it
      const S sh = shift_of<Mo>((float)rho);
      const float enc_add = -shift_value<Mo>(sh) * ph.scale;
#pragma unroll
      for (int i = 0; i < kQ; ++i) e[i] = encode_at<Mo>(f[i], i, ph, enc_add);
      L.sout[sidx(L, xw, own / sz)] = sh;
    } else {
#pragma unroll
      for (int i = 0; i < kQ; ++i) e[i] = encode<Mo>(f[i], i, ph);
    }
    const Rsrc O = make_rsrc(L.fout + (uint64_t)(xw + 1) * kQ * L.ps);
    uint32_t psb_o = psb;
    asm volatile("" : "+s"(psb_o));
#pragma unroll
    for (int i = 0; i < kQ; ++i) bstore<S, compact_store_policy<S>()>(e[i], O, own, (uint32_t)i * psb_o);
    if (xw == 0 || xw == L.lx - 1) {  // ghost copies of the first and last planes
      if (xw == 0) {
        const Rsrc G = make_rsrc(L.fout + (uint64_t)(L.lx + 1) * kQ * L.ps);
#pragma unroll
        for (int i = 0; i < kQ; ++i) bstore<S, compact_store_policy<S>()>(e[i], G, own, (uint32_t)i * psb);
      }
      if (xw == L.lx - 1) {
        const Rsrc G = make_rsrc(L.fout);
#pragma unroll
        for (int i = 0; i < kQ; ++i) bstore<S, compact_store_policy<S>()>(e[i], G, own, (uint32_t)i * psb);
      }
    }
    if (!kPlain && ph.write_macro) {
      const uint64_t self = ((uint64_t)xw * L.ly + gy) * L.lz + gz;
      L.rho[self] = (typename Mo::M)rho;
      L.pres[self] = (typename Mo::M)pressure_of<Mo>(ph, rho);
      L.u[self] = (typename Mo::M)u[0];
      L.u[(uint64_t)L.ncell + self] = (typename Mo::M)u[1];
      L.u[2ull * L.ncell + self] = (typename Mo::M)u[2];
    }
  };

  // prologue: rank regions of planes x0-2 .. x0+2, psi of planes x0-1, x0
  for (int k = -2; k <= 2; ++k) expand(x0 + k);
  __syncthreads();
  {
    const PlaneWork w = plane_work(x0 - 1);
    psi_plane(x0 - 1, false, w, entry_of(w));
  }
  {
    const PlaneWork w = plane_work(x0);
    psi_plane(x0, true, w, entry_of(w));
  }
  // With 4-byte storage, and in the plain 16-bit kernels, the list entry of
  // the next plane's pull is loaded one plane ahead, before this plane's
  // population loads, and kept in this thread's LDS word (written after the
  // pull has waited for its loads), so no wait on it precedes the pull's
  // address computation: such a wait (waiting for the entry alone, 512^3
  // FP32 3.18 -> 3.06 ms, plain FP16S 2.05 -> 1.95 ms) also waits for the
  // stores of the previous collision, as loads and stores share one
  // counter.  The general 16-bit kernels load the entry in the pull: they
  // run at 80 registers, and one more held across the pull splits its 19
  // loads into two batches (FP16S 2.11 -> 2.18 ms); the plain ones have the
  // room (72 -> 73 VGPRs).  FP64 would drop from three waves per SIMD to two.
  constexpr bool kAhead = sizeof(S) == 4 || (sizeof(S) == 2 && kPlain);  // as in psi_plane
  PlaneWork wc = plane_work(x0 + 1);
  if constexpr (kAhead) pent[tid] = entry_of(wc);

  for (int j = 0; j < n; ++j) {
    const int x = __builtin_amdgcn_readfirstlane(x0 + j);
    const bool more = j + 2 <= n;  // rank regions are needed up to plane x1 + 1
    uint2 wpre[kPre];
    if (more) expand_issue(x + 3, wpre);
    // list entries of plane x+2, pulled in the next iteration
    PlaneWork wn{};
    int en = 0;
    if (more) {
      wn = plane_work(x + 2);
      if constexpr (kAhead) en = entry_of(wn);
    }
    // this thread's interior node of plane x, pulled one iteration ago
    NT fk[kNW];
#pragma unroll
    for (int k = 0; k < kNW; ++k) fk[k] = fN[k];
    const C rho = rhoN;
    const int p = pN;
    const uint32_t own = ownN;
    const bool has = hasN;
    if constexpr (kAhead) {
      psi_plane(x + 1, true, wc, pent[tid]);
      pent[tid] = en;  // each thread reads back only its own word
    } else {
      psi_plane(x + 1, true, wc, 0);
    }
    wc = wn;
    if (more) expand_finish(x + 3, wpre);
    __syncthreads();

    // (4) collide this thread's fluid interior node of plane x
    if (has) {
      const int r = (int)__umulhi((uint32_t)p, T.magic_pz);
      const int gy = ycrd[r + 1], gz = zcrd[p - r * PZ + 1];
      C f[kQ];
      if constexpr (kPackN) {
        float osh = 0.f;
        if constexpr (kNS) {
          S raw[kQ];
#pragma unroll
          for (int i = 0; i < kQ; ++i) raw[i] = __builtin_bit_cast(S, (uint16_t)(fk[i / 2] >> (16 * (i & 1))));
          osh = decode_shifted(x, (r + 1) * RZ + (p - r * PZ + 1), raw, f);
        } else {
#pragma unroll
          for (int i = 0; i < kQ; ++i)
            f[i] = decode<Mo>(__builtin_bit_cast(S, (uint16_t)(fk[i / 2] >> (16 * (i & 1)))), i, ph);
        }
        apply_faces(x, f, gy, gz, own, osh);
      } else {
#pragma unroll
        for (int i = 0; i < kQ; ++i) f[i] = fk[i];
      }
      collide_node(x, p, f, rho, gy, gz, own);
    }
  }
}

// Occupancy the kernels without faces are held to where they fit without
// spilling: 6 waves per SIMD (80 VGPRs), three 512-thread workgroups per CU
// instead of two, for the node-shifted 16-bit kernels (81 left alone) and
// the plain FP32 kernels (83 and 82 left alone; 512^3 FP32 FAST 2.72 ->
// 2.52 ms); 8 waves (64 VGPRs, four 512-thread workgroups per CU, whose
// 40,000 B of LDS each fill the CU's 160 KiB) for the plain 16-bit kernels
// without node shifts (73 left alone; FP16S 1.94 -> 1.93 ms, stochastic
// rounding 2.07 -> 2.04 ms; they load one rank word ahead, as two would
// spill 4 VGPRs); 1 leaves the register budget to the compiler (the other
// variants would spill at 80).
template <class Mo, bool kNS, bool kFaces, bool kPlain>
constexpr bool compact_lean() {
  return !kFaces && kPlain && !kNS && sizeof(typename Mo::S) == 2;
}
template <class Mo, bool kNS, bool kFaces, bool kPlain>
constexpr int compact_min_waves() {
  if constexpr (kFaces) return 1;
  if constexpr (compact_lean<Mo, kNS, kFaces, kPlain>()) return 8;
  return (kNS || (kPlain && Mo::kStorage == MPLBM_FP32)) ? 6 : 1;
}

template <class Mo, bool kFaces, int kThreads, bool kSR = false, bool kNS = false, bool kPlain = false, int kTZ = 0>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(compact_min_waves<Mo, kNS, kFaces, kPlain>())))
void k_compact(Lattice<Mo> L, Phys<typename Mo::C> ph, CTiling T, CLists W) {
  extern __shared__ unsigned char smem_raw[];
  int b = blockIdx.x;
  if (T.xcd_remap) {
    const int per = gridDim.x >> 3;
    b = (b & 7) * per + (b >> 3);
  }
  if (b >= T.nblocks) return;  // whole padding block leaves before any barrier
  const int ntile = T.tiles_y * T.tiles_z;
  const int k = b / ntile;
  constexpr int kPre = compact_lean<Mo, kNS, kFaces, kPlain>() ? 1 : 2;
  compact_march<Mo, kFaces, kThreads, kSR, kNS, kPlain, kPre, kTZ>(L, ph, T, W, T.seg0 + k * T.seg_stride,
                                                                   b - k * ntile, smem_raw);
}

// Store population i of the fluid node of rank r in plane x of a compact
// lattice, and its copy in the ghost plane when x is the first or last plane.
template <class Mo>
__device__ __forceinline__ void compact_store_all(const Lattice<Mo>& L, typename Mo::S* lat, int i, int x,
                                                  uint32_t r, typename Mo::S v) {
  lat[cpix(L, i, x + 1, r)] = v;
  if (x == 0) lat[cpix(L, i, L.lx + 1, r)] = v;
  if (x == L.lx - 1) lat[cpix(L, i, 0, r)] = v;
}

// ----------------------------------------------------------------------------
// set-up kernels

// link word: bit 0 solid cell, bit i upstream node x - e_i solid
__global__ void k_links(const uint8_t* __restrict__ fluid, uint32_t* __restrict__ link, int lx, int ly, int lz) {
  const uint32_t n = (uint32_t)lx * ly * lz;
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= n) return;
  const uint32_t plane = (uint32_t)ly * lz;
  const int x = (int)(c / plane);
  const uint32_t r = c - (uint32_t)x * plane;
  const int y = (int)(r / (uint32_t)lz);
  const int z = (int)(r - (uint32_t)y * lz);
  uint32_t w = fluid[c] ? 0u : 1u;
#pragma unroll
  for (int i = 1; i < kQ; ++i) {
    const uint32_t u = ((uint32_t)wrapi(x - ex(i), lx) * ly + (uint32_t)wrapi(y - ey(i), ly)) * lz +
                       (uint32_t)wrapi(z - ez(i), lz);
    if (!fluid[u]) w |= 1u << i;
  }
  link[c] = w;
}

// multiphaselbm_init (include/iteration_kernel.cuh:9-29)
template <class Mo>
__global__ void k_init(Lattice<Mo> L, typename Mo::S* fother, typename Mo::S* sother, Phys<typename Mo::C> ph,
                       const double* rho_in) {
  using C = typename Mo::C;
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= L.ncell) return;
  const C rho = (C)rho_in[c];
  const bool fluid = !(L.link[c] & 1u);
  if (!L.compact) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const typename Mo::S v = encode<Mo>(weight<C>(i) * rho, i, ph);
      L.fout[L.pix(i, c)] = v;
      fother[L.pix(i, c)] = v;
    }
  } else if (fluid) {
    int x, y, z;
    cell_coords(L, c, x, y, z);
    const uint32_t r = plane_rank(L.rw, L.W, L.ly, x, y, z);
    bool shifted = false;
    if constexpr (Mo::kDev) {
      if (ph.node_shift) {
        // both lattices and both shift arrays: shift = the node's density
        const typename Mo::S sh = shift_of<Mo>((float)rho);
        const float enc_add = -shift_value<Mo>(sh) * ph.scale;
        L.sout[sidx(L, x, r)] = sh;
        sother[sidx(L, x, r)] = sh;
#pragma unroll
        for (int i = 0; i < kQ; ++i) {
          const typename Mo::S v = encode_at<Mo>(weight<C>(i) * rho, i, ph, enc_add);
          compact_store_all(L, L.fout, i, x, r, v);
          compact_store_all(L, fother, i, x, r, v);
        }
        shifted = true;
      }
    }
    if (!shifted) {
#pragma unroll
      for (int i = 0; i < kQ; ++i) {
        const typename Mo::S v = encode<Mo>(weight<C>(i) * rho, i, ph);
        compact_store_all(L, L.fout, i, x, r, v);
        compact_store_all(L, fother, i, x, r, v);
      }
    }
  }
  if (fluid) {
    L.rho[c] = (typename Mo::M)rho;
    L.pres[c] = (typename Mo::M)pressure_of<Mo>(ph, rho);
    L.psi[c] = psi_of<Mo>(ph, rho);
  } else {
    L.rho[c] = (typename Mo::M)ph.rhos;
    L.pres[c] = (typename Mo::M)0;
    L.psi[c] = psi_of<Mo>(ph, ph.rhos);
  }
  L.u[c] = 0;
  L.u[(uint64_t)L.ncell + c] = 0;
  L.u[2ull * L.ncell + c] = 0;
}

// decode the current lattice into [cell][19] doubles / encode back
// element index of population i of cell c in either layout (compact: fluid
// cells only)
// shift of fluid cell c's stored populations (node-shifted form; 0 otherwise)
template <class Mo>
__device__ __forceinline__ float own_shift(const Lattice<Mo>& L, const Phys<typename Mo::C>& ph, uint32_t c) {
  if constexpr (Mo::kDev) {
    if (ph.node_shift) {
      int x, y, z;
      cell_coords(L, c, x, y, z);
      return shift_value<Mo>(L.sin[sidx(L, x, plane_rank(L.rw, L.W, L.ly, x, y, z))]);
    }
  }
  return 0.f;
}
template <class Mo>
__device__ __forceinline__ typename Mo::C decode_own(typename Mo::S s, int i, const Phys<typename Mo::C>& ph,
                                                     float shift) {
  if constexpr (Mo::kDev) {
    if (ph.node_shift) return decode_at<Mo>(s, i, ph, shift);
  }
  return decode<Mo>(s, i, ph);
}

template <class Mo>
__device__ __forceinline__ uint64_t any_pix(const Lattice<Mo>& L, int i, uint32_t c) {
  if (!L.compact) return L.pix(i, c);
  int x, y, z;
  cell_coords(L, c, x, y, z);
  return cpix(L, i, x + 1, plane_rank(L.rw, L.W, L.ly, x, y, z));
}

template <class Mo>
__global__ void k_read_pops(Lattice<Mo> L, Phys<typename Mo::C> ph, double* out) {
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= L.ncell) return;
  if (L.compact && (L.link[c] & 1u)) {  // no storage for solid nodes
#pragma unroll
    for (int i = 0; i < kQ; ++i) out[(uint64_t)c * kQ + i] = 0.0;
    return;
  }
  const float sh = own_shift(L, ph, c);
#pragma unroll
  for (int i = 0; i < kQ; ++i) out[(uint64_t)c * kQ + i] = (double)decode_own<Mo>(L.fin[any_pix(L, i, c)], i, ph, sh);
}

template <class Mo>
__global__ void k_write_pops(Lattice<Mo> L, Phys<typename Mo::C> ph, const double* in) {
  using C = typename Mo::C;
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= L.ncell) return;
  if (!L.compact) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) L.fout[L.pix(i, c)] = encode<Mo>((C)in[(uint64_t)c * kQ + i], i, ph);
    return;
  }
  if (L.link[c] & 1u) return;
  int x, y, z;
  cell_coords(L, c, x, y, z);
  const uint32_t r = plane_rank(L.rw, L.W, L.ly, x, y, z);
  if constexpr (Mo::kDev) {
    if (ph.node_shift) {
      // shift: the node's density, summed in float in direction order
      float rho = 0.f;
#pragma unroll
      for (int i = 0; i < kQ; ++i) rho += (float)in[(uint64_t)c * kQ + i];
      const typename Mo::S sh = shift_of<Mo>(rho);
      const float enc_add = -shift_value<Mo>(sh) * ph.scale;
      L.sout[sidx(L, x, r)] = sh;
#pragma unroll
      for (int i = 0; i < kQ; ++i)
        compact_store_all(L, L.fout, i, x, r, encode_at<Mo>((C)in[(uint64_t)c * kQ + i], i, ph, enc_add));
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < kQ; ++i) compact_store_all(L, L.fout, i, x, r, encode<Mo>((C)in[(uint64_t)c * kQ + i], i, ph));
}

// Macro fields and psi of every cell from the populations of fin, for a
// context whose populations were loaded from a checkpoint: solid cells as
// k_init leaves them (rho = rhos, p = 0, psi = Psi(rhos), u = 0); a fluid
// cell gets rho = sum f (ascending i), P(rho), Psi(rho) and u = sum f e / rho.
template <class Mo>
__global__ void k_macro_from_pops(Lattice<Mo> L, Phys<typename Mo::C> ph) {
  using C = typename Mo::C;
  const uint32_t c = blockIdx.x * 256u + threadIdx.x;
  if (c >= L.ncell) return;
  if (L.link[c] & 1u) {
    L.rho[c] = (typename Mo::M)ph.rhos;
    L.pres[c] = (typename Mo::M)0;
    L.psi[c] = psi_of<Mo>(ph, ph.rhos);
    L.u[c] = 0;
    L.u[(uint64_t)L.ncell + c] = 0;
    L.u[2ull * L.ncell + c] = 0;
    return;
  }
  const float sh = own_shift(L, ph, c);
  C rho = (C)0, m[3] = {(C)0, (C)0, (C)0};
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const C f = decode_own<Mo>(L.fin[any_pix(L, i, c)], i, ph, sh);
    rho = rho + f;
    m[0] = m[0] + f * (C)ex(i);
    m[1] = m[1] + f * (C)ey(i);
    m[2] = m[2] + f * (C)ez(i);
  }
  L.rho[c] = (typename Mo::M)rho;
  L.pres[c] = (typename Mo::M)pressure_of<Mo>(ph, rho);
  L.psi[c] = psi_of<Mo>(ph, rho);
  const C inv = (C)(1.0 / (double)rho);
  L.u[c] = (typename Mo::M)(m[0] * inv);
  L.u[(uint64_t)L.ncell + c] = (typename Mo::M)(m[1] * inv);
  L.u[2ull * L.ncell + c] = (typename Mo::M)(m[2] * inv);
}

// Stored values (no decoding) of the 19 populations of every cell of plane
// x, [y][z][19], solid cells 0: gather from fin, or scatter to fout (own
// slots and the compact layout's ghost planes).
template <class Mo>
__global__ void k_plane_raw(Lattice<Mo> L, int x, typename Mo::S* buf, int scatter) {
  using S = typename Mo::S;
  const uint32_t r = blockIdx.x * 256u + threadIdx.x;
  if (r >= L.plane) return;
  const uint32_t c = (uint32_t)x * L.plane + r;
  S* b = buf + (uint64_t)r * kQ;
  const bool fluid = !(L.link[c] & 1u);
  if (!scatter) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) b[i] = fluid ? L.fin[any_pix(L, i, c)] : S{};
    return;
  }
  if (!fluid) return;
  if (!L.compact) {
#pragma unroll
    for (int i = 0; i < kQ; ++i) L.fout[L.pix(i, c)] = b[i];
    return;
  }
  const int y = (int)(r / (uint32_t)L.lz), z = (int)(r - (uint32_t)y * L.lz);
  const uint32_t rk = plane_rank(L.rw, L.W, L.ly, x, y, z);
#pragma unroll
  for (int i = 0; i < kQ; ++i) compact_store_all(L, L.fout, i, x, rk, b[i]);
}

// Node shifts of plane x, [y][z], solid cells 0: gather from sin or scatter
// to sout (node-shifted compact form).
template <class Mo>
__global__ void k_plane_shift(Lattice<Mo> L, int x, typename Mo::S* buf, int scatter) {
  using S = typename Mo::S;
  const uint32_t r = blockIdx.x * 256u + threadIdx.x;
  if (r >= L.plane) return;
  const uint32_t c = (uint32_t)x * L.plane + r;
  const bool fluid = !(L.link[c] & 1u);
  const int y = (int)(r / (uint32_t)L.lz), z = (int)(r - (uint32_t)y * L.lz);
  if (!scatter) {
    buf[r] = fluid ? L.sin[sidx(L, x, plane_rank(L.rw, L.W, L.ly, x, y, z))] : S{};
    return;
  }
  if (fluid) L.sout[sidx(L, x, plane_rank(L.rw, L.W, L.ly, x, y, z))] = buf[r];
}

// ----------------------------------------------------------------------------
// FP64 reductions: per-block partials, then one block folds them in order

constexpr int kRedFields = 16;  // see mplbm_stats64 order
constexpr int kRedThreads = 256;

template <class Mo>
__global__ __launch_bounds__(kRedThreads) void k_reduce(Lattice<Mo> L, Phys<typename Mo::C> ph, double rhol,
                                                        double rhog, double* partial) {
  __shared__ double sh[kRedThreads];
  double acc[kRedFields];
#pragma unroll
  for (int k = 0; k < kRedFields; ++k) acc[k] = 0.0;
  acc[11] = INFINITY; acc[12] = -INFINITY; acc[13] = INFINITY; acc[14] = -INFINITY;
  for (uint32_t c = blockIdx.x * kRedThreads + threadIdx.x; c < L.ncell; c += gridDim.x * kRedThreads) {
    if (L.link[c] & 1u) continue;
    double s = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
    const float sh = own_shift(L, ph, c);
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const double f = (double)decode_own<Mo>(L.fin[any_pix(L, i, c)], i, ph, sh);
      if (!(f - f == 0.0)) acc[15] += 1.0;
      s += f;
      mx += f * ex(i);
      my += f * ey(i);
      mz += f * ez(i);
    }
    acc[0] += s; acc[1] += mx; acc[2] += my; acc[3] += mz;
    acc[4] += 1.0;
    const double r = (double)L.rho[c], p = (double)L.pres[c];
    if (r > rhol) { acc[5] += 1.0; acc[7] += r; acc[9] += p; }
    else if (r < rhog) { acc[6] += 1.0; acc[8] += r; acc[10] += p; }
    acc[11] = fmin(acc[11], r); acc[12] = fmax(acc[12], r);
    acc[13] = fmin(acc[13], p); acc[14] = fmax(acc[14], p);
  }
  for (int k = 0; k < kRedFields; ++k) {
    sh[threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
        const double a = sh[threadIdx.x], bb = sh[threadIdx.x + s];
        sh[threadIdx.x] = (k >= 11 && k <= 14) ? ((k & 1) ? fmin(a, bb) : fmax(a, bb)) : a + bb;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(uint64_t)blockIdx.x * kRedFields + k] = sh[0];
    __syncthreads();
  }
}

__global__ void k_reduce_final(const double* partial, int nblocks, double* out) {
  const int k = threadIdx.x;
  if (k >= kRedFields) return;
  double a = partial[k];
  for (int b = 1; b < nblocks; ++b) {
    const double v = partial[(uint64_t)b * kRedFields + k];
    a = (k >= 11 && k <= 14) ? ((k & 1) ? fmin(a, v) : fmax(a, v)) : a + v;
  }
  out[k] = a;
}

}  // namespace mplbm

// ============================================================================
// host side

using namespace mplbm;

struct mplbm_ctx {
  mplbm_config cfg{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  uint64_t ncell = 0;
  uint64_t ps = 0;  // population sub-plane stride (elements)
  int bytes_pop = 4;
  void* lat[2] = {nullptr, nullptr};  // population lattices
  void* shift[2] = {nullptr, nullptr};  // node shifts with each lattice (MPLBM_SHIFT_NODE)
  int cur = 0;                        // lattice holding the newest post-collision state
  uint32_t* link = nullptr;
  uint8_t* mask = nullptr;
  void* rho = nullptr;
  void* pres = nullptr;
  void* u = nullptr;
  void* psi = nullptr;
  double* red_partial = nullptr;
  double* red_out = nullptr;
  int red_blocks = 0;
  bool have_geometry = false, have_fields = false;
  int64_t fluid_nodes = 0;
  int64_t device_bytes = 0;
  Tiling tiling{};
  int kernel = MPLBM_KERNEL_FUSED;
  float rho_ref = 0.f, scale = 1.f;
  std::string err;
  // compact layout
  bool compact = false;
  uint2* rw = nullptr;          // plane-rank words
  uint32_t W = 0;               // 32-wide z blocks per row (three rank words each)
  int rank_rows = 1;            // rows per block of the plane-rank order
  uint16_t* cpos = nullptr;     // work-list entries
  uint2* ctab = nullptr;        // [tile][x] list table
  size_t cpos_cap = 0, ctab_cap = 0;
  CTiling ctiling{};
  std::vector<uint8_t> hmask;   // host copy of the fluid mask (list rebuilds, checkpoints)
  size_t lat_bytes = 0;
  size_t shift_bytes = 0;  // one node-shift array
  // cluster statistics (allocated at the first mplbm_calc_cluster)
  int32_t* cl_parent = nullptr;  // union-find parents, then cluster labels
  int32_t* cl_size = nullptr;    // component sizes at roots, then cluster numbers
  uint8_t* cl_phase = nullptr;   // 0 solid or intermediate, 1 liquid, 2 vapour
  void* cl_acc = nullptr;        // ClAcc
  int32_t* cl_sid = nullptr;     // solid numbers (mplbm_set_solid_numbers)
  int64_t cl_solid_count = 0;
};

static std::string g_create_error;

static mplbm_status fail(mplbm_ctx* ctx, mplbm_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  else g_create_error = buf;
  return st;
}

#define HIP_TRY(ctx, call)                                                                      \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail((ctx), e_ == hipErrorOutOfMemory ? MPLBM_ERR_NOMEM : MPLBM_ERR_HIP, "%s: %s (%s:%d)", \
                  #call, hipGetErrorString(e_), __FILE__, __LINE__);                            \
  } while (0)

static size_t storage_bytes(int storage) {
  switch (storage) {
    case MPLBM_FP32: return 4;
    case MPLBM_FP64: return 8;
    default: return 2;
  }
}

template <class C>
static Phys<C> make_phys(const mplbm_ctx* ctx, int64_t step, int32_t reverse, uint32_t inject_type,
                         bool write_macro) {
  const mplbm_config& c = ctx->cfg;
  Phys<C> ph{};
  ph.eos.G = (C)c.G;
  ph.eos.cs2 = (C)(1.0 / 3.0);
  ph.eos.half_cs2G = ph.eos.cs2 * ph.eos.G / (C)2;
  ph.eos.T = (C)c.T_Tc * (C)0.0943;
  ph.eos.carn_star = c.carn_star != 0;
  ph.tau = (C)c.tau;
  ph.inv_tau = (C)(1.0 / (double)ph.tau);
  ph.neg_G = -ph.eos.G;
  ph.rhos = (C)c.rhos;
  ph.rhol_sp = (C)c.rhol_spinodal;
  ph.inc = (C)c.rho_increment * (C)reverse;
  ph.inject_ylim = (C)c.inject_position * (C)c.ly;
  ph.inject_now = c.rho_inject_period > 0 && ((uint64_t)step % (uint64_t)c.rho_inject_period) == 0;
  ph.inject_type = (int)inject_type;
  ph.has_body = c.body_force[0] != 0 || c.body_force[1] != 0 || c.body_force[2] != 0;
  for (int a = 0; a < 3; ++a) ph.g[a] = (C)c.body_force[a];
  ph.any_face = 0;
  for (int f = 0; f < 6; ++f) {
    ph.face[f] = c.face_type[f];
    ph.face_rho[f] = (C)c.face_rho[f];
    ph.any_face |= c.face_type[f] != 0;
  }
  ph.rho_ref = ctx->rho_ref;
  ph.node_shift = c.shift == MPLBM_SHIFT_NODE ? 1 : 0;
  ph.scale = ctx->scale;
  ph.inv_scale = 1.0f / ctx->scale;
  for (int k = 0; k < 3; ++k) {
    const int i = k == 0 ? 0 : (k == 1 ? 1 : 4);  // a direction of each weight class
    ph.dec_mul[k] = weight<float>(i) * ph.inv_scale;
    ph.dec_add[k] = weight<float>(i) * ph.rho_ref;
    ph.enc_mul[k] = inv_weight_f(i) * ph.scale;
  }
  ph.enc_add = -ph.rho_ref * ph.scale;
  ph.write_macro = write_macro ? 1 : 0;
  ph.sr = c.rounding == MPLBM_ROUND_STOCHASTIC ? 1 : 0;
  ph.sr_seed = mix32((uint32_t)step * 0x9E3779B9u + (uint32_t)((uint64_t)step >> 32) + 0x632BE5ABu);
  return ph;
}

template <class Mo>
static Lattice<Mo> make_lattice(const mplbm_ctx* ctx, int in, int out) {
  Lattice<Mo> L{};
  L.fin = static_cast<const typename Mo::S*>(ctx->lat[in]);
  L.fout = static_cast<typename Mo::S*>(ctx->lat[out]);
  L.sin = static_cast<const typename Mo::S*>(ctx->shift[in]);
  L.sout = static_cast<typename Mo::S*>(ctx->shift[out]);
  L.link = ctx->link;
  L.rho = static_cast<typename Mo::M*>(ctx->rho);
  L.pres = static_cast<typename Mo::M*>(ctx->pres);
  L.u = static_cast<typename Mo::M*>(ctx->u);
  L.psi = static_cast<typename Mo::C*>(ctx->psi);
  L.lx = (int)ctx->cfg.lx;
  L.ly = (int)ctx->cfg.ly;
  L.lz = (int)ctx->cfg.lz;
  L.plane = (uint32_t)(ctx->cfg.ly * ctx->cfg.lz);
  L.ps = (uint32_t)ctx->ps;
  L.ncell = (uint32_t)ctx->ncell;
  L.rw = ctx->rw;
  L.W = ctx->W;
  L.compact = ctx->compact ? 1 : 0;
  return L;
}

// Tile choice for the fused kernel: a whole axis goes into one tile when it
// fits (no ring, neighbours wrap inside LDS); otherwise a ring of one cell.
static mplbm_status choose_tiling(mplbm_ctx* ctx) {
  const mplbm_config& c = ctx->cfg;
  Tiling T{};
  int tz = c.tile_z, ty = c.tile_y, seg = c.seg_x;
  if (tz <= 0) tz = c.lz <= 64 ? (int)c.lz : 34;
  if (tz >= c.lz) { T.wrap_z = 1; tz = (int)c.lz; }
  if (ty <= 0) {
    // about four waves per workgroup (measured best on 512^3 packings)
    ty = std::max(1, 256 / tz);
    if (ty > 32) ty = 32;
  }
  if (ty >= c.ly && (int64_t)tz * c.ly <= 1024) { T.wrap_y = 1; ty = (int)c.ly; }
  if (ty > c.ly) ty = (int)c.ly;
  if (!T.wrap_z && tz < 3) return fail(ctx, MPLBM_ERR_ARG, "tile_z %d too small", tz);
  if (!T.wrap_y && ty < 3) return fail(ctx, MPLBM_ERR_ARG, "tile_y %d too small", ty);
  if ((int64_t)ty * tz > 1024) return fail(ctx, MPLBM_ERR_ARG, "tile %dx%d exceeds 1024 threads", ty, tz);
  T.TY = ty;
  T.TZ = tz;
  T.tiles_y = T.wrap_y ? 1 : (int)((c.ly + (ty - 2) - 1) / (ty - 2));
  T.tiles_z = T.wrap_z ? 1 : (int)((c.lz + (tz - 2) - 1) / (tz - 2));
  if (seg <= 0) {
    // long segments amortise the two warm-up planes; keep >= ~2 blocks per CU
    const int64_t tiles = (int64_t)T.tiles_y * T.tiles_z;
    seg = 32;
    while (seg > 4 && tiles * ((c.lx + seg - 1) / seg) < 1024) seg /= 2;
  }
  if (seg > c.lx) seg = (int)c.lx;
  T.seg = seg;
  T.nseg = (int)((c.lx + seg - 1) / seg);
  const int64_t nb = (int64_t)T.tiles_y * T.tiles_z * T.nseg;
  if (nb > (1LL << 30)) return fail(ctx, MPLBM_ERR_ARG, "too many blocks");
  T.nblocks = (int)nb;
  T.xcd_remap = 0;  // hardware dispatch order, as the compact kernel (6.29 -> 6.09 ms at 512^3)
  T.pull_bb = c.bounce_back == MPLBM_BB_PULL ? 1 : 0;
  ctx->tiling = T;
  return MPLBM_OK;
}

static int fused_threads(const Tiling& T) { return ((T.TY * T.TZ + 63) / 64) * 64; }

// Tile of the compact kernel: TZi x TYi interior nodes (z a multiple of 32
// where the axis allows, so a tile row covers whole rank words), segments of
// seg planes; about 1024 workgroups or more.

// LDS one workgroup may allocate on gfx950
static constexpr size_t kMaxGroupLds = 163840;

static mplbm_status choose_ctiling(mplbm_ctx* ctx) {
  const mplbm_config& c = ctx->cfg;
  CTiling T{};
  int tz = c.tile_z, ty = c.tile_y, seg = c.seg_x;
  // measured on 512^3 packings: 16 x 32 interior nodes (512 threads), and
  // 32 x 32 (1024 threads, two workgroups per CU) for 16-bit storage with
  // round-to-nearest and the global shift (FP16S 1.86 -> 1.83 ms, BF16S
  // 1.86 -> 1.84 ms, same device; with stochastic rounding 1.97 -> 1.98 ms,
  // and the node-shifted kernels' LDS would hold one such workgroup per CU)
  const bool wide16 = (c.storage == MPLBM_FP16S || c.storage == MPLBM_BF16S) &&
                      c.rounding == MPLBM_ROUND_NEAREST && c.shift != MPLBM_SHIFT_NODE;
  const int ty_def = wide16 ? 32 : 16;
  if (tz <= 0) tz = c.lz <= 32 ? (int)c.lz : 32;
  if (ty <= 0) {
    ty = c.ly <= ty_def ? (int)c.ly : ty_def;
    // a workgroup size set without a tile height bounds the default height
    if (c.block_threads > 0) ty = std::max(1, std::min(ty, c.block_threads / std::min(tz, (int)c.lz)));
  }
  if (tz > c.lz) tz = (int)c.lz;
  if (ty > c.ly) ty = (int)c.ly;
  if (tz < 1 || ty < 1) return fail(ctx, MPLBM_ERR_ARG, "tile %dx%d too small", ty, tz);
  if ((int64_t)(ty + 2) * (tz + 2) >= 65536) return fail(ctx, MPLBM_ERR_ARG, "tile %dx%d too large", ty, tz);
  T.TYi = ty;
  T.TZi = tz;
  T.PZ = tz + 2;
  T.RZ = tz + 4;
  T.PYZ = (ty + 2) * (tz + 2);
  T.RYZ = (ty + 4) * (tz + 4);
  T.magic_pz = (uint32_t)(((1ull << 32) + T.PZ - 1) / T.PZ);
  T.magic_rz = (uint32_t)(((1ull << 32) + T.RZ - 1) / T.RZ);
  T.tiles_y = (int)((c.ly + ty - 1) / ty);
  T.tiles_z = (int)((c.lz + tz - 1) / tz);
  {
    const size_t csz = c.storage == MPLBM_FP64 ? 8 : 4;
    const size_t lds = kCompactSlots * ((size_t)T.PYZ * csz + (size_t)T.RYZ * 4) + (size_t)(ty + 4 + tz + 4) * 4 +
                       2 * (size_t)T.RYZ * 4 + 1024 * 4 +
                       (c.shift == MPLBM_SHIFT_NODE ? kCompactSlots * (size_t)T.RYZ * 2 : 0);
    if (lds > kMaxGroupLds)
      return fail(ctx, MPLBM_ERR_ARG, "tile %dx%d needs %zu bytes of LDS (max %zu)", ty, tz, lds, kMaxGroupLds);
  }
  const int64_t tiles = (int64_t)T.tiles_y * T.tiles_z;
  // segments of 48 planes where they leave at least 2048 workgroups, else
  // 32, 16, 8 or 4: fewer set-up planes per marched plane against a longer
  // tail of the last workgroups (512^3 FP32 2.63 / 2.61 / 2.69 ms and FP16S
  // 1.93 / 1.90 / 1.96 ms at 32 / 48 / 64 planes, same device; 1024^3 FP32
  // 2-4% and FP16S 3-7% faster at 48 than at 32)
  if (seg <= 0) {
    seg = 48;
    while (seg > 4 && tiles * ((c.lx + seg - 1) / seg) < 2048) seg = seg == 48 ? 32 : seg / 2;
  }
  if (seg > c.lx) seg = (int)c.lx;
  T.seg = seg;
  T.nseg = (int)((c.lx + seg - 1) / seg);
  const int64_t nb = tiles * T.nseg;
  if (nb > (1LL << 30)) return fail(ctx, MPLBM_ERR_ARG, "too many blocks");
  T.nblocks = (int)nb;
  // plain dispatch order: the hardware deals consecutive workgroups (the
  // tiles of one segment) to the eight XCDs in turn, so every XCD works on
  // the same planes at once and their ring re-reads meet in the device-wide
  // last-level cache; giving each XCD a contiguous range of tiles (its own
  // segments) measured 2.7% slower at 512^3 (3.60 against 3.50 ms) and 0.8%
  // at 1024^3
  T.xcd_remap = 0;
  T.seg0 = 0;
  T.seg_stride = 1;
  if (c.block_threads != 0 && c.block_threads != 256 && c.block_threads != 512 && c.block_threads != 1024)
    return fail(ctx, MPLBM_ERR_ARG, "block_threads must be 0, 256, 512 or 1024");
  {
    const int area = ty * tz;
    const int need = area <= 256 ? 256 : (area <= 512 ? 512 : 1024);
    const int maxt = c.storage == MPLBM_FP64 ? 512 : 1024;  // FP64 needs > 128 VGPRs
    if (area > maxt)
      return fail(ctx, MPLBM_ERR_ARG, "tile %dx%d exceeds %d interior nodes for this storage", ty, tz, maxt);
    if (c.block_threads && c.block_threads < area)
      return fail(ctx, MPLBM_ERR_ARG, "block_threads %d below the tile's %d interior nodes", c.block_threads, area);
    T.threads = c.block_threads ? c.block_threads : need;
    if (T.threads > maxt) return fail(ctx, MPLBM_ERR_ARG, "block_threads %d too large for this storage", T.threads);
  }
  ctx->ctiling = T;
  return MPLBM_OK;
}

// shift_bytes: bytes of a node shift in the node-shifted form, else 0;
// threads: the workgroup size of the launch (one list-entry word per thread)
template <class C>
static size_t compact_lds_bytes(const CTiling& T, size_t shift_bytes, int threads) {
  return kCompactSlots * ((size_t)T.PYZ * sizeof(C) + (size_t)T.RYZ * 4) + (size_t)(T.TYi + 4 + T.TZi + 4) * 4 +
         2 * (size_t)T.RYZ * 4 + (size_t)threads * 4 + kCompactSlots * (size_t)T.RYZ * shift_bytes;
}

// Work lists of the compact kernel: for every tile and plane, the psi-region
// positions (r, c) of the fluid nodes the tile pulls, interior nodes first
// (the ones it collides), then the ring, each in plane-rank order so that a
// wavefront's 64 entries address few memory windows.  The region covers the
// tile's in-domain interior [1, ny] x [1, nz] and the ring of one node
// around it.
static mplbm_status build_lists(mplbm_ctx* ctx) {
  const CTiling& T = ctx->ctiling;
  const int64_t lx = ctx->cfg.lx, ly = ctx->cfg.ly, lz = ctx->cfg.lz;
  const int64_t ntile = (int64_t)T.tiles_y * T.tiles_z;
  std::vector<uint2> tab((size_t)(ntile * lx));
  std::vector<uint16_t> pos;
  pos.reserve((size_t)(ctx->fluid_nodes * 3 / 2 + 1024));
  const uint8_t* m = ctx->hmask.data();
  std::vector<uint16_t> ring;
  // plane-rank order of a cell (the same in every plane): z block, block of
  // rows, column group (0-1, 2-29, 30-31), row, column
  const int64_t yb = ctx->rank_rows;
  const int64_t nyb = (ly + yb - 1) / yb;
  auto order_key = [&](int64_t y, int64_t z) {
    const uint32_t c = (uint32_t)(z & 31);
    const uint64_t g = c < 2 ? 0 : (c >= 30 ? 2 : 1);
    return (((((uint64_t)(z >> 5) * nyb + (uint64_t)(y / yb)) * 3 + g) * (uint64_t)ly + (uint64_t)y) << 5) | c;
  };
  struct Cell {
    uint64_t key;
    uint16_t p;
    uint8_t interior;
    int64_t y, z;
  };
  std::vector<Cell> cells;
  for (int64_t t = 0; t < ntile; ++t) {
    const int64_t tyi = t / T.tiles_z, tzi = t - tyi * T.tiles_z;
    const int64_t y0 = tyi * T.TYi, z0 = tzi * T.TZi;
    const int ny = (int)std::min<int64_t>(T.TYi, ly - y0), nz = (int)std::min<int64_t>(T.TZi, lz - z0);
    cells.clear();
    for (int r = 0; r <= ny + 1; ++r) {
      const int64_t y = ((y0 - 1 + r) % ly + ly) % ly;
      for (int c = 0; c <= nz + 1; ++c) {
        const int64_t z = ((z0 - 1 + c) % lz + lz) % lz;
        const bool in = r >= 1 && r <= ny && c >= 1 && c <= nz;
        cells.push_back(Cell{order_key(y, z), (uint16_t)(r * T.PZ + c), (uint8_t)in, y, z});
      }
    }
    std::sort(cells.begin(), cells.end(),
              [](const Cell& a, const Cell& b) { return a.key != b.key ? a.key < b.key : a.p < b.p; });
    for (int64_t x = 0; x < lx; ++x) {
      const size_t first = pos.size();
      ring.clear();
      const uint8_t* pl = m + (size_t)x * ly * lz;
      for (const Cell& e : cells) {
        if (!pl[(size_t)e.y * lz + e.z]) continue;
        if (e.interior) pos.push_back(e.p);
        else ring.push_back(e.p);
      }
      const size_t nint = pos.size() - first;
      pos.insert(pos.end(), ring.begin(), ring.end());
      const size_t npull = pos.size() - first;
      if (first > 0xFFFFFFFFull) return fail(ctx, MPLBM_ERR_ARG, "work lists exceed 2^32 entries");
      tab[(size_t)(t * lx + x)] = make_uint2((uint32_t)first, (uint32_t)nint | ((uint32_t)npull << 16));
    }
  }
  pos.push_back(0);  // keeps the allocation non-empty
  if (pos.size() > ctx->cpos_cap) {
    if (ctx->cpos) HIP_TRY(ctx, hipFree(ctx->cpos));
    ctx->cpos = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&ctx->cpos, pos.size() * 2));
    ctx->cpos_cap = pos.size();
  }
  if (tab.size() > ctx->ctab_cap) {
    if (ctx->ctab) HIP_TRY(ctx, hipFree(ctx->ctab));
    ctx->ctab = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&ctx->ctab, tab.size() * sizeof(uint2)));
    ctx->ctab_cap = tab.size();
  }
  HIP_TRY(ctx, hipMemcpy(ctx->cpos, pos.data(), pos.size() * 2, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->ctab, tab.data(), tab.size() * sizeof(uint2), hipMemcpyHostToDevice));
  return MPLBM_OK;
}

// Plane ranks, sub-plane stride and lattices of the compact layout.
constexpr uint32_t kRankAlign = 64;  // ranks: one wavefront of nodes
static mplbm_status build_compact(mplbm_ctx* ctx) {
  const int64_t lx = ctx->cfg.lx, ly = ctx->cfg.ly, lz = ctx->cfg.lz;
  const uint32_t W = (uint32_t)((lz + 31) / 32);
  std::vector<uint2> rw((size_t)(lx * ly * W) * kRankGroups);
  uint64_t maxc = 0;
  const uint8_t* m = ctx->hmask.data();
  // plane rank order: 32-wide z blocks outermost, then blocks of rows as
  // high as the tile, so a tile 32 nodes wide is one contiguous range of a
  // plane.  Each range starts at a multiple of 64 ranks (the ranks between
  // are never addressed): a wavefront takes 64 consecutive interior nodes,
  // so its stores and its in-plane pulls cover whole aligned windows (512^3
  // FP32 3.50 -> 3.31 ms, FP16S 2.17 -> 2.13 ms, for ~12% more lattice
  // memory).  Within a range, columns 0-1 of its rows come first, then
  // columns 2-29, then 30-31: the neighbouring tiles' ring nodes and the
  // cross-block pulls of the edge columns read a few short ranges instead of
  // one line per row.  Tiles chosen later keep the layout (and are as fast
  // when their rows stay within these blocks).
  const int64_t yb = std::max(1, ctx->ctiling.TYi);
  if ((uint64_t)ly * W * kRankGroups >= (1ull << 27)) return fail(ctx, MPLBM_ERR_ARG, "planes too large for the compact layout");
  const uint32_t gmask[kRankGroups] = {0x3FFFFFFCu, 0x00000003u, 0xC0000000u};  // columns 2-29, 0-1, 30-31
  const uint32_t gorder[kRankGroups] = {1, 0, 2};                                // numbering order of the groups
  std::vector<uint32_t> rowbits((size_t)ly);
  for (int64_t x = 0; x < lx; ++x) {
    uint64_t r = 0;  // exceeds 32 bits only for planes the size check below refuses
    for (uint32_t w = 0; w < W; ++w) {
      for (int64_t y = 0; y < ly; ++y) {
        const uint8_t* row = m + ((size_t)x * ly + y) * lz;
        uint32_t bits = 0;
        for (int b = 0; b < 32; ++b) {
          const int64_t z = (int64_t)w * 32 + b;
          if (z < lz && row[z]) bits |= 1u << b;
        }
        rowbits[(size_t)y] = bits;
      }
      for (int64_t y0 = 0; y0 < ly; y0 += yb) {
        r = (r + kRankAlign - 1) / kRankAlign * kRankAlign;
        const int64_t y1 = std::min(ly, y0 + yb);
        for (uint32_t g : gorder) {
          for (int64_t y = y0; y < y1; ++y) {
            const uint32_t b = rowbits[(size_t)y] & gmask[g];
            rw[((size_t)(x * ly + y) * W + w) * kRankGroups + g] = make_uint2(b, (uint32_t)r);
            r += (uint64_t)__builtin_popcount(b);
          }
        }
      }
    }
    maxc = std::max<uint64_t>(maxc, r);
  }
  ctx->rank_rows = (int)yb;
  uint64_t ps = std::max<uint64_t>(64, (maxc + 63) / 64 * 64);
  int64_t pad = ctx->cfg.plane_pad;
  if (pad < 0) pad = (ps % 4096 == 0) ? 64 : 0;
  ps += (uint64_t)pad;
  // a pull addresses three planes from one 32-bit byte offset
  if (3ull * kQ * ps * (uint64_t)ctx->bytes_pop >= (1ull << 32))
    return fail(ctx, MPLBM_ERR_ARG, "planes too large for the compact layout");
  const size_t lat_bytes = (size_t)(lx + 2) * kQ * ps * ctx->bytes_pop;
  const size_t shift_bytes = ctx->cfg.shift == MPLBM_SHIFT_NODE ? (size_t)lx * ps * ctx->bytes_pop : 0;
  for (int k = 0; k < 2; ++k) {
    if (ctx->lat[k]) HIP_TRY(ctx, hipFree(ctx->lat[k]));
    ctx->lat[k] = nullptr;
    if (ctx->shift[k]) HIP_TRY(ctx, hipFree(ctx->shift[k]));
    ctx->shift[k] = nullptr;
  }
  if (ctx->rw) HIP_TRY(ctx, hipFree(ctx->rw));
  ctx->rw = nullptr;
  ctx->device_bytes -= (int64_t)(2 * ctx->lat_bytes + 2 * ctx->shift_bytes);
  HIP_TRY(ctx, hipMalloc(&ctx->lat[0], lat_bytes));
  HIP_TRY(ctx, hipMalloc(&ctx->lat[1], lat_bytes));
  if (shift_bytes) {
    for (int k = 0; k < 2; ++k) {
      HIP_TRY(ctx, hipMalloc(&ctx->shift[k], shift_bytes));
      HIP_TRY(ctx, hipMemset(ctx->shift[k], 0, shift_bytes));
    }
  }
  ctx->lat_bytes = lat_bytes;
  ctx->shift_bytes = shift_bytes;
  ctx->device_bytes += (int64_t)(2 * lat_bytes + 2 * shift_bytes);
  HIP_TRY(ctx, hipMalloc((void**)&ctx->rw, rw.size() * sizeof(uint2)));
  HIP_TRY(ctx, hipMemcpy(ctx->rw, rw.data(), rw.size() * sizeof(uint2), hipMemcpyHostToDevice));
  ctx->W = W;
  ctx->ps = ps;
  return build_lists(ctx);
}

template <class Mo>
static mplbm_status launch_step(mplbm_ctx* ctx, int64_t step, int32_t reverse, uint32_t inject_type, bool macro) {
  using C = typename Mo::C;
  const Phys<C> ph = make_phys<C>(ctx, step, reverse, inject_type, macro);
  const Lattice<Mo> L = make_lattice<Mo>(ctx, ctx->cur, 1 - ctx->cur);
  if (ctx->kernel == MPLBM_KERNEL_COMPACT) {
    const CTiling& T = ctx->ctiling;
    const CLists WL{ctx->cpos, ctx->ctab};
    const size_t shm = compact_lds_bytes<C>(T, ph.node_shift ? sizeof(typename Mo::S) : 0, T.threads);
    // the workgroup size is at least the tile's interior count (one kept
    // interior node per thread); 1024 only for non-FP64 storage
    // stochastic rounding and node shifts are separate instantiations so the
    // round-to-nearest kernels keep the fused multiply-add-and-convert of the
    // encode
    auto launch_as = [&](auto sr, auto ns, const CTiling& R, bool faces) {
      constexpr bool kSR = decltype(sr)::value;
      constexpr bool kNS = decltype(ns)::value;
      const int grid = R.xcd_remap ? ((R.nblocks + 7) / 8) * 8 : R.nblocks;
      using KernelPtr = void (*)(Lattice<Mo>, Phys<C>, CTiling, CLists);
      KernelPtr k = nullptr;
      // the plain variant (no faces) for steps that neither inject nor
      // write macro fields, with the Shan-Chen EOS
      const bool plain = !faces && !ph.inject_now && !ph.write_macro && !ph.eos.carn_star;
      // the plain kernels of 32-wide tiles (the default) with the width fixed
      const bool z32 = plain && R.TZi == 32;
      if (R.threads == 256) {
        k = faces ? k_compact<Mo, true, 256, kSR, kNS>
                  : (z32     ? k_compact<Mo, false, 256, kSR, kNS, true, 32>
                     : plain ? k_compact<Mo, false, 256, kSR, kNS, true>
                             : k_compact<Mo, false, 256, kSR, kNS>);
      } else if (R.threads == 512) {
        k = faces ? k_compact<Mo, true, 512, kSR, kNS>
                  : (z32     ? k_compact<Mo, false, 512, kSR, kNS, true, 32>
                     : plain ? k_compact<Mo, false, 512, kSR, kNS, true>
                             : k_compact<Mo, false, 512, kSR, kNS>);
      } else if constexpr (Mo::kStorage != MPLBM_FP64) {
        k = faces ? k_compact<Mo, true, 1024, kSR, kNS>
                  : (z32     ? k_compact<Mo, false, 1024, kSR, kNS, true, 32>
                     : plain ? k_compact<Mo, false, 1024, kSR, kNS, true>
                             : k_compact<Mo, false, 1024, kSR, kNS>);
      }
      if (!k) return;
      // dynamic LDS above 64 KiB (gfx950 gives a workgroup up to 160 KiB)
      if (shm > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
      hipLaunchKernelGGL(k, dim3(grid), dim3(R.threads), shm, ctx->stream, L, ph, R, WL);
    };
    auto launch = [&](const CTiling& R, bool faces) {
      if constexpr (Mo::kDev) {
        if (ph.sr) {
          if (ph.node_shift) launch_as(std::true_type{}, std::true_type{}, R, faces);
          else launch_as(std::true_type{}, std::false_type{}, R, faces);
          return;
        }
        if (ph.node_shift) {
          launch_as(std::false_type{}, std::true_type{}, R, faces);
          return;
        }
      }
      launch_as(std::false_type{}, std::false_type{}, R, faces);
    };
    const bool x_faces_only = ph.any_face && !(ph.face[2] | ph.face[3] | ph.face[4] | ph.face[5]);
    // Face rules and the skipped force only touch planes 0 and lx-1, so with
    // faces on x alone the step runs in two launches: end segments [0, e)
    // and [s*e, lx) in the face variant (segments 0 and s of length e), the
    // planes between in segments of the tiling's length in the plain
    // variant (fewer registers).  A middle segment [x0, x1) pulls planes
    // x0-1 .. x1 for its psi ring, which are no face planes when e >= 2 and
    // the last end holds at least two planes.  e is the first of 8 .. 2,
    // 9 .. 16 with s = ceil((lx-e)/e) leaving the last end 2 .. e planes and
    // the middle at least two; short ends keep the slower face variant to at
    // most 2e planes and its launch short (768x512x512 drainage FP32 3.85 ->
    // 3.71 ms, FP16S 2.90 -> 2.75 ms against end segments of the tiling's
    // 32 planes, same device).
    const int lx = (int)ctx->cfg.lx;
    int e = 0, s_end = 0;
    for (const int c : {8, 7, 6, 5, 4, 3, 2, 9, 10, 11, 12, 13, 14, 15, 16}) {
      if (!x_faces_only) break;
      const int s = (lx - 1) / c;  // ceil((lx - c) / c)
      if (lx - s * c >= 2 && (s - 1) * c >= 2) {
        e = c;
        s_end = s;
        break;
      }
    }
    if (e > 0) {
      const int ntile = T.tiles_y * T.tiles_z;
      CTiling mid = T, ends = T;
      mid.xoff = e;
      mid.xlim = s_end * e;
      mid.seg0 = 0;
      mid.seg_stride = 1;
      mid.nblocks = ntile * ((mid.xlim - e + T.seg - 1) / T.seg);
      ends.seg = e;
      ends.xoff = 0;
      ends.xlim = lx;
      ends.seg0 = 0;
      ends.seg_stride = s_end;
      ends.nblocks = ntile * 2;
      launch(mid, false);
      launch(ends, true);
    } else {
      launch(T, ph.any_face != 0);
    }
  } else if (ctx->kernel == MPLBM_KERNEL_FUSED) {
    const Tiling& T = ctx->tiling;
    const int threads = fused_threads(T);
    const int grid = T.xcd_remap ? ((T.nblocks + 7) / 8) * 8 : T.nblocks;
    const size_t shm = 4 * (size_t)T.TY * T.TZ * sizeof(C);
    // the launch bound sets the register budget: 1024-thread tiles cap a
    // thread at 128 VGPRs, 512 and 256 let the compiler keep every value in
    // registers at 3 waves per SIMD
    if (ph.any_face) {
      if (threads > 512)
        hipLaunchKernelGGL((k_fused<Mo, true, 1024>), dim3(grid), dim3(threads), shm, ctx->stream, L, ph, T);
      else if (threads > 256)
        hipLaunchKernelGGL((k_fused<Mo, true, 512>), dim3(grid), dim3(threads), shm, ctx->stream, L, ph, T);
      else
        hipLaunchKernelGGL((k_fused<Mo, true, 256>), dim3(grid), dim3(threads), shm, ctx->stream, L, ph, T);
    } else {
      if (threads > 512)
        hipLaunchKernelGGL((k_fused<Mo, false, 1024>), dim3(grid), dim3(threads), shm, ctx->stream, L, ph, T);
      else if (threads > 256)
        hipLaunchKernelGGL((k_fused<Mo, false, 512>), dim3(grid), dim3(threads), shm, ctx->stream, L, ph, T);
      else
        hipLaunchKernelGGL((k_fused<Mo, false, 256>), dim3(grid), dim3(threads), shm, ctx->stream, L, ph, T);
    }
  } else {
    const int grid = (int)((ctx->ncell + 255) / 256);
    hipLaunchKernelGGL(k_density<Mo>, dim3(grid), dim3(256), 0, ctx->stream, L, ph);
    hipLaunchKernelGGL(k_collide<Mo>, dim3(grid), dim3(256), 0, ctx->stream, L, ph);
  }
  HIP_TRY(ctx, hipGetLastError());
  ctx->cur = 1 - ctx->cur;
  return MPLBM_OK;
}

template <class F>
static mplbm_status dispatch(mplbm_ctx* ctx, F&& f) {
  switch (ctx->cfg.storage) {
    case MPLBM_FP32:
      if (ctx->cfg.math == MPLBM_MATH_FAST) return f(ModeF32Fast{});
      return f(ModeF32{});
    case MPLBM_FP64: return f(ModeF64{});
    case MPLBM_FP16S: return f(ModeH16{});
    case MPLBM_BF16S: return f(ModeB16{});
  }
  return fail(ctx, MPLBM_ERR_ARG, "bad storage");
}

extern "C" {

const char* mplbm_version(void) { return "mplbm 0.1 gfx950"; }

void mplbm_config_default(mplbm_config* c) {
  std::memset(c, 0, sizeof *c);
  c->lx = c->ly = c->lz = 400;
  c->storage = MPLBM_FP32;
  c->kernel = MPLBM_KERNEL_AUTO;
  c->tau = 1.0;
  c->G = -1.0;
  c->T_Tc = 0.7;
  c->rhol_spinodal = 0.2726;
  c->rhog_spinodal = 0.0484;
  c->rho_inject_period = 100;
  c->rho_increment = 0.0;
  c->rhos = 0.0;
  c->inject_position = 1.0;
  c->plane_pad = -1;
  c->bounce_back = MPLBM_BB_PULL;
}

const char* mplbm_last_error(const mplbm_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

mplbm_status mplbm_create(const mplbm_config* cfg, mplbm_ctx** out) {
  if (!cfg || !out) return fail(nullptr, MPLBM_ERR_ARG, "null argument");
  *out = nullptr;
  if (cfg->lx < 1 || cfg->ly < 1 || cfg->lz < 1) return fail(nullptr, MPLBM_ERR_ARG, "domain must be positive");
  const uint64_t n = (uint64_t)cfg->lx * cfg->ly * cfg->lz;
  if (n >= (1ull << 32)) return fail(nullptr, MPLBM_ERR_ARG, "domain exceeds 2^32 cells");
  if (cfg->storage < MPLBM_FP32 || cfg->storage > MPLBM_BF16S) return fail(nullptr, MPLBM_ERR_ARG, "bad storage");
  if (!(cfg->tau > 0.5)) return fail(nullptr, MPLBM_ERR_ARG, "tau must exceed 0.5");
  if (cfg->carn_star && cfg->G == 0.0) return fail(nullptr, MPLBM_ERR_ARG, "Carnahan-Starling psi needs G != 0");
  if (cfg->math != MPLBM_MATH_REFERENCE && cfg->math != MPLBM_MATH_FAST)
    return fail(nullptr, MPLBM_ERR_ARG, "bad math mode");
  if (cfg->math == MPLBM_MATH_FAST && cfg->storage == MPLBM_FP64)
    return fail(nullptr, MPLBM_ERR_UNSUPPORTED, "FP64 storage has reference arithmetic only");
  for (int a = 0; a < 3; ++a) {
    if ((cfg->face_type[2 * a] != 0) != (cfg->face_type[2 * a + 1] != 0))
      return fail(nullptr, MPLBM_ERR_ARG, "faces of one axis must both be periodic or both non-periodic");
  }
  for (int f = 0; f < 6; ++f)
    if (cfg->face_type[f] < 0 || cfg->face_type[f] > 1) return fail(nullptr, MPLBM_ERR_ARG, "bad face type");
  if (cfg->kernel < MPLBM_KERNEL_AUTO || cfg->kernel > MPLBM_KERNEL_COMPACT)
    return fail(nullptr, MPLBM_ERR_ARG, "bad kernel");
  if (cfg->rounding != MPLBM_ROUND_NEAREST && cfg->rounding != MPLBM_ROUND_STOCHASTIC)
    return fail(nullptr, MPLBM_ERR_ARG, "bad rounding mode");
  if (cfg->rounding == MPLBM_ROUND_STOCHASTIC &&
      ((cfg->storage != MPLBM_FP16S && cfg->storage != MPLBM_BF16S) ||
       (cfg->kernel != MPLBM_KERNEL_AUTO && cfg->kernel != MPLBM_KERNEL_COMPACT)))
    return fail(nullptr, MPLBM_ERR_UNSUPPORTED,
                "stochastic rounding applies to FP16S/BF16S storage with the compact kernel");
  if (cfg->shift != MPLBM_SHIFT_GLOBAL && cfg->shift != MPLBM_SHIFT_NODE)
    return fail(nullptr, MPLBM_ERR_ARG, "bad shift");
  if (cfg->shift == MPLBM_SHIFT_NODE &&
      ((cfg->storage != MPLBM_FP16S && cfg->storage != MPLBM_BF16S) ||
       (cfg->kernel != MPLBM_KERNEL_AUTO && cfg->kernel != MPLBM_KERNEL_COMPACT)))
    return fail(nullptr, MPLBM_ERR_UNSUPPORTED, "node shifts apply to FP16S/BF16S storage with the compact kernel");

  mplbm_ctx* ctx = new mplbm_ctx();
  ctx->cfg = *cfg;
  ctx->ncell = n;
  ctx->bytes_pop = (int)storage_bytes(cfg->storage);
  ctx->kernel = cfg->kernel == MPLBM_KERNEL_AUTO ? MPLBM_KERNEL_COMPACT : cfg->kernel;
  ctx->compact = ctx->kernel == MPLBM_KERNEL_COMPACT;
  {
    // sub-plane stride: ly*lz rounded up to 64 elements plus plane_pad
    // elements; a negative plane_pad picks 64 elements when the rounded plane
    // is a multiple of 4096 elements (power-of-two sizes map every
    // sub-plane of a cell to the same memory channel otherwise)
    const uint64_t plane = (uint64_t)cfg->ly * cfg->lz;
    uint64_t ps = (plane + 63) / 64 * 64;
    int64_t pad = cfg->plane_pad;
    if (pad < 0) pad = (ps % 4096 == 0) ? 64 : 0;
    ctx->ps = ps + (uint64_t)pad;
  }
  mplbm_status st = ctx->compact ? choose_ctiling(ctx) : choose_tiling(ctx);
  if (st != MPLBM_OK) {
    g_create_error = ctx->err;
    delete ctx;
    return st;
  }
  auto bail = [&](mplbm_status s) {
    g_create_error = ctx->err;
    mplbm_destroy(ctx);
    return s;
  };
#define CREATE_TRY(call)                                  \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) {                               \
      fail(ctx, e_ == hipErrorOutOfMemory ? MPLBM_ERR_NOMEM : MPLBM_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
      return bail(e_ == hipErrorOutOfMemory ? MPLBM_ERR_NOMEM : MPLBM_ERR_HIP); \
    }                                                     \
  } while (0)
  CREATE_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  CREATE_TRY(hipEventCreate(&ctx->ev0));
  CREATE_TRY(hipEventCreate(&ctx->ev1));
  // compact lattices are sized by the fluid count in mplbm_set_geometry
  const size_t lat_bytes = ctx->compact ? 0 : (size_t)cfg->lx * kQ * ctx->ps * ctx->bytes_pop;
  const size_t mb = cfg->storage == MPLBM_FP64 ? 8 : 4;  // macro / psi element size
  if (!ctx->compact) {
    CREATE_TRY(hipMalloc(&ctx->lat[0], lat_bytes));
    CREATE_TRY(hipMalloc(&ctx->lat[1], lat_bytes));
  }
  ctx->lat_bytes = lat_bytes;
  CREATE_TRY(hipMalloc((void**)&ctx->link, n * 4));
  CREATE_TRY(hipMalloc(&ctx->rho, n * mb));
  CREATE_TRY(hipMalloc(&ctx->pres, n * mb));
  CREATE_TRY(hipMalloc(&ctx->u, 3 * n * mb));
  CREATE_TRY(hipMalloc(&ctx->psi, n * mb));
  ctx->red_blocks = 1024;
  CREATE_TRY(hipMalloc((void**)&ctx->red_partial, (size_t)ctx->red_blocks * kRedFields * 8));
  CREATE_TRY(hipMalloc((void**)&ctx->red_out, kRedFields * 8));
  ctx->device_bytes = (int64_t)(2 * lat_bytes + n * 4 + 6 * n * mb + (size_t)ctx->red_blocks * kRedFields * 8);
#undef CREATE_TRY
  *out = ctx;
  return MPLBM_OK;
}

void mplbm_destroy(mplbm_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  void* bufs[] = {ctx->lat[0], ctx->lat[1], ctx->shift[0], ctx->shift[1], ctx->link, ctx->mask, ctx->rho, ctx->pres,
                  ctx->u,      ctx->psi,    ctx->red_partial, ctx->red_out, ctx->rw,       ctx->cpos,
                  ctx->ctab,   ctx->cl_parent, ctx->cl_size, ctx->cl_phase, ctx->cl_acc, ctx->cl_sid};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

mplbm_status mplbm_set_geometry(mplbm_ctx* ctx, const uint8_t* is_fluid) {
  if (!ctx || !is_fluid) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  const uint64_t n = ctx->ncell;
  if (!ctx->mask) HIP_TRY(ctx, hipMalloc((void**)&ctx->mask, n));
  std::vector<uint8_t> m(is_fluid, is_fluid + n);
  int64_t nf = 0;
  for (uint64_t c = 0; c < n; ++c) {
    m[c] = m[c] ? 1 : 0;
    nf += m[c];
  }
  HIP_TRY(ctx, hipMemcpyAsync(ctx->mask, m.data(), n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_links, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->mask, ctx->link,
                     (int)ctx->cfg.lx, (int)ctx->cfg.ly, (int)ctx->cfg.lz);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->fluid_nodes = nf;
  ctx->have_geometry = false;
  ctx->have_fields = false;
  ctx->hmask.swap(m);
  if (ctx->compact) {
    const mplbm_status st = build_compact(ctx);
    if (st != MPLBM_OK) return st;
  }
  ctx->have_geometry = true;
  return MPLBM_OK;
}

// Deviation storage: rho_ref defaults to the mean of the initial fluid
// density, the scale to a power of two that maps the largest expected
// deviation (|rho - rho_ref| up to max(rho_ref, 2) ) near 2^13 in FP16;
// BF16 has the float exponent range and needs no scale.
static void choose_deviation(mplbm_ctx* ctx, const double* rho) {
  const mplbm_config& c = ctx->cfg;
  if (c.storage != MPLBM_FP16S && c.storage != MPLBM_BF16S) {
    ctx->rho_ref = 0.f;
    ctx->scale = 1.f;
    return;
  }
  double ref = c.rho_ref;
  if (ref == 0.0) {
    double s = 0.0;
    for (uint64_t k = 0; k < ctx->ncell; ++k) s += rho[k];
    ref = s / (double)ctx->ncell;
  }
  ctx->rho_ref = (float)ref;
  double sc = c.dev_scale;
  if (sc == 0.0) {
    if (c.storage == MPLBM_BF16S) sc = 1.0;
    else {
      const double span = std::max(2.0, std::fabs(ref) + 2.0);
      sc = std::ldexp(1.0, (int)std::floor(std::log2(8192.0 / span)));
    }
  }
  ctx->scale = (float)sc;
}

mplbm_status mplbm_init_fields_f64(mplbm_ctx* ctx, const double* rho) {
  if (!ctx || !rho) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  if (!ctx->have_geometry) return fail(ctx, MPLBM_ERR_STATE, "mplbm_set_geometry must precede mplbm_init_fields");
  const uint64_t n = ctx->ncell;
  choose_deviation(ctx, rho);
  double* d_rho = nullptr;
  HIP_TRY(ctx, hipMalloc((void**)&d_rho, n * 8));
  hipError_t e = hipMemcpyAsync(d_rho, rho, n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    mplbm_status st = dispatch(ctx, [&](auto mo) -> mplbm_status {
      using Mo = decltype(mo);
      const auto ph = make_phys<typename Mo::C>(ctx, 0, 1, 0, false);
      auto L = make_lattice<Mo>(ctx, 0, 0);
      hipLaunchKernelGGL(k_init<Mo>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, L,
                         static_cast<typename Mo::S*>(ctx->lat[1]), static_cast<typename Mo::S*>(ctx->shift[1]), ph,
                         d_rho);
      if (!ctx->compact)
        hipLaunchKernelGGL(k_ghost<Mo>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, L);
      return MPLBM_OK;
    });
    (void)st;
    e = hipGetLastError();
  }
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_rho);
  if (e != hipSuccess) return fail(ctx, MPLBM_ERR_HIP, "init: %s", hipGetErrorString(e));
  if (e2 != hipSuccess) return fail(ctx, MPLBM_ERR_HIP, "init sync: %s", hipGetErrorString(e2));
  ctx->cur = 0;
  ctx->have_fields = true;
  return MPLBM_OK;
}

mplbm_status mplbm_init_fields(mplbm_ctx* ctx, const float* rho) {
  if (!ctx || !rho) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  std::vector<double> d(rho, rho + ctx->ncell);
  return mplbm_init_fields_f64(ctx, d.data());
}

mplbm_status mplbm_step(mplbm_ctx* ctx, int64_t step, int32_t reverse, uint32_t inject_type, int32_t nsteps,
                        uint32_t flags) {
  if (!ctx) return fail(ctx, MPLBM_ERR_ARG, "null context");
  if (!ctx->have_fields) return fail(ctx, MPLBM_ERR_STATE, "mplbm_init_fields must precede mplbm_step");
  if (nsteps < 0) return fail(ctx, MPLBM_ERR_ARG, "nsteps < 0");
  if (inject_type > 2) return fail(ctx, MPLBM_ERR_ARG, "inject_type must be 0, 1 or 2");
  for (int32_t s = 0; s < nsteps; ++s) {
    const bool macro = (flags & MPLBM_STEP_WRITE_MACRO) && s == nsteps - 1;
    mplbm_status st = dispatch(ctx, [&](auto mo) {
      return launch_step<decltype(mo)>(ctx, step + s, reverse, inject_type, macro);
    });
    if (st != MPLBM_OK) return st;
  }
  return MPLBM_OK;
}

mplbm_status mplbm_step_timed(mplbm_ctx* ctx, int64_t step, int32_t reverse, uint32_t inject_type,
                              int32_t nsteps, uint32_t flags, double* elapsed_ms) {
  if (!ctx) return fail(ctx, MPLBM_ERR_ARG, "null context");
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  mplbm_status st = mplbm_step(ctx, step, reverse, inject_type, nsteps, flags);
  if (st != MPLBM_OK) return st;
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  return MPLBM_OK;
}

mplbm_status mplbm_sync(mplbm_ctx* ctx) {
  if (!ctx) return fail(ctx, MPLBM_ERR_ARG, "null context");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MPLBM_OK;
}

}  // extern "C"

template <class T>
static mplbm_status read_fields_t(mplbm_ctx* ctx, T* rho, T* pres, T* u) {
  if (!ctx) return fail(ctx, MPLBM_ERR_ARG, "null context");
  if (!ctx->have_fields) return fail(ctx, MPLBM_ERR_STATE, "fields not initialised");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const uint64_t n = ctx->ncell;
  const bool f64 = ctx->cfg.storage == MPLBM_FP64;
  const size_t eb = f64 ? 8 : 4;
  std::vector<unsigned char> buf(3 * n * eb);
  std::vector<uint32_t> lk(n);
  HIP_TRY(ctx, hipMemcpy(lk.data(), ctx->link, n * 4, hipMemcpyDeviceToHost));
  auto conv = [&](const void* dev, T* dst, uint64_t count, bool is_u) -> mplbm_status {
    HIP_TRY(ctx, hipMemcpy(buf.data(), dev, count * eb, hipMemcpyDeviceToHost));
    for (uint64_t k = 0; k < count; ++k) {
      const uint64_t cell = is_u ? k % n : k;
      double v = f64 ? reinterpret_cast<const double*>(buf.data())[k] : reinterpret_cast<const float*>(buf.data())[k];
      (void)cell;
      if (is_u) {
        // device u is [3][cell]; host wants [cell][3]
        const uint64_t comp = k / n;
        dst[cell * 3 + comp] = (lk[cell] & 1u) ? (T)0 : (T)v;
      } else {
        dst[k] = (T)v;
      }
    }
    return MPLBM_OK;
  };
  mplbm_status st = MPLBM_OK;
  if (rho && (st = conv(ctx->rho, rho, n, false)) != MPLBM_OK) return st;
  if (pres && (st = conv(ctx->pres, pres, n, false)) != MPLBM_OK) return st;
  if (pres)
    for (uint64_t k = 0; k < n; ++k)
      if (lk[k] & 1u) pres[k] = (T)0;
  if (u && (st = conv(ctx->u, u, 3 * n, true)) != MPLBM_OK) return st;
  return MPLBM_OK;
}

extern "C" {

mplbm_status mplbm_read_fields(mplbm_ctx* ctx, float* rho, float* pressure, float* u) {
  return read_fields_t<float>(ctx, rho, pressure, u);
}
mplbm_status mplbm_read_fields_f64(mplbm_ctx* ctx, double* rho, double* pressure, double* u) {
  return read_fields_t<double>(ctx, rho, pressure, u);
}

mplbm_status mplbm_read_populations_f64(mplbm_ctx* ctx, double* f) {
  if (!ctx || !f) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  if (!ctx->have_fields) return fail(ctx, MPLBM_ERR_STATE, "fields not initialised");
  const uint64_t n = ctx->ncell;
  double* d = nullptr;
  HIP_TRY(ctx, hipMalloc((void**)&d, n * kQ * 8));
  mplbm_status st = dispatch(ctx, [&](auto mo) -> mplbm_status {
    using Mo = decltype(mo);
    const auto ph = make_phys<typename Mo::C>(ctx, 0, 1, 0, false);
    auto L = make_lattice<Mo>(ctx, ctx->cur, ctx->cur);
    hipLaunchKernelGGL(k_read_pops<Mo>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, L, ph, d);
    return MPLBM_OK;
  });
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(f, d, n * kQ * 8, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (st != MPLBM_OK) return st;
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(ctx, MPLBM_ERR_HIP, "read populations: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  return MPLBM_OK;
}

mplbm_status mplbm_write_populations_f64(mplbm_ctx* ctx, const double* f) {
  if (!ctx || !f) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  if (!ctx->have_fields) return fail(ctx, MPLBM_ERR_STATE, "fields not initialised");
  const uint64_t n = ctx->ncell;
  double* d = nullptr;
  HIP_TRY(ctx, hipMalloc((void**)&d, n * kQ * 8));
  hipError_t e = hipMemcpyAsync(d, f, n * kQ * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    dispatch(ctx, [&](auto mo) -> mplbm_status {
      using Mo = decltype(mo);
      const auto ph = make_phys<typename Mo::C>(ctx, 0, 1, 0, false);
      auto L = make_lattice<Mo>(ctx, ctx->cur, ctx->cur);
      hipLaunchKernelGGL(k_write_pops<Mo>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, L, ph, d);
      if (!ctx->compact)
        hipLaunchKernelGGL(k_ghost<Mo>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, L);
      return MPLBM_OK;
    });
    e = hipGetLastError();
  }
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess || e2 != hipSuccess)
    return fail(ctx, MPLBM_ERR_HIP, "write populations: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  return MPLBM_OK;
}

}  // extern "C"

// Checkpoint file: header, then plane by plane the stored values of the 19
// populations of each fluid cell in [y][z] order.
struct CheckpointHeader {
  char magic[8];          // "MPLBMCK1"
  uint32_t version;       // 1
  uint32_t storage;       // mplbm_storage
  int64_t lx, ly, lz;
  int64_t step;
  int64_t fluid_nodes;
  uint64_t mask_hash;     // FNV-1a of the fluid mask
  double rho_ref, dev_scale;
  uint32_t bytes_per_pop;
  uint32_t flags;         // kCheckpointNodeShift: node shifts follow the populations
};
constexpr uint32_t kCheckpointNodeShift = 1u;

static uint64_t mask_hash(const std::vector<uint8_t>& m) {
  uint64_t h = 1469598103934665603ull;
  for (uint8_t v : m) h = (h ^ v) * 1099511628211ull;
  return h;
}

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

// Plane by plane: the 19 stored values of every cell ([y][z][19]), or with
// shifts = true the node shifts ([y][z]), to or from host buffers.
template <class Fn>
static mplbm_status plane_raw_pass(mplbm_ctx* ctx, bool scatter, Fn&& per_plane, bool shifts = false) {
  const uint64_t plane = ctx->cfg.ly * ctx->cfg.lz;
  const size_t bytes = plane * (shifts ? 1 : kQ) * (size_t)ctx->bytes_pop;
  void* d = nullptr;
  HIP_TRY(ctx, hipMalloc(&d, bytes));
  std::vector<unsigned char> h(bytes);
  mplbm_status st = MPLBM_OK;
  for (int64_t x = 0; x < ctx->cfg.lx && st == MPLBM_OK; ++x) {
    if (scatter) {
      st = per_plane(x, h.data());
      if (st != MPLBM_OK) break;
      if (hipMemcpyAsync(d, h.data(), bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        st = fail(ctx, MPLBM_ERR_HIP, "checkpoint upload");
        break;
      }
    }
    st = dispatch(ctx, [&](auto mo) -> mplbm_status {
      using Mo = decltype(mo);
      auto L = make_lattice<Mo>(ctx, ctx->cur, ctx->cur);
      if (shifts)
        hipLaunchKernelGGL(k_plane_shift<Mo>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, ctx->stream, L,
                           (int)x, reinterpret_cast<typename Mo::S*>(d), scatter ? 1 : 0);
      else
        hipLaunchKernelGGL(k_plane_raw<Mo>, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, ctx->stream, L,
                           (int)x, reinterpret_cast<typename Mo::S*>(d), scatter ? 1 : 0);
      return MPLBM_OK;
    });
    if (st != MPLBM_OK) break;
    if (hipGetLastError() != hipSuccess) {
      st = fail(ctx, MPLBM_ERR_HIP, "checkpoint kernel");
      break;
    }
    if (!scatter) {
      if (hipMemcpyAsync(h.data(), d, bytes, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        st = fail(ctx, MPLBM_ERR_HIP, "checkpoint download");
        break;
      }
      st = per_plane(x, h.data());
    } else if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
      st = fail(ctx, MPLBM_ERR_HIP, "checkpoint scatter");
    }
  }
  (void)hipFree(d);
  return st;
}

extern "C" {

mplbm_status mplbm_save_checkpoint(mplbm_ctx* ctx, const char* path, int64_t step) {
  if (!ctx || !path) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  if (!ctx->have_fields) return fail(ctx, MPLBM_ERR_STATE, "fields not initialised");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  FileCloser fc{fopen(path, "wb")};
  if (!fc.f) return fail(ctx, MPLBM_ERR_ARG, "cannot open %s for writing", path);
  CheckpointHeader hd{};
  memcpy(hd.magic, "MPLBMCK1", 8);
  hd.version = 1;
  hd.storage = (uint32_t)ctx->cfg.storage;
  hd.lx = ctx->cfg.lx;
  hd.ly = ctx->cfg.ly;
  hd.lz = ctx->cfg.lz;
  hd.step = step;
  hd.fluid_nodes = ctx->fluid_nodes;
  hd.mask_hash = mask_hash(ctx->hmask);
  hd.rho_ref = ctx->rho_ref;
  hd.dev_scale = ctx->scale;
  hd.bytes_per_pop = (uint32_t)ctx->bytes_pop;
  const bool node_shift = ctx->cfg.shift == MPLBM_SHIFT_NODE;
  hd.flags = node_shift ? kCheckpointNodeShift : 0u;
  if (fwrite(&hd, sizeof hd, 1, fc.f) != 1) return fail(ctx, MPLBM_ERR_ARG, "write error on %s", path);
  const uint64_t plane = ctx->cfg.ly * ctx->cfg.lz;
  const size_t rec = kQ * (size_t)ctx->bytes_pop;
  mplbm_status st = plane_raw_pass(ctx, false, [&](int64_t x, unsigned char* h) -> mplbm_status {
    const uint8_t* m = ctx->hmask.data() + (size_t)x * plane;
    for (uint64_t r = 0; r < plane; ++r)
      if (m[r] && fwrite(h + r * rec, rec, 1, fc.f) != 1) return fail(ctx, MPLBM_ERR_ARG, "write error on %s", path);
    return MPLBM_OK;
  });
  if (st != MPLBM_OK || !node_shift) return st;
  // node-shifted form: then the shift of every fluid node, plane by plane
  const size_t srec = (size_t)ctx->bytes_pop;
  return plane_raw_pass(ctx, false, [&](int64_t x, unsigned char* h) -> mplbm_status {
    const uint8_t* m = ctx->hmask.data() + (size_t)x * plane;
    for (uint64_t r = 0; r < plane; ++r)
      if (m[r] && fwrite(h + r * srec, srec, 1, fc.f) != 1) return fail(ctx, MPLBM_ERR_ARG, "write error on %s", path);
    return MPLBM_OK;
  }, true);
}

mplbm_status mplbm_load_checkpoint(mplbm_ctx* ctx, const char* path, int64_t* step) {
  if (!ctx || !path) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  if (!ctx->have_geometry) return fail(ctx, MPLBM_ERR_STATE, "geometry not set");
  FileCloser fc{fopen(path, "rb")};
  if (!fc.f) return fail(ctx, MPLBM_ERR_ARG, "cannot open %s", path);
  CheckpointHeader hd{};
  if (fread(&hd, sizeof hd, 1, fc.f) != 1 || memcmp(hd.magic, "MPLBMCK1", 8) != 0 || hd.version != 1)
    return fail(ctx, MPLBM_ERR_ARG, "%s is not a checkpoint of this library", path);
  if (hd.lx != ctx->cfg.lx || hd.ly != ctx->cfg.ly || hd.lz != ctx->cfg.lz)
    return fail(ctx, MPLBM_ERR_ARG, "checkpoint domain %lldx%lldx%lld differs from the context's",
                (long long)hd.lx, (long long)hd.ly, (long long)hd.lz);
  if ((int)hd.storage != ctx->cfg.storage || (int)hd.bytes_per_pop != ctx->bytes_pop)
    return fail(ctx, MPLBM_ERR_ARG, "checkpoint storage %u differs from the context's", hd.storage);
  if (hd.fluid_nodes != ctx->fluid_nodes || hd.mask_hash != mask_hash(ctx->hmask))
    return fail(ctx, MPLBM_ERR_ARG, "checkpoint geometry differs from the context's");
  const bool node_shift = ctx->cfg.shift == MPLBM_SHIFT_NODE;
  if (((hd.flags & kCheckpointNodeShift) != 0) != node_shift)
    return fail(ctx, MPLBM_ERR_ARG, "checkpoint shift form (%s) differs from the context's",
                (hd.flags & kCheckpointNodeShift) ? "node" : "global");
  // the stored deviations are relative to the file's reference density and scale
  ctx->rho_ref = (float)hd.rho_ref;
  ctx->scale = (float)hd.dev_scale;
  const uint64_t plane = ctx->cfg.ly * ctx->cfg.lz;
  const size_t rec = kQ * (size_t)ctx->bytes_pop;
  mplbm_status st = plane_raw_pass(ctx, true, [&](int64_t x, unsigned char* h) -> mplbm_status {
    const uint8_t* m = ctx->hmask.data() + (size_t)x * plane;
    for (uint64_t r = 0; r < plane; ++r)
      if (m[r] && fread(h + r * rec, rec, 1, fc.f) != 1) return fail(ctx, MPLBM_ERR_ARG, "%s is truncated", path);
    return MPLBM_OK;
  });
  if (st != MPLBM_OK) return st;
  if (node_shift) {
    const size_t srec = (size_t)ctx->bytes_pop;
    st = plane_raw_pass(ctx, true, [&](int64_t x, unsigned char* h) -> mplbm_status {
      const uint8_t* m = ctx->hmask.data() + (size_t)x * plane;
      std::memset(h, 0, plane * srec);
      for (uint64_t r = 0; r < plane; ++r)
        if (m[r] && fread(h + r * srec, srec, 1, fc.f) != 1) return fail(ctx, MPLBM_ERR_ARG, "%s is truncated", path);
      return MPLBM_OK;
    }, true);
    if (st != MPLBM_OK) return st;
  }
  // bounce-back slots of the full layout, then every macro field and psi
  // (the two-pass kernel reads psi of solid neighbours; a context that was
  // never initialised holds no fields yet)
  st = dispatch(ctx, [&](auto mo) -> mplbm_status {
    using Mo = decltype(mo);
    const auto ph = make_phys<typename Mo::C>(ctx, 0, 1, 0, false);
    auto L = make_lattice<Mo>(ctx, ctx->cur, ctx->cur);
    const dim3 grid((unsigned)((ctx->ncell + 255) / 256));
    if (!ctx->compact) hipLaunchKernelGGL(k_ghost<Mo>, grid, dim3(256), 0, ctx->stream, L);
    hipLaunchKernelGGL(k_macro_from_pops<Mo>, grid, dim3(256), 0, ctx->stream, L, ph);
    return MPLBM_OK;
  });
  if (st != MPLBM_OK) return st;
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->have_fields = true;
  if (step) *step = hd.step;
  return MPLBM_OK;
}

mplbm_status mplbm_reduce(mplbm_ctx* ctx, mplbm_stats64* out) {
  if (!ctx || !out) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  if (!ctx->have_fields) return fail(ctx, MPLBM_ERR_STATE, "fields not initialised");
  mplbm_status st = dispatch(ctx, [&](auto mo) -> mplbm_status {
    using Mo = decltype(mo);
    const auto ph = make_phys<typename Mo::C>(ctx, 0, 1, 0, false);
    auto L = make_lattice<Mo>(ctx, ctx->cur, ctx->cur);
    hipLaunchKernelGGL(k_reduce<Mo>, dim3(ctx->red_blocks), dim3(kRedThreads), 0, ctx->stream, L, ph,
                       ctx->cfg.rhol_spinodal, ctx->cfg.rhog_spinodal, ctx->red_partial);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(64), 0, ctx->stream, ctx->red_partial, ctx->red_blocks,
                       ctx->red_out);
    return MPLBM_OK;
  });
  if (st != MPLBM_OK) return st;
  HIP_TRY(ctx, hipGetLastError());
  double h[kRedFields];
  HIP_TRY(ctx, hipMemcpyAsync(h, ctx->red_out, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  out->mass = h[0];
  out->momentum[0] = h[1];
  out->momentum[1] = h[2];
  out->momentum[2] = h[3];
  out->fluid_nodes = h[4];
  out->liquid_nodes = h[5];
  out->vapor_nodes = h[6];
  out->liquid_rho_sum = h[7];
  out->vapor_rho_sum = h[8];
  out->liquid_p_sum = h[9];
  out->vapor_p_sum = h[10];
  out->rho_min = h[11];
  out->rho_max = h[12];
  out->p_min = h[13];
  out->p_max = h[14];
  out->nonfinite = h[15];
  return MPLBM_OK;
}

mplbm_status mplbm_set_tiling(mplbm_ctx* ctx, int32_t tile_y, int32_t tile_z, int32_t seg_x) {
  if (!ctx) return fail(ctx, MPLBM_ERR_ARG, "null context");
  const mplbm_config saved = ctx->cfg;
  const Tiling saved_t = ctx->tiling;
  const CTiling saved_c = ctx->ctiling;
  ctx->cfg.tile_y = tile_y;
  ctx->cfg.tile_z = tile_z;
  ctx->cfg.seg_x = seg_x;
  mplbm_status st = ctx->compact ? choose_ctiling(ctx) : choose_tiling(ctx);
  if (st == MPLBM_OK && ctx->compact && ctx->have_geometry) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    st = build_lists(ctx);
  }
  if (st != MPLBM_OK) {
    ctx->cfg = saved;
    ctx->tiling = saved_t;
    ctx->ctiling = saved_c;
    if (ctx->compact && ctx->have_geometry) (void)build_lists(ctx);
  }
  return st;
}

mplbm_status mplbm_info_get(const mplbm_ctx* ctx, mplbm_info* out) {
  if (!ctx || !out) return MPLBM_ERR_ARG;
  out->fluid_nodes = ctx->fluid_nodes;
  out->cells = (int64_t)ctx->ncell;
  out->storage = ctx->cfg.storage;
  out->kernel = ctx->kernel;
  out->tile_y = ctx->compact ? ctx->ctiling.TYi : ctx->tiling.TY;
  out->tile_z = ctx->compact ? ctx->ctiling.TZi : ctx->tiling.TZ;
  out->seg_x = ctx->compact ? ctx->ctiling.seg : ctx->tiling.seg;
  out->bytes_per_pop = ctx->bytes_pop;
  out->rho_ref = ctx->rho_ref;
  out->dev_scale = ctx->scale;
  out->device_bytes = ctx->device_bytes;
  return MPLBM_OK;
}

}  // extern "C"

// ============================================================================
// reference-named entry points (include/iteration_host_compat.h)

#include "iteration_host_compat.h"

namespace {
struct Compat {
  mplbm_config cfg{};
  bool have_domain = false, have_params = false;
  int abort_on_error = 1;
  int status = MPLBM_OK;
  mplbm_ctx* ctx = nullptr;
  void* rho = nullptr;   // the caller's density and pressure arrays (any memory kind)
  void* pres = nullptr;
  std::vector<float> hrho, hpres;
};
Compat g_compat;

// The reference's cudaErrorCheck (include/iteration_host.h:6-15): print the
// error, then exit with its code unless the caller asked to continue.
// src: the context whose last error describes the failure, or null for
// failures recorded without one.
void compat_report(mplbm_status st, const mplbm_ctx* src = nullptr) {
  g_compat.status = st;
  if (st == MPLBM_OK) return;
  fprintf(stderr, "mplbm compat error %d: %s\n", (int)st, mplbm_last_error(src));
  if (g_compat.abort_on_error) exit((int)st);
}

// Copies density and pressure from the context into the caller's arrays.
mplbm_status compat_write_back(mplbm_ctx* ctx) {
  const uint64_t n = ctx->ncell;
  g_compat.hrho.resize(n);
  g_compat.hpres.resize(n);
  mplbm_status st = mplbm_read_fields(ctx, g_compat.hrho.data(), g_compat.hpres.data(), nullptr);
  if (st != MPLBM_OK) return st;
  HIP_TRY(ctx, hipMemcpy(g_compat.rho, g_compat.hrho.data(), n * sizeof(float), hipMemcpyDefault));
  HIP_TRY(ctx, hipMemcpy(g_compat.pres, g_compat.hpres.data(), n * sizeof(float), hipMemcpyDefault));
  return MPLBM_OK;
}

mplbm_status compat_initialization(void* rho_dd, void* pressure_dd, void* is_fluid_dd) {
  if (!rho_dd || !pressure_dd || !is_fluid_dd)
    return fail(nullptr, MPLBM_ERR_ARG, "multiphaselbm_initialization: null rho, pressure or is_fluid address");
  void* rho = *static_cast<void**>(rho_dd);
  void* pres = *static_cast<void**>(pressure_dd);
  const void* fluid = *static_cast<void**>(is_fluid_dd);
  if (!rho || !pres || !fluid)
    return fail(nullptr, MPLBM_ERR_ARG, "multiphaselbm_initialization: null rho, pressure or is_fluid array");
  if (!g_compat.have_params) return fail(nullptr, MPLBM_ERR_STATE, "transfer_parameter must come first");
  mplbm_compat_release();
  mplbm_ctx* ctx = nullptr;
  mplbm_status st = mplbm_create(&g_compat.cfg, &ctx);
  if (st != MPLBM_OK) return st;
  const uint64_t n = ctx->ncell;
  auto bail = [&](mplbm_status s) {
    g_create_error = ctx->err;
    mplbm_destroy(ctx);
    return s;
  };
  static_assert(sizeof(bool) == 1, "the reference's Boolxyz mask is one byte per cell");
  std::vector<uint8_t> m(n);
  std::vector<float> r(n);
  if (hipMemcpy(m.data(), fluid, n, hipMemcpyDefault) != hipSuccess ||
      hipMemcpy(r.data(), rho, n * sizeof(float), hipMemcpyDefault) != hipSuccess) {
    fail(ctx, MPLBM_ERR_HIP, "multiphaselbm_initialization: cannot read the caller's rho or is_fluid");
    return bail(MPLBM_ERR_HIP);
  }
  for (uint64_t k = 0; k < n; ++k) m[k] = m[k] ? 1 : 0;
  if ((st = mplbm_set_geometry(ctx, m.data())) != MPLBM_OK || (st = mplbm_init_fields(ctx, r.data())) != MPLBM_OK)
    return bail(st);
  g_compat.ctx = ctx;
  g_compat.rho = rho;
  g_compat.pres = pres;
  if ((st = compat_write_back(ctx)) != MPLBM_OK) {
    g_compat.ctx = nullptr;
    return bail(st);
  }
  return MPLBM_OK;
}
}  // namespace

extern "C" {

int mplbm_compat_set_domain(int64_t lx, int64_t ly, int64_t lz, int32_t storage) {
  if (!g_compat.have_params) mplbm_config_default(&g_compat.cfg);
  g_compat.cfg.lx = lx;
  g_compat.cfg.ly = ly;
  g_compat.cfg.lz = lz;
  g_compat.cfg.storage = storage;
  g_compat.have_domain = true;
  return MPLBM_OK;
}

void transfer_parameter(const float* tau, const float* G, const bool* carn_star, const float* T_Tc,
                        const float* rhol_spinodal, const float* rhog_spinodal, const bool* inject_density,
                        const size_t* rho_inject_period, const float* rho_increment, const float* rhos,
                        const float* inject_position, const size_t* statPeriod) {
  if (!tau || !G || !carn_star || !T_Tc || !rhol_spinodal || !rhog_spinodal || !inject_density ||
      !rho_inject_period || !rho_increment || !rhos || !inject_position || !statPeriod) {
    compat_report(fail(nullptr, MPLBM_ERR_ARG, "transfer_parameter: null argument"));
    return;
  }
  if (!g_compat.have_domain) {
    mplbm_config_default(&g_compat.cfg);
    g_compat.have_domain = true;
  }
  mplbm_config& c = g_compat.cfg;
  c.tau = *tau;
  c.G = *G;
  c.carn_star = *carn_star ? 1 : 0;
  c.T_Tc = *T_Tc;
  c.rhol_spinodal = *rhol_spinodal;
  c.rhog_spinodal = *rhog_spinodal;
  // the reference copies inject_density but never reads it on the device
  (void)inject_density;
  c.rho_inject_period = (int64_t)*rho_inject_period;
  c.rho_increment = *rho_increment;
  c.rhos = *rhos;
  c.inject_position = *inject_position;
  (void)statPeriod;  // copied to the device but unused there (iteration_host.cu:39)
  g_compat.have_params = true;
  compat_report(MPLBM_OK);
}

void multiphaselbm_initialization(void* rho_dd, void* psi_dd, void* pressure_dd, void* is_fluid_dd,
                                  void* collision_f_dd, void* stream_f_dd, void* force_dd) {
  // psi, both lattices and the force are owned by the library
  (void)psi_dd;
  (void)collision_f_dd;
  (void)stream_f_dd;
  (void)force_dd;
  compat_report(compat_initialization(rho_dd, pressure_dd, is_fluid_dd));
}

void multiphaselbm_iteration(long step, size_t statPeriod, int reverse, unsigned int inject_type) {
  mplbm_ctx* ctx = g_compat.ctx;
  if (!ctx) {
    compat_report(fail(nullptr, MPLBM_ERR_STATE, "multiphaselbm_initialization must come first"));
    return;
  }
  const bool stat = statPeriod > 0 && (uint64_t)step % statPeriod == 0;
  mplbm_status st = mplbm_step(ctx, step, reverse, inject_type, 1, stat ? MPLBM_STEP_WRITE_MACRO : 0);
  if (st == MPLBM_OK && stat) st = compat_write_back(ctx);
  compat_report(st, ctx);
}

void mplbm_compat_set_abort(int abort_on_error) { g_compat.abort_on_error = abort_on_error ? 1 : 0; }

int mplbm_compat_status(void) { return g_compat.status; }

void mplbm_compat_release(void) {
  if (g_compat.ctx) mplbm_destroy(g_compat.ctx);
  g_compat.ctx = nullptr;
  g_compat.rho = g_compat.pres = nullptr;
}

}  // extern "C"

// ============================================================================
// batched ensemble (include/mplbm_ensemble.h)

#include "mplbm_ensemble.h"

namespace mplbm {

struct EnsItem {
  int sample, seg, tyi, tzi;
};

template <class Mo>
struct EnsSample {
  Lattice<Mo> L[2];  // [parity]: L[0] reads the lattice current at table build
  Phys<typename Mo::C> ph;
  Tiling T;
};

// Work queue of a persistent ensemble launch: the next item to take and the
// number of workgroups that found the queue empty.  Items differ in cost
// (partial tiles, short segments, porosity), so workgroups take them one at
// a time instead of a fixed stride.  The last workgroup to finish clears
// both counters for the next launch, which keeps a captured graph free of
// memset nodes.
struct EnsQueue {
  uint32_t next;
  uint32_t done;
};

// Index of the workgroup's next item (the same value in every thread).
__device__ __forceinline__ int ens_take(EnsQueue* q, int* slot) {
  if (threadIdx.x == 0) *slot = (int)atomicAdd(&q->next, 1u);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*slot);
}

// After the workgroup's last take: the last workgroup clears the queue.
__device__ __forceinline__ void ens_finish(EnsQueue* q) {
  if (threadIdx.x == 0) {
    __threadfence();  // this workgroup's last take precedes its count
    if (atomicAdd(&q->done, 1u) == gridDim.x - 1) {
      atomicExch(&q->next, 0u);
      atomicExch(&q->done, 0u);
    }
  }
}

// Persistent launch over the flattened work list; every workgroup takes
// items from the work queue until it is empty.  Sample descriptors are read
// with wave-uniform addresses (scalar loads).
template <class Mo, int kBlock>
__global__ __launch_bounds__(kBlock) void k_ensemble(const EnsItem* __restrict__ items, int nitems,
                                                     const EnsSample<Mo>* __restrict__ samples, int parity,
                                                     int macro, EnsQueue* queue) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ int slot;
  for (;;) {
    const int w = ens_take(queue, &slot);
    if (w >= nitems) break;
    const EnsItem it = items[w];
    const int s = __builtin_amdgcn_readfirstlane(it.sample);
    const EnsSample<Mo>& sd = samples[s];
    Phys<typename Mo::C> ph = sd.ph;
    ph.write_macro = macro;
    march_tile<Mo, false>(sd.L[parity], ph, sd.T, __builtin_amdgcn_readfirstlane(it.seg),
                          __builtin_amdgcn_readfirstlane(it.tyi), __builtin_amdgcn_readfirstlane(it.tzi),
                          reinterpret_cast<typename Mo::C*>(smem_raw));
    __syncthreads();  // the next item reuses the LDS ring
  }
  ens_finish(queue);
}

template <class Mo>
struct EnsSampleC {
  Lattice<Mo> L[2];  // [parity], as EnsSample
  Phys<typename Mo::C> ph;
  CTiling T;
  CLists W;
};

// Compact form of the persistent launch: each item is a compact march of
// one sample's tile and x segment.  The 16-bit forms are held to 80 VGPRs
// (6 waves per SIMD, as the single-domain kernel): left to itself the
// compiler takes 88 (5 waves), and the 12 bytes per thread it spills at 80
// cost less than the wave gains (64 packings: FP16S 1.64 -> 1.53 ms per
// step; FP32 FAST held to 96 was slower, 1.695 -> 1.709 ms).
template <class Mo, bool kPlain>
constexpr int ensemble_min_waves() {
  return (sizeof(typename Mo::S) == 2 || (kPlain && Mo::kStorage == MPLBM_FP32)) ? 6 : 1;
}
// kPlain as in compact_march: launches without macro writes when no sample
// injects or uses the Carnahan-Starling EOS; kTZ as in compact_march, when
// every sample's tile is that wide
template <class Mo, int kBlock, bool kPlain = false, int kTZ = 0>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ensemble_min_waves<Mo, kPlain>())))
void k_ensemble_c(const EnsItem* __restrict__ items, int nitems,
                                                       const EnsSampleC<Mo>* __restrict__ samples, int parity,
                                                       int macro, EnsQueue* queue) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ int slot;
  for (;;) {
    const int w = ens_take(queue, &slot);
    if (w >= nitems) break;
    const EnsItem it = items[w];
    const int s = __builtin_amdgcn_readfirstlane(it.sample);
    const EnsSampleC<Mo>& sd = samples[s];
    Phys<typename Mo::C> ph = sd.ph;
    ph.write_macro = macro;
    const int tile = __builtin_amdgcn_readfirstlane(it.tyi * sd.T.tiles_z + it.tzi);
    compact_march<Mo, false, kBlock, false, false, kPlain, 2, kTZ>(sd.L[parity], ph, sd.T, sd.W,
                                                           __builtin_amdgcn_readfirstlane(it.seg), tile, smem_raw);
    __syncthreads();  // the next item reuses the LDS regions
  }
  ens_finish(queue);
}

}  // namespace mplbm

struct mplbm_ensemble {
  std::vector<mplbm_ctx*> ctx;
  bool compact = true;  // fluid-node layout and compact marching
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int block = 512;
  int storage = MPLBM_FP32, math = MPLBM_MATH_REFERENCE;
  std::vector<EnsItem> items;
  EnsItem* d_items = nullptr;
  void* d_samples = nullptr;
  EnsQueue* d_queue = nullptr;  // cleared at creation, then by each launch's last workgroup
  size_t sample_bytes = 0;
  size_t lds_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int64_t fluid_nodes = 0;
  // per-sample compact tile {tile_y, tile_z, seg_x} set by
  // mplbm_ensemble_set_tiling; all zero = fitted to the block size
  std::vector<std::array<int32_t, 3>> tile_req;
  std::string err;
};

#define ENS_HIP(e, call)                                                                          \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return ens_fail((e), e_ == hipErrorOutOfMemory ? MPLBM_ERR_NOMEM : MPLBM_ERR_HIP, "%s: %s", #call, \
                      hipGetErrorString(e_));                                                     \
  } while (0)

static mplbm_status ens_fail(mplbm_ensemble* e, mplbm_status st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  else g_create_error = buf;
  return st;
}

// Tile of one sample for a workgroup of `block` threads: an axis that fits
// wraps inside the tile, otherwise a one-cell ring.
static void ens_fit_tile(const mplbm_config& c, int block, int* ty, int* tz, int* seg) {
  int z = c.lz <= 64 ? (int)c.lz : (block >= 512 ? 66 : 34);
  if (z > block / 3) z = 34;
  int y = block / z;
  if (y >= c.ly && (int64_t)z * c.ly <= block) y = (int)c.ly;
  if (y > c.ly) y = (int)c.ly;
  *ty = y;
  *tz = z;
  *seg = (int)std::min<int64_t>(32, c.lx);
}

static mplbm_status ens_plan(mplbm_ensemble* e) {
  e->items.clear();
  e->lds_bytes = 0;
  const size_t csize = (e->storage == MPLBM_FP64) ? 8 : 4;
  for (size_t s = 0; s < e->ctx.size(); ++s) {
    mplbm_ctx* c = e->ctx[s];
    if (e->compact) {
      // interior tile of at most `block` nodes, 32 wide in z where the axis
      // allows; narrower in y until the LDS regions fit.  Fields set by
      // mplbm_ensemble_set_tiling replace the fitted ones (one attempt).
      const std::array<int32_t, 3> r = s < e->tile_req.size() ? e->tile_req[s] : std::array<int32_t, 3>{0, 0, 0};
      const bool asked = r[0] | r[1] | r[2];
      const int tz = r[1] ? r[1] : (int)std::min<int64_t>(32, c->cfg.lz);
      int ty = r[0] ? r[0] : (int)std::min<int64_t>(c->cfg.ly, std::max(1, std::min(e->block / tz, 16)));
      const int seg = r[2] ? r[2] : (int)std::min<int64_t>(32, c->cfg.lx);
      mplbm_status st = MPLBM_ERR_ARG;
      for (; ty >= 1; --ty) {
        if (std::min<int64_t>(ty, c->cfg.ly) * std::min<int64_t>(tz, c->cfg.lz) > e->block) {
          if (asked) {
            st = fail(c, MPLBM_ERR_ARG, "tile %dx%d exceeds the %d threads of a workgroup", ty, tz, e->block);
            break;
          }
          continue;
        }
        if ((st = mplbm_set_tiling(c, ty, tz, seg)) == MPLBM_OK || asked) break;
      }
      if (st != MPLBM_OK) return ens_fail(e, st, "sample %zu: %s", s, c->err.c_str());
      e->lds_bytes = std::max(e->lds_bytes, csize == 8 ? compact_lds_bytes<double>(c->ctiling, 0, e->block)
                                                       : compact_lds_bytes<float>(c->ctiling, 0, e->block));
      continue;
    }
    int ty, tz, seg;
    ens_fit_tile(c->cfg, e->block, &ty, &tz, &seg);
    mplbm_status st = mplbm_set_tiling(c, ty, tz, seg);
    if (st != MPLBM_OK) return ens_fail(e, st, "sample %zu: %s", s, c->err.c_str());
    const Tiling& T = c->tiling;
    if (T.TY * T.TZ > e->block)
      return ens_fail(e, MPLBM_ERR_ARG, "sample %zu: tile %dx%d exceeds %d threads", s, T.TY, T.TZ, e->block);
    e->lds_bytes = std::max(e->lds_bytes, 4 * (size_t)T.TY * T.TZ * csize);
  }
  // items ordered by segment, then sample, then tile: consecutive workgroups
  // march neighbouring tiles of one sample, so ring re-reads meet in L2
  auto nseg_of = [&](mplbm_ctx* c) { return e->compact ? c->ctiling.nseg : c->tiling.nseg; };
  int max_seg = 0;
  for (mplbm_ctx* c : e->ctx) max_seg = std::max(max_seg, nseg_of(c));
  for (int g = 0; g < max_seg; ++g)
    for (size_t s = 0; s < e->ctx.size(); ++s) {
      mplbm_ctx* c = e->ctx[s];
      if (g >= nseg_of(c)) continue;
      const int ty_n = e->compact ? c->ctiling.tiles_y : c->tiling.tiles_y;
      const int tz_n = e->compact ? c->ctiling.tiles_z : c->tiling.tiles_z;
      for (int a = 0; a < ty_n; ++a)
        for (int b = 0; b < tz_n; ++b) e->items.push_back(EnsItem{(int)s, g, a, b});
    }
  if (e->d_items) (void)hipFree(e->d_items);
  e->d_items = nullptr;
  ENS_HIP(e, hipMalloc((void**)&e->d_items, e->items.size() * sizeof(EnsItem)));
  ENS_HIP(e, hipMemcpy(e->d_items, e->items.data(), e->items.size() * sizeof(EnsItem), hipMemcpyHostToDevice));
  if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->graph_exec = nullptr;
  e->graph = nullptr;
  return MPLBM_OK;
}

template <class F>
static mplbm_status ens_dispatch(mplbm_ensemble* e, F&& f) {
  switch (e->storage) {
    case MPLBM_FP32:
      if (e->math == MPLBM_MATH_FAST) return f(ModeF32Fast{});
      return f(ModeF32{});
    case MPLBM_FP64: return f(ModeF64{});
    case MPLBM_FP16S: return f(ModeH16{});
    case MPLBM_BF16S: return f(ModeB16{});
  }
  return ens_fail(e, MPLBM_ERR_ARG, "bad storage");
}

// Upload the per-sample descriptor table for the lattices current now;
// inject_now follows `step`.
template <class Mo, class Tab>
static void ens_fill(mplbm_ensemble* e, int64_t step, Tab& tab) {
  for (size_t s = 0; s < e->ctx.size(); ++s) {
    mplbm_ctx* c = e->ctx[s];
    tab[s].L[0] = make_lattice<Mo>(c, c->cur, 1 - c->cur);
    tab[s].L[1] = make_lattice<Mo>(c, 1 - c->cur, c->cur);
    tab[s].ph = make_phys<typename Mo::C>(c, step, 1, 0, false);
  }
}

static mplbm_status ens_upload(mplbm_ensemble* e, int64_t step) {
  return ens_dispatch(e, [&](auto mo) -> mplbm_status {
    using Mo = decltype(mo);
    std::vector<EnsSample<Mo>> tab;
    std::vector<EnsSampleC<Mo>> tabc;
    const void* src = nullptr;
    size_t bytes = 0;
    if (e->compact) {
      tabc.resize(e->ctx.size());
      ens_fill<Mo>(e, step, tabc);
      for (size_t s = 0; s < e->ctx.size(); ++s) {
        tabc[s].T = e->ctx[s]->ctiling;
        tabc[s].W = CLists{e->ctx[s]->cpos, e->ctx[s]->ctab};
      }
      src = tabc.data();
      bytes = tabc.size() * sizeof(EnsSampleC<Mo>);
    } else {
      tab.resize(e->ctx.size());
      ens_fill<Mo>(e, step, tab);
      for (size_t s = 0; s < e->ctx.size(); ++s) tab[s].T = e->ctx[s]->tiling;
      src = tab.data();
      bytes = tab.size() * sizeof(EnsSample<Mo>);
    }
    if (e->sample_bytes != bytes) {
      if (e->d_samples) (void)hipFree(e->d_samples);
      e->d_samples = nullptr;
      ENS_HIP(e, hipMalloc(&e->d_samples, bytes));
      e->sample_bytes = bytes;
      if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
      if (e->graph) (void)hipGraphDestroy(e->graph);
      e->graph_exec = nullptr;
      e->graph = nullptr;
    }
    // stream-ordered after earlier steps; synchronous so the table can go away
    ENS_HIP(e, hipStreamSynchronize(e->stream));
    ENS_HIP(e, hipMemcpy(e->d_samples, src, bytes, hipMemcpyHostToDevice));
    return MPLBM_OK;
  });
}

static mplbm_status ens_launch(mplbm_ensemble* e, int parity, int macro) {
  return ens_dispatch(e, [&](auto mo) -> mplbm_status {
    using Mo = decltype(mo);
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int nitems = (int)e->items.size();
    const int per_cu = e->block >= 1024 ? 2 : (e->block >= 512 ? 3 : 6);
    const int grid = std::max(1, std::min(nitems, ncu * per_cu));
    if (e->compact) {
      const auto* sc = static_cast<const EnsSampleC<Mo>*>(e->d_samples);
      // the plain kernel when this launch writes no macro fields and no
      // sample injects (a zero increment adds nothing) or uses the
      // Carnahan-Starling EOS
      bool plain = !macro;
      for (const mplbm_ctx* c : e->ctx)
        plain = plain && !c->cfg.carn_star && !(c->cfg.rho_inject_period > 0 && c->cfg.rho_increment != 0.0);
      auto go = [&](auto kernel, int block) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), e->lds_bytes, e->stream, e->d_items, nitems, sc, parity,
                           macro, e->d_queue);
      };
      // every sample's tile 32 nodes wide: the plain kernel with that width
      bool z32 = plain;
      for (const mplbm_ctx* c : e->ctx) z32 = z32 && c->ctiling.TZi == 32;
      if (e->block == 512) {
        if (z32) go(k_ensemble_c<Mo, 512, true, 32>, 512);
        else if (plain) go(k_ensemble_c<Mo, 512, true>, 512);
        else go(k_ensemble_c<Mo, 512>, 512);
      } else if (e->block == 256) {
        if (z32) go(k_ensemble_c<Mo, 256, true, 32>, 256);
        else if (plain) go(k_ensemble_c<Mo, 256, true>, 256);
        else go(k_ensemble_c<Mo, 256>, 256);
      } else if constexpr (Mo::kStorage != MPLBM_FP64) {
        if (z32) go(k_ensemble_c<Mo, 1024, true, 32>, 1024);
        else if (plain) go(k_ensemble_c<Mo, 1024, true>, 1024);
        else go(k_ensemble_c<Mo, 1024>, 1024);
      }
      ENS_HIP(e, hipGetLastError());
      return MPLBM_OK;
    }
    const auto* samp = static_cast<const EnsSample<Mo>*>(e->d_samples);
    if (e->block == 1024)
      hipLaunchKernelGGL((k_ensemble<Mo, 1024>), dim3(grid), dim3(1024), e->lds_bytes, e->stream, e->d_items,
                         nitems, samp, parity, macro, e->d_queue);
    else if (e->block == 512)
      hipLaunchKernelGGL((k_ensemble<Mo, 512>), dim3(grid), dim3(512), e->lds_bytes, e->stream, e->d_items, nitems,
                         samp, parity, macro, e->d_queue);
    else
      hipLaunchKernelGGL((k_ensemble<Mo, 256>), dim3(grid), dim3(256), e->lds_bytes, e->stream, e->d_items, nitems,
                         samp, parity, macro, e->d_queue);
    ENS_HIP(e, hipGetLastError());
    return MPLBM_OK;
  });
}

extern "C" {

mplbm_status mplbm_ensemble_create(const mplbm_config* cfgs, int32_t n, int32_t block_threads,
                                   mplbm_ensemble** out) {
  if (!cfgs || !out || n <= 0) return ens_fail(nullptr, MPLBM_ERR_ARG, "bad arguments");
  *out = nullptr;
  if (block_threads == 0) block_threads = 512;
  if (block_threads != 256 && block_threads != 512 && block_threads != 1024)
    return ens_fail(nullptr, MPLBM_ERR_ARG, "block_threads must be 256, 512 or 1024");
  if (cfgs[0].kernel == MPLBM_KERNEL_TWOPASS)
    return ens_fail(nullptr, MPLBM_ERR_UNSUPPORTED, "the ensemble runs the compact or the fused kernel");
  for (int32_t s = 0; s < n; ++s)
    if (cfgs[s].rounding != MPLBM_ROUND_NEAREST)  // a replayed graph would repeat the dither of one step
      return ens_fail(nullptr, MPLBM_ERR_UNSUPPORTED, "the ensemble stores with round-to-nearest only");
  for (int32_t s = 0; s < n; ++s)
    if (cfgs[s].shift != MPLBM_SHIFT_GLOBAL)
      return ens_fail(nullptr, MPLBM_ERR_UNSUPPORTED, "the ensemble stores with the global shift only");
  const bool compact = cfgs[0].kernel != MPLBM_KERNEL_FUSED;
  if (compact && cfgs[0].storage == MPLBM_FP64 && block_threads > 512)
    return ens_fail(nullptr, MPLBM_ERR_ARG, "FP64 compact ensemble needs block_threads <= 512");
  for (int32_t s = 0; s < n; ++s) {
    if (cfgs[s].storage != cfgs[0].storage || cfgs[s].math != cfgs[0].math)
      return ens_fail(nullptr, MPLBM_ERR_ARG, "samples must share storage and arithmetic");
    for (int f = 0; f < 6; ++f)
      if (cfgs[s].face_type[f] != MPLBM_FACE_PERIODIC)
        return ens_fail(nullptr, MPLBM_ERR_UNSUPPORTED, "ensemble samples are periodic");
  }
  mplbm_ensemble* e = new mplbm_ensemble();
  e->compact = compact;
  e->block = block_threads;
  e->storage = cfgs[0].storage;
  e->math = cfgs[0].math;
  for (int32_t s = 0; s < n; ++s) {
    mplbm_config c = cfgs[s];
    c.kernel = compact ? MPLBM_KERNEL_COMPACT : MPLBM_KERNEL_FUSED;
    mplbm_ctx* ctx = nullptr;
    mplbm_status st = mplbm_create(&c, &ctx);
    if (st != MPLBM_OK) {
      const std::string msg = g_create_error;
      mplbm_ensemble_destroy(e);
      return ens_fail(nullptr, st, "sample %d: %s", s, msg.c_str());
    }
    e->ctx.push_back(ctx);
  }
  hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreate(&e->ev0);
  if (he == hipSuccess) he = hipEventCreate(&e->ev1);
  if (he == hipSuccess) he = hipMalloc((void**)&e->d_queue, sizeof(EnsQueue));
  if (he == hipSuccess) he = hipMemset(e->d_queue, 0, sizeof(EnsQueue));
  if (he != hipSuccess) {
    mplbm_ensemble_destroy(e);
    return ens_fail(nullptr, MPLBM_ERR_HIP, "stream/event/queue: %s", hipGetErrorString(he));
  }
  mplbm_status st = ens_plan(e);
  if (st != MPLBM_OK) {
    const std::string msg = e->err.empty() ? g_create_error : e->err;
    mplbm_ensemble_destroy(e);
    return ens_fail(nullptr, st, "%s", msg.c_str());
  }
  *out = e;
  return MPLBM_OK;
}

void mplbm_ensemble_destroy(mplbm_ensemble* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->d_items) (void)hipFree(e->d_items);
  if (e->d_samples) (void)hipFree(e->d_samples);
  if (e->d_queue) (void)hipFree(e->d_queue);
  for (mplbm_ctx* c : e->ctx) mplbm_destroy(c);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

const char* mplbm_ensemble_last_error(const mplbm_ensemble* e) { return e ? e->err.c_str() : g_create_error.c_str(); }

#define ENS_SAMPLE(e, s)                                                                  \
  do {                                                                                    \
    if (!(e)) return ens_fail(nullptr, MPLBM_ERR_ARG, "null ensemble");                   \
    if ((s) < 0 || (size_t)(s) >= (e)->ctx.size()) return ens_fail((e), MPLBM_ERR_ARG, "bad sample %d", (int)(s)); \
  } while (0)

mplbm_status mplbm_ensemble_set_tiling(mplbm_ensemble* e, int32_t s, int32_t tile_y, int32_t tile_z,
                                       int32_t seg_x) {
  if (!e) return ens_fail(nullptr, MPLBM_ERR_ARG, "null ensemble");
  if (s < -1 || s >= (int32_t)e->ctx.size()) return ens_fail(e, MPLBM_ERR_ARG, "bad sample %d", (int)s);
  if (!e->compact) return ens_fail(e, MPLBM_ERR_UNSUPPORTED, "per-sample tiles apply to the compact ensemble");
  if (tile_y < 0 || tile_z < 0 || seg_x < 0) return ens_fail(e, MPLBM_ERR_ARG, "negative tile");
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ens_fail(e, MPLBM_ERR_HIP, "sync");
  const auto saved = e->tile_req;
  e->tile_req.resize(e->ctx.size(), std::array<int32_t, 3>{0, 0, 0});
  for (int32_t k = (s < 0 ? 0 : s); k < (s < 0 ? (int32_t)e->ctx.size() : s + 1); ++k)
    e->tile_req[k] = {tile_y, tile_z, seg_x};
  mplbm_status st = ens_plan(e);
  if (st != MPLBM_OK) {
    const std::string msg = e->err;
    e->tile_req = saved;
    (void)ens_plan(e);
    return ens_fail(e, st, "%s", msg.c_str());
  }
  return MPLBM_OK;
}

mplbm_status mplbm_ensemble_set_geometry(mplbm_ensemble* e, int32_t s, const uint8_t* is_fluid) {
  ENS_SAMPLE(e, s);
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ens_fail(e, MPLBM_ERR_HIP, "sync");
  mplbm_status st = mplbm_set_geometry(e->ctx[s], is_fluid);
  if (st != MPLBM_OK) return ens_fail(e, st, "%s", e->ctx[s]->err.c_str());
  e->fluid_nodes = 0;
  for (mplbm_ctx* c : e->ctx) e->fluid_nodes += c->fluid_nodes;
  return MPLBM_OK;
}

mplbm_status mplbm_ensemble_init_fields(mplbm_ensemble* e, int32_t s, const float* rho) {
  ENS_SAMPLE(e, s);
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ens_fail(e, MPLBM_ERR_HIP, "sync");
  mplbm_status st = mplbm_init_fields(e->ctx[s], rho);
  if (st != MPLBM_OK) return ens_fail(e, st, "%s", e->ctx[s]->err.c_str());
  return MPLBM_OK;
}

mplbm_status mplbm_ensemble_step(mplbm_ensemble* e, int64_t step, int32_t nsteps, uint32_t flags,
                                 int32_t use_graph) {
  if (!e) return ens_fail(nullptr, MPLBM_ERR_ARG, "null ensemble");
  if (nsteps <= 0) return MPLBM_OK;
  for (mplbm_ctx* c : e->ctx)
    if (!c->have_fields) return ens_fail(e, MPLBM_ERR_STATE, "every sample needs geometry and fields first");
  bool injects = false;
  for (mplbm_ctx* c : e->ctx) injects |= c->cfg.rho_inject_period > 0 && c->cfg.rho_increment != 0.0;
  const bool macro = (flags & MPLBM_STEP_WRITE_MACRO) != 0;
  mplbm_status st;
  if (injects) {
    // the descriptor table changes with step % period: direct launches
    for (int32_t k = 0; k < nsteps; ++k) {
      if ((st = ens_upload(e, step + k)) != MPLBM_OK) return st;
      if ((st = ens_launch(e, 0, macro && k == nsteps - 1)) != MPLBM_OK) return st;
      for (mplbm_ctx* c : e->ctx) c->cur = 1 - c->cur;
    }
    return MPLBM_OK;
  }
  if ((st = ens_upload(e, step)) != MPLBM_OK) return st;
  const int32_t plain = macro ? nsteps - 1 : nsteps;
  int parity = 0;
  int32_t done = 0;
  if (use_graph && plain >= 2) {
    if (!e->graph_exec) {
      ENS_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
      mplbm_status s0 = ens_launch(e, 0, 0);
      mplbm_status s1 = s0 == MPLBM_OK ? ens_launch(e, 1, 0) : s0;
      hipError_t ce = hipStreamEndCapture(e->stream, &e->graph);
      if (s1 != MPLBM_OK) return s1;
      if (ce != hipSuccess) return ens_fail(e, MPLBM_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
      ENS_HIP(e, hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
    }
    for (; done + 2 <= plain; done += 2) ENS_HIP(e, hipGraphLaunch(e->graph_exec, e->stream));
  }
  for (; done < plain; ++done) {
    if ((st = ens_launch(e, parity ^ (done & 1), 0)) != MPLBM_OK) return st;
  }
  if (macro && (st = ens_launch(e, plain & 1, 1)) != MPLBM_OK) return st;
  if (nsteps & 1)
    for (mplbm_ctx* c : e->ctx) c->cur = 1 - c->cur;
  return MPLBM_OK;
}

mplbm_status mplbm_ensemble_step_timed(mplbm_ensemble* e, int64_t step, int32_t nsteps, uint32_t flags,
                                       int32_t use_graph, double* elapsed_ms) {
  if (!e) return ens_fail(nullptr, MPLBM_ERR_ARG, "null ensemble");
  // build tables and graph first so the timed region holds only the steps
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  ENS_HIP(e, hipEventRecord(e->ev0, e->stream));
  mplbm_status st = mplbm_ensemble_step(e, step, nsteps, flags, use_graph);
  if (st != MPLBM_OK) return st;
  ENS_HIP(e, hipEventRecord(e->ev1, e->stream));
  ENS_HIP(e, hipEventSynchronize(e->ev1));
  float ms = 0.f;
  ENS_HIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
  if (elapsed_ms) *elapsed_ms = ms;
  return MPLBM_OK;
}

mplbm_status mplbm_ensemble_read_fields(mplbm_ensemble* e, int32_t s, float* rho, float* pressure, float* u) {
  ENS_SAMPLE(e, s);
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ens_fail(e, MPLBM_ERR_HIP, "sync");
  mplbm_status st = mplbm_read_fields(e->ctx[s], rho, pressure, u);
  return st == MPLBM_OK ? st : ens_fail(e, st, "%s", e->ctx[s]->err.c_str());
}

mplbm_status mplbm_ensemble_read_populations_f64(mplbm_ensemble* e, int32_t s, double* f) {
  ENS_SAMPLE(e, s);
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ens_fail(e, MPLBM_ERR_HIP, "sync");
  mplbm_status st = mplbm_read_populations_f64(e->ctx[s], f);
  return st == MPLBM_OK ? st : ens_fail(e, st, "%s", e->ctx[s]->err.c_str());
}

mplbm_status mplbm_ensemble_reduce(mplbm_ensemble* e, int32_t s, mplbm_stats64* out) {
  ENS_SAMPLE(e, s);
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ens_fail(e, MPLBM_ERR_HIP, "sync");
  mplbm_status st = mplbm_reduce(e->ctx[s], out);
  return st == MPLBM_OK ? st : ens_fail(e, st, "%s", e->ctx[s]->err.c_str());
}

mplbm_status mplbm_ensemble_info(const mplbm_ensemble* e, int32_t s, mplbm_info* out) {
  if (!e || s < 0 || (size_t)s >= e->ctx.size()) return MPLBM_ERR_ARG;
  return mplbm_info_get(e->ctx[s], out);
}

mplbm_status mplbm_ensemble_summary(const mplbm_ensemble* e, int64_t* work_items, int64_t* fluid_nodes,
                                    int32_t* block_threads) {
  if (!e) return MPLBM_ERR_ARG;
  if (work_items) *work_items = (int64_t)e->items.size();
  if (fluid_nodes) *fluid_nodes = e->fluid_nodes;
  if (block_threads) *block_threads = e->block;
  return MPLBM_OK;
}

}  // extern "C"

// ============================================================================
// cluster statistics on the device (calc_cluster, src/multiphase_lbm.cu:225-562)
//
// Liquid (rho > rhol) and vapour (rho < rhog) fluid nodes are labelled by
// 6-neighbour periodic connectivity with a lock-free union-find whose roots
// are the smallest cell index of each component.  Components of fewer than 7
// nodes are dropped and the rest numbered 1, 2, ... in the order of their
// smallest cell index, which is the order in which the reference's scan
// (z fastest) starts its depth-first searches, so the labels are the
// reference's LiqClusters / GasClusters values.

namespace mplbm {

constexpr int kClusterMin = 7;  // smaller clusters are discarded (:330-334)
constexpr int kClThreads = 256;
constexpr int kClWaves = kClThreads / 64;

struct ClAcc {
  double liq_p, vap_p, liq_rho, vap_rho;
  unsigned long long fluid, liquid, vapor, liq_solid, liq_vapor;
  unsigned long long candidates, pairs, touching;
  unsigned long long roots[2];
  unsigned long long sum[3], size;
  unsigned int rho_min, rho_max, p_min, p_max;
  unsigned int maxr;
};

// order-preserving unsigned key of a float (min/max with integer atomics)
__host__ __device__ inline uint32_t fkey(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float fkey_value(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

template <class T>
__device__ __forceinline__ T cl_wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint32_t cl_wave_min(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t cl_wave_max(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

// block-wide sum of v added to *dst by one atomic (every thread calls it)
template <class T>
__device__ void cl_block_add(T v, T* dst, T* sh) {
  v = cl_wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int k = 0; k < kClWaves; ++k) t += sh[k];
    if (t != (T)0) atomicAdd(dst, t);
  }
  __syncthreads();
}
template <bool kMax>
__device__ void cl_block_minmax(uint32_t v, uint32_t* dst, uint32_t* sh) {
  v = kMax ? cl_wave_max(v) : cl_wave_min(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = sh[0];
    for (int k = 1; k < kClWaves; ++k) t = kMax ? max(t, sh[k]) : min(t, sh[k]);
    if (kMax) atomicMax(dst, t);
    else atomicMin(dst, t);
  }
  __syncthreads();
}

struct ClDims {
  int lx, ly, lz;
  uint32_t n, plane;
  __device__ void coords(uint32_t c, int& x, int& y, int& z) const {
    x = (int)(c / plane);
    const uint32_t r = c - (uint32_t)x * plane;
    y = (int)(r / (uint32_t)lz);
    z = (int)(r - (uint32_t)y * lz);
  }
  __device__ uint32_t at(int x, int y, int z) const {
    x = x < 0 ? x + lx : (x >= lx ? x - lx : x);
    y = y < 0 ? y + ly : (y >= ly ? y - ly : y);
    z = z < 0 ? z + lz : (z >= lz ? z - lz : z);
    return (uint32_t)x * plane + (uint32_t)y * lz + (uint32_t)z;
  }
};

// Phase of every cell, FP64 phase sums, extrema (write_stats starts from
// 1e9 / -1e-9, :1058-1059) and the boundary counts of liquid nodes: the
// first neighbour x - e_i (i = 1..18) that is solid or below rhol (:277-316).
// Every cell starts as its own union-find root with size 0.
template <class M>
__global__ __launch_bounds__(kClThreads) void k_cl_classify(ClDims D, const uint8_t* __restrict__ mask,
                                                            const M* __restrict__ rho, const M* __restrict__ pres,
                                                            float rhol, float rhog, uint8_t* __restrict__ phase,
                                                            int32_t* __restrict__ parent, int32_t* __restrict__ size,
                                                            ClAcc* acc) {
  __shared__ double shd[kClWaves];
  __shared__ unsigned long long shu[kClWaves];
  __shared__ uint32_t shk[kClWaves];
  double lp = 0, vp = 0, lr = 0, vr = 0;
  unsigned long long nf = 0, nl = 0, nv = 0, nls = 0, nlv = 0;
  uint32_t rmin = fkey(1e9f), rmax = fkey(-1e-9f), pmin = fkey(1e9f), pmax = fkey(-1e-9f);
  for (uint32_t c = blockIdx.x * kClThreads + threadIdx.x; c < D.n; c += gridDim.x * kClThreads) {
    uint8_t ph = 0;
    parent[c] = (int32_t)c;
    size[c] = 0;
    if (mask[c]) {
      const float r = (float)rho[c], p = (float)pres[c];
      ++nf;
      rmin = min(rmin, fkey(r));
      rmax = max(rmax, fkey(r));
      pmin = min(pmin, fkey(p));
      pmax = max(pmax, fkey(p));
      if (r > rhol) {
        ph = 1;
        lp += (double)p;
        lr += (double)r;
        ++nl;
        int x, y, z;
        D.coords(c, x, y, z);
        for (int i = 1; i < kQ; ++i) {
          const uint32_t u = D.at(x - ex(i), y - ey(i), z - ez(i));
          if (!mask[u]) {
            ++nls;
            break;
          }
          if ((float)rho[u] < rhol) {
            ++nlv;
            break;
          }
        }
      } else if (r < rhog) {
        ph = 2;
        vp += (double)p;
        vr += (double)r;
        ++nv;
      }
    }
    phase[c] = ph;
  }
  cl_block_add(lp, &acc->liq_p, shd);
  cl_block_add(vp, &acc->vap_p, shd);
  cl_block_add(lr, &acc->liq_rho, shd);
  cl_block_add(vr, &acc->vap_rho, shd);
  cl_block_add(nf, &acc->fluid, shu);
  cl_block_add(nl, &acc->liquid, shu);
  cl_block_add(nv, &acc->vapor, shu);
  cl_block_add(nls, &acc->liq_solid, shu);
  cl_block_add(nlv, &acc->liq_vapor, shu);
  cl_block_minmax<false>(rmin, &acc->rho_min, shk);
  cl_block_minmax<true>(rmax, &acc->rho_max, shk);
  cl_block_minmax<false>(pmin, &acc->p_min, shk);
  cl_block_minmax<true>(pmax, &acc->p_max, shk);
}

// Root of c.  Parents only ever point to smaller indices, so a stale read
// still names an ancestor; path halving stores ancestors only.
__device__ __forceinline__ int32_t cl_find(int32_t* parent, int32_t c) {
  for (;;) {
    const int32_t p = __atomic_load_n(&parent[c], __ATOMIC_RELAXED);
    if (p == c) return c;
    const int32_t g = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
    if (g != p) __atomic_store_n(&parent[c], g, __ATOMIC_RELAXED);
    c = g;
  }
}

// root of c without path updates (once the forest no longer changes)
__device__ __forceinline__ int32_t cl_root(const int32_t* parent, int32_t c) {
  for (;;) {
    const int32_t p = __atomic_load_n(&parent[c], __ATOMIC_RELAXED);
    if (p == c) return c;
    c = p;
  }
}

// union: the larger root is linked under the smaller one by compare-and-swap
__device__ __forceinline__ void cl_union(int32_t* parent, int32_t a, int32_t b) {
  for (;;) {
    a = cl_find(parent, a);
    b = cl_find(parent, b);
    if (a == b) return;
    if (a < b) {
      const int32_t t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&parent[a], a, b) == a) return;
  }
}

// one edge per axis (+x, +y, +z, periodic) between nodes of the same phase
__global__ __launch_bounds__(kClThreads) void k_cl_hook(ClDims D, const uint8_t* __restrict__ phase,
                                                        int32_t* parent) {
  for (uint32_t c = blockIdx.x * kClThreads + threadIdx.x; c < D.n; c += gridDim.x * kClThreads) {
    const uint8_t ph = phase[c];
    if (!ph) continue;
    int x, y, z;
    D.coords(c, x, y, z);
    const uint32_t nb[3] = {D.at(x + 1, y, z), D.at(x, y + 1, z), D.at(x, y, z + 1)};
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (phase[nb[k]] == ph) cl_union(parent, (int32_t)c, (int32_t)nb[k]);
  }
}

// parent := root, and component sizes counted at the roots up to
// kClusterMin (only "fewer than 7" matters).  The only store to parent[c]
// is c's root, so concurrent walks still see ancestors.  Lanes of a wave
// hold consecutive cells; each run of equal roots adds its length once.
__global__ __launch_bounds__(kClThreads) void k_cl_compress(ClDims D, const uint8_t* __restrict__ phase,
                                                            int32_t* parent, int32_t* size) {
  const int lane = threadIdx.x & 63;
  for (uint32_t base = blockIdx.x * kClThreads; base < D.n; base += gridDim.x * kClThreads) {
    const uint32_t c = base + threadIdx.x;
    int32_t r = -1;
    if (c < D.n && phase[c]) r = cl_root(parent, (int32_t)c);
    const int32_t prev = __shfl_up(r, 1, 64);
    const bool start = lane == 0 || prev != r;
    const unsigned long long starts = __ballot(start);
    if (c < D.n && phase[c]) parent[c] = r;
    if (start && r >= 0) {
      const unsigned long long above = starts & ~((2ull << lane) - 1ull);
      const int next = above ? (int)__ffsll((long long)above) - 1 : 64;
      if (__atomic_load_n(&size[r], __ATOMIC_RELAXED) < kClusterMin) atomicAdd(&size[r], next - lane);
    }
  }
}

// Roots of kept components (kClusterMin nodes or more): counted (lists
// null) or appended to the per-phase lists, dropped roots' size set to 0.
__global__ __launch_bounds__(kClThreads) void k_cl_roots(ClDims D, const uint8_t* __restrict__ phase,
                                                         const int32_t* __restrict__ parent, int32_t* size,
                                                         int32_t* list_liq, int32_t* list_gas, ClAcc* acc) {
  for (uint32_t c = blockIdx.x * kClThreads + threadIdx.x; c < D.n; c += gridDim.x * kClThreads) {
    const uint8_t ph = phase[c];
    if (!ph || parent[c] != (int32_t)c) continue;
    if (size[c] >= kClusterMin) {
      const unsigned long long k = atomicAdd(&acc->roots[ph - 1], 1ull);
      if (list_liq) (ph == 1 ? list_liq : list_gas)[k] = (int32_t)c;
    } else if (list_liq) {
      size[c] = 0;
    }
  }
}

__global__ void k_cl_assign(const int32_t* __restrict__ list, int32_t count, int32_t* size) {
  const int32_t k = blockIdx.x * kClThreads + threadIdx.x;
  if (k < count) size[list[k]] = k + 1;
}

// label := cluster number of the node's root (0: dropped or no phase)
__global__ __launch_bounds__(kClThreads) void k_cl_label(ClDims D, const uint8_t* __restrict__ phase,
                                                         int32_t* parent, const int32_t* __restrict__ size) {
  for (uint32_t c = blockIdx.x * kClThreads + threadIdx.x; c < D.n; c += gridDim.x * kClThreads)
    parent[c] = phase[c] ? size[parent[c]] : 0;
}

// Distinct (liquid cluster, solid number) pairs over the 6 neighbours of
// the nodes of kept liquid clusters (the solids_touched sets of :318-360).
// table null: count the candidate incidences only.
__global__ __launch_bounds__(kClThreads) void k_cl_pairs(ClDims D, const uint8_t* __restrict__ phase,
                                                         const int32_t* __restrict__ label,
                                                         const int32_t* __restrict__ sid,
                                                         unsigned long long* table, unsigned long long mask,
                                                         int32_t* touch, ClAcc* acc) {
  __shared__ unsigned long long shu[kClWaves];
  unsigned long long cand = 0, pairs = 0, touching = 0;
  for (uint32_t c = blockIdx.x * kClThreads + threadIdx.x; c < D.n; c += gridDim.x * kClThreads) {
    if (phase[c] != 1) continue;
    const int32_t lab = label[c];
    if (lab <= 0) continue;
    int x, y, z;
    D.coords(c, x, y, z);
    const uint32_t nb[6] = {D.at(x - 1, y, z), D.at(x, y - 1, z), D.at(x, y, z - 1),
                            D.at(x + 1, y, z), D.at(x, y + 1, z), D.at(x, y, z + 1)};
    for (int k = 0; k < 6; ++k) {
      const int32_t s = sid[nb[k]];
      if (s == 0 || phase[nb[k]] == 1) continue;
      if (!table) {
        ++cand;
        continue;
      }
      const unsigned long long key = ((unsigned long long)(uint32_t)lab << 32) | (uint32_t)s;
      unsigned long long h = (key * 0x9E3779B97F4A7C15ull) >> 17;
      for (;;) {
        h &= mask;
        const unsigned long long cur = __atomic_load_n(&table[h], __ATOMIC_RELAXED);
        if (cur == key) break;
        if (cur == ~0ull) {
          const unsigned long long old = atomicCAS(&table[h], ~0ull, key);
          if (old == ~0ull) {
            ++pairs;
            if (atomicExch(&touch[lab], 1) == 0) ++touching;
            break;
          }
          if (old == key) break;
        }
        ++h;
      }
    }
  }
  cl_block_add(cand, &acc->candidates, shu);
  cl_block_add(pairs, &acc->pairs, shu);
  cl_block_add(touching, &acc->touching, shu);
}

// Single liquid cluster (:362-533): whether it touches both faces of an
// axis (brk) and which planes along each axis it occupies.
__global__ __launch_bounds__(kClThreads) void k_cl_span(ClDims D, const uint8_t* __restrict__ phase,
                                                        const int32_t* __restrict__ label, int32_t* occ,
                                                        int32_t* brk) {
  for (uint32_t c = blockIdx.x * kClThreads + threadIdx.x; c < D.n; c += gridDim.x * kClThreads) {
    if (phase[c] != 1 || label[c] != 1) continue;
    int x, y, z;
    D.coords(c, x, y, z);
    occ[x] = 1;
    occ[D.lx + y] = 1;
    occ[D.lx + D.ly + z] = 1;
    auto one = [&](uint32_t u) { return phase[u] == 1 && label[u] == 1; };
    if (x == 0 && one(D.at(D.lx - 1, y, z))) brk[0] = 1;
    if (y == 0 && one(D.at(x, D.ly - 1, z))) brk[1] = 1;
    if (z == 0 && one(D.at(x, y, D.lz - 1))) brk[2] = 1;
  }
}

struct ClShift {
  int brk[3], vplane[3];
  float ctr[3];
};

// coordinate of a cluster node along axis a, unwrapped across the first
// free plane of a broken axis
__device__ __forceinline__ int cl_unwrap(const ClShift& S, const ClDims& D, int a, int v) {
  const int L = a == 0 ? D.lx : (a == 1 ? D.ly : D.lz);
  return v + ((S.brk[a] && v < S.vplane[a]) ? L : 0);
}

// centre sums (pass 0) or the largest distance from the centre (pass 1)
__global__ __launch_bounds__(kClThreads) void k_cl_shape(ClDims D, const uint8_t* __restrict__ phase,
                                                         const int32_t* __restrict__ label, ClShift S, int pass,
                                                         ClAcc* acc) {
  __shared__ unsigned long long shu[kClWaves];
  __shared__ uint32_t shk[kClWaves];
  unsigned long long s[3] = {0, 0, 0}, cnt = 0;
  float maxr = 0.f;
  for (uint32_t c = blockIdx.x * kClThreads + threadIdx.x; c < D.n; c += gridDim.x * kClThreads) {
    if (phase[c] != 1 || label[c] != 1) continue;
    int co[3];
    D.coords(c, co[0], co[1], co[2]);
    if (pass == 0) {
#pragma unroll
      for (int a = 0; a < 3; ++a) s[a] += (unsigned long long)cl_unwrap(S, D, a, co[a]);
      ++cnt;
    } else {
      float r2 = 0.f;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float adj = (float)cl_unwrap(S, D, a, co[a]);
        r2 += (S.ctr[a] - adj) * (S.ctr[a] - adj);
      }
      maxr = max(maxr, __builtin_sqrtf(r2));
    }
  }
  if (pass == 0) {
    cl_block_add(s[0], &acc->sum[0], shu);
    cl_block_add(s[1], &acc->sum[1], shu);
    cl_block_add(s[2], &acc->sum[2], shu);
    cl_block_add(cnt, &acc->size, shu);
  } else {
    cl_block_minmax<true>(__float_as_uint(maxr), &acc->maxr, shk);
  }
}

}  // namespace mplbm

// ----------------------------------------------------------------------------
// cluster statistics API

static unsigned cl_blocks(uint64_t n) {
  const uint64_t b = (n + kClThreads - 1) / kClThreads;
  return (unsigned)std::min<uint64_t>(std::max<uint64_t>(b, 1), 8192);
}

static mplbm_status cl_read_acc(mplbm_ctx* ctx, ClAcc* h) {
  HIP_TRY(ctx, hipMemcpyAsync(h, ctx->cl_acc, sizeof(ClAcc), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MPLBM_OK;
}

// kept roots of one phase: sorted ascending and numbered 1, 2, ... at the roots
static mplbm_status cl_number(mplbm_ctx* ctx, int32_t* d_list, int64_t count) {
  if (count == 0) return MPLBM_OK;
  std::vector<int32_t> h((size_t)count);
  HIP_TRY(ctx, hipMemcpyAsync(h.data(), d_list, (size_t)count * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  std::sort(h.begin(), h.end());
  HIP_TRY(ctx, hipMemcpyAsync(d_list, h.data(), (size_t)count * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_cl_assign, dim3((unsigned)((count + kClThreads - 1) / kClThreads)), dim3(kClThreads), 0,
                     ctx->stream, d_list, (int32_t)count, ctx->cl_size);
  HIP_TRY(ctx, hipGetLastError());
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));  // h is released on return
  return MPLBM_OK;
}

extern "C" {

mplbm_status mplbm_set_solid_numbers(mplbm_ctx* ctx, const int64_t* solid_number, int64_t solid_count) {
  if (!ctx) return fail(ctx, MPLBM_ERR_ARG, "null context");
  if (solid_count < 0) return fail(ctx, MPLBM_ERR_ARG, "negative solid_count");
  const uint64_t n = ctx->ncell;
  if (!solid_number) {
    if (ctx->cl_sid) HIP_TRY(ctx, hipFree(ctx->cl_sid));
    ctx->cl_sid = nullptr;
    ctx->cl_solid_count = solid_count;
    return MPLBM_OK;
  }
  std::vector<int32_t> h(n);
  for (uint64_t c = 0; c < n; ++c) {
    if (solid_number[c] < 0 || solid_number[c] > INT32_MAX)
      return fail(ctx, MPLBM_ERR_ARG, "solid number %lld at cell %llu outside [0, 2^31)",
                  (long long)solid_number[c], (unsigned long long)c);
    h[c] = (int32_t)solid_number[c];
  }
  if (!ctx->cl_sid) HIP_TRY(ctx, hipMalloc((void**)&ctx->cl_sid, n * 4));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->cl_sid, h.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->cl_solid_count = solid_count;
  return MPLBM_OK;
}

mplbm_status mplbm_calc_cluster(mplbm_ctx* ctx, double rhol_spinodal, double rhog_spinodal, mplbm_clusters* out,
                                 int32_t* liq_label, int32_t* gas_label) {
  if (!ctx || !out) return fail(ctx, MPLBM_ERR_ARG, "null argument");
  if (!ctx->have_fields) return fail(ctx, MPLBM_ERR_STATE, "fields not initialised");
  const uint64_t n = ctx->ncell;
  if (n >= (1ull << 31)) return fail(ctx, MPLBM_ERR_UNSUPPORTED, "cluster labels need fewer than 2^31 cells");
  if (!ctx->cl_parent) {
    HIP_TRY(ctx, hipMalloc((void**)&ctx->cl_parent, n * 4));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->cl_size, n * 4));
    HIP_TRY(ctx, hipMalloc((void**)&ctx->cl_phase, n));
    HIP_TRY(ctx, hipMalloc(&ctx->cl_acc, sizeof(ClAcc)));
  }
  const ClDims D{(int)ctx->cfg.lx, (int)ctx->cfg.ly, (int)ctx->cfg.lz, (uint32_t)n,
                 (uint32_t)(ctx->cfg.ly * ctx->cfg.lz)};
  const unsigned nb = cl_blocks(n);
  const dim3 blk(kClThreads);
  ClAcc h{};
  h.rho_min = fkey(1e9f);
  h.rho_max = fkey(-1e-9f);
  h.p_min = fkey(1e9f);
  h.p_max = fkey(-1e-9f);
  ClAcc* acc = reinterpret_cast<ClAcc*>(ctx->cl_acc);
  HIP_TRY(ctx, hipMemcpyAsync(acc, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
  // thresholds compare as the reference's floats (real rho > rhol_spinodal)
  const float rhol = (float)rhol_spinodal, rhog = (float)rhog_spinodal;
  if (ctx->cfg.storage == MPLBM_FP64)
    hipLaunchKernelGGL(k_cl_classify<double>, dim3(nb), blk, 0, ctx->stream, D, ctx->mask,
                       (const double*)ctx->rho, (const double*)ctx->pres, rhol, rhog, ctx->cl_phase, ctx->cl_parent,
                       ctx->cl_size, acc);
  else
    hipLaunchKernelGGL(k_cl_classify<float>, dim3(nb), blk, 0, ctx->stream, D, ctx->mask, (const float*)ctx->rho,
                       (const float*)ctx->pres, rhol, rhog, ctx->cl_phase, ctx->cl_parent, ctx->cl_size, acc);
  hipLaunchKernelGGL(k_cl_hook, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent);
  hipLaunchKernelGGL(k_cl_compress, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, ctx->cl_size);
  hipLaunchKernelGGL(k_cl_roots, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, ctx->cl_size,
                     (int32_t*)nullptr, (int32_t*)nullptr, acc);
  HIP_TRY(ctx, hipGetLastError());
  mplbm_status st = cl_read_acc(ctx, &h);
  if (st != MPLBM_OK) return st;
  const int64_t nroot[2] = {(int64_t)h.roots[0], (int64_t)h.roots[1]};
  int32_t* lists = nullptr;
  HIP_TRY(ctx, hipMalloc((void**)&lists, (size_t)(nroot[0] + nroot[1] + 1) * 4));
  auto run = [&]() -> mplbm_status {
    const unsigned long long zero[2] = {0, 0};
    HIP_TRY(ctx, hipMemcpyAsync(&acc->roots[0], zero, sizeof zero, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_cl_roots, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, ctx->cl_size,
                       lists, lists + nroot[0], acc);
    HIP_TRY(ctx, hipGetLastError());
    mplbm_status s2 = cl_number(ctx, lists, nroot[0]);
    if (s2 == MPLBM_OK) s2 = cl_number(ctx, lists + nroot[0], nroot[1]);
    return s2;
  };
  st = run();
  (void)hipFree(lists);
  if (st != MPLBM_OK) return st;
  hipLaunchKernelGGL(k_cl_label, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, ctx->cl_size);
  HIP_TRY(ctx, hipGetLastError());

  // cluster order and wet coordination number (:535-561): distinct
  // (cluster, solid) pairs in an open-addressing table
  if (ctx->cl_sid && nroot[0] > 0) {
    hipLaunchKernelGGL(k_cl_pairs, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, ctx->cl_sid,
                       (unsigned long long*)nullptr, 0ull, (int32_t*)nullptr, acc);
    HIP_TRY(ctx, hipGetLastError());
    if ((st = cl_read_acc(ctx, &h)) != MPLBM_OK) return st;
    if (h.candidates > 0) {
      unsigned long long cap = 1024;
      while (cap < 2 * h.candidates) cap <<= 1;
      void* tab = nullptr;
      HIP_TRY(ctx, hipMalloc(&tab, cap * 8 + (size_t)(nroot[0] + 1) * 4));
      int32_t* touch = reinterpret_cast<int32_t*>(reinterpret_cast<unsigned long long*>(tab) + cap);
      hipError_t e = hipMemsetAsync(tab, 0xFF, cap * 8, ctx->stream);
      if (e == hipSuccess) e = hipMemsetAsync(touch, 0, (size_t)(nroot[0] + 1) * 4, ctx->stream);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_cl_pairs, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent,
                           ctx->cl_sid, (unsigned long long*)tab, cap - 1, touch, acc);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      (void)hipFree(tab);
      if (e != hipSuccess) return fail(ctx, MPLBM_ERR_HIP, "cluster pairs: %s", hipGetErrorString(e));
    }
  }

  // single liquid cluster: periodic unwrapping, centre and radius (:362-533)
  double radius = 0.0;
  if (nroot[0] == 1) {
    const int64_t L[3] = {ctx->cfg.lx, ctx->cfg.ly, ctx->cfg.lz};
    const size_t nocc = (size_t)(L[0] + L[1] + L[2]);
    int32_t* occ = nullptr;
    HIP_TRY(ctx, hipMalloc((void**)&occ, (nocc + 3) * 4));
    std::vector<int32_t> ho(nocc + 3);
    hipError_t e = hipMemsetAsync(occ, 0, (nocc + 3) * 4, ctx->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_cl_span, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, occ,
                         occ + nocc);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(ho.data(), occ, (nocc + 3) * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(occ);
    if (e != hipSuccess) return fail(ctx, MPLBM_ERR_HIP, "cluster span: %s", hipGetErrorString(e));
    ClShift S{};
    int64_t off = 0;
    for (int a = 0; a < 3; ++a) {
      S.brk[a] = ho[nocc + a] != 0;
      S.vplane[a] = 0;
      if (S.brk[a])  // first plane free of the cluster
        for (int64_t k = 0; k < L[a]; ++k)
          if (!ho[(size_t)(off + k)]) {
            S.vplane[a] = (int)k;
            break;
          }
      off += L[a];
    }
    hipLaunchKernelGGL(k_cl_shape, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, S, 0, acc);
    HIP_TRY(ctx, hipGetLastError());
    if ((st = cl_read_acc(ctx, &h)) != MPLBM_OK) return st;
    // integer division of the sums as in the reference (size_t sums, :496-498)
    for (int a = 0; a < 3; ++a) S.ctr[a] = h.size ? (float)(h.sum[a] / h.size) : 0.f;
    hipLaunchKernelGGL(k_cl_shape, dim3(nb), blk, 0, ctx->stream, D, ctx->cl_phase, ctx->cl_parent, S, 1, acc);
    HIP_TRY(ctx, hipGetLastError());
    if ((st = cl_read_acc(ctx, &h)) != MPLBM_OK) return st;
    float r;
    memcpy(&r, &h.maxr, 4);
    radius = r;
  }
  if ((st = cl_read_acc(ctx, &h)) != MPLBM_OK) return st;

  if (liq_label || gas_label) {
    std::vector<int32_t> lab(n);
    std::vector<uint8_t> ph(n);
    HIP_TRY(ctx, hipMemcpyAsync(lab.data(), ctx->cl_parent, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ph.data(), ctx->cl_phase, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (uint64_t c = 0; c < n; ++c) {
      if (liq_label) liq_label[c] = ph[c] == 1 ? lab[c] : 0;
      if (gas_label) gas_label[c] = ph[c] == 2 ? lab[c] : 0;
    }
  }

  mplbm_clusters o{};
  o.min_rho = fkey_value(h.rho_min);
  o.max_rho = fkey_value(h.rho_max);
  o.min_p = fkey_value(h.p_min);
  o.max_p = fkey_value(h.p_max);
  o.fluid_count = (int64_t)h.fluid;
  o.liquid_count = (int64_t)h.liquid;
  o.vapor_count = (int64_t)h.vapor;
  if (h.liquid) {
    o.avg_liq_press = (float)(h.liq_p / (double)h.liquid);
    o.avg_liq_rho = (float)(h.liq_rho / (double)h.liquid);
    o.ratio_liquid_solid_boundary = (float)((double)h.liq_solid / (double)h.liquid);
    o.ratio_liquid_vapor_boundary = (float)((double)h.liq_vapor / (double)h.liquid);
  }
  if (h.vapor) {
    o.avg_vap_press = (float)(h.vap_p / (double)h.vapor);
    o.avg_vap_rho = (float)(h.vap_rho / (double)h.vapor);
  }
  o.saturation = h.fluid ? (double)((float)h.liquid / (float)h.fluid) : 0.0;
  o.liq_cluster_count = nroot[0];
  o.gas_cluster_count = nroot[1];
  // sums of set sizes in float as the reference accumulates them
  o.cluster_order = h.touching ? (double)((float)h.pairs / (float)h.touching) : 0.0;
  o.wet_coord_num = ctx->cl_solid_count ? (double)((float)h.pairs / (float)ctx->cl_solid_count) : (double)(float)h.pairs;
  o.cluster_radius = radius;
  *out = o;
  return MPLBM_OK;
}

}  // extern "C"

// ----------------------------------------------------------------------------
// sphere voxeliser: place_sphere (src/multiphase_lbm.cu:893-927) for a list of
// spheres, as read_positions (:1106-1126) and generate_solids (:1086-1103)
// call it once per sphere.  One workgroup per sphere: the float-stepped
// sample coordinates x - R, x - R + 1, ... <= x + R of each axis go to LDS
// (the same float additions as the reference's loop), then the workgroup's
// threads test the sample points of the box.  A covered cell is marked solid
// and its sphere number raised with an atomic maximum, so the last sphere of
// the list that covers a cell names it, as with sequential placement.

namespace mplbm {

constexpr int kVoxAxis = 1024;  // sample points per axis of one sphere's box
constexpr int kVoxThreads = 256;

__global__ __launch_bounds__(kVoxThreads) void k_voxelise(const float4* __restrict__ sph, int64_t count, int lx,
                                                          int ly, int lz, uint8_t* __restrict__ mask,
                                                          uint32_t* __restrict__ ids) {
  __shared__ float ax[3][kVoxAxis];
  __shared__ int an[3];
  for (int64_t s = blockIdx.x; s < count; s += gridDim.x) {
    const float4 q = sph[s];
    if (threadIdx.x < 3) {
      const float c = threadIdx.x == 0 ? q.x : (threadIdx.x == 1 ? q.y : q.z);
      const float lo = c - q.w, hi = c + q.w;
      int n = 0;
      // the host checked that every axis has fewer than kVoxAxis samples
      for (float p = lo; p <= hi && n < kVoxAxis; p += 1.f) ax[threadIdx.x][n++] = p;
      an[threadIdx.x] = n;
    }
    __syncthreads();
    const int nx = an[0], ny = an[1], nz = an[2];
    const int nyz = ny * nz;
    const float R2 = q.w * q.w;
    const int total = nx * nyz;
    for (int k = (int)threadIdx.x; k < total; k += kVoxThreads) {
      const int i = k / nyz;
      const int r = k - i * nyz;
      const int j = r / nz;
      const int l = r - j * nz;
      const float px = ax[0][i], py = ax[1][j], pz = ax[2][l];
      const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
      const float dist2 = dx * dx + dy * dy + dz * dz;
      if (!(dist2 < R2)) continue;
      // nearest() of include/lbm_model.cuh:58 and one periodic wrap
      int cx = (int)floor((double)px + 0.5), cy = (int)floor((double)py + 0.5), cz = (int)floor((double)pz + 0.5);
      if (cx < 0) cx += lx;
      if (cy < 0) cy += ly;
      if (cz < 0) cz += lz;
      if (cx >= lx) cx -= lx;
      if (cy >= ly) cy -= ly;
      if (cz >= lz) cz -= lz;
      if (cx < 0 || cy < 0 || cz < 0 || cx >= lx || cy >= ly || cz >= lz) continue;  // beyond one wrap
      const size_t c = ((size_t)cx * (size_t)ly + (size_t)cy) * (size_t)lz + (size_t)cz;
      mask[c] = 0;
      if (ids) atomicMax(&ids[c], (uint32_t)(s + 1));
    }
    __syncthreads();  // the next sphere rewrites ax / an
  }
}

// solid numbers of the covered cells: first_id - 1 + (list position + 1)
__global__ void k_vox_numbers(const uint32_t* __restrict__ ids, int64_t* __restrict__ sn, int64_t base, uint64_t n) {
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = ids[c];
    if (v) sn[c] = base + (int64_t)v;
  }
}

}  // namespace mplbm

// Every axis of every sphere's box must hold a float-stepped sample run that
// ends and fits k_voxelise's LDS: coordinates well inside float's integer
// range and fewer than kVoxAxis samples.
static mplbm_status check_spheres(const float* spheres, int64_t count) {
  for (int64_t s = 0; s < count; ++s) {
    const float* q = spheres + 4 * s;
    for (int a = 0; a < 3; ++a) {
      const float lo = q[a] - q[3], hi = q[a] + q[3];
      if (!std::isfinite(q[a]) || !std::isfinite(q[3]) || std::fabs(lo) > 8388608.f || std::fabs(hi) > 8388608.f ||
          (double)hi - (double)lo > (double)(kVoxAxis - 3))
        return fail(nullptr, MPLBM_ERR_ARG, "place_spheres: sphere %ld (%g, %g, %g, R %g) outside the voxeliser's "
                    "range (|coordinate| + R <= 2^23, R <= %d)", (long)s, q[0], q[1], q[2], q[3], (kVoxAxis - 3) / 2);
    }
  }
  return MPLBM_OK;
}

extern "C" {

mplbm_status mplbm_place_spheres(int64_t lx, int64_t ly, int64_t lz, const float* spheres, int64_t count,
                                 int64_t first_id, uint8_t* is_fluid, int64_t* solid_number) {
  using namespace mplbm;
  if (lx < 1 || ly < 1 || lz < 1 || lx > INT32_MAX || ly > INT32_MAX || lz > INT32_MAX)
    return fail(nullptr, MPLBM_ERR_ARG, "place_spheres: bad domain %ld x %ld x %ld", (long)lx, (long)ly, (long)lz);
  if (count < 0 || (count > 0 && !spheres) || !is_fluid)
    return fail(nullptr, MPLBM_ERR_ARG, "place_spheres: null array or negative count");
  if (count >= (int64_t)UINT32_MAX)
    return fail(nullptr, MPLBM_ERR_ARG, "place_spheres: %ld spheres in one call (max %u)", (long)count,
                UINT32_MAX - 1);
  if (const mplbm_status cs = check_spheres(spheres, count); cs != MPLBM_OK) return cs;
  const uint64_t n = (uint64_t)lx * (uint64_t)ly * (uint64_t)lz;
  uint8_t* dmask = nullptr;
  uint32_t* dids = nullptr;
  int64_t* dsn = nullptr;
  float4* dsph = nullptr;
  mplbm_status st = MPLBM_OK;
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && st == MPLBM_OK)
      st = fail(nullptr, e == hipErrorOutOfMemory ? MPLBM_ERR_NOMEM : MPLBM_ERR_HIP, "place_spheres: %s: %s", what,
                hipGetErrorString(e));
    return st == MPLBM_OK;
  };
  if (check(hipMalloc(&dmask, n), "hipMalloc mask") &&
      check(hipMemcpy(dmask, is_fluid, n, hipMemcpyHostToDevice), "copy mask") &&
      (count == 0 || (check(hipMalloc(&dsph, (size_t)count * sizeof(float4)), "hipMalloc spheres") &&
                      check(hipMemcpy(dsph, spheres, (size_t)count * sizeof(float4), hipMemcpyHostToDevice),
                            "copy spheres"))) &&
      (!solid_number || (check(hipMalloc(&dids, n * sizeof(uint32_t)), "hipMalloc numbers") &&
                         check(hipMemset(dids, 0, n * sizeof(uint32_t)), "clear numbers")))) {
    if (count > 0) {
      const unsigned grid = (unsigned)std::min<int64_t>(count, 1 << 16);
      hipLaunchKernelGGL(k_voxelise, dim3(grid), dim3(kVoxThreads), 0, 0, dsph, count, (int)lx, (int)ly, (int)lz,
                         dmask, dids);
      check(hipGetLastError(), "k_voxelise");
    }
    if (st == MPLBM_OK && check(hipMemcpy(is_fluid, dmask, n, hipMemcpyDeviceToHost), "copy mask back") &&
        solid_number && count > 0 && check(hipMalloc(&dsn, n * sizeof(int64_t)), "hipMalloc solid numbers") &&
        check(hipMemcpy(dsn, solid_number, n * sizeof(int64_t), hipMemcpyHostToDevice), "copy solid numbers")) {
      const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16);
      hipLaunchKernelGGL(k_vox_numbers, dim3(grid), dim3(256), 0, 0, dids, dsn, first_id - 1, n);
      if (check(hipGetLastError(), "k_vox_numbers"))
        check(hipMemcpy(solid_number, dsn, n * sizeof(int64_t), hipMemcpyDeviceToHost), "copy solid numbers back");
    }
  }
  (void)hipFree(dmask);
  (void)hipFree(dids);
  (void)hipFree(dsn);
  (void)hipFree(dsph);
  return st;
}

}  // extern "C"

// ============================================================================
// sphere voxeliser that allocates nothing (mplbm_place_spheres_noalloc)

namespace mplbm {

// Scratch of mplbm_place_spheres_noalloc: the per-cell sphere numbers
// (uint32, cleared by every call that needs them) followed by the staged
// sphere list (float4).  One copy per GPU, held until the process ends.
constexpr uint64_t kNoAllocMaxCells = 1ull << 24;    // 256^3 cells of sphere numbers (64 MiB)
constexpr uint64_t kNoAllocMaxSpheres = 1ull << 18;  // 262,144 spheres (4 MiB)
constexpr uint64_t kNoAllocIdsBytes = kNoAllocMaxCells * sizeof(uint32_t);
constexpr uint64_t kScratchBytes = kNoAllocIdsBytes + kNoAllocMaxSpheres * sizeof(float4);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace mplbm

namespace {
// Device address of `scratch` on each device, from hipGetSymbolAddress with
// that device current; the symbol's host-side name is not a device address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

mplbm_status scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(nullptr, MPLBM_ERR_HIP, "place_spheres_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, MPLBM_ERR_UNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(mplbm::scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, MPLBM_ERR_HIP, "place_spheres_noalloc: hipGetSymbolAddress(scratch): %s",
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MPLBM_OK;
}
}  // namespace

extern "C" {

mplbm_status mplbm_place_spheres_noalloc(int64_t lx, int64_t ly, int64_t lz, const float* spheres, int64_t count,
                                         int64_t first_id, uint8_t* is_fluid_dev, int64_t* solid_number_dev) {
  using namespace mplbm;
  if (lx < 1 || ly < 1 || lz < 1 || lx > INT32_MAX || ly > INT32_MAX || lz > INT32_MAX)
    return fail(nullptr, MPLBM_ERR_ARG, "place_spheres_noalloc: bad domain %ld x %ld x %ld", (long)lx, (long)ly,
                (long)lz);
  if (count < 0 || (count > 0 && !spheres) || !is_fluid_dev)
    return fail(nullptr, MPLBM_ERR_ARG, "place_spheres_noalloc: null array or negative count");
  const uint64_t n = (uint64_t)lx * (uint64_t)ly * (uint64_t)lz;
  if ((uint64_t)count > kNoAllocMaxSpheres)
    return fail(nullptr, MPLBM_ERR_ARG, "place_spheres_noalloc: %ld spheres exceed the scratch limit of %llu",
                (long)count, (unsigned long long)kNoAllocMaxSpheres);
  if (solid_number_dev && count > 0 && n > kNoAllocMaxCells)
    return fail(nullptr, MPLBM_ERR_ARG, "place_spheres_noalloc: %llu cells exceed the scratch limit of %llu cells "
                "with solid numbers", (unsigned long long)n, (unsigned long long)kNoAllocMaxCells);
  if (const mplbm_status cs = check_spheres(spheres, count); cs != MPLBM_OK) return cs;
  if (count == 0) return MPLBM_OK;
  void* base = nullptr;
  if (const mplbm_status sa = scratch_address(&base); sa != MPLBM_OK) return sa;
  uint32_t* ids = solid_number_dev ? static_cast<uint32_t*>(base) : nullptr;
  float4* sph = reinterpret_cast<float4*>(static_cast<unsigned char*>(base) + kNoAllocIdsBytes);
  auto hip_fail = [](hipError_t e, const char* what) {
    return fail(nullptr, MPLBM_ERR_HIP, "place_spheres_noalloc: %s: %s", what, hipGetErrorString(e));
  };
  hipError_t e = hipMemcpyAsync(sph, spheres, (size_t)count * sizeof(float4), hipMemcpyHostToDevice, 0);
  if (e != hipSuccess) return hip_fail(e, "stage spheres");
  // the sphere numbers of an earlier call are still in the scratch
  if (ids) hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16)), dim3(256),
                              0, 0, ids, n, 0u);
  hipLaunchKernelGGL(k_voxelise, dim3((unsigned)std::min<int64_t>(count, 1 << 16)), dim3(kVoxThreads), 0, 0, sph,
                     count, (int)lx, (int)ly, (int)lz, is_fluid_dev, ids);
  if (ids)
    hipLaunchKernelGGL(k_vox_numbers, dim3((unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16)), dim3(256), 0, 0,
                       ids, solid_number_dev, first_id - 1, n);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "launch");
  if ((e = hipStreamSynchronize(0)) != hipSuccess) return hip_fail(e, "synchronise");
  return MPLBM_OK;
}

}  // extern "C"
