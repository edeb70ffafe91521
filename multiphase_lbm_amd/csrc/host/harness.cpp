// Host harness of the MI355X D3Q19 Shan-Chen program: the reference's host
// application (struct MultiphaseLBM, include/multiphase_lbm.h:26-218,
// src/multiphase_lbm.cu) with runtime domain sizes, driving libmplbm
// through its C ABI instead of CUDA unified memory.
//
// Kept compatible with the reference: input tokens and their order
// semantics, the positions/sample/VTK formats, the srand(1) density
// initialisation, the float sphere voxeliser and the Stats.txt columns.
// Differences (each noted where it happens): cluster search is iterative
// (the reference recurses once per node, src/multiphase_lbm.cu:163-223),
// phase sums accumulate in double, statistics read fields after an explicit
// synchronisation, and erosion refuses a target above the packing density
// (the reference casts a negative count to size_t, :852).

#include "mplbm_host.h"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "mplbm.h"

namespace {

// last error text; the field-file writer thread may set it while the
// driver runs, so it is guarded
std::mutex g_err_mu;
std::string g_err;

int set_err(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = buf;
  return -1;
}

// D3Q19 velocities in the reference order (include/lbm_model.cuh:49-54)
const int EX[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
const int EY[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
const int EZ[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};

struct Dims {
  int64_t lx, ly, lz;
  size_t n() const { return (size_t)lx * ly * lz; }
  size_t at(int64_t x, int64_t y, int64_t z) const { return ((size_t)x * ly + y) * lz + z; }
  int64_t wx(int64_t v) const { return v < 0 ? v + lx : (v >= lx ? v - lx : v); }
  int64_t wy(int64_t v) const { return v < 0 ? v + ly : (v >= ly ? v - ly : v); }
  int64_t wz(int64_t v) const { return v < 0 ? v + lz : (v >= lz ? v - lz : v); }
};

// nearest() of include/lbm_model.cuh:58: floor(x + 0.5) with a double 0.5
inline int nearest(float x) { return (int)std::floor((double)x + 0.5); }

void place_sphere(const Dims& d, uint8_t* is_fluid, int64_t* solid_number, float x, float y, float z, float R,
                  int64_t solid_id) {
  const float xmin = x - R, ymin = y - R, zmin = z - R;
  const float xmax = x + R, ymax = y + R, zmax = z + R;
  for (float px = xmin; px <= xmax; px += 1.f) {
    for (float py = ymin; py <= ymax; py += 1.f) {
      for (float pz = zmin; pz <= zmax; pz += 1.f) {
        const float dx = px - x, dy = py - y, dz = pz - z;
        const float dist2 = dx * dx + dy * dy + dz * dz;
        const float R2 = R * R;
        if (dist2 < R2) {
          int nx = nearest(px) < 0 ? nearest(px) + (int)d.lx : nearest(px);
          int ny = nearest(py) < 0 ? nearest(py) + (int)d.ly : nearest(py);
          int nz = nearest(pz) < 0 ? nearest(pz) + (int)d.lz : nearest(pz);
          if (nx >= d.lx) nx -= (int)d.lx;
          if (ny >= d.ly) ny -= (int)d.ly;
          if (nz >= d.lz) nz -= (int)d.lz;
          if (nx < 0 || ny < 0 || nz < 0 || nx >= d.lx || ny >= d.ly || nz >= d.lz) continue;  // beyond one wrap
          const size_t c = d.at(nx, ny, nz);
          is_fluid[c] = 0;
          if (solid_number) solid_number[c] = solid_id;
        }
      }
    }
  }
}

// values of a positions file (read_positions, :1106-1126): one number per
// line, x, y, z, r of each sphere in turn, each times res as a double
// product narrowed to float (:1113-1119); an incomplete last sphere is
// dropped.  Returns the sphere count or -1.
int64_t read_positions_values(const char* path, float res, std::vector<float>& v) {
  std::ifstream f(path);
  if (!f) return set_err("cannot open positions file %s", path);
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    v.push_back((float)(std::stod(line) * (double)res));
  }
  v.resize(v.size() / 4 * 4);
  return (int64_t)(v.size() / 4);
}

// splitmix64: seeded uniform numbers for the synthetic packings
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// ---------------------------------------------------------------------------
// cluster statistics (calc_cluster, src/multiphase_lbm.cu:225-562)

struct ClusterResult {
  mplbm_cluster_stats s{};
};

int64_t flood(const Dims& d, std::vector<uint8_t>& member, std::vector<int64_t>& label, int64_t id, size_t start,
              const int64_t* solid_number, std::set<int64_t>* solids_touched, std::vector<size_t>& stack) {
  // 6-neighbour periodic flood fill (depth_first_search, :163-223)
  int64_t size = 0;
  stack.clear();
  stack.push_back(start);
  member[start] = 0;
  label[start] = id;
  while (!stack.empty()) {
    const size_t c = stack.back();
    stack.pop_back();
    ++size;
    const int64_t x = (int64_t)(c / ((size_t)d.ly * d.lz));
    const int64_t y = (int64_t)((c / d.lz) % d.ly);
    const int64_t z = (int64_t)(c % d.lz);
    const size_t nb[6] = {d.at(d.wx(x - 1), y, z), d.at(x, d.wy(y - 1), z), d.at(x, y, d.wz(z - 1)),
                          d.at(d.wx(x + 1), y, z), d.at(x, d.wy(y + 1), z), d.at(x, y, d.wz(z + 1))};
    for (size_t n : nb) {
      if (member[n]) {
        member[n] = 0;
        label[n] = id;
        stack.push_back(n);
      } else if (solids_touched && solid_number && solid_number[n] != 0) {
        solids_touched->insert(solid_number[n]);
      }
    }
  }
  return size;
}

int cluster_stats(const Dims& d, const uint8_t* is_fluid, const float* rho, const float* pres,
                  const int64_t* solid_number, double rhol, double rhog, int64_t solid_count, int64_t* liq_label_out,
                  int64_t* gas_label_out, mplbm_cluster_stats* out) {
  const size_t n = d.n();
  mplbm_cluster_stats s{};
  std::vector<uint8_t> liq(n, 0), gas(n, 0);
  std::vector<int64_t> liq_label(n, 0), gas_label(n, 0);
  int64_t fluid = 0, nl = 0, nv = 0;
  double lp = 0, vp = 0, lr = 0, vr = 0;
  float min_r = 1e9f, max_r = -1e-9f, min_p = 1e9f, max_p = -1e-9f;  // write_stats init, :1058-1059
  for (size_t c = 0; c < n; ++c) {
    if (!is_fluid[c]) continue;
    ++fluid;
    const float r = rho[c];
    min_r = std::min(r, min_r);
    max_r = std::max(r, max_r);
    min_p = std::min(pres[c], min_p);
    max_p = std::max(pres[c], max_p);
    if (r > rhol) {
      lp += pres[c]; lr += r; ++nl; liq[c] = 1;
    } else if (r < rhog) {
      vp += pres[c]; vr += r; ++nv; gas[c] = 1;
    }
  }
  s.min_rho = min_r; s.max_rho = max_r; s.min_p = min_p; s.max_p = max_p;
  s.fluid_count = fluid; s.liquid_count = nl; s.vapor_count = nv;
  if (nl) { s.avg_liq_press = (float)(lp / nl); s.avg_liq_rho = (float)(lr / nl); }
  if (nv) { s.avg_vap_press = (float)(vp / nv); s.avg_vap_rho = (float)(vr / nv); }
  s.saturation = fluid ? (double)((float)nl / (float)fluid) : 0.0;

  // liquid-solid / liquid-vapour boundary nodes (:277-316)
  int64_t nls = 0, nlv = 0;
  for (int64_t x = 0; x < d.lx; ++x)
    for (int64_t y = 0; y < d.ly; ++y)
      for (int64_t z = 0; z < d.lz; ++z) {
        const size_t c = d.at(x, y, z);
        if (!liq[c]) continue;
        for (int i = 1; i < 19; ++i) {
          const size_t u = d.at(d.wx(x - EX[i]), d.wy(y - EY[i]), d.wz(z - EZ[i]));
          if (!is_fluid[u]) { ++nls; break; }
          if (rho[u] < rhol) { ++nlv; break; }
        }
      }
  if (nl > 0) {
    s.ratio_liquid_solid_boundary = (float)((double)nls / nl);
    s.ratio_liquid_vapor_boundary = (float)((double)nlv / nl);
  }

  // clusters (:318-360); fewer than 7 nodes are discarded
  const int64_t kMin = 7;
  std::vector<size_t> stack;
  std::map<int64_t, std::set<int64_t>> cluster_solids;
  int64_t lcount = 0;
  for (size_t c = 0; c < n; ++c) {
    if (!liq[c]) continue;
    ++lcount;
    std::set<int64_t> touched;
    const int64_t sz = flood(d, liq, liq_label, lcount, c, solid_number, &touched, stack);
    if (sz < kMin) {
      // relabel to 0: walk the labels just written (cluster is tiny)
      std::vector<uint8_t> tmp;  // not needed: re-flood on labels
      std::vector<size_t> st{c};
      liq_label[c] = 0;
      while (!st.empty()) {
        const size_t q = st.back();
        st.pop_back();
        const int64_t x = (int64_t)(q / ((size_t)d.ly * d.lz)), y = (int64_t)((q / d.lz) % d.ly),
                      z = (int64_t)(q % d.lz);
        const size_t nb[6] = {d.at(d.wx(x - 1), y, z), d.at(x, d.wy(y - 1), z), d.at(x, y, d.wz(z - 1)),
                              d.at(d.wx(x + 1), y, z), d.at(x, d.wy(y + 1), z), d.at(x, y, d.wz(z + 1))};
        for (size_t m : nb)
          if (liq_label[m] == lcount) { liq_label[m] = 0; st.push_back(m); }
      }
      --lcount;
    } else if (!touched.empty()) {
      cluster_solids[lcount] = std::move(touched);
    }
  }
  int64_t gcount = 0;
  for (size_t c = 0; c < n; ++c) {
    if (!gas[c]) continue;
    ++gcount;
    const int64_t sz = flood(d, gas, gas_label, gcount, c, nullptr, nullptr, stack);
    if (sz < kMin) {
      std::vector<size_t> st{c};
      gas_label[c] = 0;
      while (!st.empty()) {
        const size_t q = st.back();
        st.pop_back();
        const int64_t x = (int64_t)(q / ((size_t)d.ly * d.lz)), y = (int64_t)((q / d.lz) % d.ly),
                      z = (int64_t)(q % d.lz);
        const size_t nb[6] = {d.at(d.wx(x - 1), y, z), d.at(x, d.wy(y - 1), z), d.at(x, y, d.wz(z - 1)),
                              d.at(d.wx(x + 1), y, z), d.at(x, d.wy(y + 1), z), d.at(x, y, d.wz(z + 1))};
        for (size_t m : nb)
          if (gas_label[m] == gcount) { gas_label[m] = 0; st.push_back(m); }
      }
      --gcount;
    }
  }
  s.liq_cluster_count = lcount;
  s.gas_cluster_count = gcount;

  // single liquid cluster: periodic unwrapping, centre and radius (:362-533)
  if (lcount == 1) {
    bool brk[3] = {false, false, false};
    for (int64_t y = 0; y < d.ly && !brk[0]; ++y)
      for (int64_t z = 0; z < d.lz; ++z)
        if (liq_label[d.at(0, y, z)] == 1 && liq_label[d.at(d.lx - 1, y, z)] == 1) { brk[0] = true; break; }
    for (int64_t x = 0; x < d.lx && !brk[1]; ++x)
      for (int64_t z = 0; z < d.lz; ++z)
        if (liq_label[d.at(x, 0, z)] == 1 && liq_label[d.at(x, d.ly - 1, z)] == 1) { brk[1] = true; break; }
    for (int64_t x = 0; x < d.lx && !brk[2]; ++x)
      for (int64_t y = 0; y < d.ly; ++y)
        if (liq_label[d.at(x, y, 0)] == 1 && liq_label[d.at(x, y, d.lz - 1)] == 1) { brk[2] = true; break; }
    // first plane free of the cluster along each broken axis
    int64_t vplane[3] = {0, 0, 0};
    const int64_t L[3] = {d.lx, d.ly, d.lz};
    for (int a = 0; a < 3; ++a) {
      if (!brk[a]) continue;
      for (int64_t k = 0; k < L[a]; ++k) {
        bool free_plane = true;
        for (int64_t p = 0; p < L[(a + 1) % 3] && free_plane; ++p)
          for (int64_t q = 0; q < L[(a + 2) % 3]; ++q) {
            int64_t co[3];
            co[a] = k; co[(a + 1) % 3] = p; co[(a + 2) % 3] = q;
            if (liq_label[d.at(co[0], co[1], co[2])] == 1) { free_plane = false; break; }
          }
        if (free_plane) { vplane[a] = k; break; }
      }
    }
    uint64_t sum[3] = {0, 0, 0}, size = 0;
    for (int64_t x = 0; x < d.lx; ++x)
      for (int64_t y = 0; y < d.ly; ++y)
        for (int64_t z = 0; z < d.lz; ++z)
          if (liq_label[d.at(x, y, z)] == 1) {
            const int64_t co[3] = {x, y, z};
            for (int a = 0; a < 3; ++a) sum[a] += (uint64_t)(co[a] + ((brk[a] && co[a] < vplane[a]) ? L[a] : 0));
            ++size;
          }
    // integer division as in the reference (size_t sums, :496-498)
    const float ctr[3] = {(float)(sum[0] / size), (float)(sum[1] / size), (float)(sum[2] / size)};
    float maxr = 0.f;
    for (int64_t x = 0; x < d.lx; ++x)
      for (int64_t y = 0; y < d.ly; ++y)
        for (int64_t z = 0; z < d.lz; ++z)
          if (liq_label[d.at(x, y, z)] == 1) {
            const int64_t co[3] = {x, y, z};
            float r2 = 0.f;
            for (int a = 0; a < 3; ++a) {
              const float adj = (float)(co[a] + ((brk[a] && co[a] < vplane[a]) ? L[a] : 0));
              r2 += (ctr[a] - adj) * (ctr[a] - adj);
            }
            maxr = std::max(maxr, std::sqrt(r2));
          }
    s.cluster_radius = maxr;
  }

  // cluster order and wet coordination number (:535-561)
  float order = 0.f;
  std::map<int64_t, std::set<int64_t>> solid_clusters;
  for (auto& kv : cluster_solids) {
    order += (float)kv.second.size();
    for (int64_t sid : kv.second) solid_clusters[sid].insert(kv.first);
  }
  if (lcount != 0 && !cluster_solids.empty()) order /= (float)cluster_solids.size();
  s.cluster_order = order;
  float wet = 0.f;
  for (auto& kv : solid_clusters) wet += (float)kv.second.size();
  if (solid_count != 0) wet /= (float)solid_count;
  s.wet_coord_num = wet;

  if (liq_label_out) std::copy(liq_label.begin(), liq_label.end(), liq_label_out);
  if (gas_label_out) std::copy(gas_label.begin(), gas_label.end(), gas_label_out);
  *out = s;
  return 0;
}

int write_vtk(const char* path, const Dims& d, const uint8_t* is_fluid, const float* rho, const float* pres,
              const int64_t* liq, const int64_t* gas, const int64_t* solid) {
  FILE* f = std::fopen(path, "w");
  if (!f) return set_err("cannot open %s", path);
  const float pas = (float)(1.0 / (double)d.lx);  // pasxyz = 1.0/lx for every axis, :953
  std::fprintf(f, "# vtk DataFile Version 2.0\n");
  std::fprintf(f, "Sortie domaine LB+LINK\n");
  std::fprintf(f, "ASCII\n");
  std::fprintf(f, "DATASET RECTILINEAR_GRID\n");
  std::fprintf(f, "DIMENSIONS %zu %zu %zu\n", (size_t)d.lx, (size_t)d.ly, (size_t)d.lz);
  std::fprintf(f, "X_COORDINATES %zu float\n", (size_t)d.lx);
  for (int64_t i = 0; i < d.lx; ++i) std::fprintf(f, "%.4e ", (double)((float)i * pas));
  std::fprintf(f, "\n");
  std::fprintf(f, "Y_COORDINATES %zu float\n", (size_t)d.ly);
  for (int64_t i = 0; i < d.ly; ++i) std::fprintf(f, "%.4e ", (double)((float)i * pas));
  std::fprintf(f, "\n");
  std::fprintf(f, "Z_COORDINATES %zu float\n", (size_t)d.lz);
  for (int64_t i = 0; i < d.lz; ++i) std::fprintf(f, "%e ", (double)((float)i * pas));
  std::fprintf(f, "\n");
  std::fprintf(f, "POINT_DATA %zu\n", d.n());
  auto block = [&](const char* head, auto value) {
    std::fprintf(f, "%s", head);
    std::fprintf(f, "LOOKUP_TABLE default\n");
    for (int64_t z = 0; z < d.lz; ++z)
      for (int64_t y = 0; y < d.ly; ++y)
        for (int64_t x = 0; x < d.lx; ++x) value(d.at(x, y, z));
  };
  block("SCALARS Zones int 1\n", [&](size_t c) { std::fprintf(f, "%d\n", is_fluid[c] ? 0 : 1); });
  block("SCALARS Density float 1\n", [&](size_t c) {
    if (is_fluid[c]) std::fprintf(f, "%.4e\n", (double)rho[c]);
    else std::fprintf(f, "0\n");
  });
  block("SCALARS Pressure float 1\n", [&](size_t c) {
    if (is_fluid[c]) std::fprintf(f, "%.4e\n", (double)pres[c]);
    else std::fprintf(f, "0\n");
  });
  block("SCALARS LiqClusters int 1\n", [&](size_t c) { std::fprintf(f, "%ld\n", (long)(liq ? liq[c] : 0)); });
  block("SCALARS GasClusters int 1\n", [&](size_t c) { std::fprintf(f, "%ld\n", (long)(gas ? gas[c] : 0)); });
  block("SCALARS Solids int 1\n", [&](size_t c) { std::fprintf(f, "%ld\n", (long)(solid ? solid[c] : 0)); });
  std::fclose(f);
  return 0;
}

// Binary VTK XML RectilinearGrid with the point arrays of write_vtk (raw
// appended data, UInt64 block sizes, little endian, x fastest).
int write_vtr(const char* path, const Dims& d, const uint8_t* is_fluid, const float* rho, const float* pres,
              const int64_t* liq, const int64_t* gas, const int64_t* solid) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_err("cannot open %s", path);
  const size_t n = d.n();
  struct Arr {
    const char* type;
    const char* name;
    size_t count, esize;
  };
  const Arr arrs[] = {{"Int32", "Zones", n, 4},        {"Float32", "Density", n, 4}, {"Float32", "Pressure", n, 4},
                      {"Int64", "LiqClusters", n, 8},  {"Int64", "GasClusters", n, 8}, {"Int64", "Solids", n, 8},
                      {"Float32", "x", (size_t)d.lx, 4}, {"Float32", "y", (size_t)d.ly, 4},
                      {"Float32", "z", (size_t)d.lz, 4}};
  size_t off[9], o = 0;
  for (int k = 0; k < 9; ++k) {
    off[k] = o;
    o += 8 + arrs[k].count * arrs[k].esize;
  }
  std::fprintf(f, "<?xml version=\"1.0\"?>\n");
  std::fprintf(f, "<VTKFile type=\"RectilinearGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
                  "header_type=\"UInt64\">\n");
  std::fprintf(f, "  <RectilinearGrid WholeExtent=\"0 %ld 0 %ld 0 %ld\">\n", (long)d.lx - 1, (long)d.ly - 1,
               (long)d.lz - 1);
  std::fprintf(f, "    <Piece Extent=\"0 %ld 0 %ld 0 %ld\">\n", (long)d.lx - 1, (long)d.ly - 1, (long)d.lz - 1);
  std::fprintf(f, "      <PointData Scalars=\"Density\">\n");
  for (int k = 0; k < 6; ++k)
    std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" format=\"appended\" offset=\"%zu\"/>\n",
                 arrs[k].type, arrs[k].name, off[k]);
  std::fprintf(f, "      </PointData>\n      <Coordinates>\n");
  for (int k = 6; k < 9; ++k)
    std::fprintf(f, "        <DataArray type=\"%s\" Name=\"%s\" format=\"appended\" offset=\"%zu\"/>\n",
                 arrs[k].type, arrs[k].name, off[k]);
  std::fprintf(f, "      </Coordinates>\n    </Piece>\n  </RectilinearGrid>\n");
  std::fprintf(f, "  <AppendedData encoding=\"raw\">\n   _");
  // point order x fastest, written in runs of lx values
  std::vector<unsigned char> buf((size_t)d.lx * 8);
  auto points = [&](size_t esize, auto value) {
    const uint64_t bytes = (uint64_t)n * esize;
    std::fwrite(&bytes, 8, 1, f);
    for (int64_t z = 0; z < d.lz; ++z)
      for (int64_t y = 0; y < d.ly; ++y) {
        for (int64_t x = 0; x < d.lx; ++x) value(d.at(x, y, z), buf.data() + (size_t)x * esize);
        std::fwrite(buf.data(), esize, (size_t)d.lx, f);
      }
  };
  auto put32 = [](unsigned char* q, const void* v) { std::memcpy(q, v, 4); };
  auto put64 = [](unsigned char* q, const void* v) { std::memcpy(q, v, 8); };
  points(4, [&](size_t c, unsigned char* q) { const int32_t v = is_fluid[c] ? 0 : 1; put32(q, &v); });
  points(4, [&](size_t c, unsigned char* q) { const float v = is_fluid[c] ? rho[c] : 0.f; put32(q, &v); });
  points(4, [&](size_t c, unsigned char* q) { const float v = is_fluid[c] ? pres[c] : 0.f; put32(q, &v); });
  points(8, [&](size_t c, unsigned char* q) { const int64_t v = liq ? liq[c] : 0; put64(q, &v); });
  points(8, [&](size_t c, unsigned char* q) { const int64_t v = gas ? gas[c] : 0; put64(q, &v); });
  points(8, [&](size_t c, unsigned char* q) { const int64_t v = solid ? solid[c] : 0; put64(q, &v); });
  const float pas = (float)(1.0 / (double)d.lx);  // write_vtk's coordinates i/lx on every axis
  for (int64_t L : {d.lx, d.ly, d.lz}) {
    const uint64_t bytes = (uint64_t)L * 4;
    std::fwrite(&bytes, 8, 1, f);
    for (int64_t i = 0; i < L; ++i) {
      const float v = (float)i * pas;
      std::fwrite(&v, 4, 1, f);
    }
  }
  std::fprintf(f, "\n  </AppendedData>\n</VTKFile>\n");
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  return bad ? set_err("write error on %s", path) : 0;
}

// ---------------------------------------------------------------------------
// the MultiphaseLBM program (include/multiphase_lbm.h:26-218)

struct MultiphaseLBM {
  mplbm_input in{};
  Dims d{400, 400, 400};
  bool allocated = false;
  std::vector<float> rho, pres;
  std::vector<uint8_t> is_fluid;
  std::vector<int64_t> solid_number, liq_label, gas_label;
  int64_t solid_count = 0;
  int64_t solids = 0;
  int64_t eroded_nodes = 0;
  float packing_density = 0.f;
  std::ofstream stat_file;

  void resize() {  // resize_gpu (src/multiphase_lbm.cu:83-157), host part
    d = Dims{in.lx, in.ly, in.lz};
    const size_t n = d.n();
    std::cout << "Setting domain size to: " << d.lx << "x" << d.ly << "x" << d.lz << std::endl;
    rho.assign(n, 0.f);
    pres.assign(n, 0.f);
    is_fluid.assign(n, 1);
    solid_number.assign(n, 0);
    liq_label.assign(n, 0);
    gas_label.assign(n, 0);
    if (in.read_densities) {
      std::cout << "Reading density data from file: " << in.densities_filename << std::endl;
      mplbm_host_read_densities(in.densities_filename, rho.data(), d.lx, d.ly, d.lz);
    } else {
      mplbm_host_rho_init(rho.data(), d.lx, d.ly, d.lz, (float)in.rho0, (float)in.IniPerturbRate);
    }
    allocated = true;
  }

  void calc_packing_density() {  // :1129-1142
    solids = 0;
    for (uint8_t f : is_fluid) solids += f ? 0 : 1;
    packing_density = (float)solids / (float)(d.n());
    std::cout << "Packing density: " << packing_density << std::endl;
  }

  std::vector<uint8_t> solid_touch_liquid() {  // find_solid_touch_liquid, :810-838
    std::vector<uint8_t> t(d.n(), 0);
    for (int64_t x = 0; x < d.lx; ++x)
      for (int64_t y = 0; y < d.ly; ++y)
        for (int64_t z = 0; z < d.lz; ++z) {
          const size_t c = d.at(x, y, z);
          if (is_fluid[c]) continue;
          for (int i = 1; i < 19; ++i)
            if (is_fluid[d.at(d.wx(x - EX[i]), d.wy(y - EY[i]), d.wz(z - EZ[i]))]) { t[c] = 1; break; }
        }
    return t;
  }

  // erode_solid_touch_liquid, :842-890.  The eroded count is the
  // reference's float expression (1 - target_density/packing_density) *
  // solids, truncated.  While it exceeds the number of solids touching the
  // fluid, whole layers of touching solids turn fluid; then each touching
  // solid turns fluid when rand()/RAND_MAX < eroded/touching (float), in the
  // reference's z, y, x scan order, so the draws follow its rand() stream.
  // One deliberate difference: the touching count is recounted after every
  // whole layer.  The reference counts it once and only decrements it; its
  // size_t counter wraps below zero in the second layer, after which the loop
  // either stops with two whole layers eroded and a random pass of rate ~0
  // (target at or beyond two layers) or never ends (target inside the second
  // layer).  A target within the first layer takes the same path in both.
  // Returns the number of solids turned fluid.
  int64_t erode() {
    std::vector<uint8_t> touch = solid_touch_liquid();
    int64_t solid_liquid_num = 0;
    for (uint8_t t : touch) solid_liquid_num += t;
    const float target = (float)in.target_density;
    if (packing_density <= 0.f || target >= packing_density) {
      std::cerr << "target_density " << target << " is not below the packing density " << packing_density
                << "; no erosion" << std::endl;
      return 0;
    }
    int64_t eroded = (int64_t)((1.0f - target / packing_density) * (float)solids);
    std::cout << "Eroded solids number: " << eroded << std::endl;
    std::cout << "Solid touch liquid number: " << solid_liquid_num << std::endl;
    int64_t done = 0;
    while (eroded > solid_liquid_num) {
      for (int64_t z = 0; z < d.lz; ++z)
        for (int64_t y = 0; y < d.ly; ++y)
          for (int64_t x = 0; x < d.lx; ++x) {
            const size_t c = d.at(x, y, z);
            if (touch[c]) { touch[c] = 0; is_fluid[c] = 1; --eroded; ++done; }
          }
      touch = solid_touch_liquid();
      solid_liquid_num = 0;
      for (uint8_t t : touch) solid_liquid_num += t;
      if (solid_liquid_num == 0) break;  // no solid left next to the fluid
    }
    if (eroded <= solid_liquid_num && eroded > 0) {
      const float rate = (float)((double)eroded / (double)solid_liquid_num);
      for (int64_t z = 0; z < d.lz; ++z)
        for (int64_t y = 0; y < d.ly; ++y)
          for (int64_t x = 0; x < d.lx; ++x) {
            const size_t c = d.at(x, y, z);
            if (touch[c] && (float)rand() / (float)RAND_MAX < rate) { touch[c] = 0; is_fluid[c] = 1; ++done; }
          }
    }
    calc_packing_density();
    return done;
  }

  int read_data(const char* path) {  // :619-745
    mplbm_host_input_default(&in);
    std::ifstream file(path);
    if (file.fail()) return set_err("Cannot open command file %s", path);
    std::string token;
    file >> token;
    while (file.good()) {
      if (token[0] == '!' || token[0] == '#') {
        std::getline(file, token);
      } else if (token == "resize") {
        resize();
      } else if (token == "read_sample") {
        std::string name;
        file >> name;
        std::cout << "Reading node ids from: " << name << std::endl;
        if (!allocated) resize();
        if (mplbm_host_read_sample(name.c_str(), is_fluid.data(), d.lx, d.ly, d.lz) != 0) return -1;
        calc_packing_density();
      } else if (token == "place_sphere") {
        float x, y, z, R;
        file >> x >> y >> z >> R;
        if (!allocated) return set_err("place_sphere before resize");
        place_sphere(d, is_fluid.data(), solid_number.data(), x, y, z, R, solid_count);
        std::cout << "Placing sphere with radius " << R << " at (" << x << ", " << y << ", " << z << ")" << std::endl;
      } else if (token == "generate_solids") {  // :1086-1103
        int num;
        float rmin, rmax, xmin, xmax, ymin, ymax, zmin, zmax;
        file >> num >> rmin >> rmax >> xmin >> xmax >> ymin >> ymax >> zmin >> zmax;
        if (!allocated) return set_err("generate_solids before resize");
        std::vector<float> v;
        for (int i = 0; i < num; ++i) {
          const float r = (rmax - rmin) * ((float)rand() / (float)RAND_MAX) + rmin;
          const float x = (xmax - xmin) * ((float)rand() / (float)RAND_MAX) + xmin;
          const float y = (ymax - ymin) * ((float)rand() / (float)RAND_MAX) + ymin;
          const float z = (zmax - zmin) * ((float)rand() / (float)RAND_MAX) + zmin;
          v.insert(v.end(), {x, y, z, r});
        }
        if (in.device_voxeliser && num > 0) {
          if (mplbm_place_spheres(d.lx, d.ly, d.lz, v.data(), num, solid_count + 1, is_fluid.data(),
                                  solid_number.data()) != MPLBM_OK)
            return set_err("%s", mplbm_last_error(nullptr));
        } else {
          for (int i = 0; i < num; ++i)
            place_sphere(d, is_fluid.data(), solid_number.data(), v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3],
                         solid_count + 1 + i);
        }
        solid_count += std::max(num, 0);
      } else if (token == "read_positions") {
        std::string name;
        file >> name;
        if (!allocated) return set_err("read_positions before resize");
        std::cout << "Reading grain positions and radii from: " << name << "\nwith resolution of: " << in.res
                  << std::endl;
        const int64_t k =
            (in.device_voxeliser ? mplbm_host_read_positions_device : mplbm_host_read_positions)(
                name.c_str(), (float)in.res, is_fluid.data(), solid_number.data(), d.lx, d.ly, d.lz, solid_count + 1);
        if (k < 0) return -1;
        solid_count += k;
        std::cout << "number of grain = " << solid_count << std::endl;
        calc_packing_density();
      } else if (!parse_scalar(file, token)) {
        std::cerr << "The token " << token << " is undefined." << std::endl;
        in.unknown_tokens++;
      }
      file >> token;
    }
    std::cout << "rho0: " << in.rho0 << std::endl;
    std::cout << "perturbation rate: " << in.IniPerturbRate << std::endl;
    std::cout << "rhos: " << in.rhos << std::endl;
    if (in.carn_star)
      std::cout << "Using Carnahan-Starling EOS with T/Tc = " << in.T_Tc << ", rhol* = " << in.rhol_spinodal
                << " and rhog* = " << in.rhog_spinodal << std::endl;
    else
      std::cout << "Using Shan-Chen EOS with G = " << in.G << std::endl;
    std::cout << "Injecting density with an increment of " << in.rho_increment << " every "
              << in.rho_inject_period << " step." << std::endl;
    std::cout << "Injection type: " << in.inject_type << std::endl;
    if (!allocated) return set_err("no domain: the input has neither resize nor read_sample");
    eroded_nodes = in.erosion ? erode() : 0;
    return 0;
  }

  // scalar tokens; returns false for an unknown token
  bool parse_scalar(std::ifstream& file, const std::string& t) { return mplbm_parse_scalar_token(file, t, in); }

  static bool mplbm_parse_scalar_token(std::istream& file, const std::string& token, mplbm_input& in);

  int init() {  // :930-944
    mkdir(in.result_folder, 0755);
    std::cout << "Resulting files will be saved in " << in.result_folder << std::endl;
    const std::string name = std::string(in.result_folder) + "/Stats.txt";
    stat_file.open(name);
    if (!stat_file) return set_err("cannot open %s", name.c_str());
    stat_file << "step\tmin_rho\tmax_rho\tmin_p\tmax_p\tavg_liq_press\tavg_vap_press\tsuction\tSr\tavg_liq_rho\t"
                 "avg_vap_rho\tliq_cluster_count\tcluster_order\twet_coord_num\tgas_cluster_count\tliquid_solid_"
                 "boundary\tliquid_vapor_boundary\tcluster_radius"
              << std::endl;
    return 0;
  }

  int run() {  // run_gpu, :565-614
    mplbm_config cfg;
    mplbm_config_default(&cfg);
    cfg.lx = d.lx; cfg.ly = d.ly; cfg.lz = d.lz;
    cfg.storage = in.storage;
    cfg.kernel = in.kernel;
    cfg.math = in.math;
    cfg.rounding = in.rounding;
    cfg.shift = in.shift;
    cfg.tau = (float)in.tau;
    cfg.G = (float)in.G;
    cfg.carn_star = in.carn_star;
    cfg.T_Tc = (float)in.T_Tc;
    cfg.rhol_spinodal = (float)in.rhol_spinodal;
    cfg.rhog_spinodal = (float)in.rhog_spinodal;
    cfg.rho_inject_period = in.rho_inject_period;
    cfg.rho_increment = (float)in.rho_increment;
    cfg.rhos = (float)in.rhos;
    cfg.inject_position = (float)in.inject_position;
    for (int f = 0; f < 6; ++f) { cfg.face_type[f] = in.face_type[f]; cfg.face_rho[f] = in.face_rho[f]; }
    for (int a = 0; a < 3; ++a) cfg.body_force[a] = in.body_force[a];
    mplbm_ctx* ctx = nullptr;
    if (mplbm_create(&cfg, &ctx) != MPLBM_OK) return set_err("mplbm_create: %s", mplbm_last_error(nullptr));
    auto check = [&](mplbm_status st) {
      if (st != MPLBM_OK) { set_err("%s", mplbm_last_error(ctx)); return false; }
      return true;
    };
    int rc = 0;
    int64_t step = in.step;
    int resume_reverse = 1;
    if (!check(mplbm_set_geometry(ctx, is_fluid.data()))) {
      mplbm_destroy(ctx);
      return -1;
    }
    if (in.read_checkpoint) {  // exact restart (the reference restarts from densities only)
      if (!check(mplbm_load_checkpoint(ctx, in.checkpoint_filename, &step))) {
        mplbm_destroy(ctx);
        return -1;
      }
      if (read_driver_state(in.checkpoint_filename, &resume_reverse) != 0) {
        mplbm_destroy(ctx);
        return -1;
      }
      std::cout << "Resuming from checkpoint " << in.checkpoint_filename << " at step " << step
                << (resume_reverse < 0 ? " (draining)" : "") << std::endl;
    } else if (!check(mplbm_init_fields(ctx, rho.data()))) {
      mplbm_destroy(ctx);
      return -1;
    }
    if (!in.host_statistics && !check(mplbm_set_solid_numbers(ctx, solid_number.data(), solid_count))) {
      mplbm_destroy(ctx);
      return -1;
    }
    std::vector<int32_t> liq32, gas32;
    // reverse = 1 injects, -1 drains (src/multiphase_lbm.cu:571-573).  The
    // reference parses max_saturation / min_saturation (:712-716) but never
    // changes reverse.  With saturation_cycle 1 the driver cycles: inject
    // until the saturation reaches max_saturation, then drain until it falls
    // to min_saturation, and so on, switching at statistics steps.  The
    // direction is driver state: it is saved beside each checkpoint and
    // restored on resume.
    int reverse = resume_reverse;
    const int64_t stat_period = in.statPeriod > 0 ? in.statPeriod : 1;
    const int64_t ck_period = in.checkpointPeriod;
    size_t vtk_num = 0, vtr_num = 0;
    int64_t fluid_nodes = 0;
    for (uint8_t f : is_fluid) fluid_nodes += f;
    // field files are written by a thread of their own while the GPU steps
    // on; rho, pres and the label arrays are not touched until it is joined
    std::thread writer;
    int writer_rc = 0;
    auto join_writer = [&]() {
      if (writer.joinable()) writer.join();
      if (writer_rc != 0) std::fprintf(stderr, "field file: %s\n", mplbm_host_last_error());
      writer_rc = 0;
    };
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t step0 = step;
    while (step < in.step_max) {
      // advance to the next statistics or checkpoint step in one call
      const int64_t next_stat = ((step + stat_period - 1) / stat_period) * stat_period;
      int64_t last = std::min(next_stat, in.step_max - 1);
      if (ck_period > 0) last = std::min(last, (step / ck_period + 1) * ck_period - 1);
      const int32_t n = (int32_t)(last - step + 1);
      const bool stat = (last % stat_period) == 0;
      const bool ckpt = ck_period > 0 && (last + 1) % ck_period == 0;
      if (!check(mplbm_step(ctx, step, reverse, in.inject_type, n, stat ? MPLBM_STEP_WRITE_MACRO : 0))) {
        rc = -1;
        break;
      }
      step = last;
      if (stat) {
        const bool vtk = in.save_vtk_ascii && in.vtkPeriod > 0 && step % in.vtkPeriod == 0;
        const bool vtr = in.save_vtr_binary && in.vtrPeriod > 0 && step % in.vtrPeriod == 0;
        const bool files = vtk || vtr;
        mplbm_cluster_stats s{};
        if (in.host_statistics) {
          join_writer();
          if (!check(mplbm_read_fields(ctx, rho.data(), pres.data(), nullptr))) { rc = -1; break; }
          cluster_stats(d, is_fluid.data(), rho.data(), pres.data(), solid_number.data(), (float)in.rhol_spinodal,
                        (float)in.rhog_spinodal, solid_count, liq_label.data(), gas_label.data(), &s);
        } else {
          // statistics on the device; fields and labels cross to the host
          // only for a field file
          if (files) {
            liq32.resize(d.n());
            gas32.resize(d.n());
          }
          mplbm_clusters g{};
          if (!check(mplbm_calc_cluster(ctx, (float)in.rhol_spinodal, (float)in.rhog_spinodal, &g,
                                        files ? liq32.data() : nullptr, files ? gas32.data() : nullptr))) {
            rc = -1;
            break;
          }
          s.min_rho = g.min_rho; s.max_rho = g.max_rho; s.min_p = g.min_p; s.max_p = g.max_p;
          s.avg_liq_press = g.avg_liq_press; s.avg_vap_press = g.avg_vap_press; s.saturation = g.saturation;
          s.avg_liq_rho = g.avg_liq_rho; s.avg_vap_rho = g.avg_vap_rho;
          s.liq_cluster_count = g.liq_cluster_count; s.gas_cluster_count = g.gas_cluster_count;
          s.cluster_order = g.cluster_order; s.wet_coord_num = g.wet_coord_num;
          s.ratio_liquid_solid_boundary = g.ratio_liquid_solid_boundary;
          s.ratio_liquid_vapor_boundary = g.ratio_liquid_vapor_boundary;
          s.cluster_radius = g.cluster_radius;
          s.fluid_count = g.fluid_count; s.liquid_count = g.liquid_count; s.vapor_count = g.vapor_count;
          if (files) {
            join_writer();
            if (!check(mplbm_read_fields(ctx, rho.data(), pres.data(), nullptr))) { rc = -1; break; }
            std::copy(liq32.begin(), liq32.end(), liq_label.begin());
            std::copy(gas32.begin(), gas32.end(), gas_label.begin());
          }
        }
        std::printf("saturation=%f, aLp=%f,aVp=%f, step=%ld\n", s.saturation, s.avg_liq_press, s.avg_vap_press,
                    (long)step);
        write_stats_row(step, s);
        if (in.saturation_cycle) {
          if (reverse == 1 && s.saturation >= in.max_saturation) {
            reverse = -1;
            std::printf("saturation %f reached max_saturation at step %ld: draining\n", s.saturation, (long)step);
          } else if (reverse == -1 && s.saturation <= in.min_saturation) {
            reverse = 1;
            std::printf("saturation %f reached min_saturation at step %ld: injecting\n", s.saturation, (long)step);
          }
        }
        if (files) {
          std::string vtk_name, vtr_name;
          char name[512];
          if (vtk) {
            std::cout << "Saving vtk ascii file number " << vtk_num << std::endl;
            std::snprintf(name, sizeof name, "%s/result%06zu.vtk", in.result_folder, vtk_num++);
            vtk_name = name;
          }
          if (vtr) {
            std::cout << "Saving vtr binary file number " << vtr_num << std::endl;
            std::snprintf(name, sizeof name, "%s/result%06zu.vtr", in.result_folder, vtr_num++);
            vtr_name = name;
          }
          join_writer();
          writer = std::thread([this, vtk_name, vtr_name, &writer_rc]() {
            int r = 0;
            if (!vtk_name.empty())
              r |= write_vtk(vtk_name.c_str(), d, is_fluid.data(), rho.data(), pres.data(), liq_label.data(),
                             gas_label.data(), solid_number.data());
            if (!vtr_name.empty())
              r |= write_vtr(vtr_name.c_str(), d, is_fluid.data(), rho.data(), pres.data(), liq_label.data(),
                             gas_label.data(), solid_number.data());
            writer_rc = r;
          });
        }
      }
      if (ckpt) {
        const std::string name = std::string(in.result_folder) + "/checkpoint.mplbm";
        if (!check(mplbm_save_checkpoint(ctx, name.c_str(), last + 1))) { rc = -1; break; }
        if (write_driver_state(name, reverse) != 0) { rc = -1; break; }
        std::cout << "Saved checkpoint at step " << last + 1 << std::endl;
      }
      ++step;
    }
    join_writer();
    mplbm_sync(ctx);
    const double td = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("GPU time for %ld steps on resolution of %ld: %f sec\n", (long)in.step_max, (long)d.lx, td);
    const double upd = (double)fluid_nodes * (double)(in.step_max - step0);
    std::printf("MFLUPS (incl. statistics): %.1f\n", td > 0 ? upd / td / 1e6 : 0.0);
    std::cout << std::endl;
    mplbm_destroy(ctx);
    return rc;
  }

  // Driver state saved beside a checkpoint (<checkpoint>.driver, one line
  // "reverse R"): the injection direction is not part of the populations.
  // A checkpoint without the file resumes injecting.
  static int write_driver_state(const std::string& ckpt, int reverse) {
    const std::string name = ckpt + ".driver";
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) return set_err("cannot write %s", name.c_str());
    const bool ok = std::fprintf(f, "reverse %d\n", reverse) > 0;
    if (std::fclose(f) != 0 || !ok) return set_err("write error on %s", name.c_str());
    return 0;
  }
  static int read_driver_state(const std::string& ckpt, int* reverse) {
    const std::string name = ckpt + ".driver";
    std::ifstream f(name);
    if (!f) return 0;
    std::string key;
    int r = 0;
    if (!(f >> key >> r) || key != "reverse" || (r != 1 && r != -1)) return set_err("%s is malformed", name.c_str());
    *reverse = r;
    return 0;
  }

  void write_stats_row(int64_t step, const mplbm_cluster_stats& s) {  // :1057-1083
    const float alp = (float)s.avg_liq_press, avp = (float)s.avg_vap_press;
    stat_file << step << "\t" << (float)s.min_rho << "\t" << (float)s.max_rho << "\t" << (float)s.min_p << "\t"
              << (float)s.max_p << "\t" << alp << "\t" << avp << "\t" << avp - alp << "\t" << (float)s.saturation
              << "\t" << (float)s.avg_liq_rho << "\t" << (float)s.avg_vap_rho << "\t" << s.liq_cluster_count << "\t"
              << (float)s.cluster_order << "\t" << (float)s.wet_coord_num << "\t" << s.gas_cluster_count << "\t"
              << (float)s.ratio_liquid_solid_boundary << "\t" << (float)s.ratio_liquid_vapor_boundary << "\t"
              << (float)s.cluster_radius << std::endl
              << std::flush;
  }
};

template <class T>
bool read_into(std::istream& f, T& v) {
  f >> v;
  return true;
}

bool MultiphaseLBM::mplbm_parse_scalar_token(std::istream& file, const std::string& token, mplbm_input& in) {
  std::string s;
  if (token == "result_folder") {
    file >> s;
    std::snprintf(in.result_folder, sizeof in.result_folder, "%s", s.c_str());
  } else if (token == "num_threads") file >> in.num_threads;
  else if (token == "step") file >> in.step;
  else if (token == "step_max") file >> in.step_max;
  else if (token == "vtkPeriod") { file >> in.vtkPeriod; in.save_vtk_ascii = 1; }
  else if (token == "vtrPeriod") { file >> in.vtrPeriod; in.save_vtr_binary = 1; }
  else if (token == "tecplotPeriod") { file >> in.tecplotPeriod; in.save_tecplot_ascii = 1; }
  else if (token == "statPeriod") file >> in.statPeriod;
  else if (token == "rho0") file >> in.rho0;
  else if (token == "rhos") file >> in.rhos;
  else if (token == "G") file >> in.G;
  else if (token == "IniPerturbRate") file >> in.IniPerturbRate;
  else if (token == "tau") file >> in.tau;
  else if (token == "resolution") file >> in.res;
  else if (token == "rho_inject_period") file >> in.rho_inject_period;
  else if (token == "rho_increment") { in.inject_density = 1; file >> in.rho_increment; }
  else if (token == "inject_type") file >> in.inject_type;
  else if (token == "inject_position") file >> in.inject_position;
  else if (token == "carn_star") in.carn_star = 1;
  else if (token == "T_Tc") file >> in.T_Tc;
  else if (token == "rhol_spinodal") file >> in.rhol_spinodal;
  else if (token == "rhog_spinodal") file >> in.rhog_spinodal;
  else if (token == "read_densities") {
    file >> s;
    std::snprintf(in.densities_filename, sizeof in.densities_filename, "%s", s.c_str());
    in.read_densities = 1;
  } else if (token == "max_saturation") { file >> in.max_saturation; in.hysteresis = 1; }
  else if (token == "min_saturation") file >> in.min_saturation;
  else if (token == "bound_difference") { double v; file >> v; }
  else if (token == "target_density") { file >> in.target_density; in.erosion = 1; }
  // runtime domain size (compile-time lx/ly/lz in include/lbm_model.cuh:4-6;
  // the example inputs name them lxSample/lySample/lzSample)
  else if (token == "lxSample" || token == "lx") file >> in.lx;
  else if (token == "lySample" || token == "ly") file >> in.ly;
  else if (token == "lzSample" || token == "lz") file >> in.lz;
  else if (token == "storage") {
    file >> s;
    in.storage = s == "fp64" ? MPLBM_FP64 : s == "fp16s" ? MPLBM_FP16S : s == "bf16s" ? MPLBM_BF16S : MPLBM_FP32;
  } else if (token == "math") {
    file >> s;
    in.math = s == "fast" ? MPLBM_MATH_FAST : MPLBM_MATH_REFERENCE;
  } else if (token == "kernel") {
    file >> s;
    in.kernel = s == "twopass"   ? MPLBM_KERNEL_TWOPASS
                : s == "fused"   ? MPLBM_KERNEL_FUSED
                : s == "compact" ? MPLBM_KERNEL_COMPACT
                                 : MPLBM_KERNEL_AUTO;
  } else if (token == "pressure_x" || token == "pressure_y" || token == "pressure_z") {
    const int a = token[9] - 'x';
    double lo, hi;
    file >> lo >> hi;
    in.face_type[2 * a] = in.face_type[2 * a + 1] = MPLBM_FACE_PRESSURE;
    in.face_rho[2 * a] = lo;
    in.face_rho[2 * a + 1] = hi;
  } else if (token == "body_force") file >> in.body_force[0] >> in.body_force[1] >> in.body_force[2];
  else if (token == "host_statistics") file >> in.host_statistics;
  else if (token == "device_voxeliser") file >> in.device_voxeliser;
  else if (token == "shift") {
    file >> s;
    in.shift = s == "node" ? MPLBM_SHIFT_NODE : MPLBM_SHIFT_GLOBAL;
  }
  else if (token == "checkpointPeriod") file >> in.checkpointPeriod;
  else if (token == "rounding") {
    file >> s;
    in.rounding = s == "stochastic" ? MPLBM_ROUND_STOCHASTIC : MPLBM_ROUND_NEAREST;
  }
  else if (token == "saturation_cycle") file >> in.saturation_cycle;
  else if (token == "read_checkpoint") {
    file >> s;
    std::snprintf(in.checkpoint_filename, sizeof in.checkpoint_filename, "%s", s.c_str());
    in.read_checkpoint = 1;
  }
  else return false;
  return true;
}

}  // namespace

extern "C" {

const char* mplbm_host_last_error(void) {
  // a copy per calling thread, stable until that thread's next call
  thread_local std::string copy;
  std::lock_guard<std::mutex> lk(g_err_mu);
  copy = g_err;
  return copy.c_str();
}

void mplbm_host_input_default(mplbm_input* in) {
  std::memset(in, 0, sizeof *in);
  in->lx = in->ly = in->lz = 400;
  std::snprintf(in->result_folder, sizeof in->result_folder, "%s", "./RESULT");
  in->num_threads = 4;
  in->step = 0;
  in->step_max = 1000;
  in->vtkPeriod = 10;
  in->statPeriod = 10;
  in->rho0 = 0.10;
  in->IniPerturbRate = 1e-3;
  in->G = -1.0;
  in->tau = 1.0;
  in->res = 1.0;
  in->T_Tc = 0.7;
  in->rhol_spinodal = 0.2726;
  in->rhog_spinodal = 0.0484;
  in->inject_type = 0;
  in->inject_position = 1.0;
  in->rho_inject_period = 100;
  in->rho_increment = 0.0;
  in->max_saturation = 0.98;
  in->min_saturation = 0.04;
  in->storage = MPLBM_FP32;
  in->kernel = MPLBM_KERNEL_AUTO;
}

int mplbm_host_parse(const char* path, mplbm_input* in) {
  mplbm_host_input_default(in);
  std::ifstream file(path);
  if (file.fail()) return set_err("Cannot open command file %s", path);
  std::string token;
  file >> token;
  while (file.good()) {
    if (token[0] == '!' || token[0] == '#') {
      std::getline(file, token);
    } else if (token == "resize") {
    } else if (token == "read_sample" || token == "read_positions") {
      std::string s;
      file >> s;
    } else if (token == "place_sphere") {
      float v[4];
      file >> v[0] >> v[1] >> v[2] >> v[3];
    } else if (token == "generate_solids") {
      float v[9];
      for (float& x : v) file >> x;
    } else if (!MultiphaseLBM::mplbm_parse_scalar_token(file, token, *in)) {
      in->unknown_tokens++;
    }
    file >> token;
  }
  return 0;
}

void mplbm_host_place_sphere(uint8_t* is_fluid, int64_t* solid_number, int64_t lx, int64_t ly, int64_t lz, float x,
                             float y, float z, float R, int64_t solid_id) {
  place_sphere(Dims{lx, ly, lz}, is_fluid, solid_number, x, y, z, R, solid_id);
}

int64_t mplbm_host_read_positions(const char* path, float res, uint8_t* is_fluid, int64_t* solid_number, int64_t lx,
                                  int64_t ly, int64_t lz, int64_t first_id) {
  std::vector<float> v;
  if (read_positions_values(path, res, v) < 0) return -1;
  const Dims d{lx, ly, lz};
  const int64_t count = (int64_t)(v.size() / 4);
  for (int64_t k = 0; k < count; ++k)
    place_sphere(d, is_fluid, solid_number, v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3], first_id + k);
  return count;
}

int64_t mplbm_host_read_positions_device(const char* path, float res, uint8_t* is_fluid, int64_t* solid_number,
                                         int64_t lx, int64_t ly, int64_t lz, int64_t first_id) {
  std::vector<float> v;
  if (read_positions_values(path, res, v) < 0) return -1;
  const int64_t count = (int64_t)(v.size() / 4);
  if (mplbm_place_spheres(lx, ly, lz, v.data(), count, first_id, is_fluid, solid_number) != MPLBM_OK)
    return set_err("%s", mplbm_last_error(nullptr));
  return count;
}

void mplbm_host_rho_init(float* rho, int64_t lx, int64_t ly, int64_t lz, float rho0, float ipr) {
  const Dims d{lx, ly, lz};
  srand(1);
  for (int64_t z = 0; z < lz; ++z)
    for (int64_t y = 0; y < ly; ++y)
      for (int64_t x = 0; x < lx; ++x)
        rho[d.at(x, y, z)] = (float)((double)rho0 * (1.0 + (double)ipr * ((double)((float)rand() / (float)RAND_MAX) - 0.5)));
}

int64_t mplbm_host_read_densities(const char* path, float* rho, int64_t lx, int64_t ly, int64_t lz) {
  std::ifstream f(path);
  if (!f) return set_err("cannot open density file %s", path);
  const Dims d{lx, ly, lz};
  std::string line;
  bool reading = false;
  int64_t k = 0;
  const int64_t n = (int64_t)d.n();
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line == "SCALARS Pressure float 1") break;
    if (line == "SCALARS Density float 1") { reading = true; continue; }
    if (!reading || line == "LOOKUP_TABLE default") continue;
    if (k >= n) return set_err("density file %s has more than %ld values", path, (long)n);
    const int64_t x = k % lx, y = (k / lx) % ly, z = k / lx / ly;
    rho[d.at(x, y, z)] = (float)std::stod(line);
    ++k;
  }
  return k;
}

int mplbm_host_read_sample(const char* path, uint8_t* is_fluid, int64_t lx, int64_t ly, int64_t lz) {
  std::ifstream f(path);
  if (!f) return set_err("Cannot open sample file %s", path);
  int64_t fx, fy, fz;
  f >> fx >> fy >> fz;
  if (!f || fx <= 0 || fy <= 0 || fz <= 0) return set_err("bad sample header in %s", path);
  const int64_t ds = fx / lx;
  if (ds < 1) return set_err("sample %ld smaller than the domain %ld", (long)fx, (long)lx);
  const Dims d{lx, ly, lz};
  std::vector<uint8_t> ids;
  std::string line;
  std::getline(f, line);  // rest of the header line
  int64_t k = 0;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    const int64_t x = k / fz / fy, y = (k / fz) % fy, z = k % fz;
    if (x % ds == 0 && y % ds == 0 && z % ds == 0) ids.push_back((uint8_t)(std::stoi(line) != 0));
    ++k;
  }
  if ((int64_t)ids.size() < (int64_t)d.n()) return set_err("sample %s has %zu kept voxels, need %zu", path, ids.size(), d.n());
  size_t q = 0;
  for (int64_t x = 0; x < lx; ++x)
    for (int64_t y = 0; y < ly; ++y)
      for (int64_t z = 0; z < lz; ++z) is_fluid[d.at(x, y, z)] = ids[q++] ? 0 : 1;  // 1 = solid in the file
  return 0;
}

int64_t mplbm_host_generate_packing(uint8_t* is_fluid, int64_t lx, int64_t ly, int64_t lz, uint64_t seed,
                                    int64_t n_max, double rmin, double rmax, double porosity_target,
                                    int32_t nonoverlap, double* spheres_out, int64_t capacity) {
  const Dims d{lx, ly, lz};
  const size_t n = d.n();
  std::fill(is_fluid, is_fluid + n, 1);
  SplitMix rng{seed};
  std::vector<double> sph;
  size_t solid = 0;
  int64_t attempts = 0;
  const int64_t max_attempts = 200 * std::max<int64_t>(n_max, 1000);
  std::vector<uint8_t> before;
  while ((int64_t)(sph.size() / 4) < n_max && attempts < max_attempts) {
    if ((double)(n - solid) / (double)n <= porosity_target) break;
    ++attempts;
    const double r = rmin + (rmax - rmin) * rng.uniform();
    const double x = lx * rng.uniform(), y = ly * rng.uniform(), z = lz * rng.uniform();
    if (nonoverlap) {
      bool ok = true;
      for (size_t k = 0; k < sph.size() && ok; k += 4) {
        double dx = std::fabs(x - sph[k]), dy = std::fabs(y - sph[k + 1]), dz = std::fabs(z - sph[k + 2]);
        dx = std::min(dx, lx - dx); dy = std::min(dy, ly - dy); dz = std::min(dz, lz - dz);
        const double rr = r + sph[k + 3];
        if (dx * dx + dy * dy + dz * dz < rr * rr) ok = false;
      }
      if (!ok) continue;
    }
    // count newly solid voxels inside the sphere's bounding box
    const int64_t id = (int64_t)(sph.size() / 4) + 1;
    size_t newly = 0;
    {
      const float xf = (float)x, yf = (float)y, zf = (float)z, rf = (float)r;
      for (float px = xf - rf; px <= xf + rf; px += 1.f)
        for (float py = yf - rf; py <= yf + rf; py += 1.f)
          for (float pz = zf - rf; pz <= zf + rf; pz += 1.f) {
            const float dx = px - xf, dy = py - yf, dz = pz - zf;
            if (dx * dx + dy * dy + dz * dz < rf * rf) {
              int ix = nearest(px), iy = nearest(py), iz = nearest(pz);
              if (ix < 0) ix += (int)lx;
              if (iy < 0) iy += (int)ly;
              if (iz < 0) iz += (int)lz;
              if (ix >= lx) ix -= (int)lx;
              if (iy >= ly) iy -= (int)ly;
              if (iz >= lz) iz -= (int)lz;
              if (ix < 0 || iy < 0 || iz < 0 || ix >= lx || iy >= ly || iz >= lz) continue;
              uint8_t& v = is_fluid[d.at(ix, iy, iz)];
              if (v) { v = 0; ++newly; }
            }
          }
    }
    (void)id;
    solid += newly;
    sph.insert(sph.end(), {x, y, z, r});
  }
  const int64_t count = (int64_t)(sph.size() / 4);
  if (spheres_out) {
    const int64_t m = std::min(count, capacity);
    std::copy(sph.begin(), sph.begin() + 4 * m, spheres_out);
  }
  return count;
}

int mplbm_host_write_positions(const char* path, const double* s, int64_t n) {
  FILE* f = std::fopen(path, "w");
  if (!f) return set_err("cannot open %s", path);
  for (int64_t k = 0; k < 4 * n; ++k) std::fprintf(f, "%.9g\n", s[k]);
  std::fclose(f);
  return 0;
}

int mplbm_host_cluster_stats(const uint8_t* is_fluid, const float* rho, const float* pres,
                             const int64_t* solid_number, int64_t lx, int64_t ly, int64_t lz, double rhol,
                             double rhog, int64_t solid_count, int64_t* liq_label, int64_t* gas_label,
                             mplbm_cluster_stats* out) {
  return cluster_stats(Dims{lx, ly, lz}, is_fluid, rho, pres, solid_number, rhol, rhog, solid_count, liq_label,
                       gas_label, out);
}

int mplbm_host_write_vtr(const char* path, const uint8_t* is_fluid, const float* rho, const float* pres,
                         const int64_t* liq, const int64_t* gas, const int64_t* solid, int64_t lx, int64_t ly,
                         int64_t lz) {
  return write_vtr(path, Dims{lx, ly, lz}, is_fluid, rho, pres, liq, gas, solid);
}

int mplbm_host_write_vtk(const char* path, const uint8_t* is_fluid, const float* rho, const float* pres,
                         const int64_t* liq, const int64_t* gas, const int64_t* solid, int64_t lx, int64_t ly,
                         int64_t lz) {
  return write_vtk(path, Dims{lx, ly, lz}, is_fluid, rho, pres, liq, gas, solid);
}

int mplbm_host_read_data(const char* input_path, mplbm_input* in, uint8_t* is_fluid, float* rho, int64_t capacity,
                         int64_t* eroded) {
  if (!input_path || !in) return set_err("null argument");
  MultiphaseLBM box;
  if (box.read_data(input_path) != 0) return -1;
  *in = box.in;
  in->lx = box.d.lx;
  in->ly = box.d.ly;
  in->lz = box.d.lz;
  const int64_t n = (int64_t)box.d.n();
  if ((is_fluid || rho) && n > capacity) return set_err("domain of %lld cells exceeds capacity %lld", (long long)n,
                                                         (long long)capacity);
  if (is_fluid) std::copy(box.is_fluid.begin(), box.is_fluid.end(), is_fluid);
  if (rho) std::copy(box.rho.begin(), box.rho.end(), rho);
  if (eroded) *eroded = box.eroded_nodes;
  return 0;
}

int mplbm_host_run(const char* input_path) {
  MultiphaseLBM box;
  std::cout << "Opening file : " << input_path << std::endl;
  if (box.read_data(input_path) != 0) {
    std::cerr << g_err << std::endl;
    return -1;
  }
  if (box.init() != 0 || box.run() != 0) {
    std::cerr << g_err << std::endl;
    return -1;
  }
  return 0;
}

}  // extern "C"
