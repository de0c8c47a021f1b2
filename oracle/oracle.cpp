// CellFlow step oracle — TEST INFRASTRUCTURE ONLY.
//
// A scalar CPU restatement of the reference particle-life step and proximity-graph
// builder.  It exists to check the HIP engine; nothing in the product links it.
// Only the test suite, the smoke check and the benchmark's CPU-baseline leg load it.
//
// Arithmetic contract (shared with the HIP kernels, fixed by the engine's spec):
//   * discrete decisions (cutoff predicate, neighbour count, edge predicate, cell
//     binning) are evaluated in FP32 with the op order written below, compiled with
//     -ffp-contract=off and explicit fmaf, so they are bit-identical to the GPU's;
//   * everything continuous (forces, integration, wrap) is evaluated in FP64.
//
// Functions and the reference lines they restate (paths relative to the reference root):
//   orc_pair_in_range   cuda-native/src/ParticleSimulation.cu:92-112  (min-image, softened
//                       distance, per-type effective radius, strict cutoff)
//   orc_step            ParticleSimulation.cu:69-166 with Jacobi semantics (every read from
//                       the pre-step state; the reference updates in place and races)
//   orc_edges           ParticleSimulation.cu:187-246 (same type, j>i, non-periodic, first
//                       2*maxC by index, stable sort by distance, keep maxC); output is the
//                       (owner, rank) order instead of the reference's atomicAdd order
//   orc_force_table     ParticleSimulation.cu:521-531 (updateForceTable transform)
//   orc_move            ParticleSimulation.cu:169-185 (moveUniverse's wrap)
//   orc_bin / orc_order the engine's own uniform-grid rule (no reference counterpart: the
//                       reference enumerates all pairs, :86)
//   orc_rollout         repeated orc_step with the neighbour-count ping-pong of :541-556
//   orc_hist            soft pair histogram of the differentiable-rollout loss (engine
//                       extension, no reference code)
//   orc_pairs           one step's neighbour lists by the FP32 predicate
//   orc_frozen_rollout  FP64 rollout with given (frozen) neighbour lists, for central finite
//                       differences of the rollout loss
//
// Pinning: every function above is checked in tests/test_oracle.py against closed forms
// (two-particle force, isolated particle, balance=1, zero table), invariants (pair
// symmetry, permutation and periodic-shift equivariance), brute force on tiny inputs
// (all-pairs vs grid enumeration, edge truncation cases built by hand) and the four
// reference preset files.  The rollout helpers are pinned the same way: orc_hist by the
// cubic B-spline's exact values at integer and half-integer offsets, its partition of
// unity and the shell sums of a simple cubic lattice; orc_pairs by brute force and by
// orc_step's counts; orc_frozen_rollout by equality with orc_rollout when given that
// rollout's own decisions, and its finite differences by step-size convergence and exact
// zeros for unused table entries.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

struct orc_params {
  float radius, dt, friction, repulsion, attraction, k, balance;
  float L[3];
  int T;
  float ratio;
  float force_multiplier;
  int max_expected_neighbors;
};

}  // extern "C"

namespace {

inline float min_image_f(float d, float L) {
  // ParticleSimulation.cu:97-102: one shift each way, in this order.
  if (d > L * 0.5f) d -= L;
  if (d < -L * 0.5f) d += L;
  return d;
}

inline double min_image_d(double d, double L) {
  if (d > L * 0.5) d -= L;
  if (d < -L * 0.5) d += L;
  return d;
}

inline float eff_radius_f(const orc_params& p, const float* rbt, unsigned ti, unsigned tj) {
  // ParticleSimulation.cu:108-110, FP32, no contraction.
  float ri = p.radius * (1.0f + rbt[ti] * p.ratio);
  float rj = p.radius * (1.0f + rbt[tj] * p.ratio);
  return (ri + rj) * 0.5f;
}

inline double eff_radius_d(const orc_params& p, const float* rbt, unsigned ti, unsigned tj) {
  double r = p.radius, rho = p.ratio;
  double ri = r * (1.0 + (double)rbt[ti] * rho);
  double rj = r * (1.0 + (double)rbt[tj] * rho);
  return (ri + rj) * 0.5;
}

// FP32 decision for one ordered pair (i sees j); returns the squared FP32 distance too.
inline bool pair_in_range_f(const orc_params& p, const float* rbt, const float* xi, const float* xj,
                            unsigned ti, unsigned tj) {
  float dx = min_image_f(xj[0] - xi[0], p.L[0]);
  float dy = min_image_f(xj[1] - xi[1], p.L[1]);
  float dz = min_image_f(xj[2] - xi[2], p.L[2]);
  float s = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  float dist = sqrtf(s + 0.0001f);  // ParticleSimulation.cu:105
  return dist < eff_radius_f(p, rbt, ti, tj);  // :112, strict
}

struct Grid {
  int n[3];
  double h[3];
  std::vector<int> start;   // cell -> first index into order
  std::vector<int> order;   // particle ids sorted by cell
};

// A coarse grid for the oracle's own accelerated enumeration.  Independent of the
// engine's grid: it only has to cover every pair within `reach_dist`.
Grid build_grid(int n, const float* pos, const float* L, double reach_dist) {
  Grid g;
  for (int a = 0; a < 3; ++a) {
    int na = (int)std::floor((double)L[a] / (reach_dist * 1.01));
    na = std::max(1, std::min(na, 512));
    g.n[a] = na;
    g.h[a] = (double)L[a] / na;
  }
  long long nc = (long long)g.n[0] * g.n[1] * g.n[2];
  while (nc > std::max<long long>(8LL * n, 64)) {
    int a = 0;
    if (g.n[1] > g.n[a]) a = 1;
    if (g.n[2] > g.n[a]) a = 2;
    g.n[a] = std::max(1, g.n[a] - 1);
    g.h[a] = (double)L[a] / g.n[a];
    nc = (long long)g.n[0] * g.n[1] * g.n[2];
  }
  std::vector<int> cell(n);
  std::vector<int> cnt(nc + 1, 0);
  for (int i = 0; i < n; ++i) {
    int c[3];
    for (int a = 0; a < 3; ++a) {
      double u = (double)pos[3 * i + a] / g.h[a];
      int ca = (u >= 0) ? (int)std::floor(u) : 0;
      c[a] = std::min(std::max(ca, 0), g.n[a] - 1);
    }
    cell[i] = c[0] + g.n[0] * (c[1] + g.n[1] * c[2]);
    cnt[cell[i] + 1]++;
  }
  for (long long c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
  g.start = cnt;
  g.order.assign(n, 0);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; ++i) g.order[fill[cell[i]]++] = i;
  return g;
}

// Visit every particle in the cells within `r` cells of particle i's cell (periodic or
// not); each cell is visited at most once.
template <class F>
void for_each_near(const Grid& g, const float* pos, int i, int r, bool periodic, F&& f) {
  int c[3];
  for (int a = 0; a < 3; ++a) {
    double u = (double)pos[3 * i + a] / g.h[a];
    int ca = (u >= 0) ? (int)std::floor(u) : 0;
    c[a] = std::min(std::max(ca, 0), g.n[a] - 1);
  }
  std::vector<int> offs[3];
  for (int a = 0; a < 3; ++a) {
    if (periodic && 2 * r + 1 >= g.n[a]) {
      for (int v = 0; v < g.n[a]; ++v) offs[a].push_back(v);
    } else {
      for (int o = -r; o <= r; ++o) {
        int v = c[a] + o;
        if (periodic) v = ((v % g.n[a]) + g.n[a]) % g.n[a];
        else if (v < 0 || v >= g.n[a]) continue;
        offs[a].push_back(v);
      }
    }
  }
  for (int z : offs[2])
    for (int y : offs[1])
      for (int x : offs[0]) {
        int cell = x + g.n[0] * (y + g.n[1] * z);
        for (int k = g.start[cell]; k < g.start[cell + 1]; ++k) f(g.order[k]);
      }
}

double max_eff_radius(const orc_params& p, const float* rbt) {
  double m = 0;
  for (int a = 0; a < p.T; ++a)
    for (int b = 0; b < p.T; ++b) m = std::max(m, (double)eff_radius_f(p, rbt, a, b));
  return m;
}

struct PairTerm {
  double f[3];
  double absnet;
};

// FP64 force of j on i for a pair already accepted by the FP32 predicate.
inline PairTerm pair_force_d(const orc_params& p, const float* table, const float* rbt,
                             const double* xi, const double* xj, unsigned ti, unsigned tj) {
  double d[3];
  for (int a = 0; a < 3; ++a) d[a] = min_image_d(xj[a] - xi[a], (double)p.L[a]);
  double s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double dist = std::sqrt(s + (double)0.0001f);
  double eff = eff_radius_d(p, rbt, ti, tj);
  double fv = table[ti * p.T + tj];                 // :116
  double r = dist / eff;                            // :119
  double net = ((double)p.repulsion * std::exp(-(double)p.k * r * r) - (double)p.attraction * r) * fv;  // :120-125
  PairTerm t;
  for (int a = 0; a < 3; ++a) t.f[a] = d[a] / dist * net;  // :128-131
  t.absnet = std::fabs(net);
  return t;
}

void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

}  // namespace

extern "C" {

// FP32 cutoff decision for the ordered pair (i, j), exactly as the engine evaluates it.
int orc_pair_in_range(const orc_params* p, const float* rbt, const float* xi, const float* xj,
                      unsigned ti, unsigned tj) {
  return pair_in_range_f(*p, rbt, xi, xj, ti, tj) ? 1 : 0;
}

// FP32 effective radius of a type pair (ParticleSimulation.cu:108-110).
float orc_eff_radius(const orc_params* p, const float* rbt, unsigned ti, unsigned tj) {
  return eff_radius_f(*p, rbt, ti, tj);
}

// One Jacobi step.  pos/vel are FP64 (an FP32 state promoted exactly, or an FP64 rollout
// state); the FP32 decisions use (float)pos.  mode 0 enumerates all j != i in ascending
// order (ParticleSimulation.cu:86); mode 1 enumerates a covering grid (same neighbour set).
// fabs_out[i] (optional) receives sum_j |net_ij|, the magnitude scale of the force sum.
int orc_step(int n, const double* pos, const double* vel, const uint32_t* type, const int32_t* prev,
             const orc_params* pp, const float* table, const float* rbt, double* pos_out,
             double* vel_out, double* acc_out, int32_t* count_out, double* fabs_out, int mode,
             int nthreads) {
  const orc_params& p = *pp;
  if (n < 0 || p.T <= 0 || p.T > 10) return 1;
  for (int i = 0; i < n; ++i)
    if (type[i] >= (uint32_t)p.T) return 1;
  set_threads(nthreads);
  std::vector<float> posf(3 * (size_t)n);
  for (size_t k = 0; k < 3 * (size_t)n; ++k) posf[k] = (float)pos[k];
  Grid g;
  double reach = max_eff_radius(p, rbt);
  const bool use_grid = (mode == 1) && reach > 0;
  if (use_grid) g = build_grid(n, posf.data(), p.L, reach);

#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const unsigned ti = type[i];
    double F[3] = {0, 0, 0};
    double fabs_sum = 0;
    int count = 0;
    auto visit = [&](int j) {
      if (j == i) return;  // :87
      const unsigned tj = type[j];
      if (!pair_in_range_f(p, rbt, &posf[3 * i], &posf[3 * j], ti, tj)) return;
      count++;  // :113
      PairTerm t = pair_force_d(p, table, rbt, &pos[3 * i], &pos[3 * j], ti, tj);
      for (int a = 0; a < 3; ++a) F[a] += t.f[a];
      fabs_sum += t.absnet;
    };
    if (use_grid) {
      for_each_near(g, posf.data(), i, 1, true, visit);
    } else if (reach > 0) {
      for (int j = 0; j < n; ++j) visit(j);
    }
    // :136-143 adaptive factor from the current and previous neighbour counts.
    double avg = (double)(count + prev[i]) * 0.5;
    double dens = std::fmin(avg / (double)p.max_expected_neighbors, 1.0);  // fminf semantics (NaN -> 1)
    double fac = 1.0 - (1.0 - (double)p.balance) * dens;
    for (int a = 0; a < 3; ++a) F[a] *= (double)p.force_multiplier * fac;
    for (int a = 0; a < 3; ++a) {
      double v = vel[3 * i + a] * (double)p.friction + F[a] * (double)p.dt;  // :149-151
      double x = pos[3 * i + a] + v * (double)p.dt;                            // :154-156
      double L = (double)p.L[a];
      x = std::fmod(x + L, L);                                                 // :159-161
      vel_out[3 * i + a] = v;
      pos_out[3 * i + a] = x;
      if (acc_out) acc_out[3 * i + a] = F[a];                                  // :146
    }
    count_out[i] = count;                                                      // :165
    if (fabs_out) fabs_out[i] = fabs_sum;
  }
  return 0;
}

// `steps` Jacobi steps in FP64 with FP32 decisions, carrying the count ping-pong
// (ParticleSimulation.cu:541-556).  pos/vel/prev are updated in place.
int orc_rollout(int n, double* pos, double* vel, const uint32_t* type, int32_t* prev,
                const orc_params* p, const float* table, const float* rbt, int steps, int mode,
                int nthreads) {
  std::vector<double> pos2(3 * (size_t)n), vel2(3 * (size_t)n);
  std::vector<int32_t> cnt(n);
  for (int s = 0; s < steps; ++s) {
    int rc = orc_step(n, pos, vel, type, prev, p, table, rbt, pos2.data(), vel2.data(), nullptr,
                      cnt.data(), nullptr, mode, nthreads);
    if (rc) return rc;
    std::memcpy(pos, pos2.data(), sizeof(double) * 3 * n);
    std::memcpy(vel, vel2.data(), sizeof(double) * 3 * n);
    std::memcpy(prev, cnt.data(), sizeof(int32_t) * n);
  }
  return 0;
}

// Engine grid rule restated: cell = clamp(floor(x * inv_h), 0, n-1) per axis in FP32 with
// inv_h = (float)n / L, linear id x-fastest.
int orc_bin(int n, const float* pos, const int* dims, const float* L, uint32_t* cell_out) {
  float inv_h[3];
  for (int a = 0; a < 3; ++a) {
    if (dims[a] < 1) return 1;
    inv_h[a] = (float)dims[a] / L[a];
  }
  for (int i = 0; i < n; ++i) {
    int c[3];
    for (int a = 0; a < 3; ++a) {
      float u = pos[3 * i + a] * inv_h[a];
      int ca = (u >= 0.0f) ? (int)floorf(u) : 0;
      if (!(u < (float)dims[a])) ca = (u >= 0.0f) ? dims[a] - 1 : 0;
      c[a] = std::min(std::max(ca, 0), dims[a] - 1);
    }
    cell_out[i] = (uint32_t)(c[0] + dims[0] * (c[1] + dims[1] * c[2]));
  }
  return 0;
}

// Canonical (cell, id) order: order_out[k] = id of the k-th particle.
int orc_order(int n, const uint32_t* cell, uint32_t* order_out) {
  std::vector<uint32_t> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return cell[a] < cell[b]; });
  std::memcpy(order_out, idx.data(), sizeof(uint32_t) * n);
  return 0;
}

// Proximity graph (ParticleSimulation.cu:187-246).  Output: pairs (i, j) in canonical
// order (owner i ascending, then distance rank).  mode 0 is the reference's own scan
// (j = i+1.. while found < 2*maxC); mode 1 uses a non-periodic grid (same result).
// Returns 0, or 3 when `capacity` pairs are too few (count_out then holds the need).
int orc_edges(int n, const float* pos, const uint32_t* type, float dist, int max_conn, int mode,
              int32_t* pairs_out, long long capacity, long long* count_out, int nthreads) {
  set_threads(nthreads);
  const float d2 = dist * dist;                 // ParticleSimulation.cu:660
  const int maxc = std::min(max_conn, 32);      // :212
  const int cap = 2 * maxc;                     // :215 (capacity of the candidate list)
  std::vector<std::vector<std::pair<float, int>>> kept(n);
  if (maxc <= 0 || n <= 1) {
    *count_out = 0;
    return 0;
  }
  auto d2of = [&](int i, int j) {
    float dx = pos[3 * j] - pos[3 * i];
    float dy = pos[3 * j + 1] - pos[3 * i + 1];
    float dz = pos[3 * j + 2] - pos[3 * i + 2];
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  };
  Grid g;
  float Lb[3] = {0, 0, 0};
  if (mode == 1) {
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) Lb[a] = std::max(Lb[a], pos[3 * i + a]);
    for (int a = 0; a < 3; ++a) Lb[a] = std::nextafter(Lb[a], 1e30f) * 1.001f + 1e-30f;
    g = build_grid(n, pos, Lb, std::max((double)dist, 1e-6));
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    std::vector<std::pair<float, int>> cand;  // (d2, j) in ascending j
    if (mode == 0) {
      for (int j = i + 1; j < n && (int)cand.size() < cap; ++j) {  // :215
        if (type[j] != type[i]) continue;                          // :219
        float s = d2of(i, j);
        if (s < d2) cand.push_back({s, j});                         // :227-231
      }
    } else {
      for_each_near(g, pos, i, 1, false, [&](int j) {
        if (j <= i || type[j] != type[i]) return;
        float s = d2of(i, j);
        if (s < d2) cand.push_back({s, j});
      });
      std::sort(cand.begin(), cand.end(), [](auto& a, auto& b) { return a.second < b.second; });
      if ((int)cand.size() > cap) cand.resize(cap);
    }
    // :235-243 stable insertion sort by distance (ties keep index order).
    std::stable_sort(cand.begin(), cand.end(), [](auto& a, auto& b) { return a.first < b.first; });
    if ((int)cand.size() > maxc) cand.resize(maxc);  // :246
    kept[i] = std::move(cand);
  }
  long long total = 0;
  for (int i = 0; i < n; ++i) total += (long long)kept[i].size();
  *count_out = total;
  if (total > capacity) return 3;
  long long w = 0;
  for (int i = 0; i < n; ++i)
    for (auto& e : kept[i]) {
      pairs_out[2 * w] = i;
      pairs_out[2 * w + 1] = e.second;
      ++w;
    }
  return 0;
}

// Triangle mesh (generateTriangleMeshKernel, ParticleSimulation.cu:279-348, with the
// driver's threshold :722-724).  Output: triangles (i, a, b) in canonical order (owner i
// ascending, then the owner's fan order).  mode 0 is the reference's own scan (j = 0.. while
// found < maxC, :306); mode 1 uses a non-periodic grid (same result).  Returns 0, or 3 when
// `capacity` triangles are too few (count_out then holds the need).
int orc_triangles(int n, const float* pos, const uint32_t* type, float dist, int max_conn,
                  int mode, int32_t* tris_out, long long capacity, long long* count_out,
                  int nthreads) {
  set_threads(nthreads);
  const float hm = 1.2f;
  const float d2 = dist * dist * (hm * hm);     // :723
  const int maxc = std::min(max_conn, 32);      // :303
  std::vector<std::vector<int>> tri(n);         // per owner: a, b of each triangle
  if (maxc <= 0 || n <= 1) {
    *count_out = 0;
    return 0;
  }
  auto d2of = [&](int i, int j) {               // :314-317
    float dx = pos[3 * j] - pos[3 * i];
    float dy = pos[3 * j + 1] - pos[3 * i + 1];
    float dz = pos[3 * j + 2] - pos[3 * i + 2];
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  };
  Grid g;
  if (mode == 1) {
    float Lb[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) Lb[a] = std::max(Lb[a], pos[3 * i + a]);
    for (int a = 0; a < 3; ++a) Lb[a] = std::nextafter(Lb[a], 1e30f) * 1.001f + 1e-30f;
    g = build_grid(n, pos, Lb, std::max(std::sqrt((double)d2), 1e-6));
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    std::vector<int> near;  // the neighbour list of :301-324, ascending index
    if (mode == 0) {
      for (int j = 0; j < n && (int)near.size() < maxc; ++j) {
        if (j == i || type[j] != type[i]) continue;
        if (d2of(i, j) < d2) near.push_back(j);
      }
    } else {
      for_each_near(g, pos, i, 1, false, [&](int j) {
        if (j == i || type[j] != type[i]) return;
        if (d2of(i, j) < d2) near.push_back(j);
      });
      std::sort(near.begin(), near.end());
      if ((int)near.size() > maxc) near.resize(maxc);
    }
    // :334-348 fan triangles, at most 64
    std::vector<int>& t = tri[i];
    for (size_t a = 0; a < near.size() && t.size() < 128; ++a)
      for (size_t b = a + 1; b < near.size() && t.size() < 128; ++b)
        if (i < near[a] && i < near[b]) {
          t.push_back(near[a]);
          t.push_back(near[b]);
        }
  }
  long long total = 0;
  for (int i = 0; i < n; ++i) total += (long long)tri[i].size() / 2;
  *count_out = total;
  if (total > capacity) return 3;
  long long w = 0;
  for (int i = 0; i < n; ++i)
    for (size_t k = 0; k < tri[i].size(); k += 2, ++w) {
      tris_out[3 * w] = i;
      tris_out[3 * w + 1] = tri[i][k];
      tris_out[3 * w + 2] = tri[i][k + 1];
    }
  return 0;
}

// Vertices of triangles (:362-418): per vertex x, y, z, nx, ny, nz, r, g, b with the flat
// normal of the triangle (cross product of the edges from its first vertex, normalised when
// longer than 1e-4) and the colour of the first vertex's type, colors[ptype % T].  FP32.
void orc_triangle_vertices(long long ntri, const int32_t* tris, const float* pos,
                           const uint32_t* type, int T, const float* colors, int ncolors,
                           float* out) {
  for (long long t = 0; t < ntri; ++t) {
    const float* v[3] = {pos + 3 * tris[3 * t], pos + 3 * tris[3 * t + 1], pos + 3 * tris[3 * t + 2]};
    const float e1x = v[1][0] - v[0][0], e1y = v[1][1] - v[0][1], e1z = v[1][2] - v[0][2];
    const float e2x = v[2][0] - v[0][0], e2y = v[2][1] - v[0][1], e2z = v[2][2] - v[0][2];
    float nx = e1y * e2z - e1z * e2y;
    float ny = e1z * e2x - e1x * e2z;
    float nz = e1x * e2y - e1y * e2x;
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    if (len > 0.0001f) {
      nx /= len;
      ny /= len;
      nz /= len;
    }
    const uint32_t c = type[tris[3 * t]] % (uint32_t)T;
    float col[3] = {0.0f, 0.0f, 0.0f};
    if ((int)c < ncolors)
      for (int k = 0; k < 3; ++k) col[k] = colors[3 * c + k];
    for (int k = 0; k < 3; ++k) {
      float* o = out + 27 * t + 9 * k;
      o[0] = v[k][0];
      o[1] = v[k][1];
      o[2] = v[k][2];
      o[3] = nx;
      o[4] = ny;
      o[5] = nz;
      o[6] = col[0];
      o[7] = col[1];
      o[8] = col[2];
    }
  }
}

// updateForceTable (ParticleSimulation.cu:521-531): clamp(tanh(raw*offset)*range+bias, -1, 1).
// The operands are float, so the float overloads of tanh/fmax/fmin apply; all in FP32.
void orc_force_table(const float* raw, int count, float range, float bias, float offset, float* out) {
  for (int i = 0; i < count; ++i) {
    float transformed = tanhf(raw[i] * offset) * range + bias;
    out[i] = fmaxf(-1.0f, fminf(1.0f, transformed));
  }
}

// moveParticlesKernel, ParticleSimulation.cu:169-185: x = fmodf(x + d + L, L) per axis, FP32,
// evaluated left to right.
void orc_move(int n, float* pos, float dx, float dy, float dz, const float* L) {
  const float d[3] = {dx, dy, dz};
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) pos[3 * i + a] = fmodf(pos[3 * i + a] + d[a] + L[a], L[a]);
}

// ---- differentiable rollout (engine extension, no reference counterpart) ---------------
//
// Soft pair histogram: h_b = (1/N) sum_i sum_{j != i} B3((r_ij - c_b) / w), c_b = (b+0.5) w,
// w = r_max / bins, r_ij the minimum-image distance (no softening), B3 the cubic B-spline
// on [-2, 2].  Loss = sum_b (h_b - target_b)^2.

static inline double bspline3(double u) {
  u = std::fabs(u);
  if (u >= 2.0) return 0.0;
  if (u >= 1.0) {
    const double t = 2.0 - u;
    return t * t * t / 6.0;
  }
  return (4.0 - 6.0 * u * u + 3.0 * u * u * u) / 6.0;
}

// FP64 soft histogram of FP64 positions (all pairs within r_max + 2w, grid enumeration).
int orc_hist(int n, const double* pos, const float* L, int bins, double r_max, double* hist,
             int nthreads) {
  set_threads(nthreads);
  const double w = r_max / bins;
  const double reach = r_max + 2.0 * w;
  std::vector<float> posf(3 * (size_t)n);
  for (size_t k = 0; k < 3 * (size_t)n; ++k) posf[k] = (float)pos[k];
  Grid g = build_grid(n, posf.data(), L, reach);
  int r = 1;
  for (int a = 0; a < 3; ++a) r = std::max(r, (int)std::ceil(reach / g.h[a] + 1e-9));
  std::vector<double> part((size_t)n * bins, 0.0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    double* hi = &part[(size_t)i * bins];
    for_each_near(g, posf.data(), i, r, true, [&](int j) {
      if (j == i) return;
      double s = 0;
      for (int a = 0; a < 3; ++a) {
        const double d = min_image_d(pos[3 * j + a] - pos[3 * i + a], (double)L[a]);
        s += d * d;
      }
      const double rij = std::sqrt(s);
      if (rij >= reach) return;
      const double u = rij / w - 0.5;
      const int b0 = (int)std::floor(u) - 1;
      for (int b = std::max(b0, 0); b <= std::min(b0 + 3, bins - 1); ++b) hi[b] += bspline3(u - b);
    });
  }
  for (int b = 0; b < bins; ++b) {
    double acc = 0;
    for (int i = 0; i < n; ++i) acc += part[(size_t)i * bins + b];
    hist[b] = acc / n;
  }
  return 0;
}

// Neighbour lists of one step by the FP32 predicate on FP32 positions (the decisions the
// engine makes for that state): CSR offsets (n+1) and neighbour indices, j ascending.
int orc_pairs(int n, const float* pos, const uint32_t* type, const orc_params* pp,
              const float* rbt, int64_t* offs, int32_t* nbr, int64_t cap, int nthreads) {
  const orc_params& p = *pp;
  set_threads(nthreads);
  double reach = max_eff_radius(p, rbt);
  std::vector<std::vector<int32_t>> lists(n);
  if (reach > 0) {
    Grid g = build_grid(n, pos, p.L, reach);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
      auto& li = lists[i];
      for_each_near(g, pos, i, 1, true, [&](int j) {
        if (j != i && pair_in_range_f(p, rbt, &pos[3 * i], &pos[3 * j], type[i], type[j]))
          li.push_back(j);
      });
      std::sort(li.begin(), li.end());
    }
  }
  offs[0] = 0;
  for (int i = 0; i < n; ++i) offs[i + 1] = offs[i] + (int64_t)lists[i].size();
  if (offs[n] > cap) return 3;
  for (int i = 0; i < n; ++i)
    std::copy(lists[i].begin(), lists[i].end(), nbr + offs[i]);
  return 0;
}

// Frozen-decision FP64 rollout: step k uses the neighbour lists (offs[k], nbr[k]) and their
// sizes as counts, whatever the continuous state does; table and ratio enter only the
// continuous force.  Returns the loss of the final state and (optionally) the final state.
int orc_frozen_rollout(int n, const double* pos0, const double* vel0, const uint32_t* type,
                       const int32_t* prev0, const orc_params* pp, const float* table_f,
                       const double* table_d, double ratio, const float* rbt, int steps,
                       const int64_t* const* offs, const int32_t* const* nbr, int bins,
                       double r_max, const double* target, double* loss_out, double* pos_out,
                       double* hist_out, int nthreads) {
  const orc_params& p = *pp;
  set_threads(nthreads);
  std::vector<double> pos(pos0, pos0 + 3 * (size_t)n), vel(vel0, vel0 + 3 * (size_t)n);
  std::vector<double> pos2(3 * (size_t)n), vel2(3 * (size_t)n);
  std::vector<int32_t> prev(prev0, prev0 + n);
  const int T = p.T;
  (void)table_f;
  for (int s = 0; s < steps; ++s) {
    const int64_t* o = offs[s];
    const int32_t* nb = nbr[s];
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
      const unsigned ti = type[i];
      double F[3] = {0, 0, 0};
      for (int64_t q = o[i]; q < o[i + 1]; ++q) {
        const int j = nb[q];
        const unsigned tj = type[j];
        double d[3];
        for (int a = 0; a < 3; ++a) d[a] = min_image_d(pos[3 * j + a] - pos[3 * i + a], (double)p.L[a]);
        const double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + (double)0.0001f);
        const double R = p.radius;
        const double eff = (R * (1.0 + (double)rbt[ti] * ratio) + R * (1.0 + (double)rbt[tj] * ratio)) * 0.5;
        const double r = dist / eff;
        const double net = ((double)p.repulsion * std::exp(-(double)p.k * r * r) - (double)p.attraction * r) * table_d[ti * T + tj];
        for (int a = 0; a < 3; ++a) F[a] += d[a] / dist * net;
      }
      const int count = (int)(o[i + 1] - o[i]);
      const double avg = (double)(count + prev[i]) * 0.5;
      const double dens = std::fmin(avg / (double)p.max_expected_neighbors, 1.0);
      const double fac = 1.0 - (1.0 - (double)p.balance) * dens;
      for (int a = 0; a < 3; ++a) {
        const double v = vel[3 * i + a] * (double)p.friction + F[a] * (double)p.force_multiplier * fac * (double)p.dt;
        const double L = (double)p.L[a];
        vel2[3 * i + a] = v;
        pos2[3 * i + a] = std::fmod(pos[3 * i + a] + v * (double)p.dt + L, L);
      }
    }
    pos.swap(pos2);
    vel.swap(vel2);
    for (int i = 0; i < n; ++i) prev[i] = (int32_t)(o[i + 1] - o[i]);
  }
  std::vector<double> h(bins);
  orc_hist(n, pos.data(), p.L, bins, r_max, h.data(), nthreads);
  double loss = 0;
  for (int b = 0; b < bins; ++b) loss += (h[b] - target[b]) * (h[b] - target[b]);
  *loss_out = loss;
  if (pos_out) std::copy(pos.begin(), pos.end(), pos_out);
  if (hist_out) std::copy(h.begin(), h.end(), hist_out);
  return 0;
}

}  // extern "C"
