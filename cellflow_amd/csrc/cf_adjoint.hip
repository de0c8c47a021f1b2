// Reverse-mode (adjoint) kernels of the differentiable rollout (no reference counterpart).
//
// The forward rollout is the ordinary FP32 step.  Its discrete decisions (which pairs are
// within the cutoff, the neighbour counts and hence the density factor) are frozen: the
// adjoint differentiates the smooth force and integration of ParticleSimulation.cu:116-161
// with respect to the positions, velocities, the force table and the ratio.  Decisions
// are recomputed bit for bit from the stored FP32 states with the same FP32 predicate.  The
// continuous state the loss and the adjoint are evaluated on is an FP64 shadow of the
// rollout along those decisions (k_shadow_step), so the gradient is exactly that of one
// FP64 function; all adjoint arithmetic is FP64.
//
// One forward step for particle i (pre-step state x, v; counts c, prev frozen):
//   F_i = sum_j f_ij,  f_ij = fv_ij g(r_ij) d_ij / dist_ij,  d_ij = minimage(x_j - x_i),
//   dist = sqrt(|d|^2 + 1e-4),  r = dist / effR_ij,  g(r) = rep exp(-k r^2) - att r,
//   v' = fr v + dt m_i F_i,  x' = wrap(x + dt v'),  m_i = fm fac(c_i + prev_i).
// Backward, with lx' = dL/dx', lv' = dL/dv':
//   mu_i = lv'_i + dt lx'_i,  phi_i = dt m_i mu_i (= dL/dF_i),  lv_i = fr mu_i,
//   lx_i = lx'_i + sum_{j in N_i} (J_ji phi_j - J_ij phi_i),  J = df/dd (symmetric),
//   dL/dfv[ti][tj] += phi_i . g d / dist,   dL/dratio += phi_i . fv dg/deffR deffR/dratio d/dist.
//
// Loss of the final state: soft pair histogram h_b = (1/N) sum_i sum_{j != i} B3((r_ij - c_b)/w)
// (cubic B-spline, c_b = (b + 1/2) w, w = r_max / bins, r_ij the minimum-image distance),
// L = sum_b (h_b - target_b)^2.

#include "cf_internal.h"

namespace cf {

__device__ __forceinline__ double bspline3(double u) {
  u = fabs(u);
  if (u >= 2.0) return 0.0;
  if (u >= 1.0) {
    const double t = 2.0 - u;
    return t * t * t / 6.0;
  }
  return (4.0 - 6.0 * u * u + 3.0 * u * u * u) / 6.0;
}

__device__ __forceinline__ double bspline3_d(double u) {  // dB3/du
  const double a = fabs(u), s = (u < 0.0) ? -1.0 : 1.0;
  if (a >= 2.0) return 0.0;
  if (a >= 1.0) {
    const double t = 2.0 - a;
    return -0.5 * t * t * s;
  }
  return (-2.0 * a + 1.5 * a * a) * s;
}

// fmod(y, L) for L > 0, exactly: for y in [0, 2L) it is y or y - L (the subtraction is exact
// there), else fmod itself.
__device__ __forceinline__ double wrap_d(double y, double L) {
  if (y >= 0.0 && y < 2.0 * L) return y >= L ? y - L : y;
  return fmod(y, L);
}

__device__ __forceinline__ double min_image_d(double d, double L) {
  if (d > L * 0.5) d -= L;
  if (d < -L * 0.5) d += L;
  return d;
}

__device__ __forceinline__ float min_image_f(float d, float L) {
  if (d > L * 0.5f) d -= L;
  if (d < -L * 0.5f) d += L;
  return d;
}

__device__ __forceinline__ int bin_axis_a(float x, float ih, int n) {
  float u = x * ih;
  if (!(u >= 0.0f)) return 0;
  if (u >= (float)n) return n - 1;
  return (int)u;
}

// ---- histogram and its gradient ------------------------------------------------------

// Fixed-order sum of `n` values with stride, one output per block (deterministic).
__global__ void __launch_bounds__(256) k_reduce_rows(const double* __restrict__ in, long stride,
                                                      int n, double scale, double* out,
                                                      int accumulate) {
  __shared__ double sm[256];
  const int row = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += in[(long)row * stride + i];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[row] = (accumulate ? out[row] : 0.0) + sm[0] * scale;
}

// ---- one backward step -----------------------------------------------------------------

// The FP64 density factor of a frozen count (the oracle's op order).
__device__ __forceinline__ double shadow_fac(const AdjConst& c, int cnt, int prev) {
  const double avg = (double)(cnt + prev) * 0.5;
  const double dens = fmin(avg / c.max_exp, 1.0);
  return 1.0 - (1.0 - c.balance) * dens;
}

// phi_i = dt m_i (lv'_i + dt lx'_i),  lv_i = fr (lv'_i + dt lx'_i) of particle k; phi is
// stored by the pre-step slot, where the pair kernel finds both ends of a pair without an
// index lookup.  The walk kernel computes it on the side (it needs nothing the walk makes).
__device__ __forceinline__ void adj_phi(const AdjArgs& a, const SimDesc& sd, int k) {
  const int p = sd.p0 + k;
  const uint32_t id = a.id[p];
  const int prev = __float_as_int(a.vel[p].w);
  const int cnt = __float_as_int(a.tvel[p].w);  // this step's count, by pre-step slot
  // the shadow's FP64 factor of the frozen counts (ParticleSimulation.cu:136-143)
  const double m = a.c.fm * shadow_fac(a.c, cnt, prev);
  const double4 lx = a.lam_x_next[id], lv = a.lam_v_next[id];
  const double dt = a.c.dt, fr = a.c.friction;
  const double mx = lv.x + dt * lx.x, my = lv.y + dt * lx.y, mz = lv.z + dt * lx.z;
  a.phi[p] = make_double4(dt * m * mx, dt * m * my, dt * m * mz, 0.0);  // by pre-step slot
  a.lam_v[id] = make_double4(fr * mx, fr * my, fr * mz, 0.0);
}

// ---- pair kernels: G lanes per particle -----------------------------------------------
//
// One lane per particle would leave most of the GPU idle at a few tens of thousands of
// particles and walk every candidate serially.  Instead the stencil's cells are split
// over G lanes (cell index c = g, g+G, ...) and the lanes' partial sums are reduced in a
// fixed order, so results are deterministic.

constexpr int kAdjG = 8;                       // lanes per particle
constexpr int kAdjBlock = 128;                 // threads per block
constexpr int kAdjPer = kAdjBlock / kAdjG;     // particles per block
constexpr int kAdjVals = 4 + kMaxTypes;        // lx, ly, lz, ratio, table row

// Visit the stencil cells c = first, first + step, ... of the (2r+1)^3 cube around
// (cx, cy, cz) (every cell once on axes with fewer than 2r+1 cells), in a fixed order:
// for_cell_runs passes each cell's slot range, for_cells_split each of its slots.  Cell
// coordinates come from float reciprocals: (c + 1/2) / m is at least 1/(2m) away from an
// integer and m <= 25, far beyond the float error.
template <class Fn>
__device__ __forceinline__ void for_cell_runs(const SimDesc& sd, const uint32_t* start, int cx,
                                              int cy, int cz, int r, int first, int step,
                                              Fn&& fn) {
  const bool gx = sd.nx < 2 * r + 1, gy = sd.ny < 2 * r + 1, gz = sd.nz < 2 * r + 1;
  const int nxo = gx ? sd.nx : 2 * r + 1, nyo = gy ? sd.ny : 2 * r + 1,
            nzo = gz ? sd.nz : 2 * r + 1;
  const int ncell = nxo * nyo * nzo, nxy = nxo * nyo;
  const float inxy = 1.0f / (float)nxy, inx = 1.0f / (float)nxo;
  for (int ci = first; ci < ncell; ci += step) {
    const int iz = (int)(((float)ci + 0.5f) * inxy);
    const int rem = ci - iz * nxy;
    const int iy = (int)(((float)rem + 0.5f) * inx);
    const int ix = rem - iy * nxo;
    int x = gx ? ix : cx + ix - r, y = gy ? iy : cy + iy - r, z = gz ? iz : cz + iz - r;
    if (x < 0) x += sd.nx; else if (x >= sd.nx) x -= sd.nx;
    if (y < 0) y += sd.ny; else if (y >= sd.ny) y -= sd.ny;
    if (z < 0) z += sd.nz; else if (z >= sd.nz) z -= sd.nz;
    const int cc = sd.c0 + x + sd.nx * (y + sd.ny * z);
    fn(start[cc], start[cc + 1]);
  }
}

template <class Fn>
__device__ __forceinline__ void for_cells_split(const SimDesc& sd, const uint32_t* start, int cx,
                                                int cy, int cz, int r, int first, int step,
                                                Fn&& fn) {
  for_cell_runs(sd, start, cx, cy, cz, r, first, step, [&](uint32_t q0, uint32_t q1) {
    for (uint32_t q = q0; q < q1; ++q) fn(q);
  });
}

// Histogram partials, G lanes per particle.  The group's lanes add each pair's four spline
// weights into the particle's bins in LDS as 64-bit fixed point, rounded once per weight:
// integer addition does not depend on order, so the result is deterministic whatever the
// lanes' interleaving.  The unit is 2^-(62 - bitlen(n)): a weight is at most 1 and a bin
// gathers at most n - 1 of them, so no sum can overflow, and at n = 20k one weight is held
// to 2^-47.  A pair is skipped early when its FP32 squared distance exceeds the reach
// squared by more than FP32 rounding can explain.
constexpr int kHistG = 32;
constexpr int kHistBlock = 256;

// The FP64 position of slot q for the histogram: the shadow's when given, else the FP32
// state's.
__device__ __forceinline__ double4 hist_pos(const HistArgs& a, int q, const float4& x) {
  if (a.x64) return a.x64[a.id[q]];
  return make_double4((double)x.x, (double)x.y, (double)x.z, 0.0);
}

__global__ void __launch_bounds__(kHistBlock) k_hist_g(HistArgs a) {
  constexpr int G = kHistG, kPer = kHistBlock / G;
  __shared__ unsigned long long hq[kPer * kMaxBins];
  const SimDesc sd = a.sims[0];
  const int o = threadIdx.x / G, g = threadIdx.x % G;
  const int k = blockIdx.x * kPer + o;
  for (int i = threadIdx.x; i < kPer * kMaxBins; i += kHistBlock) hq[i] = 0ull;
  __syncthreads();
  unsigned long long* h = &hq[o * kMaxBins];
  const double unit = ldexp(1.0, 62 - (32 - __clz(sd.n)));
  const double w = a.r_max / a.bins, reach = a.r_max + 2.0 * w;
  const float reach2f = (float)((reach + a.slack) * (reach + a.slack) * (1.0 + 1e-5));
  if (k < sd.n) {
    const int p = sd.p0 + k;
    const float4 pi = a.pos[p];
    const double4 xi = hist_pos(a, p, pi);
    const int cx = bin_axis_a(pi.x, sd.ihx, sd.nx), cy = bin_axis_a(pi.y, sd.ihy, sd.ny),
              cz = bin_axis_a(pi.z, sd.ihz, sd.nz);
    for_cells_split(sd, a.start, cx, cy, cz, a.reach, g, G, [&](uint32_t q) {
      if ((int)q == p) return;
      const float4 pj = a.pos[q];
      const float fx = min_image_f(pj.x - pi.x, sd.Lx), fy = min_image_f(pj.y - pi.y, sd.Ly),
                  fz = min_image_f(pj.z - pi.z, sd.Lz);
      if (fx * fx + fy * fy + fz * fz > reach2f) return;
      const double4 xj = hist_pos(a, q, pj);
      const double dx = min_image_d(xj.x - xi.x, sd.Lx);
      const double dy = min_image_d(xj.y - xi.y, sd.Ly);
      const double dz = min_image_d(xj.z - xi.z, sd.Lz);
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r >= reach) return;
      const double u = r / w - 0.5;
      const int b0 = (int)floor(u) - 1;
      for (int b = max(b0, 0); b <= min(b0 + 3, a.bins - 1); ++b)
        atomicAdd(&h[b], (unsigned long long)llrint(bspline3(u - b) * unit));
    });
  }
  __syncthreads();
  if (k < sd.n)
    for (int b = g; b < a.bins; b += G) a.partial[(long)b * a.stride + k] = (double)h[b] / unit;
}

// lx of the final state, G lanes per particle (fixed-order group sum by shuffles).
__global__ void __launch_bounds__(128) k_hist_grad_g(HistArgs a) {
  constexpr int G = kHistG;
  const SimDesc sd = a.sims[0];
  const int o = threadIdx.x / G, g = threadIdx.x % G;
  const int k = blockIdx.x * (128 / G) + o;
  const bool active = k < sd.n;
  const double w = a.r_max / a.bins, reach = a.r_max + 2.0 * w;
  const float reach2f = (float)((reach + a.slack) * (reach + a.slack) * (1.0 + 1e-5));
  double gx = 0.0, gy = 0.0, gz = 0.0;
  const int p = sd.p0 + k;
  float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (active) {
    pi = a.pos[p];
    const double4 xi = hist_pos(a, p, pi);
    const int cx = bin_axis_a(pi.x, sd.ihx, sd.nx), cy = bin_axis_a(pi.y, sd.ihy, sd.ny),
              cz = bin_axis_a(pi.z, sd.ihz, sd.nz);
    for_cells_split(sd, a.start, cx, cy, cz, a.reach, g, G, [&](uint32_t q) {
      if ((int)q == p) return;
      const float4 pj = a.pos[q];
      const float fx = min_image_f(pj.x - pi.x, sd.Lx), fy = min_image_f(pj.y - pi.y, sd.Ly),
                  fz = min_image_f(pj.z - pi.z, sd.Lz);
      if (fx * fx + fy * fy + fz * fz > reach2f) return;
      const double4 xj = hist_pos(a, q, pj);
      const double dx = min_image_d(xj.x - xi.x, sd.Lx);
      const double dy = min_image_d(xj.y - xi.y, sd.Ly);
      const double dz = min_image_d(xj.z - xi.z, sd.Lz);
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r >= reach || r <= 0.0) return;
      const double u = r / w - 0.5;
      const int b0 = (int)floor(u) - 1;
      double s = 0.0;
      for (int b = max(b0, 0); b <= min(b0 + 3, a.bins - 1); ++b)
        s += a.gh[b] * bspline3_d(u - b);
      s *= 2.0 / ((double)sd.n * w * r);
      gx -= s * dx;
      gy -= s * dy;
      gz -= s * dz;
    });
  }
#pragma unroll
  for (int m = G / 2; m > 0; m >>= 1) {
    gx += __shfl_xor(gx, m, 64);
    gy += __shfl_xor(gy, m, 64);
    gz += __shfl_xor(gz, m, 64);
  }
  if (active && g == 0) {
    const uint32_t id = a.id[p];
    a.lam_x[id] = make_double4(gx, gy, gz, 0.0);
    a.lam_v[id] = make_double4(0.0, 0.0, 0.0, 0.0);
  }
}

// Exclusive prefix sum over aligned groups of G lanes (g = lane index in the group); every
// lane of the wave takes part.
template <int G>
__device__ __forceinline__ uint32_t group_excl_scan(uint32_t v, int g) {
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, 64);
    if (g >= d) x += y;
  }
  return x - v;
}

constexpr int kAdjHits = kPairList;  // also the stride of the force kernel's lists

// The pair kernel's particle context and its FP32 in-range decision, shared by the walk
// kernel and the pair kernel.
struct AdjOwn {
  float4 pi;
  int p, ti, cx, cy, cz;
};

__device__ __forceinline__ AdjOwn adj_own(const AdjArgs& a, const SimDesc& sd, int k, bool active) {
  AdjOwn c;
  c.p = sd.p0 + k;
  c.pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  c.ti = -1;
  c.cx = c.cy = c.cz = 0;
  if (active) {
    c.pi = a.pos[c.p];
    c.ti = (int)__float_as_uint(c.pi.w);
    c.cx = bin_axis_a(c.pi.x, sd.ihx, sd.nx);
    c.cy = bin_axis_a(c.pi.y, sd.ihy, sd.ny);
    c.cz = bin_axis_a(c.pi.z, sd.ihz, sd.nz);
  }
  return c;
}

// the forward's FP32 decision (ParticleSimulation.cu:92-112 with the engine's op order)
__device__ __forceinline__ bool adj_in_range(const AdjArgs& a, const SimDesc& sd, const AdjOwn& c,
                                             uint32_t q) {
  if ((int)q == c.p) return false;
  const float4 pj = a.pos[q];
  const int tj = (int)__float_as_uint(pj.w);
  const float fx = min_image_f(pj.x - c.pi.x, sd.Lx), fy = min_image_f(pj.y - c.pi.y, sd.Ly),
              fz = min_image_f(pj.z - c.pi.z, sd.Lz);
  const float s2f = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
  return s2f + 1e-4f < a.thr[c.ti * a.c.T + tj];
}

// ---- the FP64 shadow of the rollout ------------------------------------------------------
//
// The forward rollout is the FP32 step; its decisions (in-range pairs, counts) define the
// function the gradient is of.  The shadow carries the same rollout in FP64 along those
// decisions: x64' = wrap(x64 + dt v64'), v64' = fr v64 + dt fm fac(c + prev) F64, with F64
// summed over the FP32 step's in-range pairs from FP64 positions (ParticleSimulation.cu:
// 92-161 with the cutoff test of the FP32 state).  The loss and every adjoint term are
// evaluated on the shadow, so the gradient is that of one smooth FP64 function of the table
// and the ratio, the function central differences of the oracle's frozen-decision rollout
// evaluate.  Shadow states are by particle index.

__global__ void __launch_bounds__(256) k_shadow_init(AdjArgs a) {
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  const float4 x = a.pos[p], v = a.vel[p];
  const uint32_t id = a.id[p];
  a.x64_next[id] = make_double4((double)x.x, (double)x.y, (double)x.z, 0.0);
  a.v64_next[id] = make_double4((double)v.x, (double)v.y, (double)v.z, 0.0);
}

// One FP64 shadow step, G lanes per particle; lane sums are reduced in a fixed order.  The
// step's in-range pairs come from the lists the FP32 step's force kernel wrote (a.hits, at
// most kAdjHits per particle, in slot order), or from a walk of the FP32 state's stencil with
// the forward's FP32 decision (no lists, or a particle with more pairs than a list holds).
__global__ void __launch_bounds__(kAdjBlock) k_shadow_step(AdjArgs a) {
  constexpr int G = kAdjG;
  const SimDesc sd = a.sims[0];
  const int o = threadIdx.x / G, g = threadIdx.x % G;
  const int k = blockIdx.x * kAdjPer + o;
  const bool active = k < sd.n;
  const AdjOwn c = adj_own(a, sd, k, active);
  const int T = a.c.T;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  int cnt = 0;
  uint32_t idi = 0;
  double4 xi = make_double4(0.0, 0.0, 0.0, 0.0);
  const double R = a.c.radius, rho = a.c.ratio;
  const double effi = active ? R * (1.0 + a.c.rbt[c.ti] * rho) : 0.0;
  // the FP64 force term of one in-range pair (the oracle's op order), partner type tj at
  // shadow position xj
  auto pair = [&](int tj, const double4& xj) {
    const double dx = min_image_d(xj.x - xi.x, sd.Lx);
    const double dy = min_image_d(xj.y - xi.y, sd.Ly);
    const double dz = min_image_d(xj.z - xi.z, sd.Lz);
    const double dist = sqrt(dx * dx + dy * dy + dz * dz + (double)1e-4f);
    const double eff = (effi + R * (1.0 + a.c.rbt[tj] * rho)) * 0.5;
    const double r = dist / eff;
    const double net = (a.c.rep * exp(-a.c.k * r * r) - a.c.att * r) * a.c.table[c.ti * T + tj];
    fx += dx / dist * net;
    fy += dy / dist * net;
    fz += dz / dist * net;
  };
  int prev = 0;
  double4 vi = make_double4(0.0, 0.0, 0.0, 0.0);
  if (active) {
    idi = a.id[c.p];
    xi = a.x64[idi];
    if (g == 0) {  // the integration's inputs, loaded before the pairs are summed
      prev = __float_as_int(a.vel[c.p].w);
      vi = a.v64[idi];
    }
    const uint32_t nh = a.hits ? a.nhits[k] : 0u;
    if (a.hits && nh <= (uint32_t)kAdjHits) {
      if (g == 0) cnt = (int)nh;
      // the lane's entries two at a time, each stage's loads issued together (list entry,
      // then type and index, then shadow position); pairs summed in entry order
      const uint32_t* lst = a.hits + (long)k * kAdjHits;
#pragma nounroll
      for (uint32_t e = g; e < nh; e += 2 * G) {
        const bool two = e + G < nh;
        const uint32_t q0 = lst[e], q1 = two ? lst[e + G] : q0;
        const int t0 = (int)__float_as_uint(a.pos[q0].w), t1 = (int)__float_as_uint(a.pos[q1].w);
        const uint32_t j0 = a.id[q0], j1 = a.id[q1];
        const double4 x0 = a.x64[j0], x1 = a.x64[j1];
        pair(t0, x0);
        if (two) pair(t1, x1);
      }
    } else {
      for_cells_split(sd, a.start, c.cx, c.cy, c.cz, sd.reach, g, G, [&](uint32_t q) {
        if (!adj_in_range(a, sd, c, q)) return;
        ++cnt;
        pair((int)__float_as_uint(a.pos[q].w), a.x64[a.id[q]]);
      });
    }
  }
#pragma unroll
  for (int m = G / 2; m > 0; m >>= 1) {
    fx += __shfl_xor(fx, m, 64);
    fy += __shfl_xor(fy, m, 64);
    fz += __shfl_xor(fz, m, 64);
    cnt += __shfl_xor(cnt, m, 64);
  }
  if (!active || g != 0) return;
  const double m = a.c.fm * shadow_fac(a.c, cnt, prev) * a.c.dt;
  const double vx = vi.x * a.c.friction + fx * m;
  const double vy = vi.y * a.c.friction + fy * m;
  const double vz = vi.z * a.c.friction + fz * m;
  a.v64_next[idi] = make_double4(vx, vy, vz, 0.0);
  a.x64_next[idi] = make_double4(wrap_d(xi.x + vx * a.c.dt + (double)sd.Lx, (double)sd.Lx),
                                 wrap_d(xi.y + vy * a.c.dt + (double)sd.Ly, (double)sd.Ly),
                                 wrap_d(xi.z + vz * a.c.dt + (double)sd.Lz, (double)sd.Lz), 0.0);
}

// Sets flag bit 3 when a shadow position lies more than slack/2 (periodic, per axis) from
// the FP32 position it shadows.
__global__ void __launch_bounds__(256) k_shadow_drift(AdjArgs a) {
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  const float4 x = a.pos[p];
  const double4 s = a.x64[a.id[p]];
  const double h = 0.5 * a.slack;
  const double dx = fabs(min_image_d(s.x - (double)x.x, sd.Lx));
  const double dy = fabs(min_image_d(s.y - (double)x.y, sd.Ly));
  const double dz = fabs(min_image_d(s.z - (double)x.z, sd.Lz));
  if (!(dx <= h && dy <= h && dz <= h)) atomicOr(a.flag, 8u);
}

// The walk: G lanes per particle take the FP32 decision over the stencil and list the
// in-range slots in LDS, which are then put in slot order (rank by counting, slots are
// distinct) and written to hits[k][0..); nhits[k] is the particle's pair count, also when
// it exceeds kAdjHits (the pair kernel then walks again).  No FP64 here, so the walk runs
// at the occupancy its few registers allow.
__global__ void __launch_bounds__(kAdjBlock) k_adj_walk(AdjArgs a) {
  constexpr int G = kAdjG;
  __shared__ uint32_t hit[kAdjPer][kAdjHits];
  __shared__ uint32_t nhit[kAdjPer];
  const SimDesc sd = a.sims[0];
  const int o = threadIdx.x / G, g = threadIdx.x % G;
  const int k = blockIdx.x * kAdjPer + o;
  const bool active = k < sd.n;
  const AdjOwn c = adj_own(a, sd, k, active);
  if (g == 0) nhit[o] = 0u;
  if (active && g == G - 1) adj_phi(a, sd, k);  // not the lead lane (g == 0) of the list
  __syncthreads();
  if (active)
    for_cells_split(sd, a.start, c.cx, c.cy, c.cz, sd.reach, g, G, [&](uint32_t q) {
      if (!adj_in_range(a, sd, c, q)) return;
      const uint32_t at = atomicAdd(&nhit[o], 1u);
      if (at < (uint32_t)kAdjHits) hit[o][at] = q;
    });
  __syncthreads();
  if (!active) return;
  const uint32_t nh = nhit[o];
  if (g == 0) a.nhits[k] = nh;
  if (nh <= (uint32_t)kAdjHits)
    for (uint32_t e = g; e < nh; e += G) {
      const uint32_t q = hit[o][e];
      uint32_t rank = 0;
      for (uint32_t e2 = 0; e2 < nh; ++e2) rank += (hit[o][e2] < q) ? 1u : 0u;
      a.hits[(long)k * kAdjHits + rank] = q;
    }
}

// One backward step's pair terms, G lanes per particle, over the walk's slot-ordered lists:
// entry e by lane e mod G.  A particle with more than kAdjHits pairs walks again and takes
// them in windows of kAdjHits in lane-major walk order (lane g's pairs after those of the
// lanes below g), one more walk per window.
// Per particle: lane sums reduced in lane order; per block: the table and ratio partials
// of its particles summed in particle order and added to one column of `partial` (row r at
// partial[r * stride + block]), which so sums the steps of the rollout in the order they are
// reversed.  Every order is fixed by the data, so the results are deterministic.
__global__ void __launch_bounds__(kAdjBlock) k_adj_pair_g(AdjArgs a) {
  constexpr int G = kAdjG;
  __shared__ double red[kAdjBlock][kAdjVals];
  __shared__ int btype[kAdjPer];
  __shared__ uint32_t hsorted[kAdjPer][kAdjHits];
  const SimDesc sd = a.sims[0];
  const int o = threadIdx.x / G, g = threadIdx.x % G;
  const int k = blockIdx.x * kAdjPer + o;
  const bool active = k < sd.n;
  const int T = a.c.T;
  double lx = 0.0, ly = 0.0, lz = 0.0, grat = 0.0;
  double gfv[kMaxTypes];
#pragma unroll
  for (int t = 0; t < kMaxTypes; ++t) gfv[t] = 0.0;
  const AdjOwn c = adj_own(a, sd, k, active);
  const int p = c.p, ti = c.ti;
  uint32_t idi = 0, nh = 0;
  double4 phi_i = make_double4(0.0, 0.0, 0.0, 0.0);
  double4 xi = make_double4(0.0, 0.0, 0.0, 0.0);
  if (active) {
    idi = a.id[p];
    phi_i = a.phi[p];
    nh = a.nhits[k];
    xi = a.x64[idi];
  }
  // the FP64 terms of one in-range pair: partner type tj, shadow position xj, factor phi_j
  auto pair = [&](int tj, const double4& xj, const double4& phi_j) {
    const double R = a.c.radius, rho = a.c.ratio, rep = a.c.rep, att = a.c.att, kk = a.c.k;
    const double eps = (double)1e-4f;
    const double dx = min_image_d(xj.x - xi.x, sd.Lx);
    const double dy = min_image_d(xj.y - xi.y, sd.Ly);
    const double dz = min_image_d(xj.z - xi.z, sd.Lz);
    const double dist = sqrt(dx * dx + dy * dy + dz * dz + eps);
    const double dR = 0.5 * R * (a.c.rbt[ti] + a.c.rbt[tj]);
    const double eff = R + dR * rho;
    const double r = dist / eff;
    const double ex = exp(-kk * r * r);
    const double gr = rep * ex - att * r;
    const double gp = -2.0 * kk * r * rep * ex - att;
    const double h = gr / dist;
    const double hp = gp / (eff * dist) - gr / (dist * dist);
    const double fij = a.c.table[ti * T + tj], fji = a.c.table[tj * T + ti];
    const double dpi = (dx * phi_i.x + dy * phi_i.y + dz * phi_i.z) / dist;
    const double dpj = (dx * phi_j.x + dy * phi_j.y + dz * phi_j.z) / dist;
    lx += fji * (h * phi_j.x + hp * dx * dpj) - fij * (h * phi_i.x + hp * dx * dpi);
    ly += fji * (h * phi_j.y + hp * dy * dpj) - fij * (h * phi_i.y + hp * dy * dpi);
    lz += fji * (h * phi_j.z + hp * dz * dpj) - fij * (h * phi_i.z + hp * dz * dpi);
    const double gd = gr * dpi;
#pragma unroll
    for (int t = 0; t < kMaxTypes; ++t) gfv[t] += (t == tj) ? gd : 0.0;
    grat += fij * (-gp * r / eff) * dR * dpi;
  };
  const bool listed = nh <= (uint32_t)kAdjHits;
  // a particle beyond the list: this lane's pairs, and where they start in lane-major order
  uint32_t nmine = 0u;
  if (active && !listed)
    for_cells_split(sd, a.start, c.cx, c.cy, c.cz, sd.reach, g, G, [&](uint32_t q) {
      if (adj_in_range(a, sd, c, q)) ++nmine;
    });
  const uint32_t first = group_excl_scan<G>(nmine, g);
  auto pair_q = [&](uint32_t q) {
    pair((int)__float_as_uint(a.pos[q].w), a.x64[a.id[q]], a.phi[q]);
  };
  if (listed) {
    // the lane's entries two at a time, each stage's loads issued together (list entry, then
    // type, index and factor, then shadow position); pairs summed in entry order
    const uint32_t* lst = a.hits + (long)k * kAdjHits;
#pragma nounroll
    for (uint32_t e = g; e < nh; e += 2 * G) {
      const bool two = e + G < nh;
      const uint32_t q0 = lst[e], q1 = two ? lst[e + G] : q0;
      const int t0 = (int)__float_as_uint(a.pos[q0].w), t1 = (int)__float_as_uint(a.pos[q1].w);
      const uint32_t j0 = a.id[q0], j1 = a.id[q1];
      const double4 f0 = a.phi[q0], f1 = a.phi[q1];
      const double4 x0 = a.x64[j0], x1 = a.x64[j1];
      pair(t0, x0, f0);
      if (two) pair(t1, x1, f1);
    }
  } else {
    // a group's lanes share one wave: LDS writes and reads between them are ordered by
    // wavefront fences, so the windows need no block barrier
    for (uint32_t w0 = 0; w0 < nh; w0 += (uint32_t)kAdjHits) {
      uint32_t idx = first;
      for_cells_split(sd, a.start, c.cx, c.cy, c.cz, sd.reach, g, G, [&](uint32_t q) {
        if (!adj_in_range(a, sd, c, q)) return;
        if (idx >= w0 && idx < w0 + (uint32_t)kAdjHits) hsorted[o][idx - w0] = q;
        ++idx;
      });
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const uint32_t nb = min(nh - w0, (uint32_t)kAdjHits);
#pragma nounroll
      for (uint32_t e = g; e < nb; e += G) pair_q(hsorted[o][e]);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // window reused by the next
    }
  }
  double* mine = red[threadIdx.x];
  mine[0] = lx;
  mine[1] = ly;
  mine[2] = lz;
  mine[3] = grat;
#pragma unroll
  for (int t = 0; t < kMaxTypes; ++t) mine[4 + t] = gfv[t];
  if (g == 0) btype[o] = active ? ti : -1;
  __syncthreads();
  // lane g of each group sums values g, g + G, ... over the group's lanes (only lane v mod G
  // touches column v of the group's first row)
  for (int v = g; v < kAdjVals; v += G) {
    double acc = 0.0;
    for (int l = 0; l < G; ++l) acc += red[o * G + l][v];
    red[o * G][v] = acc;
  }
  __syncthreads();
  if (active && g == 0) {
    const double4 lxn = a.lam_x_next[idi];
    a.lam_x[idi] = make_double4(lxn.x + red[o * G][0], lxn.y + red[o * G][1],
                                lxn.z + red[o * G][2], 0.0);
  }
  // block partials: row (a2 * T + t) sums the table term of particles of type a2, row T*T
  // the ratio term, each over the block's particles in order
  const int rows = T * T + 1;
  for (int row = threadIdx.x; row < rows; row += kAdjBlock) {
    double acc = 0.0;
    if (row < T * T) {
      const int a2 = row / T, t = row % T;
      for (int oo = 0; oo < kAdjPer; ++oo)
        if (btype[oo] == a2) acc += red[oo * G][4 + t];
    } else {
      for (int oo = 0; oo < kAdjPer; ++oo)
        if (btype[oo] >= 0) acc += red[oo * G][3];
    }
    a.partial[(long)row * a.stride + blockIdx.x] += acc;
  }
}

int adj_pair_blocks(int n) { return (n + kAdjPer - 1) / kAdjPer; }

hipError_t launch_hist(const HistArgs& a, int n, hipStream_t st) {
  constexpr int per = kHistBlock / kHistG;
  hipLaunchKernelGGL(k_hist_g, dim3((n + per - 1) / per), dim3(kHistBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_hist_grad(const HistArgs& a, int n, hipStream_t st) {
  constexpr int per = 128 / kHistG;
  hipLaunchKernelGGL(k_hist_grad_g, dim3((n + per - 1) / per), dim3(128), 0, st, a);
  return hipGetLastError();
}

// The density-factor terms alone, for steps whose lists the recompute's force kernel wrote.
__global__ void __launch_bounds__(128) k_adj_phi(AdjArgs a) {
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < sd.n) adj_phi(a, sd, k);
}

hipError_t launch_adj_step(const AdjArgs& a, int n, hipStream_t st, bool walk) {
  if (!walk) {
    hipLaunchKernelGGL(k_adj_phi, dim3((n + 127) / 128), dim3(128), 0, st, a);
    hipLaunchKernelGGL(k_adj_pair_g, dim3(adj_pair_blocks(n)), dim3(kAdjBlock), 0, st, a);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_adj_walk, dim3(adj_pair_blocks(n)), dim3(kAdjBlock), 0, st, a);
  hipLaunchKernelGGL(k_adj_pair_g, dim3(adj_pair_blocks(n)), dim3(kAdjBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_shadow_init(const AdjArgs& a, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_shadow_init, dim3((n + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_shadow_step(const AdjArgs& a, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_shadow_step, dim3((n + kAdjPer - 1) / kAdjPer), dim3(kAdjBlock), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_shadow_drift(const AdjArgs& a, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_shadow_drift, dim3((n + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_reduce_rows(const double* in, long stride, int n, int rows, double scale,
                              double* out, int accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_rows, dim3(rows), dim3(256), 0, st, in, stride, n, scale, out,
                     accumulate);
  return hipGetLastError();
}

}  // namespace cf
