// Reverse-mode (adjoint) kernels of the differentiable rollout (no reference counterpart).
//
// The forward rollout is the ordinary FP32 step.  Its discrete decisions (which pairs are
// within the cutoff, the neighbour counts and hence the density factor) are frozen: the
// adjoint differentiates the smooth force and integration of ParticleSimulation.cu:116-161
// with respect to the positions, velocities, the force table and the ratio.  Decisions
// are recomputed bit for bit from the stored FP32 states with the same FP32 predicate; all
// adjoint arithmetic is FP64.
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

// Visit every slot of the cells within `r` cells (periodic, each cell once) of cell
// (cx, cy, cz) in a fixed order.
template <class Fn>
__device__ __forceinline__ void for_cells(const SimDesc& sd, const uint32_t* start, int cx, int cy,
                                          int cz, int r, Fn&& fn) {
  const int nzo = (sd.nz < 2 * r + 1) ? sd.nz : 2 * r + 1;
  const int nyo = (sd.ny < 2 * r + 1) ? sd.ny : 2 * r + 1;
  const int nxo = (sd.nx < 2 * r + 1) ? sd.nx : 2 * r + 1;
  for (int iz = 0; iz < nzo; ++iz) {
    int z = (sd.nz < 2 * r + 1) ? iz : cz + iz - r;
    z = (z % sd.nz + sd.nz) % sd.nz;
    for (int iy = 0; iy < nyo; ++iy) {
      int y = (sd.ny < 2 * r + 1) ? iy : cy + iy - r;
      y = (y % sd.ny + sd.ny) % sd.ny;
      for (int ix = 0; ix < nxo; ++ix) {
        int x = (sd.nx < 2 * r + 1) ? ix : cx + ix - r;
        x = (x % sd.nx + sd.nx) % sd.nx;
        const int c = sd.c0 + x + sd.nx * (y + sd.ny * z);
        for (uint32_t q = start[c]; q < start[c + 1]; ++q) fn(q);
      }
    }
  }
}

// ---- histogram and its gradient ------------------------------------------------------

__global__ void __launch_bounds__(128) k_hist(HistArgs a) {
  __shared__ double part[128][kMaxBins];
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double* h = part[threadIdx.x];
  for (int b = 0; b < a.bins; ++b) h[b] = 0.0;
  if (k < sd.n) {
    const int p = sd.p0 + k;
    const float4 pi = a.pos[p];
    const int cx = bin_axis_a(pi.x, sd.ihx, sd.nx), cy = bin_axis_a(pi.y, sd.ihy, sd.ny),
              cz = bin_axis_a(pi.z, sd.ihz, sd.nz);
    const double w = a.r_max / a.bins, reach = a.r_max + 2.0 * w;
    for_cells(sd, a.start, cx, cy, cz, a.reach, [&](uint32_t q) {
      if ((int)q == p) return;
      const float4 pj = a.pos[q];
      const double dx = min_image_d((double)pj.x - (double)pi.x, sd.Lx);
      const double dy = min_image_d((double)pj.y - (double)pi.y, sd.Ly);
      const double dz = min_image_d((double)pj.z - (double)pi.z, sd.Lz);
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r >= reach) return;
      const double u = r / w - 0.5;
      const int b0 = (int)floor(u) - 1;
      for (int b = max(b0, 0); b <= min(b0 + 3, a.bins - 1); ++b) h[b] += bspline3(u - b);
    });
  }
  for (int b = 0; b < a.bins; ++b) a.partial[(long)b * a.stride + k] = (k < sd.n) ? h[b] : 0.0;
}

// lx_i = dL/dx_i of the final state: (2/N) sum_j sum_b gh_b B3'((r - c_b)/w) / w (x_i - x_j)/r
__global__ void __launch_bounds__(128) k_hist_grad(HistArgs a) {
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  const float4 pi = a.pos[p];
  const int cx = bin_axis_a(pi.x, sd.ihx, sd.nx), cy = bin_axis_a(pi.y, sd.ihy, sd.ny),
            cz = bin_axis_a(pi.z, sd.ihz, sd.nz);
  const double w = a.r_max / a.bins, reach = a.r_max + 2.0 * w;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  for_cells(sd, a.start, cx, cy, cz, a.reach, [&](uint32_t q) {
    if ((int)q == p) return;
    const float4 pj = a.pos[q];
    const double dx = min_image_d((double)pj.x - (double)pi.x, sd.Lx);
    const double dy = min_image_d((double)pj.y - (double)pi.y, sd.Ly);
    const double dz = min_image_d((double)pj.z - (double)pi.z, sd.Lz);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r >= reach || r <= 0.0) return;
    const double u = r / w - 0.5;
    const int b0 = (int)floor(u) - 1;
    double s = 0.0;
    for (int b = max(b0, 0); b <= min(b0 + 3, a.bins - 1); ++b)
      s += a.gh[b] * bspline3_d(u - b);
    s *= 2.0 / ((double)sd.n * w * r);  // d r / d x_i = -(d)/r, d = x_j - x_i
    gx -= s * dx;
    gy -= s * dy;
    gz -= s * dz;
  });
  const uint32_t id = a.id[p];
  a.lam_x[id] = make_double4(gx, gy, gz, 0.0);
  a.lam_v[id] = make_double4(0.0, 0.0, 0.0, 0.0);
}

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

// counts c_i of the step (stored as the "previous count" of the post-step state), by index
__global__ void k_adj_counts(AdjArgs a) {
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  a.cnt_by_id[a.id_next[p]] = __float_as_int(a.vel_next[p].w);
}

// phi_i = dt m_i (lv'_i + dt lx'_i),  lv_i = fr (lv'_i + dt lx'_i)
__global__ void k_adj_phi(AdjArgs a) {
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  const uint32_t id = a.id[p];
  const int prev = __float_as_int(a.vel[p].w);
  const int cnt = a.cnt_by_id[id];
  // the forward's FP32 factor (ParticleSimulation.cu:136-143)
  const float avg = (float)(cnt + prev) * 0.5f;
  const float dens = fminf(avg / sd.max_exp, 1.0f);
  const float fac = 1.0f - sd.one_minus_balance * dens;
  const double m = (double)(sd.fm * fac);
  const double4 lx = a.lam_x_next[id], lv = a.lam_v_next[id];
  const double dt = (double)sd.dt, fr = (double)sd.friction;
  const double mx = lv.x + dt * lx.x, my = lv.y + dt * lx.y, mz = lv.z + dt * lx.z;
  a.phi[id] = make_double4(dt * m * mx, dt * m * my, dt * m * mz, 0.0);
  a.lam_v[id] = make_double4(fr * mx, fr * my, fr * mz, 0.0);
}

// lx_i = lx'_i + sum_j (J_ji phi_j - J_ij phi_i); parameter partials of particle i
__global__ void __launch_bounds__(128) k_adj_pair(AdjArgs a) {
  const SimDesc sd = a.sims[0];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  const float4 pi = a.pos[p];
  const uint32_t idi = a.id[p];
  const int ti = (int)__float_as_uint(pi.w);
  const int T = a.c.T;
  const int cx = bin_axis_a(pi.x, sd.ihx, sd.nx), cy = bin_axis_a(pi.y, sd.ihy, sd.ny),
            cz = bin_axis_a(pi.z, sd.ihz, sd.nz);
  const double4 phi_i = a.phi[idi];
  double lx = 0.0, ly = 0.0, lz = 0.0;
  double gfv[kMaxTypes];
#pragma unroll
  for (int t = 0; t < kMaxTypes; ++t) gfv[t] = 0.0;
  double grat = 0.0;
  const double R = a.c.radius, rho = a.c.ratio, rep = a.c.rep, att = a.c.att, kk = a.c.k;
  const double eps = (double)1e-4f;
  for_cells(sd, a.start, cx, cy, cz, sd.reach, [&](uint32_t q) {
    if ((int)q == p) return;
    const float4 pj = a.pos[q];
    const int tj = (int)__float_as_uint(pj.w);
    // the forward's FP32 decision (ParticleSimulation.cu:92-112 with the engine's op order)
    const float fx = min_image_f(pj.x - pi.x, sd.Lx), fy = min_image_f(pj.y - pi.y, sd.Ly),
                fz = min_image_f(pj.z - pi.z, sd.Lz);
    const float s2f = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
    if (!(s2f + 1e-4f < a.thr[ti * T + tj])) return;
    const double dx = min_image_d((double)pj.x - (double)pi.x, sd.Lx);
    const double dy = min_image_d((double)pj.y - (double)pi.y, sd.Ly);
    const double dz = min_image_d((double)pj.z - (double)pi.z, sd.Lz);
    const double dist = sqrt(dx * dx + dy * dy + dz * dz + eps);
    const double dR = 0.5 * R * (a.c.rbt[ti] + a.c.rbt[tj]);  // d effR / d ratio
    const double eff = R + dR * rho;
    const double r = dist / eff;
    const double ex = exp(-kk * r * r);
    const double g = rep * ex - att * r;
    const double gp = -2.0 * kk * r * rep * ex - att;  // g'(r)
    const double h = g / dist;
    const double hp = gp / (eff * dist) - g / (dist * dist);
    const double4 phi_j = a.phi[a.id[q]];
    const double fij = a.c.table[ti * T + tj], fji = a.c.table[tj * T + ti];
    // J phi = fv (h phi + hp d (d . phi) / dist)
    const double dpi = (dx * phi_i.x + dy * phi_i.y + dz * phi_i.z) / dist;
    const double dpj = (dx * phi_j.x + dy * phi_j.y + dz * phi_j.z) / dist;
    lx += fji * (h * phi_j.x + hp * dx * dpj) - fij * (h * phi_i.x + hp * dx * dpi);
    ly += fji * (h * phi_j.y + hp * dy * dpj) - fij * (h * phi_i.y + hp * dy * dpi);
    lz += fji * (h * phi_j.z + hp * dz * dpj) - fij * (h * phi_i.z + hp * dz * dpi);
    // parameters: phi_i . (g d / dist) and phi_i . fv (dg/deffR) deffR/dratio d / dist
    const double gd = g * dpi;
#pragma unroll
    for (int t = 0; t < kMaxTypes; ++t) gfv[t] += (t == tj) ? gd : 0.0;
    const double dg_deff = -gp * r / eff;
    grat += fij * dg_deff * dR * dpi;
  });
  const double4 lxn = a.lam_x_next[idi];
  a.lam_x[idi] = make_double4(lxn.x + lx, lxn.y + ly, lxn.z + lz, 0.0);
  // partials: row (ti * T + t) for the table, row T*T for the ratio; column = particle
#pragma unroll
  for (int t = 0; t < kMaxTypes; ++t)
    if (t < T)
      for (int a2 = 0; a2 < T; ++a2)
        a.partial[(long)(a2 * T + t) * a.stride + k] = (a2 == ti) ? gfv[t] : 0.0;
  a.partial[(long)(T * T) * a.stride + k] = grat;
}

hipError_t launch_hist(const HistArgs& a, int n, hipStream_t st) {
  const int g = (n + 127) / 128;
  hipLaunchKernelGGL(k_hist, dim3(g), dim3(128), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_hist_grad(const HistArgs& a, int n, hipStream_t st) {
  const int g = (n + 127) / 128;
  hipLaunchKernelGGL(k_hist_grad, dim3(g), dim3(128), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_reduce_rows(const double* in, long stride, int n, int rows, double scale,
                              double* out, int accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_rows, dim3(rows), dim3(256), 0, st, in, stride, n, scale, out,
                     accumulate);
  return hipGetLastError();
}

hipError_t launch_adj_step(const AdjArgs& a, int n, hipStream_t st) {
  const int g = (n + 127) / 128;
  hipLaunchKernelGGL(k_adj_counts, dim3(g), dim3(128), 0, st, a);
  hipLaunchKernelGGL(k_adj_phi, dim3(g), dim3(128), 0, st, a);
  hipLaunchKernelGGL(k_adj_pair, dim3(g), dim3(128), 0, st, a);
  return hipGetLastError();
}

}  // namespace cf
