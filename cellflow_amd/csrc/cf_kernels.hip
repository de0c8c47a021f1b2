// HIP kernels of the particle-life step engine, written for gfx950 (CDNA4, wave64).
//
// One step (Jacobi: every read comes from the pre-step state A, every write goes to
// temporaries and then to state B):
//   k_force         27-cell (or 125-cell) stencil over the cell-sorted state: cutoff
//                   predicate, neighbour count, smooth force, density factor, friction-
//                   Euler integration, periodic wrap (ParticleSimulation.cu:69-166); then
//                   the new cell of the particle and its arrival rank in that cell
//                   (atomic histogram).  Optionally also the proximity-graph candidate
//                   selection of the pre-step state (ParticleSimulation.cu:214-246).
//   k_scan          decoupled look-back exclusive scan of the cell histogram (and of the
//                   per-owner edge counts), wave64 prefix scans inside each tile.
//   k_fill          scatters each particle index into its new cell range (slot =
//                   start + arrival) and writes the canonical edge list.
//   k_rank_scatter  ranks each particle by index inside its new cell and moves its state
//                   there: the new order is (cell, index), independent of atomic order.
// The state lives in HBM sorted by (cell, index) as float4 SoA, so the stencil reads
// contiguous runs of slots: for each (dy, dz) row the x-neighbour cells are adjacent in
// memory.
//
// Arithmetic contract: compiled with -ffp-contract=off; the squared distance is
// fmaf(dz, dz, fmaf(dy, dy, dx*dx)) in FP32, the cutoff test is s + 1e-4f < thr with thr
// the smallest float whose sqrtf reaches the effective radius (identical to the
// reference's sqrtf(s + 0.0001f) < effR), and binning is floor(x * (n/L)) in FP32.  The
// oracle evaluates the same FP32 expressions, so counts, cells, order and edges match it
// bit for bit.

#include <hip/hip_ext.h>

#include <type_traits>

#include "cf_internal.h"

namespace cf {

namespace {
thread_local hipEvent_t t_timer_start = nullptr, t_timer_stop = nullptr;
}

void arm_kernel_timer(hipEvent_t start, hipEvent_t stop) {
  t_timer_start = start;
  t_timer_stop = stop;
}

// hipLaunchKernelGGL, or, when arm_kernel_timer armed this thread's timer, the same launch
// through hipExtLaunchKernelGGL with the armed events (then disarmed).
#define CF_LAUNCH(kernel, grid, block, shm, st, ...)                                   \
  do {                                                                                 \
    if (t_timer_start) {                                                               \
      hipEvent_t cf_s0_ = t_timer_start, cf_s1_ = t_timer_stop;                        \
      t_timer_start = t_timer_stop = nullptr;                                          \
      hipExtLaunchKernelGGL(kernel, grid, block, shm, st, cf_s0_, cf_s1_, 0u, __VA_ARGS__); \
    } else {                                                                           \
      hipLaunchKernelGGL(kernel, grid, block, shm, st, __VA_ARGS__);                   \
    }                                                                                  \
  } while (0)

// ------------------------------------------------------------------------------------
// small device helpers

__device__ __forceinline__ int bin_axis(float x, float ih, int n) {
  // floor(x * ih) clamped to [0, n - 1], NaN to 0 (fmaxf takes the number); n - 1 < 2^24 is
  // exact in float, and u in [n - 1, n) truncates to n - 1 either way
  return (int)fminf(fmaxf(x * ih, 0.0f), (float)(n - 1));
}

__device__ __forceinline__ float min_image(float d, float L) {
  // ParticleSimulation.cu:97-102
  if (d > L * 0.5f) d -= L;
  if (d < -L * 0.5f) d += L;
  return d;
}

// fmodf(x + L, L) of ParticleSimulation.cu:159-161; the two common cases are exact
// without the library call (y - L is exact for y in [L, 2L) by Sterbenz's lemma).
__device__ __forceinline__ float wrap_pos(float x, float L) {
  float y = x + L;
  if (y >= 0.0f && y < L) return y;
  if (y >= L && y < 2.0f * L) return y - L;
  return fmodf(y, L);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  This maps
// the hardware block index to a logical one so that each XCD gets one contiguous range of
// blocks (a bijection for any grid size): neighbouring particles, whose stencils read the
// same slots, then share an L2.
constexpr int kXcds = 8;

// The same order over the first nb blocks of a grid (b < nb).
__device__ __forceinline__ int xcd_map(int b, int nb) {
  const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
  return x * q + min(x, r) + b / kXcds;
}

__device__ __forceinline__ int xcd_block() { return xcd_map((int)blockIdx.x, (int)gridDim.x); }

// Orders a wave's LDS writes before later LDS reads by other lanes of the same wave (a
// wave's LDS operations complete in issue order).  The force kernels keep every LDS table
// per wave, so their waves need no workgroup barrier and never wait for each other.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t y = __shfl_up(v, d, kWave);
    if (lane >= d) v += y;
  }
  return v;
}

// ------------------------------------------------------------------------------------
// proximity-graph candidate list of one owner: the `cap` smallest qualifying indices,
// kept in a slot-major scratch array (coalesced across the wave).

struct EdgeAcc {
  int q;            // qualifying candidates seen (stored: min(q, cap))
  uint32_t maxid;   // largest stored index
  int maxk;         // its position
};

__device__ __forceinline__ void edge_push(EdgeAcc& e, float d2, uint32_t jid, int cap, float* cd2,
                                          uint32_t* cid, long P, int p) {
  if (e.q < cap) {
    cd2[(long)e.q * P + p] = d2;
    cid[(long)e.q * P + p] = jid;
    if (e.q == 0 || jid > e.maxid) {
      e.maxid = jid;
      e.maxk = e.q;
    }
    e.q++;
  } else {
    e.q++;
    if (jid < e.maxid) {
      cd2[(long)e.maxk * P + p] = d2;
      cid[(long)e.maxk * P + p] = jid;
      uint32_t m = 0;
      int mk = 0;
      for (int k = 0; k < cap; ++k) {
        uint32_t v = cid[(long)k * P + p];
        if (k == 0 || v > m) {
          m = v;
          mk = k;
        }
      }
      e.maxid = m;
      e.maxk = mk;
    }
  }
}

// Stable sort of the kept candidates by squared distance (ties by index, which is the
// order the reference's insertion sort leaves, ParticleSimulation.cu:235-243), truncated
// to maxc (:246).  Writes estage[o * es + rank] and returns the kept count.  o is the
// owner's slot, or its index when the tile sort (which writes edges in index order) follows.
__device__ __forceinline__ int edge_finish(const EdgeAcc& e, int cap, int maxc, const float* cd2,
                                           const uint32_t* cid, uint32_t* estage, long P, int p,
                                           uint32_t o, long es) {
  const int m = min(e.q, cap);
  for (int k = 0; k < m; ++k) {
    const float dk = cd2[(long)k * P + p];
    const uint32_t ik = cid[(long)k * P + p];
    int rank = 0;
    for (int k2 = 0; k2 < m; ++k2) {
      const float d2 = cd2[(long)k2 * P + p];
      const uint32_t i2 = cid[(long)k2 * P + p];
      rank += (d2 < dk || (d2 == dk && i2 < ik)) ? 1 : 0;
    }
    if (rank < maxc) estage[(long)o * es + rank] = ik;
  }
  return min(m, maxc);
}

// New cell of an integrated particle (x, y, z) whose pre-step cell had rows/planes (cyo,
// czo).  For the counting sort: its key and arrival rank.  For the tile sort: its key, with
// kFarBit (and a place in a far list) when the pre-step cell lies outside the window of
// the new cell's tile, and the per-cell and per-tile counts.  The window test is the one
// k_tile_sort enumerates.
// TILE selects the sort at compile time for the kernels the tuner picks (a runtime flag
// there costs the stencil loop registers); the others pass a.tile_sort through.
template <bool TILE>
__device__ __forceinline__ void commit_cell(const ForceArgs& a, const SimDesc& sd, int p, int cyo,
                                            int czo, float x, float y, float z) {
  const int xn = bin_axis(x, sd.ihx, sd.nx), yn = bin_axis(y, sd.ihy, sd.ny),
            zn = bin_axis(z, sd.ihz, sd.nz);
  const uint32_t c = (uint32_t)(sd.c0 + xn + sd.nx * (yn + sd.ny * zn));
  if constexpr (!TILE) {  // the counting sort ranks by arrival
    a.key[p] = c;
    a.arrival[p] = atomicAdd(&a.hist[c], 1u);
    return;
  }
  const int ty = yn / sd.tr, y0 = ty * sd.tr;
  const int nrw = min(min(sd.tr, sd.ny - y0) + 2, sd.ny), npl = min(3, sd.nz);
  // the old row and plane relative to the window's first, cyclically (each difference lies
  // within one period either way, so one correction replaces the modulo)
  int dyw = cyo - y0 + 1, dzw = czo - zn + 1;
  if (dyw < 0) dyw += sd.ny;
  else if (dyw >= sd.ny) dyw -= sd.ny;
  if (dzw < 0) dzw += sd.nz;
  else if (dzw >= sd.nz) dzw -= sd.nz;
  const bool near = dyw < nrw && dzw < npl;
  const uint32_t tile = (uint32_t)(sd.t0 + zn * sd.nyt + ty);
  uint32_t k = c;
  if (!near) {
    k |= kFarBit;
    const uint32_t j = atomicAdd(&a.ftn[tile * kCtrStride], 1u);
    if (j < (uint32_t)kFarCap) {
      a.ftl[(long)tile * kFarCap + j] = (uint32_t)p;
    } else {
      const uint32_t jo = atomicAdd(a.far_n, 1u);
      if (jo < (uint32_t)a.P) a.far_slot[jo] = (uint32_t)p;
    }
  }
  a.key[p] = k;
  // one tile-count atomic per distinct tile among the wave's particles (they are
  // consecutive slots, so nearly always one)
  const int lane = threadIdx.x & (kWave - 1);
  bool todo = true;
  while (true) {
    const unsigned long long m = __ballot(todo);
    if (m == 0ull) break;
    const int lead = __ffsll((long long)m) - 1;
    const uint32_t lt = (uint32_t)__shfl((int)tile, lead, kWave);
    const bool same = todo && tile == lt;
    const unsigned long long ms = __ballot(same);
    if (lane == lead) atomicAdd(&a.tcount[lt * kCtrStride], (uint32_t)__popcll(ms));
    if (same) todo = false;
  }
}

// One axis of the force stencil as up to three runs of cells: the run below cell 0
// (wrapped, image shift -L), the unwrapped run, and the run above cell n-1 (wrapped,
// shift +L).  Axes with fewer than 2r+1 cells are one run over every cell (the minimum
// image is then decided per pair).  Computed with selects so nothing lands in scratch.
__device__ __forceinline__ bool axis_run(int k, int c, int r, int n, float L, int& lo, int& hi,
                                         float& sh) {
  if (n < 2 * r + 1) {
    lo = 0;
    hi = n - 1;
    sh = 0.0f;
    return k == 1;
  }
  const int a = c - r, b = c + r;
  if (k == 0) {
    lo = a + n;
    hi = n - 1;
    sh = -L;
    return a < 0;
  }
  if (k == 2) {
    lo = 0;
    hi = b - n;
    sh = L;
    return b >= n;
  }
  lo = max(a, 0);
  hi = min(b, n - 1);
  sh = 0.0f;
  return true;
}

// ------------------------------------------------------------------------------------
// k_force: the fused force + integrate + rebin-key pass (and fused edge candidates).


template <bool EDGES>
__global__ void __launch_bounds__(256) k_force(ForceArgs a) {
  __shared__ PairConst tab[4][kMaxTypes * kMaxTypes];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = blockIdx.x * (blockDim.x >> 6) + wv;
  const bool wave_ok = gw < a.nwaves;
  const int s = __builtin_amdgcn_readfirstlane(wave_ok ? (int)a.wave_sim[gw] : 0);
  const SimDesc sd = a.sims[s];
  if (wave_ok)
    for (int t = lane; t < sd.T * sd.T; t += kWave) tab[wv][t] = a.pairs[sd.tab + t];
  wave_lds_sync();
  if (!wave_ok) return;
  const int p = gw * kWave + lane;
  if (p - sd.p0 >= sd.n) return;

  const float4 pi = a.pos[p];
  const float4 vi = a.vel[p];
  const uint32_t idi = a.id[p];
  const uint32_t ti = __float_as_uint(pi.w);
  const PairConst* row = &tab[wv][ti * sd.T];
  const int cx = bin_axis(pi.x, sd.ihx, sd.nx);
  const int cy = bin_axis(pi.y, sd.ihy, sd.ny);
  const int cz = bin_axis(pi.z, sd.ihz, sd.nz);

  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  int cnt = 0;
  const bool do_edges = EDGES && a.edges && sd.edge_on;
  EdgeAcc ea;
  ea.q = 0;
  ea.maxid = 0;
  ea.maxk = 0;
  const int r = sd.reach;
  const bool gen = sd.generic != 0;
  const int ylo = (sd.ny < 2 * r + 1) ? 0 : cy - r;
  const int yhi = (sd.ny < 2 * r + 1) ? sd.ny - 1 : cy + r;
  const int zlo = (sd.nz < 2 * r + 1) ? 0 : cz - r;
  const int zhi = (sd.nz < 2 * r + 1) ? sd.nz - 1 : cz + r;

  for (int zz = zlo; zz <= zhi; ++zz) {
    int z = zz;
    float sz = 0.0f;
    if (z < 0) { z += sd.nz; sz = -sd.Lz; }
    else if (z >= sd.nz) { z -= sd.nz; sz = sd.Lz; }
    for (int yy = ylo; yy <= yhi; ++yy) {
      int y = yy;
      float sy = 0.0f;
      if (y < 0) { y += sd.ny; sy = -sd.Ly; }
      else if (y >= sd.ny) { y -= sd.ny; sy = sd.Ly; }
      const int rowc = sd.c0 + sd.nx * (y + sd.ny * z);
      for (int run = 0; run < 3; ++run) {
        int xlo, xhi;
        float sx;
        if (!axis_run(run, cx, r, sd.nx, sd.Lx, xlo, xhi, sx)) continue;
        const int q0 = (int)a.start[rowc + xlo];
        const int q1 = (int)a.start[rowc + xhi + 1];
        const bool nowrap = (sx == 0.0f) && (sy == 0.0f) && (sz == 0.0f);
        for (int q = q0; q < q1; ++q) {
          const float4 pj = a.pos[q];
          float dx, dy, dz;
          bool raw;
          if (!gen) {
            dx = (pj.x - pi.x) + sx;
            dy = (pj.y - pi.y) + sy;
            dz = (pj.z - pi.z) + sz;
            raw = nowrap;
          } else {
            const float ux = pj.x - pi.x, uy = pj.y - pi.y, uz = pj.z - pi.z;
            dx = min_image(ux, sd.Lx);
            dy = min_image(uy, sd.Ly);
            dz = min_image(uz, sd.Lz);
            raw = (dx == ux) && (dy == uy) && (dz == uz);
          }
          const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
          const uint32_t tj = __float_as_uint(pj.w);
          const float sp = s2 + 1e-4f;
          const PairConst pc = row[tj];
          if (sp < pc.thr && q != p) {
            cnt++;
            const float inv = __builtin_amdgcn_rsqf(sp);
            const float dist = sp * inv;
            const float rr = dist * pc.inv_eff;
            const float ex = __builtin_amdgcn_exp2f(sd.kexp2 * (rr * rr));
            const float net = (sd.rep * ex - sd.att * rr) * pc.fv;
            const float w = net * inv;
            fx += dx * w;
            fy += dy * w;
            fz += dz * w;
          }
          if (EDGES) {
            if (do_edges && raw && tj == ti && s2 < sd.edge_d2 && q != p) {
              const uint32_t jid = a.id[q];
              if (jid > idi) edge_push(ea, s2, jid, a.edge_cap, a.ecand_d2, a.ecand_id, a.P, p);
            }
          }
        }
      }
    }
  }

  // ParticleSimulation.cu:136-161
  const int prev = __float_as_int(vi.w);
  const float avg = (float)(cnt + prev) * 0.5f;
  const float dens = fminf(avg / sd.max_exp, 1.0f);
  const float fac = 1.0f - sd.one_minus_balance * dens;
  const float m = sd.fm * fac;
  fx *= m;
  fy *= m;
  fz *= m;
  const float vx = vi.x * sd.friction + fx * sd.dt;
  const float vy = vi.y * sd.friction + fy * sd.dt;
  const float vz = vi.z * sd.friction + fz * sd.dt;
  const float x = wrap_pos(pi.x + vx * sd.dt, sd.Lx);
  const float y = wrap_pos(pi.y + vy * sd.dt, sd.Ly);
  const float z = wrap_pos(pi.z + vz * sd.dt, sd.Lz);
  const float fin = fabsf(x) + fabsf(y) + fabsf(z) + fabsf(vx) + fabsf(vy) + fabsf(vz);
  if (!(fin <= 3.402823466e38f)) atomicOr(a.flag, 1u);

  a.tpos[p] = make_float4(x, y, z, pi.w);
  a.tvel[p] = make_float4(vx, vy, vz, __int_as_float(cnt));
  if (a.acc) a.acc[idi] = make_float4(fx, fy, fz, 0.0f);
  if (a.tile_sort) commit_cell<true>(a, sd, p, cy, cz, x, y, z);
  else commit_cell<false>(a, sd, p, cy, cz, x, y, z);

  if (EDGES) {
    if (do_edges) {
      const int e = edge_finish(ea, a.edge_cap, sd.edge_maxc, a.ecand_d2, a.ecand_id, a.estage,
                                a.P, p, a.tile_sort ? idi : (uint32_t)p, a.estride);
      if (e != 0) a.ecount[idi] = (uint32_t)e;  // zero already (the sorts clear it)
    }
  }
}

// ------------------------------------------------------------------------------------
// k_force_lds: the same pass with the stencil staged in LDS.  One wave per workgroup.
//
// The wave's 64 owners are consecutive slots, so their cells are consecutive in (z, y, x)
// order and, for each stencil row offset (dy, dz), their runs of neighbour slots fall in
// consecutive neighbour rows.  Per row offset the wave reduces its lanes' runs to one
// slot interval [A, B) (one interval per y/z wrap pattern present), copies the intervals
// into a 1024-candidate LDS tile with direct global->LDS loads (all in flight at once,
// one wait), and each lane then walks only its own runs inside the tile.  Intervals
// longer than the tile are split over several batches.

constexpr int kLdsCand = 1024;   // float4 candidates per LDS batch (16 KB)

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, kWave));
  return v;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, kWave));
  return v;
}

// Neighbour row of stencil row index k for a lane in cell (cy, cz): absolute rows when the
// axis has fewer than 2r+1 cells (minimum image per pair), else offset rows with the
// periodic image shift.  wrap = 3*ywrap + zwrap (0 none, 1 below, 2 above).
__device__ __forceinline__ void stencil_row(const SimDesc& sd, int k, int nyo, int cy, int cz,
                                            int& rowc, float& sy, float& sz, int& wrap) {
  const int r = sd.reach;
  const int iy = k % nyo, iz = k / nyo;
  int y, z, yw = 0, zw = 0;
  sy = 0.0f;
  sz = 0.0f;
  if (sd.ny < 2 * r + 1) {
    y = iy;
  } else {
    y = cy + iy - r;
    if (y < 0) { y += sd.ny; sy = -sd.Ly; yw = 1; }
    else if (y >= sd.ny) { y -= sd.ny; sy = sd.Ly; yw = 2; }
  }
  if (sd.nz < 2 * r + 1) {
    z = iz;
  } else {
    z = cz + iz - r;
    if (z < 0) { z += sd.nz; sz = -sd.Lz; zw = 1; }
    else if (z >= sd.nz) { z -= sd.nz; sz = sd.Lz; zw = 2; }
  }
  rowc = sd.c0 + sd.nx * (y + sd.ny * z);
  wrap = 3 * yw + zw;
}

// x sub-runs of a lane: [xa0, xb0] unshifted and [xa1, xb1] with shift sx1 (empty if
// xa1 > xb1).
__device__ __forceinline__ void stencil_x(const SimDesc& sd, int cx, int& xa0, int& xb0, int& xa1,
                                          int& xb1, float& sx1) {
  const int r = sd.reach;
  sx1 = 0.0f;
  xa1 = 1;
  xb1 = 0;
  if (sd.nx < 2 * r + 1) {
    xa0 = 0;
    xb0 = sd.nx - 1;
    return;
  }
  const int lo = cx - r, hi = cx + r;
  if (lo < 0) {
    xa0 = 0; xb0 = hi; xa1 = lo + sd.nx; xb1 = sd.nx - 1; sx1 = -sd.Lx;
  } else if (hi >= sd.nx) {
    xa0 = lo; xb0 = sd.nx - 1; xa1 = 0; xb1 = hi - sd.nx; sx1 = sd.Lx;
  } else {
    xa0 = lo; xb0 = hi;
  }
}

struct PairAcc {
  float fx, fy, fz;
  int cnt;
  EdgeAcc ea;
};

template <bool EDGES>
__device__ __forceinline__ void visit_tile(PairAcc& acc, const ForceArgs& a, const SimDesc& sd,
                                           const float4* __restrict__ tile, const PairConst* row,
                                           const float4& pi, uint32_t ti, uint32_t idi, int p,
                                           uint32_t lo, uint32_t hi, uint32_t base, int off,
                                           float sx, float sy, float sz, bool gen, bool do_edges) {
  const bool nowrap = (sx == 0.0f) && (sy == 0.0f) && (sz == 0.0f);
  for (uint32_t q = lo; q < hi; ++q) {
    const float4 pj = tile[off + (int)(q - base)];
    float dx, dy, dz;
    bool raw;
    if (!gen) {
      dx = (pj.x - pi.x) + sx;
      dy = (pj.y - pi.y) + sy;
      dz = (pj.z - pi.z) + sz;
      raw = nowrap;
    } else {
      const float ux = pj.x - pi.x, uy = pj.y - pi.y, uz = pj.z - pi.z;
      dx = min_image(ux, sd.Lx);
      dy = min_image(uy, sd.Ly);
      dz = min_image(uz, sd.Lz);
      raw = (dx == ux) && (dy == uy) && (dz == uz);
    }
    const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const uint32_t tj = __float_as_uint(pj.w);
    const float sp = s2 + 1e-4f;
    const PairConst pc = row[tj];
    if (sp < pc.thr && (int)q != p) {
      acc.cnt++;
      const float inv = __builtin_amdgcn_rsqf(sp);
      const float dist = sp * inv;
      const float rr = dist * pc.inv_eff;
      const float ex = __builtin_amdgcn_exp2f(sd.kexp2 * (rr * rr));
      const float net = (sd.rep * ex - sd.att * rr) * pc.fv;
      const float w = net * inv;
      acc.fx += dx * w;
      acc.fy += dy * w;
      acc.fz += dz * w;
    }
    if (EDGES) {
      if (do_edges && raw && tj == ti && s2 < sd.edge_d2 && (int)q != p) {
        const uint32_t jid = a.id[q];
        if (jid > idi) edge_push(acc.ea, s2, jid, a.edge_cap, a.ecand_d2, a.ecand_id, a.P, p);
      }
    }
  }
}

template <bool EDGES, int R>
__global__ void __launch_bounds__(64) k_force_lds(ForceArgs a) {
  constexpr int kRows = (2 * R + 1) * (2 * R + 1);
  __shared__ float4 tile[kLdsCand];
  __shared__ uint4 lruns[kRows][kWave];
  __shared__ PairConst tab[kMaxTypes * kMaxTypes];
  const int lane = threadIdx.x;
  const int gw = blockIdx.x;
  if (gw >= a.nwaves) return;  // whole workgroup
  const int s = __builtin_amdgcn_readfirstlane((int)a.wave_sim[gw]);
  const SimDesc sd = a.sims[s];
  for (int t = lane; t < sd.T * sd.T; t += kWave) tab[t] = a.pairs[sd.tab + t];
  const int p = gw * kWave + lane;
  const bool active = (p - sd.p0) < sd.n;
  float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = pi;
  uint32_t idi = 0;
  if (active) {
    pi = a.pos[p];
    vi = a.vel[p];
    idi = a.id[p];
  }
  const uint32_t ti = active ? __float_as_uint(pi.w) : 0u;
  const int cx = bin_axis(pi.x, sd.ihx, sd.nx);
  const int cy = bin_axis(pi.y, sd.ihy, sd.ny);
  const int cz = bin_axis(pi.z, sd.ihz, sd.nz);
  const int nyo = (sd.ny < 2 * R + 1) ? sd.ny : 2 * R + 1;
  const int nzo = (sd.nz < 2 * R + 1) ? sd.nz : 2 * R + 1;
  const int nrows = nyo * nzo;
  int xa0, xb0, xa1, xb1;
  float sx1;
  stencil_x(sd, cx, xa0, xb0, xa1, xb1, sx1);
  const bool has1 = xa1 <= xb1;

  // phase A: this lane's slot runs for every stencil row (independent loads)
#pragma unroll
  for (int k = 0; k < kRows; ++k) {
    if (k < nrows) {
      int rowc, wrap;
      float sy, sz;
      stencil_row(sd, k, nyo, cy, cz, rowc, sy, sz, wrap);
      uint4 rr = make_uint4(0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u);
      if (active) {
        rr.x = a.start[rowc + xa0];
        rr.y = a.start[rowc + xb0 + 1];
        if (has1) {
          rr.z = a.start[rowc + xa1];
          rr.w = a.start[rowc + xb1 + 1];
        }
      }
      lruns[k][lane] = rr;
    }
  }
  __syncthreads();

  const PairConst* row = &tab[ti * sd.T];
  const bool gen = sd.generic != 0;
  PairAcc acc;
  acc.fx = acc.fy = acc.fz = 0.0f;
  acc.cnt = 0;
  acc.ea.q = 0;
  acc.ea.maxid = 0;
  acc.ea.maxk = 0;
  const bool do_edges = EDGES && a.edges && sd.edge_on;

  // phase B: batches of staged intervals; lane j of the wave describes piece j
  int used = 0, npieces = 0;
  int my_k = 0, my_g = 0, my_off = 0;
  uint32_t my_a = 0, my_b = 0;
  auto flush = [&]() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int j = 0; j < npieces; ++j) {
      const int k = __builtin_amdgcn_readlane(my_k, j);
      const int g = __builtin_amdgcn_readlane(my_g, j);
      const uint32_t pa = (uint32_t)__builtin_amdgcn_readlane((int)my_a, j);
      const uint32_t pb = (uint32_t)__builtin_amdgcn_readlane((int)my_b, j);
      const int off = __builtin_amdgcn_readlane(my_off, j);
      int rowc, wrap;
      float sy, sz;
      stencil_row(sd, k, nyo, cy, cz, rowc, sy, sz, wrap);
      if (active && wrap == g) {
        const uint4 rr = lruns[k][lane];
        visit_tile<EDGES>(acc, a, sd, tile, row, pi, ti, idi, p, max(rr.x, pa), min(rr.y, pb), pa,
                          off, 0.0f, sy, sz, gen, do_edges);
        visit_tile<EDGES>(acc, a, sd, tile, row, pi, ti, idi, p, max(rr.z, pa), min(rr.w, pb), pa,
                          off, sx1, sy, sz, gen, do_edges);
      }
    }
    __syncthreads();
    used = 0;
    npieces = 0;
  };

  for (int k = 0; k < nrows; ++k) {
    int rowc, gk;
    float sy, sz;
    stencil_row(sd, k, nyo, cy, cz, rowc, sy, sz, gk);
    const uint4 rr = lruns[k][lane];
    unsigned long long pend = __ballot(active);
    while (pend) {
      const int lead = __ffsll((long long)pend) - 1;
      const int g = __builtin_amdgcn_readlane(gk, lead);
      const bool mine = active && gk == g;
      pend &= ~__ballot(mine);
      uint32_t A = wave_min_u32(mine ? min(rr.x, rr.z) : 0xFFFFFFFFu);
      const uint32_t B = wave_max_u32(mine ? max(rr.y, rr.w) : 0u);
      A = __builtin_amdgcn_readfirstlane(A);
      while (A < B) {
        if (used == kLdsCand || npieces == kWave) flush();
        const int take = min((int)(B - A), kLdsCand - used);
        for (int j0 = 0; j0 < take; j0 += kWave) {
          if (j0 + lane < take)
            __builtin_amdgcn_global_load_lds((const void*)(a.pos + A + j0 + lane),
                                             (__attribute__((address_space(3))) void*)&tile[used + j0],
                                             16, 0, 0);
        }
        if (lane == npieces) {
          my_k = k;
          my_g = g;
          my_a = A;
          my_b = A + (uint32_t)take;
          my_off = used;
        }
        npieces++;
        used += take;
        A += (uint32_t)take;
      }
    }
  }
  if (npieces) flush();

  if (!active) return;
  // ParticleSimulation.cu:136-161
  const int prev = __float_as_int(vi.w);
  const float avg = (float)(acc.cnt + prev) * 0.5f;
  const float dens = fminf(avg / sd.max_exp, 1.0f);
  const float fac = 1.0f - sd.one_minus_balance * dens;
  const float m = sd.fm * fac;
  const float fx = acc.fx * m, fy = acc.fy * m, fz = acc.fz * m;
  const float vx = vi.x * sd.friction + fx * sd.dt;
  const float vy = vi.y * sd.friction + fy * sd.dt;
  const float vz = vi.z * sd.friction + fz * sd.dt;
  const float x = wrap_pos(pi.x + vx * sd.dt, sd.Lx);
  const float y = wrap_pos(pi.y + vy * sd.dt, sd.Ly);
  const float z = wrap_pos(pi.z + vz * sd.dt, sd.Lz);
  const float fin = fabsf(x) + fabsf(y) + fabsf(z) + fabsf(vx) + fabsf(vy) + fabsf(vz);
  if (!(fin <= 3.402823466e38f)) atomicOr(a.flag, 1u);
  a.tpos[p] = make_float4(x, y, z, pi.w);
  a.tvel[p] = make_float4(vx, vy, vz, __int_as_float(acc.cnt));
  if (a.acc) a.acc[idi] = make_float4(fx, fy, fz, 0.0f);
  if (a.tile_sort) commit_cell<true>(a, sd, p, cy, cz, x, y, z);
  else commit_cell<false>(a, sd, p, cy, cz, x, y, z);
  if (EDGES) {
    if (do_edges) {
      const int e = edge_finish(acc.ea, a.edge_cap, sd.edge_maxc, a.ecand_d2, a.ecand_id,
                                a.estage, a.P, p, a.tile_sort ? idi : (uint32_t)p, a.estride);
      if (e != 0) a.ecount[idi] = (uint32_t)e;  // zero already (the sorts clear it)
    }
  }
}

// ------------------------------------------------------------------------------------
// k_force_grp: G lanes per particle.
//
// Setup: the particle's lanes build its run table in LDS: one or two slot runs per
// stencil row (the unshifted x-range and the periodic-wrap x-range, with the image shift
// packed into the run's length word).  Rows and x-cells that cannot hold a partner are
// left out: a cell is kept only if the particle's distance to it, less a margin of 1e-3
// cells, is below the largest cutoff of the particle's type (with a relative margin of
// 1e-3), so no pair that the FP32 predicate accepts is ever dropped.
// Main loop: the G lanes split the flattened run list cyclically (lane g takes candidates
// g, g+G, ...), so each lane walks about 1/G of the candidates and neighbouring lanes
// read neighbouring slots.  Partial forces and counts are reduced across the group; the
// particle itself is visited once (its own cell is always kept) and removed from the
// count afterwards instead of being tested per candidate.
// Edges: candidates of all G lanes are appended to a per-particle LDS list; the group
// selects the 2*maxc smallest indices and ranks them by (distance, index).  A particle
// with more than kEdgeList candidates falls back to the exact streaming selection.

constexpr int kEdgeList = 64;
constexpr uint32_t kLenMask = (1u << 26) - 1u;

template <int G>
__device__ __forceinline__ float group_sum_f(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

template <int G>
__device__ __forceinline__ int group_sum_i(int v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// gap (in cells) from fractional position f in [0, 1) to the cell at offset d, minus the
// safety margin
__device__ __forceinline__ float cell_gap(int d, float f) {
  const float g = (d > 0) ? ((float)d - f) : ((d < 0) ? ((float)(-d - 1) + f) : 0.0f);
  return fmaxf(g - 1e-3f, 0.0f);
}

// Image shift of one axis from its 2-bit code (a signed count of periods: 1 = +L, 3 = -L,
// 0 = none): exactly 0 or +-L.
__device__ __forceinline__ float shift_of(uint32_t c, float L) {
  return (float)(((int)(c << 30)) >> 30) * L;
}

// Per-particle context of the stencil walk.
struct Own {
  float4 pi;
  uint32_t ti;
  int cx, cy, cz;
  float fxr, fyr, fzr;  // fractional position inside the cell, per axis
  float cut2;           // pruning radius^2 of the particle's type
  bool gen;             // grid too small for image-shifted runs
  int nyo, nzo;         // stencil rows per axis
  // cell_gap(-k, f) * h and cell_gap(k, f) * h per axis, k = 1..2 (grids with image-shifted
  // runs only): the gaps to the k-th cells below and above
  float gxl[2], gxr[2], gyl[2], gyr[2], gzl[2], gzr[2];
};

// The gaps of Own for reach R (the same float operations as cell_gap).
template <int R>
__device__ __forceinline__ void own_gaps(Own& c, const SimDesc& sd) {
#pragma unroll
  for (int k = 1; k <= R; ++k) {
    c.gxl[k - 1] = cell_gap(-k, c.fxr) * sd.hx;
    c.gxr[k - 1] = cell_gap(k, c.fxr) * sd.hx;
    c.gyl[k - 1] = cell_gap(-k, c.fyr) * sd.hy;
    c.gyr[k - 1] = cell_gap(k, c.fyr) * sd.hy;
    c.gzl[k - 1] = cell_gap(-k, c.fzr) * sd.hz;
    c.gzr[k - 1] = cell_gap(k, c.fzr) * sd.hz;
  }
}

// cell_gap(d, f) * h for |d| <= R from the precomputed gaps (d = 0: 0).
template <int R>
__device__ __forceinline__ float gap_of(int d, const float* lo, const float* hi) {
  float g = 0.0f;
#pragma unroll
  for (int k = 1; k <= R; ++k) {
    g = (d == -k) ? lo[k - 1] : g;
    g = (d == k) ? hi[k - 1] : g;
  }
  return g;
}

// The up to two runs of stencil row r (x-neighbour cells are adjacent slots): the
// unwrapped part and the part wrapped across the x boundary, each with its image-shift
// code (x: bits 0-1, y: 2-3, z: 4-5; 1 = +L, 3 = -L).  Rows and x ends farther than the
// type's reach (with margins) are dropped.  Length 0 means no run.
template <int R>
__device__ __forceinline__ void row_runs_gen(const ForceArgs& a, const SimDesc& sd, const Own& c,
                                         int r, uint32_t& s0, uint32_t& l0, uint32_t& k0,
                                         uint32_t& s1, uint32_t& l1, uint32_t& k1) {
  s0 = l0 = k0 = s1 = l1 = k1 = 0u;
  int iy, iz;
  if (c.gen) {  // uniform: the simulation's grid decides
    iy = r % c.nyo;
    iz = r / c.nyo;
  } else {  // nyo == 2R+1: division by a constant
    iy = r % (2 * R + 1);
    iz = r / (2 * R + 1);
  }
  int y, z, dy = 0, dz = 0;
  uint32_t cyc = 0u, czc = 0u;
  if (c.gen && sd.ny < 2 * R + 1) {
    y = iy;
  } else {
    dy = iy - R;
    y = c.cy + dy;
    if (y < 0) { y += sd.ny; cyc = 3u; }
    else if (y >= sd.ny) { y -= sd.ny; cyc = 1u; }
  }
  if (c.gen && sd.nz < 2 * R + 1) {
    z = iz;
  } else {
    dz = iz - R;
    z = c.cz + dz;
    if (z < 0) { z += sd.nz; czc = 3u; }
    else if (z >= sd.nz) { z -= sd.nz; czc = 1u; }
  }
  float gyz = 0.0f;
  if (!c.gen) {
    const float gy = cell_gap(dy, c.fyr) * sd.hy, gz = cell_gap(dz, c.fzr) * sd.hz;
    gyz = fmaf(gy, gy, gz * gz);
    if (!(gyz < c.cut2)) return;
  }
  const int rowc = sd.c0 + sd.nx * (y + sd.ny * z);
  int xlo, xhi;
  if (c.gen && sd.nx < 2 * R + 1) {
    xlo = 0;
    xhi = sd.nx - 1;
  } else {
    xlo = c.cx - R;
    xhi = c.cx + R;
    if (!c.gen) {
      while (xlo < c.cx) {
        const float gx = cell_gap(xlo - c.cx, c.fxr) * sd.hx;
        if (fmaf(gx, gx, gyz) < c.cut2) break;
        ++xlo;
      }
      while (xhi > c.cx) {
        const float gx = cell_gap(xhi - c.cx, c.fxr) * sd.hx;
        if (fmaf(gx, gx, gyz) < c.cut2) break;
        --xhi;
      }
    }
  }
  const uint32_t yzc = (cyc << 2) | (czc << 4);
  int a0 = xlo, b0 = xhi, a1 = 1, b1 = 0;
  uint32_t xc1 = 0u;
  if (xlo < 0) { a0 = 0; a1 = xlo + sd.nx; b1 = sd.nx - 1; xc1 = 3u; }
  else if (xhi >= sd.nx) { b0 = sd.nx - 1; a1 = 0; b1 = xhi - sd.nx; xc1 = 1u; }
  s0 = a.start[rowc + a0];
  l0 = a.start[rowc + b0 + 1] - s0;
  k0 = yzc;
  if (a1 <= b1) {
    s1 = a.start[rowc + a1];
    l1 = a.start[rowc + b1 + 1] - s1;
    k1 = yzc | xc1;
  }
}

// row_runs on a grid with at least 2R+1 cells per axis, straight-line: the row's y and z
// gaps from Own's precomputed ones, and the x ends by counting the x offsets within reach
// (the gap grows with the offset, so those offsets are 1..K on each side and the count is
// where cell-by-cell trimming from the far end stops).  Same runs as row_runs_gen.
template <int R>
__device__ __forceinline__ void row_runs_fast(const ForceArgs& a, const SimDesc& sd,
                                              const Own& c, int r, uint32_t& s0, uint32_t& l0,
                                              uint32_t& k0, uint32_t& s1, uint32_t& l1,
                                              uint32_t& k1) {
  s0 = l0 = k0 = s1 = l1 = k1 = 0u;
  constexpr int D = 2 * R + 1;
  const int dy = r % D - R, dz = r / D - R;
  const float gy = gap_of<R>(dy, c.gyl, c.gyr), gz = gap_of<R>(dz, c.gzl, c.gzr);
  const float gyz = fmaf(gy, gy, gz * gz);
  if (!(gyz < c.cut2)) return;
  int y = c.cy + dy, z = c.cz + dz;
  uint32_t cyc = 0u, czc = 0u;
  if (y < 0) { y += sd.ny; cyc = 3u; }
  else if (y >= sd.ny) { y -= sd.ny; cyc = 1u; }
  if (z < 0) { z += sd.nz; czc = 3u; }
  else if (z >= sd.nz) { z -= sd.nz; czc = 1u; }
  const int rowc = sd.c0 + sd.nx * (y + sd.ny * z);
  int xlo = c.cx, xhi = c.cx;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    xlo -= (fmaf(c.gxl[k], c.gxl[k], gyz) < c.cut2) ? 1 : 0;
    xhi += (fmaf(c.gxr[k], c.gxr[k], gyz) < c.cut2) ? 1 : 0;
  }
  const uint32_t yzc = (cyc << 2) | (czc << 4);
  int a0 = xlo, b0 = xhi, a1 = 1, b1 = 0;
  uint32_t xc1 = 0u;
  if (xlo < 0) { a0 = 0; a1 = xlo + sd.nx; b1 = sd.nx - 1; xc1 = 3u; }
  else if (xhi >= sd.nx) { b0 = sd.nx - 1; a1 = 0; b1 = xhi - sd.nx; xc1 = 1u; }
  s0 = a.start[rowc + a0];
  l0 = a.start[rowc + b0 + 1] - s0;
  k0 = yzc;
  if (a1 <= b1) {
    s1 = a.start[rowc + a1];
    l1 = a.start[rowc + b1 + 1] - s1;
    k1 = yzc | xc1;
  }
}

template <int R>
__device__ __forceinline__ void row_runs(const ForceArgs& a, const SimDesc& sd, const Own& c,
                                         int r, uint32_t& s0, uint32_t& l0, uint32_t& k0,
                                         uint32_t& s1, uint32_t& l1, uint32_t& k1) {
  if (c.gen) row_runs_gen<R>(a, sd, c, r, s0, l0, k0, s1, l1, k1);
  else row_runs_fast<R>(a, sd, c, r, s0, l0, k0, s1, l1, k1);
}

// A type pair's cutoff and force constants folded for the pair loop: thr, rep * fv,
// att * fv * inv_eff and kexp2 * inv_eff^2.
struct PairFold {
  float thr, rep_fv, att_fv_ie, k_ie2;
};

__device__ __forceinline__ PairFold fold_pair(const PairConst& pc, const SimDesc& sd) {
  return PairFold{pc.thr, sd.rep * pc.fv, sd.att * pc.fv * pc.inv_eff,
                  sd.kexp2 * pc.inv_eff * pc.inv_eff};
}

// ParticleSimulation.cu:104-131 for a pair in range, sp = d^2 + 1e-4 (r = sqrt(sp),
// rr = r / effR): (rep e^(-k rr^2) - att rr) fv / r, the factor of the offset (dx, dy, dz).
__device__ __forceinline__ float pair_weight(const PairFold& pc, float sp) {
  const float inv = __builtin_amdgcn_rsqf(sp);
  const float ex = __builtin_amdgcn_exp2f(pc.k_ie2 * sp);
  return fmaf(pc.rep_fv, ex, -(pc.att_fv_ie * (sp * inv))) * inv;
}

template <bool EDGES>
struct GrpAcc {
  float fx, fy, fz;
  int cnt;
};

// HITS (the gradient rollout's recompute, no edges): also list each particle's in-range
// partners in slot order (ForceArgs::hits), the lists the adjoint's pair kernel reads.
template <bool EDGES, int R, int G, bool TILE, bool HITS = false>
__global__ void __launch_bounds__(256) k_force_grp(ForceArgs a) {
  constexpr int kRows = (2 * R + 1) * (2 * R + 1);
  constexpr int kRuns = 2 * kRows;
  constexpr int kOwn = kWave / G;  // particles per wave
  __shared__ PairFold tab[4][kMaxTypes * kMaxTypes];
  __shared__ uint2 runs[4][kOwn][kRuns];  // (start slot, length | shift code << 26)
  __shared__ float2 el_d[EDGES ? 4 : 1][EDGES ? kOwn : 1][EDGES ? kEdgeList : 1];
  __shared__ uint32_t el_i[EDGES ? 4 : 1][EDGES ? kOwn : 1][EDGES ? kEdgeList : 1];
  __shared__ uint32_t el_n[4][kOwn];
  __shared__ uint32_t hl[HITS ? 4 : 1][HITS ? kOwn : 1][HITS ? kPairList : 1];
  __shared__ uint32_t hn[HITS ? 4 : 1][HITS ? kOwn : 1];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = xcd_block() * (blockDim.x >> 6) + wv;  // wave index in units of kOwn slots
  const long base = (long)gw * kOwn;
  const bool wave_ok = base < (long)a.nwaves * kWave;
  const int s = __builtin_amdgcn_readfirstlane(wave_ok ? (int)a.wave_sim[base / kWave] : 0);
  const SimDesc sd = a.sims[s];
  if (wave_ok)
    for (int t = lane; t < sd.T * sd.T; t += kWave) tab[wv][t] = fold_pair(a.pairs[sd.tab + t], sd);
  const int o = lane / G, g = lane % G;
  const int p = (int)base + o;
  const bool active = wave_ok && (p - sd.p0) < sd.n;
  float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vi = pi;
  uint32_t idi = 0;
  if (active) {
    pi = a.pos[p];
    vi = a.vel[p];
    idi = a.id[p];
  }
  const uint32_t ti = active ? __float_as_uint(pi.w) : 0u;
  const int cx = bin_axis(pi.x, sd.ihx, sd.nx);
  const int cy = bin_axis(pi.y, sd.ihy, sd.ny);
  const int cz = bin_axis(pi.z, sd.ihz, sd.nz);
  const bool gen = sd.generic != 0;
  const int nyo = gen && sd.ny < 2 * R + 1 ? sd.ny : 2 * R + 1;
  const int nzo = gen && sd.nz < 2 * R + 1 ? sd.nz : 2 * R + 1;
  const int nrows = nyo * nzo;
  if (EDGES && g == 0) el_n[wv][o] = 0u;
  if (HITS && g == 0) hn[wv][o] = 0u;
  wave_lds_sync();  // pair table ready

  const PairFold* row = &tab[wv][ti * sd.T];
  // ---- run table: row_runs of each stencil row, as (start slot, length | code << 26) ----
  {
    Own c;
    c.pi = pi;
    c.ti = ti;
    c.cx = cx;
    c.cy = cy;
    c.cz = cz;
    c.gen = gen;
    c.nyo = nyo;
    c.nzo = nzo;
    // largest squared reach of this type, with margins
    c.cut2 = active ? a.pairs[sd.tab + ti * sd.T].pad : 0.0f;
    c.fxr = fminf(fmaxf(pi.x * sd.ihx - (float)cx, 0.0f), 1.0f);
    c.fyr = fminf(fmaxf(pi.y * sd.ihy - (float)cy, 0.0f), 1.0f);
    c.fzr = fminf(fmaxf(pi.z * sd.ihz - (float)cz, 0.0f), 1.0f);
    own_gaps<R>(c, sd);
    for (int r = g; r < kRows; r += G) {
      uint32_t s0 = 0u, l0 = 0u, k0 = 0u, s1 = 0u, l1 = 0u, k1 = 0u;
      if (active && r < nrows) row_runs<R>(a, sd, c, r, s0, l0, k0, s1, l1, k1);
      runs[wv][o][2 * r] = make_uint2(s0, l0 | (k0 << 26));
      runs[wv][o][2 * r + 1] = make_uint2(s1, l1 | (k1 << 26));
    }
  }
  wave_lds_sync();

  // ---- candidates -----------------------------------------------------------------
  const bool do_edges = EDGES && a.edges && sd.edge_on && active;
  const bool ein = do_edges && ((sd.edge_inside >> ti) & 1);   // edge test implies in-range
  const bool eout = do_edges && !ein;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  int cnt = 0;
  uint32_t cum = 0;
  const int nruns = active ? 2 * nrows : 0;
  for (int run = 0; run < nruns; ++run) {
    const uint2 e = runs[wv][o][run];
    const uint32_t len = e.y & kLenMask;
    if (len == 0u) continue;
    const uint32_t code = e.y >> 26;
    const uint32_t q0 = e.x;
    // this lane's first index inside the run: flattened index == g (mod G)
    uint32_t t = (uint32_t)((g - (int)(cum & (G - 1))) & (G - 1));
    cum += len;
    if (code == 0u && !gen) {
      // unshifted run: raw (non-periodic) differences
      for (; t < len; t += G) {
        const uint32_t q = q0 + t;
        const float4 pj = a.pos[q];
        const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const uint32_t tj = __float_as_uint(pj.w);
        const PairFold pc = row[tj];
        const float sp = s2 + 1e-4f;
        const bool in = sp < pc.thr;
        if (HITS && in && q != (uint32_t)p) {
          const uint32_t at = atomicAdd(&hn[wv][o], 1u);
          if (at < (uint32_t)kPairList) hl[wv][o][at] = q;
        }
        if (in) {
          cnt++;
          const float w = pair_weight(pc, sp);
          fx = fmaf(dx, w, fx);
          fy = fmaf(dy, w, fy);
          fz = fmaf(dz, w, fz);
        }
        if (EDGES) {
          if ((in ? ein : eout) && tj == ti && s2 < sd.edge_d2) {
            const uint32_t jid = a.id[q];
            if (jid > idi) {
              const uint32_t kk = atomicAdd(&el_n[wv][o], 1u);
              if (kk < (uint32_t)kEdgeList) {
                el_d[wv][o][kk] = make_float2(s2, 0.0f);
                el_i[wv][o][kk] = jid;
              }
            }
          }
        }
      }
    } else {
      // image-shifted run, or per-pair minimum image on a grid too small for shifts
      const float sx = shift_of(code & 3u, sd.Lx), sy = shift_of((code >> 2) & 3u, sd.Ly),
                  sz = shift_of((code >> 4) & 3u, sd.Lz);
      for (; t < len; t += G) {
        const uint32_t q = q0 + t;
        const float4 pj = a.pos[q];
        float dx, dy, dz;
        bool raw;
        if (!gen) {
          dx = (pj.x - pi.x) + sx;
          dy = (pj.y - pi.y) + sy;
          dz = (pj.z - pi.z) + sz;
          raw = false;
        } else {
          const float ux = pj.x - pi.x, uy = pj.y - pi.y, uz = pj.z - pi.z;
          dx = min_image(ux, sd.Lx);
          dy = min_image(uy, sd.Ly);
          dz = min_image(uz, sd.Lz);
          raw = (dx == ux) && (dy == uy) && (dz == uz);
        }
        const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const uint32_t tj = __float_as_uint(pj.w);
        const PairFold pc = row[tj];
        const float sp = s2 + 1e-4f;
        if (HITS && sp < pc.thr && q != (uint32_t)p) {
          const uint32_t at = atomicAdd(&hn[wv][o], 1u);
          if (at < (uint32_t)kPairList) hl[wv][o][at] = q;
        }
        if (sp < pc.thr) {
          cnt++;
          const float w = pair_weight(pc, sp);
          fx = fmaf(dx, w, fx);
          fy = fmaf(dy, w, fy);
          fz = fmaf(dz, w, fz);
        }
        if (EDGES) {
          if (do_edges && raw && tj == ti && s2 < sd.edge_d2) {
            const uint32_t jid = a.id[q];
            if (jid > idi) {
              const uint32_t kk = atomicAdd(&el_n[wv][o], 1u);
              if (kk < (uint32_t)kEdgeList) {
                el_d[wv][o][kk] = make_float2(s2, 0.0f);
                el_i[wv][o][kk] = jid;
              }
            }
          }
        }
      }
    }
  }
  fx = group_sum_f<G>(fx);
  fy = group_sum_f<G>(fy);
  fz = group_sum_f<G>(fz);
  cnt = group_sum_i<G>(cnt);
  if (active && 1e-4f < row[ti].thr) cnt -= 1;  // the particle itself

  if (HITS) {
    // the wave's lists are complete: each in slot order (rank by counting; slots are
    // distinct), all of a particle's partners or, beyond kPairList, only their number
    wave_lds_sync();
    const uint32_t nh = active ? hn[wv][o] : 0u;
    const long k = (long)(p - sd.p0);
    if (nh <= (uint32_t)kPairList)
      for (uint32_t e = g; e < nh; e += G) {
        const uint32_t q = hl[wv][o][e];
        uint32_t rank = 0;
        for (uint32_t e2 = 0; e2 < nh; ++e2) rank += (hl[wv][o][e2] < q) ? 1u : 0u;
        a.hits[k * kPairList + rank] = q;
      }
    if (active && g == 0) a.nhits[k] = nh;
  }

  if (EDGES) {
    // the wave's candidate lists are complete
    wave_lds_sync();
    const uint32_t n = do_edges ? el_n[wv][o] : 0u;
    const bool fits = n <= (uint32_t)kEdgeList;
    const int cap = a.edge_cap, maxc = sd.edge_maxc;
    // membership in S: fewer than cap smaller indices among the candidates
    if (fits)
      for (uint32_t k = g; k < n; k += G) {
        const uint32_t ik = el_i[wv][o][k];
        uint32_t below = 0;
        for (uint32_t k2 = 0; k2 < n; ++k2) below += (el_i[wv][o][k2] < ik) ? 1u : 0u;
        el_d[wv][o][k].y = (below < (uint32_t)cap) ? 1.0f : 0.0f;
      }
    wave_lds_sync();
    // rank of each member of S by (distance, index); keep the first maxc
    if (fits)
      for (uint32_t k = g; k < n; k += G) {
        const float2 dk = el_d[wv][o][k];
        if (dk.y == 0.0f) continue;
        const uint32_t ik = el_i[wv][o][k];
        uint32_t rank = 0;
        for (uint32_t k2 = 0; k2 < n; ++k2) {
          const float2 d2 = el_d[wv][o][k2];
          const uint32_t i2 = el_i[wv][o][k2];
          rank += (d2.y != 0.0f && (d2.x < dk.x || (d2.x == dk.x && i2 < ik))) ? 1u : 0u;
        }
        if (rank < (uint32_t)maxc)
          a.estage[(long)(TILE ? idi : (uint32_t)p) * a.estride + rank] = ik;
      }
    if (do_edges && g == 0) {
      uint32_t e;
      if (fits) {
        e = (uint32_t)min(min((int)n, cap), maxc);
      } else {
        // more candidates than the LDS list holds: exact streaming selection by one lane
        EdgeAcc ea;
        ea.q = 0;
        ea.maxid = 0;
        ea.maxk = 0;
        for (int run = 0; run < 2 * nrows; ++run) {
          const uint2 e = runs[wv][o][run];
          const uint32_t len = e.y & kLenMask, code = e.y >> 26;
          if (!gen && code != 0u) continue;  // image-shifted runs hold no edge partner
          for (uint32_t t = 0; t < len; ++t) {
            const uint32_t q = e.x + t;
            const float4 pj = a.pos[q];
            if (__float_as_uint(pj.w) != ti) continue;
            const float ux = pj.x - pi.x, uy = pj.y - pi.y, uz = pj.z - pi.z;
            if (gen && !(min_image(ux, sd.Lx) == ux && min_image(uy, sd.Ly) == uy &&
                         min_image(uz, sd.Lz) == uz))
              continue;
            const float s2 = fmaf(uz, uz, fmaf(uy, uy, ux * ux));
            if (!(s2 < sd.edge_d2)) continue;
            const uint32_t jid = a.id[q];
            if (jid > idi) edge_push(ea, s2, jid, cap, a.ecand_d2, a.ecand_id, a.P, p);
          }
        }
        e = (uint32_t)edge_finish(ea, cap, maxc, a.ecand_d2, a.ecand_id, a.estage, a.P, p,
                                  TILE ? idi : (uint32_t)p, a.estride);
      }
      if (e != 0u) a.ecount[idi] = e;  // zero already (the sorts clear it)
    }
  }

  if (!active || g != 0) return;
  // ParticleSimulation.cu:136-161
  const int prev = __float_as_int(vi.w);
  const float avg = (float)(cnt + prev) * 0.5f;
  const float dens = fminf(avg / sd.max_exp, 1.0f);
  const float fac = 1.0f - sd.one_minus_balance * dens;
  const float m = sd.fm * fac;
  fx *= m;
  fy *= m;
  fz *= m;
  const float vx = vi.x * sd.friction + fx * sd.dt;
  const float vy = vi.y * sd.friction + fy * sd.dt;
  const float vz = vi.z * sd.friction + fz * sd.dt;
  const float x = wrap_pos(pi.x + vx * sd.dt, sd.Lx);
  const float y = wrap_pos(pi.y + vy * sd.dt, sd.Ly);
  const float z = wrap_pos(pi.z + vz * sd.dt, sd.Lz);
  const float fin = fabsf(x) + fabsf(y) + fabsf(z) + fabsf(vx) + fabsf(vy) + fabsf(vz);
  if (!(fin <= 3.402823466e38f)) atomicOr(a.flag, 1u);
  a.tpos[p] = make_float4(x, y, z, pi.w);
  a.tvel[p] = make_float4(vx, vy, vz, __int_as_float(cnt));
  if (a.acc) a.acc[idi] = make_float4(fx, fy, fz, 0.0f);
  commit_cell<TILE>(a, sd, p, cy, cz, x, y, z);
}

// ------------------------------------------------------------------------------------
// k_force_list: candidate lists in LDS.
//
// k_force_grp walks every run of the stencil in every lane, which at ~20 candidates per
// particle costs more instructions than the candidates themselves.  Here each lane builds
// the runs of its own stencil row(s) in registers, a group prefix scan places them in the
// particle's flattened order, the lanes copy their slots into the wave's LDS list, and each
// particle's G lanes then take its list entries g, g+G, ... with one loop body for shifted
// and unshifted runs alike.  When the wave's particles have at most kWaveList candidates in
// all, their lists lie back to back and take one pass; otherwise each particle has 9 * G
// entries of the list and takes its candidates that many at a time, so the wave's
// particles still advance together (split by position instead, a long list before the
// split and one after it would run one after the other).  Group sums use DPP mirror
// butterflies.

constexpr int kWaveList = 640;  // candidate slots per wave
constexpr int kListEdges = 24;  // LDS edge candidates per particle (more: streaming selection)

// Sum over aligned groups of G lanes (G in 4, 8, 16, 32), every lane gets its group's sum.
template <int G>
__device__ __forceinline__ int dpp_group_sum_i(int v) {
  // quad_perm [1,0,3,2] and [2,3,0,1]: partners within a quad
  v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
  if (G >= 8) v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // half mirror
  if (G >= 16) v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // row mirror
  if (G >= 32) v += __shfl_xor(v, 16, kWave);
  return v;
}

template <int Ctrl>
__device__ __forceinline__ float dpp_add(float x) {
  return x + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), Ctrl, 0xF, 0xF, false));
}

template <int G>
__device__ __forceinline__ float dpp_group_sum_f(float x) {
  x = dpp_add<0xB1>(x);
  x = dpp_add<0x4E>(x);
  if (G >= 8) x = dpp_add<0x141>(x);
  if (G >= 16) x = dpp_add<0x140>(x);
  if (G >= 32) x += __shfl_xor(x, 16, kWave);
  return x;
}

// Exclusive prefix sum over aligned groups of G lanes (g = lane index in the group).
template <int G>
__device__ __forceinline__ uint32_t group_excl_scan(uint32_t v, int g) {
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, kWave);
    if (g >= d) x += y;
  }
  return x - v;
}

// ParticleSimulation.cu:136-161 for one particle: density factor, friction-Euler step and
// wrap; the new state by pre-step slot p, the acceleration (when stored) and the new cell
// (commit_cell).  cyo, czo: the pre-step cell's row and plane.
template <bool TILE>
__device__ __forceinline__ void integrate_commit(const ForceArgs& a, const SimDesc& sd, int p,
                                                 const float4& pi, const float4& vi,
                                                 uint32_t idi, float fx, float fy, float fz,
                                                 int cnt, int cyo, int czo) {
  const int prev = __float_as_int(vi.w);
  const float avg = (float)(cnt + prev) * 0.5f;
  const float dens = fminf(avg / sd.max_exp, 1.0f);
  const float fac = 1.0f - sd.one_minus_balance * dens;
  const float m = sd.fm * fac;
  fx *= m;
  fy *= m;
  fz *= m;
  const float vx = vi.x * sd.friction + fx * sd.dt;
  const float vy = vi.y * sd.friction + fy * sd.dt;
  const float vz = vi.z * sd.friction + fz * sd.dt;
  const float x = wrap_pos(pi.x + vx * sd.dt, sd.Lx);
  const float y = wrap_pos(pi.y + vy * sd.dt, sd.Ly);
  const float z = wrap_pos(pi.z + vz * sd.dt, sd.Lz);
  const float fin = fabsf(x) + fabsf(y) + fabsf(z) + fabsf(vx) + fabsf(vy) + fabsf(vz);
  if (!(fin <= 3.402823466e38f)) atomicOr(a.flag, 1u);
  a.tpos[p] = make_float4(x, y, z, pi.w);
  a.tvel[p] = make_float4(vx, vy, vz, __int_as_float(cnt));
  if (a.acc) a.acc[idi] = make_float4(fx, fy, fz, 0.0f);
  commit_cell<TILE>(a, sd, p, cyo, czo, x, y, z);
}

// CONS (variant 4): the integration tail of the block's 4 * kOwn particles runs once, in
// wave 0 with one lane per particle, instead of once per wave with one lane of each group
// active: each wave leaves its particles' sums in LDS, and after a block barrier wave 0
// reloads their (cache-resident) state and finishes them.  Same arithmetic, same results.
template <bool EDGES, int R, int G, bool TILE, bool CONS = false>
// With the consolidated tail, the 27-cell stencil at G = 8 fits 7 waves per SIMD (72 VGPRs,
// under 22.8 KB of LDS per block).  Without it the velocity stays live to the end and the
// kernel needs a few more registers: held to 72 it spills, and a spilling kernel's scratch
// allocation cuts its occupancy far below 7 waves (a 6x slower launch), so no hint there.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R == 1 && G == 8 && CONS ? 7 : 1, 8)))
k_force_list(ForceArgs a) {
  constexpr int kRows = (2 * R + 1) * (2 * R + 1);
  constexpr int kOwn = kWave / G;
  constexpr int kRpl = (kRows + G - 1) / G;  // stencil rows per lane
  // the pair table by field (kMaxTypes^2 words each), so the lanes of different owners,
  // which read rows of different types, fall on different banks: thr, then the force
  // constants folded per type pair (rep * fv, att * fv * inv_eff, kexp2 * inv_eff^2).
  // One table per block: a block's 4 * kOwn particles (at most 64, aligned) belong to one
  // simulation, and each wave writes the whole table itself (the same values as the other
  // waves) before it reads it, so no block barrier is needed.
  static_assert(4 * kOwn <= kWave, "a block's particles lie in one simulation's 64 slots");
  constexpr int kTT = kMaxTypes * kMaxTypes;
  __shared__ float ptab[4 * kTT];
  // list entries: slot | image-shift code << kCodeShift (slots stay below 2^kCodeShift),
  // the wave's particles one after another
  __shared__ uint32_t lslot[4][kWaveList];
  __shared__ float2 el_d[EDGES ? 4 : 1][EDGES ? kOwn : 1][EDGES ? kListEdges : 1];
  __shared__ uint32_t el_i[EDGES ? 4 : 1][EDGES ? kOwn : 1][EDGES ? kListEdges : 1];
  __shared__ uint32_t el_n[4][kOwn];
  __shared__ float4 cres[CONS ? 4 * kOwn : 1];    // force sums and count bits (CONS)
  __shared__ uint32_t cedge[CONS ? 4 * kOwn : 1];  // kept edges | active << 30 | edges << 31
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = xcd_block() * (blockDim.x >> 6) + wv;
  const long base = (long)gw * kOwn;
  const bool wave_ok = base < (long)a.nwaves * kWave;
  const int o = lane / G, g = lane % G;
  const int p = (int)base + o;
  // the slot's state loads do not wait for the descriptor: every slot below nwaves * 64
  // is allocated (a padding slot's values are loaded but never used)
  Own c;
  float4 vi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t idi = 0;
  c.pi = vi;
  if (wave_ok) {
    c.pi = a.pos[p];
    vi = a.vel[p];
    idi = a.id[p];
  }
  const int s = __builtin_amdgcn_readfirstlane(wave_ok ? (int)a.wave_sim[base / kWave] : 0);
  const SimDesc sd = a.sims[s];
  if (wave_ok)
    for (int t = lane; t < sd.T * sd.T; t += kWave) {
      const PairConst pt = a.pairs[sd.tab + t];
      ptab[t] = pt.thr;
      ptab[kTT + t] = sd.rep * pt.fv;
      ptab[2 * kTT + t] = sd.att * pt.fv * pt.inv_eff;
      ptab[3 * kTT + t] = sd.kexp2 * pt.inv_eff * pt.inv_eff;
    }
  const bool active = wave_ok && (p - sd.p0) < sd.n;
  const float4 pi = c.pi;
  c.ti = active ? __float_as_uint(pi.w) : 0u;
  const uint32_t ti = c.ti;
  c.cx = bin_axis(pi.x, sd.ihx, sd.nx);
  c.cy = bin_axis(pi.y, sd.ihy, sd.ny);
  c.cz = bin_axis(pi.z, sd.ihz, sd.nz);
  c.gen = sd.generic != 0;
  c.nyo = c.gen && sd.ny < 2 * R + 1 ? sd.ny : 2 * R + 1;
  c.nzo = c.gen && sd.nz < 2 * R + 1 ? sd.nz : 2 * R + 1;
  const int nrows = c.nyo * c.nzo;
  c.fxr = fminf(fmaxf(pi.x * sd.ihx - (float)c.cx, 0.0f), 1.0f);
  c.fyr = fminf(fmaxf(pi.y * sd.ihy - (float)c.cy, 0.0f), 1.0f);
  c.fzr = fminf(fmaxf(pi.z * sd.ihz - (float)c.cz, 0.0f), 1.0f);
  own_gaps<R>(c, sd);
  if (EDGES && g == 0) el_n[wv][o] = 0u;
  // the pruning radius straight from the pair table in memory, so the stencil rows (and
  // their cell-start loads) overlap the copy of the table into LDS
  c.cut2 = active ? a.pairs[sd.tab + ti * sd.T].pad : 0.0f;

  // ---- this lane's runs, their place in the particle's flattened candidate order ----
  uint32_t rs[2 * kRpl], rl[2 * kRpl], rk[2 * kRpl];
  uint32_t mine = 0u;
  if (R == 1 && G == 8 && !c.gen) {
    // 9 rows over 8 lanes in one pass: lane g takes row g (lane 4 row 8 in place of the
    // centre row 4), and every lane works out the centre row, which has no y or z wrap
    // and no y or z gap (so the row itself folds to its x ends), for lane 0 to keep
    static_assert(R != 1 || G != 8 || kRpl == 2, "two run pairs per lane");
#pragma unroll
    for (int k = 0; k < 2 * kRpl; ++k) rs[k] = rl[k] = rk[k] = 0u;
    if (active) {
      row_runs_fast<R>(a, sd, c, g == 4 ? 8 : g, rs[0], rl[0], rk[0], rs[1], rl[1], rk[1]);
      row_runs_fast<R>(a, sd, c, 4, rs[2], rl[2], rk[2], rs[3], rl[3], rk[3]);
      if (g != 0) rl[2] = rl[3] = 0u;
    }
#pragma unroll
    for (int k = 0; k < 2 * kRpl; ++k) mine += rl[k];
  } else {
#pragma unroll
    for (int i = 0; i < kRpl; ++i) {
      const int r = g + i * G;
      rs[2 * i] = rl[2 * i] = rk[2 * i] = rs[2 * i + 1] = rl[2 * i + 1] = rk[2 * i + 1] = 0u;
      if (active && r < nrows)
        row_runs<R>(a, sd, c, r, rs[2 * i], rl[2 * i], rk[2 * i], rs[2 * i + 1], rl[2 * i + 1],
                    rk[2 * i + 1]);
      mine += rl[2 * i] + rl[2 * i + 1];
    }
  }
  // each run as its first list entry (slot | code << kCodeShift): one register per run
#pragma unroll
  for (int k = 0; k < 2 * kRpl; ++k) rs[k] |= rk[k] << kCodeShift;
  const uint32_t first = group_excl_scan<G>(mine, g);
  const uint32_t total = (uint32_t)dpp_group_sum_i<G>((int)mine);
  // the particle's part of the wave's list: from its first position in the wave's order
  // (an inclusive scan of the particles' counts over the wave's groups, less its own) when
  // the wave's candidates fit the list at once, else kQuota entries at o * kQuota.  kQuota
  // is an odd multiple of G: each lane keeps the entries it takes in one pass (g, g+G, ...)
  // and the particles' lanes of one read fall on different banks.
  static_assert(kWaveList % kWave == 0, "list of whole rows");
  constexpr uint32_t kQuota = (uint32_t)(G * ((kWaveList / kWave - 1) | 1));
  uint32_t incl = total;
#pragma unroll
  for (int d = G; d < kWave; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += y;
  }
  const bool packed = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1) <= (uint32_t)kWaveList;
  const uint32_t obase = packed ? incl - total : (uint32_t)o * kQuota;
  const uint32_t quota = packed ? (uint32_t)kWaveList : kQuota;
  wave_lds_sync();  // pair table ready
  const float* row = &ptab[ti * sd.T];

  // ---- candidates, quota at a time -------------------------------------------------
  const bool want_edges = EDGES && a.edges && sd.edge_on;  // the same for the whole wave
  const bool do_edges = want_edges && active;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  int cnt = 0;
  for (uint32_t b0 = 0; b0 < total; b0 += quota) {
    const uint32_t b1 = b0 + quota;
    uint32_t f0 = first;
#pragma unroll
    for (int k = 0; k < 2 * kRpl; ++k) {
      const uint32_t lo = max(f0, b0), hi = min(f0 + rl[k], b1);
      // entries lo..hi-1 of run k: (slot | code << kCodeShift) + f - f0 (no carry), two per
      // store instruction
      if (lo < hi) {
        uint32_t* d = &lslot[wv][obase + (lo - b0)];
        uint32_t v = rs[k] + (lo - f0);
        uint32_t m = hi - lo;
        for (; m >= 2u; m -= 2u, d += 2, v += 2u) {
          d[0] = v;
          d[1] = v + 1u;
        }
        if (m) d[0] = v;
      }
      f0 += rl[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint32_t nb = min(total - b0, quota);
    // one copy of the loop per grid kind (the choice is the same for the whole wave), so
    // the common kind carries no per-pair minimum-image code
    auto pairs = [&](auto gen_kind) {
      constexpr bool GEN = decltype(gen_kind)::value;
      for (uint32_t k = (uint32_t)g; k < nb; k += (uint32_t)G) {
        const uint32_t e = lslot[wv][obase + k];
        const uint32_t q = e & ((1u << kCodeShift) - 1u), code = e >> kCodeShift;
        const float4 pj = a.pos[q];
        // the index travels with the position: an edge candidate then waits for no load
        const uint32_t jq = want_edges ? a.id[q] : 0u;
        float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        bool raw;
        if constexpr (!GEN) {
          raw = code == 0u;
          if (!raw) {  // image-shifted run
            dx += shift_of(code & 3u, sd.Lx);
            dy += shift_of((code >> 2) & 3u, sd.Ly);
            dz += shift_of((code >> 4) & 3u, sd.Lz);
          }
        } else {  // per-pair minimum image
          const float ux = dx, uy = dy, uz = dz;
          dx = min_image(ux, sd.Lx);
          dy = min_image(uy, sd.Ly);
          dz = min_image(uz, sd.Lz);
          raw = (dx == ux) && (dy == uy) && (dz == uz);
        }
        const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const uint32_t tj = __float_as_uint(pj.w);
        const float thr = row[tj];
        const float sp = s2 + 1e-4f;
        if (sp < thr) {
          cnt++;
          // ParticleSimulation.cu:104-131 with rr = r / effR, r = sqrt(sp):
          // (rep e^(-k rr^2) - att rr) fv / r, the constants folded per type pair
          const float rep_fv = row[kTT + tj], att_fv_ie = row[2 * kTT + tj];
          const float k_ie2 = row[3 * kTT + tj];
          const float inv = __builtin_amdgcn_rsqf(sp);
          const float ex = __builtin_amdgcn_exp2f(k_ie2 * sp);
          const float w = fmaf(rep_fv, ex, -(att_fv_ie * (sp * inv))) * inv;
          fx = fmaf(dx, w, fx);
          fy = fmaf(dy, w, fy);
          fz = fmaf(dz, w, fz);
        }
        if (EDGES) {
          // edges are non-periodic (ParticleSimulation.cu:222-227): unshifted pairs only
          if (do_edges && raw && tj == ti && s2 < sd.edge_d2) {
            const uint32_t jid = jq;
            if (jid > idi) {
              const uint32_t kk = atomicAdd(&el_n[wv][o], 1u);
              if (kk < (uint32_t)kListEdges) {
                el_d[wv][o][kk] = make_float2(s2, 0.0f);
                el_i[wv][o][kk] = jid;
              }
            }
          }
        }
      }
    };
    if (c.gen) pairs(std::true_type{});
    else pairs(std::false_type{});
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // list reused by the next chunk
  }
  fx = dpp_group_sum_f<G>(fx);
  fy = dpp_group_sum_f<G>(fy);
  fz = dpp_group_sum_f<G>(fz);
  cnt = dpp_group_sum_i<G>(cnt);
  if (active && 1e-4f < row[ti]) cnt -= 1;  // the particle itself
  // CONS: the sums go to LDS now, so they take no registers through the edge selection
  if (CONS && g == 0) cres[wv * kOwn + o] = make_float4(fx, fy, fz, __int_as_float(cnt));

  uint32_t ekept = 0u;  // kept edges of the particle (lane g == 0), passed on by CONS
  if (EDGES) {
    // the wave's candidate lists are complete
    wave_lds_sync();
    const uint32_t n = do_edges ? el_n[wv][o] : 0u;
    const bool fits = n <= (uint32_t)kListEdges;
    const int cap = a.edge_cap, maxc = sd.edge_maxc;
    if (fits)
      for (uint32_t k = g; k < n; k += G) {
        const uint32_t ik = el_i[wv][o][k];
        uint32_t below = 0;
        for (uint32_t k2 = 0; k2 < n; ++k2) below += (el_i[wv][o][k2] < ik) ? 1u : 0u;
        el_d[wv][o][k].y = (below < (uint32_t)cap) ? 1.0f : 0.0f;
      }
    wave_lds_sync();
    if (fits)
      for (uint32_t k = g; k < n; k += G) {
        const float2 dk = el_d[wv][o][k];
        if (dk.y == 0.0f) continue;
        const uint32_t ik = el_i[wv][o][k];
        uint32_t rank = 0;
        for (uint32_t k2 = 0; k2 < n; ++k2) {
          const float2 d2 = el_d[wv][o][k2];
          const uint32_t i2 = el_i[wv][o][k2];
          rank += (d2.y != 0.0f && (d2.x < dk.x || (d2.x == dk.x && i2 < ik))) ? 1u : 0u;
        }
        if (rank < (uint32_t)maxc)
          a.estage[(long)(TILE ? idi : (uint32_t)p) * a.estride + rank] = ik;
      }
    // Owners with more candidates than their LDS list holds (n counts them all): the whole
    // wave takes them one at a time, in the wave's LDS lists (free now: every fitting
    // owner's selection above is done).  Its 64 lanes walk the owner's flattened candidates
    // again, append the qualifying ones (d^2 in el_d's words, index in lslot's, up to
    // kOvfCap), keep the cap smallest indices and rank those by (d^2, index), as above.
    // More than kOvfCap qualifying: the owner's lane takes the streaming selection below.
    uint32_t ovf_e = ~0u;  // kept edges of an owner this handled (its lane g == 0)
    if constexpr (G == 8) {
      constexpr int kOvfCap = 320;
      static_assert(kOvfCap <= kWaveList, "indices fit the wave's lslot");
      static_assert(kOvfCap <= 2 * kOwn * kListEdges, "d^2 fits the wave's el_d");
      constexpr int kOvfInfo = G * 2 * kRpl * 2;  // after the runs in the wave's el_i
      static_assert(kOvfInfo + 7 <= kOwn * kListEdges, "runs and owner fit the wave's el_i");
      float* ov_d = &el_d[wv][0][0].x;
      uint32_t* ov_i = &lslot[wv][0];
      uint32_t* ov_r = &el_i[wv][0][0];
      unsigned long long left = __ballot(do_edges && g == 0 && !fits);
      if (left) wave_lds_sync();  // the fitting owners' reads of el_d, el_i are done
      while (left) {
        const int L = __ffsll((long long)left) - 1;
        left &= left - 1ull;
        const int os = L / G;
        // the owner's runs, position, type, index, candidate count and record go to LDS
        // (read where used: they take no registers through the walk)
        if (o == os)
#pragma unroll
          for (int k = 0; k < 2 * kRpl; ++k) {
            ov_r[(g * 2 * kRpl + k) * 2] = rs[k];
            ov_r[(g * 2 * kRpl + k) * 2 + 1] = rl[k];
          }
        if (lane == L) {
          ov_r[kOvfInfo] = __float_as_uint(pi.x);
          ov_r[kOvfInfo + 1] = __float_as_uint(pi.y);
          ov_r[kOvfInfo + 2] = __float_as_uint(pi.z);
          ov_r[kOvfInfo + 3] = ti;
          ov_r[kOvfInfo + 4] = idi;
          ov_r[kOvfInfo + 5] = total;
          ov_r[kOvfInfo + 6] = TILE ? idi : (uint32_t)p;
          el_n[wv][os] = 0u;
        }
        wave_lds_sync();
        const uint32_t otot = ov_r[kOvfInfo + 5];
        for (uint32_t i0 = 0; i0 < otot; i0 += (uint32_t)kWave) {
          const uint32_t i = i0 + (uint32_t)lane;
          if (i >= otot) continue;
          uint32_t off = i, ent = 0u;
          for (int r = 0; r < G * 2 * kRpl; ++r) {
            const uint32_t len = ov_r[2 * r + 1];
            if (off < len) {
              ent = ov_r[2 * r] + off;
              break;
            }
            off -= len;
          }
          const uint32_t q = ent & ((1u << kCodeShift) - 1u);
          if (!c.gen && (ent >> kCodeShift) != 0u) continue;  // image-shifted: no partner
          const float4 pj = a.pos[q];
          if (__float_as_uint(pj.w) != ov_r[kOvfInfo + 3]) continue;
          const float ux = pj.x - __uint_as_float(ov_r[kOvfInfo]),
                      uy = pj.y - __uint_as_float(ov_r[kOvfInfo + 1]),
                      uz = pj.z - __uint_as_float(ov_r[kOvfInfo + 2]);
          if (c.gen && !(min_image(ux, sd.Lx) == ux && min_image(uy, sd.Ly) == uy &&
                         min_image(uz, sd.Lz) == uz))
            continue;
          const float s2 = fmaf(uz, uz, fmaf(uy, uy, ux * ux));
          if (!(s2 < sd.edge_d2)) continue;
          const uint32_t jid = a.id[q];
          if (jid <= ov_r[kOvfInfo + 4]) continue;
          const uint32_t k = atomicAdd(&el_n[wv][os], 1u);
          if (k < (uint32_t)kOvfCap) {
            ov_d[k] = s2;
            ov_i[k] = jid;
          }
        }
        wave_lds_sync();
        const uint32_t m = el_n[wv][os];
        if (m <= (uint32_t)kOvfCap) {
          // the cap smallest indices (marked in bit 31; indices stay below 2^26)
          uint32_t keep = 0u;
          for (uint32_t k = lane, b = 0; k < m; k += kWave, ++b) {
            const uint32_t ik = ov_i[k];
            uint32_t below = 0;
            for (uint32_t k2 = 0; k2 < m; ++k2) below += (ov_i[k2] < ik) ? 1u : 0u;
            keep |= (below < (uint32_t)cap ? 1u : 0u) << b;
          }
          wave_lds_sync();
          for (uint32_t k = lane, b = 0; k < m; k += kWave, ++b)
            if ((keep >> b) & 1u) ov_i[k] |= 0x80000000u;
          wave_lds_sync();
          for (uint32_t k = lane; k < m; k += kWave) {
            const uint32_t ik = ov_i[k];
            if (!(ik & 0x80000000u)) continue;
            const float dk = ov_d[k];
            uint32_t rank = 0;
            for (uint32_t k2 = 0; k2 < m; ++k2) {
              const uint32_t i2 = ov_i[k2];
              const float d2 = ov_d[k2];
              rank += ((i2 & 0x80000000u) && (d2 < dk || (d2 == dk && i2 < ik))) ? 1u : 0u;
            }
            if (rank < (uint32_t)maxc)
              a.estage[(long)ov_r[kOvfInfo + 6] * a.estride + rank] = ik & 0x7FFFFFFFu;
          }
          if (lane == L) ovf_e = (uint32_t)min(min((int)m, cap), maxc);
        }
        wave_lds_sync();  // the LDS lists are reused for the next owner
      }
    }
    if (do_edges && g == 0) {
      uint32_t e;
      if (fits) {
        e = (uint32_t)min(min((int)n, cap), maxc);
      } else if (ovf_e != ~0u) {
        e = ovf_e;
      } else {
        // more candidates than the wave's LDS lists hold: exact streaming selection by one
        // lane, over the stencil rows again
        EdgeAcc ea;
        ea.q = 0;
        ea.maxid = 0;
        ea.maxk = 0;
        for (int r = 0; r < nrows; ++r) {
          uint32_t st[2], ln[2], cd[2];
          row_runs<R>(a, sd, c, r, st[0], ln[0], cd[0], st[1], ln[1], cd[1]);
          for (int h = 0; h < 2; ++h) {
            if (!c.gen && cd[h] != 0u) continue;  // image-shifted runs hold no edge partner
            for (uint32_t t = 0; t < ln[h]; ++t) {
              const uint32_t q = st[h] + t;
              const float4 pj = a.pos[q];
              if (__float_as_uint(pj.w) != ti) continue;
              const float ux = pj.x - pi.x, uy = pj.y - pi.y, uz = pj.z - pi.z;
              if (c.gen && !(min_image(ux, sd.Lx) == ux && min_image(uy, sd.Ly) == uy &&
                             min_image(uz, sd.Lz) == uz))
                continue;
              const float s2 = fmaf(uz, uz, fmaf(uy, uy, ux * ux));
              if (!(s2 < sd.edge_d2)) continue;
              const uint32_t jid = a.id[q];
              if (jid > idi) edge_push(ea, s2, jid, cap, a.ecand_d2, a.ecand_id, a.P, p);
            }
          }
        }
        e = (uint32_t)edge_finish(ea, cap, maxc, a.ecand_d2, a.ecand_id, a.estage, a.P, p,
                                  TILE ? idi : (uint32_t)p, a.estride);
      }
      ekept = e;
      if (!CONS && e != 0u) a.ecount[idi] = e;  // zero already (the sorts clear it)
    }
  }

  if constexpr (CONS) {
    const bool de = EDGES && a.edges && sd.edge_on && active;
    if (g == 0) cedge[wv * kOwn + o] = ekept | (active ? (1u << 30) : 0u) | (de ? (1u << 31) : 0u);
    __syncthreads();
    if (wv != 0 || lane >= 4 * kOwn) return;
    // wave 0: lane L finishes the block's particle L
    const uint32_t fl = cedge[lane];
    if (!(fl & (1u << 30))) return;
    const int pl = (int)(xcd_block() * 4 * kOwn) + lane;
    const float4 r = cres[lane];
    const float4 pl_pos = a.pos[pl];
    const float4 pl_vel = a.vel[pl];
    const uint32_t pl_id = a.id[pl];
    if (fl & (1u << 31)) {
      const uint32_t e = fl & 0x3FFFFFFFu;
      if (e != 0u) a.ecount[pl_id] = e;  // zero already (the sorts clear it)
    }
    integrate_commit<TILE>(a, sd, pl, pl_pos, pl_vel, pl_id, r.x, r.y, r.z,
                           __float_as_int(r.w), bin_axis(pl_pos.y, sd.ihy, sd.ny),
                           bin_axis(pl_pos.z, sd.ihz, sd.nz));
    return;
  }
  if (!active || g != 0) return;
  integrate_commit<TILE>(a, sd, p, pi, vi, idi, fx, fy, fz, cnt, c.cy, c.cz);
}

// ------------------------------------------------------------------------------------
// k_edges_g: the standalone proximity-graph pass (the state after a call's last step, or
// every step when the edge distance reaches beyond the force stencil).  G lanes per owner
// take the rows of the owner's non-periodic stencil (ParticleSimulation.cu:222-227: no
// minimum image) of per-axis reach edge_r*; rows and row ends farther from the owner than
// the edge distance are dropped (the same gap test as the force stencil's pruning, with its
// margins), which leaves about a fifth of the cells at the 100k workload's 0.02 over a
// 0.052 cell.  Qualifying candidates (same type :219, larger index :215, d^2 < D^2 :227) go
// to the owner's LDS list; the two-phase selection then keeps the 2*maxC smallest indices
// and ranks them by (d^2, index) (:228-246).  An owner with more than kListEdges candidates
// takes the exact streaming selection of one lane.  Kept partners are staged by owner index
// (estage[index * estride + rank]), the non-zero per-owner counts go to `ecount`, and
// k_tile_sort's edge tiles write the canonical list.

constexpr int kEdgeLanes = 8;

// Row r of an owner's non-periodic edge stencil: false when it lies outside the grid or
// farther than the pruning radius; else its cell range [cell0, cell1] (one row, x-adjacent
// cells, so one run of slots from start[cell0] to start[cell1 + 1]).
__device__ __forceinline__ bool edge_row(const SimDesc& sd, int cx, int cy, int cz, float fxr,
                                         float fyr, float fzr, float cut2, int r, int& cell0,
                                         int& cell1) {
  const int nyr = 2 * sd.edge_ry + 1;
  const int dy = r % nyr - sd.edge_ry, dz = r / nyr - sd.edge_rz;
  const int y = cy + dy, z = cz + dz;
  if (y < 0 || y >= sd.ny || z < 0 || z >= sd.nz) return false;
  const float gy = cell_gap(dy, fyr) * sd.hy, gz = cell_gap(dz, fzr) * sd.hz;
  const float gyz = fmaf(gy, gy, gz * gz);
  if (!(gyz < cut2)) return false;
  int xlo = max(cx - sd.edge_rx, 0), xhi = min(cx + sd.edge_rx, sd.nx - 1);
  while (xlo < cx) {
    const float gx = cell_gap(xlo - cx, fxr) * sd.hx;
    if (fmaf(gx, gx, gyz) < cut2) break;
    ++xlo;
  }
  while (xhi > cx) {
    const float gx = cell_gap(xhi - cx, fxr) * sd.hx;
    if (fmaf(gx, gx, gyz) < cut2) break;
    --xhi;
  }
  const int rowc = sd.c0 + sd.nx * (y + sd.ny * z);
  cell0 = rowc + xlo;
  cell1 = rowc + xhi;
  return true;
}

// An owner of the standalone graph pass: its cell, and for reach ER (> 0) the gaps to
// the k-th cells below and above per axis, cell_gap(-k, f) * h and cell_gap(k, f) * h.
template <int ER>
struct EdgeOwner {
  int cx, cy, cz;
  float fxr, fyr, fzr;
  float gxl[ER > 0 ? ER : 1], gxr[ER > 0 ? ER : 1], gyl[ER > 0 ? ER : 1],
      gyr[ER > 0 ? ER : 1], gzl[ER > 0 ? ER : 1], gzr[ER > 0 ? ER : 1];
};

// edge_row for reach ER on every axis, straight-line: the row's gaps from the owner's
// precomputed ones, and the x ends by counting the x offsets within reach (the gap grows
// with the offset, so those offsets are 1..K on each side: where edge_row's trimming from
// the far end stops).  Same result as edge_row.
template <int ER>
__device__ __forceinline__ bool edge_row_fast(const SimDesc& sd, const EdgeOwner<ER>& w,
                                              float cut2, int r, int& cell0, int& cell1) {
  constexpr int D = 2 * ER + 1;
  const int dy = r % D - ER, dz = r / D - ER;
  const int y = w.cy + dy, z = w.cz + dz;
  if (y < 0 || y >= sd.ny || z < 0 || z >= sd.nz) return false;
  const float gy = gap_of<ER>(dy, w.gyl, w.gyr), gz = gap_of<ER>(dz, w.gzl, w.gzr);
  const float gyz = fmaf(gy, gy, gz * gz);
  if (!(gyz < cut2)) return false;
  int xlo = w.cx, xhi = w.cx;
#pragma unroll
  for (int k = 0; k < ER; ++k) {
    xlo -= (fmaf(w.gxl[k], w.gxl[k], gyz) < cut2) ? 1 : 0;
    xhi += (fmaf(w.gxr[k], w.gxr[k], gyz) < cut2) ? 1 : 0;
  }
  const int rowc = sd.c0 + sd.nx * (y + sd.ny * z);
  cell0 = rowc + max(xlo, 0);
  cell1 = rowc + min(xhi, sd.nx - 1);
  return true;
}

// Runs a lane's stencil rows r = r0, r0 + G, ... (kEdgeRowBatch of them at a time, their
// cell-start loads issued together) and calls f(q, pj, jid) for each slot q, in row order.
// The batch's slot runs are walked as one sequence, kEdgeLoadBatch slots at a time, so the
// loads of several short rows are in flight together; an index is loaded with its
// position, so a qualifying candidate waits for no further load.
constexpr int kEdgeRowBatch = 3;
constexpr int kEdgeLoadBatch = 4;

// ER > 0: reach ER on every axis (edge_row_fast); ER = 0: any reach (edge_row).
template <int ER, class F>
__device__ __forceinline__ void edge_walk(const EdgeArgs& a, const SimDesc& sd,
                                          const EdgeOwner<ER>& w, float cut2, int r0, int step,
                                          int nrows, F&& f) {
  for (int rb = r0; rb < nrows; rb += kEdgeRowBatch * step) {
    int c0[kEdgeRowBatch], c1[kEdgeRowBatch];
    bool on[kEdgeRowBatch];
#pragma unroll
    for (int i = 0; i < kEdgeRowBatch; ++i) {
      const int r = rb + i * step;
      if constexpr (ER > 0)
        on[i] = r < nrows && edge_row_fast<ER>(sd, w, cut2, r, c0[i], c1[i]);
      else
        on[i] = r < nrows && edge_row(sd, w.cx, w.cy, w.cz, w.fxr, w.fyr, w.fzr, cut2, r,
                                      c0[i], c1[i]);
    }
    int qa[kEdgeRowBatch], len[kEdgeRowBatch];
#pragma unroll
    for (int i = 0; i < kEdgeRowBatch; ++i) {
      qa[i] = on[i] ? (int)a.start[c0[i]] : 0;
      len[i] = on[i] ? (int)a.start[c1[i] + 1] : 0;
    }
    int total = 0;
#pragma unroll
    for (int i = 0; i < kEdgeRowBatch; ++i) {
      len[i] -= qa[i];
      total += len[i];
    }
    for (int x0 = 0; x0 < total; x0 += kEdgeLoadBatch) {
      int qs[kEdgeLoadBatch];
      float4 pj[kEdgeLoadBatch];
      uint32_t jid[kEdgeLoadBatch];
#pragma unroll
      for (int u = 0; u < kEdgeLoadBatch; ++u) {
        const int x = x0 + u;
        int q = qa[0] + x, run = len[0];  // x's row: the first whose runs reach past x
#pragma unroll
        for (int i = 1; i < kEdgeRowBatch; ++i) {
          if (x >= run) q = qa[i] + (x - run);
          run += len[i];
        }
        qs[u] = q;
        const bool in = x < total;
        pj[u] = in ? a.pos[q] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        jid[u] = in ? a.id[q] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kEdgeLoadBatch; ++u)
        if (x0 + u < total) f(qs[u], pj[u], jid[u]);
    }
  }
}

template <int ER>
__device__ __forceinline__ void edges_owner(const EdgeArgs& a, const SimDesc& sd, int wv, int o,
                                            int g, int p, bool active, const float4& pi,
                                            uint32_t idi, float2 (*el_d)[kListEdges],
                                            uint32_t (*el_i)[kListEdges], uint32_t* el_n) {
  constexpr int G = kEdgeLanes;
  const uint32_t ti = __float_as_uint(pi.w);
  EdgeOwner<ER> w;
  w.cx = bin_axis(pi.x, sd.ihx, sd.nx);
  w.cy = bin_axis(pi.y, sd.ihy, sd.ny);
  w.cz = bin_axis(pi.z, sd.ihz, sd.nz);
  w.fxr = fminf(fmaxf(pi.x * sd.ihx - (float)w.cx, 0.0f), 1.0f);
  w.fyr = fminf(fmaxf(pi.y * sd.ihy - (float)w.cy, 0.0f), 1.0f);
  w.fzr = fminf(fmaxf(pi.z * sd.ihz - (float)w.cz, 0.0f), 1.0f);
#pragma unroll
  for (int k = 1; k <= ER; ++k) {
    w.gxl[k - 1] = cell_gap(-k, w.fxr) * sd.hx;
    w.gxr[k - 1] = cell_gap(k, w.fxr) * sd.hx;
    w.gyl[k - 1] = cell_gap(-k, w.fyr) * sd.hy;
    w.gyr[k - 1] = cell_gap(k, w.fyr) * sd.hy;
    w.gzl[k - 1] = cell_gap(-k, w.fzr) * sd.hz;
    w.gzr[k - 1] = cell_gap(k, w.fzr) * sd.hz;
  }
  // pruning radius: the edge distance with the force stencil's margins
  const float cut2 = sd.edge_d2 * 1.001f + 1e-12f;
  const int nrows = (2 * sd.edge_ry + 1) * (2 * sd.edge_rz + 1);
  if (g == 0) el_n[o] = 0u;
  wave_lds_sync();
  if (active)
    edge_walk<ER>(a, sd, w, cut2, g, G, nrows,
              [&](int, const float4& pj, uint32_t jid) {
                if (__float_as_uint(pj.w) != ti) return;        // :219
                const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if (!(s2 < sd.edge_d2)) return;                  // :227
                if (jid <= idi) return;                          // j > i, :215
                const uint32_t kk = atomicAdd(&el_n[o], 1u);
                if (kk < (uint32_t)kListEdges) {
                  el_d[o][kk] = make_float2(s2, 0.0f);
                  el_i[o][kk] = jid;
                }
              });
  wave_lds_sync();
  const uint32_t n = active ? el_n[o] : 0u;
  const bool fits = n <= (uint32_t)kListEdges;
  const int cap = a.edge_cap, maxc = sd.edge_maxc;
  // phase 1: keep the `cap` smallest indices (the reference's first 2*maxC by index, :215)
  if (fits)
    for (uint32_t k = g; k < n; k += G) {
      const uint32_t ik = el_i[o][k];
      uint32_t below = 0;
      for (uint32_t k2 = 0; k2 < n; ++k2) below += (el_i[o][k2] < ik) ? 1u : 0u;
      el_d[o][k].y = (below < (uint32_t)cap) ? 1.0f : 0.0f;
    }
  wave_lds_sync();
  // phase 2: rank the kept ones by (d^2, index), keep the first maxC (:235-246)
  if (fits)
    for (uint32_t k = g; k < n; k += G) {
      const float2 dk = el_d[o][k];
      if (dk.y == 0.0f) continue;
      const uint32_t ik = el_i[o][k];
      uint32_t rank = 0;
      for (uint32_t k2 = 0; k2 < n; ++k2) {
        const float2 d2 = el_d[o][k2];
        const uint32_t i2 = el_i[o][k2];
        rank += (d2.y != 0.0f && (d2.x < dk.x || (d2.x == dk.x && i2 < ik))) ? 1u : 0u;
      }
      if (rank < (uint32_t)maxc) a.estage[(long)idi * a.estride + rank] = ik;
    }
  if (!active || g != 0) return;
  uint32_t e;
  if (fits) {
    e = (uint32_t)min(min((int)n, cap), maxc);
  } else {
    // more candidates than the LDS list holds: exact streaming selection by this lane
    EdgeAcc ea;
    ea.q = 0;
    ea.maxid = 0;
    ea.maxk = 0;
    edge_walk<ER>(a, sd, w, cut2, 0, 1, nrows,
              [&](int, const float4& pj, uint32_t jid) {
                if (__float_as_uint(pj.w) != ti) return;
                const float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
                const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if (!(s2 < sd.edge_d2)) return;
                if (jid > idi) edge_push(ea, s2, jid, cap, a.ecand_d2, a.ecand_id, a.P, p);
              });
    e = (uint32_t)edge_finish(ea, cap, maxc, a.ecand_d2, a.ecand_id, a.estage, a.P, p, idi,
                              a.estride);
  }
  if (e) a.ecount[idi] = e;  // zero already: the edge tiles that read the counts clear them
}

__global__ void __launch_bounds__(256) k_edges_g(EdgeArgs a) {
  constexpr int G = kEdgeLanes;
  constexpr int kOwn = kWave / G;
  __shared__ float2 el_d[4][kOwn][kListEdges];
  __shared__ uint32_t el_i[4][kOwn][kListEdges];
  __shared__ uint32_t el_n[4][kOwn];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = xcd_block() * (blockDim.x >> 6) + wv;
  const long base = (long)gw * kOwn;
  const bool wave_ok = base < (long)a.nwaves * kWave;
  const int o = lane / G, g = lane % G;
  const int p = (int)base + o;
  const int s = __builtin_amdgcn_readfirstlane(wave_ok ? (int)a.wave_sim[base / kWave] : 0);
  const SimDesc sd = a.sims[s];
  const bool active = wave_ok && (p - sd.p0) < sd.n;
  float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t idi = 0;
  if (active) {
    pi = a.pos[p];
    idi = a.id[p];
  }
  if (sd.edge_rx == 1 && sd.edge_ry == 1 && sd.edge_rz == 1)
    edges_owner<1>(a, sd, wv, o, g, p, active, pi, idi, el_d[wv], el_i[wv], el_n[wv]);
  else
    edges_owner<0>(a, sd, wv, o, g, p, active, pi, idi, el_d[wv], el_i[wv], el_n[wv]);
}

// ------------------------------------------------------------------------------------
// k_tile_sort: the (cell, index) sort of one step, and the step's canonical edge list.
//
// The force pass leaves, by pre-step slot, the new state and the new cell (key), and
// counts per cell and per cell tile.  A cell tile's block needs no other block's result:
// its first slot is its simulation's first slot plus the counts of the simulation's tiles
// before it, its cells' starts follow from the per-cell counts, and its particles are
// found by reading the keys of its window (and the far list).  Inside a cell they are
// ranked by index, so the order is the canonical one whatever order the atomics arrived
// in.  An edge tile writes its owners' edges at the sum of the counts of the edge tiles
// before it (a decoupled look-back) plus a local scan.  Each cell-tile block also clears its
// own counters for the next step (the next step counts into the other parity).

constexpr int kSortThreads = 256;
constexpr int kVisitBatch = 8;  // window slots per thread with their loads in flight together
constexpr int kMoveBatch = 4;   // members per thread moved together

// Flag word of a look-back tile: launch epoch (30 bits), status (0 none, 1 the tile's own
// sum, 2 the inclusive prefix), value.
__device__ __forceinline__ unsigned long long pack_flag(uint32_t epoch, uint32_t status,
                                                        uint32_t value) {
  return ((unsigned long long)(epoch & 0x3FFFFFFFu) << 34) |
         ((unsigned long long)status << 32) | (unsigned long long)value;
}

// Single-pass decoupled look-back for tile g, in two halves, each run by one whole wave.
// lookback_publish publishes the tile's sum `total` (as its inclusive prefix when g ==
// first).  lookback_wait then reads back over tiles g-1, g-2, ... (lane l reads tile
// hi - l) until an inclusive prefix is found, publishes the tile's own inclusive prefix
// and returns the exclusive one.  Tiles are numbered by tickets drawn at block start, so
// every earlier tile's workgroup is already running and publishes its sum without waiting
// on anything; tiles before `first` count as 0.  Flags of another epoch (an earlier
// launch) read as status 0.
__device__ __forceinline__ void lookback_publish(unsigned long long* flags, int g, int first,
                                                 uint32_t epoch, uint32_t total, int lane) {
  if (lane == 0)
    __hip_atomic_store(&flags[g], pack_flag(epoch, g == first ? 2u : 1u, total),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t lookback_wait(unsigned long long* flags, int g, int first, uint32_t epoch,
                                  uint32_t total, int lane) {
  if (g == first) return 0u;
  uint32_t prefix = 0;
  int hi = g - 1;
  while (true) {
    const int j = hi - lane;
    uint32_t st = 2u, val = 0u;
    if (j >= first) {
      const unsigned long long w =
          __hip_atomic_load(&flags[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      st = ((uint32_t)(w >> 34) == epoch) ? ((uint32_t)(w >> 32) & 3u) : 0u;
      val = (uint32_t)w;
    }
    const unsigned long long incl = __ballot(st == 2u);
    const int stop = incl ? (__ffsll((long long)incl) - 1) : (kWave - 1);
    const unsigned long long need = (stop == kWave - 1) ? ~0ull : ((1ull << (stop + 1)) - 1ull);
    if (__ballot(st == 0u) & need) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t v = (lane <= stop) ? val : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, kWave);
    prefix += v;
    if (incl) break;
    hi -= kWave;
  }
  if (lane == 0)
    __hip_atomic_store(&flags[g], pack_flag(epoch, 2u, prefix + total), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

// Sums of `v` and `u` over the block, and the exclusive scan of `w` in thread order with
// its total, behind one pair of barriers.  `red` holds 12 words.
__device__ __forceinline__ void block_sums_excl_u32(uint32_t v, uint32_t u, uint32_t w,
                                                    uint32_t* red, uint32_t& vsum,
                                                    uint32_t& usum, uint32_t& wpre,
                                                    uint32_t& wtotal) {
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
  constexpr int kW = kSortThreads / kWave;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v += (uint32_t)__shfl_xor((int)v, o, kWave);
    u += (uint32_t)__shfl_xor((int)u, o, kWave);
  }
  const uint32_t incl = wave_inclusive_scan(w, lane);
  __syncthreads();
  if (lane == kWave - 1) {
    red[wv] = v;
    red[kW + wv] = u;
    red[2 * kW + wv] = incl;
  }
  __syncthreads();
  vsum = usum = wtotal = 0u;
  uint32_t pre = 0u;
#pragma unroll
  for (int k = 0; k < kW; ++k) {
    vsum += red[k];
    usum += red[kW + k];
    wtotal += red[2 * kW + k];
    if (k < wv) pre += red[2 * kW + k];
  }
  wpre = pre + incl - w;
}

// An edge tile's owners' edges.  The block draws a ticket (its edge tile, in the order the
// edge-tile blocks started), sums its owners' counts and finds the tile's place in the list
// by a decoupled look-back over the tiles before it; then it writes the tile's edges
// e = tid, tid + 256, ...: each finds its owner in LDS (the last thread whose first edge is
// at or before e, then that thread's owners in order) and its rank k = e - the owner's
// first edge, so a thread's staged-partner loads are all in flight together and the stores
// are contiguous.  `toff`, `p0s` (kSortThreads words) and `ecs` (kEdgeTile words) are LDS
// scratch.
__device__ void edge_tile(const TileSortArgs& a, uint32_t* red, uint32_t* toff, uint32_t* p0s,
                          uint32_t* ecs) {
  __shared__ uint32_t s_u, s_epoch, s_base;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  if (tid == 0) {
    const unsigned long long t = atomicAdd(a.eticket, 1ull);
    s_epoch = (uint32_t)((t / (unsigned long long)a.netiles) & 0x3FFFFFFFu);
    s_u = (uint32_t)(t % (unsigned long long)a.netiles);
  }
  // the owners' edge counts (a thread's kPer owners share one 64-slot group): loaded for the
  // tile of the block's own position while the ticket is in flight (blocks nearly always
  // start in order), and again in the rare case the ticket names another tile
  constexpr int kPer = kEdgeTile / kSortThreads;
  int u = (int)blockIdx.x - a.ntiles;
  long id0 = (long)u * kEdgeTile + (long)tid * kPer;
  uint32_t ec[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) ec[i] = (id0 + i < a.P) ? a.ecount[id0 + i] : 0u;
  __syncthreads();
  if ((int)s_u != u) {
    u = (int)s_u;
    id0 = (long)u * kEdgeTile + (long)tid * kPer;
#pragma unroll
    for (int i = 0; i < kPer; ++i) ec[i] = (id0 + i < a.P) ? a.ecount[id0 + i] : 0u;
  }
#pragma unroll
  for (int i = 0; i < kPer; ++i)  // the next pass writes only non-zero counts
    if (ec[i]) a.ecount[id0 + i] = 0u;
  // one simulation starts at slot 0; otherwise the owners' simulation's first slot
  const int p0 = (a.one_sim || id0 >= a.P) ? 0 : a.sims[a.wave_sim[id0 / kWave]].p0;
  uint32_t run = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) run += ec[i];
  uint32_t unused0, unused1, pre, total;
  block_sums_excl_u32(0u, 0u, run, red, unused0, unused1, pre, total);
  const uint32_t epoch = s_epoch;
  if (tid < kWave) lookback_publish(a.eflags, u, 0, epoch, total, lane);
  toff[tid] = pre;
  p0s[tid] = (uint32_t)p0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) ecs[tid * kPer + i] = ec[i];
  __syncthreads();
  // edges e0 + b * kSortThreads + tid, b < kB: the staged partners (loads issued together),
  // then the pairs stored at base + e
  constexpr int kB = 4;
  int own[kB];
  uint32_t ob[kB], jv[kB];
  auto gather = [&](uint32_t e0) {
    long src[kB];
#pragma unroll
    for (int b = 0; b < kB; ++b) {
      const uint32_t e = e0 + (uint32_t)(b * kSortThreads + tid);
      src[b] = 0;
      own[b] = 0;
      ob[b] = 0u;
      if (e < total) {
        int t = 0;  // toff[0] = 0 <= e; the steps reach at most index 255
#pragma unroll
        for (int st = kSortThreads / 2; st > 0; st >>= 1)
          if (toff[t + st] <= e) t += st;
        uint32_t r = e - toff[t];
        int i = 0;
#pragma unroll
        for (; i < kPer - 1; ++i) {
          const uint32_t c = ecs[t * kPer + i];
          if (r < c) break;
          r -= c;
        }
        const long id = (long)u * kEdgeTile + (long)t * kPer + i;
        src[b] = id * a.estride + r;
        own[b] = (int)(id - (long)p0s[t]);
        ob[b] = p0s[t];
      }
    }
#pragma unroll
    for (int b = 0; b < kB; ++b) {
      const uint32_t e = e0 + (uint32_t)(b * kSortThreads + tid);
      jv[b] = (e < total) ? a.estage[src[b]] : 0u;
    }
  };
  auto put = [&](uint32_t e0, uint32_t base) {
#pragma unroll
    for (int b = 0; b < kB; ++b) {
      const uint32_t e = e0 + (uint32_t)(b * kSortThreads + tid);
      if (e < total) a.edges[(long)base + e] = make_int2(own[b], (int)(jv[b] - ob[b]));
    }
  };
  // the first edges' loads go out before the look-back waits: only the stores need the base
  gather(0u);
  if (tid < kWave) {
    const uint32_t b = lookback_wait(a.eflags, u, 0, epoch, total, lane);
    if (lane == 0) s_base = b;
  }
  __syncthreads();
  const uint32_t base = s_base;
  if ((long)base + total > a.edge_room) {
    if (tid == 0) atomicOr(a.flag, 4u);
    return;
  }
  if (tid == 0 && u == a.netiles - 1) *a.etotal = base + total;  // the whole list's size
  put(0u, base);
  for (uint32_t e0 = kSortThreads * kB; e0 < total; e0 += kSortThreads * kB) {
    gather(e0);
    put(e0, base);
  }
}

__global__ void __launch_bounds__(kSortThreads) k_tile_sort(TileSortArgs a) {
  __shared__ uint32_t cnt_s[kTileCells];
  __shared__ uint32_t cofs[kTileCells + 1];
  __shared__ uint32_t fill_s[kTileCells];
  __shared__ uint32_t mslot[kTileCap];
  __shared__ uint32_t mid[kTileCap];
  __shared__ uint16_t mlc[kTileCap];
  __shared__ uint32_t aq[kTileCap];  // members in visit order (slot, index, local cell)
  __shared__ uint32_t aid[kTileCap];
  __shared__ uint16_t alc[kTileCap];
  __shared__ uint32_t red[3 * (kSortThreads / kWave)];
  __shared__ uint32_t wq0[6], wq1[6];  // the window's slot runs
  __shared__ int s_next;
  __shared__ uint32_t s_app;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  if ((int)blockIdx.x >= a.ntiles) {
    static_assert(kTileCells >= kEdgeTile && kTileCells >= kSortThreads, "edge tile scratch");
    edge_tile(a, red, cofs, fill_s, cnt_s);
    return;
  }
  // consecutive tiles (consecutive slots) share an XCD, as in the force pass
  const int t = xcd_map((int)blockIdx.x, a.ntiles);
  const SimDesc& sd = a.sims[a.one_sim ? 0 : (int)a.tile_sim[t]];
  if (tid == 0) {
    a.tcount_next[t * kCtrStride] = 0u;
    a.ftn_next[t * kCtrStride] = 0u;
    if (t == 0) *a.far_n_next = 0u;
  }
  const int lt = t - sd.t0, zt = lt / sd.nyt, y0 = (lt % sd.nyt) * sd.tr;
  const int nrow = min(sd.tr, sd.ny - y0), nc = sd.nx * nrow;
  const uint32_t cA = (uint32_t)(sd.c0 + sd.nx * (y0 + sd.ny * zt));

  // Window: planes zt-1, zt, zt+1 and rows y0-1 .. y0+nrow, cyclic and distinct; each
  // plane's rows are one or two runs of consecutive cells, so of consecutive slots.  The
  // runs' bounds are loaded first: they depend on nothing the counts below decide.
  const int np = min(3, sd.nz), nrw = min(nrow + 2, sd.ny);
  const int ys = (y0 - 1 + sd.ny) % sd.ny, yrun0 = min(nrw, sd.ny - ys);
  if (tid < 6) {
    const int ip = tid >> 1, h = tid & 1;
    const int ra = h ? 0 : ys, rn = h ? nrw - yrun0 : yrun0;
    uint32_t q0 = 0u, q1 = 0u;
    if (ip < np && rn > 0) {
      const int zp = (zt - 1 + ip + sd.nz) % sd.nz;
      const int cf = sd.c0 + sd.nx * (ra + sd.ny * zp);
      q0 = a.start_old[cf];
      q1 = a.start_old[cf + sd.nx * rn];
    }
    wq0[tid] = q0;
    wq1[tid] = q1;
  }

  // the far lists' lengths do not depend on anything below either
  const uint32_t nfar_t = min(a.ftn[t * kCtrStride], (uint32_t)kFarCap);
  const uint32_t nfar = min(*a.far_n, (uint32_t)a.P);

  // fn(valid, q, key, id) for every slot of the window and of the far lists, called by
  // every lane of the block the same number of times (valid: a slot whose key is not a far
  // mover's; far movers come from the lists, key without the far bit).  The index is
  // loaded with the key: most slots are not members, but a dependent load per member costs
  // more.  The window's six runs are taken as one sequence, kVisitBatch slots per thread
  // with all their loads in flight before any is used; fn's order does not matter (ranks
  // are by index).
  auto visit = [&](auto&& fn) {
    uint32_t wp[7];
    wp[0] = 0u;
#pragma unroll
    for (int g = 0; g < 6; ++g) wp[g + 1] = wp[g] + (wq1[g] - wq0[g]);
    const uint32_t nw = wp[6];
    for (uint32_t b = 0; b < nw; b += kVisitBatch * kSortThreads) {
      uint32_t qv[kVisitBatch], kv[kVisitBatch], iv[kVisitBatch];
#pragma unroll
      for (int u = 0; u < kVisitBatch; ++u) {
        const uint32_t i = b + (uint32_t)(u * kSortThreads + tid);
        qv[u] = 0u;
        kv[u] = kFarBit;  // nothing to visit
        iv[u] = 0u;
        if (i < nw) {
          uint32_t q = wq0[0] + i;
#pragma unroll
          for (int g = 1; g < 6; ++g)
            if (i >= wp[g]) q = wq0[g] + (i - wp[g]);  // an empty run gives way to the next
          qv[u] = q;
          kv[u] = a.key[q];
          iv[u] = a.src_id[q];
        }
      }
#pragma unroll
      for (int u = 0; u < kVisitBatch; ++u) fn(!(kv[u] & kFarBit), qv[u], kv[u], iv[u]);
    }
    for (uint32_t j0 = 0; j0 < nfar_t; j0 += kSortThreads) {
      const uint32_t j = j0 + (uint32_t)tid;
      uint32_t q = 0u, k = 0u, id = 0u;
      if (j < nfar_t) {
        q = a.ftl[(long)t * kFarCap + j];
        k = a.key[q] & ~kFarBit;
        id = a.src_id[q];
      }
      fn(j < nfar_t, q, k, id);
    }
    for (uint32_t j0 = 0; j0 < nfar; j0 += kSortThreads) {
      const uint32_t j = j0 + (uint32_t)tid;
      uint32_t q = 0u, k = 0u, id = 0u;
      if (j < nfar) {
        q = a.far_slot[j];
        k = a.key[q] & ~kFarBit;
        id = a.src_id[q];
      }
      fn(j < nfar, q, k, id);
    }
  };

  // The tile's cell counts come from its members.  Normally (the force pass counted at most
  // kTileCap particles into the tile) one visit appends the members to LDS, a wave's
  // members at one LDS atomic, and counts them per cell; else a visit only counts, and the
  // cells are then sorted a chunk at a time below.
  uint32_t part = 0;  // the particles of the simulation's tiles before this one
  for (int u = sd.t0 + tid; u < t; u += kSortThreads) part += a.tcount[u * kCtrStride];
  const uint32_t expect = a.tcount[t * kCtrStride];
  for (int i = tid; i < nc; i += kSortThreads) {
    cnt_s[i] = 0u;
    fill_s[i] = 0u;
  }
  if (tid == 0) s_app = 0u;
  __syncthreads();
  const bool placed = expect <= (uint32_t)kTileCap;  // the same for the whole block
  if (placed) {
    visit([&](bool valid, uint32_t q, uint32_t k, uint32_t id) {
      const bool mem = valid && k >= cA && k - cA < (uint32_t)nc;
      const unsigned long long bal = __ballot(mem);
      if (bal == 0ull) return;
      const int lead = __ffsll((long long)bal) - 1;
      uint32_t b0 = 0u;
      if (lane == lead) b0 = atomicAdd(&s_app, (uint32_t)__popcll(bal));
      b0 = (uint32_t)__shfl((int)b0, lead, kWave);
      if (mem) {
        const uint32_t j = b0 + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        const uint32_t lc = k - cA;
        atomicAdd(&cnt_s[lc], 1u);
        if (j < (uint32_t)kTileCap) {
          aq[j] = q;
          aid[j] = id;
          alc[j] = (uint16_t)lc;
        }
      }
    });
  } else {
    visit([&](bool valid, uint32_t, uint32_t k, uint32_t) {
      if (valid && k >= cA && k - cA < (uint32_t)nc) atomicAdd(&cnt_s[k - cA], 1u);
    });
  }
  __syncthreads();

  // A tile of at most kMoveBatch * kSortThreads appended members loads their states now,
  // entry m = u * kSortThreads + tid by thread tid: the loads are in flight during the
  // scan and the grouping below, and the same thread then moves them.
  const uint32_t napp = placed ? s_app : 0u;
  const bool preload = placed && napp <= (uint32_t)(kMoveBatch * kSortThreads);
  float4 psp[kMoveBatch], psv[kMoveBatch];
#pragma unroll
  for (int u = 0; u < kMoveBatch; ++u) {
    const uint32_t m = (uint32_t)(u * kSortThreads + tid);
    psp[u] = psv[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (preload && m < napp) {
      psp[u] = a.src_pos[aq[m]];
      psv[u] = a.src_vel[aq[m]];
    }
  }

  // first slot of the tile, then its cells' starts
  const int per = (nc + kSortThreads - 1) / kSortThreads;
  const int i0 = min(tid * per, nc), i1 = min(i0 + per, nc);
  uint32_t run = 0;
  for (int i = i0; i < i1; ++i) run += cnt_s[i];
  uint32_t before, unused, pre, M;
  block_sums_excl_u32(part, 0u, run, red, before, unused, pre, M);
  const uint32_t base = (uint32_t)sd.p0 + before;
  // The members must be what the force pass counted into the tile (and the simulation's
  // last tile must end at its last particle); if not, the tile writes nothing but clamped
  // starts, and the call fails.
  const uint32_t lim = (uint32_t)(sd.p0 + sd.n);
  const bool last = lt == sd.nz * sd.nyt - 1;
  const bool ok = M == expect && base + M <= lim && (!last || base + M == lim);
  for (int i = i0; i < i1; ++i) {
    cofs[i] = pre;
    a.start_new[cA + i] = min(base + pre, lim);
    pre += cnt_s[i];
  }
  if (tid == 0) {
    cofs[nc] = M;
    if (!ok) atomicOr(a.flag, 4u);
    if (last) {  // the simulation's pad cell, and the start of what follows it
      a.start_new[sd.c0 + sd.ncells] = lim;
      a.start_new[sd.c0 + sd.ncells + 1] =
          (uint32_t)(sd.p0 + (sd.n + kSlotAlign - 1) / kSlotAlign * kSlotAlign);
    }
  }
  __syncthreads();
  if (!ok) return;
  if (placed) {  // group the appended members by cell (M == expect <= kTileCap: all kept)
    for (uint32_t m = tid; m < M; m += kSortThreads) {
      const uint32_t lc = alc[m];
      const uint32_t at = cofs[lc] + atomicAdd(&fill_s[lc], 1u);
      mslot[at] = aq[m];
      mid[at] = aid[m];
      mlc[at] = (uint16_t)lc;
    }
  }
  if (preload) {  // rank each preloaded member by index inside its cell and move it
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kMoveBatch; ++u) {
      const uint32_t m = (uint32_t)(u * kSortThreads + tid);
      if (m < M) {
        const uint32_t lc = alc[m], me = aid[m];
        const uint32_t s0 = cofs[lc], sn = cnt_s[lc];
        uint32_t rank = 0;
        for (uint32_t e = 0; e < sn; ++e) rank += (mid[s0 + e] < me) ? 1u : 0u;
        const uint32_t dst = base + s0 + rank;
        a.dst_pos[dst] = psp[u];
        a.dst_vel[dst] = psv[u];
        a.dst_id[dst] = me;
      }
    }
    return;
  }

  // cells [ia, ib) at a time, as many as kTileCap members allow (all of them, normally)
  int ia = 0;
  while (ia < nc) {
    int ib = nc;  // every member fits (the normal case): no search
    if (M - cofs[ia] > (uint32_t)kTileCap) {
      if (tid == 0) {
        int lo = ia + 1, hi = nc;
        while (lo < hi) {
          const int m = (lo + hi + 1) >> 1;
          if (cofs[m] - cofs[ia] <= (uint32_t)kTileCap) lo = m;
          else hi = m - 1;
        }
        s_next = lo;
      }
      __syncthreads();
      ib = s_next;
    }
    const uint32_t m0 = cofs[ia], mn = cofs[ib] - m0;
    if (mn <= (uint32_t)kTileCap) {
      if (!placed) {
        for (int i = ia + tid; i < ib; i += kSortThreads) fill_s[i] = 0u;
        for (uint32_t m = tid; m < mn; m += kSortThreads) mlc[m] = 0xFFFFu;
        __syncthreads();
        const uint32_t ca = cA + (uint32_t)ia, cb = cA + (uint32_t)ib;
        visit([&](bool valid, uint32_t q, uint32_t k, uint32_t id) {
          if (!valid || k < ca || k >= cb) return;
          const uint32_t lc = k - cA;
          const uint32_t at = atomicAdd(&fill_s[lc], 1u);
          if (at >= cnt_s[lc]) {
            atomicOr(a.flag, 4u);
            return;
          }
          const uint32_t m = cofs[lc] - m0 + at;
          mslot[m] = q;
          mid[m] = id;
          mlc[m] = (uint16_t)lc;
        });
      }
      __syncthreads();
      // kMoveBatch members per thread: their state loads are in flight during the ranking
      for (uint32_t mb = 0; mb < mn; mb += kMoveBatch * kSortThreads) {
        float4 sp[kMoveBatch], sv[kMoveBatch];
        uint32_t lcv[kMoveBatch];
#pragma unroll
        for (int u = 0; u < kMoveBatch; ++u) {
          const uint32_t m = mb + (uint32_t)(u * kSortThreads + tid);
          lcv[u] = 0xFFFFu;
          sp[u] = sv[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (m < mn) {
            const uint32_t lc = mlc[m];
            if (lc >= (uint32_t)nc || fill_s[lc] != cnt_s[lc]) {
              atomicOr(a.flag, 4u);  // a cell lost a particle: leave its slots alone
            } else {
              const uint32_t q = mslot[m];
              sp[u] = a.src_pos[q];
              sv[u] = a.src_vel[q];
              lcv[u] = lc;
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kMoveBatch; ++u) {
          const uint32_t lc = lcv[u];
          if (lc != 0xFFFFu) {
            const uint32_t m = mb + (uint32_t)(u * kSortThreads + tid);
            const uint32_t s0 = cofs[lc] - m0, sn = cnt_s[lc], me = mid[m];
            uint32_t rank = 0;
            for (uint32_t e = 0; e < sn; ++e) rank += (mid[s0 + e] < me) ? 1u : 0u;
            const uint32_t dst = base + cofs[lc] + rank;
            a.dst_pos[dst] = sp[u];
            a.dst_vel[dst] = sv[u];
            a.dst_id[dst] = me;
          }
        }
      }
    } else {
      // one cell (ib == ia + 1) holds more than the LDS takes: its ids and slots go to
      // global scratch at the cell's own slot range, ranks by a pass over all of them
      const uint32_t K = cnt_s[ia], g0 = base + cofs[ia], cg = cA + (uint32_t)ia;
      if (tid == 0) {
        fill_s[ia] = 0u;
        atomicOr(a.flag, 2u);
      }
      __syncthreads();
      visit([&](bool valid, uint32_t q, uint32_t k, uint32_t id) {
        if (!valid || k != cg) return;
        const uint32_t at = atomicAdd(&fill_s[ia], 1u);
        if (at >= K) {
          atomicOr(a.flag, 4u);
          return;
        }
        a.tmp_id[g0 + at] = id;
        a.tmp_slot[g0 + at] = q;
      });
      __syncthreads();
      if (fill_s[ia] == K) {
        for (uint32_t m = tid; m < K; m += kSortThreads) {
          const uint32_t me = a.tmp_id[g0 + m];
          uint32_t rank = 0;
          for (uint32_t e = 0; e < K; ++e) rank += (a.tmp_id[g0 + e] < me) ? 1u : 0u;
          const uint32_t q = a.tmp_slot[g0 + m];
          a.dst_pos[g0 + rank] = a.src_pos[q];
          a.dst_vel[g0 + rank] = a.src_vel[q];
          a.dst_id[g0 + rank] = me;
        }
      } else if (tid == 0) {
        atomicOr(a.flag, 4u);
      }
    }
    __syncthreads();  // the LDS lists are reused by the next cells
    ia = ib;
  }
}

// ------------------------------------------------------------------------------------
// k_scan: single-pass decoupled look-back exclusive scan over up to two arrays.

__global__ void __launch_bounds__(kScanThreads) k_scan(ScanPlan plan) {
  __shared__ uint32_t smem[kScanTile + kScanTile / 32];
  __shared__ uint32_t wsum[kScanThreads / kWave];
  __shared__ uint32_t s_tile, s_epoch, s_prefix;
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wv = tid >> 6;
  if (tid == 0) {
    const unsigned long long t = atomicAdd(plan.ticket, 1ull);
    s_epoch = (uint32_t)((t / (unsigned long long)plan.tiles_total) & 0x3FFFFFFFu);
    s_tile = (uint32_t)(t % (unsigned long long)plan.tiles_total);
  }
  __syncthreads();
  const uint32_t g = s_tile;
  const uint32_t epoch = s_epoch;
  const int seg = (plan.nseg > 1 && (int)g >= plan.seg[0].tiles) ? 1 : 0;
  const int tile = (int)g - (seg ? plan.seg[0].tiles : 0);
  const uint32_t* in = plan.seg[seg].in;
  uint32_t* out = plan.seg[seg].out;
  const int n = plan.seg[seg].n;
  const long base = (long)tile * kScanTile;

#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int e = i * kScanThreads + tid;
    const long gi = base + e;
    smem[e + (e >> 5)] = (gi < n) ? in[gi] : 0u;
  }
  __syncthreads();
  uint32_t items[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int e = tid * kScanItems + k;
    items[k] = smem[e + (e >> 5)];
    sum += items[k];
  }
  const uint32_t incl = wave_inclusive_scan(sum, lane);
  if (lane == kWave - 1) wsum[wv] = incl;
  __syncthreads();
  uint32_t wpre = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / kWave; ++w) {
    const uint32_t v = wsum[w];
    if (w < wv) wpre += v;
    total += v;
  }
  const uint32_t texcl = wpre + incl - sum;

  if (wv == 0) {
    lookback_publish(plan.flags, (int)g, (int)g - tile, epoch, total, lane);
    const uint32_t prefix = lookback_wait(plan.flags, (int)g, (int)g - tile, epoch, total, lane);
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  uint32_t run = s_prefix + texcl;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int e = tid * kScanItems + k;
    smem[e + (e >> 5)] = run;
    run += items[k];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const int e = i * kScanThreads + tid;
    const long gi = base + e;
    if (gi < n) out[gi] = smem[e + (e >> 5)];
  }
  if (tid == 0 && tile == plan.seg[seg].tiles - 1) out[n] = s_prefix + total;
}

// ------------------------------------------------------------------------------------
// k_fill / k_rank_scatter: the two halves of the stable counting sort.


__global__ void __launch_bounds__(256) k_fill(SortArgs a) {
  const int gw = xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (a.edges && gw == 0 && threadIdx.x % kWave == 0) *a.etotal = a.eoff[a.P];
  if (gw >= a.nwaves) return;
  const SimDesc sd = a.sims[a.wave_sim[gw]];
  const int p = gw * kWave + (threadIdx.x & (kWave - 1));
  if (p - sd.p0 >= sd.n) return;
  const uint32_t c = a.key[p];
  const uint32_t me = a.src_id[p];
  const uint32_t arr = a.arrival[p];
  a.tmp_id[a.start[c] + arr] = me;
  a.hist[c] = 0u;  // every writer stores 0; pad cells are never keys and keep their size
  // canonical edge list of the pre-step state (offsets from the same scan as the cells)
  if (a.edges) {
    const uint32_t off = a.eoff[me];
    const int e = (int)(a.eoff[me + 1] - off);
    if (e) a.fcount[me] = 0u;  // the next step's force pass writes only non-zero counts
    for (int k = 0; k < e; ++k)
      a.edges[off + k] =
          make_int2((int)(me - sd.p0), (int)(a.estage[(long)p * a.estride + k] - sd.p0));
  }
}

__global__ void __launch_bounds__(256) k_rank_scatter(SortArgs a) {
  const int gw = xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gw >= a.nwaves) return;
  const SimDesc sd = a.sims[a.wave_sim[gw]];
  const int p = gw * kWave + (threadIdx.x & (kWave - 1));
  if (p - sd.p0 >= sd.n) return;
  const uint32_t c = a.key[p];
  const uint32_t me = a.src_id[p];
  const float4 sp = a.src_pos[p];
  const float4 sv = a.src_vel[p];
  const uint32_t s0 = a.start[c], s1 = a.start[c + 1];
  uint32_t rank = 0;
  if (sd.ident) {
    rank = a.arrival[p];  // source already in index order
  } else {
    uint32_t q = s0;
    for (; q + 4 <= s1; q += 4) {  // four independent loads per trip
      const uint32_t i0 = a.tmp_id[q], i1 = a.tmp_id[q + 1], i2 = a.tmp_id[q + 2],
                     i3 = a.tmp_id[q + 3];
      rank += (i0 < me) + (i1 < me) + (i2 < me) + (i3 < me);
    }
    for (; q < s1; ++q) rank += (a.tmp_id[q] < me) ? 1u : 0u;
  }
  const uint32_t dst = s0 + rank;
  a.dst_pos[dst] = sp;
  a.dst_vel[dst] = sv;
  a.dst_id[dst] = me;
}

// ------------------------------------------------------------------------------------
// binning of a given order (load, grid change)


__global__ void __launch_bounds__(256) k_bin(BinArgs a) {
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gw >= a.nwaves) return;
  const SimDesc sd = a.sims[a.wave_sim[gw]];
  const int p = gw * kWave + (threadIdx.x & (kWave - 1));
  if (p - sd.p0 >= sd.n) return;
  if (sd.ident) {  // one cell in index order: no atomics, the count is known
    a.key[p] = (uint32_t)sd.c0;
    a.arrival[p] = (uint32_t)(p - sd.p0);
    if (p == sd.p0) a.hist[sd.c0] = (uint32_t)sd.n;
    return;
  }
  const float4 pi = a.pos[p];
  const uint32_t c = (uint32_t)(sd.c0 + bin_axis(pi.x, sd.ihx, sd.nx) +
                                sd.nx * (bin_axis(pi.y, sd.ihy, sd.ny) +
                                         sd.ny * bin_axis(pi.z, sd.ihz, sd.nz)));
  a.key[p] = c;
  a.arrival[p] = atomicAdd(&a.hist[c], 1u);
}

// ------------------------------------------------------------------------------------
// state load / unload in the reference's 44-byte AoS layout (index order)



// Index-order AoS -> slot-order temporaries of one simulation (slot = p0 + index).
__global__ void k_load_aos(IoArgs a) {
  const SimDesc sd = a.sims[a.sim];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sd.n) return;
  const AosParticle r = a.aos[i];
  const int p = sd.p0 + i;
  const int prev = a.counts_in ? a.counts_in[i] : 0;
  a.tpos[p] = make_float4(r.px, r.py, r.pz, __uint_as_float(r.ptype));
  a.tvel[p] = make_float4(r.vx, r.vy, r.vz, __int_as_float(prev));
  a.tid[p] = (uint32_t)p;
  a.acc[p] = make_float4(r.ax, r.ay, r.az, 0.0f);
  a.pad[p] = r.pad;
}

// Universe shift of one simulation's loaded temporaries, moveParticlesKernel
// (ParticleSimulation.cu:169-185): x = fmodf(x + dx + L, L) per axis in FP32, left to right.
__global__ void k_move(IoArgs a) {
  const SimDesc sd = a.sims[a.sim];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  float4 q = a.tpos[p];
  q.x = fmodf(q.x + a.shift[0] + a.box[0], a.box[0]);
  q.y = fmodf(q.y + a.shift[1] + a.box[1], a.box[1]);
  q.z = fmodf(q.z + a.shift[2] + a.box[2], a.box[2]);
  a.tpos[p] = q;
}

// Sorted state -> index-order AoS of one simulation.
__global__ void k_unload_aos(IoArgs a) {
  const SimDesc sd = a.sims[a.sim];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  const float4 ps = a.pos[p];
  const float4 vs = a.vel[p];
  const uint32_t g = a.id[p];
  const float4 ac = a.acc[g];
  AosParticle r;
  r.px = ps.x; r.py = ps.y; r.pz = ps.z;
  r.vx = vs.x; r.vy = vs.y; r.vz = vs.z;
  r.ax = ac.x; r.ay = ac.y; r.az = ac.z;
  r.ptype = __float_as_uint(ps.w);
  r.pad = a.pad[g];
  a.aos[g - sd.p0] = r;
}

__global__ void k_get_counts(IoArgs a) {
  const SimDesc sd = a.sims[a.sim];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  a.counts_out[a.id[p] - sd.p0] = __float_as_int(a.vel[p].w);
}

__global__ void k_set_counts(IoArgs a) {
  const SimDesc sd = a.sims[a.sim];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  float4 v = a.vel[p];
  v.w = __int_as_float(a.counts_in[a.id[p] - sd.p0]);
  a.tvel[p] = v;  // the caller passes the current velocity buffer as tvel
}

__global__ void k_sort_order(IoArgs a) {
  const SimDesc sd = a.sims[a.sim];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  const float4 ps = a.pos[p];
  a.cell_out[k] = (uint32_t)(bin_axis(ps.x, sd.ihx, sd.nx) +
                             sd.nx * (bin_axis(ps.y, sd.ihy, sd.ny) +
                                      sd.ny * bin_axis(ps.z, sd.ihz, sd.nz)));
  a.index_out[k] = a.id[p] - sd.p0;
}

// ------------------------------------------------------------------------------------
// seeded initial state (replacement for ParticleSimulation.cu:21-66)

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float u24(unsigned long long seed, unsigned long long stream,
                                     unsigned long long idx) {
  const unsigned long long key = seed * 0x632BE59BD9B4E019ull + stream * 0x85EBCA77C2B2AE63ull;
  const unsigned long long base = mix64(key + 0x9E3779B97F4A7C15ull);
  const unsigned long long z = mix64(base + (idx + 1ull) * 0x9E3779B97F4A7C15ull);
  return (float)(uint32_t)(z >> 40) * 5.9604644775390625e-8f;  // 2^-24
}


__global__ void k_init(InitArgs a) {
  const SimDesc sd = a.sims[a.sim];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sd.n) return;
  const float sw = fminf(a.spawn, a.cw), sh = fminf(a.spawn, a.ch), sdp = fminf(a.spawn, a.cd);
  const float ox = (a.cw - sw) * 0.5f, oy = (a.ch - sh) * 0.5f, oz = (a.cd - sdp) * 0.5f;
  // stream 1: positions (3 per particle), stream 2: type — as cellflow_amd.configs
  const unsigned long long i3 = 3ull * (unsigned long long)i;
  const float x = ox + u24(a.seed, 1, i3) * sw;
  const float y = oy + u24(a.seed, 1, i3 + 1) * sh;
  const float z = oz + u24(a.seed, 1, i3 + 2) * sdp;
  uint32_t t = (uint32_t)(u24(a.seed, 2, (unsigned long long)i) * (float)sd.T);
  if (t >= (uint32_t)sd.T) t = (uint32_t)sd.T - 1;
  const int p = sd.p0 + i;
  a.tpos[p] = make_float4(x, y, z, __uint_as_float(t));
  a.tvel[p] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
  a.tid[p] = (uint32_t)p;
  a.acc[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  a.pad[p] = 0.0f;
}

// ------------------------------------------------------------------------------------
// line vertices (ParticleSimulation.cu:253-275 layout) from the canonical edge list


__global__ void k_slot_of(const uint32_t* __restrict__ id, uint32_t* slot_of, const SimDesc* sims,
                          int sim) {
  const SimDesc sd = sims[sim];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= sd.n) return;
  const int p = sd.p0 + k;
  slot_of[id[p] - sd.p0] = (uint32_t)p;
}

__global__ void k_line_vertices(LineArgs a) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.nedges) return;
  const int2 ij = a.edges[e];
  const float4 pi = a.pos[a.slot_of[ij.x]];
  const float4 pj = a.pos[a.slot_of[ij.y]];
  const uint32_t ci = __float_as_uint(pi.w) % (uint32_t)a.T;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if ((int)ci < a.ncolors) {
    r = a.colors[3 * ci];
    g = a.colors[3 * ci + 1];
    b = a.colors[3 * ci + 2];
  }
  float* o = a.out + e * 12;
  o[0] = pi.x; o[1] = pi.y; o[2] = pi.z; o[3] = r; o[4] = g; o[5] = b;
  o[6] = pj.x; o[7] = pj.y; o[8] = pj.z; o[9] = r; o[10] = g; o[11] = b;
}

// ------------------------------------------------------------------------------------
// Triangle mesh (generateTriangleMeshKernel, ParticleSimulation.cu:279-424), in canonical
// order instead of the reference's atomicAdd order (:353).  Per owner i: its neighbours are
// the first maxc particles j != i of its type, in index order, whose non-periodic squared
// distance is below d2 (:306-324), that is the maxc smallest such indices; its triangles are
// (i, a, b) for the pairs a < b of that list with i < a, in list order, at most 64
// (:334-348).  The owners are visited by slot over a non-periodic stencil.

__global__ void __launch_bounds__(256) k_tri_count(TriArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gw >= a.nwaves) return;
  const SimDesc sd = a.sims[a.wave_sim[gw]];
  const int p = gw * kWave + lane;
  if (p - sd.p0 >= sd.n) return;
  const float4 pi = a.pos[p];
  const uint32_t idi = a.id[p];
  const uint32_t ti = __float_as_uint(pi.w);
  const int cx = bin_axis(pi.x, sd.ihx, sd.nx);
  const int cy = bin_axis(pi.y, sd.ihy, sd.ny);
  const int cz = bin_axis(pi.z, sd.ihz, sd.nz);
  const int xlo = max(cx - a.rx, 0), xhi = min(cx + a.rx, sd.nx - 1);
  const int ylo = max(cy - a.ry, 0), yhi = min(cy + a.ry, sd.ny - 1);
  const int zlo = max(cz - a.rz, 0), zhi = min(cz + a.rz, sd.nz - 1);
  EdgeAcc ea;
  ea.q = 0;
  ea.maxid = 0;
  ea.maxk = 0;
  for (int z = zlo; z <= zhi; ++z)
    for (int y = ylo; y <= yhi; ++y) {
      const int rowc = sd.c0 + sd.nx * (y + sd.ny * z);
      const int q0 = (int)a.start[rowc + xlo];
      const int q1 = (int)a.start[rowc + xhi + 1];
      for (int q = q0; q < q1; ++q) {
        if (q == p) continue;                             // :307
        const float4 pj = a.pos[q];
        if (__float_as_uint(pj.w) != ti) continue;       // :311
        const float dx = pj.x - pi.x;                     // :314-317
        const float dy = pj.y - pi.y;
        const float dz = pj.z - pi.z;
        const float s2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        if (!(s2 < a.d2)) continue;                       // :319
        edge_push(ea, s2, a.id[q], a.maxc, a.cand_d2, a.cand_id, a.P, p);
      }
    }
  // the kept indices in ascending order (distinct, so rank by counting), and the owner's
  // triangles: pairs among the kept indices above its own
  const int m = min(ea.q, a.maxc);
  int above = 0;
  for (int k = 0; k < m; ++k) {
    const uint32_t ik = a.cand_id[(long)k * a.P + p];
    int rank = 0;
    for (int k2 = 0; k2 < m; ++k2) rank += (a.cand_id[(long)k2 * a.P + p] < ik) ? 1 : 0;
    a.sorted[(long)rank * a.P + p] = ik;
    above += (ik > idi) ? 1 : 0;
  }
  a.nsel[p] = (uint32_t)m;
  a.count[idi] = (uint32_t)min(64, above * (above - 1) / 2);
}

__global__ void __launch_bounds__(256) k_tri_write(TriArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (gw >= a.nwaves) return;
  const SimDesc sd = a.sims[a.wave_sim[gw]];
  const int p = gw * kWave + lane;
  if (p - sd.p0 >= sd.n) return;
  const uint32_t idi = a.id[p];
  const uint32_t nt = a.count[idi];
  if (nt == 0u) return;
  const int m = (int)a.nsel[p];
  int s0 = 0;  // the first kept index above the owner's
  while (s0 < m && a.sorted[(long)s0 * a.P + p] <= idi) ++s0;
  long w = (long)a.off[idi];
  uint32_t t = 0;
  for (int i1 = s0; i1 < m && t < nt; ++i1) {
    const uint32_t n1 = a.sorted[(long)i1 * a.P + p];
    for (int i2 = i1 + 1; i2 < m && t < nt; ++i2, ++t, ++w) {
      a.tris[3 * w] = (int32_t)idi - sd.p0;
      a.tris[3 * w + 1] = (int32_t)n1 - sd.p0;
      a.tris[3 * w + 2] = (int32_t)a.sorted[(long)i2 * a.P + p] - sd.p0;
    }
  }
}

// The reference's vertex layout (:362-418): per triangle three vertices of position,
// flat normal (normalised cross product of the two edges from the first vertex, left as
// is when its length is at most 1e-4) and the owner type's colour.
__global__ void k_tri_vertices(TriVertArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntris) return;
  const float4 v1 = a.pos[a.slot_of[a.tris[3 * t]]];  // local indices, like slot_of's
  const float4 v2 = a.pos[a.slot_of[a.tris[3 * t + 1]]];
  const float4 v3 = a.pos[a.slot_of[a.tris[3 * t + 2]]];
  const float e1x = v2.x - v1.x, e1y = v2.y - v1.y, e1z = v2.z - v1.z;
  const float e2x = v3.x - v1.x, e2y = v3.y - v1.y, e2z = v3.z - v1.z;
  float nx = e1y * e2z - e1z * e2y;
  float ny = e1z * e2x - e1x * e2z;
  float nz = e1x * e2y - e1y * e2x;
  const float len = sqrtf(nx * nx + ny * ny + nz * nz);
  if (len > 0.0001f) {
    nx /= len;
    ny /= len;
    nz /= len;
  }
  const uint32_t ci = __float_as_uint(v1.w) % (uint32_t)a.T;
  float r = 0.0f, g = 0.0f, b = 0.0f;
  if ((int)ci < a.ncolors) {
    r = a.colors[3 * ci];
    g = a.colors[3 * ci + 1];
    b = a.colors[3 * ci + 2];
  }
  float* o = a.out + t * 27;
  const float4 v[3] = {v1, v2, v3};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    o[9 * k + 0] = v[k].x;
    o[9 * k + 1] = v[k].y;
    o[9 * k + 2] = v[k].z;
    o[9 * k + 3] = nx;
    o[9 * k + 4] = ny;
    o[9 * k + 5] = nz;
    o[9 * k + 6] = r;
    o[9 * k + 7] = g;
    o[9 * k + 8] = b;
  }
}

// ------------------------------------------------------------------------------------
// host-side launchers

static inline int grid_for(int nwaves, int block) { return (nwaves * kWave + block - 1) / block; }

template <int R, int G>
static void launch_grp(const ForceArgs& a, bool edges, hipStream_t st) {
  const long waves = (long)a.nwaves * G;  // one wave per 64/G particle slots
  const dim3 g((unsigned)((waves + 3) / 4)), b(256);
  if (a.hits) {  // the gradient rollout's lists (never with edges)
    if (a.tile_sort) CF_LAUNCH((k_force_grp<false, R, G, true, true>), g, b, 0, st, a);
    else CF_LAUNCH((k_force_grp<false, R, G, false, true>), g, b, 0, st, a);
    return;
  }
  if (a.tile_sort) {
    if (edges) CF_LAUNCH((k_force_grp<true, R, G, true>), g, b, 0, st, a);
    else CF_LAUNCH((k_force_grp<false, R, G, true>), g, b, 0, st, a);
  } else {
    if (edges) CF_LAUNCH((k_force_grp<true, R, G, false>), g, b, 0, st, a);
    else CF_LAUNCH((k_force_grp<false, R, G, false>), g, b, 0, st, a);
  }
}

template <int R, int G>
static void launch_list(const ForceArgs& a, bool edges, hipStream_t st) {
  const long waves = (long)a.nwaves * G;  // one wave per 64/G particle slots
  const dim3 g((unsigned)((waves + 3) / 4)), b(256);
  if (a.tile_sort) {
    if (edges) CF_LAUNCH((k_force_list<true, R, G, true>), g, b, 0, st, a);
    else CF_LAUNCH((k_force_list<false, R, G, true>), g, b, 0, st, a);
  } else {
    if (edges) CF_LAUNCH((k_force_list<true, R, G, false>), g, b, 0, st, a);
    else CF_LAUNCH((k_force_list<false, R, G, false>), g, b, 0, st, a);
  }
}

template <int R, int G>
static void launch_cons(const ForceArgs& a, bool edges, hipStream_t st) {
  const long waves = (long)a.nwaves * G;  // one wave per 64/G particle slots
  const dim3 g((unsigned)((waves + 3) / 4)), b(256);
  if (a.tile_sort) {
    if (edges) CF_LAUNCH((k_force_list<true, R, G, true, true>), g, b, 0, st, a);
    else CF_LAUNCH((k_force_list<false, R, G, true, true>), g, b, 0, st, a);
  } else {
    if (edges) CF_LAUNCH((k_force_list<true, R, G, false, true>), g, b, 0, st, a);
    else CF_LAUNCH((k_force_list<false, R, G, false, true>), g, b, 0, st, a);
  }
}

hipError_t launch_force(const ForceArgs& a, bool edges, int block, int variant, int reach,
                        int lanes, hipStream_t st) {
  if (a.hits && (variant != 0 || edges)) return hipErrorInvalidValue;  // lists: variant 0 only
  if (variant == 0) {
    if (reach == 1) {
      if (lanes == 4) launch_grp<1, 4>(a, edges, st);
      else if (lanes == 8) launch_grp<1, 8>(a, edges, st);
      else if (lanes == 16) launch_grp<1, 16>(a, edges, st);
      else if (lanes == 32) launch_grp<1, 32>(a, edges, st);
      else return hipErrorInvalidValue;
    } else {
      if (lanes == 4) launch_grp<2, 4>(a, edges, st);
      else if (lanes == 8) launch_grp<2, 8>(a, edges, st);
      else if (lanes == 16) launch_grp<2, 16>(a, edges, st);
      else if (lanes == 32) launch_grp<2, 32>(a, edges, st);
      else return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if ((variant == 3 || variant == 4) && a.P > (1L << kCodeShift)) return hipErrorInvalidValue;
  if (variant == 4) {  // variant 3 with the block-consolidated tail
    if (reach == 1) {
      if (lanes == 8) launch_cons<1, 8>(a, edges, st);
      else if (lanes == 16) launch_cons<1, 16>(a, edges, st);
      else return hipErrorInvalidValue;
    } else {
      if (lanes == 8) launch_cons<2, 8>(a, edges, st);
      else if (lanes == 16) launch_cons<2, 16>(a, edges, st);
      else return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (variant == 3) {
    if (reach == 1) {
      if (lanes == 4) launch_list<1, 4>(a, edges, st);
      else if (lanes == 8) launch_list<1, 8>(a, edges, st);
      else if (lanes == 16) launch_list<1, 16>(a, edges, st);
      else if (lanes == 32) launch_list<1, 32>(a, edges, st);
      else return hipErrorInvalidValue;
    } else {
      if (lanes == 8) launch_list<2, 8>(a, edges, st);
      else if (lanes == 16) launch_list<2, 16>(a, edges, st);
      else if (lanes == 32) launch_list<2, 32>(a, edges, st);
      else return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  if (variant == 2) {
    const dim3 g(a.nwaves), b(kWave);
    if (reach == 1) {
      if (edges) CF_LAUNCH((k_force_lds<true, 1>), g, b, 0, st, a);
      else CF_LAUNCH((k_force_lds<false, 1>), g, b, 0, st, a);
    } else {
      if (edges) CF_LAUNCH((k_force_lds<true, 2>), g, b, 0, st, a);
      else CF_LAUNCH((k_force_lds<false, 2>), g, b, 0, st, a);
    }
    return hipGetLastError();
  }
  if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
  const int g = grid_for(a.nwaves, block);
  if (edges)
    CF_LAUNCH(k_force<true>, dim3(g), dim3(block), 0, st, a);
  else
    CF_LAUNCH(k_force<false>, dim3(g), dim3(block), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_edges(const EdgeArgs& a, hipStream_t st) {
  // kWave / kEdgeLanes owners per wave, four waves per block
  CF_LAUNCH(k_edges_g, dim3((a.nwaves * kEdgeLanes + 3) / 4), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_tile_sort(const TileSortArgs& a, hipStream_t st) {
  const int g = a.ntiles + a.netiles;
  if (g > 0) CF_LAUNCH(k_tile_sort, dim3(g), dim3(kSortThreads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_scan(const ScanPlan& p, hipStream_t st) {
  CF_LAUNCH(k_scan, dim3(p.tiles_total), dim3(kScanThreads), 0, st, p);
  return hipGetLastError();
}

hipError_t launch_fill(const SortArgs& a, hipStream_t st) {
  CF_LAUNCH(k_fill, dim3(grid_for(a.nwaves, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_rank_scatter(const SortArgs& a, hipStream_t st) {
  CF_LAUNCH(k_rank_scatter, dim3(grid_for(a.nwaves, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_tri_count(const TriArgs& a, hipStream_t st) {
  CF_LAUNCH(k_tri_count, dim3(grid_for(a.nwaves, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_tri_write(const TriArgs& a, hipStream_t st) {
  CF_LAUNCH(k_tri_write, dim3(grid_for(a.nwaves, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_tri_vertices(const TriVertArgs& a, hipStream_t st) {
  const long g = (a.ntris + 255) / 256;
  if (g > 0) CF_LAUNCH(k_tri_vertices, dim3((unsigned)g), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_bin(const BinArgs& a, hipStream_t st) {
  CF_LAUNCH(k_bin, dim3(grid_for(a.nwaves, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_io(int which, const IoArgs& a, int n, hipStream_t st) {
  const int g = (n + 255) / 256;
  if (g == 0) return hipSuccess;
  switch (which) {
    case 0: CF_LAUNCH(k_load_aos, dim3(g), dim3(256), 0, st, a); break;
    case 1: CF_LAUNCH(k_unload_aos, dim3(g), dim3(256), 0, st, a); break;
    case 2: CF_LAUNCH(k_get_counts, dim3(g), dim3(256), 0, st, a); break;
    case 3: CF_LAUNCH(k_set_counts, dim3(g), dim3(256), 0, st, a); break;
    case 4: CF_LAUNCH(k_sort_order, dim3(g), dim3(256), 0, st, a); break;
    case 5: CF_LAUNCH(k_move, dim3(g), dim3(256), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_init(const InitArgs& a, int n, hipStream_t st) {
  const int g = (n + 255) / 256;
  if (g == 0) return hipSuccess;
  CF_LAUNCH(k_init, dim3(g), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_slot_of(const uint32_t* id, uint32_t* slot_of, const SimDesc* sims, int sim, int n,
                          hipStream_t st) {
  const int g = (n + 255) / 256;
  if (g == 0) return hipSuccess;
  CF_LAUNCH(k_slot_of, dim3(g), dim3(256), 0, st, id, slot_of, sims, sim);
  return hipGetLastError();
}

hipError_t launch_line_vertices(const LineArgs& a, hipStream_t st) {
  const long g = (a.nedges + 255) / 256;
  if (g == 0) return hipSuccess;
  CF_LAUNCH(k_line_vertices, dim3((unsigned)g), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace cf
