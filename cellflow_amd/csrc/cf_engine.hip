// Host engine behind the C ABI (include/cellflow.h).
//
// One Engine owns the device state of one or more simulations (a plain simulation is an
// ensemble of one, so both paths run the same kernels in the same order).  All work goes
// to one non-blocking HIP stream per engine at default priority.  The reference's host
// class is ParticleSimulation (cuda-native/src/ParticleSimulation.cu:426-754); each C
// entry point at the bottom of this file names the member it replaces.

#include "../../include/cellflow.h"
#include "cf_host.h"
#include "cf_internal.h"
#include "cf_tune_cache.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

namespace cf {

// ------------------------------------------------------------------------------------
// errors

// fail() and the thread's error text live in cf_host.cpp (host-only, shared with the
// preset and tuning-cache code).

#define HIPR(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return fail(CF_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define CFR(call)            \
  do {                       \
    int rc_ = (call);        \
    if (rc_ != CF_OK) return rc_; \
  } while (0)

// ------------------------------------------------------------------------------------
// host math shared with the kernels' contract

// ParticleSimulation.cu:108-110 in FP32 (this file is compiled with -ffp-contract=off).
static float eff_radius(const cf_params& p, const float* rbt, int ti, int tj) {
  const float ri = p.radius * (1.0f + rbt[ti] * p.ratio_with_lfo);
  const float rj = p.radius * (1.0f + rbt[tj] * p.ratio_with_lfo);
  return (ri + rj) * 0.5f;
}

float cutoff_threshold(float e) {
  if (!(e > 0.0f)) return 0.0f;  // effR <= 0 (or NaN): sqrtf(s + 1e-4f) < effR never holds
  if (std::isinf(e)) return INFINITY;
  float t = e * e;
  while (!(sqrtf(t) >= e)) t = nextafterf(t, INFINITY);
  while (t > 0.0f && sqrtf(nextafterf(t, 0.0f)) >= e) t = nextafterf(t, 0.0f);
  return t;
}

static float max_cutoff(const cf_params& p, const float* rbt, int T) {
  float m = 0.0f;
  for (int a = 0; a < T; ++a)
    for (int b = 0; b < T; ++b) m = std::max(m, eff_radius(p, rbt, a, b));
  return m;
}

// The grid rule.  Cells per axis n = floor(L / (cutmax * (1 + 2^-10) / divisor)), at least
// 1, at most 1024; the total is capped near 8 cells per particle by coarsening the
// largest axis.  Any n with L/n >= that target keeps every interacting pair within
// `divisor` cells of each other, so the rule only trades speed, never correctness.
int grid_dims(const cf_params& p, const float* rbt, int T, int n, int divisor, int dims[3]) {
  const float L[3] = {p.canvas_width, p.canvas_height, p.canvas_depth};
  const float cut = max_cutoff(p, rbt, T);
  if (!std::isfinite(cut)) return fail(CF_ERR_ARG, "effective radius is not finite");
  for (int a = 0; a < 3; ++a) {
    if (!(L[a] > 0.0f) || !std::isfinite(L[a])) return fail(CF_ERR_ARG, "canvas size must be > 0");
    if (cut <= 0.0f) {
      dims[a] = 1;
      continue;
    }
    const double target = (double)cut * (1.0 + 1.0 / 1024.0) / (double)divisor;
    double q = std::floor((double)L[a] / target);
    dims[a] = (int)std::max(1.0, std::min(q, 1024.0));
  }
  const long long cap = std::max<long long>(8LL * n, 4096);
  while ((long long)dims[0] * dims[1] * dims[2] > cap) {
    int a = 0;
    if (dims[1] > dims[a]) a = 1;
    if (dims[2] > dims[a]) a = 2;
    dims[a] = std::max(1, dims[a] - 1);
  }
  return CF_OK;
}

// ------------------------------------------------------------------------------------
// scratch of the allocation-free pair histogram (cf_pair_histogram_noalloc)
//
// A fixed device array instead of memory allocated per call: the per-particle partial
// histograms (bins x stride doubles) and the reduced histogram (bins doubles).  Every GPU
// the library runs on holds its own copy until the process ends.  In host code the name
// `scratch` is not a device address: its address on each device comes from
// hipGetSymbolAddress, taken once per device when the first engine is created there, and
// only that address is passed to kernels.  Calls that overlap on one GPU share the array
// and race; nothing in it must start at zero (the kernels overwrite what they read).
constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

// The address of `scratch` on `device`, taken with that device current; the caller's
// current device is restored.
static int scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices)
    return fail(CF_ERR_DEVICE, "device %d: scratch is set up for devices 0..%d", device,
                kScratchDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return CF_OK;
  int prev = 0;
  HIPR(hipGetDevice(&prev));
  if (prev != device) HIPR(hipSetDevice(device));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess || !p)
    return fail(CF_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return CF_OK;
}

static void* scratch_ptr(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ------------------------------------------------------------------------------------
// engine

struct SimHost {
  int n = 0;
  int T = 1;
  std::vector<float> table;  // T*T final table
  std::vector<float> raw;    // T*T raw table (updateForceTable input)
  std::vector<float> rbt;    // T radius modifiers
  cf_params prm{};
  bool have_prm = false;
  int dims[3] = {1, 1, 1};
  int reach = 1;
  int p0 = 0;
  int c0 = 0;
};

struct EdgeReq {
  bool on = false;
  float distance = 0.0f;
  int maxc = 0;
};

template <class T>
static int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIPR(hipMalloc((void**)p, count * sizeof(T)));
  return CF_OK;
}

template <class T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

struct Engine {
  int kind = 0;  // 1: cf_sim, 2: cf_ensemble
  int device = 0;
  hipStream_t st = nullptr;
  std::vector<SimHost> sims;
  long P = 0;      // padded slots
  int nwaves = 0;
  long C = 0;      // cells incl. one pad cell per simulation
  long Ccap = 0;
  DevBufs b{};
  int cur = 0;
  bool load_pending = false;
  std::vector<int> grid_sig;       // dims/reach of the binning A is sorted for
  std::vector<SimDesc> hdesc;      // uploaded descriptors (kept alive for async copies)
  std::vector<PairConst> hpairs;
  std::vector<uint32_t> hhist;
  // scan plans: 0 = cells, 1 = cells + edge counts, 2 = edge counts
  unsigned long long* tickets = nullptr;
  unsigned long long* flags[3] = {nullptr, nullptr, nullptr};
  int plan_tiles[3] = {0, 0, 0};
  // edges
  int edge_cap_alloc = 0;   // candidate rows allocated (2*maxc)
  int edge_maxc_alloc = 0;
  long long edges_count = -1;  // -1: no edge list built
  // the graph the last call's final step built inside its force pass and sort (that of
  // the state before the step); -1: none (fewer than two steps, or no fused edges)
  long long step_edges_count = -1;
  int32_t* tris = nullptr;     // triangle mesh: 3 local indices per triangle
  long long tris_cap = 0;      // triangles `tris` holds
  long long tri_count = -1;    // -1: no mesh built
  float* vtx = nullptr;        // vertex expansion output (line or triangle layout)
  size_t vtx_cap = 0;          // floats `vtx` holds
  float* vcol = nullptr;       // the caller's colours on the device
  size_t vcol_cap = 0;         // floats `vcol` holds
  unsigned char* iobuf = nullptr;  // staging of the host get/set calls, kept between calls
  size_t iobuf_cap = 0;            // bytes `iobuf` holds
  int edges_maxc = 0;
  uint32_t* slot_of = nullptr;
  // tile sort of the step: per-parity counters (the step with state buffer A counts into
  // parity A and clears parity B for the next step), far list, simulation of each tile
  uint32_t* tcnt[2] = {nullptr, nullptr};
  // the edge tiles' look-back (every launch with edge tiles): ticket counter and one flag
  // word per edge tile, zeroed once with the buffers
  unsigned long long* eticket = nullptr;
  unsigned long long* eflags = nullptr;
  uint32_t* farn = nullptr;  // two words
  uint32_t* far_slot = nullptr;
  uint32_t* ftn[2] = {nullptr, nullptr};  // far movers per tile
  uint32_t* ftl = nullptr;

  uint32_t* tile_sim = nullptr;
  int ntiles = 0, ntiles_cap = 0, netiles = 0;
  // set only while the gradient rollout steps: where the force kernel lists the step's
  // in-range partners (ForceArgs::hits)
  uint32_t* fhits = nullptr;
  uint32_t* fnhits = nullptr;
  std::vector<uint32_t> htile_sim;
  cf_tuning tun{1, 0, 16, 256, 32, 1};
  uint32_t* hsmall = nullptr;  // pinned readback words
  unsigned char* hpin = nullptr;  // pinned staging of the large readbacks, kept between calls
  size_t hpin_cap = 0;            // bytes `hpin` holds
  // Captured step graphs.  A graph bakes kernel arguments (buffer pointers, sizes, tuning),
  // so the cache is keyed on them and dropped whenever a buffer is reallocated.
  struct GraphEntry {
    std::vector<long long> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<GraphEntry> graphs;
  long long buf_gen = 0;  // bumped on every reallocation of a buffer a graph references

  ~Engine() { release(); }

  void drop_graphs() {
    for (auto& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }

  void release() {
    if (st) (void)hipStreamSynchronize(st);
    drop_graphs();
    free_adj();
    for (int i = 0; i < 2; ++i) {
      dfree(b.pos[i]);
      dfree(b.vel[i]);
      dfree(b.id[i]);
      dfree(b.start[i]);
    }
    dfree(b.tpos); dfree(b.tvel); dfree(b.tid); dfree(b.key); dfree(b.arrival);
    dfree(b.hist); dfree(b.tmp_id); dfree(b.acc); dfree(b.pad);
    dfree(b.ecand_d2); dfree(b.ecand_id); dfree(b.estage); dfree(b.ecount); dfree(b.fcount); dfree(b.eoff);
    dfree(b.edges); dfree(b.sedges); dfree(b.stotal); dfree(b.gtotal); dfree(b.sims); dfree(b.wave_sim); dfree(b.pairs); dfree(b.flag);
    dfree(slot_of); dfree(tickets); dfree(tris); dfree(vtx); dfree(vcol); dfree(iobuf);
    tris_cap = 0;
    tri_count = -1;
    vtx_cap = vcol_cap = iobuf_cap = 0;
    for (int i = 0; i < 2; ++i) {
      dfree(tcnt[i]);
      dfree(ftn[i]);
    }
    dfree(farn); dfree(far_slot); dfree(tile_sim); dfree(ftl); dfree(eticket); dfree(eflags);
    for (auto& f : flags) dfree(f);
    if (hsmall) (void)hipHostFree(hsmall);
    hsmall = nullptr;
    if (hpin) (void)hipHostFree(hpin);
    hpin = nullptr;
    hpin_cap = 0;
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
  }

  int init(int nsims, const int* counts, const int* types) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return fail(CF_ERR_DEVICE, "no HIP device available");
    HIPR(hipGetDevice(&device));
    CFR(scratch_setup(device));
    HIPR(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPR(hipHostMalloc((void**)&hsmall, 64 * sizeof(uint32_t), hipHostMallocDefault));
    sims.assign(nsims, SimHost{});
    long p = 0;
    for (int s = 0; s < nsims; ++s) {
      SimHost& h = sims[s];
      h.n = counts[s];
      h.T = types[s];
      h.table.assign(h.T * h.T, 0.0f);
      h.raw.assign(h.T * h.T, 0.0f);
      h.rbt.assign(h.T, 0.0f);
      h.p0 = (int)p;
      p += ((long)h.n + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
    }
    P = std::max<long>(p, kSlotAlign);
    nwaves = (int)(P / kWave);
    for (int i = 0; i < 2; ++i) {
      CFR(dalloc(&b.pos[i], P));
      CFR(dalloc(&b.vel[i], P));
      CFR(dalloc(&b.id[i], P));
    }
    CFR(dalloc(&b.tpos, P)); CFR(dalloc(&b.tvel, P)); CFR(dalloc(&b.tid, P));
    CFR(dalloc(&b.key, P)); CFR(dalloc(&b.arrival, P)); CFR(dalloc(&b.tmp_id, P));
    CFR(dalloc(&b.acc, P)); CFR(dalloc(&b.pad, P));
    CFR(dalloc(&b.ecount, P + 1)); CFR(dalloc(&b.eoff, P + 1)); CFR(dalloc(&b.fcount, P + 1));
    CFR(dalloc(&b.sims, nsims)); CFR(dalloc(&b.wave_sim, nwaves));
    CFR(dalloc(&b.pairs, (size_t)nsims * kMaxTypes * kMaxTypes));
    CFR(dalloc(&b.flag, 4)); CFR(dalloc(&tickets, 3)); CFR(dalloc(&slot_of, P));
    CFR(dalloc(&b.stotal, 1));
    CFR(dalloc(&b.gtotal, 1));
    netiles = (int)((P + kEdgeTile - 1) / kEdgeTile);
    CFR(dalloc(&eticket, 1)); CFR(dalloc(&eflags, netiles));
    HIPR(hipMemsetAsync(eticket, 0, sizeof(unsigned long long), st));
    HIPR(hipMemsetAsync(eflags, 0, (size_t)netiles * sizeof(unsigned long long), st));
    CFR(dalloc(&farn, 2)); CFR(dalloc(&far_slot, P));
    HIPR(hipMemsetAsync(far_slot, 0, P * sizeof(uint32_t), st));
    HIPR(hipMemsetAsync(b.ecount, 0, (P + 1) * sizeof(uint32_t), st));
    HIPR(hipMemsetAsync(b.fcount, 0, (P + 1) * sizeof(uint32_t), st));
    HIPR(hipMemsetAsync(b.tpos, 0, P * sizeof(float4), st));
    HIPR(hipMemsetAsync(b.tvel, 0, P * sizeof(float4), st));
    HIPR(hipMemsetAsync(b.acc, 0, P * sizeof(float4), st));
    HIPR(hipMemsetAsync(b.pad, 0, P * sizeof(float), st));
    HIPR(hipMemsetAsync(b.flag, 0, 4 * sizeof(uint32_t), st));
    std::vector<uint32_t> ws(nwaves, 0);
    for (int s = 0; s < nsims; ++s) {
      const int w0 = sims[s].p0 / kWave;
      const int w1 = (int)((sims[s].p0 + ((long)sims[s].n + kWave - 1) / kWave * kWave) / kWave);
      for (int w = w0; w < w1 && w < nwaves; ++w) ws[w] = (uint32_t)s;
    }
    HIPR(hipMemcpyAsync(b.wave_sim, ws.data(), ws.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice, st));
    // identity indices in slot order for the initial (all zero) state
    std::vector<uint32_t> ids(P);
    for (long i = 0; i < P; ++i) ids[i] = (uint32_t)i;
    HIPR(hipMemcpyAsync(b.tid, ids.data(), P * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    HIPR(hipStreamSynchronize(st));
    load_pending = true;
    return flush();
  }

  // ---- descriptors and grid --------------------------------------------------------

  int build_descs(const EdgeReq& er, bool* grid_changed) {
    hdesc.assign(sims.size(), SimDesc{});
    hpairs.assign(sims.size() * kMaxTypes * kMaxTypes, PairConst{});
    std::vector<int> sig;
    long c = 0;
    int nt = 0;
    for (size_t s = 0; s < sims.size(); ++s) {
      SimHost& h = sims[s];
      SimDesc& d = hdesc[s];
      cf_params p = h.prm;
      if (!h.have_prm) {
        // No parameters yet: one cell holds everything (only loads/gets happen).
        p = cf_params{};
        p.radius = 0.0f;
        p.canvas_width = p.canvas_height = p.canvas_depth = 1.0f;
        p.num_particle_types = h.T;
        p.max_expected_neighbors = 1;
      }
      const int divisor = std::max(1, std::min(2, tun.cell_divisor));
      if (h.have_prm) {
        CFR(grid_dims(p, h.rbt.data(), h.T, h.n, divisor, h.dims));
        h.reach = divisor;
      } else {
        h.dims[0] = h.dims[1] = h.dims[2] = 1;
        h.reach = 1;
      }
      h.c0 = (int)c;
      d.n = h.n;
      d.p0 = h.p0;
      d.c0 = h.c0;
      d.nx = h.dims[0];
      d.ny = h.dims[1];
      d.nz = h.dims[2];
      d.ncells = d.nx * d.ny * d.nz;
      d.T = h.T;
      d.reach = h.reach;
      d.generic = (d.nx < 2 * d.reach + 1 || d.ny < 2 * d.reach + 1 || d.nz < 2 * d.reach + 1);
      d.tab = (int)(s * kMaxTypes * kMaxTypes);
      d.Lx = p.canvas_width;
      d.Ly = p.canvas_height;
      d.Lz = p.canvas_depth;
      d.ihx = (float)d.nx / d.Lx;
      d.ihy = (float)d.ny / d.Ly;
      d.ihz = (float)d.nz / d.Lz;
      d.dt = p.delta_t;
      d.friction = p.friction;
      d.fm = p.force_multiplier;
      d.one_minus_balance = 1.0f - p.balance;
      d.max_exp = (float)p.max_expected_neighbors;
      d.rep = p.repulsion;
      d.att = p.attraction;
      d.kexp2 = -p.k * 1.44269504088896340736f;
      d.hx = d.Lx / (float)d.nx;
      d.hy = d.Ly / (float)d.ny;
      d.hz = d.Lz / (float)d.nz;
      d.edge_inside = 0;
      const float ed2 = er.distance * er.distance;  // ParticleSimulation.cu:660
      for (int a = 0; a < h.T; ++a) {
        // Pruning reach of owner type a: every accepted pair has s + 1e-4f < thr, so its
        // squared distance is below thr - 1e-4 (FP32 rounding absorbed by the margin).
        double reach2 = 0.0;
        for (int bb = 0; bb < h.T; ++bb) {
          PairConst& pc = hpairs[d.tab + a * h.T + bb];
          const float e = h.have_prm ? eff_radius(p, h.rbt.data(), a, bb) : 0.0f;
          pc.thr = cutoff_threshold(e);
          pc.fv = h.table[a * h.T + bb];
          pc.inv_eff = (e > 0.0f) ? 1.0f / e : 0.0f;
          reach2 = std::max(reach2, (double)pc.thr - 1e-4);
        }
        if (er.on) reach2 = std::max(reach2, (double)ed2);
        const float cut2 = (float)(reach2 * (1.0 + 1e-3) + 1e-12);
        for (int bb = 0; bb < h.T; ++bb) hpairs[d.tab + a * h.T + bb].pad = cut2;
        if (er.on && ed2 + 1e-4f < hpairs[d.tab + a * h.T + a].thr) d.edge_inside |= (1 << a);
      }
      d.ident = h.have_prm ? 0 : 1;
      d.edge_on = er.on ? 1 : 0;
      d.edge_d2 = er.distance * er.distance;
      d.edge_maxc = er.maxc;
      const float ih[3] = {d.ihx, d.ihy, d.ihz};
      int re[3];
      for (int a = 0; a < 3; ++a) {
        const double q = std::ceil((double)er.distance * (1.0 + 1.0 / 1024.0) * (double)ih[a]);
        re[a] = (int)std::max(1.0, std::min(q, 1024.0));
      }
      d.edge_rx = re[0];
      d.edge_ry = re[1];
      d.edge_rz = re[2];
      // cell tiles: tr rows of one plane, about kTileTarget particles at this density
      const double occ = (double)std::max(h.n, 1) / (double)d.ncells;
      const int trmax = std::max(1, std::min(d.ny, kTileCells / d.nx));
      d.tr = (int)std::max(1.0, std::min((double)trmax, std::floor(kTileTarget / (d.nx * occ))));
      d.nyt = (d.ny + d.tr - 1) / d.tr;
      d.t0 = nt;
      nt += d.nz * d.nyt;
      c += (long)d.ncells + 1;
      sig.push_back(d.nx);
      sig.push_back(d.ny);
      sig.push_back(d.nz);
      sig.push_back(d.reach);
    }
    C = c;
    *grid_changed = (sig != grid_sig);
    if (*grid_changed) {
      grid_sig = sig;
      CFR(ensure_cells());
    }
    HIPR(hipStreamSynchronize(st));  // descriptors may still be read by queued work
    if (nt > ntiles_cap) {
      drop_graphs();
      ++buf_gen;
      for (int i = 0; i < 2; ++i) {
        dfree(tcnt[i]);
        CFR(dalloc(&tcnt[i], (size_t)nt * kCtrStride));
        dfree(ftn[i]);
        CFR(dalloc(&ftn[i], (size_t)nt * kCtrStride));
      }
      dfree(tile_sim);
      CFR(dalloc(&tile_sim, nt));
      dfree(ftl);
      CFR(dalloc(&ftl, (size_t)nt * kFarCap));
      HIPR(hipMemsetAsync(ftl, 0, (size_t)nt * kFarCap * sizeof(uint32_t), st));
      ntiles_cap = nt;
    }
    ntiles = nt;
    htile_sim.assign(nt, 0u);
    for (size_t s2 = 0; s2 < hdesc.size(); ++s2)
      for (int t = 0; t < hdesc[s2].nz * hdesc[s2].nyt; ++t) htile_sim[hdesc[s2].t0 + t] = (uint32_t)s2;
    if (nt > 0)
      HIPR(hipMemcpyAsync(tile_sim, htile_sim.data(), nt * sizeof(uint32_t), hipMemcpyHostToDevice,
                          st));
    // no step is in flight here: both parities of the step counters start at zero
    for (int i = 0; i < 2; ++i) {
      if (nt > 0) {
        HIPR(hipMemsetAsync(tcnt[i], 0, (size_t)nt * kCtrStride * sizeof(uint32_t), st));
        HIPR(hipMemsetAsync(ftn[i], 0, (size_t)nt * kCtrStride * sizeof(uint32_t), st));
      }
    }
    HIPR(hipMemsetAsync(farn, 0, 2 * sizeof(uint32_t), st));
    HIPR(hipMemcpyAsync(b.sims, hdesc.data(), hdesc.size() * sizeof(SimDesc),
                        hipMemcpyHostToDevice, st));
    HIPR(hipMemcpyAsync(b.pairs, hpairs.data(), hpairs.size() * sizeof(PairConst),
                        hipMemcpyHostToDevice, st));
    return CF_OK;
  }

  // Cell arrays for the current C; the histogram holds zeros and each pad cell's size.
  int ensure_cells() {
    HIPR(hipStreamSynchronize(st));
    drop_graphs();
    ++buf_gen;
    if (C > Ccap) {
      dfree(b.hist);
      dfree(b.start[0]);
      dfree(b.start[1]);
      const long cap = C + C / 4 + 64;
      CFR(dalloc(&b.hist, cap));
      CFR(dalloc(&b.start[0], cap + 1));
      CFR(dalloc(&b.start[1], cap + 1));
      Ccap = cap;
    }
    hhist.assign(C, 0u);
    for (const SimHost& h : sims) {
      const long pad = ((long)h.n + kSlotAlign - 1) / kSlotAlign * kSlotAlign - h.n;
      hhist[h.c0 + (long)h.dims[0] * h.dims[1] * h.dims[2]] = (uint32_t)pad;
    }
    HIPR(hipMemcpyAsync(b.hist, hhist.data(), C * sizeof(uint32_t), hipMemcpyHostToDevice, st));
    // scan plans: tile counts change with C, so tickets and flags restart
    const int tc = (int)((C + kScanTile - 1) / kScanTile);
    const int te = (int)((P + kScanTile - 1) / kScanTile);
    const int tiles[3] = {tc, tc + te, te};
    for (int i = 0; i < 3; ++i) {
      if (tiles[i] > plan_tiles[i]) {
        dfree(flags[i]);
        CFR(dalloc(&flags[i], tiles[i]));
      }
      plan_tiles[i] = tiles[i];
      HIPR(hipMemsetAsync(flags[i], 0, tiles[i] * sizeof(unsigned long long), st));
    }
    HIPR(hipMemsetAsync(tickets, 0, 3 * sizeof(unsigned long long), st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  ScanPlan plan(int which, uint32_t* cell_out) {
    ScanPlan sp{};
    const int tc = (int)((C + kScanTile - 1) / kScanTile);
    const int te = (int)((P + kScanTile - 1) / kScanTile);
    ScanSeg cells{b.hist, cell_out, (int)C, tc};
    // the step's fused graph (plan 1) counts into fcount; the standalone pass and the mesh
    // (plan 2) into ecount
    ScanSeg edges{which == 1 ? b.fcount : b.ecount, b.eoff, (int)P, te};
    if (which == 0) {
      sp.seg[0] = cells;
      sp.nseg = 1;
      sp.tiles_total = tc;
    } else if (which == 1) {
      sp.seg[0] = cells;
      sp.seg[1] = edges;
      sp.nseg = 2;
      sp.tiles_total = tc + te;
    } else {
      sp.seg[0] = edges;
      sp.nseg = 1;
      sp.tiles_total = te;
    }
    sp.ticket = tickets + which;
    sp.flags = flags[which];
    return sp;
  }

  // ---- sorting ---------------------------------------------------------------------

  // Bin `src` (slot order `src_id`) with the current grid and write the (cell, index)
  // sorted state into buffer `dst`, including its cell starts.
  int sort_into(const float4* src_pos, const float4* src_vel, const uint32_t* src_id, int dst) {
    BinArgs ba{src_pos, b.key, b.arrival, b.hist, b.sims, b.wave_sim, nwaves};
    HIPR(launch_bin(ba, st));
    HIPR(launch_scan(plan(0, b.start[dst]), st));
    SortArgs sa{};
    sa.src_pos = src_pos;
    sa.src_vel = src_vel;
    sa.src_id = src_id;
    sa.key = b.key;
    sa.arrival = b.arrival;
    sa.start = b.start[dst];
    sa.hist = b.hist;
    sa.tmp_id = b.tmp_id;
    sa.dst_pos = b.pos[dst];
    sa.dst_vel = b.vel[dst];
    sa.dst_id = b.id[dst];
    sa.sims = b.sims;
    sa.wave_sim = b.wave_sim;
    sa.P = P;
    sa.nwaves = nwaves;
    HIPR(launch_fill(sa, st));
    HIPR(launch_rank_scatter(sa, st));
    return CF_OK;
  }

  // Loads write index-order records of one simulation into the temporaries; the other
  // simulations' current state is copied there first, then everything is sorted once.
  int begin_load() {
    if (load_pending) return CF_OK;
    HIPR(hipMemcpyAsync(b.tpos, b.pos[cur], P * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(b.tvel, b.vel[cur], P * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(b.tid, b.id[cur], P * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    load_pending = true;
    return CF_OK;
  }

  int flush() {
    if (!load_pending) return CF_OK;
    bool changed = false;
    EdgeReq none;
    CFR(build_descs(none, &changed));
    CFR(sort_into(b.tpos, b.tvel, b.tid, cur));
    load_pending = false;
    edges_count = -1;  // a graph or mesh belongs to the state it was built on
    step_edges_count = -1;
    tri_count = -1;
    return CF_OK;
  }

  // Re-sort the current state when the grid changes (new parameters).
  int prepare(const EdgeReq& er) {
    CFR(flush());
    bool changed = false;
    CFR(build_descs(er, &changed));
    if (changed) {
      CFR(sort_into(b.pos[cur], b.vel[cur], b.id[cur], 1 - cur));
      cur = 1 - cur;
    }
    return CF_OK;
  }

  int ensure_edge_bufs(int maxc) {
    const int cap = 2 * maxc;
    if (cap > edge_cap_alloc || maxc > edge_maxc_alloc) {
      HIPR(hipStreamSynchronize(st));
      drop_graphs();
      ++buf_gen;
    }
    if (cap > edge_cap_alloc) {
      HIPR(hipStreamSynchronize(st));
      dfree(b.ecand_d2);
      dfree(b.ecand_id);
      CFR(dalloc(&b.ecand_d2, (size_t)cap * P));
      CFR(dalloc(&b.ecand_id, (size_t)cap * P));
      edge_cap_alloc = cap;
    }
    if (maxc > edge_maxc_alloc) {
      HIPR(hipStreamSynchronize(st));
      dfree(b.estage);
      CFR(dalloc(&b.estage, (size_t)maxc * P));
      // the last graph stays readable when a mesh with a longer list grows the buffers
      decltype(b.edges) grown = nullptr;
      CFR(dalloc(&grown, (size_t)maxc * P));
      if (edges_count > 0 && b.edges)
        HIPR(hipMemcpyAsync(grown, b.edges, (size_t)edges_count * sizeof(*grown),
                            hipMemcpyDeviceToDevice, st));
      HIPR(hipStreamSynchronize(st));
      dfree(b.edges);
      b.edges = grown;
      dfree(b.sedges);
      step_edges_count = -1;
      CFR(dalloc(&b.sedges, (size_t)maxc * P));
      edge_maxc_alloc = maxc;
    }
    return CF_OK;
  }

  // ---- the step ----------------------------------------------------------------------

  ForceArgs force_args(bool edges, int cap, bool write_acc = true) {
    const int A = cur;
    ForceArgs fa{};
    fa.pos = b.pos[A];
    fa.vel = b.vel[A];
    fa.id = b.id[A];
    fa.start = b.start[A];
    fa.tpos = b.tpos;
    fa.tvel = b.tvel;
    fa.key = b.key;
    fa.arrival = b.arrival;
    fa.hist = b.hist;
    fa.acc = write_acc ? b.acc : nullptr;
    fa.sims = b.sims;
    fa.wave_sim = b.wave_sim;
    fa.pairs = b.pairs;
    fa.flag = b.flag;
    fa.ecand_d2 = b.ecand_d2;
    fa.ecand_id = b.ecand_id;
    fa.estage = b.estage;
    fa.estride = edge_maxc_alloc;
    fa.ecount = b.fcount;
    fa.P = P;
    fa.nwaves = nwaves;
    fa.edges = edges ? 1 : 0;
    fa.edge_cap = cap;
    fa.tile_sort = tun.sort_variant == 1 ? 1 : 0;
    fa.tcount = tcnt[A];
    fa.far_n = farn + A;
    fa.far_slot = far_slot;
    fa.ftn = ftn[A];
    fa.ftl = ftl;
    fa.hits = fhits;
    fa.nhits = fnhits;
    return fa;
  }

  SortArgs step_sort_args(bool edges) {
    const int A = cur, B = 1 - cur;
    SortArgs sa{};
    sa.src_pos = b.tpos;
    sa.src_vel = b.tvel;
    sa.src_id = b.id[A];
    sa.key = b.key;
    sa.arrival = b.arrival;
    sa.start = b.start[B];
    sa.hist = b.hist;
    sa.tmp_id = b.tmp_id;
    sa.dst_pos = b.pos[B];
    sa.dst_vel = b.vel[B];
    sa.dst_id = b.id[B];
    sa.sims = b.sims;
    sa.wave_sim = b.wave_sim;
    if (edges) {
      sa.estage = b.estage;
      sa.estride = edge_maxc_alloc;
      sa.eoff = b.eoff;
      sa.edges = b.sedges;
      sa.etotal = b.stotal;
      sa.fcount = b.fcount;
    }
    sa.P = P;
    sa.nwaves = nwaves;
    return sa;
  }

  TileSortArgs tile_sort_args(bool edges) {
    const int A = cur, B = 1 - cur;
    TileSortArgs ta{};
    ta.src_pos = b.tpos;
    ta.src_vel = b.tvel;
    ta.src_id = b.id[A];
    ta.key = b.key;
    ta.hist = b.hist;
    ta.start_old = b.start[A];
    ta.start_new = b.start[B];
    ta.dst_pos = b.pos[B];
    ta.dst_vel = b.vel[B];
    ta.dst_id = b.id[B];
    ta.tmp_id = b.tmp_id;
    ta.tmp_slot = b.arrival;
    ta.sims = b.sims;
    ta.tile_sim = tile_sim;
    ta.tcount = tcnt[A];
    ta.tcount_next = tcnt[B];
    ta.far_n = farn + A;
    ta.far_n_next = farn + B;
    ta.far_slot = far_slot;
    ta.ftn = ftn[A];
    ta.ftn_next = ftn[B];
    ta.ftl = ftl;
    ta.flag = b.flag;
    ta.ntiles = ntiles;
    if (edges) {
      ta.ecount = b.fcount;
      ta.eticket = eticket;
      ta.eflags = eflags;
      ta.estage = b.estage;
      ta.estride = edge_maxc_alloc;
      ta.edges = b.sedges;
      ta.etotal = b.stotal;
      ta.edge_room = (long)edge_maxc_alloc * P;
      ta.netiles = netiles;
    }
    ta.wave_sim = b.wave_sim;
    ta.P = P;
    ta.one_sim = sims.size() == 1 ? 1 : 0;
    return ta;
  }

  // Times the next kernel launch into ev[2i] (start) and ev[2i+1] (end) when `ev` is given.
  static void timed(hipEvent_t* ev, int i) {
    if (ev) arm_kernel_timer(ev[2 * i], ev[2 * i + 1]);
  }

  // Kernel sequence of one step: the force pass, then either the tile sort (which also
  // writes the step's canonical edge list) or the three-pass counting sort; `ev`
  // (optional, two events per kernel) receives each kernel's start and end.  `write_acc`
  // false leaves the accelerations unstored: they are output-only
  // (ParticleSimulation.cu:146), so only the last step of a call stores them.
  int step(bool fused_edges, int cap, hipEvent_t* ev = nullptr, bool write_acc = true) {
    if ((tun.force_variant == 3 || tun.force_variant == 4) && P > (1L << kCodeShift))
      return fail(CF_ERR_ARG, "force variants 3 and 4 take at most 2^%d particle slots",
                  kCodeShift);
    const int B = 1 - cur;
    ForceArgs fa = force_args(fused_edges, cap, write_acc);
    timed(ev, 0);
    const hipError_t e0 = launch_force(fa, fused_edges, tun.block_size, tun.force_variant,
                                       std::max(1, std::min(2, tun.cell_divisor)),
                                       tun.lanes_per_particle, st);
    arm_kernel_timer(nullptr, nullptr);
    HIPR(e0);
    if (tun.sort_variant == 1) {
      timed(ev, 1);
      const hipError_t e1 = launch_tile_sort(tile_sort_args(fused_edges), st);
      arm_kernel_timer(nullptr, nullptr);
      HIPR(e1);
    } else {
      timed(ev, 1);
      const hipError_t e1 = launch_scan(plan(fused_edges ? 1 : 0, b.start[B]), st);
      arm_kernel_timer(nullptr, nullptr);
      HIPR(e1);
      SortArgs sa = step_sort_args(fused_edges);
      timed(ev, 2);
      const hipError_t e2 = launch_fill(sa, st);
      arm_kernel_timer(nullptr, nullptr);
      HIPR(e2);
      timed(ev, 3);
      const hipError_t e3 = launch_rank_scatter(sa, st);
      arm_kernel_timer(nullptr, nullptr);
      HIPR(e3);
    }
    cur = B;
    return CF_OK;
  }

  // Standalone proximity-graph pass on the current state: the candidates (k_edges_g),
  // then the canonical list by the edge tiles of k_tile_sort (launched without cell tiles).
  int edge_pass(int cap, hipEvent_t* ev = nullptr) {
    EdgeArgs ea{};
    ea.pos = b.pos[cur];
    ea.id = b.id[cur];
    ea.start = b.start[cur];
    ea.sims = b.sims;
    ea.wave_sim = b.wave_sim;
    ea.ecand_d2 = b.ecand_d2;
    ea.ecand_id = b.ecand_id;
    ea.estage = b.estage;
    ea.estride = edge_maxc_alloc;
    ea.ecount = b.fcount;
    ea.P = P;
    ea.nwaves = nwaves;
    ea.edge_cap = cap;
    timed(ev, 0);
    const hipError_t e0 = launch_edges(ea, st);
    arm_kernel_timer(nullptr, nullptr);
    HIPR(e0);
    TileSortArgs ta{};
    ta.sims = b.sims;
    ta.wave_sim = b.wave_sim;
    ta.flag = b.flag;
    ta.ntiles = 0;
    ta.ecount = b.fcount;
    ta.eticket = eticket;
    ta.eflags = eflags;
    ta.estage = b.estage;
    ta.estride = edge_maxc_alloc;
    ta.edges = b.edges;
    ta.etotal = b.gtotal;
    ta.edge_room = (long)edge_maxc_alloc * P;
    ta.netiles = netiles;
    ta.P = P;
    ta.one_sim = sims.size() == 1 ? 1 : 0;
    timed(ev, 1);
    const hipError_t e1 = launch_tile_sort(ta, st);
    arm_kernel_timer(nullptr, nullptr);
    HIPR(e1);
    return CF_OK;
  }

  // `count` steps (fused edge candidates when `fused`), replaying a captured graph of
  // 2*m steps where possible.  Even graph lengths return the state to the same buffer, so
  // one graph serves every replay.  The graph's steps store no accelerations; the last
  // step is always launched directly and stores them.
  int run_steps(int count, bool fused, int cap) {
    const int span = std::max(2, (tun.steps_per_graph + 1) / 2 * 2);
    if (tun.steps_per_graph >= 2 && count > span) {
      std::vector<long long> key = {buf_gen, cur, fused ? 1 : 0, cap, span, tun.cell_divisor,
                                    tun.force_variant, tun.lanes_per_particle, tun.block_size,
                                    tun.sort_variant, C, P};
      hipGraphExec_t exec = nullptr;
      for (auto& g : graphs)
        if (g.key == key) exec = g.exec;
      if (!exec) {
        hipGraph_t graph = nullptr;
        HIPR(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int cur0 = cur;
        int rc = CF_OK;
        for (int k = 0; k < span && rc == CF_OK; ++k) rc = step(fused, cap, nullptr, false);
        hipError_t ce = hipStreamEndCapture(st, &graph);
        cur = cur0;
        if (rc != CF_OK) {
          if (graph) (void)hipGraphDestroy(graph);
          return rc;
        }
        HIPR(ce);
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPR(ie);
        graphs.push_back(GraphEntry{key, exec});
      }
      while (count > span) {
        HIPR(hipGraphLaunch(exec, st));
        count -= span;
      }
    }
    for (int k = 0; k < count; ++k) CFR(step(fused, cap, nullptr, k == count - 1));
    return CF_OK;
  }

  bool fusable(const EdgeReq& er) const {
    if (!er.on) return false;
    for (const SimDesc& d : hdesc) {
      const bool ax = (d.nx < 2 * d.reach + 1 || d.edge_rx <= d.reach) &&
                      (d.ny < 2 * d.reach + 1 || d.edge_ry <= d.reach) &&
                      (d.nz < 2 * d.reach + 1 || d.edge_rz <= d.reach);
      if (!ax) return false;
    }
    return true;
  }

  int finish_call(bool edges_built, bool step_edges_built = false) {
    HIPR(hipMemcpyAsync(hsmall, b.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (edges_built)
      HIPR(hipMemcpyAsync(hsmall + 1, b.gtotal, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    if (step_edges_built)
      HIPR(hipMemcpyAsync(hsmall + 3, b.stotal, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    if (edges_built) edges_count = (long long)hsmall[1];
    if (step_edges_built) step_edges_count = (long long)hsmall[3];
    if (hsmall[0] & 1u) return fail(CF_ERR_NONFINITE, "a position or velocity became non-finite");
    if (hsmall[0] & 4u) return fail(CF_ERR_DEVICE, "internal: the step sort lost track of a particle");
    return CF_OK;
  }

  int simulate(int nsteps, const EdgeReq& er) {
    for (const SimHost& h : sims)
      if (!h.have_prm) return fail(CF_ERR_STATE, "parameters not set for every simulation");
    if (er.on) CFR(ensure_edge_bufs(er.maxc));
    CFR(prepare(er));
    HIPR(hipMemsetAsync(b.flag, 0, sizeof(uint32_t), st));
    const bool fuse = fusable(er);
    const int cap = 2 * er.maxc;
    edges_count = -1;
    step_edges_count = -1;
    if (nsteps > 0) tri_count = -1;  // a mesh belongs to the state it was built on
    if (er.on && !fuse) {
      for (int k = 0; k < nsteps; ++k) {
        CFR(step(false, cap));
        CFR(edge_pass(cap));
      }
    } else if (nsteps > 0) {
      // the first step has no pre-step graph to build; later steps build the graph of
      // their pre-step state (= the previous step's post-step state) inside the force pass
      CFR(step(false, cap));
      CFR(run_steps(nsteps - 1, er.on, cap));
    }
    if (er.on && (fuse || nsteps == 0)) CFR(edge_pass(cap));
    if (er.on) edges_maxc = er.maxc;
    return finish_call(er.on, er.on && fuse && nsteps >= 2);
  }

  // ---- state I/O ---------------------------------------------------------------------

  IoArgs io_args(int s) {
    IoArgs a{};
    a.pos = b.pos[cur];
    a.vel = b.vel[cur];
    a.id = b.id[cur];
    a.tpos = b.tpos;
    a.tvel = b.tvel;
    a.tid = b.tid;
    a.acc = b.acc;
    a.pad = b.pad;
    a.sims = b.sims;
    a.sim = s;
    return a;
  }

  // Device staging for `count` elements of T, shared by the get/set calls (each ends with
  // a stream synchronisation, so no two uses overlap); grows on demand.
  template <class T>
  int io_stage(T** p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (bytes > iobuf_cap) {
      HIPR(hipStreamSynchronize(st));
      dfree(iobuf);
      iobuf_cap = 0;
      CFR(dalloc(&iobuf, bytes));
      iobuf_cap = bytes;
    }
    *p = reinterpret_cast<T*>(iobuf);
    return CF_OK;
  }

  // Device-to-host copy of `bytes` into the caller's memory through the pinned staging
  // buffer (a DMA into pinned memory, then one host copy); synchronous.  Splitting it into
  // chunks whose host copies overlap the next DMA measured no faster.
  int to_host(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return CF_OK;
    if (bytes > hpin_cap) {
      HIPR(hipStreamSynchronize(st));
      if (hpin) (void)hipHostFree(hpin);
      hpin = nullptr;
      hpin_cap = 0;
      HIPR(hipHostMalloc((void**)&hpin, bytes, hipHostMallocDefault));
      hpin_cap = bytes;
    }
    HIPR(hipMemcpyAsync(hpin, src, bytes, hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    std::memcpy(dst, hpin, bytes);
    return CF_OK;
  }

  int load_aos(int s, const cf_particle* parts) {
    const SimHost& h = sims[s];
    for (int i = 0; i < h.n; ++i)
      if (parts[i].ptype >= (uint32_t)h.T)
        return fail(CF_ERR_ARG, "particle %d has type %u >= num_types %d", i, parts[i].ptype, h.T);
    CFR(begin_load());
    AosParticle* d = nullptr;
    CFR(io_stage(&d, (size_t)std::max(1, h.n)));
    HIPR(hipMemcpyAsync(d, parts, (size_t)h.n * sizeof(AosParticle), hipMemcpyHostToDevice, st));
    IoArgs a = io_args(s);
    a.aos = d;
    HIPR(launch_io(0, a, h.n, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  int unload_aos(int s, cf_particle* parts) {
    CFR(flush());
    const SimHost& h = sims[s];
    AosParticle* d = nullptr;
    CFR(io_stage(&d, (size_t)std::max(1, h.n)));
    IoArgs a = io_args(s);
    a.aos = d;
    HIPR(launch_io(1, a, h.n, st));
    return to_host(parts, d, (size_t)h.n * sizeof(AosParticle));
  }

  int get_counts(int s, int32_t* out) {
    CFR(flush());
    const SimHost& h = sims[s];
    int32_t* d = nullptr;
    CFR(io_stage(&d, (size_t)std::max(1, h.n)));
    IoArgs a = io_args(s);
    a.counts_out = d;
    HIPR(launch_io(2, a, h.n, st));
    HIPR(hipMemcpyAsync(out, d, (size_t)h.n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  int set_counts(int s, const int32_t* in) {
    CFR(flush());
    const SimHost& h = sims[s];
    int32_t* d = nullptr;
    CFR(io_stage(&d, (size_t)std::max(1, h.n)));
    HIPR(hipMemcpyAsync(d, in, (size_t)h.n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    IoArgs a = io_args(s);
    a.counts_in = d;
    a.tvel = b.vel[cur];
    HIPR(launch_io(3, a, h.n, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  int sort_order(int s, uint32_t* cells, uint32_t* index) {
    CFR(flush());
    const SimHost& h = sims[s];
    uint32_t* d = nullptr;
    CFR(io_stage(&d, 2 * (size_t)std::max(1, h.n)));
    IoArgs a = io_args(s);
    a.cell_out = d;
    a.index_out = d + h.n;
    HIPR(launch_io(4, a, h.n, st));
    HIPR(hipMemcpyAsync(cells, d, (size_t)h.n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPR(hipMemcpyAsync(index, d + h.n, (size_t)h.n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  // Shift one simulation's positions with the reference's wrap and re-bin.  The state is
  // loaded into the temporaries (slots p0..p0+n-1 hold the simulation's particles), moved
  // there and sorted again by the next flush.
  int move(int s, float dx, float dy, float dz, const float box[3]) {
    CFR(flush());
    CFR(begin_load());
    IoArgs a = io_args(s);
    a.shift[0] = dx;
    a.shift[1] = dy;
    a.shift[2] = dz;
    for (int k = 0; k < 3; ++k) a.box[k] = box[k];
    HIPR(launch_io(5, a, sims[s].n, st));
    return flush();
  }

  int init_particles(int s, uint64_t seed, const float canvas[3], float spawn) {
    CFR(begin_load());
    // descriptors must carry this simulation's T before the init kernel runs
    bool changed = false;
    EdgeReq none;
    CFR(build_descs(none, &changed));
    InitArgs a{};
    a.tpos = b.tpos;
    a.tvel = b.tvel;
    a.tid = b.tid;
    a.acc = b.acc;
    a.pad = b.pad;
    a.sims = b.sims;
    a.sim = s;
    a.seed = seed;
    a.cw = canvas[0];
    a.ch = canvas[1];
    a.cd = canvas[2];
    a.spawn = spawn;
    HIPR(launch_init(a, sims[s].n, st));
    return flush();
  }

  int get_edges(int32_t* pairs, long long capacity, long long* count) {
    CFR(flush());
    if (edges_count < 0) return fail(CF_ERR_STATE, "no proximity graph has been built");
    *count = edges_count;
    if (capacity < edges_count)
      return fail(CF_ERR_CAPACITY, "edge buffer holds %lld pairs, %lld needed", capacity,
                  edges_count);
    if (edges_count > 0)
      HIPR(hipMemcpyAsync(pairs, b.edges, (size_t)edges_count * sizeof(int2),
                          hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  int get_step_edges(int32_t* pairs, long long capacity, long long* count) {
    CFR(flush());
    if (step_edges_count < 0)
      return fail(CF_ERR_STATE, "the last call built no graph inside its steps");
    *count = step_edges_count;
    if (capacity < step_edges_count)
      return fail(CF_ERR_CAPACITY, "edge buffer holds %lld pairs, %lld needed", capacity,
                  step_edges_count);
    if (step_edges_count > 0)
      HIPR(hipMemcpyAsync(pairs, b.sedges, (size_t)step_edges_count * sizeof(int2),
                          hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  // Triangle mesh of the current state (generateTriangleMesh, ParticleSimulation.cu:689-754)
  // in canonical order: owner index ascending, then the owner's fan order.
  int tri_mesh(float distance, int max_conn) {
    tri_count = -1;
    const int maxc = std::min(max_conn, 32);  // :303
    if (maxc <= 0) {
      tri_count = 0;
      return CF_OK;
    }
    CFR(ensure_edge_bufs(maxc));
    EdgeReq none;
    CFR(prepare(none));
    const SimDesc& d = hdesc[0];
    // the counts are scanned in 32 bits: at most 64 triangles per owner must fit
    if ((long long)d.n * 64 > 0xFFFFFFFFll)
      return fail(CF_ERR_ARG, "%d particles: a triangle mesh holds at most %lld owners", d.n,
                  0xFFFFFFFFll / 64);
    const float hm = 1.2f;                             // :722, a 20% larger radius
    const float d2 = distance * distance * (hm * hm);  // :723, FP32 left to right
    TriArgs ta{};
    ta.pos = b.pos[cur];
    ta.id = b.id[cur];
    ta.start = b.start[cur];
    ta.sims = b.sims;
    ta.wave_sim = b.wave_sim;
    ta.cand_d2 = b.ecand_d2;
    ta.cand_id = b.ecand_id;
    ta.sorted = b.estage;
    ta.nsel = b.tmp_id;
    ta.count = b.ecount;
    ta.off = b.eoff;
    ta.P = P;
    ta.nwaves = nwaves;
    ta.maxc = maxc;
    ta.d2 = d2;
    // cells within reach of sqrt(d2) on each axis (non-periodic, as :314-317)
    const double reach = std::sqrt((double)d2) * (1.0 + 1.0 / 1024.0);
    const float ih[3] = {d.ihx, d.ihy, d.ihz};
    int r[3];
    for (int a = 0; a < 3; ++a) r[a] = (int)std::max(1.0, std::min(std::ceil(reach * ih[a]), 1024.0));
    ta.rx = r[0];
    ta.ry = r[1];
    ta.rz = r[2];
    HIPR(launch_tri_count(ta, st));
    HIPR(launch_scan(plan(2, nullptr), st));
    HIPR(hipMemcpyAsync(hsmall + 2, b.eoff + P, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    const long long total = (long long)hsmall[2];
    if (total > tris_cap) {
      dfree(tris);
      tris_cap = 0;
      CFR(dalloc(&tris, (size_t)(3 * total)));
      tris_cap = total;
    }
    ta.tris = tris;
    if (total > 0) HIPR(launch_tri_write(ta, st));
    HIPR(hipStreamSynchronize(st));
    tri_count = total;
    return CF_OK;
  }

  int get_triangles(int32_t* out, long long capacity, long long* count) {
    CFR(flush());
    if (tri_count < 0) return fail(CF_ERR_STATE, "no triangle mesh has been built");
    *count = tri_count;
    if (capacity < tri_count)
      return fail(CF_ERR_CAPACITY, "triangle buffer holds %lld, %lld needed", capacity, tri_count);
    if (tri_count > 0)
      HIPR(hipMemcpyAsync(out, tris, (size_t)(3 * tri_count) * sizeof(int32_t),
                          hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  // Device buffers of the vertex expansions, kept for the next call (a render loop calls
  // once per frame): `need` output floats of staging (0: the caller's device buffer takes
  // the output), and the caller's colours uploaded.
  int vertex_bufs(size_t need, const cf_color* colors, int ncolors) {
    const size_t ncf = 3 * (size_t)std::max(1, ncolors);
    if (need > vtx_cap) {
      HIPR(hipStreamSynchronize(st));
      dfree(vtx);
      vtx_cap = 0;
      CFR(dalloc(&vtx, need));
      vtx_cap = need;
    }
    if (ncf > vcol_cap) {
      HIPR(hipStreamSynchronize(st));
      dfree(vcol);
      vcol_cap = 0;
      CFR(dalloc(&vcol, ncf));
      vcol_cap = ncf;
    }
    if (ncolors > 0)
      HIPR(hipMemcpyAsync(vcol, colors, 3 * (size_t)ncolors * sizeof(float),
                          hipMemcpyHostToDevice, st));
    HIPR(launch_slot_of(b.id[cur], slot_of, b.sims, 0, sims[0].n, st));
    return CF_OK;
  }

  // `out` must be memory this engine's device can write: device or managed memory of that
  // device, or pinned host memory (hipHostMalloc).  Anything else (pageable host memory, a
  // pointer HIP does not know, another device's memory) is CF_ERR_ARG, checked on the
  // host before any kernel runs.
  int check_device_out(const void* out) {
    if (!out) return fail(CF_ERR_ARG, "null output pointer");
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, out) != hipSuccess) {
      (void)hipGetLastError();  // the query's error is not the engine's
      return fail(CF_ERR_ARG, "the output is not memory the HIP runtime knows");
    }
    const bool dev = at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
    const bool pinned = at.type == hipMemoryTypeHost && at.devicePointer != nullptr;
    if (!dev && !pinned)
      return fail(CF_ERR_ARG, "the output is not device, managed or pinned host memory");
    if (dev && at.device != device)
      return fail(CF_ERR_ARG, "the output lives on device %d, the engine on device %d",
                  at.device, device);
    return CF_OK;
  }

  // A device output must hold `bytes` from `out` to the end of its allocation, where the
  // runtime can tell (it cannot for every kind of memory; then the caller's capacity is
  // taken as given).
  int check_device_room(const void* out, size_t bytes) {
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)out) != hipSuccess) {
      (void)hipGetLastError();
      return CF_OK;
    }
    const size_t room = (size_t)((const char*)base + size - (const char*)out);
    if (room < bytes)
      return fail(CF_ERR_CAPACITY, "the output allocation holds %zu bytes from the pointer, "
                  "%zu needed", room, bytes);
    return CF_OK;
  }

  // The triangle-vertex expansion into `dev_out` (device-writable, see check_device_out)
  // when non-null, else into `out` (host memory) through the engine's staging buffer.
  int triangle_vertices(const cf_color* colors, int ncolors, float* out, float* dev_out,
                        long long cap_floats, long long* vcount) {
    CFR(flush());
    if (tri_count < 0) return fail(CF_ERR_STATE, "no triangle mesh has been built");
    *vcount = 3 * tri_count;
    const long long need = 27 * tri_count;
    if (cap_floats < need)
      return fail(CF_ERR_CAPACITY, "vertex buffer holds %lld floats, %lld needed", cap_floats, need);
    if (tri_count == 0) return CF_OK;
    if (dev_out) {
      CFR(check_device_out(dev_out));
      CFR(check_device_room(dev_out, (size_t)need * sizeof(float)));
    }
    CFR(vertex_bufs(dev_out ? 0 : (size_t)need, colors, ncolors));
    TriVertArgs va{};
    va.tris = tris;
    va.pos = b.pos[cur];
    va.slot_of = slot_of;
    va.colors = vcol;
    va.out = dev_out ? dev_out : vtx;
    va.ntris = tri_count;
    va.ncolors = ncolors;
    va.T = sims[0].T;
    HIPR(launch_tri_vertices(va, st));
    if (dev_out) {
      HIPR(hipStreamSynchronize(st));
      return CF_OK;
    }
    return to_host(out, vtx, (size_t)need * sizeof(float));
  }

  // The line-vertex expansion, into `dev_out` or `out` as triangle_vertices.
  int line_vertices(const cf_color* colors, int ncolors, float* out, float* dev_out,
                    long long cap_floats, long long* vcount) {
    CFR(flush());
    if (edges_count < 0) return fail(CF_ERR_STATE, "no proximity graph has been built");
    *vcount = 2 * edges_count;
    const long long need = 12 * edges_count;
    if (cap_floats < need)
      return fail(CF_ERR_CAPACITY, "vertex buffer holds %lld floats, %lld needed", cap_floats, need);
    if (edges_count == 0) return CF_OK;
    if (dev_out) {
      CFR(check_device_out(dev_out));
      CFR(check_device_room(dev_out, (size_t)need * sizeof(float)));
    }
    CFR(vertex_bufs(dev_out ? 0 : (size_t)need, colors, ncolors));
    LineArgs la{};
    la.edges = b.edges;
    la.pos = b.pos[cur];
    la.slot_of = slot_of;
    la.colors = vcol;
    la.out = dev_out ? dev_out : vtx;
    la.nedges = edges_count;
    la.ncolors = ncolors;
    la.T = sims[0].T;
    la.p0 = sims[0].p0;
    HIPR(launch_line_vertices(la, st));
    if (dev_out) {
      HIPR(hipStreamSynchronize(st));
      return CF_OK;
    }
    return to_host(out, vtx, (size_t)need * sizeof(float));
  }

  int profile(int nsteps, const EdgeReq& er, double* ms_out, int max_k, int* nk, char* names,
              int names_cap) {
    for (const SimHost& h : sims)
      if (!h.have_prm) return fail(CF_ERR_STATE, "parameters not set");
    if (er.on) CFR(ensure_edge_bufs(er.maxc));
    CFR(prepare(er));
    const bool fuse = fusable(er);
    const int cap = 2 * er.maxc;
    if (nsteps > 0) {  // the profiled steps advance the state
      edges_count = -1;
      tri_count = -1;
    }
    step_edges_count = -1;
    // the step's kernels (ns of them), then the standalone edge pass's two
    const bool tiles = tun.sort_variant == 1;
    const int ns = tiles ? 2 : 4;
    const char* kt[] = {"force", "sort", "edges", "edge_emit"};
    const char* kc[] = {"force", "scan", "fill", "rank_scatter", "edges", "edge_emit"};
    const char* const* kn = tiles ? kt : kc;
    const int K = ns + 2;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int cnt[6] = {0, 0, 0, 0, 0, 0};
    // Every step is queued before any is waited for, so the GPU runs the kernels back to
    // back; each kernel's interval comes from its own dispatch (start and end events of
    // hipExtLaunchKernel), so it holds that kernel alone.  The first step's kernels are
    // left out when there are others (they start on a cold state).
    const int per = 2 * K;
    std::vector<hipEvent_t> ev((size_t)(nsteps + 1) * per, nullptr);
    int rc = CF_OK;
    for (auto& e : ev)
      if (rc == CF_OK && hipEventCreate(&e) != hipSuccess)
        rc = fail(CF_ERR_DEVICE, "hipEventCreate failed");
    std::vector<int> nev(nsteps, 0);
    for (int k = 0; k < nsteps && rc == CF_OK; ++k) {
      hipEvent_t* e = &ev[(size_t)k * per];
      rc = step(er.on && fuse && k > 0, cap, e, k == nsteps - 1);
      nev[k] = ns;
      if (rc == CF_OK && er.on && !fuse) {
        rc = edge_pass(cap, e + 2 * ns);
        nev[k] = ns + 2;
      }
    }
    hipEvent_t* fe = &ev[(size_t)nsteps * per];  // the final standalone edge pass (fused case)
    if (rc == CF_OK && er.on && fuse) rc = edge_pass(cap, fe + 2 * ns);
    if (rc == CF_OK && hipStreamSynchronize(st) != hipSuccess)
      rc = fail(CF_ERR_DEVICE, "the profiled steps failed");
    auto add = [&](const hipEvent_t* e, int i) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, e[2 * i], e[2 * i + 1]) != hipSuccess)
        return fail(CF_ERR_DEVICE, "hipEventElapsedTime failed (a kernel was not timed)");
      acc[i] += ms;
      cnt[i]++;
      return CF_OK;
    };
    for (int k = (nsteps > 1 ? 1 : 0); k < nsteps && rc == CF_OK; ++k)
      for (int i = 0; i < nev[k] && rc == CF_OK; ++i) rc = add(&ev[(size_t)k * per], i);
    if (rc == CF_OK && er.on && fuse)
      for (int i = ns; i < ns + 2 && rc == CF_OK; ++i) rc = add(fe, i);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (rc != CF_OK) return rc;
    int w = 0;
    std::string nm;
    for (int i = 0; i < K && w < max_k; ++i) {
      if (!cnt[i]) continue;
      ms_out[w++] = acc[i] / cnt[i];
      if (!nm.empty()) nm += ",";
      nm += kn[i];
    }
    *nk = w;
    if (names && names_cap > 0) {
      std::snprintf(names, names_cap, "%s", nm.c_str());
    }
    if (er.on) edges_maxc = er.maxc;
    return finish_call(er.on, er.on && fuse && nsteps >= 2);
  }

  // ---- differentiable rollout -----------------------------------------------------------

  struct StateBuf {
    float4* pos = nullptr;
    float4* vel = nullptr;
    uint32_t* id = nullptr;
    uint32_t* start = nullptr;
    double4* x64 = nullptr;  // the gradient rollout's FP64 shadow, by particle index
    double4* v64 = nullptr;
  };

  int alloc_state(StateBuf& s, bool shadow = false) {
    CFR(dalloc(&s.pos, P));
    CFR(dalloc(&s.vel, P));
    CFR(dalloc(&s.id, P));
    CFR(dalloc(&s.start, C + 1));
    if (shadow) {
      CFR(dalloc(&s.x64, P));
      CFR(dalloc(&s.v64, P));
    }
    return CF_OK;
  }

  static void free_state(StateBuf& s) {
    dfree(s.pos);
    dfree(s.vel);
    dfree(s.id);
    dfree(s.start);
    dfree(s.x64);
    dfree(s.v64);
  }

  // The gradient rollout's buffers, kept from one call to the next of the same shape
  // (slots, cells, particles, segments, checkpoint interval, bins, types), so an
  // optimisation loop that calls it repeatedly allocates once.  Freed with the engine, when
  // the shape changes and after a failed call.
  struct AdjBufs {
    std::vector<StateBuf> ck, seg;
    std::vector<float4*> tv;
    StateBuf fin;
    double4* fwx[2] = {nullptr, nullptr};  // the forward's shadow states that are not kept
    double4* fwv[2] = {nullptr, nullptr};
    double4* lx[2] = {nullptr, nullptr};
    double4* lv[2] = {nullptr, nullptr};
    double4* phi = nullptr;
    // in-range lists of each segment step (the force kernel writes them during the
    // recompute with variant 0; with other variants the walk fills the first)
    std::vector<uint32_t*> hits, nhits;
    double *partial = nullptr, *hd = nullptr, *gh = nullptr, *gacc = nullptr;
    float* thr = nullptr;
    std::vector<long> shape;  // empty: nothing allocated
  } adj;

  void free_adj() {
    for (auto& s : adj.ck) free_state(s);
    for (auto& s : adj.seg) free_state(s);
    for (auto& t : adj.tv) dfree(t);
    free_state(adj.fin);
    for (int i = 0; i < 2; ++i) {
      dfree(adj.lx[i]);
      dfree(adj.lv[i]);
      dfree(adj.fwx[i]);
      dfree(adj.fwv[i]);
    }
    dfree(adj.phi);
    for (auto& x : adj.hits) dfree(x);
    for (auto& x : adj.nhits) dfree(x);
    adj.hits.clear();
    adj.nhits.clear();
    dfree(adj.partial);
    dfree(adj.hd);
    dfree(adj.gh);
    dfree(adj.gacc);
    dfree(adj.thr);
    adj.ck.clear();
    adj.seg.clear();
    adj.tv.clear();
    adj.shape.clear();
  }

  // shape = {P, C, n, segments, checkpoint interval, bins, T, partial rows, partial stride}
  int alloc_adj(const std::vector<long>& shape) {
    if (adj.shape == shape) return CF_OK;
    free_adj();
    const long n = shape[2], nseg = shape[3], ckpt = shape[4], bins = shape[5], T = shape[6],
               rows = shape[7], stride = shape[8];
    adj.ck.resize((size_t)std::max(nseg, 1L));
    adj.seg.resize((size_t)ckpt + 1);
    adj.tv.assign((size_t)ckpt, nullptr);
    for (auto& s : adj.ck) CFR(alloc_state(s, true));
    for (size_t i = 1; i < adj.seg.size(); ++i) CFR(alloc_state(adj.seg[i], true));
    for (auto& t : adj.tv) CFR(dalloc(&t, P));
    CFR(alloc_state(adj.fin));
    for (int i = 0; i < 2; ++i) {
      CFR(dalloc(&adj.lx[i], P));
      CFR(dalloc(&adj.lv[i], P));
      CFR(dalloc(&adj.fwx[i], P));
      CFR(dalloc(&adj.fwv[i], P));
    }
    CFR(dalloc(&adj.phi, P));
    adj.hits.assign((size_t)ckpt, nullptr);
    adj.nhits.assign((size_t)ckpt, nullptr);
    for (long i = 0; i < ckpt; ++i) {
      CFR(dalloc(&adj.hits[i], (size_t)std::max(n, 1L) * kPairList));
      CFR(dalloc(&adj.nhits[i], (size_t)std::max(n, 1L)));
    }
    CFR(dalloc(&adj.partial, (size_t)(rows * stride)));
    CFR(dalloc(&adj.hd, (size_t)bins));
    CFR(dalloc(&adj.gh, (size_t)bins));
    CFR(dalloc(&adj.gacc, (size_t)(T * T + 1)));
    CFR(dalloc(&adj.thr, (size_t)(T * T)));
    adj.shape = shape;
    return CF_OK;
  }

  int save_state(StateBuf& s) {
    HIPR(hipMemcpyAsync(s.pos, b.pos[cur], P * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(s.vel, b.vel[cur], P * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(s.id, b.id[cur], P * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(s.start, b.start[cur], (C + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                        st));
    return CF_OK;
  }

  int load_state(const StateBuf& s) {
    HIPR(hipMemcpyAsync(b.pos[cur], s.pos, P * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(b.vel[cur], s.vel, P * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(b.id[cur], s.id, P * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    HIPR(hipMemcpyAsync(b.start[cur], s.start, (C + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                        st));
    return CF_OK;
  }

  // `x64` (optional): FP64 positions by index that lie within slack/2 of s.pos per axis.
  HistArgs hist_args(const StateBuf& s, int bins, double r_max, double* partial, long stride,
                     const double4* x64 = nullptr, double slack = 0.0) {
    const SimDesc& d = hdesc[0];
    HistArgs h{};
    h.pos = s.pos;
    h.id = s.id;
    h.start = s.start;
    h.sims = b.sims;
    h.bins = bins;
    h.r_max = r_max;
    h.x64 = x64;
    h.slack = x64 ? slack : 0.0;
    const double reach = (r_max * (1.0 + 2.0 / bins) + h.slack) * (1.0 + 1e-3);
    const double ih[3] = {d.ihx, d.ihy, d.ihz};
    int r = 1;
    for (int a = 0; a < 3; ++a) r = std::max(r, (int)std::ceil(reach * ih[a]));
    h.reach = r;
    h.partial = partial;
    h.stride = stride;
    return h;
  }

  // Soft pair histogram of the current state (FP64, deterministic).
  int histogram(int bins, double r_max, double* hist_out) {
    if (sims.size() != 1) return fail(CF_ERR_STATE, "histograms are defined for one simulation");
    if (!sims[0].have_prm) return fail(CF_ERR_STATE, "parameters not set");
    EdgeReq none;
    CFR(prepare(none));
    const int n = sims[0].n;
    const long stride = ((long)n + 127) / 128 * 128;
    StateBuf s;
    s.pos = b.pos[cur];
    s.id = b.id[cur];
    s.start = b.start[cur];
    double *partial = nullptr, *hd = nullptr;
    CFR(dalloc(&partial, (size_t)bins * stride));
    CFR(dalloc(&hd, bins));
    HIPR(launch_hist(hist_args(s, bins, r_max, partial, stride), n, st));
    HIPR(launch_reduce_rows(partial, stride, n, bins, 1.0 / n, hd, 0, st));
    HIPR(hipMemcpyAsync(hist_out, hd, bins * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    dfree(partial);
    dfree(hd);
    return CF_OK;
  }

  // The rollout's FP64 constants (from the FP32 parameters).
  AdjConst adj_const(const SimHost& h) const {
    AdjConst c{};
    c.T = h.T;
    c.radius = h.prm.radius;
    c.ratio = h.prm.ratio_with_lfo;
    c.rep = h.prm.repulsion;
    c.att = h.prm.attraction;
    c.k = h.prm.k;
    c.dt = h.prm.delta_t;
    c.friction = h.prm.friction;
    c.fm = h.prm.force_multiplier;
    c.balance = h.prm.balance;
    c.max_exp = (double)h.prm.max_expected_neighbors;
    for (int t = 0; t < h.T; ++t) c.rbt[t] = h.rbt[t];
    for (int t = 0; t < h.T * h.T; ++t) c.table[t] = h.table[t];
    return c;
  }

  // histogram() with its partial sums and result in the device's `scratch` array: no
  // allocation.  The state must already be binned for its parameters (a load since the last
  // step would re-bin, which may allocate), and the partials must fit the array.
  int histogram_noalloc(int bins, double r_max, double* hist_out) {
    if (sims.size() != 1) return fail(CF_ERR_STATE, "histograms are defined for one simulation");
    const int n = sims[0].n;
    const long stride = ((long)n + 127) / 128 * 128;
    const size_t need = ((size_t)bins * (size_t)stride + (size_t)bins) * sizeof(double);
    if (need > kScratchBytes)
      return fail(CF_ERR_ARG, "%d particles x %d bins need %zu bytes of scratch; the "
                  "allocation-free histogram holds at most %zu", n, bins, need, kScratchBytes);
    if (!sims[0].have_prm || load_pending)
      return fail(CF_ERR_STATE, "the state is not binned: step it (cf_simulate, 0 steps "
                                "bins it) after loading, before the allocation-free histogram");
    double* partial = static_cast<double*>(scratch_ptr(device));
    if (!partial) return fail(CF_ERR_DEVICE, "no scratch on device %d", device);
    double* hd = partial + (size_t)bins * stride;
    StateBuf s;
    s.pos = b.pos[cur];
    s.id = b.id[cur];
    s.start = b.start[cur];
    HIPR(launch_hist(hist_args(s, bins, r_max, partial, stride), n, st));
    HIPR(launch_reduce_rows(partial, stride, n, bins, 1.0 / n, hd, 0, st));
    HIPR(hipMemcpyAsync(hist_out, hd, bins * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    return CF_OK;
  }

  // Gradient of the final-state histogram loss after `nsteps` steps with respect to the
  // force table and the ratio, by the adjoint method with checkpoints every `ckpt` steps.
  int rollout_grad(int nsteps, int ckpt, int bins, double r_max, const double* target,
                   double* loss_out, double* hist_out, double* gtab, double* grat) {
    if (sims.size() != 1) return fail(CF_ERR_STATE, "gradients are defined for one simulation");
    SimHost& h = sims[0];
    if (!h.have_prm) return fail(CF_ERR_STATE, "parameters not set");
    EdgeReq none;
    CFR(prepare(none));
    // the rollout advances the state: a graph or mesh of the old state no longer applies
    edges_count = -1;
    step_edges_count = -1;
    tri_count = -1;
    const int n = h.n, T = h.T;
    const long stride = ((long)n + 127) / 128 * 128;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    const int rows = std::max(bins, T * T + 1);
    int rc = CF_OK;
    const DevBufs own = b;  // the engine's state buffers (the recompute re-points them)
    auto body = [&]() -> int {
      CFR(alloc_adj({P, C, (long)n, (long)nseg, (long)ckpt, (long)bins, (long)T, (long)rows,
                     stride}));
      std::vector<StateBuf>& ck = adj.ck;
      std::vector<StateBuf>& seg = adj.seg;  // seg[0] is never allocated: it is ck[sg]
      std::vector<float4*>& tv = adj.tv;     // each segment step's new velocities + counts
      StateBuf& fin = adj.fin;
      double4** lx = adj.lx;
      double4** lv = adj.lv;
      double4* phi = adj.phi;
      // variant 0 lists each step's in-range partners during the recompute; the other
      // variants leave that to the adjoint's walk kernel
      const bool lists = tun.force_variant == 0;
      auto listing = [&](int i) {
        fhits = lists ? adj.hits[i] : nullptr;
        fnhits = lists ? adj.nhits[i] : nullptr;
      };
      double *partial = adj.partial, *hd = adj.hd, *gh = adj.gh, *gacc = adj.gacc;
      float* thr = adj.thr;
      std::vector<float> hthr(T * T);
      for (int i = 0; i < T * T; ++i) hthr[i] = hpairs[hdesc[0].tab + i].thr;
      HIPR(hipMemcpyAsync(thr, hthr.data(), T * T * sizeof(float), hipMemcpyHostToDevice, st));
      HIPR(hipMemsetAsync(b.flag, 0, sizeof(uint32_t), st));
      // Each step reads the engine's current buffers and writes the other ones (its new
      // velocities and counts to tvel, by pre-step slot).  Where a step's states are kept,
      // those are pointed at the kept buffers; the engine's own are restored after, also
      // when a call fails (see below).
      auto point = [&](int side, const StateBuf& s) {
        b.pos[side] = s.pos;
        b.vel[side] = s.vel;
        b.id[side] = s.id;
        b.start[side] = s.start;
      };
      // the FP64 shadow of the rollout (decisions of the FP32 steps, state in FP64): the
      // constants, and where the shadow state after s steps lives (checkpoints and the last
      // segment's states are kept; the others alternate between two buffers)
      AdjArgs sa{};
      sa.sims = b.sims;
      sa.thr = thr;
      sa.c = adj_const(h);
      sa.flag = b.flag;
      sa.slack = 1e-3 * std::max({h.prm.canvas_width, h.prm.canvas_height, h.prm.canvas_depth});
      const int last0 = (nseg - 1) * ckpt;
      auto shadow = [&](int s, bool vel) -> double4* {
        if (s > last0) return vel ? seg[s - last0].v64 : seg[s - last0].x64;
        if (s % ckpt == 0) return vel ? ck[s / ckpt].v64 : ck[s / ckpt].x64;
        return vel ? adj.fwv[s & 1] : adj.fwx[s & 1];
      };
      // one shadow step of an FP32 pre-step state, after the FP32 step whose force kernel
      // listed the state's in-range pairs into list buffer `li` (no lists: the shadow walks)
      auto shadow_step = [&](const StateBuf& pre, int li, const double4* x, const double4* v,
                             double4* xn, double4* vn) -> int {
        sa.pos = pre.pos;
        sa.vel = pre.vel;
        sa.id = pre.id;
        sa.start = pre.start;
        sa.hits = lists ? adj.hits[li] : nullptr;
        sa.nhits = lists ? adj.nhits[li] : nullptr;
        sa.x64 = x;
        sa.v64 = v;
        sa.x64_next = xn;
        sa.v64_next = vn;
        HIPR(launch_shadow_step(sa, n, st));
        return CF_OK;
      };
      sa.pos = b.pos[cur];
      sa.vel = b.vel[cur];
      sa.id = b.id[cur];
      sa.x64_next = shadow(0, false);
      sa.v64_next = shadow(0, true);
      HIPR(launch_shadow_init(sa, n, st));
      // forward with checkpoints; the last segment's states are kept as they are made.
      // Every step lists its in-range pairs (for the shadow; the last segment's lists are
      // also the backward's), into list buffer s mod ckpt.
      for (int s = 0; s < nsteps; ++s) {
        if (s % ckpt == 0) CFR(save_state(ck[s / ckpt]));
        StateBuf pre;
        pre.pos = b.pos[cur];
        pre.vel = b.vel[cur];
        pre.id = b.id[cur];
        pre.start = b.start[cur];
        if (s >= last0) {
          point(1 - cur, seg[s - last0 + 1]);
          b.tvel = tv[s - last0];
        }
        listing(s % ckpt);
        CFR(step(false, 0));
        CFR(shadow_step(pre, s % ckpt, shadow(s, false), shadow(s, true), shadow(s + 1, false),
                        shadow(s + 1, true)));
      }
      fhits = fnhits = nullptr;
      CFR(save_state(fin));
      b = own;
      const double4* xfin = shadow(nsteps, false);
      // the shadow stays close enough to the FP32 state for the histogram's cell reach
      sa.pos = fin.pos;
      sa.id = fin.id;
      sa.x64 = xfin;
      HIPR(launch_shadow_drift(sa, n, st));
      // loss of the final (shadow) state
      HIPR(launch_hist(hist_args(fin, bins, r_max, partial, stride, xfin, sa.slack), n, st));
      HIPR(launch_reduce_rows(partial, stride, n, bins, 1.0 / n, hd, 0, st));
      std::vector<double> hh(bins), g(bins);
      HIPR(hipMemcpyAsync(hh.data(), hd, bins * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPR(hipMemcpyAsync(hsmall, b.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      HIPR(hipStreamSynchronize(st));
      if (hsmall[0] & 1u) return fail(CF_ERR_NONFINITE, "a position or velocity became non-finite");
      if (hsmall[0] & 8u)
        return fail(CF_ERR_STATE, "the rollout's FP64 shadow strayed more than %g from the FP32 "
                    "trajectory: too chaotic for a gradient with frozen decisions", 0.5 * sa.slack);
      double loss = 0.0;
      for (int bb = 0; bb < bins; ++bb) {
        loss += (hh[bb] - target[bb]) * (hh[bb] - target[bb]);
        g[bb] = 2.0 * (hh[bb] - target[bb]);
      }
      *loss_out = loss;
      if (hist_out) std::copy(hh.begin(), hh.end(), hist_out);
      HIPR(hipMemcpyAsync(gh, g.data(), bins * sizeof(double), hipMemcpyHostToDevice, st));
      HistArgs ha = hist_args(fin, bins, r_max, partial, stride, xfin, sa.slack);
      ha.gh = gh;
      ha.lam_x = lx[0];
      ha.lam_v = lv[0];
      HIPR(launch_hist_grad(ha, n, st));
      const int nblk = adj_pair_blocks(n);
      HIPR(hipMemsetAsync(partial, 0, (size_t)(T * T + 1) * nblk * sizeof(double), st));
      AdjArgs aa{};
      aa.sims = b.sims;
      aa.thr = thr;
      aa.c = sa.c;
      aa.phi = phi;
      aa.partial = partial;
      aa.stride = nblk;
      int L = 0;  // lx[L] holds dL/dx of the state after the step being reversed
      for (int sg = nseg - 1; sg >= 0; --sg) {
        const int k0 = sg * ckpt, k1 = std::min(nsteps, k0 + ckpt);
        if (sg != nseg - 1) {  // recompute the segment (and its shadow) from its checkpoint
          point(cur, ck[sg]);
          for (int s = k0; s < k1; ++s) {
            const StateBuf& sp = (s == k0) ? ck[sg] : seg[s - k0];
            point(1 - cur, seg[s - k0 + 1]);
            b.tvel = tv[s - k0];
            listing(s - k0);
            CFR(step(false, 0));
            CFR(shadow_step(sp, s - k0, sp.x64, sp.v64, seg[s - k0 + 1].x64,
                            seg[s - k0 + 1].v64));
          }
          fhits = fnhits = nullptr;
          b = own;
        }
        for (int s = k1 - 1; s >= k0; --s) {
          const StateBuf& pre = (s == k0) ? ck[sg] : seg[s - k0];
          aa.pos = pre.pos;
          aa.vel = pre.vel;
          aa.id = pre.id;
          aa.start = pre.start;
          aa.x64 = pre.x64;
          aa.tvel = tv[s - k0];
          aa.lam_x_next = lx[L];
          aa.lam_v_next = lv[L];
          aa.lam_x = lx[1 - L];
          aa.lam_v = lv[1 - L];
          aa.hits = adj.hits[lists ? s - k0 : 0];
          aa.nhits = adj.nhits[lists ? s - k0 : 0];
          HIPR(launch_adj_step(aa, n, st, !lists));
          L = 1 - L;
        }
      }
      HIPR(launch_reduce_rows(partial, nblk, nblk, T * T + 1, 1.0, gacc, 0, st));
      std::vector<double> gg(T * T + 1);
      HIPR(hipMemcpyAsync(gg.data(), gacc, (T * T + 1) * sizeof(double), hipMemcpyDeviceToHost, st));
      CFR(load_state(fin));
      HIPR(hipMemcpyAsync(hsmall, b.flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
      HIPR(hipStreamSynchronize(st));
      std::copy(gg.begin(), gg.begin() + T * T, gtab);
      *grat = gg[T * T];
      if (hsmall[0] & 1u) return fail(CF_ERR_NONFINITE, "a position or velocity became non-finite");
      return CF_OK;
    };
    rc = body();
    (void)hipStreamSynchronize(st);
    b = own;  // the engine's own state buffers, whatever the recompute pointed them at
    fhits = fnhits = nullptr;
    if (rc != CF_OK) free_adj();
    return rc;
  }

  // ---- autotuner -----------------------------------------------------------------------

  // Bucket key of the current workload: device arch, log2 of the particle count (all
  // simulations), half-octaves of the mean occupancy a particle sees in its own cell of the
  // divisor-1 grid (sum of occupancy^2 over cells / particles; this tracks clustering, which
  // a mean density would miss), the largest type count, the edge cap and log2 of the number
  // of simulations.  Leaves the state binned for the divisor-1 grid; the next step re-sorts
  // it for the tuning in force (the canonical order depends on positions only).
  int tune_key(const EdgeReq& er, std::string* key) {
    for (const SimHost& h : sims)
      if (!h.have_prm) return fail(CF_ERR_STATE, "parameters not set for every simulation");
    const cf_tuning saved = tun;
    tun.cell_divisor = 1;
    int rc = prepare(er);
    tun = saved;
    CFR(rc);
    std::vector<uint32_t> starts(C + 1);
    HIPR(hipMemcpyAsync(starts.data(), b.start[cur], (C + 1) * sizeof(uint32_t),
                        hipMemcpyDeviceToHost, st));
    HIPR(hipStreamSynchronize(st));
    double sq = 0.0, ntot = 0.0;
    int tmax = 1;
    for (const SimHost& h : sims) {
      const long nc = (long)h.dims[0] * h.dims[1] * h.dims[2];
      for (long c = h.c0; c < h.c0 + nc; ++c) {
        const double occ = (double)starts[c + 1] - (double)starts[c];
        sq += occ * occ;
      }
      ntot += h.n;
      tmax = std::max(tmax, h.T);
    }
    const double seen = sq / std::max(1.0, ntot);
    const int nb = (int)std::floor(std::log2(std::max(1.0, ntot)));
    const int rb = (int)std::floor(2.0 * std::log2(std::max(seen, 1.0 / 1024.0)));
    const int sb = (int)std::floor(std::log2((double)sims.size()));
    const int eb = !er.on ? 0 : (er.maxc <= 16 ? 16 : 32);
    std::string arch = "unknown";
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
      arch = prop.gcnArchName;
      const size_t colon = arch.find(':');
      if (colon != std::string::npos) arch.resize(colon);
    }
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s/sims=%d/n=%d/rho=%d/T=%d/e=%d", arch.c_str(), sb, nb, rb,
                  tmax, eb);
    *key = buf;
    return CF_OK;
  }

  // State of the call (binned for `orig`), restored before every trial and at the end.
  int restore(const StateBuf& snap, const cf_tuning& orig, const EdgeReq& er) {
    tun = orig;
    bool changed = false;
    CFR(build_descs(er, &changed));
    return load_state(snap);
  }

  // Average microseconds per step over `nsteps` steps as cf_simulate runs them after its
  // first step (graph replays where the tuning asks for them), timed with HIP events on
  // the engine's stream.
  int time_steps(const EdgeReq& er, int nsteps, double* us) {
    const bool fuse = fusable(er);
    const int cap = 2 * er.maxc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPR(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      return fail(CF_ERR_DEVICE, "hipEventCreate failed");
    }
    auto body = [&]() -> int {
      HIPR(hipEventRecord(e0, st));
      if (er.on && !fuse) {
        for (int k = 0; k < nsteps; ++k) {
          CFR(step(false, cap));
          CFR(edge_pass(cap));
        }
      } else {
        CFR(run_steps(nsteps, er.on, cap));
      }
      HIPR(hipEventRecord(e1, st));
      HIPR(hipEventSynchronize(e1));
      float ms = 0.0f;
      HIPR(hipEventElapsedTime(&ms, e0, e1));
      *us = 1e3 * (double)ms / nsteps;
      return CF_OK;
    };
    const int rc = body();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
  }

  struct Trial {
    cf_tuning t;
    double us;
    bool pruned;
  };

  // Times every candidate tuning and leaves the fastest in force.  Each candidate first runs
  // one directly launched probe step; one whose probe takes over three times the fastest
  // probe so far is not timed further (its probe time is reported).
  int autotune(const EdgeReq& er, int trial_steps, cf_tuning* best_out, double* best_us,
               std::vector<Trial>* trials) {
    for (const SimHost& h : sims)
      if (!h.have_prm) return fail(CF_ERR_STATE, "parameters not set for every simulation");
    if (er.on) CFR(ensure_edge_bufs(er.maxc));
    const cf_tuning orig = tun;
    CFR(prepare(er));
    std::vector<cf_tuning> cands;
    for (int d = 1; d <= 2; ++d) {
      for (int g : {16, 8, 4, 32}) cands.push_back(cf_tuning{d, 0, g, orig.block_size,
                                                             orig.steps_per_graph, 1});
      for (int bs : {64, 128, 256}) cands.push_back(cf_tuning{d, 1, orig.lanes_per_particle, bs,
                                                              orig.steps_per_graph, 1});
      cands.push_back(cf_tuning{d, 2, orig.lanes_per_particle, orig.block_size,
                                orig.steps_per_graph, 1});
      if (P > (1L << kCodeShift)) continue;  // the list kernels' slot limit
      for (int g : {16, 8, 4, 32})
        if (g >= 4 * d)
          cands.push_back(cf_tuning{d, 3, g, orig.block_size, orig.steps_per_graph, 1});
      for (int g : {8, 16})
        cands.push_back(cf_tuning{d, 4, g, orig.block_size, orig.steps_per_graph, 1});
    }
    const size_t nforce = cands.size();  // then the best two again with the three-pass sort
    const int span = orig.steps_per_graph >= 2 ? (orig.steps_per_graph + 1) / 2 * 2 : 2;
    const int nsteps = std::max(span, (std::max(1, trial_steps) + span - 1) / span * span);
    StateBuf snap;
    float4* acc_snap = nullptr;  // the reported acceleration is state too (trials rewrite it)
    auto body = [&]() -> int {
      CFR(alloc_state(snap));
      CFR(save_state(snap));
      CFR(dalloc(&acc_snap, P));
      HIPR(hipMemcpyAsync(acc_snap, b.acc, P * sizeof(float4), hipMemcpyDeviceToDevice, st));
      double best = -1.0, min_probe = -1.0;
      cf_tuning bt = orig;
      for (size_t ci = 0;; ++ci) {
        if (ci == nforce) {
          std::vector<std::pair<double, cf_tuning>> timed;
          for (const Trial& tr : *trials)
            if (!tr.pruned) timed.push_back({tr.us, tr.t});
          std::sort(timed.begin(), timed.end(),
                    [](const auto& x, const auto& y) { return x.first < y.first; });
          for (size_t k = 0; k < timed.size() && k < 2; ++k) {
            cf_tuning c = timed[k].second;
            c.sort_variant = 0;
            cands.push_back(c);
          }
        }
        if (ci >= cands.size()) break;
        const cf_tuning c = cands[ci];
        CFR(restore(snap, orig, er));
        tun = c;
        CFR(prepare(er));
        HIPR(hipMemsetAsync(b.flag, 0, sizeof(uint32_t), st));
        CFR(step(false, 2 * er.maxc));  // the first step of a call builds no graph
        double probe = 0.0, us = 0.0;
        CFR(time_steps(er, 1, &probe));
        if (min_probe > 0.0 && probe > 3.0 * min_probe) {
          trials->push_back(Trial{c, probe, true});
          continue;
        }
        min_probe = (min_probe < 0.0) ? probe : std::min(min_probe, probe);
        CFR(time_steps(er, span, &us));  // captures the step graph
        CFR(time_steps(er, nsteps, &us));
        trials->push_back(Trial{c, us, false});
        if (best < 0.0 || us < best) {
          best = us;
          bt = c;
        }
      }
      CFR(restore(snap, orig, er));
      HIPR(hipMemcpyAsync(b.acc, acc_snap, P * sizeof(float4), hipMemcpyDeviceToDevice, st));
      HIPR(hipMemsetAsync(b.flag, 0, sizeof(uint32_t), st));
      HIPR(hipStreamSynchronize(st));
      tun = bt;
      *best_out = bt;
      *best_us = best;
      return CF_OK;
    };
    const int rc = body();
    (void)hipStreamSynchronize(st);
    if (rc != CF_OK && snap.pos) {
      // leave the call's state in place even when a trial failed
      const std::string msg = error_text();
      if (restore(snap, orig, er) == CF_OK && acc_snap)
        (void)hipMemcpyAsync(b.acc, acc_snap, P * sizeof(float4), hipMemcpyDeviceToDevice, st);
      (void)hipStreamSynchronize(st);
      error_text() = msg;
    }
    if (rc != CF_OK) tun = orig;
    free_state(snap);
    dfree(acc_snap);
    drop_graphs();
    edges_count = -1;
    step_edges_count = -1;
    return rc;
  }
};

static bool check_tuning_values(const cf_tuning& t) { return tunecache::valid(t); }

// Lookup-or-measure driver shared by cf_autotune and cf_ensemble_autotune.
static int tune_call(Engine& e, const EdgeReq& er, int trial_steps, const char* path, int flags,
                     cf_tuning* chosen, double* us_out, char* report, int report_cap) {
  std::string key;
  CFR(e.tune_key(er, &key));
  std::string why;
  std::map<std::string, tunecache::Entry> ents;
  int st = tunecache::kMissing;
  if (path) {
    st = tunecache::read_file(path, ents, why);
    if (st == tunecache::kBad)
      return fail(CF_ERR_ARG, "tuning cache %s: %s", path, why.c_str());
  }
  std::string rep = "{\"key\": \"" + tunecache::escape(key) + "\", ";
  cf_tuning t = e.tun;
  double us = 0.0;
  auto hit = ents.find(key);
  if (path && !(flags & CF_TUNE_FORCE) && hit != ents.end() && check_tuning_values(hit->second.tun)) {
    t = hit->second.tun;
    t.steps_per_graph = e.tun.steps_per_graph;
    e.tun = t;
    us = hit->second.us;
    rep += "\"cached\": true}";
  } else {
    std::vector<Engine::Trial> trials;
    CFR(e.autotune(er, trial_steps, &t, &us, &trials));
    rep += "\"cached\": false, \"trials\": [";
    for (size_t i = 0; i < trials.size(); ++i) {
      char b[256];
      const cf_tuning& c = trials[i].t;
      std::snprintf(b, sizeof b,
                    "%s{\"cell_divisor\": %d, \"force_variant\": %d, \"lanes_per_particle\": %d, "
                    "\"block_size\": %d, \"sort_variant\": %d, \"us_per_step\": %.3f, "
                    "\"pruned\": %s}",
                    i ? ", " : "", c.cell_divisor, c.force_variant, c.lanes_per_particle,
                    c.block_size, c.sort_variant, trials[i].us, trials[i].pruned ? "true" : "false");
      rep += b;
    }
    rep += "]}";
    if (path) {
      ents[key] = tunecache::Entry{t, us};
      if (!tunecache::write_file(path, ents, why))
        return fail(CF_ERR_ARG, "cannot write tuning cache %s: %s", path, why.c_str());
    }
  }
  if (chosen) *chosen = t;
  if (us_out) *us_out = us;
  if (report && report_cap > 0) std::snprintf(report, report_cap, "%s", rep.c_str());
  return CF_OK;
}

}  // namespace cf

// ------------------------------------------------------------------------------------
// handles

struct cf_sim {
  std::unique_ptr<cf::Engine> e;
};
struct cf_ensemble {
  std::unique_ptr<cf::Engine> e;
};

namespace {
std::mutex g_mu;

template <class H>
std::unordered_set<const H*>& live() {
  static std::unordered_set<const H*> s;
  return s;
}

template <class H>
int check(H* h) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!h || !live<H>().count(h)) return cf::fail(CF_ERR_HANDLE, "invalid or destroyed handle");
  return CF_OK;
}

template <class H>
void reg(H* h) {
  std::lock_guard<std::mutex> lk(g_mu);
  live<H>().insert(h);
}

template <class H>
void unreg(H* h) {
  std::lock_guard<std::mutex> lk(g_mu);
  live<H>().erase(h);
}

int check_params(const cf_params* p, int T) {
  if (!p) return cf::fail(CF_ERR_ARG, "params is null");
  if (p->num_particle_types != T)
    return cf::fail(CF_ERR_ARG, "params.num_particle_types %d != simulation types %d",
                    p->num_particle_types, T);
  const float f[] = {p->radius, p->delta_t, p->friction, p->repulsion, p->attraction, p->k,
                     p->balance, p->ratio_with_lfo, p->force_multiplier};
  for (float v : f)
    if (!std::isfinite(v)) return cf::fail(CF_ERR_ARG, "non-finite parameter");
  if (!(p->canvas_width > 0 && p->canvas_height > 0 && p->canvas_depth > 0) ||
      !std::isfinite(p->canvas_width) || !std::isfinite(p->canvas_height) ||
      !std::isfinite(p->canvas_depth))
    return cf::fail(CF_ERR_ARG, "canvas sizes must be finite and > 0");
  return CF_OK;
}

int edge_req(const cf_edge_opts* e, cf::EdgeReq* r) {
  r->on = false;
  if (!e) return CF_OK;
  if (!(e->distance >= 0.0f) || !std::isfinite(e->distance))
    return cf::fail(CF_ERR_ARG, "edge distance must be finite and >= 0");
  if (e->max_connections > CF_MAX_CONNECTIONS)
    return cf::fail(CF_ERR_ARG, "max_connections > %d", CF_MAX_CONNECTIONS);
  r->on = e->max_connections > 0;
  r->distance = e->distance;
  r->maxc = std::max(0, std::min(e->max_connections, CF_MAX_CONNECTIONS));
  if (!r->on) r->maxc = 0;
  return CF_OK;
}

int make_engine(std::unique_ptr<cf::Engine>& out, int kind, int nsims, const int* counts,
                const int* types) {
  for (int s = 0; s < nsims; ++s) {
    if (counts[s] < 1 || counts[s] > (1 << 26))
      return cf::fail(CF_ERR_ARG, "particle count %d out of range", counts[s]);
    if (types[s] < 1 || types[s] > CF_MAX_TYPES)
      return cf::fail(CF_ERR_ARG, "num_types %d out of range 1..%d", types[s], CF_MAX_TYPES);
  }
  std::unique_ptr<cf::Engine> e(new cf::Engine());
  e->kind = kind;
  CFR(e->init(nsims, counts, types));
  out = std::move(e);
  return CF_OK;
}
}  // namespace

// ------------------------------------------------------------------------------------
// C ABI

extern "C" {

int cf_version(void) { return 100; }

int cf_device_count(int* count) {
  if (!count) return cf::fail(CF_ERR_ARG, "count is null");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return cf::fail(CF_ERR_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return CF_OK;
}

float cf_cutoff_threshold(float eff_radius) { return cf::cutoff_threshold(eff_radius); }

void cf_force_table_transform(const float* raw, int count, float range, float bias, float offset,
                              float* out) {
  // ParticleSimulation.cu:524-527 (float operands: the float overloads of tanh/fmax/fmin)
  for (int i = 0; i < count; ++i) {
    const float t = tanhf(raw[i] * offset) * range + bias;
    out[i] = fmaxf(-1.0f, fminf(1.0f, t));
  }
}

int cf_grid_dims(const cf_params* params, const float* rbt, int particle_count, int divisor,
                 int dims[3]) {
  if (!params || !rbt || !dims) return cf::fail(CF_ERR_ARG, "null argument");
  if (params->num_particle_types < 1 || params->num_particle_types > CF_MAX_TYPES)
    return cf::fail(CF_ERR_ARG, "num_particle_types out of range");
  if (divisor != 1 && divisor != 2) return cf::fail(CF_ERR_ARG, "cell_divisor must be 1 or 2");
  return cf::grid_dims(*params, rbt, params->num_particle_types, particle_count, divisor, dims);
}

int cf_create(int particle_count, int num_types, cf_sim** out) {
  if (!out) return cf::fail(CF_ERR_ARG, "out is null");
  *out = nullptr;
  cf_sim* s = new cf_sim();
  int rc = make_engine(s->e, 1, 1, &particle_count, &num_types);
  if (rc != CF_OK) {
    delete s;
    return rc;
  }
  reg(s);
  *out = s;
  return CF_OK;
}

int cf_destroy(cf_sim* sim) {
  CFR(check(sim));
  unreg(sim);
  delete sim;
  return CF_OK;
}

int cf_get_particle_count(cf_sim* sim, int* out) {
  CFR(check(sim));
  if (!out) return cf::fail(CF_ERR_ARG, "out is null");
  *out = sim->e->sims[0].n;
  return CF_OK;
}

int cf_get_num_particle_types(cf_sim* sim, int* out) {
  CFR(check(sim));
  if (!out) return cf::fail(CF_ERR_ARG, "out is null");
  *out = sim->e->sims[0].T;
  return CF_OK;
}

int cf_set_particle_count(cf_sim* sim, int particle_count) {
  CFR(check(sim));
  const cf::SimHost old = sim->e->sims[0];
  if (particle_count == old.n) return CF_OK;
  std::unique_ptr<cf::Engine> fresh;
  int T = old.T;
  CFR(make_engine(fresh, 1, 1, &particle_count, &T));
  fresh->tun = sim->e->tun;
  cf::SimHost& h = fresh->sims[0];
  h.table = old.table;
  h.raw = old.raw;
  h.rbt = old.rbt;
  h.prm = old.prm;
  h.have_prm = old.have_prm;
  sim->e = std::move(fresh);
  return CF_OK;
}

int cf_set_num_particle_types(cf_sim* sim, int num_types) {
  CFR(check(sim));
  if (num_types < 1 || num_types > CF_MAX_TYPES)
    return cf::fail(CF_ERR_ARG, "num_types %d out of range 1..%d", num_types, CF_MAX_TYPES);
  cf::SimHost& h = sim->e->sims[0];
  h.T = num_types;
  h.table.assign(num_types * num_types, 0.0f);
  h.raw.assign(num_types * num_types, 0.0f);
  h.rbt.assign(num_types, 0.0f);
  h.have_prm = false;
  // types of the current particles may now be out of range: reset the state to zero
  std::vector<cf_particle> zero(h.n);
  std::memset(zero.data(), 0, zero.size() * sizeof(cf_particle));
  CFR(sim->e->load_aos(0, zero.data()));
  return sim->e->flush();
}

int cf_set_particles(cf_sim* sim, const cf_particle* particles, int count) {
  CFR(check(sim));
  if (!particles) return cf::fail(CF_ERR_ARG, "particles is null");
  if (count != sim->e->sims[0].n)
    return cf::fail(CF_ERR_ARG, "count %d != particle count %d", count, sim->e->sims[0].n);
  CFR(sim->e->load_aos(0, particles));
  return sim->e->flush();
}

int cf_get_particles(cf_sim* sim, cf_particle* particles, int count) {
  CFR(check(sim));
  if (!particles) return cf::fail(CF_ERR_ARG, "particles is null");
  if (count != sim->e->sims[0].n)
    return cf::fail(CF_ERR_ARG, "count %d != particle count %d", count, sim->e->sims[0].n);
  return sim->e->unload_aos(0, particles);
}

int cf_set_state_soa(cf_sim* sim, const float* pos, const float* vel, const uint32_t* type,
                     int count) {
  CFR(check(sim));
  if (!pos || !type) return cf::fail(CF_ERR_ARG, "pos and type are required");
  if (count != sim->e->sims[0].n)
    return cf::fail(CF_ERR_ARG, "count %d != particle count %d", count, sim->e->sims[0].n);
  std::vector<cf_particle> a(count);
  for (int i = 0; i < count; ++i) {
    cf_particle& r = a[i];
    std::memset(&r, 0, sizeof(r));
    for (int k = 0; k < 3; ++k) {
      r.pos[k] = pos[3 * i + k];
      r.vel[k] = vel ? vel[3 * i + k] : 0.0f;
    }
    r.ptype = type[i];
  }
  CFR(sim->e->load_aos(0, a.data()));
  return sim->e->flush();
}

int cf_get_state_soa(cf_sim* sim, float* pos, float* vel, uint32_t* type, int count) {
  CFR(check(sim));
  if (count != sim->e->sims[0].n)
    return cf::fail(CF_ERR_ARG, "count %d != particle count %d", count, sim->e->sims[0].n);
  std::vector<cf_particle> a(count);
  CFR(sim->e->unload_aos(0, a.data()));
  for (int i = 0; i < count; ++i) {
    for (int k = 0; k < 3; ++k) {
      if (pos) pos[3 * i + k] = a[i].pos[k];
      if (vel) vel[3 * i + k] = a[i].vel[k];
    }
    if (type) type[i] = a[i].ptype;
  }
  return CF_OK;
}

int cf_set_neighbor_counts(cf_sim* sim, const int32_t* counts, int count) {
  CFR(check(sim));
  if (!counts || count != sim->e->sims[0].n) return cf::fail(CF_ERR_ARG, "bad counts");
  return sim->e->set_counts(0, counts);
}

int cf_get_neighbor_counts(cf_sim* sim, int32_t* counts, int count) {
  CFR(check(sim));
  if (!counts || count != sim->e->sims[0].n) return cf::fail(CF_ERR_ARG, "bad counts");
  return sim->e->get_counts(0, counts);
}

int cf_init_particles(cf_sim* sim, uint64_t seed, const float canvas[3], float spawn_region) {
  CFR(check(sim));
  if (!canvas) return cf::fail(CF_ERR_ARG, "canvas is null");
  for (int a = 0; a < 3; ++a)
    if (!(canvas[a] > 0.0f) || !std::isfinite(canvas[a])) return cf::fail(CF_ERR_ARG, "bad canvas");
  if (!(spawn_region > 0.0f)) return cf::fail(CF_ERR_ARG, "spawn_region must be > 0");
  return sim->e->init_particles(0, seed, canvas, spawn_region);
}

int cf_move_universe(cf_sim* sim, float dx, float dy, float dz, const float canvas[3]) {
  CFR(check(sim));
  if (!canvas) return cf::fail(CF_ERR_ARG, "canvas is null");
  const float sh[3] = {dx, dy, dz};
  for (int k = 0; k < 3; ++k) {
    if (!(canvas[k] > 0.0f) || !std::isfinite(canvas[k]) || !std::isfinite(sh[k]))
      return cf::fail(CF_ERR_ARG, "canvas must be finite and > 0, shifts finite");
  }
  return sim->e->move(0, dx, dy, dz, canvas);
}

int cf_set_force_table(cf_sim* sim, const float* table, int num_types) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  if (!table || num_types != h.T) return cf::fail(CF_ERR_ARG, "table must be num_types^2 floats");
  h.table.assign(table, table + h.T * h.T);
  return CF_OK;
}

int cf_get_force_table(cf_sim* sim, float* table, int count) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  if (!table || count != h.T * h.T) return cf::fail(CF_ERR_ARG, "count must be num_types^2");
  std::copy(h.table.begin(), h.table.end(), table);
  return CF_OK;
}

int cf_set_raw_force_table(cf_sim* sim, const float* raw, int count) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  if (!raw || count != h.T * h.T) return cf::fail(CF_ERR_ARG, "count must be num_types^2");
  h.raw.assign(raw, raw + count);
  return CF_OK;
}

int cf_get_raw_force_table(cf_sim* sim, float* raw, int count) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  if (!raw || count != h.T * h.T) return cf::fail(CF_ERR_ARG, "count must be num_types^2");
  std::copy(h.raw.begin(), h.raw.end(), raw);
  return CF_OK;
}

int cf_update_force_table(cf_sim* sim, float range, float bias, float offset) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  cf_force_table_transform(h.raw.data(), h.T * h.T, range, bias, offset, h.table.data());
  return CF_OK;
}

int cf_set_radio_by_type(cf_sim* sim, const float* values, int count) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  if (!values || count != h.T) return cf::fail(CF_ERR_ARG, "count must be num_types");
  h.rbt.assign(values, values + count);
  return CF_OK;
}

int cf_set_radio_by_type_value(cf_sim* sim, int index, float value) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  if (index >= 0 && index < h.T) h.rbt[index] = value;  // out of range: ignored (.cu:616)
  return CF_OK;
}

int cf_get_radio_by_type(cf_sim* sim, float* values, int count) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  if (!values || count != h.T) return cf::fail(CF_ERR_ARG, "count must be num_types");
  std::copy(h.rbt.begin(), h.rbt.end(), values);
  return CF_OK;
}

int cf_rotate_radio_by_type(cf_sim* sim) {
  CFR(check(sim));
  cf::SimHost& h = sim->e->sims[0];
  std::rotate(h.rbt.rbegin(), h.rbt.rbegin() + 1, h.rbt.rend());  // .cu:596-600
  return CF_OK;
}

int cf_simulate(cf_sim* sim, const cf_params* params, int nsteps, const cf_edge_opts* edges) {
  CFR(check(sim));
  if (nsteps < 0) return cf::fail(CF_ERR_ARG, "nsteps < 0");
  CFR(check_params(params, sim->e->sims[0].T));
  cf::EdgeReq er;
  CFR(edge_req(edges, &er));
  sim->e->sims[0].prm = *params;
  sim->e->sims[0].have_prm = true;
  return sim->e->simulate(nsteps, er);
}

int cf_build_edges(cf_sim* sim, float distance, int max_connections) {
  CFR(check(sim));
  cf_edge_opts o{distance, max_connections};
  cf::EdgeReq er;
  CFR(edge_req(&o, &er));
  cf::Engine& e = *sim->e;
  if (!er.on) {
    e.edges_count = 0;
    return CF_OK;
  }
  if (!e.sims[0].have_prm) {
    // no step parameters yet: bin with a unit-cell grid spanning the particles
    return cf::fail(CF_ERR_STATE, "call cf_simulate (or set parameters) before building edges");
  }
  CFR(e.ensure_edge_bufs(er.maxc));
  CFR(e.prepare(er));
  CFR(e.edge_pass(2 * er.maxc));
  e.edges_maxc = er.maxc;
  return e.finish_call(true);
}

int cf_get_edges(cf_sim* sim, int32_t* pairs, long long capacity, long long* count) {
  CFR(check(sim));
  if (!count || (!pairs && capacity > 0)) return cf::fail(CF_ERR_ARG, "null argument");
  return sim->e->get_edges(pairs, capacity, count);
}

int cf_get_step_edges(cf_sim* sim, int32_t* pairs, long long capacity, long long* count) {
  CFR(check(sim));
  if (!count || (!pairs && capacity > 0)) return cf::fail(CF_ERR_ARG, "null argument");
  return sim->e->get_step_edges(pairs, capacity, count);
}

int cf_build_triangles(cf_sim* sim, float distance, int max_connections) {
  CFR(check(sim));
  if (!(distance >= 0.0f) || !std::isfinite(distance))
    return cf::fail(CF_ERR_ARG, "distance must be finite and >= 0");
  cf::Engine& e = *sim->e;
  if (!e.sims[0].have_prm)
    return cf::fail(CF_ERR_STATE, "call cf_simulate (or set parameters) before building a mesh");
  return e.tri_mesh(distance, max_connections);
}

int cf_get_triangles(cf_sim* sim, int32_t* tris, long long capacity, long long* count) {
  CFR(check(sim));
  if (!count || (!tris && capacity > 0)) return cf::fail(CF_ERR_ARG, "null argument");
  return sim->e->get_triangles(tris, capacity, count);
}

int cf_get_triangle_vertices(cf_sim* sim, const cf_color* colors, int ncolors, float* out,
                             long long capacity_floats, long long* vertex_count) {
  CFR(check(sim));
  if (!vertex_count || (!out && capacity_floats > 0) || (!colors && ncolors > 0))
    return cf::fail(CF_ERR_ARG, "null argument");
  return sim->e->triangle_vertices(colors, ncolors, out, nullptr, capacity_floats, vertex_count);
}

int cf_write_triangle_vertices_device(cf_sim* sim, const cf_color* colors, int ncolors,
                                      float* device_out, long long capacity_floats,
                                      long long* vertex_count) {
  CFR(check(sim));
  if (!vertex_count || (!device_out && capacity_floats > 0) || (!colors && ncolors > 0))
    return cf::fail(CF_ERR_ARG, "null argument");
  return sim->e->triangle_vertices(colors, ncolors, nullptr, device_out, capacity_floats,
                                   vertex_count);
}

int cf_get_line_vertices(cf_sim* sim, const cf_color* colors, int ncolors, float* out,
                         long long capacity_floats, long long* vertex_count) {
  CFR(check(sim));
  if (!vertex_count || (!out && capacity_floats > 0) || (!colors && ncolors > 0))
    return cf::fail(CF_ERR_ARG, "null argument");
  return sim->e->line_vertices(colors, ncolors, out, nullptr, capacity_floats, vertex_count);
}

int cf_write_line_vertices_device(cf_sim* sim, const cf_color* colors, int ncolors,
                                  float* device_out, long long capacity_floats,
                                  long long* vertex_count) {
  CFR(check(sim));
  if (!vertex_count || (!device_out && capacity_floats > 0) || (!colors && ncolors > 0))
    return cf::fail(CF_ERR_ARG, "null argument");
  return sim->e->line_vertices(colors, ncolors, nullptr, device_out, capacity_floats,
                               vertex_count);
}

int cf_get_grid(cf_sim* sim, int dims[3], int* reach) {
  CFR(check(sim));
  CFR(sim->e->flush());
  const cf::SimHost& h = sim->e->sims[0];
  for (int a = 0; a < 3; ++a) dims[a] = h.dims[a];
  if (reach) *reach = h.reach;
  return CF_OK;
}

int cf_get_sort_order(cf_sim* sim, uint32_t* cell_of_slot, uint32_t* index_of_slot, int count) {
  CFR(check(sim));
  if (!cell_of_slot || !index_of_slot || count != sim->e->sims[0].n)
    return cf::fail(CF_ERR_ARG, "bad arguments");
  return sim->e->sort_order(0, cell_of_slot, index_of_slot);
}

int cf_profile_steps(cf_sim* sim, const cf_params* params, int nsteps, const cf_edge_opts* edges,
                     double* ms_out, int max_kernels, int* nkernels, char* names_buf,
                     int names_cap) {
  CFR(check(sim));
  if (!ms_out || !nkernels || nsteps < 1) return cf::fail(CF_ERR_ARG, "bad arguments");
  CFR(check_params(params, sim->e->sims[0].T));
  cf::EdgeReq er;
  CFR(edge_req(edges, &er));
  sim->e->sims[0].prm = *params;
  sim->e->sims[0].have_prm = true;
  return sim->e->profile(nsteps, er, ms_out, max_kernels, nkernels, names_buf, names_cap);
}

using cf::check_tuning;

int cf_set_tuning(cf_sim* sim, const cf_tuning* tuning) {
  CFR(check(sim));
  CFR(check_tuning(tuning));
  sim->e->tun = *tuning;
  return CF_OK;
}

int cf_get_tuning(cf_sim* sim, cf_tuning* tuning) {
  CFR(check(sim));
  if (!tuning) return cf::fail(CF_ERR_ARG, "tuning is null");
  *tuning = sim->e->tun;
  return CF_OK;
}

int cf_autotune(cf_sim* sim, const cf_params* params, const cf_edge_opts* edges,
                int trial_steps, const char* cache_path, int flags, cf_tuning* chosen,
                double* us_per_step, char* report, int report_cap) {
  CFR(check(sim));
  if (trial_steps < 1 || (!report && report_cap > 0))
    return cf::fail(CF_ERR_ARG, "bad arguments");
  CFR(check_params(params, sim->e->sims[0].T));
  cf::EdgeReq er;
  CFR(edge_req(edges, &er));
  sim->e->sims[0].prm = *params;
  sim->e->sims[0].have_prm = true;
  return cf::tune_call(*sim->e, er, trial_steps, cache_path, flags, chosen, us_per_step, report,
                       report_cap);
}

int cf_tuning_key(cf_sim* sim, const cf_params* params, const cf_edge_opts* edges, char* key,
                  int key_cap) {
  CFR(check(sim));
  if (!key || key_cap < 1) return cf::fail(CF_ERR_ARG, "bad arguments");
  CFR(check_params(params, sim->e->sims[0].T));
  cf::EdgeReq er;
  CFR(edge_req(edges, &er));
  sim->e->sims[0].prm = *params;
  sim->e->sims[0].have_prm = true;
  std::string k;
  CFR(sim->e->tune_key(er, &k));
  if ((int)k.size() >= key_cap)
    return cf::fail(CF_ERR_CAPACITY, "key needs %d bytes", (int)k.size() + 1);
  std::snprintf(key, key_cap, "%s", k.c_str());
  return CF_OK;
}

int cf_pair_histogram(cf_sim* sim, int bins, float r_max, double* hist) {
  CFR(check(sim));
  if (!hist || bins < 1 || bins > cf::kMaxBins || !(r_max > 0.0f))
    return cf::fail(CF_ERR_ARG, "bins must be 1..%d and r_max > 0", cf::kMaxBins);
  return sim->e->histogram(bins, r_max, hist);
}

int cf_pair_histogram_noalloc(cf_sim* sim, int bins, float r_max, double* hist) {
  CFR(check(sim));
  if (!hist || bins < 1 || bins > cf::kMaxBins || !(r_max > 0.0f))
    return cf::fail(CF_ERR_ARG, "bins must be 1..%d and r_max > 0", cf::kMaxBins);
  return sim->e->histogram_noalloc(bins, r_max, hist);
}

int cf_rollout_grad(cf_sim* sim, const cf_params* params, int nsteps, int checkpoint_every,
                    int bins, float r_max, const double* target_hist, double* loss, double* hist,
                    double* grad_table, double* grad_ratio) {
  CFR(check(sim));
  if (nsteps < 1 || checkpoint_every < 1 || !target_hist || !loss || !grad_table || !grad_ratio ||
      bins < 1 || bins > cf::kMaxBins || !(r_max > 0.0f))
    return cf::fail(CF_ERR_ARG, "bad arguments");
  CFR(check_params(params, sim->e->sims[0].T));
  sim->e->sims[0].prm = *params;
  sim->e->sims[0].have_prm = true;
  return sim->e->rollout_grad(nsteps, checkpoint_every, bins, r_max, target_hist, loss, hist,
                              grad_table, grad_ratio);
}

// ---- ensemble ------------------------------------------------------------------------

int cf_ensemble_create(int nsims, const int* counts, const int* types, cf_ensemble** out) {
  if (!out || !counts || !types || nsims < 1) return cf::fail(CF_ERR_ARG, "bad arguments");
  *out = nullptr;
  cf_ensemble* e = new cf_ensemble();
  int rc = make_engine(e->e, 2, nsims, counts, types);
  if (rc != CF_OK) {
    delete e;
    return rc;
  }
  reg(e);
  *out = e;
  return CF_OK;
}

int cf_ensemble_destroy(cf_ensemble* ens) {
  CFR(check(ens));
  unreg(ens);
  delete ens;
  return CF_OK;
}

int cf_ensemble_set_sim(cf_ensemble* ens, int s, const cf_params* params, const float* table,
                        const float* rbt, const cf_particle* particles, int count) {
  CFR(check(ens));
  cf::Engine& e = *ens->e;
  if (s < 0 || s >= (int)e.sims.size()) return cf::fail(CF_ERR_ARG, "simulation index out of range");
  cf::SimHost& h = e.sims[s];
  CFR(check_params(params, h.T));
  if (!table || !rbt || !particles || count != h.n) return cf::fail(CF_ERR_ARG, "bad arguments");
  h.table.assign(table, table + h.T * h.T);
  h.rbt.assign(rbt, rbt + h.T);
  h.prm = *params;
  h.have_prm = true;
  return e.load_aos(s, particles);
}

int cf_ensemble_simulate(cf_ensemble* ens, int nsteps) {
  CFR(check(ens));
  if (nsteps < 0) return cf::fail(CF_ERR_ARG, "nsteps < 0");
  cf::EdgeReq none;
  return ens->e->simulate(nsteps, none);
}

int cf_ensemble_get_particles(cf_ensemble* ens, int s, cf_particle* particles, int count) {
  CFR(check(ens));
  cf::Engine& e = *ens->e;
  if (s < 0 || s >= (int)e.sims.size() || !particles || count != e.sims[s].n)
    return cf::fail(CF_ERR_ARG, "bad arguments");
  return e.unload_aos(s, particles);
}

int cf_ensemble_get_neighbor_counts(cf_ensemble* ens, int s, int32_t* counts, int count) {
  CFR(check(ens));
  cf::Engine& e = *ens->e;
  if (s < 0 || s >= (int)e.sims.size() || !counts || count != e.sims[s].n)
    return cf::fail(CF_ERR_ARG, "bad arguments");
  return e.get_counts(s, counts);
}

int cf_ensemble_set_tuning(cf_ensemble* ens, const cf_tuning* tuning) {
  CFR(check(ens));
  CFR(check_tuning(tuning));
  ens->e->tun = *tuning;
  return CF_OK;
}

int cf_ensemble_get_tuning(cf_ensemble* ens, cf_tuning* tuning) {
  CFR(check(ens));
  if (!tuning) return cf::fail(CF_ERR_ARG, "tuning is null");
  *tuning = ens->e->tun;
  return CF_OK;
}

int cf_ensemble_autotune(cf_ensemble* ens, int trial_steps, const char* cache_path, int flags,
                         cf_tuning* chosen, double* us_per_step, char* report, int report_cap) {
  CFR(check(ens));
  if (trial_steps < 1 || (!report && report_cap > 0))
    return cf::fail(CF_ERR_ARG, "bad arguments");
  cf::EdgeReq none;
  return cf::tune_call(*ens->e, none, trial_steps, cache_path, flags, chosen, us_per_step, report,
                       report_cap);
}

}  // extern "C"
