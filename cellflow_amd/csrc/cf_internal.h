// Internal layout shared by the HIP kernels (cf_kernels.hip) and the host engine
// (cf_engine.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cf {

constexpr int kMaxTypes = 10;          // MAX_PARTICLE_TYPES (SimulationParams.h:64)
constexpr int kWave = 64;              // CDNA wavefront
constexpr int kSlotAlign = 64;         // every simulation's slot range starts on a wave
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096 elements per scan tile
constexpr int kMaxEdgeCap = 64;        // 2 * CF_MAX_CONNECTIONS

// Step sort by cell tiles (k_tile_sort).  A cell tile is `tr` consecutive rows of one
// z-plane of one simulation.  A particle whose new cell lies in a tile was, before the
// step, in the tile's window (the tile's rows and one row either side, in the tile's plane
// and one plane either side, cyclically), unless it is a far mover: those carry kFarBit
// in their key and are listed apart.
constexpr int kTileCells = 2048;       // most cells in one tile (tr * nx)
constexpr int kTileCap = 2048;         // members a tile sorts in LDS at once
constexpr int kTileTarget = 384;       // expected members of a tile (chooses tr)
constexpr int kEdgeTile = 1024;        // owner ids per edge tile
// The list force kernels (variants 3 and 4) keep each candidate as slot | image-shift code
// << kCodeShift in LDS, so they take at most 2^kCodeShift slots.
constexpr int kCodeShift = 26;
constexpr uint32_t kFarBit = 0x80000000u;
// Tile counters sit one per 128-byte line: a line takes every atomic of the tiles it holds,
// one after another.
constexpr int kCtrStride = 32;
// Far movers are listed by destination tile (kFarCap per tile); a tile's excess goes to one
// shared overflow list that every tile reads.
constexpr int kFarCap = 256;
// In-range partners listed per particle for the adjoint's pair kernel (more: it walks again).
constexpr int kPairList = 64;

// Per-type-pair constants of one simulation, T*T entries, row = own type.
struct PairConst {
  float thr;       // smallest float t with sqrtf(t) >= effR (cutoff test: s + 1e-4f < thr)
  float fv;        // force table value (ParticleSimulation.cu:116)
  float inv_eff;   // 1 / effR (continuous part only)
  float pad;       // row-wide: largest squared reach of the owner type (pruning radius)
};

// Device descriptor of one simulation.  Every wave covers 64 slots of one simulation, so
// the descriptor is wave-uniform and lives in scalar registers.
struct SimDesc {
  int n;            // particles
  int p0;           // first slot (multiple of 64); slot p0 + k, id p0 + local index
  int c0;           // first cell in the concatenated cell space (the pad cell is c0+ncells)
  int ncells;       // nx * ny * nz
  int nx, ny, nz;
  int T;
  int generic;      // 1: some axis has fewer than 2*reach+1 cells -> per-pair minimum image
  int reach;        // force stencil reach in cells (1 or 2)
  int tab;          // offset of this simulation's PairConst table
  int edge_on;      // edges computed inside the force pass
  int ident;        // no parameters yet: one cell, slots in index order, nothing to sort
  float Lx, Ly, Lz;
  float ihx, ihy, ihz;      // n / L in FP32 (binning)
  float dt, friction, fm, one_minus_balance, max_exp;
  float rep, att, kexp2;    // repulsion, attraction, -k*log2(e)
  float hx, hy, hz;          // cell sizes L / n in FP32 (stencil pruning)
  int edge_inside;          // bit t: for owner type t the edge test implies the cutoff test
  float edge_d2;            // proximityDistance^2 in FP32 (ParticleSimulation.cu:660)
  int edge_maxc;            // min(maxConn, 32)
  int edge_rx, edge_ry, edge_rz;  // non-periodic edge stencil reach (standalone edge pass)
  int t0;                   // first cell tile of this simulation
  int tr;                   // rows per cell tile
  int nyt;                  // cell tiles per z-plane, ceil(ny / tr)
};

// Decoupled look-back scan plan: one 64-bit ticket counter and one flag word per tile.
// A flag word is (epoch:30 | status:2 | value:32); tickets never reset, so no kernel has
// to clear scan state between launches (also under HIP graph replay).
struct ScanSeg {
  const uint32_t* in;
  uint32_t* out;      // n + 1 entries, exclusive prefix, out[n] = total
  int n;
  int tiles;
};

struct ScanPlan {
  ScanSeg seg[2];
  int nseg;
  int tiles_total;
  unsigned long long* ticket;   // device counter
  unsigned long long* flags;    // tiles_total words
};

// Buffers of one engine (all device pointers).
struct DevBufs {
  // sorted state, double-buffered (A = current)
  float4* pos[2];       // x, y, z, type bits
  float4* vel[2];       // vx, vy, vz, previous neighbour count bits
  uint32_t* id[2];      // global index of the particle in each slot
  uint32_t* start[2];   // cell start slots, C + 1 entries
  // step temporaries (old-slot order)
  float4* tpos;
  float4* tvel;
  uint32_t* tid;
  uint32_t* key;        // new global cell of each old slot
  uint32_t* arrival;    // arrival rank inside the new cell (atomic order)
  uint32_t* hist;       // per-cell counts (pad cells hold their pad size)
  uint32_t* tmp_id;     // particle index per new slot, cell ranges unsorted
  float4* acc;          // acceleration by index (output-only, ParticleSimulation.cu:146)
  float* pad;           // the Particle.pad field by index, carried unchanged
  // edges
  float* ecand_d2;      // [cap][P]
  uint32_t* ecand_id;   // [cap][P]
  uint32_t* estage;     // [P][maxc] kept neighbour index by owner and rank
  uint32_t* ecount;     // [P + 1] kept edges per index (standalone pass, mesh)
  uint32_t* fcount;     // [P + 1] kept edges per index of the graph a step builds inside
                        // its force pass: zero between steps (the force pass writes the
                        // non-zero counts, the sort that reads them clears them)
  uint32_t* eoff;       // [P + 1] exclusive prefix of ecount (or of fcount)
  int2* edges;          // canonical pairs of the standalone edge pass (the current state)
  uint32_t* gtotal;     // number of pairs in `edges`
  int2* sedges;         // canonical pairs the step's sort writes (the pre-step state)
  uint32_t* stotal;     // their number
  // descriptors
  SimDesc* sims;
  uint32_t* wave_sim;   // simulation index of every 64-slot wave
  PairConst* pairs;
  uint32_t* flag;       // bit 0: non-finite state seen
};

// ---- kernel arguments ----
struct AosParticle {
  float px, py, pz, vx, vy, vz, ax, ay, az;
  uint32_t ptype;
  float pad;
};
static_assert(sizeof(AosParticle) == 44, "Particle record must be 44 bytes");

struct ForceArgs {
  const float4* __restrict__ pos;
  const float4* __restrict__ vel;
  const uint32_t* __restrict__ id;
  const uint32_t* __restrict__ start;
  float4* __restrict__ tpos;
  float4* __restrict__ tvel;
  uint32_t* __restrict__ key;
  uint32_t* __restrict__ arrival;
  uint32_t* hist;
  float4* __restrict__ acc;  // null: not stored (a step whose acceleration nobody reads)
  const SimDesc* __restrict__ sims;
  const uint32_t* __restrict__ wave_sim;
  const PairConst* __restrict__ pairs;
  uint32_t* flag;
  float* ecand_d2;
  uint32_t* ecand_id;
  uint32_t* estage;  // kept partners: owner o's rank-r partner at o * estride + r
  uint32_t* ecount;
  long P;
  long estride;     // words per owner record in estage (the largest maxc allocated)
  int nwaves;
  int edges;        // fused edge candidates requested for this launch
  int edge_cap;
  // 1: the tile sort follows (keys with kFarBit, tile counts, far lists, estage by owner
  // index); 0: the three-pass counting sort follows (arrival ranks, estage by slot)
  int tile_sort;
  // tile sort inputs (this step's counters; the sort clears the next step's)
  uint32_t* tcount;     // members per cell tile (tile t at t * kCtrStride)
  uint32_t* far_n;      // far movers in the overflow list
  uint32_t* far_slot;   // their pre-step slots
  uint32_t* ftn;        // far movers per tile (tile t at t * kCtrStride)
  uint32_t* ftl;        // their pre-step slots (tile t at t * kFarCap)
  // for the gradient rollout (variant 0, no edges; null otherwise): each particle's
  // in-range partners (pre-step slots, in slot order, at most kPairList of them, at
  // hits[k * kPairList]) and their number nhits[k], k = slot - first slot
  uint32_t* hits;
  uint32_t* nhits;
};

// The standalone proximity-graph pass (k_edges_g); k_tile_sort's edge tiles then write the
// canonical list from estage and ecount.
struct EdgeArgs {
  const float4* __restrict__ pos;
  const uint32_t* __restrict__ id;
  const uint32_t* __restrict__ start;
  const SimDesc* __restrict__ sims;
  const uint32_t* __restrict__ wave_sim;
  float* ecand_d2;      // [cap][P] by slot: scratch of the streaming selection
  uint32_t* ecand_id;
  uint32_t* estage;     // kept partners: owner index o's rank r at o * estride + r
  uint32_t* ecount;     // kept partners per owner index: non-zero counts only (zero between
                        // passes: the edge tiles clear what they read)
  long P;
  long estride;
  int nwaves;
  int edge_cap;
};

struct SortArgs {
  const float4* __restrict__ src_pos;
  const float4* __restrict__ src_vel;
  const uint32_t* __restrict__ src_id;
  const uint32_t* __restrict__ key;
  const uint32_t* __restrict__ arrival;
  const uint32_t* __restrict__ start;   // new cell starts
  uint32_t* hist;
  uint32_t* tmp_id;
  float4* __restrict__ dst_pos;
  float4* __restrict__ dst_vel;
  uint32_t* __restrict__ dst_id;
  const SimDesc* __restrict__ sims;
  const uint32_t* __restrict__ wave_sim;
  // canonical edge output (k_fill only; null when no edges this step)
  const uint32_t* __restrict__ estage;  // slot p's rank-r partner at p * estride + r
  const uint32_t* __restrict__ eoff;
  int2* __restrict__ edges;
  uint32_t* etotal;     // k_fill: receives the number of edges written (eoff[P])
  uint32_t* fcount;     // k_fill: the step's per-owner edge counts, cleared after use
  long P;
  long estride;
  int nwaves;
};

// One step's sort after the force pass (k_tile_sort): blocks [0, ntiles) each place the
// particles of one cell tile, blocks [ntiles, ntiles + netiles) each write the canonical
// edge list of one edge tile.
struct TileSortArgs {
  const float4* __restrict__ src_pos;   // new state in pre-step slot order
  const float4* __restrict__ src_vel;
  const uint32_t* __restrict__ src_id;
  const uint32_t* __restrict__ key;     // new cell (| kFarBit) by pre-step slot
  uint32_t* hist;                       // per-cell counts (read, then cleared)
  const uint32_t* __restrict__ start_old;
  uint32_t* start_new;
  float4* dst_pos;
  float4* dst_vel;
  uint32_t* dst_id;
  uint32_t* tmp_id;                     // scratch of the one-cell fallback: ids
  uint32_t* tmp_slot;                   // and their pre-step slots
  const SimDesc* __restrict__ sims;
  const uint32_t* __restrict__ tile_sim;  // simulation of each cell tile
  const uint32_t* tcount;
  uint32_t* tcount_next;
  const uint32_t* far_n;
  uint32_t* far_n_next;
  const uint32_t* far_slot;
  const uint32_t* ftn;
  uint32_t* ftn_next;
  const uint32_t* ftl;
  uint32_t* flag;                       // bit 1: a cell outgrew the LDS; bit 2: lost particle
  int ntiles;
  // edges (edge tiles; null edges: no edge tiles)
  uint32_t* ecount;                     // per-owner counts: read, then cleared
  // the edge tiles' decoupled look-back: a ticket counter that every edge-tile block of
  // every launch draws from (tile = ticket % netiles, launch epoch = ticket / netiles) and
  // one flag word per edge tile (epoch, status, value); both start at zero and are never
  // reset while netiles stays the same
  unsigned long long* eticket;
  unsigned long long* eflags;
  const uint32_t* estage;               // owner index o's rank r at o * estride + r
  long estride;
  int2* edges;
  uint32_t* etotal;                     // the last edge tile writes the edge count here
  long edge_room;                       // entries `edges` holds
  const uint32_t* __restrict__ wave_sim;
  int netiles;
  long P;
  int one_sim;                          // one simulation: tile_sim is all zero
};

struct BinArgs {
  const float4* __restrict__ pos;
  uint32_t* key;
  uint32_t* arrival;
  uint32_t* hist;
  const SimDesc* __restrict__ sims;
  const uint32_t* __restrict__ wave_sim;
  int nwaves;
};

struct IoArgs {
  AosParticle* aos;        // count records of one simulation
  const float4* pos;
  const float4* vel;
  const uint32_t* id;
  float4* tpos;
  float4* tvel;
  uint32_t* tid;
  float4* acc;
  float* pad;
  const int32_t* counts_in;   // optional prev counts (index order)
  int32_t* counts_out;
  uint32_t* cell_out;
  uint32_t* index_out;
  const SimDesc* sims;
  int sim;
  float shift[3];          // universe shift (k_move)
  float box[3];
};

struct InitArgs {
  float4* tpos;
  float4* tvel;
  uint32_t* tid;
  float4* acc;
  float* pad;
  const SimDesc* __restrict__ sims;
  int sim;
  unsigned long long seed;
  float cw, ch, cd, spawn;
};

struct LineArgs {
  const int2* __restrict__ edges;
  const float4* __restrict__ pos;
  const uint32_t* __restrict__ slot_of;   // index -> slot
  const float* __restrict__ colors;       // ncolors x 3
  float* out;
  long nedges;
  int ncolors;
  int T;
  int p0;
};

// Triangle mesh of the current state (generateTriangleMeshKernel): per owner slot, the
// kept same-type neighbours and the owner's triangle count, then the canonical triangles.
struct TriArgs {
  const float4* __restrict__ pos;
  const uint32_t* __restrict__ id;
  const uint32_t* __restrict__ start;
  const SimDesc* __restrict__ sims;
  const uint32_t* __restrict__ wave_sim;
  float* cand_d2;        // [maxc][P] scratch of the index selection
  uint32_t* cand_id;
  uint32_t* sorted;      // [maxc][P] each owner's kept neighbours, ascending
  uint32_t* nsel;        // [P] their number, by owner slot
  uint32_t* count;       // [P] triangles of each owner, by owner index
  const uint32_t* off;   // exclusive scan of count
  int32_t* tris;         // 3 indices per triangle
  long P;
  int nwaves;
  int maxc;              // min(maxConnectionsPerParticle, 32)
  float d2;              // squared distance threshold, hysteresis included
  int rx, ry, rz;        // non-periodic stencil reach in cells
};

struct TriVertArgs {
  const int32_t* __restrict__ tris;
  const float4* __restrict__ pos;
  const uint32_t* __restrict__ slot_of;   // index -> slot
  const float* __restrict__ colors;       // ncolors x 3
  float* out;                             // 27 floats per triangle
  long ntris;
  int ncolors;
  int T;
};
hipError_t launch_tri_count(const TriArgs& a, hipStream_t st);
hipError_t launch_tri_write(const TriArgs& a, hipStream_t st);
hipError_t launch_tri_vertices(const TriVertArgs& a, hipStream_t st);

// ---- differentiable rollout (cf_adjoint.hip) ----
constexpr int kMaxBins = 64;

struct AdjConst {
  int T;
  double radius, ratio, rep, att, k;
  double dt, friction, fm, balance, max_exp;  // the step's scalars as FP64 (from FP32)
  double rbt[kMaxTypes];
  double table[kMaxTypes * kMaxTypes];
};

struct HistArgs {
  const float4* pos;
  const uint32_t* id;
  const uint32_t* start;
  const SimDesc* sims;
  int bins;
  double r_max;
  int reach;            // cells covering r_max + 2 bin widths
  double* partial;      // [bins][stride] per-particle histograms
  long stride;
  const double* gh;     // dL/dh_b
  double4* lam_x;
  double4* lam_v;
  // optional FP64 positions by particle index (the rollout's FP64 shadow), used for the
  // distances instead of `pos`, which still gives the cells; `slack` bounds how far they
  // may lie from `pos` (it widens the cell reach and the FP32 pre-test)
  const double4* x64;
  double slack;
};

struct AdjArgs {
  const float4* pos;        // pre-step state S_k (sorted)
  const float4* vel;
  const uint32_t* id;
  const uint32_t* start;
  const float4* tvel;       // the step's new velocities by pre-step slot, .w = its count bits
  const SimDesc* sims;
  const float* thr;         // T*T cutoff thresholds of the forward
  AdjConst c;
  const double4* lam_x_next;
  const double4* lam_v_next;
  double4* lam_x;
  double4* lam_v;
  double4* phi;
  uint32_t* hits;           // [n][64] in-range slots of each particle, slot order (walk)
  uint32_t* nhits;          // [n] their number (more than 64: the pair kernel walks again)
  double* partial;          // [T*T + 1][stride] parameter partials, one column per block,
                            // added to (the caller zeroes them once per rollout)
  long stride;              // >= adj_pair_blocks(n)
  // the FP64 shadow of the pre-step state, by particle index, and (shadow step) of the
  // post-step state
  const double4* x64;
  const double4* v64;
  double4* x64_next;
  double4* v64_next;
  uint32_t* flag;           // shadow drift check: bit 3 when the shadow strays beyond slack/2
  double slack;
};

hipError_t launch_hist(const HistArgs& a, int n, hipStream_t st);
hipError_t launch_hist_grad(const HistArgs& a, int n, hipStream_t st);
hipError_t launch_reduce_rows(const double* in, long stride, int n, int rows, double scale,
                              double* out, int accumulate, hipStream_t st);
// One backward step: with `walk`, the FP32 walk lists each particle's in-range partners
// (and computes the density-factor terms); without, the lists in AdjArgs::hits came from
// the recompute's force kernel.
hipError_t launch_adj_step(const AdjArgs& a, int n, hipStream_t st, bool walk);
// The rollout's FP64 shadow: the pre-step FP32 state promoted (index order); one FP64 step
// with the decisions of the FP32 pre-step state; the drift check of a shadow state against
// the FP32 state it shadows.
hipError_t launch_shadow_init(const AdjArgs& a, int n, hipStream_t st);
hipError_t launch_shadow_step(const AdjArgs& a, int n, hipStream_t st);
hipError_t launch_shadow_drift(const AdjArgs& a, int n, hipStream_t st);
int adj_pair_blocks(int n);  // blocks (columns of AdjArgs::partial) of one backward step

// Per-kernel timing (cf_profile_steps): the next kernel launch of this thread through the
// launchers below records its own start and end into `start` and `stop`
// (hipExtLaunchKernel, timestamps of the dispatch itself: no marker between kernels), and
// the pair is then cleared.  Never armed during graph capture.
void arm_kernel_timer(hipEvent_t start, hipEvent_t stop);

// host-side launchers (cf_kernels.hip)
hipError_t launch_force(const ForceArgs& a, bool edges, int block, int variant, int reach,
                        int lanes, hipStream_t st);
hipError_t launch_edges(const EdgeArgs& a, hipStream_t st);
hipError_t launch_tile_sort(const TileSortArgs& a, hipStream_t st);
hipError_t launch_scan(const ScanPlan& p, hipStream_t st);
hipError_t launch_fill(const SortArgs& a, hipStream_t st);
hipError_t launch_rank_scatter(const SortArgs& a, hipStream_t st);
hipError_t launch_bin(const BinArgs& a, hipStream_t st);
hipError_t launch_io(int which, const IoArgs& a, int n, hipStream_t st);
hipError_t launch_init(const InitArgs& a, int n, hipStream_t st);
hipError_t launch_slot_of(const uint32_t* id, uint32_t* slot_of, const SimDesc* sims, int sim, int n,
                          hipStream_t st);
hipError_t launch_line_vertices(const LineArgs& a, hipStream_t st);

}  // namespace cf
