/* cellflow.h — C ABI of the MI355X particle-life step engine.
 *
 * This is the drop-in boundary for the reference's simulation-step path, the C++ class
 * `ParticleSimulation` (cuda-native/include/ParticleSimulation.cuh:10-103, implemented in
 * cuda-native/src/ParticleSimulation.cu:426-754).  Each entry point below names the
 * reference member it replaces.  Plain C types and pointers to host memory (the
 * *_device entry points alone take caller-owned device memory); every call returns
 * an int status (CF_OK on success) and never exits the process (the reference's
 * CUDA_CHECK calls exit(1), ParticleSimulation.cu:10-18).  A handle is not thread-safe;
 * separate handles are independent.
 *
 * Semantics that differ from the reference on purpose:
 *   - the step is a Jacobi update (every particle reads the pre-step state); the
 *     reference updates `particles[idx]` in place while other threads read it
 *     (ParticleSimulation.cu:89 vs :164), which makes its output schedule-dependent;
 *   - proximity-graph edges are returned as (i, j) index pairs in canonical order
 *     (owner i ascending, then distance rank) instead of the atomicAdd order of
 *     ParticleSimulation.cu:250;  cf_get_line_vertices expands them to the reference's
 *     6-float line vertices (:253-275);
 *   - max_connections in (16, 32] uses a candidate list of 2*max_connections entries,
 *     where the reference overflows its 32-entry local array (:210-215).
 */
#ifndef CELLFLOW_H
#define CELLFLOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CF_OK 0
#define CF_ERR_ARG 1        /* invalid argument (sizes, types, null pointers) */
#define CF_ERR_DEVICE 2     /* no usable GPU, or a HIP runtime error */
#define CF_ERR_CAPACITY 3   /* output buffer too small; the needed size is returned */
#define CF_ERR_NONFINITE 4  /* a position or velocity became NaN/Inf during the step */
#define CF_ERR_HANDLE 5     /* handle is null, destroyed, or of the wrong kind */
#define CF_ERR_STATE 6      /* call not valid in the current state (e.g. no edges built) */

#define CF_MAX_TYPES 10     /* MAX_PARTICLE_TYPES, cuda-native/include/SimulationParams.h:64 */
#define CF_MAX_CONNECTIONS 32

/* struct Particle, SimulationParams.h:6-12: 44 bytes, float3 members 4-byte aligned. */
typedef struct cf_particle {
  float pos[3];
  float vel[3];
  float acc[3];
  uint32_t ptype;
  float pad;
} cf_particle;

/* struct ParticleColor, SimulationParams.h:59-61. */
typedef struct cf_color {
  float r, g, b;
} cf_color;

/* The fields of SimulationParams (SimulationParams.h:14-57) that the step reads
 * (simulateParticlesKernel, ParticleSimulation.cu:69-166).  Render-only fields are not
 * part of the step and are omitted. */
typedef struct cf_params {
  float radius;                 /* SimulationParams.h:15 */
  float delta_t;                /* :16 */
  float friction;               /* :17 */
  float repulsion;              /* :18 */
  float attraction;             /* :19 */
  float k;                      /* :20 */
  float balance;                /* :21 */
  float canvas_width;           /* :22 */
  float canvas_height;          /* :23 */
  float canvas_depth;           /* :24 */
  int num_particle_types;       /* :26 */
  float ratio_with_lfo;         /* :27 */
  float force_multiplier;       /* :28 */
  int max_expected_neighbors;   /* :29 */
} cf_params;

/* Proximity-graph settings (CellFlowWidget.cpp:569-575 passes these per frame). */
typedef struct cf_edge_opts {
  float distance;               /* proximityDistance */
  int max_connections;          /* maxConnectionsPerParticle, 1..32 */
} cf_edge_opts;

/* Launch/grid choices made by the autotuner (no reference counterpart). */
typedef struct cf_tuning {
  int cell_divisor;             /* 1: cell >= max cutoff, 27-cell stencil; 2: half cells, 125 */
  int force_variant;            /* 0: G lanes per particle walking the stencil's runs;
                                   1: one lane per particle; 2: LDS-staged stencil, one wave
                                   per workgroup; 3: G lanes per particle over a per-particle
                                   candidate list in LDS (G >= 8 with cell_divisor 2); 4: 3
                                   with each workgroup's integration done by one wave (G 8
                                   or 16) */
  int lanes_per_particle;       /* G of variants 0, 3 and 4: 4, 8, 16 or 32 */
  int block_size;               /* threads per workgroup of variant 1: 64, 128, 256 */
  int steps_per_graph;          /* steps per captured HIP graph (rounded up to even);
                                   1 launches every kernel directly */
  int sort_variant;             /* the step's (cell, index) sort: 1 (default) one pass over
                                   cell tiles, each placing the particles its window holds;
                                   0 three passes (scan, fill, rank), the faster one for
                                   strongly clustered states */
} cf_tuning;

typedef struct cf_sim cf_sim;
typedef struct cf_ensemble cf_ensemble;

/* ---- lifetime -------------------------------------------------------------------- */

/* ParticleSimulation(int particleCount), ParticleSimulation.cuh:12 / .cu:427-435.
 * Allocates device buffers on the current HIP device.  The state starts at zero with
 * `num_types` types; load it with cf_set_particles or cf_init_particles (the reference
 * seeds it from time(nullptr), .cu:480). */
int cf_create(int particle_count, int num_types, cf_sim** out);
/* ~ParticleSimulation, .cu:447-449. */
int cf_destroy(cf_sim* sim);
/* setParticleCount, .cu:564-572: reallocates; neighbour counts return to zero. */
int cf_set_particle_count(cf_sim* sim, int particle_count);
/* setNumParticleTypes, .cu:574-579 (the caller reloads table, radii and state). */
int cf_set_num_particle_types(cf_sim* sim, int num_types);
int cf_get_particle_count(cf_sim* sim, int* out);      /* getParticleCount, .cuh:40 */
int cf_get_num_particle_types(cf_sim* sim, int* out); /* getNumParticleTypes, .cuh:43 */

/* ---- particle state ---------------------------------------------------------------- */

/* No reference counterpart (state only comes from initializeParticlesKernel there);
 * needed to drive the engine from a given state. Resets neighbour counts to zero. */
int cf_set_particles(cf_sim* sim, const cf_particle* particles, int count);
/* getParticleData, .cu:558-562: the full 44-byte records in index order. */
int cf_get_particles(cf_sim* sim, cf_particle* particles, int count);
/* Structure-of-arrays forms: pos/vel are count x 3 floats, type is count uint32. */
int cf_set_state_soa(cf_sim* sim, const float* pos, const float* vel, const uint32_t* type,
                     int count);
int cf_get_state_soa(cf_sim* sim, float* pos, float* vel, uint32_t* type, int count);
/* The neighbour-count ping-pong buffers d_neighborCountsA/B (.cu:456-462, :544-555):
 * the counts the next step reads as `prevNeighborCount` (.cu:136). */
int cf_set_neighbor_counts(cf_sim* sim, const int32_t* counts, int count);
int cf_get_neighbor_counts(cf_sim* sim, int32_t* counts, int count);
/* Seeded replacement for initCurandKernel + initializeParticlesKernel (.cu:21-66):
 * positions uniform in the central min(spawn_region, canvas) cube, zero velocity and
 * acceleration, type uniform in [0, num_types).  Counter-based (SplitMix64), so the same
 * seed gives the same state on every device. */
int cf_init_particles(cf_sim* sim, uint64_t seed, const float canvas[3], float spawn_region);

/* moveUniverse (ParticleSimulation.cu:585-592, moveParticlesKernel :169-185): every
 * position becomes fmodf(x + d + L, L) per axis in FP32 with L = canvas (the reference
 * passes a default-constructed SimulationParams' 8000^3 canvas and dz = 0).  Velocities,
 * counts and accelerations are kept; the state is re-binned. */
int cf_move_universe(cf_sim* sim, float dx, float dy, float dz, const float canvas[3]);

/* ---- force table and per-type radius modifiers -------------------------------------- */

/* d_forceTable upload (.cu:529-530) with a caller-given final num_types^2 table
 * (row = own type, column = other type); values are used as given (no clamp). */
int cf_set_force_table(cf_sim* sim, const float* table, int num_types);
int cf_get_force_table(cf_sim* sim, float* table, int count);
/* getRawForceTableValues (.cuh:47) is a writable pointer in the reference; here it is a
 * set/get pair on the host-side raw table. */
int cf_set_raw_force_table(cf_sim* sim, const float* raw, int count);
int cf_get_raw_force_table(cf_sim* sim, float* raw, int count);
/* updateForceTable(range, bias, offset), .cu:521-531: table = clamp(tanh(raw*offset)*range
 * + bias, -1, 1) from the raw table. */
int cf_update_force_table(cf_sim* sim, float range, float bias, float offset);
/* radioByType: initializeRadioByType upload (.cu:533-539) with given values. */
int cf_set_radio_by_type(cf_sim* sim, const float* values, int count);
/* setRadioByTypeValue, .cu:615-622: out-of-range indices are ignored (returns CF_OK). */
int cf_set_radio_by_type_value(cf_sim* sim, int index, float value);
/* getRadioByType, .cu:607-613. */
int cf_get_radio_by_type(cf_sim* sim, float* values, int count);
/* rotateRadioByType, .cu:594-605. */
int cf_rotate_radio_by_type(cf_sim* sim);

/* ---- the step and the proximity graph ------------------------------------------------ */

/* `nsteps` calls of simulate(params) (.cu:541-556).  When `edges` is non-null the
 * proximity graph of every post-step state is built too, as CellFlowWidget::paintGL does
 * after each step (CellFlowWidget.cpp:551-575); the graph of the last state is kept for
 * cf_get_edges.  Synchronous like the reference (it waits for the GPU); returns
 * CF_ERR_NONFINITE if any position or velocity became NaN/Inf. */
int cf_simulate(cf_sim* sim, const cf_params* params, int nsteps, const cf_edge_opts* edges);
/* generateProximityGraph, .cu:624-687, on the current state (the GL VBO is replaced by
 * an internal buffer read with cf_get_edges / cf_get_line_vertices). */
int cf_build_edges(cf_sim* sim, float distance, int max_connections);
/* The last built edge list: `pairs` holds capacity (i, j) int32 pairs.  Returns
 * CF_ERR_CAPACITY with *count set to the needed number when capacity is too small. */
int cf_get_edges(cf_sim* sim, int32_t* pairs, long long capacity, long long* count);
/* The reference's line-vertex layout (.cu:253-275): per edge two vertices of
 * (x, y, z, r, g, b), colour = colors[ptype_i % num_types]; *vertex_count = 2 * edges
 * (the value generateProximityGraph returns through outVertexCount). */
int cf_get_line_vertices(cf_sim* sim, const cf_color* colors, int ncolors, float* out,
                         long long capacity_floats, long long* vertex_count);

/* generateTriangleMesh, .cu:689-754 (its kernel :279-424), on the current state: for each
 * owner i, the first min(max_connections, 32) particles of its type in index order within
 * 1.2 * distance (non-periodic), and the fan triangles (i, a, b) over the pairs a < b of
 * that list with i < a, at most 64 per owner.  Triangles are kept in canonical order
 * (owner ascending, then fan order) instead of the reference's atomicAdd order (:353),
 * and read with cf_get_triangles / cf_get_triangle_vertices (the GL VBO is replaced).
 * CF_ERR_ARG for a negative or non-finite distance, or above 67,108,863 particles (the
 * triangle offsets are 32-bit); CF_ERR_STATE before the first cf_simulate. */
int cf_build_triangles(cf_sim* sim, float distance, int max_connections);
/* The last built mesh: `tris` holds capacity triangles of three int32 indices.  Returns
 * CF_ERR_CAPACITY with *count set to the needed number when capacity is too small. */
int cf_get_triangles(cf_sim* sim, int32_t* tris, long long capacity, long long* count);
/* The reference's triangle-vertex layout (.cu:362-418): per triangle three vertices of
 * (x, y, z, nx, ny, nz, r, g, b), the flat normal normalised when longer than 1e-4,
 * colour = colors[ptype_owner % num_types]; *vertex_count = 3 * triangles (the value
 * generateTriangleMesh returns through outVertexCount). */
int cf_get_triangle_vertices(cf_sim* sim, const cf_color* colors, int ncolors, float* out,
                             long long capacity_floats, long long* vertex_count);

/* The GL buffer writes of generateProximityGraph and generateTriangleMesh (.cu:633-643 map
 * the VBO, the kernels write vertices into it at :253-275 and :362-418, :681-686 unmap):
 * the same bytes as cf_get_line_vertices / cf_get_triangle_vertices, written by the GPU
 * straight into caller-owned device memory, for instance the pointer
 * hipGraphicsResourceGetMappedPointer returns for a GL buffer the caller registered and
 * mapped.  No host copy is made; the call returns once the writes are done.  `device_out`
 * must be device or managed memory of the engine's device, or pinned host memory
 * (CF_ERR_ARG otherwise, checked before any kernel runs); CF_ERR_CAPACITY when
 * capacity_floats, or the allocation behind the pointer where the runtime can tell, is too
 * small (*vertex_count is still set).  CF_ERR_STATE when no graph / mesh has been built. */
int cf_write_line_vertices_device(cf_sim* sim, const cf_color* colors, int ncolors,
                                  float* device_out, long long capacity_floats,
                                  long long* vertex_count);
int cf_write_triangle_vertices_device(cf_sim* sim, const cf_color* colors, int ncolors,
                                      float* device_out, long long capacity_floats,
                                      long long* vertex_count);

/* ---- introspection (parity checks) -------------------------------------------------- */

/* Grid of the current binning: cells per axis and the reach of the force stencil. */
int cf_get_grid(cf_sim* sim, int dims[3], int* reach);
/* The canonical (cell, index) order of the current binning: for slot k, the particle
 * index and its linear cell id (x fastest). */
int cf_get_sort_order(cf_sim* sim, uint32_t* cell_of_slot, uint32_t* index_of_slot, int count);
/* The proximity graph the last cf_simulate / cf_profile_steps call built inside its final
 * step's force pass and sort: that of the state before the final step (the graph of every
 * intermediate post-step state is built this way and replaced by the next one; the state
 * after the final step gets its graph from the standalone pass that cf_get_edges reads).
 * Same layout and order as cf_get_edges.  CF_ERR_STATE when the last call built none:
 * fewer than two steps, no edges, or an edge distance beyond the force stencil (then
 * every step uses the standalone pass). */
int cf_get_step_edges(cf_sim* sim, int32_t* pairs, long long capacity, long long* count);
/* Per-kernel average durations (ms) over `nsteps` steps timed with HIP events on the
 * engine's stream: names is a comma-separated list written into names_buf. */
int cf_profile_steps(cf_sim* sim, const cf_params* params, int nsteps, const cf_edge_opts* edges,
                     double* ms_out, int max_kernels, int* nkernels, char* names_buf,
                     int names_cap);

/* ---- differentiable rollout (no reference counterpart) ------------------------------- */

/* Soft pair histogram of the current state: h_b = (1/N) sum_i sum_{j != i} B3((r_ij - c_b)/w),
 * w = r_max / bins, c_b = (b + 1/2) w, B3 the cubic B-spline on [-2, 2], r_ij the minimum-
 * image distance.  bins <= 64.  FP64, deterministic. */
int cf_pair_histogram(cf_sim* sim, int bins, float r_max, double* hist);
/* cf_pair_histogram without allocating device memory: no counterpart in the reference.
 * For callers that cannot allocate during their loop (serving loops, harnesses that forbid
 * allocation).  The partial sums live in one 32 MiB device array per GPU, held until the
 * process ends: bins * (particles rounded up to 128) + bins doubles must fit in it
 * (CF_ERR_ARG otherwise, before anything runs; about 131,000 particles at 32 bins).  The
 * state must be binned: after loading particles, step once (cf_simulate with 0 steps bins
 * without stepping), else CF_ERR_STATE.  Same result as cf_pair_histogram, bit for bit.
 * Calls that overlap on one GPU share the array and race. */
int cf_pair_histogram_noalloc(cf_sim* sim, int bins, float r_max, double* hist);
/* Runs `nsteps` steps from the current state (leaving the simulation at the final state)
 * and returns the loss L = sum_b (h_b - target_b)^2 of the final histogram and its gradient
 * with respect to the num_types^2 force-table entries (row-major) and ratio_with_lfo, by
 * a hand-written adjoint with state checkpoints every `checkpoint_every` steps.  The
 * rollout's discrete decisions (which pairs interact, neighbour counts) are those of the
 * FP32 steps and are held fixed; the loss, the histogram and the gradient are of the FP64
 * rollout along those decisions (an FP64 shadow of the state, stepped beside the FP32
 * one), so the gradient is that of one smooth FP64 function of the table and the ratio.
 * CF_ERR_STATE when the shadow strays from the FP32 state by more than 1/2000 of the
 * largest box side (a rollout too chaotic for a gradient with frozen decisions).
 * The call's device buffers (checkpoints, one segment's states, adjoint state) stay
 * allocated for the next call of the same shape and are freed with the handle. */
int cf_rollout_grad(cf_sim* sim, const cf_params* params, int nsteps, int checkpoint_every,
                    int bins, float r_max, const double* target_hist, double* loss,
                    double* hist, double* grad_table, double* grad_ratio);

/* ---- tuning ------------------------------------------------------------------------- */

int cf_set_tuning(cf_sim* sim, const cf_tuning* tuning);
int cf_get_tuning(cf_sim* sim, cf_tuning* tuning);

/* Autotuner (no reference counterpart).  Picks the fastest cell divisor, force-kernel
 * variant, lanes per particle and block size for this simulation's workload and applies
 * it (steps_per_graph is kept).  Each candidate runs `trial_steps` timed steps (HIP events,
 * whole step) on a copy of the current state; the state and neighbour counts are restored
 * afterwards, so tuning never changes a trajectory.  An edge list built before the call
 * is discarded.
 *
 * With a non-null `cache_path` the result is looked up first under the workload's bucket
 * key (device arch, log2 particle count, half-octave density, types, edge cap; see
 * cf_tuning_key) and stored there after measuring; flags CF_TUNE_FORCE re-measures even
 * on a hit.  `report` (optional, report_cap bytes) receives a JSON object with the key,
 * whether the cache answered and the per-candidate microseconds per step (truncated to
 * report_cap - 1 bytes). */
#define CF_TUNE_FORCE 1
int cf_autotune(cf_sim* sim, const cf_params* params, const cf_edge_opts* edges,
                int trial_steps, const char* cache_path, int flags, cf_tuning* chosen,
                double* us_per_step, char* report, int report_cap);
/* The cache key cf_autotune uses for this workload. */
int cf_tuning_key(cf_sim* sim, const cf_params* params, const cf_edge_opts* edges, char* key,
                  int key_cap);
/* Host-only access to a tuning cache file.  Lookup sets *found (0/1); a missing file is
 * an empty cache; a file that is not a tuning cache is CF_ERR_ARG.  Store adds or replaces
 * one entry (written through a temporary file and a rename). */
int cf_tuning_cache_lookup(const char* cache_path, const char* key, cf_tuning* tuning,
                           double* us_per_step, int* found);
int cf_tuning_cache_store(const char* cache_path, const char* key, const cf_tuning* tuning,
                          double us_per_step);

/* ---- preset files (host only) -------------------------------------------------------- */

/* The reference UI's preset JSON (CellFlowWidget::loadPreset / savePreset,
 * CellFlowWidget.cpp:1070-1269): the fields that feed the step path.  `present` has one
 * CF_PRESET_* bit per field the file holds; fields without their bit keep the defaults of
 * SimulationParams.h:15-35 (particle_count 0).  Arrays keep at most CF_MAX_TYPES colours,
 * CF_MAX_TYPES radius modifiers and CF_MAX_TYPES^2 raw table values, with their counts. */
#define CF_PRESET_PARTICLE_COUNT (1u << 0)
#define CF_PRESET_NUM_TYPES (1u << 1)
#define CF_PRESET_RADIUS (1u << 2)
#define CF_PRESET_DELTA_T (1u << 3)
#define CF_PRESET_FRICTION (1u << 4)
#define CF_PRESET_REPULSION (1u << 5)
#define CF_PRESET_ATTRACTION (1u << 6)
#define CF_PRESET_K (1u << 7)
#define CF_PRESET_BALANCE (1u << 8)
#define CF_PRESET_FORCE_MULTIPLIER (1u << 9)
#define CF_PRESET_FORCE_RANGE (1u << 10)
#define CF_PRESET_FORCE_BIAS (1u << 11)
#define CF_PRESET_RATIO (1u << 12)
#define CF_PRESET_LFO_A (1u << 13)
#define CF_PRESET_LFO_S (1u << 14)
#define CF_PRESET_FORCE_OFFSET (1u << 15)
#define CF_PRESET_CANVAS_WIDTH (1u << 16)
#define CF_PRESET_CANVAS_HEIGHT (1u << 17)
#define CF_PRESET_CANVAS_DEPTH (1u << 18)
#define CF_PRESET_SPAWN_REGION (1u << 19)
#define CF_PRESET_COLORS (1u << 20)
#define CF_PRESET_RADIO_BY_TYPE (1u << 21)
#define CF_PRESET_RAW_FORCE_TABLE (1u << 22)
#define CF_PRESET_ALL ((1u << 23) - 1u)

typedef struct cf_preset {
  uint32_t present;
  int particle_count;           /* "PARTICLE_COUNT" */
  int num_particle_types;       /* "numParticleTypes" */
  float radius, delta_t, friction, repulsion, attraction, k, balance, force_multiplier;
  float force_range, force_bias, force_offset;       /* updateForceTable arguments */
  float ratio, lfo_a, lfo_s;    /* ratioWithLFO = ratio + lfoA*sin(2*pi*lfoS*t), :414-421 */
  float canvas_width, canvas_height, canvas_depth, spawn_region_size;
  int num_colors;               /* "particleColors": [{"r","g","b"}, ...] */
  cf_color colors[CF_MAX_TYPES];
  int num_radio;                /* "radioByType" */
  float radio_by_type[CF_MAX_TYPES];
  int num_raw;                  /* "rawForceTable", row-major T*T */
  float raw_force_table[CF_MAX_TYPES * CF_MAX_TYPES];
} cf_preset;

/* Defaults (present = 0). */
void cf_preset_init(cf_preset* preset);
/* Parse a preset file.  Stricter than QJsonValue::toDouble/toInt, which turn a wrong type
 * into 0: a field of the wrong type (or a non-integral count) is CF_ERR_ARG. */
int cf_preset_read(const char* path, cf_preset* preset);
/* Write the present fields in the layout Qt's QJsonDocument::toJson gives the reference's
 * files (sorted keys, 4-space indent, shortest round-trip numbers).  With a non-null
 * `template_path`, every other member of that file's object (render and camera settings)
 * is kept, so load + save of an unchanged preset reproduces the file byte for byte. */
int cf_preset_write(const char* path, const cf_preset* preset, const char* template_path);
/* The step parameters of a preset at LFO phase 0 (ratio_with_lfo = ratio, 400 expected
 * neighbours, SimulationParams.h:29). */
int cf_preset_params(const cf_preset* preset, cf_params* params);

/* ---- host-only helpers (no GPU needed) ---------------------------------------------- */

/* The grid rule: cells per axis for a parameter set, type radii and particle count. */
int cf_grid_dims(const cf_params* params, const float* radio_by_type, int particle_count,
                 int cell_divisor, int dims[3]);
/* Smallest float t with sqrtf(t) >= eff_radius: the engine tests s + 1e-4f < t, which is
 * the reference's sqrtf(s + 0.0001f) < effectiveRadius (.cu:105, :112) exactly. */
float cf_cutoff_threshold(float eff_radius);
/* The updateForceTable transform (.cu:521-531) on host arrays. */
void cf_force_table_transform(const float* raw, int count, float range, float bias,
                              float offset, float* out);
const char* cf_last_error_string(void);
int cf_version(void);
int cf_device_count(int* count);

/* ---- batched ensemble (no reference counterpart) ------------------------------------ */

/* nsims independent simulations stepped together by the same kernels as cf_simulate.
 * Each has its own particle count, type count, force table, radii and parameters. */
int cf_ensemble_create(int nsims, const int* particle_counts, const int* num_types,
                       cf_ensemble** out);
int cf_ensemble_destroy(cf_ensemble* ens);
int cf_ensemble_set_sim(cf_ensemble* ens, int sim, const cf_params* params, const float* table,
                        const float* radio_by_type, const cf_particle* particles, int count);
int cf_ensemble_simulate(cf_ensemble* ens, int nsteps);
int cf_ensemble_get_particles(cf_ensemble* ens, int sim, cf_particle* particles, int count);
int cf_ensemble_get_neighbor_counts(cf_ensemble* ens, int sim, int32_t* counts, int count);
int cf_ensemble_set_tuning(cf_ensemble* ens, const cf_tuning* tuning);
int cf_ensemble_get_tuning(cf_ensemble* ens, cf_tuning* tuning);
/* cf_autotune for the whole ensemble (every member with its own parameters). */
int cf_ensemble_autotune(cf_ensemble* ens, int trial_steps, const char* cache_path, int flags,
                         cf_tuning* chosen, double* us_per_step, char* report, int report_cap);

#ifdef __cplusplus
}
#endif

#endif /* CELLFLOW_H */
