// cellflow.hpp — header-only C++ drop-in for the reference's `ParticleSimulation` class
// (cuda-native/include/ParticleSimulation.cuh:10-103) over the C ABI in cellflow.h.
//
// A C++ caller written against the reference (CellFlowWidget.cpp) keeps its calls:
//   sim.simulate(params); sim.getParticleData(particles); sim.updateForceTable(r, b, o);
//   sim.getRawForceTableValues()[k] = v; sim.setRadioByTypeValue(i, v); ...
// Differences a caller sees:
//   * errors throw cellflow::Error (status code + message) instead of exit(1)
//     (ParticleSimulation.cu:10-18);
//   * generateProximityGraph writes the reference's line vertices (2 per edge, each
//     x y z r g b) into a host vector, or (the overload taking a float*) straight into
//     device memory the caller owns, such as the pointer of a GL VBO the caller registered
//     and mapped with hipGraphicsGLRegisterBuffer / hipGraphicsMapResources; the reference
//     maps the VBO inside the call (.cu:633-643);
//   * random initialisation is seeded: particles by cf_init_particles with a per-object
//     counter instead of curand seeded from time(nullptr) (.cu:480); the raw force table and
//     the radius modifiers use std::rand() exactly as the reference does (.cu:516, :535).
// Link with libcellflow.so (cellflow_amd/libcellflow.so).
#ifndef CELLFLOW_HPP
#define CELLFLOW_HPP

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "cellflow.h"

namespace cellflow {

using Particle = cf_particle;        // struct Particle, SimulationParams.h:6-12
using ParticleColor = cf_color;      // struct ParticleColor, SimulationParams.h:59-61
constexpr int MAX_PARTICLE_TYPES = CF_MAX_TYPES;

static_assert(sizeof(Particle) == 44, "Particle must keep the reference's 44-byte layout");

// The step fields of SimulationParams (SimulationParams.h:14-29) with the same defaults.
struct SimulationParams {
  float radius = 42.07f;
  float delta_t = 0.18f;
  float friction = 0.51f;
  float repulsion = 64.83f;
  float attraction = 3.06f;
  float k = 29.45f;
  float balance = 0.79f;
  float canvasWidth = 8000.0f;
  float canvasHeight = 8000.0f;
  float canvasDepth = 8000.0f;
  float spawnRegionSize = 2000.0f;
  int numParticleTypes = 6;
  float ratioWithLFO = 0.0f;
  float forceMultiplier = 2.33f;
  int maxExpectedNeighbors = 400;

  cf_params to_c() const {
    cf_params p;
    p.radius = radius;
    p.delta_t = delta_t;
    p.friction = friction;
    p.repulsion = repulsion;
    p.attraction = attraction;
    p.k = k;
    p.balance = balance;
    p.canvas_width = canvasWidth;
    p.canvas_height = canvasHeight;
    p.canvas_depth = canvasDepth;
    p.num_particle_types = numParticleTypes;
    p.ratio_with_lfo = ratioWithLFO;
    p.force_multiplier = forceMultiplier;
    p.max_expected_neighbors = maxExpectedNeighbors;
    return p;
  }
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != CF_OK) {
    const char* m = cf_last_error_string();
    throw Error(rc, std::string("cellflow error ") + std::to_string(rc) + ": " + (m ? m : ""));
  }
}

class ParticleSimulation {
 public:
  // ParticleSimulation(int), .cu:427-435: 6 types, canvas 8000^3, random particles, force
  // table and radius modifiers.
  explicit ParticleSimulation(int particleCount)
      : ParticleSimulation(particleCount, 8000.0f, 8000.0f) {}
  // ParticleSimulation(int, float, float), .cu:437-445 (depth = height).
  ParticleSimulation(int particleCount, float canvasWidth, float canvasHeight)
      : canvas_{canvasWidth, canvasHeight, canvasHeight} {
    check(cf_create(particleCount, types_, &h_));
    initializeParticles();
    initializeForceTable();
    initializeRadioByType();
  }
  ~ParticleSimulation() {
    if (h_) cf_destroy(h_);
  }
  ParticleSimulation(const ParticleSimulation&) = delete;
  ParticleSimulation& operator=(const ParticleSimulation&) = delete;
  ParticleSimulation(ParticleSimulation&& o) noexcept { *this = std::move(o); }
  ParticleSimulation& operator=(ParticleSimulation&& o) noexcept {
    if (this != &o) {
      if (h_) cf_destroy(h_);
      h_ = o.h_;
      o.h_ = nullptr;
      types_ = o.types_;
      seed_ = o.seed_;
      for (int a = 0; a < 3; ++a) canvas_[a] = o.canvas_[a];
      spawn_ = o.spawn_;
      for (int i = 0; i < MAX_PARTICLE_TYPES * MAX_PARTICLE_TYPES; ++i) raw_[i] = o.raw_[i];
    }
    return *this;
  }

  // initializeParticles(), .cu:484-492: spawn cube min(2000, canvas), centred.
  void initializeParticles() {
    check(cf_init_particles(h_, seed_++, canvas_, spawn_));
  }
  // initializeParticles(w, h), .cu:494-505.
  void initializeParticles(float canvasWidth, float canvasHeight) {
    updateCanvasDimensions(canvasWidth, canvasHeight);
    initializeParticles();
  }
  // updateCanvasDimensions, .cu:507-511 (depth follows height).
  void updateCanvasDimensions(float canvasWidth, float canvasHeight) {
    canvas_[0] = canvasWidth;
    canvas_[1] = canvasHeight;
    canvas_[2] = canvasHeight;
  }

  // initializeForceTable, .cu:513-519: raw ~ U[-1, 1] from rand(), then (0.28, -0.2, 1.0).
  void initializeForceTable() {
    for (int i = 0; i < types_ * types_; ++i)
      raw_[i] = (float)std::rand() / RAND_MAX * 2.0f - 1.0f;
    updateForceTable(0.28f, -0.20f, 1.0f);
  }
  // updateForceTable, .cu:521-531, from the raw values (possibly edited in place through
  // getRawForceTableValues, as CellFlowWidget does).
  void updateForceTable(float forceRange, float forceBias, float forceOffset) {
    check(cf_set_raw_force_table(h_, raw_, types_ * types_));
    check(cf_update_force_table(h_, forceRange, forceBias, forceOffset));
  }
  void regenerateForceTable() { initializeForceTable(); }  // .cu:581-583
  float* getRawForceTableValues() { return raw_; }          // .cuh:47

  // initializeRadioByType, .cu:533-539.
  void initializeRadioByType() {
    std::vector<float> v(types_);
    for (auto& x : v) x = (float)std::rand() / RAND_MAX * 2.0f - 1.0f;
    check(cf_set_radio_by_type(h_, v.data(), types_));
  }
  std::vector<float> getRadioByType() const {  // .cu:607-613
    std::vector<float> v(types_);
    check(cf_get_radio_by_type(h_, v.data(), types_));
    return v;
  }
  void setRadioByTypeValue(int index, float value) {  // .cu:615-622 (out of range: ignored)
    check(cf_set_radio_by_type_value(h_, index, value));
  }
  void rotateRadioByType() { check(cf_rotate_radio_by_type(h_)); }  // .cu:594-605

  // simulate, .cu:541-556: one step.
  void simulate(const SimulationParams& params) {
    const cf_params p = params.to_c();
    check(cf_simulate(h_, &p, 1, nullptr));
  }
  // Several steps in one call (graph-replayed), optionally with the proximity graph of
  // every post-step state (the last one is kept for getEdges / the vertex expansion).
  void simulate(const SimulationParams& params, int steps, const cf_edge_opts* edges = nullptr) {
    const cf_params p = params.to_c();
    check(cf_simulate(h_, &p, steps, edges));
  }

  // getParticleData, .cu:558-562.
  void getParticleData(std::vector<Particle>& particles) {
    particles.resize(getParticleCount());
    check(cf_get_particles(h_, particles.data(), (int)particles.size()));
  }
  // No reference counterpart: load a given state (neighbour counts restart at zero).
  void setParticleData(const std::vector<Particle>& particles) {
    check(cf_set_particles(h_, particles.data(), (int)particles.size()));
  }

  // setParticleCount, .cu:564-572: reallocate and re-initialise the particles.
  void setParticleCount(int count) {
    if (count != getParticleCount()) {
      check(cf_set_particle_count(h_, count));
      initializeParticles();
    }
  }
  int getParticleCount() const {
    int n = 0;
    check(cf_get_particle_count(h_, &n));
    return n;
  }
  // setNumParticleTypes, .cu:574-579: new table, radii and particles.
  void setNumParticleTypes(int types) {
    check(cf_set_num_particle_types(h_, types));
    types_ = types;
    initializeForceTable();
    initializeRadioByType();
    initializeParticles();
  }
  int getNumParticleTypes() const { return types_; }

  // moveUniverse, .cu:585-592: the reference shifts with a default SimulationParams'
  // canvas (8000^3) whatever the current one is, and dz = 0.
  void moveUniverse(float dx, float dy) {
    const SimulationParams d;
    const float box[3] = {d.canvasWidth, d.canvasHeight, d.canvasDepth};
    check(cf_move_universe(h_, dx, dy, 0.0f, box));
  }

  // generateProximityGraph, .cu:624-687, into a host vertex buffer: 2 vertices per edge of
  // (x, y, z, r, g, b), colour = particleColors[ptype_i % numTypes].
  void generateProximityGraph(std::vector<float>& vertices, int& outVertexCount,
                              float proximityDistance, int maxConnectionsPerParticle,
                              const std::vector<ParticleColor>& particleColors) {
    check(cf_build_edges(h_, proximityDistance, maxConnectionsPerParticle));
    long long vc = 0;
    const int rc = cf_get_line_vertices(h_, particleColors.data(), (int)particleColors.size(),
                                        nullptr, 0, &vc);
    if (rc != CF_OK && rc != CF_ERR_CAPACITY) check(rc);
    vertices.resize((size_t)vc * 6);
    check(cf_get_line_vertices(h_, particleColors.data(), (int)particleColors.size(),
                               vertices.data(), (long long)vertices.size(), &vc));
    outVertexCount = (int)vc;
  }
  // generateTriangleMesh, .cu:689-754, into a host vertex buffer: 3 vertices per triangle
  // of (x, y, z, nx, ny, nz, r, g, b), colour = particleColors[ptype_owner % numTypes].
  void generateTriangleMesh(std::vector<float>& vertices, int& outVertexCount,
                            float proximityDistance, int maxConnectionsPerParticle,
                            const std::vector<ParticleColor>& particleColors) {
    check(cf_build_triangles(h_, proximityDistance, maxConnectionsPerParticle));
    long long vc = 0;
    const int rc = cf_get_triangle_vertices(h_, particleColors.data(),
                                            (int)particleColors.size(), nullptr, 0, &vc);
    if (rc != CF_OK && rc != CF_ERR_CAPACITY) check(rc);
    vertices.resize((size_t)vc * 9);
    check(cf_get_triangle_vertices(h_, particleColors.data(), (int)particleColors.size(),
                                   vertices.data(), (long long)vertices.size(), &vc));
    outVertexCount = (int)vc;
  }
  // generateProximityGraph writing the line vertices into caller-owned device memory
  // (`deviceVertices`, capacityFloats floats), e.g. a mapped GL VBO: no host round trip.
  void generateProximityGraph(float* deviceVertices, long long capacityFloats,
                              int& outVertexCount, float proximityDistance,
                              int maxConnectionsPerParticle,
                              const std::vector<ParticleColor>& particleColors) {
    check(cf_build_edges(h_, proximityDistance, maxConnectionsPerParticle));
    long long vc = 0;
    check(cf_write_line_vertices_device(h_, particleColors.data(), (int)particleColors.size(),
                                        deviceVertices, capacityFloats, &vc));
    outVertexCount = (int)vc;
  }
  // generateTriangleMesh into caller-owned device memory, as above.
  void generateTriangleMesh(float* deviceVertices, long long capacityFloats,
                            int& outVertexCount, float proximityDistance,
                            int maxConnectionsPerParticle,
                            const std::vector<ParticleColor>& particleColors) {
    check(cf_build_triangles(h_, proximityDistance, maxConnectionsPerParticle));
    long long vc = 0;
    check(cf_write_triangle_vertices_device(h_, particleColors.data(),
                                            (int)particleColors.size(), deviceVertices,
                                            capacityFloats, &vc));
    outVertexCount = (int)vc;
  }
  // The last proximity graph as canonical (i, j) pairs.
  std::vector<int32_t> getEdges() {
    long long n = 0;
    const int rc = cf_get_edges(h_, nullptr, 0, &n);
    if (rc != CF_OK && rc != CF_ERR_CAPACITY) check(rc);
    std::vector<int32_t> e((size_t)n * 2);
    check(cf_get_edges(h_, e.data(), n, &n));
    return e;
  }

  std::vector<int32_t> getNeighborCounts() {
    std::vector<int32_t> c(getParticleCount());
    check(cf_get_neighbor_counts(h_, c.data(), (int)c.size()));
    return c;
  }

  // CellFlowWidget::loadPreset (CellFlowWidget.cpp:1070-1203) on this simulation, in the
  // widget's order: count (re-allocates), types (new random table and radii, as
  // setNumParticleTypes does), the step fields into `params`, radius modifiers for the
  // first min(len, T), raw table entries for the first min(len, T*T), then
  // updateForceTable(forceRange, forceBias, forceOffset).  Fields the file lacks keep their
  // values in `params`.  After a count or type change the particles are re-initialised in
  // the preset's canvas and spawn region.  Returns the preset (colours, LFO settings, ...).
  cf_preset loadPreset(const char* path, SimulationParams& params) {
    cf_preset pr;
    check(cf_preset_read(path, &pr));
    bool reinit = false;
    if ((pr.present & CF_PRESET_PARTICLE_COUNT) && pr.particle_count != getParticleCount()) {
      check(cf_set_particle_count(h_, pr.particle_count));
      reinit = true;
    }
    if (pr.present & CF_PRESET_NUM_TYPES) {
      check(cf_set_num_particle_types(h_, pr.num_particle_types));
      types_ = pr.num_particle_types;
      initializeForceTable();
      initializeRadioByType();
      reinit = true;
    }
    params.numParticleTypes = types_;
    struct F {
      uint32_t bit;
      float cf_preset::*src;
      float SimulationParams::*dst;
    };
    const F fields[] = {
        {CF_PRESET_RADIUS, &cf_preset::radius, &SimulationParams::radius},
        {CF_PRESET_DELTA_T, &cf_preset::delta_t, &SimulationParams::delta_t},
        {CF_PRESET_FRICTION, &cf_preset::friction, &SimulationParams::friction},
        {CF_PRESET_REPULSION, &cf_preset::repulsion, &SimulationParams::repulsion},
        {CF_PRESET_ATTRACTION, &cf_preset::attraction, &SimulationParams::attraction},
        {CF_PRESET_K, &cf_preset::k, &SimulationParams::k},
        {CF_PRESET_BALANCE, &cf_preset::balance, &SimulationParams::balance},
        {CF_PRESET_FORCE_MULTIPLIER, &cf_preset::force_multiplier,
         &SimulationParams::forceMultiplier},
        {CF_PRESET_CANVAS_WIDTH, &cf_preset::canvas_width, &SimulationParams::canvasWidth},
        {CF_PRESET_CANVAS_HEIGHT, &cf_preset::canvas_height, &SimulationParams::canvasHeight},
        {CF_PRESET_CANVAS_DEPTH, &cf_preset::canvas_depth, &SimulationParams::canvasDepth},
        {CF_PRESET_SPAWN_REGION, &cf_preset::spawn_region_size,
         &SimulationParams::spawnRegionSize},
        {CF_PRESET_RATIO, &cf_preset::ratio, &SimulationParams::ratioWithLFO},  // LFO phase 0
    };
    for (const F& f : fields)
      if (pr.present & f.bit) params.*f.dst = pr.*f.src;
    for (int i = 0; i < pr.num_radio && i < types_; ++i)
      setRadioByTypeValue(i, pr.radio_by_type[i]);
    for (int i = 0; i < pr.num_raw && i < types_ * types_; ++i) raw_[i] = pr.raw_force_table[i];
    updateForceTable(pr.force_range, pr.force_bias, pr.force_offset);
    if (reinit) {
      canvas_[0] = params.canvasWidth;
      canvas_[1] = params.canvasHeight;
      canvas_[2] = params.canvasDepth;
      spawn_ = params.spawnRegionSize;
      initializeParticles();
    }
    return pr;
  }

  // CellFlowWidget::savePreset (:1205-1269): the simulation's count, types, radius modifiers
  // and raw table and the step fields of `params`, with the UI fields of `ui` (force range,
  // bias and offset, ratio, LFO, colours).  A non-null `template_path` keeps every other
  // member of that file (render and camera settings).
  void savePreset(const char* path, const SimulationParams& params, cf_preset ui,
                  const char* template_path = nullptr) {
    ui.present |= CF_PRESET_PARTICLE_COUNT | CF_PRESET_NUM_TYPES | CF_PRESET_RADIUS |
                  CF_PRESET_DELTA_T | CF_PRESET_FRICTION | CF_PRESET_REPULSION |
                  CF_PRESET_ATTRACTION | CF_PRESET_K | CF_PRESET_BALANCE |
                  CF_PRESET_FORCE_MULTIPLIER | CF_PRESET_CANVAS_WIDTH | CF_PRESET_CANVAS_HEIGHT |
                  CF_PRESET_CANVAS_DEPTH | CF_PRESET_RADIO_BY_TYPE | CF_PRESET_RAW_FORCE_TABLE;
    ui.particle_count = getParticleCount();
    ui.num_particle_types = types_;
    ui.radius = params.radius;
    ui.delta_t = params.delta_t;
    ui.friction = params.friction;
    ui.repulsion = params.repulsion;
    ui.attraction = params.attraction;
    ui.k = params.k;
    ui.balance = params.balance;
    ui.force_multiplier = params.forceMultiplier;
    ui.canvas_width = params.canvasWidth;
    ui.canvas_height = params.canvasHeight;
    ui.canvas_depth = params.canvasDepth;
    const std::vector<float> rbt = getRadioByType();
    ui.num_radio = types_;
    for (int i = 0; i < types_; ++i) ui.radio_by_type[i] = rbt[i];
    ui.num_raw = types_ * types_;
    for (int i = 0; i < types_ * types_; ++i) ui.raw_force_table[i] = raw_[i];
    check(cf_preset_write(path, &ui, template_path));
  }

  // Autotuner (no reference counterpart): times the kernel variants, lanes per particle and
  // step sorts on a copy of the current state for `params` (and the edge options the caller
  // passes to simulate) and applies the fastest; the state is left as it was.  With
  // `cachePath`, a JSON cache answers later calls for the same workload bucket.  Untuned, the
  // engine runs force variant 0 at 16 lanes: at 100k particles about 1.5x slower per step.
  cf_tuning autotune(const SimulationParams& params, const cf_edge_opts* edges = nullptr,
                     const char* cachePath = nullptr, int trialSteps = 64) {
    const cf_params p = params.to_c();
    cf_tuning t{};
    check(cf_autotune(h_, &p, edges, trialSteps, cachePath, 0, &t, nullptr, nullptr, 0));
    return t;
  }

  cf_sim* handle() const { return h_; }

 private:
  cf_sim* h_ = nullptr;
  int types_ = 6;
  uint64_t seed_ = 1;
  float canvas_[3] = {8000.0f, 8000.0f, 8000.0f};
  float spawn_ = 2000.0f;
  float raw_[MAX_PARTICLE_TYPES * MAX_PARTICLE_TYPES] = {};
};

}  // namespace cellflow

#endif  // CELLFLOW_HPP
