// Drives include/cellflow.hpp the way CellFlowWidget drives the reference class
// (CellFlowWidget.cpp:410-427 step per frame, :551-575 proximity graph after the step,
// :906-917 raw-table edits + updateForceTable, :893 setRadioByTypeValue) and writes the
// inputs it ended up with and the outputs to <dir>, so a test can replay the same calls
// through another binding and compare bit for bit.  With a preset path as the second
// argument it also loads that preset, steps, and saves it to <dir>/saved.json.
//
// With "--tune" as the first argument it instead checks the class's autotune(): the state is
// left as it was, the chosen tuning is the one applied, and the tuned simulation steps.
//
// Exit status: 0 success; 1 a failed check; 10 + cellflow status on a cellflow::Error
// (12 = no device).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cellflow.hpp"

namespace {

template <class T>
void dump(const std::string& path, const T* data, size_t count) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(data, sizeof(T), count, f);
  std::fclose(f);
}

int tune_check() {
  cellflow::ParticleSimulation sim(4000, 0.5f, 0.5f);
  cellflow::SimulationParams p;
  p.radius = 0.05f;
  p.canvasWidth = p.canvasHeight = p.canvasDepth = 0.5f;
  p.ratioWithLFO = 0.3f;
  sim.simulate(p, 3);
  std::vector<cellflow::Particle> before, after;
  sim.getParticleData(before);
  const cf_tuning t = sim.autotune(p, nullptr, nullptr, 4);
  sim.getParticleData(after);
  if (before.size() != after.size() ||
      std::memcmp(before.data(), after.data(), before.size() * sizeof(before[0])) != 0) {
    std::printf("autotune changed the state\n");
    return 1;
  }
  cf_tuning applied{};
  if (cf_get_tuning(sim.handle(), &applied) != CF_OK ||
      applied.force_variant != t.force_variant ||
      applied.lanes_per_particle != t.lanes_per_particle ||
      applied.sort_variant != t.sort_variant || applied.cell_divisor != t.cell_divisor) {
    std::printf("the chosen tuning is not the applied one\n");
    return 1;
  }
  sim.simulate(p, 3);
  std::printf("tuned: variant %d, %d lanes, cell divisor %d, sort %d\n", t.force_variant,
              t.lanes_per_particle, t.cell_divisor, t.sort_variant);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  try {
    if (dir == "--tune") return tune_check();
    const int n = 6000, frames = 12;
    cellflow::ParticleSimulation sim(n, 0.6f, 0.6f);  // canvas 0.6^3; spawn cube = canvas
    cellflow::SimulationParams p;
    p.radius = 0.05f;
    p.delta_t = 0.01f;
    p.friction = 0.1f;
    p.k = 1.0f;
    p.canvasWidth = p.canvasHeight = p.canvasDepth = 0.6f;
    p.ratioWithLFO = 0.3f;
    // UI edits: one raw entry, then the transform; one radius modifier, then a rotation
    sim.getRawForceTableValues()[3] = 0.5f;
    sim.updateForceTable(0.6f, 0.1f, 1.5f);
    sim.setRadioByTypeValue(2, 0.25f);
    sim.setRadioByTypeValue(99, 1.0f);  // out of range: ignored, as in the reference
    sim.rotateRadioByType();

    std::vector<cellflow::Particle> init;
    sim.getParticleData(init);
    std::vector<float> table(36), raw(sim.getRawForceTableValues(),
                                      sim.getRawForceTableValues() + 36);
    cellflow::check(cf_get_force_table(sim.handle(), table.data(), 36));
    const std::vector<float> rbt = sim.getRadioByType();

    const std::vector<cellflow::ParticleColor> colors = {
        {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};
    std::vector<float> verts;
    int vc = 0;
    for (int f = 0; f < frames; ++f) {
      sim.simulate(p);
      sim.generateProximityGraph(verts, vc, 0.02f, 16, colors);
    }
    // generateTriangleMesh (.cu:689-754) of the last state; the graph stays readable
    std::vector<float> tverts;
    int tvc = 0;
    sim.generateTriangleMesh(tverts, tvc, 0.05f, 8, colors);
    std::vector<cellflow::Particle> fin;
    sim.getParticleData(fin);
    const std::vector<int32_t> counts = sim.getNeighborCounts();
    const std::vector<int32_t> edges = sim.getEdges();

    dump(dir + "/init.bin", init.data(), init.size());
    dump(dir + "/table.bin", table.data(), table.size());
    dump(dir + "/raw.bin", raw.data(), raw.size());
    dump(dir + "/rbt.bin", rbt.data(), rbt.size());
    dump(dir + "/final.bin", fin.data(), fin.size());
    dump(dir + "/counts.bin", counts.data(), counts.size());
    dump(dir + "/verts.bin", verts.data(), verts.size());
    dump(dir + "/edges.bin", edges.data(), edges.size());
    dump(dir + "/tverts.bin", tverts.data(), tverts.size());
    std::printf("ok n=%d frames=%d vertices=%d edges=%zu triangle vertices=%d\n", n, frames,
                vc, edges.size() / 2, tvc);

    // CellFlowWidget::loadPreset / savePreset: load a preset into a fresh simulation, run a
    // few frames, save it back with the original as the template.
    if (argc > 2) {
      cellflow::ParticleSimulation ps(1000);
      cellflow::SimulationParams pp;
      const cf_preset ui = ps.loadPreset(argv[2], pp);
      for (int f = 0; f < 3; ++f) ps.simulate(pp);
      ps.savePreset((dir + "/saved.json").c_str(), pp, ui, argv[2]);
      std::vector<float> t(ps.getNumParticleTypes() * ps.getNumParticleTypes());
      cellflow::check(cf_get_force_table(ps.handle(), t.data(), (int)t.size()));
      dump(dir + "/preset_table.bin", t.data(), t.size());
      std::printf("preset n=%d types=%d\n", ps.getParticleCount(), ps.getNumParticleTypes());
    }
    return 0;
  } catch (const cellflow::Error& e) {
    std::printf("%s\n", e.what());
    return 10 + e.code();
  }
}
