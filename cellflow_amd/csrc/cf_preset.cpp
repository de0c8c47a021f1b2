// Preset files of the reference UI (CellFlowWidget::loadPreset / savePreset,
// CellFlowWidget.cpp:1070-1269): the step-path fields, read into and written from
// cf_preset.  Host-only; no device calls.

#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include <unistd.h>

#include "../../include/cellflow.h"
#include "cf_host.h"
#include "cf_json.h"

namespace {

using cf::json::JVal;

// Scalar members: key, presence bit, float member (or int member).
struct FloatField {
  const char* key;
  uint32_t bit;
  float cf_preset::*member;
};

const FloatField kFloats[] = {
    {"radius", CF_PRESET_RADIUS, &cf_preset::radius},
    {"delta_t", CF_PRESET_DELTA_T, &cf_preset::delta_t},
    {"friction", CF_PRESET_FRICTION, &cf_preset::friction},
    {"repulsion", CF_PRESET_REPULSION, &cf_preset::repulsion},
    {"attraction", CF_PRESET_ATTRACTION, &cf_preset::attraction},
    {"k", CF_PRESET_K, &cf_preset::k},
    {"balance", CF_PRESET_BALANCE, &cf_preset::balance},
    {"forceMultiplier", CF_PRESET_FORCE_MULTIPLIER, &cf_preset::force_multiplier},
    {"forceRange", CF_PRESET_FORCE_RANGE, &cf_preset::force_range},
    {"forceBias", CF_PRESET_FORCE_BIAS, &cf_preset::force_bias},
    {"ratio", CF_PRESET_RATIO, &cf_preset::ratio},
    {"lfoA", CF_PRESET_LFO_A, &cf_preset::lfo_a},
    {"lfoS", CF_PRESET_LFO_S, &cf_preset::lfo_s},
    {"forceOffset", CF_PRESET_FORCE_OFFSET, &cf_preset::force_offset},
    {"canvasWidth", CF_PRESET_CANVAS_WIDTH, &cf_preset::canvas_width},
    {"canvasHeight", CF_PRESET_CANVAS_HEIGHT, &cf_preset::canvas_height},
    {"canvasDepth", CF_PRESET_CANVAS_DEPTH, &cf_preset::canvas_depth},
    {"spawnRegionSize", CF_PRESET_SPAWN_REGION, &cf_preset::spawn_region_size},
};

struct IntField {
  const char* key;
  uint32_t bit;
  int cf_preset::*member;
};

const IntField kInts[] = {
    {"PARTICLE_COUNT", CF_PRESET_PARTICLE_COUNT, &cf_preset::particle_count},
    {"numParticleTypes", CF_PRESET_NUM_TYPES, &cf_preset::num_particle_types},
};

int read_text(const char* path, std::string& text) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return cf::fail(CF_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
  char buf[65536];
  size_t n = 0;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) {
    text.append(buf, n);
    if (text.size() > (16u << 20)) {
      std::fclose(f);
      return cf::fail(CF_ERR_ARG, "%s is larger than 16 MiB", path);
    }
  }
  std::fclose(f);
  return CF_OK;
}

int parse_object(const char* path, JVal& root) {
  std::string text;
  const int rc = read_text(path, text);
  if (rc != CF_OK) return rc;
  cf::json::Reader r(text.data(), text.size());
  if (!r.parse(root) || root.kind != JVal::Obj)
    return cf::fail(CF_ERR_ARG, "%s is not a JSON object", path);
  return CF_OK;
}

int number_array(const JVal* v, const char* key, float* out, int cap, int* count) {
  if (v->kind != JVal::Arr) return cf::fail(CF_ERR_ARG, "\"%s\" is not an array", key);
  int n = 0;
  for (const JVal& e : v->arr) {
    if (e.kind != JVal::Num) return cf::fail(CF_ERR_ARG, "\"%s\" holds a non-number", key);
    if (n < cap) out[n++] = (float)e.num;
  }
  *count = n;
  return CF_OK;
}

}  // namespace

extern "C" {

void cf_preset_init(cf_preset* p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  // SimulationParams.h:15-35 defaults
  p->radius = 42.07f;
  p->delta_t = 0.18f;
  p->friction = 0.51f;
  p->repulsion = 64.83f;
  p->attraction = 3.06f;
  p->k = 29.45f;
  p->balance = 0.79f;
  p->force_multiplier = 2.33f;
  p->canvas_width = p->canvas_height = p->canvas_depth = 8000.0f;
  p->spawn_region_size = 2000.0f;
  p->num_particle_types = 6;
  p->force_range = 0.28f;
  p->force_bias = -0.20f;
  p->ratio = 0.0f;
  p->lfo_a = 0.0f;
  p->lfo_s = 0.1f;
  p->force_offset = 1.0f;
}

int cf_preset_read(const char* path, cf_preset* p) {
  if (!path || !p) return cf::fail(CF_ERR_ARG, "null argument");
  JVal root;
  int rc = parse_object(path, root);
  if (rc != CF_OK) return rc;
  cf_preset_init(p);
  for (const IntField& f : kInts) {
    const JVal* v = root.get(f.key);
    if (!v) continue;
    if (v->kind != JVal::Num || v->num != std::floor(v->num) || std::fabs(v->num) > 2147483647.0)
      return cf::fail(CF_ERR_ARG, "\"%s\" is not an integer", f.key);
    p->*f.member = (int)v->num;
    p->present |= f.bit;
  }
  for (const FloatField& f : kFloats) {
    const JVal* v = root.get(f.key);
    if (!v) continue;
    if (v->kind != JVal::Num) return cf::fail(CF_ERR_ARG, "\"%s\" is not a number", f.key);
    p->*f.member = (float)v->num;
    p->present |= f.bit;
  }
  if (const JVal* v = root.get("particleColors")) {
    if (v->kind != JVal::Arr) return cf::fail(CF_ERR_ARG, "\"particleColors\" is not an array");
    int n = 0;
    for (const JVal& c : v->arr) {
      if (c.kind != JVal::Obj) return cf::fail(CF_ERR_ARG, "a particle colour is not an object");
      float rgb[3] = {0.0f, 0.0f, 0.0f};
      const char* names[3] = {"r", "g", "b"};
      for (int a = 0; a < 3; ++a) {
        const JVal* x = c.get(names[a]);
        if (x && x->kind != JVal::Num)
          return cf::fail(CF_ERR_ARG, "colour component \"%s\" is not a number", names[a]);
        if (x) rgb[a] = (float)x->num;
      }
      if (n < CF_MAX_TYPES) p->colors[n++] = cf_color{rgb[0], rgb[1], rgb[2]};
    }
    p->num_colors = n;
    p->present |= CF_PRESET_COLORS;
  }
  if (const JVal* v = root.get("radioByType")) {
    rc = number_array(v, "radioByType", p->radio_by_type, CF_MAX_TYPES, &p->num_radio);
    if (rc != CF_OK) return rc;
    p->present |= CF_PRESET_RADIO_BY_TYPE;
  }
  if (const JVal* v = root.get("rawForceTable")) {
    rc = number_array(v, "rawForceTable", p->raw_force_table, CF_MAX_TYPES * CF_MAX_TYPES,
                      &p->num_raw);
    if (rc != CF_OK) return rc;
    p->present |= CF_PRESET_RAW_FORCE_TABLE;
  }
  return CF_OK;
}

int cf_preset_write(const char* path, const cf_preset* p, const char* template_path) {
  if (!path || !p) return cf::fail(CF_ERR_ARG, "null argument");
  if (p->num_colors < 0 || p->num_colors > CF_MAX_TYPES || p->num_radio < 0 ||
      p->num_radio > CF_MAX_TYPES || p->num_raw < 0 || p->num_raw > CF_MAX_TYPES * CF_MAX_TYPES)
    return cf::fail(CF_ERR_ARG, "array counts out of range");
  JVal root;
  root.kind = JVal::Obj;
  if (template_path) {
    const int rc = parse_object(template_path, root);
    if (rc != CF_OK) return rc;
  }
  using cf::json::num;
  using cf::json::set;
  for (const IntField& f : kInts)
    if (p->present & f.bit) set(root, f.key, num((double)(p->*f.member)));
  for (const FloatField& f : kFloats) {
    if (!(p->present & f.bit)) continue;
    const float x = p->*f.member;
    if (!std::isfinite(x)) return cf::fail(CF_ERR_ARG, "\"%s\" is not finite", f.key);
    set(root, f.key, num((double)x));
  }
  auto arr = [](const float* x, int n, JVal& out) -> bool {
    out = JVal();
    out.kind = JVal::Arr;
    for (int i = 0; i < n; ++i) {
      if (!std::isfinite(x[i])) return false;
      out.arr.push_back(num((double)x[i]));
    }
    return true;
  };
  if (p->present & CF_PRESET_COLORS) {
    JVal a;
    a.kind = JVal::Arr;
    for (int i = 0; i < p->num_colors; ++i) {
      JVal c;
      c.kind = JVal::Obj;
      const float rgb[3] = {p->colors[i].r, p->colors[i].g, p->colors[i].b};
      const char* names[3] = {"r", "g", "b"};
      for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(rgb[k])) return cf::fail(CF_ERR_ARG, "colour is not finite");
        set(c, names[k], num((double)rgb[k]));
      }
      a.arr.push_back(std::move(c));
    }
    set(root, "particleColors", std::move(a));
  }
  if (p->present & CF_PRESET_RADIO_BY_TYPE) {
    JVal a;
    if (!arr(p->radio_by_type, p->num_radio, a))
      return cf::fail(CF_ERR_ARG, "radioByType is not finite");
    set(root, "radioByType", std::move(a));
  }
  if (p->present & CF_PRESET_RAW_FORCE_TABLE) {
    JVal a;
    if (!arr(p->raw_force_table, p->num_raw, a))
      return cf::fail(CF_ERR_ARG, "rawForceTable is not finite");
    set(root, "rawForceTable", std::move(a));
  }
  const std::string text = cf::json::dump(root);
  const std::string tmp = std::string(path) + ".tmp." + std::to_string((long)getpid());
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return cf::fail(CF_ERR_ARG, "cannot write %s: %s", path, std::strerror(errno));
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size() && std::fflush(f) == 0;
  std::fclose(f);
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return cf::fail(CF_ERR_ARG, "cannot write %s: %s", path, std::strerror(errno));
  }
  return CF_OK;
}

int cf_preset_params(const cf_preset* p, cf_params* out) {
  if (!p || !out) return cf::fail(CF_ERR_ARG, "null argument");
  out->radius = p->radius;
  out->delta_t = p->delta_t;
  out->friction = p->friction;
  out->repulsion = p->repulsion;
  out->attraction = p->attraction;
  out->k = p->k;
  out->balance = p->balance;
  out->canvas_width = p->canvas_width;
  out->canvas_height = p->canvas_height;
  out->canvas_depth = p->canvas_depth;
  out->num_particle_types = p->num_particle_types;
  out->ratio_with_lfo = p->ratio;  // + lfoA * sin(0)
  out->force_multiplier = p->force_multiplier;
  out->max_expected_neighbors = 400;
  return CF_OK;
}

}  // extern "C"
