// Tuning cache: a small JSON file mapping a workload bucket key to the fastest tuning
// measured for it.  Host-only code; no device calls.
//
// File format (written sorted by key, so the file diffs cleanly):
//   {
//     "format": "cellflow-tuning",
//     "version": 1,
//     "entries": {
//       "gfx950/sims=0/n=16/rho=1/T=8/e=16": {"cell_divisor": 1, "force_variant": 0,
//           "lanes_per_particle": 16, "block_size": 256, "steps_per_graph": 32,
//           "sort_variant": 1, "us_per_step": 74.1},
//       ...
//     }
//   }
// Unknown top-level fields and unknown entry fields are ignored on read; an entry without
// "sort_variant" (written before it existed) reads as sort_variant 1.
#pragma once

#include <cerrno>
#include <cmath>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include <unistd.h>

#include "../../include/cellflow.h"
#include "cf_json.h"

namespace cf {
namespace tunecache {

using json::JVal;
using json::Reader;
using json::escape;

struct Entry {
  cf_tuning tun{};
  double us = 0.0;
};

static inline bool int_field(const JVal& e, const char* k, int* out) {
  const JVal* v = e.get(k);
  if (!v || v->kind != JVal::Num || v->num != std::floor(v->num) || std::fabs(v->num) > 1e9)
    return false;
  *out = (int)v->num;
  return true;
}

// The values cf_set_tuning accepts (variant 3 needs at least 8 lanes per particle for the
// 125-cell stencil: at most four stencil rows per lane).
static inline bool valid(const cf_tuning& t) {
  const bool lanes_ok = t.lanes_per_particle == 4 || t.lanes_per_particle == 8 ||
                        t.lanes_per_particle == 16 || t.lanes_per_particle == 32;
  const bool list_ok = (t.force_variant != 3 || t.lanes_per_particle >= 4 * t.cell_divisor) &&
                       (t.force_variant != 4 || t.lanes_per_particle == 8 ||
                        t.lanes_per_particle == 16);
  return (t.cell_divisor == 1 || t.cell_divisor == 2) &&
         (t.block_size == 64 || t.block_size == 128 || t.block_size == 256) &&
         (t.force_variant >= 0 && t.force_variant <= 4) && lanes_ok && list_ok &&
         t.steps_per_graph >= 1 && (t.sort_variant == 0 || t.sort_variant == 1);
}

enum Status { kOk = 0, kMissing = 1, kBad = 2 };

// Reads every entry of the file; kMissing when the file does not exist, kBad when it
// exists but is not a tuning cache.
static inline int read_file(const char* path, std::map<std::string, Entry>& out,
                            std::string& why) {
  out.clear();
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    if (errno == ENOENT) return kMissing;
    why = std::strerror(errno);
    return kBad;
  }
  std::string text;
  char buf[65536];
  size_t n = 0;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) {
    text.append(buf, n);
    if (text.size() > (64u << 20)) {
      std::fclose(f);
      why = "file larger than 64 MiB";
      return kBad;
    }
  }
  std::fclose(f);
  JVal root;
  Reader r(text.data(), text.size());
  if (!r.parse(root) || root.kind != JVal::Obj) {
    why = "not a JSON object";
    return kBad;
  }
  const JVal* fmt = root.get("format");
  if (!fmt || fmt->kind != JVal::Str || fmt->str != "cellflow-tuning") {
    why = "missing \"format\": \"cellflow-tuning\"";
    return kBad;
  }
  const JVal* ents = root.get("entries");
  if (!ents) return kOk;
  if (ents->kind != JVal::Obj) {
    why = "\"entries\" is not an object";
    return kBad;
  }
  for (const auto& kv : ents->obj) {
    Entry e;
    if (kv.second.kind != JVal::Obj) continue;
    if (!int_field(kv.second, "cell_divisor", &e.tun.cell_divisor) ||
        !int_field(kv.second, "force_variant", &e.tun.force_variant) ||
        !int_field(kv.second, "lanes_per_particle", &e.tun.lanes_per_particle) ||
        !int_field(kv.second, "block_size", &e.tun.block_size) ||
        !int_field(kv.second, "steps_per_graph", &e.tun.steps_per_graph))
      continue;  // incomplete entries are skipped (and dropped on the next store)
    e.tun.sort_variant = 1;
    if (kv.second.get("sort_variant") && !int_field(kv.second, "sort_variant", &e.tun.sort_variant))
      continue;
    if (!valid(e.tun)) continue;
    const JVal* us = kv.second.get("us_per_step");
    e.us = (us && us->kind == JVal::Num) ? us->num : 0.0;
    out[kv.first] = e;
  }
  return kOk;
}

// Writes all entries to `path` through a temporary file and a rename.
static inline bool write_file(const char* path, const std::map<std::string, Entry>& ents,
                              std::string& why) {
  std::string tmp = std::string(path) + ".tmp." + std::to_string((long)getpid());
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) {
    why = std::strerror(errno);
    return false;
  }
  std::fprintf(f, "{\n  \"format\": \"cellflow-tuning\",\n  \"version\": 1,\n  \"entries\": {");
  bool first = true;
  for (const auto& kv : ents) {
    const cf_tuning& t = kv.second.tun;
    // %d does not depend on the locale; the fraction is written with to_chars, which does
    // not either (a comma-decimal LC_NUMERIC would make "%.4f" invalid JSON)
    char us[64];
    const double usv = std::isfinite(kv.second.us) ? kv.second.us : 0.0;  // JSON has no NaN
    const std::to_chars_result r =
        std::to_chars(us, us + sizeof us - 1, usv, std::chars_format::fixed, 4);
    *r.ptr = '\0';
    std::fprintf(f,
                 "%s\n    \"%s\": {\"cell_divisor\": %d, \"force_variant\": %d, "
                 "\"lanes_per_particle\": %d, \"block_size\": %d, \"steps_per_graph\": %d, "
                 "\"sort_variant\": %d, \"us_per_step\": %s}",
                 first ? "" : ",", escape(kv.first).c_str(), t.cell_divisor, t.force_variant,
                 t.lanes_per_particle, t.block_size, t.steps_per_graph, t.sort_variant, us);
    first = false;
  }
  std::fprintf(f, "%s}\n}\n", first ? "" : "\n  ");
  const bool ok = std::fflush(f) == 0;
  std::fclose(f);
  if (!ok || std::rename(tmp.c_str(), path) != 0) {
    why = std::strerror(errno);
    std::remove(tmp.c_str());
    return false;
  }
  return true;
}

}  // namespace tunecache
}  // namespace cf
