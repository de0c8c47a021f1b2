// Host-only entry points and helpers of the C ABI: the error string, tuning validation and
// the tuning-cache file calls.  No device code.

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>

#include "cf_host.h"
#include "cf_tune_cache.h"

namespace cf {

std::string& error_text() {
  static thread_local std::string text;
  return text;
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  error_text() = buf;
  return code;
}

int check_tuning(const cf_tuning* t) {
  if (!t) return fail(CF_ERR_ARG, "tuning is null");
  if (t->cell_divisor != 1 && t->cell_divisor != 2)
    return fail(CF_ERR_ARG, "cell_divisor must be 1 or 2");
  if (t->block_size != 64 && t->block_size != 128 && t->block_size != 256)
    return fail(CF_ERR_ARG, "block_size must be 64, 128 or 256");
  if (t->force_variant < 0 || t->force_variant > 4)
    return fail(CF_ERR_ARG, "force_variant must be 0, 1, 2, 3 or 4");
  if (t->force_variant == 4 && t->lanes_per_particle != 8 && t->lanes_per_particle != 16)
    return fail(CF_ERR_ARG, "force_variant 4 needs lanes_per_particle 8 or 16");
  if (t->lanes_per_particle != 4 && t->lanes_per_particle != 8 && t->lanes_per_particle != 16 &&
      t->lanes_per_particle != 32)
    return fail(CF_ERR_ARG, "lanes_per_particle must be 4, 8, 16 or 32");
  if (t->force_variant == 3 && t->lanes_per_particle < 4 * t->cell_divisor)
    return fail(CF_ERR_ARG, "force_variant 3 needs lanes_per_particle >= 8 with cell_divisor 2");
  if (t->steps_per_graph < 1) return fail(CF_ERR_ARG, "steps_per_graph must be >= 1");
  if (t->sort_variant != 0 && t->sort_variant != 1)
    return fail(CF_ERR_ARG, "sort_variant must be 0 or 1");
  return CF_OK;
}

}  // namespace cf

extern "C" {

const char* cf_last_error_string(void) { return cf::error_text().c_str(); }

int cf_tuning_cache_lookup(const char* cache_path, const char* key, cf_tuning* tuning,
                           double* us_per_step, int* found) {
  if (!cache_path || !key || !tuning || !found) return cf::fail(CF_ERR_ARG, "null argument");
  *found = 0;
  std::map<std::string, cf::tunecache::Entry> ents;
  std::string why;
  const int st = cf::tunecache::read_file(cache_path, ents, why);
  if (st == cf::tunecache::kBad)
    return cf::fail(CF_ERR_ARG, "tuning cache %s: %s", cache_path, why.c_str());
  auto it = ents.find(key);
  if (it == ents.end() || !cf::tunecache::valid(it->second.tun)) return CF_OK;
  *tuning = it->second.tun;
  if (us_per_step) *us_per_step = it->second.us;
  *found = 1;
  return CF_OK;
}

int cf_tuning_cache_store(const char* cache_path, const char* key, const cf_tuning* tuning,
                          double us_per_step) {
  if (!cache_path || !key) return cf::fail(CF_ERR_ARG, "null argument");
  const int rc = cf::check_tuning(tuning);
  if (rc != CF_OK) return rc;
  if (!std::isfinite(us_per_step)) return cf::fail(CF_ERR_ARG, "us_per_step must be finite");
  std::map<std::string, cf::tunecache::Entry> ents;
  std::string why;
  const int st = cf::tunecache::read_file(cache_path, ents, why);
  if (st == cf::tunecache::kBad)
    return cf::fail(CF_ERR_ARG, "tuning cache %s: %s", cache_path, why.c_str());
  ents[key] = cf::tunecache::Entry{*tuning, us_per_step};
  if (!cf::tunecache::write_file(cache_path, ents, why))
    return cf::fail(CF_ERR_ARG, "cannot write tuning cache %s: %s", cache_path, why.c_str());
  return CF_OK;
}

}  // extern "C"
