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
//           "us_per_step": 74.1},
//       ...
//     }
//   }
// Unknown top-level fields and unknown entry fields are ignored on read.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include <unistd.h>

#include "../../include/cellflow.h"

namespace cf {
namespace tunecache {

struct JVal {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;

  const JVal* get(const char* k) const {
    if (kind != Obj) return nullptr;
    for (const auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

// Recursive-descent reader for RFC 8259 JSON; nesting is limited to 32 levels.
class Reader {
 public:
  Reader(const char* s, size_t n) : p_(s), end_(s + n) {}

  bool parse(JVal& out) {
    ws();
    if (!value(out, 0)) return false;
    ws();
    return p_ == end_;
  }

 private:
  const char* p_;
  const char* end_;

  void ws() {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
  }

  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if ((size_t)(end_ - p_) < n || std::strncmp(p_, w, n) != 0) return false;
    p_ += n;
    return true;
  }

  static void utf8(std::string& s, unsigned c) {
    if (c < 0x80) {
      s += (char)c;
    } else if (c < 0x800) {
      s += (char)(0xC0 | (c >> 6));
      s += (char)(0x80 | (c & 0x3F));
    } else if (c < 0x10000) {
      s += (char)(0xE0 | (c >> 12));
      s += (char)(0x80 | ((c >> 6) & 0x3F));
      s += (char)(0x80 | (c & 0x3F));
    } else {
      s += (char)(0xF0 | (c >> 18));
      s += (char)(0x80 | ((c >> 12) & 0x3F));
      s += (char)(0x80 | ((c >> 6) & 0x3F));
      s += (char)(0x80 | (c & 0x3F));
    }
  }

  bool hex4(unsigned& v) {
    if (end_ - p_ < 4) return false;
    v = 0;
    for (int i = 0; i < 4; ++i) {
      const char c = *p_++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else return false;
    }
    return true;
  }

  bool string(std::string& s) {
    if (p_ >= end_ || *p_ != '"') return false;
    ++p_;
    while (p_ < end_) {
      const char c = *p_++;
      if (c == '"') return true;
      if ((unsigned char)c < 0x20) return false;
      if (c != '\\') {
        s += c;
        continue;
      }
      if (p_ >= end_) return false;
      const char e = *p_++;
      switch (e) {
        case '"': s += '"'; break;
        case '\\': s += '\\'; break;
        case '/': s += '/'; break;
        case 'b': s += '\b'; break;
        case 'f': s += '\f'; break;
        case 'n': s += '\n'; break;
        case 'r': s += '\r'; break;
        case 't': s += '\t'; break;
        case 'u': {
          unsigned c1 = 0;
          if (!hex4(c1)) return false;
          if (c1 >= 0xD800 && c1 < 0xDC00) {  // surrogate pair
            unsigned c2 = 0;
            if (!lit("\\u") || !hex4(c2) || c2 < 0xDC00 || c2 >= 0xE000) return false;
            c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
          } else if (c1 >= 0xDC00 && c1 < 0xE000) {
            return false;
          }
          utf8(s, c1);
          break;
        }
        default: return false;
      }
    }
    return false;
  }

  bool number(double& v) {
    const char* q = p_;
    if (q < end_ && *q == '-') ++q;
    if (q >= end_) return false;
    if (*q == '0') {
      ++q;
    } else if (*q >= '1' && *q <= '9') {
      while (q < end_ && *q >= '0' && *q <= '9') ++q;
    } else {
      return false;
    }
    if (q < end_ && *q == '.') {
      ++q;
      if (q >= end_ || *q < '0' || *q > '9') return false;
      while (q < end_ && *q >= '0' && *q <= '9') ++q;
    }
    if (q < end_ && (*q == 'e' || *q == 'E')) {
      ++q;
      if (q < end_ && (*q == '+' || *q == '-')) ++q;
      if (q >= end_ || *q < '0' || *q > '9') return false;
      while (q < end_ && *q >= '0' && *q <= '9') ++q;
    }
    const std::string text(p_, q);
    v = std::strtod(text.c_str(), nullptr);
    p_ = q;
    return true;
  }

  bool value(JVal& v, int depth) {
    if (depth > 32 || p_ >= end_) return false;
    const char c = *p_;
    if (c == '{') {
      ++p_;
      v.kind = JVal::Obj;
      ws();
      if (p_ < end_ && *p_ == '}') {
        ++p_;
        return true;
      }
      while (true) {
        ws();
        std::string k;
        if (!string(k)) return false;
        ws();
        if (p_ >= end_ || *p_ != ':') return false;
        ++p_;
        ws();
        JVal item;
        if (!value(item, depth + 1)) return false;
        v.obj.emplace_back(std::move(k), std::move(item));
        ws();
        if (p_ < end_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < end_ && *p_ == '}') {
          ++p_;
          return true;
        }
        return false;
      }
    }
    if (c == '[') {
      ++p_;
      v.kind = JVal::Arr;
      ws();
      if (p_ < end_ && *p_ == ']') {
        ++p_;
        return true;
      }
      while (true) {
        ws();
        JVal item;
        if (!value(item, depth + 1)) return false;
        v.arr.push_back(std::move(item));
        ws();
        if (p_ < end_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < end_ && *p_ == ']') {
          ++p_;
          return true;
        }
        return false;
      }
    }
    if (c == '"') {
      v.kind = JVal::Str;
      return string(v.str);
    }
    if (lit("true")) {
      v.kind = JVal::Bool;
      v.b = true;
      return true;
    }
    if (lit("false")) {
      v.kind = JVal::Bool;
      return true;
    }
    if (lit("null")) return true;
    v.kind = JVal::Num;
    return number(v.num);
  }
};

struct Entry {
  cf_tuning tun{};
  double us = 0.0;
};

static inline std::string escape(const std::string& s) {
  std::string o;
  for (unsigned char c : s) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += (char)c;
    } else if (c < 0x20) {
      char b[8];
      std::snprintf(b, sizeof b, "\\u%04x", c);
      o += b;
    } else {
      o += (char)c;
    }
  }
  return o;
}

static inline bool int_field(const JVal& e, const char* k, int* out) {
  const JVal* v = e.get(k);
  if (!v || v->kind != JVal::Num || v->num != std::floor(v->num) || std::fabs(v->num) > 1e9)
    return false;
  *out = (int)v->num;
  return true;
}

// The values cf_set_tuning accepts.
static inline bool valid(const cf_tuning& t) {
  return (t.cell_divisor == 1 || t.cell_divisor == 2) &&
         (t.block_size == 64 || t.block_size == 128 || t.block_size == 256) &&
         (t.force_variant >= 0 && t.force_variant <= 2) &&
         (t.lanes_per_particle == 4 || t.lanes_per_particle == 8 || t.lanes_per_particle == 16 ||
          t.lanes_per_particle == 32) &&
         t.steps_per_graph >= 1;
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
    std::fprintf(f,
                 "%s\n    \"%s\": {\"cell_divisor\": %d, \"force_variant\": %d, "
                 "\"lanes_per_particle\": %d, \"block_size\": %d, \"steps_per_graph\": %d, "
                 "\"us_per_step\": %.4f}",
                 first ? "" : ",", escape(kv.first).c_str(), t.cell_divisor, t.force_variant,
                 t.lanes_per_particle, t.block_size, t.steps_per_graph, kv.second.us);
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
