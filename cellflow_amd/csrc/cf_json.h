// Small JSON reader and writer shared by the tuning cache and the preset files.  Host-only.
//
// Reader: RFC 8259, nesting limited to 32 levels, strings decoded to UTF-8.
// Writer: the indented layout of Qt's QJsonDocument::toJson (the format of the reference's
// preset files): keys sorted, four spaces per level, one member or element per line,
// integral numbers below 2^53 without a fraction, other numbers in the shortest form that
// reads back to the same double (Python repr's layout, which is also what the reference's
// files hold), a newline at the end.
//
// Numbers are converted with std::from_chars / std::to_chars, which do not depend on the
// C locale: a host application that sets LC_NUMERIC to a comma-decimal locale (a Qt
// application does, from the environment) still reads "0.25" as 0.25 and writes '.'.
#pragma once

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace cf {
namespace json {

// A JSON number's text [b, e) as a double, locale-free.  Magnitudes beyond the double
// range are rejected; ones below it read as a zero of the text's sign.
static inline bool parse_double(const char* b, const char* e, double& v) {
  const std::from_chars_result r = std::from_chars(b, e, v);
  if (r.ptr != e) return false;
  if (r.ec == std::errc()) return true;
  if (r.ec != std::errc::result_out_of_range) return false;
  const char* x = b;
  while (x < e && *x != 'e' && *x != 'E') ++x;
  if (x + 1 < e && x[1] == '-') {  // underflow
    v = (*b == '-') ? -0.0 : 0.0;
    return true;
  }
  return false;
}

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
    if (!parse_double(p_, q, v)) return false;
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

static inline std::string escape(const std::string& s) {
  std::string o;
  for (unsigned char c : s) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += (char)c;
    } else if (c == '\n') {
      o += "\\n";
    } else if (c == '\r') {
      o += "\\r";
    } else if (c == '\t') {
      o += "\\t";
    } else if (c == '\b') {
      o += "\\b";
    } else if (c == '\f') {
      o += "\\f";
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

// Shortest decimal significand `digits` (no trailing zeros) and exponent `x` (value =
// d.ddd... * 10^x) that read back to |v|: std::to_chars without a precision gives exactly
// that (the shortest round-trip form, correctly rounded), in scientific layout.
static inline void shortest(double v, std::string& digits, int& x) {
  char b[64];
  const std::to_chars_result r =
      std::to_chars(b, b + sizeof b, std::fabs(v), std::chars_format::scientific);
  const char* e = std::find(b, r.ptr, 'e');
  std::string d;
  for (const char* c = b; c < e; ++c)
    if (*c != '.') d += *c;
  int xe = 0;
  const char* xs = e + 1;
  if (xs < r.ptr && *xs == '+') ++xs;  // from_chars takes no leading '+'
  std::from_chars(xs, r.ptr, xe);
  while (d.size() > 1 && d.back() == '0') d.pop_back();
  digits = d;
  x = xe;
}

// Number text: integral values below 2^53 as integers, negative zero as -0.0 (it keeps its
// sign through any reader), everything else as the shortest round-trip decimal in Python's
// repr layout (fixed for 1e-4 <= |v| < 1e16, else d.ddde+XX).
static inline std::string number(double v) {
  char b[40];
  if (v == 0.0 && std::signbit(v)) return "-0.0";
  if (std::isfinite(v) && v == std::floor(v) && std::fabs(v) < 9007199254740992.0) {
    const std::to_chars_result r = std::to_chars(b, b + sizeof b, v, std::chars_format::fixed, 0);
    return std::string(b, r.ptr);
  }
  if (!std::isfinite(v)) return "null";  // not representable in JSON
  std::string d;
  int x = 0;
  shortest(v, d, x);
  std::string o = std::signbit(v) ? "-" : "";
  const int n = (int)d.size();
  if (x >= -4 && x < 16) {
    if (x >= 0) {
      if (n <= x + 1) {
        o += d + std::string(x + 1 - n, '0') + ".0";
      } else {
        o += d.substr(0, x + 1) + "." + d.substr(x + 1);
      }
    } else {
      o += "0." + std::string(-x - 1, '0') + d;
    }
  } else {
    o += d.substr(0, 1);
    if (n > 1) o += "." + d.substr(1);
    std::snprintf(b, sizeof b, "e%c%02d", x < 0 ? '-' : '+', std::abs(x));
    o += b;
  }
  return o;
}

static inline void write(const JVal& v, std::string& o, int indent) {
  const std::string pad(4 * (indent + 1), ' '), end(4 * indent, ' ');
  switch (v.kind) {
    case JVal::Null: o += "null"; break;
    case JVal::Bool: o += v.b ? "true" : "false"; break;
    case JVal::Num: o += number(v.num); break;
    case JVal::Str: o += '"' + escape(v.str) + '"'; break;
    case JVal::Arr:
      if (v.arr.empty()) {
        o += "[]";
        break;
      }
      o += "[\n";
      for (size_t i = 0; i < v.arr.size(); ++i) {
        o += pad;
        write(v.arr[i], o, indent + 1);
        o += (i + 1 < v.arr.size()) ? ",\n" : "\n";
      }
      o += end + "]";
      break;
    case JVal::Obj: {
      if (v.obj.empty()) {
        o += "{}";
        break;
      }
      std::vector<const std::pair<std::string, JVal>*> m;
      for (const auto& kv : v.obj) m.push_back(&kv);
      std::stable_sort(m.begin(), m.end(),
                       [](const auto* a, const auto* b) { return a->first < b->first; });
      o += "{\n";
      for (size_t i = 0; i < m.size(); ++i) {
        o += pad + '"' + escape(m[i]->first) + "\": ";
        write(m[i]->second, o, indent + 1);
        o += (i + 1 < m.size()) ? ",\n" : "\n";
      }
      o += end + "}";
      break;
    }
  }
}

// Whole document in the layout above.
static inline std::string dump(const JVal& v) {
  std::string o;
  write(v, o, 0);
  o += "\n";
  return o;
}

// Value constructors and object member update.
static inline JVal num(double x) {
  JVal v;
  v.kind = JVal::Num;
  v.num = x;
  return v;
}

static inline void set(JVal& obj, const std::string& key, JVal val) {
  for (auto& kv : obj.obj)
    if (kv.first == key) {
      kv.second = std::move(val);
      return;
    }
  obj.obj.emplace_back(key, std::move(val));
}

}  // namespace json
}  // namespace cf
