// geofdtd — the reference's time-domain solver command (cmd/geofdtd/main.go)
// on libgeofdtd.
//
//   geofdtd -model medium.json -ranges 3,6,10 -o fdtd.csv
//
// Same flags, defaults, messages and CSV as the Go command (main.go:43-149,
// 204-238): the medium is a layer.Golden JSON file (only its name and layers
// are used), the source a Ricker wavelet, and the CSV holds time_s and one
// fdtd_r<range> column per receiver, every value as strconv.FormatFloat(v,
// 'g', 8, 64).  -compare also synthesises each trace through the
// frequency-wavenumber solver (synthesise, main.go:168-200; gf_fk_synthesise
// on the GPU), adds fk_r<range> columns and prints the comparison table
// (summarise, main.go:240-259).
//
// Host code only: everything on the GPU goes through the C ABI.
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/geofdtd.h"
#include "gofmt.h"

namespace {

// ------------------------------------------------------------------ errors

struct Failure {
  std::string msg;
};

[[noreturn]] void fail(const std::string &m) { throw Failure{m}; }

std::string lib_error() { return gf_last_error(); }

// Go's os error text: "open <path>: no such file or directory".
std::string os_error(const char *op, const std::string &path, int err) {
  std::string e = strerror(err);
  if (!e.empty()) e[0] = (char)tolower((unsigned char)e[0]);
  return std::string(op) + " " + path + ": " + e;
}

// ---------------------------------------------------- strconv.ParseFloat

// strconv.ParseFloat(s, 64): decimal or hexadecimal floats, "inf",
// "infinity", "nan" in any case, an optional sign, underscores only between
// digits of a prefixed number (not supported here: rejected as Go rejects
// them in plain decimals).  Errors name the function as Go does.
bool parse_float(const std::string &s, double &out, std::string &err) {
  const std::string q = gofmt::quote(s);
  if (s.empty() || isspace((unsigned char)s[0]) || isspace((unsigned char)s.back())) {
    err = "strconv.ParseFloat: parsing " + q + ": invalid syntax";
    return false;
  }
  std::string t = s;
  size_t i = (t[0] == '+' || t[0] == '-') ? 1 : 0;
  std::string body = t.substr(i);
  std::string lower = body;
  for (auto &c : lower) c = (char)tolower((unsigned char)c);
  if (lower == "inf" || lower == "infinity") {
    out = (t[0] == '-') ? -INFINITY : INFINITY;
    return true;
  }
  if (lower == "nan") {
    if (i) {
      err = "strconv.ParseFloat: parsing " + q + ": invalid syntax";
      return false;
    }
    out = NAN;
    return true;
  }
  // Only digits, '.', exponent, hex forms: strtod's own extensions (nan(...),
  // leading blanks) are excluded above and by the character check.
  for (char c : body)
    if (!(isxdigit((unsigned char)c) || c == '.' || c == 'x' || c == 'X' || c == 'p' || c == 'P' || c == '+' ||
          c == '-')) {
      err = "strconv.ParseFloat: parsing " + q + ": invalid syntax";
      return false;
    }
  errno = 0;
  char *end = nullptr;
  const double v = strtod(t.c_str(), &end);
  if (end != t.c_str() + t.size() || end == t.c_str()) {
    err = "strconv.ParseFloat: parsing " + q + ": invalid syntax";
    return false;
  }
  if (errno == ERANGE && isinf(v)) {
    err = "strconv.ParseFloat: parsing " + q + ": value out of range";
    return false;
  }
  out = v;
  return true;
}

// ------------------------------------------------------ Go's flag package

struct Flag {
  enum Kind { kFloat, kInt, kString, kBool } kind;
  std::string usage;
  std::string def;  // default as Go's flag.Value.String() prints it
  double f = 0;
  long long i = 0;
  std::string s;
  bool b = false;
};

class FlagSet {
 public:
  explicit FlagSet(std::string prog) : prog_(std::move(prog)) {}

  double *Float(const char *name, double v, const char *usage) {
    Flag &f = add(name, Flag::kFloat, usage);
    f.f = v;
    f.def = gofmt::g(v);
    return &f.f;
  }
  long long *Int(const char *name, long long v, const char *usage) {
    Flag &f = add(name, Flag::kInt, usage);
    f.i = v;
    f.def = std::to_string(v);
    return &f.i;
  }
  std::string *String(const char *name, const char *v, const char *usage) {
    Flag &f = add(name, Flag::kString, usage);
    f.s = v;
    f.def = v;
    return &f.s;
  }
  bool *Bool(const char *name, bool v, const char *usage) {
    Flag &f = add(name, Flag::kBool, usage);
    f.b = v;
    f.def = v ? "true" : "false";
    return &f.b;
  }

  // flag.Parse with ExitOnError: returns the exit code to use when parsing
  // stops the program (-h: 0, a bad flag: 2), or -1 to carry on.
  int parse(int argc, char **argv) {
    for (int a = 1; a < argc; a++) {
      std::string arg = argv[a];
      if (arg.size() < 2 || arg[0] != '-') return -1;  // first non-flag ends the flags
      size_t dashes = (arg[1] == '-') ? 2 : 1;
      if (dashes == 2 && arg.size() == 2) return -1;  // "--" terminator
      std::string name = arg.substr(dashes);
      if (name.empty() || name[0] == '-' || name[0] == '=') {
        fprintf(stderr, "bad flag syntax: %s\n", arg.c_str());
        usage();
        return 2;
      }
      bool has_value = false;
      std::string value;
      size_t eq = name.find('=');
      if (eq != std::string::npos) {
        value = name.substr(eq + 1);
        name = name.substr(0, eq);
        has_value = true;
      }
      auto it = flags_.find(name);
      if (it == flags_.end()) {
        if (name == "help" || name == "h") {
          usage();
          return 0;
        }
        fprintf(stderr, "flag provided but not defined: -%s\n", name.c_str());
        usage();
        return 2;
      }
      Flag &f = it->second;
      if (f.kind == Flag::kBool) {
        if (has_value) {
          static const char *t[] = {"1", "t", "T", "TRUE", "true", "True"};
          static const char *fl[] = {"0", "f", "F", "FALSE", "false", "False"};
          bool ok = false;
          for (const char *x : t)
            if (value == x) f.b = true, ok = true;
          for (const char *x : fl)
            if (value == x) f.b = false, ok = true;
          if (!ok) return bad_value(value, name, "parse error");
        } else {
          f.b = true;
        }
        continue;
      }
      if (!has_value) {
        if (a + 1 >= argc) {
          fprintf(stderr, "flag needs an argument: -%s\n", name.c_str());
          usage();
          return 2;
        }
        value = argv[++a];
      }
      if (f.kind == Flag::kString) {
        f.s = value;
      } else if (f.kind == Flag::kFloat) {
        std::string err;
        if (!parse_float(value, f.f, err))
          return bad_value(value, name, err.find("out of range") != std::string::npos ? "value out of range"
                                                                                      : "parse error");
      } else {
        errno = 0;
        char *end = nullptr;
        int base = 10;
        const char *p = value.c_str();
        std::string digits = value;
        bool neg = false;
        if (!digits.empty() && (digits[0] == '+' || digits[0] == '-')) neg = digits[0] == '-', digits = digits.substr(1);
        if (digits.size() > 1 && digits[0] == '0') {
          char c = (char)tolower((unsigned char)digits[1]);
          if (c == 'x') base = 16, digits = digits.substr(2);
          else if (c == 'b') base = 2, digits = digits.substr(2);
          else if (c == 'o') base = 8, digits = digits.substr(2);
          else base = 8, digits = digits.substr(1);
        }
        (void)p;
        long long v = digits.empty() ? 0 : strtoll(digits.c_str(), &end, base);
        if (digits.empty() || *end != '\0' || !isalnum((unsigned char)digits[0]))
          return bad_value(value, name, "parse error");
        if (errno == ERANGE) return bad_value(value, name, "value out of range");
        f.i = neg ? -v : v;
      }
    }
    return -1;
  }

  void usage() const {
    fprintf(stderr, "Usage of %s:\n", prog_.c_str());
    for (const auto &kv : flags_) {  // std::map: sorted by name, as Go prints them
      const Flag &f = kv.second;
      std::string line = "  -" + kv.first;
      const char *type = f.kind == Flag::kFloat ? "float" : f.kind == Flag::kInt ? "int"
                         : f.kind == Flag::kString ? "string" : "";
      if (*type) line += std::string(" ") + type;
      line += line.size() <= 4 ? "\t" : "\n    \t";
      line += f.usage;
      const bool zero = (f.kind == Flag::kFloat && f.f == 0 && f.def == "0") ||
                        (f.kind == Flag::kInt && f.def == "0") || (f.kind == Flag::kString && f.def.empty()) ||
                        (f.kind == Flag::kBool && f.def == "false");
      if (!zero) line += f.kind == Flag::kString ? " (default " + gofmt::quote(f.def) + ")" : " (default " + f.def + ")";
      fprintf(stderr, "%s\n", line.c_str());
    }
  }

 private:
  Flag &add(const char *name, Flag::Kind k, const char *usage) {
    Flag &f = flags_[name];
    f.kind = k;
    f.usage = usage;
    return f;
  }
  int bad_value(const std::string &value, const std::string &name, const char *why) {
    fprintf(stderr, "invalid value %s for flag -%s: %s\n", gofmt::quote(value).c_str(), name.c_str(), why);
    usage();
    return 2;
  }

  std::string prog_;
  std::map<std::string, Flag> flags_;
};

// -------------------------------------------------------- layer.Golden JSON

// A JSON value, enough of encoding/json's grammar to read a layer.Golden file.
struct Json {
  enum Type { kNull, kBool, kNumber, kString, kArray, kObject } type = kNull;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;
};

class JsonParser {
 public:
  explicit JsonParser(const std::string &t) : t_(t) {}

  // encoding/json's syntax error texts for the common cases.
  Json parse() {
    ws();
    Json v = value("looking for beginning of value");
    ws();
    if (p_ < t_.size()) bad("after top-level value");
    return v;
  }

 private:
  [[noreturn]] void bad(const char *ctx) {
    if (p_ >= t_.size()) fail("unexpected end of JSON input");
    const char c = t_[p_];
    std::string ch = c == '\'' ? "'\\''" : c == '"' ? "'\"'" : std::string("'") + c + "'";
    fail("invalid character " + ch + " " + ctx);
  }
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) p_++;
  }
  bool lit(const char *w) {
    size_t n = strlen(w);
    if (t_.compare(p_, n, w) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  Json value(const char *ctx) {
    Json v;
    if (p_ >= t_.size()) bad(ctx);
    const char c = t_[p_];
    if (c == '{') {
      v.type = Json::kObject;
      p_++;
      ws();
      if (p_ < t_.size() && t_[p_] == '}') {
        p_++;
        return v;
      }
      for (;;) {
        ws();
        if (p_ >= t_.size() || t_[p_] != '"') bad("looking for beginning of object key string");
        std::string k = string();
        ws();
        if (p_ >= t_.size() || t_[p_] != ':') bad("after object key");
        p_++;
        ws();
        Json e = value("looking for beginning of value");
        v.obj.emplace_back(std::move(k), std::move(e));
        ws();
        if (p_ < t_.size() && t_[p_] == ',') {
          p_++;
          continue;
        }
        if (p_ < t_.size() && t_[p_] == '}') {
          p_++;
          return v;
        }
        bad("after object key:value pair");
      }
    }
    if (c == '[') {
      v.type = Json::kArray;
      p_++;
      ws();
      if (p_ < t_.size() && t_[p_] == ']') {
        p_++;
        return v;
      }
      for (;;) {
        ws();
        v.arr.push_back(value("looking for beginning of value"));
        ws();
        if (p_ < t_.size() && t_[p_] == ',') {
          p_++;
          continue;
        }
        if (p_ < t_.size() && t_[p_] == ']') {
          p_++;
          return v;
        }
        bad("after array element");
      }
    }
    if (c == '"') {
      v.type = Json::kString;
      v.str = string();
      return v;
    }
    if (c == '-' || (c >= '0' && c <= '9')) {
      v.type = Json::kNumber;
      const size_t s = p_;
      if (t_[p_] == '-') p_++;
      if (p_ >= t_.size() || !isdigit((unsigned char)t_[p_])) bad("in numeric literal");
      if (t_[p_] == '0') p_++;
      else while (p_ < t_.size() && isdigit((unsigned char)t_[p_])) p_++;
      if (p_ < t_.size() && t_[p_] == '.') {
        p_++;
        if (p_ >= t_.size() || !isdigit((unsigned char)t_[p_])) bad("after decimal point in numeric literal");
        while (p_ < t_.size() && isdigit((unsigned char)t_[p_])) p_++;
      }
      if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
        p_++;
        if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) p_++;
        if (p_ >= t_.size() || !isdigit((unsigned char)t_[p_])) bad("in exponent of numeric literal");
        while (p_ < t_.size() && isdigit((unsigned char)t_[p_])) p_++;
      }
      v.num = strtod(t_.substr(s, p_ - s).c_str(), nullptr);
      return v;
    }
    if (lit("true")) {
      v.type = Json::kBool;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.type = Json::kBool;
      return v;
    }
    if (lit("null")) return v;
    bad(ctx);
  }
  std::string string() {
    std::string out;
    p_++;  // opening quote
    while (p_ < t_.size() && t_[p_] != '"') {
      char c = t_[p_++];
      if ((unsigned char)c < 0x20) {
        p_--;
        bad("in string literal");
      }
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= t_.size()) break;
      c = t_[p_++];
      switch (c) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          if (p_ + 4 > t_.size()) bad("in \\u hexadecimal character escape");
          unsigned cp = (unsigned)strtoul(t_.substr(p_, 4).c_str(), nullptr, 16);
          p_ += 4;
          if (cp >= 0xD800 && cp < 0xDC00 && p_ + 6 <= t_.size() && t_[p_] == '\\' && t_[p_ + 1] == 'u') {
            unsigned lo = (unsigned)strtoul(t_.substr(p_ + 2, 4).c_str(), nullptr, 16);
            if (lo >= 0xDC00 && lo < 0xE000) {
              cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
              p_ += 6;
            }
          }
          if (cp < 0x80) {
            out += (char)cp;
          } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
          } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
          } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
          }
          break;
        }
        default:
          p_--;
          bad("in string escape code");
      }
    }
    if (p_ >= t_.size()) fail("unexpected end of JSON input");
    p_++;  // closing quote
    return out;
  }

  const std::string &t_;
  size_t p_ = 0;
};

// encoding/json matches object keys to struct fields exactly first, then
// case-insensitively; the last matching key wins.
const Json *field(const Json &o, const char *name) {
  const Json *hit = nullptr;
  for (const auto &kv : o.obj)
    if (kv.first == name) hit = &kv.second;
  if (hit) return hit;
  for (const auto &kv : o.obj)
    if (strcasecmp(kv.first.c_str(), name) == 0) hit = &kv.second;
  return hit;
}

double number(const Json *v, const char *path, const char *go_type) {
  if (!v || v->type == Json::kNull) return 0;
  if (v->type != Json::kNumber) {
    const char *kind = v->type == Json::kString ? "string" : v->type == Json::kBool ? "bool"
                       : v->type == Json::kArray ? "array" : "object";
    fail(std::string("json: cannot unmarshal ") + kind + " into Go struct field " + path + " of type " + go_type);
  }
  return v->num;
}

struct Golden {
  std::string name;
  std::vector<gf_layer> layers;
};

// layer.LoadGolden (internal/layer/layer.go:130-143).
Golden load_golden(const std::string &path) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) fail(os_error("open", path, errno));
  std::string text;
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
  fclose(f);
  Golden g;
  try {
    Json root = JsonParser(text).parse();
    if (root.type == Json::kObject) {
      if (const Json *nm = field(root, "name"); nm && nm->type == Json::kString) g.name = nm->str;
      const Json *ls = field(root, "layers");
      if (ls && ls->type == Json::kArray) {
        for (const Json &l : ls->arr) {
          gf_layer L{};
          if (l.type == Json::kObject) {
            L.thickness_m = number(field(l, "thickness_m"), "Golden.layers.thickness_m", "units.Metres");
            L.vp_mps = number(field(l, "vp_mps"), "Golden.layers.vp_mps", "units.SpeedMPS");
            L.vs_mps = number(field(l, "vs_mps"), "Golden.layers.vs_mps", "units.SpeedMPS");
            L.density_kgm3 = number(field(l, "density_kgm3"), "Golden.layers.density_kgm3", "units.DensityKgM3");
            L.qs = number(field(l, "qs"), "Golden.layers.qs", "float64");
          } else if (l.type != Json::kNull) {
            fail("json: cannot unmarshal value into Go struct field Golden.layers of type layer.Layer");
          }
          g.layers.push_back(L);
        }
      } else if (ls && ls->type != Json::kNull) {
        fail("json: cannot unmarshal value into Go struct field Golden.layers of type layer.Stack");
      }
    } else if (root.type != Json::kNull) {
      fail("json: cannot unmarshal value into Go value of type layer.Golden");
    }
  } catch (const Failure &e) {
    fail("layer: parsing " + path + ": " + e.msg);
  }
  if (gf_stack_validate(g.layers.data(), (int32_t)g.layers.size()) != GF_OK)
    fail("layer: " + path + ": " + lib_error());
  return g;
}

// parseRanges (main.go:151-167).
std::vector<double> parse_ranges(const std::string &s) {
  std::vector<double> out;
  size_t start = 0;
  for (;;) {
    size_t comma = s.find(',', start);
    std::string part = s.substr(start, comma == std::string::npos ? std::string::npos : comma - start);
    size_t a = 0, b = part.size();
    while (a < b && isspace((unsigned char)part[a])) a++;
    while (b > a && isspace((unsigned char)part[b - 1])) b--;
    double v;
    std::string err;
    if (!parse_float(part.substr(a, b - a), v, err)) fail("parsing range " + gofmt::quote(part) + ": " + err);
    if (v <= 0) fail("range must be positive, got " + gofmt::g(v));
    out.push_back(v);
    if (comma == std::string::npos) break;
    start = comma + 1;
  }
  if (out.empty()) fail("no ranges given");
  return out;
}

struct SimHandle {
  gf_sim *s = nullptr;
  ~SimHandle() {
    if (s) gf_sim_destroy(s);
  }
};

struct RunResult {
  double dt = 0, t0 = 0, peak = 0;
  std::vector<double> snapped;
  std::vector<double> traces;  // [receiver][step]
};

// fdtd.Run with a Progress callback (run.go:88-127), on the device in blocks
// that end where the reference reports progress (every 500 steps) or probes
// for divergence (every 256).
RunResult run_shot(const gf_model &m, const std::vector<double> &force, const std::vector<double> &ranges, int steps,
                   bool progress) {
  SimHandle h;
  if (gf_sim_create(&m, &h.s) != GF_OK) fail(lib_error());
  if (steps <= 0) fail("fdtd: a shot needs a positive number of steps");
  if (gf_sim_drive(h.s, force.data(), (int64_t)force.size()) != GF_OK) fail(lib_error());
  RunResult r;
  int32_t nr, nz;
  gf_sim_info(h.s, &r.dt, &nr, &nz);
  r.t0 = 0.5 * r.dt;
  std::vector<int32_t> cols;
  for (double rg : ranges) {
    int32_t c;
    double sn;
    gf_sim_column(h.s, rg, &c, &sn);
    if (sn >= m.max_range_m)
      fail("fdtd: receiver at " + gofmt::g(rg) + " m is outside the interior (" + gofmt::g(m.max_range_m) + " m)");
    cols.push_back(c);
    r.snapped.push_back(sn);
  }
  const size_t R = cols.size();
  r.traces.assign(R * (size_t)steps, 0.0);
  std::vector<double> block;
  int n = 0;
  while (n < steps) {
    int nxt = n;
    while (nxt < steps - 1 && nxt % 500 != 0 && nxt % 256 != 0) nxt++;
    const int len = nxt - n + 1;
    block.assign(R * (size_t)len, 0.0);
    if (gf_sim_step_n(h.s, len, cols.data(), (int32_t)R, R ? block.data() : nullptr) != GF_OK) fail(lib_error());
    for (size_t c = 0; c < R; c++)
      std::copy(block.begin() + c * len, block.begin() + (c + 1) * len, r.traces.begin() + c * steps + n);
    if (progress && nxt % 500 == 0) {
      fprintf(stderr, "\r  %3d%%", (int)(100LL * nxt / steps));
      fflush(stderr);
    }
    if (nxt % 256 == 0) {
      double p;
      if (gf_sim_max_speed(h.s, &p) != GF_OK) fail(lib_error());
      r.peak = (isnan(p) || p > r.peak) ? p : r.peak;  // math.Max propagates NaN
      if (isnan(p) || isinf(p)) fail("fdtd: diverged at step " + std::to_string(nxt) + " of " + std::to_string(steps));
    }
    n = nxt + 1;
  }
  return r;
}

// write (main.go:204-238): encoding/csv with no field needing quotes.
// write (main.go:204-238); ref, when not empty, holds the f-k traces
// [receiver][step] of -compare.
void write_csv(const std::string &path, const RunResult &r, int steps, const std::vector<double> &ref) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) fail(os_error("open", path, errno));
  std::string line = "time_s";
  char buf[64];
  for (double sn : r.snapped) {
    snprintf(buf, sizeof buf, ",fdtd_r%.2f", sn);
    line += buf;
  }
  if (!ref.empty())
    for (double sn : r.snapped) {
      snprintf(buf, sizeof buf, ",fk_r%.2f", sn);
      line += buf;
    }
  line += '\n';
  fwrite(line.data(), 1, line.size(), f);
  const size_t R = r.snapped.size();
  for (int n = 0; n < steps; n++) {
    line = gofmt::g_prec(r.t0 + (double)n * r.dt, 8);
    for (size_t c = 0; c < R; c++) {
      line += ',';
      line += gofmt::g_prec(r.traces[c * steps + n], 8);
    }
    if (!ref.empty())
      for (size_t c = 0; c < R; c++) {
        line += ',';
        line += gofmt::g_prec(ref[c * steps + n], 8);
      }
    line += '\n';
    fwrite(line.data(), 1, line.size(), f);
  }
  if (fclose(f) != 0) fail(os_error("write", path, errno));
}

// math.Max (NaN wins).
double go_max(double x, double y) {
  if (std::isinf(x) && x > 0) return x;
  if (std::isinf(y) && y > 0) return y;
  if (std::isnan(x) || std::isnan(y)) return NAN;
  return x > y ? x : y;
}

// fmt's verbs for a float with C's digits and Go's Inf/NaN spelling.
std::string go_num(const char *cfmt, int width, double x) {
  char b[64];
  if (std::isnan(x) || std::isinf(x)) {
    const char *w = std::isnan(x) ? "NaN" : x > 0 ? "+Inf" : "-Inf";
    snprintf(b, sizeof b, "%*s", width, w);
  } else {
    snprintf(b, sizeof b, cfmt, x);
  }
  return b;
}

// summarise (main.go:240-259): peaks, their ratio and the rms residual of
// each receiver against the f-k trace.
void summarise(const RunResult &r, int steps, const std::vector<double> &ref) {
  fprintf(stderr, "\n  %8s  %12s  %12s  %8s  %8s\n", "range", "grid peak", "f-k peak", "ratio", "rms");
  for (size_t i = 0; i < r.snapped.size(); i++) {
    double pa = 0, pb = 0, num = 0, den = 0;
    for (int n = 0; n < steps; n++) {
      double a = r.traces[i * steps + n], b = ref[i * steps + n];
      pa = go_max(pa, fabs(a));
      pb = go_max(pb, fabs(b));
      num += (a - b) * (a - b);
      den += b * b;
    }
    std::string line = "  " + go_num("%7.2f", 7, r.snapped[i]) + "m  " + go_num("%12.4e", 12, pa) + "  " +
                       go_num("%12.4e", 12, pb) + "  " + go_num("%8.4f", 8, pa / pb) + "  " +
                       go_num("%7.2f", 7, 100 * sqrt(num / den)) + "%\n";
    fputs(line.c_str(), stderr);
  }
  fprintf(stderr,
          "\n  the residual is dominated by grid dispersion and falls linearly\n"
          "  with -spacing; extrapolating two runs to zero spacing is what V5 asserts\n\n");
}

int run(int argc, char **argv) {
  FlagSet fs(argv[0]);
  std::string *model_path = fs.String("model", "", "layered model JSON (required)");
  std::string *out_path = fs.String("o", "fdtd.csv", "output CSV path");
  std::string *range_list = fs.String("ranges", "4,8,16", "receiver ranges in metres, comma separated");
  double *f0 = fs.Float("f0", 25, "Ricker wavelet peak frequency, Hz");
  double *band = fs.Float("band", 0, "highest frequency to resolve, Hz (default 3.5*f0)");
  double *ppw = fs.Float("ppw", 20, "grid cells across the shortest wavelength");
  double *spacing = fs.Float("spacing", 0, "grid spacing in metres, overriding -ppw");
  double *seconds = fs.Float("seconds", 0.3, "how long to run");
  double *peak = fs.Float("force", 1000, "wavelet peak force, N");
  double *q = fs.Float("q", 25, "quality factor at the reference frequency; 0 for elastic");
  double *ref_freq = fs.Float("ref-freq", 30, "frequency at which nominal velocities are phase velocities, Hz");
  double *depth = fs.Float("depth", 0, "domain depth in metres (default 1.2x the furthest range)");
  double *margin = fs.Float("margin", 1.5, "domain radius as a multiple of the furthest range");
  bool *compare = fs.Bool("compare", false, "also synthesise through the frequency-wavenumber path");
  long long *ref_samples = fs.Int("ref-samples", 0, "wavenumber samples for the reference; 0 lets the solver choose");
  const int code = fs.parse(argc, argv);
  if (code >= 0) return code;

  if (model_path->empty()) fail("-model is required");
  if (*band <= 0) *band = 3.5 * *f0;
  Golden g = load_golden(*model_path);
  std::vector<double> ranges = parse_ranges(*range_list);
  const double furthest = ranges.back();

  gf_sls relax{0, 0};
  if (*q > 0) {
    if (gf_visco_fit(*q, *ref_freq, &relax) != GF_OK) fail(lib_error());
  } else if (*compare) {
    fail("-compare needs attenuation: with no loss the Rayleigh pole sits on the wavenumber integration path "
         "and the reference does not exist");
  }

  gf_model m{};
  m.stack = g.layers.data();
  m.n_layers = (int32_t)g.layers.size();
  m.relax = relax;
  m.ref_freq_hz = *ref_freq;
  m.max_range_m = *margin * furthest;
  m.depth_m = std::max(*depth, 1.2 * furthest);
  m.dominant_freq_hz = *f0;
  if (*spacing > 0) {
    m.spacing_m = *spacing;
  } else if (gf_spacing_for(&m, *band, *ppw, &m.spacing_m) != GF_OK) {
    fail(lib_error());
  }

  double dt;
  int32_t nr, nz;
  {
    SimHandle probe;  // fdtd.New, as the reference's probe (main.go:99-104)
    if (gf_sim_create(&m, &probe.s) != GF_OK) fail(lib_error());
    gf_sim_info(probe.s, &dt, &nr, &nz);
  }
  const int steps = (int)(*seconds / dt);
  fprintf(stderr, "model %s: %d layers, spacing %.4f m, grid %dx%d, dt %.3e s, %d steps\n", g.name.c_str(),
          (int)g.layers.size(), m.spacing_m, nr, nz, dt, steps);

  std::vector<double> force((size_t)std::max(steps, 0));
  if (gf_ricker(*peak, *f0, dt, steps, force.data()) != GF_OK) fail(lib_error());
  const auto start = std::chrono::steady_clock::now();
  RunResult res = run_shot(m, force, ranges, steps, true);
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
  fprintf(stderr, "\r  grid run %.1f s\n", el);
  std::vector<double> ref;
  if (*compare) {
    // synthesise (main.go:168-200) on the instants the grid recorded, at the
    // snapped ranges, through the frequency-wavenumber solver on the GPU.
    if (*ref_samples > INT32_MAX) fail("-ref-samples is larger than the f-k solver takes");
    const auto s0 = std::chrono::steady_clock::now();
    gf_medium med{};
    med.stack = g.layers.data();
    med.n_layers = (int32_t)g.layers.size();
    med.ref_freq_hz = *ref_freq;
    med.has_relax = 1;
    med.relax = relax;
    ref.assign(res.snapped.size() * (size_t)std::max(steps, 0), 0.0);
    if (gf_fk_synthesise(&med, res.snapped.data(), (int32_t)res.snapped.size(), *peak, *f0, *band, res.dt, res.t0,
                         steps, (int32_t)std::max<long long>(*ref_samples, 0), ref.data()) != GF_OK)
      fail(lib_error());
    const double rs = std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
    fprintf(stderr, "  reference %.1f s\n", rs);
  }
  write_csv(*out_path, res, steps, ref);
  if (*compare) summarise(res, steps, ref);
  fprintf(stderr, "wrote %s\n", out_path->c_str());
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const Failure &e) {
    fprintf(stderr, "geofdtd: %s\n", e.msg.c_str());
    return 1;
  }
}
