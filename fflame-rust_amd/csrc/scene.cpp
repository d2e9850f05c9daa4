// Config loading: the reference's on-disk flame formats read into the
// descriptors this engine renders.
//
//  - FractalConfig JSON (`.fflame` / `.fae`, one object or an array of them):
//    src/config/fractal_config.rs:178-700 (fields and serde defaults),
//    :1746-2050 (from_json / from_json_multi / version-keyed migration),
//    src/scene/transforms.rs:1307-1560 (Transform serde),
//    :2567-2690 (Flame serde, legacy `final_transform`),
//    src/scene/palette.rs:839-915 (Palette serde: `indexed_colors` or `stops`).
//  - Apophysis / JWildfire XML (`.flame`): parse_flame_xml,
//    src/flame_xml.rs:41-560 (flame attributes and the config it builds),
//    :600-1000 (xform / finalxform attributes), :1014-1067 (palette).
//
// Parsing happens in two phases. Phase 1 reads every config in the text into
// a plain intermediate form, failing with FFM_ERR_PARSE where the
// reference's reader fails (bad syntax, a missing required field, a wrong
// type). Phase 2 turns the requested config into an ffm_scene and refuses,
// with FFM_ERR_UNSUPPORTED and the feature's name, anything this engine does
// not render exactly as the reference would; nothing is approximated.
//
// The JSON and XML readers below are small purpose-built parsers (no
// third-party dependency): JSON per RFC 8259, XML as far as flame files use
// it (elements, attributes, text, comments, CDATA, processing instructions,
// a DOCTYPE without internal entities).

#include <algorithm>
#include <array>
#include <clocale>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <locale.h>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "ffm.h"

namespace {

// ---------------------------------------------------------------- errors

struct Fail {
  int code;
  std::string msg;
};

[[noreturn]] void raise(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Fail{code, buf};
}

void put_err(char* err, size_t cap, const std::string& m) {
  if (!err || cap == 0) return;
  size_t n = std::min(cap - 1, m.size());
  memcpy(err, m.data(), n);
  err[n] = 0;
}

// Locale-independent number conversion (the reference's Rust parsers do not
// read the process locale).
locale_t c_locale() {
  static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
  return loc;
}

// Rust's `str::parse::<f32>()` grammar: an optional sign, then "inf",
// "infinity" or "nan" (any case), or decimal digits with an optional point
// and exponent; nothing else, no surrounding space. The value is the
// correctly rounded binary32, as Rust's.
bool parse_f32(const std::string& s, float* out) {
  size_t i = 0, n = s.size();
  if (n == 0) return false;
  if (s[0] == '+' || s[0] == '-') i = 1;
  std::string rest = s.substr(i);
  for (auto& ch : rest) ch = (char)tolower((unsigned char)ch);
  if (rest == "inf" || rest == "infinity" || rest == "nan") {
    float v = rest == "nan" ? NAN : INFINITY;
    *out = s[0] == '-' ? -v : v;
    return true;
  }
  size_t digits = 0;
  while (i < n && isdigit((unsigned char)s[i])) i++, digits++;
  if (i < n && s[i] == '.') {
    i++;
    while (i < n && isdigit((unsigned char)s[i])) i++, digits++;
  }
  if (digits == 0) return false;
  if (i < n && (s[i] == 'e' || s[i] == 'E')) {
    i++;
    if (i < n && (s[i] == '+' || s[i] == '-')) i++;
    size_t ed = 0;
    while (i < n && isdigit((unsigned char)s[i])) i++, ed++;
    if (ed == 0) return false;
  }
  if (i != n) return false;
  *out = strtof_l(s.c_str(), nullptr, c_locale());
  return true;
}

// The same grammar read as binary64 (Rust's `str::parse::<f64>()`).
bool parse_f64(const std::string& s, double* out) {
  float probe;
  if (!parse_f32(s, &probe)) return false;
  *out = strtod_l(s.c_str(), nullptr, c_locale());
  return true;
}

// Rust's `str::parse::<u32>()` / `<usize>()`: an optional '+', then digits.
bool parse_uint(const std::string& s, uint64_t max, uint64_t* out) {
  size_t i = (!s.empty() && s[0] == '+') ? 1 : 0;
  if (i >= s.size()) return false;
  uint64_t v = 0;
  for (; i < s.size(); i++) {
    if (!isdigit((unsigned char)s[i])) return false;
    v = v * 10 + (uint64_t)(s[i] - '0');
    if (v > max) return false;
  }
  *out = v;
  return true;
}

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && isspace((unsigned char)s[i])) i++;
    size_t j = i;
    while (j < s.size() && !isspace((unsigned char)s[j])) j++;
    if (j > i) out.push_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

// ------------------------------------------------------------------ JSON

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } k = NUL;
  bool b = false;
  double n = 0;
  std::string s;
  std::vector<JVal> a;
  std::vector<std::pair<std::string, JVal>> o;  // in file order

  // A repeated key reads as its last occurrence (serde_json's map).
  const JVal* get(const char* key) const {
    if (k != OBJ) return nullptr;
    for (auto it = o.rbegin(); it != o.rend(); ++it)
      if (it->first == key) return &it->second;
    return nullptr;
  }
};

class JsonReader {
 public:
  JsonReader(const char* p, size_t n) : p_(p), n_(n) {}

  JVal document() {
    JVal v = value(0);
    ws();
    if (i_ != n_) error("trailing characters");
    return v;
  }

 private:
  const char* p_;
  size_t n_, i_ = 0;

  [[noreturn]] void error(const char* what) {
    size_t line = 1, col = 1;
    for (size_t k = 0; k < i_ && k < n_; k++) {
      if (p_[k] == '\n') line++, col = 1;
      else col++;
    }
    raise(FFM_ERR_PARSE, "JSON: %s at line %zu column %zu", what, line, col);
  }
  void ws() {
    while (i_ < n_ && (p_[i_] == ' ' || p_[i_] == '\t' || p_[i_] == '\n' || p_[i_] == '\r')) i_++;
  }
  bool lit(const char* w) {
    size_t m = strlen(w);
    if (n_ - i_ >= m && memcmp(p_ + i_, w, m) == 0) {
      i_ += m;
      return true;
    }
    return false;
  }
  JVal value(int depth) {
    if (depth > 128) error("nesting too deep");
    ws();
    if (i_ >= n_) error("EOF while parsing a value");
    JVal v;
    char ch = p_[i_];
    if (ch == '{') {
      v.k = JVal::OBJ;
      i_++;
      ws();
      if (i_ < n_ && p_[i_] == '}') {
        i_++;
        return v;
      }
      for (;;) {
        ws();
        if (i_ >= n_ || p_[i_] != '"') error("expected a key string");
        std::string key = str();
        ws();
        if (i_ >= n_ || p_[i_] != ':') error("expected ':'");
        i_++;
        v.o.emplace_back(std::move(key), value(depth + 1));
        ws();
        if (i_ < n_ && p_[i_] == ',') {
          i_++;
          continue;
        }
        if (i_ < n_ && p_[i_] == '}') {
          i_++;
          return v;
        }
        error("expected ',' or '}'");
      }
    }
    if (ch == '[') {
      v.k = JVal::ARR;
      i_++;
      ws();
      if (i_ < n_ && p_[i_] == ']') {
        i_++;
        return v;
      }
      for (;;) {
        v.a.push_back(value(depth + 1));
        ws();
        if (i_ < n_ && p_[i_] == ',') {
          i_++;
          continue;
        }
        if (i_ < n_ && p_[i_] == ']') {
          i_++;
          return v;
        }
        error("expected ',' or ']'");
      }
    }
    if (ch == '"') {
      v.k = JVal::STR;
      v.s = str();
      return v;
    }
    if (lit("true")) {
      v.k = JVal::BOOL;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.k = JVal::BOOL;
      return v;
    }
    if (lit("null")) return v;
    if (ch == '-' || (ch >= '0' && ch <= '9')) {
      size_t s = i_;
      if (p_[i_] == '-') i_++;
      if (i_ < n_ && p_[i_] == '0') {
        i_++;
      } else {
        if (i_ >= n_ || !isdigit((unsigned char)p_[i_])) error("invalid number");
        while (i_ < n_ && isdigit((unsigned char)p_[i_])) i_++;
      }
      if (i_ < n_ && p_[i_] == '.') {
        i_++;
        if (i_ >= n_ || !isdigit((unsigned char)p_[i_])) error("invalid number");
        while (i_ < n_ && isdigit((unsigned char)p_[i_])) i_++;
      }
      if (i_ < n_ && (p_[i_] == 'e' || p_[i_] == 'E')) {
        i_++;
        if (i_ < n_ && (p_[i_] == '+' || p_[i_] == '-')) i_++;
        if (i_ >= n_ || !isdigit((unsigned char)p_[i_])) error("invalid number");
        while (i_ < n_ && isdigit((unsigned char)p_[i_])) i_++;
      }
      std::string tok(p_ + s, i_ - s);
      v.k = JVal::NUM;
      v.n = strtod_l(tok.c_str(), nullptr, c_locale());
      return v;
    }
    error("expected a value");
  }
  unsigned hex4() {
    if (n_ - i_ < 4) error("invalid \\u escape");
    unsigned u = 0;
    for (int k = 0; k < 4; k++) {
      char c = p_[i_++];
      u <<= 4;
      if (c >= '0' && c <= '9') u |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') u |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') u |= (unsigned)(c - 'A' + 10);
      else error("invalid \\u escape");
    }
    return u;
  }
  static void utf8(std::string& out, unsigned cp) {
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
  }
  std::string str() {
    i_++;  // opening quote
    std::string out;
    for (;;) {
      if (i_ >= n_) error("EOF while parsing a string");
      char c = p_[i_++];
      if (c == '"') return out;
      if ((unsigned char)c < 0x20) error("control character in a string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (i_ >= n_) error("EOF while parsing a string");
      char e = p_[i_++];
      switch (e) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (n_ - i_ < 6 || p_[i_] != '\\' || p_[i_ + 1] != 'u') error("lone surrogate");
            i_ += 2;
            unsigned lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) error("lone surrogate");
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            error("lone surrogate");
          }
          utf8(out, cp);
          break;
        }
        default: error("invalid escape");
      }
    }
  }
};

// ------------------------------------------------------------------- XML

struct XAttr {
  std::string key, value;  // value as written (entities left as they are)
};
struct XEvent {
  enum Kind { START, EMPTY, END, TEXT, END_OF_FILE } k = END_OF_FILE;
  std::string name;  // element name, or the text (entities resolved)
  std::vector<XAttr> attrs;
};

class XmlReader {
 public:
  XmlReader(const char* p, size_t n) : p_(p), n_(n) {}

  XEvent next() {
    for (;;) {
      if (i_ >= n_) {
        if (!open_.empty()) error("unclosed element <%s>", open_.back().c_str());
        return XEvent{};
      }
      if (p_[i_] != '<') {
        size_t s = i_;
        while (i_ < n_ && p_[i_] != '<') i_++;
        std::string t = trim(std::string(p_ + s, i_ - s));
        if (t.empty()) continue;
        XEvent e;
        e.k = XEvent::TEXT;
        e.name = unescape(t);
        return e;
      }
      if (starts("<!--")) {
        size_t e = find("-->", i_ + 4);
        i_ = e + 3;
        continue;
      }
      if (starts("<![CDATA[")) {
        size_t e = find("]]>", i_ + 9);
        XEvent ev;
        ev.k = XEvent::TEXT;
        ev.name = std::string(p_ + i_ + 9, e - i_ - 9);
        i_ = e + 3;
        if (trim(ev.name).empty()) continue;
        return ev;
      }
      if (starts("<?")) {
        i_ = find("?>", i_ + 2) + 2;
        continue;
      }
      if (starts("<!")) {  // DOCTYPE and the like; an internal subset in [ ]
        int depth = 0;
        size_t k = i_ + 2;
        for (; k < n_; k++) {
          if (p_[k] == '[') depth++;
          else if (p_[k] == ']') depth--;
          else if (p_[k] == '>' && depth <= 0) break;
        }
        if (k >= n_) error("unterminated declaration");
        i_ = k + 1;
        continue;
      }
      if (starts("</")) {
        i_ += 2;
        std::string nm = name();
        skip_ws();
        if (i_ >= n_ || p_[i_] != '>') error("malformed end tag </%s", nm.c_str());
        i_++;
        if (open_.empty() || open_.back() != nm)
          error("end tag </%s> does not match the open element", nm.c_str());
        open_.pop_back();
        XEvent ev;
        ev.k = XEvent::END;
        ev.name = nm;
        return ev;
      }
      i_++;  // '<'
      XEvent ev;
      ev.name = name();
      if (ev.name.empty()) error("malformed tag");
      for (;;) {
        skip_ws();
        if (i_ >= n_) error("unterminated tag <%s", ev.name.c_str());
        if (p_[i_] == '>') {
          i_++;
          ev.k = XEvent::START;
          open_.push_back(ev.name);
          return ev;
        }
        if (p_[i_] == '/') {
          if (i_ + 1 >= n_ || p_[i_ + 1] != '>') error("malformed tag <%s", ev.name.c_str());
          i_ += 2;
          ev.k = XEvent::EMPTY;
          return ev;
        }
        XAttr a;
        a.key = name();
        if (a.key.empty()) error("malformed attribute in <%s>", ev.name.c_str());
        skip_ws();
        if (i_ >= n_ || p_[i_] != '=') error("attribute %s without a value", a.key.c_str());
        i_++;
        skip_ws();
        if (i_ >= n_ || (p_[i_] != '"' && p_[i_] != '\'')) error("unquoted attribute %s", a.key.c_str());
        char q = p_[i_++];
        size_t s = i_;
        while (i_ < n_ && p_[i_] != q) i_++;
        if (i_ >= n_) error("unterminated attribute %s", a.key.c_str());
        a.value.assign(p_ + s, i_ - s);
        i_++;
        for (auto& b : ev.attrs)
          if (b.key == a.key) error("duplicated attribute %s in <%s>", a.key.c_str(), ev.name.c_str());
        ev.attrs.push_back(std::move(a));
      }
    }
  }

 private:
  const char* p_;
  size_t n_, i_ = 0;
  std::vector<std::string> open_;

  [[noreturn]] void error(const char* fmt, ...) {
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    raise(FFM_ERR_PARSE, "XML parse error at position %zu: %s", i_, buf);
  }
  bool starts(const char* w) const {
    size_t m = strlen(w);
    return n_ - i_ >= m && memcmp(p_ + i_, w, m) == 0;
  }
  size_t find(const char* w, size_t from) {
    size_t m = strlen(w);
    for (size_t k = from; k + m <= n_; k++)
      if (memcmp(p_ + k, w, m) == 0) return k;
    error("unterminated markup");
  }
  void skip_ws() {
    while (i_ < n_ && isspace((unsigned char)p_[i_])) i_++;
  }
  std::string name() {
    size_t s = i_;
    while (i_ < n_ && !isspace((unsigned char)p_[i_]) && p_[i_] != '>' && p_[i_] != '/' &&
           p_[i_] != '=' && p_[i_] != '<')
      i_++;
    return std::string(p_ + s, i_ - s);
  }
  static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) a++;
    while (b > a && isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
  }
  std::string unescape(const std::string& t) {
    std::string out;
    for (size_t k = 0; k < t.size(); k++) {
      if (t[k] != '&') {
        out += t[k];
        continue;
      }
      size_t e = t.find(';', k);
      if (e == std::string::npos) error("unterminated entity");
      std::string ent = t.substr(k + 1, e - k - 1);
      if (ent == "lt") out += '<';
      else if (ent == "gt") out += '>';
      else if (ent == "amp") out += '&';
      else if (ent == "quot") out += '"';
      else if (ent == "apos") out += '\'';
      else if (ent.size() > 1 && ent[0] == '#') {
        unsigned long cp = ent[1] == 'x' ? strtoul(ent.c_str() + 2, nullptr, 16)
                                         : strtoul(ent.c_str() + 1, nullptr, 10);
        if (cp < 0x80) out += (char)cp;
        else out += '?';
      } else {
        error("unknown entity &%s;", ent.c_str());
      }
      k = e;
    }
    return out;
  }
};

// ------------------------------------------------- variations this path has

struct VarInfo {
  const char* name;
  int id;                      // registry index (src/variations/defs/mod.rs:446-...)
  const char* params[3];       // user parameters, slot order
  float defaults[3];
  int n_params;
};

// The registry's own names and parameter defaults for the variations the
// kernels implement (src/variations/defs/basic.rs, advanced.rs, blur.rs),
// in registry order (the order a transform's variations take after its
// variation_order).
const VarInfo kVars[] = {
    {"linear", FFM_VAR_LINEAR, {}, {}, 0},
    {"sinusoidal", FFM_VAR_SINUSOIDAL, {}, {}, 0},
    {"spherical", FFM_VAR_SPHERICAL, {}, {}, 0},
    {"swirl", FFM_VAR_SWIRL, {}, {}, 0},
    {"horseshoe", FFM_VAR_HORSESHOE, {}, {}, 0},
    {"polar", FFM_VAR_POLAR, {}, {}, 0},
    {"disc", FFM_VAR_DISC, {}, {}, 0},
    {"hyperbolic", FFM_VAR_HYPERBOLIC, {}, {}, 0},
    {"julia", FFM_VAR_JULIA, {}, {}, 0},
    {"waves", FFM_VAR_WAVES, {}, {}, 0},
    {"julian", FFM_VAR_JULIAN, {"power", "dist"}, {2.0f, 1.0f}, 2},
    {"blob", FFM_VAR_BLOB, {"high", "low", "waves"}, {1.0f, 1.0f, 6.0f}, 3},
    {"bubble", FFM_VAR_BUBBLE, {}, {}, 0},
    {"noise", FFM_VAR_NOISE, {}, {}, 0},
    {"blur", FFM_VAR_BLUR, {}, {}, 0},
    {"gaussian_blur", FFM_VAR_GAUSSIAN_BLUR, {}, {}, 0},
    {"curl", FFM_VAR_CURL, {"c1", "c2"}, {1.0f, 0.0f}, 2},
    {"pre_blur", FFM_VAR_PRE_BLUR, {}, {}, 0},
};

const VarInfo* var_by_name(const std::string& n) {
  for (const auto& v : kVars)
    if (n == v.name) return &v;
  return nullptr;
}

// Variations that leave the plane alone: `flatten` is a post-phase
// z := 0 (src/variations/defs/depth3d.rs:50-77), whose 2D form returns its
// input, so with x and y unchanged it draws nothing here.
bool plane_noop(const std::string& n) { return n == "flatten"; }

const VarInfo* var_by_id(int id) {
  for (const auto& v : kVars)
    if (v.id == id) return &v;
  return nullptr;
}

// -------------------------------------------------- intermediate config form

struct XfIn {
  float a = 1, b = 0, c = 0, d = 1, e = 0, f = 0, g = 0;
  float weight = 1, color = 0.5f, color_speed = 0, opacity = 1;
  bool post = false;
  float pa = 1, pb = 0, pc = 0, pd = 1, pe = 0, pf = 0, pg = 0;
  bool plane_identity = true;  // yz / zx (post) plane affines
  // name -> weight, as the reference's HashMap holds them (a key present
  // with weight 0 still takes a local index)
  std::vector<std::pair<std::string, float>> vars;
  std::vector<std::string> order;                       // variation_order
  std::vector<std::pair<std::string, float>> params;    // "name.param" -> value
  std::vector<std::pair<std::string, int>> priorities;  // variation_priorities
  std::vector<int64_t> final_att;
  size_t n_linked_att = 0;

  bool has_var(const std::string& n) const {
    for (auto& v : vars)
      if (v.first == n) return true;
    return false;
  }
  float* var(const std::string& n) {
    for (auto& v : vars)
      if (v.first == n) return &v.second;
    return nullptr;
  }
  void set_param(const std::string& key, float v) {
    for (auto& p : params)
      if (p.first == key) {
        p.second = v;
        return;
      }
    params.emplace_back(key, v);
  }
};

struct SceneIn {
  std::string name = "Untitled";
  std::vector<XfIn> xforms, finals;
  size_t n_linked = 0;
  std::vector<std::vector<float>> xaos;  // Flame::xaos as written (rows may be short)
  bool solo = false;
  bool subflames = false;
  bool post_symmetry = false;
  bool custom_variations = false;

  std::string render_mode = "2d";
  bool preserve_z = false;
  bool reset_orbits = false;
  float zoom = 1, rotation = 0;
  double pan_x = 0, pan_y = 0;
  // 3D-only state; a 3D config renders through the 2D chaos game only
  // when all of it is neutral
  float cam[8] = {0};  // pitch, yaw, bank, x, y, z, perspective, dof
  float depth_fx[5] = {0};  // fog, solid, depth compensation, far fade, far fade start
  uint32_t image_w = 1920, image_h = 1080;
  float filter_radius = 0, filter_blur_edges = 0;
  float density_scale = 1;
  uint64_t max_iterations = 1000000000ull;
  std::string color_mode = "palette";
  std::vector<std::array<float, 4>> stops;
  bool palette_set = false;
  float palette_rotation = 0;
  uint32_t palette_size = 256;
  float palette_squeeze = 1, palette_squeeze_falloff = 0.5f, palette_log_strength = 0;
  std::string palette_squeeze_mode = "linear";
  bool palette_reverse = false;
  float background[3] = {0, 0, 0};
  std::string tonemap_mode = "logarithmic", highlight_mode = "clip";
  bool use_curve = true;
  std::vector<float> curve_x, curve_y;  // tone curve points; empty: the linear curve
  float exposure = 1, gamma = 4, gamma_threshold = 5, brightness = 4, vibrancy = 1;
  float white_level = 200, saturation = 1, hue_shift = 0;
  float alpha_blend_low = 0.3f, alpha_blend_high = 0.8f;
  bool levels_enabled = true;
  float levels_low = 0, levels_high = 10, levels_gamma = 0.5f;
  bool deterministic_rng = false;
  ffm_effects effects{};                 // the enabled built-in effects, parameters filled
  std::vector<std::string> unsupported;  // config features present that this path lacks
};


// ------------------------------------------------------- JSON -> SceneIn

float jnum_f32(const JVal* v, const char* what) {
  if (!v || v->k != JVal::NUM) raise(FFM_ERR_PARSE, "%s: expected a number", what);
  return (float)v->n;  // serde: the JSON number as f64, then `as f32`
}

float jopt_f32(const JVal& o, const char* key, float dflt) {
  const JVal* v = o.get(key);
  if (!v) return dflt;
  return jnum_f32(v, key);
}

bool jopt_bool(const JVal& o, const char* key, bool dflt) {
  const JVal* v = o.get(key);
  if (!v) return dflt;
  if (v->k != JVal::BOOL) raise(FFM_ERR_PARSE, "%s: expected true or false", key);
  return v->b;
}

std::string jopt_str(const JVal& o, const char* key, const std::string& dflt) {
  const JVal* v = o.get(key);
  if (!v) return dflt;
  if (v->k != JVal::STR) raise(FFM_ERR_PARSE, "%s: expected a string", key);
  return v->s;
}

uint64_t jopt_u64(const JVal& o, const char* key, uint64_t dflt) {
  const JVal* v = o.get(key);
  if (!v) return dflt;
  if (v->k != JVal::NUM || v->n < 0 || v->n != std::floor(v->n) || v->n > 1.8446744073709552e19)
    raise(FFM_ERR_PARSE, "%s: expected an unsigned integer", key);
  return (uint64_t)v->n;
}

bool nonempty_array(const JVal& o, const char* key) {
  const JVal* v = o.get(key);
  return v && v->k == JVal::ARR && !v->a.empty();
}

// The built-in effects this path runs, by registry name, with their
// parameter names in declaration order (assets/effects/<name>.rhai PARAMS).
struct FxName {
  const char* name;
  int32_t type;
  int chain;  // 1 density, 2 colour
  const char* params[FFM_MAX_EFFECT_PARAMS];
};
const FxName kEffects[] = {
    {"density_blur", FFM_FX_DENSITY_BLUR, 1, {"radius", "threshold", "falloff"}},
    {"bilateral_blur", FFM_FX_BILATERAL_BLUR, 1, {"radius", "sigma_spatial", "sigma_range"}},
    {"sharpen", FFM_FX_SHARPEN, 1, {"amount", "radius"}},
    {"vignette", FFM_FX_VIGNETTE, 2, {"intensity", "radius", "softness", "blend_mode"}},
    {"film_grain", FFM_FX_FILM_GRAIN, 2, {"intensity", "seed", "blend_mode"}},
    {"chromatic_aberration", FFM_FX_CHROMATIC_ABERRATION, 2,
     {"amount", "radial", "intensity", "blend_mode"}},
    {"hue_shift", FFM_FX_HUE_SHIFT, 2, {"offset", "intensity", "blend_mode"}},
    {"sobel_edges", FFM_FX_SOBEL_EDGES, 2,
     {"intensity", "threshold", "glow", "preserve_color", "blend_mode"}},
    {"kaleidoscope", FFM_FX_KALEIDOSCOPE, 2,
     {"segments", "rotation", "zoom", "intensity", "blend_mode", "square_mode", "edge_offset"}},
    {"plasma", FFM_FX_PLASMA, 2, {"intensity", "scale", "speed", "time", "blend_mode", "direction"}},
    {"tunnel", FFM_FX_TUNNEL, 2,
     {"speed", "rotation_speed", "distortion", "time", "intensity", "blend_mode"}},
    {"julia", FFM_FX_JULIA, 2,
     {"mode", "julia_cx", "julia_cy", "max_iter", "zoom", "center_x", "center_y", "color_mode",
      "intensity", "blend_mode"}},
    {"worley_noise", FFM_FX_WORLEY_NOISE, 2,
     {"intensity", "scale", "time", "mode", "color_mode", "blend_mode", "direction"}},
    {"simplex_noise", FFM_FX_SIMPLEX_NOISE, 2,
     {"intensity", "scale", "octaves", "time", "mode", "blend_mode", "direction"}},
    {"domain_warp", FFM_FX_DOMAIN_WARP, 2,
     {"intensity", "scale", "octaves", "time", "blend_mode", "direction"}},
    {"bloom", FFM_FX_BLOOM, 2, {"threshold", "softness", "intensity", "spread"}},
    {"distance_field", FFM_FX_DISTANCE_FIELD, 2,
     {"threshold", "spread", "mode", "intensity", "band_count"}},
    {"trails", FFM_FX_TRAILS, 2, {"persistence", "intensity"}},
};

// A config's effect list (Vec<EffectInstance>, src/effects/mod.rs:319-336:
// effect_type, enabled, params by name; a missing parameter takes the
// registry's default). Disabled effects are dropped, as the chain skips
// them (effect_chain.rs:745); an enabled one this path does not run is
// named in `unsupported`.
void json_effects(const JVal& root, const char* key, int chain, SceneIn& s) {
  const JVal* arr = root.get(key);
  if (!arr || arr->k == JVal::NUL) return;
  if (arr->k != JVal::ARR) raise(FFM_ERR_PARSE, "%s: expected an array", key);
  for (const JVal& e : arr->a) {
    if (e.k != JVal::OBJ) raise(FFM_ERR_PARSE, "%s: expected objects", key);
    const std::string type = jopt_str(e, "effect_type", "");
    const JVal* en = e.get("enabled");
    if (!en || en->k != JVal::BOOL) raise(FFM_ERR_PARSE, "%s: effect '%s' needs enabled", key, type.c_str());
    if (!en->b) continue;
    const FxName* d = nullptr;
    for (const FxName& f : kEffects)
      if (type == f.name && f.chain == chain) d = &f;
    if (!d) {
      s.unsupported.push_back(std::string(chain == 1 ? "density effect '" : "color effect '") +
                              type + "'");
      continue;
    }
    int& n = chain == 1 ? s.effects.n_density : s.effects.n_color;
    if (n == FFM_MAX_EFFECTS) {
      s.unsupported.push_back(std::string("more than 16 ") + key);
      continue;
    }
    ffm_effect fx;
    ffm_effect_defaults(d->type, &fx);
    if (const JVal* p = e.get("params"); p && p->k != JVal::NUL) {
      if (p->k != JVal::OBJ) raise(FFM_ERR_PARSE, "%s: params must be an object", key);
      for (const auto& kv : p->o)
        for (int i = 0; i < FFM_MAX_EFFECT_PARAMS && d->params[i]; i++)
          if (kv.first == d->params[i]) fx.params[i] = jnum_f32(&kv.second, d->params[i]);
    }
    (chain == 1 ? s.effects.density : s.effects.color)[n++] = fx;
  }
}

// A transform (Transform's Deserialize, src/scene/transforms.rs:1307-1560).
XfIn json_xform(const JVal& t, const char* where) {
  if (t.k != JVal::OBJ) raise(FFM_ERR_PARSE, "%s: expected a transform object", where);
  XfIn x;
  auto req = [&](const char* k) {
    const JVal* v = t.get(k);
    if (!v) raise(FFM_ERR_PARSE, "%s: missing field `%s`", where, k);
    return jnum_f32(v, k);
  };
  x.a = req("a");
  x.b = req("b");
  x.c = req("c");
  x.d = req("d");
  x.e = req("e");
  x.f = req("f");
  x.g = jopt_f32(t, "g", 0.0f);
  x.weight = req("weight");
  const JVal* vars = t.get("variations");
  if (!vars) raise(FFM_ERR_PARSE, "%s: missing field `variations`", where);
  if (vars->k == JVal::OBJ) {
    for (auto& kv : vars->o) {
      if (kv.second.k != JVal::NUM) continue;  // non-numbers are skipped, as the reference does
      float w = (float)kv.second.n;
      if (float* p = x.var(kv.first)) *p = w;
      else x.vars.emplace_back(kv.first, w);
    }
  } else if (vars->k == JVal::ARR) {
    // legacy form: weights by registry index, non-zero ones only
    for (size_t i = 0; i < vars->a.size(); i++) {
      const JVal& w = vars->a[i];
      if (w.k != JVal::NUM) continue;
      float wf = (float)w.n;
      if (std::fabs(wf) <= 1e-6f) continue;
      const VarInfo* vi = var_by_id((int)i);
      if (!vi) {
        x.vars.emplace_back("#" + std::to_string(i), wf);  // refused in phase 2
        continue;
      }
      if (float* p = x.var(vi->name)) *p = wf;
      else x.vars.emplace_back(vi->name, wf);
    }
  } else {
    raise(FFM_ERR_PARSE, "%s: variations must be an object (new format) or array (legacy format)",
          where);
  }
  if (const JVal* ps = t.get("variation_params")) {
    if (ps->k != JVal::OBJ) raise(FFM_ERR_PARSE, "%s: variation_params must be an object", where);
    for (auto& kv : ps->o) x.set_param(kv.first, jnum_f32(&kv.second, "variation_params"));
  }
  if (const JVal* pr = t.get("variation_priorities")) {
    if (pr->k != JVal::OBJ) raise(FFM_ERR_PARSE, "%s: variation_priorities must be an object", where);
    for (auto& kv : pr->o) {
      if (kv.second.k != JVal::NUM) raise(FFM_ERR_PARSE, "%s: bad variation priority", where);
      x.priorities.emplace_back(kv.first, (int)kv.second.n);
    }
  }
  if (const JVal* vo = t.get("variation_order")) {
    if (vo->k != JVal::ARR) raise(FFM_ERR_PARSE, "%s: variation_order must be a list", where);
    for (auto& s : vo->a) {
      if (s.k != JVal::STR) raise(FFM_ERR_PARSE, "%s: variation_order holds names", where);
      x.order.push_back(s.s);
    }
  }
  const JVal* col = t.get("color");
  if (!col) raise(FFM_ERR_PARSE, "%s: missing field `color`", where);
  if (col->k == JVal::NUM) {
    x.color = (float)col->n;
  } else if (col->k == JVal::ARR && col->a.size() >= 3) {
    // the old RGB form, averaged (transforms.rs:1436-1452)
    float r = col->a[0].k == JVal::NUM ? (float)col->a[0].n : 0.0f;
    float g = col->a[1].k == JVal::NUM ? (float)col->a[1].n : 0.0f;
    float b = col->a[2].k == JVal::NUM ? (float)col->a[2].n : 0.0f;
    x.color = (r + g + b) / 3.0f;
  } else {
    x.color = 0.5f;
  }
  x.color_speed = jopt_f32(t, "color_speed", 0.0f);
  x.opacity = jopt_f32(t, "opacity", 1.0f);
  x.post = jopt_bool(t, "post_affine_enabled", false);
  x.pa = jopt_f32(t, "post_a", 1.0f);
  x.pb = jopt_f32(t, "post_b", 0.0f);
  x.pc = jopt_f32(t, "post_c", 0.0f);
  x.pd = jopt_f32(t, "post_d", 1.0f);
  x.pe = jopt_f32(t, "post_e", 0.0f);
  x.pf = jopt_f32(t, "post_f", 0.0f);
  x.pg = jopt_f32(t, "post_g", 0.0f);
  static const float ident[6] = {1, 0, 0, 1, 0, 0};
  for (const char* k : {"yz_coefs", "zx_coefs", "yz_post_coefs", "zx_post_coefs"}) {
    const JVal* v = t.get(k);
    if (!v) continue;
    if (v->k != JVal::ARR || v->a.size() != 6) raise(FFM_ERR_PARSE, "%s: %s must hold 6 numbers", where, k);
    for (int q = 0; q < 6; q++)
      if (jnum_f32(&v->a[q], k) != ident[q]) x.plane_identity = false;
  }
  if (const JVal* la = t.get("linked_attachments")) {
    if (la->k != JVal::ARR) raise(FFM_ERR_PARSE, "%s: linked_attachments must be a list", where);
    x.n_linked_att = la->a.size();
  }
  if (const JVal* fa = t.get("final_attachments")) {
    if (fa->k != JVal::ARR) raise(FFM_ERR_PARSE, "%s: final_attachments must be a list", where);
    for (auto& v : fa->a) {
      if (v.k != JVal::NUM || v.n < 0 || v.n != std::floor(v.n))
        raise(FFM_ERR_PARSE, "%s: final_attachments holds indices", where);
      x.final_att.push_back((int64_t)v.n);
    }
  }
  return x;
}

// legacy_projection_to_strength (src/config/fractal_config.rs:849-860)
float projection_strength(const JVal& v) {
  if (v.k == JVal::OBJ) {
    const JVal* p = v.get("Perspective");
    const JVal* s = p ? p->get("strength") : nullptr;
    if (s && s->k == JVal::NUM) return (float)s->n;
  }
  return 0.0f;
}

// tonemap_curve: {"points": [{"x": .., "y": ..}, ..]} (ToneCurve,
// src/scene/tonemap.rs:142-149); absent, the linear curve (no points kept).
void json_curve(const JVal* c, SceneIn& s) {
  if (!c || c->k == JVal::NUL) return;
  const JVal* pts = c->get("points");
  if (!pts || pts->k != JVal::ARR) raise(FFM_ERR_PARSE, "tonemap_curve: missing field `points`");
  for (auto& p : pts->a) {
    s.curve_x.push_back(jnum_f32(p.get("x"), "tonemap_curve.points.x"));
    s.curve_y.push_back(jnum_f32(p.get("y"), "tonemap_curve.points.y"));
  }
  if (s.curve_x.size() < 2) raise(FFM_ERR_PARSE, "tonemap_curve: at least two points are needed");
}

void json_palette(const JVal& p, SceneIn& s) {
  if (p.k == JVal::NUL) return;  // `palette: null` -> the default
  if (p.k != JVal::OBJ) raise(FFM_ERR_PARSE, "palette: expected an object");
  if (!p.get("name")) raise(FFM_ERR_PARSE, "palette: missing field `name`");
  s.stops.clear();
  if (const JVal* hex = p.get("indexed_colors")) {
    // from_hex_string (src/scene/palette.rs:453-480): "RRGGBB,..." at i/255
    if (hex->k != JVal::STR) raise(FFM_ERR_PARSE, "palette: indexed_colors must be a string");
    const std::string& h = hex->s;
    size_t i = 0, idx = 0;
    for (;;) {
      size_t e = h.find(',', i);
      std::string tok = h.substr(i, e == std::string::npos ? std::string::npos : e - i);
      if (tok.size() != 6) raise(FFM_ERR_PARSE, "palette: invalid hex color '%s' at index %zu", tok.c_str(), idx);
      std::array<float, 4> st;
      st[0] = (float)idx / 255.0f;
      for (int c = 0; c < 3; c++) {
        char* end = nullptr;
        std::string two = tok.substr(2 * c, 2);
        long v = strtol(two.c_str(), &end, 16);
        if (!end || *end || !isxdigit((unsigned char)two[0]) || !isxdigit((unsigned char)two[1]))
          raise(FFM_ERR_PARSE, "palette: invalid hex color '%s'", tok.c_str());
        st[1 + c] = (float)v / 255.0f;
      }
      s.stops.push_back(st);
      idx++;
      if (e == std::string::npos) break;
      i = e + 1;
    }
  } else if (const JVal* stops = p.get("stops")) {
    if (stops->k != JVal::ARR) raise(FFM_ERR_PARSE, "palette: stops must be a list");
    for (auto& st : stops->a) {
      const JVal* pos = st.get("position");
      const JVal* col = st.get("color");
      if (!pos || !col || col->k != JVal::ARR || col->a.size() != 3)
        raise(FFM_ERR_PARSE, "palette: a stop needs `position` and a 3-number `color`");
      std::array<float, 4> v = {jnum_f32(pos, "position"), jnum_f32(&col->a[0], "color"),
                                jnum_f32(&col->a[1], "color"), jnum_f32(&col->a[2], "color")};
      s.stops.push_back(v);
    }
  } else {
    raise(FFM_ERR_PARSE, "palette: missing field `indexed_colors or stops`");
  }
  s.palette_set = true;
}

// FractalConfig::from_json_value_with_default_version + migrate_value +
// Deserialize (src/config/fractal_config.rs:1789-2050), for the fields
// this path reads.
SceneIn json_config(const JVal& root) {
  if (root.k != JVal::OBJ) raise(FFM_ERR_PARSE, "expected a config object");
  SceneIn s;
  // the version the file was written at
  uint32_t version = 0;
  if (const JVal* v = root.get("version"); v && v->k == JVal::NUM && v->n >= 0) {
    version = (uint32_t)v->n;
  } else {
    const JVal* mode = root.get("render_mode");
    const JVal* fl = root.get("flame");
    bool on_flame = fl && fl->get("render_mode");
    bool mode_str = mode && mode->k == JVal::STR;
    if ((mode_str && !on_flame) ||
        (mode_str && (mode->s == "escape" || mode->s == "simulation")))
      version = 3;
  }
  uint32_t min_reader = version;
  if (const JVal* m = root.get("min_reader"); m && m->k == JVal::NUM) min_reader = (uint32_t)m->n;
  if (min_reader > 7)
    raise(FFM_ERR_PARSE,
          "Config version %u is newer than supported version 7. Please update the application.",
          min_reader);

  const JVal* flame = root.get("flame");
  if (!flame) raise(FFM_ERR_PARSE, "missing field `flame`");
  if (flame->k != JVal::OBJ) raise(FFM_ERR_PARSE, "flame: expected an object");

  // v2 -> v3: scene-level render fields lived on the flame
  if (version < 3) {
    const JVal* rm = flame->get("render_mode");
    s.render_mode = rm && rm->k == JVal::STR ? rm->s : "2d";
    if (const JVal* ps = flame->get("perspective_strength")) s.cam[6] = jnum_f32(ps, "perspective_strength");
    else if (const JVal* pj = flame->get("projection")) s.cam[6] = projection_strength(*pj);
    s.preserve_z = jopt_bool(*flame, "preserve_z", true);
    s.depth_fx[2] = jopt_f32(*flame, "depth_density_compensation", 0.0f);
    s.depth_fx[3] = jopt_f32(*flame, "far_density_fade", 0.0f);
    s.depth_fx[4] = jopt_f32(*flame, "far_density_fade_start", 0.0f);
  } else {
    s.render_mode = jopt_str(root, "render_mode", "2d");
    s.preserve_z = jopt_bool(root, "preserve_z", true);
    s.cam[6] = jopt_f32(root, "perspective_strength", 0.0f);
    s.depth_fx[2] = jopt_f32(root, "depth_density_compensation", 0.0f);
    s.depth_fx[3] = jopt_f32(root, "far_density_fade", 0.0f);
    s.depth_fx[4] = jopt_f32(root, "far_density_fade_start", 0.0f);
  }
  if (s.render_mode == "TwoD") s.render_mode = "2d";
  if (s.render_mode == "ThreeD") s.render_mode = "3d";
  if (s.render_mode != "2d" && s.render_mode != "3d" && s.render_mode != "escape" &&
      s.render_mode != "simulation")
    raise(FFM_ERR_PARSE, "render_mode: unknown variant `%s`", s.render_mode.c_str());

  // the flame (Flame's Deserialize, src/scene/transforms.rs:2567-2690)
  s.name = jopt_str(*flame, "name", "");
  const JVal* ts = flame->get("transforms");
  if (!ts) raise(FFM_ERR_PARSE, "flame: missing field `transforms`");
  if (ts->k != JVal::ARR) raise(FFM_ERR_PARSE, "flame: transforms must be a list");
  for (size_t i = 0; i < ts->a.size(); i++) {
    std::string w = "transform " + std::to_string(i);
    s.xforms.push_back(json_xform(ts->a[i], w.c_str()));
  }
  if (const JVal* fs = flame->get("final_transforms")) {
    if (fs->k != JVal::ARR) raise(FFM_ERR_PARSE, "flame: final_transforms must be a list");
    for (size_t i = 0; i < fs->a.size(); i++) {
      std::string w = "final transform " + std::to_string(i);
      s.finals.push_back(json_xform(fs->a[i], w.c_str()));
    }
  }
  // migrate_legacy_final (transforms.rs:2257-2266): the singular final goes
  // to the end of the pool, attached to every normal transform
  if (const JVal* lf = flame->get("final_transform"); lf && lf->k != JVal::NUL) {
    int64_t idx = (int64_t)s.finals.size();
    s.finals.push_back(json_xform(*lf, "final_transform"));
    for (auto& x : s.xforms) {
      bool has = false;
      for (auto k : x.final_att) has |= k == idx;
      if (!has) x.final_att.push_back(idx);
    }
  }
  if (const JVal* lt = flame->get("linked_transforms"); lt && lt->k == JVal::ARR)
    s.n_linked = lt->a.size();
  if (const JVal* xa = flame->get("xaos"); xa && xa->k == JVal::ARR) {
    for (auto& row : xa->a) {
      if (row.k != JVal::ARR) raise(FFM_ERR_PARSE, "flame: xaos must be a matrix");
      s.xaos.emplace_back();
      for (auto& w : row.a) s.xaos.back().push_back(jnum_f32(&w, "xaos"));
    }
  }
  if (const JVal* so = flame->get("solo_transform"); so && so->k != JVal::NUL) s.solo = true;
  s.subflames = nonempty_array(*flame, "subflames");
  s.custom_variations = nonempty_array(*flame, "custom_variations");
  if (const JVal* ps = flame->get("post_symmetry"); ps && ps->k == JVal::OBJ) {
    std::string ty = jopt_str(*ps, "ty", "none");
    if (ty != "none") s.post_symmetry = true;
  }

  // view and scene
  s.reset_orbits = jopt_bool(root, "reset_orbits", false);
  s.zoom = jopt_f32(root, "zoom", 1.0f);
  if (const JVal* v = root.get("pan_x")) s.pan_x = v->k == JVal::NUM ? v->n : (raise(FFM_ERR_PARSE, "pan_x: expected a number"), 0.0);
  if (const JVal* v = root.get("pan_y")) s.pan_y = v->k == JVal::NUM ? v->n : (raise(FFM_ERR_PARSE, "pan_y: expected a number"), 0.0);
  s.rotation = jopt_f32(root, "rotation", 0.0f);
  s.cam[0] = jopt_f32(root, "camera_rotation_x", 0.0f);
  s.cam[1] = jopt_f32(root, "camera_rotation_y", 0.0f);
  s.cam[2] = jopt_f32(root, "camera_bank", 0.0f);
  s.cam[3] = jopt_f32(root, "camera_x", 0.0f);
  s.cam[4] = jopt_f32(root, "camera_y", 0.0f);
  s.cam[5] = jopt_f32(root, "camera_z", 0.0f);
  s.cam[7] = jopt_f32(root, "dof_blur_strength", 0.0f);
  s.depth_fx[0] = jopt_f32(root, "fog_strength", 0.0f);
  s.depth_fx[1] = jopt_f32(root, "solid_strength", 0.0f);
  if (const JVal* is = root.get("image_size")) {
    if (is->k != JVal::ARR || is->a.size() != 2) raise(FFM_ERR_PARSE, "image_size: expected [w, h]");
    for (int q = 0; q < 2; q++)
      if (is->a[q].k != JVal::NUM || is->a[q].n < 0 || is->a[q].n > 4294967295.0)
        raise(FFM_ERR_PARSE, "image_size: expected two unsigned integers");
    s.image_w = (uint32_t)is->a[0].n;
    s.image_h = (uint32_t)is->a[1].n;
  }
  s.filter_radius = jopt_f32(root, "filter_radius", 0.0f);
  s.filter_blur_edges = jopt_f32(root, "filter_blur_edges", 0.0f);
  s.density_scale = jopt_f32(root, "density_scale", 1.0f);
  s.max_iterations = jopt_u64(root, "max_iterations", 1000000000ull);
  s.color_mode = jopt_str(root, "color_mode", "palette");
  if (const JVal* p = root.get("palette")) json_palette(*p, s);
  s.palette_rotation = jopt_f32(root, "palette_rotation", 0.0f);
  s.palette_size = (uint32_t)jopt_u64(root, "palette_size", 256);
  s.palette_squeeze = jopt_f32(root, "palette_squeeze", 1.0f);
  s.palette_squeeze_mode = jopt_str(root, "palette_squeeze_mode", "linear");
  s.palette_squeeze_falloff = jopt_f32(root, "palette_squeeze_falloff", 0.5f);
  s.palette_log_strength = jopt_f32(root, "palette_log_strength", 0.0f);
  s.palette_reverse = jopt_bool(root, "palette_reverse", false);
  if (const JVal* bg = root.get("background_color")) {
    if (bg->k != JVal::ARR || bg->a.size() != 3) raise(FFM_ERR_PARSE, "background_color: expected [r, g, b]");
    for (int c = 0; c < 3; c++) s.background[c] = jnum_f32(&bg->a[c], "background_color");
  }
  s.tonemap_mode = jopt_str(root, "tonemap_mode", "logarithmic");
  s.highlight_mode = jopt_str(root, "highlight_mode", "clip");
  s.use_curve = jopt_bool(root, "use_curve", true);
  json_curve(root.get("tonemap_curve"), s);
  s.exposure = jopt_f32(root, "exposure", 1.0f);
  s.gamma = jopt_f32(root, "gamma", 4.0f);
  s.gamma_threshold = jopt_f32(root, "gamma_threshold", 5.0f);
  s.brightness = jopt_f32(root, "brightness", 4.0f);
  s.vibrancy = jopt_f32(root, "vibrancy", 1.0f);
  s.white_level = jopt_f32(root, "white_level", 200.0f);
  s.saturation = jopt_f32(root, "saturation", 1.0f);
  s.hue_shift = jopt_f32(root, "hue_shift", 0.0f);
  s.alpha_blend_low = jopt_f32(root, "alpha_blend_low", 0.3f);
  s.alpha_blend_high = jopt_f32(root, "alpha_blend_high", 0.8f);
  s.levels_enabled = jopt_bool(root, "levels_enabled", true);
  s.levels_low = jopt_f32(root, "levels_low", 0.0f);
  s.levels_high = jopt_f32(root, "levels_high", 10.0f);
  // v3 -> v4: a file saved at v3 omitted levels_gamma at 1.0
  s.levels_gamma = jopt_f32(root, "levels_gamma", version == 3 ? 1.0f : 0.5f);
  s.deterministic_rng = jopt_bool(root, "deterministic_rng", false);

  // features outside this path, present in the file
  s.effects.pixel_scale = 1.0f;
  json_effects(root, "density_effects", 1, s);
  json_effects(root, "color_effects", 2, s);
  if (nonempty_array(root, "custom_effects")) s.unsupported.push_back("custom effects");
  if (nonempty_array(root, "word_removals")) s.unsupported.push_back("word removals");
  if (jopt_bool(root, "cylinder_targeting", false) || jopt_bool(root, "cylinder_always", false))
    s.unsupported.push_back("cylinder targeting");
  if (jopt_bool(root, "auto_exposure", false)) s.unsupported.push_back("auto exposure");
  if (const JVal* imp = root.get("importance"); imp && imp->k == JVal::OBJ &&
                                                jopt_bool(*imp, "enabled", false))
    s.unsupported.push_back("importance sampling");
  return s;
}

// -------------------------------------------------------- XML -> SceneIn

// xform / finalxform attributes that are not variations: read above or
// carried by Apophysis / JWildfire for features outside this path. The
// reference ignores every numeric attribute that names no variation; this
// list lets the importer refuse, rather than ignore, a numeric attribute it
// does not know (most likely a variation this path lacks).
const char* const kXformIgnored[] = {
    "animate", "var_color", "material", "material_speed", "mod_gamma", "mod_gamma_speed",
    "mod_contrast", "mod_contrast_speed", "mod_saturation", "mod_saturation_speed", "mod_hue",
    "mod_hue_speed", "weight", "opacity", "enabled", "plotmode", "name", "chaos",
};

bool xform_ignored(const std::string& k) {
  for (const char* s : kXformIgnored)
    if (k == s) return true;
  return false;
}

void xml_coefs(const std::string& v, float* a, float* b, float* c, float* d, float* e, float* f) {
  // "a c b d e f": Apophysis writes the matrix column by column
  auto p = split_ws(v);
  if (p.size() < 6) return;
  float t;
  *a = parse_f32(p[0], &t) ? t : 1.0f;
  *c = parse_f32(p[1], &t) ? t : 0.0f;
  *b = parse_f32(p[2], &t) ? t : 0.0f;
  *d = parse_f32(p[3], &t) ? t : 1.0f;
  *e = parse_f32(p[4], &t) ? t : 0.0f;
  *f = parse_f32(p[5], &t) ? t : 0.0f;
}

// parse_xform_element / parse_finalxform_element (src/flame_xml.rs:600-900).
// Transform::new's defaults: weight 1, color 0.5, color_speed 0, opacity 1.
XfIn xml_xform(const XEvent& el, bool final_xf, std::vector<float>* chaos,
               std::vector<std::string>* unknown) {
  XfIn x;
  std::vector<std::pair<std::string, float>> pending;  // params, applied after the loop
  for (const XAttr& at : el.attrs) {
    const std::string& k = at.key;
    const std::string& v = at.value;
    float t;
    if (!final_xf && k == "weight") {
      x.weight = parse_f32(v, &t) ? t : 1.0f;
    } else if (k == "color") {
      auto p = split_ws(v);
      if (!p.empty() && parse_f32(p[0], &t)) x.color = t;
    } else if (k == "color_speed" || k == "symmetry") {
      x.color_speed = parse_f32(v, &t) ? t : 0.0f;
    } else if (k == "coefs") {
      xml_coefs(v, &x.a, &x.b, &x.c, &x.d, &x.e, &x.f);
    } else if (!final_xf && k == "opacity") {
      x.opacity = parse_f32(v, &t) ? t : 1.0f;
    } else if (k == "pluginColor" || k == "plugin_color") {
      // direct colour: read by direct-colour variations only, none here
    } else if (k == "post") {
      if (split_ws(v).size() >= 6) {
        x.post = true;
        xml_coefs(v, &x.pa, &x.pb, &x.pc, &x.pd, &x.pe, &x.pf);
      }
    } else if (k == "yzCoefs" || k == "zxCoefs" || k == "yzPost" || k == "zxPost") {
      auto p = split_ws(v);
      static const float ident[6] = {1, 0, 0, 1, 0, 0};
      if (p.size() >= 6)
        for (int q = 0; q < 6; q++)
          if (parse_f32(p[q], &t) && t != ident[q]) x.plane_identity = false;
    } else if (!final_xf && k == "chaos") {
      if (chaos) {
        for (auto& s : split_ws(v))
          if (parse_f32(s, &t)) chaos->push_back(t);
      }
    } else if (k.size() > 12 && k.compare(k.size() - 12, 12, "_fx_priority") == 0) {
      std::string var = k.substr(0, k.size() - 12);
      if (var_by_name(var) && parse_f32(v, &t)) x.priorities.emplace_back(var, (int)std::lround(t));
    } else if (k.rfind("subflame_wf_", 0) == 0) {
      // JWildfire subflame payloads belong to a variation this path lacks
    } else if (parse_f32(v, &t)) {
      if (const VarInfo* vi = var_by_name(k)) {
        if (t != 0.0f) {
          if (float* w = x.var(vi->name)) *w += t;
          else {
            x.order.push_back(vi->name);
            x.vars.emplace_back(vi->name, t);
          }
        }
        continue;
      }
      // "<variation>_<param>"
      bool param = false;
      size_t us = k.find('_');
      while (us != std::string::npos) {
        const VarInfo* vi = var_by_name(k.substr(0, us));
        std::string pn = k.substr(us + 1);
        if (vi)
          for (int q = 0; q < vi->n_params; q++)
            if (pn == vi->params[q]) {
              pending.emplace_back(std::string(vi->name) + "." + pn, t);
              param = true;
            }
        us = k.find('_', us + 1);
      }
      if (!param && !xform_ignored(k) && !plane_noop(k) && t != 0.0f && unknown)
        unknown->push_back(k);
    }
  }
  for (auto& p : pending) x.set_param(p.first, p.second);
  return x;
}

// parse_flame_element (src/flame_xml.rs:70-560)
SceneIn xml_flame(XmlReader& rd, const XEvent& start) {
  SceneIn s;
  // the config parse_flame_element builds (flame_xml.rs:434-560)
  s.render_mode = "3d";
  s.preserve_z = false;
  s.gamma = 2.2f;
  s.brightness = 1.0f;
  s.vibrancy = 1.0f;
  float gamma_threshold = 0.0025f;
  float scale = 100.0f, rotate = 0.0f, cam_zoom = 1.0f;
  float cam_zpos = 0.0f;
  bool cam_pos_z_set = false;
  bool sld_enabled = false;
  s.max_iterations = 1000000000ull;
  s.deterministic_rng = false;
  s.density_scale = 1.0f;
  s.filter_blur_edges = 0.0f;
  for (const XAttr& at : start.attrs) {
    const std::string& k = at.key;
    const std::string& v = at.value;
    float t;
    auto f = [&](float dflt) { return parse_f32(v, &t) ? t : dflt; };
    if (k == "name") s.name = v;
    else if (k == "size") {
      auto p = split_ws(v);
      if (p.size() == 2) {
        uint64_t w, h;
        s.image_w = parse_uint(p[0], 0xFFFFFFFFu, &w) ? (uint32_t)w : 1920;
        s.image_h = parse_uint(p[1], 0xFFFFFFFFu, &h) ? (uint32_t)h : 1080;
      }
    } else if (k == "center") {
      auto p = split_ws(v);
      if (p.size() == 2) {  // the centre is read as f64, as the config holds pan
        double dv;
        s.pan_x = parse_f64(p[0], &dv) ? dv : 0.0;
        s.pan_y = parse_f64(p[1], &dv) ? dv : 0.0;
      }
    } else if (k == "scale") scale = f(100.0f);
    else if (k == "rotate") rotate = f(0.0f);
    else if (k == "background") {
      auto p = split_ws(v);
      if (p.size() == 3)
        for (int c = 0; c < 3; c++) s.background[c] = (parse_f32(p[c], &t) ? t : 0.0f) / 255.0f;
    } else if (k == "brightness") s.brightness = f(4.0f);
    else if (k == "gamma") s.gamma = f(4.0f);
    else if (k == "vibrancy") s.vibrancy = f(1.0f);
    else if (k == "gamma_threshold") gamma_threshold = f(0.0025f);
    else if (k == "filter") s.filter_radius = f(0.0f);
    else if (k == "cam_pitch") s.cam[0] = f(0.0f);
    else if (k == "cam_yaw") s.cam[1] = f(0.0f);
    else if (k == "cam_roll") s.cam[2] = f(0.0f);
    else if (k == "cam_zpos") cam_zpos = f(0.0f);
    else if (k == "cam_pos_x") s.cam[3] = f(0.0f);
    else if (k == "cam_pos_y") s.cam[4] = f(0.0f);
    else if (k == "cam_pos_z") {
      if (parse_f32(v, &t)) {
        s.cam[5] = t;
        cam_pos_z_set = true;
      }
    } else if (k == "cam_perspective" || k == "cam_persp") s.cam[6] = std::fabs(f(0.0f));
    else if (k == "cam_dof") s.cam[7] = f(0.0f);
    else if (k == "cam_zoom") cam_zoom = f(1.0f);
    else if (k == "saturation") s.saturation = f(1.0f);
    else if (k == "white_level") s.white_level = f(200.0f);
    else if (k == "post_symmetry_type") {
      std::string u = v;
      for (auto& ch : u) ch = (char)toupper((unsigned char)ch);
      if (u == "X_AXIS" || u == "Y_AXIS" || u == "POINT") s.post_symmetry = true;
    } else if (k == "preserve_z") {
      std::string tv = v;
      while (!tv.empty() && isspace((unsigned char)tv.back())) tv.pop_back();
      while (!tv.empty() && isspace((unsigned char)tv.front())) tv.erase(tv.begin());
      s.preserve_z = !tv.empty() && tv != "0";
    } else if (k == "curves") {
      // 48 numbers: 4 rational Bezier curves of 4 (x, y, w) points; the
      // reference samples the first at t = 1/4, 1/2, 3/4 into a 5-point
      // tone curve (parse_apophysis_curves, flame_xml.rs:1084-1160)
      std::vector<float> cv;
      for (auto& q : split_ws(v))
        if (parse_f32(q, &t)) cv.push_back(t);
      if (cv.size() == 48) {
        std::vector<std::pair<float, float>> pts{{0.0f, 0.0f}};
        for (float tt : {0.25f, 0.5f, 0.75f}) {
          float sv = 1.0f - tt, s2 = sv * sv, s3 = s2 * sv, t2 = tt * tt, t3 = t2 * tt;
          float b0 = cv[2] * s3, b1 = cv[5] * 3.0f * s2 * tt, b2 = cv[8] * 3.0f * sv * t2,
                b3 = cv[11] * t3;
          float nx = b0 * cv[0] + b1 * cv[3] + b2 * cv[6] + b3 * cv[9];
          float ny = b0 * cv[1] + b1 * cv[4] + b2 * cv[7] + b3 * cv[10];
          float den = b0 + b1 + b2 + b3;
          if (std::fabs(den) < 1e-10f) continue;
          // CurvePoint::new clamps both coordinates to [0, 1]
          // (src/scene/tonemap.rs:134-139)
          pts.push_back({std::min(std::max(nx / den, 0.0f), 1.0f),
                         std::min(std::max(ny / den, 0.0f), 1.0f)});
        }
        pts.push_back({1.0f, 1.0f});
        std::stable_sort(pts.begin(), pts.end(),
                         [](const std::pair<float, float>& a, const std::pair<float, float>& b) {
                           return a.first < b.first;
                         });
        s.curve_x.clear();
        s.curve_y.clear();
        for (auto& q : pts) {
          s.curve_x.push_back(q.first);
          s.curve_y.push_back(q.second);
        }
      }
    } else if (k == "soloxform") {
      uint64_t u;
      if (parse_uint(v, UINT64_MAX / 10, &u)) s.solo = true;
    } else if (k == "sld_render_enabled") {
      sld_enabled = v == "1";
    }
  }
  if (!cam_pos_z_set) s.cam[5] = cam_zpos;
  if (sld_enabled) s.depth_fx[1] = 1.0f;
  s.zoom = (scale / 200.0f) * cam_zoom;
  s.rotation = rotate * 3.14159265358979323846f / 180.0f;
  s.gamma_threshold = gamma_threshold * 2000.0f + 50.0f;
  s.use_curve = true;

  std::vector<std::vector<float>> chaos_rows;
  std::vector<bool> chaos_has;
  std::vector<std::string> unknown;
  size_t layers_seen = 0;
  bool skip_layer = false;
  const std::string element = start.name;
  for (;;) {
    XEvent e = rd.next();
    if (e.k == XEvent::START || e.k == XEvent::EMPTY) {
      if (e.name == "layer") {
        skip_layer = layers_seen > 0;
        layers_seen++;
      } else if (e.name == "xform" && !skip_layer) {
        std::vector<float> ch;
        bool had = false;
        for (auto& a : e.attrs) had |= a.key == "chaos";
        s.xforms.push_back(xml_xform(e, false, &ch, &unknown));
        chaos_has.push_back(had && !ch.empty());
        chaos_rows.push_back(ch);
      } else if (e.name == "finalxform" && !skip_layer) {
        s.finals.push_back(xml_xform(e, true, nullptr, &unknown));
      } else if (e.name == "palette" && !skip_layer) {
        // parse_palette_element (flame_xml.rs:1014-1067): 256 RRGGBB
        std::string hex;
        if (e.k == XEvent::START) {
          for (;;) {
            XEvent t = rd.next();
            if (t.k == XEvent::TEXT) hex += t.name;
            else if (t.k == XEvent::END && t.name == "palette") break;
            else if (t.k == XEvent::END_OF_FILE) break;
          }
        }
        std::string h;
        for (char ch : hex)
          if (!isspace((unsigned char)ch)) h += ch;
        if (h.size() < 1536)
          raise(FFM_ERR_PARSE, "Invalid palette data: expected 1536 hex characters, got %zu", h.size());
        s.stops.clear();
        for (int i = 0; i < 256; i++) {
          std::array<float, 4> st;
          st[0] = (float)i / 255.0f;
          for (int c = 0; c < 3; c++) {
            char two[3] = {h[6 * i + 2 * c], h[6 * i + 2 * c + 1], 0};
            if (!isxdigit((unsigned char)two[0]) || !isxdigit((unsigned char)two[1]))
              raise(FFM_ERR_PARSE, "Invalid palette data: '%s' is not hex", two);
            st[1 + c] = (float)strtol(two, nullptr, 16) / 255.0f;
          }
          s.stops.push_back(st);
        }
        s.palette_set = true;
      }
    } else if (e.k == XEvent::END) {
      if (e.name == element) break;
      if (e.name == "layer") skip_layer = false;
    } else if (e.k == XEvent::END_OF_FILE) {
      break;
    }
  }
  if (!s.palette_set) {
    // Palette::default() -- the grayscale ramp
    s.stops = {{0.0f, 0.0f, 0.0f, 0.0f}, {1.0f, 1.0f, 1.0f, 1.0f}};
    s.palette_set = true;
  }
  // xaos from the chaos rows (flame_xml.rs:360-380): 1.0 where unset
  size_t n = s.xforms.size();
  bool any_chaos = false;
  for (size_t src = 0; src < n; src++) any_chaos |= chaos_has[src];
  if (any_chaos) {
    s.xaos.assign(n, std::vector<float>(n, 1.0f));
    for (size_t src = 0; src < n; src++)
      if (chaos_has[src])
        for (size_t dst = 0; dst < chaos_rows[src].size() && dst < n; dst++)
          s.xaos[src][dst] = chaos_rows[src][dst];
  }
  // every <finalxform>, in file order, attached to every normal transform
  for (size_t k = 0; k < s.finals.size(); k++)
    for (auto& x : s.xforms) x.final_att.push_back((int64_t)k);
  for (auto& u : unknown) s.unsupported.push_back("attribute `" + u + "` (a variation outside this path)");
  return s;
}

// ------------------------------------------------------ documents -> list

std::vector<SceneIn> read_all(const char* text, size_t len) {
  if (!text) raise(FFM_ERR_INVALID_ARGUMENT, "no config text");
  size_t i = 0;
  if (len >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB &&
      (unsigned char)text[2] == 0xBF)
    i = 3;  // UTF-8 byte-order mark
  while (i < len && isspace((unsigned char)text[i])) i++;
  std::vector<SceneIn> out;
  if (i < len && text[i] == '<') {
    XmlReader rd(text, len);
    for (;;) {
      XEvent e = rd.next();
      if (e.k == XEvent::END_OF_FILE) break;
      if (e.k == XEvent::START && (e.name == "flame" || e.name == "jwf-flame"))
        out.push_back(xml_flame(rd, e));
    }
    if (out.empty()) raise(FFM_ERR_PARSE, "No <flame> or <jwf-flame> elements found in XML");
    return out;
  }
  JsonReader jr(text, len);
  JVal doc = jr.document();
  if (doc.k == JVal::ARR) {
    for (size_t k = 0; k < doc.a.size(); k++) {
      try {
        out.push_back(json_config(doc.a[k]));
      } catch (Fail& f) {
        raise(f.code, "Error in config[%zu]: %s", k, f.msg.c_str());
      }
    }
  } else if (doc.k == JVal::OBJ) {
    out.push_back(json_config(doc));
  } else {
    raise(FFM_ERR_PARSE, "Expected JSON object or array");
  }
  return out;
}

// ----------------------------------------------------------- SceneIn -> ABI

int mode_index(const std::string& v, const char* const* names, int n) {
  for (int i = 0; i < n; i++)
    if (v == names[i]) return i;
  return -1;
}

void build_scene(const SceneIn& s, ffm_scene* out) {
  auto unsupported = [](const std::string& what) {
    raise(FFM_ERR_UNSUPPORTED, "not rendered by this engine: %s", what.c_str());
  };
  if (!s.unsupported.empty()) unsupported(s.unsupported.front());
  if (s.render_mode == "escape" || s.render_mode == "simulation")
    unsupported("render mode `" + s.render_mode + "`");
  bool is3d = s.render_mode == "3d";
  if (is3d) {
    static const char* cam_names[] = {"camera pitch", "camera yaw", "camera bank", "camera x",
                                      "camera y", "camera z", "perspective", "depth of field"};
    for (int q = 0; q < 8; q++)
      if (s.cam[q] != 0.0f) unsupported(std::string("3D ") + cam_names[q]);
    static const char* fx_names[] = {"depth fog", "solid rendering", "depth density compensation",
                                     "far density fade", "far density fade"};
    for (int q = 0; q < 5; q++)
      if (s.depth_fx[q] != 0.0f) unsupported(fx_names[q]);
    if (s.preserve_z) unsupported("preserve_z (z carried across iterations)");
  }
  if (s.color_mode != "palette" && s.color_mode != "Palette") unsupported("color mode `" + s.color_mode + "`");
  if (s.solo) unsupported("solo transform");
  if (s.subflames) unsupported("subflames");
  if (s.post_symmetry) unsupported("post symmetry");
  if (s.custom_variations) unsupported("user-written variations");
  if (s.curve_x.size() > FFM_MAX_CURVE_POINTS)
    raise(FFM_ERR_TOO_LARGE, "tone curve of %zu points: at most %d fit", s.curve_x.size(),
          FFM_MAX_CURVE_POINTS);
  for (size_t i = 0; i < s.curve_x.size(); i++)
    if (!std::isfinite(s.curve_x[i]) || !std::isfinite(s.curve_y[i]) ||
        (i && s.curve_x[i] < s.curve_x[i - 1]))
      unsupported("tone curve whose points are not finite and sorted by x");
  if (s.xforms.empty()) raise(FFM_ERR_NO_FLAME, "the flame has no transforms");
  if ((int)s.xforms.size() > FFM_MAX_XFORMS)
    raise(FFM_ERR_TOO_LARGE, "%zu transforms: at most %d fit", s.xforms.size(), FFM_MAX_XFORMS);
  if ((int)s.finals.size() > FFM_MAX_FINALS)
    raise(FFM_ERR_TOO_LARGE, "%zu final transforms: at most %d fit", s.finals.size(), FFM_MAX_FINALS);

  memset(out, 0, sizeof *out);
  ffm_flame_desc& fd = out->flame;

  // per-transform variation order (Transform::ordered_variation_names,
  // transforms.rs:668-682): variation_order, then registry order; the
  // flame's local indices dedupe over transforms, then finals
  // (active_variation_names_ordered, transforms.rs:2297-2319)
  std::vector<int> ids;
  auto absorb = [&](const XfIn& x, const char* what, size_t idx) {
    for (auto& v : x.vars)
      if (!var_by_name(v.first) && !plane_noop(v.first) && v.second != 0.0f)
        unsupported("variation `" + v.first + "` (" + what + " " + std::to_string(idx) + ")");
    // an fx priority moves an Any-phase variation to the pre or post phase
    // (resolve_phase_buckets, src/shader_builder_v2.rs:2406-2470); pre_blur
    // is pre-phase by definition and keeps its place
    for (auto& p : x.priorities)
      if (p.second != 0 && x.has_var(p.first) && var_by_name(p.first) && p.first != "pre_blur")
        unsupported("variation priority on `" + p.first + "`");
    if (!x.plane_identity) unsupported("YZ / ZX plane affines");
    if (x.n_linked_att) unsupported("linked transforms");
    std::vector<int> mine;
    for (auto& n : x.order)
      if (x.has_var(n))
        if (const VarInfo* vi = var_by_name(n))
          if (std::find(mine.begin(), mine.end(), vi->id) == mine.end()) mine.push_back(vi->id);
    for (const auto& vi : kVars)
      if (x.has_var(vi.name) && std::find(mine.begin(), mine.end(), vi.id) == mine.end())
        mine.push_back(vi.id);
    for (int id : mine)
      if (std::find(ids.begin(), ids.end(), id) == ids.end()) ids.push_back(id);
  };
  for (size_t i = 0; i < s.xforms.size(); i++) absorb(s.xforms[i], "transform", i);
  for (size_t i = 0; i < s.finals.size(); i++) absorb(s.finals[i], "final transform", i);
  if ((int)ids.size() > FFM_MAX_VARS)
    raise(FFM_ERR_TOO_LARGE, "%zu active variations: at most %d", ids.size(), FFM_MAX_VARS);
  fd.n_xforms = (int32_t)s.xforms.size();
  fd.n_finals = (int32_t)s.finals.size();
  fd.n_vars = (int32_t)ids.size();
  for (size_t k = 0; k < ids.size(); k++) fd.var_id[k] = ids[k];

  auto fill = [&](ffm_xform& d, const XfIn& x) {
    d.a = x.a, d.b = x.b, d.c = x.c, d.d = x.d, d.e = x.e, d.f = x.f;
    d.post_enabled = x.post ? 1 : 0;
    d.post_a = x.post ? x.pa : 1.0f;
    d.post_b = x.post ? x.pb : 0.0f;
    d.post_c = x.post ? x.pc : 0.0f;
    d.post_d = x.post ? x.pd : 1.0f;
    d.post_e = x.post ? x.pe : 0.0f;
    d.post_f = x.post ? x.pf : 0.0f;
    d.weight = x.weight;
    d.color = x.color;
    d.color_speed = x.color_speed;
    d.opacity = x.opacity;
    for (size_t k = 0; k < ids.size(); k++) {
      const VarInfo* vi = var_by_id(ids[k]);
      for (auto& v : x.vars)
        if (v.first == vi->name) d.var_weight[k] = v.second;
      for (int q = 0; q < vi->n_params; q++) {
        float val = vi->defaults[q];
        std::string key = std::string(vi->name) + "." + vi->params[q];
        for (auto& p : x.params)
          if (p.first == key) val = p.second;
        d.var_param[k][q] = val;
      }
    }
  };
  // Flame::has_xaos (src/scene/xaos.rs:29-41): some weight differs from 1
  // by more than 1e-6; get_xaos reads 1 past a short row
  bool has_xaos = false;
  for (auto& row : s.xaos)
    for (float w : row) has_xaos |= std::fabs(w - 1.0f) > 1e-6f;
  if (has_xaos) {
    const size_t n = s.xforms.size();
    if ((int)n > FFM_MAX_XAOS)
      raise(FFM_ERR_TOO_LARGE, "xaos on %zu transforms: at most %d", n, FFM_MAX_XAOS);
    fd.has_xaos = 1;
    for (size_t i = 0; i < n; i++)
      for (size_t j = 0; j < n; j++)
        fd.xaos[i][j] = i < s.xaos.size() && j < s.xaos[i].size() ? s.xaos[i][j] : 1.0f;
  }
  for (size_t i = 0; i < s.xforms.size(); i++) fill(fd.xforms[i], s.xforms[i]);
  for (size_t i = 0; i < s.finals.size(); i++) fill(fd.finals[i], s.finals[i]);
  if (!s.finals.empty()) {
    fd.per_transform_finals = 1;
    for (size_t i = 0; i < s.xforms.size(); i++) {
      const auto& att = s.xforms[i].final_att;
      if ((int)att.size() > FFM_MAX_FINALS)
        raise(FFM_ERR_TOO_LARGE, "transform %zu attaches %zu finals", i, att.size());
      for (size_t j = 0; j < att.size(); j++) {
        if (att[j] < 0 || att[j] >= (int64_t)s.finals.size())
          raise(FFM_ERR_PARSE, "transform %zu attaches final %lld of %zu", i, (long long)att[j],
                s.finals.size());
        fd.xforms[i].final_attach[j] = (int32_t)att[j];
      }
      fd.xforms[i].n_final_attach = (int32_t)att.size();
    }
  }

  // palette
  std::vector<std::array<float, 4>> stops = s.stops;
  if (!s.palette_set) {
    // default_palette: Palette::fire (src/scene/palette.rs:148-175)
    stops = {{0.0f, 0.0f, 0.0f, 0.0f},
             {0.25f, 0.5f, 0.0f, 0.0f},
             {0.5f, 1.0f, 0.3f, 0.0f},
             {0.75f, 1.0f, 0.8f, 0.0f},
             {1.0f, 1.0f, 1.0f, 0.8f}};
  }
  if ((int)stops.size() > FFM_MAX_STOPS)
    raise(FFM_ERR_TOO_LARGE, "%zu palette stops: at most %d", stops.size(), FFM_MAX_STOPS);
  out->n_stops = (int32_t)stops.size();
  for (size_t i = 0; i < stops.size(); i++)
    for (int c = 0; c < 4; c++) out->stops[i][c] = stops[i][c];
  if (s.palette_size < FFM_MIN_PALETTE || s.palette_size > FFM_MAX_PALETTE)
    raise(FFM_ERR_INVALID_ARGUMENT, "palette_size %u outside %d..%d", s.palette_size,
          FFM_MIN_PALETTE, FFM_MAX_PALETTE);
  out->palette_size = (int32_t)s.palette_size;
  std::string sq = s.palette_squeeze_mode;
  if (sq == "linear" || sq == "Linear") out->palette_xf.squeeze_mode = 0;
  else if (sq == "geometric" || sq == "Geometric") out->palette_xf.squeeze_mode = 1;
  else raise(FFM_ERR_PARSE, "palette_squeeze_mode: unknown variant `%s`", sq.c_str());
  // PaletteTransform::from_settings (palette.rs:560-573): the factor the mode reads
  out->palette_xf.squeeze_factor =
      out->palette_xf.squeeze_mode == 0 ? s.palette_squeeze : s.palette_squeeze_falloff;
  out->palette_xf.log_strength = s.palette_log_strength;
  out->palette_xf.rotation = s.palette_rotation;
  out->palette_xf.reverse = s.palette_reverse ? 1 : 0;

  out->zoom = s.zoom;
  out->rotation = s.rotation;
  out->pan_x = s.pan_x;
  out->pan_y = s.pan_y;
  out->filter_radius = s.filter_radius;
  out->filter_blur_edges = s.filter_blur_edges;
  out->max_iterations = s.max_iterations;
  out->reset_orbits = s.reset_orbits ? 1 : 0;
  out->deterministic_rng = s.deterministic_rng ? 1 : 0;
  out->render_3d = is3d ? 1 : 0;
  out->image_width = s.image_w;
  out->image_height = s.image_h;

  ffm_tonemap_params& tp = out->tonemap;
  static const char* tm[] = {"linear", "logarithmic", "density"};
  static const char* tm_alias[] = {"Linear", "Logarithmic", "DensityVisualization"};
  int t = mode_index(s.tonemap_mode, tm, 3);
  if (t < 0) t = mode_index(s.tonemap_mode, tm_alias, 3);
  if (t < 0) raise(FFM_ERR_PARSE, "tonemap_mode: unknown variant `%s`", s.tonemap_mode.c_str());
  tp.tonemap_mode = t;
  static const char* hm[] = {"clip", "max_norm", "reinhard", "filmic", "agx", "agx_punchy"};
  static const char* hm_alias[] = {"Clip", "MaxNorm", "Reinhard", "Filmic", "AgX", "AgXPunchy"};
  int h = mode_index(s.highlight_mode, hm, 6);
  if (h < 0) h = mode_index(s.highlight_mode, hm_alias, 6);
  if (h < 0) raise(FFM_ERR_PARSE, "highlight_mode: unknown variant `%s`", s.highlight_mode.c_str());
  if (h >= 4) unsupported("highlight mode `" + s.highlight_mode + "`");
  tp.highlight_mode = h;
  tp.use_curve = s.use_curve ? 1 : 0;
  tp.curve_n = (int32_t)s.curve_x.size();
  for (size_t i = 0; i < s.curve_x.size(); i++) {
    tp.curve_x[i] = s.curve_x[i];
    tp.curve_y[i] = s.curve_y[i];
  }
  tp.exposure = s.exposure;
  tp.gamma = s.gamma;
  tp.gamma_threshold = s.gamma_threshold;
  tp.brightness = s.brightness;
  tp.vibrancy = s.vibrancy;
  tp.white_level = s.white_level;
  tp.density_scale = s.density_scale;
  tp.saturation = s.saturation;
  tp.hue_shift = s.hue_shift;
  tp.alpha_blend_low = s.alpha_blend_low;
  tp.alpha_blend_high = s.alpha_blend_high;
  // effective_levels_enabled: a flame mode keeps its Levels
  tp.levels_enabled = s.levels_enabled ? 1 : 0;
  tp.levels_low = s.levels_low;
  tp.levels_high = s.levels_high;
  tp.levels_gamma = s.levels_gamma;
  for (int c = 0; c < 3; c++) tp.background[c] = s.background[c];
  snprintf(out->name, sizeof out->name, "%s", s.name.c_str());
  out->effects = s.effects;
  out->effects.pixel_scale = 1.0f;
}

}  // namespace

extern "C" {

int ffm_scene_count(const char* text, size_t len, int* count, char* err, size_t err_cap) {
  if (!count) return FFM_ERR_INVALID_ARGUMENT;
  try {
    *count = (int)read_all(text, len).size();
    return FFM_OK;
  } catch (Fail& f) {
    put_err(err, err_cap, f.msg);
    return f.code;
  } catch (std::bad_alloc&) {
    put_err(err, err_cap, "out of host memory");
    return FFM_ERR_OUT_OF_MEMORY;
  }
}

int ffm_scene_parse(const char* text, size_t len, int index, ffm_scene* out, char* err,
                    size_t err_cap) {
  if (!out) return FFM_ERR_INVALID_ARGUMENT;
  try {
    std::vector<SceneIn> all = read_all(text, len);
    if (index < 0 || (size_t)index >= all.size())
      raise(FFM_ERR_INVALID_ARGUMENT, "config %d of %zu", index, all.size());
    build_scene(all[(size_t)index], out);
    return FFM_OK;
  } catch (Fail& f) {
    put_err(err, err_cap, f.msg);
    return f.code;
  } catch (std::bad_alloc&) {
    put_err(err, err_cap, "out of host memory");
    return FFM_ERR_OUT_OF_MEMORY;
  }
}

int ffm_scene_job(const ffm_scene* s, int width, int height, uint64_t iterations, uint32_t ipt,
                  uint32_t flags, ffm_job* j, char* err, size_t err_cap) {
  if (!s || !j) return FFM_ERR_INVALID_ARGUMENT;
  if (width <= 0 || height <= 0) {
    put_err(err, err_cap, "width and height must be positive");
    return FFM_ERR_INVALID_ARGUMENT;
  }
  if (s->render_3d && !(flags & FFM_RENDER_3D_AS_2D)) {
    put_err(err, err_cap,
            "the config's render mode is 3d: pass FFM_RENDER_3D_AS_2D to render its flat "
            "camera through the 2D chaos game");
    return FFM_ERR_UNSUPPORTED;
  }
  // 2x antialiasing renders twice the size with 4x the samples and twice
  // the filter radius (scale_config_for_supersample,
  // src/export/supersample.rs:54-58); the caller boxes the picture down
  const bool ss = (flags & FFM_RENDER_SUPERSAMPLE_2X) != 0;
  if (ss && (width > 16384 || height > 16384)) {
    put_err(err, err_cap, "a 2x supersampled render of " + std::to_string(width) + "x" +
                              std::to_string(height) + " exceeds 32768 pixels a side");
    return FFM_ERR_TOO_LARGE;
  }
  memset(j, 0, sizeof *j);
  j->view.width = ss ? 2 * width : width;
  j->view.height = ss ? 2 * height : height;
  j->view.zoom = s->zoom;
  j->view.pan_x = (float)s->pan_x;
  j->view.pan_y = (float)s->pan_y;
  j->view.rotation = s->rotation;
  // RenderJob::new (src/renderer/render.rs:111-138): burn-in 20; ipt is
  // pyfae's default 1024; the dispatch is 128 workgroups of 64 walkers. With
  // FFM_RENDER_FILL_GPU a dispatch walks enough walkers to fill the GPU:
  // about half the budget in one dispatch, 8192 to 2^20, a multiple of 8192.
  j->ipt = ipt ? ipt : 1024;
  j->burn_in = 20;
  uint64_t target = iterations ? iterations : s->max_iterations;
  if (ss) target = target > UINT64_MAX / 4 ? UINT64_MAX : target * 4;
  j->persistent_orbits = s->reset_orbits ? 0 : 1;
  if (!(flags & FFM_RENDER_FILL_GPU)) {
    j->walkers = 8192;
  } else {
    // at least two dispatches; as few walkers (a multiple of 8192) as plot
    // the budget in that many: a walker plots ipt - burn_in samples in its
    // first dispatch and, with persistent orbits, ipt in every later one
    const uint64_t first = j->ipt > j->burn_in ? j->ipt - j->burn_in : 1;
    const uint64_t later = j->persistent_orbits ? j->ipt : first;
    const uint64_t most = 1u << 20;
    uint64_t n = 2;
    if (target / most > first + later) n = 1 + (target / most - first + later - 1) / later;
    const uint64_t each = first + (n - 1) * later;
    const uint64_t w = (target / each + 8191) / 8192 * 8192;
    j->walkers = std::min<uint64_t>(std::max<uint64_t>(w, 8192), most);
  }
  // the deterministic seed schedule 12345 + k * 2654435761
  // (src/renderer/compute_kernel.rs:3541-3554); a config without
  // deterministic_rng starts it from a random base
  j->seed_base = 12345u;
  if (!s->deterministic_rng) j->seed_base = std::random_device{}();
  j->seed_stride = 2654435761u;
  j->target_samples = target;
  j->batch_size = 4;
  j->splat_mode = FFM_SPLAT_U32_LDS;
  j->filter_radius = ss ? 2.0f * s->filter_radius : s->filter_radius;
  j->filter_blur_edges = s->filter_blur_edges;
  j->tonemap = s->tonemap;
  // transparent_mode (src/renderer/compute_kernel.rs:3771-3784)
  j->tonemap.transparent_mode = !(flags & FFM_RENDER_TRANSPARENT)  ? 0
                                : (flags & FFM_RENDER_PREMULTIPLIED) ? 2
                                                                     : 1;
  return FFM_OK;
}

}  // extern "C"
