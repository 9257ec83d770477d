// scene_file.cpp — the reference's scene-file format for the in-scope subset:
// SceneEnv::Load + CreateEntList (proj/scene_env.h:357-1990, proj/scene.cpp:38)
// producing a prt_scene_desc, with the expression evaluator of every numeric
// field (proj/Expr.h, proj/Expr.cpp) and the OBJ reader (proj/ObjLoader.cpp).
//
// Reading rules kept from the reference:
//   - a line starting with '#' is skipped, but only one: getLine reads the
//     next line and returns it whatever it holds (scene_env.h:141-149);
//     keyword lines that match nothing are ignored;
//   - keywords are whole lines compared with their '\n' (strcmp); lines are
//     read as text, so "\r\n" counts as "\n" (the reference is read in
//     Windows text mode);
//   - numeric fields go through expr_eval (atof/atoi are macros for it,
//     scene_env.h:24-25), which accepts + - * / % ^, parentheses, named
//     functions, VAR variables and, for single values, %ENV% references;
//     the fields the reference reads with sscanf("%d") / "%lf" keep that;
//   - after the file: TARGET_POS, CAMERA_MATRIX, CAMERA_ANGLE and the
//     environment overrides SAMPLING, SUPERSAMPLING, NEXTEVENTESTIMATION,
//     IMAGE_X, IMAGE_Y, LUMINANCE_CUTOFF, ENVLIGHT_R/G/B (scene_env.h:1858-1920).
// Anything outside the subset that would change the image (MLT/ERPT, media,
// lens, textures, other entity types and BRDFs, ...) fails with
// PRT_E_UNSUPPORTED and names the keyword.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prt.h"
#include "prt_internal.h"

using namespace prt;

namespace {

const double kPi = 3.14159265358979323846;
const double kRad = kPi / 180.0;  // PS_RAD
const double kInf = 1.0e128;      // PS_INF

struct Var {
  std::string name;
  double value;
};

struct ParseError {
  int code;
  std::string msg;
};

[[noreturn]] void raise(int code, const std::string& msg) { throw ParseError{code, msg}; }

// ------------------------------------------------------------ expressions
// Recursive-descent evaluator with the grammar of proj/Expr.h:
//   expr   := term (('+'|'-') term)*
//   term   := factor (('*'|'/'|'%') factor)*        a % b = a - floor(a/b)*b
//   factor := signed ('^' factor)?                    right-associative
//   signed := ('-'|'+') signed | number
//   number := constant | '(' expr ')' | name [ '(' args ')' ]
// Constants: a leading '0' selects the base (0x.. 16, 0t.. 10, 0b<n>:.. base
// n, 0. decimal, otherwise octal); the fraction digits are accumulated as an
// integer and divided once; an exponent multiplies or divides by base^n.
class Expr {
 public:
  Expr(const char* s, const std::vector<Var>& vars) : s_(s), vars_(vars) {
    const size_t n = std::strlen(s);
    if (n >= 2 && (s[n - 2] == ' ' || s[n - 2] == '\t')) raise(PRT_E_INVALID, std::string("extra blank at the end of [") + s + "]");
  }
  double expr() {
    space();
    double x = term();
    for (;;) {
      switch (*s_) {
        case '+': ++s_; x += term(); break;
        case '-': ++s_; x -= term(); break;
        default: return x;
      }
    }
  }
  const char* cur() const { return s_; }
  void next() { ++s_; }
  bool bad() const { return bad_; }

 private:
  const char* s_;
  const std::vector<Var>& vars_;
  bool bad_ = false;

  void space() { while (std::isspace((unsigned char)*s_)) ++s_; }
  void error() { bad_ = true; }
  double term() {
    double x = factor();
    for (;;) {
      switch (*s_) {
        case '*': ++s_; x *= factor(); break;
        case '/': ++s_; x /= factor(); break;
        case '%': { ++s_; const double y = factor(); x = x - std::floor(x / y) * y; break; }
        default: return x;
      }
    }
  }
  double factor() {
    const double x = signednumber();
    if (*s_ == '^') {
      ++s_;
      return std::pow(x, factor());
    }
    return x;
  }
  double signednumber() {
    space();
    if (*s_ == '-') { ++s_; return -signednumber(); }
    if (*s_ == '+') { ++s_; return signednumber(); }
    return number();
  }
  static int digit(int c) {
    return std::isdigit(c) ? c - '0' : (c >= 'a' && c <= 'z') ? c - 'a' + 10 : (c >= 'A' && c <= 'Z') ? c - 'A' + 10 : -1;
  }
  static double expt(int a, int n) {
    double t = a, x = 1.0;
    for (; n > 0; n >>= 1) {
      if (n & 1) x *= t;
      t *= t;
    }
    return x;
  }
  double posconst() {
    int base = 10;
    space();
    if (*s_ == '0') {
      ++s_;
      switch (*s_) {
        case 'b':
          ++s_;
          for (base = 0; std::isdigit((unsigned char)*s_); ++s_) base = base * 10 + *s_ - '0';
          if (*s_ != ':') error();
          else ++s_;
          break;
        case 'x': ++s_; base = 16; break;
        case 't': ++s_; base = 10; break;
        case '.': base = 10; break;
        default: base = 8; break;
      }
    }
    double x = 0.0;
    int d;
    for (; d = digit(*s_), d >= 0 && d < base; ++s_) x = x * base + d;
    if (*s_ == '.') {
      ++s_;
      double y = 1.0;
      for (; d = digit(*s_), d >= 0 && d < base; ++s_) {
        x = x * base + d;
        y *= base;
      }
      x /= y;
    }
    if (*s_ == 'e' || *s_ == 'E') {
      ++s_;
      bool pos = true;
      if (*s_ == '-') { ++s_; pos = false; }
      else if (*s_ == '+') { ++s_; }
      int e = 0;
      for (; std::isdigit((unsigned char)*s_); ++s_) e = e * 10 + *s_ - '0';
      const double y = expt(base, e);
      x = pos ? x * y : x / y;
    }
    return x;
  }
  double paren() {
    space();
    if (*s_ != '(') { error(); return 0.0; }
    ++s_;
    const double x = expr();
    space();
    if (*s_ != ')') { error(); return x; }
    ++s_;
    return x;
  }
  void paren2(double* x, double* y) {
    space();
    if (*s_ != '(') { error(); *x = *y = 0.0; return; }
    ++s_;
    *x = expr();
    space();
    if (*s_ != ',') { error(); *y = 0.0; return; }
    ++s_;
    *y = expr();
    space();
    if (*s_ != ')') { error(); return; }
    ++s_;
  }
  bool is(const char* f, size_t n, const char* name) { return std::strlen(name) == n && std::strncmp(f, name, n) == 0; }
  double number() {
    space();
    if (std::isdigit((unsigned char)*s_) || *s_ == '.') return posconst();
    if (*s_ == '(') return paren();
    if (std::isalpha((unsigned char)*s_)) {
      const char* f = s_;
      for (++s_; std::isalpha((unsigned char)*s_) || std::isdigit((unsigned char)*s_); ++s_) {
      }
      const size_t n = (size_t)(s_ - f);
      double x, y;
      if (is(f, n, "pi")) return kPi;
      if (is(f, n, "e")) return std::exp(1.0);
      if (is(f, n, "sqrt")) return std::sqrt(paren());
      if (is(f, n, "exp")) return std::exp(paren());
      if (is(f, n, "log")) return std::log(paren());
      if (is(f, n, "pow")) { paren2(&x, &y); return std::pow(x, y); }
      if (is(f, n, "sin")) return std::sin(paren());
      if (is(f, n, "cos")) return std::cos(paren());
      if (is(f, n, "tan")) return std::tan(paren());
      if (is(f, n, "asin")) return std::asin(paren());
      if (is(f, n, "acos")) return std::acos(paren());
      if (is(f, n, "atan")) return std::atan(paren());
      if (is(f, n, "atan2")) { paren2(&x, &y); return std::atan2(x, y); }
      if (is(f, n, "sind")) return std::sin(paren() * (kPi / 180.0));
      if (is(f, n, "cosd")) return std::cos(paren() * (kPi / 180.0));
      if (is(f, n, "tand")) return std::tan(paren() * (kPi / 180.0));
      if (is(f, n, "dasin")) return std::asin(paren()) * (180.0 / kPi);
      if (is(f, n, "dacos")) return std::acos(paren()) * (180.0 / kPi);
      if (is(f, n, "datan")) return std::atan(paren()) * (180.0 / kPi);
      if (is(f, n, "datan2")) { paren2(&x, &y); return std::atan2(x, y) * (180.0 / kPi); }
      if (is(f, n, "floor")) return std::floor(paren());
      if (is(f, n, "ceil")) return std::ceil(paren());
      for (const Var& v : vars_)  // first definition wins (a linear search)
        if (is(f, n, v.name.c_str())) return v.value;
      error();
      return 0.0;
    }
    error();
    return 0.0;
  }
};

// expr_eval(const char*) (Expr.cpp): one value; "%NAME%" reads the
// environment variable NAME with atof (0 when unset).
double eval1(const char* s, const std::vector<Var>& vars) {
  Expr x(s, vars);  // checks the trailing blank first, as the reference does
  if (*s == '%') {
    const char* e = std::strchr(s + 1, '%');
    if (!e) raise(PRT_E_INVALID, std::string("unterminated environment reference [") + s + "]");
    const std::string name(s + 1, e);
    const char* v = std::getenv(name.c_str());
    return v ? std::atof(v) : 0.0;
  }
  const double r = x.expr();
  if (*x.cur() == '\n') x.next();
  if (*x.cur() || x.bad()) raise(PRT_E_INVALID, std::string("bad expression [") + s + "]");
  return r;
}

// expr_eval(const char*, std::vector<double*>&): n values in one line.
void evaln(const char* s, int n, double* out, const std::vector<Var>& vars) {
  Expr x(s, vars);
  for (int i = 0; i < n; ++i) {
    out[i] = x.expr();
    if (x.bad()) break;
  }
  if (*x.cur() == '\n') x.next();
  if (*x.cur() || x.bad()) raise(PRT_E_INVALID, std::string("bad expression [") + s + "]");
}

// "key v1 v2 ...": the values after the first blank (sscanf*fs)
void evaln_after_key(const char* line, int n, double* out, const std::vector<Var>& vars) {
  const char* p = std::strchr(line, ' ');
  if (!p) raise(PRT_E_INVALID, std::string("missing values in [") + line + "]");
  evaln(p + 1, n, out, vars);
}

int scan_int(const char* s) { return (int)std::strtol(s, nullptr, 10); }

// ------------------------------------------------------------ 4x4 matrices
// Matrix4x4 of proj/matrix.h: row-major val[4*i+j]; points are rows,
// p' = p * M with the translation in row 3.
struct M4 {
  double v[16];
  static M4 identity() {
    M4 m{};
    m.v[0] = m.v[5] = m.v[10] = m.v[15] = 1.0;
    return m;
  }
  double& at(int i, int j) { return v[4 * i + j]; }
  double at(int i, int j) const { return v[4 * i + j]; }
  static M4 scale(double x, double y, double z) {
    M4 m = identity();
    m.at(0, 0) = x; m.at(1, 1) = y; m.at(2, 2) = z;
    return m;
  }
  static M4 translation(double x, double y, double z) {
    M4 m = identity();
    m.at(3, 0) = x; m.at(3, 1) = y; m.at(3, 2) = z;
    return m;
  }
  static M4 rot_x(double a) {
    M4 m = identity();
    m.at(1, 1) = m.at(2, 2) = std::cos(a);
    m.at(1, 2) = std::sin(a);
    m.at(2, 1) = -m.at(1, 2);
    return m;
  }
  static M4 rot_y(double a) {
    M4 m = identity();
    m.at(0, 0) = m.at(2, 2) = std::cos(a);
    m.at(2, 0) = std::sin(a);
    m.at(0, 2) = -m.at(2, 0);
    return m;
  }
  static M4 rot_z(double a) {
    M4 m = identity();
    m.at(0, 0) = m.at(1, 1) = std::cos(a);
    m.at(0, 1) = std::sin(a);
    m.at(1, 0) = -m.at(0, 1);
    return m;
  }
  void apply(const double in[3], double out[3]) const {
    const double x = in[0], y = in[1], z = in[2];
    out[0] = x * at(0, 0) + y * at(1, 0) + z * at(2, 0) + at(3, 0);
    out[1] = x * at(0, 1) + y * at(1, 1) + z * at(2, 1) + at(3, 1);
    out[2] = x * at(0, 2) + y * at(1, 2) + z * at(2, 2) + at(3, 2);
  }
  // vector_coordinate_transformation: transpose of the inverse (cofactors / det)
  M4 normal_matrix() const {
    const double* a = v;
    double inv[16];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    M4 r{};
    if (det == 0.0) return r;  // the reference prints an error and goes on with zeros
    det = 1.0 / det;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) r.v[4 * i + j] = inv[4 * j + i] * det;
    return r;
  }
};

double len3(const double* a) { return std::sqrt((a[0] * a[0] + a[2] * a[2]) + a[1] * a[1]); }

// ------------------------------------------------------------ OBJ meshes
struct Mesh {
  std::vector<float> v;       // 3 per vertex
  std::vector<float> vcol;    // 3 per vertex (only when every vertex line gives r g b)
  std::vector<int32_t> f;     // 3 per triangle, 0-based
  bool colours = true;
};

// The face forms of ObjLoader.cpp: "a b c [d]", "a/t b/t c/t [d/t]",
// "a//n ...", "a/t/n ..."; quads become (a,b,c) and (c,d,a).
void read_obj(const std::string& path, Mesh& m) {
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) raise(PRT_E_INVALID, "cannot open OBJ file " + path);
  char buf[4096];
  int nv_lines = 0;
  while (std::fgets(buf, sizeof(buf), fp)) {
    if (buf[0] == 'v' && buf[1] == ' ') {
      float x[6];
      const int n = std::sscanf(buf + 2, "%f %f %f %f %f %f", x, x + 1, x + 2, x + 3, x + 4, x + 5);
      if (n < 3) {
        std::fclose(fp);
        raise(PRT_E_INVALID, "bad vertex line in " + path);
      }
      m.v.insert(m.v.end(), {x[0], x[1], x[2]});
      if (n == 6) m.vcol.insert(m.vcol.end(), {x[3], x[4], x[5]});
      else m.colours = false;
      ++nv_lines;
    } else if (buf[0] == 'f' && buf[1] == ' ') {
      int idx[4], cnt = 0;
      const char* p = buf + 2;
      while (cnt < 4) {
        while (*p == ' ' || *p == '\t') ++p;
        if (!std::isdigit((unsigned char)*p) && *p != '-') break;
        char* end;
        const long a = std::strtol(p, &end, 10);
        idx[cnt++] = (int)a;
        p = end;
        while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;  // skip /t/n
      }
      if (cnt < 3) {
        std::fclose(fp);
        raise(PRT_E_INVALID, "bad face line in " + path);
      }
      for (int k = 0; k < cnt; ++k)
        if (idx[k] < 1) {
          std::fclose(fp);
          raise(PRT_E_UNSUPPORTED, "OBJ face indices must be positive (1-based) in " + path);
        }
      m.f.insert(m.f.end(), {idx[0] - 1, idx[1] - 1, idx[2] - 1});
      if (cnt == 4) m.f.insert(m.f.end(), {idx[2] - 1, idx[3] - 1, idx[0] - 1});
    }
  }
  std::fclose(fp);
  if (nv_lines == 0) m.colours = false;
  const int32_t nv = (int32_t)(m.v.size() / 3);
  for (int32_t i : m.f)
    if (i >= nv) raise(PRT_E_INVALID, "OBJ face index beyond the vertex count in " + path);
}

// ------------------------------------------------------------ lines
struct Reader {
  FILE* fp;
  char buf[1024];
  // fgets with "\r\n" -> "\n"
  bool raw() {
    if (!std::fgets(buf, sizeof(buf), fp)) return false;
    const size_t n = std::strlen(buf);
    if (n >= 2 && buf[n - 2] == '\r' && buf[n - 1] == '\n') {
      buf[n - 2] = '\n';
      buf[n - 1] = '\0';
    }
    return true;
  }
  // getLine: one '#' line is skipped (scene_env.h:141-149)
  bool line() {
    if (!raw()) return false;
    if (buf[0] == '#') return raw();
    return true;
  }
  bool is(const char* kw) const { return std::strcmp(buf, kw) == 0; }
  bool starts(const char* kw) const { return std::strncmp(buf, kw, std::strlen(kw)) == 0; }
};

std::string dir_of(const std::string& path) {
  const size_t k = path.find_last_of("/\\");
  return k == std::string::npos ? std::string() : path.substr(0, k + 1);
}

}  // namespace

struct prt_scene_file {
  prt_scene_desc desc{};
  std::vector<prt_material> mats;
  std::vector<prt_entity> ents;
  std::vector<std::vector<float>> verts, albedo;
  std::vector<std::vector<int32_t>> faces;
  std::string output = "output.ppm";
  double gamma_offset = 0.0;
  // metropolisTransport and its parameters (scene_env.h:205-208, 309-315)
  int metropolis = 0, erpt = 0;
  double mutation = 2.0, pre_sample = 0.4;
  // IMAGEDUMP (scene_env.h:168, 305, 402, 1905)
  int image_dump = 60;
};

namespace {

void load(const std::string& path, prt_scene_file& F) {
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) raise(PRT_E_INVALID, "cannot open scene file " + path);
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{fp};
  Reader R{fp, {}};
  std::vector<Var> vars;
  // SceneEnv::Default (scene_env.h:262-350) and the luminance_cutoff initialiser
  prt_camera& cam = F.desc.camera;
  prt_sampling& smp = F.desc.sampling;
  double cpos[3] = {50.0, 52.0, 220.0}, cdir[3] = {0.0, -0.04, -1.0}, cup[3] = {0.0, 1.0, 0.0};
  {  // Default() stores normalize(0, -0.04, -1): v * (1 / length)
    const double inv = 1.0 / len3(cdir);
    for (double& c : cdir) c = c * inv;
  }
  double target[3] = {kInf, kInf, kInf};
  double wsw = 30.0, wsh = 30.0, wsd = 40.0, angle = kInf;
  int W = 320, H = 240, samples = 8, supersamples = 2, depth_limit = 64, depth = 8, nee = 0;
  double cutoff = 0.025, sensor = 1.0, bg[3] = {0, 0, 0}, envl[3] = {0, 0, 0};
  int mlt = 0, erpt = 0, media = 0, tent = 0;
  double lens = -1.0;
  std::vector<M4> camera_mats;
  const std::string base_dir = dir_of(path);
  auto one = [&](const char* s) { return eval1(s, vars); };
  auto unsupported = [&](const std::string& what) { raise(PRT_E_UNSUPPORTED, what + " is outside the supported subset"); };

  while (R.line()) {
    if (R.is("gamma_offset\n")) { R.line(); F.gamma_offset = one(R.buf); continue; }
    if (R.is("luminance_cutoff\n")) { R.line(); cutoff = one(R.buf); continue; }
    if (R.is("tentfilter\n")) { R.line(); tent = (int)one(R.buf); continue; }
    if (R.is("IMAGEDUMP\n")) { R.line(); F.image_dump = std::atoi(R.buf); continue; }
    if (R.is("TIMELIMIT\n") || R.is("THREAD\n")) { R.line(); (void)one(R.buf); continue; }
    if (R.starts("VAR ")) {
      std::string name(R.buf + 4);
      if (!name.empty() && name.back() == '\n') name.pop_back();
      R.line();
      vars.push_back(Var{name, one(R.buf)});
      continue;
    }
    if (R.is("DEPTH_LIMIT\n")) { R.line(); depth_limit = (int)one(R.buf); continue; }
    if (R.is("DEPTH\n")) { R.line(); depth = (int)one(R.buf); continue; }
    if (R.is("OUTPUT\n")) {
      R.line();
      F.output = R.buf;
      if (!F.output.empty() && F.output.back() == '\n') F.output.pop_back();
      continue;
    }
    if (R.is("IMAGE\n")) { R.line(); if (std::sscanf(R.buf, "%d %d", &W, &H) != 2) raise(PRT_E_INVALID, "bad IMAGE line"); continue; }
    if (R.is("SCREEN\n")) { R.line(); double v[3]; evaln(R.buf, 3, v, vars); wsw = v[0]; wsh = v[1]; wsd = v[2]; continue; }
    if (R.is("SAMPLING\n")) { R.line(); samples = scan_int(R.buf); continue; }
    if (R.is("SUPERSAMPLING\n")) { R.line(); supersamples = scan_int(R.buf); continue; }
    if (R.is("velocity\n")) { R.line(); double v[3]; evaln(R.buf, 3, v, vars); if (v[0] != 0 || v[1] != 0 || v[2] != 0) unsupported("velocity"); continue; }
    if (R.is("use_mollweide_projection\n") || R.is("special_relativity_effects\n") ||
        R.is("use_accretion_disk_temperature\n")) {
      const std::string kw(R.buf, std::strlen(R.buf) - 1);
      R.line();
      if ((int)one(R.buf) != 0) unsupported(kw);
      continue;
    }
    if (R.is("color_doppler_factor_effect\n")) { R.line(); (void)one(R.buf); continue; }
    if (R.is("CAMERA_POS\n")) { R.line(); evaln(R.buf, 3, cpos, vars); continue; }
    if (R.is("CAMERA_DIR\n")) { R.line(); evaln(R.buf, 3, cdir, vars); continue; }
    if (R.is("TARGET_POS\n")) { R.line(); evaln(R.buf, 3, target, vars); continue; }
    if (R.is("CAMERA_UPVEC\n")) { R.line(); evaln(R.buf, 3, cup, vars); continue; }
    if (R.is("CAMERA_ANGLE\n")) { R.line(); angle = one(R.buf); continue; }
    if (R.is("CAMERA_MATRIX\n")) {
      while (R.line()) {
        if (R.buf[0] == '\n' || R.buf[0] == '\0') break;
        double v[3];
        if (R.starts("rotation ")) {
          evaln_after_key(R.buf, 3, v, vars);
          camera_mats.push_back(M4::rot_x(v[0] * kRad));
          camera_mats.push_back(M4::rot_y(v[1] * kRad));
          camera_mats.push_back(M4::rot_z(v[2] * kRad));
        } else if (R.starts("translate ")) {
          evaln_after_key(R.buf, 3, v, vars);
          camera_mats.push_back(M4::translation(v[0], v[1], v[2]));
        }
      }
      continue;
    }
    if (R.is("LENS\n")) { R.line(); double v[2]; evaln(R.buf, 2, v, vars); lens = v[0]; continue; }
    if (R.is("ENVLIGHT\n")) { R.line(); evaln(R.buf, 3, envl, vars); continue; }
    if (R.is("SCATTERING\n") || R.is("ABSORBING\n")) { R.line(); R.line(); continue; }  // media only
    if (R.is("PAHASE\n")) { R.line(); continue; }
    if (R.starts("PAHASE ")) continue;
    if (R.is("BACKGROUND\n")) { R.line(); evaln(R.buf, 3, bg, vars); continue; }
    if (R.is("nextEventEstimation\n")) { R.line(); nee = scan_int(R.buf); continue; }
    if (R.is("participatingMedia\n")) { R.line(); media = scan_int(R.buf); continue; }
    if (R.is("metropolisTransport\n")) { R.line(); mlt = scan_int(R.buf); continue; }
    if (R.is("energyRedistributionPathTracing\n")) { R.line(); erpt = scan_int(R.buf); continue; }
    if (R.is("mutation\n")) { R.line(); F.mutation = one(R.buf); continue; }
    if (R.is("pre_sample\n")) { R.line(); F.pre_sample = one(R.buf); continue; }
    // mlt_r1r2 only feeds generate_new_path's disabled local mutation (render.h:207)
    if (R.is("mlt_r1r2\n")) { R.line(); continue; }
    if (R.is("sensor_response\n")) { R.line(); sensor = std::strtod(R.buf, nullptr); continue; }
    if (R.starts("background_texture ")) unsupported("background_texture");
    if (R.starts("map_coef ") || R.starts("coef ")) continue;
    if (R.starts("BLACKHOLE")) unsupported("BLACKHOLE");
    if (R.starts("WORMHOLE")) unsupported("WORMHOLE");
    if (!R.is("OBJECT\n")) continue;

    // ---- OBJECT (scene_env.h:854-1850)
    R.line();
    enum { NONE, SPHERE, UVPLANE, POLYGON } type = NONE;
    double p[4] = {0, 0, 0, 0}, U[3] = {0, 0, 0}, V[3] = {0, 0, 0};
    std::string objfile;
    if (R.starts("sphere ")) {
      evaln_after_key(R.buf, 4, p, vars);
      type = SPHERE;
    } else if (R.starts("uvplane ")) {
      evaln_after_key(R.buf, 3, p, vars);
      for (int k = 0; k < 2; ++k) {
        R.line();
        if (R.buf[0] == 'U') evaln_after_key(R.buf, 3, U, vars);
        else evaln_after_key(R.buf, 3, V, vars);
      }
      type = UVPLANE;
    } else if (R.starts("objfile ")) {
      objfile = R.buf + 8;
      if (!objfile.empty() && objfile.back() == '\n') objfile.pop_back();
      type = POLYGON;
    } else if (R.starts("plane ") || R.starts("hemisphere ") || R.starts("disk ") || R.starts("disk2 ") ||
               R.starts("infinity_light ")) {
      unsupported(std::string("OBJECT ") + std::string(R.buf, std::strcspn(R.buf, " \n")));
    }
    // Material() defaults (material.h:60-94): a mirror
    prt_material m{};
    m.reflection_type = PRT_REFLECTION_DIFFUSE;
    m.roughness = 0.0;
    m.refractive_index = 1.5;
    int ninv = 1;
    std::vector<M4> mats;
    while (R.raw()) {  // attribute lines: fgets, a blank line ends the block
      if (R.buf[0] == '\n') break;
      double v[3];
      if (R.starts("rotation ")) {
        evaln_after_key(R.buf, 3, v, vars);
        mats.push_back(M4::rot_x(v[0] * kRad));
        mats.push_back(M4::rot_y(v[1] * kRad));
        mats.push_back(M4::rot_z(v[2] * kRad));
      } else if (R.starts("translate ")) {
        evaln_after_key(R.buf, 3, v, vars);
        mats.push_back(M4::translation(v[0], v[1], v[2]));
      } else if (R.starts("scale ")) {
        evaln_after_key(R.buf, 3, v, vars);
        mats.push_back(M4::scale(v[0], v[1], v[2]));
      } else if (R.starts("emission ")) {
        evaln_after_key(R.buf, 3, m.emission, vars);
      } else if (R.starts("color ")) {
        evaln_after_key(R.buf, 3, m.color, vars);
      } else if (R.starts("specular ") || R.starts("r_refractive_index ")) {
        // only the Ward / Phong BRDFs and dispersion formulas read these
      } else if (R.starts("refractive_index ")) {
        evaln_after_key(R.buf, 1, &m.refractive_index, vars);
      } else if (R.starts("roughness ")) {
        evaln_after_key(R.buf, 1, &m.roughness, vars);
      } else if (R.is("reflection diffuse\n")) {
        m.reflection_type = PRT_REFLECTION_DIFFUSE;
        m.roughness = 1.0;
      } else if (R.is("reflection reflection\n")) {
        m.reflection_type = PRT_REFLECTION_REFRACTION;
        if (m.refractive_index == 0.0) m.refractive_index = 1.5;
      } else if (R.is("reflection specular\n")) {
        m.reflection_type = PRT_REFLECTION_SPECULAR;
        m.roughness = 0.0;
      } else if (R.starts("reflection ")) {
        unsupported(std::string(R.buf, std::strlen(R.buf) - 1));
      } else if (R.starts("normal ")) {
        ninv = scan_int(R.buf + 7);
      } else if (R.starts("smooth ") || R.starts("shadow ") || R.starts("background ") ||
                 R.starts("parallel_light ") || R.starts("spot_light ") || R.starts("blackhole_disk ") ||
                 R.starts("media_ignore_front_back ")) {
        if (scan_int(std::strchr(R.buf, ' ') + 1) != 0) unsupported(std::string(R.buf, std::strcspn(R.buf, " ")));
      } else if (R.starts("back ")) {
        if (scan_int(R.buf + 5) != 1) unsupported("back 0");
      } else if (R.starts("texture ") || R.starts("bump_texture ") || R.starts("ibl_texture ") ||
                 R.is("MATERIAL_SYMBOL\n") || R.is("SCATTERING\n") || R.is("ABSORBING\n") || R.is("BSSRDF\n") ||
                 R.is("level\n")) {
        unsupported(std::string(R.buf, std::strcspn(R.buf, " \n")));
      }
    }
    if (type == NONE) continue;
    const int mat_index = (int)F.mats.size();
    F.mats.push_back(m);
    prt_entity e{};
    e.material = mat_index;
    e.normal_vector_inverse = ninv;
    if (type == SPHERE) {
      // Sphere::MatrixTransformation moves the centre only (sphere.h:505-511)
      double c[3] = {p[0], p[1], p[2]};
      for (const M4& t : mats) t.apply(c, c);
      e.type = PRT_ENTITY_SPHERE;
      std::memcpy(e.p0, c, sizeof(c));
      e.radius = p[3];
    } else if (type == UVPLANE) {
      // UVPlane ctor (normalised axes, lengths) then MatrixTransformation:
      // org by M, axes by the transposed inverse (plane.h:143-166, 375-389)
      const double ul = len3(U), vl = len3(V);
      if (!(ul > 0) || !(vl > 0)) raise(PRT_E_INVALID, "uvplane with a zero axis");
      double org[3] = {p[0], p[1], p[2]};
      double ua[3] = {U[0] * (1.0 / ul), U[1] * (1.0 / ul), U[2] * (1.0 / ul)};
      double va[3] = {V[0] * (1.0 / vl), V[1] * (1.0 / vl), V[2] * (1.0 / vl)};
      for (const M4& t : mats) {
        t.apply(org, org);
        const M4 g = t.normal_matrix();
        g.apply(ua, ua);
        g.apply(va, va);
        const double nu = len3(ua), nv = len3(va);
        if (std::fabs(nu - 1.0) > 1e-9 || std::fabs(nv - 1.0) > 1e-9)
          unsupported("a uvplane transform that changes its axis lengths");
      }
      e.type = PRT_ENTITY_UVPLANE;
      std::memcpy(e.p0, org, sizeof(org));
      for (int k = 0; k < 3; ++k) {
        e.p1[k] = ua[k] * ul;
        e.p2[k] = va[k] * vl;
      }
    } else {
      // Polygon::MatrixTransformation: every vertex through each matrix in
      // double, stored as float (polygon.h:749-760); face normals are
      // recomputed from the final vertices when the scene is built
      std::string file = objfile;
      if (FILE* t = std::fopen(file.c_str(), "r")) std::fclose(t);
      else file = base_dir + objfile;
      Mesh mesh;
      read_obj(file, mesh);
      for (const M4& t : mats)
        for (size_t i = 0; i + 2 < mesh.v.size(); i += 3) {
          double a[3] = {mesh.v[i], mesh.v[i + 1], mesh.v[i + 2]};
          t.apply(a, a);
          mesh.v[i] = (float)a[0]; mesh.v[i + 1] = (float)a[1]; mesh.v[i + 2] = (float)a[2];
        }
      std::vector<float> alb;
      if (mesh.colours && !mesh.vcol.empty()) {
        // per-vertex colours become per-face albedo when a face's three
        // vertices agree (the barycentric blend of polygon.h:214-225 is then
        // constant); otherwise the blend would differ
        const size_t nf = mesh.f.size() / 3;
        alb.resize(nf * 3);
        for (size_t f = 0; f < nf; ++f) {
          const float* c0 = &mesh.vcol[3 * (size_t)mesh.f[3 * f]];
          const float* c1 = &mesh.vcol[3 * (size_t)mesh.f[3 * f + 1]];
          const float* c2 = &mesh.vcol[3 * (size_t)mesh.f[3 * f + 2]];
          for (int k = 0; k < 3; ++k) {
            if (c0[k] != c1[k] || c0[k] != c2[k]) unsupported("per-vertex colours that vary across a face");
            alb[3 * f + k] = c0[k];
          }
        }
      }
      F.verts.push_back(std::move(mesh.v));
      F.faces.push_back(std::move(mesh.f));
      F.albedo.push_back(std::move(alb));
      e.type = PRT_ENTITY_POLYGON;
      e.n_vertices = (int32_t)(F.verts.back().size() / 3);
      e.n_faces = (int64_t)(F.faces.back().size() / 3);
    }
    F.ents.push_back(e);
  }

  // after the file (scene_env.h:1858-1920)
  if (target[0] < kInf) {
    double d[3] = {target[0] - cpos[0], target[1] - cpos[1], target[2] - cpos[2]};
    const double l = len3(d);
    for (int k = 0; k < 3; ++k) cdir[k] = d[k] * (1.0 / l);
  }
  if (!camera_mats.empty()) {
    double org[3] = {0, 0, 0};
    for (const M4& t : camera_mats) {
      t.apply(org, org);
      double rel[3] = {cdir[0] - org[0], cdir[1] - org[1], cdir[2] - org[2]};
      t.apply(rel, cdir);
      t.apply(cpos, cpos);
    }
  }
  const auto env = [&](const char* name) -> const char* { return std::getenv(name); };
  if (const char* s = env("SAMPLING")) samples = (int)eval1(s, vars);
  if (const char* s = env("SUPERSAMPLING")) supersamples = (int)eval1(s, vars);
  if (const char* s = env("NEXTEVENTESTIMATION")) nee = (int)eval1(s, vars);
  if (const char* s = env("IMAGE_X")) W = (int)eval1(s, vars);
  if (const char* s = env("IMAGE_Y")) H = (int)eval1(s, vars);
  if (const char* s = env("LUMINANCE_CUTOFF")) cutoff = eval1(s, vars);
  if (const char* s = env("METROPOLIS")) mlt = (int)eval1(s, vars);
  if (const char* s = env("ERPT")) erpt = (int)eval1(s, vars);
  if (const char* s = env("MUTATION")) F.mutation = eval1(s, vars);
  if (const char* s = env("PRESAMPLE")) F.pre_sample = eval1(s, vars);
  if (const char* s = env("ENVLIGHT_R")) envl[0] = eval1(s, vars);
  if (const char* s = env("ENVLIGHT_G")) envl[1] = eval1(s, vars);
  if (const char* s = env("ENVLIGHT_B")) envl[2] = eval1(s, vars);
  if (const char* s = env("IMAGEDUMP")) F.image_dump = std::atoi(s);
  F.metropolis = mlt;
  F.erpt = erpt;
  if (mlt || erpt) {  // scene_env.h:1946-1968 (tent filter and the 0.25 sensor factor live in the renderers)
    if (F.pre_sample <= 0.0) F.pre_sample = 0.4;
    if (F.mutation <= 0.0) F.mutation = 32.0;
  }
  if (media) unsupported("participatingMedia");
  if (tent) unsupported("tentfilter");
  if (lens > 0.0) unsupported("LENS");

  std::memcpy(cam.camera_position, cpos, sizeof(cpos));
  std::memcpy(cam.camera_dir, cdir, sizeof(cdir));  // normalised by prt_scene_create
  std::memcpy(cam.camera_up, cup, sizeof(cup));
  cam.world_screen_width = wsw;
  cam.world_screen_height = wsh;
  cam.world_screen_dist = wsd;
  cam.camera_angle = angle < kInf ? angle : 0.0;
  if (angle < kInf && !(angle > 0.0)) raise(PRT_E_UNSUPPORTED, "CAMERA_ANGLE <= 0");
  cam.image_width = W;
  cam.image_height = H;
  smp.samples = samples;
  smp.supersamples = supersamples;
  smp.depth = depth;
  smp.depth_limit = depth_limit;
  smp.next_event_estimation = nee;
  smp.luminance_cutoff = cutoff;
  smp.sensor_response = sensor;
  std::memcpy(smp.background_color, bg, sizeof(bg));
  std::memcpy(smp.environment_light, envl, sizeof(envl));
  // arrays
  size_t mesh = 0;
  for (prt_entity& e : F.ents)
    if (e.type == PRT_ENTITY_POLYGON) {
      e.vertices = F.verts[mesh].data();
      e.indices = F.faces[mesh].data();
      e.face_albedo = F.albedo[mesh].empty() ? nullptr : F.albedo[mesh].data();
      ++mesh;
    }
  F.desc.materials = F.mats.data();
  F.desc.n_materials = (int32_t)F.mats.size();
  F.desc.entities = F.ents.data();
  F.desc.n_entities = (int32_t)F.ents.size();
}

}  // namespace

extern "C" {

prt_scene_file* prt_scene_file_load(const char* path) {
  if (!path) {
    g_err = "null path";
    return nullptr;
  }
  prt_scene_file* f = new (std::nothrow) prt_scene_file();
  if (!f) {
    g_err = "out of host memory";
    return nullptr;
  }
  try {
    load(path, *f);
  } catch (const ParseError& e) {
    g_err = std::string(path) + ": " + e.msg;
    delete f;
    return nullptr;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory while reading " + std::string(path);
    delete f;
    return nullptr;
  }
  return f;
}

const prt_scene_desc* prt_scene_file_desc(const prt_scene_file* f) { return f ? &f->desc : nullptr; }

const char* prt_scene_file_output(const prt_scene_file* f, double* gamma_offset) {
  if (!f) return nullptr;
  if (gamma_offset) *gamma_offset = f->gamma_offset;
  return f->output.c_str();
}

int prt_scene_file_mlt(const prt_scene_file* f, prt_mlt_opts* opts) {
  if (!f) return 0;
  if (opts) {
    std::memset(opts, 0, sizeof(*opts));
    opts->mutation = f->mutation;
    opts->pre_sample = f->pre_sample;
  }
  return f->erpt ? 2 : (f->metropolis ? 1 : 0);  // main.cpp:189-196 tests ERPT first
}

int prt_scene_file_image_dump(const prt_scene_file* f) { return f ? f->image_dump : -1; }

void prt_scene_file_free(prt_scene_file* f) { delete f; }

double prt_eval_expression(const char* text, int* status) {
  std::vector<Var> none;
  try {
    const double v = eval1(text ? text : "", none);
    if (status) *status = PRT_OK;
    return v;
  } catch (const ParseError& e) {
    g_err = e.msg;
    if (status) *status = e.code;
    return 0.0;
  }
}

}  // extern "C"
