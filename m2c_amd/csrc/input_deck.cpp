// input_deck.cpp -- reader of M2C input decks for the options of this library (host code, no HIP).
//
// Syntax (parser/Parser.y:37-137, tokens parser/Lexer.l:15-77). A deck is a sequence of
//   Assignable = Symbol | IntExpr | DblExpr | "String" ;
//   under Assignable { ... }
// ended by the end of the text or by `end`. Assignable = Symbol | Assignable . Symbol |
// Assignable [ IntExpr ]. Expressions: + - * / with the usual precedence, left associative, and
// parentheses; two integers give an integer (C division truncates), otherwise a double. A sign
// belongs to the number literal, so `3 -1` is the two numbers 3 and -1 -- a syntax error, as in the
// reference. Numbers take e/E or Fortran d/D exponents; strings are "[A-z0-9._/-]*"; C and C++
// comments. A bare Symbol names a member of the innermost enclosing `under` object that has one,
// else a top-level object (Parser.y:60-70).
//
// Assignment rules of the reference's assigners (parser/Assigner.h:100-140): an integer field takes
// only an integer, a double field an integer or a double, a keyword field only one of its keywords.
// Aliases bound to one object (IoData.cpp:3299-3302, :1626-1632) share their fields.
//
// The objects follow IoData (IoData.cpp:453-590 Mesh, :843-1046 Equations and Material[i],
// :1049-1343 Space, :1345-1398 ExactRiemannSolution, :1400-1528 MultiPhase and Time, :1606-1700
// BoundaryConditions and InitialCondition, :3293-3330 the top-level objects) and map onto m2c_config.
// Options outside this library's path (output, probes, terminal visualisation, special tools) are
// listed in the message and ignored; options of the path that the library does not provide, and
// physics beyond it (embedded boundaries, lasers, ionisation, viscosity, ...), are rejected with
// M2C_ERR_UNSUPPORTED.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "deck.hpp"

namespace m2c_deck {

[[noreturn]] void fail(int code, int line, const std::string &m) {
  throw DeckError{code, (line > 0 ? "line " + std::to_string(line) + ": " : std::string()) + m};
}

// ------------------------------------------------------------------ tokens

enum Tok { T_EOF, T_END, T_UNDER, T_SYM, T_INT, T_DBL, T_STR, T_CHAR };

struct Token {
  Tok t = T_EOF;
  std::string s;      // symbol or string text
  long long i = 0;
  double d = 0.0;
  char c = 0;
  int line = 1;
};

bool is_digit(char c) { return c >= '0' && c <= '9'; }
bool is_sym0(char c) { return c == '_' || (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z'); }
bool is_sym(char c) { return is_sym0(c) || is_digit(c); }
// body of a string: [A-z0-9._/-], where A-z is the ASCII range 'A'..'z'
bool is_str(char c) { return (c >= 'A' && c <= 'z') || is_digit(c) || c == '.' || c == '/' || c == '-'; }

class Lexer {
 public:
  explicit Lexer(const char *text) : p_(text) {}

  Token next() {
    skip_blank();
    Token t;
    t.line = line_;
    const char c = *p_;
    if (!c) return t;
    if (c == '"') {
      const char *q = p_ + 1;
      while (*q && is_str(*q)) q++;
      if (*q != '"') fail(M2C_ERR_ARG, line_, "malformed string");
      t.t = T_STR;
      t.s.assign(p_ + 1, q);
      p_ = q + 1;
      return t;
    }
    if (number(t)) return t;
    if (is_sym0(c)) {
      const char *q = p_;
      while (is_sym(*q)) q++;
      t.s.assign(p_, q);
      p_ = q;
      t.t = t.s == "end" ? T_END : (t.s == "under" ? T_UNDER : T_SYM);
      return t;
    }
    if (std::strchr("-+*[]^{}=,():;./", c)) {
      t.t = T_CHAR;
      t.c = c;
      p_++;
      return t;
    }
    fail(M2C_ERR_ARG, line_, std::string("unexpected character '") + c + "'");
  }

 private:
  void skip_blank() {
    for (;;) {
      const char c = *p_;
      if (c == '\n') { line_++; p_++; }
      else if (c == ' ' || c == '\t' || c == '\r') p_++;
      else if (c == '/' && p_[1] == '/') { while (*p_ && *p_ != '\n') p_++; }
      else if (c == '/' && p_[1] == '*') {
        p_ += 2;
        while (*p_ && !(p_[0] == '*' && p_[1] == '/')) { if (*p_ == '\n') line_++; p_++; }
        if (!*p_) fail(M2C_ERR_ARG, line_, "unterminated comment");
        p_ += 2;
      } else {
        return;
      }
    }
  }

  // [+-]? digits, with an optional fraction and exponent; the longest match is taken, as flex does
  bool number(Token &t) {
    const char *q = p_;
    if (*q == '+' || *q == '-') q++;
    const char *ds = q;
    while (is_digit(*q)) q++;
    const bool digits = q > ds;
    bool dot = false, frac = false;
    if (*q == '.') {
      const char *fs = q + 1;
      const char *r = fs;
      while (is_digit(*r)) r++;
      frac = r > fs;
      if (digits || frac) { dot = true; q = r; }
    }
    if (!digits && !frac) return false;
    bool expo = false;
    size_t epos = 0;
    if (*q == 'e' || *q == 'E' || *q == 'd' || *q == 'D') {
      const char *r = q + 1;
      if (*r == '+' || *r == '-') r++;
      if (is_digit(*r)) {
        epos = (size_t)(q - p_);
        while (is_digit(*r)) r++;
        q = r;
        expo = true;
      }
    }
    std::string num(p_, q);
    if (expo) num[epos] = 'e';
    p_ = q;
    if (!dot && !expo) {
      t.t = T_INT;
      t.i = std::strtoll(num.c_str(), nullptr, 10);
    } else {
      t.t = T_DBL;
      t.d = std::strtod(num.c_str(), nullptr);
    }
    return true;
  }

  const char *p_;
  int line_ = 1;
};

// ------------------------------------------------------------------ values and the object tree

// names bound to one object in IoData
std::string canonical(const std::string &n) {
  if (n == "NavierStokesEquations") return "Equations";
  if (n == "EmbeddedBoundaryMethod") return "EmbeddedBoundary";
  if (n == "Farfield0" || n == "Farfield") return "Inlet";
  if (n == "Farfield1") return "Outlet";
  return n;
}

// top-level objects of IoData::setupCmdFileVariables (IoData.cpp:3293-3330), plus those that IoData
// registers without a parent object (IoData.cpp:1635, :1695, :1697, :2947, :3028-3030, :3209)
bool root_name(const std::string &n) {
  static const char *names[] = {"ConcurrentPrograms", "EmbeddedBoundary", "Equations", "InitialCondition",
                                "BoundaryConditions", "Mesh", "Space", "ExactRiemannSolution", "Laser",
                                "Ionization", "Time", "MultiPhase", "ReferenceMap", "Output", "SpecialTools",
                                "TerminalVisualization", "DefaultInitialState", "GeometricEntities",
                                "GeometricEntities2D", "EmbeddedSurfaces", "AeroS", "AeroF", "M2CTwin",
                                "TransientInputData"};
  for (const char *s : names)
    if (n == s) return true;
  return false;
}

// members (objects and fields) of the objects this library reads, for resolving bare names:
// 1 = a member, 0 = not a member, -1 = an object without a listed schema (it owns every name)
int member_of(const std::string &type, const std::string &n) {
  static const std::map<std::string, std::vector<const char *>> schema = {
      {"Mesh", {"Type", "X0", "Xmax", "Y0", "Ymax", "Z0", "Zmax", "NumberOfCellsX", "NumberOfCellsY",
                "NumberOfCellsZ", "ControlPointX", "ControlPointY", "ControlPointZ", "BoundaryConditionX0",
                "BoundaryConditionXmax", "BoundaryConditionY0", "BoundaryConditionYmax", "BoundaryConditionZ0",
                "BoundaryConditionZmax"}},
      {"Equations", {"Material", "MaterialTransition", "DummyState"}},
      {"Space", {"NavierStokes", "Boundaries", "LevelSet"}},
      {"NavierStokes", {"Flux", "EntropyFixCoefficient", "Reconstruction", "Smoothing"}},
      {"Reconstruction", {"Type", "SlopeNearInterface", "Limiter", "GeneralizedMinModCoefficient", "VariableType",
                          "Fixes"}},
      {"Time", {"Type", "MaxIts", "TimeStep", "CFL", "MaxTime", "ConvergenceTolerance", "LocalTimeStepping",
                "Explicit"}},
      {"BoundaryConditions", {"Inlet", "Outlet", "Wall"}},
      {"InitialCondition", {"UserDataFile", "ApplyUserDataBeforeGeometricEntities", "InterpolationFunction"}},
      {"GeometricEntities", {"Point", "Plane", "Sphere", "Parallelepiped", "Spheroid", "CylinderAndCone",
                             "CylinderWithSphericalCaps", "ArbitraryEnclosure"}},
  };
  auto it = schema.find(type);
  if (it == schema.end()) return -1;
  for (const char *s : it->second)
    if (n == s) return 1;
  return 0;
}

struct Num {
  bool isint;
  long long i;
  double d;
  double val() const { return isint ? (double)i : d; }
};

// where an Assignable points: a field `leaf` of `obj` (which a `.`, `[` or `{` turns into an object)
struct Ref {
  Node *obj;
  std::string leaf;
  int line;
  Node *as_object() { return obj->kid(leaf, line); }
};

class Parser {
 public:
  Parser(const char *text, Node *root) : lx_(text), root_(root) { adv(); }

  void deck() {
    block();
    if (cur_.t == T_CHAR && cur_.c == '}') fail(M2C_ERR_ARG, cur_.line, "'}' without 'under'");
    if (cur_.t != T_EOF && cur_.t != T_END) fail(M2C_ERR_ARG, cur_.line, "syntax error");
  }

 private:
  void adv() { cur_ = lx_.next(); }
  bool is(char c) const { return cur_.t == T_CHAR && cur_.c == c; }
  void expect(char c) {
    if (!is(c)) fail(M2C_ERR_ARG, cur_.line, std::string("syntax error: expected '") + c + "'");
    adv();
  }

  void block() {
    while (cur_.t == T_UNDER || cur_.t == T_SYM) statement();
  }

  void statement() {
    if (cur_.t == T_UNDER) {
      adv();
      Ref r = assignable();
      expect('{');
      under_.push_back(r.as_object());
      block();
      expect('}');
      under_.pop_back();
      return;
    }
    Ref r = assignable();
    expect('=');
    Value v;
    v.line = r.line;
    if (cur_.t == T_SYM) {
      v.kind = Value::SYM;
      v.s = cur_.s;
      adv();
    } else if (cur_.t == T_STR) {
      v.kind = Value::STR;
      v.s = cur_.s;
      adv();
    } else {
      const Num n = expr();
      if (n.isint) { v.kind = Value::INT; v.i = n.i; }
      else { v.kind = Value::DBL; v.d = n.d; }
    }
    expect(';');
    if (r.obj->kids.count(r.leaf)) fail(M2C_ERR_ARG, r.line, r.leaf + " is an object, not a field");
    r.obj->leaves[r.leaf] = v;
  }

  Ref assignable() {
    if (cur_.t != T_SYM) fail(M2C_ERR_ARG, cur_.line, "syntax error: expected a name");
    Ref r;
    r.line = cur_.line;
    r.leaf = canonical(cur_.s);
    adv();
    // innermost `under` object that has the name; an object whose members are not listed here
    // owns every name except the top-level ones
    r.obj = nullptr;
    for (size_t lv = under_.size(); lv-- > 0 && !r.obj;) {
      Node *u = under_[lv];
      const int mem = member_of(u->type, r.leaf);
      if (u->kids.count(r.leaf) || u->leaves.count(r.leaf) || mem == 1) r.obj = u;
      else if (mem < 0) r.obj = root_name(r.leaf) ? root_ : u;
    }
    if (!r.obj) r.obj = (under_.empty() || root_name(r.leaf)) ? root_ : under_.back();
    for (;;) {
      if (is('.')) {
        adv();
        if (cur_.t != T_SYM) fail(M2C_ERR_ARG, cur_.line, "syntax error: expected a name after '.'");
        r.obj = r.as_object();
        r.leaf = canonical(cur_.s);
        r.line = cur_.line;
        adv();
      } else if (is('[')) {
        adv();
        const Num n = expr();
        if (!n.isint) fail(M2C_ERR_ARG, cur_.line, "an index must be an integer expression");
        expect(']');
        r.obj = r.as_object();
        r.leaf = "#" + std::to_string(n.i);
      } else {
        return r;
      }
    }
  }

  Num expr() {
    Num a = term();
    for (;;) {
      if (is('+') || is('-')) {
        const char op = cur_.c;
        adv();
        a = arith(a, term(), op);
      } else if (cur_.t == T_INT || cur_.t == T_DBL) {
        fail(M2C_ERR_ARG, cur_.line, "syntax error: a signed number follows an operand (write '- 1', not '-1')");
      } else {
        return a;
      }
    }
  }
  Num term() {
    Num a = factor();
    while (is('*') || is('/')) {
      const char op = cur_.c;
      adv();
      a = arith(a, factor(), op);
    }
    return a;
  }
  Num factor() {
    if (cur_.t == T_INT) { Num n{true, cur_.i, 0.0}; adv(); return n; }
    if (cur_.t == T_DBL) { Num n{false, 0, cur_.d}; adv(); return n; }
    if (is('(')) {
      adv();
      const Num n = expr();
      expect(')');
      return n;
    }
    fail(M2C_ERR_ARG, cur_.line, "syntax error: expected a number");
  }
  Num arith(const Num &a, const Num &b, char op) {
    if (a.isint && b.isint) {
      Num r{true, 0, 0.0};
      if (op == '+') r.i = a.i + b.i;
      else if (op == '-') r.i = a.i - b.i;
      else if (op == '*') r.i = a.i * b.i;
      else {
        if (b.i == 0) fail(M2C_ERR_ARG, cur_.line, "integer division by zero");
        r.i = a.i / b.i;
      }
      return r;
    }
    const double x = a.val(), y = b.val();
    Num r{false, 0, 0.0};
    if (op == '+') r.d = x + y;
    else if (op == '-') r.d = x - y;
    else if (op == '*') r.d = x * y;
    else r.d = x / y;
    return r;
  }

  Lexer lx_;
  Node *root_;
  Token cur_;
  std::vector<Node *> under_;
};

// ------------------------------------------------------------------ mapping onto m2c_config

const char *const kBcNames[6] = {"BoundaryConditionX0", "BoundaryConditionXmax", "BoundaryConditionY0",
                                 "BoundaryConditionYmax", "BoundaryConditionZ0", "BoundaryConditionZmax"};

struct Mapper {
  std::vector<std::string> notes;

  static const Value *leaf(const Node *n, const char *name) {
    if (!n) return nullptr;
    auto it = n->leaves.find(name);
    return it == n->leaves.end() ? nullptr : &it->second;
  }
  // ClassDouble: an integer or a double
  static void getd(const Node *n, const char *name, double &out) {
    const Value *v = leaf(n, name);
    if (!v) return;
    if (v->kind == Value::DBL) out = v->d;
    else if (v->kind == Value::INT) out = (double)v->i;
    else fail(M2C_ERR_ARG, v->line, std::string(name) + " cannot be assigned a keyword or string");
  }
  // ClassInt: an integer only
  static void geti(const Node *n, const char *name, int &out) {
    const Value *v = leaf(n, name);
    if (!v) return;
    if (v->kind != Value::INT) fail(M2C_ERR_ARG, v->line, std::string(name) + " takes an integer");
    out = (int)v->i;
  }
  // ClassToken: one of `opts`; keywords in `unsupported` are valid in the reference but not provided
  static void gett(const Node *n, const char *name, int &out, std::initializer_list<std::pair<const char *, int>> opts,
                   std::initializer_list<const char *> unsupported = {}) {
    const Value *v = leaf(n, name);
    if (!v) return;
    if (v->kind != Value::SYM) fail(M2C_ERR_ARG, v->line, std::string(name) + " takes a keyword");
    for (const char *u : unsupported)
      if (v->s == u) fail(M2C_ERR_UNSUPPORTED, v->line, std::string(name) + " = " + v->s + " is not provided");
    for (auto &o : opts)
      if (v->s == o.first) { out = o.second; return; }
    fail(M2C_ERR_ARG, v->line, std::string(name) + ": keyword not understood: " + v->s);
  }
  static void state(const Node *n, double *V, int *matid = nullptr) {
    getd(n, "Density", V[0]);
    getd(n, "VelocityX", V[1]);
    getd(n, "VelocityY", V[2]);
    getd(n, "VelocityZ", V[3]);
    getd(n, "Pressure", V[4]);
    if (matid) geti(n, "MaterialID", *matid);
  }
  static void reject_if_present(const Node *n, const char *what) {
    if (n && !n->empty()) fail(M2C_ERR_UNSUPPORTED, n->line, std::string(what) + " is not provided");
  }
  // a model object whose Type must stay None
  static void reject_type(const Node *n, const char *what) {
    const Value *t = leaf(n, "Type");
    if (t && !(t->kind == Value::SYM && t->s == "None"))
      fail(M2C_ERR_UNSUPPORTED, t->line, std::string(what) + " is not provided");
  }
  void ignore_rest(const Node *n, const std::string &path, std::initializer_list<const char *> used) {
    if (!n) return;
    auto is_used = [&](const std::string &k) {
      for (const char *s : used)
        if (k == s) return true;
      return false;
    };
    for (auto &kv : n->leaves)
      if (!is_used(kv.first))
        notes.push_back("ignored " + path + kv.first + " (line " + std::to_string(kv.second.line) + ")");
    for (auto &kv : n->kids)
      if (!is_used(kv.first))
        notes.push_back("ignored " + path + kv.first + " (line " + std::to_string(kv.second->line) + ")");
  }

  void mesh(const Node *m, m2c_config &c) {
    int type = 0;
    gett(m, "Type", type, {{"ThreeDimensional", 0}}, {"Spherical", "Cylindrical"});
    getd(m, "X0", c.x0); getd(m, "Xmax", c.x1);
    getd(m, "Y0", c.y0); getd(m, "Ymax", c.y1);
    getd(m, "Z0", c.z0); getd(m, "Zmax", c.z1);
    geti(m, "NumberOfCellsX", c.nx); geti(m, "NumberOfCellsY", c.ny); geti(m, "NumberOfCellsZ", c.nz);
    // Mesh.ControlPointX/Y/Z[i] { Coordinate; CellWidth } (MeshResolution1DPointData, IoData.cpp:
    // 430-447; both 0 unless given); MeshGenerator.cpp:41-48: an axis takes either a number of cells
    // or control points, and with control points its cell count is the generator's
    const char *const cpn[3] = {"ControlPointX", "ControlPointY", "ControlPointZ"};
    int *const nn[3] = {&c.nx, &c.ny, &c.nz};
    const double lo[3] = {c.x0, c.y0, c.z0}, hi[3] = {c.x1, c.y1, c.z1};
    if (c.x1 < c.x0 || c.y1 < c.y0 || c.z1 < c.z0) fail(M2C_ERR_ARG, 0, "Mesh: the domain's dimensions are incorrect");
    for (int d = 0; d < 3; d++) {
      const Node *cp = m ? m->find(cpn[d]) : nullptr;
      const std::vector<const Node *> el = elements(cp);
      c.mesh_npts[d] = 0;
      if (el.empty()) {
        if (*nn[d] <= 0) fail(M2C_ERR_ARG, 0, "Mesh.NumberOfCellsX/Y/Z must be given (positive): unable to create the mesh");
        continue;
      }
      if (*nn[d] > 0)
        fail(M2C_ERR_ARG, cp->line, std::string("Mesh.") + cpn[d] + " together with a number of cells (conflicting input)");
      if ((int)el.size() > M2C_MAX_MESH_PTS)
        fail(M2C_ERR_UNSUPPORTED, cp->line,
             std::string("Mesh.") + cpn[d] + ": more than " + std::to_string(M2C_MAX_MESH_PTS) + " control points");
      for (size_t q = 0; q < el.size(); q++) {
        double x = 0.0, h = 0.0;
        getd(el[q], "Coordinate", x);
        getd(el[q], "CellWidth", h);
        c.mesh_pts[d][q][0] = x;
        c.mesh_pts[d][q][1] = h;
        ignore_rest(el[q], std::string("Mesh.") + cpn[d] + "[].", {"Coordinate", "CellWidth"});
      }
      c.mesh_npts[d] = (int)el.size();
      int ncell = 0;
      if (m2c_mesh_1d(lo[d], hi[d], 0, c.mesh_npts[d], &c.mesh_pts[d][0][0], 0, nullptr, nullptr, &ncell) != M2C_OK)
        fail(M2C_ERR_ARG, cp->line,
             std::string("Mesh.") + cpn[d] + ": control points outside the domain, or widths in conflict with them");
      *nn[d] = ncell;
    }
    for (int s = 0; s < 6; s++) {
      gett(m, kBcNames[s], c.bc[s],
           {{"None", 0}, {"Inlet", M2C_BC_INLET}, {"Outlet", M2C_BC_OUTLET}, {"Farfield0", M2C_BC_INLET},
            {"Farfield1", M2C_BC_OUTLET}, {"Farfield", M2C_BC_INLET}, {"Wall", M2C_BC_SLIPWALL},
            {"SlipWall", M2C_BC_SLIPWALL}, {"StickWall", M2C_BC_STICKWALL}, {"NoSlipWall", M2C_BC_STICKWALL},
            {"Symmetry", M2C_BC_SYMMETRY}},
           {"Overset"});
      // a domain face left at None stops the reference at its first boundary update
      // (SpaceOperator.cpp:1729-1733)
      if (c.bc[s] == 0) fail(M2C_ERR_ARG, 0, std::string("Mesh.") + kBcNames[s] + " is None (or not given)");
    }
    ignore_rest(m, "Mesh.", {"Type", "X0", "Xmax", "Y0", "Ymax", "Z0", "Zmax", "NumberOfCellsX", "NumberOfCellsY",
                             "NumberOfCellsZ", cpn[0], cpn[1], cpn[2], kBcNames[0], kBcNames[1], kBcNames[2], kBcNames[3], kBcNames[4],
                             kBcNames[5]});
  }

  void material(const Node *mn, m2c_material &mm) {
    std::memset(&mm, 0, sizeof(mm));
    mm.rhomin = 0.0;                        // MaterialModelData::MaterialModelData
    mm.pmin = -1.7976931348623157e308;
    mm.rhomax = mm.pmax = 1.7976931348623157e308;
    mm.failsafe_density = 1.0e-10;
    int eos = M2C_EOS_SG;
    gett(mn, "EquationOfState", eos,
         {{"StiffenedGas", M2C_EOS_SG}, {"NobleAbelStiffenedGas", M2C_EOS_NASG}, {"MieGruneisen", M2C_EOS_MG},
          {"JonesWilkinsLee", M2C_EOS_JWL}},
         {"ANEOSBirchMurnaghanDebye"});
    mm.eos = eos;
    getd(mn, "DensityCutOff", mm.rhomin);
    getd(mn, "PressureCutOff", mm.pmin);
    getd(mn, "DensityUpperLimit", mm.rhomax);
    getd(mn, "PressureUpperLimit", mm.pmax);
    getd(mn, "DensityPrescribedAtFailure", mm.failsafe_density);
    // temperature laws (tlaw = cv, T0, e0, rho0, cp, h0), read for outputs
    if (eos == M2C_EOS_SG) {                // StiffenedGasModelData defaults
      mm.p[0] = 1.4;
      const Node *e = mn->find("StiffenedGasModel");
      getd(e, "SpecificHeatRatio", mm.p[0]);
      getd(e, "PressureConstant", mm.p[1]);
      getd(e, "SpecificHeatAtConstantVolume", mm.tlaw[0]);
      getd(e, "ReferenceTemperature", mm.tlaw[1]);
      getd(e, "ReferenceSpecificInternalEnergy", mm.tlaw[2]);
      getd(e, "ReferenceDensity", mm.tlaw[3]);
      getd(e, "SpecificHeatAtConstantPressure", mm.tlaw[4]);
      getd(e, "ReferenceSpecificEnthalpy", mm.tlaw[5]);
    } else if (eos == M2C_EOS_NASG) {       // NobleAbelStiffenedGasModelData defaults
      mm.p[0] = 1.4;
      const Node *e = mn->find("NobleAbelStiffenedGasModel");
      getd(e, "SpecificHeatRatio", mm.p[0]);
      getd(e, "PressureConstant", mm.p[1]);
      getd(e, "VolumeConstant", mm.p[2]);
      getd(e, "EnergyConstant", mm.p[3]);
      getd(e, "SpecificHeatAtConstantVolume", mm.tlaw[0]);
    } else if (eos == M2C_EOS_MG) {         // MieGruneisenModelData defaults (copper, g-mm-s units)
      mm.p[0] = 8.96e-3; mm.p[1] = 3.933e6; mm.p[2] = 1.99; mm.p[3] = 1.5; mm.p[4] = 0.0;
      mm.tlaw[0] = 3.90e8;
      const Node *e = mn->find("MieGruneisenModel");
      getd(e, "ReferenceDensity", mm.p[0]);
      getd(e, "BulkSpeedOfSound", mm.p[1]);
      getd(e, "ReferenceGamma", mm.p[2]);
      getd(e, "HugoniotSlope", mm.p[3]);
      getd(e, "ReferenceSpecificInternalEnergy", mm.p[4]);
      getd(e, "SpecificHeatAtConstantVolume", mm.tlaw[0]);
      getd(e, "ReferenceTemperature", mm.tlaw[1]);
      getd(e, "SpecificHeatAtConstantPressure", mm.tlaw[4]);
      getd(e, "ReferenceSpecificEnthalpy", mm.tlaw[5]);
    } else {                                // JonesWilkinsLeeModelData defaults
      mm.p[0] = 0.28; mm.p[1] = 3.712e11; mm.p[2] = 3.23e9; mm.p[3] = 4.15; mm.p[4] = 0.95; mm.p[5] = 1.63e-3;
      const Node *e = mn->find("JonesWilkinsLeeModel");
      getd(e, "Omega", mm.p[0]);
      getd(e, "A1", mm.p[1]);
      getd(e, "A2", mm.p[2]);
      getd(e, "R1", mm.p[3]);
      getd(e, "R2", mm.p[4]);
      getd(e, "Rho0", mm.p[5]);
    }
    // ViscosityModelData / HeatDiffusionModelData (IoData.cpp defaults: mu 1.716e-5, Sutherland
    // T0 273, S 111, mu0 1.716e-5, bulk 0; diffusivity 0)
    const Node *vm = mn->find("ViscosityModel");
    gett(vm, "Type", mm.viscosity, {{"None", 0}, {"Constant", 1}, {"Sutherland", 2}}, {"Artificial"});
    double dyn = 1.716e-5, mu0 = 1.716e-5;
    mm.bulk = 0.0; mm.su_T0 = 273.0; mm.su_S = 111.0;
    getd(vm, "DynamicViscosity", dyn);
    getd(vm, "ReferenceDynamicViscosity", mu0);
    getd(vm, "BulkViscosity", mm.bulk);
    getd(vm, "ReferenceTemperature", mm.su_T0);
    getd(vm, "SutherlandConstant", mm.su_S);
    mm.mu = mm.viscosity == 2 ? mu0 : (mm.viscosity == 1 ? dyn : 0.0);
    const Node *hd = mn->find("HeatDiffusionModel");
    int htype = 0;
    gett(hd, "Type", htype, {{"None", 0}, {"Constant", 1}});
    double kap = 0.0;
    getd(hd, "Diffusivity", kap);
    mm.kappa = htype == 1 ? kap : 0.0;
    reject_type(mn->find("HyperelasticityModel"), "Material.HyperelasticityModel");
  }

  void equations(const Node *eqs, m2c_config &c) {
    const Node *mats = eqs ? eqs->find("Material") : nullptr;
    // Main.cpp:104-115: material IDs are 0 .. n-1, in any order
    const int nmat = mats ? (int)mats->kids.size() : 0;
    if (nmat == 0) fail(M2C_ERR_ARG, 0, "no Equations.Material[i] given");
    if (nmat > M2C_MAX_MAT) fail(M2C_ERR_UNSUPPORTED, 0, "more than " + std::to_string(M2C_MAX_MAT) + " materials");
    for (int m = 0; m < nmat; m++)
      if (!mats->elem(m)) fail(M2C_ERR_ARG, 0, "material indices must be 0 .. " + std::to_string(nmat - 1));
    c.nmat = nmat;
    for (int m = 0; m < nmat; m++) material(mats->elem(m), c.mat[m]);
    reject_if_present(eqs->find("MaterialTransition"), "Equations.MaterialTransition (phase transitions)");
  }

  void space(const Node *sp, m2c_config &c) {
    const Node *ns = sp ? sp->find("NavierStokes") : nullptr;
    gett(ns, "Flux", c.flux, {{"Roe", M2C_FLUX_ROE}, {"LocalLaxFriedrichs", M2C_FLUX_LLF}, {"HLLC", M2C_FLUX_HLLC},
                              {"Godunov", M2C_FLUX_GODUNOV}});
    getd(ns, "EntropyFixCoefficient", c.roe_delta);
    const Node *rec = ns ? ns->find("Reconstruction") : nullptr;
    gett(rec, "Type", c.rec_linear, {{"Constant", 0}, {"Linear", 1}});
    gett(rec, "Limiter", c.limiter, {{"None", M2C_LIM_NONE}, {"GeneralizedMinMod", M2C_LIM_GENMINMOD},
                                     {"VanAlbada", M2C_LIM_VANALBADA}});
    getd(rec, "GeneralizedMinModCoefficient", c.gminmod_alpha);
    int snz = 1;
    gett(rec, "SlopeNearInterface", snz, {{"Zero", 0}, {"NonZero", 1}});
    c.slope_zero_near_interface = snz == 0;
    int vt = 0;
    gett(rec, "VariableType", vt, {{"Primitive", 0}}, {"Conservative", "PrimitiveCharacteristic", "ConservativeCharacteristic"});
    if (rec) reject_if_present(rec->find("Fixes"), "Space.NavierStokes.Reconstruction.Fixes (first-order regions)");
    if (ns) reject_type(ns->find("Smoothing"), "Space.NavierStokes.Smoothing");

    const Node *lsa = sp ? sp->find("LevelSet") : nullptr;
    if (lsa && lsa->kids.size() > 1) fail(M2C_ERR_UNSUPPORTED, lsa->line, "more than one level set");
    if (lsa && lsa->kids.size() == 1) {
      const Node *ls = lsa->kids.begin()->second.get();
      c.ls_matid = -1;
      geti(ls, "MaterialID", c.ls_matid);
      int solver = 1;
      gett(ls, "Solver", solver, {{"FiniteDifference", 1}}, {"FiniteVolume"});
      int band = INT_MAX;
      geti(ls, "Bandwidth", band);
      if (band != INT_MAX) fail(M2C_ERR_UNSUPPORTED, 0, "narrow-band level sets (Bandwidth) are not provided");
      for (int s = 0; s < 6; s++)
        gett(ls, kBcNames[s], c.ls_bc[s], {{"None", 0}, {"ZeroNeumann", 1}, {"LinearExtrapolation", 2}, {"NonNegative", 3}});
      const Node *ri = ls->find("Reinitialization");
      geti(ri, "Frequency", c.reinit_freq);
      geti(ri, "MaxIts", c.reinit_maxits);
      getd(ri, "CFL", c.reinit_cfl);
      getd(ri, "ConvergenceTolerance", c.reinit_tol);
      int flt = 0;
      gett(ri, "FirstLayerTreatment", flt, {{"Fixed", 0}},
           {"ConstrainedMethod1", "ConstrainedMethod2", "IterativelyConstrainedMethod1", "IterativelyConstrainedMethod2"});
      double tdt = -1.0;
      getd(ri, "TimeInterval", tdt);
      if (tdt > 0.0) fail(M2C_ERR_UNSUPPORTED, 0, "Reinitialization.TimeInterval is not provided (use Frequency)");
    }
  }

  void time(const Node *ts, m2c_config &c, m2c_deck_run &run) {
    int ty = 0;
    gett(ts, "Type", ty, {{"Explicit", 0}}, {"Implicit"});
    getd(ts, "CFL", c.cfl);
    geti(ts, "MaxIts", run.max_its);
    getd(ts, "TimeStep", run.time_step);
    getd(ts, "MaxTime", run.max_time);
    double conv = -1.0;
    getd(ts, "ConvergenceTolerance", conv);
    if (conv > 0.0) fail(M2C_ERR_UNSUPPORTED, 0, "steady-state runs (Time.ConvergenceTolerance) are not provided");
    int lts = 0;
    gett(ts, "LocalTimeStepping", lts, {{"Off", 0}}, {"On"});
    const Node *ex = ts ? ts->find("Explicit") : nullptr;
    gett(ex, "Type", c.ts, {{"ForwardEuler", M2C_TS_FE}, {"RungeKutta2", M2C_TS_RK2}, {"RungeKutta3", M2C_TS_RK3}});
  }

  void multiphase(const Node *mp, m2c_config &c) {
    // the reference numbers Exact 0, Numerical 1, LocalLaxFriedrichs 2; m2c_config numbers them 1, 0, 2
    gett(mp, "Flux", c.mp_flux, {{"Exact", 1}, {"Numerical", 0}, {"LocalLaxFriedrichs", 2}});
    gett(mp, "ReconstructionAtInterface", c.mp_recon, {{"Constant", 0}, {"Linear", 1}});
    int v = 0;
    gett(mp, "RiemannNormal", c.mp_riemann_normal, {{"LevelSet", 0}, {"Mesh", 1}, {"Average", 2}});
    gett(mp, "PhaseChange", v, {{"RiemannSolution", 0}}, {"Extrapolation"});
    gett(mp, "PhaseChangeDirection", v, {{"Upwind", 1}}, {"All"});
    gett(mp, "LatentHeatTransfer", v, {{"Off", 0}}, {"On"});
    gett(mp, "ApplyFailSafeDensity", v, {{"On", 1}}, {"Off"});
    int lcf = -1;
    geti(mp, "LevelSetCorrectionFrequency", lcf);
    if (lcf > 0) fail(M2C_ERR_UNSUPPORTED, 0, "MultiPhase.LevelSetCorrectionFrequency is not provided");
    getd(mp, "ConstantReconstructionDepth", c.mp_conrec_depth);
  }

  void riemann(const Node *er, m2c_riemann_params &r) {
    geti(er, "MaxIts", r.maxIts_main);
    geti(er, "MaxItsBracketing", r.maxIts_bracket);
    geti(er, "MaxItsShock", r.maxIts_shock);
    geti(er, "IntegrationSteps", r.numSteps_rarefaction);
    getd(er, "Tolerance", r.tol_main);
    getd(er, "ToleranceShock", r.tol_shock);
    getd(er, "ToleranceRarefaction", r.tol_rarefaction);
    getd(er, "MinPressure", r.min_pressure);
    getd(er, "FailureThreshold", r.failure_threshold);
    getd(er, "PrescribedPressureUponFailure", r.pressure_at_failure);
  }

  void map(const Node *root, m2c_config &c, m2c_deck_run &run) {
    mesh(root->find("Mesh"), c);
    equations(root->find("Equations"), c);
    space(root->find("Space"), c);
    time(root->find("Time"), c, run);
    multiphase(root->find("MultiPhase"), c);
    riemann(root->find("ExactRiemannSolution"), c.riemann);
    const Node *bc = root->find("BoundaryConditions");
    if (bc) {
      state(bc->find("Inlet"), c.inlet);
      state(bc->find("Outlet"), c.outlet);
    }
    // DefaultInitialState is a top-level object (IoData.cpp:1695)
    state(root->find("DefaultInitialState"), run.initial_state, &run.initial_material);
    if (run.initial_material < 0 || run.initial_material >= c.nmat)
      fail(M2C_ERR_ARG, 0, "DefaultInitialState.MaterialID is not a material");
    const Node *ic = root->find("InitialCondition");
    if (ic) {
      const Value *f = leaf(ic, "UserDataFile");
      if (f && !f->s.empty()) fail(M2C_ERR_UNSUPPORTED, f->line, "InitialCondition.UserDataFile is not provided");
    }
    const Node *geo = root->find("GeometricEntities");
    for (const char *shape : {"Point", "Spheroid", "ArbitraryEnclosure"})
      if (geo) reject_if_present(geo->find(shape), (std::string("GeometricEntities.") + shape).c_str());
    for (const char *physics : {"EmbeddedBoundary", "EmbeddedSurfaces", "Laser", "Ionization", "ReferenceMap",
                                "ConcurrentPrograms", "GeometricEntities2D", "AeroS", "AeroF", "M2CTwin"})
      reject_if_present(root->find(physics), physics);
    ignore_rest(root, "", {"Mesh", "Equations", "Space", "Time", "MultiPhase", "ExactRiemannSolution",
                           "BoundaryConditions", "DefaultInitialState", "InitialCondition", "GeometricEntities",
                           "EmbeddedBoundary", "EmbeddedSurfaces", "Laser", "Ionization", "ReferenceMap",
                           "ConcurrentPrograms", "GeometricEntities2D", "AeroS", "AeroF", "M2CTwin"});
  }
};

// ------------------------------------------------------------------ initial conditions
//
// SpaceOperator::SetInitialCondition (SpaceOperator.cpp:696-960) and LevelSetOperator::SetInitialCondition
// (LevelSetOperator.cpp:291-556) for the shapes of GeometricEntities: the default state everywhere, then
// planes, cylinder-cones, cylinders with spherical caps, spheres and parallelepipeds, each kind in
// increasing index order, each overriding what came before inside it. The level set of material m
// starts at the domain diagonal (Main.cpp:164-166) and takes, shape by shape (only shapes whose
// initial state has material m), the signed distance to the shape wherever that is smaller in
// magnitude (negative inside). Cell centres are x0 + h/2 + i h with h = (xmax - x0)/N
// (MeshGenerator.cpp:146-158).

struct V3 {
  double x, y, z;
};
inline V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scale(double c, const V3 &a) { return {c * a.x, c * a.y, c * a.z}; }
inline double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3 &a, const V3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(const V3 &a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// unit vector as the reference forms it: multiply by the reciprocal of the norm
inline V3 unit(const V3 &a) { return scale(1.0 / norm(a), a); }

// distance from X to segment AB (GeoTools.h:53-89: projection parameter, clamped to the end points)
double seg_dist(const V3 &X, const V3 &A, const V3 &B) {
  const V3 AB = sub(B, A), AX = sub(X, A);
  const double len2 = dot(AB, AB);
  const double alpha = len2 != 0.0 ? dot(AB, AX) / len2 : 0.0;
  if (alpha > 1.0) return norm(sub(X, B));
  if (alpha < 0.0 || !std::isfinite(alpha)) return norm(sub(X, A));
  return norm(sub(add(A, scale(alpha, AB)), X));
}

struct Shape {
  int kind;            // 0 plane, 1 cylinder-cone, 2 cylinder with caps, 3 sphere, 4 parallelepiped
  double V[5];
  int id;
  // plane / cylinders / sphere
  V3 x0{0, 0, 0}, dir{0, 0, 0};
  double radius = 0.0, L = 0.0, tan_alpha = 0.0, Hmax = 0.0, H = 0.0;
  bool front = false, back = false;
  // parallelepiped: corner O and edges a, b, c; faces with outward unit normals
  V3 O{0, 0, 0}, a{0, 0, 0}, b{0, 0, 0}, c{0, 0, 0};
  V3 fo[6], fe1[6], fe2[6], fn[6];

  bool inside(const V3 &p) const {
    if (kind == 0) return dot(sub(p, x0), dir) > 0.0;
    if (kind == 3) return norm(sub(p, x0)) - radius < 0.0;
    if (kind == 4) return box_signed(p) < 0.0;
    const double x = dot(sub(p, x0), dir);
    const double r = norm(sub(sub(p, x0), scale(x, dir)));
    if (kind == 1) return (x > 0 && x < L && r < radius) || (x >= L && x < L + H && r < (L + Hmax - x) * tan_alpha);
    const double Lh = 0.5 * L;
    if (x > -Lh && x < Lh && r < radius) return true;
    if (front && x >= Lh && norm(sub(p, add(x0, scale(Lh, dir)))) < radius) return true;
    if (back && x <= -Lh && norm(sub(p, sub(x0, scale(Lh, dir)))) < radius) return true;
    return false;
  }

  // level-set update at p (LevelSetOperator.cpp:316-556)
  void levelset(const V3 &p, double &phi) const {
    if (kind == 0) {
      const double d = dot(sub(p, x0), dir);
      if (std::fabs(d) < std::fabs(phi)) phi = -d;
      return;
    }
    if (kind == 3) {
      const double d = norm(sub(p, x0)) - radius;
      if (std::fabs(d) < std::fabs(phi)) phi = d;
      return;
    }
    if (kind == 4) {
      const double d = box_signed(p);
      if (std::fabs(d) < std::fabs(phi)) phi = d;
      return;
    }
    // axisymmetric shapes: distance in the (axial, radial) half plane to the generating polyline
    const double x = dot(sub(p, x0), dir);
    const double r = norm(sub(sub(p, x0), scale(x, dir)));
    const V3 q{x, r, 0.0};
    double d;
    if (kind == 1) {
      const V3 pts[5] = {{0, 0, 0}, {0, radius, 0}, {L, radius, 0}, {L + H, (Hmax - H) * tan_alpha, 0}, {L + H, 0, 0}};
      d = 1.7976931348623157e308;
      for (int l = 0; l < 4; l++) d = std::min(d, seg_dist(q, pts[l], pts[l + 1]));
      if (d < std::fabs(phi)) phi = inside(p) ? -d : d;
      return;
    }
    const double Lh = L / 2.0;
    double rf = 0.0, rb = 0.0;
    if (back && x <= -Lh) {
      rb = norm(sub(p, sub(x0, scale(Lh, dir))));
      d = std::fabs(rb - radius);
    } else if (front && x >= Lh) {
      rf = norm(sub(p, add(x0, scale(Lh, dir))));
      d = std::fabs(rf - radius);
    } else {
      d = seg_dist(q, {-Lh, radius, 0}, {Lh, radius, 0});
      if (!back) d = std::min(d, seg_dist(q, {-Lh, 0, 0}, {-Lh, radius, 0}));
      if (!front) d = std::min(d, seg_dist(q, {Lh, radius, 0}, {Lh, 0, 0}));
    }
    if (d < std::fabs(phi)) {
      const bool in = (x > -Lh && x < Lh && r < radius) || (back && x <= -Lh && rb < radius) ||
                      (front && x >= Lh && rf < radius);
      phi = in ? -d : d;
    }
  }

  // signed distance to the parallelepiped (>0 outside; GeoTools/DistancePointToParallelepiped.h:73-110):
  // the smallest distance to its six faces, positive when p lies on the outer side of any face
  // plane. The distance to a face is the exact point-parallelogram distance.
  double box_signed(const V3 &p) const {
    double dmin = 1.7976931348623157e308;
    bool outside = false;
    for (int f = 0; f < 6; f++) {
      const double s = dot(sub(p, fo[f]), fn[f]);
      const double d = para_dist(p, f, s);
      if (d < dmin) dmin = d;
      if ((s >= 0.0 ? d : -d) > 0.0) outside = true;
    }
    return outside ? dmin : -dmin;
  }
  double para_dist(const V3 &p, int f, double s) const {
    // coordinates of the projection in the face's edge basis
    const V3 e1 = fe1[f], e2 = fe2[f];
    const V3 w = sub(sub(p, fo[f]), scale(s, fn[f]));
    const double g11 = dot(e1, e1), g12 = dot(e1, e2), g22 = dot(e2, e2);
    const double r1 = dot(w, e1), r2 = dot(w, e2);
    const double det = g11 * g22 - g12 * g12;
    const double u = (r1 * g22 - r2 * g12) / det, v = (r2 * g11 - r1 * g12) / det;
    if (u >= 0.0 && u <= 1.0 && v >= 0.0 && v <= 1.0) return std::fabs(s);
    const V3 A = fo[f], B = add(A, e1), D = add(A, e2), C = add(B, e2);
    return std::min(std::min(seg_dist(p, A, B), seg_dist(p, B, C)), std::min(seg_dist(p, C, D), seg_dist(p, D, A)));
  }
};

struct IcBuilder {
  std::vector<Shape> shapes;

  static void side_inclusion(const Node *n) {
    int side = 0, inc = 0;
    Mapper::gett(n, "Side", side, {{"Interior", 0}, {"Exterior", 1}});
    Mapper::gett(n, "Inclusion", inc, {{"Override", 0}, {"Intersection", 1}, {"Union", 2}});
    // SpaceOperator.cpp:737, :768-773, :812-817, :867-872, :900-905: interior sides, overriding inclusions
    if (side != 0) fail(M2C_ERR_ARG, n->line, "initialization only supports Side = Interior");
    if (inc != 0) fail(M2C_ERR_ARG, n->line, "initialization only supports Inclusion = Override");
  }
  static void initial_state(const Node *n, Shape &s, int nmat) {
    s.V[0] = 1.0e-6; s.V[1] = s.V[2] = s.V[3] = s.V[4] = 0.0;
    s.id = 0;
    Mapper::state(n->find("InitialState"), s.V, &s.id);
    if (s.id < 0 || s.id >= nmat) fail(M2C_ERR_ARG, n->line, "InitialState.MaterialID is not a material");
  }
  static V3 vec(const Node *n, const char *xn, const char *yn, const char *zn, V3 v) {
    Mapper::getd(n, xn, v.x);
    Mapper::getd(n, yn, v.y);
    Mapper::getd(n, zn, v.z);
    return v;
  }
  // elements of an ObjectMap in increasing index order
  static std::vector<const Node *> elements(const Node *arr) {
    std::vector<std::pair<long long, const Node *>> v;
    if (arr)
      for (auto &kv : arr->kids)
        if (kv.first[0] == '#') v.push_back({std::atoll(kv.first.c_str() + 1), kv.second.get()});
    std::sort(v.begin(), v.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    std::vector<const Node *> out;
    for (auto &e : v) out.push_back(e.second);
    return out;
  }

  void read(const Node *root, int nmat) {
    const Node *geo = root->find("GeometricEntities");
    if (!geo) return;
    for (const Node *n : elements(geo->find("Plane"))) {
      Shape s{};
      s.kind = 0;
      int inc = 0;
      Mapper::gett(n, "Inclusion", inc, {{"Override", 0}, {"Intersection", 1}, {"Union", 2}});
      if (inc != 0) fail(M2C_ERR_ARG, n->line, "initialization only supports Inclusion = Override");
      s.x0 = vec(n, "Point_x", "Point_y", "Point_z", {0, 0, 0});
      s.dir = unit(vec(n, "Normal_x", "Normal_y", "Normal_z", {0, 0, 0}));
      initial_state(n, s, nmat);
      shapes.push_back(s);
    }
    for (const Node *n : elements(geo->find("CylinderAndCone"))) {
      Shape s{};
      s.kind = 1;
      side_inclusion(n);
      s.x0 = vec(n, "BaseCenter_x", "BaseCenter_y", "BaseCenter_z", {0, 0, 0});
      s.dir = unit(vec(n, "Axis_x", "Axis_y", "Axis_z", {0, 0, 1}));
      s.radius = 1.0;
      Mapper::getd(n, "CylinderRadius", s.radius);
      Mapper::getd(n, "CylinderHeight", s.L);
      double angle = 45.0, cone_h = 0.0;
      Mapper::getd(n, "ConeOpeningAngleInDegrees", angle);
      Mapper::getd(n, "ConeHeight", cone_h);
      s.tan_alpha = std::tan(angle / 180.0 * std::acos(-1.0));
      s.Hmax = s.radius / s.tan_alpha;
      s.H = std::min(cone_h, s.Hmax);
      initial_state(n, s, nmat);
      shapes.push_back(s);
    }
    for (const Node *n : elements(geo->find("CylinderWithSphericalCaps"))) {
      Shape s{};
      s.kind = 2;
      side_inclusion(n);
      s.x0 = vec(n, "CylinderCenter_x", "CylinderCenter_y", "CylinderCenter_z", {0, 0, 0});
      s.dir = unit(vec(n, "Axis_x", "Axis_y", "Axis_z", {0, 0, 1}));
      s.radius = 1.0;
      Mapper::getd(n, "CylinderRadius", s.radius);
      Mapper::getd(n, "CylinderHeight", s.L);
      int f = 0, b = 0;
      Mapper::gett(n, "FrontSphericalCap", f, {{"Off", 0}, {"On", 1}});
      Mapper::gett(n, "BackSphericalCap", b, {{"Off", 0}, {"On", 1}});
      s.front = f == 1;
      s.back = b == 1;
      initial_state(n, s, nmat);
      shapes.push_back(s);
    }
    for (const Node *n : elements(geo->find("Sphere"))) {
      Shape s{};
      s.kind = 3;
      side_inclusion(n);
      s.x0 = vec(n, "Center_x", "Center_y", "Center_z", {0, 0, 0});
      s.radius = -1.0;
      Mapper::getd(n, "Radius", s.radius);
      initial_state(n, s, nmat);
      shapes.push_back(s);
    }
    for (const Node *n : elements(geo->find("Parallelepiped"))) {
      Shape s{};
      s.kind = 4;
      side_inclusion(n);
      s.O = vec(n, "X0", "Y0", "Z0", {0, 0, 0});
      s.a = vec(n, "Ax", "Ay", "Az", {0, 0, 0});
      s.b = vec(n, "Bx", "By", "Bz", {0, 0, 0});
      s.c = vec(n, "Cx", "Cy", "Cz", {0, 0, 0});
      // SpaceOperator.cpp:914-917: non-degenerate, right-handed edges
      if (norm(s.a) == 0 || norm(s.b) == 0 || norm(s.c) == 0 || dot(cross(s.a, s.b), s.c) <= 0.0)
        fail(M2C_ERR_ARG, n->line, "detected error in a user-specified parallelepiped");
      const V3 O = s.O, A = add(O, s.a), B = add(O, s.b), C = add(O, s.c);
      const V3 ctr = add(O, scale(0.5, add(add(s.a, s.b), s.c)));
      const V3 o[6] = {O, O, O, A, B, C};
      const V3 e1[6] = {s.a, s.a, s.b, s.b, s.a, s.a};
      const V3 e2[6] = {s.b, s.c, s.c, s.c, s.c, s.b};
      for (int f = 0; f < 6; f++) {
        s.fo[f] = o[f];
        s.fe1[f] = e1[f];
        s.fe2[f] = e2[f];
        V3 nrm = unit(cross(e1[f], e2[f]));
        if (dot(nrm, sub(o[f], ctr)) < 0.0) nrm = scale(-1.0, nrm);
        s.fn[f] = nrm;
      }
      initial_state(n, s, nmat);
      shapes.push_back(s);
    }
  }
};

}  // namespace m2c_deck

using namespace m2c_deck;

extern "C" {

int m2c_config_defaults(m2c_config *c) {
  if (!c) return M2C_ERR_ARG;
  std::memset(c, 0, sizeof(*c));
  c->nx = c->ny = c->nz = -1;       // MeshData: no cells until given; boundaries None
  c->x1 = c->y1 = c->z1 = 1.0;
  for (int s = 0; s < 6; s++) c->ls_bc[s] = 1;   // LevelSetSchemeData: zero Neumann
  c->inlet[0] = c->outlet[0] = 1.0e-6;           // StateVariable: density 1e-6, zero velocity and pressure
  c->nmat = 0;
  c->inactive_id = -1;
  c->flux = M2C_FLUX_HLLC;          // SchemeData
  c->roe_delta = 0.2;
  c->rec_linear = 1;                // ReconstructionData: linear, no limiter, generalized minmod 1.2
  c->limiter = M2C_LIM_NONE;
  c->gminmod_alpha = 1.2;
  c->slope_zero_near_interface = 0;
  c->ts = M2C_TS_RK2;               // ExplicitData
  c->cfl = 0.5;                     // TsData
  c->mp_flux = 0;                   // MultiPhaseData: numerical flux, constant reconstruction
  c->mp_recon = 0;
  m2c_riemann_params &r = c->riemann;   // ExactRiemannSolverData
  r.maxIts_main = 200; r.maxIts_bracket = 100; r.maxIts_shock = 200; r.numSteps_rarefaction = 200;
  r.tol_main = 1e-4; r.tol_shock = 1e-12; r.tol_rarefaction = 1e-8;
  r.min_pressure = -1e8; r.failure_threshold = 0.2; r.pressure_at_failure = 1e-8;
  r.surrogate_residual_tol = 1e-6;
  c->precision = M2C_PREC_FP64;
  c->device = 0;
  c->ls_matid = -1;                 // LevelSetSchemeData / LevelSetReinitializationData
  c->reinit_freq = -1;
  c->reinit_maxits = 30;
  c->reinit_cfl = 0.8;
  c->reinit_tol = 2.0e-4;
  c->mp_riemann_normal = 1;         // MultiPhaseData::riemann_normal = MESH (IoData.cpp:1408)
  c->mp_conrec_depth = 0.0;         // MultiPhaseData::conRec_depth (IoData.cpp:1416)
  return M2C_OK;
}

}  // extern "C"

namespace m2c_deck {

// parses and maps a deck and checks its shapes; returns the status and fills `report` (the error, or
// the ignored options)
int read_deck(const char *text, Node &root, m2c_config &c, m2c_deck_run &r, IcBuilder &ic, std::string &report) {
  m2c_config_defaults(&c);
  std::memset(&r, 0, sizeof(r));
  r.max_its = INT_MAX;              // TsData
  r.time_step = -1.0;
  r.max_time = 1e10;
  r.initial_state[0] = 1.0e-6;      // StateVariable
  r.initial_material = 0;
  try {
    parse(text, root);
    Mapper m;
    m.map(&root, c, r);
    ic.read(&root, c.nmat);
    for (auto &n : m.notes) report += n + "\n";
  } catch (const DeckError &e) {
    report = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    report = e.what();
    return M2C_ERR_ARG;
  }
  return M2C_OK;
}

void put_msg(const std::string &report, char *msg, int msglen) {
  if (msg && msglen > 0) {
    std::strncpy(msg, report.c_str(), (size_t)msglen - 1);
    msg[msglen - 1] = 0;
  }
}

void parse(const char *text, Node &root) {
  Parser p(text, &root);
  p.deck();
}

int read_config(const char *text, Node &root, m2c_config &c, m2c_deck_run &r, std::string &report) {
  IcBuilder ic;
  return read_deck(text, root, c, r, ic, report);
}

const Value *leaf(const Node *n, const char *name) { return Mapper::leaf(n, name); }
void getd(const Node *n, const char *name, double &out) { Mapper::getd(n, name, out); }
void geti(const Node *n, const char *name, int &out) { Mapper::geti(n, name, out); }
void gets(const Node *n, const char *name, std::string &out) {
  const Value *v = leaf(n, name);
  if (!v) return;
  if (v->kind != Value::STR) fail(M2C_ERR_ARG, v->line, std::string(name) + " takes a string");
  out = v->s;
}
void gett(const Node *n, const char *name, int &out, std::initializer_list<std::pair<const char *, int>> opts,
          std::initializer_list<const char *> unsupported) {
  Mapper::gett(n, name, out, opts, unsupported);
}
std::vector<const Node *> elements(const Node *arr) { return IcBuilder::elements(arr); }

}  // namespace m2c_deck

extern "C" {

int m2c_config_from_deck(const char *text, m2c_config *cfg, m2c_deck_run *run, char *msg, int msglen) {
  if (msg && msglen > 0) msg[0] = 0;
  if (!text || !cfg) return M2C_ERR_ARG;
  Node root;
  m2c_config c;
  m2c_deck_run r;
  IcBuilder ic;
  std::string report;
  const int rc = read_deck(text, root, c, r, ic, report);
  put_msg(report, msg, msglen);
  if (rc == M2C_OK) {
    *cfg = c;
    if (run) *run = r;
  }
  return rc;
}

int m2c_initial_condition_from_deck(const char *text, double *V, int32_t *id, double *phi, char *msg, int msglen) {
  if (msg && msglen > 0) msg[0] = 0;
  if (!text || !V || !id) return M2C_ERR_ARG;
  Node root;
  m2c_config c;
  m2c_deck_run r;
  IcBuilder ic;
  std::string report;
  const int rc = read_deck(text, root, c, r, ic, report);
  put_msg(report, msg, msglen);
  if (rc != M2C_OK) return rc;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  // cell centres: MeshGenerator's (x0 + h/2 + i h on a uniform axis)
  std::vector<double> ax[3], aw[3];
  if (int mrc = mesh_axes(c, ax, aw)) return mrc;
  const double diag = std::sqrt(std::pow(c.x1 - c.x0, 2) + std::pow(c.y1 - c.y0, 2) + std::pow(c.z1 - c.z0, 2));
  const bool want_phi = phi && c.ls_matid >= 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const V3 p{ax[0][i], ax[1][j], ax[2][k]};
        const size_t cell = ((size_t)k * ny + j) * nx + i;
        double *v = V + 5 * cell;
        for (int q = 0; q < 5; q++) v[q] = r.initial_state[q];
        int32_t m = r.initial_material;
        double ph = diag;
        for (const Shape &s : ic.shapes) {
          if (s.inside(p)) {
            for (int q = 0; q < 5; q++) v[q] = s.V[q];
            m = s.id;
          }
          if (want_phi && s.id == c.ls_matid) s.levelset(p, ph);
        }
        id[cell] = m;
        if (want_phi) phi[cell] = ph;
      }
  return M2C_OK;
}

}  // extern "C"
