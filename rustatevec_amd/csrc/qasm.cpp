// qasm.cpp -- OpenQASM 3 (and 2.0) reader: program text -> Circuit of dense gate matrices.
//
// The reference names a QASM3 front end as future work of its CLI
// (crates/qsv-cli/src/main.rs:1-2, README.md:63); this is that front end for the engine.
// Supported: the version line; `include "stdgates.inc"` / `"qelib1.inc"` (the standard
// gates are built in); `qubit[n] q;` / `qreg q[n];` (registers are concatenated in
// declaration order, q[0] of the first register is qubit 0 = the least significant index
// bit, and a gate's operands map to its qubits in order, control first, as the gate
// library's convention, crates/qsv-core/src/gate.rs:8-12); classical declarations; `gate`
// definitions (expanded inline, any depth); the modifiers `inv @`, `pow(k) @` (integer k),
// `ctrl(k) @` and `negctrl(k) @`; `gphase`; register broadcast and index ranges; parameter
// expressions (+ - * / % ** ^, pi/tau/euler, sin cos tan exp ln sqrt arcsin arccos
// arctan); `barrier` (ignored) and terminal `measure` (ignored: sample the final state; a
// gate on an already measured qubit is an error).  Anything else (classical control flow,
// reset, timing, calibration) is rejected with the line it starts on.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "ir.hpp"

namespace qsvh {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxDepth = 64;  // nesting of gate definitions

struct ParseError {
  std::string msg;
  int line;
};

// ------------------------------------------------------------------------ lexer
struct Tok {
  enum Kind { kEnd, kId, kNum, kStr, kPunct } kind = kEnd;
  std::string s;
  double v = 0;
  int line = 0;
};

std::vector<Tok> lex(const std::string& src) {
  std::vector<Tok> out;
  size_t i = 0;
  int line = 1;
  const size_t n = src.size();
  auto id_char = [](unsigned char c) { return std::isalnum(c) || c == '_' || c >= 0x80; };
  while (i < n) {
    const unsigned char c = src[i];
    if (c == '\n') {
      ++line;
      ++i;
    } else if (std::isspace(c)) {
      ++i;
    } else if (c == '/' && i + 1 < n && src[i + 1] == '/') {
      while (i < n && src[i] != '\n') ++i;
    } else if (c == '/' && i + 1 < n && src[i + 1] == '*') {
      const int start = line;
      i += 2;
      while (i + 1 < n && !(src[i] == '*' && src[i + 1] == '/')) {
        if (src[i] == '\n') ++line;
        ++i;
      }
      if (i + 1 >= n) throw ParseError{"unterminated comment", start};
      i += 2;
    } else if (std::isdigit(c) ||
               (c == '.' && i + 1 < n && std::isdigit((unsigned char)src[i + 1]))) {
      std::string num;
      while (i < n && (std::isdigit((unsigned char)src[i]) || src[i] == '.' || src[i] == '_')) {
        if (src[i] != '_') num += src[i];
        ++i;
      }
      if (i < n && (src[i] == 'e' || src[i] == 'E')) {
        size_t j = i + 1;
        if (j < n && (src[j] == '+' || src[j] == '-')) ++j;
        if (j < n && std::isdigit((unsigned char)src[j])) {
          num += src.substr(i, j - i);
          i = j;
          while (i < n && std::isdigit((unsigned char)src[i])) num += src[i++];
        }
      }
      if (i < n && id_char((unsigned char)src[i]))
        throw ParseError{"unsupported numeric literal suffix", line};
      char* end = nullptr;
      Tok t{Tok::kNum, num, std::strtod(num.c_str(), &end), line};
      if (!end || *end) throw ParseError{"bad number '" + num + "'", line};
      out.push_back(t);
    } else if (id_char(c) || c == '$') {
      size_t j = i + 1;
      while (j < n && id_char((unsigned char)src[j])) ++j;
      out.push_back({Tok::kId, src.substr(i, j - i), 0, line});
      i = j;
    } else if (c == '"' || c == '\'') {
      size_t j = i + 1;
      while (j < n && src[j] != c && src[j] != '\n') ++j;
      if (j >= n || src[j] != c) throw ParseError{"unterminated string", line};
      out.push_back({Tok::kStr, src.substr(i + 1, j - i - 1), 0, line});
      i = j + 1;
    } else {
      static const char* two[] = {"->", "**", "==", "!=", "<=", ">=", "&&", "||", "+=", "-=",
                                  "++", "<<", ">>"};
      std::string p(1, char(c));
      for (const char* t : two)
        if (i + 1 < n && src[i] == t[0] && src[i + 1] == t[1]) p = t;
      out.push_back({Tok::kPunct, p, 0, line});
      i += p.size();
    }
  }
  out.push_back({Tok::kEnd, "", 0, line});
  return out;
}

// ------------------------------------------------------------------------ matrices
// A gate application before it is pushed: m qubits (m = 0 is a global phase held in
// re[0], im[0]) and a row-major 2^m x 2^m matrix.
struct Gate {
  uint32_t m = 0;
  std::vector<uint32_t> q;
  std::vector<double> re, im;
};

Gate from_op(const Op& op) {
  Gate g;
  g.m = op.m;
  g.re = op.re;
  g.im = op.im;
  return g;
}

Gate mat2(double r0, double i0, double r1, double i1, double r2, double i2, double r3,
          double i3) {
  Gate g;
  g.m = 1;
  g.re = {r0, r1, r2, r3};
  g.im = {i0, i1, i2, i3};
  return g;
}

Gate library(const char* name, std::initializer_list<double> p) {
  Op op;
  std::vector<double> v(p);
  if (v.empty()) v.push_back(0.0);
  make_gate(name, v.data(), uint32_t(p.size()), &op);
  return from_op(op);
}

// Conjugate transpose (a global phase is conjugated).
void dagger(Gate* g) {
  const uint32_t d = 1u << g->m;
  std::vector<double> re(g->re.size()), im(g->im.size());
  for (uint32_t r = 0; r < d; ++r)
    for (uint32_t c = 0; c < d; ++c) {
      re[size_t(c) * d + r] = g->re[size_t(r) * d + c];
      im[size_t(c) * d + r] = -g->im[size_t(r) * d + c];
    }
  g->re.swap(re);
  g->im.swap(im);
}

// Controlled form on `ctl` (low internal index bits, controls first): identity except on
// the rows and columns whose control bits equal `on` (all ones for ctrl, zeros for negctrl),
// where the gate's matrix acts on the remaining bits.  A global phase becomes a diagonal
// on the controls.
Gate control(const Gate& g, const std::vector<uint32_t>& ctl, bool negated) {
  const uint32_t k = uint32_t(ctl.size());
  const uint32_t m = k + g.m;
  Gate o;
  o.m = m;
  o.q = ctl;
  o.q.insert(o.q.end(), g.q.begin(), g.q.end());
  const uint32_t d = 1u << m, dg = 1u << g.m;
  o.re.assign(size_t(d) * d, 0.0);
  o.im.assign(size_t(d) * d, 0.0);
  for (uint32_t i = 0; i < d; ++i) o.re[size_t(i) * d + i] = 1.0;
  const uint32_t on = negated ? 0u : (1u << k) - 1;
  for (uint32_t r = 0; r < dg; ++r)
    for (uint32_t c = 0; c < dg; ++c) {
      const size_t at = size_t(on | (r << k)) * d + (on | (c << k));
      o.re[at] = g.re[size_t(r) * dg + c];
      o.im[at] = g.im[size_t(r) * dg + c];
    }
  return o;
}

// ------------------------------------------------------------------------ parser
struct Reg {
  std::string name;
  uint32_t off = 0, size = 0;
};

struct GateDef {
  std::vector<std::string> params, args;
  size_t begin = 0, end = 0;  // token range of the body (between the braces)
};

struct Modifier {
  enum Kind { kInv, kPow, kCtrl, kNegCtrl } kind;
  double value = 1;  // pow exponent, or the number of controls
};

// Binding of a gate body: parameter values and argument qubits by name.
struct Scope {
  std::map<std::string, double> params;
  std::map<std::string, uint32_t> args;
};

class Parser {
 public:
  explicit Parser(const std::string& src) : t_(lex(src)) {}

  Circuit run() {
    while (peek().kind != Tok::kEnd) statement();
    Circuit c;
    c.n = nq_;
    for (Gate& g : out_) {
      if (g.m == 0) {  // global phase: the same phase on both rows of qubit 0
        if (nq_ == 0) continue;
        Gate d = mat2(g.re[0], g.im[0], 0, 0, 0, 0, g.re[0], g.im[0]);
        d.q = {0};
        g = d;
      }
      const int st = push(&c, g.m, g.q.data(), g.re.data(), g.im.data());
      if (st != kOk) throw ParseError{status_text(st), g_line(g)};
    }
    return c;
  }

 private:
  std::vector<Tok> t_;
  size_t p_ = 0;
  std::vector<Reg> qregs_;
  std::vector<std::string> cregs_;
  std::map<std::string, GateDef> defs_;
  std::vector<bool> measured_;
  uint32_t nq_ = 0;
  std::vector<Gate> out_;
  std::vector<int> out_line_;

  int g_line(const Gate& g) const {
    const size_t i = size_t(&g - out_.data());
    return i < out_line_.size() ? out_line_[i] : 0;
  }
  static std::string status_text(int st) {
    switch (st) {
      case kInvalidQubit: return "qubit index out of range";
      case kRepeatedQubit: return "the same qubit appears twice in one gate";
      case kTooWide: return "gate acts on more than 6 qubits";
      default: return "invalid gate";
    }
  }

  const Tok& peek(size_t k = 0) const { return t_[std::min(p_ + k, t_.size() - 1)]; }
  const Tok& next() {
    const Tok& t = t_[p_];
    if (p_ + 1 < t_.size()) ++p_;
    return t;
  }
  [[noreturn]] void fail(const std::string& msg) const { throw ParseError{msg, peek().line}; }
  bool is(const char* s) const {
    const Tok& t = peek();
    return (t.kind == Tok::kPunct || t.kind == Tok::kId) && t.s == s;
  }
  bool accept(const char* s) {
    if (!is(s)) return false;
    next();
    return true;
  }
  void expect(const char* s) {
    if (!accept(s)) fail(std::string("expected '") + s + "', found '" + peek().s + "'");
  }
  std::string ident() {
    if (peek().kind != Tok::kId) fail("expected a name, found '" + peek().s + "'");
    return next().s;
  }
  void skip_statement() {
    while (peek().kind != Tok::kEnd && !is(";")) next();
    expect(";");
  }

  // -------------------------------------------------------------- expressions
  double expr(const Scope* s) {
    double v = term(s);
    for (;;) {
      if (accept("+")) v += term(s);
      else if (accept("-")) v -= term(s);
      else return v;
    }
  }
  double term(const Scope* s) {
    double v = unary(s);
    for (;;) {
      if (accept("*")) v *= unary(s);
      else if (accept("/")) v /= unary(s);
      else if (accept("%")) v = std::fmod(v, unary(s));
      else return v;
    }
  }
  double unary(const Scope* s) {
    if (accept("-")) return -unary(s);
    if (accept("+")) return unary(s);
    const double b = primary(s);
    if (accept("**") || accept("^")) return std::pow(b, unary(s));
    return b;
  }
  double primary(const Scope* s) {
    const Tok& t = peek();
    if (t.kind == Tok::kNum) return next().v;
    if (accept("(")) {
      const double v = expr(s);
      expect(")");
      return v;
    }
    if (t.kind != Tok::kId) fail("expected an expression, found '" + t.s + "'");
    const std::string name = next().s;
    if (s) {
      auto it = s->params.find(name);
      if (it != s->params.end()) return it->second;
    }
    if (name == "pi" || name == "\xCF\x80") return kPi;            // pi, π
    if (name == "tau" || name == "\xCF\x84") return 2 * kPi;       // tau, τ
    if (name == "euler" || name == "\xE2\x84\x87") return std::exp(1.0);  // euler, ℇ
    static const std::map<std::string, double (*)(double)> fns = {
        {"sin", [](double x) { return std::sin(x); }},
        {"cos", [](double x) { return std::cos(x); }},
        {"tan", [](double x) { return std::tan(x); }},
        {"exp", [](double x) { return std::exp(x); }},
        {"ln", [](double x) { return std::log(x); }},
        {"sqrt", [](double x) { return std::sqrt(x); }},
        {"arcsin", [](double x) { return std::asin(x); }},
        {"arccos", [](double x) { return std::acos(x); }},
        {"arctan", [](double x) { return std::atan(x); }},
        {"asin", [](double x) { return std::asin(x); }},
        {"acos", [](double x) { return std::acos(x); }},
        {"atan", [](double x) { return std::atan(x); }},
    };
    auto f = fns.find(name);
    if (f != fns.end()) {
      expect("(");
      const double v = expr(s);
      expect(")");
      return f->second(v);
    }
    fail("unknown identifier '" + name + "' in an expression");
  }
  uint32_t index_expr(const Scope* s) {
    const double v = expr(s);
    if (!(v >= 0) || v != std::floor(v) || v > 4294967295.0)
      fail("index must be a non-negative integer");
    return uint32_t(v);
  }

  // -------------------------------------------------------------- statements
  void statement() {
    const Tok& t = peek();
    if (t.kind != Tok::kId) {
      if (accept(";")) return;
      fail("unexpected '" + t.s + "'");
    }
    const std::string& w = t.s;
    if (w == "OPENQASM") {
      next();
      if (peek().kind != Tok::kNum) fail("expected a version number");
      const double v = next().v;
      if (v < 2 || v >= 4) fail("unsupported OpenQASM version");
      expect(";");
    } else if (w == "include") {
      next();
      if (peek().kind != Tok::kStr) fail("expected a file name");
      const std::string f = next().s;
      if (f != "stdgates.inc" && f != "qelib1.inc")
        fail("include of '" + f + "': only stdgates.inc and qelib1.inc (built in) are read");
      expect(";");
    } else if (w == "qubit" || w == "qreg") {
      declare_qubits(w == "qreg");
    } else if (w == "bit" || w == "creg") {
      next();
      if (accept("[")) {
        index_expr(nullptr);
        expect("]");
      }
      cregs_.push_back(ident());
      skip_statement();
    } else if (w == "gate") {
      define_gate();
    } else if (w == "barrier") {
      skip_statement();
    } else if (w == "measure") {
      next();
      measure_operands();
    } else if (is_creg(w) && (peek(1).s == "=" || peek(1).s == "[")) {
      next();
      if (accept("[")) {
        index_expr(nullptr);
        expect("]");
      }
      expect("=");
      if (!accept("measure")) fail("only 'c = measure q' assignments are supported");
      measure_operands();
    } else if (w == "reset") {
      fail("reset is not supported (states start in |0...0>)");
    } else if (w == "opaque" || w == "if" || w == "for" || w == "while" || w == "def" ||
               w == "defcal" || w == "cal" || w == "box" || w == "delay" || w == "input" ||
               w == "output" || w == "const" || w == "let" || w == "int" || w == "uint" ||
               w == "float" || w == "angle" || w == "bool" || w == "duration" ||
               w == "stretch" || w == "complex" || w == "array" || w == "extern" ||
               w == "return" || w == "break" || w == "continue" || w == "end" ||
               w == "defcalgrammar" || w == "switch") {
      fail("unsupported statement '" + w + "'");
    } else {
      const int line = t.line;
      std::vector<Gate> gs = call(nullptr, 0);
      for (Gate& g : gs) {
        for (uint32_t q : g.q)
          if (measured_[q])
            throw ParseError{"gate on a measured qubit (mid-circuit measurement is not supported)",
                             line};
        out_.push_back(std::move(g));
        out_line_.push_back(line);
      }
    }
  }

  bool is_creg(const std::string& w) const {
    for (const std::string& c : cregs_)
      if (c == w) return true;
    return false;
  }

  void declare_qubits(bool qasm2) {
    next();
    uint32_t size = 1;
    std::string name;
    if (qasm2) {
      name = ident();
      expect("[");
      size = index_expr(nullptr);
      expect("]");
    } else {
      if (accept("[")) {
        size = index_expr(nullptr);
        expect("]");
      }
      name = ident();
    }
    expect(";");
    for (const Reg& r : qregs_)
      if (r.name == name) fail("register '" + name + "' declared twice");
    if (size == 0 || uint64_t(nq_) + size > 64) fail("register size out of range");
    qregs_.push_back({name, nq_, size});
    nq_ += size;
    measured_.resize(nq_, false);
  }

  void define_gate() {
    next();
    const std::string name = ident();
    if (name == "U" || name == "CX" || name == "gphase") fail("'" + name + "' is built in");
    GateDef d;
    if (accept("(")) {
      if (!is(")")) {
        d.params.push_back(ident());
        while (accept(",")) d.params.push_back(ident());
      }
      expect(")");
    }
    d.args.push_back(ident());
    while (accept(",")) d.args.push_back(ident());
    expect("{");
    d.begin = p_;
    int depth = 1;
    while (peek().kind != Tok::kEnd) {
      if (is("{")) ++depth;
      if (is("}") && --depth == 0) break;
      next();
    }
    d.end = p_;
    expect("}");
    defs_[name] = d;
  }

  // Qubits of one top-level operand: q (whole register), q[i] or q[a:b] / q[a:s:b].
  std::vector<uint32_t> operand() {
    const std::string name = ident();
    const Reg* reg = nullptr;
    for (const Reg& r : qregs_)
      if (r.name == name) reg = &r;
    if (!reg) fail("unknown qubit register '" + name + "'");
    std::vector<uint32_t> qs;
    if (accept("[")) {
      const uint32_t a = index_expr(nullptr);
      if (accept(":")) {
        uint32_t step = 1, b = index_expr(nullptr);
        if (accept(":")) {
          step = b;
          b = index_expr(nullptr);
        }
        if (step == 0) fail("zero range step");
        for (uint32_t i = a; i <= b; i += step) qs.push_back(i);
      } else {
        qs.push_back(a);
      }
      expect("]");
      for (uint32_t i : qs)
        if (i >= reg->size) fail("index " + std::to_string(i) + " out of range for '" + name + "'");
      for (uint32_t& i : qs) i += reg->off;
    } else {
      for (uint32_t i = 0; i < reg->size; ++i) qs.push_back(reg->off + i);
    }
    return qs;
  }

  void measure_operands() {
    for (uint32_t q : operand()) measured_[q] = true;
    if (accept("->")) {
      ident();
      if (accept("[")) {
        index_expr(nullptr);
        expect("]");
      }
    }
    expect(";");
  }

  // One gate call statement (at top level when scope is null, else inside a gate body):
  // modifiers, name, parameters, operands.  Returns the gate applications it expands to.
  std::vector<Gate> call(const Scope* scope, int depth) {
    if (depth > kMaxDepth) fail("gate definitions nested too deeply (recursive?)");
    std::vector<Modifier> mods;
    for (;;) {
      if (is("inv") && peek(1).s == "@") {
        next();
        next();
        mods.push_back({Modifier::kInv, 1});
      } else if ((is("ctrl") || is("negctrl")) && (peek(1).s == "@" || peek(1).s == "(")) {
        const bool neg = next().s == "negctrl";
        double k = 1;
        if (accept("(")) {
          k = index_expr(scope);
          expect(")");
        }
        expect("@");
        if (k < 1) fail("ctrl needs at least one control");
        mods.push_back({neg ? Modifier::kNegCtrl : Modifier::kCtrl, k});
      } else if (is("pow") && peek(1).s == "(") {
        next();
        next();
        const double k = expr(scope);
        expect(")");
        expect("@");
        if (k != std::floor(k) || std::fabs(k) > 1e6) fail("pow supports integer exponents only");
        mods.push_back({Modifier::kPow, k});
      } else {
        break;
      }
    }
    const int line = peek().line;
    const std::string name = ident();
    std::vector<double> params;
    if (accept("(")) {
      if (!is(")")) {
        params.push_back(expr(scope));
        while (accept(",")) params.push_back(expr(scope));
      }
      expect(")");
    }
    std::vector<std::vector<uint32_t>> ops;
    if (!is(";")) {
      ops.push_back(scoped_operand(scope));
      while (accept(",")) ops.push_back(scoped_operand(scope));
    }
    expect(";");
    // broadcast over whole-register operands (all of one size)
    size_t width = 1;
    for (const auto& o : ops)
      if (o.size() != 1) {
        if (width != 1 && o.size() != width)
          throw ParseError{"register operands differ in size", line};
        width = o.size();
      }
    std::vector<Gate> res;
    for (size_t i = 0; i < width; ++i) {
      std::vector<uint32_t> qs;
      for (const auto& o : ops) qs.push_back(o.size() == 1 ? o[0] : o[i]);
      std::vector<Gate> gs = apply(name, params, qs, mods, depth, line);
      res.insert(res.end(), gs.begin(), gs.end());
    }
    return res;
  }

  std::vector<uint32_t> scoped_operand(const Scope* scope) {
    if (!scope) return operand();
    const std::string a = ident();
    auto it = scope->args.find(a);
    if (it == scope->args.end()) fail("unknown gate argument '" + a + "'");
    if (is("[")) fail("gate arguments are single qubits and take no index");
    return {it->second};
  }

  // Expand a named gate with its modifiers (mods[0] outermost) on operands qs.
  std::vector<Gate> apply(const std::string& name, const std::vector<double>& params,
                          const std::vector<uint32_t>& qs, const std::vector<Modifier>& mods,
                          int depth, int line) {
    // controls are taken from the front of the operand list, outermost modifier first
    std::vector<std::vector<uint32_t>> ctl(mods.size());
    size_t pos = 0;
    for (size_t i = 0; i < mods.size(); ++i)
      if (mods[i].kind == Modifier::kCtrl || mods[i].kind == Modifier::kNegCtrl) {
        const size_t k = size_t(mods[i].value);
        if (pos + k > qs.size()) throw ParseError{"too few operands for the controls", line};
        ctl[i].assign(qs.begin() + pos, qs.begin() + pos + k);
        pos += k;
      }
    const std::vector<uint32_t> args(qs.begin() + pos, qs.end());
    std::vector<Gate> gs = base(name, params, args, depth, line);
    for (size_t i = mods.size(); i-- > 0;) {
      const Modifier& md = mods[i];
      if (md.kind == Modifier::kInv) {
        std::vector<Gate> r(gs.rbegin(), gs.rend());
        for (Gate& g : r) dagger(&g);
        gs.swap(r);
      } else if (md.kind == Modifier::kPow) {
        if (md.value < 0) {
          std::vector<Gate> r(gs.rbegin(), gs.rend());
          for (Gate& g : r) dagger(&g);
          gs.swap(r);
        }
        std::vector<Gate> r;
        for (long k = 0; k < std::labs(long(md.value)); ++k)
          r.insert(r.end(), gs.begin(), gs.end());
        gs.swap(r);
      } else {
        for (Gate& g : gs) {
          if (ctl[i].size() + g.m > kMaxGateQubits)
            throw ParseError{"controlled gate acts on more than 6 qubits", line};
          g = control(g, ctl[i], md.kind == Modifier::kNegCtrl);
        }
      }
    }
    return gs;
  }

  // A gate without modifiers: a definition from the program, else the built-in library.
  std::vector<Gate> base(const std::string& name, const std::vector<double>& p,
                         const std::vector<uint32_t>& qs, int depth, int line) {
    auto d = defs_.find(name);
    if (d != defs_.end()) {
      const GateDef& def = d->second;
      if (p.size() != def.params.size())
        throw ParseError{
            "'" + name + "' takes " + std::to_string(def.params.size()) + " parameters", line};
      if (qs.size() != def.args.size())
        throw ParseError{"'" + name + "' takes " + std::to_string(def.args.size()) + " qubits",
                         line};
      Scope s;
      for (size_t i = 0; i < p.size(); ++i) s.params[def.params[i]] = p[i];
      for (size_t i = 0; i < qs.size(); ++i) s.args[def.args[i]] = qs[i];
      const size_t saved = p_;
      p_ = def.begin;
      std::vector<Gate> out;
      while (p_ < def.end) {
        if (accept(";")) continue;
        if (is("barrier")) {
          skip_statement();
          continue;
        }
        std::vector<Gate> gs = call(&s, depth + 1);
        out.insert(out.end(), gs.begin(), gs.end());
      }
      p_ = saved;
      return out;
    }
    Gate g;
    uint32_t arity = 0, np = 0;
    if (!builtin(name, p, &g, &arity, &np)) throw ParseError{"unknown gate '" + name + "'", line};
    if (p.size() != np)
      throw ParseError{"'" + name + "' takes " + std::to_string(np) + " parameters", line};
    if (qs.size() != arity)
      throw ParseError{"'" + name + "' takes " + std::to_string(arity) + " qubits", line};
    if (name == "id" || name == "i") return {};
    g.q = qs;
    return {g};
  }

  // The standard gates: the gate library's matrices where it has the gate, otherwise the
  // textbook matrix (the same conventions: rz(t) = diag(e^{-it/2}, e^{it/2}), u3 = U).
  static bool builtin(const std::string& n, const std::vector<double>& p, Gate* g,
                      uint32_t* arity, uint32_t* np) {
    auto P = [&](size_t i) { return i < p.size() ? p[i] : 0.0; };
    static const std::map<std::string, std::pair<uint32_t, uint32_t>> kShape = {
        {"h", {1, 0}},     {"x", {1, 0}},     {"y", {1, 0}},      {"z", {1, 0}},
        {"s", {1, 0}},     {"t", {1, 0}},     {"sx", {1, 0}},     {"sdg", {1, 0}},
        {"tdg", {1, 0}},   {"sxdg", {1, 0}},  {"id", {1, 0}},     {"i", {1, 0}},
        {"rx", {1, 1}},    {"ry", {1, 1}},    {"rz", {1, 1}},     {"p", {1, 1}},
        {"phase", {1, 1}}, {"u1", {1, 1}},    {"u2", {1, 2}},     {"u3", {1, 3}},
        {"u", {1, 3}},     {"U", {1, 3}},     {"gphase", {0, 1}}, {"cx", {2, 0}},
        {"CX", {2, 0}},    {"cnot", {2, 0}},  {"cy", {2, 0}},     {"cz", {2, 0}},
        {"ch", {2, 0}},    {"csx", {2, 0}},   {"swap", {2, 0}},   {"cp", {2, 1}},
        {"cphase", {2, 1}}, {"cu1", {2, 1}},  {"crx", {2, 1}},    {"cry", {2, 1}},
        {"crz", {2, 1}},   {"rzz", {2, 1}},   {"rxx", {2, 1}},    {"ryy", {2, 1}},
        {"cu3", {2, 3}},   {"cu", {2, 4}},    {"ccx", {3, 0}},    {"toffoli", {3, 0}},
        {"ccz", {3, 0}},   {"cswap", {3, 0}}, {"c3x", {4, 0}},    {"c4x", {5, 0}},
    };
    auto it = kShape.find(n);
    if (it == kShape.end()) return false;
    *arity = it->second.first;
    *np = it->second.second;
    if (p.size() != *np) return true;  // the caller reports the count
    constexpr double r = 0.70710678118654752440;
    auto ctl1 = [](const Gate& u, uint32_t k) {
      std::vector<uint32_t> c(k);
      for (uint32_t i = 0; i < k; ++i) c[i] = i;
      Gate o = control(u, c, false);
      o.q.clear();
      return o;
    };
    if (n == "h" || n == "x" || n == "y" || n == "z" || n == "s" || n == "t" || n == "sx" ||
        n == "rx" || n == "ry" || n == "rz" || n == "p" || n == "phase" || n == "u3" ||
        n == "cx" || n == "cnot" || n == "cz" || n == "cp" || n == "cphase" || n == "swap" ||
        n == "rzz" || n == "ccx" || n == "toffoli" || n == "ccz") {
      Op op;
      std::vector<double> v(p);
      v.resize(3, 0.0);
      make_gate(n, v.data(), *np, &op);
      *g = from_op(op);
    } else if (n == "U" || n == "u") {
      *g = library("u3", {P(0), P(1), P(2)});
    } else if (n == "CX") {
      *g = library("cx", {});
    } else if (n == "id" || n == "i") {
      *g = mat2(1, 0, 0, 0, 0, 0, 1, 0);
    } else if (n == "u1") {
      *g = library("p", {P(0)});
    } else if (n == "u2") {
      *g = library("u3", {kPi / 2, P(0), P(1)});
    } else if (n == "sdg") {
      *g = mat2(1, 0, 0, 0, 0, 0, 0, -1);
    } else if (n == "tdg") {
      *g = mat2(1, 0, 0, 0, 0, 0, r, -r);
    } else if (n == "sxdg") {
      *g = mat2(0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5, -0.5);
    } else if (n == "gphase") {
      g->m = 0;
      g->re = {std::cos(P(0))};
      g->im = {std::sin(P(0))};
    } else if (n == "cy") {
      *g = ctl1(library("y", {}), 1);
    } else if (n == "ch") {
      *g = ctl1(library("h", {}), 1);
    } else if (n == "csx") {
      *g = ctl1(library("sx", {}), 1);
    } else if (n == "cu1") {
      *g = library("cp", {P(0)});
    } else if (n == "crx" || n == "cry" || n == "crz") {
      *g = ctl1(library(n.c_str() + 1, {P(0)}), 1);
    } else if (n == "cu3") {
      *g = ctl1(library("u3", {P(0), P(1), P(2)}), 1);
    } else if (n == "cu") {  // controlled e^{i gamma} U(theta, phi, lambda)
      Gate u = library("u3", {P(0), P(1), P(2)});
      const double gr = std::cos(P(3)), gi = std::sin(P(3));
      for (size_t e = 0; e < 4; ++e) {
        const double a = u.re[e], b = u.im[e];
        u.re[e] = gr * a - gi * b;
        u.im[e] = gr * b + gi * a;
      }
      *g = ctl1(u, 1);
    } else if (n == "rxx" || n == "ryy") {  // exp(-i t/2 X(x)X), exp(-i t/2 Y(x)Y)
      const double c = std::cos(P(0) / 2), s = std::sin(P(0) / 2);
      g->m = 2;
      g->re.assign(16, 0.0);
      g->im.assign(16, 0.0);
      for (int d = 0; d < 4; ++d) g->re[d * 5] = c;
      const double corner = n == "rxx" ? -s : s;  // |00><11| and |11><00|
      g->im[0 * 4 + 3] = corner;
      g->im[3 * 4 + 0] = corner;
      g->im[1 * 4 + 2] = -s;
      g->im[2 * 4 + 1] = -s;
    } else if (n == "cswap") {
      *g = ctl1(library("swap", {}), 1);
    } else if (n == "c3x") {
      *g = ctl1(library("x", {}), 3);
    } else if (n == "c4x") {
      *g = ctl1(library("x", {}), 4);
    }
    return true;
  }
};

}  // namespace

int parse_qasm(const std::string& text, Circuit* out, std::string* err) {
  try {
    Parser p(text);
    *out = p.run();
    return kOk;
  } catch (const ParseError& e) {
    if (err) *err = "line " + std::to_string(e.line) + ": " + e.msg;
    return kParse;
  }
}

}  // namespace qsvh
