// ir.cpp -- gate library, circuit validation, generators and greedy fusion grouping.
#include "ir.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace qsvh {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtHalf = 0.70710678118654752440;

// sin and cos of one angle through glibc's sincos: fixes the last bit independently of
// whether the compiler would emit separate sin/cos calls or merge them.
inline void sc(double x, double* s, double* c) { ::sincos(x, s, c); }

// A 1-qubit matrix given as four complex entries (row-major).
void set2x2(Op* op, double r0, double i0, double r1, double i1, double r2, double i2,
            double r3, double i3) {
  op->m = 1;
  op->re = {r0, r1, r2, r3};
  op->im = {i0, i1, i2, i3};
}

// controlled_1q (crates/qsv-core/src/gate.rs:72-87): internal index = bit_c + 2*bit_t;
// identity on the control-0 rows, `u` on the control-1 rows (odd indices).
void controlled(const Op& u, Op* op) {
  op->m = 2;
  op->re.assign(16, 0.0);
  op->im.assign(16, 0.0);
  op->re[0] = 1.0;
  op->re[10] = 1.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int idx = (1 + 2 * a) * 4 + (1 + 2 * b);
      op->re[idx] = u.re[a * 2 + b];
      op->im[idx] = u.im[a * 2 + b];
    }
}

// Diagonal matrix from a list of complex phases.
void diagonal(uint32_t m, const double* dr, const double* di, Op* op) {
  const uint32_t d = 1u << m;
  op->m = m;
  op->re.assign(size_t(d) * d, 0.0);
  op->im.assign(size_t(d) * d, 0.0);
  for (uint32_t i = 0; i < d; ++i) {
    op->re[size_t(i) * d + i] = dr[i];
    op->im[size_t(i) * d + i] = di[i];
  }
}

struct GateDef {
  const char* name;
  uint32_t arity;
  uint32_t n_params;
  void (*build)(const double* p, Op* op);
};

void g_h(const double*, Op* o) { set2x2(o, kSqrtHalf, 0, kSqrtHalf, 0, kSqrtHalf, 0, -kSqrtHalf, 0); }
void g_x(const double*, Op* o) { set2x2(o, 0, 0, 1, 0, 1, 0, 0, 0); }
void g_y(const double*, Op* o) { set2x2(o, 0, 0, 0, -1, 0, 1, 0, 0); }
void g_z(const double*, Op* o) { set2x2(o, 1, 0, 0, 0, 0, 0, -1, 0); }
void g_s(const double*, Op* o) { set2x2(o, 1, 0, 0, 0, 0, 0, 0, 1); }
void g_t(const double*, Op* o) { set2x2(o, 1, 0, 0, 0, 0, 0, kSqrtHalf, kSqrtHalf); }
void g_sx(const double*, Op* o) { set2x2(o, 0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5); }
void g_rx(const double* p, Op* o) {
  double s, c;
  sc(p[0] * 0.5, &s, &c);
  set2x2(o, c, 0, 0, -s, 0, -s, c, 0);
}
void g_ry(const double* p, Op* o) {
  double s, c;
  sc(p[0] * 0.5, &s, &c);
  set2x2(o, c, 0, -s, 0, s, 0, c, 0);
}
void g_rz(const double* p, Op* o) {
  double s, c;
  sc(p[0] * 0.5, &s, &c);
  set2x2(o, c, -s, 0, 0, 0, 0, c, s);
}
void g_p(const double* p, Op* o) {
  double s, c;
  sc(p[0], &s, &c);
  set2x2(o, 1, 0, 0, 0, 0, 0, c, s);
}
// OpenQASM 3 U(theta, phi, lambda) =
//   [[cos(t/2), -e^{i l} sin(t/2)], [e^{i p} sin(t/2), e^{i (p + l)} cos(t/2)]].
void g_u3(const double* p, Op* o) {
  double st, ct, sp, cp, sl, cl, spl, cpl;
  sc(p[0] * 0.5, &st, &ct);
  sc(p[1], &sp, &cp);
  sc(p[2], &sl, &cl);
  sc(p[1] + p[2], &spl, &cpl);
  set2x2(o, ct, 0, -cl * st, -sl * st, cp * st, sp * st, cpl * ct, spl * ct);
}
void g_cx(const double* p, Op* o) { Op u; g_x(p, &u); controlled(u, o); }
void g_cz(const double* p, Op* o) { Op u; g_z(p, &u); controlled(u, o); }
void g_cp(const double* p, Op* o) { Op u; g_p(p, &u); controlled(u, o); }
void g_swap(const double*, Op* o) {
  o->m = 2;
  o->re.assign(16, 0.0);
  o->im.assign(16, 0.0);
  o->re[0] = o->re[6] = o->re[9] = o->re[15] = 1.0;  // (0,0) (1,2) (2,1) (3,3)
}
// RZZ(theta) = exp(-i theta/2 Z(x)Z): phase e^{-i theta/2 * (+-1)} by the parity of the
// two internal bits (crates/qsv-core/src/gate.rs:170-180).
void g_rzz(const double* p, Op* o) {
  double dr[4], di[4];
  for (int idx = 0; idx < 4; ++idx) {
    const double zz = (((idx & 1) ^ ((idx >> 1) & 1)) == 0) ? 1.0 : -1.0;
    const double a = -0.5 * p[0] * zz;
    sc(a, &di[idx], &dr[idx]);
  }
  diagonal(2, dr, di, o);
}
// Toffoli on [c0, c1, t] (internal index c0 + 2 c1 + 4 t): exchanges indices 3 and 7
// (crates/qsv-core/tests/equivalence.rs:182-193).
void g_ccx(const double*, Op* o) {
  o->m = 3;
  o->re.assign(64, 0.0);
  o->im.assign(64, 0.0);
  for (int i = 0; i < 8; ++i) {
    const int col = (i == 3) ? 7 : (i == 7) ? 3 : i;
    o->re[i * 8 + col] = 1.0;
  }
}
void g_ccz(const double*, Op* o) {
  double dr[8], di[8];
  for (int i = 0; i < 8; ++i) {
    dr[i] = (i == 7) ? -1.0 : 1.0;
    di[i] = 0.0;
  }
  diagonal(3, dr, di, o);
}

const GateDef kGates[] = {
    {"h", 1, 0, g_h},       {"x", 1, 0, g_x},       {"y", 1, 0, g_y},
    {"z", 1, 0, g_z},       {"s", 1, 0, g_s},       {"t", 1, 0, g_t},
    {"sx", 1, 0, g_sx},     {"rx", 1, 1, g_rx},     {"ry", 1, 1, g_ry},
    {"rz", 1, 1, g_rz},     {"p", 1, 1, g_p},       {"phase", 1, 1, g_p},
    {"u3", 1, 3, g_u3},     {"cx", 2, 0, g_cx},     {"cnot", 2, 0, g_cx},
    {"cz", 2, 0, g_cz},     {"cp", 2, 1, g_cp},     {"cphase", 2, 1, g_cp},
    {"swap", 2, 0, g_swap}, {"rzz", 2, 1, g_rzz},   {"ccx", 3, 0, g_ccx},
    {"toffoli", 3, 0, g_ccx}, {"ccz", 3, 0, g_ccz},
};

const GateDef* find_gate(const std::string& name) {
  for (const GateDef& g : kGates)
    if (name == g.name) return &g;
  return nullptr;
}

// Small helper for the generators: push a library gate on up to three qubits.
void add(Circuit* c, const char* name, uint32_t a, uint32_t b = 0, uint32_t d = 0,
         double p0 = 0, double p1 = 0, double p2 = 0) {
  const uint32_t q[3] = {a, b, d};
  const double p[3] = {p0, p1, p2};
  push_gate(c, name, q, p, 3);
}

}  // namespace

bool Op::is_diagonal() const {
  const uint32_t d = dim();
  for (uint32_t r = 0; r < d; ++r)
    for (uint32_t col = 0; col < d; ++col)
      if (r != col && (re[size_t(r) * d + col] != 0.0 || im[size_t(r) * d + col] != 0.0))
        return false;
  return true;
}

int validate_qubits(uint32_t n, uint32_t m, const uint32_t* q) {
  if (m == 0) return kArity;
  if (m > kMaxGateQubits) return kTooWide;
  if (!q) return kInvalidArg;
  for (uint32_t i = 0; i < m; ++i) {
    if (q[i] >= n) return kInvalidQubit;
    for (uint32_t j = i + 1; j < m; ++j)
      if (q[i] == q[j]) return kRepeatedQubit;
  }
  return kOk;
}

int make_gate(const std::string& name, const double* params, uint32_t n_params, Op* op) {
  const GateDef* g = find_gate(name);
  if (!g) return kInvalidArg;
  if (n_params < g->n_params || (g->n_params > 0 && !params)) return kInvalidArg;
  double p[3] = {0, 0, 0};
  for (uint32_t i = 0; i < g->n_params; ++i) p[i] = params[i];
  g->build(p, op);
  return kOk;
}

int push(Circuit* c, uint32_t m, const uint32_t* q, const double* re, const double* im) {
  if (!c || !re || !im) return kInvalidArg;
  const int st = validate_qubits(c->n, m, q);
  if (st != kOk) return st;
  Op op;
  op.m = m;
  std::copy(q, q + m, op.q);
  const size_t n = size_t(1) << (2 * m);
  op.re.assign(re, re + n);
  op.im.assign(im, im + n);
  c->ops.push_back(std::move(op));
  return kOk;
}

int push_gate(Circuit* c, const char* name, const uint32_t* q, const double* params,
              uint32_t n_params) {
  if (!c || !name) return kInvalidArg;
  Op op;
  const int st = make_gate(name, params, n_params, &op);
  if (st != kOk) return st;
  const int v = validate_qubits(c->n, op.m, q);
  if (v != kOk) return v;
  std::copy(q, q + op.m, op.q);
  c->ops.push_back(std::move(op));
  return kOk;
}

Circuit gen_ghz(uint32_t n) {
  Circuit c;
  c.n = n;
  add(&c, "h", 0);
  for (uint32_t t = 1; t < n; ++t) add(&c, "cx", 0, t);
  return c;
}

Circuit gen_ghz_chain(uint32_t n) {
  Circuit c;
  c.n = n;
  add(&c, "h", 0);
  for (uint32_t i = 0; i + 1 < n; ++i) add(&c, "cx", i, i + 1);
  return c;
}

Circuit gen_qft(uint32_t n) {
  Circuit c;
  c.n = n;
  for (uint32_t j = 0; j < n; ++j) {
    add(&c, "h", j);
    for (uint32_t k = j + 1; k < n; ++k)
      add(&c, "cp", k, j, 0, kPi / static_cast<double>(1ull << (k - j)));
  }
  for (uint32_t j = 0; j < n / 2; ++j) add(&c, "swap", j, n - 1 - j);
  return c;
}

Circuit gen_random(uint32_t n, uint64_t n_gates, uint64_t seed) {
  static const char* kFixed[7] = {"h", "x", "y", "z", "s", "t", "sx"};
  static const char* kRot[4] = {"rx", "ry", "rz", "p"};
  static const char* kTwo[3] = {"cx", "cz", "swap"};
  Circuit c;
  c.n = n;
  if (n < 2) return c;
  SplitMix64 rng(seed);
  for (uint64_t g = 0; g < n_gates; ++g) {
    const uint32_t a = static_cast<uint32_t>(rng.below(n));
    const uint64_t pick = rng.below(15);
    if (pick < 7) {
      add(&c, kFixed[pick], a);
    } else if (pick < 11) {
      add(&c, kRot[pick - 7], a, 0, 0, rng.angle());
    } else {
      uint32_t b = static_cast<uint32_t>(rng.below(n));
      while (b == a) b = static_cast<uint32_t>(rng.below(n));
      const uint64_t w = rng.below(4);
      if (w < 3)
        add(&c, kTwo[w], a, b);
      else
        add(&c, "rzz", a, b, 0, rng.angle());
    }
  }
  return c;
}

Circuit gen_layered(uint32_t n, uint32_t depth, uint64_t seed) {
  Circuit c;
  c.n = n;
  SplitMix64 rng(seed);
  for (uint32_t l = 0; l < depth; ++l) {
    for (uint32_t j = 0; j < n; ++j) {
      const double t = rng.angle(), f = rng.angle(), la = rng.angle();
      add(&c, "u3", j, 0, 0, t, f, la);
    }
    for (uint32_t j = 0; j < n; ++j) add(&c, "rz", j, 0, 0, rng.angle());
    for (uint32_t j = l & 1u; j + 1 < n; j += 2) add(&c, "cz", j, j + 1);
  }
  return c;
}

Circuit gen_hea(uint32_t n, uint32_t depth, uint64_t seed) {
  Circuit c;
  c.n = n;
  SplitMix64 rng(seed);
  for (uint32_t l = 0; l < depth; ++l) {
    for (uint32_t j = 0; j < n; ++j) add(&c, "ry", j, 0, 0, rng.angle());
    for (uint32_t j = 0; j < n; ++j) add(&c, "rz", j, 0, 0, rng.angle());
    for (uint32_t j = 0; j + 1 < n; ++j) add(&c, "cx", j, j + 1);
  }
  return c;
}

Circuit gen_mix(uint32_t n, uint64_t n_gates, uint64_t seed) {
  Circuit c;
  c.n = n;
  if (n < 3) return c;
  SplitMix64 rng(seed);
  auto other = [&](uint32_t a, uint32_t b, bool two) {
    for (;;) {
      const uint32_t x = static_cast<uint32_t>(rng.below(n));
      if (x != a && (!two || x != b)) return x;
    }
  };
  for (uint64_t g = 0; g < n_gates; ++g) {
    const uint64_t r = rng.below(100);
    const uint32_t a = static_cast<uint32_t>(rng.below(n));
    if (r < 50) {
      const double t = rng.angle(), f = rng.angle(), la = rng.angle();
      add(&c, "u3", a, 0, 0, t, f, la);
    } else if (r < 80) {
      const uint32_t b = other(a, 0, false);
      add(&c, rng.below(2) == 0 ? "cx" : "cz", a, b);
    } else if (r < 95) {
      const uint32_t b = other(a, 0, false);
      add(&c, "rzz", a, b, 0, rng.angle());
    } else {
      const uint32_t b = other(a, 0, false);
      const uint32_t d = other(a, b, true);
      add(&c, rng.below(2) == 0 ? "ccx" : "ccz", a, b, d);
    }
  }
  return c;
}

std::vector<Group> greedy_groups(const Circuit& c, uint32_t max_qubits) {
  const uint32_t cap = std::max<uint32_t>(max_qubits, 1);
  std::vector<Group> out;
  Group cur;
  bool open = false;
  auto close = [&]() {
    if (!open) return;
    if (cur.end - cur.begin > 1) std::sort(cur.q, cur.q + cur.m);
    out.push_back(cur);
    open = false;
  };
  auto start = [&](uint64_t i) {
    const Op& op = c.ops[i];
    cur = Group();
    cur.begin = i;
    cur.end = i + 1;
    cur.m = op.m;
    std::copy(op.q, op.q + op.m, cur.q);
    open = true;
  };
  for (uint64_t i = 0; i < c.ops.size(); ++i) {
    const Op& op = c.ops[i];
    if (op.m > cap) {  // wider than the cap: emitted standalone (fusion.rs:49-54)
      close();
      start(i);
      close();
      continue;
    }
    if (!open) {
      start(i);
      continue;
    }
    // Union of the group's qubits with this op's, in first-seen order.
    uint32_t uni[2 * kMaxGateQubits];
    uint32_t nu = cur.m;
    std::copy(cur.q, cur.q + cur.m, uni);
    for (uint32_t j = 0; j < op.m; ++j) {
      if (std::find(uni, uni + nu, op.q[j]) == uni + nu) uni[nu++] = op.q[j];
    }
    if (nu <= cap) {
      cur.end = i + 1;
      cur.m = nu;
      std::copy(uni, uni + nu, cur.q);
    } else {
      close();
      start(i);
    }
  }
  close();
  return out;
}

}  // namespace qsvh
