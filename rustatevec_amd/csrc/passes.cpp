// passes.cpp -- learned choice of the execution form per pass.
//
// fusion::fuse (crates/qsv-core/src/fusion.rs:39-77) fixes one width for the whole circuit
// and the engine's other forms (exact tile passes, MFMA tile passes) are likewise chosen for
// the whole circuit.  Here the circuit is first cut into passes -- each a set of gates whose
// non-diagonal gates act on the qubits of one LDS tile of 2^12 amplitudes (qubits 0..3 and
// up to 8 more), diagonal gates on any qubits, a gate moving ahead only past gates it
// commutes with -- and then every pass gets its own form: exact (gate by gate in the tile,
// the state's own precision) or an MFMA tile pass of its gates fused greedily up to k = 3, 4
// or 5 qubits (the fuse rule restricted to the pass: a group stays diagonal or inside the
// tile; dense groups are split-BF16 matrix-core products).  The exact-pass and MFMA-pass
// cost-model networks price every form of every pass from its lowered ops, and the plan
// keeps the cheapest form of each pass -- or, when that mixed plan is not predicted clearly
// faster than the best plan of one form for every pass, that one-form plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <new>
#include <vector>

#include "../../include/qsvh.h"
#include "handles.hpp"
#include "mlp.hpp"
#include "program.hpp"

namespace qsvh {

Launch classify(const Op& op, int dtype);  // api.cpp
void tile_pass_features(const TileBuilder& b, const TilePassArg& pa, uint32_t n, int dtype,
                        float* f);                     // planner.cpp
void mt_pass_features(const MtPlan& plan, size_t i, uint32_t n, float* f);  // planner.cpp
int build_partition(Ctx* ctx, const Circuit& c,
                    const std::vector<std::pair<uint32_t, uint32_t>>& parts, Circuit* out);

namespace {

constexpr uint32_t kPassTileBits = 12;
constexpr uint64_t kLowMask = 15;  // qubits 0..3: 16-amplitude runs in every tile
constexpr size_t kPassLookahead = 1024;
constexpr int kTileF = 18, kMtF = 11;
constexpr uint32_t kFuseMin = 3, kFuseMax = 5;
constexpr double kMixMargin = 0.03;

uint32_t popc(uint64_t x) { return uint32_t(__builtin_popcountll(x)); }

uint64_t qubits_of(const Op& op) {
  uint64_t m = 0;
  for (uint32_t j = 0; j < op.m; ++j) m |= 1ull << op.q[j];
  return m;
}

// Phase-table entries a diagonal gate may take in an MFMA pass once fused (a fused diagonal
// group of up to kFuseMax qubits holds 2^kFuseMax entries): the pass budget of kMtTabEntries
// then holds whatever the fusion makes of its diagonal gates.
uint32_t diag_entries(const Op& op) { return 1u << std::max(op.m, kFuseMax); }

struct PassSet {
  uint64_t tmask = 0;            // the tile: kPassTileBits state bits, 0..3 among them
  std::vector<uint64_t> idx;     // the pass's gates, in the order it applies them
  bool exact_ok = true;          // every dense gate fits the exact tile lowering (m <= 3)
};

// Cut the circuit into passes.  A pass takes gates in circuit order: a diagonal gate when it
// shares no qubit with a dense gate left behind and the phase budget allows, a dense gate
// when it shares no qubit with any gate left behind and its qubits with the tile's so far fit
// kPassTileBits; the scan gives up after kPassLookahead gates left behind.  Gates moved ahead
// commute with everything they pass, so the product is unchanged.
std::vector<PassSet> schedule_pass_sets(const Circuit& c, const std::vector<char>& diag) {
  const size_t L = c.ops.size();
  const uint64_t all = c.n >= 64 ? ~0ull : (1ull << c.n) - 1;
  std::vector<char> taken(L, 0);
  std::vector<PassSet> out;
  size_t first = 0;
  while (first < L) {
    PassSet ps;
    uint64_t T = kLowMask, left_dense = 0, left_diag = 0;
    uint32_t tab = 0;
    size_t left = 0;
    for (size_t i = first; i < L && left_dense != all && left < kPassLookahead; ++i) {
      if (taken[i]) continue;
      const Op& op = c.ops[i];
      const uint64_t Q = qubits_of(op);
      bool take;
      if (diag[i])
        take = !(Q & left_dense) && tab + diag_entries(op) <= kMtTabEntries;
      else
        take = !(Q & (left_dense | left_diag)) && popc(T | Q) <= kPassTileBits;
      if (!take) {
        (diag[i] ? left_diag : left_dense) |= Q;
        ++left;
        continue;
      }
      if (diag[i])
        tab += diag_entries(op);
      else
        T |= Q;
      if (!diag[i] && op.m > 3) ps.exact_ok = false;
      ps.idx.push_back(i);
      taken[i] = 1;
    }
    for (uint32_t b = 0; b < c.n && popc(T) < kPassTileBits; ++b) T |= 1ull << b;
    ps.tmask = T;
    out.push_back(std::move(ps));
    while (first < L && taken[first]) ++first;
  }
  return out;
}

// Greedy consecutive groups of the pass's gates (positions in `sub`) up to k qubits, as
// fusion::fuse groups (fusion.rs:39-77), with one more rule: a group that is not diagonal
// keeps all its qubits in the tile (the MFMA pass applies dense groups inside the tile only).
void pass_groups(const Circuit& sub, const std::vector<char>& diag, uint64_t tmask, uint32_t k,
                 uint32_t offset, std::vector<std::pair<uint32_t, uint32_t>>* parts) {
  uint32_t begin = 0;
  uint64_t Q = 0;
  bool d = true, open = false;
  auto close = [&](uint32_t end) {
    if (open) parts->emplace_back(offset + begin, offset + end);
    open = false;
  };
  for (uint32_t j = 0; j < sub.ops.size(); ++j) {
    const Op& op = sub.ops[j];
    const uint64_t q = qubits_of(op);
    if (op.m > k) {
      close(j);
      parts->emplace_back(offset + j, offset + j + 1);
      continue;
    }
    if (open) {
      const uint64_t u = Q | q;
      const bool nd = d && diag[j];
      if (popc(u) <= k && (nd || (u & ~tmask) == 0)) {
        Q = u;
        d = nd;
        continue;
      }
      close(j);
    }
    begin = j;
    Q = q;
    d = diag[j];
    open = true;
  }
  close(uint32_t(sub.ops.size()));
}

struct PassForms {
  bool exact_ok = false;
  float exact_f[kTileF] = {0};
  bool mt_ok[kFuseMax + 1] = {false};
  float mt_f[kFuseMax + 1][kMtF] = {};
  std::vector<Op> fused[kFuseMax + 1];
};

}  // namespace

// Exact lowering of one pass (launches of dtype `dtype`) and its feature row.
static void exact_features(const std::vector<Op>& ops, uint32_t n, uint64_t tmask, int dtype,
                           float* f) {
  std::vector<Launch> in;
  in.reserve(ops.size());
  for (const Op& op : ops) in.push_back(classify(op, dtype));
  std::vector<uint64_t> take(in.size());
  for (size_t i = 0; i < take.size(); ++i) take[i] = i;
  TileBuilder bld;
  std::vector<TilePassArg> passes;
  std::vector<Launch> out;
  std::vector<uint64_t> order;
  lower_exact_pass(in, take, n, kPassTileBits, tmask, bld, passes, out, order);
  tile_pass_features(bld, passes.back(), n, dtype, f);
}

// MFMA lowering of one pass's fused ops and its feature row; false when its phase table
// exceeds the pass budget.
static bool mt_features(const std::vector<Op>& ops, uint32_t n, uint64_t tmask, float* f) {
  std::vector<Launch> in;
  in.reserve(ops.size());
  uint32_t entries = 0;
  for (const Op& op : ops) {
    in.push_back(classify(op, 1));
    if (in.back().kind == Launch::kDiag) entries += 1u << op.m;
  }
  if (entries > kMtTabEntries) return false;
  std::vector<size_t> cur(in.size());
  for (size_t i = 0; i < cur.size(); ++i) cur[i] = i;
  MtPlan plan;
  emit_mt_pass(in, cur, n, kPassTileBits, tmask, &plan);
  mt_pass_features(plan, 0, n, f);
  return true;
}

// form: -1 learned (both models), 0 every pass exact, k in 3..5 every pass MFMA at width k
// (a pass the form cannot run falls back to the other forms, cheapest first when models
// are given, else exact, else the narrowest valid width).
int plan_passes(Ctx* ctx, Mlp* tile_mlp, Mlp* mt_mlp, const Circuit& c, int form, Circuit* out,
                double* predicted_ms) {
  const uint32_t n = c.n;
  std::vector<char> diag(c.ops.size());
  for (size_t i = 0; i < c.ops.size(); ++i) diag[i] = c.ops[i].is_diagonal() ? 1 : 0;
  const std::vector<PassSet> sets = schedule_pass_sets(c, diag);
  const size_t P = sets.size();
  std::vector<PassForms> forms(P);
  // the gates in pass order, and each pass's position in that list
  Circuit ordered;
  ordered.n = n;
  std::vector<uint32_t> at(P + 1, 0);
  std::vector<char> odiag;
  for (size_t p = 0; p < P; ++p) {
    at[p] = uint32_t(ordered.ops.size());
    for (uint64_t i : sets[p].idx) {
      ordered.ops.push_back(c.ops[i]);
      odiag.push_back(diag[i]);
    }
  }
  at[P] = uint32_t(ordered.ops.size());
  // fuse every pass at every width (one device composition per width)
  for (uint32_t k = kFuseMin; k <= kFuseMax; ++k) {
    std::vector<std::pair<uint32_t, uint32_t>> parts;
    std::vector<size_t> first_part(P + 1, 0);
    for (size_t p = 0; p < P; ++p) {
      first_part[p] = parts.size();
      Circuit sub;
      sub.n = n;
      sub.ops.assign(ordered.ops.begin() + at[p], ordered.ops.begin() + at[p + 1]);
      const std::vector<char> sd(odiag.begin() + at[p], odiag.begin() + at[p + 1]);
      pass_groups(sub, sd, sets[p].tmask, k, at[p], &parts);
    }
    first_part[P] = parts.size();
    Circuit fused;
    const int rc = build_partition(ctx, ordered, parts, &fused);
    if (rc != kOk) return rc;
    for (size_t p = 0; p < P; ++p)
      forms[p].fused[k].assign(fused.ops.begin() + first_part[p],
                               fused.ops.begin() + first_part[p + 1]);
  }
  // features of every form of every pass
  for (size_t p = 0; p < P; ++p) {
    PassForms& pf = forms[p];
    pf.exact_ok = sets[p].exact_ok;
    if (pf.exact_ok) {
      const std::vector<Op> ops(ordered.ops.begin() + at[p], ordered.ops.begin() + at[p + 1]);
      exact_features(ops, n, sets[p].tmask, 1, pf.exact_f);
    }
    for (uint32_t k = kFuseMin; k <= kFuseMax; ++k)
      pf.mt_ok[k] = mt_features(pf.fused[k], n, sets[p].tmask, pf.mt_f[k]);
  }
  // predicted milliseconds of every form (infinite where a form cannot run the pass)
  const double inf = std::numeric_limits<double>::infinity();
  std::vector<double> t_exact(P, inf), t_mt((kFuseMax + 1) * P, inf);
  const double scale = std::exp2(double(n) - 30.0);
  if (tile_mlp && mt_mlp) {
    std::vector<float> xt, xm, yt, ym;
    std::vector<size_t> rt, rm;
    for (size_t p = 0; p < P; ++p) {
      if (forms[p].exact_ok) {
        xt.insert(xt.end(), forms[p].exact_f, forms[p].exact_f + kTileF);
        rt.push_back(p);
      }
      for (uint32_t k = kFuseMin; k <= kFuseMax; ++k)
        if (forms[p].mt_ok[k]) {
          xm.insert(xm.end(), forms[p].mt_f[k], forms[p].mt_f[k] + kMtF);
          rm.push_back(k * P + p);
        }
    }
    yt.resize(rt.size());
    ym.resize(rm.size());
    if (!rt.empty()) {
      const int rc = tile_mlp->predict(xt.data(), rt.size(), yt.data(), false);
      if (rc != kOk) return rc;
    }
    if (!rm.empty()) {
      const int rc = mt_mlp->predict(xm.data(), rm.size(), ym.data(), false);
      if (rc != kOk) return rc;
    }
    for (size_t r = 0; r < rt.size(); ++r) t_exact[rt[r]] = std::exp2(double(yt[r])) * scale;
    for (size_t r = 0; r < rm.size(); ++r) t_mt[rm[r]] = std::exp2(double(ym[r])) * scale;
  }
  // A learned plan that mixes forms must be predicted faster than the best one-form plan by
  // kMixMargin: consecutive passes of different kernels ran 1-4% slower back to back than
  // the sum of their times measured one by one (random mix at 24 and 28 qubits), which the
  // per-pass models do not see.  Otherwise every pass takes that one form (passes it cannot
  // run keep their cheapest form).
  if (form < 0 && tile_mlp && mt_mlp) {
    auto pass_t = [&](size_t p, int f) {
      return f == 0 ? t_exact[p] : t_mt[size_t(f) * P + p];
    };
    auto cheapest = [&](size_t p) {
      double b = t_exact[p];
      for (uint32_t k = kFuseMin; k <= kFuseMax; ++k) b = std::min(b, t_mt[k * P + p]);
      return b;
    };
    double mixed = 0, best_uniform = inf;
    int best_f = -1;
    for (size_t p = 0; p < P; ++p) mixed += cheapest(p);
    for (int f = 0; f <= int(kFuseMax); ++f) {
      if (f > 0 && f < int(kFuseMin)) continue;
      double tot = 0;
      for (size_t p = 0; p < P; ++p) {
        const bool ok = f == 0 ? forms[p].exact_ok : forms[p].mt_ok[f];
        tot += ok ? pass_t(p, f) : cheapest(p);
      }
      if (tot < best_uniform) {
        best_uniform = tot;
        best_f = f;
      }
    }
    if (best_f >= 0 && !(mixed < (1.0 - kMixMargin) * best_uniform)) form = best_f;
  }
  // the form of every pass
  out->n = n;
  out->ops.clear();
  out->plan.clear();
  double total = 0;
  for (size_t p = 0; p < P; ++p) {
    const PassForms& pf = forms[p];
    int pick = -1;
    if (form == 0 && pf.exact_ok) pick = 0;
    if (form >= int(kFuseMin) && form <= int(kFuseMax) && pf.mt_ok[form]) pick = form;
    if (pick < 0) {  // learned, or the forced form cannot run this pass
      double best = inf;
      if (pf.exact_ok && t_exact[p] < best) {
        best = t_exact[p];
        pick = 0;
      }
      for (uint32_t k = kFuseMin; k <= kFuseMax; ++k)
        if (pf.mt_ok[k] && t_mt[k * P + p] < best) {
          best = t_mt[k * P + p];
          pick = int(k);
        }
      if (pick < 0) {  // no predictions: exact, else the narrowest width that runs it
        if (pf.exact_ok) pick = 0;
        for (uint32_t k = kFuseMin; k <= kFuseMax && pick < 0; ++k)
          if (pf.mt_ok[k]) pick = int(k);
      }
      if (pick < 0) return set_err(ctx, kInvalidArg, "plan_passes: no form runs a pass");
    }
    total += pick == 0 ? t_exact[p] : t_mt[pick * P + p];
    PlanPass pp;
    pp.form = pick;
    pp.op_begin = out->ops.size();
    pp.tile_mask = sets[p].tmask;
    if (pick == 0)
      out->ops.insert(out->ops.end(), ordered.ops.begin() + at[p], ordered.ops.begin() + at[p + 1]);
    else
      out->ops.insert(out->ops.end(), pf.fused[pick].begin(), pf.fused[pick].end());
    pp.n_ops = out->ops.size() - pp.op_begin;
    out->plan.push_back(pp);
  }
  if (predicted_ms) *predicted_ms = (tile_mlp && mt_mlp) ? total : 0.0;
  return kOk;
}

// Feature row of pass i of a planned circuit for its own form (kTileF floats for an exact
// pass, kMtF for an MFMA pass).
int planned_pass_features(const Circuit& c, size_t i, float* f) {
  const PlanPass& pp = c.plan[i];
  const std::vector<Op> ops(c.ops.begin() + pp.op_begin, c.ops.begin() + pp.op_begin + pp.n_ops);
  if (pp.form == 0) {
    exact_features(ops, c.n, pp.tile_mask, 1, f);
    return kOk;
  }
  return mt_features(ops, c.n, pp.tile_mask, f) ? kOk : kInvalidArg;
}

// The launches of a planned circuit's passes (Program::launches holds one classified launch
// per op on entry): exact passes lowered into the program's tile passes, MFMA passes into its
// MFMA tile passes, both uploaded.
int plan_planned_passes(Program* p, const Circuit& c) {
  State& st = *p->st;
  if (c.plan.empty()) return set_err(st.ctx, kInvalidArg, "the circuit has no pass plan");
  uint64_t covered = 0;
  bool any_mt = false;
  for (const PlanPass& pp : c.plan) {
    if (pp.op_begin != covered) return set_err(st.ctx, kInvalidArg, "pass plan out of order");
    covered += pp.n_ops;
    any_mt = any_mt || pp.form > 0;
    if (popc(pp.tile_mask) != kPassTileBits || (pp.tile_mask & kLowMask) != kLowMask)
      return set_err(st.ctx, kInvalidArg, "pass plan: bad tile");
  }
  if (covered != c.ops.size())
    return set_err(st.ctx, kInvalidArg, "pass plan does not cover the circuit (ops added?)");
  if (st.dtype != 1 || st.n < kPassTileBits)
    return set_err(st.ctx, kDtype, "planned passes run on complex64 states of >= 12 qubits");
  if (any_mt && p->precision != QSVH_PREC_MFMA_BF16X3)
    return set_err(st.ctx, kInvalidArg, "MFMA passes run split-BF16 groups (MFMA_BF16X3)");
  const std::vector<Launch> in = std::move(p->launches);
  p->launches.clear();
  p->order.clear();
  TileBuilder bld;
  MtPlan mt;
  for (const PlanPass& pp : c.plan) {
    if (pp.n_ops == 0) continue;
    if (pp.form == 0) {
      std::vector<uint64_t> take(pp.n_ops);
      for (uint64_t j = 0; j < pp.n_ops; ++j) {
        take[j] = pp.op_begin + j;
        const Launch& l = in[take[j]];
        if (l.kind != Launch::kDiag && l.m > 3)
          return set_err(st.ctx, kInvalidArg, "exact pass with a gate wider than 3 qubits");
        uint64_t qm = 0;
        for (uint32_t q = 0; q < l.m; ++q) qm |= 1ull << l.q[q];
        if (l.kind != Launch::kDiag && (qm & ~pp.tile_mask))
          return set_err(st.ctx, kInvalidArg, "exact pass gate outside its tile");
      }
      lower_exact_pass(in, take, st.n, kPassTileBits, pp.tile_mask, bld, p->tile_passes,
                       p->launches, p->order);
    } else {
      std::vector<size_t> cur(pp.n_ops);
      uint32_t entries = 0;
      for (uint64_t j = 0; j < pp.n_ops; ++j) {
        cur[j] = size_t(pp.op_begin + j);
        const Launch& l = in[cur[j]];
        uint64_t qm = 0;
        for (uint32_t q = 0; q < l.m; ++q) qm |= 1ull << l.q[q];
        if (l.kind == Launch::kDiag)
          entries += 1u << l.m;
        else if (qm & ~pp.tile_mask)
          return set_err(st.ctx, kInvalidArg, "MFMA pass dense group outside its tile");
      }
      if (entries > kMtTabEntries)
        return set_err(st.ctx, kInvalidArg, "MFMA pass phase table over 2048 entries");
      const size_t before = mt.order.size();
      emit_mt_pass(in, cur, st.n, kPassTileBits, pp.tile_mask, &mt);
      p->order.insert(p->order.end(), mt.order.begin() + before, mt.order.end());
      p->launches.push_back(mt.launches.back());
    }
  }
  p->mt_passes = std::move(mt.passes);
  auto upload = [&](const void* src, size_t bytes, void** dst) {
    hipError_t e = hipMalloc(dst, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_err(st.ctx, e, "pass buffers alloc");
    p->owned.push_back(*dst);
    if (bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    return hip_err(st.ctx, e, "pass buffers upload");
  };
  int rc = kOk;
  if (!bld.ops.empty()) {
    rc = upload(bld.ops.data(), bld.ops.size() * sizeof(TileOp),
                reinterpret_cast<void**>(&p->tile_ops));
    if (rc == kOk)
      rc = upload(bld.diags.data(), bld.diags.size() * sizeof(TileDiag),
                  reinterpret_cast<void**>(&p->tile_diags));
    if (rc == kOk) {
      const std::vector<float> f(bld.pool.begin(), bld.pool.end());
      rc = upload(f.data(), f.size() * sizeof(float), &p->tile_pool);
    }
  }
  if (rc == kOk && !mt.ops.empty()) {
    void* d = nullptr;
    rc = upload(mt.ops.data(), mt.ops.size() * sizeof(MtOp), &d);
    p->mt_ops = static_cast<MtOp*>(d);
    if (rc == kOk) rc = upload(mt.pool.data(), mt.pool.size(), &p->mt_pool);
  }
  return rc;
}

}  // namespace qsvh

using namespace qsvh;

extern "C" {

int qsvh_plan_passes(qsvh_ctx* h, qsvh_mlp* tile_model, qsvh_mlp* mt_model,
                     const qsvh_circuit* in, int form, qsvh_circuit** out,
                     double* predicted_ms) {
  if (!h || !in || !out) return kInvalidArg;
  *out = nullptr;
  if (form != -1 && form != 0 && (form < int(kFuseMin) || form > int(kFuseMax)))
    return set_err(&h->c, kInvalidArg, "plan_passes: form is -1 (learned), 0 or 3..5");
  if (form == -1 && (!tile_model || !mt_model))
    return set_err(&h->c, kInvalidArg, "plan_passes: the learned form needs both models");
  if (tile_model && tile_model->m.F != kTileF)
    return set_err(&h->c, kInvalidArg, "plan_passes: tile model has 18 inputs");
  if (mt_model && mt_model->m.F != kMtF)
    return set_err(&h->c, kInvalidArg, "plan_passes: MFMA-tile model has 11 inputs");
  if (in->c.n < kPassTileBits)
    return set_err(&h->c, kInvalidArg, "plan_passes: circuits of at least 12 qubits");
  qsvh_circuit* res = new (std::nothrow) qsvh_circuit;
  if (!res) return kOom;
  const int rc = plan_passes(&h->c, tile_model ? &tile_model->m : nullptr,
                             mt_model ? &mt_model->m : nullptr, in->c, form, &res->c,
                             predicted_ms);
  if (rc != kOk) {
    delete res;
    return rc;
  }
  *out = res;
  return kOk;
}

int qsvh_circuit_plan_len(const qsvh_circuit* c, size_t* n_passes) {
  if (!c || !n_passes) return kInvalidArg;
  *n_passes = c->c.plan.size();
  return kOk;
}

int qsvh_circuit_plan_pass(const qsvh_circuit* c, size_t i, int* form, uint64_t* op_begin,
                           uint64_t* n_ops, uint64_t* tile_mask, float* features) {
  if (!c || i >= c->c.plan.size()) return kInvalidArg;
  const PlanPass& pp = c->c.plan[i];
  if (form) *form = pp.form;
  if (op_begin) *op_begin = pp.op_begin;
  if (n_ops) *n_ops = pp.n_ops;
  if (tile_mask) *tile_mask = pp.tile_mask;
  if (features) {
    if (pp.op_begin + pp.n_ops > c->c.ops.size()) return kInvalidArg;
    return planned_pass_features(c->c, i, features);
  }
  return kOk;
}

}  // extern "C"
