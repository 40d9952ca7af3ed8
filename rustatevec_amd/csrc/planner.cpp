// planner.cpp -- learned fusion planning.
//
// fusion::fuse (crates/qsv-core/src/fusion.rs:39-77) grows each group greedily up to a
// fixed max_qubits.  Here the boundaries and the width of every group are chosen by
// dynamic programming over the circuit's contiguous partitions: every run of consecutive
// gates whose qubit union has <= max_qubits qubits (and <= `window` gates) is a candidate
// group; its cost is the device time predicted by the cost-model network from the group's
// features (width, qubit positions, diagonal flag, density of the composed matrix, state
// size), all candidates being evaluated in one batched forward pass on the device; the
// plan minimises the predicted total.  The greedy plans for every max_qubits are scored by
// the same model and kept if they predict lower (greedy runs can exceed the window).
// plan_form then prices the other execution forms with their own networks -- exact tile
// passes and greedy fusion run as MFMA tile passes -- and keeps the cheapest.
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

constexpr int kFeatures = 12;
constexpr int kTileFeatures = 18;
// The pass models see the state size up to 28 qubits: complex64 states from 2^27 amplitudes
// up live in HBM beyond the 256 MiB infinity cache, where a pass's time per amplitude no
// longer depends on n, and the models are trained on 24..28-qubit states -- a larger n
// would only ask them to extrapolate.
constexpr uint32_t kPassFeatureMaxN = 28;

// Feature vector of one tile pass (the ops schedule_tiles lowered for it): the state's
// dtype, tile bits and size (n capped at kPassFeatureMaxN), then counts (/16) of diagonal gates by lowering variant
// (phase uniform per tile, indexed by lane bits, by register bits only, m > 2), of lane-run
// 1-qubit members (lane or register target; general or cheap pair variant), register
// blocks and their members, LDS ops (standalone 1-qubit and dense 2-3 qubit), permutation
// ops, lane runs, and members with per-amplitude conditions.
void tile_pass_features(const TileBuilder& b, const TilePassArg& pa, uint32_t n, int dtype,
                        float* f) {
  float cnt[kTileFeatures] = {0};
  auto diag = [&](const TileDiag& d) {
    const uint32_t k0 = d.code & 15u;  // the lowering variant's first index bit
    cnt[d.m > 2 ? 6 : k0 == 0 ? 3 : k0 == 1 ? 4 : 5] += 1.f;
  };
  const uint32_t end = pa.op_begin + pa.n_ops;
  for (uint32_t o = pa.op_begin; o < end; ++o) {
    const TileOp& op = b.ops[o];
    if (op.kind == 4 || op.kind == 3) {
      cnt[op.kind == 4 ? 16 : 11] += 1.f;
      for (uint32_t u = 1; u <= op.n_diag; ++u) {
        const TileOp& g = b.ops[o + u];
        if (g.kind == 1) {
          for (uint32_t d = 0; d < g.n_diag; ++d) diag(b.diags[g.off + d]);
          continue;
        }
        if (g.kind == 5) {  // controlled phases: lane- or register-bit index variants
          cnt[g.lq[0] >= kTileLaneBits ? 5 : 4] += float(g.n_diag);
          continue;
        }
        if (g.kind == 6) {  // phases on two in-tile bits: register-only or with a lane bit
          for (uint32_t d = 0; d < g.n_diag; ++d)
            cnt[g.lq[0] >= kTileLaneBits && b.diags[g.off + d].oq >= kTileLaneBits ? 5 : 4] += 1.f;
          continue;
        }
        const bool cheap = g.real != 0;
        if (op.kind == 3)
          cnt[cheap ? 13 : 12] += 1.f;
        else
          cnt[g.lq[0] >= kTileLaneBits ? (cheap ? 10 : 9) : (cheap ? 8 : 7)] += 1.f;
        if (op.kind == 4 && g.insm[0] != 0) cnt[17] += 1.f;
      }
      o += op.n_diag;
    } else if (op.kind == 1) {
      for (uint32_t d = 0; d < op.n_diag; ++d) diag(b.diags[op.off + d]);
    } else {
      cnt[op.kind == 0 ? 14 : 15] += 1.f;
    }
  }
  f[0] = float(dtype);
  f[1] = float(pa.tb) / 13.f;
  f[2] = float(std::min(n, kPassFeatureMaxN)) / 32.f;
  for (int i = 3; i < kTileFeatures; ++i) f[i] = cnt[i] / 16.f;
}

// The tile schedule of circuit c (dtype) at tb tile bits: the feature rows of its passes
// and the launches it leaves outside passes.
void tile_schedule_features(const Circuit& c, int dtype, uint32_t tb, std::vector<float>* rows,
                            std::vector<Launch>* other) {
  std::vector<Launch> in;
  in.reserve(c.ops.size());
  for (const Op& op : c.ops) in.push_back(classify(op, dtype));
  std::vector<Launch> perm;  // the trailing qubit permutation, run after the passes
  strip_perm_suffix(in, c.n, &perm);
  TileBuilder bld;
  std::vector<TilePassArg> passes;
  std::vector<Launch> out;
  std::vector<uint64_t> order;
  schedule_tiles(in, c.n, tb, bld, passes, out, order);
  rows->assign(passes.size() * kTileFeatures, 0.f);
  for (size_t i = 0; i < passes.size(); ++i)
    tile_pass_features(bld, passes[i], c.n, dtype, rows->data() + i * kTileFeatures);
  other->clear();
  for (Launch& l : out)
    if (l.kind != Launch::kTile) other->push_back(std::move(l));
  for (Launch& l : perm) other->push_back(std::move(l));
}

constexpr int kMtFeatures = 11;

// Feature vector of one MFMA tile pass (complex64, 2^12-amplitude tiles): counts of its
// dense groups by width (4, 5, 6; /16), single diagonal ops (/32), merged diagonal runs
// (/16) and their members (/64), the entries of the per-tile tables the merged runs build
// and of the pass's phase table (/2048), flags for a 5- or 6-qubit widest group (the
// kernel's occupancy steps) and the state size n/32 (n capped at kPassFeatureMaxN).
void mt_pass_features(const MtPlan& plan, size_t i, uint32_t n, float* f) {
  const MtPassArg& pa = plan.passes[i];
  std::fill(f, f + kMtFeatures, 0.f);
  for (uint32_t j = pa.op_begin; j < pa.op_begin + pa.n_ops; ++j) {
    const MtOp& o = plan.ops[j];
    if (o.kind == 0) {
      f[o.k - 4] += 1.f / 16;
    } else if (o.kind == 1) {
      f[3] += 1.f / 32;
    } else if (o.kind == 2) {
      f[4] += 1.f / 16;
      f[6] += float(1u << o.k) / 2048;
    } else {
      f[5] += 1.f / 64;
    }
  }
  f[7] = float(pa.tab_entries) / 2048;
  f[8] = pa.kmax == 5 ? 1.f : 0.f;
  f[9] = pa.kmax == 6 ? 1.f : 0.f;
  f[10] = float(std::min(n, kPassFeatureMaxN)) / 32.f;
}

void mt_schedule_features(const Circuit& c, std::vector<float>* rows, std::vector<Launch>* other) {
  std::vector<Launch> in;
  in.reserve(c.ops.size());
  for (const Op& op : c.ops) in.push_back(classify(op, 1));
  std::vector<Launch> perm;  // the trailing qubit permutation, placed as the program places it
  std::vector<uint32_t> dest;
  strip_perm_suffix(in, c.n, &perm, &dest);
  uint64_t k = in.size();
  if (!perm.empty())
    k = hoist_perm(in, dest, [&](const std::vector<Launch>& v) {
      return mt_launch_count(v, c.n);
    });
  MtPlan plan;
  build_mt_plan(std::vector<Launch>(in.begin(), in.begin() + k), c.n, 12, &plan);
  build_mt_plan(std::vector<Launch>(in.begin() + k, in.end()), c.n, 12, &plan);
  rows->assign(plan.passes.size() * kMtFeatures, 0.f);
  for (size_t i = 0; i < plan.passes.size(); ++i)
    mt_pass_features(plan, i, c.n, rows->data() + i * kMtFeatures);
  other->clear();
  for (Launch& l : plan.launches)
    if (l.kind != Launch::kMfmaTile) other->push_back(std::move(l));
  for (Launch& l : perm) other->push_back(std::move(l));
}

void group_features(uint32_t n, uint32_t k, const uint32_t* q, bool diag, double density,
                    float* f);

// Group-model features of a launch no pass holds; a bit-swap pass (one read and one write
// of every amplitude) is priced as a dense 1-qubit gate on the top qubit.
static void other_features(const Launch& l, uint32_t n, float* f) {
  if (l.kind == Launch::kBitSwap) {
    const uint32_t top = n ? n - 1 : 0;
    group_features(n, 1, &top, false, 1.0, f);
    return;
  }
  uint32_t q[kMaxGateQubits];
  std::copy(l.q, l.q + l.m, q);
  std::sort(q, q + l.m);
  group_features(n, l.m, q, l.kind == Launch::kDiag, 1.0, f);
}

// Feature vector of a group: k qubits (ascending), diagonal flag, density of the composed
// matrix (nonzeros / 4^k) and the state size n.
void group_features(uint32_t n, uint32_t k, const uint32_t* q, bool diag, double density,
                    float* f) {
  const float fn = float(std::max<uint32_t>(n, 1));
  f[0] = float(k) / 6.f;
  f[1] = float(n) / 32.f;
  f[2] = diag ? 1.f : 0.f;
  f[3] = float(density);
  for (uint32_t i = 0; i < 6; ++i) f[4 + i] = i < k ? float(q[i] + 1) / fn : 0.f;
  f[10] = k ? float(std::min<uint32_t>(q[0], 8)) / 8.f : 0.f;  // lowest qubit (coalescing)
  f[11] = (!diag && k >= 4) ? 1.f : 0.f;                      // matrix-core path
}

namespace {

// Nonzero pattern of a composed group over its qubits in first-seen order: row r is the
// bitmask of nonzero columns (k <= 6, so 64 rows of 64 bits).
struct Pattern {
  uint32_t k = 0;
  uint32_t q[kMaxGateQubits] = {0, 0, 0, 0, 0, 0};
  uint64_t row[64] = {1};

  int pos(uint32_t qb) const {
    for (uint32_t i = 0; i < k; ++i)
      if (q[i] == qb) return int(i);
    return -1;
  }
  // Append qubit qb (identity on it).
  void grow(uint32_t qb) {
    const uint32_t d = 1u << k;
    for (uint32_t r = 0; r < d; ++r) row[r + d] = row[r] << d;
    q[k++] = qb;
  }
  // Left-multiply by the pattern of op (nonzero entries of its matrix).
  void apply(const Op& op) {
    uint32_t local[kMaxGateQubits];
    for (uint32_t j = 0; j < op.m; ++j) local[j] = uint32_t(pos(op.q[j]));
    const uint32_t d = 1u << k, sub = 1u << op.m;
    uint64_t out[64];
    for (uint32_t r = 0; r < d; ++r) {
      uint32_t rs = 0;
      for (uint32_t j = 0; j < op.m; ++j) rs |= ((r >> local[j]) & 1u) << j;
      uint64_t acc = 0;
      for (uint32_t t = 0; t < sub; ++t) {
        const size_t e = size_t(rs) * sub + t;
        if (op.re[e] == 0.0 && op.im[e] == 0.0) continue;
        uint32_t kr = r;
        for (uint32_t j = 0; j < op.m; ++j)
          kr = (kr & ~(1u << local[j])) | (((t >> j) & 1u) << local[j]);
        acc |= row[kr];
      }
      out[r] = acc;
    }
    for (uint32_t r = 0; r < d; ++r) row[r] = out[r];
  }
  double density() const {
    const uint32_t d = 1u << k;
    uint64_t nz = 0;
    for (uint32_t r = 0; r < d; ++r) nz += uint64_t(__builtin_popcountll(row[r]));
    return double(nz) / double(uint64_t(d) * d);
  }
};

struct Candidate {
  uint32_t begin, end;
  float f[kFeatures];
};

Candidate make_candidate(const Circuit& c, uint32_t b, uint32_t e, const Pattern& pt, bool diag) {
  Candidate cd;
  cd.begin = b;
  cd.end = e;
  uint32_t q[kMaxGateQubits];
  std::copy(pt.q, pt.q + pt.k, q);
  std::sort(q, q + pt.k);
  group_features(c.n, pt.k, q, diag, pt.density(), cd.f);
  return cd;
}

bool op_diag(const Op& op) { return op.is_diagonal(); }

// Predicted milliseconds of every candidate (targets are log2 of ms scaled to 2^30
// amplitudes).
int predict_ms(Mlp& mlp, uint32_t n, const std::vector<Candidate>& cands, std::vector<double>* ms) {
  std::vector<float> x(cands.size() * kFeatures), y(cands.size());
  for (size_t i = 0; i < cands.size(); ++i)
    std::copy(cands[i].f, cands[i].f + kFeatures, x.begin() + i * kFeatures);
  if (cands.empty()) return kOk;
  const int rc = mlp.predict(x.data(), cands.size(), y.data(), false);
  if (rc != kOk) return rc;
  ms->resize(cands.size());
  for (size_t i = 0; i < cands.size(); ++i)
    (*ms)[i] = std::exp2(double(y[i])) * std::exp2(double(n) - 30.0);
  return kOk;
}

// Candidates for one greedy plan (fusion.rs:39-77) at cap `max`.
std::vector<Candidate> greedy_candidates(const Circuit& c, uint32_t max) {
  std::vector<Candidate> out;
  for (const Group& g : greedy_groups(c, max)) {
    Pattern pt;
    bool diag = true;
    for (uint64_t i = g.begin; i < g.end; ++i) {
      const Op& op = c.ops[i];
      for (uint32_t j = 0; j < op.m; ++j)
        if (pt.pos(op.q[j]) < 0) pt.grow(op.q[j]);
      pt.apply(op);
      diag = diag && op_diag(op);
    }
    out.push_back(make_candidate(c, uint32_t(g.begin), uint32_t(g.end), pt, diag));
  }
  return out;
}

}  // namespace

// Build the circuit of a partition: single-op groups pass through, the others are
// composed in f64 on the device and emitted on their sorted qubits (fusion.rs:79-99).
int build_partition(Ctx* ctx, const Circuit& c, const std::vector<std::pair<uint32_t, uint32_t>>& parts,
                    Circuit* out) {
  std::vector<ComposeGroup> cg;
  std::vector<ComposeGate> gates;
  std::vector<double> mats;
  std::vector<std::vector<uint32_t>> qs(parts.size());
  uint64_t out_off = 0;
  for (size_t gi = 0; gi < parts.size(); ++gi) {
    const uint32_t b = parts[gi].first, e = parts[gi].second;
    if (e - b == 1) continue;
    std::vector<uint32_t>& q = qs[gi];
    for (uint32_t i = b; i < e; ++i)
      for (uint32_t j = 0; j < c.ops[i].m; ++j)
        if (std::find(q.begin(), q.end(), c.ops[i].q[j]) == q.end()) q.push_back(c.ops[i].q[j]);
    std::sort(q.begin(), q.end());
    ComposeGroup x;
    x.k = uint32_t(q.size());
    x.first_gate = uint32_t(gates.size());
    x.n_gates = e - b;
    x.out_off = out_off;
    out_off += uint64_t(1) << (2 * x.k);
    for (uint32_t i = b; i < e; ++i) {
      const Op& op = c.ops[i];
      ComposeGate y;
      y.m = op.m;
      for (uint32_t j = 0; j < op.m; ++j)
        y.local[j] = uint32_t(std::find(q.begin(), q.end(), op.q[j]) - q.begin());
      y.mat_off = mats.size() / 2;
      for (size_t k = 0; k < op.re.size(); ++k) {
        mats.push_back(op.re[k]);
        mats.push_back(op.im[k]);
      }
      gates.push_back(y);
    }
    cg.push_back(x);
  }
  std::vector<double> composed;
  const int rc = compose_groups(ctx, cg, gates, mats, &composed);
  if (rc != kOk) return rc;
  out->n = c.n;
  out->ops.clear();
  size_t k = 0;
  for (size_t gi = 0; gi < parts.size(); ++gi) {
    const uint32_t b = parts[gi].first, e = parts[gi].second;
    if (e - b == 1) {
      out->ops.push_back(c.ops[b]);
      continue;
    }
    const ComposeGroup& x = cg[k++];
    Op op;
    op.m = x.k;
    std::copy(qs[gi].begin(), qs[gi].end(), op.q);
    const size_t nn = size_t(1) << (2 * x.k);
    op.re.resize(nn);
    op.im.resize(nn);
    for (size_t i = 0; i < nn; ++i) {
      op.re[i] = composed[2 * (x.out_off + i)];
      op.im[i] = composed[2 * (x.out_off + i) + 1];
    }
    out->ops.push_back(std::move(op));
  }
  return kOk;
}

int plan_fusion(Ctx* ctx, Mlp& mlp, const Circuit& c, uint32_t max_qubits, uint32_t window,
                Circuit* out, double* predicted_ms, int* chosen) {
  const uint32_t L = uint32_t(c.ops.size());
  // candidate groups: every run [b, e) with <= max_qubits qubits and <= window gates
  std::vector<Candidate> cands;
  for (uint32_t b = 0; b < L; ++b) {
    Pattern pt;
    bool diag = true;
    for (uint32_t e = b; e < L && e - b < window; ++e) {
      const Op& op = c.ops[e];
      uint32_t extra = 0;
      for (uint32_t j = 0; j < op.m; ++j)
        if (pt.pos(op.q[j]) < 0) ++extra;
      if (pt.k + extra > max_qubits) {
        // an op wider than the cap is a group of its own, as in fusion.rs:49-54
        if (e == b) {
          for (uint32_t j = 0; j < op.m; ++j) pt.grow(op.q[j]);
          pt.apply(op);
          cands.push_back(make_candidate(c, b, b + 1, pt, op_diag(op)));
        }
        break;
      }
      for (uint32_t j = 0; j < op.m; ++j)
        if (pt.pos(op.q[j]) < 0) pt.grow(op.q[j]);
      pt.apply(op);
      diag = diag && op_diag(op);
      cands.push_back(make_candidate(c, b, e + 1, pt, diag));
    }
  }
  std::vector<double> cost;
  int rc = predict_ms(mlp, c.n, cands, &cost);
  if (rc != kOk) return rc;
  // shortest path over positions 0..L
  std::vector<double> best(L + 1, std::numeric_limits<double>::infinity());
  std::vector<uint32_t> from(L + 1, 0);
  best[0] = 0;
  std::vector<std::vector<size_t>> ending(L + 1);
  for (size_t i = 0; i < cands.size(); ++i) ending[cands[i].end].push_back(i);
  for (uint32_t e = 1; e <= L; ++e)
    for (size_t i : ending[e]) {
      const double v = best[cands[i].begin] + cost[i];
      if (v < best[e]) {
        best[e] = v;
        from[e] = uint32_t(i);
      }
    }
  // Every position is reachable (each op has at least its own singleton candidate), so a
  // non-finite total means a prediction was not finite: the greedy plans decide then.
  std::vector<std::pair<uint32_t, uint32_t>> parts;
  double total = L ? best[L] : 0.0;
  if (std::isfinite(total)) {
    for (uint32_t e = L; e > 0;) {
      const Candidate& cd = cands[from[e]];
      parts.emplace_back(cd.begin, cd.end);
      e = cd.begin;
    }
    std::reverse(parts.begin(), parts.end());
  }
  int pick = 0;  // 0: dynamic programme; m > 0: greedy at max_qubits m
  for (uint32_t m = 1; m <= max_qubits; ++m) {
    const std::vector<Candidate> g = greedy_candidates(c, m);
    std::vector<double> gc;
    rc = predict_ms(mlp, c.n, g, &gc);
    if (rc != kOk) return rc;
    double t = 0;
    for (double x : gc) t += x;
    // (the last greedy plan is taken whatever it predicts when nothing finite came before)
    if (t < total || (m == max_qubits && L && parts.empty())) {
      total = t;
      pick = int(m);
      parts.clear();
      for (const Candidate& cd : g) parts.emplace_back(cd.begin, cd.end);
    }
  }
  if (predicted_ms) *predicted_ms = total;
  if (chosen) *chosen = pick;
  return build_partition(ctx, c, parts, out);
}

// Learned choice of the execution form: the fused plan (plan_fusion, matrix-core groups)
// against tile passes at every allowed tile size, each priced by its cost model -- pass
// times by the tile model, launches a tile cannot hold by the group model -- keeping the
// lowest predicted total.  Tiled: *out is the circuit unchanged and *tile_qubits > 0.
int plan_form(Ctx* ctx, Mlp& group_mlp, Mlp& tile_mlp, Mlp* mt_mlp, const Circuit& c,
              int dtype, uint32_t max_qubits, uint32_t window, Circuit* out, int* tile_qubits,
              int* mfma_tiles, double* predicted_ms, int* chosen) {
  double fused_ms = 0;
  int fused_pick = 0;
  int rc = plan_fusion(ctx, group_mlp, c, max_qubits, window, out, &fused_ms, &fused_pick);
  if (rc != kOk) return rc;
  double best = fused_ms;
  int best_tb = 0;
  const uint32_t lo = dtype == 0 ? 10 : 11, hi = dtype == 0 ? 12 : 13;
  for (uint32_t tb = lo; tb <= hi && tb <= c.n; ++tb) {
    std::vector<float> rows, y;
    std::vector<Launch> other;
    tile_schedule_features(c, dtype, tb, &rows, &other);
    const size_t np = rows.size() / kTileFeatures;
    y.resize(np);
    if (np) {
      rc = tile_mlp.predict(rows.data(), np, y.data(), false);
      if (rc != kOk) return rc;
    }
    double t = 0;
    for (size_t i = 0; i < np; ++i) t += std::exp2(double(y[i])) * std::exp2(double(c.n) - 30.0);
    std::vector<Candidate> cands;
    for (const Launch& l : other) {
      Candidate cd;
      other_features(l, c.n, cd.f);
      cands.push_back(cd);
    }
    std::vector<double> oc;
    rc = predict_ms(group_mlp, c.n, cands, &oc);
    if (rc != kOk) return rc;
    for (double x : oc) t += x;
    if (t < best) {
      best = t;
      best_tb = int(tb);
    }
  }
  // greedy fusion at k = 4, 5 run as MFMA tile passes (complex64): passes priced by the
  // MFMA-tile model, launches no pass holds by the group model
  int best_mt = 0;
  Circuit mt_best;
  if (mt_mlp && dtype == 1 && c.n >= 12) {
    for (uint32_t k = 4; k <= std::min<uint32_t>(5, max_qubits); ++k) {
      std::vector<std::pair<uint32_t, uint32_t>> parts;
      for (const Group& g : greedy_groups(c, k))
        parts.emplace_back(uint32_t(g.begin), uint32_t(g.end));
      Circuit fused;
      rc = build_partition(ctx, c, parts, &fused);
      if (rc != kOk) return rc;
      std::vector<float> rows, y;
      std::vector<Launch> other;
      mt_schedule_features(fused, &rows, &other);
      const size_t np = rows.size() / kMtFeatures;
      y.resize(np);
      if (np) {
        rc = mt_mlp->predict(rows.data(), np, y.data(), false);
        if (rc != kOk) return rc;
      }
      double t = 0;
      for (size_t i = 0; i < np; ++i) t += std::exp2(double(y[i])) * std::exp2(double(c.n) - 30.0);
      std::vector<Candidate> cands;
      for (const Launch& l : other) {
        Candidate cd;
        other_features(l, c.n, cd.f);
        cands.push_back(cd);
      }
      std::vector<double> oc;
      rc = predict_ms(group_mlp, c.n, cands, &oc);
      if (rc != kOk) return rc;
      for (double x : oc) t += x;
      if (t < best) {
        best = t;
        best_mt = int(k);
        best_tb = 0;
        mt_best = std::move(fused);
      }
    }
  }
  if (best_mt > 0) *out = std::move(mt_best);
  else if (best_tb > 0) *out = c;
  if (tile_qubits) *tile_qubits = best_tb;
  if (mfma_tiles) *mfma_tiles = best_mt > 0 ? 12 : 0;
  if (predicted_ms) *predicted_ms = best;
  if (chosen) *chosen = best_mt > 0 ? 100 + best_mt : best_tb > 0 ? -best_tb : fused_pick;
  return kOk;
}

}  // namespace qsvh

using namespace qsvh;

extern "C" {

int qsvh_group_features(uint32_t n, uint32_t k, const uint32_t* qubits, int diagonal,
                        double density, float* out) {
  if (!out || k > kMaxGateQubits || (k && !qubits)) return kInvalidArg;
  uint32_t q[kMaxGateQubits];
  std::copy(qubits, qubits + k, q);
  std::sort(q, q + k);
  group_features(n, k, q, diagonal != 0, density, out);
  return kOk;
}

int qsvh_plan_fusion(qsvh_ctx* h, qsvh_mlp* model, const qsvh_circuit* in, uint32_t max_qubits,
                     uint32_t window, qsvh_circuit** out, double* predicted_ms, int* chosen) {
  if (!h || !model || !in || !out) return kInvalidArg;
  *out = nullptr;
  if (max_qubits < 1 || max_qubits > kMaxGateQubits)
    return set_err(&h->c, kTooWide, "plan: max_qubits must be 1..6");
  if (model->m.F != kFeatures) return set_err(&h->c, kInvalidArg, "plan: model has 12 inputs");
  qsvh_circuit* res = new (std::nothrow) qsvh_circuit;
  if (!res) return kOom;
  const int rc = plan_fusion(&h->c, model->m, in->c, max_qubits, window ? window : 48, &res->c,
                             predicted_ms, chosen);
  if (rc != kOk) {
    delete res;
    return rc;
  }
  *out = res;
  return kOk;
}

int qsvh_tile_pass_features(const qsvh_circuit* c, int dtype, uint32_t tile_qubits, float* out,
                            size_t max_passes, size_t* n_passes, size_t* n_other) {
  if (!c || (dtype != 0 && dtype != 1) || (max_passes && !out)) return kInvalidArg;
  const uint32_t lo = dtype == 0 ? 10 : 11, hi = dtype == 0 ? 12 : 13;
  if (tile_qubits < lo || tile_qubits > hi || c->c.n < tile_qubits) return kInvalidArg;
  std::vector<float> rows;
  std::vector<Launch> other;
  tile_schedule_features(c->c, dtype, tile_qubits, &rows, &other);
  const size_t np = rows.size() / kTileFeatures;
  std::copy(rows.begin(), rows.begin() + std::min(np, max_passes) * kTileFeatures, out);
  if (n_passes) *n_passes = np;
  if (n_other) *n_other = other.size();
  return kOk;
}

int qsvh_mt_pass_features(const qsvh_circuit* c, float* out, size_t max_passes,
                          size_t* n_passes, size_t* n_other) {
  if (!c || (max_passes && !out) || c->c.n < 12) return kInvalidArg;
  std::vector<float> rows;
  std::vector<Launch> other;
  mt_schedule_features(c->c, &rows, &other);
  const size_t np = rows.size() / kMtFeatures;
  std::copy(rows.begin(), rows.begin() + std::min(np, max_passes) * kMtFeatures, out);
  if (n_passes) *n_passes = np;
  if (n_other) *n_other = other.size();
  return kOk;
}

int qsvh_plan_form(qsvh_ctx* h, qsvh_mlp* group_model, qsvh_mlp* tile_model,
                   qsvh_mlp* mt_model, const qsvh_circuit* in, int dtype, uint32_t max_qubits,
                   uint32_t window, qsvh_circuit** out, int* tile_qubits, int* mfma_tiles,
                   double* predicted_ms, int* chosen) {
  if (!h || !group_model || !tile_model || !in || !out || (dtype != 0 && dtype != 1))
    return kInvalidArg;
  *out = nullptr;
  if (max_qubits < 1 || max_qubits > kMaxGateQubits)
    return set_err(&h->c, kTooWide, "plan: max_qubits must be 1..6");
  if (group_model->m.F != kFeatures) return set_err(&h->c, kInvalidArg, "plan: group model has 12 inputs");
  if (tile_model->m.F != kTileFeatures)
    return set_err(&h->c, kInvalidArg, "plan: tile model has 18 inputs");
  if (mt_model && mt_model->m.F != kMtFeatures)
    return set_err(&h->c, kInvalidArg, "plan: MFMA-tile model has 11 inputs");
  qsvh_circuit* res = new (std::nothrow) qsvh_circuit;
  if (!res) return kOom;
  const int rc = plan_form(&h->c, group_model->m, tile_model->m, mt_model ? &mt_model->m : nullptr,
                           in->c, dtype, max_qubits, window ? window : 48, &res->c, tile_qubits,
                           mfma_tiles, predicted_ms, chosen);
  if (rc != kOk) {
    delete res;
    return rc;
  }
  *out = res;
  return kOk;
}

}  // extern "C"
