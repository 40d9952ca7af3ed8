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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <new>
#include <vector>

#include "../../include/qsvh.h"
#include "handles.hpp"
#include "mlp.hpp"

namespace qsvh {

constexpr int kFeatures = 12;

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
      if (pt.k + extra > max_qubits) break;
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
  std::vector<std::pair<uint32_t, uint32_t>> parts;
  for (uint32_t e = L; e > 0;) {
    const Candidate& cd = cands[from[e]];
    parts.emplace_back(cd.begin, cd.end);
    e = cd.begin;
  }
  std::reverse(parts.begin(), parts.end());
  double total = L ? best[L] : 0.0;
  int pick = 0;  // 0: dynamic programme; m > 0: greedy at max_qubits m
  for (uint32_t m = 1; m <= max_qubits; ++m) {
    const std::vector<Candidate> g = greedy_candidates(c, m);
    std::vector<double> gc;
    rc = predict_ms(mlp, c.n, g, &gc);
    if (rc != kOk) return rc;
    double t = 0;
    for (double x : gc) t += x;
    if (t < total) {
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

}  // extern "C"
