// batch.cpp -- many independent circuits (uneven qubit counts) executed together: one
// persistent launch per pass index, captured once in a HIP graph and replayed; the tile
// size and the persistent grid are chosen by timing candidates on the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/qsvh.h"
#include "engine.hpp"
#include "ir.hpp"
#include "program.hpp"
#include "handles.hpp"


namespace qsvh {

Launch classify(const Op& op, int dtype);  // api.cpp

struct Batch {
  Ctx* ctx = nullptr;
  int dtype = 1;
  std::vector<Circuit> circuits;
  std::vector<uint64_t> off;  // state offsets (amplitudes)
  uint64_t total = 0;
  void* amp = nullptr;        // every circuit's interleaved state, back to back
  // plan
  int tb = 0;
  int blocks = 0;
  std::vector<std::vector<BatchPass>> launches;  // pass j of every circuit
  std::vector<std::vector<uint64_t>> order;      // per circuit
  std::vector<uint32_t> part_off, n_part;
  BatchPass* d_passes = nullptr;
  std::vector<uint64_t> launch_first;  // index of launch j's first pass in d_passes
  TileOp* d_ops = nullptr;
  TileDiag* d_diags = nullptr;
  void* d_pool = nullptr;
  double* d_partials = nullptr;
  uint32_t* d_part_off = nullptr;
  uint32_t* d_n_part = nullptr;
  double* d_out = nullptr;
  unsigned long long* d_counter = nullptr;
  hipGraphExec_t graph = nullptr;

  size_t elem() const { return dtype == 0 ? sizeof(double) : sizeof(float); }
  State view(size_t i) const {
    State s;
    s.ctx = ctx;
    s.n = circuits[i].n;
    s.dtype = dtype;
    s.amp = static_cast<char*>(amp) + off[i] * 2 * elem();
    return s;
  }
  void free_plan() {
    if (graph) hipGraphExecDestroy(graph);
    graph = nullptr;
    for (void* p : {static_cast<void*>(d_passes), static_cast<void*>(d_ops),
                    static_cast<void*>(d_diags), d_pool, static_cast<void*>(d_partials),
                    static_cast<void*>(d_part_off), static_cast<void*>(d_n_part)})
      hipFree(p);
    d_passes = nullptr;
    d_ops = nullptr;
    d_diags = nullptr;
    d_pool = nullptr;
    d_partials = nullptr;
    d_part_off = nullptr;
    d_n_part = nullptr;
    launches.clear();
    order.clear();
  }
};

namespace {

template <typename X>
int upload(Ctx* ctx, const std::vector<X>& v, X** dst) {
  const size_t bytes = std::max<size_t>(v.size() * sizeof(X), 16);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), bytes);
  if (e != hipSuccess) return hip_err(ctx, e, "batch alloc");
  if (!v.empty()) e = hipMemcpy(*dst, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice);
  return hip_err(ctx, e, "batch upload");
}

// Per-tile cost estimate used to order work items (largest first): ops x tile size.
uint64_t pass_cost(const BatchPass& p) { return uint64_t(p.n_ops + 1) << p.tb; }

int plan(Batch* b, int tb, int blocks) {
  b->free_plan();
  b->tb = tb;
  b->blocks = blocks;
  TileBuilder bld;
  bld.free_runs = false;  // the batch kernel runs the interpreter without free lane runs
  std::vector<TilePassArg> args;
  uint32_t n_partials = 0;
  b->order.resize(b->circuits.size());
  b->part_off.assign(b->circuits.size(), 0);
  b->n_part.assign(b->circuits.size(), 0);
  for (size_t c = 0; c < b->circuits.size(); ++c) {
    const Circuit& circ = b->circuits[c];
    std::vector<Launch> in;
    in.reserve(circ.ops.size());
    for (const Op& op : circ.ops) in.push_back(classify(op, b->dtype));
    const uint32_t tbc = std::min<uint32_t>(uint32_t(tb), circ.n);
    std::vector<Launch> out;
    const size_t first_arg = args.size();
    schedule_tiles(in, circ.n, tbc, bld, args, out, b->order[c]);
    for (const Launch& l : out)
      if (l.kind != Launch::kTile)
        return set_err(b->ctx, kTooWide, "batch circuits take gates acting on <= 3 qubits");
    if (out.empty()) {  // no gates: one empty pass still builds |0> and <Z_0>
      TilePassArg pa{};
      pa.tb = tbc;
      for (uint32_t j = 0; j < tbc; ++j) pa.ts[j] = pa.tsort[j] = j;
      pa.op_begin = uint32_t(bld.ops.size());
      pa.n_ops = 0;
      Launch l;
      l.kind = Launch::kTile;
      l.arena_off = args.size();
      args.push_back(pa);
      out.push_back(l);
    }
    for (size_t j = 0; j < out.size(); ++j) {
      const TilePassArg& pa = args[out[j].arena_off];
      BatchPass bp{};
      bp.state_off = b->off[c];
      bp.n = circ.n;
      bp.tb = pa.tb;
      std::copy(pa.ts, pa.ts + 16, bp.ts);
      std::copy(pa.tsort, pa.tsort + 16, bp.tsort);
      bp.op_begin = pa.op_begin;
      bp.n_ops = pa.n_ops;
      bp.n_tiles = uint64_t(1) << (circ.n - pa.tb);
      bp.flags = (j == 0 ? kBatchInit : 0u) | (j + 1 == out.size() ? kBatchExpZ0 : 0u);
      if (j + 1 == out.size()) {
        bp.part_off = n_partials;
        b->part_off[c] = n_partials;
        b->n_part[c] = uint32_t(bp.n_tiles);
        n_partials += uint32_t(bp.n_tiles);
      }
      if (b->launches.size() <= j) b->launches.resize(j + 1);
      b->launches[j].push_back(bp);
    }
    (void)first_arg;
  }
  // order each launch's passes by decreasing tile cost, then number the work items
  std::vector<BatchPass> all;
  b->launch_first.clear();
  for (auto& L : b->launches) {
    std::stable_sort(L.begin(), L.end(),
                     [](const BatchPass& x, const BatchPass& y) { return pass_cost(x) > pass_cost(y); });
    uint64_t w = 0;
    for (BatchPass& p : L) {
      p.tile_begin = w;
      w += p.n_tiles;
    }
    b->launch_first.push_back(all.size());
    all.insert(all.end(), L.begin(), L.end());
  }
  int rc = upload(b->ctx, all, &b->d_passes);
  if (rc == kOk) rc = upload(b->ctx, bld.ops, &b->d_ops);
  if (rc == kOk) rc = upload(b->ctx, bld.diags, &b->d_diags);
  if (rc == kOk) {
    if (b->dtype == 0) {
      double* d = nullptr;
      rc = upload(b->ctx, bld.pool, &d);
      b->d_pool = d;
    } else {
      std::vector<float> f(bld.pool.begin(), bld.pool.end());
      float* d = nullptr;
      rc = upload(b->ctx, f, &d);
      b->d_pool = d;
    }
  }
  if (rc == kOk) rc = upload(b->ctx, std::vector<double>(std::max<uint32_t>(n_partials, 1)), &b->d_partials);
  if (rc == kOk) rc = upload(b->ctx, b->part_off, &b->d_part_off);
  if (rc == kOk) rc = upload(b->ctx, b->n_part, &b->d_n_part);
  return rc;
}

hipError_t issue(Batch* b) {
  hipStream_t s = b->ctx->stream;
  for (size_t j = 0; j < b->launches.size(); ++j) {
    const auto& L = b->launches[j];
    uint64_t work = 0;
    for (const BatchPass& p : L) work += p.n_tiles;
    const hipError_t e = launch_batch_pass(b->dtype, b->d_passes + b->launch_first[j],
                                           uint32_t(L.size()), work, b->d_counter, b->amp,
                                           b->d_ops, b->d_diags, b->d_pool,
                                           b->d_partials, b->blocks, s);
    if (e != hipSuccess) return e;
  }
  return launch_batch_reduce(b->d_part_off, b->d_n_part, uint32_t(b->circuits.size()),
                             b->d_partials, b->d_out, s);
}

int timed(Batch* b, double* ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return set_err(b->ctx, kHip, "event");
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return set_err(b->ctx, kHip, "event");
  }
  hipError_t e = hipEventRecord(e0, b->ctx->stream);
  if (e == hipSuccess) e = b->graph ? hipGraphLaunch(b->graph, b->ctx->stream) : issue(b);
  if (e == hipSuccess) e = hipEventRecord(e1, b->ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float f = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) return hip_err(b->ctx, e, "batch time");
  *ms = f;
  return kOk;
}

int capture(Batch* b) {
  hipStream_t s = b->ctx->stream;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_err(b->ctx, e, "batch capture");
  const hipError_t le = issue(b);
  e = hipStreamEndCapture(s, &g);
  if (le != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return hip_err(b->ctx, le, "batch capture launch");
  }
  if (e != hipSuccess) return hip_err(b->ctx, e, "batch end capture");
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  return hip_err(b->ctx, e, "batch graph instantiate");
}

int default_blocks(const Batch* b) { return std::max(b->ctx->cus, 1) * 2; }

}  // namespace
}  // namespace qsvh

using namespace qsvh;

struct qsvh_batch {
  Batch b;
};

extern "C" {

int qsvh_batch_create(qsvh_ctx* h, const qsvh_circuit* const* circuits, size_t count, int dtype,
                      qsvh_batch** out) {
  if (!h || !circuits || !out || count == 0) return kInvalidArg;
  *out = nullptr;
  if (dtype != QSVH_C128 && dtype != QSVH_C64) return set_err(&h->c, kDtype, "unknown dtype");
  qsvh_batch* bh = new (std::nothrow) qsvh_batch;
  if (!bh) return kOom;
  Batch& b = bh->b;
  b.ctx = &h->c;
  b.dtype = dtype;
  for (size_t i = 0; i < count; ++i) {
    if (!circuits[i] || circuits[i]->c.n < 4 || circuits[i]->c.n > 30) {
      delete bh;
      return set_err(&h->c, kInvalidArg, "batch circuits need 4..30 qubits");
    }
    b.circuits.push_back(circuits[i]->c);
    b.off.push_back(b.total);
    b.total += uint64_t(1) << circuits[i]->c.n;
  }
  hipError_t e = hipMalloc(&b.amp, b.total * 2 * b.elem());
  if (e == hipSuccess) e = hipMalloc(&b.d_out, count * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b.d_counter), 64);
  if (e != hipSuccess) {
    hipFree(b.amp);
    hipFree(b.d_out);
    hipFree(b.d_counter);
    delete bh;
    return hip_err(&h->c, e, "batch alloc");
  }
  const int rc = plan(&b, dtype == QSVH_C128 ? 12 : 13, default_blocks(&b));
  if (rc != kOk) {
    qsvh_batch_destroy(bh);
    return rc;
  }
  *out = bh;
  return kOk;
}

void qsvh_batch_destroy(qsvh_batch* bh) {
  if (!bh) return;
  Batch& b = bh->b;
  hipStreamSynchronize(b.ctx->stream);
  b.free_plan();
  hipFree(b.amp);
  hipFree(b.d_out);
  hipFree(b.d_counter);
  delete bh;
}

int qsvh_batch_configure(qsvh_batch* bh, int tile_qubits, int blocks) {
  if (!bh) return kInvalidArg;
  Batch& b = bh->b;
  const int lo = b.dtype == QSVH_C128 ? 10 : 11, hi = b.dtype == QSVH_C128 ? 12 : 13;
  if (tile_qubits < lo || tile_qubits > hi || blocks < 0 || blocks > (1 << 20))
    return set_err(b.ctx, kInvalidArg, "batch: tile_qubits out of range or bad grid");
  hipStreamSynchronize(b.ctx->stream);
  return plan(&b, tile_qubits, blocks ? blocks : default_blocks(&b));
}

int qsvh_batch_autotune(qsvh_batch* bh, int* tile_qubits, int* blocks, double* best_ms) {
  if (!bh) return kInvalidArg;
  Batch& b = bh->b;
  const int lo = b.dtype == QSVH_C128 ? 10 : 11, hi = b.dtype == QSVH_C128 ? 12 : 13;
  const int cus = std::max(b.ctx->cus, 1);
  double best = 1e300;
  int best_tb = hi, best_bl = default_blocks(&b);
  for (int tb = lo; tb <= hi; ++tb) {
    for (int mult : {1, 2, 4, 8}) {
      int rc = qsvh_batch_configure(bh, tb, cus * mult);
      if (rc != kOk) return rc;
      double t = 0, m = 1e300;
      for (int rep = 0; rep < 3; ++rep) {  // the best of three damps launch noise
        rc = timed(&b, &t);
        if (rc != kOk) return rc;
        m = std::min(m, t);
      }
      if (m < best) {
        best = m;
        best_tb = tb;
        best_bl = cus * mult;
      }
    }
  }
  const int rc = qsvh_batch_configure(bh, best_tb, best_bl);
  if (rc != kOk) return rc;
  if (tile_qubits) *tile_qubits = best_tb;
  if (blocks) *blocks = best_bl;
  if (best_ms) *best_ms = best;
  return kOk;
}

int qsvh_batch_run(qsvh_batch* bh) {
  if (!bh) return kInvalidArg;
  Batch& b = bh->b;
  if (!b.graph) {
    const int rc = capture(&b);
    if (rc != kOk) return rc;
  }
  return hip_err(b.ctx, hipGraphLaunch(b.graph, b.ctx->stream), "batch run");
}

int qsvh_batch_time(qsvh_batch* bh, double* ms) {
  if (!bh || !ms) return kInvalidArg;
  Batch& b = bh->b;
  if (!b.graph) {
    const int rc = capture(&b);
    if (rc != kOk) return rc;
  }
  return timed(&b, ms);
}

int qsvh_batch_passes(const qsvh_batch* bh, size_t i, uint32_t* passes) {
  if (!bh || !passes) return kInvalidArg;
  const Batch& b = bh->b;
  if (i >= b.circuits.size()) return kInvalidArg;
  uint32_t n = 0;
  for (const auto& L : b.launches)
    for (const BatchPass& p : L)
      if (p.state_off == b.off[i]) {
        ++n;
        break;
      }
  *passes = n;
  return kOk;
}

int qsvh_batch_info(const qsvh_batch* bh, size_t* count, int* tile_qubits, int* blocks,
                    uint64_t* n_launches, uint64_t* total_amplitudes) {
  if (!bh) return kInvalidArg;
  const Batch& b = bh->b;
  if (count) *count = b.circuits.size();
  if (tile_qubits) *tile_qubits = b.tb;
  if (blocks) *blocks = b.blocks;
  if (n_launches) *n_launches = b.launches.size();
  if (total_amplitudes) *total_amplitudes = b.total;
  return kOk;
}

int qsvh_batch_expectation_z0(qsvh_batch* bh, double* out) {
  if (!bh || !out) return kInvalidArg;
  Batch& b = bh->b;
  hipError_t e = hipStreamSynchronize(b.ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpy(out, b.d_out, b.circuits.size() * sizeof(double), hipMemcpyDeviceToHost);
  return hip_err(b.ctx, e, "batch expectation");
}

int qsvh_batch_download(qsvh_batch* bh, size_t i, void* re, void* im) {
  if (!bh || !re || !im || i >= bh->b.circuits.size()) return kInvalidArg;
  Batch& b = bh->b;
  const State s = b.view(i);
  hipError_t e = hipStreamSynchronize(b.ctx->stream);
  if (e == hipSuccess) e = download_split(s, re, im);
  return hip_err(b.ctx, e, "batch download");
}

int qsvh_batch_probabilities(qsvh_batch* bh, size_t i, void* host_out) {
  if (!bh || !host_out || i >= bh->b.circuits.size()) return kInvalidArg;
  Batch& b = bh->b;
  const State s = b.view(i);
  const size_t bytes = size_t(s.dim()) * s.elem();
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, bytes);
  if (e != hipSuccess) return hip_err(b.ctx, e, "batch probabilities alloc");
  e = launch_probs(s, d, b.ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(host_out, d, bytes, hipMemcpyDeviceToHost, b.ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b.ctx->stream);
  hipFree(d);
  return hip_err(b.ctx, e, "batch probabilities");
}

int qsvh_batch_order(const qsvh_batch* bh, size_t i, uint64_t* order, size_t cap, size_t* n) {
  if (!bh || !n || i >= bh->b.circuits.size()) return kInvalidArg;
  const std::vector<uint64_t>& o = bh->b.order[i];
  *n = o.size();
  for (size_t k = 0; k < o.size() && k < cap && order; ++k) order[k] = o[k];
  return kOk;
}

}  // extern "C"
