// api.cpp -- the C ABI (include/qsvh.h): contexts, device states, circuits, fusion and
// programs (a circuit bound to a state: classified launches, a device matrix arena
// uploaded once, optionally a captured HIP graph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "../../include/qsvh.h"
#include "engine.hpp"
#include "ir.hpp"
#include "program.hpp"
#include "handles.hpp"


namespace qsvh {

int set_err(Ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

int hip_err(Ctx* ctx, hipError_t e, const char* where) {
  if (e == hipSuccess) return kOk;
  return set_err(ctx, e == hipErrorOutOfMemory ? kOom : kHip,
                 std::string(where) + ": " + hipGetErrorString(e));
}

// Round an f64 matrix to the state's precision (DenseGate<R> is built with
// Cplx::from_f64, crates/qsv-core/src/complex.rs:43-50) and keep it as f64, interleaved.
std::vector<double> rounded(const Op& op, int dtype) {
  const size_t n = op.re.size();
  std::vector<double> g(2 * n);
  for (size_t e = 0; e < n; ++e) {
    g[2 * e] = dtype == 0 ? op.re[e] : double(float(op.re[e]));
    g[2 * e + 1] = dtype == 0 ? op.im[e] : double(float(op.im[e]));
  }
  return g;
}

bool diag_of(const std::vector<double>& g, uint32_t m) {
  const uint32_t d = 1u << m;
  for (uint32_t r = 0; r < d; ++r)
    for (uint32_t c = 0; c < d; ++c)
      if (r != c &&
          (g[2 * (size_t(r) * d + c)] != 0.0 || g[2 * (size_t(r) * d + c) + 1] != 0.0))
        return false;
  return true;
}

// CpuBackend::apply dispatch (crates/qsv-core/src/backend/cpu.rs:281-290):
// diagonal -> 1q -> mq, with the gate in the state's precision.
Launch classify(const Op& op, int dtype) {
  Launch l;
  l.m = op.m;
  std::copy(op.q, op.q + op.m, l.q);
  std::vector<double> g = rounded(op, dtype);
  if (diag_of(g, op.m)) {
    l.kind = Launch::kDiag;
    const uint32_t d = 1u << op.m;
    l.mat.resize(2 * d);
    for (uint32_t s = 0; s < d; ++s) {
      l.mat[2 * s] = g[2 * (size_t(s) * d + s)];
      l.mat[2 * s + 1] = g[2 * (size_t(s) * d + s) + 1];
    }
  } else if (op.m == 1) {
    l.kind = Launch::k1q;
    l.mat = std::move(g);
  } else {
    l.kind = Launch::kMq;
    l.mat = std::move(g);
  }
  return l;
}

hipError_t issue(const Program& p, const Launch& l, const Tuning& t, hipStream_t s) {
  const State& st = *p.st;
  const char* base = static_cast<const char*>(p.arena);
  const void* dmat = base ? base + l.arena_off * st.elem() : nullptr;
  switch (l.kind) {
    case Launch::k1q:
      return launch_1q(st, l.q[0], l.mat.data(), t, s);
    case Launch::kDiag:
      return launch_diag(st, l.m, l.q, l.mat.data(), t, s);
    case Launch::kMq:
      return launch_mq(st, l.m, l.q, l.mat.data(), dmat, t, s);
    case Launch::kFusedMfma:
      return launch_fused_mfma(st, l.m, l.q, static_cast<const char*>(p.mfma) + l.mfma_off,
                               p.precision, t, s);
    case Launch::kTile:
      return launch_tile(st, p.tile_passes[l.arena_off], p.tile_ops, p.tile_diags, p.tile_pool,
                         t.tile_blocks, s);
    case Launch::kMfmaTile:
      return launch_mfma_tiles(st, p.mt_passes[l.arena_off], p.mt_ops, p.mt_pool, t, s);
    case Launch::kCtrl1q:
      return launch_ctrl1q(st, l.m, l.q, l.fix, l.mat.data(), t, s);
    case Launch::kDiagSub:
      return launch_diag_sub(st, l.m, l.q, l.fix, l.mat.data(), t, s);
    case Launch::kSwapSub:
      return launch_swap_sub(st, l.q, t, s);
    case Launch::kBitSwap:
      return launch_bitswap(st, l.img.data(), t, s);
    default:
      return hipErrorInvalidValue;
  }
}

// Upload, once, the matrices of launches that read them from device memory (m >= 4),
// converted to the state's precision.
int build_arena(Program* p) {
  State& st = *p->st;
  Ctx* ctx = st.ctx;
  std::vector<double> host;
  for (Launch& l : p->launches) {
    if (l.kind == Launch::kMq && l.m >= 4) {
      l.arena_off = host.size();
      host.insert(host.end(), l.mat.begin(), l.mat.end());
    }
  }
  if (host.empty()) return kOk;
  p->arena_bytes = host.size() * st.elem();
  hipError_t e = hipMalloc(&p->arena, p->arena_bytes);
  if (e != hipSuccess) return hip_err(ctx, e, "arena alloc");
  if (st.dtype == 0) {
    e = hipMemcpy(p->arena, host.data(), p->arena_bytes, hipMemcpyHostToDevice);
  } else {
    std::vector<float> f(host.begin(), host.end());
    e = hipMemcpy(p->arena, f.data(), p->arena_bytes, hipMemcpyHostToDevice);
  }
  return hip_err(ctx, e, "arena upload");
}

hipError_t issue_all(Program* p) {
  State& st = *p->st;
  for (const Launch& l : p->launches) {
    const hipError_t e = issue(*p, l, st.ctx->tune, st.ctx->stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int program_launch(Program* p) {
  Ctx* ctx = p->st->ctx;
  if (p->launches.empty()) return kOk;  // an empty circuit: nothing to run or capture
  if (!p->use_graph) return hip_err(ctx, issue_all(p), "launch");
  if (!p->graph) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_err(ctx, e, "begin capture");
    const hipError_t le = issue_all(p);
    e = hipStreamEndCapture(ctx->stream, &g);
    if (le != hipSuccess) {
      if (g) hipGraphDestroy(g);
      return hip_err(ctx, le, "capture launch");
    }
    if (e != hipSuccess) return hip_err(ctx, e, "end capture");
    e = hipGraphInstantiate(&p->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_err(ctx, e, "graph instantiate");
  }
  return hip_err(ctx, hipGraphLaunch(p->graph, ctx->stream), "graph launch");
}

int program_build(State* st, const Circuit& c, const qsvh_program_opts* opts, Program* p) {
  p->st = st;
  p->use_graph = opts && opts->use_graph;
  p->precision = opts ? opts->precision : 0;
  if (c.n != st->n)
    return set_err(st->ctx, kInvalidArg, "circuit and state qubit counts differ");
  if (p->precision != QSVH_PREC_EXACT && st->dtype != QSVH_C64)
    return set_err(st->ctx, kDtype, "MFMA precisions need a complex64 state");
  p->launches.reserve(c.ops.size());
  for (const Op& op : c.ops) p->launches.push_back(classify(op, st->dtype));
  if (opts && opts->planned) {
    if (opts->tile_qubits > 0 || opts->mfma_tiles > 0 || opts->subspace)
      return set_err(st->ctx, kInvalidArg,
                     "planned excludes tile_qubits, mfma_tiles and subspace");
    const int rc = plan_planned_passes(p, c);
    if (rc != kOk) return rc;
    return build_arena(p);
  }
  const int mt = opts ? opts->mfma_tiles : 0;
  if (mt > 0 && opts->tile_qubits > 0)
    return set_err(st->ctx, kInvalidArg, "tile_qubits and mfma_tiles exclude each other");
  if (mt > 0 && opts->subspace)
    return set_err(st->ctx, kInvalidArg, "subspace launches do not combine with mfma_tiles");
  // Tile passes and fused groups leave a trailing qubit permutation (a QFT's SWAP network)
  // to bit-swap passes: one or two HBM round trips instead of passes that need each
  // swapped pair in one tile, and exact data movement instead of matrix-core products.
  // The per-gate forms keep every gate a launch of its own.
  // The MFMA tile form may run the swaps earlier (PermTail), where its passes come out
  // fewer (fused QFT(30) at k = 4: 5 passes instead of 6, 43.3 -> 42.2 ms).  The exact tile
  // form keeps them at the end: moved past the swaps, QFT's controlled-phase ladders index
  // lane bits instead of bits outside the tile, and 3 passes took longer than 4 (93.4 ->
  // 96.3 ms in complex128).
  const uint64_t n_ops = p->launches.size();
  const int tq = opts ? opts->tile_qubits : 0;
  PermTail pt;
  uint64_t tail = n_ops, body = n_ops;
  if (opts && (tq > 0 || mt > 0 || p->precision != QSVH_PREC_EXACT)) {
    tail = strip_perm_suffix(p->launches, st->n, &pt.swaps, &pt.dest);
    pt.k = body = p->launches.size();
    if (!pt.swaps.empty() && mt > 0 && st->n >= 12)
      pt.k = hoist_perm(p->launches, pt.dest, [&](const std::vector<Launch>& v) {
        return mt_launch_count(v, st->n);
      });
  }
  int rc = plan_tiles(p, tq, &pt);
  if (rc != kOk) return rc;
  rc = mt > 0 ? plan_mfma_tiles(p, mt, &pt) : plan_mfma(p);
  if (rc != kOk) return rc;
  if (opts && opts->subspace) plan_subspace(p);
  if (!pt.swaps.empty() && pt.k != body)
    return set_err(st->ctx, kInvalidArg, "internal: moved bit swaps left unplaced");
  for (uint64_t i = tail; i < n_ops; ++i) p->order.push_back(i);
  for (Launch& l : pt.swaps) p->launches.push_back(std::move(l));  // not placed by a planner
  return build_arena(p);
}

void program_free(Program* p) {
  // Device buffers may still be read by queued launches: wait before releasing them.
  // A program that owns nothing (gates of m <= 3 travel as kernel arguments) returns
  // without synchronising, so qsvh_apply stays asynchronous.
  const bool owns = p->graph || p->arena || !p->owned.empty();
  if (owns && p->st && p->st->ctx) hipStreamSynchronize(p->st->ctx->stream);
  if (p->graph) hipGraphExecDestroy(p->graph);
  hipFree(p->arena);
  release_extra(p);
  p->graph = nullptr;
  p->arena = nullptr;
}


// Host SoA <-> device interleaved: chunks of up to 2^24 amplitudes are split (joined) on the
// device between the state and a device staging pair, and the host copies move whole re / im
// runs straight to (from) the caller's arrays -- no per-amplitude host loop.
template <typename T>
static hipError_t split_t(const State& s, T* re, T* im) {
  const uint64_t dim = s.dim();
  const uint64_t chunk = std::min<uint64_t>(dim, uint64_t(1) << 24);
  hipStream_t str = s.ctx->stream;
  T* d = nullptr;
  hipError_t e = hipMalloc(&d, 2 * chunk * sizeof(T));
  if (e != hipSuccess) return e;
  for (uint64_t a = 0; a < dim && e == hipSuccess; a += chunk) {
    const uint64_t c = std::min(chunk, dim - a);
    e = launch_split(s, a, c, d, d + chunk, str);
    if (e == hipSuccess) e = hipMemcpyAsync(re + a, d, c * sizeof(T), hipMemcpyDeviceToHost, str);
    if (e == hipSuccess)
      e = hipMemcpyAsync(im + a, d + chunk, c * sizeof(T), hipMemcpyDeviceToHost, str);
  }
  const hipError_t f = hipStreamSynchronize(str);
  hipFree(d);
  return e != hipSuccess ? e : f;
}

template <typename T>
static hipError_t join_t(const State& s, const T* re, const T* im) {
  const uint64_t dim = s.dim();
  const uint64_t chunk = std::min<uint64_t>(dim, uint64_t(1) << 24);
  hipStream_t str = s.ctx->stream;
  T* d = nullptr;
  hipError_t e = hipMalloc(&d, 2 * chunk * sizeof(T));
  if (e != hipSuccess) return e;
  for (uint64_t a = 0; a < dim && e == hipSuccess; a += chunk) {
    const uint64_t c = std::min(chunk, dim - a);
    e = hipMemcpyAsync(d, re + a, c * sizeof(T), hipMemcpyHostToDevice, str);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d + chunk, im + a, c * sizeof(T), hipMemcpyHostToDevice, str);
    if (e == hipSuccess) e = launch_join(s, a, c, d, d + chunk, str);
  }
  const hipError_t f = hipStreamSynchronize(str);
  hipFree(d);
  return e != hipSuccess ? e : f;
}

hipError_t download_split(const State& s, void* re, void* im) {
  return s.dtype == 0 ? split_t(s, static_cast<double*>(re), static_cast<double*>(im))
                      : split_t(s, static_cast<float*>(re), static_cast<float*>(im));
}

hipError_t upload_join(const State& s, const void* re, const void* im) {
  return s.dtype == 0
             ? join_t(s, static_cast<const double*>(re), static_cast<const double*>(im))
             : join_t(s, static_cast<const float*>(re), static_cast<const float*>(im));
}

}  // namespace qsvh

using namespace qsvh;

// ======================================================================= C ABI
extern "C" {

int qsvh_abi_version(void) { return QSVH_ABI_VERSION; }

int qsvh_ctx_create(int device, qsvh_ctx** out) {
  if (!out) return kInvalidArg;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return kNoDevice;
  qsvh_ctx* h = new (std::nothrow) qsvh_ctx;
  if (!h) return kOom;
  Ctx& c = h->c;
  c.device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c.d_partials, Ctx::kMaxPartials * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c.d_result, sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&c.h_result, sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c.cus = prop.multiProcessorCount;
    // the allocation-free sampler's scratch address on this device (a failure here only
    // disables qsvh_sample_noalloc, which reports it)
    c.sample_scratch_err = sample_noalloc_scratch(device, &c.sample_scratch);
  }
  if (e != hipSuccess) {
    qsvh_ctx_destroy(h);
    return e == hipErrorOutOfMemory ? kOom : kHip;
  }
  *out = h;
  return kOk;
}

void qsvh_ctx_destroy(qsvh_ctx* h) {
  if (!h) return;
  Ctx& c = h->c;
  if (c.stream) hipStreamSynchronize(c.stream);
  hipFree(c.d_partials);
  hipFree(c.d_result);
  if (c.h_result) hipHostFree(c.h_result);
  if (c.stream) hipStreamDestroy(c.stream);
  delete h;
}

const char* qsvh_last_error(const qsvh_ctx* h) { return h ? h->c.last_error.c_str() : ""; }

int qsvh_synchronize_ctx(qsvh_ctx* h) {
  if (!h) return kInvalidArg;
  return hip_err(&h->c, hipStreamSynchronize(h->c.stream), "synchronize");
}

int qsvh_tune_set(qsvh_ctx* h, const char* key, int64_t value) {
  if (!h || !key) return kInvalidArg;
  Tuning& t = h->c.tune;
  const std::string k(key);
  if (k == "max_blocks" && value >= 0 && value <= (1ll << 30)) {
    t.stream_max_blocks = int(value);
  } else if (k == "unroll" && (value == 1 || value == 2 || value == 4)) {
    t.unroll = int(value);
  } else if (k == "nontemporal" && value >= 0 && value <= 2) {
    t.nontemporal = int(value);
  } else if (k == "mfma_blocks" && value >= 0 && value <= (1ll << 20)) {
    t.mfma_blocks = int(value);
  } else if (k == "mfma_layout" && value >= 0 && value <= 2) {
    t.mfma_layout = int(value);
  } else if (k == "tile_blocks" && value >= 0 && value <= (1ll << 20)) {
    t.tile_blocks = int(value);
  } else if (k == "contig" && (value == 0 || value == 1)) {
    t.contig = int(value);
  } else if (k == "xcd" && (value == 0 || value == 1)) {
    t.xcd = int(value);
  } else if (k == "mq_lane" && (value == 0 || value == 1)) {
    t.mq_lane = int(value);
  } else if (k == "diag_blocks" && value >= 0 && value <= (1ll << 20)) {
    t.diag_blocks = int(value);
  } else {
    return set_err(&h->c, kInvalidArg, "unknown tuning key or value: " + k);
  }
  return kOk;
}

int qsvh_tune_get(qsvh_ctx* h, const char* key, int64_t* value) {
  if (!h || !key || !value) return kInvalidArg;
  const Tuning& t = h->c.tune;
  const std::string k(key);
  if (k == "max_blocks") *value = t.stream_max_blocks;
  else if (k == "unroll") *value = t.unroll;
  else if (k == "nontemporal") *value = t.nontemporal;
  else if (k == "mfma_blocks") *value = t.mfma_blocks;
  else if (k == "mfma_layout") *value = t.mfma_layout;
  else if (k == "tile_blocks") *value = t.tile_blocks;
  else if (k == "contig") *value = t.contig;
  else if (k == "xcd") *value = t.xcd;
  else if (k == "mq_lane") *value = t.mq_lane;
  else if (k == "diag_blocks") *value = t.diag_blocks;
  else return set_err(&h->c, kInvalidArg, "unknown tuning key: " + k);
  return kOk;
}

int qsvh_device_info(qsvh_ctx* h, char* name, size_t cap, int* cus, uint64_t* hbm) {
  if (!h) return kInvalidArg;
  hipDeviceProp_t prop;
  const hipError_t e = hipGetDeviceProperties(&prop, h->c.device);
  if (e != hipSuccess) return hip_err(&h->c, e, "device properties");
  if (name && cap) {
    std::strncpy(name, prop.name, cap - 1);
    name[cap - 1] = '\0';
  }
  if (cus) *cus = prop.multiProcessorCount;
  if (hbm) *hbm = prop.totalGlobalMem;
  return kOk;
}

// ------------------------------------------------------------------ gates / circuits
int qsvh_gate_matrix(const char* name, const double* params, uint32_t n_params, double* re,
                     double* im, uint32_t* m) {
  if (!name || !re || !im || !m) return kInvalidArg;
  Op op;
  const int st = make_gate(name, params, n_params, &op);
  if (st != kOk) return st;
  std::copy(op.re.begin(), op.re.end(), re);
  std::copy(op.im.begin(), op.im.end(), im);
  *m = op.m;
  return kOk;
}

int qsvh_circuit_create(uint32_t n_qubits, qsvh_circuit** out) {
  if (!out) return kInvalidArg;
  *out = new (std::nothrow) qsvh_circuit;
  if (!*out) return kOom;
  (*out)->c.n = n_qubits;
  return kOk;
}

void qsvh_circuit_destroy(qsvh_circuit* c) { delete c; }

int qsvh_circuit_push(qsvh_circuit* c, uint32_t m, const uint32_t* qubits, const double* re,
                      const double* im) {
  if (!c) return kInvalidArg;
  return push(&c->c, m, qubits, re, im);
}

int qsvh_circuit_push_gate(qsvh_circuit* c, const char* name, const uint32_t* qubits,
                           const double* params, uint32_t n_params) {
  if (!c) return kInvalidArg;
  return push_gate(&c->c, name, qubits, params, n_params);
}

int qsvh_circuit_from_qasm(const char* text, size_t len, qsvh_circuit** out, char* err,
                           size_t err_cap) {
  if (err && err_cap) err[0] = 0;
  if (!text || !out) return kInvalidArg;
  *out = nullptr;
  std::string msg;
  auto* c = new (std::nothrow) qsvh_circuit;
  if (!c) return kOom;
  const int st = parse_qasm(std::string(text, len), &c->c, &msg);
  if (st != kOk) {
    delete c;
    if (err && err_cap) {
      const size_t k = std::min(msg.size(), err_cap - 1);
      std::memcpy(err, msg.data(), k);
      err[k] = 0;
    }
    return st;
  }
  *out = c;
  return kOk;
}

uint32_t qsvh_circuit_n_qubits(const qsvh_circuit* c) { return c ? c->c.n : 0; }
size_t qsvh_circuit_len(const qsvh_circuit* c) { return c ? c->c.ops.size() : 0; }

int qsvh_circuit_op(const qsvh_circuit* c, size_t i, uint32_t* m, uint32_t* qubits, double* re,
                    double* im) {
  if (!c || i >= c->c.ops.size()) return kInvalidArg;
  const Op& op = c->c.ops[i];
  if (m) *m = op.m;
  if (qubits) std::copy(op.q, op.q + op.m, qubits);
  if (re) std::copy(op.re.begin(), op.re.end(), re);
  if (im) std::copy(op.im.begin(), op.im.end(), im);
  return kOk;
}

static int wrap(Circuit&& c, qsvh_circuit** out) {
  if (!out) return kInvalidArg;
  *out = new (std::nothrow) qsvh_circuit;
  if (!*out) return kOom;
  (*out)->c = std::move(c);
  return kOk;
}

int qsvh_circuit_ghz(uint32_t n, qsvh_circuit** out) {
  if (n < 1) return kInvalidArg;
  return wrap(gen_ghz(n), out);
}
int qsvh_circuit_ghz_chain(uint32_t n, qsvh_circuit** out) {
  if (n < 1) return kInvalidArg;
  return wrap(gen_ghz_chain(n), out);
}
int qsvh_circuit_qft(uint32_t n, qsvh_circuit** out) {
  if (n < 1) return kInvalidArg;
  return wrap(gen_qft(n), out);
}
int qsvh_circuit_random(uint32_t n, uint64_t n_gates, uint64_t seed, qsvh_circuit** out) {
  if (n < 2) return kInvalidArg;  // circuits.rs:80
  return wrap(gen_random(n, n_gates, seed), out);
}
int qsvh_circuit_layered(uint32_t n, uint32_t depth, uint64_t seed, qsvh_circuit** out) {
  if (n < 1) return kInvalidArg;
  return wrap(gen_layered(n, depth, seed), out);
}
int qsvh_circuit_hea(uint32_t n, uint32_t depth, uint64_t seed, qsvh_circuit** out) {
  if (n < 1) return kInvalidArg;
  return wrap(gen_hea(n, depth, seed), out);
}
int qsvh_circuit_hea_batch(size_t count, uint32_t depth, uint64_t seed, qsvh_circuit** out) {
  if (!out) return kInvalidArg;
  SplitMix64 rng(seed);
  for (size_t c = 0; c < count; ++c) {
    const uint32_t n = 10 + uint32_t(rng.below(13));
    const uint64_t s = rng.next_u64();
    const int rc = wrap(gen_hea(n, depth, s), &out[c]);
    if (rc != kOk) {
      for (size_t k = 0; k < c; ++k) qsvh_circuit_destroy(out[k]);
      return rc;
    }
  }
  return kOk;
}

int qsvh_circuit_mix(uint32_t n, uint64_t n_gates, uint64_t seed, qsvh_circuit** out) {
  if (n < 3) return kInvalidArg;
  return wrap(gen_mix(n, n_gates, seed), out);
}

int qsvh_fuse(qsvh_ctx* h, const qsvh_circuit* in, uint32_t max_qubits, qsvh_circuit** out) {
  if (!h || !in || !out) return kInvalidArg;
  *out = nullptr;
  if (max_qubits > kMaxGateQubits)
    return set_err(&h->c, kTooWide, "fusion groups are limited to 6 qubits");
  const Circuit& c = in->c;
  const std::vector<Group> groups = greedy_groups(c, max_qubits);
  std::vector<ComposeGroup> cg;
  std::vector<ComposeGate> gates;
  std::vector<double> mats;
  uint64_t out_off = 0;
  for (const Group& g : groups) {
    if (g.end - g.begin == 1) continue;
    ComposeGroup x;
    x.k = g.m;
    x.first_gate = uint32_t(gates.size());
    x.n_gates = uint32_t(g.end - g.begin);
    x.out_off = out_off;
    out_off += uint64_t(1) << (2 * g.m);
    for (uint64_t i = g.begin; i < g.end; ++i) {
      const Op& op = c.ops[i];
      ComposeGate y;
      y.m = op.m;
      for (uint32_t j = 0; j < op.m; ++j)
        y.local[j] = uint32_t(std::find(g.q, g.q + g.m, op.q[j]) - g.q);
      y.mat_off = mats.size() / 2;
      for (size_t e = 0; e < op.re.size(); ++e) {
        mats.push_back(op.re[e]);
        mats.push_back(op.im[e]);
      }
      gates.push_back(y);
    }
    cg.push_back(x);
  }
  std::vector<double> composed;
  const int rc = compose_groups(&h->c, cg, gates, mats, &composed);
  if (rc != kOk) return rc;
  Circuit res;
  res.n = c.n;
  size_t gi = 0;
  for (const Group& g : groups) {
    if (g.end - g.begin == 1) {
      res.ops.push_back(c.ops[g.begin]);
      continue;
    }
    const ComposeGroup& x = cg[gi++];
    Op op;
    op.m = g.m;
    std::copy(g.q, g.q + g.m, op.q);
    const size_t n = size_t(1) << (2 * g.m);
    op.re.resize(n);
    op.im.resize(n);
    for (size_t e = 0; e < n; ++e) {
      op.re[e] = composed[2 * (x.out_off + e)];
      op.im[e] = composed[2 * (x.out_off + e) + 1];
    }
    res.ops.push_back(std::move(op));
  }
  return wrap(std::move(res), out);
}

int qsvh_fuse_groups(const qsvh_circuit* in, uint32_t max_qubits, uint64_t* begin,
                     uint64_t* end, uint32_t* m, uint32_t* qubits, size_t cap,
                     size_t* n_groups) {
  if (!in || !n_groups) return kInvalidArg;
  if (max_qubits > kMaxGateQubits) return kTooWide;
  const std::vector<Group> groups = greedy_groups(in->c, max_qubits);
  *n_groups = groups.size();
  for (size_t g = 0; g < groups.size() && g < cap; ++g) {
    if (begin) begin[g] = groups[g].begin;
    if (end) end[g] = groups[g].end;
    if (m) m[g] = groups[g].m;
    if (qubits)
      for (uint32_t j = 0; j < kMaxGateQubits; ++j) qubits[6 * g + j] = j < groups[g].m ? groups[g].q[j] : 0;
  }
  return kOk;
}

// ------------------------------------------------------------------ states
int qsvh_state_alloc(qsvh_ctx* h, uint32_t n, int dtype, qsvh_state** out) {
  if (!h || !out) return kInvalidArg;
  *out = nullptr;
  if (dtype != QSVH_C128 && dtype != QSVH_C64) return set_err(&h->c, kDtype, "unknown dtype");
  if (n > 40) return set_err(&h->c, kOom, "state too large");
  qsvh_state* st = new (std::nothrow) qsvh_state;
  if (!st) return kOom;
  State& s = st->s;
  s.ctx = &h->c;
  s.n = n;
  s.dtype = dtype;
  hipError_t e = hipSetDevice(h->c.device);
  if (e == hipSuccess) e = hipMalloc(&s.amp, s.bytes());
  if (e != hipSuccess) {
    delete st;
    return hip_err(&h->c, e, "state alloc");
  }
  *out = st;
  return kOk;
}

void qsvh_state_free(qsvh_state* st) {
  if (!st) return;
  hipStreamSynchronize(st->s.ctx->stream);
  hipFree(st->s.amp);
  delete st;
}

uint32_t qsvh_state_n_qubits(const qsvh_state* st) { return st ? st->s.n : 0; }
int qsvh_state_dtype(const qsvh_state* st) { return st ? st->s.dtype : -1; }

int qsvh_init_basis(qsvh_state* st, uint64_t basis) {
  if (!st) return kInvalidArg;
  State& s = st->s;
  if (basis >= s.dim()) return set_err(s.ctx, kInvalidArg, "basis index out of range");
  return hip_err(s.ctx, launch_init_basis(s, basis, s.ctx->stream), "init_basis");
}

int qsvh_apply(qsvh_state* st, uint32_t m, const uint32_t* qubits, const double* re,
               const double* im) {
  if (!st || !re || !im) return kInvalidArg;
  State& s = st->s;
  Circuit c;
  c.n = s.n;
  const int v = push(&c, m, qubits, re, im);
  if (v != kOk) return set_err(s.ctx, v, "invalid gate");
  Program p;
  const int rc = program_build(&s, c, nullptr, &p);
  if (rc == kOk) {
    const int lr = program_launch(&p);
    program_free(&p);  // waits for the launch only when it owns device matrices
    return lr;
  }
  program_free(&p);
  return rc;
}

int qsvh_amplitude(qsvh_state* st, uint64_t index, double* re, double* im) {
  if (!st || !re || !im) return kInvalidArg;
  State& s = st->s;
  if (index >= s.dim()) return set_err(s.ctx, kInvalidArg, "amplitude index out of range");
  const size_t es = s.elem();
  alignas(16) unsigned char b[16];
  hipStream_t str = s.ctx->stream;
  hipError_t e = hipMemcpyAsync(b, static_cast<char*>(s.amp) + 2 * index * es, 2 * es,
                                hipMemcpyDeviceToHost, str);
  if (e == hipSuccess) e = hipStreamSynchronize(str);
  if (e != hipSuccess) return hip_err(s.ctx, e, "amplitude");
  if (s.dtype == 0) {
    std::memcpy(re, b, 8);
    std::memcpy(im, b + 8, 8);
  } else {
    float f[2];
    std::memcpy(f, b, 8);
    *re = f[0];
    *im = f[1];
  }
  return kOk;
}

int qsvh_probabilities_device(qsvh_state* st, void* dout) {
  if (!st || !dout) return kInvalidArg;
  return hip_err(st->s.ctx, launch_probs(st->s, dout, st->s.ctx->stream), "probabilities");
}

int qsvh_probabilities(qsvh_state* st, void* host_out) {
  if (!st || !host_out) return kInvalidArg;
  State& s = st->s;
  void* d = nullptr;
  const size_t bytes = size_t(s.dim()) * s.elem();
  hipError_t e = hipMalloc(&d, bytes);
  if (e != hipSuccess) return hip_err(s.ctx, e, "probabilities alloc");
  e = launch_probs(s, d, s.ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(host_out, d, bytes, hipMemcpyDeviceToHost, s.ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s.ctx->stream);
  hipFree(d);
  return hip_err(s.ctx, e, "probabilities");
}

static int reduce(State& s, uint64_t mask, double* out) {
  Ctx* c = s.ctx;
  hipError_t e = launch_reduce(s, mask, c->d_partials, Ctx::kMaxPartials, c->d_result, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->h_result, c->d_result, sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_err(c, e, "reduction");
  *out = *c->h_result;
  return kOk;
}

int qsvh_norm_sqr(qsvh_state* st, double* out) {
  if (!st || !out) return kInvalidArg;
  return reduce(st->s, 0, out);
}

int qsvh_expectation_z(qsvh_state* st, uint64_t mask, double* out) {
  if (!st || !out) return kInvalidArg;
  return reduce(st->s, mask & (st->s.dim() - 1), out);
}

int qsvh_sample(qsvh_state* st, uint64_t shots, uint64_t seed, uint64_t* out) {
  if (!st || (!out && shots > 0)) return kInvalidArg;
  State& s = st->s;
  if (shots > (uint64_t(1) << 31)) return set_err(s.ctx, kInvalidArg, "at most 2^31 shots per call");
  const uint64_t nc = sample_chunks(s.dim());
  if (nc > (uint64_t(1) << 23)) return set_err(s.ctx, kInvalidArg, "sampling supports up to 34 qubits");
  const uint64_t nb = sample_blocks(s.dim());
  const size_t bytes =
      sizeof(double) * (s.dim() + 2 * nc + 1 + nb) + sizeof(uint64_t) * shots;
  void* buf = nullptr;
  hipError_t e = hipMalloc(&buf, bytes);
  if (e != hipSuccess) return hip_err(s.ctx, e, "sample alloc");
  double* local = static_cast<double*>(buf);
  double* totals = local + s.dim();
  double* offset = totals + nc;
  double* bsum = offset + nc + 1;
  uint64_t* dout = reinterpret_cast<uint64_t*>(bsum + nb);
  e = launch_sample(s, shots, seed, local, totals, offset, bsum, dout, s.ctx->stream);
  if (e == hipSuccess && shots > 0)
    e = hipMemcpyAsync(out, dout, sizeof(uint64_t) * shots, hipMemcpyDeviceToHost,
                       s.ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s.ctx->stream);
  hipFree(buf);
  return hip_err(s.ctx, e, "sample");
}

int qsvh_sample_noalloc(qsvh_state* st, uint64_t shots, uint64_t seed, uint64_t* device_out) {
  if (!st || (!device_out && shots > 0)) return kInvalidArg;
  State& s = st->s;
  Ctx* c = s.ctx;
  if (shots > (uint64_t(1) << 31)) return set_err(c, kInvalidArg, "at most 2^31 shots per call");
  if (s.n > kNoAllocMaxQubits)
    return set_err(c, kInvalidArg,
                   "qsvh_sample_noalloc: at most 22 qubits (its scratch holds the CDF of 2^22 "
                   "amplitudes); use qsvh_sample");
  if (!c->sample_scratch)
    return set_err(c, kHip, std::string("sampler scratch address: ") +
                                hipGetErrorString(c->sample_scratch_err));
  // the layout of qsvh_sample's buffer, inside the static scratch
  const uint64_t nc = sample_chunks(s.dim()), nb = sample_blocks(s.dim());
  if (s.dim() + 2 * nc + 1 + nb > sample_noalloc_doubles())
    return set_err(c, kInvalidArg, "qsvh_sample_noalloc: state larger than the scratch");
  double* local = c->sample_scratch;
  double* totals = local + s.dim();
  double* offset = totals + nc;
  double* bsum = offset + nc + 1;
  return hip_err(c, launch_sample(s, shots, seed, local, totals, offset, bsum, device_out,
                                  c->stream),
                 "sample_noalloc");
}

int qsvh_download(qsvh_state* st, void* re, void* im) {
  if (!st || !re || !im) return kInvalidArg;
  State& s = st->s;
  hipError_t e = hipStreamSynchronize(s.ctx->stream);
  if (e == hipSuccess) e = download_split(s, re, im);
  return hip_err(s.ctx, e, "download");
}

int qsvh_upload(qsvh_state* st, const void* re, const void* im) {
  if (!st || !re || !im) return kInvalidArg;
  State& s = st->s;
  hipError_t e = hipStreamSynchronize(s.ctx->stream);
  if (e == hipSuccess) e = upload_join(s, re, im);
  return hip_err(s.ctx, e, "upload");
}

int qsvh_init_random_normal(qsvh_state* st, uint64_t seed) {
  if (!st) return kInvalidArg;
  State& s = st->s;
  Ctx* c = s.ctx;
  return hip_err(c, launch_init_random(s, seed, c->d_partials, Ctx::kMaxPartials, c->d_result,
                                       c->stream),
                 "init_random_normal");
}

int qsvh_compare(qsvh_state* a, qsvh_state* b, double* max_abs_err, double* fidelity) {
  if (!a || !b) return kInvalidArg;
  Ctx* c = a->s.ctx;
  if (a->s.n != b->s.n) return set_err(c, kInvalidArg, "compare: qubit counts differ");
  if (a->s.ctx != b->s.ctx) {
    // Different contexts: order b's queue before a's by waiting for it.
    const hipError_t e = hipStreamSynchronize(b->s.ctx->stream);
    if (e != hipSuccess) return hip_err(c, e, "compare");
  }
  double* d_out = nullptr;
  hipError_t e = hipMalloc(&d_out, 5 * sizeof(double));
  if (e != hipSuccess) return hip_err(c, e, "compare alloc");
  double h[5] = {0, 0, 0, 0, 0};
  e = launch_compare(a->s, b->s, c->d_partials, d_out, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(d_out);
  if (e != hipSuccess) return hip_err(c, e, "compare");
  if (max_abs_err) *max_abs_err = h[0];
  if (fidelity) {
    const double den = h[3] * h[4];
    *fidelity = den > 0 ? (h[1] * h[1] + h[2] * h[2]) / den : 0.0;
  }
  return kOk;
}

int qsvh_synchronize(qsvh_state* st) {
  if (!st) return kInvalidArg;
  return hip_err(st->s.ctx, hipStreamSynchronize(st->s.ctx->stream), "synchronize");
}

int qsvh_execute(qsvh_ctx* h, const qsvh_circuit* c, int dtype, qsvh_state** out) {
  if (!h || !c || !out) return kInvalidArg;
  int rc = qsvh_state_alloc(h, c->c.n, dtype, out);
  if (rc != kOk) return rc;
  rc = qsvh_init_basis(*out, 0);
  if (rc == kOk) {
    Program p;
    rc = program_build(&(*out)->s, c->c, nullptr, &p);
    if (rc == kOk) rc = program_launch(&p);
    program_free(&p);
  }
  if (rc != kOk) {
    qsvh_state_free(*out);
    *out = nullptr;
  }
  return rc;
}

// ------------------------------------------------------------------ programs
int qsvh_program_create(qsvh_state* st, const qsvh_circuit* c, const qsvh_program_opts* opts,
                        qsvh_program** out) {
  if (!st || !c || !out) return kInvalidArg;
  *out = nullptr;
  qsvh_program* p = new (std::nothrow) qsvh_program;
  if (!p) return kOom;
  const int rc = program_build(&st->s, c->c, opts, &p->p);
  if (rc != kOk) {
    program_free(&p->p);
    delete p;
    return rc;
  }
  *out = p;
  return kOk;
}

int qsvh_program_launch(qsvh_program* p) {
  if (!p) return kInvalidArg;
  return program_launch(&p->p);
}

int qsvh_program_time(qsvh_program* p, double* ms) {
  if (!p || !ms) return kInvalidArg;
  Ctx* c = p->p.st->ctx;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return set_err(c, kHip, "event");
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return set_err(c, kHip, "event");
  }
  hipError_t e = hipEventRecord(a, c->stream);
  int rc = kOk;
  if (e == hipSuccess) rc = program_launch(&p->p);
  if (e == hipSuccess && rc == kOk) e = hipEventRecord(b, c->stream);
  if (e == hipSuccess && rc == kOk) e = hipEventSynchronize(b);
  float f = 0.f;
  if (e == hipSuccess && rc == kOk) e = hipEventElapsedTime(&f, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (rc != kOk) return rc;
  if (e != hipSuccess) return hip_err(c, e, "program time");
  *ms = f;
  return kOk;
}

int qsvh_program_num_kernels(const qsvh_program* p, uint64_t* out) {
  if (!p || !out) return kInvalidArg;
  *out = p->p.launches.size();
  return kOk;
}

int qsvh_program_time_kernels(qsvh_program* h, double* ms, size_t cap, size_t* n) {
  if (!h || !n || (cap && !ms)) return kInvalidArg;
  Program* p = &h->p;
  Ctx* c = p->st->ctx;
  const size_t L = p->launches.size();
  *n = L;
  std::vector<hipEvent_t> ev(L + 1, nullptr);
  hipError_t e = hipSuccess;
  for (size_t j = 0; j <= L && e == hipSuccess; ++j) e = hipEventCreate(&ev[j]);
  for (size_t j = 0; j < L && e == hipSuccess; ++j) {
    e = hipEventRecord(ev[j], c->stream);
    if (e == hipSuccess) e = issue(*p, p->launches[j], c->tune, c->stream);
  }
  if (e == hipSuccess) e = hipEventRecord(ev[L], c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev[L]);
  for (size_t j = 0; j < L && j < cap && e == hipSuccess; ++j) {
    float f = 0.f;
    e = hipEventElapsedTime(&f, ev[j], ev[j + 1]);
    ms[j] = f;
  }
  for (hipEvent_t x : ev)
    if (x) hipEventDestroy(x);
  return hip_err(c, e, "program kernel times");
}

int qsvh_program_order(const qsvh_program* p, uint64_t* order, size_t cap, size_t* n) {
  if (!p || !n) return kInvalidArg;
  *n = p->p.order.size();
  for (size_t i = 0; i < p->p.order.size() && i < cap && order; ++i) order[i] = p->p.order[i];
  return kOk;
}

void qsvh_program_destroy(qsvh_program* p) {
  if (!p) return;
  program_free(&p->p);
  delete p;
}

static int time_launches(Ctx* c, uint32_t reps, double* ms, const std::function<hipError_t()>& f) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return set_err(c, kHip, "event");
  if (hipEventCreate(&b) != hipSuccess) {
    hipEventDestroy(a);
    return set_err(c, kHip, "event");
  }
  hipError_t e = f();  // warm-up
  if (e == hipSuccess) e = hipEventRecord(a, c->stream);
  for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) e = f();
  if (e == hipSuccess) e = hipEventRecord(b, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (e != hipSuccess) return hip_err(c, e, "timing");
  *ms = double(t) / double(reps ? reps : 1);
  return kOk;
}

int qsvh_time_gate(qsvh_state* st, uint32_t m, const uint32_t* qubits, const double* re,
                   const double* im, uint32_t reps, double* ms) {
  if (!st || !re || !im || !ms || reps == 0) return kInvalidArg;
  State& s = st->s;
  Circuit c;
  c.n = s.n;
  const int v = push(&c, m, qubits, re, im);
  if (v != kOk) return set_err(s.ctx, v, "invalid gate");
  Program p;
  int rc = program_build(&s, c, nullptr, &p);
  if (rc == kOk)
    rc = time_launches(s.ctx, reps, ms, [&]() { return issue_all(&p); });
  program_free(&p);
  return rc;
}

int qsvh_time_copy(qsvh_state* st, uint32_t reps, double* ms) {
  if (!st || !ms || reps == 0) return kInvalidArg;
  State& s = st->s;
  void* dst = nullptr;
  hipError_t e = hipMalloc(&dst, s.bytes());
  if (e != hipSuccess) return hip_err(s.ctx, e, "copy alloc");
  const int rc = time_launches(s.ctx, reps, ms,
                               [&]() { return launch_copy(s, dst, s.ctx->tune, s.ctx->stream); });
  hipStreamSynchronize(s.ctx->stream);
  hipFree(dst);
  return rc;
}

}  // extern "C"
