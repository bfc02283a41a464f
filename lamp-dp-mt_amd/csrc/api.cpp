// api.cpp -- C ABI of libdpx.so (declared in include/dpx.h).
//
// Host side of the backend: owns the device, a stream, the resident Turner
// tables and grow-only DP scratch, turns the reference-style entry points
// (one instance or many) into one persistent fill launch plus one finalize
// launch, and copies answers and traces back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/dpx.h"
#include "engine.h"
#include "rna_model.h"

using namespace dpx;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool ensure(size_t bytes) {
    if (bytes <= cap && p) return true;
    if (p) { (void)hipFree(p); p = nullptr; cap = 0; }
    size_t want = std::max<size_t>(bytes, 256);
    if (hipMalloc(&p, want) != hipSuccess) { p = nullptr; return false; }
    cap = want;
    return true;
  }
  void release() { if (p) (void)hipFree(p); p = nullptr; cap = 0; }
};

}  // namespace

struct dpx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::mutex mu;
  int cu_count = 0;
  int occ[3] = {0, 0, 0};
  // energy model
  bool has_rna = false;
  HostRna* rna = nullptr;
  DevBuf d_rna, d_hl;
  int hl_n = -1;
  // scratch
  DevBuf tab, queue, dep, ctrl, inst, seeds, codes, dims, results, traces, lens;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  float t_fill = 0, t_fin = 0;
  // SW scratch
  DevBuf sw_a, sw_b, sw_bt, sw_row, sw_flag, sw_best, sw_out, sw_tr;
};

namespace {

int fail(dpx_ctx* c, int code, const std::string& msg) {
  c->err = msg;
  return code;
}

int hip_fail(dpx_ctx* c, hipError_t e, const char* what) {
  return fail(c, DPX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return hip_fail(c, _e, #call); \
  } while (0)

int encode_rna(dpx_ctx* c, const char* seq, int n, uint8_t* out) {
  for (int q = 0; q < n; ++q) {
    switch (seq[q]) {
      case 'A': case 'a': out[q] = 1; break;
      case 'C': case 'c': out[q] = 2; break;
      case 'G': case 'g': out[q] = 3; break;
      case 'U': case 'u': case 'T': case 't': out[q] = 4; break;
      default: return fail(c, DPX_EINVAL, std::string("symbol outside ACGU at position ") + std::to_string(q));
    }
  }
  return DPX_OK;
}

int ensure_hl(dpx_ctx* c, int n_max) {
  if (n_max <= c->hl_n) return DPX_OK;
  const int len = std::max(n_max, 64) + 1;
  std::vector<int32_t> hl(len);
  for (int l = 0; l < len; ++l) hl[l] = l < 3 ? 0 : c->rna->hl_ent(l);
  if (!c->d_hl.ensure(sizeof(int32_t) * len)) return fail(c, DPX_ENOMEM, "hairpin table");
  HIPCHK(hipMemcpy(c->d_hl.p, hl.data(), sizeof(int32_t) * len, hipMemcpyHostToDevice));
  c->hl_n = len - 1;
  return DPX_OK;
}

int nt_of(int g) { return g == G_ZUKER ? NT_ZU : g == G_NUSSINOV ? NT_NU : NT_MC; }

// One fill + finalize launch over `count` instances of grammar g.
// lens[q]: instance size; data: codes (RNA) or rows|cols (chain) per instance.
int run_tri(dpx_ctx* c, int g, int count, const std::vector<int>& lens, const std::vector<int64_t>& data_off,
            const void* data, size_t data_bytes, int min_loop, int32_t* results, dpx_trace_t* traces,
            const int64_t* trace_offs, int32_t* trace_lens) {
  if (count <= 0) return DPX_OK;
  const int NTv = nt_of(g);
  std::vector<TriInst> inst(count);
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  // seed the biggest instances first: their critical path is the longest
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return lens[x] > lens[y]; });
  int64_t tiles = 0, trace_total = 0;
  int n_max = 0;
  const bool want_trace = traces != nullptr;
  for (int q = 0; q < count; ++q) {
    TriInst& t = inst[q];
    t.n = lens[q];
    t.nb = (t.n + 1 + T - 1) / T;
    t.seq_off = data_off[q];
    t.tile0 = tiles;
    tiles += tiles_of(t.nb);
    t.trace_off = trace_total;
    t.trace_cap = want_trace ? 4 * t.n + 8 : 0;
    trace_total += t.trace_cap;
    t.out = q;
    n_max = std::max(n_max, t.n);
  }
  if (tiles >= (int64_t)1 << 31) return fail(c, DPX_ENOMEM, "too many tiles in one launch");
  std::vector<unsigned long long> seeds;
  seeds.reserve(tiles);
  for (int q : order)
    for (int I = 0; I < inst[q].nb; ++I) seeds.push_back(pack_item(q, I, I));
  const size_t tab_bytes = (size_t)tiles * NTv * T * T * sizeof(int32_t);
  if (!c->tab.ensure(tab_bytes) || !c->queue.ensure(tiles * 8) || !c->dep.ensure(tiles * 4) || !c->ctrl.ensure(64) ||
      !c->inst.ensure(sizeof(TriInst) * count) || !c->seeds.ensure(seeds.size() * 8) || !c->codes.ensure(data_bytes + 64) ||
      !c->results.ensure(sizeof(int32_t) * count) || !c->lens.ensure(sizeof(int32_t) * count) ||
      (want_trace && !c->traces.ensure(sizeof(dpx_trace_t) * std::max<int64_t>(trace_total, 1))))
    return fail(c, DPX_ENOMEM, "device allocation for the DP tables failed (" + std::to_string(tab_bytes >> 20) + " MiB)");
  if (g == G_ZUKER) {
    int r = ensure_hl(c, n_max);
    if (r) return r;
    if (n_max + 1 > 16000) return fail(c, DPX_EINVAL, "Zuker sequences are limited to 15999 bases");
  }
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(c->inst.p, inst.data(), sizeof(TriInst) * count, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->seeds.p, seeds.data(), seeds.size() * 8, hipMemcpyHostToDevice, s));
  if (data_bytes) HIPCHK(hipMemcpyAsync(c->codes.p, data, data_bytes, hipMemcpyHostToDevice, s));
  TriArgs a{};
  a.inst = (const TriInst*)c->inst.p;
  a.n_inst = count;
  a.total_tiles = (int32_t)tiles;
  a.queue = (unsigned long long*)c->queue.p;
  a.ctrl = (int32_t*)c->ctrl.p;
  a.dep = (int32_t*)c->dep.p;
  a.tab = (int32_t*)c->tab.p;
  a.codes = (const uint8_t*)c->codes.p;
  a.dims = (const int32_t*)c->codes.p;
  a.rna = c->d_rna.p;
  a.hl_ent = (const int32_t*)c->d_hl.p;
  a.min_loop = min_loop;
  a.results = (int32_t*)c->results.p;
  a.traces = c->traces.p;
  a.trace_lens = (int32_t*)c->lens.p;
  a.want_trace = want_trace ? 1 : 0;
  if (!c->occ[g]) c->occ[g] = fill_occupancy(g);
  const int per_cu = std::max(1, c->occ[g]);
  const int grid = (int)std::min<int64_t>(tiles, (int64_t)per_cu * c->cu_count);
  HIPCHK(hipEventRecord(c->ev[0], s));
  if (launch_fill(g, a, (const unsigned long long*)c->seeds.p, (int)seeds.size(), grid, s)) return hip_fail(c, hipGetLastError(), "fill launch");
  HIPCHK(hipEventRecord(c->ev[1], s));
  if (launch_finalize(g, a, count, n_max, s)) return hip_fail(c, hipGetLastError(), "finalize launch");
  HIPCHK(hipEventRecord(c->ev[2], s));
  std::vector<int32_t> res(count), tl(count);
  int32_t ctrl[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(ctrl, c->ctrl.p, sizeof ctrl, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(res.data(), c->results.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
  if (want_trace) HIPCHK(hipMemcpyAsync(tl.data(), c->lens.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  (void)hipEventElapsedTime(&c->t_fill, c->ev[0], c->ev[1]);
  (void)hipEventElapsedTime(&c->t_fin, c->ev[1], c->ev[2]);
  if (ctrl[2]) return fail(c, DPX_EHIP, "tile queue wait timed out (scheduler error)");
  int rc = DPX_OK;
  for (int q = 0; q < count; ++q) results[q] = res[q];
  if (want_trace) {
    for (int q = 0; q < count; ++q) {
      const int64_t cap = trace_offs[q + 1] - trace_offs[q];
      if (tl[q] < 0) { trace_lens[q] = -1; rc = fail(c, DPX_ETRACE, "backtrace failed"); continue; }
      trace_lens[q] = tl[q];
      if (tl[q] > cap) { rc = fail(c, DPX_ETRACE, "trace buffer too small"); continue; }
      HIPCHK(hipMemcpyAsync(traces + trace_offs[q], (const dpx_trace_t*)c->traces.p + inst[q].trace_off,
                            sizeof(dpx_trace_t) * tl[q], hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
  }
  return rc;
}

int rna_batch(dpx_ctx* c, int g, const char* seqs, const int64_t* offs, int count, int min_loop, int32_t* results,
              dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens) {
  if (count < 0 || (count > 0 && (!seqs || !offs || !results))) return fail(c, DPX_EINVAL, "null argument");
  if (traces && (!trace_offs || !trace_lens)) return fail(c, DPX_EINVAL, "trace offsets missing");
  if (g == G_ZUKER && !c->has_rna) return fail(c, DPX_ENOPARAMS, "dpx_set_rna_params was not called");
  std::vector<int> lens(count);
  std::vector<int64_t> off(count);
  const int64_t total = count ? offs[count] - offs[0] : 0;
  std::vector<uint8_t> codes(total + count + 64);
  int64_t pos = 0;
  for (int q = 0; q < count; ++q) {
    const int64_t n = offs[q + 1] - offs[q];
    if (n < 0 || n > 30000) return fail(c, DPX_EINVAL, "sequence length out of range");
    lens[q] = (int)n;
    off[q] = pos;
    int r = encode_rna(c, seqs + offs[q], (int)n, codes.data() + pos);
    if (r) return r;
    pos += n + 1;  // one spare slot per instance (st scratch shares the offsets)
  }
  return run_tri(c, g, count, lens, off, codes.data(), (size_t)pos, min_loop, results, traces, trace_offs, trace_lens);
}

}  // namespace

extern "C" {

int dpx_create(dpx_ctx** out, int device) {
  if (!out) return DPX_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return DPX_EHIP;
  if (device < 0 || device >= ndev) return DPX_EINVAL;
  dpx_ctx* c = new dpx_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DPX_EHIP;
  }
  (void)hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  *out = c;
  return DPX_OK;
}

void dpx_destroy(dpx_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  DevBuf* bufs[] = {&c->d_rna, &c->d_hl, &c->tab, &c->queue, &c->dep, &c->ctrl, &c->inst, &c->seeds, &c->codes,
                    &c->dims, &c->results, &c->traces, &c->lens, &c->sw_a, &c->sw_b, &c->sw_bt, &c->sw_row,
                    &c->sw_flag, &c->sw_best, &c->sw_out, &c->sw_tr};
  for (DevBuf* b : bufs) b->release();
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c->rna;
  delete c;
}

const char* dpx_last_error(const dpx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dpx_set_rna_params(dpx_ctx* c, const char* path) {
  if (!c || !path) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  HostRna* h = new HostRna();
  char msg[256] = {0};
  if (load_par_file(path, h, msg, sizeof msg)) {
    delete h;
    return fail(c, DPX_EPARAMS, msg);
  }
  delete c->rna;
  c->rna = h;
  if (!c->d_rna.ensure(sizeof(DevRna))) return fail(c, DPX_ENOMEM, "energy tables");
  HIPCHK(hipMemcpy(c->d_rna.p, &h->dev, sizeof(DevRna), hipMemcpyHostToDevice));
  c->hl_n = -1;
  c->has_rna = true;
  return ensure_hl(c, 4096);
}

int dpx_zuker_batch(dpx_ctx* c, const char* seqs, const int64_t* offs, int count, int32_t* results, dpx_trace_t* traces,
                    const int64_t* trace_offs, int32_t* trace_lens) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  return rna_batch(c, G_ZUKER, seqs, offs, count, 0, results, traces, trace_offs, trace_lens);
}

int dpx_nussinov_batch(dpx_ctx* c, const char* seqs, const int64_t* offs, int count, int min_loop, int32_t* results,
                       dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens) {
  if (!c) return DPX_EINVAL;
  if (min_loop < 0) return fail(c, DPX_EINVAL, "min_loop < 0");
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  return rna_batch(c, G_NUSSINOV, seqs, offs, count, min_loop, results, traces, trace_offs, trace_lens);
}

static int single_rna(dpx_ctx* c, int g, const char* seq, int n, int min_loop, int32_t* score, dpx_trace_t* trace,
                      int trace_cap, int* trace_len) {
  if (!c) return DPX_EINVAL;
  if (!seq || n < 0 || !score || (trace && (!trace_len || trace_cap < 0))) return fail(c, DPX_EINVAL, "bad argument");
  int64_t offs[2] = {0, n};
  int64_t toffs[2] = {0, trace_cap};
  int32_t tl = 0, res = 0;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  int r = rna_batch(c, g, seq, offs, 1, min_loop, &res, trace, toffs, &tl);
  *score = res;
  if (trace) *trace_len = tl;
  return r;
}

int dpx_nussinov(dpx_ctx* c, const char* seq, int n, int min_loop, int32_t* score, dpx_trace_t* trace, int trace_cap,
                 int* trace_len) {
  if (c && min_loop < 0) return fail(c, DPX_EINVAL, "min_loop < 0");
  return single_rna(c, G_NUSSINOV, seq, n, min_loop, score, trace, trace_cap, trace_len);
}

int dpx_zuker(dpx_ctx* c, const char* seq, int n, int32_t* mfe, dpx_trace_t* trace, int trace_cap, int* trace_len) {
  return single_rna(c, G_ZUKER, seq, n, 0, mfe, trace, trace_cap, trace_len);
}

int dpx_matchain(dpx_ctx* c, const int32_t* rows, const int32_t* cols, int n, int32_t* cost, dpx_trace_t* trace,
                 int trace_cap, int* trace_len) {
  if (!c) return DPX_EINVAL;
  if (!rows || !cols || n < 1 || !cost || (trace && (!trace_len || trace_cap < 0))) return fail(c, DPX_EINVAL, "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  std::vector<int32_t> dims(2 * (size_t)n);
  std::copy(rows, rows + n, dims.begin());
  std::copy(cols, cols + n, dims.begin() + n);
  std::vector<int> lens{n};
  std::vector<int64_t> off{0};
  int64_t toffs[2] = {0, trace_cap};
  int32_t tl = 0, res = 0;
  int r = run_tri(c, G_MATCHAIN, 1, lens, off, dims.data(), dims.size() * sizeof(int32_t), 0, &res, trace, toffs, &tl);
  *cost = res;
  if (trace) *trace_len = tl;
  return r;
}

int dpx_swat(dpx_ctx* c, const char* a, int n1, const char* b, int n2, const dpx_swat_scores* sc, int32_t* score,
             int32_t* best, int32_t* best_i, int32_t* best_j, int trace_from, dpx_trace_t* trace, int trace_cap,
             int* trace_len) {
  if (!c) return DPX_EINVAL;
  if (!a || !b || n1 < 0 || n2 < 0 || !sc || !score || (trace && (!trace_len || trace_cap < 0)) ||
      (trace_from != 0 && trace_from != 1))
    return fail(c, DPX_EINVAL, "bad argument");
  if ((int64_t)(n1 + 1) * (n2 + 1) >= ((int64_t)1 << 32)) return fail(c, DPX_EINVAL, "matrix too large");
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  const int W = n2 + 1;
  // row 0 (i == 0): g1 = 0 [4]; g2 chain [7/8/9]; M = first of g1 [2], g2 if strictly greater [3]
  std::vector<int2> row0(W);
  std::vector<uint8_t> r0m(W), r0g2(W);
  int Mprev = 0, G2prev = 0, best0 = -1, best0j = 0;
  for (int j = 0; j < W; ++j) {
    int g2, rg2;
    if (j == 0) { g2 = 0; rg2 = 7; }
    else {
      const int o = std::max(0, Mprev + sc->gap_open), e = std::max(0, G2prev + sc->gap_ext);
      rg2 = e > o ? 9 : 8; g2 = e > o ? e : o;
    }
    int M = 0, rm = j == 0 ? 0 : 2;
    if (g2 > M) { M = g2; rm = 3; }
    row0[j] = make_int2(M, 0);
    r0m[j] = (uint8_t)rm; r0g2[j] = (uint8_t)rg2;
    if (M > best0) { best0 = M; best0j = j; }
    Mprev = M; G2prev = g2;
  }
  SwArgs A{};
  A.n1 = n1; A.n2 = n2;
  A.match = sc->match; A.mismatch = sc->mismatch; A.open = sc->gap_open; A.ext = sc->gap_ext;
  A.nbands = (n1 + 63) / 64;
  A.nwords = (n2 + 64 + 7) / 8;
  const size_t bt_bytes = (size_t)A.nbands * A.nwords * 64 * 4;
  if (!c->sw_a.ensure(n1 + 64) || !c->sw_b.ensure(n2 + 64) || !c->sw_row.ensure((size_t)(A.nbands + 2) * W * 8 + 2 * (size_t)W + 64) ||
      !c->sw_flag.ensure(4 * (A.nbands + 2) + 64) || !c->sw_best.ensure(64) || !c->sw_bt.ensure(bt_bytes) ||
      (trace && !c->sw_tr.ensure(sizeof(dpx_trace_t) * (size_t)(trace_cap + 1))))
    return fail(c, DPX_ENOMEM, "device allocation for the alignment failed");
  hipStream_t s = c->stream;
  char* rowbase = (char*)c->sw_row.p;
  A.hrow = (int2*)rowbase;
  int2* d_row0 = (int2*)(rowbase + (size_t)(A.nbands + 1) * W * 8);  // scratch after the rows
  uint8_t* d_r0m = (uint8_t*)(d_row0 + W);
  uint8_t* d_r0g2 = d_r0m + W;
  A.a = (const uint8_t*)c->sw_a.p;
  A.b = (const uint8_t*)c->sw_b.p;
  A.prog = (int*)c->sw_flag.p;
  A.ctrl = A.prog + A.nbands + 1;
  A.bt = (uint32_t*)c->sw_bt.p;
  A.best = (unsigned long long*)c->sw_best.p;
  A.out = (int*)((char*)c->sw_best.p + 8);
  if (n1) HIPCHK(hipMemcpyAsync(c->sw_a.p, a, n1, hipMemcpyHostToDevice, s));
  if (n2) HIPCHK(hipMemcpyAsync(c->sw_b.p, b, n2, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_row0, row0.data(), sizeof(int2) * W, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_r0m, r0m.data(), W, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_r0g2, r0g2.data(), W, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(c->ev[0], s));
  if (A.nbands > 0) {
    static int occ = 0;
    if (!occ) occ = std::max(1, sw_occupancy());
    const int grid = std::min(A.nbands, occ * c->cu_count);
    if (sw_launch(A, d_row0, grid, s)) return hip_fail(c, hipGetLastError(), "sw launch");
  }
  HIPCHK(hipEventRecord(c->ev[1], s));
  unsigned long long key = 0;
  int mlast = 0, flags[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(&key, A.best, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&mlast, A.out, 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(flags, A.ctrl, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (A.nbands > 0 && flags[1]) return fail(c, DPX_EHIP, "band wait timed out (scheduler error)");
  *score = n1 == 0 ? row0[n2].x : mlast;
  int bv = best0, bi = 0, bj = best0j;
  if (A.nbands > 0) {
    const int dv = (int)(key >> 32);
    const uint32_t lin = 0xffffffffu - (uint32_t)(key & 0xffffffffu);
    if (dv > bv) { bv = dv; bi = (int)(lin / (uint32_t)W); bj = (int)(lin % (uint32_t)W); }
  }
  if (best) *best = bv;
  if (best_i) *best_i = bi;
  if (best_j) *best_j = bj;
  if (trace) {
    const int i0 = trace_from ? bi : n1, j0 = trace_from ? bj : n2;
    int* d_len = (int*)((char*)c->sw_best.p + 16);
    if (sw_trace_launch(A, d_r0m, d_r0g2, i0, j0, c->sw_tr.p, trace_cap, d_len, s)) return hip_fail(c, hipGetLastError(), "sw trace");
    HIPCHK(hipEventRecord(c->ev[2], s));
    int tl = 0;
    HIPCHK(hipMemcpyAsync(&tl, d_len, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (tl < 0) { *trace_len = -1; return fail(c, DPX_ETRACE, "trace buffer too small"); }
    HIPCHK(hipMemcpy(trace, c->sw_tr.p, sizeof(dpx_trace_t) * tl, hipMemcpyDeviceToHost));
    *trace_len = tl;
    (void)hipEventElapsedTime(&c->t_fin, c->ev[1], c->ev[2]);
  } else {
    c->t_fin = 0;
  }
  (void)hipEventElapsedTime(&c->t_fill, c->ev[0], c->ev[1]);
  return DPX_OK;
}

int dpx_last_timing(const dpx_ctx* c, float* fill_ms, float* finalize_ms) {
  if (!c) return DPX_EINVAL;
  if (fill_ms) *fill_ms = c->t_fill;
  if (finalize_ms) *finalize_ms = c->t_fin;
  return DPX_OK;
}

}  // extern "C"
