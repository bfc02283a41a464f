// api.cpp -- C ABI of libdpx.so (declared in include/dpx.h).
//
// Host side of the backend.  A context owns the device, the resident Turner
// tables and one workspace (stream, events, pinned staging) for the
// synchronous entry points.  Every computation is a Job: prepared on the
// host, given one contiguous block of device memory laid out by the job
// itself, launched asynchronously on a workspace (copies in, fill kernel,
// finalize kernel, copies out into pinned memory, one completion event) and
// harvested once that event has completed.  The server (continuous batching)
// runs jobs on several workspaces at once, carving their device blocks out of
// one arena whose size is the HBM budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/dpx.h"
#include "adp.h"
#include "engine.h"
#include "rna_model.h"

using namespace dpx;

namespace {

constexpr int kMaxZuker = 14000;      // finalize keeps the st column (n+1 int32), the sequence and the energy tables' head in LDS (87 KiB at this n)
constexpr int kMaxTri = 30000;        // Nussinov / matrix chain
constexpr int kHairpinTable = 32768;  // hairpin entropy entries kept on the device

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

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

struct HostBuf {  // pinned host memory, so that the copies are asynchronous
  void* p = nullptr;
  size_t cap = 0;
  bool ensure(size_t bytes) {
    if (bytes <= cap && p) return true;
    if (p) { (void)hipHostFree(p); p = nullptr; cap = 0; }
    size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 16);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) { p = nullptr; return false; }
    cap = want;
    return true;
  }
  void release() { if (p) (void)hipHostFree(p); p = nullptr; cap = 0; }
};

struct Workspace {
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  HostBuf h_in, h_out;
  bool init() {
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return false;
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) return false;
    return true;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    h_in.release();
    h_out.release();
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

enum { JOB_TRI = 0, JOB_SW = 1, JOB_SWB = 2 };
// device block of a triangular job
enum {
  TL_TAB, TL_QUEUE, TL_DEP, TL_CTRL, TL_INST, TL_SEEDS, TL_DATA, TL_SPEC, TL_RES, TL_LENS, TL_TRACES,
  TL_MP, TL_N
};
// device block of an alignment job
enum { SL_A, SL_B, SL_ROWS, SL_FLAGS, SL_BEST, SL_BT, SL_CP, SL_TR, SL_N };

// One alignment of a batched launch (JOB_SWB): its inputs, row 0 and where
// its parts lie in the job's device block.
struct SwPairJob {
  std::string a, b;
  dpx_swat_scores sc{};
  int trace_from = 0, cap = 0;
  std::vector<int2> row0;
  std::vector<uint8_t> r0m, r0g2;
  int row0_best = 0, row0_bestj = 0, row0_axiom = 0;
  int nbands = 0, band0 = 0, nslots = 0, nwords = 0;
  size_t in_a = 0, in_b = 0, in_row0 = 0;  // inputs (offsets in the block)
  size_t ring = 0, codes = 0;              // boundary-row ring, traceback codes
  int64_t tr0 = 0;                         // first trace record in the batch's trace region
  int64_t cells = 0;
};

struct Job {
  int kind = JOB_TRI;
  size_t off[TL_N] = {0};
  size_t dev_bytes = 0;  // size of the job's device block
  // triangular grammars
  int g = G_ZUKER, count = 0, n_max = 0, min_loop = 0;
  bool want_trace = false;
  int64_t tiles = 0, trace_total = 0;
  // split middle products (launches of few instances): queue items, partial
  // slots and each tile's first slot (-1: not split)
  bool mp_split = false;
  int64_t items = 0, mp_slots = 0;
  std::vector<int32_t> mp_base;
  std::vector<TriInst> inst;  // longest first (TriInst::out is the caller's index)
  std::vector<int64_t> trace_off_by_out;
  std::vector<unsigned long long> seeds;
  std::vector<uint8_t> data;  // codes (RNA) or rows|cols (int32, matrix chain)
  // Smith-Waterman
  std::string a, b;
  dpx_swat_scores sc{};
  int trace_from = 0, sw_cap = 0;
  int sw_seg = 0;  // bands per checkpoint segment (0: traceback codes of every band kept)
  std::vector<int2> row0;
  std::vector<uint8_t> r0m, r0g2;
  int row0_best = 0, row0_bestj = 0, row0_axiom = 0;
  // batched Smith-Waterman (JOB_SWB): the alignments and the block's regions
  std::vector<SwPairJob> swb;
  int swb_bands = 0;
  bool swb_trace = false;
  size_t swb_desc = 0, swb_bandtab = 0, swb_claim = 0, swb_in0 = 0, swb_in1 = 0, swb_out = 0, swb_tr = 0;
  // outputs: views into the workspace's pinned buffer, valid after the event
  const int32_t* res = nullptr;
  const int32_t* lens = nullptr;
  const int32_t* flags = nullptr;
  const dpx_trace_t* tr = nullptr;
  float t_fill = 0, t_fin = 0;
  int64_t cells = 0;
};

}  // namespace

struct dpx_ctx {
  int device = 0;
  std::string err;
  std::mutex mu;
  int cu_count = 0;
  int occ[N_GRAMMARS] = {1, 1, 1, 1};
  int sw_occ = 1;
  bool has_rna = false;
  HostRna* rna = nullptr;
  DevBuf d_rna, d_hl;
  Workspace ws;
  DevBuf block;  // device block of the synchronous entry points (grow-only)
  float t_fill = 0, t_fin = 0;
  std::atomic<int> servers{0};  // servers and plans hold the energy tables' address
  std::atomic<int> plans{0};
  // launch shape per request kind (DPX_REQ_*) and length bucket; read by
  // server threads at launch time, hence its own lock
  std::mutex tune_mu;
  dpx_tuning tune[6][DPX_TUNE_BUCKETS] = {};
  int64_t sw_code_budget = (int64_t)32 << 30;  // traceback codes of one alignment before checkpointing
};

// Scratch of dpx_zuker_noalloc: the device block (tile records, queue,
// inputs, results, trace) of one Zuker instance, n <= kNoAllocMaxZuker.  Each
// device holds its own copy until the process ends; host code uses the
// address hipGetSymbolAddress returns on that device (g_scratch_ptr), never
// the array's name.
constexpr size_t kScratchBytes = (size_t)16 << 20;
constexpr int kNoAllocMaxZuker = 1024;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;                     // guards g_scratch_ptr / g_scratch_rc
void* g_scratch_ptr[kMaxDevices] = {};       // address of `scratch` on each device
hipError_t g_scratch_rc[kMaxDevices] = {};   // result of hipGetSymbolAddress there
std::mutex g_scratch_call_mu[kMaxDevices];   // one dpx_zuker_noalloc at a time per device

// Called by dpx_create with `device` current: takes the scratch address of
// the device once.
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_rc[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = p;
}

int fail(std::string& err, int code, const std::string& msg) {
  err = msg;
  return code;
}

int hip_fail(std::string& err, hipError_t e, const char* what) {
  return fail(err, DPX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(call)                                       \
  do {                                                     \
    hipError_t _e = (call);                                \
    if (_e != hipSuccess) return hip_fail(err, _e, #call); \
  } while (0)

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int encode_rna(std::string& err, const char* seq, int n, uint8_t* out) {
  for (int q = 0; q < n; ++q) {
    switch (seq[q]) {
      case 'A': case 'a': out[q] = 1; break;
      case 'C': case 'c': out[q] = 2; break;
      case 'G': case 'g': out[q] = 3; break;
      case 'U': case 'u': case 'T': case 't': out[q] = 4; break;
      default: return fail(err, DPX_EINVAL, std::string("symbol outside ACGU at position ") + std::to_string(q));
    }
  }
  return DPX_OK;
}

int nt_of(int g) { return g == G_ZUKER ? NT_ZU : g == G_NUSSINOV ? NT_NU : g == G_RNAFOLD ? NT_RF : NT_MC; }
int64_t tri_cells(int n) { return (int64_t)(n + 1) * (n + 2) / 2; }
int tri_trace_cap(int n) { return 4 * n + 8; }
int max_len(int g) { return g == G_ZUKER ? kMaxZuker : kMaxTri; }
// grammar of a folding request kind (DPX_REQ_NUSSINOV / ZUKER / RNAFOLD)
int grammar_of(int kind) {
  return kind == DPX_REQ_NUSSINOV ? G_NUSSINOV : kind == DPX_REQ_RNAFOLD ? G_RNAFOLD : G_ZUKER;
}
int kind_of(int g) {
  return g == G_NUSSINOV ? DPX_REQ_NUSSINOV : g == G_RNAFOLD ? DPX_REQ_RNAFOLD : g == G_MATCHAIN ? DPX_REQ_MATCHAIN
                                                                                                   : DPX_REQ_ZUKER;
}
// length bucket: n <= 64 -> 0, (64,128] -> 1, ... (4096, inf) -> 7
int tune_bucket(int64_t n) {
  int b = 0;
  for (int64_t hi = 64; n > hi && b < DPX_TUNE_BUCKETS - 1; hi *= 2) ++b;
  return b;
}

// ---------------------------------------------------------------- launch shape

dpx_tuning get_tune(dpx_ctx* c, int kind, int bucket) {
  std::lock_guard<std::mutex> lk(c->tune_mu);
  return c->tune[kind][bucket];
}

// the length bucket holding most of a triangular job's cells
int job_bucket(const Job& j) {
  int64_t w[DPX_TUNE_BUCKETS] = {};
  for (const TriInst& t : j.inst) w[tune_bucket(t.n)] += tri_cells(t.n);
  return (int)(std::max_element(w, w + DPX_TUNE_BUCKETS) - w);
}

// persistent workgroups of a triangular launch: tuned workgroups per CU
// (0 = the kernel's occupancy) times the CU count, at most one per tile
int tri_grid_for(dpx_ctx* c, int g, int bucket, int64_t tiles) {
  const dpx_tuning t = get_tune(c, kind_of(g), bucket);
  const int occ = c->occ[g];
  const int w = t.wg_per_cu >= 1 && t.wg_per_cu <= occ ? t.wg_per_cu : occ;
  return (int)std::min<int64_t>(tiles, (int64_t)w * c->cu_count);
}

int tri_grid(dpx_ctx* c, const Job& j) { return tri_grid_for(c, j.g, job_bucket(j), j.tiles); }

// ---------------------------------------------------------------- layouts

// Device block of a triangular launch over instances with the given totals;
// mp_slots > 0: the middle products are split (queue of `items` slots).
size_t tri_layout(int g, int64_t tiles, int count, size_t data_bytes, int64_t trace_total, size_t* off,
                  int64_t items = 0, int64_t mp_slots = 0) {
  size_t sz[TL_N];
  sz[TL_TAB] = (size_t)tiles * nt_of(g) * T * T * sizeof(int32_t);
  sz[TL_QUEUE] = (size_t)std::max(tiles, items) * 8;
  sz[TL_DEP] = (size_t)tiles * 4;
  sz[TL_CTRL] = 64;
  sz[TL_INST] = sizeof(TriInst) * (size_t)count;
  sz[TL_SEEDS] = (size_t)tiles * 8;
  sz[TL_DATA] = data_bytes + 64;
  // special-hairpin energies by sequence position (RNA energy grammars)
  sz[TL_SPEC] = (g == G_ZUKER || g == G_RNAFOLD) ? 3 * sizeof(int16_t) * data_bytes : 0;
  sz[TL_RES] = 4 * (size_t)count;
  sz[TL_LENS] = 4 * (size_t)count;
  sz[TL_TRACES] = sizeof(dpx_trace_t) * (size_t)trace_total;
  sz[TL_MP] = mp_slots ? (size_t)(mp_part_offset(tiles) + mp_slots * T * T) * sizeof(int32_t) : 0;
  size_t at = 0;
  for (int k = 0; k < TL_N; ++k) {
    if (off) off[k] = at;
    at += a256(sz[k]);
  }
  return at;
}

// footprint of one instance launched alone; a batch needs at most the sum
size_t tri_footprint(int g, int n, bool trace) {
  const int nb = (n + 1 + T - 1) / T;
  const size_t data = g == G_MATCHAIN ? 8 * (size_t)n : (size_t)n + 1;
  return tri_layout(g, tiles_of(nb), 1, data, trace ? tri_trace_cap(n) : 0, nullptr);
}

// boundary-row ring of an alignment: more slots than bands can be in flight
constexpr int kSwMaxGrid = 4096;
int sw_slots(int nbands) {
  const int s = std::min(nbands + 1, kSwMaxGrid + 2);
  return s % SW_TAGS == 0 ? s + 1 : s;  // consecutive users of a slot carry different tags
}
// bands per checkpoint segment when the codes of every band would not fit the
// budget: about sqrt(nbands), balancing checkpoint rows against one segment's codes
int sw_segment(int n1, int n2, int64_t code_budget) {
  const int64_t nbands = (n1 + 63) / 64, nwords = sw_code_words(n2);
  if (nbands * nwords * 64 * 4 <= code_budget) return 0;
  return std::max(1, (int)std::ceil(std::sqrt((double)nbands)));
}

size_t sw_layout(int n1, int n2, int cap, int seg, size_t* off) {
  const size_t nbands = (n1 + 63) / 64, nwords = sw_code_words(n2), W = (size_t)n2 + 1;
  const size_t slots = sw_slots((int)nbands);
  size_t sz[SL_N];
  sz[SL_A] = (size_t)n1 + 64;
  sz[SL_B] = (size_t)n2 + 64;
  sz[SL_ROWS] = slots * (W + 1) * 8 + W * 8 + 2 * W + 64;  // boundary-row ring, then row 0 and its rule bytes
  sz[SL_FLAGS] = 64;                         // two ctrl blocks
  sz[SL_BEST] = 128;                         // best, axiom, pick, trace state, recompute scratch
  sz[SL_BT] = cap > 0 ? (seg ? std::min((size_t)seg, nbands) : nbands) * nwords * 64 * 4 : 0;
  sz[SL_CP] = seg ? (nbands / seg + 1) * W * 8 : 0;
  sz[SL_TR] = sizeof(dpx_trace_t) * (size_t)(cap + 1);
  size_t at = 0;
  for (int k = 0; k < SL_N; ++k) {
    if (off) off[k] = at;
    at += a256(sz[k]);
  }
  return at;
}

int sw_trace_cap(int n1, int n2) { return 2 * (n1 + n2) + 8; }

// ---------------------------------------------------------------- jobs

// Split middle products pay off while a launch has too few ready tiles to
// fill the chip: launches of at most this many instances split them, when
// the partial results fit kMpSplitBytes.
constexpr int kMpSplitMax = 8;
constexpr int64_t kMpSplitBytes = (int64_t)4 << 30;

// per-tile partial slots of the split middle products; false if not split
bool mp_plan(Job& j) {
  int64_t extra = 0, slots = 0;
  for (const TriInst& t : j.inst)
    for (int d = 0; d < t.nb; ++d) {
      const int P = mp_parts(d);
      if (P > 1) {
        extra += (int64_t)(t.nb - d) * (P - 1);
        slots += (int64_t)(t.nb - d) * P;
      }
    }
  if (!slots || slots * T * T * 4 > kMpSplitBytes || j.tiles + extra >= ((int64_t)1 << 31)) return false;
  j.mp_base.assign(j.tiles, -1);
  int32_t at = 0;
  for (const TriInst& t : j.inst)
    for (int I = 0; I < t.nb; ++I)
      for (int J = I; J < t.nb; ++J) {
        const int P = mp_parts(J - I);
        if (P > 1) {
          j.mp_base[t.tile0 + tile_index(t.nb, I, J)] = at;
          at += P;
        }
      }
  j.mp_split = true;
  j.items = j.tiles + extra;
  j.mp_slots = slots;
  return true;
}

// Triangular job over `count` instances: lens[q] and the per-instance data
// (codes for RNA, rows|cols for the chain) at data_off[q] within `data`.
// split_ok: the launch may split middle products (mp_plan).
int tri_prepare(std::string& err, Job& j, int g, const std::vector<int>& lens, const std::vector<int64_t>& data_off,
                std::vector<uint8_t>&& data, int min_loop, bool want_trace, bool split_ok = true) {
  const int count = (int)lens.size();
  j.kind = JOB_TRI;
  j.g = g;
  j.count = count;
  j.min_loop = min_loop;
  j.want_trace = want_trace;
  j.data = std::move(data);
  j.inst.assign(count, TriInst{});
  j.tiles = j.trace_total = 0;
  j.n_max = 0;
  j.cells = 0;
  for (int q = 0; q < count; ++q) {
    TriInst& t = j.inst[q];
    if (lens[q] < 0 || lens[q] > max_len(g)) return fail(err, DPX_EINVAL, "instance size out of range");
    t.n = lens[q];
    t.nb = (t.n + 1 + T - 1) / T;
    t.seq_off = data_off[q];
    t.tile0 = j.tiles;
    j.tiles += tiles_of(t.nb);
    t.trace_off = j.trace_total;
    t.trace_cap = want_trace ? tri_trace_cap(t.n) : 0;
    j.trace_total += t.trace_cap;
    t.out = q;
    j.n_max = std::max(j.n_max, t.n);
    j.cells += tri_cells(t.n);
  }
  if (j.tiles >= ((int64_t)1 << 31)) return fail(err, DPX_ENOMEM, "too many tiles in one launch");
  // longest instances first: their chain of tile diagonals is the longest,
  // so they are seeded first, and the finalize launch (one workgroup per
  // instance, in array order) starts them first too
  j.trace_off_by_out.resize(count);
  for (int q = 0; q < count; ++q) j.trace_off_by_out[q] = j.inst[q].trace_off;
  std::stable_sort(j.inst.begin(), j.inst.end(), [](const TriInst& x, const TriInst& y) { return x.n > y.n; });
  j.seeds.clear();
  j.seeds.reserve(j.tiles);
  for (int q = 0; q < count; ++q)
    for (int I = 0; I < j.inst[q].nb; ++I) j.seeds.push_back(pack_item(q, I, I));
  j.mp_split = false;
  j.items = j.tiles;
  j.mp_slots = 0;
  j.mp_base.clear();
  if (split_ok && count <= kMpSplitMax) mp_plan(j);
  j.dev_bytes = tri_layout(g, j.tiles, count, j.data.size(), j.trace_total, j.off, j.items, j.mp_slots);
  return DPX_OK;
}

// kernel arguments of a triangular job whose device block starts at dev
TriArgs tri_args(const dpx_ctx* c, const Job& j, char* dev) {
  TriArgs a{};
  a.inst = (const TriInst*)(dev + j.off[TL_INST]);
  a.n_inst = j.count;
  a.total_tiles = (int32_t)j.tiles;
  a.total_items = (int32_t)(j.mp_split ? j.items : j.tiles);
  a.mp_split = j.mp_split ? 1 : 0;
  if (j.mp_split) a.mp = (int32_t*)(dev + j.off[TL_MP]);
  a.queue = (unsigned long long*)(dev + j.off[TL_QUEUE]);
  a.ctrl = (int32_t*)(dev + j.off[TL_CTRL]);
  a.dep = (int32_t*)(dev + j.off[TL_DEP]);
  a.tab = (int32_t*)(dev + j.off[TL_TAB]);
  a.codes = (const uint8_t*)(dev + j.off[TL_DATA]);
  a.dims = (const int32_t*)(dev + j.off[TL_DATA]);
  a.n_codes = (int64_t)j.data.size();
  a.spec = (j.g == G_ZUKER || j.g == G_RNAFOLD) ? (int16_t*)(dev + j.off[TL_SPEC]) : nullptr;
  a.rna = c->d_rna.p;
  a.hl_ent = (const int32_t*)c->d_hl.p;
  a.min_loop = j.min_loop;
  a.results = (int32_t*)(dev + j.off[TL_RES]);
  a.traces = dev + j.off[TL_TRACES];
  a.trace_lens = (int32_t*)(dev + j.off[TL_LENS]);
  a.want_trace = j.want_trace ? 1 : 0;
  return a;
}

int tri_launch(dpx_ctx* c, std::string& err, Workspace& w, char* dev, Job& j) {
  const int g = j.g, count = j.count;
  const size_t in_bytes = sizeof(TriInst) * count + j.seeds.size() * 8 + j.data.size() + 64 + j.mp_base.size() * 4;
  const size_t out_bytes = 8 * (size_t)count + 64 + (j.want_trace ? sizeof(dpx_trace_t) * (size_t)j.trace_total : 0);
  if (!w.h_in.ensure(in_bytes) || !w.h_out.ensure(out_bytes)) return fail(err, DPX_ENOMEM, "pinned host staging");
  // stage the inputs in pinned memory, then copy them asynchronously
  char* p_inst = (char*)w.h_in.p;
  std::memcpy(p_inst, j.inst.data(), sizeof(TriInst) * count);
  char* p_seeds = p_inst + sizeof(TriInst) * count;
  std::memcpy(p_seeds, j.seeds.data(), j.seeds.size() * 8);
  char* p_data = p_seeds + j.seeds.size() * 8;
  if (!j.data.empty()) std::memcpy(p_data, j.data.data(), j.data.size());
  char* p_base = p_data + ((j.data.size() + 63) & ~(size_t)63);
  if (j.mp_split) std::memcpy(p_base, j.mp_base.data(), j.mp_base.size() * 4);
  hipStream_t s = w.stream;
  HIPCHK(hipMemcpyAsync(dev + j.off[TL_INST], p_inst, sizeof(TriInst) * count, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dev + j.off[TL_SEEDS], p_seeds, j.seeds.size() * 8, hipMemcpyHostToDevice, s));
  if (!j.data.empty()) HIPCHK(hipMemcpyAsync(dev + j.off[TL_DATA], p_data, j.data.size(), hipMemcpyHostToDevice, s));
  if (j.mp_split)
    HIPCHK(hipMemcpyAsync(dev + j.off[TL_MP] + j.tiles * 4, p_base, j.mp_base.size() * 4, hipMemcpyHostToDevice, s));
  const TriArgs a = tri_args(c, j, dev);
  const int grid = tri_grid(c, j);
  HIPCHK(hipEventRecord(w.ev[0], s));
  if (launch_fill(g, a, (const unsigned long long*)(dev + j.off[TL_SEEDS]), (int)j.seeds.size(), grid, s))
    return fail(err, DPX_EHIP, "fill launch failed");
  HIPCHK(hipEventRecord(w.ev[1], s));
  if (launch_finalize(g, a, count, j.n_max, s)) return fail(err, DPX_EHIP, "finalize launch failed");
  HIPCHK(hipEventRecord(w.ev[2], s));
  char* hout = (char*)w.h_out.p;
  int32_t* o_res = (int32_t*)hout;
  int32_t* o_lens = o_res + count;
  int32_t* o_flags = o_lens + count;
  dpx_trace_t* o_tr = (dpx_trace_t*)(((uintptr_t)(o_flags + 4) + 15) & ~(uintptr_t)15);
  HIPCHK(hipMemcpyAsync(o_res, a.results, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(o_flags, a.ctrl, sizeof(int32_t) * 4, hipMemcpyDeviceToHost, s));
  if (j.want_trace) {
    HIPCHK(hipMemcpyAsync(o_lens, a.trace_lens, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(o_tr, a.traces, sizeof(dpx_trace_t) * j.trace_total, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipEventRecord(w.ev[3], s));
  j.res = o_res;
  j.lens = o_lens;
  j.flags = o_flags;
  j.tr = o_tr;
  return DPX_OK;
}

// Checks an alignment's size and scores and computes its row 0 (i == 0):
// g1 = 0 [4]; g2 [7/8/9]; M = g1 [2], g2 when strictly greater [3].
int sw_row0(std::string& err, int n1, int n2, const dpx_swat_scores& sc, std::vector<int2>& row0,
            std::vector<uint8_t>& r0m, std::vector<uint8_t>& r0g2, int* best, int* bestj, int* axiom) {
  // the best cell travels as one 64-bit key: a 36-bit cell index and a
  // 28-bit score (bounded by the largest gain per step times the steps)
  if ((int64_t)(n1 + 1) * (n2 + 1) > (int64_t)SW_KEY_MASK) return fail(err, DPX_EINVAL, "matrix too large (over 2^36 cells)");
  const int64_t gain = std::max({0, sc.match, sc.gap_open, sc.gap_ext});
  if (gain * ((int64_t)n1 + n2 + 1) >= ((int64_t)1 << (64 - SW_KEY_BITS - 1)))
    return fail(err, DPX_EINVAL, "scores too large for the best-cell key");
  const int W = n2 + 1;
  row0.resize(W);
  r0m.resize(W);
  r0g2.resize(W);
  int Mprev = 0, G2prev = 0;
  *best = -1;
  *bestj = 0;
  for (int x = 0; x < W; ++x) {
    int g2, rg2;
    if (x == 0) {
      g2 = 0;
      rg2 = 7;
    } else {
      const int o = std::max(0, Mprev + sc.gap_open), e = std::max(0, G2prev + sc.gap_ext);
      rg2 = e > o ? 9 : 8;
      g2 = e > o ? e : o;
    }
    int M = 0, rm = x == 0 ? 0 : 2;
    if (g2 > M) { M = g2; rm = 3; }
    row0[x] = make_int2(M, 0);
    r0m[x] = (uint8_t)rm;
    r0g2[x] = (uint8_t)rg2;
    if (M > *best) { *best = M; *bestj = x; }
    Mprev = M;
    G2prev = g2;
  }
  *axiom = row0[n2].x;
  return DPX_OK;
}

// seg > 0: keep checkpoint rows every seg bands instead of every band's
// traceback codes (only with a trace; the synchronous entry point's choice)
int sw_prepare(std::string& err, Job& j, const char* a, int n1, const char* b, int n2, const dpx_swat_scores& sc,
               int trace_from, int trace_cap, int seg = 0) {
  const int r = sw_row0(err, n1, n2, sc, j.row0, j.r0m, j.r0g2, &j.row0_best, &j.row0_bestj, &j.row0_axiom);
  if (r) return r;
  j.kind = JOB_SW;
  j.a.assign(a, n1);
  j.b.assign(b, n2);
  j.sc = sc;
  j.trace_from = trace_from;
  j.sw_cap = trace_cap;
  j.sw_seg = trace_cap > 0 ? seg : 0;
  j.cells = (int64_t)(n1 + 1) * (n2 + 1);
  j.dev_bytes = sw_layout(n1, n2, trace_cap, j.sw_seg, j.off);
  return DPX_OK;
}

// ---- batched alignments (JOB_SWB) ----
// Block: descriptors (SwPair per alignment) | band -> alignment table |
// claim counter | inputs (per alignment a, b, row 0 and its rule bytes) |
// outputs (per alignment 32 ints: [0..5] pick, [8..13] trace state,
// [16..19] ctrl, [20..21] best-cell key, [22] M(n1,n2)) | trace records |
// per alignment its boundary-row ring and traceback codes.  Descriptors,
// table and inputs are one upload, outputs one download.
constexpr size_t kSwbOut = 32 * sizeof(int32_t);
size_t a16(size_t x) { return (x + 15) & ~(size_t)15; }
size_t swb_in_bytes(int n1, int n2) { return a16(n1) + a16(n2) + a16(sizeof(int2) * (n2 + 1)) + 2 * a16(n2 + 1); }
size_t swb_work_bytes(int n1, int n2, int cap) {
  const int nbands = (n1 + 63) / 64;
  const size_t ring = (size_t)sw_slots(nbands) * (n2 + 2) * 8;
  const size_t codes = cap > 0 ? (size_t)nbands * sw_code_words(n2) * 64 * 4 : 0;
  return a256(ring) + a256(codes);
}
// upper bound of one alignment's share of a batch's block (the server's footprint)
size_t swb_pair_bytes(int n1, int n2, int cap) {
  return a256(sizeof(SwPair)) + a256(4 * (size_t)((n1 + 63) / 64)) + a256(swb_in_bytes(n1, n2)) + a256(kSwbOut) +
         a256(sizeof(dpx_trace_t) * (size_t)cap) + swb_work_bytes(n1, n2, cap) + 256;
}

// Lays out a batch whose alignments are in j.swb (sizes, caps set).
void swb_layout(Job& j) {
  const int np = (int)j.swb.size();
  j.swb_bands = 0;
  int64_t tr = 0;
  size_t in = 0;
  for (SwPairJob& q : j.swb) {
    const int n1 = (int)q.a.size(), n2 = (int)q.b.size();
    q.nbands = (n1 + 63) / 64;
    q.band0 = j.swb_bands;
    j.swb_bands += q.nbands;
    q.nslots = sw_slots(q.nbands);
    q.nwords = sw_code_words(n2);
    q.tr0 = tr;
    tr += q.cap;
    q.in_a = in;
    q.in_b = q.in_a + a16(n1);
    q.in_row0 = q.in_b + a16(n2);
    in += swb_in_bytes(n1, n2);
  }
  size_t at = 0;
  j.swb_desc = at;
  at += a256(sizeof(SwPair) * (size_t)np);
  j.swb_bandtab = at;
  at += a256(4 * (size_t)j.swb_bands);
  j.swb_in0 = 0;  // the upload starts at the descriptors
  j.swb_claim = at;
  at += 256;
  const size_t inputs = at;
  for (SwPairJob& q : j.swb) {
    q.in_a += inputs;
    q.in_b += inputs;
    q.in_row0 += inputs;
  }
  at += a256(in);
  j.swb_in1 = at;  // end of the upload
  j.swb_out = at;
  at += a256(kSwbOut * (size_t)np);
  j.swb_tr = at;
  at += a256(sizeof(dpx_trace_t) * (size_t)tr);
  for (SwPairJob& q : j.swb) {
    const int n2 = (int)q.b.size();
    q.ring = at;
    at += a256((size_t)q.nslots * (n2 + 2) * 8);
    q.codes = at;
    at += q.cap > 0 ? a256((size_t)q.nbands * q.nwords * 64 * 4) : 0;
  }
  j.dev_bytes = at;
}

// A batch of alignments (every one traced, or none).
int swb_add(std::string& err, Job& j, const std::string& a, const std::string& b, const dpx_swat_scores& sc,
            int trace_from, int cap) {
  SwPairJob q;
  const int r = sw_row0(err, (int)a.size(), (int)b.size(), sc, q.row0, q.r0m, q.r0g2, &q.row0_best, &q.row0_bestj,
                        &q.row0_axiom);
  if (r) return r;
  q.a = a;
  q.b = b;
  q.sc = sc;
  q.trace_from = trace_from;
  q.cap = cap;
  q.cells = (int64_t)(a.size() + 1) * (b.size() + 1);
  j.kind = JOB_SWB;
  j.cells += q.cells;
  j.swb.push_back(std::move(q));
  return DPX_OK;
}

int sw_grid(dpx_ctx* c, int nbands, int n) {
  // tuned: bands in flight per CU (0 = occupancy); never more than the ring allows
  const dpx_tuning t = get_tune(c, DPX_REQ_SWAT, tune_bucket(n));
  const int w = t.wg_per_cu >= 1 && t.wg_per_cu <= c->sw_occ ? t.wg_per_cu : c->sw_occ;
  return std::max(1, std::min(std::min(nbands, w * c->cu_count), sw_slots(nbands) - 1));
}

int sw_chunk(dpx_ctx* c, int n) {
  const dpx_tuning t = get_tune(c, DPX_REQ_SWAT, tune_bucket(n));
  return t.chunk == 16 || t.chunk == 64 ? t.chunk : 32;
}

int sw_launch_job(dpx_ctx* c, std::string& err, Workspace& w, char* dev, Job& j) {
  const int n1 = (int)j.a.size(), n2 = (int)j.b.size(), W = n2 + 1;
  SwArgs A{};
  A.n1 = n1;
  A.n2 = n2;
  A.match = j.sc.match;
  A.mismatch = j.sc.mismatch;
  A.open = j.sc.gap_open;
  A.ext = j.sc.gap_ext;
  A.nbands = (n1 + 63) / 64;
  A.nwords = sw_code_words(n2);
  A.nslots = sw_slots(A.nbands);
  const bool want_trace = j.sw_cap > 0;
  const size_t in_bytes = (size_t)n1 + n2 + sizeof(int2) * W + 2 * (size_t)W + 64;
  const size_t out_bytes = 64 + (want_trace ? sizeof(dpx_trace_t) * (size_t)j.sw_cap : 0);
  if (!w.h_in.ensure(in_bytes) || !w.h_out.ensure(out_bytes)) return fail(err, DPX_ENOMEM, "pinned host staging");
  hipStream_t s = w.stream;
  char* rows = dev + j.off[SL_ROWS];
  A.hrow = (unsigned long long*)rows;
  int2* d_row0 = (int2*)(rows + (size_t)A.nslots * (W + 1) * 8);
  uint8_t* d_r0m = (uint8_t*)(d_row0 + W);
  uint8_t* d_r0g2 = d_r0m + W;
  A.a = (const uint8_t*)(dev + j.off[SL_A]);
  A.b = (const uint8_t*)(dev + j.off[SL_B]);
  A.ctrl = (int*)(dev + j.off[SL_FLAGS]);  // [0] claims, [1] time-out; [4], [5] for recomputed segments
  uint32_t* d_codes = want_trace ? (uint32_t*)(dev + j.off[SL_BT]) : nullptr;
  A.bt = j.sw_seg ? nullptr : d_codes;
  A.cp = j.sw_seg ? (int2*)(dev + j.off[SL_CP]) : nullptr;
  A.cp_every = j.sw_seg ? j.sw_seg : 1;
  char* bestblk = dev + j.off[SL_BEST];
  A.best = (unsigned long long*)bestblk;
  A.out = (int*)(bestblk + 8);
  int* d_pick = (int*)(bestblk + 16);   // 6 ints, see sw_pick
  int* d_state = (int*)(bestblk + 48);  // SW_STATE ints
  char* h = (char*)w.h_in.p;
  std::memcpy(h, j.a.data(), n1);
  std::memcpy(h + n1, j.b.data(), n2);
  char* h_row0 = h + ((n1 + n2 + 15) & ~15);
  std::memcpy(h_row0, j.row0.data(), sizeof(int2) * W);
  std::memcpy(h_row0 + sizeof(int2) * W, j.r0m.data(), W);
  std::memcpy(h_row0 + sizeof(int2) * W + W, j.r0g2.data(), W);
  if (n1) HIPCHK(hipMemcpyAsync((void*)A.a, h, n1, hipMemcpyHostToDevice, s));
  if (n2) HIPCHK(hipMemcpyAsync((void*)A.b, h + n1, n2, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d_row0, h_row0, sizeof(int2) * W + 2 * (size_t)W, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(w.ev[0], s));
  const int n = std::max(n1, n2), chunk = sw_chunk(c, n);
  if (A.nbands > 0 && sw_launch(A, d_row0, sw_grid(c, A.nbands, n), chunk, s))
    return fail(err, DPX_EHIP, "alignment launch failed");
  HIPCHK(hipEventRecord(w.ev[1], s));
  void* d_tr = want_trace ? dev + j.off[SL_TR] : nullptr;
  if (sw_finish_launch(A, j.row0_best, j.row0_bestj, j.row0_axiom, j.trace_from, d_pick, d_r0m, d_r0g2, d_tr, j.sw_cap,
                       d_state, s))
    return fail(err, DPX_EHIP, "traceback launch failed");
  if (want_trace && j.sw_seg) {
    // Checkpointed traceback: segment by segment from the end of the walk,
    // recompute the segment's codes (its rows, the columns up to the walk's)
    // from the checkpoint row above it, then walk through them.
    const int seg_rows = 64 * j.sw_seg;
    int st[SW_STATE], late = 0;
    for (;;) {
      HIPCHK(hipMemcpyAsync(st, d_state, sizeof st, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (late) return fail(err, DPX_EHIP, "band wait timed out in a recomputed segment (scheduler error)");
      if (st[4] || !st[5]) break;
      const int i = st[0], jc = st[1];
      if (i == 0) {  // row 0: its rules are known, no codes needed
        if (sw_trace_launch(A, d_r0m, d_r0g2, 0, d_tr, j.sw_cap, d_state, s))
          return fail(err, DPX_EHIP, "traceback launch failed");
        continue;
      }
      const int sg = (i - 1) / seg_rows, base = sg * seg_rows;
      SwArgs B = A;
      B.a = A.a + base;
      B.n1 = std::min(n1, base + seg_rows) - base;
      B.n2 = jc;
      B.nbands = (B.n1 + 63) / 64;
      B.nwords = sw_code_words(jc);
      B.nslots = sw_slots(B.nbands);
      B.ctrl = A.ctrl + 4;
      B.bt = d_codes;
      B.cp = nullptr;
      B.best = (unsigned long long*)(bestblk + 80);
      B.out = (int*)(bestblk + 88);
      const int2* top = sg == 0 ? d_row0 : A.cp + (size_t)sg * W;
      if (sw_launch(B, top, sw_grid(c, B.nbands, n), chunk, s) ||
          sw_trace_launch(B, d_r0m, d_r0g2, base, d_tr, j.sw_cap, d_state, s))
        return fail(err, DPX_EHIP, "traceback recompute failed");
      HIPCHK(hipMemcpyAsync(&late, B.ctrl + 1, sizeof late, hipMemcpyDeviceToHost, s));
    }
  }
  HIPCHK(hipEventRecord(w.ev[2], s));
  char* ho = (char*)w.h_out.p;
  // [0..5] sw_pick output, [6] trace length (set at harvest), [7] band-wait
  // time-out flag, [8..13] trace state, trace records at +64
  HIPCHK(hipMemcpyAsync(ho, d_pick, 6 * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(ho + 32, d_state, SW_STATE * sizeof(int), hipMemcpyDeviceToHost, s));
  if (A.nbands > 0) HIPCHK(hipMemcpyAsync(ho + 28, A.ctrl + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  else std::memset(ho + 28, 0, 4);
  if (want_trace) HIPCHK(hipMemcpyAsync(ho + 64, d_tr, sizeof(dpx_trace_t) * j.sw_cap, hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(w.ev[3], s));
  j.res = (const int32_t*)ho;
  j.tr = (const dpx_trace_t*)(ho + 64);
  return DPX_OK;
}

// A batch of alignments: one upload, reset + one persistent fill over the
// bands of all of them, best cell and traceback per alignment, one download.
int swb_launch_job(dpx_ctx* c, std::string& err, Workspace& w, char* dev, Job& j) {
  const int np = (int)j.swb.size();
  int64_t trtot = 0;
  for (const SwPairJob& q : j.swb) trtot += q.cap;
  const size_t up = j.swb_in1, outb = a256(kSwbOut * (size_t)np);
  const size_t trb = j.swb_trace ? sizeof(dpx_trace_t) * (size_t)trtot : 0;
  if (!w.h_in.ensure(up) || !w.h_out.ensure(outb + trb + 64)) return fail(err, DPX_ENOMEM, "pinned host staging");
  char* h = (char*)w.h_in.p;
  SwPair* hp = (SwPair*)(h + j.swb_desc);
  int32_t* hbt = (int32_t*)(h + j.swb_bandtab);
  int n_max = 0;
  size_t max_words = 1;
  for (int p = 0; p < np; ++p) {
    const SwPairJob& q = j.swb[p];
    const int n1 = (int)q.a.size(), n2 = (int)q.b.size(), W = n2 + 1;
    int* outp = (int*)(dev + j.swb_out + kSwbOut * (size_t)p);
    SwPair d{};
    SwArgs& A = d.A;
    A.n1 = n1;
    A.n2 = n2;
    A.match = q.sc.match;
    A.mismatch = q.sc.mismatch;
    A.open = q.sc.gap_open;
    A.ext = q.sc.gap_ext;
    A.nbands = q.nbands;
    A.nwords = q.nwords;
    A.nslots = q.nslots;
    A.hrow = (unsigned long long*)(dev + q.ring);
    A.a = (const uint8_t*)(dev + q.in_a);
    A.b = (const uint8_t*)(dev + q.in_b);
    A.ctrl = outp + 16;  // [16] (unused claim counter), [17] band-wait time-out
    A.bt = q.cap > 0 ? (uint32_t*)(dev + q.codes) : nullptr;
    A.cp = nullptr;
    A.cp_every = 1;
    A.best = (unsigned long long*)(outp + 20);
    A.out = outp + 22;
    d.band0 = q.band0;
    d.row0_best = q.row0_best;
    d.row0_bestj = q.row0_bestj;
    d.row0_axiom = q.row0_axiom;
    d.trace_from = q.trace_from;
    d.cap = q.cap;
    d.row0 = (const int2*)(dev + q.in_row0);
    d.r0m = (const uint8_t*)(dev + q.in_row0 + a16(sizeof(int2) * W));
    d.r0g2 = d.r0m + a16(W);
    d.pick = outp;
    d.state = outp + 8;
    d.tr = q.cap > 0 ? dev + j.swb_tr + sizeof(dpx_trace_t) * (size_t)q.tr0 : nullptr;
    hp[p] = d;
    for (int b = 0; b < q.nbands; ++b) hbt[q.band0 + b] = p;
    std::memcpy(h + q.in_a, q.a.data(), n1);
    std::memcpy(h + q.in_b, q.b.data(), n2);
    std::memcpy(h + q.in_row0, q.row0.data(), sizeof(int2) * W);
    std::memcpy(h + q.in_row0 + a16(sizeof(int2) * W), q.r0m.data(), W);
    std::memcpy(h + q.in_row0 + a16(sizeof(int2) * W) + a16(W), q.r0g2.data(), W);
    n_max = std::max(n_max, std::max(n1, n2));
    max_words = std::max(max_words, (size_t)q.nslots * (n2 + 2));
  }
  hipStream_t s = w.stream;
  HIPCHK(hipMemcpyAsync(dev, h, up, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(w.ev[0], s));
  const SwPair* d_pairs = (const SwPair*)(dev + j.swb_desc);
  if (sw_batch_launch(d_pairs, np, (const int32_t*)(dev + j.swb_bandtab), j.swb_bands, (int*)(dev + j.swb_claim),
                      max_words, j.swb_trace, j.swb_trace, sw_grid(c, std::max(1, j.swb_bands), n_max),
                      sw_chunk(c, n_max), s))
    return fail(err, DPX_EHIP, "batched alignment launch failed");
  HIPCHK(hipEventRecord(w.ev[1], s));
  if (sw_batch_finish(d_pairs, np, j.swb_trace, s)) return fail(err, DPX_EHIP, "batched traceback launch failed");
  HIPCHK(hipEventRecord(w.ev[2], s));
  char* ho = (char*)w.h_out.p;
  HIPCHK(hipMemcpyAsync(ho, dev + j.swb_out, kSwbOut * (size_t)np, hipMemcpyDeviceToHost, s));
  if (trb) HIPCHK(hipMemcpyAsync(ho + outb, dev + j.swb_tr, trb, hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(w.ev[3], s));
  j.res = (const int32_t*)ho;
  j.tr = (const dpx_trace_t*)(ho + outb);
  return DPX_OK;
}

int job_launch(dpx_ctx* c, std::string& err, Workspace& w, char* dev, Job& j) {
  return j.kind == JOB_TRI  ? tri_launch(c, err, w, dev, j)
         : j.kind == JOB_SW ? sw_launch_job(c, err, w, dev, j)
                            : swb_launch_job(c, err, w, dev, j);
}

// after ev[3] has completed: timings and the kernels' time-out flags
int job_harvest(std::string& err, Workspace& w, Job& j) {
  (void)hipEventElapsedTime(&j.t_fill, w.ev[0], w.ev[1]);
  (void)hipEventElapsedTime(&j.t_fin, w.ev[1], w.ev[2]);
  if (j.kind == JOB_TRI && j.flags[2]) return fail(err, DPX_EHIP, "tile queue wait timed out (scheduler error)");
  if (j.kind == JOB_SW) {
    int32_t* r = (int32_t*)j.res;
    if (r[7]) return fail(err, DPX_EHIP, "band wait timed out (scheduler error)");
    r[6] = j.sw_cap > 0 ? (r[8 + 5] ? r[8 + 3] : -1) : 0;  // trace length from the trace state
  }
  if (j.kind == JOB_SWB) {
    for (size_t p = 0; p < j.swb.size(); ++p) {
      int32_t* r = (int32_t*)j.res + 32 * p;
      if (r[17]) return fail(err, DPX_EHIP, "band wait timed out (scheduler error)");
      r[6] = j.swb[p].cap > 0 ? (r[8 + 5] ? r[8 + 3] : -1) : 0;
    }
  }
  return DPX_OK;
}

// synchronous run on the context's workspace and device block
int run_sync(dpx_ctx* c, Job& j) {
  std::string& err = c->err;
  if (!c->block.ensure(j.dev_bytes))
    return fail(err, DPX_ENOMEM, "device allocation for the DP tables failed (" + std::to_string(j.dev_bytes >> 20) + " MiB)");
  int r = job_launch(c, err, c->ws, (char*)c->block.p, j);
  if (r) return r;
  HIPCHK(hipEventSynchronize(c->ws.ev[3]));
  r = job_harvest(err, c->ws, j);
  c->t_fill = j.t_fill;
  c->t_fin = j.t_fin;
  return r;
}

// results and traces of a finished triangular job into the caller's buffers
int tri_copy_out(std::string& err, const Job& j, int32_t* results, dpx_trace_t* traces, const int64_t* trace_offs,
                 int32_t* trace_lens) {
  int rc = DPX_OK;
  for (int q = 0; q < j.count; ++q) results[q] = j.res[q];
  if (!j.want_trace || !traces) return rc;
  for (int q = 0; q < j.count; ++q) {
    const int64_t cap = trace_offs[q + 1] - trace_offs[q];
    const int tl = j.lens[q];
    trace_lens[q] = tl;
    if (tl < 0) { rc = fail(err, DPX_ETRACE, "backtrace failed"); continue; }
    if (tl > cap) { rc = fail(err, DPX_ETRACE, "trace buffer too small"); continue; }
    std::memcpy(traces + trace_offs[q], j.tr + j.trace_off_by_out[q], sizeof(dpx_trace_t) * tl);
  }
  return rc;
}

int rna_job(std::string& err, Job& j, int g, const char* seqs, const int64_t* offs, int count, int min_loop, bool trace) {
  std::vector<int> lens(count);
  std::vector<int64_t> off(count);
  const int64_t total = count ? offs[count] - offs[0] : 0;
  std::vector<uint8_t> codes((size_t)std::max<int64_t>(total, 0) + count + 64);
  int64_t pos = 0;
  for (int q = 0; q < count; ++q) {
    const int64_t n = offs[q + 1] - offs[q];
    if (n < 0 || n > max_len(g)) return fail(err, DPX_EINVAL, "sequence length out of range");
    lens[q] = (int)n;
    off[q] = pos;
    int r = encode_rna(err, seqs + offs[q], (int)n, codes.data() + pos);
    if (r) return r;
    pos += n + 1;
  }
  codes.resize(pos);
  return tri_prepare(err, j, g, lens, off, std::move(codes), min_loop, trace);
}

int rna_batch(dpx_ctx* c, int g, const char* seqs, const int64_t* offs, int count, int min_loop, int32_t* results,
              dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens) {
  std::string& err = c->err;
  if (count < 0 || (count > 0 && (!seqs || !offs || !results))) return fail(err, DPX_EINVAL, "null argument");
  if (traces && (!trace_offs || !trace_lens)) return fail(err, DPX_EINVAL, "trace offsets missing");
  if (g != G_NUSSINOV && !c->has_rna) return fail(err, DPX_ENOPARAMS, "dpx_set_rna_params was not called");
  if (count == 0) return DPX_OK;
  Job j;
  int r = rna_job(err, j, g, seqs, offs, count, min_loop, traces != nullptr);
  if (!r) r = run_sync(c, j);
  if (r) return r;
  return tri_copy_out(err, j, results, traces, trace_offs, trace_lens);
}

std::vector<uint8_t> chain_data(const int32_t* rows, const int32_t* cols, int n) {
  std::vector<uint8_t> d(2 * (size_t)n * sizeof(int32_t));
  std::memcpy(d.data(), rows, sizeof(int32_t) * n);
  std::memcpy(d.data() + sizeof(int32_t) * n, cols, sizeof(int32_t) * n);
  return d;
}

}  // namespace

// ================================================================== server

namespace {

struct Request {
  uint64_t id = 0;
  int kind = 0;
  int trace = 0, trace_from = 0, min_loop = 0;
  std::string s1, s2;
  std::vector<int32_t> rows, cols;
  dpx_swat_scores sc{};
  double t_submit = 0;
  size_t footprint = 0;
  int64_t cells = 0;
  int size_class = 0;
  int passed = 0;  // launches started ahead of it while it waited for memory
};

// alignments batched into one launch by the server at most
constexpr int kSwMaxBatch = 64;

// A request that does not fit the arena's free memory may be passed over by
// later requests that do, at most this many times; then it blocks the queue
// until memory frees up, so a large request cannot starve.
constexpr int kMaxPassOver = 16;

struct Finished {
  dpx_result r{};
  std::vector<dpx_trace_t> trace;
};

struct Slot {
  Workspace ws;
  Job job;
  std::vector<Request> reqs;
  bool busy = false;
  size_t block_off = 0;
  double t_launch = 0;
  std::string err;
};

// First-fit allocator of 256-byte aligned ranges of the arena.
struct Arena {
  DevBuf mem;
  std::map<size_t, size_t> free_;  // offset -> size
  size_t largest() const {
    size_t m = 0;
    for (auto& f : free_) m = std::max(m, f.second);
    return m;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = a256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = a256(bytes);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

}  // namespace

struct dpx_server {
  dpx_ctx* ctx = nullptr;
  dpx_server_config cfg{};
  Arena arena;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> pending;
  std::deque<uint64_t> ready;
  std::unordered_map<uint64_t, Finished> done;
  std::unordered_map<uint64_t, std::vector<dpx_trace_t>> traces;
  std::thread th;
  std::atomic<bool> stop{false};
  uint64_t next_id = 1;
  dpx_server_stats st{};
};

namespace {

int size_class(int n) {
  int c = 0;
  while ((32 << c) < n) ++c;
  return c;
}

void publish(dpx_server* S, Finished&& f) {
  std::lock_guard<std::mutex> lk(S->mu);
  if (f.r.status == DPX_OK) {
    S->st.completed++;
    S->st.cells_done += f.r.cells;
  } else {
    S->st.failed++;
  }
  const uint64_t id = f.r.id;
  S->done[id] = std::move(f);
  S->ready.push_back(id);
  S->cv_done.notify_all();
}

Finished base_result(const Request& q, const Slot& sl, int status) {
  Finished f;
  f.r.id = q.id;
  f.r.kind = q.kind;
  f.r.status = status;
  f.r.t_submit = q.t_submit;
  f.r.t_launch = sl.t_launch;
  f.r.batch_size = (int32_t)sl.reqs.size();
  f.r.cells = q.cells;
  f.r.footprint = (int64_t)q.footprint;
  return f;
}

// build the job of the requests in `sl` (server thread)
int prepare_slot(dpx_server* S, Slot& sl) {
  Job& j = sl.job;
  j = Job();
  const Request& r0 = sl.reqs.front();
  if (r0.kind == DPX_REQ_SWAT) {  // one batched launch (a batch shares the trace option)
    for (const Request& r : sl.reqs) {
      const int rc = swb_add(sl.err, j, r.s1, r.s2, r.sc, r.trace_from,
                             r.trace ? sw_trace_cap((int)r.s1.size(), (int)r.s2.size()) : 0);
      if (rc) return rc;
    }
    j.swb_trace = r0.trace != 0;
    swb_layout(j);
    return DPX_OK;
  }
  const int count = (int)sl.reqs.size();
  std::vector<int> lens(count);
  std::vector<int64_t> off(count);
  std::vector<uint8_t> data;
  if (r0.kind == DPX_REQ_MATCHAIN) {
    for (int q = 0; q < count; ++q) {
      const Request& r = sl.reqs[q];
      lens[q] = (int)r.rows.size();
      off[q] = (int64_t)(data.size() / 4);  // matrix chain offsets count int32 dims
      std::vector<uint8_t> d = chain_data(r.rows.data(), r.cols.data(), lens[q]);
      data.insert(data.end(), d.begin(), d.end());
    }
    return tri_prepare(sl.err, j, G_MATCHAIN, lens, off, std::move(data), 0, r0.trace != 0, false);
  }
  for (int q = 0; q < count; ++q) {
    const Request& r = sl.reqs[q];
    lens[q] = (int)r.s1.size();
    off[q] = (int64_t)data.size();
    const size_t at = data.size();
    data.resize(at + r.s1.size() + 1);
    int rc = encode_rna(sl.err, r.s1.data(), lens[q], data.data() + at);
    if (rc) return rc;
  }
  const int g = grammar_of(r0.kind);
  // footprints admitted by the server are those of unsplit launches
  return tri_prepare(sl.err, j, g, lens, off, std::move(data), r0.min_loop, r0.trace != 0, false);
}

void harvest_slot(dpx_server* S, Slot& sl, int status) {
  Job& j = sl.job;
  const int rc = status ? status : job_harvest(sl.err, sl.ws, j);
  const int batch = (int)sl.reqs.size();
  for (int q = 0; q < batch; ++q) {
    const Request& rq = sl.reqs[q];
    Finished f = base_result(rq, sl, rc);
    f.r.device_ms = j.t_fill + j.t_fin;
    if (!rc) {
      int tl = 0;
      const dpx_trace_t* src = nullptr;
      if (j.kind == JOB_SW) {
        f.r.score = j.res[0];
        f.r.best = j.res[1];
        f.r.best_i = j.res[2];
        f.r.best_j = j.res[3];
        tl = j.sw_cap > 0 ? j.res[6] : 0;
        src = j.tr;
      } else if (j.kind == JOB_SWB) {
        const int32_t* r = j.res + 32 * q;
        f.r.score = r[0];
        f.r.best = r[1];
        f.r.best_i = r[2];
        f.r.best_j = r[3];
        tl = j.swb[q].cap > 0 ? r[6] : 0;
        src = j.tr + j.swb[q].tr0;
      } else {
        f.r.score = j.res[q];
        tl = j.want_trace ? j.lens[q] : 0;
        src = j.tr + j.trace_off_by_out[q];
      }
      f.r.trace_len = tl;
      if (tl < 0) f.r.status = DPX_ETRACE;
      else if (tl > 0) f.trace.assign(src, src + tl);
    }
    f.r.t_done = now_s();
    publish(S, std::move(f));
  }
  S->arena.release(sl.block_off, j.dev_bytes);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->st.bytes_in_flight -= (int64_t)a256(j.dev_bytes);
  }
  sl.busy = false;
  sl.reqs.clear();
}

// Take the next batch for a free slot: the oldest waiting request that fits
// the largest free range of the arena decides kind, options and size class
// (requests ahead of it that do not fit are passed over, each at most
// kMaxPassOver times); later requests of the same kind, options and a
// neighbouring size class join in arrival order while the batch fits.
// Returns false when nothing can start.
bool take_batch(dpx_server* S, Slot& sl) {
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->pending.empty()) return false;
  const size_t room = S->arena.largest();
  auto hit = S->pending.begin();
  for (; hit != S->pending.end() && hit->footprint > room; ++hit)
    if (hit->passed >= kMaxPassOver) return false;  // it waits for memory; nothing passes it any more
  if (hit == S->pending.end()) return false;  // wait until launches in flight release memory
  for (auto it = S->pending.begin(); it != hit; ++it) ++it->passed;
  const Request& h = *hit;
  const int cap = h.kind == DPX_REQ_SWAT ? std::min(S->cfg.max_batch, kSwMaxBatch) : S->cfg.max_batch;
  size_t used = 0;
  const int kind = h.kind, trace = h.trace, ml = h.min_loop, cls = h.size_class;
  for (auto it = hit; it != S->pending.end() && (int)sl.reqs.size() < cap;) {
    const bool same = it->kind == kind && it->trace == trace && it->min_loop == ml && std::abs(it->size_class - cls) <= 1;
    if (same && used + it->footprint <= room) {
      used += it->footprint;
      sl.reqs.push_back(std::move(*it));
      it = S->pending.erase(it);
    } else {
      ++it;
    }
  }
  S->st.pending = (int64_t)S->pending.size();
  return true;
}

void server_loop(dpx_server* S) {
  (void)hipSetDevice(S->ctx->device);
  while (!S->stop.load()) {
    bool progress = false;
    for (Slot& sl : S->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.ws.ev[3]);
      if (q == hipErrorNotReady) continue;
      harvest_slot(S, sl, q == hipSuccess ? DPX_OK : hip_fail(sl.err, q, "launch in flight"));
      progress = true;
    }
    for (Slot& sl : S->slots) {
      if (sl.busy) continue;
      if (!take_batch(S, sl)) break;
      progress = true;
      sl.t_launch = now_s();
      int rc = prepare_slot(S, sl);
      if (!rc && !S->arena.alloc(sl.job.dev_bytes, &sl.block_off)) rc = fail(sl.err, DPX_ENOMEM, "arena exhausted");
      if (!rc) {
        rc = job_launch(S->ctx, sl.err, sl.ws, (char*)S->arena.mem.p + sl.block_off, sl.job);
        if (rc) {
          // work queued before the failure may still run: let the stream drain first
          (void)hipStreamSynchronize(sl.ws.stream);
          S->arena.release(sl.block_off, sl.job.dev_bytes);
        }
      }
      if (!rc) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->st.launches++;
        S->st.bytes_in_flight += (int64_t)a256(sl.job.dev_bytes);
        S->st.peak_bytes_in_flight = std::max(S->st.peak_bytes_in_flight, S->st.bytes_in_flight);
        S->st.max_batch_seen = std::max<int64_t>(S->st.max_batch_seen, (int64_t)sl.reqs.size());
        sl.busy = true;
      } else {
        for (const Request& q : sl.reqs) {
          Finished f = base_result(q, sl, rc);
          f.r.t_done = now_s();
          publish(S, std::move(f));
        }
        sl.reqs.clear();
      }
    }
    if (!progress) {
      bool any_busy = false;
      for (Slot& sl : S->slots) any_busy |= sl.busy;
      std::unique_lock<std::mutex> lk(S->mu);
      // with launches in flight, poll their events every 50 us; otherwise sleep until a submission
      S->cv_work.wait_for(lk, std::chrono::microseconds(any_busy ? 50 : 20000));
    }
  }
  for (Slot& sl : S->slots)
    if (sl.busy) {
      const hipError_t q = hipEventSynchronize(sl.ws.ev[3]);
      harvest_slot(S, sl, q == hipSuccess ? DPX_OK : hip_fail(sl.err, q, "launch in flight"));
    }
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
  if (hipSetDevice(device) != hipSuccess || !c->ws.init()) {
    c->ws.release();
    delete c;
    return DPX_EHIP;
  }
  (void)hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
  scratch_setup(device);
  for (int g = 0; g < N_GRAMMARS; ++g) c->occ[g] = std::max(1, fill_occupancy(g));
  c->sw_occ = std::max(1, sw_occupancy());
  *out = c;
  return DPX_OK;
}

void dpx_destroy(dpx_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  c->ws.release();
  c->block.release();
  c->d_rna.release();
  c->d_hl.release();
  delete c->rna;
  delete c;
}

const char* dpx_last_error(const dpx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dpx_set_rna_params(dpx_ctx* c, const char* path) {
  if (!c || !path) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (c->servers.load() > 0 || c->plans.load() > 0)
    return fail(err, DPX_EBUSY, "energy parameters cannot change while a server or a plan exists");
  (void)hipSetDevice(c->device);
  HostRna* h = new HostRna();
  char msg[256] = {0};
  if (load_par_file(path, h, msg, sizeof msg)) {
    delete h;
    return fail(err, DPX_EPARAMS, msg);
  }
  delete c->rna;
  c->rna = h;
  c->has_rna = false;
  std::vector<int32_t> hl(kHairpinTable);
  for (int l = 0; l < kHairpinTable; ++l) hl[l] = l < 3 ? 0 : h->hl_ent(l);
  if (!c->d_rna.ensure(sizeof(DevRna)) || !c->d_hl.ensure(sizeof(int32_t) * kHairpinTable))
    return fail(err, DPX_ENOMEM, "energy tables");
  HIPCHK(hipMemcpy(c->d_rna.p, &h->dev, sizeof(DevRna), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_hl.p, hl.data(), sizeof(int32_t) * kHairpinTable, hipMemcpyHostToDevice));
  c->has_rna = true;
  return DPX_OK;
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
  std::lock_guard<std::mutex> lk(c->mu);
  if (min_loop < 0) return fail(c->err, DPX_EINVAL, "min_loop < 0");
  (void)hipSetDevice(c->device);
  return rna_batch(c, G_NUSSINOV, seqs, offs, count, min_loop, results, traces, trace_offs, trace_lens);
}

static int single_rna(dpx_ctx* c, int g, const char* seq, int n, int min_loop, int32_t* score, dpx_trace_t* trace,
                      int trace_cap, int* trace_len) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!seq || n < 0 || !score || (trace && (!trace_len || trace_cap < 0)))
    return fail(c->err, DPX_EINVAL, "bad argument");
  if (min_loop < 0) return fail(c->err, DPX_EINVAL, "min_loop < 0");
  int64_t offs[2] = {0, n};
  int64_t toffs[2] = {0, trace_cap};
  int32_t tl = 0, res = 0;
  (void)hipSetDevice(c->device);
  int r = rna_batch(c, g, seq, offs, 1, min_loop, &res, trace, toffs, &tl);
  *score = res;
  if (trace) *trace_len = tl;
  return r;
}

int dpx_nussinov(dpx_ctx* c, const char* seq, int n, int min_loop, int32_t* score, dpx_trace_t* trace, int trace_cap,
                 int* trace_len) {
  return single_rna(c, G_NUSSINOV, seq, n, min_loop, score, trace, trace_cap, trace_len);
}

int dpx_zuker(dpx_ctx* c, const char* seq, int n, int32_t* mfe, dpx_trace_t* trace, int trace_cap, int* trace_len) {
  return single_rna(c, G_ZUKER, seq, n, 0, mfe, trace, trace_cap, trace_len);
}

int dpx_rnafold(dpx_ctx* c, const char* seq, int n, int32_t* mfe, dpx_trace_t* trace, int trace_cap, int* trace_len) {
  return single_rna(c, G_RNAFOLD, seq, n, 0, mfe, trace, trace_cap, trace_len);
}

// dpx_zuker on the device's static scratch: no device or pinned-host
// allocation (inputs staged from the stack, results copied straight into the
// caller's buffers).  Same kernels and launch shape as dpx_zuker.
int dpx_zuker_noalloc(dpx_ctx* c, const char* seq, int n, int32_t* mfe, dpx_trace_t* trace, int trace_cap,
                      int* trace_len) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (!seq || n < 0 || !mfe || (trace && (!trace_len || trace_cap < 0))) return fail(err, DPX_EINVAL, "bad argument");
  if (n > kNoAllocMaxZuker)
    return fail(err, DPX_EINVAL, "dpx_zuker_noalloc: n = " + std::to_string(n) + " exceeds its maximum of " +
                                     std::to_string(kNoAllocMaxZuker) + " (16 MiB of static scratch per device)");
  if (!c->has_rna) return fail(err, DPX_ENOPARAMS, "dpx_set_rna_params was not called");
  if (c->device >= kMaxDevices) return fail(err, DPX_EINVAL, "dpx_zuker_noalloc: device index too large");
  char* dev = nullptr;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    dev = (char*)g_scratch_ptr[c->device];
    if (!dev)
      return fail(err, DPX_EHIP, std::string("dpx_zuker_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                                     hipGetErrorString(g_scratch_rc[c->device]));
  }
  std::lock_guard<std::mutex> dev_lk(g_scratch_call_mu[c->device]);
  (void)hipSetDevice(c->device);
  // inputs on the stack: codes (followed by a 0, as in a batch), the
  // instance record and the diagonal tiles that seed the queue
  uint8_t codes[kNoAllocMaxZuker + 64];
  int r = encode_rna(err, seq, n, codes);
  if (r) return r;
  codes[n] = 0;
  const int nb = (n + 1 + T - 1) / T;
  unsigned long long seeds[(kNoAllocMaxZuker + 1 + T - 1) / T];
  for (int I = 0; I < nb; ++I) seeds[I] = pack_item(0, I, I);
  Job j;
  j.kind = JOB_TRI;
  j.g = G_ZUKER;
  j.count = 1;
  j.n_max = n;
  j.want_trace = trace != nullptr;
  j.tiles = tiles_of(nb);
  j.trace_total = trace ? tri_trace_cap(n) : 0;
  j.dev_bytes = tri_layout(G_ZUKER, j.tiles, 1, (size_t)n + 1, j.trace_total, j.off);
  if (j.dev_bytes > kScratchBytes) return fail(err, DPX_EINVAL, "dpx_zuker_noalloc: instance exceeds the scratch");
  TriInst inst{};
  inst.n = n;
  inst.nb = nb;
  inst.seq_off = 0;
  inst.tile0 = 0;
  inst.trace_off = 0;
  inst.trace_cap = (int32_t)j.trace_total;
  inst.out = 0;
  TriArgs a = tri_args(c, j, dev);
  a.n_codes = (int64_t)n + 1;
  hipStream_t s = c->ws.stream;
  HIPCHK(hipMemcpyAsync(dev + j.off[TL_INST], &inst, sizeof inst, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dev + j.off[TL_SEEDS], seeds, sizeof(unsigned long long) * nb, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dev + j.off[TL_DATA], codes, (size_t)n + 1, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(c->ws.ev[0], s));
  // the fill's reset kernel clears the queue, counters and flags the last
  // call left in the scratch; every table cell is written before it is read
  if (launch_fill(G_ZUKER, a, (const unsigned long long*)(dev + j.off[TL_SEEDS]), nb,
                  tri_grid_for(c, G_ZUKER, tune_bucket(n), j.tiles), s))
    return fail(err, DPX_EHIP, "fill launch failed");
  HIPCHK(hipEventRecord(c->ws.ev[1], s));
  if (launch_finalize(G_ZUKER, a, 1, n, s)) return fail(err, DPX_EHIP, "finalize launch failed");
  HIPCHK(hipEventRecord(c->ws.ev[2], s));
  int32_t out[8] = {0};  // [0] MFE, [1] trace length, [4..7] ctrl (queue cursors, time-out flag)
  HIPCHK(hipMemcpyAsync(&out[0], a.results, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (trace) HIPCHK(hipMemcpyAsync(&out[1], a.trace_lens, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&out[4], a.ctrl, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipEventRecord(c->ws.ev[3], s));
  HIPCHK(hipEventSynchronize(c->ws.ev[3]));
  (void)hipEventElapsedTime(&c->t_fill, c->ws.ev[0], c->ws.ev[1]);
  (void)hipEventElapsedTime(&c->t_fin, c->ws.ev[1], c->ws.ev[2]);
  if (out[4 + 2]) return fail(err, DPX_EHIP, "tile queue wait timed out (scheduler error)");
  *mfe = out[0];
  if (trace) {
    const int tl = out[1];
    *trace_len = tl;
    if (tl < 0) return fail(err, DPX_ETRACE, "backtrace failed");
    if (tl > trace_cap) return fail(err, DPX_ETRACE, "trace buffer too small");
    HIPCHK(hipMemcpy(trace, a.traces, sizeof(dpx_trace_t) * (size_t)tl, hipMemcpyDeviceToHost));
  }
  return DPX_OK;
}

int dpx_rnafold_batch(dpx_ctx* c, const char* seqs, const int64_t* offs, int count, int32_t* results,
                      dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  (void)hipSetDevice(c->device);
  return rna_batch(c, G_RNAFOLD, seqs, offs, count, 0, results, traces, trace_offs, trace_lens);
}

int dpx_matchain(dpx_ctx* c, const int32_t* rows, const int32_t* cols, int n, int32_t* cost, dpx_trace_t* trace,
                 int trace_cap, int* trace_len) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (!rows || !cols || n < 1 || n > kMaxTri || !cost || (trace && (!trace_len || trace_cap < 0)))
    return fail(err, DPX_EINVAL, "bad argument");
  (void)hipSetDevice(c->device);
  Job j;
  std::vector<int> lens{n};
  std::vector<int64_t> off{0};
  int r = tri_prepare(err, j, G_MATCHAIN, lens, off, chain_data(rows, cols, n), 0, trace != nullptr);
  if (!r) r = run_sync(c, j);
  if (r) return r;
  int64_t toffs[2] = {0, trace_cap};
  int32_t tl = 0;
  r = tri_copy_out(err, j, cost, trace, toffs, &tl);
  if (trace) *trace_len = tl;
  return r;
}

int dpx_swat(dpx_ctx* c, const char* a, int n1, const char* b, int n2, const dpx_swat_scores* sc, int32_t* score,
             int32_t* best, int32_t* best_i, int32_t* best_j, int trace_from, dpx_trace_t* trace, int trace_cap,
             int* trace_len) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (!a || !b || n1 < 0 || n2 < 0 || !sc || !score || (trace && (!trace_len || trace_cap < 0)) ||
      (trace_from != 0 && trace_from != 1))
    return fail(err, DPX_EINVAL, "bad argument");
  (void)hipSetDevice(c->device);
  Job j;
  int r = sw_prepare(err, j, a, n1, b, n2, *sc, trace_from, trace ? trace_cap : 0,
                     sw_segment(n1, n2, c->sw_code_budget));
  if (!r) r = run_sync(c, j);
  if (r) return r;
  *score = j.res[0];
  if (best) *best = j.res[1];
  if (best_i) *best_i = j.res[2];
  if (best_j) *best_j = j.res[3];
  if (trace) {
    const int tl = j.res[6];
    *trace_len = tl;
    if (tl < 0) return fail(err, DPX_ETRACE, "trace buffer too small");
    std::memcpy(trace, j.tr, sizeof(dpx_trace_t) * tl);
  }
  return DPX_OK;
}

// ---------------------------------------------------------------- plans

}  // extern "C"

struct dpx_plan {
  dpx_ctx* ctx = nullptr;
  Job job;
  DevBuf block;
  Workspace ws;
  // two graphs replayed back to back on the plan's stream: queue reset +
  // persistent fill, then finalize; events between them time the kernels
  hipGraph_t graph = nullptr, graph_fin = nullptr;
  hipGraphExec_t exec = nullptr, exec_fin = nullptr;
  bool ran = false;
  float fill_ms = -1, fin_ms = -1;  // kernel times of the last replay
};

namespace {

void plan_free(dpx_plan* P) {
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  if (P->exec_fin) (void)hipGraphExecDestroy(P->exec_fin);
  if (P->graph_fin) (void)hipGraphDestroy(P->graph_fin);
  P->ws.release();
  P->block.release();
}

// Upload the job's inputs once and capture reset + fill + finalize as a graph.
int plan_build(dpx_plan* P) {
  dpx_ctx* c = P->ctx;
  std::string& err = c->err;
  Job& j = P->job;
  if (!P->ws.init()) return fail(err, DPX_EHIP, "stream creation failed");
  if (!P->block.ensure(j.dev_bytes))
    return fail(err, DPX_ENOMEM, "device allocation for the DP tables failed (" + std::to_string(j.dev_bytes >> 20) + " MiB)");
  char* dev = (char*)P->block.p;
  HIPCHK(hipMemcpy(dev + j.off[TL_INST], j.inst.data(), sizeof(TriInst) * j.count, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dev + j.off[TL_SEEDS], j.seeds.data(), j.seeds.size() * 8, hipMemcpyHostToDevice));
  if (!j.data.empty()) HIPCHK(hipMemcpy(dev + j.off[TL_DATA], j.data.data(), j.data.size(), hipMemcpyHostToDevice));
  if (j.mp_split)
    HIPCHK(hipMemcpy(dev + j.off[TL_MP] + j.tiles * 4, j.mp_base.data(), j.mp_base.size() * 4, hipMemcpyHostToDevice));
  const TriArgs a = tri_args(c, j, dev);
  const int grid = tri_grid(c, j);
  hipStream_t s = P->ws.stream;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int r1 = launch_fill(j.g, a, (const unsigned long long*)(dev + j.off[TL_SEEDS]), (int)j.seeds.size(), grid, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  P->graph = g;
  if (r1 || e != hipSuccess) return fail(err, DPX_EHIP, "graph capture of the fill launch failed");
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int r2 = launch_finalize(j.g, a, j.count, j.n_max, s);
  g = nullptr;
  e = hipStreamEndCapture(s, &g);
  P->graph_fin = g;
  if (r2 || e != hipSuccess) return fail(err, DPX_EHIP, "graph capture of the finalize launch failed");
  HIPCHK(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
  HIPCHK(hipGraphInstantiate(&P->exec_fin, P->graph_fin, nullptr, nullptr, 0));
  return DPX_OK;
}

}  // namespace

extern "C" {

int dpx_plan_create(dpx_ctx* c, int kind, const char* seqs, const int64_t* offs, int count, int min_loop, int want_trace,
                    dpx_plan** out) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (!out || count < 1 || !seqs || !offs || min_loop < 0 ||
      (kind != DPX_REQ_ZUKER && kind != DPX_REQ_NUSSINOV && kind != DPX_REQ_RNAFOLD))
    return fail(err, DPX_EINVAL, "bad argument");
  *out = nullptr;
  const int g = grammar_of(kind);
  if (g != G_NUSSINOV && !c->has_rna) return fail(err, DPX_ENOPARAMS, "dpx_set_rna_params was not called");
  (void)hipSetDevice(c->device);
  dpx_plan* P = new dpx_plan();
  P->ctx = c;
  int r = rna_job(err, P->job, g, seqs, offs, count, g == G_NUSSINOV ? min_loop : 0, want_trace != 0);
  if (!r) r = plan_build(P);
  if (r) {
    plan_free(P);
    delete P;
    return r;
  }
  c->plans++;
  *out = P;
  return DPX_OK;
}

int dpx_plan_run(dpx_plan* P, float* device_ms) {
  if (!P) return DPX_EINVAL;
  dpx_ctx* c = P->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  (void)hipSetDevice(c->device);
  hipStream_t s = P->ws.stream;
  HIPCHK(hipEventRecord(P->ws.ev[0], s));
  HIPCHK(hipGraphLaunch(P->exec, s));
  HIPCHK(hipEventRecord(P->ws.ev[2], s));
  HIPCHK(hipGraphLaunch(P->exec_fin, s));
  HIPCHK(hipEventRecord(P->ws.ev[1], s));
  int32_t flags[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(flags, (char*)P->block.p + P->job.off[TL_CTRL], sizeof flags, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, P->ws.ev[0], P->ws.ev[1]);
  if (device_ms) *device_ms = ms;
  // fill = queue reset + persistent fill kernel; finalize = the second graph
  P->fill_ms = -1;
  P->fin_ms = -1;
  float f = 0, z = 0;
  if (hipEventElapsedTime(&f, P->ws.ev[0], P->ws.ev[2]) == hipSuccess &&
      hipEventElapsedTime(&z, P->ws.ev[2], P->ws.ev[1]) == hipSuccess) {
    P->fill_ms = f;
    P->fin_ms = z;
  }
  if (flags[2]) return fail(err, DPX_EHIP, "tile queue wait timed out (scheduler error)");
  P->ran = true;
  return DPX_OK;
}

int dpx_plan_timing(dpx_plan* P, float* fill_ms, float* finalize_ms) {
  if (!P || !fill_ms || !finalize_ms) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(P->ctx->mu);
  if (!P->ran || P->fill_ms < 0) return fail(P->ctx->err, DPX_EHIP, "no timed replay");
  *fill_ms = P->fill_ms;
  *finalize_ms = P->fin_ms;
  return DPX_OK;
}

int dpx_plan_results(dpx_plan* P, int32_t* results, dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens) {
  if (!P) return DPX_EINVAL;
  dpx_ctx* c = P->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (!results || (traces && (!trace_offs || !trace_lens))) return fail(err, DPX_EINVAL, "bad argument");
  if (!P->ran) return fail(err, DPX_EINVAL, "dpx_plan_run was not called");
  (void)hipSetDevice(c->device);
  Job& j = P->job;
  const char* dev = (const char*)P->block.p;
  const size_t out_bytes = 8 * (size_t)j.count + 64 + (j.want_trace ? sizeof(dpx_trace_t) * (size_t)j.trace_total : 0);
  if (!P->ws.h_out.ensure(out_bytes)) return fail(err, DPX_ENOMEM, "pinned host staging");
  int32_t* o_res = (int32_t*)P->ws.h_out.p;
  int32_t* o_lens = o_res + j.count;
  dpx_trace_t* o_tr = (dpx_trace_t*)(((uintptr_t)(o_lens + j.count + 4) + 15) & ~(uintptr_t)15);
  HIPCHK(hipMemcpy(o_res, dev + j.off[TL_RES], sizeof(int32_t) * j.count, hipMemcpyDeviceToHost));
  if (j.want_trace) {
    HIPCHK(hipMemcpy(o_lens, dev + j.off[TL_LENS], sizeof(int32_t) * j.count, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(o_tr, dev + j.off[TL_TRACES], sizeof(dpx_trace_t) * j.trace_total, hipMemcpyDeviceToHost));
  }
  j.res = o_res;
  j.lens = o_lens;
  j.tr = o_tr;
  return tri_copy_out(err, j, results, traces, trace_offs, trace_lens);
}

void dpx_plan_destroy(dpx_plan* P) {
  if (!P) return;
  (void)hipSetDevice(P->ctx->device);
  plan_free(P);
  P->ctx->plans--;
  delete P;
}

int dpx_last_timing(const dpx_ctx* c, float* fill_ms, float* finalize_ms) {
  if (!c) return DPX_EINVAL;
  if (fill_ms) *fill_ms = c->t_fill;
  if (finalize_ms) *finalize_ms = c->t_fin;
  return DPX_OK;
}

double dpx_now(void) { return now_s(); }

// ================================================================== tuning

int dpx_tune_bucket(int n) { return n < 0 ? -1 : tune_bucket(n); }

int dpx_set_swat_trace_budget(dpx_ctx* c, int64_t bytes) {
  if (!c || bytes < 0) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  c->sw_code_budget = bytes;
  return DPX_OK;
}

int dpx_get_tuning(dpx_ctx* c, int kind, int bucket, dpx_tuning* out) {
  if (!c || !out || kind < DPX_REQ_NUSSINOV || kind > DPX_REQ_RNAFOLD || bucket < 0 || bucket >= DPX_TUNE_BUCKETS)
    return DPX_EINVAL;
  *out = get_tune(c, kind, bucket);
  return DPX_OK;
}

int dpx_set_tuning(dpx_ctx* c, int kind, int bucket, const dpx_tuning* t) {
  if (!c || !t || kind < DPX_REQ_NUSSINOV || kind > DPX_REQ_RNAFOLD || bucket < 0 || bucket >= DPX_TUNE_BUCKETS)
    return DPX_EINVAL;
  if (t->wg_per_cu < 0 || (t->chunk != 0 && t->chunk != 16 && t->chunk != 32 && t->chunk != 64)) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->tune_mu);
  c->tune[kind][bucket] = *t;
  return DPX_OK;
}

}  // extern "C"

namespace {

struct Mix64 {  // splitmix64: seeded synthetic inputs of the tuning runs
  uint64_t x;
  uint64_t next() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// lengths of bucket b: (32 << b, 64 << b], bucket 0 from 16
void bucket_range(int b, int* lo, int* hi) {
  *lo = b == 0 ? 16 : (32 << b) + 1;
  *hi = 64 << b;
}

// A seeded synthetic launch of `kind` with lengths in bucket b, about
// 4e7 cells (at least one instance): the shape the tuning is chosen for.
int tune_job(std::string& err, int kind, int b, int max_n, Job& j, int* count_out) {
  int lo, hi;
  bucket_range(b, &lo, &hi);
  hi = std::min(hi, max_n);
  lo = std::min(lo, hi);
  Mix64 r{0x7475ull * 1000003ull + (uint64_t)kind * 131ull + (uint64_t)b};
  auto len = [&]() { return (int)std::lround(std::exp(std::log((double)lo) + r.unit() * (std::log((double)hi) - std::log((double)lo)))); };
  if (kind == DPX_REQ_SWAT) {
    const int n1 = hi, n2 = hi;
    std::string a(n1, 'A'), bb(n2, 'A');
    for (auto* sq : {&a, &bb})
      for (char& ch : *sq) ch = "ACGT"[r.next() & 3];
    *count_out = 1;
    return sw_prepare(err, j, a.data(), n1, bb.data(), n2, dpx_swat_scores{2, -3, -5, -2}, 1, sw_trace_cap(n1, n2));
  }
  std::vector<int> lens;
  std::vector<int64_t> off;
  std::vector<uint8_t> data;
  int64_t cells = 0;
  while ((cells < 40000000 && (int)lens.size() < 4096) || lens.empty()) {
    const int n = len();
    lens.push_back(n);
    // matrix chain offsets count int32 dims, RNA offsets bytes
    off.push_back(kind == DPX_REQ_MATCHAIN ? (int64_t)(data.size() / 4) : (int64_t)data.size());
    cells += tri_cells(n);
    if (kind == DPX_REQ_MATCHAIN) {
      std::vector<int32_t> rc(2 * (size_t)n);
      int32_t prev = 1 + (int32_t)(r.next() % 1000);
      for (int q = 0; q < n; ++q) {  // matrix q is rows[q] x cols[q], chained
        const int32_t nx = 1 + (int32_t)(r.next() % 1000);
        rc[q] = prev;
        rc[n + q] = nx;
        prev = nx;
      }
      const size_t at = data.size();
      data.resize(at + rc.size() * 4);
      std::memcpy(data.data() + at, rc.data(), rc.size() * 4);
    } else {
      for (int q = 0; q < n; ++q) {  // 50% GC
        const double u = r.unit();
        data.push_back(u < 0.25 ? 3 : u < 0.5 ? 2 : u < 0.75 ? 1 : 4);
      }
      data.push_back(0);
    }
  }
  *count_out = (int)lens.size();
  const int g = kind == DPX_REQ_MATCHAIN ? G_MATCHAIN : grammar_of(kind);
  return tri_prepare(err, j, g, lens, off, std::move(data), kind == DPX_REQ_NUSSINOV ? 3 : 0, true);
}

// everything a launch returns, to check that the shape does not change it
std::vector<int32_t> job_outputs(const Job& j) {
  std::vector<int32_t> v;
  if (j.kind == JOB_SW) {
    v.assign(j.res, j.res + 7);
    const int32_t* t = (const int32_t*)j.tr;
    if (j.res[6] > 0) v.insert(v.end(), t, t + 5 * (size_t)j.res[6]);
    return v;
  }
  v.assign(j.res, j.res + j.count);
  v.insert(v.end(), j.lens, j.lens + j.count);
  const int32_t* t = (const int32_t*)j.tr;
  v.insert(v.end(), t, t + 5 * (size_t)j.trace_total);
  return v;
}

}  // namespace

extern "C" {

int dpx_autotune(dpx_ctx* c, int kind, int bucket_lo, int bucket_hi, int reps, dpx_tune_report* report,
                 int report_cap, int* report_len) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (kind < DPX_REQ_NUSSINOV || kind > DPX_REQ_RNAFOLD || bucket_lo < 0 || bucket_hi >= DPX_TUNE_BUCKETS ||
      bucket_lo > bucket_hi || reps < 1 || (report && report_cap < 0))
    return fail(err, DPX_EINVAL, "bad argument");
  if ((kind == DPX_REQ_ZUKER || kind == DPX_REQ_RNAFOLD) && !c->has_rna)
    return fail(err, DPX_ENOPARAMS, "dpx_set_rna_params was not called");
  (void)hipSetDevice(c->device);
  const int g = kind == DPX_REQ_MATCHAIN ? G_MATCHAIN : kind == DPX_REQ_SWAT ? -1 : grammar_of(kind);
  // candidate shapes: workgroups per CU (triangular), steps per chunk (alignment)
  std::vector<dpx_tuning> cands;
  if (kind == DPX_REQ_SWAT) {
    for (int ch : {16, 32, 64}) cands.push_back(dpx_tuning{0, ch, {0, 0}});
  } else {
    for (int w : {1, 2, 3, 4, 6, 8})
      if (w <= c->occ[g]) cands.push_back(dpx_tuning{w, 0, {0, 0}});
  }
  const int max_n = kind == DPX_REQ_SWAT ? 8192 : std::min(8192, max_len(g));
  int nrep = 0;
  for (int b = bucket_lo; b <= bucket_hi; ++b) {
    Job j;
    int count = 0;
    int r = tune_job(err, kind, b, max_n, j, &count);
    if (r) return r;
    const dpx_tuning saved = get_tune(c, kind, b);
    std::vector<int32_t> want;
    float best_ms = 0;
    int best = -1, first_rep = nrep;
    for (size_t q = 0; q < cands.size(); ++q) {
      {
        std::lock_guard<std::mutex> tl(c->tune_mu);
        c->tune[kind][b] = cands[q];
      }
      float ms = 0;
      for (int k = 0; k < reps; ++k) {
        r = run_sync(c, j);
        if (r) {
          std::lock_guard<std::mutex> tl(c->tune_mu);
          c->tune[kind][b] = saved;
          return r;
        }
        const float t = j.t_fill + j.t_fin;
        ms = k == 0 ? t : std::min(ms, t);
      }
      std::vector<int32_t> got = job_outputs(j);
      if (q == 0) want = std::move(got);
      else if (got != want) {
        std::lock_guard<std::mutex> tl(c->tune_mu);
        c->tune[kind][b] = saved;
        size_t at = 0;
        while (at < got.size() && at < want.size() && got[at] == want[at]) ++at;
        return fail(err, DPX_EHIP, "results differ between launch shapes (bucket " + std::to_string(b) + ", shape " +
                                       std::to_string(q) + " vs 0, first difference at output word " +
                                       std::to_string(at) + " of " + std::to_string(want.size()) + ": " +
                                       std::to_string(at < got.size() ? got[at] : 0) + " vs " +
                                       std::to_string(at < want.size() ? want[at] : 0) + ")");
      }
      if (best < 0 || ms < best_ms) { best_ms = ms; best = (int)q; }
      if (report && nrep < report_cap) {
        dpx_tune_report& e = report[nrep];
        e.kind = kind;
        e.bucket = b;
        bucket_range(b, &e.n_lo, &e.n_hi);
        e.n_hi = std::min(e.n_hi, max_n);
        e.count = count;
        e.t = cands[q];
        e.ms = ms;
        e.chosen = 0;
      }
      ++nrep;
    }
    {
      std::lock_guard<std::mutex> tl(c->tune_mu);
      c->tune[kind][b] = cands[best];
    }
    if (report && first_rep + best < report_cap) report[first_rep + best].chosen = 1;
  }
  if (report_len) *report_len = nrep;
  return DPX_OK;
}

// ---------------------------------------------------------------- server

int dpx_server_create(dpx_ctx* c, const dpx_server_config* cfg, dpx_server** out) {
  if (!c) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  std::string& err = c->err;
  if (!cfg || !out || cfg->slots < 1 || cfg->slots > 8 || cfg->max_batch < 1 || cfg->hbm_budget < (1 << 20) ||
      cfg->max_pending < 1)
    return fail(err, DPX_EINVAL, "bad server configuration");
  *out = nullptr;
  (void)hipSetDevice(c->device);
  dpx_server* S = new dpx_server();
  S->ctx = c;
  S->cfg = *cfg;
  const size_t budget = a256((size_t)cfg->hbm_budget);
  if (!S->arena.mem.ensure(budget)) {
    delete S;
    return fail(err, DPX_ENOMEM, "cannot reserve the HBM budget");
  }
  S->arena.free_[0] = budget;
  S->slots.resize(cfg->slots);
  for (Slot& sl : S->slots)
    if (!sl.ws.init()) {
      for (Slot& x : S->slots) x.ws.release();
      S->arena.mem.release();
      delete S;
      return fail(err, DPX_EHIP, "stream creation failed");
    }
  c->servers++;
  S->th = std::thread(server_loop, S);
  *out = S;
  return DPX_OK;
}

void dpx_server_destroy(dpx_server* S) {
  if (!S) return;
  S->stop.store(true);
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  (void)hipSetDevice(S->ctx->device);
  for (Slot& sl : S->slots) sl.ws.release();
  S->arena.mem.release();
  S->ctx->servers--;
  delete S;
}

int dpx_server_submit(dpx_server* S, const dpx_request* req, uint64_t* id) {
  if (!S || !req || !id) return DPX_EINVAL;
  Request q;
  q.kind = req->kind;
  q.trace = (req->flags & DPX_SUBMIT_TRACE) ? 1 : 0;
  q.trace_from = (req->flags & DPX_SUBMIT_TRACE_FROM_BEST) ? 1 : 0;
  const int n1 = req->n1, n2 = req->n2;
  if (q.kind == DPX_REQ_ZUKER || q.kind == DPX_REQ_NUSSINOV || q.kind == DPX_REQ_RNAFOLD) {
    const int g = grammar_of(q.kind);
    if (!req->a || n1 < 0 || n1 > max_len(g) || req->min_loop < 0) return DPX_EINVAL;
    if (g != G_NUSSINOV && !S->ctx->has_rna) return DPX_ENOPARAMS;
    for (int x = 0; x < n1; ++x)
      if (!req->a[x] || !std::strchr("ACGUTacgut", req->a[x])) return DPX_EINVAL;
    q.s1.assign(req->a, n1);
    q.min_loop = g == G_NUSSINOV ? req->min_loop : 0;
    q.footprint = tri_footprint(g, n1, q.trace);
    q.cells = tri_cells(n1);
    q.size_class = size_class(n1);
  } else if (q.kind == DPX_REQ_MATCHAIN) {
    if (!req->rows || !req->cols || n1 < 1 || n1 > kMaxTri) return DPX_EINVAL;
    q.rows.assign(req->rows, req->rows + n1);
    q.cols.assign(req->cols, req->cols + n1);
    q.footprint = tri_footprint(G_MATCHAIN, n1, q.trace);
    q.cells = tri_cells(n1);
    q.size_class = size_class(n1);
  } else if (q.kind == DPX_REQ_SWAT) {
    if (!req->a || !req->b || n1 < 0 || n2 < 0 || (int64_t)(n1 + 1) * (n2 + 1) >= ((int64_t)1 << 32)) return DPX_EINVAL;
    q.s1.assign(req->a, n1);
    q.s2.assign(req->b, n2);
    q.sc = req->sc;
    q.footprint = swb_pair_bytes(n1, n2, q.trace ? sw_trace_cap(n1, n2) : 0);
    q.cells = (int64_t)(n1 + 1) * (n2 + 1);
    q.size_class = size_class(std::max(n1, n2));
  } else {
    return DPX_EINVAL;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  if (q.footprint > a256((size_t)S->cfg.hbm_budget)) {
    S->st.rejected++;
    return DPX_ENOMEM;
  }
  if ((int64_t)S->pending.size() >= S->cfg.max_pending) {
    S->st.rejected++;
    return DPX_EBUSY;
  }
  q.t_submit = now_s();
  q.id = S->next_id++;
  *id = q.id;
  S->st.submitted++;
  S->pending.push_back(std::move(q));
  S->st.pending = (int64_t)S->pending.size();
  S->cv_work.notify_all();
  return DPX_OK;
}

int dpx_server_poll(dpx_server* S, dpx_result* out, int max, int wait_us, int* got) {
  if (!S || !got || max < 0 || (max > 0 && !out)) return DPX_EINVAL;
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->ready.empty() && wait_us > 0)
    S->cv_done.wait_for(lk, std::chrono::microseconds(wait_us), [&] { return !S->ready.empty(); });
  int n = 0;
  while (n < max && !S->ready.empty()) {
    const uint64_t id = S->ready.front();
    S->ready.pop_front();
    auto it = S->done.find(id);
    if (it == S->done.end()) continue;
    out[n++] = it->second.r;
    if (!it->second.trace.empty()) S->traces[id] = std::move(it->second.trace);
    S->done.erase(it);
  }
  *got = n;
  return n ? DPX_OK : DPX_EAGAIN;
}

int dpx_server_trace(dpx_server* S, uint64_t id, dpx_trace_t* buf, int cap, int* len) {
  if (!S || !len) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->traces.find(id);
  if (it == S->traces.end()) {
    *len = 0;
    return DPX_EAGAIN;
  }
  *len = (int)it->second.size();
  if (buf) {
    if (cap < *len) return DPX_ETRACE;
    std::memcpy(buf, it->second.data(), sizeof(dpx_trace_t) * it->second.size());
  }
  S->traces.erase(it);
  return DPX_OK;
}

int dpx_server_get_stats(dpx_server* S, dpx_server_stats* out) {
  if (!S || !out) return DPX_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->st;
  return DPX_OK;
}

// ---------------------------------------------------------------- generic grammars

struct dpx_grammar {
  dpx::adp::Grammar* g = nullptr;
  int device = 0;
  std::mutex mu;
};

int dpx_grammar_create(dpx_ctx* c, const char* spec, dpx_grammar** out) {
  if (!c || !spec || !out) return DPX_EINVAL;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c->err, DPX_EHIP, "hipSetDevice");
  int code = DPX_EINVAL;
  dpx::adp::Grammar* g = dpx::adp::compile(spec, true, c->err, &code);
  if (!g) return code;
  dpx_grammar* h = new dpx_grammar();
  h->g = g;
  h->device = c->device;
  *out = h;
  return DPX_OK;
}

void dpx_grammar_destroy(dpx_grammar* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  dpx::adp::destroy(h->g);
  delete h;
}

int dpx_grammar_check(const char* spec, const char** src, const char** msg) {
  static thread_local std::string s_src, s_err;
  s_src.clear();
  s_err.clear();
  if (src) *src = s_src.c_str();
  if (msg) *msg = s_err.c_str();
  dpx::adp::Grammar* g = dpx::adp::compile(spec, false, s_err);
  if (msg) *msg = s_err.c_str();
  if (!g) return DPX_EINVAL;
  s_src = dpx::adp::source(g);
  dpx::adp::destroy(g);
  if (src) *src = s_src.c_str();
  return DPX_OK;
}

int dpx_grammar_get_info(const dpx_grammar* h, dpx_grammar_info* out) {
  if (!h || !out) return DPX_EINVAL;
  dpx::adp::info(h->g, out);
  return DPX_OK;
}

const char* dpx_grammar_source(const dpx_grammar* h) { return h ? dpx::adp::source(h->g).c_str() : nullptr; }

int dpx_grammar_parse(dpx_ctx* c, dpx_grammar* h, const void* in1, int n1, const void* in2, int n2, void* answer,
                      int* found, dpx_trace_t* trace, int cap, int* len) {
  if (!c || !h) return DPX_EINVAL;
  if (h->device != c->device) return fail(c->err, DPX_EINVAL, "grammar was compiled for another device");
  std::lock_guard<std::mutex> lk(c->mu);
  std::lock_guard<std::mutex> lk2(h->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c->err, DPX_EHIP, "hipSetDevice");
  float ms = 0;
  const int rc = dpx::adp::parse(h->g, c->ws.stream, in1, n1, in2, n2, answer, found, trace, cap, len,
                                 c->has_rna ? c->d_rna.p : nullptr, c->has_rna ? (const int*)c->d_hl.p : nullptr, &ms, c->err);
  c->t_fill = ms;
  c->t_fin = 0;
  return rc;
}

int dpx_grammar_parse_batch(dpx_ctx* c, dpx_grammar* h, const void* in1, const int64_t* offs1, const void* in2,
                            const int64_t* offs2, int count, void* answers, int32_t* found, dpx_trace_t* traces,
                            const int64_t* trace_offs, int32_t* trace_lens) {
  if (!c || !h) return DPX_EINVAL;
  if (h->device != c->device) return fail(c->err, DPX_EINVAL, "grammar was compiled for another device");
  std::lock_guard<std::mutex> lk(c->mu);
  std::lock_guard<std::mutex> lk2(h->mu);
  if (hipSetDevice(c->device) != hipSuccess) return fail(c->err, DPX_EHIP, "hipSetDevice");
  float ms = 0;
  const int rc = dpx::adp::parse_batch(h->g, c->ws.stream, count, (const char*)in1, offs1, (const char*)in2, offs2, answers,
                                       found, traces, trace_offs, trace_lens, c->has_rna ? c->d_rna.p : nullptr,
                                       c->has_rna ? (const int*)c->d_hl.p : nullptr, &ms, c->err);
  c->t_fill = ms;
  c->t_fin = 0;
  return rc;
}

}  // extern "C"
