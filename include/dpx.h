/*
 * dpx.h -- C ABI of the MI355X DP backend (libdpx.so).
 *
 * Drop-in replacement for the native layer that the reference's CodeGen
 * backend generates and loads over JNI.  The reference emits, per grammar and
 * per size, a library with
 *     void my_init(input_t* in1, input_t* in2); void my_solve();
 *     int  my_backtrack(trace_t** trace, unsigned* size); void my_free();
 * (src/main/scala/v4/CodeGen.scala:291-358, frozen copies in bench2/zu.c:95-111
 * and bench2/nu.c:61-73) reached from Java_CompWrapperN_parse/backtrack
 * (CodeGen.scala:366-390).  This ABI keeps the same inputs, the same answer
 * and the same backtrace records (rule ids and split positions of the
 * reference's `analyze`, in the reference BFS order), but:
 *   - sizes are runtime arguments (the reference bakes M_W/M_H in),
 *   - a context handle replaces static globals and keeps energy tables and DP
 *     scratch resident on the GPU between calls,
 *   - errors are status codes (the reference calls exit(), CodeGen.scala:261),
 *   - batches of instances and a serving queue are added.
 * Plain C types only; all pointers are host pointers.
 */
#ifndef DPX_H
#define DPX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpx_ctx dpx_ctx;

/* Status codes (0 = success). */
enum {
  DPX_OK = 0,
  DPX_EINVAL = 1,    /* bad argument (size, symbol outside ACGU/ACGT, null pointer) */
  DPX_ENOMEM = 2,    /* device or host allocation failed / over the HBM budget */
  DPX_EHIP = 3,      /* HIP runtime error */
  DPX_ENOPARAMS = 4, /* RNA grammar used before dpx_set_rna_params */
  DPX_EPARAMS = 5,   /* parameter file unreadable or malformed */
  DPX_ETRACE = 6,    /* trace buffer too small (trace_len holds the size needed) */
  DPX_EBUSY = 7,     /* serving queue full */
  DPX_EAGAIN = 8     /* no finished request yet */
};

/* One backtrace record: the reference trace_t {short i,j,rule; short pos[]}
 * (CodeGen.scala:239; bench2/zu.h:17; bench2/nu.h:10) widened to int32.
 * rule is the winning subrule id; pos[] its split indices (0 where unused). */
typedef struct { int32_t i, j, rule, pos[2]; } dpx_trace_t;

/* Context: one per device.  dpx_create selects the HIP device and a stream. */
int dpx_create(dpx_ctx** out, int device);
void dpx_destroy(dpx_ctx* ctx);
/* Text of the last error on this context (never NULL). */
const char* dpx_last_error(const dpx_ctx* ctx);

/* Load a ViennaRNA v2.0 .par file (semantics of src/librna/vienna/vienna.c:
 * read_parameter_file + get_scaled_parameters at 37 C) and upload the scaled
 * tables to the device.  Replaces the generated rna_init (CodeGen.scala:417-425)
 * and ADPParsers.setParams (ADPParsers.scala:85-88). */
int dpx_set_rna_params(dpx_ctx* ctx, const char* par_path);

/* ---- single instance (parse + backtrack of one input) --------------------
 * trace may be NULL (parse only, the reference's CompWrapperN.parse); else it
 * must hold trace_cap records and *trace_len receives the count.  Sequences
 * are ASCII letters, case-insensitive. */

/* Nussinov78 (v4/examples/Nussinov.scala:49-76; bench2/nu_kern.h).  Rules
 * 0 nil, 1 left, 2 right, 3 pair, 4 split(pos k).  min_loop = 0 is the
 * reference grammar (a pair needs i+2<=j); min_loop = m > 0 additionally asks
 * for m unpaired bases inside every pair. */
int dpx_nussinov(dpx_ctx* ctx, const char* seq, int n, int min_loop, int32_t* score,
                 dpx_trace_t* trace, int trace_cap, int* trace_len);

/* Zuker MFE (v4/examples/Zuker.scala:64-147; bench2/zu_kern.h): tables cl
 * (rules 0-5), ml (6-7), ml1 (8-11), st (12-14); energies in dcal/mol. */
int dpx_zuker(dpx_ctx* ctx, const char* seq, int n, int32_t* mfe,
              dpx_trace_t* trace, int trace_cap, int* trace_len);

/* dpx_zuker without any allocation: no reference counterpart.  For callers
 * that cannot allocate device or pinned host memory (code running beside a
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and traces as dpx_zuker, for n <= 1024: the DP
 * tables live in a static 16 MiB scratch array per device (held until the
 * process ends), inputs are staged from the stack.  n > 1024 gives
 * DPX_EINVAL.  Calls on one device are serialised by the library; the
 * scratch keeps the last call's contents (every call clears what it needs). */
int dpx_zuker_noalloc(dpx_ctx* ctx, const char* seq, int n, int32_t* mfe,
                      dpx_trace_t* trace, int trace_cap, int* trace_len);

/* RNAfold MFE grammar (v4/examples/RNAFold.scala:20-64): tables Q (rules
 * 0 QM, 1 Q(i,j-1), 2 Q(i+1,j), 3 Qp), QM (rule 4 split(pos k)), Qp (rules
 * 5 hairpin, 6 stack, 7 interior loop (pos k, l) with both unpaired
 * stretches < 30, 8 multiloop); energies in dcal/mol; 999999 when the
 * grammar admits no structure (Q(0,n) empty). */
int dpx_rnafold(dpx_ctx* ctx, const char* seq, int n, int32_t* mfe,
                dpx_trace_t* trace, int trace_cap, int* trace_len);

/* Smith-Waterman with affine gaps, two-track grammar (v4/examples/
 * SWatAffine.scala:22-69; bench/dp_swat/comp1.cu:50-132) with parametric
 * scores: pair = M + (a==b ? match : mismatch); open = max(0, x + gap_open);
 * extend = max(0, x + gap_ext).  Rules: M 0-3, g1 4-6, g2 7-9.
 * *score receives the reference axiom M(n1,n2); *best the local score
 * max_ij M(i,j) at (*best_i,*best_j), the first maximum in row-major order.
 * trace_from: 0 = trace from (n1,n2) like the reference, 1 = from the best
 * cell.  Sequences are letters ACGT/ACGU. */
typedef struct { int32_t match, mismatch, gap_open, gap_ext; } dpx_swat_scores;
/* Traceback memory of dpx_swat: when the 4-bit codes of every cell
 * (about n1*n2/2 bytes) would exceed `bytes`, only checkpoint rows every
 * ~sqrt(n1/64) bands are kept (8 bytes per column each) and the traceback
 * recomputes one segment at a time from its checkpoint: the same trace, for
 * about one more fill (the checkpoint strategy of docs/report/design.tex:
 * 226-262).  Default 32 GiB.  No reference counterpart. */
int dpx_set_swat_trace_budget(dpx_ctx* ctx, int64_t bytes);
int dpx_swat(dpx_ctx* ctx, const char* a, int n1, const char* b, int n2, const dpx_swat_scores* sc,
             int32_t* score, int32_t* best, int32_t* best_i, int32_t* best_j, int trace_from,
             dpx_trace_t* trace, int trace_cap, int* trace_len);

/* Matrix-chain ordering, Int algebra (v4/report/Benchmarks.scala:18-27,
 * v4/examples/MatrixMult2.scala:16-27): M(i,j) = min_k M(i,k)+M(k,j) +
 * rows[i]*rows[k]*cols[j-1] with JVM int32 wrap-around.  Rules 0 single,
 * 1 mult(pos k).  matrix m is rows[m] x cols[m]. */
int dpx_matchain(dpx_ctx* ctx, const int32_t* rows, const int32_t* cols, int n, int32_t* cost,
                 dpx_trace_t* trace, int trace_cap, int* trace_len);

/* ---- batches: many instances in one persistent launch ---------------------
 * seqs holds all sequences back to back; instance q is seqs[offs[q]..offs[q+1]).
 * results[q] receives the MFE.  If traces != NULL, instance q's trace is
 * written at traces + trace_offs[q] (capacity trace_offs[q+1]-trace_offs[q])
 * and its length in trace_lens[q].  No reference counterpart (the reference
 * runs one synchronous instance per call, CodeGen.scala:309-310). */
int dpx_zuker_batch(dpx_ctx* ctx, const char* seqs, const int64_t* offs, int count, int32_t* results,
                    dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens);
int dpx_nussinov_batch(dpx_ctx* ctx, const char* seqs, const int64_t* offs, int count, int min_loop,
                       int32_t* results, dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens);
int dpx_rnafold_batch(dpx_ctx* ctx, const char* seqs, const int64_t* offs, int count, int32_t* results,
                      dpx_trace_t* traces, const int64_t* trace_offs, int32_t* trace_lens);

/* Timing of the last synchronous call on this context: device time (ms) of
 * the fill kernel (persistent tile queue / alignment bands) and of the
 * finalize kernels (answer, st column, backtrace). */
int dpx_last_timing(const dpx_ctx* ctx, float* fill_ms, float* finalize_ms);

/* ---- plans: a batch uploaded once and replayed as a HIP graph -------------
 * No reference counterpart.  dpx_plan_create copies the batch (Zuker,
 * Nussinov or RNAfold, as DPX_REQ_ZUKER / NUSSINOV / RNAFOLD) to the device, reserves its
 * DP tables and captures the queue reset with the persistent fill launch, and
 * the finalize launch, as HIP graphs; dpx_plan_run replays them (synchronous,
 * *device_ms = device time of the replay) and dpx_plan_results copies the
 * answers of the last replay out (same conventions as dpx_zuker_batch).
 * dpx_set_rna_params returns DPX_EBUSY while a plan exists. */
typedef struct dpx_plan dpx_plan;
int dpx_plan_create(dpx_ctx* ctx, int kind, const char* seqs, const int64_t* offs, int count, int min_loop,
                    int want_trace, dpx_plan** out);
int dpx_plan_run(dpx_plan* plan, float* device_ms);
/* Device time of the last replay's fill (queue reset + persistent fill
 * kernel) and finalize launches: the plan replays them as two graphs on its
 * stream, with events recorded before, between and after. */
int dpx_plan_timing(dpx_plan* plan, float* fill_ms, float* finalize_ms);
int dpx_plan_results(dpx_plan* plan, int32_t* results, dpx_trace_t* traces, const int64_t* trace_offs,
                     int32_t* trace_lens);
void dpx_plan_destroy(dpx_plan* plan);

/* Seconds on the host's monotonic clock, the time base of dpx_result. */
double dpx_now(void);

/* ---- launch shape: tuning table and autotuner ------------------------------
 * No reference counterpart (the reference fixes its launch shape when it
 * generates the code, CodeGen.scala:277-278,547).  The tile edge (32) and the
 * waves per tile (4) are fixed by the kernels' data layout and LDS budget;
 * what varies per grammar (request kind DPX_REQ_*) and length bucket is:
 *   wg_per_cu  persistent workgroups per compute unit of the fill launch
 *              (0 = as many as fit, 1..that number)
 *   chunk      Smith-Waterman only: steps a band computes between two
 *              handoffs with its neighbours (16, 32 or 64; 0 = 32).
 * Length bucket b holds n in (32 << b, 64 << b] (b = 0 from n = 0, the last
 * bucket everything above 4096); n is the sequence / chain length, for SW
 * max(n1, n2).  A triangular batch takes the tuning of the bucket holding
 * most of its cells.  Results never depend on the launch shape. */
#define DPX_TUNE_BUCKETS 8
typedef struct { int32_t wg_per_cu, chunk, reserved[2]; } dpx_tuning;
int dpx_tune_bucket(int n);
int dpx_get_tuning(dpx_ctx* ctx, int kind, int bucket, dpx_tuning* out);
int dpx_set_tuning(dpx_ctx* ctx, int kind, int bucket, const dpx_tuning* t);

/* Times every candidate shape of `kind` on seeded synthetic launches of each
 * bucket in [bucket_lo, bucket_hi] (about 4e7 cells each, or one instance;
 * device time of fill + finalize, best of `reps`), checks that all shapes
 * return identical results (else DPX_EHIP), and keeps the fastest in the
 * context's table.  One report entry per candidate (chosen = 1 for the
 * winner); *report_len receives the number of candidates timed. */
typedef struct {
  int32_t kind, bucket, n_lo, n_hi, count;
  dpx_tuning t;
  float ms;
  int32_t chosen;
} dpx_tune_report;
int dpx_autotune(dpx_ctx* ctx, int kind, int bucket_lo, int bucket_hi, int reps, dpx_tune_report* report,
                 int report_cap, int* report_len);

/* ---- serving: continuous batching -----------------------------------------
 * No reference counterpart (the reference solves one problem per loaded
 * library, CodeGen.scala:287-325).  A server owns `slots` streams and one
 * device arena of hbm_budget bytes.  Submitted requests wait in arrival
 * order; whenever a slot is free the oldest waiting request starts a launch,
 * joined by later requests of the same kind, options and neighbouring size
 * class (sizes within a factor of about 4), up to max_batch, while their DP
 * footprint fits the arena.  Launches on different slots run concurrently on
 * the device, so requests arriving while others run are admitted as soon as
 * a slot and memory free up.  Requests whose footprint alone exceeds the
 * budget are refused at submission with DPX_ENOMEM.  Results are identical
 * to the synchronous entry points for the same input.
 * The context must outlive its servers, and dpx_set_rna_params returns
 * DPX_EBUSY while a server exists. */
typedef struct dpx_server dpx_server;

enum { DPX_REQ_NUSSINOV = 1, DPX_REQ_ZUKER = 2, DPX_REQ_SWAT = 3, DPX_REQ_MATCHAIN = 4, DPX_REQ_RNAFOLD = 5 };
enum { DPX_SUBMIT_TRACE = 1, DPX_SUBMIT_TRACE_FROM_BEST = 2 };

typedef struct {
  int64_t hbm_budget;   /* bytes of device memory for the DP matrices in flight (>= 1 MiB) */
  int32_t slots;        /* launches in flight at once, one stream each (1..8) */
  int32_t max_batch;    /* most requests in one triangular launch */
  int32_t max_pending;  /* waiting requests beyond this are refused with DPX_EBUSY */
  int32_t reserved;
} dpx_server_config;

typedef struct {
  int32_t kind;             /* DPX_REQ_* */
  int32_t flags;            /* DPX_SUBMIT_* */
  const char* a; int32_t n1;  /* RNA sequence; SW first sequence; matrix chain: n1 = chain length */
  const char* b; int32_t n2;  /* SW second sequence */
  const int32_t* rows; const int32_t* cols;  /* matrix chain dimensions, n1 each */
  dpx_swat_scores sc;       /* SW scores */
  int32_t min_loop;         /* Nussinov */
} dpx_request;

typedef struct {
  uint64_t id;
  int32_t kind, status;             /* status: DPX_OK or the error of the launch */
  int32_t score;                    /* MFE / Nussinov score / chain cost / SW axiom M(n1,n2) */
  int32_t best, best_i, best_j;     /* SW only (see dpx_swat) */
  int32_t trace_len, batch_size;    /* records kept for dpx_server_trace; requests in the launch */
  int64_t cells;                    /* DP cells of this request */
  int64_t footprint;                /* device bytes admitted for it */
  double t_submit, t_launch, t_done;  /* dpx_now() at submission, launch and completion */
  float device_ms;                  /* device time of the launch that carried it */
  int32_t reserved;
} dpx_result;

typedef struct {
  int64_t submitted, completed, failed, rejected, launches, cells_done;
  int64_t bytes_in_flight, peak_bytes_in_flight, max_batch_seen, pending;
} dpx_server_stats;

int dpx_server_create(dpx_ctx* ctx, const dpx_server_config* cfg, dpx_server** out);
/* Finishes the launches in flight, drops waiting requests, frees the arena. */
void dpx_server_destroy(dpx_server* srv);
/* Copies the request's inputs; *id identifies it in dpx_result. */
int dpx_server_submit(dpx_server* srv, const dpx_request* req, uint64_t* id);
/* Up to max finished requests, waiting up to wait_us for the first one;
 * DPX_EAGAIN when none finished. */
int dpx_server_poll(dpx_server* srv, dpx_result* out, int max, int wait_us, int* got);
/* The backtrace of a polled request that asked for one.  buf == NULL drops
 * it; cap smaller than the trace gives DPX_ETRACE and keeps it. */
int dpx_server_trace(dpx_server* srv, uint64_t id, dpx_trace_t* buf, int cap, int* len);
int dpx_server_get_stats(dpx_server* srv, dpx_server_stats* out);

/* ---- Generic ADP grammars, compiled to HIP at run time ------------------
 * Replaces the reference's per-grammar code generation and JNI library
 * (CodeGen.genTab/genKern/genBT/parseC/backtrackC, CodeGen.scala:100-252,
 * 291-358, 565-568) for grammars other than the built-in ones.  The grammar
 * is a JSON tree of the reference's combinators (BaseParsers.scala,
 * ADPParsers.scala, TTParsers.scala):
 *   {"tracks": 1|2, "alphabet": C type, "alphabet_bytes": n, "answer": C type,
 *    "answer_bytes": n, "header": C declarations, "h": "max"|"min"|"sum"|
 *    "count"|{"maxBy": fn}|{"minBy": fn}, "empty": C literal (optional, the
 *    reference's cudaEmpty), "window": w (single track), "axiom": name,
 *    "functions": {name: {"args": "a,b", "body": "return a+b;"}},
 *    "tabulations": {name: {"always_valid": bool, "p": parser}}}
 *   parser: {"tab": name} | {"term": "empty"|"emptyi"|"el"|"eli"|"seq"|
 *    "empty1"|"empty2"|"el1"|"el2", "min": m, "max": M (seq)} |
 *    {"or": [p, p, ...]} | {"cat": [p, p], "track": 0|1|2 (~, -~, ~-),
 *    "hint": [minl, maxl, minr, maxr] (optional)} | {"map": p, "f": fn} |
 *    {"filter": p, "f": fn} | {"agg": p, "h": h}
 * Function bodies are C/C++ (`_in1[k]`, `_in2[k]` read the inputs, M_W and
 * M_H are the table sizes as in the reference's generated code; argument
 * types are deduced).  With "rna": true (single track, input = base codes
 * A=1 C=2 G=3 U=4 as RNASignature.convert) the bodies may call the librna
 * functions of librna_impl.h (hl_energy, bl_energy, br_energy, il_energy,
 * sr_energy, termau_energy, ext_mismatch_energy, ml_mismatch_energy,
 * ml_energy, ul_energy, bp_index/NO_BP) over the context's parameter file.  Rule ids, split positions, the backtrace order and
 * ties (first strictly better candidate in loop order) are the reference's
 * generated code's; trace records hold up to 2 split positions.  "sum" and
 * "count" (BaseParsers.scala:385-392) add up a cell's candidates (count:
 * their number) and have no backtrace: answer and found flag only, trace
 * length 0 (BaseParsers.scala:150).  An "agg" below a "cat" or "map" is a
 * nested aggregation (max/min/maxBy/minBy): the best candidate of the union
 * of its alternatives, as the Scala parsers aggregate it
 * (BaseParsers.scala:228), under the rule ids of the reference's
 * normalisation (CodeGen.scala:31-36); the winner's splits join the rule's
 * positions.  Nested sum/count are rejected. */
typedef struct dpx_grammar dpx_grammar;
typedef struct {
  int32_t tracks, n_tabs, n_rules, max_pos, axiom_id, window;
  int32_t answer_bytes, alphabet_bytes;
  int32_t loops;       /* 1 if some rule has a split loop (O(n) work per cell) */
  int32_t rna;         /* 1 if the algebra calls the librna energy functions */
  int64_t code_bytes;  /* size of the compiled code object */
} dpx_grammar_info;
/* Analyses, generates and compiles (hipRTC, gfx950) the grammar and loads it
 * on the context's device.  Errors (JSON, unsupported construct, compiler
 * log) are in dpx_last_error(ctx). */
int dpx_grammar_create(dpx_ctx* ctx, const char* spec_json, dpx_grammar** out);
void dpx_grammar_destroy(dpx_grammar* g);
/* Generates and compiles without a device: *src points at the generated HIP
 * source until the next call on this thread; *msg at the error, if any. */
int dpx_grammar_check(const char* spec_json, const char** src, const char** msg);
int dpx_grammar_get_info(const dpx_grammar* g, dpx_grammar_info* out);
const char* dpx_grammar_source(const dpx_grammar* g);
/* One parse: in1[n1] (and in2[n2] for two tracks) of alphabet_bytes each.
 * answer receives answer_bytes; *found = 0 when the axiom has no parse
 * (then no trace).  Records are the reference's (i, j, rule, pos) in BFS
 * order; DPX_ETRACE when trace_cap is too small. */
int dpx_grammar_parse(dpx_ctx* ctx, dpx_grammar* g, const void* in1, int n1, const void* in2, int n2,
                      void* answer, int* found, dpx_trace_t* trace, int trace_cap, int* trace_len);
/* A batch: instance q reads in1 elements [offs1[q], offs1[q+1]) (and in2 by
 * offs2 for two tracks), answers + q * answer_bytes receives its answer,
 * found[q] its flag; with traces != NULL its trace goes to traces +
 * trace_offs[q] (capacity trace_offs[q+1] - trace_offs[q]), length in
 * trace_lens[q].  One launch per diagonal covers every instance still on it
 * (largest first).  No reference counterpart (one instance per call there). */
int dpx_grammar_parse_batch(dpx_ctx* ctx, dpx_grammar* g, const void* in1, const int64_t* offs1, const void* in2,
                            const int64_t* offs2, int count, void* answers, int32_t* found, dpx_trace_t* traces,
                            const int64_t* trace_offs, int32_t* trace_lens);

#ifdef __cplusplus
}
#endif
#endif
