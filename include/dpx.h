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

/* Smith-Waterman with affine gaps, two-track grammar (v4/examples/
 * SWatAffine.scala:22-69; bench/dp_swat/comp1.cu:50-132) with parametric
 * scores: pair = M + (a==b ? match : mismatch); open = max(0, x + gap_open);
 * extend = max(0, x + gap_ext).  Rules: M 0-3, g1 4-6, g2 7-9.
 * *score receives the reference axiom M(n1,n2); *best the local score
 * max_ij M(i,j) at (*best_i,*best_j), the first maximum in row-major order.
 * trace_from: 0 = trace from (n1,n2) like the reference, 1 = from the best
 * cell.  Sequences are letters ACGT/ACGU. */
typedef struct { int32_t match, mismatch, gap_open, gap_ext; } dpx_swat_scores;
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

/* Batch timing: the last batch call's device time (ms) for the fill
 * (persistent tile kernel) and the finalize (st + backtrace) kernels. */
int dpx_last_timing(const dpx_ctx* ctx, float* fill_ms, float* finalize_ms);

#ifdef __cplusplus
}
#endif
#endif
