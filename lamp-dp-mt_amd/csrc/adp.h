// adp.h -- generic ADP grammars compiled to HIP at run time (internal).
//
// A grammar arrives as a JSON tree of the reference's parser combinators
// (tabulate, terminals, |, ~, -~, ~-, ^^, filter, aggregate with C bodies
// for the algebra functions).  compile() runs the reference's analysis
// (yield sizes, evaluation order, rule ids), flattens every tabulation into
// its alternatives, generates the cell bodies and the backtrace as HIP
// source and compiles it for gfx950 with hipRTC.  parse() fills the tables
// diagonal by diagonal and walks the backtrace on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/dpx.h"

namespace dpx {
namespace adp {

struct Grammar;

// load == false: generate and compile only (no device needed); on failure
// *code is DPX_EINVAL (spec, analysis, compiler) or DPX_EHIP (loading)
Grammar* compile(const char* spec, bool load, std::string& err, int* code = nullptr);
void destroy(Grammar* g);
const std::string& source(const Grammar* g);
void info(const Grammar* g, dpx_grammar_info* out);
int parse(Grammar* g, hipStream_t s, const void* in1, int n1, const void* in2, int n2, void* answer, int* found,
          dpx_trace_t* tr, int cap, int* len, const void* rna, const int* hl, float* fill_ms, std::string& err);
// count instances: instance q reads in1[off1[q]..off1[q+1]) (and in2 by off2
// for two tracks), elements of alphabet_bytes; answers q * answer_bytes;
// trace of q at tr + tr_off[q] with capacity tr_off[q+1] - tr_off[q];
// rna / hl: the energy tables and hairpin entropies (grammars with "rna")
int parse_batch(Grammar* g, hipStream_t s, int count, const char* in1, const int64_t* off1, const char* in2,
                const int64_t* off2, void* answers, int32_t* found, dpx_trace_t* tr, const int64_t* tr_off,
                int32_t* tr_len, const void* rna, const int* hl, float* fill_ms, std::string& err);

}  // namespace adp
}  // namespace dpx
