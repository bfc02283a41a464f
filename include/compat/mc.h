/* mc.h -- declarations of the reference's generated per-grammar ABI
 * (Benchmarks.scala:18-27 (MatrixGrammar2)) over libdpx.so: input_t, trace_t, trace_len[] and
 * my_init/my_solve/my_backtrack/my_free.  A harness includes it in place of
 * the generated header; the definitions are lamp-dp-mt_amd/compat/mc.c,
 * built with the same -DSIZE (see dpx_compat.h). */
#ifndef DPX_COMPAT_MATCHAIN_H
#define DPX_COMPAT_MATCHAIN_H
#define DPX_COMPAT_MATCHAIN
#define DPX_COMPAT_DECLS_ONLY
#include "../dpx_compat.h"
#undef DPX_COMPAT_DECLS_ONLY
#endif
