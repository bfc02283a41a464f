/* The replacement of the reference's generated mc translation unit
 * (Benchmarks.scala:18-27 (MatrixGrammar2)): my_init/my_solve/my_backtrack/my_free over libdpx.so.
 * Build with the harness's -DSIZE and -I include, link -ldpx. */
#define DPX_COMPAT_MATCHAIN
#include "dpx_compat.h"
