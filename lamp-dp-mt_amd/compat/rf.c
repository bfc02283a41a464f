/* The replacement of the reference's generated rf translation unit
 * (RNAFold.scala:20-64): my_init/my_solve/my_backtrack/my_free over libdpx.so.
 * Build with the harness's -DSIZE and -I include, link -ldpx. */
#define DPX_COMPAT_RNAFOLD
#include "dpx_compat.h"
