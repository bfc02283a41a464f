/* The replacement of the reference's generated nu translation unit
 * (bench2/nu.h): my_init/my_solve/my_backtrack/my_free over libdpx.so.
 * Build with the harness's -DSIZE and -I include, link -ldpx. */
#define DPX_COMPAT_NUSSINOV
#include "dpx_compat.h"
