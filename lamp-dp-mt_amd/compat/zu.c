/* The replacement of the reference's generated zu translation unit
 * (bench2/zu.h): my_init/my_solve/my_backtrack/my_free over libdpx.so.
 * Build with the harness's -DSIZE and -I include, link -ldpx. */
#define DPX_COMPAT_ZUKER
#include "dpx_compat.h"
