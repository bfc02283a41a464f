/* The replacement of the reference's generated sw translation unit
 * (bench/dp_swat/comp1.h): my_init/my_solve/my_backtrack/my_free over libdpx.so.
 * Build with the harness's -DSIZE and -I include, link -ldpx. */
#define DPX_COMPAT_SWAT
#include "dpx_compat.h"
