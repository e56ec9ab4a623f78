// Workgroup tile shapes of the LDS-tiled stencil kernels, as chosen by
// fv3core_amd/autotune.py on one MI355X (six C96 tiles x 79 levels, the
// acoustic step).  A shape changes only how the work is split, never the
// arithmetic, so every choice returns the same bits.  The tuner builds
// variants by defining these macros on the compiler line and rewrites the
// defaults below with the fastest.
#pragma once

// delnflux_kernel: interface rows per block, waves per block
#ifndef FV3_DN_TJ
#define FV3_DN_TJ 8
#endif
#ifndef FV3_DN_NY
#define FV3_DN_NY 4
#endif

// delnflux_kernel, tall shape for FP32 launches: interface rows per block,
// waves per block
#ifndef FV3_DN_TJ_TALL
#define FV3_DN_TJ_TALL 25
#endif
#ifndef FV3_DN_NY_TALL
#define FV3_DN_NY_TALL 4
#endif

// a2b_interp_kernel: corner rows per block, waves per block
#ifndef FV3_AB_TJ
#define FV3_AB_TJ 16
#endif
#ifndef FV3_AB_NY
#define FV3_AB_NY 8
#endif
