// tuned.hpp -- written by ns3dperiodic_amd/autotune.py (select); do not edit by hand.
// Timed on AMD Radeon Graphics (gfx950) at N = 256 (RK3 step, dt = 0.0019); a value replaces the baseline only when its step is faster by more than 0.5%.
#pragma once
// NS3D_EXINV64 = 4 (kept; fp64 step 4.4196 ms summed)
// NS3D_EYINV64 = 4 (kept; fp64 step 4.4196 ms summed)
// NS3D_EXFWD64 = 8 (kept; fp64 step 4.4196 ms summed)
// NS3D_PXINV64 = 8 (kept; fp64 step 4.4196 ms summed)
// NS3D_PYINV64 = 8 (kept; fp64 step 4.4196 ms summed)
// NS3D_PZT64 = 128 (kept; fp64 step 4.4196 ms summed)
// NS3D_ZSWZ = 16 (kept; fp64 step 4.4196 ms summed)
// NS3D_ZILP = 1 (kept; fp64 step 4.4196 ms summed)
// NS3D_XFWD_ILP = 2 (kept; fp64, fp32, fp32s step 10.0375 ms summed)
// NS3D_EXINV32 = 8 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_EYINV32 = 4 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_PXINV32 = 8 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_PYINV32 = 8 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_PYFWD32 = 16 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_PXFWD32 = 8 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_PZT32 = 128 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_ZPLAN32 = 8 (kept; fp32, fp32s step 5.6179 ms summed)
// NS3D_ZSWZ32 = 16 (kept; fp32, fp32s step 5.6179 ms summed)
#define NS3D_EXINV64 4
#define NS3D_EYINV64 4
#define NS3D_EXFWD64 8
#define NS3D_PXINV64 8
#define NS3D_PYINV64 8
#define NS3D_PZT64 128
#define NS3D_ZSWZ 16
#define NS3D_ZILP 1
#define NS3D_XFWD_ILP 2
#define NS3D_EXINV32 8
#define NS3D_EYINV32 4
#define NS3D_PXINV32 8
#define NS3D_PYINV32 8
#define NS3D_PYFWD32 16
#define NS3D_PXFWD32 8
#define NS3D_PZT32 128
#define NS3D_ZPLAN32 8
#define NS3D_ZSWZ32 16
