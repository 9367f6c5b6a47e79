// tuned.hpp -- written by ns3dperiodic_amd/autotune.py (select), then NS3D_ZSWZ set from per-pass timings (its note).
// Timed on AMD Radeon Graphics (gfx950): RK3 steps at 256^3 (power-of-two parameters) and 192^3 (NS3D_R3_*); a value replaces the baseline only when its step is faster by more than 0.5%.
#pragma once
// NS3D_EXINV64 = 4 (kept; fp64 step at 256^3 4.3406 ms summed)
// NS3D_EYINV64 = 4 (kept; fp64 step at 256^3 4.3406 ms summed)
// NS3D_EXFWD64 = 8 (kept; fp64 step at 256^3 4.3406 ms summed)
// NS3D_PXINV64 = 8 (kept; fp64 step at 256^3 4.3406 ms summed)
// NS3D_PYINV64 = 8 (kept; fp64 step at 256^3 4.3406 ms summed)
// NS3D_PZT64 = 128 (kept; fp64 step at 256^3 4.3406 ms summed)
// NS3D_ZSWZ = 8 (per-pass A/B/A: FP64 z pass -1.6% / -1.3% / -2.9% at 128^3 / 256^3 / 512^3 against 16, bitwise-identical; below the 0.5% step margin)
// NS3D_ZILP = 1 (kept; fp64 step at 256^3 4.3406 ms summed)
// NS3D_XFWD_ILP = 2 (kept; fp64, fp32, fp32s, fp32i step at 256^3 12.4749 ms summed)
// NS3D_EXINV32 = 8 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_EYINV32 = 4 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_PXINV32 = 8 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_PYINV32 = 8 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_PYFWD32 = 16 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_PXFWD32 = 8 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_PZT32 = 128 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_ZPLAN32 = 8 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_ZSWZ32 = 16 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_EXFWD32 = 4 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_ZILP32 = 0 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_ZWAVE = 1 (kept; fp64, fp32, fp32s, fp32i step at 256^3 12.4749 ms summed)
// NS3D_WAVES_Z32 = 2 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_WAVES_XFWD32 = 2 (kept; fp32, fp32s, fp32i step at 256^3 8.1343 ms summed)
// NS3D_R3_PXINV64 = 4 (kept; fp64 step at 192^3 2.3676 ms summed)
// NS3D_R3_PYINV64 = 4 (kept; fp64 step at 192^3 2.3676 ms summed)
// NS3D_R3_PZT64 = 256 (kept; fp64 step at 192^3 2.3676 ms summed)
// NS3D_R3_ZSWZ = 0 (kept; fp64 step at 192^3 2.3676 ms summed)
// NS3D_R3_PXFWD32 = 16 (kept; fp32, fp32s, fp32i step at 192^3 4.1344 ms summed)
// NS3D_R3_PYINV32 = 4 (kept; fp32, fp32s, fp32i step at 192^3 4.1344 ms summed)
#define NS3D_EXINV64 4
#define NS3D_EYINV64 4
#define NS3D_EXFWD64 8
#define NS3D_PXINV64 8
#define NS3D_PYINV64 8
#define NS3D_PZT64 128
#define NS3D_ZSWZ 8
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
#define NS3D_EXFWD32 4
#define NS3D_ZILP32 0
#define NS3D_ZWAVE 1
#define NS3D_WAVES_Z32 2
#define NS3D_WAVES_XFWD32 2
#define NS3D_R3_PXINV64 4
#define NS3D_R3_PYINV64 4
#define NS3D_R3_PZT64 256
#define NS3D_R3_ZSWZ 0
#define NS3D_R3_PXFWD32 16
#define NS3D_R3_PYINV32 4
