// Shared host/device helpers for libnepx (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include "../../include/nepx.h"

namespace nepx {

// ---- error reporting (host) ------------------------------------------------
void set_error(const char* fmt, ...);

#define NEPX_HIP_CHECK(expr)                                                      \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::nepx::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                              __FILE__, __LINE__);                                \
            return NEPX_EHIP;                                                     \
        }                                                                         \
    } while (0)

#define NEPX_LAUNCH_CHECK() NEPX_HIP_CHECK(hipGetLastError())

#define NEPX_REQUIRE(cond, code, ...)      \
    do {                                   \
        if (!(cond)) {                     \
            ::nepx::set_error(__VA_ARGS__); \
            return (code);                 \
        }                                  \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---- launch tuning (host) ----------------------------------------------------
nepx_tuning tuning();        // current values with defaults filled in
int device_cus();            // compute units of the current device

// ---- compile-time helpers --------------------------------------------------
constexpr double csqrt(double x) {
    double r = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 64; ++i) r = 0.5 * (r + x / r);
    return r;
}
constexpr double cfact(int n) {
    double r = 1.0;
    for (int k = 2; k <= n; ++k) r *= k;
    return r;
}

// ---- list flags (nepx_nlist.flags) -----------------------------------------------
enum { F_REBUILD = 0, F_OVERFLOW = 1, F_PL = 2, F_P = 3, F_BUILDS = 4, F_MAXROW = 5, F_FROZEN = 6, F_SOON = 7,
       F_LONGSYS = 8, F_TAG = 9, F_RTAG = 10 };

// ---- XCD-contiguous block order ---------------------------------------------
// The dispatcher deals a grid's blocks to the 8 XCDs in turn (block b to XCD b % 8),
// and each XCD has its own L2.  Kernels whose neighbouring blocks share data take
// their logical block from here: XCD x then works through one contiguous range of
// logical blocks (the first nb % 8 XCDs take one block more), so the data that
// neighbouring blocks share stays in one L2.  A bijection of [0, nb) for any nb.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_contiguous_block(int b, int nb) {
    const int per = nb / kXcds, extra = nb % kXcds;
    const int x = b % kXcds, r = b / kXcds;
    return x * per + (x < extra ? x : extra) + r;
}

// ---- per-component force / torque clip -------------------------------------
// v > c -> c, v < -c -> -c, anything else (including NaN) unchanged: the
// masked assignments of ForceEvaluator.py:125-131.  c <= 0 disables the clip.
__device__ __forceinline__ float clipf(float v, float c) {
    if (!(c > 0.f)) return v;
    return v > c ? c : (v < -c ? -c : v);
}

// ---- quaternion helpers (float, device) ------------------------------------
// R(q) v = (w^2-|u|^2) v + 2 w (u x v) + 2 (u.v) u  as a 3x3 matrix, written
// out so that it equals the reference formula (PosToPtsKernels.py:54-71) for
// any q, unit or not.
struct Mat3 {
    float m[9];
};

__device__ __forceinline__ Mat3 quat_matrix(float w, float x, float y, float z) {
    const float a = w * w - (x * x + y * y + z * z);
    Mat3 R;
    R.m[0] = a + 2.f * x * x;
    R.m[1] = 2.f * (x * y - w * z);
    R.m[2] = 2.f * (x * z + w * y);
    R.m[3] = 2.f * (x * y + w * z);
    R.m[4] = a + 2.f * y * y;
    R.m[5] = 2.f * (y * z - w * x);
    R.m[6] = 2.f * (x * z - w * y);
    R.m[7] = 2.f * (y * z + w * x);
    R.m[8] = a + 2.f * z * z;
    return R;
}

__device__ __forceinline__ void mat_vec(const Mat3& R, float bx, float by, float bz, float& ox,
                                        float& oy, float& oz) {
    ox = R.m[0] * bx + R.m[1] * by + R.m[2] * bz;
    oy = R.m[3] * bx + R.m[4] * by + R.m[5] * bz;
    oz = R.m[6] * bx + R.m[7] * by + R.m[8] * bz;
}

}  // namespace nepx
