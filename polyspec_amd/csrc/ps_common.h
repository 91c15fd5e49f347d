// Shared definitions for the PolySpec MI355X (gfx950) kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

namespace ps {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
const char* last_error();

#define PS_CHECK_HIP(expr)                                                        \
    do {                                                                          \
        hipError_t _e = (expr);                                                   \
        if (_e != hipSuccess) {                                                   \
            ::ps::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
            return -2;                                                            \
        }                                                                         \
    } while (0)

#define PS_REQUIRE(cond, msg)                 \
    do {                                      \
        if (!(cond)) {                        \
            ::ps::set_error(msg);             \
            return -1;                        \
        }                                     \
    } while (0)

#define PS_LAUNCH_CHECK()                                                          \
    do {                                                                           \
        hipError_t _e = hipGetLastError();                                         \
        if (_e != hipSuccess) {                                                    \
            ::ps::set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
            return -2;                                                             \
        }                                                                          \
    } while (0)

// ----------------------------------------------------------------- complex
template <typename T>
struct cplx {
    T re, im;
};

template <typename T> struct vec2;
template <> struct vec2<float> { typedef float2 type; };
template <> struct vec2<double> { typedef double2 type; };

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
    return {a.re, -a.im};
}

__device__ __forceinline__ void sincospi_t(float x, float* s, float* c) { sincospif(x, s, c); }
__device__ __forceinline__ void sincospi_t(double x, double* s, double* c) { sincospi(x, s, c); }

// Scaled-exponent representation of the Legendre recursion values:
// true value = scaled * 2^(64 k); rescale when |scaled| exceeds 2^60.
template <typename T> struct scale_consts;
template <> struct scale_consts<float> {
    static constexpr float DN = 5.42101086242752217e-20f;  // 2^-64
    static constexpr float THR = 1.152921504606846976e18f; // 2^60
    static constexpr int KMIN = -1;  // k below which a value cannot matter in FP32
};
template <> struct scale_consts<double> {
    static constexpr double DN = 5.42101086242752217003726400434970855712890625e-20;
    static constexpr double THR = 1152921504606846976.0;
    static constexpr int KMIN = -2;
};

template <typename T>
__device__ __forceinline__ T scale_factor(int k) {
    if (k == 0) return T(1);
    if (k == -1) return scale_consts<T>::DN;
    if (scale_consts<T>::KMIN <= -2 && k == -2) return scale_consts<T>::DN * scale_consts<T>::DN;
    return T(0);
}

// healpy m-major index lm = m(2 lmax + 1 - m)/2 + l  ->  (l, m).  The m block
// starts at s(m) = m(2 lmax + 3 - m)/2; m is the largest value with s(m) <= lm.
// The first guess comes from an FP32 square root (within a few units of m at any lmax below
// 2^14); the two loops then move it to the exact m.
__device__ __forceinline__ void lm_decode(long lm, int lmax, int* l, int* m) {
    const float B = 2.0f * lmax + 3.0f;
    int mm = (int)floorf((B - sqrtf(fmaxf(B * B - 8.0f * (float)lm, 0.0f))) * 0.5f);
    mm = min(max(mm, 0), lmax);
    while (mm > 0 && (long)mm * (2 * lmax + 3 - mm) / 2 > lm) mm--;
    while (mm < lmax && (long)(mm + 1) * (2 * lmax + 2 - mm) / 2 <= lm) mm++;
    *m = mm;
    *l = (int)(lm - (long)mm * (2 * lmax + 1 - mm) / 2);
}

// Remap a linear block id so that each XCD (8 per MI355X; the dispatcher sends
// block p to XCD p % 8) receives a contiguous range of logical blocks.
__device__ __forceinline__ unsigned xcd_remap(unsigned p, unsigned nblocks) {
    const unsigned nx = 8;
    unsigned per = (nblocks + nx - 1) / nx;
    unsigned x = p % nx, q = p / nx;
    unsigned full = nblocks - (per - 1) * nx;  // XCDs that get 'per' blocks
    unsigned base = (x < full) ? x * per : full * per + (x - full) * (per - 1);
    return base + q;
}

}  // namespace ps
