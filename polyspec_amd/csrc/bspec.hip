// Binned-bispectrum pixel reductions (BSpecBin numerator, bspec_bin.py:339-403).
//
// For triplets t = (i1, i2, i3) of H-map indices,
//   S_t = sum_p [A_i1 B_i2 C_i3 + A_i2 B_i3 C_i1 + A_i3 B_i1 C_i2](p)
// with A, B, C stacks of complex maps (nmap, npix).  The cubic term of the estimator uses
// (A, B, C) = (p1 H of the data, p1 H of the data, m2 H of the data), run with the shared
// stack loaded once; the linear term sums the three assignments in which one of the three
// maps comes from the data and two from a simulation, run as one four-stack pass per
// simulation.  One block owns a range of pixels and a chunk of triplets: the stacks are
// staged tile by tile in LDS, every thread accumulates its triplets over the tile in
// registers and over tiles in FP64, and writes one partial per (block, triplet); a second
// pass adds the partials of all pixel ranges in a fixed order (bit-reproducible).
#include <algorithm>

#include "polyspec_hip.h"
#include "ps_common.h"

namespace ps {

template <typename T> struct bs_tile { static constexpr int TP = 16; };  // pixels per LDS tile
template <> struct bs_tile<double> { static constexpr int TP = 8; };
constexpr int BS_TPMAX = 16;    // largest tile (alignment of the pixel ranges)
constexpr int BS_TPT = 8;       // triplet slots per thread
constexpr int BS_NT = 256;      // threads per block
constexpr int BS_MAXMAP = 128;  // maps per stack

// Stack modes: TRIPLE (A, B, C) three stacks; CUBIC (A = B, C) two stacks, each value
// loaded once; LINEAR (P, Q, M, N) four stacks with the linear-term integrand
//   sum_cyc [(P_a Q_b + Q_a P_b) N_c + Q_a Q_b M_c]
// = the three assignments (P, Q, N), (Q, P, N), (Q, Q, M) of the TRIPLE mode in one pass.
enum BsMode { BS_TRIPLE = 0, BS_CUBIC = 1, BS_LINEAR = 2 };
template <int MODE> struct bs_nstack { static constexpr int N = MODE == BS_TRIPLE ? 3 : MODE == BS_CUBIC ? 2 : 4; };

// two consecutive complex values from a 16-byte-aligned LDS address
__device__ __forceinline__ void bs_ld2(const cplx<float>* p, cplx<float> (&v)[2]) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = cplx<float>{q.x, q.y};
    v[1] = cplx<float>{q.z, q.w};
}
__device__ __forceinline__ void bs_ld2(const cplx<double>* p, cplx<double> (&v)[2]) {
    const double2 q0 = reinterpret_cast<const double2*>(p)[0], q1 = reinterpret_cast<const double2*>(p)[1];
    v[0] = cplx<double>{q0.x, q0.y};
    v[1] = cplx<double>{q1.x, q1.y};
}

template <typename T>
__device__ __forceinline__ cplx<T> cfma(cplx<T> acc, cplx<T> a, cplx<T> b) {
    acc.re += a.re * b.re - a.im * b.im;
    acc.im += a.re * b.im + a.im * b.re;
    return acc;
}

// One block owns a pixel range and a chunk of BS_NT * BS_TPT triplets: the stacks are
// staged tile by tile in LDS as [map][pixel] (a map's TP pixels contiguous, read as wide
// LDS loads), each thread accumulates its triplets over the tile in FP32 (FP64 kernels:
// FP64) and over tiles in FP64, and writes one partial per (block, triplet).
template <typename T, int MODE>
__global__ __launch_bounds__(BS_NT) void k_bspec_sums(const cplx<T>* __restrict__ X0, const cplx<T>* __restrict__ X1,
                                                    const cplx<T>* __restrict__ X2, const cplx<T>* __restrict__ X3,
                                                    int nmap, long npix, long pix_per_block,
                                                    const int* __restrict__ trip, int ntrip,
                                                    cplx<double>* __restrict__ part) {
    constexpr int TP = bs_tile<T>::TP;
    constexpr int NS = bs_nstack<MODE>::N;
    extern __shared__ __align__(16) unsigned char bs_smem[];
    cplx<T>* sm = reinterpret_cast<cplx<T>*>(bs_smem);
    const int SS = TP * nmap;  // one stack's tile
    const int tid = threadIdx.x;
    const int tchunk = blockIdx.y;
    const long p0 = (long)blockIdx.x * pix_per_block, p1 = min(npix, p0 + pix_per_block);
    int idx[BS_TPT][3];
    bool on[BS_TPT];
#pragma unroll
    for (int k = 0; k < BS_TPT; k++) {
        const int t = tchunk * BS_NT * BS_TPT + tid + BS_NT * k;
        on[k] = t < ntrip;
        const int tc = on[k] ? t : 0;
        idx[k][0] = trip[3 * tc];
        idx[k][1] = trip[3 * tc + 1];
        idx[k][2] = trip[3 * tc + 2];
    }
    cplx<double> acc[BS_TPT];
#pragma unroll
    for (int k = 0; k < BS_TPT; k++) acc[k] = cplx<double>{0.0, 0.0};
    const cplx<T>* src[4] = {X0, X1, X2, X3};
    for (long pt = p0; pt < p1; pt += TP) {
        const int np = (int)min((long)TP, p1 - pt);
        __syncthreads();
        // stage: sm[s * SS + map * TP + j] = X_s[map][pt + j]  (coalesced along j)
        for (int e = tid; e < nmap * TP; e += BS_NT) {
            const int mp = e / TP, j = e - mp * TP;
            const bool ok = j < np;
            const size_t g = (size_t)mp * npix + pt + j;
#pragma unroll
            for (int s = 0; s < NS; s++) sm[s * SS + e] = ok ? src[s][g] : cplx<T>{T(0), T(0)};
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < BS_TPT; k++) {
            if (!on[k]) continue;
            const int a = idx[k][0] * TP, b = idx[k][1] * TP, c = idx[k][2] * TP;
            cplx<T> sacc{T(0), T(0)};
            // two pixels per step: every row offset is a multiple of TP (>= 8) complex
            // values, so each pair is one 16-byte-aligned 128-bit LDS load (two for FP64)
#pragma unroll 2
            for (int j = 0; j < TP; j += 2) {
                const cplx<T>* r = sm + j;
                if constexpr (MODE == BS_TRIPLE) {
                    cplx<T> A0[2], A1[2], A2[2], B0[2], B1[2], B2[2], C0[2], C1[2], C2[2];
                    bs_ld2(r + a, A0); bs_ld2(r + b, A1); bs_ld2(r + c, A2);
                    bs_ld2(r + SS + a, B0); bs_ld2(r + SS + b, B1); bs_ld2(r + SS + c, B2);
                    bs_ld2(r + 2 * SS + a, C0); bs_ld2(r + 2 * SS + b, C1); bs_ld2(r + 2 * SS + c, C2);
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        sacc = cfma(sacc, cmul(A0[u], B1[u]), C2[u]);
                        sacc = cfma(sacc, cmul(A1[u], B2[u]), C0[u]);
                        sacc = cfma(sacc, cmul(A2[u], B0[u]), C1[u]);
                    }
                } else if constexpr (MODE == BS_CUBIC) {
                    cplx<T> A0[2], A1[2], A2[2], C0[2], C1[2], C2[2];
                    bs_ld2(r + a, A0); bs_ld2(r + b, A1); bs_ld2(r + c, A2);
                    bs_ld2(r + SS + a, C0); bs_ld2(r + SS + b, C1); bs_ld2(r + SS + c, C2);
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        sacc = cfma(sacc, cmul(A0[u], A1[u]), C2[u]);
                        sacc = cfma(sacc, cmul(A1[u], A2[u]), C0[u]);
                        sacc = cfma(sacc, cmul(A2[u], A0[u]), C1[u]);
                    }
                } else {
                    cplx<T> P0[2], P1[2], P2[2], Q0[2], Q1[2], Q2[2], M0[2], M1[2], M2[2], N0[2], N1[2], N2[2];
                    bs_ld2(r + a, P0); bs_ld2(r + b, P1); bs_ld2(r + c, P2);
                    bs_ld2(r + SS + a, Q0); bs_ld2(r + SS + b, Q1); bs_ld2(r + SS + c, Q2);
                    bs_ld2(r + 2 * SS + a, M0); bs_ld2(r + 2 * SS + b, M1); bs_ld2(r + 2 * SS + c, M2);
                    bs_ld2(r + 3 * SS + a, N0); bs_ld2(r + 3 * SS + b, N1); bs_ld2(r + 3 * SS + c, N2);
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        // cyclic orders (a, b, c), (b, c, a), (c, a, b)
                        sacc = cfma(sacc, cfma(cmul(P0[u], Q1[u]), Q0[u], P1[u]), N2[u]);
                        sacc = cfma(sacc, cmul(Q0[u], Q1[u]), M2[u]);
                        sacc = cfma(sacc, cfma(cmul(P1[u], Q2[u]), Q1[u], P2[u]), N0[u]);
                        sacc = cfma(sacc, cmul(Q1[u], Q2[u]), M0[u]);
                        sacc = cfma(sacc, cfma(cmul(P2[u], Q0[u]), Q2[u], P0[u]), N1[u]);
                        sacc = cfma(sacc, cmul(Q2[u], Q0[u]), M1[u]);
                    }
                }
            }
            acc[k].re += (double)sacc.re;
            acc[k].im += (double)sacc.im;
        }
    }
#pragma unroll
    for (int k = 0; k < BS_TPT; k++) {
        const int t = tchunk * BS_NT * BS_TPT + tid + BS_NT * k;
        if (t < ntrip) part[(size_t)blockIdx.x * ntrip + t] = acc[k];
    }
}

// out[t] (+)= scale * sum_b part[b][t]: 32 lanes per triplet, lane q adds blocks q, q + 32, ...
// in order, then a fixed butterfly over the 32 lanes (bit-reproducible)
__global__ void k_bspec_sum(const cplx<double>* __restrict__ part, int nblk, int ntrip, double scale,
                            int accumulate, cplx<double>* __restrict__ out) {
    const int q = threadIdx.x & 31;
    const long t = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const bool ok = t < ntrip;
    double re = 0.0, im = 0.0;
    if (ok)
        for (int b = q; b < nblk; b += 32) {
            re += part[(size_t)b * ntrip + t].re;
            im += part[(size_t)b * ntrip + t].im;
        }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        re += __shfl_xor(re, o, 32);
        im += __shfl_xor(im, o, 32);
    }
    if (!ok || q != 0) return;
    if (accumulate) {
        out[t].re += scale * re;
        out[t].im += scale * im;
    } else {
        out[t] = cplx<double>{scale * re, scale * im};
    }
}

constexpr int BS_PIX_BLOCKS = 1024;

}  // namespace ps

using namespace ps;

extern "C" {

long long ps_bspec_scratch_bytes(int ntrip) {
    return ntrip < 1 ? -1 : (long long)BS_PIX_BLOCKS * ntrip * (long long)sizeof(cplx<double>);
}

}  // extern "C"

template <typename T, int MODE>
static int bs_launch(const void* X0, const void* X1, const void* X2, const void* X3, int nmap, long npix,
                     const int* trip, int ntrip, cplx<double>* part, long ppb, unsigned nblk, hipStream_t st) {
    constexpr int TP = bs_tile<T>::TP;
    const size_t lds = (size_t)bs_nstack<MODE>::N * TP * nmap * sizeof(cplx<T>);
    if (lds > 65536)
        PS_CHECK_HIP(hipFuncSetAttribute((const void*)k_bspec_sums<T, MODE>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const unsigned nchunk = (unsigned)((ntrip + BS_NT * BS_TPT - 1) / (BS_NT * BS_TPT));
    hipLaunchKernelGGL((k_bspec_sums<T, MODE>), dim3(nblk, nchunk), dim3(BS_NT), lds, st, (const cplx<T>*)X0,
                       (const cplx<T>*)X1, (const cplx<T>*)X2, (const cplx<T>*)X3, nmap, npix, ppb, trip, ntrip, part);
    PS_LAUNCH_CHECK();
    return 0;
}

static int bs_run(int precision, int mode, const void* X0, const void* X1, const void* X2, const void* X3, int nmap,
                  long long npix, const int* triplets, int ntrip, double scale, void* scratch, void* out,
                  int accumulate, hipStream_t st) {
    const long ppb = ((npix + BS_PIX_BLOCKS - 1) / BS_PIX_BLOCKS + BS_TPMAX - 1) / BS_TPMAX * BS_TPMAX;
    const unsigned nblk = (unsigned)((npix + ppb - 1) / ppb);
    cplx<double>* part = (cplx<double>*)scratch;
    int rc;
    if (precision == 32)
        rc = mode == BS_TRIPLE ? bs_launch<float, BS_TRIPLE>(X0, X1, X2, X3, nmap, npix, triplets, ntrip, part, ppb, nblk, st)
           : mode == BS_CUBIC  ? bs_launch<float, BS_CUBIC>(X0, X1, X2, X3, nmap, npix, triplets, ntrip, part, ppb, nblk, st)
                               : bs_launch<float, BS_LINEAR>(X0, X1, X2, X3, nmap, npix, triplets, ntrip, part, ppb, nblk, st);
    else
        rc = mode == BS_TRIPLE ? bs_launch<double, BS_TRIPLE>(X0, X1, X2, X3, nmap, npix, triplets, ntrip, part, ppb, nblk, st)
           : mode == BS_CUBIC  ? bs_launch<double, BS_CUBIC>(X0, X1, X2, X3, nmap, npix, triplets, ntrip, part, ppb, nblk, st)
                               : bs_launch<double, BS_LINEAR>(X0, X1, X2, X3, nmap, npix, triplets, ntrip, part, ppb, nblk, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bspec_sum, dim3((unsigned)(((long)ntrip * 32 + 255) / 256)), dim3(256), 0, st,
                       (const cplx<double>*)part, (int)nblk, ntrip, scale, accumulate, (cplx<double>*)out);
    PS_LAUNCH_CHECK();
    return 0;
}

extern "C" {

int ps_bspec_triple_sums(int precision, const void* A, const void* B, const void* C, int nmap, long long npix,
                         const int* triplets, int ntrip, double scale, void* scratch, long long scratch_bytes,
                         void* out, int accumulate, void* stream) {
    PS_REQUIRE(A && B && C && triplets && scratch && out, "ps_bspec_triple_sums: NULL argument");
    PS_REQUIRE(nmap >= 1 && nmap <= BS_MAXMAP, "ps_bspec_triple_sums: nmap must be in [1, 128]");
    PS_REQUIRE(npix >= 1 && ntrip >= 1, "ps_bspec_triple_sums: empty input");
    PS_REQUIRE(precision == 32 || precision == 64, "ps_bspec_triple_sums: precision must be 32 or 64");
    PS_REQUIRE(scratch_bytes >= ps_bspec_scratch_bytes(ntrip), "ps_bspec_triple_sums: scratch too small");
    // A = B (the cubic term): each value of the shared stack is staged and loaded once
    if (A == B) return bs_run(precision, BS_CUBIC, A, C, nullptr, nullptr, nmap, npix, triplets, ntrip, scale, scratch,
                              out, accumulate, (hipStream_t)stream);
    return bs_run(precision, BS_TRIPLE, A, B, C, nullptr, nmap, npix, triplets, ntrip, scale, scratch, out, accumulate,
                  (hipStream_t)stream);
}

int ps_bspec_linear_sums(int precision, const void* P, const void* Q, const void* M, const void* N, int nmap,
                         long long npix, const int* triplets, int ntrip, double scale, void* scratch,
                         long long scratch_bytes, void* out, int accumulate, void* stream) {
    PS_REQUIRE(P && Q && M && N && triplets && scratch && out, "ps_bspec_linear_sums: NULL argument");
    PS_REQUIRE(nmap >= 1 && nmap <= BS_MAXMAP, "ps_bspec_linear_sums: nmap must be in [1, 128]");
    PS_REQUIRE(npix >= 1 && ntrip >= 1, "ps_bspec_linear_sums: empty input");
    PS_REQUIRE(precision == 32 || precision == 64, "ps_bspec_linear_sums: precision must be 32 or 64");
    PS_REQUIRE(scratch_bytes >= ps_bspec_scratch_bytes(ntrip), "ps_bspec_linear_sums: scratch too small");
    return bs_run(precision, BS_LINEAR, P, Q, M, N, nmap, npix, triplets, ntrip, scale, scratch, out, accumulate,
                  (hipStream_t)stream);
}

}  // extern "C"
