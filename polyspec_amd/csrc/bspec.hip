// Binned-bispectrum pixel reductions (BSpecBin numerator, bspec_bin.py:339-403).
//
// For triplets t = (i1, i2, i3) of H-map indices,
//   S_t = sum_p [A_i1 B_i2 C_i3 + A_i2 B_i3 C_i1 + A_i3 B_i1 C_i2](p)
// with A, B, C stacks of complex maps (nmap, npix).  The cubic term of the estimator uses
// (A, B, C) = (p1 H of the data, p1 H of the data, m2 H of the data); the linear term sums
// the three assignments in which one of the three maps comes from the data and two from
// a simulation.  One block owns a range of pixels and a chunk of triplets: the three stacks
// are staged tile by tile in LDS (TP pixels x nmap maps each), every thread accumulates
// its triplets over the tile in FP32 registers and over tiles in FP64, and writes one
// partial per (block, triplet); a second pass adds the partials of all pixel ranges in a
// fixed order (bit-reproducible).
#include <algorithm>

#include "polyspec_hip.h"
#include "ps_common.h"

namespace ps {

template <typename T> struct bs_tile { static constexpr int TP = 16; };  // pixels per LDS tile (48 KiB)
template <> struct bs_tile<double> { static constexpr int TP = 8; };
constexpr int BS_TPMAX = 16;    // largest tile (alignment of the pixel ranges)
constexpr int BS_TPT = 8;       // triplets per thread
constexpr int BS_NT = 256;      // threads per block
constexpr int BS_MAXMAP = 128;  // maps per stack held in LDS

template <typename T>
__global__ __launch_bounds__(BS_NT) void k_bspec_triple(const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ B,
                                                      const cplx<T>* __restrict__ C, int nmap, long npix,
                                                      long pix_per_block, const int* __restrict__ trip, int ntrip,
                                                      cplx<double>* __restrict__ part) {
    constexpr int BS_TP = bs_tile<T>::TP;
    __shared__ cplx<T> sA[BS_MAXMAP * BS_TP], sB[BS_MAXMAP * BS_TP], sC[BS_MAXMAP * BS_TP];
    const int tid = threadIdx.x;
    const int tchunk = blockIdx.y;  // triplets tchunk * BS_NT * BS_TPT + tid + BS_NT k
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
    for (long pt = p0; pt < p1; pt += BS_TP) {
        const int np = (int)min((long)BS_TP, p1 - pt);
        __syncthreads();
        // stage tile: sX[map * BS_TP + j] = X[map][pt + j]
        for (int e = tid; e < nmap * BS_TP; e += BS_NT) {
            const int mp = e / BS_TP, j = e - mp * BS_TP;
            const bool ok = j < np;
            const size_t g = (size_t)mp * npix + pt + j;
            sA[e] = ok ? A[g] : cplx<T>{T(0), T(0)};
            sB[e] = ok ? B[g] : cplx<T>{T(0), T(0)};
            sC[e] = ok ? C[g] : cplx<T>{T(0), T(0)};
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < BS_TPT; k++) {
            if (!on[k]) continue;
            const int a = idx[k][0] * BS_TP, b = idx[k][1] * BS_TP, c = idx[k][2] * BS_TP;
            cplx<T> s{T(0), T(0)};
#pragma unroll 4
            for (int j = 0; j < BS_TP; j++) {
                s = cadd(s, cmul(cmul(sA[a + j], sB[b + j]), sC[c + j]));
                s = cadd(s, cmul(cmul(sA[b + j], sB[c + j]), sC[a + j]));
                s = cadd(s, cmul(cmul(sA[c + j], sB[a + j]), sC[b + j]));
            }
            acc[k].re += (double)s.re;
            acc[k].im += (double)s.im;
        }
    }
#pragma unroll
    for (int k = 0; k < BS_TPT; k++) {
        const int t = tchunk * BS_NT * BS_TPT + tid + BS_NT * k;
        if (t < ntrip) part[(size_t)blockIdx.x * ntrip + t] = acc[k];
    }
}

// out[t] (+)= scale * sum_b part[b][t]
__global__ void k_bspec_sum(const cplx<double>* __restrict__ part, int nblk, int ntrip, double scale,
                            int accumulate, cplx<double>* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntrip) return;
    double re = 0.0, im = 0.0;
    for (int b = 0; b < nblk; b++) {
        re += part[(size_t)b * ntrip + t].re;
        im += part[(size_t)b * ntrip + t].im;
    }
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

int ps_bspec_triple_sums(int precision, const void* A, const void* B, const void* C, int nmap, long long npix,
                         const int* triplets, int ntrip, double scale, void* scratch, long long scratch_bytes,
                         void* out, int accumulate, void* stream) {
    PS_REQUIRE(A && B && C && triplets && scratch && out, "ps_bspec_triple_sums: NULL argument");
    PS_REQUIRE(nmap >= 1 && nmap <= BS_MAXMAP, "ps_bspec_triple_sums: nmap must be in [1, 128]");
    PS_REQUIRE(npix >= 1 && ntrip >= 1, "ps_bspec_triple_sums: empty input");
    PS_REQUIRE(scratch_bytes >= ps_bspec_scratch_bytes(ntrip), "ps_bspec_triple_sums: scratch too small");
    hipStream_t st = (hipStream_t)stream;
    const long ppb = ((npix + BS_PIX_BLOCKS - 1) / BS_PIX_BLOCKS + BS_TPMAX - 1) / BS_TPMAX * BS_TPMAX;
    const unsigned nblk = (unsigned)((npix + ppb - 1) / ppb);
    const unsigned nchunk = (unsigned)((ntrip + BS_NT * BS_TPT - 1) / (BS_NT * BS_TPT));
    cplx<double>* part = (cplx<double>*)scratch;
    if (precision == 32)
        hipLaunchKernelGGL(k_bspec_triple<float>, dim3(nblk, nchunk), dim3(BS_NT), 0, st, (const cplx<float>*)A,
                           (const cplx<float>*)B, (const cplx<float>*)C, nmap, (long)npix, ppb, triplets, ntrip, part);
    else if (precision == 64)
        hipLaunchKernelGGL(k_bspec_triple<double>, dim3(nblk, nchunk), dim3(BS_NT), 0, st, (const cplx<double>*)A,
                           (const cplx<double>*)B, (const cplx<double>*)C, nmap, (long)npix, ppb, triplets, ntrip,
                           part);
    else
        PS_REQUIRE(false, "ps_bspec_triple_sums: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bspec_sum, dim3((ntrip + 255) / 256), dim3(256), 0, st, (const cplx<double>*)part, (int)nblk,
                       ntrip, scale, accumulate, (cplx<double>*)out);
    PS_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
