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
#include <vector>

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

// One block's work: the pixels [p0, p1) of stacks src against the triplet chunk tchunk of
// BS_NT * TPT triplets.  The stacks are staged tile by tile (TP pixels) in LDS as
// [map][pixel] (a map's TP pixels contiguous, read as wide LDS loads), each thread
// accumulates its TPT triplets over the tile in FP32 (FP64 kernels: FP64) and over tiles in
// FP64, and writes one partial per triplet to part_row.
template <typename T, int MODE, int TP, int TPT>
__device__ __forceinline__ void bs_block(const cplx<T>* const (&src)[4], int nmap, long npix, long p0, long p1,
                                         const int* __restrict__ trip, int ntrip, int tchunk,
                                         cplx<double>* __restrict__ part_row, cplx<T>* sm) {
    constexpr int NS = bs_nstack<MODE>::N;
    const int SS = TP * nmap;  // one stack's tile
    const int tid = threadIdx.x;
    int idx[TPT][3];
    bool on[TPT];
#pragma unroll
    for (int k = 0; k < TPT; k++) {
        const int t = tchunk * BS_NT * TPT + tid + BS_NT * k;
        on[k] = t < ntrip;
        const int tc = on[k] ? t : 0;
        idx[k][0] = trip[3 * tc];
        idx[k][1] = trip[3 * tc + 1];
        idx[k][2] = trip[3 * tc + 2];
    }
    cplx<double> acc[TPT];
#pragma unroll
    for (int k = 0; k < TPT; k++) acc[k] = cplx<double>{0.0, 0.0};
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
        for (int k = 0; k < TPT; k++) {
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
    for (int k = 0; k < TPT; k++) {
        const int t = tchunk * BS_NT * TPT + tid + BS_NT * k;
        if (t < ntrip) part_row[t] = acc[k];
    }
}

// One block per (pixel range blockIdx.x, triplet chunk blockIdx.y) of one set of stacks.
template <typename T, int MODE>
__global__ __launch_bounds__(BS_NT) void k_bspec_sums(const cplx<T>* __restrict__ X0, const cplx<T>* __restrict__ X1,
                                                    const cplx<T>* __restrict__ X2, const cplx<T>* __restrict__ X3,
                                                    int nmap, long npix, long pix_per_block,
                                                    const int* __restrict__ trip, int ntrip,
                                                    cplx<double>* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char bs_smem[];
    const cplx<T>* const src[4] = {X0, X1, X2, X3};
    const long p0 = (long)blockIdx.x * pix_per_block, p1 = min(npix, p0 + pix_per_block);
    bs_block<T, MODE, bs_tile<T>::TP, BS_TPT>(src, nmap, npix, p0, p1, trip, ntrip, blockIdx.y,
                                             part + (size_t)blockIdx.x * ntrip, reinterpret_cast<cplx<T>*>(bs_smem));
}

// ------------------------------------------------------------ multi-segment launch
// A segment is one set of stacks with its triplets and mode (e.g. the cubic term of one
// data set, or the linear term of one simulation, at any resolution); segments that add
// into one output form a group.  One launch covers every segment.  Every segment is cut into
// pixel ranges of one common length, so blocks of different segments carry similar work
// whatever their resolution; the launch visits segments heaviest mode first (a launch-order
// table of segment index and first block, searched by binary search), and blocks go to the
// XCDs round-robin in pixel-range order, so every segment and every triplet chunk is spread
// over the whole chip.
struct BsSegDev {
    const void* X[4];
    const int* trip;
    long long npix, ppb, part0;
    int nmap, ntrip, mode, nchunk;
    unsigned nblk;  // pixel ranges
};
struct BsGroupDev {
    cplx<double>* out;
    double scale;
    long long t0;  // first global triplet index of the group
    int ntrip, accumulate, seg0, nseg;
};
struct BsOrderDev {
    unsigned blk0;  // first block of this launch entry
    int seg;        // segment index (group order)
};

template <typename T, int TP, int TPT>
__global__ __launch_bounds__(BS_NT) void k_bspec_sums_multi(const BsSegDev* __restrict__ segs,
                                                          const BsOrderDev* __restrict__ order, int nseg,
                                                          cplx<double>* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char bs_smem[];
    const unsigned b = blockIdx.x;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (order[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const BsOrderDev O = order[lo];
    const BsSegDev S = segs[O.seg];
    const unsigned local = b - O.blk0;
    // chunk-major: consecutive blocks take consecutive pixel ranges of one triplet chunk, so the
    // round-robin XCD dispatch does not pair a chunk with a fixed set of XCDs
    const int tchunk = (int)(local / S.nblk);
    const long pb = (long)(local - (unsigned)tchunk * S.nblk);
    const long p0 = pb * S.ppb, p1 = min((long)S.npix, p0 + (long)S.ppb);
    const cplx<T>* const src[4] = {(const cplx<T>*)S.X[0], (const cplx<T>*)S.X[1], (const cplx<T>*)S.X[2],
                                   (const cplx<T>*)S.X[3]};
    cplx<double>* row = part + S.part0 + pb * S.ntrip;
    cplx<T>* sm = reinterpret_cast<cplx<T>*>(bs_smem);
    if (S.mode == BS_TRIPLE)
        bs_block<T, BS_TRIPLE, TP, TPT>(src, S.nmap, S.npix, p0, p1, S.trip, S.ntrip, tchunk, row, sm);
    else if (S.mode == BS_CUBIC)
        bs_block<T, BS_CUBIC, TP, TPT>(src, S.nmap, S.npix, p0, p1, S.trip, S.ntrip, tchunk, row, sm);
    else
        bs_block<T, BS_LINEAR, TP, TPT>(src, S.nmap, S.npix, p0, p1, S.trip, S.ntrip, tchunk, row, sm);
}

// out_g[t] (+)= scale_g * sum over the group's segments (in order) of their block partials:
// 32 lanes per (group, triplet), lane q adds blocks q, q + 32, ... of each segment in turn,
// then a fixed butterfly over the 32 lanes (bit-reproducible)
__global__ void k_bspec_sum_multi(const BsSegDev* __restrict__ segs, const BsGroupDev* __restrict__ groups,
                                  int ngroup, long long ttot, const cplx<double>* __restrict__ part) {
    const int q = threadIdx.x & 31;
    const long long T = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
    const bool ok = T < ttot;
    int g = 0;
    if (ok) {
        int hi = ngroup - 1;
        while (g < hi) {
            const int mid = (g + hi + 1) >> 1;
            if (groups[mid].t0 <= T) g = mid;
            else hi = mid - 1;
        }
    }
    double re = 0.0, im = 0.0;
    long long t = 0;
    if (ok) {
        const BsGroupDev G = groups[g];
        t = T - G.t0;
        for (int s = G.seg0; s < G.seg0 + G.nseg; s++) {
            const long long base = segs[s].part0 + t;
            const unsigned nb = segs[s].nblk;
            const int nt = segs[s].ntrip;
            for (unsigned b = q; b < nb; b += 32) {
                re += part[base + (long long)b * nt].re;
                im += part[base + (long long)b * nt].im;
            }
        }
    }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
        re += __shfl_xor(re, o, 32);
        im += __shfl_xor(im, o, 32);
    }
    if (!ok || q != 0) return;
    const BsGroupDev G = groups[g];
    if (G.accumulate) {
        G.out[t].re += G.scale * re;
        G.out[t].im += G.scale * im;
    } else {
        G.out[t] = cplx<double>{G.scale * re, G.scale * im};
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

// ------------------------------------------------------------ multi-segment host side
namespace {

struct BsShape {
    int tp, tpt;    // pixels per LDS tile, triplet slots per thread
    double pscale;  // blocks per launch relative to the built-in target
};
// Launch-shape candidates; id 0 is the built-in shape.
const BsShape BS_SHAPES32[] = {{16, 8, 1.0}, {16, 8, 0.5}, {16, 8, 2.0}, {16, 4, 1.0}, {16, 4, 2.0},
                               {32, 8, 1.0}, {32, 8, 0.5}, {32, 4, 1.0}, {32, 4, 2.0}};
const BsShape BS_SHAPES64[] = {{8, 8, 1.0}, {8, 8, 0.5}, {8, 8, 2.0}, {8, 4, 1.0}, {8, 4, 2.0},
                               {16, 8, 1.0}, {16, 4, 1.0}, {16, 4, 2.0}};
constexpr int BS_NSHAPE32 = sizeof(BS_SHAPES32) / sizeof(BsShape);
constexpr int BS_NSHAPE64 = sizeof(BS_SHAPES64) / sizeof(BsShape);
constexpr long BS_MULTI_ALIGN = 32;      // pixel ranges start at multiples of the largest tile
constexpr long BS_MULTI_TARGET = 8192;   // blocks per launch the built-in rule aims at
constexpr long BS_MULTI_MINPIX = 256;    // shortest pixel range
constexpr size_t BS_MAX_LDS = 163840;    // LDS per workgroup on gfx950

int bs_nshape(int precision) { return precision == 32 ? BS_NSHAPE32 : BS_NSHAPE64; }
const BsShape& bs_shape(int precision, int id) { return precision == 32 ? BS_SHAPES32[id] : BS_SHAPES64[id]; }
int bs_modestacks(int mode) { return mode == BS_TRIPLE ? 3 : mode == BS_CUBIC ? 2 : 4; }
size_t bs_align256(size_t n) { return (n + 255) / 256 * 256; }

struct BsLayout {
    std::vector<BsSegDev> seg;    // ordered by group (stable)
    std::vector<BsGroupDev> grp;
    std::vector<BsOrderDev> ord;  // launch order: heaviest mode first
    unsigned long long nblk = 0;  // blocks of the main launch
    long long part_elems = 0, ttot = 0;
    size_t lds = 0;
};

// Validates the segments and lays them out for one shape.  groups may be NULL (sizing and
// tuning: no reduction is launched).
int bs_layout(int precision, int nseg, const ps_bspec_segment* segs, int ngroup, const ps_bspec_group* groups,
              const BsShape& sh, BsLayout& L) {
    PS_REQUIRE(nseg >= 1 && segs, "ps_bspec_sums_multi: no segments");
    PS_REQUIRE(ngroup >= 1, "ps_bspec_sums_multi: no groups");
    std::vector<int> gtrip(ngroup, -1), gcount(ngroup, 0);
    long long chunks = 0;
    const size_t esz = precision == 32 ? sizeof(cplx<float>) : sizeof(cplx<double>);
    for (int i = 0; i < nseg; i++) {
        const ps_bspec_segment& a = segs[i];
        PS_REQUIRE(a.mode >= 0 && a.mode <= 2, "ps_bspec_sums_multi: segment mode must be 0, 1 or 2");
        PS_REQUIRE(a.nmap >= 1 && a.nmap <= BS_MAXMAP, "ps_bspec_sums_multi: nmap must be in [1, 128]");
        PS_REQUIRE(a.npix >= 1 && a.ntrip >= 1, "ps_bspec_sums_multi: empty segment");
        PS_REQUIRE(a.group >= 0 && a.group < ngroup, "ps_bspec_sums_multi: segment group out of range");
        PS_REQUIRE(a.triplets != nullptr, "ps_bspec_sums_multi: NULL triplets");
        for (int k = 0; k < bs_modestacks(a.mode); k++) PS_REQUIRE(a.X[k] != nullptr, "ps_bspec_sums_multi: NULL stack");
        PS_REQUIRE(gtrip[a.group] < 0 || gtrip[a.group] == a.ntrip,
                   "ps_bspec_sums_multi: segments of one group must have the same number of triplets");
        gtrip[a.group] = a.ntrip;
        gcount[a.group]++;
        chunks += (a.ntrip + BS_NT * sh.tpt - 1) / (BS_NT * sh.tpt);
        L.lds = std::max(L.lds, (size_t)bs_modestacks(a.mode) * sh.tp * a.nmap * esz);
    }
    for (int g = 0; g < ngroup; g++) PS_REQUIRE(gcount[g] > 0, "ps_bspec_sums_multi: a group has no segment");
    PS_REQUIRE(L.lds <= BS_MAX_LDS, "ps_bspec_sums_multi: tile does not fit the LDS (too many maps for this shape)");
    // one pixel-range length for every segment: the pixel-chunk work of the launch split into
    // about BS_MULTI_TARGET / pscale blocks
    long double work = 0;
    for (int i = 0; i < nseg; i++)
        work += (long double)segs[i].npix * ((segs[i].ntrip + BS_NT * sh.tpt - 1) / (BS_NT * sh.tpt));
    const long want = (long)(work / (BS_MULTI_TARGET * sh.pscale)) + 1;
    const long ppb = std::max(BS_MULTI_MINPIX, (want + BS_MULTI_ALIGN - 1) / BS_MULTI_ALIGN * BS_MULTI_ALIGN);
    L.seg.clear();
    L.grp.assign(ngroup, BsGroupDev{});
    long long part = 0, t0 = 0;
    for (int g = 0; g < ngroup; g++) {
        BsGroupDev& G = L.grp[g];
        G.out = groups ? (cplx<double>*)groups[g].out : nullptr;
        G.scale = groups ? groups[g].scale : 1.0;
        G.accumulate = groups ? groups[g].accumulate : 0;
        G.ntrip = gtrip[g];
        G.t0 = t0;
        G.seg0 = (int)L.seg.size();
        G.nseg = gcount[g];
        t0 += gtrip[g];
        for (int i = 0; i < nseg; i++) {
            const ps_bspec_segment& a = segs[i];
            if (a.group != g) continue;
            BsSegDev d;
            for (int k = 0; k < 4; k++) d.X[k] = a.X[k];
            d.trip = a.triplets;
            d.npix = a.npix;
            d.ppb = ppb;
            d.nblk = (unsigned)((a.npix + ppb - 1) / ppb);
            d.nmap = a.nmap;
            d.ntrip = a.ntrip;
            d.mode = a.mode;
            d.nchunk = (a.ntrip + BS_NT * sh.tpt - 1) / (BS_NT * sh.tpt);
            d.part0 = part;
            part += (long long)d.nblk * a.ntrip;
            L.seg.push_back(d);
        }
    }
    // launch order: linear-term segments (4 stacks) first, then triple, then cubic
    std::vector<int> idx(L.seg.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = (int)i;
    auto weight = [&](int i) { return L.seg[i].mode == BS_LINEAR ? 0 : L.seg[i].mode == BS_TRIPLE ? 1 : 2; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return weight(a) < weight(b); });
    unsigned long long blk = 0;
    L.ord.clear();
    for (int i : idx) {
        L.ord.push_back(BsOrderDev{(unsigned)blk, i});
        blk += (unsigned long long)L.seg[i].nblk * L.seg[i].nchunk;
        PS_REQUIRE(blk < (1ULL << 31), "ps_bspec_sums_multi: too many blocks for one launch");
    }
    L.nblk = blk;
    L.part_elems = part;
    L.ttot = t0;
    return 0;
}

size_t bs_table_bytes(int nseg, int ngroup) {
    return bs_align256(nseg * sizeof(BsSegDev)) + bs_align256(ngroup * sizeof(BsGroupDev)) +
           bs_align256(nseg * sizeof(BsOrderDev));
}

template <typename T, int TP, int TPT>
int bs_multi_main_t(const BsLayout& L, const BsSegDev* dseg, const BsOrderDev* dord, cplx<double>* part,
                    hipStream_t st) {
    if (L.lds > 65536)
        PS_CHECK_HIP(hipFuncSetAttribute((const void*)k_bspec_sums_multi<T, TP, TPT>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)L.lds));
    hipLaunchKernelGGL((k_bspec_sums_multi<T, TP, TPT>), dim3((unsigned)L.nblk), dim3(BS_NT), L.lds, st, dseg, dord,
                       (int)L.seg.size(), part);
    PS_LAUNCH_CHECK();
    return 0;
}

int bs_multi_main(int precision, const BsShape& sh, const BsLayout& L, const BsSegDev* dseg, const BsOrderDev* dord,
                  cplx<double>* part, hipStream_t st) {
    if (precision == 32) {
        if (sh.tp == 16) return sh.tpt == 8 ? bs_multi_main_t<float, 16, 8>(L, dseg, dord, part, st)
                                            : bs_multi_main_t<float, 16, 4>(L, dseg, dord, part, st);
        return sh.tpt == 8 ? bs_multi_main_t<float, 32, 8>(L, dseg, dord, part, st)
                           : bs_multi_main_t<float, 32, 4>(L, dseg, dord, part, st);
    }
    if (sh.tp == 8) return sh.tpt == 8 ? bs_multi_main_t<double, 8, 8>(L, dseg, dord, part, st)
                                       : bs_multi_main_t<double, 8, 4>(L, dseg, dord, part, st);
    return sh.tpt == 8 ? bs_multi_main_t<double, 16, 8>(L, dseg, dord, part, st)
                       : bs_multi_main_t<double, 16, 4>(L, dseg, dord, part, st);
}

// Copies the layout's tables to the front of scratch; returns the partials' address.
int bs_upload(const BsLayout& L, void* scratch, long long scratch_bytes, hipStream_t st, BsSegDev** dseg,
              BsGroupDev** dgrp, BsOrderDev** dord, cplx<double>** part) {
    const size_t s0 = bs_align256(L.seg.size() * sizeof(BsSegDev));
    const size_t s1 = s0 + bs_align256(L.grp.size() * sizeof(BsGroupDev));
    const size_t tb = bs_table_bytes((int)L.seg.size(), (int)L.grp.size());
    PS_REQUIRE((long long)(tb + L.part_elems * sizeof(cplx<double>)) <= scratch_bytes,
               "ps_bspec_sums_multi: scratch too small");
    unsigned char* base = (unsigned char*)scratch;
    *dseg = (BsSegDev*)base;
    *dgrp = (BsGroupDev*)(base + s0);
    *dord = (BsOrderDev*)(base + s1);
    *part = (cplx<double>*)(base + tb);
    PS_CHECK_HIP(hipMemcpyAsync(*dseg, L.seg.data(), L.seg.size() * sizeof(BsSegDev), hipMemcpyHostToDevice, st));
    PS_CHECK_HIP(hipMemcpyAsync(*dgrp, L.grp.data(), L.grp.size() * sizeof(BsGroupDev), hipMemcpyHostToDevice, st));
    PS_CHECK_HIP(hipMemcpyAsync(*dord, L.ord.data(), L.ord.size() * sizeof(BsOrderDev), hipMemcpyHostToDevice, st));
    return 0;
}

}  // namespace

extern "C" {

int ps_bspec_multi_shape_count(int precision) {
    if (precision != 32 && precision != 64) {
        set_error("ps_bspec_multi_shape_count: precision must be 32 or 64");
        return -1;
    }
    return bs_nshape(precision);
}

long long ps_bspec_multi_scratch_bytes(int precision, int nseg, const ps_bspec_segment* segs, int ngroup) {
    if (precision != 32 && precision != 64) {
        set_error("ps_bspec_multi_scratch_bytes: precision must be 32 or 64");
        return -1;
    }
    long long most = -1;
    for (int id = 0; id < bs_nshape(precision); id++) {
        BsLayout L;
        if (bs_layout(precision, nseg, segs, ngroup, nullptr, bs_shape(precision, id), L)) {
            if (id == 0) return -1;  // invalid input (the error is set)
            continue;                // this shape's tile does not fit the LDS
        }
        most = std::max(most, L.part_elems);
    }
    return (long long)bs_table_bytes(nseg, ngroup) + most * (long long)sizeof(cplx<double>);
}

int ps_bspec_sums_multi(int precision, int nseg, const ps_bspec_segment* segs, int ngroup,
                        const ps_bspec_group* groups, int shape, void* scratch, long long scratch_bytes,
                        void* stream) {
    PS_REQUIRE(precision == 32 || precision == 64, "ps_bspec_sums_multi: precision must be 32 or 64");
    PS_REQUIRE(groups && scratch, "ps_bspec_sums_multi: NULL argument");
    if (shape < 0) shape = 0;
    PS_REQUIRE(shape < bs_nshape(precision), "ps_bspec_sums_multi: invalid launch shape");
    for (int g = 0; g < ngroup; g++) PS_REQUIRE(groups[g].out != nullptr, "ps_bspec_sums_multi: NULL output");
    const BsShape& sh = bs_shape(precision, shape);
    BsLayout L;
    int rc = bs_layout(precision, nseg, segs, ngroup, groups, sh, L);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    BsSegDev* dseg;
    BsGroupDev* dgrp;
    BsOrderDev* dord;
    cplx<double>* part;
    rc = bs_upload(L, scratch, scratch_bytes, st, &dseg, &dgrp, &dord, &part);
    if (rc) return rc;
    rc = bs_multi_main(precision, sh, L, dseg, dord, part, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bspec_sum_multi, dim3((unsigned)((L.ttot * 32 + 255) / 256)), dim3(256), 0, st, dseg, dgrp,
                       ngroup, L.ttot, part);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_bspec_multi_autotune(int precision, int nseg, const ps_bspec_segment* segs, int ngroup, void* scratch,
                            long long scratch_bytes, int reps, double* ms, int* best, void* stream) {
    PS_REQUIRE(precision == 32 || precision == 64, "ps_bspec_multi_autotune: precision must be 32 or 64");
    PS_REQUIRE(scratch && ms && best && reps >= 1, "ps_bspec_multi_autotune: bad argument");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    PS_CHECK_HIP(hipEventCreate(&e0));
    PS_CHECK_HIP(hipEventCreate(&e1));
    int rc = 0, pick = 0;
    double tbest = 0.0;
    for (int id = 0; id < bs_nshape(precision) && rc == 0; id++) {
        ms[id] = -1.0;
        BsLayout L;
        if (bs_layout(precision, nseg, segs, ngroup, nullptr, bs_shape(precision, id), L)) {
            if (id == 0) rc = -1;
            continue;
        }
        BsSegDev* dseg;
        BsGroupDev* dgrp;
        BsOrderDev* dord;
        cplx<double>* part;
        if ((rc = bs_upload(L, scratch, scratch_bytes, st, &dseg, &dgrp, &dord, &part))) break;
        if ((rc = bs_multi_main(precision, bs_shape(precision, id), L, dseg, dord, part, st))) break;  // untimed
        hipEventRecord(e0, st);
        for (int r = 0; r < reps && rc == 0; r++)
            rc = bs_multi_main(precision, bs_shape(precision, id), L, dseg, dord, part, st);
        hipEventRecord(e1, st);
        if (rc) break;
        if (hipEventSynchronize(e1) != hipSuccess) {
            set_error("ps_bspec_multi_autotune: event synchronisation failed");
            rc = -2;
            break;
        }
        float t = 0.f;
        hipEventElapsedTime(&t, e0, e1);
        ms[id] = t / reps;
        if (id == 0 || ms[id] < tbest) {
            pick = id;
            tbest = ms[id];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // keep the built-in shape unless the fastest is at least 3% faster
    *best = (rc == 0 && ms[0] > 0.0 && tbest < 0.97 * ms[0]) ? pick : 0;
    return rc;
}

}  // extern "C"
