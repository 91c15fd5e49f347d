// Learned CG preconditioner: kernels of a 9-point-stencil convolutional network on
// HEALPix maps (forward, backward, Adam).  FP32 throughout, FP64 for the loss.
//
// A stencil layer  Y[b,p,:] = act(bias + sum_k W_k^T X[b, nbr(p,k), :])  is computed as
// an im2col gather  Xcol[(b,p), (k,c)] = X[b, nbr(p,k), c]  followed by the GEMM
// Y = Xcol W  (W stored (9 cin, cout) row-major).  Backward: dXcol = dY W^T (GEMM)
// followed by a gather through the reverse stencil table (each pixel collects the
// (p, k) slots that read it, so no atomics), and dW = Xcol^T dY (split-K GEMM with a
// fixed-order partial reduction).
#include <math.h>

#include "polyspec_hip.h"
#include "ps_common.h"

namespace ps {

__global__ void k_im2col9(const float* __restrict__ X, const int* __restrict__ nbr, int B, int npix, int cin,
                          float* __restrict__ Xcol) {
    const long n = (long)B * npix * 9 * cin;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int c = (int)(i % cin);
        const long r = i / cin;
        const int k = (int)(r % 9);
        const long bp = r / 9;
        const int p = (int)(bp % npix);
        const long b = bp / npix;
        const int q = nbr[(long)p * 9 + k];
        Xcol[i] = (q >= 0) ? X[(b * npix + q) * cin + c] : 0.f;
    }
}

// dX[b, q, c] = sum_j G[(b, p_j), (k_j, c)] over the reverse entries rev[q, j] = p_j * 9 + k_j
// (j < rw, -1 = unused)
__global__ void k_col2im9(const float* __restrict__ G, const int* __restrict__ rev, int rw, int B, int npix, int cin,
                          float* __restrict__ dX) {
    const long n = (long)B * npix * cin;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int c = (int)(i % cin);
        const long bq = i / cin;
        const int q = (int)(bq % npix);
        const long b = bq / npix;
        float s = 0.f;
        for (int j = 0; j < rw; j++) {
            const int e = rev[(long)q * rw + j];
            if (e < 0) continue;
            const int p = e / 9, k = e % 9;
            s += G[((b * npix + p) * 9 + k) * (long)cin + c];
        }
        dX[i] = s;
    }
}

// C[z] (M x N) = op(A) (M x K-range z) * op(B) (K-range z x N); 64x64x16 tiles staged in LDS, each
// wave one 32 x 32 quarter of the tile on the matrix cores (v_mfma_f32_32x32x2f32: A[i][k] at lane
// i + 32 k, B[k][j] at lane j + 32 k, D[i][j] at lane j + 32 ((i / 4) % 2), register
// 4 (i / 8) + i % 4).  With gridDim.z == 1 the epilogue adds bias[n] and applies ReLU if requested.
typedef float v16f_nn __attribute__((ext_vector_type(16)));

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ A, long lda,
                                              const float* __restrict__ Bm, long ldb, float* __restrict__ C, long ldc,
                                              const float* __restrict__ bias, int relu, int klen) {
    __shared__ float As[16][68];
    __shared__ float Bs[16][68];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, li32 = lane & 31;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // this wave's quarter of the tile
    const int bm = blockIdx.x * 64, bn = blockIdx.y * 64;
    const int k0 = blockIdx.z * klen;
    const int k1 = min(K, k0 + klen);
    v16f_nn acc = v16f_nn{0};
    for (int kk = k0; kk < k1; kk += 16) {
        for (int e = tid; e < 64 * 16; e += 256) {
            // A tile (64 x 16): the contiguous index of the stored matrix varies fastest
            const int mi = TA ? e % 64 : e / 16, ki = TA ? e / 64 : e % 16;
            const int m = bm + mi, k = kk + ki;
            float av = 0.f;
            if (m < M && k < k1) av = TA ? A[(long)k * lda + m] : A[(long)m * lda + k];
            As[ki][mi] = av;
            // B tile (16 x 64)
            const int kb = TB ? e % 16 : e / 64, nb = TB ? e / 16 : e % 64;
            const int k2 = kk + kb, n = bn + nb;
            float bv = 0.f;
            if (n < N && k2 < k1) bv = TB ? Bm[(long)n * ldb + k2] : Bm[(long)k2 * ldb + n];
            Bs[kb][nb] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 16; k += 2)
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(As[k + h][wm + li32], Bs[k + h][wn + li32], acc, 0, 0, 0);
        __syncthreads();
    }
    float* Cz = C + (long)blockIdx.z * M * ldc;
    const int n = bn + wn + li32;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int m = bm + wm + 8 * (v >> 2) + 4 * h + (v & 3);
        if (m >= M || n >= N) continue;
        float val = acc[v];
        if (gridDim.z == 1) {
            if (bias) val += bias[n];
            if (relu) val = fmaxf(val, 0.f);
        }
        Cz[(long)m * ldc + n] = val;
    }
}

// out[i] = sum_z part[z][i] in fixed order
__global__ void k_sum_splits(const float* __restrict__ part, int nz, long n, float* __restrict__ out) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double s = 0.0;
        for (int z = 0; z < nz; z++) s += part[(long)z * n + i];
        out[i] = (float)s;
    }
}

__global__ void k_relu_bwd(long n, const float* __restrict__ Y, float* __restrict__ dY) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (!(Y[i] > 0.f)) dY[i] = 0.f;
}

// db[n] = sum_m dY[m, n] (two stages: column partials over row chunks, then fixed-order sum)
__global__ void k_colsum_part(int M, int N, const float* __restrict__ dY, float* __restrict__ part, int rows) {
    const int n = blockIdx.y * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const int m0 = blockIdx.x * rows, m1 = min(M, m0 + rows);
    double s = 0.0;
    for (int m = m0; m < m1; m++) s += dY[(long)m * N + n];
    part[(long)blockIdx.x * N + n] = (float)s;
}

__global__ void k_adam(long n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// per-sample scaled squared error: loss = sum_b s_b sum_i (Y - T)^2, dY = 2 s_b (Y - T)
__global__ __launch_bounds__(256) void k_mse(long per, int B, const float* __restrict__ Y, const float* __restrict__ Tg,
                                             const double* __restrict__ s, float* __restrict__ dY,
                                             double* __restrict__ part) {
    __shared__ double sh[256];
    const long n = per * B;
    double acc = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double sb = s[i / per];
        const double d = (double)Y[i] - (double)Tg[i];
        acc += sb * d * d;
        if (dY) dY[i] = (float)(2.0 * sb * d);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_d(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Y[b, p, j] = w[p] X[b, p, j]  (in place allowed)
__global__ void k_pixel_scale(long n, int npix, int c, const float* __restrict__ w, const float* X, float* Y) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        Y[i] = w[(i / c) % npix] * X[i];
}

static unsigned grid1d(long n) { return (unsigned)std::min<long>((n + 255) / 256, 16384); }

// ------------------------------------------------ inference on bf16 matrix cores
// A stencil layer for inference, Y[r, n] = act(bias[n] + sum_{k < 9, c < cin} X[b npix + nbr(p, k), c]
// W[k cin + c, n]) (times gate[p] when given), r = b npix + p, on v_mfma_f32_32x32x16_bf16 with
// FP32 accumulation and the stencil gather fused into the operand loads (no im2col buffer).  X is
// bf16 with cin = 16 CPH channels per row (the caller pads with zeros); Y is bf16 (the next
// layer's input) or FP32.  Wp holds the weights in bf16 as (32 NT, 9 cin) row-major.  One wave
// computes 32 rows x 32 NT columns; lane l supplies row (A) or column (B) l % 32 and the 8
// consecutive k values 8 (l / 32) .. + 7 of each 16-wide chunk of k, the same k for both
// operands, so the products pair up whatever order the matrix core takes the 16 k values of a
// chunk in.  A neighbour's whole row segment (CPH chunks) is one batch of loads, issued a
// neighbour ahead of its MFMAs.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int NT, int CPH>
__global__ __launch_bounds__(256) void k_conv9_bf16(const __bf16* __restrict__ X, const int* __restrict__ nbr, int B,
                                                    int npix, const __bf16* __restrict__ Wp,
                                                    const float* __restrict__ bias, int cout, int relu,
                                                    const float* __restrict__ gate, void* __restrict__ Y, int ybf) {
    constexpr int CIN = 16 * CPH, K = 9 * CIN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long M = (long)B * npix;
    const long r0 = ((long)blockIdx.x * 4 + wave) * 32;
    if (r0 >= M) return;  // a wave past the last row (the kernel has no barriers)
    const int li = lane & 31, h = lane >> 5;
    const long r = r0 + li;
    const bool rv = r < M;
    const long bb = rv ? r / npix : 0;
    const int p = rv ? (int)(r - bb * npix) : 0;
    const long rowbase = bb * npix;
    v16f_nn acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[nt] = v16f_nn{0};
    v8bf xa[CPH], xn[CPH], wa[CPH][NT], wn[CPH][NT];
    auto load = [&](int k, v8bf (&x)[CPH], v8bf (&w)[CPH][NT]) {
        const int q = rv ? nbr[(long)p * 9 + k] : -1;
        const v8bf* row = reinterpret_cast<const v8bf*>(X + (rowbase + (q >= 0 ? q : 0)) * (long)CIN);
#pragma unroll
        for (int c = 0; c < CPH; c++) x[c] = q >= 0 ? row[2 * c + h] : v8bf{0};
#pragma unroll
        for (int c = 0; c < CPH; c++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++)
                w[c][nt] = *reinterpret_cast<const v8bf*>(Wp + (long)(li + 32 * nt) * K + k * CIN + 16 * c + 8 * h);
    };
    load(0, xa, wa);
    for (int k = 0; k < 9; k++) {
        if (k + 1 < 9) load(k + 1, xn, wn);
#pragma unroll
        for (int c = 0; c < CPH; c++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++)
                acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa[c], wa[c][nt], acc[nt], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < CPH; c++) {
            xa[c] = xn[c];
#pragma unroll
            for (int nt = 0; nt < NT; nt++) wa[c][nt] = wn[c][nt];
        }
    }
    // accumulator entry i of lane l: row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32 of the tile
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
        const int col = li + 32 * nt;
        if (col >= cout) continue;
        const float bc = bias[col];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const long row = r0 + 8 * (i >> 2) + 4 * h + (i & 3);
            if (row < M) {
                float v = acc[nt][i] + bc;
                if (relu) v = fmaxf(v, 0.f);
                if (gate) v *= gate[row % npix];
                if (ybf)
                    reinterpret_cast<__bf16*>(Y)[row * cout + col] = (__bf16)v;
                else
                    reinterpret_cast<float*>(Y)[row * cout + col] = v;
            }
        }
    }
}

}  // namespace ps

using namespace ps;

extern "C" {

int ps_nn_conv9_bf16(const void* X, const int* nbr, int B, int npix, int cin, const void* Wp, const float* bias,
                     int cout, int relu, const float* gate, void* Y, int y_bf16, void* stream) {
    PS_REQUIRE(X && nbr && Wp && bias && Y, "ps_nn_conv9_bf16: NULL argument");
    PS_REQUIRE(B >= 1 && npix >= 1 && cout >= 1 && cout <= 64, "ps_nn_conv9_bf16: bad sizes (cout <= 64)");
    PS_REQUIRE(cin == 16 || cin == 32 || cin == 64, "ps_nn_conv9_bf16: cin must be 16, 32 or 64 (pad with zeros)");
    PS_REQUIRE((((uintptr_t)Wp | (uintptr_t)X) & 15) == 0, "ps_nn_conv9_bf16: X and Wp must be 16-byte aligned");
    const long M = (long)B * npix;
    const unsigned nblk = (unsigned)((M + 127) / 128);
    const hipStream_t st = (hipStream_t)stream;
    const __bf16* x = (const __bf16*)X;
    const __bf16* w = (const __bf16*)Wp;
#define PS_CONV9(NT, CPH)                                                                                \
    hipLaunchKernelGGL((k_conv9_bf16<NT, CPH>), dim3(nblk), dim3(256), 0, st, x, nbr, B, npix, w, bias, cout, \
                       relu, gate, Y, y_bf16)
    if (cout <= 32) {
        if (cin == 16) PS_CONV9(1, 1);
        else if (cin == 32) PS_CONV9(1, 2);
        else PS_CONV9(1, 4);
    } else {
        if (cin == 16) PS_CONV9(2, 1);
        else if (cin == 32) PS_CONV9(2, 2);
        else PS_CONV9(2, 4);
    }
#undef PS_CONV9
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_nn_im2col9(const float* X, const int* nbr, int B, int npix, int cin, float* Xcol, void* stream) {
    PS_REQUIRE(X && nbr && Xcol && B >= 1 && npix >= 1 && cin >= 1, "ps_nn_im2col9: bad argument");
    hipLaunchKernelGGL(k_im2col9, dim3(grid1d((long)B * npix * 9 * cin)), dim3(256), 0, (hipStream_t)stream, X, nbr,
                       B, npix, cin, Xcol);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_nn_col2im9(const float* G, const int* rev, int rw, int B, int npix, int cin, float* dX, void* stream) {
    PS_REQUIRE(G && rev && dX && rw >= 1 && B >= 1 && npix >= 1 && cin >= 1, "ps_nn_col2im9: bad argument");
    hipLaunchKernelGGL(k_col2im9, dim3(grid1d((long)B * npix * cin)), dim3(256), 0, (hipStream_t)stream, G, rev, rw,
                       B, npix, cin, dX);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_nn_gemm(int transA, int transB, int M, int N, int K, const float* A, long long lda, const float* B,
               long long ldb, float* C, long long ldc, const float* bias, int relu, int ksplit, float* work,
               void* stream) {
    PS_REQUIRE(A && B && C && M >= 1 && N >= 1 && K >= 1, "ps_nn_gemm: bad argument");
    PS_REQUIRE(ksplit >= 1 && (ksplit == 1 || work), "ps_nn_gemm: split-K needs a work buffer of ksplit*M*N floats");
    PS_REQUIRE(ksplit == 1 || (!bias && !relu), "ps_nn_gemm: bias/relu only without split-K");
    hipStream_t st = (hipStream_t)stream;
    const int klen = ((K + ksplit - 1) / ksplit + 15) / 16 * 16;
    const int nz = (K + klen - 1) / klen;
    dim3 grid((M + 63) / 64, (N + 63) / 64, nz);
    float* out = (nz == 1) ? C : work;
    const long ldo = (nz == 1) ? ldc : N;
#define PS_GEMM_LAUNCH(TA_, TB_)                                                                                  \
    hipLaunchKernelGGL((k_gemm<TA_, TB_>), grid, dim3(256), 0, st, M, N, K, A, (long)lda, B, (long)ldb, out, ldo, \
                       bias, relu, klen)
    if (transA && transB) PS_GEMM_LAUNCH(true, true);
    else if (transA) PS_GEMM_LAUNCH(true, false);
    else if (transB) PS_GEMM_LAUNCH(false, true);
    else PS_GEMM_LAUNCH(false, false);
#undef PS_GEMM_LAUNCH
    PS_LAUNCH_CHECK();
    if (nz > 1) {
        PS_REQUIRE(ldc == N, "ps_nn_gemm: split-K output must be dense (ldc == N)");
        hipLaunchKernelGGL(k_sum_splits, dim3(grid1d((long)M * N)), dim3(256), 0, st, (const float*)work, nz,
                           (long)M * N, C);
        PS_LAUNCH_CHECK();
    }
    return 0;
}

int ps_nn_relu_backward(long long n, const float* Y, float* dY, void* stream) {
    PS_REQUIRE(Y && dY && n >= 1, "ps_nn_relu_backward: bad argument");
    hipLaunchKernelGGL(k_relu_bwd, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, (long)n, Y, dY);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_nn_colsum(int M, int N, const float* dY, float* work, float* out, void* stream) {
    PS_REQUIRE(dY && work && out && M >= 1 && N >= 1 && N <= 1024, "ps_nn_colsum: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const int rows = 1024;
    const int nch = (M + rows - 1) / rows;
    hipLaunchKernelGGL(k_colsum_part, dim3(nch, (N + 63) / 64), dim3(64), 0, st, M, N, dY, work, rows);
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_splits, dim3(grid1d(N)), dim3(256), 0, st, (const float*)work, nch, (long)N, out);
    PS_LAUNCH_CHECK();
    return 0;
}

long long ps_nn_colsum_work_floats(int M, int N) { return (long long)((M + 1023) / 1024) * N; }

int ps_nn_adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
               int t, void* stream) {
    PS_REQUIRE(p && g && m && v && n >= 1 && t >= 1, "ps_nn_adam: bad argument");
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, (long)n, p, g, m, v, lr, b1, b2,
                       eps, c1, c2);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_nn_mse(long long per_sample, int B, const float* Y, const float* T, const double* scale, float* dY,
              double* work, double* loss, void* stream) {
    PS_REQUIRE(Y && T && scale && work && loss && per_sample >= 1 && B >= 1, "ps_nn_mse: bad argument");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_mse, dim3(512), dim3(256), 0, st, (long)per_sample, B, Y, T, scale, dY, work);
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(256), 0, st, (const double*)work, 512, loss);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_nn_pixel_scale(int B, int npix, int c, const float* w, const float* X, float* Y, void* stream) {
    PS_REQUIRE(w && X && Y && B >= 1 && npix >= 1 && c >= 1, "ps_nn_pixel_scale: bad argument");
    const long n = (long)B * npix * c;
    hipLaunchKernelGGL(k_pixel_scale, dim3(grid1d(n)), dim3(256), 0, (hipStream_t)stream, n, npix, c, w, X, Y);
    PS_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"
