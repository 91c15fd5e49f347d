// Radial-node pixel and harmonic kernels of the gNL-loc / tauNL-loc estimators
// (the n = 0 loop nests of polyspec/cython/tauNL_utils.pyx and ideal_fisher.pyx).
//
// All of these are HBM-bound streams; the design rules are: one pass over
// each input, coalesced 1-D grid-stride access, FP64 accumulation, and a
// fixed-order two-stage reduction (per-block partials, then one block) so
// that results are bit-reproducible run to run.
#include <math.h>

#include "polyspec_hip.h"
#include <algorithm>
#include <mutex>

#include "ps_common.h"

namespace ps {

static constexpr int RED_BLOCKS = 1024;  // fixed partial count => deterministic order
static constexpr int RED_THREADS = 256;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    // fixed-order tree over 256 threads
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ part, int n, double* out,
                                                      int accumulate) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = accumulate ? (*out + s) : s;
}

// ---------------------------------------------------------- radial weighting
// tauNL_utils.apply_fl_weights (tauNL_utils.pyx:80-98) writing the l-major
// synthesis input: out_t[idx(l,m) * ld + r] = w sum_{X<npol} f[l, X, r] h[X, idx(l,m)]
// for lmin <= l <= lmax, zero otherwise (and zero in padding columns r >= nr).
// f: (lmax_f+1, nslot, nr) real; h: (npol, nlm) complex (full m-major set).
__device__ __forceinline__ int l_of_lm(long lm, int lmax) {
    int l, m;
    lm_decode(lm, lmax, &l, &m);
    return l;
}

template <typename T>
__global__ void k_weight_alm_lmajor(const T* __restrict__ f, int nslot, int nr, const cplx<T>* __restrict__ h,
                                    int npol, int lmax, int lmin, T w, int ld, cplx<T>* __restrict__ out,
                                    int col_offset) {
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlm * nr) return;
    long lm;
    int c;
    if (nlm * nr < (1L << 31)) {  // 32-bit index arithmetic
        const unsigned iu = (unsigned)i;
        lm = (long)(iu / (unsigned)nr);
        c = (int)(iu - (unsigned)lm * (unsigned)nr);
    } else {
        lm = i / nr;
        c = (int)(i % nr);
    }
    const int l = l_of_lm(lm, lmax);
    cplx<T> v{T(0), T(0)};
    if (l >= lmin) {
        const cplx<T> h0 = h[lm];
        const T f0 = f[((long)l * nslot + 0) * nr + c];
        v.re = f0 * h0.re;
        v.im = f0 * h0.im;
        if (npol > 1) {
            const cplx<T> h1 = h[nlm + lm];
            const T f1 = f[((long)l * nslot + 1) * nr + c];
            v.re += f1 * h1.re;
            v.im += f1 * h1.im;
        }
        v.re *= w;
        v.im *= w;
    }
    out[lm * ld + col_offset + c] = v;
}

// The P and Q weightings of one field in one pass: columns col_offset + r take the P table
// fP with weight wP, columns col_offset + nr + r the Q table fQ with weight wQ (both
// (lmax_f+1, nslot, nr)), so each l-major row is written whole by one launch.  Each work item
// is V consecutive columns of one row (V = 4 when nr, ld and col_offset are multiples of 4 and
// the output is 32-byte aligned: one index decode and one pair of h loads per four columns, and
// 16-byte stores); every column is computed as in the one-column form, so both agree bit for bit.
// IL: the columns interleaved, P_r at col_offset + 2r and Q_r at col_offset + 2r + 1.
template <typename T, int V, bool IL>
__global__ void k_weight_alm_pq_lmajor(const T* __restrict__ fP, const T* __restrict__ fQ, int nslot, int nr,
                                       const cplx<T>* __restrict__ h, int npol, int lmax, int lmin, T wP, T wQ,
                                       int ld, cplx<T>* __restrict__ out, int col_offset) {
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const int ng = 2 * nr / V;
    const long tot = nlm * ng;
    // a fixed grid strides over the (lm, column group) items: blocks of single stores would be
    // bound by the workgroup dispatch rate
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long)gridDim.x * blockDim.x) {
        long lm;
        int g;
        if (tot < (1L << 31)) {  // 32-bit index arithmetic
            const unsigned iu = (unsigned)i;
            lm = (long)(iu / (unsigned)ng);
            g = (int)(iu - (unsigned)lm * (unsigned)ng);
        } else {
            lm = i / ng;
            g = (int)(i % ng);
        }
        const int c2 = g * V;
        const int l = l_of_lm(lm, lmax);
        cplx<T> v[V];
#pragma unroll
        for (int k = 0; k < V; k++) v[k] = cplx<T>{T(0), T(0)};
        if (l >= lmin) {
            const cplx<T> h0 = h[lm];
            const cplx<T> h1 = npol > 1 ? h[nlm + lm] : cplx<T>{T(0), T(0)};
            if constexpr (IL) {
                // column c2 + k is P (k even) or Q (k odd) of node (c2 + k) / 2; c2 is even
#pragma unroll
                for (int k = 0; k < V; k++) {
                    const bool isQ = ((c2 + k) & 1) != 0;
                    const T w = isQ ? wQ : wP;
                    const T* f0 = (isQ ? fQ : fP) + (long)l * nslot * nr + ((c2 + k) >> 1);
                    v[k].re = f0[0] * h0.re;
                    v[k].im = f0[0] * h0.im;
                    if (npol > 1) {
                        v[k].re += f0[nr] * h1.re;
                        v[k].im += f0[nr] * h1.im;
                    }
                    v[k].re *= w;
                    v[k].im *= w;
                }
            } else {
                const bool isQ = c2 >= nr;  // a group never straddles P and Q (nr % V == 0)
                const int c = isQ ? c2 - nr : c2;
                const T* f = isQ ? fQ : fP;
                const T w = isQ ? wQ : wP;
                const T* f0 = f + ((long)l * nslot + 0) * nr + c;
                const T* f1 = f0 + nr;
#pragma unroll
                for (int k = 0; k < V; k++) {
                    v[k].re = f0[k] * h0.re;
                    v[k].im = f0[k] * h0.im;
                    if (npol > 1) {
                        v[k].re += f1[k] * h1.re;
                        v[k].im += f1[k] * h1.im;
                    }
                    v[k].re *= w;
                    v[k].im *= w;
                }
            }
        }
        cplx<T>* o = out + lm * ld + col_offset + c2;
        if constexpr (V == 4 && sizeof(T) == 4) {
            float4* o4 = reinterpret_cast<float4*>(o);
            o4[0] = make_float4(v[0].re, v[0].im, v[1].re, v[1].im);
            o4[1] = make_float4(v[2].re, v[2].im, v[3].re, v[3].im);
        } else if constexpr (V == 4) {
            double2* o2 = reinterpret_cast<double2*>(o);
#pragma unroll
            for (int k = 0; k < 4; k++) o2[k] = make_double2(v[k].re, v[k].im);
        } else {
            o[0] = v[0];
        }
    }
}

// ------------------------------------------------------------- gNL-loc sums
// sum_r w_r sum_p [c4 P^3 Q + c2 (P^2 U + P Q V)], P, Q, U, V: (nr, npix).
// c2 = 0 gives tauNL_utils.gnl_loc_sum (tauNL_utils.pyx:260-270) times c4; U = sum_s P_s Q_s and
// V = sum_s P_s^2 give the collapsed disconnected sum of gnl_loc_disc_sum
// (tauNL_utils.pyx:275-285) over the simulations s.
template <typename T>
__global__ __launch_bounds__(256) void k_gnl_fused(const T* __restrict__ P, const T* __restrict__ Q,
                                                   const double* __restrict__ U, const double* __restrict__ V,
                                                   const double* __restrict__ w, int nr, long npix, double c4,
                                                   double c2, double* __restrict__ part) {
    __shared__ double sh[256];
    double acc = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long px = (long)blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += stride) {
        for (int r = 0; r < nr; r++) {
            const long i = (long)r * npix + px;
            const double p = (double)P[i], q = (double)Q[i];
            double t = c4 * p * p * p * q;
            if (U) t += c2 * (p * p * U[i] + p * q * V[i]);
            acc += w[r] * t;
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// The same sum with four consecutive pixels per thread and step (128-bit loads of P and Q,
// two of U and V) and the radial loop unrolled by two, for rows whose starts are aligned
// (npix a multiple of 4, 16-byte aligned P and Q, 32-byte aligned U and V).  The pixel loop of
// every thread visits groups g, g + G, ... (G threads) in order; the last npix % 4 pixels do
// not occur (npix is a multiple of 4).
template <typename T> struct vec4_of;
template <> struct vec4_of<float> { typedef float4 type; };
template <> struct vec4_of<double> { typedef double4 type; };

template <typename T>
__global__ __launch_bounds__(256) void k_gnl_fused4(const T* __restrict__ P, const T* __restrict__ Q,
                                                    const double* __restrict__ U, const double* __restrict__ V,
                                                    const double* __restrict__ w, int nr, long npix, double c4,
                                                    double c2, double* __restrict__ part) {
    typedef typename vec4_of<T>::type V4;
    __shared__ double sh[256];
    double acc = 0.0;
    const long ng = npix >> 2;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += stride) {
#pragma unroll 2
        for (int r = 0; r < nr; r++) {
            const long i = (long)r * ng + g;
            const V4 p4 = reinterpret_cast<const V4*>(P)[i], q4 = reinterpret_cast<const V4*>(Q)[i];
            const double p[4] = {(double)p4.x, (double)p4.y, (double)p4.z, (double)p4.w};
            const double q[4] = {(double)q4.x, (double)q4.y, (double)q4.z, (double)q4.w};
            double t = 0.0;
            if (U) {
                const double4 u4 = reinterpret_cast<const double4*>(U)[i], v4 = reinterpret_cast<const double4*>(V)[i];
                const double u[4] = {u4.x, u4.y, u4.z, u4.w}, v[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
                for (int k = 0; k < 4; k++) t += c4 * p[k] * p[k] * p[k] * q[k] + c2 * (p[k] * p[k] * u[k] + p[k] * q[k] * v[k]);
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) t += c4 * p[k] * p[k] * p[k] * q[k];
            }
            acc += w[r] * t;
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// tauNL_utils.gnl_loc_disc_sum (tauNL_utils.pyx:275-285):
// sum_r w_r sum_p (P1^2 P2 Q2 + P2^2 P1 Q1)
template <typename T>
__global__ __launch_bounds__(256) void k_gnl_disc(const T* __restrict__ P1, const T* __restrict__ P2,
                                                  const T* __restrict__ Q1, const T* __restrict__ Q2,
                                                  const double* __restrict__ w, int nr, long npix,
                                                  double* __restrict__ part) {
    __shared__ double sh[256];
    double acc = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long px = (long)blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += stride) {
        for (int r = 0; r < nr; r++) {
            const long i = (long)r * npix + px;
            const double p1 = (double)P1[i], p2 = (double)P2[i];
            acc += w[r] * (p1 * p1 * p2 * (double)Q2[i] + p2 * p2 * p1 * (double)Q1[i]);
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// U += sum over the batch of P_s Q_s, V += sum of P_s^2; P, Q: (nsim, n), U, V: (n)
template <typename T>
__global__ void k_mc_accumulate(const T* __restrict__ P, const T* __restrict__ Q, int nsim, long n,
                                double* __restrict__ U, double* __restrict__ V) {
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double u = U[i], v = V[i];
        for (int s = 0; s < nsim; s++) {
            const double p = (double)P[(long)s * n + i], q = (double)Q[(long)s * n + i];
            u += p * q;
            v += p * p;
        }
        U[i] = u;
        V[i] = v;
    }
}

// ------------------------------------------------------- pixel products
// mode 0: out = a*b                   (ideal_fisher.multiplyRR, ideal_fisher.pyx:100-108)
// mode 1: out = a*b + c*d             (multiplyRR_sym, ideal_fisher.pyx:112-120)
// mode 2: out = a*b*c                 (tauNL_utils.multiply, tauNL_utils.pyx:103-111)
// mode 3: out = a*(2 b c + a d)       (tauNL_utils.multiply_asym, tauNL_utils.pyx:116-124)
// mode 4: out = a*b + c*d with c batched like a (multiplyRR_sym of each simulation's maps
//         (a, c) with the shared data maps (b, d): Q_s P + P_s Q, tspec_template.py:1294-1303)
template <typename T>
__global__ void k_products(int mode, const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
                           const T* __restrict__ d, long n, T scale, T* __restrict__ out) {
    // blockIdx.y: map stack of a batch (a and out n elements apart; b, c, d shared)
    a += (long)blockIdx.y * n;
    out += (long)blockIdx.y * n;
    if (mode == 4) c += (long)blockIdx.y * n;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        T v;
        switch (mode) {
            case 0: v = a[i] * b[i]; break;
            case 1:
            case 4: v = a[i] * b[i] + c[i] * d[i]; break;
            case 2: v = a[i] * b[i] * c[i]; break;
            default: v = a[i] * (T(2) * b[i] * c[i] + a[i] * d[i]); break;
        }
        out[i] = v * scale;
    }
}

// ------------------------------------------------------------- radial sums
// tauNL_utils.radial_sum (tauNL_utils.pyx:220-235) over the full m-major set with
// l < lmin zeroed: out[X, lm] = scale sum_r lm_map[r, lm] f[l, X, r] w_r (X < npol),
// lm_map: (nr, nlm); f: (lmax+1, nslot, nr).  accumulate != 0 adds into out.
template <typename T>
__global__ void k_radial_sum(const cplx<T>* __restrict__ lmmap, int nr, int lmax, int lmin,
                             const T* __restrict__ f, int nslot, int npol, const double* __restrict__ w,
                             double scale, int accumulate, cplx<T>* __restrict__ out) {
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const long lm = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lm >= nlm) return;
    const int l = l_of_lm(lm, lmax);
    for (int X = 0; X < npol; X++) {
        double re = 0.0, im = 0.0;
        if (l >= lmin) {
            for (int r = 0; r < nr; r++) {
                const cplx<T> v = lmmap[(long)r * nlm + lm];
                const double fw = (double)f[((long)l * nslot + X) * nr + r] * w[r];
                re += (double)v.re * fw;
                im += (double)v.im * fw;
            }
        }
        cplx<T>* o = out + (long)X * nlm + lm;
        if (accumulate) {
            o->re += (T)(re * scale);
            o->im += (T)(im * scale);
        } else {
            *o = cplx<T>{(T)(re * scale), (T)(im * scale)};
        }
    }
}

// -------------------------------------------------------------- tau_sum_n0
// tauNL_utils.tau_sum_n0 (tauNL_utils.pyx:604-646) restated as per-L Gram
// matrices: tau = sum_L sum_{j,k} F_L(j,k) w_j w_k / (sqrt(4pi) sqrt(2n1+1))
//                 sum_{M} (1+[M>0]) Re(A_j(L,M) conj(B_k(L,M))).
// A, B: (nmu, nr, nlm(Lmaxs)) complex in m-major order over 0..Lmaxs.
// One block per L; each thread owns Gram entries (j, k).
template <typename T>
__global__ __launch_bounds__(256) void k_tau_gram(const cplx<T>* __restrict__ A, const cplx<T>* __restrict__ Bm,
                                                  int nmu, int nr, int Lmaxs, int Lmin, int Lmax,
                                                  const double* __restrict__ FLL, const double* __restrict__ w,
                                                  double pref, double* __restrict__ part) {
    __shared__ double sh[256];
    const int L = Lmin + blockIdx.x;
    double acc = 0.0;
    if (L <= Lmax) {
        const long nlm = (long)(Lmaxs + 1) * (Lmaxs + 2) / 2;
        for (int jk = threadIdx.x; jk < nr * nr; jk += blockDim.x) {
            const int j = jk / nr, k = jk % nr;
            if (w[j] == 0.0 || w[k] == 0.0) continue;
            double g = 0.0;
            for (int mu = 0; mu < nmu; mu++) {
                const cplx<T>* a = A + ((long)mu * nr + j) * nlm;
                const cplx<T>* b = Bm + ((long)mu * nr + k) * nlm;
                for (int M = 0; M <= L; M++) {
                    const long id = (long)M * (2 * Lmaxs + 1 - M) / 2 + L;
                    const double re = (double)a[id].re * (double)b[id].re + (double)a[id].im * (double)b[id].im;
                    g += (M == 0 ? 1.0 : 2.0) * re;
                }
            }
            acc += FLL[((long)L * nr + j) * nr + k] * w[j] * w[k] * g;
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc * pref;
}

// The same sum on FP64 matrix cores, with A and B l-major and batched over simulations:
// At[idx(L,M) * ld + mu * nr + j] (idx the m-major index over 0..Lmaxs; mu < nmu, j < nr).
//   tau = pref sum_L sum_{j,k} F_L(j,k) w_j w_k G_L(j,k),
//   G_L(j,k) = sum_mu sum_M (1+[M>0]) Re(A_mu,j(L,M) conj(B_mu,k(L,M))),
// i.e. the sum over mu of tau_sum_n0 of each (A_mu, B_mu) pair.  G_L = X^T Y is a real GEMM
// with K = (M, mu, re/im): v_mfma_f64_16x16x4f64 takes two (M, mu) entries per instruction
// (lanes 16q + c: q = 0,1 the re, im parts of the first entry, q = 2,3 of the second; c the
// column within a 16-wide tile of j or k).  Lane layout (verified on gfx950 for the Fisher
// derivative GEMM): A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k, D[i][j] at lane
// j + 16 (i % 4), register i / 4.  One block per L, heaviest (largest L) first; the four waves
// take alternate entry pairs and contract their partial G_L with F_L in the epilogue (the
// contraction is linear in G), so no Gram matrix is stored.  NT = ceil(nr / 16) tiles per side.
// Optional shift: Mt (l-major, [idx(L,M) * nr + j]) is subtracted from every A_mu (shiftA) and/or
// B_mu (shiftB) column block as the values are widened to FP64, i.e. the sum is taken over
// (A_mu - M, B_mu - M).
typedef double ps_dbl4 __attribute__((ext_vector_type(4)));

template <typename T, int NT, bool SAME>
__global__ __launch_bounds__(256) void k_tau_gram_mfma(const cplx<T>* __restrict__ At, const cplx<T>* __restrict__ Bt,
                                                       long ld, int nmu, int nr, int Lmaxs, int Lmax,
                                                       const double* __restrict__ FLL, const double* __restrict__ w,
                                                       const cplx<T>* __restrict__ Mt, int shiftA, int shiftB,
                                                       double pref, double* __restrict__ part) {
    __shared__ double sh[256];
    const int L = Lmax - (int)blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, q = lane >> 4, half = q >> 1;
    const bool im = (q & 1) != 0;
    ps_dbl4 acc[NT][NT];
#pragma unroll
    for (int ti = 0; ti < NT; ti++)
#pragma unroll
        for (int tk = 0; tk < NT; tk++) acc[ti][tk] = ps_dbl4{0.0, 0.0, 0.0, 0.0};
    const long nent = (long)(L + 1) * nmu;  // entries (M, mu), mu fastest
    const long rowL = L;                     // idx(L, M) = M (2 Lmaxs + 1 - M) / 2 + L
    auto fetch = [&](long e, double* a, double* b) {
        if (e < nent) {
            const int M = (int)(e / nmu), mu = (int)(e - (long)M * nmu);
            const long row = (long)M * (2 * Lmaxs + 1 - M) / 2 + rowL;
            const cplx<T>* ra = At + row * ld + (long)mu * nr;
            const cplx<T>* rb = Bt + row * ld + (long)mu * nr;
            const cplx<T>* rm = Mt ? Mt + row * nr : nullptr;
            const double wm = M == 0 ? 1.0 : 2.0;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int j = 16 * t + c;
                double va = 0.0, vb = 0.0;
                if (j < nr) {
                    double sh = 0.0;
                    if (rm) {
                        const cplx<T> z = rm[j];
                        sh = (double)(im ? z.im : z.re);
                    }
                    const cplx<T> x = ra[j];
                    va = (double)(im ? x.im : x.re) - (shiftA ? sh : 0.0);
                    if constexpr (SAME) {
                        vb = va;
                    } else {
                        const cplx<T> y = rb[j];
                        vb = (double)(im ? y.im : y.re) - (shiftB ? sh : 0.0);
                    }
                }
                a[t] = va;
                b[t] = wm * vb;
            }
        } else {
#pragma unroll
            for (int t = 0; t < NT; t++) a[t] = b[t] = 0.0;
        }
    };
    double a0[NT], b0[NT], a1[NT], b1[NT];
    long e = 2 * wave + half;
    fetch(e, a0, b0);
    for (; e - half < nent; e += 8) {
        fetch(e + 8, a1, b1);  // next pair of entries in flight while this one multiplies
#pragma unroll
        for (int ti = 0; ti < NT; ti++)
#pragma unroll
            for (int tk = 0; tk < NT; tk++)
                acc[ti][tk] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[ti], b0[tk], acc[ti][tk], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            a0[t] = a1[t];
            b0[t] = b1[t];
        }
    }
    // epilogue: register v of tile (ti, tk) holds G_L(16 ti + 4 v + q, 16 tk + c)
    double sum = 0.0;
#pragma unroll
    for (int ti = 0; ti < NT; ti++)
#pragma unroll
        for (int tk = 0; tk < NT; tk++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = 16 * ti + 4 * v + q, k = 16 * tk + c;
                if (j < nr && k < nr) sum += FLL[((long)L * nr + j) * nr + k] * w[j] * w[k] * acc[ti][tk][v];
            }
    sum = block_sum(sum, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = sum * pref;
}

// Per-node form of k_radial_sum (ideal_fisher.integrate_pq_deriv, ideal_fisher.pyx:395-407, the
// weight derivative of the tauNL ideal Fisher): out[r, X, lm] (+)= scale w_r f[l, X, r] lm_map[r, lm]
// for l >= lmin, else 0.  out: (nr, npol, nlm).  grid (ceil(nlm / 256), nr).
template <typename T>
__global__ void k_radial_scatter(const cplx<T>* __restrict__ lmmap, int nr, int lmax, int lmin,
                                 const T* __restrict__ f, int nslot, int npol, const double* __restrict__ w,
                                 double scale, int accumulate, cplx<T>* __restrict__ out) {
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const long lm = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (lm >= nlm) return;
    const int l = l_of_lm(lm, lmax);
    const cplx<T> v = lmmap[(long)r * nlm + lm];
    for (int X = 0; X < npol; X++) {
        const double fw = l >= lmin ? (double)f[((long)l * nslot + X) * nr + r] * w[r] * scale : 0.0;
        cplx<T>* o = out + ((long)r * npol + X) * nlm + lm;
        const double re = (double)v.re * fw, im = (double)v.im * fw;
        if (accumulate) {
            o->re += (T)re;
            o->im += (T)im;
        } else {
            *o = cplx<T>{(T)re, (T)im};
        }
    }
}

// ---------------------------------------------------------- outer product
// ideal_fisher.outer_product (ideal_fisher.pyx:423-459), sym = False:
// fish[a,b] = Re sum_j [Q0a* Q0b + Q1a* Q1b + 9(Q2a* Q2b + Q3a* Q3b)
//                       - 3(Q0a* Q3b + Q1a* Q2b + Q3a* Q0b + Q2a* Q1b)] / (24 * 48)
// Qa: (4, nta, nin), Qb: (4, ntb, nin).  grid (RED_BLOCKS, nta*ntb) -> partials.
template <typename T>
__global__ __launch_bounds__(256) void k_outer(const cplx<T>* __restrict__ Qa, const cplx<T>* __restrict__ Qb,
                                               int nta, int ntb, long nin, double* __restrict__ part) {
    __shared__ double sh[256];
    const int ab = blockIdx.y, ia = ab / ntb, ib = ab % ntb;
    double acc = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < nin; j += stride) {
        cplx<T> a[4], b[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            a[c] = Qa[((long)c * nta + ia) * nin + j];
            b[c] = Qb[((long)c * ntb + ib) * nin + j];
        }
        auto rdot = [](cplx<T> x, cplx<T> y) { return (double)x.re * (double)y.re + (double)x.im * (double)y.im; };
        acc += rdot(a[0], b[0]) + rdot(a[1], b[1]) + 9.0 * (rdot(a[2], b[2]) + rdot(a[3], b[3])) -
               3.0 * (rdot(a[0], b[3]) + rdot(a[1], b[2]) + rdot(a[3], b[0]) + rdot(a[2], b[1]));
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(long)ab * gridDim.x + blockIdx.x] = acc / (24.0 * 48.0);
}

__global__ void k_sum_rows(const double* __restrict__ part, int nrow, int ncolp, double* __restrict__ out,
                           int accumulate) {
    __shared__ double sh[256];
    const int row = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < ncolp; i += 256) s += part[(long)row * ncolp + i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[row] = accumulate ? out[row] + s : s;
}


// ------------------------------------------------------ F_L convolution
// tauNL_utils.convolve_F_n000 (tauNL_utils.pyx:966-992, inds = all nodes) feeding
// TSpecTemplate._convolve_F (tspec_template.py:2155-2162), written as the
// l-major synthesis input of the following to_map_vec:
// out_t[lm * ld + j] = sum_{k: w_k != 0} F_L(j, k) w_k PQ[k, lm] / sqrt(4 pi)
// for Lmin <= L <= Lmax, else 0.  PQ: (nr, nlm(Lmaxs)) complex; FLL (Lmaxs+1, nr, nr).
template <typename T>
__global__ void k_convolve_F(const cplx<T>* __restrict__ PQ, int nr, int Lmaxs, int Lmin, int Lmax,
                             const double* __restrict__ FLL, const double* __restrict__ w, int ld,
                             cplx<T>* __restrict__ out) {
    const long nlm = (long)(Lmaxs + 1) * (Lmaxs + 2) / 2;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlm * nr) return;
    const long lm = i / nr;
    const int j = (int)(i % nr);
    const int L = l_of_lm(lm, Lmaxs);
    double re = 0.0, im = 0.0;
    if (L >= Lmin && L <= Lmax) {
        const double* F = FLL + ((long)L * nr + j) * nr;
        for (int k = 0; k < nr; k++) {
            if (w[k] == 0.0) continue;
            const double fw = F[k] * w[k];
            const cplx<T> v = PQ[(long)k * nlm + lm];
            re += fw * (double)v.re;
            im += fw * (double)v.im;
        }
    }
    const double c = 0.28209479177387814;  // 1/sqrt(4 pi)
    out[lm * ld + j] = cplx<T>{(T)(re * c), (T)(im * c)};
}

}  // namespace ps

using namespace ps;

// =============================================================== C ABI
template <typename T>
static int weight_alm_impl(const void* f, int nslot, int nr, const void* h, int npol, int lmax, int lmin, double w,
                           int ld, void* out, int col_offset, void* stream) {
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const long tot = nlm * nr;
    hipLaunchKernelGGL(k_weight_alm_lmajor<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, (const T*)f, nslot, nr, (const cplx<T>*)h, npol, lmax, lmin, (T)w, ld,
                       (cplx<T>*)out, col_offset);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T, bool IL>
static int weight_alm_pq_impl(const void* fP, const void* fQ, int nslot, int nr, const void* h, int npol, int lmax,
                              int lmin, double wP, double wQ, int ld, void* out, int col_offset, void* stream) {
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    // four-column groups: within P or within Q (nr % 4 == 0), or two interleaved (P, Q) pairs
    const bool vec = (IL ? nr % 2 == 0 : nr % 4 == 0) && ld % 4 == 0 && col_offset % 4 == 0 &&
                     (uintptr_t)out % 32 == 0;
    const long tot = nlm * 2 * nr / (vec ? 4 : 1);
    const long nblk = std::min<long>((tot + 255) / 256, 8192);
    if (vec)
        hipLaunchKernelGGL((k_weight_alm_pq_lmajor<T, 4, IL>), dim3((unsigned)nblk), dim3(256), 0,
                           (hipStream_t)stream, (const T*)fP, (const T*)fQ, nslot, nr, (const cplx<T>*)h, npol, lmax,
                           lmin, (T)wP, (T)wQ, ld, (cplx<T>*)out, col_offset);
    else
        hipLaunchKernelGGL((k_weight_alm_pq_lmajor<T, 1, IL>), dim3((unsigned)nblk), dim3(256), 0,
                           (hipStream_t)stream, (const T*)fP, (const T*)fQ, nslot, nr, (const cplx<T>*)h, npol, lmax,
                           lmin, (T)wP, (T)wQ, ld, (cplx<T>*)out, col_offset);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int reduce_launch_fused(const void* P, const void* Q, const void* U, const void* V, const double* w, int nr,
                               long long npix, double c4, double c2, double* part, double* out, int accumulate,
                               hipStream_t st) {
    const bool vec = npix % 4 == 0 && ((uintptr_t)P | (uintptr_t)Q) % (4 * sizeof(T)) == 0 &&
                     (!U || (((uintptr_t)U | (uintptr_t)V) % 32 == 0));
    if (vec)
        hipLaunchKernelGGL(k_gnl_fused4<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, (const T*)P, (const T*)Q,
                           (const double*)U, (const double*)V, w, nr, (long)npix, c4, c2, part);
    else
        hipLaunchKernelGGL(k_gnl_fused<T>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, (const T*)P, (const T*)Q,
                           (const double*)U, (const double*)V, w, nr, (long)npix, c4, c2, part);
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (const double*)part, RED_BLOCKS, out, accumulate);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T, int NT>
static void launch_tau_mfma(const void* At, const void* Bt, long ld, int nmu, int nr, int Lmaxs, int Lmin, int Lmax,
                            const double* FLL, const double* w, const void* Mt, int shA, int shB, double pref,
                            double* scratch, hipStream_t st) {
    const dim3 grid((unsigned)(Lmax - Lmin + 1));
    if (At == Bt && shA == shB)
        hipLaunchKernelGGL((k_tau_gram_mfma<T, NT, true>), grid, dim3(256), 0, st, (const cplx<T>*)At,
                           (const cplx<T>*)Bt, ld, nmu, nr, Lmaxs, Lmax, FLL, w, (const cplx<T>*)Mt, shA, shB, pref,
                           scratch);
    else
        hipLaunchKernelGGL((k_tau_gram_mfma<T, NT, false>), grid, dim3(256), 0, st, (const cplx<T>*)At,
                           (const cplx<T>*)Bt, ld, nmu, nr, Lmaxs, Lmax, FLL, w, (const cplx<T>*)Mt, shA, shB, pref,
                           scratch);
}

template <typename T>
static void launch_tau_mfma_nt(int nt, const void* At, const void* Bt, long ld, int nmu, int nr, int Lmaxs, int Lmin,
                               int Lmax, const double* FLL, const double* w, const void* Mt, int shA, int shB,
                               double pref, double* scratch, hipStream_t st) {
    switch (nt) {
        case 1: launch_tau_mfma<T, 1>(At, Bt, ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, Mt, shA, shB, pref, scratch, st); break;
        case 2: launch_tau_mfma<T, 2>(At, Bt, ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, Mt, shA, shB, pref, scratch, st); break;
        case 3: launch_tau_mfma<T, 3>(At, Bt, ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, Mt, shA, shB, pref, scratch, st); break;
        default: launch_tau_mfma<T, 4>(At, Bt, ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, Mt, shA, shB, pref, scratch, st); break;
    }
}

// ------------------------------------------------- allocation-free tau sum
// Partial sums (one per L) of ps_tau_sum_lmajor_noalloc live in this file-scope device array
// instead of a caller workspace.  Host code never passes the symbol itself: its device address,
// which differs per GPU, comes from hipGetSymbolAddress on the first call on each device (that
// device current) and is cached per device.  Every call on a GPU shares that GPU's copy, so
// calls that overlap on one GPU (concurrent streams or threads) race.
static constexpr int TAU_NOALLOC_MAX_L = 8192;  // 64 KiB of partials per device
static __device__ __align__(256) double scratch[TAU_NOALLOC_MAX_L];
static constexpr int TAU_NOALLOC_MAX_DEVICES = 64;
static double* g_scratch_ptr[TAU_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int tau_noalloc_scratch(double** out) {
    int dev = -1;
    PS_CHECK_HIP(hipGetDevice(&dev));
    PS_REQUIRE(dev >= 0 && dev < TAU_NOALLOC_MAX_DEVICES, "ps_tau_sum_lmajor_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        PS_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PS_REQUIRE(ptr != nullptr, "ps_tau_sum_lmajor_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = (double*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" {

long long ps_reduce_scratch_bytes(void) { return (long long)RED_BLOCKS * 64 * sizeof(double); }

int ps_weight_alm_lmajor(int precision, const void* f, int nslot, int nr, const void* h, int npol, int lmax,
                         int lmin, double weight, int ld, int col_offset, void* out_t, void* stream) {
    PS_REQUIRE(f && h && out_t, "ps_weight_alm_lmajor: NULL argument");
    PS_REQUIRE(nr >= 1 && col_offset >= 0 && col_offset + nr <= ld, "ps_weight_alm_lmajor: bad nr/ld/col_offset");
    PS_REQUIRE(npol >= 1 && npol <= 2 && nslot >= npol, "ps_weight_alm_lmajor: npol must be 1 or 2 and <= nslot");
    PS_REQUIRE(lmin >= 0 && lmax >= lmin, "ps_weight_alm_lmajor: bad l range");
    if (precision == 32) return weight_alm_impl<float>(f, nslot, nr, h, npol, lmax, lmin, weight, ld, out_t, col_offset, stream);
    PS_REQUIRE(precision == 64, "ps_weight_alm_lmajor: precision must be 32 or 64");
    return weight_alm_impl<double>(f, nslot, nr, h, npol, lmax, lmin, weight, ld, out_t, col_offset, stream);
}

int ps_weight_alm_pq_lmajor(int precision, const void* fP, const void* fQ, int nslot, int nr, const void* h,
                            int npol, int lmax, int lmin, double wP, double wQ, int ld, int col_offset, void* out_t,
                            void* stream) {
    PS_REQUIRE(fP && fQ && h && out_t, "ps_weight_alm_pq_lmajor: NULL argument");
    PS_REQUIRE(nr >= 1 && col_offset >= 0 && col_offset + 2 * nr <= ld, "ps_weight_alm_pq_lmajor: bad nr/ld/col_offset");
    PS_REQUIRE(npol >= 1 && npol <= 2 && nslot >= npol, "ps_weight_alm_pq_lmajor: npol must be 1 or 2 and <= nslot");
    PS_REQUIRE(lmin >= 0 && lmax >= lmin, "ps_weight_alm_pq_lmajor: bad l range");
    if (precision == 32)
        return weight_alm_pq_impl<float, false>(fP, fQ, nslot, nr, h, npol, lmax, lmin, wP, wQ, ld, out_t, col_offset,
                                                stream);
    PS_REQUIRE(precision == 64, "ps_weight_alm_pq_lmajor: precision must be 32 or 64");
    return weight_alm_pq_impl<double, false>(fP, fQ, nslot, nr, h, npol, lmax, lmin, wP, wQ, ld, out_t, col_offset,
                                             stream);
}

int ps_weight_alm_pq_interleaved_lmajor(int precision, const void* fP, const void* fQ, int nslot, int nr,
                                        const void* h, int npol, int lmax, int lmin, double wP, double wQ, int ld,
                                        int col_offset, void* out_t, void* stream) {
    PS_REQUIRE(fP && fQ && h && out_t, "ps_weight_alm_pq_interleaved_lmajor: NULL argument");
    PS_REQUIRE(nr >= 1 && col_offset >= 0 && col_offset + 2 * nr <= ld,
               "ps_weight_alm_pq_interleaved_lmajor: bad nr/ld/col_offset");
    PS_REQUIRE(npol >= 1 && npol <= 2 && nslot >= npol,
               "ps_weight_alm_pq_interleaved_lmajor: npol must be 1 or 2 and <= nslot");
    PS_REQUIRE(lmin >= 0 && lmax >= lmin, "ps_weight_alm_pq_interleaved_lmajor: bad l range");
    if (precision == 32)
        return weight_alm_pq_impl<float, true>(fP, fQ, nslot, nr, h, npol, lmax, lmin, wP, wQ, ld, out_t, col_offset,
                                               stream);
    PS_REQUIRE(precision == 64, "ps_weight_alm_pq_interleaved_lmajor: precision must be 32 or 64");
    return weight_alm_pq_impl<double, true>(fP, fQ, nslot, nr, h, npol, lmax, lmin, wP, wQ, ld, out_t, col_offset,
                                            stream);
}

int ps_gnl_loc_fused_sum(int precision, const void* P, const void* Q, const void* U, const void* V,
                         const double* w, int nr, long long npix, double c4, double c2, double* scratch,
                         double* out, int accumulate, void* stream) {
    PS_REQUIRE(P && Q && w && scratch && out, "ps_gnl_loc_fused_sum: NULL argument");
    PS_REQUIRE((U == nullptr) == (V == nullptr), "ps_gnl_loc_fused_sum: U and V must both be given or both NULL");
    PS_REQUIRE(nr >= 1 && npix >= 1, "ps_gnl_loc_fused_sum: empty input");
    hipStream_t st = (hipStream_t)stream;
    if (precision == 32) return reduce_launch_fused<float>(P, Q, U, V, w, nr, npix, c4, U ? c2 : 0.0, scratch, out, accumulate, st);
    PS_REQUIRE(precision == 64, "ps_gnl_loc_fused_sum: precision must be 32 or 64");
    return reduce_launch_fused<double>(P, Q, U, V, w, nr, npix, c4, U ? c2 : 0.0, scratch, out, accumulate, st);
}

int ps_gnl_loc_disc_sum(int precision, const void* P1, const void* P2, const void* Q1, const void* Q2,
                        const double* w, int nr, long long npix, double* scratch, double* out, int accumulate,
                        void* stream) {
    PS_REQUIRE(P1 && P2 && Q1 && Q2 && w && scratch && out, "ps_gnl_loc_disc_sum: NULL argument");
    PS_REQUIRE(nr >= 1 && npix >= 1, "ps_gnl_loc_disc_sum: empty input");
    hipStream_t st = (hipStream_t)stream;
    if (precision == 32)
        hipLaunchKernelGGL(k_gnl_disc<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, (const float*)P1,
                           (const float*)P2, (const float*)Q1, (const float*)Q2, w, nr, (long)npix, scratch);
    else if (precision == 64)
        hipLaunchKernelGGL(k_gnl_disc<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st, (const double*)P1,
                           (const double*)P2, (const double*)Q1, (const double*)Q2, w, nr, (long)npix, scratch);
    else
        PS_REQUIRE(false, "ps_gnl_loc_disc_sum: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (const double*)scratch, RED_BLOCKS, out, accumulate);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_mc_accumulate(int precision, const void* P, const void* Q, int nsim, long long n, void* U, void* V,
                     void* stream) {
    PS_REQUIRE(P && Q && U && V && nsim >= 1 && n >= 1, "ps_mc_accumulate: bad argument");
    hipStream_t st = (hipStream_t)stream;
    unsigned nb = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    if (precision == 32)
        hipLaunchKernelGGL(k_mc_accumulate<float>, dim3(nb), dim3(256), 0, st, (const float*)P, (const float*)Q, nsim,
                           (long)n, (double*)U, (double*)V);
    else if (precision == 64)
        hipLaunchKernelGGL(k_mc_accumulate<double>, dim3(nb), dim3(256), 0, st, (const double*)P, (const double*)Q,
                           nsim, (long)n, (double*)U, (double*)V);
    else
        PS_REQUIRE(false, "ps_mc_accumulate: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

static int products_launch(int precision, int mode, int nbatch, const void* a, const void* b, const void* c,
                           const void* d, long n, double scale, void* out, hipStream_t st) {
    PS_REQUIRE(mode >= 0 && mode <= 4, "ps_pixel_products: mode must be 0..4");
    PS_REQUIRE(a && b && out, "ps_pixel_products: NULL argument");
    PS_REQUIRE(!(mode == 1 || mode == 3 || mode == 4) || (c && d), "ps_pixel_products: c and d required for this mode");
    PS_REQUIRE(!(mode == 2) || c, "ps_pixel_products: c required for mode 2");
    PS_REQUIRE(nbatch >= 1 && nbatch <= 65535, "ps_pixel_products: nbatch must be 1..65535");
    const dim3 grid((unsigned)std::min<long>((n + 255) / 256, nbatch == 1 ? 16384 : 2048), (unsigned)nbatch);
    if (precision == 32)
        hipLaunchKernelGGL(k_products<float>, grid, dim3(256), 0, st, mode, (const float*)a, (const float*)b,
                           (const float*)c, (const float*)d, n, (float)scale, (float*)out);
    else if (precision == 64)
        hipLaunchKernelGGL(k_products<double>, grid, dim3(256), 0, st, mode, (const double*)a, (const double*)b,
                           (const double*)c, (const double*)d, n, scale, (double*)out);
    else
        PS_REQUIRE(false, "ps_pixel_products: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_pixel_products(int precision, int mode, const void* a, const void* b, const void* c, const void* d,
                      long long n, double scale, void* out, void* stream) {
    return products_launch(precision, mode, 1, a, b, c, d, (long)n, scale, out, (hipStream_t)stream);
}

int ps_pixel_products_batch(int precision, int mode, int nbatch, const void* a, const void* b, const void* c,
                            const void* d, long long n, double scale, void* out, void* stream) {
    return products_launch(precision, mode, nbatch, a, b, c, d, (long)n, scale, out, (hipStream_t)stream);
}

int ps_radial_sum(int precision, const void* lm_map, int nr, int lmax, int lmin, const void* f, int nslot, int npol,
                  const double* w, double scale, int accumulate, void* out, void* stream) {
    PS_REQUIRE(lm_map && f && w && out, "ps_radial_sum: NULL argument");
    PS_REQUIRE(nr >= 1 && npol >= 1 && npol <= nslot, "ps_radial_sum: bad nr/npol");
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    hipStream_t st = (hipStream_t)stream;
    unsigned nb = (unsigned)((nlm + 255) / 256);
    if (precision == 32)
        hipLaunchKernelGGL(k_radial_sum<float>, dim3(nb), dim3(256), 0, st, (const cplx<float>*)lm_map, nr, lmax, lmin,
                           (const float*)f, nslot, npol, w, scale, accumulate, (cplx<float>*)out);
    else if (precision == 64)
        hipLaunchKernelGGL(k_radial_sum<double>, dim3(nb), dim3(256), 0, st, (const cplx<double>*)lm_map, nr, lmax,
                           lmin, (const double*)f, nslot, npol, w, scale, accumulate, (cplx<double>*)out);
    else
        PS_REQUIRE(false, "ps_radial_sum: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_radial_scatter(int precision, const void* lm_map, int nr, int lmax, int lmin, const void* f, int nslot,
                      int npol, const double* w, double scale, int accumulate, void* out, void* stream) {
    PS_REQUIRE(lm_map && f && w && out, "ps_radial_scatter: NULL argument");
    PS_REQUIRE(nr >= 1 && nr <= 65535 && npol >= 1 && npol <= nslot, "ps_radial_scatter: bad nr/npol");
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)((nlm + 255) / 256), (unsigned)nr);
    if (precision == 32)
        hipLaunchKernelGGL(k_radial_scatter<float>, grid, dim3(256), 0, st, (const cplx<float>*)lm_map, nr, lmax, lmin,
                           (const float*)f, nslot, npol, w, scale, accumulate, (cplx<float>*)out);
    else if (precision == 64)
        hipLaunchKernelGGL(k_radial_scatter<double>, grid, dim3(256), 0, st, (const cplx<double>*)lm_map, nr, lmax,
                           lmin, (const double*)f, nslot, npol, w, scale, accumulate, (cplx<double>*)out);
    else
        PS_REQUIRE(false, "ps_radial_scatter: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_tau_sum_n0(int precision, const void* A, const void* B, int nmu, int nr, int Lmaxs, int Lmin, int Lmax,
                  const double* FLL, const double* w, int n1, double* scratch, double* out, int accumulate,
                  void* stream) {
    PS_REQUIRE(A && B && FLL && w && scratch && out, "ps_tau_sum_n0: NULL argument");
    PS_REQUIRE(Lmin >= 0 && Lmax >= Lmin && Lmax <= Lmaxs, "ps_tau_sum_n0: need 0 <= Lmin <= Lmax <= Lmaxs");
    PS_REQUIRE(Lmax - Lmin + 1 <= RED_BLOCKS * 64, "ps_tau_sum_n0: L range too large for the scratch");
    hipStream_t st = (hipStream_t)stream;
    const int nL = Lmax - Lmin + 1;
    const double pref = 1.0 / (sqrt(4.0 * M_PI) * sqrt(2.0 * n1 + 1.0));
    if (precision == 32)
        hipLaunchKernelGGL(k_tau_gram<float>, dim3(nL), dim3(256), 0, st, (const cplx<float>*)A, (const cplx<float>*)B,
                           nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, pref, scratch);
    else if (precision == 64)
        hipLaunchKernelGGL(k_tau_gram<double>, dim3(nL), dim3(256), 0, st, (const cplx<double>*)A,
                           (const cplx<double>*)B, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, pref, scratch);
    else
        PS_REQUIRE(false, "ps_tau_sum_n0: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (const double*)scratch, nL, out, accumulate);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_tau_sum_lmajor(int precision, const void* At, const void* Bt, long long ld, int nmu, int nr, int Lmaxs,
                      int Lmin, int Lmax, const double* FLL, const double* w, int n1, double* scratch, double* out,
                      int accumulate, void* stream) {
    return ps_tau_sum_lmajor_shift(precision, At, Bt, ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, n1, nullptr, 0, 0,
                                   scratch, out, accumulate, stream);
}

int ps_tau_sum_lmajor_shift(int precision, const void* At, const void* Bt, long long ld, int nmu, int nr, int Lmaxs,
                            int Lmin, int Lmax, const double* FLL, const double* w, int n1, const void* Mt,
                            int shiftA, int shiftB, double* scratch, double* out, int accumulate, void* stream) {
    PS_REQUIRE(At && Bt && FLL && w && scratch && out, "ps_tau_sum_lmajor: NULL argument");
    PS_REQUIRE(Mt || (!shiftA && !shiftB), "ps_tau_sum_lmajor_shift: a shift needs Mt");
    PS_REQUIRE(nr >= 1 && nr <= 64, "ps_tau_sum_lmajor: need 1 <= nr <= 64");
    PS_REQUIRE(nmu >= 1 && ld >= (long long)nmu * nr, "ps_tau_sum_lmajor: need ld >= nmu * nr");
    PS_REQUIRE(Lmin >= 0 && Lmax >= Lmin && Lmax <= Lmaxs, "ps_tau_sum_lmajor: need 0 <= Lmin <= Lmax <= Lmaxs");
    PS_REQUIRE(Lmax - Lmin + 1 <= RED_BLOCKS * 64, "ps_tau_sum_lmajor: L range too large for the scratch");
    hipStream_t st = (hipStream_t)stream;
    const int nL = Lmax - Lmin + 1, nt = (nr + 15) / 16;
    const double pref = 1.0 / (sqrt(4.0 * M_PI) * sqrt(2.0 * n1 + 1.0));
    const void* M = (shiftA || shiftB) ? Mt : nullptr;
    if (precision == 32)
        launch_tau_mfma_nt<float>(nt, At, Bt, (long)ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, M, shiftA, shiftB, pref,
                                  scratch, st);
    else if (precision == 64)
        launch_tau_mfma_nt<double>(nt, At, Bt, (long)ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, M, shiftA, shiftB, pref,
                                   scratch, st);
    else
        PS_REQUIRE(false, "ps_tau_sum_lmajor: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, (const double*)scratch, nL, out, accumulate);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_tau_sum_lmajor_noalloc(int precision, const void* At, const void* Bt, long long ld, int nmu, int nr, int Lmaxs,
                              int Lmin, int Lmax, const double* FLL, const double* w, int n1, double* out,
                              int accumulate, void* stream) {
    PS_REQUIRE(Lmin >= 0 && Lmax >= Lmin, "ps_tau_sum_lmajor_noalloc: need 0 <= Lmin <= Lmax");
    PS_REQUIRE(Lmax - Lmin + 1 <= TAU_NOALLOC_MAX_L,
               "ps_tau_sum_lmajor_noalloc: Lmax - Lmin + 1 exceeds the 8192 partial sums of the built-in scratch");
    double* part = nullptr;
    const int rc = tau_noalloc_scratch(&part);
    if (rc) return rc;
    return ps_tau_sum_lmajor(precision, At, Bt, ld, nmu, nr, Lmaxs, Lmin, Lmax, FLL, w, n1, part, out, accumulate,
                             stream);
}

int ps_outer_product(int precision, const void* Qa, const void* Qb, int nta, int ntb, long long nin, double* scratch,
                     double* fish, int accumulate, void* stream) {
    PS_REQUIRE(Qa && Qb && scratch && fish, "ps_outer_product: NULL argument");
    PS_REQUIRE(nta >= 1 && ntb >= 1 && nta * ntb <= 64, "ps_outer_product: need 1 <= nta*ntb <= 64");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(RED_BLOCKS, nta * ntb);
    if (precision == 32)
        hipLaunchKernelGGL(k_outer<float>, grid, dim3(256), 0, st, (const cplx<float>*)Qa, (const cplx<float>*)Qb, nta,
                           ntb, (long)nin, scratch);
    else if (precision == 64)
        hipLaunchKernelGGL(k_outer<double>, grid, dim3(256), 0, st, (const cplx<double>*)Qa, (const cplx<double>*)Qb,
                           nta, ntb, (long)nin, scratch);
    else
        PS_REQUIRE(false, "ps_outer_product: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_rows, dim3(nta * ntb), dim3(256), 0, st, (const double*)scratch, nta * ntb, RED_BLOCKS,
                       fish, accumulate);
    PS_LAUNCH_CHECK();
    return 0;
}

int ps_convolve_F_lmajor(int precision, const void* PQ, int nr, int Lmaxs, int Lmin, int Lmax, const double* FLL,
                         const double* w, int ld, void* out_t, void* stream) {
    PS_REQUIRE(PQ && FLL && w && out_t, "ps_convolve_F_lmajor: NULL argument");
    PS_REQUIRE(nr >= 1 && ld >= nr, "ps_convolve_F_lmajor: need ld >= nr >= 1");
    const long nlm = (long)(Lmaxs + 1) * (Lmaxs + 2) / 2;
    const long tot = nlm * nr;
    hipStream_t st = (hipStream_t)stream;
    unsigned nb = (unsigned)((tot + 255) / 256);
    if (precision == 32)
        hipLaunchKernelGGL(k_convolve_F<float>, dim3(nb), dim3(256), 0, st, (const cplx<float>*)PQ, nr, Lmaxs, Lmin,
                           Lmax, FLL, w, ld, (cplx<float>*)out_t);
    else if (precision == 64)
        hipLaunchKernelGGL(k_convolve_F<double>, dim3(nb), dim3(256), 0, st, (const cplx<double>*)PQ, nr, Lmaxs,
                           Lmin, Lmax, FLL, w, ld, (cplx<double>*)out_t);
    else
        PS_REQUIRE(false, "ps_convolve_F_lmajor: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

}  // extern "C"

