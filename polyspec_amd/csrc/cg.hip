// Harmonic-space vector kernels of the S^-1 filters (polyspec/weighting.py) and
// of PolySpec.applyAinv (polyspec/base.py:424-446).  Vectors are (nf, nlm)
// complex in healpy m-major order; per-l tables are FP64 (nf, nf, lmax+1) or
// (nf, lmax+1).  All HBM-bound; CG inner products reduce in FP64 with a fixed
// order.  CG scalars (alpha, beta) stay on the device so an iteration needs no
// host round trip except the convergence test.
#include <math.h>

#include "polyspec_hip.h"
#include "ps_common.h"

namespace ps {

// out[a] = post[a,l] sum_b mat[a,b,l] pre[b,l] in[b]   (accumulate: out += ...);
// blockIdx.y selects a vector of a batch (vectors in_bstride / out_bstride elements apart)
template <typename T>
__global__ void k_lmat(int nf, int lmax, const double* __restrict__ mat, const double* __restrict__ pre,
                       const double* __restrict__ post, const cplx<T>* __restrict__ in, long in_stride,
                       cplx<T>* __restrict__ out, long out_stride, double alpha, int accumulate, long in_bstride,
                       long out_bstride) {
    in += (long)blockIdx.y * in_bstride;
    out += (long)blockIdx.y * out_bstride;
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const long lm = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lm >= nlm) return;
    int l, m;
    lm_decode(lm, lmax, &l, &m);
    const int L1 = lmax + 1;
    double xr[3], xi[3];
    for (int b = 0; b < nf; b++) {
        const cplx<T> v = in[(long)b * in_stride + lm];
        const double p = pre ? pre[(long)b * L1 + l] : 1.0;
        xr[b] = (double)v.re * p;
        xi[b] = (double)v.im * p;
    }
    for (int a = 0; a < nf; a++) {
        double yr = 0.0, yi = 0.0;
        for (int b = 0; b < nf; b++) {
            const double M = mat ? mat[((long)a * nf + b) * L1 + l] : (a == b ? 1.0 : 0.0);
            yr += M * xr[b];
            yi += M * xi[b];
        }
        const double q = (post ? post[(long)a * L1 + l] : 1.0) * alpha;
        cplx<T>* o = out + (long)a * out_stride + lm;
        if (accumulate) {
            o->re = (T)((double)o->re + yr * q);
            o->im = (T)((double)o->im + yi * q);
        } else {
            *o = cplx<T>{(T)(yr * q), (T)(yi * q)};
        }
    }
}

static constexpr int DOT_BLOCKS = 512;

__device__ __forceinline__ double block_sum256(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Re sum_{a, lm} x[a,lm] conj(y[a,lm]) (1 + [m > 0])   (weighting.py:255)
// (blockIdx.y: vector of a batch, nf * nlm elements apart; partial sums gridDim.x apart)
template <typename T>
__global__ __launch_bounds__(256) void k_mdot(int nf, int lmax, const cplx<T>* __restrict__ x,
                                              const cplx<T>* __restrict__ y, double* __restrict__ part) {
    __shared__ double sh[256];
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const long n = nlm * nf;
    x += (long)blockIdx.y * n;
    y += (long)blockIdx.y * n;
    part += (long)blockIdx.y * gridDim.x;
    double acc = 0.0;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const long lm = i % nlm;
        const double w = (lm <= lmax) ? 1.0 : 2.0;  // the first lmax+1 entries are m = 0
        acc += w * ((double)x[i].re * (double)y[i].re + (double)x[i].im * (double)y[i].im);
    }
    acc = block_sum256(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// one block per vector of the batch: out[b] = sum of its n partial sums, fixed order
__global__ void k_dot_finish(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double sh[256];
    part += (long)blockIdx.x * n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    s = block_sum256(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// CG step (weighting.py:329-332): alpha = rz / pAp; x += alpha p; r -= alpha Ap.
// blockIdx.y: system of a batch (vectors n elements apart, its own scalars); systems with
// active[b] == 0 (converged) are left as they are.
template <typename T>
__global__ void k_cg_xr(long n, const double* __restrict__ rz, const double* __restrict__ pAp,
                        const int* __restrict__ active, const cplx<T>* __restrict__ p,
                        const cplx<T>* __restrict__ Ap, cplx<T>* __restrict__ x, cplx<T>* __restrict__ r) {
    const int bs = blockIdx.y;
    if (active && !active[bs]) return;
    const double alpha = rz[bs] / pAp[bs];
    const long off = (long)bs * n;
    p += off;
    Ap += off;
    x += off;
    r += off;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        x[i].re = (T)((double)x[i].re + alpha * (double)p[i].re);
        x[i].im = (T)((double)x[i].im + alpha * (double)p[i].im);
        r[i].re = (T)((double)r[i].re - alpha * (double)Ap[i].re);
        r[i].im = (T)((double)r[i].im - alpha * (double)Ap[i].im);
    }
}

// p = z + beta p with beta = num / den (weighting.py:337-339); flexible variants pass
// the Polak-Ribiere numerator.
template <typename T>
__global__ void k_cg_p(long n, const double* __restrict__ num, const double* __restrict__ den,
                       const int* __restrict__ active, const cplx<T>* __restrict__ z, cplx<T>* __restrict__ p) {
    const int bs = blockIdx.y;
    if (active && !active[bs]) return;
    const double beta = num[bs] / den[bs];
    z += (long)bs * n;
    p += (long)bs * n;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        p[i].re = (T)((double)z[i].re + beta * (double)p[i].re);
        p[i].im = (T)((double)z[i].im + beta * (double)p[i].im);
    }
}

}  // namespace ps

using namespace ps;

static int lmat_launch(int precision, int nbatch, int nf, int lmax, const double* mat, const double* pre,
                       const double* post, const void* in, long in_stride, long in_bstride, void* out,
                       long out_stride, long out_bstride, double alpha, int accumulate, hipStream_t st) {
    PS_REQUIRE(nf >= 1 && nf <= 3, "lmat_apply: nf must be 1..3");
    PS_REQUIRE(nbatch >= 1 && nbatch <= 65535, "lmat_apply: nbatch must be 1..65535");
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    const dim3 grid((unsigned)((nlm + 255) / 256), (unsigned)nbatch);
    if (precision == 32)
        hipLaunchKernelGGL(k_lmat<float>, grid, dim3(256), 0, st, nf, lmax, mat, pre, post, (const cplx<float>*)in,
                           in_stride, (cplx<float>*)out, out_stride, alpha, accumulate, in_bstride, out_bstride);
    else if (precision == 64)
        hipLaunchKernelGGL(k_lmat<double>, grid, dim3(256), 0, st, nf, lmax, mat, pre, post, (const cplx<double>*)in,
                           in_stride, (cplx<double>*)out, out_stride, alpha, accumulate, in_bstride, out_bstride);
    else
        PS_REQUIRE(false, "lmat_apply: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

// partial-sum blocks per vector: DOT_BLOCKS for one vector, fewer per vector for batches
static int dot_blocks(int nbatch) { return nbatch == 1 ? DOT_BLOCKS : 64; }

static int mdot_launch(int precision, int nbatch, int nf, int lmax, const void* x, const void* y, double* scratch,
                       double* out, hipStream_t st) {
    PS_REQUIRE(nbatch >= 1 && nbatch <= 65535, "mdot: nbatch must be 1..65535");
    const int nbk = dot_blocks(nbatch);
    const dim3 grid((unsigned)nbk, (unsigned)nbatch);
    if (precision == 32)
        hipLaunchKernelGGL(k_mdot<float>, grid, dim3(256), 0, st, nf, lmax, (const cplx<float>*)x,
                           (const cplx<float>*)y, scratch);
    else if (precision == 64)
        hipLaunchKernelGGL(k_mdot<double>, grid, dim3(256), 0, st, nf, lmax, (const cplx<double>*)x,
                           (const cplx<double>*)y, scratch);
    else
        PS_REQUIRE(false, "mdot: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dot_finish, dim3((unsigned)nbatch), dim3(256), 0, st, (const double*)scratch, nbk, out);
    PS_LAUNCH_CHECK();
    return 0;
}

static int cg_xr_launch(int precision, int nbatch, long n, const double* rz, const double* pAp, const int* active,
                        const void* p, const void* Ap, void* x, void* r, hipStream_t st) {
    PS_REQUIRE(nbatch >= 1 && nbatch <= 65535, "cg_update_xr: nbatch must be 1..65535");
    const dim3 grid((unsigned)std::min<long>((n + 255) / 256, nbatch == 1 ? 4096 : 1024), (unsigned)nbatch);
    if (precision == 32)
        hipLaunchKernelGGL(k_cg_xr<float>, grid, dim3(256), 0, st, n, rz, pAp, active, (const cplx<float>*)p,
                           (const cplx<float>*)Ap, (cplx<float>*)x, (cplx<float>*)r);
    else if (precision == 64)
        hipLaunchKernelGGL(k_cg_xr<double>, grid, dim3(256), 0, st, n, rz, pAp, active, (const cplx<double>*)p,
                           (const cplx<double>*)Ap, (cplx<double>*)x, (cplx<double>*)r);
    else
        PS_REQUIRE(false, "cg_update_xr: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

static int cg_p_launch(int precision, int nbatch, long n, const double* num, const double* den, const int* active,
                       const void* z, void* p, hipStream_t st) {
    PS_REQUIRE(nbatch >= 1 && nbatch <= 65535, "cg_update_p: nbatch must be 1..65535");
    const dim3 grid((unsigned)std::min<long>((n + 255) / 256, nbatch == 1 ? 4096 : 1024), (unsigned)nbatch);
    if (precision == 32)
        hipLaunchKernelGGL(k_cg_p<float>, grid, dim3(256), 0, st, n, num, den, active, (const cplx<float>*)z,
                           (cplx<float>*)p);
    else if (precision == 64)
        hipLaunchKernelGGL(k_cg_p<double>, grid, dim3(256), 0, st, n, num, den, active, (const cplx<double>*)z,
                           (cplx<double>*)p);
    else
        PS_REQUIRE(false, "cg_update_p: precision must be 32 or 64");
    PS_LAUNCH_CHECK();
    return 0;
}

extern "C" {

int ps_lmat_apply(int precision, int nf, int lmax, const double* mat, const double* pre, const double* post,
                  const void* in, long long in_stride, void* out, long long out_stride, double alpha, int accumulate,
                  void* stream) {
    PS_REQUIRE(in && out, "ps_lmat_apply: NULL argument");
    return lmat_launch(precision, 1, nf, lmax, mat, pre, post, in, (long)in_stride, 0, out, (long)out_stride, 0,
                       alpha, accumulate, (hipStream_t)stream);
}

int ps_lmat_apply_batch(int precision, int nbatch, int nf, int lmax, const double* mat, const double* pre,
                        const double* post, const void* in, void* out, double alpha, int accumulate, void* stream) {
    PS_REQUIRE(in && out, "ps_lmat_apply_batch: NULL argument");
    const long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    return lmat_launch(precision, nbatch, nf, lmax, mat, pre, post, in, nlm, nf * nlm, out, nlm, nf * nlm, alpha,
                       accumulate, (hipStream_t)stream);
}

int ps_mdot(int precision, int nf, int lmax, const void* x, const void* y, double* scratch, double* out,
            void* stream) {
    PS_REQUIRE(x && y && scratch && out, "ps_mdot: NULL argument");
    return mdot_launch(precision, 1, nf, lmax, x, y, scratch, out, (hipStream_t)stream);
}

int ps_mdot_batch(int precision, int nbatch, int nf, int lmax, const void* x, const void* y, double* scratch,
                  double* out, void* stream) {
    PS_REQUIRE(x && y && scratch && out, "ps_mdot_batch: NULL argument");
    return mdot_launch(precision, nbatch, nf, lmax, x, y, scratch, out, (hipStream_t)stream);
}

int ps_cg_update_xr(int precision, long long n, const double* rz, const double* pAp, const void* p, const void* Ap,
                    void* x, void* r, void* stream) {
    PS_REQUIRE(rz && pAp && p && Ap && x && r, "ps_cg_update_xr: NULL argument");
    return cg_xr_launch(precision, 1, (long)n, rz, pAp, nullptr, p, Ap, x, r, (hipStream_t)stream);
}

int ps_cg_update_xr_batch(int precision, int nbatch, long long n, const double* rz, const double* pAp,
                          const int* active, const void* p, const void* Ap, void* x, void* r, void* stream) {
    PS_REQUIRE(rz && pAp && active && p && Ap && x && r, "ps_cg_update_xr_batch: NULL argument");
    return cg_xr_launch(precision, nbatch, (long)n, rz, pAp, active, p, Ap, x, r, (hipStream_t)stream);
}

int ps_cg_update_p(int precision, long long n, const double* num, const double* den, const void* z, void* p,
                   void* stream) {
    PS_REQUIRE(num && den && z && p, "ps_cg_update_p: NULL argument");
    return cg_p_launch(precision, 1, (long)n, num, den, nullptr, z, p, (hipStream_t)stream);
}

int ps_cg_update_p_batch(int precision, int nbatch, long long n, const double* num, const double* den,
                         const int* active, const void* z, void* p, void* stream) {
    PS_REQUIRE(num && den && active && z && p, "ps_cg_update_p_batch: NULL argument");
    return cg_p_launch(precision, nbatch, (long)n, num, den, active, z, p, (hipStream_t)stream);
}

long long ps_mdot_scratch_bytes(void) { return (long long)DOT_BLOCKS * sizeof(double); }

long long ps_mdot_batch_scratch_bytes(int nbatch) {
    if (nbatch < 1) return -1;
    return (long long)dot_blocks(nbatch) * nbatch * sizeof(double);
}

}  // extern "C"
