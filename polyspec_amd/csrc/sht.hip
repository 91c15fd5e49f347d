// HEALPix spin-0 / spin-s spherical harmonic transforms for MI355X (gfx950).
//
// Replaces ducc0.sht.synthesis / adjoint_synthesis as PolySpec calls them
// (polyspec/base.py:147-151, 172-176, 193, 206, 225, 230-231, 253, 282).
//
// Two stages per direction:
//   Legendre stage  alm <-> ring phases F_m(ring) for m = 0..mmax.  One thread
//     per north/south ring pair, one block per (m, batch chunk, ring block):
//     the normalised spin-weighted Legendre recursion runs in registers, the
//     alm of the chunk are wave-uniform (scalar loads) and the batch is
//     accumulated in registers.  Seeds are built in FP64 from log2 tables and
//     carried as scaled value * 2^(64 k) so that sin^m(theta) never underflows.
//   Ring stage      ring phases <-> pixels: one block per (ring, map); m is
//     folded modulo nphi in LDS, then an in-LDS radix-2 FFT (nphi a power of
//     two) or a Bluestein chirp-z transform over a power-of-two length.
//
// Layouts: phases [map*ncomp][mmax+1][nrings] complex; maps [map][ncomp][npix];
// alm (user) [map][ncomp][nlm]; alm (Legendre input) [nlm][ld] with the
// (map, comp) columns contiguous for each (l, m).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <map>
#include <vector>

#include "ps_common.h"
#include "sht.h"

namespace ps {

// ---------------------------------------------------------------- plan

static long nlm_of(int lmax, int mmax) {
    return (long)(mmax + 1) * (lmax + 1) - (long)mmax * (mmax + 1) / 2;
}

static int ilog2(long v) {
    int r = 0;
    while ((1L << r) < v) r++;
    return r;
}

// host radix-2 FFT (double), natural order in and out; sign -1 forward
static void host_fft(std::vector<std::complex<double>>& a, int sign) {
    long n = (long)a.size();
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (long len = 2; len <= n; len <<= 1) {
        for (long i = 0; i < n; i += len)
            for (long k = 0; k < len / 2; k++) {
                double ang = sign * 2.0 * M_PI * (double)k / (double)len;
                std::complex<double> w(cos(ang), sin(ang));
                std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

template <typename T>
static int upload(T** dst, const std::vector<T>& src) {
    PS_CHECK_HIP(hipMalloc((void**)dst, sizeof(T) * std::max<size_t>(1, src.size())));
    if (!src.empty())
        PS_CHECK_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return 0;
}

int ShtPlan::build() {
    const int N = nside;
    nrings = 4 * N - 1;
    npairs = 2 * N;
    npix = 12L * N * N;
    nlm = nlm_of(lmax, mmax);
    ncomp = spin ? 2 : 1;

    // ---- rings
    std::vector<RingInfo> rings(nrings);
    std::map<int, int> blue_off;  // nphi -> offset into the Bluestein table
    long blue_total = 0;
    maxM = 1;
    for (int i = 1; i <= nrings; i++) {
        int ir = (i > 2 * N) ? 4 * N - i : i;
        RingInfo& r = rings[i - 1];
        long st;
        if (ir < N) {
            r.nphi = 4 * ir;
            r.phi_p = 1;
            r.phi_q = 4 * ir;
            st = 2L * ir * (ir - 1);
        } else {
            r.nphi = 4 * N;
            r.phi_p = (((ir - N) & 1) == 0) ? 1 : 0;
            r.phi_q = 4 * N;
            st = 2L * N * (N - 1) + (long)(ir - N) * 4 * N;
        }
        if (i > 2 * N) st = npix - st - r.nphi;
        r.start = st;
        bool pow2 = (r.nphi & (r.nphi - 1)) == 0;
        long M = pow2 ? r.nphi : (1L << ilog2(2L * r.nphi - 1));
        r.M = (int)M;
        r.logM = ilog2(M);
        r.voff = -1;
        if (!pow2) {
            auto it = blue_off.find(r.nphi);
            if (it == blue_off.end()) {
                blue_off[r.nphi] = (int)blue_total;
                r.voff = blue_total;
                blue_total += M;
            } else {
                r.voff = it->second;
            }
        }
        maxM = std::max<long>(maxM, M);
    }
    // twiddles e^{-2 pi i k / maxM}, k < maxM/2
    std::vector<double> tw((size_t)maxM);  // interleaved re, im
    for (long k = 0; k < maxM / 2; k++) {
        double s, c;
        sincos(-2.0 * M_PI * (double)k / (double)maxM, &s, &c);
        tw[2 * k] = c;
        tw[2 * k + 1] = s;
    }
    // Bluestein kernels V = FFT_M(v), v_t = exp(-i pi t^2 / n) wrapped; stored
    // in bit-reversed order (the pointwise product is taken between a DIF
    // forward transform and a DIT inverse transform).
    std::vector<double> blue((size_t)std::max<long>(1, 2 * blue_total));
    for (auto& kv : blue_off) {
        int n = kv.first;
        long M = 1L << ilog2(2L * n - 1);
        std::vector<std::complex<double>> v(M, 0.0);
        for (long t = 0; t < n; t++) {
            long t2 = (t * t) % (2L * n);
            double s, c;
            sincos(-M_PI * (double)t2 / (double)n, &s, &c);
            v[t] = std::complex<double>(c, s);
            if (t > 0) v[M - t] = v[t];
        }
        host_fft(v, -1);
        int lg = ilog2(M);
        for (long k = 0; k < M; k++) {
            unsigned rk = 0;
            for (int b = 0; b < lg; b++)
                if (k & (1L << b)) rk |= 1u << (lg - 1 - b);
            blue[2 * (kv.second + rk)] = v[k].real();
            blue[2 * (kv.second + rk) + 1] = v[k].imag();
        }
    }

    // ---- ring pairs (north ring jp, south ring nrings-1-jp)
    std::vector<double> cth(npairs), l2s(npairs), l2sh(npairs), l2ch(npairs);
    for (int jp = 0; jp < npairs; jp++) {
        int ir = jp + 1;
        double z, omz;
        if (ir < N) {
            omz = (double)ir * ir / (3.0 * N * N);
            z = 1.0 - omz;
        } else {
            z = 2.0 * (2 * N - ir) / (3.0 * N);
            omz = 1.0 - z;
        }
        double sth = sqrt(omz * (1.0 + z));
        cth[jp] = z;
        l2s[jp] = log2(sth);
        // sin^2(theta/2) = (1 - z)/2, cos^2(theta/2) = (1 + z)/2
        l2sh[jp] = 0.5 * log2(0.5 * omz);
        l2ch[jp] = 0.5 * log2(0.5 * (1.0 + z));
    }

    // ---- seeds per m: log2 f_{l0} = L + A log2 sin(theta/2) + B log2 cos(theta/2)
    const int s = spin;
    const int nsg = spin ? 2 : 1;
    std::vector<double> seedL((size_t)nsg * (mmax + 1)), seedS((size_t)nsg * (mmax + 1));
    std::vector<int> seedA((size_t)nsg * (mmax + 1)), seedB((size_t)nsg * (mmax + 1));
    const double LOG2E = 1.0 / M_LN2;
    for (int m = 0; m <= mmax; m++) {
        for (int g = 0; g < nsg; g++) {
            int ss = (g == 0) ? s : -s;
            double lv;
            int A, B;
            double sg;
            if (s == 0) {
                // (-1)^m sqrt((2m+1)/(4pi) (2m)!/(4^m m!^2)) (2 sin(t/2) cos(t/2))^m
                lv = 0.5 * (log((2.0 * m + 1.0) / (4.0 * M_PI)) + lgamma(2.0 * m + 1.0) -
                            2.0 * lgamma(m + 1.0) - 2.0 * m * log(2.0)) * LOG2E + m;
                A = m;
                B = m;
                sg = (m & 1) ? -1.0 : 1.0;
            } else if (m >= s) {
                lv = 0.5 * (log((2.0 * m + 1.0) / (4.0 * M_PI)) + lgamma(2.0 * m + 1.0) -
                            lgamma(m + s + 1.0) - lgamma(m - s + 1.0)) * LOG2E;
                A = (ss > 0) ? m + s : m - s;
                B = (ss > 0) ? m - s : m + s;
                sg = (m & 1) ? -1.0 : 1.0;
            } else {
                lv = 0.5 * (log((2.0 * s + 1.0) / (4.0 * M_PI)) + lgamma(2.0 * s + 1.0) -
                            lgamma(s + m + 1.0) - lgamma(s - m + 1.0)) * LOG2E;
                A = (ss > 0) ? s + m : s - m;
                B = (ss > 0) ? s - m : s + m;
                sg = (ss > 0) ? ((m & 1) ? -1.0 : 1.0) : ((s & 1) ? -1.0 : 1.0);
            }
            seedL[(size_t)g * (mmax + 1) + m] = lv;
            seedA[(size_t)g * (mmax + 1) + m] = A;
            seedB[(size_t)g * (mmax + 1) + m] = B;
            seedS[(size_t)g * (mmax + 1) + m] = sg;
        }
    }

    // ---- recursion coefficients per (l, m): f_{l+1} = alpha (x + beta) f_l - gamma f_{l-1}
    std::vector<double> coef((size_t)nlm * 4, 0.0);
    for (int m = 0; m <= mmax; m++) {
        long base = (long)m * (2 * lmax + 1 - m) / 2;
        for (int l = m; l <= lmax; l++) {
            double lp = l + 1.0, dl = l, ds = s;
            double den = sqrt((lp * lp - (double)m * m) * (lp * lp - ds * ds));
            double al = (2.0 * dl + 1.0) * lp * sqrt((2.0 * dl + 3.0) / (2.0 * dl + 1.0)) / den;
            double be = 0.0, ga = 0.0;
            if (l > 0) {
                be = (s == 0) ? 0.0 : (double)m * s / (dl * lp);
                double t = (dl * dl - (double)m * m) * (dl * dl - ds * ds);
                ga = (t > 0) ? lp / dl * sqrt(t) / den * sqrt((2.0 * dl + 3.0) / (2.0 * dl - 1.0)) : 0.0;
            }
            double* c = &coef[(size_t)(base + l) * 4];
            c[0] = al;
            c[1] = be;
            c[2] = ga;
        }
    }

    // ---- upload
    if (upload(&d_rings, rings)) return -2;
    if (upload(&d_l2s, l2s)) return -2;
    if (upload(&d_l2sh, l2sh)) return -2;
    if (upload(&d_l2ch, l2ch)) return -2;
    if (upload(&d_seedL, seedL)) return -2;
    if (upload(&d_seedS, seedS)) return -2;
    if (upload(&d_seedA, seedA)) return -2;
    if (upload(&d_seedB, seedB)) return -2;
    if (upload((double**)&d_cth, cth)) return -2;
    if (upload((double**)&d_coef, coef)) return -2;
    if (prec == 4) {
        std::vector<float> tf(tw.begin(), tw.end());
        if (upload((float**)&d_tw, tf)) return -2;
        std::vector<float> bf(blue.begin(), blue.end());
        if (upload((float**)&d_blue, bf)) return -2;
    } else {
        if (upload((double**)&d_tw, tw)) return -2;
        if (upload((double**)&d_blue, blue)) return -2;
    }
    return 0;
}

ShtPlan::~ShtPlan() {
    void* ptrs[] = {d_rings, d_l2s, d_l2sh, d_l2ch, d_seedL, d_seedS, d_seedA,
                    d_seedB, d_cth, d_tw, d_blue, d_coef};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

// ------------------------------------------------------------ device views

template <typename T>
struct ShtDev {
    int nside, lmax, mmax, spin, nrings, npairs;
    long npix, nlm;
    const RingInfo* rings;
    const double *l2s, *l2sh, *l2ch, *seedL, *seedS;
    const int *seedA, *seedB;
    const double* cth;  // recursion runs in FP64 for both precisions
    const cplx<T>* tw;
    const cplx<T>* blue;
    const double* coef;  // (alpha, beta, gamma, 0) per idx(l, m)
    long maxM;
};

template <typename T>
static ShtDev<T> dev_view(const ShtPlan& p) {
    ShtDev<T> d;
    d.nside = p.nside;
    d.lmax = p.lmax;
    d.mmax = p.mmax;
    d.spin = p.spin;
    d.nrings = p.nrings;
    d.npairs = p.npairs;
    d.npix = p.npix;
    d.nlm = p.nlm;
    d.rings = p.d_rings;
    d.l2s = p.d_l2s;
    d.l2sh = p.d_l2sh;
    d.l2ch = p.d_l2ch;
    d.seedL = p.d_seedL;
    d.seedS = p.d_seedS;
    d.seedA = p.d_seedA;
    d.seedB = p.d_seedB;
    d.cth = (const double*)p.d_cth;
    d.tw = (const cplx<T>*)p.d_tw;
    d.blue = (const cplx<T>*)p.d_blue;
    d.coef = (const double*)p.d_coef;
    d.maxM = p.maxM;
    return d;
}

// Recursion state of one normalised Legendre function for one ring.  The
// recursion itself always runs in FP64: in FP32 the rounding of cos(theta) and
// of the coefficients shifts the recurrence frequency near the poles and the
// error grows like l * theta-resolution (measured 4e-4 at lmax 512); only the
// batch accumulation uses the storage precision T.
struct Rec {
    double f, fm1, cf;
    int k;
};

__device__ __forceinline__ void rec_seed(Rec& r, double l2v, double sign) {
    int k = 0;
    if (l2v < 0.0) {
        k = (int)floor(l2v * (1.0 / 64.0)) + 1;
        if (k > 0) k = 0;
    }
    r.f = sign * exp2(l2v - 64.0 * (double)k);
    r.fm1 = 0.0;
    r.k = k;
    r.cf = scale_factor<double>(k);
}

__device__ __forceinline__ void rec_step(Rec& r, double ax, double g) {
    // ax = alpha * (x + beta), g = gamma
    const double fn = ax * r.f - g * r.fm1;
    r.fm1 = r.f;
    r.f = fn;
    if (r.k < 0 && fabs(fn) > scale_consts<double>::THR) {
        r.f *= scale_consts<double>::DN;
        r.fm1 *= scale_consts<double>::DN;
        r.k += 1;
        r.cf = scale_factor<double>(r.k);
    }
}

__device__ __forceinline__ int m_of_q(int q, int mmax) {
    // interleave heavy (small m) and light (large m) blocks: 0, mmax, 1, mmax-1, ...
    return (q & 1) ? (mmax - (q >> 1)) : (q >> 1);
}

// -------------------------------------------------- Legendre synthesis, spin 0
// alm_t: [nlm][ld] complex, ld >= nchunk*BC (padding columns are read, never used)
template <typename T, int BC>
__global__ __launch_bounds__(256) void k_leg_synth_s0(ShtDev<T> d, int ncol, int ld,
                                                      const cplx<T>* __restrict__ alm_t,
                                                      cplx<T>* __restrict__ phase) {
    const int nrb = (d.npairs + 255) / 256;
    const int nchunk = (ncol + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = lb % nrb;
    const int chunk = (lb / nrb) % nchunk;
    const int m = m_of_q(lb / (nrb * nchunk), d.mmax);
    const int jp = rb * 256 + threadIdx.x;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;

    const double x = d.cth[jpc];
    Rec r;
    rec_seed(r, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);

    cplx<T> ev[BC], od[BC];
#pragma unroll
    for (int j = 0; j < BC; j++) ev[j] = od[j] = cplx<T>{T(0), T(0)};

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const cplx<T>* a = alm_t + (size_t)chunk * BC;
    const double* cf4 = d.coef + 4 * base;
    int l = m;
    for (; l + 1 <= d.lmax; l += 2) {
        if (__any(r.k >= scale_consts<T>::KMIN)) {
            const T v = (T)(r.f * r.cf);
            const cplx<T>* al = a + (size_t)(base + l) * ld;
#pragma unroll
            for (int j = 0; j < BC; j++) {
                ev[j].re += v * al[j].re;
                ev[j].im += v * al[j].im;
            }
        }
        rec_step(r, cf4[4 * l] * x, cf4[4 * l + 2]);
        if (__any(r.k >= scale_consts<T>::KMIN)) {
            const T v = (T)(r.f * r.cf);
            const cplx<T>* al = a + (size_t)(base + l + 1) * ld;
#pragma unroll
            for (int j = 0; j < BC; j++) {
                od[j].re += v * al[j].re;
                od[j].im += v * al[j].im;
            }
        }
        rec_step(r, cf4[4 * (l + 1)] * x, cf4[4 * (l + 1) + 2]);
    }
    if (l <= d.lmax) {  // trailing even-parity term
        const T v = (T)(r.f * r.cf);
        const cplx<T>* al = a + (size_t)(base + l) * ld;
#pragma unroll
        for (int j = 0; j < BC; j++) {
            ev[j].re += v * al[j].re;
            ev[j].im += v * al[j].im;
        }
    }
    if (jp >= d.npairs) return;
    const int rn = jp, rs = d.nrings - 1 - jp;
    const size_t pstride = (size_t)(d.mmax + 1) * d.nrings;
#pragma unroll
    for (int j = 0; j < BC; j++) {
        const int col = chunk * BC + j;
        if (col < ncol) {
            cplx<T>* ph = phase + (size_t)col * pstride + (size_t)m * d.nrings;
            ph[rn] = cadd(ev[j], od[j]);
            if (rs != rn) ph[rs] = csub(ev[j], od[j]);
        }
    }
}

// -------------------------------------------------- Legendre synthesis, spin s
// alm_t: [nlm][ld] complex with columns (map b, comp c) at 2 b + c (G, C).
template <typename T, int BC>
__global__ __launch_bounds__(256) void k_leg_synth_sp(ShtDev<T> d, int nmap, int ld,
                                                      const cplx<T>* __restrict__ alm_t,
                                                      cplx<T>* __restrict__ phase) {
    const int nrb = (d.npairs + 255) / 256;
    const int nchunk = (nmap + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = lb % nrb;
    const int chunk = (lb / nrb) % nchunk;
    const int m = m_of_q(lb / (nrb * nchunk), d.mmax);
    const int jp = rb * 256 + threadIdx.x;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    const int M1 = d.mmax + 1;

    const double x = d.cth[jpc];
    Rec rp, rm;
    rec_seed(rp, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
    rec_seed(rm, d.seedL[M1 + m] + d.seedA[M1 + m] * d.l2sh[jpc] + d.seedB[M1 + m] * d.l2ch[jpc],
             d.seedS[M1 + m]);

    // acc[par][j][0..3] = sum G lam+, C lam+, G lam-, C lam-
    cplx<T> acc[2][BC][4];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int j = 0; j < BC; j++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[p][j][c] = cplx<T>{T(0), T(0)};

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const cplx<T>* a = alm_t + (size_t)chunk * BC * 2;
    const double* cf4 = d.coef + 4 * base;
    const int l0 = m > d.spin ? m : d.spin;
    for (int l = l0; l <= d.lmax; l++) {
        const int par = (l + m) & 1;
        if (__any((rp.k >= scale_consts<T>::KMIN) || (rm.k >= scale_consts<T>::KMIN))) {
            const double fp = rp.f * rp.cf, fm = rm.f * rm.cf;
            const T lp = (T)(0.5 * (fp + fm)), ln = (T)(0.5 * (fp - fm));
            const cplx<T>* al = a + (size_t)(base + l) * ld;
            if (par == 0) {
#pragma unroll
                for (int j = 0; j < BC; j++) {
                    cplx<T> g = al[2 * j], c = al[2 * j + 1];
                    acc[0][j][0].re += lp * g.re; acc[0][j][0].im += lp * g.im;
                    acc[0][j][1].re += lp * c.re; acc[0][j][1].im += lp * c.im;
                    acc[0][j][2].re += ln * g.re; acc[0][j][2].im += ln * g.im;
                    acc[0][j][3].re += ln * c.re; acc[0][j][3].im += ln * c.im;
                }
            } else {
#pragma unroll
                for (int j = 0; j < BC; j++) {
                    cplx<T> g = al[2 * j], c = al[2 * j + 1];
                    acc[1][j][0].re += lp * g.re; acc[1][j][0].im += lp * g.im;
                    acc[1][j][1].re += lp * c.re; acc[1][j][1].im += lp * c.im;
                    acc[1][j][2].re += ln * g.re; acc[1][j][2].im += ln * g.im;
                    acc[1][j][3].re += ln * c.re; acc[1][j][3].im += ln * c.im;
                }
            }
        }
        const double al_ = cf4[4 * l], be = cf4[4 * l + 1], ga = cf4[4 * l + 2];
        rec_step(rp, al_ * (x + be), ga);
        rec_step(rm, al_ * (x - be), ga);
    }
    if (jp >= d.npairs) return;
    const int rn = jp, rs = d.nrings - 1 - jp;
    const size_t pstride = (size_t)M1 * d.nrings;
#pragma unroll
    for (int j = 0; j < BC; j++) {
        const int b = chunk * BC + j;
        if (b >= nmap) continue;
        cplx<T>* phR = phase + (size_t)(2 * b) * pstride + (size_t)m * d.nrings;
        cplx<T>* phI = phase + (size_t)(2 * b + 1) * pstride + (size_t)m * d.nrings;
        // R = -(G lam+ + i C lam-), I = -(C lam+ - i G lam-); south: lam+ -> (-1)^(l+m),
        // lam- -> -(-1)^(l+m)
        const cplx<T> GPe = acc[0][j][0], GPo = acc[1][j][0], CPe = acc[0][j][1], CPo = acc[1][j][1];
        const cplx<T> GMe = acc[0][j][2], GMo = acc[1][j][2], CMe = acc[0][j][3], CMo = acc[1][j][3];
        cplx<T> t;
        // north
        t = cadd(CMe, CMo);
        phR[rn] = cplx<T>{-(GPe.re + GPo.re) + t.im, -(GPe.im + GPo.im) - t.re};
        t = cadd(GMe, GMo);
        phI[rn] = cplx<T>{-(CPe.re + CPo.re) - t.im, -(CPe.im + CPo.im) + t.re};
        if (rs != rn) {
            t = csub(CMo, CMe);
            phR[rs] = cplx<T>{-(GPe.re - GPo.re) + t.im, -(GPe.im - GPo.im) - t.re};
            t = csub(GMo, GMe);
            phI[rs] = cplx<T>{-(CPe.re - CPo.re) - t.im, -(CPe.im - CPo.im) + t.re};
        }
    }
}

// ----------------------------------------------------------- ring FFTs (LDS)

__device__ __forceinline__ int bitrev(int k, int logM) {
    return logM ? (int)(__brev((unsigned)k) >> (32 - logM)) : 0;
}

// Decimation in frequency: natural order in, bit-reversed out.
template <typename T>
__device__ void fft_dif(cplx<T>* a, int M, const cplx<T>* __restrict__ tw, long maxM, bool inv) {
    for (int h = M >> 1; h >= 1; h >>= 1) {
        const long ts = maxM / (2L * h);
        for (int t = threadIdx.x; t < (M >> 1); t += blockDim.x) {
            const int pos = t & (h - 1);
            const int i = ((t - pos) << 1) + pos, j = i + h;
            cplx<T> w = tw[pos * ts];
            if (inv) w.im = -w.im;
            const cplx<T> u = a[i], v = a[j];
            a[i] = cadd(u, v);
            a[j] = cmul(csub(u, v), w);
        }
        __syncthreads();
    }
}

// Decimation in time: bit-reversed order in, natural out.
template <typename T>
__device__ void fft_dit(cplx<T>* a, int M, const cplx<T>* __restrict__ tw, long maxM, bool inv) {
    for (int h = 1; h < M; h <<= 1) {
        const long ts = maxM / (2L * h);
        for (int t = threadIdx.x; t < (M >> 1); t += blockDim.x) {
            const int pos = t & (h - 1);
            const int i = ((t - pos) << 1) + pos, j = i + h;
            cplx<T> w = tw[pos * ts];
            if (inv) w.im = -w.im;
            const cplx<T> u = a[i], v = cmul(a[j], w);
            a[i] = cadd(u, v);
            a[j] = csub(u, v);
        }
        __syncthreads();
    }
}

template <typename T>
__device__ __forceinline__ cplx<T> expi_pi(long num, long den) {
    // exp(i pi num/den) with exact argument reduction
    long r = num % (2 * den);
    if (r < 0) r += 2 * den;
    T s, c;
    sincospi_t((T)r / (T)den, &s, &c);
    return cplx<T>{c, s};
}

// phases [col][mmax+1][nrings] -> maps [col][npix]
template <typename T>
__global__ __launch_bounds__(256) void k_ring_synth(ShtDev<T> d, const cplx<T>* __restrict__ phase,
                                                    T* __restrict__ map, long map_stride) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int ring = lb % d.nrings;
    const int col = lb / d.nrings;
    const RingInfo ri = d.rings[ring];
    const int n = ri.nphi, M = ri.M, logM = ri.logM;
    const bool pow2 = ri.voff < 0;
    const cplx<T>* F = phase + ((size_t)col * (d.mmax + 1)) * d.nrings + ring;
    const long phq = ri.phi_q, php = ri.phi_p;

    if (!pow2)
        for (int j = threadIdx.x; j < M; j += blockDim.x) buf[j] = cplx<T>{T(0), T(0)};
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        cplx<T> c{T(0), T(0)};
        for (int m = j; m <= d.mmax; m += n) {
            cplx<T> g = cmul(F[(size_t)m * d.nrings], expi_pi<T>(m * php, phq));
            c = cadd(c, g);
        }
        for (int m = (j == 0) ? n : n - j; m <= d.mmax; m += n) {
            cplx<T> g = cmul(F[(size_t)m * d.nrings], expi_pi<T>(m * php, phq));
            c = cadd(c, cconj(g));
        }
        if (!pow2) c = cmul(c, expi_pi<T>((long)j * j, n));  // Bluestein pre-chirp
        buf[j] = c;
    }
    __syncthreads();
    T* out = map + (size_t)col * map_stride + ri.start;
    if (pow2) {
        fft_dif(buf, M, d.tw, d.maxM, true);
        for (int k = threadIdx.x; k < n; k += blockDim.x) out[k] = buf[bitrev(k, logM)].re;
    } else {
        fft_dif(buf, M, d.tw, d.maxM, false);
        const cplx<T>* V = d.blue + ri.voff;
        for (int k = threadIdx.x; k < M; k += blockDim.x) buf[k] = cmul(buf[k], V[k]);
        __syncthreads();
        fft_dit(buf, M, d.tw, d.maxM, true);
        const T invM = T(1) / (T)M;
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            cplx<T> v = cmul(buf[k], expi_pi<T>((long)k * k, n));
            out[k] = v.re * invM;
        }
    }
}

// maps [col][npix] -> phases [col][mmax+1][nrings]: sum_k f_k exp(-i m phi_k)
template <typename T>
__global__ __launch_bounds__(256) void k_ring_adjoint(ShtDev<T> d, const T* __restrict__ map,
                                                      long map_stride, cplx<T>* __restrict__ phase) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int ring = lb % d.nrings;
    const int col = lb / d.nrings;
    const RingInfo ri = d.rings[ring];
    const int n = ri.nphi, M = ri.M, logM = ri.logM;
    const bool pow2 = ri.voff < 0;
    const T* in = map + (size_t)col * map_stride + ri.start;
    const long phq = ri.phi_q, php = ri.phi_p;

    for (int k = threadIdx.x; k < M; k += blockDim.x) {
        cplx<T> v{T(0), T(0)};
        if (k < n) {
            v.re = in[k];
            if (!pow2) v = cmul(v, expi_pi<T>((long)k * k, n));
        }
        buf[k] = v;
    }
    __syncthreads();
    if (pow2) {
        fft_dif(buf, M, d.tw, d.maxM, false);
    } else {
        // g_j = sum_k f_k exp(+2 pi i jk/n) by Bluestein; C_j = conj(g_j)
        fft_dif(buf, M, d.tw, d.maxM, false);
        const cplx<T>* V = d.blue + ri.voff;
        for (int k = threadIdx.x; k < M; k += blockDim.x) buf[k] = cmul(buf[k], V[k]);
        __syncthreads();
        fft_dit(buf, M, d.tw, d.maxM, true);
    }
    cplx<T>* P = phase + ((size_t)col * (d.mmax + 1)) * d.nrings + ring;
    const T invM = T(1) / (T)M;
    for (int m = threadIdx.x; m <= d.mmax; m += blockDim.x) {
        const int j = m % n;
        cplx<T> C;
        if (pow2) {
            C = buf[bitrev(j, logM)];
        } else {
            cplx<T> g = cmul(buf[j], expi_pi<T>((long)j * j, n));
            C = cplx<T>{g.re * invM, -g.im * invM};
        }
        P[(size_t)m * d.nrings] = cmul(C, expi_pi<T>(-(long)m * php, phq));
    }
}

// ------------------------------------------------ Legendre adjoint, spin 0
// phases [col][mmax+1][nrings] -> alm [col][nlm] (scaled). One block per (m, chunk);
// thread t owns ring pairs t + 256 r, r < NRT.
template <typename T, int NRT, int BC, int LC>
__global__ __launch_bounds__(256) void k_leg_adj_s0(ShtDev<T> d, int ncol, const cplx<T>* __restrict__ phase,
                                                    cplx<T>* __restrict__ alm, long alm_stride, T scale) {
    __shared__ T red[LC * BC * 2][257];
    const int nchunk = (ncol + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int m = m_of_q(lb / nchunk, d.mmax);
    const int tid = threadIdx.x;
    const size_t pstride = (size_t)(d.mmax + 1) * d.nrings;

    double x[NRT];
    Rec r[NRT];
    cplx<T> ev[NRT][BC], od[NRT][BC];
#pragma unroll
    for (int q = 0; q < NRT; q++) {
        const int jp = tid + 256 * q;
        const bool ok = jp < d.npairs;
        const int jpc = ok ? jp : 0;
        x[q] = d.cth[jpc];
        rec_seed(r[q], d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
        if (!ok) r[q].k = -1000000, r[q].cf = 0.0, r[q].f = 0.0;
        const int rn = jpc, rs = d.nrings - 1 - jpc;
#pragma unroll
        for (int j = 0; j < BC; j++) {
            const int col = chunk * BC + j;
            cplx<T> pn{T(0), T(0)}, ps{T(0), T(0)};
            if (ok && col < ncol) {
                const cplx<T>* P = phase + (size_t)col * pstride + (size_t)m * d.nrings;
                pn = P[rn];
                if (rs != rn) ps = P[rs];
            }
            ev[q][j] = cadd(pn, ps);
            od[q][j] = csub(pn, ps);
        }
    }
    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const double* cf4 = d.coef + 4 * base;
    for (int l0 = m; l0 <= d.lmax; l0 += LC) {
        cplx<T> acc[LC][BC];
#pragma unroll
        for (int c = 0; c < LC; c++)
#pragma unroll
            for (int j = 0; j < BC; j++) acc[c][j] = cplx<T>{T(0), T(0)};
#pragma unroll
        for (int c = 0; c < LC; c++) {
            const int l = l0 + c;
            if (l <= d.lmax) {
                const double al = cf4[4 * l], ga = cf4[4 * l + 2];
#pragma unroll
                for (int q = 0; q < NRT; q++) {
                    const T v = (T)(r[q].f * r[q].cf);
#pragma unroll
                    for (int j = 0; j < BC; j++) {
                        const cplx<T> e = (c & 1) ? od[q][j] : ev[q][j];
                        acc[c][j].re += v * e.re;
                        acc[c][j].im += v * e.im;
                    }
                    rec_step(r[q], al * x[q], ga);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < LC; c++)
#pragma unroll
            for (int j = 0; j < BC; j++) {
                red[(c * BC + j) * 2][tid] = acc[c][j].re;
                red[(c * BC + j) * 2 + 1][tid] = acc[c][j].im;
            }
        __syncthreads();
        // 256 threads reduce LC*BC*2 rows of 256 partial sums
        constexpr int NROW = LC * BC * 2;
        constexpr int TPR = 256 / NROW;  // threads per row
        {
            const int row = tid / TPR, part = tid % TPR;
            T s = T(0);
            for (int i = part; i < 256; i += TPR) s += red[row][i];
            // combine the TPR partials of a row (consecutive lanes) with shuffles
            for (int o = TPR / 2; o >= 1; o >>= 1) s += __shfl_down(s, o, TPR);
            __syncthreads();
            if (part == 0) red[row][256] = s;
        }
        __syncthreads();
        if (tid < LC * BC) {
            const int c = tid / BC, j = tid % BC;
            const int l = l0 + c, col = chunk * BC + j;
            if (l <= d.lmax && col < ncol)
                alm[(size_t)col * alm_stride + base + l] =
                    cplx<T>{red[(c * BC + j) * 2][256] * scale, red[(c * BC + j) * 2 + 1][256] * scale};
        }
        __syncthreads();
    }
}

// ------------------------------------------------ Legendre adjoint, spin s
// phases [(b, R|I)][mmax+1][nrings] -> alm [(b, G|C)][nlm] (scaled)
template <typename T, int NRT, int BC, int LC>
__global__ __launch_bounds__(256) void k_leg_adj_sp(ShtDev<T> d, int nmap, const cplx<T>* __restrict__ phase,
                                                    cplx<T>* __restrict__ alm, long alm_stride, T scale) {
    constexpr int NROW = LC * BC * 4;
    __shared__ T red[NROW][257];
    const int nchunk = (nmap + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int m = m_of_q(lb / nchunk, d.mmax);
    const int tid = threadIdx.x;
    const int M1 = d.mmax + 1;
    const size_t pstride = (size_t)M1 * d.nrings;

    double x[NRT];
    Rec rp[NRT], rm[NRT];
    // per ring / map: Rsum, Rdif, Isum, Idif
    cplx<T> in4[NRT][BC][4];
#pragma unroll
    for (int q = 0; q < NRT; q++) {
        const int jp = tid + 256 * q;
        const bool ok = jp < d.npairs;
        const int jpc = ok ? jp : 0;
        x[q] = d.cth[jpc];
        rec_seed(rp[q], d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
        rec_seed(rm[q], d.seedL[M1 + m] + d.seedA[M1 + m] * d.l2sh[jpc] + d.seedB[M1 + m] * d.l2ch[jpc],
                 d.seedS[M1 + m]);
        if (!ok) {
            rp[q].k = rm[q].k = -1000000;
            rp[q].cf = rm[q].cf = 0.0;
            rp[q].f = rm[q].f = 0.0;
        }
        const int rn = jpc, rs = d.nrings - 1 - jpc;
#pragma unroll
        for (int j = 0; j < BC; j++) {
            const int b = chunk * BC + j;
            cplx<T> Rn{T(0), T(0)}, Rs{T(0), T(0)}, In{T(0), T(0)}, Is{T(0), T(0)};
            if (ok && b < nmap) {
                const cplx<T>* PR = phase + (size_t)(2 * b) * pstride + (size_t)m * d.nrings;
                const cplx<T>* PI = phase + (size_t)(2 * b + 1) * pstride + (size_t)m * d.nrings;
                Rn = PR[rn];
                In = PI[rn];
                if (rs != rn) {
                    Rs = PR[rs];
                    Is = PI[rs];
                }
            }
            in4[q][j][0] = cadd(Rn, Rs);
            in4[q][j][1] = csub(Rn, Rs);
            in4[q][j][2] = cadd(In, Is);
            in4[q][j][3] = csub(In, Is);
        }
    }
    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const double* cf4 = d.coef + 4 * base;
    const int lstart = m > d.spin ? m : d.spin;
    // coefficients with m <= l < spin vanish identically
    for (int i = tid; i < (lstart - m) * BC * 2; i += blockDim.x) {
        const int l = m + i / (BC * 2), b = chunk * BC + (i / 2) % BC, g = i % 2;
        if (l <= d.lmax && b < nmap) alm[(size_t)(2 * b + g) * alm_stride + base + l] = cplx<T>{T(0), T(0)};
    }
    for (int l0 = lstart; l0 <= d.lmax; l0 += LC) {
        cplx<T> acc[LC][BC][2];  // G, C
#pragma unroll
        for (int c = 0; c < LC; c++)
#pragma unroll
            for (int j = 0; j < BC; j++) acc[c][j][0] = acc[c][j][1] = cplx<T>{T(0), T(0)};
#pragma unroll
        for (int c = 0; c < LC; c++) {
            const int l = l0 + c;
            if (l <= d.lmax) {
                const bool even = ((l + m) & 1) == 0;
                const double al = cf4[4 * l], be = cf4[4 * l + 1], ga = cf4[4 * l + 2];
#pragma unroll
                for (int q = 0; q < NRT; q++) {
                    const double fp = rp[q].f * rp[q].cf, fm = rm[q].f * rm[q].cf;
                    const T lp = (T)(0.5 * (fp + fm)), ln = (T)(0.5 * (fp - fm));
#pragma unroll
                    for (int j = 0; j < BC; j++) {
                        const cplx<T> Rp = even ? in4[q][j][0] : in4[q][j][1];
                        const cplx<T> Ip = even ? in4[q][j][2] : in4[q][j][3];
                        const cplx<T> Rm = even ? in4[q][j][1] : in4[q][j][0];
                        const cplx<T> Im = even ? in4[q][j][3] : in4[q][j][2];
                        // G += -(lp Rp + i ln Im), C += -(lp Ip - i ln Rm)
                        acc[c][j][0].re += -lp * Rp.re + ln * Im.im;
                        acc[c][j][0].im += -lp * Rp.im - ln * Im.re;
                        acc[c][j][1].re += -lp * Ip.re - ln * Rm.im;
                        acc[c][j][1].im += -lp * Ip.im + ln * Rm.re;
                    }
                    rec_step(rp[q], al * (x[q] + be), ga);
                    rec_step(rm[q], al * (x[q] - be), ga);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < LC; c++)
#pragma unroll
            for (int j = 0; j < BC; j++)
#pragma unroll
                for (int g = 0; g < 2; g++) {
                    red[((c * BC + j) * 2 + g) * 2][tid] = acc[c][j][g].re;
                    red[((c * BC + j) * 2 + g) * 2 + 1][tid] = acc[c][j][g].im;
                }
        __syncthreads();
        constexpr int TPR = 256 / NROW;
        {
            const int row = tid / TPR, part = tid % TPR;
            T s = T(0);
            for (int i = part; i < 256; i += TPR) s += red[row][i];
            for (int o = TPR / 2; o >= 1; o >>= 1) s += __shfl_down(s, o, TPR);
            __syncthreads();
            if (part == 0) red[row][256] = s;
        }
        __syncthreads();
        if (tid < LC * BC * 2) {
            const int c = tid / (BC * 2), j = (tid / 2) % BC, g = tid % 2;
            const int l = l0 + c, b = chunk * BC + j;
            if (l <= d.lmax && b < nmap) {
                const int row = ((c * BC + j) * 2 + g) * 2;
                alm[(size_t)(2 * b + g) * alm_stride + base + l] =
                    cplx<T>{red[row][256] * scale, red[row + 1][256] * scale};
            }
        }
        __syncthreads();
    }
}

// --------------------------------------------------------- layout helpers
// alm [col][nlm] (row stride alm_stride) -> alm_t [nlm][ld]; padding columns zeroed
template <typename T>
__global__ void k_alm_to_lmajor(const cplx<T>* __restrict__ alm, long alm_stride, int ncol, long nlm,
                                int ld, cplx<T>* __restrict__ out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nlm * ld) return;
    long lm = i / ld;
    int col = (int)(i % ld);
    out[i] = (col < ncol) ? alm[(size_t)col * alm_stride + lm] : cplx<T>{T(0), T(0)};
}

// ---------------------------------------------------------------- drivers

template <typename T>
static size_t lds_bytes(const ShtPlan& p) {
    return (size_t)p.maxM * sizeof(cplx<T>);
}

template <typename T, int BC>
static int launch_synth_s0(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase, hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    const int nrb = (p.npairs + 255) / 256, nchunk = (ncol + BC - 1) / BC;
    const unsigned nb = (unsigned)(nrb * nchunk * (p.mmax + 1));
    hipLaunchKernelGGL((k_leg_synth_s0<T, BC>), dim3(nb), dim3(256), 0, st, d, ncol, ld,
                       (const cplx<T>*)alm_t, (cplx<T>*)phase);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T, int BC>
static int launch_synth_sp(const ShtPlan& p, int nmap, int ld, const void* alm_t, void* phase, hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    const int nrb = (p.npairs + 255) / 256, nchunk = (nmap + BC - 1) / BC;
    const unsigned nb = (unsigned)(nrb * nchunk * (p.mmax + 1));
    hipLaunchKernelGGL((k_leg_synth_sp<T, BC>), dim3(nb), dim3(256), 0, st, d, nmap, ld,
                       (const cplx<T>*)alm_t, (cplx<T>*)phase);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int launch_ring_synth(const ShtPlan& p, int ncol, const void* phase, void* map, long map_stride,
                             hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    size_t sh = lds_bytes<T>(p);
    if (sh > 65536)
        PS_CHECK_HIP(hipFuncSetAttribute((const void*)k_ring_synth<T>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    hipLaunchKernelGGL(k_ring_synth<T>, dim3((unsigned)(p.nrings * ncol)), dim3(256), sh, st, d,
                       (const cplx<T>*)phase, (T*)map, map_stride);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int launch_ring_adjoint(const ShtPlan& p, int ncol, const void* map, long map_stride, void* phase,
                               hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    size_t sh = lds_bytes<T>(p);
    if (sh > 65536)
        PS_CHECK_HIP(hipFuncSetAttribute((const void*)k_ring_adjoint<T>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    hipLaunchKernelGGL(k_ring_adjoint<T>, dim3((unsigned)(p.nrings * ncol)), dim3(256), sh, st, d,
                       (const T*)map, map_stride, (cplx<T>*)phase);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T, int NRT>
static int launch_adj_nrt(const ShtPlan& p, int ncol, const void* phase, void* alm, long alm_stride,
                          double scale, hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    // LDS reduction buffer: LC * BC * (2 | 4) rows of 257 reals, kept under 33 KB
    if (p.spin == 0) {
        constexpr int BC = 2;
        constexpr int LC = sizeof(T) == 8 ? 4 : 8;
        const int nchunk = (ncol + BC - 1) / BC;
        hipLaunchKernelGGL((k_leg_adj_s0<T, NRT, BC, LC>), dim3((unsigned)(nchunk * (p.mmax + 1))), dim3(256), 0,
                           st, d, ncol, (const cplx<T>*)phase, (cplx<T>*)alm, alm_stride, (T)scale);
    } else {
        constexpr int BC = (NRT >= 8 || sizeof(T) == 8) ? 1 : 2;
        constexpr int LC = 4;
        const int nmap = ncol / 2;
        const int nchunk = (nmap + BC - 1) / BC;
        hipLaunchKernelGGL((k_leg_adj_sp<T, NRT, BC, LC>), dim3((unsigned)(nchunk * (p.mmax + 1))), dim3(256), 0,
                           st, d, nmap, (const cplx<T>*)phase, (cplx<T>*)alm, alm_stride, (T)scale);
    }
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int launch_adj(const ShtPlan& p, int ncol, const void* phase, void* alm, long alm_stride, double scale,
                      hipStream_t st) {
    const int nrt = (p.npairs + 255) / 256;
    if (nrt <= 1) return launch_adj_nrt<T, 1>(p, ncol, phase, alm, alm_stride, scale, st);
    if (nrt <= 2) return launch_adj_nrt<T, 2>(p, ncol, phase, alm, alm_stride, scale, st);
    if (nrt <= 4) return launch_adj_nrt<T, 4>(p, ncol, phase, alm, alm_stride, scale, st);
    if (nrt <= 8) return launch_adj_nrt<T, 8>(p, ncol, phase, alm, alm_stride, scale, st);
    set_error("adjoint: nside above 1024 is not supported");
    return -1;
}

// synthesis BC (batch columns per block)
static constexpr int SYN_BC_S0_F = 16, SYN_BC_S0_D = 8, SYN_BC_SP_F = 4, SYN_BC_SP_D = 2;

int sht_synth_lmajor_ld_multiple(const ShtPlan& p) {
    if (p.spin == 0) return p.prec == 4 ? SYN_BC_S0_F : SYN_BC_S0_D;
    return 2 * (p.prec == 4 ? SYN_BC_SP_F : SYN_BC_SP_D);
}

size_t sht_phase_bytes(const ShtPlan& p, int nmap) {
    return (size_t)nmap * p.ncomp * (p.mmax + 1) * p.nrings * 2 * p.prec;
}

// Column-chunk size of the synthesis kernels: the smallest instantiated chunk covering
// the batch (a single map uses BC = 1 instead of wasting the padding columns).
template <typename T>
static int launch_synth_s0_any(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase, hipStream_t st) {
    constexpr int BCMAX = sizeof(T) == 4 ? SYN_BC_S0_F : SYN_BC_S0_D;
    if (ncol <= 1) return launch_synth_s0<T, 1>(p, ncol, ld, alm_t, phase, st);
    if (ncol <= 2) return launch_synth_s0<T, 2>(p, ncol, ld, alm_t, phase, st);
    if (ncol <= 4) return launch_synth_s0<T, 4>(p, ncol, ld, alm_t, phase, st);
    if (BCMAX > 8 && ncol <= 8) return launch_synth_s0<T, 8>(p, ncol, ld, alm_t, phase, st);
    return launch_synth_s0<T, BCMAX>(p, ncol, ld, alm_t, phase, st);
}

template <typename T>
static int launch_synth_sp_any(const ShtPlan& p, int nmap, int ld, const void* alm_t, void* phase, hipStream_t st) {
    constexpr int BCMAX = sizeof(T) == 4 ? SYN_BC_SP_F : SYN_BC_SP_D;
    if (nmap <= 1) return launch_synth_sp<T, 1>(p, nmap, ld, alm_t, phase, st);
    if (BCMAX > 2 && nmap <= 2) return launch_synth_sp<T, 2>(p, nmap, ld, alm_t, phase, st);
    return launch_synth_sp<T, BCMAX>(p, nmap, ld, alm_t, phase, st);
}

int sht_synthesis_lmajor(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* map, long map_stride,
                         void* phase, hipStream_t st) {
    int rc = sht_legendre_synthesis_lmajor(p, nmap, alm_t, ld, phase, st);
    if (rc) return rc;
    return sht_ring_synthesis(p, nmap, phase, map, map_stride, st);
}

int sht_legendre_synthesis_lmajor(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* phase,
                                  hipStream_t st) {
    const int ncol = nmap * p.ncomp;
    if (p.prec == 4)
        return p.spin == 0 ? launch_synth_s0_any<float>(p, ncol, ld, alm_t, phase, st)
                           : launch_synth_sp_any<float>(p, nmap, ld, alm_t, phase, st);
    return p.spin == 0 ? launch_synth_s0_any<double>(p, ncol, ld, alm_t, phase, st)
                       : launch_synth_sp_any<double>(p, nmap, ld, alm_t, phase, st);
}

int sht_ring_synthesis(const ShtPlan& p, int nmap, const void* phase, void* map, long map_stride, hipStream_t st) {
    const int ncol = nmap * p.ncomp;
    if (p.prec == 4) return launch_ring_synth<float>(p, ncol, phase, map, map_stride, st);
    return launch_ring_synth<double>(p, ncol, phase, map, map_stride, st);
}

int sht_alm_to_lmajor(const ShtPlan& p, int ncol, const void* alm, long alm_stride, int ld, void* alm_t,
                      hipStream_t st) {
    long tot = p.nlm * (long)ld;
    unsigned nb = (unsigned)((tot + 255) / 256);
    if (p.prec == 4)
        hipLaunchKernelGGL(k_alm_to_lmajor<float>, dim3(nb), dim3(256), 0, st, (const cplx<float>*)alm,
                           alm_stride, ncol, p.nlm, ld, (cplx<float>*)alm_t);
    else
        hipLaunchKernelGGL(k_alm_to_lmajor<double>, dim3(nb), dim3(256), 0, st, (const cplx<double>*)alm,
                           alm_stride, ncol, p.nlm, ld, (cplx<double>*)alm_t);
    PS_LAUNCH_CHECK();
    return 0;
}

int sht_adjoint(const ShtPlan& p, int nmap, const void* map, long map_stride, void* alm, long alm_stride,
                double scale, void* phase, hipStream_t st) {
    const int ncol = nmap * p.ncomp;
    int rc;
    if (p.prec == 4) {
        rc = launch_ring_adjoint<float>(p, ncol, map, map_stride, phase, st);
        if (rc) return rc;
        return launch_adj<float>(p, ncol, phase, alm, alm_stride, scale, st);
    }
    rc = launch_ring_adjoint<double>(p, ncol, map, map_stride, phase, st);
    if (rc) return rc;
    return launch_adj<double>(p, ncol, phase, alm, alm_stride, scale, st);
}

}  // namespace ps
