// HEALPix spin-0 / spin-s spherical harmonic transforms for MI355X (gfx950).
//
// Replaces ducc0.sht.synthesis / adjoint_synthesis as PolySpec calls them
// (polyspec/base.py:147-151, 172-176, 193, 206, 225, 230-231, 253, 282).
//
// Two stages per direction:
//   Legendre stage  alm <-> ring phases F_m(ring) for m = 0..mmax.  The
//     normalised spin-weighted Legendre recursion runs in FP64 (seeds built from
//     log2 tables and carried as scaled value * 2^(64 k), so that sin^m(theta)
//     never underflows); the contraction with the batch columns runs on the
//     matrix cores (v_mfma_f32_32x32x2_f32) for FP32 batches, with alm rows and
//     recursion coefficients staged in LDS, or in registers on the vector ALUs
//     for narrow batches and FP64.
//   Ring stage      ring phases <-> pixels: m is folded modulo nphi, then a
//     self-sorting (Stockham) radix-16 FFT runs in LDS with the 16-point DFTs in
//     registers (nphi a power of two), or a Bluestein chirp-z transform over a
//     power-of-two length (the polar rings of length 4i).  Two real columns
//     share one complex transform in both directions.  The synthesis can stage
//     a block's phase rows in LDS before folding, and can add the gNL-loc pixel
//     integrand instead of writing the maps.
//
// Layouts: synthesis phases in tiles of 16 rings, [ring/16][mmax+1][ring%16][map*ncomp]
// complex, already rotated by exp(i m phi0(ring)); adjoint phases [map*ncomp][mmax+1][nrings]; maps
// [map][ncomp][npix]; alm (user) [map][ncomp][nlm]; alm (Legendre input) [nlm][ld]
// with the (map, comp) columns contiguous for each (l, m).
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
    std::map<int, int> chirp_off;  // nphi -> offset into the chirp table
    long chirp_total = 0;
    blue_total = 0;
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
        r.coff = -1;
        if (!pow2) {
            auto it = blue_off.find(r.nphi);
            if (it == blue_off.end()) {
                blue_off[r.nphi] = (int)blue_total;
                chirp_off[r.nphi] = (int)chirp_total;
                r.voff = blue_total;
                r.coff = chirp_total;
                blue_total += M;
                chirp_total += r.nphi;
            } else {
                r.voff = it->second;
                r.coff = chirp_off[r.nphi];
            }
        }
        maxM = std::max<long>(maxM, M);
    }
    // ring lists for the synthesis FFTs: power-of-two rings first, then Bluestein rings
    std::vector<int> rlist;
    maxM_pow2 = maxM_blue = 1;
    for (int i = 0; i < nrings; i++)
        if (rings[i].voff < 0) {
            rlist.push_back(i);
            maxM_pow2 = std::max<long>(maxM_pow2, rings[i].M);
        }
    n_pow2 = (int)rlist.size();
    for (int i = 0; i < nrings; i++)
        if (rings[i].voff >= 0) {
            rlist.push_back(i);
            maxM_blue = std::max<long>(maxM_blue, rings[i].M);
        }
    // twiddles e^{-2 pi i k / maxM}, k < maxM/2
    std::vector<double> tw((size_t)maxM);  // interleaved re, im
    for (long k = 0; k < maxM / 2; k++) {
        double s, c;
        sincos(-2.0 * M_PI * (double)k / (double)maxM, &s, &c);
        tw[2 * k] = c;
        tw[2 * k + 1] = s;
    }
    // Bluestein kernels V = FFT_M(v), v_t = exp(-i pi t^2 / n) wrapped; stored in
    // bit-reversed order (for the product taken between a DIF forward transform and a DIT
    // inverse transform), then again in natural order at offset blue_total (for the
    // self-sorting transforms).
    std::vector<double> blue((size_t)std::max<long>(1, 4 * blue_total));
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
            blue[2 * (blue_total + kv.second + k)] = v[k].real();
            blue[2 * (blue_total + kv.second + k) + 1] = v[k].imag();
        }
    }

    // Bluestein chirps exp(i pi t^2 / n), t < n, per distinct ring length (exact argument
    // reduction t^2 mod 2n, evaluated in FP64): the pre- and post-multiplications of the ring
    // FFTs read them instead of evaluating sincospi per element
    std::vector<double> chirp((size_t)std::max<long>(1, 2 * chirp_total));
    for (auto& kv : chirp_off) {
        const int n = kv.first;
        for (long t = 0; t < n; t++) {
            double s, c;
            sincos(M_PI * (double)((t * t) % (2L * n)) / (double)n, &s, &c);
            chirp[2 * (kv.second + t)] = c;
            chirp[2 * (kv.second + t) + 1] = s;
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
            // the -s recursion is carried as (-1)^s (-s)lambda: sY_{l,-m} = (-1)^(s+m) conj(-sY_lm),
            // so the m < 0 half of a real spin field pairs +s with (-1)^s (-s)
            if (g == 1 && (s & 1)) sg = -sg;
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
    if (upload(&d_rlist, rlist)) return -2;
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
        std::vector<float> cf(chirp.begin(), chirp.end());
        if (upload((float**)&d_chirp, cf)) return -2;
    } else {
        if (upload((double**)&d_tw, tw)) return -2;
        if (upload((double**)&d_blue, blue)) return -2;
        if (upload((double**)&d_chirp, chirp)) return -2;
    }
    return 0;
}

ShtPlan::~ShtPlan() {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    void* ptrs[] = {d_rings, d_l2s, d_l2sh, d_l2ch, d_seedL, d_seedS, d_seedA,
                    d_seedB, d_cth, d_tw, d_blue, d_coef, d_rlist, d_dead, d_chirp};
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
    const cplx<T>* blue;   // bit-reversed order; natural order from blue + blue_nat
    long blue_nat;
    const cplx<T>* chirp;  // exp(i pi t^2 / nphi) of the Bluestein rings, RingInfo::coff
    const double* coef;  // (alpha, beta, gamma, 0) per idx(l, m)
    const unsigned char* dead;  // (mmax+1) x npairs, see ShtPlan::d_dead
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
    d.dead = p.d_dead;
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
    d.blue_nat = p.blue_total;
    d.chirp = (const cplx<T>*)p.d_chirp;
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
    // ax = alpha * (x + beta), g = gamma; the scale exponent is updated by rec_rescale
    const double fn = ax * r.f - g * r.fm1;
    r.fm1 = r.f;
    r.f = fn;
}

// Scale-exponent update, called every few steps (at most 8) instead of after each one.
// Scaling by 2^-64 is exact, so the values f 2^(64 k) do not depend on when it happens;
// between calls f can grow past THR by the growth of those steps (below 2^6 per step),
// far from overflow since the true values stay O(1).  A lagging k is at most one below
// its per-step value, so accumulation tests use KMIN - 1.
__device__ __forceinline__ void rec_rescale(Rec& r) {
    const bool up = (r.k < 0) && (fmax(fabs(r.f), fabs(r.fm1)) > scale_consts<double>::THR);
    const int e = up ? -64 : 0;
    r.f = ldexp(r.f, e);
    r.fm1 = ldexp(r.fm1, e);
    r.k += up ? 1 : 0;
    // 2^(64 k) for k >= -2 (exact), 0 below: one ldexp and one select, no branches
    r.cf = r.k >= -2 ? ldexp(1.0, 64 * r.k) : 0.0;
}

__device__ __forceinline__ int m_of_q(int q, int mmax) {
    // interleave heavy (small m) and light (large m) blocks: 0, mmax, 1, mmax-1, ...
    return (q & 1) ? (mmax - (q >> 1)) : (q >> 1);
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

// exp(i pi num/den) for 0 <= num < 2^31
template <typename T>
__device__ __forceinline__ cplx<T> expi_pi32(int num, int den) {
    const int r = num % (2 * den);
    T s, c;
    sincospi_t((T)r / (T)den, &s, &c);
    return cplx<T>{c, s};
}

// Synthesis phase layout: tiles of PRT consecutive rings, [ring / PRT][m][ring % PRT][col].
// A wave of the Legendre stage (64 consecutive rings, one m) writes a few compact
// regions; the ring stage reads one ring's m rows PRT * ncol elements apart.
constexpr int PRT = 16;
__host__ __device__ __forceinline__ size_t syn_phase_off(int ring, int m, int mmax, int ncol) {
    return (((size_t)(ring / PRT) * (mmax + 1) + m) * PRT + (ring % PRT)) * (size_t)ncol;
}

// Recursion step without the rescale test: valid once k == 0, after which the
// normalised functions only grow to their O(1) oscillating regime.
__device__ __forceinline__ void rec_fast(Rec& r, double ax, double g) {
    const double fn = ax * r.f - g * r.fm1;
    r.fm1 = r.f;
    r.f = fn;
}

// l values per LDS-staged chunk of recursion coefficients and alm.  Even, so that the
// chunk-local index li has the parity of l - l_start.
constexpr int LCH = 32;

// -------------------------------------------------- Legendre synthesis, spin 0
// alm_t: [nlm][ld] complex, ld >= nchunk*BC (padding columns are read, never used).
// One block per (m, column chunk, 256 ring pairs); the block stages LCH rows of alm
// and the (alpha, gamma) coefficients in LDS, every lane reads them by broadcast.
template <typename T, int BC>
__global__ __launch_bounds__(256) void k_leg_synth_s0(ShtDev<T> d, int ncol, int ld,
                                                      const cplx<T>* __restrict__ alm_t,
                                                      cplx<T>* __restrict__ phase) {
    __shared__ cplx<T> sa[LCH * BC];
    __shared__ double sco[LCH * 2];
    const int nrb = (d.npairs + 255) / 256;
    const int nchunk = (ncol + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = lb % nrb;
    const int chunk = (lb / nrb) % nchunk;
    const int m = m_of_q(lb / (nrb * nchunk), d.mmax);
    const int tid = threadIdx.x;
    const int jp = rb * 256 + tid;
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
    if constexpr (BC >= 8) {
        // wide chunks: the alm row is read straight from memory by wave-uniform (scalar)
        // loads; staging it in LDS would cost every lane a full row of LDS reads per l
        int l = m;
        for (; l + 1 <= d.lmax; l += 2) {
            if (__any(r.k >= scale_consts<T>::KMIN - 1)) {
                const T v = (T)(r.f * r.cf);
                const cplx<T>* al = a + (size_t)(base + l) * ld;
#pragma unroll
                for (int j = 0; j < BC; j++) {
                    ev[j].re += v * al[j].re;
                    ev[j].im += v * al[j].im;
                }
            }
            rec_step(r, cf4[4 * l] * x, cf4[4 * l + 2]);
            if (__any(r.k >= scale_consts<T>::KMIN - 1)) {
                const T v = (T)(r.f * r.cf);
                const cplx<T>* al = a + (size_t)(base + l + 1) * ld;
#pragma unroll
                for (int j = 0; j < BC; j++) {
                    od[j].re += v * al[j].re;
                    od[j].im += v * al[j].im;
                }
            }
            rec_step(r, cf4[4 * (l + 1)] * x, cf4[4 * (l + 1) + 2]);
            if (r.k < 0) rec_rescale(r);
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
    } else {
        // chunk staging: thread t holds alm element t (row t / BC) and the (alpha, gamma) of
        // row t; the next chunk is fetched while the current one is processed
        static_assert(LCH * BC <= 256, "one alm element per thread per chunk");
        cplx<T> pa{T(0), T(0)};
        double pal = 0.0, pga = 0.0;
        auto fetch = [&](int l0n) {
            const int nn = min(LCH, d.lmax + 1 - l0n);
            if (tid < nn * BC) pa = a[(size_t)(base + l0n + tid / BC) * ld + tid % BC];
            if (tid < nn) {
                pal = cf4[4 * (l0n + tid)];
                pga = cf4[4 * (l0n + tid) + 2];
            }
        };
        fetch(m);
        bool fast = false;
        for (int l0 = m; l0 <= d.lmax; l0 += LCH) {
            const int n = min(LCH, d.lmax + 1 - l0);
            __syncthreads();
            if (tid < n * BC) sa[tid] = pa;
            if (tid < n) {
                sco[2 * tid] = pal;
                sco[2 * tid + 1] = pga;
            }
            __syncthreads();
            if (l0 + LCH <= d.lmax) fetch(l0 + LCH);
            if (!fast) fast = __all(r.k == 0);
            if (fast) {
                int li = 0;
                for (; li + 1 < n; li += 2) {
                    T v = (T)r.f;
    #pragma unroll
                    for (int j = 0; j < BC; j++) {
                        const cplx<T> al = sa[li * BC + j];
                        ev[j].re += v * al.re;
                        ev[j].im += v * al.im;
                    }
                    rec_fast(r, sco[2 * li] * x, sco[2 * li + 1]);
                    v = (T)r.f;
    #pragma unroll
                    for (int j = 0; j < BC; j++) {
                        const cplx<T> al = sa[(li + 1) * BC + j];
                        od[j].re += v * al.re;
                        od[j].im += v * al.im;
                    }
                    rec_fast(r, sco[2 * li + 2] * x, sco[2 * li + 3]);
                }
                if (li < n) {  // odd count: the last l of the transform, even parity
                    const T v = (T)r.f;
    #pragma unroll
                    for (int j = 0; j < BC; j++) {
                        const cplx<T> al = sa[li * BC + j];
                        ev[j].re += v * al.re;
                        ev[j].im += v * al.im;
                    }
                }
            } else {
                // groups of 4 values of l: the parity of l and the exponent update are fixed per
                // unrolled step (no per-l branches); a partial last group is not rescaled
                for (int li0 = 0; li0 < n; li0 += 4) {
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const int li = li0 + c;
                        if (li < n) {
                            if (__any(r.k >= scale_consts<T>::KMIN - 1)) {
                                const T v = (T)(r.f * r.cf);
#pragma unroll
                                for (int j = 0; j < BC; j++) {
                                    const cplx<T> al = sa[li * BC + j];
                                    cplx<T>& acc = (c & 1) ? od[j] : ev[j];
                                    acc.re += v * al.re;
                                    acc.im += v * al.im;
                                }
                            }
                            rec_step(r, sco[2 * li] * x, sco[2 * li + 1]);
                        }
                    }
                    if (li0 + 3 < n) rec_rescale(r);
                }
            }
        }
    }
    if (jp >= d.npairs) return;
    const int rn = jp, rs = d.nrings - 1 - jp;
    const RingInfo ri = d.rings[rn];
    const cplx<T> rot = expi_pi<T>((long)m * ri.phi_p, ri.phi_q);  // exp(i m phi0), same for both rings
    cplx<T>* phn = phase + syn_phase_off(rn, m, d.mmax, ncol) + chunk * BC;
    cplx<T>* phs = phase + syn_phase_off(rs, m, d.mmax, ncol) + chunk * BC;
#pragma unroll
    for (int j = 0; j < BC; j++) {
        if (chunk * BC + j < ncol) {
            phn[j] = cmul(cadd(ev[j], od[j]), rot);
            if (rs != rn) phs[j] = cmul(csub(ev[j], od[j]), rot);
        }
    }
}

// ------------------------------------------------------------ dead (m, ring) pairs
// One thread per (m, ring pair): runs the recursion(s) of the synthesis kernels and records
// whether the true magnitude |f| 2^(64 k) stays below 2^-200 for every l.  Such values are
// exactly zero after the FP32 scaling (k <= -3) and below 1e-60 in FP64, so a block whose
// ring pairs are all dead at its m may skip its l loop.
__global__ void k_mark_dead(ShtDev<double> d, unsigned char* __restrict__ dead) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)(d.mmax + 1) * d.npairs) return;
    const int m = (int)(idx / d.npairs), jp = (int)(idx % d.npairs);
    const int M1 = d.mmax + 1;
    const double x = d.cth[jp];
    Rec rp, rm;
    rec_seed(rp, d.seedL[m] + d.seedA[m] * d.l2sh[jp] + d.seedB[m] * d.l2ch[jp], d.seedS[m]);
    const bool sp = d.spin != 0;
    if (sp)
        rec_seed(rm, d.seedL[M1 + m] + d.seedA[M1 + m] * d.l2sh[jp] + d.seedB[M1 + m] * d.l2ch[jp], d.seedS[M1 + m]);
    else
        rm = rp;
    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const int lstart = m > d.spin ? m : d.spin;
    bool alive = false;
    for (int l = lstart; l <= d.lmax && !alive; l++) {
        const double lg = fmax(log2(fabs(rp.f)) + 64.0 * rp.k, log2(fabs(rm.f)) + 64.0 * rm.k);
        alive = lg >= -200.0;
        const double* c = d.coef + 4 * (base + l);
        rec_step(rp, c[0] * (x + c[1]), c[2]);
        if (sp) rec_step(rm, c[0] * (x - c[1]), c[2]);
        else rm = rp;
        if (((l - lstart) & 3) == 3) {
            rec_rescale(rp);
            if (sp) rec_rescale(rm);
            else rm = rp;
        }
    }
    dead[idx] = alive ? 0 : 1;
}

int sht_mark_dead(ShtPlan& p) {
    const size_t n = (size_t)(p.mmax + 1) * p.npairs;
    PS_CHECK_HIP(hipMalloc((void**)&p.d_dead, n));
    ShtDev<double> d = dev_view<double>(p);
    hipLaunchKernelGGL(k_mark_dead, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, d, p.d_dead);
    PS_LAUNCH_CHECK();
    PS_CHECK_HIP(hipStreamSynchronize(0));
    return 0;
}

// ------------------------------------- Legendre synthesis, spin 0, FP32 matrix cores
// The contraction F_m(ring, col) = sum_l lam_lm(ring) a_lm(col) of one m for 32 ring pairs
// and 16 complex (32 real) columns runs on v_mfma_f32_32x32x2f32: A = lam (32 rings x 2
// values of l), B = alm (2 values of l x 32 real columns), D += A B.  Even and odd l - m
// accumulate separately (north = even + odd, south = even - odd), so one MFMA takes
// l and l + 2.  Lane layout (A, B: one value per lane, D: 16 per lane):
//   A[i][k] at lane i + 32 k;  B[k][j] at lane j + 32 k;
//   D[i][j] at lane j + 32 ((i / 4) % 2), register 4 (i / 8) + i % 4.
// Lanes i and i + 32 run the same FP64 recursion (ring i) and feed l, l + 1 (k = 0) and
// l + 2, l + 3 (k = 1).  One block = 4 waves = 128 ring pairs, one m, NT tiles of 16 complex
// columns: every recursion value feeds 2 NT MFMAs, so the FP64 recursion is shared by up to
// 80 columns.  The chunk's alm rows are staged in LDS.
typedef float v16f __attribute__((ext_vector_type(16)));

template <int NT>
__global__ __launch_bounds__(256) void k_leg_synth_s0_mfma(ShtDev<float> d, int ncol, int ld,
                                                           const cplx<float>* __restrict__ alm_t,
                                                           cplx<float>* __restrict__ phase, int rb0, int nrbs) {
    constexpr int BCM = 16 * NT;   // complex columns per block
    constexpr int RW = 2 * BCM;    // floats per staged row
    __shared__ __align__(16) float sa[LCH * RW];
    __shared__ double sco[LCH * 2];
    __shared__ cplx<float> rot[128];
    // ring-pair blocks rb0 .. rb0 + nrbs - 1 of 128 pairs each
    const int nchunk = (ncol + BCM - 1) / BCM;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = rb0 + (int)(lb % nrbs);
    const int chunk = (lb / nrbs) % nchunk;
    const int m = m_of_q(lb / (nrbs * nchunk), d.mmax);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, li32 = lane & 31;
    const int jp = rb * 128 + wave * 32 + li32;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    if (tid < 128) {
        const int jr = min(rb * 128 + tid, d.npairs - 1);
        const RingInfo ri = d.rings[jr];
        rot[tid] = expi_pi<float>((long)m * ri.phi_p, ri.phi_q);  // exp(i m phi0)
    }

    const double x = d.cth[jpc];
    Rec r;
    rec_seed(r, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
    v16f Dev[NT], Dod[NT];
#pragma unroll
    for (int c = 0; c < NT; c++) Dev[c] = Dod[c] = v16f{0};

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const float* a = reinterpret_cast<const float*>(alm_t + (size_t)chunk * BCM);
    const double* cf4 = d.coef + 4 * base;
    // chunk staging: float4 e of thread t is q = t + 256 e = (row q / (2 NT * 4), float4 q % (8 NT));
    // columns at or past ld read as zero; the next chunk is fetched while the current one runs
    static_assert(LCH * RW == 4 * 256 * NT, "NT float4 per thread per chunk");
    float4 pa[NT];
    double pal = 0.0, pga = 0.0;
    auto fetch = [&](int l0n) {
        const int nn = min(LCH, d.lmax + 1 - l0n);
#pragma unroll
        for (int e = 0; e < NT; e++) {
            const int q = tid + 256 * e, li = q / (8 * NT), c4 = q % (8 * NT);
            const bool ok = li < nn && chunk * BCM + 2 * c4 < ld;
            pa[e] = ok ? *reinterpret_cast<const float4*>(a + (size_t)(base + l0n + li) * 2 * ld + 4 * c4)
                       : float4{0.f, 0.f, 0.f, 0.f};
        }
        if (tid < nn) {
            pal = cf4[4 * (l0n + tid)];
            pga = cf4[4 * (l0n + tid) + 2];
        }
    };
    // a block whose ring pairs are all dead at this m writes zero phases without the l loop
    const bool skip = __syncthreads_and(d.dead[(size_t)m * d.npairs + jpc]);
    // stage the chunk at l0s (alm rows and coefficients into LDS) and fetch the next one
    auto stage = [&](int l0s, int ns) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < NT; e++) *reinterpret_cast<float4*>(sa + 4 * (tid + 256 * e)) = pa[e];
        if (tid < ns) {
            sco[2 * tid] = pal;
            sco[2 * tid + 1] = pga;
        }
        __syncthreads();
        if (l0s + LCH <= d.lmax) fetch(l0s + LCH);
    };
    // one group of 4 values of l: two MFMAs per column tile (even and odd l - m)
    auto mma = [&](int g, float aev, float aod) {
#pragma unroll
        for (int c = 0; c < NT; c++) {
            const float bev = sa[(g + 2 * h) * RW + 32 * c + li32];
            const float bod = sa[(g + 1 + 2 * h) * RW + 32 * c + li32];
            Dev[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(aev, bev, Dev[c], 0, 0, 0);
            Dod[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(aod, bod, Dod[c], 0, 0, 0);
        }
    };
    // The l loop runs in three phases, so that each inner loop is a loop of its own (the
    // accumulators then stay in the same registers from chunk to chunk): chunks while some
    // lane still carries a scale exponent, full chunks once every lane is scaled in, and the
    // last partial chunk.  Every wave stages every chunk once, so the block's barriers match
    // even when its waves switch phase at different chunks.
    int l0 = m;
    bool staged = false;
    if (!skip) {
        fetch(m);
        for (; l0 <= d.lmax; l0 += LCH) {
            const int n = min(LCH, d.lmax + 1 - l0);
            stage(l0, n);
            if (__all(r.k == 0)) {
                staged = true;  // this chunk and the rest run scaled in
                break;
            }
            if (n == LCH) {
                // straight-line groups, one exponent update per group
#pragma unroll 2
                for (int g = 0; g < LCH; g += 4) {
                    double v[4];
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        v[c] = r.f * r.cf;
                        rec_step(r, sco[2 * (g + c)] * x, sco[2 * (g + c) + 1]);
                    }
                    rec_rescale(r);
                    if (!__any(r.k >= scale_consts<float>::KMIN - 1)) continue;  // all values below FP32 range
                    mma(g, (float)(h ? v[2] : v[0]), (float)(h ? v[3] : v[1]));
                }
            } else {
                for (int g = 0; g < n; g += 4) {
                    // recursion values at l0+g .. l0+g+3 (zero past the end of the chunk)
                    double v[4];
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const int li = g + c;
                        if (li < n) {
                            v[c] = r.f * r.cf;
                            rec_step(r, sco[2 * li] * x, sco[2 * li + 1]);
                        } else {
                            v[c] = 0.0;
                        }
                    }
                    rec_rescale(r);
                    if (!__any(r.k >= scale_consts<float>::KMIN - 1)) continue;
                    const int le = g + 2 * h, lo = g + 1 + 2 * h;
#pragma unroll
                    for (int c = 0; c < NT; c++) {
                        const float bev = le < n ? sa[le * RW + 32 * c + li32] : 0.f;
                        const float bod = lo < n ? sa[lo * RW + 32 * c + li32] : 0.f;
                        Dev[c] = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? v[2] : v[0]), bev, Dev[c], 0, 0, 0);
                        Dod[c] = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? v[3] : v[1]), bod, Dod[c], 0, 0, 0);
                    }
                }
            }
        }
        // full chunks, every scale exponent zero: straight-line groups
        for (; l0 + LCH <= d.lmax + 1; l0 += LCH) {
            if (!staged) stage(l0, LCH);
            staged = false;
#pragma unroll 2
            for (int g = 0; g < LCH; g += 4) {
                const double v0 = r.f;
                rec_fast(r, sco[2 * g] * x, sco[2 * g + 1]);
                const double v1 = r.f;
                rec_fast(r, sco[2 * g + 2] * x, sco[2 * g + 3]);
                const double v2 = r.f;
                rec_fast(r, sco[2 * g + 4] * x, sco[2 * g + 5]);
                const double v3 = r.f;
                rec_fast(r, sco[2 * g + 6] * x, sco[2 * g + 7]);
                mma(g, (float)(h ? v2 : v0), (float)(h ? v3 : v1));
            }
        }
        // the last partial chunk, scaled in
        if (l0 <= d.lmax) {
            const int n = d.lmax + 1 - l0;
            if (!staged) stage(l0, n);
            for (int g = 0; g < n; g += 4) {
                double v[4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int li = g + c;
                    if (li < n) {
                        v[c] = r.f;
                        rec_fast(r, sco[2 * li] * x, sco[2 * li + 1]);
                    } else {
                        v[c] = 0.0;
                    }
                }
                const int le = g + 2 * h, lo = g + 1 + 2 * h;
#pragma unroll
                for (int c = 0; c < NT; c++) {
                    const float bev = le < n ? sa[le * RW + 32 * c + li32] : 0.f;
                    const float bod = lo < n ? sa[lo * RW + 32 * c + li32] : 0.f;
                    Dev[c] = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? v[2] : v[0]), bev, Dev[c], 0, 0, 0);
                    Dod[c] = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? v[3] : v[1]), bod, Dod[c], 0, 0, 0);
                }
            }
        }
    }
    // D[i][j]: ring i of the wave, real column j = lane % 32 of tile c (complex column j / 2)
    const bool re = (li32 & 1) == 0;
#pragma unroll
    for (int c = 0; c < NT; c++) {
        const int col = chunk * BCM + 16 * c + (li32 >> 1);
#pragma unroll
        for (int vv = 0; vv < 16; vv++) {
            const int i = 8 * (vv >> 2) + 4 * h + (vv & 3);
            const int jr = rb * 128 + wave * 32 + i;
            const float e = Dev[c][vv], o = Dod[c][vv];
            float nv = e + o, sv = e - o;
            const float nv2 = __shfl_xor(nv, 1), sv2 = __shfl_xor(sv, 1);  // the other part
            const cplx<float> w = rot[wave * 32 + i];
            // (re + i im)(w.re + i w.im)
            const float no = re ? nv * w.re - nv2 * w.im : nv2 * w.im + nv * w.re;
            const float so = re ? sv * w.re - sv2 * w.im : sv2 * w.im + sv * w.re;
            if (jr < d.npairs && col < ncol) {
                const int rn = jr, rs = d.nrings - 1 - jr;
                reinterpret_cast<float*>(phase + syn_phase_off(rn, m, d.mmax, ncol) + col)[re ? 0 : 1] = no;
                if (rs != rn)
                    reinterpret_cast<float*>(phase + syn_phase_off(rs, m, d.mmax, ncol) + col)[re ? 0 : 1] = so;
            }
        }
    }
}

// ------------------- Legendre synthesis, spin 0, FP32 matrix cores, recursion per wave in registers
// As k_leg_synth_s0_mfma (each wave runs the FP64 recursion of its 32 ring pairs in both lane
// halves and feeds the matrix cores from registers), but one block covers NT column tiles (all
// 80 columns of the Nside-512 numerator at NT = 5), so each recursion value feeds 2 NT MFMAs, and
// the l loop is one compact loop over double-buffered LDS chunks: the alm rows and the recursion
// coefficients of the next chunk are copied global -> LDS by global_load_lds_dwordx4 (16 bytes
// per lane, no staging registers) while the current chunk runs; one barrier per chunk.  Rows past
// the end of the transform and columns past ld read clamped, finite entries: such rows only ever
// meet zero lambda values and such columns land in outputs that are not stored.
typedef __attribute__((address_space(3))) void ps_lds_void;

template <int NT>
__global__ __launch_bounds__(256, 2) void k_leg_synth_s0_wr(ShtDev<float> d, int ncol, int ld,
                                                           const cplx<float>* __restrict__ alm_t,
                                                           cplx<float>* __restrict__ phase, int rb0, int nrbs) {
    constexpr int BCM = 16 * NT;
    constexpr int RW = 2 * BCM;
    // double buffers as distinct LDS objects: the compiler then sees that the asynchronous copies
    // into one buffer cannot alias the reads of the other, and does not wait for the copies
    // before each read (with one [2][...] array it waited at every group of four l)
    __shared__ __align__(16) float sa0[LCH * RW], sa1[LCH * RW];
    __shared__ __align__(16) double sco0[LCH * 4], sco1[LCH * 4];  // (alpha, beta, gamma, 0) per l
    __shared__ cplx<float> rot[128];
    const int nchunk = (ncol + BCM - 1) / BCM;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = rb0 + (int)(lb % nrbs);
    const int chunk = (lb / nrbs) % nchunk;
    const int m = __builtin_amdgcn_readfirstlane(m_of_q(lb / (nrbs * nchunk), d.mmax));
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int h = lane >> 5, li32 = lane & 31;
    const int jp = rb * 128 + wave * 32 + li32;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    if (tid < 128) {
        const int jr = min(rb * 128 + tid, d.npairs - 1);
        const RingInfo ri = d.rings[jr];
        rot[tid] = expi_pi<float>((long)m * ri.phi_p, ri.phi_q);
    }
    v16f Dev[NT], Dod[NT];
#pragma unroll
    for (int c = 0; c < NT; c++) Dev[c] = Dod[c] = v16f{0};
    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const float* a = reinterpret_cast<const float*>(alm_t + (size_t)chunk * BCM);
    const int skip = __builtin_amdgcn_readfirstlane(__syncthreads_and(d.dead[(size_t)m * d.npairs + jpc]));
    if (!skip) {
        const double x = d.cth[jpc];
        Rec r;
        rec_seed(r, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
        const int nl = d.lmax + 1 - m;
        const int nch = (nl + LCH - 1) / LCH;
        static_assert(LCH * RW == 4 * 256 * NT, "NT 16-byte copies per thread per chunk");
        static_assert(LCH * 4 * 8 == 64 * 16, "one wave copies a chunk's coefficients");
        // chunk c -> buffers (dsa, dsco): thread tid copies the 16-byte pieces q = tid + 256 e of the rows
        auto stage = [&](int c, float* dsa, double* dsco) {
            const int l0n = m + c * LCH, nn = min(LCH, nl - c * LCH);
#pragma unroll
            for (int e = 0; e < NT; e++) {
                const int q = tid + 256 * e, li = min(q / (8 * NT), nn - 1), c4 = q % (8 * NT);
                const int col = min(chunk * BCM + 2 * c4, ld - 2);  // ld is even
                const float* src = reinterpret_cast<const float*>(alm_t) + ((size_t)(base + l0n + li) * ld + col) * 2;
                __builtin_amdgcn_global_load_lds((const void*)src, (ps_lds_void*)&dsa[4 * (256 * e + 64 * wave)], 16, 0,
                                                 0);
            }
            if (wave == 0) {
                const int l = min(l0n + (lane >> 1), d.lmax);
                const char* src = reinterpret_cast<const char*>(d.coef + 4 * (base + l)) + 16 * (lane & 1);
                __builtin_amdgcn_global_load_lds((const void*)src, (ps_lds_void*)dsco, 16, 0, 0);
            }
        };
        stage(0, sa0, sco0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        bool fast = __all(r.k == 0);
        // one chunk from (sab, co); the copies of the next chunk go to the other buffers
        auto run_chunk = [&](int c, const float* __restrict__ sab, const double* __restrict__ co) {
            const int n = __builtin_amdgcn_readfirstlane(min(LCH, nl - c * LCH));
            for (int g = 0; g < n; g += 4) {
                // the B operands first: their LDS latency then passes under the recursion
                float bev[NT], bod[NT];
#pragma unroll
                for (int cc = 0; cc < NT; cc++) {
                    bev[cc] = sab[(g + 2 * h) * RW + 32 * cc + li32];
                    bod[cc] = sab[(g + 1 + 2 * h) * RW + 32 * cc + li32];
                }
                double v[4];
                bool live = true;
                if (g + 4 <= n) {  // a full group: four steps, no tail tests
                    double ca[4], cg[4];
#pragma unroll
                    for (int cc = 0; cc < 4; cc++) {
                        ca[cc] = co[4 * (g + cc)] * x;
                        cg[cc] = co[4 * (g + cc) + 2];
                    }
                    if (fast) {
#pragma unroll
                        for (int cc = 0; cc < 4; cc++) {
                            v[cc] = r.f;
                            rec_fast(r, ca[cc], cg[cc]);
                        }
                    } else {
#pragma unroll
                        for (int cc = 0; cc < 4; cc++) {
                            v[cc] = r.f * r.cf;
                            rec_step(r, ca[cc], cg[cc]);
                        }
                    }
                } else {  // the last, partial group of the transform
#pragma unroll
                    for (int cc = 0; cc < 4; cc++) {
                        v[cc] = g + cc < n ? r.f * r.cf : 0.0;
                        if (g + cc < n) rec_step(r, co[4 * (g + cc)] * x, co[4 * (g + cc) + 2]);
                    }
                }
                if (!fast) {
                    rec_rescale(r);
                    fast = __all(r.k == 0);
                    live = __any(r.k >= scale_consts<float>::KMIN - 1);  // all below FP32 range: skip
                }
                if (!live) continue;
                const float aev = (float)(h ? v[2] : v[0]), aod = (float)(h ? v[3] : v[1]);
#pragma unroll
                for (int cc = 0; cc < NT; cc++) {
                    Dev[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(aev, bev[cc], Dev[cc], 0, 0, 0);
                    Dod[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(aod, bod[cc], Dod[cc], 0, 0, 0);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);  // the next chunk's copies have landed in LDS
            __syncthreads();
        };
        for (int c = 0; c < nch; c += 2) {
            if (c + 1 < nch) stage(c + 1, sa1, sco1);
            run_chunk(c, sa0, sco0);
            if (c + 1 < nch) {
                if (c + 2 < nch) stage(c + 2, sa0, sco0);
                run_chunk(c + 1, sa1, sco1);
            }
        }
    }
    const bool re = (li32 & 1) == 0;
#pragma unroll
    for (int c = 0; c < NT; c++) {
        const int col = chunk * BCM + 16 * c + (li32 >> 1);
#pragma unroll
        for (int vv = 0; vv < 16; vv++) {
            const int i = 8 * (vv >> 2) + 4 * h + (vv & 3);
            const int jr = rb * 128 + wave * 32 + i;
            const float e = Dev[c][vv], o = Dod[c][vv];
            float nv = e + o, sv = e - o;
            const float nv2 = __shfl_xor(nv, 1), sv2 = __shfl_xor(sv, 1);
            const cplx<float> w = rot[wave * 32 + i];
            const float no = re ? nv * w.re - nv2 * w.im : nv2 * w.im + nv * w.re;
            const float so = re ? sv * w.re - sv2 * w.im : sv2 * w.im + sv * w.re;
            if (jr < d.npairs && col < ncol) {
                const int rn = jr, rs = d.nrings - 1 - jr;
                reinterpret_cast<float*>(phase + syn_phase_off(rn, m, d.mmax, ncol) + col)[re ? 0 : 1] = no;
                if (rs != rn)
                    reinterpret_cast<float*>(phase + syn_phase_off(rs, m, d.mmax, ncol) + col)[re ? 0 : 1] = so;
            }
        }
    }
}

// ------------------------------------- Legendre synthesis, spin s, FP32 matrix cores
// The tiling of the spin-0 kernel with two recursions per ring (+s and (-1)^s (-s)):
// A = lam+ = (f+ + f-)/2 or lam- = (f+ - f-)/2, B = the alm rows of 8 maps (16 complex
// columns, map b and comp c (G = 0, C = 1) at 2 b + c, i.e. real column j = 4 b + 2 c + part).
// Four accumulators: lam+ / lam- times even / odd l + m + s.  With q = j % 4 the output at
// lane j of map b is
//   q = 0: Re R = -(G_re lam+ - C_im lam-)     q = 1: Im R = -(G_im lam+ + C_re lam-)
//   q = 2: Re I = -(C_re lam+ + G_im lam-)     q = 3: Im I = -(C_im lam+ - G_re lam-)
// i.e. -(D+[j] - D-[j ^ 3]) for q = 0, 3 and -(D+[j] + D-[j ^ 3]) for q = 1, 2; the south
// ring takes lam+ -> (-1)^(l+m+s) lam+, lam- -> -(-1)^(l+m+s) lam-.  R and I land in the
// complex columns 2 b and 2 b + 1 of the phase buffer, the layout of the FP64 / few-map kernel.
template <bool GONLY>
__global__ __launch_bounds__(256) void k_leg_synth_sp_mfma(ShtDev<float> d, int ncol, int ld,
                                                           const cplx<float>* __restrict__ alm_t,
                                                           cplx<float>* __restrict__ phase) {
    constexpr int BCM = 16;
    __shared__ __align__(16) float sa[LCH * 2 * BCM];
    __shared__ double sco[LCH * 3];
    __shared__ cplx<float> rot[128];
    const int nrb = (d.npairs + 127) / 128;
    const int nchunk = (ncol + BCM - 1) / BCM;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = lb % nrb;
    const int chunk = (lb / nrb) % nchunk;
    const int m = m_of_q(lb / (nrb * nchunk), d.mmax);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, li32 = lane & 31;
    const int jp = rb * 128 + wave * 32 + li32;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    const int M1 = d.mmax + 1;
    if (tid < 128) {
        const int jr = min(rb * 128 + tid, d.npairs - 1);
        const RingInfo ri = d.rings[jr];
        rot[tid] = expi_pi<float>((long)m * ri.phi_p, ri.phi_q);  // exp(i m phi0)
    }

    const double x = d.cth[jpc];
    Rec rp, rm;
    rec_seed(rp, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
    rec_seed(rm, d.seedL[M1 + m] + d.seedA[M1 + m] * d.l2sh[jpc] + d.seedB[M1 + m] * d.l2ch[jpc],
             d.seedS[M1 + m]);
    v16f Pev = {0}, Pod = {0}, Mev = {0}, Mod = {0};

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const float* a = reinterpret_cast<const float*>(alm_t + (size_t)chunk * BCM);
    const double* cf4 = d.coef + 4 * base;
    const int lstart = m > d.spin ? m : d.spin;
    // chunk staging as in the spin-0 kernel; columns at or past ld (the last chunk of a
    // batch whose ld is not a multiple of 16) read as zero
    static_assert(LCH * 2 * BCM == 4 * 256, "one float4 per thread per chunk");
    const bool col_ok = chunk * BCM + 2 * (tid & 7) < ld;
    float4 pa;
    double pal = 0.0, pbe = 0.0, pga = 0.0;
    auto fetch = [&](int l0n) {
        const int nn = min(LCH, d.lmax + 1 - l0n);
        const int li = tid >> 3;
        pa = (li < nn && col_ok)
                 ? *reinterpret_cast<const float4*>(a + (size_t)(base + l0n + li) * 2 * ld + 4 * (tid & 7))
                 : float4{0.f, 0.f, 0.f, 0.f};
        if (tid < nn) {
            pal = cf4[4 * (l0n + tid)];
            pbe = cf4[4 * (l0n + tid) + 1];
            pga = cf4[4 * (l0n + tid) + 2];
        }
    };
    // a block whose ring pairs are all dead at this m writes zero phases without the l loop
    const bool skip = __syncthreads_and(d.dead[(size_t)m * d.npairs + jpc]);
    if (!skip) fetch(lstart);
    bool fast = false;
    for (int l0 = lstart; l0 <= d.lmax && !skip; l0 += LCH) {
        const int n = min(LCH, d.lmax + 1 - l0);
        __syncthreads();
        *reinterpret_cast<float4*>(sa + 4 * tid) = pa;
        if (tid < n) {
            sco[3 * tid] = pal;
            sco[3 * tid + 1] = pbe;
            sco[3 * tid + 2] = pga;
        }
        __syncthreads();
        if (l0 + LCH <= d.lmax) fetch(l0 + LCH);
        if (!fast) fast = __all(rp.k == 0 && rm.k == 0);
        if (fast && n == LCH) {
            // every scale exponent is zero and the chunk is full: straight-line groups
#pragma unroll 2
            for (int g = 0; g < LCH; g += 4) {
                double vp[4], vm[4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    vp[c] = 0.5 * (rp.f + rm.f);
                    vm[c] = 0.5 * (rp.f - rm.f);
                    const double al = sco[3 * (g + c)], be = sco[3 * (g + c) + 1], ga = sco[3 * (g + c) + 2];
                    rec_fast(rp, al * (x + be), ga);
                    rec_fast(rm, al * (x - be), ga);
                }
                const float bev = sa[(g + 2 * h) * 2 * BCM + li32];
                const float bod = sa[(g + 1 + 2 * h) * 2 * BCM + li32];
                Pev = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vp[2] : vp[0]), bev, Pev, 0, 0, 0);
                Pod = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vp[3] : vp[1]), bod, Pod, 0, 0, 0);
                Mev = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vm[2] : vm[0]), bev, Mev, 0, 0, 0);
                Mod = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vm[3] : vm[1]), bod, Mod, 0, 0, 0);
            }
            continue;
        }
        if (!fast && n == LCH) {
            // full chunk while some lane still carries a scale exponent: straight-line groups,
            // one exponent update per group (the general loop below, without its bounds tests)
#pragma unroll 2
            for (int g = 0; g < LCH; g += 4) {
                double vp[4], vm[4];
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const double fp = rp.f * rp.cf, fm = rm.f * rm.cf;
                    vp[c] = 0.5 * (fp + fm);
                    vm[c] = 0.5 * (fp - fm);
                    const double al = sco[3 * (g + c)], be = sco[3 * (g + c) + 1], ga = sco[3 * (g + c) + 2];
                    rec_step(rp, al * (x + be), ga);
                    rec_step(rm, al * (x - be), ga);
                }
                rec_rescale(rp);
                rec_rescale(rm);
                if (!__any((rp.k >= scale_consts<float>::KMIN - 1) || (rm.k >= scale_consts<float>::KMIN - 1)))
                    continue;  // all values below FP32 range
                const float bev = sa[(g + 2 * h) * 2 * BCM + li32];
                const float bod = sa[(g + 1 + 2 * h) * 2 * BCM + li32];
                Pev = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vp[2] : vp[0]), bev, Pev, 0, 0, 0);
                Pod = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vp[3] : vp[1]), bod, Pod, 0, 0, 0);
                Mev = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vm[2] : vm[0]), bev, Mev, 0, 0, 0);
                Mod = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vm[3] : vm[1]), bod, Mod, 0, 0, 0);
            }
            continue;
        }
        for (int g = 0; g < n; g += 4) {
            // recursion values at l0+g .. l0+g+3 (zero past the end of the chunk)
            double vp[4], vm[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int li = g + c;
                if (li < n) {
                    const double fp = fast ? rp.f : rp.f * rp.cf, fm = fast ? rm.f : rm.f * rm.cf;
                    vp[c] = 0.5 * (fp + fm);
                    vm[c] = 0.5 * (fp - fm);
                    const double al = sco[3 * li], be = sco[3 * li + 1], ga = sco[3 * li + 2];
                    if (fast) {
                        rec_fast(rp, al * (x + be), ga);
                        rec_fast(rm, al * (x - be), ga);
                    } else {
                        rec_step(rp, al * (x + be), ga);
                        rec_step(rm, al * (x - be), ga);
                    }
                } else {
                    vp[c] = vm[c] = 0.0;
                }
            }
            if (!fast) {
                rec_rescale(rp);
                rec_rescale(rm);
                // all values below FP32 range
                if (!__any((rp.k >= scale_consts<float>::KMIN - 1) || (rm.k >= scale_consts<float>::KMIN - 1)))
                    continue;
            }
            const int le = g + 2 * h, lo = g + 1 + 2 * h;
            const float bev = le < n ? sa[le * 2 * BCM + li32] : 0.f;
            const float bod = lo < n ? sa[lo * 2 * BCM + li32] : 0.f;
            Pev = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vp[2] : vp[0]), bev, Pev, 0, 0, 0);
            Pod = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vp[3] : vp[1]), bod, Pod, 0, 0, 0);
            Mev = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vm[2] : vm[0]), bev, Mev, 0, 0, 0);
            Mod = __builtin_amdgcn_mfma_f32_32x32x2f32((float)(h ? vm[3] : vm[1]), bod, Mod, 0, 0, 0);
        }
    }
    // ev accumulators hold l - l_start even, i.e. l + m + s even unless l_start + m + s is odd
    const bool swp = ((lstart + m + d.spin) & 1) != 0;
    const bool re = (li32 & 1) == 0;
    if constexpr (GONLY) {
        // C = 0 and the tile's 32 real columns are the G of 16 maps (map chunk*16 + li32/2, part
        // li32 % 2): Re R = -D+[G re], Im R = -D+[G im], Re I = -D-[G im], Im I = +D-[G re],
        // the lam- terms taken from the partner lane (the other part of the same map)
        const int map = chunk * BCM + (li32 >> 1);
        const int ncp = 2 * ncol;  // phase columns: R and I of every map
        const float sgi = re ? -1.f : 1.f;
#pragma unroll
        for (int vv = 0; vv < 16; vv++) {
            const int i = 8 * (vv >> 2) + 4 * h + (vv & 3);
            const int jr = rb * 128 + wave * 32 + i;
            const float pe = swp ? Pod[vv] : Pev[vv], po = swp ? Pev[vv] : Pod[vv];
            const float me = swp ? Mod[vv] : Mev[vv], mo = swp ? Mev[vv] : Mod[vv];
            const float nR = -(pe + po), sR = -(pe - po);
            const float nI = sgi * __shfl_xor(me + mo, 1), sI = sgi * __shfl_xor(mo - me, 1);
            const float nR2 = __shfl_xor(nR, 1), sR2 = __shfl_xor(sR, 1);  // the other part
            const float nI2 = __shfl_xor(nI, 1), sI2 = __shfl_xor(sI, 1);
            const cplx<float> w = rot[wave * 32 + i];
            const float noR = re ? nR * w.re - nR2 * w.im : nR2 * w.im + nR * w.re;
            const float soR = re ? sR * w.re - sR2 * w.im : sR2 * w.im + sR * w.re;
            const float noI = re ? nI * w.re - nI2 * w.im : nI2 * w.im + nI * w.re;
            const float soI = re ? sI * w.re - sI2 * w.im : sI2 * w.im + sI * w.re;
            if (jr < d.npairs && map < ncol) {
                const int rn = jr, rs = d.nrings - 1 - jr;
                float* pn = reinterpret_cast<float*>(phase + syn_phase_off(rn, m, d.mmax, ncp) + 2 * map);
                pn[re ? 0 : 1] = noR;
                pn[re ? 2 : 3] = noI;
                if (rs != rn) {
                    float* ps = reinterpret_cast<float*>(phase + syn_phase_off(rs, m, d.mmax, ncp) + 2 * map);
                    ps[re ? 0 : 1] = soR;
                    ps[re ? 2 : 3] = soI;
                }
            }
        }
        return;
    }
    const int col = chunk * BCM + (li32 >> 1);
    const int q = li32 & 3;
    const float sg = (q == 0 || q == 3) ? -1.f : 1.f;
#pragma unroll
    for (int vv = 0; vv < 16; vv++) {
        const int i = 8 * (vv >> 2) + 4 * h + (vv & 3);
        const int jr = rb * 128 + wave * 32 + i;
        const float pe = swp ? Pod[vv] : Pev[vv], po = swp ? Pev[vv] : Pod[vv];
        const float me = swp ? Mod[vv] : Mev[vv], mo = swp ? Mev[vv] : Mod[vv];
        const float mn3 = __shfl_xor(me + mo, 3), ms3 = __shfl_xor(mo - me, 3);
        const float nv = -((pe + po) + sg * mn3), sv = -((pe - po) + sg * ms3);
        const float nv2 = __shfl_xor(nv, 1), sv2 = __shfl_xor(sv, 1);  // the other part
        const cplx<float> w = rot[wave * 32 + i];
        const float no = re ? nv * w.re - nv2 * w.im : nv2 * w.im + nv * w.re;
        const float so = re ? sv * w.re - sv2 * w.im : sv2 * w.im + sv * w.re;
        if (jr < d.npairs && col < ncol) {
            const int rn = jr, rs = d.nrings - 1 - jr;
            reinterpret_cast<float*>(phase + syn_phase_off(rn, m, d.mmax, ncol) + col)[re ? 0 : 1] = no;
            if (rs != rn) reinterpret_cast<float*>(phase + syn_phase_off(rs, m, d.mmax, ncol) + col)[re ? 0 : 1] = so;
        }
    }
}

// ----------------------- Legendre synthesis, spin s, FP32 matrix cores, wide form
// k_leg_synth_sp_mfma's tiling and output combination with the structure of k_leg_synth_s0_wr:
// one block covers NT tiles of 16 complex columns (8 maps of (G, C), or with GONLY the G of 16
// maps), so the two FP64 recursions of a ring pair (lam+ and lam-) feed 4 NT MFMAs per group of
// four l instead of 4; the alm rows and recursion coefficients of the next chunk are copied
// global -> LDS by global_load_lds_dwordx4 while the current chunk runs (one barrier per chunk).
// Rows past the end of the transform and columns past ld read clamped, finite entries: such rows
// only meet zero recursion values and such columns land in outputs that are not stored.
template <int NT, bool GONLY>
__global__ __launch_bounds__(256) void k_leg_synth_sp_wr(ShtDev<float> d, int ncol, int ld,
                                                         const cplx<float>* __restrict__ alm_t,
                                                         cplx<float>* __restrict__ phase) {
    constexpr int BCM = 16 * NT;
    constexpr int RW = 2 * BCM;
    __shared__ __align__(16) float sa0[LCH * RW], sa1[LCH * RW];
    __shared__ __align__(16) double sco0[LCH * 4], sco1[LCH * 4];  // (alpha, beta, gamma, 0) per l
    __shared__ cplx<float> rot[128];
    const int nrb = (d.npairs + 127) / 128;
    const int nchunk = (ncol + BCM - 1) / BCM;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = (int)(lb % nrb);
    const int chunk = (lb / nrb) % nchunk;
    const int m = __builtin_amdgcn_readfirstlane(m_of_q(lb / (nrb * nchunk), d.mmax));
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int h = lane >> 5, li32 = lane & 31;
    const int jp = rb * 128 + wave * 32 + li32;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    const int M1 = d.mmax + 1;
    if (tid < 128) {
        const int jr = min(rb * 128 + tid, d.npairs - 1);
        const RingInfo ri = d.rings[jr];
        rot[tid] = expi_pi<float>((long)m * ri.phi_p, ri.phi_q);  // exp(i m phi0)
    }
    v16f Pev[NT], Pod[NT], Mev[NT], Mod[NT];
#pragma unroll
    for (int c = 0; c < NT; c++) Pev[c] = Pod[c] = Mev[c] = Mod[c] = v16f{0};
    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const int lstart = m > d.spin ? m : d.spin;
    const int skip = __builtin_amdgcn_readfirstlane(__syncthreads_and(d.dead[(size_t)m * d.npairs + jpc]));
    if (!skip) {
        const double x = d.cth[jpc];
        Rec rp, rm;
        rec_seed(rp, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
        rec_seed(rm, d.seedL[M1 + m] + d.seedA[M1 + m] * d.l2sh[jpc] + d.seedB[M1 + m] * d.l2ch[jpc],
                 d.seedS[M1 + m]);
        const int nl = d.lmax + 1 - lstart;
        const int nch = (nl + LCH - 1) / LCH;
        static_assert(LCH * RW == 4 * 256 * NT, "NT 16-byte copies per thread per chunk");
        static_assert(LCH * 4 * 8 == 64 * 16, "one wave copies a chunk's coefficients");
        auto stage = [&](int c, float* dsa, double* dsco) {
            const int l0n = lstart + c * LCH, nn = min(LCH, nl - c * LCH);
#pragma unroll
            for (int e = 0; e < NT; e++) {
                const int q = tid + 256 * e, li = min(q / (8 * NT), nn - 1), c4 = q % (8 * NT);
                const int col = min(chunk * BCM + 2 * c4, ld - 2);  // ld is even
                const float* src = reinterpret_cast<const float*>(alm_t) + ((size_t)(base + l0n + li) * ld + col) * 2;
                __builtin_amdgcn_global_load_lds((const void*)src, (ps_lds_void*)&dsa[4 * (256 * e + 64 * wave)], 16, 0,
                                                 0);
            }
            if (wave == 0) {
                const int l = min(l0n + (lane >> 1), d.lmax);
                const char* src = reinterpret_cast<const char*>(d.coef + 4 * (base + l)) + 16 * (lane & 1);
                __builtin_amdgcn_global_load_lds((const void*)src, (ps_lds_void*)dsco, 16, 0, 0);
            }
        };
        stage(0, sa0, sco0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        bool fast = __all(rp.k == 0 && rm.k == 0);
        auto run_chunk = [&](int c, const float* __restrict__ sab, const double* __restrict__ co) {
            const int n = __builtin_amdgcn_readfirstlane(min(LCH, nl - c * LCH));
            for (int g = 0; g < n; g += 4) {
                float bev[NT], bod[NT];
#pragma unroll
                for (int cc = 0; cc < NT; cc++) {
                    bev[cc] = sab[(g + 2 * h) * RW + 32 * cc + li32];
                    bod[cc] = sab[(g + 1 + 2 * h) * RW + 32 * cc + li32];
                }
                double vp[4], vm[4];
                bool live = true;
                if (g + 4 <= n) {  // a full group: four steps of each recursion, no tail tests
                    double cp[4], cm[4], cg[4];
#pragma unroll
                    for (int cc = 0; cc < 4; cc++) {
                        const double al = co[4 * (g + cc)], be = co[4 * (g + cc) + 1];
                        cp[cc] = al * (x + be);
                        cm[cc] = al * (x - be);
                        cg[cc] = co[4 * (g + cc) + 2];
                    }
                    if (fast) {
#pragma unroll
                        for (int cc = 0; cc < 4; cc++) {
                            vp[cc] = 0.5 * (rp.f + rm.f);
                            vm[cc] = 0.5 * (rp.f - rm.f);
                            rec_fast(rp, cp[cc], cg[cc]);
                            rec_fast(rm, cm[cc], cg[cc]);
                        }
                    } else {
#pragma unroll
                        for (int cc = 0; cc < 4; cc++) {
                            const double fp = rp.f * rp.cf, fm = rm.f * rm.cf;
                            vp[cc] = 0.5 * (fp + fm);
                            vm[cc] = 0.5 * (fp - fm);
                            rec_step(rp, cp[cc], cg[cc]);
                            rec_step(rm, cm[cc], cg[cc]);
                        }
                    }
                } else {  // the last, partial group of the transform
#pragma unroll
                    for (int cc = 0; cc < 4; cc++) {
                        if (g + cc < n) {
                            const double fp = rp.f * rp.cf, fm = rm.f * rm.cf;
                            vp[cc] = 0.5 * (fp + fm);
                            vm[cc] = 0.5 * (fp - fm);
                            const double al = co[4 * (g + cc)], be = co[4 * (g + cc) + 1], ga = co[4 * (g + cc) + 2];
                            rec_step(rp, al * (x + be), ga);
                            rec_step(rm, al * (x - be), ga);
                        } else {
                            vp[cc] = vm[cc] = 0.0;
                        }
                    }
                }
                if (!fast) {
                    rec_rescale(rp);
                    rec_rescale(rm);
                    fast = __all(rp.k == 0 && rm.k == 0);
                    // all values below FP32 range: skip the products
                    live = __any((rp.k >= scale_consts<float>::KMIN - 1) || (rm.k >= scale_consts<float>::KMIN - 1));
                }
                if (!live) continue;
                const float ape = (float)(h ? vp[2] : vp[0]), apo = (float)(h ? vp[3] : vp[1]);
                const float ame = (float)(h ? vm[2] : vm[0]), amo = (float)(h ? vm[3] : vm[1]);
#pragma unroll
                for (int cc = 0; cc < NT; cc++) {
                    Pev[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(ape, bev[cc], Pev[cc], 0, 0, 0);
                    Pod[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(apo, bod[cc], Pod[cc], 0, 0, 0);
                    Mev[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(ame, bev[cc], Mev[cc], 0, 0, 0);
                    Mod[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(amo, bod[cc], Mod[cc], 0, 0, 0);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);  // the next chunk's copies have landed in LDS
            __syncthreads();
        };
        for (int c = 0; c < nch; c += 2) {
            if (c + 1 < nch) stage(c + 1, sa1, sco1);
            run_chunk(c, sa0, sco0);
            if (c + 1 < nch) {
                if (c + 2 < nch) stage(c + 2, sa0, sco0);
                run_chunk(c + 1, sa1, sco1);
            }
        }
    }
    // the output combination of k_leg_synth_sp_mfma, tile by tile
    const bool swp = ((lstart + m + d.spin) & 1) != 0;
    const bool re = (li32 & 1) == 0;
#pragma unroll
    for (int cc = 0; cc < NT; cc++) {
        if constexpr (GONLY) {
            const int map = chunk * BCM + 16 * cc + (li32 >> 1);
            const int ncp = 2 * ncol;
            const float sgi = re ? -1.f : 1.f;
#pragma unroll
            for (int vv = 0; vv < 16; vv++) {
                const int i = 8 * (vv >> 2) + 4 * h + (vv & 3);
                const int jr = rb * 128 + wave * 32 + i;
                const float pe = swp ? Pod[cc][vv] : Pev[cc][vv], po = swp ? Pev[cc][vv] : Pod[cc][vv];
                const float me = swp ? Mod[cc][vv] : Mev[cc][vv], mo = swp ? Mev[cc][vv] : Mod[cc][vv];
                const float nR = -(pe + po), sR = -(pe - po);
                const float nI = sgi * __shfl_xor(me + mo, 1), sI = sgi * __shfl_xor(mo - me, 1);
                const float nR2 = __shfl_xor(nR, 1), sR2 = __shfl_xor(sR, 1);
                const float nI2 = __shfl_xor(nI, 1), sI2 = __shfl_xor(sI, 1);
                const cplx<float> w = rot[wave * 32 + i];
                const float noR = re ? nR * w.re - nR2 * w.im : nR2 * w.im + nR * w.re;
                const float soR = re ? sR * w.re - sR2 * w.im : sR2 * w.im + sR * w.re;
                const float noI = re ? nI * w.re - nI2 * w.im : nI2 * w.im + nI * w.re;
                const float soI = re ? sI * w.re - sI2 * w.im : sI2 * w.im + sI * w.re;
                if (jr < d.npairs && map < ncol) {
                    const int rn = jr, rs = d.nrings - 1 - jr;
                    float* pn = reinterpret_cast<float*>(phase + syn_phase_off(rn, m, d.mmax, ncp) + 2 * map);
                    pn[re ? 0 : 1] = noR;
                    pn[re ? 2 : 3] = noI;
                    if (rs != rn) {
                        float* ps = reinterpret_cast<float*>(phase + syn_phase_off(rs, m, d.mmax, ncp) + 2 * map);
                        ps[re ? 0 : 1] = soR;
                        ps[re ? 2 : 3] = soI;
                    }
                }
            }
        } else {
            const int col = chunk * BCM + 16 * cc + (li32 >> 1);
            const int q = li32 & 3;
            const float sg = (q == 0 || q == 3) ? -1.f : 1.f;
#pragma unroll
            for (int vv = 0; vv < 16; vv++) {
                const int i = 8 * (vv >> 2) + 4 * h + (vv & 3);
                const int jr = rb * 128 + wave * 32 + i;
                const float pe = swp ? Pod[cc][vv] : Pev[cc][vv], po = swp ? Pev[cc][vv] : Pod[cc][vv];
                const float me = swp ? Mod[cc][vv] : Mev[cc][vv], mo = swp ? Mev[cc][vv] : Mod[cc][vv];
                const float mn3 = __shfl_xor(me + mo, 3), ms3 = __shfl_xor(mo - me, 3);
                const float nv = -((pe + po) + sg * mn3), sv = -((pe - po) + sg * ms3);
                const float nv2 = __shfl_xor(nv, 1), sv2 = __shfl_xor(sv, 1);
                const cplx<float> w = rot[wave * 32 + i];
                const float no = re ? nv * w.re - nv2 * w.im : nv2 * w.im + nv * w.re;
                const float so = re ? sv * w.re - sv2 * w.im : sv2 * w.im + sv * w.re;
                if (jr < d.npairs && col < ncol) {
                    const int rn = jr, rs = d.nrings - 1 - jr;
                    reinterpret_cast<float*>(phase + syn_phase_off(rn, m, d.mmax, ncol) + col)[re ? 0 : 1] = no;
                    if (rs != rn)
                        reinterpret_cast<float*>(phase + syn_phase_off(rs, m, d.mmax, ncol) + col)[re ? 0 : 1] = so;
                }
            }
        }
    }
}

// ------------------------------- Legendre syntheses, spin 0 and spin s, FP64 matrix cores
// F_m(ring, col) = sum_l lambda_l(ring) a_lm(col) on v_mfma_f64_16x16x4f64 (A[i][k] at lane
// i + 16 k, B[k][j] at lane j + 16 k, D[i][j] at lane j + 16 (i % 4), register i / 4).  The
// recursion values are staged in LDS as in the FP64 adjoints: per tile of 32 values of l each
// lane writes its ring's values into rows [parity][(l - l0) / 2][ring ^ swizzle(row)], and a
// wave issues one MFMA per 16 rings, parity and 4 values of l of that parity, A = lambda
// (16 rings x 4 l), B = the tile's alm rows (4 l x 16 real columns), which the block stages in
// LDS (the next tile's rows are loaded into registers while the current one runs).  Even and
// odd l - l_start accumulate separately: north = even + odd, south = even - odd.
typedef double v4d_t __attribute__((ext_vector_type(4)));
constexpr int SYN64_L = 32;  // values of l per tile

// Spin 0: one block = 4 waves = 256 ring pairs (one per lane), one m, 8 complex columns.
__global__ __launch_bounds__(256) void k_leg_synth_s0_mfma64(ShtDev<double> d, int ncol, int ld,
                                                             const cplx<double>* __restrict__ alm_t,
                                                             cplx<double>* __restrict__ phase) {
    __shared__ __align__(16) double T[4 * 2 * 16 * 64];  // lambda tiles [wave][parity][row][64]
    __shared__ __align__(16) double sa[SYN64_L * 16];   // alm rows [l - l0][real column]
    __shared__ double sco[SYN64_L * 2];                 // alpha, gamma
    __shared__ cplx<double> rot[256];
    const int nrb = (d.npairs + 255) / 256;
    const int nchunk = (ncol + 7) / 8;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = lb % nrb;
    const int chunk = (lb / nrb) % nchunk;
    const int m = m_of_q(lb / (nrb * nchunk), d.mmax);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k4 = lane >> 4, j16 = lane & 15;
    const int jp = rb * 256 + tid;  // this lane's ring pair
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    {
        const RingInfo ri = d.rings[jpc];
        rot[tid] = expi_pi<double>((long)m * ri.phi_p, ri.phi_q);  // exp(i m phi0)
    }
    const double x = d.cth[jpc];
    Rec r;
    rec_seed(r, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
    v4d_t D[4][2];
#pragma unroll
    for (int rg = 0; rg < 4; rg++) D[rg][0] = D[rg][1] = v4d_t{0.0, 0.0, 0.0, 0.0};

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const double* cf4 = d.coef + 4 * base;
    // staging: thread t holds complex column t % 8 of tile row t / 8 (columns past ld read as zero)
    const int srow = tid >> 3, scol = tid & 7;
    const bool scol_ok = chunk * 8 + scol < ld;
    cplx<double> pa{0.0, 0.0};
    double pal = 0.0, pga = 0.0;
    auto fetch = [&](int l0n) {
        const int nn = min(SYN64_L, d.lmax + 1 - l0n);
        pa = (srow < nn && scol_ok) ? alm_t[(size_t)(base + l0n + srow) * ld + chunk * 8 + scol]
                                    : cplx<double>{0.0, 0.0};
        if (tid < nn) {
            pal = cf4[4 * (l0n + tid)];
            pga = cf4[4 * (l0n + tid) + 2];
        }
    };
    // a block whose ring pairs are all dead at this m writes zero phases without the l loop
    const bool skip = __syncthreads_and(jp >= d.npairs || d.dead[(size_t)m * d.npairs + jpc]);
    double* Tw = T + wave * 2048;
    bool fast = false;
    if (!skip) fetch(m);
    for (int l0 = m; l0 <= d.lmax && !skip; l0 += SYN64_L) {
        const int n = min(SYN64_L, d.lmax + 1 - l0);
        __syncthreads();  // every read of the previous tile is done
        sa[srow * 16 + 2 * scol] = pa.re;
        sa[srow * 16 + 2 * scol + 1] = pa.im;
        if (tid < SYN64_L) {
            sco[2 * tid] = pal;
            sco[2 * tid + 1] = pga;
        }
        __syncthreads();
        if (l0 + SYN64_L <= d.lmax) fetch(l0 + SYN64_L);
        if (!fast) fast = __all(r.k == 0);
        if (fast && n == SYN64_L) {
#pragma unroll 4
            for (int c = 0; c < SYN64_L; c++) {
                const int i = c >> 1;
                Tw[(c & 1) * 1024 + i * 64 + (lane ^ (4 * i))] = r.f;
                rec_fast(r, sco[2 * c] * x, sco[2 * c + 1]);
            }
        } else if (n == SYN64_L) {
            // full tile while some lane still carries a scale exponent
#pragma unroll 4
            for (int c = 0; c < SYN64_L; c++) {
                const int i = c >> 1;
                Tw[(c & 1) * 1024 + i * 64 + (lane ^ (4 * i))] = r.f * r.cf;
                rec_fast(r, sco[2 * c] * x, sco[2 * c + 1]);
                if ((c & 3) == 3) rec_rescale(r);
            }
        } else {
            for (int c = 0; c < SYN64_L; c++) {
                const int i = c >> 1;
                double v = 0.0;
                if (c < n) {
                    v = fast ? r.f : r.f * r.cf;
                    rec_fast(r, sco[2 * c] * x, sco[2 * c + 1]);
                    if (!fast && (c & 3) == 3) rec_rescale(r);
                }
                Tw[(c & 1) * 1024 + i * 64 + (lane ^ (4 * i))] = v;
            }
        }
        __syncthreads();  // the tiles are complete
#pragma unroll
        for (int par = 0; par < 2; par++)
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int row = 4 * kk + k4;  // (l - l0) / 2 of this lane's k
                const double b = sa[(2 * row + par) * 16 + j16];
#pragma unroll
                for (int rg = 0; rg < 4; rg++) {
                    const double a = Tw[par * 1024 + row * 64 + ((16 * rg + j16) ^ (4 * row))];
                    D[rg][par] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, D[rg][par], 0, 0, 0);
                }
            }
    }
    // D[rg][par][v] at lane (j16, k4): ring pair 16 rg + 4 v + k4 of the wave, real column j16
    const bool re = (j16 & 1) == 0;
    const int col = chunk * 8 + (j16 >> 1);
#pragma unroll
    for (int rg = 0; rg < 4; rg++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int ir = wave * 64 + 16 * rg + 4 * v + k4;
            const int jr = rb * 256 + ir;
            const double e = D[rg][0][v], o = D[rg][1][v];
            const double nv = e + o, sv = e - o;
            const double nv2 = __shfl_xor(nv, 1), sv2 = __shfl_xor(sv, 1);  // the other part
            const cplx<double> w = rot[ir];
            const double no = re ? nv * w.re - nv2 * w.im : nv2 * w.im + nv * w.re;
            const double so = re ? sv * w.re - sv2 * w.im : sv2 * w.im + sv * w.re;
            if (jr < d.npairs && col < ncol) {
                const int rn = jr, rs = d.nrings - 1 - jr;
                reinterpret_cast<double*>(phase + syn_phase_off(rn, m, d.mmax, ncol) + col)[re ? 0 : 1] = no;
                if (rs != rn)
                    reinterpret_cast<double*>(phase + syn_phase_off(rs, m, d.mmax, ncol) + col)[re ? 0 : 1] = so;
            }
        }
}

// Spin s: lanes 0-31 of a wave run f+ and lanes 32-63 f- of the wave's 32 ring pairs; the
// accumulators hold F+- = sum_l f+- a_lm, and lam+- = (f+ +- f-) / 2 are formed at the end
// (D+- = (F+ +- F-) / 2), which then combine as in k_leg_synth_sp_mfma.  One block = 4 waves =
// 128 ring pairs, one m, 4 maps (8 complex columns: G and C of each map).
__global__ __launch_bounds__(256) void k_leg_synth_sp_mfma64(ShtDev<double> d, int ncol, int ld,
                                                             const cplx<double>* __restrict__ alm_t,
                                                             cplx<double>* __restrict__ phase) {
    __shared__ __align__(16) double T[4 * 2 * 2 * 16 * 32];  // f tiles [wave][sign][parity][row][32]
    __shared__ __align__(16) double sa[SYN64_L * 16];
    __shared__ double sco[SYN64_L * 3];
    __shared__ cplx<double> rot[128];
    const int nrb = (d.npairs + 127) / 128;
    const int nchunk = (ncol + 7) / 8;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = lb % nrb;
    const int chunk = (lb / nrb) % nchunk;
    const int m = m_of_q(lb / (nrb * nchunk), d.mmax);
    const int M1 = d.mmax + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, li32 = lane & 31;
    const int k4 = lane >> 4, j16 = lane & 15;
    const int jp = rb * 128 + wave * 32 + li32;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    if (tid < 128) {
        const int jr = min(rb * 128 + tid, d.npairs - 1);
        const RingInfo ri = d.rings[jr];
        rot[tid] = expi_pi<double>((long)m * ri.phi_p, ri.phi_q);  // exp(i m phi0)
    }
    const double x = d.cth[jpc];
    const double sgb = h ? -1.0 : 1.0;
    Rec r;
    rec_seed(r, d.seedL[h * M1 + m] + d.seedA[h * M1 + m] * d.l2sh[jpc] + d.seedB[h * M1 + m] * d.l2ch[jpc],
             d.seedS[h * M1 + m]);
    v4d_t F[2][2][2];  // [sign][parity][ring group]
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) F[a][b][0] = F[a][b][1] = v4d_t{0.0, 0.0, 0.0, 0.0};

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const double* cf4 = d.coef + 4 * base;
    const int lstart = m > d.spin ? m : d.spin;
    const int srow = tid >> 3, scol = tid & 7;
    const bool scol_ok = chunk * 8 + scol < ld;
    cplx<double> pa{0.0, 0.0};
    double pal = 0.0, pbe = 0.0, pga = 0.0;
    auto fetch = [&](int l0n) {
        const int nn = min(SYN64_L, d.lmax + 1 - l0n);
        pa = (srow < nn && scol_ok) ? alm_t[(size_t)(base + l0n + srow) * ld + chunk * 8 + scol]
                                    : cplx<double>{0.0, 0.0};
        if (tid < nn) {
            pal = cf4[4 * (l0n + tid)];
            pbe = cf4[4 * (l0n + tid) + 1];
            pga = cf4[4 * (l0n + tid) + 2];
        }
    };
    const bool skip = __syncthreads_and(jp >= d.npairs || d.dead[(size_t)m * d.npairs + jpc]);
    double* Tw = T + wave * 2048 + h * 1024;  // this lane's sign: [parity][row][32]
    const double* Ta = T + wave * 2048;
    bool fast = false;
    if (!skip) fetch(lstart);
    for (int l0 = lstart; l0 <= d.lmax && !skip; l0 += SYN64_L) {
        const int n = min(SYN64_L, d.lmax + 1 - l0);
        __syncthreads();  // every read of the previous tile is done
        sa[srow * 16 + 2 * scol] = pa.re;
        sa[srow * 16 + 2 * scol + 1] = pa.im;
        if (tid < SYN64_L) {
            sco[3 * tid] = pal;
            sco[3 * tid + 1] = pbe;
            sco[3 * tid + 2] = pga;
        }
        __syncthreads();
        if (l0 + SYN64_L <= d.lmax) fetch(l0 + SYN64_L);
        if (!fast) fast = __all(r.k == 0);
        if (fast && n == SYN64_L) {
#pragma unroll 4
            for (int c = 0; c < SYN64_L; c++) {
                const int i = c >> 1;
                Tw[(c & 1) * 512 + i * 32 + (li32 ^ ((4 * i) & 28))] = r.f;
                rec_fast(r, sco[3 * c] * (x + sgb * sco[3 * c + 1]), sco[3 * c + 2]);
            }
        } else if (n == SYN64_L) {
#pragma unroll 4
            for (int c = 0; c < SYN64_L; c++) {
                const int i = c >> 1;
                Tw[(c & 1) * 512 + i * 32 + (li32 ^ ((4 * i) & 28))] = r.f * r.cf;
                rec_fast(r, sco[3 * c] * (x + sgb * sco[3 * c + 1]), sco[3 * c + 2]);
                if ((c & 3) == 3) rec_rescale(r);
            }
        } else {
            for (int c = 0; c < SYN64_L; c++) {
                const int i = c >> 1;
                double v = 0.0;
                if (c < n) {
                    v = fast ? r.f : r.f * r.cf;
                    rec_fast(r, sco[3 * c] * (x + sgb * sco[3 * c + 1]), sco[3 * c + 2]);
                    if (!fast && (c & 3) == 3) rec_rescale(r);
                }
                Tw[(c & 1) * 512 + i * 32 + (li32 ^ ((4 * i) & 28))] = v;
            }
        }
        __syncthreads();  // the tiles are complete
#pragma unroll
        for (int par = 0; par < 2; par++)
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int row = 4 * kk + k4;
                const double b = sa[(2 * row + par) * 16 + j16];
#pragma unroll
                for (int sg = 0; sg < 2; sg++)
#pragma unroll
                    for (int rg = 0; rg < 2; rg++) {
                        const double a = Ta[sg * 1024 + par * 512 + row * 32 + ((16 * rg + j16) ^ ((4 * row) & 28))];
                        F[sg][par][rg] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, F[sg][par][rg], 0, 0, 0);
                    }
            }
    }
    // even accumulators hold l - l_start even, i.e. l + m + s even unless l_start + m + s is odd
    const bool swp = ((lstart + m + d.spin) & 1) != 0;
    const int col = chunk * 8 + (j16 >> 1);
    const bool re = (j16 & 1) == 0;
    const int q = j16 & 3;
    const double sgn = (q == 0 || q == 3) ? -1.0 : 1.0;
#pragma unroll
    for (int rg = 0; rg < 2; rg++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int ir = wave * 32 + 16 * rg + 4 * v + k4;
            const int jr = rb * 128 + ir;
            const double Pe = 0.5 * (F[0][0][rg][v] + F[1][0][rg][v]), Po = 0.5 * (F[0][1][rg][v] + F[1][1][rg][v]);
            const double Me = 0.5 * (F[0][0][rg][v] - F[1][0][rg][v]), Mo = 0.5 * (F[0][1][rg][v] - F[1][1][rg][v]);
            const double pe = swp ? Po : Pe, po = swp ? Pe : Po;
            const double me = swp ? Mo : Me, mo = swp ? Me : Mo;
            const double mn3 = __shfl_xor(me + mo, 3), ms3 = __shfl_xor(mo - me, 3);
            const double nv = -((pe + po) + sgn * mn3), sv = -((pe - po) + sgn * ms3);
            const double nv2 = __shfl_xor(nv, 1), sv2 = __shfl_xor(sv, 1);  // the other part
            const cplx<double> w = rot[ir];
            const double no = re ? nv * w.re - nv2 * w.im : nv2 * w.im + nv * w.re;
            const double so = re ? sv * w.re - sv2 * w.im : sv2 * w.im + sv * w.re;
            if (jr < d.npairs && col < ncol) {
                const int rn = jr, rs = d.nrings - 1 - jr;
                reinterpret_cast<double*>(phase + syn_phase_off(rn, m, d.mmax, ncol) + col)[re ? 0 : 1] = no;
                if (rs != rn)
                    reinterpret_cast<double*>(phase + syn_phase_off(rs, m, d.mmax, ncol) + col)[re ? 0 : 1] = so;
            }
        }
}

// -------------------------------------------------- Legendre synthesis, spin s
// alm_t: [nlm][ld] complex with columns (map b, comp c) at 2 b + c (G, C).
template <typename T, int BC>
__device__ __forceinline__ void synth_sp_acc(cplx<T> (&acc)[BC][4], const cplx<T>* __restrict__ sa, T lp, T ln) {
#pragma unroll
    for (int j = 0; j < BC; j++) {
        const cplx<T> g = sa[2 * j], c = sa[2 * j + 1];
        acc[j][0].re += lp * g.re; acc[j][0].im += lp * g.im;
        acc[j][1].re += lp * c.re; acc[j][1].im += lp * c.im;
        acc[j][2].re += ln * g.re; acc[j][2].im += ln * g.im;
        acc[j][3].re += ln * c.re; acc[j][3].im += ln * c.im;
    }
}

template <typename T, int BC>
__global__ __launch_bounds__(256) void k_leg_synth_sp(ShtDev<T> d, int nmap, int ld,
                                                      const cplx<T>* __restrict__ alm_t,
                                                      cplx<T>* __restrict__ phase) {
    __shared__ cplx<T> sa[LCH * 2 * BC];
    __shared__ double sco[LCH * 3];
    const int nrb = (d.npairs + 255) / 256;
    const int nchunk = (nmap + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = lb % nrb;
    const int chunk = (lb / nrb) % nchunk;
    const int m = m_of_q(lb / (nrb * nchunk), d.mmax);
    const int tid = threadIdx.x;
    const int jp = rb * 256 + tid;
    const int jpc = jp < d.npairs ? jp : d.npairs - 1;
    const int M1 = d.mmax + 1;

    const double x = d.cth[jpc];
    Rec rp, rm;
    rec_seed(rp, d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
    rec_seed(rm, d.seedL[M1 + m] + d.seedA[M1 + m] * d.l2sh[jpc] + d.seedB[M1 + m] * d.l2ch[jpc],
             d.seedS[M1 + m]);

    // acc[q][j][0..3] = sum G lam+, C lam+, G lam-, C lam- over l with li parity q
    cplx<T> acc0[BC][4], acc1[BC][4];
#pragma unroll
    for (int j = 0; j < BC; j++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc0[j][c] = acc1[j][c] = cplx<T>{T(0), T(0)};

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const cplx<T>* a = alm_t + (size_t)chunk * BC * 2;
    const double* cf4 = d.coef + 4 * base;
    const int lstart = m > d.spin ? m : d.spin;
    if constexpr (BC >= 4) {
        // wide chunks: alm rows straight from memory by wave-uniform (scalar) loads
        for (int l = lstart; l <= d.lmax; l++) {
            if (__any((rp.k >= scale_consts<T>::KMIN - 1) || (rm.k >= scale_consts<T>::KMIN - 1))) {
                const double fp = rp.f * rp.cf, fm = rm.f * rm.cf;
                const T lp = (T)(0.5 * (fp + fm)), ln = (T)(0.5 * (fp - fm));
                const cplx<T>* al = a + (size_t)(base + l) * ld;
                if ((l - lstart) & 1)
                    synth_sp_acc<T, BC>(acc1, al, lp, ln);
                else
                    synth_sp_acc<T, BC>(acc0, al, lp, ln);
            }
            const double* c = cf4 + 4 * l;
            rec_step(rp, c[0] * (x + c[1]), c[2]);
            rec_step(rm, c[0] * (x - c[1]), c[2]);
            if (((l - lstart) & 3) == 3) {
                rec_rescale(rp);
                rec_rescale(rm);
            }
        }
    } else {
        // chunk staging: thread t holds alm elements t and t + 256 (rows / (2 BC)) and the
        // (alpha, beta, gamma) of row t; the next chunk is fetched while the current one runs
        static_assert(LCH * 2 * BC <= 512, "two alm elements per thread per chunk");
        cplx<T> pa0{T(0), T(0)}, pa1{T(0), T(0)};
        double pal = 0.0, pbe = 0.0, pga = 0.0;
        auto fetch = [&](int l0n) {
            const int nn = min(LCH, d.lmax + 1 - l0n);
            if (tid < nn * 2 * BC) pa0 = a[(size_t)(base + l0n + tid / (2 * BC)) * ld + tid % (2 * BC)];
            if (tid + 256 < nn * 2 * BC)
                pa1 = a[(size_t)(base + l0n + (tid + 256) / (2 * BC)) * ld + (tid + 256) % (2 * BC)];
            if (tid < nn) {
                pal = cf4[4 * (l0n + tid)];
                pbe = cf4[4 * (l0n + tid) + 1];
                pga = cf4[4 * (l0n + tid) + 2];
            }
        };
        fetch(lstart);
        bool fast = false;
        for (int l0 = lstart; l0 <= d.lmax; l0 += LCH) {
            const int n = min(LCH, d.lmax + 1 - l0);
            __syncthreads();
            if (tid < n * 2 * BC) sa[tid] = pa0;
            if (tid + 256 < n * 2 * BC) sa[tid + 256] = pa1;
            if (tid < n) {
                sco[3 * tid] = pal;
                sco[3 * tid + 1] = pbe;
                sco[3 * tid + 2] = pga;
            }
            __syncthreads();
            if (l0 + LCH <= d.lmax) fetch(l0 + LCH);
            if (!fast) fast = __all(rp.k == 0 && rm.k == 0);
            if (fast) {
                int li = 0;
                for (; li + 1 < n; li += 2) {
                    synth_sp_acc<T, BC>(acc0, sa + li * 2 * BC, (T)(0.5 * (rp.f + rm.f)), (T)(0.5 * (rp.f - rm.f)));
                    double al = sco[3 * li], be = sco[3 * li + 1], ga = sco[3 * li + 2];
                    rec_fast(rp, al * (x + be), ga);
                    rec_fast(rm, al * (x - be), ga);
                    synth_sp_acc<T, BC>(acc1, sa + (li + 1) * 2 * BC, (T)(0.5 * (rp.f + rm.f)),
                                        (T)(0.5 * (rp.f - rm.f)));
                    al = sco[3 * li + 3], be = sco[3 * li + 4], ga = sco[3 * li + 5];
                    rec_fast(rp, al * (x + be), ga);
                    rec_fast(rm, al * (x - be), ga);
                }
                if (li < n)
                    synth_sp_acc<T, BC>(acc0, sa + li * 2 * BC, (T)(0.5 * (rp.f + rm.f)), (T)(0.5 * (rp.f - rm.f)));
            } else {
                // groups of 4 values of l: the accumulator and the exponent update are fixed per
                // unrolled step (no per-l branches); a partial last group is not rescaled
                for (int li0 = 0; li0 < n; li0 += 4) {
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        const int li = li0 + c;
                        if (li < n) {
                            if (__any((rp.k >= scale_consts<T>::KMIN - 1) || (rm.k >= scale_consts<T>::KMIN - 1))) {
                                const double fp = rp.f * rp.cf, fm = rm.f * rm.cf;
                                const T lp = (T)(0.5 * (fp + fm)), ln = (T)(0.5 * (fp - fm));
                                synth_sp_acc<T, BC>((c & 1) ? acc1 : acc0, sa + li * 2 * BC, lp, ln);
                            }
                            const double al = sco[3 * li], be = sco[3 * li + 1], ga = sco[3 * li + 2];
                            rec_step(rp, al * (x + be), ga);
                            rec_step(rm, al * (x - be), ga);
                        }
                    }
                    if (li0 + 3 < n) {
                        rec_rescale(rp);
                        rec_rescale(rm);
                    }
                }
            }
        }
    }
    if (jp >= d.npairs) return;
    const int rn = jp, rs = d.nrings - 1 - jp;
    const RingInfo ri = d.rings[rn];
    const cplx<T> rot = expi_pi<T>((long)m * ri.phi_p, ri.phi_q);  // exp(i m phi0)
    const int ncol = 2 * nmap;
    // li parity 0 holds l + m + s even unless l_start + m + s is odd; with the (-1)^s in the
    // -s recursion, sY(pi - theta) = (-1)^(l+m) (-s)Y(theta) gives lam+- the south sign (-1)^(l+m+s)
    const bool swp = ((lstart + m + d.spin) & 1) != 0;
#pragma unroll
    for (int j = 0; j < BC; j++) {
        const int b = chunk * BC + j;
        if (b >= nmap) continue;
        cplx<T> R[2], I[2];  // north, south
        // R = -(G lam+ + i C lam-), I = -(C lam+ - i G lam-); south: lam+ -> (-1)^(l+m+s),
        // lam- -> -(-1)^(l+m+s)
        const cplx<T> GPe = swp ? acc1[j][0] : acc0[j][0], GPo = swp ? acc0[j][0] : acc1[j][0];
        const cplx<T> CPe = swp ? acc1[j][1] : acc0[j][1], CPo = swp ? acc0[j][1] : acc1[j][1];
        const cplx<T> GMe = swp ? acc1[j][2] : acc0[j][2], GMo = swp ? acc0[j][2] : acc1[j][2];
        const cplx<T> CMe = swp ? acc1[j][3] : acc0[j][3], CMo = swp ? acc0[j][3] : acc1[j][3];
        cplx<T> t;
        t = cadd(CMe, CMo);
        R[0] = cplx<T>{-(GPe.re + GPo.re) + t.im, -(GPe.im + GPo.im) - t.re};
        t = cadd(GMe, GMo);
        I[0] = cplx<T>{-(CPe.re + CPo.re) - t.im, -(CPe.im + CPo.im) + t.re};
        t = csub(CMo, CMe);
        R[1] = cplx<T>{-(GPe.re - GPo.re) + t.im, -(GPe.im - GPo.im) - t.re};
        t = csub(GMo, GMe);
        I[1] = cplx<T>{-(CPe.re - CPo.re) - t.im, -(CPe.im - CPo.im) + t.re};
        cplx<T>* phn = phase + syn_phase_off(rn, m, d.mmax, ncol) + 2 * b;
        phn[0] = cmul(R[0], rot);
        phn[1] = cmul(I[0], rot);
        if (rs != rn) {
            cplx<T>* phs = phase + syn_phase_off(rs, m, d.mmax, ncol) + 2 * b;
            phs[0] = cmul(R[1], rot);
            phs[1] = cmul(I[1], rot);
        }
    }
}

// ----------------------------------------------------------- ring FFTs (LDS)

__device__ __forceinline__ int bitrev(int k, int logM) {
    return logM ? (int)(__brev((unsigned)k) >> (32 - logM)) : 0;
}

// LDS element index of entry i of a transform buffer: bits 2..5 of i are XORed into bits
// 0..3 (a permutation inside each aligned group of 16 entries, so no padding).  With
// radix-4 butterflies a quarter-wave of 16 lanes then touches 32 distinct 4-byte banks in
// every pass, the short-stride passes included (a pad slot per 32 entries left the last
// passes 2-4 way conflicted).  Buffers are spaced fft_pitch(M) apart.
__device__ __forceinline__ int pidx(int i) { return i ^ ((i >> 2) & 15); }
__host__ __device__ __forceinline__ int fft_pitch(int M) { return M + 1; }

// Batched transforms over NB padded buffers of length M <= 256 * 4 * Q (Q <= FFT_MAXQ: the
// butterfly-group slots each thread holds, a template parameter so that the twiddle registers
// of a launch are sized for its longest transform).  Two
// radix-2 stages are fused into one radix-2^2 pass in registers (half the LDS traffic,
// same bit-reversed ordering); with log2 M odd one plain radix-2 stage remains.  The
// twiddles of a pass are loaded for all of a thread's butterflies before any of them
// runs.  tw[k] = exp(-2 pi i k / maxM); the inverse transform uses conjugates.
constexpr int FFT_MAXQ = 8;

__device__ __forceinline__ int ilog2_dev(int v) { return 31 - __clz(v); }

template <typename T>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a, bool inv) {  // a * (-i), or a * i for inv
    return inv ? cplx<T>{-a.im, a.re} : cplx<T>{a.im, -a.re};
}

// one radix-2 stage of half-size h over all buffers (DIF when !dit)
template <typename T, int NB, int Q>
__device__ __forceinline__ void fft_r2_stage(cplx<T>* a, int M, int h, const cplx<T>* __restrict__ tw, long maxM,
                                             bool inv, bool dit) {
    const int half = M >> 1, pitch = fft_pitch(M);
    const long ts = maxM / (2L * h);
    cplx<T> w[2 * Q];
#pragma unroll
    for (int k = 0; k < 2 * Q; k++) {
        const int u = threadIdx.x + 256 * k;
        if (u < half) {
            w[k] = tw[(u & (h - 1)) * ts];
            if (inv) w[k].im = -w[k].im;
        }
    }
#pragma unroll
    for (int k = 0; k < 2 * Q; k++) {
        const int u = threadIdx.x + 256 * k;
        if (u < half) {
            const int pos = u & (h - 1);
            const int i = pidx(((u - pos) << 1) + pos), j = pidx(((u - pos) << 1) + pos + h);
#pragma unroll
            for (int q = 0; q < NB; q++) {
                cplx<T>* b = a + q * pitch;
                if (dit) {
                    const cplx<T> x = b[i], y = cmul(b[j], w[k]);
                    b[i] = cadd(x, y);
                    b[j] = csub(x, y);
                } else {
                    const cplx<T> x = b[i], y = b[j];
                    b[i] = cadd(x, y);
                    b[j] = cmul(csub(x, y), w[k]);
                }
            }
        }
    }
    __syncthreads();
}

// Twiddles of the fused DIF pass with stage half-size h: wA[k] for butterfly group
// v = threadIdx.x + 256 k (stage h; stage h/2 derives its own from them).
template <typename T, int Q>
__device__ __forceinline__ void dif_twiddles(cplx<T> (&wA)[Q], int h, int quarter,
                                             const cplx<T>* __restrict__ tw, long maxM, bool inv) {
    const int hh = h >> 1;
    const long ts = maxM / (2L * h);
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int v = threadIdx.x + 256 * k;
        if (v < quarter) {
            wA[k] = tw[(v & (hh - 1)) * ts];
            if (inv) wA[k].im = -wA[k].im;
        }
    }
}

// Decimation in frequency: natural order in, bit-reversed out.  Fused stages (h, h/2).  The
// next pass's twiddles are loaded before the current pass's butterflies, so their global-memory
// latency overlaps the LDS work and the barrier.
template <typename T, int NB, int Q>
__device__ void fft_dif_nb(cplx<T>* a, int M, const cplx<T>* __restrict__ tw, long maxM, bool inv) {
    const int quarter = M >> 2, pitch = fft_pitch(M);
    int h = M >> 1;
    cplx<T> wA[Q];
    if (h >= 2) dif_twiddles<T, Q>(wA, h, quarter, tw, maxM, inv);
    for (; h >= 2; h >>= 2) {
        const int hh = h >> 1;
        cplx<T> wN[Q];
        if ((h >> 2) >= 2) dif_twiddles<T, Q>(wN, h >> 2, quarter, tw, maxM, inv);
#pragma unroll
        for (int k = 0; k < Q; k++) {
            const int v = threadIdx.x + 256 * k;
            if (v < quarter) {
                const int pos = v & (hh - 1);
                const int base = (v - pos) * 4 + pos;  // block of 2h entries, h/2 quads each
                const int i0 = pidx(base), i1 = pidx(base + hh), i2 = pidx(base + h), i3 = pidx(base + h + hh);
                const cplx<T> wa = wA[k], wb = mul_mi(wa, inv), wc = cmul(wa, wa);
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    cplx<T>* b = a + q * pitch;
                    const cplx<T> e0 = b[i0], e1 = b[i1], e2 = b[i2], e3 = b[i3];
                    const cplx<T> a0 = cadd(e0, e2), a2 = cmul(csub(e0, e2), wa);
                    const cplx<T> a1 = cadd(e1, e3), a3 = cmul(csub(e1, e3), wb);
                    b[i0] = cadd(a0, a1);
                    b[i1] = cmul(csub(a0, a1), wc);
                    b[i2] = cadd(a2, a3);
                    b[i3] = cmul(csub(a2, a3), wc);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < Q; k++) wA[k] = wN[k];
    }
    if (h == 1) fft_r2_stage<T, NB, Q>(a, M, 1, tw, maxM, inv, false);
}

// Twiddles of the fused DIT pass with stage half-size h: wA (stage h) and wB (stage 2h) for
// butterfly group v = threadIdx.x + 256 k.
template <typename T, int Q>
__device__ __forceinline__ void dit_twiddles(cplx<T> (&wA)[Q], cplx<T> (&wB)[Q], int h, int quarter,
                                             const cplx<T>* __restrict__ tw, long maxM, bool inv) {
    const long ts = maxM / (2L * h);  // stage h; stage 2h uses ts / 2
#pragma unroll
    for (int k = 0; k < Q; k++) {
        const int v = threadIdx.x + 256 * k;
        if (v < quarter) {
            const int pos = v & (h - 1);
            wA[k] = tw[pos * ts];
            wB[k] = tw[pos * (ts >> 1)];
            if (inv) {
                wA[k].im = -wA[k].im;
                wB[k].im = -wB[k].im;
            }
        }
    }
}

// Decimation in time: bit-reversed order in, natural out.  Fused stages (h, 2h); the next
// pass's twiddles are loaded before the current pass's butterflies.
template <typename T, int NB, int Q>
__device__ void fft_dit_nb(cplx<T>* a, int M, const cplx<T>* __restrict__ tw, long maxM, bool inv) {
    const int quarter = M >> 2, pitch = fft_pitch(M);
    int h = 1;
    if (ilog2_dev(M) & 1) {
        fft_r2_stage<T, NB, Q>(a, M, 1, tw, maxM, inv, true);
        h = 2;
    }
    cplx<T> wA[Q], wB[Q];
    if (2 * h < M) dit_twiddles<T, Q>(wA, wB, h, quarter, tw, maxM, inv);
    for (; 2 * h < M; h <<= 2) {
        cplx<T> nA[Q], nB[Q];
        if (8 * h < M) dit_twiddles<T, Q>(nA, nB, 4 * h, quarter, tw, maxM, inv);
#pragma unroll
        for (int k = 0; k < Q; k++) {
            const int v = threadIdx.x + 256 * k;
            if (v < quarter) {
                const int pos = v & (h - 1);
                const int base = (v - pos) * 4 + pos;  // block of 4h entries, h quads each
                const int i0 = pidx(base), i1 = pidx(base + h), i2 = pidx(base + 2 * h), i3 = pidx(base + 3 * h);
                const cplx<T> wa = wA[k], wb = wB[k], wc = mul_mi(wb, inv);
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    cplx<T>* b = a + q * pitch;
                    const cplx<T> e0 = b[i0], e1 = cmul(b[i1], wa), e2 = b[i2], e3 = cmul(b[i3], wa);
                    const cplx<T> a0 = cadd(e0, e1), a1 = csub(e0, e1), a2 = cadd(e2, e3), a3 = csub(e2, e3);
                    const cplx<T> t2 = cmul(a2, wb), t3 = cmul(a3, wc);
                    b[i0] = cadd(a0, t2);
                    b[i2] = csub(a0, t2);
                    b[i1] = cadd(a1, t3);
                    b[i3] = csub(a1, t3);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < Q; k++) {
            wA[k] = nA[k];
            wB[k] = nB[k];
        }
    }
}

// ------------------------------------------- self-sorting (Stockham) ring FFTs (LDS)

// LDS element index of entry i of a self-sorting transform buffer: bits 4..7 of i XORed into
// bits 0..3.  The first pass writes entries 16 apart across lanes; the XOR puts 16 such lanes
// on 16 distinct 8-byte slots (all 32 4-byte banks).  Every other access touches runs of 16
// consecutive aligned entries, which the XOR only permutes.
__device__ __forceinline__ int sidx(int i) { return i ^ ((i >> 4) & 15); }

// cos(2 pi k / 16), sin(2 pi k / 16)
__host__ __device__ constexpr double ur16_cos(int k) {
    constexpr double c[5] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173, 0.0};
    k &= 15;
    return k <= 4 ? c[k] : k <= 8 ? -c[8 - k] : k <= 12 ? -c[k - 8] : c[16 - k];
}
__host__ __device__ constexpr double ur16_sin(int k) { return ur16_cos(k + 12); }
__host__ __device__ constexpr int brev_c(int r, int lr) {
    int o = 0;
    for (int b = 0; b < lr; b++)
        if (r & (1 << b)) o |= 1 << (lr - 1 - b);
    return o;
}

// In-register DFT of R = 2, 4, 8 or 16 points (kernel exp(-2 pi i jk / R), or exp(+...) for
// inv): radix-2 decimation in frequency with compile-time twiddles; v[p] ends holding output
// bitrev_R(p).
template <typename T, int R>
__device__ __forceinline__ void dft_reg(cplx<T> (&v)[R], bool inv) {
#pragma unroll
    for (int h = R / 2; h >= 1; h >>= 1) {
#pragma unroll
        for (int b = 0; b < R; b += 2 * h) {
#pragma unroll
            for (int p = 0; p < h; p++) {
                const cplx<T> x = v[b + p], y = v[b + p + h];
                v[b + p] = cadd(x, y);
                cplx<T> dl = csub(x, y);
                if (p > 0) {
                    if (2 * p == h) {
                        dl = mul_mi(dl, inv);
                    } else {
                        const int k = 8 * p / h;
                        const T c = (T)ur16_cos(k), sn = (T)ur16_sin(k);
                        dl = cmul(dl, cplx<T>{c, inv ? sn : -sn});
                    }
                }
                v[b + p + h] = dl;
            }
        }
    }
}

// One self-sorting pass of radix R over NB buffers of length M, after sub-transforms of length
// Ns: butterfly j < M/R reads entries j + r M/R, multiplies entry r by w^(r (j mod Ns)) with
// w = exp(-/+ 2 pi i / (Ns R)), transforms the R values in registers and hands output r to
// out(q, (j - j mod Ns) R + j mod Ns + r Ns, value).  All reads come before the barrier and all
// writes after it, so the pass runs in place.  Powers w^(2^e) are table entries (indices below
// maxM / 2); the others are products of at most three of them.  U: butterflies per thread
// (M / R <= 256 U).
template <typename T, int NB, int R, int U, typename Out>
__device__ __forceinline__ void stk_pass(const cplx<T>* a, int M, int Ns, const cplx<T>* __restrict__ tw, long maxM,
                                         bool inv, Out&& out) {
    constexpr int LR = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
    const int nb = M / R, pitch = fft_pitch(M);
    cplx<T> v[U][NB][R];
    cplx<T> w[U][LR];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int j = threadIdx.x + 256 * u;
        if (j < nb) {
#pragma unroll
            for (int q = 0; q < NB; q++)
#pragma unroll
                for (int r = 0; r < R; r++) v[u][q][r] = a[q * pitch + sidx(j + r * nb)];
            if (Ns > 1) {
                const long ts = maxM / ((long)Ns * R);
                const long k = j & (Ns - 1);
#pragma unroll
                for (int e = 0; e < LR; e++) {
                    cplx<T> t = tw[(k << e) * ts];
                    if (inv) t.im = -t.im;
                    w[u][e] = t;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int j = threadIdx.x + 256 * u;
        if (j < nb) {
            const int k = j & (Ns - 1);
            if (Ns > 1) {
                cplx<T> pw[R];
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const int hb = 1 << (31 - __builtin_clz((unsigned)r));
                    pw[r] = hb == r ? w[u][31 - __builtin_clz((unsigned)r)] : cmul(pw[hb], pw[r - hb]);
                }
#pragma unroll
                for (int q = 0; q < NB; q++)
#pragma unroll
                    for (int r = 1; r < R; r++) v[u][q][r] = cmul(v[u][q][r], pw[r]);
            }
#pragma unroll
            for (int q = 0; q < NB; q++) dft_reg<T, R>(v[u][q], inv);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < NB; q++)
#pragma unroll
                for (int r = 0; r < R; r++) out(q, base + r * Ns, v[u][q][brev_c(r, LR)]);
        }
    }
}

// Self-sorting LDS FFT over NB buffers of length M = 2^logM, natural order in and out (entry i
// at sidx(i)): radix-16 passes, the last of radix 2^(logM - 4 p) in 2..16.  Every pass but the
// last writes back to the buffers; the last hands its outputs to last(q, index, value) and the
// caller synchronises after it.  M <= 1024 Q.
template <typename T, int NB, int Q, typename Last>
__device__ __forceinline__ void fft_stk(cplx<T>* a, int M, int logM, const cplx<T>* __restrict__ tw, long maxM,
                                        bool inv, Last&& last) {
    constexpr int U16 = Q >= 4 ? Q / 4 : 1, U8 = Q >= 2 ? Q / 2 : 1, U4 = Q, U2 = 2 * Q;
    const int pitch = fft_pitch(M);
    auto back = [&](int q, int i, cplx<T> z) { a[q * pitch + sidx(i)] = z; };
    int Ns = 1, rem = logM;
    for (; rem > 4; rem -= 4, Ns <<= 4) {
        stk_pass<T, NB, 16, U16>(a, M, Ns, tw, maxM, inv, back);
        __syncthreads();
    }
    if (rem == 4) stk_pass<T, NB, 16, U16>(a, M, Ns, tw, maxM, inv, last);
    else if (rem == 3) stk_pass<T, NB, 8, U8>(a, M, Ns, tw, maxM, inv, last);
    else if (rem == 2) stk_pass<T, NB, 4, U4>(a, M, Ns, tw, maxM, inv, last);
    else stk_pass<T, NB, 2, U2>(a, M, Ns, tw, maxM, inv, last);
}

// phases [ring][mmax+1][ncol], already rotated by exp(i m phi0) -> maps [col][npix].
// One block per (ring of ring_list, group of 2 NB columns).  Columns 2q and 2q+1 of the
// group share one complex transform: with A_j, B_j their Hermitian folds
// (A_j = sum_{m = j mod n} F_m + sum_{m = -j mod n, m > 0} conj(F_m), Im F_0 dropped),
// the inverse DFT of A + i B is f_2q + i f_2q+1, both real.  m is folded modulo nphi in
// rounds of nphi consecutive values (each round hits every j once, so no two threads
// add to one element), direct and mirrored terms in separate passes.  Then an LDS
// radix-2^2 FFT (nphi a power of two) or a Bluestein chirp-z transform; with STK the
// self-sorting radix-16 FFT instead, whose last pass writes the maps (or, forward in Bluestein,
// multiplies by the chirp kernel's transform).
template <typename T, int NB, int Q, bool STK>
__global__ __launch_bounds__(256) void k_ring_synth(ShtDev<T> d, const int* __restrict__ ring_list, int ncol,
                                                    const cplx<T>* __restrict__ phase, T* __restrict__ map,
                                                    long map_stride) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
    constexpr int G = 2 * NB;
    const int ngrp = (ncol + G - 1) / G;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int grp = lb % ngrp;
    const int ring = ring_list[lb / ngrp];
    const RingInfo ri = d.rings[ring];
    const int n = ri.nphi, M = ri.M, logM = ri.logM;
    const bool pow2 = ri.voff < 0;
    const int tid = threadIdx.x;
    const int c0 = grp * G;
    const cplx<T>* P = phase + syn_phase_off(ring, 0, d.mmax, ncol) + c0;
    const size_t mstride = (size_t)PRT * ncol;
    __shared__ cplx<T> part[256];

    const int pitch = fft_pitch(M);
    if (!pow2) {  // Bluestein: the input is zero-padded from n to M
        for (int i = tid; i < NB * pitch; i += 256) buf[i] = cplx<T>{T(0), T(0)};
        __syncthreads();  // the fold below writes entries that other waves zero here
    }
    // Fold, owner computes: item (j, q) sums m = j + k n (direct) and m = (n - j) + k n or,
    // for j = 0, n + k n (mirrored, conjugated), split over S slices of k when the ring
    // has fewer than 256 items; slices are added in a fixed order.
    const int nitem = n * NB;
    const bool nofold = n > d.mmax;  // no aliasing: one direct and one mirrored term at most
    const int S = (nofold || nitem >= 256) ? 1 : 256 / nitem;
    const int step = S * n;
    if (nofold) {
        // no aliasing (n > mmax): item (j, q) takes m = j and the mirror m = n - j at most.
        // Batches of FB items issue all their loads before any use, so the global-memory
        // latency is paid once per batch rather than once per item.
        constexpr int FB = 4;
        for (int i0 = tid; i0 < nitem; i0 += 256 * FB) {
            cplx<T> A[FB], B[FB], Am[FB], Bm[FB], ch[FB];
#pragma unroll
            for (int u = 0; u < FB; u++) {
                const int i = i0 + 256 * u;
                const int j = i / NB, q = i - (i / NB) * NB;
                const bool in = i < nitem;
                const bool hasA = in && c0 + 2 * q < ncol, hasB = in && c0 + 2 * q + 1 < ncol;
                const cplx<T>* src = P + 2 * q;
                const bool dir = j <= d.mmax, mir = j > 0 && n - j <= d.mmax;
                const cplx<T> z{T(0), T(0)};
                A[u] = (hasA && dir) ? src[(size_t)j * mstride] : z;
                B[u] = (hasB && dir) ? src[(size_t)j * mstride + 1] : z;
                Am[u] = (hasA && mir) ? src[(size_t)(n - j) * mstride] : z;
                Bm[u] = (hasB && mir) ? src[(size_t)(n - j) * mstride + 1] : z;
                if (!pow2 && in) ch[u] = d.chirp[ri.coff + j];
            }
#pragma unroll
            for (int u = 0; u < FB; u++) {
                const int i = i0 + 256 * u;
                if (i >= nitem) continue;
                const int j = i / NB, q = i - (i / NB) * NB;
                cplx<T> a = A[u], b = B[u];
                if (j == 0) a.im = b.im = T(0);
                cplx<T> acc{a.re - b.im, a.im + b.re};  // A + i B
                acc.re += Am[u].re + Bm[u].im;          // conj(A) + i conj(B) of the mirror
                acc.im += Bm[u].re - Am[u].im;
                if (!pow2) acc = cmul(acc, ch[u]);  // Bluestein pre-chirp
                buf[q * pitch + (STK ? sidx(j) : pidx(j))] = acc;
            }
        }
    } else
    for (int i = tid; i < nitem * S; i += 256) {
        const int sl = i / nitem, r = i - sl * nitem;
        const int j = r / NB, q = r - j * NB;
        const bool hasA = c0 + 2 * q < ncol, hasB = c0 + 2 * q + 1 < ncol;
        const cplx<T>* src = P + 2 * q;
        cplx<T> acc{T(0), T(0)};
        for (int m = j + sl * n; m <= d.mmax; m += step) {
            cplx<T> A{T(0), T(0)}, B{T(0), T(0)};
            if (hasA) A = src[(size_t)m * mstride];
            if (hasB) B = src[(size_t)m * mstride + 1];
            if (m == 0) A.im = B.im = T(0);
            acc.re += A.re - B.im;  // A + i B
            acc.im += A.im + B.re;
        }
        for (int m = (j ? n - j : n) + sl * n; m <= d.mmax; m += step) {
            cplx<T> A{T(0), T(0)}, B{T(0), T(0)};
            if (hasA) A = src[(size_t)m * mstride];
            if (hasB) B = src[(size_t)m * mstride + 1];
            acc.re += A.re + B.im;  // conj(A) + i conj(B)
            acc.im += B.re - A.im;
        }
        if (S == 1) {
            if (!pow2) acc = cmul(acc, d.chirp[ri.coff + j]);  // Bluestein pre-chirp
            buf[q * pitch + (STK ? sidx(j) : pidx(j))] = acc;
        } else {
            part[i] = acc;
        }
    }
    if (S > 1) {
        __syncthreads();
        for (int r = tid; r < nitem; r += 256) {
            cplx<T> acc = part[r];
            for (int sl = 1; sl < S; sl++) acc = cadd(acc, part[sl * nitem + r]);
            const int j = r / NB, q = r - j * NB;
            if (!pow2) acc = cmul(acc, d.chirp[ri.coff + j]);
            buf[q * pitch + (STK ? sidx(j) : pidx(j))] = acc;
        }
    }
    __syncthreads();
    const int nc = min(G, ncol - c0);
    if constexpr (STK) {
        if (pow2) {
            fft_stk<T, NB, Q>(buf, M, logM, d.tw, d.maxM, true, [&](int q, int k, cplx<T> z) {
                if (2 * q < nc) map[(size_t)(c0 + 2 * q) * map_stride + ri.start + k] = z.re;
                if (2 * q + 1 < nc) map[(size_t)(c0 + 2 * q + 1) * map_stride + ri.start + k] = z.im;
            });
        } else {
            const cplx<T>* V = d.blue + d.blue_nat + ri.voff;
            fft_stk<T, NB, Q>(buf, M, logM, d.tw, d.maxM, false,
                              [&](int q, int i, cplx<T> z) { buf[q * pitch + sidx(i)] = cmul(z, V[i]); });
            __syncthreads();
            const T invM = T(1) / (T)M;
            fft_stk<T, NB, Q>(buf, M, logM, d.tw, d.maxM, true, [&](int q, int k, cplx<T> z) {
                if (k < n) {
                    z = cmul(z, d.chirp[ri.coff + k]);
                    if (2 * q < nc) map[(size_t)(c0 + 2 * q) * map_stride + ri.start + k] = z.re * invM;
                    if (2 * q + 1 < nc) map[(size_t)(c0 + 2 * q + 1) * map_stride + ri.start + k] = z.im * invM;
                }
            });
        }
        return;
    }
    if (pow2) {
        fft_dif_nb<T, NB, Q>(buf, M, d.tw, d.maxM, true);
        for (int i = tid; i < NB * n; i += 256) {
            const int q = i >> logM, k = i & (M - 1);  // n = M for power-of-two rings
            const cplx<T> z = buf[q * pitch + pidx(bitrev(k, logM))];
            if (2 * q < nc) map[(size_t)(c0 + 2 * q) * map_stride + ri.start + k] = z.re;
            if (2 * q + 1 < nc) map[(size_t)(c0 + 2 * q + 1) * map_stride + ri.start + k] = z.im;
        }
    } else {
        fft_dif_nb<T, NB, Q>(buf, M, d.tw, d.maxM, false);
        // pointwise product with the chirp kernel's transform; loads issued in batches of 8
        const cplx<T>* V = d.blue + ri.voff;
        for (int i0 = tid; i0 < NB * M; i0 += 256 * 8) {
            cplx<T> v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + 256 * u;
                if (i < NB * M) v[u] = V[i & (M - 1)];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + 256 * u;
                if (i < NB * M) {
                    cplx<T>& e = buf[(i >> logM) * pitch + pidx(i & (M - 1))];
                    e = cmul(e, v[u]);
                }
            }
        }
        __syncthreads();
        fft_dit_nb<T, NB, Q>(buf, M, d.tw, d.maxM, true);
        const T invM = T(1) / (T)M;
        // post-chirp and output: each chirp value is loaded once and serves the NB buffers
        for (int k0 = tid; k0 < n; k0 += 256 * 8) {
            cplx<T> ch[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int k = k0 + 256 * u;
                if (k < n) ch[u] = d.chirp[ri.coff + k];
            }
#pragma unroll
            for (int q = 0; q < NB; q++) {
                T* outA = map + (size_t)(c0 + 2 * q) * map_stride + ri.start;
                T* outB = outA + map_stride;
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int k = k0 + 256 * u;
                    if (k < n) {
                        const cplx<T> z = cmul(buf[q * pitch + pidx(k)], ch[u]);
                        if (2 * q < nc) outA[k] = z.re * invM;
                        if (2 * q + 1 < nc) outB[k] = z.im * invM;
                    }
                }
            }
        }
    }
}

// Ring synthesis with the block's phase rows staged in LDS (shapes RING_SG | NB).  Same work
// split, fold and transforms as k_ring_synth<T, NB, Q, true>, but the input is gathered first:
// every load of the block's rows (m = 0 .. mmax, its 2 NB columns) is issued before any is
// used, SG_LOADS per thread in flight, and stored to a raw LDS array [m][2 NB]; the Hermitian
// fold then reads LDS only, into registers, and once every thread has folded the FFT buffers
// take the raw array's place.  (The fold of k_ring_synth reads global memory in dependent
// rounds, several per item on rings shorter than mmax, so its waves mostly waited on memory.)
// Bluestein pre-chirps are loaded together with the rows.  BLUE: the launch's ring class
// (Bluestein rings, nphi <= (M + 1) / 2, or power-of-two rings, nphi = M); IT: fold items per
// thread for the class's longest ring.
constexpr int SG_LOADS = 16;

template <int NB, int Q, bool BLUE>
__host__ __device__ constexpr int sg_items() {
    return BLUE ? 2 * NB * Q : 4 * NB * Q;
}

template <typename T>
__device__ __forceinline__ void ld_pair(const cplx<T>* p, cplx<T>& a, cplx<T>& b) {
    if constexpr (sizeof(T) == 4) {
        const float4 v = *reinterpret_cast<const float4*>(p);  // p is 16-byte aligned
        a = cplx<T>{v.x, v.y};
        b = cplx<T>{v.z, v.w};
    } else {
        a = p[0];
        b = p[1];
    }
}

// The gNL-loc sum fused into the ring synthesis (GNL): the columns are interleaved as
// (P_r, Q_r), so the complex transform of column pair r gives P_r + i Q_r on the ring, and
// instead of writing the maps each block adds w_r [c4 P^3 Q + c2 (P^2 U_r + P Q V_r)] over its
// pixels (FP64; U, V may be null) and writes the sum to part[block].
struct GnlArgs {
    const double* U = nullptr;  // (nr, npix) Monte-Carlo sums, or null
    const double* V = nullptr;
    const double* w = nullptr;  // (nr,) radial weights
    // (nr,) or null: the columns hold a = s_r P_r and b = Q_r / s_r (magnitudes balanced, since the
    // two real columns of one complex FFT share its absolute rounding error), inv_s2 = 1 / s_r^2:
    // P Q = a b and P^2 = a^2 inv_s2
    const double* inv_s2 = nullptr;
    double c4 = 0.0, c2 = 0.0;
    double* part = nullptr;     // one partial sum per block of the launch
    long npix = 0;
};

template <typename T, int NB, int Q, bool BLUE, bool GNL>
__global__ __launch_bounds__(256) void k_ring_synth_sg(ShtDev<T> d, const int* __restrict__ ring_list, int ncol,
                                                       const cplx<T>* __restrict__ phase, T* __restrict__ map,
                                                       long map_stride, GnlArgs g) {
    extern __shared__ __align__(16) unsigned char sg_smem[];
    cplx<T>* lds = reinterpret_cast<cplx<T>*>(sg_smem);  // raw rows, then the FFT buffers
    constexpr int G = 2 * NB, IT = sg_items<NB, Q, BLUE>();
    const int ngrp = (ncol + G - 1) / G;
    const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int grp = lb % ngrp;
    const int ring = ring_list[lb / ngrp];
    const RingInfo ri = d.rings[ring];
    const int n = ri.nphi, M = ri.M, logM = ri.logM;
    const int tid = threadIdx.x;
    const int c0 = grp * G, gv = min(G, ncol - c0);
    const cplx<T>* P = phase + syn_phase_off(ring, 0, d.mmax, ncol) + c0;
    const size_t mstride = (size_t)PRT * ncol;
    const int nm = d.mmax + 1;
    __shared__ cplx<T> part[256];

    const int nitem = n * NB;
    const bool nofold = n > d.mmax;  // no aliasing: one direct and one mirrored term at most
    const int S = (nofold || nitem >= 256) ? 1 : 256 / nitem;

    // Bluestein pre-chirps of the items this thread writes (item tid + 256 u: ring entry j = item / NB)
    cplx<T> ch[IT];
    if constexpr (BLUE) {
#pragma unroll
        for (int u = 0; u < IT; u++) {
            const int i = tid + 256 * u;
            if (i < nitem) ch[u] = d.chirp[ri.coff + i / NB];
        }
    }
    // 1. raw rows: entry e = m G + c holds column c0 + c of order m
    const int total = nm * G;
    for (int e0 = 0; e0 < total; e0 += 256 * SG_LOADS) {
        cplx<T> v[SG_LOADS];
#pragma unroll
        for (int u = 0; u < SG_LOADS; u++) {
            const int e = e0 + tid + 256 * u;
            const int m = e / G, c = e % G;
            v[u] = (e < total && c < gv) ? P[(size_t)m * mstride + c] : cplx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < SG_LOADS; u++) {
            const int e = e0 + tid + 256 * u;
            if (e < total) lds[e] = v[u];
        }
    }
    __syncthreads();

    // 2. fold, owner computes: item (j, q) sums m = j + k n (direct) and m = (n - j) + k n or,
    // for j = 0, n + k n (mirrored, conjugated), split over S slices of k when the ring has
    // fewer than 256 items (then one item per thread; slices are added in a fixed order)
    cplx<T> acc[IT];
    const int step = S * n;
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int i = tid + 256 * u;
        acc[u] = cplx<T>{T(0), T(0)};
        if (i < nitem * S) {
            const int sl = i / nitem, r = i - sl * nitem;
            const int j = r / NB, q = r % NB;
            const cplx<T>* rq = lds + 2 * q;
            cplx<T> s{T(0), T(0)};
            for (int m = j + sl * n; m < nm; m += step) {
                cplx<T> A, B;
                ld_pair(rq + m * G, A, B);
                if (m == 0) A.im = B.im = T(0);
                s.re += A.re - B.im;  // A + i B
                s.im += A.im + B.re;
            }
            for (int m = (j ? n - j : n) + sl * n; m < nm; m += step) {
                cplx<T> A, B;
                ld_pair(rq + m * G, A, B);
                s.re += A.re + B.im;  // conj(A) + i conj(B)
                s.im += B.re - A.im;
            }
            acc[u] = s;
        }
    }
    if (S > 1) {
        part[tid] = acc[0];
        __syncthreads();  // also: every read of the raw rows is done
        if (tid < nitem) {
            cplx<T> s = part[tid];
            for (int sl = 1; sl < S; sl++) s = cadd(s, part[sl * nitem + tid]);
            acc[0] = s;
        }
    } else {
        __syncthreads();  // every read of the raw rows is done before the buffers overwrite them
    }
    // 3. the FFT buffers (natural order at sidx), Bluestein pre-chirp and zero padding
    const int pitch = fft_pitch(M);
#pragma unroll
    for (int u = 0; u < IT; u++) {
        const int i = tid + 256 * u;
        if (i < nitem) {
            const int j = i / NB, q = i % NB;
            cplx<T> a = acc[u];
            if constexpr (BLUE) a = cmul(a, ch[u]);
            lds[q * pitch + sidx(j)] = a;
        }
    }
    if constexpr (BLUE) {
        const int pad = M - n;
        for (int i = tid; i < NB * pad; i += 256) {
            const int q = i / pad, j = n + (i - q * pad);
            lds[q * pitch + sidx(j)] = cplx<T>{T(0), T(0)};
        }
    }
    __syncthreads();
    const int nc = min(G, ncol - c0);
    double gacc = 0.0;
    // one output value pair (columns c0 + 2q, c0 + 2q + 1) at ring entry k
    auto emit = [&](int q, int k, T a, T b) {
        if constexpr (GNL) {
            if (2 * q + 1 < nc) {
                // w_r [c4 P^3 Q + c2 (P^2 U + P Q V)] = P^2 (w c4 P Q + w c2 U) + w c2 P Q V
                const int r = (c0 >> 1) + q;
                const double wc4 = g.w[r] * g.c4, wc2 = g.w[r] * g.c2;
                const double P = (double)a, pq = P * (double)b, p2 = P * P * (g.inv_s2 ? g.inv_s2[r] : 1.0);
                if (g.U) {
                    const long off = (long)r * g.npix + ri.start;  // uniform across the block
                    const double* Ur = g.U + off;
                    const double* Vr = g.V + off;
                    gacc = fma(p2, fma(wc4, pq, wc2 * Ur[k]), gacc);
                    gacc = fma(wc2 * pq, Vr[k], gacc);
                } else {
                    gacc = fma(p2, wc4 * pq, gacc);
                }
            }
        } else {
            if (2 * q < nc) map[(size_t)(c0 + 2 * q) * map_stride + ri.start + k] = a;
            if (2 * q + 1 < nc) map[(size_t)(c0 + 2 * q + 1) * map_stride + ri.start + k] = b;
        }
    };
    if constexpr (!BLUE) {
        fft_stk<T, NB, Q>(lds, M, logM, d.tw, d.maxM, true, [&](int q, int k, cplx<T> z) { emit(q, k, z.re, z.im); });
    } else {
        const cplx<T>* V = d.blue + d.blue_nat + ri.voff;
        fft_stk<T, NB, Q>(lds, M, logM, d.tw, d.maxM, false,
                          [&](int q, int i, cplx<T> z) { lds[q * pitch + sidx(i)] = cmul(z, V[i]); });
        __syncthreads();
        const T invM = T(1) / (T)M;
        fft_stk<T, NB, Q>(lds, M, logM, d.tw, d.maxM, true, [&](int q, int k, cplx<T> z) {
            if (k < n) {
                z = cmul(z, d.chirp[ri.coff + k]);
                emit(q, k, z.re * invM, z.im * invM);
            }
        });
    }
    if constexpr (GNL) {
        // fixed-order block sum: lanes by shuffles, then the four waves
        double s = gacc;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        double* red = reinterpret_cast<double*>(part);  // the fold's partials are no longer read
        __syncthreads();
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) g.part[lb] = (red[0] + red[1]) + (red[2] + red[3]);
    }
}

// maps [col][npix] -> phases [col][mmax+1][nrings]: sum_k f_k exp(-i m phi_k).  One block per
// (ring, pair of columns a, b): the two real rings share one complex transform of
// z = f_a + i f_b (loads issued in batches of 8), by the self-sorting radix-16 LDS FFT
// (power-of-two rings) or a Bluestein chirp-z transform built on it, giving D_j = A_j + i B_j
// with A, B the DFTs of f_a, f_b.  Since f_a, f_b are real, A_j = (D_j + conj D_{n-j}) / 2 and
// B_j = (D_j - conj D_{n-j}) / 2i; then the m rows with the phi0 rotation.  An odd last column
// runs with f_b = 0.  Q: butterfly slots per thread for the plan's longest transform
// (M <= 1024 Q).
template <typename T, int Q>
__global__ __launch_bounds__(256) void k_ring_adjoint(ShtDev<T> d, int ncol, const T* __restrict__ map,
                                                      long map_stride, cplx<T>* __restrict__ phase) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw);
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int ring = lb % d.nrings;
    const int ca = 2 * (int)(lb / d.nrings);
    const bool hasB = ca + 1 < ncol;
    const RingInfo ri = d.rings[ring];
    const int n = ri.nphi, M = ri.M, logM = ri.logM;
    const bool pow2 = ri.voff < 0;
    const int tid = threadIdx.x;
    const T* inA = map + (size_t)ca * map_stride + ri.start;
    const T* inB = inA + map_stride;
    const long phq = ri.phi_q, php = ri.phi_p;

    // Bluestein evaluates g_j = sum_k w_k exp(+2 pi i jk/n) and takes conj(g_j); with
    // w = conj(z) = f_a - i f_b that is D_j.
    const T sgnB = pow2 ? T(1) : T(-1);
    for (int k0 = tid; k0 < M; k0 += 256 * 8) {
        T x[8], y[8];
        cplx<T> ch[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + 256 * u;
            x[u] = k < n ? inA[k] : T(0);
            y[u] = (hasB && k < n) ? inB[k] : T(0);
            if (!pow2 && k < n) ch[u] = d.chirp[ri.coff + k];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int k = k0 + 256 * u;
            if (k < M) {
                cplx<T> v{x[u], sgnB * y[u]};
                if (!pow2 && k < n) v = cmul(v, ch[u]);
                buf[sidx(k)] = v;
            }
        }
    }
    __syncthreads();
    auto back = [&](int, int i, cplx<T> z) { buf[sidx(i)] = z; };
    if (pow2) {
        fft_stk<T, 1, Q>(buf, M, logM, d.tw, d.maxM, false, back);
    } else {
        // the forward transform's last pass multiplies by the chirp kernel's transform
        const cplx<T>* V = d.blue + d.blue_nat + ri.voff;
        fft_stk<T, 1, Q>(buf, M, logM, d.tw, d.maxM, false,
                         [&](int, int i, cplx<T> z) { buf[sidx(i)] = cmul(z, V[i]); });
        __syncthreads();
        fft_stk<T, 1, Q>(buf, M, logM, d.tw, d.maxM, true, back);
    }
    __syncthreads();
    cplx<T>* PA = phase + ((size_t)ca * (d.mmax + 1)) * d.nrings + ring;
    cplx<T>* PB = PA + (size_t)(d.mmax + 1) * d.nrings;
    const T invM = T(1) / (T)M;
    for (int m0 = tid; m0 <= d.mmax; m0 += 256 * 8) {
        cplx<T> c1[8], c2[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int m = m0 + 256 * u;
            if (!pow2 && m <= d.mmax) {
                const int j = m % n;
                c1[u] = d.chirp[ri.coff + j];
                c2[u] = d.chirp[ri.coff + (j ? n - j : 0)];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int m = m0 + 256 * u;
            if (m <= d.mmax) {
                const int j = m % n, j2 = j ? n - j : 0;
                cplx<T> D1 = buf[sidx(j)], D2 = buf[sidx(j2)];
                if (!pow2) {
                    const cplx<T> g1 = cmul(D1, c1[u]), g2 = cmul(D2, c2[u]);
                    D1 = cplx<T>{g1.re * invM, -g1.im * invM};
                    D2 = cplx<T>{g2.re * invM, -g2.im * invM};
                }
                const cplx<T> rot = expi_pi<T>(-(long)m * php, phq);
                const cplx<T> A{T(0.5) * (D1.re + D2.re), T(0.5) * (D1.im - D2.im)};
                PA[(size_t)m * d.nrings] = cmul(A, rot);
                if (hasB) {
                    const cplx<T> B{T(0.5) * (D1.im + D2.im), T(0.5) * (D2.re - D1.re)};
                    PB[(size_t)m * d.nrings] = cmul(B, rot);
                }
            }
        }
    }
}

// ------------------------------------------------ Legendre adjoint (reduction)
// Block-wide sum of V partial values per thread: red[v][t] holds thread t's value v;
// TPV = NT / V consecutive threads sum one row and combine with shuffles.  The row
// pitch NT + TPV keeps the strided reads free of bank conflicts.  The caller writes red
// and calls this between two barriers; the sum of row v is returned to thread
// t = v * TPV.
template <typename T, int NT, int V>
__device__ __forceinline__ T red_rows(T (*red)[NT + NT / V], int tid) {
    constexpr int TPV = NT / V;
    static_assert(TPV >= 1 && TPV <= 64 && (TPV & (TPV - 1)) == 0, "bad reduction shape");
    const int row = tid / TPV, part = tid % TPV;
    T s = T(0);
#pragma unroll 4
    for (int i = part; i < NT; i += TPV) s += red[row][i];
#pragma unroll
    for (int o = TPV / 2; o >= 1; o >>= 1) s += __shfl_down(s, o, TPV);
    return s;
}

// One chunk of LC values of l for NRT ring pairs (spin 0): acc[c] += sum_q lam_l(q) E_q,
// E = even / odd phase combination by the parity of c.  FAST: every recursion has k == 0
// (no rescale test, no scale factor); FULL: all LC values of l exist.
template <typename T, int NRT, int BC, int LC, bool FAST, bool FULL>
__device__ __forceinline__ void adj_s0_chunk(cplx<T> (&acc)[LC][BC], Rec (&r)[NRT], const double (&x)[NRT],
                                             const cplx<T> (&ev)[NRT][BC], const cplx<T> (&od)[NRT][BC],
                                             const double* __restrict__ sco, int nvalid) {
#pragma unroll
    for (int c = 0; c < LC; c++) {
        if (FULL || c < nvalid) {
            const double al = sco[2 * c], ga = sco[2 * c + 1];
#pragma unroll
            for (int q = 0; q < NRT; q++) {
                const T v = FAST ? (T)r[q].f : (T)(r[q].f * r[q].cf);
#pragma unroll
                for (int j = 0; j < BC; j++) {
                    const cplx<T> e = (c & 1) ? od[q][j] : ev[q][j];
                    acc[c][j].re += v * e.re;
                    acc[c][j].im += v * e.im;
                }
                rec_fast(r[q], al * x[q], ga);
            }
        }
    }
    if (!FAST)
#pragma unroll
        for (int q = 0; q < NRT; q++) rec_rescale(r[q]);
}

// ------------------------------------------------ Legendre adjoint, spin 0
// phases [col][mmax+1][nrings] -> alm [col][nlm] (scaled), or, with nsplit > 1, the
// unscaled partial sums over ring-pair slice s into part[s][col][nlm].  One block of
// NW waves per (m, column chunk, ring slice); thread t owns ring pairs
// s * pps + t + 64 NW q (q < NRT) whose recursions run as independent chains; the
// partial sums over a thread's rings are reduced across the block every LC values of l.
template <typename T, int NRT, int NW, int BC, int LC>
__global__ __launch_bounds__(64 * NW) void k_leg_adj_s0(ShtDev<T> d, int ncol, const cplx<T>* __restrict__ phase,
                                                        cplx<T>* __restrict__ alm, long alm_stride, T scale,
                                                        int nsplit, int pps) {
    constexpr int NT = 64 * NW;
    constexpr int V = LC * BC * 2;
    constexpr int TPV = NT / V;
    static_assert(LCH % LC == 0 && (LC & 1) == 0, "LC must be even and divide LCH");
    __shared__ T red[V][NT + TPV];
    __shared__ double sco[LCH * 2];
    const int nchunk = (ncol + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int sl = (lb / nchunk) % nsplit;
    const int m = m_of_q(lb / (nchunk * nsplit), d.mmax);
    const int tid = threadIdx.x;
    const size_t pstride = (size_t)(d.mmax + 1) * d.nrings;
    const int jp0 = sl * pps, jp1 = min(d.npairs, jp0 + pps);
    cplx<T>* out = nsplit > 1 ? alm + (size_t)sl * ncol * alm_stride : alm;
    const T osc = nsplit > 1 ? T(1) : scale;

    double x[NRT];
    Rec r[NRT];
    cplx<T> ev[NRT][BC], od[NRT][BC];
#pragma unroll
    for (int q = 0; q < NRT; q++) {
        const int jp = jp0 + tid + NT * q;
        const bool ok = jp < jp1;
        const int jpc = ok ? jp : jp0;
        x[q] = d.cth[jpc];
        rec_seed(r[q], d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
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
            // idle slots carry zero phases: their (finite) recursion values never matter
            ev[q][j] = cadd(pn, ps);
            od[q][j] = csub(pn, ps);
        }
    }
    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const double* cf4 = d.coef + 4 * base;
    // coefficients of the next chunk are fetched into registers while the current one runs
    double pal = 0.0, pga = 0.0;
    auto fetch = [&](int l0n) {
        if (tid < min(LCH, d.lmax + 1 - l0n)) {
            pal = cf4[4 * (l0n + tid)];
            pga = cf4[4 * (l0n + tid) + 2];
        }
    };
    fetch(m);
    for (int l0 = m; l0 <= d.lmax; l0 += LCH) {
        const int n = min(LCH, d.lmax + 1 - l0);
        __syncthreads();
        if (tid < n) {
            sco[2 * tid] = pal;
            sco[2 * tid + 1] = pga;
        }
        __syncthreads();
        if (l0 + LCH <= d.lmax) fetch(l0 + LCH);
        bool allk0 = true;
#pragma unroll
        for (int q = 0; q < NRT; q++) allk0 = allk0 && (r[q].k == 0);
        const bool fast = __all(allk0);
        for (int lc = 0; lc < n; lc += LC) {
            cplx<T> acc[LC][BC];
#pragma unroll
            for (int c = 0; c < LC; c++)
#pragma unroll
                for (int j = 0; j < BC; j++) acc[c][j] = cplx<T>{T(0), T(0)};
            const int nv = n - lc;
            if (nv >= LC) {
                if (fast) adj_s0_chunk<T, NRT, BC, LC, true, true>(acc, r, x, ev, od, sco + 2 * lc, nv);
                else adj_s0_chunk<T, NRT, BC, LC, false, true>(acc, r, x, ev, od, sco + 2 * lc, nv);
            } else {
                if (fast) adj_s0_chunk<T, NRT, BC, LC, true, false>(acc, r, x, ev, od, sco + 2 * lc, nv);
                else adj_s0_chunk<T, NRT, BC, LC, false, false>(acc, r, x, ev, od, sco + 2 * lc, nv);
            }
#pragma unroll
            for (int c = 0; c < LC; c++)
#pragma unroll
                for (int j = 0; j < BC; j++) {
                    red[(c * BC + j) * 2][tid] = acc[c][j].re;
                    red[(c * BC + j) * 2 + 1][tid] = acc[c][j].im;
                }
            __syncthreads();
            const T sum = red_rows<T, NT, V>(red, tid);
            if (tid % TPV == 0) {
                const int row = tid / TPV;
                const int c = row / (2 * BC), j = (row >> 1) % BC, part = row & 1;
                const int l = l0 + lc + c, col = chunk * BC + j;
                if (c < nv && col < ncol)
                    reinterpret_cast<T*>(out + (size_t)col * alm_stride + base + l)[part] = sum * osc;
            }
            __syncthreads();
        }
    }
}

// ------------------------------------- Legendre adjoint, spin 0, FP32 matrix cores
// For one m, alm(l, col) = sum over ring pairs of lam_lm(ring) (P_north + P_south)(col) for
// even l - m and lam_lm(ring) (P_north - P_south)(col) for odd l - m: a contraction over
// rings, run on v_mfma_f32_32x32x2f32 with A = lam (32 values of l of one parity x 2 rings),
// B = the summed / differenced phases (2 rings x 32 real columns), D(l, col) += A B (lane
// layout as in the synthesis kernel).  Each lane runs the FP64 recursion of one ring pair
// for a tile of 64 values of l and stores the values in its wave's LDS tile
// [parity][l / 2][ring], rows XOR-swizzled by l / 2 so that the row-wise stores and the
// transposed operand loads both hit 32 distinct banks per half-wave.  The B operands (32 ring
// pairs x 2 per lane) stay in registers for the whole l loop.  One block = 4 waves = 256
// ring pairs (a ring slice), one m, 16 complex columns; the waves' sums meet in LDS and
// slices write partial sums that k_adj_sum_splits adds in a fixed order.
constexpr int ADJM_PAIRS = 256;

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(256) void k_leg_adj_s0_mfma(ShtDev<float> d, int ncol,
                                                         const cplx<float>* __restrict__ phase,
                                                         cplx<float>* __restrict__ out, long out_stride, float osc,
                                                         int nslice) {
    __shared__ __align__(16) float T[4 * 2 * 32 * 64];  // lam tiles, then the waves' sums [w][par][i][33]
    const int nchunk = (ncol + 15) / 16;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int sl = (lb / nchunk) % nslice;
    const int m = m_of_q(lb / (nchunk * nslice), d.mmax);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, li32 = lane & 31;
    const int wring = sl * ADJM_PAIRS + wave * 64;  // first ring pair of this wave
    cplx<float>* dst = nslice > 1 ? out + (size_t)sl * ncol * out_stride : out;

    // recursion of ring pair wring + lane
    const int jr = wring + lane;
    const int jrc = jr < d.npairs ? jr : d.npairs - 1;
    const double x = d.cth[jrc];
    Rec r;
    rec_seed(r, d.seedL[m] + d.seedA[m] * d.l2sh[jrc] + d.seedB[m] * d.l2ch[jrc], d.seedS[m]);

    // B operands: ring pair wring + 2 s + h, real column li32 (complex column li32 / 2)
    float bev[32], bod[32];
    {
        const int col = chunk * 16 + (li32 >> 1), part = li32 & 1;
        const float* P = reinterpret_cast<const float*>(phase + (size_t)col * (d.mmax + 1) * d.nrings +
                                                        (size_t)m * d.nrings);
#pragma unroll
        for (int s2 = 0; s2 < 32; s2++) {
            const int jb = wring + 2 * s2 + h;
            float pn = 0.f, ps = 0.f;
            if (jb < d.npairs && col < ncol) {
                const int rs = d.nrings - 1 - jb;
                pn = P[2 * jb + part];
                if (rs != jb) ps = P[2 * rs + part];
            }
            bev[s2] = pn + ps;
            bod[s2] = pn - ps;
        }
    }

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const double* cf4 = d.coef + 4 * base;
    // a block whose ring pairs are all dead at this m writes zeros without the recursion
    const bool skip = __syncthreads_and(jr >= d.npairs || d.dead[(size_t)m * d.npairs + jrc]);
    float* Tw = T + wave * 4096;
    bool fast = false;
    for (int l0 = m; l0 <= d.lmax; l0 += 64) {
        const int n = min(64, d.lmax + 1 - l0);
        v16f De = v16f{0}, Do = v16f{0};
        if (!skip) {
            // lane c holds (alpha, gamma) of l0 + c
            double pal = 0.0, pga = 0.0;
            if (lane < n) {
                pal = cf4[4 * (l0 + lane)];
                pga = cf4[4 * (l0 + lane) + 2];
            }
            if (!fast) fast = __all(r.k == 0);
            if (fast && n == 64) {
#pragma unroll 4
                for (int c = 0; c < 64; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 2048 + i * 64 + (lane ^ i)] = (float)r.f;
                    rec_fast(r, readlane_f64(pal, c) * x, readlane_f64(pga, c));
                }
            } else if (n == 64) {
                // full tile while some lane still carries a scale exponent: the loop below
                // without its bounds tests, one exponent update per 4 values of l
#pragma unroll 4
                for (int c = 0; c < 64; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 2048 + i * 64 + (lane ^ i)] = (float)(r.f * r.cf);
                    rec_fast(r, readlane_f64(pal, c) * x, readlane_f64(pga, c));
                    if ((c & 3) == 3) rec_rescale(r);
                }
            } else {
                for (int c = 0; c < 64; c++) {
                    const int i = c >> 1;
                    float v = 0.f;
                    if (c < n) {
                        v = fast ? (float)r.f : (float)(r.f * r.cf);
                        rec_fast(r, readlane_f64(pal, c) * x, readlane_f64(pga, c));
                        if (!fast && (c & 3) == 3) rec_rescale(r);
                    }
                    Tw[(c & 1) * 2048 + i * 64 + (lane ^ i)] = v;
                }
            }
            __syncthreads();  // the wave's tile is complete
#pragma unroll
            for (int s2 = 0; s2 < 32; s2++) {
                const int rl = 2 * s2 + h;
                const float aev = Tw[li32 * 64 + (rl ^ li32)];
                const float aod = Tw[2048 + li32 * 64 + (rl ^ li32)];
                De = __builtin_amdgcn_mfma_f32_32x32x2f32(aev, bev[s2], De, 0, 0, 0);
                Do = __builtin_amdgcn_mfma_f32_32x32x2f32(aod, bod[s2], Do, 0, 0, 0);
            }
        }
        __syncthreads();  // every wave has read its tile: reuse the space for the sums
        // D[i][j] at lane j + 32 ((i / 4) % 2), register 4 (i / 8) + i % 4; i = (l - l0) / 2
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int i = 8 * (v >> 2) + 4 * h + (v & 3);
            T[((wave * 2) * 32 + i) * 33 + li32] = De[v];
            T[((wave * 2 + 1) * 32 + i) * 33 + li32] = Do[v];
        }
        __syncthreads();
        {
            // thread: real column j, l = l0 + 8 (tid % 8) + u
            const int j = tid >> 3, cc = chunk * 16 + (j >> 1), pp = j & 1;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int ll = (tid & 7) * 8 + u, par = ll & 1, i = ll >> 1;
                float sum = T[((0 * 2 + par) * 32 + i) * 33 + j];
                sum += T[((1 * 2 + par) * 32 + i) * 33 + j];
                sum += T[((2 * 2 + par) * 32 + i) * 33 + j];
                sum += T[((3 * 2 + par) * 32 + i) * 33 + j];
                if (ll < n && cc < ncol)
                    reinterpret_cast<float*>(dst + (size_t)cc * out_stride + base + l0 + ll)[pp] = sum * osc;
            }
        }
        __syncthreads();  // the sums are read before the next tile overwrites them
    }
}

// ------------------------------------- Legendre adjoint, spin s, FP32 matrix cores
// Per map, with R, I the phases of the two real maps summed (+) and differenced (-) over
// north / south rings, lam+- = (f+ +- f-) / 2 the two recursions' combinations, and for
// l + m + s even (Rp, Rm, Ip, Im) = (R+, R-, I+, I-), for odd (R-, R+, I-, I+):
//   G += -(lam+ Rp + i lam- Im),   C += -(lam+ Ip - i lam- Rm).
// Written as f+ B+ + f- B- with B+- = (Bp +- Bn) / 2, where Bp holds the lam+ factors and
// Bn the lam- factors of the four real outputs (G re, G im, C re, C im) of a map:
//   even: Bp = -(R+re, R+im, I+re, I+im), Bn = (I-im, -I-re, -R-im, R-re)
//   odd:  Bp = -(R-re, R-im, I-re, I-im), Bn = (I+im, -I+re, -R+im, R+re)
// so both recursions feed the matrix cores directly: lanes 0-31 of a wave run f+ and lanes
// 32-63 f- of the wave's 32 ring pairs, and per 64 values of l a wave issues 64
// v_mfma_f32_32x32x2f32 (A = f+- of 32 values of l of one parity x 2 rings, B = B+- of 2 rings
// x 32 real columns = 8 maps).  One block = 8 waves = 256 ring pairs, one m, 8 maps.
// NC > 1: NC chunks of 8 maps per block share the recursion tiles (their MFMAs run from the same
// A operands, each chunk's sums pass through LDS in turn); per chunk the products accumulate in
// the order of NC = 1, so the results are the same.
constexpr int ADJM_SP_PAIRS = 256;

template <int NC>
__global__ __launch_bounds__(512) void k_leg_adj_sp_mfma(ShtDev<float> d, int nmap,
                                                         const cplx<float>* __restrict__ phase,
                                                         cplx<float>* __restrict__ out, long out_stride, float osc,
                                                         int nslice) {
    // f tiles [wave][sign][parity][l / 2][ring ^ (l / 2)], then the waves' sums [w][par][i][33]
    __shared__ __align__(16) float T[8 * 4 * 32 * 32];
    const int nchunk = (nmap + 8 * NC - 1) / (8 * NC);
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int sl = (lb / nchunk) % nslice;
    const int m = m_of_q(lb / (nchunk * nslice), d.mmax);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, li32 = lane & 31;
    const int M1 = d.mmax + 1;
    const int wring = sl * ADJM_SP_PAIRS + wave * 32;  // first ring pair of this wave
    cplx<float>* dst = nslice > 1 ? out + (size_t)sl * 2 * nmap * out_stride : out;

    // recursion f+ (h = 0) or f- (h = 1) of ring pair wring + li32
    const int jr = wring + li32;
    const int jrc = jr < d.npairs ? jr : d.npairs - 1;
    const double x = d.cth[jrc];
    const double sgb = h ? -1.0 : 1.0;
    Rec r;
    rec_seed(r, d.seedL[h * M1 + m] + d.seedA[h * M1 + m] * d.l2sh[jrc] + d.seedB[h * M1 + m] * d.l2ch[jrc],
             d.seedS[h * M1 + m]);

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const int lstart = m > d.spin ? m : d.spin;
    const int p0 = (lstart + m + d.spin) & 1;  // tile rows of parity p0 are the "odd" type

    // B operands of ring pair wring + 2 s + h, real column li32 = 4 (map) + q:
    // bp0 / bm0 for rows of even l - lstart, bp1 / bm1 for odd
    float bp0[NC][16], bm0[NC][16], bp1[NC][16], bm1[NC][16];
#pragma unroll
    for (int cc = 0; cc < NC; cc++) {
        const int b = (chunk * NC + cc) * 8 + (li32 >> 2), q = li32 & 3;
        const size_t pstride = (size_t)M1 * d.nrings;
        const cplx<float>* PR = phase + (size_t)(2 * b) * pstride + (size_t)m * d.nrings;
        const cplx<float>* PI = PR + pstride;
#pragma unroll
        for (int s2 = 0; s2 < 16; s2++) {
            const int jb = wring + 2 * s2 + h;
            cplx<float> Rn{0.f, 0.f}, Rs{0.f, 0.f}, In{0.f, 0.f}, Is{0.f, 0.f};
            if (jb < d.npairs && b < nmap) {
                const int rs = d.nrings - 1 - jb;
                Rn = PR[jb];
                In = PI[jb];
                if (rs != jb) {
                    Rs = PR[rs];
                    Is = PI[rs];
                }
            }
            const cplx<float> Ra = cadd(Rn, Rs), Rd = csub(Rn, Rs), Ia = cadd(In, Is), Id = csub(In, Is);
            float pe, ne, po, no;  // Bp, Bn of the even and the odd type
            switch (q) {
                case 0: pe = -Ra.re; ne = Id.im; po = -Rd.re; no = Ia.im; break;
                case 1: pe = -Ra.im; ne = -Id.re; po = -Rd.im; no = -Ia.re; break;
                case 2: pe = -Ia.re; ne = -Rd.im; po = -Id.re; no = -Ra.im; break;
                default: pe = -Ia.im; ne = Rd.re; po = -Id.im; no = Ra.re; break;
            }
            // row parity 0 (l = lstart + even) is the even type when p0 == 0
            const float p_0 = p0 ? po : pe, n_0 = p0 ? no : ne, p_1 = p0 ? pe : po, n_1 = p0 ? ne : no;
            bp0[cc][s2] = 0.5f * (p_0 + n_0);
            bm0[cc][s2] = 0.5f * (p_0 - n_0);
            bp1[cc][s2] = 0.5f * (p_1 + n_1);
            bm1[cc][s2] = 0.5f * (p_1 - n_1);
        }
    }

    // coefficients with m <= l < spin vanish identically
    for (int i = tid; i < (lstart - m) * 16 * NC; i += 512) {
        const int l = m + i / (16 * NC), b = chunk * 8 * NC + (i % (16 * NC)) / 2, g = i % 2;
        if (l <= d.lmax && b < nmap) dst[(size_t)(2 * b + g) * out_stride + base + l] = cplx<float>{0.f, 0.f};
    }
    const double* cf4 = d.coef + 4 * base;
    const bool skip = __syncthreads_and(jr >= d.npairs || d.dead[(size_t)m * d.npairs + jrc]);
    float* Tw = T + wave * 4096 + h * 2048;  // this lane's sign
    const float* Ta = T + wave * 4096;
    bool fast = false;
    for (int l0 = lstart; l0 <= d.lmax; l0 += 64) {
        const int n = min(64, d.lmax + 1 - l0);
        v16f D0[NC], D1[NC];
#pragma unroll
        for (int cc = 0; cc < NC; cc++) D0[cc] = D1[cc] = v16f{0};
        if (!skip) {
            double pal = 0.0, pbe = 0.0, pga = 0.0;
            if (lane < n) {
                pal = cf4[4 * (l0 + lane)];
                pbe = cf4[4 * (l0 + lane) + 1];
                pga = cf4[4 * (l0 + lane) + 2];
            }
            if (!fast) fast = __all(r.k == 0);
            if (fast && n == 64) {
#pragma unroll 4
                for (int c = 0; c < 64; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 1024 + i * 32 + (li32 ^ i)] = (float)r.f;
                    rec_fast(r, readlane_f64(pal, c) * (x + sgb * readlane_f64(pbe, c)), readlane_f64(pga, c));
                }
            } else if (n == 64) {
                // full tile while some lane still carries a scale exponent: the loop below
                // without its bounds tests, one exponent update per 4 values of l
#pragma unroll 4
                for (int c = 0; c < 64; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 1024 + i * 32 + (li32 ^ i)] = (float)(r.f * r.cf);
                    rec_fast(r, readlane_f64(pal, c) * (x + sgb * readlane_f64(pbe, c)), readlane_f64(pga, c));
                    if ((c & 3) == 3) rec_rescale(r);
                }
            } else {
                for (int c = 0; c < 64; c++) {
                    const int i = c >> 1;
                    float v = 0.f;
                    if (c < n) {
                        v = fast ? (float)r.f : (float)(r.f * r.cf);
                        rec_fast(r, readlane_f64(pal, c) * (x + sgb * readlane_f64(pbe, c)), readlane_f64(pga, c));
                        if (!fast && (c & 3) == 3) rec_rescale(r);
                    }
                    Tw[(c & 1) * 1024 + i * 32 + (li32 ^ i)] = v;
                }
            }
            __syncthreads();  // the wave's tiles are complete
#pragma unroll
            for (int s2 = 0; s2 < 16; s2++) {
                const int rl = 2 * s2 + h, sw = li32 * 32 + (rl ^ li32);
                const float a0 = Ta[sw], a1 = Ta[2048 + sw], a2 = Ta[1024 + sw], a3 = Ta[2048 + 1024 + sw];
#pragma unroll
                for (int cc = 0; cc < NC; cc++) {
                    D0[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bp0[cc][s2], D0[cc], 0, 0, 0);
                    D0[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bm0[cc][s2], D0[cc], 0, 0, 0);
                    D1[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, bp1[cc][s2], D1[cc], 0, 0, 0);
                    D1[cc] = __builtin_amdgcn_mfma_f32_32x32x2f32(a3, bm1[cc][s2], D1[cc], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int cc = 0; cc < NC; cc++) {
            // every wave has read its tiles (cc = 0) or the previous chunk's sums: reuse the space
            __syncthreads();
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int i = 8 * (v >> 2) + 4 * h + (v & 3);
                T[((wave * 2) * 32 + i) * 33 + li32] = D0[cc][v];
                T[((wave * 2 + 1) * 32 + i) * 33 + li32] = D1[cc][v];
            }
            __syncthreads();
            // thread: real column j = 4 (map) + q, l = l0 + 4 (tid % 16) + u
            const int j = tid >> 4, b = (chunk * NC + cc) * 8 + (j >> 2), g = (j >> 1) & 1, pp = j & 1;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int ll = (tid & 15) * 4 + u, par = ll & 1, i = ll >> 1;
                float sum = 0.f;
#pragma unroll
                for (int w = 0; w < 8; w++) sum += T[((w * 2 + par) * 32 + i) * 33 + j];
                if (ll < n && b < nmap)
                    reinterpret_cast<float*>(dst + (size_t)(2 * b + g) * out_stride + base + l0 + ll)[pp] = sum * osc;
            }
        }
        __syncthreads();  // the sums are read before the next tile overwrites them
    }
}

// ------------------------------- Legendre adjoints, spin 0 and spin s, FP64 matrix cores
// The matrix-core adjoints above in FP64, on v_mfma_f64_16x16x4f64 (lane layout as in the
// Fisher-derivative GEMM: A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k, D[i][j] at lane
// j + 16 (i % 4), register i / 4).  Per tile of 32 values of l each lane writes its ring's FP64
// recursion values into LDS rows [parity][(l - l0) / 2][ring], the ring index XOR-swizzled by
// 4 ((l - l0) / 2) so that the row-wise stores and the reads of the A operands (16 rows x 4
// consecutive rings) spread over the banks; a wave then issues one MFMA per 4 rings and parity,
// A = lambda (16 values of l of one parity x 4 rings), B = the north/south sums or differences
// (4 rings x 16 real columns), held in registers for the whole l loop.  The waves' sums meet in
// LDS; 256-ring-pair slices write partial sums that k_adj_sum_splits adds in a fixed order.
constexpr int ADJ64_L = 32;  // values of l per tile

// Spin 0: one block = 4 waves = 256 ring pairs (one per lane), one m, 8 complex columns.
__global__ __launch_bounds__(256) void k_leg_adj_s0_mfma64(ShtDev<double> d, int ncol,
                                                           const cplx<double>* __restrict__ phase,
                                                           cplx<double>* __restrict__ out, long out_stride,
                                                           double osc, int nslice) {
    // lambda tiles [wave][parity][i][64], then the waves' sums [wave][parity][i][17]
    __shared__ __align__(16) double T[4 * 2 * 16 * 64];
    const int nchunk = (ncol + 7) / 8;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int sl = (lb / nchunk) % nslice;
    const int m = m_of_q(lb / (nchunk * nslice), d.mmax);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k4 = lane >> 4, j16 = lane & 15;
    const int wring = sl * ADJM_PAIRS + wave * 64;  // first ring pair of this wave
    cplx<double>* dst = nslice > 1 ? out + (size_t)sl * ncol * out_stride : out;

    // recursion of ring pair wring + lane
    const int jr = wring + lane;
    const int jrc = jr < d.npairs ? jr : d.npairs - 1;
    const double x = d.cth[jrc];
    Rec r;
    rec_seed(r, d.seedL[m] + d.seedA[m] * d.l2sh[jrc] + d.seedB[m] * d.l2ch[jrc], d.seedS[m]);

    // B operands: ring pair wring + 4 s + k4, real column j16 (complex column j16 / 2)
    double bev[16], bod[16];
    {
        const int col = chunk * 8 + (j16 >> 1), part = j16 & 1;
        const double* P = reinterpret_cast<const double*>(phase + (size_t)col * (d.mmax + 1) * d.nrings +
                                                          (size_t)m * d.nrings);
#pragma unroll
        for (int s = 0; s < 16; s++) {
            const int jb = wring + 4 * s + k4;
            double pn = 0.0, ps = 0.0;
            if (jb < d.npairs && col < ncol) {
                const int rs = d.nrings - 1 - jb;
                pn = P[2 * jb + part];
                if (rs != jb) ps = P[2 * rs + part];
            }
            bev[s] = pn + ps;
            bod[s] = pn - ps;
        }
    }

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const double* cf4 = d.coef + 4 * base;
    // a block whose ring pairs are all dead at this m writes zeros without the recursion
    const bool skip = __syncthreads_and(jr >= d.npairs || d.dead[(size_t)m * d.npairs + jrc]);
    double* Tw = T + wave * 2048;
    bool fast = false;
    for (int l0 = m; l0 <= d.lmax; l0 += ADJ64_L) {
        const int n = min(ADJ64_L, d.lmax + 1 - l0);
        v4d_t De = v4d_t{0.0, 0.0, 0.0, 0.0}, Do = v4d_t{0.0, 0.0, 0.0, 0.0};
        if (!skip) {
            // lane c holds (alpha, gamma) of l0 + c
            double pal = 0.0, pga = 0.0;
            if (lane < n) {
                pal = cf4[4 * (l0 + lane)];
                pga = cf4[4 * (l0 + lane) + 2];
            }
            if (!fast) fast = __all(r.k == 0);
            if (fast && n == ADJ64_L) {
#pragma unroll 4
                for (int c = 0; c < ADJ64_L; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 1024 + i * 64 + (lane ^ (4 * i))] = r.f;
                    rec_fast(r, readlane_f64(pal, c) * x, readlane_f64(pga, c));
                }
            } else if (n == ADJ64_L) {
                // full tile while some lane still carries a scale exponent
#pragma unroll 4
                for (int c = 0; c < ADJ64_L; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 1024 + i * 64 + (lane ^ (4 * i))] = r.f * r.cf;
                    rec_fast(r, readlane_f64(pal, c) * x, readlane_f64(pga, c));
                    if ((c & 3) == 3) rec_rescale(r);
                }
            } else {
                for (int c = 0; c < ADJ64_L; c++) {
                    const int i = c >> 1;
                    double v = 0.0;
                    if (c < n) {
                        v = fast ? r.f : r.f * r.cf;
                        rec_fast(r, readlane_f64(pal, c) * x, readlane_f64(pga, c));
                        if (!fast && (c & 3) == 3) rec_rescale(r);
                    }
                    Tw[(c & 1) * 1024 + i * 64 + (lane ^ (4 * i))] = v;
                }
            }
            __syncthreads();  // the wave's tile is complete
#pragma unroll
            for (int s = 0; s < 16; s++) {
                const int sw = j16 * 64 + ((4 * s + k4) ^ (4 * j16));
                De = __builtin_amdgcn_mfma_f64_16x16x4f64(Tw[sw], bev[s], De, 0, 0, 0);
                Do = __builtin_amdgcn_mfma_f64_16x16x4f64(Tw[1024 + sw], bod[s], Do, 0, 0, 0);
            }
        }
        __syncthreads();  // every wave has read its tile: reuse the space for the sums
        // D[i][j] at lane j + 16 (i % 4), register i / 4; i = (l - l0) / 2
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = 4 * v + k4;
            T[((wave * 2) * 16 + i) * 17 + j16] = De[v];
            T[((wave * 2 + 1) * 16 + i) * 17 + j16] = Do[v];
        }
        __syncthreads();
        {
            // thread: real column j, l = l0 + 2 (tid % 16) + u
            const int j = tid >> 4, cc = chunk * 8 + (j >> 1), pp = j & 1;
#pragma unroll
            for (int u = 0; u < 2; u++) {
                const int ll = (tid & 15) * 2 + u, par = ll & 1, i = ll >> 1;
                double sum = T[((0 * 2 + par) * 16 + i) * 17 + j];
                sum += T[((1 * 2 + par) * 16 + i) * 17 + j];
                sum += T[((2 * 2 + par) * 16 + i) * 17 + j];
                sum += T[((3 * 2 + par) * 16 + i) * 17 + j];
                if (ll < n && cc < ncol)
                    reinterpret_cast<double*>(dst + (size_t)cc * out_stride + base + l0 + ll)[pp] = sum * osc;
            }
        }
        __syncthreads();  // the sums are read before the next tile overwrites them
    }
}

// Spin s: as k_leg_adj_sp_mfma (lanes 0-31 of a wave run f+ and lanes 32-63 f- of the wave's
// 32 ring pairs; D += f+ B+ + f- B-), one block = 8 waves = 256 ring pairs, one m, 4 maps
// (16 real columns: G re, G im, C re, C im of each map).
// NC > 1: NC chunks of 4 maps per block share the recursion tiles; per chunk the products
// accumulate in the order of NC = 1.
template <int NC>
__global__ __launch_bounds__(512) void k_leg_adj_sp_mfma64(ShtDev<double> d, int nmap,
                                                           const cplx<double>* __restrict__ phase,
                                                           cplx<double>* __restrict__ out, long out_stride,
                                                           double osc, int nslice) {
    // f tiles [wave][sign][parity][i][32], then the waves' sums [wave][parity][i][17]
    __shared__ __align__(16) double T[8 * 2 * 2 * 16 * 32];
    const int nchunk = (nmap + 4 * NC - 1) / (4 * NC);
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int sl = (lb / nchunk) % nslice;
    const int m = m_of_q(lb / (nchunk * nslice), d.mmax);
    const int M1 = d.mmax + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int h = lane >> 5, li32 = lane & 31;
    const int k4 = lane >> 4, j16 = lane & 15;
    const int wring = sl * ADJM_SP_PAIRS + wave * 32;  // first ring pair of this wave
    cplx<double>* dst = nslice > 1 ? out + (size_t)sl * 2 * nmap * out_stride : out;

    // recursion f+ (h = 0) or f- (h = 1) of ring pair wring + li32
    const int jr = wring + li32;
    const int jrc = jr < d.npairs ? jr : d.npairs - 1;
    const double x = d.cth[jrc];
    const double sgb = h ? -1.0 : 1.0;
    Rec r;
    rec_seed(r, d.seedL[h * M1 + m] + d.seedA[h * M1 + m] * d.l2sh[jrc] + d.seedB[h * M1 + m] * d.l2ch[jrc],
             d.seedS[h * M1 + m]);

    const long base = (long)m * (2 * d.lmax + 1 - m) / 2;
    const int lstart = m > d.spin ? m : d.spin;
    const int p0 = (lstart + m + d.spin) & 1;  // tile rows of parity p0 are the "odd" type

    // B operands of ring pair wring + 4 s + k4, real column j16 = 4 (map) + q
    double bp0[NC][8], bm0[NC][8], bp1[NC][8], bm1[NC][8];
#pragma unroll
    for (int cc = 0; cc < NC; cc++) {
        const int b = (chunk * NC + cc) * 4 + (j16 >> 2), q = j16 & 3;
        const size_t pstride = (size_t)M1 * d.nrings;
        const cplx<double>* PR = phase + (size_t)(2 * b) * pstride + (size_t)m * d.nrings;
        const cplx<double>* PI = PR + pstride;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            const int jb = wring + 4 * s + k4;
            cplx<double> Rn{0.0, 0.0}, Rs{0.0, 0.0}, In{0.0, 0.0}, Is{0.0, 0.0};
            if (jb < d.npairs && b < nmap) {
                const int rs = d.nrings - 1 - jb;
                Rn = PR[jb];
                In = PI[jb];
                if (rs != jb) {
                    Rs = PR[rs];
                    Is = PI[rs];
                }
            }
            const cplx<double> Ra = cadd(Rn, Rs), Rd = csub(Rn, Rs), Ia = cadd(In, Is), Id = csub(In, Is);
            double pe, ne, po, no;  // Bp, Bn of the even and the odd type
            switch (q) {
                case 0: pe = -Ra.re; ne = Id.im; po = -Rd.re; no = Ia.im; break;
                case 1: pe = -Ra.im; ne = -Id.re; po = -Rd.im; no = -Ia.re; break;
                case 2: pe = -Ia.re; ne = -Rd.im; po = -Id.re; no = -Ra.im; break;
                default: pe = -Ia.im; ne = Rd.re; po = -Id.im; no = Ra.re; break;
            }
            const double p_0 = p0 ? po : pe, n_0 = p0 ? no : ne, p_1 = p0 ? pe : po, n_1 = p0 ? ne : no;
            bp0[cc][s] = 0.5 * (p_0 + n_0);
            bm0[cc][s] = 0.5 * (p_0 - n_0);
            bp1[cc][s] = 0.5 * (p_1 + n_1);
            bm1[cc][s] = 0.5 * (p_1 - n_1);
        }
    }

    // coefficients with m <= l < spin vanish identically
    for (int i = tid; i < (lstart - m) * 8 * NC; i += 512) {
        const int l = m + i / (8 * NC), b = chunk * 4 * NC + (i % (8 * NC)) / 2, g = i % 2;
        if (l <= d.lmax && b < nmap) dst[(size_t)(2 * b + g) * out_stride + base + l] = cplx<double>{0.0, 0.0};
    }
    const double* cf4 = d.coef + 4 * base;
    const bool skip = __syncthreads_and(jr >= d.npairs || d.dead[(size_t)m * d.npairs + jrc]);
    double* Tw = T + wave * 2048 + h * 1024;  // this lane's sign
    const double* Ta = T + wave * 2048;
    bool fast = false;
    for (int l0 = lstart; l0 <= d.lmax; l0 += ADJ64_L) {
        const int n = min(ADJ64_L, d.lmax + 1 - l0);
        v4d_t D0[NC], D1[NC];
#pragma unroll
        for (int cc = 0; cc < NC; cc++) D0[cc] = D1[cc] = v4d_t{0.0, 0.0, 0.0, 0.0};
        if (!skip) {
            double pal = 0.0, pbe = 0.0, pga = 0.0;
            if (lane < n) {
                pal = cf4[4 * (l0 + lane)];
                pbe = cf4[4 * (l0 + lane) + 1];
                pga = cf4[4 * (l0 + lane) + 2];
            }
            if (!fast) fast = __all(r.k == 0);
            if (fast && n == ADJ64_L) {
#pragma unroll 4
                for (int c = 0; c < ADJ64_L; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 512 + i * 32 + (li32 ^ ((4 * i) & 28))] = r.f;
                    rec_fast(r, readlane_f64(pal, c) * (x + sgb * readlane_f64(pbe, c)), readlane_f64(pga, c));
                }
            } else if (n == ADJ64_L) {
#pragma unroll 4
                for (int c = 0; c < ADJ64_L; c++) {
                    const int i = c >> 1;
                    Tw[(c & 1) * 512 + i * 32 + (li32 ^ ((4 * i) & 28))] = r.f * r.cf;
                    rec_fast(r, readlane_f64(pal, c) * (x + sgb * readlane_f64(pbe, c)), readlane_f64(pga, c));
                    if ((c & 3) == 3) rec_rescale(r);
                }
            } else {
                for (int c = 0; c < ADJ64_L; c++) {
                    const int i = c >> 1;
                    double v = 0.0;
                    if (c < n) {
                        v = fast ? r.f : r.f * r.cf;
                        rec_fast(r, readlane_f64(pal, c) * (x + sgb * readlane_f64(pbe, c)), readlane_f64(pga, c));
                        if (!fast && (c & 3) == 3) rec_rescale(r);
                    }
                    Tw[(c & 1) * 512 + i * 32 + (li32 ^ ((4 * i) & 28))] = v;
                }
            }
            __syncthreads();  // the wave's tiles are complete
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int sw = j16 * 32 + ((4 * s + k4) ^ ((4 * j16) & 28));
                const double a0 = Ta[sw], a1 = Ta[1024 + sw], a2 = Ta[512 + sw], a3 = Ta[1024 + 512 + sw];
#pragma unroll
                for (int cc = 0; cc < NC; cc++) {
                    D0[cc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bp0[cc][s], D0[cc], 0, 0, 0);
                    D0[cc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bm0[cc][s], D0[cc], 0, 0, 0);
                    D1[cc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, bp1[cc][s], D1[cc], 0, 0, 0);
                    D1[cc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a3, bm1[cc][s], D1[cc], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int cc = 0; cc < NC; cc++) {
            // every wave has read its tiles (cc = 0) or the previous chunk's sums: reuse the space
            __syncthreads();
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 4 * v + k4;
                T[((wave * 2) * 16 + i) * 17 + j16] = D0[cc][v];
                T[((wave * 2 + 1) * 16 + i) * 17 + j16] = D1[cc][v];
            }
            __syncthreads();
            // thread: real column j = 4 (map) + q, l = l0 + tid % 32
            const int j = tid >> 5, b = (chunk * NC + cc) * 4 + (j >> 2), g = (j >> 1) & 1, pp = j & 1;
            const int ll = tid & 31, par = ll & 1, i = ll >> 1;
            double sum = 0.0;
#pragma unroll
            for (int w = 0; w < 8; w++) sum += T[((w * 2 + par) * 16 + i) * 17 + j];
            if (ll < n && b < nmap)
                reinterpret_cast<double*>(dst + (size_t)(2 * b + g) * out_stride + base + l0 + ll)[pp] = sum * osc;
        }
        __syncthreads();  // the sums are read before the next tile overwrites them
    }
}

// One chunk of LC values of l for NRT ring pairs (spin s); p0 = parity of l_start + m + s.
template <typename T, int NRT, int BC, int LC, bool FAST, bool FULL>
__device__ __forceinline__ void adj_sp_chunk(cplx<T> (&acc)[LC][BC][2], Rec (&rp)[NRT], Rec (&rm)[NRT],
                                             const double (&x)[NRT], const cplx<T> (&in4)[NRT][BC][4],
                                             const double* __restrict__ sco, int nvalid, int p0) {
#pragma unroll
    for (int c = 0; c < LC; c++) {
        if (FULL || c < nvalid) {
            const bool even = ((c + p0) & 1) == 0;  // l + m + s even
            const double al = sco[3 * c], be = sco[3 * c + 1], ga = sco[3 * c + 2];
#pragma unroll
            for (int q = 0; q < NRT; q++) {
                const double fp = FAST ? rp[q].f : rp[q].f * rp[q].cf;
                const double fm = FAST ? rm[q].f : rm[q].f * rm[q].cf;
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
                rec_fast(rp[q], al * (x[q] + be), ga);
                rec_fast(rm[q], al * (x[q] - be), ga);
            }
        }
    }
    if (!FAST)
#pragma unroll
        for (int q = 0; q < NRT; q++) {
            rec_rescale(rp[q]);
            rec_rescale(rm[q]);
        }
}

// ------------------------------------------------ Legendre adjoint, spin s
// phases [(b, R|I)][mmax+1][nrings] -> alm [(b, G|C)][nlm] (scaled); ring slices and
// partial sums as for spin 0.
template <typename T, int NRT, int NW, int BC, int LC>
__global__ __launch_bounds__(64 * NW) void k_leg_adj_sp(ShtDev<T> d, int nmap, const cplx<T>* __restrict__ phase,
                                                        cplx<T>* __restrict__ alm, long alm_stride, T scale,
                                                        int nsplit, int pps) {
    constexpr int NT = 64 * NW;
    constexpr int V = LC * BC * 4;
    constexpr int TPV = NT / V;
    static_assert(LCH % LC == 0 && (LC & 1) == 0, "LC must be even and divide LCH");
    __shared__ T red[V][NT + TPV];
    __shared__ double sco[LCH * 3];
    const int nchunk = (nmap + BC - 1) / BC;
    unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = lb % nchunk;
    const int sl = (lb / nchunk) % nsplit;
    const int m = m_of_q(lb / (nchunk * nsplit), d.mmax);
    const int tid = threadIdx.x;
    const int M1 = d.mmax + 1;
    const size_t pstride = (size_t)M1 * d.nrings;
    const int jp0 = sl * pps, jp1 = min(d.npairs, jp0 + pps);
    cplx<T>* out = nsplit > 1 ? alm + (size_t)sl * 2 * nmap * alm_stride : alm;
    const T osc = nsplit > 1 ? T(1) : scale;

    double x[NRT];
    Rec rp[NRT], rm[NRT];
    // per ring / map: Rsum, Rdif, Isum, Idif
    cplx<T> in4[NRT][BC][4];
#pragma unroll
    for (int q = 0; q < NRT; q++) {
        const int jp = jp0 + tid + NT * q;
        const bool ok = jp < jp1;
        const int jpc = ok ? jp : jp0;
        x[q] = d.cth[jpc];
        rec_seed(rp[q], d.seedL[m] + d.seedA[m] * d.l2sh[jpc] + d.seedB[m] * d.l2ch[jpc], d.seedS[m]);
        rec_seed(rm[q], d.seedL[M1 + m] + d.seedA[M1 + m] * d.l2sh[jpc] + d.seedB[M1 + m] * d.l2ch[jpc],
                 d.seedS[M1 + m]);
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
    for (int i = tid; i < (lstart - m) * BC * 2; i += NT) {
        const int l = m + i / (BC * 2), b = chunk * BC + (i / 2) % BC, g = i % 2;
        if (l <= d.lmax && b < nmap) out[(size_t)(2 * b + g) * alm_stride + base + l] = cplx<T>{T(0), T(0)};
    }
    const int p0 = (lstart + m + d.spin) & 1;
    double pal = 0.0, pbe = 0.0, pga = 0.0;
    auto fetch = [&](int l0n) {
        if (tid < min(LCH, d.lmax + 1 - l0n)) {
            pal = cf4[4 * (l0n + tid)];
            pbe = cf4[4 * (l0n + tid) + 1];
            pga = cf4[4 * (l0n + tid) + 2];
        }
    };
    fetch(lstart);
    for (int l0 = lstart; l0 <= d.lmax; l0 += LCH) {
        const int n = min(LCH, d.lmax + 1 - l0);
        __syncthreads();
        if (tid < n) {
            sco[3 * tid] = pal;
            sco[3 * tid + 1] = pbe;
            sco[3 * tid + 2] = pga;
        }
        __syncthreads();
        if (l0 + LCH <= d.lmax) fetch(l0 + LCH);
        bool allk0 = true;
#pragma unroll
        for (int q = 0; q < NRT; q++) allk0 = allk0 && (rp[q].k == 0) && (rm[q].k == 0);
        const bool fast = __all(allk0);
        for (int lc = 0; lc < n; lc += LC) {
            cplx<T> acc[LC][BC][2];  // G, C
#pragma unroll
            for (int c = 0; c < LC; c++)
#pragma unroll
                for (int j = 0; j < BC; j++) acc[c][j][0] = acc[c][j][1] = cplx<T>{T(0), T(0)};
            const int nv = n - lc;
            if (nv >= LC) {
                if (fast) adj_sp_chunk<T, NRT, BC, LC, true, true>(acc, rp, rm, x, in4, sco + 3 * lc, nv, p0);
                else adj_sp_chunk<T, NRT, BC, LC, false, true>(acc, rp, rm, x, in4, sco + 3 * lc, nv, p0);
            } else {
                if (fast) adj_sp_chunk<T, NRT, BC, LC, true, false>(acc, rp, rm, x, in4, sco + 3 * lc, nv, p0);
                else adj_sp_chunk<T, NRT, BC, LC, false, false>(acc, rp, rm, x, in4, sco + 3 * lc, nv, p0);
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
            const T sum = red_rows<T, NT, V>(red, tid);
            if (tid % TPV == 0) {
                const int row = tid / TPV;
                const int c = row / (4 * BC), j = (row >> 2) % BC, g = (row >> 1) & 1, part = row & 1;
                const int l = l0 + lc + c, b = chunk * BC + j;
                if (c < nv && b < nmap)
                    reinterpret_cast<T*>(out + (size_t)(2 * b + g) * alm_stride + base + l)[part] = sum * osc;
            }
            __syncthreads();
        }
    }
}

// alm[col][lm] = scale * sum_s part[s][col][lm] (fixed order)
template <typename T>
__global__ void k_adj_sum_splits(const cplx<T>* __restrict__ part, int nsplit, int ncol, long nlm,
                                 cplx<T>* __restrict__ alm, long alm_stride, T scale) {
    const long n = (long)ncol * nlm;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        cplx<T> s{T(0), T(0)};
        for (int k = 0; k < nsplit; k++) s = cadd(s, part[(size_t)k * n + i]);
        const long col = i / nlm, lm = i - col * nlm;
        alm[(size_t)col * alm_stride + lm] = cplx<T>{s.re * scale, s.im * scale};
    }
}

// --------------------------------------------------------- layout helpers
// alm [col][nlm] (row stride alm_stride) -> alm_t [nlm][ld]; padding columns zeroed.  A block
// transposes a tile of 64 values of lm x 16 columns through LDS: the loads run along lm (one
// 64-element run per column), the stores along the columns of each lm row.
constexpr int TRL = 64, TRC = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_alm_to_lmajor(const cplx<T>* __restrict__ alm, long alm_stride, int ncol,
                                                       long nlm, int ld, cplx<T>* __restrict__ out) {
    __shared__ cplx<T> tile[TRC][TRL + 1];
    const long lm0 = (long)blockIdx.x * TRL;
    const int c0 = blockIdx.y * TRC;
    const int tid = threadIdx.x;
#pragma unroll
    for (int k = 0; k < TRL * TRC / 256; k++) {
        const int e = tid + 256 * k, c = e / TRL, t = e % TRL;  // t fastest: coalesced along lm
        const long lm = lm0 + t;
        const int col = c0 + c;
        tile[c][t] = (col < ncol && lm < nlm) ? alm[(size_t)col * alm_stride + lm] : cplx<T>{T(0), T(0)};
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TRL * TRC / 256; k++) {
        const int e = tid + 256 * k, t = e / TRC, c = e % TRC;  // c fastest: coalesced along a row
        const long lm = lm0 + t;
        const int col = c0 + c;
        if (lm < nlm && col < ld) out[lm * ld + col] = tile[c][t];
    }
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

// Column pairs per block: as many FFT buffers as fit 64 KB of LDS for the ring class.  Q: the
// butterfly slots per thread for the class's longest transform (maxM <= 1024 Q).  STK: the
// self-sorting radix-16 transforms instead of the radix-2^2 ones.
template <typename T, int NB, int Q, bool STK>
static int launch_ring_synth_nbq(const ShtPlan& p, const int* list, int nlist, long maxM, int ncol,
                                 const void* phase, void* map, long map_stride, hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    const size_t sh = (size_t)NB * fft_pitch((int)maxM) * sizeof(cplx<T>);
    if (sh + 256 * sizeof(cplx<T>) > 65536)  // dynamic buffers plus the static slice partials
        PS_CHECK_HIP(hipFuncSetAttribute((const void*)k_ring_synth<T, NB, Q, STK>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    const int ngrp = (ncol + 2 * NB - 1) / (2 * NB);
    hipLaunchKernelGGL((k_ring_synth<T, NB, Q, STK>), dim3((unsigned)(nlist * ngrp)), dim3(256), sh, st, d, list,
                       ncol, (const cplx<T>*)phase, (T*)map, map_stride);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T, int NB, bool STK>
static int launch_ring_synth_nb(const ShtPlan& p, const int* list, int nlist, long maxM, int ncol,
                                const void* phase, void* map, long map_stride, hipStream_t st) {
    if (maxM <= 1024)
        return launch_ring_synth_nbq<T, NB, 1, STK>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
    if (maxM <= 2048)
        return launch_ring_synth_nbq<T, NB, 2, STK>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
    if constexpr (NB <= 2) {
        if (maxM <= 4096)
            return launch_ring_synth_nbq<T, NB, 4, STK>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
    }
    if constexpr (NB == 1)
        return launch_ring_synth_nbq<T, NB, FFT_MAXQ, STK>(p, list, nlist, maxM, ncol, phase, map, map_stride,
                                                           st);
    set_error("ring synthesis: transform too long for this many buffers per block");
    return -1;
}

template <typename T, int NB, int Q, bool BLUE, bool GNL = false>
static int launch_ring_synth_sgq(const ShtPlan& p, const int* list, int nlist, int ncol, const void* phase,
                                 void* map, long map_stride, hipStream_t st, size_t sh, const GnlArgs& g = GnlArgs()) {
    ShtDev<T> d = dev_view<T>(p);
    if (sh + 256 * sizeof(cplx<T>) > 65536)
        PS_CHECK_HIP(hipFuncSetAttribute((const void*)k_ring_synth_sg<T, NB, Q, BLUE, GNL>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    const int ngrp = (ncol + 2 * NB - 1) / (2 * NB);
    hipLaunchKernelGGL((k_ring_synth_sg<T, NB, Q, BLUE, GNL>), dim3((unsigned)(nlist * ngrp)), dim3(256), sh, st, d,
                       list, ncol, (const cplx<T>*)phase, (T*)map, map_stride, g);
    PS_LAUNCH_CHECK();
    return 0;
}

static int ring_q_of(long maxM) { return maxM <= 1024 ? 1 : maxM <= 2048 ? 2 : maxM <= 4096 ? 4 : 8; }

// LDS of the staged ring synthesis: the raw rows or the FFT buffers, whichever is larger
static size_t ring_sg_lds(const ShtPlan& p, long maxM, int nb) {
    const size_t c = 2 * (size_t)p.prec;
    return std::max((size_t)(p.mmax + 1) * 2 * nb * c, (size_t)nb * (size_t)(maxM + 1) * c);
}

template <typename T, int NB, bool BLUE>
static int launch_ring_synth_sg(const ShtPlan& p, const int* list, int nlist, long maxM, int ncol,
                                const void* phase, void* map, long map_stride, hipStream_t st) {
    const size_t sh = ring_sg_lds(p, maxM, NB);
    switch (ring_q_of(maxM)) {
        case 1: return launch_ring_synth_sgq<T, NB, 1, BLUE>(p, list, nlist, ncol, phase, map, map_stride, st, sh);
        case 2: return launch_ring_synth_sgq<T, NB, 2, BLUE>(p, list, nlist, ncol, phase, map, map_stride, st, sh);
        case 4:
            if constexpr (NB == 1)
                return launch_ring_synth_sgq<T, 1, 4, BLUE>(p, list, nlist, ncol, phase, map, map_stride, st, sh);
            break;
        case 8:
            if constexpr (NB == 1 && BLUE)
                return launch_ring_synth_sgq<T, 1, 8, BLUE>(p, list, nlist, ncol, phase, map, map_stride, st, sh);
            break;
        default: break;
    }
    set_error("ring synthesis: staged shape not valid for this ring class");
    return -1;
}

// Ring-synthesis shape id: NB FFT buffers of two real columns each per block (1, 2 or 4), plus
// RING_STK for the self-sorting radix-16 transforms (NB 1 or 2) instead of the radix-2^2 ones,
// or RING_SG for the self-sorting transforms with the phase rows staged in LDS (NB 1 or 2).
// The built-in rule takes the staged transforms where they are valid, else the self-sorting
// ones: two buffers per block for power-of-two rings when they fit 64 KB of LDS and the batch
// can use them, one for Bluestein rings (two transforms per block already).
constexpr int RING_STK = 8, RING_SG = 16;

static bool ring_nb_valid(const ShtPlan& p, long maxM, int ncol, bool blue, int id);

static int ring_nb_rule(const ShtPlan& p, long maxM, int ncol, bool blue) {
    // the staged kernels wherever their LDS and item budget allow (measured on MI355X, Nside
    // 256 - 1024, 1 - 128 columns, both precisions: 10 - 50 % faster than the unstaged ones);
    // two buffers per block for power-of-two rings, one for Bluestein rings
    if (!blue && ncol >= 4 && ring_nb_valid(p, maxM, ncol, blue, RING_SG | 2)) return RING_SG | 2;
    if (ring_nb_valid(p, maxM, ncol, blue, RING_SG | 1)) return RING_SG | 1;
    const long per = maxM * 2L * p.prec;
    if (!blue && ncol >= 4 && 2 * per <= 65536) return RING_STK | 2;
    return RING_STK | 1;
}

static bool ring_nb_valid(const ShtPlan& p, long maxM, int ncol, bool blue, int id) {
    if (id >= 0 && (id & RING_SG)) {
        // two buffers within 64 KB of LDS; one buffer up to the CU's 160 KB (the Bluestein rings
        // of Nside 1024, transforms of 8192 points, take the dynamic-LDS attribute)
        const int nb = id & ~RING_SG, q = ring_q_of(maxM);
        const int items = (blue ? 2 : 4) * nb * q;
        const size_t lds = ring_sg_lds(p, maxM, nb) + 256 * 2 * (size_t)p.prec;
        return (nb == 1 || (nb == 2 && ncol > 2)) && items <= 16 && (q < 8 || blue) &&
               lds <= (nb == 1 ? (size_t)160 * 1024 : (size_t)65536);
    }
    const int prec = p.prec;
    const int nb = id & ~RING_STK;
    if (id < 0 || ((id & RING_STK) && nb > 2)) return false;
    const long per = maxM * 2L * prec;
    // one buffer always (beyond 64 KB it takes the dynamic-LDS attribute); more only within 64 KB
    return nb == 1 || ((nb == 2 || nb == 4) && nb * per <= 65536 && ncol > nb);
}

template <typename T>
static int launch_ring_synth_list(const ShtPlan& p, const int* list, int nlist, long maxM, int ncol,
                                  const void* phase, void* map, long map_stride, hipStream_t st, int nb) {
    if (nlist == 0) return 0;
    if (maxM > 4L * 256 * FFT_MAXQ) {
        set_error("ring synthesis: FFT length above 8192 is not supported");
        return -1;
    }
    // d_rlist lists the power-of-two rings first: a list starting past them is the Bluestein class
    const bool blue = list != p.d_rlist || p.n_pow2 == 0;
    if (nb < 0) nb = ring_nb_rule(p, maxM, ncol, blue);
    if (!ring_nb_valid(p, maxM, ncol, blue, nb)) {
        set_error("ring synthesis: invalid launch shape");
        return -1;
    }
    switch (nb) {
        case 4: return launch_ring_synth_nb<T, 4, false>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
        case 2: return launch_ring_synth_nb<T, 2, false>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
        case RING_STK | 1:
            return launch_ring_synth_nb<T, 1, true>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
        case RING_STK | 2:
            return launch_ring_synth_nb<T, 2, true>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
        case RING_SG | 1:
            return blue ? launch_ring_synth_sg<T, 1, true>(p, list, nlist, maxM, ncol, phase, map, map_stride, st)
                        : launch_ring_synth_sg<T, 1, false>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
        case RING_SG | 2:
            return blue ? launch_ring_synth_sg<T, 2, true>(p, list, nlist, maxM, ncol, phase, map, map_stride, st)
                        : launch_ring_synth_sg<T, 2, false>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
        default: return launch_ring_synth_nb<T, 1, false>(p, list, nlist, maxM, ncol, phase, map, map_stride, st);
    }
}

template <typename T>
static int launch_ring_synth(const ShtPlan& p, int ncol, const void* phase, void* map, long map_stride,
                             hipStream_t st, int nb_pow2, int nb_blue) {
    int rc = launch_ring_synth_list<T>(p, p.d_rlist, p.n_pow2, p.maxM_pow2, ncol, phase, map, map_stride, st,
                                       nb_pow2);
    if (rc) return rc;
    return launch_ring_synth_list<T>(p, p.d_rlist + p.n_pow2, p.nrings - p.n_pow2, p.maxM_blue, ncol, phase, map,
                                     map_stride, st, nb_blue);
}

template <typename T, int Q>
static int launch_ring_adjoint_q(const ShtPlan& p, int ncol, const void* map, long map_stride, void* phase,
                                 hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    size_t sh = lds_bytes<T>(p);
    if (sh > 65536)
        PS_CHECK_HIP(hipFuncSetAttribute((const void*)k_ring_adjoint<T, Q>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    hipLaunchKernelGGL((k_ring_adjoint<T, Q>), dim3((unsigned)(p.nrings * ((ncol + 1) / 2))), dim3(256), sh, st,
                       d, ncol, (const T*)map, map_stride, (cplx<T>*)phase);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int launch_ring_adjoint(const ShtPlan& p, int ncol, const void* map, long map_stride, void* phase,
                               hipStream_t st) {
    if (p.maxM > 4L * 256 * FFT_MAXQ) {
        set_error("ring adjoint: FFT length above 8192 is not supported");
        return -1;
    }
    if (p.maxM <= 1024) return launch_ring_adjoint_q<T, 1>(p, ncol, map, map_stride, phase, st);
    if (p.maxM <= 2048) return launch_ring_adjoint_q<T, 2>(p, ncol, map, map_stride, phase, st);
    if (p.maxM <= 4096) return launch_ring_adjoint_q<T, 4>(p, ncol, map, map_stride, phase, st);
    return launch_ring_adjoint_q<T, FFT_MAXQ>(p, ncol, map, map_stride, phase, st);
}

// Ring slices of the adjoint: with few batch columns the blocks of small m (the longest
// recursions) set the kernel time, so their rings are split over nsplit blocks whose
// partial sums a second pass adds in fixed order.
int sht_adjoint_nsplit(const ShtPlan& p, int nmap) {
    const bool many = p.spin == 0 ? nmap * p.ncomp > 2 : nmap > 1;
    if (many) return 1;
    int ns = 1;
    while (ns < 8 && (p.npairs + 2 * ns - 1) / (2 * ns) >= 64) ns *= 2;
    return ns;
}

size_t sht_adjoint_work_bytes(const ShtPlan& p, int nmap) {
    const size_t ph = (sht_phase_bytes(p, nmap) + 255) & ~(size_t)255;
    const int ns = sht_adjoint_nsplit(p, nmap);
    // ring-slice partial sums: of the vector kernels' split, or of the matrix-core kernel
    int np = ns > 1 ? ns : 0;
    const int nsm = (p.npairs + ADJM_PAIRS - 1) / ADJM_PAIRS;
    if (nsm > 1) np = std::max(np, nsm);
    return ph + (size_t)np * nmap * p.ncomp * p.nlm * 2 * p.prec;
}

template <typename T, int NRT, int NW, int BC>
static int launch_adj_cfg(const ShtPlan& p, int ncol, const void* phase, void* alm, long alm_stride,
                          double scale, int ns, int pps, hipStream_t st) {
    ShtDev<T> d = dev_view<T>(p);
    if (p.spin == 0) {
        constexpr int LC = 8;
        const int nchunk = (ncol + BC - 1) / BC;
        hipLaunchKernelGGL((k_leg_adj_s0<T, NRT, NW, BC, LC>), dim3((unsigned)(nchunk * ns * (p.mmax + 1))),
                           dim3(64 * NW), 0, st, d, ncol, (const cplx<T>*)phase, (cplx<T>*)alm, alm_stride,
                           (T)scale, ns, pps);
    } else {
        constexpr int LC = 4;
        const int nmap = ncol / 2;
        const int nchunk = (nmap + BC - 1) / BC;
        hipLaunchKernelGGL((k_leg_adj_sp<T, NRT, NW, BC, LC>), dim3((unsigned)(nchunk * ns * (p.mmax + 1))),
                           dim3(64 * NW), 0, st, d, nmap, (const cplx<T>*)phase, (cplx<T>*)alm, alm_stride,
                           (T)scale, ns, pps);
    }
    PS_LAUNCH_CHECK();
    return 0;
}

// Adjoint Legendre shapes: NRT ring pairs per thread, NW waves per block, BC batch columns
// (spin 0) or maps (spin s) per block; a block covers 64 NRT NW ring pairs of its ring slice.
struct AdjShape {
    int nrt, nw, bc;
    bool s0, sp;  // instantiated for spin 0 / spin s
};
static constexpr AdjShape ADJ_SHAPES[] = {
    {1, 1, 2, true, false}, {2, 1, 2, true, false}, {4, 1, 2, true, true},  {8, 1, 2, true, false},
    {8, 2, 2, true, false}, {8, 4, 2, true, false}, {4, 2, 2, true, true},  {4, 4, 2, true, true},
    {4, 8, 2, true, true},  {1, 1, 1, false, true}, {2, 1, 1, false, true}, {4, 1, 1, false, true},
    {8, 1, 1, false, true}, {8, 2, 1, false, true}, {8, 4, 1, false, true}};
static constexpr int N_ADJ_SHAPES = (int)(sizeof(ADJ_SHAPES) / sizeof(ADJ_SHAPES[0]));
// shape id of the matrix-core adjoint: 256 ring pairs per block and 16 complex columns (spin 0) or
// 8 maps (spin s) in FP32, 8 complex columns or 4 maps in FP64
static constexpr int ADJ_SHAPE_MFMA = N_ADJ_SHAPES;
// the spin-s matrix-core kernels with two chunks of maps per block (8 maps a chunk in FP32, 4 in FP64)
static constexpr int ADJ_SHAPE_MFMA2 = N_ADJ_SHAPES + 1;

static bool adj_mfma_ok(const ShtPlan& p) { return p.prec == 4 || p.prec == 8; }

static bool adj_shape_valid(const ShtPlan& p, int id, int pps) {
    if (id == ADJ_SHAPE_MFMA) return adj_mfma_ok(p);
    if (id == ADJ_SHAPE_MFMA2) return adj_mfma_ok(p) && p.spin > 0;
    if (id < 0 || id >= N_ADJ_SHAPES) return false;
    const AdjShape& a = ADJ_SHAPES[id];
    return (p.spin == 0 ? a.s0 : a.sp) && 64 * a.nrt * a.nw >= pps;
}

// Built-in rule.  Latency (few columns, ring slices): up to 8 ring pairs per thread, one column
// chunk of 2 (spin 0) or 1 map (spin s) per block.  Throughput (many columns): 4 ring pairs per
// thread and more waves per block.
static int adj_shape_rule(const ShtPlan& p, int ncol, bool many, int pps) {
    // batches: the matrix-core kernel (FP32 as for the synthesis; FP64 from 8 real columns).
    // FP32 spin 0 from 3 columns on plans of 1024 ring pairs and more (Nside >= 512; measured,
    // the whole adjoint of 4 / 8 maps: Nside 1024, lmax 2000, 8.3 / 10.0 ms on the vector kernel
    // against 4.9 / 5.3 ms; Nside 512, lmax 1024, 0.93 / 1.13 against 0.77 / 0.88 ms; at Nside 256
    // the vector kernel stays ahead, 0.18 / 0.22 against 0.23 / 0.25 ms)
    if (p.prec == 4 && p.spin == 0 && ncol > 2 && p.npairs >= 1024) return ADJ_SHAPE_MFMA;
    // FP32 spin s: two chunks of 8 maps per block from 9 maps on plans of 1024 ring pairs and
    // more, from 16 maps on smaller ones (measured, whole adjoint: 64 maps at Nside 1024, lmax 2048,
    // 104.1 against 80.9 ms; 16 at Nside 512, lmax 1535, 7.84 against 6.10 ms; 13 at Nside 512,
    // lmax 1024, 3.78 against 2.96 ms; but 9 at Nside 256, 0.63 against 0.76 ms: the second chunk
    // nearly empty and half the blocks)
    if (p.prec == 4 && p.spin > 0 && ncol / 2 > 8 && (p.npairs >= 1024 || ncol / 2 >= 16)) return ADJ_SHAPE_MFMA2;
    if (p.prec == 4 && (p.spin == 0 ? ncol > 8 : ncol / 2 > 2)) return ADJ_SHAPE_MFMA;
    // FP64 spin s likewise with chunks of 4 maps (whole adjoint: 64 maps at Nside 1024, lmax 2048,
    // 203.3 against 159.0 ms; 16 at Nside 512, lmax 1535, 14.19 against 11.08 ms; 6 at Nside 512,
    // lmax 1024, 3.51 against 2.77 ms; at Nside 256, 16 maps 1.29 against 1.07 ms but 5 maps 0.61
    // against 0.77 ms)
    if (p.prec == 8 && p.spin > 0 && ncol / 2 > 4 && (p.npairs >= 1024 || ncol / 2 >= 8)) return ADJ_SHAPE_MFMA2;
    if (p.prec == 8 && (p.spin == 0 ? ncol > 4 : ncol / 2 > 1)) return ADJ_SHAPE_MFMA;
    const int lvl = pps <= 64 ? 0 : pps <= 128 ? 1 : pps <= 256 ? 2 : pps <= 512 ? 3 : pps <= 1024 ? 4 : 5;
    if (many) {
        static constexpr int thr[6] = {2, 2, 2, 6, 7, 8};
        return thr[lvl];
    }
    static constexpr int lat_s0[6] = {0, 1, 2, 3, 4, 5}, lat_sp[6] = {9, 10, 11, 12, 13, 14};
    return p.spin == 0 ? lat_s0[lvl] : lat_sp[lvl];
}

template <typename T>
static int launch_adj_shape(const ShtPlan& p, int id, int ncol, const void* phase, void* dst, long dstride,
                            double scale, int ns, int pps, hipStream_t st) {
    if (!adj_shape_valid(p, id, pps)) {
        set_error("adjoint: invalid launch shape");
        return -1;
    }
#define PS_ADJ_CASE(I, NRT, NW, BC) \
    case I: return launch_adj_cfg<T, NRT, NW, BC>(p, ncol, phase, dst, dstride, scale, ns, pps, st);
    switch (id) {
        PS_ADJ_CASE(0, 1, 1, 2)
        PS_ADJ_CASE(1, 2, 1, 2)
        PS_ADJ_CASE(2, 4, 1, 2)
        PS_ADJ_CASE(3, 8, 1, 2)
        PS_ADJ_CASE(4, 8, 2, 2)
        PS_ADJ_CASE(5, 8, 4, 2)
        PS_ADJ_CASE(6, 4, 2, 2)
        PS_ADJ_CASE(7, 4, 4, 2)
        PS_ADJ_CASE(8, 4, 8, 2)
        PS_ADJ_CASE(9, 1, 1, 1)
        PS_ADJ_CASE(10, 2, 1, 1)
        PS_ADJ_CASE(11, 4, 1, 1)
        PS_ADJ_CASE(12, 8, 1, 1)
        PS_ADJ_CASE(13, 8, 2, 1)
        PS_ADJ_CASE(14, 8, 4, 1)
    }
#undef PS_ADJ_CASE
    set_error("adjoint: invalid launch shape");
    return -1;
}

static_assert(ADJM_PAIRS == ADJM_SP_PAIRS, "both matrix-core adjoints slice rings alike");
static int adj_mfma_nslice(const ShtPlan& p) { return (p.npairs + ADJM_PAIRS - 1) / ADJM_PAIRS; }

// nc: chunks of maps per block of the spin-s kernels (1 or 2)
template <typename T>
static int launch_adj_mfma(const ShtPlan& p, int ncol, const void* phase, void* alm, long alm_stride,
                           double scale, void* part, hipStream_t st, int nc = 1) {
    ShtDev<T> d = dev_view<T>(p);
    const int nslice = adj_mfma_nslice(p);
    void* dst = nslice > 1 ? part : alm;
    const long dstride = nslice > 1 ? p.nlm : alm_stride;
    const T osc = nslice > 1 ? T(1) : (T)scale;
    if constexpr (sizeof(T) == 4) {
        if (p.spin == 0) {
            const int nchunk = (ncol + 15) / 16;
            hipLaunchKernelGGL(k_leg_adj_s0_mfma, dim3((unsigned)(nchunk * nslice * (p.mmax + 1))), dim3(256), 0,
                               st, d, ncol, (const cplx<float>*)phase, (cplx<float>*)dst, dstride, osc, nslice);
        } else {
            const int nmap = ncol / 2, nchunk = (nmap + 8 * nc - 1) / (8 * nc);
            const unsigned nb = (unsigned)(nchunk * nslice * (p.mmax + 1));
            if (nc == 2)
                hipLaunchKernelGGL(k_leg_adj_sp_mfma<2>, dim3(nb), dim3(512), 0, st, d, nmap, (const cplx<float>*)phase,
                                   (cplx<float>*)dst, dstride, osc, nslice);
            else
                hipLaunchKernelGGL(k_leg_adj_sp_mfma<1>, dim3(nb), dim3(512), 0, st, d, nmap, (const cplx<float>*)phase,
                                   (cplx<float>*)dst, dstride, osc, nslice);
        }
    } else {
        if (p.spin == 0) {
            const int nchunk = (ncol + 7) / 8;
            hipLaunchKernelGGL(k_leg_adj_s0_mfma64, dim3((unsigned)(nchunk * nslice * (p.mmax + 1))), dim3(256), 0,
                               st, d, ncol, (const cplx<double>*)phase, (cplx<double>*)dst, dstride, osc, nslice);
        } else {
            const int nmap = ncol / 2, nchunk = (nmap + 4 * nc - 1) / (4 * nc);
            const unsigned nb = (unsigned)(nchunk * nslice * (p.mmax + 1));
            if (nc == 2)
                hipLaunchKernelGGL(k_leg_adj_sp_mfma64<2>, dim3(nb), dim3(512), 0, st, d, nmap,
                                   (const cplx<double>*)phase, (cplx<double>*)dst, dstride, osc, nslice);
            else
                hipLaunchKernelGGL(k_leg_adj_sp_mfma64<1>, dim3(nb), dim3(512), 0, st, d, nmap,
                                   (const cplx<double>*)phase, (cplx<double>*)dst, dstride, osc, nslice);
        }
    }
    PS_LAUNCH_CHECK();
    if (nslice == 1) return 0;
    const long n = (long)ncol * p.nlm;
    hipLaunchKernelGGL(k_adj_sum_splits<T>, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0,
                       st, (const cplx<T>*)part, nslice, ncol, p.nlm, (cplx<T>*)alm, alm_stride, (T)scale);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int launch_adj(const ShtPlan& p, int ncol, const void* phase, void* alm, long alm_stride, double scale,
                      void* part, hipStream_t st, int shape) {
    if (p.npairs > 2048) {
        set_error("adjoint: nside above 1024 is not supported");
        return -1;
    }
    const int nmap = ncol / p.ncomp;
    const int ns = sht_adjoint_nsplit(p, nmap);
    const int pps = (p.npairs + ns - 1) / ns;
    void* dst = ns > 1 ? part : alm;
    const long dstride = ns > 1 ? p.nlm : alm_stride;
    const bool many = p.spin == 0 ? ncol > 2 : nmap > 1;
    if (shape < 0) shape = adj_shape_rule(p, ncol, many, pps);
    if (shape == ADJ_SHAPE_MFMA || shape == ADJ_SHAPE_MFMA2) {
        if (!adj_shape_valid(p, shape, pps)) {
            set_error("adjoint: invalid launch shape");
            return -1;
        }
        return launch_adj_mfma<T>(p, ncol, phase, alm, alm_stride, scale, part, st, shape == ADJ_SHAPE_MFMA2 ? 2 : 1);
    }
    int rc = launch_adj_shape<T>(p, shape, ncol, phase, dst, dstride, scale, ns, pps, st);
    if (rc || ns == 1) return rc;
    const long n = (long)ncol * p.nlm;
    hipLaunchKernelGGL(k_adj_sum_splits<T>, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, st,
                       (const cplx<T>*)part, ns, ncol, p.nlm, (cplx<T>*)alm, alm_stride, (T)scale);
    PS_LAUNCH_CHECK();
    return 0;
}

// synthesis BC (batch columns per block)
static constexpr int SYN_BC_S0_F = 16, SYN_BC_S0_D = 8, SYN_BC_SP_F = 4, SYN_BC_SP_D = 2;

int sht_synth_lmajor_ld_multiple(const ShtPlan& p) {
    if (p.spin == 0) return p.prec == 4 ? SYN_BC_S0_F : SYN_BC_S0_D;
    return 2 * (p.prec == 4 ? SYN_BC_SP_F : SYN_BC_SP_D);
}

size_t sht_phase_bytes(const ShtPlan& p, int nmap) {
    // synthesis tiles cover the ring count rounded up to PRT (the adjoint layout is smaller)
    const size_t rings = (size_t)(p.nrings + PRT - 1) / PRT * PRT;
    return (size_t)nmap * p.ncomp * (p.mmax + 1) * rings * 2 * p.prec;
}

// Column-chunk size of the synthesis kernels: the smallest instantiated chunk covering
// the batch (a single map uses BC = 1 instead of wasting the padding columns).
static int launch_synth_s0_mfma(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase,
                                hipStream_t st, int rb0 = 0, int rb1 = -1) {
    // One tile of 16 complex columns per block.  Sharing the recursion over up to 5 tiles
    // (NT = 5, 418 registers, one wave per SIMD) measured slower at Nside 512 / 80 maps
    // (3.69 ms against 3.57 ms), and so did v_mfma_f32_32x32x1f32 with one ring per lane
    // (no duplicated recursion, 4.57 ms): the kernel is bound by its VALU instruction stream
    // (about 32 VALU instructions per MFMA) and latency, not by the recursion's arithmetic.
    ShtDev<float> d = dev_view<float>(p);
    constexpr int NT = 1;
    const int nrb = (p.npairs + 127) / 128, nchunk = (ncol + 16 * NT - 1) / (16 * NT);
    if (rb1 < 0) rb1 = nrb;  // ring-pair blocks [rb0, rb1) of 128 pairs
    const unsigned nb = (unsigned)((rb1 - rb0) * nchunk * (p.mmax + 1));
    hipLaunchKernelGGL(k_leg_synth_s0_mfma<NT>, dim3(nb), dim3(256), 0, st, d, ncol, ld,
                       (const cplx<float>*)alm_t, (cplx<float>*)phase, rb0, rb1 - rb0);
    PS_LAUNCH_CHECK();
    return 0;
}

// The wide matrix-core synthesis (shape SYN_SHAPE_WR, k_leg_synth_s0_wr): NT = the column tiles
// of one block, up to 5 (80 columns); wider batches take several column chunks.
constexpr int SYN_SHAPE_WR = 65;
// the wide form with an explicit tile count: SYN_SHAPE_WR_NT + nt, nt = 2 .. 4 (fewer registers and
// more waves per SIMD, the recursion repeated for each column chunk); autotuner candidates
constexpr int SYN_SHAPE_WR_NT = 80;
static bool synth_is_wr(int shape) {
    return shape == SYN_SHAPE_WR || (shape >= SYN_SHAPE_WR_NT + 2 && shape <= SYN_SHAPE_WR_NT + 4);
}
static int synth_lt_nt(int ncol) {
    const int tiles = (ncol + 15) / 16;
    int nt = tiles <= 5 ? tiles : 5;  // wider batches: the tile count of least padding, largest first
    for (int c = 5; tiles > 5 && c >= 3; c--)
        if ((tiles + c - 1) / c * c < (tiles + nt - 1) / nt * nt) nt = c;
    return nt;
}
static int launch_synth_s0_wr(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase, hipStream_t st,
                              int rb0 = 0, int rb1 = -1, int shape = SYN_SHAPE_WR) {
    ShtDev<float> d = dev_view<float>(p);
    const int nt = shape == SYN_SHAPE_WR ? synth_lt_nt(ncol) : shape - SYN_SHAPE_WR_NT;
    const int nrb = (p.npairs + 127) / 128, nchunk = (ncol + 16 * nt - 1) / (16 * nt);
    if (rb1 < 0) rb1 = nrb;
    const unsigned nb = (unsigned)((rb1 - rb0) * nchunk * (p.mmax + 1));
    const cplx<float>* a = (const cplx<float>*)alm_t;
    cplx<float>* ph = (cplx<float>*)phase;
    switch (nt) {
        case 1: hipLaunchKernelGGL(k_leg_synth_s0_wr<1>, dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph, rb0, rb1 - rb0); break;
        case 2: hipLaunchKernelGGL(k_leg_synth_s0_wr<2>, dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph, rb0, rb1 - rb0); break;
        case 3: hipLaunchKernelGGL(k_leg_synth_s0_wr<3>, dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph, rb0, rb1 - rb0); break;
        case 4: hipLaunchKernelGGL(k_leg_synth_s0_wr<4>, dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph, rb0, rb1 - rb0); break;
        default: hipLaunchKernelGGL(k_leg_synth_s0_wr<5>, dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph, rb0, rb1 - rb0); break;
    }
    PS_LAUNCH_CHECK();
    return 0;
}
// Legendre-synthesis shapes: 0 = the matrix-core kernel (FP32, or FP64), SYN_SHAPE_WR = its wide form
// (FP32), otherwise the batch columns (spin 0) or maps (spin s) per block of the vector
// kernel.  Measured at Nside 256 and 512 (spin 0): the wide form is fastest from 24 columns up
// (1.2-1.4x), the 16-column one from 5 to 16 columns, the vector kernel up to 4.  Spin s: the wide
// form from two tiles of 16 complex columns (synth_sp_nt).
template <typename T>
static int synth_shape_rule(const ShtPlan& p, int ncol) {
    if (p.spin == 0) {
        constexpr int BCMAX = sizeof(T) == 4 ? SYN_BC_S0_F : SYN_BC_S0_D;
        if (sizeof(T) == 4 && ncol > 16) return SYN_SHAPE_WR;
        if (ncol > 4) return 0;  // the matrix-core kernel (FP32, or its FP64 form)
        return ncol <= 1 ? 1 : ncol <= 2 ? 2 : ncol <= 4 ? 4 : (BCMAX > 8 && ncol <= 8) ? 8 : BCMAX;
    }
    constexpr int BCMAX = sizeof(T) == 4 ? SYN_BC_SP_F : SYN_BC_SP_D;
    const int nmap = ncol / p.ncomp;
    if (sizeof(T) == 4 && ncol > 16) return SYN_SHAPE_WR;  // the wide matrix-core form
    if (nmap > 2) return 0;
    return nmap <= 1 ? 1 : (BCMAX > 2 && nmap <= 2) ? 2 : BCMAX;
}

static int launch_synth_mfma64(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase, hipStream_t st);

template <typename T>
static int launch_synth_s0_any(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase, hipStream_t st,
                               int shape) {
    if (shape < 0) shape = synth_shape_rule<T>(p, ncol);
    switch (shape) {
        case 0:
            if constexpr (sizeof(T) == 4) return launch_synth_s0_mfma(p, ncol, ld, alm_t, phase, st);
            else return launch_synth_mfma64(p, ncol, ld, alm_t, phase, st);
        case SYN_SHAPE_WR:
        case SYN_SHAPE_WR_NT + 2:
        case SYN_SHAPE_WR_NT + 3:
        case SYN_SHAPE_WR_NT + 4:
            if constexpr (sizeof(T) == 4) return launch_synth_s0_wr(p, ncol, ld, alm_t, phase, st, 0, -1, shape);
            break;
        case 1: return launch_synth_s0<T, 1>(p, ncol, ld, alm_t, phase, st);
        case 2: return launch_synth_s0<T, 2>(p, ncol, ld, alm_t, phase, st);
        case 4: return launch_synth_s0<T, 4>(p, ncol, ld, alm_t, phase, st);
        case 8: return launch_synth_s0<T, 8>(p, ncol, ld, alm_t, phase, st);
        case 16:
            if constexpr (sizeof(T) == 4) return launch_synth_s0<T, 16>(p, ncol, ld, alm_t, phase, st);
            break;
    }
    set_error("synthesis: invalid launch shape");
    return -1;
}

// FP64 matrix-core syntheses: 256 (spin 0) or 128 (spin s) ring pairs and 8 complex columns per block
static int launch_synth_mfma64(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase,
                               hipStream_t st) {
    ShtDev<double> d = dev_view<double>(p);
    const int pairs = p.spin == 0 ? 256 : 128;
    const int nrb = (p.npairs + pairs - 1) / pairs, nchunk = (ncol + 7) / 8;
    const unsigned nb = (unsigned)(nrb * nchunk * (p.mmax + 1));
    if (p.spin == 0)
        hipLaunchKernelGGL(k_leg_synth_s0_mfma64, dim3(nb), dim3(256), 0, st, d, ncol, ld,
                           (const cplx<double>*)alm_t, (cplx<double>*)phase);
    else
        hipLaunchKernelGGL(k_leg_synth_sp_mfma64, dim3(nb), dim3(256), 0, st, d, ncol, ld,
                           (const cplx<double>*)alm_t, (cplx<double>*)phase);
    PS_LAUNCH_CHECK();
    return 0;
}

static int launch_synth_sp_mfma(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase,
                                hipStream_t st) {
    ShtDev<float> d = dev_view<float>(p);
    const int nrb = (p.npairs + 127) / 128, nchunk = (ncol + 15) / 16;
    const unsigned nb = (unsigned)(nrb * nchunk * (p.mmax + 1));
    hipLaunchKernelGGL(k_leg_synth_sp_mfma<false>, dim3(nb), dim3(256), 0, st, d, ncol, ld,
                       (const cplx<float>*)alm_t, (cplx<float>*)phase);
    PS_LAUNCH_CHECK();
    return 0;
}

static int shape_of(const ShtPlan& p, int ncol, int stage);

// The wide spin-s form (k_leg_synth_sp_wr): nt = 2 .. 4 tiles of 16 complex columns per block;
// ncol counts complex columns, or maps with GONLY.  synth_sp_nt: the built-in tile count, 2 from
// two tiles of columns up (1: the one-tile kernel).  Measured (FP32, the one-tile kernel against
// nt = 2 / 3): the gradient-only syntheses of 60 maps at Nside 1024, lmax 2000, 49.0 against
// 41.1 / 59.2 ms, 40 maps at Nside 256 1.12 against 1.02 ms, the (G, C) synthesis of 64 maps at
// Nside 1024, lmax 2048, 88.0 against 71.3 / 86.5 ms.  At nt = 2 the kernel holds 215 VGPRs,
// two waves per SIMD as the one-tile kernel, and runs each recursion step for twice the MFMAs;
// at nt = 3 and 4 it needs more than 256 registers and one wave per SIMD no longer hides the
// recursion's latency (tuner candidates only).
static int synth_sp_nt(int ncol) { return (ncol + 15) / 16 >= 2 ? 2 : 1; }
template <bool GONLY>
static int launch_synth_sp_wr(const ShtPlan& p, int ncol, int ld, const void* alm_t, void* phase, hipStream_t st,
                              int nt) {
    ShtDev<float> d = dev_view<float>(p);
    const int nrb = (p.npairs + 127) / 128, nchunk = (ncol + 16 * nt - 1) / (16 * nt);
    const unsigned nb = (unsigned)(nrb * nchunk * (p.mmax + 1));
    const cplx<float>* a = (const cplx<float>*)alm_t;
    cplx<float>* ph = (cplx<float>*)phase;
    switch (nt) {
        case 2: hipLaunchKernelGGL((k_leg_synth_sp_wr<2, GONLY>), dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph); break;
        case 3: hipLaunchKernelGGL((k_leg_synth_sp_wr<3, GONLY>), dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph); break;
        case 4: hipLaunchKernelGGL((k_leg_synth_sp_wr<4, GONLY>), dim3(nb), dim3(256), 0, st, d, ncol, ld, a, ph); break;
        default: set_error("synthesis: invalid wide tile count"); return -1;
    }
    PS_LAUNCH_CHECK();
    return 0;
}

// Spin-s synthesis of (G, C = 0) from G alone: alm_t (nlm x ld, column b = G of map b) -> ring
// phases of the R and I components (2 nmap columns); FP32 plans of spin >= 1.  The Legendre shape
// of the plan's shape table for this batch (ps_sht_set_shapes) picks the one-tile kernel (0) or
// the wide form (SYN_SHAPE_WR, SYN_SHAPE_WR_NT + nt); otherwise the built-in rule.
int sht_legendre_synthesis_grad(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* phase,
                                hipStream_t st) {
    if (p.prec != 4 || p.spin < 1) {
        set_error("gradient-only synthesis: FP32 spin-s plans only");
        return -1;
    }
    const int shape = shape_of(p, nmap * p.ncomp, SHT_STAGE_SYNTH);
    int nt = synth_sp_nt(nmap);
    if (shape == 0) nt = 1;
    else if (shape >= SYN_SHAPE_WR_NT + 2 && shape <= SYN_SHAPE_WR_NT + 4) nt = shape - SYN_SHAPE_WR_NT;
    if (nt >= 2) return launch_synth_sp_wr<true>(p, nmap, ld, alm_t, phase, st, nt);
    ShtDev<float> d = dev_view<float>(p);
    const int nrb = (p.npairs + 127) / 128, nchunk = (nmap + 15) / 16;
    const unsigned nb = (unsigned)(nrb * nchunk * (p.mmax + 1));
    hipLaunchKernelGGL(k_leg_synth_sp_mfma<true>, dim3(nb), dim3(256), 0, st, d, nmap, ld,
                       (const cplx<float>*)alm_t, (cplx<float>*)phase);
    PS_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int launch_synth_sp_any(const ShtPlan& p, int nmap, int ld, const void* alm_t, void* phase, hipStream_t st,
                               int shape) {
    if (shape < 0) shape = synth_shape_rule<T>(p, nmap * p.ncomp);
    switch (shape) {
        case 0:
            if constexpr (sizeof(T) == 4) return launch_synth_sp_mfma(p, 2 * nmap, ld, alm_t, phase, st);
            else return launch_synth_mfma64(p, 2 * nmap, ld, alm_t, phase, st);
        case SYN_SHAPE_WR:
        case SYN_SHAPE_WR_NT + 2:
        case SYN_SHAPE_WR_NT + 3:
        case SYN_SHAPE_WR_NT + 4:
            if constexpr (sizeof(T) == 4) {
                const int nt = shape == SYN_SHAPE_WR ? synth_sp_nt(2 * nmap) : shape - SYN_SHAPE_WR_NT;
                if (nt < 2) return launch_synth_sp_mfma(p, 2 * nmap, ld, alm_t, phase, st);
                return launch_synth_sp_wr<false>(p, 2 * nmap, ld, alm_t, phase, st, nt);
            }
            break;
        case 1: return launch_synth_sp<T, 1>(p, nmap, ld, alm_t, phase, st);
        case 2: return launch_synth_sp<T, 2>(p, nmap, ld, alm_t, phase, st);
        case 4:
            if constexpr (sizeof(T) == 4) return launch_synth_sp<T, 4>(p, nmap, ld, alm_t, phase, st);
            break;
    }
    set_error("synthesis: invalid launch shape");
    return -1;
}

static int shape_of(const ShtPlan& p, int ncol, int stage) {
    const ShtShapes* s = p.shapes_for(ncol);
    return s ? s->id[stage] : -1;
}

// n: batch columns (spin 0) or maps (spin s)
static int legendre_synthesis(const ShtPlan& p, int n, const void* alm_t, int ld, void* phase, hipStream_t st,
                              int shape) {
    if (p.prec == 4)
        return p.spin == 0 ? launch_synth_s0_any<float>(p, n, ld, alm_t, phase, st, shape)
                           : launch_synth_sp_any<float>(p, n, ld, alm_t, phase, st, shape);
    return p.spin == 0 ? launch_synth_s0_any<double>(p, n, ld, alm_t, phase, st, shape)
                       : launch_synth_sp_any<double>(p, n, ld, alm_t, phase, st, shape);
}

static int ring_synthesis(const ShtPlan& p, int ncol, const void* phase, void* map, long map_stride, hipStream_t st,
                          int nb_pow2, int nb_blue) {
    if (p.prec == 4) return launch_ring_synth<float>(p, ncol, phase, map, map_stride, st, nb_pow2, nb_blue);
    return launch_ring_synth<double>(p, ncol, phase, map, map_stride, st, nb_pow2, nb_blue);
}

// one ring class of the ring synthesis (maps of stride npix)
static int ring_synth_list(const ShtPlan& p, const int* list, int nlist, long maxM, int ncol, const void* phase,
                           void* map, hipStream_t st, int nb) {
    if (p.prec == 4) return launch_ring_synth_list<float>(p, list, nlist, maxM, ncol, phase, map, p.npix, st, nb);
    return launch_ring_synth_list<double>(p, list, nlist, maxM, ncol, phase, map, p.npix, st, nb);
}

static int legendre_adjoint(const ShtPlan& p, int ncol, const void* phase, void* alm, long alm_stride,
                            double scale, void* part, hipStream_t st, int shape) {
    if (p.prec == 4) return launch_adj<float>(p, ncol, phase, alm, alm_stride, scale, part, st, shape);
    return launch_adj<double>(p, ncol, phase, alm, alm_stride, scale, part, st, shape);
}

// Matrix-core spin-0 synthesis with the Bluestein ring FFTs overlapped: with Nside a power of
// two every Bluestein ring is a polar ring (nphi = 4 i, i < Nside), whose phases come from ring
// pairs below Nside - 1 only.  The Legendre stage runs in two launches, the ring-pair blocks
// holding the polar pairs first; the Bluestein rings start on the plan's side stream as soon as
// those are done and run beside the second launch (the matrix-core kernel is bound by its VALU
// stream, the ring FFTs by LDS latency), and the power-of-two rings follow on `st`, which then
// waits for the side stream.  The two launches write disjoint phase entries.  Returns 1 when
// the plan or batch does not qualify (nothing launched).
static int synthesis_overlapped(const ShtPlan& p, int ncol, const void* alm_t, int ld, void* map, long map_stride,
                                void* phase, hipStream_t st) {
    const int nrb = (p.npairs + 127) / 128;
    const int rbA = (p.nside - 1 + 127) / 128;  // blocks holding ring pairs 0 .. Nside - 2
    const bool pow2_nside = (p.nside & (p.nside - 1)) == 0;
    int shape = shape_of(p, ncol, SHT_STAGE_SYNTH);
    if (shape < 0) shape = synth_shape_rule<float>(p, ncol);  // no tuned or set shape: the built-in rule
    if (p.prec != 4 || p.spin != 0 || !pow2_nside || p.nrings == p.n_pow2 || rbA < 1 || rbA >= nrb ||
        (shape != 0 && !synth_is_wr(shape)))
        return 1;
    auto legendre = [&](int r0, int r1) {
        return synth_is_wr(shape) ? launch_synth_s0_wr(p, ncol, ld, alm_t, phase, st, r0, r1, shape)
                                     : launch_synth_s0_mfma(p, ncol, ld, alm_t, phase, st, r0, r1);
    };
    if (!p.side) {
        PS_CHECK_HIP(hipStreamCreateWithFlags(&p.side, hipStreamNonBlocking));
        PS_CHECK_HIP(hipEventCreateWithFlags(&p.ev_fork, hipEventDisableTiming));
        PS_CHECK_HIP(hipEventCreateWithFlags(&p.ev_join, hipEventDisableTiming));
    }
    int rc = legendre(0, rbA);
    if (rc) return rc;
    PS_CHECK_HIP(hipEventRecord(p.ev_fork, st));
    if ((rc = legendre(rbA, nrb))) return rc;
    PS_CHECK_HIP(hipStreamWaitEvent(p.side, p.ev_fork, 0));
    // once the side stream is forked it is always joined back into `st`, whatever fails in
    // between: no unjoined fork under capture, no side-stream work unordered with `st`
    rc = launch_ring_synth_list<float>(p, p.d_rlist + p.n_pow2, p.nrings - p.n_pow2, p.maxM_blue, ncol, phase, map,
                                       map_stride, p.side, shape_of(p, ncol, SHT_STAGE_RING_BLUE));
    const hipError_t ej = hipEventRecord(p.ev_join, p.side);
    if (rc == 0 && ej == hipSuccess)
        rc = launch_ring_synth_list<float>(p, p.d_rlist, p.n_pow2, p.maxM_pow2, ncol, phase, map, map_stride, st,
                                           shape_of(p, ncol, SHT_STAGE_RING_POW2));
    const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(st, p.ev_join, 0) : ej;
    if (rc) return rc;
    PS_CHECK_HIP(ew);
    return 0;
}

int sht_synthesis_lmajor(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* map, long map_stride,
                         void* phase, hipStream_t st) {
    const int rc0 = synthesis_overlapped(p, nmap * p.ncomp, alm_t, ld, map, map_stride, phase, st);
    if (rc0 <= 0) return rc0;
    int rc = sht_legendre_synthesis_lmajor(p, nmap, alm_t, ld, phase, st);
    if (rc) return rc;
    return sht_ring_synthesis(p, nmap, phase, map, map_stride, st);
}

int sht_legendre_synthesis_lmajor(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* phase,
                                  hipStream_t st) {
    const int ncol = nmap * p.ncomp;
    // the spin-0 kernels take batch columns, the spin-s kernels maps
    return legendre_synthesis(p, p.spin == 0 ? ncol : nmap, alm_t, ld, phase, st,
                              shape_of(p, ncol, SHT_STAGE_SYNTH));
}

// ------------------------------------------------ ring synthesis fused with the gNL-loc sum

// Staged shape of one ring class for the fused sum: the plan's shape when it is a staged one,
// else the staged rule; -1 when no staged kernel fits the class.
static int ring_gnl_shape(const ShtPlan& p, int ncol, bool blue) {
    const long maxM = blue ? p.maxM_blue : p.maxM_pow2;
    const int id = shape_of(p, ncol, blue ? SHT_STAGE_RING_BLUE : SHT_STAGE_RING_POW2);
    if (id >= 0 && (id & RING_SG) && ring_nb_valid(p, maxM, ncol, blue, id)) return id;
    if (!blue && ring_nb_valid(p, maxM, ncol, blue, RING_SG | 2)) return RING_SG | 2;
    if (ring_nb_valid(p, maxM, ncol, blue, RING_SG | 1)) return RING_SG | 1;
    return -1;
}

template <bool BLUE>
static int launch_ring_gnl_class(const ShtPlan& p, const int* list, int nlist, long maxM, int ncol,
                                 const void* phase, hipStream_t st, int id, const GnlArgs& g) {
    const int nb = id & ~RING_SG, q = ring_q_of(maxM);
    const size_t sh = ring_sg_lds(p, maxM, nb);
    if (nb == 2 && q == 1)
        return launch_ring_synth_sgq<float, 2, 1, BLUE, true>(p, list, nlist, ncol, phase, nullptr, 0, st, sh, g);
    if (nb == 2 && q == 2)
        return launch_ring_synth_sgq<float, 2, 2, BLUE, true>(p, list, nlist, ncol, phase, nullptr, 0, st, sh, g);
    if (nb == 1 && q == 1)
        return launch_ring_synth_sgq<float, 1, 1, BLUE, true>(p, list, nlist, ncol, phase, nullptr, 0, st, sh, g);
    if (nb == 1 && q == 2)
        return launch_ring_synth_sgq<float, 1, 2, BLUE, true>(p, list, nlist, ncol, phase, nullptr, 0, st, sh, g);
    if (nb == 1 && q == 4)
        return launch_ring_synth_sgq<float, 1, 4, BLUE, true>(p, list, nlist, ncol, phase, nullptr, 0, st, sh, g);
    if constexpr (BLUE) {
        if (nb == 1 && q == 8)
            return launch_ring_synth_sgq<float, 1, 8, true, true>(p, list, nlist, ncol, phase, nullptr, 0, st, sh, g);
    }
    set_error("fused gNL sum: no staged ring kernel for this ring class");
    return -1;
}

// Fixed-order sums of a partials array: each block adds GS_PER consecutive entries (a tree over
// its threads), then one block adds the block sums.
constexpr int GS_PER = 2048;

__global__ __launch_bounds__(256) void k_gnl_parts_level1(const double* __restrict__ part, long n,
                                                          double* __restrict__ out) {
    __shared__ double sh[256];
    const long base = (long)blockIdx.x * GS_PER;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < GS_PER / 256; k++) {
        const long i = base + threadIdx.x + 256 * k;
        s += i < n ? part[i] : 0.0;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_gnl_parts_level2(const double* __restrict__ part, int n, double* out,
                                                          int accumulate) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = accumulate ? *out + sh[0] : sh[0];
}

// Partial sums the fused launch writes (per block of both ring classes, then per level-1 block);
// -1 when the plan has no staged kernel for one of its ring classes.
long sht_ring_gnl_parts(const ShtPlan& p, int nr) {
    const int ncol = 2 * nr;
    long n = 0;
    if (p.n_pow2 > 0) {
        const int id = ring_gnl_shape(p, ncol, false);
        if (id < 0) return -1;
        const int nb = id & ~RING_SG;
        n += (long)p.n_pow2 * ((ncol + 2 * nb - 1) / (2 * nb));
    }
    if (p.nrings > p.n_pow2) {
        const int id = ring_gnl_shape(p, ncol, true);
        if (id < 0) return -1;
        const int nb = id & ~RING_SG;
        n += (long)(p.nrings - p.n_pow2) * ((ncol + 2 * nb - 1) / (2 * nb));
    }
    return n + (n + GS_PER - 1) / GS_PER;
}

int sht_ring_gnl(const ShtPlan& p, int nr, const void* phase, const double* U, const double* V, const double* w,
                 const double* inv_s2, double c4, double c2, double* part, long nparts, double* out, int accumulate,
                 hipStream_t st) {
    if (p.prec != 4 || p.spin != 0) {
        set_error("fused gNL sum: needs an FP32 spin-0 plan");
        return -1;
    }
    if (nr < 1 || (U == nullptr) != (V == nullptr) || !w || !phase || !part || !out) {
        set_error("fused gNL sum: bad arguments");
        return -1;
    }
    const long need = sht_ring_gnl_parts(p, nr);
    if (need < 0) {
        set_error("fused gNL sum: no staged ring kernel fits this plan");
        return -1;
    }
    if (nparts < need) {
        set_error("fused gNL sum: partials buffer too small");
        return -1;
    }
    const int ncol = 2 * nr;
    GnlArgs g;
    g.U = U;
    g.V = V;
    g.w = w;
    g.inv_s2 = inv_s2;
    g.c4 = c4;
    g.c2 = c2;
    g.npix = p.npix;
    g.part = part;
    long n = 0;
    if (p.n_pow2 > 0) {
        const int id = ring_gnl_shape(p, ncol, false), nb = id & ~RING_SG;
        const int rc = launch_ring_gnl_class<false>(p, p.d_rlist, p.n_pow2, p.maxM_pow2, ncol, phase, st, id, g);
        if (rc) return rc;
        n += (long)p.n_pow2 * ((ncol + 2 * nb - 1) / (2 * nb));
    }
    if (p.nrings > p.n_pow2) {
        const int id = ring_gnl_shape(p, ncol, true), nb = id & ~RING_SG;
        g.part = part + n;
        const int rc = launch_ring_gnl_class<true>(p, p.d_rlist + p.n_pow2, p.nrings - p.n_pow2, p.maxM_blue, ncol,
                                                   phase, st, id, g);
        if (rc) return rc;
        n += (long)(p.nrings - p.n_pow2) * ((ncol + 2 * nb - 1) / (2 * nb));
    }
    const int nb1 = (int)((n + GS_PER - 1) / GS_PER);
    hipLaunchKernelGGL(k_gnl_parts_level1, dim3((unsigned)nb1), dim3(256), 0, st, (const double*)part, n, part + n);
    PS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gnl_parts_level2, dim3(1), dim3(256), 0, st, (const double*)(part + n), nb1, out, accumulate);
    PS_LAUNCH_CHECK();
    return 0;
}

int sht_ring_synthesis(const ShtPlan& p, int nmap, const void* phase, void* map, long map_stride, hipStream_t st) {
    const int ncol = nmap * p.ncomp;
    return ring_synthesis(p, ncol, phase, map, map_stride, st, shape_of(p, ncol, SHT_STAGE_RING_POW2),
                          shape_of(p, ncol, SHT_STAGE_RING_BLUE));
}

int sht_alm_to_lmajor(const ShtPlan& p, int ncol, const void* alm, long alm_stride, int ld, void* alm_t,
                      hipStream_t st) {
    const dim3 grid((unsigned)((p.nlm + TRL - 1) / TRL), (unsigned)((ld + TRC - 1) / TRC));
    if (p.prec == 4)
        hipLaunchKernelGGL(k_alm_to_lmajor<float>, grid, dim3(256), 0, st, (const cplx<float>*)alm,
                           alm_stride, ncol, p.nlm, ld, (cplx<float>*)alm_t);
    else
        hipLaunchKernelGGL(k_alm_to_lmajor<double>, grid, dim3(256), 0, st, (const cplx<double>*)alm,
                           alm_stride, ncol, p.nlm, ld, (cplx<double>*)alm_t);
    PS_LAUNCH_CHECK();
    return 0;
}

int sht_adjoint(const ShtPlan& p, int nmap, const void* map, long map_stride, void* alm, long alm_stride,
                double scale, void* work, hipStream_t st) {
    const int ncol = nmap * p.ncomp;
    void* phase = work;
    void* part = (char*)work + ((sht_phase_bytes(p, nmap) + 255) & ~(size_t)255);
    int rc = p.prec == 4 ? launch_ring_adjoint<float>(p, ncol, map, map_stride, phase, st)
                         : launch_ring_adjoint<double>(p, ncol, map, map_stride, phase, st);
    if (rc) return rc;
    return legendre_adjoint(p, ncol, phase, alm, alm_stride, scale, part, st, shape_of(p, ncol, SHT_STAGE_ADJ));
}

// ------------------------------------------------------------ launch-shape tuning

int sht_default_shape(const ShtPlan& p, int nmap, int stage) {
    const int ncol = nmap * p.ncomp;
    switch (stage) {
        case SHT_STAGE_SYNTH:
            return p.prec == 4 ? synth_shape_rule<float>(p, ncol) : synth_shape_rule<double>(p, ncol);
        case SHT_STAGE_RING_POW2: return ring_nb_rule(p, p.maxM_pow2, ncol, false);
        case SHT_STAGE_RING_BLUE: return ring_nb_rule(p, p.maxM_blue, ncol, true);
        case SHT_STAGE_ADJ: {
            const int ns = sht_adjoint_nsplit(p, nmap);
            const bool many = p.spin == 0 ? ncol > 2 : nmap > 1;
            return adj_shape_rule(p, ncol, many, (p.npairs + ns - 1) / ns);
        }
    }
    return -1;
}

int sht_shape_candidates(const ShtPlan& p, int nmap, int stage, int* ids, int max) {
    const int ncol = nmap * p.ncomp;
    int buf[64], n = 0;
    auto add = [&](int id) {
        if (n < 64) buf[n++] = id;
    };
    switch (stage) {
        case SHT_STAGE_SYNTH: {
            // vector kernels whose chunk is between half and all of the (power-of-two rounded)
            // batch, capped at the widest instantiated chunk; the matrix-core kernel in FP32
            const int w = p.spin == 0 ? ncol : nmap;
            const int bcmax = p.spin == 0 ? (p.prec == 4 ? SYN_BC_S0_F : SYN_BC_S0_D)
                                          : (p.prec == 4 ? SYN_BC_SP_F : SYN_BC_SP_D);
            int top = 1;
            while (top < w && top < bcmax) top *= 2;
            add(0);  // the matrix-core kernel (FP32 or FP64)
            if (p.prec == 4 && p.spin == 0) add(SYN_SHAPE_WR);
            if (p.prec == 4 && p.spin == 0 && ncol > 16)
                for (int nt = 2; nt < synth_lt_nt(ncol) && nt <= 4; nt++) add(SYN_SHAPE_WR_NT + nt);
            if (p.prec == 4 && p.spin > 0 && ncol > 16) {  // ncol = 2 nmap complex columns
                add(SYN_SHAPE_WR);
                for (int nt = 3; nt <= 4 && 16 * (nt - 1) < ncol; nt++) add(SYN_SHAPE_WR_NT + nt);
            }
            for (int bc = std::max(1, top / 2); bc <= top; bc *= 2) add(bc);
            break;
        }
        case SHT_STAGE_RING_POW2:
        case SHT_STAGE_RING_BLUE: {
            const long maxM = stage == SHT_STAGE_RING_POW2 ? p.maxM_pow2 : p.maxM_blue;
            const bool blue = stage == SHT_STAGE_RING_BLUE;
            for (int nb = 1; nb <= 4; nb *= 2)
                if (ring_nb_valid(p, maxM, ncol, blue, nb)) add(nb);
            for (int nb = 1; nb <= 2; nb *= 2)
                if (ring_nb_valid(p, maxM, ncol, blue, RING_STK | nb)) add(RING_STK | nb);
            for (int nb = 1; nb <= 2; nb *= 2)
                if (ring_nb_valid(p, maxM, ncol, blue, RING_SG | nb)) add(RING_SG | nb);
            break;
        }
        case SHT_STAGE_ADJ: {
            // shapes covering the ring slice with at most 4x idle threads
            const int ns = sht_adjoint_nsplit(p, nmap);
            const int pps = (p.npairs + ns - 1) / ns;
            for (int id = 0; id < N_ADJ_SHAPES; id++) {
                const int cover = 64 * ADJ_SHAPES[id].nrt * ADJ_SHAPES[id].nw;
                if (adj_shape_valid(p, id, pps) && (cover <= 4 * pps || cover == 64)) add(id);
            }
            if (adj_mfma_ok(p)) add(ADJ_SHAPE_MFMA);
            if (p.spin > 0 && nmap > (p.prec == 4 ? 8 : 4)) add(ADJ_SHAPE_MFMA2);
            break;
        }
        default: set_error("launch shapes: stage must be 0..3"); return -1;
    }
    // the built-in choice is always a candidate
    const int def = sht_default_shape(p, nmap, stage);
    bool has_def = false;
    for (int i = 0; i < n; i++) has_def = has_def || buf[i] == def;
    if (!has_def) add(def);
    for (int i = 0; ids && i < n && i < max; i++) ids[i] = buf[i];
    return n;
}

static bool shape_allowed(const ShtPlan& p, int nmap, int stage, int id) {
    int ids[64];
    const int n = sht_shape_candidates(p, nmap, stage, ids, 64);
    for (int i = 0; i < n && i < 64; i++)
        if (ids[i] == id) return true;
    return id == sht_default_shape(p, nmap, stage);
}

int sht_set_shapes(ShtPlan& p, int nmap, const int* ids) {
    const int ncol = nmap * p.ncomp;
    for (int st = 0; st < SHT_NSTAGE; st++)
        if (ids[st] >= 0 && !shape_allowed(p, nmap, st, ids[st])) {
            set_error("launch shapes: shape " + std::to_string(ids[st]) + " is not a candidate of stage " +
                      std::to_string(st) + " for this batch");
            return -1;
        }
    ShtShapes* e = nullptr;
    for (int i = 0; i < p.nshapes; i++)
        if (p.shapes[i].ncol == ncol) e = &p.shapes[i];
    if (!e) {
        // a full table drops the oldest batch width, which returns to the built-in rule
        if (p.nshapes == SHT_MAX_SHAPES) {
            for (int i = 1; i < SHT_MAX_SHAPES; i++) p.shapes[i - 1] = p.shapes[i];
            p.nshapes--;
        }
        e = &p.shapes[p.nshapes++];
    }
    e->ncol = ncol;
    for (int st = 0; st < SHT_NSTAGE; st++) e->id[st] = ids[st];
    return 0;
}

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t sht_autotune_work_bytes(const ShtPlan& p, int nmap) {
    const int ncol = nmap * p.ncomp;
    const int mult = sht_synth_lmajor_ld_multiple(p);
    const long ld = (ncol + mult - 1) / mult * mult;
    return al256((size_t)ld * p.nlm * 2 * p.prec) + al256((size_t)ncol * p.npix * p.prec) +
           al256((size_t)ncol * p.nlm * 2 * p.prec) +
           al256(std::max(sht_phase_bytes(p, nmap), sht_adjoint_work_bytes(p, nmap)));
}

// Bounded synthetic coefficients in [-0.5, 0.5) (a multiplicative hash of the index); the
// padding columns of the l-major layout are zero.
template <typename T>
__global__ void k_tune_fill(cplx<T>* x, long n, int ld, int ncol) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const unsigned h = (unsigned)i * 2654435761u;
        const bool live = (int)(i % ld) < ncol;
        x[i] = live ? cplx<T>{T((h >> 8) * (1.0 / 16777216.0) - 0.5), T(((h ^ 0x9e3779b9u) >> 8) * (1.0 / 16777216.0) - 0.5)}
                    : cplx<T>{T(0), T(0)};
    }
}

namespace {
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace

// one untimed launch, then reps launches between two events; milliseconds per launch
template <class F>
static int time_shape(F&& launch, int reps, hipStream_t st, EventPair& ev, double* ms) {
    int rc = launch();
    if (rc) return rc;
    PS_CHECK_HIP(hipEventRecord(ev.a, st));
    for (int i = 0; i < reps; i++)
        if ((rc = launch())) return rc;
    PS_CHECK_HIP(hipEventRecord(ev.b, st));
    PS_CHECK_HIP(hipEventSynchronize(ev.b));
    float t = 0.f;
    PS_CHECK_HIP(hipEventElapsedTime(&t, ev.a, ev.b));
    *ms = (double)t / reps;
    return 0;
}

int sht_autotune(ShtPlan& p, int nmap, void* work, size_t work_bytes, int reps, hipStream_t st, double* times) {
    if (work_bytes < sht_autotune_work_bytes(p, nmap)) {
        set_error("autotune: workspace too small");
        return -1;
    }
    if (reps < 1) reps = 3;
    const int ncol = nmap * p.ncomp;
    const int mult = sht_synth_lmajor_ld_multiple(p);
    const int ld = (ncol + mult - 1) / mult * mult;
    char* w = (char*)work;
    void* alm_t = w;
    w += al256((size_t)ld * p.nlm * 2 * p.prec);
    void* map = w;
    w += al256((size_t)ncol * p.npix * p.prec);
    void* alm = w;
    w += al256((size_t)ncol * p.nlm * 2 * p.prec);
    void* phase = w;
    void* part = (char*)phase + al256(sht_phase_bytes(p, nmap));

    const long nel = (long)ld * p.nlm;
    const unsigned nbf = (unsigned)std::min<long>((nel + 255) / 256, 16384);
    if (p.prec == 4)
        hipLaunchKernelGGL(k_tune_fill<float>, dim3(nbf), dim3(256), 0, st, (cplx<float>*)alm_t, nel, ld, ncol);
    else
        hipLaunchKernelGGL(k_tune_fill<double>, dim3(nbf), dim3(256), 0, st, (cplx<double>*)alm_t, nel, ld, ncol);
    PS_LAUNCH_CHECK();

    EventPair ev;
    PS_CHECK_HIP(hipEventCreate(&ev.a));
    PS_CHECK_HIP(hipEventCreate(&ev.b));
    int chosen[SHT_NSTAGE];
    const int n_syn = p.spin == 0 ? ncol : nmap;
    for (int stage = 0; stage < SHT_NSTAGE; stage++) {
        int ids[64];
        const int nc = std::min(64, sht_shape_candidates(p, nmap, stage, ids, 64));
        const int def = sht_default_shape(p, nmap, stage);
        chosen[stage] = def;
        if (times) times[2 * stage] = times[2 * stage + 1] = 0.0;
        const bool empty = (stage == SHT_STAGE_RING_POW2 && p.n_pow2 == 0) ||
                           (stage == SHT_STAGE_RING_BLUE && p.nrings == p.n_pow2);
        if (empty) continue;
        if (stage == SHT_STAGE_ADJ) {
            // phases of the adjoint layout, once, from the synthesised maps
            int rc = p.prec == 4 ? launch_ring_adjoint<float>(p, ncol, map, p.npix, phase, st)
                                 : launch_ring_adjoint<double>(p, ncol, map, p.npix, phase, st);
            if (rc) return rc;
        }
        auto run = [&](int id) -> int {
            switch (stage) {
                case SHT_STAGE_SYNTH: return legendre_synthesis(p, n_syn, alm_t, ld, phase, st, id);
                case SHT_STAGE_RING_POW2:
                    return ring_synth_list(p, p.d_rlist, p.n_pow2, p.maxM_pow2, ncol, phase, map, st, id);
                case SHT_STAGE_RING_BLUE:
                    return ring_synth_list(p, p.d_rlist + p.n_pow2, p.nrings - p.n_pow2, p.maxM_blue, ncol,
                                           phase, map, st, id);
                default: return legendre_adjoint(p, ncol, phase, alm, p.nlm, 1.0, part, st, id);
            }
        };
        double t_def = 0.0;
        int rc = time_shape([&] { return run(def); }, reps, st, ev, &t_def);
        if (rc) return rc;
        double t_best = t_def;
        int best = def;
        for (int i = 0; i < nc; i++) {
            if (ids[i] == def) continue;
            double t = 0.0;
            if ((rc = time_shape([&] { return run(ids[i]); }, reps, st, ev, &t))) return rc;
            if (t < t_best) {
                t_best = t;
                best = ids[i];
            }
        }
        if (best != def && t_best < 0.97 * t_def) chosen[stage] = best;
        else t_best = t_def;
        if (times) {
            times[2 * stage] = t_best;
            times[2 * stage + 1] = t_def;
        }
        if (stage == SHT_STAGE_SYNTH) {
            // the ring stages read phases of the chosen synthesis
            if ((rc = run(chosen[stage]))) return rc;
        }
    }
    return sht_set_shapes(p, nmap, chosen);
}

}  // namespace ps
