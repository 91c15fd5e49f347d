/*
 * FP64 CPU oracle for the HEALPix spherical harmonic transforms on PolySpec's
 * hot path.  TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP
 * kernels are compared against (and the timed CPU baseline); it is never part
 * of the product path.
 *
 * What it restates (the reference calls ducc0 for these; ducc0 is a third-party
 * C++ library that is not vendored in the reference tree, version unpinned):
 *   - ducc0.sht.synthesis(alm, spin, lmax, mmax, **Healpix_Base(nside,"RING").sht_info())
 *       polyspec/base.py:172-176, 225, 230-231, 282
 *   - ducc0.sht.adjoint_synthesis(map, spin, ...)   (unscaled; the caller
 *       multiplies by pi/(3 nside^2) = 4 pi/Npix)   polyspec/base.py:147-151, 193, 206, 253
 *
 * Algorithm (the published libsharp/ducc method, Reinecke & Seljebotn 2013):
 *   HEALPix RING geometry (Gorski et al. 2005, ApJ 622:759, sec. 4);
 *   per-m three-term l-recursion of the normalised spin-weighted Legendre
 *   functions, seeded at l0 = max(m, s) in log space and carried with an
 *   integer power-of-two exponent so that sin^m(theta) seeds never underflow;
 *   north/south ring pairs share one recursion through the (-1)^(l+m) parity;
 *   per-ring DFTs of length nphi (4i polar, 4 nside belt) by a mixed-radix
 *   Cooley-Tukey FFT, with m folded modulo nphi (aliasing) when m > nphi/2.
 *
 * Conventions (HEALPix / healpy / ducc, spin s > 0):
 *   sY_lm is the Goldberg et al. (1967) spin-weighted harmonic (Condon-Shortley
 *   phase at s = 0).  A spin-s pair of real maps (R, I) is synthesised from
 *   gradient/curl coefficients (G, C) as
 *       R + iI = sum_lm -(G + iC)_lm  sY_lm ,
 *   i.e. for s = 2: Q + iU = -sum (a_E + i a_B) 2Y_lm.  The adjoint is the exact
 *   transpose of synthesis under <a,b> = sum_{m>=0} (1+[m>0]) Re(a* b).
 *
 * Layouts: alm complex128 (nmaps, ncomp, nlm) with healpy m-major index
 * idx(l,m) = m(2 lmax + 1 - m)/2 + l; maps float64 (nmaps, ncomp, 12 nside^2)
 * in RING order; ncomp = 1 for spin 0, 2 for spin > 0.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------ geometry */

/* Ring i (1-based, 1 .. 4 nside - 1) of the RING scheme. */
void orc_ring_info(int nside, int i, double *z, double *sth, double *phi0,
                   int *nphi, long *start)
{
    long npix = 12L * nside * nside;
    int ir = (i > 2 * nside) ? 4 * nside - i : i; /* north-hemisphere ring */
    double zz, st, p0;
    int np;
    long st0;
    if (ir < nside) {
        double omz = (double)ir * ir / (3.0 * nside * nside); /* 1 - z */
        zz = 1.0 - omz;
        st = sqrt(omz * (2.0 - omz));
        np = 4 * ir;
        p0 = M_PI / (4.0 * ir);
        st0 = 2L * ir * (ir - 1);
    } else {
        zz = 2.0 * (2 * nside - ir) / (3.0 * nside);
        st = sqrt((1.0 - zz) * (1.0 + zz));
        np = 4 * nside;
        p0 = (((ir - nside) & 1) == 0) ? M_PI / (4.0 * nside) : 0.0;
        st0 = 2L * nside * (nside - 1) + (long)(ir - nside) * 4 * nside;
    }
    if (i > 2 * nside) {
        zz = -zz;
        st0 = npix - st0 - np;
    }
    *z = zz;
    *sth = st;
    *phi0 = p0;
    *nphi = np;
    *start = st0;
}

static long nlm_of(int lmax, int mmax)
{
    return (long)(mmax + 1) * (lmax + 1) - (long)mmax * (mmax + 1) / 2;
}

static long idx_lm(int lmax, int l, int m)
{
    return (long)m * (2 * lmax + 1 - m) / 2 + l;
}

/* ---------------------------------------------------------------------- FFT */

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; (long)p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* out[k] = sum_j in[j*is] * w^(jk) for k < n, w = exp(sign 2 pi i / n).
 * wt is the table exp(sign 2 pi i t / N) for t < N, N a multiple of n. */
static void fft_rec(int n, const cplx *in, long is, cplx *out, const cplx *wt,
                    int N, cplx *tmp)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_factor(n), q = n / p;
    long ws = N / n;
    for (int r = 0; r < p; r++)
        fft_rec(q, in + r * is, is * p, out + (long)r * q, wt, N, tmp);
    /* butterflies: X[k + s q] = sum_r w^(r (k + s q)) Y_r[k] */
    for (int k = 0; k < q; k++) {
        for (int r = 0; r < p; r++) tmp[r] = out[(long)r * q + k];
        for (int s = 0; s < p; s++) {
            cplx acc = 0;
            long kk = k + (long)s * q;
            for (int r = 0; r < p; r++) acc += tmp[r] * wt[((long)r * kk * ws) % N];
            out[kk] = acc;
        }
    }
}

/* sign = +1: out_k = sum_j in_j exp(+2 pi i jk/n); sign = -1 the forward DFT. */
static void cdft(int n, const cplx *in, cplx *out, int sign, cplx *wt, cplx *tmp)
{
    for (int t = 0; t < n; t++) wt[t] = cexp(sign * 2.0 * M_PI * I * t / n);
    fft_rec(n, in, 1, out, wt, n, tmp);
}

/* ------------------------------------------------------ Legendre recursion */

/* Recursion state of one normalised spin-weighted Legendre function
 * f_l = lambda_{l m}^{(sign s)}(theta) for fixed m, carried as value * 2^(64 k). */
typedef struct {
    double fm1, f;  /* f_{l-1}, f_l in scaled units */
    int k;          /* scale index: true value = scaled * 2^(64 k) */
} rec_t;

#define SCALE_LOG2 64
static const double SC_DN = 5.42101086242752217003726400434970855712890625e-20; /* 2^-64 */
static const double SC_THR = 1152921504606846976.0;          /* 2^60 */

static double scale_factor(int k)
{
    if (k == 0) return 1.0;
    if (k == -1) return SC_DN;
    if (k == -2) return SC_DN * SC_DN;
    return 0.0;
}

/* log|seed| and sign of f_{l0} for spin ss (may be negative), l0 = max(m, |ss|). */
static void seed_log(int m, int ss, double theta, double *logv, double *sgn)
{
    int s = abs(ss);
    double lsh = log(sin(0.5 * theta)), lch = log(cos(0.5 * theta));
    if (s == 0) {
        /* lambda_mm = (-1)^m sqrt((2m+1)/(4pi) (2m)!/(4^m m!^2)) sin^m theta */
        double lv = 0.5 * (log((2.0 * m + 1.0) / (4.0 * M_PI)) + lgamma(2.0 * m + 1.0) -
                           2.0 * lgamma(m + 1.0) - 2.0 * m * log(2.0));
        lv += m * log(sin(theta));
        *logv = lv;
        *sgn = (m & 1) ? -1.0 : 1.0;
        return;
    }
    if (m >= s) {
        double lv = 0.5 * (log((2.0 * m + 1.0) / (4.0 * M_PI)) + lgamma(2.0 * m + 1.0) -
                           lgamma(m + s + 1.0) - lgamma(m - s + 1.0));
        if (ss > 0)
            lv += (m + s) * lsh + (m - s) * lch;
        else
            lv += (m - s) * lsh + (m + s) * lch;
        *logv = lv;
        *sgn = (m & 1) ? -1.0 : 1.0;
    } else {
        double lv = 0.5 * (log((2.0 * s + 1.0) / (4.0 * M_PI)) + lgamma(2.0 * s + 1.0) -
                           lgamma(s + m + 1.0) - lgamma(s - m + 1.0));
        if (ss > 0) {
            lv += (s + m) * lsh + (s - m) * lch;
            *sgn = (m & 1) ? -1.0 : 1.0;
        } else {
            lv += (s - m) * lsh + (s + m) * lch;
            *sgn = (s & 1) ? -1.0 : 1.0;
        }
        *logv = lv;
    }
}

static void rec_init(rec_t *r, int m, int ss, double theta)
{
    double lv, sg;
    seed_log(m, ss, theta, &lv, &sg);
    double l2 = lv / M_LN2;
    int k = 0;
    if (l2 < 0) k = (int)floor(l2 / SCALE_LOG2) + 1;
    if (k > 0) k = 0;
    r->fm1 = 0.0;
    r->f = sg * exp2(l2 - (double)SCALE_LOG2 * k);
    r->k = k;
}

/* Coefficients of f_{l+1} = alpha (x + beta) f_l - gamma f_{l-1} (Wigner-d
 * three-term recursion, normalised by sqrt((2l+1)/4pi)); ss is the signed spin. */
static void rec_coef(int l, int m, int ss, double *alpha, double *beta, double *gam)
{
    double s = (double)abs(ss), lp = l + 1.0, dl = l;
    double den = sqrt((lp * lp - (double)m * m) * (lp * lp - s * s));
    *alpha = (2.0 * dl + 1.0) * lp * sqrt((2.0 * dl + 3.0) / (2.0 * dl + 1.0)) / den;
    if (l == 0) {
        *beta = 0.0;
        *gam = 0.0;
        return;
    }
    *beta = (ss == 0) ? 0.0 : (double)m * ss / (dl * lp);
    *gam = lp / dl * sqrt((dl * dl - (double)m * m) * (dl * dl - s * s)) / den *
           sqrt((2.0 * dl + 3.0) / (2.0 * dl - 1.0));
}

static void rec_step(rec_t *r, double x, double alpha, double beta, double gam)
{
    double fn = alpha * (x + beta) * r->f - gam * r->fm1;
    r->fm1 = r->f;
    r->f = fn;
    if (r->k < 0 && fabs(fn) > SC_THR) {
        r->f *= SC_DN;
        r->fm1 *= SC_DN;
        r->k += 1;
    }
}

/* --------------------------------------------------------------- transforms */

typedef struct {
    int nside, nrings, npairs;
    double *theta, *phi0;
    int *nphi;
    long *start;
} geom_t;

static void geom_build(geom_t *g, int nside)
{
    g->nside = nside;
    g->nrings = 4 * nside - 1;
    g->npairs = 2 * nside; /* north rings 1..2nside; the last is the equator */
    g->theta = malloc(sizeof(double) * g->nrings);
    g->phi0 = malloc(sizeof(double) * g->nrings);
    g->nphi = malloc(sizeof(int) * g->nrings);
    g->start = malloc(sizeof(long) * g->nrings);
    for (int i = 1; i <= g->nrings; i++) {
        double z, st;
        orc_ring_info(nside, i, &z, &st, &g->phi0[i - 1], &g->nphi[i - 1], &g->start[i - 1]);
        g->theta[i - 1] = atan2(st, z);
    }
}

static void geom_free(geom_t *g)
{
    free(g->theta);
    free(g->phi0);
    free(g->nphi);
    free(g->start);
}

/* Synthesis phases -> one ring of real map values. ph: (mmax+1) phases for the ring. */
static void ring_synth(int nphi, double phi0, int mmax, const cplx *ph, double *out,
                       cplx *c, cplx *f, cplx *wt, cplx *tmp)
{
    for (int j = 0; j < nphi; j++) c[j] = 0;
    for (int m = 0; m <= mmax; m++) {
        cplx g = ph[m] * cexp(I * m * phi0);
        c[m % nphi] += g;
        if (m > 0) c[(nphi - m % nphi) % nphi] += conj(g);
    }
    cdft(nphi, c, f, +1, wt, tmp);
    for (int k = 0; k < nphi; k++) out[k] = creal(f[k]);
}

/* One ring of real map values -> adjoint phases sum_k f_k exp(-i m phi_k). */
static void ring_adjoint(int nphi, double phi0, int mmax, const double *in, cplx *ph,
                         cplx *c, cplx *f, cplx *wt, cplx *tmp)
{
    for (int k = 0; k < nphi; k++) c[k] = in[k];
    cdft(nphi, c, f, -1, wt, tmp);
    for (int m = 0; m <= mmax; m++) ph[m] = f[m % nphi] * cexp(-I * m * phi0);
}

/* Spin-0 or spin-s synthesis. Returns 0 on success. */
int orc_sht_synthesis(int nside, int lmax, int mmax, int spin, int nmaps,
                      const cplx *alm, double *map, int nthreads)
{
    if (nside < 1 || lmax < 0 || mmax < 0 || mmax > lmax || spin < 0 || nmaps < 1) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    geom_t g;
    geom_build(&g, nside);
    int ncomp = spin ? 2 : 1;
    long nlm = nlm_of(lmax, mmax), npix = 12L * nside * nside;
    long nph = (long)g.nrings * (mmax + 1);
    cplx *ph = calloc((size_t)nmaps * ncomp * nph, sizeof(cplx));

#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m <= mmax; m++) {
        int l0 = (m > spin) ? m : spin;
        for (int jp = 0; jp < g.npairs; jp++) {
            int in_ = jp, is_ = g.nrings - 1 - jp;
            double th = g.theta[in_], x = cos(th);
            for (int b = 0; b < nmaps; b++) {
                if (spin == 0) {
                    const cplx *a = alm + (long)b * nlm;
                    cplx ev = 0, od = 0;
                    rec_t r;
                    rec_init(&r, m, 0, th);
                    for (int l = m; l <= lmax; l++) {
                        double v = r.f * scale_factor(r.k);
                        if (((l - m) & 1) == 0) ev += v * a[idx_lm(lmax, l, m)];
                        else od += v * a[idx_lm(lmax, l, m)];
                        double al, be, ga;
                        rec_coef(l, m, 0, &al, &be, &ga);
                        rec_step(&r, x, al, be, ga);
                    }
                    ph[(long)b * nph + (long)in_ * (mmax + 1) + m] = ev + od;
                    if (is_ != in_) ph[(long)b * nph + (long)is_ * (mmax + 1) + m] = ev - od;
                } else {
                    const cplx *G = alm + ((long)b * 2) * nlm;
                    const cplx *C = alm + ((long)b * 2 + 1) * nlm;
                    /* accumulators [parity][G lam+, C lam+, G lam-, C lam-] */
                    cplx acc[2][4] = {{0}};
                    rec_t rp, rm;
                    rec_init(&rp, m, spin, th);
                    rec_init(&rm, m, -spin, th);
                    /* sY_{l,-m} = (-1)^(s+m) conj(-sY_lm): the m < 0 half of a real
                       field pairs +s with (-1)^s (-s); and sY(pi - theta) = (-1)^(l+m)
                       (-s)Y(theta), so the south parity of lam+- carries (-1)^s too */
                    const double nsg = (spin & 1) ? -1.0 : 1.0;
                    for (int l = l0; l <= lmax; l++) {
                        double fp = rp.f * scale_factor(rp.k), fm = nsg * rm.f * scale_factor(rm.k);
                        double lp = 0.5 * (fp + fm), lmn = 0.5 * (fp - fm);
                        int par = (l + m + spin) & 1;
                        cplx gv = G[idx_lm(lmax, l, m)], cv = C[idx_lm(lmax, l, m)];
                        acc[par][0] += lp * gv;
                        acc[par][1] += lp * cv;
                        acc[par][2] += lmn * gv;
                        acc[par][3] += lmn * cv;
                        double al, be, ga;
                        rec_coef(l, m, spin, &al, &be, &ga);
                        rec_step(&rp, x, al, be, ga);
                        rec_coef(l, m, -spin, &al, &be, &ga);
                        rec_step(&rm, x, al, be, ga);
                    }
                    /* north: lam(+-) as is; south: lam+ -> (-1)^(l+m+s) lam+,
                       lam- -> -(-1)^(l+m+s) lam- */
                    cplx Rn = -(acc[0][0] + acc[1][0] + I * (acc[0][3] + acc[1][3]));
                    cplx In = -(acc[0][1] + acc[1][1] - I * (acc[0][2] + acc[1][2]));
                    cplx Rs = -(acc[0][0] - acc[1][0] + I * (acc[1][3] - acc[0][3]));
                    cplx Is = -(acc[0][1] - acc[1][1] - I * (acc[1][2] - acc[0][2]));
                    ph[((long)b * 2) * nph + (long)in_ * (mmax + 1) + m] = Rn;
                    ph[((long)b * 2 + 1) * nph + (long)in_ * (mmax + 1) + m] = In;
                    if (is_ != in_) {
                        ph[((long)b * 2) * nph + (long)is_ * (mmax + 1) + m] = Rs;
                        ph[((long)b * 2 + 1) * nph + (long)is_ * (mmax + 1) + m] = Is;
                    }
                }
            }
        }
    }

    int maxn = 4 * nside;
#pragma omp parallel
    {
        cplx *c = malloc(sizeof(cplx) * maxn), *f = malloc(sizeof(cplx) * maxn);
        cplx *wt = malloc(sizeof(cplx) * maxn), *tmp = malloc(sizeof(cplx) * maxn);
#pragma omp for schedule(dynamic, 4)
        for (long t = 0; t < (long)nmaps * ncomp * g.nrings; t++) {
            long bc = t / g.nrings;
            int ir = (int)(t % g.nrings);
            ring_synth(g.nphi[ir], g.phi0[ir], mmax, ph + bc * nph + (long)ir * (mmax + 1),
                       map + bc * npix + g.start[ir], c, f, wt, tmp);
        }
        free(c);
        free(f);
        free(wt);
        free(tmp);
    }
    free(ph);
    geom_free(&g);
    return 0;
}

/* Unscaled adjoint of orc_sht_synthesis. */
int orc_sht_adjoint(int nside, int lmax, int mmax, int spin, int nmaps, const double *map,
                    cplx *alm, int nthreads)
{
    if (nside < 1 || lmax < 0 || mmax < 0 || mmax > lmax || spin < 0 || nmaps < 1) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    geom_t g;
    geom_build(&g, nside);
    int ncomp = spin ? 2 : 1;
    long nlm = nlm_of(lmax, mmax), npix = 12L * nside * nside;
    long nph = (long)g.nrings * (mmax + 1);
    cplx *ph = calloc((size_t)nmaps * ncomp * nph, sizeof(cplx));
    int maxn = 4 * nside;
#pragma omp parallel
    {
        cplx *c = malloc(sizeof(cplx) * maxn), *f = malloc(sizeof(cplx) * maxn);
        cplx *wt = malloc(sizeof(cplx) * maxn), *tmp = malloc(sizeof(cplx) * maxn);
#pragma omp for schedule(dynamic, 4)
        for (long t = 0; t < (long)nmaps * ncomp * g.nrings; t++) {
            long bc = t / g.nrings;
            int ir = (int)(t % g.nrings);
            ring_adjoint(g.nphi[ir], g.phi0[ir], mmax, map + bc * npix + g.start[ir],
                         ph + bc * nph + (long)ir * (mmax + 1), c, f, wt, tmp);
        }
        free(c);
        free(f);
        free(wt);
        free(tmp);
    }
    memset(alm, 0, sizeof(cplx) * (size_t)nmaps * ncomp * nlm);

#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m <= mmax; m++) {
        int l0 = (m > spin) ? m : spin;
        for (int jp = 0; jp < g.npairs; jp++) {
            int in_ = jp, is_ = g.nrings - 1 - jp;
            double th = g.theta[in_], x = cos(th);
            for (int b = 0; b < nmaps; b++) {
                if (spin == 0) {
                    cplx pn = ph[(long)b * nph + (long)in_ * (mmax + 1) + m];
                    cplx ps = (is_ != in_) ? ph[(long)b * nph + (long)is_ * (mmax + 1) + m] : 0;
                    cplx ev = pn + ps, od = pn - ps;
                    cplx *a = alm + (long)b * nlm;
                    rec_t r;
                    rec_init(&r, m, 0, th);
                    for (int l = m; l <= lmax; l++) {
                        double v = r.f * scale_factor(r.k);
                        a[idx_lm(lmax, l, m)] += v * ((((l - m) & 1) == 0) ? ev : od);
                        double al, be, ga;
                        rec_coef(l, m, 0, &al, &be, &ga);
                        rec_step(&r, x, al, be, ga);
                    }
                } else {
                    cplx Rn = ph[((long)b * 2) * nph + (long)in_ * (mmax + 1) + m];
                    cplx In = ph[((long)b * 2 + 1) * nph + (long)in_ * (mmax + 1) + m];
                    cplx Rs = 0, Is = 0;
                    if (is_ != in_) {
                        Rs = ph[((long)b * 2) * nph + (long)is_ * (mmax + 1) + m];
                        Is = ph[((long)b * 2 + 1) * nph + (long)is_ * (mmax + 1) + m];
                    }
                    /* sums for lam+ (sign (-1)^(l+m+s) south) and lam- (sign -(-1)^(l+m+s)) */
                    cplx Rsum = Rn + Rs, Rdif = Rn - Rs, Isum = In + Is, Idif = In - Is;
                    cplx *G = alm + ((long)b * 2) * nlm, *C = alm + ((long)b * 2 + 1) * nlm;
                    rec_t rp, rm;
                    rec_init(&rp, m, spin, th);
                    rec_init(&rm, m, -spin, th);
                    /* sY_{l,-m} = (-1)^(s+m) conj(-sY_lm): the m < 0 half of a real
                       field pairs +s with (-1)^s (-s); and sY(pi - theta) = (-1)^(l+m)
                       (-s)Y(theta), so the south parity of lam+- carries (-1)^s too */
                    const double nsg = (spin & 1) ? -1.0 : 1.0;
                    for (int l = l0; l <= lmax; l++) {
                        double fp = rp.f * scale_factor(rp.k), fm = nsg * rm.f * scale_factor(rm.k);
                        double lp = 0.5 * (fp + fm), lmn = 0.5 * (fp - fm);
                        int even = (((l + m + spin) & 1) == 0);
                        cplx Rp = even ? Rsum : Rdif, Ip = even ? Isum : Idif; /* with lam+ */
                        cplx Rm = even ? Rdif : Rsum, Im = even ? Idif : Isum; /* with lam- */
                        long id = idx_lm(lmax, l, m);
                        G[id] += -(lp * Rp + I * lmn * Im);
                        C[id] += -(lp * Ip - I * lmn * Rm);
                        double al, be, ga;
                        rec_coef(l, m, spin, &al, &be, &ga);
                        rec_step(&rp, x, al, be, ga);
                        rec_coef(l, m, -spin, &al, &be, &ga);
                        rec_step(&rm, x, al, be, ga);
                    }
                }
            }
        }
    }
    free(ph);
    geom_free(&g);
    return 0;
}

/* Values of the normalised spin-weighted Legendre function sY_lm(theta, 0)
 * for l = 0..lmax at fixed m and signed spin ss (zero below l0); used by the
 * tests to pin the recursion against closed forms. */
void orc_spin_legendre(int lmax, int m, int ss, double theta, double *out)
{
    int l0 = (m > abs(ss)) ? m : abs(ss);
    for (int l = 0; l <= lmax; l++) out[l] = 0.0;
    rec_t r;
    rec_init(&r, m, ss, theta);
    double x = cos(theta);
    for (int l = l0; l <= lmax; l++) {
        out[l] = r.f * scale_factor(r.k);
        double al, be, ga;
        rec_coef(l, m, ss, &al, &be, &ga);
        rec_step(&r, x, al, be, ga);
    }
}
