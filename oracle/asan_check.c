/* Memory-safety driver for the FP64 CPU oracle (test infrastructure only).
 *
 * Built by `make -C oracle asan` with AddressSanitizer and UndefinedBehaviorSanitizer (no
 * OpenMP: the sanitizers then see every access in one thread), it calls each orc_* entry point
 * on small inputs whose buffers are allocated at exactly the sizes the entry point documents, so
 * that a read or write one element past any of them aborts the run.  Shapes cover the edge cases
 * the estimators reach: Nside 1 and 2 (rings shorter than mmax), lmax above 3 Nside - 1, spin 0
 * and 2, one and several maps, radial-weight vectors with zeros, empty L ranges.
 * Prints "asan_check ok" and exits 0 when every output is finite. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

typedef double complex cplx;

int orc_sht_synthesis(int nside, int lmax, int mmax, int spin, int nmaps, const cplx *alm,
                      double *map, int nthreads);
int orc_sht_adjoint(int nside, int lmax, int mmax, int spin, int nmaps, const double *map,
                    cplx *alm, int nthreads);
void orc_spin_legendre(int lmax, int m, int ss, double theta, double *out);
void orc_ring_info(int nside, int i, double *z, double *sth, double *phi0, int *nphi, long *start);
void orc_apply_fl_weights(const double *flX, int nslot, int nr, const cplx *h, int npol, long nl,
                          const int *ls, double weight, cplx *out);
double orc_gnl_loc_sum(const double *w, const double *P, const double *Q, int nr, long npix);
double orc_gnl_loc_disc_sum(const double *w, const double *P1, const double *P2, const double *Q1,
                            const double *Q2, int nr, long npix);
void orc_radial_sum(const cplx *lm, int nr, long nlm, const double *w, const double *flXs,
                    int npol, const int *ls, cplx *out);
double orc_tau_sum_n0(const cplx *a1, const cplx *a2, int nmu, long nLM, int nr,
                      const double *FLLs, const double *w, const int *Ls, const int *Ms, int Lmin,
                      int Lmax, int n1);
void orc_convolve_F_n000(const cplx *PQ, long nLM, int nr, const double *FLL, const double *w,
                         const int *Ls, const int *L_indices, cplx *out);
void orc_outer_product(const cplx *Qa, const cplx *Qb, int nta, int ntb, long nin, int sym,
                       double *fish);
void orc_integrate_pq(const double *flXs, int npol, int nr, const cplx *P, long nlm,
                      const int *ls, const double *w, cplx *out);

static unsigned long long rng_state = 12345;
static double urand(void) /* splitmix64 -> [-1, 1) */
{
    unsigned long long z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-52 - 1.0;
}

static int failures = 0;
#define CHECK(cond, ...)                                                                           \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            fprintf(stderr, __VA_ARGS__);                                                          \
            fputc('\n', stderr);                                                                   \
            failures++;                                                                            \
        }                                                                                          \
    } while (0)

static double *dvec(long n)
{
    double *v = malloc(sizeof(double) * (n > 0 ? n : 1));
    for (long i = 0; i < n; i++) v[i] = urand();
    return v;
}
static cplx *cvec(long n)
{
    cplx *v = malloc(sizeof(cplx) * (n > 0 ? n : 1));
    for (long i = 0; i < n; i++) v[i] = urand() + I * urand();
    return v;
}
static int dfinite(const double *v, long n)
{
    for (long i = 0; i < n; i++)
        if (!isfinite(v[i])) return 0;
    return 1;
}
static int cfinite(const cplx *v, long n)
{
    for (long i = 0; i < n; i++)
        if (!isfinite(creal(v[i])) || !isfinite(cimag(v[i]))) return 0;
    return 1;
}
static long nlm_of(int lmax) { return (long)(lmax + 1) * (lmax + 2) / 2; }
/* l of each (l, m) entry in the m-major order the SHTs use */
static int *ls_of(int lmax)
{
    int *ls = malloc(sizeof(int) * nlm_of(lmax));
    long i = 0;
    for (int m = 0; m <= lmax; m++)
        for (int l = m; l <= lmax; l++) ls[i++] = l;
    return ls;
}

static void check_sht(int nside, int lmax, int spin, int nmaps)
{
    int ncomp = spin ? 2 : 1;
    long npix = 12L * nside * nside, nlm = nlm_of(lmax);
    cplx *alm = cvec((long)nmaps * ncomp * nlm);
    for (int k = 0; k < nmaps * ncomp; k++) /* m = 0 entries are real */
        for (int l = 0; l <= lmax; l++) alm[k * nlm + l] = creal(alm[k * nlm + l]);
    double *map = malloc(sizeof(double) * nmaps * ncomp * npix);
    cplx *back = malloc(sizeof(cplx) * nmaps * ncomp * nlm);
    int rc = orc_sht_synthesis(nside, lmax, lmax, spin, nmaps, alm, map, 1);
    CHECK(rc == 0 && dfinite(map, (long)nmaps * ncomp * npix), "synthesis nside=%d lmax=%d spin=%d",
          nside, lmax, spin);
    rc = orc_sht_adjoint(nside, lmax, lmax, spin, nmaps, map, back, 1);
    CHECK(rc == 0 && cfinite(back, (long)nmaps * ncomp * nlm), "adjoint nside=%d lmax=%d spin=%d",
          nside, lmax, spin);
    /* mmax below lmax: the entry points then read and write only the m <= mmax entries */
    if (lmax >= 2) {
        int mmax = lmax / 2;
        long nlm_m = 0;
        for (int m = 0; m <= mmax; m++) nlm_m += lmax + 1 - m;
        cplx *a2 = cvec((long)nmaps * ncomp * nlm_m);
        cplx *b2 = malloc(sizeof(cplx) * nmaps * ncomp * nlm_m);
        for (int k = 0; k < nmaps * ncomp; k++)
            for (int l = 0; l <= lmax; l++) a2[k * nlm_m + l] = creal(a2[k * nlm_m + l]);
        rc = orc_sht_synthesis(nside, lmax, mmax, spin, nmaps, a2, map, 1);
        CHECK(rc == 0 && dfinite(map, (long)nmaps * ncomp * npix), "synthesis mmax=%d", mmax);
        rc = orc_sht_adjoint(nside, lmax, mmax, spin, nmaps, map, b2, 1);
        CHECK(rc == 0 && cfinite(b2, (long)nmaps * ncomp * nlm_m), "adjoint mmax=%d", mmax);
        free(a2);
        free(b2);
    }
    CHECK(orc_sht_synthesis(0, lmax, lmax, spin, nmaps, alm, map, 1) != 0, "nside 0 accepted");
    CHECK(orc_sht_synthesis(nside, lmax, lmax + 1, spin, nmaps, alm, map, 1) != 0,
          "mmax > lmax accepted");
    free(alm);
    free(map);
    free(back);
}

static void check_geometry(int nside)
{
    long total = 0;
    for (int i = 1; i <= 4 * nside - 1; i++) { /* rings are numbered from 1 */
        double z, sth, phi0;
        int nphi;
        long start;
        orc_ring_info(nside, i, &z, &sth, &phi0, &nphi, &start);
        CHECK(start == total && nphi > 0 && fabs(z) < 1 && sth > 0, "ring %d of nside %d", i, nside);
        total += nphi;
    }
    CHECK(total == 12L * nside * nside, "rings of nside %d cover %ld pixels", nside, total);
    for (int ss = -2; ss <= 2; ss += 2)
        for (int m = 0; m <= 6; m++) {
            double *out = malloc(sizeof(double) * 9);
            orc_spin_legendre(8, m, ss, 0.7, out);
            CHECK(dfinite(out, 9), "spin_legendre m=%d s=%d", m, ss);
            free(out);
        }
}

static void check_kernels(int lmax, int nr, int npol)
{
    long nlm = nlm_of(lmax), npix = 48;
    int *ls = ls_of(lmax);
    double *w = dvec(nr);
    w[nr / 2] = 0.0; /* the tau sums skip nodes of zero weight */
    int nslot = 3;
    double *flX = dvec((long)(lmax + 1) * nslot * nr);
    cplx *h = cvec((long)npol * nlm);
    cplx *o1 = malloc(sizeof(cplx) * nr * nlm);
    orc_apply_fl_weights(flX, nslot, nr, h, npol, nlm, ls, 0.5, o1);
    CHECK(cfinite(o1, (long)nr * nlm), "apply_fl_weights");

    double *P = dvec(nr * npix), *Q = dvec(nr * npix), *P2 = dvec(nr * npix), *Q2 = dvec(nr * npix);
    CHECK(isfinite(orc_gnl_loc_sum(w, P, Q, nr, npix)), "gnl_loc_sum");
    CHECK(isfinite(orc_gnl_loc_disc_sum(w, P, P2, Q, Q2, nr, npix)), "gnl_loc_disc_sum");

    double *flXs = dvec((long)(lmax + 1) * npol * nr);
    cplx *lm = cvec((long)nr * nlm);
    cplx *o2 = malloc(sizeof(cplx) * npol * nlm);
    orc_radial_sum(lm, nr, nlm, w, flXs, npol, ls, o2);
    CHECK(cfinite(o2, (long)npol * nlm), "radial_sum");
    cplx *Pi = cvec(nlm * nr);
    orc_integrate_pq(flXs, npol, nr, Pi, nlm, ls, w, o2);
    CHECK(cfinite(o2, (long)npol * nlm), "integrate_pq");

    /* tau sums over (L, M) entries with L up to Lmax: F_L tables of nLmax + 1 rows */
    int Lmax = lmax / 2, nmu = 3;
    long nLM = nlm_of(Lmax);
    int *Ls = ls_of(Lmax), *Ms = malloc(sizeof(int) * nLM);
    for (int m = 0, i = 0; m <= Lmax; m++)
        for (int L = m; L <= Lmax; L++) Ms[i++] = m;
    double *FLL = dvec((long)(Lmax + 1) * nr * nr);
    cplx *a1 = cvec((long)nmu * nLM * nr), *a2 = cvec((long)nmu * nLM * nr);
    CHECK(isfinite(orc_tau_sum_n0(a1, a2, nmu, nLM, nr, FLL, w, Ls, Ms, 1, Lmax, 0)), "tau_sum_n0");
    CHECK(orc_tau_sum_n0(a1, a2, nmu, nLM, nr, FLL, w, Ls, Ms, Lmax + 1, Lmax, 2) == 0.0,
          "tau_sum_n0 with an empty L range");
    int *Li = malloc(sizeof(int) * nLM);
    long nout = 0;
    for (long i = 0; i < nLM; i++) Li[i] = (i % 3 == 1) ? -1 : (int)nout++;
    cplx *PQ = cvec(nLM * nr), *o3 = malloc(sizeof(cplx) * nout * nr);
    orc_convolve_F_n000(PQ, nLM, nr, FLL, w, Ls, Li, o3);
    CHECK(cfinite(o3, nout * nr), "convolve_F_n000");

    int nta = 3, ntb = 2;
    long nin = 37;
    cplx *Qa = cvec(4L * nta * nin), *Qb = cvec(4L * ntb * nin);
    double *fish = malloc(sizeof(double) * nta * ntb), *fs = malloc(sizeof(double) * nta * nta);
    orc_outer_product(Qa, Qb, nta, ntb, nin, 0, fish);
    orc_outer_product(Qa, Qa, nta, nta, nin, 1, fs);
    CHECK(dfinite(fish, nta * ntb) && dfinite(fs, nta * nta), "outer_product");
    for (int i = 0; i < nta; i++)
        for (int j = 0; j < nta; j++)
            CHECK(fs[i * nta + j] == fs[j * nta + i], "outer_product sym (%d, %d)", i, j);

    free(ls); free(w); free(flX); free(h); free(o1); free(P); free(Q); free(P2); free(Q2);
    free(flXs); free(lm); free(o2); free(Pi); free(Ls); free(Ms); free(FLL); free(a1); free(a2);
    free(Li); free(PQ); free(o3); free(Qa); free(Qb); free(fish); free(fs);
}

int main(void)
{
    static const int sht_cases[][4] = {/* nside, lmax, spin, nmaps */
                                       {1, 2, 0, 1},  {1, 4, 2, 1},  {2, 5, 0, 2},  {2, 9, 2, 2},
                                       {4, 11, 0, 1}, {4, 16, 2, 3}, {8, 23, 0, 2}, {8, 15, 2, 1},
                                       {3, 8, 0, 1}};
    for (size_t k = 0; k < sizeof(sht_cases) / sizeof(sht_cases[0]); k++)
        check_sht(sht_cases[k][0], sht_cases[k][1], sht_cases[k][2], sht_cases[k][3]);
    for (int nside = 1; nside <= 8; nside *= 2) check_geometry(nside);
    check_geometry(3);
    check_kernels(6, 4, 1);
    check_kernels(9, 5, 2);
    check_kernels(2, 1, 2);
    if (failures) {
        fprintf(stderr, "asan_check: %d failures\n", failures);
        return 1;
    }
    puts("asan_check ok");
    return 0;
}
