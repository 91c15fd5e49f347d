/*
 * FP64 CPU oracle for the radial-node pixel kernels of the gNL-loc / tauNL-loc
 * estimators.  TEST INFRASTRUCTURE ONLY: the checker the HIP kernels are
 * compared against and the timed CPU baseline; never on the product path.
 *
 * Each function restates one Cython loop nest of the reference (cited per
 * function).  Accumulation is FP64 in a fixed order (the reference's OpenMP
 * reductions are order-nondeterministic and built with -ffast-math,
 * polyspec/cython/setup.py:14,33, so they are not bit-reproducible anyway).
 * Complex arrays are interleaved (re, im) double pairs.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* tauNL_utils.apply_fl_weights, polyspec/cython/tauNL_utils.pyx:80-98:
 * out[r, i] = w * sum_{X < npol} flX[l_i, X, r] h[X, i]. flX (lmax+1, nslot, nr). */
void orc_apply_fl_weights(const double *flX, int nslot, int nr, const cplx *h, int npol,
                          long nl, const int *ls, double weight, cplx *out)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < nr; r++)
        for (long i = 0; i < nl; i++) {
            long l = ls[i];
            cplx acc = flX[(l * nslot + 0) * nr + r] * h[i];
            if (npol > 1) acc += flX[(l * nslot + 1) * nr + r] * h[nl + i];
            out[(long)r * nl + i] = acc * weight;
        }
}

/* tauNL_utils.gnl_loc_sum, tauNL_utils.pyx:260-270: sum_r w_r sum_p P^3 Q. */
double orc_gnl_loc_sum(const double *w, const double *P, const double *Q, int nr, long npix)
{
    double out = 0.0;
    for (int r = 0; r < nr; r++) {
        double t = 0.0;
        const double *p = P + (long)r * npix, *q = Q + (long)r * npix;
#pragma omp parallel for reduction(+ : t) schedule(static)
        for (long i = 0; i < npix; i++) t += p[i] * p[i] * p[i] * q[i];
        out += w[r] * t;
    }
    return out;
}

/* tauNL_utils.gnl_loc_disc_sum, tauNL_utils.pyx:275-285:
 * sum_r w_r sum_p (P1^2 P2 Q2 + P2^2 P1 Q1). */
double orc_gnl_loc_disc_sum(const double *w, const double *P1, const double *P2,
                            const double *Q1, const double *Q2, int nr, long npix)
{
    double out = 0.0;
    for (int r = 0; r < nr; r++) {
        double t = 0.0;
        long o = (long)r * npix;
#pragma omp parallel for reduction(+ : t) schedule(static)
        for (long i = 0; i < npix; i++)
            t += P1[o + i] * P1[o + i] * P2[o + i] * Q2[o + i] +
                 P2[o + i] * P2[o + i] * P1[o + i] * Q1[o + i];
        out += w[r] * t;
    }
    return out;
}

/* tauNL_utils.radial_sum, tauNL_utils.pyx:220-235:
 * out[X, i] = sum_r lm[r, i] flXs[l_i, X, r] w_r, for X < npol (npol = flXs.shape[1]). */
void orc_radial_sum(const cplx *lm, int nr, long nlm, const double *w, const double *flXs,
                    int npol, const int *ls, cplx *out)
{
    for (int X = 0; X < npol; X++) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nlm; i++) {
            long l = ls[i];
            cplx acc = 0;
            for (int r = 0; r < nr; r++) acc += lm[(long)r * nlm + i] * flXs[(l * npol + X) * nr + r] * w[r];
            out[(long)X * nlm + i] = acc;
        }
    }
}

/* tauNL_utils.tau_sum_n0, tauNL_utils.pyx:604-646.  arr (nmu, nLM, nr) complex,
 * FLLs (nLmax+1, nr, nr), Ls/Ms (nLM).  Returns
 * sum_{j,k: w_j w_k != 0} sum_{i: Lmin<=L_i<=Lmax} (1+[M_i>0]) F_{L_i}(j,k)/sqrt(4pi)
 *   w_j w_k / sqrt(2 n1 + 1) Re sum_mu arr1[mu,i,j] conj(arr2[mu,i,k]). */
double orc_tau_sum_n0(const cplx *a1, const cplx *a2, int nmu, long nLM, int nr,
                      const double *FLLs, const double *w, const int *Ls, const int *Ms,
                      int Lmin, int Lmax, int n1)
{
    double out = 0.0;
#pragma omp parallel for reduction(+ : out) schedule(static)
    for (int idx = 0; idx < nr * nr; idx++) {
        int j = idx / nr, k = idx % nr;
        if (w[j] == 0.0 || w[k] == 0.0) continue;
        double part = 0.0;
        for (long i = 0; i < nLM; i++) {
            int L = Ls[i];
            if (L < Lmin || L > Lmax) continue;
            double mw = (Ms[i] == 0) ? 1.0 : 2.0;
            double t = mw * FLLs[((long)L * nr + j) * nr + k] / sqrt(4.0 * M_PI) * w[j] * w[k] /
                       sqrt(2.0 * n1 + 1.0);
            for (int mu = 0; mu < nmu; mu++)
                part += t * creal(a1[((long)mu * nLM + i) * nr + j] * conj(a2[((long)mu * nLM + i) * nr + k]));
        }
        out += part;
    }
    return out;
}

/* tauNL_utils.convolve_F_n000, tauNL_utils.pyx:966-992 (with inds = arange(nr)):
 * out[L_indices[ip], j] = sum_k F[Ls[ip], j, k] w_k PQ[ip, k] / sqrt(4 pi)
 * for ip with L_indices[ip] != -1. FLL (nLmax+1, nr, nr) is FLLs[nmax_F]. */
void orc_convolve_F_n000(const cplx *PQ, long nLM, int nr, const double *FLL,
                         const double *w, const int *Ls, const int *L_indices, cplx *out)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nr; j++)
        for (long ip = 0; ip < nLM; ip++) {
            int i = L_indices[ip];
            if (i == -1) continue;
            int Lp = Ls[ip];
            cplx acc = 0;
            for (int k = 0; k < nr; k++) {
                if (w[k] == 0.0) continue;
                acc += FLL[((long)Lp * nr + j) * nr + k] * w[k] * PQ[ip * nr + k] / sqrt(4.0 * M_PI);
            }
            out[(long)i * nr + j] = acc;
        }
}

/* ideal_fisher.outer_product, polyspec/cython/ideal_fisher.pyx:423-459.
 * Q4 (4, n_t, n_in) complex -> fish (n_ta, n_tb). */
void orc_outer_product(const cplx *Qa, const cplx *Qb, int nta, int ntb, long nin, int sym,
                       double *fish)
{
    memset(fish, 0, sizeof(double) * nta * ntb);
    for (int ia = 0; ia < nta; ia++)
        for (int ib = 0; ib < ntb; ib++) {
            if (sym && ia > ib) continue;
            double acc = 0.0;
#define QA(c, j) Qa[((long)(c) * nta + ia) * nin + (j)]
#define QB(c, j) Qb[((long)(c) * ntb + ib) * nin + (j)]
#pragma omp parallel for reduction(+ : acc) schedule(static)
            for (long j = 0; j < nin; j++) {
                cplx t = conj(QA(0, j)) * QB(0, j) + conj(QA(1, j)) * QB(1, j) +
                         9.0 * (conj(QA(2, j)) * QB(2, j) + conj(QA(3, j)) * QB(3, j)) -
                         3.0 * (conj(QA(0, j)) * QB(3, j) + conj(QA(1, j)) * QB(2, j) +
                                conj(QA(3, j)) * QB(0, j) + conj(QA(2, j)) * QB(1, j));
                acc += creal(t);
            }
#undef QA
#undef QB
            fish[ia * ntb + ib] += acc / 24.0 / 48.0;
            if (sym && ia != ib) fish[ib * ntb + ia] += acc / 24.0 / 48.0;
        }
}

/* ideal_fisher.integrate_pq, ideal_fisher.pyx:361-375 (inds = arange(nr)):
 * out[X, i] += sum_r 0.5 w_r flXs[l_i, X, r] P[i, r]. */
void orc_integrate_pq(const double *flXs, int npol, int nr, const cplx *P, long nlm,
                      const int *ls, const double *w, cplx *out)
{
    for (int X = 0; X < npol; X++) {
#pragma omp parallel for schedule(static)
        for (long i = 0; i < nlm; i++) {
            long l = ls[i];
            cplx acc = 0;
            for (int r = 0; r < nr; r++) acc += 0.5 * w[r] * flXs[(l * npol + X) * nr + r] * P[i * nr + r];
            out[(long)X * nlm + i] += acc;
        }
    }
}
