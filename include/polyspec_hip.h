/*
 * polyspec_hip.h -- C ABI of the MI355X (gfx950) backend for PolySpec's hot path.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on error
 *     (-1 invalid argument, -2 HIP runtime error); ps_last_error() returns
 *     the message of the last failure on the calling thread.  No call aborts.
 *   - All array arguments are device pointers owned by the caller; nothing is
 *     allocated by the compute entry points.  Complex arrays are interleaved
 *     (re, im); "precision" 32 means float / complex64, 64 means double /
 *     complex128, fixed per plan.
 *   - `stream` is a hipStream_t (NULL = default stream).  Entry points only
 *     enqueue work; the *_sum functions write their FP64 result to a device
 *     scalar so that no call synchronises.
 *   - Reductions are deterministic (fixed order) and accumulate in FP64.
 *
 * Reference interfaces replaced are cited per function (paths relative to the
 * PolySpec repository root).
 */
#ifndef POLYSPEC_HIP_H
#define POLYSPEC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PS_ABI_VERSION 1

const char *ps_last_error(void);
int ps_version(void);

/* Select the HIP device for the calling thread. */
int ps_init(int device);

/* ------------------------------------------------------------------ SHTs
 * Replace ducc0.sht.synthesis / adjoint_synthesis with
 * **ducc0.healpix.Healpix_Base(nside,"RING").sht_info() geometry, as called by
 * PolySpec.to_map / to_lm / to_map_vec / to_lm_vec / to_map_spin / to_lm_spin
 * (polyspec/base.py:70-74, 147-151, 172-176, 193, 206, 225, 230-231, 253, 282).
 *
 * A plan fixes (nside, lmax, mmax, spin, precision).  spin 0: one complex alm
 * row per map; spin s > 0: two rows (gradient G, curl C) per map producing two
 * real maps (R, I) with R + iI = -sum (G + iC)_lm sY_lm (HEALPix convention;
 * for s = 2 that is Q + iU = -sum (a_E + i a_B) 2Y_lm).
 * alm layout: [nmaps][ncomp][nlm], healpy m-major index m(2 lmax+1-m)/2 + l.
 * map layout: [nmaps][ncomp][12 nside^2] RING order.
 */
typedef struct ps_sht_plan ps_sht_plan;

int ps_sht_plan_create(int nside, int lmax, int mmax, int spin, int precision, ps_sht_plan **out);
int ps_sht_plan_destroy(ps_sht_plan *plan);
int ps_sht_plan_info(const ps_sht_plan *plan, long long *npix, long long *nlm, int *ncomp);

/* Bytes of device workspace needed for nmaps maps (lmajor_input = 1 for
 * ps_sht_synthesis_lmajor and ps_sht_adjoint). */
long long ps_sht_workspace_bytes(const ps_sht_plan *plan, int nmaps, int lmajor_input);

/* Column multiple required of `ld` by ps_sht_synthesis_lmajor. */
int ps_sht_lmajor_multiple(const ps_sht_plan *plan);

/* alm -> map (ducc0.sht.synthesis; base.py:172-176, 225, 230-231, 282). */
int ps_sht_synthesis(const ps_sht_plan *plan, int nmaps, const void *alm, void *map, void *work,
                     long long work_bytes, void *stream);

/* Spin-s synthesis of (G, C = 0) from G alone: G (nmaps, nlm) complex m-major -> maps
 * (nmaps, 2, npix), i.e. ps_sht_synthesis of (G, 0) without the C half of the Legendre work.
 * It replaces compute_spin_transform_map's synthesis of the bin-filtered h (bspec_bin.py:216-228,
 * 335-337 through base.py:260-288), whose curl part is zero.  FP32 plans of spin >= 1; work:
 * ps_sht_workspace_bytes(plan, nmaps, 0) bytes.  The Legendre shape set for nmaps by
 * ps_sht_set_shapes (0: one tile of 16 maps per block; the wide shapes: 2-4 tiles sharing the
 * recursions) applies; by default the wide form from 17 maps on. */
int ps_sht_synthesis_grad(const ps_sht_plan *plan, int nmaps, const void *G, void *map, void *work,
                          long long work_bytes, void *stream);

/* ps_sht_synthesis with alm given l-major: alm_t[idx(l,m) * ld + col], col = map*ncomp + comp,
 * ld >= nmaps*ncomp a multiple of ps_sht_lmajor_multiple(); padding columns must be zero.
 * FP32 spin-0 batches on the matrix-core shape (Nside a power of two) fork the Bluestein-ring
 * FFTs onto the plan's own side stream and join it back into `stream` before returning (event
 * waits only: the call stays asynchronous and capturable into a HIP graph). */
int ps_sht_synthesis_lmajor(const ps_sht_plan *plan, int nmaps, const void *alm_t, int ld, void *map,
                            void *work, long long work_bytes, void *stream);

/* Transpose ncol m-major alm rows [col][nlm] into the l-major layout alm_t[idx(l,m) * ld + col]
 * (columns ncol..ld-1 zeroed), in the plan's precision.  No reference counterpart: the layout
 * change that lets batched consumers (the l-major synthesis input above, ps_tau_sum_lmajor
 * below) read whole rows of columns; it replaces the transposes of
 * `self.base.to_lm_vec(...).T[self.nLminfilt]` (tspec_template.py:1294-1303). */
int ps_sht_alm_to_lmajor(const ps_sht_plan *plan, int ncol, const void *alm, int ld, void *alm_t, void *stream);

/* The two stages of ps_sht_synthesis_lmajor, for callers that schedule or time them
 * separately: the Legendre stage writes ring phases F_m(ring) exp(i m phi0(ring)) in
 * tiles of 16 rings, [ring / 16][mmax+1][ring % 16][nmaps*ncomp] complex
 * (ps_sht_phase_bytes bytes); the ring stage turns them into maps by per-ring FFTs. */
int ps_sht_phase_bytes(const ps_sht_plan *plan, int nmaps, long long *bytes);
int ps_sht_legendre_synthesis_lmajor(const ps_sht_plan *plan, int nmaps, const void *alm_t, int ld, void *phase,
                                     void *stream);
int ps_sht_ring_synthesis(const ps_sht_plan *plan, int nmaps, const void *phase, void *map, void *stream);

/* map -> alm: scale * sum_p map_p sY*_lm(p) (ducc0.sht.adjoint_synthesis times
 * `scale`; PolySpec passes scale = pi/(3 nside^2) = 4 pi/Npix, base.py:147-151, 193, 206, 253). */
int ps_sht_adjoint(const ps_sht_plan *plan, int nmaps, const void *map, void *alm, double scale, void *work,
                   long long work_bytes, void *stream);

/* Launch-shape tuning (the plan's stages: 0 Legendre synthesis, 1 ring FFTs of power-of-two
 * rings, 2 ring FFTs of Bluestein rings, 3 Legendre adjoint).  No reference counterpart: ducc0
 * picks its own blocking inside ducc0.sht (base.py:172-176, 193 call it), so these tune what
 * the transforms above launch for one batch width nmaps*ncomp; results of every shape agree
 * with the built-in choice to rounding.  Shapes are stored in the plan: set or tune them
 * before, not during, transforms on other threads.
 *
 * Ring-stage ids: FFT buffers (of two real columns each) per block, 1, 2 or 4, plus 8 for the
 * self-sorting radix-16 transforms (9, 10) instead of the radix-2^2 ones (the built-in choice).
 * candidates: the shape ids of `stage` (count returned; at most max_ids written).
 * get_shapes: shapes[4] in use for nmaps (tuned or set: returns 1; built-in rule: 0).
 * set_shapes: shapes[4], -1 keeps the built-in rule for a stage.
 * autotune: times every candidate (one untimed launch, then `reps` launches between HIP events)
 * on synthetic data in `work` (ps_sht_autotune_work_bytes bytes) and keeps the fastest when it
 * beats the built-in choice by 3%; synchronises `stream` (not capturable).  times (optional,
 * 8 doubles): per stage, ms per launch of the kept and of the built-in shape. */
int ps_sht_shape_candidates(const ps_sht_plan *plan, int nmaps, int stage, int *ids, int max_ids);
int ps_sht_get_shapes(const ps_sht_plan *plan, int nmaps, int *shapes);
int ps_sht_set_shapes(ps_sht_plan *plan, int nmaps, const int *shapes);
long long ps_sht_autotune_work_bytes(const ps_sht_plan *plan, int nmaps);
int ps_sht_autotune(ps_sht_plan *plan, int nmaps, void *work, long long work_bytes, int reps, double *times,
                    void *stream);


/* ------------------------------------------------------- radial kernels
 * n = 0 loop nests of polyspec/cython/tauNL_utils.pyx and ideal_fisher.pyx.
 * `precision` (32 | 64) selects float/complex64 or double/complex128 data;
 * weights w and F tables are always double.  Reductions write one double to
 * the device pointer `out` (added to it when accumulate != 0) and need a
 * device scratch of ps_reduce_scratch_bytes() bytes.
 * Radial-profile tables f are laid out (lmax+1, nslot, nr) as plLXs[nmax] /
 * qlXs in the reference (tspec_template.py:623-630). */
long long ps_reduce_scratch_bytes(void);

/* tauNL_utils.apply_fl_weights (tauNL_utils.pyx:80-98) fused with the zero fill
 * below lmin of TSpecTemplate._compute_weighted_maps (tspec_template.py:799-808),
 * emitting the l-major synthesis input of ps_sht_synthesis_lmajor:
 * out_t[idx(l,m)*ld + col_offset + r] = weight * sum_{X<npol} f[l,X,r] h[X,idx(l,m)].
 * h: (npol, nlm(lmax)) complex. */
int ps_weight_alm_lmajor(int precision, const void *f, int nslot, int nr, const void *h, int npol, int lmax,
                         int lmin, double weight, int ld, int col_offset, void *out_t, void *stream);
/* The two weightings of one filtered field in one launch (TSpecTemplate._compute_weighted_maps,
 * tspec_template.py:799-855, for the 'p00' and 'q' maps of _apply_all_filters): columns
 * col_offset + r as ps_weight_alm_lmajor with (fP, wP), columns col_offset + nr + r with
 * (fQ, wQ); fP and fQ share nslot and nr (rows l <= lmax are read).  Each row of out_t is written by
 * one launch, not half by each of two. */
int ps_weight_alm_pq_lmajor(int precision, const void *fP, const void *fQ, int nslot, int nr, const void *h,
                            int npol, int lmax, int lmin, double wP, double wQ, int ld, int col_offset,
                            void *out_t, void *stream);
/* The same weightings with the columns interleaved: P_r at col_offset + 2r, Q_r at
 * col_offset + 2r + 1 (the layout ps_sht_ring_gnl reads). */
int ps_weight_alm_pq_interleaved_lmajor(int precision, const void *fP, const void *fQ, int nslot, int nr,
                                        const void *h, int npol, int lmax, int lmin, double wP, double wQ, int ld,
                                        int col_offset, void *out_t, void *stream);

/* c4 * tauNL_utils.gnl_loc_sum (tauNL_utils.pyx:260-270) plus, when U and V are
 * given, c2 * sum_r w_r sum_p (P^2 U + P Q V): with U = sum_s P_s Q_s and
 * V = sum_s P_s^2 this equals c2 * sum_s gnl_loc_disc_sum(w, P, P_s, Q, Q_s)
 * (tauNL_utils.pyx:275-285, summed at tspec_template.py:1565-1575).
 * P, Q in the working precision; U, V are FP64 (double*) in both precisions. */
int ps_gnl_loc_fused_sum(int precision, const void *P, const void *Q, const void *U, const void *V,
                         const double *w, int nr, long long npix, double c4, double c2, double *scratch,
                         double *out, int accumulate, void *stream);

/* The same sum without the maps: the ring stage of the numerator's synthesis fused with it.
 * phase holds the Legendre-stage output (ps_sht_legendre_synthesis_lmajor) of 2*nr columns
 * interleaved as (a_r, b_r) = (s_r P_r, Q_r / s_r) (ps_weight_alm_pq_interleaved_lmajor with the
 * tables scaled by s_r and 1 / s_r); inv_s2[r] = 1 / s_r^2 (NULL: s_r = 1).  The two columns of a
 * pair share one complex FFT and so its absolute rounding error: s_r should make their magnitudes
 * alike.  *out (added when accumulate) = c4 * gnl_loc_sum(w, P, Q)
 * [+ c2 * sum_r w_r sum_p (P^2 U + P Q V)] over the maps that ps_sht_ring_synthesis would produce
 * (tspec_template.py:1391 and 1565-1575 through tauNL_utils.pyx:260-285), each ring block adding
 * its pixels in FP64, fixed order.  FP32 spin-0 plans; work: ps_sht_ring_gnl_work_bytes bytes,
 * which is -1 when no staged ring kernel fits the plan (then use ps_sht_ring_synthesis and
 * ps_gnl_loc_fused_sum). */
long long ps_sht_ring_gnl_work_bytes(const ps_sht_plan *plan, int nr);
int ps_sht_ring_gnl(const ps_sht_plan *plan, int nr, const void *phase, const double *U, const double *V,
                    const double *w, const double *inv_s2, double c4, double c2, void *work, long long work_bytes,
                    double *out, int accumulate, void *stream);

/* tauNL_utils.gnl_loc_disc_sum (tauNL_utils.pyx:275-285). */
int ps_gnl_loc_disc_sum(int precision, const void *P1, const void *P2, const void *Q1, const void *Q2,
                        const double *w, int nr, long long npix, double *scratch, double *out, int accumulate,
                        void *stream);

/* U += sum_s P_s Q_s, V += sum_s P_s^2 over nsim stacked (P_s, Q_s) of n values
 * (the Monte-Carlo sums that collapse the per-simulation loop of
 * tspec_template.py:1542-1580).  P, Q in the working precision; the sums U, V are
 * FP64 (double*) in both precisions. */
int ps_mc_accumulate(int precision, const void *P, const void *Q, int nsim, long long n, void *U, void *V,
                     void *stream);

/* Elementwise map products times scale: mode 0 a*b (ideal_fisher.multiplyRR,
 * ideal_fisher.pyx:100-108), 1 a*b + c*d (multiplyRR_sym, 112-120),
 * 2 a*b*c (tauNL_utils.multiply, tauNL_utils.pyx:103-111),
 * 3 a*(2 b c + a d) (tauNL_utils.multiply_asym, 116-124). */
int ps_pixel_products(int precision, int mode, const void *a, const void *b, const void *c, const void *d,
                      long long n, double scale, void *out, void *stream);
/* The same over a batch: a and out are (nbatch, n), b, c, d (n) shared by every
 * stack (the weight maps of the batched CG below); mode 4 is mode 1 with c batched like a,
 * out[s] = a[s]*b + c[s]*d (multiplyRR_sym of each simulation's (Q_s, P_s) with the data's
 * (P, Q), tspec_template.py:1294-1303 as called at 1665-1687). */
int ps_pixel_products_batch(int precision, int mode, int nbatch, const void *a, const void *b, const void *c,
                            const void *d, long long n, double scale, void *out, void *stream);

/* tauNL_utils.radial_sum (tauNL_utils.pyx:220-235) over the full m-major set,
 * zero below lmin: out[X, lm] (+)= scale sum_r lm_map[r, lm] f[l,X,r] w_r, X < npol. */
int ps_radial_sum(int precision, const void *lm_map, int nr, int lmax, int lmin, const void *f, int nslot, int npol,
                  const double *w, double scale, int accumulate, void *out, void *stream);

/* The per-node terms of that sum, ideal_fisher.integrate_pq_deriv (ideal_fisher.pyx:395-407) as
 * used for the tauNL weight derivative (tspec_template.py:2385, 2436):
 * out[r, X, lm] (+)= scale w_r f[l,X,r] lm_map[r, lm] for l >= lmin, else 0; out (nr, npol, nlm),
 * nr <= 65535. */
int ps_radial_scatter(int precision, const void *lm_map, int nr, int lmax, int lmin, const void *f, int nslot,
                      int npol, const double *w, double scale, int accumulate, void *out, void *stream);

/* tauNL_utils.tau_sum_n0 (tauNL_utils.pyx:604-646) as per-L Gram matrices.
 * A, B: (nmu, nr, nlm(Lmaxs)) complex m-major over L = 0..Lmaxs (the transpose
 * of the reference's (nmu, nLM, nr) layout); FLL (Lmaxs+1, nr, nr) = FLLs[nmax_F]. */
int ps_tau_sum_n0(int precision, const void *A, const void *B, int nmu, int nr, int Lmaxs, int Lmin, int Lmax,
                  const double *FLL, const double *w, int n1, double *scratch, double *out, int accumulate,
                  void *stream);

/* tauNL_utils.tau_sum_n0 (tauNL_utils.pyx:604-646), summed over a batch of nmu (A, B) pairs, with
 * A and B l-major: At[idx(L,M) * ld + mu * nr + j] over L = 0..Lmaxs (ld >= nmu * nr, nr <= 64):
 *   out (+)= sum_mu tau_sum_n0(A_mu, B_mu)
 * = pref sum_L sum_{j,k} F_L(j,k) w_j w_k sum_mu sum_M (1+[M>0]) Re(A_mu,j(L,M) conj(B_mu,k(L,M))),
 * pref = 1/(sqrt(4 pi) sqrt(2 n1 + 1)).  Per-L Gram matrices on FP64 matrix cores (FP32 data is
 * widened), F_L contraction fused; replaces the per-simulation tau_sum_n0 calls of the
 * Monte-Carlo terms (tspec_template.py:1665-1687).  At == Bt is the symmetric case tau(A, A). */
int ps_tau_sum_lmajor(int precision, const void *At, const void *Bt, long long ld, int nmu, int nr, int Lmaxs,
                      int Lmin, int Lmax, const double *FLL, const double *w, int n1, double *scratch, double *out,
                      int accumulate, void *stream);
/* The same sum over shifted maps, sum_mu tau_sum_n0(A_mu - s_A M, B_mu - s_B M) with s_A = shiftA,
 * s_B = shiftB in {0, 1} and Mt (nlm(Lmaxs), nr) l-major in the data precision, the shift applied
 * as the values are widened to FP64.  No reference counterpart: the batched tauNL-loc numerator
 * subtracts the Monte-Carlo mean field M = mean_s A_ss from every self-product A map, which
 * leaves t4 + t2 + t0 (tspec_template.py:1420-1432, 1655-1687) unchanged in exact arithmetic but
 * turns its terms from ~1e3 sigma (masked sky) into ~sigma, so FP32 A maps suffice. */
int ps_tau_sum_lmajor_shift(int precision, const void *At, const void *Bt, long long ld, int nmu, int nr, int Lmaxs,
                            int Lmin, int Lmax, const double *FLL, const double *w, int n1, const void *Mt, int shiftA,
                            int shiftB, double *scratch, double *out, int accumulate, void *stream);

/* ps_tau_sum_lmajor without a scratch argument and without allocating: the per-L partial sums
 * go to a device array built into the library (8192 doubles per GPU, held until the process
 * ends), so Lmax - Lmin + 1 <= 8192 (larger ranges are rejected with an error).  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Calls that overlap on one GPU share that array and race. */
int ps_tau_sum_lmajor_noalloc(int precision, const void *At, const void *Bt, long long ld, int nmu, int nr,
                              int Lmaxs, int Lmin, int Lmax, const double *FLL, const double *w, int n1, double *out,
                              int accumulate, void *stream);

/* ideal_fisher.outer_product (ideal_fisher.pyx:423-459, sym = 0): Qa (4, nta, nin),
 * Qb (4, ntb, nin) complex -> fish (nta, ntb) double (added when accumulate != 0). */
int ps_outer_product(int precision, const void *Qa, const void *Qb, int nta, int ntb, long long nin, double *scratch,
                     double *fish, int accumulate, void *stream);

/* tauNL_utils.convolve_F_n000 (tauNL_utils.pyx:966-992, all radial nodes) as used by
 * TSpecTemplate._convolve_F (tspec_template.py:2155-2162), written l-major for the
 * following synthesis: out_t[idx(L,M)*ld + j] = sum_{k: w_k != 0} F_L(j,k) w_k PQ[k, idx(L,M)] / sqrt(4 pi)
 * for Lmin <= L <= Lmax, else 0.  PQ: (nr, nlm(Lmaxs)) complex; FLL (Lmaxs+1, nr, nr). */
int ps_convolve_F_lmajor(int precision, const void *PQ, int nr, int Lmaxs, int Lmin, int Lmax, const double *FLL,
                         const double *w, int ld, void *out_t, void *stream);

/* ------------------------------------------------- S^-1 filter vectors
 * Harmonic vectors are (nf, nlm(lmax)) complex (nf = 1 or 3: T or T,E,B). */

/* out[a] (+)= alpha * post[a,l] sum_b mat[a,b,l] pre[b,l] in[b] for every (l, m):
 * the per-l 3x3 products of weighting.py (_apply_Cinv 279-281, _apply_M harmonic
 * 304-305, x0 174, applySinv_ideal 200, 214) and PolySpec.applyAinv (base.py:444).
 * mat (nf,nf,lmax+1), pre/post (nf,lmax+1) double, any of them NULL = identity.
 * In-place (in == out) is allowed. Row strides in complex elements. */
int ps_lmat_apply(int precision, int nf, int lmax, const double *mat, const double *pre, const double *post,
                  const void *in, long long in_stride, void *out, long long out_stride, double alpha, int accumulate,
                  void *stream);

/* CGD.dot (weighting.py:255): Re sum x conj(y) (1 + [m > 0]) -> *out (device double).
 * scratch: ps_mdot_scratch_bytes() bytes. */
long long ps_mdot_scratch_bytes(void);
int ps_mdot(int precision, int nf, int lmax, const void *x, const void *y, double *scratch, double *out,
            void *stream);

/* CGD.step updates (weighting.py:329-339) with device scalars:
 * x += (rz/pAp) p, r -= (rz/pAp) Ap;   p = z + (num/den) p. n = complex elements. */
int ps_cg_update_xr(int precision, long long n, const double *rz, const double *pAp, const void *p, const void *Ap,
                    void *x, void *r, void *stream);
int ps_cg_update_p(int precision, long long n, const double *num, const double *den, const void *z, void *p,
                   void *stream);

/* Batched CG: nbatch independent systems solved in lock step (the reference solves one
 * map at a time, weighting.py:148-184 called per simulation; these run the same
 * iteration for every map of a batch).  Vectors are (nbatch, nf, nlm) contiguous;
 * scalars are per system (double[nbatch]); active[b] == 0 leaves system b unchanged
 * (converged).  mdot_batch: out[b] = dot of system b; scratch ps_mdot_batch_scratch_bytes. */
int ps_lmat_apply_batch(int precision, int nbatch, int nf, int lmax, const double *mat, const double *pre,
                        const double *post, const void *in, void *out, double alpha, int accumulate, void *stream);
long long ps_mdot_batch_scratch_bytes(int nbatch);
int ps_mdot_batch(int precision, int nbatch, int nf, int lmax, const void *x, const void *y, double *scratch,
                  double *out, void *stream);
int ps_cg_update_xr_batch(int precision, int nbatch, long long n, const double *rz, const double *pAp,
                          const int *active, const void *p, const void *Ap, void *x, void *r, void *stream);
int ps_cg_update_p_batch(int precision, int nbatch, long long n, const double *num, const double *den,
                         const int *active, const void *z, void *p, void *stream);

/* -------------------------------------------- binned bispectrum (BSpecBin)
 * Pixel reductions of the binned-bispectrum numerator (bspec_bin.py:339-403, the
 * _analyze3 / _analyze1 map products summed with mp.Pool):
 *   out[t] (+)= scale * sum_p [A_i1 B_i2 C_i3 + A_i2 B_i3 C_i1 + A_i3 B_i1 C_i2](p)
 * for triplets t = (i1, i2, i3) (int32, ntrip x 3) of map indices into the complex
 * stacks A, B, C (nmap x npix, nmap <= 128, FP32 or FP64 complex); out is complex FP64.
 * A == B (the cubic term, bspec_bin.py:339-364) runs a two-stack pass.
 * scratch: ps_bspec_scratch_bytes(ntrip) bytes. */
long long ps_bspec_scratch_bytes(int ntrip);
int ps_bspec_triple_sums(int precision, const void *A, const void *B, const void *C, int nmap, long long npix,
                         const int *triplets, int ntrip, double scale, void *scratch, long long scratch_bytes,
                         void *out, int accumulate, void *stream);
/* The linear term of one simulation (bspec_bin.py:366-403: the three _analyze1 calls with
 * (data, sim, sim), (sim, data, sim) and (sim, sim, data) H maps) in one pass:
 *   out[t] (+)= scale * sum_p sum_cyc [(P_a Q_b + Q_a P_b) N_c + Q_a Q_b M_c](p)
 * over the cyclic orders (a, b, c) of (i1, i2, i3), with P / M the data's (+1)H / (-2)H
 * stacks and Q / N the simulation's; equal to ps_bspec_triple_sums on (P, Q, N), (Q, P, N)
 * and (Q, Q, M) added.  Same layouts and scratch as ps_bspec_triple_sums. */
int ps_bspec_linear_sums(int precision, const void *P, const void *Q, const void *M, const void *N, int nmap,
                         long long npix, const int *triplets, int ntrip, double scale, void *scratch,
                         long long scratch_bytes, void *out, int accumulate, void *stream);

/* Every pixel reduction of one or several BSpecBin numerators in one launch: the cubic term
 * of each data set and the linear term of every simulation (bspec_bin.py:339-403, whose
 * mp.Pool map products and per-simulation loop this replaces), at any mix of resolutions.
 * A segment is one set of device stacks (nmap x npix complex, FP32 or FP64 as precision)
 * with its device triplets (ntrip x 3 int32) and mode:
 *   0 = ps_bspec_triple_sums on (X[0], X[1], X[2]); 1 = the same with A = B = X[0], C = X[1];
 *   2 = ps_bspec_linear_sums on (X[0], X[1], X[2], X[3]) = (P, Q, M, N).
 * Segments of one group (same ntrip) add into the group's out (ntrip complex FP64):
 *   out (+)= scale * sum over its segments, summed in segment order, bit-reproducible.
 * shape: launch-shape id in [0, ps_bspec_multi_shape_count(precision)) (pixel tile, triplets
 * per thread, pixel ranges per segment), -1 = built-in; ps_bspec_multi_autotune times every
 * candidate on the given segments (partials only; no group output is written) and returns the
 * fastest in *best (the built-in one unless another is 3% faster).  scratch:
 * ps_bspec_multi_scratch_bytes(...) bytes, enough for every shape.  The segment table is copied
 * from host memory on the stream (not for graph capture). */
typedef struct ps_bspec_segment {
    int mode;
    int nmap;
    long long npix;
    const void *X[4];
    const int *triplets;
    int ntrip;
    int group;
} ps_bspec_segment;
typedef struct ps_bspec_group {
    void *out;
    double scale;
    int accumulate;
} ps_bspec_group;
int ps_bspec_multi_shape_count(int precision);
long long ps_bspec_multi_scratch_bytes(int precision, int nseg, const ps_bspec_segment *segs, int ngroup);
int ps_bspec_sums_multi(int precision, int nseg, const ps_bspec_segment *segs, int ngroup,
                        const ps_bspec_group *groups, int shape, void *scratch, long long scratch_bytes,
                        void *stream);
int ps_bspec_multi_autotune(int precision, int nseg, const ps_bspec_segment *segs, int ngroup, void *scratch,
                            long long scratch_bytes, int reps, double *ms, int *best, void *stream);

/* -------------------------------------------- radial-node optimizer
 * Exact gNL-loc Fisher-matrix derivative (fisher_deriv_gNL_loc, ideal_fisher.pyx:750-787,
 * called from _compute_fisher_derivatives, tspec_template.py:3858-3870), all FP64:
 *   D(r, s) = pref w_r w_s sum_mu w_mu [S^PP^3 S^QQ + 3 S^PQ S^QP S^PP^2](mu; r, s),
 *   S^XY(mu; r, s) = sum_l legs[mu][l] (2(lmin+l)+1) sum_ij X[l][i][r] Cinv[i][j][l] Y[l][j][s],
 *   pref = 6 / (4 pi)^2 (9/25)^2 / 2.
 * P, Q: (nl, npol, nr) for l = lmin .. lmin+nl-1 (npol <= 3); Cinv: (npol, npol, nl), the
 * beam-weighted inverse total spectrum; legs: (nmu, nl) Legendre P_l(mu_k); w_mu: (nmu)
 * quadrature weights; weights: (nr); out: (nr, nr), symmetric.  All device pointers.
 * work: ps_fisher_deriv_work_bytes(nl, npol, nr, nmu) bytes (4 nl nr^2 doubles dominate).
 */
long long ps_fisher_deriv_work_bytes(int nl, int npol, int nr, int nmu);
int ps_fisher_deriv_gnl_loc(const double *P, const double *Q, int nl, int npol, int nr, const double *Cinv,
                            const double *legs, const double *w_mu, int nmu, const double *weights, int lmin,
                            void *work, long long work_bytes, double *out, void *stream);

/* -------------------------------------------- learned CG preconditioner
 * Kernels of a 9-point-stencil convolutional network on HEALPix maps (FP32),
 * used as the nonlinear preconditioner of a flexible CG (no counterpart in the
 * reference, whose CG uses the 'harmonic' / 'pseudo_inverse' preconditioners of
 * weighting.py:301-324).  Maps are channels-last (B, npix, C); a layer's weights
 * are W (9*cin, cout) row-major, its stencil table nbr (npix, 9) (-1 = absent). */

/* Xcol[(b,p), k*cin + c] = X[b, nbr[p,k], c] (0 where absent). */
int ps_nn_im2col9(const float *X, const int *nbr, int B, int npix, int cin, float *Xcol, void *stream);

/* dX[b,q,c] = sum_j G[(b,p_j), k_j*cin + c] over rev[q, j] = p_j*9 + k_j (j < rw, -1 unused). */
int ps_nn_col2im9(const float *G, const int *rev, int rw, int B, int npix, int cin, float *dX, void *stream);

/* C (M x N) = op(A) op(B) (+ bias, ReLU when ksplit == 1); op = transpose when trans* != 0.
 * ksplit > 1 splits K and needs work of ksplit*M*N floats (then C must be dense, ldc = N). */
int ps_nn_gemm(int transA, int transB, int M, int N, int K, const float *A, long long lda, const float *B,
               long long ldb, float *C, long long ldc, const float *bias, int relu, int ksplit, float *work,
               void *stream);

/* dY *= (Y > 0). */
int ps_nn_relu_backward(long long n, const float *Y, float *dY, void *stream);

/* out[n] = sum_m dY[m, n]; work of ps_nn_colsum_work_floats(M, N) floats. */
long long ps_nn_colsum_work_floats(int M, int N);
int ps_nn_colsum(int M, int N, const float *dY, float *work, float *out, void *stream);

/* Adam step t (>= 1) on n parameters. */
int ps_nn_adam(long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
               int t, void *stream);

/* loss = sum_b scale[b] sum_i (Y - T)^2 over B samples of per_sample values (FP64, device scalar);
 * dY = 2 scale[b] (Y - T) if dY != NULL.  work: 512 doubles. */
int ps_nn_mse(long long per_sample, int B, const float *Y, const float *T, const double *scale, float *dY,
              double *work, double *loss, void *stream);

/* Inference form of one stencil layer on bf16 matrix cores: Y = act(im2col9(X) W + b), times gate[p]
 * per pixel when gate is given, with the gather fused into the operand loads and FP32
 * accumulation.  X: bf16 (B npix, cin) with cin = 16, 32 or 64 (pad the channels with zeros);
 * Wp: W in bf16 as (32 ceil(cout / 32), 9 cin) row-major (element [n][k cin + c] = W[k cin + c, n],
 * zero rows for padded channels); cout <= 64; Y: bf16 (y_bf16 = 1) or FP32 (B npix, cout).  No
 * counterpart in the reference (the learned preconditioner is new). */
int ps_nn_conv9_bf16(const void *X, const int *nbr, int B, int npix, int cin, const void *Wp, const float *bias,
                     int cout, int relu, const float *gate, void *Y, int y_bf16, void *stream);

/* Y[b,p,j] = w[p] X[b,p,j] (per-pixel gate; Y == X allowed). */
int ps_nn_pixel_scale(int B, int npix, int c, const float *w, const float *X, float *Y, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* POLYSPEC_HIP_H */
