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

/* Same, with alm given l-major: alm_t[idx(l,m) * ld + col], col = map*ncomp + comp,
 * ld >= nmaps*ncomp a multiple of ps_sht_lmajor_multiple(); padding columns must be zero. */
int ps_sht_synthesis_lmajor(const ps_sht_plan *plan, int nmaps, const void *alm_t, int ld, void *map,
                            void *work, long long work_bytes, void *stream);

/* map -> alm: scale * sum_p map_p sY*_lm(p) (ducc0.sht.adjoint_synthesis times
 * `scale`; PolySpec passes scale = pi/(3 nside^2) = 4 pi/Npix, base.py:147-151, 193, 206, 253). */
int ps_sht_adjoint(const ps_sht_plan *plan, int nmaps, const void *map, void *alm, double scale, void *work,
                   long long work_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* POLYSPEC_HIP_H */
