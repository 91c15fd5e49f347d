// Internal interface of the HEALPix SHT kernels (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace ps {

struct RingInfo {
    long start;  // first pixel of the ring (RING order)
    int nphi;    // pixels in the ring
    int phi_p;   // phi0 = pi * phi_p / phi_q
    int phi_q;
    int M;       // FFT length (nphi if a power of two, else the Bluestein length)
    int logM;
    int voff;    // offset of the Bluestein kernel in the table, -1 for power-of-two rings
    int coff;    // offset of the chirp exp(i pi t^2 / nphi), t < nphi, in the chirp table (Bluestein rings)
};

// Launch shapes for one batch width (ncol = nmaps * ncomp), chosen by timing in sht_autotune
// or set by the caller; -1 keeps the built-in rule for that stage.
enum { SHT_STAGE_SYNTH = 0, SHT_STAGE_RING_POW2 = 1, SHT_STAGE_RING_BLUE = 2, SHT_STAGE_ADJ = 3, SHT_NSTAGE = 4 };
struct ShtShapes {
    int ncol = 0;
    int id[SHT_NSTAGE] = {-1, -1, -1, -1};
};
constexpr int SHT_MAX_SHAPES = 16;

struct ShtPlan {
    int nside = 0, lmax = 0, mmax = 0, spin = 0, prec = 4;  // prec: bytes per real (4 or 8)
    int nrings = 0, npairs = 0, ncomp = 1;
    long npix = 0, nlm = 0, maxM = 1;
    long maxM_pow2 = 1, maxM_blue = 1;  // largest FFT length of the power-of-two / Bluestein rings
    long blue_total = 0;                // Bluestein kernel entries per order (d_blue holds two orders)
    int n_pow2 = 0;                     // power-of-two rings, listed first in d_rlist
    int* d_rlist = nullptr;             // ring indices: power-of-two rings, then Bluestein rings
    int device = 0;
    RingInfo* d_rings = nullptr;
    double *d_l2s = nullptr, *d_l2sh = nullptr, *d_l2ch = nullptr, *d_seedL = nullptr, *d_seedS = nullptr;
    int *d_seedA = nullptr, *d_seedB = nullptr;
    void *d_cth = nullptr, *d_tw = nullptr, *d_blue = nullptr, *d_coef = nullptr, *d_chirp = nullptr;
    // dead[m * npairs + jp] = 1: |lambda_lm| (both spin signs) < 2^-200 at ring pair jp for every l,
    // so the pair contributes nothing at order m (the evanescent region of high m near the poles)
    unsigned char* d_dead = nullptr;
    // per-batch-width launch shapes (not thread-safe against concurrent transforms on this plan)
    ShtShapes shapes[SHT_MAX_SHAPES];
    int nshapes = 0;
    // side stream and events of the overlapped spin-0 synthesis (created on first use, on the
    // plan's device; like the shapes, not for concurrent transforms on this plan)
    mutable hipStream_t side = nullptr;
    mutable hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    const ShtShapes* shapes_for(int ncol) const {
        for (int i = 0; i < nshapes; i++)
            if (shapes[i].ncol == ncol) return &shapes[i];
        return nullptr;
    }
    int build();
    ~ShtPlan();
};

// fills p.d_dead (one kernel on the null stream, synchronised); called once after build()
int sht_mark_dead(ShtPlan& p);
int sht_synth_lmajor_ld_multiple(const ShtPlan& p);
size_t sht_phase_bytes(const ShtPlan& p, int nmap);
int sht_adjoint_nsplit(const ShtPlan& p, int nmap);
size_t sht_adjoint_work_bytes(const ShtPlan& p, int nmap);
int sht_synthesis_lmajor(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* map, long map_stride,
                         void* phase, hipStream_t st);
int sht_legendre_synthesis_lmajor(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* phase,
                                  hipStream_t st);
int sht_ring_synthesis(const ShtPlan& p, int nmap, const void* phase, void* map, long map_stride, hipStream_t st);
// Legendre stage of the spin-s synthesis of (G, C = 0) from G alone (alm_t column b = G of map b);
// the phases of the R and I components of nmap maps.  FP32 plans of spin >= 1.
int sht_legendre_synthesis_grad(const ShtPlan& p, int nmap, const void* alm_t, int ld, void* phase,
                                hipStream_t st);
// Ring synthesis of 2 nr interleaved columns (a_r, b_r) = (s_r P_r, Q_r / s_r) fused with the
// gNL-loc sum sum_r w_r sum_pix [c4 P^3 Q + c2 (P^2 U_r + P Q V_r)] (U, V (nr, npix) FP64 or both
// null; inv_s2 = 1 / s_r^2 or null for s_r = 1) into *out (FP64, added when accumulate); part:
// sht_ring_gnl_parts(p, nr) doubles.  FP32 spin-0 plans whose ring classes have staged kernels
// (sht_ring_gnl_parts >= 0).
long sht_ring_gnl_parts(const ShtPlan& p, int nr);
int sht_ring_gnl(const ShtPlan& p, int nr, const void* phase, const double* U, const double* V, const double* w,
                 const double* inv_s2, double c4, double c2, double* part, long nparts, double* out, int accumulate,
                 hipStream_t st);
int sht_alm_to_lmajor(const ShtPlan& p, int ncol, const void* alm, long alm_stride, int ld, void* alm_t,
                      hipStream_t st);
// work: sht_adjoint_work_bytes(p, nmap) bytes (ring phases, then ring-slice partial sums)
int sht_adjoint(const ShtPlan& p, int nmap, const void* map, long map_stride, void* alm, long alm_stride,
                double scale, void* work, hipStream_t st);

// Launch-shape candidates of one stage for nmap maps (ids usable by sht_set_shapes); returns
// their count, writing at most `max` ids.  The built-in rule's choice is reported by
// sht_default_shape.
int sht_shape_candidates(const ShtPlan& p, int nmap, int stage, int* ids, int max);
int sht_default_shape(const ShtPlan& p, int nmap, int stage);
// ids[SHT_NSTAGE]; -1 entries keep the built-in rule
int sht_set_shapes(ShtPlan& p, int nmap, const int* ids);
size_t sht_autotune_work_bytes(const ShtPlan& p, int nmap);
// Times every candidate of every stage on synthetic data in `work` (reps launches each, after
// one untimed launch) and keeps the fastest; a candidate replaces the built-in choice only when
// it is at least 3% faster.  Synchronises the stream.  times (optional): per stage, the
// winning and the built-in shape's milliseconds per launch.
int sht_autotune(ShtPlan& p, int nmap, void* work, size_t work_bytes, int reps, hipStream_t st,
                 double* times);

}  // namespace ps
