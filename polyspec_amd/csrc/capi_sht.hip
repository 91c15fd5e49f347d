// C ABI: library state, errors and the SHT entry points.
#include <mutex>
#include <new>
#include <string>

#include <algorithm>

#include "polyspec_hip.h"
#include "ps_common.h"
#include "sht.h"

namespace ps {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }
}  // namespace ps

struct ps_sht_plan {
    ps::ShtPlan p;
};

using namespace ps;

extern "C" {

const char* ps_last_error(void) { return ps::last_error(); }

int ps_version(void) { return PS_ABI_VERSION; }

int ps_init(int device) {
    int n = 0;
    PS_CHECK_HIP(hipGetDeviceCount(&n));
    PS_REQUIRE(device >= 0 && device < n, "ps_init: device index out of range");
    PS_CHECK_HIP(hipSetDevice(device));
    return 0;
}

int ps_sht_plan_create(int nside, int lmax, int mmax, int spin, int precision, ps_sht_plan** out) {
    PS_REQUIRE(out != nullptr, "ps_sht_plan_create: out is NULL");
    *out = nullptr;
    PS_REQUIRE(nside >= 1 && nside <= 1024, "ps_sht_plan_create: nside must be in [1, 1024]");
    PS_REQUIRE(lmax >= 0 && lmax <= 4 * nside, "ps_sht_plan_create: lmax must be in [0, 4 nside]");
    PS_REQUIRE(mmax >= 0 && mmax <= lmax, "ps_sht_plan_create: mmax must be in [0, lmax]");
    PS_REQUIRE(spin >= 0 && spin <= 3, "ps_sht_plan_create: spin must be in [0, 3]");
    PS_REQUIRE(precision == 32 || precision == 64, "ps_sht_plan_create: precision must be 32 or 64");
    ps_sht_plan* h = new (std::nothrow) ps_sht_plan();
    PS_REQUIRE(h != nullptr, "ps_sht_plan_create: out of host memory");
    h->p.nside = nside;
    h->p.lmax = lmax;
    h->p.mmax = mmax;
    h->p.spin = spin;
    h->p.prec = precision / 8;
    hipGetDevice(&h->p.device);
    int rc = h->p.build();
    if (!rc) rc = sht_mark_dead(h->p);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

int ps_sht_plan_destroy(ps_sht_plan* plan) {
    delete plan;
    return 0;
}

int ps_sht_plan_info(const ps_sht_plan* plan, long long* npix, long long* nlm, int* ncomp) {
    PS_REQUIRE(plan != nullptr, "ps_sht_plan_info: plan is NULL");
    if (npix) *npix = plan->p.npix;
    if (nlm) *nlm = plan->p.nlm;
    if (ncomp) *ncomp = plan->p.ncomp;
    return 0;
}

int ps_sht_lmajor_multiple(const ps_sht_plan* plan) {
    if (!plan) return -1;
    return sht_synth_lmajor_ld_multiple(plan->p);
}

long long ps_sht_workspace_bytes(const ps_sht_plan* plan, int nmaps, int lmajor_input) {
    if (!plan || nmaps < 1) return -1;
    size_t b = sht_phase_bytes(plan->p, nmaps);
    b = (b + 255) & ~(size_t)255;
    if (!lmajor_input) {
        int mult = sht_synth_lmajor_ld_multiple(plan->p);
        long ncol = (long)nmaps * plan->p.ncomp;
        long ld = (ncol + mult - 1) / mult * mult;
        b += (size_t)ld * plan->p.nlm * 2 * plan->p.prec;
    } else {
        b = std::max(b, sht_adjoint_work_bytes(plan->p, nmaps));
    }
    return (long long)b;
}

int ps_sht_synthesis(const ps_sht_plan* plan, int nmaps, const void* alm, void* map, void* work,
                     long long work_bytes, void* stream) {
    PS_REQUIRE(plan && alm && map && work, "ps_sht_synthesis: NULL argument");
    PS_REQUIRE(nmaps >= 1, "ps_sht_synthesis: nmaps must be >= 1");
    const ShtPlan& p = plan->p;
    long long need = ps_sht_workspace_bytes(plan, nmaps, 0);
    PS_REQUIRE(work_bytes >= need, "ps_sht_synthesis: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    size_t phb = (sht_phase_bytes(p, nmaps) + 255) & ~(size_t)255;
    void* phase = work;
    void* alm_t = (char*)work + phb;
    int mult = sht_synth_lmajor_ld_multiple(p);
    int ncol = nmaps * p.ncomp;
    int ld = (ncol + mult - 1) / mult * mult;
    int rc = sht_alm_to_lmajor(p, ncol, alm, p.nlm, ld, alm_t, st);
    if (rc) return rc;
    return sht_synthesis_lmajor(p, nmaps, alm_t, ld, map, p.npix, phase, st);
}

int ps_sht_synthesis_grad(const ps_sht_plan* plan, int nmaps, const void* G, void* map, void* work,
                          long long work_bytes, void* stream) {
    PS_REQUIRE(plan && G && map && work, "ps_sht_synthesis_grad: NULL argument");
    PS_REQUIRE(nmaps >= 1, "ps_sht_synthesis_grad: nmaps must be >= 1");
    const ShtPlan& p = plan->p;
    PS_REQUIRE(p.prec == 4 && p.spin >= 1, "ps_sht_synthesis_grad: FP32 plans of spin >= 1 only");
    PS_REQUIRE(work_bytes >= ps_sht_workspace_bytes(plan, nmaps, 0), "ps_sht_synthesis_grad: workspace too small");
    hipStream_t st = (hipStream_t)stream;
    const size_t phb = (sht_phase_bytes(p, nmaps) + 255) & ~(size_t)255;
    void* phase = work;
    void* alm_t = (char*)work + phb;
    // G of map b in column b, rows of even length (16-byte copies); the kernel reads columns past
    // ld as zero, and the (G, C) workspace holds 2 nmaps >= ld columns
    const int ld = (nmaps + 1) / 2 * 2;
    int rc = sht_alm_to_lmajor(p, nmaps, G, p.nlm, ld, alm_t, st);
    if (rc) return rc;
    if ((rc = sht_legendre_synthesis_grad(p, nmaps, alm_t, ld, phase, st))) return rc;
    return sht_ring_synthesis(p, nmaps, phase, map, p.npix, st);
}

int ps_sht_alm_to_lmajor(const ps_sht_plan* plan, int ncol, const void* alm, int ld, void* alm_t, void* stream) {
    PS_REQUIRE(plan && alm && alm_t, "ps_sht_alm_to_lmajor: NULL argument");
    PS_REQUIRE(ncol >= 1 && ld >= ncol && ld <= 65535 * 16, "ps_sht_alm_to_lmajor: need 1 <= ncol <= ld <= 1048560");
    return sht_alm_to_lmajor(plan->p, ncol, alm, plan->p.nlm, ld, alm_t, (hipStream_t)stream);
}

int ps_sht_synthesis_lmajor(const ps_sht_plan* plan, int nmaps, const void* alm_t, int ld, void* map,
                            void* work, long long work_bytes, void* stream) {
    PS_REQUIRE(plan && alm_t && map && work, "ps_sht_synthesis_lmajor: NULL argument");
    PS_REQUIRE(nmaps >= 1, "ps_sht_synthesis_lmajor: nmaps must be >= 1");
    const ShtPlan& p = plan->p;
    int mult = sht_synth_lmajor_ld_multiple(p);
    PS_REQUIRE(ld >= nmaps * p.ncomp && ld % mult == 0,
               "ps_sht_synthesis_lmajor: ld must cover nmaps*ncomp columns and be a multiple of "
               "ps_sht_lmajor_multiple()");
    PS_REQUIRE(work_bytes >= ps_sht_workspace_bytes(plan, nmaps, 1), "ps_sht_synthesis_lmajor: workspace too small");
    return sht_synthesis_lmajor(p, nmaps, alm_t, ld, map, p.npix, work, (hipStream_t)stream);
}

int ps_sht_phase_bytes(const ps_sht_plan* plan, int nmaps, long long* bytes) {
    PS_REQUIRE(plan && bytes && nmaps >= 1, "ps_sht_phase_bytes: bad argument");
    *bytes = (long long)sht_phase_bytes(plan->p, nmaps);
    return 0;
}

int ps_sht_legendre_synthesis_lmajor(const ps_sht_plan* plan, int nmaps, const void* alm_t, int ld, void* phase,
                                     void* stream) {
    PS_REQUIRE(plan && alm_t && phase, "ps_sht_legendre_synthesis_lmajor: NULL argument");
    const ShtPlan& p = plan->p;
    int mult = sht_synth_lmajor_ld_multiple(p);
    PS_REQUIRE(nmaps >= 1 && ld >= nmaps * p.ncomp && ld % mult == 0,
               "ps_sht_legendre_synthesis_lmajor: bad nmaps / ld");
    return sht_legendre_synthesis_lmajor(p, nmaps, alm_t, ld, phase, (hipStream_t)stream);
}

int ps_sht_ring_synthesis(const ps_sht_plan* plan, int nmaps, const void* phase, void* map, void* stream) {
    PS_REQUIRE(plan && phase && map && nmaps >= 1, "ps_sht_ring_synthesis: bad argument");
    return sht_ring_synthesis(plan->p, nmaps, phase, map, plan->p.npix, (hipStream_t)stream);
}

long long ps_sht_ring_gnl_work_bytes(const ps_sht_plan* plan, int nr) {
    PS_REQUIRE(plan && nr >= 1, "ps_sht_ring_gnl_work_bytes: bad argument");
    const long n = sht_ring_gnl_parts(plan->p, nr);
    PS_REQUIRE(n >= 0, "ps_sht_ring_gnl_work_bytes: no staged ring kernel fits this plan");
    return (long long)n * (long long)sizeof(double);
}

int ps_sht_ring_gnl(const ps_sht_plan* plan, int nr, const void* phase, const double* U, const double* V,
                    const double* w, const double* inv_s2, double c4, double c2, void* work, long long work_bytes,
                    double* out, int accumulate, void* stream) {
    PS_REQUIRE(plan && phase && w && work && out && nr >= 1, "ps_sht_ring_gnl: bad argument");
    return sht_ring_gnl(plan->p, nr, phase, U, V, w, inv_s2, c4, c2, (double*)work,
                        (long)(work_bytes / (long long)sizeof(double)), out, accumulate, (hipStream_t)stream);
}

int ps_sht_adjoint(const ps_sht_plan* plan, int nmaps, const void* map, void* alm, double scale, void* work,
                   long long work_bytes, void* stream) {
    PS_REQUIRE(plan && alm && map && work, "ps_sht_adjoint: NULL argument");
    PS_REQUIRE(nmaps >= 1, "ps_sht_adjoint: nmaps must be >= 1");
    const ShtPlan& p = plan->p;
    PS_REQUIRE(work_bytes >= ps_sht_workspace_bytes(plan, nmaps, 1), "ps_sht_adjoint: workspace too small");
    return sht_adjoint(p, nmaps, map, p.npix, alm, p.nlm, scale, work, (hipStream_t)stream);
}

int ps_sht_shape_candidates(const ps_sht_plan* plan, int nmaps, int stage, int* ids, int max_ids) {
    PS_REQUIRE(plan && nmaps >= 1, "ps_sht_shape_candidates: bad argument");
    PS_REQUIRE(stage >= 0 && stage < SHT_NSTAGE, "ps_sht_shape_candidates: stage must be 0..3");
    return sht_shape_candidates(plan->p, nmaps, stage, ids, ids ? max_ids : 0);
}

int ps_sht_get_shapes(const ps_sht_plan* plan, int nmaps, int* shapes) {
    PS_REQUIRE(plan && shapes && nmaps >= 1, "ps_sht_get_shapes: bad argument");
    const ShtShapes* s = plan->p.shapes_for(nmaps * plan->p.ncomp);
    for (int st = 0; st < SHT_NSTAGE; st++) {
        const int id = s ? s->id[st] : -1;
        shapes[st] = id >= 0 ? id : sht_default_shape(plan->p, nmaps, st);
    }
    return s ? 1 : 0;
}

int ps_sht_set_shapes(ps_sht_plan* plan, int nmaps, const int* shapes) {
    PS_REQUIRE(plan && shapes && nmaps >= 1, "ps_sht_set_shapes: bad argument");
    return sht_set_shapes(plan->p, nmaps, shapes);
}

long long ps_sht_autotune_work_bytes(const ps_sht_plan* plan, int nmaps) {
    if (!plan || nmaps < 1) return -1;
    return (long long)sht_autotune_work_bytes(plan->p, nmaps);
}

int ps_sht_autotune(ps_sht_plan* plan, int nmaps, void* work, long long work_bytes, int reps, double* times,
                    void* stream) {
    PS_REQUIRE(plan && work && nmaps >= 1, "ps_sht_autotune: bad argument");
    PS_REQUIRE(work_bytes >= 0, "ps_sht_autotune: negative workspace size");
    return sht_autotune(plan->p, nmaps, work, (size_t)work_bytes, reps, (hipStream_t)stream, times);
}

}  // extern "C"
