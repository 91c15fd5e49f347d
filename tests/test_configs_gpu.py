"""Every benchmark configuration at its own size, on the GPU.

At these sizes the FP64 CPU oracle is too slow, so the checks are between the batched FP32 path
and the FP64 kernels (themselves pinned to the oracle at Nside <= 256 and, at full size, by the
adjointness identity in test_sht_gpu.py), with the tolerances the targets set:
  1. Nside 1024, lmax 2048, 64 TEB a_lm sets: spin-0 + spin-2 synthesis and adjoint, complex64 vs
     complex128, max |da| / max |a| <= 1e-6 per field (the a_lm target).
  2. Nside 512, T+E, lmin 2, lmax 1024, 40 radial nodes, |b| < 20 deg cut + 200 holes of 0.5 deg
     apodised over 1 deg: the batched FP32 gNL-loc and tauNL-loc numerators against the reference's
     per-simulation FP64 loop (batch_sims=False) within 1e-4 sigma, sigma = 1 / sqrt(F) (the
     amplitude target), and the FP32 Fisher contribution against FP64 within 1e-4.
  3. Nside 512 CG S^-1 (base lmax 1535) over that mask with lognormal noise: FP32 and FP64 SHTs
     give the same early statistic history, stopping iterations within 3, and solutions at equal
     iteration counts within FP32 CG's reach.
  4. BSpecBin TTT/TTE/TEE/EEE, both parities, 20 log-spaced bins, one simulation each at Nside 256 /
     512 / 1024 (lmax 512 / 1024 / 2000) in one mixed-resolution launch: FP32 against the FP64
     per-(bin, field) path.
"""
import numpy as np
import pytest
import torch

from estimator_setup import spectra

pytestmark = pytest.mark.gpu


def _log_bins(lmax, nbins):
    """nbins log-spaced l-bin edges on [2, lmax + 1], strictly increasing integers."""
    out = [2]
    for x in np.geomspace(2, lmax + 1, nbins + 1)[1:]:
        out.append(max(int(round(x)), out[-1] + 1))
    out[-1] = lmax + 1
    return out


def test_cfg1_teb_roundtrip_complex64_vs_complex128(gpu):
    from polyspec_amd._abi import ShtPlan
    from polyspec_amd.base import getlm
    nside, lmax, nsets = 1024, 2048, 64
    print(f"cfg1 round trip nside={nside} lmax={lmax} sets={nsets}", flush=True)
    ls, ms = getlm(lmax)
    lf = np.maximum(ls, 1).astype(np.float64)
    tt = np.where(ls >= 2, 2 * np.pi * 1e3 / (lf * (lf + 1)), 0.)
    ee = 0.05 * tt
    c21 = np.where(tt > 0, 0.3 * np.sqrt(tt * ee) / np.sqrt(np.where(tt > 0, tt, 1.)), 0.)
    c22 = np.sqrt(np.maximum(ee - c21 ** 2, 0.))
    d = lambda x: torch.from_numpy(x).to(gpu)
    g = torch.Generator(device=gpu).manual_seed(11)
    m0 = d(ms == 0)

    def gauss():
        z = torch.randn((nsets, len(ls)), dtype=torch.complex128, device=gpu, generator=g)
        z[:, m0] = (z[:, m0].real * np.sqrt(2.)).to(torch.complex128)
        return z
    g1, g2 = gauss(), gauss()
    aT = (d(np.sqrt(tt)) * g1).unsqueeze(1).contiguous()
    aEB = torch.zeros((nsets, 2, len(ls)), dtype=torch.complex128, device=gpu)
    aEB[:, 0] = d(c21) * g1 + d(c22) * g2
    del g1, g2
    sc = 4 * np.pi / (12 * nside ** 2)
    out = {}
    for prec in (64, 32):
        p0, p2 = ShtPlan(nside, lmax, 0, prec), ShtPlan(nside, lmax, 2, prec)
        mT, mQU = p0.synthesis(aT.to(p0.cdtype)), p2.synthesis(aEB.to(p2.cdtype))
        out[prec] = (mT, mQU, p0.adjoint(mT, sc), p2.adjoint(mQU, sc))
        del p0, p2
    torch.cuda.synchronize()

    def rel(x32, x64, norm=None):
        ref = x64 if norm is None else norm
        return float((x32.to(x64.dtype) - x64).abs().max() / ref.abs().max())
    eT, eE = rel(out[32][2], out[64][2]), rel(out[32][3][:, 0], out[64][3][:, 0])
    eB = rel(out[32][3][:, 1], out[64][3][:, 1], norm=out[64][3][:, 0])
    emT, emQU = rel(out[32][0], out[64][0]), rel(out[32][1], out[64][1])
    print("cfg1: alm T %.2e E %.2e B/E %.2e; maps T %.2e QU %.2e" % (eT, eE, eB, emT, emQU), flush=True)
    assert eT <= 1e-6 and eE <= 1e-6 and eB <= 1e-6
    assert emT <= 3e-6 and emQU <= 3e-6


def _cfg2_template(prec, batch, templates, mask, nside=512, lmax=1024, nr=40):
    from scipy.special import roots_legendre, spherical_jn
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    raw, tot, beam = spectra(nside, True)
    base = PolySpec(nside, dict(tot), beam=beam, pol=True, precision=prec)
    base.grf_device = True
    k = np.geomspace(1e-4, 0.2, 200)
    Tl = np.zeros((lmax + 1, len(k)))
    for l in range(2, lmax + 1):
        Tl[l] = spherical_jn(l, 14000. * k)
    x, wg = roots_legendre(nr)
    r = 600. * x + 14000.
    w = r ** 2 * wg * 600.
    W = Weightings(base, mask, dict(raw))
    tm = TSpecTemplate(base, mask, W.applySinv_ideal, templates, k, np.asarray([Tl, Tl]), 2, lmax, r_values=r,
                       r_weights={t: w for t in templates}, batch_sims=batch)
    return base, tm, tot


@pytest.mark.parametrize("templates", [['gNL-loc', 'tauNL-loc'], ['gNL-loc']])
def test_cfg2_numerators_and_fisher_fp32_batched_vs_fp64_reference_loop(gpu, templates):
    """With gNL-loc alone the batched FP32 path takes the fused ring-stage sum (no P/Q maps), whose
    magnitude balancing of the P and Q columns this pins at the config size."""
    from polyspec_amd.masks import galactic_cut_mask
    nside, lmax, nr, npairs, ndata = 512, 1024, 40, 2, 4
    print(f"cfg2 numerators nside={nside} lmax={lmax} nr={nr} pairs={npairs}", flush=True)
    mask = galactic_cut_mask(nside)
    base64, tm64, tot = _cfg2_template(64, False, templates, mask)
    mk = torch.from_numpy(mask).to(gpu)
    maps = [(base64.generate_data(seed=int(7e6) + i) * mk).contiguous() for i in range(ndata + 2 * npairs)]
    sims = [(maps[ndata + 2 * i], maps[ndata + 2 * i + 1]) for i in range(npairs)]
    tm64.load_sims(lambda ii: sims[ii], npairs, input_type='map', preload=True)
    n64 = np.array([tm64.Tl_numerator(maps[i]) for i in range(ndata)])
    a_maps = [base64.generate_data(seed=int(2e9) + i, output_type='harmonic', deconvolve_beam=True) for i in range(2)]
    f64 = tm64.fisher_contribution_from_grfs(a_maps)
    del tm64, base64
    torch.cuda.empty_cache()
    print("cfg2: fp64 reference loop done", flush=True)
    # on this masked sky tauNL-loc's t4, t2 and t0 are each about 2300 sigma (the mask's mean field
    # in A_dd, A_ss); the batched FP32 path takes its tau sums over A - M, M the simulations' mean
    # field, which leaves the sum unchanged and its terms near sigma
    base32, tm32, _ = _cfg2_template(32, True, templates, mask)
    assert tm32._gnl_fused_ok() == (templates == ['gNL-loc'])
    tm32.load_sims(lambda ii: (sims[ii][0].float(), sims[ii][1].float()), npairs, input_type='map')
    n32 = np.array([tm32.Tl_numerator(maps[i].float()) for i in range(ndata)])
    f32 = tm32.fisher_contribution_from_grfs(a_maps)
    # amplitude errors in units of sigma = 1 / sqrt(F): |dN| / F * sqrt(F) = |dN| / sqrt(F)
    sig = np.sqrt(np.diag(f64))
    diff = np.abs(n32 - n64).max(axis=0) / sig
    print("cfg2: numerators |fp32 - fp64| / sqrt(F): %s (scatter over data maps / sqrt(F): %s); Fisher rel %.2e" %
          (diff, n64.std(axis=0, ddof=1) / sig, np.abs(f32 - f64).max() / np.abs(f64).max()), flush=True)
    assert np.all(diff <= 1e-4), diff
    assert np.abs(f32 - f64).max() <= 1e-4 * np.abs(f64).max()


def test_cfg3_cg_fp32_vs_fp64_histories(gpu):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.masks import galactic_cut_mask
    from polyspec_amd.weighting import CGD, Weightings
    nside = 512
    print(f"cfg3 CG nside={nside}", flush=True)
    mask = galactic_cut_mask(nside)
    masks = np.repeat(mask[None], 3, 0)
    raw, tot, beam = spectra(nside, True)
    # lognormal noise variance: sigma0^2 exp(g), g a GRF smoothed on 5 deg with std 0.5
    b64 = PolySpec(nside, dict(tot), beam=beam, pol=True, precision=64)
    L = b64.lmax
    l = np.arange(L + 1)
    cl = np.exp(-l * (l + 1) * (np.radians(5.) / np.sqrt(8 * np.log(2))) ** 2)
    gm = b64.to_map(b64.synalm(11, {'TT': cl, 'EE': 0 * cl, 'TE': 0 * cl, 'BB': 0 * cl}, L))[0]
    gm = (gm - gm.mean()) / gm.std() * 0.5
    v = (30. * np.pi / 10800.) ** 2 / b64.A_pix * np.exp(gm)
    ncov = np.stack([v, 2 * v, 2 * v])
    sig = b64.to_map(b64.synalm(8, {k: np.asarray(x) * beam[0] ** 2 for k, x in raw.items()}, L))
    rng = np.random.default_rng(5)
    data = (sig + np.sqrt(ncov) * rng.standard_normal(ncov.shape)) * masks
    hist, sols = {}, {}
    n_cmp = 40  # solutions compared at one iteration count, so only rounding separates them
    for prec in (64, 32):
        base = b64 if prec == 64 else PolySpec(nside, dict(tot), beam=beam, pol=True, precision=32)
        W = Weightings(base, masks, dict(raw), noise_cov=ncov)
        d = torch.from_numpy(data).to(gpu, base.rdtype)
        cg = CGD(W, d, x0=W.wiener_start(d), preconditioner='harmonic')

        def keep(c, prec=prec):
            if c.i == n_cmp:
                sols[prec] = (c.x + c.x0).to(torch.complex128)
        cg.callback = keep
        cg.run(nstep=200, thresh=1e-6)
        hist[prec] = np.array([h[1] for h in cg.history])
    n = min(len(hist[32]), len(hist[64]))
    drift = np.abs(np.log(hist[32][:n] / hist[64][:n]))
    sol = float((sols[32] - sols[64]).abs().max() / sols[64].abs().max())
    print("cfg3: stop at %d (fp64) / %d (fp32); max |log ratio| of rz/rz0 over the first 20 iterations: %.2e, "
          "first %d: %.2e; solutions at iteration %d: max rel %.2e" %
          (len(hist[64]), len(hist[32]), drift[:20].max(), n_cmp, drift[:n_cmp].max(), n_cmp, sol), flush=True)
    # FP32 transforms change the statistic in its late, flat part (near the threshold the stopping
    # iteration is sensitive to it), not in the early iterations
    assert abs(len(hist[32]) - len(hist[64])) <= 3
    assert drift[:20].max() < 1e-4
    assert sol < 1e-3


def test_cfg4_bspec_mixed_resolutions_fp32_vs_fp64(gpu):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.bspec_bin import BSpecBin, Bl_numerator_multi
    from polyspec_amd.masks import galactic_cut_mask
    from polyspec_amd.weighting import Weightings
    print("cfg4 bspec nside 256/512/1024", flush=True)
    fields = ['TTT', 'TTE', 'TEE', 'EEE']
    res = {}
    for prec, batch in ((32, True), (64, False)):
        analyses, datas = [], []
        for nside in (256, 512, 1024):
            lmax = min(2 * nside, 2000)
            raw, tot, beam = spectra(nside, True)
            mask = galactic_cut_mask(nside)
            base = PolySpec(nside, dict(tot), beam=beam, pol=True, precision=prec)
            base.grf_device = True
            W = Weightings(base, mask, dict(raw))
            bs = BSpecBin(base, mask, W.applySinv_ideal, _log_bins(lmax, 20), fields=fields, parity='both',
                          batch_maps=batch)
            mk = torch.from_numpy(mask).to(gpu, base.rdtype)
            maps = [(base.generate_data(seed=int(9e6) + 10 * nside + i) * mk).contiguous() for i in range(2)]
            bs.load_sims(lambda i, maps=maps: maps[1 + i], 1, preload=True)
            analyses.append(bs)
            datas.append(maps[0])
        if batch:
            res[prec] = Bl_numerator_multi(analyses, datas)
        else:
            res[prec] = [bs.Bl_numerator(d) for bs, d in zip(analyses, datas)]
        del analyses, datas
        torch.cuda.empty_cache()
    for i, nside in enumerate((256, 512, 1024)):
        a, b = np.asarray(res[32][i]), np.asarray(res[64][i])
        err = np.abs(a - b).max() / np.abs(b).max()
        print("cfg4: nside %d: %d entries, max |fp32 - fp64| / max |B| %.2e" % (nside, b.size, err), flush=True)
        assert err < 1e-5
