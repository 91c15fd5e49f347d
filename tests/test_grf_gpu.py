"""PolySpec.synalm / generate_data (the role of healpy.synalm at base.py:332-334; this generator's
random stream is its own): the empirical spectra of many seeded draws against the input C_l.

For a Gaussian a_lm with covariance C_l, C^XY_l-hat = [Re a^X_l0 a^Y*_l0 + 2 sum_{m>0} Re a^X_lm a^Y*_lm]
/ (2l + 1) has mean C^XY_l and variance (C^XX_l C^YY_l + (C^XY_l)^2) / (2l + 1) (auto: 2 C_l^2 / (2l+1));
the standardised deviations z over every (l, seed) then have mean 0 and mean square 1.  A swapped
Cholesky factor, a missing sqrt(2) on m = 0 or a dropped cross term moves one of these by many
standard errors.  Both generators are checked: host numpy and device torch."""
import numpy as np
import pytest
import torch

from estimator_setup import spectra

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("device_grf", [False, True])
def test_synalm_spectra(gpu, device_grf):
    from polyspec_amd.base import PolySpec, getlm
    nside, nseed = 16, 200
    print(f"synalm statistics device={device_grf}", flush=True)
    raw, tot, beam = spectra(nside, True)
    base = PolySpec(nside, dict(tot), beam=beam, pol=True, precision=64)
    base.grf_device = device_grf
    lmax = base.lmax
    ls, ms = getlm(lmax)
    names = ['TT', 'EE', 'BB', 'TE']
    pairs = {'TT': (0, 0), 'EE': (1, 1), 'BB': (2, 2), 'TE': (0, 1)}
    C = {k: np.asarray(tot[k])[:lmax + 1] for k in names}
    w = np.where(ms > 0, 2., 1.)
    lsel = np.arange(2, lmax + 1)
    z = {k: [] for k in names}
    for seed in range(nseed):
        a = base.synalm(seed, dict(tot), lmax)
        a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
        assert np.all(a[:, ms == 0].imag == 0)
        for k in names:
            i, j = pairs[k]
            c = np.bincount(ls, weights=w * (a[i] * np.conj(a[j])).real, minlength=lmax + 1) / (2 * np.arange(lmax + 1) + 1)
            var = (C[k[0] * 2][lsel] * C[k[1] * 2][lsel] + C[k][lsel] ** 2) / (2 * lsel + 1)
            z[k].append((c[lsel] - C[k][lsel]) / np.sqrt(var))
    n = nseed * len(lsel)
    for k in names:
        zz = np.concatenate(z[k])
        mean, ms2 = zz.mean(), np.mean(zz ** 2)
        print("%s: mean z %+.4f (se %.4f), mean z^2 %.4f" % (k, mean, 1 / np.sqrt(n), ms2), flush=True)
        assert abs(mean) < 5. / np.sqrt(n)
        assert abs(ms2 - 1) < 0.05
    # generate_data: the same draw, beam-deconvolved and truncated at a lower lmax
    h = base.generate_data(seed=7, output_type='harmonic', deconvolve_beam=True, lmax=20)
    a = base.synalm(7, dict(tot), lmax)
    h = h.cpu().numpy() if isinstance(h, torch.Tensor) else np.asarray(h)
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    sel = ls <= 20
    b = np.asarray(base.beam)[:, ls[sel]]
    assert np.allclose(h * b, a[:, sel], rtol=1e-12, atol=0)
