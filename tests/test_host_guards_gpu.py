"""Host-side guards in front of the device kernels: operands whose extents disagree are refused
before any launch, a change of radial nodes drops every simulation product built on the old
nodes, and graph capture refuses an S^-1 that synchronises with the host."""
import numpy as np
import pytest
import torch

import oracle
from estimator_setup import grf_alm, radial_nodes, spectra, transfer
from oracle.tspec import OracleTSpec

pytestmark = pytest.mark.gpu


def test_wrappers_reject_mismatched_stacks(gpu):
    from polyspec_amd import kernels as K
    print("wrapper shape guards", flush=True)
    nr, npix = 4, 768
    P = torch.ones((nr, npix), dtype=torch.float32, device=gpu)
    Q = torch.ones_like(P)
    U = torch.zeros((nr - 1, npix), dtype=torch.float64, device=gpu)
    V = torch.zeros((nr - 1, npix), dtype=torch.float64, device=gpu)
    w = torch.ones(nr, dtype=torch.float64)
    out = torch.full((), 7., dtype=torch.float64, device=gpu)
    with pytest.raises(AssertionError):  # U, V sized for fewer radial nodes than P
        K.gnl_loc_fused_sum(P, Q, w, 1., U, V, 1., out=out)
    with pytest.raises(AssertionError):  # one weight too many
        K.gnl_loc_fused_sum(P, Q, torch.ones(nr + 1, dtype=torch.float64), 1., out=out)
    with pytest.raises(AssertionError):  # Q of another precision
        K.gnl_loc_fused_sum(P, Q.double(), w, 1., out=out)
    with pytest.raises(AssertionError):
        K.gnl_loc_disc_sum(w, P, P[:-1].contiguous(), Q, Q, out=out)
    with pytest.raises(AssertionError):
        K.mc_accumulate(P, Q[:, :-1].contiguous(), U, V)
    with pytest.raises(AssertionError):
        K.multiplyRR(P, Q[:-1].contiguous())
    with pytest.raises(AssertionError):
        K.multiply_asym(P, Q, Q, None)
    lm = torch.zeros((nr, 10), dtype=torch.complex64, device=gpu)  # nlm(3) = 10
    f = torch.zeros((4, 1, nr), dtype=torch.float32, device=gpu)
    with pytest.raises(AssertionError):  # table shorter than lmax + 1
        K.radial_sum(lm, w, f[:3].contiguous(), 3, 2, 1)
    with pytest.raises(AssertionError):  # A and B of different radial counts
        K.tau_sum_n0(lm[None], lm[None, :-1].contiguous(), np.zeros((4, nr, nr)), w, 3, 1, 3, out=out)
    torch.cuda.synchronize()
    assert float(out) == 7.  # nothing was launched


def _template(nside, lmax, nr, templates, **kw):
    from polyspec_amd.base import PolySpec
    from polyspec_amd.tspec_template import TSpecTemplate
    from polyspec_amd.weighting import Weightings
    raw, tot, beam = spectra(nside, False)
    base = PolySpec(nside, dict(tot), beam=beam, pol=False, precision=64)
    mask = np.ones(12 * nside ** 2)
    W = Weightings(base, mask, dict(raw))
    k, Tl = transfer(lmax, False)
    r, w = radial_nodes(nr)
    tm = TSpecTemplate(base, mask, kw.pop('applySinv', W.applySinv_ideal), templates, k, Tl, 2, lmax, r_values=r,
                       r_weights={t: w for t in templates}, **kw)
    return tm, tot, W


def test_new_radial_nodes_drop_old_simulations(gpu):
    """After the radial nodes change (as optimize_radial_sampling_1d does) the simulations must be
    loaded again: a numerator with the old Monte-Carlo sums is refused, and after reloading it
    equals the oracle on the new nodes."""
    print("radial nodes change: nside=16 lmax=24", flush=True)
    nside, lmax = 16, 24
    templates = ['gNL-loc', 'tauNL-loc']
    tm, tot, _ = _template(nside, lmax, 3, templates)
    rng = np.random.default_rng(3)
    L = 3 * nside - 1
    data = oracle.to_map(grf_alm(rng, tot, L, False), nside, L)
    sims = [(grf_alm(rng, tot, lmax, False), grf_alm(rng, tot, lmax, False)) for _ in range(2)]
    tm.load_sims(lambda ii: sims[ii], 2, input_type='harmonic')
    tm.Tl_numerator(data)
    r, w = radial_nodes(5)  # more nodes than the stacks were built for
    tm.r_arr, tm.N_r = r, len(r)
    tm.r_weights = {t: w for t in templates}
    tm._prepare_templates()
    with pytest.raises(Exception, match="bias simulations"):
        tm.Tl_numerator(data)
    tm.load_sims(lambda ii: sims[ii], 2, input_type='harmonic')
    got = tm.Tl_numerator(data)
    orc = OracleTSpec(nside, 2, lmax, tm.Lmin, tm.Lmax, tot, tm.base.beam, False, templates, tm.plLXs[0], tm.qlXs,
                      tm.FLLs[0], tm.r_weights)
    ref = orc.numerator(data, sims, input_type='map', sim_type='harmonic')
    ref4 = orc.numerator(data, sims, input_type='map', sim_type='harmonic', include_disconnected_term=False)
    for i in range(len(templates)):
        assert abs(got[i] - ref[i]) <= 1e-8 * abs(ref4[i]), (templates[i], got[i], ref[i])


def test_graph_capture_refuses_unmarked_applysinv(gpu):
    print("graph capture guard", flush=True)
    nside, lmax = 16, 24
    holder = {}
    tm, tot, W = _template(nside, lmax, 3, ['gNL-loc'], capture_graph=True,
                           applySinv=lambda *a, **k: holder['W'].applySinv_ideal(*a, **k))
    holder['W'] = W
    rng = np.random.default_rng(4)
    sims = [(grf_alm(rng, tot, lmax, False), grf_alm(rng, tot, lmax, False))]
    tm.load_sims(lambda ii: sims[ii], 1, input_type='harmonic')
    L = 3 * nside - 1
    data = torch.from_numpy(oracle.to_map(grf_alm(rng, tot, L, False), nside, L)).to(gpu).reshape(1, -1)
    with pytest.raises(Exception, match="capturable"):
        tm.Tl_numerator(data)
