"""ps_tau_sum_lmajor_noalloc (the allocation-free variant whose partial sums live in a device
array built into the library) against ps_tau_sum_lmajor with a caller scratch, through the C ABI:
two calls in a row, the second on a smaller input, and the size limit enforced before any launch."""

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

def _case(gpu, rng, prec, Lmaxs, nr, nmu):
    Ls, _ = oracle.getlm(Lmaxs)
    nlm = len(Ls)
    cdt = torch.complex128 if prec == 64 else torch.complex64
    ld = nmu * nr + 3
    At = torch.from_numpy(rng.standard_normal((nlm, ld)) + 1j * rng.standard_normal((nlm, ld))).to(gpu, cdt)
    Bt = torch.from_numpy(rng.standard_normal((nlm, ld)) + 1j * rng.standard_normal((nlm, ld))).to(gpu, cdt)
    F = torch.from_numpy(rng.standard_normal((Lmaxs + 1, nr, nr))).to(gpu)
    w = torch.from_numpy(rng.random(nr)).to(gpu)
    return At, Bt, F, w, ld

@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_tau_sum_matches_workspace_entry_point(gpu, prec):
    from polyspec_amd._abi import load, ptr, stream_handle
    from polyspec_amd.kernels import reduce_scratch
    lib = load()
    rng = np.random.default_rng(prec)
    print(f"no-alloc tau sum prec={prec}", flush=True)
    for Lmaxs, nr, nmu, Lmin in ((47, 20, 3, 1), (23, 5, 1, 2)):  # the second call is smaller
        At, Bt, F, w, ld = _case(gpu, rng, prec, Lmaxs, nr, nmu)
        for same in (False, True):
            B = At if same else Bt
            ref = torch.full((), 0.25, dtype=torch.float64, device=gpu)
            got = torch.full((), 0.25, dtype=torch.float64, device=gpu)
            assert lib.ps_tau_sum_lmajor(prec, ptr(At), ptr(B), ld, nmu, nr, Lmaxs, Lmin, Lmaxs - 1, ptr(F), ptr(w),
                                         0, ptr(reduce_scratch(gpu)), ptr(ref), 1, stream_handle()) == 0
            assert lib.ps_tau_sum_lmajor_noalloc(prec, ptr(At), ptr(B), ld, nmu, nr, Lmaxs, Lmin, Lmaxs - 1, ptr(F),
                                                 ptr(w), 0, ptr(got), 1, stream_handle()) == 0
            torch.cuda.synchronize()
            assert float(got) == float(ref), (Lmaxs, same, float(got), float(ref))
    # beyond the built-in scratch: refused on the host, nothing launched
    out = torch.full((), 3.0, dtype=torch.float64, device=gpu)
    rc = lib.ps_tau_sum_lmajor_noalloc(prec, ptr(At), ptr(At), ld, 1, 1, 9000, 0, 8999, ptr(F), ptr(w), 0, ptr(out),
                                       0, stream_handle())
    assert rc < 0 and b"8192" in lib.ps_last_error()
    torch.cuda.synchronize()
    assert float(out) == 3.0
