"""Binned bispectrum numerator on the HIP backend (polyspec/bspec_bin.py:18-420).

Mirrors BSpecBin's interface: the bin and triangle bookkeeping (bspec_bin.py:34-176,
186-214) is the reference's logic restated; the H maps (bspec_bin.py:177-184, 216-228,
335-337) are batched spin-1 and spin-2 syntheses of all bins and fields at once on the
device; the map products summed over pixels by mp.Pool (_analyze3 / _analyze1,
bspec_bin.py:340-403) are one ps_bspec_triple_sums launch for the cubic term and one
ps_bspec_linear_sums launch (the three (data, sim) assignments together) per simulation; or,
with fused_sums or through Bl_numerator_multi (several analyses, e.g. one per resolution), one
ps_bspec_sums_multi launch with autotuned tiles in which the cubic term and the linear term of
every preloaded simulation of every analysis are segments.

H maps: for a harmonic field h (nf, nlm) and bin b with mask 1_b(l),
  (+1)H_b = sum_lm 1Y_lm 1_b conj(h)      (compute_spin_transform_map(., +1))
  (-2)H_b = sum_lm -2Y_lm 1_b conj(h)     (compute_spin_transform_map(., -2))
which with the HEALPix spin conventions of base.py:260-288 are the spin-1 / spin-2
syntheses of (G, C) = (-1_b conj(h), 0), taking Q + iU and Q - iU respectively.  The
transforms run at the largest binned l (the coefficients above it are zero).
"""
import ctypes

import numpy as np
import torch

from . import kernels as K
from ._abi import BspecGroup, BspecSegment, check, load, ptr, stream_handle

_MODE = {3: 0, 2: 1, 4: 2}  # stacks -> ps_bspec_segment mode (triple, cubic (A = B), linear)
_multi_scratch = {}
_multi_shapes = {}


def _segment_table(segments, ngroup):
    """segments: [(stacks, triplets, group)] -> ctypes array of ps_bspec_segment."""
    arr = (BspecSegment * len(segments))()
    for e, (stacks, trip, g) in zip(arr, segments):
        e.mode = _MODE[len(stacks)]
        e.nmap = int(stacks[0].shape[0])
        e.npix = int(stacks[0].shape[-1])
        for k, X in enumerate(stacks):
            assert X.dtype == stacks[0].dtype and X.shape == stacks[0].shape and X.is_contiguous()
            e.X[k] = X.data_ptr()
        e.triplets = trip.data_ptr()
        e.ntrip = int(trip.shape[0])
        assert 0 <= g < ngroup
        e.group = int(g)
    return arr


def _scratch(device, nbytes):
    key = torch.device(device).index or 0
    buf = _multi_scratch.get(key)
    if buf is None or buf.numel() < nbytes:
        _multi_scratch[key] = None
        buf = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
        _multi_scratch[key] = buf
    return buf


def multi_autotune(segments, ngroup, reps=2):
    """Time every launch shape of ps_bspec_sums_multi on these segments (partials only) and
    return (best shape id, {shape id: ms per launch})."""
    lib = load()
    prec = 32 if segments[0][0][0].dtype == torch.complex64 else 64
    arr = _segment_table(segments, ngroup)
    nb = int(lib.ps_bspec_multi_scratch_bytes(prec, len(segments), ctypes.cast(arr, ctypes.c_void_p), ngroup))
    check(min(nb, 0))
    scr = _scratch(segments[0][0][0].device, nb)
    n = lib.ps_bspec_multi_shape_count(prec)
    ms = (ctypes.c_double * n)()
    best = ctypes.c_int()
    check(lib.ps_bspec_multi_autotune(prec, len(segments), ctypes.cast(arr, ctypes.c_void_p), ngroup, ptr(scr), nb,
                                      int(reps), ms, ctypes.byref(best), stream_handle()))
    return best.value, {i: ms[i] for i in range(n)}


def multi_sums(segments, outs, shape=-1, accumulate=False):
    """One ps_bspec_sums_multi launch.  segments: [(stacks, triplets, group)], stacks being
    (A, B, C) (triple), (A, C) (cubic, A = B) or (P, Q, M, N) (linear term); outs: one
    (ntrip, 2) FP64 tensor per group, overwritten with the sum of its segments (accumulate:
    the sum is added to it)."""
    lib = load()
    prec = 32 if segments[0][0][0].dtype == torch.complex64 else 64
    ng = len(outs)
    arr = _segment_table(segments, ng)
    grp = (BspecGroup * ng)()
    for g, o in zip(grp, outs):
        assert o.dtype == torch.float64 and o.is_contiguous()
        g.out, g.scale, g.accumulate = o.data_ptr(), 1.0, int(bool(accumulate))
    nb = int(lib.ps_bspec_multi_scratch_bytes(prec, len(segments), ctypes.cast(arr, ctypes.c_void_p), ng))
    check(min(nb, 0))
    scr = _scratch(segments[0][0][0].device, nb)
    check(lib.ps_bspec_sums_multi(prec, len(segments), ctypes.cast(arr, ctypes.c_void_p), ng,
                                  ctypes.cast(grp, ctypes.c_void_p), int(shape), ptr(scr), nb, stream_handle()))


def _tuned_shape(segments, ngroup, enabled):
    """Launch shape for these segments: autotuned once per set of segment kinds (stack count,
    stack shape and dtype, triplet count), on one segment of each kind -- the tuning run does not
    depend on how many segments of each kind a launch holds, so neither does the key (rounds of
    streamed simulations of different sizes share one tuning)."""
    if not enabled:
        return -1
    kinds = sorted(set((len(st), tuple(st[0].shape), str(st[0].dtype), int(tr.shape[0])) for st, tr, _ in segments))
    key = tuple(kinds)
    if key not in _multi_shapes:
        # a launch whose kinds are a subset of a tuned launch's (the later rounds of streamed
        # simulations, once one resolution has run out) takes that launch's shape
        for k, v in _multi_shapes.items():
            if set(key) <= set(k):
                return v[0]
        sub, seen = [], set()
        for st, tr, g in segments:
            k = (len(st), tuple(st[0].shape), str(st[0].dtype), int(tr.shape[0]))
            if k not in seen:
                seen.add(k)
                sub.append((st, tr, len(sub)))
        _multi_shapes[key] = multi_autotune(sub, len(sub))
    return _multi_shapes[key][0]


def _numerator_segments(bs, data, include_linear_term, g_cub, g_lin):
    """S^-1, H maps and the sum segments of one analysis (groups g_cub: cubic, g_lin: linear):
    the cubic term and the linear term of every preloaded simulation.  Returns (segments,
    (p1, m2)), the data's H maps being what streamed simulations pair with."""
    h = bs.applySinv(data, input_type='map')
    p1, m2 = bs.H_maps(h)
    segs = [((p1, m2), bs._trip, g_cub)]
    if include_linear_term and bs.preload:
        for ii in range(bs.N_it):
            segs.append(((p1, bs.p1_H_maps[ii], m2, bs.m2_H_maps[ii]), bs._trip, g_lin))
    return segs, (p1, m2)


def _stream_rounds(streamed, budget):
    """Segments of the streamed simulations in rounds: each round takes the next simulation of
    every analysis in turn (round-robin over resolutions) until the round's H maps would exceed
    `budget` bytes (at least one simulation per round), then makes the H maps of each analysis's
    share of the round (BSpecBin._load_sims_data: one batched S^-1 for them where S^-1 allows)."""
    nxt = [0] * len(streamed)
    while any(nxt[k] < st[0].N_it for k, st in enumerate(streamed)):
        take, used, added = [0] * len(streamed), 0, True
        while added:
            added = False
            for k, (bs, p1, m2, g) in enumerate(streamed):
                per = 2 * p1.numel() * p1.element_size()
                if nxt[k] + take[k] >= bs.N_it or (used and used + per > budget):
                    continue
                take[k] += 1
                used += per
                added = True
        segs = []
        for k, (bs, p1, m2, g) in enumerate(streamed):
            idx = list(range(nxt[k], nxt[k] + take[k]))
            nxt[k] += take[k]
            for sp1, sm2 in (bs._load_sims_data(idx) if idx else []):
                segs.append(((p1, sp1, m2, sm2), bs._trip, g))
        yield segs


def Bl_numerator_multi(analyses, datas, include_linear_term=True, sim_memory_gb=None):
    """Numerators of several BSpecBin analyses (e.g. one per resolution) with every pixel
    reduction -- the cubic term of each data map and the linear term of every simulation of
    every analysis -- in ps_bspec_sums_multi launches: one launch for the cubic terms and the
    preloaded simulations, and, for analyses whose simulations are streamed (load_sims /
    generate_sims with preload=False, as a simulation set too large for device memory needs),
    one launch per round of simulations whose H maps fit `sim_memory_gb` (default: half the
    free device memory), each adding to the previous ones.  Returns the list of numerators,
    each as BSpecBin.Bl_numerator returns it."""
    assert len(analyses) == len(datas) and len(analyses) >= 1
    segs, outs, gids, streamed = [], [], [], []
    for bs, data in zip(analyses, datas):
        if not hasattr(bs, 'sym_factor'):
            bs._compute_symmetry_factor()
        if include_linear_term and not hasattr(bs, 'preload'):
            raise Exception("Need to generate or specify bias simulations!")
        ntrip = bs._trip.shape[0]
        g_cub = len(outs)
        g_lin = g_cub + 1 if include_linear_term else -1
        outs += [torch.zeros((ntrip, 2), dtype=torch.float64, device=bs.base.device)
                 for _ in range(2 if include_linear_term else 1)]
        gids.append((g_cub, g_lin))
        s_, (p1, m2) = _numerator_segments(bs, data, include_linear_term, g_cub, g_lin)
        segs += s_
        if include_linear_term and not bs.preload:
            streamed.append((bs, p1, m2, g_lin))
    tune = all(bs.base.autotune for bs in analyses)
    ev = analyses[0].kernel_events

    def launch(sg, acc):
        # the launch's groups: those its segments add to (a group needs at least one segment)
        used = sorted({g for _, _, g in sg})
        remap = {g: i for i, g in enumerate(used)}
        sg = [(st, tr, remap[g]) for st, tr, g in sg]
        shape = _tuned_shape(sg, len(used), tune)
        if ev is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        multi_sums(sg, [outs[g] for g in used], shape, accumulate=acc)
        if ev is not None:
            e1.record()
            ev.append(('multi', e0, e1))

    launch(segs, False)
    if streamed:
        if sim_memory_gb is None:
            free, _ = torch.cuda.mem_get_info(analyses[0].base.device)
            budget = 0.5 * free
        else:
            budget = float(sim_memory_gb) * 2 ** 30
        for rs in _stream_rounds(streamed, budget):
            launch(rs, True)
            del rs
    res = []
    for bs, (g_cub, g_lin) in zip(analyses, gids):
        b3 = bs.base.A_pix * bs._assemble(outs[g_cub]) / 3.
        if include_linear_term:
            b1 = -bs.base.A_pix * bs._assemble(outs[g_lin]) / 3. / bs.N_it
        else:
            print("No linear correction applied!")
            b1 = np.zeros(bs.N_b)
        res.append(b3 + b1)
    return res


class BSpecBin:
    """Binned bispectrum estimation (bspec_bin.py:18-33): base, mask, applySinv, l_bins, ..."""

    def __init__(self, base, mask, applySinv, l_bins, l_bins_squeeze=[], include_partial_triangles=False,
                 fields=['TTT', 'TTE', 'TEE', 'EEE', 'TBB', 'EBB', 'TTB', 'TEB', 'EEB', 'BBB'], parity='both',
                 batch_maps=True):
        """batch_maps=True: every (bin, field) H map of a field in one spin-1 and one spin-2
        synthesis launch.  batch_maps=False: one transform per (bin, field), as the reference's
        compute_spin_transform_map loop (bspec_bin.py:336-337) issues them."""
        self.base = base
        self.batch_maps = batch_maps
        self.kernel_events = None  # set to a list to time every triple-sum launch
        # fused_sums: with preloaded simulations (or no linear term) the cubic term and every
        # simulation's linear term are segments of one ps_bspec_sums_multi launch.  Off by
        # default: for one analysis the per-simulation launches measured 9% faster (Nside 256,
        # 16 simulations); Bl_numerator_multi, which joins analyses of several resolutions into
        # one launch, always uses it.
        self.fused_sums = False
        # streamed simulations filtered together by a batched S^-1 (Bl_numerator_multi's rounds):
        # at most this many per batch, which bounds the batch's maps and transform workspace
        self.sinv_batch = 16
        self.mask = mask
        self.applySinv = applySinv
        self.l_bins = np.asarray(l_bins)
        self.pol = self.base.pol
        self.parity = parity
        self.fields = list(fields)
        self.min_l = np.min(l_bins)
        self.Nl = len(l_bins) - 1
        if len(l_bins_squeeze) > 0:
            self.l_bins_squeeze = np.asarray(l_bins_squeeze)
            self.Nl_squeeze = len(l_bins_squeeze) - 1
            for l_edge in self.l_bins:
                assert l_edge in self.l_bins_squeeze, "Squeezed bins must contain all the unsqueezed bins!"
        else:
            self.l_bins_squeeze = self.l_bins.copy()
            self.Nl_squeeze = self.Nl
        self.include_partial_triangles = include_partial_triangles
        self.beam = self.base.beam
        self.beam_lm = self.base.beam_lm
        for f in self.fields:
            assert f in ['TTT', 'TTE', 'TEE', 'EEE', 'TBB', 'EBB', 'TTB', 'TEB', 'EEB', 'BBB'], \
                "Unknown field '%s' supplied!" % f
        assert len(self.fields) == len(np.unique(self.fields)), "Duplicate fields supplied!"
        if not self.pol and self.fields != ['TTT']:
            print("## Polarization mode not turned on; setting fields to TT only!")
            self.fields = ['TTT']
        assert parity in ['even', 'odd', 'both'], "Parity must be one of 'even', 'odd' or 'both'!"
        if not self.pol and parity != 'even':
            print("Caution: scalar 3-point functions can't probe parity-odd primordial physics!")
        self.chi_arr = [1] if parity == 'even' else [-1] if parity == 'odd' else [1, -1]
        if np.max(self.l_bins_squeeze) > base.lmax:
            raise Exception("Maximum l is larger than HEALPix resolution!")
        if np.max(self.l_bins) > base.lmax // 2:
            print("## Caution: Maximum l is greater than HEALPix-lmax/2; this might cause boundary effects.")
        print("Binning: %d bins in [%d, %d]" % (self.Nl, self.min_l, np.max(self.l_bins)))
        if self.Nl_squeeze != self.Nl:
            print("Squeezed binning: %d bins in [%d, %d]" % (self.Nl_squeeze, self.min_l, np.max(self.l_bins_squeeze)))
        self.ones_mask = bool(np.std(self.mask) < 1e-12 and np.abs(np.mean(self.mask) - 1) < 1e-12)
        print("Mask: ones" if self.ones_mask else "Mask: spatially varying")

        # list of bins (bspec_bin.py:92-112)
        self.all_bins, self.good_bins = [], []
        for u in self.fields:
            u1, u2, u3 = [self.base.indices[u[i]] for i in range(3)]
            p_u = int(np.prod([self.base.parities[u[i]] for i in range(3)]))
            for bin1 in range(self.Nl_squeeze):
                for bin2 in range(self.Nl_squeeze):
                    if u1 == u2 and bin2 < bin1:
                        continue
                    for bin3 in range(self.Nl_squeeze):
                        if u2 == u3 and bin3 < bin2:
                            continue
                        if not self._check_bin(bin1, bin2, bin3):
                            continue
                        if min([bin1, bin2, bin3]) >= self.Nl:
                            continue
                        self.all_bins.append([u1, u2, u3, bin1, bin2, bin3, p_u])
                        self.good_bins.append(np.asarray([bin1, bin2, bin3]))
        self.all_bins = np.asarray(self.all_bins, dtype=np.int64).reshape(-1, 7)
        gb = np.asarray(self.good_bins).reshape(-1, 3)
        self.is_good = lambda bins: bool(np.any((gb[:, 0] == bins[0]) & (gb[:, 1] == bins[1]) & (gb[:, 2] == bins[2])))

        # device set-up: H-map transforms at the largest binned l
        self.nf = 1 + 2 * self.pol
        self.lmax_b = int(np.max(self.l_bins_squeeze)) - 1
        L = self.base.lmax
        ls = self.base.l_arr
        dev = self.base.device
        self.ell_bins = [(ls >= self.l_bins_squeeze[b]) & (ls < self.l_bins_squeeze[b + 1])
                         for b in range(self.Nl_squeeze)]
        sel = np.concatenate([np.arange(m * (2 * L + 1 - m) // 2 + m, m * (2 * L + 1 - m) // 2 + self.lmax_b + 1)
                              for m in range(self.lmax_b + 1)])
        self._sel = torch.from_numpy(sel).to(dev)
        lb = ls[sel]
        binmask = np.stack([((lb >= self.l_bins_squeeze[b]) & (lb < self.l_bins_squeeze[b + 1])).astype(float)
                            for b in range(self.Nl_squeeze)])
        self._binmask = torch.from_numpy(binmask).to(dev, self.base.rdtype)  # (Nb, nlm_b)
        # map-stack index of (bin, field) = bin * nf + field
        ab = self.all_bins
        trip = np.stack([ab[:, 3] * self.nf + ab[:, 0], ab[:, 4] * self.nf + ab[:, 1], ab[:, 5] * self.nf + ab[:, 2]], 1)
        self._trip = torch.from_numpy(np.ascontiguousarray(trip.astype(np.int32))).to(dev)
        self.nmap = self.Nl_squeeze * self.nf
        assert self.nmap <= 128, "at most 128 (bin, field) maps per stack"
        self._scratch = None

    # ------------------------------------------------------ bookkeeping
    def _check_bin(self, bin1, bin2, bin3):
        """Triangle condition on all modes of the bin or on its centre (bspec_bin.py:126-147)."""
        if self.include_partial_triangles:
            l1s = np.arange(self.l_bins_squeeze[bin1], self.l_bins_squeeze[bin1 + 1])[:, None, None]
            l2s = np.arange(self.l_bins_squeeze[bin2], self.l_bins_squeeze[bin2 + 1])[None, :, None]
            l3s = np.arange(self.l_bins_squeeze[bin3], self.l_bins_squeeze[bin3 + 1])[None, None, :]
            filt = (l3s >= abs(l1s - l2s)) * (l3s <= l1s + l2s)
            return 1 if np.sum(filt) > 0 else 0
        l1 = 0.5 * (self.l_bins_squeeze[bin1] + self.l_bins_squeeze[bin1 + 1])
        l2 = 0.5 * (self.l_bins_squeeze[bin2] + self.l_bins_squeeze[bin2 + 1])
        l3 = 0.5 * (self.l_bins_squeeze[bin3] + self.l_bins_squeeze[bin3 + 1])
        return 0 if (l3 < abs(l1 - l2) or l3 > l1 + l2) else 1

    def _compute_symmetry_factor(self):
        """Degeneracy of each bin, stacked over parities (bspec_bin.py:149-175)."""
        self.sym_factor_all = []
        for chi in self.chi_arr:
            sym = []
            for u1, u2, u3, bin1, bin2, bin3, p_u in self.all_bins:
                if bin1 == bin2 and u1 == u2 and bin2 == bin3 and u2 == u3:
                    sym.append(6)
                elif bin1 == bin2 and u1 == u2:
                    sym.append(2)
                elif bin2 == bin3 and u2 == u3:
                    sym.append(2)
                else:
                    sym.append(1)
            self.sym_factor_all.append(np.asarray(sym))
        self.sym_factor = np.concatenate(self.sym_factor_all)
        self.N_b = len(self.sym_factor)
        print("Using a total of %d bins" % self.N_b)

    def get_ells(self, field='TTT'):
        """Central (l1, l2, l3) of each bin of one field (bspec_bin.py:186-214)."""
        assert field in self.fields, "Specified field '%s' not in inputs!" % field
        u1, u2, u3 = [self.base.indices[field[i]] for i in range(3)]
        l1s, l2s, l3s = [], [], []
        for bin1 in range(self.Nl_squeeze):
            l1 = 0.5 * (self.l_bins_squeeze[bin1] + self.l_bins_squeeze[bin1 + 1])
            for bin2 in range(self.Nl_squeeze):
                if u1 == u2 and bin2 < bin1:
                    continue
                l2 = 0.5 * (self.l_bins_squeeze[bin2] + self.l_bins_squeeze[bin2 + 1])
                for bin3 in range(self.Nl_squeeze):
                    if u2 == u3 and bin3 < bin2:
                        continue
                    l3 = 0.5 * (self.l_bins_squeeze[bin3] + self.l_bins_squeeze[bin3 + 1])
                    if not self.is_good([bin1, bin2, bin3]):
                        continue
                    if min([bin1, bin2, bin3]) >= self.Nl:
                        continue
                    l1s.append(l1)
                    l2s.append(l2)
                    l3s.append(l3)
        return l1s, l2s, l3s

    # ------------------------------------------------------------ H maps
    def H_maps(self, h_lm):
        """(+1)H and (-2)H maps of a harmonic field (nf, nlm(base lmax)) for every bin:
        two (Nb * nf, Npix) complex stacks, row bin * nf + field."""
        h = h_lm if isinstance(h_lm, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(h_lm))
        h = h.to(self.base.device, self.base.cdtype).reshape(self.nf, -1)
        hb = torch.conj(h.index_select(-1, self._sel))  # (nf, nlm_b)
        Nb, nf = self.Nl_squeeze, self.nf
        G = -(self._binmask[:, None, :] * hb[None, :, :]).reshape(Nb * nf, 1, -1)
        P1, P2 = self.base.plan(self.lmax_b, 1), self.base.plan(self.lmax_b, 2)
        if self.batch_maps:
            # C = 0: the gradient-only syntheses skip the curl half of the Legendre work
            G2 = G.reshape(Nb * nf, -1).contiguous()
            QU1 = P1.synthesis_grad(G2)
            QU2 = P2.synthesis_grad(G2)
        else:
            GC = torch.cat([G, torch.zeros_like(G)], dim=1).contiguous()  # (Nb nf, 2, nlm_b)
            QU1 = torch.empty((Nb * nf, 2, P1.npix), dtype=self.base.rdtype, device=GC.device)
            QU2 = torch.empty_like(QU1)
            for k in range(Nb * nf):
                P1.synthesis(GC[k:k + 1], out=QU1[k:k + 1])
                P2.synthesis(GC[k:k + 1], out=QU2[k:k + 1])
        self.base.n_SHTs_reverse += 2 * Nb * nf
        p1 = torch.complex(QU1[:, 0], QU1[:, 1]).contiguous()
        m2 = torch.complex(QU2[:, 0], -QU2[:, 1]).contiguous()
        return p1, m2

    def _process_sim(self, sim, input_type='map'):
        """S^-1 then H maps of one simulation (bspec_bin.py:216-228)."""
        return self.H_maps(self.applySinv(sim, input_type=input_type))

    def _raw_sim(self, ii):
        """Simulation ii of the streamed set and its input type ('map' or 'harmonic')."""
        fn, kind = self._sim_loader
        return fn(ii) if kind is None else (fn(ii), kind)

    def load_sim_data(self, ii):
        """H maps (p1, m2) of streamed simulation ii (load_sims / generate_sims, preload=False)."""
        return self._process_sim(*self._raw_sim(ii))

    def _sinv_batch(self):
        """The batched form of applySinv, when it is Weightings.applySinv_ideal; else None."""
        W = getattr(self.applySinv, '__self__', None)
        fn = getattr(type(W), 'applySinv_ideal', None)
        if fn is None or getattr(self.applySinv, '__func__', None) is not fn:
            return None
        return W.applySinv_ideal_batch

    def _load_sims_data(self, idx):
        """load_sim_data of the streamed simulations idx.  With S^-1 = Weightings.applySinv_ideal
        and pixel-space simulations, the filter of up to sinv_batch of them runs as one batched
        adjoint per spin (matrix-core kernels) rather than one single-map adjoint pair each, with the
        built-in launch shapes (no tuning pass for a new batch width inside a streamed run); the H
        maps are made per simulation.  A load_sim_data set on the instance is called per index."""
        batch = self._sinv_batch() if len(idx) > 1 and 'load_sim_data' not in self.__dict__ else None
        if batch is None:
            return [self.load_sim_data(i) for i in idx]
        out = []
        for c in range(0, len(idx), self.sinv_batch):
            raw = [self._raw_sim(i) for i in idx[c:c + self.sinv_batch]]
            if len(raw) == 1 or any(kind != 'map' for _, kind in raw):
                out += [self._process_sim(m, input_type=kind) for m, kind in raw]
                continue
            maps = torch.stack([torch.as_tensor(m).to(self.base.device, self.base.rdtype).reshape(-1, self.base.Npix)
                                for m, _ in raw])
            del raw
            h = batch(maps, tune=False)
            del maps
            out += [self.H_maps(h[j]) for j in range(h.shape[0])]
        return out

    def _sums(self, stacks, scale, out, accumulate):
        """ps_bspec_triple_sums on (A, B, C) = stacks, or ps_bspec_linear_sums on (P, Q, M, N)."""
        ntrip = self._trip.shape[0]
        need = int(load().ps_bspec_scratch_bytes(ntrip))
        if self._scratch is None or self._scratch.numel() < need:
            self._scratch = torch.empty(need, dtype=torch.uint8, device=self.base.device)
        A = stacks[0]
        for X in stacks:
            assert X.dtype == A.dtype and X.shape == (self.nmap, A.shape[-1]) and X.is_contiguous()
        prec = 32 if A.dtype == torch.complex64 else 64
        if self.kernel_events is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        fn = load().ps_bspec_triple_sums if len(stacks) == 3 else load().ps_bspec_linear_sums
        check(fn(prec, *[ptr(X) for X in stacks], self.nmap, A.shape[-1], ptr(self._trip), ntrip, float(scale),
                 ptr(self._scratch), need, ptr(out), int(accumulate), stream_handle()))
        if self.kernel_events is not None:
            e1.record()
            self.kernel_events.append((len(stacks), e0, e1))

    def _assemble(self, S):
        """Per parity: Re S, or Re(i S) = -Im S when p_u chi = -1 (bspec_bin.py:358-364)."""
        S = S.cpu().numpy()
        out = []
        for chi in self.chi_arr:
            odd = self.all_bins[:, 6] * chi == -1
            out.append(np.where(odd, -S[:, 1], S[:, 0]))
        return np.concatenate(out) / self.sym_factor

    # ----------------------------------------------------- simulations
    def load_sims(self, load_sim, N_sims, verb=False, input_type='map', preload=True):
        """Monte-Carlo simulations for the linear term (bspec_bin.py:230-265)."""
        self.N_it = N_sims
        print("Using %d Monte Carlo simulations" % self.N_it)
        self.preload = preload
        if preload:
            self.p1_H_maps, self.m2_H_maps = [], []
            for ii in range(self.N_it):
                if verb:
                    print("Loading bias simulation %d of %d" % (ii + 1, self.N_it))
                p1, m2 = self._process_sim(load_sim(ii), input_type=input_type)
                self.p1_H_maps.append(p1)
                self.m2_H_maps.append(m2)
        else:
            if verb:
                print("No preloading; simulations will be loaded and accessed at runtime.")
            self.__dict__.pop('load_sim_data', None)
            self._sim_loader = (load_sim, input_type)

    def generate_sims(self, N_sim, Cl_input=[], b_input=None, add_B=False, remove_mean=True, verb=False,
                      preload=True):
        """Gaussian simulations with the survey mask (bspec_bin.py:267-315; add_B is not provided)."""
        if add_B or b_input is not None:
            raise NotImplementedError("generate_sims(add_B=True) is not provided by the hip backend")
        if len(Cl_input) == 0:
            Cl_input = self.base.Cl_tot
        mask_d = torch.as_tensor(np.asarray(self.mask)).to(self.base.device, self.base.rdtype)
        base, ones_mask = self.base, self.ones_mask  # not self: the loader is stored on self

        def sim(ii):
            if ones_mask:
                return base.generate_data(int(1e5) + ii, Cl_input=Cl_input, deconvolve_beam=False,
                                          remove_mean=remove_mean, output_type='harmonic'), 'harmonic'
            m = base.generate_data(int(1e5) + ii, Cl_input=Cl_input, deconvolve_beam=False,
                                   remove_mean=remove_mean)
            m = torch.as_tensor(m).to(base.device, base.rdtype)
            return (mask_d * m).contiguous(), 'map'

        self.N_it = N_sim
        print("Using %d Monte Carlo simulations" % self.N_it)
        self.preload = preload
        if preload:
            self.p1_H_maps, self.m2_H_maps = [], []
            for ii in range(self.N_it):
                if verb:
                    print("Generating bias simulation %d of %d" % (ii + 1, self.N_it))
                s, kind = sim(ii)
                p1, m2 = self._process_sim(s, input_type=kind)
                self.p1_H_maps.append(p1)
                self.m2_H_maps.append(m2)
        else:
            self.__dict__.pop('load_sim_data', None)
            self._sim_loader = (sim, None)

    # -------------------------------------------------------- estimator
    def Bl_numerator(self, data, include_linear_term=True, verb=False):
        """Unwindowed binned-bispectrum numerator (bspec_bin.py:317-420): cubic term and, with
        simulations, the linear term, ordered [parity][bin] as the reference's."""
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        if (not hasattr(self, 'preload')) and include_linear_term:
            raise Exception("Need to generate or specify bias simulations!")
        if self.fused_sums and (not include_linear_term or self.preload):
            # cubic term and every simulation's linear term in one launch
            return Bl_numerator_multi([self], [data], include_linear_term)[0]
        h = self.applySinv(data, input_type='map')
        if verb:
            print("Computing H maps")
        p1, m2 = self.H_maps(h)
        ntrip = self._trip.shape[0]
        S3 = torch.zeros((ntrip, 2), dtype=torch.float64, device=self.base.device)
        self._sums((p1, p1, m2), 1.0, S3, 0)
        b3 = self.base.A_pix * self._assemble(S3) / 3.
        b1 = np.zeros(self.N_b)
        if not include_linear_term:
            print("No linear correction applied!")
        else:
            S1 = torch.zeros_like(S3)
            for ii in range(self.N_it):
                if (ii + 1) % 5 == 0 and verb:
                    print("Computing b_1 piece from simulation %d" % (ii + 1))
                if self.preload:
                    sp1, sm2 = self.p1_H_maps[ii], self.m2_H_maps[ii]
                else:
                    sp1, sm2 = self.load_sim_data(ii)
                # (p1, sp1, sm2) + (sp1, p1, sm2) + (sp1, sp1, m2) in one pass
                self._sums((p1, sp1, m2, sm2), 1.0, S1, 1)
            b1 = -self.base.A_pix * self._assemble(S1) / 3. / self.N_it
        return b3 + b1
