"""The template library at its full size on the GPU: the 40-atom orthorhombic
cell, r = 2.0 1/A, 200,000 orientations of the super-Fibonacci SO(3) grid,
256x256 fast patterns at 0.015 1/A/px, sigma 1 px, normalised.

The batch takes the band renderer (plan kernel + band writer), as the library
benchmark does.  Checked at full size: every render status, the spot-list
capacity, finite values in [0, 1] and every pattern's maximum (1, or 0 for a
pattern with nothing in frame); bit equality with the splat renderer
(dsx_render_fast) over the first, a middle and the last 4,096 patterns; and
48 orientations spread over the batch against the FP64 oracle (spot sets,
intensities and patterns, tests/parity.py rules).
"""

import ctypes

import numpy as np
import pytest

from oracle import diffsims_oracle as O
from tests import parity as P

pytestmark = pytest.mark.gpu


def test_full_library_through_the_band_renderer():
    import torch

    from diffsims_amd import _lib, engine, workloads

    print("library: 200,000 x 256x256, band renderer", flush=True)
    dev = torch.device("cuda", torch.cuda.current_device())
    wl = workloads.template_library(200_000)
    rr = wl.reciprocal_radius[0]
    H, W = wl.shape
    tab = engine.build_phase_table(wl.phases[0], rr, wl.with_direct_beam, "lobato", device=dev)
    Rt = engine.rotation_matrices_to_device(wl.rotations[0], dev)
    n = int(Rt.shape[0])
    assert n == 200_000
    assert engine.fast_renderer(n, H, W) == "dsx_render_fast_bands"
    probe = engine.simulate(tab, Rt[:4096], wl.settings, rr)
    cap = int(max(probe.cap, int(probe.ncand.max()) + 64))
    batch = engine.empty_batch(tab, n, cap, dev)
    engine.simulate_into(tab, Rt, wl.settings, batch)
    need = int(batch.ncand.max())
    if need > cap:
        cap = need + 64
        batch = engine.empty_batch(tab, n, cap, dev)
        engine.simulate_into(tab, Rt, wl.settings, batch)
    assert int(batch.ncand.max()) <= cap and int(batch.count.max()) <= cap
    out = torch.full((n, H, W), float("nan"), dtype=torch.float32, device=dev)
    status = torch.full((n,), -1, dtype=torch.int32, device=dev)
    engine.render_spots(batch.spots, batch.count, cap, wl.shape, wl.sigma, wl.calibration,
                        out=out, status=status)
    engine.check_render_status(status)
    assert int(status.min()) == 0

    # whole-batch properties
    assert not bool(torch.isnan(out).any())
    assert float(out.amin()) >= 0.0
    mx = out.amax(dim=(1, 2))
    empty = mx == 0
    assert bool((((mx - 1.0).abs() <= 1e-6) | empty).all())
    assert int(empty.sum()) <= n // 100

    # the band renderer equals the splat renderer bit for bit
    L = _lib.load(require_gpu=True)
    p = engine.render_params(wl.shape, wl.sigma, wl.calibration)
    stream = _lib.stream_handle(dev)
    m = 4096
    ref = torch.empty((m, H, W), dtype=torch.float32, device=dev)
    st = torch.full((m,), -1, dtype=torch.int32, device=dev)
    for o in (0, n // 2, n - m):
        print(f"library: splat renderer on patterns {o}..{o + m - 1}", flush=True)
        _lib.check(L.dsx_render_fast(_lib.ptr(batch.spots[o:o + m]), _lib.ptr(batch.count[o:o + m]),
                                     cap, m, ctypes.byref(p), _lib.ptr(ref), _lib.ptr(st), stream),
                   "dsx_render_fast")
        torch.cuda.synchronize()
        assert int(st.max()) == 0
        assert torch.equal(ref, out[o:o + m]), o

    # a spread sample against the FP64 oracle
    print("library: 48 orientations against the oracle", flush=True)
    sel = np.unique(np.concatenate([np.linspace(0, n - 1, 40).round(),
                                    np.random.default_rng(7).integers(0, n, 8)]).astype(np.int64))
    t = torch.as_tensor(sel, device=dev)
    idx = batch.idx[t].cpu().numpy()
    sp = batch.spots[t].cpu().numpy()
    cnt = batch.count[t].cpu().numpy()
    pats = out[t].cpu().numpy()
    p_or = P.oracle_params(wl.settings, rr, wl.with_direct_beam)
    otab = O.phase_tables(P.phase_spec(wl.phases[0]), p_or)
    assert np.array_equal(tab.hkl_host(), otab.hkl)
    k = 1 / O.electron_wavelength(p_or.kv)
    stats = {}
    worst = 0.0
    for j, o in enumerate(sel):
        c = int(cnt[j])
        ref_o, _ = P.check_orientation(otab, wl.rotations[0][o], k, p_or, idx[j, :c], sp[j, :c, 3],
                                       stats)
        rp = P.flipped_reference_pattern(ref_o, sp[j, :c, :2], wl.shape, wl.sigma, wl.calibration)
        worst = max(worst, P.pattern_error(pats[j], rp))
    assert worst <= P.TOL_P, worst
    assert stats.get("carved_orientations", 0) <= 2, stats
