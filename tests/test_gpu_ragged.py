"""Ragged multi-phase batching (one persistent simulate launch, one render
launch per method, hipGraph replay) against the single-phase path and the
FP64 oracle (GPU).

Parity: per phase the ragged spot lists equal the single-phase kernel's bit
for bit (same kernel code, same arithmetic); splat-rendered patterns equal
dsx_render_fast's bit for bit; dense-rendered patterns equal it to 1e-6 of
the pattern maximum (different summation order) and the FP64 oracle to the
usual 1e-5 with the flipped-pixel carve-out.
"""

import numpy as np
import pytest
import torch

from oracle import diffsims_oracle as O
from tests import parity as P

pytestmark = pytest.mark.gpu


def _phases():
    """Six phases of very different size, a few with > 1024 in-frame spots."""
    specs = [P.al_c0(), O.PhaseSpec((5.431,) * 3, (90.0,) * 3, ["Si"] * 8, P.si_positions(),
                                     np.ones(8), np.full(8, 0.4))]
    rng = np.random.default_rng(5)
    for i, (n_atoms, edges) in enumerate([(3, 4.0), (12, 9.0), (40, 14.0), (7, 22.0)]):
        abc = tuple(float(v) for v in rng.uniform(0.8, 1.2, 3) * edges)
        ang = tuple(float(v) for v in rng.uniform(80, 110, 3))
        syms = [["O", "Si", "Fe", "Mg", "Cu"][j] for j in rng.integers(0, 5, n_atoms)]
        specs.append(O.PhaseSpec(abc, ang, syms, rng.uniform(0, 1, (n_atoms, 3)), np.ones(n_atoms),
                                 rng.uniform(0.2, 1.5, n_atoms)))
    return specs


def _library(shape=(128, 128), cal=0.03, n_orient=(37, 5, 64, 23, 41, 30), radii=None, sigma=1.0):
    from diffsims_amd import engine
    from diffsims_amd.ragged import RaggedLibrary

    specs = _phases()
    radii = radii or [1.0, 1.4, 1.8, 1.2, 1.6, 1.9]
    rots = [O.bunge_matrix(O.random_euler(n, 100 + i)) for i, n in enumerate(n_orient)]
    st = engine.SimSettings(accelerating_voltage=200, max_excitation_error=0.02,
                            shape_model="lorentzian", minimum_intensity=1e-20)
    lib = RaggedLibrary([P.to_arrays(s, f"p{i}") for i, s in enumerate(specs)], rots, radii, st,
                        shape, sigma, cal)
    lib.run()
    lib.check()
    torch.cuda.synchronize()
    return lib, specs, rots, radii, st


def test_ragged_equals_single_phase_path():
    from diffsims_amd import _lib, engine

    print("ragged vs single-phase", flush=True)
    lib, specs, rots, radii, st = _library()
    methods = [s.method for s in lib.slots]
    print("  render methods per phase:", methods, flush=True)
    for i, s in enumerate(lib.slots):
        idx, sp, cnt = lib.phase_spots(i)
        b = engine.simulate(s.table, engine.rotation_matrices_to_device(rots[i]), st, radii[i])
        c = cnt.cpu().numpy()
        assert np.array_equal(c, b.count.cpu().numpy()), i
        for o in range(s.n):
            assert np.array_equal(idx[o, :c[o]].cpu().numpy(), b.idx[o, :c[o]].cpu().numpy())
            assert torch.equal(sp[o, :c[o]], b.spots[o, :c[o]])
        ref = engine.render(b, lib.shape, 1.0, lib.calibration, status=None) \
            if b.count.max() <= _lib.load().dsx_render_max_spots() else None
        got = lib.phase_patterns(i)
        if ref is None:
            continue                                  # the splat path cannot hold it
        if s.method == _lib.RENDER_SPLAT:
            assert torch.equal(got, ref), i
        else:
            assert float((got - ref).abs().max()) <= 1e-6, i


def test_dense_render_matches_oracle():
    """Phases rendered densely (many spots) against the FP64 oracle."""
    from diffsims_amd import _lib

    print("ragged dense vs oracle", flush=True)
    lib, specs, rots, radii, st = _library()
    k = 1 / O.electron_wavelength(200)
    checked = 0
    for i, s in enumerate(lib.slots):
        p = O.SimParams(kv=200, reciprocal_radius=radii[i], max_excitation_error=0.02,
                        shape_model="lorentzian", with_direct_beam=False)
        otab = O.phase_tables(specs[i], p)
        idx, sp, cnt = lib.phase_spots(i)
        pats = lib.phase_patterns(i).cpu().numpy()
        for o in range(0, s.n, max(1, s.n // 6)):
            c = int(cnt[o])
            ref, _ = P.check_orientation(otab, rots[i][o], k, p, idx[o, :c].cpu().numpy(),
                                         sp[o, :c, 3].cpu().numpy())
            rp = P.flipped_reference_pattern(ref, sp[o, :c, :2].cpu().numpy(), lib.shape, 1.0,
                                             lib.calibration)
            err = P.pattern_error(pats[o], rp)
            assert err <= P.TOL_P, (i, o, err)
            checked += 1
    assert checked >= 20
    assert any(s.method in (_lib.RENDER_DENSE, _lib.RENDER_DENSE_INPLACE) for s in lib.slots)


def test_dense_kernel_many_spots():
    """Force the dense method on every phase (the 22 A cell has thousands of
    spots per pattern) and compare with the oracle."""
    from diffsims_amd import _lib

    print("dense renderer forced", flush=True)
    lib, specs, rots, radii, st = _library(cal=0.03)
    for s in lib.slots:
        s.method = _lib.RENDER_DENSE
    lib.render_bpc[_lib.RENDER_DENSE] = 1
    lib._lists()
    lib.out.zero_()
    lib.run()
    lib.check()
    i = 5
    p = O.SimParams(kv=200, reciprocal_radius=radii[i], max_excitation_error=0.02,
                    shape_model="lorentzian", with_direct_beam=False)
    otab = O.phase_tables(specs[i], p)
    idx, sp, cnt = lib.phase_spots(i)
    assert int(cnt.max()) > 1024
    pats = lib.phase_patterns(i).cpu().numpy()
    k = 1 / O.electron_wavelength(200)
    stats = {}
    for o in range(lib.slots[i].n):
        c = int(cnt[o])
        ref, _ = P.check_orientation(otab, rots[i][o], k, p, idx[o, :c].cpu().numpy(),
                                     sp[o, :c, 3].cpu().numpy(), stats)
        rp = P.flipped_reference_pattern(ref, sp[o, :c, :2].cpu().numpy(), lib.shape, 1.0,
                                         lib.calibration)
        assert P.pattern_error(pats[o], rp) <= P.TOL_P, o
    assert stats["carved_orientations"] <= 2, stats


@pytest.mark.parametrize("sigma", [0.3, 0.5, 0.7, 1.0, 1.2])
def test_inplace_dense_equals_two_buffer_dense(sigma):
    """The in-place dense kernel (one LDS buffer, two patterns per CU) gives
    the two-buffer dense kernel's bits on every pattern, tap radii 1..5."""
    from diffsims_amd import _lib

    print(f"in-place dense vs two-buffer dense, sigma {sigma}", flush=True)
    lib, *_ = _library(cal=0.03, sigma=sigma)
    assert lib.L.dsx_render_dense_inplace_fits(128, 128, lib.render.radius) == 1
    outs = []
    for m in (_lib.RENDER_DENSE, _lib.RENDER_DENSE_INPLACE):
        for s in lib.slots:
            s.method = m
        lib.render_bpc[m] = 2 if m == _lib.RENDER_DENSE_INPLACE else 1
        lib._lists()
        lib.out.fill_(-7.0)
        lib.status.fill_(5)
        lib.run()
        torch.cuda.synchronize()
        assert int(lib.status.abs().max()) == 0
        outs.append(lib.out.clone())
    assert int(lib.count.max()) > 1024
    assert torch.equal(outs[0], outs[1])


def test_graph_replay_matches_eager():
    """Capture the ragged forward in a hipGraph; every replay (with eager work
    and synchronisation in between) reproduces the eager result."""
    print("ragged hipGraph replay", flush=True)
    lib, *_ = _library()
    eager = lib.out.clone()
    eager_cnt = lib.count.clone()
    lib.capture()
    for r in range(3):
        lib.out.zero_()
        lib.count.fill_(-1)
        torch.cuda.synchronize()
        lib.replay()
        torch.cuda.synchronize()
        assert torch.equal(lib.count, eager_cnt), r
        assert torch.equal(lib.out, eager), r
        _ = (lib.out.sum() * 2).item()               # eager work between replays


def test_autotuner_cache_roundtrip(tmp_path):
    from diffsims_amd.autotune import Autotuner

    f = str(tmp_path / "tune.json")
    calls = []
    t = Autotuner(reps=1, cache_file=f)
    x = torch.zeros(1 << 20, device="cuda")

    def run(c):
        calls.append(c)
        for _ in range(c):
            x.add_(1.0)

    best = t.best(("k", 1), [8, 1, 4], run)
    assert best == 1
    n = len(calls)
    t2 = Autotuner(reps=1, cache_file=f)
    assert t2.best(("k", 1), [8, 1, 4], run) == 1
    assert len(calls) == n                            # served from the cache file
    # a cached winner that is not among the candidates is timed again
    assert t2.best(("k", 1), [8, 4], run) == 4
    assert len(calls) > n
