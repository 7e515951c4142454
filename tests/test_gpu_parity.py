"""HIP path vs the FP64 oracle on the same seeded inputs (GPU).

Every call goes through the C ABI (diffsims_amd.engine binds it with ctypes).
Parity rules are in tests/parity.py.
"""

import warnings

import numpy as np
import pytest

from oracle import diffsims_oracle as O
from tests import parity as P

pytestmark = pytest.mark.gpu


def eng():
    from diffsims_amd import engine
    return engine


def _oracle_params(settings, rr, wdb, scattering="lobato"):
    return O.SimParams(kv=settings.accelerating_voltage, reciprocal_radius=rr,
                       max_excitation_error=settings.max_excitation_error,
                       shape_factor_width=settings.shape_factor_width,
                       shape_model=settings.shape_model,
                       shape_kwargs=({"minima_number": settings.minima_number}
                                     if settings.shape_model in ("sinc", "sin2c", "atanc") else {}),
                       precession_angle=settings.precession_angle,
                       approximate_precession=settings.approximate_precession,
                       minimum_intensity=settings.minimum_intensity, with_direct_beam=wdb,
                       scattering_params=scattering)


def run_parity(ph, settings, rr, wdb, n_orient, seed, scattering="lobato", cap=None,
               render=None):
    e = eng()
    tab = e.build_phase_table(P.to_arrays(ph), rr, wdb, scattering)
    p = _oracle_params(settings, rr, wdb, scattering)
    if settings.shape_model == "binary":
        p.shape_model = O.sf_binary
    otab = O.phase_tables(ph, p)
    assert np.array_equal(tab.hkl_host(), otab.hkl)
    M = O.bunge_matrix(O.random_euler(n_orient, seed))
    b = e.simulate(tab, e.rotation_matrices_to_device(M), settings, rr, cap=cap)
    cnt = b.count.cpu().numpy()
    idx = b.idx.cpu().numpy()
    sp = b.spots.cpu().numpy()
    k = 1 / O.electron_wavelength(settings.accelerating_voltage)
    stats = {}
    refs = []
    for o in range(n_orient):
        # refs[o] is the FP64 reference of the GPU's survivor set (equal to the
        # oracle's own set unless a row within 1e-6 of the cut went the other way)
        ref, _ = P.check_orientation(otab, M[o], k, p, idx[o, :cnt[o]], sp[o, :cnt[o], 3], stats)
        refs.append(ref)
    assert stats["carved_orientations"] <= max(2, n_orient // 50), stats
    return tab, b, M, refs, stats


# -- a2, a6-a9: tables ---------------------------------------------------------

@pytest.mark.parametrize("ph_fn, rr, wdb", [
    (P.al_c0, 1.0, False), (P.al_c0, 2.5, True), (P.ortho40, 2.0, False),
    (lambda: P.triclinic(7, 11), 1.8, True),
    (lambda: O.PhaseSpec((3.073, 3.073, 10.053), (90, 90, 120), ["Si", "Si", "C", "C"],
                         np.array([[0, 0, 0], [.33, .667, .25], [0, 0, .188], [.333, .667, .438]]),
                         np.ones(4), np.zeros(4)), 2.0, False)])
def test_tables_match_oracle(ph_fn, rr, wdb):
    ph = ph_fn()
    tab = eng().build_phase_table(P.to_arrays(ph), rr, wdb, "lobato")
    otab = O.phase_tables(ph, O.SimParams(reciprocal_radius=rr, with_direct_beam=wdb))
    assert np.array_equal(tab.hkl_host(), otab.hkl)           # bit-exact, same order
    F2 = tab.F2.cpu().numpy().astype(np.float64)
    assert np.abs(F2 - otab.F2).max() <= 1e-6 * otab.F2.max()
    g4 = tab.g4.cpu().numpy().astype(np.float64)
    assert np.abs(g4[:, :3] - otab.g).max() <= 1e-6 * max(1.0, np.abs(otab.g).max())


@pytest.mark.parametrize("scattering", ["xtables", None])
def test_structure_factor_other_tables(scattering):
    ph = P.triclinic(9, 12)
    tab = eng().build_phase_table(P.to_arrays(ph), 1.5, True, scattering)
    otab = O.phase_tables(ph, O.SimParams(reciprocal_radius=1.5, with_direct_beam=True,
                                          scattering_params=scattering))
    F2 = tab.F2.cpu().numpy().astype(np.float64)
    assert np.abs(F2 - otab.F2).max() <= 1e-6 * otab.F2.max()


def test_forbidden_reflections_are_exactly_zero():
    tab = eng().build_phase_table(P.to_arrays(P.al_c0()), 1.5, False, "lobato")
    hkl = tab.hkl_host()
    F2 = tab.F2.cpu().numpy()
    mixed = (hkl % 2).sum(axis=1) % 3 != 0       # not all even / all odd
    assert mixed.any() and np.all(F2[mixed] == 0.0) and np.all(F2[~mixed] > 0)


def test_unknown_element_warns_and_scatters_nothing():
    ph = O.PhaseSpec((3.5236,) * 3, (90,) * 3, ["UnOb"], np.zeros((1, 3)), np.ones(1), np.zeros(1))
    with pytest.warns(UserWarning, match="not found in scattering parameter library"):
        tab = eng().build_phase_table(P.to_arrays(ph), 1.0, True, "lobato")
    assert float(tab.F2.abs().max()) == 0.0
    tab = eng().build_phase_table(P.to_arrays(ph), 1.0, True, None)
    assert np.allclose(tab.F2.cpu().numpy(), 1.0)


# -- a3-a5, a10: per-orientation simulation --------------------------------------

def test_small_case_all_orientations():
    """FCC Al, 500 seeded orientations, linear, no direct beam (forbidden-only
    orientations keep their intersecting reflections with I = 0)."""
    st = eng().SimSettings(accelerating_voltage=200, max_excitation_error=0.01, shape_model="linear")
    tab, b, M, refs, stats = run_parity(P.al_c0(), st, 1.0, False, 500, 20261202)
    forbidden_only = sum(1 for r in refs if r.idx.size and r.max_intensity == 0.0)
    assert 150 < forbidden_only < 350                 # ~49 % of random orientations
    assert stats["max_rel_err"] <= P.TOL_I


def test_large_library_cell_subset():
    st = eng().SimSettings(accelerating_voltage=200, max_excitation_error=0.02,
                           shape_model="lorentzian")
    tab, b, M, refs, stats = run_parity(P.ortho40(), st, 2.0, False, 192, 7)
    assert 150 < np.mean([r.idx.size for r in refs]) < 400
    assert stats["max_rel_err"] <= P.TOL_I


@pytest.mark.parametrize("model", ["linear", "lorentzian", "sinc", "sin2c", "atanc", "binary"])
def test_shape_models(model):
    st = eng().SimSettings(accelerating_voltage=300, max_excitation_error=0.03,
                           shape_factor_width=0.025, shape_model=model, minima_number=4)
    run_parity(P.triclinic(12, 21), st, 1.6, True, 96, 31)


def test_precession_lorentzian():
    st = eng().SimSettings(accelerating_voltage=200, max_excitation_error=0.01,
                           shape_model="lorentzian", precession_angle=1.0)
    tab, b, M, refs, stats = run_parity(P.triclinic(10, 41), st, 1.8, False, 64, 5)
    assert stats["max_rel_err"] <= P.TOL_I


@pytest.mark.parametrize("model", ["linear", "lorentzian", "sinc", "sin2c", "atanc", "binary"])
def test_full_precession_average(model):
    """approximate_precession=False: the device average of every built-in
    model against scipy.integrate.quad per spot in the oracle."""
    st = eng().SimSettings(accelerating_voltage=200, max_excitation_error=0.02,
                           shape_factor_width=0.025, shape_model=model, minima_number=4,
                           precession_angle=1.2, approximate_precession=False)
    tab, b, M, refs, stats = run_parity(P.triclinic(6, 43), st, 1.4, False, 6, 7)
    assert stats.get("max_rel_err", 0.0) <= P.TOL_I
    assert np.mean([r.idx.size for r in refs]) > 20


def test_cap_overflow_rerun_gives_the_same_result():
    st = eng().SimSettings(max_excitation_error=0.02)
    tab, b, M, refs, stats = run_parity(P.ortho40(), st, 2.0, False, 32, 9, cap=16)
    assert b.cap >= max(r.n_intersect for r in refs)


def test_direct_beam_rows():
    st = eng().SimSettings(accelerating_voltage=300, max_excitation_error=0.01)
    si = O.PhaseSpec((5.431,) * 3, (90,) * 3, ["Si"] * 8, P.si_positions(), np.ones(8), np.zeros(8))
    tab = eng().build_phase_table(P.to_arrays(si), 5.0, True, "lobato")
    b = eng().simulate(tab, eng().rotation_matrices_to_device(np.eye(3)[None]), st, 5.0)
    n = int(b.count[0])
    assert n == 70                                     # test_simulation_generator.py:137-143
    hkl = tab.hkl_host()[b.idx[0, :n].cpu().numpy()]
    assert (np.abs(hkl).sum(axis=1) == 0).sum() == 2   # (000) twice, last
    assert np.abs(hkl[-2:]).sum() == 0


# -- a12, a13: rendering ---------------------------------------------------------

def _render_parity(ph, settings, rr, n, seed, shape, sigma, cal):
    tab, b, M, refs, _ = run_parity(ph, settings, rr, False, n, seed)
    pats = eng().render(b, shape, sigma, cal).cpu().numpy()
    sp = b.spots.cpu().numpy()
    cnt = b.count.cpu().numpy()
    worst = 0.0
    for o, ref in enumerate(refs):
        assert np.array_equal(b.idx[o, :cnt[o]].cpu().numpy(), ref.idx)
        rp = P.flipped_reference_pattern(ref, sp[o, :cnt[o], :2], shape, sigma, cal)
        worst = max(worst, P.pattern_error(pats[o], rp))
    assert worst <= P.TOL_P, worst
    return worst


def test_render_fast_small_case():
    st = eng().SimSettings(max_excitation_error=0.01, shape_model="linear")
    _render_parity(P.al_c0(), st, 1.0, 500, 20261202, (144, 144), 1.0, 0.01)


def test_render_fast_band_path():
    # above the small-batch threshold engine.render takes the band renderer
    assert eng().fast_renderer(1200, 144, 144) == "dsx_render_fast_bands"
    st = eng().SimSettings(max_excitation_error=0.01, shape_model="linear")
    _render_parity(P.al_c0(), st, 1.0, 1200, 20261203, (144, 144), 1.0, 0.01)


def test_render_fast_large_cell():
    st = eng().SimSettings(max_excitation_error=0.02, shape_model="lorentzian")
    _render_parity(P.ortho40(), st, 2.0, 64, 3, (256, 256), 1.0, 0.015)


@pytest.mark.parametrize("sigma", [0.6, 2.5, 10.0])
def test_render_fast_sigmas(sigma):
    st = eng().SimSettings(max_excitation_error=0.02, shape_model="lorentzian")
    _render_parity(P.triclinic(6, 8), st, 1.5, 24, 4, (96, 128), sigma, 0.02)


def test_render_subpixel():
    st = eng().SimSettings(max_excitation_error=0.02, shape_model="lorentzian")
    ph = P.triclinic(6, 8)
    tab, b, M, refs, _ = run_parity(ph, st, 1.5, False, 24, 4)
    shape, sigma, cal, clip = (100, 120), 1.3, 0.02, 1e-3
    cnt = b.count.cpu().numpy()
    for normalize in (True, False):
        pats = eng().render(b, shape, sigma, cal, fast=False, normalize=normalize,
                            clip_threshold=clip).cpu().numpy()
        for o, ref in enumerate(refs):
            assert np.array_equal(b.idx[o, :cnt[o]].cpu().numpy(), ref.idx)
            raw = O.render_subpixel(ref.xyz[:, :2], ref.intensity, shape, sigma, cal,
                                    normalize=False, clip_threshold=clip)
            rp = raw / raw.max() if (normalize and raw.max() > 0) else raw
            # a clipped box whose edge moves by one pixel under FP32 coordinates
            # changes pixels whose value is <= clip_threshold
            edge = clip / raw.max() if (normalize and raw.max() > 0) else clip
            assert np.abs(pats[o] - rp).max() <= P.TOL_P * max(rp.max(), 1e-300) + edge


def test_render_in_plane_angle_and_mirror():
    st = eng().SimSettings(max_excitation_error=0.02, shape_model="lorentzian")
    ph = P.triclinic(6, 8)
    tab, b, M, refs, _ = run_parity(ph, st, 1.5, False, 24, 4)
    shape, sigma, cal = (128, 128), 1.0, 0.02
    compared = 0
    for angle, mirrored in [(30.0, False), (-75.0, True), (0.0, True)]:
        pats = eng().render(b, shape, sigma, cal, in_plane_angle=angle, mirrored=mirrored).cpu().numpy()
        cnt = b.count.cpu().numpy()
        for o, ref in enumerate(refs):
            assert np.array_equal(b.idx[o, :cnt[o]].cpu().numpy(), ref.idx)
            X, Y = O.pixel_coords(ref.xyz[:, :2], shape, cal, (64, 64), angle, mirrored)
            fr = np.concatenate([X - np.floor(X), Y - np.floor(Y)])
            if np.any((fr < P.PIX_EPS) | (fr > 1 - P.PIX_EPS)):
                continue
            rp = O.render_fast(ref.xyz[:, :2], ref.intensity, shape, sigma, cal, angle, mirrored)
            assert np.abs(pats[o] - rp).max() <= P.TOL_P * max(rp.max(), 1e-300)
            compared += 1
    assert compared >= 60


def test_render_no_spots_and_zero_intensity():
    e = eng()
    import torch
    dev = torch.device("cuda")
    spots = torch.zeros((3, 4, 4), dtype=torch.float32, device=dev)
    spots[0, 0] = torch.tensor([10.0, 10.0, 0.0, 5.0])      # outside the frame
    spots[1, :2, 3] = 0.0                                     # zero intensities
    count = torch.tensor([1, 2, 0], dtype=torch.int32, device=dev)
    out = e.render_spots(spots, count, 4, (50, 50), 1.0, 0.001).cpu().numpy()
    assert np.all(out == 0.0) and not np.isnan(out).any()


@pytest.mark.parametrize("case", ["equal", "dominant", "border", "clustered"])
def test_render_fast_maximum_search(case):
    """The fast renderer finds each pattern's maximum from tile bounds and
    falls back to a full pass when too many tiles remain; both must give the
    oracle's normalised pattern.  Spots sit at pixel centres (no truncation
    ambiguity) with intensity layouts that drive either branch."""
    import torch

    e = eng()
    rng = np.random.default_rng({"equal": 1, "dominant": 2, "border": 3, "clustered": 4}[case])
    n_pat, n_spots, shape, cal, sigma = 24, 300, (200, 160), 0.01, 1.0
    H, W = shape
    cx, cy = W // 2, H // 2
    px = rng.integers(0, W, (n_pat, n_spots))
    py = rng.integers(0, H, (n_pat, n_spots))
    if case == "border":
        px[:, ::2] = rng.choice([0, 1, W - 2, W - 1], (n_pat, (n_spots + 1) // 2))
        py[:, 1::2] = rng.choice([0, 1, H - 2, H - 1], (n_pat, n_spots // 2))
    if case == "clustered":
        px = np.clip(W // 3 + rng.integers(-6, 7, (n_pat, n_spots)), 0, W - 1)
        py = np.clip(H // 2 + rng.integers(-6, 7, (n_pat, n_spots)), 0, H - 1)
    I = {"equal": np.ones((n_pat, n_spots)),
         "dominant": rng.exponential(1.0, (n_pat, n_spots)) * (1 + 50 * (np.arange(n_spots) == 7)),
         "border": rng.exponential(1.0, (n_pat, n_spots)),
         "clustered": rng.uniform(0.5, 1.0, (n_pat, n_spots))}[case]
    x = (px + 0.5 - cx) * cal
    y = (py + 0.5 - cy) * cal
    spots = np.zeros((n_pat, n_spots, 4), np.float32)
    spots[..., 0], spots[..., 1], spots[..., 3] = x, y, I
    dev = torch.device("cuda")
    st = torch.from_numpy(spots).to(dev)
    cnt = torch.full((n_pat,), n_spots, dtype=torch.int32, device=dev)
    out = e.render_spots(st, cnt, n_spots, shape, sigma, cal).cpu().numpy()
    for o in range(n_pat):
        X = (spots[o, :, 0] / np.float32(cal) + np.float32(cx)).astype(np.float64)
        Y = (spots[o, :, 1] / np.float32(cal) + np.float32(cy)).astype(np.float64)
        rp = P.render_from_pixels(X, Y, spots[o, :, 3], shape, sigma)
        assert np.abs(out[o] - rp).max() <= P.TOL_P * rp.max(), (case, o)


@pytest.mark.parametrize("axes", [False, True])
def test_polar_templates_on_device(axes):
    """dsx_polar_templates against the oracle's polar_flatten on the same spot
    lists (FP64 coordinates); nearest-index ties within 1e-6 of a midpoint may
    fall on either side."""
    import torch

    e = eng()
    st = e.SimSettings(max_excitation_error=0.02, shape_model="lorentzian")
    ph = P.triclinic(6, 8)
    tab, b, M, refs, _ = run_parity(ph, st, 1.5, False, 40, 9)
    cnt = b.count.cpu().numpy()
    ok = [o for o in range(len(refs)) if np.array_equal(b.idx[o, :cnt[o]].cpu().numpy(), refs[o].idx)]
    lists = [(refs[o].xyz, refs[o].intensity) for o in ok]
    radial = np.linspace(0, 1.6, 81) if axes else None
    az = np.linspace(-np.pi, np.pi, 361) if axes else None
    r_ref, t_ref, i_ref = O.polar_flatten(lists, radial, az)
    r, t, I = e.polar_templates(b.spots[ok], b.count[ok], b.cap, radial, az,
                                width=r_ref.shape[1])
    r, t, I = r.cpu().numpy(), t.cpu().numpy(), I.cpu().numpy()
    assert np.abs(I - i_ref).max() <= P.TOL_I * max(i_ref.max(), 1e-300)
    if not axes:
        assert np.abs(r - r_ref).max() <= 1e-6 and np.abs(t - t_ref).max() <= 1e-5
        return
    for row in range(len(ok)):
        xyz = refs[ok[row]].xyz
        rv = np.hypot(xyz[:, 0], xyz[:, 1])
        tv = np.arctan2(xyz[:, 1], xyz[:, 0])
        near = np.any(np.abs(np.abs(rv[:, None] - (radial[:-1] + radial[1:])[None] / 2)) < 1e-6) or \
            np.any(np.abs(np.abs(tv[:, None] - (az[:-1] + az[1:])[None] / 2)) < 1e-6)
        if near:
            continue
        assert np.array_equal(r[row], r_ref[row]) and np.array_equal(t[row], t_ref[row]), row


@pytest.mark.parametrize("shape,cal", [((512, 480), 0.01), ((100, 120), 0.03)])
def test_render_fast_many_spots(shape, cal):
    """Fast rendering of patterns with more than 1,024 spots: the windowed
    strip renderer (512x480) and the whole-pattern dense renderer (100x120),
    against the oracle (flipped-pixel carve-out as usual)."""
    e = eng()
    rng = np.random.default_rng(12)
    ph = O.PhaseSpec((19.5, 20.5, 21.0), (88.0, 96.0, 91.0), ["O", "Si", "Fe", "Mg"],
                     rng.uniform(0, 1, (4, 3)), np.ones(4), rng.uniform(0.3, 1.0, 4))
    st = e.SimSettings(max_excitation_error=0.03, shape_model="lorentzian")
    tab, b, M, refs, _ = run_parity(ph, st, 1.2, False, 4, 31)
    cnt = b.count.cpu().numpy()
    assert b.cap > 1024 and cnt.min() > 1024
    pats = e.render(b, shape, 1.0, cal).cpu().numpy()
    sp = b.spots.cpu().numpy()
    for o, ref in enumerate(refs):
        assert np.array_equal(b.idx[o, :cnt[o]].cpu().numpy(), ref.idx)
        rp = P.flipped_reference_pattern(ref, sp[o, :cnt[o], :2], shape, 1.0, cal)
        assert P.pattern_error(pats[o], rp) <= P.TOL_P, o


def test_render_workspace_on_two_streams():
    """Large fast patterns (windowed strip renderer, 512x480) and large
    sub-pixel patterns (chunked renderer) take scratch from the caching
    allocator per launch: renders launched on two side streams at once, three
    times over, equal the render on the current stream."""
    import torch

    e = eng()
    rng = np.random.default_rng(12)
    ph = O.PhaseSpec((19.5, 20.5, 21.0), (88.0, 96.0, 91.0), ["O", "Si", "Fe", "Mg"],
                     rng.uniform(0, 1, (4, 3)), np.ones(4), rng.uniform(0.3, 1.0, 4))
    st = e.SimSettings(max_excitation_error=0.03, shape_model="lorentzian")
    tab, b, M, refs, _ = run_parity(ph, st, 1.2, False, 4, 31)
    assert b.cap > 1024 and int(b.count.min()) > 1024
    for shape, fast in (((512, 480), True), ((256, 256), False)):
        print(f"two streams: shape {shape} fast={fast}", flush=True)
        ref = e.render(b, shape, 1.0, 0.01, fast=fast)
        ref2 = e.render(b, shape, 1.0, 0.01, fast=fast)
        tol = 0.0 if torch.equal(ref, ref2) else 1e-6 * float(ref.abs().max())
        streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        outs = [torch.full_like(ref, -1.0) for _ in streams]
        stats = [torch.zeros(b.n, dtype=torch.int32, device=ref.device) for _ in streams]
        torch.cuda.synchronize()
        for _ in range(3):
            for out, status, s in zip(outs, stats, streams):
                e.render_spots(b.spots, b.count, b.cap, shape, 1.0, 0.01, fast=fast, out=out,
                               stream=s.cuda_stream, status=status)
        torch.cuda.synchronize()
        for out, status in zip(outs, stats):
            e.check_render_status(status)
            assert float((out - ref).abs().max()) <= tol


def test_stream_handle_is_the_current_stream():
    """The handle every entry point receives is torch's current stream, on
    the default stream and inside a stream context, however the device is
    named."""
    import torch

    from diffsims_amd import _lib

    idx = torch.cuda.current_device()
    names = (None, idx, f"cuda:{idx}", torch.device("cuda", idx), torch.device("cuda"), "cuda")
    side = torch.cuda.Stream()
    for s in (torch.cuda.current_stream(), side):
        with torch.cuda.stream(s):
            want = torch.cuda.current_stream(idx).cuda_stream
            assert want == s.cuda_stream
            for d in names:
                assert _lib.stream_handle(d) == want, d
    assert side.cuda_stream != 0
