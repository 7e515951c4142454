"""Line index of a reflection table (dsx_hkl_lines) and the line sweep of the
spot kernel (GPU).

Parity: the line index equals a numpy restatement of its definition (runs of
rows with equal (h, k) and l falling by one; |g|^2 bounds from g4) exactly.
dsx_simulate with the line index returns the same spot lists, counts,
candidate counts and maxima, bit for bit, as the full sweep without it, for
every shape model, with precession, with direct-beam rows, for an irregular
caller-given table, for the beam along and exactly perpendicular to a
reciprocal axis, and in the ragged launch.  The full sweep itself is held to
the FP64 oracle by the parity tests of the spot kernel.
"""

import dataclasses

import numpy as np
import pytest
import torch

from oracle import diffsims_oracle as O
from tests import parity as P

pytestmark = pytest.mark.gpu


def lines_numpy(hkl: np.ndarray, w: np.ndarray) -> np.ndarray:
    """int32 [n_lines, 8] {first, count, h, k, l, min w, max w (float bits), 0}."""
    n = len(hkl)
    if n == 0:
        return np.zeros((0, 8), np.int32)
    cont = np.zeros(n, bool)
    cont[1:] = ((hkl[1:, 0] == hkl[:-1, 0]) & (hkl[1:, 1] == hkl[:-1, 1])
                & (hkl[1:, 2] == hkl[:-1, 2] - 1))
    first = np.flatnonzero(~cont)
    count = np.diff(np.append(first, n))
    w = np.asarray(w, np.float32)
    out = np.zeros((len(first), 8), np.int32)
    out[:, 0] = first
    out[:, 1] = count
    out[:, 2:5] = hkl[first, :3]
    out[:, 5] = np.minimum.reduceat(w, first).view(np.int32)
    out[:, 6] = np.maximum.reduceat(w, first).view(np.int32)
    return out


def hexagonal():
    return O.PhaseSpec((3.21, 3.21, 5.21), (90.0, 90.0, 120.0), ["Mg", "Mg"],
                       np.array([[1 / 3, 2 / 3, 0.25], [2 / 3, 1 / 3, 0.75]]), np.ones(2),
                       np.full(2, 0.6))


def special_rotations():
    """Exact matrices: beam along c*, along a, and exactly perpendicular to
    c* (90 degrees about x and y), plus near-perpendicular Euler angles."""
    ex = [np.eye(3),
          np.array([[1.0, 0, 0], [0, 0, 1.0], [0, -1.0, 0]]),
          np.array([[0, 0, -1.0], [0, 1.0, 0], [1.0, 0, 0]]),
          np.array([[0, 1.0, 0], [-1.0, 0, 0], [0, 0, 1.0]])]
    near = O.bunge_matrix(np.radians(np.array([[0.0, 90.0, 0.0], [30.0, 90.0, 45.0],
                                               [0.0, 89.9999, 0.0], [17.0, 54.7356, 45.0]])))
    return np.concatenate([np.stack(ex), near])


def rotations(n, seed):
    return np.concatenate([special_rotations(), O.bunge_matrix(O.random_euler(n, seed))])


def assert_same_batches(a, b):
    assert a.cap == b.cap
    assert torch.equal(a.count, b.count)
    assert torch.equal(a.ncand, b.ncand)
    assert torch.equal(a.maxI, b.maxI)
    keep = torch.arange(a.cap, device=a.count.device)[None, :] < a.count[:, None].long()
    assert torch.equal(torch.where(keep, a.idx, 0), torch.where(keep, b.idx, 0))
    assert torch.equal(torch.where(keep[..., None], a.spots, 0.0),
                       torch.where(keep[..., None], b.spots, 0.0))


def uses_lines(tab):
    """The kernel takes the line sweep when lines average at least
    DSX_LINE_SWEEP_MIN_ROWS rows."""
    from diffsims_amd import _lib
    return tab.lines is not None and tab.n >= _lib.LINE_SWEEP_MIN_ROWS * tab.n_lines


def sweep_and_lines(tab, M, st, rr):
    from diffsims_amd import engine

    assert uses_lines(tab), (tab.n, tab.n_lines)
    Rt = engine.rotation_matrices_to_device(M)
    with_lines = engine.simulate(tab, Rt, st, rr)
    full = engine.simulate(dataclasses.replace(tab, lines=None), Rt, st, rr)
    torch.cuda.synchronize()
    return with_lines, full


def check_lines(tab):
    ref = lines_numpy(tab.hkl[:, :3].cpu().numpy(), tab.g4[:, 3].cpu().numpy())
    got = tab.lines.cpu().numpy()
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)
    return ref


@pytest.mark.parametrize("name,spec,rr,wdb", [
    ("al", P.al_c0, 3.0, False),
    ("ortho40", P.ortho40, 2.0, True),
    ("hexagonal", hexagonal, 2.5, False),
    ("triclinic", lambda: P.triclinic(12, 7), 1.8, True),
])
def test_line_index_matches_its_definition(name, spec, rr, wdb):
    from diffsims_amd import engine

    print(f"line index {name}", flush=True)
    tab = engine.build_phase_table(P.to_arrays(spec()), rr, wdb)
    ref = check_lines(tab)
    # the enumeration's box columns are the lines: (000) splits the (0, 0)
    # column, each direct-beam row stands alone
    assert ref[:, 1].sum() == tab.n
    assert len(ref) >= len(np.unique(ref[:, 2:4], axis=0))


def test_line_index_of_an_irregular_table():
    """A caller's own table: shuffled blocks, gaps, repeats and single rows."""
    from diffsims_amd import engine

    print("line index irregular table", flush=True)
    rng = np.random.default_rng(11)
    box = np.stack(np.meshgrid(np.arange(3, -4, -1), np.arange(3, -4, -1), np.arange(40, -41, -1),
                               indexing="ij"), -1).reshape(-1, 3)
    keep = rng.random(len(box)) > 0.015
    rows = box[keep]
    blocks = np.array_split(rows, 12)
    order = rng.permutation(len(blocks))
    hkl = np.concatenate([blocks[i] for i in order] + [rows[:5], rows[:5]])
    tab = engine.table_from_hkl(hkl, (5.1, 6.3, 7.7), (88.0, 95.0, 101.0))
    check_lines(tab)
    st = engine.SimSettings(accelerating_voltage=120.0, max_excitation_error=0.03)
    a, b = sweep_and_lines(tab, rotations(300, 3), st, 2.5)
    assert int(a.count.sum()) > 0
    assert_same_batches(a, b)


def test_line_index_of_an_empty_table():
    from diffsims_amd import engine

    dev = torch.device("cuda", 0)
    hkl = torch.zeros((0, 4), dtype=torch.int32, device=dev)
    g4 = torch.zeros((0, 4), dtype=torch.float32, device=dev)
    lines = engine.build_lines(hkl, g4)
    assert lines.shape == (0, 8)


@pytest.mark.parametrize("model,kw", [
    ("lorentzian", {}),
    ("linear", {}),
    ("sinc", {"minima_number": 3.0}),
    ("sin2c", {}),
    ("atanc", {}),
    ("binary", {}),
    ("lorentzian", {"precession_angle": 1.0}),
    ("linear", {"precession_angle": 0.5, "approximate_precession": False}),
    ("binary", {"precession_angle": 0.8}),
    ("lorentzian", {"minimum_intensity": 1e-3}),
    ("lorentzian", {"accelerating_voltage": 10.0, "max_excitation_error": 0.1}),
    ("sinc", {"accelerating_voltage": 300.0, "max_excitation_error": 0.005,
              "shape_factor_width": 0.02}),
])
def test_line_sweep_equals_full_sweep(model, kw):
    from diffsims_amd import engine

    print(f"line sweep {model} {kw}", flush=True)
    kw = dict(kw)
    st = engine.SimSettings(shape_model=model, max_excitation_error=kw.pop("max_excitation_error", 0.02),
                            **kw)
    for spec, rr, wdb, seed in ((P.ortho40, 2.5, True, 1), (hexagonal, 2.5, False, 2),
                                (lambda: P.triclinic(20, 9, edges=(6, 14)), 1.6, False, 3)):
        tab = engine.build_phase_table(P.to_arrays(spec()), rr, wdb)
        a, b = sweep_and_lines(tab, rotations(256, seed), st, rr)
        assert int(a.count.sum()) > 0
        assert_same_batches(a, b)


def test_line_sweep_at_library_size():
    """The large template library's phase at reciprocal radius 2.5: 8,192
    orientations, both sweeps."""
    from diffsims_amd import engine

    print("line sweep library size", flush=True)
    st = engine.SimSettings(max_excitation_error=0.02)
    tab = engine.build_phase_table(P.to_arrays(P.ortho40()), 2.5, False)
    a, b = sweep_and_lines(tab, O.bunge_matrix(O.random_euler(8192, 77)), st, 2.5)
    assert float(a.count.float().mean()) > 100
    assert_same_batches(a, b)


def test_ragged_line_sweep_equals_full_sweep():
    """The ragged launch (every phase with its line index) against the
    single-phase full sweep."""
    from diffsims_amd import engine
    from diffsims_amd.ragged import RaggedLibrary

    print("ragged line sweep", flush=True)
    specs = [P.al_c0(), hexagonal(), P.ortho40(), P.triclinic(9, 4, edges=(8, 16))]
    # rows per line 6.4, 19.1, 8.8, 29.0: phases 1 and 3 take the line sweep, 0 and 2 the full one
    radii = [1.2, 2.5, 1.2, 1.4]
    rots = [rotations(n, 40 + i) for i, n in enumerate((33, 70, 129, 12))]
    st = engine.SimSettings(accelerating_voltage=200, max_excitation_error=0.02)
    lib = RaggedLibrary([P.to_arrays(s, f"p{i}") for i, s in enumerate(specs)], rots, radii, st,
                        (64, 64), 1.0, 0.05)
    lib.run()
    lib.check()
    torch.cuda.synchronize()
    assert [uses_lines(s.table) for s in lib.slots] == [False, True, False, True]
    for i, s in enumerate(lib.slots):
        idx, sp, cnt = lib.phase_spots(i)
        b = engine.simulate(dataclasses.replace(s.table, lines=None),
                            engine.rotation_matrices_to_device(rots[i]), st, radii[i])
        c = cnt.cpu().numpy()
        assert np.array_equal(c, b.count.cpu().numpy()), i
        for o in range(s.n):
            assert np.array_equal(idx[o, :c[o]].cpu().numpy(), b.idx[o, :c[o]].cpu().numpy())
            assert torch.equal(sp[o, :c[o]], b.spots[o, :c[o]])


@pytest.mark.parametrize("seed", range(12))
def test_line_sweep_random_cells(seed):
    """Random triclinic cells, radii, voltages, excitation errors, shape
    models and precession: the line sweep equals the full sweep bit for bit."""
    from diffsims_amd import engine

    rng = np.random.default_rng(900 + seed)
    edges = (float(rng.uniform(3, 10)), float(rng.uniform(10, 24)))
    spec = P.triclinic(int(rng.integers(1, 30)), 1000 + seed, edges=edges)
    model = ["lorentzian", "linear", "sinc", "sin2c", "atanc", "binary"][seed % 6]
    kw = {"accelerating_voltage": float(rng.choice([30.0, 100.0, 200.0, 300.0])),
          "max_excitation_error": float(rng.uniform(0.005, 0.05))}
    if seed % 3 == 0:
        kw["precession_angle"] = float(rng.uniform(0.3, 2.0))
    st = engine.SimSettings(shape_model=model, **kw)
    rr = float(rng.uniform(1.6, 2.6))
    print(f"random cell {seed}: {spec.abc} {model} {kw} r={rr:.2f}", flush=True)
    tab = engine.build_phase_table(P.to_arrays(spec), rr, bool(seed % 2))
    if not uses_lines(tab):
        pytest.skip(f"lines average {tab.n / tab.n_lines:.1f} rows: the kernel keeps the full sweep")
    a, b = sweep_and_lines(tab, rotations(384, 2000 + seed), st, rr)
    assert_same_batches(a, b)
