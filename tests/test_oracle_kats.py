"""The FP64 oracle against the reference's golden data and known-answer tests,
closed forms and brute force (CPU only).

Each test names the reference test or source it pins.  Some tests plant a
mistake in a copy of an oracle function and check that the pinning data
catches it, so that the pins are shown to be sensitive.
"""

import os

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import diffsims_oracle as O
from tests.parity import FCC_POS, si_positions

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "old_simulation.npy")


# -- a1: wavelength (diffsims/tests/utils/test_sim_utils.py:49-60) ------------

@pytest.mark.parametrize("kv, lam", [(100, 0.0370143659), (200, 0.0250793403),
                                     (300, 0.0196874888), ("inf", 0)])
def test_wavelength_kat(kv, lam):
    np.testing.assert_almost_equal(O.electron_wavelength(kv), lam)


def test_wavelength_closed_form_limits():
    # non-relativistic limit h / sqrt(2 m e E) at low voltage within 1e-5
    from scipy.constants import e, h, m_e
    lam = O.electron_wavelength(0.01)
    assert abs(lam - h / np.sqrt(2 * m_e * e * 10.0) * 1e10) / lam < 1e-5


# -- a2: hkl box (reciprocal_lattice_vector.py:1077-1142 and its tests) --------

def _rec(a, b, c, al=90, be=90, ga=90):
    return O.lattice_basis(a, b, c, al, be, ga)[1]


@pytest.mark.parametrize("include_zero", [True, False])
@pytest.mark.parametrize("d, size", [(2, 18), (1, 92), (0.5, 750)])
def test_ferrite_sizes(d, size, include_zero):
    # test_reciprocal_lattice_vector.py:76-87
    hkl = O.hkl_box(_rec(2.8665, 2.8665, 2.8665), d, include_zero)
    assert hkl.shape[0] == size + include_zero
    if include_zero:
        assert np.array_equal(hkl[-1], [0, 0, 0])


def test_ferrite_first_row():
    # test_reciprocal_lattice_vector.py:118-124: repr starts with [1, 1, 0]
    assert np.array_equal(O.hkl_box(_rec(2.8665, 2.8665, 2.8665), 2)[0], [1, 1, 0])


def test_al_docstring_counts_and_order():
    # reciprocal_lattice_vector.py:1107-1127 (Al a = 4.04)
    rec = _rec(4.04, 4.04, 4.04)
    h = O.hkl_box(rec, 0.7)
    assert h.shape[0] == 798
    assert np.array_equal(h[0], [5, 2, 2]) and np.array_equal(h[1], [5, 2, 1])
    assert np.array_equal(h[2], [5, 2, 0]) and np.array_equal(h[-1], [-5, -2, -2])
    assert O.hkl_box(rec, 1.0).shape[0] == 256
    d = 1 / O.rnorm(h, rec)
    assert abs(d.min() - 0.7032737300610338) < 1e-12


@pytest.mark.parametrize("d, box", [(0.5, [6, 6, 21]), (1, [3, 3, 11]), (1.5, [2, 2, 7])])
def test_sic_box(d, box):
    # test_get_hkl.py:52-64: max hkl + 1 equals the deprecated bound
    h = O.hkl_box(_rec(3.073, 3.073, 10.053, 90, 90, 120), d)
    assert np.array_equal(h.max(axis=0) + 1, box)


def test_sic_d3_order():
    # test_reciprocal_lattice_vector.py:147-155
    h = O.hkl_box(_rec(3.073, 3.073, 10.053, 90, 90, 120), 3)
    assert np.array_equal(h[:, 2], [3, 2, 1, -1, -2, -3]) and not h[:, :2].any()


def test_box_drops_in_sphere_reflections_for_oblique_cells():
    # the box rule, not the sphere: SiC d_min = 0.5 keeps 2688 of 2764 in-sphere
    rec = _rec(3.073, 3.073, 10.053, 90, 90, 120)
    h = O.hkl_box(rec, 0.5)
    assert h.shape[0] == 2688
    big = np.stack(np.meshgrid(*[np.arange(-30, 31)] * 3, indexing="ij"), -1).reshape(-1, 3)
    big = big[np.any(big != 0, axis=1)]
    with np.errstate(divide="ignore"):
        assert int(np.sum(1 / O.rnorm(big, rec) >= 0.5)) == 2764


def test_cartesian_order_equals_hkl_order():
    # orix sorts by Cartesian (x, y, z) descending; in the orix frame that is
    # the descending (h, k, l) order for any cell
    for abc, ang in [((3.0, 4.0, 5.0), (80, 95, 105)), ((3.073, 3.073, 10.053), (90, 90, 120))]:
        _, rec = O.lattice_basis(*abc, *ang)
        h = O.hkl_box(rec, 0.8)
        g = O.g_cartesian(h, rec)
        order = np.lexsort((-g[:, 2], -g[:, 1], -g[:, 0]))
        assert np.array_equal(order, np.arange(h.shape[0]))


@pytest.mark.parametrize("hkl, g", [((1, 0, 0), 0.34885749), ((2, 0, 0), 0.69771498),
                                    ((1, 1, 0), 0.493359), ((2, 1, 0), 0.78006907),
                                    ((2, 2, 0), 0.98671799), ((1, 1, 1), 0.604238),
                                    ((2, 1, 1), 0.85452285)])
def test_ferrite_gspacing(hkl, g):
    # test_reciprocal_lattice_vector.py:157-187
    # np.allclose tolerances, as the reference test uses
    assert np.allclose(O.rnorm(np.array(hkl), _rec(2.8665, 2.8665, 2.8665)), g)


def test_orix_frame_exact_zeros():
    base, rec = O.lattice_basis(3.0, 4.0, 5.0, 90, 90, 90)
    assert np.count_nonzero(base - np.diag(np.diag(base))) == 0
    base, rec = O.lattice_basis(2.464, 2.464, 6.711, 90, 90, 120)
    assert base[0, 1] == base[0, 2] == base[1, 2] == 0.0 and rec[0, 1] == rec[0, 2] == 0.0
    assert base[1, 0] == -0.5 * 2.464  # degree-exact cos(120)


def _frame_error(base, rec, abc, ang):
    """How far a (base, rec) pair is from the orix frame of the cell: the
    largest of the errors in the lengths and cosines recovered from the basis
    rows, the components that must vanish (a || x, b in the xy-plane, c* || z),
    the sign conditions (a_x, b_y, c*_z > 0, right-handed basis) and the
    duality base @ rec = I.  The frame is fixed by these conditions alone
    (test_simulation_generator.py:311 states x || a and z || c*; y follows
    from right-handedness), so none of them restates the oracle's formulas."""
    a, b, c = (np.asarray(base[i], dtype=np.float64) for i in range(3))
    errs = []
    for v, L in zip((a, b, c), abc):
        errs.append(abs(np.linalg.norm(v) - L) / L)
    for (u, v), A in (((b, c), ang[0]), ((a, c), ang[1]), ((a, b), ang[2])):
        cos = np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v))
        errs.append(abs(cos - np.cos(np.radians(A))))
    errs += [abs(a[1]), abs(a[2]), abs(b[2]), abs(rec[0, 2]), abs(rec[1, 2])]
    errs += [0.0 if a[0] > 0 else 1.0, 0.0 if b[1] > 0 else 1.0, 0.0 if rec[2, 2] > 0 else 1.0,
             0.0 if np.linalg.det(base) > 0 else 1.0]
    errs.append(np.abs(np.asarray(base) @ np.asarray(rec) - np.eye(3)).max())
    return max(errs)


def _oblique_cells():
    """Monoclinic and triclinic cells: hand-picked ones and the cells of the
    workloads (the 64-phase ragged batch, the 24-atom monoclinic refinement
    cell, the 500-atom triclinic precession cell)."""
    from diffsims_amd import workloads

    cells = [((3.0, 4.0, 5.0), (90.0, 103.4, 90.0)), ((5.1, 6.3, 7.7), (80.0, 95.0, 105.0)),
             ((9.0, 4.0, 12.0), (60.0, 70.0, 110.0)), ((4.0, 4.0, 4.0), (109.47, 109.47, 109.47)),
             ((7.0, 8.0, 9.0), (90.0, 90.0, 97.0)), ((7.0, 8.0, 9.0), (84.0, 90.0, 90.0))]
    rw = workloads.refinement_monoclinic(8)
    cells.append((rw.phase.abc, rw.phase.angles))
    cells.append((rw.truth_phase.abc, rw.truth_phase.angles))
    lp = workloads.large_precession(4)
    cells.append((lp["phase"].abc, lp["phase"].angles))
    rl = workloads.ragged_library(orient_scale=1e-4)
    cells += [(p.abc, p.angles) for p in rl.phases]
    return cells


def test_oblique_frame_recovers_cell():
    # monoclinic/triclinic frame: lengths and cosines recovered to 1e-12,
    # x || a, b in the xy-plane, c* || +z, right-handed, rec = inv(base)
    cells = _oblique_cells()
    assert len(cells) >= 70
    for abc, ang in cells:
        base, rec = O.lattice_basis(*abc, *ang)
        assert _frame_error(base, rec, abc, ang) < 1e-12, (abc, ang)


def _mutant_basis(abc, ang, which):
    """The oblique basis with one planted mistake (test of the pin's sensitivity)."""
    a, b, c = abc
    ca, cb, cg = (np.cos(np.radians(x)) for x in ang)
    sg = np.sin(np.radians(ang[2]))
    a21, a22, a31 = b * cg, b * sg, c * cb
    a32 = c * (ca - cb * cg) / sg
    if which == "a32_sign":
        a32 = c * (ca + cb * cg) / sg
    elif which == "a32_no_cb":
        a32 = c * ca / sg
    elif which == "a31_gamma":
        a31 = c * cg
    elif which == "alpha_beta_swapped":
        a31 = c * ca
        a32 = c * (cb - ca * cg) / sg
    a33 = c * np.sqrt(max(1.0 - (a31 / c) ** 2 - (a32 / c) ** 2, 0.0))
    base = np.array([[a, 0.0, 0.0], [a21, a22, 0.0], [a31, a32, a33]])
    if which == "left_handed":
        base[2, 2] = -base[2, 2]
    return base, np.linalg.inv(base)


@pytest.mark.parametrize("which", ["a32_sign", "a32_no_cb", "a31_gamma", "alpha_beta_swapped",
                                   "left_handed"])
def test_oblique_frame_catches_planted_mistakes(which):
    abc, ang = (5.1, 6.3, 7.7), (80.0, 95.0, 105.0)
    base, rec = _mutant_basis(abc, ang, which)
    assert _frame_error(base, rec, abc, ang) > 1e-3
    # and the unmutated closed form of the same construction passes
    base, rec = _mutant_basis(abc, ang, "none")
    assert _frame_error(base, rec, abc, ang) < 1e-12


def test_oblique_g_spacings_from_metric():
    # |g| = 1/d with d from the textbook triclinic d-spacing formula
    # (reciprocal metric = inverse of the direct metric tensor)
    for abc, ang in _oblique_cells()[:12]:
        _, rec = O.lattice_basis(*abc, *ang)
        a, b, c = abc
        ca, cb, cg = (np.cos(np.radians(x)) for x in ang)
        G = np.array([[a * a, a * b * cg, a * c * cb], [a * b * cg, b * b, b * c * ca],
                      [a * c * cb, b * c * ca, c * c]])
        Gs = np.linalg.inv(G)
        hkl = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, -2, 3], [-4, 1, 2], [2, 2, -1]])
        want = np.sqrt(np.einsum("ni,ij,nj->n", hkl, Gs, hkl))
        assert np.abs(O.rnorm(hkl, rec) - want).max() / want.max() < 1e-13


# -- a3: rotations -------------------------------------------------------------

def test_quaternion_matrix_matches_bunge():
    e = O.random_euler(200, 3)
    M = O.bunge_matrix(e)
    assert np.abs(O.matrix_from_quat(O.quat_from_bunge(e)) - M).max() < 1e-14
    assert np.abs(np.einsum("nij,nkj->nik", M, M) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-14


def test_rotation_recovers_hkl():
    # test_diffracting_vector.py:47-54: rotating vectors and basis recovers hkl
    _, rec = O.lattice_basis(3.0, 4.0, 5.0, 80, 95, 105)
    hkl = np.array([[1, 2, 3], [-2, 0, 1]])
    M = O.bunge_matrix(np.array([0.3, 1.1, -0.4]))
    g_lab = O.g_cartesian(hkl, rec) @ M
    back = (g_lab @ M.T) @ np.linalg.inv(rec).T
    assert np.abs(back - hkl).max() < 1e-12


# -- a5: shape factors (shape_factor_models.py) ----------------------------------

def test_linear_kat():
    # test_simulation_generator.py:106-110
    s = np.array([-2, -1, -0.5, 0, 0.5, 1, 2])
    np.testing.assert_allclose(O.sf_linear(s, 1), [0, 0, 0.5, 1, 0.5, 0, 0])


def test_model_special_values():
    assert O.sf_sinc(np.array([0.0]), 0.1)[0] == 0.0        # :95-101
    assert O.sf_sin2c(np.array([0.0]), 0.1)[0] == 0.0
    assert O.sf_atanc(np.array([0.0]), 0.1)[0] == 1.0       # :147-151
    s = np.linspace(-0.05, 0.05, 11)
    np.testing.assert_allclose(O.sf_lorentzian(s, 0.02), 1 / (1 + (np.pi * s / 0.02) ** 2))


def test_lorentzian_precession_zero_angle():
    s = np.linspace(-0.05, 0.05, 11)
    np.testing.assert_allclose(O.sf_lorentzian_precession(s, 0.02, 1.0, 0.0),
                               (2 / 0.02) * O.sf_lorentzian(s, 0.02))


def test_lorentzian_precession_is_scaled_exact_average():
    # eq. 10 of Palatinus et al. equals (2 sigma / pi) times the exact
    # precession average of the Lorentzian; check against quad
    w, phi = 0.02, np.deg2rad(1.0)
    for s, r in [(0.003, 0.5), (-0.01, 1.7), (0.0, 2.5)]:
        avg = quad(lambda t: O.sf_lorentzian(s + r * phi * np.cos(t), w), 0, 2 * np.pi,
                   limit=200)[0] / (2 * np.pi)
        sigma = np.pi / w
        assert abs(O.sf_lorentzian_precession(s, w, r, phi) - 2 * sigma / np.pi * avg) < 1e-8 * (
            2 * sigma / np.pi)


# -- a6-a9: scattering and structure factors ------------------------------------

def test_ni_hex_kat():
    # test_sim_utils.py:330-353: Ni hex (3, 3, 5, 120), r = 0.2 -> only (000),
    # lobato I = 43.0979, None -> 1
    base, rec = O.lattice_basis(3, 3, 5, 90, 90, 120)
    coeffs, _ = O.atom_coeffs(["Ni"], "lobato")
    F = O.structure_factor(np.zeros((1, 3)), np.zeros(1), np.zeros((1, 3)), np.ones(1),
                           np.zeros(1), coeffs, "lobato")
    assert abs(abs(F[0]) ** 2 - 43.0979) < 1e-4
    F = O.structure_factor(np.zeros((1, 3)), np.zeros(1), np.zeros((1, 3)), np.ones(1),
                           np.zeros(1), coeffs, None)
    assert abs(F[0]) ** 2 == 1.0


def test_unknown_element_kat():
    # test_sim_utils.py:370-394: unknown element -> f = 0 (lobato), 1 (None)
    coeffs, missing = O.atom_coeffs(["UnOb"], "lobato")
    assert missing == ["UnOb"] and not coeffs.any()
    assert O.element_of("Fe2+") == "Fe" and O.element_of("UnOb") == "UnOb"


def _brute_F(hkl, g, xyz, occ, B, coeffs, table):
    out = np.zeros(len(hkl), dtype=complex)
    for i, h in enumerate(hkl):
        for j in range(len(xyz)):
            if table == "lobato":
                f = sum(a * (2 + b * g[i] ** 2) / (1 + b * g[i] ** 2) ** 2 for a, b in coeffs[j])
            else:
                f = sum(a * np.exp(-b * g[i] ** 2 / 4) for a, b in coeffs[j])
            out[i] += f * occ[j] * np.exp(-0.25 * g[i] ** 2 * B[j]) * np.exp(
                2j * np.pi * np.dot(h, xyz[j]))
    return out


@pytest.mark.parametrize("table", ["lobato", "xtables"])
def test_structure_factor_brute_force(table):
    rng = np.random.default_rng(5)
    syms = ["C", "O", "Fe", "Si", "Mg"]
    xyz = rng.uniform(0, 1, (5, 3))
    occ = rng.uniform(0.5, 1, 5)
    B = rng.uniform(0.2, 1.0, 5)
    coeffs, _ = O.atom_coeffs(syms, table)
    hkl = rng.integers(-4, 5, (20, 3))
    _, rec = O.lattice_basis(4.1, 5.2, 6.3, 85, 95, 100)
    g = O.rnorm(hkl, rec)
    for mode in ("reference", "exact"):
        F = O.structure_factor(hkl, g, xyz, occ, B, coeffs, table, mode)
        assert np.abs(F - _brute_F(hkl, g, xyz, occ, B, coeffs, table)).max() < 1e-10


def test_friedel_and_extinctions():
    rng = np.random.default_rng(6)
    coeffs, _ = O.atom_coeffs(["Si", "O", "O"], "lobato")
    xyz = rng.uniform(0, 1, (3, 3))
    hkl = rng.integers(-5, 6, (30, 3))
    g = np.linalg.norm(hkl, axis=1) / 5.0
    F1 = O.structure_factor(hkl, g, xyz, np.ones(3), np.full(3, 0.4), coeffs, "lobato")
    F2 = O.structure_factor(-hkl, g, xyz, np.ones(3), np.full(3, 0.4), coeffs, "lobato")
    assert np.abs(np.abs(F1) - np.abs(F2)).max() < 1e-12
    # FCC: mixed parity is extinct; diamond: all even with h+k+l = 4n+2 too
    cf, _ = O.atom_coeffs(["Al"] * 4, "lobato")
    fcc = np.array([[1, 0, 0], [2, 1, 0], [1, 1, 0], [3, 2, 1]])
    F = O.structure_factor(fcc, np.ones(4), FCC_POS, np.ones(4), np.zeros(4), cf, "lobato", "exact")
    assert np.all(F == 0)
    cs, _ = O.atom_coeffs(["Si"] * 8, "lobato")
    dia = np.array([[2, 0, 0], [2, 2, 2], [4, 2, 0]])
    F = O.structure_factor(dia, np.ones(3), si_positions(), np.ones(8), np.zeros(8), cs, "lobato",
                           "exact")
    assert np.all(F == 0)
    # the reference's own arithmetic leaves round-off there
    Fr = O.structure_factor(fcc, np.ones(4), FCC_POS, np.ones(4), np.zeros(4), cf, "lobato",
                            "reference")
    assert np.all(np.abs(Fr) < 1e-13)


# -- a4, a10, a11: full per-orientation KATs (test_simulation_generator.py) -------

def _si(a=5.431):
    return O.PhaseSpec((a, a, a), (90, 90, 90), ["Si"] * 8, si_positions(), np.ones(8), np.zeros(8))


@pytest.mark.parametrize("prec, approx, expected", [(0, True, 70), (0.5, True, 250),
                                                    (0.5, False, 250)])
def test_si_spot_counts(prec, approx, expected):
    # :137-161 (300 kV, r = 5, identity, with the direct beam, lorentzian)
    ph = _si()
    base, rec = ph.basis()
    p = O.SimParams(kv=300, reciprocal_radius=5.0, precession_angle=prec, approximate_precession=approx)
    rows = O.hkl_box(rec, 1 / 5.0, include_zero=True)
    k = 1 / O.electron_wavelength(300)
    assert len(O.simulate_one_reference(ph, rows, rec, np.eye(3), k, p).idx) == expected
    if approx:
        assert len(O.simulate_one(O.phase_tables(ph, p), np.eye(3), k, p).idx) == expected


def test_si_count_needs_the_duplicated_direct_beam():
    # planted mistake: without the second (000) row the 70-spot KAT fails
    ph = _si()
    base, rec = ph.basis()
    p = O.SimParams(kv=300, reciprocal_radius=5.0, with_direct_beam=False)
    rows = O.hkl_box(rec, 1 / 5.0, include_zero=True)
    k = 1 / O.electron_wavelength(300)
    assert len(O.simulate_one_reference(ph, rows, rec, np.eye(3), k, p).idx) == 69


def test_direct_beam_strongest_and_absent():
    # :177-208
    ph = _si()
    base, rec = ph.basis()
    k = 1 / O.electron_wavelength(300)
    p = O.SimParams(kv=300, reciprocal_radius=0.5, with_direct_beam=True)
    tab = O.phase_tables(ph, p)
    sp = O.simulate_one(tab, np.eye(3), k, p)
    zero = np.all(tab.hkl[sp.idx] == 0, axis=1)
    assert zero.sum() == 2 and sp.intensity[zero].max() >= sp.intensity.max()
    p = O.SimParams(kv=300, reciprocal_radius=0.5, with_direct_beam=False)
    tab = O.phase_tables(ph, p)
    sp = O.simulate_one(tab, np.eye(3), k, p)
    assert not np.all(tab.hkl[sp.idx] == 0, axis=1).any()


def test_appropriate_scaling():
    # :170-186: the (220) position halves when the cell doubles
    k = 1 / O.electron_wavelength(300)
    out = []
    for a in (5.0, 10.0):
        ph = _si(a)
        p = O.SimParams(kv=300, reciprocal_radius=5.0)
        tab = O.phase_tables(ph, p)
        sp = O.simulate_one(tab, np.eye(3), k, p)
        rows = tab.hkl[sp.idx]
        i = int(np.nonzero(np.all(rows == [2, 2, 0], axis=1))[0][0])
        out.append(sp.xyz[i])
    np.testing.assert_allclose(out[0], 2 * out[1])


# -- a12, a13: golden pattern and renderer KATs -------------------------------

def _graphite():
    return O.PhaseSpec((2.464, 2.464, 6.711), (90, 90, 120), ["C"] * 4,
                       np.array([[0, 0, .25], [0, 0, .75], [1 / 3, 2 / 3, .25], [2 / 3, 1 / 3, .75]]),
                       np.ones(4), np.zeros(4))


def _golden_pattern(M_fn=O.bunge_matrix):
    ph = _graphite()
    p = O.SimParams(kv=200, reciprocal_radius=1.6768, max_excitation_error=0.1,
                    shape_model="lorentzian", with_direct_beam=False)
    M = M_fn(np.deg2rad(np.array([0.0, 90.0, 90.0])))
    k = 1 / O.electron_wavelength(200)
    sp = O.simulate_one(O.phase_tables(ph, p, "reference"), M, k, p)
    return O.render_fast(sp.xyz[:, :2], sp.intensity, (128, 128), 1.4, 0.0262)


def test_golden_old_simulation():
    # test_simulation_generator.py:283-343 (asserts atol 1e-8)
    old = np.load(GOLDEN)
    assert np.abs(_golden_pattern() - old).max() < 1e-10


def test_golden_catches_transposed_rotation():
    # planted mistake: g_lab = M g instead of M^T g must fail the golden test
    old = np.load(GOLDEN)
    bad = _golden_pattern(lambda e: O.bunge_matrix(e).T)
    assert np.abs(bad - old).max() > 0.1


def test_render_total_intensity_and_edges():
    # test_detector_functions.py:236-306
    p = O.render_fast(np.array([[10.0 - 25, 10.0 - 25]]), np.array([100.0]), (50, 50), 3, 1.0,
                      normalize=False)
    assert abs(p.sum() - 100) < 1e-9
    # float coordinates in a corner lose intensity, integer ones keep it
    out = O.subpixel_gaussian(np.array([[0.3, 0.1]]), np.array([100.0]), (50, 50), 3)
    assert out.sum() < 100
    # low intensity below the clip threshold renders nothing
    out = O.subpixel_gaussian(np.array([[10.0, 10.0], [20.0, 30.0]]), np.ones(2), (50, 50), 1)
    assert out.sum() == 0.0
    # large intensity: the clipped box keeps > 99.9 % of the total
    out = O.subpixel_gaussian(np.array([[10.0, 10.0]]), np.array([1000.0]), (50, 50), 1)
    assert out.sum() / 1000 > 0.999


@pytest.mark.parametrize("sigma", [1.7, 2.5, 3.2])
def test_subpixel_normalisation_at_sigma(sigma):
    # detector_functions.py:333-358: each spot is I/(2 pi sigma^2) exp(-d^2/2 sigma^2)
    # over its clipped box.  With a tiny clip threshold and the spot far from
    # the edges the pixel sum is I (the lattice sum of a normalised Gaussian of
    # sigma >= 1.7 px equals its integral to exp(-2 pi^2 sigma^2) < 1e-24), and
    # the peak at an integer centre is I/(2 pi sigma^2): away from sigma = 1 a
    # prefactor in sigma instead of sigma^2 fails both.
    I = 1000.0
    for cx, cy in ((40.0, 40.0), (40.3, 39.6)):
        out = O.subpixel_gaussian(np.array([[cx, cy]]), np.array([I]), (80, 80), sigma,
                                  clip_threshold=1e-9)
        assert abs(out.sum() / I - 1) < 1e-6, (sigma, cx, cy)
    out = O.subpixel_gaussian(np.array([[40.0, 40.0]]), np.array([I]), (80, 80), sigma,
                              clip_threshold=1e-9)
    assert abs(out[40, 40] - I / (2 * np.pi * sigma ** 2)) < 1e-12 * I
    # the box edge: the first pixel past rad = sqrt(2 sigma^2 ln(A/clip)) is empty
    A = I / (2 * np.pi * sigma ** 2)
    rad = np.sqrt(2 * sigma ** 2 * np.log(A / 1e-3))
    out = O.subpixel_gaussian(np.array([[40.0, 40.0]]), np.array([I]), (80, 80), sigma,
                              clip_threshold=1e-3)
    r_in, r_out = int(np.floor(rad)), int(np.floor(rad)) + 1
    assert out[40, 40 + r_in] > 0 and out[40, 40 + r_out] == 0.0


def test_gaussian_filter_is_reflected_tap_sum():
    # the renderer's closed form: gaussian_filter of a delta = outer product of
    # reflected 1-D taps (checked against scipy on small frames, edges included)
    from scipy import ndimage as ndi
    sigma = 1.4
    R = int(4 * sigma + 0.5)
    x = np.arange(-R, R + 1)
    w = np.exp(-0.5 * x ** 2 / sigma ** 2)
    w /= w.sum()

    def K(y, c, n):
        t = 0.0
        for img in (c, -1 - c, 2 * n - 1 - c):
            if abs(img - y) <= R:
                t += w[R + abs(img - y)]
        return t
    H, W = 9, 12
    for (cy, cx) in [(0, 0), (4, 6), (8, 11), (1, 10)]:
        d = np.zeros((H, W))
        d[cy, cx] = 1.0
        ref = ndi.gaussian_filter(d, sigma)
        mine = np.array([[K(y, cy, H) * K(xx, cx, W) for xx in range(W)] for y in range(H)])
        assert np.abs(ref - mine).max() < 1e-15


def test_polar_flatten_oracle_nearest_index_rules():
    """get_closest semantics: ties go to the upper neighbour, values past either
    end clamp, and the last index is dropped by polar_flatten."""
    ax = np.array([0.0, 1.0, 2.0, 3.0])
    v = np.array([-5.0, 0.4, 0.5, 0.6, 1.5, 2.49, 2.5, 9.0])
    assert O.nearest_indices(ax, v).tolist() == [0, 0, 1, 1, 2, 2, 3, 3]
    xyz = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0], [3.0, 0.0, 0.0]])
    r, t, I = O.polar_flatten([(xyz, np.array([1.0, 2.0, 3.0])), (xyz[:1], np.array([5.0]))],
                              ax, np.linspace(-np.pi, np.pi, 9))
    assert r.dtype.kind == "i" and r.shape == (2, 3)
    # the spot at r = 3 (index 3 = last) is dropped; theta pi/2 -> index 6 of 9
    assert r[0].tolist() == [1, 2, 0] and t[0].tolist() == [4, 6, 0] and I[0].tolist() == [1, 2, 0]
    assert I[1].tolist() == [5, 0, 0]
    rr, tt, _ = O.polar_flatten([(xyz, np.ones(3))])
    assert np.allclose(rr[0], [1, 2, 3]) and np.allclose(tt[0], [0, np.pi / 2, 0])


def test_precession_integral_closed_forms():
    """The converged precession average equals the closed forms the device
    uses: Re[1/sqrt((1 + i sigma s)^2 + sigma^2 A^2)] for the Lorentzian and
    the piecewise-exact average of the linear model."""
    w, phi = 0.02, np.deg2rad(1.0)
    sg = np.pi / w
    rng = np.random.default_rng(3)
    s = rng.uniform(-0.03, 0.03, 12)
    r = rng.uniform(0.2, 2.5, 12)
    A = r * phi
    lor = O.sf_precession_integral(s, r, phi, O.sf_lorentzian, w)
    closed = np.array([(1 / np.sqrt(complex((1 + 1j * sg * si) ** 2 + (sg * ai) ** 2))).real
                       for si, ai in zip(s, A)])
    assert np.allclose(lor, closed, rtol=1e-9, atol=0)
    lin = O.sf_precession_integral(s, r, phi, O.sf_linear, w)
    t = np.linspace(0, np.pi, 200001)
    brute = [np.trapezoid(np.maximum(1 - np.abs(si + ai * np.cos(t)) / w, 0), t) / np.pi
             for si, ai in zip(s, A)]
    assert np.allclose(lin, brute, rtol=1e-7, atol=1e-12)


# ---------------------------------------------------------------------------
# beam-direction grids: the reference's own known answers
# ---------------------------------------------------------------------------

def test_sphere_mesh_kats():
    """test_sphere_mesh_generators.py:48-78: mesh sizes at 10 deg, zero sum,
    no duplicates, and the four directions every cube mesh contains."""
    uv = O.uv_sphere_vertices(10)
    assert uv.shape == (614, 3)
    assert abs(uv.sum()) < 1e-7
    assert np.unique(uv, axis=0).shape[0] == 614
    want = np.array([[1, 1, 1], [1, 0, 0], [1, 1, 0], [-1, 1, -1]], dtype=float)
    want = np.round(want / np.linalg.norm(want, axis=1)[:, None], 13).tolist()
    for grid_type, n in (("normalized", 866), ("spherified_edge", 602), ("spherified_corner", 866)):
        g = O.cube_mesh_vertices(10, grid_type)
        assert g.shape == (n, 3)
        assert abs(g.sum()) < 1e-7
        assert np.unique(g, axis=0).shape[0] == n
        rounded = np.round(g, 13).tolist()
        for w in want:
            assert w in rounded
    with pytest.raises(ValueError):
        O.cube_mesh_vertices(10, "non_existant")


@pytest.mark.parametrize("crystal_system, size", [
    ("cubic", 300), ("hexagonal", 1050), ("trigonal", 1657), ("tetragonal", 852),
    ("orthorhombic", 1657), ("monoclinic", 6441), ("triclinic", 12698)])
def test_beam_directions_grid_size_kat(crystal_system, size):
    """test_rotation_list_generator.py:80-94 (default mesh, 2 deg)."""
    assert O.beam_directions_grid(crystal_system, 2).shape == (size, 3)


def test_directions_to_euler_kat():
    """test_sphere_mesh_generators.py:124-143."""
    v = np.array([[1, 0, 0], [0, 1, 0], [0, 1, 1], [1, 0, 1]], dtype=float)
    v /= np.linalg.norm(v, axis=1)[:, None]
    assert np.allclose(O.directions_to_euler(v), [[0, 90, 90], [0, 90, 0], [0, 45, 0], [0, 45, 90]])


def test_beam_directions_euler_is_the_direction():
    """Closed form: with g_lab = M^T g, the crystal direction along the beam
    (lab z) of each grid orientation is M e_z, and every such direction lies
    in the cubic standard triangle 001-101-111."""
    e = O.beam_directions_grid("cubic", 3)
    M = O.bunge_matrix(np.deg2rad(e))
    d = M[:, :, 2]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0)
    assert np.all(d[:, 2] >= d[:, 0] - 1e-12) and np.all(d[:, 0] >= d[:, 1] - 1e-12)
    assert np.all(d[:, 1] >= -1e-12)


# ---------------------------------------------------------------------------
# icosahedral and random sphere meshes
# ---------------------------------------------------------------------------

def test_icosahedral_mesh_kat():
    """test_sphere_mesh_generators.py:115-121: 642 unique unit vertices summing
    to zero at 10 degrees (level 8 = 10 n^2 + 2)."""
    g = O.icosahedral_mesh_vertices(10)
    assert g.shape == (642, 3)
    np.testing.assert_almost_equal(np.sum(g), 0)
    assert np.unique(g, axis=0).shape[0] == 642
    assert np.allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-15)
    n, seen = O.icosahedral_level(10)
    assert n == 8 and seen[-1][1] <= 10 < seen[-2][1]


def test_icosahedral_bare_angle_is_arctan2():
    """Neighbouring corners of an icosahedron subtend arctan(2); the refinement
    starts there and its first level is the bare icosahedron again."""
    _, seen = O.icosahedral_level(40)
    assert abs(seen[0][1] - np.degrees(np.arctan(2.0))) < 1e-12
    assert seen[1] == (1, seen[0][1])
    assert [lv for lv, _ in seen] == [0, 1, 2]
    with pytest.raises(ValueError):
        O.icosahedral_level(64)


@pytest.mark.parametrize("n", [2, 3, 4, 7])
def test_icosahedral_points_are_the_face_lattice(n):
    """Every flat vertex is f0 + (f2 - f0) a/n + (f1 - f0) b/n for some face and
    a + b <= n (brute force over all faces), each lattice point appears once,
    and the count is 10 n^2 + 2."""
    v = O.icosa_refined_vertices(n)
    C = O.ICOSA_CORNERS.astype(np.float64)
    pts = set()
    for f in O.ICOSA_FACES:
        for a in range(n + 1):
            for b in range(n + 1 - a):
                p = C[f[0]] + (C[f[2]] - C[f[0]]) * a / n + (C[f[1]] - C[f[0]]) * b / n
                pts.add(tuple(np.round(p, 9)))
    got = [tuple(np.round(r, 9)) for r in v]
    assert len(v) == 10 * n * n + 2 == len(pts) == len(set(got))
    assert set(got) == pts


def test_icosahedral_order_kat():
    """Row order: corners, then edge points in edge order from the lower
    corner, then face interiors with i (weight on f2) outer and j (weight on f1)
    inner.  Spelled out for level 4, face 0 = (0, 11, 5)."""
    n = 4
    v = O.icosa_refined_vertices(n)
    C = O.ICOSA_CORNERS.astype(np.float64)
    assert np.array_equal(v[:12], C)
    e0 = O.icosa_edge_order()[0]
    assert np.allclose(v[12], 0.75 * C[e0[0]] + 0.25 * C[e0[1]], atol=1e-15)
    assert np.allclose(v[14], 0.25 * C[e0[0]] + 0.75 * C[e0[1]], atol=1e-15)
    first_face = 12 + 30 * (n - 1)
    rows = [(1, 1), (1, 2), (2, 1)]  # (i, j) for n = 4: 3 interior points per face
    for k, (i, j) in enumerate(rows):
        want = C[0] * (1 - i / n - j / n) + C[11] * (j / n) + C[5] * (i / n)
        assert np.allclose(v[first_face + k], want, atol=1e-15), (i, j)
    # planted mistake: weights on f1 and f2 swapped moves the second point
    wrong = C[0] * (1 - 1 / n - 2 / n) + C[11] * (1 / n) + C[5] * (2 / n)
    assert not np.allclose(v[first_face + 1], wrong, atol=1e-6)


def test_icosahedral_edge_order_is_the_set_order():
    """The 30 edges are the sorted corner pairs of the faces, each once."""
    e = O.icosa_edge_order()
    assert len(e) == 30 == len(set(e)) and all(p < q for p, q in e)
    want = {tuple(sorted(p)) for f in O.ICOSA_FACES for p in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0]))}
    assert set(e) == want


def test_icosahedral_angles_fall_with_level():
    _, seen = O.icosahedral_level(3)
    ang = [a for lv, a in seen[1:]]
    assert all(b < a for a, b in zip(ang, ang[1:]))


def test_random_sphere_mesh_kats():
    """test_sphere_mesh_generators.py:33-45."""
    g = O.random_sphere_vertices(1)
    assert g.shape == (10313, 3)
    g7 = O.random_sphere_vertices(3, seed=7)
    assert np.allclose(g7, O.random_sphere_vertices(3, seed=7))
    assert not np.allclose(g7, O.random_sphere_vertices(3, seed=8))
    assert np.allclose(np.linalg.norm(g7, axis=1), 1.0)


@pytest.mark.parametrize("mesh", ["icosahedral", "random"])
@pytest.mark.parametrize("crystal_system", ["cubic", "hexagonal", "trigonal", "tetragonal",
                                            "orthorhombic", "monoclinic", "triclinic"])
def test_beam_directions_grid_other_meshes_run(crystal_system, mesh):
    """test_rotation_list_generator.py:53-77 for the icosahedral and random
    meshes: every crystal system at 5 degrees."""
    g = O.beam_directions_grid(crystal_system, 5, mesh=mesh)
    assert g.ndim == 2 and g.shape[1] == 3 and g.shape[0] > 0
