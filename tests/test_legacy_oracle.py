"""Pins of the oracle's legacy-generator restatement and the legacy host
containers (CPU).

The legacy DiffractionGenerator works in diffpy's standard frame (a* || x,
c || z, b in the yz-plane), rotates with transforms3d's rzxz Euler matrices
and takes its reflections from get_points_in_sphere.  None of diffpy or
transforms3d can be imported here; the frame is pinned by the conditions that
define it, the rotation by the identities of intrinsic z-x-z rotations, the
reflection set by brute force, and the generator by the reference's own
known answers (diffsims/tests/generators/test_diffraction_generator.py:171-199).
"""

import math

import numpy as np
import pytest

from oracle import diffsims_oracle as O

CELLS = [(5.0, 6.0, 7.0, 80.0, 95.0, 100.0), (5.431, 5.431, 5.431, 90.0, 90.0, 90.0),
         (3.0, 3.0, 5.0, 90.0, 90.0, 120.0), (19.5, 20.5, 21.0, 88.0, 96.0, 91.0),
         (8.1, 5.2, 9.3, 90.0, 104.0, 90.0)]


@pytest.mark.parametrize("cell", CELLS)
def test_legacy_frame_conditions(cell):
    """Rows give back a, b, c and the three angles; c || +z, b in the
    yz-plane, a* || +x; right-handed; equals the closed form the product
    code uses (diffpy's stdbase written with the reciprocal cell)."""
    from diffsims_amd import engine

    a, b, c, al, be, ga = cell
    B = O.diffpy_base_from_conditions(*cell)
    n = np.linalg.norm(B, axis=1)
    assert np.allclose(n, [a, b, c], rtol=0, atol=1e-12)
    ang = lambda u, v: math.degrees(math.acos(u @ v / np.linalg.norm(u) / np.linalg.norm(v)))
    assert abs(ang(B[1], B[2]) - al) < 1e-10
    assert abs(ang(B[0], B[2]) - be) < 1e-10
    assert abs(ang(B[0], B[1]) - ga) < 1e-10
    assert B[2, 0] == 0 and B[2, 1] == 0 and B[1, 0] == 0
    astar = np.linalg.inv(B)[:, 0]
    assert abs(astar[1]) < 1e-15 and abs(astar[2]) < 1e-15 and astar[0] > 0
    assert np.linalg.det(B) > 0
    assert np.abs(engine.diffpy_standard_base(cell[:3], cell[3:]) - B).max() < 1e-12 * max(cell[:3])
    # planted mistake: the orix frame (x || a) differs for an oblique cell
    if al != 90 or be != 90 or ga != 90:
        base, _ = O.lattice_basis(*cell)
        assert np.abs(base - B).max() > 1e-3


def test_rzxz_identities():
    """Intrinsic z-x'-z'': identity at zero, Rz(a + c) with no tilt, Rx(b)
    alone, composition M(a,b,c) = M(a,0,0) M(0,b,0) M(0,0,c), orthonormal;
    the static-axes order Rz(c) Rx(b) Rz(a) differs (planted mistake)."""
    d = np.deg2rad
    assert np.allclose(O.rzxz_matrix(0, 0, 0), np.eye(3))
    a, b, c = d(23.0), d(61.0), d(-140.0)
    Rz = lambda t: O.rzxz_matrix(t, 0, 0)
    Rx = lambda t: O.rzxz_matrix(0, t, 0)
    assert np.allclose(O.rzxz_matrix(a, 0, c), Rz(a + c), atol=1e-15)
    assert np.allclose(Rx(b), [[1, 0, 0], [0, math.cos(b), -math.sin(b)],
                               [0, math.sin(b), math.cos(b)]], atol=1e-15)
    M = O.rzxz_matrix(a, b, c)
    assert np.allclose(M, Rz(a) @ Rx(b) @ Rz(c), atol=1e-15)
    assert np.allclose(M @ M.T, np.eye(3), atol=1e-15) and abs(np.linalg.det(M) - 1) < 1e-14
    assert not np.allclose(M, Rz(c) @ Rx(b) @ Rz(a), atol=1e-3)
    from diffsims_amd.legacy import euler2mat_rzxz
    e = np.array([[23.0, 61.0, -140.0], [0.0, 90.0, 0.0], [300.0, 10.0, 45.0]])
    for row, Md in zip(e, euler2mat_rzxz(e)):
        assert np.abs(Md - O.rzxz_matrix(*np.deg2rad(row))).max() < 1e-15


@pytest.mark.parametrize("cell", CELLS[:4])
def test_points_in_sphere_brute_force(cell):
    """Every point of a larger box with |g| < r (Cartesian norm) and inside
    the reference's box, in ascending lexicographic order, (000) once."""
    r = 1.3
    hkl, g, d = O.legacy_points_in_sphere(cell[:3], cell[3:], r)
    B = O.diffpy_base_from_conditions(*cell)
    rec = np.linalg.inv(B)
    H = [int(np.floor(r / v)) for v in np.sqrt(np.sum(rec * rec, axis=0))]
    big = np.stack(np.meshgrid(*[np.arange(-h - 3, h + 4) for h in H], indexing="ij"),
                   -1).reshape(-1, 3)
    gb = big @ rec.T
    inb = np.all(np.abs(big) <= np.array(H), axis=1) & (np.linalg.norm(gb, axis=1) < r - 1e-12)
    assert set(map(tuple, big[inb])) <= set(map(tuple, hkl))
    assert len(hkl) - inb.sum() <= 0 + np.sum(np.abs(np.linalg.norm(hkl @ rec.T, axis=1) - r) < 1e-9)
    keys = [tuple(h) for h in hkl]
    assert keys == sorted(keys)
    assert sum(k == (0, 0, 0) for k in keys) == 1
    assert np.allclose(g, hkl @ rec.T) and np.all(d < r)


def test_points_in_sphere_strict_boundary():
    """Cubic a = 5, r = 1: (5,0,0) lies exactly on the sphere and is left out."""
    hkl, _, _ = O.legacy_points_in_sphere((5.0,) * 3, (90.0,) * 3, 1.0)
    keys = set(map(tuple, hkl))
    assert (4, 0, 0) in keys and (4, 2, 2) in keys
    assert (5, 0, 0) not in keys and (4, 3, 0) not in keys and (3, 4, 0) not in keys


def _si(a=5.431):
    return O.PhaseSpec((a,) * 3, (90.0,) * 3, ["Si"] * 8,
                       np.array([[0, 0, 0], [0.5, 0, 0.5], [0, 0.5, 0.5], [0.5, 0.5, 0],
                                 [0.25, 0.25, 0.25], [0.75, 0.25, 0.75], [0.25, 0.75, 0.75],
                                 [0.75, 0.75, 0.25]], dtype=float), np.ones(8), np.zeros(8))


def test_legacy_reference_kats():
    """test_diffraction_generator.py:171-199 on the oracle: (2,2,0) present
    for a = 5 and a = 10 with twice the coordinates, the direct beam the
    strongest spot."""
    p = O.SimParams(kv=300, reciprocal_radius=5.0, max_excitation_error=1e-2,
                    shape_model="lorentzian", with_direct_beam=True)
    k = 1 / O.electron_wavelength(300)
    res = []
    for a in (5.0, 10.0):
        tab = O.legacy_tables(_si(a), p)
        sp = O.legacy_simulate_one(tab, (0, 0, 0), k, p)
        idx = [tuple(tab.hkl[i]) for i in sp.idx]
        assert (2, 2, 0) in idx
        res.append(sp.xyz[idx.index((2, 2, 0))])
        c = idx.index((0, 0, 0))
        assert np.all(sp.intensity[c] >= sp.intensity)
    assert np.allclose(res[0], res[1] * 2)


def test_legacy_profile_families():
    """Si profile: families by permutation and sign, multiplicity x |F|^2,
    (111) the strongest line and (200) extinct; the hexagonal lattice keeps
    the reference's four-row label step."""
    p = O.SimParams(kv=300, reciprocal_radius=1.0, with_direct_beam=True)
    x, y, labels = O.legacy_profile(_si(), p)
    assert labels[int(np.argmax(y))] in ("111",) and abs(max(y) - 100) < 1e-12
    assert "200" not in labels and "000" not in labels
    assert np.all(np.diff(sorted(x)) >= 0)
    ni = O.PhaseSpec((3.0, 3.0, 5.0), (90.0, 90.0, 120.0), ["Ni"], np.zeros((1, 3)), np.ones(1),
                     np.zeros(1))
    x, y, labels = O.legacy_profile(ni, p, hexagonal=True)
    assert len(labels) <= 4


def test_unique_families_matches_oracle_grouping():
    from diffsims_amd.legacy import unique_families

    hkl, _, _ = O.legacy_points_in_sphere((5.431,) * 3, (90.0,) * 3, 1.2)
    fam = unique_families(hkl.tolist())
    assert sum(fam.values()) == len(hkl)
    assert fam[(0, 0, 0)] == 1 and fam[(1, 1, 1)] == 8 and fam[(2, 0, 0)] == 6
    assert fam[(3, 1, 1)] == 24


def test_legacy_containers():
    from diffsims_amd.legacy import (DiffractionLibrary, DiffractionSimulation,
                                     StructureLibrary)

    coords = np.array([[0.0, 0.0, 0.0], [0.1, 0.2, 0.0], [-0.3, 0.1, 0.0]])
    s = DiffractionSimulation(coords, indices=np.zeros((3, 3)), intensities=np.array([9.0, 1, 2]))
    assert len(s) == 2 and s.size == 2                         # direct beam masked
    s2 = DiffractionSimulation(coords, indices=np.zeros((3, 3)), intensities=np.ones(3),
                               with_direct_beam=True)
    assert len(s2) == 3
    with pytest.raises(Exception, match="calibration"):
        _ = s.calibrated_coordinates
    s.calibration = 0.05
    assert np.allclose(s.calibrated_coordinates, coords[1:, :2] / 0.05)
    s.offset = np.array([0.1, 0.0])
    assert np.allclose(s.calibrated_coordinates, (coords[1:, :2] + [0.1, 0]) / 0.05)
    with pytest.raises(ValueError):
        s.calibration = 0
    with pytest.raises(ValueError):
        s.calibration = (1, 2, 3)
    with pytest.raises(ValueError):
        DiffractionSimulation(coords, indices=np.zeros((2, 3)), intensities=np.ones(3))
    assert len(s[0]) == 1
    t = s.deepcopy()
    t.rotate_shift_coordinates(90)
    assert np.allclose(t.coordinates[:, :2], np.stack([-coords[1:, 1], coords[1:, 0]], 1))
    with pytest.raises(ValueError):
        StructureLibrary(["a", "b"], [1], [[(0, 0, 0)]])
    with pytest.raises(ValueError):
        StructureLibrary(["a"], [1], [[(0, 0, 0)], [(1, 1, 1)]])
    sl = StructureLibrary(["a", "b"], [1, 2], [[(0, 0, 0)], [(0, 0, 0), (1, 2, 3)]])
    assert sl.get_library_size() == 3
    lib = DiffractionLibrary()
    lib["a"] = {"simulations": [s], "orientations": [(0, 0, 0), (10, 20, 30)],
                "pixel_coords": [np.zeros((2, 2))], "intensities": [np.array([3.0, 4.0])]}
    e = lib.get_library_entry("a")
    assert e["pattern_norm"] == 5.0
    with pytest.raises(ValueError):
        lib.get_library_entry(angle=(0, 0, 0))
    with pytest.raises(ValueError):
        lib.get_library_entry("a", angle=(5, 5, 5))


# ---------------------------------------------------------------------------
# point groups of the 1-D profile (host)
# ---------------------------------------------------------------------------

ORDERS = {"1": 1, "-1": 2, "2": 2, "m": 2, "2/m": 4, "222": 4, "mm2": 4, "mmm": 8, "4": 4,
          "-4": 4, "4/m": 8, "422": 8, "4mm": 8, "-42m": 8, "4/mmm": 16, "3": 3, "-3": 6,
          "32": 6, "3m": 6, "-3m": 12, "6": 6, "-6": 6, "6/m": 12, "622": 12, "6mm": 12,
          "-6m2": 12, "6/mmm": 24, "23": 12, "m-3": 24, "432": 24, "-43m": 24, "m-3m": 48}


@pytest.mark.parametrize("pg", sorted(ORDERS))
def test_point_group_orders_and_closure(pg):
    """Group orders of the 32 crystallographic point groups; closed under
    products; orthogonal; proper rotations only where the symbol says so."""
    from diffsims_amd.symmetry import point_group_matrices

    G = point_group_matrices(pg)
    assert G.shape == (ORDERS[pg], 3, 3)
    keys = {tuple(np.round(g, 8).ravel()) for g in G}
    for a in G:
        assert np.allclose(a @ a.T, np.eye(3), atol=1e-12)
        for b in G:
            assert tuple(np.round(a @ b, 8).ravel()) in keys
    proper = pg in ("1", "2", "222", "4", "422", "3", "32", "6", "622", "23", "432")
    assert all(np.linalg.det(g) > 0 for g in G) == proper


def test_space_group_to_point_group():
    from diffsims_amd.symmetry import point_group_of_space_group

    assert [point_group_of_space_group(n) for n in (1, 2, 14, 62, 139, 166, 194, 225, 227)] == \
        ["1", "-1", "2/m", "mmm", "4/mmm", "-3m", "6/mmm", "m-3m", "m-3m"]
    with pytest.raises(ValueError):
        point_group_of_space_group(231)


def test_families_al_multiplicities():
    """The docstring KAT of ReciprocalLatticeVector.multiplicity
    (reciprocal_lattice_vector.py:303-310): Al (m-3m) {111} 8, {200} 6; every
    vector in exactly one family; the lattice maps onto itself."""
    from diffsims_amd.symmetry import point_group_matrices, symmetry_families

    base, rec = O.lattice_basis(4.04, 4.04, 4.04, 90, 90, 90)
    hkl = O.hkl_box(rec, 0.7)
    assert len(hkl) == 798
    g = hkl @ rec.T
    rep, mult = symmetry_families(g, "m-3m")
    assert mult.sum() == len(hkl)
    fam = {tuple(sorted(np.abs(hkl[r]))): m for r, m in zip(rep, mult)}
    assert fam[(1, 1, 1)] == 8 and fam[(0, 0, 2)] == 6 and fam[(1, 1, 3)] == 24
    for op in point_group_matrices("m-3m"):
        img = (g @ op.T) @ np.linalg.inv(rec).T
        assert np.allclose(img, np.rint(img), atol=1e-9)
    # hexagonal: 6/mmm orbits of (100) and (001)
    base, rec = O.lattice_basis(3.0, 3.0, 5.0, 90, 90, 120)
    hkl = O.hkl_box(rec, 0.9)
    rep, mult = symmetry_families(hkl @ rec.T, "6/mmm")
    fam = {tuple(hkl[r]): m for r, m in zip(rep, mult)}
    m100 = [m for k, m in fam.items() if sorted(np.abs(k)) in ([0, 0, 1], [0, 1, 1])
            and k[2] == 0]
    assert m100 == [6]                     # the orbit can leave the index box: no sum check
    assert set(m for k, m in fam.items() if k[0] == 0 and k[1] == 0) == {2}
