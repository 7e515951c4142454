"""The SO(3) grid restatement of the oracle against closed forms, the
reference's own test of these grids and the mathematics of the maps (CPU
only), and the host side of the package's SO(3) grid functions.

The reference samples these grids with orix (absent here), on the cubochoric
grid of Rosca, Morawiec & De Graef (2014).  What fixes the restatement:
* the cube -> ball map is volume-preserving (Jacobian determinant 1), sends
  the cube's surface onto the ball of radius (3 pi / 4)^(1/3), is continuous
  across the faces of its six pyramids and odd; a wrong constant fails the
  first two (planted below);
* the homochoric angle solves |h|^3 = (3/4)(w - sin w);
* the Euler conversions invert each other and reproduce the passive Bunge
  matrix (which the golden pattern pins);
* fundamental zones: the group orders, tiling of the grid by the zone's
  images, and Mackenzie's maximal disorientation angles;
* get_grid_around_beam_direction: test_rotation_list_generator.py:47-54.
"""
import numpy as np
import pytest

from oracle import diffsims_oracle as O


def _jacobian_det(f, p, h=1e-6):
    J = np.zeros((3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        J[:, k] = (f(p + e)[0] - f(p - e)[0]) / (2 * h)
    return np.linalg.det(J)


def _pyramid_points(rng, n):
    """Points spread over all six pyramids of the cube (away from faces)."""
    a = O.CUBO_SEMI_EDGE * 0.97
    pts = rng.uniform(-a, a, size=(n, 3))
    dom = rng.integers(0, 3, n)
    sgn = rng.choice([-1.0, 1.0], n)
    mag = np.abs(pts).max(axis=1) * rng.uniform(1.0, 1.02, n)
    pts[np.arange(n), dom] = sgn * np.minimum(mag, a)
    return pts


def test_cubochoric_map_preserves_volume():
    rng = np.random.default_rng(11)
    for p in _pyramid_points(rng, 300):
        assert abs(_jacobian_det(O.cubochoric_to_homochoric, p) - 1.0) < 1e-6, p


def test_cubochoric_map_sends_the_cube_surface_onto_the_ball():
    rng = np.random.default_rng(12)
    a = O.CUBO_SEMI_EDGE
    pts = rng.uniform(-a, a, size=(2000, 3))
    pts[np.arange(2000), rng.integers(0, 3, 2000)] = rng.choice([-a, a], 2000)
    r = np.linalg.norm(O.cubochoric_to_homochoric(pts), axis=1)
    assert np.abs(r - O.HOMO_RADIUS).max() < 1e-13
    # face centres, edges and corners keep their directions
    for v in ([0, 0, a], [a, 0, 0], [0, -a, 0], [a, a, 0], [a, a, a], [-a, a, -a]):
        h = O.cubochoric_to_homochoric(np.array(v))[0]
        u = np.asarray(v) / np.linalg.norm(v)
        np.testing.assert_allclose(h, O.HOMO_RADIUS * u, atol=1e-13)
    # the volume of the cube is pi^2, the volume of the ball too
    assert abs((2 * a) ** 3 - 4 / 3 * np.pi * O.HOMO_RADIUS ** 3) < 1e-12


def test_cubochoric_map_is_continuous_and_odd():
    rng = np.random.default_rng(13)
    a = O.CUBO_SEMI_EDGE * 0.9
    b = rng.uniform(-a, a, size=(2000, 3))
    eps = 1e-12
    for i, j in ((0, 2), (1, 2), (0, 1)):  # pyramid faces |x| = |z|, |y| = |z|, |x| = |y|
        c = b.copy()
        c[:, i] = np.abs(c[:, j]) * np.sign(c[:, i])
        d = np.zeros(3)
        d[i] = eps
        jump = O.cubochoric_to_homochoric(c + d) - O.cubochoric_to_homochoric(c - d)
        assert np.abs(jump).max() < 1e-9
    np.testing.assert_allclose(O.cubochoric_to_homochoric(-b), -O.cubochoric_to_homochoric(b),
                               atol=1e-15)
    # the three axes: h = (6/pi)^(1/3) c on the axis
    for k in range(3):
        v = np.zeros(3)
        v[k] = 0.4
        np.testing.assert_allclose(O.cubochoric_to_homochoric(v)[0], (6 / np.pi) ** (1 / 3) * v,
                                   atol=1e-15)


def test_cubochoric_pins_catch_a_planted_constant(monkeypatch):
    """A 0.1 % error in the ball radius inside the map breaks volume
    preservation and moves the cube's corners off the body diagonals."""
    rng = np.random.default_rng(14)
    good = O.HOMO_RADIUS
    monkeypatch.setattr(O, "HOMO_RADIUS", good * 1.001)
    p = _pyramid_points(rng, 20)
    dets = [_jacobian_det(O.cubochoric_to_homochoric, x) for x in p]
    assert max(abs(d - 1) for d in dets) > 1e-3
    a = O.CUBO_SEMI_EDGE
    h = O.cubochoric_to_homochoric(np.array([a, a, a]))[0]
    assert np.abs(h / np.linalg.norm(h) - 1 / np.sqrt(3)).max() > 1e-5


def test_homochoric_angle_solves_its_equation():
    rng = np.random.default_rng(15)
    hn = np.concatenate([rng.uniform(0, O.HOMO_RADIUS, 5000), [1e-8, 1e-3, 0.2, 0.25,
                                                               O.HOMO_RADIUS]])
    u = rng.normal(size=(hn.size, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    axis, w = O.homochoric_to_axis_angle(u * hn[:, None])
    np.testing.assert_allclose(axis, u, atol=1e-14)
    lhs = 0.75 * np.where(w < 0.5, O._omega_minus_sin(w), w - np.sin(w))
    np.testing.assert_allclose(lhs, hn ** 3, rtol=1e-13, atol=0)
    assert abs(w[-1] - np.pi) < 1e-14
    np.testing.assert_allclose(w[hn < 1e-3], 2 * hn[hn < 1e-3], rtol=1e-6)
    # the series of w - sin w against the direct form where both are accurate
    t = np.linspace(0.3, 0.5, 50)
    np.testing.assert_allclose(O._omega_minus_sin(t - 1e-12), t - 1e-12 - np.sin(t - 1e-12),
                               rtol=1e-13)


def test_euler_quaternion_round_trip_and_branches():
    e = O.random_euler(4000, 16)
    e = e[(e[:, 1] > 1e-3) & (e[:, 1] < np.pi - 1e-3)]
    back = O.bunge_from_quat(O.quat_from_bunge(e))
    d = (back - e + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(d).max() < 1e-10
    assert (back[:, [0, 2]] >= 0).all() and (back[:, [0, 2]] < 2 * np.pi).all()
    # any quaternion: the angles give back its passive Bunge matrix
    rng = np.random.default_rng(17)
    q = rng.normal(size=(3000, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    np.testing.assert_allclose(O.bunge_matrix(O.bunge_from_quat(q)), O.matrix_from_quat(q),
                               atol=1e-12)
    # Phi = 0: the whole rotation about z lands in phi1
    for p1, p2 in ((0.3, 1.1), (5.0, 4.0), (0.0, 0.0)):
        got = O.bunge_from_quat(O.quat_from_bunge(np.array([p1, 0.0, p2])))[0]
        np.testing.assert_allclose(got, [np.mod(p1 + p2, 2 * np.pi), 0, 0], atol=1e-13)
    # the two degenerate branches on exact quaternions: (cos t, 0, 0, sin t) is
    # (-2t, 0, 0) and (0, cos t, sin t, 0) is (2t, pi, 0), both mod 2 pi
    for t in (0.2, 1.3, 2.9, -0.7):
        for q, want in (([np.cos(t), 0, 0, np.sin(t)], [np.mod(-2 * t, 2 * np.pi), 0, 0]),
                        ([0, np.cos(t), np.sin(t), 0], [np.mod(2 * t, 2 * np.pi), np.pi, 0])):
            got = O.bunge_from_quat(np.array(q))
            np.testing.assert_allclose(got[0], want, atol=1e-13)
            np.testing.assert_allclose(O.bunge_matrix(got)[0], O.matrix_from_quat(np.array(q)),
                                       atol=1e-13)


def test_grid_around_beam_direction_kat():
    """test_rotation_list_generator.py:47-54, and the closed form: a rotation
    about z by w after Bunge (phi1, Phi, phi2) is (phi1 - w, Phi, phi2)."""
    rows = O.rotation_list(O.grid_around_beam_direction((0, 90, 0), resolution=2,
                                                        angular_range=(0, 9)))
    assert len(rows) == 5 and isinstance(rows[0], tuple)
    assert np.allclose([r[1] for r in rows], 90)
    e = O.grid_around_beam_direction((10, 40, 70), 7.5)
    w = np.deg2rad(np.arange(0, 360, 7.5))
    want = np.stack([np.mod(np.deg2rad(10) - w, 2 * np.pi), np.full_like(w, np.deg2rad(40)),
                     np.full_like(w, np.deg2rad(70))], axis=1)
    d = (e - want + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(d).max() < 1e-12


# Mackenzie's maximal disorientation angles of the proper groups (degrees)
MACKENZIE = {"432": 62.7994, "622": 93.8377, "422": 98.4211, "32": 104.4775, "222": 120.0}
ORDERS = {"1": 1, "2": 2, "222": 4, "4": 4, "422": 8, "3": 3, "32": 6, "6": 6, "622": 12,
          "23": 12, "432": 24}


@pytest.mark.parametrize("name", sorted(ORDERS))
def test_fundamental_zone_tiles_the_grid(name):
    ops = O.proper_group_quaternions(name)
    assert len(ops) == ORDERS[name]
    # a group: products stay in the set (up to sign)
    prod = O.quat_multiply(ops[:, None, :], ops[None, :, :]).reshape(-1, 4)
    dots = np.abs(prod @ ops.T).max(axis=1)
    assert np.abs(dots - 1).max() < 1e-12
    q = O.cubochoric_quaternions(10)
    margin = O.fundamental_zone_margin(q, ops)
    inside = margin >= -O.FZ_EPS
    # every grid rotation has an equivalent q s^-1 inside the zone
    conj = ops * np.array([1, -1, -1, -1])
    best = np.max([O.fundamental_zone_margin(O.quat_multiply(q, s[None, :]), ops) for s in conj],
                  axis=0)
    assert (best >= -1e-12).all()
    # the zone holds 1/|G| of the (uniform) grid up to its boundary rows
    frac = inside.sum() / q.shape[0]
    assert abs(frac * len(ops) - 1) < 0.12, frac


@pytest.mark.parametrize("name", sorted(MACKENZIE))
def test_fundamental_zone_maximal_angle(name):
    ops = O.proper_group_quaternions(name)
    q = O.cubochoric_quaternions(O.cubochoric_steps(3.0))
    q = q[O.fundamental_zone_margin(q, ops) >= -O.FZ_EPS]
    ang = np.rad2deg(2 * np.arccos(np.clip(np.abs(q[:, 0]), 0, 1)))
    assert ang.max() <= MACKENZIE[name] + 1e-6
    assert ang.max() >= MACKENZIE[name] - 3.0


def test_reference_grid_calls():
    """test_rotation_list_generator.py:30-44: non-empty lists of tuples."""
    for rows in (O.rotation_list(O.local_grid(30, (0, 1, 0), 35)),
                 O.rotation_list(O.fundamental_zone_grid(20, 20))):
        assert isinstance(rows, list) and len(rows) > 0 and isinstance(rows[0], tuple)


def test_local_grid_stays_within_its_width():
    for width in (10.0, 35.0):
        e = O.local_grid(4.0, None, width)
        q = O.quat_from_bunge(e)
        assert (2 * np.rad2deg(np.arccos(np.clip(np.abs(q[:, 0]), 0, 1))) < width + 1e-9).all()
        c = O.quat_from_bunge(np.deg2rad([20.0, 30.0, 40.0]))
        qc = O.quat_from_bunge(O.local_grid(4.0, (20.0, 30.0, 40.0), width))
        rel = O.quat_multiply(c * np.array([1, -1, -1, -1]), qc)
        assert (2 * np.rad2deg(np.arccos(np.clip(np.abs(rel[:, 0]), 0, 1))) < width + 1e-9).all()
        assert len(qc) == len(q)


def test_cubochoric_steps_meet_the_resolution():
    for res in (0.5, 1.0, 2.0, 3.0, 10.0, 20.0):
        n = O.cubochoric_steps(res)
        for m, ok in ((n, True), (n - 1, False)):
            if m < 1:
                continue
            q = O.quat_from_axis_angle(*O.homochoric_to_axis_angle(
                O.cubochoric_to_homochoric(np.array([O.CUBO_SEMI_EDGE / m, 0, 0]))))
            step = np.rad2deg(2 * np.arccos(q[0, 0]))
            assert (step <= res + 1e-9) == ok, (res, m, step)


# -- host side of the package ----------------------------------------------------

def test_host_groups_and_euler_match_the_oracle():
    from diffsims_amd import symmetry
    from diffsims_amd.rotation import Rotation, bunge_from_quaternions

    for name in ORDERS:
        a = symmetry.rotation_group_quaternions(name)
        b = O.proper_group_quaternions(name)
        assert len(a) == len(b)
        assert np.abs(np.abs(a @ b.T).max(axis=1) - 1).max() < 1e-12
    rng = np.random.default_rng(18)
    q = rng.normal(size=(500, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    np.testing.assert_allclose(bunge_from_quaternions(q), O.bunge_from_quat(q), atol=1e-13)
    e = O.random_euler(100, 19)
    np.testing.assert_allclose(Rotation.from_euler(e).to_euler(),
                               O.bunge_from_quat(O.quat_from_bunge(e)), atol=1e-13)
    assert symmetry.rotation_group_of_space_group(20) == "222"
    assert symmetry.rotation_group_of_space_group(225) == "432"
    assert symmetry.rotation_group_of_space_group(186) == "622"


def test_host_grid_around_beam_direction():
    from diffsims_amd import generators

    got = generators.get_grid_around_beam_direction((0, 90, 0), resolution=2,
                                                    angular_range=(0, 9))
    assert len(got) == 5 and isinstance(got[0], tuple)
    assert np.allclose([x[1] for x in got], 90)
    for beam, res, rng in (((10, 40, 70), 7.5, (0, 360)), ((123.4, 77.7, 301.2), 1, (5, 95))):
        assert generators.get_grid_around_beam_direction(beam, res, rng) == \
            O.rotation_list(O.grid_around_beam_direction(beam, res, rng))
    assert generators.get_list_from_orix(np.deg2rad([[10.004, 20.0, 30.0]])) == \
        [(10.0, 20.0, 30.0)]
    assert generators.cubochoric_steps(2.0) == O.cubochoric_steps(2.0)
    with pytest.raises(ValueError):
        generators.cubochoric_steps(0.1)


def test_so3_grid_queries_through_the_abi():
    from diffsims_amd import _lib

    L = _lib.load()
    assert L.dsx_so3_grid_candidates(8) == 16 ** 3
    assert L.dsx_so3_grid_candidates(0) == -1 and L.dsx_so3_grid_candidates(1025) == -1
    nb = (16 ** 3 + 255) // 256
    assert L.dsx_so3_grid_workspace_ints(8) == ((nb + 1) & ~1) + 2 * nb
