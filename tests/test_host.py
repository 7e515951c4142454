"""Host-side logic of the drop-in surface (CPU only, no kernel launches).

Mirrors the container and argument tests of
diffsims/tests/simulations/test_simulations2d.py and
diffsims/tests/generators/test_simulation_generator.py.
"""

import numpy as np
import pytest

from diffsims_amd import (Atom, DiffractingVector, Lattice, Phase, Rotation, Simulation2D,
                          SimulationGenerator, Structure, get_closest, shape_factor_models as sfm,
                          uniform_so3_grid)
from diffsims_amd.crystallography import highest_hkl, lattice_setup, phase_arrays
from oracle import diffsims_oracle as O


def al_phase():
    return Phase(name="al", space_group=225, structure=Structure(
        atoms=[Atom("al", [0, 0, 0])], lattice=Lattice(0.405, 0.405, 0.405, 90, 90, 90)))


# -- crystallography ------------------------------------------------------------

@pytest.mark.parametrize("abc, ang", [((4.05,) * 3, (90,) * 3), ((2.464, 2.464, 6.711), (90, 90, 120)),
                                      ((5.1, 6.2, 7.3), (81, 97, 104))])
def test_lattice_setup_matches_oracle_bitwise(abc, ang):
    base, rec = lattice_setup(abc, ang)
    ob, orec = O.lattice_basis(*abc, *ang)
    assert np.array_equal(base, ob) and np.array_equal(rec, orec)
    for d in (0.3, 0.5, 1.0):
        assert np.array_equal(highest_hkl(rec, d), O.highest_hkl(orec, d))


def test_phase_arrays_debye_waller():
    p = Phase("x", structure=Structure([Atom("Fe2+", [0, 0, 0]), Atom("O", [.5, .5, .5])],
                                       Lattice(3, 3, 3)))
    pa = phase_arrays(p, {"Fe": 0.7})
    assert pa.B.tolist() == [0.7, 0.0] and pa.symbols == ["Fe2+", "O"]
    pa = phase_arrays(p, [0.1, 0.2])
    assert pa.B.tolist() == [0.1, 0.2]
    with pytest.raises(ValueError):
        phase_arrays(p, [0.1])


class _RealignedLattice(Lattice):
    """A diffpy-style lattice after orix's re-alignment: ``stdbase`` is the
    basis the atoms were first placed in (the cell in another orientation),
    ``base`` the aligned basis (x || a, z || c*), ``recbase = inv(base)``."""

    def __init__(self, abc, ang, Q):
        super().__init__(*abc, *ang)
        self._aligned = lattice_setup(abc, ang)[0]
        self.stdbase = self._aligned @ Q
        self.baserot = np.linalg.inv(self.stdbase) @ self._aligned

    @property
    def base(self):
        return self._aligned

    @property
    def recbase(self):
        return np.linalg.inv(self._aligned)


def test_phase_arrays_undo_orix_realignment():
    # sim_utils.py:288-291: positions of a re-aligned structure are mapped back
    # with inv(stdbase @ recbase); orix keeps the Cartesian positions, so the
    # re-aligned fractional coordinates are xyz0 @ stdbase @ inv(base)
    abc, ang = (5.1, 6.3, 7.7), (80.0, 95.0, 105.0)
    Q = O.bunge_matrix(np.array([0.3, 1.1, -0.7]))
    lat = _RealignedLattice(abc, ang, Q)
    xyz0 = np.random.default_rng(4).uniform(0, 1, (5, 3))
    moved = xyz0 @ lat.stdbase @ np.linalg.inv(lat.base)
    assert np.abs(moved - xyz0).max() > 0.05                 # the re-alignment moved them
    p = Phase("x", structure=Structure([Atom("Fe", v) for v in moved], lat))
    pa = phase_arrays(p)
    assert np.abs(pa.xyz - xyz0).max() < 1e-12
    # a lattice never re-aligned (stdbase = base) keeps the positions
    lat2 = _RealignedLattice(abc, ang, np.eye(3))
    pa = phase_arrays(Phase("y", structure=Structure([Atom("Fe", v) for v in xyz0], lat2)))
    assert np.abs(pa.xyz - xyz0).max() < 1e-12
    # this package's own lattice has no stdbase: positions as given
    pa = phase_arrays(Phase("z", structure=Structure([Atom("Fe", v) for v in xyz0],
                                                     Lattice(*abc, *ang))))
    assert np.array_equal(pa.xyz, xyz0)


def test_phase_arrays_reject_unrecoverable_base_rotation():
    class Rotated(Lattice):
        baserot = O.bunge_matrix(np.array([0.2, 0.4, 0.1]))

    p = Phase("x", structure=Structure([Atom("Fe", [0.1, 0.2, 0.3])], Rotated(3, 4, 5)))
    with pytest.raises(ValueError, match="base rotation"):
        phase_arrays(p)


# -- rotations --------------------------------------------------------------------

def test_rotation_conventions_match_oracle():
    e = O.random_euler(50, 2)
    r = Rotation.from_euler(e)
    assert np.abs(r.to_matrix() - O.bunge_matrix(e)).max() < 1e-14
    assert np.abs(Rotation.from_matrix(r.to_matrix()).data - r.data).max() < 1e-12
    np.testing.assert_allclose(r.to_euler(), np.mod(e, 2 * np.pi) * [1, 0, 1] + e * [0, 1, 0],
                               atol=1e-10)
    v = np.array([0.3, -1.2, 0.8])
    one = r[0]
    assert np.allclose(one * v, one.to_matrix()[0] @ v)
    assert np.allclose(~one * v, one.to_matrix()[0].T @ v)
    assert np.allclose((one * ~one).to_matrix(), np.eye(3))


def test_rotation_helpers():
    r = Rotation.from_axes_angles([1, 0, 0], (0, 15, 30, 45), degrees=True)
    assert r.size == 4 and r.shape == (4,)
    assert np.allclose(r[1] * np.array([0, 1, 0]), [0, np.cos(np.pi / 12), np.sin(np.pi / 12)])
    assert Rotation.identity().size == 1 and len(list(r)) == 4
    g = uniform_so3_grid(1000)
    assert g.size == 1000 and np.allclose(np.linalg.norm(g.data, axis=1), 1)
    # near-uniform: |w| of a uniform quaternion has density ~ sqrt(1 - w^2),
    # mean 4 / (3 pi)
    assert abs(np.mean(np.abs(g.data[:, 0])) - 4 / (3 * np.pi)) < 0.002


# -- DiffractingVector ------------------------------------------------------------

def test_diffracting_vector_surface():
    p = al_phase()
    dv = DiffractingVector(p, xyz=[[1, 0, 0], [1, 1, 1]])
    assert dv.size == 2 and np.isnan(dv.intensity).all()
    dv.intensity = 1
    assert dv.intensity.tolist() == [1, 1]
    with pytest.raises(ValueError):
        dv.intensity = [1, 2, 3]
    with pytest.raises(ValueError):
        DiffractingVector(p, xyz=[[1, 0, 0]], intensity=[1, 2])
    assert dv[1].size == 1 and dv[0:2].size == 2
    r, t = dv.to_flat_polar()
    np.testing.assert_allclose(r, [1, np.sqrt(2)])
    np.testing.assert_allclose(t, [0, np.pi / 4])
    hkl = DiffractingVector(p, hkl=[[1, 1, 1], [2, 0, 0]])
    np.testing.assert_allclose(hkl.hkl, [[1, 1, 1], [2, 0, 0]])
    rot = hkl.rotate_with_basis(Rotation.from_euler([[0.2, 0.4, 0.6]]))
    np.testing.assert_allclose(rot.hkl, [[1, 1, 1], [2, 0, 0]], atol=1e-12)
    with pytest.raises(ValueError):
        hkl.rotate_with_basis(Rotation.random(2, seed=0))


# -- Simulation2D containers (test_simulations2d.py) -------------------------------

@pytest.fixture
def single():
    p = al_phase()
    coords = DiffractingVector(p, xyz=[[1, 0, 0], [2, 0, 0], [3, 3, 0], [-4, 0, 0], [-5, 0, 0],
                                       [-6, 0, 0], [-7, 0, 0], [-8, 0, 0]],
                               intensity=[1, 2, 3, 4, 5, 6, 7, 8])
    return Simulation2D(phases=p, simulation_generator=SimulationGenerator(200), coordinates=coords,
                        rotations=Rotation.from_axes_angles([1, 0, 0], 45, degrees=True))


def test_single_simulation(single):
    rotation, phase, coords = single.get_simulation(0)
    assert isinstance(rotation, Rotation) and phase == 0
    with pytest.raises(ValueError):
        single.iphase[0]
    with pytest.raises(ValueError):
        single.irot[0]
    assert sum(1 for _ in single) == 1 and single._num_rotations() == 1
    r, t, i = single.polar_flatten_simulations()
    assert r.shape == t.shape == i.shape == (1, 8)
    r, t, i = single.polar_flatten_simulations(radial_axes=np.linspace(0, 7, 5),
                                               azimuthal_axes=np.linspace(0, 2 * np.pi, 10))
    assert np.all(r[:, 6:] == 0) and np.all(i[:, 6:] == 0)
    assert single.deepcopy() is not single


def test_init_failures():
    p = al_phase()
    gen = SimulationGenerator(200)
    coords = DiffractingVector(p, xyz=[[1, 0, 0], [1, 1, 1]])
    rot = Rotation.from_axes_angles([1, 0, 0], 45, degrees=True)
    with pytest.raises(ValueError):
        Simulation2D(phases=p, simulation_generator=gen, coordinates=[coords, coords], rotations=rot)
    rot2 = Rotation.from_axes_angles([1, 0, 0], (0, 45), degrees=True)
    with pytest.raises(ValueError):
        Simulation2D(phases=p, simulation_generator=gen, coordinates=[coords] * 3, rotations=rot2)
    with pytest.raises(ValueError):
        Simulation2D(phases=[p, p], simulation_generator=gen,
                     coordinates=[[coords, coords], [coords, coords]], rotations=[rot, rot])
    with pytest.raises(ValueError):
        Simulation2D(phases=[p, p], simulation_generator=gen,
                     coordinates=[[coords] * 2] * 3, rotations=[rot, rot, rot])


def _multi():
    p = al_phase()
    p2 = p.deepcopy()
    p2.name = "al2"
    rot = Rotation.from_axes_angles([1, 0, 0], (0, 15, 30, 45), degrees=True)
    coords = DiffractingVector(p, xyz=[[1, 0, 0], [0, -0.3, 0], [1 / 0.405, 1 / -0.405, 0],
                                       [0.1, -0.1, -0.3]])
    coords.intensity = 1
    return Simulation2D(phases=[p, p2], simulation_generator=SimulationGenerator(200),
                        coordinates=[[coords] * 4, [coords] * 4], rotations=[rot, rot])


def test_multi_phase_multi_rotation():
    sim = _multi()
    assert isinstance(sim.phases, np.ndarray) and isinstance(sim.coordinates, np.ndarray)
    for i in range(8):
        rot, ph, coords = sim.get_simulation(i)
        assert ph == (0 if i < 4 else 1)
    assert sim.iphase[0].rotations.size == 4 and sim.iphase["al"].phases.name == "al"
    with pytest.raises(ValueError):
        sim.iphase[3.1]
    assert sim.irot[0].rotations.size == 2 and sim.irot[0:2].rotations.size == 2
    assert sum(1 for _ in sim) == 8
    r, t, i = sim.polar_flatten_simulations()
    assert r.shape == (8, 4)
    assert isinstance(sim.rotate_shift_coordinates(angle=0.1), DiffractingVector)


def test_single_phase_multi_rotation():
    p = al_phase()
    rot = Rotation.from_axes_angles([1, 0, 0], (0, 15, 30, 45), degrees=True)
    coords = DiffractingVector(p, xyz=[[1, 0, 0], [0, 1, 0], [1, 1, 0], [1, 1, 1]],
                               intensity=[1, 2, 3, 4])
    sim = Simulation2D(phases=p, simulation_generator=SimulationGenerator(200),
                       coordinates=[coords] * 4, rotations=rot)
    np.testing.assert_array_equal(sim.get_current_rotation_matrix(), rot[0].to_matrix()[0])
    with pytest.raises(ValueError):
        sim.iphase[0]
    s = sim.irot[0]
    assert s.rotations.size == 1 and s.coordinates.size == 4
    assert sim.irot[0:2].coordinates.size == 2
    assert sum(1 for _ in sim) == 4


def test_transformed_coordinates_keep_integer_dtype():
    p = al_phase()
    sim = Simulation2D(phases=p, simulation_generator=None, rotations=Rotation.identity(),
                       coordinates=DiffractingVector(p, xyz=np.array([[3, 4, 0]]), intensity=[1]))
    t = sim._get_transformed_coordinates(0, (5, 5), units="pixel", calibration=2)
    assert np.issubdtype(t.data.dtype, np.integer)


def test_get_closest():
    np.testing.assert_array_equal(get_closest([0, 1, 2, 3], np.array([0.4, 0.6, 2.9, 5])), [0, 1, 3, 3])


# -- SimulationGenerator arguments (test_simulation_generator.py:95-130, :240-260) --

def test_generator_init_and_errors():
    g = SimulationGenerator(300)
    assert g.scattering_params == "lobato" and g.precession_angle == 0
    assert g.shape_factor_model is sfm.lorentzian and g.approximate_precession
    assert g.minimum_intensity == 1e-20
    assert repr(g) == ("SimulationGenerator(accelerating_voltage=300, scattering_params=lobato, "
                       "approximate_precession=True)")
    assert SimulationGenerator(300, shape_factor_model="linear").shape_factor_model is sfm.linear
    assert SimulationGenerator(300, shape_factor_model=sfm.binary).shape_factor_model is sfm.binary
    for sp in ("lobato", "xtables", None):
        SimulationGenerator(300, scattering_params=sp)
    with pytest.raises(NotImplementedError):
        SimulationGenerator(300, scattering_params="_empty")
    with pytest.raises(NotImplementedError):
        SimulationGenerator(300, shape_factor_model="dracula")
    assert abs(SimulationGenerator(200).wavelength - 0.0250793403) < 1e-9
    assert SimulationGenerator(precession_angle=-0.5).precession_angle == 0.5


def test_host_shape_models_match_oracle():
    s = np.linspace(-0.05, 0.05, 41)
    for name in ("linear", "lorentzian", "sinc", "sin2c", "atanc"):
        np.testing.assert_allclose(getattr(sfm, name)(s, 0.02), O.SHAPE_MODELS[name](s, 0.02),
                                   rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(sfm.lorentzian_precession(s, 0.02, 1.3, 0.01),
                               O.sf_lorentzian_precession(s, 0.02, 1.3, 0.01), rtol=1e-12)


def test_extreme_rows_hold_the_largest_g():
    """The refinement's largest |g| over the convex hull's rows equals the
    maximum over the whole hkl set for arbitrary bases (ball-shaped sets as
    the box enumeration makes them, a planar set, a line and tiny sets, which
    keep every row)."""
    from diffsims_amd.differentiable import _extreme_rows

    rng = np.random.default_rng(5)
    _, rec = O.lattice_basis(7.3, 9.1, 11.7, 83.0, 101.0, 95.0)
    ball = O.hkl_box(rec, 1.0 / 1.6).astype(np.float64)
    plane = np.array([[h, k, 0] for h in range(-6, 7) for k in range(-5, 6)], dtype=np.float64)
    line = np.array([[0, 0, l] for l in range(-9, 10)], dtype=np.float64)
    tiny = np.array([[1, 0, 0], [0, 1, 0], [1, 1, 1]], dtype=np.float64)
    for rows in (ball, plane, line, tiny):
        ext = _extreme_rows(rows)
        assert ext.shape[0] <= rows.shape[0]
        for _ in range(20):
            r = rec @ (np.eye(3) + 0.3 * rng.normal(size=(3, 3)))
            full = np.max(np.sum((rows @ r.T) ** 2, axis=1))
            assert np.max(np.sum((ext @ r.T) ** 2, axis=1)) == full
    assert _extreme_rows(ball).shape[0] < ball.shape[0] // 5


def test_fast_renderer_choice():
    """Small batches of patterns up to 256x256 take the splat kernel alone,
    everything else the band renderer (the same output either way)."""
    from diffsims_amd.engine import fast_renderer

    assert fast_renderer(500, 144, 144) == "dsx_render_fast"
    assert fast_renderer(767, 256, 256) == "dsx_render_fast"
    assert fast_renderer(768, 256, 256) == "dsx_render_fast_bands"
    assert fast_renderer(200000, 256, 256) == "dsx_render_fast_bands"
    assert fast_renderer(300, 512, 512) == "dsx_render_fast_bands"
    assert fast_renderer(10, 256, 257) == "dsx_render_fast_bands"


# -- bench contract ---------------------------------------------------------------

def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before any
    device work."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
