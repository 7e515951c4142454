"""The drop-in SimulationGenerator / Simulation2D surface on the GPU, mirroring
the reference's own tests (diffsims/tests/generators/test_simulation_generator.py,
diffsims/tests/simulations/test_simulations2d.py) and checked against the oracle."""

import os
import re

import numpy as np
import pytest

from oracle import diffsims_oracle as O
from tests import parity as P

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "old_simulation.npy")


def make_phase(a=5.431):
    from diffsims_amd import Atom, Lattice, Phase, Structure
    latt = Lattice(a, a, a, 90, 90, 90)
    atoms = [Atom(atype="Si", xyz=p, lattice=latt) for p in P.si_positions()]
    return Phase(structure=Structure(atoms=atoms, lattice=latt), space_group=227)


@pytest.fixture
def gen300():
    from diffsims_amd import SimulationGenerator
    return SimulationGenerator(300)


def test_matching_results(gen300):
    assert gen300.calculate_diffraction2d(make_phase(), reciprocal_radius=5.0).coordinates.size == 70


def test_precession_simple():
    from diffsims_amd import SimulationGenerator
    g = SimulationGenerator(300, precession_angle=0.5, approximate_precession=True)
    assert g.calculate_diffraction2d(make_phase(), reciprocal_radius=5.0).coordinates.size == 250


def test_precession_full_integral():
    from diffsims_amd import SimulationGenerator
    g = SimulationGenerator(300, precession_angle=0.5, approximate_precession=False)
    assert g.calculate_diffraction2d(make_phase(), reciprocal_radius=5.0).coordinates.size == 250


def test_custom_shape_function_matches_oracle():
    # test_simulation_generator.py:163-168 expects 52 for this callable; the
    # restatement of the code gives 36 (parity unpinned); the drop-in must agree
    # with the oracle
    from diffsims_amd import SimulationGenerator

    def local_excite(s, w, t):
        return (np.sin(t) * s) / w
    g = SimulationGenerator(300, shape_factor_model=local_excite, t=0.2)
    sim = g.calculate_diffraction2d(make_phase(), reciprocal_radius=5.0)
    ph = O.PhaseSpec((5.431,) * 3, (90,) * 3, ["Si"] * 8, P.si_positions(), np.ones(8), np.zeros(8))
    base, rec = ph.basis()
    p = O.SimParams(kv=300, reciprocal_radius=5.0, shape_model=local_excite,
                    shape_kwargs={"t": 0.2})
    ref = O.simulate_one_reference(ph, O.hkl_box(rec, 0.2, True), rec, np.eye(3),
                                   1 / O.electron_wavelength(300), p)
    assert sim.coordinates.size == ref.idx.size == 36
    np.testing.assert_allclose(np.sort(sim.coordinates.intensity), np.sort(ref.intensity),
                               rtol=1e-6)


def test_appropriate_scaling(gen300):
    d = gen300.calculate_diffraction2d(make_phase(5), reciprocal_radius=5.0)
    big = gen300.calculate_diffraction2d(make_phase(10), reciprocal_radius=5.0)
    i1 = [tuple(i) for i in np.rint(d.coordinates.hkl).astype(int)]
    i2 = [tuple(i) for i in np.rint(big.coordinates.hkl).astype(int)]
    assert (2, 2, 0) in i1 and (2, 2, 0) in i2
    c1 = d.coordinates[i1.index((2, 2, 0))]
    c2 = big.coordinates[i2.index((2, 2, 0))]
    assert np.allclose(c1.data, c2.data * 2)


def test_direct_beam_strongest_and_absent(gen300):
    d = gen300.calculate_diffraction2d(make_phase(), reciprocal_radius=0.5, with_direct_beam=True)
    idx = [tuple(i) for i in np.rint(d.coordinates.hkl).astype(int)]
    c = idx.index((0, 0, 0))
    assert np.all(d.coordinates.intensity[c] >= d.coordinates.intensity)
    d = gen300.calculate_diffraction2d(make_phase(), reciprocal_radius=0.5, with_direct_beam=False)
    assert (0, 0, 0) not in [tuple(i) for i in np.rint(d.coordinates.hkl).astype(int)]


def test_shape_factor_width_changes_result(gen300):
    t1 = gen300.calculate_diffraction2d(make_phase(), max_excitation_error=0.02)
    t2 = gen300.calculate_diffraction2d(make_phase(), max_excitation_error=0.4)
    assert np.sum(t1.coordinates.intensity) != np.sum(t2.coordinates.intensity)


def test_multiphase_multirotation_equals_single_runs(gen300):
    from diffsims_amd import Rotation
    rot = Rotation.from_euler([[0, 0, 0], [0.1, 0.1, 0.1]])
    rot2 = Rotation.from_euler([[0, 0, 0], [0.1, 0.1, 0.1], [0.2, 0.2, 0.2]])
    sim = gen300.calculate_diffraction2d([make_phase(5), make_phase(10)], rotation=[rot, rot2])
    assert sim.has_multiple_phases and sim._num_rotations() == [2, 3]
    single = gen300.calculate_diffraction2d(make_phase(10), rotation=rot2)
    for a, b in zip(sim.coordinates[1], single.coordinates):
        assert np.array_equal(a.hkl, b.hkl) and np.array_equal(a.intensity, b.intensity)
    with pytest.raises(ValueError):
        gen300.calculate_diffraction2d([make_phase(5), make_phase(10)], rotation=[rot])


def test_progressbar(capsys):
    from diffsims_amd import Rotation, SimulationGenerator
    gen = SimulationGenerator()
    p1, p2 = make_phase(), make_phase()
    p1.name, p2.name = "A", "B"
    rots = Rotation.random(10, seed=1)
    gen.calculate_diffraction2d([p1, p2], [rots, rots], show_progressbar=False)
    assert capsys.readouterr().err == ""
    gen.calculate_diffraction2d([p1, p2], [rots, rots], show_progressbar=True)
    err = capsys.readouterr().err
    assert re.findall(r"A: 100%\|█+\| 10/10", err) and re.findall(r"B: 100%\|█+\| 10/10", err)


def test_golden_pattern_through_the_drop_in():
    # test_simulation_generator.py:283-343, in FP32 on the GPU
    from diffsims_amd import Atom, Lattice, Phase, Rotation, SimulationGenerator, Structure
    latt = Lattice(2.464, 2.464, 6.711, 90, 90, 120)
    atoms = [Atom("C", [0, 0, .25], lattice=latt), Atom("C", [0, 0, .75], lattice=latt),
             Atom("C", [1 / 3, 2 / 3, .25], lattice=latt), Atom("C", [2 / 3, 1 / 3, .75], lattice=latt)]
    p = Phase("Graphite", point_group="6/mmm", structure=Structure(atoms=atoms, lattice=latt))
    gen = SimulationGenerator(accelerating_voltage=200, scattering_params="lobato",
                              precession_angle=0, shape_factor_model="lorentzian",
                              approximate_precession=True, minimum_intensity=1e-20)
    sim = gen.calculate_diffraction2d(phase=p, rotation=Rotation.from_euler([[0, 90, 90]], degrees=True),
                                      reciprocal_radius=1.6768, max_excitation_error=0.1,
                                      with_direct_beam=False)
    new = sim.get_diffraction_pattern(shape=(128, 128), sigma=1.4, calibration=0.0262)
    assert np.abs(new - np.load(GOLDEN)).max() < 1e-5


def test_fast_vs_slow_integer_and_float_coordinates():
    # test_simulations2d.py:133-180
    from diffsims_amd import (Atom, DiffractingVector, Lattice, Phase, Rotation, Simulation2D,
                              SimulationGenerator, Structure)
    al = Phase(name="al", space_group=225, structure=Structure(
        atoms=[Atom("al", [0, 0, 0])], lattice=Lattice(0.405, 0.405, 0.405, 90, 90, 90)))
    gen = SimulationGenerator()
    rot = Rotation.identity()
    xyz = np.asarray([[0, 0, 0]])
    i = np.array([30])
    int_sim = Simulation2D(phases=al, simulation_generator=gen, rotations=rot,
                           coordinates=DiffractingVector(al, xyz=xyz.astype(int), intensity=i))
    float_sim = Simulation2D(phases=al, simulation_generator=gen, rotations=rot,
                             coordinates=DiffractingVector(al, xyz=xyz.astype(float), intensity=i))
    kw = dict(shape=(11, 11), sigma=1, calibration=1, normalize=False, clip_threshold=0.01)
    fast = int_sim.get_diffraction_pattern(fast=True, **kw)
    slow = int_sim.get_diffraction_pattern(fast=False, **kw)
    assert np.array_equal(fast, slow)
    ffast = float_sim.get_diffraction_pattern(fast=True, **kw)
    fslow = float_sim.get_diffraction_pattern(fast=False, **kw)
    assert np.allclose(ffast, fast)
    assert np.all((fslow - ffast) < kw["clip_threshold"])
    kw["shape"] = (10, 10)
    assert not np.allclose(float_sim.get_diffraction_pattern(fast=True, **kw),
                           float_sim.get_diffraction_pattern(fast=False, **kw))
    # against the restated reference renderers
    from scipy import ndimage as ndi
    d = np.zeros((11, 11))
    d[5, 5] = 30
    assert np.abs(fast - ndi.gaussian_filter(d, 1)).max() < 1e-5 * 30


def test_no_spots_and_normalisation():
    # test_simulations2d.py:457-470
    from diffsims_amd import (Atom, DiffractingVector, Lattice, Phase, Rotation, Simulation2D,
                              SimulationGenerator, Structure)
    al = Phase(name="al", space_group=225, structure=Structure(
        atoms=[Atom("al", [0, 0, 0])], lattice=Lattice(0.405, 0.405, 0.405, 90, 90, 90)))
    rot = Rotation.from_axes_angles([1, 0, 0], (0, 15, 30, 45), degrees=True)
    coords = DiffractingVector(al, xyz=[[1, 0, 0], [0, -0.3, 0], [1 / 0.405, 1 / -0.405, 0],
                                        [0.1, -0.1, -0.3]])
    coords.intensity = 1
    sim = Simulation2D(phases=[al, al], simulation_generator=SimulationGenerator(200),
                       coordinates=[[coords] * 4, [coords] * 4], rotations=[rot, rot])
    pat = sim.get_diffraction_pattern(shape=(50, 50), calibration=0.001)
    assert pat.shape == (50, 50) and np.max(pat) == 0
    pat = sim.get_diffraction_pattern(shape=(512, 512), calibration=0.01)
    assert pat.shape == (512, 512) and np.max(pat) == 1
    many = sim.get_diffraction_patterns(shape=(64, 64), sigma=2, calibration=0.02, as_numpy=True)
    assert many.shape == (8, 64, 64)


def test_get_intersecting_reflections(gen300):
    from diffsims_amd import DiffractingVector, Rotation
    phase = make_phase()
    hkl = O.hkl_box(O.lattice_basis(5.431, 5.431, 5.431, 90, 90, 90)[1], 0.5, True)
    recip = DiffractingVector(phase, hkl=hkl)
    vecs, h, sf = gen300.get_intersecting_reflections(recip, Rotation.identity(),
                                                      gen300.wavelength, 0.01)
    k = 1 / gen300.wavelength
    g = vecs.data
    r = np.sqrt(np.sum(g[:, :2] ** 2, 1))
    s = (-np.sqrt(k ** 2 - r ** 2) + k) - g[:, 2]
    assert np.all(np.abs(s) < 0.01)
    np.testing.assert_allclose(sf, O.sf_lorentzian(s, 0.01), rtol=1e-6)
    assert (np.abs(h).sum(axis=1) == 0).sum() == 2
