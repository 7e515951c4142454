"""The legacy generator API on the GPU against the oracle's restatement and
the reference's own known answers
(diffsims/tests/generators/test_diffraction_generator.py,
test_library_generator.py).

Parity as for the current path (tests/parity.py): the reflection set of
get_points_in_sphere equals the oracle's bit for bit and in order; survivor
sets equal except rows within 1e-6 of the cut; intensities within 1e-5 of the
pattern maximum; coordinates within 1e-9 1/A (FP64 on both sides).
"""

import numpy as np
import pytest

from oracle import diffsims_oracle as O
from tests import parity as P

pytestmark = pytest.mark.gpu


def _si_structure(a=5.431):
    from diffsims_amd import Atom, Lattice, Structure

    latt = Lattice(a, a, a, 90, 90, 90)
    atoms = []
    for x, y, z in [[0, 0, 0], [0.5, 0, 0.5], [0, 0.5, 0.5], [0.5, 0.5, 0]]:
        atoms.append(Atom(atype="Si", xyz=[x, y, z], lattice=latt))
        atoms.append(Atom(atype="Si", xyz=[x + 0.25, y + 0.25, z + 0.25], lattice=latt))
    return Structure(atoms=atoms, lattice=latt)


def _triclinic_structure(seed=3):
    from diffsims_amd import Atom, Lattice, Structure

    rng = np.random.default_rng(seed)
    latt = Lattice(6.2, 7.1, 8.3, 84.0, 97.0, 103.0)
    syms = ["O", "Si", "Fe", "Mg", "Cu"]
    atoms = [Atom(atype=syms[i % 5], xyz=rng.uniform(0, 1, 3), lattice=latt) for i in range(9)]
    return Structure(atoms=atoms, lattice=latt)


def _spec(structure, dwf=None):
    from diffsims_amd.legacy import structure_arrays

    return P.phase_spec(structure_arrays(structure, dwf))


MODELS = ["lorentzian", "linear", "atanc", "sinc", "sin2c"]


@pytest.mark.parametrize("model", MODELS)
def test_ed_data_matches_oracle(model):
    from diffsims_amd import engine
    from diffsims_amd.legacy import DiffractionGenerator, euler2mat_rzxz

    gen = DiffractionGenerator(300, shape_factor_model=model)
    stats = {}
    for structure, rr, angles in ((_si_structure(), 3.0, [(0, 0, 0), (12, 34, 56), (90, 45, 10)]),
                                  (_triclinic_structure(), 2.0, [(0, 0, 0), (200, 71, -33)])):
        spec = _spec(structure)
        p = O.SimParams(kv=300, reciprocal_radius=rr, max_excitation_error=1e-2,
                        shape_model=model, with_direct_beam=True)
        otab = O.legacy_tables(spec, p)
        tab = gen._table(structure, rr, {})
        assert np.array_equal(tab.hkl_host(), otab.hkl)
        M = euler2mat_rzxz(angles)
        Rt = engine.rotation_matrices_to_device(np.transpose(M, (0, 2, 1)))
        b = engine.simulate(tab, Rt, gen._settings(1e-2, None), rr)
        k = 1 / O.electron_wavelength(300)
        for o, e in enumerate(angles):
            c = int(b.count[o])
            ref, _ = P.check_orientation(otab, M[o].T, k, p, b.idx[o, :c].cpu().numpy(),
                                         b.spots[o, :c, 3].cpu().numpy(), stats)
            sim = gen.calculate_ed_data(structure, rr, rotation=e, with_direct_beam=True)
            assert np.array_equal(sim.indices, otab.hkl[b.idx[o, :c].cpu().numpy()])
            assert np.abs(sim.coordinates - ref.xyz).max() <= 1e-9
    assert stats["orientations"] == 5


def test_ed_data_reference_kats():
    """test_diffraction_generator.py:136-213 through the device."""
    from diffsims_amd.legacy import DiffractionGenerator

    gen = DiffractionGenerator(300)
    si = _si_structure()
    d = gen.calculate_ed_data(si, reciprocal_radius=5.0)
    assert len(d.indices) == len(d.coordinates) == len(d.intensities)
    sims = [gen.calculate_ed_data(_si_structure(a), reciprocal_radius=5.0) for a in (5, 10)]
    idx = [[tuple(i) for i in s.indices] for s in sims]
    assert (2, 2, 0) in idx[0] and (2, 2, 0) in idx[1]
    assert np.allclose(sims[0].coordinates[idx[0].index((2, 2, 0))],
                       sims[1].coordinates[idx[1].index((2, 2, 0))] * 2)
    central = idx[0].index((0, 0, 0))
    assert np.all(sims[0].intensities[central] >= sims[0].intensities)
    t1 = gen.calculate_ed_data(si, 2, max_excitation_error=0.02)
    t2 = gen.calculate_ed_data(si, 2, max_excitation_error=0.4)
    assert np.sum(t1.intensities) != np.sum(t2.intensities)
    nb = gen.calculate_ed_data(si, 5.0, with_direct_beam=False)
    assert (0, 0, 0) not in [tuple(i) for i in nb.indices] and len(nb) == len(d) - 1
    for g in (DiffractionGenerator(300, precession_angle=0.5, approximate_precession=True),
              DiffractionGenerator(300, precession_angle=0.5, approximate_precession=False)):
        e = g.calculate_ed_data(si, reciprocal_radius=5.0)
        assert len(e.indices) == len(e.coordinates) == len(e.intensities) > 0

    def local_excite(excitation_error, maximum_excitation_error, t):
        return (np.sin(t) * excitation_error) / maximum_excitation_error

    c = DiffractionGenerator(300, shape_factor_model=local_excite, t=0.2)
    e = c.calculate_ed_data(si, reciprocal_radius=5.0)
    assert len(e.indices) == len(e.coordinates) == len(e.intensities)
    with pytest.raises(NotImplementedError):
        DiffractionGenerator(300, shape_factor_model="dracula")
    with pytest.raises(NotImplementedError):
        DiffractionGenerator(300, scattering_params="_empty")


def test_precession_matches_oracle():
    from diffsims_amd import engine
    from diffsims_amd.legacy import DiffractionGenerator, euler2mat_rzxz

    gen = DiffractionGenerator(200, precession_angle=1.0)
    structure = _triclinic_structure(5)
    rr = 1.5
    p = O.SimParams(kv=200, reciprocal_radius=rr, max_excitation_error=1e-2,
                    precession_angle=1.0, with_direct_beam=True)
    otab = O.legacy_tables(_spec(structure), p)
    tab = gen._table(structure, rr, {})
    angles = [(0, 0, 0), (33, 66, 99)]
    M = euler2mat_rzxz(angles)
    Rt = engine.rotation_matrices_to_device(np.transpose(M, (0, 2, 1)))
    b = engine.simulate(tab, Rt, gen._settings(1e-2, None), rr)
    k = 1 / O.electron_wavelength(200)
    for o in range(2):
        c = int(b.count[o])
        P.check_orientation(otab, M[o].T, k, p, b.idx[o, :c].cpu().numpy(),
                            b.spots[o, :c, 3].cpu().numpy())


def test_profile_matches_oracle():
    from diffsims_amd import Atom, Lattice, Structure
    from diffsims_amd.legacy import DiffractionGenerator, ProfileSimulation

    gen = DiffractionGenerator(300)
    si = _si_structure()
    prof = gen.calculate_profile_data(si, reciprocal_radius=1.0)
    assert isinstance(prof, ProfileSimulation)
    p = O.SimParams(kv=300, reciprocal_radius=1.0, with_direct_beam=True)
    x, y, labels = O.legacy_profile(_spec(si), p)
    assert prof.hkls == labels
    assert np.allclose(prof.magnitudes, x, rtol=1e-12, atol=0)
    assert np.abs(np.asarray(prof.intensities) - y).max() <= 1e-5 * 100
    latt = Lattice(3, 3, 5, 90, 90, 120)
    hexa = Structure(atoms=[Atom(atype="Ni", xyz=[0, 0, 0], lattice=latt)], lattice=latt)
    hp = gen.calculate_profile_data(structure=hexa, reciprocal_radius=1.0)
    x, y, labels = O.legacy_profile(_spec(hexa), p, hexagonal=True)
    assert isinstance(hp, ProfileSimulation) and hp.hkls == labels
    assert np.abs(np.asarray(hp.intensities) - y).max() <= 1e-5 * 100


def test_library_equals_single_simulations():
    """get_diffraction_library (library_generator.py:51-153): one launch per
    phase gives every orientation's single calculate_ed_data; pixel
    coordinates rint(calibrated + half_shape); library entry lookup."""
    from diffsims_amd.legacy import (DiffractionGenerator, DiffractionLibrary,
                                     DiffractionLibraryGenerator, StructureLibrary)

    gen = DiffractionGenerator(300)
    structures = [_si_structure(), _triclinic_structure(7)]
    orientations = [[(0, 0, 0), (0, 0, 45), (10, 20, 30)], [(5, 15, 25), (90, 90, 90)]]
    sl = StructureLibrary(["Si", "tri"], structures, orientations)
    lib = DiffractionLibraryGenerator(gen).get_diffraction_library(
        sl, calibration=0.015, reciprocal_radius=1.5, half_shape=(72, 72), with_direct_beam=False)
    assert isinstance(lib, DiffractionLibrary) and lib.identifiers == ["Si", "tri"]
    for name, st, ori in zip(["Si", "tri"], structures, orientations):
        for i, e in enumerate(ori):
            sim = gen.calculate_ed_data(st, 1.5, rotation=e, with_direct_beam=False)
            got = lib[name]["simulations"][i]
            assert np.array_equal(got.indices, sim.indices)
            assert np.array_equal(got.coordinates, sim.coordinates)
            assert np.array_equal(lib[name]["intensities"][i], sim.intensities)
            sim.calibration = 0.015
            px = np.rint(sim.calibrated_coordinates[:, :2] + (72, 72)).astype(int)
            assert np.array_equal(lib[name]["pixel_coords"][i], px)
    entry = lib.get_library_entry("Si", angle=(0, 0, 45))
    assert entry["Sim"] is lib["Si"]["simulations"][1]


def test_legacy_pattern_matches_oracle_render():
    """DiffractionSimulation.get_diffraction_pattern (the GPU fast renderer on
    calibrated coordinates) against scipy's filter on the FP64 pixels."""
    from diffsims_amd.legacy import DiffractionGenerator

    gen = DiffractionGenerator(300)
    for sigma, shape in ((1.0, (144, 144)), (10, (256, 256))):
        pat = gen.calculate_ed_data(_si_structure(), 1.5, rotation=(10, 20, 30))
        pat.calibration = 0.01
        got = pat.get_diffraction_pattern(shape=shape, sigma=sigma)
        H, W = shape
        xy = pat._get_transformed_coordinates(0, (W // 2, H // 2), False, units="pixel")
        # spots within 1e-4 px of a pixel edge: the pixel the GPU's FP32 took
        g32 = pat.calibrated_coordinates.astype(np.float32)
        for c, cen in ((0, W // 2), (1, H // 2)):
            fr = xy[:, c] - np.floor(xy[:, c])
            edge = (fr < P.PIX_EPS) | (fr > 1 - P.PIX_EPS)
            xy[edge, c] = (g32[edge, c] + np.float32(cen)).astype(np.float64)
        want = P.render_from_pixels(xy[:, 0], xy[:, 1], pat.intensities, shape, sigma)
        assert P.pattern_error(got, want) <= P.TOL_P, sigma


@pytest.mark.parametrize("is_hex", [True, False])
def test_calculate_diffraction1d(is_hex):
    """test_simulation_generator.py:225-249: a Si phase (space group 227),
    optionally with its lattice made hexagonal; equal-length arrays and hkil
    indices for the hexagonal lattice.  Intensities = |F|^2 x orbit size of
    each family (FP64 oracle F), scaled to 100."""
    from diffsims_amd import Phase, Simulation1D, SimulationGenerator
    from diffsims_amd.symmetry import symmetry_families

    st = _si_structure()
    phase = Phase("Si", space_group=227, structure=st)
    if is_hex:
        lat = phase.structure.lattice
        lat.a = lat.b
        lat.alpha, lat.beta, lat.gamma = 90.0, 90.0, 120.0
    gen = SimulationGenerator(300)
    sim = gen.calculate_diffraction1d(phase, 0.5)
    assert isinstance(sim, Simulation1D)
    assert len(sim.intensities) == len(sim.reciprocal_spacing) == len(sim.hkl)
    assert all(len(h) == (4 if is_hex else 3) for h in sim.hkl)
    assert abs(max(sim.intensities) - 100) < 1e-9 and sim.reciprocal_radius == 0.5
    # the same families and intensities from the oracle's FP64 F
    lat = phase.structure.lattice
    base, rec = O.lattice_basis(lat.a, lat.b, lat.c, lat.alpha, lat.beta, lat.gamma)
    hkl = O.hkl_box(rec, 0.7)
    g = hkl @ rec.T
    rep, mult = symmetry_families(g, "m-3m")
    spec = _spec(st)
    coeffs, _ = O.atom_coeffs(spec.symbols, "lobato")
    gn = np.linalg.norm(g[rep], axis=1)
    F = O.structure_factor(hkl[rep], gn, np.asarray(spec.xyz), np.asarray(spec.occ),
                           np.asarray(spec.B), coeffs, "lobato", phase_mode="exact")
    I = np.abs(F) ** 2 * mult
    I = I / I.max() * 100
    keep = I > 1e-3
    assert np.allclose(sim.reciprocal_spacing, gn[keep], rtol=1e-12, atol=0)
    assert np.abs(np.asarray(sim.intensities) - I[keep]).max() <= 1e-5 * 100
    h3 = np.asarray(sim.hkl)[:, [0, 1, 3]] if is_hex else np.asarray(sim.hkl)
    assert np.array_equal(h3, hkl[rep][keep])


def test_calculate_diffraction1d_al_kat():
    """Al (Fm-3m): the {111} and {200} lines carry multiplicities 8 and 6
    (reciprocal_lattice_vector.py:303-310), so I(111)/I(200) = 8 |F111|^2 /
    (6 |F200|^2) with the oracle's F."""
    from diffsims_amd import Atom, Lattice, Phase, SimulationGenerator, Structure

    latt = Lattice(4.04, 4.04, 4.04, 90, 90, 90)
    st = Structure([Atom("Al", p, lattice=latt) for p in
                    ([0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5])], lattice=latt)
    sim = SimulationGenerator(200).calculate_diffraction1d(Phase("al", point_group="m-3m",
                                                                  structure=st))
    fam = {tuple(sorted(np.abs(h).astype(int))): i for h, i in zip(sim.hkl, sim.intensities)}
    spec = _spec(st)
    coeffs, _ = O.atom_coeffs(spec.symbols, "lobato")
    _, rec = O.lattice_basis(4.04, 4.04, 4.04, 90, 90, 90)
    hk = np.array([[1, 1, 1], [2, 0, 0]])
    F = O.structure_factor(hk, np.linalg.norm(hk @ rec.T, axis=1), np.asarray(spec.xyz),
                           np.asarray(spec.occ), np.asarray(spec.B), coeffs, "lobato",
                           phase_mode="exact")
    want = 8 * abs(F[0]) ** 2 / (6 * abs(F[1]) ** 2)
    assert abs(fam[(1, 1, 1)] / fam[(0, 0, 2)] - want) <= 1e-5 * want
