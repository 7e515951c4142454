"""Beam-direction grids on the device against the FP64 oracle: the same rows
in the same order (the crop decisions are exact: |n . v| is either 0 or far
from the -1e-13 tolerance), angles within 1e-9 degrees (the device's FP64
tan/acos and numpy's may differ in the last bit)."""
import numpy as np
import pytest
import torch

from oracle import diffsims_oracle as O

pytestmark = pytest.mark.gpu

SYSTEMS = ["cubic", "hexagonal", "trigonal", "tetragonal", "orthorhombic", "monoclinic",
           "triclinic"]
MESHES = ["uv_sphere", "normalized_cube", "spherified_cube_edge", "spherified_cube_corner"]


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("crystal_system", SYSTEMS)
def test_grid_matches_oracle(crystal_system, mesh):
    from diffsims_amd import generators

    for res in (5.0, 2.0):
        want = O.beam_directions_grid(crystal_system, res, mesh=mesh)
        got = generators.get_beam_directions_grid(crystal_system, res, mesh=mesh)
        assert got.shape == want.shape, (crystal_system, mesh, res)
        assert np.abs(got - want).max() <= 1e-9, (crystal_system, mesh, res)


def test_grid_size_kats_on_the_device():
    """test_rotation_list_generator.py:80-94 through the device path."""
    from diffsims_amd import generators

    sizes = {"cubic": 300, "hexagonal": 1050, "trigonal": 1657, "tetragonal": 852,
             "orthorhombic": 1657, "monoclinic": 6441, "triclinic": 12698}
    for cs, n in sizes.items():
        assert generators.get_beam_directions_grid(cs, 2).shape == (n, 3)


def test_fine_grid_and_rotations_on_the_device():
    """A 0.25 degree triclinic grid stays on the device and feeds the rotation
    matrices directly; spot-checked rows against the oracle."""
    from diffsims_amd import differentiable as D
    from diffsims_amd import generators

    g = generators.beam_directions_grid_device("triclinic", 0.25)
    want = O.beam_directions_grid("triclinic", 0.25)
    assert g.shape == want.shape and g.is_cuda
    idx = np.random.default_rng(3).integers(0, want.shape[0], 4000)
    assert np.abs(g.cpu().numpy()[idx] - want[idx]).max() <= 1e-9
    Rt = D.euler_to_rt(torch.deg2rad(g[:1000]).contiguous())
    M = O.bunge_matrix(np.deg2rad(want[:1000]))
    assert np.abs(Rt.cpu().numpy().reshape(-1, 3, 3) - M.transpose(0, 2, 1)).max() <= 1e-12


def test_grid_errors():
    from diffsims_amd import generators

    with pytest.raises(NotImplementedError):
        generators.get_beam_directions_grid("cubic", 10, mesh="invalid")
    with pytest.raises(ValueError):
        generators.get_beam_directions_grid("cubic", 70, mesh="icosahedral")
    with pytest.raises(KeyError):
        generators.get_beam_directions_grid("quasicrystal", 10)


@pytest.mark.parametrize("crystal_system", SYSTEMS)
def test_icosahedral_grid_matches_oracle(crystal_system):
    from diffsims_amd import generators

    for res in (5.0, 2.0):
        want = O.beam_directions_grid(crystal_system, res, mesh="icosahedral")
        got = generators.get_beam_directions_grid(crystal_system, res, mesh="icosahedral")
        assert got.shape == want.shape, (crystal_system, res)
        assert np.abs(got - want).max() <= 1e-9, (crystal_system, res)


def test_icosahedral_levels_and_vertices():
    """The refinement level equals the reference loop's (oracle with a k-d
    tree) over resolutions from 20 to 0.9 degrees; the device vertices equal
    the oracle's rows; 642 vertices at 10 degrees
    (test_sphere_mesh_generators.py:115-121)."""
    from diffsims_amd import generators

    for res in (20.0, 10.0, 7.3, 5.0, 3.1, 2.0, 1.37, 0.9):
        n, _ = O.icosahedral_level(res)
        assert generators.icosahedral_level(res) == n, res
    for res in (10.0, 2.0, 0.9):
        got = generators.get_icosahedral_mesh_vertices(res)
        want = O.icosahedral_mesh_vertices(res)
        assert got.shape == want.shape
        assert np.abs(got - want).max() <= 1e-15, res
    g = generators.get_icosahedral_mesh_vertices(10)
    assert g.shape == (642, 3) and abs(g.sum()) < 1e-7


def test_icosahedral_fine_level_on_the_device():
    """A 0.2 degree mesh (level ~ 380, 1.4 M vertices): the level search runs
    every level on the device; the level's angle brackets the resolution
    (oracle k-d tree on the chosen level and the one before it)."""
    from diffsims_amd import generators

    n = generators.icosahedral_level(0.2)
    assert O.max_nn_angle(O.icosa_refined_vertices(n)) <= 0.2 < \
        O.max_nn_angle(O.icosa_refined_vertices(n - 1))
    g = generators.beam_directions_grid_device("cubic", 0.2, mesh="icosahedral")
    want = O.beam_directions_grid("cubic", 0.2, vertices=O.icosahedral_vertices_at_level(n))
    assert g.shape == want.shape
    assert np.abs(g.cpu().numpy() - want).max() <= 1e-9


@pytest.mark.parametrize("crystal_system", SYSTEMS)
def test_random_mesh_crop_matches_oracle(crystal_system):
    """Seeded random mesh (numpy's generator, as the reference): the device
    crop and angles equal the oracle's for the same points."""
    from diffsims_amd import generators

    v = generators.get_random_sphere_vertices(2.0, seed=11)
    assert np.array_equal(v, O.random_sphere_vertices(2.0, seed=11))
    planes = generators.crop_planes(crystal_system)
    got = generators.directions_to_euler_device(v, planes).cpu().numpy()
    want = O.beam_directions_grid(crystal_system, 2.0, mesh="random", vertices=v)
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-9
    g = generators.get_beam_directions_grid(crystal_system, 5.0, mesh="random")
    assert g.ndim == 2 and g.shape[1] == 3 and g.shape[0] > 0


def test_vectors_to_euler_kat_on_the_device():
    """test_sphere_mesh_generators.py:124-143 through the device."""
    from diffsims_amd import generators

    v = np.array([[1, 0, 0], [0, 1, 0], [0, 1, 1], [1, 0, 1]], dtype=np.float64)
    v = v / np.linalg.norm(v, axis=1)[:, None]
    want = np.array([[0, 90, 90], [0, 90, 0], [0, 45, 0], [0, 45, 90]], dtype=np.float64)
    assert np.allclose(generators.beam_directions_grid_to_euler(v), want)
