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
    with pytest.raises(NotImplementedError):
        generators.get_beam_directions_grid("cubic", 10, mesh="icosahedral")
    with pytest.raises(KeyError):
        generators.get_beam_directions_grid("quasicrystal", 10)
