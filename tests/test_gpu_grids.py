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


# -- SO(3) cubochoric grids (rotation_list_generators.py:85-134) ---------------

def _walk_match(got, want, ambiguous):
    """Device rows against oracle rows, both in grid order: every oracle row
    is matched in order except rows within the fundamental-zone tolerance
    band, which either side may keep.  Angles compared modulo 2 pi."""
    i = 0
    for k in range(want.shape[0]):
        if i < got.shape[0]:
            d = (got[i] - want[k] + np.pi) % (2 * np.pi) - np.pi
            if np.abs(d).max() <= 1e-9:
                i += 1
                continue
        assert ambiguous[k], (k, want[k], got[i] if i < got.shape[0] else None)
    assert i == got.shape[0]


@pytest.mark.parametrize("space_group", [1, 4, 20, 75, 89, 143, 150, 168, 177, 195, 207, 225])
def test_fundamental_zone_grid_matches_oracle(space_group):
    from diffsims_amd import generators, symmetry

    for res in (15.0, 7.0):
        n = O.cubochoric_steps(res)
        q = O.cubochoric_quaternions(n)
        ops = O.proper_group_quaternions(O.proper_group_of_space_group(space_group))
        margin = O.fundamental_zone_margin(q, ops)
        keep = margin >= -O.FZ_EPS
        want = O.bunge_from_quat(q[keep])
        amb = np.abs(margin[keep] + O.FZ_EPS) < 1e-13
        got = generators.fundamental_zone_grid_device(res, space_group).cpu().numpy()
        print(f"sg {space_group} res {res}: {got.shape[0]} rows, {int(amb.sum())} on the boundary",
              flush=True)
        assert symmetry.rotation_group_of_space_group(space_group) == \
            O.proper_group_of_space_group(space_group)
        _walk_match(got, want, amb)


def test_local_grid_matches_oracle():
    from diffsims_amd import generators

    for res, center, width in ((5.0, None, 20.0), (3.0, (12.0, 34.0, 56.0), 15.0),
                               (30.0, (0, 1, 0), 35.0)):
        want = O.local_grid(res, center, width)
        got = generators.local_grid_device(res, center, width).cpu().numpy()
        assert got.shape == want.shape, (res, center, width)
        d = (got - want + np.pi) % (2 * np.pi) - np.pi
        assert np.abs(d).max() <= 1e-9


def test_reference_grid_calls_on_the_device():
    """test_rotation_list_generator.py:30-44 through the drop-in functions,
    against the oracle's rounded lists (entries may differ by one unit of
    the last decimal where the FP64 angle sits on a rounding tie)."""
    from diffsims_amd import generators

    for got, want in ((generators.get_local_grid(resolution=30, center=(0, 1, 0), grid_width=35),
                       O.rotation_list(O.local_grid(30, (0, 1, 0), 35))),
                      (generators.get_fundamental_zone_grid(space_group=20, resolution=20),
                       O.rotation_list(O.fundamental_zone_grid(20, 20)))):
        assert isinstance(got, list) and len(got) > 0 and isinstance(got[0], tuple)
        assert len(got) == len(want)
        d = (np.array(got) - np.array(want) + 180.0) % 360.0 - 180.0
        assert np.abs(d).max() <= 0.01 + 1e-9
    with pytest.raises(ValueError):
        generators.get_fundamental_zone_grid(resolution=20)


def test_fine_fundamental_zone_on_the_device():
    """0.5 degree cubic zone: 610^3 grid points, 1/24 of them kept (up to the
    boundary rows), every kept rotation within Mackenzie's 62.8 degrees."""
    import time

    from diffsims_amd import generators

    n = generators.cubochoric_steps(0.5)
    torch.cuda.synchronize()
    t = time.perf_counter()
    e = generators.fundamental_zone_grid_device(0.5, 225)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    total = (2 * n) ** 3
    print(f"cubic zone at 0.5 deg: {e.shape[0]} of {total} grid points in {dt * 1e3:.1f} ms",
          flush=True)
    assert abs(e.shape[0] * 24 / total - 1) < 0.01
    # the angle of each kept rotation from its passive Bunge matrix
    sub = e[:: max(1, e.shape[0] // 200000)]
    M = torch.from_numpy(O.bunge_matrix(sub.cpu().numpy()))
    cosw = (M.diagonal(dim1=-2, dim2=-1).sum(-1) - 1) / 2
    ang = torch.rad2deg(torch.arccos(cosw.clamp(-1, 1)))
    assert float(ang.max()) <= 62.7994 + 1e-6
