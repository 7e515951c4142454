"""Beam-direction grids of template libraries on the device.

``get_beam_directions_grid`` mirrors
diffsims.generators.rotation_list_generators.get_beam_directions_grid
(diffsims/generators/rotation_list_generators.py:176-267): a sphere mesh
(diffsims/generators/sphere_mesh_generators.py), cropped to the crystal
system's stereographic triangle, as Bunge angles in degrees with phi1 = 0.
The mesh, the crop and the angles are computed on the device in the
reference's row order (``dsx_beam_directions`` for the UV sphere and the cube
meshes, ``dsx_beam_directions_icosahedral`` for the icosahedral mesh, whose
refinement level comes from ``dsx_icosahedral_nn_dots``);
``beam_directions_grid_device`` keeps the result on the GPU (for
``Rotation.from_euler`` / the spot kernel without a host copy).  The random
mesh draws its normal deviates from numpy's generator on the host, as the
reference does (the draws are the definition of that mesh), and is cropped
and converted on the device (``dsx_beam_directions_from_vertices``).

The SO(3) grids of the same module (rotation_list_generators.py:58-173),
which the reference samples with orix, are built on the device too
(``dsx_so3_grid``): the cubochoric grid (Rosca, Morawiec & De Graef 2014)
cropped to a fundamental zone or to a ball about the identity, as Bunge
angles.  ``get_grid_around_beam_direction`` composes a few hundred
quaternions at most and stays on the host.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib

MESHES = {"uv_sphere": _lib.MESH_UV_SPHERE, "normalized_cube": _lib.MESH_NORMALIZED_CUBE,
          "spherified_cube_edge": _lib.MESH_SPHERIFIED_EDGE,
          "spherified_cube_corner": _lib.MESH_SPHERIFIED_CORNER}

# corners of the stereographic triangles, rotation_list_generators.py:48-55
CRYSTAL_TRIANGLES = {
    "cubic": [(0, 0, 1), (1, 1, 1), (1, 0, 1)],
    "hexagonal": [(0, 0, 0, 1), (9, 1, -10, 0), (2, -1, -1, 0)],
    "trigonal": [(0, 0, 0, 1), (-2, 1, 1, 0), (-1, 2, -1, 0)],
    "tetragonal": [(0, 0, 1), (1, 0, 0), (1, 1, 0)],
    "orthorhombic": [(0, 0, 1), (-1, 0, 0), (0, 1, 0)],
    "monoclinic": [(0, -1, 0), (0, 0, 1), (0, 1, 0)],
}


def _uvtw_to_uvw(d):
    """Four-index direction -> three-index (diffsims/utils/sim_utils.py:496-510)."""
    u, v, t, w = d
    u, v, w = 2 * u + v, 2 * v + u, w
    g = math.gcd(math.gcd(u, v), w)
    return tuple(int(x / g) for x in (u, v, w))


def crop_planes(crystal_system: str):
    """(normal, sign) integer rows of the triangle test s (n . v) >= -1e-13
    (rotation_list_generators.py:230-265).  Triclinic keeps the sphere; the
    monoclinic branch keeps only x >= -1e-13 (its first test's result is
    overwritten in the reference)."""
    if crystal_system == "triclinic":
        return []
    if crystal_system == "monoclinic":
        return [(1, 0, 0, 1)]
    a, b, c = CRYSTAL_TRIANGLES[crystal_system]
    if len(a) == 4:
        a, b, c = _uvtw_to_uvw(a), _uvtw_to_uvw(b), _uvtw_to_uvw(c)
    rows = []
    for p, q, r in ((a, b, c), (b, c, a), (c, a, b)):
        n = (p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0])
        rows.append(n + (n[0] * r[0] + n[1] * r[1] + n[2] * r[2],))
    return rows


# icosahedron faces of get_icosahedral_mesh_vertices (sphere_mesh_generators.py:413-434)
ICOSA_FACES = ((0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
               (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
               (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1))


def icosahedral_edge_order():
    """The icosahedron's 30 edges as sorted corner pairs in the order the
    reference refines them: it collects them in a Python set, face by face
    (sphere_mesh_generators.py:224-232), and iterates that set, so the order
    is CPython's set order for these tuples (fixed: ints hash to themselves
    and tuple hashing is not randomised).  The same construction gives the
    same order here."""
    edges = set()
    for f in ICOSA_FACES:
        for p, q in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            edges.add(tuple(sorted((p, q))))
    return list(edges)


def _edge_array():
    flat = [int(v) for e in icosahedral_edge_order() for v in e]
    return (ctypes.c_int * 60)(*flat)


def icosahedral_level(resolution: float, device=None) -> int:
    """Refinement level of get_icosahedral_mesh_vertices(resolution)
    (sphere_mesh_generators.py:432-438): the first level >= 2 whose largest
    nearest-neighbour angle (degrees(arccos) of dsx_icosahedral_nn_dots) is
    <= resolution.  Levels are evaluated on the device in batches.  At a
    resolution the bare icosahedron (63.43 deg) already meets, the reference's
    refinement loop never runs and it fails; ValueError here."""
    import torch

    from . import engine

    L = _lib.load(require_gpu=True)
    dev = engine._device(device)
    res = float(resolution)
    if not res > 0:
        raise ValueError(f"bad resolution {resolution}")
    edges = _edge_array()
    stream = _lib.stream_handle(dev)
    first, last = 1, min(max(8, int(math.ceil(1.3 * 63.5 / res)) + 2), 16384)
    while True:
        dots = torch.empty(last - first + 1, dtype=torch.float64, device=dev)
        _lib.check(L.dsx_icosahedral_nn_dots(first, last, edges, _lib.ptr(dots), stream),
                   "dsx_icosahedral_nn_dots")
        ang = np.rad2deg(np.arccos(dots.cpu().numpy()))
        if first == 1 and not ang[0] > res:
            raise ValueError(f"resolution {resolution} is not finer than the icosahedron's own "
                             f"{ang[0]:.4f} degrees; the reference's refinement loop never runs")
        for i, a in enumerate(ang):
            if first + i >= 2 and not a > res:
                return first + i
        if last >= 16384:
            raise ValueError(f"resolution {resolution} needs more than 16384 refinement levels")
        first, last = last + 1, min(2 * last, 16384)


def _grid_out(dev, n_cand: int):
    import torch

    L = _lib.load()
    work = torch.empty(max(1, int(L.dsx_grid_workspace_ints(int(n_cand)))), dtype=torch.int32,
                       device=dev)
    out = torch.empty((max(n_cand, 0), 3), dtype=torch.float64, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    return work, out, count


def _planes_arg(planes):
    flat = [int(v) for row in planes for v in row]
    return (ctypes.c_int * max(1, len(flat)))(*flat)


def icosahedral_mesh_device(resolution: float, device=None):
    """get_icosahedral_mesh_vertices (sphere_mesh_generators.py:376-449) as a
    float64 torch tensor [10 n^2 + 2, 3] of unit vectors on the device."""
    from . import engine

    L = _lib.load(require_gpu=True)
    dev = engine._device(device)
    level = icosahedral_level(resolution, dev)
    n = int(L.dsx_icosahedral_vertices(level))
    work, out, count = _grid_out(dev, n)
    _lib.check(L.dsx_beam_directions_icosahedral(level, _edge_array(), 0, _planes_arg([]), 1,
                                                 _lib.ptr(out), n, _lib.ptr(count),
                                                 _lib.ptr(work), work.numel(),
                                                 _lib.stream_handle(dev)),
               "dsx_beam_directions_icosahedral")
    return out


def get_icosahedral_mesh_vertices(resolution: float) -> np.ndarray:
    """Drop-in for the reference function (device mesh copied to the host)."""
    return icosahedral_mesh_device(resolution).cpu().numpy()


def get_random_sphere_vertices(resolution: float, seed=None) -> np.ndarray:
    """get_random_sphere_vertices (sphere_mesh_generators.py:452-483):
    int((360/resolution)^2 / (4 pi)) points from numpy's default_rng(seed)
    standard normal triples, each divided by its norm.  The draws are numpy's
    generator by definition (a seeded call must give the reference's points),
    so this mesh is drawn on the host."""
    number = int(1 / (4 * np.pi) * (360 / resolution) ** 2)
    rng = np.random.default_rng(seed=seed) if seed is not None else np.random.default_rng()
    xyz = rng.normal(size=(number, 3))
    return (xyz.T / np.linalg.norm(xyz, axis=1)).T


def directions_to_euler_device(vectors, planes=(), device=None):
    """Rows of ``vectors`` (n, 3) kept by the crop ``planes`` (crop_planes
    rows) as Bunge angles in degrees, on the device (torch float64 [m, 3]);
    with no planes this is beam_directions_grid_to_euler
    (sphere_mesh_generators.py:486-526)."""
    import torch

    from . import engine

    L = _lib.load(require_gpu=True)
    dev = engine._device(device)
    v = torch.as_tensor(np.ascontiguousarray(np.asarray(vectors, dtype=np.float64).reshape(-1, 3)))
    v = v.to(dev)
    n = int(v.shape[0])
    work, out, count = _grid_out(dev, n)
    _lib.check(L.dsx_beam_directions_from_vertices(_lib.ptr(v), n, len(planes), _planes_arg(planes),
                                                   _lib.ptr(out), n, _lib.ptr(count),
                                                   _lib.ptr(work), work.numel(),
                                                   _lib.stream_handle(dev)),
               "dsx_beam_directions_from_vertices")
    return out[:int(count.item())]


def beam_directions_grid_to_euler(vectors) -> np.ndarray:
    """Drop-in for the reference function (sphere_mesh_generators.py:486-526)."""
    return directions_to_euler_device(vectors).cpu().numpy()


def beam_directions_grid_device(crystal_system: str, resolution: float,
                                mesh: str = "spherified_cube_edge", device=None):
    """The grid as a float64 torch tensor [N, 3] (degrees) on the device."""
    import torch

    from . import engine

    if mesh not in MESHES and mesh not in ("icosahedral", "random"):
        raise NotImplementedError(
            f"The mesh {mesh} is not recognized. Please use: uv_sphere, normalized_cube, "
            f"spherified_cube_edge, spherified_cube_corner, icosahedral, random")
    planes = crop_planes(crystal_system)  # KeyError for an unknown system, as the reference
    if mesh == "random":
        return directions_to_euler_device(get_random_sphere_vertices(resolution), planes, device)
    if mesh == "icosahedral":
        L = _lib.load(require_gpu=True)
        dev = engine._device(device)
        level = icosahedral_level(resolution, dev)
        n = int(L.dsx_icosahedral_vertices(level))
        work, out, count = _grid_out(dev, n)
        _lib.check(L.dsx_beam_directions_icosahedral(
            level, _edge_array(), len(planes), _planes_arg(planes), 0, _lib.ptr(out), n,
            _lib.ptr(count), _lib.ptr(work), work.numel(), _lib.stream_handle(dev)),
            "dsx_beam_directions_icosahedral")
        return out[:int(count.item())]
    res = float(resolution)
    if mesh in ("normalized_cube", "spherified_cube_edge") and crystal_system == "hexagonal":
        res = res / np.sqrt(2)  # rotation_list_generators.py:212-215
    L = _lib.load(require_gpu=True)
    dev = engine._device(device)
    m = MESHES[mesh]
    n_cand = int(L.dsx_beam_directions_candidates(m, res))
    if n_cand < 0:
        raise ValueError(f"bad resolution {resolution}")
    work = torch.empty(max(1, int(L.dsx_beam_directions_workspace_ints(m, res))),
                       dtype=torch.int32, device=dev)
    out = torch.empty((n_cand, 3), dtype=torch.float64, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    flat = [int(v) for row in planes for v in row]
    arr = (ctypes.c_int * max(1, len(flat)))(*flat)
    _lib.check(L.dsx_beam_directions(m, res, len(planes), arr, _lib.ptr(out), n_cand,
                                     _lib.ptr(count), _lib.ptr(work), work.numel(),
                                     _lib.stream_handle(dev)), "dsx_beam_directions")
    return out[:int(count.item())]


def get_beam_directions_grid(crystal_system: str, resolution: float,
                             mesh: str = "spherified_cube_edge") -> np.ndarray:
    """Drop-in for the reference function: the device grid copied to a numpy
    array (N, 3) of Bunge angles in degrees."""
    return beam_directions_grid_device(crystal_system, resolution, mesh).cpu().numpy()


# ---- SO(3) grids (rotation_list_generators.py:58-173) ------------------------

_CUBO_SEMI_EDGE = 0.5 * np.pi ** (2.0 / 3.0)


def cubochoric_steps(resolution: float) -> int:
    """Semi-edge steps of the cubochoric grid for a neighbour distance in
    degrees: a step a/n of the cube is 2 (6/pi)^(1/3) a / n radians of
    rotation next to the identity, and n is the smallest count for which
    that does not exceed the resolution."""
    rad = np.deg2rad(float(resolution))
    if not rad > 0:
        raise ValueError(f"resolution must be positive, got {resolution}")
    n = max(1, int(np.ceil(2.0 * (6.0 / np.pi) ** (1.0 / 3.0) * _CUBO_SEMI_EDGE / rad - 1e-9)))
    if n > 1024:
        raise ValueError(f"resolution {resolution} deg needs {n} > 1024 semi-edge steps")
    return n


def _quat_from_bunge_deg(euler_deg) -> np.ndarray:
    """orix Rotation.from_euler of Bunge angles in degrees."""
    from .rotation import Rotation

    return Rotation.from_euler(np.deg2rad(np.asarray(euler_deg, dtype=np.float64).reshape(3))).data[0]


def _center_quaternion(center):
    """get_local_grid's centre: a Bunge tuple in degrees, or an object
    holding unit quaternions in ``.data`` (an orix Rotation)."""
    if center is None:
        return None
    if isinstance(center, tuple):
        return _quat_from_bunge_deg(center)
    data = getattr(center, "data", None)
    if data is not None and np.asarray(data).reshape(-1, 4).shape[0] == 1:
        return np.asarray(data, dtype=np.float64).reshape(4)
    raise TypeError("center must be a tuple of Bunge angles in degrees or a single rotation "
                    "with unit quaternions in .data")


def _so3_grid_device(n: int, mode: int, ops, half: float, center, device):
    import torch

    from . import engine

    L = _lib.load(require_gpu=True)
    dev = engine._device(device)
    ops = np.ascontiguousarray(np.asarray(ops, dtype=np.float64).reshape(-1, 4))
    ops_arr = (ctypes.c_double * max(1, ops.size))(*ops.ravel().tolist())
    c_arr = None if center is None else (ctypes.c_double * 4)(*np.asarray(center, float).tolist())
    work = torch.empty(int(L.dsx_so3_grid_workspace_ints(n)), dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = _lib.stream_handle(dev)
    _lib.check(L.dsx_so3_grid(n, mode, ops.shape[0], ops_arr, half, c_arr, None, 0,
                              _lib.ptr(count), _lib.ptr(work), work.numel(), stream),
               "dsx_so3_grid")
    m = int(count.item())
    out = torch.empty((max(m, 1), 3), dtype=torch.float64, device=dev)
    _lib.check(L.dsx_so3_grid(n, mode, ops.shape[0], ops_arr, half, c_arr, _lib.ptr(out), m,
                              _lib.ptr(count), _lib.ptr(work), work.numel(), stream),
               "dsx_so3_grid")
    return out[:m]


def fundamental_zone_grid_device(resolution: float = 2, space_group: int = None, device=None):
    """Rotations of the cubochoric grid inside the fundamental zone of the
    space group's proper point group (orix get_sample_fundamental) as Bunge
    angles in radians, a float64 torch tensor [n, 3] on the device."""
    from . import symmetry

    if space_group is None:
        raise ValueError("get_fundamental_zone_grid needs a space group number: the "
                         "fundamental zone is taken from space_group alone")
    ops = symmetry.rotation_group_quaternions(symmetry.rotation_group_of_space_group(space_group))
    return _so3_grid_device(cubochoric_steps(resolution), _lib.SO3_FUNDAMENTAL_ZONE, ops[1:],
                            0.0, None, device)


def local_grid_device(resolution: float = 2, center=None, grid_width: float = 10, device=None):
    """Rotations of the cubochoric grid by less than grid_width degrees,
    premultiplied by the centre rotation (orix get_sample_local), as Bunge
    angles in radians, a float64 torch tensor [n, 3] on the device."""
    half = float(np.deg2rad(grid_width / 2))
    return _so3_grid_device(cubochoric_steps(resolution), _lib.SO3_LOCAL, np.zeros((0, 4)), half,
                            _center_quaternion(center), device)


def get_list_from_orix(grid, rounding: int = 2):
    """get_list_from_orix (rotation_list_generators.py:58-82): Bunge angles
    in degrees, rounded to `rounding` decimals, one tuple per rotation.
    ``grid`` is anything with ``to_euler()`` (radians) or an array / tensor
    [n, 3] of Bunge angles in radians."""
    if hasattr(grid, "to_euler"):
        z = grid.to_euler()
    elif hasattr(grid, "detach"):
        z = grid.detach().cpu().numpy()
    else:
        z = grid
    rows = np.asarray(z, dtype=np.float64).reshape(-1, 3).tolist()
    return [tuple(np.round(np.rad2deg(r), decimals=rounding)) for r in rows]


def get_fundamental_zone_grid(resolution=2, point_group=None, space_group=None):
    """Drop-in for get_fundamental_zone_grid (rotation_list_generators.py:
    85-106): a list of Bunge tuples (degrees, 2 decimals) inside the
    fundamental zone.  As in the reference, only ``space_group`` selects the
    zone (``point_group`` is accepted and not used)."""
    return get_list_from_orix(fundamental_zone_grid_device(resolution, space_group), rounding=2)


def get_local_grid(resolution=2, center=None, grid_width=10):
    """Drop-in for get_local_grid (rotation_list_generators.py:109-134)."""
    return get_list_from_orix(local_grid_device(resolution, center, grid_width), rounding=2)


def get_grid_around_beam_direction(beam_rotation, resolution, angular_range=(0, 360)):
    """Drop-in for get_grid_around_beam_direction (rotation_list_generators.py:
    137-173): the beam rotation (Bunge degrees) times rotations about z by
    arange(start, stop, resolution) degrees, as Bunge tuples.  A few hundred
    rotations: host FP64."""
    from .rotation import Rotation

    beam = Rotation.from_euler(np.deg2rad(np.asarray(beam_rotation, dtype=np.float64)))
    ang = np.deg2rad(np.arange(start=angular_range[0], stop=angular_range[1], step=resolution))
    axes = np.repeat([[0, 0, 1]], ang.shape[0], axis=0)
    return get_list_from_orix(beam * Rotation.from_axes_angles(axes, ang), rounding=2)
