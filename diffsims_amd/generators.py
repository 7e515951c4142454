"""Beam-direction grids of template libraries on the device.

``get_beam_directions_grid`` mirrors
diffsims.generators.rotation_list_generators.get_beam_directions_grid
(diffsims/generators/rotation_list_generators.py:176-267): a sphere mesh
(diffsims/generators/sphere_mesh_generators.py), cropped to the crystal
system's stereographic triangle, as Bunge angles in degrees with phi1 = 0.
The mesh, the crop and the angles are computed by ``dsx_beam_directions`` in
the reference's row order; ``beam_directions_grid_device`` keeps the result on
the GPU (for ``Rotation.from_euler`` / the spot kernel without a host copy).
The icosahedral and random meshes are not built on the device (the first
refines until a nearest-neighbour angle test passes, the second draws from
numpy's generator); they raise NotImplementedError here.
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


def beam_directions_grid_device(crystal_system: str, resolution: float,
                                mesh: str = "spherified_cube_edge", device=None):
    """The grid as a float64 torch tensor [N, 3] (degrees) on the device."""
    import torch

    from . import engine

    if mesh in ("icosahedral", "random"):
        raise NotImplementedError(f"the {mesh} mesh is not generated on the device")
    if mesh not in MESHES:
        raise NotImplementedError(
            f"The mesh {mesh} is not recognized. Please use: uv_sphere, normalized_cube, "
            f"spherified_cube_edge, spherified_cube_corner, icosahedral, random")
    planes = crop_planes(crystal_system)  # KeyError for an unknown system, as the reference
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
