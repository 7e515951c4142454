"""Scattering tables and the C ABI (CPU only: nothing here launches a kernel)."""

import ast
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from diffsims_amd import _lib
from diffsims_amd.scattering import symbol, table

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diffsims_amd.h")
REF = "/root/reference"


def test_table_sizes_and_known_values():
    lob, xt = table("lobato"), table("xtables")
    assert len(lob) == 103 and len(xt) == 98
    assert symbol(26) == "Fe" and symbol(1) == "H" and symbol(103) == "Lr"
    # Ni f(0) = 2 sum a (the 43.0979 = f(0)^2 KAT of test_sim_utils.py:330-353)
    a = lob["Ni"][0::2]
    assert abs((2 * a.sum()) ** 2 - 43.0979) < 1e-4
    # ITC H row (atomic_scattering_params.py, first entry)
    np.testing.assert_allclose(xt["H"], [0.0088, 0.1152, 0.0449, 1.0867, 0.1481, 4.9755,
                                         0.2356, 16.5591, 0.0914, 43.2743])


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
@pytest.mark.parametrize("name, rel", [("lobato", "diffsims/utils/lobato_scattering_params.py"),
                                       ("xtables", "diffsims/utils/atomic_scattering_params.py")])
def test_tables_match_reference_text(name, rel):
    src = open(os.path.join(REF, rel)).read()
    node = next(n for n in ast.parse(src).body if isinstance(n, ast.Assign))
    rows = ast.literal_eval(node.value.args[0])
    t = table(name)
    assert [r[0] for r in rows] == list(t.keys())
    for el, pairs in rows:
        flat = np.array([v for p in pairs for v in p] + [0.0] * (10 - 2 * len(pairs)))
        assert np.array_equal(flat, t[el]), el


def _declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dsx_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_are_bound():
    assert sorted(_lib.EXPORTED) == _declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared_symbols():
        assert hasattr(lib, name), name
    assert lib.dsx_version() == 100
    assert lib.dsx_render_max_spots() >= 256
    assert _lib.last_error() == ""


def test_entry_points_reject_bad_arguments_without_a_gpu():
    lib = _lib.load()
    h = (ctypes.c_int * 3)(-1, 0, 0)
    m = (ctypes.c_double * 6)(*([1.0] * 6))
    assert lib.dsx_hkl_enumerate(h, m, 1.0, None, 0, 1, 1, None) == _lib.DSX_EINVAL
    assert "bad box" in _lib.last_error()
    p = _lib.SimParams()
    assert lib.dsx_simulate(None, None, None, 0, None, 0, None, 1, ctypes.byref(p), 0, None, None,
                            None, None, None, None) == _lib.DSX_EINVAL
    assert lib.dsx_simulate(None, None, None, 0, None, 3, None, 1, ctypes.byref(p), 1, 1, 1, 1, 1,
                            1, None) == _lib.DSX_EINVAL
    assert lib.dsx_hkl_lines(None, None, -1, None, 0, 1, None, 0, None) == _lib.DSX_EINVAL
    assert lib.dsx_hkl_lines(1, 1, 1000, None, 0, 1, 1, 1, None) == _lib.DSX_EINVAL
    assert "workspace" in _lib.last_error()
    assert lib.dsx_hkl_lines_workspace_ints(1000) == 2 * 4
    assert lib.dsx_hkl_lines_workspace_ints(-1) == -1
    rp = _lib.RenderParams()
    rp.height = 0
    assert lib.dsx_render_fast(1, 1, 1, 1, ctypes.byref(rp), 1, None, None) == _lib.DSX_EINVAL
    assert lib.dsx_render_fast(None, None, 1, 0, ctypes.byref(rp), None, None, None) == _lib.DSX_OK


def test_struct_layouts_match_the_header(tmp_path):
    """Compile a probe against the header with gcc and compare sizes/offsets
    with the ctypes mirrors."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "diffsims_amd.h"\n'
        "int main(void){\n"
        'printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(dsx_sim_params), offsetof(dsx_sim_params, rec),'
        " sizeof(dsx_render_params), offsetof(dsx_render_params, taps), sizeof(dsx_phase_desc),"
        " offsetof(dsx_phase_desc, n_lines), offsetof(dsx_phase_desc, lines));\n"
        "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals == [ctypes.sizeof(_lib.SimParams), _lib.SimParams.rec.offset,
                    ctypes.sizeof(_lib.RenderParams), _lib.RenderParams.taps.offset,
                    ctypes.sizeof(_lib.PhaseDesc), _lib.PhaseDesc.n_lines.offset,
                    _lib.PhaseDesc.lines.offset]


def test_beam_direction_mesh_sizes_on_the_host():
    """dsx_beam_directions_candidates (host arithmetic) equals the oracle's
    mesh sizes: every cube-mesh vertex is a candidate; the UV sphere drops
    all but one point per pole from its psi x theta product."""
    from oracle import diffsims_oracle as O

    lib = _lib.load()
    for res in (10.0, 3.0, 1.3):
        for mesh, grid_type in ((_lib.MESH_NORMALIZED_CUBE, "normalized"),
                                (_lib.MESH_SPHERIFIED_EDGE, "spherified_edge"),
                                (_lib.MESH_SPHERIFIED_CORNER, "spherified_corner")):
            assert lib.dsx_beam_directions_candidates(mesh, res) == O.cube_mesh_vertices(res, grid_type).shape[0]
        n_theta, n_psi = int(np.ceil(180 / res)) + 1, int(np.ceil(360 / res))
        assert lib.dsx_beam_directions_candidates(_lib.MESH_UV_SPHERE, res) == n_theta * n_psi
        assert O.uv_sphere_vertices(res).shape[0] == n_psi * (n_theta - 2) + 2
    assert lib.dsx_beam_directions_candidates(7, 1.0) == -1
    assert lib.dsx_beam_directions_candidates(_lib.MESH_UV_SPHERE, 0.0) == -1
    text = open(HEADER).read()
    for name in ("UV_SPHERE", "NORMALIZED_CUBE", "SPHERIFIED_EDGE", "SPHERIFIED_CORNER"):
        v = int(re.search(rf"#define DSX_MESH_{name} (\d+)", text).group(1))
        assert v == getattr(_lib, f"MESH_{name}")
    v = int(re.search(r"#define DSX_LINE_SWEEP_MIN_ROWS (\d+)", text).group(1))
    assert v == _lib.LINE_SWEEP_MIN_ROWS


def test_crop_planes_match_the_oracle_triangles():
    """The integer crop rows equal the oracle's triangle normals and signs."""
    from oracle import diffsims_oracle as O
    from diffsims_amd import generators

    for cs, corners in O.CRYSTAL_TRIANGLES.items():
        rows = generators.crop_planes(cs)
        if cs == "monoclinic":
            assert rows == [(1, 0, 0, 1)]
            continue
        a, b, c = corners
        if len(a) == 4:
            a, b, c = O.uvtw_to_uvw(a), O.uvtw_to_uvw(b), O.uvtw_to_uvw(c)
        for row, (p, q, r) in zip(rows, ((a, b, c), (b, c, a), (c, a, b))):
            n = np.cross(p, q)
            assert row == tuple(int(x) for x in n) + (int(np.dot(n, r)),)
    assert generators.crop_planes("triclinic") == []
