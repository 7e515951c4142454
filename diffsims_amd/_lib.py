"""ctypes binding of the C ABI in ``include/diffsims_amd.h``.

The native library is the only compute path: when it is missing, or no GPU is
visible, every device entry point raises instead of computing anything on the
CPU.  torch is imported before the library is loaded so that the library binds
to the HIP runtime torch already loaded (same soname), keeping one runtime per
process.
"""

from __future__ import annotations

import ctypes
import os

LIB_PATH = os.environ.get(
    "DIFFSIMS_AMD_LIB",
    os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdiffsims_amd.so"))

DSX_OK = 0
DSX_EINVAL = -1
DSX_ELAUNCH = -2
DSX_ECAPACITY = -3

SF_LINEAR = 0
SF_LORENTZIAN = 1
SF_SINC = 2
SF_SIN2C = 3
SF_ATANC = 4
SF_BINARY = 5
SF_LORENTZIAN_PRECESSION = 6
SF_PRECESSION_AVG = 16

SCATTER_NONE = 0
SCATTER_LOBATO = 1
SCATTER_XTABLES = 2

MAX_TAP_RADIUS = 64

ADJ_L2 = 0
ADJ_PATTERN_GRAD = 1
ADJ_FORWARD = 2

RENDER_SPLAT = 0
RENDER_DENSE = 1
RENDER_DENSE_INPLACE = 2

LINE_SWEEP_MIN_ROWS = 10  # DSX_LINE_SWEEP_MIN_ROWS
MESH_UV_SPHERE = 0
MESH_NORMALIZED_CUBE = 1
MESH_SPHERIFIED_EDGE = 2
MESH_SPHERIFIED_CORNER = 3

SO3_FUNDAMENTAL_ZONE = 0
SO3_LOCAL = 1

# Every symbol the header declares (checked by the CPU test suite).
EXPORTED = (
    "dsx_version",
    "dsx_last_error",
    "dsx_hkl_workspace_ints",
    "dsx_hkl_enumerate",
    "dsx_hkl_points_in_sphere",
    "dsx_structure_factor",
    "dsx_stage_atoms",
    "dsx_hkl_lines_workspace_ints",
    "dsx_hkl_lines",
    "dsx_beam_directions_candidates",
    "dsx_beam_directions_workspace_ints",
    "dsx_beam_directions",
    "dsx_grid_workspace_ints",
    "dsx_icosahedral_vertices",
    "dsx_icosahedral_nn_dots",
    "dsx_beam_directions_icosahedral",
    "dsx_beam_directions_from_vertices",
    "dsx_so3_grid_candidates",
    "dsx_so3_grid_workspace_ints",
    "dsx_so3_grid",
    "dsx_simulate",
    "dsx_render_max_spots",
    "dsx_render_fast",
    "dsx_render_subpixel",
    "dsx_render_adjoint",
    "dsx_render_adjoint_workspace_ints",
    "dsx_render_adjoint_noalloc_max_ints",
    "dsx_render_adjoint_noalloc",
    "dsx_structure_factor_adjoint",
    "dsx_simulate_tile_size",
    "dsx_simulate_ragged",
    "dsx_render_dense_max_pixels",
    "dsx_render_fast_ragged",
    "dsx_render_dense_inplace_fits",
    "dsx_polar_templates",
    "dsx_render_fast_large_workspace_ints",
    "dsx_render_fast_large",
    "dsx_render_fast_bands_workspace_ints",
    "dsx_render_fast_bands",
    "dsx_euler_to_rt",
    "dsx_euler_to_rt_backward",
)


class SimParams(ctypes.Structure):
    """dsx_sim_params."""
    _fields_ = [
        ("k", ctypes.c_double),
        ("max_excitation_error", ctypes.c_double),
        ("shape_width", ctypes.c_double),
        ("shape_model", ctypes.c_int),
        ("minima_number", ctypes.c_double),
        ("precession_angle", ctypes.c_double),
        ("minimum_intensity", ctypes.c_double),
        ("g_max", ctypes.c_double),
        ("rec", ctypes.c_double * 9),
    ]


class PhaseDesc(ctypes.Structure):
    """dsx_phase_desc (packed on the host, copied to the device as bytes)."""
    _fields_ = [
        ("hkl", ctypes.c_void_p),
        ("g4", ctypes.c_void_p),
        ("F2", ctypes.c_void_p),
        ("n_refl", ctypes.c_int),
        ("cap", ctypes.c_int),
        ("rec", ctypes.c_double * 9),
        ("g_max", ctypes.c_double),
        ("Rt", ctypes.c_void_p),
        ("n_orient", ctypes.c_int),
        ("n_lines", ctypes.c_int),
        ("idx", ctypes.c_void_p),
        ("spots", ctypes.c_void_p),
        ("count", ctypes.c_void_p),
        ("ncand", ctypes.c_void_p),
        ("maxI", ctypes.c_void_p),
        ("lines", ctypes.c_void_p),
    ]


class RenderParams(ctypes.Structure):
    """dsx_render_params."""
    _fields_ = [
        ("height", ctypes.c_int),
        ("width", ctypes.c_int),
        ("calibration", ctypes.c_float),
        ("center_x", ctypes.c_float),
        ("center_y", ctypes.c_float),
        ("cos_angle", ctypes.c_float),
        ("sin_angle", ctypes.c_float),
        ("mirrored", ctypes.c_int),
        ("normalize", ctypes.c_int),
        ("sigma", ctypes.c_float),
        ("clip_threshold", ctypes.c_float),
        ("radius", ctypes.c_int),
        ("taps", ctypes.c_float * (MAX_TAP_RADIUS + 1)),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_SIGS = {
    "dsx_version": (_I, []),
    "dsx_last_error": (ctypes.c_char_p, []),
    "dsx_hkl_workspace_ints": (ctypes.c_size_t, [ctypes.POINTER(_I)]),
    "dsx_hkl_enumerate": (_I, [ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double),
                               ctypes.c_double, _P, _I, _P, _P, _P]),
    "dsx_hkl_points_in_sphere": (_I, [ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double),
                                      ctypes.c_double, _P, _I, _P, _P, _P]),
    "dsx_structure_factor": (_I, [_P, _I, ctypes.POINTER(ctypes.c_double), _P, _P, _P, _P, _I,
                                  _P, _I, _I, _P, _P, _P, _P]),
    "dsx_stage_atoms": (_I, [_P, _P, _P, _I, _P, _P, _P]),
    "dsx_beam_directions_candidates": (ctypes.c_longlong, [_I, ctypes.c_double]),
    "dsx_beam_directions_workspace_ints": (ctypes.c_longlong, [_I, ctypes.c_double]),
    "dsx_beam_directions": (_I, [_I, ctypes.c_double, _I, ctypes.POINTER(ctypes.c_int), _P,
                                 ctypes.c_longlong, _P, _P, ctypes.c_longlong, _P]),
    "dsx_grid_workspace_ints": (ctypes.c_longlong, [ctypes.c_longlong]),
    "dsx_icosahedral_vertices": (ctypes.c_longlong, [_I]),
    "dsx_icosahedral_nn_dots": (_I, [_I, _I, ctypes.POINTER(ctypes.c_int), _P, _P]),
    "dsx_beam_directions_icosahedral": (_I, [_I, ctypes.POINTER(ctypes.c_int), _I,
                                             ctypes.POINTER(ctypes.c_int), _I, _P,
                                             ctypes.c_longlong, _P, _P, ctypes.c_longlong, _P]),
    "dsx_beam_directions_from_vertices": (_I, [_P, ctypes.c_longlong, _I,
                                               ctypes.POINTER(ctypes.c_int), _P,
                                               ctypes.c_longlong, _P, _P, ctypes.c_longlong,
                                               _P]),
    "dsx_so3_grid_candidates": (ctypes.c_longlong, [_I]),
    "dsx_so3_grid_workspace_ints": (ctypes.c_longlong, [_I]),
    "dsx_so3_grid": (_I, [_I, _I, _I, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                          ctypes.POINTER(ctypes.c_double), _P, ctypes.c_longlong, _P, _P,
                          ctypes.c_longlong, _P]),
    "dsx_hkl_lines_workspace_ints": (ctypes.c_longlong, [_I]),
    "dsx_hkl_lines": (_I, [_P, _P, _I, _P, _I, _P, _P, ctypes.c_longlong, _P]),
    "dsx_simulate": (_I, [_P, _P, _P, _I, _P, _I, _P, _I, ctypes.POINTER(SimParams), _I, _P, _P,
                          _P, _P, _P, _P]),
    "dsx_render_max_spots": (_I, []),
    "dsx_render_fast": (_I, [_P, _P, _I, _I, ctypes.POINTER(RenderParams), _P, _P, _P]),
    "dsx_render_subpixel": (_I, [_P, _P, _I, _I, ctypes.POINTER(RenderParams), _P, _P, _P]),
    "dsx_render_adjoint": (_I, [_P, _P, _P, _I, _I, _P, _P, _I, _P, ctypes.POINTER(SimParams),
                                ctypes.POINTER(RenderParams), _I, _P, ctypes.c_longlong, _P, _P,
                                _P, _P, _P, _P, ctypes.c_longlong, _P, _P]),
    "dsx_render_adjoint_workspace_ints": (ctypes.c_longlong, [_I, _I, _I, _I, _I]),
    "dsx_render_adjoint_noalloc_max_ints": (ctypes.c_longlong, []),
    "dsx_render_adjoint_noalloc": (_I, [_P, _P, _P, _I, _I, _P, _P, _I, _P,
                                        ctypes.POINTER(SimParams), ctypes.POINTER(RenderParams),
                                        _I, _P, ctypes.c_longlong, _P, _P, _P, _P, _P, _P, _P]),
    "dsx_structure_factor_adjoint": (_I, [_P, _I, ctypes.POINTER(ctypes.c_double), _P, _P, _P, _P,
                                          _I, _P, _I, _I, _P, _P, _P, _P, _P]),
    "dsx_simulate_tile_size": (_I, []),
    "dsx_simulate_ragged": (_I, [_P, _I, _P, _I, ctypes.POINTER(SimParams), _I, _P]),
    "dsx_render_dense_max_pixels": (_I, []),
    "dsx_render_dense_inplace_fits": (_I, [_I, _I, _I]),
    "dsx_render_fast_large_workspace_ints": (ctypes.c_longlong, [_I]),
    "dsx_render_fast_bands_workspace_ints": (ctypes.c_longlong, [_I, _I, _I]),
    "dsx_render_fast_bands": (_I, [_P, _P, _I, _I, ctypes.POINTER(RenderParams), _P,
                                   ctypes.c_longlong, _P, _P, _P]),
    "dsx_euler_to_rt": (_I, [_P, _I, _P, _P]),
    "dsx_euler_to_rt_backward": (_I, [_P, _P, _I, _P, _P]),
    "dsx_render_fast_large": (_I, [_P, _P, _I, _I, ctypes.POINTER(RenderParams), _P,
                                   ctypes.c_longlong, _P, _P, _P]),
    "dsx_polar_templates": (_I, [_P, _P, _I, _I, _I, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "dsx_render_fast_ragged": (_I, [_P, _P, _P, _P, _P, _I, ctypes.POINTER(RenderParams), _I, _I,
                                    _P, _P, _P]),
}

_lib = None
_gpu_seen = False  # a GPU was visible once (it stays visible for the process)


def load(require_gpu: bool = False):
    """Load the native library (once).  Raises RuntimeError when it is absent."""
    global _lib, _gpu_seen
    if _lib is None:
        import torch  # noqa: F401  (binds the library to torch's HIP runtime)

        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"diffsims_amd native library not found at {LIB_PATH}; build it with "
                "`python -m diffsims_amd.build`")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    if require_gpu and not _gpu_seen:
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("diffsims_amd needs a ROCm GPU; none is visible")
        _gpu_seen = True
    return _lib


def last_error() -> str:
    return load().dsx_last_error().decode(errors="replace")


def check(status: int, what: str) -> None:
    """Map a DSX status to the Python exception the reference would raise."""
    if status == DSX_OK:
        return
    msg = f"{what}: {last_error()}"
    if status == DSX_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def ptr(t) -> int:
    """Device pointer of a torch tensor (0 for None)."""
    return 0 if t is None else t.data_ptr()


def stream_handle(device=None) -> int:
    """hipStream_t of torch's current stream on ``device`` (a torch.device,
    device string or index; the current device when None), as an int.  Read
    with torch's raw-stream accessor, which returns the same handle as
    ``torch.cuda.current_stream(device).cuda_stream`` without building a
    Stream object (a few microseconds per launch on the refinement path)."""
    import torch

    if device is None:
        idx = torch.cuda.current_device()
    elif isinstance(device, int):
        idx = device
    else:
        d = device if isinstance(device, torch.device) else torch.device(device)
        idx = d.index if d.index is not None else torch.cuda.current_device()
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is None:  # a torch build without the accessor
        return torch.cuda.current_stream(idx).cuda_stream
    return raw(idx)
