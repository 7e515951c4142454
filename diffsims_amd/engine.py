"""Device pipeline: per-phase reflection tables, per-orientation spot lists and
rendered patterns, all resident in HBM as torch tensors.

This is the array-level layer under the drop-in ``SimulationGenerator`` /
``Simulation2D`` classes; every computation runs in the HIP kernels of
``libdiffsims_amd.so`` (no CPU path).

Layout in HBM:
  PhaseTable   hkl  int32 [N, 4]   {h, k, l, 0}, reference order, then the
                                    direct-beam rows when with_direct_beam
               g4   float32 [N, 4] {gx, gy, gz, |g|^2} (crystal frame, 1/A)
               F2   float32 [N]    |F(hkl)|^2
  SpotBatch    idx  int32 [n, cap]     row into the table
               spots float32 [n, cap, 4] {x, y, z, I} lab frame
               count/ncand int32 [n], maxI float32 [n]
  patterns     float32 [n, H, W]
"""

from __future__ import annotations

import ctypes
import math
import warnings
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib
from .crystallography import (PhaseArrays, element_of, highest_hkl, lattice_setup,
                              reciprocal_metric)
from .scattering import electron_wavelength, table

SHAPE_MODEL_IDS = {
    "linear": _lib.SF_LINEAR,
    "lorentzian": _lib.SF_LORENTZIAN,
    "sinc": _lib.SF_SINC,
    "sin2c": _lib.SF_SIN2C,
    "atanc": _lib.SF_ATANC,
    "binary": _lib.SF_BINARY,
}

SCATTER_IDS = {None: _lib.SCATTER_NONE, "lobato": _lib.SCATTER_LOBATO,
               "xtables": _lib.SCATTER_XTABLES}

# Fast rendering of small batches: below SPLAT_MAX_PATTERNS patterns of at
# most SPLAT_MAX_PIXELS pixels the persistent splat kernel (dsx_render_fast,
# one launch) finishes before the band renderer (dsx_render_fast_bands: four
# launches, and a plan wave per pattern whose latency is not hidden by a
# small grid).  Measured with HIP events, both giving the same bits:
# 500 x 144x144 with ~40 spots 26 us (splat) against 38 us (bands), 500 x
# 256x256 with ~260 spots 47 us against 58 us; at 1,000 patterns the two are
# equal, and from 2,000 on the band writer is faster (32,768 x 256x256:
# 1.63 ms against 1.33 ms).  Larger patterns always take the band renderer,
# since a few hundred persistent blocks cannot keep the stores in flight.
SPLAT_MAX_PATTERNS = 768
SPLAT_MAX_PIXELS = 256 * 256


def fast_renderer(n: int, height: int, width: int) -> str:
    """Which entry point ``render_spots`` uses for ``n`` fast patterns of at
    most dsx_render_max_spots() spots: "dsx_render_fast" or
    "dsx_render_fast_bands" (the same output bit for bit)."""
    if n < SPLAT_MAX_PATTERNS and height * width <= SPLAT_MAX_PIXELS:
        return "dsx_render_fast"
    return "dsx_render_fast_bands"


def _device(device=None) -> torch.device:
    _lib.load(require_gpu=True)
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    d = torch.device(device)
    if d.type != "cuda":
        raise ValueError(f"diffsims_amd runs on ROCm GPUs only, got device {d}")
    return d


@dataclass
class PhaseTable:
    """Device-resident reflection table of one phase."""
    hkl: torch.Tensor
    g4: torch.Tensor
    F2: torch.Tensor
    F: torch.Tensor
    n_box: int              # rows from the hkl box (before direct-beam rows)
    with_direct_beam: bool
    g_max: float
    rec: np.ndarray         # FP64 reciprocal basis (columns a*, b*, c*)
    base: np.ndarray
    name: str = ""
    lines: Optional[torch.Tensor] = None  # line index (dsx_hkl_lines), int32 [n_lines, 8]

    @property
    def n(self) -> int:
        return int(self.hkl.shape[0])

    @property
    def n_lines(self) -> int:
        return 0 if self.lines is None else int(self.lines.shape[0])

    def hkl_host(self) -> np.ndarray:
        return self.hkl[:, :3].cpu().numpy()


def _fixed_point(frac: np.ndarray) -> np.ndarray:
    """Fractional coordinates as uint32 turns: round(frac * 2^32) mod 2^32."""
    f = np.asarray(frac, dtype=np.float64)
    f = f - np.floor(f)
    return (np.round(f * 4294967296.0).astype(np.uint64) & 0xFFFFFFFF).astype(np.uint32)


def build_lines(hkl: torch.Tensor, g4: torch.Tensor, n_lines: Optional[int] = None,
                stream=None) -> torch.Tensor:
    """Line index of a reflection table (dsx_hkl_lines): int32 [n_lines, 8].
    With ``n_lines`` known (same hkl rows as an earlier call) nothing is read
    back; otherwise the count is read once."""
    L = _lib.load(require_gpu=True)
    dev = hkl.device
    n = int(hkl.shape[0])
    stream = _lib.stream_handle(dev) if stream is None else stream
    wn = int(L.dsx_hkl_lines_workspace_ints(n))
    work = torch.empty(max(wn, 1), dtype=torch.int32, device=dev)
    count = torch.empty(1, dtype=torch.int32, device=dev)
    if n_lines is None:
        _lib.check(L.dsx_hkl_lines(_lib.ptr(hkl), _lib.ptr(g4), n, None, 0, _lib.ptr(count),
                                   _lib.ptr(work), wn, stream), "dsx_hkl_lines")
        n_lines = int(count.item())
    lines = torch.empty((n_lines, 8), dtype=torch.int32, device=dev)
    _lib.check(L.dsx_hkl_lines(_lib.ptr(hkl), _lib.ptr(g4), n, _lib.ptr(lines), n_lines,
                               _lib.ptr(count), _lib.ptr(work), wn, stream), "dsx_hkl_lines")
    return lines


def build_phase_table(pa: PhaseArrays, reciprocal_radius: float, with_direct_beam: bool,
                      scattering_params: Optional[str] = "lobato", device=None,
                      warn_missing: bool = True, d_min: Optional[float] = None) -> PhaseTable:
    """hkl set + g vectors + |F|^2 of a phase on the GPU.

    Follows DiffractingVector.from_min_dspacing(phase, 1/reciprocal_radius,
    include_zero_vector=with_direct_beam) and the structure-factor part of
    get_kinematical_intensities (reciprocal_lattice_vector.py:1077-1142,
    sim_utils.py:256-354); the extra (000) row the reference stacks on in
    get_intersecting_reflections (simulation_generator.py:351-353) is included.
    ``d_min`` (default 1 / reciprocal_radius) is the box's d-spacing bound.
    """
    dev = _device(device)
    L = _lib.load()
    if not reciprocal_radius > 0:
        raise ValueError("reciprocal_radius must be positive")
    if d_min is None:
        d_min = 1.0 / reciprocal_radius
    base, rec = lattice_setup(pa.abc, pa.angles)
    metric = reciprocal_metric(rec)
    H = highest_hkl(rec, d_min)
    h3 = (ctypes.c_int * 3)(*[int(v) for v in H])
    m6 = (ctypes.c_double * 6)(*metric.tolist())
    stream = _lib.stream_handle(dev)
    work = torch.empty(int(L.dsx_hkl_workspace_ints(h3)), dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(L.dsx_hkl_enumerate(h3, m6, d_min, None, 0, _lib.ptr(count), _lib.ptr(work), stream),
               "dsx_hkl_enumerate")
    n_box = int(count.item())
    n_extra = 2 if with_direct_beam else 0
    hkl = torch.zeros((n_box + n_extra, 4), dtype=torch.int32, device=dev)
    _lib.check(L.dsx_hkl_enumerate(h3, m6, d_min, _lib.ptr(hkl), n_box, _lib.ptr(count),
                                   _lib.ptr(work), stream), "dsx_hkl_enumerate")

    g4, F2, F = _structure_factor(pa, hkl, rec, scattering_params, dev, warn_missing, stream)
    N = n_box + n_extra
    g_max = float(torch.sqrt(g4[:, 3].max()).item()) if N else 0.0
    return PhaseTable(hkl, g4, F2, F, n_box, with_direct_beam, g_max, rec, base, pa.name,
                      build_lines(hkl, g4, stream=stream))


def _structure_factor(pa: PhaseArrays, hkl: torch.Tensor, rec: np.ndarray,
                      scattering_params: Optional[str], dev, warn_missing: bool, stream):
    """g vectors and |F|^2, F of the rows of ``hkl`` (int32 [N, 4]) for the
    phase's atoms (dsx_structure_factor); g = rec hkl."""
    L = _lib.load()
    # atoms sorted by element so the kernel evaluates f once per element
    n_atoms = len(pa.symbols)
    elements = [element_of(s) for s in pa.symbols]
    if scattering_params is not None:
        tab = table(scattering_params)
    else:
        tab = {}
    uniq = sorted(set(elements))
    coeff = np.zeros((max(len(uniq), 1), 10), dtype=np.float64)
    for i, el in enumerate(uniq):
        if scattering_params is None:
            continue
        if el in tab:
            coeff[i] = tab[el]
        elif warn_missing:
            sym = pa.symbols[elements.index(el)]
            warnings.warn(f"Element {el} from atom type symbol {sym} not found in scattering "
                          "parameter library.", UserWarning)
    el_idx = np.array([uniq.index(e) for e in elements], dtype=np.int32)
    order = np.argsort(el_idx, kind="stable")
    xyz_fx = np.zeros((max(n_atoms, 1), 4), dtype=np.uint32)
    if n_atoms:
        xyz_fx[:n_atoms, :3] = _fixed_point(pa.xyz[order])
    occ = np.asarray(pa.occ, dtype=np.float32)[order] if n_atoms else np.zeros(1, np.float32)
    B = np.asarray(pa.B, dtype=np.float32)[order] if n_atoms else np.zeros(1, np.float32)
    el_sorted = el_idx[order] if n_atoms else np.zeros(1, np.int32)

    t_xyz = torch.from_numpy(xyz_fx.view(np.int32)).to(dev)
    t_el = torch.from_numpy(np.ascontiguousarray(el_sorted)).to(dev)
    t_occ = torch.from_numpy(np.ascontiguousarray(occ)).to(dev)
    t_B = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    t_coeff = torch.from_numpy(coeff).to(dev)
    N = int(hkl.shape[0])
    g4 = torch.empty((N, 4), dtype=torch.float32, device=dev)
    F2 = torch.empty(N, dtype=torch.float32, device=dev)
    F = torch.empty((N, 2), dtype=torch.float32, device=dev)
    rec9 = (ctypes.c_double * 9)(*np.asarray(rec, dtype=np.float64).reshape(-1).tolist())
    _lib.check(L.dsx_structure_factor(_lib.ptr(hkl), N, rec9, _lib.ptr(t_xyz), _lib.ptr(t_el),
                                      _lib.ptr(t_occ), _lib.ptr(t_B), n_atoms, _lib.ptr(t_coeff),
                                      len(uniq), SCATTER_IDS[scattering_params], _lib.ptr(g4),
                                      _lib.ptr(F2), _lib.ptr(F), stream), "dsx_structure_factor")
    return g4, F2, F


def diffpy_standard_base(abc, angles) -> np.ndarray:
    """Rows a, b, c of diffpy.structure's standard Cartesian base (a* || x,
    c || z, b in the yz-plane), the frame of the legacy DiffractionGenerator
    (its structures are diffpy lattices with no base rotation):
    a = (1/a*, -cos(gamma*)/(sin(gamma*) a*), a cos(beta)),
    b = (0, b sin(alpha), b cos(alpha)), c = (0, 0, c)."""
    from .crystallography import cosd, sind

    a, b, c = (float(v) for v in abc)
    al, be, ga = (float(v) for v in angles)
    ca, cb, cg = cosd(al), cosd(be), cosd(ga)
    sa, sb, sg = sind(al), sind(be), sind(ga)
    vunit = np.sqrt(1.0 + 2.0 * ca * cb * cg - ca * ca - cb * cb - cg * cg)
    ar = sa / (a * vunit)
    cgr = (ca * cb - cg) / (sa * sb)
    sgr = np.sqrt(1.0 - cgr * cgr)
    return np.array([[1.0 / ar, -cgr / sgr / ar, cb * a],
                     [0.0, b * sa, b * ca],
                     [0.0, 0.0, c]], dtype=np.float64)


def build_legacy_table(pa: PhaseArrays, reciprocal_radius: float,
                       scattering_params: Optional[str] = "lobato", device=None,
                       warn_missing: bool = True) -> PhaseTable:
    """Reflection table of the legacy DiffractionGenerator
    (diffsims/generators/diffraction_generator.py:233-242): the reciprocal
    lattice points of get_points_in_sphere (sim_utils.py:436-474) -- box
    floor(r / |a*|) etc., ascending (h, k, l), |g| < r, (000) included --
    with g in diffpy's standard frame (diffpy_standard_base) and |F|^2."""
    dev = _device(device)
    L = _lib.load()
    if not reciprocal_radius > 0:
        raise ValueError("reciprocal_radius must be positive")
    base = diffpy_standard_base(pa.abc, pa.angles)
    rec = np.linalg.inv(base)                       # columns a*, b*, c*
    metric = reciprocal_metric(rec)
    lengths = np.sqrt(np.sum(rec * rec, axis=0))    # |a*|, |b*|, |c*|
    H = [int(np.floor(reciprocal_radius / v)) for v in lengths]
    h3 = (ctypes.c_int * 3)(*H)
    m6 = (ctypes.c_double * 6)(*metric.tolist())
    stream = _lib.stream_handle(dev)
    work = torch.empty(max(1, int(L.dsx_hkl_workspace_ints(h3))), dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.check(L.dsx_hkl_points_in_sphere(h3, m6, float(reciprocal_radius), None, 0,
                                          _lib.ptr(count), _lib.ptr(work), stream),
               "dsx_hkl_points_in_sphere")
    N = int(count.item())
    hkl = torch.zeros((N, 4), dtype=torch.int32, device=dev)
    _lib.check(L.dsx_hkl_points_in_sphere(h3, m6, float(reciprocal_radius), _lib.ptr(hkl), N,
                                          _lib.ptr(count), _lib.ptr(work), stream),
               "dsx_hkl_points_in_sphere")
    g4, F2, F = _structure_factor(pa, hkl, rec, scattering_params, dev, warn_missing, stream)
    g_max = float(torch.sqrt(g4[:, 3].max()).item()) if N else 0.0
    return PhaseTable(hkl, g4, F2, F, N, False, g_max, rec, base, pa.name, None)


def structure_factor_table(pa: PhaseArrays, hkl: np.ndarray, rec: np.ndarray,
                           scattering_params: Optional[str] = "lobato", device=None,
                           warn_missing: bool = True) -> PhaseTable:
    """|F|^2 (and g) of given integer indices for a phase, g = rec hkl."""
    dev = _device(device)
    h = np.zeros((len(hkl), 4), dtype=np.int32)
    h[:, :3] = np.rint(np.asarray(hkl, dtype=np.float64)).astype(np.int32).reshape(-1, 3)
    t_hkl = torch.from_numpy(h).to(dev)
    g4, F2, F = _structure_factor(pa, t_hkl, rec, scattering_params, dev, warn_missing,
                                  _lib.stream_handle(dev))
    g_max = float(torch.sqrt(g4[:, 3].max()).item()) if len(h) else 0.0
    return PhaseTable(t_hkl, g4, F2, F, len(h), False, g_max, rec, None, pa.name, None)


def table_from_hkl(hkl: np.ndarray, abc, angles, F2: Optional[np.ndarray] = None,
                   device=None) -> PhaseTable:
    """Reflection table for given integer indices (g vectors on the GPU);
    |F|^2 = 1 unless given.  Used for a caller's own reflection set
    (get_intersecting_reflections)."""
    dev = _device(device)
    L = _lib.load()
    base, rec = lattice_setup(abc, angles)
    h = np.zeros((len(hkl), 4), dtype=np.int32)
    h[:, :3] = np.asarray(hkl, dtype=np.int32).reshape(-1, 3)
    t_hkl = torch.from_numpy(h).to(dev)
    N = h.shape[0]
    g4 = torch.empty((N, 4), dtype=torch.float32, device=dev)
    F2t = torch.empty(N, dtype=torch.float32, device=dev)
    rec9 = (ctypes.c_double * 9)(*rec.reshape(-1).tolist())
    _lib.check(L.dsx_structure_factor(_lib.ptr(t_hkl), N, rec9, None, None, None, None, 0, None, 0,
                                      _lib.SCATTER_NONE, _lib.ptr(g4), _lib.ptr(F2t), None,
                                      _lib.stream_handle(dev)), "dsx_structure_factor")
    if F2 is None:
        F2t.fill_(1.0)
    else:
        F2t.copy_(torch.as_tensor(np.asarray(F2, dtype=np.float32)))
    g_max = float(torch.sqrt(g4[:, 3].max()).item()) if N else 0.0
    return PhaseTable(t_hkl, g4, F2t, None, N, False, g_max, rec, base, "", build_lines(t_hkl, g4))


@dataclass
class SimSettings:
    """Physical settings of SimulationGenerator + calculate_diffraction2d."""
    accelerating_voltage: float = 200.0
    max_excitation_error: float = 1e-2
    shape_factor_width: Optional[float] = None
    shape_model: str = "lorentzian"
    minima_number: float = 5.0
    precession_angle: float = 0.0        # degrees
    approximate_precession: bool = True
    minimum_intensity: float = 1e-20

    def params(self, g_max: float, rec: np.ndarray) -> _lib.SimParams:
        lam = electron_wavelength(self.accelerating_voltage)
        if lam == 0:
            raise ValueError("infinite accelerating voltage is not supported by the simulation")
        width = self.shape_factor_width if self.shape_factor_width is not None else self.max_excitation_error
        phi = abs(self.precession_angle)
        if phi > 0 and self.shape_model != "binary":
            # the precession band test applies to every model; the shape factor
            # is the reference's approximate precession Lorentzian, or with
            # approximate_precession=False the full average of the model
            if self.approximate_precession:
                model = _lib.SF_LORENTZIAN_PRECESSION
            else:
                if self.shape_model not in SHAPE_MODEL_IDS:
                    raise NotImplementedError(f"shape factor model {self.shape_model!r} has no GPU kernel")
                model = _lib.SF_PRECESSION_AVG + SHAPE_MODEL_IDS[self.shape_model]
        else:
            if self.shape_model not in SHAPE_MODEL_IDS:
                raise NotImplementedError(f"shape factor model {self.shape_model!r} has no GPU kernel")
            model = SHAPE_MODEL_IDS[self.shape_model]
        p = _lib.SimParams()
        p.k = 1.0 / lam
        p.max_excitation_error = self.max_excitation_error
        p.shape_width = width
        p.shape_model = model
        p.minima_number = self.minima_number
        p.precession_angle = math.radians(phi)
        p.minimum_intensity = self.minimum_intensity
        p.g_max = g_max
        p.rec[:] = np.asarray(rec, dtype=np.float64).reshape(-1).tolist()
        return p


@dataclass
class SpotBatch:
    """Per-orientation spot lists on the device (see module docstring)."""
    table: PhaseTable
    idx: torch.Tensor
    spots: torch.Tensor
    count: torch.Tensor
    ncand: torch.Tensor
    maxI: torch.Tensor
    cap: int

    @property
    def n(self) -> int:
        return int(self.count.shape[0])


def rotation_matrices_to_device(matrices: np.ndarray, device=None) -> torch.Tensor:
    """(n, 3, 3) orix to_matrix() -> float64 [n, 9] rows of M^T on the device."""
    M = np.asarray(matrices, dtype=np.float64).reshape(-1, 3, 3)
    Rt = np.ascontiguousarray(np.transpose(M, (0, 2, 1)).reshape(-1, 9), dtype=np.float64)
    return torch.from_numpy(Rt).to(_device(device))


def expected_spots(tab: PhaseTable, settings: SimSettings, reciprocal_radius: float) -> int:
    """Mean number of intersecting reflections for a random orientation.

    The reflections fill a ball of radius R (the reciprocal radius); the Ewald
    shell is, near the origin, a slab |z| <= me + rho sin(phi) (rho the distance
    from the beam axis, phi the precession angle), which holds the fraction
    1.5 me / R + sin(phi) of the ball."""
    rr = max(reciprocal_radius, 1e-6)
    frac = 1.5 * settings.max_excitation_error / rr + math.sin(math.radians(abs(settings.precession_angle)))
    return int(math.ceil(tab.n * min(1.0, frac)))


def default_cap(tab: PhaseTable, settings: SimSettings, reciprocal_radius: float) -> int:
    """Spot-list capacity per orientation: generous for small lists (their
    relative spread is large), 25 % above the mean for large ones."""
    e = expected_spots(tab, settings, reciprocal_radius)
    cap = 2 * e + 128 if e < 4096 else int(1.25 * e) + 256
    return max(1, min(max(tab.n, 1), cap))


def simulate(tab: PhaseTable, Rt: torch.Tensor, settings: SimSettings, reciprocal_radius: float,
             cap: Optional[int] = None, stream=None) -> SpotBatch:
    """Spot lists of every orientation (rows of Rt) against one phase table."""
    L = _lib.load(require_gpu=True)
    dev = Rt.device
    n = int(Rt.shape[0])
    if Rt.dtype != torch.float64 or Rt.dim() != 2 or Rt.shape[1] != 9 or not Rt.is_contiguous():
        raise ValueError("Rt must be a contiguous float64 [n, 9] tensor")
    params = settings.params(tab.g_max, tab.rec)
    if cap is None:
        cap = default_cap(tab, settings, reciprocal_radius)
    stream = _lib.stream_handle(dev) if stream is None else stream
    for attempt in range(2):
        idx = torch.empty((n, cap), dtype=torch.int32, device=dev)
        spots = torch.empty((n, cap, 4), dtype=torch.float32, device=dev)
        count = torch.empty(n, dtype=torch.int32, device=dev)
        ncand = torch.empty(n, dtype=torch.int32, device=dev)
        maxI = torch.empty(n, dtype=torch.float32, device=dev)
        _lib.check(L.dsx_simulate(_lib.ptr(tab.hkl), _lib.ptr(tab.g4), _lib.ptr(tab.F2), tab.n,
                                  _lib.ptr(tab.lines), tab.n_lines, _lib.ptr(Rt), n,
                                  ctypes.byref(params), cap, _lib.ptr(idx), _lib.ptr(spots),
                                  _lib.ptr(count), _lib.ptr(ncand), _lib.ptr(maxI), stream),
                   "dsx_simulate")
        need = int(ncand.max().item()) if n else 0
        if need <= cap:
            break
        cap = need
    return SpotBatch(tab, idx, spots, count, ncand, maxI, cap)


def gaussian_taps(sigma: float):
    """Half of scipy.ndimage's normalised Gaussian kernel (truncate 4): taps[0..R]."""
    R = int(4.0 * float(sigma) + 0.5)
    x = np.arange(-R, R + 1)
    phi = np.exp(-0.5 / (float(sigma) ** 2) * x ** 2)
    phi = phi / phi.sum()
    return R, phi[R:]


def render_params(shape, sigma, calibration, direct_beam_position=None, in_plane_angle=0.0,
                  mirrored=False, normalize=True, fast=True, clip_threshold=1.0) -> _lib.RenderParams:
    H, W = int(shape[0]), int(shape[1])
    if direct_beam_position is None:
        direct_beam_position = (W // 2, H // 2) if fast else ((W - 1) / 2, (H - 1) / 2)
    p = _lib.RenderParams()
    p.height, p.width = H, W
    p.calibration = calibration
    p.center_x, p.center_y = float(direct_beam_position[0]), float(direct_beam_position[1])
    a = math.radians(in_plane_angle)
    p.cos_angle, p.sin_angle = (1.0, 0.0) if in_plane_angle == 0 else (math.cos(a), math.sin(a))
    p.mirrored = int(bool(mirrored))
    p.normalize = int(bool(normalize))
    p.sigma = float(sigma)
    p.clip_threshold = float(clip_threshold)
    if fast:
        R, taps = gaussian_taps(sigma)
        if R > _lib.MAX_TAP_RADIUS:
            raise NotImplementedError(f"Gaussian radius {R} (sigma={sigma}) exceeds {_lib.MAX_TAP_RADIUS}")
        p.radius = R
        for i, t in enumerate(taps):
            p.taps[i] = float(t)
    return p


_RENDER_PARAMS_CACHE: dict = {}


def _cached_render_params(shape, sigma, calibration, direct_beam_position, in_plane_angle,
                          mirrored, normalize, fast, clip_threshold) -> _lib.RenderParams:
    """render_params for render_spots, built once per distinct argument set
    (the Gaussian taps cost several microseconds per call; the structs stay
    private to this module and the library only reads them)."""
    key = (int(shape[0]), int(shape[1]), float(sigma), float(calibration),
           None if direct_beam_position is None else tuple(float(v) for v in direct_beam_position),
           float(in_plane_angle), bool(mirrored), bool(normalize), bool(fast),
           float(clip_threshold))
    p = _RENDER_PARAMS_CACHE.get(key)
    if p is None:
        p = render_params(shape, sigma, calibration, direct_beam_position, in_plane_angle,
                          mirrored, normalize, fast, clip_threshold)
        if len(_RENDER_PARAMS_CACHE) >= 256:
            _RENDER_PARAMS_CACHE.clear()
        _RENDER_PARAMS_CACHE[key] = p
    return p


def render_spots(spots: torch.Tensor, count: torch.Tensor, cap: int, shape, sigma, calibration,
                 direct_beam_position=None, in_plane_angle=0.0, mirrored=False, normalize=True,
                 fast=True, clip_threshold=1.0, out: Optional[torch.Tensor] = None,
                 stream=None, status: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Render n spot lists (float4 {x, y, z, I}, stride cap) to [n, H, W] float32.

    Without ``status`` the call synchronises and raises if a pattern had more
    distinct in-frame spots than the renderer holds; with a caller-provided
    int32 [n] ``status`` tensor it only launches (check it with
    :func:`check_render_status`).
    """
    L = _lib.load(require_gpu=True)
    dev = spots.device
    n = int(count.shape[0])
    H, W = int(shape[0]), int(shape[1])
    if out is None:
        out = torch.empty((n, H, W), dtype=torch.float32, device=dev)
    elif out.shape != (n, H, W) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 [n, H, W] tensor")
    sync_check = status is None
    if status is None:
        status = torch.zeros(n, dtype=torch.int32, device=dev)
    p = _cached_render_params(shape, sigma, calibration, direct_beam_position, in_plane_angle,
                              mirrored, normalize, fast, clip_threshold)
    stream = _lib.stream_handle(dev) if stream is None else stream
    if not fast and cap > int(L.dsx_render_max_spots()):
        # more spots than the render LDS holds: the adjoint kernel's forward
        # mode renders them chunk by chunk (same pixels as dsx_render_subpixel)
        work = _workspace(dev, int(L.dsx_render_adjoint_workspace_ints(int(cap), int(n), H, W, 0)),
                          stream)
        sim = _lib.SimParams()
        sim.k, sim.shape_width, sim.shape_model = 1.0, 1.0, _lib.SF_LINEAR
        _lib.check(L.dsx_render_adjoint(
            None, _lib.ptr(spots), _lib.ptr(count), cap, n, None, None, 0, None, ctypes.byref(sim),
            ctypes.byref(p), _lib.ADJ_FORWARD, None, 0, _lib.ptr(out), None, None, None, None,
            _lib.ptr(work), work.numel(), _lib.ptr(status), stream), "dsx_render_adjoint")
    elif fast and cap > int(L.dsx_render_max_spots()):
        # more spots than the splat renderer holds: whole pattern in LDS when
        # it fits, otherwise strip windows with spot bins in a workspace
        if H * W <= int(L.dsx_render_dense_max_pixels()):
            lst = torch.arange(n, dtype=torch.int32, device=dev)
            off = torch.arange(n, dtype=torch.int64, device=dev) * cap
            capv = torch.full((n,), cap, dtype=torch.int32, device=dev)
            method = (_lib.RENDER_DENSE_INPLACE if L.dsx_render_dense_inplace_fits(H, W, p.radius)
                      else _lib.RENDER_DENSE)
            _lib.check(L.dsx_render_fast_ragged(
                _lib.ptr(spots), _lib.ptr(off), _lib.ptr(capv), _lib.ptr(count), _lib.ptr(lst), n,
                ctypes.byref(p), method, 0, _lib.ptr(out), _lib.ptr(status), stream),
                "dsx_render_fast_ragged")
        else:
            work = _workspace(dev, int(L.dsx_render_fast_large_workspace_ints(int(cap))), stream)
            _lib.check(L.dsx_render_fast_large(
                _lib.ptr(spots), _lib.ptr(count), cap, n, ctypes.byref(p), _lib.ptr(work),
                work.numel(), _lib.ptr(out), _lib.ptr(status), stream), "dsx_render_fast_large")
    elif fast and fast_renderer(n, H, W) == "dsx_render_fast":
        _lib.check(L.dsx_render_fast(_lib.ptr(spots), _lib.ptr(count), cap, n, ctypes.byref(p),
                                     _lib.ptr(out), _lib.ptr(status), stream), "dsx_render_fast")
    elif fast:
        # plan kernel + one-shot band writer (dsx_render_fast's output, written
        # at the store rate of a one-shot grid)
        work = _workspace(dev, int(L.dsx_render_fast_bands_workspace_ints(n, H, W)), stream)
        _lib.check(L.dsx_render_fast_bands(
            _lib.ptr(spots), _lib.ptr(count), cap, n, ctypes.byref(p), _lib.ptr(work), work.numel(),
            _lib.ptr(out), _lib.ptr(status), stream), "dsx_render_fast_bands")
    else:
        _lib.check(L.dsx_render_subpixel(_lib.ptr(spots), _lib.ptr(count), cap, n, ctypes.byref(p),
                                         _lib.ptr(out), _lib.ptr(status), stream), "dsx_render")
    if sync_check:
        check_render_status(status)
    return out


def _workspace(dev, ints: int, stream) -> torch.Tensor:
    """Device scratch of ``ints`` int32 words for one launch on ``stream``.

    Taken from torch's caching allocator per call (no hipMalloc once its pool
    holds a block of the size), so launches on different streams never share
    scratch.  When the launch stream is not torch's current stream the block
    is recorded on it, so the allocator hands it out again only after the
    kernel is done."""
    w = torch.empty(max(1, int(ints)), dtype=torch.int32, device=dev)
    if int(stream) != torch.cuda.current_stream(dev).cuda_stream:
        w.record_stream(torch.cuda.ExternalStream(int(stream), device=dev))
    return w


def check_render_status(status: torch.Tensor) -> None:
    """Raise if any rendered pattern overflowed the renderer's spot capacity."""
    if status.numel() and int(status.max().item()) != 0:
        bad = int(torch.nonzero(status)[0].item())
        raise RuntimeError(f"pattern {bad} has more than {_lib.load().dsx_render_max_spots()} "
                           "distinct in-frame spots; the renderer cannot hold it")


def simulate_into(tab: PhaseTable, Rt: torch.Tensor, settings: SimSettings, out: SpotBatch,
                  stream=None) -> None:
    """Launch-only variant of :func:`simulate` into preallocated buffers (no
    synchronisation; the caller checks ``out.ncand.max() <= out.cap``)."""
    L = _lib.load(require_gpu=True)
    params = settings.params(tab.g_max, tab.rec)
    stream = _lib.stream_handle(Rt.device) if stream is None else stream
    _lib.check(L.dsx_simulate(_lib.ptr(tab.hkl), _lib.ptr(tab.g4), _lib.ptr(tab.F2), tab.n,
                              _lib.ptr(tab.lines), tab.n_lines, _lib.ptr(Rt), int(Rt.shape[0]), ctypes.byref(params), out.cap,
                              _lib.ptr(out.idx), _lib.ptr(out.spots), _lib.ptr(out.count),
                              _lib.ptr(out.ncand), _lib.ptr(out.maxI), stream), "dsx_simulate")


def empty_batch(tab: PhaseTable, n: int, cap: int, device) -> SpotBatch:
    dev = torch.device(device)
    return SpotBatch(tab, torch.empty((n, cap), dtype=torch.int32, device=dev),
                     torch.empty((n, cap, 4), dtype=torch.float32, device=dev),
                     torch.empty(n, dtype=torch.int32, device=dev),
                     torch.empty(n, dtype=torch.int32, device=dev),
                     torch.empty(n, dtype=torch.float32, device=dev), cap)


def polar_templates(spots: torch.Tensor, count: torch.Tensor, cap: int, radial_axes=None,
                    azimuthal_axes=None, width: Optional[int] = None, stream=None):
    """(r, theta, intensity) templates of every spot list on the device:
    Simulation2D.polar_flatten_simulations (simulation2d.py:313-355) without a
    host round trip.  Returns torch tensors of shape [n, width] (width = the
    largest spot count unless given): float32 r and theta, or int32 nearest
    indices when both axes are given (entries past the last index dropped)."""
    L = _lib.load(require_gpu=True)
    dev = spots.device
    n = int(count.shape[0])
    if width is None:
        width = int(count.max().item()) if n else 0
    axes = radial_axes is not None and azimuthal_axes is not None
    I = torch.zeros((n, width), dtype=torch.float32, device=dev)
    n_out = torch.zeros(n, dtype=torch.int32, device=dev)
    stream = _lib.stream_handle(dev) if stream is None else stream
    if axes:
        ra = torch.as_tensor(np.asarray(radial_axes, dtype=np.float64)).to(dev)
        az = torch.as_tensor(np.asarray(azimuthal_axes, dtype=np.float64)).to(dev)
        r = torch.zeros((n, width), dtype=torch.int32, device=dev)
        t = torch.zeros((n, width), dtype=torch.int32, device=dev)
        _lib.check(L.dsx_polar_templates(_lib.ptr(spots), _lib.ptr(count), cap, n, width,
                                         _lib.ptr(ra), int(ra.numel()), _lib.ptr(az),
                                         int(az.numel()), None, None, _lib.ptr(r), _lib.ptr(t),
                                         _lib.ptr(I), _lib.ptr(n_out), stream), "dsx_polar_templates")
    else:
        r = torch.zeros((n, width), dtype=torch.float32, device=dev)
        t = torch.zeros((n, width), dtype=torch.float32, device=dev)
        _lib.check(L.dsx_polar_templates(_lib.ptr(spots), _lib.ptr(count), cap, n, width, None, 0,
                                         None, 0, _lib.ptr(r), _lib.ptr(t), None, None, _lib.ptr(I),
                                         _lib.ptr(n_out), stream), "dsx_polar_templates")
    return r, t, I


def render(batch: SpotBatch, shape, sigma, calibration, **kw) -> torch.Tensor:
    return render_spots(batch.spots, batch.count, batch.cap, shape, sigma, calibration, **kw)
