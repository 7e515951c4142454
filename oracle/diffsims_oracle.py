"""FP64 CPU oracle for the kinematical 2-D diffraction-pattern path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py`` and by nothing else.  The product (``diffsims_amd``) never imports,
calls or falls back to it.

It is a plain NumPy/SciPy restatement of the algorithm of diffsims 0.7.0
(pure Python, GPL-3.0; its code is not copied here).  Every function cites the
reference ``file:line`` it follows.  The reference itself cannot be imported in
this environment (orix, diffpy.structure, numba and transforms3d are absent), so
the restatement is pinned by the reference's own golden data and known-answer
tests, by closed forms and by brute force; the tests in ``tests/test_oracle*.py``
hold those checks.

Parity pins (what fixes each function):
  electron_wavelength        λ KATs (tests/utils/test_sim_utils.py:49-60)
  lattice_basis / rnorm      ferrite g-spacings (test_reciprocal_lattice_vector.py:157-187),
                             golden old_simulation.npy (hexagonal frame); oblique
                             (monoclinic / triclinic) frame: the cell recovered from
                             the basis rows, x || a, b in the xy-plane, c* || +z,
                             right-handed, base @ rec = I (the conditions stated at
                             test_simulation_generator.py:311 fix the frame), with
                             planted mistakes in a31/a32; rnorm = the textbook
                             triclinic d-spacing from the inverse direct metric
  highest_hkl / hkl_box      ferrite 18/92/750, SiC box, Al 798/256 + order KATs
  bunge_matrix / quat*       golden old_simulation.npy; closed-form orthogonality;
                             quaternion<->matrix round trip
  structure factor / f(g)    Ni 43.0979 KAT, unknown-element KAT, brute-force F,
                             Friedel |F(g)| = |F(-g)|, FCC/diamond extinctions
  intersecting / shape       Si 70-spot and precession 250-spot KATs, linear KAT,
                             closed forms, scipy.integrate.quad for the full
                             precession integral
  render_fast                golden old_simulation.npy (uses scipy.ndimage, as the
                             reference does), total-intensity KAT
  render_subpixel            detector KATs (test_detector_functions.py:160-306),
                             fast-vs-slow equality KAT (test_simulations2d.py:133-180),
                             normalisation at sigma = 1.7, 2.5, 3.2 (pixel sum = I,
                             peak = I / (2 pi sigma^2), the clip box edge)
  beam_directions_grid       mesh sizes and contained directions
                             (test_sphere_mesh_generators.py:48-78), grid sizes per
                             crystal system (test_rotation_list_generator.py:80-94),
                             vector -> Euler KAT (test_sphere_mesh_generators.py:124-143)
  icosahedral mesh           642 unique vertices summing to 0 at 10 deg
                             (test_sphere_mesh_generators.py:115-121); vertex count
                             10 n^2 + 2; corners and edge points on the flat
                             icosahedron; nearest-neighbour angle of the bare
                             icosahedron = arctan(2) (63.435 deg); level n+1 finer
  random mesh                10313 points at 1 deg, seeds 7/7/8
                             (test_sphere_mesh_generators.py:33-45)
  cubochoric SO(3) grids     cube -> ball map: Jacobian 1, surface onto the ball,
                             faces / edges / corners on their directions,
                             continuous and odd (planted constant error fails);
                             the homochoric angle's equation; Euler <-> quaternion
                             round trip and the passive Bunge matrix; group
                             closure, the zone tiling the grid, Mackenzie's
                             maximal disorientations; beam-direction grid KAT
                             (test_rotation_list_generator.py:47-54) and closed form
Parity unpinned (no reference test and no closed form fixes it):
  * the order of the Cartesian sort for oblique cells when rounding could tie
    (exact arithmetic gives descending lexicographic (h,k,l); see hkl_box);
  * orix's Euler/quaternion branch handling beyond the golden case;
  * the d == d_min equality case of the hkl box;
  * the custom-callable shape-model count (test_custom_shape_func expects 52);
  * orix's resolution -> cubochoric steps rule, grid loop order, zone boundary
    tolerance and space group -> proper group table (cubochoric_steps,
    proper_group_of_space_group).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

# --------------------------------------------------------------------------
# a1: wavelength  (diffsims/utils/sim_utils.py:57-79)
# --------------------------------------------------------------------------


def electron_wavelength(kv) -> float:
    """Relativistic electron wavelength in Angstrom for a voltage in kV.

    lambda = h / sqrt(2 m_e e E (1 + e E / (2 m_e c^2))) * 1e10, E in volts,
    CODATA constants from scipy.constants (sim_utils.py:57-79).
    """
    from scipy.constants import c, e, h, m_e

    if kv in (np.inf, "inf"):
        return 0.0
    E = kv * 1e3
    return h / math.sqrt(2 * m_e * e * E * (1 + (e / (2 * m_e * c * c)) * E)) * 1e10


# --------------------------------------------------------------------------
# Lattice in the orix frame (x || a, y in the ab-plane, z || c*)
# --------------------------------------------------------------------------


def cosd(x: float) -> float:
    """cos of an angle in degrees, exact where the cosine is 0, +-1/2 or +-1."""
    r = math.fmod(float(x), 360.0)
    if r < 0:
        r += 360.0
    exact = {0.0: 1.0, 60.0: 0.5, 90.0: 0.0, 120.0: -0.5, 180.0: -1.0,
             240.0: -0.5, 270.0: 0.0, 300.0: 0.5}
    if r in exact:
        return exact[r]
    return math.cos(math.radians(x))


def sind(x: float) -> float:
    """sin of an angle in degrees, exact where the sine is 0, +-1/2 or +-1."""
    return cosd(90.0 - float(x))


def lattice_basis(a, b, c, alpha, beta, gamma):
    """Direct basis rows (a, b, c) and reciprocal basis in the orix frame.

    The reference aligns every structure with x || a and z || c* (stated at
    diffsims/tests/generators/test_simulation_generator.py:311; undone for atom
    positions at diffsims/utils/sim_utils.py:290-291).  The basis is lower
    triangular in that frame; its inverse is written in closed form so that the
    structurally-zero entries are exact zeros.

    Returns (base, rec) with ``base`` rows = direct vectors in Cartesian
    Angstrom and ``rec`` = inv(base), whose *columns* are a*, b*, c*; the
    Cartesian reciprocal vector of integer indices hkl is ``hkl @ rec.T``.
    """
    ca, cb, cg = cosd(alpha), cosd(beta), cosd(gamma)
    sg = sind(gamma)
    a21, a22 = b * cg, b * sg
    a31 = c * cb
    a32 = c * (ca - cb * cg) / sg
    a33 = c * math.sqrt(1.0 - cb * cb - ((ca - cb * cg) / sg) ** 2)
    base = np.array([[a, 0.0, 0.0], [a21, a22, 0.0], [a31, a32, a33]])
    rec = np.zeros((3, 3))
    rec[0, 0] = 1.0 / a
    rec[1, 1] = 1.0 / a22
    rec[2, 2] = 1.0 / a33
    rec[1, 0] = -a21 / (a * a22)
    rec[2, 1] = -a32 / (a22 * a33)
    rec[2, 0] = (a21 * a32 - a22 * a31) / (a * a22 * a33)
    return base, rec


def reciprocal_metric(rec: np.ndarray) -> np.ndarray:
    """G* = (a*, b*, c*) Gram matrix; 1/d^2 = hkl G* hkl."""
    return rec.T @ rec


def rnorm(hkl: np.ndarray, rec: np.ndarray) -> np.ndarray:
    """|g| of integer/float indices (diffpy Lattice.rnorm, used at
    reciprocal_lattice_vector.py:440 and :1135).

    Evaluated as the explicit quadratic form
    q = h^2 G00 + k^2 G11 + l^2 G22 + 2 (h k G01 + h l G02 + k l G12),
    in this summation order, so the GPU enumeration kernel can reproduce it
    bit for bit.
    """
    G = reciprocal_metric(rec)
    h = np.asarray(hkl, dtype=np.float64)
    h0, h1, h2 = h[..., 0], h[..., 1], h[..., 2]
    q = h0 * h0 * G[0, 0]
    q = q + h1 * h1 * G[1, 1]
    q = q + h2 * h2 * G[2, 2]
    cross = h0 * h1 * G[0, 1]
    cross = cross + h0 * h2 * G[0, 2]
    cross = cross + h1 * h2 * G[1, 2]
    q = q + 2.0 * cross
    return np.sqrt(q)


# --------------------------------------------------------------------------
# a2: hkl box rule  (reciprocal_lattice_vector.py:1077-1142, get_hkl.py:49-57)
# --------------------------------------------------------------------------


def highest_hkl(rec: np.ndarray, d_min: float) -> np.ndarray:
    """Per-axis box bound: the largest n with d(n e_i) > d_min.

    Follows the bound loop of diffsims/crystallography/get_hkl.py:49-57 (which
    returns one more than this; orix's variant used at
    reciprocal_lattice_vector.py:1131 subtracts it again, as
    diffsims/tests/crystallography/test_get_hkl.py:52-64 shows).
    """
    out = np.zeros(3, dtype=np.int64)
    for i in range(3):
        e = np.zeros(3)
        d = d_min + 1.0
        while d > d_min:
            e[i] += 1
            d = 1.0 / rnorm(e, rec)
        out[i] = int(e[i]) - 1
    return out


def hkl_box(rec: np.ndarray, d_min: float, include_zero: bool = False) -> np.ndarray:
    """Integer hkl of ``from_min_dspacing`` in the reference's order.

    Box [-H,H]x[-K,K]x[-L,L] (H from :func:`highest_hkl`), keep
    1/rnorm >= d_min, drop (000), order as orix ``Vector3d.unique`` leaves it
    (descending lexicographic on Cartesian x, y, z;
    reciprocal_lattice_vector.py:1131-1141, :1509-1563).  In the orix frame
    g_x = h/a, g_y depends on (h, k) increasing in k, g_z increasing in l, so
    that order is descending lexicographic (h, k, l) in exact arithmetic; this
    function uses the integer order.  (000) is appended last when
    ``include_zero`` (:1139-1141).
    """
    H = highest_hkl(rec, d_min)
    hs = np.arange(H[0], -H[0] - 1, -1)
    ks = np.arange(H[1], -H[1] - 1, -1)
    ls = np.arange(H[2], -H[2] - 1, -1)
    hkl = np.stack(np.meshgrid(hs, ks, ls, indexing="ij"), axis=-1).reshape(-1, 3)
    with np.errstate(divide="ignore"):
        d = 1.0 / rnorm(hkl, rec)
    keep = (d >= d_min) & np.any(hkl != 0, axis=1)
    hkl = hkl[keep]
    if include_zero:
        hkl = np.vstack([hkl, np.zeros((1, 3), dtype=hkl.dtype)])
    return hkl


def g_cartesian(hkl: np.ndarray, rec: np.ndarray) -> np.ndarray:
    """Cartesian reciprocal vectors (1/Angstrom) of indices hkl."""
    return np.asarray(hkl, dtype=np.float64) @ rec.T


# --------------------------------------------------------------------------
# a3: rotations  (_diffracting_vector.py:127-161; orix Rotation conventions)
# --------------------------------------------------------------------------


def bunge_matrix(euler: np.ndarray) -> np.ndarray:
    """Passive Bunge (ZXZ) matrices M(phi1, Phi, phi2), shape (..., 3, 3).

    The reference rotates vectors with ``~rotation * miller``
    (_diffracting_vector.py:160), i.e. g_lab = M^T g_crystal with M =
    ``Rotation.to_matrix()``; pinned by the golden pattern
    (test_simulation_generator.py:283-343).
    """
    e = np.asarray(euler, dtype=np.float64)
    p1, P, p2 = e[..., 0], e[..., 1], e[..., 2]
    c1, s1 = np.cos(p1), np.sin(p1)
    c, s = np.cos(P), np.sin(P)
    c2, s2 = np.cos(p2), np.sin(p2)
    M = np.empty(e.shape[:-1] + (3, 3))
    M[..., 0, 0] = c1 * c2 - s1 * s2 * c
    M[..., 0, 1] = s1 * c2 + c1 * s2 * c
    M[..., 0, 2] = s2 * s
    M[..., 1, 0] = -c1 * s2 - s1 * c2 * c
    M[..., 1, 1] = -s1 * s2 + c1 * c2 * c
    M[..., 1, 2] = c2 * s
    M[..., 2, 0] = s1 * s
    M[..., 2, 1] = -c1 * s
    M[..., 2, 2] = c
    return M


def quat_from_bunge(euler: np.ndarray) -> np.ndarray:
    """Unit quaternions (w, x, y, z) of Bunge angles, Rowenhorst et al. (2015)
    convention with P = +1 and a non-negative scalar part (orix
    ``Rotation.from_euler``)."""
    e = np.asarray(euler, dtype=np.float64)
    p1, P, p2 = e[..., 0], e[..., 1], e[..., 2]
    sigma = 0.5 * (p1 + p2)
    delta = 0.5 * (p1 - p2)
    c, s = np.cos(P / 2), np.sin(P / 2)
    q = np.stack([c * np.cos(sigma), -s * np.cos(delta), -s * np.sin(delta),
                  -c * np.sin(sigma)], axis=-1)
    q = np.where(q[..., :1] < 0, -q, q)
    return q


def matrix_from_quat(q: np.ndarray) -> np.ndarray:
    """Rotation matrix of a unit quaternion, R v = q v q^-1 (what orix
    ``Rotation.to_matrix`` returns); equals :func:`bunge_matrix` for quaternions
    from :func:`quat_from_bunge`."""
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    qq = w * w - (x * x + y * y + z * z)
    M = np.empty(q.shape[:-1] + (3, 3))
    M[..., 0, 0] = qq + 2 * x * x
    M[..., 0, 1] = 2 * (x * y - w * z)
    M[..., 0, 2] = 2 * (x * z + w * y)
    M[..., 1, 0] = 2 * (y * x + w * z)
    M[..., 1, 1] = qq + 2 * y * y
    M[..., 1, 2] = 2 * (y * z - w * x)
    M[..., 2, 0] = 2 * (z * x - w * y)
    M[..., 2, 1] = 2 * (z * y + w * x)
    M[..., 2, 2] = qq + 2 * z * z
    return M


def random_euler(n: int, seed: int) -> np.ndarray:
    """Seeded Haar-uniform Bunge angles: phi1, phi2 ~ U[0, 2pi), Phi = arccos U[-1, 1]."""
    rng = np.random.default_rng(seed)
    p1 = rng.uniform(0, 2 * np.pi, n)
    P = np.arccos(rng.uniform(-1, 1, n))
    p2 = rng.uniform(0, 2 * np.pi, n)
    return np.stack([p1, P, p2], axis=1)


# --------------------------------------------------------------------------
# a5: shape-factor models  (diffsims/utils/shape_factor_models.py)
# --------------------------------------------------------------------------


def sf_binary(s, w):
    """shape_factor_models.py:33-49: unit factor."""
    return np.ones_like(np.asarray(s, dtype=np.float64))


def sf_linear(s, w):
    """shape_factor_models.py:52-73: max(1 - |s|/w, 0)."""
    sf = 1 - np.abs(s) / w
    return np.maximum(sf, 0.0)


def sf_sinc(s, w, minima_number=5):
    """shape_factor_models.py:76-101: |sin(x)/x|, x = pi*n*s/w, and 0 at s = 0."""
    s = np.asarray(s, dtype=np.float64)
    fac = np.pi * minima_number / w
    num = np.sin(fac * s)
    den = fac * s
    out = np.zeros_like(num)
    np.divide(num, den, out=out, where=den != 0)
    return np.abs(out)


def sf_sin2c(s, w, minima_number=5):
    """shape_factor_models.py:104-123: sinc^2 (0 at s = 0)."""
    return sf_sinc(s, w, minima_number) ** 2


def sf_atanc(s, w, minima_number=5):
    """shape_factor_models.py:126-151: atan(x)/x with x = pi*n*s/|w|, 1 at s = 0."""
    s = np.asarray(s, dtype=np.float64)
    fac = np.pi * minima_number / np.abs(w)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.arctan(fac * s) / (fac * s)
    return np.nan_to_num(r, nan=1.0)


def sf_lorentzian(s, w):
    """shape_factor_models.py:154-180: sigma/(pi (sigma^2 s^2 + 1)) * w, sigma = pi/w."""
    sigma = np.pi / w
    return sigma / (np.pi * (sigma ** 2 * np.asarray(s) ** 2 + 1)) * w


def sf_lorentzian_precession(s, w, r, phi):
    """shape_factor_models.py:183-219 (Palatinus et al. 2019, eq. 10)."""
    sigma = np.pi / w
    u = sigma ** 2 * (r ** 2 * phi ** 2 - s ** 2) + 1
    z = np.sqrt(u ** 2 + 4 * sigma ** 2 * s ** 2)
    return (sigma / np.pi) * np.sqrt(2 * (u + z) / z ** 2)


def sf_precession_integral(s, r, phi, model: Callable, w, **kw):
    """shape_factor_models.py:222-268: (1/2pi) * integral_0^2pi model(s + r phi cos t) dt,
    with scipy.integrate.quad per spot as in the reference.

    The reference calls quad with its defaults (50 subintervals); across the
    kinks of |sinc| that leaves up to ~1e-3 relative error (measured against
    the converged integral).  The restatement integrates the same definition
    to convergence (2000 subintervals, epsrel 1e-12), so it is the exact
    average the reference approximates."""
    import warnings

    from scipy.integrate import IntegrationWarning, quad

    s = np.asarray(s, dtype=np.float64)
    out = np.zeros(s.shape)
    for i, (si, ri) in enumerate(zip(s, np.asarray(r, dtype=np.float64))):
        def integrand(t, si=si, ri=ri):
            return model(si + ri * phi * np.cos(t), w, **kw)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", IntegrationWarning)
            val = quad(integrand, 0, 2 * np.pi, limit=2000, epsabs=1e-14, epsrel=1e-12)[0]
        out[i] = (1 / (2 * np.pi)) * val
    return out


SHAPE_MODELS = {
    "linear": sf_linear,
    "atanc": sf_atanc,
    "sinc": sf_sinc,
    "sin2c": sf_sin2c,
    "lorentzian": sf_lorentzian,
}


# --------------------------------------------------------------------------
# a6-a9: scattering factors and structure factor  (sim_utils.py:165-354)
# --------------------------------------------------------------------------


def load_table(name: str) -> dict:
    """Element -> (5, 2) coefficient array from the package's JSON data file."""
    import json
    import os

    path = os.path.join(os.path.dirname(__file__), "..", "diffsims_amd", "data",
                        f"scattering_{name}.json")
    d = json.load(open(path))
    return {el: np.asarray(c, dtype=np.float64).reshape(5, 2)
            for el, c in zip(d["elements"], d["coeffs"])}


def element_of(symbol: str) -> str:
    """Alphabetic head of a CIF atom-type symbol
    (structure_factor/atomic_scattering_parameters.py:145-164)."""
    i = 0
    while i < len(symbol) and symbol[i].isalpha() and symbol[i].isascii():
        i += 1
    return symbol[:i]


def atom_coeffs(symbols: Sequence[str], table: Optional[str]):
    """(n, 5, 2) coefficients, zeros (and a miss flag) for unknown elements
    (sim_utils.py:165-224)."""
    coeffs = np.zeros((len(symbols), 5, 2))
    missing = []
    if table is None:
        return coeffs, missing
    t = load_table(table)
    for i, s in enumerate(symbols):
        el = element_of(s)
        if el in t:
            coeffs[i] = t[el]
        else:
            missing.append(el)
    return coeffs, missing


def atomic_f(g2: np.ndarray, coeffs: np.ndarray, table: Optional[str]) -> np.ndarray:
    """f_j(g^2), shape (M, n) (sim_utils.py:227-253, :281-286).

    lobato: sum_i a_i (2 + b_i g^2) / (1 + b_i g^2)^2; xtables: sum_i a_i
    exp(-b_i g^2 / 4); None: 1.
    """
    g2 = np.asarray(g2, dtype=np.float64)
    if table is None:
        return np.ones((g2.shape[0], coeffs.shape[0]))
    bg = g2[:, None, None] * coeffs[None, :, :, 1]
    if table == "lobato":
        f = (2 + bg) * (1 / np.square(1 + bg))
    elif table == "xtables":
        f = np.exp(-0.25 * bg)
    else:
        raise NotImplementedError(table)
    return np.sum(coeffs[None, :, :, 0] * f, axis=-1)


def _exact_phase_expi(t: np.ndarray) -> np.ndarray:
    """exp(2 pi i t) with t reduced to its fractional part first; exact
    for t in (1/4) Z (the kinematic-extinction cases)."""
    fr = t - np.round(t)                     # in [-1/2, 1/2]
    c = np.cos(2 * np.pi * fr)
    s = np.sin(2 * np.pi * fr)
    q = fr * 4
    on = q == np.round(q)
    qi = np.round(q).astype(np.int64) % 4
    c = np.where(on, np.choose(qi, [1.0, 0.0, -1.0, 0.0]), c)
    s = np.where(on, np.choose(qi, [0.0, 1.0, 0.0, -1.0]), s)
    return c + 1j * s


def structure_factor(hkl, g, xyz, occ, B, coeffs, table, phase_mode="reference"):
    """Complex F(hkl) = sum_j f_j occ_j exp(2 pi i hkl.x_j - g^2 B_j / 4)
    (sim_utils.py:256-304).

    ``g`` is |g| in 1/Angstrom; ``B`` is per atom (the reference looks B up per
    element from ``debye_waller_factors`` with default 0, sim_utils.py:220, which
    is the special case of a per-atom array that is constant per element).

    phase_mode "reference" evaluates exp(2 pi i x) as the reference does (so
    kinematically forbidden reflections carry ~1e-16 round-off); "exact"
    reduces hkl.x to its fractional part first, so forbidden reflections are
    exactly zero (the GPU kernel's behaviour).
    """
    hkl = np.asarray(hkl, dtype=np.float64)
    g2_all = np.asarray(g, dtype=np.float64) ** 2
    xyz = np.asarray(xyz, dtype=np.float64)
    occ = np.asarray(occ, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    out = np.empty(hkl.shape[0], dtype=np.complex128)
    # rows in blocks so the (rows, atoms) temporaries stay small for big cells
    step = max(1, (1 << 22) // max(1, xyz.shape[0]))
    for r0 in range(0, hkl.shape[0], step):
        g2 = g2_all[r0:r0 + step]
        f = atomic_f(g2, coeffs, table)                  # (M, n)
        t = hkl[r0:r0 + step] @ xyz.T                    # (M, n) turns
        dw = np.exp(-0.25 * np.outer(g2, B))
        if phase_mode == "reference":
            ph = np.exp(2j * np.pi * t)
        else:
            ph = _exact_phase_expi(t)
        out[r0:r0 + step] = np.sum(f * occ * dw * ph, axis=-1)
    return out


# --------------------------------------------------------------------------
# a4, a10: Ewald intersection and the relative cut
#   (simulation_generator.py:198-241, :319-412)
# --------------------------------------------------------------------------


@dataclass
class PhaseSpec:
    """Arrays describing a phase: lattice parameters and per-atom data."""
    abc: Sequence[float]
    angles: Sequence[float]
    symbols: Sequence[str]
    xyz: np.ndarray
    occ: np.ndarray
    B: np.ndarray

    def basis(self):
        return lattice_basis(*self.abc, *self.angles)


@dataclass
class SimParams:
    kv: float = 200.0
    reciprocal_radius: float = 1.0
    max_excitation_error: float = 1e-2
    shape_factor_width: Optional[float] = None
    shape_model: object = "lorentzian"       # name or callable
    shape_kwargs: dict = field(default_factory=dict)
    precession_angle: float = 0.0             # degrees
    approximate_precession: bool = True
    minimum_intensity: float = 1e-20
    with_direct_beam: bool = True
    scattering_params: Optional[str] = "lobato"


@dataclass
class PhaseTables:
    hkl: np.ndarray          # (N, 3) int, including the duplicated (000) rows
    g: np.ndarray            # (N, 3) Cartesian crystal frame
    gnorm: np.ndarray        # (N,)
    F2: np.ndarray           # (N,) |F|^2 (exact-phase mode)
    n_box: int               # rows before the appended direct-beam copies


def phase_tables(ph: PhaseSpec, p: SimParams, phase_mode="exact") -> PhaseTables:
    """Per-phase hkl set, g vectors and |F|^2 (what the GPU computes once)."""
    base, rec = ph.basis()
    hkl = hkl_box(rec, 1.0 / p.reciprocal_radius, include_zero=p.with_direct_beam)
    n_box = hkl.shape[0]
    if p.with_direct_beam:
        # second (000) row: simulation_generator.py:351-353
        hkl = np.vstack([hkl, np.zeros((1, 3), dtype=hkl.dtype)])
    g = g_cartesian(hkl, rec)
    gn = rnorm(hkl, rec)
    coeffs, _ = atom_coeffs(ph.symbols, p.scattering_params)
    F = structure_factor(hkl, gn, ph.xyz, ph.occ, ph.B, coeffs, p.scattering_params,
                         phase_mode)
    return PhaseTables(hkl, g, gn, (F * F.conjugate()).real, n_box)


def shape_factor(s, r, p: SimParams):
    """Shape factor dispatch (simulation_generator.py:387-411)."""
    w = p.shape_factor_width if p.shape_factor_width is not None else p.max_excitation_error
    if p.precession_angle == 0:
        model = p.shape_model if callable(p.shape_model) else SHAPE_MODELS[p.shape_model]
        return model(s, w, **p.shape_kwargs)
    phi = np.deg2rad(abs(p.precession_angle))
    if p.approximate_precession:
        return sf_lorentzian_precession(s, w, r, phi)
    model = p.shape_model if callable(p.shape_model) else SHAPE_MODELS[p.shape_model]
    return sf_precession_integral(s, r, phi, model, w, **p.shape_kwargs)


def intersect(g_lab: np.ndarray, k: float, p: SimParams):
    """Excitation error, in-plane radius and the intersection mask
    (simulation_generator.py:355-375)."""
    r = np.sqrt(np.sum(np.square(g_lab[:, :2]), axis=1))
    z = g_lab[:, 2]
    z_sphere = -np.sqrt(k ** 2 - r ** 2) + k
    s = z_sphere - z
    me = p.max_excitation_error
    if p.precession_angle == 0:
        mask = np.abs(s) < me
    else:
        phi = np.deg2rad(abs(p.precession_angle))
        Pz, Pt = k * np.cos(phi), k * np.sin(phi)
        with np.errstate(invalid="ignore"):
            up = Pz - np.sqrt(k ** 2 - (r + Pt) ** 2)
            do = Pz - np.sqrt(k ** 2 - (r - Pt) ** 2)
        mask = (z - me <= up) & (z + me >= do)
    return s, r, mask


@dataclass
class Spots:
    idx: np.ndarray          # indices into the phase hkl table
    xyz: np.ndarray          # (S, 3) lab-frame 1/Angstrom
    intensity: np.ndarray    # (S,)
    n_intersect: int         # spots before the relative cut
    max_intensity: float


def simulate_one(tab: PhaseTables, M: np.ndarray, k: float, p: SimParams) -> Spots:
    """One orientation of the reference's hot loop with |F|^2 from the table
    (simulation_generator.py:211-241).  An empty intersection returns no spots
    (the reference raises ValueError from np.max at :237)."""
    g_lab = tab.g @ M                          # (M^T g)^T = g^T M
    s, r, mask = intersect(g_lab, k, p)
    idx = np.nonzero(mask)[0]
    if idx.size == 0:
        return Spots(idx, np.zeros((0, 3)), np.zeros(0), 0, 0.0)
    sf = shape_factor(s[idx], r[idx], p)
    I = sf * tab.F2[idx]
    mx = float(np.max(I))
    keep = I > mx * p.minimum_intensity
    if mx == 0.0:
        # Every intersecting reflection is kinematically forbidden: the
        # reference keeps them all (their FP64 round-off |F|^2 passes its
        # relative cut); this build keeps them all with intensity exactly 0.
        keep = np.ones_like(keep)
    return Spots(idx[keep], g_lab[idx[keep]], I[keep], int(idx.size), mx)


def simulate_one_reference(ph: PhaseSpec, hkl_box_rows: np.ndarray, rec: np.ndarray,
                           M: np.ndarray, k: float, p: SimParams) -> Spots:
    """The reference's per-orientation loop body with F recomputed for the
    intersecting reflections (sim_utils.py:307-354 called at
    simulation_generator.py:227-234), reference-mode phases.  Used for the KAT
    counts and as the CPU baseline."""
    hkl = hkl_box_rows
    if p.with_direct_beam:
        hkl = np.vstack([hkl, np.zeros((1, 3), dtype=hkl.dtype)])
    g_lab = g_cartesian(hkl, rec) @ M
    s, r, mask = intersect(g_lab, k, p)
    idx = np.nonzero(mask)[0]
    if idx.size == 0:
        return Spots(idx, np.zeros((0, 3)), np.zeros(0), 0, 0.0)
    sf = shape_factor(s[idx], r[idx], p)
    gn = np.sqrt(np.sum(g_lab[idx] ** 2, axis=1))
    coeffs, _ = atom_coeffs(ph.symbols, p.scattering_params)
    F = structure_factor(hkl[idx], gn, ph.xyz, ph.occ, ph.B, coeffs,
                         p.scattering_params, "reference")
    I = sf * (F * F.conjugate()).real
    mx = float(np.max(I))
    keep = I > mx * p.minimum_intensity
    return Spots(idx[keep], g_lab[idx[keep]], I[keep], int(idx.size), mx)


# --------------------------------------------------------------------------
# a12, a13: rendering  (simulation2d.py:261-285, :357-442;
#                       detector_functions.py:251-359)
# --------------------------------------------------------------------------


def pixel_coords(xy: np.ndarray, shape, calibration, center, in_plane_angle=0.0,
                 mirrored=False):
    """Polar round trip of Simulation2D._get_transformed_coordinates
    (simulation2d.py:261-285) in pixel units."""
    x = np.asarray(xy[:, 0], dtype=np.float64) / calibration
    y = np.asarray(xy[:, 1], dtype=np.float64) / calibration
    mf = -1 if mirrored else 1
    theta = mf * np.arctan2(y, x) + np.deg2rad(in_plane_angle)
    rd = np.sqrt(x ** 2 + y ** 2)
    return rd * np.cos(theta) + center[0], rd * np.sin(theta) + center[1]


def render_fast(xy, intensity, shape, sigma, calibration, in_plane_angle=0.0,
                mirrored=False, normalize=True, center=None):
    """fast=True rendering (simulation2d.py:405-442, detector_functions.py:293-300):
    truncate to int pixels, last write wins, scipy.ndimage.gaussian_filter
    (reflect, truncate 4), divide by the maximum."""
    from scipy import ndimage as ndi

    H, W = shape
    if center is None:
        center = (W // 2, H // 2)
    X, Y = pixel_coords(xy, shape, calibration, center, in_plane_angle, mirrored)
    inf = (X >= 0) & (X < W) & (Y >= 0) & (Y < H)
    I = np.asarray(intensity, dtype=np.float64)[inf]
    if I.size == 0:
        return np.zeros(shape)
    xi = X[inf].astype(int)
    yi = Y[inf].astype(int)
    out = np.zeros(shape)
    out[yi, xi] = I
    out = ndi.gaussian_filter(out, sigma)
    if normalize:
        mx = np.max(out)
        if mx == 0:
            return out
        out = out / mx
    return out


def subpixel_gaussian(coords_xy, intensity, shape, sigma, clip_threshold=1.0):
    """detector_functions.py:314-359: per-spot clipped box of
    I/(2 pi s^2) exp(-d^2 / 2 s^2); a spot is skipped when the box radius is NaN."""
    H, W = shape
    out = np.zeros(shape)
    pref = 1 / (2 * np.pi * sigma ** 2)
    ep = -1 / (2 * sigma ** 2)
    yy, xx = np.indices(shape)
    for (cx, cy), I in zip(np.asarray(coords_xy, dtype=np.float64)[:, :2],
                           np.asarray(intensity, dtype=np.float64)):
        with np.errstate(invalid="ignore", divide="ignore"):
            radius = np.sqrt(np.log(clip_threshold / (pref * I)) / ep)
        if np.isnan(radius):
            continue
        y0 = max(0, int(np.ceil(cy - radius)))
        y1 = min(H, int(np.floor(cy + radius + 1)))
        x0 = max(0, int(np.ceil(cx - radius)))
        x1 = min(W, int(np.floor(cx + radius + 1)))
        if y1 <= y0 or x1 <= x0:
            continue
        d2 = (yy[y0:y1, x0:x1] - cy) ** 2 + (xx[y0:y1, x0:x1] - cx) ** 2
        out[y0:y1, x0:x1] += I * pref * np.exp(ep * d2)
    return out


def render_subpixel(xy, intensity, shape, sigma, calibration, in_plane_angle=0.0,
                    mirrored=False, normalize=True, clip_threshold=1.0, center=None):
    """fast=False rendering: centre ((W-1)/2, (H-1)/2), float coordinates,
    :func:`subpixel_gaussian`, optional normalisation (simulation2d.py:405-442)."""
    H, W = shape
    if center is None:
        center = ((W - 1) / 2, (H - 1) / 2)
    X, Y = pixel_coords(xy, shape, calibration, center, in_plane_angle, mirrored)
    inf = (X >= 0) & (X < W) & (Y >= 0) & (Y < H)
    I = np.asarray(intensity, dtype=np.float64)[inf]
    if I.size == 0:
        return np.zeros(shape)
    out = subpixel_gaussian(np.stack([X[inf], Y[inf]], 1), I, shape, sigma, clip_threshold)
    if normalize:
        mx = np.max(out)
        if mx == 0:
            return out
        out = out / mx
    return out


# --------------------------------------------------------------------------
# Reference-shaped library loop (the CPU baseline)
# --------------------------------------------------------------------------


def simulate_library(ph: PhaseSpec, rotations_M: np.ndarray, p: SimParams,
                     render: Optional[dict] = None):
    """calculate_diffraction2d for one phase and many orientations, followed by
    get_diffraction_pattern per orientation when ``render`` is given
    (simulation_generator.py:198-241 + simulation2d.py:357-442).  Returns the
    list of :class:`Spots` and, if rendering, an (n, H, W) float64 array."""
    base, rec = ph.basis()
    rows = hkl_box(rec, 1.0 / p.reciprocal_radius, include_zero=p.with_direct_beam)
    k = 1.0 / electron_wavelength(p.kv)
    spots = []
    pats = None
    if render is not None:
        pats = np.zeros((len(rotations_M),) + tuple(render["shape"]))
    for i, M in enumerate(rotations_M):
        sp = simulate_one_reference(ph, rows, rec, M, k, p)
        spots.append(sp)
        if render is not None:
            pats[i] = render_fast(sp.xyz[:, :2], sp.intensity, render["shape"],
                                  render["sigma"], render["calibration"])
    return spots, pats


# --------------------------------------------------------------------------
# Polar templates (simulation2d.py:313-355, 767-781; _diffracting_vector.py:186-194)
# --------------------------------------------------------------------------


def nearest_indices(axis, values):
    """For each value, the index of the closest axis entry; the lower neighbour
    wins only when strictly closer (get_closest, simulation2d.py:767-781)."""
    axis = np.asarray(axis, dtype=np.float64)
    values = np.asarray(values, dtype=np.float64)
    out = np.empty(values.shape, dtype=np.int64)
    for i, v in enumerate(values.ravel()):
        hi = int(np.sum(axis < v))            # first index with axis >= v
        if hi == len(axis):
            out.flat[i] = hi - 1
            continue
        lo = max(hi - 1, 0)
        out.flat[i] = hi - 1 if abs(v - axis[lo]) < abs(v - axis[hi]) else hi
    return out


def polar_flatten(spot_lists, radial_axes=None, azimuthal_axes=None):
    """(r, theta, intensity) arrays (n, max spots) of lab-frame spot lists
    [(xyz (S, 3), intensity (S,))] as polar_flatten_simulations builds them."""
    width = max(len(I) for _, I in spot_lists)
    n = len(spot_lists)
    r_t, t_t, i_t = np.zeros((n, width)), np.zeros((n, width)), np.zeros((n, width))
    for i, (xyz, I) in enumerate(spot_lists):
        xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        r = np.sqrt(xyz[:, 0] ** 2 + xyz[:, 1] ** 2)
        t = np.arctan2(xyz[:, 1], xyz[:, 0])
        I = np.asarray(I, dtype=np.float64)
        if radial_axes is not None and azimuthal_axes is not None:
            r = nearest_indices(radial_axes, r)
            t = nearest_indices(azimuthal_axes, t)
            keep = (r < len(radial_axes) - 1) & (t < len(azimuthal_axes) - 1)
            r, t, I = r[keep], t[keep], I[keep]
        r_t[i, :len(r)] = r
        t_t[i, :len(t)] = t
        i_t[i, :len(I)] = I
    if radial_axes is not None and azimuthal_axes is not None:
        r_t, t_t = r_t.astype(int), t_t.astype(int)
    return r_t, t_t, i_t


# ---------------------------------------------------------------------------
# beam-direction grids (rotation lists for template libraries)
# ---------------------------------------------------------------------------

def uv_sphere_vertices(resolution: float) -> np.ndarray:
    """UV (spherical-coordinate) mesh: get_uv_sphere_mesh_vertices
    (diffsims/generators/sphere_mesh_generators.py:42-93) with
    vectorised_spherical_polars_to_cartesians (diffsims/utils/vector_utils.py:82-105):
    psi outer, theta inner, one pole point each (psi == 0)."""
    n_theta = int(np.ceil(180 / resolution)) + 1
    n_psi = int(np.ceil(360 / resolution))
    psi = np.linspace(0, 2 * np.pi, num=n_psi, endpoint=False)
    theta = np.linspace(0, np.pi, num=n_theta, endpoint=True)
    P = np.repeat(psi, n_theta)
    T = np.tile(theta, n_psi)
    keep = (np.abs(T) > 0) | (P == 0)
    P, T = P[keep], T[keep]
    keep = (np.abs(T) < np.deg2rad(180)) | (P == 0)
    P, T = P[keep], T[keep]
    return np.stack([np.cos(P) * np.sin(T), np.sin(P) * np.sin(T), np.cos(T)], axis=1)


def cube_mesh_vertices(resolution: float, grid_type: str = "spherified_corner") -> np.ndarray:
    """Cube mesh projected on the sphere: get_cube_mesh_vertices
    (sphere_mesh_generators.py:96-197): per-face grid coordinates i (linear,
    equal angles towards the edges, or equal angles towards the corners), the
    six faces in the order bottom, top, east, west, south, north, then the two
    missing corners; rows normalised."""
    if grid_type == "normalized":
        steps = np.ceil(1 / np.tan(np.deg2rad(resolution)))
        i = np.arange(-steps, steps) / steps
    elif grid_type == "spherified_edge":
        steps = np.ceil(np.rad2deg(np.deg2rad(45)) / resolution)
        i = np.tan(np.arange(-steps, steps) * (np.arctan(1.0) / steps))
    elif grid_type == "spherified_corner":
        steps = np.ceil(np.arccos(1 / np.sqrt(3)) / np.deg2rad(resolution))
        i = np.tan(np.arange(-steps, steps) * (np.arctan(np.sqrt(2)) / steps)) / np.sqrt(2)
    else:
        raise ValueError(f"grid type {grid_type} not a valid grid type")
    x, y = np.meshgrid(i, i)
    x, y = x.ravel(), y.ravel()
    z = np.ones(x.shape[0])
    faces = [(-x, -y, -z), (x, y, z), (z, x, -y), (-z, -x, y), (x, -z, y), (-x, z, -y)]
    v = np.vstack([np.stack(f, axis=1) for f in faces] + [np.array([[-1.0, 1, 1], [1, -1, -1]])])
    return v / np.linalg.norm(v, axis=1)[:, None]


# icosahedron of get_icosahedral_mesh_vertices (sphere_mesh_generators.py:394-431)
_GOLDEN = (1.0 + np.sqrt(5.0)) / 2.0
ICOSA_CORNERS = np.array(
    [[-1, _GOLDEN, 0], [1, _GOLDEN, 0], [-1, -_GOLDEN, 0], [1, -_GOLDEN, 0],
     [0, -1, _GOLDEN], [0, 1, _GOLDEN], [0, -1, -_GOLDEN], [0, 1, -_GOLDEN],
     [_GOLDEN, 0, -1], [_GOLDEN, 0, 1], [-_GOLDEN, 0, -1], [-_GOLDEN, 0, 1]])
ICOSA_FACES = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
               (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
               (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]


def icosa_edge_order(faces=ICOSA_FACES):
    """The 30 edges as sorted corner pairs, in the order the reference visits
    them: the iteration order of a Python set filled face by face with the
    pairs (f0,f1), (f1,f2), (f2,f0) (_compose_from_faces,
    sphere_mesh_generators.py:224-232).  CPython's set order for tuples of
    small ints is fixed by their hashes (no hash randomisation for ints)."""
    edges = set()
    for f in faces:
        for p, q in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            edges.add((min(p, q), max(p, q)))
    return list(edges)


def icosa_refined_vertices(n: int) -> np.ndarray:
    """Vertices of the icosahedron with every face split into n^2 triangles
    (_compose_from_faces, sphere_mesh_generators.py:200-325; only the vertex
    array is returned there): the 12 corners; then per edge in
    icosa_edge_order(), the n-1 points (1-t) c0 + t c1 with
    t = linspace(1/n, 1, n-1, endpoint=False); then per face (f0, f1, f2) the
    interior points for i = 1..n-1 (outer), j = 1..n-i-1 (inner) with
    barycentric weights (1 - i/n - j/n, j/n, i/n) on (f0, f1, f2), combined
    by a matrix product as the reference's np.dot."""
    C = ICOSA_CORNERS.astype(np.float64)
    parts = [C]
    t = np.linspace(1 / n, 1.0, n - 1, endpoint=False)
    for i0, i1 in icosa_edge_order():
        parts.append(np.outer(1 - t, C[i0]) + np.outer(t, C[i1]))
    if n > 1:
        ii = [i for i in range(1, n) for _ in range(n - i - 1)]
        jj = [j for i in range(1, n) for j in range(1, n - i)]
        b = np.array([ii, jj], dtype=np.float64) / n
        bary = np.array([1.0 - b[0] - b[1], b[1], b[0]])
        for f in ICOSA_FACES:
            corner_verts = np.array([C[k] for k in f])
            parts.append(np.dot(corner_verts.T, bary).T)
    return np.concatenate(parts)


def max_nn_angle(v: np.ndarray) -> float:
    """Largest angle (degrees) between a mesh point and its nearest neighbour
    (_get_max_grid_angle / _get_angles_between_nn_gridpoints /
    _get_first_nearest_neighbors, sphere_mesh_generators.py:328-372): rows
    normalised, nearest neighbour by a k-d tree (k = 2, leaf size 50)."""
    from scipy.spatial import cKDTree

    u = (v.T / np.linalg.norm(v, axis=1)).T
    nn = cKDTree(u, 50).query(u, k=2)[1][:, 1]
    return float(np.max(np.rad2deg(np.arccos(np.sum(u * u[nn], axis=1)))))


def icosahedral_level(resolution: float):
    """Refinement level n the reference settles on: the corners' angle, then
    levels 1, 2, ... until the largest nearest-neighbour angle is <= the
    resolution (get_icosahedral_mesh_vertices, :432-438).  Returns n and the
    angle of every level visited ([level, angle] rows, the corners first as
    level 0).  A resolution at or above the corners' angle leaves the
    reference without vertices (its loop never runs): ValueError here."""
    angle = max_nn_angle(ICOSA_CORNERS.astype(np.float64))
    seen = [(0, angle)]
    n = 1
    if not angle > resolution:
        raise ValueError("resolution too coarse for the icosahedral mesh")
    while angle > resolution:
        angle = max_nn_angle(icosa_refined_vertices(n))
        seen.append((n, angle))
        n += 1
    return n - 1, seen


def icosahedral_vertices_at_level(n: int) -> np.ndarray:
    """The refined icosahedron of level n with rows divided by the square root
    of the row-wise dot product (einsum), as get_icosahedral_mesh_vertices
    returns it (sphere_mesh_generators.py:446-449)."""
    v = icosa_refined_vertices(n)
    norms = np.sqrt(np.einsum("ij,ij->i", v, v))
    return (v.T / norms.T).T


def icosahedral_mesh_vertices(resolution: float) -> np.ndarray:
    """get_icosahedral_mesh_vertices (sphere_mesh_generators.py:376-449) at
    the level icosahedral_level(resolution) settles on."""
    n, _ = icosahedral_level(resolution)
    return icosahedral_vertices_at_level(n)


def random_sphere_vertices(resolution: float, seed=None) -> np.ndarray:
    """get_random_sphere_vertices (sphere_mesh_generators.py:452-483):
    int((360/res)^2 / (4 pi)) standard-normal triples of
    numpy.random.default_rng(seed), normalised."""
    number = int(1 / (4 * np.pi) * (360 / resolution) ** 2)
    rng = np.random.default_rng(seed=seed) if seed is not None else np.random.default_rng()
    xyz = rng.normal(size=(number, 3))
    return (xyz.T / np.linalg.norm(xyz, axis=1)).T


def directions_to_euler(v: np.ndarray) -> np.ndarray:
    """beam_directions_grid_to_euler (sphere_mesh_generators.py:486-526): Bunge
    angles in degrees with phi1 = 0, Phi = acos(z/|v|), phi2 = 90 deg minus
    the signed azimuth (sign of y, of x where y == 0; 0 on the pole)."""
    v = np.asarray(v, dtype=np.float64)
    norm = np.linalg.norm(v, axis=1)
    Phi = np.arccos(v[:, 2] / norm)
    nproj = np.linalg.norm(v[:, :2], axis=1)
    sign = np.sign(v[:, 1])
    sign[v[:, 1] == 0] = np.sign(v[v[:, 1] == 0, 0])
    with np.errstate(invalid="ignore", divide="ignore"):
        phi2 = sign * np.nan_to_num(np.arccos(v[:, 0] / nproj))
    return np.rad2deg(np.stack([np.zeros(len(v)), Phi, np.pi / 2 - phi2], axis=1))


# stereographic triangles, rotation_list_generators.py:48-55
CRYSTAL_TRIANGLES = {
    "cubic": [(0, 0, 1), (1, 1, 1), (1, 0, 1)],
    "hexagonal": [(0, 0, 0, 1), (9, 1, -10, 0), (2, -1, -1, 0)],
    "trigonal": [(0, 0, 0, 1), (-2, 1, 1, 0), (-1, 2, -1, 0)],
    "tetragonal": [(0, 0, 1), (1, 0, 0), (1, 1, 0)],
    "orthorhombic": [(0, 0, 1), (-1, 0, 0), (0, 1, 0)],
    "monoclinic": [(0, -1, 0), (0, 0, 1), (0, 1, 0)],
}


def uvtw_to_uvw(d):
    """Four-index direction -> three-index (diffsims/utils/sim_utils.py:496-510)."""
    u, v, t, w = d
    u, v, w = 2 * u + v, 2 * v + u, w
    g = math.gcd(math.gcd(u, v), w)
    return tuple(int(x / g) for x in (u, v, w))


def beam_directions_grid(crystal_system: str, resolution: float,
                         mesh: str = "spherified_cube_edge", vertices=None) -> np.ndarray:
    """get_beam_directions_grid (diffsims/generators/rotation_list_generators.py:
    176-267): the mesh, then the crop to the crystal system's stereographic
    triangle (each side: the sign of the triangle's opposite corner,
    tolerance -1e-13), then directions_to_euler.  The monoclinic branch keeps
    only its second test (x >= -1e-13): its first result is overwritten
    (rotation_list_generators.py:236-241).  mesh="random" takes its points
    from ``vertices`` when given (the reference draws them unseeded)."""
    if vertices is not None:
        v = np.asarray(vertices, dtype=np.float64)
    elif mesh == "random":
        v = random_sphere_vertices(resolution)
    elif mesh == "icosahedral":
        v = icosahedral_mesh_vertices(resolution)
    elif mesh == "uv_sphere":
        v = uv_sphere_vertices(resolution)
    elif mesh == "spherified_cube_corner":
        v = cube_mesh_vertices(resolution, "spherified_corner")
    elif mesh in ("normalized_cube", "spherified_cube_edge"):
        if crystal_system == "hexagonal":
            resolution = resolution / np.sqrt(2)
        v = cube_mesh_vertices(resolution, "normalized" if mesh == "normalized_cube"
                               else "spherified_edge")
    else:
        raise NotImplementedError(f"mesh {mesh}")
    eps = -1e-13
    if crystal_system == "triclinic":
        return directions_to_euler(v)
    if crystal_system == "monoclinic":
        return directions_to_euler(v[np.dot(np.array([1, 0, 0]), v.T) >= eps])
    a, b, c = CRYSTAL_TRIANGLES[crystal_system]
    if len(a) == 4:
        a, b, c = uvtw_to_uvw(a), uvtw_to_uvw(b), uvtw_to_uvw(c)
    for p, q, r in ((a, b, c), (b, c, a), (c, a, b)):
        n = np.cross(p, q)
        v = v[np.dot(n, r) * np.dot(n, v.T) >= eps]
    return directions_to_euler(v)


# ---------------------------------------------------------------------------
# legacy generator (DiffractionGenerator / get_points_in_sphere / profiles)
# ---------------------------------------------------------------------------

def diffpy_base_from_conditions(a, b, c, alpha, beta, gamma) -> np.ndarray:
    """Rows a, b, c of the frame the legacy generator works in (diffpy's
    standard base), built from its defining conditions: c along +z, b in the
    yz-plane with b_y > 0, a with a.c = ac cos(beta), a.b = ab cos(gamma),
    |a| = a and a_x > 0 (so a* || +x)."""
    ca, cb, cg = cosd(alpha), cosd(beta), cosd(gamma)
    sa = sind(alpha)
    cvec = np.array([0.0, 0.0, c])
    bvec = np.array([0.0, b * sa, b * ca])
    az = a * cb
    ay = a * (cg - cb * ca) / sa
    ax = math.sqrt(a * a - ay * ay - az * az)
    return np.array([[ax, ay, az], bvec, cvec])


def legacy_points_in_sphere(abc, angles, radius):
    """get_points_in_sphere (diffsims/utils/sim_utils.py:436-474) on the
    reciprocal lattice of the legacy frame: H = floor(r / |a*|) etc., the
    product of ascending ranges (h outer, l inner), kept where |g| < r
    (|g| from the quadratic form of rnorm), (000) included.  Returns hkl
    (n, 3) int, g (n, 3) Cartesian and |g|."""
    base = diffpy_base_from_conditions(*abc, *angles)
    rec = np.linalg.inv(base)
    lengths = np.sqrt(np.sum(rec * rec, axis=0))
    H = [int(np.floor(radius / v)) for v in lengths]
    grid = np.stack(np.meshgrid(np.arange(-H[0], H[0] + 1), np.arange(-H[1], H[1] + 1),
                                np.arange(-H[2], H[2] + 1), indexing="ij"), -1).reshape(-1, 3)
    d = rnorm(grid, rec)
    keep = d < radius
    hkl = grid[keep]
    return hkl, g_cartesian(hkl, rec), d[keep]


def rzxz_matrix(ai, aj, ak) -> np.ndarray:
    """transforms3d euler2mat(ai, aj, ak, 'rzxz') (radians): rotating axes
    z, x', z'', written out as the entries of Rz(ai) Rx(aj) Rz(ak)."""
    ca, sa, cb, sb, cc, sc = (math.cos(ai), math.sin(ai), math.cos(aj), math.sin(aj),
                              math.cos(ak), math.sin(ak))
    return np.array([[ca * cc - sa * cb * sc, -ca * sc - sa * cb * cc, sa * sb],
                     [sa * cc + ca * cb * sc, -sa * sc + ca * cb * cc, -ca * sb],
                     [sb * sc, sb * cc, cb]])


def legacy_tables(ph: PhaseSpec, p: SimParams) -> PhaseTables:
    """Legacy reflection set with g and |F|^2 (diffraction_generator.py:233-242)."""
    hkl, g, gn = legacy_points_in_sphere(ph.abc, ph.angles, p.reciprocal_radius)
    coeffs, _ = atom_coeffs(ph.symbols, p.scattering_params)
    F = structure_factor(hkl, gn, np.asarray(ph.xyz, float), np.asarray(ph.occ, float),
                         np.asarray(ph.B, float), coeffs, p.scattering_params, phase_mode="exact")
    return PhaseTables(hkl, g, gn, np.abs(F) ** 2, hkl.shape[0])


def legacy_simulate_one(tab: PhaseTables, euler_deg, k: float, p: SimParams) -> Spots:
    """calculate_ed_data for one orientation (diffraction_generator.py:192-331):
    g_lab = R g with R = rzxz_matrix of the angles in degrees, then the same
    intersection, shape factor and relative cut as simulate_one."""
    R = rzxz_matrix(*np.deg2rad(np.asarray(euler_deg, dtype=np.float64)))
    return simulate_one(tab, R.T, k, p)


def legacy_profile(ph: PhaseSpec, p: SimParams, minimum_intensity=1e-3, hexagonal=False):
    """calculate_profile_data (diffraction_generator.py:334-417) with
    get_intensities_params / get_unique_families (sim_utils.py:101-136,
    513-556): families by permutation of |h|,|k|,|l| in first-appearance
    order, representative = the largest member, intensity |F|^2 x
    multiplicity; returns (magnitudes, intensities, labels)."""
    hkl, g, gn = legacy_points_in_sphere(ph.abc, ph.angles, p.reciprocal_radius)
    fam = {}
    for row in hkl.tolist():
        hit = None
        for key in fam:
            if sorted(abs(v) for v in key) == sorted(abs(v) for v in row):
                hit = key
                break
        if hit is None:
            fam[tuple(row)] = [row]
        else:
            fam[hit].append(row)
    reps = np.array([sorted(v)[-1] for v in fam.values()])
    mult = np.array([len(v) for v in fam.values()], dtype=float)
    d_of = {tuple(h): d for h, d in zip(hkl.tolist(), gn)}
    g_reps = np.array([d_of[tuple(r)] for r in reps.tolist()])
    coeffs, _ = atom_coeffs(ph.symbols, p.scattering_params)
    F = structure_factor(reps, g_reps, np.asarray(ph.xyz, float), np.asarray(ph.occ, float),
                         np.asarray(ph.B, float), coeffs, p.scattering_params, phase_mode="exact")
    I = np.abs(F) ** 2 * mult
    rows = (reps[0], reps[1], -reps[0] - reps[1], reps[2]) if hexagonal else reps.tolist()
    labels = ["".join(str(int(x)) for x in xs) for xs in rows]
    peaks = {}
    for lab, i, d in zip(labels, I, g_reps):
        peaks[lab] = [i, d]
    mx = max(v[0] for v in peaks.values())
    x, y, keep = [], [], []
    for lab, (i, d) in peaks.items():
        if i / mx * 100 > minimum_intensity and lab != "000":
            x.append(d)
            y.append(i)
            keep.append(lab)
    return np.array(x), np.asarray(y) / max(y) * 100, keep


# ---------------------------------------------------------------------------
# SO(3) rotation grids: get_fundamental_zone_grid / get_local_grid /
# get_grid_around_beam_direction / get_list_from_orix
# (diffsims/generators/rotation_list_generators.py:58-173).  The reference
# samples SO(3) with orix (absent here) on the cubochoric grid of Rosca,
# Morawiec & De Graef, Modelling Simul. Mater. Sci. Eng. 22 (2014) 075013;
# Euler <-> quaternion conversions follow Rowenhorst et al., Modelling Simul.
# Mater. Sci. Eng. 23 (2015) 083501, with P = +1.
# ---------------------------------------------------------------------------

CUBO_SEMI_EDGE = 0.5 * np.pi ** (2.0 / 3.0)   # cube of volume pi^2
HOMO_RADIUS = (0.75 * np.pi) ** (1.0 / 3.0)   # ball of volume pi^2


def cubochoric_to_homochoric(cu) -> np.ndarray:
    """Volume-preserving map of the cube [-a, a]^3, a = pi^(2/3) / 2, onto
    the homochoric ball of radius (3 pi / 4)^(1/3) (Rosca et al. 2014): the
    pyramid of the dominant axis (z, then x, then y on ties) is rotated onto
    +-z; the square cross-section at height Z is mapped area-preservingly onto
    the Lambert-azimuthal preimage of the curved square (angle pi/12 * Y/X),
    then lifted onto the sphere of radius sqrt(6/pi) Z."""
    cu = np.asarray(cu, dtype=np.float64).reshape(-1, 3)
    a = np.abs(cu)
    pz = np.maximum(a[:, 0], a[:, 1]) <= a[:, 2]
    px = ~pz & (np.maximum(a[:, 2], a[:, 1]) <= a[:, 0])
    py = ~pz & ~px
    s = cu.copy()
    s[px] = cu[px][:, [1, 2, 0]]
    s[py] = cu[py][:, [2, 0, 1]]
    sc = (np.pi / 6.0) ** (1.0 / 6.0)
    beta = 0.5 * np.pi ** (5.0 / 6.0) / 6.0 ** (1.0 / 6.0)
    prek = HOMO_RADIUS * 2.0 ** 0.25 / beta
    pref = np.sqrt(6.0 / np.pi)
    X, Y, Z = (sc * s).T
    out = np.zeros_like(s)
    on_axis = (X == 0) & (Y == 0)
    out[on_axis, 2] = pref * Z[on_axis]
    g = ~on_axis
    Xg, Yg, Zg = X[g], Y[g], Z[g]
    ylex = np.abs(Yg) <= np.abs(Xg)
    num = np.where(ylex, Yg, Xg)
    den = np.where(ylex, Xg, Yg)
    t = np.pi / 12.0 * num / den
    c, sn = np.cos(t), np.sin(t)
    q = prek * den / np.sqrt(np.sqrt(2.0) - c)
    u = (np.sqrt(2.0) * c - 1.0) * q   # along the dominant in-plane axis
    v = np.sqrt(2.0) * sn * q
    T1 = np.where(ylex, u, v)
    T2 = np.where(ylex, v, u)
    rho2 = T1 * T1 + T2 * T2
    lift = np.sqrt(1.0 - np.pi * rho2 / (24.0 * Zg * Zg))
    out[g, 0] = T1 * lift
    out[g, 1] = T2 * lift
    out[g, 2] = pref * Zg - np.sqrt(np.pi) * rho2 / np.sqrt(24.0) / Zg
    res = out.copy()
    res[px] = out[px][:, [2, 0, 1]]
    res[py] = out[py][:, [1, 2, 0]]
    return res


def _omega_minus_sin(w):
    """w - sin(w) without cancellation for small w (Taylor series below 0.5)."""
    w = np.asarray(w, dtype=np.float64)
    w2 = w * w
    series = w * w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0 * (1.0 - w2 / 72.0 * (
        1.0 - w2 / 110.0 * (1.0 - w2 / 156.0)))))
    return np.where(w < 0.5, series, w - np.sin(w))


NEWTON_STEPS = 6


def homochoric_to_axis_angle(ho):
    """Unit axis and angle of homochoric vectors: |h|^3 = (3/4)(w - sin w),
    solved by NEWTON_STEPS steps of Newton's method from w0 = 2|h| (below the
    root; the function is convex on [0, pi], so the iteration then converges
    from above: four steps reach the last bit from any start on [0, pi])."""
    ho = np.asarray(ho, dtype=np.float64).reshape(-1, 3)
    hn = np.linalg.norm(ho, axis=1)
    t = (4.0 / 3.0) * (hn * hn * hn)
    w = np.minimum(2.0 * hn, np.pi)
    for _ in range(NEWTON_STEPS):
        f = _omega_minus_sin(w) - t
        d = 2.0 * np.sin(0.5 * w) ** 2
        step = np.where(d > 0, f / np.where(d > 0, d, 1.0), 0.0)
        w = np.clip(w - step, 0.0, np.pi)
    axis = np.zeros_like(ho)
    nz = hn > 0
    axis[nz] = ho[nz] / hn[nz, None]
    axis[~nz] = (0.0, 0.0, 1.0)
    return axis, w


def quat_from_axis_angle(axis, w) -> np.ndarray:
    axis = np.asarray(axis, dtype=np.float64).reshape(-1, 3)
    w = np.asarray(w, dtype=np.float64).reshape(-1)
    return np.concatenate([np.cos(0.5 * w)[:, None], np.sin(0.5 * w)[:, None] * axis], axis=1)


def quat_multiply(p, q) -> np.ndarray:
    """Hamilton product p q (rotation q first, then p)."""
    p = np.asarray(p, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    a1, b1, c1, d1 = np.moveaxis(p, -1, 0)
    a2, b2, c2, d2 = np.moveaxis(q, -1, 0)
    return np.stack([a1 * a2 - b1 * b2 - c1 * c2 - d1 * d2,
                     a1 * b2 + b1 * a2 + c1 * d2 - d1 * c2,
                     a1 * c2 - b1 * d2 + c1 * a2 + d1 * b2,
                     a1 * d2 + b1 * c2 - c1 * b2 + d1 * a2], axis=-1)


def bunge_from_quat(q) -> np.ndarray:
    """Bunge angles (radians) of quaternions, Rowenhorst et al. (2015) eq.
    A.14 with P = +1, including its two degenerate branches (Phi = 0 and
    Phi = pi); negative phi1 / phi2 are moved into [0, 2 pi), Phi into
    [0, pi]: what orix ``Rotation.to_euler`` returns."""
    q = np.asarray(q, dtype=np.float64).reshape(-1, 4)
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    q03 = w * w + z * z
    q12 = x * x + y * y
    chi = np.sqrt(q03 * q12)
    out = np.zeros((q.shape[0], 3))
    m1 = (chi == 0) & (q12 == 0)
    m2 = (chi == 0) & (q03 == 0) & ~m1
    m3 = ~(m1 | m2)
    out[m1, 0] = np.arctan2(-2.0 * w[m1] * z[m1], w[m1] ** 2 - z[m1] ** 2)
    out[m2, 0] = np.arctan2(2.0 * x[m2] * y[m2], x[m2] ** 2 - y[m2] ** 2)
    out[m2, 1] = np.pi
    cm = chi[m3]
    W, X, Y, Z = w[m3], x[m3], y[m3], z[m3]
    out[m3, 0] = np.arctan2((X * Z - W * Y) / cm, (-W * X - Y * Z) / cm)
    out[m3, 1] = np.arctan2(2.0 * cm, q03[m3] - q12[m3])
    out[m3, 2] = np.arctan2((W * Y + X * Z) / cm, (Y * Z - W * X) / cm)
    for k, period in ((0, 2 * np.pi), (1, np.pi), (2, 2 * np.pi)):
        neg = out[:, k] < 0
        out[neg, k] = np.mod(out[neg, k] + 100 * np.pi, period)
    return out


def cubochoric_steps(resolution: float) -> int:
    """Semi-edge steps N for a requested neighbour distance (degrees): the
    grid spacing a/N maps to 2 (6/pi)^(1/3) a / N radians of rotation next
    to the identity; N is the smallest count whose spacing there does not
    exceed the resolution.  (orix's own resolution -> N rule is not in the
    reference and is parity unpinned.)"""
    rad = np.deg2rad(float(resolution))
    if not rad > 0:
        raise ValueError(f"resolution must be positive, got {resolution}")
    return max(1, int(np.ceil(2.0 * (6.0 / np.pi) ** (1.0 / 3.0) * CUBO_SEMI_EDGE / rad - 1e-9)))


def cubochoric_quaternions(n: int) -> np.ndarray:
    """The cubochoric grid with n semi-edge steps: coordinates i a/n for
    i = -n+1 .. n on each axis (the -a faces are the +a faces' antipodes),
    x outermost and z innermost, as unit quaternions with w >= 0."""
    ax = np.arange(-n + 1, n + 1, dtype=np.float64) * (CUBO_SEMI_EDGE / n)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    cu = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    axis, w = homochoric_to_axis_angle(cubochoric_to_homochoric(cu))
    return quat_from_axis_angle(axis, w)


def _axis_rot(axis, n):
    axis = np.asarray(axis, dtype=np.float64) / np.linalg.norm(axis)
    return quat_from_axis_angle(axis, 2 * np.pi / n)[0]


# generators of the 11 proper point groups, axes as orix sets them: n-fold
# axes along z, two-fold axes along x, three-folds along <111>
_PROPER_GENERATORS = {
    "1": [], "2": [(0, 0, 1, 2)], "222": [(0, 0, 1, 2), (1, 0, 0, 2)],
    "4": [(0, 0, 1, 4)], "422": [(0, 0, 1, 4), (1, 0, 0, 2)],
    "3": [(0, 0, 1, 3)], "32": [(0, 0, 1, 3), (1, 0, 0, 2)],
    "6": [(0, 0, 1, 6)], "622": [(0, 0, 1, 6), (1, 0, 0, 2)],
    "23": [(0, 0, 1, 2), (1, 0, 0, 2), (1, 1, 1, 3)], "432": [(0, 0, 1, 4), (1, 1, 1, 3)],
}

# point group -> rotation group of its Laue class
_LAUE_PROPER = {"1": "1", "-1": "1", "2": "2", "m": "2", "2/m": "2", "222": "222",
                "mm2": "222", "mmm": "222", "4": "4", "-4": "4", "4/m": "4", "422": "422",
                "4mm": "422", "-42m": "422", "4/mmm": "422", "3": "3", "-3": "3", "32": "32",
                "3m": "32", "-3m": "32", "6": "6", "-6": "6", "6/m": "6", "622": "622",
                "6mm": "622", "-6m2": "622", "6/mmm": "622", "23": "23", "m-3": "23",
                "432": "432", "-43m": "432", "m-3m": "432"}


def proper_group_quaternions(name: str) -> np.ndarray:
    """All rotations of a proper point group as quaternions (w >= 0), by
    closure of the generators."""
    gens = [_axis_rot(g[:3], g[3]) for g in _PROPER_GENERATORS[name]]
    elems = [np.array([1.0, 0.0, 0.0, 0.0])]
    frontier = list(elems)

    def key(q):
        q = q if (q[0] > 1e-9 or (abs(q[0]) <= 1e-9 and tuple(np.round(q[1:], 9)) > (0, 0, 0))) else -q
        return tuple(np.round(q, 9))

    seen = {key(elems[0])}
    while frontier:
        nxt = []
        for a in frontier:
            for g in gens:
                m = quat_multiply(g, a)
                k = key(m)
                if k not in seen:
                    seen.add(k)
                    elems.append(m)
                    nxt.append(m)
        frontier = nxt
    return np.array(elems)


def proper_group_of_space_group(number: int) -> str:
    from_sg = [(1, "1"), (2, "-1"), (5, "2"), (9, "m"), (15, "2/m"), (24, "222"), (46, "mm2"),
               (74, "mmm"), (80, "4"), (82, "-4"), (88, "4/m"), (98, "422"), (110, "4mm"),
               (122, "-42m"), (142, "4/mmm"), (146, "3"), (148, "-3"), (155, "32"),
               (161, "3m"), (167, "-3m"), (173, "6"), (174, "-6"), (176, "6/m"), (182, "622"),
               (186, "6mm"), (190, "-6m2"), (194, "6/mmm"), (199, "23"), (206, "m-3"),
               (214, "432"), (220, "-43m"), (230, "m-3m")]
    n = int(number)
    if not 1 <= n <= 230:
        raise ValueError(f"space group number {number} is not in 1..230")
    for last, pg in from_sg:
        if n <= last:
            return _LAUE_PROPER[pg]
    raise AssertionError


FZ_EPS = 1e-12


def fundamental_zone_margin(q, ops) -> np.ndarray:
    """|w| - max_s |q . s| over the group's non-identity rotations: >= 0 when
    q is the member of its class q s closest to the identity (the
    fundamental zone of orix's OrientationRegion, whose faces are the
    bisectors between the identity and each s)."""
    q = np.asarray(q, dtype=np.float64).reshape(-1, 4)
    others = [s for s in ops if not np.allclose(np.abs(s), [1, 0, 0, 0])]
    if not others:
        return np.full(q.shape[0], np.inf)
    d = np.abs(q @ np.array(others).T).max(axis=1)
    return np.abs(q[:, 0]) - d


def fundamental_zone_grid(resolution: float, space_group: int) -> np.ndarray:
    """get_fundamental_zone_grid (rotation_list_generators.py:85-106) before
    the rounding of get_list_from_orix: Bunge angles (radians) of the
    cubochoric grid rows inside the fundamental zone of the space group's
    proper point group (faces inclusive by FZ_EPS)."""
    q = cubochoric_quaternions(cubochoric_steps(resolution))
    ops = proper_group_quaternions(proper_group_of_space_group(space_group))
    keep = fundamental_zone_margin(q, ops) >= -FZ_EPS
    return bunge_from_quat(q[keep])


def local_grid(resolution: float, center=None, grid_width: float = 10.0) -> np.ndarray:
    """get_local_grid (rotation_list_generators.py:109-134): cubochoric rows
    whose rotation angle is below grid_width (arccos(w) < grid_width / 2),
    premultiplied by the centre rotation (Bunge degrees) when one is given."""
    q = cubochoric_quaternions(cubochoric_steps(resolution))
    half = np.deg2rad(grid_width / 2.0)
    q = q[np.arccos(q[:, 0]) < half]
    if center is not None:
        c = quat_from_bunge(np.deg2rad(np.asarray(center, dtype=np.float64)))
        q = quat_multiply(c[None, :], q)
    return bunge_from_quat(q)


def grid_around_beam_direction(beam_rotation, resolution, angular_range=(0, 360)) -> np.ndarray:
    """get_grid_around_beam_direction (rotation_list_generators.py:137-173):
    the beam rotation (Bunge degrees) times rotations about z by
    arange(start, stop, resolution) degrees."""
    b = quat_from_bunge(np.deg2rad(np.asarray(beam_rotation, dtype=np.float64)))
    ang = np.deg2rad(np.arange(start=angular_range[0], stop=angular_range[1], step=resolution))
    axes = np.repeat([[0.0, 0.0, 1.0]], ang.shape[0], axis=0)
    q = quat_multiply(b[None, :], quat_from_axis_angle(axes, ang))
    return bunge_from_quat(q)


def rotation_list(euler_rad, rounding: int = 2):
    """get_list_from_orix (rotation_list_generators.py:58-82): rows in degrees
    rounded to `rounding` decimals, as tuples."""
    return [tuple(np.round(np.rad2deg(r), decimals=rounding))
            for r in np.asarray(euler_rad, dtype=np.float64).reshape(-1, 3).tolist()]
