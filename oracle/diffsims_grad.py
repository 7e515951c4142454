"""FP64 reference for gradients of the sub-pixel pattern path.

TEST INFRASTRUCTURE ONLY.  Like ``diffsims_oracle`` this module is a checker:
only ``tests/`` may import it; the product never does.

diffsims computes no gradients, so there is no reference function to restate
here.  This module restates the forward path of ``diffsims_oracle`` --
``phase_tables`` + ``simulate_one`` + ``render_subpixel`` and an L2 loss
(diffsims/generators/simulation_generator.py:211-241, :319-412;
diffsims/utils/sim_utils.py:256-354; diffsims/simulations/simulation2d.py:357-442;
diffsims/pattern/detector_functions.py:314-359) -- in torch float64 on the CPU
with the discrete structure frozen (which reflections survive, which pixels
each clip box covers), so that torch autograd gives its exact derivative.

Pins (tests/test_grad_oracle.py):
  * its loss and patterns equal the NumPy oracle's at the same parameters;
  * its gradients equal central finite differences (h = 1e-4) of the NumPy
    oracle's loss, recomputed from scratch at every shifted point;
  * its rotation and reciprocal-basis helpers equal ``bunge_matrix`` and
    ``lattice_basis`` of the NumPy oracle (they are built differently: matrix
    products of elementary rotations and a numerical inverse).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import diffsims_oracle as O

_EXACT = {0.0: 1.0, 60.0: 0.5, 90.0: 0.0, 120.0: -0.5, 180.0: -1.0,
          240.0: -0.5, 270.0: 0.0, 300.0: 0.5}


def _cosd(x: torch.Tensor) -> torch.Tensor:
    c = torch.cos(x * (math.pi / 180.0))
    r = math.fmod(float(x.detach()), 360.0) % 360.0
    if r in _EXACT:
        c = c + (_EXACT[r] - c).detach()
    return c


def reciprocal_basis(abc: torch.Tensor, angles: torch.Tensor) -> torch.Tensor:
    """inv(base) of the x || a, z || c* direct basis (rows a, b, c)."""
    a, b, c = abc[0], abc[1], abc[2]
    ca, cb, cg = _cosd(angles[0]), _cosd(angles[1]), _cosd(angles[2])
    sg = _cosd(90.0 - angles[2])
    t = (ca - cb * cg) / sg
    z = torch.zeros((), dtype=torch.float64)
    base = torch.stack([torch.stack([a, z, z]),
                        torch.stack([b * cg, b * sg, z]),
                        torch.stack([c * cb, c * t, c * torch.sqrt(1 - cb * cb - t * t)])])
    return torch.linalg.inv(base)


def _rz(t):
    c, s = torch.cos(t), torch.sin(t)
    o, z = torch.ones_like(t), torch.zeros_like(t)
    return torch.stack([torch.stack([c, s, z], -1), torch.stack([-s, c, z], -1),
                        torch.stack([z, z, o], -1)], -2)


def _rx(t):
    c, s = torch.cos(t), torch.sin(t)
    o, z = torch.ones_like(t), torch.zeros_like(t)
    return torch.stack([torch.stack([o, z, z], -1), torch.stack([z, c, s], -1),
                        torch.stack([z, -s, c], -1)], -2)


def bunge(euler: torch.Tensor) -> torch.Tensor:
    """Passive Bunge matrix as the product Rz(phi2) Rx(Phi) Rz(phi1)."""
    return _rz(euler[..., 2]) @ _rx(euler[..., 1]) @ _rz(euler[..., 0])


def shape_factor(s, r, p: O.SimParams):
    """Torch float64 versions of the built-in shape models (shape_factor_models.py)."""
    w = p.shape_factor_width if p.shape_factor_width is not None else p.max_excitation_error
    if p.precession_angle:
        phi = math.radians(abs(p.precession_angle))
        sg = math.pi / w
        u = sg ** 2 * (r ** 2 * phi ** 2 - s ** 2) + 1
        z = torch.sqrt(u ** 2 + 4 * sg ** 2 * s ** 2)
        return (sg / math.pi) * torch.sqrt(2 * (u + z) / z ** 2)
    m = p.shape_model
    n = p.shape_kwargs.get("minima_number", 5)
    if m == "linear":
        return torch.clamp(1 - torch.abs(s) / w, min=0.0)
    if m == "lorentzian":
        return 1.0 / (1.0 + (math.pi * s / w) ** 2)
    if m in ("sinc", "sin2c"):
        x = math.pi * n / w * s
        v = torch.where(x != 0, torch.sin(x) / torch.where(x != 0, x, torch.ones_like(x)),
                        torch.zeros_like(x))
        return torch.abs(v) if m == "sinc" else v * v
    if m == "atanc":
        x = math.pi * n / abs(w) * s
        safe = torch.where(x != 0, x, torch.ones_like(x))
        return torch.where(x != 0, torch.atan(safe) / safe, torch.ones_like(x))
    raise NotImplementedError(m)


@dataclass
class FrozenPattern:
    idx: np.ndarray          # survivor rows (the oracle's simulate_one)
    boxes: np.ndarray        # (S, 4) y0, y1, x0, x1 (half-open) or -1 where skipped
    argmax: int = -1         # first pixel of the pattern maximum (normalisation)


@dataclass
class Problem:
    """A differentiable test problem: one phase, n orientations, a detector."""
    ph: O.PhaseSpec
    p: O.SimParams
    shape: tuple
    sigma: float
    calibration: float
    clip: float = 1.0
    normalize: bool = True
    in_plane_angle: float = 0.0
    mirrored: bool = False


def _pixel_xy(xy, pr: Problem):
    H, W = pr.shape
    cx, cy = (W - 1) / 2, (H - 1) / 2
    x = xy[:, 0] / pr.calibration
    y = xy[:, 1] / pr.calibration
    a = math.radians(pr.in_plane_angle)
    ca, sa = math.cos(a), math.sin(a)
    if pr.mirrored:
        return x * ca + y * sa + cx, x * sa - y * ca + cy
    return x * ca - y * sa + cx, x * sa + y * ca + cy


def numpy_patterns(pr: Problem, M: np.ndarray, xyz=None, B=None, abc=None, angles=None):
    """The NumPy oracle's patterns at the given parameters (every discrete
    choice recomputed), plus the frozen structure of that evaluation."""
    ph = O.PhaseSpec(abc if abc is not None else pr.ph.abc,
                     angles if angles is not None else pr.ph.angles, pr.ph.symbols,
                     xyz if xyz is not None else pr.ph.xyz, pr.ph.occ,
                     B if B is not None else pr.ph.B)
    tab = O.phase_tables(ph, pr.p)
    k = 1 / O.electron_wavelength(pr.p.kv)
    pats, frozen = [], []
    for o in range(M.shape[0]):
        sp = O.simulate_one(tab, M[o], k, pr.p)
        pat, fz = freeze(pr, sp)
        pats.append(pat)
        frozen.append(fz)
    return np.array(pats), frozen, tab


def freeze(pr: Problem, sp: O.Spots, rows: Optional[np.ndarray] = None):
    """The NumPy oracle's sub-pixel pattern of one spot list and its frozen
    structure: each in-frame spot's clip box (detector_functions.py:333-358)
    and the first pixel of the maximum.  ``rows`` (default ``sp.idx``) are the
    table rows the frozen pattern refers to."""
    H, W = pr.shape
    pat = O.render_subpixel(sp.xyz[:, :2], sp.intensity, pr.shape, pr.sigma, pr.calibration,
                            pr.in_plane_angle, pr.mirrored, pr.normalize, pr.clip)
    X, Y = O.pixel_coords(sp.xyz[:, :2], pr.shape, pr.calibration, ((W - 1) / 2, (H - 1) / 2),
                          pr.in_plane_angle, pr.mirrored)
    boxes = np.full((len(sp.idx), 4), -1, dtype=np.int64)
    pref = 1 / (2 * np.pi * pr.sigma ** 2)
    ep = -1 / (2 * pr.sigma ** 2)
    for j in range(len(sp.idx)):
        if not (0 <= X[j] < W and 0 <= Y[j] < H):
            continue
        with np.errstate(invalid="ignore", divide="ignore"):
            rad = np.sqrt(np.log(pr.clip / (pref * sp.intensity[j])) / ep)
        if np.isnan(rad):
            continue
        y0, y1 = max(0, int(np.ceil(Y[j] - rad))), min(H, int(np.floor(Y[j] + rad + 1)))
        x0, x1 = max(0, int(np.ceil(X[j] - rad))), min(W, int(np.floor(X[j] + rad + 1)))
        if y1 > y0 and x1 > x0:
            boxes[j] = (y0, y1, x0, x1)
    idx = sp.idx if rows is None else np.asarray(rows)
    return pat, FrozenPattern(idx, boxes, int(np.argmax(pat)))


def numpy_loss(pr: Problem, M, target, **kw) -> float:
    pats, _, _ = numpy_patterns(pr, M, **kw)
    return float(np.sum((pats - target) ** 2))


def torch_patterns(pr: Problem, Rt: torch.Tensor, rec: torch.Tensor, xyz: torch.Tensor,
                   B: torch.Tensor, frozen: List[FrozenPattern], hkl: np.ndarray):
    """Patterns as a differentiable function of (Rt [n,3,3] = M^T, rec, xyz, B)
    with the frozen structure; float64 throughout."""
    p = pr.p
    hkl_t = torch.as_tensor(np.asarray(hkl, dtype=np.float64))
    g = hkl_t @ rec.T                                             # (N, 3)
    g2 = (g * g).sum(1)
    coeffs, _ = O.atom_coeffs(pr.ph.symbols, p.scattering_params)
    c = torch.as_tensor(coeffs)                                   # (n, 5, 2)
    bg = g2[:, None, None] * c[None, :, :, 1]
    if p.scattering_params == "lobato":
        f = (c[None, :, :, 0] * (2 + bg) / (1 + bg) ** 2).sum(-1)
    elif p.scattering_params == "xtables":
        f = (c[None, :, :, 0] * torch.exp(-0.25 * bg)).sum(-1)
    else:
        f = torch.ones(g2.shape[0], len(pr.ph.symbols), dtype=torch.float64)
    occ = torch.as_tensor(np.asarray(pr.ph.occ, dtype=np.float64))
    ph = 2 * math.pi * (hkl_t @ xyz.T)
    amp = f * occ * torch.exp(-0.25 * g2[:, None] * B[None, :])
    Fr = (amp * torch.cos(ph)).sum(1)
    Fi = (amp * torch.sin(ph)).sum(1)
    F2 = Fr * Fr + Fi * Fi
    k = 1 / O.electron_wavelength(p.kv)
    H, W = pr.shape
    pref = 1 / (2 * math.pi * pr.sigma ** 2)
    ep = -1 / (2 * pr.sigma ** 2)
    out = []
    for o, fz in enumerate(frozen):
        Q = torch.zeros((H, W), dtype=torch.float64)
        keep = np.nonzero(fz.boxes[:, 0] >= 0)[0] if len(fz.idx) else np.zeros(0, np.int64)
        if keep.size:
            rows = fz.idx[keep]
            gl = g[rows] @ Rt[o].T                                  # M^T g
            r2 = gl[:, 0] ** 2 + gl[:, 1] ** 2
            s = (k - torch.sqrt(k * k - r2)) - gl[:, 2]
            I = F2[rows] * shape_factor(s, torch.sqrt(r2), p)
            X, Y = _pixel_xy(gl[:, :2], pr)
            # every clip box as a masked patch of the largest box size
            bx = fz.boxes[keep]
            hmax = int((bx[:, 1] - bx[:, 0]).max())
            wmax = int((bx[:, 3] - bx[:, 2]).max())
            dy = np.arange(hmax)[None, :, None]
            dx = np.arange(wmax)[None, None, :]
            yy = bx[:, 0][:, None, None] + dy + 0 * dx
            xx = bx[:, 2][:, None, None] + dx + 0 * dy
            m = (yy < bx[:, 1][:, None, None]) & (xx < bx[:, 3][:, None, None])
            sel = np.nonzero(m)
            ys = torch.as_tensor(yy[sel])
            xs = torch.as_tensor(xx[sel])
            j = torch.as_tensor(sel[0])
            d2 = (ys.double() - Y[j]) ** 2 + (xs.double() - X[j]) ** 2
            Q = Q.index_put((ys, xs), I[j] * pref * torch.exp(ep * d2), accumulate=True)
        if pr.normalize and float(Q.detach().max()) > 0:
            Q = Q / Q.max()
        out.append(Q)
    return torch.stack(out)


def torch_loss(pr, Rt, rec, xyz, B, frozen, hkl, target) -> torch.Tensor:
    P = torch_patterns(pr, Rt, rec, xyz, B, frozen, hkl)
    return ((P - torch.as_tensor(target, dtype=torch.float64)) ** 2).sum()
