"""Differentiable pattern simulation: hand-written adjoint kernels behind
``torch.autograd.Function``.

diffsims computes no gradients; this module makes the rendered sub-pixel
patterns of the pattern path (``calculate_diffraction2d`` followed by
``get_diffraction_pattern(fast=False)``, simulation_generator.py:134-259,
simulation2d.py:357-442) differentiable with respect to

* the orientation of every pattern (``Rt``: float64 [n, 9], rows of M^T with
  M = orix ``to_matrix()``; build it from Euler angles with
  :func:`bunge_matrices` / :func:`rt_from_matrices` to get Euler gradients),
* the reciprocal basis ``rec`` (float64 [3, 3], CPU; build it from
  a, b, c, alpha, beta, gamma with :func:`reciprocal_basis`),
* the fractional atom positions ``xyz`` (float64 [n_atoms, 3]) and
* the Debye-Waller factors ``B`` (float64 [n_atoms], A^2).

The discrete structure (the hkl set, which reflections intersect the Ewald
shell, which pixels each spot's clip box covers) is fixed by the forward
evaluation and is piecewise constant in the parameters, as in any
differentiable renderer.

Two functions:

* :meth:`RefinementModel.l2_loss` -- sum over patterns and pixels of
  (P - T)^2.  When gradients are needed the forward launch computes the loss
  and every gradient in one pass (fused forward+backward): the render adjoint
  re-renders each pattern in LDS, forms 2 (P - T) there and pulls it back to
  the spots without writing the pattern to HBM.
* :meth:`RefinementModel.patterns` -- the rendered patterns; its backward
  re-runs the spot kernel and the render adjoint with the incoming dL/dP.

Both recompute F(hkl) and the excitation errors in the backward instead of
storing them per orientation.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib, engine
from .crystallography import (PhaseArrays, element_of, highest_hkl, lattice_setup,
                              reciprocal_metric)
from .scattering import table

# Long spot lists hand their per-spot gathered sums (n x cap x 16 B) from the
# render adjoint to its chain rule; calls are split to keep that bounded.
GRAD_SUMS_BYTES = 8 << 30

__all__ = ["bunge_matrices", "rt_from_matrices", "euler_to_rt", "reciprocal_basis",
           "RefinementModel", "leaf_parameters", "forward_args"]

_EXACT_COS = {0.0: 1.0, 60.0: 0.5, 90.0: 0.0, 120.0: -0.5, 180.0: -1.0,
              240.0: -0.5, 270.0: 0.0, 300.0: 0.5}


# ---------------------------------------------------------------------------
# differentiable parameterisations (torch, float64)
# ---------------------------------------------------------------------------


def bunge_matrices(euler: torch.Tensor) -> torch.Tensor:
    """Passive Bunge (ZXZ) matrices M(phi1, Phi, phi2) from radians, shape
    (..., 3, 3); M equals orix ``Rotation.from_euler(euler).to_matrix()``."""
    p1, P, p2 = euler.unbind(-1)
    c1, s1 = torch.cos(p1), torch.sin(p1)
    c, s = torch.cos(P), torch.sin(P)
    c2, s2 = torch.cos(p2), torch.sin(p2)
    rows = [
        torch.stack([c1 * c2 - s1 * s2 * c, s1 * c2 + c1 * s2 * c, s2 * s], -1),
        torch.stack([-c1 * s2 - s1 * c2 * c, -s1 * s2 + c1 * c2 * c, c2 * s], -1),
        torch.stack([s1 * s, -c1 * s, c], -1),
    ]
    return torch.stack(rows, -2)


def rt_from_matrices(M: torch.Tensor) -> torch.Tensor:
    """(n, 3, 3) orientation matrices -> contiguous [n, 9] rows of M^T, the
    layout the kernels read (g_lab = M^T g)."""
    return M.transpose(-1, -2).reshape(-1, 9).contiguous()


class _EulerToRt(torch.autograd.Function):
    @staticmethod
    def forward(ctx, euler):
        e = euler.detach().contiguous()
        rt = torch.empty((e.shape[0], 9), dtype=torch.float64, device=e.device)
        L = _lib.load(require_gpu=True)
        _lib.check(L.dsx_euler_to_rt(_lib.ptr(e), int(e.shape[0]), _lib.ptr(rt),
                                     _lib.stream_handle(e.device)), "dsx_euler_to_rt")
        ctx.save_for_backward(e)
        return rt

    @staticmethod
    def backward(ctx, g):
        (e,) = ctx.saved_tensors
        g = g.contiguous().to(torch.float64)
        de = torch.empty_like(e)
        L = _lib.load()
        _lib.check(L.dsx_euler_to_rt_backward(_lib.ptr(e), _lib.ptr(g), int(e.shape[0]),
                                              _lib.ptr(de), _lib.stream_handle(e.device)),
                   "dsx_euler_to_rt_backward")
        return de


def euler_to_rt(euler: torch.Tensor) -> torch.Tensor:
    """Bunge angles [n, 3] (radians, float64, on the device) -> Rt [n, 9] in one
    kernel, with a hand-written backward; equals
    ``rt_from_matrices(bunge_matrices(euler))``."""
    if euler.dtype != torch.float64 or euler.dim() != 2 or euler.shape[1] != 3 or not euler.is_cuda:
        raise ValueError("euler must be a float64 [n, 3] device tensor")
    return _EulerToRt.apply(euler)


def _cosd(x: torch.Tensor) -> torch.Tensor:
    """cos of degrees with the exact values of the lattice setup at multiples
    of 30/60/90 degrees; the derivative is that of cos."""
    c = torch.cos(torch.deg2rad(x))
    r = math.fmod(float(x.detach()), 360.0)
    if r < 0:
        r += 360.0
    if r in _EXACT_COS:
        c = c + (_EXACT_COS[r] - c).detach()
    return c


def _basis_and_jacobian(abc: np.ndarray, angles: np.ndarray):
    """Direct basis rows (x || a, z || c*) and d base / d(a, b, c, alpha, beta,
    gamma) (angles in degrees), shape (3, 3, 6)."""
    a, b, c = (float(v) for v in abc)
    al, be, ga = (math.radians(float(v)) for v in angles)
    ex = [_EXACT_COS.get(math.fmod(float(v), 360.0) % 360.0) for v in angles]
    ca = ex[0] if ex[0] is not None else math.cos(al)
    cb = ex[1] if ex[1] is not None else math.cos(be)
    cg = ex[2] if ex[2] is not None else math.cos(ga)
    sa, sb = math.sin(al), math.sin(be)
    sgx = _EXACT_COS.get(math.fmod(90.0 - float(angles[2]), 360.0) % 360.0)
    sg = sgx if sgx is not None else math.sin(ga)
    t = (ca - cb * cg) / sg
    root = math.sqrt(1.0 - cb * cb - t * t)
    base = np.array([[a, 0.0, 0.0], [b * cg, b * sg, 0.0], [c * cb, c * t, c * root]])
    J = np.zeros((3, 3, 6))
    d = math.pi / 180.0                           # per degree
    J[0, 0, 0] = 1.0
    J[1, 0, 1], J[1, 1, 1] = cg, sg
    J[2, 0, 2], J[2, 1, 2], J[2, 2, 2] = cb, t, root
    # t = (ca - cb cg) / sg
    dt = np.array([-sa / sg, sb * cg / sg, (cb * sg) / sg - (ca - cb * cg) * math.cos(ga) / sg ** 2])
    dt *= d
    droot = np.array([-t * dt[0], (cb * sb * d - t * dt[1]), -t * dt[2]]) / root
    J[1, 0, 5], J[1, 1, 5] = -b * math.sin(ga) * d, b * math.cos(ga) * d
    J[2, 0, 4] = -c * sb * d
    J[2, 1, 3:6] = c * dt
    J[2, 2, 3:6] = c * droot
    return base, J


class _ReciprocalBasis(torch.autograd.Function):
    """rec = inv(base) with d rec = -rec (d base) rec, so
    dL/d base = -rec^T G rec^T for G = dL/d rec."""

    @staticmethod
    def forward(ctx, abc, angles):
        a_np, g_np = abc.detach().cpu().numpy(), angles.detach().cpu().numpy()
        _, J = _basis_and_jacobian(a_np, g_np)
        # the same closed-form inverse the tables are built with
        rec = lattice_setup(tuple(a_np.tolist()), tuple(g_np.tolist()))[1]
        ctx.rec, ctx.J = rec, J
        return torch.as_tensor(rec, dtype=torch.float64)

    @staticmethod
    def backward(ctx, G):
        rec, J = ctx.rec, ctx.J
        dbase = -rec.T @ G.detach().cpu().numpy() @ rec.T
        g = np.tensordot(dbase, J, axes=([0, 1], [0, 1]))
        return (torch.as_tensor(g[:3], dtype=torch.float64),
                torch.as_tensor(g[3:], dtype=torch.float64))


def reciprocal_basis(abc, angles) -> torch.Tensor:
    """rec (3x3, columns a*, b*, c*; g = rec @ hkl) in the x || a, z || c*
    frame from a, b, c (A) and alpha, beta, gamma (degrees), as CPU float64
    tensors; differentiable (closed-form Jacobian of the lower-triangular
    direct basis and d inv(B) = -inv(B) dB inv(B))."""
    abc = torch.as_tensor(abc, dtype=torch.float64)
    angles = torch.as_tensor(angles, dtype=torch.float64)
    return _ReciprocalBasis.apply(abc, angles)


def leaf_parameters(phase: PhaseArrays, euler: np.ndarray, device=None,
                    requires_grad: bool = True) -> dict:
    """Leaf tensors of a refinement: Bunge angles (radians, on the device),
    a/b/c and alpha/beta/gamma (CPU), positions and B (on the device)."""
    dev = engine._device(device)

    def leaf(v, d):
        return torch.tensor(np.asarray(v, dtype=np.float64), dtype=torch.float64, device=d,
                            requires_grad=requires_grad)

    return {"euler": leaf(euler, dev), "abc": leaf(phase.abc, "cpu"),
            "angles": leaf(phase.angles, "cpu"), "xyz": leaf(phase.xyz, dev),
            "B": leaf(phase.B, dev)}


def forward_args(leaves: dict):
    """(Rt, rec, xyz, B) for RefinementModel from :func:`leaf_parameters`."""
    return (euler_to_rt(leaves["euler"]),
            reciprocal_basis(leaves["abc"], leaves["angles"]), leaves["xyz"], leaves["B"])


# ---------------------------------------------------------------------------
# the model
# ---------------------------------------------------------------------------


def _extreme_rows(hkl: np.ndarray) -> np.ndarray:
    """Rows of an hkl set among which max |rec hkl| lies for every basis rec:
    the vertices of the set's convex hull (|rec h|^2 is convex in h, so its
    maximum over a finite set is taken at an extreme point).  The largest |g|
    of a step then costs the hull's few thousand rows instead of the whole
    table (the 500-atom cell: 3,500 of 1.56 million rows, 0.2 ms instead of
    42 ms on the host, with the same value).  All rows when no hull can be
    built (fewer than four rows, rows in one plane, or no scipy)."""
    if hkl.shape[0] < 8:
        return hkl
    try:
        from scipy.spatial import ConvexHull
        return np.ascontiguousarray(hkl[ConvexHull(hkl).vertices])
    except Exception:  # qhull rejects degenerate sets; any failure keeps every row
        return hkl


# The line index follows the lattice, so a refinement step rebuilds it (three
# launches) before the spot kernel can use it.  That pays on tables whose lines
# average at least LINE_INDEX_LONG_ROWS rows whatever the batch (the 500-atom
# cell, 109 rows per line, in slices of 8,192 orientations), and on lines of
# DSX_LINE_SWEEP_MIN_ROWS or more only for large batches: the 24-atom
# refinement (14.1 rows per line, 4,096 orientations) measured 0.92 ms per
# fused step without the index and 1.01 ms with it, while the 40-atom library
# (14.6 rows per line, 200,000 orientations) gains 0.64 ms from it.
LINE_INDEX_LONG_ROWS = 16
LINE_INDEX_MIN_ORIENTATIONS = 16384


@dataclass
class _Tables:
    hkl: torch.Tensor   # int32 [N, 4]
    g4: torch.Tensor
    F2: torch.Tensor
    F: torch.Tensor
    xyz_fx: torch.Tensor
    B: torch.Tensor
    rec: np.ndarray
    g_max: float
    lines: Optional[torch.Tensor] = None  # line index for the spot kernel


class RefinementModel:
    """Fixed part of a differentiable simulation: one phase's elements and
    occupancies, the hkl set (from the initial lattice and
    ``reciprocal_radius``), the physical settings and the detector."""

    def __init__(self, phase: PhaseArrays, reciprocal_radius: float, settings: engine.SimSettings,
                 shape, sigma: float, calibration: float, clip_threshold: float = 1.0,
                 normalize: bool = True, in_plane_angle: float = 0.0, mirrored: bool = False,
                 direct_beam_position=None, scattering_params: Optional[str] = "lobato",
                 cap: Optional[int] = None, device=None):
        self.dev = engine._device(device)
        self.L = _lib.load(require_gpu=True)
        if not settings.approximate_precession and settings.precession_angle:
            raise NotImplementedError("gradients of the full precession integral are not implemented")
        self.settings = settings
        self.rr = float(reciprocal_radius)
        self.shape = (int(shape[0]), int(shape[1]))
        self.sigma = float(sigma)
        self.calibration = float(calibration)
        self.render = engine.render_params(self.shape, sigma, calibration, direct_beam_position,
                                           in_plane_angle, mirrored, normalize, fast=False,
                                           clip_threshold=clip_threshold)
        self.render_kw = dict(direct_beam_position=direct_beam_position,
                              in_plane_angle=in_plane_angle, mirrored=mirrored,
                              normalize=normalize, fast=False, clip_threshold=clip_threshold)
        self.scatter = scattering_params
        # hkl set of the initial lattice, on the device
        init = engine.build_phase_table(phase, reciprocal_radius, False, scattering_params,
                                        device=self.dev)
        self.hkl = init.hkl
        self.hkl_np = init.hkl[:, :3].cpu().numpy().astype(np.float64)
        # the rows that can hold the largest |g| for any lattice (see _extreme_rows)
        self.hkl_extreme = _extreme_rows(self.hkl_np)
        self.n_refl = init.n
        # the line index's rows are fixed by the hkl set, its |g|^2 bounds follow
        # the lattice; built per step only where the spot kernel uses it
        self.n_lines = init.n_lines
        self.use_lines = init.n_lines > 0 and init.n >= _lib.LINE_SWEEP_MIN_ROWS * init.n_lines
        # atoms sorted by element (the structure-factor kernel's order)
        elements = [element_of(s) for s in phase.symbols]
        uniq = sorted(set(elements))
        tab = table(scattering_params) if scattering_params is not None else {}
        coeff = np.zeros((max(len(uniq), 1), 10), dtype=np.float64)
        for i, el in enumerate(uniq):
            if el in tab:
                coeff[i] = tab[el]
        el_idx = np.array([uniq.index(e) for e in elements], dtype=np.int32)
        self.order = torch.from_numpy(np.argsort(el_idx, kind="stable").astype(np.int64))
        self.n_atoms = len(elements)
        self.n_elem = len(uniq)
        self.elem = torch.from_numpy(np.ascontiguousarray(el_idx[self.order.numpy()])).to(self.dev)
        self.occ = torch.from_numpy(np.ascontiguousarray(
            np.asarray(phase.occ, np.float32)[self.order.numpy()])).to(self.dev)
        self.coeff = torch.from_numpy(coeff).to(self.dev)
        self.order_dev = self.order.to(self.dev)
        inv = torch.argsort(self.order)
        self.inv_order_dev = inv.to(self.dev)
        # the atoms' gradient words in the caller's order (dxyz, 3 per atom,
        # then dB), read from the sorted accumulator by one gather
        na = self.n_atoms
        self.grad_perm_dev = torch.cat([(inv[:, None] * 3 + torch.arange(3)).reshape(-1),
                                        3 * na + inv]).to(self.dev)
        if cap is None:
            cap = engine.default_cap(init, settings, reciprocal_radius)
        self.cap = int(cap)

    # -- forward pieces ------------------------------------------------------

    def tables(self, rec: torch.Tensor, xyz: torch.Tensor, B: torch.Tensor) -> _Tables:
        """g vectors and |F|^2 of the fixed hkl set for the given parameters."""
        rec_np = np.ascontiguousarray(rec.detach().cpu().numpy(), dtype=np.float64).reshape(3, 3)
        stream = _lib.stream_handle(self.dev)
        na = self.n_atoms
        if na:
            # element order, uint32 turns and float B in one launch
            fx = torch.empty((na, 4), dtype=torch.int32, device=self.dev)
            Bs = torch.empty(na, dtype=torch.float32, device=self.dev)
            xs = xyz.detach().to(self.dev, torch.float64).contiguous()
            Bd = B.detach().to(self.dev, torch.float64).contiguous()
            _lib.check(self.L.dsx_stage_atoms(_lib.ptr(xs), _lib.ptr(Bd), _lib.ptr(self.order_dev),
                                              na, _lib.ptr(fx), _lib.ptr(Bs), stream),
                       "dsx_stage_atoms")
        else:
            fx = torch.zeros((1, 4), dtype=torch.int32, device=self.dev)
            Bs = torch.zeros(1, dtype=torch.float32, device=self.dev)
        N = self.n_refl
        g4 = torch.empty((N, 4), dtype=torch.float32, device=self.dev)
        F2 = torch.empty(N, dtype=torch.float32, device=self.dev)
        F = torch.empty((N, 2), dtype=torch.float32, device=self.dev)
        rec9 = (ctypes.c_double * 9)(*rec_np.reshape(-1).tolist())
        _lib.check(self.L.dsx_structure_factor(
            _lib.ptr(self.hkl), N, rec9, _lib.ptr(fx), _lib.ptr(self.elem), _lib.ptr(self.occ),
            _lib.ptr(Bs), self.n_atoms, _lib.ptr(self.coeff), self.n_elem,
            engine.SCATTER_IDS[self.scatter], _lib.ptr(g4), _lib.ptr(F2), _lib.ptr(F),
            stream), "dsx_structure_factor")
        g_max = float(np.sqrt(np.max(np.sum((self.hkl_extreme @ rec_np.T) ** 2, axis=1)))) if N else 0.0
        return _Tables(self.hkl, g4, F2, F, fx, Bs, rec_np, g_max)

    def _phase_table(self, t: _Tables) -> engine.PhaseTable:
        return engine.PhaseTable(t.hkl, t.g4, t.F2, t.F, self.n_refl, False, t.g_max, t.rec,
                                 None, "", t.lines)

    def spots(self, t: _Tables, Rt: torch.Tensor, batch: Optional[engine.SpotBatch] = None,
              check: bool = True) -> engine.SpotBatch:
        """Spot lists of every orientation (dsx_simulate)."""
        Rt = Rt.detach()
        long_lines = self.n_refl >= LINE_INDEX_LONG_ROWS * self.n_lines
        if t.lines is None and self.use_lines and (
                long_lines or int(Rt.shape[0]) >= LINE_INDEX_MIN_ORIENTATIONS):
            t.lines = engine.build_lines(self.hkl, t.g4, self.n_lines, _lib.stream_handle(self.dev))
        tab = self._phase_table(t)
        if batch is None:
            batch = engine.empty_batch(tab, int(Rt.shape[0]), self.cap, self.dev)
        engine.simulate_into(tab, Rt.contiguous(), self.settings, batch)
        if check:
            need = int(batch.ncand.max().item()) if batch.n else 0
            if need > batch.cap:
                self.cap = need + 64
                batch = engine.empty_batch(tab, batch.n, self.cap, self.dev)
                engine.simulate_into(tab, Rt.contiguous(), self.settings, batch)
        return batch

    def _sim_params(self, t: _Tables) -> _lib.SimParams:
        return self.settings.params(t.g_max, t.rec)

    def workspace(self, cap: int, n: int, gradients: bool = True) -> torch.Tensor:
        """Scratch of dsx_render_adjoint: per-spot sums between its render and
        chain-rule passes and, for spot lists longer than the render LDS
        holds, per-block strip bins and fixed-point pixels."""
        H, W = self.shape
        need = max(1, int(self.L.dsx_render_adjoint_workspace_ints(int(cap), int(n), H, W,
                                                                   int(gradients))))
        if getattr(self, "_work", None) is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.int32, device=self.dev)
        return self._work

    def adjoint(self, t: _Tables, Rt: torch.Tensor, batch: engine.SpotBatch, mode: int,
                target: Optional[torch.Tensor], want_grad: bool,
                pattern: Optional[torch.Tensor] = None, status: Optional[torch.Tensor] = None,
                atoms: bool = True, lattice: bool = True):
        """The render adjoint (+ the structure-factor adjoint when gradients of
        the atoms or the lattice are wanted).  ``target`` is [n, H, W] or one
        [H, W] pattern shared by every orientation.  With ``want_grad`` dL/dRt
        is always computed; ``atoms`` adds dL/dxyz and dL/dB, ``lattice``
        dL/drec.  Returns (loss[n], dRt, dF2, dxyz, dB, drec), None where not
        computed."""
        n = batch.n
        H, W = self.shape
        Rt = Rt.detach().contiguous()
        if pattern is not None and not pattern.is_contiguous():
            raise ValueError("pattern must be a contiguous float32 [n, H, W] tensor")
        stride = 0
        if target is not None:
            if target.dtype != torch.float32 or not target.is_contiguous():
                raise ValueError("target must be a contiguous float32 tensor")
            if tuple(target.shape) == (n, H, W):
                stride = H * W
            elif tuple(target.shape) != (H, W):
                raise ValueError(f"target must be [n, {H}, {W}] or [{H}, {W}]")
        loss = torch.empty(n, dtype=torch.float32, device=self.dev)
        dRt = dF2 = drec = None
        na = max(self.n_atoms, 1)
        table = want_grad and (atoms or lattice)  # gradients through |F|^2
        if want_grad:
            dRt = torch.empty((n, 9), dtype=torch.float64, device=self.dev)
        if table:
            # the accumulated gradients, zeroed by one fill: dF2 [n_refl] |
            # dxyz [na, 3] | dB [na] (float32), then drec [9] (float64) at an
            # 8-byte aligned offset
            nf = self.n_refl + 4 * na
            nf += nf & 1
            words = torch.zeros(nf + 18, dtype=torch.int32, device=self.dev)
            acc = words[:nf].view(torch.float32)
            dF2 = acc[:self.n_refl]
            drec = words[nf:].view(torch.float64)
        sync_status = status is None
        if status is None:
            status = torch.zeros(n, dtype=torch.int32, device=self.dev)
        sp = self._sim_params(t)
        stream = _lib.stream_handle(self.dev)
        # long spot lists hand n * cap per-spot sums (16 B each) from the
        # render to the chain rule: keep that under GRAD_SUMS_BYTES per call
        step = n
        if want_grad and batch.cap > int(self.L.dsx_render_max_spots()):
            step = max(1, min(n, GRAD_SUMS_BYTES // (16 * batch.cap)))
        work = self.workspace(batch.cap, step, want_grad)
        # per-pattern buffers are contiguous rows: a call's first pattern is
        # addressed by pointer offsets (no tensor views per call)
        cap = batch.cap
        rows = ((batch.idx, 4 * cap), (batch.spots, 16 * cap), (batch.count, 4), (Rt, 72),
                (target, 4 * stride), (pattern, 4 * H * W), (loss, 4), (dRt, 72), (status, 4))
        base = [(_lib.ptr(x), b if x is not None else 0) for x, b in rows]
        for o in range(0, n, step):
            e = min(n, o + step)
            a = [p + o * b if p else 0 for p, b in base]
            _lib.check(self.L.dsx_render_adjoint(
                a[0], a[1], a[2], cap, e - o, _lib.ptr(t.hkl), _lib.ptr(t.F2), self.n_refl,
                a[3], ctypes.byref(sp), ctypes.byref(self.render), mode, a[4], stride, a[5],
                a[6], a[7], _lib.ptr(dF2), _lib.ptr(drec if lattice else None), _lib.ptr(work),
                work.numel(), a[8], stream), "dsx_render_adjoint")
        if sync_status:
            engine.check_render_status(status)
        dxyz = dB = None
        if table:
            o = self.n_refl
            dxyz_s = acc[o:o + 3 * na].view(na, 3)
            dB_s = acc[o + 3 * na:o + 4 * na]
            rec9 = (ctypes.c_double * 9)(*t.rec.reshape(-1).tolist())
            _lib.check(self.L.dsx_structure_factor_adjoint(
                _lib.ptr(t.hkl), self.n_refl, rec9, _lib.ptr(t.xyz_fx), _lib.ptr(self.elem),
                _lib.ptr(self.occ), _lib.ptr(t.B), self.n_atoms, _lib.ptr(self.coeff), self.n_elem,
                engine.SCATTER_IDS[self.scatter], _lib.ptr(dF2), _lib.ptr(dxyz_s), _lib.ptr(dB_s),
                _lib.ptr(drec), stream), "dsx_structure_factor_adjoint")
            if atoms:
                # float64 in the caller's atom order, one gather and one conversion
                g = acc[o:].index_select(0, self.grad_perm_dev).to(torch.float64)
                dxyz = g[:3 * self.n_atoms].view(self.n_atoms, 3)
                dB = g[3 * self.n_atoms:]
            if not lattice:
                drec = None
        return loss, dRt, dF2, dxyz, dB, drec

    # -- public differentiable functions --------------------------------------

    def l2_loss(self, Rt: torch.Tensor, rec: torch.Tensor, xyz: torch.Tensor, B: torch.Tensor,
                target: torch.Tensor) -> torch.Tensor:
        """sum_patterns sum_pixels (P - T)^2 as a float64 CPU scalar (fetched
        with the step's one synchronisation; the backward then needs none)."""
        return _L2Loss.apply(Rt, rec, xyz, B, self, target)

    def patterns(self, Rt: torch.Tensor, rec: torch.Tensor, xyz: torch.Tensor,
                 B: torch.Tensor) -> torch.Tensor:
        """Rendered sub-pixel patterns [n, H, W] (float32), differentiable."""
        return _Patterns.apply(Rt, rec, xyz, B, self)


def _grads_like(ctx_inputs, dRt, drec, dxyz, dB):
    Rt, rec, xyz, B = ctx_inputs
    out = []
    for need, g, ref in ((True, dRt, Rt), (True, drec, rec), (True, dxyz, xyz), (True, dB, B)):
        out.append(None if g is None else g.to(ref.device, ref.dtype).reshape(ref.shape))
    return out


def _grow_if_needed(model: "RefinementModel", batch: engine.SpotBatch, status: torch.Tensor) -> bool:
    """One synchronisation per step: True when a spot list overflowed its
    capacity (the caller re-runs with the grown capacity); raises when a
    pattern could not be rendered."""
    if batch.n == 0:
        return False
    flags = torch.stack([batch.ncand.max(), status.max()]).cpu()
    if int(flags[1]) != 0:
        engine.check_render_status(status)
    if int(flags[0]) > batch.cap:
        model.cap = int(flags[0]) + 64
        return True
    return False


def _fetch_step(batch: engine.SpotBatch, status: torch.Tensor, loss: torch.Tensor,
                drec: Optional[torch.Tensor]) -> np.ndarray:
    """The step's one device->host copy (the per-pattern losses, candidate
    counts and statuses and dL/drec, gathered as 32-bit words by one
    concatenation), reduced on the host to [sum of the losses (float64),
    max ncand, max status, dL/drec (9, when given)] as float64."""
    n = batch.n
    parts = [loss.view(torch.int32), batch.ncand, status]
    if drec is not None:
        parts.append(drec.view(torch.int32))
    w = torch.cat(parts).cpu().numpy()
    head = [w[:n].view(np.float32).sum(dtype=np.float64), w[n:2 * n].max(), w[2 * n:3 * n].max()]
    tail = w[3 * n:].view(np.float64) if drec is not None else np.zeros(0)
    return np.concatenate([np.array(head, dtype=np.float64), tail])


class _L2Loss(torch.autograd.Function):
    """Forward and backward in one pass (the adjoint runs with the loss); the
    loss is returned as a CPU float64 scalar fetched with the step's single
    synchronisation, so the backward needs none."""

    @staticmethod
    def forward(ctx, Rt, rec, xyz, B, model: RefinementModel, target):
        need = ctx.needs_input_grad
        want = any(need[:4])
        atoms, lattice = need[2] or need[3], need[1]
        t = model.tables(rec, xyz, B)
        while True:
            batch = model.spots(t, Rt, check=False)
            if batch.n == 0:
                host, dRt, dxyz, dB = np.zeros(12), Rt.detach().new_zeros(0, 9), None, None
                if want:
                    dxyz = torch.zeros_like(xyz.detach(), dtype=torch.float32)
                    dB = torch.zeros_like(B.detach(), dtype=torch.float32)
                break
            status = torch.empty(batch.n, dtype=torch.int32, device=model.dev)  # every entry written
            loss, dRt, _, dxyz, dB, drec = model.adjoint(t, Rt.detach(), batch, _lib.ADJ_L2, target,
                                                         want, status=status, atoms=atoms,
                                                         lattice=lattice)
            host = _fetch_step(batch, status, loss, drec)
            if int(host[2]) != 0:
                engine.check_render_status(status)
            if int(host[1]) <= batch.cap:
                break
            model.cap = int(host[1]) + 64
        if want:
            drec_h = torch.from_numpy(np.array(host[3:12])) if lattice else None
            g = _grads_like((Rt, rec, xyz, B), dRt, drec_h, dxyz, dB)
            ctx.nones = [x is None for x in g]
            ctx.save_for_backward(*[x for x in g if x is not None])
        return torch.tensor(float(host[0]), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gl):
        saved = iter(ctx.saved_tensors)
        g = [None if none else next(saved) for none in ctx.nones]
        s = float(gl)  # a CPU scalar: no synchronisation
        if s != 1.0:
            g = [None if x is None else x * s for x in g]
        return (*g, None, None)


class _Patterns(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Rt, rec, xyz, B, model: RefinementModel):
        t = model.tables(rec, xyz, B)
        while True:
            batch = model.spots(t, Rt, check=False)
            status = torch.zeros(batch.n, dtype=torch.int32, device=model.dev)
            out = engine.render_spots(batch.spots, batch.count, batch.cap, model.shape, model.sigma,
                                      model.calibration, status=status, **model.render_kw)
            if not _grow_if_needed(model, batch, status):
                break
        ctx.model = model
        ctx.save_for_backward(Rt.detach(), rec.detach(), xyz.detach(), B.detach())
        return out

    @staticmethod
    def backward(ctx, gP):
        model: RefinementModel = ctx.model
        Rt, rec, xyz, B = ctx.saved_tensors
        # recompute the tables and spot lists instead of keeping them alive
        t = model.tables(rec, xyz, B)
        while True:
            batch = model.spots(t, Rt, check=False)
            status = torch.zeros(batch.n, dtype=torch.int32, device=model.dev)
            _, dRt, _, dxyz, dB, drec = model.adjoint(t, Rt, batch, _lib.ADJ_PATTERN_GRAD,
                                                      gP.contiguous().to(torch.float32), True,
                                                      status=status)
            if not _grow_if_needed(model, batch, status):
                break
        g = _grads_like((Rt, rec, xyz, B), dRt, drec, dxyz, dB)
        return (*g, None)
