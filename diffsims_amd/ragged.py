"""Ragged multi-phase batching: many phases with uneven reflection and atom
counts, each with its own orientation set, simulated in ONE persistent launch
and rendered in one launch per render method, the whole forward captured in a
hipGraph (torch.cuda.CUDAGraph) and replayed.

This is the batched form of calculate_diffraction2d's per-phase loop
(diffsims/generators/simulation_generator.py:198-242) followed by
get_diffraction_pattern for every (phase, orientation)
(diffsims/simulations/simulation2d.py:357-442); the results equal the
single-phase path phase by phase.

Layout in HBM (one pool per quantity, phases back to back):
  Rt     float64 [n_total, 9]          orientation rows of M^T
  idx    int32   [sum_p n_p cap_p]     per phase [n_p][cap_p]
  spots  float32 [sum_p n_p cap_p, 4]
  count/ncand/maxI [n_total]           global pattern order (phase-major)
  out    float32 [n_total, H, W]
Work items of the simulate launch are (phase, first orientation) tiles,
ordered by decreasing cost (reflections per tile) so the persistent waves
finish together.  Launch widths and the per-phase render method come from
:class:`diffsims_amd.autotune.Autotuner`.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib, engine
from .autotune import Autotuner
from .crystallography import PhaseArrays

__all__ = ["RaggedLibrary"]


@dataclass
class _PhaseSlot:
    table: engine.PhaseTable
    n: int            # orientations
    first: int        # first global pattern id
    cap: int
    pool_off: int     # first spot of this phase in the pool
    method: int = _lib.RENDER_SPLAT


class RaggedLibrary:
    """A multi-phase template library on one GPU.

    ``phases``/``rotations``/``reciprocal_radius`` are per-phase lists
    (rotations as (n, 3, 3) orix ``to_matrix()`` arrays).  ``run()`` launches
    the whole forward on the current stream; ``capture()`` records it in a
    hipGraph and ``replay()`` re-launches it with no host work.
    """

    def __init__(self, phases: Sequence[PhaseArrays], rotations: Sequence[np.ndarray],
                 reciprocal_radius: Sequence[float], settings: engine.SimSettings, shape,
                 sigma: float, calibration: float, with_direct_beam: bool = False,
                 scattering_params: Optional[str] = "lobato", device=None,
                 tuner: Optional[Autotuner] = None, cap_margin: float = 1.3):
        if not (len(phases) == len(rotations) == len(reciprocal_radius)) or not phases:
            raise ValueError("phases, rotations and reciprocal_radius must be non-empty lists "
                             "of the same length")
        self.dev = engine._device(device)
        self.L = _lib.load(require_gpu=True)
        self.settings = settings
        self.shape = (int(shape[0]), int(shape[1]))
        self.sigma = float(sigma)
        self.calibration = float(calibration)
        self.render = engine.render_params(self.shape, sigma, calibration)
        self.tuner = tuner if tuner is not None else Autotuner()
        self.tile = int(self.L.dsx_simulate_tile_size())
        self.graph = None

        tables = []
        for pa, rr in zip(phases, reciprocal_radius):
            tables.append(engine.build_phase_table(pa, rr, with_direct_beam, scattering_params,
                                                   device=self.dev))
        self.slots: List[_PhaseSlot] = []
        first = 0
        for tab, R, rr in zip(tables, rotations, reciprocal_radius):
            n = int(np.asarray(R).reshape(-1, 3, 3).shape[0])
            exp = engine.expected_spots(tab, settings, rr)
            cap = max(16, min(max(tab.n, 1), int(math.ceil(cap_margin * exp)) + 64))
            self.slots.append(_PhaseSlot(tab, n, first, cap, 0))
            first += n
        self.n_total = first
        self.Rt = engine.rotation_matrices_to_device(
            np.concatenate([np.asarray(R, dtype=np.float64).reshape(-1, 3, 3) for R in rotations]),
            self.dev)
        self.count = torch.zeros(self.n_total, dtype=torch.int32, device=self.dev)
        self.ncand = torch.zeros(self.n_total, dtype=torch.int32, device=self.dev)
        self.maxI = torch.zeros(self.n_total, dtype=torch.float32, device=self.dev)
        self.status = torch.zeros(self.n_total, dtype=torch.int32, device=self.dev)
        self.out = torch.empty((self.n_total,) + self.shape, dtype=torch.float32, device=self.dev)
        self._allocate_pool()
        self._build_items()
        self.tune()

    # -- layout ---------------------------------------------------------------

    def _allocate_pool(self):
        off = 0
        for s in self.slots:
            s.pool_off = off
            off += s.n * s.cap
        self.pool_size = off
        self.idx = torch.empty(off, dtype=torch.int32, device=self.dev)
        self.spots = torch.empty((off, 4), dtype=torch.float32, device=self.dev)
        spot_off = np.empty(self.n_total, dtype=np.int64)
        capv = np.empty(self.n_total, dtype=np.int32)
        for s in self.slots:
            spot_off[s.first:s.first + s.n] = s.pool_off + s.cap * np.arange(s.n, dtype=np.int64)
            capv[s.first:s.first + s.n] = s.cap
        self.spot_off = torch.from_numpy(spot_off).to(self.dev)
        self.capv = torch.from_numpy(capv).to(self.dev)
        descs = (_lib.PhaseDesc * len(self.slots))()
        rt_row = self.Rt.element_size() * 9
        for d, s in zip(descs, self.slots):
            t = s.table
            d.hkl, d.g4, d.F2 = t.hkl.data_ptr(), t.g4.data_ptr(), t.F2.data_ptr()
            d.n_refl, d.cap = t.n, s.cap
            for i, v in enumerate(np.asarray(t.rec, dtype=np.float64).reshape(-1)):
                d.rec[i] = float(v)
            d.g_max = t.g_max
            d.Rt = self.Rt.data_ptr() + rt_row * s.first
            d.n_orient = s.n
            d.idx = self.idx.data_ptr() + 4 * s.pool_off
            d.spots = self.spots.data_ptr() + 16 * s.pool_off
            d.count = self.count.data_ptr() + 4 * s.first
            d.ncand = self.ncand.data_ptr() + 4 * s.first
            d.maxI = self.maxI.data_ptr() + 4 * s.first
            d.n_lines = t.n_lines
            d.lines = t.lines.data_ptr() if t.lines is not None else None
        raw = np.frombuffer(bytes(descs), dtype=np.uint8)
        self.descs = torch.from_numpy(raw.copy()).to(self.dev)
        self.graph = None

    def _build_items(self):
        """(phase, first orientation) tiles, most expensive phases first."""
        order = sorted(range(len(self.slots)), key=lambda i: -self.slots[i].table.n)
        items = [np.stack([np.full(len(o0), i, dtype=np.int32), o0], 1)
                 for i in order
                 for o0 in [np.arange(0, self.slots[i].n, self.tile, dtype=np.int32)]]
        self.items = torch.from_numpy(np.ascontiguousarray(np.concatenate(items))).to(self.dev)

    def _lists(self):
        by = {}
        for s in self.slots:
            by.setdefault(s.method, []).append(np.arange(s.first, s.first + s.n, dtype=np.int32))
        self.render_lists = {m: torch.from_numpy(np.concatenate(v)).to(self.dev)
                             for m, v in by.items()}

    # -- launches -------------------------------------------------------------

    def _simulate(self, blocks_per_cu: int):
        params = self.settings.params(0.0, np.eye(3))
        _lib.check(self.L.dsx_simulate_ragged(
            _lib.ptr(self.descs), len(self.slots), _lib.ptr(self.items), int(self.items.shape[0]),
            ctypes.byref(params), int(blocks_per_cu), _lib.stream_handle(self.dev)),
            "dsx_simulate_ragged")

    def _render(self, method: int, lst: torch.Tensor, blocks_per_cu: int):
        _lib.check(self.L.dsx_render_fast_ragged(
            _lib.ptr(self.spots), _lib.ptr(self.spot_off), _lib.ptr(self.capv),
            _lib.ptr(self.count), _lib.ptr(lst), int(lst.shape[0]), ctypes.byref(self.render),
            int(method), int(blocks_per_cu), _lib.ptr(self.out), _lib.ptr(self.status),
            _lib.stream_handle(self.dev)), "dsx_render_fast_ragged")

    def run(self):
        """Launch the whole forward (no host synchronisation)."""
        self._simulate(self.sim_bpc)
        for m, lst in self.render_lists.items():
            self._render(m, lst, self.render_bpc[m])

    def check(self):
        """Synchronise; grow capacities and re-run if a spot list overflowed;
        raise if a pattern could not be rendered."""
        ncand = self.ncand.cpu().numpy()
        grown = False
        for s in self.slots:
            need = int(ncand[s.first:s.first + s.n].max()) if s.n else 0
            if need > s.cap:
                s.cap = need + 16
                grown = True
        if grown:
            self._allocate_pool()
            self.run()
            self.check()
            return
        st = self.status.cpu().numpy()
        if st.any():
            # a phase the tuner sampled as splat-safe has a pattern with more
            # distinct in-frame spots than the splat renderer holds: render
            # that phase densely
            H, W = self.shape
            if H * W > int(self.L.dsx_render_dense_max_pixels()):
                bad = int(np.nonzero(st)[0][0])
                raise RuntimeError(f"pattern {bad} overflowed the splat renderer")
            dense = self._dense_method()
            for s in self.slots:
                if st[s.first:s.first + s.n].any():
                    s.method = dense
            self.render_bpc.setdefault(dense, 2 if dense == _lib.RENDER_DENSE_INPLACE else 1)
            self._lists()
            self.graph = None
            self.status.zero_()
            self.run()
            self.check()

    def _dense_method(self) -> int:
        """The dense renderer for this shape: the in-place kernel where it
        applies (same bits, two patterns per CU), else the two-buffer one."""
        H, W = self.shape
        return (_lib.RENDER_DENSE_INPLACE
                if self.L.dsx_render_dense_inplace_fits(H, W, self.render.radius)
                else _lib.RENDER_DENSE)

    # -- hipGraph ---------------------------------------------------------------

    def capture(self):
        """Record run() in a hipGraph on a side stream (torch.cuda.CUDAGraph)."""
        self.run()
        self.check()
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    # -- autotuning -------------------------------------------------------------

    def tune(self):
        """Pick the simulate launch width, and per phase bucket the render
        method and width, by timing candidates on this device."""
        # one-wave blocks per CU of the spot kernel's launch: the work items
        # are in decreasing cost, so a launch far wider than what stays
        # resident (at most one item per wave) lets the dispatcher hand the
        # costly items out first and fill in behind them (64-phase batch:
        # 21.6 ms at 64 waves per CU, 18.2 ms at one item per wave)
        self.sim_bpc = self.tuner.best(
            ("simulate_ragged", len(self.slots), int(self.items.shape[0])),
            [32, 64, 128, 256, 512, 1024, 4096], lambda bpc: self._simulate(bpc))
        torch.cuda.synchronize(self.dev)
        self.check_spots_only()
        max_pix = int(self.L.dsx_render_dense_max_pixels())
        H, W = self.shape
        self.render_bpc = {}
        buckets = {}
        for i, s in enumerate(self.slots):
            key = ("render", H, W, self.render.radius, int(math.log2(max(s.cap, 1))))
            buckets.setdefault(key, []).append(i)
        for key, members in buckets.items():
            rep = max(members, key=lambda i: self.slots[i].n)
            s = self.slots[rep]
            n_sample = min(s.n, 2048)
            lst = torch.arange(s.first, s.first + n_sample, dtype=torch.int32, device=self.dev)
            cands = [(_lib.RENDER_SPLAT, b) for b in (1, 2, 3)]
            if H * W <= max_pix:
                cands += [(_lib.RENDER_DENSE, 1)]
            if self.L.dsx_render_dense_inplace_fits(H, W, self.render.radius):
                cands += [(_lib.RENDER_DENSE_INPLACE, b) for b in (1, 2)]

            def feasible(c, lst=lst):
                self.status.zero_()
                self._render(c[0], lst, c[1])
                return int(self.status.max().item()) == 0

            m, b = self.tuner.best(key, cands, lambda c, lst=lst: self._render(c[0], lst, c[1]),
                                   feasible=feasible)
            for i in members:
                self.slots[i].method = m
            self.render_bpc[m] = self.render_bpc.get(m, b)
        self.status.zero_()
        self._lists()

    def check_spots_only(self):
        ncand = self.ncand.cpu().numpy()
        grown = False
        for s in self.slots:
            need = int(ncand[s.first:s.first + s.n].max()) if s.n else 0
            if need > s.cap:
                s.cap = need + 16
                grown = True
        if grown:
            self._allocate_pool()
            self._simulate(self.sim_bpc)
            torch.cuda.synchronize(self.dev)

    # -- results ----------------------------------------------------------------

    def phase_patterns(self, i: int) -> torch.Tensor:
        s = self.slots[i]
        return self.out[s.first:s.first + s.n]

    def phase_spots(self, i: int):
        """(idx [n, cap], spots [n, cap, 4], count [n]) views of phase i."""
        s = self.slots[i]
        sl = slice(s.pool_off, s.pool_off + s.n * s.cap)
        return (self.idx[sl].view(s.n, s.cap), self.spots[sl].view(s.n, s.cap, 4),
                self.count[s.first:s.first + s.n])
