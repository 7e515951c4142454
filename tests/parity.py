"""Shared fixtures and parity checks between the HIP path and the FP64 oracle.

Parity rules (stated once, used by every GPU test):
* hkl sets: bit-exact, same order.
* survivor index lists per orientation: identical, except reflections whose
  FP64 excitation error lies within 1e-6 of the cut (|s| vs max_excitation_error)
  may differ; forbidden-only orientations keep every intersecting reflection
  with I = 0 on both sides.
* intensities: |I_gpu - I_fp64| <= 1e-5 * max(I_fp64) per orientation.
* rendered patterns (normalised): |P_gpu - P_fp64| <= 1e-5 * max(P_fp64);
  in fast mode a spot whose FP64 pixel coordinate lies within 1e-4 px of an
  integer may land on the neighbouring pixel, and the oracle is then re-rendered
  with the GPU's pixel for that spot ("compare with the flipped pixel").
"""

from __future__ import annotations

import numpy as np

from oracle import diffsims_oracle as O

TOL_I = 1e-5
TOL_P = 1e-5
CUT_EPS = 1e-6
PIX_EPS = 1e-4

FCC_POS = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])


def si_positions():
    pos = []
    for c in [[0, 0, 0], [0.5, 0, 0.5], [0, 0.5, 0.5], [0.5, 0.5, 0]]:
        pos.append(c)
        pos.append([c[0] + 0.25, c[1] + 0.25, c[2] + 0.25])
    return np.array(pos, dtype=np.float64)


def al_c0():
    """FCC Al a=4.05, 4 atoms, B = 0.5 (the 500-orientation small case)."""
    return O.PhaseSpec((4.05,) * 3, (90.0,) * 3, ["Al"] * 4, FCC_POS.copy(), np.ones(4),
                       np.full(4, 0.5))


def ortho40(seed=20261203):
    """Orthorhombic cell, a,b,c ~ U[4,12], 40 atoms, Z in {6,8,14,26},
    B ~ U[0.3,1] (the large single-phase template library case)."""
    rng = np.random.default_rng(seed)
    abc = tuple(rng.uniform(4, 12, 3))
    sym = {6: "C", 8: "O", 14: "Si", 26: "Fe"}
    Z = rng.choice([6, 8, 14, 26], 40)
    return O.PhaseSpec(abc, (90.0,) * 3, [sym[int(z)] for z in Z], rng.uniform(0, 1, (40, 3)),
                       np.ones(40), rng.uniform(0.3, 1.0, 40))


def triclinic(n_atoms, seed, edges=(4, 9), angles=(80, 110)):
    rng = np.random.default_rng(seed)
    abc = tuple(rng.uniform(*edges, 3))
    ang = tuple(rng.uniform(*angles, 3))
    syms = ["C", "N", "O", "Si", "Fe", "Cu", "Mg", "Al"]
    return O.PhaseSpec(abc, ang, [syms[i] for i in rng.integers(0, len(syms), n_atoms)],
                       rng.uniform(0, 1, (n_atoms, 3)), rng.uniform(0.5, 1.0, n_atoms),
                       rng.uniform(0.2, 1.5, n_atoms))


def to_arrays(ph: O.PhaseSpec, name=""):
    from diffsims_amd.crystallography import PhaseArrays
    return PhaseArrays(tuple(ph.abc), tuple(ph.angles), list(ph.symbols), np.asarray(ph.xyz),
                       np.asarray(ph.occ), np.asarray(ph.B), name)


def excitation_errors(otab: O.PhaseTables, M: np.ndarray, k: float):
    g_lab = otab.g @ M
    r = np.sqrt(np.sum(g_lab[:, :2] ** 2, axis=1))
    return (-np.sqrt(k ** 2 - r ** 2) + k) - g_lab[:, 2]


def compare_spot_lists(gpu_idx, gpu_I, ref: O.Spots, s_all, me, stats=None):
    """Check one orientation; returns the number of carved-out reflections."""
    g = np.asarray(gpu_idx)
    r = np.asarray(ref.idx)
    if np.array_equal(g, r):
        if r.size:
            scale = max(float(np.max(ref.intensity)), 1e-300)
            err = float(np.max(np.abs(np.asarray(gpu_I, np.float64) - ref.intensity))) / scale
            if ref.max_intensity == 0.0:
                err = float(np.max(np.abs(gpu_I)))
            assert err <= TOL_I, f"intensity error {err:.3g} > {TOL_I}"
            if stats is not None:
                stats["max_rel_err"] = max(stats.get("max_rel_err", 0.0), err)
        return 0
    diff = np.setxor1d(g, r)
    near = np.abs(np.abs(s_all[diff]) - me) < CUT_EPS
    assert near.all(), f"index sets differ away from the cut: {diff[~near][:10]}"
    return int(diff.size)


def flipped_reference_pattern(ref: O.Spots, gpu_xy, shape, sigma, cal):
    """Oracle fast render with FP64 pixel coordinates, but for spots within
    PIX_EPS of a pixel boundary the pixel the GPU used (gpu_xy: the GPU's
    lab-frame x, y for the same spots)."""
    H, W = shape
    cx, cy = W // 2, H // 2
    X, Y = O.pixel_coords(ref.xyz[:, :2], shape, cal, (cx, cy))
    Xg = np.asarray(gpu_xy[:, 0], np.float32) / np.float32(cal) + np.float32(cx)
    Yg = np.asarray(gpu_xy[:, 1], np.float32) / np.float32(cal) + np.float32(cy)
    for arr, arrg in ((X, Xg), (Y, Yg)):
        fr = arr - np.floor(arr)
        edge = (fr < PIX_EPS) | (fr > 1 - PIX_EPS)
        arr[edge] = np.asarray(arrg, np.float64)[edge]
    return render_from_pixels(X, Y, ref.intensity, shape, sigma)


def render_from_pixels(X, Y, I, shape, sigma, normalize=True):
    """The reference's fast render from given pixel coordinates: in-frame test,
    truncation, last write wins, scipy gaussian_filter, normalisation
    (simulation2d.py:416-442, detector_functions.py:293-300)."""
    from scipy import ndimage as ndi

    H, W = shape
    inf = (X >= 0) & (X < W) & (Y >= 0) & (Y < H)
    I = np.asarray(I, np.float64)[inf]
    if I.size == 0:
        return np.zeros(shape)
    out = np.zeros(shape)
    out[Y[inf].astype(int), X[inf].astype(int)] = I
    out = ndi.gaussian_filter(out, sigma)
    if normalize and out.max() > 0:
        out = out / out.max()
    return out
