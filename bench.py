"""Throughput of the kinematical pattern-simulation path on one MI355X.

One step = simulate + render the whole orthorhombic 40-atom template library
(200,000 orientations, 256x256 FP32 patterns) with every input resident in HBM:
the per-orientation kernel (excitation error, shape factor, cut, compaction)
followed by the renderer, writing all patterns to one [n, 256, 256] tensor.

Further legs (each with its own roofline figure, CPU baseline and FP64 error):
the plain path at its small size (FCC Al, 500 orientations), the fused
forward+backward refinement, the 64-phase ragged batch and the 500-atom
precession cell.

Prints one JSON line (see --help).  The CPU baselines are the FP64
restatements of the reference's loop timed on bounded samples on this host,
one thread, with the CPU model and core count stated.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "patterns/s: batched fwd & fwd+bwd vs plain fwd, 1 MI355X, rel err ≤1e-5 vs FP64"
HBM_PEAK_GBS = 8000.0


PROFILE = os.path.join(ROOT, "profiles", "render_bands_ortho40_library.md")
REFINE_PROFILE = os.path.join(ROOT, "profiles", "render_adjoint_mono24_refinement_counters.md")
RAGGED_PROFILE = os.path.join(ROOT, "profiles", "render_dense_inplace_ragged_64phase.md")
SMALL_PROFILE = os.path.join(ROOT, "profiles", "render_fast_fcc_al_small_kernel_trace.md")
LARGE_PROFILE = os.path.join(ROOT, "profiles", "render_adjoint_large_tri500_counters.md")


def measured_traffic(render_bytes, profile=PROFILE):
    """HBM bytes per launch from a FETCH_SIZE/WRITE_SIZE measurement committed
    in profiles/ (by default the library's render stage: plan kernel, band
    writer and the splat kernel for the patterns the plan lists, taken on the
    200,000-orientation library), scaled to this run's algorithmic bytes;
    None when the summary is absent."""
    import re
    try:
        text = open(profile).read()
        t = float(re.search(r"traffic_bytes_per_launch: (\d+)", text).group(1))
        a = float(re.search(r"algorithmic_bytes_per_launch: (\d+)", text).group(1))
    except (OSError, AttributeError):
        return None
    return round(t / a * render_bytes)


def host_cpu():
    """CPU model and the number of logical CPUs this process sees."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "os_cpu_count": os.cpu_count()}


def cpu_baseline_entry(value, sample, seconds, kind="port", unit="patterns/s"):
    """The cpu_baseline object: single-thread FP64 timing on this host."""
    return {"value": round(value, 3), "unit": unit, "cores": 1, "kind": kind,
            "sample": f"{sample} ({seconds:.1f} s, 1 thread)", **host_cpu()}


def spot_accuracy(otab, M, p, idx, spots, count, pats, shape, sigma, calibration, fast=True,
                  clip=1.0):
    """Max error of GPU spot lists and patterns against the FP64 oracle for
    the given orientations (rows of ``idx``/``spots``/``pats``, host arrays):
    survivor sets equal except rows within 1e-6 of the cut (counted; the
    pattern is then compared on the GPU's set), intensities and patterns
    relative to each pattern's maximum; fast renders with the flipped-pixel
    rule, sub-pixel renders against the clip-box envelope.  Raises on a
    difference away from the cut."""
    from oracle import diffsims_oracle as O
    from tests import parity as P

    k = 1 / O.electron_wavelength(p.kv)
    stats = {}
    worst = 0.0
    for j in range(len(M)):
        c = int(count[j])
        ref, _ = P.check_orientation(otab, M[j], k, p, idx[j, :c], spots[j, :c, 3], stats)
        if fast:
            rp = P.flipped_reference_pattern(ref, spots[j, :c, :2], shape, sigma, calibration)
            err = P.pattern_error(pats[j], rp)
        else:
            nom, lo, hi = P.subpixel_envelope(ref.xyz[:, :2], ref.intensity, shape, sigma,
                                              calibration, clip)
            err = P.envelope_error(pats[j], nom, lo, hi)
        worst = max(worst, err)
    return {"max_rel_err_vs_fp64": worst, "max_intensity_rel_err": stats.get("max_rel_err", 0.0),
            "patterns_compared": len(M), "carved_orientations": stats.get("carved_orientations", 0),
            "carved_rows_at_cut": stats.get("carved_rows", 0)}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--orientations", type=int, default=200_000)
    ap.add_argument("--cpu-sample", type=int, default=8000,
                    help="orientations in the CPU-baseline sample (0 disables it)")
    ap.add_argument("--refine-steps", type=int, default=10,
                    help="timed steps of the fused forward+backward leg (0 disables it)")
    ap.add_argument("--ragged-steps", type=int, default=3,
                    help="timed hipGraph replays of the 64-phase ragged batch (0 disables it)")
    ap.add_argument("--ragged-scale", type=float, default=1.0,
                    help="fraction of the ragged batch's orientations to use")
    ap.add_argument("--accuracy-patterns", type=int, default=1024,
                    help="library orientations compared with the FP64 oracle (0 disables it)")
    ap.add_argument("--plain-steps", type=int, default=100,
                    help="timed steps of the 500-orientation FCC Al plain path (0 disables it)")
    ap.add_argument("--large-chunks", type=int, default=2,
                    help="65,536-orientation chunks of the 500-atom precession case to run "
                         "fused forward+backward (16 = the full 10^6 orientations; 0 disables)")
    ap.add_argument("--grid-resolution", type=float, default=0.1,
                    help="degrees of the triclinic beam-direction grid leg (0 disables it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference "
                         "kernel for scaling the HBM byte counters of a profiler) and exit")
    return ap.parse_args()


def cpu_baseline(wl, n_sample):
    """Reference-shaped FP64 loop (the oracle: per orientation rotate, cut, F of
    the intersecting rows, shape factor, relative cut, then the
    gaussian_filter render) on the first n_sample orientations, one thread."""
    from threadpoolctl import threadpool_limits

    from oracle import diffsims_oracle as O
    from tests import parity as P

    pa = wl.phases[0]
    p = P.oracle_params(wl.settings, wl.reciprocal_radius[0], wl.with_direct_beam)
    M = wl.rotations[0][:n_sample]
    render = {"shape": wl.shape, "sigma": wl.sigma, "calibration": wl.calibration}
    with threadpool_limits(limits=1):
        O.simulate_library(P.phase_spec(pa), M[:8], p, render)          # warm caches
        t0 = time.perf_counter()
        O.simulate_library(P.phase_spec(pa), M, p, render)
        dt = time.perf_counter() - t0
    return n_sample / dt, dt


def library_accuracy(wl, batch, pats, n_compare):
    """FP64 error of n_compare orientations spread over the whole batch."""
    from oracle import diffsims_oracle as O
    from tests import parity as P

    p = P.oracle_params(wl.settings, wl.reciprocal_radius[0], wl.with_direct_beam)
    otab = O.phase_tables(P.phase_spec(wl.phases[0]), p)
    n = batch.n
    sel = np.unique(np.linspace(0, n - 1, min(n, n_compare)).round().astype(np.int64))
    t = __import__("torch").as_tensor(sel, device=batch.idx.device)
    return spot_accuracy(otab, wl.rotations[0][sel], p, batch.idx[t].cpu().numpy(),
                         batch.spots[t].cpu().numpy(), batch.count[t].cpu().numpy(),
                         pats[t].cpu().numpy(), wl.shape, wl.sigma, wl.calibration)


def plain_leg(dev, steps, warmup):
    """The plain path at its small CPU-runnable size: FCC Al, 500 random
    orientations, 144x144, simulate + render on the GPU against the same
    500 orientations through the oracle's reference-shaped loop on one core;
    every pattern compared with FP64."""
    import torch
    from threadpoolctl import threadpool_limits

    from diffsims_amd import engine, workloads
    from oracle import diffsims_oracle as O
    from tests import parity as P

    wl = workloads.small_fcc_al(500)
    rr = wl.reciprocal_radius[0]
    tab = engine.build_phase_table(wl.phases[0], rr, wl.with_direct_beam, "lobato", device=dev)
    Rt = engine.rotation_matrices_to_device(wl.rotations[0], dev)
    n = Rt.shape[0]
    probe = engine.simulate(tab, Rt, wl.settings, rr)
    batch = engine.empty_batch(tab, n, probe.cap, dev)
    H, W = wl.shape
    out = torch.empty((n, H, W), dtype=torch.float32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        engine.simulate_into(tab, Rt, wl.settings, batch)
        engine.render_spots(batch.spots, batch.count, batch.cap, wl.shape, wl.sigma,
                            wl.calibration, out=out, status=status)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record(stream)
        engine.simulate_into(tab, Rt, wl.settings, batch)
        e1.record(stream)
        engine.render_spots(batch.spots, batch.count, batch.cap, wl.shape, wl.sigma,
                            wl.calibration, out=out, status=status)
        e2.record(stream)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    engine.check_render_status(status)
    assert int(batch.ncand.max()) <= batch.cap
    ren_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    sim_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    render_bytes = n * (H * W * 4 + 4) + float(batch.count.float().sum()) * 16
    achieved = render_bytes / (ren_ms * 1e-3) / 1e9
    p = P.oracle_params(wl.settings, rr, wl.with_direct_beam)
    render = {"shape": wl.shape, "sigma": wl.sigma, "calibration": wl.calibration}
    with threadpool_limits(limits=1):
        O.simulate_library(P.phase_spec(wl.phases[0]), wl.rotations[0][:8], p, render)
        c0 = time.perf_counter()
        O.simulate_library(P.phase_spec(wl.phases[0]), wl.rotations[0], p, render)
        cdt = time.perf_counter() - c0
    otab = O.phase_tables(P.phase_spec(wl.phases[0]), p)
    acc = spot_accuracy(otab, wl.rotations[0], p, batch.idx.cpu().numpy(),
                        batch.spots.cpu().numpy(), batch.count.cpu().numpy(), out.cpu().numpy(),
                        wl.shape, wl.sigma, wl.calibration)
    return {
        "workload": "FCC Al a=4.05 A, 4 atoms, B=0.5; 200 kV, r=1.0 1/A, max_exc 0.01, linear, "
                    "no direct beam; 500 random orientations; 144x144 f32 at 0.01 1/A/px, sigma 1 "
                    "px, fast render, normalised",
        "n_patterns": n, "n_hkl": tab.n,
        "patterns_per_s": round(n / dt, 1), "ms_per_step": round(1e3 * dt, 4),
        "kernel_ms": {"simulate": round(sim_ms, 4), "render": round(ren_ms, 4)},
        "roofline": {"bound": "hbm", "kernel": f"render stage ({engine.fast_renderer(n, H, W)})",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     # the counters measured the splat kernel, which renders this batch
                     "traffic": (measured_traffic(render_bytes, SMALL_PROFILE)
                                 if engine.fast_renderer(n, H, W) == "dsx_render_fast" else None)},
        "cpu_baseline": cpu_baseline_entry(n / cdt, "all 500 orientations, FP64 NumPy "
                                           "restatement of the reference loop incl. render", cdt),
        "accuracy": acc,
    }


def refinement_leg(dev, steps, warmup):
    """Fused forward+backward of the L2 refinement loss (24-atom monoclinic
    cell, 4,096 256x256 sub-pixel targets) against the plain forward of the
    same patterns; both in patterns/s.  Also: the render adjoint timed alone
    with HIP events (roofline: the targets it reads), the FP64 error of the
    forward patterns on a sample, and the FP64 torch restatement of the same
    fused step on one core (CPU baseline)."""
    import torch
    from threadpoolctl import threadpool_limits

    from diffsims_amd import _lib
    from diffsims_amd import differentiable as D
    from diffsims_amd import workloads
    from oracle import diffsims_grad as G
    from oracle import diffsims_oracle as O
    from tests import parity as P

    rw = workloads.refinement_monoclinic()
    model = D.RefinementModel(rw.phase, rw.reciprocal_radius, rw.settings, rw.shape, rw.sigma,
                              rw.calibration, clip_threshold=1.0, device=dev)
    with torch.no_grad():
        truth = D.leaf_parameters(rw.truth_phase, rw.truth_euler, dev, requires_grad=False)
        T = model.patterns(*D.forward_args(truth))
        mx = T.amax(dim=(1, 2), keepdim=True)
        gen = torch.Generator(device=dev).manual_seed(rw.seed)
        T = (T + rw.noise * mx * torch.randn(T.shape, generator=gen, device=dev)).contiguous()
    leaves = D.leaf_parameters(rw.phase, rw.euler, dev)
    n = rw.euler.shape[0]
    H, W = rw.shape

    def fwd_bwd():
        for t in leaves.values():
            t.grad = None
        L = model.l2_loss(*D.forward_args(leaves), T)
        L.backward()
        return L

    def fwd():
        with torch.no_grad():
            return model.patterns(*D.forward_args(leaves))

    out = {}
    for name, fn in (("fwd_bwd", fwd_bwd), ("fwd", fwd)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / steps
    L = fwd_bwd()
    torch.cuda.synchronize()
    # the fused step's device work alone: render adjoint (+ chain rule) and
    # structure-factor adjoint, timed with events on the launch stream
    with torch.no_grad():
        Rt, rec, xyz, B = D.forward_args(leaves)
        t = model.tables(rec, xyz, B)
        batch = model.spots(t, Rt)
        st = torch.zeros(batch.n, dtype=torch.int32, device=dev)
        stream = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        model.adjoint(t, Rt, batch, _lib.ADJ_L2, T, True, status=st)
        e0.record(stream)
        for _ in range(steps):
            model.adjoint(t, Rt, batch, _lib.ADJ_L2, T, True, status=st)
        e1.record(stream)
        torch.cuda.synchronize()
        adj_ms = e0.elapsed_time(e1) / steps
        adj_bytes = n * (H * W * 4 + 4) + float(batch.count.float().sum()) * 16 + n * 72
        achieved = adj_bytes / (adj_ms * 1e-3) / 1e9
        # FP64 error of the forward patterns on 32 orientations
        pats = model.patterns(Rt, rec, xyz, B)
        sel = np.arange(0, n, n // 32)[:32]
        ts = torch.as_tensor(sel, device=dev)
        p = P.oracle_params(rw.settings, rw.reciprocal_radius, False)
        otab = O.phase_tables(P.phase_spec(rw.phase), p)
        acc = spot_accuracy(otab, O.bunge_matrix(rw.euler[sel]), p, batch.idx[ts].cpu().numpy(),
                            batch.spots[ts].cpu().numpy(), batch.count[ts].cpu().numpy(),
                            pats[ts].cpu().numpy(), rw.shape, rw.sigma, rw.calibration,
                            fast=False, clip=1.0)
        del pats
    # CPU: the FP64 torch restatement of the fused step (forward, loss,
    # gradients w.r.t. Euler angles, a/b/c/angles, positions and B)
    pr = G.Problem(P.phase_spec(rw.phase), p, rw.shape, rw.sigma, rw.calibration, clip=1.0)
    n_cpu = 8
    Tn = T[:n_cpu].cpu().numpy().astype(np.float64)
    with threadpool_limits(limits=1):
        torch.set_num_threads(1)
        c0 = time.perf_counter()
        for o in range(n_cpu):
            eu = rw.euler[o:o + 1]
            pats64, frozen, tab64 = G.numpy_patterns(pr, O.bunge_matrix(eu))
            e, abc, ang, x64, B64 = (torch.tensor(np.asarray(v, dtype=np.float64),
                                                  requires_grad=True)
                                     for v in (eu, pr.ph.abc, pr.ph.angles, pr.ph.xyz, pr.ph.B))
            L64 = G.torch_loss(pr, G.bunge(e).transpose(-1, -2), G.reciprocal_basis(abc, ang),
                               x64, B64, frozen, tab64.hkl, Tn[o:o + 1])
            L64.backward()
        cdt = time.perf_counter() - c0
        torch.set_num_threads(os.cpu_count() or 1)
    return {
        "workload": "24-atom monoclinic cell (a,b,c~U[5,10] A, beta~U[95,115] deg, Z in "
                    "{8,12,14,26}, B~U[0.3,1]), r=1.5 1/A, max_exc 0.02, lorentzian, 200 kV; "
                    f"{n} targets 256x256 sub-pixel (0.015 1/A/px, sigma 1 px, clip 1) from the "
                    "perturbed truth + 1% noise; L2 loss, gradients w.r.t. Euler angles, a/b/c/beta, "
                    "positions and B",
        "n_patterns": n, "n_hkl": model.n_refl,
        "fwd_bwd_patterns_per_s": round(n / out["fwd_bwd"], 1),
        "fwd_patterns_per_s": round(n / out["fwd"], 1),
        "fwd_bwd_over_fwd": round(out["fwd"] / out["fwd_bwd"], 3),
        "ms_fwd_bwd": round(1e3 * out["fwd_bwd"], 3), "ms_fwd": round(1e3 * out["fwd"], 3),
        "ms_adjoint_device": round(adj_ms, 4),
        "loss": float(L.detach()),
        "dtype": "f32 (FP64 tables and excitation errors)",
        "roofline": {"bound": "hbm", "kernel": "dsx_render_adjoint + dsx_structure_factor_adjoint "
                     "(timed together)", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(adj_bytes, REFINE_PROFILE)},
        "cpu_baseline": cpu_baseline_entry(
            n_cpu / cdt, f"first {n_cpu} of {n} patterns, FP64 torch restatement of the fused "
            "forward+backward (oracle forward, frozen structure, autograd)", cdt),
        "accuracy": acc,
    }


def ragged_leg(dev, steps, scale):
    """64-phase ragged batch (about 10^6 orientations, 128x128) as one
    persistent simulate launch plus one render launch per method, captured in
    a hipGraph; whole-batch patterns/s over graph replays.  The simulate and
    render launches are also timed alone (eagerly, with events); two
    orientations per phase are compared with FP64 and the oracle's loop runs
    eight orientations per phase on one core."""
    import torch
    from threadpoolctl import threadpool_limits

    from diffsims_amd import _lib, workloads
    from diffsims_amd.ragged import RaggedLibrary
    from oracle import diffsims_oracle as O
    from tests import parity as P

    wl = workloads.ragged_library(orient_scale=scale)
    t0 = time.perf_counter()
    lib = RaggedLibrary(wl.phases, wl.rotations, wl.reciprocal_radius, wl.settings, wl.shape,
                        wl.sigma, wl.calibration, with_direct_beam=False, device=dev)
    lib.run()
    lib.check()
    lib.capture()
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    lib.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    # the launches alone, eagerly (the same kernels the graph replays)
    stream = torch.cuda.current_stream(dev)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 + len(lib.render_lists))]
    ev[0].record(stream)
    lib._simulate(lib.sim_bpc)
    ev[1].record(stream)
    kernel_ms = {}
    for j, (m, lst) in enumerate(lib.render_lists.items()):
        lib._render(m, lst, lib.render_bpc[m])
        ev[2 + j].record(stream)
    torch.cuda.synchronize()
    kernel_ms["simulate"] = ev[0].elapsed_time(ev[1])
    names = {_lib.RENDER_SPLAT: "render_splat", _lib.RENDER_DENSE: "render_dense",
             _lib.RENDER_DENSE_INPLACE: "render_dense_inplace"}
    for j, m in enumerate(lib.render_lists):
        kernel_ms[names[m]] = ev[1 + j].elapsed_time(ev[2 + j])
    n = lib.n_total
    H, W = lib.shape
    n_hkl = [s.table.n for s in lib.slots]
    counts = lib.count.float()
    # dominant render launch: patterns written + spot lists and counts read
    dom = max((m for m in lib.render_lists), key=lambda m: kernel_ms[names[m]])
    lst = lib.render_lists[dom].long()
    dom_bytes = lst.numel() * (H * W * 4 + 4) + float(counts[lst].sum()) * 16
    achieved = dom_bytes / (kernel_ms[names[dom]] * 1e-3) / 1e9
    # FP64 error: two orientations per phase; CPU: eight per phase
    k = 1 / O.electron_wavelength(wl.settings.accelerating_voltage)
    acc_parts = []
    cpu_n, cpu_s = 0, 0.0
    render = {"shape": wl.shape, "sigma": wl.sigma, "calibration": wl.calibration}
    for i, pa in enumerate(wl.phases):
        s = lib.slots[i]
        rr = wl.reciprocal_radius[i]
        p = P.oracle_params(wl.settings, rr, False)
        picks = np.unique(np.array([0, s.n - 1]))
        otab = P.sparse_oracle_table(P.phase_spec(pa), p, wl.rotations[i][picks])
        assert np.array_equal(otab.hkl, s.table.hkl_host()), f"phase {i}: hkl table"
        idx, sp, cnt = lib.phase_spots(i)
        tp = torch.as_tensor(picks, device=dev)
        acc_parts.append(spot_accuracy(otab, wl.rotations[i][picks], p, idx[tp].cpu().numpy(),
                                       sp[tp].cpu().numpy(), cnt[tp].cpu().numpy(),
                                       lib.phase_patterns(i)[tp].cpu().numpy(), wl.shape,
                                       wl.sigma, wl.calibration))
        with threadpool_limits(limits=1):
            c0 = time.perf_counter()
            O.simulate_library(P.phase_spec(pa), wl.rotations[i][:8], p, render)
            cpu_s += time.perf_counter() - c0
        cpu_n += min(8, s.n)
    acc = {"max_rel_err_vs_fp64": max(a["max_rel_err_vs_fp64"] for a in acc_parts),
           "max_intensity_rel_err": max(a["max_intensity_rel_err"] for a in acc_parts),
           "patterns_compared": sum(a["patterns_compared"] for a in acc_parts),
           "carved_orientations": sum(a["carved_orientations"] for a in acc_parts),
           "carved_rows_at_cut": sum(a["carved_rows_at_cut"] for a in acc_parts)}
    res = {
        "workload": "64 triclinic phases (atoms log-U[1,200], edges U[3,25] A, angles U[80,110] deg, "
                    "Z U[3,42], B U[0.2,1.5]), each 2k-50k random orientations (log-U) and r "
                    "U[1.0,2.5] 1/A; 200 kV, max_exc 0.02, lorentzian; 128x128 f32 at 0.04 1/A/px, "
                    "sigma 1 px, fast render, normalised; one ragged simulate launch + one render "
                    "launch per method, hipGraph replay",
        "n_phases": len(lib.slots), "n_patterns": n,
        "n_hkl_min": min(n_hkl), "n_hkl_max": max(n_hkl),
        "mean_spots": round(float(counts.mean()), 1), "max_spots": int(lib.count.max()),
        "patterns_per_s": round(n / dt, 1), "ms_per_step": round(1e3 * dt, 3),
        "kernel_ms": {k_: round(v, 3) for k_, v in kernel_ms.items()},
        "setup_s_incl_tables_and_tuning": round(setup, 2),
        "render_phases": {"splat": sum(s.method == _lib.RENDER_SPLAT for s in lib.slots),
                          "dense": sum(s.method == _lib.RENDER_DENSE for s in lib.slots),
                          "dense_inplace": sum(s.method == _lib.RENDER_DENSE_INPLACE
                                               for s in lib.slots)},
        "tuned": {"simulate_blocks_per_cu": lib.sim_bpc,
                  "render_blocks_per_cu": {str(k_): v for k_, v in lib.render_bpc.items()}},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     # the counters measured the in-place dense renderer
                     "traffic": (measured_traffic(dom_bytes, RAGGED_PROFILE)
                                 if dom == _lib.RENDER_DENSE_INPLACE else None)},
        "cpu_baseline": cpu_baseline_entry(
            cpu_n / cpu_s, f"first 8 orientations of each of the 64 phases ({cpu_n} patterns), "
            "FP64 NumPy restatement of the reference loop incl. render", cpu_s),
        "accuracy": acc,
    }
    del lib
    torch.cuda.empty_cache()
    return res


def grid_leg(dev, resolution):
    """Beam-direction grid of a template library on the device (triclinic,
    spherified-cube-edge mesh): directions/s of dsx_beam_directions (its
    parity with the oracle is a GPU test)."""
    import torch

    from diffsims_amd import generators

    g = generators.beam_directions_grid_device("triclinic", resolution, device=dev)
    torch.cuda.synchronize()
    reps = 5
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        generators.beam_directions_grid_device("triclinic", resolution, device=dev)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    n = int(g.shape[0])
    return {"workload": f"triclinic beam-direction grid, spherified_cube_edge mesh, {resolution} deg",
            "directions": n, "ms_per_grid": round(ms, 3),
            "directions_per_s": round(n / (ms * 1e-3), 1)}


def large_precession_leg(dev, n_chunks):
    """500-atom triclinic cell, r = 3.0 (~1.6e6 reflections), precession 1 deg,
    512x512 sub-pixel patterns; each 65,536-orientation chunk computes the summed
    L2 loss against one target and its orientation and B gradients.  Also the
    forward alone (the chunk's patterns rendered, in slices of 16,384), the
    FP64 error of one pattern and the oracle's loop on one orientation."""
    import torch
    from threadpoolctl import threadpool_limits

    from diffsims_amd import differentiable as D
    from diffsims_amd import workloads
    from oracle import diffsims_oracle as O
    from tests import parity as P

    d = workloads.large_precession()
    pa = d["phase"]
    t0 = time.perf_counter()
    model = D.RefinementModel(pa, d["reciprocal_radius"], d["settings"], d["shape"], d["sigma"],
                              d["calibration"], clip_threshold=d["clip"], device=dev)
    setup = time.perf_counter() - t0
    with torch.no_grad():
        T = model.patterns(*D.forward_args(D.leaf_parameters(pa, d["target_euler"], dev,
                                                             requires_grad=False)))[0].contiguous()
    chunk = d["chunk"]
    H, W = d["shape"]
    base = D.leaf_parameters(pa, d["euler"][:chunk], dev)
    for name in ("abc", "angles", "xyz"):
        base[name].requires_grad_(False)
    stream = torch.cuda.current_stream(dev)
    times, dev_ms, losses = [], [], []
    n_done = 0
    for c in range(n_chunks):
        e = d["euler"][c * chunk:(c + 1) * chunk]
        leaves = dict(base)
        leaves["euler"] = torch.tensor(e, dtype=torch.float64, device=dev, requires_grad=True)
        leaves["B"] = base["B"].detach().clone().requires_grad_(True)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(stream)
        L = model.l2_loss(*D.forward_args(leaves), T)
        L.backward()
        ev1.record(stream)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        dev_ms.append(ev0.elapsed_time(ev1))
        losses.append(float(L.detach()))
        n_done += e.shape[0]
    steady = times[1:] if len(times) > 1 else times
    per_chunk = float(np.mean(steady))
    # the forward alone: the first chunk's patterns, 8,192 at a time
    # 16,384 orientations are 4,096 spot-kernel tiles, one per resident wave
    # (a slice of 8,192 left half the GPU idle in the spot kernel)
    sl = 16384
    with torch.no_grad():
        args = D.forward_args(D.leaf_parameters(pa, d["euler"][:sl], dev, requires_grad=False))
        model.patterns(*args)                                   # warm (capacity, allocator)
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        for c in range(0, chunk, sl):
            a = D.forward_args(D.leaf_parameters(pa, d["euler"][c:c + sl], dev,
                                                 requires_grad=False))
            pats = model.patterns(*a)
            del pats
        torch.cuda.synchronize()
        fwd_s = time.perf_counter() - f0
        # spot counts and the FP64 error of one pattern
        t = model.tables(*args[1:])
        b = model.spots(t, args[0])
        mean_spots = float(b.count.float().mean())
        one = model.patterns(args[0][:1], *args[1:])[0].cpu().numpy()
        p = P.oracle_params(d["settings"], d["reciprocal_radius"], False)
        M1 = O.bunge_matrix(d["euler"][:1])
        otab = P.sparse_oracle_table(P.phase_spec(pa), p, M1, hkl=model.hkl_np.astype(np.int64))
        acc = spot_accuracy(otab, M1, p, b.idx[:1].cpu().numpy(), b.spots[:1].cpu().numpy(),
                            b.count[:1].cpu().numpy(), one[None], d["shape"], d["sigma"],
                            d["calibration"], fast=False, clip=d["clip"])
    # CPU: the reference-shaped loop (F of the intersecting rows, sub-pixel
    # render) for one orientation
    ph = P.phase_spec(pa)
    _, rec = ph.basis()
    rows = O.hkl_box(rec, 1.0 / d["reciprocal_radius"])
    k = 1 / O.electron_wavelength(200)
    with threadpool_limits(limits=1):
        c0 = time.perf_counter()
        sp = O.simulate_one_reference(ph, rows, rec, M1[0], k, p)
        O.render_subpixel(sp.xyz[:, :2], sp.intensity, d["shape"], d["sigma"], d["calibration"],
                          clip_threshold=d["clip"])
        cdt = time.perf_counter() - c0
    chunk_bytes = chunk * (mean_spots * 16 + 4 + 72) + H * W * 4
    achieved = chunk_bytes / (float(np.mean(dev_ms[1:] if len(dev_ms) > 1 else dev_ms)) * 1e-3) / 1e9
    return {
        "workload": "500-atom triclinic cell (edges U[15,30] A, angles U[80,110] deg, Z U[1,30], "
                    "B U[0.5,2]), r=3.0 1/A, precession 1 deg (precession-integrated lorentzian), "
                    "max_exc 0.02, 200 kV; chunks of 65,536 random orientations, 512x512 sub-pixel "
                    "(0.012 1/A/px, sigma 1 px, clip 1); summed L2 loss vs one target, orientation "
                    "and B gradients",
        "n_hkl": model.n_refl, "chunks": n_chunks, "orientations": n_done,
        "mean_spots": round(mean_spots, 1), "spot_cap": model.cap,
        "ms_per_chunk_first": round(1e3 * times[0], 1),
        "ms_per_chunk": round(1e3 * per_chunk, 1),
        "fwd_bwd_patterns_per_s": round(chunk / per_chunk, 1) if n_done >= chunk else None,
        "fwd_patterns_per_s": round(chunk / fwd_s, 1),
        "fwd_bwd_over_fwd": round(fwd_s / per_chunk, 3),
        "setup_s": round(setup, 2), "loss_first_chunk": losses[0],
        "roofline": {"bound": "hbm", "kernel": "fused chunk (render_adjoint_large_kernel dominant)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": measured_traffic(chunk_bytes, LARGE_PROFILE),
                     "note": "algorithmic bytes: the spot lists read once and the shared target; "
                             "the measured traffic adds the spot lists' writing and the adjoint's "
                             "intermediate pixels and per-spot sums. The kernel is bound by LDS "
                             "and VALU work per spot, not by HBM"},
        "cpu_baseline": cpu_baseline_entry(
            1 / cdt, "1 orientation, FP64 NumPy restatement of the reference loop (F of the "
            "intersecting rows, precession shape factor) incl. sub-pixel render, forward only",
            cdt),
        "accuracy": acc,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from diffsims_amd import engine, workloads

    dev = torch.device("cuda", 0)
    if args.calibrate_copy > 0:
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        # the same buffers timed with HIP events: what a plain copy and a plain
        # store stream reach on this GPU (ceilings for the HBM-bound kernels)
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        for _ in range(3):
            dst.copy_(src)
        e[1].record()
        for _ in range(3):
            dst.fill_(2.0)
        e[2].record()
        torch.cuda.synchronize()
        copy_ms, fill_ms = e[0].elapsed_time(e[1]) / 3, e[1].elapsed_time(e[2]) / 3
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n,
                          "copy_GBps": round(8 * n / copy_ms / 1e6, 1),
                          "fill_GBps": round(4 * n / fill_ms / 1e6, 1)}), flush=True)
        return
    wl = workloads.template_library(args.orientations)
    pa = wl.phases[0]
    rr = wl.reciprocal_radius[0]
    tab = engine.build_phase_table(pa, rr, wl.with_direct_beam, "lobato", device=dev)
    Rt = engine.rotation_matrices_to_device(wl.rotations[0], dev)
    n = Rt.shape[0]
    # capacity from one calibration pass (ncand is exact even on overflow)
    probe = engine.simulate(tab, Rt[:4096], wl.settings, rr)
    cap = int(max(probe.cap, int(probe.ncand.max()) + 64))
    batch = engine.empty_batch(tab, n, cap, dev)
    H, W = wl.shape
    out = torch.empty((n, H, W), dtype=torch.float32, device=dev)
    status = torch.zeros(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)

    def step():
        engine.simulate_into(tab, Rt, wl.settings, batch)
        engine.render_spots(batch.spots, batch.count, batch.cap, wl.shape, wl.sigma,
                            wl.calibration, out=out, status=status)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if int(batch.ncand.max()) > cap:             # re-size once if the grid needs more
        cap = int(batch.ncand.max()) + 64
        batch = engine.empty_batch(tab, n, cap, dev)
        step()
        torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record(stream)
        engine.simulate_into(tab, Rt, wl.settings, batch)
        e1.record(stream)
        engine.render_spots(batch.spots, batch.count, batch.cap, wl.shape, wl.sigma,
                            wl.calibration, out=out, status=status)
        e2.record(stream)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    engine.check_render_status(status)
    assert int(batch.ncand.max()) <= cap
    sim_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    ren_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    counts = batch.count.float()
    mean_spots = float(counts.mean())
    # algorithmic bytes of one render stage: every pattern written once, its
    # spot list and count read once.  The stage is the plan kernel, the band
    # writer and the splat kernel for listed patterns (dsx_render_fast_bands),
    # timed together by the events around engine.render_spots
    render_bytes = n * (H * W * 4) + float(counts.sum()) * 16 + n * 4
    achieved = render_bytes / (ren_ms * 1e-3) / 1e9
    value = args.steps * n / dt
    res = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "patterns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "orthorhombic 40-atom template library (a,b,c~U[4,12] A, Z in {6,8,14,26}, "
                        "B~U[0.3,1]): r=2.0 1/A, max_exc 0.02, lorentzian, 200 kV, no direct beam, "
                        f"{n} orientations (super-Fibonacci SO(3) grid), {H}x{W} f32 patterns at "
                        f"{wl.calibration} 1/A/px, sigma {wl.sigma} px, fast render, normalised",
            "n_orientations": n, "n_hkl": tab.n, "mean_spots": round(mean_spots, 2),
            "mode": "batched forward (simulate + render)",
            "kernel_ms": {"simulate": round(sim_ms, 3), "render": round(ren_ms, 3)},
            "render_kernels": "render_plan_kernel + render_band_kernel (+ render_kernel for "
                              "patterns the plan lists)",
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(render_bytes)},
    }
    if args.accuracy_patterns > 0:
        acc = library_accuracy(wl, batch, out, args.accuracy_patterns)
        res["config"]["max_rel_err_vs_fp64"] = acc["max_rel_err_vs_fp64"]
        res["config"]["accuracy"] = acc
    if args.cpu_sample > 0:
        v, t = cpu_baseline(wl, args.cpu_sample)
        res["cpu_baseline"] = cpu_baseline_entry(
            v, f"first {args.cpu_sample} of {n} orientations, FP64 NumPy restatement of the "
            "reference loop incl. render", t)
    del out, batch, status, Rt
    torch.cuda.empty_cache()
    errors = {"C1 library": res["config"].get("max_rel_err_vs_fp64")}
    if args.plain_steps > 0:
        res["plain"] = plain_leg(dev, args.plain_steps, args.warmup)
        errors["C0 plain"] = res["plain"]["accuracy"]["max_rel_err_vs_fp64"]
        torch.cuda.empty_cache()
    if args.refine_steps > 0:
        res["fwd_bwd"] = refinement_leg(dev, args.refine_steps, args.warmup)
        res["fwd_bwd"]["fwd_bwd_over_library_fwd"] = round(
            res["fwd_bwd"]["fwd_bwd_patterns_per_s"] / value, 4)
        errors["C3 refinement (forward patterns)"] = res["fwd_bwd"]["accuracy"]["max_rel_err_vs_fp64"]
        torch.cuda.empty_cache()
    if args.ragged_steps > 0:
        res["ragged"] = ragged_leg(dev, args.ragged_steps, args.ragged_scale)
        errors["C2 ragged"] = res["ragged"]["accuracy"]["max_rel_err_vs_fp64"]
        torch.cuda.empty_cache()
    if args.large_chunks > 0:
        res["large_precession"] = large_precession_leg(dev, args.large_chunks)
        errors["C4 large precession"] = res["large_precession"]["accuracy"]["max_rel_err_vs_fp64"]
        torch.cuda.empty_cache()
    if args.grid_resolution > 0:
        res["beam_grid"] = grid_leg(dev, args.grid_resolution)
    res["max_rel_err_vs_fp64_per_config"] = errors
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
