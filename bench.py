"""Throughput of the kinematical pattern-simulation path on one MI355X.

One step = simulate + render the whole orthorhombic 40-atom template library
(200,000 orientations, 256x256 FP32 patterns) with every input resident in HBM:
the per-orientation kernel (excitation error, shape factor, cut, compaction)
followed by the renderer, writing all patterns to one [n, 256, 256] tensor.

Prints one JSON line (see --help).  The CPU baseline is the FP64 NumPy
restatement of the reference's loop timed on a bounded sample on this host.
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


PROFILE = os.path.join(ROOT, "profiles", "render_fast_ortho40_library_kernel_trace.md")


def measured_traffic(render_bytes):
    """HBM bytes of one render launch: the FETCH_SIZE/WRITE_SIZE measurement
    committed in profiles/ (taken on the 200,000-orientation library), scaled
    to this run's algorithmic bytes; None when the summary is absent."""
    import re
    try:
        text = open(PROFILE).read()
        t = float(re.search(r"traffic_bytes_per_launch: (\d+)", text).group(1))
        a = float(re.search(r"algorithmic_bytes_per_launch: (\d+)", text).group(1))
    except (OSError, AttributeError):
        return None
    return round(t / a * render_bytes)


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
    ap.add_argument("--large-chunks", type=int, default=1,
                    help="65,536-orientation chunks of the 500-atom precession case to run "
                         "fused forward+backward (16 = the full 10^6 orientations; 0 disables)")
    ap.add_argument("--grid-resolution", type=float, default=0.1,
                    help="degrees of the triclinic beam-direction grid leg (0 disables it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (the reference "
                         "kernel for scaling the HBM byte counters of a profiler) and exit")
    return ap.parse_args()


def cpu_baseline(wl, n_sample, batch, pats):
    """Reference-shaped FP64 loop (the oracle) on the first n_sample
    orientations, single thread; its patterns are then compared with the GPU's
    for the same orientations: max |P - P64| / max(P64), with the carve-outs
    of tests/parity.py (index sets equal; spots within 1e-4 px of a pixel
    edge compared with the flipped pixel)."""
    from threadpoolctl import threadpool_limits

    from oracle import diffsims_oracle as O
    from tests import parity as P

    pa = wl.phases[0]
    ph = O.PhaseSpec(pa.abc, pa.angles, pa.symbols, pa.xyz, pa.occ, pa.B)
    st = wl.settings
    p = O.SimParams(kv=st.accelerating_voltage, reciprocal_radius=wl.reciprocal_radius[0],
                    max_excitation_error=st.max_excitation_error, shape_model=st.shape_model,
                    minimum_intensity=st.minimum_intensity, with_direct_beam=wl.with_direct_beam)
    M = wl.rotations[0][:n_sample]
    render = {"shape": wl.shape, "sigma": wl.sigma, "calibration": wl.calibration}
    with threadpool_limits(limits=1):
        O.simulate_library(ph, M[:8], p, render)          # warm caches
        t0 = time.perf_counter()
        spots, _ = O.simulate_library(ph, M, p, render)
        dt = time.perf_counter() - t0
    # accuracy of the GPU patterns on the same sample (exact-phase |F|^2 oracle)
    otab = O.phase_tables(ph, p)
    k = 1 / O.electron_wavelength(st.accelerating_voltage)
    worst, compared = 0.0, 0
    for o in range(0, n_sample, max(1, n_sample // 64)):
        ref = O.simulate_one(otab, M[o], k, p)
        c = int(batch.count[o])
        if not np.array_equal(batch.idx[o, :c].cpu().numpy(), ref.idx):
            continue
        rp = P.flipped_reference_pattern(ref, batch.spots[o, :c, :2].cpu().numpy(), wl.shape,
                                         wl.sigma, wl.calibration)
        worst = max(worst, float(np.abs(pats[o].cpu().numpy() - rp).max() / max(rp.max(), 1e-300)))
        compared += 1
    return n_sample / dt, dt, worst, compared


def refinement_leg(dev, steps, warmup):
    """Fused forward+backward of the L2 refinement loss (24-atom monoclinic
    cell, 4,096 256x256 sub-pixel targets) against the plain forward of the
    same patterns; both in patterns/s."""
    import torch

    from diffsims_amd import differentiable as D
    from diffsims_amd import workloads

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
        "loss": float(L.detach()),
        "dtype": "f32 (FP64 tables and excitation errors)",
    }


def ragged_leg(dev, steps, scale):
    """64-phase ragged batch (about 10^6 orientations, 128x128) as one
    persistent simulate launch plus one render launch per method, captured in
    a hipGraph; whole-batch patterns/s over graph replays."""
    import torch

    from diffsims_amd import _lib, workloads
    from diffsims_amd.ragged import RaggedLibrary

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
    n = lib.n_total
    n_hkl = [s.table.n for s in lib.slots]
    counts = lib.count.float()
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
        "setup_s_incl_tables_and_tuning": round(setup, 2),
        "render_phases": {"splat": sum(s.method == _lib.RENDER_SPLAT for s in lib.slots),
                          "dense": sum(s.method == _lib.RENDER_DENSE for s in lib.slots)},
        "tuned": {"simulate_blocks_per_cu": lib.sim_bpc,
                  "render_blocks_per_cu": {str(k): v for k, v in lib.render_bpc.items()}},
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
    L2 loss against one target and its orientation and B gradients."""
    import torch

    from diffsims_amd import differentiable as D
    from diffsims_amd import workloads

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
    base = D.leaf_parameters(pa, d["euler"][:chunk], dev)
    for name in ("abc", "angles", "xyz"):
        base[name].requires_grad_(False)
    times, spots, losses = [], [], []
    n_done = 0
    for c in range(n_chunks):
        e = d["euler"][c * chunk:(c + 1) * chunk]
        leaves = dict(base)
        leaves["euler"] = torch.tensor(e, dtype=torch.float64, device=dev, requires_grad=True)
        leaves["B"] = base["B"].detach().clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L = model.l2_loss(*D.forward_args(leaves), T)
        L.backward()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        losses.append(float(L.detach()))
        n_done += e.shape[0]
    # mean spots of the last chunk (its spot lists are rebuilt for the count)
    t = model.tables(*D.forward_args(leaves)[1:])
    b = model.spots(t, D.forward_args(leaves)[0].detach())
    mean_spots = float(b.count.float().mean())
    steady = times[1:] if len(times) > 1 else times
    per_chunk = float(np.mean(steady))
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
        "setup_s": round(setup, 2), "loss_first_chunk": losses[0],
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
    # algorithmic bytes of one render launch: every pattern written once, its
    # spot list and count read once
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
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(render_bytes)},
    }
    if args.cpu_sample > 0:
        v, t, err, compared = cpu_baseline(wl, args.cpu_sample, batch, out)
        res["cpu_baseline"] = {"value": round(v, 2), "unit": "patterns/s", "cores": 1,
                               "kind": "port",
                               "sample": f"first {args.cpu_sample} of {n} orientations, FP64 NumPy "
                                         f"restatement of the reference loop incl. render "
                                         f"({t:.1f} s, 1 thread)"}
        res["config"]["max_rel_err_vs_fp64"] = err
        res["config"]["accuracy_patterns"] = compared
    del out, batch, status, Rt
    torch.cuda.empty_cache()
    if args.refine_steps > 0:
        res["fwd_bwd"] = refinement_leg(dev, args.refine_steps, args.warmup)
        res["fwd_bwd"]["fwd_bwd_over_library_fwd"] = round(
            res["fwd_bwd"]["fwd_bwd_patterns_per_s"] / value, 4)
        torch.cuda.empty_cache()
    if args.ragged_steps > 0:
        res["ragged"] = ragged_leg(dev, args.ragged_steps, args.ragged_scale)
        torch.cuda.empty_cache()
    if args.large_chunks > 0:
        res["large_precession"] = large_precession_leg(dev, args.large_chunks)
    if args.grid_resolution > 0:
        res["beam_grid"] = grid_leg(dev, args.grid_resolution)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
