"""The 64-phase ragged batch itself (GPU): ``workloads.ragged_library`` with
its real phases (1-200 atoms, 3.5k-365k reflections, patterns of thousands
of spots) at a reduced orientation count, through ``RaggedLibrary`` with the
autotuner's own render-method choice and a hipGraph replay.

Per phase: the hkl table equals the oracle's box bit for bit; at least four
orientations are compared with the FP64 oracle (survivor sets, intensities
and rendered 128x128 patterns, with the carve-outs of tests/parity.py counted,
not skipped).  |F|^2 is evaluated in FP64 for the rows that intersect the
compared orientations (and the rows at the cut), which is what the pattern
depends on.
"""

import numpy as np
import pytest
import torch

from oracle import diffsims_oracle as O
from tests import parity as P

pytestmark = pytest.mark.gpu


def test_ragged_library_64_phases():
    from diffsims_amd import _lib, workloads
    from diffsims_amd.ragged import RaggedLibrary

    wl = workloads.ragged_library(orient_scale=0.003)
    print(f"64-phase ragged batch: {wl.n_patterns} orientations", flush=True)
    lib = RaggedLibrary(wl.phases, wl.rotations, wl.reciprocal_radius, wl.settings, wl.shape,
                        wl.sigma, wl.calibration, with_direct_beam=False)
    lib.run()
    lib.check()
    torch.cuda.synchronize()
    eager = lib.out.clone()
    eager_cnt = lib.count.clone()
    lib.capture()
    lib.out.zero_()
    lib.count.fill_(-1)
    torch.cuda.synchronize()
    lib.replay()
    torch.cuda.synchronize()
    assert torch.equal(lib.count, eager_cnt) and torch.equal(lib.out, eager)
    methods = [s.method for s in lib.slots]
    print(f"  render methods: {methods.count(_lib.RENDER_SPLAT)} splat, "
          f"{methods.count(_lib.RENDER_DENSE)} dense; max spots {int(lib.count.max())}", flush=True)

    k = 1 / O.electron_wavelength(wl.settings.accelerating_voltage)
    stats = {}
    worst = 0.0
    n_refl = []
    for i, pa in enumerate(wl.phases):
        s = lib.slots[i]
        rr = wl.reciprocal_radius[i]
        M = wl.rotations[i]
        picks = np.unique(np.linspace(0, s.n - 1, 4).round().astype(int))
        assert picks.size == 4, (i, s.n)
        p = P.oracle_params(wl.settings, rr, False)
        otab = P.sparse_oracle_table(P.phase_spec(pa), p, M[picks])
        assert np.array_equal(s.table.hkl_host(), otab.hkl), i
        n_refl.append(len(otab.hkl))
        idx, sp, cnt = lib.phase_spots(i)
        idx, sp, cnt = idx.cpu().numpy(), sp.cpu().numpy(), cnt.cpu().numpy()
        pats = lib.phase_patterns(i).cpu().numpy()
        for o in picks:
            c = int(cnt[o])
            ref, _ = P.check_orientation(otab, M[o], k, p, idx[o, :c], sp[o, :c, 3], stats)
            rp = P.flipped_reference_pattern(ref, sp[o, :c, :2], wl.shape, wl.sigma,
                                             wl.calibration)
            err = P.pattern_error(pats[o], rp)
            assert err <= P.TOL_P, (i, int(o), err)
            worst = max(worst, err)
    print(f"  {stats['orientations']} orientations over 64 phases ({min(n_refl)}-{max(n_refl)} "
          f"reflections): max intensity error {stats['max_rel_err']:.2e}, max pattern error "
          f"{worst:.2e}; carved {stats['carved_orientations']} orientations / "
          f"{stats['carved_rows']} rows at the cut", flush=True)
    assert stats["orientations"] == 4 * 64
    assert max(n_refl) > 300_000
    assert stats["carved_orientations"] <= stats["orientations"] // 20
