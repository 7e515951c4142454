"""The allocation-free render adjoint (dsx_render_adjoint_noalloc, static
device scratch) against dsx_render_adjoint with a caller workspace (GPU).

Both run the same kernels on the same inputs, so the loss and the patterns
must be bit-identical.  The gradients are sums over a pattern's spots in
the order the spots took their LDS slots (an atomic counter), partly in
FP32 (the long-list kernel's gathered records, dL/d|F|^2's atomics), so
they agree to rounding: 1e-6 of each gradient's largest component (the
gradient tests' own bar is 5e-5).  Two calls in a row,
the second with fewer patterns, for a small-pattern case and a long-list
(chunked) case; a size beyond the static scratch is rejected.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import diffsims_oracle as O
from tests.test_grad_oracle import mono_problem
from tests.test_gpu_adjoint import _big_problem, _model, _params

pytestmark = pytest.mark.gpu


def _call(model, t, Rt, b, n, target, noalloc):
    from diffsims_amd import _lib

    L = _lib.load()
    H, W = model.shape
    loss = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    pat = torch.full((n, H, W), -1.0, dtype=torch.float32, device="cuda")
    dRt = torch.full((n, 9), -1.0, dtype=torch.float64, device="cuda")
    dF2 = torch.zeros(model.n_refl, dtype=torch.float32, device="cuda")
    drec = torch.zeros(9, dtype=torch.float64, device="cuda")
    status = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    sim = model._sim_params(t)
    common = (_lib.ptr(b.idx), _lib.ptr(b.spots), _lib.ptr(b.count), b.cap, n, _lib.ptr(t.hkl),
              _lib.ptr(t.F2), model.n_refl, _lib.ptr(Rt), ctypes.byref(sim),
              ctypes.byref(model.render), _lib.ADJ_L2, _lib.ptr(target), H * W, _lib.ptr(pat),
              _lib.ptr(loss), _lib.ptr(dRt), _lib.ptr(dF2), _lib.ptr(drec))
    stream = _lib.stream_handle()
    if noalloc:
        _lib.check(L.dsx_render_adjoint_noalloc(*common, _lib.ptr(status), stream),
                   "dsx_render_adjoint_noalloc")
    else:
        work = torch.empty(max(1, int(L.dsx_render_adjoint_workspace_ints(b.cap, n, H, W, 1))),
                           dtype=torch.int32, device="cuda")
        _lib.check(L.dsx_render_adjoint(*common, _lib.ptr(work), work.numel(), _lib.ptr(status),
                                        stream), "dsx_render_adjoint")
    torch.cuda.synchronize()
    assert int(status.abs().max()) == 0
    return loss, pat, dRt, dF2, drec


def _compare(a, b):
    loss_a, pat_a, dRt_a, dF2_a, drec_a = a
    loss_b, pat_b, dRt_b, dF2_b, drec_b = b
    assert torch.equal(loss_a, loss_b)
    assert torch.equal(pat_a, pat_b)
    for x, y, tol in ((dRt_a, dRt_b, 1e-6), (drec_a, drec_b, 1e-6), (dF2_a, dF2_b, 1e-6)):
        scale = max(float(y.abs().max()), 1e-30)
        assert float((x - y).abs().max()) <= tol * scale


@pytest.mark.parametrize("case", ["small", "long_lists"])
def test_noalloc_equals_workspace_entry(case):
    pr = mono_problem() if case == "small" else _big_problem()
    sizes = (96, 40) if case == "small" else (6, 2)
    print(f"no-alloc adjoint, {case} spot lists, {sizes} patterns", flush=True)
    # long lists: a capacity whose per-CU chunk workspace fits the static scratch
    model = _model(pr) if case == "small" else _model(pr, cap=2600)
    eu = O.random_euler(sizes[0], 77)
    _, args = _params(pr, eu, "cuda")
    Rt = args[0].detach().contiguous()
    t = model.tables(*args[1:])
    b = model.spots(t, Rt)
    from diffsims_amd import _lib
    L = _lib.load()
    H, W = model.shape
    assert int(L.dsx_render_adjoint_workspace_ints(b.cap, sizes[0], H, W, 1)) <= \
        int(L.dsx_render_adjoint_noalloc_max_ints())
    if case == "long_lists":
        assert b.cap > int(L.dsx_render_max_spots())
    gen = torch.Generator("cuda").manual_seed(3)
    target = torch.rand((sizes[0],) + model.shape, device="cuda", generator=gen) * 0.3
    for n in sizes:  # second call smaller: stale scratch from the first must not leak
        ref = _call(model, t, Rt, b, n, target, noalloc=False)
        got = _call(model, t, Rt, b, n, target, noalloc=True)
        _compare(got, ref)
        assert float(ref[0].sum()) > 0


def test_noalloc_rejects_sizes_beyond_the_scratch():
    from diffsims_amd import _lib

    L = _lib.load()
    max_ints = int(L.dsx_render_adjoint_noalloc_max_ints())
    assert max_ints == 16 << 20
    cap = 1000
    n = max_ints // (4 * cap) + 16
    assert int(L.dsx_render_adjoint_workspace_ints(cap, n, 64, 64, 1)) > max_ints
    sim = _lib.SimParams()
    rp = _lib.RenderParams()
    rp.height, rp.width = 64, 64
    with pytest.raises(ValueError, match="static scratch"):
        _lib.check(L.dsx_render_adjoint_noalloc(
            None, None, None, cap, n, None, None, 0, None, ctypes.byref(sim), ctypes.byref(rp),
            _lib.ADJ_L2, None, 0, None, None, None, None, None, None, _lib.stream_handle()),
            "dsx_render_adjoint_noalloc")
