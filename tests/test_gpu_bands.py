"""Band renderer (plan kernel + one-shot band writer) against the splat renderer (GPU).

dsx_render_fast_bands must give dsx_render_fast's output bit for bit: the plan
reproduces the splat kernel's fixed-point scale and exact maximum, and the band
writer adds the same fixed-point terms.  Both are called through the C ABI on
the same seeded spot lists; the outputs start as NaN so that every pixel must
be written.  The spot lists mix in-frame and out-of-frame spots, spots on the
same pixel (last write wins), spots within the tap radius of the edges
(reflected taps) and the shapes whose last band is partial.  The library
parity tests (test_gpu_parity.py) check the same path against the FP64
oracle through engine.render_spots.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from diffsims_amd import _lib as lib
    return lib, lib.load(require_gpu=True)


def _spot_lists(n, cap, H, W, mean, seed, edge=0.0, dup=0.0, margin=2.0, R=4):
    """Spot lists in pixel units around the default centre (calibration 1)."""
    import torch
    rng = np.random.default_rng(seed)
    cnt = np.minimum(rng.poisson(mean, n), cap).astype(np.int32)
    X = rng.uniform(-margin, W + margin, (n, cap))
    Y = rng.uniform(-margin, H + margin, (n, cap))
    if edge > 0:
        e = rng.random((n, cap)) < edge
        side = rng.integers(0, 4, (n, cap))
        near = rng.uniform(0.0, R + 1.0, (n, cap))
        X = np.where(e & (side == 0), near, X)
        X = np.where(e & (side == 1), W - near, X)
        Y = np.where(e & (side == 2), near, Y)
        Y = np.where(e & (side == 3), H - near, Y)
    if dup > 0:
        # a later spot on the pixel of an earlier one (other sub-pixel offset)
        for i in range(n):
            for j in range(1, int(cnt[i])):
                if rng.random() < dup:
                    k = int(rng.integers(0, j))
                    X[i, j] = np.floor(X[i, k]) + rng.uniform(0.05, 0.95)
                    Y[i, j] = np.floor(Y[i, k]) + rng.uniform(0.05, 0.95)
    I = rng.lognormal(0.0, 1.5, (n, cap))
    sp = np.zeros((n, cap, 4), np.float32)
    sp[..., 0] = X - W // 2
    sp[..., 1] = Y - H // 2
    sp[..., 3] = I
    return torch.from_numpy(sp).cuda(), torch.from_numpy(cnt).cuda()


def _render_both(spots, cnt, cap, shape, sigma, normalize=True, angle=0.0, mirrored=False):
    """(splat output, band output, splat status, band status, patterns the
    last round listed for the splat kernel)."""
    import torch
    from diffsims_amd import engine
    lib, L = _lib()
    n = int(cnt.shape[0])
    H, W = shape
    p = engine.render_params(shape, sigma, 1.0, None, angle, mirrored, normalize, True)
    ref = torch.full((n, H, W), float("nan"), dtype=torch.float32, device="cuda")
    out = torch.full((n, H, W), float("nan"), dtype=torch.float32, device="cuda")
    st_ref = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    st = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    stream = lib.stream_handle(spots.device)
    lib.check(L.dsx_render_fast(lib.ptr(spots), lib.ptr(cnt), cap, n, ctypes.byref(p), lib.ptr(ref),
                                lib.ptr(st_ref), stream), "dsx_render_fast")
    ws = int(L.dsx_render_fast_bands_workspace_ints(n, H, W))
    work = torch.full((ws,), -7, dtype=torch.int32, device="cuda")
    lib.check(L.dsx_render_fast_bands(lib.ptr(spots), lib.ptr(cnt), cap, n, ctypes.byref(p),
                                      lib.ptr(work), ws, lib.ptr(out), lib.ptr(st), stream),
              "dsx_render_fast_bands")
    torch.cuda.synchronize()
    return ref, out, st_ref, st, int(work[0].item())


def _assert_same(ref, out, st_ref, st):
    import torch
    assert not torch.isnan(out).any()
    bad = (ref != out).flatten(1).any(1)
    assert not bool(bad.any()), f"patterns differ: {torch.nonzero(bad).flatten()[:8].tolist()}"
    assert torch.equal(st_ref, st)


def test_library_like_patterns_bit_identical():
    print("bands: 3000 x 256x256, sigma 1, ~260 spots", flush=True)
    spots, cnt = _spot_lists(3000, 400, 256, 256, 260, seed=1, edge=0.05, dup=0.02)
    ref, out, st_ref, st, listed = _render_both(spots, cnt, 400, (256, 256), 1.0)
    _assert_same(ref, out, st_ref, st)
    assert listed < 300, listed  # the band path renders most patterns itself


@pytest.mark.parametrize("sigma", [0.3, 0.6, 1.0, 1.4, 2.0])
@pytest.mark.parametrize("normalize", [True, False])
def test_tap_radii_partial_bands_and_edges(sigma, normalize):
    R = int(4 * sigma + 0.5)
    print(f"bands: 400 x 144x144, sigma {sigma} (R={R}), normalize {normalize}", flush=True)
    spots, cnt = _spot_lists(400, 256, 144, 144, 120, seed=10 + R, edge=0.2, dup=0.05, R=R)
    ref, out, st_ref, st, listed = _render_both(spots, cnt, 256, (144, 144), sigma, normalize)
    _assert_same(ref, out, st_ref, st)
    assert listed < 400


def test_in_plane_angle_and_mirror():
    print("bands: 500 x 256x192, angle 23, mirrored", flush=True)
    spots, cnt = _spot_lists(500, 320, 256, 192, 200, seed=3, edge=0.05, dup=0.03)
    for angle, mirrored in ((23.0, False), (23.0, True), (0.0, True)):
        ref, out, st_ref, st, _ = _render_both(spots, cnt, 320, (256, 192), 1.0, angle=angle,
                                               mirrored=mirrored)
        _assert_same(ref, out, st_ref, st)


def test_large_patterns_and_long_lists():
    # lists above the plan's 512 spots go to the splat kernel; the rest stay
    print("bands: 300 x 512x512, ~480 spots, cap 700", flush=True)
    spots, cnt = _spot_lists(300, 700, 512, 512, 480, seed=4, edge=0.02, dup=0.01)
    ref, out, st_ref, st, listed = _render_both(spots, cnt, 700, (512, 512), 1.0)
    _assert_same(ref, out, st_ref, st)
    n_long = int((cnt > 512).sum().item())
    assert n_long > 0 and listed >= n_long


def test_degenerate_lists():
    import torch
    print("bands: empty, out-of-frame, zero-intensity and single-spot lists", flush=True)
    spots, cnt = _spot_lists(64, 32, 128, 128, 20, seed=5)
    cnt[0:8] = 0                          # no spot
    spots[8:16, :, 0] += 1000.0           # every spot out of frame
    spots[16:24, :, 3] = 0.0              # only zero intensities (maximum 0)
    cnt[24:32] = 1                        # one spot
    spots[32:40, :, 3] = 2.5              # equal intensities
    ref, out, st_ref, st, _ = _render_both(spots, cnt, 32, (128, 128), 1.0)
    _assert_same(ref, out, st_ref, st)
    assert torch.equal(out[0:16], torch.zeros_like(out[0:16]))


@pytest.mark.parametrize("shape", [(100, 132), (250, 36), (17, 1020)])
def test_partial_tiles_and_counts_above_capacity(shape):
    # sides that are not multiples of the 16-pixel tile (partial tiles in the
    # bounds and in the candidate evaluation), and counts above the list
    # capacity (only the first cap spots exist)
    import torch
    H, W = shape
    print(f"bands: 300 x {H}x{W}, partial tiles", flush=True)
    spots, cnt = _spot_lists(300, 96, H, W, 60, seed=20 + W, edge=0.2, dup=0.05)
    cnt[::7] = 96 + 40
    ref, out, st_ref, st, listed = _render_both(spots, cnt, 96, shape, 1.0)
    _assert_same(ref, out, st_ref, st)
    assert listed < 300


def test_dense_rows_read_the_row_sorted_list():
    # most spots in a 12-row stripe: bands with more spots than their slot
    # holds take the band's window of the row-sorted list instead
    print("bands: 200 x 256x256, 400 spots, most in 12 rows", flush=True)
    spots, cnt = _spot_lists(200, 400, 256, 256, 380, seed=12, dup=0.02)
    rng = np.random.default_rng(13)
    stripe = rng.random((200, 400)) < 0.8
    y = rng.uniform(120.0, 132.0, (200, 400)).astype(np.float32) - 128.0
    import torch
    spots[:, :, 1] = torch.where(torch.from_numpy(stripe).cuda(), torch.from_numpy(y).cuda(),
                                 spots[:, :, 1])
    ref, out, st_ref, st, listed = _render_both(spots, cnt, 400, (256, 256), 1.0)
    _assert_same(ref, out, st_ref, st)
    assert listed < 200


def test_rounds_of_patterns():
    # five plan/writer rounds of at most 16,384 patterns (the last partial),
    # each reusing the workspace after the previous round's bands are written
    print("bands: 70000 x 32x32", flush=True)
    spots, cnt = _spot_lists(70000, 8, 32, 32, 4, seed=6, dup=0.05)
    ref, out, st_ref, st, _ = _render_both(spots, cnt, 8, (32, 32), 1.0)
    _assert_same(ref, out, st_ref, st)


@pytest.mark.parametrize("shape,sigma", [((128, 128), 3.0), ((128, 130), 1.0)])
def test_shapes_outside_the_band_path(shape, sigma):
    # tap radius 12, and a width that is not a multiple of 4: the splat kernel alone
    print(f"bands: 50 x {shape}, sigma {sigma}", flush=True)
    spots, cnt = _spot_lists(50, 64, shape[0], shape[1], 40, seed=7)
    ref, out, st_ref, st, _ = _render_both(spots, cnt, 64, shape, sigma)
    _assert_same(ref, out, st_ref, st)


def test_workspace_query_and_rejection():
    import torch
    lib, L = _lib()
    from diffsims_amd import engine
    assert L.dsx_render_fast_bands_workspace_ints(-1, 8, 8) == -1
    assert L.dsx_render_fast_bands_workspace_ints(10, 256, 256) > 0
    spots, cnt = _spot_lists(4, 8, 64, 64, 4, seed=8)
    p = engine.render_params((64, 64), 1.0, 1.0)
    out = torch.empty((4, 64, 64), dtype=torch.float32, device="cuda")
    work = torch.empty(16, dtype=torch.int32, device="cuda")
    rc = L.dsx_render_fast_bands(lib.ptr(spots), lib.ptr(cnt), 8, 4, ctypes.byref(p), lib.ptr(work),
                                 16, lib.ptr(out), None, lib.stream_handle(spots.device))
    assert rc == lib.DSX_EINVAL
    assert b"workspace" in L.dsx_last_error()


def test_render_spots_same_bits_across_the_small_batch_threshold():
    # render_spots takes the splat kernel alone below 768 patterns and the
    # band renderer above: a batch and its leading part give the same pixels
    import torch
    from diffsims_amd import engine
    print("render_spots: 800 and 700 x 256x256 across the dispatch threshold", flush=True)
    spots, cnt = _spot_lists(800, 400, 256, 256, 260, seed=30, edge=0.05, dup=0.02)
    assert engine.fast_renderer(800, 256, 256) == "dsx_render_fast_bands"
    assert engine.fast_renderer(700, 256, 256) == "dsx_render_fast"
    big = engine.render_spots(spots, cnt, 400, (256, 256), 1.0, 1.0)
    small = engine.render_spots(spots[:700].contiguous(), cnt[:700].contiguous(), 400, (256, 256),
                                1.0, 1.0)
    torch.cuda.synchronize()
    assert torch.equal(big[:700], small)
