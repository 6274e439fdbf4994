"""The no-allocation render (us3d_ug_render_static) through the C ABI: its image, alpha, depth and aux equal those
of us3d_ug_map + us3d_ug_forward with a caller workspace, bit for bit, on two calls in a row (the second with a
smaller batch, so the static scratch holds the first call's contents when it starts), with the frames' own grids
and with explicit pixels; a batch whose workspace cannot fit is refused before any device work."""

import ctypes

import numpy as np
import pytest
import torch

from tests.ug_helpers import cfg_for, small_case
from us3d import _lib
from us3d.batch import CapacityHint, SweepSpec
from us3d.scenes import linear_pixels
from us3d.ultragauss import UGBatch

pytestmark = pytest.mark.gpu


def _render_both(dev, scene, frames, cfg, explicit):
    G = scene.means.shape[0]
    b = UGBatch(frames, [SweepSpec(0, G)], G, cfg.settings(), dev, CapacityHint())
    p = [t.to(dev).float().contiguous() for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    pix = torch.cat([linear_pixels(f) for f in frames]).to(dev).contiguous() if explicit else None
    m = b.map(*p)
    ref = b.forward(m, pix)  # image, alpha, depth, aux with a caller workspace
    n = b.layout.n_pixels
    outs = [torch.full((n,), -7.0, device=dev) for _ in range(3)] + [torch.full((n, 4), -7.0, device=dev)]
    overflow = torch.full((1,), -1, dtype=torch.int32, device=dev)
    plan = b.tables.plan(torch.empty(1, dtype=torch.uint8, device=dev), 1, 1)
    plan.workspace, plan.workspace_bytes, plan.k_capacity, plan.v_capacity = None, 0, 0, 0  # ignored by the static variant
    L = _lib.lib()
    _lib.check(L.us3d_ug_render_static(plan, *[_lib.ptr(t) for t in p], _lib.ptr(pix), b.settings,
                                       *[_lib.ptr(t) for t in outs], _lib.ptr(overflow), _lib.stream_ptr(dev)),
               "us3d_ug_render_static")
    torch.cuda.synchronize()
    return ref, outs, int(overflow.item()), int(m.stats.n_instances)


def test_render_static_equals_map_and_forward_on_two_calls(dev):
    scene, frames = small_case(n=4096, n_frames=3)
    cfg = cfg_for(frames[0])
    ref, got, ovf, K = _render_both(dev, scene, frames, cfg, explicit=False)
    print(f"[no-alloc] call 1: 3 frames 128x128, 4096 Gaussians, grid pixels, K={K}", flush=True)
    assert ovf == 0 and K > 1000
    assert float(ref[0].abs().max()) > 0.05
    for r, g in zip(ref, got):
        assert torch.equal(r, g)
    small, frames2 = small_case(n=1500, n_frames=1, seed=5, H=96, W=80)
    ref2, got2, ovf2, K2 = _render_both(dev, small, frames2, cfg_for(frames2[0]), explicit=True)
    print(f"[no-alloc] call 2: 1 frame 96x80, 1500 Gaussians, explicit pixels, K={K2}", flush=True)
    assert ovf2 == 0 and 0 < K2 < K
    assert float(ref2[0].abs().max()) > 0.05
    for r, g in zip(ref2, got2):
        assert torch.equal(r, g)


def test_render_static_refuses_a_batch_beyond_the_scratch(dev):
    p = _lib.Plan()
    p.n_gauss, p.n_frames, p.n_sweeps, p.n_tiles, p.n_pairs, p.n_pixels = 1 << 20, 512, 1, 1 << 18, 1 << 29, 1 << 27
    dummy = ctypes.create_string_buffer(64)
    for f in ("frames", "frame_tile_base", "frame_pixel_base", "sweeps"):
        setattr(p, f, ctypes.addressof(dummy))
    st = _lib.UGSettings(chi2_threshold=7.815, diag_beta=1e-4, tile_size=16)
    L = _lib.lib()
    assert L.us3d_ug_render_static(p, None, None, None, None, None, st, None, None, None, None, None, None) == 3
    assert b"static scratch" in L.us3d_last_error()
