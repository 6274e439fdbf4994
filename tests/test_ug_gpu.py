"""Parity of the HIP slice splat (libus3d via the C ABI) with the oracle.

Bars: tile keys / sort order / tile ranges / pixel lists bit-exact; FP32 image, alpha and depth within 1e-5
relative L2; gradients within 1e-4 relative L2 of the FP64 oracle (per parameter group).
"""

import numpy as np
import pytest
import torch

from oracle import fallback as ofb
from oracle import ug as oug
from tests.parity import note
from tests.ug_helpers import cfg_for, gpu_render, oracle_frame, rel_l2, small_case
from us3d.batch import CapacityHint, SweepSpec
from us3d.scenes import linear_pixels, make_scene, straight_sweep

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.mark.parametrize("ts", [16, 8, 32])
@pytest.mark.parametrize("sort_by_depth", [True, False])
def test_mapper_keys_sort_ranges_bitexact(dev, ts, sort_by_depth):
    scene, frames = small_case(n_frames=3)
    for fr in frames:
        out = gpu_render(scene, [fr], cfg_for(fr, tile_size=ts, mapper_sort_by_depth=sort_by_depth), dev)
        view = out["batch"].sorted_view(out["batch"].map(*[t.to(dev) for t in (
            scene.means, scene.tril, scene.opacity, scene.intensity)]))[0]
        _g, _p, mr, _l = oracle_frame(scene, fr, oug.Semantics(), ts=ts, sort_by_depth=sort_by_depth)
        assert mr.keys.size > 100
        np.testing.assert_array_equal(view["keys"].cpu().numpy().astype(np.uint64), mr.keys)
        np.testing.assert_array_equal(view["values"].cpu().numpy(), mr.vals)
        np.testing.assert_array_equal(view["tile_start"].cpu().numpy(), mr.tile_start)
        np.testing.assert_array_equal(view["tile_end"].cpu().numpy(), mr.tile_end)


def test_pixel_lists_bitexact(dev):
    scene, frames = small_case(n_frames=2)
    for fr in frames:
        out = gpu_render(scene, [fr], cfg_for(fr), dev)
        b = out["batch"]
        m = b.map(*[t.to(dev) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)])
        counts, offsets, indices = b.pixel_lists(m, linear_pixels(fr).to(dev))
        _g, _p, _mr, (c_ref, o_ref, i_ref) = oracle_frame(scene, fr, oug.Semantics())
        np.testing.assert_array_equal(counts.cpu().numpy(), c_ref)
        np.testing.assert_array_equal(offsets.cpu().numpy(), o_ref)
        np.testing.assert_array_equal(indices.cpu().numpy(), i_ref)


@pytest.mark.parametrize("ts", [16, 8, 32])
def test_forward_backward_parity_cuda_semantics(dev, ts):
    scene, frames = small_case(n_frames=2)
    rng = np.random.default_rng(5)
    sem = oug.Semantics()
    for fr in frames:
        P = fr.height * fr.width
        dI = rng.normal(size=P).astype(np.float32)
        dA = rng.normal(size=P).astype(np.float32) * 0.1
        dD = rng.normal(size=P).astype(np.float32) * 0.01
        out = gpu_render(scene, [fr], cfg_for(fr, tile_size=ts), dev, grads=(dI, dA, dD))
        geom, pix, _mr, lists = oracle_frame(scene, fr, sem)
        args = (scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(), scene.intensity.numpy(), pix, geom, lists,
                sem)
        img32, a32, d32, _ = oug.forward(*args)
        img64, a64, d64, _ = oug.forward(*args, fp64=True)
        got = {k: out[k].detach().cpu().numpy() for k in ("image", "alpha", "depth")}
        case = "UltraGauss, explicit pixels, tiles 8/16/32"
        assert note(case, "image vs FP32", rel_l2(got["image"], img32), IMG_TOL) < IMG_TOL
        assert note(case, "alpha vs FP32", rel_l2(got["alpha"], a32), IMG_TOL) < IMG_TOL
        assert note(case, "depth vs FP32", rel_l2(got["depth"], d32), IMG_TOL) < IMG_TOL
        assert note(case, "image vs FP64", rel_l2(got["image"], img64), IMG_TOL) < IMG_TOL
        note(case, "alpha vs FP64", rel_l2(got["alpha"], a64), IMG_TOL)
        note(case, "depth vs FP64", rel_l2(got["depth"], d64), IMG_TOL)
        ref = oug.backward_f64(*args, dI, dA, dD, gate=True)
        for name, g, r in zip(("means", "tril", "opacity", "intensity"), out["grads"], ref):
            err = note(case, f"d{name} vs FP64", rel_l2(g, r), GRAD_TOL)
            assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"
        # the reference's backward omits the alpha_hat > 1e-8 gate; the difference is below the bar too
        ref_q = oug.backward_f64(*args, dI, dA, dD, gate=False)
        for g, r in zip(out["grads"], ref_q):
            assert rel_l2(g, r) < GRAD_TOL


def test_fallback_semantics_vs_torch_fallback(dev):
    """The named oracle: the reference's PyTorch fallback renderer, restated, in float64 with autograd."""
    scene = make_scene(512, 64.0, seed=3)
    fr = straight_sweep(3, 48, 40, 0.3, 64.0, 10.0, 5.0, seed=4)[1]
    P = fr.height * fr.width
    rng = np.random.default_rng(1)
    dI = rng.normal(size=P).astype(np.float32)
    out = gpu_render(scene, [fr], cfg_for(fr, semantics="fallback"), dev, grads=(dI, None, None))
    ts = [t.double().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    img, alpha = ofb.render(*ts, linear_pixels(fr).double(), fr.height, fr.width)
    (img.reshape(-1) * torch.from_numpy(dI).double()).sum().backward()
    case = "UltraGauss fallback semantics vs torch fallback"
    assert note(case, "image vs FP64", rel_l2(out["image"].detach().cpu().numpy(), img.detach().numpy()),
                IMG_TOL) < IMG_TOL
    assert note(case, "alpha vs FP64", rel_l2(out["alpha"].detach().cpu().numpy(), alpha.detach().numpy()),
                IMG_TOL) < IMG_TOL
    for name, g, t in zip(("means", "tril", "opacity", "intensity"), out["grads"], ts):
        err = note(case, f"d{name} vs FP64", rel_l2(g, t.grad.numpy()), GRAD_TOL)
        assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"


def test_batch_equals_per_frame_for_explicit_and_grid_pixels(dev):
    """Batched equals per frame bit for bit, with explicit pixel positions and with the frames' own grids."""
    scene, frames = small_case(n_frames=4)
    cfg = cfg_for(frames[0])
    rng = np.random.default_rng(2)
    P = sum(f.height * f.width for f in frames)
    dI = rng.normal(size=P).astype(np.float32)
    for pixels in ("explicit", "procedural"):
        batched = gpu_render(scene, frames, cfg, dev, pixels=pixels, grads=(dI, None, None))
        off = 0
        acc = [np.zeros_like(g) for g in batched["grads"]]
        for fr in frames:
            n = fr.height * fr.width
            single = gpu_render(scene, [fr], cfg, dev, pixels=pixels, grads=(dI[off:off + n], None, None))
            np.testing.assert_array_equal(single["image"].detach().cpu().numpy(),
                                          batched["image"].detach().cpu().numpy()[off:off + n])
            acc = [a + g for a, g in zip(acc, single["grads"])]
            off += n
        for a, b in zip(acc, batched["grads"]):
            assert rel_l2(b, a) < 1e-6


@pytest.mark.parametrize("ts", [16, 8, 32])
def test_grid_pixels_plane_form_parity(dev, ts):
    """Frames rendered on their own grids (pixels = NULL) evaluate q in the plane form of each pair. The plane
    coefficients the preprocess leaves in the workspace equal the oracle's bit for bit for every visible pair; the
    image, alpha and depth match the oracle's plane-form forward within 1e-5 (FP32 and FP64), and the gradients
    its FP64 adjoint within 1e-4, with and without alpha/depth gradients."""
    scene, frames = small_case(n_frames=2)
    rng = np.random.default_rng(15)
    sem = oug.Semantics()
    G = scene.means.shape[0]
    for k, fr in enumerate(frames):
        P = fr.height * fr.width
        dI = rng.normal(size=P).astype(np.float32)
        dA = rng.normal(size=P).astype(np.float32) * 0.1
        dD = rng.normal(size=P).astype(np.float32) * 0.01
        full = k == 1
        cfg = cfg_for(fr, tile_size=ts)
        out = gpu_render(scene, [fr], cfg, dev, pixels="procedural", grads=(dI, dA if full else None,
                                                                             dD if full else None))
        b = out["batch"]
        m = b.map(*[t.to(dev) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)])
        geom, pix, mr, lists = oracle_frame(scene, fr, sem, ts=ts)
        plane = oug.plane_form(scene.means.numpy(), scene.tril.numpy(), geom)
        off = b.layout_offsets(m.k_capacity, m.v_capacity)
        vis = mr.count > 0
        assert vis.sum() > 100 and int(m.stats.n_visible) == int(vis.sum())
        # one frame: the visible pairs are its Gaussians with a tile, in index order
        V = int(vis.sum())
        sect = m.workspace[off.pair_sect: off.pair_sect + 64 * V].view(torch.float32).view(V, 16).cpu().numpy()
        got_plane = np.stack([sect[:, 8], sect[:, 9], sect[:, 10], sect[:, 4], sect[:, 5], sect[:, 6]], 1)
        np.testing.assert_array_equal(got_plane.view(np.uint32), plane[vis].view(np.uint32))
        np.testing.assert_array_equal(sect[:, 11].view(np.int32), np.flatnonzero(vis))  # each pair's Gaussian
        args = (scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(), scene.intensity.numpy(), pix, geom,
                lists, sem)
        img32, a32, d32, _ = oug.forward(*args, plane=plane)
        img64, a64, d64, _ = oug.forward(*args, fp64=True, plane=plane)
        got = {key: out[key].detach().cpu().numpy() for key in ("image", "alpha", "depth")}
        assert np.abs(img32).max() > 0.05
        case = "UltraGauss, frames' grids (plane form), tiles 8/16/32"
        for key, r32, r64 in (("image", img32, img64), ("alpha", a32, a64), ("depth", d32, d64)):
            assert note(case, f"{key} vs FP32", rel_l2(got[key], r32), IMG_TOL) < IMG_TOL, key
            assert note(case, f"{key} vs FP64", rel_l2(got[key], r64), IMG_TOL) < IMG_TOL, key
        ref = oug.backward_f64(*args, dI, dA if full else None, dD if full else None, gate=True, plane=plane)
        for name, g, r in zip(("means", "tril", "opacity", "intensity"), out["grads"], ref):
            err = note(case, f"d{name} vs FP64", rel_l2(g, r), GRAD_TOL)
            assert err < GRAD_TOL, f"ts {ts} full {full} {name}: rel L2 {err:.3e}"


@pytest.mark.parametrize("ts", [16, 8, 32])
def test_lds_batch_sizes_give_the_same_image_and_gradients(dev, ts):
    """The splat kernels stage 64, 128 or 256 Gaussians per LDS batch (fwd_lds_batch, bwd_lds_batch; 0 = the
    work-group's thread count). Every size gives the same image, alpha, depth and aux bit for bit (each pixel sums its Gaussians in list
    order whatever the batching), on the frames' grids and from explicit pixels, and gradients that differ only by
    FP32 rounding of the per-Gaussian sums; one size meets the FP64 oracle within 1e-4. 30k Gaussians put 300-700
    into a tile's list, so every size takes several batches per tile."""
    from dataclasses import replace

    scene, frames = small_case(n=30000, n_frames=2)
    rng = np.random.default_rng(16)
    P = sum(f.height * f.width for f in frames)
    grads = (rng.normal(size=P).astype(np.float32), rng.normal(size=P).astype(np.float32) * 0.1, None)
    cfg = cfg_for(frames[0], tile_size=ts)
    sem = oug.Semantics()
    for pixels in ("procedural", "explicit"):
        ref = gpu_render(scene, frames, cfg, dev, pixels=pixels, grads=grads)
        assert int(ref["batch"].map(*[t.to(dev) for t in (scene.means, scene.tril, scene.opacity,
                                                           scene.intensity)]).stats.max_segment) > 256
        for nb in (64, 128, 256):
            out = gpu_render(scene, frames, replace(cfg, fwd_lds_batch=nb, bwd_lds_batch=nb), dev, pixels=pixels,
                             grads=grads)
            for key in ("image", "alpha", "depth"):
                assert torch.equal(out[key], ref[key]), f"{pixels} lds_batch {nb}: {key} differs"
            for name, a, b in zip(("means", "tril", "opacity", "intensity"), out["grads"], ref["grads"]):
                assert rel_l2(a, b) < 1e-5, f"{pixels} lds_batch {nb}: {name} rel L2 {rel_l2(a, b):.3e}"
    # one size against the oracle, frame by frame (explicit pixels: the oracle's general form of q)
    off = 0
    out = gpu_render(scene, frames, replace(cfg, fwd_lds_batch=64, bwd_lds_batch=64), dev, pixels="explicit",
                     grads=(grads[0], None, None))
    acc = [np.zeros_like(g) for g in out["grads"]]
    for fr in frames:
        n = fr.height * fr.width
        geom, pix, mr, lists = oracle_frame(scene, fr, sem, ts=ts)
        args = (scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(), scene.intensity.numpy(), pix, geom,
                lists, sem)
        r = oug.backward_f64(*args, grads[0][off:off + n], None, None, gate=True)
        acc = [a + np.asarray(x, np.float64) for a, x in zip(acc, r)]
        off += n
    for name, g, r in zip(("means", "tril", "opacity", "intensity"), out["grads"], acc):
        assert rel_l2(g, r) < GRAD_TOL, f"ts {ts} {name}: rel L2 {rel_l2(g, r):.3e}"


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_work_queue_launch_equals_one_workgroup_per_tile(dev, ts):
    """work_queue: persistent workgroups (as many as the device holds) take the (frame, tile) items from per-XCD
    queues instead of one workgroup per tile. 8 frames of 512x512 give 2k-33k tiles, several per resident workgroup,
    some empty. Every tile is rendered once by some workgroup, so the image, alpha, depth and every gradient equal
    the one-workgroup-per-tile launch bit for bit, on the frames' grids and from explicit pixels; a captured training
    step with the queues replays equal to the eager step without them."""
    from dataclasses import replace

    from us3d.train_step import SliceSplatTrainer

    scene, frames = small_case(n=30000, n_frames=8, H=512, W=512, spacing=0.1)
    rng = np.random.default_rng(21)
    P = sum(f.height * f.width for f in frames)
    grads = (rng.normal(size=P).astype(np.float32), rng.normal(size=P).astype(np.float32) * 0.1, None)
    cfg = cfg_for(frames[0], tile_size=ts)
    for pixels in ("procedural", "explicit"):
        ref = gpu_render(scene, frames, replace(cfg, work_queue=False), dev, pixels=pixels, grads=grads)
        out = gpu_render(scene, frames, replace(cfg, work_queue=True), dev, pixels=pixels, grads=grads)
        assert float(ref["image"].detach().abs().max()) > 0
        for key in ("image", "alpha", "depth"):
            assert torch.equal(out[key], ref[key]), f"{pixels}: {key} differs"
        for name, a, b in zip(("means", "tril", "opacity", "intensity"), out["grads"], ref["grads"]):
            assert np.array_equal(a, b), f"{pixels}: d{name} differs"
    if ts != 16:
        return
    with torch.no_grad():
        gt = gpu_render(make_scene(30000, 64.0, seed=5), frames, cfg, dev, pixels="procedural")["image"].detach()
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]
    a = SliceSplatTrainer(*init, frames, gt, replace(cfg, work_queue=True), use_graph=True)
    b = SliceSplatTrainer(*init, frames, gt, replace(cfg, work_queue=False), use_graph=False)
    for _ in range(3):
        a.step()
        b.step()
        torch.cuda.synchronize()
        for x, y in zip(a.state(), b.state()):
            assert torch.equal(x, y)
        assert torch.equal(a.losses(), b.losses())


def test_bitwise_deterministic(dev):
    scene, frames = small_case(n_frames=3)
    cfg = cfg_for(frames[0])
    P = sum(f.height * f.width for f in frames)
    dI = np.random.default_rng(3).normal(size=P).astype(np.float32)
    a = gpu_render(scene, frames, cfg, dev, grads=(dI, None, None))
    b = gpu_render(scene, frames, cfg, dev, grads=(dI, None, None))
    for x, y in zip(a["grads"], b["grads"]):
        np.testing.assert_array_equal(x, y)


def test_capacity_overflow_retry(dev):
    scene, frames = small_case(n_frames=2)
    cfg = cfg_for(frames[0])
    ref = gpu_render(scene, frames, cfg, dev)
    hint = CapacityHint(floor=16, growth=1.25)
    hint.k_last = 0
    tiny = gpu_render(scene, frames, cfg, dev, hint=hint)
    np.testing.assert_array_equal(ref["image"].detach().cpu().numpy(), tiny["image"].detach().cpu().numpy())


def test_empty_and_invisible_scenes(dev):
    scene, frames = small_case(n=64, n_frames=1)
    scene.means[:] = 1.0e4  # far outside every slice
    fr = frames[0]
    P = fr.height * fr.width
    out = gpu_render(scene, [fr], cfg_for(fr), dev, grads=(np.ones(P, np.float32), None, None))
    assert float(out["image"].detach().abs().max()) == 0.0
    np.testing.assert_array_equal(out["alpha"].detach().cpu().numpy(), np.float32(1e-6))
    for g in out["grads"]:
        assert np.all(g == 0)


def test_nonfinite_gaussians_are_culled(dev):
    scene, frames = small_case(n=256, n_frames=1)
    fr = frames[0]
    clean = gpu_render(scene, [fr], cfg_for(fr), dev)
    scene.means[7] = float("nan")
    scene.tril[11, 0] = float("inf")
    scene.opacity[7] = 0.0
    out = gpu_render(scene, [fr], cfg_for(fr), dev)
    assert torch.isfinite(out["image"]).all()
    geom, pix, _mr, lists = oracle_frame(scene, fr, oug.Semantics())
    img, _a, _d, _s = oug.forward(scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(),
                                  scene.intensity.numpy(), pix, geom, lists, oug.Semantics())
    assert rel_l2(out["image"].detach().cpu().numpy(), img) < IMG_TOL
    assert clean["image"].shape == out["image"].shape


def test_config2_frame_parity(dev):
    """One frame of the 200k-Gaussian, 512x512 freehand workload against the oracle (FP32 image, FP64 grads)."""
    from us3d.scenes import config2

    scene, _target, frames = config2(n_frames=256)
    fr = frames[128]
    P = fr.height * fr.width
    dI = np.random.default_rng(9).normal(size=P).astype(np.float32) / P
    out = gpu_render(scene, [fr], cfg_for(fr), dev, pixels="procedural", grads=(dI, None, None))
    sem = oug.Semantics()
    geom, pix, mr, lists = oracle_frame(scene, fr, sem)
    plane = oug.plane_form(scene.means.numpy(), scene.tril.numpy(), geom)  # rendered on the frame's own grid
    args = (scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(), scene.intensity.numpy(), pix, geom, lists,
            sem)
    img, _a, _d, _s = oug.forward(*args, plane=plane)
    img64, _a, _d, _s = oug.forward(*args, fp64=True, plane=plane)
    assert mr.keys.size > 100_000
    case = "UltraGauss config-2 frame (200k Gaussians, 512x512)"
    got = out["image"].detach().cpu().numpy()
    assert note(case, "image vs FP32", rel_l2(got, img), IMG_TOL) < IMG_TOL
    assert note(case, "image vs FP64", rel_l2(got, img64), IMG_TOL) < IMG_TOL
    ref = oug.backward_f64(*args, dI, None, None, gate=True, plane=plane)
    # the reference's own precision: its backward restated in FP32 (naive per-pair accumulation)
    ref32 = oug.backward_f64(*args, dI, None, None, gate=True, fp32=True, plane=plane)
    for name, g, r, r32 in zip(("means", "tril", "opacity", "intensity"), out["grads"], ref, ref32):
        err = note(case, f"d{name} vs FP64", rel_l2(g, r), GRAD_TOL)
        note("UltraGauss config-2 frame: FP32 restatement", f"d{name} vs FP64", rel_l2(r32, r), GRAD_TOL)
        assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"


def test_multi_sweep_batch_equals_per_sweep(dev):
    """Sweeps of uneven Gaussian count, frame count and frame size rendered as ONE batch give bit-identical
    images and parameter gradients to rendering every sweep on its own."""
    from dataclasses import replace

    from us3d.scenes import multi_sweep
    from us3d.ultragauss import render_frames

    specs = [(700, 3, 64, 48), (300, 2, 32, 80), (1200, 1, 48, 48)]
    scene, _target, frames, sweeps = multi_sweep(specs, cube_mm=64.0, seed=7, length_mm=30.0, spacing=0.3)
    assert [f.sweep for f in frames] == [0, 0, 0, 1, 1, 2]
    cfg = cfg_for(frames[0])
    P = sum(f.height * f.width for f in frames)
    dI = torch.from_numpy(np.random.default_rng(4).normal(size=P).astype(np.float32)).to(dev)
    params = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity,
                                                                 scene.intensity)]
    image, _alpha, _depth = render_frames(*params, None, frames, cfg, sweeps=sweeps)
    assert float(image.detach().abs().max()) > 0
    (image * dI).sum().backward()
    off = 0
    for s, sw in enumerate(sweeps):
        fs = [replace(f, sweep=0) for f in frames if f.sweep == s]
        sub = [t.detach()[sw.g_begin:sw.g_begin + sw.g_count].clone().requires_grad_(True) for t in params]
        img_s, _a, _d = render_frames(*sub, None, fs, cfg, sweeps=[SweepSpec(0, sw.g_count)])
        n = sum(f.height * f.width for f in fs)
        assert torch.equal(img_s, image.detach()[off:off + n]), f"sweep {s}: image differs"
        (img_s * dI[off:off + n]).sum().backward()
        for a, b in zip(sub, params):
            assert torch.equal(a.grad, b.grad[sw.g_begin:sw.g_begin + sw.g_count]), f"sweep {s}: gradient differs"
        off += n


def test_sweeps_that_leave_gaussians_out(dev):
    """A batch whose sweep table leaves Gaussians out (here only the middle sweep's frames and range, so rows
    before and after it belong to no sweep of the batch, as in a chunk of a training step): the image equals that
    sweep rendered on its own bit for bit, its Gaussians' gradients equal that render's, and every Gaussian outside
    the range gets exactly zero (the per-Gaussian reduction finds no sweep for it and never indexes the pair
    table)."""
    from dataclasses import replace

    from us3d.scenes import multi_sweep
    from us3d.ultragauss import render_frames

    specs = [(700, 2, 64, 48), (300, 2, 32, 80), (500, 1, 48, 48)]
    scene, _target, frames, sweeps = multi_sweep(specs, cube_mm=64.0, seed=9, length_mm=30.0, spacing=0.3)
    cfg = cfg_for(frames[0])
    sw = sweeps[1]
    fs = [replace(f, sweep=0) for f in frames if f.sweep == 1]
    P = sum(f.height * f.width for f in fs)
    dI = torch.from_numpy(np.random.default_rng(5).normal(size=P).astype(np.float32)).to(dev)
    full = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    image, _a, _d = render_frames(*full, None, fs, cfg, sweeps=[sw])
    (image * dI).sum().backward()
    sub = [t.detach()[sw.g_begin:sw.g_begin + sw.g_count].clone().requires_grad_(True) for t in full]
    img_s, _a, _d = render_frames(*sub, None, fs, cfg, sweeps=[SweepSpec(0, sw.g_count)])
    (img_s * dI).sum().backward()
    assert float(img_s.detach().abs().max()) > 0.05
    assert torch.equal(image.detach(), img_s.detach())
    for a, b in zip(full, sub):
        assert float(b.grad.abs().max()) > 0
        assert torch.equal(a.grad[sw.g_begin:sw.g_begin + sw.g_count], b.grad)
        assert not bool(a.grad[:sw.g_begin].any()) and not bool(a.grad[sw.g_begin + sw.g_count:].any())


@pytest.mark.parametrize("ts", [16, 8])
def test_mapper_long_tile_lists_bitexact(dev, ts):
    """Dense clusters put ~700, ~2500 and ~6000 Gaussians into single tiles, so every tile-sort path runs
    (one-wavefront register sorts, the two-wavefront sort of 513..1024 keys, the workgroup merge sort of up to
    4096 keys, the chunked merge beyond); keys, order and ranges must still equal the oracle's bit for bit."""
    from us3d.scenes import Scene, precision_tril_from_scales_quats

    fr = straight_sweep(1, 128, 128, 0.3, 64.0, 10.0, 0.0, seed=1)[0]
    rng = np.random.default_rng(8)
    parts = []
    for n, (u, v) in zip((700, 2500, 6000), ((20, 30), (70, 64), (100, 100))):
        lat = (u - fr.col_offset) * fr.delta_lat + rng.normal(0.0, 0.4, n)
        ax = fr.near_mm + v * fr.delta_ax + rng.normal(0.0, 0.4, n)
        ele = rng.normal(0.0, 0.1, n)
        o, a_l, a_a, a_e = (f.numpy().astype(np.float64) for f in (fr.origin, fr.axis_lat, fr.axis_ax, fr.axis_ele))
        parts.append(o + lat[:, None] * a_l + ax[:, None] * a_a + ele[:, None] * a_e)
    means = torch.from_numpy(np.concatenate(parts).astype(np.float32))
    G = means.shape[0]
    g = torch.Generator().manual_seed(3)
    scales = 0.3 * torch.exp(0.2 * torch.randn((G, 3), generator=g))
    quats = torch.randn((G, 4), generator=g)
    quats = quats / quats.norm(dim=1, keepdim=True)
    scene = Scene(means, scales, quats, torch.full((G,), 0.5), torch.full((G,), 0.5),
                  precision_tril_from_scales_quats(scales, quats, 1e-4))
    out = gpu_render(scene, [fr], cfg_for(fr, tile_size=ts), dev, pixels="procedural")
    view = out["batch"].sorted_view(out["batch"].map(*[t.to(dev) for t in (
        scene.means, scene.tril, scene.opacity, scene.intensity)]))[0]
    _g, _p, mr, _l = oracle_frame(scene, fr, oug.Semantics(), ts=ts)
    seg = (mr.tile_end - mr.tile_start)[mr.tile_start >= 0]
    assert seg.max() > 4096 and ((seg > 1024) & (seg <= 4096)).any() and ((seg > 512) & (seg <= 1024)).any()
    np.testing.assert_array_equal(view["keys"].cpu().numpy().astype(np.uint64), mr.keys)
    np.testing.assert_array_equal(view["values"].cpu().numpy(), mr.vals)
    np.testing.assert_array_equal(view["tile_start"].cpu().numpy(), mr.tile_start)
    np.testing.assert_array_equal(view["tile_end"].cpu().numpy(), mr.tile_end)


@pytest.mark.parametrize("shadow_delta", [0.0, 0.05])
def test_shadow_branch_attenuates_rows_by_running_transmittance(dev, shadow_delta):
    """UltraGaussRenderer with enable_shadow (api.py:37-42): image * T with T = exp(-cumsum(alpha * delta))
    down the rows, delta = shadow_delta or 1/H. Checked against a running product of per-row factors in
    FP64 on the CPU from the unshadowed render, and the shadowed render must stay differentiable."""
    from us3d.ultragauss import SliceGeometry, UltraGaussRenderer

    scene, frames = small_case(n=2048, n_frames=1, H=64, W=48)
    fr = frames[0]
    geo = SliceGeometry(fr.origin.to(dev), fr.axis_lat.to(dev), fr.axis_ax.to(dev), fr.axis_ele.to(dev))
    pix = linear_pixels(fr).to(dev)
    args = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    plain = UltraGaussRenderer(cfg_for(fr))
    shadow = UltraGaussRenderer(cfg_for(fr, enable_shadow=True, shadow_delta=shadow_delta))
    img0, alpha0, _d0, _ = plain(*args, pix, geo)
    img1, alpha1, _d1, _ = shadow(*args, pix, geo)
    a = alpha0.detach().cpu().double().numpy()
    assert a.max() > 0.5
    delta = shadow_delta if shadow_delta > 0 else 1.0 / fr.height
    T = np.ones(fr.width)
    want = np.empty_like(a)
    for r in range(fr.height):
        T = T * np.exp(-a[r] * delta)
        want[r] = img0.detach().cpu().double().numpy()[r] * T
    np.testing.assert_array_equal(alpha1.detach().cpu().numpy(), alpha0.detach().cpu().numpy())
    assert rel_l2(img1.detach().cpu().numpy(), want) < 1e-6
    assert float((img0 - img1).detach().abs().max()) > 1e-3
    img1.sum().backward()
    for t in args:
        assert t.grad is not None and torch.isfinite(t.grad).all()
    assert float(args[2].grad.abs().max()) > 0


@pytest.mark.parametrize("semantics", ["cuda", "fallback"])
def test_fan_frames_match_brute_force_oracle(dev, semantics):
    """Fan (sector) frames, geometry 1: the mapper assigns tiles from the polar extent of each Gaussian's
    marginal box. Rendered from explicit fan pixel positions, image, alpha and gradients must equal the oracle's
    evaluation of EVERY (pixel, Gaussian) pair with the same canonical decisions (each pixel lists all
    Gaussians), so no contributing pair may be missing from the tile lists. Two frames in one batch."""
    from us3d.scenes import fan_pixels, fan_sweep
    from us3d.ultragauss import render_frames

    box = [64.0, 64.0, 64.0]
    scene = make_scene(1200, 64.0, seed=21, box_mm=box)
    frames = fan_sweep(2, 48, 64, 60.0, 8.0, 56.0, box, 16.0, seed=5)
    sem = oug.Semantics() if semantics == "cuda" else oug.Semantics.fallback()
    cfg = cfg_for(frames[0], semantics=semantics)
    pix = [fan_pixels(f) for f in frames]
    P = [f.height * f.width for f in frames]
    rng = np.random.default_rng(6)
    dI = rng.normal(size=sum(P)).astype(np.float32)
    dA = (rng.normal(size=sum(P)) * 0.1).astype(np.float32)
    ts = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    image, alpha, _depth = render_frames(*ts, torch.cat(pix).to(dev), frames, cfg,
                                         sweeps=[SweepSpec(0, scene.means.shape[0])], hint=CapacityHint())
    ((image * torch.from_numpy(dI).to(dev)).sum() + (alpha * torch.from_numpy(dA).to(dev)).sum()).backward()
    G = scene.means.shape[0]
    params = (scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(), scene.intensity.numpy())
    refs = [np.zeros((G, 3)), np.zeros((G, 6)), np.zeros(G), np.zeros(G)]
    off = 0
    for f, px in zip(frames, pix):
        n = f.height * f.width
        lists = (np.full(n, G, np.int32), (np.arange(n, dtype=np.int64) * G).astype(np.int32),
                 np.tile(np.arange(G, dtype=np.int32), n))
        geom = oug.FrameGeom.from_spec(f)
        args = params + (px.numpy(), geom, lists, sem)
        img32, a32, _d, _s = oug.forward(*args)
        assert np.abs(img32).max() > 1e-2
        got_i = image[off:off + n].detach().cpu().numpy()
        got_a = alpha[off:off + n].detach().cpu().numpy()
        case = f"UltraGauss fan frames vs all-pairs oracle ({semantics})"
        assert note(case, "image vs FP32", rel_l2(got_i, img32), IMG_TOL) < IMG_TOL
        assert note(case, "alpha vs FP32", rel_l2(got_a, a32), IMG_TOL) < IMG_TOL
        for acc, r in zip(refs, oug.backward_f64(*args, dI[off:off + n], dA[off:off + n], None, gate=True)):
            acc += r
        off += n
    for name, t, r in zip(("means", "tril", "opacity", "intensity"), ts, refs):
        assert np.abs(r).max() > 0, name
        err = note(case, f"d{name} vs FP64", rel_l2(t.grad.detach().cpu().numpy(), r), GRAD_TOL)
        assert err < GRAD_TOL, f"{semantics} {name}: rel L2 {err:.3e}"


@pytest.mark.parametrize("n_frames", [1, 20, 260])
def test_emit_split_modes_bitexact(dev, n_frames):
    """The emit splits a batch by its frame count: parts of the frame's visible-pair list with global cursors below
    16 frames, bands of tile rows with LDS cursors below 256, one workgroup per frame beyond. In every mode the sorted
    keys and tile ranges equal the oracle's for sampled frames. The visible pairs are exactly the pairs with a tile,
    ranked Gaussian-major (vis_pair increasing; counts, offsets and rectangles per rank), the frame-major list holds
    each frame's ranks in Gaussian order, and the emit indices formed from an instance's rank and tile are a
    permutation of [0, K) whose row-major position in the pair's tile rectangle is exactly the tile holding it."""
    scene = make_scene(2000, 64.0, seed=13)
    frames = straight_sweep(n_frames, 128, 48, 0.3, 64.0, 40.0, 5.0, seed=14)
    cfg = cfg_for(frames[0])
    out = gpu_render(scene, frames, cfg, dev, pixels="procedural")
    b = out["batch"]
    m = b.map(*[t.to(dev) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)])
    views = b.sorted_view(m)
    counts = {}
    for f in sorted({0, n_frames // 2, n_frames - 1}):
        _g, _p, mr, _l = oracle_frame(scene, frames[f], oug.Semantics())
        np.testing.assert_array_equal(views[f]["keys"].cpu().numpy().astype(np.uint64), mr.keys)
        np.testing.assert_array_equal(views[f]["values"].cpu().numpy(), mr.vals)
        np.testing.assert_array_equal(views[f]["tile_start"].cpu().numpy(), mr.tile_start)
        np.testing.assert_array_equal(views[f]["tile_end"].cpu().numpy(), mr.tile_end)
        counts[f] = mr.count
    lay = b.layout
    off = b.layout_offsets(m.k_capacity, m.v_capacity)
    ws = m.workspace.cpu()
    K = int(m.stats.n_instances)
    V = int(m.stats.n_visible)
    i32 = lambda o, n: ws[o: o + 4 * n].view(torch.int32).numpy().astype(np.int64)  # noqa: E731
    key = ws[off.seg_key: off.seg_key + 8 * K].view(torch.int64).numpy()
    val = key & 0xFFFFFFFF  # each instance's visible pair v (the key's low word)
    tile_off = i32(off.tile_off, lay.n_tiles + 1)
    vis_pair = i32(off.vis_pair, V)
    pair_count = i32(off.pair_count, V)
    pair_off = i32(off.pair_off, V + 1)
    rect = ws[off.pair_rect: off.pair_rect + 16 * V].view(torch.int32).numpy().reshape(-1, 4).astype(np.int64)
    vis_list = i32(off.vis_list, V)
    pair_g = i32(off.pair_g, V)
    assert K > 300 * min(n_frames, 3) and 0 < V <= K
    # ranks: dense pairs with a tile, Gaussian-major; the sampled frames' visible Gaussians are the oracle's
    assert np.all(np.diff(vis_pair) > 0) and vis_pair[-1] < lay.n_pairs
    vg, vf = vis_pair // n_frames, vis_pair % n_frames  # one sweep: pair = g * n_frames + frame
    for f, cnt in counts.items():
        np.testing.assert_array_equal(np.sort(vg[vf == f]), np.flatnonzero(cnt > 0))
    np.testing.assert_array_equal(pair_count, (rect[:, 2] - rect[:, 0]) * (rect[:, 3] - rect[:, 1]))
    assert np.all(pair_count > 0) and pair_off[0] == 0
    np.testing.assert_array_equal(np.diff(pair_off), pair_count)
    assert pair_off[V] == K
    # the frame-major list: frame by frame, each frame's ranks in Gaussian order
    order = np.lexsort((vg, vf))
    np.testing.assert_array_equal(vis_list, order)
    # every rank's Gaussian is its dense pair's; an instance lies in its pair's frame and rectangle, and the emit
    # indices from (rank, tile) are a permutation of [0, K)
    np.testing.assert_array_equal(pair_g, vg)
    tile = np.repeat(np.arange(lay.n_tiles), np.diff(tile_off))  # global tile of each instance
    frame = np.searchsorted(np.asarray(lay.frame_tile_base), tile, side="right") - 1
    np.testing.assert_array_equal(vf[val], frame)
    tiles_x = np.asarray(lay.tiles_x)[frame]
    lt = tile - np.asarray(lay.frame_tile_base)[frame]
    tx, ty = lt % tiles_x, lt // tiles_x
    r = rect[val]
    assert np.all((tx >= r[:, 0]) & (tx < r[:, 2]) & (ty >= r[:, 1]) & (ty < r[:, 3]))
    e = pair_off[val] + (ty - r[:, 1]) * (r[:, 2] - r[:, 0]) + (tx - r[:, 0])
    np.testing.assert_array_equal(np.sort(e), np.arange(K))
    # each visible pair's record carries its emit base (first emit index, x0, y0, width) for the adjoint
    base = ws[off.pair_sect: off.pair_sect + 64 * V].view(torch.int32).numpy().reshape(V, 16)[:, 12:16]
    np.testing.assert_array_equal(base, np.stack([pair_off[:V], rect[:, 0], rect[:, 1], rect[:, 2] - rect[:, 0]], 1))


def test_fan_grid_in_kernel_matches_explicit_positions(dev):
    """pixels = NULL on fan frames: the kernels evaluate the polar grid (sincos of the beam angle) themselves.
    It matches rendering from the host's restatement of that grid (fan_pixels) to within the last-bit rounding
    of sin/cos (image and gradients, relative L2 < 1e-5)."""
    from us3d.scenes import fan_pixels, fan_sweep
    from us3d.ultragauss import render_frames

    box = [64.0, 64.0, 64.0]
    scene = make_scene(1500, 64.0, seed=23, box_mm=box)
    frames = fan_sweep(3, 64, 80, 60.0, 8.0, 56.0, box, 20.0, seed=7)
    cfg = cfg_for(frames[0])
    dI = torch.from_numpy(np.random.default_rng(2).normal(size=sum(f.height * f.width for f in frames))
                          .astype(np.float32)).to(dev)
    outs = []
    for pix in (torch.cat([fan_pixels(f) for f in frames]).to(dev), None):
        ts = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
        image, _a, _d = render_frames(*ts, pix, frames, cfg, sweeps=[SweepSpec(0, scene.means.shape[0])],
                                      hint=CapacityHint())
        (image * dI).sum().backward()
        outs.append((image.detach().cpu().numpy(), [t.grad.detach().cpu().numpy() for t in ts]))
    (img_e, g_e), (img_p, g_p) = outs
    assert np.abs(img_e).max() > 1e-2
    assert rel_l2(img_p, img_e) < 1e-5
    for a, b in zip(g_p, g_e):
        assert rel_l2(a, b) < 1e-5


@pytest.mark.parametrize("ts", [16, 32, 8])
def test_sub_block_masks_keep_every_contributing_pair(dev, ts):
    """The splat kernels skip, per 8x8 sub-block, the listed Gaussians whose plane section cannot reach the
    sub-block's pixels (the section ellipse of largest level over the sub-block's elevation range, moved along
    its centre segment, tested exactly against the pixels' bounding rectangle). Strongly anisotropic Gaussians
    (scales ~0.03-5 mm) packed around two oblique frames, rendered from explicit pixel positions jittered in
    and out of the plane: every pixel must match the oracle's evaluation of its candidate list, so no pair with
    q <= chi2 at a pixel may be skipped (a skipped pair moves alpha by its alpha_hat)."""
    from us3d.scenes import Scene, precision_tril_from_scales_quats

    frames = straight_sweep(2, 96, 80, 0.1, 64.0, 6.0, 25.0, seed=13)
    rng = np.random.default_rng(11)
    G = 2500
    base = make_scene(G, 64.0, seed=12, log_scale_mean=float(np.log(0.4)), log_scale_std=1.0)
    fr0 = frames[0]
    o, al, aa, ae = (np.asarray(v, np.float64) for v in (fr0.origin, fr0.axis_lat, fr0.axis_ax, fr0.axis_ele))
    lat = rng.uniform(-fr0.col_offset * fr0.delta_lat - 1.0, (fr0.width - fr0.col_offset) * fr0.delta_lat + 1.0, G)
    ax = rng.uniform(fr0.near_mm - 1.0, fr0.near_mm + fr0.height * fr0.delta_ax + 1.0, G)
    ele = rng.uniform(-4.0, 4.0, G)
    means = torch.from_numpy((o + lat[:, None] * al + ax[:, None] * aa + ele[:, None] * ae).astype(np.float32))
    scales = base.scales.clamp(0.03, 5.0)
    scene = Scene(means, scales, base.quats, base.opacity, base.intensity,
                  precision_tril_from_scales_quats(scales, base.quats))
    pix = []
    for f in frames:
        n = f.height * f.width
        jit = rng.uniform(-1.0, 1.0, size=(n, 3)) * np.array([0.03, 0.03, 0.15])
        a3 = np.stack([np.asarray(v, np.float64) for v in (f.axis_lat, f.axis_ax, f.axis_ele)])
        pix.append((linear_pixels(f).double() + torch.from_numpy(jit @ a3)).float())
    sem = oug.Semantics()
    cfg = cfg_for(frames[0], tile_size=ts)
    P = [f.height * f.width for f in frames]
    dI = rng.normal(size=sum(P)).astype(np.float32)
    dA = (rng.normal(size=sum(P)) * 0.1).astype(np.float32)
    from us3d.ultragauss import render_frames

    tens = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    image, alpha, _depth = render_frames(*tens, torch.cat(pix).to(dev), frames, cfg,
                                         sweeps=[SweepSpec(0, G)], hint=CapacityHint())
    ((image * torch.from_numpy(dI).to(dev)).sum() + (alpha * torch.from_numpy(dA).to(dev)).sum()).backward()
    params = (scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(), scene.intensity.numpy())
    refs = [np.zeros((G, 3)), np.zeros((G, 6)), np.zeros(G), np.zeros(G)]
    off = 0
    for f, px in zip(frames, pix):
        n = f.height * f.width
        geom = oug.FrameGeom.from_spec(f)
        mr = oug.map_frame(params[0], params[1], geom, chi2=sem.chi2, beta=sem.beta, ts=ts)
        lists = oug.pixel_lists(params[0], params[1], px.numpy(), geom, mr, chi2=sem.chi2, beta=sem.beta, ts=ts)
        assert int(lists[0].sum()) > 20 * n  # a dense case: many candidates per pixel
        args = params + (px.numpy(), geom, lists, sem)
        img32, a32, _d, _s = oug.forward(*args)
        got_i = image[off:off + n].detach().cpu().numpy()
        got_a = alpha[off:off + n].detach().cpu().numpy()
        assert np.all(np.abs(got_a - a32) <= 2e-5 * np.maximum(1.0, np.abs(a32)))
        assert np.abs(got_i - img32).max() <= 2e-5
        for acc, r in zip(refs, oug.backward_f64(*args, dI[off:off + n], dA[off:off + n], None, gate=True)):
            acc += r
        off += n
    for name, t, r in zip(("means", "tril", "opacity", "intensity"), tens, refs):
        err = rel_l2(t.grad.detach().cpu().numpy(), r)
        assert err < GRAD_TOL, f"ts {ts} {name}: rel L2 {err:.3e}"


def test_config5_fan_gradients_vs_fallback_on_64_frames(dev):
    """Config 5 at full size: 2M Gaussians in a 200 x 150 x 150 mm box, 640 x 480 fan frames (60 degrees,
    40-160 mm), fallback semantics (chi2 = 120, the reference fallback renderer's; the only reference semantics
    defined for arbitrary pixel positions). A fixed 64-frame subset of the 1000-frame sweep (every 15th frame)
    is rendered on the GPU in batches of 4 frames (30-60M instances per frame) with the loss gradient nonzero
    on 16 fixed pixels per frame. The oracle evaluates those pixels against every Gaussian with q <= 120; the
    candidates come from an FP64 prefilter with a 1 % margin, and a pair beyond it contributes at most
    o * exp(-60) in the fallback renderer and nothing in the kernels. Image at the sampled pixels and the
    accumulated gradients of all four parameter groups must match."""
    from oracle.fallback import build_precision
    from us3d.scenes import fan_pixels, fan_sweep
    from us3d.ultragauss import render_frames

    box = [200.0, 150.0, 150.0]
    scene = make_scene(2_000_000, 150.0, seed=0, box_mm=box)
    frames = fan_sweep(1000, 480, 640, 60.0, 40.0, 160.0, box, 150.0, seed=2)[::15][:64]
    G, n_pix, n_samp = scene.means.shape[0], 480 * 640, 16
    sem = oug.Semantics.fallback()
    cfg = cfg_for(frames[0], semantics="fallback")
    rng = np.random.default_rng(7)
    samples = [np.sort(rng.choice(n_pix, n_samp, replace=False)) for _ in frames]
    dI = [rng.normal(size=n_samp).astype(np.float32) for _ in frames]
    tens = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    hint = CapacityHint()
    got_img = []
    for b0 in range(0, len(frames), 4):
        fb = frames[b0:b0 + 4]
        image, _alpha, _depth = render_frames(*tens, None, fb, cfg, sweeps=[SweepSpec(0, G)], hint=hint)
        g = torch.zeros(len(fb) * n_pix, dtype=torch.float32)
        for k in range(len(fb)):
            g[k * n_pix + samples[b0 + k]] = torch.from_numpy(dI[b0 + k])
        (image * g.to(dev)).sum().backward()
        img = image.detach().cpu().numpy()
        got_img += [img[k * n_pix + samples[b0 + k]] for k in range(len(fb))]
        del image, _alpha, _depth
    lam = build_precision(scene.tril.to(dev).double(), sem.beta)
    means64 = scene.means.to(dev).double()
    params = (scene.means.numpy(), scene.tril.numpy(), scene.opacity.numpy(), scene.intensity.numpy())
    refs = [np.zeros((G, 3)), np.zeros((G, 6)), np.zeros(G), np.zeros(G)]
    ref_img, gpu_img = [], []
    for k, f in enumerate(frames):
        px = fan_pixels(f)[samples[k]]
        d = px.to(dev).double()[:, None, :] - means64[None]
        q = (torch.einsum("pgi,gij->pgj", d, lam) * d).sum(-1)
        keep = q <= sem.chi2 * 1.01
        counts = keep.sum(1).to(torch.int32).cpu().numpy()
        indices = keep.nonzero()[:, 1].to(torch.int32).cpu().numpy()
        offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32)
        args = params + (px.numpy(), oug.FrameGeom.from_spec(f), (counts, offsets, indices), sem)
        img32, _a, _d, _s = oug.forward(*args)
        ref_img.append(img32)
        gpu_img.append(got_img[k])
        for acc, r in zip(refs, oug.backward_f64(*args, dI[k], None, None, gate=True)):
            acc += r
    img_err = rel_l2(np.concatenate(gpu_img), np.concatenate(ref_img))
    errs = {name: rel_l2(t.grad.detach().cpu().numpy(), r)
            for name, t, r in zip(("means", "tril", "opacity", "intensity"), tens, refs)}
    case = "UltraGauss config 5: 64 fan frames, 2M Gaussians"
    note(case, "image vs FP32 fallback", img_err, IMG_TOL)
    for k, v in errs.items():
        note(case, f"d{k} vs FP64 fallback", v, GRAD_TOL)
    assert img_err < IMG_TOL
    for name, r in zip(errs, refs):
        assert np.abs(r).max() > 0, name
        assert errs[name] < GRAD_TOL, f"config 5 {name}: rel L2 {errs[name]:.3e}"


def test_fallback_semantics_vs_reference_fallback_golden(dev):
    """The reference's own fallback renderer output (tests/golden/fallback_ref_24x32_g384.npz, written by
    _render_ultragauss_torch_fallback, pipeline_god_file.py:556-603, over 384 anisotropic Gaussians): the HIP
    render with fallback semantics, from the explicit pixel array and from the frame's own grid, and the
    gradients w.r.t. the raw parameters through the same sigmoids (pipeline_god_file.py:369-373)."""
    from tests.test_cpu_fallback_golden import load_fallback_golden
    from us3d.batch import FrameSpec
    from us3d.ultragauss import render_frames

    z, H, W = load_fallback_golden()
    _h, _w, dlat, dax, near, col = (float(x) for x in z["frame"])
    ax = torch.from_numpy(z["axes"])
    fr = FrameSpec(origin=torch.from_numpy(z["origin"]), axis_lat=ax[0], axis_ax=ax[1], axis_ele=ax[2], height=H,
                   width=W, delta_lat=dlat, delta_ax=dax, near_mm=near, col_offset=col)
    cfg = cfg_for(fr, semantics="fallback", diag_beta=float(z["beta"]))
    dI = torch.from_numpy(z["dI"]).to(dev)
    dA = torch.from_numpy(z["dA"]).to(dev)
    case = "UltraGauss fallback semantics vs reference fallback output"
    for pixels in ("explicit", "grid"):
        raw = [torch.from_numpy(z[k]).to(dev).requires_grad_(True)
               for k in ("means", "raw_tril", "raw_opacity", "raw_intensity")]
        pix = torch.from_numpy(z["pixels"]).to(dev) if pixels == "explicit" else None
        image, alpha, _depth = render_frames(raw[0], raw[1], torch.sigmoid(raw[2]), torch.sigmoid(raw[3]), pix,
                                             [fr], cfg, sweeps=[SweepSpec(0, raw[0].shape[0])], hint=CapacityHint())
        ((image * dI).sum() + (alpha * dA).sum()).backward()
        assert note(case, f"image ({pixels} pixels)", rel_l2(image.detach().cpu().numpy(), z["image"]),
                    IMG_TOL) < IMG_TOL
        assert note(case, f"alpha ({pixels} pixels)", rel_l2(alpha.detach().cpu().numpy(), z["alpha"]),
                    IMG_TOL) < IMG_TOL
        for r, k in zip(raw, ("d_means", "d_raw_tril", "d_raw_opacity", "d_raw_intensity")):
            err = note(case, f"{k} ({pixels} pixels)", rel_l2(r.grad.cpu().numpy(), z[k]), GRAD_TOL)
            assert err < GRAD_TOL, f"{k}: rel L2 {err:.3e}"
