"""Parity of the GaussianPile us_focus slice splat (libus3d us3d_gp_* through the C ABI) with the oracle.

Bars: the image within 1e-5 relative L2 of the FP32 oracle (same footprints, tile lists and decisions; FP32 sums
in list order) and within 2e-5 * max(1, |ref|) at every pixel; the gradients of all five parameter groups within
1e-4 relative L2 of torch autograd through the FP64 oracle, each frame's gradient sanitised as the reference
sanitises every render call's (non-finite -> 0, magnitudes clamped to 1e3; us_focus.cu:129-137).
"""

import numpy as np
import pytest
import torch

from oracle import gp
from tests.parity import note
from tests.ug_helpers import rel_l2
from us3d.batch import CapacityHint, FrameSpec, SweepSpec
from us3d.gaussianpile import GaussianPlieRenderConfig, GaussianPlieRenderer, frame_from_pose, render_frames

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
GRAD_TOL = 1e-4


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _frames(n, H=64, W=80, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        R = _rot(rng)
        origin = torch.tensor(rng.uniform(-5, 5, 3) + np.array([0.0, 0.0, 0.4 * k]), dtype=torch.float32)
        out.append(FrameSpec(origin=origin, axis_lat=torch.tensor(R[:, 0], dtype=torch.float32),
                             axis_ax=torch.tensor(R[:, 1], dtype=torch.float32),
                             axis_ele=torch.tensor(R[:, 2], dtype=torch.float32), height=H, width=W,
                             delta_lat=0.25, delta_ax=0.2, near_mm=1.0, col_offset=(W - 1) * 0.5))
    return out


def _scene(fr: FrameSpec, G=400, seed=1):
    rng = np.random.default_rng(seed)
    o, al, aa, ae = (fr.origin.double().numpy(), fr.axis_lat.double().numpy(), fr.axis_ax.double().numpy(),
                     fr.axis_ele.double().numpy())
    lat = rng.uniform(-fr.col_offset * fr.delta_lat - 3, (fr.width - fr.col_offset) * fr.delta_lat + 3, G)
    ax = rng.uniform(fr.near_mm - 3, fr.near_mm + fr.height * fr.delta_ax + 3, G)
    ele = rng.uniform(-3, 3, G)
    means = o + lat[:, None] * al + ax[:, None] * aa + ele[:, None] * ae
    scales = np.exp(rng.uniform(np.log(0.3), np.log(2.5), (G, 3)))
    quats = rng.normal(size=(G, 4))
    opac = 1 / (1 + np.exp(-rng.normal(size=G)))
    inten = rng.uniform(0.2, 1.0, G)
    return [torch.tensor(a, dtype=torch.float32) for a in (means, scales, quats, opac, inten)]


def _oframe(fr: FrameSpec):
    return gp.GPFrame(origin=fr.origin, axis_lat=fr.axis_lat, axis_ax=fr.axis_ax, axis_ele=fr.axis_ele, H=fr.height,
                      W=fr.width, delta_lat=fr.delta_lat, delta_ax=fr.delta_ax, near=fr.near_mm,
                      col_offset=fr.col_offset)


def _cfg(fr, ts=16, **kw):
    return GaussianPlieRenderConfig(H=fr.height, W=fr.width, lateral_spacing_mm_per_pix=fr.delta_lat,
                                    axial_spacing_mm_per_pix=fr.delta_ax, axial_near_mm=fr.near_mm,
                                    lateral_col_offset=fr.col_offset, tile_size_px=ts, **kw)


@pytest.mark.parametrize("ts, H, W", [(16, 64, 80), (8, 64, 80), (32, 64, 80), (8, 70, 84), (16, 70, 84),
                                       (32, 70, 84)])
def test_gp_forward_and_gradients_match_oracle(dev, ts, H, W):
    """H = 70, W = 84 leave the last tile row and column partly outside the frame at every tile size (the
    forward's multi-row threads guard those rows)."""
    fr = _frames(1, H=H, W=W)[0]
    params = _scene(fr)
    cfg = _cfg(fr, ts, radius_sigma=3.0, sigma_z_mm=1.2)
    st = gp.GPSettings(k_sigma=3.0, sigma_ele=1.2, det_eps=1e-6, weight_min=1e-6, tile_size=ts)
    ofr = _oframe(fr)
    rng = np.random.default_rng(ts)
    dI = torch.tensor(rng.normal(size=(fr.height, fr.width)), dtype=torch.float32)
    tens = [p.to(dev).clone().requires_grad_(True) for p in params]
    image = render_frames(tens[0], tens[1], tens[2], tens[3], [fr], cfg, intensity=tens[4])
    (image * dI.reshape(-1).to(dev)).sum().backward()
    got = image.detach().cpu().reshape(fr.height, fr.width)
    ref32 = gp.render(*params, ofr, st)
    assert float(ref32.abs().max()) > 0.05
    case = "GaussianPile us_focus, tiles 8/16/32"
    assert note(case, "image vs FP32", rel_l2(got.numpy(), ref32.numpy()), IMG_TOL) < IMG_TOL
    assert np.all(np.abs(got.numpy() - ref32.numpy()) <= 2e-5 * np.maximum(1.0, np.abs(ref32.numpy())))
    p64 = [p.double().clone().requires_grad_(True) for p in params]
    img64 = gp.render(*p64[:4], p64[4], ofr, st, fp64=True)
    note(case, "image vs FP64", rel_l2(got.numpy(), img64.detach().numpy()), IMG_TOL)
    refs = [gp.sanitize_grad(r) for r in torch.autograd.grad((img64 * dI.double()).sum(), p64)]
    for name, t, r in zip(("means", "scales", "quats", "opacity", "intensity"), tens, refs):
        assert float(r.abs().max()) > 0, name
        err = note(case, f"d{name} vs FP64", rel_l2(t.grad.detach().cpu().numpy(), r.numpy()), GRAD_TOL)
        assert err < GRAD_TOL, f"ts {ts} {name}: rel L2 {err:.3e}"


def test_gp_gradients_are_sanitised_per_frame(dev):
    """Small Gaussians (0.06-0.25 mm, so det2^-1/2 makes them bright) over two frames: some per-frame gradients
    exceed the reference's clamp of 1e3, and the renderer's gradient equals the sum over frames of each frame's
    clamped gradient (the reference clamps every render call's gradients, us_focus.cu:129-137, 694-752), not the
    clamp of the sum; within 1e-4 rel. L2 of the FP64 oracle per parameter group."""
    frames = _frames(2, H=48, W=64, seed=41)
    frames[1] = FrameSpec(**{**frames[0].__dict__, "origin": frames[0].origin + 0.05 * frames[0].axis_ele})
    rng = np.random.default_rng(42)
    G = 60
    means, _s, quats, opac, inten = _scene(frames[0], G=G, seed=43)
    scales = torch.tensor(np.exp(rng.uniform(np.log(0.06), np.log(0.25), (G, 3))), dtype=torch.float32)
    params = [means, scales, quats, opac, inten]
    cfg = _cfg(frames[0], 16, radius_sigma=3.0, sigma_z_mm=1.0)
    st = gp.GPSettings(k_sigma=3.0, sigma_ele=1.0, det_eps=1e-6, weight_min=1e-6, tile_size=16)
    n = frames[0].height * frames[0].width
    dI = torch.tensor(rng.normal(size=(2, n)), dtype=torch.float32)
    tens = [p.to(dev).clone().requires_grad_(True) for p in params]
    image = render_frames(tens[0], tens[1], tens[2], tens[3], frames, cfg, intensity=tens[4])
    (image * dI.reshape(-1).to(dev)).sum().backward()
    p64 = [p.double().clone().requires_grad_(True) for p in params]
    H, W = frames[0].height, frames[0].width

    def frame_loss(k):
        return lambda: (gp.render(*p64[:4], p64[4], _oframe(frames[k]), st, fp64=True)
                        * dI[k].double().reshape(H, W)).sum()

    refs = gp.frame_gradients(p64, [frame_loss(0), frame_loss(1)])
    raw = [torch.autograd.grad(frame_loss(k)(), p64) for k in range(2)]
    clamped = sum(int((r.abs() > gp.GRAD_CLAMP).sum()) for rk in raw for r in rk)
    assert clamped > 0, "no per-frame gradient reached the clamp"
    summed = [gp.sanitize_grad(a + b) for a, b in zip(*raw)]
    assert any(float((a - b).abs().max()) > 1.0 for a, b in zip(refs, summed)), "per-frame clamp not exercised"
    for name, t, r in zip(("means", "scales", "quats", "opacity", "intensity"), tens, refs):
        err = rel_l2(t.grad.detach().cpu().numpy(), r.numpy())
        assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"


def test_gp_batch_equals_per_frame_and_is_deterministic(dev):
    frames = _frames(3, seed=4)
    params = _scene(frames[1], G=600, seed=5)
    cfg = _cfg(frames[0])
    tens = [p.to(dev) for p in params]
    batched, batch = render_frames(tens[0], tens[1], tens[2], tens[3], frames, cfg, intensity=tens[4],
                                   return_batch=True)
    again = render_frames(tens[0], tens[1], tens[2], tens[3], frames, cfg, intensity=tens[4])
    assert torch.equal(batched, again)
    # a batch kept from an earlier call (its frame tables already on the device) renders the same image
    reused = render_frames(tens[0], tens[1], tens[2], tens[3], frames, cfg, intensity=tens[4], batch=batch)
    assert torch.equal(batched, reused)
    with pytest.raises(ValueError):
        render_frames(tens[0], tens[1], tens[2], tens[3], frames[:2], cfg, intensity=tens[4], batch=batch)
    n = frames[0].height * frames[0].width
    for k, fr in enumerate(frames):
        one = render_frames(tens[0], tens[1], tens[2], tens[3], [fr], cfg, intensity=tens[4])
        assert torch.equal(one, batched[k * n:(k + 1) * n])
    assert float(batched.abs().max()) > 0.05


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_gp_work_queue_equals_one_workgroup_per_tile(dev, ts):
    """work_queue: persistent workgroups take the (frame, tile) items from per-XCD queues. 8 frames of 512x512 give
    2k-33k tiles, several per resident workgroup, many empty; the image and every gradient equal the
    one-workgroup-per-tile launch's bit for bit."""
    from dataclasses import replace

    frames = _frames(8, H=512, W=512, seed=11)
    params = _scene(frames[0], G=3000, seed=12)
    cfg = _cfg(frames[0], ts=ts)
    dI = torch.from_numpy(np.random.default_rng(13).normal(size=8 * 512 * 512).astype(np.float32)).to(dev)
    outs = []
    for wq in (False, True):
        tens = [p.to(dev).clone().requires_grad_(True) for p in params]
        img = render_frames(tens[0], tens[1], tens[2], tens[3], frames, replace(cfg, work_queue=wq),
                            intensity=tens[4])
        (img * dI).sum().backward()
        outs.append((img.detach(), [t.grad.detach() for t in tens]))
    assert float(outs[0][0].abs().max()) > 0.05
    assert torch.equal(outs[0][0], outs[1][0])
    for name, a, b in zip(("means", "scales", "quats", "opacity", "intensity"), outs[0][1], outs[1][1]):
        assert torch.equal(a, b), f"d{name} differs"


def test_gp_facade_backends(dev):
    fr = _frames(1, seed=7)[0]
    params = _scene(fr, G=300, seed=8)
    pose = torch.eye(4)
    pose[:3, 0], pose[:3, 1], pose[:3, 2], pose[:3, 3] = fr.axis_lat, fr.axis_ax, fr.axis_ele, fr.origin
    cfg = _cfg(fr, sigma_z_mm=1.0)
    means, scales, quats, opac = (p.to(dev) for p in params[:4])
    out = GaussianPlieRenderer(cfg)(means, scales, quats, opac, pose.to(dev))
    assert out.shape == (fr.height, fr.width)
    direct = render_frames(means, scales, quats, opac, [frame_from_pose(pose, cfg)], cfg)
    assert torch.equal(out.reshape(-1), direct)
    # the UltraGauss backend: chi2 = radius_sigma^2, no depth sort, unit intensities (api.py:26-50)
    cfg_u = _cfg(fr, backend="ultragauss_lowmem")
    from us3d.scenes import linear_pixels

    pix = linear_pixels(fr).to(dev)
    img_u = GaussianPlieRenderer(cfg_u)(means, scales, quats, opac, pose.to(dev), pixel_positions_world=pix,
                                        origin_world=fr.origin.to(dev), axis_lateral_world=fr.axis_lat.to(dev),
                                        axis_axial_world=fr.axis_ax.to(dev),
                                        axis_elevational_world=fr.axis_ele.to(dev))
    assert img_u.shape == (fr.height, fr.width) and float(img_u.max()) > 0.05


def test_gp_config2_full_size_sampled_pixels_and_gradients(dev):
    """The bench workload at full size (200k Gaussians, 256 frames of 512x512, one batch): 4 frames x 32
    sampled pixels against the brute-force oracle over every Gaussian (no tile lists, so a pair the tile
    mapper dropped would show), and the gradients of a weighted sum of those pixels against torch autograd
    through the FP64 brute force. Bars: 2e-5 * max(1, |ref|) per pixel, 1e-4 rel. L2 per parameter group."""
    from us3d.scenes import config2

    scene, _target, frames = config2()
    G = scene.means.shape[0]
    fr0 = frames[0]
    H, W = fr0.height, fr0.width
    cfg = GaussianPlieRenderConfig(H=H, W=W, lateral_spacing_mm_per_pix=fr0.delta_lat,
                                   axial_spacing_mm_per_pix=fr0.delta_ax, axial_near_mm=fr0.near_mm,
                                   lateral_col_offset=fr0.col_offset)
    base = [t.float() for t in (scene.means, scene.scales, scene.quats, scene.opacity)]
    params = [t.to(dev).clone().requires_grad_(True) for t in base]
    img = render_frames(*params, frames, cfg)
    assert img.numel() == len(frames) * H * W
    rng = np.random.default_rng(11)
    sel = [0, 85, 170, 255]
    n = 32
    rows = torch.tensor(rng.integers(0, H, size=(len(sel), n)))
    cols = torch.tensor(rng.integers(0, W, size=(len(sel), n)))
    wts = torch.tensor(rng.normal(size=(len(sel), n)), dtype=torch.float32)
    idx = torch.cat([f * H * W + rows[k] * W + cols[k] for k, f in enumerate(sel)])
    (img[idx.to(dev)] * wts.reshape(-1).to(dev)).sum().backward()
    got = img.detach()[idx.to(dev)].cpu()
    st = gp.GPSettings(tile_size=16)
    one32 = torch.ones(G)
    p64 = [t.double().clone().requires_grad_(True) for t in base]
    one64 = torch.ones(G, dtype=torch.float64)
    ref32, losses64 = [], []
    for k, f in enumerate(sel):
        of = _oframe(frames[f])
        ref32.append(gp.render_pixels(*base, one32, of, st, rows[k], cols[k]))
        losses64.append((lambda of, k: lambda: (gp.render_pixels(*p64, one64, of, st, rows[k], cols[k], fp64=True)
                                                 * wts[k].double()).sum())(of, k))
    ref32 = torch.cat(ref32)
    assert float(ref32.abs().max()) > 0.05
    assert bool(torch.all((got - ref32).abs() <= 2e-5 * torch.clamp(ref32.abs(), min=1.0)))
    case = "GaussianPile config 2: 4 frames x 32 pixels, 200k"
    note(case, "pixels vs FP32 brute force", rel_l2(got.numpy(), ref32.numpy()), IMG_TOL)
    refs = gp.frame_gradients(p64, losses64)
    for name, t, ref in zip(("means", "scales", "quats", "opacity"), params, refs):
        assert float(ref.abs().max()) > 0, name
        err = note(case, f"d{name} vs FP64", float((t.grad.detach().cpu().double() - ref).norm() / ref.norm()),
                   GRAD_TOL)
        assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"


def _trainer_setup(dev):
    from us3d.train_step import GaussianPileTrainer

    frames = _frames(2, seed=21)
    means, scales, quats, opac, _ = _scene(frames[0], G=500, seed=22)
    cfg = _cfg(frames[0])
    tm, ts_, tq, to_, _ = _scene(frames[0], G=500, seed=23)
    with torch.no_grad():
        gt = render_frames(tm.to(dev), ts_.to(dev), tq.to(dev), to_.to(dev), frames, cfg).clone()
    raw = [means, torch.log(scales), quats * 1.7, torch.log(opac / (1 - opac))]
    return GaussianPileTrainer, frames, cfg, gt, [t.to(dev) for t in raw]


def test_gp_trainer_gradients_match_torch_autograd_through_the_activations(dev):
    """The GaussianPile training step passes raw values: the pack applies sigmoid(opacity) and
    max(exp(scale), 1e-5), the kernels normalise the quaternions, the reduce applies the chain rules
    (pipeline_god_file.py:360-370). The packed values agree with torch's activations to a few ulps. Taking the
    packed values themselves (an ulp of difference can move a pixel across a box edge), loss and gradients equal
    torch autograd of the same per-frame loss through the renderer at those values, followed by the chain rules
    d raw_s = d s * s (s > 1e-5) and d raw_o = d o * o (1 - o) in torch."""
    from us3d.pipeline import simple_ssim

    Trainer, frames, cfg, gt, raw = _trainer_setup(dev)
    # the separate loss kernels, whose dL/dimage is checked here; the fused path is checked against them below
    tr = Trainer(*raw, frames, gt, cfg, use_graph=False, lr=0.0, fused_loss=False)
    tr.step()
    torch.cuda.synchronize()
    G = raw[0].shape[0]
    off = tr.batch.layout_offsets(tr.k_capacity, tr.v_capacity)
    rec = tr.workspace[off.records: off.records + 48 * G].view(torch.float32).view(G, 12).clone()
    o_pk, s_pk = rec[:, 3], rec[:, 4:7]
    s_t = torch.exp(raw[1]).clamp_min(1e-5)
    o_t = torch.sigmoid(raw[3])
    assert float(((s_pk - s_t).abs() / s_t).max()) < 1e-6
    assert float(((o_pk - o_t).abs() / o_t).max()) < 1e-6
    ps = [raw[0].clone().requires_grad_(True), s_pk.clone().requires_grad_(True),
          raw[2].clone().requires_grad_(True), o_pk.clone().requires_grad_(True)]
    img = render_frames(ps[0], ps[1], ps[2], ps[3], frames, cfg)
    # the fused loss: its value and dL/dimage against torch autograd of the per-frame loss
    x_all = img.detach().clone().requires_grad_(True)
    P = img.numel() // len(frames)
    per = []
    for f in range(len(frames)):
        x, y = x_all[f * P:(f + 1) * P], gt[f * P:(f + 1) * P]
        per.append((x - y).abs().mean() + 0.2 * (1 - simple_ssim(x, y)))
    loss = torch.stack(per).mean()
    loss.backward()
    assert abs(float(tr.losses().mean()) - float(loss.detach())) < 1e-5 * max(1.0, abs(float(loss.detach())))
    d_img = tr.d_image.detach().clone()
    assert float((d_img.double() - x_all.grad.double()).norm() / x_all.grad.double().norm()) < 1e-4
    # the chain rules: the renderer's adjoint at the packed values, driven by the same dL/dimage
    img.backward(d_img)
    refs = [ps[0].grad, torch.where(s_pk > 1e-5, ps[1].grad * s_pk, torch.zeros_like(s_pk)), ps[2].grad,
            ps[3].grad * o_pk * (1 - o_pk)]
    for name, g, ref in zip(("means", "raw_scales", "raw_quats", "raw_opacity"), tr.grads, refs):
        ref = ref.detach().double()
        assert float(ref.abs().max()) > 0, name
        err = float((g.detach().double() - ref).norm() / ref.norm())
        assert err < 1e-5, f"{name}: rel L2 {err:.3e}"


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_gp_loss_fused_into_the_splat_equals_the_separate_loss_kernels(dev, ts):
    """us3d_gp_forward_loss / us3d_gp_backward_loss (loss sums in the forward's epilogue, dL/dimage formed while the
    adjoint stages its pixels) against us3d_gp_forward + us3d_image_loss + us3d_gp_backward: the same losses and
    gradients up to summation order, at every tile size (one or four wavefronts per tile)."""
    from dataclasses import replace

    Trainer, frames, cfg, gt, raw = _trainer_setup(dev)
    cfg = replace(cfg, tile_size_px=ts)
    a = Trainer(*raw, frames, gt, cfg, use_graph=False, lr=0.0, fused_loss=True)
    b = Trainer(*raw, frames, gt, cfg, use_graph=False, lr=0.0, fused_loss=False)
    a.step()
    b.step()
    torch.cuda.synchronize()
    la, lb = a.losses().double().cpu().numpy(), b.losses().double().cpu().numpy()
    assert np.all(lb > 0)
    assert float(np.linalg.norm(la - lb) / np.linalg.norm(lb)) < 1e-6, (la, lb)
    for name, ga, gb in zip(("means", "raw_scales", "raw_quats", "raw_opacity"), a.grads, b.grads):
        ref = gb.detach().double()
        assert float(ref.abs().max()) > 0, name
        err = float((ga.detach().double() - ref).norm() / ref.norm())
        assert err < 1e-6, f"{name}: rel L2 {err:.3e}"


def test_gp_trainer_graph_replays_equal_eager_steps(dev):
    """Three captured steps (replays, with eager reads between them) against three eager steps: the same
    parameters bit for bit after every step, and the loss decreases with lr 1e-2."""
    Trainer, frames, cfg, gt, raw = _trainer_setup(dev)
    eager = Trainer(*raw, frames, gt, cfg, use_graph=False, lr=1e-2)
    graph = Trainer(*raw, frames, gt, cfg, use_graph=True, lr=1e-2)
    losses = []
    for _ in range(3):
        eager.step()
        graph.step()
        torch.cuda.synchronize()
        for a, b in zip(eager.state(), graph.state()):
            assert torch.equal(a, b)
        losses.append(float(graph.losses().mean()))
    assert losses[-1] < losses[0]


def test_gp_two_sweeps_in_one_batch_equal_separate_batches(dev):
    """Two sweeps with their own Gaussians and frames (3 and 2 frames, 400 and 300 Gaussians) in one batch:
    each sweep's image equals its own batch bit for bit, and each Gaussian's gradient equals the one from its
    own sweep's batch (the reduce walks the pairs of the Gaussian's sweep only)."""
    fa = _frames(3, seed=31)
    fb = [FrameSpec(**{**f.__dict__, "sweep": 1}) for f in _frames(2, seed=32)]
    pa = _scene(fa[1], G=400, seed=33)
    pb = _scene(fb[0], G=300, seed=34)
    cfg = _cfg(fa[0])
    both = [torch.cat([a, b]).to(dev).requires_grad_(True) for a, b in zip(pa[:4], pb[:4])]
    img = render_frames(*both, fa + fb, cfg, sweeps=[SweepSpec(0, 400), SweepSpec(400, 300)])
    n = fa[0].height * fa[0].width
    w = torch.linspace(-1.0, 1.0, img.numel(), device=dev)
    (img * w).sum().backward()
    sep_a = [t.to(dev).requires_grad_(True) for t in pa[:4]]
    sep_b = [t.to(dev).requires_grad_(True) for t in pb[:4]]
    ia = render_frames(*sep_a, fa, cfg)
    ib = render_frames(*[t for t in sep_b], [FrameSpec(**{**f.__dict__, "sweep": 0}) for f in fb], cfg)
    assert torch.equal(img[:3 * n], ia) and torch.equal(img[3 * n:], ib)
    assert float(ia.detach().abs().max()) > 0.05 and float(ib.detach().abs().max()) > 0.05
    (ia * w[:3 * n]).sum().backward()
    (ib * w[3 * n:]).sum().backward()
    for t, a, b in zip(both, sep_a, sep_b):
        assert torch.equal(t.grad[:400], a.grad) and torch.equal(t.grad[400:], b.grad)


def test_gp_sweep_table_that_leaves_gaussians_out(dev):
    """Only the second sweep's frames and Gaussian range [400, 700) in the batch: rows 0..399 belong to no sweep
    of it. The image equals that sweep's own batch bit for bit, its Gaussians' gradients equal that batch's, and
    the other rows get exactly zero."""
    fb = [FrameSpec(**{**f.__dict__, "sweep": 0}) for f in _frames(2, seed=32)]
    pa = _scene(_frames(3, seed=31)[1], G=400, seed=33)
    pb = _scene(fb[0], G=300, seed=34)
    cfg = _cfg(fb[0])
    both = [torch.cat([a, b]).to(dev).requires_grad_(True) for a, b in zip(pa[:4], pb[:4])]
    img = render_frames(*both, fb, cfg, sweeps=[SweepSpec(400, 300)])
    w = torch.linspace(-1.0, 1.0, img.numel(), device=dev)
    (img * w).sum().backward()
    sep = [t.to(dev).requires_grad_(True) for t in pb[:4]]
    ib = render_frames(*sep, fb, cfg)
    (ib * w).sum().backward()
    assert float(ib.detach().abs().max()) > 0.05
    assert torch.equal(img, ib)
    for t, b in zip(both, sep):
        assert float(b.grad.abs().max()) > 0
        assert torch.equal(t.grad[400:], b.grad) and not bool(t.grad[:400].any())
