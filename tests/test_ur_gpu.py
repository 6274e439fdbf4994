"""Parity of the HIP UltraG-Ray renderer (libus3d via the C ABI) with the oracle.

Bars: ray counts/offsets, sorted Gaussian indices and depths bit-exact; FP32 B-mode image within 1e-5 relative
L2 of the FP32 oracle (and of the FP64 oracle); gradients within 1e-4 relative L2 of the FP64 oracle per
parameter group.
"""

import numpy as np
import pytest

from oracle import ur as our
from tests.parity import note
from tests.ug_helpers import rel_l2
from tests.ur_helpers import gpu_ray_render, oracle_frame, ray_cfg, ray_settings, small_ray_case

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5
GRAD_TOL = 1e-4


def _oracle(scene, fr, cfg, delta=None):
    geo = oracle_frame(fr, cfg, delta)
    st = ray_settings(cfg)
    params = [t.numpy() for t in scene.params()]
    rm = our.map_frame(params[0], params[1], geo, st)
    return geo, st, params, rm


@pytest.mark.parametrize("variant", ["default", "whole_ray", "local_window"])
def test_ray_mapper_forward_backward_parity(dev, variant):
    kw = {"default": {}, "whole_ray": {"quadrature_window_radius_mm": 0.0},
          "local_window": {"local_transmission_window_only": True, "local_transmission_max_window_mm": 6.0}}[variant]
    scene, frames = small_ray_case(n_frames=2)
    rng = np.random.default_rng(4)
    for fr in frames:
        cfg = ray_cfg(fr, **kw)
        P = fr.width * fr.height
        dB = (rng.normal(size=P) * 1e-2).astype(np.float32)
        out = gpu_ray_render(scene, [fr], cfg, dev, d_image=dB)
        geo, st, params, rm = _oracle(scene, fr, cfg)
        assert rm.indices.size > 200
        lists = out["batch"].ray_lists(out["batch"].map(*[t.to(dev) for t in scene.params()]))[0]
        np.testing.assert_array_equal(lists["counts"].cpu().numpy(), rm.counts)
        np.testing.assert_array_equal(lists["offsets"].cpu().numpy(), rm.offsets)
        np.testing.assert_array_equal(lists["indices"].cpu().numpy(), rm.indices)
        np.testing.assert_array_equal(lists["sorted_depths"].cpu().numpy().astype(np.float32), rm.sorted_depths)
        img32, _aux = our.forward(params, geo, rm, st)
        img64 = our.forward(params, geo, rm, st, fp64=True)
        got = out["image"].detach().cpu().numpy()
        assert np.abs(img64).max() > 1e-3
        case = "UltraG-Ray B-mode, 3 window modes"
        assert note(case, "image vs FP32", rel_l2(got, img32), IMG_TOL) < IMG_TOL
        assert note(case, "image vs FP64", rel_l2(got, img64), IMG_TOL) < IMG_TOL
        ref = our.backward_f64(params, geo, rm, st, dB)
        for name, g, r in zip(("means", "scales", "quats", "atten", "sh"), out["grads"], ref):
            assert np.abs(r).max() > 0, name
            err = note(case, f"d{name} vs FP64", rel_l2(g, r), GRAD_TOL)
            assert err < GRAD_TOL, f"{variant} {name}: rel L2 {err:.3e}"


def test_ray_ops_offsets_and_batch(dev):
    scene, frames = small_ray_case(n_frames=3)
    cfg = ray_cfg(frames[0])
    rng = np.random.default_rng(6)
    deltas = [rng.uniform(-1.5, 1.5, size=f.width).astype(np.float32) for f in frames]
    P = sum(f.width * f.height for f in frames)
    dB = (rng.normal(size=P) * 1e-2).astype(np.float32)
    batched = gpu_ray_render(scene, frames, cfg, dev, deltas=deltas, d_image=dB)
    off = 0
    acc = None
    for fr, de in zip(frames, deltas):
        n = fr.width * fr.height
        single = gpu_ray_render(scene, [fr], cfg, dev, deltas=[de], d_image=dB[off:off + n])
        np.testing.assert_array_equal(single["image"].detach().cpu().numpy(),
                                      batched["image"].detach().cpu().numpy()[off:off + n])
        geo, st, params, rm = _oracle(scene, fr, cfg, de)
        img32, _ = our.forward(params, geo, rm, st)
        assert rel_l2(single["image"].detach().cpu().numpy(), img32) < IMG_TOL
        acc = single["grads"] if acc is None else [a + g for a, g in zip(acc, single["grads"])]
        off += n
    for a, b in zip(acc, batched["grads"]):
        assert rel_l2(b, a) < 1e-6


def test_ray_grad_threshold_and_determinism(dev):
    scene, frames = small_ray_case(n_frames=1)
    fr = frames[0]
    P = fr.width * fr.height
    dB = np.full(P, 5e-7, np.float32)  # below the reference's 1e-6 cut: every sample is skipped
    out = gpu_ray_render(scene, [fr], ray_cfg(fr), dev, d_image=dB)
    for g in out["grads"]:
        assert np.all(g == 0)
    cfg0 = ray_cfg(fr, grad_threshold=0.0)
    a = gpu_ray_render(scene, [fr], cfg0, dev, d_image=dB)
    b = gpu_ray_render(scene, [fr], cfg0, dev, d_image=dB)
    assert any(np.abs(g).max() > 0 for g in a["grads"])
    for x, y in zip(a["grads"], b["grads"]):
        np.testing.assert_array_equal(x, y)


def test_ray_empty_scene(dev):
    scene, frames = small_ray_case(n=32, n_frames=1)
    scene.means[:] = -1.0e4
    fr = frames[0]
    out = gpu_ray_render(scene, [fr], ray_cfg(fr), dev, d_image=np.ones(fr.width * fr.height, np.float32))
    assert float(out["image"].detach().abs().max()) == 0.0
    for g in out["grads"]:
        assert np.all(g == 0)


def test_ray_samples_in_decreasing_depth_take_the_general_path(dev):
    """Each scanline's samples given far-to-near (depth not monotone along the sample index): the adjoint cannot
    use its finished-window suffix sums and walks every sample; images and gradients must still match."""
    from us3d.scenes import ray_geometry

    scene, frames = small_ray_case(n_frames=1)
    fr = frames[0]
    cfg = ray_cfg(fr)
    pix = ray_geometry(fr)[0].flip(1).contiguous()  # (W, H, 3), samples reversed along each scanline
    P = fr.width * fr.height
    dB = (np.random.default_rng(12).normal(size=P) * 1e-2).astype(np.float32)
    out = gpu_ray_render(scene, [fr], cfg, dev, d_image=dB, pixels=[pix])
    geo = oracle_frame(fr, cfg, pixels=pix)
    st = ray_settings(cfg)
    params = [t.numpy() for t in scene.params()]
    rm = our.map_frame(params[0], params[1], geo, st)
    img32, _aux = our.forward(params, geo, rm, st)
    assert rel_l2(out["image"].detach().cpu().numpy(), img32) < IMG_TOL
    ref = our.backward_f64(params, geo, rm, st, dB)
    for name, g, r in zip(("means", "scales", "quats", "atten", "sh"), out["grads"], ref):
        err = rel_l2(g, r)
        assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"


@pytest.mark.parametrize("H", [1000, 1300])
@pytest.mark.parametrize("window", [2.0, 0.0])
def test_ray_long_rays_with_several_hit_batches(dev, H, window):
    """About 580 hits per scanline (three batches of 256 in the adjoint) over 1000 samples (one staged chunk
    shared by every batch) or 1300 samples (two chunks, restaged per batch); forward and adjoint must match the
    oracle either way."""
    scene, frames = small_ray_case(n=3000, n_frames=1, W=4, H=H, d_ax=0.04)
    fr = frames[0]
    cfg = ray_cfg(fr, quadrature_window_radius_mm=window)
    P = fr.width * fr.height
    dB = (np.random.default_rng(5).normal(size=P) * 1e-2).astype(np.float32)
    out = gpu_ray_render(scene, [fr], cfg, dev, d_image=dB)
    geo, st, params, rm = _oracle(scene, fr, cfg)
    assert rm.counts.min() > 2 * 256
    img32, _aux = our.forward(params, geo, rm, st)
    case = "UltraG-Ray, ~580 hits per scanline, 1000/1300 samples"
    assert note(case, "image vs FP32", rel_l2(out["image"].detach().cpu().numpy(), img32), IMG_TOL) < IMG_TOL
    ref = our.backward_f64(params, geo, rm, st, dB)
    for name, g, r in zip(("means", "scales", "quats", "atten", "sh"), out["grads"], ref):
        assert np.abs(r).max() > 0, name
        err = note(case, f"d{name} vs FP64", rel_l2(g, r), GRAD_TOL)
        assert err < GRAD_TOL, f"H={H} Rw={window} {name}: rel L2 {err:.3e}"


def test_ray_log_scales_are_activated_in_the_pack(dev):
    """scales_are_log: the pack applies max(exp(x), 1e-5) to log-scales (pipeline_god_file.py:360-361) and the
    reduce returns d/dx = d/ds * exp(x) where exp(x) >= 1e-5, else 0. Same image as passing the activated scales,
    gradients equal to the chain applied on the host; one clamped axis gets a zero gradient."""
    import torch

    from us3d.batch import CapacityHint, SweepSpec
    from us3d.scenes import ray_geometry
    from us3d.ultragray import URBatch
    from tests.ur_helpers import ray_spec

    scene, frames = small_ray_case(n_frames=1)
    fr = frames[0]
    cfg = ray_cfg(fr)
    means, scales, quats, atten, sh = [t.to(dev).float().contiguous() for t in scene.params()]
    log_s = torch.log(scales)
    log_s[0, 1] = -20.0  # exp = 2e-9: clamped to 1e-5
    act = torch.exp(log_s).clamp_min(1e-5)
    pix, ro, rd = (t.to(dev).contiguous() for t in ray_geometry(fr))
    pix = pix.reshape(-1, 3).contiguous()
    off = torch.zeros(fr.width, device=dev)
    P = fr.width * fr.height
    dB = torch.from_numpy((np.random.default_rng(8).normal(size=P) * 1e-2).astype(np.float32)).to(dev)
    G = means.shape[0]
    out = {}
    for log_mode in (0, 1):
        st = cfg.settings()
        st.scales_are_log = log_mode
        batch = URBatch([ray_spec(fr, cfg)], [SweepSpec(0, G)], G, st, dev, CapacityHint())
        m = batch.map(means, log_s if log_mode else act, quats, atten, sh)
        image, aux = batch.forward(m, pix, ro, rd, off)
        grads = batch.backward(m, pix, ro, rd, off, aux, dB)
        out[log_mode] = (image.clone(), [g.clone() for g in grads])
    img_lin, g_lin = out[0]
    img_log, g_log = out[1]
    assert float(img_lin.abs().max()) > 1e-3
    assert rel_l2(img_log.cpu().numpy(), img_lin.cpu().numpy()) < 1e-6
    want = g_lin[1] * act * (torch.exp(log_s) >= 1e-5).float()
    assert float(g_log[1][0, 1]) == 0.0
    assert rel_l2(g_log[1].cpu().numpy(), want.cpu().numpy()) < 1e-5
    for a, b in zip(g_log[:1] + g_log[2:], g_lin[:1] + g_lin[2:]):
        assert rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-5


def test_ray_sweep_table_that_leaves_gaussians_out(dev):
    """The scene's Gaussians sit at rows [100, 100 + G) of arrays padded with 100 rows before and 60 after, and
    the sweep table names that range only: the image equals the unpadded batch's bit for bit, the scene rows'
    gradients equal its gradients, and the padding rows (in no sweep of the batch) get exactly zero."""
    import torch

    from us3d.batch import CapacityHint, SweepSpec
    from us3d.ultragray import URBatch
    from tests.ur_helpers import ray_spec

    scene, frames = small_ray_case(n_frames=2)
    cfg = ray_cfg(frames[0])
    specs = [ray_spec(f, cfg) for f in frames]
    params = [t.to(dev).float().contiguous() for t in scene.params()]
    G = params[0].shape[0]
    pads = [(torch.rand((100,) + t.shape[1:], device=dev), torch.rand((60,) + t.shape[1:], device=dev))
            for t in params]
    padded = [torch.cat([a, t, b]).contiguous() for t, (a, b) in zip(params, pads)]
    off = torch.zeros(sum(s.width for s in specs), device=dev)
    plain = URBatch(specs, [SweepSpec(0, G)], G, cfg.settings(), dev, CapacityHint())
    gapped = URBatch(specs, [SweepSpec(100, G)], G + 160, cfg.settings(), dev, CapacityHint())
    m0, m1 = plain.map(*params), gapped.map(*padded)
    img0, aux0 = plain.forward(m0, None, None, None, off)
    img1, aux1 = gapped.forward(m1, None, None, None, off)
    assert float(img0.abs().max()) > 1e-3
    assert torch.equal(img0, img1)
    dB = torch.from_numpy((np.random.default_rng(6).normal(size=img0.numel()) * 1e-2).astype(np.float32)).to(dev)
    g0 = plain.backward(m0, None, None, None, off, aux0, dB)
    g1 = gapped.backward(m1, None, None, None, off, aux1, dB)
    for a, b in zip(g0, g1):
        assert float(a.abs().max()) > 0
        assert torch.equal(a, b[100:100 + G])
        assert not bool(b[:100].any()) and not bool(b[100 + G:].any())


def test_ray_linear_probe_geometry_in_kernel_equals_explicit_arrays(dev):
    """pixels / ray_origins / ray_dirs = NULL: the kernels evaluate the linear probe's samples and rays from the
    frame table with the FP32 operations of the pipeline's grid (pipeline_god_file.py:292-333), two frames, nonzero
    elevational offsets. On the grid the forward evaluates the window integral and the echo as one-dimensional
    quadratics along the scanline and the adjoint forms the echo sums from the same quadratic, so image, aux and
    gradients equal the explicit arrays' to FP32 rounding (image and aux 1e-6, gradients 1e-5), and the grid's image
    meets the oracle's FP32 and FP64 passes within 1e-5 and its gradients the FP64 oracle's within 1e-4."""
    import torch

    from us3d.batch import CapacityHint, SweepSpec
    from us3d.scenes import ray_geometry
    from us3d.ultragray import URBatch
    from tests.ur_helpers import ray_spec

    scene, frames = small_ray_case(n_frames=2)
    cfg = ray_cfg(frames[0])
    specs = [ray_spec(f, cfg) for f in frames]
    geos = [ray_geometry(s, dev) for s in specs]
    pix = torch.cat([g[0].reshape(-1, 3) for g in geos]).contiguous()
    ro = torch.cat([g[1] for g in geos]).contiguous()
    rd = torch.cat([g[2] for g in geos]).contiguous()
    W = sum(s.width for s in specs)
    off = torch.linspace(-0.4, 0.4, W, device=dev)
    params = [t.to(dev).float().contiguous() for t in scene.params()]
    G = params[0].shape[0]
    batch = URBatch(specs, [SweepSpec(0, G)], G, cfg.settings(), dev, CapacityHint())
    m = batch.map(*params)
    img_e, aux_e = batch.forward(m, pix, ro, rd, off)
    img_p, aux_p = batch.forward(m, None, None, None, off)
    assert float(img_e.abs().max()) > 1e-3
    case = "UltraG-Ray on the probe's grid (echo line form)"
    assert note(case, "image vs explicit arrays", rel_l2(img_p.cpu().numpy(), img_e.cpu().numpy()), 1e-6) < 1e-6
    for c, name in enumerate(("T", "E", "den")):
        a, b = aux_p[:, c].cpu().numpy(), aux_e[:, c].cpu().numpy()
        assert note(case, f"aux {name} vs explicit arrays", rel_l2(a, b), 1e-6) < 1e-6, name
    dB = torch.from_numpy((np.random.default_rng(3).normal(size=img_e.numel()) * 1e-2).astype(np.float32)).to(dev)
    g_e = batch.backward(m, pix, ro, rd, off, aux_e, dB)
    g_p = batch.backward(m, None, None, None, off, aux_p, dB)
    for name, a, b in zip(("means", "scales", "quats", "atten", "sh"), g_e, g_p):
        assert float(a.abs().max()) > 0
        assert note(case, f"d{name} vs explicit arrays", rel_l2(b.cpu().numpy(), a.cpu().numpy()), 1e-5) < 1e-5, name
    # and against the FP64 oracle, frame by frame (the oracle sums the frames' gradients)
    off_np = off.cpu().numpy()
    refs = None
    p0 = 0
    r0 = 0
    for fr, sp in zip(frames, specs):
        n = sp.width * sp.height
        geo, st, params_np, rm = _oracle(scene, fr, cfg, off_np[r0:r0 + sp.width])
        img32, _aux = our.forward(params_np, geo, rm, st)
        img64 = our.forward(params_np, geo, rm, st, fp64=True)
        got = img_p[p0:p0 + n].cpu().numpy()
        assert note(case, "image vs FP32", rel_l2(got, img32), IMG_TOL) < IMG_TOL
        assert note(case, "image vs FP64", rel_l2(got, img64), IMG_TOL) < IMG_TOL
        r = our.backward_f64(params_np, geo, rm, st, dB[p0:p0 + n].cpu().numpy())
        refs = r if refs is None else [x + y for x, y in zip(refs, r)]
        p0 += n
        r0 += sp.width
    for name, g, r in zip(("means", "scales", "quats", "atten", "sh"), g_p, refs):
        err = note(case, f"d{name} vs FP64", rel_l2(g.cpu().numpy(), r), GRAD_TOL)
        assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"


def test_ray_rotated_anisotropic_gaussians(dev):
    """Strongly anisotropic Gaussians (scales about (3, 1, 0.5) mm) under composed rotations q_a (x) q_b about
    random axes (never symmetric): mapper bit-exact, image and gradients against the oracle, whose rotation and
    covariance conventions the CPU tests pin to Rodrigues' formula and the whitening of shared.cuh:37-48."""
    import torch

    from tests.ur_helpers import axis_quat, hamilton

    G = 400
    scene, frames = small_ray_case(n=G, n_frames=1)
    rng = np.random.default_rng(21)
    q = [hamilton(axis_quat(rng.normal(size=3), rng.uniform(15, 165)), axis_quat(rng.normal(size=3),
                                                                                  rng.uniform(15, 165)))
         for _ in range(G)]
    scene.quats = torch.tensor(np.array(q), dtype=torch.float32)
    scene.scales = torch.tensor(np.array([3.0, 1.0, 0.5]) * np.exp(rng.normal(0.0, 0.1, (G, 3))), dtype=torch.float32)
    fr = frames[0]
    cfg = ray_cfg(fr)
    P = fr.width * fr.height
    dB = (rng.normal(size=P) * 1e-2).astype(np.float32)
    out = gpu_ray_render(scene, [fr], cfg, dev, d_image=dB)
    geo, st, params, rm = _oracle(scene, fr, cfg)
    assert rm.indices.size > 200
    lists = out["batch"].ray_lists(out["batch"].map(*[t.to(dev) for t in scene.params()]))[0]
    np.testing.assert_array_equal(lists["indices"].cpu().numpy(), rm.indices)
    img32, _aux = our.forward(params, geo, rm, st)
    got = out["image"].detach().cpu().numpy()
    assert np.abs(img32).max() > 1e-3
    case = "UltraG-Ray, rotated anisotropic Gaussians"
    assert note(case, "image vs FP32", rel_l2(got, img32), IMG_TOL) < IMG_TOL
    ref = our.backward_f64(params, geo, rm, st, dB)
    for name, g, r in zip(("means", "scales", "quats", "atten", "sh"), out["grads"], ref):
        assert np.abs(r).max() > 0, name
        err = note(case, f"d{name} vs FP64", rel_l2(g, r), GRAD_TOL)
        assert err < GRAD_TOL, f"{name}: rel L2 {err:.3e}"


def test_ray_echo_weight_of_a_rotated_gaussian_is_the_closed_form(dev):
    """On the GPU, one Gaussian (s = 3, 1, 0.5 mm, the composed rotation, tau = 1): the echo weight sum the kernel
    stores per sample (aux den) equals exp(-1/2 |S^-1 R^T (p - mu)|^2) with R from Rodrigues' formula, at points
    along each principal axis and at random points."""
    import torch

    from tests.ur_helpers import composed_rotation
    from us3d.batch import CapacityHint, FrameSpec, SweepSpec
    from us3d.ultragray import URBatch, UltragrayRenderConfig

    q, R = composed_rotation()
    s = np.array([3.0, 1.0, 0.5])
    mu = np.array([0.0, 30.0, 0.0])
    rng = np.random.default_rng(3)
    pts = [mu + t * R[:, k] for k in range(3) for t in (-1.5 * s[k], -0.5 * s[k], 0.7 * s[k], 1.2 * s[k])]
    pts = np.array(pts + [mu + rng.uniform(-2.0, 2.0, 3) for _ in range(12)])
    H = pts.shape[0]
    v = lambda *a: torch.tensor(a, dtype=torch.float32)  # noqa: E731
    fr = FrameSpec(v(0, 0, 0), v(1, 0, 0), v(0, 1, 0), v(0, 0, 1), H, 1, 1.0, 0.1, 0.0, 0.0)
    cfg = UltragrayRenderConfig(image_width=1, image_height=H, width_mm=1.0, near_mm=0.0, ops_max_offset_mm=0.0)
    f32 = lambda a: torch.tensor(np.asarray(a), dtype=torch.float32, device=dev).contiguous()  # noqa: E731
    params = [f32(mu[None]), f32(s[None]), f32(q[None]), f32([60.0]), f32([[0.5, 0, 0, 0]])]
    batch = URBatch([fr], [SweepSpec(0, 1)], 1, cfg.settings(), dev, CapacityHint())
    m = batch.map(*params)
    _img, aux = batch.forward(m, f32(pts), f32([[0.0, 0.0, 0.0]]), f32([[0.0, 1.0, 0.0]]), f32([0.0]))
    aux = aux.cpu().numpy()
    z = (pts - mu) @ R / s
    np.testing.assert_allclose(aux[:, 0], 1.0)
    np.testing.assert_allclose(aux[:, 2], np.exp(-0.5 * (z * z).sum(1)), rtol=2e-5, atol=1e-12)
