"""GPU parity at the sizes of the two batched workloads that the smaller tests do not reach.

UltraG-Ray at its benchmark density (200k Gaussians in a 120 mm cube, 128 scanlines x 1024 samples at
0.8 x 0.1 mm, about 2,400 hits per scanline, so every ray list takes the long-segment sort and the adjoint
restages its chunk per 256-hit batch): one frame, with the mapper's ray lists bit-exact against the oracle
(ops/ultragray_renderer/src/mapper/core.cu:3-129), the whole FP32 image within 1e-5 rel. L2 of the oracle's FP32
forward (src/shader/forward.cu:22-119), the image on four scanlines within 1e-5 of its FP64 forward, and the
gradients of a loss on those four scanlines within 1e-4 of its FP64 adjoint (src/shader/backward.cu:26-325).

The multi-sweep slice-splat step at full size (8 sweeps, 50k-500k Gaussians, 32-400 frames of 256x256, 512x512 and
640x480): two captured replays equal two eager steps bit for bit, and one frame of every sweep rendered by the
first step equals the oracle's FP32 forward (ops/ultragauss_renderer/src/shader/forward.cu:17-76) within 1e-5
rel. L2, with that sweep's Gaussians only.
"""

import numpy as np
import pytest
import torch

from oracle import ug as oug
from oracle import ur as our
from tests.parity import note
from tests.ug_helpers import rel_l2
from tests.ur_helpers import oracle_frame, ray_settings, ray_spec

pytestmark = pytest.mark.gpu


def test_config3_frame_ray_lists_image_and_sampled_gradients(dev):
    from us3d.batch import CapacityHint, SweepSpec
    from us3d.scenes import config3
    from us3d.ultragray import URBatch, UltragrayRenderConfig

    scene, frames = config3()
    fr = frames[len(frames) // 2]
    G = scene.means.shape[0]
    H, W = fr.height, fr.width
    cfg = UltragrayRenderConfig(image_width=W, image_height=H, width_mm=W * fr.delta_lat, near_mm=fr.near_mm,
                                lateral_col_offset=fr.col_offset, ops_max_offset_mm=0.0)
    print(f"[config3] G={G} frame {len(frames) // 2} {W}x{H}", flush=True)
    params = [t.to(dev).float().contiguous() for t in scene.params()]
    batch = URBatch([ray_spec(fr, cfg)], [SweepSpec(0, G)], G, cfg.settings(), dev, CapacityHint())
    m = batch.map(*params)
    off = torch.zeros(W, device=dev)
    image, aux = batch.forward(m, None, None, None, off)
    rays = [9, 47, 80, 118]
    rng = np.random.default_rng(31)
    dB = np.zeros((H, W), np.float32)
    dB[:, rays] = (rng.normal(size=(H, len(rays))) * 1e-2).astype(np.float32)
    grads = [g.cpu().numpy().astype(np.float64) for g in
             batch.backward(m, None, None, None, off, aux, torch.from_numpy(dB.reshape(-1)).to(dev))]
    lists = batch.ray_lists(m)[0]
    torch.cuda.synchronize()

    geo = oracle_frame(fr, cfg)
    st = ray_settings(cfg)
    p_np = [t.numpy() for t in scene.params()]
    rm = our.map_frame(p_np[0], p_np[1], geo, st)
    print(f"[config3] oracle hits {rm.indices.size}, max per ray {rm.counts.max()}", flush=True)
    assert rm.counts.min() > 1024 and rm.indices.size > 200_000
    np.testing.assert_array_equal(lists["counts"].cpu().numpy(), rm.counts)
    np.testing.assert_array_equal(lists["offsets"].cpu().numpy(), rm.offsets)
    np.testing.assert_array_equal(lists["indices"].cpu().numpy(), rm.indices)
    np.testing.assert_array_equal(lists["sorted_depths"].cpu().numpy().astype(np.float32), rm.sorted_depths)
    got = image.cpu().numpy().reshape(H, W)
    img32, _aux = our.forward(p_np, geo, rm, st)
    img32 = img32.reshape(H, W)
    assert np.abs(img32).max() > 1e-3
    err32 = rel_l2(got, img32)
    # FP64 forward and adjoint on the sampled scanlines only (the other rays' lists emptied)
    keep = np.zeros(W, bool)
    keep[rays] = True
    rm.counts = np.where(keep, rm.counts, 0).astype(np.int32)
    img64 = our.forward(p_np, geo, rm, st, fp64=True).reshape(H, W)
    err64 = rel_l2(got[:, rays], img64[:, rays])
    ref = our.backward_f64(p_np, geo, rm, st, dB.reshape(-1))
    errs = {name: rel_l2(g, r) for name, g, r in zip(("means", "scales", "quats", "atten", "sh"), grads, ref)}
    case = "UltraG-Ray config-3 frame (200k Gaussians, 128x1024)"
    note(case, "image vs FP32", err32, 1e-5)
    note(case, "image vs FP64 (4 scanlines)", err64, 1e-5)
    for k, v in errs.items():
        note(case, f"d{k} vs FP64 (4 scanlines)", v, 1e-4)
    assert err32 < 1e-5 and err64 < 1e-5
    for name, r in zip(errs, ref):
        assert np.abs(r).max() > 0, name
        assert errs[name] < 1e-4, f"{name}: rel L2 {errs[name]:.3e}"


def test_config4_captured_step_equals_eager_and_sampled_frames_match_the_oracle(dev):
    from us3d.batch import CapacityHint
    from us3d.scenes import config4, linear_pixels
    from us3d.train_step import SliceSplatTrainer
    from us3d.ultragauss import UltraGaussRenderConfig, render_frames

    scene, target, frames, sweeps = config4()
    F = len(frames)
    G = scene.means.shape[0]
    print(f"[config4] {len(sweeps)} sweeps, {G} Gaussians, {F} frames", flush=True)
    f0 = frames[0]
    cfg = UltraGaussRenderConfig(H=f0.height, W=f0.width, lateral_spacing_mm_per_pix=0.2,
                                 axial_spacing_mm_per_pix=0.2, axial_near_mm=0.0, fixed_bg_alpha=1e-6, tile_size=16)
    with torch.no_grad():
        gt = render_frames(*[t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)], None,
                           frames, cfg, sweeps=sweeps, hint=CapacityHint())[0].clone()
    del target
    torch.cuda.empty_cache()
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]
    # eager steps first (states kept on the host), then the captured steps: one workspace on the device at a time
    eager = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=False, lr=1e-3, check_every=10 ** 9)
    states, losses = [], []
    for step in range(3):
        eager.step()
        torch.cuda.synchronize()
        if step == 0:
            # the first step's images render the initial parameters; its packed records hold the activated
            # opacity and intensity the kernels used
            image0 = eager.image.detach().cpu()
            off = eager.batch.layout_offsets(eager.k_capacity, eager.v_capacity)
            rec = eager.workspace[off.records: off.records + 48 * G].view(torch.float32).view(G, 12).cpu()
            o_pk, c_pk = rec[:, 3].numpy(), rec[:, 4].numpy()
        states.append([t.cpu() for t in eager.state()])
        losses.append(eager.losses().cpu())
    assert not eager.overflowed()
    base = eager.batch.layout.frame_pixel_base
    ws_gib = eager.workspace.numel() / 2 ** 30
    del eager
    torch.cuda.empty_cache()
    graph = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=True, lr=1e-3, check_every=10 ** 9)
    for step in range(3):
        graph.step()
        torch.cuda.synchronize()
        for x, y in zip(states[step], graph.state()):
            assert torch.equal(x, y.cpu())
        assert torch.equal(losses[step], graph.losses().cpu())
    assert not graph.overflowed()
    print(f"[config4] 3 graph replays equal 3 eager steps bit for bit; workspace {ws_gib:.1f} GiB", flush=True)
    del graph
    torch.cuda.empty_cache()

    sem = oug.Semantics()
    means, tril = scene.means.numpy(), scene.tril.numpy()
    for s, sw in enumerate(sweeps):
        ks = [k for k, f in enumerate(frames) if f.sweep == s]
        k = ks[len(ks) // 2]
        fr = frames[k]
        g0, g1 = sw.g_begin, sw.g_begin + sw.g_count
        geom = oug.FrameGeom.from_spec(fr)
        pix = linear_pixels(fr).numpy()
        mr = oug.map_frame(means[g0:g1], tril[g0:g1], geom)
        lists = oug.pixel_lists(means[g0:g1], tril[g0:g1], pix, geom, mr)
        plane = oug.plane_form(means[g0:g1], tril[g0:g1], geom)  # the trainer renders the frames' own grids
        img, _a, _d, _s = oug.forward(means[g0:g1], tril[g0:g1], o_pk[g0:g1], c_pk[g0:g1], pix, geom, lists, sem,
                                      plane=plane)
        got = image0[base[k]:base[k + 1]].cpu().numpy()
        err = note("UltraGauss config 4: one frame per sweep, 8 sweeps", "image vs FP32", rel_l2(got, img), 1e-5)
        assert np.abs(img).max() > 0.05
        assert err < 1e-5
