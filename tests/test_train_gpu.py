"""The graph-captured training step: every replay equals the eager launch sequence bit for bit; the fused
loss and its gradient match torch autograd of the reference loss; the Adam kernel matches torch.optim.Adam;
an overflowing capacity is detected, skipped by Adam and retried."""

from dataclasses import replace

import numpy as np
import pytest
import torch

from tests.ug_helpers import cfg_for, rel_l2, small_case
from us3d import _lib
from us3d.train_step import SliceSplatTrainer
from us3d.ultragauss import render_frames

pytestmark = pytest.mark.gpu


def _setup(dev, n_frames=3):
    scene, frames = small_case(n_frames=n_frames)
    target, _ = small_case(n_frames=n_frames, seed=11)
    cfg = cfg_for(frames[0])
    with torch.no_grad():
        gt = render_frames(*[t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)], None,
                           frames, cfg)[0]
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]
    return frames, gt, cfg, init


def test_graph_replays_equal_eager_steps(dev):
    frames, gt, cfg, init = _setup(dev)
    a = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=True)
    b = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=False)
    for _ in range(4):
        a.step()
        torch.cuda.synchronize()
        b.step()
        torch.cuda.synchronize()
        for x, y in zip(a.state(), b.state()):
            assert torch.equal(x, y)
        assert torch.equal(a.losses(), b.losses())
    assert int(a.step_count.item()) == 4


def test_fused_loss_and_gradient_match_torch(dev):
    from us3d.pipeline import simple_ssim

    frames, gt, cfg, init = _setup(dev)
    P = gt.numel()
    F = len(frames)
    pred = torch.rand(P, device=dev)
    loss = torch.empty(F, device=dev)
    d = torch.empty(P, device=dev)
    tmp = torch.empty(int(_lib.lib().us3d_image_loss_tmp_bytes(F)) // 8, dtype=torch.float64, device=dev)
    base = torch.tensor([i * (P // F) for i in range(F + 1)], dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().us3d_image_loss(pred.data_ptr(), gt.data_ptr(), base.data_ptr(), F, 1.0, 0.2, 1.0 / F,
                                          loss.data_ptr(), d.data_ptr(), tmp.data_ptr(),
                                          _lib.stream_ptr(dev)), "us3d_image_loss")
    x = pred.detach().double().cpu().requires_grad_(True)
    y = gt.detach().double().cpu()
    per = []
    for f in range(F):
        xs, ys = x[f * (P // F):(f + 1) * (P // F)], y[f * (P // F):(f + 1) * (P // F)]
        per.append((xs - ys).abs().mean() + 0.2 * (1 - simple_ssim(xs, ys)))
    torch.stack(per).mean().backward()
    assert rel_l2(loss.cpu().numpy(), torch.stack(per).detach().numpy()) < 1e-6
    assert rel_l2(d.cpu().numpy(), x.grad.numpy()) < 1e-5


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("chunk", [None, 2])
def test_loss_fused_into_the_splat_equals_the_separate_loss_kernels(dev, tile, chunk):
    """us3d_ug_forward_loss / us3d_ug_backward_loss (loss sums in the forward's epilogue, the loss gradient formed in
    the adjoint's pixel staging) against us3d_ug_forward + us3d_image_loss + us3d_ug_backward: the same losses and
    gradients up to the FP64 summation order, at every tile size (one or four wavefronts per tile) and with the
    frames split into chunks (target and loss offsets per chunk)."""
    frames, gt, cfg, init = _setup(dev, n_frames=5)
    cfg = replace(cfg, tile_size=tile)
    a = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=False, lr=0.0, chunk_frames=chunk, fused_loss=True)
    b = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=False, lr=0.0, chunk_frames=chunk, fused_loss=False)
    a.step()
    b.step()
    torch.cuda.synchronize()
    la, lb = a.losses().double().cpu().numpy(), b.losses().double().cpu().numpy()
    assert np.all(lb > 0)
    assert rel_l2(la, lb) < 1e-6, (la, lb)
    for name, ga, gb in zip(("means", "tril", "raw_opacity", "raw_intensity"), a.grads, b.grads):
        ref = gb.detach().cpu().numpy()
        assert np.abs(ref).max() > 0, name
        err = rel_l2(ga.detach().cpu().numpy(), ref)
        assert err < 1e-6, f"{name}: rel L2 {err:.3e}"


def test_adam_kernel_matches_torch_adam(dev):
    g = torch.Generator(device="cpu").manual_seed(0)
    p0 = torch.randn(10_000, generator=g).to(dev)
    p_ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([p_ref], lr=1e-3)
    p = p0.clone()
    m, v = torch.zeros_like(p), torch.zeros_like(p)
    step = torch.zeros(1, dtype=torch.int32, device=dev)
    for k in range(5):
        grad = torch.randn(10_000, generator=g).to(dev)
        p_ref.grad = grad.clone()
        opt.step()
        L = _lib.lib()
        s = _lib.stream_ptr(dev)
        _lib.check(L.us3d_adam_step(p.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr(), p.numel(),
                                    step.data_ptr(), None, 1e-3, 0.9, 0.999, 1e-8, s), "adam")
        _lib.check(L.us3d_adam_advance(step.data_ptr(), None, s), "advance")
        torch.testing.assert_close(p, p_ref.detach(), rtol=1e-6, atol=1e-7)
    assert int(step.item()) == 5


def test_capacity_overflow_is_skipped_and_retried(dev):
    frames, gt, cfg, init = _setup(dev)
    ref = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=True)
    t = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=True)
    t._allocate((64, 16))  # far below the instance and visible-pair counts
    # the visible pairs overflow first (no instances are counted then), then the instances: each time Adam skips the
    # step, the capacity grows, the graph is re-captured and the step repeated
    t.step()
    ref.step()
    assert t.k_capacity > 64 and t.v_capacity > 16
    for x, y in zip(t.state(), ref.state()):
        assert torch.equal(x, y)
    assert int(t.step_count.item()) == 1


def test_multi_sweep_graph_step_equals_eager(dev):
    """A captured step over several sweeps of uneven size (frame sizes differ too) replays bit-identically to
    the eager launch sequence."""
    from us3d.scenes import multi_sweep

    specs = [(500, 2, 64, 48), (900, 3, 48, 48)]
    scene, target, frames, sweeps = multi_sweep(specs, cube_mm=64.0, seed=3, length_mm=30.0, spacing=0.3)
    cfg = cfg_for(frames[0])
    with torch.no_grad():
        gt = render_frames(*[t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)], None,
                           frames, cfg, sweeps=sweeps)[0]
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]
    a = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=True)
    b = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=False)
    for _ in range(3):
        a.step()
        torch.cuda.synchronize()
        b.step()
        torch.cuda.synchronize()
        for x, y in zip(a.state(), b.state()):
            assert torch.equal(x, y)
    assert a.losses().shape[0] == len(frames)


def test_chunked_step_equals_the_whole_batch_step(dev):
    """chunk_frames renders a step's frames in chunks (here 2 + 2 + 1 over two sweeps, a chunk crossing the sweep
    boundary) with one workspace sized for a chunk: the per-frame losses equal the whole-batch step's bit for bit,
    the parameters after each step agree to float rounding (the chunks' gradients are summed in another order), the
    captured chunked step replays bit-identically to its eager sequence, and an overflow in any chunk is skipped
    and retried."""
    from us3d.scenes import multi_sweep

    specs = [(700, 3, 64, 48), (900, 2, 48, 48)]
    scene, target, frames, sweeps = multi_sweep(specs, cube_mm=64.0, seed=13, length_mm=30.0, spacing=0.3)
    cfg = cfg_for(frames[0])
    with torch.no_grad():
        gt = render_frames(*[t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)], None,
                           frames, cfg, sweeps=sweeps)[0]
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]
    whole = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=False)
    eager = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=False, chunk_frames=2)
    graph = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=True, chunk_frames=2)
    assert len(eager.chunks) == 3 and eager.workspace.numel() < whole.workspace.numel()
    for step in range(3):
        for t in (whole, eager, graph):
            t.step()
        torch.cuda.synchronize()
        for x, y, z in zip(whole.state(), eager.state(), graph.state()):
            assert torch.equal(y, z)
            assert float((x - y).abs().max()) <= 1e-5 * max(1.0, float(x.abs().max()))
        assert torch.equal(eager.losses(), graph.losses())
        if step == 0:  # same parameters: the same images, so the same per-frame losses
            assert torch.equal(whole.losses(), eager.losses())
    # an overflow in a chunk: Adam skips the step, the capacity grows, the step is repeated
    small = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=True, chunk_frames=2)
    ref = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=True, chunk_frames=2)
    small._allocate((64, 16))
    small.step()
    ref.step()
    assert small.k_capacity > 64 and small.v_capacity > 16 and int(small.step_count.item()) == 1
    for x, y in zip(small.state(), ref.state()):
        assert torch.equal(x, y)


def test_autotuner_picks_and_caches_a_tile_size(dev, tmp_path, monkeypatch):
    """The autotuner times every tile size at every LDS batch size (per kernel), keeps a tile size and a batch size
    for the forward and for the adjoint, caches the choice, and its choice trains like the defaults: the step with
    the tuned configuration gives the same per-frame losses as the default configuration's on its first step."""
    from us3d.autotune import CANDIDATES, LDS_BATCHES, apply, tune_tile_size
    from us3d.scenes import multi_sweep

    monkeypatch.setenv("US3D_CACHE_DIR", str(tmp_path))
    specs = [(600, 3, 64, 64), (400, 2, 48, 64)]
    scene, target, frames, sweeps = multi_sweep(specs, cube_mm=64.0, seed=5, length_mm=30.0, spacing=0.3)
    cfg = cfg_for(frames[0])
    with torch.no_grad():
        gt = render_frames(*[t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)], None,
                           frames, cfg, sweeps=sweeps)[0]
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]
    first = tune_tile_size(*init, frames, gt, cfg, sample_frames=4, reps=1, sweeps=sweeps)
    assert first["tile_size"] in CANDIDATES and not first["cached"]
    assert first["fwd_lds_batch"] in LDS_BATCHES and first["bwd_lds_batch"] in LDS_BATCHES
    assert set(first["ms_per_frame"]) == {str(c) for c in CANDIDATES}
    assert all(set(v) == {str(b) for b in LDS_BATCHES} for v in first["lds_batch_ms"].values())
    again = tune_tile_size(*init, frames, gt, cfg, sample_frames=4, reps=1, sweeps=sweeps)
    assert again["cached"] and all(again[k] == first[k] for k in ("tile_size", "fwd_lds_batch", "bwd_lds_batch"))
    tuned = apply(cfg, first)
    a = SliceSplatTrainer(*init, frames, gt, replace(cfg, tile_size=tuned.tile_size), sweeps=sweeps, use_graph=False)
    b = SliceSplatTrainer(*init, frames, gt, tuned, sweeps=sweeps, use_graph=False)
    a.step()
    b.step()
    assert torch.equal(a.losses(), b.losses())


def test_trainer_gradients_match_torch_autograd_through_the_activations(dev):
    """The step passes the opacity/intensity logits to the library: the pack applies the sigmoid and the reduce
    its chain (pipeline_god_file.py:369-373). Its gradients equal torch autograd of the same loss through
    torch.sigmoid and the differentiable renderer (which is given the activated values)."""
    from us3d.pipeline import simple_ssim

    frames, gt, cfg, init = _setup(dev)
    tr = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=False, lr=0.0)
    tr.step()
    torch.cuda.synchronize()
    ps = [t.detach().clone().requires_grad_(True) for t in init]
    img = render_frames(ps[0], ps[1], torch.sigmoid(ps[2]), torch.sigmoid(ps[3]), None, frames, cfg)[0]
    F = len(frames)
    P = img.numel() // F
    per = []
    for f in range(F):
        x, y = img[f * P:(f + 1) * P], gt.reshape(-1)[f * P:(f + 1) * P]
        per.append((x - y).abs().mean() + 0.2 * (1 - simple_ssim(x, y)))
    loss = torch.stack(per).mean()
    loss.backward()
    assert abs(float(tr.losses().mean()) - float(loss.detach())) < 1e-5 * max(1.0, abs(float(loss.detach())))
    for name, g, p in zip(("means", "tril", "raw_opacity", "raw_intensity"), tr.grads, ps):
        ref = p.grad.detach().cpu().numpy()
        assert np.abs(ref).max() > 0, name
        err = rel_l2(g.detach().cpu().numpy(), ref)
        assert err < 1e-4, f"{name}: rel L2 {err:.3e}"
