"""Pins of the GaussianPile us_focus oracle (CPU only).

The oracle restates the reference's CUDA footprint, which forms the world covariance as R diag(s^2) R from the
rotation's row-major entries (plane_cond2d_device.cuh:48-69). The package's own reference math
(gaussianplie_renderer/reference.py:54-115) forms R diag(s^2) R^T; the two agree exactly when R is symmetric
(identity and half-turns), so on such rotations, with the epsilons small and the k-sigma box wide, the
oracle must equal that reference math. Its FP64 restatement must be differentiable with gradients that
central differences confirm.
"""

import math

import numpy as np
import pytest
import torch

from oracle import gp
from us3d.scenes import make_scene


def _frame():
    return gp.GPFrame(origin=torch.tensor([0.0, 0.0, 10.0]), axis_lat=torch.tensor([1.0, 0.0, 0.0]),
                      axis_ax=torch.tensor([0.0, 1.0, 0.0]), axis_ele=torch.tensor([0.0, 0.0, 1.0]), H=40, W=32,
                      delta_lat=0.4, delta_ax=0.3, near=0.0, col_offset=-1.0)


def test_oracle_equals_reference_math_for_symmetric_rotations():
    G = 200
    sc = make_scene(G, 20.0, seed=3)
    g = torch.Generator().manual_seed(1)
    n = torch.randn(G, 3, generator=g)
    n = n / n.norm(dim=1, keepdim=True)
    q = torch.cat([torch.zeros(G, 1), n], 1)  # half-turns: R = 2 n n^T - I, symmetric
    q[::3] = torch.tensor([1.0, 0.0, 0.0, 0.0])
    fr = _frame()
    st = gp.GPSettings(k_sigma=8.0, sigma_ele=1.5, det_eps=1e-12, weight_min=0.0)
    img = gp.render(sc.means.double(), sc.scales.double(), q.double(), sc.opacity.double(),
                    torch.ones(G, dtype=torch.float64), fr, st, fp64=True)
    ref = gp.render_reference_math(sc.means, sc.scales, q, sc.opacity, fr, 1.5)
    assert float(ref.abs().max()) > 0.1
    assert float((img - ref).norm() / ref.norm()) < 1e-6


def test_oracle_fp32_and_fp64_agree_and_general_rotations_differ_from_rt_form():
    G = 150
    sc = make_scene(G, 20.0, seed=4)
    fr = _frame()
    st = gp.GPSettings(k_sigma=3.0, sigma_ele=1.0)
    args = (sc.means, sc.scales, sc.quats, sc.opacity)
    img32 = gp.render(*args, torch.ones(G), fr, st)
    img64 = gp.render(*(t.double() for t in args), torch.ones(G, dtype=torch.float64), fr, st, fp64=True)
    assert float((img32.double() - img64).norm() / img64.norm()) < 1e-5
    # with general rotations the reference's CUDA covariance (R S^2 R) is not the R S^2 R^T of reference.py
    ref = gp.render_reference_math(*args, fr, 1.0)
    assert float((img64 - ref).norm() / ref.norm()) > 1e-2


def test_oracle_fp64_gradients_match_central_differences():
    G = 6
    sc = make_scene(G, 6.0, seed=5)
    fr = gp.GPFrame(origin=torch.tensor([-3.0, -3.0, 3.0]), axis_lat=torch.tensor([1.0, 0.0, 0.0]),
                    axis_ax=torch.tensor([0.0, 1.0, 0.0]), axis_ele=torch.tensor([0.0, 0.0, 1.0]), H=16, W=16,
                    delta_lat=0.4, delta_ax=0.4, near=0.0, col_offset=0.0)
    st = gp.GPSettings(k_sigma=30.0, sigma_ele=1.2, weight_min=0.0)  # a wide box: no decision near a pixel
    ps = [t.double().clone().requires_grad_(True) for t in (sc.means, sc.scales, sc.quats, sc.opacity)]
    inten = torch.rand(G, dtype=torch.float64, generator=torch.Generator().manual_seed(2)).requires_grad_(True)
    w = torch.randn(fr.H, fr.W, dtype=torch.float64, generator=torch.Generator().manual_seed(3))

    def loss(*xs):
        return (gp.render(*xs[:4], xs[4], fr, st, fp64=True) * w).sum()

    L = loss(*ps, inten)
    grads = torch.autograd.grad(L, ps + [inten])
    h = 1e-6
    rng = np.random.default_rng(0)
    for k, (p, gr) in enumerate(zip(ps + [inten], grads)):
        flat = p.detach().reshape(-1)
        for idx in rng.choice(flat.numel(), size=min(6, flat.numel()), replace=False):
            xs_p = [t.detach().clone() for t in ps + [inten]]
            xs_m = [t.detach().clone() for t in ps + [inten]]
            xs_p[k].reshape(-1)[idx] += h
            xs_m[k].reshape(-1)[idx] -= h
            fd = (float(loss(*xs_p)) - float(loss(*xs_m))) / (2 * h)
            an = float(gr.reshape(-1)[idx])
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)), (k, int(idx), fd, an)


def test_brute_force_pixels_equal_the_tile_list_render():
    """render_pixels (every Gaussian, no tile lists) against render (tile lists): the same FP32 image bit for
    bit and the same FP64 image, so the brute force can stand in for the tile-list oracle at sizes where only
    sampled pixels are affordable."""
    G = 150
    sc = make_scene(G, 20.0, seed=4)
    fr = _frame()
    st = gp.GPSettings()
    args = (sc.means, sc.scales, sc.quats, sc.opacity)
    rows = torch.arange(fr.H).repeat_interleave(fr.W)
    cols = torch.arange(fr.W).repeat(fr.H)
    img = gp.render(*args, torch.ones(G), fr, st)
    assert float(img.max()) > 0.05
    assert torch.equal(gp.render_pixels(*args, torch.ones(G), fr, st, rows, cols).reshape(fr.H, fr.W), img)
    a64 = [t.double() for t in args]
    one = torch.ones(G, dtype=torch.float64)
    img64 = gp.render(*a64, one, fr, st, fp64=True)
    bf64 = gp.render_pixels(*a64, one, fr, st, rows, cols, fp64=True).reshape(fr.H, fr.W)
    assert float((bf64 - img64).abs().max()) < 1e-12
