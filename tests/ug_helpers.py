"""Shared helpers of the UltraGauss tests: run the HIP path and the oracle on the same seeded inputs."""

from __future__ import annotations

import numpy as np
import torch

from oracle import ug as oug
from us3d.batch import CapacityHint, SweepSpec
from us3d.scenes import config1, linear_pixels, make_scene, straight_sweep
from us3d.ultragauss import UltraGaussRenderConfig, render_frames


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def small_case(n=4096, seed=0, n_frames=4, H=128, W=128, spacing=0.3):
    scene = make_scene(n, 64.0, seed=seed)
    frames = straight_sweep(n_frames, H, W, spacing, 64.0, 40.0, 5.0, seed=seed + 2)
    return scene, frames


def cfg_for(fr, **kw) -> UltraGaussRenderConfig:
    base = dict(H=fr.height, W=fr.width, lateral_spacing_mm_per_pix=fr.delta_lat,
                axial_spacing_mm_per_pix=fr.delta_ax, axial_near_mm=fr.near_mm, lateral_col_offset=fr.col_offset,
                fixed_bg_alpha=1e-6)
    base.update(kw)
    return UltraGaussRenderConfig(**base)


def gpu_render(scene, frames, cfg, dev, pixels="explicit", grads=None, hint=None):
    """Render frames as one batch; optionally backpropagate `grads` = (dI, dA, dD) flat arrays."""
    ts = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    pix = torch.cat([linear_pixels(f) for f in frames]).to(dev) if pixels == "explicit" else None
    image, alpha, depth, batch = render_frames(*ts, pix, frames, cfg, sweeps=[SweepSpec(0, scene.means.shape[0])],
                                               hint=hint or CapacityHint(), return_batch=True)
    out = {"image": image, "alpha": alpha, "depth": depth, "batch": batch}
    if grads is not None:
        dI, dA, dD = (None if g is None else torch.as_tensor(g, dtype=torch.float32, device=dev) for g in grads)
        loss = (image * dI).sum()
        if dA is not None:
            loss = loss + (alpha * dA).sum()
        if dD is not None:
            loss = loss + (depth * dD).sum()
        loss.backward()
        out["grads"] = [t.grad.detach().cpu().numpy().astype(np.float64) for t in ts]
    return out


def oracle_frame(scene, fr, sem: oug.Semantics, ts=16, sort_by_depth=True):
    geom = oug.FrameGeom.from_spec(fr)
    pix = linear_pixels(fr).numpy()
    m, t = scene.means.numpy(), scene.tril.numpy()
    mr = oug.map_frame(m, t, geom, chi2=sem.chi2, beta=sem.beta, sort_by_depth=sort_by_depth, ts=ts)
    lists = oug.pixel_lists(m, t, pix, geom, mr, chi2=sem.chi2, beta=sem.beta, ts=ts)
    return geom, pix, mr, lists
