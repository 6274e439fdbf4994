"""Shared helpers of the UltraG-Ray tests: HIP path and oracle on the same seeded inputs."""

from __future__ import annotations

import math

import numpy as np
import torch

from oracle import ur as our
from us3d.batch import CapacityHint, FrameSpec, SweepSpec
from us3d.scenes import make_ray_scene, ray_geometry, straight_sweep
from us3d.ultragray import UltragrayRenderConfig, render_ray_frames


def small_ray_case(n=512, seed=0, n_frames=2, W=16, H=96, d_lat=0.8, d_ax=0.25):
    scene = make_ray_scene(n, 48.0, seed=seed, log_scale_mean=np.log(2.5))
    frames = straight_sweep(n_frames, H, W, d_lat, 48.0, 6.0, 5.0, seed=seed + 2)
    for f in frames:
        f.delta_ax = d_ax
        f.origin = f.origin + torch.tensor([0.0, (H * d_lat - H * d_ax) / 2.0, 0.0])
    return scene, frames


def ray_cfg(fr: FrameSpec, **kw) -> UltragrayRenderConfig:
    base = dict(image_width=fr.width, image_height=fr.height, width_mm=fr.width * fr.delta_lat, near_mm=fr.near_mm,
                lateral_col_offset=fr.col_offset, ops_max_offset_mm=0.0)
    base.update(kw)
    return UltragrayRenderConfig(**base)


def ray_spec(fr: FrameSpec, cfg: UltragrayRenderConfig) -> FrameSpec:
    return FrameSpec(fr.origin, fr.axis_lat, fr.axis_ax, fr.axis_ele, fr.height, fr.width, cfg.pixel_spacing(), fr.delta_ax,
                     fr.near_mm, fr.col_offset, fr.sweep)


def oracle_frame(fr: FrameSpec, cfg: UltragrayRenderConfig, delta=None, pixels=None):
    pix, ro, rd = ray_geometry(fr)
    if pixels is not None:
        pix = pixels
    n = lambda t: t.detach().cpu().numpy().astype(np.float32)  # noqa: E731
    W = fr.width
    return our.RayFrame(origin=n(fr.origin), axis_lat=n(fr.axis_lat), axis_ax=n(fr.axis_ax), axis_ele=n(fr.axis_ele),
                        W=W, H=fr.height, width_mm=float(cfg.width_mm), near_mm=float(cfg.near_mm),
                        col_offset=float(cfg.lateral_col_offset), pixels=n(pix), ray_o=n(ro), ray_d=n(rd),
                        delta=np.zeros(W, np.float32) if delta is None else np.asarray(delta, np.float32))


def ray_settings(cfg: UltragrayRenderConfig) -> our.RaySettings:
    return our.RaySettings(k_sigma=cfg.k_sigma, Rw=cfg.quadrature_window_radius_mm,
                           elev_margin=cfg.elev_cutoff_margin_mm, local_only=cfg.local_transmission_window_only,
                           local_max=cfg.local_transmission_max_window_mm, grad_threshold=cfg.grad_threshold)


def gpu_ray_render(scene, frames, cfg, dev, deltas=None, d_image=None, hint=None, pixels=None):
    """pixels: optional list of per-frame (W, H, 3) sample positions replacing the linear grid."""
    params = [t.to(dev).clone().requires_grad_(True) for t in scene.params()]
    geos = [ray_geometry(f) for f in frames]
    if pixels is None:
        pix = torch.cat([g[0].reshape(-1, 3) for g in geos]).to(dev)
    else:
        pix = torch.cat([p.reshape(-1, 3) for p in pixels]).to(dev)
    ro = torch.cat([g[1] for g in geos]).to(dev)
    rd = torch.cat([g[2] for g in geos]).to(dev)
    W = sum(f.width for f in frames)
    off = torch.zeros(W, device=dev) if deltas is None else torch.as_tensor(np.concatenate(deltas), device=dev)
    specs = [ray_spec(f, cfg) for f in frames]
    image, batch = render_ray_frames(*params, pix, ro, rd, off, specs, cfg,
                                     sweeps=[SweepSpec(0, scene.means.shape[0])], hint=hint or CapacityHint(),
                                     return_batch=True)
    out = {"image": image, "batch": batch}
    if d_image is not None:
        (image * torch.as_tensor(d_image, dtype=torch.float32, device=dev)).sum().backward()
        out["grads"] = [p.grad.detach().cpu().numpy().astype(np.float64) for p in params]
    return out


# Rotations built independently of the oracle: Rodrigues' formula and the Hamilton product of (w, x, y, z)
# quaternions (a unit quaternion (cos a/2, sin a/2 n) is the rotation by a about n; q_a (x) q_b is R_a R_b).


def rodrigues(axis, deg):
    n = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    a = math.radians(deg)
    return np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * K @ K


def axis_quat(axis, deg):
    n = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    a = math.radians(deg) / 2
    return np.array([math.cos(a), *(math.sin(a) * n)])


def hamilton(p, q):
    w1, x1, y1, z1 = p
    w2, x2, y2, z2 = q
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def composed_rotation():
    """30 degrees about (0, 0, 1) after 20 degrees about (1, 1, 0): q = q_a (x) q_b <-> R = R_a R_b."""
    qa, qb = axis_quat([0, 0, 1], 30), axis_quat([1, 1, 0], 20)
    R = rodrigues([0, 0, 1], 30) @ rodrigues([1, 1, 0], 20)
    assert np.abs(R - R.T).max() > 0.3  # far from symmetric: R and R^T are different rotations
    return hamilton(qa, qb), R
