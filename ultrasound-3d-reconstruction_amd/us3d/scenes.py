"""Seeded synthetic scenes and probe sweeps for the benchmark workloads.

Gaussians: means uniform in a cube, per-axis log-scales N(log 2 mm, 0.3), rotations uniform on S^3
(normalised N(0, I_4), w-first), intensity sigmoid(N(0,1)), opacity sigmoid(N(0,1)). The slice splat takes
the precision parameterisation of the reference (pipeline_god_file.py:380-389): Lambda = R diag(1/s^2) R^T
factored as L L^T (Cholesky), raw diag = sqrt(L_ii - beta), raw off-diagonals = L_ij, in the order
(l11, l12, l13, l22, l23, l33).

Sweeps: a linear probe (lateral x, axial y, elevational z at rest). The straight sweep translates along
the elevational axis with uniform tilts; the freehand sweep is a Gaussian random walk in translation and
rotation.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List

import numpy as np
import torch

from .batch import FrameSpec, SweepSpec


@dataclass
class Scene:
    means: torch.Tensor       # (G, 3)
    scales: torch.Tensor      # (G, 3)
    quats: torch.Tensor       # (G, 4) w, x, y, z (unit)
    opacity: torch.Tensor     # (G,)
    intensity: torch.Tensor   # (G,)
    tril: torch.Tensor        # (G, 6) raw precision factors

    def to(self, device) -> "Scene":
        return Scene(*(t.to(device) for t in (self.means, self.scales, self.quats, self.opacity, self.intensity,
                                               self.tril)))


def quat_to_rot(q: torch.Tensor) -> torch.Tensor:
    """(G,4) w,x,y,z -> (G,3,3) rotation, columns = principal axes."""
    q = q / q.norm(dim=-1, keepdim=True).clamp_min(1e-12)
    w, x, y, z = q.unbind(-1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y),
    ], dim=-1)
    return R.reshape(q.shape[:-1] + (3, 3))


def precision_tril_from_scales_quats(scales: torch.Tensor, quats: torch.Tensor, beta: float = 1e-4) -> torch.Tensor:
    """Raw (l11, l12, l13, l22, l23, l33) with L L^T = R diag(1/s^2) R^T and l_ii = raw_ii^2 + beta."""
    R = quat_to_rot(quats.double())
    lam = R @ torch.diag_embed(1.0 / scales.double() ** 2) @ R.transpose(-1, -2)
    L = torch.linalg.cholesky(lam)
    diag = torch.diagonal(L, dim1=-2, dim2=-1)
    raw_diag = (diag - beta).clamp_min(1e-12).sqrt()
    out = torch.stack([raw_diag[:, 0], L[:, 1, 0], L[:, 2, 0], raw_diag[:, 1], L[:, 2, 1], raw_diag[:, 2]], dim=1)
    return out.float()


def make_scene(n: int, cube_mm: float, seed: int, box_mm=None, beta: float = 1e-4,
               log_scale_mean: float = math.log(2.0), log_scale_std: float = 0.3) -> Scene:
    g = torch.Generator().manual_seed(int(seed))
    box = torch.tensor(box_mm if box_mm is not None else [cube_mm] * 3, dtype=torch.float32)
    means = torch.rand((n, 3), generator=g) * box
    scales = torch.exp(log_scale_mean + log_scale_std * torch.randn((n, 3), generator=g))
    quats = torch.randn((n, 4), generator=g)
    quats = quats / quats.norm(dim=1, keepdim=True)
    intensity = torch.sigmoid(torch.randn(n, generator=g))
    opacity = torch.sigmoid(torch.randn(n, generator=g))
    tril = precision_tril_from_scales_quats(scales, quats, beta)
    return Scene(means, scales, quats, opacity, intensity, tril)


def _rot(axis: np.ndarray, ang: float) -> np.ndarray:
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _frame(origin, R, H, W, d_lat, d_ax, near=0.0) -> FrameSpec:
    R = np.asarray(R, dtype=np.float64)
    t = lambda v: torch.tensor(np.asarray(v, dtype=np.float32))  # noqa: E731
    return FrameSpec(origin=t(origin), axis_lat=t(R[:, 0]), axis_ax=t(R[:, 1]), axis_ele=t(R[:, 2]), height=H,
                     width=W, delta_lat=d_lat, delta_ax=d_ax, near_mm=near, col_offset=(W - 1) * 0.5)


def straight_sweep(n_frames: int, H: int, W: int, spacing: float, cube_mm: float, length_mm: float, tilt_deg: float,
                   seed: int) -> List[FrameSpec]:
    """Straight sweep along the elevational axis through the cube centre, uniform +/-tilt about lat and ax."""
    rng = np.random.default_rng(seed)
    c = cube_mm / 2
    depth = H * spacing
    frames = []
    for i in range(n_frames):
        z = c - length_mm / 2 + length_mm * (i / max(1, n_frames - 1))
        R = _rot(np.array([1.0, 0, 0]), math.radians(rng.uniform(-tilt_deg, tilt_deg))) @ \
            _rot(np.array([0, 1.0, 0]), math.radians(rng.uniform(-tilt_deg, tilt_deg)))
        origin = np.array([c, c - depth / 2, z])
        frames.append(_frame(origin, R, H, W, spacing, spacing))
    return frames


def freehand_sweep(n_frames: int, H: int, W: int, spacing: float, cube_mm: float, length_mm: float, seed: int,
                   jitter_mm: float = 0.1, rot_sigma_deg: float = 2.0, spacing_ax=None) -> List[FrameSpec]:
    """Freehand sweep: translation = cumulative N(length/n along elevation, jitter); rotation = cumulative
    per-frame N(0, rot_sigma) about the three probe axes. spacing is lateral (and axial unless spacing_ax)."""
    rng = np.random.default_rng(seed)
    d_ax = spacing if spacing_ax is None else spacing_ax
    c = cube_mm / 2
    depth = H * d_ax
    pos = np.array([c, c - depth / 2, c - length_mm / 2])
    R = np.eye(3)
    step = length_mm / n_frames
    frames = []
    for _ in range(n_frames):
        frames.append(_frame(pos.copy(), R.copy(), H, W, spacing, d_ax))
        pos = pos + R[:, 2] * step + rng.normal(0.0, jitter_mm, 3)
        for ax in range(3):
            R = R @ _rot(np.eye(3)[ax], math.radians(rng.normal(0.0, rot_sigma_deg)))
        u, _, vt = np.linalg.svd(R)
        R = u @ vt
    return frames


def linear_pixels(frame: FrameSpec, device=None) -> torch.Tensor:
    """(H*W, 3) pixel grid, the same torch ops as NpzUltrasoundDataset._pixel_positions
    (pipeline_god_file.py:292-306)."""
    dev = device or frame.origin.device
    u = torch.arange(frame.width, device=dev, dtype=torch.float32)
    v = torch.arange(frame.height, device=dev, dtype=torch.float32)
    lateral_mm = (u - float(frame.col_offset)) * float(frame.delta_lat)
    axial_mm = float(frame.near_mm) + v * float(frame.delta_ax)
    yy, xx = torch.meshgrid(axial_mm, lateral_mm, indexing="ij")
    o, lat, ax = (t.to(dev, torch.float32) for t in (frame.origin, frame.axis_lat, frame.axis_ax))
    pix = o.view(1, 1, 3) + xx[..., None] * lat.view(1, 1, 3) + yy[..., None] * ax.view(1, 1, 3)
    return pix.reshape(-1, 3).contiguous()


def fan_pixels(frame: FrameSpec, device=None) -> torch.Tensor:
    """(H*W, 3) pixel grid of a fan frame (geometry 1): angle t = (u - c) * d_t, radius r = near + v * d_r,
    p = apex + r sin(t) a_lat + r cos(t) a_ax -- the grid the kernels evaluate for such frames when no pixel
    array is given, up to the last-bit rounding of sin/cos."""
    dev = device or frame.origin.device
    u = torch.arange(frame.width, device=dev, dtype=torch.float32)
    v = torch.arange(frame.height, device=dev, dtype=torch.float32)
    t = (u - float(frame.col_offset)) * float(frame.delta_lat)
    r = float(frame.near_mm) + v * float(frame.delta_ax)
    rr, tt = torch.meshgrid(r, t, indexing="ij")
    lat_mm, ax_mm = rr * torch.sin(tt), rr * torch.cos(tt)
    o, lat, ax = (x.to(dev, torch.float32) for x in (frame.origin, frame.axis_lat, frame.axis_ax))
    pix = o.view(1, 1, 3) + lat_mm[..., None] * lat.view(1, 1, 3) + ax_mm[..., None] * ax.view(1, 1, 3)
    return pix.reshape(-1, 3).contiguous()


def fan_sweep(n_frames: int, H: int, W: int, fan_deg: float, r0: float, r1: float, box_mm, length_mm: float,
              seed: int, jitter_mm: float = 0.1, rot_sigma_deg: float = 1.0) -> List[FrameSpec]:
    """Freehand sweep of a fan (sector) probe: W beams over fan_deg, H depth samples over [r0, r1) mm from the
    apex; the apex rides above the box's y = 0 face so the sector's centre line crosses the box's middle, and
    moves along elevation (z) with jitter and small cumulative rotations."""
    rng = np.random.default_rng(seed)
    bx, by, bz = (float(b) for b in box_mm)
    pos = np.array([bx / 2, by / 2 - (r0 + r1) / 2, bz / 2 - length_mm / 2])
    R = np.eye(3)
    step = length_mm / n_frames
    frames = []
    for _ in range(n_frames):
        f = _frame(pos.copy(), R.copy(), H, W, math.radians(fan_deg) / W, (r1 - r0) / H, near=r0)
        f.geometry = 1
        frames.append(f)
        pos = pos + R[:, 2] * step + rng.normal(0.0, jitter_mm, 3)
        for a in range(3):
            R = R @ _rot(np.eye(3)[a], math.radians(rng.normal(0.0, rot_sigma_deg)))
        u_, _, vt = np.linalg.svd(R)
        R = u_ @ vt
    return frames


# ---------------------------------------------------------------- benchmark workloads

def config1(device="cpu"):
    """4,096 Gaussians in a 64 mm cube; 16 frames 128x128 at 0.3 mm, straight 40 mm sweep, +/-5 deg tilt."""
    scene = make_scene(4096, 64.0, seed=0).to(device)
    target = make_scene(4096, 64.0, seed=1).to(device)
    frames = straight_sweep(16, 128, 128, 0.3, 64.0, 40.0, 5.0, seed=2)
    return scene, target, frames


def config2(device="cpu", n_frames: int = 256, n_gauss: int = 200_000):
    """200k Gaussians in a 120 mm cube; 256 frames 512x512 at 0.2 mm along a 150 mm freehand sweep."""
    scene = make_scene(n_gauss, 120.0, seed=0).to(device)
    target = make_scene(n_gauss, 120.0, seed=1).to(device)
    frames = freehand_sweep(n_frames, 512, 512, 0.2, 120.0, 150.0, seed=2)
    return scene, target, frames


def multi_sweep(specs, cube_mm: float = 120.0, seed: int = 0, length_mm: float = 150.0, spacing: float = 0.2):
    """Several sweeps in one batch. specs: [(n_gauss, n_frames, H, W), ...]. Sweep s gets its own Gaussians
    (scene seed 100 s + seed, target seed 100 s + seed + 1) and its own freehand path (seed 100 s + seed + 2).
    Returns the concatenated scene and target, the frames (grouped by sweep, f.sweep = s) and the sweeps'
    Gaussian row ranges."""
    scenes, targets, frames, sweeps = [], [], [], []
    g0 = 0
    for s, (G, F, H, W) in enumerate(specs):
        base = 100 * s + seed
        scenes.append(make_scene(G, cube_mm, seed=base))
        targets.append(make_scene(G, cube_mm, seed=base + 1))
        for f in freehand_sweep(F, H, W, spacing, cube_mm, length_mm, seed=base + 2):
            f.sweep = s
            frames.append(f)
        sweeps.append(SweepSpec(g0, G))
        g0 += G

    def cat(sc):
        return Scene(*(torch.cat([getattr(x, k) for x in sc]) for k in
                       ("means", "scales", "quats", "opacity", "intensity", "tril")))

    return cat(scenes), cat(targets), frames, sweeps


def config4_specs(seed: int = 4, n_sweeps: int = 8, scale: float = 1.0):
    """Sweep sizes of the batched multi-sweep workload: Gaussians log-uniform in [50k, 500k], frames uniform in
    32..400, frame size from {256x256, 512x512, 640x480 (W x H)}; `scale` shrinks Gaussian and frame counts."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_sweeps):
        G = int(round(math.exp(rng.uniform(math.log(50_000), math.log(500_000))) * scale))
        F = int(rng.integers(32, 401))
        H, W = [(256, 256), (512, 512), (480, 640)][int(rng.integers(0, 3))]
        out.append((max(G, 1), max(int(round(F * scale)), 1), H, W))
    return out


def config4(seed: int = 4, scale: float = 1.0):
    """8 sweeps of uneven size (config4_specs) with config-2 Gaussian and pose distributions."""
    return multi_sweep(config4_specs(seed, scale=scale), seed=seed)


def config5(n_frames: int = 1000, n_gauss: int = 2_000_000, H: int = 480, W: int = 640):
    """Fan-probe workload: Gaussians (config-2 distributions) in a 200 x 150 x 150 mm box; n_frames fan frames of
    W beams over 60 degrees and H depth samples over 40-160 mm, swept 150 mm along elevation; target scene from
    seed 1."""
    box = [200.0, 150.0, 150.0]
    scene = make_scene(n_gauss, 150.0, seed=0, box_mm=box)
    target = make_scene(n_gauss, 150.0, seed=1, box_mm=box)
    frames = fan_sweep(n_frames, H, W, 60.0, 40.0, 160.0, box, 150.0, seed=2)
    return scene, target, frames


# ---------------------------------------------------------------- UltraG-Ray scenes and scanline geometry

@dataclass
class RayScene:
    means: torch.Tensor       # (G, 3)
    scales: torch.Tensor      # (G, 3)
    quats: torch.Tensor       # (G, 4)
    atten: torch.Tensor       # (G,) raw logits of tau
    sh: torch.Tensor          # (G, 4)

    def params(self):
        return self.means, self.scales, self.quats, self.atten, self.sh

    def to(self, device) -> "RayScene":
        return RayScene(*(t.to(device) for t in self.params()))


def attenuation_to_raw_tau(alpha_db_cm_mhz: torch.Tensor, path_mm: torch.Tensor, f_mhz: float = 5.0) -> torch.Tensor:
    """raw = logit(tau), tau = 10^(-alpha f L / 20): amplitude transmission across a path of L mm (0.1 L cm)."""
    loss_db = alpha_db_cm_mhz * f_mhz * (0.1 * path_mm)
    tau = torch.pow(10.0, -loss_db / 20.0).clamp(1e-6, 1 - 1e-6)
    return torch.log(tau / (1 - tau))


def make_ray_scene(n: int, cube_mm: float, seed: int, log_scale_mean: float = math.log(2.0),
                   log_scale_std: float = 0.3) -> RayScene:
    """Echogenicity sigmoid(N(0,1)) as SH c0 (c1..c3 = 0); attenuation U(0.3, 1.0) dB/cm/MHz at 5 MHz over a
    path of 2 max(s), mapped to the raw tau logit the renderer learns."""
    g = torch.Generator().manual_seed(int(seed))
    means = torch.rand((n, 3), generator=g) * cube_mm
    scales = torch.exp(log_scale_mean + log_scale_std * torch.randn((n, 3), generator=g))
    quats = torch.randn((n, 4), generator=g)
    quats = quats / quats.norm(dim=1, keepdim=True)
    echo = torch.sigmoid(torch.randn(n, generator=g))
    alpha = 0.3 + 0.7 * torch.rand(n, generator=g)
    raw = attenuation_to_raw_tau(alpha, 2.0 * scales.max(dim=1).values)
    sh = torch.zeros((n, 4))
    sh[:, 0] = echo
    return RayScene(means, scales, quats, raw.float(), sh)


def ray_geometry(frame: FrameSpec, device=None):
    """pixels (W, H, 3), ray origins (W, 3), ray dirs (W, 3) of a linear probe, from the pipeline's grid
    (pipeline_god_file.py:292-333; the UltraG-Ray branch permutes the grid to (W, H, 3), :517)."""
    dev = device or frame.origin.device
    grid = linear_pixels(frame, dev).reshape(frame.height, frame.width, 3).permute(1, 0, 2).contiguous()
    u = torch.arange(frame.width, device=dev, dtype=torch.float32)
    lateral_mm = (u - float(frame.col_offset)) * float(frame.delta_lat)
    o, lat, ax = (t.to(dev, torch.float32) for t in (frame.origin, frame.axis_lat, frame.axis_ax))
    ray_o = o.view(1, 3) + lateral_mm[:, None] * lat.view(1, 3)
    ray_d = ax.view(1, 3).expand(frame.width, 3).contiguous()
    return grid, ray_o.contiguous(), ray_d


def config3(device="cpu", n_frames: int = 128, n_gauss: int = 200_000):
    """200k Gaussians (120 mm cube); 128 frames of 128 scanlines x 1024 samples at 0.8 mm x 0.1 mm."""
    scene = make_ray_scene(n_gauss, 120.0, seed=0).to(device)
    frames = freehand_sweep(n_frames, 1024, 128, 0.8, 120.0, 150.0, seed=2, spacing_ax=0.1)
    return scene, frames
