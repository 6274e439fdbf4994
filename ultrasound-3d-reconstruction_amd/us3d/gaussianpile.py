"""GaussianPile slice renderer on MI355X: config, renderer facade, autograd function and batch driver.

Mirrors the reference package ops/gaussianplie_renderer/gaussianplie_renderer/ (types.py:5-48, api.py:10-154,
autograd.py:17-246): the same config fields and validation, the same facade signature
forward(gauss_means, gauss_scales, gauss_quats, gauss_alphas, pose_probe_to_world, ...) -> image (H, W) and the
same two backends:
  "ours_focus"        -- the us_focus slice splat (libus3d us3d_gp_*: per-pair focus footprint, tile mapper,
                         additive FP32 splat, exact adjoint), intensities fixed to one as the facade does;
  "ultragauss_lowmem" -- the UltraGauss splat with chi2 = radius_sigma^2, no depth sort, background alpha 1e-6,
                         beta = cov_eps and unit intensities (api.py:26-50), run by us3d.ultragauss.
The reference's "ours_focus" path reads its pixel grid from a slice-grid pack built by the absent module
ours_renderer.grid_geometry; here the linear-probe grid comes from the config (H, W, spacings, near,
lateral_col_offset) and set_slice_grid_geometry only checks and stores its argument.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import torch
import torch.nn as nn
import torch.nn.functional as F

from . import _lib
from .batch import CapacityHint, DeviceTables, FrameSpec, Mapped, SweepSpec, new_workspace, plan_layout, read_stats


@dataclass
class GaussianPlieRenderConfig:
    """ops/gaussianplie_renderer/gaussianplie_renderer/types.py:5-48 (+ work_queue: the forward and the adjoint run as
    persistent workgroups over per-XCD (frame, tile) queues instead of one workgroup per tile; same outputs)."""

    H: int
    W: int
    lateral_spacing_mm_per_pix: float
    axial_spacing_mm_per_pix: float
    axial_near_mm: float
    backend: str = "ours_focus"
    tile_size_px: int = 16
    radius_sigma: float = 3.0
    sigma_z_mm: float = 1.0
    prefilter_elevational_radius_mm: float = 0.0
    det_eps: float = 1.0e-6
    cov_eps: float = 1.0e-6
    opacity_cull_threshold: float = 1.0e-6
    render_chunk_size: int = 256
    enable_input_checks: bool = False
    lateral_col_offset: Optional[float] = None
    work_queue: bool = True

    def __post_init__(self) -> None:
        if self.H <= 0 or self.W <= 0:
            raise ValueError(f"H/W must be positive, got H={self.H}, W={self.W}.")
        if self.lateral_spacing_mm_per_pix <= 0.0:
            raise ValueError("lateral_spacing_mm_per_pix must be > 0.")
        if self.axial_spacing_mm_per_pix <= 0.0:
            raise ValueError("axial_spacing_mm_per_pix must be > 0.")
        if self.lateral_col_offset is None:
            self.lateral_col_offset = float(self.W - 1) * 0.5
        if self.backend not in ("ours_focus", "ultragauss_lowmem"):
            raise ValueError("backend must be 'ours_focus' or 'ultragauss_lowmem'.")
        if self.tile_size_px <= 0:
            raise ValueError("tile_size_px must be > 0.")
        if self.radius_sigma <= 0.0:
            raise ValueError("radius_sigma must be > 0.")
        if self.sigma_z_mm <= 0.0:
            raise ValueError("sigma_z_mm must be > 0.")
        if self.prefilter_elevational_radius_mm < 0.0:
            raise ValueError("prefilter_elevational_radius_mm must be >= 0.")
        if self.det_eps <= 0.0 or self.cov_eps <= 0.0:
            raise ValueError("det_eps and cov_eps must be > 0.")
        if self.opacity_cull_threshold < 0.0:
            raise ValueError("opacity_cull_threshold must be >= 0.")
        if self.render_chunk_size <= 0:
            raise ValueError("render_chunk_size must be > 0.")

    def settings(self) -> _lib.GPSettings:
        """The reference's us_focus parameter vector (autograd.py:41-55) as the library's settings."""
        if self.tile_size_px not in (8, 16, 32):
            raise ValueError("tile_size_px must be 8, 16 or 32 for the MI355X kernels.")
        s = _lib.GPSettings()
        s.k_sigma = float(self.radius_sigma)
        s.sigma_ele = float(self.sigma_z_mm)
        s.det_eps = float(self.det_eps)
        s.weight_min = max(float(self.opacity_cull_threshold), 1.0e-12)
        s.tile_size = int(self.tile_size_px)
        s.work_queue = 1 if self.work_queue else 0
        return s


def frame_from_pose(pose_probe_to_world: torch.Tensor, cfg: GaussianPlieRenderConfig, sweep: int = 0) -> FrameSpec:
    """Frame of a 4x4 probe-to-world pose: origin = translation, axes = the normalised rotation columns
    (lateral, axial, elevational), as the reference's kernels read them (forward.cu:145-151)."""
    p = pose_probe_to_world.detach().to(torch.float32).cpu()
    axes = [F.normalize(p[:3, k], dim=0, eps=1e-6) for k in range(3)]
    return FrameSpec(origin=p[:3, 3].clone(), axis_lat=axes[0], axis_ax=axes[1], axis_ele=axes[2], height=cfg.H,
                     width=cfg.W, delta_lat=float(cfg.lateral_spacing_mm_per_pix),
                     delta_ax=float(cfg.axial_spacing_mm_per_pix), near_mm=float(cfg.axial_near_mm),
                     col_offset=float(cfg.lateral_col_offset), sweep=sweep)


class GPBatch:
    """Map/forward/backward of one batch of frames over the C ABI, with overflow-driven capacity growth."""

    def __init__(self, frames: Sequence[FrameSpec], sweeps: Sequence[SweepSpec], n_gauss: int,
                 settings: _lib.GPSettings, device: torch.device, hint: Optional[CapacityHint] = None):
        # the GaussianPile kernels evaluate the linear-probe grid only (the reference's slice grid,
        # slice_grid_geometry.cuh:38-45); fan frames are an UltraGauss geometry
        if any(int(f.geometry) != 0 for f in frames):
            raise ValueError("the GaussianPile renderer takes linear-probe frames only (geometry 0)")
        if not (float(settings.weight_min) > 0.0):
            raise ValueError("weight_min must be > 0 (the reference floors it at 1e-12)")
        self.settings = settings
        self.device = device
        self.layout = plan_layout(frames, sweeps, n_gauss, settings.tile_size)
        self.tables = DeviceTables(frames, self.layout, device)
        self.hint = hint or CapacityHint()

    def workspace_bytes(self, k_cap: int, v_cap: int) -> int:
        lay = self.layout
        return int(_lib.lib().us3d_gp_workspace_bytes(lay.n_gauss, lay.n_pairs, lay.n_tiles, k_cap, v_cap))

    def layout_offsets(self, k_cap: int, v_cap: int) -> _lib.Layout:
        lay = self.layout
        out = _lib.Layout()
        _lib.check(_lib.lib().us3d_gp_workspace_layout(lay.n_gauss, lay.n_pairs, lay.n_tiles, k_cap, v_cap, out),
                   "us3d_gp_workspace_layout")
        return out

    def map(self, means, scales, quats, opacity, intensity=None) -> Mapped:
        L = _lib.lib()
        n_pairs = self.layout.n_pairs
        k_cap = int(self.hint.capacity(n_pairs))
        v_cap = int(self.hint.v_capacity(n_pairs))
        while True:
            ws = new_workspace(self.workspace_bytes(k_cap, v_cap), self.device)
            m = Mapped(self.tables, ws, k_cap, v_cap)
            _lib.check(L.us3d_gp_map(m.plan(), _lib.ptr(means), _lib.ptr(scales), _lib.ptr(quats), _lib.ptr(opacity),
                                     _lib.ptr(intensity), self.settings, _lib.stream_ptr(self.device)), "us3d_gp_map")
            st = read_stats(ws)
            m.stats = st
            self.hint.update(st.n_instances, st.n_visible)
            if not st.overflow:
                return m
            k_cap, v_cap = self.hint.regrow(k_cap, v_cap, st, n_pairs)

    def forward(self, m: Mapped) -> torch.Tensor:
        image = torch.empty(self.layout.n_pixels, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().us3d_gp_forward(m.plan(), self.settings, _lib.ptr(image), _lib.stream_ptr(self.device)),
                   "us3d_gp_forward")
        return image

    def backward(self, m: Mapped, d_image: torch.Tensor):
        G = self.layout.n_gauss
        dev = self.device
        d_means = torch.empty((G, 3), dtype=torch.float32, device=dev)
        d_scales = torch.empty((G, 3), dtype=torch.float32, device=dev)
        d_quats = torch.empty((G, 4), dtype=torch.float32, device=dev)
        d_opac = torch.empty(G, dtype=torch.float32, device=dev)
        d_inten = torch.empty(G, dtype=torch.float32, device=dev)
        _lib.check(_lib.lib().us3d_gp_backward(
            m.plan(), self.settings, _lib.ptr(d_image), _lib.ptr(d_means), _lib.ptr(d_scales), _lib.ptr(d_quats),
            _lib.ptr(d_opac), _lib.ptr(d_inten), _lib.stream_ptr(dev)), "us3d_gp_backward")
        return d_means, d_scales, d_quats, d_opac, d_inten


class _GPRenderFunction(torch.autograd.Function):
    """autograd.py:58-208 for the us_focus backend: forward maps and splats, backward runs the adjoint."""

    @staticmethod
    def forward(ctx, batch: GPBatch, means, scales, quats, opacity, intensity):
        m = batch.map(means, scales, quats, opacity, intensity)
        image = batch.forward(m)
        ctx.batch = batch
        ctx.mapped = m
        return image

    @staticmethod
    def backward(ctx, d_image):
        batch: GPBatch = ctx.batch
        if d_image is None:
            d_image = torch.zeros(batch.layout.n_pixels, dtype=torch.float32, device=batch.device)
        d_means, d_scales, d_quats, d_opac, d_inten = batch.backward(ctx.mapped, d_image.contiguous().to(torch.float32))
        return None, d_means, d_scales, d_quats, d_opac, d_inten


def _f32(t: torch.Tensor, device) -> torch.Tensor:
    return t.to(device=device, dtype=torch.float32).contiguous()


def render_frames(means, scales, quats, opacity, frames: Sequence[FrameSpec], cfg: GaussianPlieRenderConfig,
                  intensity: Optional[torch.Tensor] = None, sweeps: Optional[Sequence[SweepSpec]] = None,
                  hint: Optional[CapacityHint] = None, return_batch: bool = False, batch: Optional[GPBatch] = None):
    """Batched us_focus render: every frame of `frames` (over one or more sweeps) in one map/forward/backward.
    Returns the concatenated images (sum of H*W over frames,), differentiable w.r.t. the Gaussian tensors.
    `batch`: a GPBatch built earlier for these same frames, sweeps and settings (one returned with
    return_batch=True, or constructed directly); reusing it skips rebuilding and uploading the frame tables."""
    dev = means.device
    if dev.type != "cuda":
        raise RuntimeError("us3d.gaussianpile renders on the GPU only; there is no CPU path.")
    G = means.shape[0]
    if sweeps is None:
        sweeps = [SweepSpec(0, G)]
    if batch is None:
        batch = GPBatch(frames, sweeps, G, cfg.settings(), dev, hint)
    elif batch.layout.n_gauss != G or batch.layout.n_frames != len(frames) or batch.device != dev:
        raise ValueError("render_frames: the given batch was built for other frames, Gaussians or device")
    if intensity is None:
        intensity = torch.ones(G, dtype=torch.float32, device=dev)
    image = _GPRenderFunction.apply(batch, _f32(means, dev), _f32(scales, dev), _f32(quats, dev),
                                    _f32(opacity.reshape(-1), dev), _f32(intensity.reshape(-1), dev))
    return (image, batch) if return_batch else image


def validate_render_inputs(gauss_means, gauss_scales, gauss_quats, gauss_opacities, pose_probe_to_world) -> None:
    """check.py:10-29: the shape checks the facade runs when config.enable_input_checks is set."""
    n = gauss_means.shape[0]
    if gauss_means.ndim != 2 or tuple(gauss_means.shape) != (n, 3):
        raise ValueError(f"gauss_means must be (N, 3), got {tuple(gauss_means.shape)}.")
    if tuple(gauss_scales.shape) != (n, 3):
        raise ValueError(f"gauss_scales must be (N, 3), got {tuple(gauss_scales.shape)}.")
    if tuple(gauss_quats.shape) != (n, 4):
        raise ValueError(f"gauss_quats must be (N, 4), got {tuple(gauss_quats.shape)}.")
    if tuple(gauss_opacities.shape) not in ((n,), (n, 1)):
        raise ValueError(f"gauss_opacities must be (N,) or (N, 1), got {tuple(gauss_opacities.shape)}.")
    if tuple(pose_probe_to_world.shape) != (4, 4):
        raise ValueError(f"pose_probe_to_world must be (4, 4), got {tuple(pose_probe_to_world.shape)}.")


class GaussianPlieRenderer(nn.Module):
    """api.py:10-154: GaussianPile facade taking (mean, scale, quat, alpha) and a probe-to-world pose."""

    def __init__(self, config: GaussianPlieRenderConfig):
        super().__init__()
        self.config = config
        self._slice_grid_geometry = None
        self._hint = CapacityHint()

    @staticmethod
    def _quat_wxyz_to_matrix(quats: torch.Tensor) -> torch.Tensor:
        """api.py:53-70."""
        q = F.normalize(quats, dim=-1, eps=1e-8)
        w, x, y, z = q.unbind(dim=-1)
        rot = torch.stack([
            torch.stack([w * w + x * x - y * y - z * z, 2.0 * (x * y - w * z), 2.0 * (x * z + w * y)], -1),
            torch.stack([2.0 * (x * y + w * z), w * w - x * x + y * y - z * z, 2.0 * (y * z - w * x)], -1),
            torch.stack([2.0 * (x * z - w * y), 2.0 * (y * z + w * x), w * w - x * x - y * y + z * z], -1)], -2)
        return rot

    def _precision_tril_from_scales_quats(self, scales: torch.Tensor, quats: torch.Tensor) -> torch.Tensor:
        """api.py:72-94: raw precision factors (diagonal sqrt(chol_ii - beta)) for the UltraGauss backend."""
        scales = scales.clamp_min(float(self.config.cov_eps))
        rot = self._quat_wxyz_to_matrix(quats)
        precision = rot @ torch.diag_embed(1.0 / (scales * scales)) @ rot.transpose(1, 2)
        chol = torch.linalg.cholesky(precision)
        beta = float(self.config.cov_eps)
        diag_raw = (torch.diagonal(chol, dim1=1, dim2=2) - beta).clamp_min(1e-8).sqrt()
        return torch.stack([diag_raw[:, 0], chol[:, 1, 0], chol[:, 2, 0], diag_raw[:, 1], chol[:, 2, 1],
                            diag_raw[:, 2]], dim=1)

    def forward(self, gauss_means, gauss_scales, gauss_quats, gauss_alphas, pose_probe_to_world,
                pixel_positions_world=None, origin_world=None, axis_lateral_world=None, axis_axial_world=None,
                axis_elevational_world=None) -> torch.Tensor:
        cfg = self.config
        if cfg.enable_input_checks:
            validate_render_inputs(gauss_means, gauss_scales, gauss_quats, gauss_alphas,
                                   torch.as_tensor(pose_probe_to_world))
        if cfg.backend == "ultragauss_lowmem":
            if (pixel_positions_world is None or origin_world is None or axis_lateral_world is None
                    or axis_axial_world is None or axis_elevational_world is None):
                raise RuntimeError("ultragauss_lowmem backend requires pixel positions and slice geometry axes.")
            from .ultragauss import SliceGeometry, UltraGaussRenderConfig, UltraGaussRenderer

            ucfg = UltraGaussRenderConfig(
                H=cfg.H, W=cfg.W, lateral_spacing_mm_per_pix=cfg.lateral_spacing_mm_per_pix,
                axial_spacing_mm_per_pix=cfg.axial_spacing_mm_per_pix, axial_near_mm=cfg.axial_near_mm,
                lateral_col_offset=float(cfg.lateral_col_offset), chi2_threshold=float(cfg.radius_sigma) ** 2,
                mapper_sort_by_depth=False, fixed_bg_alpha=1.0e-6, fixed_bg_intensity=0.0,
                diag_beta=float(cfg.cov_eps), enable_shadow=False, shadow_delta=0.0)
            tril = self._precision_tril_from_scales_quats(gauss_scales, gauss_quats)
            pred, _alpha, _depth, _extras = UltraGaussRenderer(ucfg)(
                gauss_means_world=gauss_means, gauss_precision_tril=tril, gauss_opacities=gauss_alphas,
                gauss_intensities=torch.ones_like(gauss_alphas), pixel_positions_world=pixel_positions_world,
                geometry=SliceGeometry(origin_world=origin_world, axis_lateral_world=axis_lateral_world,
                                       axis_axial_world=axis_axial_world,
                                       axis_elevational_world=axis_elevational_world),
                return_intermediates=False)
            return pred
        if torch.is_tensor(pose_probe_to_world) and pose_probe_to_world.requires_grad:
            # the reference's backward can return d_pose (autograd.py:173-211); this build does not compute it
            raise NotImplementedError("GaussianPlieRenderer: gradients w.r.t. the probe pose are not computed; "
                                      "pass a pose that does not require grad")
        fr = frame_from_pose(pose_probe_to_world, cfg)
        image = render_frames(gauss_means, gauss_scales, gauss_quats, gauss_alphas, [fr], cfg, hint=self._hint)
        return image.view(cfg.H, cfg.W)

    def set_slice_grid_geometry(self, slice_grid_geometry: torch.Tensor) -> None:
        """api.py:150-154. The linear-probe grid of the config is used; the pack is checked for consistency
        where it carries the reference layout (probe type 0, H, W, spacings, near, column offset)."""
        g = torch.as_tensor(slice_grid_geometry).detach().float().cpu().reshape(-1)
        if g.numel() >= 7:
            cfg = self.config
            want = [0.0, cfg.H, cfg.W, cfg.lateral_spacing_mm_per_pix, cfg.axial_spacing_mm_per_pix,
                    cfg.axial_near_mm, cfg.lateral_col_offset]
            if any(abs(float(a) - float(b)) > 1e-4 * max(1.0, abs(float(b))) for a, b in zip(g[:7].tolist(), want)):
                raise ValueError("slice_grid_geometry differs from the config's linear-probe grid")
        self._slice_grid_geometry = slice_grid_geometry
