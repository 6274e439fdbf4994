"""UltraGauss slice splat on MI355X: renderer module, autograd function and batch driver.

Mirrors the reference package ops/ultragauss_renderer/ultragauss_renderer/ (types.py:6-40, api.py:10-55,
autograd.py:26-158): same config fields and validation, same forward signature, same outputs
(image, alpha, depth and the per-pixel candidate lists). The compute runs in libus3d (HIP, gfx950); there is
no CPU path.

Semantics switch (`semantics`):
  "cuda"     -- the reference CUDA renderer: pairs with q > chi2 or alpha_hat <= 1e-8 are dropped,
                alpha = sum alpha_hat + bg (forward.cu:17-76);
  "fallback" -- the reference PyTorch fallback (pipeline_god_file.py:556-603): no alpha_hat cut, the
                exponent clamp at -60 becomes chi2 = 120 (dropped pairs weigh <= o*e^-60), alpha = sum w.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import torch
import torch.nn as nn

from . import _lib
from .batch import CapacityHint, DeviceTables, FrameSpec, Mapped, SweepSpec, new_workspace, plan_layout, read_stats

FALLBACK_CHI2 = 120.0  # exp(clamp(-q/2, min=-60)) saturates at q = 120


@dataclass
class SliceGeometry:
    """ops/ultragauss_renderer/ultragauss_renderer/types.py:6-11"""

    origin_world: torch.Tensor
    axis_lateral_world: torch.Tensor
    axis_axial_world: torch.Tensor
    axis_elevational_world: torch.Tensor


@dataclass
class UltraGaussRenderConfig:
    """ops/ultragauss_renderer/ultragauss_renderer/types.py:14-40 (+ tile_size, the LDS batch sizes, the work queue
    and semantics). fwd_lds_batch / bwd_lds_batch: Gaussians the forward splat / the adjoint stage in LDS per batch,
    64, 128 or 256, or 0 for the kernel's thread count; the image does not depend on them and the gradients only
    through FP32 rounding (the autotuner picks them with the tile size). work_queue: the forward and the adjoint run
    as persistent workgroups taking (frame, tile) items from per-XCD queues instead of one workgroup per tile (same
    outputs)."""

    H: int
    W: int
    lateral_spacing_mm_per_pix: float
    axial_spacing_mm_per_pix: float
    axial_near_mm: float
    chi2_threshold: float = 7.815
    mapper_sort_by_depth: bool = True
    fixed_bg_alpha: float = 0.0
    fixed_bg_intensity: float = 0.0
    diag_beta: float = 1e-4
    enable_shadow: bool = False
    shadow_delta: float = 0.0
    lateral_col_offset: Optional[float] = None
    tile_size: int = 16
    semantics: str = "cuda"
    fwd_lds_batch: int = 0
    bwd_lds_batch: int = 0
    work_queue: bool = True

    def __post_init__(self) -> None:
        if self.H <= 0 or self.W <= 0:
            raise ValueError("H/W must be positive.")
        if self.lateral_spacing_mm_per_pix <= 0.0 or self.axial_spacing_mm_per_pix <= 0.0:
            raise ValueError("pixel spacing must be > 0.")
        if self.chi2_threshold <= 0.0:
            raise ValueError("chi2_threshold must be > 0.")
        if self.diag_beta <= 0.0:
            raise ValueError("diag_beta must be > 0.")
        if self.tile_size not in (8, 16, 32):
            raise ValueError("tile_size must be 8, 16 or 32.")
        if self.semantics not in ("cuda", "fallback"):
            raise ValueError("semantics must be 'cuda' or 'fallback'.")
        if self.fwd_lds_batch not in (0, 64, 128, 256) or self.bwd_lds_batch not in (0, 64, 128, 256):
            raise ValueError("fwd_lds_batch and bwd_lds_batch must be 0, 64, 128 or 256.")
        if self.lateral_col_offset is None:
            self.lateral_col_offset = float(self.W - 1) * 0.5

    def settings(self) -> _lib.UGSettings:
        s = _lib.UGSettings()
        if self.semantics == "fallback":
            s.chi2_threshold = FALLBACK_CHI2
            s.alpha_min = -1.0
            s.bg_alpha = 1e-6
            s.bg_intensity = 0.0
            s.alpha_excludes_bg = 1
        else:
            s.chi2_threshold = float(self.chi2_threshold)
            s.alpha_min = 1e-8
            s.bg_alpha = float(self.fixed_bg_alpha)
            s.bg_intensity = float(self.fixed_bg_intensity)
            s.alpha_excludes_bg = 0
        s.diag_beta = float(self.diag_beta)
        s.sort_by_depth = 1 if self.mapper_sort_by_depth else 0
        s.tile_size = int(self.tile_size)
        s.fwd_lds_batch = int(self.fwd_lds_batch)
        s.bwd_lds_batch = int(self.bwd_lds_batch)
        s.work_queue = 1 if self.work_queue else 0
        return s


def _f32(t: torch.Tensor, device: torch.device) -> torch.Tensor:
    return t.to(device=device, dtype=torch.float32).contiguous()


class UGBatch:
    """Map/forward/backward of one batch of frames over the C ABI, with overflow-driven capacity growth."""

    def __init__(self, frames: Sequence[FrameSpec], sweeps: Sequence[SweepSpec], n_gauss: int,
                 settings: _lib.UGSettings, device: torch.device, hint: Optional[CapacityHint] = None):
        self.settings = settings
        self.device = device
        self.layout = plan_layout(frames, sweeps, n_gauss, settings.tile_size)
        self.tables = DeviceTables(frames, self.layout, device)
        self.hint = hint or CapacityHint()

    def workspace_bytes(self, k_cap: int, v_cap: int) -> int:
        lay = self.layout
        return int(_lib.lib().us3d_ug_workspace_bytes(lay.n_gauss, lay.n_pairs, lay.n_tiles, k_cap, v_cap))

    def layout_offsets(self, k_cap: int, v_cap: int) -> _lib.Layout:
        lay = self.layout
        out = _lib.Layout()
        _lib.check(_lib.lib().us3d_ug_workspace_layout(lay.n_gauss, lay.n_pairs, lay.n_tiles, k_cap, v_cap, out),
                   "us3d_ug_workspace_layout")
        return out

    def map(self, means, tril, opacity, intensity, k_capacity: Optional[int] = None, v_capacity: Optional[int] = None,
            check: bool = True) -> Mapped:
        L = _lib.lib()
        n_pairs = self.layout.n_pairs
        k_cap = int(k_capacity or self.hint.capacity(n_pairs))
        v_cap = int(v_capacity or self.hint.v_capacity(n_pairs))
        while True:
            ws = new_workspace(self.workspace_bytes(k_cap, v_cap), self.device)
            m = Mapped(self.tables, ws, k_cap, v_cap)
            plan = m.plan()
            _lib.check(L.us3d_ug_map(plan, _lib.ptr(means), _lib.ptr(tril), _lib.ptr(opacity), _lib.ptr(intensity),
                                     self.settings, _lib.stream_ptr(self.device)), "us3d_ug_map")
            if not check:
                return m
            st = read_stats(ws)
            m.stats = st
            self.hint.update(st.n_instances, st.n_visible)
            if not st.overflow:
                return m
            k_cap, v_cap = self.hint.regrow(k_cap, v_cap, st, n_pairs)

    def forward(self, m: Mapped, pixels: torch.Tensor):
        n = self.layout.n_pixels
        image = torch.empty(n, dtype=torch.float32, device=self.device)
        alpha = torch.empty_like(image)
        depth = torch.empty_like(image)
        aux = torch.empty((n, 4), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().us3d_ug_forward(m.plan(), _lib.ptr(pixels), self.settings, _lib.ptr(image),
                                              _lib.ptr(alpha), _lib.ptr(depth), _lib.ptr(aux),
                                              _lib.stream_ptr(self.device)), "us3d_ug_forward")
        return image, alpha, depth, aux

    def backward(self, m: Mapped, pixels, tril, aux, d_image, d_alpha=None, d_depth=None):
        G = self.layout.n_gauss
        dev = self.device
        d_means = torch.empty((G, 3), dtype=torch.float32, device=dev)
        d_tril = torch.empty((G, 6), dtype=torch.float32, device=dev)
        d_opac = torch.empty(G, dtype=torch.float32, device=dev)
        d_inten = torch.empty(G, dtype=torch.float32, device=dev)
        _lib.check(_lib.lib().us3d_ug_backward(
            m.plan(), _lib.ptr(pixels), _lib.ptr(tril), self.settings, _lib.ptr(aux), _lib.ptr(d_image),
            _lib.ptr(d_alpha), _lib.ptr(d_depth), _lib.ptr(d_means), _lib.ptr(d_tril), _lib.ptr(d_opac),
            _lib.ptr(d_inten), _lib.stream_ptr(dev)), "us3d_ug_backward")
        return d_means, d_tril, d_opac, d_inten

    def pixel_lists(self, m: Mapped, pixels: torch.Tensor):
        """Per-pixel candidate CSR (counts, offsets, indices) of the reference mapper (debug extras)."""
        L = _lib.lib()
        n = self.layout.n_pixels
        dev = self.device
        counts = torch.empty(n, dtype=torch.int32, device=dev)
        offsets = torch.empty(n, dtype=torch.int32, device=dev)
        s = _lib.stream_ptr(dev)
        _lib.check(L.us3d_ug_pixel_lists(m.plan(), _lib.ptr(pixels), self.settings, 0, _lib.ptr(counts), None, None, s),
                   "us3d_ug_pixel_lists(count)")
        tmp = torch.empty(int(L.us3d_scan_tmp_bytes(n)), dtype=torch.uint8, device=dev)
        total = torch.zeros(1, dtype=torch.int64, device=dev)
        _lib.check(L.us3d_scan_i32(_lib.ptr(counts), _lib.ptr(offsets), n, _lib.ptr(total), _lib.ptr(tmp), s),
                   "us3d_scan_i32")
        nnz = int(total.item())
        indices = torch.empty(max(nnz, 0), dtype=torch.int32, device=dev)
        if nnz > 0:
            _lib.check(L.us3d_ug_pixel_lists(m.plan(), _lib.ptr(pixels), self.settings, 1, None, _lib.ptr(offsets),
                                             _lib.ptr(indices), s), "us3d_ug_pixel_lists(fill)")
        return counts, offsets, indices

    def sorted_view(self, m: Mapped) -> Dict[str, torch.Tensor]:
        """Reference-format mapper outputs per frame: keys (tile<<32 | depth key), values, tile ranges."""
        off = self.layout_offsets(m.k_capacity, m.v_capacity)
        ws = m.workspace
        K = int(read_stats(ws).n_instances)
        lay = self.layout
        V = int(read_stats(ws).n_visible)
        seg_key = ws[off.seg_key: off.seg_key + 8 * K].view(torch.int64)  # depth key << 32 | visible pair v
        pair_g = ws[off.pair_g: off.pair_g + 4 * V].view(torch.int32).to(torch.int64)
        tile_off = ws[off.tile_off: off.tile_off + 4 * (lay.n_tiles + 1)].view(torch.int32).to(torch.int64)
        out = {}
        sw_rows = lay.sweep_rows
        for f in range(lay.n_frames):
            t0, t1 = lay.frame_tile_base[f], lay.frame_tile_base[f + 1]
            s0, s1 = int(tile_off[t0]), int(tile_off[t1])
            counts = (tile_off[t0 + 1: t1 + 1] - tile_off[t0:t1])
            tiles = torch.repeat_interleave(torch.arange(t1 - t0, device=ws.device, dtype=torch.int64), counts)
            kk = seg_key[s0:s1]
            dk = (kk >> 32) & 0xFFFFFFFF
            g = pair_g[kk & 0xFFFFFFFF]
            sweep = int(self.tables.frames[f, 22].item())
            g_begin = int(sw_rows[sweep, 0])
            start = torch.where(counts > 0, tile_off[t0:t1] - s0, torch.full_like(counts, -1))
            end = torch.where(counts > 0, tile_off[t0 + 1: t1 + 1] - s0, torch.full_like(counts, -1))
            out[f] = {"keys": (tiles << 32) | dk, "values": (g - g_begin).to(torch.int32),
                      "tile_start": start.to(torch.int32), "tile_end": end.to(torch.int32)}
        return out


class _RenderFunction(torch.autograd.Function):
    """ops/ultragauss_renderer/ultragauss_renderer/autograd.py:26-158: forward renders, backward runs the
    recompute-based adjoint kernel. Gradients w.r.t. means, precision_tril, opacities and intensities."""

    @staticmethod
    def forward(ctx, batch: UGBatch, means, tril, opacity, intensity, pixels):
        m = batch.map(means, tril, opacity, intensity)
        image, alpha, depth, aux = batch.forward(m, pixels)
        ctx.batch = batch
        ctx.mapped = m
        ctx.save_for_backward(tril, pixels, aux)
        ctx.mark_non_differentiable(aux)
        return image, alpha, depth, aux

    @staticmethod
    def backward(ctx, d_image, d_alpha, d_depth, d_aux):
        del d_aux
        tril, pixels, aux = ctx.saved_tensors
        batch: UGBatch = ctx.batch
        z = lambda t: None if t is None else t.contiguous().to(torch.float32)  # noqa: E731
        if d_image is None:
            d_image = torch.zeros(batch.layout.n_pixels, dtype=torch.float32, device=batch.device)
        d_means, d_tril, d_opac, d_inten = batch.backward(ctx.mapped, pixels, tril, aux, z(d_image), z(d_alpha),
                                                          z(d_depth))
        return None, d_means, d_tril, d_opac, d_inten, None


def render_frames(means: torch.Tensor, tril: torch.Tensor, opacity: torch.Tensor, intensity: torch.Tensor,
                  pixels: torch.Tensor, frames: Sequence[FrameSpec], config: UltraGaussRenderConfig,
                  sweeps: Optional[Sequence[SweepSpec]] = None, hint: Optional[CapacityHint] = None,
                  return_batch: bool = False):
    """Differentiable render of several frames in one batch (one launch sequence for all of them).

    means (G,3), tril (G,6), opacity (G,), intensity (G,), pixels (sum H*W, 3) concatenated per frame, or
    None for the linear-probe grid of each frame computed in-kernel. Returns flat (image, alpha, depth) of
    length sum H*W."""
    dev = means.device
    if dev.type != "cuda":
        raise RuntimeError("us3d renders on the GPU only (no CPU fallback); move the tensors to cuda")
    G = means.shape[0]
    if sweeps is None:
        sweeps = [SweepSpec(0, G)]
    batch = UGBatch(frames, sweeps, G, config.settings(), dev, hint)
    args = [_f32(t, dev) for t in (means.reshape(G, 3), tril.reshape(G, 6), opacity.reshape(G), intensity.reshape(G))]
    pix = None
    if pixels is not None:
        pix = _f32(pixels.reshape(-1, 3), dev)
        if pix.shape[0] != batch.layout.n_pixels:
            raise ValueError(f"pixels has {pix.shape[0]} rows, the frames have {batch.layout.n_pixels} pixels")
    image, alpha, depth, _aux = _RenderFunction.apply(batch, *args, pix)
    if return_batch:
        return image, alpha, depth, batch
    return image, alpha, depth


class UltraGaussRenderer(nn.Module):
    """ops/ultragauss_renderer/ultragauss_renderer/api.py:10-55 (same forward signature and outputs)."""

    def __init__(self, config: UltraGaussRenderConfig, hint: Optional[CapacityHint] = None):
        super().__init__()
        self.config = config
        self.hint = hint or CapacityHint()

    def forward(self, gauss_means_world, gauss_precision_tril, gauss_opacities, gauss_intensities,
                pixel_positions_world, geometry: SliceGeometry, return_intermediates: bool = False,
                ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, Dict[str, torch.Tensor]]:
        cfg = self.config
        dev = gauss_means_world.device
        frame = FrameSpec(geometry.origin_world, geometry.axis_lateral_world, geometry.axis_axial_world,
                          geometry.axis_elevational_world, int(cfg.H), int(cfg.W),
                          float(cfg.lateral_spacing_mm_per_pix), float(cfg.axial_spacing_mm_per_pix),
                          float(cfg.axial_near_mm), float(cfg.lateral_col_offset))
        if pixel_positions_world.reshape(-1, 3).shape[0] != cfg.H * cfg.W:
            raise RuntimeError("n_pixels must equal H*W")
        image, alpha, depth, batch = render_frames(gauss_means_world, gauss_precision_tril, gauss_opacities,
                                                   gauss_intensities, pixel_positions_world, [frame], cfg,
                                                   hint=self.hint, return_batch=True)
        if cfg.enable_shadow:
            # api.py:37-42: T = exp(-cumsum(alpha * delta) down the rows), image *= T
            delta = cfg.shadow_delta if cfg.shadow_delta > 0.0 else 1.0 / float(cfg.H)
            trans = torch.exp(-torch.cumsum(alpha.view(cfg.H, cfg.W) * delta, dim=0))
            image = (image.view(cfg.H, cfg.W) * trans).reshape(-1)
        inter: Dict[str, torch.Tensor] = {}
        if return_intermediates:
            # the mapped workspace of this call is held by the autograd graph; rebuild the lists from a
            # fresh (no-grad) mapping so the extras do not depend on autograd internals
            with torch.no_grad():
                G = gauss_means_world.shape[0]
                m = batch.map(*[_f32(t, dev) for t in (gauss_means_world.reshape(G, 3), gauss_precision_tril.reshape(G, 6),
                                                       gauss_opacities.reshape(G), gauss_intensities.reshape(G))])
                counts, offsets, indices = batch.pixel_lists(m, _f32(pixel_positions_world.reshape(-1, 3), dev))
            inter = {"pixel_gauss_counts": counts, "pixel_gauss_offsets": offsets, "pixel_gauss_indices": indices}
        return image.view(cfg.H, cfg.W), alpha.view(cfg.H, cfg.W), depth.view(cfg.H, cfg.W), inter
