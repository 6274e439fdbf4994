"""UltraG-Ray B-mode ray renderer on MI355X: renderer module, autograd function and batch driver.

Mirrors ops/ultragray_renderer/ultragray_renderer/ (types.py:6-43, api.py:8-39, autograd.py:89-287): same
config fields and validation, same forward signature (pixel_positions (W, H, 3), rays (W, 3)) and output
image (H, W); gradients w.r.t. means, scales, quats, attenuations (raw logits) and sh_coeffs. The OPS
elevational jitter draws torch.rand(W) on every call, as the reference does (autograd.py:128-131), so the
RNG stream advances identically. Optional per-stage timing with ULTRAGRAY_PROFILE_EVERY /
ULTRAGRAY_PROFILE_WARMUP (autograd.py:13-70).
"""

from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch
import torch.nn as nn

from . import _lib
from .batch import CapacityHint, DeviceTables, FrameSpec, Mapped, SweepSpec, new_workspace, plan_layout, read_stats


@dataclass
class ProbeGeometry:
    """ops/ultragray_renderer/ultragray_renderer/types.py:6-11"""

    origin_world: torch.Tensor
    axis_lateral_world: torch.Tensor
    axis_elevational_world: torch.Tensor
    axis_axial_world: torch.Tensor


@dataclass
class UltragrayRenderConfig:
    """ops/ultragray_renderer/ultragray_renderer/types.py:14-43 (+ grad_threshold)."""

    image_width: int
    image_height: int
    width_mm: float
    near_mm: float
    elev_cutoff_margin_mm: float = 0.0
    k_sigma: float = 3.0
    ops_max_offset_mm: float = 2.0
    quadrature_window_radius_mm: float = 2.0
    local_transmission_window_only: bool = False
    local_transmission_max_window_mm: float = 17.0
    elev_beam_width_mm: float = 2.0
    lateral_col_offset: Optional[float] = None
    grad_threshold: float = 1e-6  # the reference skips samples with |dL/dB| < 1e-6 (backward.cu:170)

    def __post_init__(self) -> None:
        if self.image_width <= 0 or self.image_height <= 0:
            raise ValueError("image_width/image_height must be positive.")
        if self.width_mm <= 0.0:
            raise ValueError("width_mm must be > 0.")
        if self.k_sigma <= 0.0:
            raise ValueError("k_sigma must be > 0.")
        if self.ops_max_offset_mm < 0.0:
            raise ValueError("ops_max_offset_mm must be >= 0.")
        if self.quadrature_window_radius_mm < 0.0:
            raise ValueError("quadrature_window_radius_mm must be >= 0.")
        if self.local_transmission_max_window_mm <= 0.0:
            raise ValueError("local_transmission_max_window_mm must be > 0.")
        if self.lateral_col_offset is None:
            self.lateral_col_offset = float(self.image_width - 1) * 0.5

    def settings(self) -> _lib.URSettings:
        s = _lib.URSettings()
        s.k_sigma = float(self.k_sigma)
        s.quadrature_window_radius_mm = float(self.quadrature_window_radius_mm)
        s.elev_cutoff_margin_mm = float(self.elev_cutoff_margin_mm)
        s.local_transmission_max_window_mm = float(self.local_transmission_max_window_mm)
        s.grad_threshold = float(self.grad_threshold)
        s.local_transmission_window_only = 1 if self.local_transmission_window_only else 0
        return s

    def pixel_spacing(self) -> float:
        """width_mm / W evaluated in FP32, as the reference mapper does (wrapper_mapper.cu:21)."""
        return float(np.float32(self.width_mm) / np.float32(self.image_width))


def ray_frame(geometry: ProbeGeometry, cfg: UltragrayRenderConfig, sweep: int = 0) -> FrameSpec:
    return FrameSpec(geometry.origin_world, geometry.axis_lateral_world, geometry.axis_axial_world,
                     geometry.axis_elevational_world, int(cfg.image_height), int(cfg.image_width),
                     cfg.pixel_spacing(), 0.0, float(cfg.near_mm), float(cfg.lateral_col_offset), sweep)


def ops_offsets(n_rays: int, max_offset_mm: float, device) -> torch.Tensor:
    """OPS elevational jitter per ray, asin(2u-1)/(pi/2) * max_offset (autograd.py:128-131)."""
    u = torch.rand(n_rays, device=device, dtype=torch.float32)
    return (torch.asin(2.0 * u - 1.0) / (math.pi / 2.0) * max_offset_mm).contiguous()


_PROFILE_STEP = 0


def _profile_every() -> int:
    try:
        return int(os.environ.get("ULTRAGRAY_PROFILE_EVERY", "0"))
    except ValueError:
        return 0


def _profile_warmup() -> int:
    try:
        return int(os.environ.get("ULTRAGRAY_PROFILE_WARMUP", "0"))
    except ValueError:
        return 0


def _timed(fn, enabled: bool):
    if not enabled:
        return fn(), None
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    start.record()
    out = fn()
    end.record()
    torch.cuda.synchronize()
    return out, ((time.perf_counter() - w0) * 1e3, start.elapsed_time(end))


class URBatch:
    """Map/forward/backward of a batch of UltraG-Ray frames over the C ABI."""

    def __init__(self, frames: Sequence[FrameSpec], sweeps: Sequence[SweepSpec], n_gauss: int,
                 settings: _lib.URSettings, device: torch.device, hint: Optional[CapacityHint] = None):
        if any(int(f.geometry) != 0 for f in frames):
            raise ValueError("the UltraG-Ray renderer takes linear-probe frames only (geometry 0)")
        self.settings = settings
        self.device = device
        self.layout = plan_layout(frames, sweeps, n_gauss, None)
        self.tables = DeviceTables(frames, self.layout, device)
        self.hint = hint or CapacityHint()

    def workspace_bytes(self, k_cap: int, v_cap: int) -> int:
        lay = self.layout
        return int(_lib.lib().us3d_ur_workspace_bytes(lay.n_gauss, lay.n_pairs, lay.n_tiles, k_cap, v_cap))

    def map(self, means, scales, quats, atten, sh) -> Mapped:
        L = _lib.lib()
        n_pairs = self.layout.n_pairs
        k_cap = self.hint.capacity(n_pairs)
        v_cap = self.hint.v_capacity(n_pairs)
        while True:
            ws = new_workspace(self.workspace_bytes(k_cap, v_cap), self.device)
            m = Mapped(self.tables, ws, k_cap, v_cap)
            _lib.check(L.us3d_ur_map(m.plan(), _lib.ptr(means), _lib.ptr(scales), _lib.ptr(quats), _lib.ptr(atten),
                                     _lib.ptr(sh), self.settings, _lib.stream_ptr(self.device)), "us3d_ur_map")
            st = read_stats(ws)
            m.stats = st
            self.hint.update(st.n_instances, st.n_visible)
            if not st.overflow:
                return m
            k_cap, v_cap = self.hint.regrow(k_cap, v_cap, st, n_pairs)

    def forward(self, m: Mapped, pixels, ray_o, ray_d, ray_off):
        n = self.layout.n_pixels
        image = torch.empty(n, dtype=torch.float32, device=self.device)
        aux = torch.empty((n, 4), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().us3d_ur_forward(m.plan(), _lib.ptr(pixels), _lib.ptr(ray_o), _lib.ptr(ray_d),
                                              _lib.ptr(ray_off), self.settings, _lib.ptr(image), _lib.ptr(aux),
                                              _lib.stream_ptr(self.device)), "us3d_ur_forward")
        return image, aux

    def backward(self, m: Mapped, pixels, ray_o, ray_d, ray_off, aux, d_image):
        G = self.layout.n_gauss
        dev = self.device
        outs = [torch.empty(s, dtype=torch.float32, device=dev) for s in ((G, 3), (G, 3), (G, 4), (G,), (G, 4))]
        _lib.check(_lib.lib().us3d_ur_backward(
            m.plan(), _lib.ptr(pixels), _lib.ptr(ray_o), _lib.ptr(ray_d), _lib.ptr(ray_off), self.settings,
            _lib.ptr(aux), _lib.ptr(d_image), *[_lib.ptr(o) for o in outs], _lib.stream_ptr(dev)), "us3d_ur_backward")
        return outs

    def ray_lists(self, m: Mapped):
        """Reference-format mapper outputs per frame: ray counts/offsets, sorted Gaussian indices and depths."""
        off = _lib.Layout()
        lay = self.layout
        _lib.check(_lib.lib().us3d_ur_workspace_layout(lay.n_gauss, lay.n_pairs, lay.n_tiles, m.k_capacity,
                                                       m.v_capacity, off),
                   "us3d_ur_workspace_layout")
        ws = m.workspace
        K = int(read_stats(ws).n_instances)
        V = int(read_stats(ws).n_visible)
        seg_key = ws[off.seg_key: off.seg_key + 8 * K].view(torch.int64)  # depth key << 32 | visible pair v
        pair_g = ws[off.pair_g: off.pair_g + 4 * V].view(torch.int32).to(torch.int64)
        tile_off = ws[off.tile_off: off.tile_off + 4 * (lay.n_tiles + 1)].view(torch.int32).to(torch.int64)
        out = []
        for f in range(lay.n_frames):
            t0, t1 = lay.frame_tile_base[f], lay.frame_tile_base[f + 1]
            s0, s1 = int(tile_off[t0]), int(tile_off[t1])
            counts = tile_off[t0 + 1: t1 + 1] - tile_off[t0:t1]
            offsets = torch.where(counts > 0, tile_off[t0:t1] - s0, torch.zeros_like(counts))
            kk = seg_key[s0:s1]
            g_begin = int(lay.sweep_rows[int(self.tables.frames[f, 22].item()), 0])
            out.append({"counts": counts.to(torch.int32), "offsets": offsets.to(torch.int32),
                        "indices": (pair_g[kk & 0xFFFFFFFF] - g_begin).to(torch.int32),
                        "sorted_depths": ((kk >> 32) & 0xFFFFFFFF).to(torch.float32) * 0.001})
        return out


class _UltragrayRenderFunction(torch.autograd.Function):
    """ops/ultragray_renderer/ultragray_renderer/autograd.py:89-287."""

    @staticmethod
    def forward(ctx, batch: URBatch, means, scales, quats, atten, sh, pixels, ray_o, ray_d, ray_off):
        global _PROFILE_STEP
        _PROFILE_STEP += 1
        every = _profile_every()
        prof = every > 0 and _PROFILE_STEP > _profile_warmup() and _PROFILE_STEP % every == 0
        m, t_map = _timed(lambda: batch.map(means, scales, quats, atten, sh), prof)
        (image, aux), t_fwd = _timed(lambda: batch.forward(m, pixels, ray_o, ray_d, ray_off), prof)
        if prof:
            st = m.stats
            print(f"[ULTRAGRAY_PROFILE] step={_PROFILE_STEP} phase=forward mapper_wall_ms={t_map[0]:.3f} "
                  f"mapper_gpu_ms={t_map[1]:.3f} shader_fwd_wall_ms={t_fwd[0]:.3f} shader_fwd_gpu_ms={t_fwd[1]:.3f} "
                  f"total_hits={st.n_instances} max_hits_per_ray={st.max_segment}", flush=True)
        ctx.batch, ctx.mapped, ctx.prof, ctx.step = batch, m, prof, _PROFILE_STEP
        ctx.save_for_backward(pixels, ray_o, ray_d, ray_off, aux)
        return image

    @staticmethod
    def backward(ctx, d_image):
        pixels, ray_o, ray_d, ray_off, aux = ctx.saved_tensors
        batch: URBatch = ctx.batch
        grads, t_bwd = _timed(lambda: batch.backward(ctx.mapped, pixels, ray_o, ray_d, ray_off, aux,
                                                     d_image.contiguous().to(torch.float32)), ctx.prof)
        if ctx.prof:
            print(f"[ULTRAGRAY_PROFILE] step={ctx.step} phase=backward shader_bwd_wall_ms={t_bwd[0]:.3f} "
                  f"shader_bwd_gpu_ms={t_bwd[1]:.3f}", flush=True)
        d_means, d_scales, d_quats, d_atten, d_sh = grads
        return None, d_means, d_scales, d_quats, d_atten, d_sh, None, None, None, None


def _f32(t, dev):
    return t.to(device=dev, dtype=torch.float32).contiguous()


def render_ray_frames(means, scales, quats, attenuations, sh, pixels, ray_origins, ray_dirs, ray_offsets,
                      frames: Sequence[FrameSpec], cfg: UltragrayRenderConfig,
                      sweeps: Optional[Sequence[SweepSpec]] = None, hint: Optional[CapacityHint] = None,
                      return_batch: bool = False):
    """Differentiable render of several UltraG-Ray frames as one batch. pixels: per frame (W, H, 3)
    concatenated (flat); ray_origins/ray_dirs (sum W, 3); ray_offsets (sum W,). Returns the flat image
    (per frame (H, W) row-major, concatenated)."""
    dev = means.device
    if dev.type != "cuda":
        raise RuntimeError("us3d renders on the GPU only (no CPU fallback); move the tensors to cuda")
    G = means.shape[0]
    sweeps = sweeps or [SweepSpec(0, G)]
    batch = URBatch(frames, sweeps, G, cfg.settings(), dev, hint)
    args = [_f32(t, dev) for t in (means.reshape(G, 3), scales.reshape(G, 3), quats.reshape(G, 4),
                                   attenuations.reshape(G), sh.reshape(G, 4))]
    lay = batch.layout
    pix = _f32(pixels.reshape(-1, 3), dev)
    if pix.shape[0] != lay.n_pixels:
        raise ValueError(f"pixels has {pix.shape[0]} rows, the frames have {lay.n_pixels} samples")
    ro, rd = _f32(ray_origins.reshape(-1, 3), dev), _f32(ray_dirs.reshape(-1, 3), dev)
    off = _f32(ray_offsets.reshape(-1), dev)
    image = _UltragrayRenderFunction.apply(batch, *args, pix, ro, rd, off)
    return (image, batch) if return_batch else image


class UltragrayRenderer(nn.Module):
    """ops/ultragray_renderer/ultragray_renderer/api.py:8-39."""

    def __init__(self, config: UltragrayRenderConfig, hint: Optional[CapacityHint] = None):
        super().__init__()
        self.config = config
        self.hint = hint or CapacityHint()

    def forward(self, means, scales, quats, attenuations, sh_coeffs, pixel_positions, ray_origins, ray_dirs,
                geometry: ProbeGeometry) -> torch.Tensor:
        cfg = self.config
        W, H = int(cfg.image_width), int(cfg.image_height)
        if tuple(pixel_positions.shape[:2]) != (W, H):
            raise ValueError(f"pixel_positions must be (W, H, 3) = ({W}, {H}, 3)")
        off = ops_offsets(W, float(cfg.ops_max_offset_mm), means.device)
        image = render_ray_frames(means, scales, quats, attenuations, sh_coeffs, pixel_positions, ray_origins,
                                  ray_dirs, off, [ray_frame(geometry, cfg)], cfg, hint=self.hint)
        return image.view(H, W)
