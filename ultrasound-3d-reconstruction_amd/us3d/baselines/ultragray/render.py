"""Drop-in for baselines/ultragray/render.py of the reference (same dataclasses, same function).

`render_ultragray(gaussians, geometry, settings) -> RenderOutput(image (H, W), depth=None, extras={})`,
rendered by the MI355X kernels of libus3d. Field names, order, defaults and meaning follow
baselines/ultragray/render.py:15-119.
"""

from dataclasses import dataclass, field
from typing import Dict, Optional

import torch

from ...batch import CapacityHint
from ...ultragray import ProbeGeometry, UltragrayRenderConfig, UltragrayRenderer


@dataclass
class RaySliceInput:
    pixel_positions: torch.Tensor  # (W, H, 3) mm
    ray_origins: torch.Tensor  # (W, 3)
    ray_dirs: torch.Tensor  # (W, 3) unit
    probe_origin_world: torch.Tensor  # (3,)
    axis_lateral_world: torch.Tensor  # (3,)
    axis_axial_world: torch.Tensor  # (3,)
    axis_elevational_world: torch.Tensor  # (3,)


@dataclass
class UltraGrayInput:
    means: torch.Tensor  # (G, 3) mm
    scales: torch.Tensor  # (G, 3) mm
    quats: torch.Tensor  # (G, 4) w, x, y, z
    attenuations: torch.Tensor  # (G,) raw logits, tau = sigmoid
    sh_coeffs: torch.Tensor  # (G, 4): c0 + c1 dx + c2 dy + c3 dz


@dataclass
class UltraGrayRenderSettings:
    image_width: int
    image_height: int
    width_mm: float
    near_mm: float
    lateral_col_offset: float = 0.0
    elev_cutoff_margin_mm: float = 0.0
    k_sigma: float = 3.0
    ops_max_offset_mm: float = 0.0
    quadrature_window_radius_mm: float = 2.0
    local_transmission_window_only: bool = False
    local_transmission_max_window_mm: float = 17.0


@dataclass
class RenderOutput:
    image: torch.Tensor
    depth: Optional[torch.Tensor] = None
    extras: Dict[str, torch.Tensor] = field(default_factory=dict)


_HINTS: Dict[tuple, CapacityHint] = {}


def render_ultragray(gaussians: UltraGrayInput, geometry: RaySliceInput,
                     settings: UltraGrayRenderSettings) -> RenderOutput:
    """Render one B-mode frame (baselines/ultragray/render.py:81-119)."""
    cfg = UltragrayRenderConfig(
        image_width=int(settings.image_width),
        image_height=int(settings.image_height),
        width_mm=float(settings.width_mm),
        near_mm=float(settings.near_mm),
        lateral_col_offset=float(settings.lateral_col_offset),
        elev_cutoff_margin_mm=float(settings.elev_cutoff_margin_mm),
        k_sigma=float(settings.k_sigma),
        ops_max_offset_mm=float(settings.ops_max_offset_mm),
        quadrature_window_radius_mm=float(settings.quadrature_window_radius_mm),
        local_transmission_window_only=bool(settings.local_transmission_window_only),
        local_transmission_max_window_mm=float(settings.local_transmission_max_window_mm),
    )
    key = (cfg.image_width, cfg.image_height, int(gaussians.means.shape[0]))
    renderer = UltragrayRenderer(cfg, hint=_HINTS.setdefault(key, CapacityHint()))
    image = renderer(
        means=gaussians.means,
        scales=gaussians.scales,
        quats=gaussians.quats,
        attenuations=gaussians.attenuations,
        sh_coeffs=gaussians.sh_coeffs,
        pixel_positions=geometry.pixel_positions,
        ray_origins=geometry.ray_origins,
        ray_dirs=geometry.ray_dirs,
        geometry=ProbeGeometry(
            origin_world=geometry.probe_origin_world,
            axis_lateral_world=geometry.axis_lateral_world,
            axis_elevational_world=geometry.axis_elevational_world,
            axis_axial_world=geometry.axis_axial_world,
        ),
    )
    return RenderOutput(image=image, depth=None, extras={})
