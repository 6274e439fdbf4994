"""Adapter with the signatures of the reference's baselines/gaussianpile/render.py:16-150: the same dataclasses
(SliceGeometryInput, GaussianPileInput, GaussianPileRenderSettings, RenderOutput) and render_gaussianpile, on the
MI355X GaussianPile renderer (us3d.gaussianpile).

The wiring is kept as the reference's (render.py:97-134): the facade receives geometry.pose_world_to_probe in
its pose_probe_to_world argument, whose columns the renderer reads as the probe axes and translation
(forward.cu:145-151). The linear-probe grid comes from the settings; the reference builds a slice-grid pack
for it through ours_renderer.grid_geometry, a module its tree does not contain.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, Optional

import torch

from ...gaussianpile import GaussianPlieRenderConfig, GaussianPlieRenderer


@dataclass
class SliceGeometryInput:
    pixel_positions_world: torch.Tensor   # (H*W, 3) mm
    probe_origin_world: torch.Tensor      # (3,) mm
    axis_lateral_world: torch.Tensor      # (3,) unit
    axis_axial_world: torch.Tensor        # (3,) unit
    axis_elevational_world: torch.Tensor  # (3,) unit
    pose_world_to_probe: torch.Tensor     # (4, 4)


@dataclass
class GaussianPileInput:
    means: torch.Tensor   # (G, 3) mm
    scales: torch.Tensor  # (G, 3) mm
    quats: torch.Tensor   # (G, 4) (w, x, y, z)
    alphas: torch.Tensor  # (G,) or (G, 1)


@dataclass
class GaussianPileRenderSettings:
    image_height: int
    image_width: int
    lateral_spacing_mm_per_pix: float
    axial_spacing_mm_per_pix: float
    axial_near_mm: float
    lateral_col_offset: float = 0.0
    backend: str = "ours_focus"
    tile_size_px: int = 16
    radius_sigma: float = 3.0
    sigma_z_mm: float = 1.0
    prefilter_elevational_radius_mm: float = 0.0
    det_eps: float = 1.0e-6
    cov_eps: float = 1.0e-6
    opacity_cull_threshold: float = 1.0e-6
    render_chunk_size: int = 250_000
    enable_input_checks: bool = False


@dataclass
class RenderOutput:
    image: torch.Tensor
    depth: Optional[torch.Tensor] = None
    extras: Dict[str, torch.Tensor] = field(default_factory=dict)


def render_gaussianpile(gaussians: GaussianPileInput, geometry: SliceGeometryInput,
                        settings: GaussianPileRenderSettings) -> RenderOutput:
    """render.py:85-134: one frame; image (H, W), depth a zero tensor of the same shape."""
    cfg = GaussianPlieRenderConfig(
        H=int(settings.image_height), W=int(settings.image_width),
        lateral_spacing_mm_per_pix=float(settings.lateral_spacing_mm_per_pix),
        axial_spacing_mm_per_pix=float(settings.axial_spacing_mm_per_pix),
        axial_near_mm=float(settings.axial_near_mm), lateral_col_offset=float(settings.lateral_col_offset),
        backend=str(settings.backend), tile_size_px=int(settings.tile_size_px),
        radius_sigma=float(settings.radius_sigma), sigma_z_mm=float(settings.sigma_z_mm),
        prefilter_elevational_radius_mm=float(settings.prefilter_elevational_radius_mm),
        det_eps=float(settings.det_eps), cov_eps=float(settings.cov_eps),
        opacity_cull_threshold=float(settings.opacity_cull_threshold),
        render_chunk_size=int(settings.render_chunk_size), enable_input_checks=bool(settings.enable_input_checks))
    renderer = GaussianPlieRenderer(cfg)
    image = renderer(gaussians.means, gaussians.scales, gaussians.quats, gaussians.alphas.reshape(-1),
                     geometry.pose_world_to_probe, pixel_positions_world=geometry.pixel_positions_world,
                     origin_world=geometry.probe_origin_world, axis_lateral_world=geometry.axis_lateral_world,
                     axis_axial_world=geometry.axis_axial_world,
                     axis_elevational_world=geometry.axis_elevational_world)
    return RenderOutput(image=image, depth=torch.zeros_like(image), extras={})
