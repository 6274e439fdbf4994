"""Drop-in for baselines/ultragauss/render.py of the reference (same dataclasses, same function).

`render_ultragauss(gaussians, geometry, settings) -> RenderOutput(image (H,W), depth (H,W),
extras{alpha_image (H,W), pixel_gauss_counts, pixel_gauss_offsets, pixel_gauss_indices})`, rendered by the
MI355X kernels of libus3d. Field names, order, defaults and meaning follow
baselines/ultragauss/render.py:16-114; `return_pixel_lists` (default True, as the reference always returns
the lists) and `tile_size` / `semantics` are additions.
"""

from dataclasses import dataclass, field
from typing import Dict, Optional

import torch

from ...batch import CapacityHint
from ...ultragauss import SliceGeometry, UltraGaussRenderConfig, UltraGaussRenderer


@dataclass
class ProbeSliceInput:
    pixel_positions_world: torch.Tensor  # (H*W, 3) mm
    probe_origin_world: torch.Tensor  # (3,)
    axis_lateral_world: torch.Tensor  # (3,)
    axis_axial_world: torch.Tensor  # (3,)
    axis_elevational_world: torch.Tensor  # (3,)


@dataclass
class UltraGaussInput:
    means: torch.Tensor  # (G, 3) mm
    precision_tril: torch.Tensor  # (G, 6): l11, l12, l13, l22, l23, l33 raw factors
    opacities: torch.Tensor  # (G,)
    intensities: torch.Tensor  # (G,)


@dataclass
class UltraGaussRenderSettings:
    image_height: int
    image_width: int
    lateral_spacing_mm_per_pix: float
    axial_spacing_mm_per_pix: float
    axial_near_mm: float
    lateral_col_offset: float = 0.0
    chi2_threshold: float = 7.815
    mapper_sort_by_depth: bool = True
    fixed_bg_alpha: float = 1.0e-6
    fixed_bg_intensity: float = 0.0
    diag_beta: float = 1.0e-4
    enable_shadow: bool = False
    shadow_delta: float = 0.0
    return_pixel_lists: bool = True
    tile_size: int = 16
    semantics: str = "cuda"


@dataclass
class RenderOutput:
    image: torch.Tensor
    depth: Optional[torch.Tensor] = None
    extras: Dict[str, torch.Tensor] = field(default_factory=dict)


_HINTS: Dict[tuple, CapacityHint] = {}


def render_ultragauss(gaussians: UltraGaussInput, geometry: ProbeSliceInput,
                      settings: UltraGaussRenderSettings) -> RenderOutput:
    """Render one probe slice (baselines/ultragauss/render.py:75-114)."""
    cfg = UltraGaussRenderConfig(
        H=int(settings.image_height),
        W=int(settings.image_width),
        lateral_spacing_mm_per_pix=float(settings.lateral_spacing_mm_per_pix),
        axial_spacing_mm_per_pix=float(settings.axial_spacing_mm_per_pix),
        axial_near_mm=float(settings.axial_near_mm),
        lateral_col_offset=float(settings.lateral_col_offset),
        chi2_threshold=float(settings.chi2_threshold),
        mapper_sort_by_depth=bool(settings.mapper_sort_by_depth),
        fixed_bg_alpha=float(settings.fixed_bg_alpha),
        fixed_bg_intensity=float(settings.fixed_bg_intensity),
        diag_beta=float(settings.diag_beta),
        enable_shadow=bool(settings.enable_shadow),
        shadow_delta=float(settings.shadow_delta),
        tile_size=int(settings.tile_size),
        semantics=str(settings.semantics),
    )
    key = (cfg.H, cfg.W, cfg.tile_size, cfg.semantics, int(gaussians.means.shape[0]))
    renderer = UltraGaussRenderer(cfg, hint=_HINTS.setdefault(key, CapacityHint()))
    image, alpha, depth, inter = renderer(
        gaussians.means,
        gaussians.precision_tril,
        gaussians.opacities,
        gaussians.intensities,
        geometry.pixel_positions_world,
        SliceGeometry(
            origin_world=geometry.probe_origin_world,
            axis_lateral_world=geometry.axis_lateral_world,
            axis_axial_world=geometry.axis_axial_world,
            axis_elevational_world=geometry.axis_elevational_world,
        ),
        return_intermediates=bool(settings.return_pixel_lists),
    )
    inter["alpha_image"] = alpha
    return RenderOutput(image=image, depth=depth, extras=inter)
