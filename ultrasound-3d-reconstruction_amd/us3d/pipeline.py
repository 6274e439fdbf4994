"""Training loop counterpart of the reference's single-file pipeline (pipeline_god_file.py).

Same configuration keys (GodConfig, :60-132) and overrides (--config X.json --set k=v, :135-161, :773-783),
same .npz data contract (DATA_FORMAT.md; NpzUltrasoundDataset, :172-333), same parameter container and
activations (GaussianParameters, :340-399), same per-iteration loop (one random frame, L1 + w * (1 - global
SSIM), Adam with one learning rate, CSV/eval/checkpoint callbacks, :640-770). Rendering goes through the
dispatcher, whose `ultragauss`, `ultragray` and `gaussianpile` methods are the MI355X kernels of this package (the drop-in
selected by `method=`); there is no CPU fallback renderer -- a machine without the HIP library fails loudly.
Methods of the reference outside this build (slice_adapted_3dgs, ultrags_baseline) raise.
"""

from __future__ import annotations

import argparse
import csv
import json
import math
import random
import time
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Any, Dict, Iterable, List, Optional, Tuple

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F


@dataclass
class GodConfig:
    method: str = "ultragauss"
    data_npz: str = "dataset.npz"
    output_dir: str = "runs/debug"
    device: str = "cuda"
    seed: int = 42
    iterations: int = 1000
    log_every: int = 10
    eval_every: int = 100
    save_every: int = 500
    num_init: int = 100000
    learning_rate: float = 1.0e-3
    loss_l1_weight: float = 1.0
    loss_ssim_weight: float = 0.0
    train_fraction: float = 0.8
    image_height: int = 256
    image_width: int = 256
    delta_lateral_mm: float = 0.2
    delta_axial_mm: float = 0.2
    delta_elevational_mm: float = 0.5
    near_mm: float = 0.0
    far_mm: float = 51.2
    width_mm: float = 51.2
    center_lateral: bool = True
    init_scale_mm: float = 0.5
    init_opacity: float = 0.1
    init_intensity: float = 0.5
    sigma_z_mm: Optional[float] = None
    slice_thickness_mm: Optional[float] = None
    ultragauss_chi2_threshold: float = 7.815
    ultragauss_diag_beta: float = 1.0e-4
    ultragray_k_sigma: float = 3.0
    ultragray_ops_max_offset_mm: float = 0.0
    ultragray_quadrature_window_radius_mm: float = 2.0
    ultrags_tanfovx: float = 0.5
    ultrags_tanfovy: float = 0.5
    ultrags_active_sh_degree: int = 0


def _parse_scalar(v: str) -> Any:
    low = v.lower()
    if low in ("true", "false"):
        return low == "true"
    try:
        return float(v) if ("." in v or "e" in low) else int(v)
    except ValueError:
        return v


def load_config(path: Optional[str], overrides: List[str]) -> GodConfig:
    data: Dict[str, Any] = {}
    if path:
        data.update(json.loads(Path(path).read_text(encoding="utf-8")))
    for item in overrides:
        if "=" not in item:
            raise ValueError(f"override must be key=value, got: {item}")
        k, v = item.split("=", 1)
        data[k] = _parse_scalar(v)
    cfg = GodConfig(**data)
    if cfg.sigma_z_mm is None:
        cfg.sigma_z_mm = cfg.delta_elevational_mm
    if cfg.slice_thickness_mm is None:
        cfg.slice_thickness_mm = cfg.delta_elevational_mm
    return cfg


def set_seed(seed: int) -> None:
    random.seed(seed)
    np.random.seed(seed)
    torch.manual_seed(seed)
    if torch.cuda.is_available():
        torch.cuda.manual_seed_all(seed)


def _unit(x: torch.Tensor) -> torch.Tensor:
    return x / torch.clamp(torch.linalg.norm(x), min=1.0e-8)


class NpzUltrasoundDataset:
    """Frames of a .npz sweep (DATA_FORMAT.md): images (N,H,W) [uint8 or [0,1]], optional per-frame probe origin
    and axes, pixel positions, rays, poses; missing geometry falls back to a linear probe at the origin."""

    def __init__(self, cfg: GodConfig, split: str):
        self.cfg = cfg
        self.raw = np.load(cfg.data_npz, allow_pickle=False)
        images = self.raw["images"].astype(np.float32)
        if images.max() > 1.5:
            images = images / 255.0
        if images.ndim == 3:
            images = images[:, None]
        if images.ndim != 4:
            raise ValueError("images must be (N,H,W) or (N,1,H,W)")
        self.images = images
        self.n, _, self.h, self.w = images.shape
        cfg.image_height, cfg.image_width = int(self.h), int(self.w)
        order = np.arange(self.n)
        cut = max(1, min(self.n - 1, int(round(self.n * float(cfg.train_fraction))))) if self.n > 1 else 1
        self.indices = order[:cut] if split == "train" else order[cut:]
        if len(self.indices) == 0:
            self.indices = order[:1]

    def __len__(self) -> int:
        return int(len(self.indices))

    def random_frame(self, device: torch.device) -> Dict[str, torch.Tensor]:
        return self.get_frame(int(np.random.choice(self.indices)), device)

    def iter_eval(self, device: torch.device, max_frames: int = 8) -> Iterable[Dict[str, torch.Tensor]]:
        for idx in self.indices[:max_frames]:
            yield self.get_frame(int(idx), device)

    def _per_frame(self, name: str, idx: int, default: np.ndarray) -> np.ndarray:
        if name not in self.raw:
            return default.astype(np.float32)
        arr = self.raw[name].astype(np.float32)
        return arr[idx] if arr.ndim >= 2 and arr.shape[0] == self.n else arr

    def get_frame(self, idx: int, device: torch.device) -> Dict[str, torch.Tensor]:
        cfg = self.cfg
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=torch.float32)  # noqa: E731
        origin = t(self._per_frame("probe_origin_world", idx, np.zeros(3)))
        lat = _unit(t(self._per_frame("axis_lateral_world", idx, np.array([1.0, 0, 0]))))
        ax = _unit(t(self._per_frame("axis_axial_world", idx, np.array([0, 1.0, 0]))))
        ele = _unit(t(self._per_frame("axis_elevational_world", idx, np.array([0, 0, 1.0]))))
        centre = (self.w - 1) * 0.5 if cfg.center_lateral else 0.0
        if "pixel_positions_world" in self.raw:
            pix = t(self.raw["pixel_positions_world"][idx].reshape(-1, 3))
        else:
            u = torch.arange(self.w, device=device, dtype=torch.float32)
            v = torch.arange(self.h, device=device, dtype=torch.float32)
            lat_mm = (u - centre) * float(cfg.delta_lateral_mm)
            ax_mm = float(cfg.near_mm) + v * float(cfg.delta_axial_mm)
            yy, xx = torch.meshgrid(ax_mm, lat_mm, indexing="ij")
            pix = (origin.view(1, 1, 3) + xx[..., None] * lat.view(1, 1, 3) + yy[..., None] * ax.view(1, 1, 3))
            pix = pix.reshape(-1, 3).contiguous()
        if "ray_origins" in self.raw:
            ray_o = t(self.raw["ray_origins"][idx])
        else:
            u = torch.arange(self.w, device=device, dtype=torch.float32)
            ray_o = origin.view(1, 3) + ((u - centre) * float(cfg.delta_lateral_mm))[:, None] * lat.view(1, 3)
        ray_d = t(self.raw["ray_dirs"][idx]) if "ray_dirs" in self.raw else ax.view(1, 3).expand(self.w, 3).contiguous()
        if "pose_world_to_probe" in self.raw:  # pipeline_god_file.py:308-318
            arr = self.raw["pose_world_to_probe"]
            pose = t(arr[idx] if arr.ndim == 3 else arr)
        else:
            rot = torch.stack([lat, ax, ele], dim=0)
            pose = torch.eye(4, device=device, dtype=torch.float32)
            pose[:3, :3] = rot
            pose[:3, 3] = -rot.matmul(origin)
        return {"image": t(self.images[idx]), "frame_index": torch.tensor(idx, device=device),
                "pixel_positions_world": pix, "probe_origin_world": origin, "axis_lateral_world": lat,
                "axis_axial_world": ax, "axis_elevational_world": ele, "ray_origins": ray_o, "ray_dirs": ray_d,
                "pose_world_to_probe": pose}

    def initial_points(self, device: torch.device, num_points: int) -> Tuple[torch.Tensor, torch.Tensor]:
        if "init_means" in self.raw:
            means = torch.from_numpy(self.raw["init_means"].astype(np.float32)).to(device)
            if "init_intensities" in self.raw:
                inten = torch.from_numpy(self.raw["init_intensities"].astype(np.float32)).to(device).reshape(-1, 1)
            else:
                inten = torch.full((means.shape[0], 1), float(self.cfg.init_intensity), device=device)
            return means, inten
        pts, vals = [], []
        have = 0
        while have < num_points:
            fr = self.random_frame(device)
            pix, img = fr["pixel_positions_world"], fr["image"].reshape(-1, 1)
            take = min(num_points - have, pix.shape[0])
            ids = torch.randint(0, pix.shape[0], (take,), device=device)
            jitter = torch.randn((take, 1), device=device) * (0.25 * float(self.cfg.delta_elevational_mm))
            pts.append(pix.index_select(0, ids) + jitter * fr["axis_elevational_world"].view(1, 3))
            vals.append(img.index_select(0, ids).clamp(0.0, 1.0))
            have += take
        return torch.cat(pts), torch.cat(vals)


def _logit(x: torch.Tensor) -> torch.Tensor:
    x = x.clamp(1.0e-6, 1.0 - 1.0e-6)
    return torch.log(x / (1.0 - x))


class GaussianParameters(nn.Module):
    """Raw parameters and activations of pipeline_god_file.py:340-399 (the two renderers' subset)."""

    def __init__(self, cfg: GodConfig, init_means: torch.Tensor, init_intensities: torch.Tensor):
        super().__init__()
        n, dev = init_means.shape[0], init_means.device
        log_s = math.log(float(cfg.init_scale_mm))
        self.means = nn.Parameter(init_means.float())
        self.raw_scales_3d = nn.Parameter(torch.full((n, 3), log_s, device=dev))
        quat = torch.zeros((n, 4), device=dev)
        quat[:, 0] = 1.0
        self.raw_quats = nn.Parameter(quat)
        self.raw_opacities = nn.Parameter(_logit(torch.full((n, 1), float(cfg.init_opacity), device=dev)))
        self.raw_intensities = nn.Parameter(_logit(init_intensities.float().reshape(n, 1).clamp(1.0e-4, 1.0 - 1.0e-4)))
        diag = math.sqrt(max(1.0 / max(float(cfg.init_scale_mm), 1.0e-4) - float(cfg.ultragauss_diag_beta), 1.0e-4))
        tril = torch.zeros((n, 6), device=dev)
        tril[:, 0] = tril[:, 3] = tril[:, 5] = diag
        self.raw_precision_tril = nn.Parameter(tril)
        self.sh_gray4 = nn.Parameter(torch.cat([init_intensities.float().reshape(n, 1), torch.zeros((n, 3), device=dev)], 1))

    def scales_3d(self) -> torch.Tensor:
        return torch.exp(self.raw_scales_3d).clamp_min(1.0e-5)

    def quats(self) -> torch.Tensor:
        return F.normalize(self.raw_quats, dim=-1, eps=1.0e-8)

    def opacities(self) -> torch.Tensor:
        return torch.sigmoid(self.raw_opacities)

    def intensities(self) -> torch.Tensor:
        return torch.sigmoid(self.raw_intensities)


class RenderDispatcher:
    """Maps GaussianParameters + a frame dict to a renderer call, selected by cfg.method
    (RenderDispatcher, pipeline_god_file.py:402-536)."""

    SUPPORTED = ("ultragauss", "ultragray", "gaussianpile")

    def __init__(self, cfg: GodConfig):
        self.cfg = cfg
        self.method = cfg.method
        if self.method not in self.SUPPORTED:
            raise ValueError(f"method {self.method!r} is not provided by this build (supported: {self.SUPPORTED})")
        from .baselines.gaussianpile import render as gp_render
        from .baselines.ultragauss import render as ug_render
        from .baselines.ultragray import render as ur_render

        self.mod = {"ultragauss": ug_render, "ultragray": ur_render, "gaussianpile": gp_render}[self.method]

    def render(self, params: GaussianParameters, frame: Dict[str, torch.Tensor]):
        cfg, m = self.cfg, self.mod
        col = (cfg.image_width - 1) * 0.5 if cfg.center_lateral else 0.0
        if self.method == "ultragauss":
            return m.render_ultragauss(
                m.UltraGaussInput(params.means, params.raw_precision_tril, params.opacities().reshape(-1),
                                  params.intensities().reshape(-1)),
                m.ProbeSliceInput(frame["pixel_positions_world"], frame["probe_origin_world"],
                                  frame["axis_lateral_world"], frame["axis_axial_world"],
                                  frame["axis_elevational_world"]),
                m.UltraGaussRenderSettings(
                    image_height=cfg.image_height, image_width=cfg.image_width,
                    lateral_spacing_mm_per_pix=cfg.delta_lateral_mm, axial_spacing_mm_per_pix=cfg.delta_axial_mm,
                    axial_near_mm=cfg.near_mm, lateral_col_offset=col, chi2_threshold=cfg.ultragauss_chi2_threshold,
                    diag_beta=cfg.ultragauss_diag_beta, return_pixel_lists=False),
            )
        if self.method == "gaussianpile":  # pipeline_god_file.py:435-455
            return m.render_gaussianpile(
                m.GaussianPileInput(params.means, params.scales_3d(), params.quats(), params.opacities().reshape(-1)),
                m.SliceGeometryInput(frame["pixel_positions_world"], frame["probe_origin_world"],
                                     frame["axis_lateral_world"], frame["axis_axial_world"],
                                     frame["axis_elevational_world"], frame["pose_world_to_probe"]),
                m.GaussianPileRenderSettings(
                    image_height=cfg.image_height, image_width=cfg.image_width,
                    lateral_spacing_mm_per_pix=cfg.delta_lateral_mm, axial_spacing_mm_per_pix=cfg.delta_axial_mm,
                    axial_near_mm=cfg.near_mm, lateral_col_offset=col,
                    # load_config's default (pipeline_god_file.py:146-147) for configs built directly
                    sigma_z_mm=float(cfg.sigma_z_mm if cfg.sigma_z_mm is not None else cfg.delta_elevational_mm)),
            )
        pix = frame["pixel_positions_world"].reshape(cfg.image_height, cfg.image_width, 3).permute(1, 0, 2).contiguous()
        return m.render_ultragray(
            m.UltraGrayInput(params.means, params.scales_3d(), params.quats(), params.raw_opacities.reshape(-1),
                             params.sh_gray4),
            m.RaySliceInput(pix, frame["ray_origins"], frame["ray_dirs"], frame["probe_origin_world"],
                            frame["axis_lateral_world"], frame["axis_axial_world"], frame["axis_elevational_world"]),
            m.UltraGrayRenderSettings(
                image_width=cfg.image_width, image_height=cfg.image_height, width_mm=cfg.width_mm,
                near_mm=cfg.near_mm, lateral_col_offset=col, k_sigma=cfg.ultragray_k_sigma,
                ops_max_offset_mm=cfg.ultragray_ops_max_offset_mm,
                quadrature_window_radius_mm=cfg.ultragray_quadrature_window_radius_mm),
        )


def _canon(x: torch.Tensor) -> torch.Tensor:
    if x.ndim == 2:
        x = x.unsqueeze(0)
    if x.ndim == 3 and x.shape[0] != 1:
        x = x[:1]
    return x.float()


def simple_ssim(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """Single-window global SSIM (pipeline_god_file.py:628-637)."""
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    mx, my = x.mean(), y.mean()
    vx, vy = ((x - mx) ** 2).mean(), ((y - my) ** 2).mean()
    vxy = ((x - mx) * (y - my)).mean()
    return ((2 * mx * my + c1) * (2 * vxy + c2)) / ((mx ** 2 + my ** 2 + c1) * (vx + vy + c2))


def image_loss(pred: torch.Tensor, gt: torch.Tensor, cfg: GodConfig) -> Tuple[torch.Tensor, Dict[str, float]]:
    """w_L1 * mean|pred - gt| + w_SSIM * (1 - SSIM) (pipeline_god_file.py:606-625)."""
    pred = _canon(pred)
    gt = _canon(gt).to(dtype=pred.dtype, device=pred.device)
    if pred.shape != gt.shape:
        pred = pred.reshape_as(gt)
    l1 = torch.mean(torch.abs(pred - gt))
    total = float(cfg.loss_l1_weight) * l1
    ssim_loss = torch.zeros((), device=pred.device)
    if cfg.loss_ssim_weight > 0:
        ssim_loss = 1.0 - simple_ssim(pred, gt)
        total = total + float(cfg.loss_ssim_weight) * ssim_loss
    return total, {"l1": float(l1.detach().item()), "ssim_loss": float(ssim_loss.detach().item()),
                   "total": float(total.detach().item())}


class Callback:
    def on_start(self, ctx: Dict[str, Any]) -> None:
        pass

    def on_step_end(self, ctx: Dict[str, Any]) -> None:
        pass

    def on_finish(self, ctx: Dict[str, Any]) -> None:
        pass


class CsvLoggerCallback(Callback):
    FIELDS = ["iteration", "split", "loss_total", "loss_l1", "loss_ssim", "num_gaussians", "seconds"]

    def on_start(self, ctx):
        self.f = open(Path(ctx["cfg"].output_dir) / "scalars.csv", "w", newline="", encoding="utf-8")
        self.w = csv.DictWriter(self.f, fieldnames=self.FIELDS)
        self.w.writeheader()

    def on_step_end(self, ctx):
        it = int(ctx["iteration"])
        if it % int(ctx["cfg"].log_every) != 0:
            return
        losses = ctx["losses"]
        row = {"iteration": it, "split": "train", "loss_total": losses["total"], "loss_l1": losses["l1"],
               "loss_ssim": losses["ssim_loss"], "num_gaussians": int(ctx["params"].means.shape[0]),
               "seconds": time.time() - ctx["start_time"]}
        self.w.writerow(row)
        self.f.flush()
        print(f"[iter {it}] loss={row['loss_total']:.6f} l1={row['loss_l1']:.6f} n={row['num_gaussians']}", flush=True)

    def on_finish(self, ctx):
        self.f.close()


class EvalCallback(Callback):
    def on_step_end(self, ctx):
        cfg = ctx["cfg"]
        it = int(ctx["iteration"])
        if cfg.eval_every <= 0 or it % int(cfg.eval_every) != 0:
            return
        vals = []
        with torch.no_grad():
            for frame in ctx["test_data"].iter_eval(ctx["device"], max_frames=4):
                out = ctx["renderer"].render(ctx["params"], frame)
                vals.append(image_loss(out.image, frame["image"], cfg)[1]["l1"])
        mean_l1 = float(np.mean(vals)) if vals else float("nan")
        with open(Path(cfg.output_dir) / "eval.csv", "a", encoding="utf-8") as f:
            f.write(f"{it},{mean_l1}\n")
        print(f"[eval {it}] mean_l1={mean_l1:.6f}", flush=True)


class CheckpointCallback(Callback):
    def on_start(self, ctx):
        self._save(ctx, "initial.pt")

    def on_step_end(self, ctx):
        cfg = ctx["cfg"]
        it = int(ctx["iteration"])
        if cfg.save_every > 0 and it % int(cfg.save_every) == 0:
            self._save(ctx, f"iter_{it:06d}.pt")

    def on_finish(self, ctx):
        self._save(ctx, "final.pt")

    def _save(self, ctx, name):
        torch.save({"config": asdict(ctx["cfg"]), "iteration": int(ctx.get("iteration", 0)),
                    "params": ctx["params"].state_dict(), "optimizer": ctx["optimizer"].state_dict()},
                   Path(ctx["cfg"].output_dir) / name)


def train(cfg: GodConfig) -> Dict[str, Any]:
    out_dir = Path(cfg.output_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    (out_dir / "used_config.json").write_text(json.dumps(asdict(cfg), ensure_ascii=False, indent=2), encoding="utf-8")
    set_seed(cfg.seed)
    if not torch.cuda.is_available():
        raise RuntimeError("the MI355X renderers need a HIP device; there is no CPU fallback in this build")
    device = torch.device(cfg.device)
    train_data = NpzUltrasoundDataset(cfg, "train")
    test_data = NpzUltrasoundDataset(cfg, "test")
    means, inten = train_data.initial_points(device, int(cfg.num_init))
    params = GaussianParameters(cfg, means, inten).to(device)
    renderer = RenderDispatcher(cfg)
    optimizer = torch.optim.Adam(params.parameters(), lr=float(cfg.learning_rate))
    callbacks: List[Callback] = [CsvLoggerCallback(), EvalCallback(), CheckpointCallback()]
    ctx: Dict[str, Any] = {"cfg": cfg, "device": device, "train_data": train_data, "test_data": test_data,
                           "params": params, "renderer": renderer, "optimizer": optimizer, "iteration": 0,
                           "losses": {}, "start_time": time.time()}
    for cb in callbacks:
        cb.on_start(ctx)
    for it in range(1, int(cfg.iterations) + 1):
        ctx["iteration"] = it
        frame = train_data.random_frame(device)
        optimizer.zero_grad(set_to_none=True)
        out = renderer.render(params, frame)
        loss, parts = image_loss(out.image, frame["image"], cfg)
        loss.backward()
        optimizer.step()
        ctx["losses"] = parts
        for cb in callbacks:
            cb.on_step_end(ctx)
    for cb in callbacks:
        cb.on_finish(ctx)
    return ctx


def main(argv: Optional[List[str]] = None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=str, default=None, help="JSON config path")
    ap.add_argument("--set", nargs="*", default=[], help="overrides, e.g. method=ultragauss iterations=10")
    args = ap.parse_args(argv)
    train(load_config(args.config, args.set))


if __name__ == "__main__":
    main()
