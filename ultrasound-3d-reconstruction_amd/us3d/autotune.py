"""Autotuner for the slice splat's launch shape: tile size (8, 16 or 32 pixels; the per-tile Gaussian batch staged in
LDS is one Gaussian per thread of the tile's workgroup, so it follows the tile size).

A candidate is timed on a subset of the batch's frames with the eager training step (map, forward, loss,
adjoint, Adam on scratch copies of the parameters, so tuning does not change them) and the fastest one is kept.
Decisions are cached per (frame size, Gaussian count bucket, frames bucket, semantics) in
$US3D_CACHE_DIR/autotune.json (default ~/.cache/us3d).
"""

from __future__ import annotations

import json
import math
import os
import time
from dataclasses import replace
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import torch

from .batch import FrameSpec, SweepSpec
from .ultragauss import UltraGaussRenderConfig

CANDIDATES = (8, 16, 32)


def _cache_path() -> Path:
    return Path(os.environ.get("US3D_CACHE_DIR", Path.home() / ".cache" / "us3d")) / "autotune.json"


def _key(frames: Sequence[FrameSpec], n_gauss: int, semantics: str) -> str:
    sizes = sorted({(f.height, f.width) for f in frames})
    shape = "+".join(f"{h}x{w}" for h, w in sizes)
    return f"{shape}|g2^{round(math.log2(max(n_gauss, 1)))}|f2^{round(math.log2(len(frames)))}|{semantics}"


def _sample(frames: Sequence[FrameSpec], sweeps: Optional[Sequence[SweepSpec]], target: torch.Tensor, n: int):
    """n frames spread over the batch, their sweeps renumbered compactly, and their target pixels."""
    idx = sorted({int(i * len(frames) / n) for i in range(n)})
    offs = [0]
    for f in frames:
        offs.append(offs[-1] + f.height * f.width)
    flat = target.reshape(-1)
    used: List[int] = []
    sub: List[FrameSpec] = []
    for i in idx:
        s = frames[i].sweep
        if s not in used:
            used.append(s)
        sub.append(replace(frames[i], sweep=used.index(s)))
    tgt = torch.cat([flat[offs[i]:offs[i + 1]] for i in idx])
    sub_sweeps = None if sweeps is None else [sweeps[s] for s in used]
    return sub, sub_sweeps, tgt


def _load() -> Dict[str, dict]:
    p = _cache_path()
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _store(table: Dict[str, dict]) -> None:
    p = _cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(table, indent=1, sort_keys=True))
    except OSError:
        pass


def tune_tile_size(means, tril, raw_opacity, raw_intensity, frames: Sequence[FrameSpec], target: torch.Tensor,
                   cfg: UltraGaussRenderConfig, sample_frames: int = 16, reps: int = 2, use_cache: bool = True,
                   candidates=CANDIDATES, sweeps: Optional[Sequence[SweepSpec]] = None) -> Dict[str, object]:
    """Returns {"tile_size": best, "ms_per_frame": {ts: ms}, "cached": bool}. `sweeps` as for the trainer (frames
    grouped by sweep); the sample keeps each sampled frame's sweep."""
    from .train_step import SliceSplatTrainer

    key = _key(frames, means.shape[0], cfg.semantics)
    table = _load() if use_cache else {}
    if use_cache and key in table:
        hit = dict(table[key])
        hit["cached"] = True
        return hit
    n = min(sample_frames, len(frames))
    sub, sub_sweeps, tgt = _sample(frames, sweeps, target, n)
    timings: Dict[int, float] = {}
    for ts in candidates:
        c = replace(cfg, tile_size=ts)
        tr = SliceSplatTrainer(means, tril, raw_opacity, raw_intensity, sub, tgt, c, sweeps=sub_sweeps,
                               use_graph=False, check_every=1)
        tr.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            tr.step()
        torch.cuda.synchronize()
        timings[ts] = (time.perf_counter() - t0) / (reps * len(sub)) * 1e3
        del tr
    best = min(timings, key=timings.get)
    out = {"tile_size": best, "ms_per_frame": {str(k): round(v, 4) for k, v in timings.items()}, "cached": False}
    if use_cache:
        table[key] = {k: v for k, v in out.items() if k != "cached"}
        _store(table)
    return out
