"""Autotuner for the slice splat's launch shape: the tile size (8, 16 or 32 pixels) and, per tile size, the number
of Gaussians the forward splat and the adjoint each stage in LDS per batch (64, 128 or 256).

Every candidate runs on a subset of the batch's frames with the eager training step (map, forward, loss, adjoint,
Adam on scratch copies of the parameters, so tuning does not change them), timed per kernel with the library's
HIP-event stage timing. For each tile size the LDS batch sizes are tried with the same mapping; the forward and the
adjoint take their own fastest size (neither kernel's time depends on the other's), and the tile size with the
least kernel time per frame, at its best batch sizes, is kept. Decisions are cached per (frame size, Gaussian
count bucket, frames bucket, semantics) in $US3D_CACHE_DIR/autotune.json (default ~/.cache/us3d).
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
LDS_BATCHES = (64, 128, 256)
FWD, BWD = "ug_forward_kernel", "ug_backward_kernel"


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
    """Returns {"tile_size", "fwd_lds_batch", "bwd_lds_batch", "ms_per_frame": {ts: kernel ms per frame at its
    best batch sizes}, "lds_batch_ms": {ts: {batch: [forward ms, adjoint ms] per frame}}, "cached": bool}.
    `sweeps` as for the trainer (frames grouped by sweep); the sample keeps each sampled frame's sweep.
    apply(cfg, result) gives the configuration with the choice."""
    from . import _lib
    from .train_step import SliceSplatTrainer

    key = _key(frames, means.shape[0], cfg.semantics)
    table = _load() if use_cache else {}
    if use_cache and key in table and "fwd_lds_batch" in table[key]:
        hit = dict(table[key])
        hit["cached"] = True
        return hit
    n = min(sample_frames, len(frames))
    sub, sub_sweeps, tgt = _sample(frames, sweeps, target, n)
    per_frame: Dict[int, float] = {}
    lds: Dict[int, Dict[int, List[float]]] = {}
    pick: Dict[int, tuple] = {}
    for ts in candidates:
        c = replace(cfg, tile_size=ts, fwd_lds_batch=0, bwd_lds_batch=0)
        tr = SliceSplatTrainer(means, tril, raw_opacity, raw_intensity, sub, tgt, c, sweeps=sub_sweeps,
                               use_graph=False, check_every=1)
        tr.step()
        lds[ts] = {}
        other = 0.0
        for b in LDS_BATCHES:
            for chunk in tr.chunks:
                chunk[4].settings.fwd_lds_batch = b
                chunk[4].settings.bwd_lds_batch = b
            tr.step()  # warm-up at this size
            torch.cuda.synchronize()
            _lib.stage_timing_read()
            _lib.stage_timing(True)
            try:
                for _ in range(reps):
                    tr.step()
                torch.cuda.synchronize()
            finally:
                _lib.stage_timing(False)
            st = _lib.stage_timing_read()
            scale = 1.0 / (reps * len(sub))
            lds[ts][b] = [st.get(FWD, (0, 0.0))[1] * scale, st.get(BWD, (0, 0.0))[1] * scale]
            other = max(other, sum(v[1] for k, v in st.items() if k not in (FWD, BWD)) * scale)
        bf = min(LDS_BATCHES, key=lambda b: lds[ts][b][0])
        bb = min(LDS_BATCHES, key=lambda b: lds[ts][b][1])
        pick[ts] = (bf, bb)
        per_frame[ts] = other + lds[ts][bf][0] + lds[ts][bb][1]
        del tr
    best = min(per_frame, key=per_frame.get)
    out = {"tile_size": best, "fwd_lds_batch": pick[best][0], "bwd_lds_batch": pick[best][1],
           "ms_per_frame": {str(k): round(v, 4) for k, v in per_frame.items()},
           "lds_batch_ms": {str(t): {str(b): [round(x, 4) for x in v] for b, v in d.items()} for t, d in lds.items()},
           "cached": False}
    if use_cache:
        table[key] = {k: v for k, v in out.items() if k != "cached"}
        _store(table)
    return out


def apply(cfg: UltraGaussRenderConfig, tuned: Dict[str, object]) -> UltraGaussRenderConfig:
    """cfg with the tuned tile size and LDS batch sizes."""
    return replace(cfg, tile_size=int(tuned["tile_size"]), fwd_lds_batch=int(tuned["fwd_lds_batch"]),
                   bwd_lds_batch=int(tuned["bwd_lds_batch"]))
