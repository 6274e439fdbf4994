"""Batch descriptors shared by both renderers: frames, sweeps and the us3d_plan handed to the C ABI.

A batch is a list of frames, each rendered from one sweep's Gaussians (a contiguous row range of the
concatenated Gaussian arrays). The frame table is assembled on the device from the geometry tensors, so
building it never waits on the GPU.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib

FRAME_WORDS = 24  # us3d_frame: 16 float fields + 8 int32 fields


@dataclass
class FrameSpec:
    """Geometry of one frame. Vectors are (3,) tensors (any device); scalars are python numbers."""

    origin: torch.Tensor
    axis_lat: torch.Tensor
    axis_ax: torch.Tensor
    axis_ele: torch.Tensor
    height: int
    width: int
    delta_lat: float
    delta_ax: float
    near_mm: float
    col_offset: float
    sweep: int = 0
    # 0: linear probe grid; 1: fan probe (UltraGauss only): origin = apex, delta_lat = radians per column,
    # col_offset = column at angle 0, near_mm = radius of row 0, delta_ax = radial step
    geometry: int = 0


@dataclass
class SweepSpec:
    g_begin: int
    g_count: int


@dataclass
class BatchLayout:
    """Host-side integers of a batch (everything the C ABI needs except device tensors)."""

    n_gauss: int
    n_frames: int
    n_sweeps: int
    n_tiles: int
    n_pairs: int
    n_pixels: int
    frame_tile_base: List[int]
    frame_pixel_base: List[int]
    frame_ray_base: List[int]
    tiles_x: List[int]
    tiles_y: List[int]
    sweep_rows: np.ndarray  # (n_sweeps, 6) int32 = g_begin, g_count, f_begin, f_count, pair_base lo/hi


def plan_layout(frames: Sequence[FrameSpec], sweeps: Sequence[SweepSpec], n_gauss: int, tile_size: Optional[int],
                ) -> BatchLayout:
    """Tile and pair numbering. tile_size=None lays out UltraG-Ray scanlines (one segment per ray)."""
    if not frames:
        raise ValueError("a batch needs at least one frame")
    # frames of a sweep must be contiguous and sweeps in order (pairs are numbered sweep by sweep)
    order = [f.sweep for f in frames]
    if order != sorted(order):
        raise ValueError("frames must be grouped by sweep in ascending sweep order")
    f_begin = {}
    f_count = {}
    for i, s in enumerate(order):
        f_begin.setdefault(s, i)
        f_count[s] = f_count.get(s, 0) + 1
    rows = np.zeros((len(sweeps), 6), dtype=np.int32)
    pair_base = 0
    g_end = 0
    for s, sw in enumerate(sweeps):
        if sw.g_begin < 0 or sw.g_count < 0 or sw.g_begin + sw.g_count > n_gauss:
            raise ValueError(f"sweep {s}: Gaussian range out of bounds")
        # the kernels find a Gaussian's sweep by its position in the table; Gaussians in no range get zero
        # gradients
        if sw.g_begin < g_end:
            raise ValueError(f"sweep {s}: Gaussian ranges must be disjoint and in ascending order")
        g_end = sw.g_begin + sw.g_count
        fb, fc = f_begin.get(s, 0), f_count.get(s, 0)
        rows[s, :4] = (sw.g_begin, sw.g_count, fb, fc)
        rows[s, 4:6] = np.array([pair_base], dtype=np.int64).view(np.int32)
        pair_base += sw.g_count * fc
    tile_base, pixel_base, ray_base, tx, ty = [0], [0], [0], [], []
    for f in frames:
        if f.height <= 0 or f.width <= 0:
            raise ValueError("frame height/width must be positive")
        if tile_size is None:
            nx, ny = 1, f.width
        else:
            nx, ny = -(-f.width // tile_size), -(-f.height // tile_size)
        tx.append(nx)
        ty.append(ny)
        tile_base.append(tile_base[-1] + nx * ny)
        pixel_base.append(pixel_base[-1] + f.height * f.width)
        ray_base.append(ray_base[-1] + f.width)
    if tile_base[-1] >= 2**31 or pixel_base[-1] >= 2**31 or pair_base >= 2**31:
        raise ValueError("batch too large for 32-bit tile/pixel/pair indices")
    return BatchLayout(
        n_gauss=n_gauss, n_frames=len(frames), n_sweeps=len(sweeps), n_tiles=tile_base[-1], n_pairs=pair_base,
        n_pixels=pixel_base[-1], frame_tile_base=tile_base, frame_pixel_base=pixel_base, frame_ray_base=ray_base,
        tiles_x=tx, tiles_y=ty, sweep_rows=rows,
    )


def frame_table(frames: Sequence[FrameSpec], lay: BatchLayout, device: torch.device) -> torch.Tensor:
    """(n_frames, 25) 4-byte words laid out as us3d_frame, assembled on `device` without a host sync."""
    f32 = torch.float32
    vec = torch.stack(
        [torch.cat([f.origin.reshape(3), f.axis_lat.reshape(3), f.axis_ax.reshape(3), f.axis_ele.reshape(3)]).to(
            device=device, dtype=f32) for f in frames]
    )
    scal = torch.tensor([[f.delta_lat, f.delta_ax, f.near_mm, f.col_offset] for f in frames], dtype=f32)
    ints = np.array(
        [[f.height, f.width, lay.tiles_x[i], lay.tiles_y[i], lay.frame_tile_base[i], lay.frame_pixel_base[i], f.sweep,
          lay.frame_ray_base[i], int(f.geometry)] for i, f in enumerate(frames)],
        dtype=np.int32,
    )
    # concatenate as raw 32-bit words (int32 views) so the integer fields keep their bits exactly
    host = torch.cat([scal.view(torch.int32), torch.from_numpy(ints)], dim=1)
    return torch.cat([vec.contiguous().view(torch.int32), host.to(device)], dim=1).contiguous()


def _frame_rows_host(frames: Sequence[FrameSpec], lay: BatchLayout) -> np.ndarray:
    """(n_frames, 25) int32 words laid out as us3d_frame, from frames whose vectors are host tensors."""
    n = len(frames)
    vec = np.empty((n, 16), dtype=np.float32)
    for i, f in enumerate(frames):
        for k, v in enumerate((f.origin, f.axis_lat, f.axis_ax, f.axis_ele)):
            vec[i, 3 * k:3 * k + 3] = v.detach().reshape(3).to(torch.float32).numpy()
        vec[i, 12:16] = (f.delta_lat, f.delta_ax, f.near_mm, f.col_offset)
    ints = np.array(
        [[f.height, f.width, lay.tiles_x[i], lay.tiles_y[i], lay.frame_tile_base[i], lay.frame_pixel_base[i], f.sweep,
          lay.frame_ray_base[i], int(f.geometry)] for i, f in enumerate(frames)],
        dtype=np.int32,
    ).reshape(n, 9)
    return np.concatenate([vec.view(np.int32), ints], axis=1)


class DeviceTables:
    """Device copies of the frame/sweep/prefix tables of one batch (kept alive by the owner). When every frame
    vector is a host tensor the four tables travel in ONE host-to-device copy of a packed int32 buffer: sweep
    rows first (us3d_sweep holds an int64, so they sit at offset 0), then the frame rows, the tile bases and
    the pixel bases; each table is a view of that buffer."""

    def __init__(self, frames: Sequence[FrameSpec], lay: BatchLayout, device: torch.device):
        i32 = torch.int32
        on_host = all(v.device.type == "cpu" for f in frames for v in (f.origin, f.axis_lat, f.axis_ax, f.axis_ele))
        if on_host:
            rows = _frame_rows_host(frames, lay)
            parts = [lay.sweep_rows.reshape(-1).astype(np.int32), rows.reshape(-1),
                     np.asarray(lay.frame_tile_base, dtype=np.int32).reshape(-1),
                     np.asarray(lay.frame_pixel_base, dtype=np.int32).reshape(-1)]
            buf = torch.from_numpy(np.concatenate(parts)).to(device)
            o1 = parts[0].size
            o2 = o1 + parts[1].size
            o3 = o2 + parts[2].size
            self.sweeps = buf[:o1].view(lay.sweep_rows.shape)
            self.frames = buf[o1:o2].view(len(frames), 25)
            self.tile_base = buf[o2:o3]
            self.pixel_base = buf[o3:]
            self._buf = buf
        else:
            self.frames = frame_table(frames, lay, device)
            self.tile_base = torch.tensor(lay.frame_tile_base, dtype=i32).to(device)
            self.pixel_base = torch.tensor(lay.frame_pixel_base, dtype=i32).to(device)
            self.sweeps = torch.from_numpy(lay.sweep_rows.copy()).to(device)
        self.lay = lay

    def plan(self, workspace: torch.Tensor, k_capacity: int, v_capacity: int) -> _lib.Plan:
        lay = self.lay
        p = _lib.Plan()
        p.n_gauss = lay.n_gauss
        p.n_frames = lay.n_frames
        p.n_sweeps = lay.n_sweeps
        p.n_tiles = lay.n_tiles
        p.n_pairs = lay.n_pairs
        p.n_pixels = lay.n_pixels
        p.k_capacity = int(k_capacity)
        p.v_capacity = int(v_capacity)
        p.frames = _lib.ptr(self.frames)
        p.frame_tile_base = _lib.ptr(self.tile_base)
        p.frame_pixel_base = _lib.ptr(self.pixel_base)
        p.sweeps = _lib.ptr(self.sweeps)
        p.workspace = _lib.ptr(workspace)
        p.workspace_bytes = int(workspace.numel())
        return p


def new_workspace(nbytes: int, device) -> torch.Tensor:
    """A workspace of `nbytes` bytes with its us3d_stats block zeroed, so the sticky overflow_any flag starts
    clear (every map call resets the rest of the block itself)."""
    ws = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
    ws[: ctypes.sizeof(_lib.Stats)].zero_()
    return ws


def read_stats(workspace: torch.Tensor) -> _lib.Stats:
    """Copy the us3d_stats block at workspace offset 0 to the host (synchronises the stream)."""
    raw = workspace[: ctypes.sizeof(_lib.Stats)].cpu().numpy().tobytes()
    return _lib.Stats.from_buffer_copy(raw)


I32_LIMIT = (1 << 31) - 1


@dataclass
class CapacityHint:
    """Remembers the last instance and visible-pair counts so the next workspace rarely overflows."""

    k_last: int = 0
    v_last: int = 0
    growth: float = 1.25
    floor: int = 1 << 16

    def capacity(self, n_pairs: int) -> int:
        guess = max(self.floor, int(self.k_last * self.growth) + 1024, min(4 * n_pairs, 1 << 22))
        return min(guess, I32_LIMIT)

    def v_capacity(self, n_pairs: int) -> int:
        """Visible (frame, Gaussian) pairs: at most n_pairs."""
        guess = max(1 << 12, int(self.v_last * self.growth) + 1024, min(n_pairs, 1 << 18))
        return max(1, min(guess, n_pairs, I32_LIMIT))

    def update(self, k: int, v: int = 0) -> None:
        self.k_last = max(self.k_last, int(k))
        self.v_last = max(self.v_last, int(v))

    def regrow(self, k_cap: int, v_cap: int, st: _lib.Stats, n_pairs: int) -> Tuple[int, int]:
        """Capacities for the retry of a map that overflowed: the visible pairs and the instances it counted, with
        headroom (a map whose visible pairs overflowed counts no instances, so K may need a second retry)."""
        if st.n_visible > v_cap:
            v_cap = max(1, min(int(st.n_visible * self.growth) + 1024, n_pairs, I32_LIMIT))
        if st.n_instances > k_cap:
            if st.n_instances >= I32_LIMIT:
                raise RuntimeError("instance count exceeds 2^31; split the batch")
            k_cap = min(int(st.n_instances * self.growth) + 1024, I32_LIMIT)
        return k_cap, v_cap


@dataclass
class Mapped:
    """Workspace of a mapped batch plus what forward/backward need to find it."""

    tables: DeviceTables
    workspace: torch.Tensor
    k_capacity: int
    v_capacity: int
    stats: Optional[_lib.Stats] = None
    extra: dict = field(default_factory=dict)

    def plan(self) -> _lib.Plan:
        return self.tables.plan(self.workspace, self.k_capacity, self.v_capacity)
