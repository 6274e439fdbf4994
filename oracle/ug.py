"""Oracle (test infrastructure only): ctypes front end of ug_oracle.cpp, one frame at a time.

Functions mirror the reference stages: preprocess (core.cu:46-116), emit + stable sort + tile ranges
(core.cu:118-164, wrapper_mapper.cu:97-152), per-pixel lists (core.cu:166-287), forward (forward.cu:17-76)
in FP32/FP64 and backward (backward.cu:17-131) in FP64.
"""

from __future__ import annotations

import ctypes
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_F = ctypes.c_float


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        L.ug_oracle_preprocess.argtypes = [_I, _P, _P, _P, _P, _P, _P, _I, _I, _F, _F, _F, _F, _F, _F, _I, _I, _P, _P, _P]
        L.ug_oracle_emit_sort.argtypes = [_I, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P]
        L.ug_oracle_pixel_lists.argtypes = [_I, _P, _P, _P, _I, _I, _P, _P, _F, _F, _I, _P, _P, _P, _I, _P, _P, _P]
        grid = [_P, _I, _F, _F, _F, _F]  # plane form (or NULL), W, col_offset, delta_lat, near, delta_ax
        fwd = [_I, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _F, _F, _F, _F, _F, _I, _P, _P, _P, _P] + grid
        L.ug_oracle_forward_f32.argtypes = fwd
        L.ug_oracle_forward_f64.argtypes = fwd
        L.ug_oracle_backward_f64.argtypes = [_I, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _F, _F, _F, _F, _F, _I,
                                             _P, _P, _P, _P, _P, _P, _P] + grid
        L.ug_oracle_plane_form.argtypes = [_I, _P, _P, _P, _P, _P, _P, _F, _P]
        L.ug_oracle_backward_f32.argtypes = L.ug_oracle_backward_f64.argtypes
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class FrameGeom:
    origin: np.ndarray
    axis_lat: np.ndarray
    axis_ax: np.ndarray
    axis_ele: np.ndarray
    H: int
    W: int
    delta_lat: float
    delta_ax: float
    near_mm: float
    col_offset: float

    @staticmethod
    def from_spec(spec) -> "FrameGeom":
        n = lambda t: np.asarray(t.detach().cpu().numpy(), dtype=np.float32).reshape(3)  # noqa: E731
        return FrameGeom(n(spec.origin), n(spec.axis_lat), n(spec.axis_ax), n(spec.axis_ele), int(spec.height),
                         int(spec.width), float(spec.delta_lat), float(spec.delta_ax), float(spec.near_mm),
                         float(spec.col_offset))


def plane_form(means, tril, fr: FrameGeom, beta=1e-4) -> np.ndarray:
    """(G, 6) float32 (cx, cy, K, p, qq, s): each Gaussian's plane form of q on the frame's own pixel grid, as the
    kernels set it up per visible pair. Passing it to forward / backward_f64 (with the frame's grid) makes their
    q <= chi2 decisions, and the FP32 values of q, those of a frame rendered on its own grid (pixels = NULL)."""
    L = lib()
    means = _c(means, np.float32)
    tril = _c(tril, np.float32)
    out = np.zeros((means.shape[0], 6), np.float32)
    L.ug_oracle_plane_form(means.shape[0], _p(means), _p(tril), _p(_c(fr.origin, np.float32)),
                           _p(_c(fr.axis_lat, np.float32)), _p(_c(fr.axis_ax, np.float32)),
                           _p(_c(fr.axis_ele, np.float32)), beta, _p(out))
    return out


def _grid_args(plane, fr: "FrameGeom"):
    if plane is None:
        return [None, 0, 0.0, 0.0, 0.0, 0.0]
    return [_p(plane), fr.W, fr.col_offset, fr.delta_lat, fr.near_mm, fr.delta_ax]


@dataclass
class MapResult:
    count: np.ndarray
    rect: np.ndarray
    dk: np.ndarray
    keys: np.ndarray
    vals: np.ndarray
    tile_start: np.ndarray
    tile_end: np.ndarray


def map_frame(means, tril, fr: FrameGeom, chi2=7.815, beta=1e-4, sort_by_depth=True, ts=16) -> MapResult:
    L = lib()
    means = _c(means, np.float32)
    tril = _c(tril, np.float32)
    G = means.shape[0]
    count = np.zeros(G, np.int32)
    rect = np.zeros((G, 4), np.int32)
    dk = np.zeros(G, np.uint32)
    L.ug_oracle_preprocess(G, _p(means), _p(tril), _p(_c(fr.origin, np.float32)), _p(_c(fr.axis_lat, np.float32)),
                           _p(_c(fr.axis_ax, np.float32)), _p(_c(fr.axis_ele, np.float32)), fr.H, fr.W, fr.delta_lat,
                           fr.delta_ax, fr.near_mm, fr.col_offset, chi2, beta, int(sort_by_depth), ts, _p(count),
                           _p(rect), _p(dk))
    K = int(count.astype(np.int64).sum())
    keys = np.zeros(K, np.uint64)
    vals = np.zeros(K, np.int32)
    n_tiles = (-(-fr.W // ts)) * (-(-fr.H // ts))
    start = np.zeros(n_tiles, np.int32)
    end = np.zeros(n_tiles, np.int32)
    L.ug_oracle_emit_sort(G, _p(count), _p(rect), _p(dk), fr.H, fr.W, ts, int(sort_by_depth), _p(keys), _p(vals),
                          _p(start), _p(end))
    return MapResult(count, rect, dk, keys, vals, start, end)


def pixel_lists(means, tril, pixels, fr: FrameGeom, mr: MapResult, chi2=7.815, beta=1e-4, ts=16):
    L = lib()
    means = _c(means, np.float32)
    tril = _c(tril, np.float32)
    pixels = _c(pixels, np.float32)
    G = means.shape[0]
    P = fr.H * fr.W
    counts = np.zeros(P, np.int32)
    args = (G, _p(means), _p(tril), _p(pixels), fr.H, fr.W, _p(_c(fr.axis_lat, np.float32)),
            _p(_c(fr.axis_ax, np.float32)), chi2, beta, ts, _p(mr.vals), _p(mr.tile_start), _p(mr.tile_end))
    L.ug_oracle_pixel_lists(*args, 0, _p(counts), None, None)
    offsets = np.zeros(P, np.int32)
    offsets[1:] = np.cumsum(counts)[:-1]
    indices = np.zeros(int(counts.sum()), np.int32)
    L.ug_oracle_pixel_lists(*args, 1, None, _p(offsets), _p(indices))
    return counts, offsets, indices


@dataclass
class Semantics:
    chi2: float = 7.815
    beta: float = 1e-4
    alpha_min: float = 1e-8
    bg: float = 1e-6
    bg_intensity: float = 0.0
    alpha_excludes_bg: bool = False

    @staticmethod
    def fallback(beta=1e-4) -> "Semantics":
        return Semantics(chi2=120.0, beta=beta, alpha_min=-1.0, bg=1e-6, bg_intensity=0.0, alpha_excludes_bg=True)


def forward(means, tril, opac, inten, pixels, fr: FrameGeom, lists, sem: Semantics, fp64=False, plane=None):
    """FP32 (canonical) or FP64 forward; the FP64 path takes the parameter arrays in float64. plane: the frame's
    plane forms (plane_form) when it is rendered on its own grid."""
    L = lib()
    counts, offsets, indices = lists
    P = counts.shape[0]
    dt = np.float64 if fp64 else np.float32
    image, alpha, depth = np.zeros(P, dt), np.zeros(P, dt), np.zeros(P, dt)
    sums = np.zeros((P, 3), dt)
    fn = L.ug_oracle_forward_f64 if fp64 else L.ug_oracle_forward_f32
    plane_c = None if plane is None else _c(plane, np.float32)  # kept alive across the call
    fn(int(np.asarray(means).shape[0]), _p(_c(means, dt)), _p(_c(tril, dt)), _p(_c(opac, dt)),
       _p(_c(inten, dt)), _p(_c(pixels, dt)), P, _p(_c(fr.origin, np.float32)),
       _p(_c(fr.axis_ele, np.float32)), _p(_c(counts, np.int32)), _p(_c(offsets, np.int32)), _p(_c(indices, np.int32)),
       sem.chi2, sem.beta, sem.alpha_min, sem.bg, sem.bg_intensity, int(sem.alpha_excludes_bg), _p(image), _p(alpha),
       _p(depth), _p(sums), *_grid_args(plane_c, fr))
    return image, alpha, depth, sums


def backward_f64(means, tril, opac, inten, pixels, fr: FrameGeom, lists, sem: Semantics, d_image, d_alpha=None,
                 d_depth=None, gate=True, fp32=False, plane=None):
    """FP64 restatement of the reference backward (FP32 with fp32=True: the reference's precision, naive
    per-pair accumulation in pixel order)."""
    L = lib()
    counts, offsets, indices = lists
    P = counts.shape[0]
    G = int(np.asarray(means).shape[0])
    f8 = np.float32 if fp32 else np.float64
    dm, dt, do, di = np.zeros((G, 3), f8), np.zeros((G, 6), f8), np.zeros(G, f8), np.zeros(G, f8)
    opt = lambda a: None if a is None else _c(a, np.float32)  # noqa: E731
    gA, gD = opt(d_alpha), opt(d_depth)
    fn = L.ug_oracle_backward_f32 if fp32 else L.ug_oracle_backward_f64
    plane_c = None if plane is None else _c(plane, np.float32)  # kept alive across the call
    fn(G, _p(_c(means, f8)), _p(_c(tril, f8)), _p(_c(opac, f8)),
                             _p(_c(inten, f8)), _p(_c(pixels, f8)), P, _p(_c(fr.origin, np.float32)),
                             _p(_c(fr.axis_ele, np.float32)), _p(_c(counts, np.int32)), _p(_c(offsets, np.int32)),
                             _p(_c(indices, np.int32)), sem.chi2, sem.beta, sem.alpha_min, sem.bg, sem.bg_intensity,
                             int(gate), _p(_c(d_image, np.float32)), _p(gA), _p(gD), _p(dm), _p(dt), _p(do), _p(di),
                             *_grid_args(plane_c, fr))
    return dm, dt, do, di
