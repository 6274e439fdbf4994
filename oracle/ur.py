"""Oracle (test infrastructure only): ctypes front end of ur_oracle.cpp (UltraG-Ray), one frame at a time.

Stages mirror the reference: preprocess (core.cu:3-68), emit + stable 64-bit sort + CSR (core.cu:70-129,
wrapper_mapper.cu:21-127), per-pixel forward (forward.cu:22-119) in FP32/FP64 and backward
(backward.cu:26-325) in FP64.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import ug as _ug

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float


def lib():
    L = _ug.lib()
    if not getattr(L, "_ur_ready", False):
        L.ur_oracle_preprocess.argtypes = [_I, _P, _P, _P, _P, _P, _P, _I, _F, _F, _F, _F, _F, _P, _P, _P, _P]
        L.ur_oracle_emit_sort.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P]
        args = [_I, _P, _P, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _F, _I, _F]
        L.ur_oracle_forward_f32.argtypes = args + [_P, _P]
        L.ur_oracle_forward_f64.argtypes = args + [_P, _P]
        L.ur_oracle_backward_f64.argtypes = args + [_P, _F, _P, _P, _P, _P, _P]
        L.ur_oracle_window_centres.argtypes = [_I, _P, _P, _P, _I, _P, _P, _P, _P, _P]
        L._ur_ready = True
    return L


_c = _ug._c
_p = _ug._p


@dataclass
class RayFrame:
    """One UltraG-Ray frame: probe axes, scanline geometry, pixel grid (W, H, 3)."""

    origin: np.ndarray
    axis_lat: np.ndarray
    axis_ax: np.ndarray
    axis_ele: np.ndarray
    W: int
    H: int
    width_mm: float
    near_mm: float
    col_offset: float
    pixels: np.ndarray     # (W, H, 3)
    ray_o: np.ndarray      # (W, 3)
    ray_d: np.ndarray      # (W, 3)
    delta: np.ndarray      # (W,) OPS offsets


@dataclass
class RaySettings:
    k_sigma: float = 3.0
    Rw: float = 2.0
    elev_margin: float = 0.0
    local_only: bool = False
    local_max: float = 17.0
    grad_threshold: float = 1e-6


@dataclass
class RayMap:
    count: np.ndarray
    rmin: np.ndarray
    rmax: np.ndarray
    dq: np.ndarray
    counts: np.ndarray
    offsets: np.ndarray
    indices: np.ndarray
    sorted_depths: np.ndarray


def map_frame(means, scales, fr: RayFrame, st: RaySettings) -> RayMap:
    L = lib()
    means = _c(means, np.float32)
    scales = _c(scales, np.float32)
    G = means.shape[0]
    count, rmin, rmax = np.zeros(G, np.int32), np.zeros(G, np.int32), np.zeros(G, np.int32)
    dq = np.zeros(G, np.uint32)
    f = lambda a: _p(_c(a, np.float32))  # noqa: E731
    L.ur_oracle_preprocess(G, _p(means), _p(scales), f(fr.origin), f(fr.axis_lat), f(fr.axis_ele), f(fr.axis_ax), fr.W,
                           fr.width_mm, fr.near_mm, fr.col_offset, st.elev_margin, st.k_sigma, _p(count), _p(rmin),
                           _p(rmax), _p(dq))
    K = int(count.astype(np.int64).sum())
    counts, offsets = np.zeros(fr.W, np.int32), np.zeros(fr.W, np.int32)
    indices, depths = np.zeros(K, np.int32), np.zeros(K, np.float32)
    L.ur_oracle_emit_sort(G, fr.W, _p(count), _p(rmin), _p(rmax), _p(dq), _p(counts), _p(offsets), _p(indices),
                          _p(depths))
    return RayMap(count, rmin, rmax, dq, counts, offsets, indices, depths)


def _args(params, fr: RayFrame, rm: RayMap, st: RaySettings, dt):
    means, scales, quats, atten, sh = (_c(a, dt) for a in params)
    f = lambda a: _p(_c(a, np.float32))  # noqa: E731
    keep = (means, scales, quats, atten, sh)
    return keep, [means.shape[0], _p(means), _p(scales), _p(quats), _p(atten), _p(sh), fr.W, fr.H, f(fr.pixels),
                  f(fr.ray_o), f(fr.ray_d), f(fr.axis_ele), f(fr.delta), _p(rm.counts), _p(rm.offsets),
                  _p(rm.indices), _p(rm.sorted_depths), st.k_sigma, st.Rw, int(st.local_only), st.local_max]


def forward(params, fr: RayFrame, rm: RayMap, st: RaySettings, fp64=False, tc_frozen=None):
    """image (H, W) row-major flat; FP32 also returns aux (H*W, 3) = (T, num, den)."""
    L = lib()
    dt = np.float64 if fp64 else np.float32
    keep, a = _args(params, fr, rm, st, dt)
    image = np.zeros(fr.H * fr.W, dt)
    if fp64:
        tc = None if tc_frozen is None else _c(tc_frozen, np.float64)
        L.ur_oracle_forward_f64(*a, _p(tc), _p(image))
        return image
    aux = np.zeros((fr.H * fr.W, 3), np.float32)
    L.ur_oracle_forward_f32(*a, _p(image), _p(aux))
    del keep
    return image, aux


def window_centres(params, fr: RayFrame):
    L = lib()
    means, scales, quats = (_c(a, np.float64) for a in params[:3])
    tc = np.zeros((fr.W, means.shape[0]), np.float64)
    f = lambda a: _p(_c(a, np.float32))  # noqa: E731
    L.ur_oracle_window_centres(means.shape[0], _p(means), _p(scales), _p(quats), fr.W, f(fr.ray_o), f(fr.ray_d),
                               f(fr.axis_ele), f(fr.delta), _p(tc))
    return tc


def backward_f64(params, fr: RayFrame, rm: RayMap, st: RaySettings, d_image):
    L = lib()
    keep, a = _args(params, fr, rm, st, np.float64)
    G = keep[0].shape[0]
    out = [np.zeros((G, 3)), np.zeros((G, 3)), np.zeros((G, 4)), np.zeros(G), np.zeros((G, 4))]
    L.ur_oracle_backward_f64(*a, _p(_c(d_image, np.float32)), st.grad_threshold, *[_p(o) for o in out])
    return out
