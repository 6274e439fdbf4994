"""ctypes binding of libus3d.so (the C ABI declared in include/us3d.h).

torch is imported first on purpose: the ROCm torch wheel ships its own libamdhip64.so.7, and loading it
before libus3d makes the library bind to that same HIP runtime (same SONAME), so torch's streams and
device pointers are valid inside the library.

There is no fallback: if the library is missing, importing anything that renders raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (must precede loading libus3d, see module docstring)

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libus3d.so"

US3D_OK = 0


class Frame(ctypes.Structure):
    _fields_ = [
        ("origin", ctypes.c_float * 3),
        ("axis_lat", ctypes.c_float * 3),
        ("axis_ax", ctypes.c_float * 3),
        ("axis_ele", ctypes.c_float * 3),
        ("delta_lat", ctypes.c_float),
        ("delta_ax", ctypes.c_float),
        ("near_mm", ctypes.c_float),
        ("col_offset", ctypes.c_float),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("tiles_x", ctypes.c_int32),
        ("tiles_y", ctypes.c_int32),
        ("tile_base", ctypes.c_int32),
        ("pixel_base", ctypes.c_int32),
        ("sweep", ctypes.c_int32),
        ("ray_base", ctypes.c_int32),
        ("geometry", ctypes.c_int32),
    ]


class Sweep(ctypes.Structure):
    _fields_ = [
        ("g_begin", ctypes.c_int32),
        ("g_count", ctypes.c_int32),
        ("f_begin", ctypes.c_int32),
        ("f_count", ctypes.c_int32),
        ("pair_base", ctypes.c_int64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_instances", ctypes.c_int64),
        ("max_segment", ctypes.c_int64),
        ("overflow", ctypes.c_int32),
        ("n_sorted_big", ctypes.c_int32),
        ("overflow_any", ctypes.c_int32),
        ("reserved_i32", ctypes.c_int32),
        ("n_visible", ctypes.c_int64),
        ("reserved", ctypes.c_int64 * 3),
    ]


class Plan(ctypes.Structure):
    _fields_ = [
        ("n_gauss", ctypes.c_int32),
        ("n_frames", ctypes.c_int32),
        ("n_sweeps", ctypes.c_int32),
        ("n_tiles", ctypes.c_int32),
        ("n_pairs", ctypes.c_int64),
        ("n_pixels", ctypes.c_int64),
        ("k_capacity", ctypes.c_int64),
        ("v_capacity", ctypes.c_int64),
        ("frames", ctypes.c_void_p),
        ("frame_tile_base", ctypes.c_void_p),
        ("frame_pixel_base", ctypes.c_void_p),
        ("sweeps", ctypes.c_void_p),
        ("workspace", ctypes.c_void_p),
        ("workspace_bytes", ctypes.c_int64),
    ]


class UGSettings(ctypes.Structure):
    _fields_ = [
        ("chi2_threshold", ctypes.c_float),
        ("diag_beta", ctypes.c_float),
        ("bg_alpha", ctypes.c_float),
        ("bg_intensity", ctypes.c_float),
        ("alpha_min", ctypes.c_float),
        ("sort_by_depth", ctypes.c_int32),
        ("tile_size", ctypes.c_int32),
        ("alpha_excludes_bg", ctypes.c_int32),
        ("raw_activations", ctypes.c_int32),
        ("fwd_lds_batch", ctypes.c_int32),
        ("bwd_lds_batch", ctypes.c_int32),
        ("work_queue", ctypes.c_int32),
    ]


class URSettings(ctypes.Structure):
    _fields_ = [
        ("k_sigma", ctypes.c_float),
        ("quadrature_window_radius_mm", ctypes.c_float),
        ("elev_cutoff_margin_mm", ctypes.c_float),
        ("local_transmission_max_window_mm", ctypes.c_float),
        ("grad_threshold", ctypes.c_float),
        ("local_transmission_window_only", ctypes.c_int32),
        ("scales_are_log", ctypes.c_int32),
        ("reserved1", ctypes.c_int32),
    ]


class GPSettings(ctypes.Structure):
    _fields_ = [
        ("k_sigma", ctypes.c_float),
        ("sigma_ele", ctypes.c_float),
        ("det_eps", ctypes.c_float),
        ("weight_min", ctypes.c_float),
        ("tile_size", ctypes.c_int32),
        ("raw_activations", ctypes.c_int32),
        ("work_queue", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


_LAYOUT_FIELDS = [
    "stats", "queue", "pair_bits", "word_cnt", "word_off", "fm_bits", "fm_cnt", "fm_off", "vis_pair",
    "pair_count", "pair_off", "pair_rect", "pair_dk", "pair_g", "vis_list", "tile_count", "tile_off",
    "tile_cursor", "seg_key", "seg_key_tmp", "records", "partial", "scan_tmp",
    "inst_mask", "pair_sect", "total",
]


class Layout(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in _LAYOUT_FIELDS]


# Every symbol include/us3d.h declares, with its ctypes signature.
_VP = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F = ctypes.c_float
_PP = ctypes.POINTER(Plan)

SIGNATURES = {
    "us3d_ug_workspace_bytes": (_I64, [_I64, _I64, _I64, _I64, _I64]),
    "us3d_ug_workspace_layout": (ctypes.c_int, [_I64, _I64, _I64, _I64, _I64, ctypes.POINTER(Layout)]),
    "us3d_ug_map": (ctypes.c_int, [_PP, _VP, _VP, _VP, _VP, ctypes.POINTER(UGSettings), _VP]),
    "us3d_ug_forward": (ctypes.c_int, [_PP, _VP, ctypes.POINTER(UGSettings), _VP, _VP, _VP, _VP, _VP]),
    "us3d_ug_backward": (
        ctypes.c_int,
        [_PP, _VP, _VP, ctypes.POINTER(UGSettings), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "us3d_ug_loss_tmp_bytes": (_I64, [_PP]),
    "us3d_ug_forward_loss": (
        ctypes.c_int,
        [_PP, _VP, ctypes.POINTER(UGSettings), _VP, _VP, _VP, _VP, _VP, _F, _F, _F, _VP, _VP, _VP],
    ),
    "us3d_ug_backward_loss": (
        ctypes.c_int,
        [_PP, _VP, _VP, ctypes.POINTER(UGSettings), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "us3d_ug_pixel_lists": (ctypes.c_int, [_PP, _VP, ctypes.POINTER(UGSettings), ctypes.c_int, _VP, _VP, _VP, _VP]),
    "us3d_ug_render_static": (
        ctypes.c_int,
        [_PP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(UGSettings), _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "us3d_ur_workspace_bytes": (_I64, [_I64, _I64, _I64, _I64, _I64]),
    "us3d_ur_workspace_layout": (ctypes.c_int, [_I64, _I64, _I64, _I64, _I64, ctypes.POINTER(Layout)]),
    "us3d_ur_map": (ctypes.c_int, [_PP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(URSettings), _VP]),
    "us3d_ur_forward": (ctypes.c_int, [_PP, _VP, _VP, _VP, _VP, ctypes.POINTER(URSettings), _VP, _VP, _VP]),
    "us3d_ur_backward": (
        ctypes.c_int,
        [_PP, _VP, _VP, _VP, _VP, ctypes.POINTER(URSettings), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "us3d_gp_workspace_bytes": (_I64, [_I64, _I64, _I64, _I64, _I64]),
    "us3d_gp_workspace_layout": (ctypes.c_int, [_I64, _I64, _I64, _I64, _I64, ctypes.POINTER(Layout)]),
    "us3d_gp_map": (ctypes.c_int, [_PP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(GPSettings), _VP]),
    "us3d_gp_forward": (ctypes.c_int, [_PP, ctypes.POINTER(GPSettings), _VP, _VP]),
    "us3d_gp_backward": (ctypes.c_int, [_PP, ctypes.POINTER(GPSettings), _VP, _VP, _VP, _VP, _VP, _VP, _VP]),
    "us3d_gp_loss_tmp_bytes": (_I64, [_PP]),
    "us3d_gp_forward_loss": (ctypes.c_int, [_PP, ctypes.POINTER(GPSettings), _VP, _VP, _F, _F, _F, _VP, _VP, _VP]),
    "us3d_gp_backward_loss": (
        ctypes.c_int,
        [_PP, ctypes.POINTER(GPSettings), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP],
    ),
    "us3d_scan_tmp_bytes": (_I64, [_I64]),
    "us3d_scan_i32": (ctypes.c_int, [_VP, _VP, _I64, _VP, _VP, _VP]),
    "us3d_image_loss_tmp_bytes": (_I64, [_I32]),
    "us3d_image_loss": (ctypes.c_int, [_VP, _VP, _VP, _I32, _F, _F, _F, _VP, _VP, _VP, _VP]),
    "us3d_adam_step": (ctypes.c_int, [_VP, _VP, _VP, _VP, _I64, _VP, _VP, _F, _F, _F, _F, _VP]),
    "us3d_adam_advance": (ctypes.c_int, [_VP, _VP, _VP]),
    "us3d_fill_i32": (ctypes.c_int, [_VP, _I64, _I32, _VP]),
    "us3d_stage_timing": (ctypes.c_int, [ctypes.c_int]),
    "us3d_stage_timing_read": (ctypes.c_int, [ctypes.c_char_p, _I64]),
    "us3d_last_error": (ctypes.c_char_p, []),
    "us3d_version": (ctypes.c_int, []),
}

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """Load libus3d.so once; raise loudly if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(hipcc --offload-arch=gfx950). There is no CPU fallback."
            )
        handle = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != US3D_OK:
        msg = lib().us3d_last_error().decode("utf-8", "replace")
        raise RuntimeError(f"[us3d] {what} failed (code {rc}): {msg}")


def ptr(t) -> int:
    """Device pointer of a tensor (0 for None)."""
    if t is None:
        return 0
    return int(t.data_ptr())


def stage_timing(enable: bool) -> None:
    check(lib().us3d_stage_timing(1 if enable else 0), "us3d_stage_timing")


def stage_timing_read() -> dict:
    """{kernel: (launches, total_ms)} since the last read (synchronises on the recorded events)."""
    L = lib()
    need = L.us3d_stage_timing_read(None, 0)
    buf = ctypes.create_string_buffer(max(int(need), 1) + 64)
    check(L.us3d_stage_timing_read(buf, len(buf)), "us3d_stage_timing_read")
    out = {}
    for line in buf.value.decode().splitlines():
        name, n, ms = line.split()
        out[name] = (int(n), float(ms))
    return out


def stream_ptr(device: torch.device | None = None) -> int:
    return int(torch.cuda.current_stream(device).cuda_stream)
