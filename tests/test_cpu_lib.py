"""CPU checks of the C ABI library and the host logic (no compute calls: there is no GPU here)."""

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from us3d import _lib
from us3d.batch import FrameSpec, SweepSpec, frame_table, plan_layout

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "us3d.h"


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(us3d_\w+)\s*\(", text, flags=re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 15
    lib = _lib.lib()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in us3d.h but not exported"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures and header disagree"
    assert lib.us3d_version() >= 1


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof of every ABI struct, from a C program compiled against us3d.h, equal the ctypes mirror."""
    structs = {"us3d_frame": _lib.Frame, "us3d_sweep": _lib.Sweep, "us3d_stats": _lib.Stats, "us3d_plan": _lib.Plan,
               "us3d_ug_settings": _lib.UGSettings, "us3d_ur_settings": _lib.URSettings, "us3d_layout": _lib.Layout,
               "us3d_gp_settings": _lib.GPSettings}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "us3d.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _t in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in filter(None, out):
        c, f, v = line.split()
        got[(c, f)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _t in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, f"{cname}.{fname}"


def test_workspace_layout_is_aligned_and_monotone():
    lib = _lib.lib()
    out = _lib.Layout()
    assert lib.us3d_ug_workspace_layout(1000, 3000, 64, 12345, 700, out) == 0
    offs = [getattr(out, n) for n, _ in _lib.Layout._fields_]
    assert all(o % 256 == 0 for o in offs)
    assert offs == sorted(offs)
    assert out.total == lib.us3d_ug_workspace_bytes(1000, 3000, 64, 12345, 700)
    assert lib.us3d_ur_workspace_bytes(1000, 3000, 64, 12345, 700) > out.total  # 64 B records / partials
    # per-pair state is two bits per (frame, Gaussian) pair and their word scans; records are per visible pair
    assert out.vis_pair - out.pair_bits < 0.65 * 3000 + 12 * 64 + 6 * 256  # 0.625 B per pair, 12 B per tile
    assert out.tile_count - out.vis_pair >= 700 * 36
    big = lib.us3d_ug_workspace_bytes(1000, 3_000_000, 64, 12345, 700) - out.total
    assert big < 0.65 * (3_000_000 - 3000)


def _frame(H=64, W=48, sweep=0):
    v = lambda *a: torch.tensor(a, dtype=torch.float32)  # noqa: E731
    return FrameSpec(v(1, 2, 3), v(1, 0, 0), v(0, 1, 0), v(0, 0, 1), H, W, 0.2, 0.3, 1.5, (W - 1) / 2, sweep)


def test_plan_layout_numbering():
    frames = [_frame(sweep=0), _frame(32, 40, sweep=0), _frame(16, 16, sweep=1)]
    sweeps = [SweepSpec(0, 100), SweepSpec(100, 50)]
    lay = plan_layout(frames, sweeps, 150, 16)
    assert lay.tiles_x == [3, 3, 1] and lay.tiles_y == [4, 2, 1]
    assert lay.frame_tile_base == [0, 12, 18, 19]
    assert lay.frame_pixel_base == [0, 64 * 48, 64 * 48 + 32 * 40, 64 * 48 + 32 * 40 + 256]
    assert lay.n_pairs == 100 * 2 + 50 * 1
    rows = lay.sweep_rows
    assert rows[:, :4].tolist() == [[0, 100, 0, 2], [100, 50, 2, 1]]
    pair_base = rows[:, 4:6].copy().view(np.int64).reshape(-1).tolist()
    assert pair_base == [0, 200]
    ray = plan_layout(frames, sweeps, 150, None)
    assert ray.n_tiles == 48 + 40 + 16 and ray.frame_ray_base == [0, 48, 88, 104]
    with pytest.raises(ValueError):
        plan_layout([_frame(sweep=1), _frame(sweep=0)], sweeps, 150, 16)
    with pytest.raises(ValueError):
        plan_layout(frames, [SweepSpec(0, 100), SweepSpec(100, 51)], 150, 16)
    # the sweep table is ordered by g_begin with disjoint ranges; gaps are allowed (those Gaussians are in no frame)
    with pytest.raises(ValueError):
        plan_layout(frames, [SweepSpec(100, 50), SweepSpec(0, 100)], 150, 16)
    with pytest.raises(ValueError):
        plan_layout(frames, [SweepSpec(0, 100), SweepSpec(99, 50)], 150, 16)
    gap = plan_layout(frames, [SweepSpec(10, 80), SweepSpec(120, 20)], 150, 16)
    assert gap.n_pairs == 80 * 2 + 20 * 1


def test_chunks_of_a_step_keep_their_sweeps_in_order():
    """frame_chunks splits a step's frames into runs; each run keeps the sweeps its frames use, renumbered in the
    original order, so its sweep table is ordered by g_begin (Gaussians of the other sweeps are in no frame of the
    run)."""
    from us3d.train_step import frame_chunks

    frames = [_frame(sweep=0)] * 3 + [_frame(sweep=1)] * 2 + [_frame(sweep=2)]
    sweeps = [SweepSpec(0, 100), SweepSpec(100, 50), SweepSpec(150, 30)]
    out = frame_chunks(frames, sweeps, 2)
    assert [(a, b) for a, b, _f, _s in out] == [(0, 2), (2, 4), (4, 6)]
    assert [[f.sweep for f in fs] for _a, _b, fs, _s in out] == [[0, 0], [0, 1], [0, 1]]
    assert [[s.g_begin for s in ss] for _a, _b, _f, ss in out] == [[0], [0, 100], [100, 150]]
    for _a, _b, fs, ss in out:
        plan_layout(fs, ss, 180, 16)
    assert [(a, b) for a, b, _f, _s in frame_chunks(frames, sweeps, None)] == [(0, 6)]


def test_frame_table_bits():
    frames = [_frame(), _frame(32, 40)]
    frames[1].geometry = 1
    lay = plan_layout(frames, [SweepSpec(0, 10)], 10, 16)
    tab = frame_table(frames, lay, torch.device("cpu"))
    n = ctypes.sizeof(_lib.Frame)
    assert n == 100 and tab.shape == (2, n // 4) and tab.dtype == torch.int32
    buf = tab.numpy().tobytes()
    f0 = _lib.Frame.from_buffer_copy(buf[:n])
    f1 = _lib.Frame.from_buffer_copy(buf[n:2 * n])
    assert (f1.height, f1.width, f1.tiles_x, f1.tiles_y, f1.tile_base, f1.pixel_base) == (32, 40, 3, 2, 12, 64 * 48)
    assert abs(f1.delta_ax - 0.3) < 1e-7 and abs(f1.near_mm - 1.5) < 1e-7 and list(f1.origin) == [1.0, 2.0, 3.0]
    assert (f0.geometry, f1.geometry) == (0, 1)


def test_render_without_gpu_fails_loudly():
    from us3d.ultragauss import UltraGaussRenderConfig, render_frames

    fr = _frame()
    cfg = UltraGaussRenderConfig(H=fr.height, W=fr.width, lateral_spacing_mm_per_pix=0.2,
                                 axial_spacing_mm_per_pix=0.3, axial_near_mm=1.5)
    z = torch.zeros
    with pytest.raises(RuntimeError, match="GPU only"):
        render_frames(z(4, 3), z(4, 6), z(4), z(4), None, [fr], cfg)


def test_config_validation_mirrors_reference():
    from us3d.ultragauss import UltraGaussRenderConfig
    from us3d.ultragray import UltragrayRenderConfig

    with pytest.raises(ValueError):
        UltraGaussRenderConfig(H=0, W=4, lateral_spacing_mm_per_pix=1, axial_spacing_mm_per_pix=1, axial_near_mm=0)
    with pytest.raises(ValueError):
        UltraGaussRenderConfig(H=4, W=4, lateral_spacing_mm_per_pix=1, axial_spacing_mm_per_pix=1, axial_near_mm=0,
                               diag_beta=0.0)
    assert UltraGaussRenderConfig(H=4, W=5, lateral_spacing_mm_per_pix=1, axial_spacing_mm_per_pix=1,
                                  axial_near_mm=0).lateral_col_offset == 2.0
    with pytest.raises(ValueError):  # the splat kernels stage 64, 128 or 256 Gaussians per LDS batch
        UltraGaussRenderConfig(H=4, W=4, lateral_spacing_mm_per_pix=1, axial_spacing_mm_per_pix=1, axial_near_mm=0,
                               bwd_lds_batch=96)
    st = UltraGaussRenderConfig(H=4, W=4, lateral_spacing_mm_per_pix=1, axial_spacing_mm_per_pix=1,
                                axial_near_mm=0, fwd_lds_batch=128, bwd_lds_batch=64).settings()
    assert (st.fwd_lds_batch, st.bwd_lds_batch) == (128, 64)
    with pytest.raises(ValueError):
        UltragrayRenderConfig(image_width=4, image_height=4, width_mm=0.0, near_mm=0.0)
    with pytest.raises(ValueError):
        UltragrayRenderConfig(image_width=4, image_height=4, width_mm=1.0, near_mm=0.0, ops_max_offset_mm=-1.0)
    c = UltragrayRenderConfig(image_width=128, image_height=4, width_mm=102.4, near_mm=0.0)
    assert c.pixel_spacing() == float(np.float32(102.4) / np.float32(128))


def test_autotune_sample_renumbers_sweeps_and_slices_targets():
    from us3d.autotune import _sample

    sizes = [(4, 3), (4, 3), (2, 5), (2, 5), (6, 1)]
    frames = [_frame(H=h, W=w, sweep=s) for (h, w), s in zip(sizes, [0, 0, 1, 1, 2])]
    n = [h * w for h, w in sizes]
    target = torch.cat([torch.full((k,), float(i)) for i, k in enumerate(n)])
    sweeps = [SweepSpec(0, 10), SweepSpec(10, 5), SweepSpec(15, 7)]
    sub, sub_sweeps, tgt = _sample(frames, sweeps, target, 3)
    picked = [0, 1, 3]  # int(i * 5 / 3) for i in 0..2
    assert [f.sweep for f in sub] == [0, 0, 1]
    assert sub_sweeps == [sweeps[0], sweeps[1]]
    assert torch.equal(tgt, torch.cat([torch.full((n[i],), float(i)) for i in picked]))


def test_ray_renderer_rejects_fan_frames():
    """Fan frames (geometry 1) are a slice-splat geometry; the ray renderer refuses them before any device work."""
    from us3d.ultragray import URBatch

    fr = _frame()
    fr.geometry = 1
    with pytest.raises(ValueError):
        URBatch([fr], [SweepSpec(0, 10)], 10, _lib.URSettings(), torch.device("cpu"))


def test_render_static_refuses_an_oversized_batch_on_the_host():
    """The no-allocation render checks that the batch's workspace fits its static scratch before touching the
    device (US3D_ERR_CAPACITY); plan pointers are placeholders that are never dereferenced."""
    p = _lib.Plan()
    p.n_gauss, p.n_frames, p.n_sweeps, p.n_tiles, p.n_pairs, p.n_pixels = 1 << 20, 512, 1, 1 << 18, 1 << 29, 1 << 27
    dummy = ctypes.create_string_buffer(64)
    for f in ("frames", "frame_tile_base", "frame_pixel_base", "sweeps"):
        setattr(p, f, ctypes.addressof(dummy))
    st = _lib.UGSettings(chi2_threshold=7.815, diag_beta=1e-4, tile_size=16)
    L = _lib.lib()
    assert L.us3d_ug_render_static(p, None, None, None, None, None, st, None, None, None, None, None, None) == 3
    assert b"static scratch" in L.us3d_last_error()
