#!/usr/bin/env python
"""Forward+backward frames/s of the differentiable slice splat on one MI355X.

Headline workload: 200k anisotropic Gaussians in a 120 mm cube; 256 frames of 512x512 at 0.2 mm along a 150 mm
freehand sweep; loss = L1 + 0.2 (1 - global SSIM) against the same sweep rendered from a second seeded scene; Adam
step (lr 1e-3) over means, raw precision factors and opacity/intensity logits. One step = all 256 frames forward +
backward + Adam. "batched" renders the whole sweep as one batch (one captured launch sequence per step); "plain"
renders it frame by frame through the same API with torch autograd and torch.optim.Adam. `value` is the batched
frames/s; the plain rate, the roofline of the dominant kernel and the CPU baseline are reported beside it, and the
other workloads (UltraG-Ray, multi-sweep, fan probe, GaussianPile) as nested objects of the same line.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload slice|ray|multisweep|fan|gaussianpile]
"""

from __future__ import annotations

import argparse
import copy
import json
import os
import platform
import re
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "ultrasound-3d-reconstruction_amd"), str(ROOT)]

HBM_PEAK_GBS = 8000.0
# LDS: 128 bytes per clock per CU (the CDNA3 LDS data path), 256 CUs at 2.4 GHz
LDS_PEAK_GBS = 128 * 256 * 2.4
CLOCK_GHZ = 2.4

METRIC = "fwd+bwd frames/s per sweep, batched vs plain on 1 GPU; % of HBM roofline"


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=256)
    ap.add_argument("--gaussians", type=int, default=200_000)
    ap.add_argument("--tile-size", type=int, default=16)
    ap.add_argument("--work-queue", type=int, choices=(0, 1), default=1,
                    help="slice: 1 runs the forward and the adjoint as persistent workgroups over per-XCD (frame, tile) "
                         "queues, 0 as one workgroup per tile (both are timed in the stage-timing leg either way)")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="headline only: skip the nested ray / multisweep / fan / gaussianpile measurements")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU-baseline runs (median), after one warm-up")
    ap.add_argument("--cpu-sample-rows", type=int, default=2)
    ap.add_argument("--timing-steps", type=int, default=2)
    ap.add_argument("--eager", action="store_true", help="batched step without hipGraph capture")
    ap.add_argument("--separate-loss", action="store_true",
                    help="training steps with the image loss as its own kernels instead of fused into the splats")
    ap.add_argument("--workload", choices=("slice", "ray", "multisweep", "fan", "gaussianpile"), default="slice",
                    help="slice: the slice splat training step (the headline metric); ray: UltraG-Ray B-mode "
                         "forward+backward, 200k Gaussians, --ray-frames frames of 128 scanlines x 1024 samples; "
                         "multisweep: 8 sweeps of uneven size in one captured training step, tile size from the "
                         "autotuner; fan: --fan-frames frames of a 1000-frame fan-probe sweep over 2M Gaussians, trained "
                         "in chunks of --fan-chunk frames inside one step; "
                         "gaussianpile: the GaussianPile us_focus splat on the config-2 scene and sweep")
    ap.add_argument("--ms-scale", type=float, default=1.0,
                    help="multisweep: scale factor on the sweeps' Gaussian and frame counts")
    ap.add_argument("--ray-frames", type=int, default=128)
    ap.add_argument("--fan-frames", type=int, default=1000, help="fan: frames of the sweep in one step (all 1000)")
    ap.add_argument("--fan-chunk", type=int, default=100,
                    help="fan: frames rendered per chunk inside the step (bounds the workspace; 0 = one batch)")
    ap.add_argument("--fan-gaussians", type=int, default=2_000_000)
    ap.add_argument("--fan-cpu-rows", type=int, default=2)
    ap.add_argument("--cpu-sample-rays", type=int, default=8)
    ap.add_argument("--rocprof-summary", nargs=2, metavar=("MODE", "DIR"),
                    help="print a markdown summary of a rocprofv3 output directory: MODE 'trace' (per-kernel "
                         "durations of --kernel-trace) or 'pmc' (per-dispatch means of --pmc counters), then exit")
    ap.add_argument("--rocprof-report", nargs="+", metavar="ARG",
                    help="KERNEL TRACE_DIR PMC_DIR [PMC_DIR ...]: print the body of the kernel's profile summary (its "
                         "average duration in the trace, traffic, VALU, LDS figures and counter sums), then exit")
    ap.add_argument("--rocprof-derive", nargs="+", metavar="ARG",
                    help="KERNEL DIR [DIR ...]: print the derived figures of one kernel (traffic, LDS bytes, VALU "
                         "issue and active lanes) from rocprofv3 --pmc directories of the same command, then exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (reference dispatch for HBM counters)")
    return ap.parse_args(argv)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def ssim_global(x, y):
    """simple_ssim of pipeline_god_file.py:628-637, per frame over (F, H, W)."""
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    mx, my = x.mean(dim=(1, 2)), y.mean(dim=(1, 2))
    xc, yc = x - mx[:, None, None], y - my[:, None, None]
    vx, vy, vxy = (xc * xc).mean(dim=(1, 2)), (yc * yc).mean(dim=(1, 2)), (xc * yc).mean(dim=(1, 2))
    return ((2 * mx * my + c1) * (2 * vxy + c2)) / ((mx ** 2 + my ** 2 + c1) * (vx + vy + c2))


def frame_losses(pred, gt):
    """Per-frame loss of image_loss (pipeline_god_file.py:606-617) with w_L1 = 1, w_SSIM = 0.2."""
    l1 = (pred - gt).abs().mean(dim=(1, 2))
    return l1 + 0.2 * (1.0 - ssim_global(pred, gt))


def host_cpu() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        m = re.search(r"Model name:\s*(.+)", out)
        return m.group(1).strip() if m else platform.processor()
    except OSError:
        return platform.processor() or "unknown"


# ---------------------------------------------------------------- CPU baseline (the oracle's fallback restatement)

def _fallback_fwd_bwd(scene, target, pix, rows, width):
    """One forward + autograd backward of the restated PyTorch fallback renderer (pipeline_god_file.py:556-603)
    over `pix`, L1 against the target scene's render of the same pixels; returns the seconds it took."""
    from oracle import fallback as ofb

    ts = [t.clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    t0 = time.perf_counter()
    img, _ = ofb.render(*ts, pix, rows, width)
    (img - target).abs().mean().backward()
    return time.perf_counter() - t0


def cpu_baseline(n_rows: int, runs: int):
    """The oracle's restatement of the reference's PyTorch fallback renderer (every pixel against every Gaussian,
    2048 x 512 chunks, autograd backward), on all host threads, one warm-up then the median of `runs` timed runs:
      - the benchmark workload, bounded: n_rows full rows of the middle frame of the 256-frame sweep against all
        200k Gaussians; value = sampled fraction of a 512x512 frame / median time (the fallback's cost is pixels x
        Gaussians, so this is the rate of the whole workload);
      - the reference's CPU-sized case complete (4,096 Gaussians, 16 frames of 128x128): frames/s, and the per-pair
        rate it gives with the benchmark workload's frames/s extrapolated from it."""
    import torch

    from oracle import fallback as ofb
    from us3d.scenes import config1, config2, linear_pixels

    torch.set_num_threads(os.cpu_count() or 1)
    scene, target, frames = config2()
    fr = frames[len(frames) // 2]
    pix = linear_pixels(fr).reshape(fr.height, fr.width, 3)
    r0 = fr.height // 2 - n_rows // 2
    pix = pix[r0:r0 + n_rows].reshape(-1, 3).contiguous()
    with torch.no_grad():
        gt, _ = ofb.render(target.means, target.tril, target.opacity, target.intensity, pix, n_rows, fr.width)
    _fallback_fwd_bwd(scene, gt, pix, n_rows, fr.width)
    t2 = statistics.median(_fallback_fwd_bwd(scene, gt, pix, n_rows, fr.width) for _ in range(runs))
    G2 = scene.means.shape[0]
    pairs2 = pix.shape[0] * G2

    s1, t1s, f1 = config1()
    pix1 = torch.cat([linear_pixels(f) for f in f1]).contiguous()
    H1, W1 = f1[0].height, f1[0].width
    with torch.no_grad():
        gt1, _ = ofb.render(t1s.means, t1s.tril, t1s.opacity, t1s.intensity, pix1, len(f1) * H1, W1)
    _fallback_fwd_bwd(s1, gt1, pix1, len(f1) * H1, W1)
    times1 = [_fallback_fwd_bwd(s1, gt1, pix1, len(f1) * H1, W1) for _ in range(runs)]
    t1 = statistics.median(times1)
    pairs1 = pix1.shape[0] * s1.means.shape[0]
    rate1 = pairs1 / t1
    full2 = fr.height * fr.width * G2
    return {
        "value": n_rows / fr.height / t2, "unit": "frames/s", "cores": torch.get_num_threads(), "kind": "port",
        "sample": f"{n_rows} rows x {fr.width} pixels of frame {len(frames) // 2} of the benchmark sweep against all "
                  f"{G2} Gaussians: the restated PyTorch fallback renderer, forward + autograd backward, one warm-up "
                  f"then the median of {runs} runs ({t2:.2f} s, {pairs2 / t2 / 1e6:.0f} M pair evaluations/s); "
                  f"frames/s = sampled fraction of a 512x512 frame / time; host CPU: {host_cpu()}",
        "config1": {"frames_per_s": len(f1) / t1, "median_s": t1, "runs_s": [round(t, 3) for t in times1],
                    "workload": f"{s1.means.shape[0]} Gaussians, {len(f1)} frames {H1}x{W1}, complete",
                    "pairs_per_s": rate1,
                    "benchmark_frames_per_s_extrapolated": rate1 / full2},
    }


def parity_config2_frame(dev) -> dict:
    """Checker run beside the CPU baseline: the middle frame of the benchmark sweep (200k Gaussians, 512x512 on the
    frame's own grid, tile 16, the bench's image model) rendered and differentiated on the GPU through the C ABI,
    against the oracle's FP32 (canonical order) and FP64 passes on the same inputs: relative L2 error of the image,
    the alpha and every gradient group, for dL = <dI, image> + <dA, alpha> with seeded dI, dA."""
    import numpy as np
    import torch

    from oracle import ug as oug
    from us3d.batch import CapacityHint, SweepSpec
    from us3d.scenes import config2, linear_pixels
    from us3d.ultragauss import UltraGaussRenderConfig, render_frames

    def rel(a, b):
        a = np.asarray(a, np.float64).reshape(-1)
        b = np.asarray(b, np.float64).reshape(-1)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    t0 = time.perf_counter()
    scene, _target, frames = config2()
    fr = frames[len(frames) // 2]
    P = fr.height * fr.width
    rng = np.random.default_rng(9)
    dI = rng.normal(size=P).astype(np.float32) / P
    dA = rng.normal(size=P).astype(np.float32) / P
    cfg = UltraGaussRenderConfig(H=fr.height, W=fr.width, lateral_spacing_mm_per_pix=fr.delta_lat,
                                 axial_spacing_mm_per_pix=fr.delta_ax, axial_near_mm=fr.near_mm,
                                 lateral_col_offset=fr.col_offset, fixed_bg_alpha=1e-6)
    ts = [t.to(dev).clone().requires_grad_(True) for t in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    image, alpha, _depth = render_frames(*ts, None, [fr], cfg, sweeps=[SweepSpec(0, scene.means.shape[0])],
                                         hint=CapacityHint())
    ((image * torch.from_numpy(dI).to(dev)).sum() + (alpha * torch.from_numpy(dA).to(dev)).sum()).backward()
    torch.cuda.synchronize()
    sem = oug.Semantics()
    geom = oug.FrameGeom.from_spec(fr)
    pix = linear_pixels(fr).numpy()
    m, t = scene.means.numpy(), scene.tril.numpy()
    mr = oug.map_frame(m, t, geom, chi2=sem.chi2, beta=sem.beta, sort_by_depth=True, ts=16)
    lists = oug.pixel_lists(m, t, pix, geom, mr, chi2=sem.chi2, beta=sem.beta, ts=16)
    plane = oug.plane_form(m, t, geom)
    args = (m, t, scene.opacity.numpy(), scene.intensity.numpy(), pix, geom, lists, sem)
    img32, a32, _d, _s = oug.forward(*args, plane=plane)
    img64, a64, _d, _s = oug.forward(*args, fp64=True, plane=plane)
    refs = oug.backward_f64(*args, dI, dA, None, gate=True, plane=plane)
    got_i, got_a = image.detach().cpu().numpy(), alpha.detach().cpu().numpy()
    out = {"workload": f"frame {len(frames) // 2} of the benchmark sweep, {scene.means.shape[0]} Gaussians, "
                       f"{fr.height}x{fr.width}, {int(mr.keys.size)} instances",
           "image_vs_fp32": rel(got_i, img32), "image_vs_fp64": rel(got_i, img64),
           "alpha_vs_fp32": rel(got_a, a32), "alpha_vs_fp64": rel(got_a, a64)}
    for name, tt, r in zip(("means", "tril", "opacity", "intensity"), ts, refs):
        out[f"d{name}_vs_fp64"] = rel(tt.grad.detach().cpu().numpy(), r)
    out["bars"] = {"image_alpha": 1e-5, "gradients": 1e-4}
    out["ok"] = bool(max(out["image_vs_fp32"], out["image_vs_fp64"], out["alpha_vs_fp32"], out["alpha_vs_fp64"]) < 1e-5
                     and max(out[f"d{n}_vs_fp64"] for n in ("means", "tril", "opacity", "intensity")) < 1e-4)
    out["wall_s"] = round(time.perf_counter() - t0, 1)
    return out


# ---------------------------------------------------------------- rocprof summaries and the figures read from them

def _short(name: str) -> str:
    for k in ("void us3d::", "us3d::", "(anonymous namespace)::"):
        name = name.replace(k, "")
    return name.split("(")[0][:70]


def _pmc_rows(directory: str):
    import collections
    import glob
    import sqlite3

    agg = collections.defaultdict(list)
    for db in glob.glob(f"{directory}/**/*.db", recursive=True):
        con = sqlite3.connect(db)
        for n, c, v in con.execute("select name, counter_name, counter_value from pmc_events"):
            agg[(_short(n), c)].append(v)
    return agg


def rocprof_summary(mode: str, directory: str) -> None:
    """Markdown tables from the SQLite databases rocprofv3 writes (kernel durations or PMC counter means)."""
    import collections
    import glob
    import sqlite3

    if mode == "trace":
        agg = collections.defaultdict(list)
        for db in glob.glob(f"{directory}/**/*.db", recursive=True):
            con = sqlite3.connect(db)
            for n, t0, t1 in con.execute("select name, start, end from kernels"):
                agg[_short(n)].append(t1 - t0)
        tot = sum(sum(v) for v in agg.values()) or 1
        print("| kernel | calls | total ms | avg ms | min ms | max ms | % |")
        print("|---|---|---|---|---|---|---|")
        for n, v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:20]:
            print(f"| {n} | {len(v)} | {sum(v) / 1e6:.3f} | {sum(v) / len(v) / 1e6:.4f} | {min(v) / 1e6:.4f} | "
                  f"{max(v) / 1e6:.4f} | {100 * sum(v) / tot:.1f} |")
    else:
        agg = _pmc_rows(directory)
        print("| kernel | counter | rows (dispatches x counter instances) | mean per row |")
        print("|---|---|---|---|")
        for (n, c), v in sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:60]:
            print(f"| {n} | {c} | {len(v)} | {sum(v) / len(v):.1f} |")


def rocprof_derive(kernel: str, dirs, quiet: bool = False):
    """Per-dispatch figures of one kernel from --pmc runs of the same command. SQ counters are summed over their
    instances per dispatch (rows / dispatches instances each); dispatches = the rows of a GRBM_GUI_ACTIVE or
    FETCH_SIZE counter divided by its instances. Prints `key: value` lines that bench.py reads back:
      fetch_kib, write_kib            FETCH_SIZE / WRITE_SIZE (KiB per dispatch, raw)
      valu_insts, lds_insts           wave instructions per dispatch
      valu_issue_frac                 VALU instructions x 4 cycles / (1024 SIMDs x active cycles)
      valu_active_threads             SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU: lanes busy per VALU instruction
      lds_bytes                       (SQ_INSTS_LDS_LOAD_BANDWIDTH + _STORE_BANDWIDTH) x 64 bytes per dispatch
      lds_bank_conflict_frac          SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE
      active_cycles                   GRBM_GUI_ACTIVE (mean over XCDs)"""
    per = {}
    dispatches = None
    for d in dirs:
        for (n, c), v in _pmc_rows(d).items():
            if n != kernel:
                continue
            per[c] = v
    if "GRBM_GUI_ACTIVE" in per:
        dispatches = len(per["GRBM_GUI_ACTIVE"]) // 8
    elif "FETCH_SIZE" in per:
        dispatches = len(per["FETCH_SIZE"])
    if not dispatches:
        raise SystemExit(f"no dispatch count for {kernel}")

    def total(c):
        return sum(per[c]) / dispatches if c in per else None

    out = {"dispatches": dispatches}
    if "FETCH_SIZE" in per:
        out["fetch_kib"] = total("FETCH_SIZE")
    if "WRITE_SIZE" in per:
        out["write_kib"] = total("WRITE_SIZE")
    if "GRBM_GUI_ACTIVE" in per:
        out["active_cycles"] = sum(per["GRBM_GUI_ACTIVE"]) / len(per["GRBM_GUI_ACTIVE"])
    if "SQ_INSTS_VALU" in per:
        out["valu_insts"] = total("SQ_INSTS_VALU")
        if "active_cycles" in out:
            out["valu_issue_frac"] = out["valu_insts"] * 4 / (1024 * out["active_cycles"])
    if "SQ_INSTS_LDS" in per:
        out["lds_insts"] = total("SQ_INSTS_LDS")
    if "SQ_THREAD_CYCLES_VALU" in per and "SQ_ACTIVE_INST_VALU" in per:
        out["valu_active_threads"] = total("SQ_THREAD_CYCLES_VALU") / total("SQ_ACTIVE_INST_VALU")
    if "SQ_INSTS_LDS_LOAD_BANDWIDTH" in per and "SQ_INSTS_LDS_STORE_BANDWIDTH" in per:
        out["lds_bytes"] = 64 * (total("SQ_INSTS_LDS_LOAD_BANDWIDTH") + total("SQ_INSTS_LDS_STORE_BANDWIDTH"))
    if "SQ_LDS_BANK_CONFLICT" in per and "SQ_LDS_IDX_ACTIVE" in per:
        out["lds_bank_conflict_frac"] = total("SQ_LDS_BANK_CONFLICT") / max(total("SQ_LDS_IDX_ACTIVE"), 1.0)
    if not quiet:
        for k, v in out.items():
            print(f"{k}: {v:.6g}" if isinstance(v, float) else f"{k}: {v}")
    return out, {c: (len(v), sum(v) / dispatches) for c, v in per.items()}


def rocprof_report(kernel: str, trace_dir: str, pmc_dirs) -> None:
    """The body of a kernel's profile summary (profiles/<kernel>_<workload>_pmc.md): its average duration in the
    kernel trace, the figures bench.py reads back (HBM traffic scaled by the 1 GiB copy calibration, VALU issue and
    lanes, LDS bytes and bank conflicts) and the per-dispatch counter sums."""
    import glob
    import sqlite3

    durs = []
    for db in glob.glob(f"{trace_dir}/**/*.db", recursive=True):
        con = sqlite3.connect(db)
        durs += [t1 - t0 for n, t0, t1 in con.execute("select name, start, end from kernels") if _short(n) == kernel]
    if not durs:
        raise SystemExit(f"{kernel} is not in the kernel trace")
    avg_ms = sum(durs) / len(durs) / 1e6
    cal = (ROOT / "profiles" / "copy_1GiB_calibration_pmc.md").read_text()
    m = re.search(r"FETCH_SIZE scale = ([0-9.]+), WRITE_SIZE scale = ([0-9.]+)", cal)
    sf, sw = float(m.group(1)), float(m.group(2))
    out, rows = rocprof_derive(kernel, pmc_dirs, quiet=True)
    print(f"Per dispatch ({out['dispatches']} dispatches; kernel trace: {len(durs)} launches, average {avg_ms:.4f} ms):")
    print()
    fig = {}
    if "fetch_kib" in out and "write_kib" in out:
        fig["fetch_bytes_per_launch"] = out["fetch_kib"] * 1024 * sf
        fig["write_bytes_per_launch"] = out["write_kib"] * 1024 * sw
        fig["traffic_bytes_per_launch"] = fig["fetch_bytes_per_launch"] + fig["write_bytes_per_launch"]
    fig["avg_launch_ms"] = avg_ms
    for k in ("valu_issue_frac", "valu_active_threads", "lds_bytes", "lds_bank_conflict_frac"):
        if k in out:
            fig[k] = out[k]
    if "lds_bytes" in out:
        fig["lds_GBs"] = out["lds_bytes"] / (avg_ms * 1e-3) / 1e9
    for k in ("traffic_bytes_per_launch", "fetch_bytes_per_launch", "write_bytes_per_launch", "avg_launch_ms",
              "valu_issue_frac", "valu_active_threads", "lds_bytes", "lds_GBs", "lds_bank_conflict_frac"):
        if k in fig:
            print(f"{k}: {fig[k]:.6g}")
    print()
    print(f"- traffic = FETCH_SIZE x 1024 x {sf} + WRITE_SIZE x 1024 x {sw} (KiB per dispatch, scaled by the copy "
          "calibration, copy_1GiB_calibration_pmc.md).")
    print("- valu_issue_frac = SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x GRBM_GUI_ACTIVE cycles): the nominal issue load "
          "at one wave64 instruction per 4 cycles per SIMD; above 1 some instructions retire faster than that "
          "(transcendentals and moves co-issue), so it reads as 'VALU-issue bound', not as an exact utilisation.")
    print("- valu_active_threads = SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU: lanes busy per VALU instruction (of 64).")
    print("- lds_bytes = (SQ_INSTS_LDS_LOAD_BANDWIDTH + SQ_INSTS_LDS_STORE_BANDWIDTH) x 64 B; lds_GBs over the trace's "
          "average launch; LDS peak 128 B/clk/CU x 256 CUs x 2.4 GHz = 78643 GB/s.")
    print("- lds_bank_conflict_frac = SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE.")
    print()
    print("| counter | rows (dispatches x instances) | sum per dispatch |")
    print("|---|---|---|")
    for c, (n, v) in sorted(rows.items()):
        unit = " (KiB)" if c in ("FETCH_SIZE", "WRITE_SIZE") else ""
        print(f"| {c}{unit} | {n} | {v:.6g} |")


def profile_figures(kernel: str, workload_tag: str):
    """Figures of a kernel from its committed rocprofv3 summary (profiles/<kernel>_<workload>_pmc.md), if one exists:
    every `key: value` line of the file."""
    name = re.sub(r"[<>, ]+", "_", kernel).strip("_")
    p = ROOT / "profiles" / f"{name}_{workload_tag}_pmc.md"
    if not p.exists():
        return {}
    out = {}
    for line in p.read_text().splitlines():
        m = re.match(r"^([a-z_]+):\s*([-0-9.eE+]+)\s*$", line.strip())
        if m:
            out[m.group(1)] = float(m.group(2))
    return out


def reset_peak(dev):
    """Start the peak-memory count; the allocator's statistics exist once the device is initialised, which a
    workload run on its own has not done yet."""
    import torch

    torch.cuda.init()
    torch.cuda.reset_peak_memory_stats(dev)


def measure_copy_peak():
    """Device-to-device copy bandwidth (read + write bytes / time) of a 1 GiB buffer, timed with HIP events on
    the current stream: the measured HBM figure to set beside the nominal 8 TB/s."""
    import torch

    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(5):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    gbs = 5 * 2 * (1 << 30) / (t0.elapsed_time(t1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return round(gbs, 1)


def copy_calibration():
    """A plain copy kernel of known size: reads and writes exactly 2^30 bytes per dispatch."""
    import torch

    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_dispatch": 1 << 30, "copy_bytes_written_per_dispatch": 1 << 30}))


def stage_timing(run):
    """Per-kernel {name: (launches, total_ms)} of `run()` with a HIP event pair around every launch."""
    import torch

    from us3d import _lib

    torch.cuda.synchronize()
    _lib.stage_timing_read()
    _lib.stage_timing(True)
    run()
    torch.cuda.synchronize()
    _lib.stage_timing(False)
    return _lib.stage_timing_read()


def roofline(stages, launch_bytes, profile_key, workload_tag):
    """HBM roofline of the dominant kernel by total time: algorithmic bytes per launch / the average launch duration
    measured with HIP events on the launch stream; measured HBM traffic, VALU issue and lanes, and LDS throughput from
    the kernel's committed rocprofv3 summary (profile_key: its name there) when there is one."""
    dom, (launches, total_ms) = max(stages.items(), key=lambda kv: kv[1][1])
    avg_ms = total_ms / max(launches, 1)
    alg = launch_bytes.get(dom)
    achieved = alg / (avg_ms * 1e-3) / 1e9 if alg else None
    fig = profile_figures(profile_key(dom), workload_tag) if profile_key else {}
    out = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
           "traffic": fig.get("traffic_bytes_per_launch"), "avg_launch_ms": round(avg_ms, 4),
           "algorithmic_bytes_per_launch": alg,
           "stage_ms_per_launch": {k: round(v[1] / max(v[0], 1), 3)
                                   for k, v in sorted(stages.items(), key=lambda kv: -kv[1][1])}}
    if "valu_issue_frac" in fig:
        out["valu_issue_frac"] = round(fig["valu_issue_frac"], 3)
    if "valu_active_threads" in fig:
        out["valu_active_threads"] = round(fig["valu_active_threads"], 1)
    if "lds_bytes" in fig:
        lds_gbs = fig["lds_bytes"] / (avg_ms * 1e-3) / 1e9
        out["lds_GBs"] = round(lds_gbs, 1)
        out["lds_peak_GBs"] = round(LDS_PEAK_GBS, 1)
        out["lds_frac"] = round(lds_gbs / LDS_PEAK_GBS, 4)
    if "lds_bank_conflict_frac" in fig:
        out["lds_bank_conflict_frac"] = round(fig["lds_bank_conflict_frac"], 4)
    return out


def visible_pairs(trainer) -> int:
    """Visible (frame, Gaussian) pairs of the trainer's last map: its us3d_stats.n_visible."""
    from us3d.batch import read_stats

    return int(read_stats(trainer.workspace).n_visible)


# ---------------------------------------------------------------- workloads

def bench_slice(args):
    """The headline: the slice splat training step over the 256-frame config-2 sweep."""
    import torch

    from us3d.batch import CapacityHint, read_stats
    from us3d.scenes import config2
    from us3d.train_step import SliceSplatTrainer
    from us3d.ultragauss import UltraGaussRenderConfig, render_frames

    dev = torch.device("cuda:0")
    F, G, ts = args.frames, args.gaussians, args.tile_size
    scene, target, frames = config2(n_frames=F, n_gauss=G)
    H, W = frames[0].height, frames[0].width
    cfg = UltraGaussRenderConfig(H=H, W=W, lateral_spacing_mm_per_pix=0.2, axial_spacing_mm_per_pix=0.2,
                                 axial_near_mm=0.0, fixed_bg_alpha=1e-6, tile_size=ts, work_queue=bool(args.work_queue))
    hint = CapacityHint()
    with torch.no_grad():
        tgt = [t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)]
        gt = render_frames(*tgt, None, frames, cfg, hint=CapacityHint())[0].view(F, H, W).clone()
    del tgt

    def logit(p):
        return torch.log(p / (1 - p))

    means = torch.nn.Parameter(scene.means.to(dev))
    tril = torch.nn.Parameter(scene.tril.to(dev))
    raw_o = torch.nn.Parameter(logit(scene.opacity.clamp(1e-6, 1 - 1e-6)).to(dev))
    raw_c = torch.nn.Parameter(logit(scene.intensity.clamp(1e-6, 1 - 1e-6)).to(dev))
    init = [p.detach() for p in (means, tril, raw_o, raw_c)]

    reset_peak(dev)
    trainer = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=not args.eager, check_every=10 ** 9,
                                fused_loss=not args.separate_loss)
    dt_b = timed(trainer.step, args.steps, args.warmup)
    if trainer.overflowed():  # sticky over every replay since the workspace was allocated
        raise RuntimeError("instance capacity overflowed inside the timed region; the measurement is invalid")
    fps_b = F * args.steps / dt_b
    peak = torch.cuda.max_memory_allocated(dev)
    ws_bytes = trainer.workspace.numel()
    log(f"slice batched ({'eager' if args.eager else 'graph'}): {dt_b / args.steps * 1e3:.2f} ms/step, "
        f"{fps_b:.1f} frames/s, loss {float(trainer.losses().mean()):.5f}, workspace {ws_bytes / 2**30:.2f} GiB")
    del trainer
    torch.cuda.empty_cache()

    # per-kernel durations of the same launch sequence (eager, HIP event pair around every launch)
    tt = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=False, check_every=10 ** 9,
                           fused_loss=not args.separate_loss)
    tt.step()

    def run():
        for _ in range(args.timing_steps):
            tt.step()

    stages = stage_timing(run)
    K = int(read_stats(tt.workspace).n_instances)
    g_vis = visible_pairs(tt)
    P = F * H * W
    del tt
    torch.cuda.empty_cache()
    # the splat kernels in the other launch mode, same step
    from dataclasses import replace as _replace

    to = SliceSplatTrainer(*init, frames, gt, _replace(cfg, work_queue=not cfg.work_queue), use_graph=False,
                           check_every=10 ** 9, fused_loss=not args.separate_loss)
    to.step()

    def run_other():
        for _ in range(args.timing_steps):
            to.step()

    stages_other = stage_timing(run_other)
    del to
    torch.cuda.empty_cache()
    modes = {}
    for wq, st_ in ((cfg.work_queue, stages), (not cfg.work_queue, stages_other)):
        modes["work_queue" if wq else "one_group_per_tile"] = {
            k: round(st_[k][1] / max(st_[k][0], 1), 3) for k in ("ug_forward_kernel", "ug_backward_kernel") if k in st_}

    fps_p = None
    plain_kernel_ms = None
    batched_kernel_ms = sum(v[1] for v in stages.values()) / args.timing_steps
    if not args.no_plain:
        opt = torch.optim.Adam([means, tril, raw_o, raw_c], lr=1e-3)

        def step_plain():
            opt.zero_grad(set_to_none=True)
            for i, fr in enumerate(frames):
                img = render_frames(means, tril, torch.sigmoid(raw_o), torch.sigmoid(raw_c), None, [fr], cfg,
                                    hint=hint)[0]
                (frame_losses(img.view(1, H, W), gt[i:i + 1]).sum() / F).backward()
            opt.step()

        dt_p = timed(step_plain, args.plain_steps, 1)
        fps_p = F * args.plain_steps / dt_p
        log(f"slice plain: {dt_p / args.plain_steps * 1e3:.1f} ms/step, {fps_p:.1f} frames/s")
        # the library's kernel time inside one plain step (HIP events around its launches), so the plain/batched
        # ratio can be split into kernel time and the per-frame host work (tables, workspace, autograd, torch Adam)
        st_p = stage_timing(step_plain)
        plain_kernel_ms = sum(v[1] for v in st_p.values())

    launch_bytes = {
        # algorithmic HBM bytes per launch over the batch: the per-frame terms 96 G + 80 G_vis + 96 K + 24 P of
        # the slice splat split by the kernel that moves them (K tile instances, P pixels, G_vis visible pairs)
        "ug_backward_kernel": 36 * K + 8 * P + 32 * g_vis,
        "ug_forward_kernel": 36 * K + 16 * P,
        "sort_small_segments_kernel": 24 * K,
        "emit_instances_kernel": 28 * g_vis + 12 * K,
        "ug_cull_kernel": 48 * G + F * G // 8,
        "ug_setup_kernel": 4 * g_vis + 48 * G + 80 * g_vis,
        "ug_reduce_kernel": 48 * K + 48 * G,
    }
    tag = f"{F}x{H}x{W}_g{G}"
    full = (F, G, H, W, ts) == (256, 200_000, 512, 512, 16)
    # the template instances the step launches: (tile, depth/alpha terms, frames' grids, LDS batch)
    fb = cfg.fwd_lds_batch or {8: 64, 16: 128, 32: 256}[ts]
    bb = cfg.bwd_lds_batch or {8: 64, 16: 256, 32: 256}[ts]
    key = (lambda k: {"ug_backward_kernel": f"ug_backward_kernel<{ts}, false, true, {bb}>",
                      "ug_forward_kernel": f"ug_forward_kernel<{ts}, false, true, {fb}>"}.get(k, k)) if full else None
    roof = roofline(stages, launch_bytes, key, tag)
    roof["measured_copy_GBs"] = measure_copy_peak()
    return {
        "metric": METRIC, "value": fps_b, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt_b / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"slice splat fwd+bwd+Adam, {G} Gaussians in a 120 mm cube, {F} freehand frames "
                               f"{H}x{W} at 0.2 mm, L1+0.2(1-SSIM), batched {'eager' if args.eager else 'hipGraph'} "
                               f"step, tile {ts}, "
                               f"{'per-XCD work queues' if cfg.work_queue else 'one workgroup per tile'}",
                   "frames": F, "gaussians": G, "tile_size": ts, "instances": K, "visible_pairs": g_vis,
                   "work_queue": bool(cfg.work_queue)},
        "plain_frames_per_s": fps_p, "batched_over_plain": (fps_b / fps_p) if fps_p else None,
        "kernel_ms_per_step": {"batched": batched_kernel_ms, "plain": plain_kernel_ms,
                               "plain_host_ms": (dt_p / args.plain_steps * 1e3 - plain_kernel_ms)
                               if plain_kernel_ms is not None else None},
        "iters_per_s": args.steps / dt_b, "peak_hbm_bytes": int(peak), "workspace_bytes": int(ws_bytes),
        "roofline": roof, "splat_ms_per_launch_by_launch_mode": modes,
    }


def ray_cpu_baseline(scene, frame, cfg, n_rays: int):
    """The oracle's CPU restatement of the reference B-mode renderer (FP32 forward + FP64 adjoint, one thread) on
    n_rays scanlines of one frame: the mapper runs over the whole frame, the shader over the sampled rays."""
    import numpy as np

    from oracle import ur as our
    from us3d.scenes import ray_geometry

    pix, ro, rd = (t.detach().cpu().numpy().astype(np.float32) for t in ray_geometry(frame, "cpu"))
    n = lambda t: t.detach().cpu().numpy().astype(np.float32)  # noqa: E731
    W = frame.width
    st = our.RaySettings(k_sigma=cfg.k_sigma, Rw=cfg.quadrature_window_radius_mm, grad_threshold=cfg.grad_threshold)
    fr = our.RayFrame(origin=n(frame.origin), axis_lat=n(frame.axis_lat), axis_ax=n(frame.axis_ax),
                      axis_ele=n(frame.axis_ele), W=W, H=frame.height, width_mm=float(cfg.width_mm),
                      near_mm=float(cfg.near_mm), col_offset=float(cfg.lateral_col_offset), pixels=pix, ray_o=ro,
                      ray_d=rd, delta=np.zeros(W, np.float32))
    params = [n(t) for t in scene.params()]
    t0 = time.perf_counter()
    rm = our.map_frame(params[0], params[1], fr, st)
    r0 = W // 2 - n_rays // 2
    keep = np.zeros(W, bool)
    keep[r0:r0 + n_rays] = True
    rm.counts = np.where(keep, rm.counts, 0).astype(np.int32)
    hits = int(rm.counts.sum())
    img, _aux = our.forward(params, fr, rm, st)
    our.backward_f64(params, fr, rm, st, np.sign(img).astype(np.float32))
    dt = time.perf_counter() - t0
    return {"value": (n_rays / W) / dt, "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"{n_rays} of {W} scanlines x {frame.height} samples of the middle frame ({hits} ray-Gaussian "
                      f"hits), all {params[0].shape[0]} Gaussians mapped: oracle FP32 forward + FP64 adjoint, one "
                      f"thread, {dt:.1f} s; frames/s = sampled fraction of a frame / time"}


def bench_ray(args):
    """UltraG-Ray B-mode forward+backward over a batch of frames (200k Gaussians, 128 x 1024 scanline frames)."""
    import torch

    from us3d.batch import CapacityHint, FrameSpec, SweepSpec
    from us3d.scenes import config3, make_ray_scene
    from us3d.ultragray import URBatch, UltragrayRenderConfig

    dev = torch.device("cuda:0")
    F, G = args.ray_frames, args.gaussians
    scene, frames = config3(n_frames=F, n_gauss=G)
    target = make_ray_scene(G, 120.0, seed=1)
    f0 = frames[0]
    H, W = f0.height, f0.width
    cfg = UltragrayRenderConfig(image_width=W, image_height=H, width_mm=W * f0.delta_lat, near_mm=f0.near_mm,
                                lateral_col_offset=f0.col_offset, ops_max_offset_mm=0.0)
    specs = [FrameSpec(f.origin, f.axis_lat, f.axis_ax, f.axis_ele, H, W, cfg.pixel_spacing(), f.delta_ax, f.near_mm,
                       f.col_offset, 0) for f in frames]
    # the linear probe's samples and rays are evaluated inside the kernels from the frame table (no (P,3) array)
    pix = ro = rd = None
    off = torch.zeros(F * W, device=dev)
    params = [t.to(dev).float().contiguous() for t in scene.params()]
    tparams = [t.to(dev).float().contiguous() for t in target.params()]

    def run(batch, p):
        m = batch.map(*p)
        image, aux = batch.forward(m, pix, ro, rd, off)
        return m, image, aux

    reset_peak(dev)
    batch = URBatch(specs, [SweepSpec(0, G)], G, cfg.settings(), dev, CapacityHint())
    with torch.no_grad():
        _m, gt, _a = run(batch, tparams)
        gt = gt.clone()

    def step():
        m, image, aux = run(batch, params)
        return batch.backward(m, pix, ro, rd, off, aux, torch.sign(image - gt))  # L1 (sum) loss gradient

    dt_b = timed(step, args.steps, args.warmup)
    fps_b = F * args.steps / dt_b
    peak = torch.cuda.max_memory_allocated(dev)
    log(f"ray batched: {dt_b / args.steps * 1e3:.1f} ms/step, {fps_b:.1f} frames/s")
    holder = {}

    def once():
        m, image, aux = run(batch, params)
        batch.backward(m, pix, ro, rd, off, aux, torch.sign(image - gt))
        holder["K"] = int(m.stats.n_instances)

    stages = stage_timing(once)
    K = holder["K"]
    P = F * H * W

    fps_p = None
    if not args.no_plain:
        plain = [URBatch([specs[i]], [SweepSpec(0, G)], G, cfg.settings(), dev, CapacityHint()) for i in range(F)]

        def step_plain():
            for i, b in enumerate(plain):
                rs = slice(i * W, (i + 1) * W)
                mm = b.map(*params)
                img, ax = b.forward(mm, None, None, None, off[rs])
                b.backward(mm, None, None, None, off[rs], ax, torch.sign(img - gt[i * W * H:(i + 1) * W * H]))

        dt_p = timed(step_plain, args.plain_steps, 1)
        fps_p = F * args.plain_steps / dt_p
        log(f"ray plain: {dt_p / args.plain_steps * 1e3:.1f} ms/step, {fps_p:.1f} frames/s")

    launch_bytes = {"ur_forward_kernel": 64 * K + 8 * P, "ur_backward_kernel": 64 * K + 8 * P}
    full = (F, G, W, H) == (128, 200_000, 128, 1024)
    # the bench renders on the probe's grid: the <true> instances of the ray kernels
    roof = roofline(stages, launch_bytes, (lambda k: f"{k}<true>") if full else None, f"ray_{F}x{W}x{H}_g{G}")
    cpu = None if args.no_cpu_baseline else ray_cpu_baseline(scene, frames[F // 2], cfg, args.cpu_sample_rays)
    return {
        "metric": "UltraG-Ray fwd+bwd frames/s per sweep, batched vs plain on 1 GPU; % of HBM roofline",
        "value": fps_b, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt_b / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"UltraG-Ray B-mode map + forward + adjoint, {G} Gaussians in a 120 mm cube, {F} "
                               f"freehand frames of {W} scanlines x {H} samples (0.8 mm x 0.1 mm), L1 (sum) loss "
                               f"gradient, batched eager", "frames": F, "gaussians": G, "ray_hits": K},
        "plain_frames_per_s": fps_p, "batched_over_plain": (fps_b / fps_p) if fps_p else None,
        "peak_hbm_bytes": int(peak), "roofline": roof, "cpu_baseline": cpu,
    }


def bench_fan(args):
    """Config 5: 2M Gaussians in a 200 x 150 x 150 mm box, a 1000-frame sweep of a 60-degree fan probe (640 beams x
    480 depth samples over 40-160 mm). One step = one captured training step over --fan-frames consecutive frames
    of the sweep (all 1000 by default; pixel grid evaluated in the kernels), rendered in chunks of --fan-chunk
    frames inside the step (map -> forward -> loss -> adjoint per chunk, gradients summed, one Adam update), so the
    workspace holds one chunk's pairs and instances; reports frames/s, iterations/s and peak HBM."""
    import torch

    from us3d.batch import CapacityHint
    from us3d.scenes import config5, fan_pixels
    from us3d.train_step import SliceSplatTrainer
    from us3d.ultragauss import UltraGaussRenderConfig, render_frames

    dev = torch.device("cuda:0")
    scene, target, sweep = config5(n_gauss=args.fan_gaussians)
    F = min(args.fan_frames, len(sweep))
    f0 = len(sweep) // 2 - F // 2
    frames = sweep[f0:f0 + F]
    chunk = args.fan_chunk if args.fan_chunk and args.fan_chunk < F else None
    G = scene.means.shape[0]
    H, W = frames[0].height, frames[0].width
    P = F * H * W
    cfg = UltraGaussRenderConfig(H=H, W=W, lateral_spacing_mm_per_pix=frames[0].delta_lat,
                                 axial_spacing_mm_per_pix=frames[0].delta_ax, axial_near_mm=frames[0].near_mm,
                                 fixed_bg_alpha=1e-6, tile_size=16)
    reset_peak(dev)
    step_gt = chunk or F
    with torch.no_grad():  # the target, rendered chunk by chunk as well
        tgt = [t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)]
        gt = torch.cat([render_frames(*tgt, None, frames[a:a + step_gt], cfg, hint=CapacityHint())[0].clone()
                        for a in range(0, F, step_gt)])
        del tgt
    torch.cuda.empty_cache()
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]

    trainer = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=not args.eager, check_every=10 ** 9,
                                fused_loss=not args.separate_loss,
                                chunk_frames=chunk)
    dt = timed(trainer.step, args.steps, args.warmup)
    if trainer.overflowed():
        raise RuntimeError("instance capacity overflowed inside the timed region; the measurement is invalid")
    fps = F * args.steps / dt
    peak = torch.cuda.max_memory_allocated(dev)
    ws_bytes = trainer.workspace.numel()
    n_chunks = len(trainer.chunks)
    log(f"fan batched: {dt / args.steps * 1e3:.1f} ms/step over {F} frames in {n_chunks} chunk(s), {fps:.1f} "
        f"frames/s, peak HBM {peak / 2**30:.1f} GiB, workspace {ws_bytes / 2**30:.1f} GiB")
    del trainer
    torch.cuda.empty_cache()
    tt = SliceSplatTrainer(*init, frames, gt, cfg, use_graph=False, check_every=10 ** 9,
                           fused_loss=not args.separate_loss, chunk_frames=chunk)
    tt.step()
    stages = stage_timing(tt.step)
    K, g_vis = 0, 0  # instances and visible pairs summed over the chunks
    with torch.no_grad():
        for c in tt.chunks:
            m = c[4].map(tt.means, tt.tril, tt.raw_o, tt.raw_c)
            K += int(m.stats.n_instances)
            g_vis += int(m.stats.n_visible)
            del m
    del tt
    torch.cuda.empty_cache()
    # algorithmic bytes per launch: one launch renders one chunk, so the per-launch figure is the chunk average
    launch_bytes = {"ug_backward_kernel": (36 * K + 8 * P + 32 * g_vis) // n_chunks,
                    "ug_forward_kernel": (36 * K + 16 * P) // n_chunks}
    roof = roofline(stages, launch_bytes, None, "")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import fallback as ofb

        torch.set_num_threads(os.cpu_count() or 1)
        fr = frames[F // 2]
        rows = args.fan_cpu_rows
        pix = fan_pixels(fr).reshape(H, W, 3)[H // 2 - rows // 2:H // 2 - rows // 2 + rows].reshape(-1, 3)
        with torch.no_grad():
            ref, _ = ofb.render(target.means, target.tril, target.opacity, target.intensity, pix, rows, W)
        dtc = _fallback_fwd_bwd(scene, ref, pix, rows, W)
        cpu = {"value": rows * W / (H * W) / dtc, "unit": "frames/s", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": f"{rows} rows x {W} pixels of fan frame {f0 + F // 2} against all {G} Gaussians, forward + "
                         f"autograd backward of the restated PyTorch fallback renderer, {dtc:.1f} s "
                         f"({rows * W * G / dtc / 1e6:.0f} M pair evaluations/s); frames/s = sampled fraction of a "
                         f"{W}x{H} frame / time"}
    return {
        "metric": "fwd+bwd frames/s of a fan-probe sweep in one captured training step on 1 GPU; % of HBM roofline",
        "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "iters_per_s": args.steps / dt, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"slice splat fwd+bwd+Adam, {G} Gaussians in a 200x150x150 mm box, frames "
                               f"{f0}..{f0 + F - 1} of a 1000-frame 60-degree fan sweep ({W} beams x {H} samples, "
                               f"40-160 mm), L1+0.2(1-SSIM), batched {'eager' if args.eager else 'hipGraph'} step "
                               f"in {n_chunks} chunk(s) of {chunk or F} frames, tile 16, pixel grid evaluated in "
                               f"the kernels",
                   "frames": F, "chunk_frames": chunk or F, "gaussians": G, "pixels": P, "instances": K,
                   "visible_pairs": g_vis},
        "peak_hbm_bytes": int(peak), "workspace_bytes": int(ws_bytes), "roofline": roof, "cpu_baseline": cpu,
    }


def bench_gaussianpile(args):
    """The GaussianPile us_focus slice splat (the second slice model) on the config-2 scene: 200k Gaussians
    (means, log-scales, quaternions, opacity logits) and the 256-frame 512x512 freehand sweep. One step = the
    captured training step of GaussianPileTrainer over all frames: map + forward, L1 + 0.2 (1 - SSIM) against
    a target rendered from the seed-1 scene, adjoint, Adam. "plain": the same API frame by frame with torch
    activations, torch autograd and torch.optim.Adam."""
    import torch

    from us3d.batch import CapacityHint, SweepSpec, read_stats
    from us3d.gaussianpile import GaussianPlieRenderConfig, GPBatch, render_frames
    from us3d.scenes import config2
    from us3d.train_step import GaussianPileTrainer

    dev = torch.device("cuda:0")
    scene, target, frames = config2(n_frames=args.frames, n_gauss=args.gaussians)
    F = len(frames)
    G = scene.means.shape[0]
    H, W = frames[0].height, frames[0].width
    P = F * H * W
    fr0 = frames[0]
    cfg = GaussianPlieRenderConfig(H=H, W=W, lateral_spacing_mm_per_pix=fr0.delta_lat,
                                   axial_spacing_mm_per_pix=fr0.delta_ax, axial_near_mm=fr0.near_mm,
                                   lateral_col_offset=fr0.col_offset, tile_size_px=args.tile_size)
    hint = CapacityHint()
    reset_peak(dev)
    with torch.no_grad():
        gt = render_frames(*[t.to(dev) for t in (target.means, target.scales, target.quats, target.opacity)], frames,
                           cfg, hint=hint).clone()
    op = scene.opacity.clamp(1e-6, 1 - 1e-6)
    raw = [t.to(dev).float().contiguous() for t in (scene.means, torch.log(scene.scales), scene.quats,
                                                     torch.log(op / (1 - op)))]
    tr = GaussianPileTrainer(*raw, frames, gt, cfg, lr=1e-3, use_graph=not args.eager, check_every=10 ** 9,
                             fused_loss=not args.separate_loss)
    dt = timed(tr.step, args.steps, args.warmup)
    if tr.overflowed():  # sticky over every replay since the workspace was allocated
        raise RuntimeError("instance capacity overflowed inside the timed region; the measurement is invalid")
    fps = F * args.steps / dt
    peak = torch.cuda.max_memory_allocated(dev)
    log(f"gaussianpile batched: {dt / args.steps * 1e3:.1f} ms/step, {fps:.1f} frames/s")
    del tr
    torch.cuda.empty_cache()
    fps_plain = None
    if not args.no_plain:
        n_plain = min(8, F)
        n = H * W
        ps = [t.clone().requires_grad_(True) for t in raw]
        opt = torch.optim.Adam(ps, lr=1e-3)
        singles = [GPBatch([frames[k]], [SweepSpec(0, G)], G, cfg.settings(), dev, hint) for k in range(n_plain)]

        def plain():
            for k in range(n_plain):
                opt.zero_grad(set_to_none=True)
                img = render_frames(ps[0], torch.exp(ps[1]).clamp_min(1e-5),
                                    torch.nn.functional.normalize(ps[2], dim=-1, eps=1e-8), torch.sigmoid(ps[3]),
                                    [frames[k]], cfg, hint=hint, batch=singles[k])
                frame_losses(img.view(1, H, W), gt[k * n:(k + 1) * n].view(1, H, W)).mean().backward()
                opt.step()

        dtp = timed(plain, 1, 1)
        fps_plain = n_plain / dtp
        log(f"gaussianpile plain (per frame): {fps_plain:.1f} frames/s")
    tr_t = GaussianPileTrainer(*raw, frames, gt, cfg, lr=1e-3, use_graph=False, check_every=10 ** 9,
                               fused_loss=not args.separate_loss)
    tr_t.step()
    stages = stage_timing(tr_t.step)
    K = int(read_stats(tr_t.workspace).n_instances)
    g_vis = visible_pairs(tr_t)
    del tr_t
    torch.cuda.empty_cache()
    # algorithmic bytes per launch: per instance its 8 B sort key, 32 B footprint and 4 B intensity gathered
    # (adjoint: also its 4 B emit index read and its 32 B record-space gradient written), per pixel the 4 B
    # image written (forward) or its 4 B gradient read (adjoint)
    launch_bytes = {"gp_forward_kernel": 44 * K + 4 * P, "gp_backward_kernel": 80 * K + 4 * P}
    full = (F, G, H, W) == (256, 200_000, 512, 512)
    roof = roofline(stages, launch_bytes, (lambda k: f"{k}<{args.tile_size}>") if full else None, f"{F}x{H}x{W}_g{G}")
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import gp as ogp

        torch.set_num_threads(os.cpu_count() or 1)
        fr = frames[F // 2]
        rows = 32
        band = ogp.GPFrame(origin=fr.origin, axis_lat=fr.axis_lat, axis_ax=fr.axis_ax, axis_ele=fr.axis_ele, H=rows,
                           W=W, delta_lat=fr.delta_lat, delta_ax=fr.delta_ax,
                           near=fr.near_mm + (H // 2) * fr.delta_ax, col_offset=fr.col_offset)
        ps = [t.double().clone().requires_grad_(True) for t in (scene.means, scene.scales, scene.quats, scene.opacity)]
        t0 = time.perf_counter()
        imgc = ogp.render(*ps, torch.ones(G, dtype=torch.float64), band, ogp.GPSettings(tile_size=args.tile_size),
                          fp64=True)
        imgc.sum().backward()
        dtc = time.perf_counter() - t0
        cpu = {"value": rows / H / dtc, "unit": "frames/s", "cores": torch.get_num_threads(), "kind": "port",
               "sample": f"{rows} rows x {W} pixels (two tile rows) of frame {F // 2} against all {G} Gaussians: the "
                         f"oracle's tile lists and forward, FP64 with torch autograd for the adjoint, {dtc:.1f} s; "
                         f"frames/s = sampled fraction of a frame / time"}
    return {
        "metric": "GaussianPile us_focus training-step frames/s per sweep, batched vs plain on 1 GPU; % of HBM roofline",
        "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"GaussianPile us_focus training step (map + forward + L1+0.2(1-SSIM) + adjoint + "
                               f"Adam, raw activations fused), {G} Gaussians in a 120 mm cube, {F} freehand frames "
                               f"{H}x{W} at 0.2 mm, radius_sigma 3, sigma_z 1 mm, batched "
                               f"{'eager' if args.eager else 'hipGraph'}, tile {args.tile_size}",
                   "frames": F, "gaussians": G, "instances": K, "visible_pairs": g_vis},
        "plain_frames_per_s": fps_plain, "batched_over_plain": (fps / fps_plain) if fps_plain else None,
        "peak_hbm_bytes": int(peak), "roofline": roof, "cpu_baseline": cpu,
    }


def bench_multisweep(args):
    """8 sweeps (Gaussians log-uniform in 50k..500k, 32..400 frames, frame size 256x256 / 512x512 / 640x480) in
    ONE hipGraph-captured training step; the tile size and the forward's and adjoint's LDS batch sizes are picked by
    the autotuner on a frame sample first."""
    from dataclasses import replace

    import torch

    from us3d.autotune import apply, tune_tile_size
    from us3d.batch import CapacityHint, read_stats
    from us3d.scenes import config4, config4_specs
    from us3d.train_step import SliceSplatTrainer
    from us3d.ultragauss import UltraGaussRenderConfig, render_frames

    dev = torch.device("cuda:0")
    specs = config4_specs(scale=args.ms_scale)
    scene, target, frames, sweeps = config4(scale=args.ms_scale)
    F = len(frames)
    G = scene.means.shape[0]
    P = sum(f.height * f.width for f in frames)
    f0 = frames[0]
    cfg = UltraGaussRenderConfig(H=f0.height, W=f0.width, lateral_spacing_mm_per_pix=0.2,
                                 axial_spacing_mm_per_pix=0.2, axial_near_mm=0.0, fixed_bg_alpha=1e-6, tile_size=16)
    reset_peak(dev)
    with torch.no_grad():
        gt = render_frames(*[t.to(dev) for t in (target.means, target.tril, target.opacity, target.intensity)], None,
                           frames, cfg, sweeps=sweeps, hint=CapacityHint())[0].clone()
    del target
    torch.cuda.empty_cache()
    logit = lambda p: torch.log(p / (1 - p))  # noqa: E731
    init = [scene.means.to(dev), scene.tril.to(dev), logit(scene.opacity).to(dev), logit(scene.intensity).to(dev)]
    tune = tune_tile_size(*init, frames, gt, cfg, sample_frames=32, reps=2, sweeps=sweeps, use_cache=False)
    cfg = apply(cfg, tune)
    log(f"multisweep autotuner: {tune}")
    torch.cuda.empty_cache()
    trainer = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=not args.eager, check_every=10 ** 9,
                                fused_loss=not args.separate_loss)
    dt = timed(trainer.step, args.steps, args.warmup)
    if trainer.overflowed():
        raise RuntimeError("instance capacity overflowed inside the timed region; the measurement is invalid")
    fps = F * args.steps / dt
    peak = torch.cuda.max_memory_allocated(dev)
    ws_bytes = trainer.workspace.numel()
    log(f"multisweep batched ({'eager' if args.eager else 'graph'}): {dt / args.steps * 1e3:.1f} ms/step, "
        f"{fps:.1f} frames/s, workspace {ws_bytes / 2**30:.1f} GiB")
    del trainer
    torch.cuda.empty_cache()
    tt = SliceSplatTrainer(*init, frames, gt, cfg, sweeps=sweeps, use_graph=False, check_every=10 ** 9,
                           fused_loss=not args.separate_loss)
    tt.step()
    stages = stage_timing(tt.step)
    K = int(read_stats(tt.workspace).n_instances)
    g_vis = visible_pairs(tt)
    del tt
    torch.cuda.empty_cache()
    launch_bytes = {"ug_backward_kernel": 36 * K + 8 * P + 32 * g_vis, "ug_forward_kernel": 36 * K + 16 * P}
    roof = roofline(stages, launch_bytes, None, "")
    return {
        "metric": "fwd+bwd frames/s of a multi-sweep batch in one captured training step on 1 GPU; % of HBM roofline",
        "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"slice splat fwd+bwd+Adam over {len(sweeps)} sweeps (Gaussians, frames, HxW) = "
                               f"{specs}, {G} Gaussians and {F} frames in total, L1+0.2(1-SSIM), batched "
                               f"{'eager' if args.eager else 'hipGraph'} step, tile {cfg.tile_size}, LDS batches "
                               f"{cfg.fwd_lds_batch} (forward) / {cfg.bwd_lds_batch} (adjoint) Gaussians (autotuned)",
                   "frames": F, "gaussians": G, "pixels": P, "tile_size": cfg.tile_size, "instances": K,
                   "visible_pairs": g_vis, "autotune_ms_per_frame": tune["ms_per_frame"],
                   "autotune_lds_batch_ms": tune["lds_batch_ms"]},
        "megapixels_per_s": P * args.steps / dt / 1e6, "peak_hbm_bytes": int(peak), "workspace_bytes": int(ws_bytes),
        "roofline": roof, "cpu_baseline": None,
    }


WORKLOADS = {"ray": bench_ray, "multisweep": bench_multisweep, "fan": bench_fan, "gaussianpile": bench_gaussianpile}


def nested(name, args):
    """One of the other workloads, short (1 warm-up, 3 timed steps, no plain or CPU legs), as a compact object."""
    import torch

    sub = copy.copy(args)
    sub.steps, sub.warmup, sub.no_plain, sub.no_cpu_baseline = 3, 1, True, True
    t0 = time.perf_counter()
    try:
        d = WORKLOADS[name](sub)
    except Exception as e:  # reported in the line; the headline stands on its own
        torch.cuda.empty_cache()
        return {"error": f"{type(e).__name__}: {e}"}
    torch.cuda.empty_cache()
    r = d["roofline"]
    out = {"metric": d["metric"], "value": d["value"], "unit": d["unit"], "ms_per_step": d["ms_per_step"],
           "steps": sub.steps, "warmup": sub.warmup, "workload": d["config"]["workload"],
           "frames": d["config"].get("frames"), "gaussians": d["config"].get("gaussians"),
           "dominant_kernel": r["kernel"], "dominant_avg_launch_ms": r["avg_launch_ms"], "roofline_frac": r["frac"],
           "stage_ms_per_launch": r["stage_ms_per_launch"], "peak_hbm_bytes": d.get("peak_hbm_bytes"),
           "wall_s": round(time.perf_counter() - t0, 1)}
    for k in ("workspace_bytes", "megapixels_per_s", "iters_per_s"):
        if k in d:
            out[k] = d[k]
    if "chunk_frames" in d["config"]:
        out["chunk_frames"] = d["config"]["chunk_frames"]
    if "tile_size" in d["config"]:
        out["tile_size"] = d["config"]["tile_size"]
    return out


def main():
    args = parse()
    if args.rocprof_summary:
        rocprof_summary(*args.rocprof_summary)
        return
    if args.rocprof_derive:
        rocprof_derive(args.rocprof_derive[0], args.rocprof_derive[1:])
        return
    if args.rocprof_report:
        rocprof_report(args.rocprof_report[0], args.rocprof_report[1], args.rocprof_report[2:])
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)
    if args.workload != "slice":
        print(json.dumps(WORKLOADS[args.workload](args)), flush=True)
        return
    line = bench_slice(args)
    if not args.no_workloads:
        line["workloads"] = {}
        for name in ("ray", "multisweep", "fan", "gaussianpile"):
            log(f"nested workload {name} ...")
            line["workloads"][name] = nested(name, args)
    line["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(args.cpu_sample_rows, args.cpu_runs)
    if not args.no_cpu_baseline:  # the oracle leg: its checker figures beside its timing
        import torch

        log("parity of one config-2 frame against the oracle ...")
        line["parity"] = parity_config2_frame(torch.device("cuda:0"))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
