"""Oracle (test infrastructure only): the GaussianPile "us_focus" slice renderer, restated.

Follows the reference's CUDA path for shader weight mode US_FOCUS with the tile mapper and additive
compositing, as the facade configures it (gaussianplie_renderer/autograd.py:17-55, 228-246):
  footprint  us_focus_device.cuh:176-300 (compute_us_focus_footprint_mm), plane_cond2d_device.cuh:14-18,48-100
             (covariance jitter, world covariance, probe-frame covariance), shared.cuh:6-34 (quaternion);
  mapper     mapper/core_tile.cu:25-170 (kernel_ours_count_tile_rects: no elevation cull in this mode,
             rect from the footprint, depth key 0, so each tile lists its Gaussians in index order);
  shader     shader/forward.cu:27-111,114-210 (candidate weight o * slice_response * det^-1/2, per pixel the
             k-sigma box test, quad >= 0, weight >= weight_min, additive FP32 sum in list order);
  pixel grid slice_grid_geometry.cuh:38-45 (lateral (col - col_offset) * d_lat, axial row * d_ax + near).
With the facade's parameters the elevational focus width is the constant sigma_z, lateral focus is off,
det_gamma = 0.5 and nothing is inflated.

Two restatements share the formulas:
  * footprint_f32 / render_f32: torch float32, elementwise operations in the order the kernels use (no fused
    multiply-adds in the footprint), so the structural decisions (footprint validity, tile rectangles, box
    tests) match the kernels bit for bit; the per-pixel weight b exp(-quad/2) is formed as the kernels form
    it, b exp2(e) with e from the inverse covariance pre-scaled by -log2(e)/2 and three fused multiply-adds,
    so a weight can differ from the kernels' only by exp2's last-ulp rounding;
  * render_f64: torch float64 and differentiable, the same decisions taken from the FP32 pass (a pixel-pair
    either contributes or not), so torch autograd gives the reference gradients.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

JITTER = 1e-6  # plane_cond2d_device.cuh:9, added to the probe-frame covariance diagonal
GRAD_CLAMP = 1e3  # kGradClamp, plane_cond2d_device.cuh:12


@dataclass
class GPSettings:
    k_sigma: float = 3.0          # radius_sigma
    sigma_ele: float = 1.0        # sigma_z_mm
    det_eps: float = 1e-6
    weight_min: float = 1e-6      # max(opacity_cull_threshold, 1e-12)
    tile_size: int = 16


@dataclass
class GPFrame:
    origin: torch.Tensor   # (3,)
    axis_lat: torch.Tensor  # (3,) unit
    axis_ax: torch.Tensor
    axis_ele: torch.Tensor
    H: int
    W: int
    delta_lat: float
    delta_ax: float
    near: float
    col_offset: float


def _quat_rot(q: torch.Tensor):
    """shared.cuh:16-34: (w, x, y, z) normalised by 1/sqrt(max(1e-12, |q|^2)); returns the 9 entries row-major."""
    qq = q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1] + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3]
    nq = 1.0 / torch.sqrt(torch.clamp(qq, min=1e-12))
    w, x, y, z = q[:, 0] * nq, q[:, 1] * nq, q[:, 2] * nq, q[:, 3] * nq
    xx, yy, zz = x * x, y * y, z * z
    xy, xz, yz = x * y, x * z, y * z
    wx, wy, wz = w * x, w * y, w * z
    one = torch.ones_like(w)
    return [one - 2.0 * (yy + zz), 2.0 * (xy - wz), 2.0 * (xz + wy),
            2.0 * (xy + wz), one - 2.0 * (xx + zz), 2.0 * (yz - wx),
            2.0 * (xz - wy), 2.0 * (yz + wx), one - 2.0 * (xx + yy)]


def _inv3(M):
    """Cofactor inverse (us_focus_device.cuh:108-145); returns (inverse entries, det)."""
    a00, a01, a02, a10, a11, a12, a20, a21, a22 = M
    c00 = a11 * a22 - a12 * a21
    c01 = a02 * a21 - a01 * a22
    c02 = a01 * a12 - a02 * a11
    c10 = a12 * a20 - a10 * a22
    c11 = a00 * a22 - a02 * a20
    c12 = a02 * a10 - a00 * a12
    c20 = a10 * a21 - a11 * a20
    c21 = a01 * a20 - a00 * a21
    c22 = a00 * a11 - a01 * a10
    det = a00 * c00 + a01 * c10 + a02 * c20
    inv_det = 1.0 / det
    return [c00 * inv_det, c01 * inv_det, c02 * inv_det, c10 * inv_det, c11 * inv_det, c12 * inv_det,
            c20 * inv_det, c21 * inv_det, c22 * inv_det], det


def footprint(means, scales, quats, opacity, fr: GPFrame, st: GPSettings):
    """Per-Gaussian footprint on the frame, in the dtype of the inputs. Returns a dict of (G,) tensors:
    mu2 (lat, ax), inv2 (00, 01, 11), r (lat, ax), base weight b, validity `ok`."""
    dt = means.dtype
    o = fr.origin.to(dt)
    al, aa, ae = fr.axis_lat.to(dt), fr.axis_ax.to(dt), fr.axis_ele.to(dt)
    d = [means[:, 0] - o[0], means[:, 1] - o[1], means[:, 2] - o[2]]
    mu_lat = d[0] * al[0] + d[1] * al[1] + d[2] * al[2]
    mu_ax = d[0] * aa[0] + d[1] * aa[1] + d[2] * aa[2]
    mu_ele = d[0] * ae[0] + d[1] * ae[1] + d[2] * ae[2]
    R = _quat_rot(quats)
    s = [torch.clamp(scales[:, i], min=1e-6) for i in range(3)]
    # Sigma_w as plane_cond2d_device.cuh:48-69 forms it from the row-major rotation entries: M[i][k] = R[i][k]
    # s_k^2, Sigma_w[i][j] = sum_k M[i][k] R[k][j], i.e. R diag(s^2) R (not R diag(s^2) R^T), kept as written
    Mx = [R[k] * s[k % 3] * s[k % 3] for k in range(9)]
    Sw = [None] * 9
    for i in range(3):
        for j in range(3):
            Sw[i * 3 + j] = Mx[i * 3 + 0] * R[j] + Mx[i * 3 + 1] * R[3 + j] + Mx[i * 3 + 2] * R[6 + j]
    # Sc = A^T Sigma_w A with A = [a_lat a_ax a_ele] (plane_cond2d_device.cuh:71-100), plus the jitter
    axes = (al, aa, ae)
    cols = []
    for a in axes:
        cols.append([Sw[0] * a[0] + Sw[1] * a[1] + Sw[2] * a[2],
                     Sw[3] * a[0] + Sw[4] * a[1] + Sw[5] * a[2],
                     Sw[6] * a[0] + Sw[7] * a[1] + Sw[8] * a[2]])
    Sc = [None] * 9
    for i, a in enumerate(axes):
        for j in range(3):
            c = cols[j]
            Sc[i * 3 + j] = a[0] * c[0] + a[1] * c[1] + a[2] * c[2]
    Sc[0] = Sc[0] + JITTER
    Sc[4] = Sc[4] + JITTER
    Sc[8] = Sc[8] + JITTER
    Lc, det_c = _inv3(Sc)
    sig = torch.tensor(float(st.sigma_ele), dtype=dt)
    sig2 = sig * sig
    Le = list(Lc)
    Le[8] = Le[8] + 1.0 / torch.clamp(sig2, min=1e-12)
    Se, det_e = _inv3(Le)
    t0 = Lc[0] * mu_lat + Lc[1] * mu_ax + Lc[2] * mu_ele
    t1 = Lc[3] * mu_lat + Lc[4] * mu_ax + Lc[5] * mu_ele
    t2 = Lc[6] * mu_lat + Lc[7] * mu_ax + Lc[8] * mu_ele
    m0 = Se[0] * t0 + Se[1] * t1 + Se[2] * t2
    m1 = Se[3] * t0 + Se[4] * t1 + Se[5] * t2
    s00 = Se[0]
    s01 = 0.5 * (Se[1] + Se[3])
    s11 = Se[4]
    det2 = s00 * s11 - s01 * s01
    inv_det2 = 1.0 / det2
    i00, i01, i11 = s11 * inv_det2, -s01 * inv_det2, s00 * inv_det2
    cross = torch.abs(s01)
    r_lat = st.k_sigma * torch.sqrt(torch.clamp(s00 + cross, min=1e-12))
    r_ax = st.k_sigma * torch.sqrt(torch.clamp(s11 + cross, min=1e-12))
    denom = torch.clamp(Sc[8] + sig2 + st.det_eps, min=1e-12)
    sr = torch.exp(-0.5 * mu_ele * mu_ele / denom)
    norm = 1.0 / torch.sqrt(det2 + st.det_eps)
    b = opacity * sr * norm
    ok = ((det_c > 1e-12) & torch.isfinite(det_c) & (det_e > 1e-12) & torch.isfinite(det_e)
          & (det2 > st.det_eps) & torch.isfinite(det2) & torch.isfinite(b))
    return {"m0": m0, "m1": m1, "i00": i00, "i01": i01, "i11": i11, "r_lat": r_lat, "r_ax": r_ax, "b": b,
            "ok": ok, "mu_ele": mu_ele}


def tile_lists(fp, fr: GPFrame, st: GPSettings):
    """core_tile.cu:138-170 on the FP32 footprint: each tile's list of Gaussian indices, ascending (the stable
    sort by tile with depth key 0). Pairs whose peak weight b is below weight_min are left out: every weight of
    such a pair is below weight_min, so the reference's shader returns 0 for it at every pixel."""
    ts = st.tile_size
    tw, th = -(-fr.W // ts), -(-fr.H // ts)
    f32 = torch.float32
    u = fp["m0"] / torch.tensor(fr.delta_lat, dtype=f32) + torch.tensor(fr.col_offset, dtype=f32)
    v = (fp["m1"] - torch.tensor(fr.near, dtype=f32)) / torch.tensor(fr.delta_ax, dtype=f32)
    ru = torch.clamp(fp["r_lat"] / torch.tensor(max(1e-8, fr.delta_lat), dtype=f32), min=1.0)
    rv = torch.clamp(fp["r_ax"] / torch.tensor(max(1e-8, fr.delta_ax), dtype=f32), min=1.0)
    tsf = torch.tensor(float(ts), dtype=f32)
    x0 = torch.floor((u - ru) / tsf)
    y0 = torch.floor((v - rv) / tsf)
    x1 = torch.ceil((u + ru) / tsf)
    y1 = torch.ceil((v + rv) / tsf)
    lists = [[] for _ in range(tw * th)]
    keep = fp["ok"] & (fp["b"] >= st.weight_min)
    for g in torch.nonzero(keep).flatten().tolist():
        if not all(math.isfinite(float(t[g])) for t in (x0, y0, x1, y1)):
            continue
        gx0, gy0 = max(0, int(x0[g])), max(0, int(y0[g]))
        gx1, gy1 = min(tw, int(x1[g])), min(th, int(y1[g]))
        for ty in range(gy0, gy1):
            for tx in range(gx0, gx1):
                lists[ty * tw + tx].append(g)
    return lists, tw, th


# -log2(e)/2 in FP32: the kernels evaluate exp(-quad/2) as exp2(e) with e formed from the pre-scaled inverse
EXP2_SCALE = -0.72134752044448170


def _fma32(a, b, c):
    """FP32 fused multiply-add: the exact product and sum in float64, one rounding to float32 (a second
    rounding after float64's can differ from a single one only at an exact float64 tie, which these
    operands practically never reach)."""
    return (a.double() * b.double() + c.double()).float()


def _pixel_grid(fr: GPFrame, dtype):
    cols = torch.arange(fr.W, dtype=torch.float32)
    rows = torch.arange(fr.H, dtype=torch.float32)
    xl = (cols - torch.tensor(fr.col_offset, dtype=torch.float32)) * torch.tensor(fr.delta_lat, dtype=torch.float32)
    xa = rows * torch.tensor(fr.delta_ax, dtype=torch.float32) + torch.tensor(fr.near, dtype=torch.float32)
    return xl.to(dtype), xa.to(dtype)


def render(means, scales, quats, opacity, intensity, fr: GPFrame, st: GPSettings, fp64=False):
    """Image (H, W). FP32: the kernels' decisions and FP32 sums in list order. fp64=True: the same pixel-pair set
    (decided in FP32), values in float64 from the float64 inputs, differentiable by torch autograd."""
    f32 = torch.float32
    p32 = [t.detach().to(f32) for t in (means, scales, quats, opacity)]
    fp32 = footprint(*p32, fr, st)
    lists, tw, th = tile_lists(fp32, fr, st)
    xl32, xa32 = _pixel_grid(fr, f32)
    if fp64:
        # footprints of the listed Gaussians only: an invalid footprint (a singular or indefinite matrix) of
        # an unlisted one would otherwise put inf * 0 into the backward pass
        listed = sorted({g for lst in lists for g in lst})
        pos = {g: i for i, g in enumerate(listed)}
        idx = torch.tensor(listed, dtype=torch.long)
        sub = footprint(means[idx], scales[idx], quats[idx], opacity[idx], fr, st) if listed else {}
        fpv = {k: (lambda v: (lambda g: v[pos[g]]))(v) for k, v in sub.items()}
        xl, xa = _pixel_grid(fr, torch.float64)
        inten = intensity
        out = torch.zeros((fr.H, fr.W), dtype=torch.float64)
    else:
        fpv = fp32
        xl, xa = xl32, xa32
        inten = intensity.detach().to(f32)
        out = torch.zeros((fr.H, fr.W), dtype=f32)
    ts = st.tile_size
    wmin = torch.tensor(st.weight_min, dtype=f32)
    C = torch.tensor(EXP2_SCALE, dtype=f32)
    c00, c01, c11 = fp32["i00"] * C, fp32["i01"] * C, fp32["i11"] * C
    tiles = []
    for t, lst in enumerate(lists):
        if not lst:
            continue
        ty, tx = divmod(t, tw)
        ys, xs = slice(ty * ts, min(fr.H, (ty + 1) * ts)), slice(tx * ts, min(fr.W, (tx + 1) * ts))
        acc = torch.zeros((ys.stop - ys.start, xs.stop - xs.start), dtype=out.dtype)
        L32 = xl32[xs][None, :]
        A32 = xa32[ys][:, None]
        Lv = xl[xs][None, :]
        Av = xa[ys][:, None]
        for g in lst:
            dl32 = L32 - fp32["m0"][g]
            da32 = A32 - fp32["m1"][g]
            box = (torch.abs(dl32) <= fp32["r_lat"][g]) & (torch.abs(da32) <= fp32["r_ax"][g])
            # e = C quad with (l, q) = C inv2 d, in the kernels' order: l = fma(c00, dl, c01 da),
            # q = fma(c01, dl, c11 da), e = fma(dl, l, da q); w = b exp2(e). e <= 0 is quad >= 0 (NaN fails
            # it); an infinite quad gives w = 0 < weight_min.
            l32 = _fma32(c00[g], dl32, c01[g] * da32)
            k32 = _fma32(c01[g], dl32, c11[g] * da32)
            e32 = _fma32(dl32, l32, da32 * k32)
            w32 = fp32["b"][g] * torch.exp2(e32)
            m = box & (e32 <= 0) & (w32 >= wmin)
            if not bool(m.any()):
                continue
            if fp64:
                m0, m1, b = fpv["m0"](g), fpv["m1"](g), fpv["b"](g)
                i00, i01, i11 = fpv["i00"](g), fpv["i01"](g), fpv["i11"](g)
                dl = Lv - m0
                da = Av - m1
                q = dl * (i00 * dl + i01 * da) + da * (i01 * dl + i11 * da)
                w = b * torch.exp(-0.5 * torch.where(m, q, torch.zeros_like(q)))  # no overflow off the mask
                acc = acc + torch.where(m, inten[g] * w, torch.zeros_like(w))
            else:
                acc = acc + torch.where(m, inten[g] * w32, torch.zeros_like(w32))
        tiles.append((ys, xs, acc))
    if fp64:
        rows = []
        full = [[None] * tw for _ in range(th)]
        for ys, xs, acc in tiles:
            full[ys.start // ts][xs.start // ts] = acc
        for ty in range(th):
            row = []
            for tx in range(tw):
                h = min(fr.H, (ty + 1) * ts) - ty * ts
                w_ = min(fr.W, (tx + 1) * ts) - tx * ts
                row.append(full[ty][tx] if full[ty][tx] is not None else torch.zeros((h, w_), dtype=torch.float64))
            rows.append(torch.cat(row, dim=1))
        return torch.cat(rows, dim=0)
    for ys, xs, acc in tiles:
        out[ys, xs] = acc
    return out


def render_reference_math(means, scales, quats, opacity, fr: GPFrame, sigma_ele: float):
    """The model as gaussianplie_renderer/reference.py:54-115 states it (focus_project + an untruncated additive
    sum), in float64 with the epsilons at zero: the limit the CUDA footprint approaches as its jitter and
    det_eps vanish and its k-sigma truncation widens."""
    dt = torch.float64
    R9 = _quat_rot(quats.to(dt))
    R = torch.stack([torch.stack(R9[0:3], -1), torch.stack(R9[3:6], -1), torch.stack(R9[6:9], -1)], 1)
    S = torch.diag_embed(scales.to(dt) ** 2)
    cov = R @ S @ R.transpose(1, 2)
    A = torch.stack([fr.axis_lat, fr.axis_ax, fr.axis_ele], 1).to(dt)  # columns
    cov_c = A.T @ cov @ A
    mu_c = (means.to(dt) - fr.origin.to(dt)) @ A
    lc = torch.linalg.inv(cov_c)
    le = lc.clone()
    le[:, 2, 2] = le[:, 2, 2] + 1.0 / sigma_ele ** 2
    se = torch.linalg.inv(le)
    mu_e = (se @ (lc @ mu_c.unsqueeze(-1))).squeeze(-1)
    q_c = (mu_c.unsqueeze(1) @ lc @ mu_c.unsqueeze(-1)).reshape(-1)
    q_e = (mu_e.unsqueeze(1) @ le @ mu_e.unsqueeze(-1)).reshape(-1)
    c2 = se[:, :2, :2]
    det = c2[:, 0, 0] * c2[:, 1, 1] - c2[:, 0, 1] * c2[:, 1, 0]
    alpha = opacity.to(dt) * torch.exp(-0.5 * (q_c - q_e)) / torch.sqrt(det)
    inv2 = torch.linalg.inv(c2)
    xl, xa = _pixel_grid(fr, dt)
    grid = torch.stack(torch.meshgrid(xa, xl, indexing="ij")[::-1], -1).reshape(-1, 2)
    out = torch.zeros(grid.shape[0], dtype=dt)
    for g in range(means.shape[0]):
        dd = grid - mu_e[g, :2]
        q = (dd * (dd @ inv2[g].T)).sum(-1)
        out = out + alpha[g] * torch.exp(-0.5 * q)
    return out.reshape(fr.H, fr.W)


def render_pixels(means, scales, quats, opacity, intensity, fr: GPFrame, st: GPSettings, rows, cols,
                  fp64=False):
    """Values at the pixels (rows[k], cols[k]) of the frame by brute force over EVERY Gaussian, without tile
    lists: a pair contributes where the mapper would list it (valid footprint, b >= weight_min) and the
    pixel passes the shader's tests (box, quad >= 0, weight >= weight_min, decided in FP32 as the kernels
    decide). A tile mapper that drops a contributing pair disagrees with this at some pixel. FP32 sums in
    index order (the tile lists' order); fp64=True: float64 values from the float64 inputs on the same pair
    set, differentiable by torch autograd. Returns (n,)."""
    f32 = torch.float32
    rows = torch.as_tensor(rows, dtype=torch.long)
    cols = torch.as_tensor(cols, dtype=torch.long)
    p32 = [t.detach().to(f32) for t in (means, scales, quats, opacity)]
    fp = footprint(*p32, fr, st)
    keep = fp["ok"] & (fp["b"] >= st.weight_min)
    idx = torch.nonzero(keep).flatten()
    xl32 = (cols.to(f32) - torch.tensor(fr.col_offset, dtype=f32)) * torch.tensor(fr.delta_lat, dtype=f32)
    xa32 = rows.to(f32) * torch.tensor(fr.delta_ax, dtype=f32) + torch.tensor(fr.near, dtype=f32)
    C = torch.tensor(EXP2_SCALE, dtype=f32)
    g = {k: v[idx][None, :] for k, v in fp.items() if k != "ok"}
    c00, c01, c11 = g["i00"] * C, g["i01"] * C, g["i11"] * C
    dl32 = xl32[:, None] - g["m0"]
    da32 = xa32[:, None] - g["m1"]
    box = (torch.abs(dl32) <= g["r_lat"]) & (torch.abs(da32) <= g["r_ax"])
    l32 = _fma32(c00, dl32, c01 * da32)
    k32 = _fma32(c01, dl32, c11 * da32)
    e32 = _fma32(dl32, l32, da32 * k32)
    w32 = g["b"] * torch.exp2(e32)
    m = box & (e32 <= 0) & (w32 >= torch.tensor(st.weight_min, dtype=f32))  # (n, len(idx))
    if not fp64:
        inten = intensity.detach().to(f32)[idx][None, :]
        terms = torch.where(m, inten * w32, torch.zeros_like(w32))
        out = torch.zeros(rows.numel(), dtype=f32)
        for j in range(terms.shape[1]):  # sequential FP32 sum in index order
            out = out + terms[:, j]
        return out
    used = torch.nonzero(m.any(dim=0)).flatten()
    gi = idx[used]
    sub = footprint(means[gi], scales[gi], quats[gi], opacity[gi], fr, st)
    mm = m[:, used]
    # pixel positions as the kernels form them (FP32), then float64
    dl = xl32.double()[:, None] - sub["m0"][None, :]
    da = xa32.double()[:, None] - sub["m1"][None, :]
    q = dl * (sub["i00"] * dl + sub["i01"] * da) + da * (sub["i01"] * dl + sub["i11"] * da)
    w = sub["b"][None, :] * torch.exp(-0.5 * torch.where(mm, q, torch.zeros_like(q)))
    return torch.where(mm, intensity[gi][None, :] * w, torch.zeros_like(w)).sum(dim=1)


def sanitize_grad(g: torch.Tensor) -> torch.Tensor:
    """us_focus_sanitize_grad (us_focus.cu:129-137): a non-finite value becomes 0, a magnitude above kGradClamp
    becomes +-kGradClamp (copysign). The reference applies it to every per-Gaussian gradient of each render call,
    i.e. of each frame (us_focus.cu:694-752)."""
    g = torch.where(torch.isfinite(g), g, torch.zeros_like(g))
    return torch.where(g.abs() > GRAD_CLAMP, torch.sign(g) * GRAD_CLAMP, g)


def frame_gradients(params64, frame_losses):
    """The reference's gradients of a multi-frame loss: sum over frames of the sanitised gradient of each frame's
    own loss (one render call per frame in the reference). params64: the float64 leaf tensors; frame_losses: a
    list of callables, each returning one frame's scalar loss from params64. Returns one tensor per parameter."""
    total = [torch.zeros_like(p) for p in params64]
    for fl in frame_losses:
        gs = torch.autograd.grad(fl(), params64, allow_unused=True)
        for t, g in zip(total, gs):
            if g is not None:
                t += sanitize_grad(g)
    return total
