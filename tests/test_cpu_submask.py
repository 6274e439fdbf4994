"""Conservativeness of the splat kernels' sub-block skipping (CPU only).

The forward and the adjoint skip, per 8x8 sub-block, the Gaussians whose plane section cannot reach the
sub-block's pixels (`sub_mask` with the per-pair setup of `pair_section`, ug.hip). This test restates those
steps in FP32 operation by operation (FMAs as single roundings of the exact product-sum) on random oblique
frames, strongly anisotropic Gaussians placed up to 1.3 footprints away from a pixel of the sub-block and,
for half the cases, pixels jittered out of the plane, then evaluates q for all 64 pixels as the kernels do
(`quad_form`). A sub-block holding a pixel with q <= chi2 must never be skipped. It also checks that the
exact section test is tighter than the bounding box of the section ellipse it replaced.
"""

import numpy as np
import pytest

f32 = np.float32
CHI2 = 7.815


def _fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(f32)


def _dot3(ax, ay, az, bx, by, bz):
    return _fma(az, bz, _fma(ay, by, (ax * bx).astype(f32)))


def _rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def _cases(seed, n, off_plane):
    rng = np.random.default_rng(seed)
    axes = _rotations(rng, n)
    al, aa, ae = (axes[:, :, k].astype(f32) for k in range(3))
    origin = rng.uniform(-150, 150, size=(n, 3)).astype(f32)
    dl = rng.uniform(0.05, 0.5, n).astype(f32)
    da = rng.uniform(0.05, 0.5, n).astype(f32)
    col = rng.uniform(0, 300, n).astype(f32)
    near = rng.uniform(0, 20, n).astype(f32)
    bx, by = rng.integers(0, 60, n), rng.integers(0, 60, n)
    ix = (bx[:, None] * 8 + (np.arange(64) % 8)[None]).astype(f32)
    iy = (by[:, None] * 8 + (np.arange(64) // 8)[None]).astype(f32)
    # pixel_position of the linear grid
    lat = ((ix - col[:, None]).astype(f32) * dl[:, None]).astype(f32)
    ax = (near[:, None] + (iy * da[:, None]).astype(f32)).astype(f32)
    P = [((origin[:, None, k] + (lat * al[:, None, k]).astype(f32)).astype(f32)
          + (ax * aa[:, None, k]).astype(f32)).astype(f32) for k in range(3)]
    if off_plane:
        jit = (rng.uniform(-0.3, 0.3, size=(n, 64)) * (rng.uniform(size=(n, 1)) < 0.5)).astype(f32)
        P = [(P[k] + (jit * ae[:, None, k]).astype(f32)).astype(f32) for k in range(3)]
    # sub_bounds
    dx, dy, dz = ((P[k] - origin[:, None, k]).astype(f32) for k in range(3))
    plat = _dot3(dx, dy, dz, al[:, None, 0], al[:, None, 1], al[:, None, 2])
    pax = _dot3(dx, dy, dz, aa[:, None, 0], aa[:, None, 1], aa[:, None, 2])
    pele = _dot3(dx, dy, dz, ae[:, None, 0], ae[:, None, 1], ae[:, None, 2])
    lo_l, hi_l, lo_a, hi_a = plat.min(1), plat.max(1), pax.min(1), pax.max(1)
    lo_e, hi_e = pele.min(1), pele.max(1)
    m = np.maximum(np.maximum(np.abs(lo_l), np.abs(hi_l)), np.maximum(np.abs(lo_a), np.abs(hi_a)))
    pb = (f32(1e-3) + (f32(4e-6) * m).astype(f32)).astype(f32)
    lo_l, hi_l, lo_a, hi_a, lo_e, hi_e = (v.astype(f32) for v in (lo_l - pb, hi_l + pb, lo_a - pb, hi_a + pb,
                                                                    lo_e - pb, hi_e + pb))
    # Gaussians near the sub-block
    sc = np.exp(rng.uniform(np.log(0.03), np.log(8.0), size=(n, 3)))
    rg = _rotations(rng, n)
    lam = rg @ (np.eye(3)[None] / (sc[:, None, :] ** 2)) @ rg.transpose(0, 2, 1)
    Lc = np.linalg.cholesky(lam)
    l11, l12, l13, l22, l23, l33 = (Lc[:, i, j].astype(f32) for i, j in ((0, 0), (1, 0), (2, 0), (1, 1), (2, 1),
                                                                           (2, 2)))
    k = rng.integers(0, 64, n)
    cen = np.stack([P[j][np.arange(n), k] for j in range(3)], 1).astype(np.float64)
    dirn = rng.normal(size=(n, 3))
    dirn /= np.linalg.norm(dirn, axis=1, keepdims=True)
    mu = (cen + dirn * (np.sqrt(CHI2) * sc.max(1) * rng.uniform(0, 1.3, n))[:, None]).astype(f32)
    # quad_form at every pixel
    ddx, ddy, ddz = ((P[j] - mu[:, None, j]).astype(f32) for j in range(3))
    y0 = _fma(l13[:, None], ddz, _fma(l12[:, None], ddy, (l11[:, None] * ddx).astype(f32)))
    y1 = _fma(l23[:, None], ddz, (l22[:, None] * ddy).astype(f32))
    y2 = (l33[:, None] * ddz).astype(f32)
    contrib = (_fma(y2, y2, _fma(y1, y1, (y0 * y0).astype(f32))) <= f32(CHI2)).any(1)
    # pair_section (FP64 setup)
    vx, vy, vz = ((mu[:, j] - origin[:, j]).astype(f32) for j in range(3))
    lat_g = _dot3(vx, vy, vz, al[:, 0], al[:, 1], al[:, 2])
    ax_g = _dot3(vx, vy, vz, aa[:, 0], aa[:, 1], aa[:, 2])
    ele_g = _dot3(vx, vy, vz, ae[:, 0], ae[:, 1], ae[:, 2])
    Ld = [v.astype(np.float64) for v in (l11, l12, l13, l22, l23, l33)]

    def rows(v):
        v = v.astype(np.float64)
        return (v[:, 0] * Ld[0] + v[:, 1] * Ld[1] + v[:, 2] * Ld[2], v[:, 1] * Ld[3] + v[:, 2] * Ld[4],
                v[:, 2] * Ld[5])

    A, B, C = rows(al), rows(aa), rows(ae)
    M00, M01, M11 = (sum(x * y for x, y in zip(U, V)) for U, V in ((A, A), (A, B), (B, B)))
    b0, b1, cz = (sum(x * y for x, y in zip(U, V)) for U, V in ((A, C), (B, C), (C, C)))
    det = M00 * M11 - M01 * M01
    regular = det > 1e-8 * M00 * M11
    w0, w1 = (M11 * b0 - M01 * b1) / det, (M00 * b1 - M01 * b0) / det
    kk = np.maximum(0.0, (cz - (b0 * w0 + b1 * w1)) * (1 - 1e-5)).astype(f32)
    u_lat, u_ax, t_x, t_y = (-w0).astype(f32), (-w1).astype(f32), (M01 / M11).astype(f32), (M01 / M00).astype(f32)
    p = np.sqrt(M00)
    q_, s_, p = (M01 / p).astype(f32), np.sqrt(det / M00).astype(f32), p.astype(f32)
    # sub_mask, section branch
    sq = f32(np.sqrt(f32(CHI2)))
    chi2w = ((sq * sq).astype(f32) * f32(1.002)).astype(f32)
    pad = (f32(1e-3) + (f32(4e-6) * (np.abs(lat_g) + np.abs(ax_g)).astype(f32)).astype(f32)).astype(f32)
    d_lo, d_hi = (lo_e - ele_g).astype(f32), (hi_e - ele_g).astype(f32)
    d0 = np.minimum(np.maximum(f32(0), d_lo), d_hi)
    rem = (chi2w - ((kk * d0).astype(f32) * d0).astype(f32)).astype(f32)
    cl0, cl1 = ((lat_g + (u_lat * d).astype(f32)).astype(f32) for d in (d_lo, d_hi))
    ca0, ca1 = ((ax_g + (u_ax * d).astype(f32)).astype(f32) for d in (d_lo, d_hi))
    x0 = ((lo_l - np.maximum(cl0, cl1)).astype(f32) - pad).astype(f32)
    x1 = ((hi_l - np.minimum(cl0, cl1)).astype(f32) + pad).astype(f32)
    y0_ = ((lo_a - np.maximum(ca0, ca1)).astype(f32) - pad).astype(f32)
    y1_ = ((hi_a - np.minimum(ca0, ca1)).astype(f32) + pad).astype(f32)
    xn, yn = np.minimum(np.maximum(f32(0), x0), x1), np.minimum(np.maximum(f32(0), y0_), y1_)
    ya = np.minimum(np.maximum((-t_x * xn).astype(f32), y0_), y1_)
    xb = np.minimum(np.maximum((-t_y * yn).astype(f32), x0), x1)
    u1, v1 = ((p * xn).astype(f32) + (q_ * ya).astype(f32)).astype(f32), (s_ * ya).astype(f32)
    u2, v2 = ((p * xb).astype(f32) + (q_ * yn).astype(f32)).astype(f32), (s_ * yn).astype(f32)
    f = np.minimum(((u1 * u1).astype(f32) + (v1 * v1).astype(f32)).astype(f32),
                   ((u2 * u2).astype(f32) + (v2 * v2).astype(f32)).astype(f32))
    hit = (rem > 0) & (f <= (rem * f32(1.004)).astype(f32))
    # the bounding box of the section ellipse (the earlier form of the test), for comparison
    i00 = ((M11 / det).astype(f32) * f32(1.0001)).astype(f32)
    i11 = ((M00 / det).astype(f32) * f32(1.0001)).astype(f32)
    hl = (np.sqrt(np.maximum(rem, 0) * i00) * f32(1.002) + pad).astype(f32)
    ha = (np.sqrt(np.maximum(rem, 0) * i11) * f32(1.002) + pad).astype(f32)
    box = ((rem > 0) & (np.maximum(cl0, cl1) + hl >= lo_l) & (np.minimum(cl0, cl1) - hl <= hi_l)
           & (np.maximum(ca0, ca1) + ha >= lo_a) & (np.minimum(ca0, ca1) - ha <= hi_a))
    return contrib[regular], hit[regular], box[regular]


@pytest.mark.parametrize("off_plane", [False, True])
def test_sub_block_section_test_never_drops_a_contributing_sub_block(off_plane):
    contrib, hit, box = _cases(seed=3 + off_plane, n=150_000, off_plane=off_plane)
    assert contrib.sum() > 30_000  # the cases straddle the footprint boundary
    assert int((contrib & ~hit).sum()) == 0
    assert not np.any(hit & ~box)           # never looser than the bounding box
    assert hit.sum() < 0.95 * box.sum()     # and tighter in aggregate
