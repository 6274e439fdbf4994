"""Pins of the oracle itself (CPU only): independent implementations, closed forms, brute force and finite
differences, chosen so that a dropped term, a wrong sign or index, or a transposed operand fails."""

import math

import numpy as np
import pytest
import torch

from oracle import fallback as ofb
from oracle import ug as oug
from oracle import ur as our
from tests.ug_helpers import rel_l2
from tests.ur_helpers import axis_quat as _axis_quat, composed_rotation as _composed_rotation, rodrigues as _rodrigues
from us3d.scenes import linear_pixels, make_scene, precision_tril_from_scales_quats, quat_to_rot, straight_sweep


def _small(n=512, seed=3, H=48, W=40):
    scene = make_scene(n, 64.0, seed=seed)
    fr = straight_sweep(3, H, W, 0.3, 64.0, 10.0, 5.0, seed=seed + 1)[1]
    return scene, fr, oug.FrameGeom.from_spec(fr), linear_pixels(fr).numpy()


def test_fallback_semantics_match_independent_torch_fallback():
    """chi2 = 120 / no alpha cut restatement vs the PyTorch-fallback restatement (einsum, autograd), FP64."""
    scene, fr, geom, pix = _small()
    sem = oug.Semantics.fallback()
    m, t = scene.means.numpy(), scene.tril.numpy()
    mr = oug.map_frame(m, t, geom, chi2=sem.chi2)
    lists = oug.pixel_lists(m, t, pix, geom, mr, chi2=sem.chi2)
    args = (m, t, scene.opacity.numpy(), scene.intensity.numpy(), pix, geom, lists, sem)
    img, alpha, _d, _s = oug.forward(*args, fp64=True)
    ts = [x.double().requires_grad_(True) for x in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    fimg, falpha = ofb.render(*ts, linear_pixels(fr).double(), fr.height, fr.width)
    assert rel_l2(img, fimg.detach().numpy()) < 1e-9
    assert rel_l2(alpha, falpha.detach().numpy()) < 1e-9
    dI = np.random.default_rng(1).normal(size=img.size).astype(np.float32)
    (fimg.reshape(-1) * torch.from_numpy(dI).double()).sum().backward()
    ref = oug.backward_f64(*args, dI, gate=True)
    for g, x in zip(ref, ts):
        assert rel_l2(g, x.grad.numpy()) < 1e-8


def test_mapper_invariants_and_stable_order():
    scene, fr, geom, _pix = _small(n=2048, H=128, W=128)
    mr = oug.map_frame(scene.means.numpy(), scene.tril.numpy(), geom)
    K = mr.keys.size
    assert K == int(mr.count.sum()) and K > 500
    assert np.all(np.diff(mr.keys.astype(np.uint64)) >= 0)
    tile = (mr.keys >> np.uint64(32)).astype(np.int64)
    # within equal keys, Gaussian index ascends (stable sort of the emit order)
    same = mr.keys[1:] == mr.keys[:-1]
    assert np.all(mr.vals[1:][same] > mr.vals[:-1][same])
    # ranges partition [0, K)
    nz = mr.tile_start >= 0
    assert np.array_equal(np.sort(mr.tile_start[nz]), mr.tile_start[nz][np.argsort(mr.tile_start[nz])])
    assert int((mr.tile_end[nz] - mr.tile_start[nz]).sum()) == K
    for tt in np.unique(tile):
        idx = np.nonzero(tile == tt)[0]
        assert mr.tile_start[tt] == idx[0] and mr.tile_end[tt] == idx[-1] + 1
    # each (tile, Gaussian) once, and exactly the tiles of each Gaussian's rectangle
    pairs = set(zip(tile.tolist(), mr.vals.tolist()))
    assert len(pairs) == K
    tiles_x = -(-fr.width // 16)
    for g in np.nonzero(mr.count)[0][:200]:
        x0, y0, x1, y1 = mr.rect[g]
        for y in range(y0, y1):
            for x in range(x0, x1):
                assert (y * tiles_x + x, g) in pairs


def test_mapper_radii_closed_form_and_conservative_cover():
    """Tile rectangles follow the closed-form marginal radius sqrt(chi2 a^T Lambda^-1 a) (FP64); every
    (pixel, Gaussian) pair with q <= chi2 (brute force, FP64) is in the pixel's candidate list."""
    scene, fr, geom, pix = _small(n=300, H=64, W=64)
    chi2, beta = 7.815, 1e-4
    m, t = scene.means.numpy().astype(np.float64), scene.tril.numpy().astype(np.float64)
    mr = oug.map_frame(scene.means.numpy(), scene.tril.numpy(), geom)
    L = np.zeros((len(m), 3, 3))
    L[:, 0, 0] = t[:, 0] ** 2 + beta
    L[:, 1, 0], L[:, 2, 0] = t[:, 1], t[:, 2]
    L[:, 1, 1] = t[:, 3] ** 2 + beta
    L[:, 2, 1] = t[:, 4]
    L[:, 2, 2] = t[:, 5] ** 2 + beta
    lam = L @ L.transpose(0, 2, 1)
    cov = np.linalg.inv(lam)
    a = {k: getattr(geom, k).astype(np.float64) for k in ("axis_lat", "axis_ax", "axis_ele")}
    r = {k: np.sqrt(chi2 * np.einsum("i,gij,j->g", v, cov, v)) for k, v in a.items()}
    v = m - geom.origin.astype(np.float64)
    ele = v @ a["axis_ele"]
    u = (v @ a["axis_lat"]) / geom.delta_lat + geom.col_offset
    vv = (v @ a["axis_ax"] - geom.near_mm) / geom.delta_ax
    ru = np.maximum(1.0, r["axis_lat"] / geom.delta_lat)
    rv = np.maximum(1.0, r["axis_ax"] / geom.delta_ax)
    bounds = np.stack([(u - ru) / 16, (vv - rv) / 16, (u + ru) / 16, (vv + rv) / 16], 1)
    near_edge = np.abs(bounds - np.round(bounds)).min(1) < 1e-4
    exp_rect = np.stack([np.floor(bounds[:, 0]), np.floor(bounds[:, 1]), np.ceil(bounds[:, 2]), np.ceil(bounds[:, 3])], 1)
    exp_rect = np.stack([np.maximum(0, exp_rect[:, 0]), np.maximum(0, exp_rect[:, 1]), np.minimum(4, exp_rect[:, 2]),
                         np.minimum(4, exp_rect[:, 3])], 1)
    vis = (np.abs(ele) <= r["axis_ele"]) & (exp_rect[:, 0] < exp_rect[:, 2]) & (exp_rect[:, 1] < exp_rect[:, 3])
    ok = ~near_edge & (np.abs(np.abs(ele) - r["axis_ele"]) > 1e-5)
    assert ok.sum() > 250
    assert np.array_equal(vis[ok], mr.count[ok] > 0)
    assert np.array_equal(exp_rect[ok & vis].astype(np.int32), mr.rect[ok & vis])
    counts, offsets, indices = oug.pixel_lists(scene.means.numpy(), scene.tril.numpy(), pix, geom, mr)
    d = pix.astype(np.float64)[:, None, :] - m[None, :, :]
    q = np.einsum("pgi,gij,pgj->pg", d, lam, d)
    hits = np.argwhere(q <= chi2 * (1 - 1e-6))
    assert len(hits) > 1000
    for p, g in hits:
        assert g in indices[offsets[p]:offsets[p] + counts[p]]


def test_single_gaussian_closed_form():
    """One isotropic Gaussian centred on a pixel: image = o c / (o + bg), alpha = o + bg, depth = elevation."""
    fr = straight_sweep(1, 16, 16, 0.5, 64.0, 0.0, 0.0, seed=0)[0]
    pix = linear_pixels(fr).numpy()
    centre = pix[8 * 16 + 8].astype(np.float32)
    means = centre[None, :]
    tril = np.array([[1.0, 0, 0, 1.0, 0, 1.0]], np.float32)
    o, c = np.array([0.7], np.float32), np.array([0.3], np.float32)
    geom = oug.FrameGeom.from_spec(fr)
    mr = oug.map_frame(means, tril, geom)
    lists = oug.pixel_lists(means, tril, pix, geom, mr)
    img, alpha, depth, _ = oug.forward(means, tril, o, c, pix, geom, lists, oug.Semantics(), fp64=True)
    p = 8 * 16 + 8
    of, cf, bg = float(np.float32(0.7)), float(np.float32(0.3)), float(np.float32(1e-6))
    assert img[p] == pytest.approx(of * cf / (of + bg), rel=1e-12)
    assert alpha[p] == pytest.approx(of + bg, rel=1e-12)
    e = float(np.dot(centre.astype(np.float64) - geom.origin, geom.axis_ele))
    assert depth[p] == pytest.approx(e, abs=1e-6)
    # a neighbour one pixel (0.5 mm) away along x sees w = exp(-q/2) with q = ((1+1e-4) * 0.5)^2
    w = 0.7 * math.exp(-0.5 * ((1 + 1e-4) * 0.5) ** 2)
    assert alpha[p + 1] == pytest.approx(w + 1e-6, rel=1e-6)


def test_cuda_semantics_adjoint_matches_finite_differences():
    scene, fr, geom, pix = _small(n=96, H=32, W=32)
    sem = oug.Semantics()
    m, t = scene.means.numpy(), scene.tril.numpy()
    mr = oug.map_frame(m, t, geom)
    lists = oug.pixel_lists(m, t, pix, geom, mr)
    params = [x.numpy().astype(np.float64) for x in (scene.means, scene.tril, scene.opacity, scene.intensity)]
    rng = np.random.default_rng(0)
    P = pix.shape[0]
    dI, dA, dD = (rng.normal(size=P).astype(np.float32) * s for s in (1.0, 0.1, 0.01))
    grads = oug.backward_f64(*params, pix, geom, lists, sem, dI, dA, dD, gate=True)

    def loss(ps):
        img, alpha, depth, _ = oug.forward(*ps, pix, geom, lists, sem, fp64=True)
        return float(img @ dI + alpha @ dA + depth @ dD)

    for k, g in enumerate(grads):
        flat = g.reshape(-1)
        for j in np.argsort(-np.abs(flat))[:5]:
            h = 1e-6 * max(1.0, abs(params[k].reshape(-1)[j]))
            pp = [p.copy() for p in params]
            pp[k].reshape(-1)[j] += h
            pm = [p.copy() for p in params]
            pm[k].reshape(-1)[j] -= h
            fd = (loss(pp) - loss(pm)) / (2 * h)
            assert fd == pytest.approx(flat[j], rel=1e-5, abs=1e-9)


def test_scene_precision_factor_roundtrip():
    scene = make_scene(64, 10.0, seed=7)
    R = quat_to_rot(scene.quats.double())
    lam = R @ torch.diag_embed(1.0 / scene.scales.double() ** 2) @ R.transpose(-1, -2)
    t = precision_tril_from_scales_quats(scene.scales, scene.quats).double()
    L = torch.zeros((64, 3, 3), dtype=torch.float64)
    L[:, 0, 0] = t[:, 0] ** 2 + 1e-4
    L[:, 1, 0], L[:, 2, 0], L[:, 2, 1] = t[:, 1], t[:, 2], t[:, 4]
    L[:, 1, 1] = t[:, 3] ** 2 + 1e-4
    L[:, 2, 2] = t[:, 5] ** 2 + 1e-4
    assert torch.allclose(L @ L.transpose(-1, -2), lam, rtol=1e-5, atol=1e-6)


# ---------------------------------------------------------------- UltraG-Ray oracle


def _ray_case(n=64, W=8, H=48, **kw):
    from tests.ur_helpers import oracle_frame, ray_cfg, ray_settings, small_ray_case

    scene, frames = small_ray_case(n=n, n_frames=1, W=W, H=H)
    fr = frames[0]
    cfg = ray_cfg(fr, grad_threshold=0.0, **kw)
    geo = oracle_frame(fr, cfg)
    st = ray_settings(cfg)
    params = [x.numpy().astype(np.float64) for x in scene.params()]
    return params, geo, st, our.map_frame(params[0], params[1], geo, st)


@pytest.mark.parametrize("kw", [{}, {"quadrature_window_radius_mm": 0.0}])
def test_ray_adjoint_matches_finite_differences_with_frozen_windows(kw):
    params, geo, st, rm = _ray_case(**kw)
    assert rm.indices.size > 40
    dB = np.random.default_rng(0).normal(size=geo.W * geo.H).astype(np.float32)
    grads = our.backward_f64(params, geo, rm, st, dB)
    tc = our.window_centres(params, geo)

    def loss(ps):
        return float(our.forward(ps, geo, rm, st, fp64=True, tc_frozen=tc) @ dB.astype(np.float64))

    for k, g in enumerate(grads):
        flat = g.reshape(-1)
        for j in np.argsort(-np.abs(flat))[:5]:
            h = 1e-6 * max(1.0, abs(params[k].reshape(-1)[j]))
            pp = [p.copy() for p in params]
            pp[k].reshape(-1)[j] += h
            pm = [p.copy() for p in params]
            pm[k].reshape(-1)[j] -= h
            fd = (loss(pp) - loss(pm)) / (2 * h)
            assert fd == pytest.approx(flat[j], rel=1e-5, abs=1e-8)


def test_ray_special_cases():
    params, geo, st, rm = _ray_case()
    # tau = 1 (raw -> +inf): transmission is 1 everywhere, B = (1 - e^-S) * E
    p_open = [p.copy() for p in params]
    p_open[3][:] = 60.0
    img, aux = our.forward(p_open, geo, rm, st)
    assert np.allclose(aux[:, 0], 1.0)
    # no hits on a ray -> 0
    far = [p.copy() for p in params]
    far[0][:] = -1e4
    rm_far = our.map_frame(far[0], far[1], geo, st)
    assert rm_far.indices.size == 0
    img_far = our.forward(far, geo, rm_far, st, fp64=True)
    assert np.all(img_far == 0.0)


def test_gl3_transmission_of_one_gaussian():
    """One unit isotropic Gaussian centred on a ray at depth 30 mm, tau = 1/2, window R = 2: a sample at 40 mm
    sees T = 1/2 + 1/2 e^-psi with psi the 3-point Gauss-Legendre rule on [28, 32] --
    2 (5/9 e^{-(2 xi)^2/2} * 2 + 8/9) -- itself within 3% of the exact sqrt(2 pi) erf(sqrt 2)."""
    W, H = 1, 2
    geo = our.RayFrame(origin=np.zeros(3, np.float32), axis_lat=np.array([1, 0, 0], np.float32),
                       axis_ax=np.array([0, 1, 0], np.float32), axis_ele=np.array([0, 0, 1], np.float32), W=W, H=H,
                       width_mm=1.0, near_mm=0.0, col_offset=0.0,
                       pixels=np.array([[[0, 40, 0], [0, 29, 0]]], np.float32), ray_o=np.zeros((1, 3), np.float32),
                       ray_d=np.array([[0, 1, 0]], np.float32), delta=np.zeros(1, np.float32))
    params = [np.array([[0, 30, 0]], np.float64), np.ones((1, 3)), np.array([[1.0, 0, 0, 0]]), np.zeros(1),
              np.array([[0.4, 0, 0, 0]])]
    st = our.RaySettings(k_sigma=3.0, Rw=2.0)
    rm = our.map_frame(params[0], params[1], geo, st)
    _img, aux = our.forward(params, geo, rm, st)
    xi = 0.7745966692
    psi = 2.0 * (0.5555555556 * math.exp(-0.5 * (2 * xi) ** 2) * 2 + 0.8888888889)
    # aux is indexed by y * W + x: sample (x=0, y=0) at depth 40 mm is past the window, (0, 1) at 29 mm is in
    # front of the Gaussian centre, so it is not attenuated
    assert aux[0, 0] == pytest.approx(0.5 + 0.5 * math.exp(-psi), rel=1e-6)
    assert aux[1, 0] == 1.0
    exact = math.sqrt(2 * math.pi) * math.erf(math.sqrt(2))
    assert abs(psi - exact) / exact < 0.03


# ---------------------------------------------------------------- UltraG-Ray rotation / covariance conventions
# The reference's covariance.cuh (quaternion -> R, Sigma^-1 from (s, R)) is absent from its tree. What fixes the
# conventions: the repo-wide quaternion layout (w, x, y, z) of gaussianplie_renderer/include/shader/shared.cuh:15-34
# and reference.py:7-23; the transmission integral's whitening o_local = R^T (o - mu) / s
# (ultragray_renderer/include/shader/shared.cuh:37-48), which makes R's columns the principal axes; and the echo
# weight must describe the same Gaussian. These checks build R independently (Rodrigues' formula and the Hamilton
# product) and use a rotation that is not symmetric, so an R <-> R^T swap or an echo covariance that disagrees with
# the whitening fails them.


def _one_ray_frame(pixels, ro, rd, axes):
    al, aa, ae = axes
    H = pixels.shape[0]
    return our.RayFrame(origin=np.zeros(3, np.float32), axis_lat=al.astype(np.float32),
                        axis_ax=aa.astype(np.float32), axis_ele=ae.astype(np.float32), W=1, H=H, width_mm=1.0,
                        near_mm=0.0, col_offset=0.0, pixels=pixels.reshape(1, H, 3).astype(np.float32),
                        ray_o=ro.reshape(1, 3).astype(np.float32), ray_d=rd.reshape(1, 3).astype(np.float32),
                        delta=np.zeros(1, np.float32))


def test_ray_echo_weight_is_the_gaussian_with_principal_axes_the_columns_of_R():
    """One anisotropic Gaussian (s = 3, 1, 0.5 mm) under the composed rotation, tau = 1 (transmission 1): the echo
    weight the oracle accumulates (aux den) at sample p equals exp(-1/2 |S^-1 R^T (p - mu)|^2) -- the whitening of
    shared.cuh:37-48 -- at points along each principal axis R e_k (where it is exp(-t^2 / 2 s_k^2)) and at
    random points within the echo range."""
    q, R = _composed_rotation()
    s = np.array([3.0, 1.0, 0.5])
    mu = np.array([0.0, 30.0, 0.0])
    rng = np.random.default_rng(3)
    pts = [mu + t * R[:, k] for k in range(3) for t in (-1.5 * s[k], -0.5 * s[k], 0.7 * s[k], 1.2 * s[k])]
    pts += [mu + rng.uniform(-2.0, 2.0, 3) for _ in range(12)]
    pts = np.array(pts)
    axes = (np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]))
    geo = _one_ray_frame(pts, np.zeros(3), axes[1], axes)
    params = [mu[None], s[None], q[None], np.array([60.0]), np.array([[0.5, 0, 0, 0]])]
    st = our.RaySettings(k_sigma=3.0, Rw=2.0)
    rm = our.map_frame(params[0], params[1], geo, st)
    assert rm.indices.tolist() == [0]
    _img, aux = our.forward(params, geo, rm, st)
    z = (pts - mu) @ R / s  # S^-1 R^T (p - mu), row per point
    want = np.exp(-0.5 * (z * z).sum(1))
    assert np.allclose(aux[:, 0], 1.0)
    np.testing.assert_allclose(aux[:, 2], want, rtol=2e-6, atol=1e-12)
    for k in range(3):  # on the principal axes: the 1-D Gaussian of that axis's scale
        t = np.array([-1.5, -0.5, 0.7, 1.2]) * s[k]
        np.testing.assert_allclose(aux[4 * k:4 * k + 4, 2], np.exp(-0.5 * (t / s[k]) ** 2), rtol=2e-6)
    # the transposed rotation describes a different Gaussian at these points
    zt = (pts - mu) @ R.T / s
    assert np.abs(np.exp(-0.5 * (zt * zt).sum(1)) - want).max() > 0.1


def test_ray_transmission_whitening_uses_the_same_principal_axes():
    """A ray through the centre of the same Gaussian along its long principal axis R e_0 (s_0 = 3 mm), tau = 1/2,
    window R_w = 2 mm: along the ray the whitened distance is (t - t_c) / s_0, so a sample past the window sees
    T = 1/2 + 1/2 e^-psi with psi the 3-point Gauss-Legendre rule 2 (5/9 e^{-(2 xi / 3)^2 / 2} 2 + 8/9); a sample
    in front of the centre is not attenuated."""
    q, R = _composed_rotation()
    s = np.array([3.0, 1.0, 0.5])
    d = R[:, 0]
    e1 = np.cross(d, [0.0, 0.0, 1.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(e1, d)
    mu = 30.0 * d
    geo = _one_ray_frame(np.array([40.0 * d, 29.0 * d]), np.zeros(3), d, (e1, d, e2))
    params = [mu[None], s[None], q[None], np.array([0.0]), np.array([[0.5, 0, 0, 0]])]
    st = our.RaySettings(k_sigma=3.0, Rw=2.0)
    rm = our.map_frame(params[0], params[1], geo, st)
    _img, aux = our.forward(params, geo, rm, st)
    xi = 0.7745966692
    psi = 2.0 * (0.5555555556 * math.exp(-0.5 * (2 * xi / 3.0) ** 2) * 2 + 0.8888888889)
    assert aux[0, 0] == pytest.approx(0.5 + 0.5 * math.exp(-psi), rel=1e-6)
    assert aux[1, 0] == 1.0
    # whitened with R in place of R^T, the ray's direction would be S^-1 R d, whose length is not 1/s_0, so psi
    # would differ
    assert abs(np.linalg.norm(R @ d / s) - 1.0 / 3.0) > 0.01


def test_ray_quaternion_sign_and_scale_invariance_and_gradient_orthogonality():
    """R(q) = R(-q) = R(2q) (a rotation is fixed by the unit quaternion up to sign): the FP64 images are identical,
    and since the image is invariant under q -> lambda q its gradient satisfies dL/dq . q = 0,
    dL/d(-q) = -dL/dq and dL/d(2q) = dL/dq / 2."""
    params, geo, st, rm = _ray_case(n=48)
    dB = np.random.default_rng(5).normal(size=geo.W * geo.H).astype(np.float32)
    base = our.forward(params, geo, rm, st, fp64=True)
    g = our.backward_f64(params, geo, rm, st, dB)
    assert np.abs(g[2]).max() > 0
    for lam in (-1.0, 2.0):
        p2 = [p.copy() for p in params]
        p2[2] = p2[2] * lam
        assert np.array_equal(our.forward(p2, geo, rm, st, fp64=True), base)
        g2 = our.backward_f64(p2, geo, rm, st, dB)
        np.testing.assert_allclose(g2[2], g[2] / lam, rtol=1e-9, atol=1e-12 * np.abs(g[2]).max())
    dots = np.abs((g[2] * params[2]).sum(1))
    assert dots.max() <= 1e-9 * np.abs(g[2]).max() * np.abs(params[2]).max()


def test_oracle_quaternion_matrix_is_the_rodrigues_rotation():
    """The echo weight along many directions pins the full matrix: for the composed rotation the oracle's weight
    equals the closed form built from Rodrigues' R only (checked above); here the same holds for 20 random
    rotations given as unnormalised quaternions (q = 1.7 * (cos a/2, sin a/2 n)) against R = exp(a [n]x)."""
    rng = np.random.default_rng(8)
    axes = (np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0]))
    s = np.array([2.5, 1.2, 0.6])
    mu = np.array([0.0, 30.0, 0.0])
    st = our.RaySettings(k_sigma=3.0, Rw=2.0)
    for _ in range(20):
        n = rng.normal(size=3)
        ang = rng.uniform(10, 170)
        q = 1.7 * _axis_quat(n, ang)
        R = _rodrigues(n, ang)
        pts = mu + rng.uniform(-2.0, 2.0, (8, 3))
        geo = _one_ray_frame(pts, np.zeros(3), axes[1], axes)
        params = [mu[None], s[None], q[None], np.array([60.0]), np.array([[0.5, 0, 0, 0]])]
        rm = our.map_frame(params[0], params[1], geo, st)
        _img, aux = our.forward(params, geo, rm, st)
        z = (pts - mu) @ R / s
        np.testing.assert_allclose(aux[:, 2], np.exp(-0.5 * (z * z).sum(1)), rtol=2e-5, atol=1e-12)


# ---------------------------------------------------------------- plane form of q (frames on their own grids)


def test_plane_form_is_the_mahalanobis_distance_on_the_plane():
    """The oracle's plane form (cx, cy, K, p, qq, s) restates the kernels' reparametrisation of q on a frame's pixel
    plane. Pinned by closed forms of a Gaussian restricted to a plane with unit normal n = a_ele: the minimum of the
    Mahalanobis distance over the plane is K = ele^2 / (n^T Sigma n) (ele = (mu - o).n), attained at the conditional
    point x* = mu - Sigma n ele / (n^T Sigma n), so (cx, cy) = ((x* - o).a_lat, (x* - o).a_ax); and at any point of
    the plane q equals (x - mu)^T Lambda (x - mu) computed directly in FP64."""
    rng = np.random.default_rng(17)
    scene = make_scene(300, 30.0, seed=19)
    fr = straight_sweep(3, 64, 64, 0.3, 30.0, 10.0, 8.0, seed=20)[1]
    geom = oug.FrameGeom.from_spec(fr)
    means, tril = scene.means.numpy(), scene.tril.numpy()
    plane = oug.plane_form(means, tril, geom, beta=1e-4)
    o, al, aa, ae = (np.asarray(v, np.float64) for v in (geom.origin, geom.axis_lat, geom.axis_ax, geom.axis_ele))
    t = tril.astype(np.float64)
    L = np.zeros((len(means), 3, 3))
    L[:, 0, 0], L[:, 1, 0], L[:, 2, 0] = t[:, 0] ** 2 + 1e-4, t[:, 1], t[:, 2]
    L[:, 1, 1], L[:, 2, 1], L[:, 2, 2] = t[:, 3] ** 2 + 1e-4, t[:, 4], t[:, 5] ** 2 + 1e-4
    lam = L @ L.transpose(0, 2, 1)
    sig = np.linalg.inv(lam)
    mu = means.astype(np.float64)
    ele = (mu - o) @ ae
    sn = np.einsum("i,gij,j->g", ae, sig, ae)
    K = ele ** 2 / sn
    xs = mu - (sig @ ae) * (ele / sn)[:, None]
    near = np.abs(ele) < 3.0 * np.sqrt(sn)  # the visible pairs: the plane crosses the Gaussian
    assert near.sum() > 20
    np.testing.assert_allclose(plane[near, 2], K[near], rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(plane[near, 0], ((xs - o) @ al)[near], rtol=0, atol=2e-4)
    np.testing.assert_allclose(plane[near, 1], ((xs - o) @ aa)[near], rtol=0, atol=2e-4)
    # q at random plane points: the plane form (FP32 fixed order) against the direct quadratic form in FP64
    for g in np.nonzero(near)[0][:40]:
        cx, cy, Kg, p, qq, s = (float(v) for v in plane[g])
        for _ in range(5):
            lat = float(np.float32(cx + rng.normal() * 2.0))
            ax = float(np.float32(cy + rng.normal() * 2.0))
            x = o + lat * al + ax * aa - mu[g]
            q_ref = float(x @ lam[g] @ x)
            xf, yf = np.float32(lat) - np.float32(cx), np.float32(ax) - np.float32(cy)
            u = np.float32(p) * xf + np.float32(qq) * yf
            v = np.float32(s) * yf
            q = float(u * u + v * v + np.float32(Kg))
            assert q == pytest.approx(q_ref, rel=2e-5, abs=2e-5)


def test_plane_form_decisions_follow_the_grid_and_change_only_boundary_pairs():
    """Rendered on its own grid, a frame's FP32 oracle image (plane-form decisions and values) agrees with its FP64
    evaluation on the same pair set within 1e-6, and with the explicit-pixel semantics (q from the world positions)
    up to the few pairs whose q lies within rounding of chi2."""
    fr = straight_sweep(2, 64, 64, 0.3, 40.0, 10.0, 5.0, seed=3)[1]
    sc = make_scene(2000, 40.0, seed=4)
    geom = oug.FrameGeom.from_spec(fr)
    pix = linear_pixels(fr).numpy()
    m, t = sc.means.numpy(), sc.tril.numpy()
    mr = oug.map_frame(m, t, geom)
    lists = oug.pixel_lists(m, t, pix, geom, mr)
    plane = oug.plane_form(m, t, geom)
    sem = oug.Semantics()
    args = (m, t, sc.opacity.numpy(), sc.intensity.numpy(), pix, geom, lists, sem)
    img_p, _, _, _ = oug.forward(*args, plane=plane)
    img_p64, _, _, _ = oug.forward(*args, fp64=True, plane=plane)
    img_e, _, _, _ = oug.forward(*args)
    assert np.abs(img_p).max() > 0.05
    assert rel_l2(img_p, img_p64) < 1e-6
    assert rel_l2(img_p, img_e) < 1e-4
