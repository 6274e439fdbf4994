"""The oracle's restatement of the reference's PyTorch fallback renderer against that renderer's own output.

tests/golden/fallback_ref_24x32_g384.npz holds one 24x32 linear-probe frame rendered by
_render_ultragauss_torch_fallback (pipeline_god_file.py:556-603) over 384 anisotropic Gaussians, with the
gradients of L = <dI, image> + <dA, alpha> w.r.t. the reference's raw parameters (means, raw precision factors,
raw opacity and intensity logits through the sigmoids of pipeline_god_file.py:369-373);
tests/golden/make_fallback_fixture.py wrote it. Bars: the float32 restatement within 1e-7 relative L2 of the
fixture (it reproduces it bit for bit: the same FP32 operations in the same order), float64 within 1e-6
(measured 5e-7: the fixture's own FP32 rounding)."""

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import fallback as ofb
from tests.ug_helpers import rel_l2

GOLDEN = Path(__file__).resolve().parent / "golden" / "fallback_ref_24x32_g384.npz"


def load_fallback_golden():
    z = np.load(GOLDEN)
    H, W = int(z["frame"][0]), int(z["frame"][1])
    return z, H, W


def _oracle(z, H, W, dtype):
    raw = [torch.from_numpy(z[k]).to(dtype).requires_grad_(True)
           for k in ("means", "raw_tril", "raw_opacity", "raw_intensity")]
    img, alpha = ofb.render(raw[0], raw[1], torch.sigmoid(raw[2]), torch.sigmoid(raw[3]),
                            torch.from_numpy(z["pixels"]).to(dtype), H, W, beta=float(z["beta"]))
    loss = (img.reshape(-1) * torch.from_numpy(z["dI"]).to(dtype)).sum() + \
        (alpha.reshape(-1) * torch.from_numpy(z["dA"]).to(dtype)).sum()
    loss.backward()
    return img.detach().numpy().reshape(-1), alpha.detach().numpy().reshape(-1), [r.grad.numpy() for r in raw]


@pytest.mark.parametrize("dtype,bar", [(torch.float32, 1e-7), (torch.float64, 1e-6)])
def test_fallback_restatement_matches_reference_output(dtype, bar):
    z, H, W = load_fallback_golden()
    assert z["image"].size == H * W and z["means"].shape == (384, 3)
    # the scene is not degenerate: the image spans a range, alpha piles up, every gradient group is non-zero
    assert z["image"].max() - z["image"].min() > 0.1 and z["alpha"].max() > 1.0
    img, alpha, grads = _oracle(z, H, W, dtype)
    assert rel_l2(img, z["image"]) < bar
    assert rel_l2(alpha, z["alpha"]) < bar
    for g, k in zip(grads, ("d_means", "d_raw_tril", "d_raw_opacity", "d_raw_intensity")):
        assert np.linalg.norm(z[k]) > 0
        assert rel_l2(g, z[k]) < 10 * bar, k


def test_fallback_golden_frame_is_the_linear_probe_grid():
    """The fixture's pixels are the reference's linear-probe grid (pipeline_god_file.py:292-306) for the stored
    origin, axes and spacings, so the HIP path can render the same frame from its grid parameters."""
    z, H, W = load_fallback_golden()
    _h, _w, dlat, dax, near, col = z["frame"]
    xs, ys = np.meshgrid(np.arange(W), np.arange(H))
    grid = z["origin"][None, None].astype(np.float64) + ((xs - col) * dlat)[..., None] * z["axes"][0] + \
        (near + ys * dax)[..., None] * z["axes"][1]
    np.testing.assert_allclose(grid.reshape(-1, 3), z["pixels"], atol=2e-6)
    np.testing.assert_allclose(z["axes"] @ z["axes"].T, np.eye(3), atol=1e-6)
