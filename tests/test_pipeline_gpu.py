"""The training-loop counterpart end to end on the reference's own sweep data (3-frame fixture cut from
data/synthetic_liver_l1_single_sweep.npz), with the smoke configuration of
configs/synthetic_liver_l1_ultragauss_smoke.json, for each renderer method this build provides."""

import json
from pathlib import Path

import numpy as np
import pytest
import torch

from us3d.pipeline import GodConfig, RenderDispatcher, image_loss, train

pytestmark = pytest.mark.gpu

FIXTURE = Path(__file__).resolve().parent / "golden" / "liver_3frames.npz"

SMOKE = dict(data_npz=str(FIXTURE), device="cuda", seed=42, eval_every=0, save_every=0, log_every=1, num_init=2048,
             learning_rate=0.001, loss_l1_weight=1.0, loss_ssim_weight=0.2, train_fraction=0.67, image_height=448,
             image_width=224, delta_lateral_mm=0.35714285714285715, delta_axial_mm=0.3125,
             delta_elevational_mm=0.2685, near_mm=0.0, far_mm=140.0, width_mm=80.0, center_lateral=True,
             init_scale_mm=0.8, init_opacity=0.1, init_intensity=0.5)


@pytest.mark.parametrize("method", ["ultragauss", "ultragray", "gaussianpile"])
def test_pipeline_trains_on_reference_sweep(dev, tmp_path, method):
    cfg = GodConfig(method=method, output_dir=str(tmp_path / method), iterations=12, **SMOKE)
    ctx = train(cfg)
    rows = (tmp_path / method / "scalars.csv").read_text().strip().splitlines()
    assert len(rows) == 13
    losses = [float(r.split(",")[2]) for r in rows[1:]]
    assert all(np.isfinite(losses))
    assert (tmp_path / method / "final.pt").exists()
    used = json.loads((tmp_path / method / "used_config.json").read_text())
    assert used["method"] == method
    # gradients reached every rendered parameter group
    params = ctx["params"]
    assert params.means.grad is not None and torch.isfinite(params.means.grad).all()
    frame = ctx["test_data"].get_frame(int(ctx["test_data"].indices[0]), dev)
    with torch.no_grad():
        out = RenderDispatcher(cfg).render(params, frame)
    assert out.image.shape[-2:] == (448, 224)
    loss, parts = image_loss(out.image, frame["image"], cfg)
    assert np.isfinite(parts["total"])


def test_unsupported_method_raises(dev):
    with pytest.raises(ValueError):
        RenderDispatcher(GodConfig(method="slice_adapted_3dgs", **SMOKE))
