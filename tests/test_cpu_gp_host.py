"""Host logic of the GaussianPile facade (CPU only): the input checks of check.py:10-29 with the reference's
messages, and the loud refusals where this build differs (pose gradients, a CPU device)."""

import pytest
import torch

from us3d.gaussianpile import GaussianPlieRenderConfig, GaussianPlieRenderer, validate_render_inputs


def _inputs(n=5):
    return torch.zeros(n, 3), torch.ones(n, 3), torch.tensor([[1.0, 0.0, 0.0, 0.0]]).repeat(n, 1), torch.ones(n)


@pytest.mark.parametrize("which, shape, msg", [
    (0, (5, 2), "gauss_means must be (N, 3)"),
    (1, (5, 4), "gauss_scales must be (N, 3)"),
    (2, (5, 3), "gauss_quats must be (N, 4)"),
    (3, (5, 2), "gauss_opacities must be (N,) or (N, 1)"),
])
def test_input_checks_match_reference_messages(which, shape, msg):
    args = list(_inputs())
    args[which] = torch.zeros(shape)
    with pytest.raises(ValueError, match=msg.replace("(", r"\(").replace(")", r"\)")):
        validate_render_inputs(*args, torch.eye(4))
    validate_render_inputs(*_inputs(), torch.eye(4))  # well-formed: no error
    m, s, q, o = _inputs()
    validate_render_inputs(m, s, q, o.reshape(-1, 1), torch.eye(4))
    with pytest.raises(ValueError, match="pose_probe_to_world must be"):
        validate_render_inputs(m, s, q, o, torch.eye(3))


def test_facade_checks_inputs_when_enabled_and_refuses_pose_gradients_and_cpu():
    cfg = GaussianPlieRenderConfig(H=16, W=16, lateral_spacing_mm_per_pix=0.2, axial_spacing_mm_per_pix=0.2,
                                   axial_near_mm=0.0, enable_input_checks=True)
    r = GaussianPlieRenderer(cfg)
    m, s, q, o = _inputs()
    with pytest.raises(ValueError, match="gauss_scales"):
        r(m, torch.ones(5, 2), q, o, torch.eye(4))
    with pytest.raises(NotImplementedError, match="probe pose"):
        r(m, s, q, o, torch.eye(4).requires_grad_(True))
    with pytest.raises(RuntimeError, match="GPU only"):
        r(m, s, q, o, torch.eye(4))
