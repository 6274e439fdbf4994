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


def _gp_frame(geometry=0):
    from us3d.batch import FrameSpec

    v = lambda *a: torch.tensor(a, dtype=torch.float32)  # noqa: E731
    return FrameSpec(v(0, 0, 0), v(1, 0, 0), v(0, 1, 0), v(0, 0, 1), 16, 16, 0.2, 0.2, 0.0, 7.5, 0, geometry)


def test_batch_refuses_fan_frames_and_a_zero_weight_min():
    """The GaussianPile kernels evaluate the linear-probe grid only, so fan frames are refused before any device
    work; weight_min must be > 0 (the reference floors it at 1e-12)."""
    from us3d import _lib
    from us3d.batch import SweepSpec
    from us3d.gaussianpile import GPBatch

    st = GaussianPlieRenderConfig(H=16, W=16, lateral_spacing_mm_per_pix=0.2, axial_spacing_mm_per_pix=0.2,
                                  axial_near_mm=0.0).settings()
    with pytest.raises(ValueError, match="linear-probe"):
        GPBatch([_gp_frame(geometry=1)], [SweepSpec(0, 4)], 4, st, torch.device("cpu"))
    st0 = _lib.GPSettings(k_sigma=3.0, sigma_ele=1.0, det_eps=1e-6, weight_min=0.0, tile_size=16)
    with pytest.raises(ValueError, match="weight_min"):
        GPBatch([_gp_frame()], [SweepSpec(0, 4)], 4, st0, torch.device("cpu"))


def test_c_abi_rejects_a_zero_weight_min_before_any_device_work():
    """us3d_gp_map validates the settings after the plan and before any HIP call: weight_min = 0 is an argument
    error (at 0 a listed pair with an underflowed base weight would make the adjoint divide 0 by 0). The same call
    with weight_min = 1e-6 passes the settings and fails on the NULL parameter pointers instead, so it is the
    setting that was refused. Plan pointers are placeholders that are never dereferenced."""
    import ctypes

    from us3d import _lib

    L = _lib.lib()
    p = _lib.Plan()
    p.n_gauss, p.n_frames, p.n_sweeps, p.n_tiles, p.n_pairs, p.n_pixels, p.k_capacity = 4, 1, 1, 1, 4, 256, 64
    p.v_capacity = 4
    dummy = ctypes.create_string_buffer(64)
    for f in ("frames", "frame_tile_base", "frame_pixel_base", "sweeps", "workspace"):
        setattr(p, f, ctypes.addressof(dummy))
    p.workspace_bytes = 1 << 40
    st = _lib.GPSettings(k_sigma=3.0, sigma_ele=1.0, det_eps=1e-6, weight_min=0.0, tile_size=16)
    assert L.us3d_gp_map(p, None, None, None, None, None, st, None) == 1  # US3D_ERR_ARGUMENT
    assert b"weight_min" in L.us3d_last_error()
    st.weight_min = 1e-6
    assert L.us3d_gp_map(p, None, None, None, None, None, st, None) == 1
    assert b"pointer is NULL" in L.us3d_last_error()


def test_oracle_sanitise_follows_the_reference_definition():
    """us_focus_sanitize_grad (us_focus.cu:129-137): non-finite -> 0, |g| > 1e3 -> copysign(1e3, g), else g."""
    from oracle import gp

    g = torch.tensor([float("nan"), float("inf"), -float("inf"), 1e3, -1e3, 1000.5, -2e5, 3.25, -0.0],
                     dtype=torch.float64)
    want = torch.tensor([0.0, 0.0, 0.0, 1e3, -1e3, 1e3, -1e3, 3.25, 0.0], dtype=torch.float64)
    assert torch.equal(gp.sanitize_grad(g), want)
