"""Golden vectors of the reference's own PyTorch fallback renderer (_render_ultragauss_torch_fallback,
pipeline_god_file.py:556-603), the correctness oracle the north star names. Its FP32 behaviour (pixel/Gaussian
chunking, einsum order, the exp clamp) is pinned by nothing but the function itself, so this script runs it once,
in a container where the reference is mounted, and stores inputs and outputs as a fixture:
    python tests/golden/make_fallback_fixture.py /root/reference
It loads pipeline_god_file.py as a module (its import only extends sys.path and defines classes and functions),
builds the reference's own GaussianParameters for a small seeded anisotropic scene, sets its means, raw precision
factors and raw opacity / intensity logits, renders one 24x32 linear-probe frame through the fallback on the CPU and
back-propagates L = <dI, image> + <dA, alpha>. Saved (float32): the raw parameters, pixel positions, dI, dA, the
image and alpha, and the gradients w.r.t. the raw parameters (through the reference's sigmoid activations)."""

import importlib.util
import sys
from pathlib import Path

import numpy as np
import torch

ref_root = Path(sys.argv[1])
spec = importlib.util.spec_from_file_location("reference_pipeline", ref_root / "pipeline_god_file.py")
ref = importlib.util.module_from_spec(spec)
sys.modules["reference_pipeline"] = ref
spec.loader.exec_module(ref)

H, W, G = 24, 32, 384
rng = np.random.default_rng(20261204)
origin = np.array([1.0, -2.0, 0.5], np.float64)
# probe axes: a rotation about a tilted axis, columns lateral / axial / elevational
ang, axis = 0.35, np.array([0.3, -0.5, 0.81])
axis /= np.linalg.norm(axis)
K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
a_lat, a_ax, a_ele = R[:, 0], R[:, 1], R[:, 2]
dlat, dax, near = 0.3, 0.3, 2.0
col = (W - 1) / 2
xs, ys = np.meshgrid(np.arange(W), np.arange(H))
lat = (xs - col) * dlat
axl = near + ys * dax
pixels = (origin[None, None] + lat[..., None] * a_lat + axl[..., None] * a_ax).reshape(-1, 3).astype(np.float32)

# Gaussians around the frame: in-plane spread over the image, +-2.5 mm out of plane
u = rng.uniform(-col * dlat - 1.0, col * dlat + 1.0, G)
v = rng.uniform(near - 1.0, near + H * dax + 1.0, G)
e = rng.uniform(-2.5, 2.5, G)
means = (origin + u[:, None] * a_lat + v[:, None] * a_ax + e[:, None] * a_ele).astype(np.float32)
raw_tril = np.zeros((G, 6), np.float32)
raw_tril[:, [0, 3, 5]] = rng.uniform(0.55, 1.4, (G, 3))  # l_ii = raw^2 + beta: scales ~0.5-3 mm
raw_tril[:, [1, 2, 4]] = rng.normal(0.0, 0.35, (G, 3))
raw_o = rng.normal(0.0, 1.0, G).astype(np.float32)
raw_c = rng.normal(0.0, 1.0, G).astype(np.float32)
dI = rng.normal(0.0, 1.0, H * W).astype(np.float32)
dA = (rng.normal(0.0, 1.0, H * W) * 0.1).astype(np.float32)

cfg = ref.GodConfig()
cfg.image_height, cfg.image_width = H, W
params = ref.GaussianParameters(cfg, torch.from_numpy(means), torch.full((G,), 0.5))
with torch.no_grad():
    params.raw_precision_tril.copy_(torch.from_numpy(raw_tril))
    params.raw_opacities.copy_(torch.from_numpy(raw_o).view(G, 1))
    params.raw_intensities.copy_(torch.from_numpy(raw_c).view(G, 1))
out = ref._render_ultragauss_torch_fallback(params, {"pixel_positions_world": torch.from_numpy(pixels)}, cfg)
image, alpha = out.image.reshape(-1), out.extras["alpha_image"].reshape(-1)
((image * torch.from_numpy(dI)).sum() + (alpha * torch.from_numpy(dA)).sum()).backward()

dst = Path(__file__).resolve().parent / "fallback_ref_24x32_g384.npz"
np.savez_compressed(
    dst, means=means, raw_tril=raw_tril, raw_opacity=raw_o, raw_intensity=raw_c, pixels=pixels, dI=dI, dA=dA,
    image=image.detach().numpy(), alpha=alpha.detach().numpy(), d_means=params.means.grad.numpy(),
    d_raw_tril=params.raw_precision_tril.grad.numpy(), d_raw_opacity=params.raw_opacities.grad.reshape(-1).numpy(),
    d_raw_intensity=params.raw_intensities.grad.reshape(-1).numpy(), beta=np.float32(cfg.ultragauss_diag_beta),
    frame=np.array([H, W, dlat, dax, near, col], np.float64), origin=origin.astype(np.float32),
    axes=np.stack([a_lat, a_ax, a_ele]).astype(np.float32))
print(dst, "image range", float(image.detach().min()), float(image.detach().max()), "alpha max", float(alpha.detach().max()))
