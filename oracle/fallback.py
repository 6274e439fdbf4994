"""Oracle (test infrastructure only): the reference's PyTorch fallback renderer, restated.

Follows _render_ultragauss_torch_fallback, pipeline_god_file.py:556-603: Lambda = L L^T from the raw factors
(l_ii = raw_ii^2 + beta), q = d^T Lambda d, w = exp(clamp(-q/2, -60, 0)) * opacity, image =
clamp(sum w c / max(sum w + 1e-6, 1e-8), 0, 1) as (1, H, W), alpha = sum w as (1, H, W); no culling, pixel
chunks x Gaussian chunks. Runs on CPU in float32 or float64; autograd gives the gradients.
"""

from __future__ import annotations

import torch


def build_precision(tril: torch.Tensor, beta: float) -> torch.Tensor:
    L = torch.zeros((tril.shape[0], 3, 3), dtype=tril.dtype, device=tril.device)
    L[:, 0, 0] = tril[:, 0].square() + beta
    L[:, 1, 0] = tril[:, 1]
    L[:, 2, 0] = tril[:, 2]
    L[:, 1, 1] = tril[:, 3].square() + beta
    L[:, 2, 1] = tril[:, 4]
    L[:, 2, 2] = tril[:, 5].square() + beta
    return L @ L.transpose(-1, -2)


def render(means, tril, opacity, intensity, pixels, H, W, beta=1e-4, pix_chunk=2048, gauss_chunk=512):
    """Returns image (1,H,W) and alpha (1,H,W), differentiable w.r.t. all four Gaussian tensors."""
    lam = build_precision(tril, beta)
    out, alphas = [], []
    for p0 in range(0, pixels.shape[0], pix_chunk):
        p = pixels[p0:p0 + pix_chunk]
        num = torch.zeros(p.shape[0], dtype=pixels.dtype)
        den = torch.zeros_like(num)
        for g0 in range(0, means.shape[0], gauss_chunk):
            g1 = min(g0 + gauss_chunk, means.shape[0])
            d = p[:, None, :] - means[None, g0:g1, :]
            q = torch.einsum("pgi,gij,pgj->pg", d, lam[g0:g1], d)
            w = torch.exp(torch.clamp(-0.5 * q, min=-60.0, max=0.0)) * opacity[g0:g1].view(1, -1)
            num = num + (w * intensity[g0:g1].view(1, -1)).sum(dim=1)
            den = den + w.sum(dim=1)
        out.append(num / torch.clamp(den + 1.0e-6, min=1.0e-8))
        alphas.append(den)
    image = torch.cat(out).view(H, W).unsqueeze(0).clamp(0.0, 1.0)
    alpha = torch.cat(alphas).view(1, H, W)
    return image, alpha
