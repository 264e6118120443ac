"""CPU reference of the learned guiding network and its training loss.

TEST INFRASTRUCTURE ONLY (the checker of lmgpu_guide_*): a torch FP32
restatement used with torch autograd and torch.optim.Adam, which pin the
GPU's hand-written forward, backward and Adam. BF16 rounding is applied at
the same points as the GPU kernels (inputs, weights, hidden activations),
with straight-through gradients.

The network has no counterpart in the reference renderer; it is the guide of
the MI355X path tracer (see lm_amd.guide for the packed parameter layout).
"""
from __future__ import annotations

import math

import numpy as np
import torch

N_LOBES = 8
KAPPA_MIN, KAPPA_MAX = 0.05, 4000.0


def bf16(x: torch.Tensor) -> torch.Tensor:
    """Round to BF16 (nearest even) in the forward pass, identity gradient."""
    r = x.detach().to(torch.bfloat16).to(torch.float32)
    return x + (r - x.detach())


def encode(pos, nrm, wo, box_lo, box_hi) -> torch.Tensor:
    """78 features: sin/cos(2^j pi q) for j < 12 per axis (q = pos in [0,1]^3,
    as the FP32 product with the reciprocal box extent), then normal and
    outgoing direction."""
    lo = np.asarray(box_lo, dtype=np.float32)
    inv = np.float32(1.0) / (np.asarray(box_hi, dtype=np.float32) - lo)
    q = (pos - torch.from_numpy(lo)) * torch.from_numpy(inv)
    pi32 = torch.tensor(np.float32(math.pi))
    feats = []
    for a in range(3):
        f = pi32 * q[:, a:a + 1]
        for j in range(12):
            arg = f * (2.0 ** j)
            feats += [torch.sin(arg), torch.cos(arg)]
    return torch.cat(feats + [nrm, wo], dim=1)


def unpack(p: np.ndarray):
    from lm_amd.guide import unpack_params
    return [(torch.tensor(W.copy()), torch.tensor(b.copy())) for W, b in unpack_params(p)]


def forward(layers, x):
    """Raw 40 outputs with BF16 operands / FP32 accumulation."""
    h = bf16(x)
    for li, (W, b) in enumerate(layers):
        h = h @ bf16(W).T + b
        if li < 3:
            h = bf16(torch.relu(h))
    return h


def decode(y):
    pi = torch.softmax(y[:, :N_LOBES], dim=1)
    kappa = torch.clamp(torch.nn.functional.softplus(y[:, N_LOBES:2 * N_LOBES]), KAPPA_MIN, KAPPA_MAX)
    mu = y[:, 2 * N_LOBES:].reshape(-1, N_LOBES, 3)
    mu = mu / mu.norm(dim=2, keepdim=True)
    return pi, kappa, mu


def vmf_pdf(kappa, mu, w):
    """kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)); w [n,3]."""
    c = (mu * w[:, None, :]).sum(-1)
    return kappa / (2 * math.pi * (-torch.expm1(-2 * kappa))) * torch.exp(kappa * (c - 1))


def mixture_pdf(pi, kappa, mu, w):
    return (pi * vmf_pdf(kappa, mu, w)).sum(1)


def loss_fn(layers, records, box_lo, box_hi, wclamp=50.0):
    """-(1/n) sum_i wt_i log p(wi_i), wt = min(w/mean(w), clamp), w = lum(L)/pdf."""
    r = records
    x = encode(r[:, 0:3], r[:, 3:6], r[:, 6:9], box_lo, box_hi)
    y = forward(layers, x)
    pi, kappa, mu = decode(y)
    wi = r[:, 9:12] / r[:, 9:12].norm(dim=1, keepdim=True)
    p = mixture_pdf(pi, kappa, mu, wi)
    L = 0.2126 * r[:, 13] + 0.7152 * r[:, 14] + 0.0722 * r[:, 15]
    w = torch.where((r[:, 12] > 0) & (L > 0), L / r[:, 12], torch.zeros_like(L))
    wt = torch.clamp(w / w.mean(), max=wclamp).detach()
    return -(wt * torch.log(p)).mean()
