"""CPU reference for the differentiable superposition RMSD (TEST INFRASTRUCTURE).

Only tests/ and the cpu_baseline leg of bench.py use this module, as the
checker. The RMSD of the optimal superposition is computed with an SVD
(Kabsch; the reflection case handled by flipping the smallest singular value,
the rule of patches/Super.c.patch:40-64) in float64 torch on the CPU, and its
gradient is torch's autograd through that SVD: an independent route to the
same derivative as the device kernel's closed form.
"""

from __future__ import annotations

import torch


def kabsch_rmsd(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """RMSD of the superposition of a onto b; a, b: (k, 3) float64 CPU tensors."""
    A = a - a.mean(dim=0, keepdim=True)
    B = b - b.mean(dim=0, keepdim=True)
    X = B.T @ A
    S = torch.linalg.svdvals(X)
    det = torch.linalg.det(X)
    sgn = torch.where(det < 0, -1.0, 1.0).to(a.dtype)
    s = S[0] + S[1] + sgn * S[2]
    e = (A * A).sum() + (B * B).sum() - 2.0 * s
    return torch.sqrt(torch.clamp(e, min=0.0) / a.shape[0])


def kabsch_rmsd_grad(a, b):
    """(rmsd, d rmsd / d a, d rmsd / d b) by autograd through the SVD."""
    a = torch.as_tensor(a, dtype=torch.float64).clone().requires_grad_(True)
    b = torch.as_tensor(b, dtype=torch.float64).clone().requires_grad_(True)
    r = kabsch_rmsd(a, b)
    ga, gb = torch.autograd.grad(r, (a, b))
    return float(r), ga.numpy(), gb.numpy()
