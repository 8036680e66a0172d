"""Differentiable superposition RMSD on the GPU.

``superposition_rmsd(a, b)`` returns, for every pair of point sets, the RMSD
of the least-squares superposition of ``a`` onto ``b`` -- the quantity of
``Superposition_rmsd`` (``include/jess/super.pxd:1-14``), computed in FP64
from a one-sided Jacobi SVD of the correlation and the residuals of the
optimal rotation (the search decides hits with the reference's eigenvalue
formula instead; the two agree to rounding on well-conditioned sets) -- and
``backward`` runs a hand-written FP64 gradient kernel for both coordinate
sets. PyTorch supplies only the device
tensors, the stream and the autograd plumbing; the arithmetic is in the HIP
library (``jam_rmsd_forward`` / ``jam_rmsd_backward``).
"""

from __future__ import annotations

import ctypes
import weakref
from typing import Optional

import torch

from . import _native

__all__ = ["SuperpositionRMSD", "superposition_rmsd"]


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _check_inputs(a: torch.Tensor, b: torch.Tensor, offsets: torch.Tensor) -> None:
    for name, t in (("a", a), ("b", b)):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"`{name}` must be a tensor on a GPU (the computation runs on the device only)")
        if t.dtype != torch.float64:
            raise TypeError(f"`{name}` must be float64 (found {t.dtype})")
        if t.dim() != 2 or t.shape[1] != 3:
            raise ValueError(f"`{name}` must have shape (points, 3), found {tuple(t.shape)}")
    if a.shape != b.shape:
        raise ValueError(f"`a` and `b` differ in shape: {tuple(a.shape)} and {tuple(b.shape)}")
    if a.device != b.device or offsets.device != a.device:
        raise ValueError("`a`, `b` and `offsets` must be on the same device")
    _check_offsets(offsets, a.shape[0])


# the offsets tensor last found valid: (weak reference, its version counter,
# pointer, size and the number of points it was checked against)
_last_valid = None


def _check_offsets(offsets: torch.Tensor, n_points: int) -> None:
    """Every row belongs to exactly one set: the offsets start at 0, never
    decrease and end at the number of points (one read of a flag from the
    device), so that the kernels write every gradient row. A call with the
    same offsets tensor as the last valid one, not modified since (torch's
    version counter, which in-place writes through any view advance), and
    the same number of points skips the device read."""
    global _last_valid
    if offsets.dtype != torch.int64 or offsets.dim() != 1 or offsets.numel() < 1:
        raise ValueError("`offsets` must be a 1-D int64 tensor of n_sets + 1 entries")
    key = (offsets._version, offsets.data_ptr(), offsets.numel(), n_points)
    last = _last_valid
    if last is not None and last[0]() is offsets and last[1] == key:
        return
    ok = (offsets[0] == 0) & (offsets[-1] == n_points)
    if offsets.numel() > 1:
        ok = ok & torch.all(offsets[1:] >= offsets[:-1])
    if not bool(ok):
        raise ValueError("`offsets` must start at 0, never decrease and end at the number of points")
    _last_valid = (weakref.ref(offsets), key)


class SuperpositionRMSD(torch.autograd.Function):
    """RMSD per point set; set s owns rows ``offsets[s]:offsets[s+1]`` of ``a`` and ``b``."""

    @staticmethod
    def forward(ctx, a: torch.Tensor, b: torch.Tensor, offsets: torch.Tensor) -> torch.Tensor:
        a = a.contiguous()
        b = b.contiguous()
        offsets = offsets.contiguous()
        _check_inputs(a, b, offsets)
        n_sets = offsets.numel() - 1
        rmsd = torch.empty(n_sets, dtype=torch.float64, device=a.device)
        # the optimal rotations are kept for the backward pass when it will run
        need_grad = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        rot = torch.empty((n_sets, 9), dtype=torch.float64, device=a.device) if need_grad else None
        L = _native.lib()
        with torch.cuda.device(a.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(a.device).cuda_stream)
            _native._check(L.jam_rmsd_forward(_ptr(a), _ptr(b), _ptr(offsets), n_sets, a.shape[0], _ptr(rmsd),
                                              _ptr(rot), stream))
        ctx.save_for_backward(a, b, offsets, rmsd, rot)
        return rmsd

    @staticmethod
    def backward(ctx, grad_rmsd: torch.Tensor):
        a, b, offsets, rmsd, rot = ctx.saved_tensors
        need_a, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_a or need_b):
            return None, None, None
        g = grad_rmsd.to(torch.float64).contiguous()
        # the offsets were checked in forward: every row belongs to a set of at
        # least one point, and the kernel writes every such row
        ga = torch.empty_like(a) if need_a else None
        gb = torch.empty_like(b) if need_b else None
        L = _native.lib()
        with torch.cuda.device(a.device):
            stream = ctypes.c_void_p(torch.cuda.current_stream(a.device).cuda_stream)
            _native._check(L.jam_rmsd_backward(_ptr(a), _ptr(b), _ptr(offsets), offsets.numel() - 1, a.shape[0],
                                               _ptr(rmsd), _ptr(rot), _ptr(g), _ptr(ga), _ptr(gb), stream))
        return ga, gb, None


def superposition_rmsd(a: torch.Tensor, b: torch.Tensor, offsets: Optional[torch.Tensor] = None) -> torch.Tensor:
    """RMSD of the superposition of ``a`` onto ``b``, differentiable in both.

    ``a`` and ``b`` are float64 GPU tensors, either ``(n_sets, k, 3)`` (sets of
    equal size) or ``(points, 3)`` with ``offsets`` (int64, ``n_sets + 1``
    entries) delimiting the sets. Returns ``(n_sets,)``.
    """
    if offsets is None:
        if a.dim() != 3:
            raise ValueError("without `offsets`, `a` and `b` must have shape (n_sets, k, 3)")
        n_sets, k = a.shape[0], a.shape[1]
        offsets = torch.arange(0, n_sets + 1, dtype=torch.int64, device=a.device) * k
        a = a.reshape(-1, 3)
        b = b.reshape(-1, 3)
    return SuperpositionRMSD.apply(a, b, offsets)
