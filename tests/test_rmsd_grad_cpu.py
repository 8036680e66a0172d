"""The CPU reference of the differentiable RMSD (oracle/kabsch_torch.py:
autograd through an SVD) pinned against the RMSD oracle and against central
finite differences, on the generator's sets (including near-collinear k = 3)
and on cases whose answer is known exactly."""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle  # noqa: E402
from oracle.kabsch_torch import kabsch_rmsd, kabsch_rmsd_grad  # noqa: E402
from pyjess_amd import synthetic  # noqa: E402


def sets(n, seed, **kw):
    a, b, off = synthetic.superposition_pairs(np.random.default_rng(seed), n, **kw)
    return [(a[off[s]:off[s + 1]], b[off[s]:off[s + 1]]) for s in range(n)]


def rmsd_cpu(A, B):
    return float(kabsch_rmsd(torch.as_tensor(A), torch.as_tensor(B)))


def fd_grad(f, A, B, h=1e-6):
    ga = np.zeros_like(A)
    gb = np.zeros_like(B)
    for i in range(len(A)):
        for d in range(3):
            for X, G, first in ((A, ga, True), (B, gb, False)):
                Xp, Xm = X.copy(), X.copy()
                Xp[i, d] += h
                Xm[i, d] -= h
                fp = f(Xp, B) if first else f(A, Xp)
                fm = f(Xm, B) if first else f(A, Xm)
                G[i, d] = (fp - fm) / (2 * h)
    return ga, gb


def exact_rmsd(A, B, digits=50):
    """The RMSD in 50-digit arithmetic (mpmath SVD)."""
    import mpmath

    mpmath.mp.dps = digits
    n = len(A)
    Am = [[mpmath.mpf(float(x)) for x in r] for r in A]
    Bm = [[mpmath.mpf(float(x)) for x in r] for r in B]
    ca = [sum(r[d] for r in Am) / n for d in range(3)]
    cb = [sum(r[d] for r in Bm) / n for d in range(3)]
    Ac = [[r[d] - ca[d] for d in range(3)] for r in Am]
    Bc = [[r[d] - cb[d] for d in range(3)] for r in Bm]
    X = mpmath.matrix(3, 3)
    for i in range(n):
        for j in range(3):
            for k in range(3):
                X[j, k] += Bc[i][j] * Ac[i][k]
    S = mpmath.svd_r(X, compute_uv=False)
    s = S[0] + S[1] + (S[2] if mpmath.det(X) >= 0 else -S[2])
    e = sum(x * x for r in Ac for x in r) + sum(x * x for r in Bc for x in r) - 2 * s
    return float(mpmath.sqrt(max(e, 0) / n))


def test_reference_rmsd_equals_oracle_on_well_conditioned_sets():
    # the reference's eigenvalue formula loses digits when X^T X has a tiny
    # eigenvalue (k = 3 sets are always planar); from 6 points on it is exact
    # to rounding
    for A, B in sets(80, 1, collinear=0.0, k_min=6):
        want = oracle.superpose(A, B)[0]
        assert abs(rmsd_cpu(A, B) - want) <= 1e-12 * max(1.0, want)


def test_reference_rmsd_equals_50_digit_value():
    for A, B in sets(16, 7, collinear=0.3, k_min=3, k_max=5):
        want = exact_rmsd(A, B)
        assert abs(rmsd_cpu(A, B) - want) <= 1e-13 * max(1.0, want)


def test_reference_gradient_matches_finite_differences():
    worst = 0.0
    for A, B in sets(40, 2, collinear=0.25):
        _, ga, gb = kabsch_rmsd_grad(A, B)
        fa, fb = fd_grad(rmsd_cpu, A, B)
        g = np.concatenate([ga, gb]).ravel()
        f = np.concatenate([fa, fb]).ravel()
        worst = max(worst, np.linalg.norm(g - f) / np.linalg.norm(f))
    assert worst < 1e-7


def test_rigid_copy_has_zero_rmsd():
    A, _ = sets(1, 3, collinear=0.0, k_min=8, k_max=8)[0]
    R = synthetic.random_rotations(np.random.default_rng(4), 1)[0]
    B = A @ R.T + np.array([3.0, -2.0, 7.0])
    assert rmsd_cpu(A, B) < 1e-6


def test_mirror_image_needs_a_proper_rotation():
    # b = mirror of a: the best proper rotation cannot undo a reflection, so
    # the RMSD is that of S = s1 + s2 - s3 (the reference's reflection rule)
    A, _ = sets(1, 5, collinear=0.0, k_min=10, k_max=10)[0]
    B = A * np.array([1.0, 1.0, -1.0])
    want = oracle.superpose(A, B)[0]
    assert want > 0.1
    assert math.isclose(rmsd_cpu(A, B), want, rel_tol=1e-12)


def test_gradient_is_translation_invariant_and_sums_to_zero():
    A, B = sets(1, 6, collinear=0.0, k_min=7, k_max=7)[0]
    _, ga, gb = kabsch_rmsd_grad(A, B)
    _, ga2, gb2 = kabsch_rmsd_grad(A + 5.0, B - 3.0)
    assert np.allclose(ga, ga2, atol=1e-12) and np.allclose(gb, gb2, atol=1e-12)
    assert np.allclose(ga.sum(0), 0.0, atol=1e-12) and np.allclose(gb.sum(0), 0.0, atol=1e-12)


@pytest.mark.parametrize("off,n", [([1, 3], 3), ([0, 5, 3], 5), ([0, 2], 3), ([0, 4], 3), ([-1, 3], 3)])
def test_malformed_offsets_raise_before_any_launch(off, n):
    """Offsets that do not start at 0, decrease, or do not end at the number of
    points are rejected on the host, before the kernels could see them."""
    from pyjess_amd import autograd

    with pytest.raises(ValueError):
        autograd._check_offsets(torch.tensor(off, dtype=torch.int64), n)


def test_valid_offsets_pass_and_host_tensors_are_refused():
    from pyjess_amd import autograd

    autograd._check_offsets(torch.tensor([0, 3, 3, 7], dtype=torch.int64), 7)
    with pytest.raises(TypeError):
        autograd.superposition_rmsd(torch.zeros((2, 3, 3), dtype=torch.float64),
                                    torch.zeros((2, 3, 3), dtype=torch.float64))


def test_unchanged_offsets_are_checked_once_and_changes_are_seen():
    """The last valid offsets tensor is not read again while unmodified; an
    in-place write (through a view too), another point count or another
    tensor is checked again."""
    from pyjess_amd import autograd

    off = torch.tensor([0, 3, 3, 7], dtype=torch.int64)
    autograd._check_offsets(off, 7)
    autograd._check_offsets(off, 7)   # cached
    with pytest.raises(ValueError):
        autograd._check_offsets(off, 8)
    off[1:3].fill_(9)                 # in place, through a view: now decreasing
    with pytest.raises(ValueError):
        autograd._check_offsets(off, 7)
    other = torch.tensor([0, 5, 2, 7], dtype=torch.int64)
    with pytest.raises(ValueError):
        autograd._check_offsets(other, 7)
