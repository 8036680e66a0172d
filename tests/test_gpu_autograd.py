"""The differentiable superposition RMSD on the GPU (pyjess_amd.autograd):
forward against the CPU references, backward against central finite
differences (h = 1e-6, within 1e-7 relative) and against autograd through an
SVD on the CPU, including near-collinear and planar sets."""

import numpy as np
import pytest

from oracle import oracle
from pyjess_amd import synthetic

from tests.test_rmsd_grad_cpu import exact_rmsd

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def device():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def pairs(n, seed, **kw):
    return synthetic.superposition_pairs(np.random.default_rng(seed), n, **kw)


def run(a, b, off, grad=True, weights=None):
    from pyjess_amd.autograd import superposition_rmsd

    dev = device()
    ta = torch.tensor(a, device=dev, requires_grad=grad)
    tb = torch.tensor(b, device=dev, requires_grad=grad)
    r = superposition_rmsd(ta, tb, torch.tensor(off, device=dev))
    if not grad:
        return r.detach().cpu().numpy(), None, None
    w = torch.ones_like(r) if weights is None else torch.tensor(weights, device=dev)
    (r * w).sum().backward()
    return r.detach().cpu().numpy(), ta.grad.cpu().numpy(), tb.grad.cpu().numpy()


def fd(a, b, off, h=1e-6):
    """Central differences of every set's RMSD, all perturbed copies in one
    device call of the forward kernel."""
    n_sets = len(off) - 1
    pa, pb, po, where = [], [], [0], []
    for s in range(n_sets):
        A, B = a[off[s]:off[s + 1]], b[off[s]:off[s + 1]]
        for side in (0, 1):
            for i in range(len(A)):
                for d in range(3):
                    for sign in (1.0, -1.0):
                        A2, B2 = A.copy(), B.copy()
                        (A2 if side == 0 else B2)[i, d] += sign * h
                        pa.append(A2)
                        pb.append(B2)
                        po.append(po[-1] + len(A))
                        where.append((s, side, i, d, sign))
    r, _, _ = run(np.concatenate(pa), np.concatenate(pb), np.array(po, dtype=np.int64), grad=False)
    ga, gb = np.zeros_like(a), np.zeros_like(b)
    for (s, side, i, d, sign), v in zip(where, r):
        g = ga if side == 0 else gb
        g[off[s] + i, d] += sign * v / (2 * h)
    return ga, gb


def per_set_rel(g, f, off):
    out = []
    for s in range(len(off) - 1):
        x = np.concatenate([g[0][off[s]:off[s + 1]], g[1][off[s]:off[s + 1]]]).ravel()
        y = np.concatenate([f[0][off[s]:off[s + 1]], f[1][off[s]:off[s + 1]]]).ravel()
        out.append(np.linalg.norm(x - y) / np.linalg.norm(y))
    return np.array(out)


def test_forward_matches_references():
    a, b, off = pairs(300, 1, collinear=0.2)
    r, _, _ = run(a, b, off, grad=False)
    for s in range(len(off) - 1):
        A, B = a[off[s]:off[s + 1]], b[off[s]:off[s + 1]]
        if len(A) >= 6:
            # the reference's eigenvalue formula is exact to rounding here
            want = oracle.superpose(A, B)[0]
            assert abs(r[s] - want) <= 1e-12 * max(1.0, want)
        if s % 15 == 0 or len(A) == 3:
            want = exact_rmsd(A, B)
            assert abs(r[s] - want) <= 1e-13 * max(1.0, want), (s, len(A))


def test_backward_matches_finite_differences():
    a, b, off = pairs(120, 2, collinear=0.25)
    _, ga, gb = run(a, b, off)
    rel = per_set_rel((ga, gb), fd(a, b, off), off)
    assert rel.max() < 1e-7, rel.max()


def test_backward_matches_autograd_through_svd():
    from oracle.kabsch_torch import kabsch_rmsd_grad

    # near-collinear sets ten times tighter than the default, planar k = 3 sets
    a, b, off = pairs(150, 3, collinear=0.3, scatter=2e-3)
    w = np.random.default_rng(9).uniform(-2.0, 2.0, size=len(off) - 1)
    r, ga, gb = run(a, b, off, weights=w)
    for s in range(len(off) - 1):
        A, B = a[off[s]:off[s + 1]], b[off[s]:off[s + 1]]
        rr, ta, tb = kabsch_rmsd_grad(A, B)
        assert abs(r[s] - rr) <= 1e-12 * max(1.0, rr)
        x = np.concatenate([ga[off[s]:off[s + 1]], gb[off[s]:off[s + 1]]]).ravel()
        y = w[s] * np.concatenate([ta, tb]).ravel()
        assert np.linalg.norm(x - y) <= 1e-10 * np.linalg.norm(y), s


def test_equal_size_sets_and_one_sided_gradient():
    from pyjess_amd.autograd import superposition_rmsd

    dev = device()
    a, b, off = pairs(64, 4, k_min=7, k_max=7, collinear=0.0)
    ta = torch.tensor(a.reshape(64, 7, 3), device=dev, requires_grad=True)
    tb = torch.tensor(b.reshape(64, 7, 3), device=dev)
    r = superposition_rmsd(ta, tb)
    r.sum().backward()
    r2, ga, _ = run(a, b, off)
    assert np.array_equal(r.detach().cpu().numpy(), r2)
    assert np.array_equal(ta.grad.cpu().numpy().reshape(-1, 3), ga)


def test_rigid_copy_and_empty_inputs():
    a, _, off = pairs(4, 5, k_min=8, k_max=8, collinear=0.0)
    R = synthetic.random_rotations(np.random.default_rng(1), 1)[0]
    b = a @ R.T + 1.0
    r, ga, gb = run(a, b, off)
    assert np.all(r < 1e-6) and np.all(np.isfinite(ga)) and np.all(np.isfinite(gb))
    b2 = a.copy()
    r, ga, gb = run(a, b2, off)
    assert np.all(r < 1e-12)
    r, _, _ = run(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(1, dtype=np.int64), grad=False)
    assert r.shape == (0,)


def test_host_tensors_are_refused():
    from pyjess_amd.autograd import superposition_rmsd

    device()
    a = torch.zeros(2, 3, 3, dtype=torch.float64)
    with pytest.raises(TypeError):
        superposition_rmsd(a, a)


def test_large_sets_take_the_unstaged_path():
    from oracle.kabsch_torch import kabsch_rmsd_grad

    # tiles of 64 sets with more than 768 points in all are read straight from
    # global memory; mix them with small-set tiles that are staged in LDS
    big_a, big_b, big_off = pairs(70, 11, k_min=20, k_max=40, collinear=0.0)
    small_a, small_b, small_off = pairs(130, 12)
    a = np.concatenate([big_a, small_a])
    b = np.concatenate([big_b, small_b])
    off = np.concatenate([big_off, small_off[1:] + big_off[-1]])
    w = np.random.default_rng(3).uniform(-1.0, 1.0, size=len(off) - 1)
    r, ga, gb = run(a, b, off, weights=w)
    for s in range(len(off) - 1):
        A, B = a[off[s]:off[s + 1]], b[off[s]:off[s + 1]]
        # the SVD-based checker's |A|^2 + |B|^2 - 2 S loses digits on sets far
        # from the origin: the RMSD is checked against its 50-digit value
        want = exact_rmsd(A, B)
        assert abs(r[s] - want) <= 1e-13 * max(1.0, want), s
        _, ta, tb = kabsch_rmsd_grad(A, B)
        x = np.concatenate([ga[off[s]:off[s + 1]], gb[off[s]:off[s + 1]]]).ravel()
        y = w[s] * np.concatenate([ta, tb]).ravel()
        assert np.linalg.norm(x - y) <= 1e-10 * np.linalg.norm(y), s
