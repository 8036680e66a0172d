"""Pins of the oracle from mathematics rather than from the reference's files:
superposition against an SVD Kabsch solution and exact rigid motions, the
scanner against brute-force enumeration of the match definition on tiny
inputs, and the kd-tree arrangement against the median-split invariant."""

import itertools

import numpy as np
import pytest

from oracle import oracle
from pyjess_amd import synthetic


def kabsch(a, b):
    """Proper rotation R minimising sum |R(a - ca) - (b - cb)|^2, and its RMSD."""
    ca, cb = a.mean(0), b.mean(0)
    H = (a - ca).T @ (b - cb)
    U, S, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    D = np.diag([1.0, 1.0, d])
    R = Vt.T @ D @ U.T
    r = np.sqrt(max(((a - ca) @ R.T - (b - cb)) .__pow__(2).sum(), 0.0) / len(a))
    return R, r


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("k", [3, 4, 6, 9, 17, 24])
def test_superpose_matches_svd_kabsch(k):
    rng = np.random.default_rng(k)
    for _ in range(50):
        a = rng.normal(scale=5.0, size=(k, 3))
        b = a @ random_rotation(rng).T + rng.normal(size=3) * 10 + rng.normal(scale=0.5, size=(k, 3))
        rmsd, M, c0, c1 = oracle.superpose(a, b)
        R, r = kabsch(a, b)
        # three points are coplanar: the smallest eigenvalue of X^T X is
        # rounding noise, whose square root enters sumE (the reference's
        # eigenvalue formula), so the RMSD is good to ~1e-6 only
        tol = 1e-5 if k == 3 else 1e-7
        assert rmsd == pytest.approx(r, rel=tol, abs=tol)
        if k > 3:
            # with three points X^T X has rank 2 and the rotation formula is
            # undefined ("fails if XX has rank<3" in the reference); only the
            # RMSD is meaningful there
            assert np.allclose(M, R, atol=1e-8)
        assert np.allclose(c0, a.mean(0), atol=1e-12) and np.allclose(c1, b.mean(0), atol=1e-12)


def test_superpose_exact_copy():
    rng = np.random.default_rng(1)
    a = rng.normal(scale=5.0, size=(9, 3))
    R = random_rotation(rng)
    b = a @ R.T + np.array([3.0, -2.0, 7.0])
    rmsd, M, _, _ = oracle.superpose(a, b)
    assert rmsd < 1e-6
    assert np.allclose(M, R, atol=1e-9)
    assert np.linalg.det(M) == pytest.approx(1.0, abs=1e-12)


def test_superpose_mirror_image_gives_proper_rotation():
    """det(X) < 0 triggers the reflection rule: the RMSD is the best proper
    rotation's, not the mirror's (zero)."""
    rng = np.random.default_rng(2)
    a = rng.normal(scale=5.0, size=(8, 3))
    b = a * np.array([1.0, 1.0, -1.0])
    rmsd, M, _, _ = oracle.superpose(a, b)
    R, r = kabsch(a, b)
    assert r > 0.5
    assert rmsd == pytest.approx(r, rel=1e-7)
    assert np.linalg.det(M) == pytest.approx(1.0, abs=1e-9)


def test_superpose_translation_invariance():
    rng = np.random.default_rng(3)
    a = rng.normal(scale=5.0, size=(7, 3))
    b = rng.normal(scale=5.0, size=(7, 3))
    r1 = oracle.superpose(a, b)[0]
    r2 = oracle.superpose(a + 100.0, b - 50.0)[0]
    assert r1 == pytest.approx(r2, rel=1e-9)


def test_superpose_single_point_rotation_is_nan():
    """X^T X of rank < 3 has no inverse square root: the rotation is not finite."""
    rmsd, M, _, _ = oracle.superpose(np.array([[1.0, 2.0, 3.0]]), np.array([[4.0, 5.0, 6.0]]))
    assert rmsd == 0.0
    assert not np.all(np.isfinite(M))


def tiny_case(seed, n_res=12):
    rng = np.random.Generator(np.random.PCG64(seed))
    mol = synthetic.protein(rng, n_res, n_chains=2 if seed % 2 else 1)
    tpls = [synthetic.planted_template(rng, mol, 2, 2, jitter=0.4, radius=12.0, p_flexible=0.5,
                                       tpl_id="T%d" % i) for i in range(3)]
    tpls.append(synthetic.random_template(rng, 2, 2, p_flexible=0.5))
    return mol, tpls


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("ignore_chain", [None, "residues", "atoms"])
def test_scanner_equals_bruteforce(seed, ignore_chain):
    mol, tpls = tiny_case(seed)
    stream = oracle.query_stream(mol, tpls, 2.0, 2.0, ignore_chain=ignore_chain)
    for t in range(len(tpls)):
        got = sorted(m.atoms for m in stream if m.template == t)
        want = sorted(oracle.bruteforce(mol, tpls, t, 2.0, 2.0, ignore_chain=ignore_chain))
        assert got == want
        assert len(set(got)) == len(got)


@pytest.mark.parametrize("seed", range(4))
def test_reorder_keeps_the_match_set(seed):
    mol, tpls = tiny_case(100 + seed, n_res=30)
    a = sorted((m.template, m.atoms) for m in oracle.query_stream(mol, tpls, 2.5, 2.5, reorder=True))
    b = sorted((m.template, m.atoms) for m in oracle.query_stream(mol, tpls, 2.5, 2.5, reorder=False))
    assert a == b


def check_median_split(coords, order, lo, n, axis):
    if n <= 1:
        return
    split = n // 2
    left = coords[[order[i] for i in range(lo, lo + split)], axis]
    right = coords[[order[i] for i in range(lo + split, lo + n)], axis]
    assert left.max() <= right.min()
    check_median_split(coords, order, lo, split, (axis + 1) % 3)
    check_median_split(coords, order, lo + split, n - split, (axis + 1) % 3)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 100, 257])
def test_kd_order_is_a_median_split(n):
    rng = np.random.Generator(np.random.PCG64(n))
    mol = synthetic.protein(rng, max(3, n // 4 + 3))
    atoms = rng.choice(len(mol), size=min(n, len(mol)), replace=False)
    order = oracle.kd_order(mol, atoms)
    assert sorted(order) == list(range(len(atoms)))
    check_median_split(mol.xyz[atoms], order, 0, len(atoms), 0)


def test_kd_order_with_ties():
    """Equal coordinates (3-decimal data) still give a valid median split."""
    rng = np.random.Generator(np.random.PCG64(7))
    mol = synthetic.protein(rng, 40)
    mol.xyz = np.round(mol.xyz / 2.0) * 2.0       # many ties on every axis
    atoms = np.arange(len(mol))
    order = oracle.kd_order(mol, atoms)
    check_median_split(mol.xyz[atoms], order, 0, len(atoms), 0)
