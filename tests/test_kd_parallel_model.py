"""The parallel kd arrangement of the device (presort every axis, cut each range
in half on axis d%3, stably partition the other axes) is the reference's serial
QuickSelect arrangement whenever no tie straddles a split point; when one does,
the device falls back to the serial algorithm. This checks the claim on the
CPU with a numpy model of the parallel algorithm against the oracle."""

import numpy as np
import pytest

from oracle import oracle
from pyjess_amd import synthetic


def parallel_model(c):
    """Returns the arrangement, or None when a tie straddles a split point."""
    n = len(c)
    perm = [sorted(range(n), key=lambda i: (c[i, a], i)) for a in range(3)]
    sb = [0] * n
    se = [n] * n
    d = 0
    while True:
        a = d % 3
        side = {}
        split = False
        for p in range(n):
            s, e = sb[p], se[p]
            if e - s < 2:
                continue
            m = s + (e - s) // 2
            side[perm[a][p]] = p < m
            if p == m:
                split = True
                if c[perm[a][m - 1], a] == c[perm[a][m], a]:
                    return None
        if not split:
            break
        for b in range(3):
            if b == a:
                continue
            new = list(perm[b])
            for p in range(n):
                s, e = sb[p], se[p]
                if e - s < 2:
                    continue
                m = s + (e - s) // 2
                lc = sum(1 for q in range(s, p) if side[perm[b][q]])
                np_ = s + lc if side[perm[b][p]] else m + (p - s - lc)
                new[np_] = perm[b][p]
            perm[b] = new
        for p in range(n):
            s, e = sb[p], se[p]
            if e - s < 2:
                continue
            m = s + (e - s) // 2
            if p < m:
                se[p] = m
            else:
                sb[p] = m
        d += 1
    assert perm[0] == perm[1] == perm[2]
    return perm[0]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 16, 33, 100, 257])
@pytest.mark.parametrize("decimals", [3, 0])
def test_parallel_model_equals_quickselect(n, decimals):
    rng = np.random.Generator(np.random.PCG64(n * 10 + decimals))
    mol = synthetic.protein(rng, max(3, n // 4 + 3))
    if decimals == 0:
        mol.xyz = np.round(mol.xyz)       # many ties
    atoms = rng.choice(len(mol), size=min(n, len(mol)), replace=False)
    got = parallel_model(mol.xyz[atoms])
    want = oracle.kd_order(mol, atoms)
    if got is None:
        assert decimals == 0                  # only coarse data ties at a split
    else:
        assert got == want


def test_tie_free_sets_are_common_on_pdb_precision_data():
    rng = np.random.Generator(np.random.PCG64(99))
    mol = synthetic.protein(rng, 120)
    ok = 0
    for _ in range(30):
        atoms = rng.choice(len(mol), size=60, replace=False)
        got = parallel_model(mol.xyz[atoms])
        if got is not None:
            assert got == oracle.kd_order(mol, atoms)
            ok += 1
    assert ok >= 25
