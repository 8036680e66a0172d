"""Model of the device's exact parallel kd arrangement, checked on the CPU
against the oracle's serial QuickSelect (patches/qselect.h.patch:50-116).

A Lomuto partition of a window [l, r] with the pivot at r sends the elements
<= pivot to [l, l+k) in their original order. The others behave as a FIFO
queue: each one is pushed when scanned, and each "<=" element scanned after the
first ">" one pops the queue front and pushes it back. The queue's final
content fills [l+k+1, r-1] from its second element on, and its front goes to r
when the pivot is swapped into l+k. Push number t of a re-push is the q-th pop,
which returns push number q < t: pointer jumping resolves every push to an
element in O(log) rounds. The model applies this to all ranges of a tree level
at once, as the kernel does."""

import numpy as np
import pytest

from oracle import oracle
from pyjess_amd import synthetic


def lomuto_parallel(a, c, l, r):
    """Parallel formulation of one partition of a[l..r] (pivot already at r)."""
    piv = a[r]
    b = [c[a[i]] <= c[piv] for i in range(l, r)]
    ones = np.cumsum([0] + b)              # exclusive prefix
    zeros_before = [(i - l) - ones[i - l] for i in range(l, r)]
    rot = [b[i - l] and zeros_before[i - l] > 0 for i in range(l, r)]
    rots = np.cumsum([0] + rot)
    k = int(ones[-1])
    m = (r - l) - k
    n_rot = int(rots[-1])
    src = [None] * (m + n_rot + 1)         # 1-based push numbers
    for i in range(l, r):
        j = i - l
        t = zeros_before[j] + int(rots[j]) + 1
        if not b[j]:
            src[t] = ("elem", a[i])
        elif rot[j]:
            src[t] = ("ptr", int(rots[j]) + 1)
    # pointer jumping
    changed = True
    while changed:
        changed = False
        new = list(src)
        for t in range(1, len(src)):
            if src[t] is not None and src[t][0] == "ptr":
                new[t] = src[src[t][1]]
                changed = True
        src = new
    out = list(a)
    for i in range(l, r):
        j = i - l
        if b[j]:
            out[l + int(ones[j])] = a[i]
    out[l + k] = piv
    for q in range(1, m + 1):
        e = src[n_rot + q][1]
        out[r if q == 1 else l + k + q - 1] = e
    return out, l + k


def lomuto_serial(a, c, l, r):
    a = list(a)
    store = l
    for i in range(l, r):
        if c[a[i]] <= c[a[r]]:
            a[i], a[store] = a[store], a[i]
            store += 1
    a[store], a[r] = a[r], a[store]
    return a, store


def median3(a, c, l, r):
    mid = (l + r + 1) // 2
    def cmp(i, j):
        return int(c[a[i]] > c[a[j]]) - int(c[a[i]] < c[a[j]])
    if (cmp(l, mid) > 0) != (cmp(l, r) > 0):
        return l
    if (cmp(mid, l) < 0) != (cmp(mid, r) < 0):
        return mid
    return r


def arrange_levelwise(coords):
    """All ranges of a level run their QuickSelect together, one partition
    round at a time, as in the kernel."""
    n = len(coords)
    a = list(range(n))
    ranges = [(0, n)]
    d = 0
    while ranges:
        axis = coords[:, d % 3]
        state = {}
        for lo, hi in ranges:
            if hi - lo >= 2:
                state[lo] = [lo, hi - 1, lo + (hi - lo) // 2, False]   # wl, wr, k, done
        while any(not s[3] for s in state.values()):
            for lo, s in state.items():
                if s[3]:
                    continue
                wl, wr, k, _ = s
                p = median3(a, axis, wl, wr)
                a[wr], a[p] = a[p], a[wr]
            for lo, s in state.items():
                if s[3]:
                    continue
                wl, wr, k, _ = s
                a, p = lomuto_parallel(a, axis, wl, wr)
                if p == k:
                    s[3] = True
                elif k < p:
                    s[1] = p - 1
                else:
                    s[0] = p + 1
                if not s[3] and s[0] == s[1]:
                    s[3] = True
        nxt = []
        for lo, hi in ranges:
            if hi - lo >= 2:
                m = lo + (hi - lo) // 2
                nxt += [(lo, m), (m, hi)]
        ranges = nxt
        d += 1
    return a


@pytest.mark.parametrize("seed", range(30))
def test_parallel_lomuto_equals_serial(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    c = rng.integers(0, 6, size=n).astype(float)      # heavy ties
    a = list(rng.permutation(n))
    l = int(rng.integers(0, n - 1))
    r = int(rng.integers(l + 1, n))
    assert lomuto_parallel(a, c, l, r) == lomuto_serial(a, c, l, r)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 100, 301])
@pytest.mark.parametrize("decimals", [3, 0])
def test_levelwise_exact_arrangement(n, decimals):
    rng = np.random.Generator(np.random.PCG64(n + 1000 * decimals))
    mol = synthetic.protein(rng, max(3, n // 4 + 3))
    if decimals == 0:
        mol.xyz = np.round(mol.xyz)
    atoms = rng.choice(len(mol), size=min(n, len(mol)), replace=False)
    assert arrange_levelwise(mol.xyz[atoms]) == oracle.kd_order(mol, atoms)
