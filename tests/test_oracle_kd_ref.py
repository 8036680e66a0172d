"""The oracle's kd-tree leaf order against leaf orders computed with the
reference's own QuickSelect (patches/qselect.h.patch, built into oracle/_ref by
``make -C oracle ref``; vectors written by tests/golden/make_kd_ref_vectors.py).

The sets are full of equal coordinates (integer grids, duplicated points, a
constant axis, all points equal), where the swap sequence of the median-of-3
Lomuto QuickSelect decides which equal points go left of each median; that
order is the order in which the join meets candidates, hence the hit order."""

import os

import numpy as np
import pytest

from oracle import oracle

from conftest import golden

VEC = np.load(golden("kd_ref_orders.npz"))


def ref_sets():
    off = VEC["off"]
    for s in range(len(off) - 1):
        a, b = int(off[s]), int(off[s + 1])
        yield str(VEC["style"][s]), VEC["q"][a:b].astype(np.float64) / 8.0, VEC["order"][a:b].astype(np.int64)


def mol_of(xyz):
    n = xyz.shape[0]
    return oracle.MolView(xyz, np.full((n, 4), ord("_"), np.uint8), np.full((n, 4), 0, np.uint8),
                          np.zeros(n, np.uint8), np.zeros(n, np.int32), np.arange(n, dtype=np.int32))


def test_vectors_cover_ties_and_sizes():
    sizes = np.diff(VEC["off"])
    assert len(sizes) >= 200
    assert sizes.min() == 2 and sizes.max() == 8193
    # both sides of the device's size limits: one position per thread up to
    # 1,024 entries, LDS up to 4,096 and 8,192, global scratch above
    assert {1024, 1025, 4096, 4097, 8192, 8193} <= set(sizes.tolist())
    styles = set(str(s) for s in VEC["style"])
    assert {"int5", "grid2", "int_chain", "flat_x", "dup_pairs", "same", "binary", "distinct"} <= styles


@pytest.mark.parametrize("style", ["int5", "grid2", "int_chain", "flat_x", "dup_pairs", "same", "binary",
                                   "distinct"])
def test_oracle_kd_order_equals_reference_qselect(style):
    n_sets = 0
    for st, xyz, want in ref_sets():
        if st != style:
            continue
        n = xyz.shape[0]
        got = oracle.kd_order(mol_of(xyz), np.arange(n, dtype=np.int32))
        assert got == want.tolist(), (style, n)
        n_sets += 1
    assert n_sets > 0


def test_reference_orders_are_permutations_and_median_splits():
    """The vectors themselves: a permutation per set, and at the root the left
    half holds no x above the right half's smallest x (QuickSelect's property)."""
    for _, xyz, order in ref_sets():
        n = xyz.shape[0]
        assert sorted(order.tolist()) == list(range(n))
        h = n // 2
        assert xyz[order[:h], 0].max() <= xyz[order[h:], 0].min()
