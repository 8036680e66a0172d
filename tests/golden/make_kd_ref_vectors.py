"""Writes tests/golden/kd_ref_orders.npz: kd-tree leaf orders computed with the
reference's own QuickSelect (patches/qselect.h.patch), for point sets full of
equal coordinates, where the exact swap sequence of QuickSelect decides which
of the equal points go left of each median.

The library comes from ``make -C oracle ref`` (oracle/kd_ref.c compiled
against the header the reference's patch creates); it needs the reference
checkout, so this script runs where that checkout is, and the vectors are
committed. Run from the repository root: ``python tests/golden/make_kd_ref_vectors.py``.

Each set is stored as int16 coordinates q (the point is q / 8 Angstrom, exact
in binary floating point) and its leaf order (left subtree first) as int16.
"""

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ORACLE = os.path.join(ROOT, "oracle")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kd_ref_orders.npz")

SMALL = [2, 3, 4, 5, 6, 7, 8, 9, 11, 16, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128, 129]
LARGE = [256, 257, 513, 627, 1024, 1025, 2000]
HUGE_TIES = [512, 1024, 2000]
# around the device's LDS limits (4,096 and 8,192 entries), appended after
# the sets above so that those keep their coordinates
BIG = {"int5": [4096, 4097, 6000, 8192, 8193], "int_chain": [4096, 4097, 6000, 8192, 8193],
       "dup_pairs": [4097, 8192]}


def walk(rng, n, step=3.8):
    """A chain of points `step` apart in random directions (protein-like)."""
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.cumsum(d * step, axis=0)


def style(rng, name, n):
    """int16 coordinates (units of 1/8 Angstrom) of one set."""
    if name == "int5":          # integers 0..5: ties everywhere
        q = rng.integers(0, 6, size=(n, 3)) * 8
    elif name == "grid2":       # a chain rounded to multiples of 2 Angstrom
        q = np.round(walk(rng, n) / 2.0) * 16
    elif name == "int_chain":   # a chain rounded to whole Angstroms
        q = np.round(walk(rng, n)) * 8
    elif name == "flat_x":      # one axis constant, the others spread out
        q = rng.integers(-20000, 20001, size=(n, 3))
        q[:, 0] = 40
    elif name == "dup_pairs":   # every point twice, shuffled
        h = (n + 1) // 2
        base = rng.integers(-400, 401, size=(h, 3))
        q = np.concatenate([base, base])[:n]
        q = q[rng.permutation(n)]
    elif name == "same":        # every point equal
        q = np.full((n, 3), 24)
    elif name == "binary":      # coordinates 0 or 1 Angstrom
        q = rng.integers(0, 2, size=(n, 3)) * 8
    elif name == "distinct":    # control: almost no ties
        q = rng.integers(-20000, 20001, size=(n, 3))
    else:
        raise ValueError(name)
    return np.asarray(q, dtype=np.int16)


def plan():
    sets = []
    for name in ["int5", "grid2", "int_chain", "flat_x", "dup_pairs", "same", "binary", "distinct"]:
        for n in SMALL:
            sets.append((name, n))
        if name in ("int5", "grid2", "int_chain", "flat_x", "dup_pairs"):
            for n in LARGE:
                sets.append((name, n))
        if name in ("same", "binary"):
            for n in HUGE_TIES:
                sets.append((name, n))
    for name, sizes in BIG.items():
        for n in sizes:
            sets.append((name, n))
    return sets


def main():
    subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)
    lib = ctypes.CDLL(os.path.join(ORACLE, "_ref", "libkd_ref.so"))
    lib.kdref_leaf_order.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    lib.kdref_leaf_order.restype = ctypes.c_int32
    rng = np.random.Generator(np.random.PCG64(20261203))
    names, sizes, coords, orders = [], [], [], []
    for name, n in plan():
        q = style(rng, name, n)
        xyz = np.ascontiguousarray(q.astype(np.float64) / 8.0)
        out = np.zeros(n, dtype=np.int32)
        if lib.kdref_leaf_order(xyz.ctypes.data, n, out.ctypes.data) != n:
            raise RuntimeError("kdref_leaf_order failed")
        assert sorted(out.tolist()) == list(range(n))
        names.append(name)
        sizes.append(n)
        coords.append(q)
        orders.append(out.astype(np.int16))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    np.savez_compressed(OUT, style=np.array(names), off=off, q=np.concatenate(coords),
                        order=np.concatenate(orders))
    print(f"{len(sizes)} sets, {int(off[-1])} points -> {OUT}")


if __name__ == "__main__":
    main()
