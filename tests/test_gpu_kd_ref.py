"""The device's kd-tree leaf order (jam_kd_order: the passes that order every
candidate set of a query) against leaf orders computed with the reference's
own QuickSelect (patches/qselect.h.patch; vectors in
tests/golden/kd_ref_orders.npz, written by tests/golden/make_kd_ref_vectors.py).

The sets are full of equal coordinates, so most of them go through the exact
pass (the swap sequence of median-of-3 Lomuto QuickSelect); they span every
size tier of the device (up to 64 entries in one wave, up to 1024 with one
position per thread, larger in global scratch)."""

import numpy as np
import pytest

from pyjess_amd import _native

from conftest import golden

pytestmark = pytest.mark.gpu

VEC = np.load(golden("kd_ref_orders.npz"))


def test_device_kd_order_equals_reference_qselect():
    off = VEC["off"].astype(np.int64)
    xyz = VEC["q"].astype(np.float64) / 8.0
    want = VEC["order"].astype(np.int32)
    print(f"kd_order: {len(off) - 1} sets, {int(off[-1])} points", flush=True)
    got = _native.kd_order(xyz, off)
    bad = [s for s in range(len(off) - 1) if not np.array_equal(got[off[s]:off[s + 1]], want[off[s]:off[s + 1]])]
    assert not bad, [(str(VEC["style"][s]), int(off[s + 1] - off[s])) for s in bad[:10]]


def test_device_kd_order_sets_one_by_one_and_repeated():
    """Each set alone (a different batch shape), twice: the same order."""
    off = VEC["off"].astype(np.int64)
    xyz = VEC["q"].astype(np.float64) / 8.0
    want = VEC["order"].astype(np.int32)
    for s in range(0, len(off) - 1, 7):
        a, b = int(off[s]), int(off[s + 1])
        one = np.array([0, b - a], dtype=np.int64)
        for _ in range(2):
            assert np.array_equal(_native.kd_order(xyz[a:b], one), want[a:b]), (str(VEC["style"][s]), b - a)


def test_device_kd_order_rejects_bad_offsets():
    xyz = np.zeros((4, 3))
    with pytest.raises(ValueError):
        _native.kd_order(xyz, np.array([0, 3], dtype=np.int64))
    with pytest.raises(_native.NativeError):
        _native._check(_native.lib().jam_kd_order(_native.context().handle, _native._ptr(xyz),
                                                  _native._ptr(np.array([0, 3, 2, 4], dtype=np.int64)), 3,
                                                  _native._ptr(np.zeros(4, dtype=np.int32))))
