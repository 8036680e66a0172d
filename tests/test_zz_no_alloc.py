"""jam_superpose_batch_noalloc -- superposition with its inputs and outputs
staged in a static device array instead of allocated buffers -- against
jam_superpose_batch, through the C ABI: identical results (bit for bit) on two
calls in a row, the second smaller than the first, and a clear error past the
array's size."""

import ctypes

import numpy as np
import pytest

from pyjess_amd import _native, synthetic

pytestmark = pytest.mark.gpu


def _call(fn, a, b):
    n, k, _ = a.shape
    rmsd = np.full(n, -1.0)
    rot = np.full((n, 9), -1.0)
    c0 = np.full((n, 3), -1.0)
    c1 = np.full((n, 3), -1.0)
    ctx = _native.context()
    with ctx.lock:
        rc = fn(ctx.handle, _native._ptr(a), _native._ptr(b), ctypes.c_int32(k), ctypes.c_int64(n),
                _native._ptr(rmsd), _native._ptr(rot), _native._ptr(c0), _native._ptr(c1))
    return rc, (rmsd, rot, c0, c1)


def _sets(rng, n, k):
    a, b, off = synthetic.superposition_pairs(rng, n, k_min=k, k_max=k, collinear=0.0, n_proteins=1)
    assert np.array_equal(np.diff(off), np.full(n, k))
    return (np.ascontiguousarray(a.reshape(n, k, 3)), np.ascontiguousarray(b.reshape(n, k, 3)))


def test_noalloc_equals_allocating_entry_point_on_two_calls():
    L = _native.lib()
    rng = np.random.Generator(np.random.PCG64(77))
    big = _sets(rng, 40000, 6)      # ~16.6 MB of the 32 MiB array
    small = _sets(rng, 3000, 9)     # smaller: stale contents of the first call must not show
    print("noalloc: 40000 x 6 then 3000 x 9 point sets", flush=True)
    want = [_call(L.jam_superpose_batch, *x) for x in (big, small)]
    got = [_call(L.jam_superpose_batch_noalloc, *x) for x in (big, small)]
    for (rw, w), (rg, g) in zip(want, got):
        assert rw == 0 and rg == 0, _native.lib().jam_last_error()
        for x, y in zip(w, g):
            assert np.array_equal(x, y, equal_nan=True)
    # results are real superpositions, not untouched outputs
    assert np.all(got[1][1][0] >= 0.0) and np.all(np.isfinite(got[1][1][0]))


def test_noalloc_rejects_a_batch_beyond_its_scratch():
    L = _native.lib()
    k = 12
    per_set = 8 * (6 * k + 16)
    n = (32 << 20) // per_set + 1
    a = np.zeros((n, k, 3))
    rc, _ = _call(L.jam_superpose_batch_noalloc, a, a)
    assert rc == _native.JAM_ETOOBIG
    assert b"scratch" in L.jam_last_error()
