"""The event stream of jam_query: delivered in report order (molecule,
template in scan order, rank), kept in device memory until the next query on
the context, and refused with JAM_ESTALE after that."""

import ctypes

import numpy as np
import pytest

from pyjess_amd import Jess, synthetic
from pyjess_amd import _native

pytestmark = pytest.mark.gpu


def device():
    if not _native.device_available():
        pytest.skip("no GPU")


def test_events_arrive_in_report_order():
    device()
    mols, tpls, kw = synthetic.workload("batch-4k", seed=3, scale=6 / 4096)
    jess = Jess(tpls)
    res = jess._device_query(mols, kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"],
                             0, True)
    T = len(tpls)
    assert len(res.pair) > 100
    scan = (res.pair // T) * T + (T - 1 - res.pair % T)     # molecule, then reverse template order
    assert np.all(np.diff(scan) >= 0)
    same = np.diff(scan) == 0
    assert np.all(np.diff(res.rank)[same] > 0)
    # every event's rank is below its pair's number of complete matches
    assert np.all(res.rank < res.pair_matches[res.pair])


def test_result_is_stale_after_the_next_query():
    device()
    mols, tpls, kw = synthetic.workload("plain-2k", seed=1)
    jess = Jess(tpls)
    ctx = _native.context()
    tset = jess._device_templates(ctx)
    mset = _native.MoleculeSet(ctx, mols)
    L = _native.lib()
    p = _native.Params(kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"], 0, 1)
    h1, h2 = ctypes.c_void_p(), ctypes.c_void_p()
    with ctx.lock:
        _native._check(L.jam_query(ctx.handle, tset.handle, mset.handle, ctypes.byref(p), ctypes.byref(h1)))
        try:
            n = L.jam_result_n_events(h1)
            assert n > 0
            first = np.zeros(n, dtype=np.float64)
            assert L.jam_result_copy(h1, None, None, None, _native._ptr(first), None, None, None, None) == 0
            _native._check(L.jam_query(ctx.handle, tset.handle, mset.handle, ctypes.byref(p), ctypes.byref(h2)))
            try:
                again = np.zeros(n, dtype=np.float64)
                assert L.jam_result_copy(h1, None, None, None, _native._ptr(again), None, None, None,
                                         None) == _native.JAM_ESTALE
                assert L.jam_result_copy(h2, None, None, None, _native._ptr(again), None, None, None, None) == 0
                assert np.array_equal(first, again)
            finally:
                L.jam_result_free(h2)
        finally:
            L.jam_result_free(h1)


def test_fresh_context_grows_and_reruns():
    # a new context starts with small buffers: the first large query finds out
    # on the device that its sets, items or frontier do not fit, and is run
    # again after the host grows them; small and large queries then alternate
    device()
    from oracle import oracle
    from bench import pair_digests

    ctx = _native.Context(0)
    big_mols, big_tpls, kw = synthetic.workload("batch-4k", seed=5, scale=24 / 4096)
    small_mols, small_tpls, _ = synthetic.workload("plain-2k", seed=5)
    q = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    runs = [(small_mols, small_tpls), (big_mols, big_tpls), (small_mols, small_tpls), (big_mols, big_tpls)]
    for mols, tpls in runs:
        tset = _native.TemplateSet(ctx, tpls)
        mset = _native.MoleculeSet(ctx, mols)
        res = _native.query(tset, mset, *q, 0, True)
        cm, hc, dg = oracle.batch([oracle.MolView.of(m) for m in mols], oracle.TplView(tpls), *q, n_threads=4)
        T = len(tpls)
        sel = res.flag == 1
        gpu_hc, gpu_dg = pair_digests(res.pair[sel], res.atoms[sel], [len(t) for t in tpls], T, len(mols) * T)
        assert np.array_equal(res.pair_matches, cm)
        assert np.array_equal(gpu_hc, hc)
        assert np.array_equal(gpu_dg, dg.astype(np.uint64))


def test_device_pair_digests_equal_the_oracle():
    """jam_result_pair_digests (per-pair hit counts and digests computed on the
    device from the events in HBM, as bench.py checks the library workload)
    against the oracle's, with the molecules numbered from an offset."""
    device()
    import ctypes

    from oracle import oracle

    mols, tpls, kw = synthetic.workload("batch-4k", seed=9, scale=12 / 4096)
    q = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    T = len(tpls)
    cm, hc, dg = oracle.batch([oracle.MolView.of(m) for m in mols], oracle.TplView(tpls), *q, n_threads=4)
    assert hc.sum() > 0
    # the last 7 molecules as a batch of their own, numbered from 5
    ctx = _native.context()
    tset = _native.TemplateSet(ctx, tpls)
    mset = _native.MoleculeSet(ctx, mols[5:])
    L = _native.lib()
    p = _native.Params(*q, 0, 1)
    h = ctypes.c_void_p()
    with ctx.lock:
        _native._check(L.jam_query(ctx.handle, tset.handle, mset.handle, ctypes.byref(p), ctypes.byref(h)))
        try:
            n = L.jam_result_n_pairs(h)
            gh = np.zeros(n, dtype=np.int64)
            gd = np.zeros(n, dtype=np.uint64)
            _native._check(L.jam_result_pair_digests(h, 5, T, _native._ptr(gh), _native._ptr(gd)))
        finally:
            L.jam_result_free(h)
    assert np.array_equal(gh, hc[5 * T:])
    assert np.array_equal(gd, dg[5 * T:].astype(np.uint64))
