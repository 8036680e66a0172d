"""The hit-set digest bench.py uses for parity (vectorised over hits) equals
the oracle's own per-pair digest and the scalar definition, on a small
synthetic batch (no GPU: the hits come from the oracle's replay)."""

import numpy as np

from oracle import oracle
from pyjess_amd import synthetic

import bench


def test_vectorised_digest_matches_oracle_and_scalar_definition():
    mols, tpls, kw = synthetic.workload("batch-4k", seed=2, scale=3 / 4096)
    q = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    T = len(tpls)
    cm, hc, dg = oracle.batch([oracle.MolView.of(m) for m in mols], oracle.TplView(tpls), *q, n_threads=2)
    pairs, rows = [], []
    width = max(len(t) for t in tpls)
    for mi, m in enumerate(mols):
        stream = oracle.query_stream(m, tpls, kw["distance_cutoff"], kw["max_dynamic_distance"])
        for h in stream:
            if h.rmsd <= kw["rmsd_threshold"] and not h.nan:
                pairs.append(mi * T + h.template)
                rows.append(list(h.atoms) + [-1] * (width - len(h.atoms)))
    assert len(pairs) > 10
    pairs = np.array(pairs, dtype=np.int32)
    rows = np.array(rows, dtype=np.int32)
    counts, digests = bench.pair_digests(pairs, rows, [len(t) for t in tpls], T, len(mols) * T)
    assert np.array_equal(counts, hc)
    assert np.array_equal(digests, dg.astype(np.uint64))
    # the scalar definition, summed mod 2^64
    want = [0] * (len(mols) * T)
    for p, r in zip(pairs, rows):
        m, t = divmod(int(p), T)
        want[p] = (want[p] + bench.hit_digest(t, m, r[:len(tpls[t])])) & bench._M64
    assert [int(x) for x in digests] == want


def test_bench_refuses_more_than_one_gpu():
    # one GPU per run: --gpus above 1 exits with status 2 and says why,
    # before anything touches a device
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=root)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout.strip() == ""


def test_native_oracle_build_gives_the_same_results():
    """The host-tuned build used to time the CPU baseline (-march=native,
    contraction still off) counts and digests the same hits as the portable
    build used as the checker."""
    mols, tpls, kw = synthetic.workload("batch-4k", seed=3, scale=2 / 4096)
    q = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    mv = [oracle.MolView.of(m) for m in mols]
    tv = oracle.TplView(tpls)
    a = oracle.batch(mv, tv, *q, n_threads=2)
    b = oracle.batch(mv, tv, *q, n_threads=2, native=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert a[1].sum() > 0
