"""The host-side policy of the product (events in report order, pyjess_amd._jess._policy)
against the literal loop of Query.__next__ (_jess.pyx:2197-2283) restated in the
oracle, on random streams of complete matches with hits, NaN rotations and
sub-threshold matches, for every best_match / max_candidates combination."""

import math
import types

import numpy as np
import pytest

from oracle import oracle
from pyjess_amd import Jess, Template, TemplateAtom, synthetic
from pyjess_amd._jess import _policy


def two_atom_template(i):
    atoms = [TemplateAtom(chain_id="A", residue_number=1, x=0.0, y=0.0, z=float(k), residue_names=["ALA"],
                          atom_names=["CA"], match_mode=0) for k in range(2)]
    return Template(atoms, id=f"T{i}")


def random_stream(rng, n_templates, thr):
    stream = []
    uid = 0
    for scan in range(n_templates):
        t = n_templates - 1 - scan            # scan order is reverse insertion order
        for _ in range(int(rng.integers(0, 25))):
            r = float(rng.uniform(0.0, 3.0))
            if rng.random() < 0.15:
                r = float(rng.choice([0.5, 1.0]))   # exact ties for best_match
            nan = rng.random() < 0.1
            rot = (math.nan,) * 9 if nan else (1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0)
            stream.append(oracle.Match(scan, t, (uid, uid + 1), r, rot, (0.0,) * 3, (0.0,) * 3))
            uid += 2
    return stream


def events_of(stream, thr, n_templates):
    pos = {}
    rows = []
    for m in stream:
        rank = pos.get(m.template, 0)
        pos[m.template] = rank + 1
        if m.rmsd <= thr:
            rows.append((m.template, rank, 2 if m.nan else 1, m.rmsd, m.rotation, m.atoms))
    # the stream is in scan order with ranks increasing: the order in which
    # the device delivers its events
    res = types.SimpleNamespace()
    res.pair = np.array([r[0] for r in rows], dtype=np.int32)
    res.rank = np.array([r[1] for r in rows], dtype=np.int64)
    res.flag = np.array([r[2] for r in rows], dtype=np.int32)
    res.rmsd = np.array([r[3] for r in rows], dtype=np.float64)
    res.rotation = np.array([r[4] for r in rows], dtype=np.float64).reshape(-1, 9)
    res.c0 = np.zeros((len(rows), 3))
    res.c1 = np.zeros((len(rows), 3))
    res.atoms = np.array([r[5] for r in rows], dtype=np.int32).reshape(-1, 2)
    # complete matches per pair (before the RMSD threshold), as the device reports them
    res.pair_matches = np.array([pos.get(t, 0) for t in range(n_templates)], dtype=np.int64)
    return res


@pytest.mark.parametrize("seed", range(40))
@pytest.mark.parametrize("best_match", [False, True])
@pytest.mark.parametrize("max_candidates", [None, 0, 1, 2, 3, 7])
def test_policy_matches_query_loop(seed, best_match, max_candidates):
    rng = np.random.default_rng(seed)
    thr = 2.0
    n_t = int(rng.integers(1, 6))
    stream = random_stream(rng, n_t, thr)
    want, warns = oracle.replay(stream, thr, best_match, max_candidates)
    jess = Jess([two_atom_template(i) for i in range(n_t)])
    mol = synthetic.protein(np.random.Generator(np.random.PCG64(0)), 400)
    mc = -1 if max_candidates is None else max_candidates
    got = _policy(jess, [mol], events_of(stream, thr, n_t), best_match, mc)[0][0]
    got_hits = [h for h in got if h is not None]
    assert len(got) - len(got_hits) == warns
    assert [(jess._templates.index(h.template()), tuple(int(a.serial) - 1 for a in h.atoms(False)))
            for h in got_hits] == [(m.template, m.atoms) for m in want]
    assert [h.rmsd for h in got_hits] == [m.rmsd for m in want]


@pytest.mark.parametrize("seed", range(30))
@pytest.mark.parametrize("best_match", [False, True])
@pytest.mark.parametrize("max_candidates", [None, 0, 1, 2, 3, 7])
def test_lazy_hit_lists_match_query_loop(seed, best_match, max_candidates):
    """The lazy lists of `query_batch` (keep_none=False: NaN rotations left
    out and counted; without best_match the candidate cap is applied only to
    the pairs with more than max_candidates + 1 complete matches) against the
    same loop, for two molecules in one batch."""
    rng = np.random.default_rng(1000 + seed)
    thr = 2.0
    n_t = int(rng.integers(1, 6))
    streams = [random_stream(rng, n_t, thr) for _ in range(2)]
    parts = [events_of(s, thr, n_t) for s in streams]
    res = types.SimpleNamespace()
    for name in ("rank", "flag", "rmsd", "rotation", "c0", "c1", "atoms"):
        setattr(res, name, np.concatenate([getattr(r, name) for r in parts]))
    res.pair = np.concatenate([r.pair + m * n_t for m, r in enumerate(parts)]).astype(np.int32)
    res.pair_matches = np.concatenate([r.pair_matches for r in parts])
    jess = Jess([two_atom_template(i) for i in range(n_t)])
    mol = synthetic.protein(np.random.Generator(np.random.PCG64(0)), 400)
    mc = -1 if max_candidates is None else max_candidates
    lists, nan = _policy(jess, [mol, mol], res, best_match, mc, keep_none=False)
    for m, stream in enumerate(streams):
        want, warns = oracle.replay(stream, thr, best_match, max_candidates)
        got = list(lists[m])
        assert int(nan[m]) == warns
        assert [(jess._templates.index(h.template()), tuple(int(a.serial) - 1 for a in h.atoms(False)))
                for h in got] == [(w.template, w.atoms) for w in want]
        assert [h.rmsd for h in got] == [w.rmsd for w in want]
