"""`HitList`, the lazy per-molecule hit sequence of `Jess.query_batch`, on
hand-made result arrays (no GPU)."""

import collections.abc

import numpy as np
import pytest

from pyjess_amd import Molecule, Template, TemplateAtom
from pyjess_amd._jess import HitList, _Block

from conftest import golden


def make_list():
    mol = Molecule.load(golden("1AMY.pdb"))
    tpls = [Template.load(golden("1.3.3.tpl")), Template.load(golden("4.1.2.tpl"))]
    n_ev, T = 6, 2
    rot = np.arange(n_ev * 9, dtype=np.float64).reshape(n_ev, 9)
    c0 = np.arange(n_ev * 3, dtype=np.float64).reshape(n_ev, 3)
    c1 = -c0
    atoms = np.arange(n_ev * 8, dtype=np.int32).reshape(n_ev, 8)
    rmsd = np.linspace(0.1, 0.6, n_ev)
    pair = np.array([0, 1, 1, 0, 1, 1], dtype=np.int32)   # molecule 0, templates 0/1
    events = np.array([1, 3, 4], dtype=np.int64)
    return HitList(_Block(rot, c0, c1, atoms), rmsd, pair, events, T, tpls, mol), rot, rmsd, tpls, mol


def test_hitlist_is_a_sequence_of_hits_made_on_access():
    hl, rot, rmsd, tpls, mol = make_list()
    assert isinstance(hl, collections.abc.Sequence)
    assert len(hl) == 3 and bool(hl)
    hits = list(hl)
    assert [h.rmsd for h in hits] == [rmsd[1], rmsd[3], rmsd[4]]
    assert [h.template() for h in hits] == [tpls[1], tpls[0], tpls[1]]
    assert hits[1]._rotation == rot[3].tolist()
    assert hits[0].molecule() is mol or hits[0]._molecule is mol
    assert hl[-1].rmsd == rmsd[4]
    assert [h.rmsd for h in hl[1:]] == [rmsd[3], rmsd[4]]
    with pytest.raises(IndexError):
        hl[3]
    with pytest.raises(IndexError):
        hl[-4]


def test_hitlist_equality_and_empty():
    hl, *_ = make_list()
    again, *_ = make_list()
    assert hl == hl
    assert hl != []
    assert repr(hl) == "<HitList of 3 hits>"


def test_split_events_cuts_by_molecule_and_drops_nan():
    from pyjess_amd._hits import split_events

    T = 3
    pair = np.array([0, 2, 4, 5, 9, 10, 11], dtype=np.int32)   # molecules 0, 0, 1, 1, 3, 3, 3
    flag = np.array([1, 2, 1, 1, 1, 2, 1], dtype=np.int32)
    ev, bounds, nan = split_events(pair, flag, None, T, 5)
    assert ev.tolist() == [0, 2, 3, 4, 6]
    assert bounds.tolist() == [0, 1, 3, 3, 5, 5]
    assert nan.tolist() == [1, 0, 0, 1, 0]
    ev, bounds, nan = split_events(pair, flag, np.array([1, 3, 6], dtype=np.int64), T, 5)
    assert ev.tolist() == [3, 6]
    assert bounds.tolist() == [0, 0, 1, 1, 2, 2]
    assert nan.tolist() == [1, 0, 0, 0, 0]


def test_hit_lists_equal_one_hitlist_per_molecule():
    """`hit_lists` (all molecules' lists from one pass over the result
    arrays) gives the lists `HitList` gives one by one."""
    from pyjess_amd._hits import hit_lists, split_events

    mol = Molecule.load(golden("1AMY.pdb"))
    mols = [mol, mol, mol, mol]
    tpls = [Template.load(golden("1.3.3.tpl")), Template.load(golden("4.1.2.tpl"))]
    T = 2
    pair = np.array([0, 1, 1, 4, 5, 6, 7, 7], dtype=np.int32)   # molecules 0, 0, 0, 2, 2, 3, 3, 3
    flag = np.array([1, 1, 2, 1, 1, 1, 2, 1], dtype=np.int32)
    n_ev = len(pair)
    block = _Block(np.zeros((n_ev, 9)), np.zeros((n_ev, 3)), np.zeros((n_ev, 3)), np.zeros((n_ev, 8), np.int32))
    rmsd = np.linspace(0.1, 0.8, n_ev)
    ev, bounds, nan = split_events(pair, flag, None, T, len(mols))
    got = hit_lists(block, rmsd, pair, ev, bounds, T, tpls, mols)
    want = [HitList(block, rmsd, pair, ev[bounds[m]:bounds[m + 1]], T, tpls, mols[m]) for m in range(len(mols))]
    assert [len(x) for x in got] == [2, 0, 2, 2]
    assert got == want
    for g, w in zip(got, want):
        assert [(h.rmsd, h.template()) for h in g] == [(h.rmsd, h.template()) for h in w]
    with pytest.raises(ValueError):
        hit_lists(block, rmsd, pair, ev, bounds[:-1], T, tpls, mols)
    bad = bounds.copy()
    bad[2] = len(ev) + 1
    with pytest.raises(ValueError):
        hit_lists(block, rmsd, pair, ev, bad, T, tpls, mols)


def test_hit_list_ranges_equal_the_explicit_rows():
    """With every event chosen, `split_ranges` + `hit_list_ranges` (a range
    of rows per molecule, explicit rows only where NaN-rotation events are
    left out) give the lists `split_events` + `hit_lists` give."""
    from pyjess_amd._hits import hit_list_ranges, hit_lists, split_events, split_ranges

    mol = Molecule.load(golden("1AMY.pdb"))
    mols = [mol] * 5
    tpls = [Template.load(golden("1.3.3.tpl")), Template.load(golden("4.1.2.tpl")),
            Template.load(golden("1.3.3.tpl"))]
    T = 3
    rng = np.random.default_rng(7)
    m_of = np.sort(rng.choice([0, 1, 3, 4], size=60)).astype(np.int32)   # molecule 2: no events
    pair = (m_of * T + rng.integers(0, T, size=60)).astype(np.int32)
    flag = np.ones(60, dtype=np.int32)
    flag[[0, 5, 6, 59]] = 2                                              # NaN rotations
    flag[m_of == 3] = 1                                                  # molecule 3: none
    n_ev = len(pair)
    block = _Block(np.zeros((n_ev, 9)), np.zeros((n_ev, 3)), np.zeros((n_ev, 3)), np.zeros((n_ev, 8), np.int32))
    rmsd = rng.random(n_ev)
    ev, bounds, nan = split_events(pair, flag, None, T, len(mols))
    want = hit_lists(block, rmsd, pair, ev, bounds, T, tpls, mols)
    b2, nan2, nan_ev = split_ranges(pair, flag, T, len(mols))
    got = hit_list_ranges(block, rmsd, pair, b2, nan_ev, T, tpls, mols)
    assert np.array_equal(nan, nan2)
    assert list(nan_ev) == [i for i in range(n_ev) if flag[i] == 2]
    assert [len(x) for x in got] == [len(x) for x in want]
    assert len(got[2]) == 0 and not got[2]
    assert got == want and want == got
    for g, w in zip(got, want):
        # the RMSDs are distinct: they name the rows
        assert [(h.rmsd, h.template()) for h in g] == [(h.rmsd, h.template()) for h in w]
        if len(w):
            assert g[-1].rmsd == w[-1].rmsd and [h.rmsd for h in g[1:]] == [h.rmsd for h in w[1:]]
        with pytest.raises(IndexError):
            g[len(w)]


def test_split_ranges_rejects_events_out_of_molecule_order():
    from pyjess_amd._hits import split_ranges

    flag = np.ones(4, dtype=np.int32)
    with pytest.raises(ValueError):
        split_ranges(np.array([3, 0, 4, 5], dtype=np.int32), flag, 3, 2)   # molecules 1, 0, 1, 1
    with pytest.raises(ValueError):
        split_ranges(np.array([0, 1, 3, 9], dtype=np.int32), flag, 3, 2)   # molecule 3 of 2
    bounds, nan, nan_ev = split_ranges(np.zeros(0, dtype=np.int32), np.zeros(0, dtype=np.int32), 3, 2)
    assert list(bounds) == [0, 0, 0] and list(nan) == [0, 0] and len(nan_ev) == 0
