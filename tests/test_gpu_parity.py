"""Parity of the device path (libjess_amd.so on gfx950) with the CPU oracle and
the reference's golden data. Every query goes through the C ABI."""

import math
import warnings

import numpy as np
import pytest

from oracle import oracle
from pyjess_amd import Jess, Molecule, Template, TemplateAtom, synthetic
from pyjess_amd import _native

from conftest import DOCTEST_DUMP, DOCTEST_RMSD, golden, parse_hit_file

pytestmark = pytest.mark.gpu

def key(mol, h):
    return tuple(int(a.serial) for a in h.atoms(transform=False))


def same_bits(x, y) -> bool:
    """Equal FP64 values (NaN equal to NaN): both sides compute with the same
    operations in the same order, without contraction."""
    return np.array_equal(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64), equal_nan=True)


def assert_same_hits(mol, tpls, got, want):
    """Device hits equal oracle hits: same templates, atoms and order; RMSD,
    rotation and centroids bit for bit."""
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g.template() is tpls[w.template]
        assert key(mol, g) == tuple(int(mol.serial[a]) for a in w.atoms)
        assert same_bits(g.rmsd, w.rmsd)
        assert same_bits(np.ravel(g._rotation), w.rotation)
        assert same_bits(g._centre[0], w.c0)
        assert same_bits(g._centre[1], w.c1)


def test_device_is_gfx950():
    assert _native.device_available()


def test_fixture_1amy_133_in_jess_order(mol_1amy, tpl_133):
    blocks = parse_hit_file(golden("1AMY+1.3.3.txt"))
    hits = list(Jess([tpl_133]).query(mol_1amy, 2, 4, 4, reorder=False))
    assert len(hits) == 156
    for h, (head, lines) in zip(hits, blocks):
        assert h.rmsd == pytest.approx(float(head[1]), abs=5e-4)
        assert h.determinant == pytest.approx(float(head[4]), abs=0.05)
        assert h.log_evalue == pytest.approx(float(head[6]), abs=5e-3)
        for a, line in zip(h.atoms(), lines):
            assert a.serial == int(line[6:11])
            assert a.name == line[12:16].strip()
            assert a.residue_name == line[17:21].strip()
            assert a.chain_id == line[21:23].strip()
            assert a.residue_number == int(line[22:26])
            assert a.x == pytest.approx(float(line[30:38]), abs=5.01e-4)
            assert a.y == pytest.approx(float(line[38:46]), abs=5.01e-4)
            assert a.z == pytest.approx(float(line[46:54]), abs=5.01e-4)
            assert a.occupancy == pytest.approx(float(line[54:60]), abs=1e-3)
            assert a.temperature_factor == pytest.approx(float(line[60:66]), abs=1e-3)
    want = oracle.hits(mol_1amy, [tpl_133], 2, 4, 4, reorder=False)
    assert_same_hits(mol_1amy, [tpl_133], hits, want)


def test_fixture_1amy_133_reorder(mol_1amy, tpl_133):
    hits = list(Jess([tpl_133]).query(mol_1amy, 2, 4, 4, reorder=True))
    want = oracle.hits(mol_1amy, [tpl_133], 2, 4, 4, reorder=True)
    assert_same_hits(mol_1amy, [tpl_133], hits, want)
    blocks = parse_hit_file(golden("1AMY+1.3.3.txt"))
    assert {key(mol_1amy, h) for h in hits} == {tuple(int(l[6:11]) for l in lines) for _, lines in blocks}


def test_doctests(mol_1amy, tpl_133, tpl_412):
    jess = Jess([tpl_133, tpl_412])
    hits = list(jess.query(mol_1amy, 2, 2, 2))
    assert [h.template().id for h in hits] == ["2om2"] * 13
    assert all(repr(h.rmsd).startswith(p) for h, p in zip(hits, DOCTEST_RMSD))
    best = list(jess.query(mol_1amy, 2, 2, 2, best_match=True))
    assert len(best) == 1 and repr(best[0].rmsd).startswith("1.071")
    assert next(jess.query(mol_1amy, 2, 2, 2)).dumps(format="pdb") == DOCTEST_DUMP


@pytest.mark.parametrize("ignore_chain,count", [(None, 0), ("residues", 2), ("atoms", 7)])
def test_ignore_chain(mol_5ayx, tpl_412, ignore_chain, count):
    hits = list(Jess([tpl_412]).query(mol_5ayx, 3, 3, 3, ignore_chain=ignore_chain))
    assert len(hits) == count
    assert_same_hits(mol_5ayx, [tpl_412], hits,
                     oracle.hits(mol_5ayx, [tpl_412], 3, 3, 3, ignore_chain=ignore_chain))


@pytest.mark.parametrize("best_match", [False, True])
@pytest.mark.parametrize("max_candidates", [None, 0, 1, 3, 10])
def test_policies(mol_1amy, tpl_133, tpl_412, best_match, max_candidates):
    tpls = [tpl_133, tpl_412, tpl_133]
    got = list(Jess(tpls).query(mol_1amy, 2, 4, 4, best_match=best_match, max_candidates=max_candidates))
    want = oracle.hits(mol_1amy, tpls, 2, 4, 4, best_match=best_match, max_candidates=max_candidates)
    assert_same_hits(mol_1amy, tpls, got, want)


def test_max_candidates_negative(mol_1amy, tpl_133):
    with pytest.raises(ValueError):
        list(Jess([tpl_133]).query(mol_1amy, 2, 5, 5, max_candidates=-1))


@pytest.mark.parametrize("name", ["plain-2k", "plain-8k"])
@pytest.mark.parametrize("reorder", [True, False])
def test_synthetic_workloads(name, reorder):
    mols, tpls, kw = synthetic.workload(name, seed=11)
    mol = mols[0]
    got = list(Jess(tpls).query(mol, kw["rmsd_threshold"], kw["distance_cutoff"],
                                kw["max_dynamic_distance"], max_candidates=kw["max_candidates"],
                                reorder=reorder))
    want = oracle.hits(mol, tpls, kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"],
                       max_candidates=kw["max_candidates"], reorder=reorder)
    assert len(want) > 0
    assert_same_hits(mol, tpls, got, want)


@pytest.mark.parametrize("ignore_chain,best_match,max_candidates",
                         [(None, False, 100000), ("residues", False, 3), (None, True, 100000), ("atoms", True, 2)])
def test_library_style(ignore_chain, best_match, max_candidates):
    # the library-scale shape at a size the oracle replays in seconds: 20
    # chains, templates of 2-8 residues with symmetric mode-1 atom pairs,
    # 2.5 A cutoff
    rng = np.random.Generator(np.random.PCG64(21))
    mol = synthetic.protein_with_atoms(rng, 12000, n_chains=20, mol_id="LIBT")
    tpls = synthetic.template_set(rng, [mol], 150, 0.2, (2, 8), (2, 3), p_symmetric=0.2)
    assert any(a.match_mode == 1 for t in tpls for a in t)
    got = list(Jess(tpls).query(mol, 2.0, 2.5, 2.5, max_candidates=max_candidates, ignore_chain=ignore_chain,
                                best_match=best_match))
    want = oracle.hits(mol, tpls, 2.0, 2.5, 2.5, max_candidates=max_candidates, ignore_chain=ignore_chain,
                       best_match=best_match)
    assert len(want) > 50
    assert_same_hits(mol, tpls, got, want)


@pytest.mark.parametrize("ignore_chain", [None, "residues", "atoms"])
def test_multichain_flexible(ignore_chain):
    rng = np.random.Generator(np.random.PCG64(5))
    mol = synthetic.protein(rng, 600, n_chains=4)
    tpls = synthetic.template_set(rng, [mol], 60, 0.5, (2, 5), (2, 3), jitter=0.4, p_flexible=0.6)
    got = list(Jess(tpls).query(mol, 2.0, 2.0, 2.5, ignore_chain=ignore_chain))
    want = oracle.hits(mol, tpls, 2.0, 2.0, 2.5, ignore_chain=ignore_chain)
    assert len(want) > 0
    assert_same_hits(mol, tpls, got, want)


@pytest.mark.parametrize("max_candidates,best_match", [(None, False), (3, False), (1, False), (3, True)])
def test_query_batch_equals_single_queries(max_candidates, best_match):
    """query_batch (lazy lists; a candidate cap without best_match applied to
    the pairs it can cut only) equals the oracle and single queries, also
    with caps small enough to cut pairs."""
    mols, tpls, kw = synthetic.workload("batch-4k", seed=3, scale=6 / 4096)
    args = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    mc = kw["max_candidates"] if max_candidates is None else max_candidates
    jess = Jess(tpls)
    batch = jess.query_batch(mols, *args, max_candidates=mc, best_match=best_match)
    assert len(batch) == len(mols)
    if max_candidates is not None:
        # some pair has more complete matches than the cap lets through
        assert int(jess._last_result.pair_matches.max()) > mc + 1
    total = 0
    for mol, got in zip(mols, batch):
        want = oracle.hits(mol, tpls, *args, max_candidates=mc, best_match=best_match)
        assert_same_hits(mol, tpls, got, want)
        single = list(jess.query(mol, *args, max_candidates=mc, best_match=best_match))
        assert [key(mol, h) for h in single] == [key(mol, h) for h in got]
        total += len(want)
    assert total > 0


def test_query_batch_hits_survive_later_queries():
    """query_batch leaves the hits' rotations, centroids and atoms in a device
    copy of the events until the first of them is read; later queries, which
    overwrite the context's own event buffer, leave that copy alone."""
    mols, tpls, kw = synthetic.workload("batch-4k", seed=5, scale=4 / 4096)
    args = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    jess = Jess(tpls)
    first = jess.query_batch(mols, *args)
    second = jess.query_batch(mols[::-1], *args, best_match=True)
    single = [list(jess.query(mol, *args)) for mol in mols]
    assert sum(len(h) for h in first) > 0
    for mol, got, want in zip(mols, first, single):
        assert len(got) == len(want)
        for g, w in zip(got, want):
            assert g.template() is w.template() and key(mol, g) == key(mol, w)
            assert same_bits(g.rmsd, w.rmsd)
            assert same_bits(np.ravel(g._rotation), np.ravel(w._rotation))
            assert same_bits(np.ravel(g._centre), np.ravel(w._centre))
            assert np.array_equal(np.asarray(g.transformation), np.asarray(w.transformation))
    for mol, got in zip(mols[::-1], second):
        want = list(jess.query(mol, *args, best_match=True))
        assert [key(mol, h) for h in got] == [key(mol, h) for h in want]


def test_superpose_batch_bit_exact():
    rng = np.random.default_rng(0)
    for k in (3, 4, 6, 12, 24):
        n = 500
        a = rng.normal(scale=8.0, size=(n, k, 3))
        b = a + rng.normal(scale=0.7, size=(n, k, 3))
        b[: n // 2] = b[: n // 2] @ np.diag([1.0, -1.0, 1.0])   # reflections
        rmsd, rot, c0, c1 = _native.superpose_batch(a, b)
        for i in range(0, n, 7):
            r, M, x0, x1 = oracle.superpose(a[i], b[i])
            assert rmsd[i] == r
            assert np.array_equal(rot[i], M, equal_nan=True)
            assert np.array_equal(c0[i], x0) and np.array_equal(c1[i], x1)


def test_empty_inputs(mol_1amy, tpl_133):
    assert list(Jess().query(mol_1amy, 2, 2, 2)) == []
    assert list(Jess([tpl_133]).query(Molecule(), 2, 2, 2)) == []
    tiny = mol_1amy[:5]                    # no THR/ASN/SER: an empty candidate set
    assert list(Jess([tpl_133]).query(tiny, 2, 4, 4)) == []
    assert Jess([tpl_133]).query_batch([], 2, 4, 4) == []


def test_single_atom_template_and_nan_rotation(mol_1amy):
    """A 1-atom template: the first candidate is stepped over and every other
    match has a singular superposition, reported as a NaN warning (as pyjess)."""
    t = Template([TemplateAtom(chain_id="A", residue_number=1, x=0.0, y=0.0, z=0.0, residue_names=["CYS"],
                               atom_names=["SG"], match_mode=0)], id="one")
    stream = oracle.query_stream(mol_1amy, [t], 2, 2)
    n_cys = sum(1 for i in range(len(mol_1amy)) if bytes(mol_1amy.names[i]) == b"_SG_")
    assert len(stream) == n_cys - 1
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        hits = list(Jess([t]).query(mol_1amy, 2, 2, 2))
    assert hits == []
    assert len(w) == oracle.replay(stream, 2)[1] == n_cys - 1


def test_unsupported_mode_raises(mol_1amy):
    t = Template([TemplateAtom(chain_id="A", residue_number=1, x=0.0, y=0.0, z=0.0, residue_names=["CYS"],
                               atom_names=["SG"], match_mode=2)], id="m2")
    with pytest.raises(NotImplementedError):
        list(Jess([t]).query(mol_1amy, 2, 2, 2))


def test_repeat_queries_are_deterministic(mol_1amy, tpl_133, tpl_412):
    jess = Jess([tpl_133, tpl_412])
    first = [(h.rmsd, key(mol_1amy, h)) for h in jess.query(mol_1amy, 2, 4, 4)]
    for _ in range(3):
        assert [(h.rmsd, key(mol_1amy, h)) for h in jess.query(mol_1amy, 2, 4, 4)] == first


def trp_template(mol, rng, tid):
    """Mode-3 carbons of a TRP (a set of every TRP carbon: > 1024 entries on a
    big protein) plus two exact atoms of a residue nearby."""
    starts = synthetic._residue_starts(mol)
    trp = [s for s in starts if bytes(mol.resnames[s][:3]) == b"TRP"]
    s0 = trp[int(rng.integers(len(trp)))]
    res = synthetic._residue_atoms(mol, s0)
    ca = mol.xyz[s0 + 1]
    near = [s for s in starts if s != s0 and np.linalg.norm(mol.xyz[s + 1] - ca) < 9.0]
    s1 = near[int(rng.integers(len(near)))]
    atoms = []
    for a, mode in ((res[1], 3), (res[4], 3), (res[5], 3)):
        nm = bytes(mol.names[a]).decode().strip("_")
        atoms.append(TemplateAtom(chain_id="A", residue_number=int(mol.resseq[a]), x=float(mol.xyz[a, 0]),
                                  y=float(mol.xyz[a, 1]), z=float(mol.xyz[a, 2]), residue_names=["TRP"],
                                  atom_names=[nm], match_mode=mode))
    for a in synthetic._residue_atoms(mol, s1)[:2]:
        nm = bytes(mol.names[a]).decode().strip("_")
        rn = bytes(mol.resnames[a][:3]).decode()
        atoms.append(TemplateAtom(chain_id="A", residue_number=int(mol.resseq[a]), x=float(mol.xyz[a, 0]),
                                  y=float(mol.xyz[a, 1]), z=float(mol.xyz[a, 2]), residue_names=[rn],
                                  atom_names=[nm], match_mode=0))
    return Template(atoms, id=tid)


@pytest.mark.parametrize("decimals", [3, 0])
@pytest.mark.parametrize("reorder", [False, True])
def test_kd_order_large_sets_and_ties(decimals, reorder):
    """Sets larger than the LDS capacity, and coordinates with many equal values
    (decimals=0), still enumerate in the reference order."""
    rng = np.random.Generator(np.random.PCG64(21))
    mol = synthetic.protein(rng, 2600)
    if decimals == 0:
        mol.xyz = np.round(mol.xyz)
    tpls = [trp_template(mol, rng, f"W{i}") for i in range(6)]
    tpls += synthetic.template_set(rng, [mol], 20, 0.5, (2, 4), (2, 3), jitter=0.3, p_flexible=0.8)
    got = list(Jess(tpls).query(mol, 2.0, 1.5, 1.5, reorder=reorder))
    want = oracle.hits(mol, tpls, 2.0, 1.5, 1.5, reorder=reorder)
    assert len(want) > 0
    assert_same_hits(mol, tpls, got, want)


@pytest.mark.parametrize("reorder", [True, False])
def test_kd_ties_in_batched_queries(reorder):
    """Several molecules with integer coordinates (candidate sets of 257 to
    1024 entries full of ties) in one batched query: the tie handling of the
    batched kd passes, in order, against the oracle molecule by molecule."""
    rng = np.random.Generator(np.random.PCG64(33))
    mols = []
    for n in (2600, 3100, 1800):
        m = synthetic.protein(rng, n)
        m.xyz = np.round(m.xyz)
        mols.append(m)
    tpls = [trp_template(mols[0], rng, f"W{i}") for i in range(4)]
    tpls += [trp_template(mols[1], rng, f"V{i}") for i in range(4)]
    tpls += synthetic.template_set(rng, mols, 24, 0.5, (2, 4), (2, 3), jitter=0.3, p_flexible=0.8)
    print(f"batched ties: {[len(m) for m in mols]} atoms x {len(tpls)} templates, reorder={reorder}", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        batch = Jess(tpls).query_batch(mols, 2.0, 1.5, 1.5, reorder=reorder)
    total = 0
    for mol, got in zip(mols, batch):
        want = oracle.hits(mol, tpls, 2.0, 1.5, 1.5, reorder=reorder)
        assert_same_hits(mol, tpls, got, want)
        total += len(want)
    assert total > 0
