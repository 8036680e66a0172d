"""PDB ingestion on the device (jam_molecules_parse_pdb, `Molecule.load(...,
device=True)`, `Molecule.load_many`) against the host parser, which follows
the reference's PDBMoleculeParser and Atom.loads (src/pyjess/_io.py:68-103):
every atom field, the header, the residue-name index (the host's is the
stable sort of jam_molecules_create, patches/ResIndex.c.patch:34-119) and the
hits of queries."""

import io
import warnings

import numpy as np
import pytest

from pyjess_amd import Jess, Molecule, Template, _native, synthetic

from conftest import golden

pytestmark = pytest.mark.gpu

FIELDS = ("xyz", "names", "resnames", "chains", "resseq", "serial", "altloc", "icode", "occupancy", "tempfactor",
          "segid", "element", "charge")


def assert_same_molecule(a, b):
    assert len(a) == len(b)
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert x.dtype == y.dtype, f
        if x.dtype.kind == "f":     # bit for bit: the signs of zeros too
            x, y = x.view(np.dtype(f"u{x.itemsize}")), y.view(np.dtype(f"u{y.itemsize}"))
        assert np.array_equal(x, y), f
    assert (a.id, a.date, a.name) == (b.id, b.date, b.name)


def assert_same_index(dev_mol, host_mol):
    ctx = _native.context()
    got = dev_mol._device[id(ctx)].residue_index()
    want = _native.MoleculeSet(ctx, [host_mol]).residue_index()
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def hits_of(mol, tpls, *args, **kw):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        return [(h.template().id, round(h.rmsd, 12), tuple(a.serial for a in h.atoms(transform=False)))
                for h in Jess(tpls).query(mol, *args, **kw)]


def test_1amy_device_parse_equals_host_parse():
    host = Molecule.load(golden("1AMY.pdb"), device=False)
    dev = Molecule.load(golden("1AMY.pdb"), device=True)
    print(f"1AMY: {len(host)} atoms", flush=True)
    assert_same_molecule(dev, host)
    assert_same_index(dev, host)
    tpls = [Template.load(golden("1.3.3.tpl")), Template.load(golden("4.1.2.tpl"))]
    assert hits_of(dev, tpls, 2, 4, 4) == hits_of(host, tpls, 2, 4, 4)


def test_5ayx_and_flags():
    for skip in (False, True):
        host = Molecule.load(golden("5ayx.EF.pdb"), device=False, skip_hetatm=skip)
        dev = Molecule.load(golden("5ayx.EF.pdb"), device=True, skip_hetatm=skip)
        assert_same_molecule(dev, host)
        assert_same_index(dev, host)


def test_library_assembly_through_pdb_text():
    """A 100,000-atom, 20-chain assembly written as PDB text and read back by
    both parsers; the same hits against library templates. Serial numbers
    wrap at 99,999 as five columns force them to, so the residue-name index
    has buckets whose serials are out of order."""
    mols, tpls, kw = synthetic.workload("library-64", seed=0, scale=1 / 64)
    m = mols[0]
    wrapped = np.arange(len(m), dtype=np.int32) % 99999 + 1
    m = Molecule.from_arrays(m.xyz, m.names, m.resnames, m.chains, m.resseq, wrapped, id=m.id)
    text = m.dumps()
    print(f"assembly: {len(m)} atoms, {len(text)} bytes of PDB text", flush=True)
    host = Molecule.load(io.StringIO(text), device=False)
    dev = Molecule.load(io.StringIO(text), device=True)
    assert len(dev) >= 100_000
    assert_same_molecule(dev, host)
    assert_same_index(dev, host)
    sub = tpls[::50]
    q = (kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"])
    got = hits_of(dev, sub, *q)
    assert len(got) > 0 and got == hits_of(host, sub, *q)


EDGE = (
    "HEADER    HYDROLASE                               01-JAN-99   1ABC              \n"
    "ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N  \n"
    "ATOM      2  CA  ALA A   1      11.639   6.071  -5.147  1.00  0.00           C  \r\n"
    "ATOM      3  C   ALA A   1      12.000   6.000  -5.000\r"
    "HETATM    4 ZN    ZN B   2      -1.5e1   +6.25   0.000  0.50 12.50          ZN2+\n"
    "ATOM      9  O   GLY A   3       1.000   2.000   3.000  1.00 20.00           O1-\n"
    "ATOM      5  OXT GLY A   3       1.100   2.100   3.100  1.00 20.00      SEG1 O+2\n"
    "ATOM      7  CB  gly A   3       0.100  -0.000   1e-3        30.00           C12\n"
    "TER       8      GLY A   3\n"
    "ENDMDL\n"
    "ATOM     10  N   SER A   4       5.000   5.000   5.000  1.00  0.00           N  \n"
)


@pytest.mark.parametrize("skip_hetatm", [False, True])
@pytest.mark.parametrize("ignore_endmdl", [False, True])
def test_edge_records(tmp_path, skip_hetatm, ignore_endmdl):
    """CRLF and lone-CR line ends, short lines, exponents and signs, blank
    occupancy, charges read backwards, a lower-case residue name, serial
    numbers out of order (the index sorts them), HETATM and ENDMDL flags.
    The host reads the file in text mode, where a lone CR ends a line too."""
    path = tmp_path / "edge.pdb"
    path.write_bytes(EDGE.encode())
    kw = dict(skip_hetatm=skip_hetatm, ignore_endmdl=ignore_endmdl)
    host = Molecule.load(path, device=False, **kw)
    dev = Molecule.load(path, device=True, **kw)
    assert len(host) == 7 - skip_hetatm + ignore_endmdl
    assert_same_molecule(dev, host)
    assert_same_index(dev, host)


@pytest.mark.parametrize("text", [
    "data_1ABC\nATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N\n",
    "ATOM      1  N   ALA A   1      11.104   6.1x4  -6.504  1.00  0.00           N\n",
    "ATOM     1a  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N\n",
    "ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.0z           N\n",
])
def test_bad_texts_raise(text):
    with pytest.raises(ValueError):
        Molecule.load(io.StringIO(text), device=False)
    with pytest.raises(ValueError):
        Molecule.load(io.StringIO(text), device=True)


def test_load_many_batch_equals_single_loads_and_queries():
    rng = np.random.Generator(np.random.PCG64(11))
    mols = [synthetic.protein(rng, n) for n in (700, 2500, 1300)]
    texts = [m.dumps() for m in mols]
    # serial numbers shuffled in one of them: its buckets need the sort by serial
    lines = texts[1].splitlines(keepends=True)
    atom_lines = [ln for ln in lines if ln.startswith("ATOM")]
    perm = rng.permutation(len(atom_lines))
    texts[1] = "".join(ln[:6] + f"{int(p) + 1:5d}" + ln[11:] for ln, p in zip(atom_lines, perm))
    batch = Molecule.load_many([io.StringIO(t) for t in texts])
    host = [Molecule.load(io.StringIO(t), device=False) for t in texts]
    for d, h in zip(batch, host):
        assert_same_molecule(d, h)
    # the batch's residue index equals the host's, molecule by molecule
    ctx = _native.context()
    from pyjess_amd import _jess

    bset = _jess._last_batch[3]
    perm_b, off_b, key_b, beg_b, end_b = bset.residue_index()
    want = _native.MoleculeSet(ctx, host).residue_index()
    for g, w in zip((perm_b, off_b, key_b, beg_b, end_b), want):
        assert np.array_equal(g, w)
    tpls = synthetic.template_set(rng, mols, 30, 0.5, (2, 4), (2, 3), jitter=0.3)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        got = Jess(tpls).query_batch(batch, 2.0, 1.5, 1.5)
        want_hits = Jess(tpls).query_batch(host, 2.0, 1.5, 1.5)
    key = lambda hs: [(h.template().id, h.rmsd, tuple(a.serial for a in h.atoms(transform=False))) for h in hs]
    assert [key(h) for h in got] == [key(h) for h in want_hits]
    assert sum(len(h) for h in got) > 0


def test_empty_and_atomless_texts(tmp_path):
    path = tmp_path / "edge.pdb"
    path.write_bytes(EDGE.encode())
    batch = Molecule.load_many([io.StringIO(""), io.StringIO("REMARK nothing here\n"), path])
    assert [len(m) for m in batch[:2]] == [0, 0]
    assert_same_molecule(batch[2], Molecule.load(path, device=False))
