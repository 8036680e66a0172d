"""Parsing and formats of the data model, after the reference's test_atom.py,
test_template_atom.py and test_template.py (src/pyjess/tests/)."""

import io
import pickle
import textwrap

import numpy as np
import pytest

from pyjess_amd import Atom, Mat4, Molecule, Template, TemplateAtom

from conftest import golden

LINE = "ATOM     39  CA  PRO A 469     -14.948   2.091  10.228  1.00 27.71           C"


def test_atom_load():
    a = Atom.loads(LINE)
    assert (a.serial, a.name, a.residue_name, a.chain_id, a.residue_number) == (39, "CA", "PRO", "A", 469)
    assert a.x == pytest.approx(-14.948) and a.y == pytest.approx(2.091) and a.z == pytest.approx(10.228)
    assert a.occupancy == pytest.approx(1.00) and a.temperature_factor == pytest.approx(27.71, abs=1e-5)
    assert (a.segment, a.element, a.charge) == ("", "C", 0)


def test_atom_init_consistency():
    loaded = Atom.loads(LINE)
    created = Atom(serial=loaded.serial, name=loaded.name, residue_name=loaded.residue_name,
                   chain_id=loaded.chain_id, altloc=loaded.altloc, insertion_code=loaded.insertion_code,
                   residue_number=loaded.residue_number, x=loaded.x, y=loaded.y, z=loaded.z,
                   occupancy=loaded.occupancy, temperature_factor=loaded.temperature_factor,
                   element=loaded.element)
    assert loaded == created
    assert hash(loaded) == hash(created)


def test_atom_invalid():
    kw = dict(serial=1, name="N", residue_name="GLN", chain_id="A", residue_number=3, x=0, y=0, z=0)
    for key, bad in (("chain_id", "too long"), ("name", "too long"), ("residue_name", "too long")):
        with pytest.raises(ValueError):
            Atom(**{**kw, key: bad})


def test_atom_dumps_roundtrip_and_pickle():
    a = Atom.loads(LINE)
    assert a.dumps().strip() == LINE.strip()
    assert pickle.loads(pickle.dumps(a)) == a


def test_template_atom_load():
    atom = TemplateAtom.loads("ATOM      1  NE  ARG A1136       3.953   0.597  -1.721 K")
    assert atom.match_mode == 1
    assert atom.atom_names == ("NE",)
    assert atom.residue_names == ("ARG", "LYS")
    assert atom.chain_id == "A"
    assert atom.residue_number == 1136
    assert (atom.x, atom.y, atom.z) == (3.953, 0.597, -1.721)


def test_template_atom_dumps():
    atom = TemplateAtom(chain_id="A", residue_number=1136, x=3.953, y=0.597, z=-1.721,
                        residue_names=["ARG", "LYS"], atom_names=["NE"], distance_weight=0.00,
                        match_mode=1)
    assert atom.dumps() == "ATOM      1  NE  ARG A1136       3.953   0.597  -1.721 K     0.00"
    copy = TemplateAtom.loads(atom.dumps())
    assert copy == atom


def test_template_atom_invalid():
    kw = dict(atom_names=["NE"], residue_names=["ARG"], chain_id="A", residue_number=1, x=0.0, y=0.0,
              z=0.0, match_mode=0)
    with pytest.raises(ValueError):
        TemplateAtom(**{**kw, "match_mode": 2000})
    with pytest.raises(ValueError):
        TemplateAtom(**{**kw, "match_mode": -10})
    with pytest.raises(ValueError):
        TemplateAtom(**{**kw, "residue_names": ["something"]})
    a = TemplateAtom(**kw)
    assert eval(repr(a)) == a
    assert pickle.loads(pickle.dumps(a)) == a


TEMPLATE = textwrap.dedent(
    """
    ATOM      0  NZ  LYS A1030       1.431  -9.717  -1.722
    ATOM      0  CG  ASP A1132       0.000   0.000   0.000
    ATOM      0  OD1 ASP A1132       0.595   1.085   0.000
    ATOM      0  OD2 ASP A1132       0.609  -1.111   0.000
    ATOM      1  NE  ARG A1136       3.953   0.597  -1.721 K
    ATOM      0  CG  ASN A1137       0.736  -1.869  -4.261
    ATOM      0  OD1 ASN A1137       1.760  -2.435  -4.344
    ATOM      0  ND2 ASN A1137       0.264  -1.477  -3.134
    ATOM      0  CG  ASP A1150      -0.010  -5.408  -2.070
    ATOM      0  OD1 ASP A1150      -0.587  -5.769  -1.153
    ATOM      0  OD2 ASP A1150       1.085  -5.212  -2.011
    """
).strip()


def test_template_loads():
    t = Template.loads(TEMPLATE)
    assert len(t) == 11 and t.dimension == 5
    assert t[0].residue_names == ("LYS",) and t[0].atom_names == ("NZ",)
    assert t[1].atom_names == ("CG",)
    assert t[2].residue_number == 1132 and t[-1].residue_number == 1150


def test_template_dumps_roundtrip():
    t1 = Template.loads(TEMPLATE, id="tpl1")
    t2 = Template.loads(t1.dumps())
    assert len(t1) == len(t2) and t1 == t2


def test_template_dimension_multiple_chains():
    assert Template.load(golden("1sur.qry")).dimension == 4


def test_fixture_templates():
    t = Template.load(golden("1.3.3.tpl"))
    assert t.id == "3r6v" and len(t) == 9 and t.dimension == 3
    assert [a.match_mode for a in t] == [3, 3, 1, 3, 1, 1, 3, 3, 1]
    assert t[0].residue_names == ("THR", "THR")


def test_molecule_load(mol_1amy):
    assert len(mol_1amy) == 3339 and mol_1amy.id == "1AMY"
    a = mol_1amy[0]
    assert (a.serial, a.name, a.residue_name) == (1, "N", "GLN")
    sub = mol_1amy[:10]
    assert len(sub) == 10 and sub[3] == mol_1amy[3]


def test_molecule_conserved(mol_1amy):
    c = mol_1amy.conserved(10.0)
    assert 0 < len(c) < len(mol_1amy)
    assert all(a.temperature_factor >= 10.0 for a in c)


def test_mat4_matmul():
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=(4, 4)), rng.normal(size=(4, 4))
    assert np.allclose(np.asarray(Mat4(a) @ b), a @ b, rtol=0, atol=1e-12)


def test_molecule_arrays_are_read_only():
    """A molecule does not change once built (as the reference's cannot), so
    its device copies stay valid: in-place edits raise."""
    mol = Molecule.load(golden("1AMY.pdb"))
    for attr in ("xyz", "names", "resnames", "chains", "resseq", "serial"):
        arr = getattr(mol, attr)
        assert not arr.flags.writeable
        with pytest.raises(ValueError):
            arr[0] = arr[1]


def test_molecule_from_arrays_copies_caller_arrays():
    xyz = np.zeros((2, 3))
    names = np.frombuffer(b"_CA__CB_", dtype=np.uint8).reshape(2, 4).copy()
    resnames = np.frombuffer(b"ALA\0ALA\0", dtype=np.uint8).reshape(2, 4).copy()
    chains = np.array([[32, 65], [32, 65]], dtype=np.uint8)
    mol = Molecule.from_arrays(xyz, names, resnames, chains, np.array([1, 1], dtype=np.int32))
    xyz[0, 0] = 5.0
    assert mol.xyz[0, 0] == 0.0


def test_replacing_an_array_drops_device_copies():
    mol = Molecule.load(golden("1AMY.pdb"))
    mol._device[123] = object()
    v0 = mol.__dict__.get("_version", 0)
    mol.xyz = np.round(mol.xyz)
    assert mol._device == {}
    assert mol.__dict__["_version"] == v0 + 1
    assert not mol.xyz.flags.writeable


def _dumps_atom_by_atom(mol):
    return "".join(atom.dumps() for atom in mol)


def test_molecule_dumps_equals_atom_dumps(mol_1amy):
    """Molecule.dumps writes every atom's columns at once; its text is the
    join of Atom.dumps, rows with numbers too wide for their columns, NaN
    or rounding ties included (those are formatted one by one)."""
    from pyjess_amd import synthetic

    assert mol_1amy.dumps(write_header=False) == _dumps_atom_by_atom(mol_1amy)
    rng = np.random.Generator(np.random.PCG64(5))
    m = synthetic.protein(rng, 120)
    n = len(m)
    xyz = np.round(rng.uniform(-999, 9999, (n, 3)), 3)
    xyz[:20, 0] = [0.0, -0.0, -0.0004, -0.0005, 0.0005, 0.0015, 1.0005, 2.0625, -0.0625, 9999.9994,
                   9999.9996, -999.9994, -999.9996, np.nan, np.inf, -np.inf, 1e20, 123.4565, 5e-324, 12345.5]
    serial = np.arange(n, dtype=np.int32) * 997 - 50
    resseq = m.resseq.copy()
    resseq[:4] = [-999, -1000, 9999, 10000]
    occ = rng.uniform(0, 200, n).astype(np.float32)
    occ[:3] = [0.125, np.nan, 1000.0]
    m2 = Molecule.from_arrays(xyz, m.names, m.resnames, m.chains, resseq, serial, id="ODD1",
                              occupancy=occ, tempfactor=rng.uniform(-99, 999, n))
    text = m2.dumps()
    assert text.splitlines(keepends=True)[0].startswith("HEADER")
    assert text.split("\n", 1)[1] == _dumps_atom_by_atom(m2)
    assert Molecule().dumps() == Molecule().dumps(write_header=True)
    assert Molecule().dumps(write_header=False) == ""


@pytest.mark.parametrize("ignore_endmdl", [False, True])
def test_header_fields_from_bytes_equal_the_text_parse(tmp_path, ignore_endmdl):
    """The device loader takes id, date and name from the bytes of a text
    (no full decode); they equal what the line-by-line parser reads, line
    ends of every kind, a HEADER after ENDMDL and tags inside lines included."""
    from pyjess_amd._model import _pdb_header

    atom = "ATOM      1  N   ALA A   1      11.104   6.134  -6.504  1.00  0.00           N  "
    h1 = "HEADER    HYDROLASE                               01-JAN-99   1ABC              "
    h2 = "HEADER    OTHER NAME                              02-FEB-03   2XYZ"
    texts = [
        h1 + "\n" + atom + "\n",
        atom + "\r" + h1 + "\r\n" + atom + "\r",
        atom + "\nENDMDL\n" + h1 + "\n",
        "REMARK HEADER not a record\n" + atom + "\n" + "xHEADER\n",
        "",
        h2,
        "HEADER\n" + h1 + "\n",
    ]
    for k, text in enumerate(texts):
        path = tmp_path / f"t{k}.pdb"
        path.write_bytes(text.encode())
        host = Molecule.load(path, device=False, ignore_endmdl=ignore_endmdl)
        got = _pdb_header(text.encode(), None, None, None, ignore_endmdl)
        assert got == (host.id, host.date, host.name), text
    assert _pdb_header(h1.encode(), "X", None, "N", False)[::2] == ("X", "N")
    with pytest.raises(UnicodeDecodeError):
        _pdb_header(b"REMARK \xff\n", None, None, None, False)
