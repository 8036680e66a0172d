"""Data model of the drop-in API: `Atom`, `Molecule`, `TemplateAtom`, `Template`, `Mat4`.

Mirrors the classes of ``pyjess._jess`` (reference ``src/pyjess/_jess.pyx``)
with the same constructor arguments, properties, parsing and text formats.
Molecules keep their atoms as structure-of-arrays (numpy), the layout the
device path uploads to HBM; `Atom` objects are created on access.

Encodings follow the reference:

* atom names are 4 bytes, blanks and NUL replaced by ``_``, names shorter than
  4 characters get a leading ``_`` (``encode_token``, ``_jess.pyx:133-140``,
  ``Atom.__init__`` ``_jess.pyx:1052-1056``);
* residue names are 3 bytes + NUL (``Atom.h`` ``resName[4]``);
* a chain identifier of one character ``c`` is stored as ``(' ', c)``, an
  empty one as ``(' ', '0')`` (``_jess.pyx:1041-1050``).
"""

from __future__ import annotations

import datetime
import io
import math
import os
import re
import typing
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

__all__ = ["Atom", "Molecule", "TemplateAtom", "Template", "Mat4"]

# one-letter -> three-letter residue codes of template lines
# (TessAtom_convertResidue, patches/TessAtom.c.patch:22-56)
_ONE_TO_THREE = {
    "A": "ALA", "C": "CYS", "D": "ASP", "E": "GLU", "F": "PHE", "G": "GLY",
    "H": "HIS", "I": "ILE", "K": "LYS", "L": "LEU", "M": "MET", "N": "ASN",
    "P": "PRO", "O": "PYL", "Q": "GLN", "R": "ARG", "S": "SER", "T": "THR",
    "U": "SEC", "V": "VAL", "W": "TRP", "Y": "TYR", "X": "XXX",
}
# three-letter -> one-letter, used when writing template lines
# (AA3/AA1 of _jess.pyx:142-163, unknown names become 'X')
_THREE_TO_ONE = {
    "ALA": "A", "CYS": "C", "ASP": "D", "GLU": "E", "PHE": "F", "GLY": "G",
    "HIS": "H", "ILE": "I", "LYS": "K", "LEU": "L", "MET": "M", "ASN": "N",
    "PRO": "P", "GLN": "Q", "ARG": "R", "SER": "S", "THR": "T", "VAL": "V",
    "TRP": "W", "TYR": "Y", "XXX": "X",
}

_VALID_MODES = set(range(-1, 9)) | set(range(100, 108))


def _encode(raw: bytes, n: int) -> bytes:
    """``encode_token``: pad with NUL to ``n`` bytes, blanks/NUL become ``_``."""
    raw = raw[:n].ljust(n, b"\0")
    return bytes(0x5F if c in (0x20, 0) else c for c in raw)


def _encode_atom_name(name: str) -> bytes:
    raw = bytearray(name, "ascii")
    if len(raw) > 4:
        raise ValueError(f"Invalid atom name: {name!r}")
    if len(raw) < 4:
        raw.insert(0, ord("_"))
    return _encode(bytes(raw), 4)


def _encode_resname(name) -> bytes:
    raw = name.encode("ascii") if isinstance(name, str) else bytes(name)
    if len(raw) > 3:
        raise ValueError(f"Invalid residue name: {name!r}")
    return _encode(raw, 3) + b"\0"


def _chain_bytes(chain_id: str) -> Tuple[int, int]:
    if len(chain_id) > 2:
        raise ValueError(f"Invalid chain ID: {chain_id!r}")
    if len(chain_id) == 2:
        return ord(chain_id[0]), ord(chain_id[1])
    if len(chain_id) == 1:
        return ord(" "), ord(chain_id[0])
    return ord(" "), ord("0")


def _decode(raw: bytes) -> str:
    """``decode_token``: ``_`` and NUL become blanks."""
    return "".join(" " if c in (0x5F, 0) else chr(c) for c in raw)


# texts of at least this many bytes are parsed on the device by Molecule.load
DEVICE_PARSE_BYTES = 1 << 20


def _read_bytes(file) -> bytes:
    """The bytes of a path or of a file object (text is encoded as UTF-8)."""
    if isinstance(file, (str, os.PathLike)):
        with open(file, "rb") as f:
            return f.read()
    data = file.read()
    return data.encode("utf-8") if isinstance(data, str) else bytes(data)


def _record_starts(data: bytes, tag: bytes, end: int):
    """Offsets of the lines of ``data[:end]`` that begin with ``tag`` (lines
    end at LF, CR LF or a lone CR, as a text-mode read splits them)."""
    i = data.find(tag, 0, end)
    while i >= 0:
        if i == 0 or data[i - 1] in (10, 13):
            yield i
        i = data.find(tag, i + 1, end)


def _pdb_header(data: bytes, id_, date, name, ignore_endmdl: bool):
    """id, date and name from the HEADER records read before the parse stops
    (the HEADER branch of ``PDBMoleculeParser.load``, ``src/pyjess/_io.py:81-92``),
    found in the bytes of a PDB text without decoding all of it; a text that
    is not valid UTF-8 raises `UnicodeDecodeError` as a text-mode read does."""
    if not data.isascii():
        data.decode("utf-8")
    end = len(data)
    if not ignore_endmdl:
        end = next(_record_starts(data, b"ENDMDL", end), end)
    for i in _record_starts(data, b"HEADER", end):
        j = len(data)
        for sep in (b"\n", b"\r"):
            k = data.find(sep, i)
            if 0 <= k < j:
                j = k
        line = data[i:j].decode("utf-8")
        if id_ is None:
            id_ = line[62:66].strip() or None
        if date is None:
            t = line[50:59].strip()
            if t:
                date = datetime.datetime.strptime(t, "%d-%b-%y").date()
        if name is None:
            name = line[10:50].strip() or None
    return id_, date, name


def _fmt(value: float, spec: str) -> str:
    return format(value, spec)


def _fixed_columns(values: np.ndarray, width: int, decimals: int) -> Tuple[np.ndarray, np.ndarray]:
    """``format(v, f"{width}.{decimals}f")`` (``decimals`` 0: ``format(v,
    f"{width}d")`` of integers) of every value, as an (n, width) byte array,
    and the rows it cannot write exactly: a number wider than its columns, a
    NaN or infinity, or a float whose scaled value lies within 1e-6 of a
    rounding tie (those need the exact decimal expansion that `format` does;
    below 10**8 the scaling itself errs by less than 1.2e-8)."""
    n = values.shape[0]
    if decimals:
        v = values.astype(np.float64)
        scaled = np.abs(v) * float(10 ** decimals)
        with np.errstate(invalid="ignore"):
            slow = ~(scaled < 1e8) | (np.abs(scaled - np.floor(scaled) - 0.5) < 1e-6)
        mag = np.where(slow, 0.0, np.rint(np.where(slow, 0.0, scaled))).astype(np.int64)
        neg = np.signbit(v)
    else:
        v = values.astype(np.int64)
        mag, neg = np.abs(v), v < 0
        slow = mag < 0                         # the most negative int64
        mag = np.where(slow, 0, mag)
    ip = mag // (10 ** decimals)
    digits = np.ones(n, dtype=np.int64)
    t = ip // 10
    while (t > 0).any():
        digits += t > 0
        t //= 10
    point = decimals + 1 if decimals else 0
    slow |= digits + neg + point > width
    out = np.full((n, width), ord(" "), dtype=np.uint8)
    m = np.where(slow, 0, mag)
    col = width - 1
    for _ in range(decimals):
        out[:, col] = 48 + m % 10
        m //= 10
        col -= 1
    if decimals:
        out[:, col] = ord(".")
        col -= 1
    for j in range(col + 1):
        has = j < digits
        out[has, col - j] = 48 + m[has] % 10
        m //= 10
    rows = np.nonzero(neg & ~slow)[0]
    out[rows, col - digits[rows]] = ord("-")
    return out, slow


def _decoded(tokens: np.ndarray) -> np.ndarray:
    """``_decode`` of a byte array: ``_`` and NUL become blanks."""
    return np.where((tokens == 0x5F) | (tokens == 0), np.uint8(32), tokens).astype(np.uint8)


# --------------------------------------------------------------------------
# Mat4
# --------------------------------------------------------------------------

class Mat4:
    """A 4x4 matrix of double precision (``_jess.pyx:177-240``)."""

    __slots__ = ("_data",)

    def __init__(self, data):
        arr = np.array(data, dtype=np.float64)
        if arr.shape != (4, 4):
            raise ValueError("invalid dimensions")
        self._data = arr

    @classmethod
    def _wrap(cls, arr: np.ndarray) -> "Mat4":
        m = cls.__new__(cls)
        m._data = arr
        return m

    @property
    def shape(self):
        return (4, 4)

    def __array__(self, dtype=None, copy=None):
        return self._data if dtype is None else self._data.astype(dtype)

    def __getitem__(self, index):
        return self._data[index]

    def __len__(self):
        return 4

    def __iter__(self):
        return iter(self._data)

    def __eq__(self, other):
        try:
            return bool(np.array_equal(self._data, np.asarray(other, dtype=np.float64)))
        except Exception:
            return NotImplemented

    def __repr__(self):
        return f"Mat4({self._data.tolist()!r})"

    def __reduce__(self):
        return type(self), (self._data.tolist(),)

    def __matmul__(self, other) -> "Mat4":
        o = np.asarray(other, dtype=np.float64)
        if o.shape != (4, 4):
            raise ValueError(f"dimension error: expected shape (4, 4), got {o.shape!r}")
        # same accumulation order as Mat4.__matmul__ (_jess.pyx:228-238)
        res = [[0.0] * 4 for _ in range(4)]
        a = self._data.tolist()
        b = o.tolist()
        for i in range(4):
            for j in range(4):
                s = 0.0
                for k in range(4):
                    s += a[i][k] * b[k][j]
                res[i][j] = s
        return Mat4._wrap(np.array(res, dtype=np.float64))


def _as_mat16(matrix) -> list:
    m = np.asarray(matrix, dtype=np.float64)
    if m.shape != (4, 4):
        raise ValueError(f"dimension error: expected shape (4, 4), got {tuple(m.shape)!r}")
    return [float(v) for v in m.reshape(16)]


def _apply(mat16: list, x: float, y: float, z: float) -> Tuple[float, float, float]:
    """Homogeneous transform, accumulated as ``Atom._transform`` (``_jess.pyx:1291-1304``)."""
    v = (x, y, z, 1.0)
    out = []
    for i in range(3):
        s = 0.0
        for j in range(4):
            s += mat16[4 * i + j] * v[j]
        out.append(s)
    return out[0], out[1], out[2]


# --------------------------------------------------------------------------
# Atom
# --------------------------------------------------------------------------

class Atom:
    """A single atom in a molecule (``_jess.pyx:909-1330``)."""

    __slots__ = (
        "serial", "_name", "_altloc", "_resname", "_chain1", "_chain2",
        "residue_number", "_icode", "x", "y", "z", "_occupancy", "_tempfactor",
        "_segid", "_element", "charge",
    )

    def __init__(
        self,
        *,
        serial: int,
        name: str,
        residue_name: str,
        chain_id: str,
        residue_number: int,
        x: float,
        y: float,
        z: float,
        occupancy: float = 0.0,
        temperature_factor: float = 0.0,
        altloc: str = " ",
        insertion_code: str = " ",
        segment: str = "",
        element: str = "",
        charge: int = 0,
    ):
        if len(name) > 4:
            raise ValueError(f"Invalid atom name: {name!r}")
        if len(residue_name) > 3:
            raise ValueError(f"Invalid residue name: {residue_name!r}")
        if len(segment) > 4:
            raise ValueError(f"Invalid segment: {segment!r}")
        if len(element) > 2:
            raise ValueError(f"Invalid element: {element!r}")
        c1, c2 = _chain_bytes(chain_id)
        self.serial = int(serial)
        self._name = _encode_atom_name(name)
        self._altloc = ord(altloc)
        self._resname = _encode_resname(residue_name)
        self._chain1 = c1
        self._chain2 = c2
        self.residue_number = int(residue_number)
        self._icode = ord(insertion_code)
        self.x = float(x)
        self.y = float(y)
        self.z = float(z)
        self._occupancy = float(np.float32(occupancy))
        self._tempfactor = float(np.float32(temperature_factor))
        self._segid = _encode(segment.encode("ascii"), 4)
        self._element = _encode(element.encode("ascii").rjust(2, b"\0"), 2)
        self.charge = int(charge)

    # -- parsing ---------------------------------------------------------

    @classmethod
    def loads(cls, text) -> "Atom":
        """Parse a PDB ``ATOM``/``HETATM`` record (fixed columns)."""
        if isinstance(text, (bytes, bytearray)):
            text = bytes(text).decode("utf-8")
        line = text.rstrip("\r\n")
        if not line.startswith(("ATOM", "HETATM")):
            raise ValueError(f"Failed to parse atom: {text!r}")
        line = line.ljust(80)
        try:
            serial = int(line[6:11])
            resseq = int(line[22:26])
            x = float(line[30:38])
            y = float(line[38:46])
            z = float(line[46:54])
        except ValueError:
            raise ValueError(f"Failed to parse atom: {text!r}") from None

        def _num(field, default=0.0):
            field = field.strip()
            return float(field) if field else default

        atom = cls.__new__(cls)
        atom.serial = serial
        atom._name = _encode(line[12:16].encode("ascii"), 4)
        atom._altloc = ord(line[16])
        atom._resname = _encode(line[17:20].encode("ascii"), 3) + b"\0"
        atom._chain1 = ord(line[20])
        atom._chain2 = ord(line[21])
        atom.residue_number = resseq
        atom._icode = ord(line[26])
        atom.x, atom.y, atom.z = x, y, z
        atom._occupancy = float(np.float32(_num(line[54:60])))
        atom._tempfactor = float(np.float32(_num(line[60:66])))
        atom._segid = _encode(line[72:76].encode("ascii"), 4)
        atom._element = _encode(line[76:78].encode("ascii"), 2)
        charge = line[78:80].strip()
        try:
            atom.charge = int(charge[::-1]) if charge else 0
        except ValueError:
            atom.charge = 0
        return atom

    @classmethod
    def load(cls, file) -> "Atom":
        return cls.loads(file.read())

    # -- properties ------------------------------------------------------

    @property
    def name(self) -> str:
        return self._name.decode("ascii").strip("_")

    @property
    def residue_name(self) -> str:
        return self._resname[:3].decode("ascii").strip("_")

    @property
    def altloc(self) -> str:
        return chr(self._altloc)

    @property
    def insertion_code(self) -> str:
        return chr(self._icode)

    @property
    def chain_id(self) -> str:
        return (chr(self._chain1) + chr(self._chain2)).strip()

    @property
    def segment(self) -> str:
        return self._segid.decode("ascii").strip("_")

    @property
    def element(self) -> str:
        return self._element.decode("ascii").strip("_")

    @property
    def occupancy(self) -> float:
        return self._occupancy

    @property
    def temperature_factor(self) -> float:
        return self._tempfactor

    # -- protocol --------------------------------------------------------

    def _state(self) -> dict:
        return {
            "serial": self.serial,
            "name": self.name,
            "altloc": self.altloc,
            "residue_name": self.residue_name,
            "chain_id": self.chain_id,
            "residue_number": self.residue_number,
            "insertion_code": self.insertion_code,
            "x": self.x,
            "y": self.y,
            "z": self.z,
            "temperature_factor": self.temperature_factor,
            "occupancy": self.occupancy,
            "segment": self.segment,
            "element": self.element,
            "charge": self.charge,
        }

    def __repr__(self):
        args = ", ".join(f"{k}={v!r}" for k, v in self._state().items())
        return f"{type(self).__name__}({args})"

    def __eq__(self, other):
        if not isinstance(other, Atom):
            return NotImplemented
        return self._state() == other._state()

    def __hash__(self):
        return hash(tuple(self._state().values()))

    def __reduce__(self):
        import functools
        return functools.partial(type(self), **self._state()), ()

    def __copy__(self):
        return self.copy()

    def copy(self) -> "Atom":
        new = Atom.__new__(type(self))
        for slot in Atom.__slots__:
            setattr(new, slot, getattr(self, slot))
        return new

    # -- output ----------------------------------------------------------

    def _line(self, x: float, y: float, z: float) -> str:
        # "ATOM  %5i%5s%c%-3s%c%c%4i%-4c%8.3f%8.3f%8.3f%6.2f%6.2f          %2s\n"
        return (
            "ATOM  "
            + f"{self.serial:5d}"
            + f"{_decode(self._name):>5s}"
            + chr(self._altloc)
            + f"{_decode(self._resname[:3]):<3s}"
            + chr(self._chain1)
            + chr(self._chain2)
            + f"{self.residue_number:4d}"
            + f"{chr(self._icode):<4s}"
            + f"{x:8.3f}{y:8.3f}{z:8.3f}"
            + f"{self._occupancy:6.2f}{self._tempfactor:6.2f}"
            + "          "
            + f"{_decode(self._element):>2s}\n"
        )

    def dumps(self, format: str = "pdb") -> str:
        if format != "pdb":
            raise ValueError(f"invalid value for `format` argument: {format!r}")
        return self._line(self.x, self.y, self.z)

    def dump(self, file, format: str = "pdb") -> None:
        file.write(self.dumps(format=format))

    def transform(self, matrix) -> "Atom":
        mat = _as_mat16(matrix)
        new = self.copy()
        new.x, new.y, new.z = _apply(mat, self.x, self.y, self.z)
        return new


# --------------------------------------------------------------------------
# Molecule
# --------------------------------------------------------------------------

_MOL_ARRAYS = frozenset(("xyz", "names", "resnames", "chains", "resseq", "serial", "altloc", "icode",
                         "occupancy", "tempfactor", "segid", "element", "charge"))


def _frozen(a: np.ndarray) -> np.ndarray:
    """A read-only array that no one else can write: `a` itself when it owns
    its data and is already read-only, else a read-only copy."""
    if not (isinstance(a, np.ndarray) and a.flags.owndata and not a.flags.writeable):
        a = np.array(a, copy=True)
        a.setflags(write=False)
    return a


class Molecule:
    """A molecule, stored as structure-of-arrays (``_jess.pyx:243-907``).

    The residue-name index of the reference (``ResIndex_create``, built in
    ``Molecule.__init__`` at ``_jess.pyx:682``) is built by the native library
    when the molecule is first uploaded to a device.

    Like the reference's, a molecule does not change once built: its arrays
    are read-only, so the device copies made for queries stay valid. Assigning
    a new array to an attribute replaces that data (with a private read-only
    copy) and drops the device copies.
    """

    def __setattr__(self, name, value):
        if name in _MOL_ARRAYS:
            value = _frozen(value)
            dev = self.__dict__.get("_device")
            if dev:
                dev.clear()
            object.__setattr__(self, "_version", self.__dict__.get("_version", 0) + 1)
        object.__setattr__(self, name, value)

    def __init__(self, atoms: Iterable[Atom] = (), id: Optional[str] = None,
                 date: Optional[datetime.date] = None, name: Optional[str] = None):
        if date is not None and not isinstance(date, datetime.date):
            raise TypeError(f"expected datetime.date or None, found {type(date).__name__} ")
        atoms = list(atoms)
        n = len(atoms)
        self._id = id
        self._date = date
        self._name = name
        self._device = {}
        xyz = np.empty((n, 3), dtype=np.float64)
        names = np.empty((n, 4), dtype=np.uint8)
        resnames = np.empty((n, 4), dtype=np.uint8)
        chains = np.empty((n, 2), dtype=np.uint8)
        resseq = np.empty(n, dtype=np.int32)
        serial = np.empty(n, dtype=np.int32)
        altloc = np.empty(n, dtype=np.uint8)
        icode = np.empty(n, dtype=np.uint8)
        occupancy = np.empty(n, dtype=np.float32)
        tempfactor = np.empty(n, dtype=np.float32)
        segid = np.empty((n, 4), dtype=np.uint8)
        element = np.empty((n, 2), dtype=np.uint8)
        charge = np.empty(n, dtype=np.int32)
        for i, a in enumerate(atoms):
            if not isinstance(a, Atom):
                raise TypeError(f"expected Atom, found {type(a).__name__}")
            xyz[i] = (a.x, a.y, a.z)
            names[i] = np.frombuffer(a._name, dtype=np.uint8)
            resnames[i] = np.frombuffer(a._resname, dtype=np.uint8)
            chains[i] = (a._chain1, a._chain2)
            resseq[i] = a.residue_number
            serial[i] = a.serial
            altloc[i] = a._altloc
            icode[i] = a._icode
            occupancy[i] = a._occupancy
            tempfactor[i] = a._tempfactor
            segid[i] = np.frombuffer(a._segid, dtype=np.uint8)
            element[i] = np.frombuffer(a._element, dtype=np.uint8)
            charge[i] = a.charge
        for attr, arr in (("xyz", xyz), ("names", names), ("resnames", resnames), ("chains", chains),
                          ("resseq", resseq), ("serial", serial), ("altloc", altloc), ("icode", icode),
                          ("occupancy", occupancy), ("tempfactor", tempfactor), ("segid", segid),
                          ("element", element), ("charge", charge)):
            arr.setflags(write=False)
            setattr(self, attr, arr)

    @classmethod
    def from_arrays(cls, xyz, names, resnames, chains, resseq, serial=None, *,
                    id: Optional[str] = None, occupancy=None, tempfactor=None,
                    element=None) -> "Molecule":
        """Build a molecule directly from arrays (encoded names, see module doc)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        mol = cls.__new__(cls)
        mol._id = id
        mol._date = None
        mol._name = None
        mol._device = {}
        mol.xyz = xyz
        mol.names = np.ascontiguousarray(names, dtype=np.uint8).reshape(n, 4)
        mol.resnames = np.ascontiguousarray(resnames, dtype=np.uint8).reshape(n, 4)
        mol.chains = np.ascontiguousarray(chains, dtype=np.uint8).reshape(n, 2)
        mol.resseq = np.ascontiguousarray(resseq, dtype=np.int32).reshape(n)
        mol.serial = (np.arange(1, n + 1, dtype=np.int32) if serial is None
                      else np.ascontiguousarray(serial, dtype=np.int32).reshape(n))
        mol.altloc = np.full(n, ord(" "), dtype=np.uint8)
        mol.icode = np.full(n, ord(" "), dtype=np.uint8)
        mol.occupancy = (np.ones(n, dtype=np.float32) if occupancy is None
                         else np.asarray(occupancy, dtype=np.float32))
        mol.tempfactor = (np.zeros(n, dtype=np.float32) if tempfactor is None
                          else np.asarray(tempfactor, dtype=np.float32))
        mol.segid = np.full((n, 4), ord("_"), dtype=np.uint8)
        mol.element = (np.full((n, 2), ord("_"), dtype=np.uint8) if element is None
                       else np.ascontiguousarray(element, dtype=np.uint8).reshape(n, 2))
        mol.charge = np.zeros(n, dtype=np.int32)
        return mol

    # -- parsing ---------------------------------------------------------

    @classmethod
    def loads(cls, text: str, format: str = "pdb", *, id: Optional[str] = None,
              date=None, name: Optional[str] = None, ignore_endmdl: bool = False,
              use_author: bool = False, skip_hetatm: bool = False) -> "Molecule":
        return cls.load(io.StringIO(text), format=format, id=id, date=date, name=name,
                        ignore_endmdl=ignore_endmdl, skip_hetatm=skip_hetatm)

    @classmethod
    def load(cls, file, format: str = "detect", *, id: Optional[str] = None, date=None,
             name: Optional[str] = None, ignore_endmdl: bool = False,
             use_author: bool = False, skip_hetatm: bool = False, device: Optional[bool] = None) -> "Molecule":
        """Load a PDB file (``PDBMoleculeParser.load``, ``src/pyjess/_io.py:68-103``).

        ``device``: parse the records on the GPU (`Molecule.load_many`), which
        also builds the molecule's device copy and residue-name index there;
        None (the default) does so for texts of ``DEVICE_PARSE_BYTES`` or more
        when a device is present. The two parsers give the same molecule.
        mmCIF input needs ``gemmi``, which the reference also imports lazily;
        it is not available here and raises `ValueError`.
        """
        if format not in ("detect", "pdb", "cif"):
            raise ValueError(f"invalid value for `format` argument: {format!r}")
        if format == "cif":
            raise ValueError("mmCIF parsing is not supported: load the structure from PDB text")
        if device is False:
            if isinstance(file, (str, os.PathLike)):
                with open(file) as f:
                    return cls._load_pdb(f, id, date, name, ignore_endmdl, skip_hetatm)
            return cls._load_pdb(file, id, date, name, ignore_endmdl, skip_hetatm)
        data = _read_bytes(file)
        if device is None:
            from . import _native

            device = len(data) >= DEVICE_PARSE_BYTES and _native.device_available()
        if not device:
            text = data.decode("utf-8")
            return cls._load_pdb(io.StringIO(text, newline=None), id, date, name, ignore_endmdl, skip_hetatm)
        return cls._load_many_bytes([data], [(id, date, name)], ignore_endmdl, skip_hetatm)[0]

    @classmethod
    def load_many(cls, files, *, ignore_endmdl: bool = False, skip_hetatm: bool = False) -> typing.List["Molecule"]:
        """Load many PDB files at once, their records parsed on the GPU in one
        pass (``jam_molecules_parse_pdb``), each molecule's residue-name index
        built there too. The molecules are the ones `Molecule.load` reads,
        and their device copy is kept: a `Jess.query_batch` over the same
        list, or a `Jess.query` of one that was loaded alone, uploads nothing."""
        datas = [_read_bytes(f) for f in files]
        return cls._load_many_bytes(datas, [(None, None, None)] * len(datas), ignore_endmdl, skip_hetatm)

    @classmethod
    def _load_many_bytes(cls, datas, headers, ignore_endmdl, skip_hetatm) -> typing.List["Molecule"]:
        from . import _native

        ctx = _native.context()
        # the header fields come from the text (decoding also rejects what a
        # text-mode read would)
        heads = [_pdb_header(d, *h, ignore_endmdl) for d, h in zip(datas, headers)]
        batch = _native.ParsedMolecules(ctx, datas, skip_hetatm=skip_hetatm, ignore_endmdl=ignore_endmdl)
        off = batch.atom_offsets()
        a = batch.arrays()
        mols = []
        for k, (id_, date, name) in enumerate(heads):
            s0, s1 = int(off[k]), int(off[k + 1])
            mol = cls.from_arrays(a["xyz"][s0:s1], a["names"][s0:s1], a["resnames"][s0:s1], a["chains"][s0:s1],
                                  a["resseq"][s0:s1], a["serial"][s0:s1], id=id_,
                                  occupancy=a["occupancy"][s0:s1], tempfactor=a["tempfactor"][s0:s1],
                                  element=a["element"][s0:s1])
            mol._date, mol._name = date, name
            mol.altloc = a["altloc"][s0:s1]
            mol.icode = a["icode"][s0:s1]
            mol.segid = a["segid"][s0:s1]
            mol.charge = a["charge"][s0:s1]
            mols.append(mol)
        if len(mols) == 1:
            mols[0]._device[id(ctx)] = batch
        else:
            from . import _jess

            _jess._last_batch[:] = [ctx, list(mols), [m.__dict__.get("_version", 0) for m in mols], batch]
        return mols

    @classmethod
    def _load_pdb(cls, f, id_, date, name, ignore_endmdl, skip_hetatm) -> "Molecule":
        atoms = []
        for line in f:
            if line.startswith("HEADER"):
                if id_ is None:
                    id_ = line[62:66].strip() or None
                if date is None:
                    text = line[50:59].strip()
                    if text:
                        date = datetime.datetime.strptime(text, "%d-%b-%y").date()
                if name is None:
                    name = line[10:50].strip() or None
            elif line.startswith("ATOM"):
                atoms.append(Atom.loads(line))
            elif line.startswith("HETATM") and not skip_hetatm:
                atoms.append(Atom.loads(line))
            elif line.startswith("ENDMDL"):
                if not ignore_endmdl:
                    break
            elif line.lower().startswith(("data_", "loop_")):
                raise ValueError("mmCIF data tags found, file is not in PDB format")
        return cls(atoms, id=id_, date=date, name=name)

    # -- protocol --------------------------------------------------------

    def __len__(self) -> int:
        return int(self.xyz.shape[0])

    def _atom(self, i: int) -> Atom:
        a = Atom.__new__(Atom)
        a.serial = int(self.serial[i])
        a._name = bytes(self.names[i])
        a._altloc = int(self.altloc[i])
        a._resname = bytes(self.resnames[i])
        a._chain1 = int(self.chains[i, 0])
        a._chain2 = int(self.chains[i, 1])
        a.residue_number = int(self.resseq[i])
        a._icode = int(self.icode[i])
        a.x, a.y, a.z = (float(v) for v in self.xyz[i])
        a._occupancy = float(self.occupancy[i])
        a._tempfactor = float(self.tempfactor[i])
        a._segid = bytes(self.segid[i])
        a._element = bytes(self.element[i])
        a.charge = int(self.charge[i])
        return a

    def __getitem__(self, index):
        n = len(self)
        if isinstance(index, slice):
            return type(self)([self._atom(i) for i in range(*index.indices(n))],
                              id=self.id, date=self.date, name=self.name)
        i = int(index)
        if i < 0:
            i += n
        if i < 0 or i >= n:
            raise IndexError(index)
        return self._atom(i)

    def __iter__(self):
        for i in range(len(self)):
            yield self._atom(i)

    def __eq__(self, other):
        if not isinstance(other, Molecule):
            return NotImplemented
        if self._id != other._id or len(self) != len(other):
            return False
        return all(x == y for x, y in zip(self, other))

    def __hash__(self):
        return hash((self._id, *(hash(x) for x in self)))

    def __reduce__(self):
        import functools
        return functools.partial(type(self), list(self), id=self.id, name=self.name,
                                 date=self.date), ()

    @property
    def id(self) -> Optional[str]:
        return self._id

    @property
    def date(self):
        return self._date

    @property
    def name(self) -> Optional[str]:
        return self._name

    def _subset(self, keep: np.ndarray) -> "Molecule":
        mol = Molecule.from_arrays(self.xyz[keep], self.names[keep], self.resnames[keep],
                                   self.chains[keep], self.resseq[keep], self.serial[keep],
                                   id=self._id)
        mol._date, mol._name = self._date, self._name
        for attr in ("altloc", "icode", "occupancy", "tempfactor", "segid", "element", "charge"):
            setattr(mol, attr, np.ascontiguousarray(getattr(self, attr)[keep]))
        return mol

    def conserved(self, cutoff: float = 0.0) -> "Molecule":
        """Atoms whose temperature factor is at least ``cutoff`` (``_jess.pyx:768-791``)."""
        if cutoff <= 0.0:
            return self.copy()
        return self._subset(self.tempfactor >= cutoff)

    def copy(self) -> "Molecule":
        return self._subset(np.ones(len(self), dtype=bool))

    def transform(self, matrix) -> "Molecule":
        mat = _as_mat16(matrix)
        new = self.copy()
        out = np.empty_like(self.xyz)
        for i in range(len(self)):
            out[i] = _apply(mat, *self.xyz[i].tolist())
        out.setflags(write=False)
        new.xyz = out
        return new

    def dumps(self, format: str = "pdb", write_header: bool = True) -> str:
        if format != "pdb":
            raise ValueError(f"invalid value for `format` argument: {format!r}")
        head = ""
        if write_header:
            name = (self.name or "")[:40]
            id_ = (self.id or "")[:4]
            date = "" if self.date is None else self.date.strftime("%d-%b-%y").upper()
            head = f"{'HEADER':<6}{'':4}{name:40}{date:9}{'':3}{id_:4}\n"
        return head + self._atom_records()

    def _atom_records(self) -> str:
        """``"".join(atom.dumps() for atom in self)``, the columns of all atoms
        written at once; a row with a field that `_fixed_columns` cannot
        write exactly is formatted by `Atom.dumps`."""
        n = len(self)
        if n == 0:
            return ""
        buf = np.full((n, 79), ord(" "), dtype=np.uint8)
        buf[:, 0:6] = np.frombuffer(b"ATOM  ", dtype=np.uint8)
        slow = np.zeros(n, dtype=bool)
        for col, width, dec, values in ((6, 5, 0, self.serial), (22, 4, 0, self.resseq),
                                        (30, 8, 3, self.xyz[:, 0]), (38, 8, 3, self.xyz[:, 1]),
                                        (46, 8, 3, self.xyz[:, 2]), (54, 6, 2, self.occupancy),
                                        (60, 6, 2, self.tempfactor)):
            cols, s = _fixed_columns(values, width, dec)
            buf[:, col:col + width] = cols
            slow |= s
        buf[:, 12:16] = _decoded(self.names)
        buf[:, 16] = self.altloc
        buf[:, 17:20] = _decoded(self.resnames[:, :3])
        buf[:, 20:22] = self.chains
        buf[:, 26] = self.icode
        buf[:, 76:78] = _decoded(self.element)
        buf[:, 78] = ord("\n")
        text = buf.tobytes().decode("latin-1")     # one character per byte, as chr() gives
        if not slow.any():
            return text
        lines = [text[79 * i:79 * i + 79] for i in range(n)]
        for i in np.nonzero(slow)[0]:
            lines[i] = self._atom(int(i)).dumps()
        return "".join(lines)

    def dump(self, file, format: str = "pdb", write_header: bool = True) -> None:
        file.write(self.dumps(format=format, write_header=write_header))


# --------------------------------------------------------------------------
# TemplateAtom
# --------------------------------------------------------------------------

class TemplateAtom:
    """A single template atom (``_jess.pyx:1332-1748``)."""

    __slots__ = ("_mode", "_resseq", "_chain1", "_chain2", "_names", "_resnames",
                 "x", "y", "z", "_weight")

    def __init__(self, *, chain_id: str, residue_number: int, x: float, y: float, z: float,
                 residue_names, atom_names, distance_weight: float = 0.0,
                 match_mode: int = 0):
        match_mode = int(match_mode)
        if match_mode not in _VALID_MODES:
            raise ValueError(f"Invalid match mode: {match_mode!r}")
        if len(chain_id) > 2:
            raise ValueError(f"Invalid chain ID: {chain_id!r}")
        self._mode = match_mode
        self._resseq = int(residue_number)
        self._chain1, self._chain2 = _chain_bytes(chain_id)
        self.x, self.y, self.z = float(x), float(y), float(z)
        self._weight = float(distance_weight)
        names = []
        for nm in atom_names:
            raw = bytearray(nm, "ascii") if isinstance(nm, str) else bytearray(nm)
            if len(raw) > 4:
                raise ValueError(f"Invalid atom name: {nm!r}")
            if len(raw) <= 3:
                raw.insert(0, ord("_"))
            names.append(_encode(bytes(raw), 4))
        self._names = tuple(names)
        self._resnames = tuple(_encode_resname(r) for r in residue_names)

    @classmethod
    def loads(cls, text) -> "TemplateAtom":
        """Parse a template line: the serial field holds the match mode, the
        columns after the coordinates list extra residue names as one-letter
        codes, then the distance weight (layout of ``TemplateAtom.dump``,
        ``_jess.pyx:1684-1703``)."""
        if isinstance(text, (bytes, bytearray)):
            text = bytes(text).decode("utf-8")
        line = text.rstrip("\r\n")
        if not line.startswith(("ATOM", "HETATM")):
            raise ValueError(f"Failed to parse template atom: {text!r}")
        padded = line.ljust(80)
        try:
            mode = int(padded[6:11])
            resseq = int(padded[22:26])
            x = float(padded[30:38])
            y = float(padded[38:46])
            z = float(padded[46:54])
        except ValueError:
            raise ValueError(f"Failed to parse template atom: {text!r}") from None
        if mode not in _VALID_MODES:
            raise ValueError(f"Invalid match mode: {mode!r}")
        resnames = [_encode(padded[17:20].encode("ascii"), 3) + b"\0"]
        for c in padded[55:61]:
            if c == " ":
                continue
            three = _ONE_TO_THREE.get(c)
            if three is None:
                raise ValueError(f"Failed to parse template atom: {text!r}")
            resnames.append(three.encode("ascii") + b"\0")
        weight_field = padded[61:].strip()
        try:
            weight = float(weight_field.split()[0]) if weight_field else 0.0
        except ValueError:
            raise ValueError(f"Failed to parse template atom: {text!r}") from None
        atom = cls.__new__(cls)
        atom._mode = mode
        atom._resseq = resseq
        atom._chain1 = ord(padded[20])
        atom._chain2 = ord(padded[21])
        atom.x, atom.y, atom.z = x, y, z
        atom._weight = weight
        atom._names = (_encode(padded[12:16].encode("ascii"), 4),)
        atom._resnames = tuple(resnames)
        return atom

    @classmethod
    def load(cls, file) -> "TemplateAtom":
        if isinstance(file, (str, os.PathLike)):
            with open(file) as f:
                return cls.loads(f.read())
        return cls.loads(file.read())

    @property
    def match_mode(self) -> int:
        return self._mode

    @property
    def residue_number(self) -> int:
        return self._resseq

    @property
    def chain_id(self) -> str:
        if self._chain1 == ord(" "):
            return chr(self._chain2)
        return (chr(self._chain1) + chr(self._chain2)).strip()

    @property
    def atom_names(self) -> Tuple[str, ...]:
        return tuple(n.replace(b"_", b"").decode() for n in self._names)

    @property
    def residue_names(self) -> Tuple[str, ...]:
        return tuple(r[:3].replace(b"_", b"").replace(b"\0", b"").decode() for r in self._resnames)

    @property
    def distance_weight(self) -> float:
        return self._weight

    def _state(self) -> dict:
        return {
            "chain_id": self.chain_id,
            "residue_number": self.residue_number,
            "x": self.x,
            "y": self.y,
            "z": self.z,
            "residue_names": self.residue_names,
            "atom_names": self.atom_names,
            "distance_weight": self.distance_weight,
            "match_mode": self.match_mode,
        }

    def __repr__(self):
        args = ", ".join(f"{k}={v!r}" for k, v in self._state().items())
        return f"{type(self).__name__}({args})"

    def __eq__(self, other):
        if not isinstance(other, TemplateAtom):
            return NotImplemented
        return self._state() == other._state()

    def __hash__(self):
        return hash(tuple(self._state().values()))

    def __reduce__(self):
        import functools
        return functools.partial(type(self), **self._state()), ()

    def __copy__(self):
        return self.copy()

    def copy(self) -> "TemplateAtom":
        new = TemplateAtom.__new__(type(self))
        for slot in TemplateAtom.__slots__:
            setattr(new, slot, getattr(self, slot))
        return new

    def dumps(self) -> str:
        name = _decode(self._names[0]) if self._names else "    "
        resname = _decode(self._resnames[0][:3]) if self._resnames else "   "
        line = (
            f"ATOM  {self._mode:5d} {name:>4s} {resname:>3s}"
            f"{chr(self._chain1)}{chr(self._chain2)}{self._resseq:4d}    "
            f"{self.x:8.3f}{self.y:8.3f}{self.z:8.3f}"
        )
        extra = [" "] * 7
        for k in range(1, len(self._resnames)):
            if k < 7:
                extra[k] = _THREE_TO_ONE.get(self._resnames[k][:3].decode("ascii"), "X")
        return line + "".join(extra) + f"{self._weight:4.2f}"

    def dump(self, file) -> None:
        file.write(self.dumps())

    def transform(self, matrix) -> "TemplateAtom":
        mat = _as_mat16(matrix)
        new = self.copy()
        new.x, new.y, new.z = _apply(mat, self.x, self.y, self.z)
        return new


# --------------------------------------------------------------------------
# Template
# --------------------------------------------------------------------------

class Template:
    """A template: a sequence of `TemplateAtom` (``_jess.pyx:1751-2088``)."""

    def __init__(self, atoms: Iterable[TemplateAtom] = (), id: Optional[str] = None):
        self._atoms = tuple(a.copy() for a in atoms)
        for a in self._atoms:
            if not isinstance(a, TemplateAtom):
                raise TypeError(f"expected TemplateAtom, found {type(a).__name__}")
        self._id = id
        # number of distinct (resSeq, chainID1, chainID2), _jess.pyx:1903-1912
        self._dim = len({(a._resseq, a._chain1, a._chain2) for a in self._atoms})

    @classmethod
    def loads(cls, text: str, id: Optional[str] = None) -> "Template":
        return cls.load(io.StringIO(text), id=id)

    @classmethod
    def load(cls, file, id: Optional[str] = None) -> "Template":
        if isinstance(file, (str, os.PathLike)):
            with open(file) as f:
                return cls._load(f, id)
        return cls._load(file, id)

    @classmethod
    def _load(cls, f, id_) -> "Template":
        atoms = []
        for line in f:
            if line.startswith("ATOM"):
                atoms.append(TemplateAtom.loads(line))
            elif id_ is None and line.startswith("REMARK PDB_ID"):
                id_ = line.split(" ", maxsplit=2)[2].strip()
        return cls(atoms, id=id_)

    def __len__(self) -> int:
        return len(self._atoms)

    def __getitem__(self, index):
        if isinstance(index, slice):
            return type(self)(self._atoms[index], id=self.id)
        return self._atoms[index]

    def __iter__(self):
        return iter(self._atoms)

    def __eq__(self, other):
        if not isinstance(other, Template):
            return NotImplemented
        return (self.id == other.id and self.dimension == other.dimension
                and len(self) == len(other) and all(x == y for x, y in zip(self, other)))

    def __hash__(self):
        return hash((self.id, *(hash(x) for x in self)))

    def __reduce__(self):
        return type(self), (list(self._atoms), self.id)

    def __copy__(self):
        return self.copy()

    @property
    def id(self) -> Optional[str]:
        return self._id

    @property
    def dimension(self) -> int:
        return self._dim

    def copy(self) -> "Template":
        return type(self)(self._atoms, id=self._id)

    def dumps(self) -> str:
        out = io.StringIO()
        self.dump(out)
        return out.getvalue()

    def dump(self, file, write_id: bool = True) -> None:
        if write_id and self.id is not None:
            file.write(f"REMARK PDB_ID {self.id}\n")
        for atom in self._atoms:
            atom.dump(file)
            file.write("\n")

    def transform(self, matrix) -> "Template":
        return type(self)([a.transform(matrix) for a in self._atoms], id=self._id)
