"""Python driver of the CPU oracle.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only to check
or to time the reference algorithm; the product path never calls it.

``query_stream`` returns every complete match of a query, in the order the
reference engine produces them, with its superposition. ``replay`` restates the
host-side policy loop of ``Query.__next__`` (reference
``src/pyjess/_jess.pyx:2197-2283``) literally over that stream.
"""

from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjess_oracle.so")

# The kd-tree child push order of KdTreeQuery_next is not visible in the
# reference patches; the ordered hit list of tests/golden/1AMY+1.3.3.txt fixes
# it (tests/test_oracle_fixtures.py::test_kd_push_order_is_pinned).
KD_RIGHT_FIRST = 1

_IGNORE = {None: 0, "residues": 1, "atoms": 2}


class JoMol(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("xyz", ctypes.c_void_p),
        ("name", ctypes.c_void_p),
        ("resname", ctypes.c_void_p),
        ("chain2", ctypes.c_void_p),
        ("resseq", ctypes.c_void_p),
        ("serial", ctypes.c_void_p),
    ]


class JoTpls(ctypes.Structure):
    _fields_ = [
        ("n_templates", ctypes.c_int32),
        ("atom_off", ctypes.c_void_p),
        ("mode", ctypes.c_void_p),
        ("resseq", ctypes.c_void_p),
        ("chain2", ctypes.c_void_p),
        ("pos", ctypes.c_void_p),
        ("weight", ctypes.c_void_p),
        ("name_off", ctypes.c_void_p),
        ("names", ctypes.c_void_p),
        ("resname_off", ctypes.c_void_p),
        ("resnames", ctypes.c_void_p),
    ]


class JoParams(ctypes.Structure):
    _fields_ = [
        ("distance_cutoff", ctypes.c_double),
        ("max_dynamic_distance", ctypes.c_double),
        ("ignore_chain", ctypes.c_int32),
        ("reorder", ctypes.c_int32),
        ("kd_right_first", ctypes.c_int32),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc), return the library path."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


_native_lib = None
NATIVE_PATH = os.path.join(HERE, "_native", "libjess_oracle.so")


def native_lib():
    """The same oracle compiled on this host for this host (-march=native,
    still without contraction: the same results), for timing the CPU
    baseline. Rebuilt once per process, since the library may have been built
    on another machine."""
    global _native_lib
    if _native_lib is None:
        subprocess.run(["make", "-s", "-C", HERE, "-B", "native"], check=True)
        _native_lib = _declare(ctypes.CDLL(NATIVE_PATH))
    return _native_lib


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def _declare(L):
    """Argument and result types of the exported functions."""
    P = ctypes.POINTER
    L.jo_query.argtypes = [P(JoMol), P(JoTpls), P(JoParams), ctypes.c_void_p, ctypes.c_int32,
                           P(ctypes.c_void_p)]
    L.jo_query.restype = ctypes.c_int32
    L.jo_stream_count.argtypes = [ctypes.c_void_p]
    L.jo_stream_count.restype = ctypes.c_int64
    L.jo_stream_atom_total.argtypes = [ctypes.c_void_p]
    L.jo_stream_atom_total.restype = ctypes.c_int64
    L.jo_stream_copy.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 7
    L.jo_stream_copy.restype = None
    L.jo_stream_free.argtypes = [ctypes.c_void_p]
    L.jo_stream_free.restype = None
    L.jo_superpose.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32] + [ctypes.c_void_p] * 4
    L.jo_superpose.restype = None
    L.jo_bruteforce.argtypes = [P(JoMol), P(JoTpls), ctypes.c_int32, P(JoParams), ctypes.c_void_p,
                                ctypes.c_int64]
    L.jo_bruteforce.restype = ctypes.c_int64
    L.jo_candidates.argtypes = [P(JoMol), P(JoTpls), ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_void_p, ctypes.c_int32]
    L.jo_candidates.restype = ctypes.c_int32
    L.jo_kd_order.argtypes = [P(JoMol), ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    L.jo_kd_order.restype = ctypes.c_int32
    L.jo_batch.argtypes = [P(JoMol), ctypes.c_int32, P(JoTpls), P(JoParams), ctypes.c_double,
                           ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.jo_batch.restype = ctypes.c_int32
    L.jo_work.argtypes = [P(JoMol), P(JoTpls), P(JoParams), ctypes.c_void_p]
    L.jo_work.restype = ctypes.c_int32
    return L


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class MolView:
    """Keeps the arrays of one molecule alive behind a `JoMol`."""

    def __init__(self, xyz, names, resnames, chain2, resseq, serial):
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = self.xyz.shape[0]
        self.names = np.ascontiguousarray(names, dtype=np.uint8).reshape(n, 4)
        self.resnames = np.ascontiguousarray(resnames, dtype=np.uint8).reshape(n, 4)
        self.chain2 = np.ascontiguousarray(chain2, dtype=np.uint8).reshape(n)
        self.resseq = np.ascontiguousarray(resseq, dtype=np.int32).reshape(n)
        self.serial = np.ascontiguousarray(serial, dtype=np.int32).reshape(n)
        self.c = JoMol(n, _ptr(self.xyz), _ptr(self.names), _ptr(self.resnames), _ptr(self.chain2),
                       _ptr(self.resseq), _ptr(self.serial))

    @classmethod
    def of(cls, molecule) -> "MolView":
        """From any object exposing the arrays of `pyjess_amd.Molecule`."""
        return cls(molecule.xyz, molecule.names, molecule.resnames, molecule.chains[:, 1],
                   molecule.resseq, molecule.serial)


class TplView:
    """Flattened templates behind a `JoTpls` (atoms read through the public API)."""

    def __init__(self, templates):
        off, mode, resseq, chain2, pos, weight = [0], [], [], [], [], []
        name_off, names, rn_off, rns = [0], [], [0], []
        for t in templates:
            for a in t:
                mode.append(a.match_mode)
                resseq.append(a.residue_number)
                # raw 4-byte encodings when the object carries them
                chain2.append(a._chain2)
                pos.append((a.x, a.y, a.z))
                weight.append(a.distance_weight)
                names.extend(bytes(nm) for nm in a._names)
                name_off.append(len(names))
                rns.extend(bytes(r) for r in a._resnames)
                rn_off.append(len(rns))
            off.append(len(mode))
        self.n_templates = len(off) - 1
        self.atom_off = np.array(off, dtype=np.int32)
        self.mode = np.array(mode, dtype=np.int32)
        self.resseq = np.array(resseq, dtype=np.int32)
        self.chain2 = np.array(chain2, dtype=np.uint8)
        self.pos = np.array(pos, dtype=np.float64).reshape(-1, 3)
        self.weight = np.array(weight, dtype=np.float64)
        self.name_off = np.array(name_off, dtype=np.int32)
        self.names = np.frombuffer(b"".join(names) or b"\0", dtype=np.uint8).copy()
        self.resname_off = np.array(rn_off, dtype=np.int32)
        self.resnames = np.frombuffer(b"".join(rns) or b"\0", dtype=np.uint8).copy()
        self.c = JoTpls(self.n_templates, _ptr(self.atom_off), _ptr(self.mode), _ptr(self.resseq),
                        _ptr(self.chain2), _ptr(self.pos), _ptr(self.weight), _ptr(self.name_off),
                        _ptr(self.names), _ptr(self.resname_off), _ptr(self.resnames))


@dataclass
class Match:
    scan: int               # position in the scan order
    template: int           # index of the template in the list given
    atoms: Tuple[int, ...]  # molecule atom index per template atom
    rmsd: float
    rotation: Tuple[float, ...]
    c0: Tuple[float, ...]
    c1: Tuple[float, ...]

    @property
    def nan(self) -> bool:
        return any(math.isnan(v) for v in self.rotation)


def params(distance_cutoff, max_dynamic_distance, ignore_chain=None, reorder=True,
           kd_right_first=None) -> JoParams:
    if kd_right_first is None:
        kd_right_first = KD_RIGHT_FIRST
    return JoParams(float(distance_cutoff), float(max_dynamic_distance), _IGNORE[ignore_chain],
                    int(bool(reorder)), int(kd_right_first))


def query_stream(molecule, templates, distance_cutoff, max_dynamic_distance, ignore_chain=None,
                 reorder=True, kd_right_first=None) -> List[Match]:
    """All complete matches, templates scanned in reverse insertion order
    (Jess_addTemplate prepends, patches/Jess.c.patch:63-72)."""
    L = lib()
    mv = molecule if isinstance(molecule, MolView) else MolView.of(molecule)
    tv = templates if isinstance(templates, TplView) else TplView(templates)
    order = np.arange(tv.n_templates - 1, -1, -1, dtype=np.int32)
    p = params(distance_cutoff, max_dynamic_distance, ignore_chain, reorder, kd_right_first)
    h = ctypes.c_void_p()
    rc = L.jo_query(ctypes.byref(mv.c), ctypes.byref(tv.c), ctypes.byref(p), _ptr(order),
                    int(tv.n_templates), ctypes.byref(h))
    try:
        if rc != 0:
            raise RuntimeError(f"oracle query failed with code {rc}")
        n = L.jo_stream_count(h)
        na = L.jo_stream_atom_total(h)
        tpl = np.zeros(max(n, 1), dtype=np.int32)
        aoff = np.zeros(n + 1, dtype=np.int64)
        atoms = np.zeros(max(na, 1), dtype=np.int32)
        rmsd = np.zeros(max(n, 1))
        rot = np.zeros((max(n, 1), 9))
        c0 = np.zeros((max(n, 1), 3))
        c1 = np.zeros((max(n, 1), 3))
        L.jo_stream_copy(h, _ptr(tpl), _ptr(aoff), _ptr(atoms), _ptr(rmsd), _ptr(rot), _ptr(c0), _ptr(c1))
    finally:
        L.jo_stream_free(h)
    scan_pos = {int(t): i for i, t in enumerate(order)}
    out = []
    for i in range(n):
        t = int(tpl[i])
        out.append(Match(scan_pos[t], t, tuple(int(x) for x in atoms[aoff[i]:aoff[i + 1]]),
                         float(rmsd[i]), tuple(rot[i].tolist()), tuple(c0[i].tolist()),
                         tuple(c1[i].tolist())))
    return out


def replay(stream: Sequence[Match], rmsd_threshold: float, best_match: bool = False,
           max_candidates: Optional[int] = None) -> Tuple[List[Match], int]:
    """The loop of ``Query.__next__`` (``_jess.pyx:2197-2283``) over a stream.

    Returns the yielded hits and the number of NaN-rotation warnings.
    """
    mc = -1 if max_candidates is None else int(max_candidates)
    if mc < -1:
        raise ValueError(f"invalid value for `max_candidates` argument: {max_candidates!r}")
    out: List[Match] = []
    warnings = 0
    n = len(stream)
    pos = 0
    cur = -1
    partial = False
    counter = 0
    while True:                       # one iteration per __next__ call
        tpl = None
        hit = None
        hit_rmsd = math.inf
        hit_tpl = None
        found = False
        while True:                   # while self._advance()
            if partial:
                partial = False
            else:
                if pos >= n:
                    break
                cur = pos
                pos += 1
            prev_tpl = tpl
            c = stream[cur]
            tpl = c.scan
            if found and hit_tpl != tpl:
                partial = True        # self._rewind()
                break
            if c.rmsd <= rmsd_threshold and c.rmsd < hit_rmsd:
                if c.nan:
                    warnings += 1
                else:
                    hit, hit_rmsd, hit_tpl, found = c, c.rmsd, tpl, True
            if tpl != prev_tpl:
                counter = 0
            else:
                counter += 1
            if mc != -1 and counter > mc:
                counter = 0
                while pos < n and stream[pos].scan == tpl:   # JessQuery_nextTemplate
                    pos += 1
            if found and not best_match:
                break
        if not found:
            return out, warnings
        out.append(hit)


def hits(molecule, templates, rmsd_threshold, distance_cutoff, max_dynamic_distance, *,
         max_candidates=None, ignore_chain=None, best_match=False, reorder=True,
         kd_right_first=None) -> List[Match]:
    """`Jess(templates).query(...)` as the reference would yield it."""
    stream = query_stream(molecule, templates, distance_cutoff, max_dynamic_distance,
                          ignore_chain, reorder, kd_right_first)
    return replay(stream, rmsd_threshold, best_match, max_candidates)[0]


def superpose(a, b):
    L = lib()
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 3)
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1, 3)
    rmsd = np.zeros(1)
    rot = np.zeros(9)
    c0 = np.zeros(3)
    c1 = np.zeros(3)
    L.jo_superpose(_ptr(a), _ptr(b), a.shape[0], _ptr(rmsd), _ptr(rot), _ptr(c0), _ptr(c1))
    return float(rmsd[0]), rot.reshape(3, 3), c0, c1


def bruteforce(molecule, templates, t, distance_cutoff, max_dynamic_distance, ignore_chain=None,
               cap=1 << 16):
    L = lib()
    mv = MolView.of(molecule) if not isinstance(molecule, MolView) else molecule
    tv = TplView(templates) if not isinstance(templates, TplView) else templates
    n = int(tv.atom_off[t + 1] - tv.atom_off[t])
    out = np.zeros(max(cap * n, 1), dtype=np.int32)
    p = params(distance_cutoff, max_dynamic_distance, ignore_chain)
    cnt = L.jo_bruteforce(ctypes.byref(mv.c), ctypes.byref(tv.c), int(t), ctypes.byref(p), _ptr(out), cap)
    if cnt < 0:
        raise RuntimeError(f"oracle brute force failed with code {cnt}")
    return [tuple(int(x) for x in out[i * n:(i + 1) * n]) for i in range(min(cnt, cap))]


def candidates(molecule, templates, t, k):
    L = lib()
    mv = MolView.of(molecule) if not isinstance(molecule, MolView) else molecule
    tv = TplView(templates) if not isinstance(templates, TplView) else templates
    out = np.zeros(max(mv.xyz.shape[0], 1), dtype=np.int32)
    c = L.jo_candidates(ctypes.byref(mv.c), ctypes.byref(tv.c), int(t), int(k), _ptr(out), out.shape[0])
    if c < 0:
        raise RuntimeError(f"oracle candidates failed with code {c}")
    return [int(x) for x in out[:c]]


def kd_order(molecule, set_atoms):
    L = lib()
    mv = MolView.of(molecule) if not isinstance(molecule, MolView) else molecule
    sa = np.ascontiguousarray(set_atoms, dtype=np.int32)
    out = np.zeros(max(sa.shape[0], 1), dtype=np.int32)
    L.jo_kd_order(ctypes.byref(mv.c), _ptr(sa), sa.shape[0], _ptr(out))
    return [int(x) for x in out[:sa.shape[0]]]


def batch(mol_views: Sequence[MolView], templates, rmsd_threshold, distance_cutoff,
          max_dynamic_distance, ignore_chain=None, reorder=True, n_threads=1, native=False):
    """Every (template, molecule) pair on ``n_threads`` host threads (with
    ``native``, the build for this host).

    Returns per-pair arrays (molecule-major) of complete-match counts, hit
    counts and order-independent hit digests."""
    L = native_lib() if native else lib()
    tv = templates if isinstance(templates, TplView) else TplView(templates)
    arr = (JoMol * len(mol_views))(*[mv.c for mv in mol_views])
    npairs = len(mol_views) * tv.n_templates
    cm = np.zeros(max(npairs, 1), dtype=np.int64)
    hc = np.zeros(max(npairs, 1), dtype=np.int64)
    dg = np.zeros(max(npairs, 1), dtype=np.uint64)
    p = params(distance_cutoff, max_dynamic_distance, ignore_chain, reorder)
    rc = L.jo_batch(arr, len(mol_views), ctypes.byref(tv.c), ctypes.byref(p), float(rmsd_threshold),
                    int(n_threads), _ptr(cm), _ptr(hc), _ptr(dg))
    if rc != 0:
        raise RuntimeError(f"oracle batch failed with code {rc}")
    return cm[:npairs], hc[:npairs], dg[:npairs]


def work(molecule, templates, distance_cutoff, max_dynamic_distance, ignore_chain=None, reorder=True):
    """Algorithmic work of one query on the reference algorithm: dict with
    ``inbox_tests`` (sum over opened levels L of L x candidates inside the join
    box), ``levels`` (opened levels), ``complete`` (complete matches) and
    ``complete_atoms`` (sum of their sizes)."""
    L = lib()
    mv = MolView.of(molecule) if not isinstance(molecule, MolView) else molecule
    tv = TplView(templates) if not isinstance(templates, TplView) else templates
    out = np.zeros(4, dtype=np.int64)
    p = params(distance_cutoff, max_dynamic_distance, ignore_chain, reorder)
    rc = L.jo_work(ctypes.byref(mv.c), ctypes.byref(tv.c), ctypes.byref(p), _ptr(out))
    if rc != 0:
        raise RuntimeError(f"oracle work count failed with code {rc}")
    return dict(inbox_tests=int(out[0]), levels=int(out[1]), complete=int(out[2]),
                complete_atoms=int(out[3]))
