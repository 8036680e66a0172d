"""ctypes binding of the native library (``libjess_amd.so``, C ABI in ``include/jess_amd.h``).

The library must be built (``python pyjess_amd/_build.py``) and a gfx950 GPU
must be visible for queries; there is no CPU fallback, every failure raises.
"""

from __future__ import annotations

import ctypes
import importlib.util
import os
import sys
import threading
from typing import List, Optional, Sequence

import numpy as np

from . import _build

JAM_OK = 0
JAM_ENOMEM = -1
JAM_EINVAL = -2
JAM_EMODE = -3
JAM_EDEVICE = -4
JAM_ETOOBIG = -5
JAM_ESTALE = -6
MAX_TEMPLATE_ATOMS = 32

_SYMBOLS = (
    "jam_context_create", "jam_context_free", "jam_templates_create", "jam_templates_free",
    "jam_molecules_create", "jam_molecules_free", "jam_molecules_parse_pdb",
    "jam_molecules_sizes", "jam_molecules_copy_atoms", "jam_molecules_residue_index", "jam_query", "jam_result_n_events",
    "jam_result_stride", "jam_result_n_pairs", "jam_result_copy", "jam_result_pair_matches",
    "jam_result_timings", "jam_result_pair_digests", "jam_result_join_kernels", "jam_context_set_profiling", "jam_result_level_partials", "jam_rmsd_forward",
    "jam_rmsd_backward", "jam_result_free", "jam_superpose_batch",
    "jam_superpose_batch_noalloc", "jam_kd_order", "jam_result_keep", "jam_events_copy", "jam_events_free",
    "jam_last_error",
    "jam_version", "jam_device_available", "jam_synchronize",
)


class NativeError(RuntimeError):
    """An error reported by the native library."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (code {code})")
        self.code = code


class UnsupportedModeError(NativeError, NotImplementedError):
    """A template uses a match mode that the device path does not implement."""


class MoleculeDesc(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int32),
        ("xyz", ctypes.c_void_p),
        ("names", ctypes.c_void_p),
        ("resnames", ctypes.c_void_p),
        ("chain2", ctypes.c_void_p),
        ("resseq", ctypes.c_void_p),
        ("serial", ctypes.c_void_p),
    ]


class TemplateDesc(ctypes.Structure):
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


class Params(ctypes.Structure):
    _fields_ = [
        ("rmsd_threshold", ctypes.c_double),
        ("distance_cutoff", ctypes.c_double),
        ("max_dynamic_distance", ctypes.c_double),
        ("ignore_chain", ctypes.c_int32),
        ("reorder", ctypes.c_int32),
    ]


_lib = None
_lib_lock = threading.Lock()


def library_path() -> str:
    return _build.LIB_PATH


def _one_hip_runtime() -> None:
    """PyTorch for ROCm bundles its own HIP runtime under the same soname as
    the system's (libamdhip64.so.7); whichever is loaded first serves the
    whole process, and torch cannot run on the system one. When torch is
    installed it is therefore imported before the library is opened, so that
    the library and torch (the differentiable RMSD, user code) share one
    runtime whatever the import order."""
    if "torch" not in sys.modules and importlib.util.find_spec("torch") is not None:
        import torch  # noqa: F401


def lib():
    """Loads ``libjess_amd.so``; raises if it has not been built."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        path = library_path()
        if not os.path.exists(path):
            raise ImportError(
                f"native library {path} is missing: build it with `python pyjess_amd/_build.py`"
            )
        _one_hip_runtime()
        L = ctypes.CDLL(path)
        vp, i32, i64, P = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER
        L.jam_context_create.argtypes = [i32, P(vp)]
        L.jam_context_create.restype = i32
        L.jam_context_free.argtypes = [vp]
        L.jam_context_free.restype = None
        L.jam_templates_create.argtypes = [vp, P(TemplateDesc), P(vp)]
        L.jam_templates_create.restype = i32
        L.jam_templates_free.argtypes = [vp]
        L.jam_templates_free.restype = None
        L.jam_molecules_create.argtypes = [vp, P(MoleculeDesc), i32, P(vp)]
        L.jam_molecules_create.restype = i32
        L.jam_molecules_free.argtypes = [vp]
        L.jam_molecules_free.restype = None
        L.jam_query.argtypes = [vp, vp, vp, P(Params), P(vp)]
        L.jam_query.restype = i32
        L.jam_result_n_events.argtypes = [vp]
        L.jam_result_n_events.restype = i64
        L.jam_result_stride.argtypes = [vp]
        L.jam_result_stride.restype = i32
        L.jam_result_n_pairs.argtypes = [vp]
        L.jam_result_n_pairs.restype = i32
        L.jam_result_copy.argtypes = [vp] * 9
        L.jam_result_copy.restype = i32
        L.jam_result_pair_matches.argtypes = [vp, vp]
        L.jam_result_pair_matches.restype = i32
        L.jam_result_timings.argtypes = [vp, vp]
        L.jam_result_timings.restype = i32
        L.jam_result_join_kernels.argtypes = [vp, vp, vp]
        L.jam_result_join_kernels.restype = i32
        L.jam_result_level_partials.argtypes = [vp, vp]
        L.jam_result_level_partials.restype = i32
        L.jam_rmsd_forward.argtypes = [vp, vp, vp, i64, i64, vp, vp, vp]
        L.jam_rmsd_forward.restype = i32
        L.jam_rmsd_backward.argtypes = [vp, vp, vp, i64, i64, vp, vp, vp, vp, vp, vp]
        L.jam_rmsd_backward.restype = i32
        L.jam_context_set_profiling.argtypes = [vp, i32]
        L.jam_context_set_profiling.restype = i32
        L.jam_result_free.argtypes = [vp]
        L.jam_result_free.restype = None
        L.jam_superpose_batch.argtypes = [vp, vp, vp, i32, i64, vp, vp, vp, vp]
        L.jam_superpose_batch.restype = i32
        L.jam_superpose_batch_noalloc.argtypes = [vp, vp, vp, i32, i64, vp, vp, vp, vp]
        L.jam_superpose_batch_noalloc.restype = i32
        L.jam_result_pair_digests.argtypes = [vp, i32, i32, vp, vp]
        L.jam_result_pair_digests.restype = i32
        L.jam_molecules_parse_pdb.argtypes = [vp, P(ctypes.c_char_p), vp, i32, i32, P(vp)]
        L.jam_molecules_parse_pdb.restype = i32
        L.jam_molecules_sizes.argtypes = [vp, vp, vp]
        L.jam_molecules_sizes.restype = i32
        L.jam_molecules_copy_atoms.argtypes = [vp, P(AtomFields)]
        L.jam_molecules_copy_atoms.restype = i32
        L.jam_molecules_residue_index.argtypes = [vp, vp, vp, vp, vp, vp]
        L.jam_molecules_residue_index.restype = i32
        L.jam_result_keep.argtypes = [vp, P(vp)]
        L.jam_result_keep.restype = i32
        L.jam_events_copy.argtypes = [vp] * 9
        L.jam_events_copy.restype = i32
        L.jam_events_free.argtypes = [vp]
        L.jam_events_free.restype = None
        L.jam_kd_order.argtypes = [vp, vp, vp, i32, vp]
        L.jam_kd_order.restype = i32
        L.jam_last_error.argtypes = []
        L.jam_last_error.restype = ctypes.c_char_p
        L.jam_version.argtypes = []
        L.jam_version.restype = ctypes.c_char_p
        L.jam_device_available.argtypes = []
        L.jam_device_available.restype = i32
        L.jam_synchronize.argtypes = [vp]
        L.jam_synchronize.restype = i32
        _lib = L
        return L


def _check(rc: int) -> None:
    if rc != JAM_OK:
        msg = lib().jam_last_error().decode("utf-8", "replace")
        if rc == JAM_EMODE:
            raise UnsupportedModeError(rc, msg)
        if rc == JAM_ENOMEM:
            raise MemoryError(msg)
        raise NativeError(rc, msg)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------

class Context:
    """One device context (stream + workspace). Queries on one context are serialised."""

    def __init__(self, device: int = 0):
        L = lib()
        h = ctypes.c_void_p()
        _check(L.jam_context_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device
        self.lock = threading.Lock()

    def set_profiling(self, on: bool) -> None:
        """Records HIP events around every join kernel launch (Result.join_kernel_ms)."""
        _check(lib().jam_context_set_profiling(self.handle, 1 if on else 0))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            _lib.jam_context_free(h)
            self.handle = None


_contexts = {}
_ctx_lock = threading.Lock()


def context(device: int = 0) -> Context:
    with _ctx_lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = _contexts[device] = Context(device)
        return ctx


# ---------------------------------------------------------------------------
# templates and molecules
# ---------------------------------------------------------------------------

class TemplateSet:
    """Device copy of a list of templates."""

    def __init__(self, ctx: Context, templates: Sequence):
        off, mode, resseq, chain2, pos, weight = [0], [], [], [], [], []
        name_off, names, rn_off, rns = [0], [], [0], []
        for t in templates:
            if len(t) > MAX_TEMPLATE_ATOMS:
                raise ValueError(f"templates of more than {MAX_TEMPLATE_ATOMS} atoms are not supported")
            for a in t:
                mode.append(a._mode)
                resseq.append(a._resseq)
                chain2.append(a._chain2)
                pos.append((a.x, a.y, a.z))
                weight.append(a._weight)
                names.extend(a._names)
                name_off.append(len(names))
                rns.extend(a._resnames)
                rn_off.append(len(rns))
            off.append(len(mode))
        self.n_templates = len(off) - 1
        arrays = dict(
            atom_off=np.array(off, dtype=np.int32),
            mode=np.array(mode or [0], dtype=np.int32),
            resseq=np.array(resseq or [0], dtype=np.int32),
            chain2=np.array(chain2 or [0], dtype=np.uint8),
            pos=np.array(pos or [(0.0, 0.0, 0.0)], dtype=np.float64),
            weight=np.array(weight or [0.0], dtype=np.float64),
            name_off=np.array(name_off, dtype=np.int32),
            names=np.frombuffer(b"".join(names) or b"\0\0\0\0", dtype=np.uint8).copy(),
            resname_off=np.array(rn_off, dtype=np.int32),
            resnames=np.frombuffer(b"".join(rns) or b"\0\0\0\0", dtype=np.uint8).copy(),
        )
        desc = TemplateDesc(self.n_templates, *(_ptr(arrays[k]) for k in (
            "atom_off", "mode", "resseq", "chain2", "pos", "weight", "name_off", "names",
            "resname_off", "resnames")))
        h = ctypes.c_void_p()
        _check(lib().jam_templates_create(ctx.handle, ctypes.byref(desc), ctypes.byref(h)))
        self.ctx = ctx
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            _lib.jam_templates_free(h)
            self.handle = None


class AtomFields(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "xyz", "names", "resnames", "chains", "resseq", "serial", "altloc", "icode", "occupancy", "tempfactor",
        "segid", "element", "charge")]


JAM_PDB_SKIP_HETATM = 1
JAM_PDB_IGNORE_ENDMDL = 2


class MoleculeSet:
    """Device copy of a batch of molecules."""

    def __init__(self, ctx: Context, molecules: Sequence):
        descs = (MoleculeDesc * max(1, len(molecules)))()
        keep = []
        for i, m in enumerate(molecules):
            arrs = (
                np.ascontiguousarray(m.xyz, dtype=np.float64),
                np.ascontiguousarray(m.names, dtype=np.uint8),
                np.ascontiguousarray(m.resnames, dtype=np.uint8),
                np.ascontiguousarray(m.chains[:, 1], dtype=np.uint8),
                np.ascontiguousarray(m.resseq, dtype=np.int32),
                np.ascontiguousarray(m.serial, dtype=np.int32),
            )
            keep.append(arrs)
            descs[i] = MoleculeDesc(len(m), *(_ptr(a) for a in arrs))
        h = ctypes.c_void_p()
        _check(lib().jam_molecules_create(ctx.handle, descs, len(molecules), ctypes.byref(h)))
        self.ctx = ctx
        self.handle = h
        self.n_molecules = len(molecules)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib is not None:
            _lib.jam_molecules_free(h)
            self.handle = None

    def atom_offsets(self) -> np.ndarray:
        off = np.zeros(self.n_molecules + 1, dtype=np.int64)
        _check(lib().jam_molecules_sizes(self.handle, _ptr(off), None))
        return off

    def residue_index(self):
        """(perm, res_off, res_key, res_begin, res_end) of the batch."""
        nb = ctypes.c_int64(0)
        off = self.atom_offsets()
        _check(lib().jam_molecules_sizes(self.handle, None, ctypes.byref(nb)))
        perm = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
        res_off = np.zeros(self.n_molecules + 1, dtype=np.int32)
        key = np.zeros(max(nb.value, 1), dtype=np.uint32)
        begin = np.zeros(max(nb.value, 1), dtype=np.int32)
        end = np.zeros(max(nb.value, 1), dtype=np.int32)
        _check(lib().jam_molecules_residue_index(self.handle, _ptr(perm), _ptr(res_off), _ptr(key), _ptr(begin),
                                                 _ptr(end)))
        return perm[:off[-1]], res_off, key[:nb.value], begin[:nb.value], end[:nb.value]


class ParsedMolecules(MoleculeSet):
    """A batch of molecules parsed from PDB texts on the device
    (jam_molecules_parse_pdb): molecule k of the batch is text k."""

    def __init__(self, ctx: Context, texts: Sequence[bytes], skip_hetatm: bool = False,
                 ignore_endmdl: bool = False):
        # the library reads every text where it lies (c_char_p points into
        # the bytes object), during the call
        texts = [t if isinstance(t, bytes) else bytes(t) for t in texts]
        ptrs = (ctypes.c_char_p * max(len(texts), 1))(*texts)
        lengths = np.array([len(t) for t in texts] or [0], dtype=np.int64)
        flags = (JAM_PDB_SKIP_HETATM if skip_hetatm else 0) | (JAM_PDB_IGNORE_ENDMDL if ignore_endmdl else 0)
        h = ctypes.c_void_p()
        with ctx.lock:
            rc = lib().jam_molecules_parse_pdb(ctx.handle, ptrs, _ptr(lengths), len(texts), flags, ctypes.byref(h))
        if rc == JAM_EINVAL:
            raise ValueError(lib().jam_last_error().decode("utf-8", "replace"))
        _check(rc)
        self.ctx = ctx
        self.handle = h
        self.n_molecules = len(texts)

    def arrays(self) -> dict:
        """Every atom field of the batch, as numpy arrays (all molecules)."""
        n = int(self.atom_offsets()[-1])
        m = max(n, 1)
        a = {
            "xyz": np.empty((m, 3), dtype=np.float64),
            "names": np.empty((m, 4), dtype=np.uint8),
            "resnames": np.empty((m, 4), dtype=np.uint8),
            "chains": np.empty((m, 2), dtype=np.uint8),
            "resseq": np.empty(m, dtype=np.int32),
            "serial": np.empty(m, dtype=np.int32),
            "altloc": np.empty(m, dtype=np.uint8),
            "icode": np.empty(m, dtype=np.uint8),
            "occupancy": np.empty(m, dtype=np.float32),
            "tempfactor": np.empty(m, dtype=np.float32),
            "segid": np.empty((m, 4), dtype=np.uint8),
            "element": np.empty((m, 2), dtype=np.uint8),
            "charge": np.empty(m, dtype=np.int32),
        }
        f = AtomFields(*(_ptr(a[k]) for k in ("xyz", "names", "resnames", "chains", "resseq", "serial", "altloc",
                                              "icode", "occupancy", "tempfactor", "segid", "element", "charge")))
        with self.ctx.lock:
            _check(lib().jam_molecules_copy_atoms(self.handle, ctypes.byref(f)))
        return {k: v[:n] for k, v in a.items()}


class KeptEvents:
    """A device copy of a query's events (jam_result_keep), which later
    queries on the context leave alone; its device memory is released when
    the object goes."""

    def __init__(self, ctx: "Context", handle, n: int, stride: int):
        L = lib()
        h = ctypes.c_void_p()
        _check(L.jam_result_keep(handle, ctypes.byref(h)))
        self.ctx = ctx
        self.handle = h
        self.n = n
        self.stride = stride

    def ranks(self):
        """The rank of every event, copied to host memory."""
        rank = np.empty(max(self.n, 1), dtype=np.int64)
        with self.ctx.lock:
            _check(lib().jam_events_copy(self.handle, None, _ptr(rank), None, None, None, None, None, None))
        return rank[:self.n]

    def heavy(self):
        """(rotation, c0, c1, atoms) of every event, copied to host memory."""
        m = max(self.n, 1)
        rot = np.empty((m, 9), dtype=np.float64)
        c0 = np.empty((m, 3), dtype=np.float64)
        c1 = np.empty((m, 3), dtype=np.float64)
        atoms = np.empty((m, max(self.stride, 1)), dtype=np.int32)
        with self.ctx.lock:
            _check(lib().jam_events_copy(self.handle, None, None, None, None, _ptr(rot), _ptr(c0), _ptr(c1),
                                         _ptr(atoms)))
        n = self.n
        return rot[:n], c0[:n], c1[:n], atoms[:n]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.jam_events_free(h)
            self.handle = None


class LazyBlock:
    """The rotation, centroids and matched atoms of a batch's events, left in
    their device copy until a hit first reads one of them; then all are
    copied at once and the device copy is released."""

    __slots__ = ("_kept", "_arrays", "_lock")

    def __init__(self, kept: KeptEvents):
        self._kept = kept
        self._arrays = None
        self._lock = threading.Lock()

    def _load(self):
        if self._arrays is None:
            with self._lock:
                if self._arrays is None:
                    self._arrays = self._kept.heavy()
                    self._kept = None
        return self._arrays

    @property
    def rotation(self):
        return self._load()[0]

    @property
    def c0(self):
        return self._load()[1]

    @property
    def c1(self):
        return self._load()[2]

    @property
    def atoms(self):
        return self._load()[3]


class Result:
    """Events of one device query, as numpy arrays. With ``keep``, only the
    fields every hit list reads (pair, flag, rmsd) are copied now; the others
    stay in a device copy of the events (`LazyBlock`), the ranks (which only
    a policy that can drop events reads) until `ranks` is called."""

    __slots__ = ("pair", "rank", "flag", "rmsd", "rotation", "c0", "c1", "atoms", "pair_matches",
                 "timings_ms", "join_kernel_ms", "join_launches", "level_partials", "n_pairs", "stride", "kept")

    def __init__(self, handle, ctx: Optional["Context"] = None, keep: bool = False):
        L = lib()
        n = L.jam_result_n_events(handle)
        self.stride = L.jam_result_stride(handle)
        self.n_pairs = L.jam_result_n_pairs(handle)
        m = max(n, 1)
        # every row is written by the copy
        self.pair = np.empty(m, dtype=np.int32)
        self.flag = np.empty(m, dtype=np.int32)
        self.rmsd = np.empty(m, dtype=np.float64)
        self.kept = None
        if keep:
            self.kept = KeptEvents(ctx, handle, n, self.stride)
            self.rotation = self.c0 = self.c1 = self.atoms = None
            self.rank = None
            _check(L.jam_result_copy(handle, _ptr(self.pair), None, _ptr(self.flag), _ptr(self.rmsd), None, None,
                                     None, None))
            names = ("pair", "flag", "rmsd")
        else:
            self.rank = np.empty(m, dtype=np.int64)
            self.rotation = np.empty((m, 9), dtype=np.float64)
            self.c0 = np.empty((m, 3), dtype=np.float64)
            self.c1 = np.empty((m, 3), dtype=np.float64)
            self.atoms = np.empty((m, max(self.stride, 1)), dtype=np.int32)
            _check(L.jam_result_copy(handle, _ptr(self.pair), _ptr(self.rank), _ptr(self.flag),
                                     _ptr(self.rmsd), _ptr(self.rotation), _ptr(self.c0), _ptr(self.c1),
                                     _ptr(self.atoms)))
            names = ("pair", "rank", "flag", "rmsd", "rotation", "c0", "c1", "atoms")
        for name in names:
            setattr(self, name, getattr(self, name)[:n])
        self.pair_matches = np.zeros(max(self.n_pairs, 1), dtype=np.int64)
        _check(L.jam_result_pair_matches(handle, _ptr(self.pair_matches)))
        self.pair_matches = self.pair_matches[:self.n_pairs]
        self.timings_ms = np.zeros(5, dtype=np.float64)
        _check(L.jam_result_timings(handle, _ptr(self.timings_ms)))
        ms = ctypes.c_double(0.0)
        nl = ctypes.c_int32(0)
        _check(L.jam_result_join_kernels(handle, ctypes.byref(ms), ctypes.byref(nl)))
        self.join_kernel_ms = ms.value       # summed k_join durations (HIP events)
        self.join_launches = nl.value
        self.level_partials = np.zeros(32, dtype=np.int64)   # partials expanded per join level
        _check(L.jam_result_level_partials(handle, _ptr(self.level_partials)))

    def ranks(self) -> np.ndarray:
        """The events' ranks (from the device copy when they were not copied
        with the other fields; read them before the copy is handed on)."""
        if self.rank is None:
            if self.kept is None:
                raise RuntimeError("the ranks were not copied and the device copy is gone")
            self.rank = self.kept.ranks()
        return self.rank


def query(tset: TemplateSet, mset: MoleculeSet, rmsd_threshold: float, distance_cutoff: float,
          max_dynamic_distance: float, ignore_chain: int, reorder: bool, keep: bool = False) -> Result:
    """One `jam_query`; with ``keep`` the rotations, centroids and atoms stay
    in a device copy until read (`Result`)."""
    L = lib()
    ctx = tset.ctx
    p = Params(float(rmsd_threshold), float(distance_cutoff), float(max_dynamic_distance),
               int(ignore_chain), int(bool(reorder)))
    h = ctypes.c_void_p()
    with ctx.lock:
        _check(L.jam_query(ctx.handle, tset.handle, mset.handle, ctypes.byref(p), ctypes.byref(h)))
        try:
            # the events live in the context's device buffers until the next query
            return Result(h, ctx, keep)
        finally:
            L.jam_result_free(h)


def superpose_batch(a: np.ndarray, b: np.ndarray, device: int = 0):
    """Batched superposition of point sets a, b with shape (n_sets, k, 3)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape or a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("expected two arrays of shape (n_sets, k, 3)")
    n, k, _ = a.shape
    rmsd = np.zeros(max(n, 1))
    rot = np.zeros((max(n, 1), 9))
    c0 = np.zeros((max(n, 1), 3))
    c1 = np.zeros((max(n, 1), 3))
    ctx = context(device)
    with ctx.lock:
        _check(lib().jam_superpose_batch(ctx.handle, _ptr(a), _ptr(b), int(k), int(n), _ptr(rmsd),
                                         _ptr(rot), _ptr(c0), _ptr(c1)))
    return rmsd[:n], rot[:n].reshape(n, 3, 3), c0[:n], c1[:n]


def kd_order(xyz: np.ndarray, off: np.ndarray, device: int = 0) -> np.ndarray:
    """Leaf orders (left subtree first) of the kd-trees the reference builds
    over point sets: set s is xyz[off[s]:off[s+1]]; element off[s] + i of the
    result is the set-local index of the point at leaf i."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    off = np.ascontiguousarray(off, dtype=np.int64)
    if off.ndim != 1 or off.shape[0] < 1 or off[0] != 0 or off[-1] != xyz.shape[0]:
        raise ValueError("off must start at 0 and end at the number of points")
    out = np.zeros(max(xyz.shape[0], 1), dtype=np.int32)
    ctx = context(device)
    with ctx.lock:
        _check(lib().jam_kd_order(ctx.handle, _ptr(xyz), _ptr(off), int(off.shape[0] - 1), _ptr(out)))
    return out[:xyz.shape[0]]


def synchronize(device: int = 0) -> None:
    """Waits for all work on the device (bracketing timed regions)."""
    _check(lib().jam_synchronize(context(device).handle))


def device_available() -> bool:
    return bool(lib().jam_device_available())


def version() -> str:
    return lib().jam_version().decode()
