# cython: language_level=3, boundscheck=False, wraparound=False
"""`Hit`, the hit lists (eager for `Query`, lazy for `Jess.query_batch`) and
the iteration of `Query`, compiled (reference ``src/pyjess/_jess.pyx:2090-2621``).

A query returns thousands of hits; each is a small object that refers to a
row of its query's result arrays (rotation, centroids, matched atoms), read
when asked for. Building them here keeps the per-hit cost of the drop-in API
to an allocation.
"""

cimport cython

import io
import math
import warnings

import numpy as np

from ._model import Atom, Mat4, Molecule, Template, _apply

# TessTemplate_logE (patches/TessTemplate.c.patch:164-170):
# logE = logA + alpha * rmsd + beta * dim + ln(n). logA is visible in the
# reference; alpha and beta are fixed by the 156 logE values of the
# 1AMY+1.3.3 fixture (the only pair consistent with all of them at their
# printed precision).
LOG_A = -8.5
ALPHA = 2.5
BETA = -0.7


class Block:
    """Row-major result arrays shared by the hits of one query."""

    __slots__ = ("rotation", "c0", "c1", "atoms")

    def __init__(self, rotation, c0, c1, atoms):
        self.rotation = rotation
        self.c0 = c0
        self.c1 = c1
        self.atoms = atoms


@cython.freelist(8192)
@cython.no_gc
cdef class Hit:
    """A hit between a query molecule and a template (``_jess.pyx:2286-2621``).

    Hits are not tracked by the cyclic garbage collector: a hit refers only to
    its result block, template and molecule, none of which refers back to
    hits, and a query can return millions of them (tracking them makes the
    collector's passes over young objects the dominant cost of building the
    list)."""

    cdef readonly double rmsd
    cdef object _block
    cdef Py_ssize_t _row
    cdef object _template
    cdef object _molecule

    def __init__(self, rmsd, rotation, c0, c1, atoms, template, molecule):
        self.rmsd = float(rmsd)
        self._block = Block(np.asarray(rotation, dtype=np.float64).reshape(1, 9),
                            np.asarray(c0, dtype=np.float64).reshape(1, 3),
                            np.asarray(c1, dtype=np.float64).reshape(1, 3),
                            np.asarray(list(atoms), dtype=np.int64).reshape(1, -1))
        self._row = 0
        self._template = template
        self._molecule = molecule

    @property
    def _rotation(self):
        return self._block.rotation[self._row].tolist()

    @property
    def _centre(self):
        return self._block.c0[self._row].tolist(), self._block.c1[self._row].tolist()

    @property
    def _index(self):
        return self._block.atoms[self._row, :len(self._template)].tolist()

    @property
    def _atoms(self):
        mol = self._molecule
        return [mol._atom(k) for k in self._index]

    def __reduce__(self):
        c0, c1 = self._centre
        return Hit, (self.rmsd, self._rotation, c0, c1, list(self._index), self._template, self._molecule)

    def __repr__(self):
        return f"<Hit rmsd={self.rmsd!r} template={self._template.id!r}>"

    @property
    def transformation(self):
        """Matrix taking molecule coordinates into the template frame (``_jess.pyx:2343-2377``)."""
        M = self._rotation
        c, v = self._centre
        Mc = np.eye(4)
        Mc[0:3, 3] = [-c[0], -c[1], -c[2]]
        Mr = np.eye(4)
        Mr[0:3, 0:3] = np.array(M).reshape(3, 3)
        Mv = np.eye(4)
        Mv[0:3, 3] = v
        return Mat4._wrap(Mv) @ Mat4._wrap(Mr) @ Mat4._wrap(Mc)

    @property
    def inverse_transformation(self):
        """Matrix taking template coordinates into the molecule frame (``_jess.pyx:2379-2413``)."""
        M = self._rotation
        c, v = self._centre
        Mv = np.eye(4)
        Mv[0:3, 3] = [-v[0], -v[1], -v[2]]
        Mr = np.eye(4)
        Mr[0:3, 0:3] = np.array(M).reshape(3, 3).T
        Mc = np.eye(4)
        Mc[0:3, 3] = c
        return Mat4._wrap(Mc) @ Mat4._wrap(Mr) @ Mat4._wrap(Mv)

    @property
    def determinant(self):
        p = self._rotation
        det = 0.0
        det += p[0] * (p[4] * p[8] - p[5] * p[7])
        det -= p[1] * (p[3] * p[8] - p[5] * p[6])
        det += p[2] * (p[3] * p[7] - p[4] * p[6])
        return det

    @property
    def log_evalue(self):
        n = len(self._molecule)
        return LOG_A + ALPHA * self.rmsd + BETA * float(self._template.dimension) + math.log(float(n))

    @property
    def evalue(self):
        return math.exp(self.log_evalue)

    def atoms(self, transform=True):
        if not transform:
            return self._atoms
        trans = self.transformation
        return [a.transform(trans) for a in self._atoms]

    def molecule(self, transform=False):
        if not transform:
            return self._molecule
        return self._molecule.transform(self.transformation)

    def template(self, transform=False):
        if not transform:
            return self._template
        return self._template.transform(self.inverse_transformation)

    def dumps(self, format="pdb", transform=True):
        out = io.StringIO()
        self.dump(out, format=format, transform=transform)
        return out.getvalue()

    def dump(self, file, format="pdb", transform=True):
        if format != "pdb":
            raise ValueError(f"invalid value for `format` argument: {format!r}")
        if self._template.id is None:
            raise RuntimeError("cannot dump `Hit` where `self.template.id` is `None`")
        if self._molecule.id is None:
            raise RuntimeError("cannot dump `Hit` where `self.molecule.id` is `None`")
        file.write("REMARK ")
        file.write(self._molecule.id)
        file.write(f" {self.rmsd:5.3f} ")
        file.write(self._template.id)
        file.write(f" Det={self.determinant:4,.1f} log(E)~ {self.log_evalue:4.2f}\n")
        mat = [float(v) for v in np.asarray(self.transformation).reshape(16)]
        for a in self._atoms:
            if transform:
                x, y, z = _apply(mat, a.x, a.y, a.z)
            else:
                x, y, z = a.x, a.y, a.z
            file.write(a._line(x, y, z))
        file.write("ENDMDL\n")


def build_hits(list out, const long long[:] chosen, const int[:] mol_of, const int[:] tpl_of,
               const int[:] flag, const double[:] rmsd, object block, list templates, list molecules,
               bint keep_none=True):
    """Appends to out[molecule] the hit of every chosen event, in the given
    order. A NaN rotation appends None when ``keep_none`` (the iterator warns
    when it reaches it) and is skipped otherwise. Returns the number of NaN
    rotations met per molecule."""
    cdef Py_ssize_t i, e, n = chosen.shape[0]
    cdef int mi
    cdef Hit h
    cdef list lst
    nan = np.zeros(len(out), dtype=np.int64)
    cdef long long[:] nn = nan
    for i in range(n):
        e = chosen[i]
        mi = mol_of[e]
        lst = <list>out[mi]
        if flag[e] == 2:
            nn[mi] += 1
            if keep_none:
                lst.append(None)
            continue
        h = Hit.__new__(Hit)
        h.rmsd = rmsd[e]
        h._block = block
        h._row = e
        h._template = templates[tpl_of[e]]
        h._molecule = molecules[mi]
        lst.append(h)
    return nan


cdef class HitList:
    """The hits of one molecule of `Jess.query_batch`, in pyjess order: a
    read-only sequence over rows of the batch's shared result arrays (a list
    of rows, or a contiguous range of them). A `Hit` object is made when an
    element is read (a library-scale batch has tens of millions of hits, most
    of which a caller never touches); reading the same element twice gives two
    equal-valued `Hit` objects."""

    cdef object _block
    cdef const double[:] _rmsd
    cdef const int[:] _pair
    cdef const long long[:] _events   # the rows, unless _range
    cdef bint _range                   # rows _start ... _start + _n - 1
    cdef Py_ssize_t _start
    cdef Py_ssize_t _n
    cdef int _n_templates
    cdef list _templates
    cdef object _molecule

    def __init__(self, block, const double[:] rmsd, const int[:] pair, const long long[:] events,
                 int n_templates, list templates, molecule):
        self._block = block
        self._rmsd = rmsd
        self._pair = pair
        self._events = events
        self._range = False
        self._start = 0
        self._n = events.shape[0]
        self._n_templates = n_templates
        self._templates = templates
        self._molecule = molecule

    cdef inline Py_ssize_t _row(self, Py_ssize_t k):
        return self._start + k if self._range else self._events[k]

    cdef Hit _hit(self, Py_ssize_t k):
        cdef Py_ssize_t e = self._row(k)
        cdef Hit h = Hit.__new__(Hit)
        h.rmsd = self._rmsd[e]
        h._block = self._block
        h._row = e
        h._template = self._templates[self._pair[e] % self._n_templates]
        h._molecule = self._molecule
        return h

    def __len__(self):
        return self._n

    def __getitem__(self, index):
        cdef Py_ssize_t n = self._n, k
        if isinstance(index, slice):
            return [self._hit(k) for k in range(*index.indices(n))]
        k = index
        if k < 0:
            k += n
        if k < 0 or k >= n:
            raise IndexError("hit index out of range")
        return self._hit(k)

    def __iter__(self):
        return _HitListIter(self)

    def __bool__(self):
        return self._n > 0

    def __eq__(self, other):
        try:
            n = len(other)
        except TypeError:
            return NotImplemented
        if n != len(self):
            return False
        cdef HitList o
        cdef Py_ssize_t k
        if isinstance(other, HitList):
            # the same rows of the same result arrays
            o = <HitList>other
            if o._block is not self._block:
                return False
            for k in range(n):
                if o._row(k) != self._row(k):
                    return False
            return True
        return all(a == b for a, b in zip(self, other))

    def __repr__(self):
        return f"<HitList of {len(self)} hits>"


def hit_lists(block, const double[:] rmsd, const int[:] pair, const long long[:] events,
              const long long[:] bounds, int n_templates, list templates, list molecules):
    """One `HitList` per molecule, molecule m holding events[bounds[m]:bounds[m + 1]]
    (the result arrays are taken once for all the lists)."""
    cdef Py_ssize_t m, n = len(molecules)
    cdef HitList h
    cdef list out = []
    if bounds.shape[0] < n + 1:
        raise ValueError("bounds must hold one more entry than there are molecules")
    for m in range(n):
        if not (0 <= bounds[m] <= bounds[m + 1] <= events.shape[0]):
            raise ValueError("bounds out of range")
        h = HitList.__new__(HitList)
        h._block = block
        h._rmsd = rmsd
        h._pair = pair
        h._events = events[bounds[m]:bounds[m + 1]]
        h._range = False
        h._start = 0
        h._n = bounds[m + 1] - bounds[m]
        h._n_templates = n_templates
        h._templates = templates
        h._molecule = molecules[m]
        out.append(h)
    return out


cdef Py_ssize_t _lower_bound(const long long[:] v, Py_ssize_t lo, Py_ssize_t hi, long long x):
    """First index in [lo, hi) of the sorted v whose value is >= x."""
    cdef Py_ssize_t mid
    while lo < hi:
        mid = lo + (hi - lo) // 2
        if v[mid] < x:
            lo = mid + 1
        else:
            hi = mid
    return lo


@cython.cdivision(True)
def split_ranges(const int[:] pair, const int[:] flag, int n_templates, int n_mols):
    """Every event's molecule range when no policy drops events: the events
    come grouped by molecule in order (pair // n_templates never decreases),
    so molecule m's events are bounds[m]:bounds[m+1], found by bisection
    (the first and last event of every range are checked to belong to it;
    events out of that order elsewhere are not detected). Also the
    NaN-rotation events (flag 2, which are left out of the hit lists), in
    order, and their number per molecule."""
    cdef Py_ssize_t n = pair.shape[0], i, lo, hi, mid, k = 0, cap = 256
    cdef int m
    if n_templates <= 0:
        raise ValueError("no templates")
    bounds = np.zeros(n_mols + 1, dtype=np.int64)
    cdef long long[:] b = bounds
    for m in range(n_mols + 1):
        lo, hi = 0, n
        while lo < hi:
            mid = lo + (hi - lo) // 2
            if pair[mid] // n_templates < m:
                lo = mid + 1
            else:
                hi = mid
        b[m] = lo
    b[n_mols] = n
    if n > 0 and b[0] != 0:
        raise ValueError("events are not grouped by molecule")
    for m in range(n_mols):
        if b[m + 1] > b[m] and (pair[b[m]] // n_templates != m or pair[b[m + 1] - 1] // n_templates != m):
            raise ValueError("events are not grouped by molecule")
    nan_ev = np.empty(cap, dtype=np.int64)
    cdef long long[:] ne = nan_ev
    for i in range(n):
        if flag[i] == 2:
            if k == cap:
                cap *= 2
                nan_ev = np.resize(nan_ev, cap)
                ne = nan_ev
            ne[k] = i
            k += 1
    nan_ev = nan_ev[:k]
    ne = nan_ev
    nan = np.zeros(n_mols, dtype=np.int64)
    cdef long long[:] nn = nan
    for m in range(n_mols):
        nn[m] = _lower_bound(ne, 0, k, b[m + 1]) - _lower_bound(ne, 0, k, b[m])
    return bounds, nan, nan_ev


@cython.cdivision(True)
def pair_ranges(const int[:] pair, int n_templates, const long long[:] pairs):
    """[starts[j], ends[j]): the events of pair pairs[j] in a stream ordered
    by molecule, then template in scan order (reverse index): the key
    pair - t + (n_templates - 1 - t), t = pair % n_templates, never decreases,
    so each pair's events are found by bisection (an empty range when it has
    none; ValueError when the ends of a range belong to another pair)."""
    cdef Py_ssize_t n = pair.shape[0], j, lo, hi, mid, k = pairs.shape[0]
    cdef long long target, key, p
    cdef int t
    if n_templates <= 0:
        raise ValueError("no templates")
    starts = np.zeros(k, dtype=np.int64)
    ends = np.zeros(k, dtype=np.int64)
    cdef long long[:] st = starts
    cdef long long[:] en = ends
    for j in range(k):
        p = pairs[j]
        if p < 0:
            raise ValueError("negative pair")
        target = p - p % n_templates + (n_templates - 1 - p % n_templates)
        lo, hi = 0, n
        while lo < hi:
            mid = lo + (hi - lo) // 2
            t = pair[mid] % n_templates
            key = pair[mid] - t + (n_templates - 1 - t)
            if key < target:
                lo = mid + 1
            else:
                hi = mid
        st[j] = lo
        hi = n
        while lo < hi:
            mid = lo + (hi - lo) // 2
            t = pair[mid] % n_templates
            key = pair[mid] - t + (n_templates - 1 - t)
            if key <= target:
                lo = mid + 1
            else:
                hi = mid
        en[j] = lo
        if en[j] > st[j] and (pair[st[j]] != p or pair[en[j] - 1] != p):
            raise ValueError("events are not in report order")
    return starts, ends


def hit_list_ranges(block, const double[:] rmsd, const int[:] pair, const long long[:] bounds,
                    const long long[:] skip, int n_templates, list templates, list molecules):
    """One `HitList` per molecule holding its events bounds[m]:bounds[m + 1]
    except those in ``skip`` (sorted: the NaN-rotation events, and any hit a
    policy drops): a range of rows when it has none of them, else the list of
    its other rows."""
    cdef Py_ssize_t m, n = len(molecules), b0, b1, j0, j1, e, w, k = skip.shape[0]
    cdef HitList h
    cdef list out = []
    cdef long long[:] r
    if bounds.shape[0] < n + 1:
        raise ValueError("bounds must hold one more entry than there are molecules")
    for m in range(n):
        b0, b1 = bounds[m], bounds[m + 1]
        if not (0 <= b0 <= b1 <= rmsd.shape[0] and b1 <= pair.shape[0]):
            raise ValueError("bounds out of range")
        j0 = _lower_bound(skip, 0, k, b0)
        j1 = _lower_bound(skip, j0, k, b1)
        h = HitList.__new__(HitList)
        h._block = block
        h._rmsd = rmsd
        h._pair = pair
        h._n_templates = n_templates
        h._templates = templates
        h._molecule = molecules[m]
        if j1 == j0:
            h._range = True
            h._start = b0
            h._n = b1 - b0
        else:
            rows = np.empty(b1 - b0 - (j1 - j0), dtype=np.int64)
            r = rows
            w = 0
            for e in range(b0, b1):
                if j0 < j1 and skip[j0] == e:
                    j0 += 1
                    continue
                r[w] = e
                w += 1
            h._events = rows
            h._range = False
            h._start = 0
            h._n = w
        out.append(h)
    return out


cdef class _HitListIter:
    cdef HitList _list
    cdef Py_ssize_t _k

    def __init__(self, HitList lst):
        self._list = lst
        self._k = 0

    def __iter__(self):
        return self

    def __next__(self):
        if self._k >= self._list._n:
            raise StopIteration
        self._k += 1
        return self._list._hit(self._k - 1)


def split_events(const int[:] pair, const int[:] flag, chosen, int n_templates, int n_mols):
    """The chosen events (all of them when ``chosen`` is None) cut by molecule
    for `HitList`: the hit events in order, each molecule's range of them
    (``bounds[m]:bounds[m+1]``) and its number of NaN-rotation events, which
    are left out. One pass; events arrive grouped by molecule."""
    cdef Py_ssize_t n = pair.shape[0] if chosen is None else len(chosen)
    cdef const long long[:] ch
    if chosen is not None:
        ch = chosen
    ev = np.empty(max(n, 1), dtype=np.int64)
    bounds = np.zeros(n_mols + 1, dtype=np.int64)
    nan = np.zeros(n_mols, dtype=np.int64)
    cdef long long[:] o = ev
    cdef long long[:] b = bounds
    cdef long long[:] nn = nan
    cdef Py_ssize_t i, w = 0
    cdef long long e
    cdef int m, cur = 0
    for i in range(n):
        e = i if chosen is None else ch[i]
        m = pair[e] // n_templates
        while cur < m:
            cur += 1
            b[cur] = w
        if flag[e] == 2:
            nn[m] += 1
            continue
        o[w] = e
        w += 1
    while cur < n_mols:
        cur += 1
        b[cur] = w
    return ev[:w], bounds, nan


def select_events(const int[:] pair, const long long[:] rank, const int[:] flag, const double[:] rmsd,
                  bint best_match, long long mc):
    """Indices of the events that ``Query.__next__`` reaches
    (``_jess.pyx:2238-2276``), in order: the hits it yields and the
    NaN-rotation matches it warns about.

    The events arrive grouped by pair in report order, ranks increasing
    within a pair. The candidate counter restarts at every template and
    after every yielded hit; when it exceeds ``mc`` (-1: no limit) the rest of
    the template is skipped after the current match. With ``best_match``,
    the first match of least RMSD (strict ``<``) is kept per template."""
    cdef Py_ssize_t n = pair.shape[0], i = 0, j, e, w = 0, best
    out = np.empty(max(n, 1), dtype=np.int64)
    cdef long long[:] o = out
    cdef int p
    cdef long long rk, c, seg
    cdef double best_rmsd
    while i < n:
        p = pair[i]
        j = i
        while j < n and pair[j] == p:
            j += 1
        if best_match:
            best = -1
            best_rmsd = math.inf
            for e in range(i, j):
                rk = rank[e]
                if mc != -1 and rk > mc + 1:
                    break
                if rmsd[e] < best_rmsd:
                    if flag[e] == 2:
                        o[w] = e        # warning only
                        w += 1
                    else:
                        best = e
                        best_rmsd = rmsd[e]
                if mc != -1 and rk == mc + 1:
                    break
            if best >= 0:
                o[w] = best
                w += 1
        else:
            seg = 0
            for e in range(i, j):
                c = rank[e] - seg
                if mc != -1 and c > mc + 1:
                    break
                o[w] = e
                w += 1
                if flag[e] != 2:
                    seg = rank[e] + 1
                if mc != -1 and c == mc + 1:
                    break
        i = j
    return out[:w]


cdef class HitIterator:
    """Iteration of `Query` over its hits in pyjess order; NaN-rotation
    entries (None) become warnings, as in ``Query.__next__``
    (``_jess.pyx:2239-2251``). Subclasses provide ``_run()``, which returns the
    hit list the first time the iterator is advanced."""

    cdef public object _hits
    cdef public Py_ssize_t _pos

    def __iter__(self):
        return self

    def __next__(self):
        if self._hits is None:
            self._hits = self._run()
        cdef list hits = self._hits
        cdef Py_ssize_t n = len(hits)
        while self._pos < n:
            item = hits[self._pos]
            self._pos += 1
            if item is None:
                warnings.warn("Jess returned a superposition matrix with NaN values", UserWarning, stacklevel=2)
                continue
            return item
        raise StopIteration
