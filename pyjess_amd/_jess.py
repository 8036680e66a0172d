"""`Jess`, `Query` and `Hit`: the drop-in query API (reference ``src/pyjess/_jess.pyx:2090-2869``).

A query runs the whole search on the GPU in one call of the native library
(every template against the molecule), the first time the `Query` iterator is
advanced. The device returns "events": complete matches whose RMSD passes the
threshold (hits, and matches whose rotation has a NaN), each with its rank in
the reference engine's enumeration order. `Query` then applies the policies of
the reference's ``Query.__next__`` (``_jess.pyx:2197-2283``) over the events in
that order, so hits, their order, ``best_match`` and ``max_candidates`` behave
as in pyjess.
"""

from __future__ import annotations

import functools
import io
import warnings
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _native
from ._model import Atom, Mat4, Molecule, Template, _apply

__all__ = ["Jess", "Query", "Hit"]

_IGNORE = {None: 0, "residues": 1, "atoms": 2}
_IGNORE_NAMES = {0: None, 1: "residues", 2: "atoms"}


def _warn_nan():
    warnings.warn("Jess returned a superposition matrix with NaN values", UserWarning, stacklevel=3)


from ._hits import Block as _Block, Hit, HitIterator as _HitIterator, build_hits as _build_hits  # noqa: E402
from ._hits import HitList, hit_list_ranges as _hit_list_ranges, hit_lists as _hit_lists  # noqa: E402
from ._hits import select_events as _select_events, split_events as _split_events  # noqa: E402
from ._hits import pair_ranges as _pair_ranges, split_ranges as _split_ranges  # noqa: E402

import collections.abc as _abc  # noqa: E402

_abc.Sequence.register(HitList)


class Query(_HitIterator):
    """An iterator over the hits of a query (``_jess.pyx:2090-2283``)."""

    def __init__(self):
        raise TypeError("Query objects are created by Jess.query")

    @classmethod
    def _create(cls, jess: "Jess", molecule: Molecule, rmsd_threshold: float, distance_cutoff: float,
                max_dynamic_distance: float, max_candidates, ignore_chain, best_match: bool,
                reorder: bool) -> "Query":
        q = cls.__new__(cls)
        q.jess = jess
        q.molecule = molecule
        q.rmsd_threshold = float(rmsd_threshold)
        q.best_match = bool(best_match)
        q.max_candidates = max_candidates
        q.ignore_chain = ignore_chain
        q._distance_cutoff = float(distance_cutoff)
        q._max_dynamic_distance = float(max_dynamic_distance)
        q._reorder = bool(reorder)
        q._hits = None
        q._pos = 0
        return q

    @property
    def ignore_chain(self):
        return _IGNORE_NAMES[self._ignore_chain]

    @ignore_chain.setter
    def ignore_chain(self, ignore_chain):
        if ignore_chain not in _IGNORE:
            raise ValueError(f"invalid value for `ignore_chain`: {ignore_chain!r}")
        self._ignore_chain = _IGNORE[ignore_chain]

    @property
    def max_candidates(self):
        return None if self._max_candidates == -1 else self._max_candidates

    @max_candidates.setter
    def max_candidates(self, max_candidates):
        if max_candidates is None:
            self._max_candidates = -1
        elif max_candidates >= 0:
            self._max_candidates = int(max_candidates)
        else:
            raise ValueError(f"invalid value for `max_candidates` argument: {max_candidates!r}")

    def _run(self) -> list:
        if len(self.jess) == 0 or len(self.molecule) == 0:
            return []
        res = self.jess._device_query([self.molecule], self.rmsd_threshold, self._distance_cutoff,
                                      self._max_dynamic_distance, self._ignore_chain, self._reorder)
        return _policy(self.jess, [self.molecule], res, self.best_match, self._max_candidates)[0][0]


def _cut_events(res, pair, flag, rmsd, T: int, mc: int, nan_ev, nan):
    """The max_candidates cap without best_match, for lazy hit lists: the
    counter of ``Query.__next__`` counts the complete matches since the last
    yielded hit and can only pass ``mc + 1`` in a pair with more complete
    matches than that, so only those pairs go through `select_events`. Returns
    the events to leave out of the hit lists (every NaN-rotation event, and
    the hits the cap drops) and the NaN warnings per molecule (a NaN event
    the cap drops gives none)."""
    pairs = np.flatnonzero(np.asarray(res.pair_matches) > mc + 1).astype(np.int64)
    if len(pairs) == 0:
        return nan_ev, nan
    starts, ends = _pair_ranges(pair, T, pairs)
    if not np.any(ends > starts):
        return nan_ev, nan
    rank = np.ascontiguousarray(res.ranks() if hasattr(res, "ranks") else res.rank, dtype=np.int64)
    nan = nan.copy()
    leave = [nan_ev]
    for p, a, b in zip(pairs.tolist(), starts.tolist(), ends.tolist()):
        if a == b:
            continue
        kept = _select_events(pair[a:b], rank[a:b], flag[a:b], rmsd[a:b], False, mc)
        mask = np.ones(b - a, dtype=bool)
        mask[kept] = False
        dropped = np.flatnonzero(mask) + a
        nan[p // T] -= int(np.count_nonzero(flag[dropped] == 2))
        leave.append(dropped[flag[dropped] != 2])
    return np.unique(np.concatenate(leave)).astype(np.int64), nan


def _policy(jess: "Jess", molecules: Sequence[Molecule], res: "_native.Result", best_match: bool,
            max_candidates: int, keep_none: bool = True):
    """Hits per molecule, in pyjess order, from the device events.

    The device delivers the events in report order: by molecule, then
    template in scan order (reverse insertion order: ``Jess_addTemplate``
    prepends, ``patches/Jess.c.patch:63-72``), then rank, the position of the
    complete match in the reference enumeration. The candidate counter of
    ``Query.__next__`` restarts at every template and after every yielded
    hit; when it exceeds ``max_candidates`` the rest of the template is
    skipped (after the current match, which is still reported if it is a
    hit). NaN rotations produce ``None`` entries, turned into warnings when
    iterated (``keep_none``), or are dropped. Returns the hit lists and the
    number of NaN rotations per molecule.
    """
    T = len(jess)
    out: List[list] = [[] for _ in molecules]
    n = len(res.rmsd)
    if n == 0:
        return out, np.zeros(len(molecules), dtype=np.int64)
    mc = max_candidates
    pair = np.ascontiguousarray(res.pair, dtype=np.int32)
    # the candidate counter of a template never exceeds its number of complete
    # matches, so a cap at or above the largest one never cuts anything
    if mc != -1 and len(res.pair_matches) and mc >= int(res.pair_matches.max()):
        mc = -1
    # without best_match, lazy lists apply a candidate cap to the few pairs it
    # can cut, after cutting the events into molecules (_cut_events)
    cut_late = not best_match and mc != -1 and not keep_none
    if (not best_match and mc == -1) or cut_late:
        # no policy can drop an event: every event is reported, in order
        chosen = None if not keep_none else np.arange(n, dtype=np.int64)
    else:
        rank = res.ranks() if hasattr(res, "ranks") else res.rank
        chosen = _select_events(pair, np.ascontiguousarray(rank, dtype=np.int64),
                                np.ascontiguousarray(res.flag, dtype=np.int32),
                                np.ascontiguousarray(res.rmsd, dtype=np.float64), best_match, mc)
    if cut_late and hasattr(res, "ranks") and len(res.pair_matches) and int(np.max(res.pair_matches)) > mc + 1:
        res.ranks()   # the cap reads them (_cut_events): before the device copy is handed on
    if getattr(res, "kept", None) is not None:
        # the block owns the device copy from here (released once read)
        block = _native.LazyBlock(res.kept)
        res.kept = None
    else:
        block = _Block(res.rotation, res.c0, res.c1, res.atoms)
    if not keep_none:
        # lazy lists: the chosen hits (NaN rotations dropped and counted), cut
        # by molecule (events are grouped by molecule, in order); with every
        # event chosen, each molecule's hits are a range of the events
        flag = np.ascontiguousarray(res.flag, dtype=np.int32)
        rmsd = np.ascontiguousarray(res.rmsd, dtype=np.float64)
        if chosen is None:
            bounds, nan, skip = _split_ranges(pair, flag, T, len(molecules))
            if cut_late:
                skip, nan = _cut_events(res, pair, flag, rmsd, T, mc, skip, nan)
            out = _hit_list_ranges(block, rmsd, pair, bounds, skip, T, list(jess._templates), list(molecules))
            return out, nan
        hit_ev, bounds, nan = _split_events(pair, flag, chosen, T, len(molecules))
        out = _hit_lists(block, rmsd, pair, np.ascontiguousarray(hit_ev, dtype=np.int64),
                         np.ascontiguousarray(bounds, dtype=np.int64), T, list(jess._templates), list(molecules))
        return out, nan
    nan = _build_hits(out, chosen, (pair // T).astype(np.int32), (pair % T).astype(np.int32),
                      np.ascontiguousarray(res.flag, dtype=np.int32),
                      np.ascontiguousarray(res.rmsd, dtype=np.float64), block, list(jess._templates),
                      list(molecules), keep_none)
    return out, nan


class Jess:
    """A handle to run Jess over a list of templates (``_jess.pyx:2623-2869``)."""

    def __init__(self, templates: Iterable[Template] = ()):
        self._templates = tuple(templates)
        for t in self._templates:
            if not isinstance(t, Template):
                raise TypeError(f"expected Template, found {type(t).__name__}")
        self._tset = None
        self._last_result = None

    def __len__(self) -> int:
        return len(self._templates)

    def __bool__(self) -> bool:
        return len(self._templates) > 0

    def __getitem__(self, index):
        if isinstance(index, slice):
            return type(self)(self._templates[index])
        return self._templates[index]

    def __eq__(self, other):
        if not isinstance(other, Jess):
            return NotImplemented
        return self._templates == other._templates

    def __hash__(self):
        return hash((Jess, self._templates))

    def __reduce__(self):
        return type(self), (self._templates,)

    def __copy__(self):
        return self.copy()

    def copy(self) -> "Jess":
        return type(self)(self._templates)

    def _device_templates(self, ctx) -> "_native.TemplateSet":
        if self._tset is None or self._tset.ctx is not ctx:
            self._tset = _native.TemplateSet(ctx, self._templates)
        return self._tset

    def _device_query(self, molecules: Sequence[Molecule], rmsd_threshold, distance_cutoff,
                      max_dynamic_distance, ignore_chain: int, reorder: bool,
                      keep: bool = False) -> "_native.Result":
        ctx = _native.context()
        tset = self._device_templates(ctx)
        if len(molecules) == 1:
            mol = molecules[0]
            mset = mol._device.get(id(ctx))
            if mset is None:
                mset = mol._device[id(ctx)] = _native.MoleculeSet(ctx, [mol])
        else:
            mset = _batch_cache(ctx, molecules)
        res = _native.query(tset, mset, rmsd_threshold, distance_cutoff, max_dynamic_distance,
                            ignore_chain, reorder, keep)
        self._last_result = res     # kernel timings and match counts of the latest query
        return res

    def query(self, molecule: Molecule, rmsd_threshold: float, distance_cutoff: float,
              max_dynamic_distance: float, *, max_candidates=None, ignore_chain=None,
              best_match: bool = False, reorder: bool = True) -> Query:
        """Scan for templates matching the given molecule (``_jess.pyx:2758-2869``)."""
        ignore_chain = _normalize_ignore_chain(ignore_chain)
        if not isinstance(molecule, Molecule):
            raise TypeError(f"expected Molecule, found {type(molecule).__name__}")
        return Query._create(self, molecule, rmsd_threshold, distance_cutoff, max_dynamic_distance,
                             max_candidates, ignore_chain, best_match, reorder)

    def query_batch(self, molecules: Sequence[Molecule], rmsd_threshold: float, distance_cutoff: float,
                    max_dynamic_distance: float, *, max_candidates=None, ignore_chain=None,
                    best_match: bool = False, reorder: bool = True) -> List["HitList"]:
        """All hits of many molecules in one device pass.

        Equivalent to ``[list(self.query(m, ...)) for m in molecules]`` (the
        ThreadPool-over-molecules pattern of the reference README) but runs
        every (template, molecule) pair in a single launch chain. Each
        molecule's hits come as a `HitList`, a sequence that makes each `Hit`
        when it is read; the rotations, centroids and matched atoms of all
        hits are copied from the device when the first of them is read.
        """
        ignore_chain = _normalize_ignore_chain(ignore_chain)
        molecules = list(molecules)
        if max_candidates is None:
            mc = -1
        elif max_candidates >= 0:
            mc = int(max_candidates)
        else:
            raise ValueError(f"invalid value for `max_candidates` argument: {max_candidates!r}")
        if not molecules or len(self) == 0:
            return [[] for _ in molecules]
        # the hits' rotations, centroids and atoms stay in a device copy of
        # the events until one of them is read
        res = self._device_query(molecules, rmsd_threshold, distance_cutoff, max_dynamic_distance,
                                 _IGNORE[ignore_chain], reorder, keep=True)
        hits, nan = _policy(self, molecules, res, best_match, mc, keep_none=False)
        for _ in range(int(np.count_nonzero(nan))):
            _warn_nan()
        return hits


# (context, molecules, their data versions, device copy) of the latest batch
_last_batch = [None, None, None, None]


def _batch_cache(ctx, molecules: Sequence[Molecule]) -> "_native.MoleculeSet":
    """Device copy of a batch of molecules; the latest one is kept, so a batch
    queried again (against other templates, or repeatedly) is not re-uploaded.
    The cache holds the molecule objects, which keeps their identities valid,
    and the version of each one's data (bumped when an array is replaced)."""
    ctx0, mols0, vers0, mset0 = _last_batch
    vers = [m.__dict__.get("_version", 0) for m in molecules]
    if ctx0 is ctx and mols0 is not None and len(mols0) == len(molecules) and \
            all(a is b for a, b in zip(mols0, molecules)) and vers == vers0:
        return mset0
    mset = _native.MoleculeSet(ctx, molecules)
    _last_batch[:] = [ctx, list(molecules), vers, mset]
    return mset


def _normalize_ignore_chain(ignore_chain):
    if ignore_chain is True:
        warnings.warn(
            "`ignore_chain` parameter expects string parameters to specificy the mode since "
            "PyJess v0.7.0. Use `ignore_chain='atoms'` instead of `ignore_chain=True`",
            DeprecationWarning, stacklevel=3)
        return "atoms"
    if ignore_chain is False:
        warnings.warn(
            "`ignore_chain` parameter expects string parameters to specificy the mode since "
            "PyJess v0.7.0. Use `ignore_chain=None` instead of `ignore_chain=False`",
            DeprecationWarning, stacklevel=3)
        return None
    if ignore_chain not in _IGNORE:
        raise ValueError(f"invalid value for `ignore_chain`: {ignore_chain!r}")
    return ignore_chain
