"""One assembly of the library-scale workload at its real size -- 100,000+
atoms in 20 chains, templates of 2-8 residues with symmetric mode-1 atom
pairs, 2.5 A cutoff -- against the oracle in order, under each policy of
test_library_style. Every tenth of the workload's 7,000 templates keeps the
oracle's replay to seconds."""

import warnings

import numpy as np
import pytest

from oracle import oracle
from pyjess_amd import Jess, synthetic

from test_gpu_parity import assert_same_hits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def assembly():
    mols, tpls, kw = synthetic.workload("library-64", seed=0, scale=1 / 64)
    mol = mols[0]
    assert len(mol) >= 100_000
    return mol, tpls[::10], kw


_streams = {}


def _stream(mol, tpls, kw, ignore_chain):
    if ignore_chain not in _streams:
        _streams[ignore_chain] = oracle.query_stream(mol, tpls, kw["distance_cutoff"], kw["max_dynamic_distance"],
                                                     ignore_chain)
    return _streams[ignore_chain]


@pytest.mark.parametrize("ignore_chain,best_match,max_candidates",
                         [(None, False, 100000), ("residues", False, 3), (None, True, 100000), ("atoms", True, 2)])
def test_library_assembly_in_order(assembly, ignore_chain, best_match, max_candidates):
    mol, tpls, kw = assembly
    print(f"library assembly: {len(mol)} atoms x {len(tpls)} templates, ignore_chain={ignore_chain}, "
          f"best_match={best_match}, max_candidates={max_candidates}", flush=True)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        got = list(Jess(tpls).query(mol, kw["rmsd_threshold"], kw["distance_cutoff"], kw["max_dynamic_distance"],
                                    max_candidates=max_candidates, ignore_chain=ignore_chain,
                                    best_match=best_match))
    stream = _stream(mol, tpls, kw, ignore_chain)
    want = oracle.replay(stream, kw["rmsd_threshold"], best_match, max_candidates)[0]
    assert len(want) > 100
    assert_same_hits(mol, tpls, got, want)
