"""Shared fixtures. Tests marked ``gpu`` need a gfx950 GPU and the built library."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libjess_amd.so")


def golden(name: str) -> str:
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def mol_1amy():
    from pyjess_amd import Molecule
    return Molecule.load(golden("1AMY.pdb"))


@pytest.fixture(scope="session")
def mol_5ayx():
    from pyjess_amd import Molecule
    return Molecule.load(golden("5ayx.EF.pdb"))


@pytest.fixture(scope="session")
def tpl_133():
    from pyjess_amd import Template
    return Template.load(golden("1.3.3.tpl"))


@pytest.fixture(scope="session")
def tpl_412():
    from pyjess_amd import Template
    return Template.load(golden("4.1.2.tpl"))


def parse_hit_file(path: str):
    """Blocks of a Jess output file: (header fields, atom lines)."""
    with open(path) as f:
        blocks = [b for b in f.read().split("REMARK") if b.strip()]
    out = []
    for b in blocks:
        lines = b.strip().splitlines()
        out.append((lines[0].split(), lines[1:-1]))
    return out


# Module doctest of the reference (src/pyjess/_jess.pyx:19-70): hits of
# Jess([1.3.3, 4.1.2]).query(1AMY, 2, 2, 2), as printed (prefixes).
DOCTEST_RMSD = ["1.4386", "1.4877", "1.4376", "1.5284", "1.4863", "1.4369", "1.4790", "1.1414",
                "1.0755", "1.1973", "1.1353", "1.0711", "1.1494"]
# First hit of the Jess doctest (src/pyjess/_jess.pyx:2650-2660)
DOCTEST_DUMP = """REMARK 1AMY 1.439 2om2 Det= 1.0 log(E)~ 1.11
ATOM    729  CA  THR A  94      34.202 -24.426   8.851  1.00  2.00           C
ATOM    732  CB  THR A  94      35.157 -23.467   8.101  1.00  4.66           C
ATOM    733  OG1 THR A  94      36.338 -23.247   8.871  1.00  9.85           O
ATOM    746  CD  GLU A  96      41.454 -29.509   8.013  1.00 24.05           C
ATOM    748  OE2 GLU A  96      42.536 -29.680   7.441  1.00 34.44           O
ATOM    747  OE1 GLU A  96      41.212 -28.521   8.708  1.00 18.56           O
ATOM    437  CZ  ARG A  55      44.471 -26.619  10.181  1.00  8.51           C
ATOM    436  NE  ARG A  55      44.334 -27.346  11.290  1.00  9.05           N
ATOM    438  NH1 ARG A  55      43.590 -26.751   9.179  1.00 13.17           N
ENDMDL
"""
