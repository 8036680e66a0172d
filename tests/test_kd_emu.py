"""The device kd kernels (pyjess_amd/csrc/jam_kd.h) run on the CPU under an
emulation of the HIP features they use (tests/emu/), built with
AddressSanitizer and UBSan, every buffer and every workgroup's LDS filled
with 0xFF bytes before the kernels run, in the order the library launches
them (bucketing, then per size tier the fast pass and the exact pass over the
tied sets, or the exact pass alone for the 257-1,024-entry tiers of a single
molecule). Their leaf orders must equal the reference QuickSelect's
(tests/golden/kd_ref_orders.npz, written with the reference's own
patches/qselect.h.patch), and no pass may reach its loop bound.

A workgroup runs as one host thread per lane, so large sets are slow: the
default selection covers every tie style at 9, 33, 65 and 129 entries, two
257-entry sets (tiers 0-3, fast and exact paths) and, in the batch launch
order, one 1,025-entry set (the LDS exact pass of 1,025-8,192 entries).
JAM_EMU_FULL=1 runs every set up to 8,192 entries in both launch orders (an
hour or more)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
BIN = os.path.join(EMU, "_build", "kd_emu")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")


@pytest.fixture(scope="module")
def emulator():
    proc = subprocess.run(["make", "-s", "-C", EMU], capture_output=True, text=True)
    if proc.returncode != 0:
        pytest.fail("building the kd emulator failed:\n" + proc.stdout + proc.stderr)
    return BIN


def vectors():
    z = np.load(golden("kd_ref_orders.npz"))
    return z["style"], z["off"], z["q"], z["order"]


def selection(full: bool, one_molecule: bool):
    style, off, _, _ = vectors()
    n = np.diff(off)
    if full:
        return [i for i in range(len(n)) if int(n[i]) <= 8192]
    sel = [i for i in range(len(n)) if int(n[i]) in (9, 33, 65, 129)]
    sel += [i for i in range(len(n)) if int(n[i]) == 257 and style[i] in ("int5", "dup_pairs")]
    if one_molecule:
        return [i for i in sel if int(n[i]) >= 129]
    sel += [i for i in range(len(n)) if int(n[i]) == 1025 and style[i] == "int5"]
    return sel


def run(binary, sets, one_molecule, tmp_path):
    off = np.zeros(len(sets) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(xyz) for xyz in sets])
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([len(sets), int(one_molecule)], dtype=np.int32).tofile(f)
        off.tofile(f)
        np.concatenate(sets).astype(np.float64).tofile(f)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    proc = subprocess.run([binary, str(src), str(dst)], capture_output=True, text=True, env=env, timeout=3600)
    assert proc.returncode == 0, proc.stderr[-4000:]
    out = np.fromfile(dst, dtype=np.int32)
    return [out[off[k]:off[k + 1]] for k in range(len(sets))], int(out[-1])


@pytest.mark.parametrize("one_molecule", [False, True])
def test_emulated_kd_kernels_give_the_reference_leaf_order(emulator, tmp_path, one_molecule):
    style, off, q, order = vectors()
    sel = selection(os.environ.get("JAM_EMU_FULL") == "1", one_molecule)
    sets = [q[off[i]:off[i + 1]].astype(np.float64) / 8.0 for i in sel]
    got, kd_err = run(emulator, sets, one_molecule, tmp_path)
    assert kd_err == 0
    bad = [(str(style[i]), int(off[i + 1] - off[i])) for i, g in zip(sel, got)
           if not np.array_equal(g, order[off[i]:off[i + 1]])]
    assert not bad, bad
