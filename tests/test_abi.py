"""The C ABI library loads and exports every function include/jess_amd.h declares.
No compute call: this runs without a GPU."""

import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT


def declared_functions():
    with open(os.path.join(ROOT, "include", "jess_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jam_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for required in ("jam_context_create", "jam_templates_create", "jam_molecules_create", "jam_query",
                     "jam_result_copy", "jam_superpose_batch", "jam_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from pyjess_amd import _native
    path = _native.library_path()
    assert os.path.exists(path), "build the library first: python pyjess_amd/_build.py"
    lib = ctypes.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (jam_[a-z0-9_]+)", out))
    assert set(declared_functions()) <= exported
    assert set(_native._SYMBOLS) == set(declared_functions())


def test_library_is_built_for_gfx950():
    from pyjess_amd import _native
    out = subprocess.run(["strings", _native.library_path()], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_version_and_errors_without_device():
    from pyjess_amd import _native
    L = _native.lib()
    assert L.jam_version().decode().startswith("pyjess-amd")
    h = ctypes.c_void_p()
    # invalid arguments are reported, not crashed on
    assert L.jam_query(None, None, None, None, ctypes.byref(h)) == _native.JAM_EINVAL
    assert L.jam_last_error()
