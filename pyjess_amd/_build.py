"""Compiles the package's native code in-tree: ``libjess_amd.so`` (hipcc,
gfx950) and the ``_hits`` extension (Cython, then the host C compiler).

Run ``python pyjess_amd/_build.py`` (or ``__graft_entry__.build()``). Both are
written next to this file so that they travel with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB_NAME = "libjess_amd.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["jam_api.hip"]
HEADERS = ["jam_common.h", "jam_grad.h", "jam_kd.h", "jam_kernels.h", "jam_super.h", "jam_wave.h"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build pyjess_amd")


def flags() -> list:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # FP64 decisions must match the CPU path bit for bit: no contraction
        # into fused multiply-add, no fast-math
        "-ffp-contract=off",
        "-fno-fast-math",
        # packed FP32 instructions are kept out of the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
    ]


def build_hits(verbose: bool = False) -> str:
    """Cythonize ``_hits.pyx`` and compile it into an extension module."""
    from Cython.Build import cythonize

    pyx = os.path.join(HERE, "_hits.pyx")
    c_file = os.path.join(HERE, "_hits.c")
    cythonize([pyx], language_level=3, quiet=not verbose, force=True)
    out = os.path.join(HERE, "_hits" + sysconfig.get_config_var("EXT_SUFFIX"))
    cc = os.environ.get("CC") or shutil.which("gcc") or "cc"
    cmd = [cc, "-O2", "-shared", "-fPIC", "-fno-strict-aliasing", f"-I{sysconfig.get_paths()['include']}",
           c_file, "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"building the _hits extension failed with exit code {proc.returncode}")
    os.replace(out + ".tmp", out)
    return out


def build(verbose: bool = False) -> str:
    """Compile every source into ``libjess_amd.so`` and the ``_hits``
    extension; returns the library's path."""
    build_hits(verbose)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc()] + flags() + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        sys.stderr.write(proc.stderr)
        raise RuntimeError(f"hipcc failed with exit code {proc.returncode}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
