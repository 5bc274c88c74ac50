"""Build the gfx950 HIP library libugks_hip.so in-tree with hipcc.

Usage: python -m ugks2d_high_order_amd.build [--force]
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libugks_hip.so")
SOURCES = ["ugks_operator.hip", "ugks_implicit.hip", "ugks_explicit.hip", "ugks_adjoint.hip", "ugks_api.hip"]
HEADERS = ["ugks_device.h", "ugks_recon.h", "ugks_state.h", "ugks_prof.h", "ugks_node.h", "ugks_adjoint.h"]
FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "couette_driver")
ARCH = os.environ.get("UGKS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.isabs(c) and os.path.exists(c):
            return c
    return "hipcc"


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
        # packed FP32 instructions are disabled for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(HERE, "..", "include"),
        "-Wno-unused-result",
    ] + [f"-D{d}" for d in os.environ.get("UGKS_DEFINES", "").split()]


def _stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    deps = [os.path.join(CSRC, s) for s in SOURCES + HEADERS] + [os.path.join(HERE, "..", "include", "ugks_hip.h")]
    return any(os.path.getmtime(d) > t for d in deps)


def build(force: bool = True, verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 into libugks_hip.so (always when force)."""
    if not force and not _stale():
        return LIB
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    procs = []
    for s in SOURCES:
        o = os.path.join(build_dir, s.replace(".hip", ".o"))
        cmd = [hipcc(), "-c", os.path.join(CSRC, s), "-o", o] + flags()
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), s))
        objs.append(o)
    failed = []
    for p, s in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((s, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError("hipcc failed:\n" + msg)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def flang() -> str | None:
    for c in ("/opt/rocm/lib/llvm/bin/amdflang", "/opt/rocm/llvm/bin/amdflang"):
        if os.path.exists(c):
            return c
    return None


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran binding module and the Couette driver against
    libugks_hip.so (skipped when no Fortran compiler is installed)."""
    fc = flang()
    if fc is None:
        return None
    bdir = os.path.join(FORTRAN, "_build")
    os.makedirs(bdir, exist_ok=True)
    mod_o = os.path.join(bdir, "ugks_hip_binding.o")
    cmds = [
        [fc, "-O2", "-c", os.path.join(FORTRAN, "ugks_hip_binding.f90"), "-o", mod_o, "-module-dir", bdir],
        [fc, "-O2", "-I", bdir, os.path.join(FORTRAN, "couette_driver.f90"), mod_o, "-L", HERE, "-lugks_hip",
         "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    return DRIVER


if __name__ == "__main__":
    build(force=True, verbose="-v" in sys.argv)
    print(LIB)
    d = build_fortran(verbose="-v" in sys.argv)
    print(d or "no Fortran compiler: driver not built")
