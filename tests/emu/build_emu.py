"""Build tests/emu/libugks_emu.so: the HIP library's sources compiled for the
CPU against tests/emu/hip_emu.h.  TEST INFRASTRUCTURE ONLY."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "ugks2d_high_order_amd", "csrc")
LIB = os.path.join(HERE, "libugks_emu.so")
SRCS = ["ugks_operator.hip", "ugks_implicit.hip", "ugks_explicit.hip", "ugks_adjoint.hip", "ugks_api.hip"]


def build(force=False):
    deps = [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(HERE, "hip_emu.h")]
    if not force and os.path.exists(LIB) and all(os.path.getmtime(d) <= os.path.getmtime(LIB) for d in deps):
        return LIB
    cmd = ["g++", "-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-pthread", "-include",
           os.path.join(HERE, "hip_emu.h"), "-I", os.path.join(ROOT, "include"), "-x", "c++"]
    cmd += [os.path.join(CSRC, s) for s in SRCS] + ["-o", LIB]
    subprocess.check_call(cmd)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
