"""CPU check of the reverse S2CN2 step (ugks_cn_adjoint_device): the
library's own HIP sources compiled for the CPU against the thread emulation
(tests/emu), against the reverse-mode derivative of the FP64 torch
restatement of the step (s2cn2_torch, itself pinned to the oracle's step).

Case: 8x16 cells x 16^2 velocities, Ma 2 walls, Kn 0.1, omega 0.81; the state
is the discrete equilibrium times exp(0.5 N(0,1)) per (cell, velocity), so
the WENO5-AO weights are far from linear and the positivity limiter acts on
thousands of (cell, velocity) pairs; random adjoints of the next state and of
the wall fluxes; workgroups of 128 threads.  The adjoint of W, h, b agrees to
1e-12 normwise relative and the gradients with respect to mu_ref, Pr and the
wall speed to 1e-11 relative (measured: 1e-15..5e-13)."""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def test_emulated_reverse_step_matches_torch_reverse_mode():
    out = subprocess.run([sys.executable, os.path.join(HERE, "emu", "run_adjoint.py"), "--quick"],
                         capture_output=True, text=True, timeout=900)
    assert out.returncode == 0, out.stderr[-2000:]
    r = json.loads(out.stdout.strip().splitlines()[-1])
    for name, v in r.items():
        assert v["limited"] > 1000, (name, v)
        assert max(v["W"], v["h"], v["b"]) < 1e-12, (name, v)
        assert max(v["U"], v["mu"], v["pr"]) < 1e-11, (name, v)
