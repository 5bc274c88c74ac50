"""CPU check of the Explicit-entry kernels' logic: the library's own HIP
sources, compiled for the CPU against a thread emulation of workgroups,
barriers and LDS (tests/emu), against the oracle on the reduced grid
(8x16 cells x 16^2 velocities) with workgroups of 64 and 128 threads.
Emulated device memory starts as NaN, so a read of anything no kernel wrote
shows up.  States agree to 1e-13 normwise relative over two steps, dt to
1e-15 and the residual res(1:4) to 1e-8 (see test_gpu_explicit for why
res is looser)."""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def test_emulated_explicit_kernels_match_oracle():
    out = subprocess.run([sys.executable, os.path.join(HERE, "emu", "run_explicit.py")], capture_output=True,
                         text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    r = json.loads(out.stdout.strip().splitlines()[-1])
    for wg, v in r.items():
        assert max(v["state"]) < 1e-13, (wg, v)
        assert max(v["res_dt"][1::2]) < 1e-15, (wg, v)
        assert max(v["res_dt"][0::2]) < 1e-8, (wg, v)
