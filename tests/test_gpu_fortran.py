"""A Fortran host drives the HIP step through the iso_c_binding module.

ugks2d_high_order_amd/fortran/couette_driver (built by build_fortran with
amdflang) runs the shipped Couette case (8x64 cells, the reference's 32x32
velocity table, omega 0.5) from the cold start for 3 iterations and prints
the residual norms of each accepted state.  They must equal the oracle's
single-rank norms to 1e-11 relative (macro components against the largest
component), the same tolerance as the Python-host parity tests.
"""
import os
import subprocess

import numpy as np
import pytest

from ugks2d_high_order_amd import build as B, config

pytestmark = pytest.mark.gpu


def test_fortran_host_matches_oracle(oracle_mod, release, tmp_path):
    if not os.path.exists(B.DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    case = config.shipped_case()
    tab = tmp_path / "table.bin"
    np.concatenate([release["u"], release["v"], release["w"]]).astype("<f8").tofile(tab)
    n_it = 3
    print("fortran driver: shipped case 8x64x32^2, 3 iterations", flush=True)
    out = subprocess.run([B.DRIVER, str(tab), str(case.nx), str(case.ny), str(case.nu), str(case.nv),
                          repr(case.omega), repr(case.mu_ref), str(n_it)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [l.split() for l in out.stdout.splitlines() if l.startswith("iter")]
    assert len(rows) == n_it
    o = oracle_mod.OracleSolver(case, release["u"], release["v"], release["w"], threads=8)
    o.cold_start(1.0, 0.0, case.t_init)
    for it in range(n_it):
        st, acc, d = o.iteration()
        assert st == 0 and acc
        o.operator()
        n = o.norms()
        f = np.array([float(x) for x in rows[it][2:]])
        mscale = np.max(np.abs(n[0:4]))
        assert np.max(np.abs(f[0:4] - n[0:4])) <= 1e-11 * mscale, (it, f[0:4], n[0:4])
        assert np.max(np.abs(f[4:6] - n[12:14]) / n[12:14]) <= 1e-11, (it, f[4:6], n[12:14])
        assert abs(f[6] - d[2]) <= 1e-13, (it, f[6], d[2])
