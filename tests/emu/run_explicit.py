"""Runs the low-order Explicit entry of the emulated library (the HIP sources
compiled for the CPU) against the oracle on the reduced grid and prints one
JSON line of normwise relative differences.  TEST INFRASTRUCTURE ONLY: a
separate process, so the emulated library never replaces the HIP library in
a test process."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import build_emu  # noqa: E402
import oracle  # noqa: E402
import ugks2d_high_order_amd as U  # noqa: E402
from ugks2d_high_order_amd import config, velocity  # noqa: E402


def nrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def main():
    U.load_library(build_emu.build())
    oracle.build()
    case = config.small_case()
    tab = velocity.table(case.table, case.table_scale)
    out = {}
    for wg in (64, 128):
        g = U.UGKS(case, *tab)
        g.set_tuning(U.UGKS.TUNE_EXPLICIT_WG, wg)
        o = oracle.OracleSolver(case, *tab, threads=4)
        g.explicit_cold_start()
        o.explicit_cold_start()
        errs = [max(nrel(x, y) for x, y in zip(g.state(), o.state()))]
        res = []
        for _ in range(2):
            rg, dtg = g.explicit_step()
            st, ro, dto = o.explicit_step()
            assert st == 0
            errs.append(max(nrel(x, y) for x, y in zip(g.state(), o.state())))
            res.append(float(np.max(np.abs(rg - ro) / np.maximum(np.abs(ro), 1e-300))))
            res.append(abs(dtg / dto - 1))
        out[str(wg)] = dict(state=errs, res_dt=res)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
