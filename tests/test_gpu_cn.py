"""S2CN2 explicit step of the HIP library (ugks_cn_step) against the oracle.

Starting from a non-equilibrium state (three HighOrderImplicit oracle
iterations from the cold start), three S2CN2 steps on the reduced grid and
two on the shipped case must agree with the oracle's steps: states W, h, b
to 1e-12 normwise relative, the CN and moment gate quantities below the
reference gate 2e-12 on both sides, and the conservation ledger (totals and
cumulative wall balance) to 1e-12 relative to max(1, |value|).
"""
import numpy as np
import pytest

from ugks2d_high_order_amd import config, velocity

pytestmark = pytest.mark.gpu


def _nrel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _run(case, tab, oracle_mod, n_steps):
    import ugks2d_high_order_amd as U
    o = oracle_mod.OracleSolver(case, *tab, threads=8)
    o.cold_start(1.0, 0.0, case.t_init)
    for _ in range(3):
        st, acc, _ = o.iteration()
        assert st == 0 and acc
    S = o.state()
    g = U.UGKS(case, *tab)
    g.set_state(*S)
    for it in range(n_steps):
        acc_g, dg = g.cn_step()
        st, acc_o, do = o.cn_step()
        assert st == 0 and acc_o and acc_g, (it, dg, do)
        e = [_nrel(x, y) for x, y in zip(g.state(), o.state())]
        print("cn step", it + 1, "W/h/b", e, "cnres", dg[2], do[2], "drift", dg[12], do[12], flush=True)
        assert max(e) < 1e-12, (it, e)
        assert dg[2] <= 2e-12 and dg[3] <= 2e-12 and dg[12] <= 2e-12
        assert abs(dg[13] - do[13]) <= 1e-15 * do[13]
    tg, bg = g.cn_ledger()
    to, bo = o.cn_ledger()
    assert np.max(np.abs(tg - to) / np.maximum(1, np.abs(to))) < 1e-12
    assert np.max(np.abs(bg - bo) / np.maximum(1, np.abs(bo))) < 1e-12


def test_cn_steps_match_oracle_reduced(oracle_mod):
    case = config.small_case()
    _run(case, velocity.table(case.table, case.table_scale), oracle_mod, 3)


def test_cn_steps_match_oracle_shipped(oracle_mod, release):
    case = config.shipped_case()
    _run(case, (release["u"], release["v"], release["w"]), oracle_mod, 2)
