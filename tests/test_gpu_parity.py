"""Parity of the HIP library (through its C ABI) with the oracle.

Cases: a reduced Couette grid (8x16 cells x 16x16 velocities, the shipped
Ma3 physics) for speed, and the shipped case (8x64 x 32x32).  Tolerances are
normwise relative (max |gpu - oracle| / max |oracle| per field):
  * states W, h, b and observables after every pseudo-time step: 1e-12;
  * face fluxes: 1e-12; streaming+collision rates: 1e-10 (they are flux
    differences over dx*dt, which amplifies last-bit differences of the
    fluxes by the ratio |F|/|dF|).
"""
import numpy as np
import pytest

from ugks2d_high_order_amd import config, velocity

pytestmark = pytest.mark.gpu


def _nrel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _gpu():
    import ugks2d_high_order_amd as U
    return U


def _pair(case, oracle_mod, table=None):
    U = _gpu()
    if table is None:
        u, v, w = velocity.table(case.table, case.table_scale)
    else:
        u, v, w = table
    g = U.UGKS(case, u, v, w)
    o = oracle_mod.OracleSolver(case, u, v, w, threads=8)
    return g, o


def test_cold_start_matches_oracle(oracle_mod):
    case = config.small_case()
    g, o = _pair(case, oracle_mod)
    g.cold_start()
    o.cold_start(1.0, 0.0, case.t_init)
    Wg, hg, bg = g.state()
    Wo, ho, bo = o.state()
    assert _nrel(Wg, Wo) < 1e-15
    assert _nrel(hg, ho) < 1e-13 and _nrel(bg, bo) < 1e-13


def _compare_operator(g, o, label):
    og = g.evolution()
    o.operator()
    oo = o.operator_outputs()
    assert abs(og["dt"] - oo["dt"]) <= 1e-15 * oo["dt"], label
    errs = {
        "xflux_h": _nrel(og["xflux_h"], oo["xflux_h"]), "xflux_b": _nrel(og["xflux_b"], oo["xflux_b"]),
        "yflux_h": _nrel(og["yflux_h"], oo["yflux_h"]), "yflux_b": _nrel(og["yflux_b"], oo["yflux_b"]),
        "rate_h": _nrel(og["rate_h"], oo["stream_h"] + oo["coll_h"]),
        "rate_b": _nrel(og["rate_b"], oo["stream_b"] + oo["coll_b"]),
        "rate_w": _nrel(og["rate_w"], oo["rate_w"]),
    }
    print(label, errs, flush=True)
    for k in ("xflux_h", "xflux_b", "yflux_h", "yflux_b"):
        assert errs[k] < 1e-12, (label, k, errs[k])
    for k in ("rate_h", "rate_b", "rate_w"):
        assert errs[k] < 1e-10, (label, k, errs[k])
    ng, _wg = g.norms()
    no = o.norms()
    # macro residual components against the largest component (a component
    # of a uniform state is round-off), the h/b norms against themselves
    mscale = max(float(np.max(np.abs(no[0:4]))), 1e-300)
    for idx in (0, 1, 2, 3):
        assert abs(ng[idx] - no[idx]) <= 1e-11 * mscale, (label, idx, ng[idx], no[idx])
    for idx in (12, 13, 16, 17):
        assert abs(ng[idx] - no[idx]) <= 1e-11 * abs(no[idx]), (label, idx, ng[idx], no[idx])


def test_operator_parity_initial_and_nonuniform(oracle_mod):
    case = config.small_case()
    g, o = _pair(case, oracle_mod)
    o.cold_start(1.0, 0.0, case.t_init)
    W, h, b = o.state()
    g.set_state(W, h, b)
    _compare_operator(g, o, "initial")
    # a non-uniform state: the oracle's state after one iteration
    st, acc, _ = o.iteration()
    assert st == 0 and acc
    W, h, b = o.state()
    g.set_state(W, h, b)
    _compare_operator(g, o, "after one step")


def test_iterations_match_oracle(oracle_mod):
    case = config.small_case()
    g, o = _pair(case, oracle_mod)
    o.cold_start(1.0, 0.0, case.t_init)
    g.set_state(*o.state())
    for it in range(5):
        acc_g, dg = g.implicit_step()
        st, acc_o, do = o.iteration()
        assert acc_g and acc_o and st == 0
        Wg, hg, bg = g.state()
        Wo, ho, bo = o.state()
        e = (_nrel(Wg, Wo), _nrel(hg, ho), _nrel(bg, bo))
        fg = g.observables()
        fo, _ = o.observables()
        ef = [_nrel(fg[:, q], fo[:, q]) for q in range(10)]
        print("iter", it + 1, "W/h/b", e, "fields", max(ef), "scale-1", dg[2], do[2], flush=True)
        assert max(e) < 1e-12, (it, e)
        assert max(ef) < 1e-12, (it, ef)
        assert abs(dg[2] - do[2]) < 1e-13


def test_shipped_case_initial_norms_and_two_steps(oracle_mod, release):
    case = config.shipped_case()
    g, o = _pair(case, oracle_mod, table=(release["u"], release["v"], release["w"]))
    g.cold_start()
    g.evolution(outputs=False)
    n, walls = g.norms()
    meta = release["meta"]
    assert abs(n[16] / meta["reference_h"] - 1) < 1e-13
    assert abs(n[17] / meta["reference_b"] - 1) < 1e-13
    o.cold_start(1.0, 0.0, case.t_init)
    for it in range(2):
        acc, _ = g.implicit_step()
        st, acc_o, _ = o.iteration()
        assert acc and acc_o
        Wg, hg, bg = g.state()
        Wo, ho, bo = o.state()
        e = (_nrel(Wg, Wo), _nrel(hg, ho), _nrel(bg, bo))
        print("shipped iter", it + 1, e, flush=True)
        assert max(e) < 1e-12


@pytest.mark.parametrize("nx,ny", [(64, 6), (72, 6)])
def test_pr2_paths_match_oracle(oracle_mod, nx, ny):
    """PR(2) runs in one wavefront (lane per column) up to 64 columns and in
    one workgroup beyond; both follow the oracle's hyperplane Gauss-Seidel."""
    case = config.small_case(nx=nx, ny=ny)
    g, o = _pair(case, oracle_mod)
    o.cold_start(1.0, 0.0, case.t_init)
    g.set_state(*o.state())
    print("pr2 path", nx, "x", ny, flush=True)
    for it in range(3):
        acc_g, dg = g.implicit_step()
        st, acc_o, do = o.iteration()
        assert acc_g and acc_o and st == 0
        Wg, hg, bg = g.state()
        Wo, ho, bo = o.state()
        e = (_nrel(Wg, Wo), _nrel(hg, ho), _nrel(bg, bo))
        print("iter", it + 1, e, flush=True)
        assert max(e) < 1e-12, (it, e)
