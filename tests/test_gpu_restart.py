"""Complete restarts of the HighOrderImplicit run on the GPU.

  * 2 + 2 = 4: a run that writes the reference's restart files after two
    updates and a new instance that loads them and takes two more end in the
    same state, bit for bit, as four updates in one go (the reference's own
    restart-equivalence check: two + two complete restart updates byte-identical
    to four consecutive ones), with the same residual norms, mass target and
    correction ledger.
  * CPU -> GPU: restart files written from the oracle's state after two
    iterations of the shipped case (cold start, M0 of the cold state, the
    references of its first evaluation) are loaded by a GPU instance, whose
    next update matches the oracle's third to 1e-12 relative.
  * MassStatistics on the device equals a host evaluation of the same state,
    and the ledger equals the sum of the rescale increments.
"""
import math

import numpy as np
import pytest

from ugks2d_high_order_amd import config, restart as R, velocity

pytestmark = pytest.mark.gpu


def _nrel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _iterate(g, n):
    """n iterations of the Solve_UGKS loop: evaluate the accepted state, then update."""
    for _ in range(n):
        g.evolution(outputs=False)
        acc, _ = g.implicit_step()
        assert acc


def test_restart_two_plus_two_equals_four(tmp_path, release):
    import ugks2d_high_order_amd as U
    case = config.shipped_case()
    tab = (release["u"], release["v"], release["w"])
    print("restart 2+2: shipped case 8x64x32^2", flush=True)
    a = U.UGKS(case, *tab)
    a.cold_start()
    _iterate(a, 4)
    a.evolution(outputs=False)
    na, wa = a.norms()
    Wa, ha, ba = a.state()
    m0a, la = a.mass_target()

    b = U.UGKS(case, *tab)
    b.cold_start()
    _iterate(b, 2)
    b.evolution(outputs=False)
    R.save_restart(b, str(tmp_path), step=2)
    b.close()
    c = U.UGKS(case, *tab)
    step, epoch = R.load_restart(c, str(tmp_path))
    assert (step, epoch) == (2, 0)
    _iterate(c, 2)
    c.evolution(outputs=False)
    nc, wc = c.norms()
    Wc, hc, bc = c.state()
    m0c, lc = c.mass_target()
    assert np.array_equal(Wa, Wc) and np.array_equal(ha, hc) and np.array_equal(ba, bc)
    assert np.array_equal(na, nc) and np.array_equal(wa, wc)
    assert m0a == m0c and np.array_equal(la, lc)
    a.close(); c.close()


def test_cpu_restart_continues_on_the_gpu(tmp_path, oracle_mod, release):
    import ugks2d_high_order_amd as U
    case = config.shipped_case()
    u, v, w = release["u"], release["v"], release["w"]
    print("restart CPU->GPU: shipped case, oracle 2 iterations then 1 on each side", flush=True)
    o = oracle_mod.OracleSolver(case, u, v, w, threads=8)
    o.cold_start(1.0, 0.0, case.t_init)
    W0, _, _ = o.state()
    m0 = 0.0
    for c in range(case.nx * case.ny):          # mass of the cold state, in cell order
        m0 += case.dx * case.dy * W0[c, 0]
    o.operator()
    ref_h, ref_b = o.norms()[22:24]
    for _ in range(2):
        st, acc, _ = o.iteration()
        assert st == 0 and acc
    W, h, b = o.state()
    R.write_restart(str(tmp_path), case, W, h, b, u, w, m0, np.zeros(4), ref_h, ref_b, step=2, np_y=8)
    g = U.UGKS(case, u, v, w)
    assert R.load_restart(g, str(tmp_path)) == (2, 0)
    g.evolution(outputs=False)
    o.operator()
    ng, _ = g.norms()
    no = o.norms()
    assert _nrel(ng[[0, 1, 2, 3, 12, 13]], no[[0, 1, 2, 3, 12, 13]]) < 1e-12
    assert ng[22] == ref_h and ng[23] == ref_b
    acc, dg = g.implicit_step()
    st, acc_o, do = o.iteration()
    assert acc and acc_o and st == 0
    Wg, hg, bg = g.state()
    Wo, ho, bo = o.state()
    e = (_nrel(Wg, Wo), _nrel(hg, ho), _nrel(bg, bo))
    print("after the restarted update W/h/b", e, flush=True)
    assert max(e) < 1e-12
    assert dg[6] == m0
    g.close()


def test_mass_statistics_and_ledger():
    import ugks2d_high_order_amd as U
    case = config.small_case()
    u, v, w = velocity.table(case.table, case.table_scale)
    print("mass statistics: small case", flush=True)
    g = U.UGKS(case, u, v, w)
    g.cold_start()
    m0, led = g.mass_target()
    assert math.isnan(m0) and np.all(led == 0.0)
    assert all(math.isnan(x) for x in g.references())
    inc = 0.0
    for _ in range(3):
        g.evolution(outputs=False)
        acc, d = g.implicit_step()
        assert acc
        inc += d[5] - d[4]
    m0, led = g.mass_target()
    assert m0 == d[6]
    assert abs(led[0] - inc) <= 1e-15 * max(1.0, abs(m0))
    t, m, ok = g.mass_statistics()
    W, h, b = g.state()
    area = case.dx * case.dy
    mm = np.stack([h @ w, h @ (w * u), h @ (w * v), 0.5 * (h @ (w * (u * u + v * v)) + b @ w)], -1)
    err = np.max(np.abs(W - mm) / np.maximum(np.abs(W), 1.0))
    e = W[:, 3] - 0.5 * (W[:, 1] ** 2 + W[:, 2] ** 2) / W[:, 0]
    nh = np.max((np.maximum(-h, 0) @ w) / np.maximum(np.abs(h) @ w, np.finfo(float).tiny))
    nb = np.max((np.maximum(-b, 0) @ w) / np.maximum(np.abs(b) @ w, np.finfo(float).tiny))
    want_t = np.concatenate([area * W.sum(0), [area * mm[:, 0].sum()]])
    # sums in another order: 1e-13 of the sum of magnitudes (momentum totals cancel)
    scale_t = np.concatenate([area * np.abs(W).sum(0), [area * np.abs(mm[:, 0]).sum()]])
    assert np.all(np.abs(t - want_t) <= 1e-13 * scale_t)
    # the moment error is itself round-off: compare to the round-off of the moment sums
    assert abs(m[0] - err) <= 1e-13
    assert abs(m[1] - nh) <= 1e-13 and abs(m[2] - nb) <= 1e-13
    assert m[3] == W[:, 0].min() and abs(m[4] - e.min()) <= 1e-14 * abs(e.min())
    assert ok == (err <= case.moment_tol)
    # references and target survive a restart scalar round trip; a cold start clears them
    g.inherit_references(0.5, 0.25)
    assert g.references() == (0.5, 0.25)
    g.set_mass_target(2.0, [1.0, 2.0, 3.0, 4.0])
    assert g.mass_target()[0] == 2.0 and np.array_equal(g.mass_target()[1], [1.0, 2.0, 3.0, 4.0])
    g.cold_start()
    assert math.isnan(g.mass_target()[0]) and math.isnan(g.references()[0])
    g.close()


def test_solve_loop_writes_the_reference_files(tmp_path, release):
    """solve_ugks with an output directory leaves what the reference's run
    leaves after Implicit_MaxIter = 2 on the shipped case: one
    sparse_residual.csv / consecutive_states.csv row per accepted state with
    that state's residual norms, ResidualImplicit.plt, R5_status.txt
    ("2 2 0 F iteration_limit", as the release run) and a complete restart
    that loads back into the final state bit for bit."""
    import ugks2d_high_order_amd as U
    case = config.shipped_case()
    tab = (release["u"], release["v"], release["w"])
    print("solve loop with output: shipped case, 2 iterations", flush=True)
    g = U.UGKS(case, *tab)
    g.cold_start()
    out = tmp_path / "run"
    out.mkdir()
    s, hist = U.solve_ugks(case, max_iter=2, solver=g, output=str(out))
    lines = (out / "sparse_residual.csv").read_text().splitlines()
    assert lines[0] == R.SPARSE_HEADER and len(lines) == 3
    for row, h in zip(lines[1:], hist):
        f = row.split(",")
        assert f[:3] == [str(h["iter"]), str(h["iter"]), "accepted_convergence_verification"]
        v = np.array([float(t) for t in f[3:]])
        sci = v[12:34]
        assert np.array_equal(np.delete(sci, 18), np.delete(h["norms"][:22], 18))
        assert np.array_equal(v[34:50], h["walls"])
    assert (out / "R5_status.txt").read_text().splitlines()[1] == "2 2 0 F iteration_limit"
    assert len((out / "ResidualImplicit.plt").read_text().splitlines()) == 3
    assert (out / "consecutive_states.csv").read_text().splitlines()[1:] == ["1,F,0", "2,F,0"]
    W, h, b = g.state()
    g2 = U.UGKS(case, *tab)
    assert R.load_restart(g2, str(out)) == (2, 0)
    W2, h2, b2 = g2.state()
    assert np.array_equal(W, W2) and np.array_equal(h, h2) and np.array_equal(b, b2)
    g.close(); g2.close()


def test_state_planes_ragged():
    """ugks_get_state_planes / ugks_set_state_planes on 135 cells (not a
    multiple of the 64-cell tile) x 256 velocities: the planes are the exact
    transpose of the [cell, Nvel] state, in both directions."""
    import ugks2d_high_order_amd as U
    case = config.small_case(nx=9, ny=15)
    u, v, w = velocity.table(case.table, case.table_scale)
    print("state planes: 9x15 cells x 16x16 velocities", flush=True)
    g = U.UGKS(case, u, v, w)
    rng = np.random.default_rng(11)
    nc, nv = 135, 256
    W, h, b = rng.normal(size=(nc, 4)), rng.normal(size=(nc, nv)), rng.normal(size=(nc, nv))
    g.set_state(W, h, b)
    W2, P = g.state_planes()
    assert np.array_equal(W2, W) and np.array_equal(P[0], h.T) and np.array_equal(P[1], b.T)
    W3, P3 = rng.normal(size=(nc, 4)), rng.normal(size=(2, nv, nc))
    g.set_state_planes(W3, P3)
    W4, h4, b4 = g.state()
    assert np.array_equal(W4, W3) and np.array_equal(h4, P3[0].T) and np.array_equal(b4, P3[1].T)
    g.close()
