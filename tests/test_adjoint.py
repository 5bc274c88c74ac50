"""The differentiable S2CN2 restatement and the checkpointed adjoint (CPU).

- The torch restatement of one S2CN2 step (s2cn2_torch) equals the oracle's
  step (or_cn_step, pinned in test_oracle_cn) from a non-equilibrium state:
  W, h, b to 1e-13 normwise relative, the face fluxes to 1e-12.
- The checkpointed reverse sweep (adjoint.trajectory_objective, 4 steps with
  a checkpoint every 2) gives the same gradients as reverse-mode
  differentiation straight through the 4 steps (1e-13).
- Central differences with h = 1e-6 (relative to each parameter; along a
  seeded relative perturbation of the initial distributions) agree with the
  adjoint gradients with respect to the wall speed, mu_ref, Pr and f0 to
  1e-6 relative.

Case: 8x16 cells, 16^2 velocities, Ma = 2 walls, Kn = 0.1, omega = 0.81,
initial f = discrete equilibrium x (1 + 1e-3 N(0,1)), seeded.
"""
import math

import numpy as np
import pytest
import torch

from ugks2d_high_order_amd import adjoint as A
from ugks2d_high_order_amd import config, velocity
from ugks2d_high_order_amd.s2cn2_torch import S2CN2Model


def _f(x):
    return torch.tensor(x, dtype=torch.float64)


@pytest.fixture(scope="module")
def setup():
    torch.set_num_threads(8)
    U0 = config.wall_speed(2.0, 2.0)
    case = config.small_case(wall_ux=(-U0, U0), mu_ref=0.1 / math.sqrt(math.pi), omega=0.81)
    u, v, w = velocity.table(case.table, case.table_scale)
    M = S2CN2Model(case, u, v, w)
    return case, (u, v, w), M, U0


def test_torch_step_equals_oracle(oracle_mod):
    case = config.small_case()
    tab = velocity.table(case.table, case.table_scale)
    o = oracle_mod.OracleSolver(case, *tab, threads=8)
    o.cold_start(1.0, 0.0, case.t_init)
    for _ in range(3):
        st, acc, _ = o.iteration()
        assert st == 0 and acc
    W, h, b = o.state()
    o.operator()
    oo = o.operator_outputs()
    M = S2CN2Model(case, *tab)
    nx, ny, n = case.nx, case.ny, tab[0].size
    Wt, ht, bt = (torch.tensor(a).reshape(ny, nx, -1) for a in (W, h, b))
    xh, _xb, xw, yh, _yb, yw = M.fluxes(Wt, ht, bt, oo["dt"], (_f(case.wall_ux[0]), _f(case.wall_ux[1])), _f(case.pr),
                                        _f(case.mu_ref))

    def nrel(a, b_):
        return float(np.max(np.abs(a - b_)) / np.max(np.abs(b_)))

    assert nrel(xh.reshape(-1, n).numpy(), oo["xflux_h"]) < 1e-12
    assert nrel(yh.reshape(-1, n).numpy(), oo["yflux_h"]) < 1e-12
    assert nrel(xw.reshape(-1, 4).numpy(), oo["xflux_w"]) < 1e-12
    assert nrel(yw.reshape(-1, 4).numpy(), oo["yflux_w"]) < 1e-12
    Wn, hn, bn, _fw, dt = M.step(Wt, ht, bt, case.wall_ux[1], case.mu_ref, case.pr)
    st, acc, d = o.cn_step()
    assert st == 0 and acc and abs(dt - d[13]) <= 1e-15 * dt
    Wo, ho, bo = o.state()
    e = [nrel(Wn.reshape(-1, 4).numpy(), Wo), nrel(hn.reshape(-1, n).numpy(), ho), nrel(bn.reshape(-1, n).numpy(), bo)]
    assert max(e) < 1e-13, e


def test_checkpointed_adjoint_matches_direct_and_finite_differences(setup):
    case, _tab, M, U0 = setup
    h0, b0 = A.perturbed_equilibrium(M)
    K, every = 4, 2
    hh, bb = h0.clone().requires_grad_(True), b0.clone().requires_grad_(True)
    p = [_f(U0).requires_grad_(True), _f(case.mu_ref).requires_grad_(True), _f(case.pr).requires_grad_(True)]
    J = A.trajectory_objective(M, hh, bb, *p, K, every, backend="torch")
    J.backward()
    # straight through the steps
    hd, bd = h0.clone().requires_grad_(True), b0.clone().requires_grad_(True)
    q = [_f(U0).requires_grad_(True), _f(case.mu_ref).requires_grad_(True), _f(case.pr).requires_grad_(True)]
    W, h, b = A.initial_state(M, hd, bd), hd, bd
    Jd = 0.0
    for _ in range(K):
        W, h, b, fw, _dt = M.step(W, h, b, *q)
        Jd = Jd + A.wall_objective(M, fw, q[0])
    Jd.backward()
    assert abs(J.item() - Jd.item()) <= 1e-14 * abs(Jd.item())
    for a, c in zip(p, q):
        assert abs(a.grad.item() - c.grad.item()) <= 1e-13 * abs(c.grad.item())
    assert float((hh.grad - hd.grad).abs().max() / hd.grad.abs().max()) < 1e-13

    def Jof(U, mu, pr, h_, b_):
        with torch.no_grad():
            return A.trajectory_objective(M, h_, b_, U, mu, pr, K, every, backend="torch").item()

    base = [U0, case.mu_ref, case.pr]
    for i in range(3):
        eps = 1e-6 * abs(base[i])
        bp, bm = list(base), list(base)
        bp[i] += eps
        bm[i] -= eps
        fd = (Jof(*bp, h0, b0) - Jof(*bm, h0, b0)) / (2 * eps)
        g = p[i].grad.item()
        print("parameter", i, "FD", fd, "adjoint", g, flush=True)
        assert abs(fd - g) <= 1e-6 * abs(g)
    rng = np.random.Generator(np.random.PCG64(7))
    dh = torch.tensor(rng.standard_normal(tuple(h0.shape))) * h0
    db = torch.tensor(rng.standard_normal(tuple(b0.shape))) * b0
    eps = 1e-6
    fd = (Jof(*base, h0 + eps * dh, b0 + eps * db) - Jof(*base, h0 - eps * dh, b0 - eps * db)) / (2 * eps)
    g = ((hh.grad * dh).sum() + (bb.grad * db).sum()).item()
    print("f0 direction FD", fd, "adjoint", g, flush=True)
    assert abs(fd - g) <= 1e-6 * abs(g)
