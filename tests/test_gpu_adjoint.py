"""The checkpointed adjoint over HIP forward steps (ugks_cn_step).

- The HIP forward trajectory (4 S2CN2 steps) and the torch restatement's
  trajectory on the GPU give the same objective (1e-12 relative) and the same
  checkpointed states (1e-12 normwise relative).
- Gradients with the HIP forward equal those with the torch forward (1e-10
  relative: the recomputed states differ by round-off only).
- Central differences of the HIP forward's objective (h = 1e-6 relative to
  each parameter, and along a seeded relative perturbation of f0) agree with
  the adjoint gradients to 1e-6 relative.

Case: 8x16 cells x 16^2 velocities, Ma = 2 walls, Kn = 0.1, omega = 0.81,
initial f = discrete equilibrium x (1 + 1e-3 N(0,1)), seeded; K = 4,
checkpoints every 2 steps.
"""
import math

import numpy as np
import pytest
import torch

from ugks2d_high_order_amd import adjoint as A
from ugks2d_high_order_amd import config, velocity
from ugks2d_high_order_amd.s2cn2_torch import S2CN2Model

pytestmark = pytest.mark.gpu

K, EVERY = 4, 2


@pytest.fixture(scope="module")
def setup():
    U0 = config.wall_speed(2.0, 2.0)
    case = config.small_case(wall_ux=(-U0, U0), mu_ref=0.1 / math.sqrt(math.pi), omega=0.81)
    tab = velocity.table(case.table, case.table_scale)
    M = S2CN2Model(case, *tab, device="cuda")
    h0, b0 = A.perturbed_equilibrium(M)
    return case, M, U0, h0, b0


def _grads(M, case, U0, h0, b0, backend):
    f = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")  # noqa: E731
    hh, bb = h0.clone().requires_grad_(True), b0.clone().requires_grad_(True)
    p = [f(U0).requires_grad_(True), f(case.mu_ref).requires_grad_(True), f(case.pr).requires_grad_(True)]
    J = A.trajectory_objective(M, hh, bb, *p, K, EVERY, backend=backend)
    J.backward()
    return J.item(), [x.grad.item() for x in p], hh.grad, bb.grad


def test_hip_and_torch_trajectories_agree(setup):
    case, M, U0, h0, b0 = setup
    W0 = A.initial_state(M, h0, b0)
    params = (M.f64(U0), M.f64(case.mu_ref), M.f64(case.pr))
    kh, fh, dh = A.HipStepper(M).run(W0, h0, b0, params, K, keep_every=EVERY)
    kt, ft, dt = A.TorchStepper(M).run(W0, h0, b0, params, K, keep_every=EVERY)
    assert all(abs(a - b) <= 1e-15 * b for a, b in zip(dh, dt)), (dh, dt)
    Jh = sum(A.wall_objective(M, x, params[0]) for x in fh).item()
    Jt = sum(A.wall_objective(M, x, params[0]) for x in ft).item()
    print("J hip", Jh, "J torch", Jt, flush=True)
    assert abs(Jh - Jt) <= 1e-12 * abs(Jt)
    for a, b in zip(kh, kt):
        for x, y in zip(a, b):
            e = float((x - y).abs().max() / y.abs().max())
            assert e < 1e-12, e


def test_hip_adjoint_matches_torch_adjoint_and_finite_differences(setup):
    case, M, U0, h0, b0 = setup
    Jh, gh, hgh, bgh = _grads(M, case, U0, h0, b0, "hip")
    Jt, gt, hgt, bgt = _grads(M, case, U0, h0, b0, "torch")
    print("grads hip", gh, "torch", gt, flush=True)
    for a, b in zip(gh, gt):
        assert abs(a - b) <= 1e-10 * abs(b)
    assert float((hgh - hgt).abs().max() / hgt.abs().max()) < 1e-10
    assert float((bgh - bgt).abs().max() / bgt.abs().max()) < 1e-10

    def Jof(U, mu, pr, h_, b_):
        with torch.no_grad():
            return A.trajectory_objective(M, h_, b_, U, mu, pr, K, EVERY, backend="hip").item()

    base = [U0, case.mu_ref, case.pr]
    for i in range(3):
        eps = 1e-6 * abs(base[i])
        bp, bm = list(base), list(base)
        bp[i] += eps
        bm[i] -= eps
        fd = (Jof(*bp, h0, b0) - Jof(*bm, h0, b0)) / (2 * eps)
        print("parameter", i, "FD", fd, "adjoint", gh[i], flush=True)
        assert abs(fd - gh[i]) <= 1e-6 * abs(gh[i])
    rng = np.random.Generator(np.random.PCG64(7))
    dh = torch.tensor(rng.standard_normal(tuple(h0.shape)), device="cuda") * h0
    db = torch.tensor(rng.standard_normal(tuple(b0.shape)), device="cuda") * b0
    eps = 1e-6
    fd = (Jof(*base, h0 + eps * dh, b0 + eps * db) - Jof(*base, h0 - eps * dh, b0 - eps * db)) / (2 * eps)
    g = ((hgh * dh).sum() + (bgh * db).sum()).item()
    print("f0 direction FD", fd, "adjoint", g, flush=True)
    assert abs(fd - g) <= 1e-6 * abs(g)
