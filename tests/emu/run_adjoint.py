"""Runs the reverse S2CN2 step of the emulated library (the HIP sources
compiled for the CPU) and the reverse-mode derivative of the FP64 torch
restatement of the same step on the same inputs, and prints one JSON line of
normwise relative differences.  TEST INFRASTRUCTURE ONLY: a separate process,
so the emulated library never replaces the HIP library in a test process."""
import ctypes as C
import json
import math
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import build_emu  # noqa: E402
import ugks2d_high_order_amd as U  # noqa: E402
from ugks2d_high_order_amd import adjoint as A  # noqa: E402
from ugks2d_high_order_amd import config, velocity  # noqa: E402
from ugks2d_high_order_amd.s2cn2_torch import S2CN2Model  # noqa: E402


def nrel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def case_and_inputs(seed=5, ny=16, rough=0.0):
    """Small Couette case at Ma 2, Kn 0.1; the state is the discrete
    equilibrium times (1 + 0.03 N(0,1)), or times exp(rough N(0,1)) when
    rough > 0 (sharp enough for the positivity limiter to act)."""
    U0 = config.wall_speed(2.0, 2.0)
    case = config.small_case(ny=ny, wall_ux=(-U0, U0), mu_ref=0.1 / math.sqrt(math.pi), omega=0.81)
    tab = velocity.table(case.table, case.table_scale)
    M = S2CN2Model(case, *tab)
    h0, b0 = A.perturbed_equilibrium(M, noise=3e-2, seed=seed)
    if rough > 0.0:
        r = np.random.Generator(np.random.PCG64(seed + 7))
        h0 = h0 * torch.tensor(np.exp(rough * r.standard_normal(tuple(h0.shape))))
        b0 = b0 * torch.tensor(np.exp(rough * r.standard_normal(tuple(b0.shape))))
    W0 = A.initial_state(M, h0, b0)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    nc, n = case.nx * case.ny, M.n
    lam = (rng.standard_normal((nc, 4)), rng.standard_normal((nc, n)) * 1e-1, rng.standard_normal((nc, n)) * 1e-1)
    gwall = rng.standard_normal((2, case.nx, 4))
    return case, tab, M, U0, (W0, h0, b0), lam, gwall


def torch_reference(case, M, U0, state, lam, gwall):
    ny, nx, n = case.ny, case.nx, M.n
    W = state[0].clone().requires_grad_(True)
    h = state[1].clone().requires_grad_(True)
    b = state[2].clone().requires_grad_(True)
    p = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in (U0, case.mu_ref, case.pr)]
    wn, hn, bn, fw, _dt = M.step(W, h, b, *p)
    L = ((torch.tensor(lam[0]).reshape(ny, nx, 4) * wn).sum() + (torch.tensor(lam[1]).reshape(ny, nx, n) * hn).sum()
         + (torch.tensor(lam[2]).reshape(ny, nx, n) * bn).sum() + (torch.tensor(gwall) * fw).sum())
    L.backward()
    return (W.grad.reshape(-1, 4).numpy(), h.grad.reshape(-1, n).numpy(), b.grad.reshape(-1, n).numpy(),
            [x.grad.item() for x in p])


def hip_reverse(case, tab, state, lam, gwall, wg=64):
    g = U.UGKS(case, *tab)
    g.set_tuning(U.UGKS.TUNE_VEL_WG, wg)
    nc, n = case.nx * case.ny, tab[0].size
    arrs = [np.ascontiguousarray(x.reshape(nc, -1).numpy() if torch.is_tensor(x) else x, dtype=np.float64)
            for x in state]
    lams = [np.ascontiguousarray(x, dtype=np.float64) for x in lam]
    outs = [np.zeros((nc, 4)), np.zeros((nc, n)), np.zeros((nc, n))]
    ptr = lambda a: a.ctypes.data_as(C.c_void_p).value  # noqa: E731
    gp = g.cn_adjoint_device(*[ptr(a) for a in arrs], *[ptr(a) for a in lams], gwall, *[ptr(a) for a in outs])
    return outs, gp, g


def limiter_active(M, state):
    """Number of (cell, velocity) pairs whose positivity factor is below 1."""
    with torch.no_grad():
        hf, _, _, hv = M.reconstruct(state[1])
        bf, _, _, bv = M.reconstruct(state[2])
    neg = (hf < 0).flatten(-2).any(-1) | (hv < 0).flatten(-2).any(-1) | (bf < 0).flatten(-2).any(-1) | \
        (bv < 0).flatten(-2).any(-1)
    return int(neg.sum())


def main():
    U.load_library(build_emu.build())
    torch.set_num_threads(4)
    out = {}
    cases = (("rough_wg128", 128, 0.5),) if "--quick" in sys.argv else (("smooth_wg64", 64, 0.0),
                                                                          ("rough_wg128", 128, 0.5))
    for name, wg, rough in cases:
        case, tab, M, U0, state, lam, gwall = case_and_inputs(rough=rough)
        gW, gh, gb, gp = torch_reference(case, M, U0, state, lam, gwall)
        (hW, hh, hb), hp, _g = hip_reverse(case, tab, state, lam, gwall, wg)
        hU = -hp[2] + hp[3]
        out[name] = dict(limited=limiter_active(M, state), W=nrel(hW, gW), h=nrel(hh, gh), b=nrel(hb, gb),
                            U=abs(hU - gp[0]) / abs(gp[0]), mu=abs(hp[0] - gp[1]) / abs(gp[1]),
                            pr=abs(hp[1] - gp[2]) / abs(gp[2]))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
