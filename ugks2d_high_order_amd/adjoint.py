"""Gradients of the integrated wall shear plus wall heat flux of K S2CN2
steps, with respect to the wall speed, the reference viscosity (mu_ref, the
reference collision time), the Prandtl number and the initial distributions.

Objective (wall diagnostics of PublishWallDiagnostics,
src/mod_S2NativeAdapter.f90:1075-1097): per wall face, rate = F/dt with F the
face's macro flux (x dt, per unit length), shear = -rate(Px) and heat =
rate(E) - U_w rate(Px); J = sum over steps, walls and faces of
(shear + heat) * dx * dt.

Forward: the HIP S2CN2 step (ugks_cn_step) on the GPU, or the FP64 torch
restatement (s2cn2_torch) anywhere; the states at every `every`-th step are
kept (sqrt(K) checkpointing for every = sqrt(K)).
Backward: per segment, the states are recomputed from the segment's
checkpoint with the forward stepper, then the steps are reversed one at a
time: each step is re-run in the torch restatement with a graph and its
vector-Jacobian product taken (the closures by the implicit-function
theorem, the rest by reverse-mode differentiation of the restatement).  Peak
memory is one step's graph plus `every` states.

The reverse sweep differentiates the torch restatement, which agrees with
the HIP step to round-off; it is not a hand-written HIP adjoint.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .s2cn2_torch import S2CN2Model


def wall_objective(model: S2CN2Model, wall_fw, wall_u):
    """sum over both walls and their faces of (shear + heat) dx dt for one
    step; wall_fw [2, nx, 4] (x dt, per unit length), walls at -U, +U."""
    uw = torch.stack([-wall_u, wall_u])[:, None]
    return model.dx * (wall_fw[..., 3] - (1.0 + uw) * wall_fw[..., 1]).sum()


class TorchStepper:
    """Forward stepper on the torch restatement (no graph)."""

    def __init__(self, model: S2CN2Model):
        self.model = model

    def run(self, W, h, b, params, nsteps, keep_every=1):
        """Advance nsteps; returns the list of kept states (every keep_every
        steps, the start included), the per-step wall fluxes and dts."""
        m = self.model
        wall_u, mu_ref, pr = params
        kept, fws, dts = [(W, h, b)], [], []
        with torch.no_grad():
            for k in range(nsteps):
                W, h, b, fw, dt = m.step(W, h, b, wall_u, mu_ref, pr)
                fws.append(fw)
                dts.append(dt)
                if (k + 1) % keep_every == 0:
                    kept.append((W, h, b))
        return kept, fws, dts


class HipStepper:
    """Forward stepper on the HIP library (ugks_cn_step); the state crosses
    to the host only at kept steps."""

    def __init__(self, model: S2CN2Model):
        self.model = model

    def run(self, W, h, b, params, nsteps, keep_every=1):
        from . import UGKS
        m = self.model
        wall_u, mu_ref, pr = (float(x) for x in params)
        case = dataclasses.replace(m.case, wall_ux=(-wall_u, wall_u), mu_ref=mu_ref, pr=pr)
        g = UGKS(case, m.u.cpu().numpy(), m.v.cpu().numpy(), m.w.cpu().numpy())
        dev, sh = m.dev, (m.ny, m.nx)

        def put(W_, h_, b_):
            g.set_state(W_.reshape(-1, 4).cpu().numpy(), h_.reshape(-1, m.n).cpu().numpy(),
                        b_.reshape(-1, m.n).cpu().numpy())

        def get():
            Wn, hn, bn = g.state()
            t = lambda a, k: torch.as_tensor(a, device=dev).reshape(*sh, k)  # noqa: E731
            return t(Wn, 4), t(hn, m.n), t(bn, m.n)

        put(W, h, b)
        kept, fws, dts = [(W, h, b)], [], []
        for k in range(nsteps):
            acc, d = g.cn_step()
            if not acc:
                raise FloatingPointError(f"S2CN2 step {k} rejected by the gates (diag {d.tolist()})")
            fws.append(torch.as_tensor(g.wall_fluxes(), device=dev))
            dts.append(float(d[13]))
            if (k + 1) % keep_every == 0:
                kept.append(get())
        g.close()
        return kept, fws, dts


def initial_state(model: S2CN2Model, h0, b0):
    """W of the initial distributions (their discrete moments)."""
    return model.moments(h0, b0)


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h0, b0, wall_u, mu_ref, pr, model, stepper, nsteps, every):
        params = (wall_u.detach(), mu_ref.detach(), pr.detach())
        with torch.no_grad():
            W0 = initial_state(model, h0, b0)
        kept, fws, dts = stepper.run(W0, h0.detach(), b0.detach(), params, nsteps, keep_every=every)
        J = sum(wall_objective(model, fw, params[0]) for fw in fws)
        ctx.model, ctx.stepper, ctx.nsteps, ctx.every = model, stepper, nsteps, every
        ctx.kept, ctx.dts, ctx.params = kept, dts, params
        ctx.save_for_backward(h0, b0)
        return J.detach()

    @staticmethod
    def backward(ctx, gJ):
        model, stepper, K, every = ctx.model, ctx.stepper, ctx.nsteps, ctx.every
        h0, b0 = ctx.saved_tensors
        p0 = [p.detach().clone().requires_grad_(True) for p in ctx.params]
        lam = None
        gp = [torch.zeros_like(p) for p in p0]
        for seg in reversed(range(0, K, every)):
            n = min(every, K - seg)
            W, h, b = ctx.kept[seg // every]
            states, _, _ = stepper.run(W, h, b, ctx.params, n - 1, keep_every=1) if n > 1 else ([(W, h, b)], 0, 0)
            for k in reversed(range(n)):
                Wk, hk, bk = (x.detach().clone().requires_grad_(True) for x in states[k])
                with torch.enable_grad():
                    Wn, hn, bn, fw, _ = model.step(Wk, hk, bk, p0[0], p0[1], p0[2], dt=ctx.dts[seg + k])
                    tot = gJ * wall_objective(model, fw, p0[0])
                    if lam is not None:
                        tot = tot + (lam[0] * Wn).sum() + (lam[1] * hn).sum() + (lam[2] * bn).sum()
                    grads = torch.autograd.grad(tot, [Wk, hk, bk] + p0, allow_unused=True)
                lam = [g if g is not None else torch.zeros_like(x) for g, x in zip(grads[:3], (Wk, hk, bk))]
                for i in range(3):
                    if grads[3 + i] is not None:
                        gp[i] = gp[i] + grads[3 + i]
        # the initial W is the moments of (h0, b0)
        with torch.enable_grad():
            hh = h0.detach().clone().requires_grad_(True)
            bb = b0.detach().clone().requires_grad_(True)
            W0 = initial_state(model, hh, bb)
            gh, gb = torch.autograd.grad((lam[0] * W0).sum(), [hh, bb])
        return lam[1] + gh, lam[2] + gb, gp[0], gp[1], gp[2], None, None, None, None


def trajectory_objective(model: S2CN2Model, h0, b0, wall_u, mu_ref, pr, nsteps: int, every: int,
                         backend: str = "hip"):
    """J(h0, b0, wall speed, mu_ref, Pr) over `nsteps` S2CN2 steps from the
    initial distributions (W0 = their moments), differentiable with
    torch.autograd; checkpoints every `every` steps."""
    stepper = HipStepper(model) if backend == "hip" else TorchStepper(model)
    f = lambda x: torch.as_tensor(x, dtype=torch.float64, device=model.dev)  # noqa: E731
    return _Trajectory.apply(h0, b0, f(wall_u), f(mu_ref), f(pr), model, stepper, nsteps, every)


def perturbed_equilibrium(model: S2CN2Model, rho=1.0, t_native=2.0, noise=1e-3, seed=20261203):
    """Initial distributions of the adjoint workload: the discrete equilibrium
    of the rest gas (rho, 0, 0, T) times (1 + noise * N(0,1)), seeded."""
    ck = model.ck
    w = torch.tensor([[rho, 0.0, 0.0, rho * (ck + 2) / 4.0 * t_native]], dtype=torch.float64, device=model.dev)
    th, ok = model.closure_newton(w)
    assert bool(ok.all())
    q, qb, *_ = model.q_state(th)
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (model.ny, model.nx, model.n)
    xi_h = torch.as_tensor(rng.standard_normal(shape), device=model.dev)
    xi_b = torch.as_tensor(rng.standard_normal(shape), device=model.dev)
    return q[0] * (1.0 + noise * xi_h), qb[0] * (1.0 + noise * xi_b)
