"""Gradients of the integrated wall shear plus wall heat flux of K S2CN2
steps, with respect to the wall speed, the reference viscosity (mu_ref, the
reference collision time), the Prandtl number and the initial distributions.

Objective (wall diagnostics of PublishWallDiagnostics,
src/mod_S2NativeAdapter.f90:1075-1097): per wall face, rate = F/dt with F the
face's macro flux (x dt, per unit length), shear = -rate(Px) and heat =
rate(E) - U_w rate(Px); J = sum over steps, walls and faces of
(shear + heat) * dx * dt.

trajectory_objective() is a torch.autograd.Function with sqrt(K) step
checkpointing: the forward keeps the state every `every` steps; the backward
recomputes one segment at a time from its checkpoint and reverses its steps
in the opposite order.

backend "hip": forward steps are ugks_cn_step and reverse steps are
ugks_cn_adjoint_device, the hand-written HIP adjoint of the step; states and
adjoints stay in device memory (torch tensors on the model's device).
backend "torch": both directions run the FP64 torch restatement
(s2cn2_torch); its reverse-mode derivative is the slow reference the HIP
adjoint is checked against.
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
    """Forward and reverse S2CN2 steps on the HIP library, one instance per
    parameter set; states and adjoints are FP64 tensors on the model's device
    passed by address (ugks_set_state_device, ugks_cn_adjoint_device)."""

    def __init__(self, model: S2CN2Model):
        self.model = model
        self._g = None
        self._key = None

    def _instance(self, params):
        from . import UGKS
        key = tuple(float(x) for x in params)
        if self._g is None or self._key != key:
            if self._g is not None:
                self._g.close()
            m = self.model
            wall_u, mu_ref, pr = key
            case = dataclasses.replace(m.case, wall_ux=(-wall_u, wall_u), mu_ref=mu_ref, pr=pr)
            self._g = UGKS(case, m.u.cpu().numpy(), m.v.cpu().numpy(), m.w.cpu().numpy())
            self._key = key
        return self._g

    def _flat(self, W, h, b):
        """Contiguous FP64 copies on the model's device; torch's pending work
        on them is finished before the library (its own stream) reads them."""
        m = self.model
        f = lambda x, k: x.detach().to(device=m.dev, dtype=torch.float64).reshape(-1, k).contiguous()  # noqa: E731
        out = f(W, 4), f(h, m.n), f(b, m.n)
        if m.dev.type == "cuda":
            torch.cuda.current_stream(m.dev).synchronize()
        return out

    def _shape(self, W, h, b):
        m = self.model
        return W.reshape(m.ny, m.nx, 4), h.reshape(m.ny, m.nx, m.n), b.reshape(m.ny, m.nx, m.n)

    def _get(self, g):
        m = self.model
        nc = m.nx * m.ny
        W = torch.empty(nc, 4, dtype=torch.float64, device=m.dev)
        h = torch.empty(nc, m.n, dtype=torch.float64, device=m.dev)
        b = torch.empty(nc, m.n, dtype=torch.float64, device=m.dev)
        if m.dev.type == "cuda":
            torch.cuda.current_stream(m.dev).synchronize()
        g.get_state_device(W.data_ptr(), h.data_ptr(), b.data_ptr())
        return self._shape(W, h, b)

    def run(self, W, h, b, params, nsteps, keep_every=1):
        """Advance nsteps; returns the kept states (every keep_every steps,
        the start included), the per-step wall fluxes and dts."""
        g = self._instance(params)
        Wf, hf, bf = self._flat(W, h, b)
        g.set_state_device(Wf.data_ptr(), hf.data_ptr(), bf.data_ptr())
        kept, fws, dts = [self._shape(Wf, hf, bf)], [], []
        for k in range(nsteps):
            acc, d = g.cn_step()
            if not acc:
                raise FloatingPointError(f"S2CN2 step {k} rejected by the gates (diag {d.tolist()})")
            fws.append(torch.as_tensor(g.wall_fluxes(), device=self.model.dev))
            dts.append(float(d[13]))
            if (k + 1) % keep_every == 0:
                kept.append(self._get(g))
        return kept, fws, dts

    def reverse(self, state, lam, gwall, params):
        """Adjoint of the state a step starts from, given the adjoint of its
        result and of its wall fluxes; also d/d(mu_ref, Pr, bottom and top
        wall speed)."""
        g = self._instance(params)
        S = self._flat(*state)
        Lm = self._flat(*lam)
        out = [torch.empty_like(x) for x in S]
        if self.model.dev.type == "cuda":
            torch.cuda.current_stream(self.model.dev).synchronize()
        gp = g.cn_adjoint_device(*[x.data_ptr() for x in S], *[x.data_ptr() for x in Lm],
                                 np.asarray(gwall.detach().cpu(), dtype=np.float64),
                                 *[x.data_ptr() for x in out])
        return self._shape(*out), gp

    def reverse_segment(self, state, n, lam, gwalls, params):
        """Adjoint of the state n steps before the state whose adjoint is
        lam, running and reversing the n steps on the device in one call
        (ugks_cn_adjoint_segment); gwalls: the n steps' wall-flux adjoints."""
        g = self._instance(params)
        S = self._flat(*state)
        Lm = self._flat(*lam)
        out = [torch.empty_like(x) for x in S]
        gw = np.stack([np.asarray(x.detach().cpu(), dtype=np.float64) for x in gwalls])
        if self.model.dev.type == "cuda":
            torch.cuda.current_stream(self.model.dev).synchronize()
        gp = g.cn_adjoint_segment(*[x.data_ptr() for x in S], n, *[x.data_ptr() for x in Lm], gw,
                                  *[x.data_ptr() for x in out])
        return self._shape(*out), gp

    def close(self):
        if self._g is not None:
            self._g.close()
            self._g = None


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
        ctx.kept, ctx.dts, ctx.params, ctx.fws = kept, dts, params, fws
        ctx.save_for_backward(h0, b0)
        return J.detach()

    @staticmethod
    def backward(ctx, gJ):
        model, stepper = ctx.model, ctx.stepper
        h0, b0 = ctx.saved_tensors
        if isinstance(stepper, HipStepper):
            lam, gp = _reverse_hip(ctx, gJ)
        else:
            lam, gp = _reverse_torch(ctx, gJ)
        # the initial W is the moments of (h0, b0)
        with torch.enable_grad():
            hh = h0.detach().clone().requires_grad_(True)
            bb = b0.detach().clone().requires_grad_(True)
            W0 = initial_state(model, hh, bb)
            gh, gb = torch.autograd.grad((lam[0] * W0).sum(), [hh, bb])
        return lam[1] + gh, lam[2] + gb, gp[0], gp[1], gp[2], None, None, None, None


def _wall_adjoint(model, fw, wall_u, gJ):
    """Adjoint of one step's wall fluxes fw [2, nx, 4] and the direct
    wall-speed derivative of that step's objective term."""
    dx = model.dx
    g = torch.zeros_like(fw)
    U = float(wall_u)
    g[0, :, 1] = -dx * (1.0 - U)
    g[1, :, 1] = -dx * (1.0 + U)
    g[:, :, 3] = dx
    dU = dx * (fw[0, :, 1].sum() - fw[1, :, 1].sum())
    return gJ * g, gJ * dU


def _reverse_hip(ctx, gJ):
    """Reverse sweep with the HIP reverse step (ugks_cn_adjoint_device)."""
    model, stepper, K, every = ctx.model, ctx.stepper, ctx.nsteps, ctx.every
    params = ctx.params
    gJ = float(gJ)
    lam = [torch.zeros(model.ny, model.nx, 4, dtype=torch.float64, device=model.dev),
           torch.zeros(model.ny, model.nx, model.n, dtype=torch.float64, device=model.dev),
           torch.zeros(model.ny, model.nx, model.n, dtype=torch.float64, device=model.dev)]
    gU = gmu = gpr = 0.0
    for seg in reversed(range(0, K, every)):
        n = min(every, K - seg)
        gws = []
        for k in range(n):
            gw, dU = _wall_adjoint(model, ctx.fws[seg + k], params[0], gJ)
            gws.append(gw)
            gU += float(dU)
        lam, gp = stepper.reverse_segment(ctx.kept[seg // every], n, lam, gws, params)
        gU += -gp[2] + gp[3]
        gmu += gp[0]
        gpr += gp[1]
    f = lambda x: torch.tensor(x, dtype=torch.float64, device=model.dev)  # noqa: E731
    return lam, (f(gU), f(gmu), f(gpr))


def _reverse_torch(ctx, gJ):
    """Reverse sweep through the torch restatement (reverse-mode autograd of
    each recomputed step)."""
    model, stepper, K, every = ctx.model, ctx.stepper, ctx.nsteps, ctx.every
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
    return lam, gp


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
