"""Differentiable FP64 restatement of one S2CN2 explicit step in PyTorch.

The step is the one the HIP library runs in ugks_cn_step (advance_step with
the physical gas, src/mod_S2D1.f90:1302-1435 via CartesianAdvance,
src/mod_S2CartesianMPI.f90:235-277): WENO5-AO reconstruction of h and b with
the positivity limiter, UGKS face nodes with the discrete-equilibrium
closure, multiscale diffuse walls, volume targets at the old and at the
streamed state, and the Crank-Nicolson Shakhov update with variable tau.
Everything is vectorised over cells, faces and velocities and runs on
whatever device the tensors live on.

The discrete-equilibrium closure theta(w) (solve_q_first,
src/mod_CorrectedWF2DClosure.f90:330-392) is a Newton solve; it is computed
without a graph and differentiated by the implicit-function theorem:
moments(q(theta)) = w gives d theta = J^{-1} dw with J the moment Jacobian,
so its vector-Jacobian product is J^{-T} g (ClosureTheta).  Everything
downstream of theta (q(theta), its theta-derivatives, J, the projections and
directions) is written in differentiable tensor operations.

Used by ugks2d_high_order_amd.adjoint as the reverse-mode model of the HIP
forward step; its forward agrees with ugks_cn_step to round-off.
"""
from __future__ import annotations

import math
from fractions import Fraction

import torch

PI = math.pi
NEWTON_TOL = 5.0e-13
WENO_EPS = 1.0e-8
WENO_GAMMA = (0.01125, 0.1275, 0.01125, 0.85)
MARGIN = 64.0 * 2.220446049250313e-16
GNODE = (-0.5 * math.sqrt(3.0 / 5.0), 0.0, 0.5 * math.sqrt(3.0 / 5.0))
GW = (5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0)


# ---------------------------------------------------------------------------
# exact reconstruction maps
# ---------------------------------------------------------------------------
def _average_map(offsets):
    """Polynomial coefficients (degrees 0..n-1) whose averages over the unit
    cells at `offsets` are the unit samples: map[s][d], exact rationals."""
    n = len(offsets)
    A = [[(Fraction(2 * o + 1, 2) ** (d + 1) - Fraction(2 * o - 1, 2) ** (d + 1)) / (d + 1) for d in range(n)]
         for o in offsets]
    # invert A by Gauss-Jordan over the rationals
    M = [row[:] + [Fraction(int(r == c)) for c in range(n)] for r, row in enumerate(A)]
    for col in range(n):
        piv = next(r for r in range(col, n) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        p = M[col][col]
        M[col] = [x / p for x in M[col]]
        for r in range(n):
            if r != col and M[r][col] != 0:
                f = M[r][col]
                M[r] = [a - f * b for a, b in zip(M[r], M[col])]
    inv = [row[n:] for row in M]          # coefficients c = inv @ samples
    return [[float(inv[d][s]) for d in range(n)] for s in range(n)]


def _maps(device):
    cm = torch.zeros(4, 5, 5, dtype=torch.float64)
    for c, offs in enumerate(((-2, -1, 0), (-1, 0, 1), (0, 1, 2))):
        m = _average_map(offs)
        for s in range(3):
            for d in range(3):
                cm[c, s + c, d] = m[s][d]
    cm[3] = torch.tensor(_average_map((-2, -1, 0, 1, 2)), dtype=torch.float64)
    wm = torch.zeros(5, 5, 5, dtype=torch.float64)   # [shift][sample][degree], samples at s - shift
    for sh in range(5):
        wm[sh] = torch.tensor(_average_map(tuple(s - sh for s in range(5))), dtype=torch.float64)
    return cm.to(device), wm.to(device)


# ---------------------------------------------------------------------------
# WENO5-AO on stencils q[..., 5]
# ---------------------------------------------------------------------------
def weno_weights(q):
    """Nonlinear weights fw[..., 4] (weno_coefficients, src/mod_S2D1Reconstruction.f90:171-204)."""
    q0, q1, q2, q3, q4 = q.unbind(-1)
    b0 = 13.0 / 12.0 * (q0 - 2 * q1 + q2) ** 2 + 0.25 * (q0 - 4 * q1 + 3 * q2) ** 2
    b1 = 13.0 / 12.0 * (q1 - 2 * q2 + q3) ** 2 + 0.25 * (q1 - q3) ** 2
    b2 = 13.0 / 12.0 * (q2 - 2 * q3 + q4) ** 2 + 0.25 * (3 * q2 - 4 * q3 + q4) ** 2
    b3 = (231153 * q2 * q2 + 104963 * q1 * q1 + 6908 * q0 * q0 - 38947 * q0 * q3 + 104963 * q3 * q3
          + q1 * (-51001 * q0 + 179098 * q3 - 38947 * q4) - 3 * q2 * (99692 * q1 - 22641 * q0 + 99692 * q3
                                                                    - 22641 * q4)
          + 8209 * q0 * q4 - 51001 * q3 * q4 + 6908 * q4 * q4) / 5040.0
    tau = (torch.abs(b3 - b0) + torch.abs(b3 - b1) + torch.abs(b3 - b2)) / 3.0
    al = [g * (1.0 + (tau / (b + WENO_EPS)) ** 2) for g, b in zip(WENO_GAMMA, (b0, b1, b2, b3))]
    sa = al[0] + al[1] + al[2] + al[3]
    om = [a / sa for a in al]
    f3 = om[3] / WENO_GAMMA[3]
    return torch.stack([om[0] - f3 * WENO_GAMMA[0], om[1] - f3 * WENO_GAMMA[1], om[2] - f3 * WENO_GAMMA[2], f3], -1)


def weno_coefs(q, cmap):
    """Candidate polynomial coefficients [..., 4, 5] (quadratics padded)."""
    return torch.einsum("...s,csd->...cd", q, cmap)


def poly_at(coef, xi):
    """Value and derivative of polynomials coef[..., 5] at xi."""
    val = coef[..., 4]
    for d in (3, 2, 1, 0):
        val = val * xi + coef[..., d]
    der = 4.0 * coef[..., 4]
    for d in (3, 2, 1):
        der = der * xi + d * coef[..., d]
    return val, der


def weno_at(fw, coef, xi):
    v, d = poly_at(coef, xi)
    return (fw * v).sum(-1), (fw * d).sum(-1)


def wall_at(q, wmap_sh, xi):
    """One-sided quartic through five samples (map per stencil shift)."""
    coef = torch.einsum("...s,...sd->...d", q, wmap_sh)
    return poly_at(coef, xi)


# ---------------------------------------------------------------------------
# model
# ---------------------------------------------------------------------------
class S2CN2Model:
    """Geometry, velocity table and reconstruction maps of one Couette case.

    case: config.CouetteCase; u, v, w: the velocity table (Nvel, u fastest).
    Physical parameters that are differentiated (wall speed, mu_ref, Pr) are
    passed to step() as tensors."""

    def __init__(self, case, u, v, w, device="cpu"):
        self.case = case
        self.dev = torch.device(device)
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=self.dev)  # noqa: E731
        self.u, self.v, self.w = t(u), t(v), t(w)
        self.nx, self.ny, self.n = case.nx, case.ny, self.u.numel()
        self.dx, self.dy = case.dx, case.dy
        self.ck = case.ck
        self.omega = case.omega
        self.cmap, self.wmap = _maps(self.dev)
        ny = self.ny
        rb, near, sh = [], [], []
        for j in range(ny):
            if j <= 2:
                rb.append(0); near.append(True); sh.append(j)
            elif j >= ny - 3:
                rb.append(ny - 5); near.append(True); sh.append(j - (ny - 5))
            else:
                rb.append(j - 2); near.append(False); sh.append(0)
        self.rows = torch.tensor([[r + s for s in range(5)] for r in rb], device=self.dev)      # [ny, 5]
        self.near = torch.tensor(near, device=self.dev)
        self.wm_row = self.wmap[torch.tensor(sh, device=self.dev)]                              # [ny, 5, 5]
        self.cols = (torch.arange(self.nx, device=self.dev)[:, None] + torch.arange(-2, 3, device=self.dev)) % self.nx
        self.umax = float(self.u.max())
        self.vmax = float(self.v.max())

    def f64(self, x):
        """A parameter as an FP64 tensor on the model's device (keeps its graph)."""
        return torch.as_tensor(x, dtype=torch.float64, device=self.dev)

    # -- moments -------------------------------------------------------------
    def moments(self, h, b, u=None, v=None):
        u = self.u if u is None else u
        v = self.v if v is None else v
        w = self.w
        return torch.stack([(w * h).sum(-1), (w * u * h).sum(-1), (w * v * h).sum(-1),
                            (0.5 * w * ((u * u + v * v) * h + b)).sum(-1)], -1)

    def s2_prim(self, W):
        rho = W[..., 0]
        ux, vy = W[..., 1] / rho, W[..., 2] / rho
        eint = W[..., 3] - 0.5 * rho * (ux * ux + vy * vy)
        p = 2.0 / (self.ck + 2) * eint
        return rho, ux, vy, p / rho

    def heat(self, W, h, b):
        _rho, ux, vy, _th = self.s2_prim(W)
        cx = self.u - ux[..., None]
        cy = self.v - vy[..., None]
        c2 = cx * cx + cy * cy
        return torch.stack([(0.5 * self.w * cx * (c2 * h + b)).sum(-1), (0.5 * self.w * cy * (c2 * h + b)).sum(-1)], -1)

    def tau(self, W, mu_ref):
        pr = 2.0 / (self.ck + 2) * (W[..., 3] - 0.5 * (W[..., 1] ** 2 + W[..., 2] ** 2) / W[..., 0])
        lam = 0.5 * W[..., 0] / pr
        return mu_ref * (1.0 / lam) ** self.omega / (0.5 * W[..., 0] / lam)

    # -- closure ---------------------------------------------------------------
    def q_state(self, th):
        """q(theta) = rho/(pi T) exp(-c^2/T), qb = ck/2 T q, their theta
        derivatives [..., 4, n] and the moment Jacobian J[..., r, a]."""
        r, t1, t2, T = th.unbind(-1)
        du = self.u - t1[..., None]
        dv = self.v - t2[..., None]
        c2 = du * du + dv * dv
        q = r[..., None] / (PI * T[..., None]) * torch.exp(-c2 / T[..., None])
        qb = 0.5 * self.ck * T[..., None] * q
        iT = 1.0 / T[..., None]
        l = torch.stack([1.0 / r[..., None].expand_as(q), 2.0 * du * iT, 2.0 * dv * iT, -iT + c2 * iT * iT], -2)
        lb = l + torch.stack([torch.zeros_like(q), torch.zeros_like(q), torch.zeros_like(q), iT.expand_as(q)], -2)
        qt = q[..., None, :] * l
        qbt = qb[..., None, :] * lb
        J = self.moments(qt, qbt).transpose(-1, -2)     # [..., r, a]
        return q, qb, qt, qbt, J

    def closure_newton(self, w):
        """solve_q_first on a batch w[B, 4] (no graph); returns theta, ok."""
        with torch.no_grad():
            rho = w[:, 0]
            ux, vy = w[:, 1] / rho, w[:, 2] / rho
            temp = 4.0 * (w[:, 3] / rho - 0.5 * (ux * ux + vy * vy)) / (self.ck + 2)
            th = torch.stack([rho, ux, vy, temp], -1)

            def resid(t):
                q, qb, _qt, _qbt, J = self.q_state(t)
                r = self.moments(q, qb) - w
                s = torch.clamp(torch.abs(w), min=1.0)
                return r, (torch.abs(r) / s).amax(-1), J

            res, nrm, J = resid(th)
            ok = torch.ones_like(nrm, dtype=torch.bool)
            for _ in range(12):
                act = nrm > NEWTON_TOL
                if not bool(act.any()):
                    break
                step = torch.linalg.solve(J, -res)
                alpha = torch.ones_like(nrm)
                acc = torch.zeros_like(act)
                nth, nres, nn, nJ = th.clone(), res.clone(), nrm.clone(), J.clone()
                while True:
                    trial = act & ~acc & (alpha >= 9.5367431640625e-07)
                    if not bool(trial.any()):
                        break
                    t = th + alpha[:, None] * step
                    valid = trial & (t[:, 0] > 0) & (t[:, 3] > 0) & torch.isfinite(t).all(-1)
                    tt = torch.where(valid[:, None], t, th)
                    r2, n2, J2 = resid(tt)
                    good = valid & torch.isfinite(r2).all(-1) & ((n2 < nrm) | (n2 <= NEWTON_TOL))
                    nth = torch.where(good[:, None], tt, nth)
                    nres = torch.where(good[:, None], r2, nres)
                    nn = torch.where(good, n2, nn)
                    nJ = torch.where(good[:, None, None], J2, nJ)
                    acc = acc | good
                    alpha = torch.where(trial & ~good, 0.5 * alpha, alpha)
                ok = ok & ~(act & ~acc)
                th, res, nrm, J = nth, nres, nn, nJ
            ok = ok & (nrm <= NEWTON_TOL)
        return th, ok

    # -- reconstruction -----------------------------------------------------------
    def reconstruct(self, F):
        """cell_faces9 for every cell of F[ny, nx, K]: face[4,3], dn[4,3],
        dt[4,3] (per unit cell width) and vol[3,3] (x node, y node), each
        [ny, nx, K, ...]."""
        cm = self.cmap
        qx = F[:, self.cols, :].permute(0, 1, 3, 2)                        # [ny, nx, K, 5] row j
        fw = weno_weights(qx)
        co = weno_coefs(qx, cm)
        xv = [weno_at(fw, co, x) for x in (-0.5, 0.5)]
        rows = [weno_at(fw, co, g)[0] for g in GNODE]
        qy = F[self.rows].permute(0, 2, 3, 1)                              # [ny, nx, K, 5] column of row j's stencil
        fwy = weno_weights(qy)
        coy = weno_coefs(qy, cm)
        wmr = self.wm_row[:, None, None]                                   # [ny, 1, 1, 5, 5]
        near = self.near[:, None, None]
        yv = []
        for x in (-0.5, 0.5):
            a_v, a_d = weno_at(fwy, coy, x)
            w_v, w_d = wall_at(qy, wmr, x)
            yv.append((torch.where(near, w_v, a_v), torch.where(near, w_d, a_d)))

        def gather_rows(t):   # t[ny, nx, K] (values of cell (row, i)) -> [ny, nx, K, 5] over the row stencil
            return t[self.rows].permute(0, 2, 3, 1)

        def gather_cols(t):   # -> [ny, nx, K, 5] over columns i-2..i+2 of row j
            return t[:, self.cols, :].permute(0, 1, 3, 2)

        def companion(q, c, along_y):
            fwq = weno_weights(q)
            cq = weno_coefs(q, cm)
            cc = weno_coefs(c, cm)
            val, der, cval = [], [], []
            for g in GNODE:
                v1, d1 = weno_at(fwq, cq, g)
                v2, _ = weno_at(fwq, cc, g)
                if along_y:
                    wv, wd = wall_at(q, wmr, g)
                    wc, _ = wall_at(c, wmr, g)
                    v1, d1, v2 = torch.where(near, wv, v1), torch.where(near, wd, d1), torch.where(near, wc, v2)
                val.append(v1); der.append(d1); cval.append(v2)
            return torch.stack(val, -1), torch.stack(der, -1), torch.stack(cval, -1)

        face, dn, dt = [], [], []
        for s in range(2):   # x faces: the x lines of the row stencil, reconstructed in y
            v_, d_, c_ = companion(gather_rows(xv[s][0]), gather_rows(xv[s][1]), True)
            face.append(v_); dt.append(d_); dn.append(c_)
        for s in range(2):   # y faces: the y lines of the column stencil, reconstructed in x
            v_, d_, c_ = companion(gather_cols(yv[s][0]), gather_cols(yv[s][1]), False)
            face.append(v_); dt.append(d_); dn.append(c_)
        vol = []
        for a in range(3):
            q = gather_rows(rows[a])
            fq = weno_weights(q)
            cq = weno_coefs(q, cm)
            col = []
            for g in GNODE:
                av, _ = weno_at(fq, cq, g)
                wv, _ = wall_at(q, wmr, g)
                col.append(torch.where(near, wv, av))
            vol.append(torch.stack(col, -1))
        return torch.stack(face, -2), torch.stack(dn, -2), torch.stack(dt, -2), torch.stack(vol, -2)

    def limited(self, h, b):
        """Reconstruct h and b, apply limit_cell_outputs (src/mod_S2D1.f90:315-345)."""
        hf, hn, ht, hv = self.reconstruct(h)
        bf, bn, bt, bv = self.reconstruct(b)
        hm, bm = h[..., None, None], b[..., None, None]
        keep = 1.0 - MARGIN

        def cand(m, f):
            neg = f < 0.0
            den = torch.where(neg, m - f, torch.ones_like(f))   # no inf in the unused branch (its gradient is 0)
            return torch.where(neg, keep * m / den, torch.full_like(f, 2.0)).flatten(-2).amin(-1)

        theta = torch.ones_like(h)
        for m, f in ((hm, hf), (bm, bf), (hm, hv), (bm, bv)):
            theta = torch.minimum(theta, cand(m, f))
        theta = torch.clamp(theta, 0.0, 1.0)
        theta = torch.where((h <= 0.0) | (b <= 0.0), torch.zeros_like(theta), theta)
        th = theta[..., None, None]
        sn = torch.tensor([self.dx, self.dx, self.dy, self.dy], dtype=torch.float64, device=self.dev)[:, None]
        st = torch.tensor([self.dy, self.dy, self.dx, self.dx], dtype=torch.float64, device=self.dev)[:, None]
        out = {}
        for name, m, f, dn_, dt_ in (("h", hm, hf, hn, ht), ("b", bm, bf, bn, bt)):
            out[name] = (m + th * (f - m), th * dn_ / sn, th * dt_ / st)   # [ny, nx, n, 4, 3]
        return out

    # -- closure pieces ----------------------------------------------------------
    def closure(self, w):
        """theta(w) for w[..., 4] with the implicit-function backward."""
        shape = w.shape[:-1]
        th = ClosureTheta.apply(w.reshape(-1, 4), self)
        return th.reshape(*shape, 4)

    def direction(self, qt, qbt, J, m):
        """closure_first_direction: sum_a dq/dtheta_a (J^{-1} m)_a."""
        td = torch.linalg.solve(J, m[..., None])[..., 0]
        return (qt * td[..., :, None]).sum(-2), (qbt * td[..., :, None]).sum(-2)

    def make_mke(self, w, th, pr):
        """make_mke_prepared: discrete Maxwellian q(theta) and the projected
        Shakhov kernels kx, ky (src/mod_S2D1.f90:110-151)."""
        q, qb, qt, qbt, J = self.q_state(th)
        rho, ux, vy, thg = self.s2_prim(w)
        lam = 0.5 / thg
        coeff = (0.8 * (1.0 - pr) * lam * lam / rho)[..., None]
        cx = self.u - ux[..., None]
        cy = self.v - vy[..., None]
        c2 = cx * cx + cy * cy
        shh = 2.0 * lam[..., None] * c2 - 4.0
        shb = shh + 2.0 / self.ck
        k = []
        for c in (cx, cy):
            zh = coeff * c * shh * q
            zb = coeff * c * shb * qb
            dh, db = self.direction(qt, qbt, J, self.moments(zh, zb))
            k.append((zh - dh, zb - db))
        return q, qb, k[0], k[1], (qt, qbt, J)

    # -- time step -----------------------------------------------------------------
    def time_step(self, W):
        with torch.no_grad():
            rho, ux, vy, thg = self.s2_prim(W)
            gamma = (self.ck + 4) / (self.ck + 2)
            lam = 0.5 * rho / (thg * rho)
            sos = torch.sqrt(0.5 * gamma / lam)
            pu = torch.clamp(torch.abs(ux) + sos, min=self.umax)
            pv = torch.clamp(torch.abs(vy) + sos, min=self.vmax)
            dtl = self.case.cfl * self.dx * self.dy / (self.dy * pu + self.dx * pv)
            return float(dtl.min())

    # -- UGKS face node ------------------------------------------------------------
    def face_node(self, axis, dt, L, R, mnl, mnr, mtg, pr, mu_ref):
        """ugks_node (src/mod_S2D1.f90:421-521) for all faces of one axis at
        one Gauss node.  L, R: dicts of (h, b) face data (value, dn, dt)
        [F, n]; mnl, mnr, mtg [F, 4].  Returns fh, fb [F, n], fw [F, 4]."""
        vn = self.u if axis == 1 else self.v
        vt = self.v if axis == 1 else self.u
        up = vn >= 0.0
        sel = lambda a, b: torch.where(up, a, b)  # noqa: E731
        f0h, f0b = sel(L["h"][0], R["h"][0]), sel(L["b"][0], R["b"][0])
        fnh, fnb = sel(L["h"][1], R["h"][1]), sel(L["b"][1], R["b"][1])
        fth, ftb = sel(L["h"][2], R["h"][2]), sel(L["b"][2], R["b"][2])
        w0 = self.moments(f0h, f0b)
        th = self.closure(w0)
        mh, mb, kx, ky, (qt, qbt, J) = self.make_mke(w0, th, pr)
        # analytic Maxwellian of w0
        rho = w0[..., 0]
        ux, vy = w0[..., 1] / rho, w0[..., 2] / rho
        tna = 4.0 * (w0[..., 3] / rho - 0.5 * (ux * ux + vy * vy)) / (self.ck + 2)
        lama = (1.0 / tna)[..., None]
        c2 = (self.u - ux[..., None]) ** 2 + (self.v - vy[..., None]) ** 2
        eah = rho[..., None] * lama / PI * torch.exp(-lama * c2)
        eab = 0.5 * self.ck * tna[..., None] * eah
        q0 = self.heat(w0, f0h - eah, f0b - eab)
        s0h = mh + kx[0] * q0[..., 0:1] + ky[0] * q0[..., 1:2]
        s0b = mb + kx[1] * q0[..., 0:1] + ky[1] * q0[..., 1:2]
        nlh, nlb = self.direction(qt, qbt, J, mnl)
        nrh, nrb = self.direction(qt, qbt, J, mnr)
        tgh, tgb = self.direction(qt, qbt, J, mtg)
        slh = vn * sel(nlh, nrh) + vt * tgh
        slb = vn * sel(nlb, nrb) + vt * tgb
        wtime = -self.moments(slh, slb)
        tmh, tmb = self.direction(qt, qbt, J, wtime)
        tau = self.tau(w0, mu_ref)[..., None]
        ex = torch.exp(-dt / tau)
        m3 = tau * (1.0 - ex)
        m4 = -tau * dt * ex + tau * m3
        m0 = dt - m3
        m1 = -tau * m0 + m4
        m2 = 0.5 * dt * dt - tau * m0
        fh = vn * (m0 * s0h + m1 * slh + m2 * tmh + m3 * f0h - m4 * (vn * fnh + vt * fth))
        fb = vn * (m0 * s0b + m1 * slb + m2 * tmb + m3 * f0b - m4 * (vn * fnb + vt * ftb))
        return fh, fb, self.moments(fh, fb)

    # -- multiscale diffuse wall ---------------------------------------------------
    def wall_node(self, side, dt, Wc, hin, bin_, hn, bn, ht, bt, ux, theta, pr, mu_ref):
        """s2w2_wall_node (src/mod_S2MultiscaleWall.f90:45-184) for the nx
        faces of one wall at one Gauss node.  Wc [nx, 4]; face data [nx, n];
        ux: wall speed (tensor); theta: wall temperature T_native/2."""
        ck = self.ck
        u, v, w = self.u, self.v, self.w
        p = 2.0 / (ck + 2) * (Wc[:, 3] - 0.5 * (Wc[:, 1] ** 2 + Wc[:, 2] ** 2) / Wc[:, 0])
        lam = 0.5 / theta
        rg = p / theta
        zero = torch.zeros_like(rg)
        prim = (rg, zero, -ux + zero, lam + zero)
        wc = torch.stack([Wc[:, 0], Wc[:, 2], -Wc[:, 1], Wc[:, 3]], -1)
        wg = torch.stack([rg, zero, -rg * ux, rg * ((ck + 2) / (4.0 * lam) + 0.5 * ux * ux)], -1)
        an = micro_slope(ck, prim, (wg - wc) / (side * 0.5 * self.dy))
        vn, vt = v, -u
        v2 = vn * vn + vt * vt
        swt = torch.stack([(w * -ht).sum(-1), (w * -vn * ht).sum(-1), (w * -vt * ht).sum(-1),
                           (w * -0.5 * (v2 * ht + bt)).sum(-1)], -1)
        ww = torch.stack([rg, rg * ux, zero, rg * ((ck + 2) * theta / 2.0 + ux * ux / 2.0)], -1)
        th = self.closure(ww)
        mh, mb, _qt, _qbt, _J = self.q_state(th)
        at = micro_slope(ck, prim, swt)
        mu_, mv_, mxi = calc_moment(ck, prim)
        m1 = moment_auvxi(an, mu_, mv_, mxi, 1, 0)
        m2 = moment_auvxi(at, mu_, mv_, mxi, 0, 1)
        att = micro_slope(ck, prim, -rg[:, None] * (m1 + m2))
        wtau = torch.stack([wg[:, 0], -wg[:, 2], wg[:, 1], wg[:, 3]], -1)
        mt = wall_time_coeffs(dt, self.tau(wtau, mu_ref))
        if float(pr.detach()) != 1.0:
            vtl = -u + ux
            v2l = vn * vn + vtl * vtl
            hl0 = (0.5 * w * vn * (v2l * hin + bin_)).sum(-1)
            hl1 = (0.5 * w * vtl * (v2l * hin + bin_)).sum(-1)
            coeff = (0.8 * (1.0 - pr) * lam * lam / rg)[:, None]
            qdot = vn * hl0[:, None] + vtl * hl1[:, None]
            shapeh = 2.0 * lam * v2l - 4.0
            corrh = coeff * qdot * shapeh * mh
            corrb = coeff * qdot * (shapeh + 2.0 / ck) * mb
        else:
            corrh = torch.zeros_like(mh)
            corrb = torch.zeros_like(mb)
        mx2 = mxi[2][:, None]

        def expand(a, X, Y):
            return (a[:, 0:1] + a[:, 1:2] * vn + a[:, 2:3] * vt) * X + 0.5 * a[:, 3:4] * (v2 * X + Y)

        ginh = (mt[0][:, None] * mh + mt[1][:, None] * (vn * expand(an, mh, mb) + vt * expand(at, mh, mb))
                + mt[2][:, None] * expand(att, mh, mb) + mt[3][:, None] * hin - mt[4][:, None] * (vn * hn + u * ht)
                + mt[0][:, None] * corrh)
        ginb = (mt[0][:, None] * mb + mt[1][:, None] * (vn * expand(an, mb, mx2 * mh) + vt * expand(at, mb, mx2 * mh))
                + mt[2][:, None] * expand(att, mb, mx2 * mh) + mt[3][:, None] * bin_
                - mt[4][:, None] * (vn * bn + u * bt) + mt[0][:, None] * corrb)
        out = side * vn > 0.0
        wallh = mh / rg[:, None]
        incidence = torch.where(out, w * vn * ginh, torch.zeros_like(ginh)).sum(-1)
        reflection = torch.where(out, torch.zeros_like(ginh), dt * w * vn * wallh).sum(-1)
        rhow = (-incidence / reflection)[:, None]
        fh = torch.where(out, vn * ginh, dt * vn * rhow * wallh)
        fb = torch.where(out, vn * ginb, dt * vn * rhow * (mb / rg[:, None]))
        return fh, fb, self.moments(fh, fb)

    # -- volume targets ------------------------------------------------------------
    def targets(self, W, pr, with_kappa):
        """build_targets_of (src/mod_S2D1.f90:1246-1300): Gauss-averaged
        discrete Maxwellian and projected Shakhov kernels of the 3x3 volume
        nodes of the WENO tensor of W (mean-corrected); kappa = heat(W, kx),
        heat(W, ky)."""
        _f, _dn, _dt, vol = self.reconstruct(W)                  # vol [ny, nx, 4(m), 3(a), 3(b)]
        gw = torch.tensor(GW, dtype=torch.float64, device=self.dev)
        wt2 = gw[:, None] * gw[None, :]                           # [a, b]
        avg = (vol * wt2).sum((-1, -2))                           # [ny, nx, 4]
        wq = vol.permute(0, 1, 3, 4, 2) + (W - avg)[:, :, None, None, :]   # [ny, nx, 3, 3, 4]
        th = self.closure(wq)
        q, qb, kx, ky, _ = self.make_mke(wq, th, pr)
        ws = wt2[None, None, :, :, None]
        mh = (ws * q).sum((2, 3))
        mb = (ws * qb).sum((2, 3))
        kxh, kxb = (ws * kx[0]).sum((2, 3)), (ws * kx[1]).sum((2, 3))
        kyh, kyb = (ws * ky[0]).sum((2, 3)), (ws * ky[1]).sum((2, 3))
        kap = None
        if with_kappa:
            kap = torch.cat([self.heat(W, kxh, kxb), self.heat(W, kyh, kyb)], -1)   # [ny, nx, 4]
        return mh, mb, (kxh, kxb), (kyh, kyb), kap

    # -- fluxes of a state ---------------------------------------------------------
    def fluxes(self, W, h, b, dt, wall_ux, pr, mu_ref):
        """compute_flux: x-face and y-face micro and macro fluxes (x dt, per
        unit length), walls included.  Returns xh, xb [ny, nx, n] (east face of
        each cell), yh, yb [ny+1, nx, n] (south face of each row), xw, yw."""
        wall_ux = tuple(self.f64(x) for x in wall_ux)
        pr, mu_ref = self.f64(pr), self.f64(mu_ref)
        rec = self.limited(h, b)
        wf, wn, wt, _vol = self.reconstruct(W)                    # [ny, nx, 4, 4(side), 3]
        nx, ny = self.nx, self.ny
        east = torch.roll(torch.arange(nx, device=self.dev), -1)
        xh = xb = xw = yh = yb = yw = 0.0
        for a in range(3):
            L = {k: tuple(t[:, :, :, 1, a] for t in rec[k]) for k in ("h", "b")}
            R = {k: tuple(t[:, east, :, 0, a] for t in rec[k]) for k in ("h", "b")}
            mnl = wn[:, :, :, 1, a] / self.dx
            mnr = wn[:, east, :, 0, a] / self.dx
            mtg = 0.5 * (wt[:, :, :, 1, a] + wt[:, east, :, 0, a]) / self.dy
            fh, fb, fw = self.face_node(1, dt, L, R, mnl, mnr, mtg, pr, mu_ref)
            xh, xb, xw = xh + GW[a] * fh, xb + GW[a] * fb, xw + GW[a] * fw
            L = {k: tuple(t[:-1, :, :, 3, a] for t in rec[k]) for k in ("h", "b")}
            R = {k: tuple(t[1:, :, :, 2, a] for t in rec[k]) for k in ("h", "b")}
            mnl = wn[:-1, :, :, 3, a] / self.dy
            mnr = wn[1:, :, :, 2, a] / self.dy
            mtg = 0.5 * (wt[:-1, :, :, 3, a] + wt[1:, :, :, 2, a]) / self.dx
            fh, fb, fw = self.face_node(2, dt, L, R, mnl, mnr, mtg, pr, mu_ref)
            yh, yb, yw = yh + GW[a] * fh, yb + GW[a] * fb, yw + GW[a] * fw
        walls = []
        for top in (0, 1):
            j, sd, side = (ny - 1, 3, 1.0) if top else (0, 2, -1.0)
            acc_h = acc_b = acc_w = 0.0
            for a in range(3):
                d = {k: tuple(t[j, :, :, sd, a] for t in rec[k]) for k in ("h", "b")}
                fh, fb, fw = self.wall_node(side, dt, W[j], d["h"][0], d["b"][0], d["h"][1], d["b"][1], d["h"][2],
                                            d["b"][2], wall_ux[top], 0.5 * self.case.wall_tn[top], pr, mu_ref)
                acc_h, acc_b, acc_w = acc_h + GW[a] * fh, acc_b + GW[a] * fb, acc_w + GW[a] * fw
            walls.append((acc_h, acc_b, acc_w))
        yh = torch.cat([walls[0][0][None], yh, walls[1][0][None]], 0)
        yb = torch.cat([walls[0][1][None], yb, walls[1][1][None]], 0)
        yw = torch.cat([walls[0][2][None], yw, walls[1][2][None]], 0)
        return xh, xb, xw, yh, yb, yw

    # -- one S2CN2 step --------------------------------------------------------------
    def step(self, W, h, b, wall_u, mu_ref, pr, dt=None):
        """One S2CN2 step (or_cn_step without the accept/reject gates).
        W [ny, nx, 4], h, b [ny, nx, n]; wall_u: wall speed U (walls at -U, +U);
        returns (W', h', b', wall_fw [2, nx, 4], dt)."""
        if dt is None:
            dt = self.time_step(W)
        wall_u, mu_ref, pr = (self.f64(x) for x in (wall_u, mu_ref, pr))
        wall_ux = (-wall_u, wall_u)
        ck = self.ck
        xh, xb, xw, yh, yb, yw = self.fluxes(W, h, b, dt, wall_ux, pr, mu_ref)
        west = torch.roll(torch.arange(self.nx, device=self.dev), 1)
        dx, dy, area = self.dx, self.dy, self.dx * self.dy

        def div(x, y):
            return (x[:, west] * dy - x * dy + y[:-1] * dx - y[1:] * dx) / area

        wn = W + div(xw, yw)
        hs = h + div(xh, yh)
        bs = b + div(xb, yb)
        mh, mb, kx, ky, _ = self.targets(W, pr, False)
        mhn, mbn, kxn, kyn, kap = self.targets(wn, pr, True)
        beta_old = (dt / (2.0 * self.tau(W, mu_ref)))[..., None]
        beta = (dt / (2.0 * self.tau(wn, mu_ref)))[..., None]
        rold = self.heat(W, h - mh, b - mb)
        rh = hs + beta_old * (mh + kx[0] * rold[..., 0:1] + ky[0] * rold[..., 1:2] - h)
        rb = bs + beta_old * (mb + kx[1] * rold[..., 0:1] + ky[1] * rold[..., 1:2] - b)
        rhs = self.heat(wn, rh, rb) - self.heat(wn, mhn, mbn)
        be = beta[..., 0]
        a11 = 1.0 + be - be * kap[..., 0]
        a21 = -be * kap[..., 1]
        a12 = -be * kap[..., 2]
        a22 = 1.0 + be - be * kap[..., 3]
        det = a11 * a22 - a12 * a21
        r0 = (rhs[..., 0] * a22 - a12 * rhs[..., 1]) / det
        r1 = (a11 * rhs[..., 1] - rhs[..., 0] * a21) / det
        eqh = mhn + kxn[0] * r0[..., None] + kyn[0] * r1[..., None]
        eqb = mbn + kxn[1] * r0[..., None] + kyn[1] * r1[..., None]
        hn = (rh + beta * eqh) / (1.0 + beta)
        bn = (rb + beta * eqb) / (1.0 + beta)
        _ = ck
        return wn, hn, bn, torch.stack([yw[0], yw[-1]], 0), dt


# ---------------------------------------------------------------------------
# scalar helpers of the wall (vectorised over faces)
# ---------------------------------------------------------------------------
def micro_slope(ck, prim, sw):
    """MicroSlope, src/mod_S2MultiscaleWall.f90:240-252; prim tuple of [F], sw [F, 4]."""
    r, U, V, lam = prim
    s0, s1, s2, s3 = sw.unbind(-1)
    a3 = 4.0 * lam * lam / ((ck + 2) * r) * (2.0 * s3 - 2.0 * U * s1 - 2.0 * V * s2
                                             + s0 * (U * U + V * V - 0.5 * (ck + 2) / lam))
    a2 = 2.0 * lam / r * (s2 - V * s0) - V * a3
    a1 = 2.0 * lam / r * (s1 - U * s0) - U * a3
    a0 = s0 / r - U * a1 - V * a2 - 0.5 * (U * U + V * V + 0.5 * (ck + 2) / lam) * a3
    return torch.stack([a0, a1, a2, a3], -1)


def calc_moment(ck, prim):
    """CalcMoment, src/mod_S2MultiscaleWall.f90:261-292 (lists of [F] tensors)."""
    _r, U, V, lam = prim
    sl = torch.sqrt(lam)
    g = 0.5 * torch.exp(-lam * U * U) / torch.sqrt(PI * lam)
    mul = [0.5 * torch.special.erfc(-sl * U)]
    mur = [0.5 * torch.special.erfc(sl * U)]
    mul.append(U * mul[0] + g)
    mur.append(U * mur[0] - g)
    for i in range(2, 7):
        mul.append(U * mul[i - 1] + 0.5 * (i - 1) * mul[i - 2] / lam)
        mur.append(U * mur[i - 1] + 0.5 * (i - 1) * mur[i - 2] / lam)
    mu = [a + b for a, b in zip(mul, mur)]
    mv = [torch.ones_like(U), V + 0.0 * U]
    for i in range(2, 6):
        mv.append(V * mv[i - 1] + 0.5 * (i - 1) * mv[i - 2] / lam)
    mxi = [torch.ones_like(U), 0.5 * ck / lam, ck * (ck + 2.0) / (4.0 * lam * lam)]
    return mu, mv, mxi


def _uvxi(mu, mv, mxi, al, be, de):
    return torch.stack([mu[al] * mv[be] * mxi[de // 2], mu[al + 1] * mv[be] * mxi[de // 2],
                        mu[al] * mv[be + 1] * mxi[de // 2],
                        0.5 * (mu[al + 2] * mv[be] * mxi[de // 2] + mu[al] * mv[be + 2] * mxi[de // 2]
                               + mu[al] * mv[be] * mxi[(de + 2) // 2])], -1)


def moment_auvxi(a, mu, mv, mxi, al, be):
    """Moment_auvxi, src/mod_S2MultiscaleWall.f90:323-336."""
    t = [_uvxi(mu, mv, mxi, al, be, 0), _uvxi(mu, mv, mxi, al + 1, be, 0), _uvxi(mu, mv, mxi, al, be + 1, 0),
         _uvxi(mu, mv, mxi, al + 2, be, 0), _uvxi(mu, mv, mxi, al, be + 2, 0), _uvxi(mu, mv, mxi, al, be, 2)]
    return (a[:, 0:1] * t[0] + a[:, 1:2] * t[1] + a[:, 2:3] * t[2] + 0.5 * a[:, 3:4] * t[3]
            + 0.5 * a[:, 3:4] * t[4] + 0.5 * a[:, 3:4] * t[5])


def wall_time_coeffs(dt, tau):
    """s2w2_time_coefficients (src/mod_S2MultiscaleWall.f90:18-43): the
    24-term series for dt/tau < 0.1, the closed form otherwise; tau [F]."""
    x = dt / tau
    # closed form
    ex = torch.exp(-x)
    c3 = tau * (1.0 - ex)
    c4 = -tau * dt * ex + tau * c3
    c0 = dt - c3
    c1 = -tau * c0 + c4
    c2 = 0.5 * dt * dt - tau * c0
    # series
    t = torch.ones_like(x)
    s = [torch.zeros_like(x) for _ in range(5)]
    for n in range(1, 25):
        t = t * (-x) / n
        s[3] = s[3] - tau * t
        if n >= 2:
            s[0] = s[0] + tau * t
            s[4] = s[4] + tau * tau * (n - 1) * t
        if n >= 3:
            s[1] = s[1] + tau * tau * (n - 2) * t
            s[2] = s[2] - tau * tau * t
    small = x < 0.1
    return [torch.where(small, a, b) for a, b in zip(s, (c0, c1, c2, c3, c4))]


class ClosureTheta(torch.autograd.Function):
    """theta(w) of the discrete conservative equilibrium; backward by the
    implicit-function theorem: g_w = J(theta)^{-T} g_theta."""

    @staticmethod
    def forward(ctx, w, model):
        th, ok = model.closure_newton(w.detach())
        if not bool(ok.all()):
            raise FloatingPointError("discrete-equilibrium Newton closure failed")
        with torch.no_grad():
            J = model.q_state(th)[4]
        ctx.save_for_backward(J)
        return th

    @staticmethod
    def backward(ctx, g):
        (J,) = ctx.saved_tensors
        return torch.linalg.solve(J.transpose(-1, -2), g[..., None])[..., 0], None
