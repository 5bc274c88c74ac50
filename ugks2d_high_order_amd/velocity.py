"""Discrete velocity tables for the S2 operator.

The reference's structured tables (src/mod_VelocityTables.f90:36-121) mirror an
n-node half-range Gauss-Hermite rule for the weight exp(-x^2) on [0, inf) onto
the negative axis, then change variable u = sqrt(2)*scale*x, so that the
table weight is sqrt(2)*scale*exp(x^2)*w_x and integrates f(u) du directly.
The 2-D table is the tensor product with the u index fastest,
k = i + nu*j (src/mod_VelocityTables.f90:116-118).

'half16' and 'half32' are the 8- and 16-node half-range rules the reference
tabulates; 'half64' (32 nodes per half line) is the 64x64 grid of the large
Couette workload, which the reference cannot select (it only knows 16 and 32
points) and which this module generates from the same family.

The half-range rule is computed here from the moments of the weight
(mu_k = Gamma((k+1)/2)/2) with the Stieltjes/Lanczos recurrence in high
precision (mpmath) and the Golub-Welsch eigenproblem; results are rounded to
float64 only at the end.
"""
from __future__ import annotations

import functools
import os

import numpy as np

_CACHE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


@functools.lru_cache(maxsize=None)
def half_range_gauss_hermite(n: int, dps: int = 160):
    """n-node Gauss rule for int_0^inf exp(-x^2) f(x) dx: (nodes, weights) ascending."""
    import mpmath as mp

    with mp.workdps(dps):
        # Discretized Stieltjes procedure on a fine Gauss-Laguerre-like grid is
        # avoided: the modified Chebyshev algorithm on the exact moments is
        # stable enough at this precision for n <= 64.
        m = 2 * n
        mom = [mp.gamma(mp.mpf(k + 1) / 2) / 2 for k in range(m)]
        # Chebyshev algorithm (Gautschi, "Orthogonal Polynomials", Alg. 2.1)
        a = [mp.mpf(0)] * n
        bcoef = [mp.mpf(0)] * n
        sig_prev = [mp.mpf(0)] * m
        sig = [mom[l] for l in range(m)]
        a[0] = mom[1] / mom[0]
        bcoef[0] = mom[0]
        for k in range(1, n):
            sig_new = [mp.mpf(0)] * m
            for l in range(k, m - k):
                sig_new[l] = sig[l + 1] - a[k - 1] * sig[l] - bcoef[k - 1] * sig_prev[l]
            a[k] = sig_new[k + 1] / sig_new[k] - sig[k] / sig[k - 1]
            bcoef[k] = sig_new[k] / sig[k - 1]
            sig_prev, sig = sig, sig_new
        J = mp.zeros(n, n)
        for i in range(n):
            J[i, i] = a[i]
            if i + 1 < n:
                J[i, i + 1] = J[i + 1, i] = mp.sqrt(bcoef[i + 1])
        E, Q = mp.eigsy(J)
        nodes = [E[i] for i in range(n)]
        weights = [mom[0] * Q[0, i] ** 2 for i in range(n)]
        order = sorted(range(n), key=lambda i: nodes[i])
        x = np.array([float(nodes[i]) for i in order])
        w = np.array([float(weights[i]) for i in order])
        # exp(x^2)*w in high precision, for the change of variable
        ew = np.array([float(mp.exp(nodes[i] ** 2) * weights[i]) for i in order])
    return x, w, ew


def structured_table(n_half: int, scale: float = 2.0):
    """2-D table (u, v, weight), each of length (2*n_half)**2, u index fastest."""
    x, _w, ew = half_range_gauss_hermite(n_half)
    raw = np.concatenate([-x[::-1], x])
    rew = np.concatenate([ew[::-1], ew])
    ue = np.sqrt(2.0) * scale * raw
    we = np.sqrt(2.0) * scale * rew
    n = ue.size
    u = np.tile(ue, n)
    v = np.repeat(ue, n)
    weight = np.outer(we, we).reshape(-1)   # weight[j*n+i] = we[j]*we[i]
    return u, v, weight


def table(name: str, scale: float = 2.0):
    """Named table: 'half16', 'half32' or 'half64' (points per axis)."""
    per_axis = {"half16": 16, "half32": 32, "half64": 64}[name]
    path = os.path.join(_CACHE_DIR, f"{name}_scale{scale:g}.npz")
    if os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        return z["u"], z["v"], z["w"]
    u, v, w = structured_table(per_axis // 2, scale)
    try:
        os.makedirs(_CACHE_DIR, exist_ok=True)
        np.savez(path, u=u, v=v, w=w)
    except OSError:
        pass
    return u, v, w


def write_structured_file(path: str, u, v, w, nu: int, nv: int) -> None:
    """Text table in the reference's StructuredVelocityFile format
    (src/mod_InitializeFlow.f90:480-492): header 'Nu Nv', then 'u v w' per k."""
    with open(path, "w") as f:
        f.write(f"{nu:12d}{nv:12d}\n")
        for a, b, c in zip(u, v, w):
            f.write(f"{a:26.17E}{b:26.17E}{c:26.17E}\n")


def read_structured_file(path: str):
    with open(path) as f:
        nu, nv = (int(t) for t in f.readline().split())
        data = np.loadtxt(f)
    assert data.shape == (nu * nv, 3)
    return nu, nv, data[:, 0].copy(), data[:, 1].copy(), data[:, 2].copy()
