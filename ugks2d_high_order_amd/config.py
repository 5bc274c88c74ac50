"""Couette case parameters of the HighOrderImplicit entry.

Mirrors the subset of the reference's Config.ini namelists
(&config_para/&boundary_para, src/mod_ImportPara.f90:141-173) that the
high-order implicit pseudo-time step reads, with the derived constants of
set_const_para (src/mod_ImportPara.f90:911-989): gamma = (ck+4)/(ck+2),
Kn = sqrt(pi)*mu_ref, wall T_native per side (ResolveS2WallTemperatures,
src/mod_ImportPara.f90:275-298).  The mesh is the uniform Cartesian Couette
mesh of S2CartesianMesh on [0,1]^2 (src/mod_InitializeFlow.f90:278-290):
dx = 1/nx, dy = 1/ny, x periodic, diffuse walls at y = 0 and y = 1.
"""
from __future__ import annotations

import dataclasses
import math
import re


@dataclasses.dataclass
class CouetteCase:
    nx: int = 8
    ny: int = 64
    nu: int = 32
    nv: int = 32
    table: str = "half32"          # velocity table name (velocity.table)
    table_scale: float = 2.0       # S2VelocityScale
    ck: int = 1
    pr: float = 2.0 / 3.0
    omega: float = 0.5
    mu_ref: float = 0.05637282460904075
    cfl: float = 0.1
    ncfl: float = 1000.0
    macro_sweeps: int = 2
    micro_sweeps: int = 1
    mass_rescale: bool = True
    wall_ux: tuple = (-3.872983346207417, 3.872983346207417)   # WallUx(3:4)
    wall_tn: tuple = (2.0, 2.0)                                # wall T_native (Tw)
    t_init: float = 2.0                                        # Tinf (T_native)
    min_alpha: float = 1.0e-6
    moment_tol: float = 1.0e-12
    residual_tol: float = 1.0e-8
    max_iter: int = 200

    @property
    def dx(self) -> float:
        return 1.0 / self.nx

    @property
    def dy(self) -> float:
        return 1.0 / self.ny

    @property
    def nvel(self) -> int:
        return self.nu * self.nv

    @property
    def gamma(self) -> float:
        return (self.ck + 4) / (self.ck + 2)

    @property
    def kn(self) -> float:
        return math.sqrt(math.pi) * self.mu_ref

    def cv_per_iteration(self) -> int:
        return self.nx * self.ny * self.nu * self.nv


def wall_speed(mach: float, t_native: float, ck: int = 1) -> float:
    """Wall speed for a wall Mach number on the wall sound speed:
    c = sqrt(gamma*theta), theta = T_native/2 (p = rho*T_native/2)."""
    gamma = (ck + 4) / (ck + 2)
    return mach * math.sqrt(gamma * 0.5 * t_native)


def shipped_case(**kw) -> CouetteCase:
    """The reference's shipped HighOrderImplicit Couette case (Config.ini)."""
    return dataclasses.replace(CouetteCase(), **kw)


def couette_ma3_kn01(**kw) -> CouetteCase:
    """8x64 cells x 32x32 velocities, Ma=3 walls, Kn=0.1, omega=0.81, Pr=2/3."""
    c = CouetteCase(omega=0.81, mu_ref=0.1 / math.sqrt(math.pi), max_iter=200)
    return dataclasses.replace(c, **kw)


def couette_large(**kw) -> CouetteCase:
    """64x512 cells x 64x64 velocities, Ma=3 walls, Kn=0.05, 500 iterations."""
    u = wall_speed(3.0, 2.0)
    c = CouetteCase(nx=64, ny=512, nu=64, nv=64, table="half64", omega=0.81,
                    mu_ref=0.05 / math.sqrt(math.pi), wall_ux=(-u, u), max_iter=500)
    return dataclasses.replace(c, **kw)


def small_case(**kw) -> CouetteCase:
    """Reduced grid for fast CPU parity: 8x16 cells x 16x16 velocities."""
    c = CouetteCase(nx=8, ny=16, nu=16, nv=16, table="half16")
    return dataclasses.replace(c, **kw)


def batched_cases(n: int = 64, seed: int = 20261203, nx: int = 8, ny_choices=(32, 64, 96, 128), nu: int = 32,
                  table: str = "half32", t_init: float = 2.0, **kw) -> list:
    """Ragged batch of closed Couette instances drawn from a seed (NumPy
    PCG64): ny uniform over ny_choices, Kn log-uniform in [0.01, 1]
    (mu_ref = Kn/sqrt(pi)), wall Mach uniform in [0.5, 3] on the wall sound
    speed, and wall/initial temperature ratio uniform in [0.8, 1.2] with the
    initial gas at t_init.  Walls move at -U (bottom) and +U (top)."""
    import numpy as np
    rng = np.random.Generator(np.random.PCG64(seed))
    ny = rng.choice(np.asarray(ny_choices), size=n)
    kn = np.exp(rng.uniform(math.log(0.01), math.log(1.0), size=n))
    ma = rng.uniform(0.5, 3.0, size=n)
    tr = rng.uniform(0.8, 1.2, size=n)
    out = []
    for i in range(n):
        tw = float(tr[i]) * t_init
        u = wall_speed(float(ma[i]), tw)
        c = CouetteCase(nx=nx, ny=int(ny[i]), nu=nu, nv=nu, table=table, omega=0.81,
                        mu_ref=float(kn[i]) / math.sqrt(math.pi), wall_ux=(-u, u), wall_tn=(tw, tw), t_init=t_init,
                        max_iter=300)
        out.append(dataclasses.replace(c, **kw))
    return out


_NAMELIST_KEYS = {
    "nx_all": ("nx", int), "ny_all": ("ny", int), "nu": ("nu", int), "nv": ("nv", int),
    "ck": ("ck", int), "pr": ("pr", float), "omega": ("omega", float), "couettemuref": ("mu_ref", float),
    "cfl": ("cfl", float), "implicit_ncfl": ("ncfl", float), "implicit_macrosweeps": ("macro_sweeps", int),
    "implicit_microsweeps": ("micro_sweeps", int), "implicit_maxiter": ("max_iter", int),
    "implicit_minalpha": ("min_alpha", float), "implicit_momenttol": ("moment_tol", float),
    "implicit_residualtol": ("residual_tol", float), "tw": ("_tw", float), "tinf": ("t_init", float),
    "s2velocitytable": ("table", str), "s2velocityscale": ("table_scale", float),
    "closedcouettemassrescale": ("mass_rescale", bool),
}


def _parse_value(raw: str, kind):
    raw = raw.strip().rstrip(",")
    if kind is bool:
        return raw.lower().strip(".") in ("true", "t")
    if kind is str:
        return raw.strip("'\"")
    raw = raw.lower().replace("d", "e")
    return kind(float(raw)) if kind is int else float(raw)


def from_config_ini(text: str) -> CouetteCase:
    """Read the keys above from a reference Config.ini (Fortran namelist text)."""
    case = CouetteCase()
    tw = None
    wall = None
    for line in text.splitlines():
        line = line.split("!")[0].strip()
        m = re.match(r"([A-Za-z_0-9]+)\s*=\s*(.+)$", line)
        if not m:
            continue
        key, val = m.group(1).lower(), m.group(2)
        if key == "wallux":
            parts = [float(t.lower().replace("d", "e")) for t in val.split()]
            wall = (parts[2], parts[3])
            continue
        if key not in _NAMELIST_KEYS:
            continue
        attr, kind = _NAMELIST_KEYS[key]
        v = _parse_value(val, kind)
        if attr == "_tw":
            tw = v
        else:
            setattr(case, attr, v)
    if tw is not None:
        case.wall_tn = (tw, tw)
    if wall is not None:
        case.wall_ux = wall
    return case


def adjoint_case(**kw) -> CouetteCase:
    """Adjoint workload: 16x128 cells x 32x32 velocities, walls at -U, +U
    with U at Ma = 2 on the wall sound speed, Kn = 0.1, omega = 0.81."""
    u = wall_speed(2.0, 2.0)
    c = CouetteCase(nx=16, ny=128, nu=32, nv=32, table="half32", omega=0.81,
                    mu_ref=0.1 / math.sqrt(math.pi), wall_ux=(-u, u))
    return dataclasses.replace(c, **kw)
