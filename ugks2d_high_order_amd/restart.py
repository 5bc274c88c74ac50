"""On-disk restart and monitor formats of the UGKS2D Couette runs.

A GPU run restarts from the files a CPU run of the reference leaves, and
leaves files the reference restarts from (resume.sh copies exactly these):

  Mesh2d.dat              Fortran sequential: (nx+1, ny+1), then the node
                          coordinates x(i,j), y(i,j)    (src/mod_InitializeFlow.f90:277-310)
  flow2d.dat              Fortran sequential: 1 / (nx, ny, 4) / W(i,j,1:4)
                          (write_restart, src/mod_Output.f90:571-676;
                          InitFlow_Restart, src/mod_InitializeFlow.f90:973-1057)
  vdf2d.dat               Fortran sequential: 1 / (nx, ny, 2 Nvel) / h(i,j,k), b(i,j,k)
                          (src/mod_Output.f90:744-812; src/mod_InitializeFlow.f90:1136-1190)
  grid2d.grd              Fortran sequential: 1 / (nx, ny) / cell centres (visualisation only)
  closed_mass_rank<r>.bin stream: version 4, nx, ny, Nvel, LAP, cumulative step,
                          M0, correction ledger(4), RestartIdentity(31), then
                          W, h, b of every cell of the rank's block with its
                          LAP = 2 halo (WriteMassRestart / ReadMassRestart,
                          src/mod_ClosedCouetteMass.f90:315-393)
  residual_reference.in   text: epoch, reference_h, reference_b,
                          Implicit_ResidualTol, Implicit_MomentTol
                          (WriteRestartReference, src/mod_Output.f90:58-76;
                          StartR5Control, src/mod_R5Observe.f90:163-190)

and the monitors a run appends to:

  sparse_residual.csv, consecutive_states.csv (WriteMonitorRows, src/mod_R5Observe.f90:322-355)
  ResidualImplicit.plt and the "Implicit iter=" line (WriteImplicitResidual,
                          src/mod_ImplicitDiagnostics.f90:360-397)
  R5_status.txt           (FinalR5State, src/mod_R5Observe.f90:466-481)
  Flow2d_<iter>.plt, Residual2d_<iter>.plt  Tecplot BLOCK fields of write_flow
                          (src/mod_Output.f90:83-118,262-280); their numbers
                          are written as ES25.16E3, four to a line, where the
                          reference writes list-directed (the values agree,
                          the spacing is the compiler's own)

Layouts.  The state of an instance is W[c, 4], h[c, Nvel], b[c, Nvel] with
c = j*nx + i (i along the walls); the reference's arrays are column-major
with i fastest, so W(i,j,m) is W.reshape(ny, nx, 4)[j, i, m].  Fortran
sequential records are framed the way gfortran writes them: a 4-byte
little-endian byte count before and after each record; a record longer than
2147483639 bytes is split into subrecords whose leading count is negative
when another subrecord follows and whose trailing count is negative when
one precedes.  Stream files have no framing.
"""
from __future__ import annotations

import dataclasses
import glob
import math
import os
import re
import struct

import numpy as np

LAP = 2                       # halo width of the reference's blocks (mod_Globaldata.f90:95)
MASS_VERSION = 4              # closed_mass_rank*.bin layout version
MAX_SUBRECORD = 2147483639    # largest subrecord gfortran writes


class RestartError(RuntimeError):
    """A restart file is missing, malformed or belongs to another run."""


# ---------------------------------------------------------------------------
# Fortran sequential records
# ---------------------------------------------------------------------------
def _view(p):
    if isinstance(p, (bytes, bytearray, memoryview)):
        return memoryview(p).cast("B")
    return memoryview(np.ascontiguousarray(p)).cast("B")


def write_record(f, parts, total: int | None = None, max_subrecord: int | None = None):
    """Write one sequential record made of the concatenated byte images of
    `parts` (numpy arrays or bytes).  With `total` (the record's byte count)
    `parts` may be a generator: the parts are then streamed one at a time."""
    max_subrecord = max_subrecord or MAX_SUBRECORD
    if total is None:
        parts = [_view(p) for p in parts]
        total = sum(len(v) for v in parts)
    nsub = max(1, -(-total // max_subrecord))
    sub, left = 0, min(max_subrecord, total)
    f.write(struct.pack("<i", -left if nsub > 1 else left))
    written = 0
    for p in parts:
        v = _view(p)
        off = 0
        while off < len(v):
            if left == 0:
                # close this subrecord and open the next
                f.write(struct.pack("<i", -max_subrecord if sub > 0 else max_subrecord))
                sub += 1
                left = min(max_subrecord, total - written)
                f.write(struct.pack("<i", -left if sub + 1 < nsub else left))
            take = min(left, len(v) - off)
            f.write(v[off:off + take])
            off += take
            left -= take
            written += take
    if written != total or left != 0:
        raise RestartError("record parts do not add up to the declared size")
    last = min(max_subrecord, total - max_subrecord * (nsub - 1))
    f.write(struct.pack("<i", -last if sub > 0 else last))


def read_record(f) -> bytes:
    """Read one sequential record (joining its subrecords)."""
    chunks, first = [], True
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise RestartError("unexpected end of file in a record marker")
        n = struct.unpack("<i", head)[0]
        data = f.read(abs(n))
        tail = f.read(4)
        if len(data) != abs(n) or len(tail) < 4:
            raise RestartError("truncated record")
        t = struct.unpack("<i", tail)[0]
        if abs(t) != abs(n) or (t < 0) == first:
            raise RestartError("inconsistent record markers")
        chunks.append(data)
        first = False
        if n >= 0:
            return b"".join(chunks) if len(chunks) > 1 else chunks[0]


def read_real_record(f, n: int):
    """Read one sequential record that holds n doubles straight into an array
    (the subrecords are read in place, with no joined copy)."""
    out = np.empty(n, dtype="<f8")
    buf = memoryview(out).cast("B")
    pos, first = 0, True
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise RestartError("unexpected end of file in a record marker")
        m = struct.unpack("<i", head)[0]
        if pos + abs(m) > len(buf):
            raise RestartError(f"record holds more than {8 * n} bytes, expected {8 * n}")
        if f.readinto(buf[pos:pos + abs(m)]) != abs(m):
            raise RestartError("truncated record")
        pos += abs(m)
        tail = f.read(4)
        if len(tail) < 4:
            raise RestartError("truncated record")
        t = struct.unpack("<i", tail)[0]
        if abs(t) != abs(m) or (t < 0) == first:
            raise RestartError("inconsistent record markers")
        first = False
        if m >= 0:
            break
    if pos != len(buf):
        raise RestartError(f"record holds {pos} bytes, expected {8 * n}")
    return out


def read_records(path: str):
    with open(path, "rb") as f:
        size = os.fstat(f.fileno()).st_size
        out = []
        while f.tell() < size:
            out.append(read_record(f))
    return out


def _ints(rec: bytes):
    if len(rec) % 4:
        raise RestartError("integer record of odd length")
    return struct.unpack("<%di" % (len(rec) // 4), rec)


def _reals(rec: bytes, n: int):
    if len(rec) != 8 * n:
        raise RestartError(f"record holds {len(rec)} bytes, expected {8 * n}")
    return np.frombuffer(rec, dtype="<f8")


# ---------------------------------------------------------------------------
# Mesh2d.dat, flow2d.dat, vdf2d.dat, grid2d.grd
# ---------------------------------------------------------------------------
def read_mesh2d(path: str):
    """Node coordinates x[j, i], y[j, i] (j = 0..ny, i = 0..nx) of Mesh2d.dat."""
    recs = read_records(path)
    if len(recs) < 2:
        raise RestartError("Mesh2d.dat needs a size record and a coordinate record")
    n0, m0 = _ints(recs[0])[:2]
    xy = _reals(recs[1], 2 * n0 * m0)
    return xy[:n0 * m0].reshape(m0, n0).copy(), xy[n0 * m0:].reshape(m0, n0).copy()


def write_mesh2d(path: str, x, y):
    x = np.asarray(x, dtype="<f8")
    y = np.asarray(y, dtype="<f8")
    m0, n0 = x.shape
    with open(path, "wb") as f:
        write_record(f, [np.array([n0, m0], dtype="<i4")])
        write_record(f, [x, y])


def cartesian_mesh(x, y):
    """(nx, ny, dx, dy) of a uniform Cartesian node grid; the closed-Couette
    path requires one (InitializeClosedMass, src/mod_ClosedCouetteMass.f90:212-223)."""
    ny, nx = x.shape[0] - 1, x.shape[1] - 1
    if not (np.all(np.abs(x - x[0:1, :]) <= 1e-14) and np.all(np.abs(y - y[:, 0:1]) <= 1e-14)):
        raise RestartError("requires verified Cartesian Couette geometry")
    dxs, dys = np.diff(x[0]), np.diff(y[:, 0])
    if np.ptp(dxs) > 1e-14 * abs(dxs[0]) or np.ptp(dys) > 1e-14 * abs(dys[0]):
        raise RestartError("the GPU path takes uniform cells only")
    return nx, ny, float(x[0, -1] - x[0, 0]) / nx, float(y[-1, 0] - y[0, 0]) / ny


def couette_nodes(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0):
    """Nodes of the uniform nx x ny mesh of [0,lx] x [0,ly] (x[j, i], y[j, i])."""
    xi = np.arange(nx + 1, dtype=np.float64) * (lx / nx)
    yj = np.arange(ny + 1, dtype=np.float64) * (ly / ny)
    return np.broadcast_to(xi, (ny + 1, nx + 1)).copy(), np.broadcast_to(yj[:, None], (ny + 1, nx + 1)).copy()


def cell_centres(x, y):
    """0.25 (four corners), summed in the reference's order (src/mod_InitializeFlow.f90:306-310)."""
    xc = 0.25 * (((x[:-1, :-1] + x[:-1, 1:]) + x[1:, :-1]) + x[1:, 1:])
    yc = 0.25 * (((y[:-1, :-1] + y[:-1, 1:]) + y[1:, :-1]) + y[1:, 1:])
    return xc, yc


def write_flow2d(path: str, W, nx: int, ny: int):
    W = np.asarray(W, dtype=np.float64).reshape(ny, nx, 4)
    with open(path, "wb") as f:
        write_record(f, [np.array([1], dtype="<i4")])
        write_record(f, [np.array([nx, ny, 4], dtype="<i4")])
        write_record(f, [np.ascontiguousarray(W.transpose(2, 0, 1), dtype="<f8")])


def read_flow2d(path: str, nx: int, ny: int):
    """W[c, 4] of flow2d.dat; the reference's conformity check on the sizes."""
    if not os.path.exists(path):
        raise RestartError("Restart error: file flow2d.dat does not exist.")
    with open(path, "rb") as f:
        read_record(f)
        n0 = _ints(read_record(f))
        if tuple(n0[:3]) != (nx, ny, 4):
            raise RestartError(f"Restart error: data in flow2d.dat does not conform. (nx0, ny0, nvar0) = {n0[:3]}")
        a = _reals(read_record(f), 4 * nx * ny).reshape(4, ny, nx)
    return np.ascontiguousarray(a.transpose(1, 2, 0)).reshape(nx * ny, 4)


def write_vdf2d(path: str, h, b, nx: int, ny: int, planes=None):
    """vdf2d.dat of the structured-quadrature branch.  Either `planes`
    ([2, Nvel, cell]: h then b as velocity planes, the file's order, e.g. from
    UGKS.state_planes) is written as is, or h and b ([cell, Nvel]) are
    transposed 64 velocity planes at a time (no full transposed copy)."""
    if planes is not None:
        planes = np.asarray(planes)
        nv = planes.shape[1]
        parts = [np.ascontiguousarray(planes, dtype="<f8")]
    else:
        nv = h.shape[-1]
        h = np.asarray(h).reshape(ny * nx, nv)
        b = np.asarray(b).reshape(ny * nx, nv)
        parts = (np.ascontiguousarray(a[:, k0:k0 + 64].T, dtype="<f8") for a in (h, b) for k0 in range(0, nv, 64))
    with open(path, "wb") as f:
        write_record(f, [np.array([1], dtype="<i4")])
        write_record(f, [np.array([nx, ny, 2 * nv], dtype="<i4")])
        write_record(f, parts, total=8 * 2 * nv * nx * ny)


def read_vdf2d_planes(path: str, nx: int, ny: int, nvel: int):
    """The third record of vdf2d.dat as planes[2, Nvel, cell] (h then b)."""
    if not os.path.exists(path):
        raise RestartError("Restart error: file vdf2d.dat does not exist.")
    with open(path, "rb") as f:
        read_record(f)
        n0 = _ints(read_record(f))
        if tuple(n0[:3]) != (nx, ny, 2 * nvel):
            raise RestartError(f"Restart error: data in vdf2d.dat does not conform. (nx0, ny0, nvar0) = {n0[:3]}")
        return read_real_record(f, 2 * nvel * nx * ny).reshape(2, nvel, nx * ny)


def read_vdf2d(path: str, nx: int, ny: int, nvel: int):
    """(h[c, Nvel], b[c, Nvel]) of vdf2d.dat."""
    a = read_vdf2d_planes(path, nx, ny, nvel)
    return np.ascontiguousarray(a[0].T), np.ascontiguousarray(a[1].T)


def write_grid2d(path: str, x, y):
    xc, yc = cell_centres(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
    ny, nx = xc.shape
    with open(path, "wb") as f:
        write_record(f, [np.array([1], dtype="<i4")])
        write_record(f, [np.array([nx, ny], dtype="<i4")])
        write_record(f, [xc.astype("<f8"), yc.astype("<f8")])


# ---------------------------------------------------------------------------
# closed_mass_rank<r>.bin
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class RunIdentity:
    """Configuration a complete restart is bound to (RestartIdentity,
    src/mod_ClosedCouetteMass.f90:315-323).  The keys the Couette case does
    not otherwise carry take the shipped case's values."""
    nx: int
    ny: int
    np_x: int
    np_y: int
    nu: int
    nv: int
    cfl: float
    ncfl: float
    ck: int
    pr: float
    omega: float
    tinf: float
    tw: float
    mu_ref: float
    wall_ux: tuple            # WallUx(1:4): left, right, bottom, top
    bc_type: tuple = (10, 10, 2, 2)
    flag_vislaw: int = 1
    fluxwall_type: int = 1
    acc: float = 1.0
    bg_geometry: int = 0
    quadrature: int = 2
    zhu_moving_wall: bool = True
    init_halo_resync: bool = True
    sum_w: float = 0.0        # sum(wVel)
    sum_wuu: float = 0.0      # sum(wVel*uVel*uVel)

    def vector(self):
        return np.array([self.nx, self.ny, self.np_x, self.np_y, self.nu, self.nv, self.cfl, self.ncfl, self.ck,
                         self.pr, self.omega, self.tinf, self.tw, self.mu_ref, *self.wall_ux, *self.bc_type,
                         self.flag_vislaw, self.fluxwall_type, self.acc, self.bg_geometry, self.quadrature,
                         1.0 if self.zhu_moving_wall else 0.0, 1.0 if self.init_halo_resync else 0.0,
                         self.sum_w, self.sum_wuu], dtype=np.float64)


def sequential_sum(a):
    """Left-to-right double sum (the order of a serial Fortran sum)."""
    s = 0.0
    for x in np.asarray(a, dtype=np.float64).ravel():
        s += float(x)
    return s


def run_identity(case, u, w, np_x: int = 1, np_y: int = 1, tw: float | None = None, **keys):
    """RestartIdentity of a Couette case run on an np_x x np_y decomposition."""
    u = np.asarray(u, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    return RunIdentity(nx=case.nx, ny=case.ny, np_x=np_x, np_y=np_y, nu=case.nu, nv=case.nv, cfl=case.cfl,
                       ncfl=case.ncfl, ck=case.ck, pr=case.pr, omega=case.omega, tinf=case.t_init,
                       tw=case.wall_tn[0] if tw is None else tw, mu_ref=case.mu_ref,
                       wall_ux=(0.0, 0.0, case.wall_ux[0], case.wall_ux[1]),
                       sum_w=sequential_sum(w), sum_wuu=sequential_sum(w * u * u), **keys)


def blocks(nx_all: int, ny_all: int, np_x: int, np_y: int):
    """(rank, origin_i, origin_j, ni, nj) of the reference's decomposition
    (1-based origins; the first mod(N, NP) blocks hold one more cell,
    src/mod_MPI.f90:36-50)."""
    out = []
    for kj in range(np_y):
        for ki in range(np_x):
            ni = nx_all // np_x + (1 if ki < nx_all % np_x else 0)
            nj = ny_all // np_y + (1 if kj < ny_all % np_y else 0)
            oi = (nx_all // np_x) * ki + min(ki, nx_all % np_x) + 1
            oj = (ny_all // np_y) * kj + min(kj, ny_all % np_y) + 1
            out.append((kj * np_x + ki, oi, oj, ni, nj))
    return out


def _runs(idx):
    """(position, start, stop) of the runs of consecutive values in idx."""
    out, p0 = [], 0
    for p in range(1, len(idx) + 1):
        if p == len(idx) or idx[p] != idx[p - 1] + 1:
            out.append((p0, int(idx[p0]), int(idx[p - 1]) + 1))
            p0 = p
    return out


def write_mass_restart(directory: str, W, h, b, identity: RunIdentity, m0: float, ledger, step: int):
    """closed_mass_rank<r>.bin of every rank of identity's decomposition.
    Each holds the cells (i, j) = 1-LAP..ni+LAP x 1-LAP..nj+LAP of its block
    as the reference holds them after its halo exchanges: x wraps
    periodically (BC 10/10 and the corner exchange), rows beyond the walls are
    zero (the wall-exterior slots are zero-initialised and never written,
    src/mod_InitializeFlow.f90:116-125).  The cells are streamed one halo row
    at a time through one reused row buffer."""
    nx_all, ny_all = identity.nx, identity.ny
    nv = h.shape[-1]
    Wg = np.asarray(W, dtype=np.float64).reshape(ny_all, nx_all, 4)
    hg = np.asarray(h, dtype=np.float64).reshape(ny_all, nx_all, nv)
    bg = np.asarray(b, dtype=np.float64).reshape(ny_all, nx_all, nv)
    ledger = np.asarray(ledger, dtype="<f8").reshape(4)
    for r, oi, oj, ni, nj in blocks(nx_all, ny_all, identity.np_x, identity.np_y):
        I = (np.arange(oi - LAP, oi + ni + LAP) - 1) % nx_all
        runs = _runs(I)
        row = np.empty((ni + 2 * LAP, 4 + 2 * nv), dtype="<f8")
        zero = np.zeros_like(row)
        with open(os.path.join(directory, f"closed_mass_rank{r}.bin"), "wb") as f:
            f.write(np.array([MASS_VERSION, ni, nj, nv, LAP, step], dtype="<i4").tobytes())
            f.write(np.array([m0], dtype="<f8").tobytes())
            f.write(ledger.tobytes())
            f.write(identity.vector().astype("<f8").tobytes())
            for j in range(oj - LAP - 1, oj + nj + LAP - 1):      # global 0-based rows
                if not 0 <= j < ny_all:
                    f.write(memoryview(zero).cast("B"))
                    continue
                for p, a, e in runs:
                    q = p + e - a
                    row[p:q, 0:4] = Wg[j, a:e]
                    row[p:q, 4:4 + nv] = hg[j, a:e]
                    row[p:q, 4 + nv:] = bg[j, a:e]
                f.write(memoryview(row).cast("B"))


@dataclasses.dataclass
class MassRestart:
    step: int
    m0: float
    ledger: np.ndarray
    identity: np.ndarray


def read_mass_restart(directory: str, W, h, b, identity: RunIdentity) -> MassRestart:
    """Read every closed_mass_rank<r>.bin of identity's decomposition and
    apply ReadMassRestart's checks: header, identity, valid M0, finite ledger,
    the physical cells equal to the flow2d/vdf2d state, and M0, ledger and
    step equal on all ranks."""
    nv = h.shape[-1]
    hdr_n, nid = 6 * 4 + 5 * 8, 31
    want = identity.vector()
    res = None
    Wg = np.asarray(W).reshape(identity.ny, identity.nx, 4)
    hg = np.asarray(h).reshape(identity.ny, identity.nx, nv)
    bg = np.asarray(b).reshape(identity.ny, identity.nx, nv)
    for r, oi, oj, ni, nj in blocks(identity.nx, identity.ny, identity.np_x, identity.np_y):
        path = os.path.join(directory, f"closed_mass_rank{r}.bin")
        if not os.path.exists(path):
            raise RestartError("missing, mismatched, or invalid M0-bearing restart; refusing to redefine M0")
        ncell = (ni + 2 * LAP) * (nj + 2 * LAP)
        need = hdr_n + 8 * nid + 8 * ncell * (4 + 2 * nv)
        size = os.path.getsize(path)
        if size != need:
            raise RestartError(f"{os.path.basename(path)}: {size} bytes, expected {need}")
        with open(path, "rb") as f:
            raw = f.read(hdr_n + 8 * nid)
            cells = np.fromfile(f, dtype="<f8")
        head = np.frombuffer(raw[:24], dtype="<i4")
        m0 = float(np.frombuffer(raw[24:32], dtype="<f8")[0])
        ledger = np.frombuffer(raw[32:64], dtype="<f8").copy()
        ident = np.frombuffer(raw[64:64 + 8 * nid], dtype="<f8")
        ok = (tuple(head[:5]) == (MASS_VERSION, ni, nj, nv, LAP) and np.array_equal(ident, want)
              and math.isfinite(m0) and m0 > 0.0 and np.all(np.isfinite(ledger)))
        if not ok:
            raise RestartError("missing, mismatched, or invalid M0-bearing restart; refusing to redefine M0")
        cells = cells.reshape(nj + 2 * LAP, ni + 2 * LAP, 4 + 2 * nv)
        inner = cells[LAP:LAP + nj, LAP:LAP + ni]
        sl = (slice(oj - 1, oj - 1 + nj), slice(oi - 1, oi - 1 + ni))
        if not (np.array_equal(inner[..., :4], Wg[sl]) and np.array_equal(inner[..., 4:4 + nv], hg[sl])
                and np.array_equal(inner[..., 4 + nv:], bg[sl])):
            raise RestartError("missing, mismatched, or invalid M0-bearing restart; refusing to redefine M0")
        this = MassRestart(step=int(head[5]), m0=m0, ledger=ledger, identity=ident.copy())
        if res is None:
            res = this
        elif res.m0 != this.m0:
            raise RestartError("restart M0 differs between ranks")
        elif not np.array_equal(res.ledger, this.ledger) or res.step != this.step:
            raise RestartError("restart ledger or origin step differs across ranks")
    return res


def mass_restart_ranks(directory: str):
    """Rank numbers of the closed_mass_rank<r>.bin files in a directory."""
    out = []
    for p in glob.glob(os.path.join(directory, "closed_mass_rank*.bin")):
        m = re.fullmatch(r"closed_mass_rank(\d+)\.bin", os.path.basename(p))
        if m:
            out.append(int(m.group(1)))
    return sorted(out)


# ---------------------------------------------------------------------------
# Fortran edit descriptors used by the text files
# ---------------------------------------------------------------------------
def _special(x: float, w: int):
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        t = ("Infinity" if w >= 9 else "Inf")
        return (("-" if x < 0 else ("+" if w > len(t) else "")) + t).rjust(w)
    return None


def _digits(x: float, nsig: int):
    """(sign, mantissa digits string of nsig digits, decimal exponent of the
    leading digit), correctly rounded."""
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign, "0" * nsig, 0
    m, e = f"{abs(x):.{nsig - 1}e}".split("e")
    return sign, m.replace(".", ""), int(e)


def fmt_es(x: float, w: int = 25, d: int = 16, e: int = 3) -> str:
    """Fortran ESw.dEe (one digit before the point, d after, e exponent digits)."""
    s = _special(x, w)
    if s is not None:
        return s
    sign, dig, ex = _digits(x, d + 1)
    t = f"{sign}{dig[0]}.{dig[1:]}E{'+' if ex >= 0 else '-'}{abs(ex):0{e}d}"
    return t.rjust(w) if len(t) <= w else "*" * w


def fmt_e1p(x: float, w: int, d: int) -> str:
    """Fortran 1P,Ew.d (scale factor 1: one digit before the point, d after;
    exponent E+dd, or +ddd without the letter beyond 99)."""
    s = _special(x, w)
    if s is not None:
        return s
    sign, dig, ex = _digits(x, d + 1)
    if abs(ex) <= 99:
        tail = f"E{'+' if ex >= 0 else '-'}{abs(ex):02d}"
    else:
        tail = f"{'+' if ex >= 0 else '-'}{abs(ex):03d}"
    t = f"{sign}{dig[0]}.{dig[1:]}{tail}"
    return t.rjust(w) if len(t) <= w else "*" * w


def fmt_i(n: int, w: int) -> str:
    t = str(int(n))
    return t.rjust(w) if len(t) <= w else "*" * w


# ---------------------------------------------------------------------------
# residual_reference.in
# ---------------------------------------------------------------------------
def write_residual_reference(path: str, epoch: int, ref_h: float, ref_b: float, residual_tol: float,
                             moment_tol: float):
    """'(i0,1x,4(es25.16e3,1x))' (src/mod_Output.f90:72-75)."""
    with open(path, "w") as f:
        f.write(f"{int(epoch)} " + "".join(fmt_es(v) + " " for v in (ref_h, ref_b, residual_tol, moment_tol)) + "\n")


def read_residual_reference(path: str, residual_tol: float | None = None, moment_tol: float | None = None,
                            origin: int | None = None):
    """(epoch, ref_h, ref_b, rtol, mtol), with StartR5Control's checks
    (src/mod_R5Observe.f90:176-187) for whichever of the expectations are given."""
    try:
        with open(path) as f:
            tok = f.read().replace(",", " ").split()
        epoch = int(tok[0])
        vals = [float(t.replace("D", "E").replace("d", "e")) for t in tok[1:5]]
    except (OSError, ValueError, IndexError) as exc:
        raise RestartError(f"cannot read residual reference inheritance file: {exc}") from None
    ref_h, ref_b, rtol, mtol = vals
    if (residual_tol is not None and rtol != residual_tol) or (moment_tol is not None and mtol != moment_tol):
        raise RestartError("reference inheritance convergence thresholds mismatch")
    if origin is not None and epoch > origin:
        raise RestartError("reference epoch after checkpoint")
    return epoch, ref_h, ref_b, rtol, mtol


# ---------------------------------------------------------------------------
# Complete restart of a GPU instance
# ---------------------------------------------------------------------------
def write_restart(directory: str, case, W, h, b, u, w, m0: float, ledger, ref_h: float, ref_b: float,
                  step: int, epoch: int = 0, np_x: int = 1, np_y: int = 1, nodes=None,
                  identity: RunIdentity | None = None, planes=None):
    """Write the complete restart of a HighOrderImplicit run after `step`
    cumulative updates (Output_restart_only, src/mod_Output.f90:38-56):
    Mesh2d.dat, flow2d.dat, vdf2d.dat, grid2d.grd, closed_mass_rank*.bin for
    an np_x x np_y decomposition and residual_reference.in (epoch kept)."""
    os.makedirs(directory, exist_ok=True)
    if nodes is None:
        nodes = couette_nodes(case.nx, case.ny, case.dx * case.nx, case.dy * case.ny)
    ident = identity or run_identity(case, u, w, np_x=np_x, np_y=np_y)
    write_mesh2d(os.path.join(directory, "Mesh2d.dat"), *nodes)
    write_flow2d(os.path.join(directory, "flow2d.dat"), W, case.nx, case.ny)
    write_vdf2d(os.path.join(directory, "vdf2d.dat"), h, b, case.nx, case.ny, planes=planes)
    write_grid2d(os.path.join(directory, "grid2d.grd"), *nodes)
    write_mass_restart(directory, W, h, b, ident, m0, ledger, step)
    write_residual_reference(os.path.join(directory, "residual_reference.in"), epoch, ref_h, ref_b,
                             case.residual_tol, case.moment_tol)
    return ident


def save_restart(solver, directory: str, step: int, epoch: int = 0, np_x: int = 1, np_y: int = 1, nodes=None,
                 identity: RunIdentity | None = None):
    """write_restart of a GPU instance's current state, mass target, ledger
    and residual references (the operator must have been evaluated once).
    vdf2d.dat is written from the device's velocity planes (state_planes), the
    companions from the [cell, Nvel] state."""
    W, planes = solver.state_planes()
    _, h, b = solver.state()
    m0, ledger = solver.mass_target()
    if not math.isfinite(m0):
        m0 = float(solver.mass_statistics()[0][0])
    ref_h, ref_b = solver.references()
    if not (math.isfinite(ref_h) and math.isfinite(ref_b)):
        raise RestartError("no residual references yet: evaluate the operator before saving")
    return write_restart(directory, solver.case, W, h, b, solver.u, solver.w, m0, ledger, ref_h, ref_b, step,
                         epoch, np_x, np_y, nodes, identity, planes=planes)


def load_restart(solver, directory: str, identity: RunIdentity | None = None):
    """Restart a GPU instance from a complete restart directory (Iflag_InitFlow
    = 2: InitFlow_Restart + ReadMassRestart + StartR5Control): the state from
    flow2d.dat/vdf2d.dat, M0 and the ledger from the closed_mass companions
    (whose decomposition is taken from the files present), the residual epoch
    from residual_reference.in.  Returns (cumulative step, epoch).

    The distributions go to the device as velocity planes (set_state_planes,
    transposed there), and the companions are checked against the [cell, Nvel]
    state read back from the device.  After a RestartError the instance holds
    the files' state but not their mass target or references: cold-start it
    or load a valid restart before stepping."""
    case = solver.case
    mesh = os.path.join(directory, "Mesh2d.dat")
    if os.path.exists(mesh):
        nx, ny, dx, dy = cartesian_mesh(*read_mesh2d(mesh))
        if (nx, ny) != (case.nx, case.ny) or abs(dx - case.dx) > 1e-14 * dx or abs(dy - case.dy) > 1e-14 * dy:
            raise RestartError("Mesh2d.dat does not match the case")
    W = read_flow2d(os.path.join(directory, "flow2d.dat"), case.nx, case.ny)
    planes = read_vdf2d_planes(os.path.join(directory, "vdf2d.dat"), case.nx, case.ny, case.nu * case.nv)
    if identity is None:
        ranks = mass_restart_ranks(directory)
        if not ranks:
            raise RestartError("Missing R6 mass companions")
        with open(os.path.join(directory, f"closed_mass_rank{ranks[0]}.bin"), "rb") as f:
            f.seek(64 + 8 * 2)
            np_x, np_y = (int(v) for v in np.frombuffer(f.read(16), dtype="<f8"))
        identity = run_identity(case, solver.u, solver.w, np_x=np_x, np_y=np_y)
    solver.set_state_planes(W, planes)
    del planes
    _, h, b = solver.state()
    mr = read_mass_restart(directory, W, h, b, identity)
    epoch, ref_h, ref_b, _, _ = read_residual_reference(os.path.join(directory, "residual_reference.in"),
                                                        case.residual_tol, case.moment_tol, origin=mr.step)
    solver.set_mass_target(mr.m0, mr.ledger)
    solver.inherit_references(ref_h, ref_b)
    return mr.step, epoch


# ---------------------------------------------------------------------------
# Monitors
# ---------------------------------------------------------------------------
SPARSE_HEADER = ("local_state,cumulative_state,source,original_gate,moment,M_W,M_h"
                 + "".join(f",R{k}" for k in range(1, 9)) + "".join(f",science_{k}" for k in range(1, 23))
                 + "".join(f",flux_{k}" for k in range(1, 17)) + ",reference_h,reference_b,c_minus_1,mass_rescale_enabled")


def science_pack(norms24, moment_error: float):
    """The 22 diagnostics of a state (SciencePack order = ugks_norms[0:22])
    with the moment error of its mass statistics in place 19."""
    s = np.array(norms24[:22], dtype=np.float64)
    s[18] = moment_error
    return s


def residual8(norms24):
    """PhysicalDiagnostic's residual(8): macro absolute (4), sqrt of the
    summed squared h/b rates (2), raw h/b norms (2)."""
    n = np.asarray(norms24, dtype=np.float64)
    return np.concatenate([n[8:12], n[20:22], n[16:18]])


def state_passes(norms24, moment_error: float, residual_tol: float, moment_tol: float) -> bool:
    """The original gate of an accepted state (src/mod_R5Observe.f90:311-312)."""
    sci = science_pack(norms24, moment_error)
    return bool(np.all(sci[0:4] <= residual_tol) and np.all(sci[12:14] <= residual_tol) and moment_error <= moment_tol)


def sparse_residual_row(local_state: int, cumulative: int, source: str, norms24, walls16, totals5, metrics5,
                        c_minus_1: float, residual_tol: float, moment_tol: float, rescale: bool = True) -> str:
    gate = state_passes(norms24, metrics5[0], residual_tol, moment_tol)
    vals = [1.0 if gate else 0.0, metrics5[0], totals5[0], totals5[4], *residual8(norms24),
            *science_pack(norms24, metrics5[0]), *np.asarray(walls16, dtype=np.float64), norms24[22], norms24[23],
            c_minus_1 if rescale else 0.0, 1.0 if rescale else 0.0]
    return f"{local_state},{cumulative},{source}," + ",".join(fmt_es(float(v)) for v in vals)


def append_sparse_residual(directory: str, row: str):
    path = os.path.join(directory, "sparse_residual.csv")
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write(SPARSE_HEADER + "\n")
        f.write(row + "\n")


def append_consecutive_state(directory: str, cumulative: int, passed: bool, streak_after: int):
    path = os.path.join(directory, "consecutive_states.csv")
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write("cumulative_state,pass,streak_after_evaluation\n")
        f.write(f"{cumulative},{'T' if passed else 'F'},{streak_after}\n")


RESIDUAL_PLT_HEADER = (
    "VARIABLES = iter dtPhysicalMin nCFL alpha rhoResidual rhoUResidual rhoVResidual energyResidual hResidual "
    "bResidual momentError massDefect retries wallTime prSeconds lusgsSeconds correctionExchangeSeconds "
    "evolutionSeconds retryTotal rhoRelative rhoURelative rhoVRelative energyRelative rhoAbsolute rhoUAbsolute "
    "rhoVAbsolute energyAbsolute hRelative bRelative hAbsolute bAbsolute microForwardSeconds "
    "microBackwardSeconds microForwardBackwardExchangeSeconds c_minus_1 mass_rescale_enabled")


def residual_plt_row(it: int, dt: float, ncfl: float, alpha: float, norms24, moment_error: float,
                     timing8=(0.0,) * 8, retries: int = 0, retry_total: int = 0, c_minus_1: float = 0.0,
                     rescale: bool = True) -> str:
    """One row of ResidualImplicit.plt,
    '(I12,1P,11E18.9,I8,5E18.9,I12,15E18.9,2ES25.16E3)'."""
    n = np.asarray(norms24, dtype=np.float64)
    a = [dt, ncfl, alpha, *n[0:4], n[12], n[13], moment_error, n[19]]
    bpart = list(timing8[0:5])
    c = [*n[4:8], *n[8:12], n[14], n[15], n[16], n[17], *timing8[5:8]]
    return (fmt_i(it, 12) + "".join(fmt_e1p(float(v), 18, 9) for v in a) + fmt_i(retries, 8)
            + "".join(fmt_e1p(float(v), 18, 9) for v in bpart) + fmt_i(retry_total, 12)
            + "".join(fmt_e1p(float(v), 18, 9) for v in c) + fmt_es(c_minus_1) + fmt_es(1.0 if rescale else 0.0))


def append_residual_plt(directory: str, row: str, first: bool):
    """Append a row; `first` starts a new file (the reference replaces it
    on the first row of a run), a missing file also gets the header."""
    path = os.path.join(directory, "ResidualImplicit.plt")
    first = first or not os.path.exists(path)
    with open(path, "w" if first else "a") as f:
        if first:
            f.write(RESIDUAL_PLT_HEADER + "\n")
        f.write(row + "\n")


def implicit_iter_line(it: int, ncfl: float, alpha: float, norms24, moment_error: float, wall: float,
                       retries: int = 0, retry_total: int = 0) -> str:
    """The per-iteration stdout line, '(A,I0,1P,11(A,E12.4),2(A,I0))'."""
    n = np.asarray(norms24, dtype=np.float64)
    names = [" nCFL=", " alpha=", " Rrho=", " RrhoU=", " RrhoV=", " RE=", " Rh=", " Rb=", " moment=", " mass=", " wall="]
    vals = [ncfl, alpha, n[0], n[1], n[2], n[3], n[12], n[13], moment_error, n[19], wall]
    return (f"Implicit iter={int(it)}" + "".join(k + fmt_e1p(float(v), 12, 4) for k, v in zip(names, vals))
            + f" retries={int(retries)} retryTotal={int(retry_total)}")


def write_r5_status(directory: str, local_steps: int, cumulative: int, streak: int, gate: bool, reason: str):
    with open(os.path.join(directory, "R5_status.txt"), "w") as f:
        f.write("local_steps cumulative_state verified_streak final_original_gate stop_reason\n")
        f.write(f"{local_steps} {cumulative} {streak} {'T' if gate else 'F'} {reason}\n")


# ---------------------------------------------------------------------------
# Tecplot fields (write_flow)
# ---------------------------------------------------------------------------
def _fortran_a(text: str, w: int = 20) -> str:
    """Aw output of a shorter string: right-justified in w columns."""
    return text.rjust(w) if len(text) <= w else text[:w]


def _zone_line(nx: int, ny: int) -> str:
    """'(a20,I8,a20,I8,a20)' of the ZONE record."""
    return _fortran_a("ZONE  I = ") + fmt_i(nx, 8) + _fortran_a(", J = ") + fmt_i(ny, 8) + _fortran_a(", DATAPACKING=BLOCK")


def _block_values(f, arrays):
    vals = np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in arrays])
    for k in range(0, vals.size, 4):
        f.write("".join(fmt_es(float(v)) for v in vals[k:k + 4]) + "\n")


def write_flow_plt(path: str, nodes, fields10, nx: int, ny: int):
    """Flow2d_<iter>.plt: cell centres and the 10 fields r, u, v, T, p, sxx,
    sxy, syy, qx, qy per cell (ugks_observables order), i fastest."""
    xc, yc = cell_centres(*[np.asarray(a, dtype=np.float64) for a in nodes])
    F = np.asarray(fields10, dtype=np.float64).reshape(ny, nx, 10)
    with open(path, "w") as f:
        f.write(" VARIABLES = x, y, r, u, v, T, p, sxx, sxy, syy, qx, qy\n")
        f.write(_zone_line(nx, ny) + "\n")
        _block_values(f, [xc, yc] + [F[..., k] for k in range(10)])


def write_residual2d_plt(path: str, nodes, res_cell4, nx: int, ny: int):
    """Residual2d_<iter>.plt: the per-cell residuals |dW| A/dt (4) and their
    log10 (floored at the single-precision tiny, as the reference)."""
    xc, yc = cell_centres(*[np.asarray(a, dtype=np.float64) for a in nodes])
    Rc = np.asarray(res_cell4, dtype=np.float64).reshape(ny, nx, 4)
    smv = 1.17549435082228750797e-38
    with open(path, "w") as f:
        f.write(" VARIABLES = x, y, resRho, resRhoU, resRhoV, resE, logResRho, logResRhoU, logResRhoV, logResE\n")
        f.write(_zone_line(nx, ny) + "\n")
        _block_values(f, [xc, yc] + [Rc[..., k] for k in range(4)] + [np.log10(np.maximum(Rc[..., k], smv)) for k in range(4)])


def read_tecplot_block(path: str):
    """(variable names, nx, ny, {name: array[ny, nx]}) of a BLOCK file
    written by write_flow_plt / write_residual2d_plt (or by the reference)."""
    with open(path) as f:
        head = f.readline()
        zone = f.readline()
        rest = f.read().split()
    names = [t.strip() for t in head.split("=", 1)[1].split(",")]
    m = re.search(r"I\s*=\s*(\d+)\s*,\s*J\s*=\s*(\d+)", zone)
    nx, ny = int(m.group(1)), int(m.group(2))
    vals = np.array([float(t.replace("D", "E")) for t in rest])
    if vals.size != len(names) * nx * ny:
        raise RestartError(f"{path}: {vals.size} values for {len(names)} variables on {nx}x{ny}")
    return names, nx, ny, {n: vals[k * nx * ny:(k + 1) * nx * ny].reshape(ny, nx) for k, n in enumerate(names)}
