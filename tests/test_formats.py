"""On-disk restart and monitor formats, pinned without a GPU.

What pins them:
  * files of the reference's own runs (tests/golden/formats, written by
    tests/golden/make_golden.py): the shipped Mesh2d.dat (a Fortran
    sequential file) must read as the 8x64 uniform mesh and be rewritten
    byte for byte; every number of the release runs' sparse_residual.csv
    (216 values), the "Implicit iter=" stdout lines and R5_status.txt must be
    reproduced character for character by the formatters;
  * a Fortran runtime (tests/fortran/format_probe.f90 built with amdflang):
    flow2d.dat, vdf2d.dat and a closed-mass companion with its halo must be
    byte-identical to the Python writers, and ES25.16E3 / 1P,E18.9 /
    1P,E12.4 output of edge values (zero, -0, subnormal, huge, 3-digit
    exponents, rounding carries) must agree;
  * round trips and the reference readers' rejections (sizes, identity,
    tampered cells, thresholds, epochs).
"""
import os
import struct
import subprocess
import types

import numpy as np
import pytest

from ugks2d_high_order_amd import build as B, config, restart as R

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "formats")
PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "format_probe.f90")


def test_mesh2d_matches_the_reference_file(tmp_path):
    path = os.path.join(GOLD, "Mesh2d.dat")
    x, y = R.read_mesh2d(path)
    assert R.cartesian_mesh(x, y) == (8, 64, 0.125, 1.0 / 64)
    X, Y = R.couette_nodes(8, 64)
    out = tmp_path / "Mesh2d.dat"
    R.write_mesh2d(str(out), X, Y)
    assert out.read_bytes() == open(path, "rb").read()


@pytest.mark.parametrize("run", ["high_implicit", "low_implicit"])
def test_monitor_text_matches_release_runs(run, tmp_path):
    lines = open(os.path.join(GOLD, f"{run}_sparse_residual.csv")).read().splitlines()
    assert lines[0] == R.SPARSE_HEADER
    n = 0
    for row in lines[1:]:
        fields = row.split(",")
        for t in fields[3:]:
            assert R.fmt_es(float(t)) == t
            n += 1
    assert n == 2 * 54
    # the row writer rebuilds each release row from the state's diagnostics:
    # norms (science pack + references), wall fluxes, totals, metrics, c-1;
    # R1..R8 and the gate are derived, science_19 is the moment error
    for row in lines[1:]:
        f = row.split(",")
        v = np.array([float(t) for t in f[3:]])
        norms = np.zeros(24)
        norms[:22] = v[12:34]
        norms[22:24] = v[50:52]
        totals = np.array([v[2], 0, 0, 0, v[3]])
        metrics = np.array([v[1], 0, 0, 0, 0])
        assert R.sparse_residual_row(int(f[0]), int(f[1]), f[2], norms, v[34:50], totals, metrics, v[52],
                                     residual_tol=1e-8, moment_tol=1e-12) == row
    # the stdout line of the last iteration
    line = open(os.path.join(GOLD, f"{run}_implicit_iter.txt")).read().rstrip("\n")
    import re
    vals = [float(t) for t in re.findall(r"=\s*([-+0-9.]+E[-+]\d+)", line)]
    n24 = np.zeros(24)
    n24[0:4] = vals[2:6]
    n24[12], n24[13], n24[19] = vals[6], vals[7], vals[9]
    assert R.implicit_iter_line(2, vals[0], vals[1], n24, vals[8], vals[10]) == line
    R.write_r5_status(str(tmp_path), 2, 2, 0, False, "iteration_limit")
    assert (tmp_path / "R5_status.txt").read_text() == open(os.path.join(GOLD, f"{run}_R5_status.txt")).read()


def _probe_state(nx=3, ny=4, nv=2):
    W = np.zeros((ny, nx, 4))
    h = np.zeros((ny, nx, nv))
    for j in range(ny):
        for i in range(nx):
            W[j, i] = [(i + 1) + 10 * (j + 1) + 100 * (m + 1) + 0.5 for m in range(4)]
            h[j, i] = [(i + 1) + 10 * (j + 1) + 100 * (k + 1) + 0.25 for k in range(nv)]
    return W.reshape(-1, 4), h.reshape(-1, nv), -h.reshape(-1, nv)


EDGE = [0.0, -0.0, 1.0, -1.0, 0.1, 1 / 3, 2 / 3, 1e-300, 5e-324, 1.7976931348623157e308, 123456789.123,
        9.9999999999999995e-8, 0.99999999995, 9.99995, 1.5e100, 2.5e-100, 3.872983346207417, 1.0032882105403651,
        -7.25e-5, 4.4081e-01]


def test_fortran_runtime_writes_the_same_bytes(tmp_path):
    fc = B.flang()
    if fc is None:
        pytest.skip("no Fortran compiler (amdflang) in this image")
    exe = tmp_path / "probe"
    subprocess.run([fc, "-O1", "-o", str(exe), PROBE], check=True, capture_output=True, timeout=300)
    with open(tmp_path / "values.txt", "w") as f:
        for v in EDGE:
            f.write(struct.pack(">d", v).hex().upper() + "\n")
    subprocess.run([str(exe)], cwd=tmp_path, check=True, capture_output=True, timeout=60)
    W, h, b = _probe_state()
    out = tmp_path / "py"
    out.mkdir()
    R.write_flow2d(str(out / "flow2d.dat"), W, 3, 4)
    R.write_vdf2d(str(out / "vdf2d.dat"), h, b, 3, 4)
    ident = types.SimpleNamespace(nx=3, ny=4, np_x=1, np_y=1, vector=lambda: 0.5 * np.arange(1, 32, dtype=float))
    R.write_mass_restart(str(out), W, h, b, ident, 1.25, [0.5, -0.5, 0.0, 1e-300], 7)
    for mine, ref in (("flow2d.dat", "flow2d.dat"), ("vdf2d.dat", "vdf2d.dat"), ("closed_mass_rank0.bin", "mass.bin")):
        assert (out / mine).read_bytes() == (tmp_path / ref).read_bytes(), mine
    assert (tmp_path / "zone.txt").read_text() == R._zone_line(3, 4) + "\n"
    for v, line in zip(EDGE, (tmp_path / "numbers.txt").read_text().splitlines()):
        assert "|".join((R.fmt_es(v), R.fmt_e1p(v, 18, 9), R.fmt_e1p(v, 12, 4))) == line, v


@pytest.mark.parametrize("total", [0, 5, 16, 17, 40])
def test_subrecords(total):
    import io
    data = bytes(range(total))
    for lim in (8, 16, 1 << 20):
        f = io.BytesIO()
        R.write_record(f, [data[:3], data[3:]], max_subrecord=lim)
        raw = f.getvalue()
        f.seek(0)
        assert R.read_record(f) == data and f.read() == b""
        if total % 8 == 0:
            f.seek(0)
            assert R.read_real_record(f, total // 8).tobytes() == data and f.read() == b""
            f.seek(0)
            with pytest.raises(R.RestartError, match="expected"):
                R.read_real_record(f, total // 8 + 1)
        # markers: leading negative while more follow, trailing negative after the first
        nsub = max(1, -(-total // lim))
        pos, heads, tails = 0, [], []
        for s in range(nsub):
            n = struct.unpack("<i", raw[pos:pos + 4])[0]
            heads.append(n)
            pos += 4 + abs(n)
            tails.append(struct.unpack("<i", raw[pos:pos + 4])[0])
            pos += 4
        assert pos == len(raw)
        assert [x < 0 for x in heads] == [s + 1 < nsub for s in range(nsub)]
        assert [x < 0 for x in tails] == [s > 0 for s in range(nsub)]


def test_state_files_round_trip_and_reject(tmp_path, monkeypatch):
    rng = np.random.default_rng(5)
    nx, ny, nv = 5, 7, 12
    W, h, b = rng.normal(size=(nx * ny, 4)), rng.normal(size=(nx * ny, nv)), rng.normal(size=(nx * ny, nv))
    monkeypatch.setattr(R, "MAX_SUBRECORD", 200)        # force subrecords in the big records
    R.write_flow2d(str(tmp_path / "flow2d.dat"), W, nx, ny)
    R.write_vdf2d(str(tmp_path / "vdf2d.dat"), h, b, nx, ny)
    assert np.array_equal(R.read_flow2d(str(tmp_path / "flow2d.dat"), nx, ny), W)
    h2, b2 = R.read_vdf2d(str(tmp_path / "vdf2d.dat"), nx, ny, nv)
    assert np.array_equal(h2, h) and np.array_equal(b2, b)
    # velocity planes (the device's plane transpose) write the same bytes
    R.write_vdf2d(str(tmp_path / "vdf2d_planes.dat"), None, None, nx, ny, planes=np.stack([h.T, b.T]))
    assert (tmp_path / "vdf2d_planes.dat").read_bytes() == (tmp_path / "vdf2d.dat").read_bytes()
    assert np.array_equal(R.read_vdf2d_planes(str(tmp_path / "vdf2d.dat"), nx, ny, nv), np.stack([h.T, b.T]))
    with pytest.raises(R.RestartError, match="flow2d.dat does not conform"):
        R.read_flow2d(str(tmp_path / "flow2d.dat"), nx, ny + 1)
    with pytest.raises(R.RestartError, match="vdf2d.dat does not conform"):
        R.read_vdf2d(str(tmp_path / "vdf2d.dat"), nx, ny, nv + 1)
    with pytest.raises(R.RestartError, match="does not exist"):
        R.read_flow2d(str(tmp_path / "missing.dat"), nx, ny)
    # grid2d.grd holds the cell centres
    X, Y = R.couette_nodes(nx, ny)
    R.write_grid2d(str(tmp_path / "grid2d.grd"), X, Y)
    recs = R.read_records(str(tmp_path / "grid2d.grd"))
    c = np.frombuffer(recs[2], dtype="<f8").reshape(2, ny, nx)
    assert np.allclose(c[0], (np.arange(nx) + 0.5) / nx, rtol=0, atol=1e-15)
    assert np.allclose(c[1], ((np.arange(ny) + 0.5) / ny)[:, None], rtol=0, atol=1e-15)


def test_blocks_match_the_reference_decomposition():
    # cases/high_implicit on 8 ranks printed origins 1, 9, ..., 57 with 8x8 cells
    assert R.blocks(8, 64, 1, 8) == [(r, 1, 8 * r + 1, 8, 8) for r in range(8)]
    # uneven: the first mod(N, NP) blocks hold one more cell and the blocks tile the grid
    for nx, ny, px, py in ((7, 10, 2, 3), (8, 64, 3, 5), (5, 5, 5, 1)):
        cover = np.zeros((ny, nx), dtype=int)
        for r, oi, oj, ni, nj in R.blocks(nx, ny, px, py):
            cover[oj - 1:oj - 1 + nj, oi - 1:oi - 1 + ni] += 1
            assert ni in (nx // px, nx // px + 1) and nj in (ny // py, ny // py + 1)
        assert np.all(cover == 1)


def _case_and_state(nx=4, ny=12, nu=4):
    case = config.CouetteCase(nx=nx, ny=ny, nu=nu, nv=nu, table=f"half{nu}")
    rng = np.random.default_rng(11)
    nv = nu * nu
    W = 1.0 + 0.1 * rng.random((nx * ny, 4))
    h = rng.random((nx * ny, nv))
    b = rng.random((nx * ny, nv))
    u = rng.normal(size=nv)
    w = rng.random(nv)
    return case, W, h, b, u, w


@pytest.mark.parametrize("np_x,np_y", [(1, 1), (1, 8), (2, 5)])
def test_mass_companions_round_trip_and_checks(tmp_path, np_x, np_y):
    case, W, h, b, u, w = _case_and_state()
    ident = R.run_identity(case, u, w, np_x=np_x, np_y=np_y)
    ledger = np.array([1e-17, -2e-16, 0.0, 3e-15])
    R.write_mass_restart(str(tmp_path), W, h, b, ident, 1.0, ledger, 42)
    assert R.mass_restart_ranks(str(tmp_path)) == list(range(np_x * np_y))
    mr = R.read_mass_restart(str(tmp_path), W, h, b, ident)
    assert mr.step == 42 and mr.m0 == 1.0 and np.array_equal(mr.ledger, ledger)
    assert np.array_equal(mr.identity, ident.vector())
    # halo of the last rank: x wraps, rows beyond the top wall are zero
    r, oi, oj, ni, nj = R.blocks(case.nx, case.ny, np_x, np_y)[-1]
    raw = np.fromfile(tmp_path / f"closed_mass_rank{r}.bin", dtype=np.uint8)[64 + 8 * 31:]
    cells = np.frombuffer(raw.tobytes(), dtype="<f8").reshape(nj + 4, ni + 4, 4 + 2 * h.shape[1])
    Wg = W.reshape(case.ny, case.nx, 4)
    assert np.all(cells[-2:] == 0.0)
    assert np.array_equal(cells[2, 0, :4], Wg[oj - 1, (oi - 3) % case.nx])
    assert np.array_equal(cells[2, -1, :4], Wg[oj - 1, (oi - 1 + ni + 1) % case.nx])
    # the reference's rejections
    with pytest.raises(R.RestartError):          # a different run
        R.read_mass_restart(str(tmp_path), W, h, b, R.run_identity(case, u, w * 1.0000001, np_x=np_x, np_y=np_y))
    W2 = W.copy()
    W2[3, 1] += 1e-15
    with pytest.raises(R.RestartError):          # the companion does not match the flow pair
        R.read_mass_restart(str(tmp_path), W2, h, b, ident)
    if np_x * np_y > 1:
        p = tmp_path / "closed_mass_rank1.bin"
        raw = bytearray(p.read_bytes())
        raw[24:32] = struct.pack("<d", 1.5)
        p.write_bytes(bytes(raw))
        with pytest.raises(R.RestartError, match="M0 differs"):
            R.read_mass_restart(str(tmp_path), W, h, b, ident)


def test_residual_reference_round_trip_and_checks(tmp_path):
    p = str(tmp_path / "residual_reference.in")
    ref_h, ref_b = 0.44080987218372391, 5.3786e-1 / 3
    R.write_residual_reference(p, 0, ref_h, ref_b, 1e-8, 1e-12)
    text = open(p).read()
    assert text == "0   " + R.fmt_es(ref_h).lstrip() + "   " + R.fmt_es(ref_b).lstrip() + \
        "   1.0000000000000000E-008   9.9999999999999998E-013 \n"
    assert R.read_residual_reference(p, 1e-8, 1e-12, origin=3) == (0, ref_h, ref_b, 1e-8, 1e-12)
    with pytest.raises(R.RestartError, match="thresholds mismatch"):
        R.read_residual_reference(p, 1e-9, 1e-12)
    R.write_residual_reference(p, 5, ref_h, ref_b, 1e-8, 1e-12)
    with pytest.raises(R.RestartError, match="epoch after checkpoint"):
        R.read_residual_reference(p, origin=4)


def test_complete_restart_directory_on_the_host(tmp_path):
    """write_restart + the readers load_restart uses, without a device."""
    case, W, h, b, u, w = _case_and_state()
    R.write_restart(str(tmp_path), case, W, h, b, u, w, 1.0, np.zeros(4), 0.5, 0.25, step=3, np_y=4)
    assert sorted(os.listdir(tmp_path)) == sorted(["Mesh2d.dat", "flow2d.dat", "vdf2d.dat", "grid2d.grd",
                                                   "residual_reference.in"]
                                                  + [f"closed_mass_rank{r}.bin" for r in range(4)])
    assert R.cartesian_mesh(*R.read_mesh2d(str(tmp_path / "Mesh2d.dat")))[:2] == (case.nx, case.ny)
    W2 = R.read_flow2d(str(tmp_path / "flow2d.dat"), case.nx, case.ny)
    h2, b2 = R.read_vdf2d(str(tmp_path / "vdf2d.dat"), case.nx, case.ny, h.shape[1])
    mr = R.read_mass_restart(str(tmp_path), W2, h2, b2, R.run_identity(case, u, w, np_y=4))
    assert mr.step == 3
    assert R.read_residual_reference(str(tmp_path / "residual_reference.in"), case.residual_tol,
                                     case.moment_tol, origin=mr.step)[1:3] == (0.5, 0.25)


def test_tecplot_fields_round_trip(tmp_path):
    nx, ny = 5, 3
    nodes = R.couette_nodes(nx, ny, 2.0, 1.0)
    F = np.random.default_rng(2).normal(size=(nx * ny, 10))
    R.write_flow_plt(str(tmp_path / "Flow2d_7.plt"), nodes, F, nx, ny)
    names, mx, my, d = R.read_tecplot_block(str(tmp_path / "Flow2d_7.plt"))
    assert names == ["x", "y", "r", "u", "v", "T", "p", "sxx", "sxy", "syy", "qx", "qy"] and (mx, my) == (nx, ny)
    for k, n in enumerate(names[2:]):
        assert np.array_equal(d[n].ravel(), F[:, k])
    assert np.allclose(d["x"][0], (np.arange(nx) + 0.5) * 2.0 / nx, rtol=0, atol=1e-15)
    res = np.abs(F[:, :4])
    res[0, 0] = 0.0
    R.write_residual2d_plt(str(tmp_path / "Residual2d_7.plt"), nodes, res, nx, ny)
    names, _, _, d = R.read_tecplot_block(str(tmp_path / "Residual2d_7.plt"))
    assert np.array_equal(d["resRhoV"].ravel(), res[:, 2])
    assert d["logResRho"][0, 0] == np.log10(1.17549435082228750797e-38)
