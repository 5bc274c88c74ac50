"""Write the golden fixtures of the UGKS2D HighOrderImplicit path.

Run once in a container that holds the reference checkout (default
/root/reference).  It reads only data files the reference ships and writes:

  couette_release.npz  - the shipped 32x32 velocity table
                         (velocity_half32_scale2.dat: u, v, weight per k)
  couette_explicit_release.json - the residuals res(1:4) and dtPhysical the
                         8-rank release run of cases/low_explicit printed
                         after its second iteration (docs/release_checks/
                         low_explicit/stdout.log, 7 significant digits)
  couette_release.json - reference-held scalars of the 8-rank release run
                         of cases/high_implicit (Config.ini as shipped):
                         initial raw micro-residual norms reference_h/b
                         (sparse_residual.csv columns reference_h/reference_b,
                         identical on both rows), the two accepted-state rows
                         (science_*, flux_*, R*, c_minus_1), and the printed
                         case constants (stdout.log).

  formats/               - data files of the reference's runs that pin the
                         on-disk and monitor formats byte for byte:
                         Mesh2d.dat (a Fortran sequential file of the shipped
                         8x64 mesh), the release runs' sparse_residual.csv and
                         R5_status.txt (cases/high_implicit, low_implicit) and
                         their "Implicit iter=" stdout lines.

Nothing from the reference's source code is copied; the fixtures are data.
"""
from __future__ import annotations

import csv
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    path = os.path.join(REF, "velocity_half32_scale2.dat")
    with open(path) as f:
        nu, nv = (int(t) for t in f.readline().split())
        data = np.loadtxt(f)
    np.savez(os.path.join(HERE, "couette_release.npz"), nu=nu, nv=nv, u=data[:, 0], v=data[:, 1], w=data[:, 2])

    rows = []
    with open(os.path.join(REF, "docs/release_checks/high_implicit/sparse_residual.csv")) as f:
        rd = csv.reader(f)
        header = [h.strip() for h in next(rd)]
        for r in rd:
            rows.append({k: (v.strip() if k == "source" else float(v)) for k, v in zip(header, r)})
    out = {
        "source": {
            "velocity_table": "velocity_half32_scale2.dat",
            "residual_rows": "docs/release_checks/high_implicit/sparse_residual.csv",
            "case": "cases/high_implicit/Config.ini (NP_X=1, NP_Y=8; 8 MPI ranks)",
        },
        "case": {"nx": 8, "ny": 64, "nu": nu, "nv": nv, "omega": 0.5, "mu_ref": 0.05637282460904075,
                 "pr": 2.0 / 3.0, "ck": 1, "cfl": 0.1, "ncfl": 1000.0, "tw": 2.0,
                 "wall_ux": [-3.872983346207417, 3.872983346207417]},
        "reference_h": rows[0]["reference_h"],
        "reference_b": rows[0]["reference_b"],
        "rows": rows,
        "printed_kn": 0.09991823,
        "persistent_M0": 1.0,
    }
    with open(os.path.join(HERE, "couette_release.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(rows), "rows; reference_h/b =", out["reference_h"], out["reference_b"])

    ex = {"source": {"stdout": "docs/release_checks/low_explicit/stdout.log",
                     "case": "cases/low_explicit/Config.ini (TimeIntegration_method Explicit, NP_Y=8; 8 MPI ranks)"}}
    with open(os.path.join(REF, "docs/release_checks/low_explicit/stdout.log")) as f:
        for line in f:
            if "iter,simTime" in line:
                ex["iteration"] = int(line.split(":")[1].split()[0])
            elif "dtPhysical:" in line:
                ex["dt"] = float(line.split(":")[1])
            elif line.strip().startswith("res:"):
                ex["res"] = [float(t) for t in line.split(":")[1].split()]
    with open(os.path.join(HERE, "couette_explicit_release.json"), "w") as f:
        json.dump(ex, f, indent=1)
    print("explicit release:", ex)


def formats():
    out = os.path.join(HERE, "formats")
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(REF, "cases/high_implicit/Mesh2d.dat"), "rb") as f:
        data = f.read()
    with open(os.path.join(out, "Mesh2d.dat"), "wb") as f:
        f.write(data)
    for run in ("high_implicit", "low_implicit"):
        base = os.path.join(REF, "docs/release_checks", run)
        for name in ("sparse_residual.csv", "R5_status.txt"):
            with open(os.path.join(base, name), "rb") as f:
                data = f.read()
            with open(os.path.join(out, f"{run}_{name}"), "wb") as f:
                f.write(data)
        with open(os.path.join(base, "stdout.log")) as f:
            lines = [l for l in f if l.startswith("Implicit iter=")]
        with open(os.path.join(out, f"{run}_implicit_iter.txt"), "w") as f:
            f.writelines(lines)
    print("format fixtures in", out)


if __name__ == "__main__":
    main()
    formats()
