"""Writes tests/golden/sedov_analytic_2d.json: the analytic Sedov blast solution
(F.X. Timmes' sedov3_qp, quad precision, as the reference's own comparison in
src/sedov_blast/sedov_plot.py:27-39 calls it) for the 2-D case of
src/sedov_blast/test_sedov_blast_2d.ini: cylindrical geometry, uniform
density 1, gamma 1.4, blast energy 0.311357, at t = 0.1 (the shock then lies
well inside the unit box).

The solver is the reference's Fortran source compiled in place by
oracle/Makefile into oracle/_ref/sedov3_qp; run this script in the container
that holds the reference:

    make -C oracle && python tests/golden/make_sedov_analytic.py
"""
from __future__ import annotations

import json
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOLVER = HERE.parent.parent / "oracle" / "_ref" / "sedov3_qp"

NPTS = 1000          # the solver's maximum (nmax), points on [0, 1.2]
EBLAST = 0.311357
GEOM = 2             # cylindrical
OMEGA = 0.0          # uniform ambient density
GAMMA = 1.4
TIME = 0.1
RMAX = 0.75          # points kept: the profile's half-diagonal is 0.7071


def main() -> None:
    if not SOLVER.exists():
        raise SystemExit(f"{SOLVER} missing: run `make -C oracle` where the reference is present")
    with tempfile.TemporaryDirectory() as d:
        out = Path(d) / "sedov.dat"
        res = subprocess.run([str(SOLVER), str(NPTS), str(EBLAST), str(GEOM), str(OMEGA),
                              str(GAMMA), str(TIME), str(out)],
                             capture_output=True, text=True, check=True)
        header = res.stdout
        rows = [line.split() for line in out.read_text().splitlines()[2:] if line.strip()]
    r = [float(x[1]) for x in rows]
    den = [float(x[2]) for x in rows]
    pres = [float(x[4]) for x in rows]
    keep = [i for i, x in enumerate(r) if x <= RMAX]
    # shock radius and post-shock state printed by the solver ("r2 = ...")
    toks = header.replace("=", " = ").split()
    def after(name: str) -> float:
        return float(toks[toks.index(name) + 2])
    data = {
        "source": "sedov3_qp (F.X. Timmes), reference src/sedov_blast/sedov_fortran",
        "params": {"nstep": NPTS, "eblast": EBLAST, "xgeom": GEOM, "omega": OMEGA,
                   "gamma": GAMMA, "time": TIME, "rho0": 1.0},
        "alpha": after("alpha"), "r2": after("r2"), "rho2": after("rho2"), "p2": after("p2"),
        "r": [r[i] for i in keep],
        "density": [den[i] for i in keep],
        "pressure": [pres[i] for i in keep],
    }
    dst = HERE / "sedov_analytic_2d.json"
    dst.write_text(json.dumps(data, indent=0) + "\n")
    print(f"wrote {dst}: {len(keep)} points, r2={data['r2']}, rho2={data['rho2']}")


if __name__ == "__main__":
    main()
