"""Writes tests/golden/golden.json from the reference itself.

Runs the reference's Fortran modules, compiled from source by
``make -C oracle ref`` into oracle/_ref/ and driven by oracle/ref_harness.f90,
on the inputs of tests/golden/cases.py, and records digests, exact samples
and norms of every output.  Usage (in a container holding the reference):

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "oracle"))
import cases as CS  # noqa: E402
import oracle as O  # noqa: E402


def summary(a):
    a = np.asarray(a, dtype=np.float64)
    flat = a.ravel(order="F")
    idx = np.linspace(0, flat.size - 1, num=min(8, flat.size)).astype(int)
    return {
        "sha256": CS.digest(a),
        "shape": list(a.shape),
        "l2": float(np.sqrt(np.sum(flat * flat))),
        "samples": {str(int(i)): float(flat[i]).hex() for i in idx},
    }


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/ref_harness is missing: run `make -C oracle ref` first")
    out = {"generator": "oracle/ref_harness (reference modules compiled with amdflang)",
           "mesh": {}, "optr": {}, "tend": {}, "run": {}}
    for (nx, ny, nz, p) in CS.MESH_CASES:
        r = O.ref_mesh(nx, ny, nz, p)
        out["mesh"][CS.case_key(nx, ny, nz, p)] = {
            k: CS.digest(r[k]) for k in ("x1D", "D1D", "Lift", "VMapM", "VMapP",
                                         "halo_src_map", "normal_fn", "Escale",
                                         "Fscale", "pos_en")}
    for p in range(1, 16):
        ix, iy, iz, fb = CS.optr_inputs(p)
        for t in ("GENERAL", "OPT1"):
            ox, oy, oz, ol = O.ref_optr(p, t, ix, iy, iz, fb)
            out["optr"][CS.case_key(p, t)] = {"x": summary(ox), "y": summary(oy),
                                              "z": summary(oz), "lift": summary(ol)}
    # the OPT1 module's directional kernels and gradlike (the reference's
    # gradlike z output carries its loop slip, mod_dg_optr_kernel_opt1.F90.erb:208)
    out["optr2"] = {}
    for p in range(1, 16):
        a = CS.optr_inputs(p)[0]
        res = O.ref_optr2(p, a)
        out["optr2"][CS.case_key(p)] = {
            k: summary(v) for k, v in zip(("dirX", "dirY", "dirZ", "gradX", "gradY", "gradZ"),
                                          res)}
    for (ne, p) in CS.TEND_CASES:
        q, u, v, w = CS.tend_inputs(ne, p)
        for t in ("GENERAL", "OPT1"):
            d = O.ref_run(*ne, p, q, u, v, w, 1.0, 1, t, mode="tend")
            out["tend"][CS.case_key(*ne, p, t)] = summary(d)
    for (ne, p, nstep, vel, kind) in CS.RUN_CASES:
        m = O.Mesh(*ne, p)
        q, u, v, w, dt = CS.run_inputs(ne, p, vel, kind, m.pos_en)
        for t in ("GENERAL", "OPT1"):
            qf, mm, _, _ = O.ref_run(*ne, p, q, u, v, w, dt, nstep, t)
            out["run"][CS.case_key(*ne, p, nstep, t)] = {
                "q": summary(qf), "dt": float(dt).hex(),
                "minmax": [[float(a).hex() for a in col] for col in mm.T]}
    # the reference program itself (main.f90 unmodified): its istep/min/max
    # lines (main.f90:76-79) for the configurations of CS.STOCK_RUNS
    out["stock"] = {}
    prog = os.path.join(REPO, "oracle", "_ref", "scale-dg_extraction")
    for name, cfg in CS.STOCK_RUNS.items():
        with O._workdir() as d:
            with open(os.path.join(d, "input.conf"), "w") as f:
                f.write(CS.namelist_text(cfg))
            r = subprocess.run([prog, "input.conf"], cwd=d, capture_output=True, text=True,
                               check=True, env=dict(os.environ, OMP_NUM_THREADS="4"))
        lines = [l for l in r.stdout.splitlines() if l[:8].strip().isdigit()]
        out["stock"][name] = {"lines": lines}
    with open(CS.GOLDEN_JSON, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", CS.GOLDEN_JSON)


if __name__ == "__main__":
    main()
