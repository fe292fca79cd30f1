"""Golden-vector cases shared by make_golden.py and the tests.

Inputs are built only from SplitMix64 integers and IEEE +,-,* so that they are
bit-identical on every CPU (no libm calls whose last ulp may differ between
platforms).  Outputs are compared through SHA-256 digests of their raw
little-endian float64 bytes, plus a few exact samples (float.hex) and norms
for diagnosis.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "scale-dg-kernel-extraction_amd"))
from dgk.ic import SplitMix64, cfl_dt  # noqa: E402

GOLDEN_JSON = os.path.join(HERE, "golden.json")


def digest(a) -> str:
    a = np.asarray(a)
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def uniform(seed, shape, lo=-1.0, hi=1.0):
    n = int(np.prod(shape))
    U = SplitMix64(seed).uniform(n).reshape(shape, order="F")
    return np.asfortranarray(lo + (hi - lo) * U)


def bump_field(pos, seed, noise=1e-3):
    """Smooth compact bump prod_d (1 - (2x_d - 1)^2)^2 plus seeded noise."""
    q = np.ones(pos.shape[:2], order="F")
    for d in range(3):
        s = 2.0 * pos[:, :, d] - 1.0
        t = 1.0 - s * s
        q = q * (t * t)
    return np.asfortranarray(q + noise * uniform(seed, q.shape))


MESH_CASES = [(3, 4, 5, 3), (2, 3, 2, 8), (4, 3, 2, 7), (2, 2, 2, 15)]

OPTR_NELEM = 3


def optr_inputs(p):
    Nq = p + 1
    Np, Nf = Nq ** 3, 6 * Nq * Nq
    s = 7000 + 10 * p
    return (uniform(s + 1, (Np, OPTR_NELEM)), uniform(s + 2, (Np, OPTR_NELEM)),
            uniform(s + 3, (Np, OPTR_NELEM)), uniform(s + 4, (Nf, OPTR_NELEM)))


# Tendency cases: rough random state and per-node random velocities (every
# face has VelM != VelP), one advect3d_eq_cal_tend.
TEND_CASES = [((8, 8, 8), 3), ((4, 3, 2), 7), ((2, 3, 2), 8), ((2, 2, 2), 11),
              ((2, 1, 2), 15),
              # every other order once (extents of 1: an element is its own
              # neighbour across that axis)
              ((3, 2, 2), 1), ((2, 3, 3), 2), ((3, 3, 2), 4), ((2, 2, 3), 5), ((2, 3, 2), 6),
              ((2, 2, 2), 9), ((2, 2, 1), 10), ((1, 2, 2), 12), ((2, 1, 1), 13), ((1, 1, 2), 14)]


def tend_inputs(ne, p):
    Nq = p + 1
    Np, Ne = Nq ** 3, ne[0] * ne[1] * ne[2]
    s = 9000 + 100 * p + ne[0]
    q = uniform(s + 1, (Np, Ne))
    u = uniform(s + 2, (Np, Ne))
    v = uniform(s + 3, (Np, Ne), -0.5, 0.75)
    w = uniform(s + 4, (Np, Ne), 0.1, 1.0)
    return q, u, v, w


# Trajectory cases: (mesh, p, nstep, velocity, kind).  "bump" is the smooth
# workload-like field (the 8^3 p3 case of the CPU configuration), "rand" a
# rough random field.
RUN_CASES = [
    ((8, 8, 8), 3, 100, (1.0, 0.5, 0.25), "bump"),
    ((4, 4, 4), 7, 20, (1.0, -0.7, 0.3), "bump"),
    ((3, 2, 2), 8, 10, (0.5, 1.0, -0.25), "rand"),
    ((2, 2, 3), 11, 5, (-0.3, 0.6, 0.9), "rand"),
    ((2, 2, 2), 15, 3, (0.8, 0.6, -0.4), "rand"),
    ((3, 3, 3), 1, 30, (0.7, -0.4, 0.2), "bump"),
    ((3, 2, 3), 5, 10, (-0.6, 0.3, 0.8), "rand"),
    ((2, 1, 2), 13, 3, (0.4, -0.9, 0.5), "rand"),
]


def run_inputs(ne, p, vel, kind, pos):
    Np, Ne = pos.shape[0], pos.shape[1]
    s = 5000 + 100 * p + ne[0]
    if kind == "bump":
        q = bump_field(pos, s)
    else:
        q = uniform(s, (Np, Ne))
    u = np.full((Np, Ne), vel[0], order="F")
    v = np.full((Np, Ne), vel[1], order="F")
    w = np.full((Np, Ne), vel[2], order="F")
    dt = cfl_dt(p, vel, ne)
    return q, u, v, w, dt


# The reference program's own configuration file knobs for its stock run
# (NeX=NeY=NeZ=4, PolyOrder=7, OPT1, dt=1e-3, nstep=1000, output every 100)
# and a second, GENERAL-order case with a non-unit velocity.
STOCK_RUNS = {
    "stock_4cube_p7_opt1": dict(NeX=4, NeY=4, NeZ=4, PolyOrder=7, opt="OPT1", dt="1.0D-3",
                                nstep=1000, output_interval=100, vel=None),
    "3x5x2_p4_general": dict(NeX=3, NeY=5, NeZ=2, PolyOrder=4, opt="GENERAL", dt="2.0D-3",
                             nstep=60, output_interval=20, vel=("0.6D0", "-1.1D0", "0.35D0")),
}


def namelist_text(cfg, opt=None):
    """&PARAM_ADVECT3D text for a STOCK_RUNS entry (main.f90:103-107)."""
    lines = ["&PARAM_ADVECT3D",
             f"  NeX = {cfg['NeX']}, NeY = {cfg['NeY']}, NeZ = {cfg['NeZ']},",
             f"  PolyOrder = {cfg['PolyOrder']},",
             f'  DGOptrKernel_OptType = "{opt or cfg["opt"]}"',
             f"  dt = {cfg['dt']},",
             f"  nstep = {cfg['nstep']}, output_interval = {cfg['output_interval']},"]
    if cfg["vel"]:
        lines.append(f"  vel_x = {cfg['vel'][0]}, vel_y = {cfg['vel'][1]}, vel_z = {cfg['vel'][2]},")
    lines.append("/")
    return "\n".join(lines) + "\n"


def case_key(*parts):
    return "_".join(str(x).replace(" ", "") for x in parts)
