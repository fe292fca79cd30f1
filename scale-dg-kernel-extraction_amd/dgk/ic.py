"""Seeded synthetic inputs for the advection workloads.

The reference hard-codes q = sin(2 pi x) sin(2 pi y) sin(2 pi z) and constant
namelist velocities (main.f90:146-157).  The workloads below replace that
initial field; every field is generated here once, on the host, from the node
coordinates pos_en(Np,Ne,3) (mod_mesh.f90:332-353), and the same arrays are fed
to every path that is compared.

Random numbers come from SplitMix64 (Steele, Lea and Flood, 2014) with the top
53 bits mapped to [0,1), so streams are identical on every platform.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream; ``uniform(n)`` returns n doubles in [0,1)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & _M64)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            inc = np.uint64(0x9E3779B97F4A7C15)
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * inc
            self.state = self.state + np.uint64(n) * inc
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def cfl_dt(p: int, vel, ne_axes, dom=(1.0, 1.0, 1.0), cfl: float = 0.1) -> float:
    """dt = CFL / ((2p+1) * sum_d |u_d| / dx_d)."""
    s = sum(abs(float(u)) / (float(L) / n) for u, n, L in zip(vel, ne_axes, dom))
    return cfl / ((2 * p + 1) * s)


def gaussian_noise(pos, seed: int, width2: float = 0.01, noise: float = 1e-3,
                   centre=(0.5, 0.5, 0.5)):
    """q = exp(-|x-c|^2 / width2) + noise * U(-1,1) per DOF (pos: (Np,Ne,3))."""
    r2 = ((pos[:, :, 0] - centre[0]) ** 2 + (pos[:, :, 1] - centre[1]) ** 2
          + (pos[:, :, 2] - centre[2]) ** 2)
    q = np.exp(-r2 / width2)
    if noise:
        U = SplitMix64(seed).uniform(q.size).reshape(q.shape, order="F")
        q = q + noise * (2.0 * U - 1.0)
    return np.asfortranarray(q)


def sine_modes(pos, seed: int, nmodes: int = 8, kmax: int = 4):
    """q = sum_m A_m sin(2 pi k_m.x + phi_m), k_m in {1..kmax}^3,
    A_m ~ U[0.1,1], phi_m ~ U[0,2 pi)."""
    g = SplitMix64(seed)
    q = np.zeros(pos.shape[:2], order="F")
    for _ in range(nmodes):
        r = g.uniform(5)
        k = [1 + min(int(r[d] * kmax), kmax - 1) for d in range(3)]
        A = 0.1 + 0.9 * r[3]
        phi = 2.0 * math.pi * r[4]
        q += A * np.sin(2.0 * math.pi * (k[0] * pos[:, :, 0] + k[1] * pos[:, :, 1]
                                         + k[2] * pos[:, :, 2]) + phi)
    return np.asfortranarray(q)


def blob(pos, centre, width):
    """Gaussian blob exp(-|x-c|^2 / w^2) (periodic images ignored)."""
    r2 = ((pos[:, :, 0] - centre[0]) ** 2 + (pos[:, :, 1] - centre[1]) ** 2
          + (pos[:, :, 2] - centre[2]) ** 2)
    return np.asfortranarray(np.exp(-r2 / (width * width)))


def constant(shape, value: float):
    return np.full(shape, float(value), order="F")


# Workload table: (NeX,NeY,NeZ), p, velocity, initial field, seed, nstep.
WORKLOADS = {
    "c1_8cube_p3": dict(ne=(8, 8, 8), p=3, vel=(1.0, 0.5, 0.25), ic="gauss", seed=1001, nstep=100),
    "c2_32cube_p7": dict(ne=(32, 32, 32), p=7, vel=(1.0, -0.7, 0.3), ic="sine4", seed=1002, nstep=100),
    "c3_64cube_p7": dict(ne=(64, 64, 64), p=7, vel=(1.0, -0.7, 0.3), ic="sine4", seed=1002, nstep=100),
    "c5_16cube_p15": dict(ne=(16, 16, 16), p=15, vel=(0.8, 0.6, -0.4), ic="sine8", seed=1005, nstep=50),
    # single-instance timing points for the orders of the ensemble workload
    "p11_24cube": dict(ne=(24, 24, 24), p=11, vel=(0.6, -0.5, 0.7), ic="sine4", seed=1011, nstep=20),
    "p5_32cube": dict(ne=(32, 32, 32), p=5, vel=(0.6, -0.5, 0.7), ic="sine4", seed=1015, nstep=20),
    "p3_64cube": dict(ne=(64, 64, 64), p=3, vel=(0.6, -0.5, 0.7), ic="sine4", seed=1013, nstep=20),
}

# The ensemble workload: 64 instances drawn by ensemble_workload(64, 1004).
ENSEMBLE = dict(n=64, seed=1004, nstep=50)


def ensemble_workload(n: int = 64, seed: int = 1004, orders=(3, 5, 7, 11),
                      ne_choices=(8, 12, 16, 24), cfl: float = 0.1):
    """Seeded ensemble of periodic unit-cube instances: per instance the order
    from ``orders``, each axis's element count from ``ne_choices``, a constant
    velocity with components U[-1,1], a Gaussian blob with centre U[0,1]^3 and
    width U[0.05,0.2], and dt at the given CFL.  Draws per instance, in this
    order: order, NeX, NeY, NeZ, u, v, w, centre x, y, z, width."""
    g = SplitMix64(seed)
    out = []
    for _ in range(n):
        r = g.uniform(11)
        p = orders[min(int(r[0] * len(orders)), len(orders) - 1)]
        ne = tuple(ne_choices[min(int(r[1 + d] * len(ne_choices)), len(ne_choices) - 1)]
                   for d in range(3))
        vel = tuple(float(2.0 * r[4 + d] - 1.0) for d in range(3))
        centre = tuple(float(r[7 + d]) for d in range(3))
        width = float(0.05 + 0.15 * r[10])
        out.append(dict(p=p, ne=ne, vel=vel, centre=centre, width=width,
                        dt=cfl_dt(p, vel, ne, cfl=cfl)))
    return out


def initial_field(kind: str, pos, seed: int):
    if kind == "gauss":
        return gaussian_noise(pos, seed)
    if kind == "sine4":
        return sine_modes(pos, seed, kmax=4)
    if kind == "sine8":
        return sine_modes(pos, seed, kmax=8)
    raise ValueError(f"unknown initial field {kind!r}")
