"""The device path on the workload configurations.

8^3 p=3 (the CPU configuration) and the first steps of 32^3 p=7 are checked
bit for bit against the oracle on the same seeded fields.  At 64^3 p=7 (134 M
DOF, too large for the oracle in a test) the checks are size-independent:
discrete mass conservation, agreement with the exact advected solution, the
FP64 variants against one another, and the tiled and linear element orders
against each other bit for bit.
"""
import math

import numpy as np
import pytest

import dgk
import oracle as O
from conftest import say
from dgk import ic

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_run(ne, p, q, vel, dt, nstep, typeid):
    m = O.Mesh(*ne, p)
    F = []
    for a in (q, *(np.full_like(q, c) for c in vel)):
        f = m.field()
        f[:, :m.Ne] = a
        m.update_halo(f)
        F.append(f)
    m.run(*F, dt, nstep, typeid)
    return F[0][:, :m.Ne]


def lgl_weights(x1D):
    p = len(x1D) - 1
    P = np.polynomial.legendre.Legendre.basis(p)(x1D)
    return 2.0 / (p * (p + 1) * P * P)


@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_cpu_configuration_bit_exact(gpu, tname):
    """8^3 elements, p=3, Gaussian + 1e-3 noise, u=(1,0.5,0.25), CFL 0.1,
    100 steps."""
    spec = ic.WORKLOADS["c1_8cube_p3"]
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.initial_field(spec["ic"], m["pos_en"], spec["seed"])
    dt = ic.cfl_dt(p, vel, ne)
    tid = dgk.optr_kernel_setup(tname)
    want = oracle_run(ne, p, q, vel, dt, spec["nstep"], O.TYPEIDS[tname])
    say(f"[c1] {tname}")
    s = dgk.Solver.structured(*ne, p, D, L, typeid=tid)
    s.set_fields(q, *(np.full_like(q, c) for c in vel))
    s.run(dt, spec["nstep"])
    assert np.array_equal(s.get_field(), want)
    s.close()


def test_32cube_p7_first_steps(gpu):
    """32^3 elements, p=7, sine modes: 2 steps bit-exact (exact variant) and
    within 1e-12 (factored lift)."""
    spec = ic.WORKLOADS["c2_32cube_p7"]
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.initial_field(spec["ic"], m["pos_en"], spec["seed"])
    dt = ic.cfl_dt(p, vel, ne)
    want = oracle_run(ne, p, q, vel, dt, 2, O.OPT1)
    for flags, exact in ((0, True), (dgk.FLAG_FACTORED_LIFT, False)):
        say(f"[c2] flags={flags}")
        s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
        s.set_fields(q, *(np.full_like(q, c) for c in vel))
        s.run(dt, 2)
        got = s.get_field()
        if exact:
            assert np.array_equal(got, want)
        else:
            assert rel_l2(got, want) <= 1e-12
        s.close()


@pytest.fixture(scope="module")
def c3_state():
    spec = ic.WORKLOADS["c3_64cube_p7"]
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    pos = m["pos_en"]
    q = ic.initial_field(spec["ic"], pos, spec["seed"])
    return dict(spec=spec, ne=ne, p=p, vel=vel, x1D=x1D, D=D, L=L, q=q, pos=pos,
                dt=ic.cfl_dt(p, vel, ne))


def _run(st, flags, variant=None, nstep=None):
    s = dgk.Solver.structured(*st["ne"], st["p"], st["D"], st["L"], flags=flags)
    if variant is not None:
        s.set_variant(variant)
    q = st["q"]
    s.set_fields(q, *(np.full_like(q, c) for c in st["vel"]))
    s.run(st["dt"], nstep or st["spec"]["nstep"])
    out = s.get_field()
    s.close()
    return out


def test_64cube_p7_properties(gpu, c3_state):
    st = c3_state
    nstep = st["spec"]["nstep"]
    say("[c3] factored, 100 steps")
    qf = _run(st, dgk.FLAG_FACTORED_LIFT)
    assert np.isfinite(qf).all()
    # discrete mass (LGL quadrature) is conserved to round-off
    w = lgl_weights(st["x1D"])
    w3 = np.einsum("i,j,k->ijk", w, w, w).ravel(order="F")
    m0 = float(np.dot(w3, st["q"].sum(axis=1)))
    m1 = float(np.dot(w3, qf.sum(axis=1)))
    scale = float(np.dot(w3, np.abs(st["q"]).sum(axis=1)))
    assert abs(m1 - m0) <= 1e-13 * scale
    # the field has moved by u*T, as the exact solution has
    T = nstep * st["dt"]
    g = ic.SplitMix64(st["spec"]["seed"])
    pos = st["pos"]
    exact = np.zeros_like(qf)
    for _ in range(8):  # the same draws as ic.sine_modes(kmax=4)
        r = g.uniform(5)
        k = [1 + min(int(r[d] * 4), 3) for d in range(3)]
        A, phi = 0.1 + 0.9 * r[3], 2.0 * math.pi * r[4]
        arg = sum(k[d] * (pos[:, :, d] - st["vel"][d] * T) for d in range(3))
        exact += A * np.sin(2.0 * math.pi * arg + phi)
    err = rel_l2(qf, exact)
    moved = rel_l2(st["q"], exact)
    assert err <= 1e-7, err
    assert moved > 1e-2, moved
    # the bit-identical variant differs from the factored one by rounding only
    say("[c3] exact, 100 steps")
    qe = _run(st, 0)
    assert rel_l2(qf, qe) <= 1e-10


def test_64cube_p7_element_orders_agree(gpu, c3_state):
    """Tiled and linear element orders and the store flavour change the
    schedule, never the bits."""
    st = c3_state
    say("[c3] tiled vs linear order, 10 steps")
    a = _run(st, 0, variant=1, nstep=10)   # non-temporal stores, tiled
    b = _run(st, 0, variant=3, nstep=10)   # non-temporal stores, linear
    c = _run(st, 0, variant=0, nstep=10)   # plain stores, tiled
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)
