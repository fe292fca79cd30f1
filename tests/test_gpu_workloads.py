"""The device path on the workload configurations.

8^3 p=3 (the CPU configuration), the 100 steps of 32^3 p=7 and the first
steps of 16^3 p=15 are checked bit for bit against the oracle on the same
seeded fields.  At 64^3 p=7 (134 M DOF, too large for the oracle in a
test) the checks are size-independent: discrete mass conservation,
agreement with the exact advected solution, the FP64 variants against one
another, the tiled and linear element orders against each other bit for
bit, and the mixed-precision form against the bit-identical one per step.
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


def _workload(name):
    spec = ic.WORKLOADS[name]
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.initial_field(spec["ic"], m["pos_en"], spec["seed"])
    return spec, ne, p, vel, x1D, D, L, q, ic.cfl_dt(p, vel, ne)


def test_32cube_p7_full_run(gpu):
    """32^3 elements, p=7, sine modes, u=(1,-0.7,0.3), CFL 0.1: all 100
    steps bit-exact against the oracle (exact variant); the factored lift
    within 1e-12 after one step and 1e-10 after 100."""
    spec, ne, p, vel, x1D, D, L, q, dt = _workload("c2_32cube_p7")
    nstep = spec["nstep"]
    say("[c2] oracle, 100 steps")
    want = oracle_run(ne, p, q, vel, dt, nstep, O.OPT1)
    say("[c2] exact, 100 steps")
    s = dgk.Solver.structured(*ne, p, D, L)
    s.set_fields(q, *(np.full_like(q, c) for c in vel))
    s.run(dt, nstep)
    assert np.array_equal(s.get_field(), want)
    say("[c2] factored, 1 and 100 steps")
    f = dgk.Solver.structured(*ne, p, D, L, flags=dgk.FLAG_FACTORED_LIFT)
    e1 = dgk.Solver.structured(*ne, p, D, L)
    for a in (f, e1):
        a.set_fields(q, *(np.full_like(q, c) for c in vel))
        a.run(dt, 1)
    assert f.diagnostics(x1D, ref=e1)["rel_l2"] <= 1e-12
    f.run(dt, nstep - 1)
    assert rel_l2(f.get_field(), want) <= 1e-10
    for a in (s, f, e1):
        a.close()


def test_16cube_p15(gpu):
    """The high-order configuration: 16^3 elements, p=15 (4096 nodes per
    element), sine modes with wavenumbers 1..8, u=(0.8,0.6,-0.4), 50 steps.
    Exact variant bit-exact against the oracle for 2 steps; the factored
    lift, the FMA form with uniform velocity and the mixed form against the
    exact variant on the device: FP64 forms <= 1e-12 after one step and
    <= 1e-10 after 50, the mixed form <= 1e-9 after 50 (measured 1.7e-10;
    at kh ~ pi many lines' remainders exceed the FP32 threshold and take the
    FP64 contraction)."""
    spec, ne, p, vel, x1D, D, L, q, dt = _workload("c5_16cube_p15")
    nstep = spec["nstep"]
    say("[c5] oracle, 2 steps")
    want = oracle_run(ne, p, q, vel, dt, 2, O.OPT1)
    uvw = [np.full_like(q, c) for c in vel]
    say("[c5] exact, 2 steps")
    e = dgk.Solver.structured(*ne, p, D, L)
    e.set_fields(q, *uvw)
    e.run(dt, 2)
    assert np.array_equal(e.get_field(), want)
    e.set_fields(q, *uvw)
    others = []
    for flags, bound in ((dgk.FLAG_FACTORED_LIFT, 1e-10),
                         (dgk.FLAG_FMA | dgk.FLAG_FACTORED_LIFT | dgk.FLAG_CONST_VELOCITY, 1e-10),
                         (dgk.FLAG_MIXED, 1e-9)):
        say(f"[c5] flags={flags}")
        s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
        s.set_fields(q, *uvw)
        others.append((s, flags, bound))
    e.run(dt, 1)
    for s, flags, _ in others:
        s.run(dt, 1)
        err = s.diagnostics(x1D, ref=e)["rel_l2"]
        assert err <= (1e-12 if not flags & dgk.FLAG_MIXED else 1e-9), (flags, err)
    e.run(dt, nstep - 1)
    for s, flags, bound in others:
        s.run(dt, nstep - 1)
        d = s.diagnostics(x1D, ref=e)
        say(f"[c5] flags={flags} rel-L2 after {nstep} steps {d['rel_l2']:.3e}")
        assert 0.0 < d["rel_l2"] <= bound, (flags, d["rel_l2"])
        s.close()
    e.close()


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


def test_64cube_p7_mixed_precision(gpu, c3_state):
    """The mixed-precision configuration: 64^3 p=7, DGK_FLAG_MIXED against
    the bit-identical FP64 variant on the device, compared after every one of
    the 100 steps: rel-L2 <= 1e-9 at every step (measured 1.3e-14 after one
    step, 9.3e-14 after 100: the lines' remainders after their cubic part are
    ~1e-4 of the field here), mass conserved to round-off, and the uniform-
    velocity mixed form gives the same bits as the per-node one."""
    st = c3_state
    q, vel, dt, x1D = st["q"], st["vel"], st["dt"], st["x1D"]
    nstep = st["spec"]["nstep"]
    say("[c3] mixed vs exact FP64, per step")
    sv = []
    for flags in (dgk.FLAG_MIXED, 0, dgk.FLAG_MIXED | dgk.FLAG_CONST_VELOCITY):
        s = dgk.Solver.structured(*st["ne"], st["p"], st["D"], st["L"], flags=flags)
        s.set_fields(q, *(np.full_like(q, c) for c in vel))
        sv.append(s)
    m0 = sv[0].diagnostics(x1D)
    errs = []
    for _ in range(nstep):
        for s in sv:
            s.step(dt)
        for s in sv:
            s.synchronize()
        errs.append(sv[0].diagnostics(x1D, ref=sv[1])["rel_l2"])
    say("[c3] mixed rel-L2 per step: " + " ".join(f"{e:.2e}" for e in errs[::10]) +
        f" ... {errs[-1]:.3e}")
    assert all(0.0 < e <= 1e-9 for e in errs), max(errs)
    m1 = sv[0].diagnostics(x1D)
    assert abs(m1["mass"] - m0["mass"]) <= 1e-12 * m0["abs_mass"]
    assert sv[2].diagnostics(x1D, ref=sv[0])["rel_l2"] == 0.0
    for s in sv:
        s.close()
