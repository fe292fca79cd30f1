"""The batched ensemble (dgk_ensemble_*): many periodic instances of uneven
order, mesh and velocity advanced by one persistent launch per RK stage.

Every instance must give the bits of the oracle's run of that instance alone
(OPT1 summation order) and of a single-instance device solver; the factored
lift, the constant-velocity form, the tuned grid split and the graph replay
change the schedule or the rounding only where stated.
"""
import numpy as np
import pytest

import dgk
import oracle as O
from conftest import say
from dgk import ic

pytestmark = pytest.mark.gpu

# Uneven small instances: every supported order, ragged work items (element
# counts that are not multiples of the elements per item), mixed velocities.
SMALL = [
    dict(p=1, ne=(3, 4, 5)),
    dict(p=3, ne=(4, 3, 5)),
    dict(p=5, ne=(3, 3, 4)),
    dict(p=7, ne=(3, 5, 3)),
    dict(p=11, ne=(2, 3, 2)),
    dict(p=3, ne=(2, 2, 3)),
    dict(p=7, ne=(2, 2, 2)),
]


def _setup(specs, seed):
    g = ic.SplitMix64(seed)
    out = []
    for s in specs:
        r = g.uniform(7)
        vel = tuple(float(2.0 * r[d] - 1.0) for d in range(3))
        centre = tuple(float(r[3 + d]) for d in range(3))
        width = float(0.05 + 0.15 * r[6])
        x1D, D, L = dgk.read_operator(s["p"])
        m = dgk.mesh_setup(*s["ne"], s["p"], x1D=x1D, tables=False)
        q = ic.blob(m["pos_en"], centre, width) + 0.01 * np.sin(6.0 * m["pos_en"][:, :, 0])
        q = np.asfortranarray(q)
        out.append(dict(s, vel=vel, q=q, D=D, L=L,
                        dt=ic.cfl_dt(s["p"], vel, s["ne"])))
    return out


def _ensemble(inst, flags=0):
    ens = dgk.Ensemble(inst, flags=flags)
    for i, s in enumerate(inst):
        ens.set_fields(i, s["q"], *(np.full_like(s["q"], c) for c in s["vel"]))
    ens.set_dt([s["dt"] for s in inst])
    return ens


def _oracle(s, nstep):
    m = O.Mesh(*s["ne"], s["p"])
    F = []
    for a in (s["q"], *(np.full_like(s["q"], c) for c in s["vel"])):
        f = m.field()
        f[:, :m.Ne] = a
        m.update_halo(f)
        F.append(f)
    m.run(*F, s["dt"], nstep, O.OPT1)
    return F[0][:, :m.Ne]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def small():
    return _setup(SMALL, 4242)


def test_ensemble_matches_oracle_bitwise(gpu, small):
    nstep = 3
    say("[ens] exact, 3 steps")
    ens = _ensemble(small)
    assert ens.query(4) == 5  # Nq 2, 4, 6, 8, 12
    ens.run(nstep)
    for i, s in enumerate(small):
        want = _oracle(s, nstep)
        got = ens.get_field(i)
        assert np.array_equal(got, want), (i, s["p"], s["ne"], rel_l2(got, want))
    ens.close()


def test_ensemble_matches_single_solver(gpu, small):
    """Each instance equals a single-instance device run with the same flags
    (FMA contractions included)."""
    nstep = 2
    for flags in (0, dgk.FLAG_FMA):
        say(f"[ens] vs Solver flags={flags}")
        ens = _ensemble(small, flags)
        ens.run(nstep)
        for i, s in enumerate(small):
            sv = dgk.Solver.structured(*s["ne"], s["p"], s["D"], s["L"], flags=flags)
            sv.set_fields(s["q"], *(np.full_like(s["q"], c) for c in s["vel"]))
            sv.run(s["dt"], nstep)
            assert np.array_equal(ens.get_field(i), sv.get_field()), (i, flags)
            sv.close()
        ens.close()


def test_ensemble_const_velocity_same_bits(gpu, small):
    say("[ens] constant-velocity form")
    a = _ensemble(small)
    b = _ensemble(small, dgk.FLAG_CONST_VELOCITY)
    a.run(2)
    b.run(2)
    for i in range(len(small)):
        assert np.array_equal(a.get_field(i), b.get_field(i)), i
    a.close()
    b.close()


def test_ensemble_factored_lift(gpu, small):
    """The 1D-factored lift is exact to rounding: within 1e-12 of the oracle
    per instance after 3 steps."""
    say("[ens] factored lift")
    ens = _ensemble(small, dgk.FLAG_FACTORED_LIFT | dgk.FLAG_CONST_VELOCITY)
    ens.run(3)
    for i, s in enumerate(small):
        assert rel_l2(ens.get_field(i), _oracle(s, 3)) <= 1e-12, i
    ens.close()


def test_ensemble_tune_keeps_state_and_bits(gpu, small):
    say("[ens] tune")
    a = _ensemble(small)
    b = _ensemble(small)
    ms = b.tune(2)
    assert ms.shape == (5,) and (ms > 0).all()
    assert b.query(5) == 1
    cls = b.classes
    assert [c[0] for c in cls] == [2, 4, 6, 8, 12]
    assert all(wg > 0 and wg % 8 == 0 for _, wg in cls)
    nt, epb = b.shape  # the tile shape the tuning chose
    most = [nt // (nq * nq) for nq, _ in cls]
    assert nt in (192, 256) and all(1 <= e <= m for e, m in zip(epb, most)), (nt, epb)
    for i in range(len(small)):  # tuning left the state alone
        assert np.array_equal(b.get_field(i), small[i]["q"])
    a.run(2)
    b.run(2)
    for i in range(len(small)):
        assert np.array_equal(a.get_field(i), b.get_field(i)), i
    a.close()
    b.close()


# most elements per item of Nq 2, 4, 6, 8, 12 at 192 and 256 threads
MOST = {192: [48, 12, 5, 3, 1], 256: [64, 16, 7, 4, 1]}


@pytest.mark.parametrize("flags", [0, dgk.FLAG_FMA | dgk.FLAG_FACTORED_LIFT |
                                   dgk.FLAG_CONST_VELOCITY, dgk.FLAG_MIXED])
def test_ensemble_tile_shapes_keep_bits(gpu, small, flags):
    """Every tile shape gives the bits of the default one: both workgroup
    widths, each class at the most, about half and one element per item,
    and uneven choices (ragged items at every class)."""
    ref = _ensemble(small, flags)
    assert ref.shape == (192, MOST[192])
    ref.run(2)
    want = [ref.get_field(i) for i in range(len(small))]
    ref.close()
    for nt in (192, 256):
        most = MOST[nt]
        for name, epb in (("most", most), ("half", [(m + 1) // 2 for m in most]),
                          ("one", [1] * 5), ("uneven", [max(1, m - 1) for m in most])):
            say(f"[ens] shape flags={flags} threads={nt} {name} {epb}")
            ens = _ensemble(small, flags)
            ens.set_shape(nt, epb)
            assert ens.shape == (nt, epb)
            assert ens.query(1) == sum(-(-int(np.prod(s["ne"])) // epb[[1, 3, 5, 7, 11].index(
                s["p"])]) for s in small)
            assert all(wg > 0 and wg % 8 == 0 for _, wg in ens.classes)
            ens.run(2)
            for i in range(len(small)):
                assert np.array_equal(ens.get_field(i), want[i]), (nt, name, i)
            ens.close()


def test_ensemble_tile_shape_errors(gpu, small):
    """Shapes outside what the workgroup holds are refused and leave the
    shape as it was; an instance of order 13 rules out 192 threads."""
    ens = _ensemble(small)
    ens.set_shape(256, [0, 8, 0, 2, 0])
    assert ens.shape == (256, [64, 8, 7, 2, 1])
    for nt, epb in ((128, None), (512, None), (192, [49, 1, 1, 1, 1]), (256, [1, 1, 1, 1, 2]),
                    (192, [1, 1, 1, 1, -1])):
        with pytest.raises(dgk.DGKError):
            ens.set_shape(nt, epb)
        assert ens.shape == (256, [64, 8, 7, 2, 1])
    with pytest.raises(ValueError):
        ens.set_shape(192, [1, 1])
    ens.set_shape()
    assert ens.shape == (192, MOST[192])
    ens.close()
    inst = _setup([dict(p=13, ne=(2, 2, 2)), dict(p=3, ne=(2, 2, 2))], 7)
    ens = _ensemble(inst)
    assert ens.shape == (256, [16, 1])
    with pytest.raises(dgk.DGKError):
        ens.set_shape(192)
    ens.close()


def test_ensemble_step_graph_vs_run(gpu, small):
    """Graph replays (step) and the stage-timing launches advance the same
    state: 3 stages of time_stages equal one step."""
    say("[ens] step vs time_stages")
    a = _ensemble(small)
    b = _ensemble(small)
    for _ in range(2):
        a.step()
    b.time_stages(6)
    for i in range(len(small)):
        assert np.array_equal(a.get_field(i), b.get_field(i)), i
    a.close()
    b.close()


@pytest.mark.parametrize("pmax,threads", [(12, 192), (15, 256)])
def test_ensemble_every_order(gpu, pmax, threads):
    """One instance of every order up to pmax (ragged meshes) in one
    ensemble: each gives the oracle's bits (OPT1 order) after 2 steps, exact
    and with the uniform-velocity form.  Orders up to 12 run 192-thread
    workgroups; an instance of order 13..15 switches them to 256."""
    specs = [dict(p=p, ne=(2 + p % 3, 2 + (p + 1) % 2, 2 + p % 2)) for p in range(1, pmax + 1)]
    inst = _setup(specs, 99)
    want = [_oracle(s, 2) for s in inst]
    for flags in (0, dgk.FLAG_CONST_VELOCITY):
        say(f"[ens] every order to p={pmax}, flags={flags}")
        ens = _ensemble(inst, flags)
        assert ens.query(4) == pmax and ens.query(6) == threads
        ens.run(2)
        for i, s in enumerate(inst):
            assert np.array_equal(ens.get_field(i), want[i]), (s["p"], flags)
        ens.close()


def test_ensemble_mixed_precision(gpu, small):
    """DGK_FLAG_MIXED in the ensemble: every instance within the mixed-
    precision target of the oracle after 2 steps (<= 1e-9 rel-L2), and the
    instances of order 1, 3, 7 equal to a single mixed-precision solver,
    whose arithmetic is the same."""
    say("[ens] mixed precision")
    ens = _ensemble(small, dgk.FLAG_MIXED)
    ens.run(2)
    errs = []
    for i, s in enumerate(small):
        got = ens.get_field(i)
        err = rel_l2(got, _oracle(s, 2))
        errs.append(err)
        if s["p"] in (1, 3, 7):
            sv = dgk.Solver.structured(*s["ne"], s["p"], s["D"], s["L"], flags=dgk.FLAG_MIXED)
            sv.set_fields(s["q"], *(np.full_like(s["q"], c) for c in s["vel"]))
            sv.run(s["dt"], 2)
            assert np.array_equal(got, sv.get_field()), i
            sv.close()
    say("[ens] mixed rel-L2 after 2 steps: " + " ".join(f"{e:.2e}" for e in errs))
    assert all(e <= 1e-9 for e in errs), errs
    ens.close()
    with pytest.raises(dgk.DGKError) as e:
        dgk.Ensemble(small, flags=dgk.FLAG_MIXED | dgk.FLAG_FMA)
    assert e.value.code == dgk.ERR_ARG


def test_ensemble_diagnostics(gpu, small):
    """Per-instance device diagnostics equal those of a single solver in
    the same state (same reduction, same order), and the error against that
    solver's field is zero."""
    ens = _ensemble(small)
    ens.run(2)
    for i, s in enumerate(small):
        say(f"[ens] diagnostics instance {i}")
        x1D, _, _ = dgk.read_operator(s["p"])
        sv = dgk.Solver.structured(*s["ne"], s["p"], s["D"], s["L"])
        sv.set_fields(s["q"], *(np.full_like(s["q"], c) for c in s["vel"]))
        sv.run(s["dt"], 2)
        want = sv.diagnostics(x1D)
        got = ens.diagnostics(i, x1D, ref=sv.field_ptr())
        for k in ("min", "max", "mass", "abs_mass", "sum_sq"):
            assert got[k] == want[k], (i, k)
        assert got["max_err"] == 0.0 and got["rel_l2"] == 0.0, i
        sv.close()
    ens.close()


def test_ensemble_errors(gpu):
    with pytest.raises(dgk.DGKError) as e:
        dgk.Ensemble([dict(p=16, ne=(2, 2, 2))])
    assert e.value.code == dgk.ERR_ORDER
    inst = _setup([dict(p=3, ne=(2, 2, 2))], 7)
    ens = dgk.Ensemble(inst)
    with pytest.raises(dgk.DGKError) as e:
        ens.step()
    assert e.value.code == dgk.ERR_STATE
    with pytest.raises(ValueError):
        ens.set_fields(0, inst[0]["q"][:, :3], *([inst[0]["q"]] * 3))
    ens.close()
    cv = dgk.Ensemble(inst, flags=dgk.FLAG_CONST_VELOCITY)
    q = inst[0]["q"]
    u = np.full_like(q, 0.5)
    u[3, 2] = 0.25
    with pytest.raises(dgk.DGKError) as e:
        cv.set_fields(0, q, u, u, u)
    assert e.value.code == dgk.ERR_ARG
    cv.close()


def test_ensemble_workload_properties(gpu):
    """The 64-instance workload (150 M DOF) for its 50 steps: finite, mass
    conserved per instance, p=3 instances bit-exact against the oracle; the
    mixed-precision ensemble on the same workload within 1e-9 rel-L2 of the
    FP64 ensemble for every instance after the 50 steps."""
    inst = ic.ensemble_workload(**{k: v for k, v in ic.ENSEMBLE.items() if k in ("n", "seed")})
    nstep = ic.ENSEMBLE["nstep"]
    say(f"[ens] workload, {len(inst)} instances, {nstep} steps")
    fields = []
    for s in inst:
        x1D, _, _ = dgk.read_operator(s["p"])
        m = dgk.mesh_setup(*s["ne"], s["p"], x1D=x1D, tables=False)
        fields.append((ic.blob(m["pos_en"], s["centre"], s["width"]), x1D))

    def build(flags):
        ens = dgk.Ensemble(inst, flags=flags)
        for i, s in enumerate(inst):
            q = fields[i][0]
            ens.set_fields(i, q, *(np.full_like(q, c) for c in s["vel"]))
        ens.set_dt([s["dt"] for s in inst])
        ens.tune(2)
        ens.run(nstep)
        return ens

    ens = build(dgk.FLAG_CONST_VELOCITY)
    say("[ens] mixed precision workload")
    mix = build(dgk.FLAG_CONST_VELOCITY | dgk.FLAG_MIXED)
    checked = 0
    worst = 0.0
    for i, s in enumerate(inst):
        got = ens.get_field(i)
        assert np.isfinite(got).all(), i
        x1D = fields[i][1]
        pp = len(x1D) - 1
        P = np.polynomial.legendre.Legendre.basis(pp)(x1D)
        w = 2.0 / (pp * (pp + 1) * P * P)
        w3 = np.einsum("i,j,k->ijk", w, w, w).ravel(order="F")
        q0 = fields[i][0]
        scale = float(np.dot(w3, np.abs(q0).sum(axis=1)))
        assert abs(float(np.dot(w3, got.sum(axis=1) - q0.sum(axis=1)))) <= 1e-13 * scale, i
        err = rel_l2(mix.get_field(i), got)
        worst = max(worst, err)
        assert 0.0 < err <= 1e-9, (i, s["p"], err)
        if s["p"] == 3 and checked < 2:
            want = _oracle(dict(s, q=np.asfortranarray(q0)), nstep)
            assert np.array_equal(got, want), i
            checked += 1
    say(f"[ens] mixed vs FP64 worst instance rel-L2 {worst:.3e}")
    assert checked == 2
    ens.close()
    mix.close()


def test_ensemble_non_unit_domains(gpu):
    """Instances on box domains of other lengths than 1 (per-axis element
    widths differ between instances and axes) give the oracle's bits."""
    specs = [dict(p=3, ne=(4, 3, 5), dom=(2.0, 0.5, 1.5)),
             dict(p=7, ne=(3, 2, 3), dom=(0.25, 3.0, 1.0)),
             dict(p=11, ne=(2, 2, 3), dom=(1.0, 1.0, 4.0))]
    g = ic.SplitMix64(4343)
    inst = []
    for s in specs:
        r = g.uniform(3)
        vel = tuple(float(2.0 * r[d] - 1.0) for d in range(3))
        x1D, D, L = dgk.read_operator(s["p"])
        m = dgk.mesh_setup(*s["ne"], s["p"], x1D=x1D, dom=s["dom"], tables=False)
        q = np.asfortranarray(np.sin(2.0 * np.pi * m["pos_en"][:, :, 0] / s["dom"][0]) +
                              0.5 * np.cos(2.0 * np.pi * m["pos_en"][:, :, 2] / s["dom"][2]))
        inst.append(dict(s, vel=vel, q=q, dt=ic.cfl_dt(s["p"], vel, s["ne"], dom=s["dom"])))
    nstep = 3
    say("[ens] non-unit domains, 3 steps")
    ens = _ensemble(inst)
    ens.run(nstep)
    for i, s in enumerate(inst):
        m = O.Mesh(*s["ne"], s["p"], dom=s["dom"])
        F = []
        for a in (s["q"], *(np.full_like(s["q"], c) for c in s["vel"])):
            f = m.field()
            f[:, :m.Ne] = a
            m.update_halo(f)
            F.append(f)
        m.run(*F, s["dt"], nstep, O.OPT1)
        assert np.array_equal(ens.get_field(i), F[0][:, :m.Ne]), i
    ens.close()
