"""The matrix-core stage (DGK_FLAG_MFMA, p=15 with uniform velocity): FP64
v_mfma_f64_16x16x4_f64 contractions, and with DGK_FLAG_MIXED FP32
v_mfma_f32_16x16x4_f32 contractions of each line's remainder after an FP64
degree-7 fit (the fit and its derivative as FP64 MFMA products).

Fused arithmetic, so checked against the oracle's bits within the FMA
variants' bounds (FP64: <= 1e-12 rel-L2 per step, <= 1e-10 after nstep) and
the mixed-precision target (<= 1e-9 rel-L2 after nstep), and against the
bit-identical variant on the 16^3 p=15 workload; every tuning variant (four
or eight wavefronts per element, element order) gives the same bits."""
import numpy as np
import pytest

import dgk
import oracle as O
from conftest import say
from dgk import ic
from golden import cases as CS

pytestmark = pytest.mark.gpu

CV = dgk.FLAG_CONST_VELOCITY


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _oracle(ne, p, q, vel, dt, nstep):
    m = O.Mesh(*ne, p)
    F = []
    for a in (q, *(np.full_like(q, c) for c in vel)):
        f = m.field()
        f[:, :m.Ne] = a
        m.update_halo(f)
        F.append(f)
    m.run(*F, dt, nstep, O.OPT1)
    return F[0][:, :m.Ne]


@pytest.mark.parametrize("ne", [(2, 3, 2), (3, 1, 4)])
def test_small_meshes_against_oracle(gpu, ne):
    p, vel = 15, (0.8, 0.6, -0.4)
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = CS.bump_field(m["pos_en"], 31, noise=0.01)
    dt = ic.cfl_dt(p, vel, ne)
    for nstep, tol64, tol32 in ((1, 1e-12, 1e-9), (5, 1e-10, 1e-9)):
        want = _oracle(ne, p, q, vel, dt, nstep)
        for flags, tol in ((dgk.FLAG_MFMA | CV, tol64), (dgk.FLAG_MFMA | dgk.FLAG_MIXED | CV, tol32)):
            say(f"[mfma] {ne} flags={flags} steps={nstep}")
            s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
            s.set_fields(q, *(np.full_like(q, c) for c in vel))
            s.run(dt, nstep)
            err = rel_l2(s.get_field(), want)
            assert 0.0 < err <= tol, (flags, nstep, err)
            s.close()


def test_workload_and_variants(gpu):
    """16^3 p=15 sine modes to k=8: FP64 MFMA within 1e-12 of the
    bit-identical variant after one step and 1e-10 after 50; the FP32 form
    within 1e-10 after one step and the 1e-9 target after 50; every tuning
    variant the same bits."""
    spec = ic.WORKLOADS["c5_16cube_p15"]
    ne, p, vel = spec["ne"], spec["p"], spec["vel"]
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.initial_field(spec["ic"], m["pos_en"], spec["seed"])
    dt = ic.cfl_dt(p, vel, ne)
    uvw = [np.full_like(q, c) for c in vel]
    e = dgk.Solver.structured(*ne, p, D, L)
    e.set_fields(q, *uvw)
    runs = []
    for flags, b1, b50 in ((dgk.FLAG_MFMA | CV, 1e-12, 1e-10),
                           (dgk.FLAG_MFMA | dgk.FLAG_MIXED | CV, 1e-10, 1e-9)):
        s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
        s.set_fields(q, *uvw)
        runs.append((s, flags, b1, b50))
    e.run(dt, 1)
    for s, flags, b1, _ in runs:
        s.run(dt, 1)
        assert s.diagnostics(x1D, ref=e)["rel_l2"] <= b1, flags
    e.run(dt, spec["nstep"] - 1)
    for s, flags, _, b50 in runs:
        s.run(dt, spec["nstep"] - 1)
        err = s.diagnostics(x1D, ref=e)["rel_l2"]
        say(f"[mfma] flags={flags} rel-L2 after {spec['nstep']} steps {err:.3e}")
        assert 0.0 < err <= b50, (flags, err)
    # tuning variants: the same bits after 2 steps from the initial field
    for s, flags, _, _ in runs:
        outs = []
        for v in range(s.num_variants):
            s.set_variant(v)
            s.set_fields(q, *uvw)
            s.run(dt, 2)
            outs.append(s.get_field())
        for v, o in enumerate(outs[1:], 1):
            assert np.array_equal(o, outs[0]), (flags, v)
        s.close()
    e.close()


def test_refused_where_not_built(gpu):
    for p, flags in ((7, dgk.FLAG_MFMA | CV), (15, dgk.FLAG_MFMA),
                     (15, dgk.FLAG_MFMA | CV | dgk.FLAG_NODE_KERNEL)):
        _, D, L = dgk.read_operator(p)
        with pytest.raises(dgk.DGKError) as e:
            dgk.Solver.structured(2, 2, 2, p, D, L, flags=flags)
        assert e.value.code == dgk.ERR_ARG, (p, flags)
