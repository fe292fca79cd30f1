"""The oracle (oracle/dg_oracle.c) against the reference's own outputs.

tests/golden/golden.json was written by tests/golden/make_golden.py from the
reference's Fortran modules compiled from source (oracle/_ref).  Every check
here is bit-exact: SHA-256 of the raw float64 output.
"""
import numpy as np
import pytest

import cases as CS
import oracle as O


@pytest.mark.parametrize("mesh", CS.MESH_CASES, ids=lambda m: CS.case_key(*m))
def test_mesh_tables(golden, mesh):
    m = O.Mesh(*mesh)
    g = golden["mesh"][CS.case_key(*mesh)]
    for k in ("x1D", "D1D", "Lift", "VMapM", "VMapP", "halo_src_map", "normal_fn",
              "Escale", "Fscale", "pos_en"):
        assert CS.digest(getattr(m, k)) == g[k], k


@pytest.mark.parametrize("p", range(1, 16))
@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_operators(golden, p, tname):
    Nq = p + 1
    x1D, D, L = O.read_operator(p)
    Dt = np.asfortranarray(D.T)
    ix, iy, iz, fb = CS.optr_inputs(p)
    outs = {k: np.zeros((Nq ** 3, CS.OPTR_NELEM), order="F") for k in "xyzl"}
    for e in range(CS.OPTR_NELEM):
        ox, oy, oz = O.divlike(Nq, O.TYPEIDS[tname], D, Dt, ix[:, e].copy(),
                               iy[:, e].copy(), iz[:, e].copy())
        outs["x"][:, e], outs["y"][:, e], outs["z"][:, e] = ox, oy, oz
        outs["l"][:, e] = O.lift(Nq, L, fb[:, e].copy())
    g = golden["optr"][CS.case_key(p, tname)]
    for k, gk in (("x", "x"), ("y", "y"), ("z", "z"), ("l", "lift")):
        assert CS.digest(outs[k]) == g[gk]["sha256"], (p, tname, k)


def _per_block(fn, nblocks, shape_np):
    """Runs fn(e) -> tuple of (Np,) arrays for e < nblocks into (Np,nblocks)
    arrays."""
    outs = None
    for e in range(nblocks):
        r = fn(e)
        if outs is None:
            outs = [np.zeros((shape_np, nblocks), order="F") for _ in r]
        for o, v in zip(outs, r):
            o[:, e] = v
    return outs


def _gradz_as_shipped(a, Dt, Nq):
    """What the reference's gradlike z loop computes
    (mod_dg_optr_kernel_opt1.F90.erb:208-213): its first term is
    in(i,1,k)*Mat_tr(1,j), the rest in(i,j,kk)*Mat_tr(kk,k), left to right.
    A model of that slip, used only to show what the reference's output is."""
    A = a.reshape((Nq, Nq, Nq), order="F")
    out = np.empty_like(A)
    for k in range(Nq):
        for j in range(Nq):
            s = A[:, 0, k] * Dt[0, j]
            for kk in range(1, Nq):
                s = s + A[:, j, kk] * Dt[kk, k]
            out[:, j, k] = s
    return out.ravel(order="F")


@pytest.mark.parametrize("p", range(1, 16))
def test_directional_and_gradlike(golden, p):
    """dirX/dirY/dirZ and gradlike x, y equal the reference bit for bit; the
    reference's gradlike z output is its loop slip (reproduced exactly by a
    model of it), not the z derivative the oracle computes."""
    Nq = p + 1
    Np = Nq ** 3
    _, D, _ = O.read_operator(p)
    Dt = np.asfortranarray(D.T)
    a = CS.optr_inputs(p)[0]
    n = a.shape[1]
    g = golden["optr2"][CS.case_key(p)]
    dx, dy, dz = _per_block(lambda e: (O.matvec_dir(Nq, 0, D, a[:, e].copy()),
                                       O.matvec_dir(Nq, 1, Dt, a[:, e].copy()),
                                       O.matvec_dir(Nq, 2, Dt, a[:, e].copy())), n, Np)
    assert CS.digest(dx) == g["dirX"]["sha256"]
    assert CS.digest(dy) == g["dirY"]["sha256"]
    assert CS.digest(dz) == g["dirZ"]["sha256"]
    gx, gy, gz = _per_block(lambda e: O.gradlike(Nq, D, Dt, a[:, e].copy()), n, Np)
    assert CS.digest(gx) == g["gradX"]["sha256"]
    assert CS.digest(gy) == g["gradY"]["sha256"]
    assert CS.digest(gz) == CS.digest(dz)  # the intended z derivative
    assert CS.digest(gz) != g["gradZ"]["sha256"]
    shipped = _per_block(lambda e: (_gradz_as_shipped(a[:, e], Dt, Nq),), n, Np)[0]
    assert CS.digest(shipped) == g["gradZ"]["sha256"]


@pytest.mark.parametrize("p", range(1, 16))
def test_divlike_var5(golden, p):
    """var5 (the intended form) is divlike_dirXYZ in OPT1's order on each of
    five fields: fed the three golden divlike blocks as fields 1-3 (and
    again as 4-5), its outputs equal the reference's divlike_dirXYZ_P<p>."""
    Nq = p + 1
    Np = Nq ** 3
    _, D, _ = O.read_operator(p)
    Dt = np.asfortranarray(D.T)
    ix, iy, iz, _ = CS.optr_inputs(p)
    order = [0, 1, 2, 0, 1]
    v5 = [np.concatenate([a[:, e] for e in order]) for a in (ix, iy, iz)]
    ox, oy, oz = O.divlike_var5(Nq, D, Dt, *v5)
    g = golden["optr"][CS.case_key(p, "OPT1")]
    for out, k in ((ox, "x"), (oy, "y"), (oz, "z")):
        blocks = np.asfortranarray(out.reshape((Np, 5), order="F")[:, :3])
        assert CS.digest(blocks) == g[k]["sha256"], (p, k)
        assert np.array_equal(out[3 * Np:], out[:2 * Np])


def test_split_order_is_visible(golden):
    """OPT1's three partial sums (p > 7) round differently from GENERAL's
    running sum, so a restatement that ignored the split would fail the
    digests above: the golden outputs of the two orders differ for p > 7."""
    for p in range(8, 16):
        g, o = golden["optr"][CS.case_key(p, "GENERAL")], golden["optr"][CS.case_key(p, "OPT1")]
        assert any(g[k]["sha256"] != o[k]["sha256"] for k in "xyz"), p
    for p in range(1, 8):  # same order up to the sign of an exact zero
        g, o = golden["optr"][CS.case_key(p, "GENERAL")], golden["optr"][CS.case_key(p, "OPT1")]
        assert all(g[k]["l2"] == o[k]["l2"] for k in "xyz"), p


@pytest.mark.parametrize("case", CS.TEND_CASES, ids=lambda c: CS.case_key(*c[0], c[1]))
@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_tendency(golden, case, tname):
    ne, p = case
    m = O.Mesh(*ne, p)
    q, u, v, w = CS.tend_inputs(ne, p)
    F = []
    for a in (q, u, v, w):
        f = m.field()
        f[:, :m.Ne] = a
        m.update_halo(f)
        F.append(f)
    d = m.cal_tend(*F, typeid=O.TYPEIDS[tname])[:, :m.Ne]
    assert CS.digest(d) == golden["tend"][CS.case_key(*ne, p, tname)]["sha256"]


@pytest.mark.parametrize("case", CS.RUN_CASES, ids=lambda c: CS.case_key(*c[0], c[1], c[2]))
@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_trajectory(golden, case, tname):
    ne, p, nstep, vel, kind = case
    m = O.Mesh(*ne, p)
    q, u, v, w, dt = CS.run_inputs(ne, p, vel, kind, m.pos_en)
    g = golden["run"][CS.case_key(*ne, p, nstep, tname)]
    assert float(dt).hex() == g["dt"]
    F = []
    for a in (q, u, v, w):
        f = m.field()
        f[:, :m.Ne] = a
        m.update_halo(f)
        F.append(f)
    _, mm = m.run(*F, dt, nstep, O.TYPEIDS[tname], minmax=True)
    assert CS.digest(F[0][:, :m.Ne]) == g["q"]["sha256"]
    assert [[float(a).hex() for a in col] for col in mm.T] == g["minmax"]


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_live_reference_matches_oracle():
    """When the compiled reference is present, one more case outside the
    fixtures: 5x3x4 elements, p=6, 7 steps, both orders."""
    ne, p, nstep, vel = (5, 3, 4), 6, 7, (0.3, -1.0, 0.45)
    m = O.Mesh(*ne, p)
    q = CS.bump_field(m.pos_en, 4242, noise=0.05)
    u = np.full_like(q, vel[0]); v = np.full_like(q, vel[1]); w = np.full_like(q, vel[2])
    dt = 1.7e-3
    for tname in ("GENERAL", "OPT1"):
        qr, mmr, _, _ = O.ref_run(*ne, p, q, u, v, w, dt, nstep, tname)
        F = []
        for a in (q, u, v, w):
            f = m.field()
            f[:, :m.Ne] = a
            m.update_halo(f)
            F.append(f)
        _, mm = m.run(*F, dt, nstep, O.TYPEIDS[tname], minmax=True)
        assert np.array_equal(F[0][:, :m.Ne], qr)
        assert np.array_equal(mm, mmr)
