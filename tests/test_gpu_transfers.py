"""Host <-> device transfers of whole fields.

Fields above 16 MiB go through two pinned 64 MiB staging buffers (filled and
emptied by several host threads while the DMA of the other buffer runs):
a field of three chunks, the last one partial, must arrive and come back
bit for bit, for the solver (dgk_set_fields / dgk_get_field) and for an
ensemble instance; a small field takes the direct copy.
"""
import numpy as np
import pytest

import dgk
from conftest import say

pytestmark = pytest.mark.gpu

# 45 x 41 x 19 elements at p=7: 35,055 x 512 doubles = 143.6 MB per field,
# two full 64 MiB chunks and a partial one
NE = (45, 41, 19)
P = 7


def _fields(ne, p, seed):
    rng = np.random.default_rng(seed)
    n = (p + 1) ** 3
    ncell = ne[0] * ne[1] * ne[2]
    return [np.asfortranarray(rng.standard_normal((n, ncell))) for _ in range(4)]


def test_solver_fields_round_trip(gpu):
    x1D, D, L = dgk.read_operator(P)
    q, u, v, w = _fields(NE, P, 31)
    assert q.nbytes > 2 * (64 << 20)
    say(f"[xfer] solver {NE} p={P}: {q.nbytes / 2**20:.1f} MiB per field")
    s = dgk.Solver.structured(*NE, P, D, L)
    s.set_fields(q, u, v, w)
    Ne = s.Ne
    for which, want in ((dgk.FIELD_Q, q), (dgk.FIELD_U, u), (dgk.FIELD_V, v), (dgk.FIELD_W, w)):
        got = s.get_field(which)
        assert np.array_equal(got[:, :Ne], want), which
    # a second upload over the first (the staging buffers are reused)
    s.set_fields(w, v, u, q)
    assert np.array_equal(s.get_field(dgk.FIELD_Q)[:, :Ne], w)
    assert np.array_equal(s.get_field(dgk.FIELD_W)[:, :Ne], q)
    s.close()


def test_small_field_direct_copy(gpu):
    x1D, D, L = dgk.read_operator(3)
    ne = (5, 4, 3)
    q, u, v, w = _fields(ne, 3, 32)
    s = dgk.Solver.structured(*ne, 3, D, L)
    s.set_fields(q, u, v, w)
    assert np.array_equal(s.get_field(dgk.FIELD_Q)[:, :s.Ne], q)
    s.close()


def test_ensemble_instance_round_trip(gpu):
    specs = [dict(p=P, ne=NE), dict(p=3, ne=(4, 5, 6))]
    e = dgk.Ensemble(specs)
    big = _fields(NE, P, 33)
    small = _fields((4, 5, 6), 3, 34)
    say(f"[xfer] ensemble: {big[0].nbytes / 2**20:.1f} MiB instance")
    e.set_fields(0, *big)
    e.set_fields(1, *small)
    assert np.array_equal(e.get_field(0), big[0])
    assert np.array_equal(e.get_field(1), small[0])
    e.close()


def test_uniform_velocity_check_on_a_large_field(gpu):
    """DGK_FLAG_CONST_VELOCITY checks the velocity uniform on the host, the
    scan split between threads: a single differing node anywhere (here in
    the last thread's part and in the first), or a NaN, is refused."""
    x1D, D, L = dgk.read_operator(P)
    q = _fields(NE, P, 35)[0]
    vel = (0.5, -0.25, 0.75)
    uvw = [np.full_like(q, c) for c in vel]
    s = dgk.Solver.structured(*NE, P, D, L, flags=dgk.FLAG_CONST_VELOCITY)
    s.set_fields(q, *uvw)
    for d, (col, node, val) in enumerate(((q.shape[1] - 1, q.shape[0] - 1, 0.5000001),
                                          (0, 1, -0.2),
                                          (q.shape[1] // 2, 7, np.nan))):
        bad = [a.copy(order="F") for a in uvw]
        bad[d][node, col] = val
        with pytest.raises(dgk.DGKError) as e:
            s.set_fields(q, *bad)
        assert e.value.code == dgk.ERR_ARG, d
    s.close()


def test_stateless_tendency_multichunk(gpu):
    """The stateless drop-in (dgk_advect3d_eq_cal_tend: host arrays in, dqdt
    out, per-node tables) on fields of three staging chunks equals the
    oracle's tendency bit for bit."""
    import oracle as O
    ne, p = (45, 41, 19), 7
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p)
    om = O.Mesh(*ne, p)
    rng = np.random.default_rng(36)
    F = []
    for k in range(4):
        f = om.field()
        f[:, :om.Ne] = rng.standard_normal((m["Np"], om.Ne)) if k == 0 else \
            0.5 + 0.25 * rng.standard_normal((m["Np"], om.Ne))
        om.update_halo(f)
        F.append(f)
    assert F[0].nbytes > 2 * (64 << 20)
    say(f"[xfer] stateless tendency {ne} p={p}")
    want = om.cal_tend(*F, typeid=O.OPT1)
    got = dgk.advect3d_eq_cal_tend(*F, D, np.asfortranarray(D.T), L, m["VMapM"], m["VMapP"],
                                   m["normal_fn"], m["Escale"], m["Fscale"], p + 1, m["Np"],
                                   m["NfpTot"], m["Ne"], m["NeA"], typeid=dgk.TYPE_OPT1)
    assert np.array_equal(got, want[:, :m["Ne"]])
