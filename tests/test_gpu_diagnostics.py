"""On-device diagnostics (dgk_diagnostics) and the variant autotuner
(dgk_autotune).

The reductions are checked against numpy on the downloaded field: min and
max exactly, the sums to the rounding of a different summation order; the
mass also against the oracle's closed form for a constant field.  The
autotuner must pick a listed variant, leave the state as it was and keep
the reference's bits.
"""
import numpy as np
import pytest

import dgk
import oracle as O
from conftest import say
from golden import cases as CS

pytestmark = pytest.mark.gpu


def _solver(ne, p, flags=0, seed=77):
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = CS.bump_field(m["pos_en"], seed, noise=0.01)
    vel = (0.6, -0.35, 0.8)
    s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
    s.set_fields(q, *(np.full_like(q, c) for c in vel))
    return s, x1D, D, L, q, vel


def _host_diag(q, x1D, ne, ref=None):
    w = dgk.lgl_weights(x1D)
    w3 = np.einsum("i,j,k->ijk", w, w, w).ravel(order="F")
    J = 1.0 / np.prod([2.0 * n for n in ne])  # unit cube: (1/n)^3 / 8
    d = dict(min=q.min(), max=q.max(), mass=J * float(w3 @ q.sum(axis=1)),
             abs_mass=J * float(w3 @ np.abs(q).sum(axis=1)), sum_sq=float((q * q).sum()))
    if ref is not None:
        d.update(rel_l2=float(np.linalg.norm(q - ref) / np.linalg.norm(ref)),
                 max_err=float(np.abs(q - ref).max()))
    return d


@pytest.mark.parametrize("p,ne", [(3, (4, 3, 5)), (7, (3, 4, 2)), (11, (2, 2, 3))])
def test_diagnostics_match_host(gpu, p, ne):
    s, x1D, D, L, q, vel = _solver(ne, p)
    dt = 0.002
    s.run(dt, 3)
    got = s.diagnostics(x1D)
    qh = s.get_field()
    want = _host_diag(qh, x1D, ne)
    say(f"[diag] p={p} {ne}")
    assert got["min"] == want["min"] and got["max"] == want["max"]
    for k in ("mass", "abs_mass", "sum_sq"):
        assert abs(got[k] - want[k]) <= 1e-13 * abs(want[k]), k
    # a second solver with FMA contractions as the reference field
    r, *_ = _solver(ne, p, flags=dgk.FLAG_FMA)
    r.run(dt, 3)
    got = s.diagnostics(x1D, ref=r)
    want = _host_diag(qh, x1D, ne, ref=r.get_field())
    assert got["rel_l2"] > 0.0
    assert abs(got["rel_l2"] - want["rel_l2"]) <= 1e-8 * want["rel_l2"]
    assert got["max_err"] == want["max_err"]
    # against itself: exactly zero
    assert s.diagnostics(x1D, ref=s)["rel_l2"] == 0.0
    s.close()
    r.close()


def test_diagnostics_constant_mass_and_tables_path(gpu):
    """A constant field c on the unit cube has mass c exactly up to rounding
    (the LGL weights sum to 2 per axis); the per-node table path (Escale per
    node) gives the compressed path's values."""
    ne, p, c = (3, 2, 4), 5, 0.625
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D)
    q = np.full((m["Np"], m["Ne"]), c, order="F")
    vel = (0.3, 0.2, -0.1)
    s = dgk.Solver.structured(*ne, p, D, L)
    s.set_fields(q, *(np.full_like(q, v) for v in vel))
    a = s.diagnostics(x1D)
    assert abs(a["mass"] - c) <= 1e-14
    t = dgk.Solver.from_tables(p + 1, m["Ne"], m["NeA"], D, L, m["VMapM"], m["VMapP"],
                               m["normal_fn"], m["Escale"], m["Fscale"], m["halo_src_map"],
                               flags=dgk.FLAG_FORCE_TABLES)
    t.set_fields(q, *(np.full_like(q, v) for v in vel))
    b = t.diagnostics(x1D)
    for k in ("min", "max", "mass", "abs_mass", "sum_sq"):
        assert abs(a[k] - b[k]) <= 1e-14 * max(1.0, abs(a[k])), k
    s.close()
    t.close()


@pytest.mark.parametrize("p,ne", [(3, (4, 3, 9)), (7, (3, 2, 9)), (11, (2, 3, 2))])
def test_autotune_keeps_state_and_bits(gpu, p, ne):
    s, x1D, D, L, q, vel = _solver(ne, p)
    dt, nstep = 0.001, 3
    before = s.get_field()
    v, ms = s.autotune(dt, nrep=3)
    say(f"[autotune] p={p} {ne} -> variant {v}, {ms * 1e3:.1f} us/stage")
    assert 0 <= v < s.num_variants and ms > 0.0
    assert s.variant == v
    assert np.array_equal(s.get_field(), before)
    s.run(dt, nstep)
    om = O.Mesh(*ne, p)
    F = []
    for a in (q, *(np.full_like(q, c) for c in vel)):
        f = om.field()
        f[:, :om.Ne] = a
        om.update_halo(f)
        F.append(f)
    om.run(*F, dt, nstep, O.OPT1)
    assert np.array_equal(s.get_field(), F[0][:, :om.Ne])
    s.close()


@pytest.mark.parametrize("ne,p,dom", [((3, 4, 5), 3, (1.0, 1.0, 1.0)), ((1, 2, 3), 7, (2.0, 0.5, 1.5)),
                                      ((4, 1, 1), 2, (1.0, 3.0, 0.25)), ((2, 2, 2), 11, (1.0, 1.0, 1.0))])
def test_device_mesh_build_equals_host(gpu, ne, p, dom):
    """dgk_mesh_setup_device writes the host build's tables bit for bit,
    including the halo numbering of the reference's serial pass."""
    torch = gpu
    say(f"[mesh device] {ne} p={p} dom={dom}")
    x1D, _, _ = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, dom=dom)
    keys = {"VMapM": torch.int32, "VMapP": torch.int32, "halo_src_map": torch.int32,
            "normal_fn": torch.float64, "Escale": torch.float64, "Fscale": torch.float64,
            "pos_en": torch.float64}
    dev = {k: torch.full((m[k].size,), -7, dtype=t, device="cuda") for k, t in keys.items()}
    dgk.mesh_setup_device(*ne, p, x1D, {k: v.data_ptr() for k, v in dev.items()}, dom=dom)
    for k in keys:
        got = dev[k].cpu().numpy()
        want = np.asarray(m[k]).ravel(order="F")
        assert got.dtype == want.dtype and np.array_equal(got, want), k


def test_nonfinite_state_is_reported(gpu):
    """One NaN seeded in q: dgk_run(minmax) still runs its steps and fills
    the per-step line, which shows the NaN (the reductions keep it), and
    returns DGK_ERR_NONFINITE naming the first step; the diagnostics show the
    NaN as min and max.  Without minmax nothing is checked (status 0)."""
    import ctypes as C
    s, x1D, _, _, q, vel = _solver((3, 2, 2), 3)
    bad = q.copy(order="F")
    bad[5, 7] = np.nan
    dt = 1e-3
    for with_mm in (True, False):
        say(f"[nonfinite] minmax={with_mm}")
        s.set_fields(bad, *(np.full_like(bad, c) for c in vel))
        mm = np.zeros((2, 3), order="F")
        rc = dgk.lib().dgk_run(s._h, dt, 3, mm.ctypes.data_as(C.c_void_p) if with_mm else None)
        if with_mm:
            assert rc == dgk.ERR_NONFINITE
            assert b"after step 1" in dgk.lib().dgk_last_error()
            assert np.isnan(mm).all()
        else:
            assert rc == 0
        d = s.diagnostics(x1D)
        assert np.isnan(d["min"]) and np.isnan(d["max"])
    # a finite state after the failure: status 0 again
    s.set_fields(q, *(np.full_like(q, c) for c in vel))
    mm = s.run(dt, 3, minmax=True)
    assert np.isfinite(mm).all()
    s.close()


def test_time_stages_whole_steps_and_autotune_state(gpu):
    """dgk_time_stages times whole SSP-RK3 steps only (a count that is not a
    multiple of 3 is refused), and the autotuner leaves the state as it was
    also for a stage kernel with a single variant (thread-per-node)."""
    s, x1D, D, L, q, vel = _solver((3, 3, 2), 5, flags=dgk.FLAG_NODE_KERNEL)
    assert s.num_variants == 1
    with pytest.raises(dgk.DGKError) as e:
        s.time_stages(1e-3, 4)
    assert e.value.code == dgk.ERR_ARG
    before = s.get_field()
    var, ms = s.autotune(1e-3, nrep=4)
    assert var == 0 and ms > 0.0
    assert np.array_equal(s.get_field(), before)
    s.close()
