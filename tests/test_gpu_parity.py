"""Device path (libdgk.so on the MI355X) against the reference.

Exact variants (default, DGOptrKernel_OptType "HIP") must reproduce the
reference's FP64 results bit for bit: digests equal to the golden fixtures
written from the compiled reference, and arrays equal to the oracle's on the
same inputs.  FMA variants must stay within 1e-12 relative L2 after one step
and 1e-10 after nstep.
"""
import numpy as np
import pytest

import cases as CS
import dgk
import oracle as O
from conftest import full_field, say
from dgk import ic

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def oracle_fields(m, q, u, v, w):
    F = []
    for a in (q, u, v, w):
        f = m.field()
        f[:, :m.Ne] = a
        m.update_halo(f)
        F.append(f)
    return F


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).cuda()


def from_dev(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


# ---------------------------------------------------------------------------
# batched element operators


@pytest.mark.parametrize("p", range(1, 16))
@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_batched_operators_bit_exact(gpu, golden, p, tname):
    torch = gpu
    say(f"[operators] p={p} {tname}")
    Nq = p + 1
    Np = Nq ** 3
    _, D, L = dgk.read_operator(p)
    ix, iy, iz, fb = CS.optr_inputs(p)
    n = CS.OPTR_NELEM
    dD, dDt, dL = to_dev(torch, D), to_dev(torch, np.asfortranarray(D.T)), to_dev(torch, L)
    dix, diy, diz, dfb = (to_dev(torch, a) for a in (ix, iy, iz, fb))
    ox, oy, oz, ol = (torch.zeros(Np * n, dtype=torch.float64, device="cuda") for _ in range(4))
    tid = dgk.optr_kernel_setup(tname)
    dgk.divlike_batched(Nq, n, tid, dD.data_ptr(), dDt.data_ptr(), dix.data_ptr(),
                        diy.data_ptr(), diz.data_ptr(), ox.data_ptr(), oy.data_ptr(),
                        oz.data_ptr())
    dgk.lift_batched(Nq, n, dL.data_ptr(), dfb.data_ptr(), ol.data_ptr())
    torch.cuda.synchronize()
    g = golden["optr"][CS.case_key(p, tname)]
    for t, k in ((ox, "x"), (oy, "y"), (oz, "z"), (ol, "lift")):
        assert CS.digest(from_dev(t, (Np, n))) == g[k]["sha256"], k


@pytest.mark.parametrize("p", [1, 3, 5, 7, 12, 15])
def test_batched_operators_many_elements(gpu, golden, p):
    """Large batches: the lift walks its elements in at most 4096 persistent
    workgroups (several elements each at low order), the other operators
    take one workgroup per element.  With the golden elements repeated past
    4096 lift workgroups' worth, every element's outputs equal the golden
    elements' bits."""
    torch = gpu
    Nq = p + 1
    Np, Nft = Nq ** 3, 6 * Nq * Nq
    epb = 256 // Np if Np <= 256 else 1  # elements per lift workgroup
    n0 = CS.OPTR_NELEM
    reps = (4096 * epb + 2048) // n0 + 1
    n = n0 * reps
    say(f"[operators many] p={p} nelem={n}")
    _, D, L = dgk.read_operator(p)
    ix, iy, iz, fb = CS.optr_inputs(p)
    tile = lambda a: torch.from_numpy(np.tile(np.asfortranarray(a).ravel(order="F"), reps)).cuda()
    dD, dDt, dL = to_dev(torch, D), to_dev(torch, np.asfortranarray(D.T)), to_dev(torch, L)
    dix, diy, diz, dfb = (tile(a) for a in (ix, iy, iz, fb))
    assert dix.numel() == Np * n and dfb.numel() == Nft * n
    outs = [torch.zeros(Np * n, dtype=torch.float64, device="cuda") for _ in range(7)]
    tid = dgk.optr_kernel_setup("OPT1")
    dgk.divlike_batched(Nq, n, tid, dD.data_ptr(), dDt.data_ptr(), dix.data_ptr(),
                        diy.data_ptr(), diz.data_ptr(), *(o.data_ptr() for o in outs[:3]))
    dgk.lift_batched(Nq, n, dL.data_ptr(), dfb.data_ptr(), outs[3].data_ptr())
    dgk.matvec_dir_batched(Nq, n, 0, dD.data_ptr(), dix.data_ptr(), outs[4].data_ptr())
    dgk.gradlike_batched(Nq, n, dD.data_ptr(), dDt.data_ptr(), dix.data_ptr(),
                         outs[5].data_ptr(), outs[6].data_ptr(), outs[4].data_ptr())
    torch.cuda.synchronize()
    g, g2 = golden["optr"][CS.case_key(p, "OPT1")], golden["optr2"][CS.case_key(p)]
    want = [g["x"], g["y"], g["z"], g["lift"], None, g2["gradX"], g2["gradY"]]
    for o, w in zip(outs, want):
        if w is None:
            continue
        blocks = o.view(reps, n0 * Np)
        assert torch.equal(blocks, blocks[:1].expand(reps, -1))
        assert CS.digest(from_dev(blocks[0], (Np, n0))) == w["sha256"]
    # dirX was overwritten by gradlike's z output: check dirX alone
    dgk.matvec_dir_batched(Nq, n, 0, dD.data_ptr(), dix.data_ptr(), outs[4].data_ptr())
    torch.cuda.synchronize()
    blocks = outs[4].view(reps, n0 * Np)
    assert torch.equal(blocks, blocks[:1].expand(reps, -1))
    assert CS.digest(from_dev(blocks[0], (Np, n0))) == g2["dirX"]["sha256"]


@pytest.mark.parametrize("p", range(1, 16))
def test_batched_directional_gradlike_var5(gpu, golden, p):
    """dirX/dirY/dirZ and gradlike x, y bit-exact against the reference;
    gradlike z against the oracle's stated form (the reference's z loop
    slips, mod_dg_optr_kernel_opt1.F90.erb:208); var5 against the
    reference's divlike blocks (its own y index slips, :400)."""
    torch = gpu
    say(f"[operators2] p={p}")
    Nq = p + 1
    Np = Nq ** 3
    _, D, _ = dgk.read_operator(p)
    Dt = np.asfortranarray(D.T)
    ix, iy, iz, _ = CS.optr_inputs(p)
    n = CS.OPTR_NELEM
    g = golden["optr2"][CS.case_key(p)]
    dD, dDt, da = to_dev(torch, D), to_dev(torch, Dt), to_dev(torch, ix)
    outs = [torch.zeros(Np * n, dtype=torch.float64, device="cuda") for _ in range(6)]
    for d, M in ((0, dD), (1, dDt), (2, dDt)):
        dgk.matvec_dir_batched(Nq, n, d, M.data_ptr(), da.data_ptr(), outs[d].data_ptr())
    dgk.gradlike_batched(Nq, n, dD.data_ptr(), dDt.data_ptr(), da.data_ptr(),
                         *(o.data_ptr() for o in outs[3:]))
    torch.cuda.synchronize()
    res = [from_dev(o, (Np, n)) for o in outs]
    for r, k in zip(res[:5], ("dirX", "dirY", "dirZ", "gradX", "gradY")):
        assert CS.digest(r) == g[k]["sha256"], k
    for e in range(n):
        want = O.gradlike(Nq, D, Dt, ix[:, e].copy())[2]
        assert np.array_equal(res[5][:, e], want)
    assert CS.digest(res[5]) != g["gradZ"]["sha256"]
    # var5: fields (block 0, 1, 2, 0, 1) of each of two elements
    order = [0, 1, 2, 0, 1]
    v5 = [np.concatenate([a[:, b] for _ in range(2) for b in order]) for a in (ix, iy, iz)]
    din = [to_dev(torch, a) for a in v5]
    dout = [torch.zeros(Np * 10, dtype=torch.float64, device="cuda") for _ in range(3)]
    dgk.divlike_var5_batched(Nq, 2, dD.data_ptr(), dDt.data_ptr(), *(t.data_ptr() for t in din),
                             *(t.data_ptr() for t in dout))
    torch.cuda.synchronize()
    go = golden["optr"][CS.case_key(p, "OPT1")]
    for t, k in zip(dout, "xyz"):
        r = from_dev(t, (Np, 10))
        assert CS.digest(np.asfortranarray(r[:, :3])) == go[k]["sha256"], k
        assert CS.digest(np.asfortranarray(r[:, 5:8])) == go[k]["sha256"], k
    with pytest.raises(dgk.DGKError):
        dgk.matvec_dir_batched(Nq, n, 3, dD.data_ptr(), da.data_ptr(), outs[0].data_ptr())


# ---------------------------------------------------------------------------
# tendency (advect3d_eq_cal_tend)


# orders with a pencil stage kernel (every Nq; 2, 4, 8 on one wavefront, the
# others on multi-wavefront workgroups)
PENCIL_NQ = tuple(range(2, 17))

def _solver_variants(ne, p, D, L, typeid, flags=0):
    m = dgk.mesh_setup(*ne, p)
    s = dgk.Solver.structured(*ne, p, D, L, typeid=typeid, flags=flags)
    assert s.kernel_kind == ("pencil" if p + 1 in PENCIL_NQ else "node")
    yield "structured-" + s.kernel_kind, s
    if p + 1 in PENCIL_NQ:
        s = dgk.Solver.structured(*ne, p, D, L, typeid=typeid, flags=flags | dgk.FLAG_NODE_KERNEL)
        assert s.kernel_kind == "node"
        yield "structured-node", s
    args = (p + 1, m["Ne"], m["NeA"], D, L, m["VMapM"], m["VMapP"], m["normal_fn"],
            m["Escale"], m["Fscale"], m["halo_src_map"])
    s = dgk.Solver.from_tables(*args, typeid=typeid, flags=flags)
    assert s.compressed
    yield "tables-compressed", s
    s = dgk.Solver.from_tables(*args, typeid=typeid, flags=flags | dgk.FLAG_FORCE_TABLES)
    assert not s.compressed
    yield "tables", s


@pytest.mark.parametrize("case", CS.TEND_CASES, ids=lambda c: CS.case_key(*c[0], c[1]))
@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_tendency_bit_exact(gpu, golden, case, tname):
    ne, p = case
    _, D, L = dgk.read_operator(p)
    q, u, v, w = CS.tend_inputs(ne, p)
    want = golden["tend"][CS.case_key(*ne, p, tname)]["sha256"]
    tid = dgk.optr_kernel_setup(tname)
    for name, s in _solver_variants(ne, p, D, L, tid):
        say(f"[tend] {ne} p={p} {tname} {name}")
        s.set_fields(q, u, v, w)
        assert CS.digest(s.cal_tend()) == want, name
        s.close()
    say(f"[tend] {ne} p={p} {tname} stateless drop-in")
    m = dgk.mesh_setup(*ne, p)
    F = [full_field(a, m["NeA"]) for a in (q, u, v, w)]
    # the stateless call expects halos filled, as the reference's caller does
    om = O.Mesh(*ne, p)
    for f in F:
        om.update_halo(f)
    d = dgk.advect3d_eq_cal_tend(*F, D, np.asfortranarray(D.T), L, m["VMapM"], m["VMapP"],
                                 m["normal_fn"], m["Escale"], m["Fscale"], p + 1, m["Np"],
                                 m["NfpTot"], m["Ne"], m["NeA"], typeid=tid)
    assert CS.digest(d) == want


def _lift_cases(L, p):
    """Lifts the pencil kernels' packed form must reproduce bit for bit or
    refuse: (name, Lift, packs).  The packed form holds each value as its
    group's smallest bit pattern plus a byte, a group being one face's
    values with one normal index, e.g. face -z (f=4) at depth k: L[:,:,k,4]."""
    Nq = p + 1
    near = L.copy()  # moved by up to 200 units in the last place, within groups
    rng = np.random.default_rng(5)
    bits = near.view(np.int64)
    bits += rng.integers(0, 200, size=bits.shape)  # away from zero, either sign
    wide = L.copy()  # one value of group (f=4, k=1) 300 units away
    wide.view(np.int64)[Nq // 2, 0, 1, 4] += 300
    sign = L.copy()  # one value of group (f=1, i=Nq-1) with the other sign
    sign[Nq - 1, 1, 0, 1] = -sign[Nq - 1, 1, 0, 1]
    return [("near", near, True), ("wide", wide, False), ("sign", sign, False)]


@pytest.mark.parametrize("p", [3, 7])
def test_lift_packing_is_lossless_or_refused(gpu, p):
    """The pencil kernels read the dense Lift as byte offsets from group
    bases.  A Lift whose groups span up to 255 units in the last place runs
    on them and gives the oracle's bits with that Lift; one that does not
    pack (a group 300 units wide, a group of mixed signs) runs on the node
    kernel, bit-exact too, and the forms that need the pencil kernel ask for
    the factored lift."""
    ne = (3, 2, 2)
    _, D, L = dgk.read_operator(p)
    q, u, v, w = CS.tend_inputs(ne, p)
    om = O.Mesh(*ne, p)
    F = [full_field(a, om.NeA) for a in (q, u, v, w)]
    for f in F:
        om.update_halo(f)
    for name, Lx, packs in _lift_cases(L, p):
        say(f"[lift pack] p={p} {name}")
        om.Lift = np.asfortranarray(Lx)
        want = om.cal_tend(*F, O.OPT1)[:, :om.Ne]
        s = dgk.Solver.structured(*ne, p, D, Lx, typeid=dgk.TYPE_OPT1)
        assert s.kernel_kind == ("pencil" if packs else "node"), name
        s.set_fields(q, u, v, w)
        assert np.array_equal(s.cal_tend(), want), name
        s.close()
        if packs:
            continue
        with pytest.raises(dgk.DGKError, match="FACTORED_LIFT"):
            dgk.Solver.structured(*ne, p, D, Lx, typeid=dgk.TYPE_OPT1,
                                  flags=dgk.FLAG_CONST_VELOCITY)
        s = dgk.Solver.structured(*ne, p, D, Lx, typeid=dgk.TYPE_OPT1,
                                  flags=dgk.FLAG_CONST_VELOCITY | dgk.FLAG_FACTORED_LIFT)
        assert s.kernel_kind == "pencil"
        s.close()


# ---------------------------------------------------------------------------
# trajectories (main.f90 step loop)


@pytest.mark.parametrize("case", CS.RUN_CASES, ids=lambda c: CS.case_key(*c[0], c[1], c[2]))
@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_trajectory_bit_exact(gpu, golden, case, tname):
    ne, p, nstep, vel, kind = case
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q, u, v, w, dt = CS.run_inputs(ne, p, vel, kind, m["pos_en"])
    g = golden["run"][CS.case_key(*ne, p, nstep, tname)]
    tid = dgk.optr_kernel_setup(tname)
    for name, s in _solver_variants(ne, p, D, L, tid):
        say(f"[run] {ne} p={p} nstep={nstep} {tname} {name}")
        s.set_fields(q, u, v, w)
        mm = s.run(dt, nstep, minmax=True)
        assert CS.digest(s.get_field()) == g["q"]["sha256"], name
        assert [[float(a).hex() for a in col] for col in mm.T] == g["minmax"], name
        s.close()


FMA_VARIANTS = [("fma", dgk.FLAG_FMA), ("fma-node", dgk.FLAG_FMA | dgk.FLAG_NODE_KERNEL),
                ("fma-factored", dgk.FLAG_FMA | dgk.FLAG_FACTORED_LIFT),
                ("factored", dgk.FLAG_FACTORED_LIFT)]


@pytest.mark.parametrize("case", CS.RUN_CASES, ids=lambda c: CS.case_key(*c[0], c[1], c[2]))
@pytest.mark.parametrize("variant", FMA_VARIANTS, ids=lambda v: v[0])
def test_trajectory_fma_tolerance(gpu, case, variant):
    """FMA contractions and/or the factored lift: rel-L2 <= 1e-12 after one
    step, <= 1e-10 after nstep against the oracle (OPT1 order)."""
    ne, p, nstep, vel, kind = case
    name, flags = variant
    if (flags & dgk.FLAG_FACTORED_LIFT) and p + 1 not in PENCIL_NQ:
        pytest.skip("the factored lift exists in the pencil kernel only")
    _, D, L = dgk.read_operator(p)
    om = O.Mesh(*ne, p)
    q, u, v, w, dt = CS.run_inputs(ne, p, vel, kind, om.pos_en)
    for steps, tol in ((1, 1e-12), (nstep, 1e-10)):
        say(f"[{name}] {ne} p={p} steps={steps}")
        F = oracle_fields(om, q, u, v, w)
        om.run(*F, dt, steps, O.OPT1)
        s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
        s.set_fields(q, u, v, w)
        s.run(dt, steps)
        err = rel_l2(s.get_field(), F[0][:, :om.Ne])
        assert err <= tol, err
        assert err > 0.0  # a different rounding from the reference's
        s.close()


def test_permuted_element_numbering(gpu):
    """A valid mesh whose elements are renumbered: the compressed form (a
    6-neighbour table) accepts any numbering, and the per-node table path is
    forced as well; both equal the oracle run on the same permuted tables bit
    for bit."""
    ne, p, nstep = (3, 2, 2), 4, 4
    om = O.Mesh(*ne, p)
    Np, Ne, NeA = om.Np, om.Ne, om.NeA
    perm = np.random.default_rng(5).permutation(Ne)  # old element e -> new perm[e]

    def remap(idx1):  # 1-based flat indices into (Np,NeA)
        i = idx1 - 1
        own = i < Np * Ne
        e, n = i // Np, i % Np
        out = i.copy()
        out[own] = n[own] + perm[e[own]] * Np
        return (out + 1).astype(np.int32)

    inv = np.argsort(perm)
    VM = np.asfortranarray(remap(om.VMapM)[:, inv])
    VP = np.asfortranarray(remap(om.VMapP)[:, inv])
    nrm = np.asfortranarray(om.normal_fn[:, inv, :])
    Es = np.asfortranarray(om.Escale[:, inv, :])
    Fs = np.asfortranarray(om.Fscale[:, inv])
    halo = remap(om.halo_src_map)
    q0 = CS.uniform(77, (Np, Ne))
    vel = [CS.uniform(78 + d, (Np, Ne)) for d in range(3)]
    qn, un, vn, wn = (np.asfortranarray(a[:, inv]) for a in (q0, *vel))
    # oracle on the permuted tables
    om.VMapM, om.VMapP, om.normal_fn, om.Escale, om.Fscale, om.halo_src_map = VM, VP, nrm, Es, Fs, halo
    F = oracle_fields(om, qn, un, vn, wn)
    dt = 2e-3
    om.run(*F, dt, nstep, O.OPT1)
    for flags, compressed in ((0, True), (dgk.FLAG_FORCE_TABLES, False)):
        say(f"[permuted] flags={flags}")
        s = dgk.Solver.from_tables(p + 1, Ne, NeA, om.D1D, om.Lift, VM, VP, nrm, Es, Fs, halo,
                                   flags=flags)
        assert s.compressed == compressed
        s.set_fields(qn, un, vn, wn)
        s.run(dt, nstep)
        assert np.array_equal(s.get_field(), F[0][:, :Ne])
        s.close()


def test_stretched_mesh_keeps_tables(gpu):
    """Element-dependent Escale/Fscale (a mesh stretched along x) cannot be
    compressed to per-axis constants: the per-node tables are used and the
    result equals the oracle's on the same tables."""
    ne, p, nstep = (4, 2, 3), 3, 3
    om = O.Mesh(*ne, p)
    Ne = om.Ne
    ix = np.arange(Ne) % ne[0]
    widths = np.array([0.1, 0.2, 0.3, 0.4])  # element widths along x, sum 1
    Es = om.Escale.copy(order="F")
    Es[:, :, 0] = (2.0 / widths[ix])[None, :]
    Fs = om.Fscale.copy(order="F")
    Nfp = om.Nfp
    for f in (1, 3):  # +x, -x faces
        Fs[f * Nfp:(f + 1) * Nfp, :] = (2.0 / widths[ix])[None, :]
    om.Escale, om.Fscale = Es, Fs
    q = CS.uniform(91, (om.Np, Ne))
    vel = [np.full((om.Np, Ne), c, order="F") for c in (0.6, -0.3, 0.2)]
    F = oracle_fields(om, q, *vel)
    dt = 1e-3
    om.run(*F, dt, nstep, O.OPT1)
    say("[stretched] tables path")
    s = dgk.Solver.from_tables(p + 1, Ne, om.NeA, om.D1D, om.Lift, om.VMapM, om.VMapP,
                               om.normal_fn, Es, Fs, om.halo_src_map)
    assert not s.compressed
    s.set_fields(q, *vel)
    s.run(dt, nstep)
    assert np.array_equal(s.get_field(), F[0][:, :Ne])
    s.close()


@pytest.mark.parametrize("p,ne", [(1, (33, 34, 9)), (2, (40, 3, 8)), (3, (65, 2, 8)),
                                  (7, (35, 2, 8)), (11, (2, 33, 8))])
def test_tiled_order_with_several_tiles(gpu, p, ne):
    """Extents above 32 elements are split into several (uneven) tiles per
    axis, with tile slots outside the mesh (33 -> 2 x 17, 65 -> 3 x 22):
    every element is still updated once, with the reference's bits."""
    nstep, vel, kind = 2, (0.4, -0.3, 0.6), "rand"
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q, u, v, w, dt = CS.run_inputs(ne, p, vel, kind, m["pos_en"])
    om = O.Mesh(*ne, p)
    F = oracle_fields(om, q, u, v, w)
    om.run(*F, dt, nstep, O.OPT1)
    s = dgk.Solver.structured(*ne, p, D, L)
    for var in (0, 1):
        say(f"[tiles] {ne} p={p} variant={var}")
        s.set_variant(var)
        assert s.query(9) == 1  # tiled order
        s.set_fields(q, u, v, w)
        s.run(dt, nstep)
        assert np.array_equal(s.get_field(), F[0][:, :om.Ne]), var
    s.close()


@pytest.mark.parametrize("p", range(1, 16))
@pytest.mark.parametrize("ne", [(3, 2, 9), (4, 5, 3)])
def test_every_tuning_variant_is_bit_exact(gpu, p, ne):
    """All code shapes the autotuner may pick (tiled and linear element
    orders included: NeZ = 9 gives uneven z-slabs per XCD) give the
    reference's bits."""
    nstep, vel, kind = 3, (0.7, 0.2, -0.5), "rand"
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q, u, v, w, dt = CS.run_inputs(ne, p, vel, kind, m["pos_en"])
    om = O.Mesh(*ne, p)
    F = oracle_fields(om, q, u, v, w)
    om.run(*F, dt, nstep, O.OPT1)
    s = dgk.Solver.structured(*ne, p, D, L)
    for var in range(s.num_variants):
        say(f"[variant] {ne} p={p} variant={var}")
        s.set_variant(var)
        assert s.variant == var
        assert s.query(9) == (1 if ne[2] >= 8 and not var & 2 else 0)
        s.set_fields(q, u, v, w)
        s.run(dt, nstep)
        assert np.array_equal(s.get_field(), F[0][:, :om.Ne]), var
    s.close()


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 7, 8, 9, 11, 13, 15])
@pytest.mark.parametrize("factored", [False, True])
def test_constant_velocity_form(gpu, p, factored):
    """DGK_FLAG_CONST_VELOCITY streams no velocity and gives the same bits as
    the per-node velocity path; a non-uniform velocity is refused."""
    ne, nstep, vel = (3, 4, 8), 4, (0.45, -0.8, 0.3)
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = CS.bump_field(m["pos_en"], 123, noise=0.01)
    dt = ic.cfl_dt(p, vel, ne)
    uvw = [np.full_like(q, c) for c in vel]
    base = dgk.FLAG_FACTORED_LIFT if factored else 0
    outs = []
    for flags in (base, base | dgk.FLAG_CONST_VELOCITY):
        say(f"[const velocity] p={p} flags={flags}")
        s = dgk.Solver.structured(*ne, p, D, L, flags=flags)
        s.set_fields(q, *uvw)
        s.run(dt, nstep)
        outs.append(s.get_field())
        s.close()
    assert np.array_equal(outs[0], outs[1])
    s = dgk.Solver.structured(*ne, p, D, L, flags=dgk.FLAG_CONST_VELOCITY)
    # the uniform velocity is written on the device, halo slots included
    s.set_fields(q, *uvw)
    for which, c in zip((dgk.FIELD_U, dgk.FIELD_V, dgk.FIELD_W), vel):
        f = s.get_field(which, owned=False)
        assert f.shape == (s.Np, s.NeA) and np.all(f == c)
    bad = uvw[1].copy(order="F")
    bad[5, 2] += 1e-3
    with pytest.raises(dgk.DGKError) as e:
        s.set_fields(q, uvw[0], bad, uvw[2])
    assert e.value.code == dgk.ERR_ARG
    s.close()


def test_graph_step_equals_stagewise(gpu):
    """Replays of the captured time step, with eager work and host syncs in
    between, equal the stage-by-stage launches and dgk_run bit for bit."""
    ne, p = (4, 3, 5), 5
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.gaussian_noise(m["pos_en"], 31)
    vel = [np.full_like(q, c) for c in (0.4, 0.9, -0.2)]
    dt = ic.cfl_dt(p, (0.4, 0.9, -0.2), ne)
    a = dgk.Solver.structured(*ne, p, D, L)
    b = dgk.Solver.structured(*ne, p, D, L)
    a.set_fields(q, *vel)
    b.set_fields(q, *vel)
    for it in range(5):
        say(f"[graph] replay {it}")
        a.step(dt)
        a.synchronize()
        for st in (1, 2, 3):
            b.rk_stage(st, dt)
        b.synchronize()
        _ = b.cal_tend()  # eager work between replays
        assert np.array_equal(a.get_field(), b.get_field()), it
    # dgk_run launches the stages on the stream: the same bits as the replays
    c = dgk.Solver.structured(*ne, p, D, L)
    c.set_fields(q, *vel)
    c.run(dt, 2)
    c.run(dt, 3)
    assert np.array_equal(a.get_field(), c.get_field())
    a.close()
    b.close()
    c.close()


def test_q0_placement_and_stage_sequence(gpu):
    """Stage 1 keeps the buffer it read as q0 instead of copying it: q0 then
    holds the state the step started from through the later stages, q and
    q0 never share a buffer, and any sequence of stages combines with that
    q0 -- each stage equals q_new = a q0 + b (q + dt dqdt) evaluated on the
    host from the device tendency (mod_common.f90:29-30, main.f90:53-73)."""
    ne, p = (3, 4, 2), 3
    rk_a, rk_b = (0.0, 0.75, 1.0 / 3.0), (1.0, 0.25, 2.0 / 3.0)
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, x1D=x1D, tables=False)
    q = ic.gaussian_noise(m["pos_en"], 47)
    vel = [np.full_like(q, c) for c in (0.3, -0.8, 0.5)]
    dt = ic.cfl_dt(p, (0.3, -0.8, 0.5), ne)
    s = dgk.Solver.structured(*ne, p, D, L)
    t = dgk.Solver.structured(*ne, p, D, L)
    s.set_fields(q, *vel)
    cur, q0 = s.get_field(), None
    for n, st in enumerate((1, 2, 2, 3, 1, 3, 3)):
        say(f"[q0] stage {st} ({n})")
        t.set_fields(cur, *vel)
        dq = t.cal_tend()
        if st == 1:
            q0 = cur
        want = rk_a[st - 1] * q0 + rk_b[st - 1] * (cur + dt * dq)
        s.rk_stage(st, dt)
        cur = s.get_field()
        assert np.array_equal(cur, want), (n, st)
        assert np.array_equal(s.get_field(dgk.FIELD_Q0), q0), (n, st)
        assert s.field_ptr(dgk.FIELD_Q) != s.field_ptr(dgk.FIELD_Q0)
    s.close()
    t.close()


def test_state_errors(gpu):
    _, D, L = dgk.read_operator(2)
    s = dgk.Solver.structured(2, 2, 2, 2, D, L)
    with pytest.raises(dgk.DGKError) as e:
        s.step(1e-3)
    assert e.value.code == dgk.ERR_STATE
    s.set_fields(*(np.ones((27, 8), order="F") for _ in range(4)))
    with pytest.raises(dgk.DGKError):
        s.rk_stage(4, 1e-3)
    with pytest.raises(ValueError):
        s.set_fields(*(np.ones((27, 7), order="F") for _ in range(4)))
    s.close()


@pytest.mark.parametrize("case", CS.RUN_CASES,
                         ids=lambda c: CS.case_key(*c[0], c[1], c[2]))
def test_mixed_precision_error(gpu, case):
    """DGK_FLAG_MIXED against the oracle.  Each line's cubic part is
    differentiated in FP64 and the remainder contracted in FP32 with the
    operator split hi + lo; a line whose remainder exceeds 1/16 of the line
    takes the FP64 contraction; lift, flux differences, residual and RK are
    FP64.  Every case within the mixed-precision target (<= 1e-9 after
    nstep, <= 5e-10 after one step).  Smooth fields ("bump") at p >= 4 go
    through FP32 and are not FP64's bits (> 1e-14); at p <= 3 the cubic fit
    is the whole line, no FP32 is left and the result stays at FP64 rounding
    (<= 1e-13).  Uniform random node values ("rand") mostly have no small
    remainder and take the FP64 contraction.  Measured: bump p=7 4e-12 /
    1.1e-11, p <= 3 <= 1e-15, rand 1e-15 .. 3e-12."""
    ne, p, nstep, vel, kind = case
    _, D, L = dgk.read_operator(p)
    om = O.Mesh(*ne, p)
    q, u, v, w, dt = CS.run_inputs(ne, p, vel, kind, om.pos_en)
    fp32 = kind == "bump" and p >= 4
    bounds = ((1, 5e-10), (nstep, 1e-9)) if p >= 4 else ((1, 1e-13), (nstep, 1e-13))
    for steps, tol in bounds:
        say(f"[mixed] {ne} p={p} steps={steps}")
        F = oracle_fields(om, q, u, v, w)
        om.run(*F, dt, steps, O.OPT1)
        s = dgk.Solver.structured(*ne, p, D, L, flags=dgk.FLAG_MIXED)
        s.set_fields(q, u, v, w)
        s.run(dt, steps)
        err = rel_l2(s.get_field(), F[0][:, :om.Ne])
        assert err <= tol, err
        if fp32:
            assert err > 1e-14, err
        s.close()


def test_mixed_precision_refused_where_not_built(gpu):
    for p, flags in ((7, dgk.FLAG_MIXED | dgk.FLAG_FMA), (5, dgk.FLAG_MIXED | dgk.FLAG_FMA),
                     (7, dgk.FLAG_MIXED | dgk.FLAG_NODE_KERNEL)):
        _, D, L = dgk.read_operator(p)
        with pytest.raises(dgk.DGKError) as e:
            dgk.Solver.structured(2, 2, 2, p, D, L, flags=flags)
        assert e.value.code == dgk.ERR_ARG


@pytest.mark.parametrize("ne,p,dom", [((3, 4, 5), 7, (2.0, 0.5, 1.5)),
                                      ((4, 3, 2), 11, (0.25, 3.0, 1.0)),
                                      ((5, 2, 3), 3, (1.0, 1.0, 4.0))])
def test_non_unit_domain_bit_exact(gpu, ne, p, dom):
    """Box domains of other lengths than 1 (elements of unequal widths per
    axis: Escale and Fscale differ per axis): every stage-kernel path gives
    the oracle's bits after a few steps."""
    nstep, vel = 3, (0.7, -0.4, 0.9)
    x1D, D, L = dgk.read_operator(p)
    m = dgk.mesh_setup(*ne, p, dom=dom, x1D=x1D)
    q = CS.bump_field(m["pos_en"], 321, noise=0.01)
    dt = ic.cfl_dt(p, vel, ne, dom=dom)
    uvw = [np.full_like(q, c) for c in vel]
    om = O.Mesh(*ne, p, dom=dom)
    F = []
    for a in (q, *uvw):
        f = om.field()
        f[:, :om.Ne] = a
        om.update_halo(f)
        F.append(f)
    om.run(*F, dt, nstep, O.OPT1)
    want = F[0][:, :om.Ne]
    solvers = [("structured", dgk.Solver.structured(*ne, p, D, L, dom=dom)),
               ("tables", dgk.Solver.from_tables(p + 1, m["Ne"], m["NeA"], D, L, m["VMapM"],
                                                 m["VMapP"], m["normal_fn"], m["Escale"],
                                                 m["Fscale"], m["halo_src_map"],
                                                 flags=dgk.FLAG_FORCE_TABLES)),
               ("structured-node", dgk.Solver.structured(*ne, p, D, L, dom=dom,
                                                         flags=dgk.FLAG_NODE_KERNEL))]
    for name, s in solvers:
        say(f"[dom] {ne} p={p} dom={dom} {name}")
        s.set_fields(q, *uvw)
        s.run(dt, nstep)
        assert np.array_equal(s.get_field(), want), name
        s.close()
