"""The oracle against what the mathematics fixes, independently of the
reference's code: LGL nodes and weights, exact polynomial differentiation,
the exact-mass lift, the upwind flux for constant velocity, discrete mass
conservation and the SSP-RK3 amplification polynomial.  Each check is chosen
so that a plausible slip (wrong axis, transposed operand, wrong sign, wrong
face, wrong RK coefficient) fails it.
"""
import numpy as np
import numpy.polynomial.legendre as leg
import pytest

import oracle as O


def lgl_nodes_weights(p):
    """Legendre-Gauss-Lobatto nodes (+-1 and the roots of P_p') and weights
    w_i = 2 / (p (p+1) P_p(x_i)^2)."""
    inner = leg.Legendre.basis(p).deriv().roots() if p > 1 else np.array([])
    x = np.concatenate([[-1.0], np.sort(inner.real), [1.0]])
    Pp = leg.Legendre.basis(p)(x)
    return x, 2.0 / (p * (p + 1) * Pp * Pp)


def exact_mass(x):
    """M_ab = int_{-1}^{1} l_a l_b dx on the Lagrange basis at nodes x, by
    Gauss-Legendre quadrature exact for degree 2p."""
    n = len(x)
    g, gw = leg.leggauss(n + 1)
    # barycentric form of l_a(t) (stable for every order used here)
    lam = np.array([1.0 / np.prod([x[a] - x[b] for b in range(n) if b != a]) for a in range(n)])
    T = lam[None, :] / (g[:, None] - x[None, :])
    L = T / T.sum(axis=1, keepdims=True)  # (quad pt, a)
    return (L * gw[:, None]).T @ L


@pytest.mark.parametrize("p", range(1, 16))
def test_lgl_nodes(p):
    x1D, _, _ = O.read_operator(p)
    x, _ = lgl_nodes_weights(p)
    assert np.max(np.abs(x1D - x)) < 1e-14


@pytest.mark.parametrize("p", range(1, 16))
def test_derivative_of_monomials(p):
    """D1D(i,l) = l_l'(x_i): D x^m = m x^(m-1) exactly for m <= p."""
    x1D, D, _ = O.read_operator(p)
    for m in range(p + 1):
        want = m * x1D ** (m - 1) if m else np.zeros_like(x1D)
        assert np.max(np.abs(D @ x1D ** m - want)) < 2e-12 * max(1, m), m


@pytest.mark.parametrize("p", range(1, 16))
def test_lift_is_exact_mass_inverse(p):
    """Lift(:,:,:,f) = (M^-1 e_-+)(normal index) x 1 x 1 with the exact (not
    lumped) 1D mass matrix; faces 1 -y, 2 +x, 3 +y, 4 -x, 5 -z, 6 +z."""
    x1D, _, L = O.read_operator(p)
    Nq = p + 1
    Mi = np.linalg.inv(exact_mass(x1D))
    lm, lp = Mi[:, 0], Mi[:, -1]
    one = np.ones(Nq)
    want = np.zeros_like(L)
    want[..., 0] = np.einsum("i,j,k->ijk", one, lm, one)
    want[..., 1] = np.einsum("i,j,k->ijk", lp, one, one)
    want[..., 2] = np.einsum("i,j,k->ijk", one, lp, one)
    want[..., 3] = np.einsum("i,j,k->ijk", lm, one, one)
    want[..., 4] = np.einsum("i,j,k->ijk", one, one, lm)
    want[..., 5] = np.einsum("i,j,k->ijk", one, one, lp)
    assert np.max(np.abs(L - want)) < 2e-13 * np.max(np.abs(want))


@pytest.mark.parametrize("p", [1, 3, 7, 8, 11, 15])
@pytest.mark.parametrize("tname", ["GENERAL", "OPT1"])
def test_divlike_differentiates_each_axis(p, tname):
    """out_x/out_y/out_z of divlike are d/dx, d/dy, d/dz of the element's
    tensor polynomial x^a y^b z^c (a,b,c <= p): pins the axis of every sweep."""
    x1D, D, _ = O.read_operator(p)
    Nq = p + 1
    X, Y, Z = np.meshgrid(x1D, x1D, x1D, indexing="ij")
    rng = np.random.default_rng(p)
    for _ in range(3):
        a, b, c = (int(v) for v in rng.integers(0, p + 1, 3))
        f = np.asfortranarray(X ** a * Y ** b * Z ** c).ravel(order="F")
        ox, oy, oz = O.divlike(Nq, O.TYPEIDS[tname], D, np.asfortranarray(D.T), f, f, f)
        for out, want in ((ox, a * X ** max(a - 1, 0) * Y ** b * Z ** c),
                          (oy, b * X ** a * Y ** max(b - 1, 0) * Z ** c),
                          (oz, c * X ** a * Y ** b * Z ** max(c - 1, 0))):
            assert np.max(np.abs(out - want.ravel(order="F"))) < 5e-12 * (1 + p * p)


def _fields(m, q, u, v, w):
    F = []
    for a in (q, u, v, w):
        f = m.field()
        f[:, :m.Ne] = a
        m.update_halo(f)
        F.append(f)
    return F


def test_flux_constant_velocity_closed_form():
    """For constant velocity the reference flux reduces to
    Fscale * min(n.u, 0) * (qP - qM) (upwind: only inflow faces carry it)."""
    m = O.Mesh(3, 4, 2, 4)
    rng = np.random.default_rng(3)
    vel = (0.7, -1.3, 0.4)
    q = np.asfortranarray(rng.standard_normal((m.Np, m.Ne)))
    Q, U, V, W = _fields(m, q, *(np.full((m.Np, m.Ne), c, order="F") for c in vel))
    fl = m.elembnd_flux(Q, U, V, W)
    qf = Q.ravel(order="F")
    qM, qP = qf[m.VMapM - 1], qf[m.VMapP - 1]
    vn = sum(m.normal_fn[:, :, d] * vel[d] for d in range(3))
    want = m.Fscale * np.minimum(vn, 0.0) * (qP - qM)
    assert np.max(np.abs(fl - want)) <= 4e-16 * np.max(np.abs(want)) * 8
    # outflow faces carry no flux (up to the rounding of the reference's form)
    assert np.max(np.abs(fl[vn > 0])) <= 1e-14 * np.max(np.abs(want))


@pytest.mark.parametrize("ne,p", [((3, 4, 2), 3), ((2, 3, 2), 7), ((2, 2, 2), 11)])
def test_discrete_mass_conservation(ne, p):
    """sum_e J sum_n w_i w_j w_k dqdt = 0 to round-off (exact mass lift and
    single-valued upwind flux); per-node random velocity."""
    m = O.Mesh(*ne, p)
    rng = np.random.default_rng(p)
    F = _fields(m, *(np.asfortranarray(rng.uniform(-1, 1, (m.Np, m.Ne))) for _ in range(4)))
    d = m.cal_tend(*F)[:, :m.Ne]
    _, w = lgl_nodes_weights(p)
    W = np.einsum("i,j,k->ijk", w, w, w).ravel(order="F")
    rate = np.sum(W[:, None] * d)
    assert abs(rate) < 1e-13 * np.sum(np.abs(W[:, None] * d))


def test_constant_state_has_no_tendency():
    m = O.Mesh(2, 3, 4, 5)
    F = _fields(m, *(np.full((m.Np, m.Ne), c, order="F") for c in (1.5, 0.3, -0.8, 1.1)))
    d = m.cal_tend(*F)[:, :m.Ne]
    assert np.max(np.abs(d)) < 1e-12 * 1.5 * 1.1 * np.max(m.Escale)


def test_rk3_amplification_polynomial():
    """One oracle step of the linear system dq/dt = L q equals
    (I + dt L + dt^2 L^2/2 + dt^3 L^3/6) q: pins the SSP-RK3 coefficients."""
    m = O.Mesh(3, 3, 2, 3)
    rng = np.random.default_rng(7)
    q = np.asfortranarray(rng.uniform(-1, 1, (m.Np, m.Ne)))
    vel = [np.full((m.Np, m.Ne), c, order="F") for c in (0.9, -0.4, 0.6)]
    dt = 4e-3

    def L(x):
        X, U, V, W = _fields(m, x, *vel)
        return np.asfortranarray(m.cal_tend(X, U, V, W)[:, :m.Ne])

    q1 = L(q); q2 = L(q1); q3 = L(q2)
    want = q + dt * q1 + dt ** 2 / 2 * q2 + dt ** 3 / 6 * q3
    Q, U, V, W = _fields(m, q, *vel)
    m.run(Q, U, V, W, dt, 1)
    got = Q[:, :m.Ne]
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(q))
    # a wrong stage coefficient shows at the dt^2 level:
    assert np.max(np.abs(dt ** 2 / 2 * q2)) > 1e-8


def test_halo_slots_copy_their_sources():
    """Every VMapP entry, resolved through the halo map, is the node of the
    neighbour's opposite face that coincides in space modulo the period."""
    m = O.Mesh(3, 2, 4, 2)
    pos = m.pos_en.reshape((m.Np * m.Ne, 3), order="F")
    iM = m.VMapM.ravel(order="F") - 1
    iP = m.VMapP.ravel(order="F") - 1
    owned = m.Np * m.Ne
    halo = iP >= owned
    iP = iP.copy()
    iP[halo] = m.halo_src_map[iP[halo] - owned] - 1
    d = pos[iM] - pos[iP]
    d = d - np.round(d)  # periodic unit cube
    assert np.max(np.abs(d)) < 1e-14
    assert halo.sum() == m.NhaloNode

