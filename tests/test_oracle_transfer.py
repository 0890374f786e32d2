"""Pins the oracle's transfer and constitutive steps with closed forms and invariants that do not
restate its code: B-spline moment identities for P2G and G2P, hand-built grid nodes for the grid
update, textbook stress laws and the Drucker-Prager cone for the materials, the Lame relations for
the material table, and the reference's own whole-substep tests (physics.test.js,
lifecycle.test.js, verify/index.js). CPU only."""
import numpy as np
import pytest

import oracle_mpm as O

FP = float(1 << 22)
DIMS = [16, 16, 16]
WATER = {"kind": 0, "density": 1.0, "stiffness": 0.5, "viscosity": 0.1}
JELLY = {"kind": 1, "density": 1.0, "stiffness": 0.35, "poisson": 0.3}
SNOW = {"kind": 2, "density": 0.4, "stiffness": 0.4, "poisson": 0.2, "hardening": 10, "compression": 0.025,
        "stretch": 0.0075}
SAND = {"kind": 3, "density": 1.6, "stiffness": 0.6, "poisson": 0.3, "friction": 30}


def _scene(mats, n=1, dims=DIMS):
    sc = O.Scene(dims, max(n, 1), 2, 4)
    tbl = np.zeros((8, 12), np.float32)
    for k, m in enumerate(mats):
        tbl[k] = O.material_row(m, 2)
    sc.set_materials(tbl)
    return sc, tbl


def _record(x, v=(0, 0, 0), C=None, tau=None, F=None, J=1.0, mat=0):
    r = np.zeros(32, np.float32)
    r[0:3] = x
    r[3] = np.array([mat], np.uint32).view(np.float32)[0]
    r[4:7] = v
    r[7] = J
    C = np.zeros((3, 3)) if C is None else np.asarray(C)
    F = np.eye(3) if F is None else np.asarray(F)
    tau = np.zeros((3, 3)) if tau is None else np.asarray(tau)
    for j in range(3):
        r[8 + 4 * j:11 + 4 * j] = C[:, j]
        r[20 + 4 * j:23 + 4 * j] = F[:, j]
    r[11], r[15], r[19] = tau[0, 0], tau[1, 1], tau[2, 2]
    r[23], r[27], r[31] = tau[0, 1], tau[0, 2], tau[1, 2]
    return r


def _C(rec):
    return np.stack([rec[8:11], rec[12:15], rec[16:19]], axis=1).astype(np.float64)


def _F(rec):
    return np.stack([rec[20:23], rec[24:27], rec[28:31]], axis=1).astype(np.float64)


def _tau(rec):
    t = np.zeros((3, 3))
    t[0, 0], t[1, 1], t[2, 2] = rec[11], rec[15], rec[19]
    t[0, 1] = t[1, 0] = rec[23]
    t[0, 2] = t[2, 0] = rec[27]
    t[1, 2] = t[2, 1] = rec[31]
    return t


def _nodes(dims=DIMS):
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float64)  # node = x + X*(y + Y*z)


# ---- P2G (mpm.js:306-401) -------------------------------------------------------------------------

@pytest.mark.parametrize("x", [(7.3, 8.6, 6.2), (5.5, 5.5, 5.5), (9.02, 4.97, 8.51)])
def test_p2g_single_particle_moments(x):
    """One particle: grid mass = m, momentum = m v, first moment sum_i q_i (x_i - x_p)^T = A / 4 with
    A = -vol4 tau + m C (quadratic B-spline: sum w = 1, sum w d = 0, sum w d d^T = I/4). A
    non-symmetric C and a full tau catch a transposed or sign-flipped affine term."""
    rng = np.random.default_rng(1)
    C = (rng.random((3, 3)) - 0.5) * 0.4
    tau = (rng.random((3, 3)) - 0.5) * 0.3
    tau = tau + tau.T
    v = np.array([0.11, -0.23, 0.07])
    sc, tbl = _scene([WATER])
    sc.write(_record(x, v, C, tau)[None])
    sc.set_count(1)
    sc.p2g()
    g = sc.grid().astype(np.float64) / FP
    m = float(tbl[0, 1])
    assert m == pytest.approx(1.0 / 8.0)
    nodes = _nodes()
    assert g[:, 3].sum() == pytest.approx(m, abs=30 / FP)
    np.testing.assert_allclose(g[:, :3].sum(0), m * v, atol=30 / FP)
    d = nodes - np.asarray(x, np.float32).astype(np.float64)
    moment = g[:, :3].T @ d  # sum_i q_i d_i^T
    A = -0.5 * tau + m * C   # vol4 = 4 / ppc^3 = 0.5
    np.testing.assert_allclose(4 * moment, A, atol=2e-5)
    base = np.floor(np.asarray(x, np.float32) - 0.5)
    nz = nodes[g[:, 3] != 0]
    assert len(nz) >= 8 and np.all(nz >= base) and np.all(nz <= base + 2)  # the 3^3 stencil only


def test_p2g_order_free_and_additive():
    """Fixed-point sums: P2G of a set equals the sum of P2Gs of its parts, in any order."""
    rng = np.random.default_rng(2)
    n = 300
    recs = np.stack([_record(rng.uniform(4, 11, 3), rng.uniform(-0.3, 0.3, 3), (rng.random((3, 3)) - 0.5) * 0.1)
                     for _ in range(n)])
    sc, _ = _scene([WATER], n)
    sc.write(recs)
    sc.set_count(n)
    sc.p2g()
    whole = sc.grid().astype(np.int64)
    parts = np.zeros_like(whole)
    for sl in (slice(0, 100), slice(100, 300)):
        s2, _ = _scene([WATER], n)
        s2.write(recs[sl][::-1].copy())
        s2.set_count(sl.stop - sl.start)
        s2.p2g()
        parts += s2.grid().astype(np.int64)
    assert np.array_equal(whole, parts)


# ---- grid update (mpm.js:414-471) -----------------------------------------------------------------

def _grid_with(node_vals, dims=DIMS):
    """node_vals: {(x, y, z): (momentum xyz, mass)} in float units -> i32 words."""
    w = np.zeros((dims[0] * dims[1] * dims[2], 4), np.int32)
    for (x, y, z), (p, m) in node_vals.items():
        i = x + dims[0] * (y + dims[1] * z)
        w[i, :3] = np.rint(np.asarray(p) * FP)
        w[i, 3] = np.rint(m * FP)
    return w


def _v(sc, x, y, z, dims=DIMS):
    return sc.grid()[x + dims[0] * (y + dims[1] * z)].view(np.float32).astype(np.float64)


def test_grid_update_interior_walls_floor():
    g = np.array([0.0, -0.0027, 0.001], np.float32)
    vals = {
        (8, 8, 8): ((0.05, 0.02, -0.01), 0.5),      # interior: v = p/m + g
        (1, 8, 8): ((-0.05, 0.02, 0.01), 0.5),      # x wall, moving out: v.x -> 0
        (14, 8, 8): ((0.05, 0.02, 0.01), 0.5),      # far x wall (x > X - bound), moving out
        (8, 8, 1): ((0.01, 0.02, -0.03), 0.5),      # z wall
        (8, 1, 8): ((0.06, -0.04, 0.08), 0.5),      # floor with Coulomb friction
        (8, 2, 8): ((0.001, -0.3, 0.0), 0.5),       # floor, friction stops it
        (5, 5, 5): ((0.0, 0.0, 0.0), 0.0),          # empty node
    }
    sc, _ = _scene([WATER])
    sc.write_grid(_grid_with(vals))
    sc.grid_update(g, 0.4)
    gg = g.astype(np.float64)
    np.testing.assert_allclose(_v(sc, 8, 8, 8)[:3], np.array([0.1, 0.04, -0.02]) + gg, rtol=1e-6, atol=1e-7)
    assert _v(sc, 8, 8, 8)[3] == pytest.approx(0.5)
    v = _v(sc, 1, 8, 8)
    assert v[0] == 0 and v[1] == pytest.approx(0.04 + gg[1], abs=1e-7)
    assert _v(sc, 14, 8, 8)[0] == 0
    assert _v(sc, 8, 8, 1)[2] == 0
    # floor: vn = -(v.y); tangential scaled by max(0, 1 - mu vn / |vt|)
    pre = np.array([0.12, -0.08, 0.16]) + gg
    vn, vt = -pre[1], np.hypot(pre[0], pre[2])
    f = max(0.0, 1 - 0.4 * vn / vt)
    np.testing.assert_allclose(_v(sc, 8, 1, 8)[:3], [pre[0] * f, 0, pre[2] * f], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(_v(sc, 8, 2, 8)[:3], [0, 0, 0], atol=1e-7)
    assert np.all(_v(sc, 5, 5, 5) == 0)


def test_grid_update_sphere_collider_removes_inflow():
    """A sphere collider (mpm.js:419-466): inflow along the normal is removed, tangential motion is
    reduced by Coulomb friction, outflow is untouched."""
    col = np.zeros(12, np.float32)
    col[0] = np.array([0], np.uint32).view(np.float32)[0]
    col[1:4] = (8, 8, 8)
    col[4] = 3.0     # radius
    col[7] = 0.2     # friction
    sc, _ = _scene([WATER])
    sc.set_colliders(col[None], 1)
    vals = {(11, 8, 8): ((-0.1, 0.05, 0.0), 1.0),    # on the surface, moving inward
            (8, 11, 8): ((0.0, 0.1, 0.0), 1.0)}      # on the surface, moving outward
    sc.write_grid(_grid_with(vals))
    sc.grid_update(np.zeros(3, np.float32), 0.4)
    v = _v(sc, 11, 8, 8)[:3]
    assert v[0] == pytest.approx(0.0, abs=1e-6)            # normal (x) inflow removed
    assert 0 < v[1] < 0.05                                  # tangential reduced by friction
    np.testing.assert_allclose(_v(sc, 8, 11, 8)[:3], [0, 0.1, 0], atol=1e-6)


# ---- G2P (mpm.js:474-560) -------------------------------------------------------------------------

def _linear_field(u, M, x0, dims=DIMS):
    nodes = _nodes(dims)
    V = u + (nodes - x0) @ M.T
    w = np.zeros((len(nodes), 4), np.float32)
    w[:, :3] = V
    w[:, 3] = 1.0
    return w.view(np.int32)


@pytest.mark.parametrize("mat", [WATER, JELLY, SNOW, SAND])
def test_g2p_linear_field_gives_exact_velocity_and_C(mat):
    """A linear grid velocity V(x) = u + M (x - x0) is reproduced exactly by the quadratic B-spline:
    v = V(x_p) and C = 4 sum w V d^T = M (a non-symmetric M pins the outer-product orientation)."""
    rng = np.random.default_rng(5)
    M = (rng.random((3, 3)) - 0.5) * 0.02
    u = np.array([0.05, -0.03, 0.02])
    x0 = np.array([8.0, 8.0, 8.0])
    xp = np.array([7.37, 8.81, 8.12], np.float32)
    sc, _ = _scene([mat])
    sc.write(_record(xp)[None])
    sc.set_count(1)
    sc.write_grid(_linear_field(u, M, x0))
    sc.g2p(0.9)
    r = sc.read(1)[0]
    vexp = u + M @ (xp.astype(np.float64) - x0)
    np.testing.assert_allclose(r[4:7], vexp, atol=2e-7)
    np.testing.assert_allclose(_C(r), M, atol=2e-6)
    np.testing.assert_allclose(r[0:3], xp + vexp, atol=2e-6)


def _g2p_one(mat, F=np.eye(3), J=1.0, M=None):
    M = np.zeros((3, 3)) if M is None else M
    sc, tbl = _scene([mat])
    xp = np.array([8.2, 7.6, 8.4], np.float32)
    sc.write(_record(xp, F=F, J=J)[None])
    sc.set_count(1)
    sc.write_grid(_linear_field(np.zeros(3), M, np.array([8.0, 8.0, 8.0])))
    sc.g2p(0.9)
    return sc.read(1)[0], tbl[0].astype(np.float64)


def test_g2p_fluid_J_and_stress():
    """Fluid (mpm.js:517-522): J <- clamp(J (1 + tr C)), tau = K min(J - 1, 0.02) I + visc J (C + C^T), F = I."""
    M = np.array([[0.01, 0.004, 0.0], [-0.002, 0.005, 0.003], [0.001, 0.0, -0.02]])
    r, row = _g2p_one(WATER, J=0.97, M=M)
    J = np.clip(0.97 * (1 + np.trace(M)), 0.6, 1.4)
    assert r[7] == pytest.approx(J, rel=1e-5)
    tau = row[2] * min(J - 1, 0.02) * np.eye(3) + row[4] * J * (M + M.T)
    np.testing.assert_allclose(_tau(r), tau, atol=1e-6)
    np.testing.assert_allclose(_F(r), np.eye(3), atol=0)


def test_g2p_jelly_neo_hookean():
    """Jelly (mpm.js:524-528): F <- (I + C) F, tau = mu (F F^T - I) + lambda ln J I (Kirchhoff stress of
    compressible Neo-Hookean)."""
    rng = np.random.default_rng(8)
    F0 = np.eye(3) + (rng.random((3, 3)) - 0.5) * 0.3
    M = (rng.random((3, 3)) - 0.5) * 0.04
    r, row = _g2p_one(JELLY, F=F0, M=M)
    F = (np.eye(3) + M) @ F0
    np.testing.assert_allclose(_F(r), F, atol=3e-6)
    mu, lam = row[2], row[3]
    tau = mu * (F @ F.T - np.eye(3)) + lam * np.log(np.linalg.det(F)) * np.eye(3)
    np.testing.assert_allclose(_tau(r), tau, rtol=1e-4, atol=1e-5 * np.abs(tau).max())


def test_g2p_snow_clamps_singular_values():
    """Snow (mpm.js:530-542): singular values of the trial F clamped to [1 - theta_c, 1 + theta_s],
    Jp <- clamp(Jp Jold / Jnew, .6, 20), hardening e^{xi (1 - Jp)}, tau = 2 mu (F - R) F^T + lambda J (J - 1) I."""
    rng = np.random.default_rng(9)
    Q1, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Q2, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    F0 = Q1 @ np.diag([1.06, 0.99, 0.95]) @ Q2.T   # outside the elastic band on two axes
    r, row = _g2p_one(SNOW, F=F0, J=1.0)
    U, s, Vt = np.linalg.svd(F0)
    sc = np.clip(s, 1 - row[6], 1 + row[7])
    Fe = U @ np.diag(sc) @ Vt
    Jp = np.clip(1.0 * np.prod(s) / np.prod(sc), 0.6, 20)
    np.testing.assert_allclose(_F(r), Fe, atol=2e-5)
    assert r[7] == pytest.approx(Jp, rel=1e-4)
    np.testing.assert_allclose(np.linalg.svd(_F(r), compute_uv=False), sc, atol=2e-5)
    h = np.exp(row[5] * (1 - Jp))
    mu, lam, J = row[2] * h, row[3] * h, np.prod(sc)
    R = U @ Vt
    tau = 2 * mu * (Fe - R) @ Fe.T + lam * J * (J - 1) * np.eye(3)
    np.testing.assert_allclose(_tau(r), tau, atol=5e-5 * max(1.0, np.abs(tau).max()))


def test_g2p_sand_drucker_prager_cone():
    """Sand (mpm.js:544-555), in log-strain space: a compressive strain outside the Drucker-Prager
    cone is projected onto it (|dev eps| = -(3 lambda + 2 mu) / (2 mu) tr(eps) alpha, tr(eps) kept);
    one inside the cone is unchanged; extension (tr eps >= 0) returns F to a rotation."""
    rng = np.random.default_rng(10)
    Q1, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Q2, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    Q1 *= np.sign(np.linalg.det(Q1))
    Q2 *= np.sign(np.linalg.det(Q2))
    sv = lambda F: np.sort(np.linalg.svd(F, compute_uv=False))[::-1]  # noqa: E731
    # outside the cone: projected
    F0 = Q1 @ np.diag([1.10, 0.98, 0.85]) @ Q2.T
    r, row = _g2p_one(SAND, F=F0)
    mu, lam, alpha = row[2], row[3], row[8]
    eps0, eps = np.log(sv(F0)), np.log(sv(_F(r)))
    radius = -(3 * lam + 2 * mu) / (2 * mu) * eps0.sum() * alpha
    assert np.linalg.norm(eps0 - eps0.mean()) > radius + 0.05
    assert eps.sum() == pytest.approx(eps0.sum(), abs=2e-5)
    assert np.linalg.norm(eps - eps.mean()) == pytest.approx(radius, abs=5e-5)
    # projection is along the deviatoric direction
    d0, d1 = eps0 - eps0.mean(), eps - eps.mean()
    assert np.dot(d0, d1) / (np.linalg.norm(d0) * np.linalg.norm(d1)) == pytest.approx(1.0, abs=1e-6)
    # inside the cone: elastic, F unchanged
    F2 = Q1 @ np.diag([0.98, 0.93, 0.9]) @ Q2.T
    r2, _ = _g2p_one(SAND, F=F2)
    np.testing.assert_allclose(_F(r2), F2, atol=3e-6)
    # extension: back to a rotation
    F1 = Q1 @ np.diag([1.05, 1.02, 0.99]) @ Q2.T
    r1, _ = _g2p_one(SAND, F=F1)
    np.testing.assert_allclose(sv(_F(r1)), 1.0, atol=3e-6)
    # Kirchhoff stress of the Hencky model: tau = U diag(2 mu eps + lambda tr eps) U^T
    U, s, Vt = np.linalg.svd(_F(r))
    e = np.log(s)
    tau = U @ np.diag(2 * mu * e + lam * e.sum()) @ U.T
    np.testing.assert_allclose(_tau(r), tau, atol=5e-5 * max(1.0, np.abs(tau).max()))


def test_g2p_speed_clamp_and_position_bounds():
    """|v| > vmax is rescaled to vmax; x is clamped to [bound - 1, dims - bound + 1] (mpm.js:497-503)."""
    sc, _ = _scene([WATER])
    xp = np.array([12.9, 8.0, 8.0], np.float32)
    sc.write(_record(xp)[None])
    sc.set_count(1)
    sc.write_grid(_linear_field(np.array([2.0, 0.0, 0.0]), np.zeros((3, 3)), np.zeros(3)))
    sc.g2p(0.9)
    r = sc.read(1)[0]
    assert r[4] == pytest.approx(0.9, rel=1e-6)
    assert r[0] == pytest.approx(12.9 + 0.9, rel=1e-6)
    sc.write(_record(np.array([13.6, 8.0, 8.0], np.float32))[None])
    sc.g2p(0.9)
    assert sc.read(1)[0][0] == 14.0  # 13.6 + 0.9 clamped to dims - bound + 1 = 14
    sc.write_grid(_linear_field(np.array([-2.0, 0.0, 0.0]), np.zeros((3, 3)), np.zeros(3)))
    sc.write(_record(np.array([2.5, 8.0, 8.0], np.float32))[None])
    sc.g2p(0.9)
    assert sc.read(1)[0][0] == 2.0   # 2.5 - 0.9 clamped to bound - 1 = 2
