"""Pins of the oracle's KKT setup, Schur system, preconditioner, PCG and dz.

The Schur/PCG/dz chain reduces to a library routine: the dense equality-constrained QP
    [G  C^T] [dz]   [-g]
    [C   0 ] [ l] = [ c]
with G = blkdiag(Q~_0, R_0, ..., Q~_{N-1}), rows C_0 = [-I 0 ...], C_{k+1} = [.. A_k B_k -I ..]
solved by numpy.linalg.solve; S = -C G^-1 C^T, gamma = c + C G^-1 g.
"""
import numpy as np
import pytest

from conftest import rel_err
from gato_amd import workloads as W


def cfg_for(O, plant, N, **kw):
    base = dict(plant=plant, N=N, dt=0.01, q_cost=100.0, qd_cost=1e-3, u_cost=1e-3, N_cost=100.0,
                q_lim_cost=0.0, vel_lim_cost=0.0, ctrl_lim_cost=0.0)
    base.update(kw)
    return O.make_cfg(**base)


def perturbed(plant, N, B=1, seed=0, scale=0.05):
    wl = W.drop_in_figure8(plant, N=N, B=B, seed=seed)
    rng = np.random.default_rng(seed + 100)
    xu = wl.xu + rng.normal(0, scale, wl.xu.shape)
    return wl, xu


def dense_kkt(kkt, rho, nq):
    N, nx, nu = kkt["q"].shape[0], kkt["q"].shape[1], kkt["r"].shape[1]
    nz = (nx + nu) * N - nu
    off = lambda k: k * (nx + nu)
    G = np.zeros((nz, nz))
    g = np.zeros(nz)
    for k in range(N):
        Qt = kkt["Q"][k].copy()
        Qt[np.arange(nq), np.arange(nq)] += rho
        G[off(k):off(k) + nx, off(k):off(k) + nx] = Qt
        g[off(k):off(k) + nx] = kkt["q"][k]
        if k < N - 1:
            G[off(k) + nx:off(k) + nx + nu, off(k) + nx:off(k) + nx + nu] = kkt["R"][k]
            g[off(k) + nx:off(k) + nx + nu] = kkt["r"][k]
    C = np.zeros((N * nx, nz))
    C[:nx, :nx] = -np.eye(nx)
    for k in range(N - 1):
        r0 = (k + 1) * nx
        C[r0:r0 + nx, off(k):off(k) + nx] = kkt["A"][k]
        C[r0:r0 + nx, off(k) + nx:off(k) + nx + nu] = kkt["B"][k]
        C[r0:r0 + nx, off(k + 1):off(k + 1) + nx] = -np.eye(nx)
    return G, g, C, kkt["c"].reshape(-1)


def step(O, plant, x, u, dt, fext=None):
    n = O.nq(plant)
    a = O.fd(plant, x[:n], x[n:], u, fext)
    return np.concatenate([x[:n] + dt * x[n:] + 0.5 * dt * dt * a, x[n:] + dt * a])


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_integrator_jacobians_and_defects(O, plant):
    N = 6
    cfg = cfg_for(O, plant, N)
    wl, xu = perturbed(plant, N)
    nq, nx, nu = O.dims(cfg)
    fext = np.array([0.1, -0.2, 0.3, 1.0, -2.0, 0.5])
    kkt = O.setup_kkt(cfg, xu[0], wl.xs[0], wl.ref[0], fext=fext)
    st = nx + nu
    for k in range(N - 1):
        x, u = xu[0, k * st:k * st + nx], xu[0, k * st + nx:k * st + st]
        xn = xu[0, (k + 1) * st:(k + 1) * st + nx]
        np.testing.assert_allclose(kkt["c"][k + 1], xn - step(O, plant, x, u, cfg.dt, fext), atol=1e-12)
        h = 1e-6
        A = np.stack([(step(O, plant, x + h * e, u, cfg.dt, fext) - step(O, plant, x - h * e, u, cfg.dt, fext)) / (2 * h)
                      for e in np.eye(nx)], 1)
        Bn = np.stack([(step(O, plant, x, u + h * e, cfg.dt, fext) - step(O, plant, x, u - h * e, cfg.dt, fext)) / (2 * h)
                       for e in np.eye(nu)], 1)
        np.testing.assert_allclose(kkt["A"][k], A, atol=1e-8)
        np.testing.assert_allclose(kkt["B"][k], Bn, atol=1e-9)
    np.testing.assert_allclose(kkt["c"][0], xu[0, :nx] - wl.xs[0], atol=1e-15)


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_cost_gradient_is_merit_gradient(O, plant):
    # q_k, r_k are the gradient of the stage cost that the merit evaluates (mu = 0), barriers active.
    N = 5
    cfg = cfg_for(O, plant, N, q_lim_cost=0.01, vel_lim_cost=0.02, ctrl_lim_cost=0.003, qd_cost=0.05, u_cost=0.01)
    wl, xu = perturbed(plant, N, seed=3, scale=0.02)
    nq, nx, nu = O.dims(cfg)
    kkt = O.setup_kkt(cfg, xu[0], wl.xs[0], wl.ref[0])
    st = nx + nu

    def phi(z):
        return O.merit(cfg, z, np.zeros_like(z), 0.0, wl.xs[0], wl.ref[0], 0.0)

    h = 1e-6
    for k in range(N - 1):
        for i in range(st):
            e = np.zeros_like(xu[0])
            e[k * st + i] = h
            d = (phi(xu[0] + e) - phi(xu[0] - e)) / (2 * h)
            ana = kkt["q"][k][i] if i < nx else kkt["r"][k][i - nx]
            assert abs(d - ana) < 1e-6 * max(1.0, abs(ana)), (k, i, d, ana)


def test_cost_hessian_structure_iiwa(O):
    # Quirk Q5 (iiwa14_plant.cuh:393-396): with only the end-effector term, the position block of
    # Q_k is w g g^T = q q^T / w, not J^T J. The diagonal barrier and velocity / control terms equal
    # the second derivatives of the stage cost.
    N = 4
    cfg = cfg_for(O, "iiwa14", N, q_cost=50.0)
    wl, xu = perturbed("iiwa14", N, seed=4)
    kkt = O.setup_kkt(cfg, xu[0], wl.xs[0], wl.ref[0])
    for k in range(N - 1):
        qq = kkt["q"][k][:7]
        np.testing.assert_allclose(kkt["Q"][k][:7, :7], np.outer(qq, qq) / 50.0, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(np.diag(kkt["Q"][k][7:, 7:]), 1e-3, rtol=1e-12)
        assert np.abs(kkt["Q"][k][:7, 7:]).max() == 0.0
    # Gauss-Newton option: J^T J
    cfg_gn = cfg_for(O, "iiwa14", N, q_cost=50.0, hess_mode=1)
    kg = O.setup_kkt(cfg_gn, xu[0], wl.xs[0], wl.ref[0])
    st = 21
    for k in range(N - 1):
        _, J = O.fk("iiwa14", xu[0, k * st:k * st + 7])
        np.testing.assert_allclose(kg["Q"][k][:7, :7], 50.0 * J.T @ J, rtol=1e-10, atol=1e-13)
    # barrier curvature = second derivative of the merit stage cost
    cfg_b = cfg_for(O, "iiwa14", N, q_cost=0.0, N_cost=0.0, q_lim_cost=0.01, vel_lim_cost=0.02, ctrl_lim_cost=0.003)
    kb = O.setup_kkt(cfg_b, xu[0], wl.xs[0], wl.ref[0])

    def phi(z):
        return O.merit(cfg_b, z, np.zeros_like(z), 0.0, wl.xs[0], wl.ref[0], 0.0)

    h = 1e-4
    for k in range(N - 1):
        for i in range(21):
            e = np.zeros_like(xu[0])
            e[k * st + i] = h
            d2 = (phi(xu[0] + e) - 2 * phi(xu[0]) + phi(xu[0] - e)) / h ** 2
            ana = kb["Q"][k][i, i] if i < 14 else kb["R"][k][i - 14, i - 14]
            assert abs(d2 - ana) < 1e-4 * max(1.0, abs(ana)), (k, i, d2, ana)


def test_terminal_block_quirk(O):
    # Quirk Q4 (iiwa14_plant.cuh:427-450): Q_{N-1}, q_{N-1} come from x_{N-2} with ref_{N-1} and
    # q_cost, so they do not move with x_{N-1}; terminal_mode=1 evaluates the merit's terminal cost.
    N = 5
    cfg = cfg_for(O, "iiwa14", N)
    wl, xu = perturbed("iiwa14", N, seed=5)
    z = xu[0].copy()
    k0 = O.setup_kkt(cfg, z, wl.xs[0], wl.ref[0])
    z2 = z.copy()
    z2[(N - 1) * 21:] += 0.1
    k1 = O.setup_kkt(cfg, z2, wl.xs[0], wl.ref[0])
    np.testing.assert_array_equal(k0["q"][N - 1], k1["q"][N - 1])
    z3 = z.copy()
    z3[(N - 2) * 21:(N - 2) * 21 + 14] += 0.1
    k3 = O.setup_kkt(cfg, z3, wl.xs[0], wl.ref[0])
    assert np.abs(k3["q"][N - 1] - k0["q"][N - 1]).max() > 1e-3
    cfg_t = cfg_for(O, "iiwa14", N, terminal_mode=1)
    kt = O.setup_kkt(cfg_t, z, wl.xs[0], wl.ref[0])

    def phi(zz):
        return O.merit(cfg_t, zz, np.zeros_like(zz), 0.0, wl.xs[0], wl.ref[0], 0.0)

    h = 1e-6
    for i in range(14):
        e = np.zeros_like(z)
        e[(N - 1) * 21 + i] = h
        d = (phi(z + e) - phi(z - e)) / (2 * h)
        assert abs(d - kt["q"][N - 1][i]) < 1e-6 * max(1.0, abs(d))


@pytest.mark.parametrize("plant,rho", [("iiwa14", 1e-3), ("iiwa14", 0.1), ("indy7", 1e-2)])
def test_schur_and_preconditioner_match_dense_kkt(O, plant, rho):
    N = 6
    cfg = cfg_for(O, plant, N)
    wl, xu = perturbed(plant, N, seed=7)
    nq, nx, nu = O.dims(cfg)
    kkt = O.setup_kkt(cfg, xu[0], wl.xs[0], wl.ref[0])
    sch = O.schur(cfg, rho, kkt)
    G, g, C, c = dense_kkt(kkt, rho, nq)
    S = -C @ np.linalg.solve(G, C.T)
    for k in range(N):
        blk = S[k * nx:(k + 1) * nx, k * nx:(k + 1) * nx]
        assert rel_err(sch["Sd"][k], blk) < 1e-10
        if k:
            assert rel_err(sch["So"][k], S[k * nx:(k + 1) * nx, (k - 1) * nx:k * nx]) < 1e-10
    assert rel_err(sch["gamma"].reshape(-1), c + C @ np.linalg.solve(G, g)) < 1e-10
    # preconditioner (schur_linsys.cuh:150-164, 214-260): D-block inverses and -D^-1 O D^-1
    E = np.diag(np.r_[np.ones(nq), np.zeros(nq)])
    for k in range(N):
        D = sch["Sd"][k] if k else None
        if k == 0:
            expect = -np.linalg.inv(-sch["Sd"][0])  # P_00 = -Q~_0, S_00 = -Q~_0^-1
        else:
            expect = -np.linalg.inv(-D + rho * E)  # quirk Q8: theta + rho diag(1_nq, 0)
        assert rel_err(sch["Pd"][k], expect) < 1e-9
        if k:
            assert rel_err(sch["Po"][k], -sch["Pd"][k] @ sch["So"][k] @ sch["Pd"][k - 1]) < 1e-10


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_pcg_and_dz_solve_the_dense_kkt(O, plant):
    N = 6
    rho = 0.05
    cfg = cfg_for(O, plant, N)
    wl, xu = perturbed(plant, N, seed=8)
    nq, nx, nu = O.dims(cfg)
    kkt = O.setup_kkt(cfg, xu[0], wl.xs[0], wl.ref[0])
    sch = O.schur(cfg, rho, kkt)
    G, g, C, c = dense_kkt(kkt, rho, nq)
    K = np.block([[G, C.T], [C, np.zeros((C.shape[0], C.shape[0]))]])
    sol = np.linalg.solve(K, np.concatenate([-g, c]))
    nz = G.shape[0]
    lam, it = O.pcg(nx, N, sch, np.zeros(N * nx), 1e-30, 5000)
    assert 0 < it <= 3 * N * nx
    # at exit the reference's stopping rule |r^T P r| < 1e-6 + eps |rho_0| holds (pcg.cuh:134)
    S = -C @ np.linalg.solve(G, C.T)
    P = np.zeros_like(S)
    for k in range(N):
        P[k * nx:(k + 1) * nx, k * nx:(k + 1) * nx] = sch["Pd"][k]
        if k:
            P[k * nx:(k + 1) * nx, (k - 1) * nx:k * nx] = sch["Po"][k]
            P[(k - 1) * nx:k * nx, k * nx:(k + 1) * nx] = sch["Po"][k].T
    r = sch["gamma"].reshape(-1) - S @ lam
    assert abs(r @ P @ r) < 1e-6
    # one iteration fewer does not meet it
    lam_m, _ = O.pcg(nx, N, sch, np.zeros(N * nx), 1e-30, it - 1)
    rm = sch["gamma"].reshape(-1) - S @ lam_m
    assert abs(rm @ P @ rm) >= 1e-6 * 0.999
    d, qres, rres = O.dz(cfg, sch, kkt, sol[nz:])
    assert rel_err(d, sol[:nz]) < 1e-8
    # stationarity residual = -G~ dz (schur_linsys.cuh:406-430)
    nu_ = nu
    for k in range(N):
        Qt = kkt["Q"][k].copy()
        Qt[np.arange(nq), np.arange(nq)] += rho
        np.testing.assert_allclose(qres[k], -Qt @ d[k * (nx + nu_):k * (nx + nu_) + nx], rtol=1e-7, atol=1e-9)


def test_pcg_warm_start_and_early_exit(O):
    N = 5
    cfg = cfg_for(O, "iiwa14", N)
    wl, xu = perturbed("iiwa14", N, seed=9)
    kkt = O.setup_kkt(cfg, xu[0], wl.xs[0], wl.ref[0])
    sch = O.schur(cfg, 0.05, kkt)
    lam, it = O.pcg(14, N, sch, np.zeros(N * 14), 1e-12, 500)
    assert it > 0
    # warm start at the solution: |r^T z| < 1e-6 exits with zero iterations and lambda untouched (pcg.cuh:111-116)
    lam2, it2 = O.pcg(14, N, sch, lam, 1e-12, 500)
    assert it2 == 0
    np.testing.assert_array_equal(lam2, lam)
    # max_pcg_iters bounds the loop
    _, it3 = O.pcg(14, N, sch, np.zeros(N * 14), 0.0, 3)
    assert it3 == 3
