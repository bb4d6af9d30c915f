"""Pins of the oracle's SQP loop (bsqp.cuh:103-197, merit.cuh, line_search.cuh) by invariants of
the reference's line search, rho schedule, batch bookkeeping and merit definition."""
import numpy as np
import pytest

from gato_amd import workloads as W


def fig8_cfg(O, N=8, iters=4, **kw):
    base = dict(plant="iiwa14", N=N, dt=0.01, max_sqp_iters=iters, kkt_tol=1e-4, max_pcg_iters=100,
                q_cost=100.0, qd_cost=1e-3, u_cost=1e-3, N_cost=100.0, q_lim_cost=0.0)
    base.update(kw)
    return O.make_cfg(**base)


def test_merit_definition(O):
    # phi = sum_k l_k + mu (sum_k |c_{k+1}|_1 + |x_0 - x_s|_1), recomputed from the dynamics and FK.
    N = 5
    cfg = fig8_cfg(O, N=N, q_lim_cost=0.01, vel_lim_cost=0.01, ctrl_lim_cost=0.001)
    wl = W.drop_in_figure8(N=N, B=1, seed=2)
    rng = np.random.default_rng(0)
    z = wl.xu[0] + rng.normal(0, 0.02, wl.xu.shape[1])
    qlo, qhi, vlo, vhi, ulo, uhi = O.limits("iiwa14")
    bar = lambda v, lo, hi: -np.log(np.maximum(v - lo, 1e-10)) - np.log(np.maximum(hi - v, 1e-10))
    mu = 7.0
    tot = 0.0
    for k in range(N):
        x = z[k * 21:k * 21 + 14]
        p, _ = O.fk("iiwa14", x[:7])
        w = 100.0
        l = 0.5 * w * np.sum((p - wl.ref[0, 6 * k:6 * k + 3]) ** 2) + 0.5 * 1e-3 * np.sum(x[7:] ** 2)
        l += 0.01 * bar(x[:7], qlo, qhi).sum() + 0.01 * bar(x[7:], vlo, vhi).sum()
        if k < N - 1:
            u = z[k * 21 + 14:k * 21 + 21]
            l += 0.5 * 1e-3 * np.sum(u ** 2) + 0.001 * bar(u, ulo, uhi).sum()
            xn = z[(k + 1) * 21:(k + 1) * 21 + 14]
            a = O.fd("iiwa14", x[:7], x[7:], u)
            step = np.concatenate([x[:7] + 0.01 * x[7:] + 0.5e-4 * a, x[7:] + 0.01 * a])
            l += mu * np.abs(xn - step).sum()
        else:
            l += mu * np.abs(z[:14] - wl.xs[0]).sum()
        tot += l
    got = O.merit(cfg, z, np.zeros_like(z), 0.0, wl.xs[0], wl.ref[0], mu)
    assert abs(got - tot) < 1e-9 * abs(tot)
    # alpha moves along dz
    dz = rng.normal(0, 0.01, z.shape)
    assert abs(O.merit(cfg, z, dz, 0.5, wl.xs[0], wl.ref[0], mu) - O.merit(cfg, z + 0.5 * dz, np.zeros_like(z), 0.0,
                                                                          wl.xs[0], wl.ref[0], mu)) < 1e-9 * abs(tot)


def test_line_search_and_rho_schedule(O):
    cfg = fig8_cfg(O, iters=6)
    wl = W.drop_in_figure8(N=8, B=6, seed=3)
    r = O.solve(cfg, wl.xu, wl.xs, wl.ref, mu=10.0, pcg_tol=1e-5, rho0=1e-3)
    steps = r["ls_step"][: r["ls_num_iters"]]
    ok = {-1.0} | {1.0 / (1 << j) for j in range(8)}
    assert set(np.unique(steps)).issubset(ok)
    for b in range(wl.B):
        # accepted steps strictly decrease phi_cur; failures keep it (line_search.cuh:59-97)
        prev = r["initial_merit"][b]
        rho, drho = 1e-3, 1.0
        for i in range(r["ls_num_iters"]):
            m = r["ls_min_merit"][i, b]
            if steps[i, b] > 0:
                assert m < prev
            else:
                assert m == prev
            prev = m
            succ = steps[i, b] > 0
            drho = min(drho / 1.2, 1 / 1.2) if succ else max(drho * 1.2, 1.2)
            rho = min(max(rho * drho, 1e-8), 10.0)
        assert abs(r["rho"][b] - rho) < 1e-12 * rho
        assert r["drho"][b] == 1.0  # reset to its init value after the solve (bsqp.cuh:189)
        # final merit is the merit of the returned trajectory
        fm = O.merit(cfg, r["XU"][b], np.zeros_like(r["XU"][b]), 0.0, wl.xs[b], wl.ref[b], 10.0)
        assert abs(fm - r["final_merit"][b]) < 1e-9 * abs(fm)


def test_reference_exit_rule_is_batch_uniform(O):
    # quirk Q10: every problem's sqp_iters grows by one per loop iteration; quirk Q9/Q11: a problem
    # whose PCG took 0 iterations is "converged"; the loop breaks before the line search once all are.
    N = 6
    cfg = fig8_cfg(O, N=N, iters=5)
    wl = W.drop_in_figure8(N=N, B=3, seed=4)
    r = O.solve(cfg, wl.xu, wl.xs, wl.ref, mu=10.0, pcg_tol=1e-5, rho0=1e-3)
    assert np.all(r["sqp_iters"] == r["n_iters"])
    # a warm start at the solution: identical problems, lambda from a tight solve, zero PCG work
    cfg_t = fig8_cfg(O, N=N, iters=1)
    r2 = O.solve(cfg_t, wl.xu[:1], wl.xs[:1], wl.ref[:1], mu=10.0, pcg_tol=1e-14, rho0=1e-3)
    lam = np.repeat(r2["lam"], 2, 0)
    xu = np.repeat(wl.xu[:1], 2, 0)
    r3 = O.solve(cfg_t, xu, np.repeat(wl.xs[:1], 2, 0), np.repeat(wl.ref[:1], 2, 0), lam=lam, mu=10.0, pcg_tol=1e-14,
                 rho0=1e-3)
    assert np.all(r3["pcg_iters"][0] == 0)
    assert np.all(r3["kkt_converged"] == 1)
    assert r3["ls_num_iters"] == 0  # broke before the line search
    np.testing.assert_array_equal(r3["XU"], xu)


def test_kkt_exit_rule_freezes_converged_problems(O):
    # KKT rule: a problem stops once max(|stationarity|, |defect|) <= kkt_tol; with a huge tolerance
    # every problem stops after its first evaluation and the trajectory is untouched.
    cfg = fig8_cfg(O, iters=4, exit_rule=O.EXIT_KKT, kkt_tol=1e9)
    wl = W.drop_in_figure8(N=8, B=2, seed=5)
    r = O.solve(cfg, wl.xu, wl.xs, wl.ref, mu=10.0)
    assert np.all(r["sqp_iters"] == 1) and np.all(r["kkt_converged"] == 1)
    np.testing.assert_array_equal(r["XU"], wl.xu)
    # with tolerance 0 nobody converges and counts are the iteration limit
    cfg0 = fig8_cfg(O, iters=3, exit_rule=O.EXIT_KKT, kkt_tol=0.0)
    r0 = O.solve(cfg0, wl.xu, wl.xs, wl.ref, mu=10.0)
    assert np.all(r0["sqp_iters"] == 3) and np.all(r0["kkt_converged"] == 0)


def test_joint_space_tracking_descends(O):
    # Joint-space quadratic cost (the plain SQP case): every step is accepted and the KKT residual
    # falls by an order of magnitude within four iterations.
    wl = W.joint_sinusoid(N=8, B=2, seed=0)
    cfg = O.make_cfg(N=8, dt=0.01, max_sqp_iters=5, q_cost=0.0, qd_cost=0.0, u_cost=1e-3, N_cost=0.0,
                     q_lim_cost=0.0, exit_rule=O.EXIT_KKT, kkt_tol=1e-4, terminal_mode=1)
    r = O.solve(cfg, wl.xu, wl.xs, wl.ref, Qd=wl.Qd, Rd=wl.Rd, xref=wl.xref, mu=10.0, pcg_tol=1e-5, rho0=1e-3)
    assert np.all(r["ls_step"][:4] > 0)
    assert np.all(r["kkt_res"][3] < 0.1 * r["kkt_res"][0])
    assert np.all(r["final_merit"] < r["initial_merit"])
