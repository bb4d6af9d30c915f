"""Uneven batching and launch shapes on the GPU.

A mixed solve runs each (plant, N) class as a branch of one hipGraph, with the same kernels and launch
shape as a stand-alone solver of that class, so its results must equal the stand-alone solves bit for bit.
The persistent mode runs every problem of every class through the same per-problem code as a stand-alone fused
solve, so it too equals those solves bit for bit at the same knot-group count.
Other launch shapes change only the grouping of partial sums (PCG dot products), so they agree to rounding.
"""
import numpy as np
import pytest
import torch

from gato_amd import workloads as W

pytestmark = pytest.mark.gpu

SCALARS = ("dt", "max_sqp_iters", "kkt_tol", "max_pcg_iters", "pcg_tol", "solve_ratio", "mu", "q_cost", "qd_cost",
           "u_cost", "N_cost", "q_lim_cost", "vel_lim_cost", "ctrl_lim_cost", "rho")


def _classes():
    # (plant, N, B): one-knot-per-group PCG (N <= 32), two and four knots per group (N = 40, 128)
    spec = [("iiwa14", 8, 6), ("indy7", 16, 5), ("iiwa14", 40, 3), ("indy7", 128, 2)]
    wls = [W.drop_in_figure8(p, N=n, B=b, seed=30 + i) for i, (p, n, b) in enumerate(spec)]
    for wl in wls:
        wl.params["max_sqp_iters"] = 3
    return spec, wls


def _dev(wl, dev):
    t = lambda a: torch.tensor(a, dtype=torch.float32, device=dev).contiguous()
    return t(wl.xu), t(wl.xs), t(wl.ref)


@pytest.mark.parametrize("tune", [False, True])
def test_mixed_solve_equals_standalone_solves(gpu, tune):
    from gato_amd.mixed import MixedSolver
    from gato_amd.solver import Solver

    spec, wls = _classes()
    kw = {k: wls[0].params[k] for k in SCALARS}
    m = MixedSolver(spec, autotune=tune, **kw)
    dev = m.device
    ins = [_dev(wl, dev) for wl in wls]
    st = m.solve_device([x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins], stats=True)
    shapes = m.launch_shapes()
    for i, ((p, n, b), wl) in enumerate(zip(spec, wls)):
        s = Solver(p, n, b, mode="graph", **kw)
        s.set_launch_shape(shapes[i][0])
        xu, xs, ref = _dev(wl, dev)
        ref_st = s.solve_device(xu, xs, ref, stats=True)
        assert torch.equal(ins[i][0], xu), (i, shapes[i])
        for key in ("final_merit", "initial_merit", "pcg_iters", "ls_step_size", "sqp_iters"):
            np.testing.assert_array_equal(st[i][key], ref_st[key], err_msg=f"class {i} {key}")
        np.testing.assert_array_equal(m.class_solver(i).get_state()[0], s.get_state()[0])
    # a second solve reuses the captured graph and repeats the first from the same state
    for i in range(len(spec)):
        m.class_solver(i).reset_dual()
        m.class_solver(i).reset_rho()
    ins2 = [_dev(wl, dev) for wl in wls]
    m.solve_device([x[0] for x in ins2], [x[1] for x in ins2], [x[2] for x in ins2])
    for a, b in zip(ins, ins2):
        assert torch.equal(a[0], b[0])
    assert m.last_time_us() > 0


@pytest.mark.parametrize("groups", [32, 16])
def test_persistent_mixed_kernel_equals_standalone_fused_solves(gpu, groups):
    # One k_fused_mixed launch over eight classes (both plants; N = 2 to 128, the shortest with a single problem; one
    # to eight knots per knot group, three of them at N = 96 and 32 groups, run as four passes), against stand-alone
    # fused solves of each class. Where the class's own default knot-group count equals the persistent kernel's
    # (N = 32, 96 and 128 at 32 groups, N = 16 at 16) the results are equal bit for bit; N = 8 and N = 2 run with more groups than knots there (idle groups add zeros to the sums),
    # checked to rounding against their default shapes. A second launch after resetting the classes' state repeats the first exactly (the work-list
    # counters reset themselves).
    from gato_amd.mixed import MixedSolver
    from gato_amd.solver import Solver

    spec = [("iiwa14", 32, 6), ("indy7", 128, 3), ("indy7", 32, 5), ("iiwa14", 128, 2), ("iiwa14", 8, 4),
            ("indy7", 16, 3), ("indy7", 2, 1), ("iiwa14", 96, 2)]
    wls = [W.drop_in_figure8(p, N=n, B=b, seed=50 + i) for i, (p, n, b) in enumerate(spec)]
    for wl in wls:
        wl.params["max_sqp_iters"] = 4
    kw = {k: wls[0].params[k] for k in SCALARS}
    m = MixedSolver(spec, mode="fused", groups=groups, **kw)
    dev = m.device
    ins = [_dev(wl, dev) for wl in wls]
    st = m.solve_device([x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins], stats=True)
    for i, ((p, n, b), wl) in enumerate(zip(spec, wls)):
        s = Solver(p, n, b, mode="fused", **kw)
        default = s.launch_shape()[0]
        exact = default == groups
        if not exact and groups < default:
            s.set_launch_shape(groups)
            exact = True
        xu, xs, ref = _dev(wl, dev)
        ref_st = s.solve_device(xu, xs, ref, stats=True)
        if exact:
            assert torch.equal(ins[i][0], xu), (i, groups)
            for key in ("final_merit", "initial_merit", "pcg_iters", "ls_step_size", "sqp_iters", "kkt_residual"):
                np.testing.assert_array_equal(st[i][key], ref_st[key], err_msg=f"class {i} {key}")
            np.testing.assert_array_equal(m.class_solver(i).get_state()[0], s.get_state()[0])
        else:
            a, r = ins[i][0].cpu().numpy(), xu.cpu().numpy()
            d = np.linalg.norm(a - r, axis=1) / np.linalg.norm(r, axis=1)
            assert d.max() < 1e-5, (i, d.max())
            assert np.all(np.abs(st[i]["pcg_iters"] - ref_st["pcg_iters"]) <= 1), i
            np.testing.assert_array_equal(st[i]["sqp_iters"], ref_st["sqp_iters"])
    for i in range(len(spec)):
        m.class_solver(i).reset_dual()
        m.class_solver(i).reset_rho()
    ins2 = [_dev(wl, dev) for wl in wls]
    m.solve_device([x[0] for x in ins2], [x[1] for x in ins2], [x[2] for x in ins2])
    for a, b in zip(ins, ins2):
        assert torch.equal(a[0], b[0])
    assert m.last_time_us() > 0
    with pytest.raises(Exception):
        m.set_mode("fused", groups=8)  # N = 128 over 8 groups is 16 passes


def test_persistent_kernel_tuner_keeps_state_and_inputs(gpu):
    # groups = 0 with autotuning on: the first solve times every valid knot-group count (8, 16 and 32 here) on
    # scratch copies of the trajectories from the classes' saved state, keeps the fastest and restores the state, so
    # its results equal those of a solver fixed at the chosen count from the start, bit for bit.
    from gato_amd.mixed import MixedSolver

    spec = [("iiwa14", 32, 6), ("indy7", 64, 3), ("iiwa14", 16, 4)]
    wls = [W.drop_in_figure8(p, N=n, B=b, seed=70 + i) for i, (p, n, b) in enumerate(spec)]
    for wl in wls:
        wl.params["max_sqp_iters"] = 3
    kw = {k: wls[0].params[k] for k in SCALARS}
    m1 = MixedSolver(spec, mode="fused", **kw)
    ins1 = [_dev(wl, m1.device) for wl in wls]
    keep = [x[0].clone() for x in ins1]
    st1 = m1.solve_device([x[0] for x in ins1], [x[1] for x in ins1], [x[2] for x in ins1], stats=True)
    G = m1.fused_groups()
    assert G in (8, 16, 32)
    m2 = MixedSolver(spec, mode="fused", groups=G, **kw)
    ins2 = [_dev(wl, m2.device) for wl in wls]
    for a, b in zip(keep, ins2):
        assert torch.equal(a, b[0])  # the tuner left the inputs as they were before the solve
    st2 = m2.solve_device([x[0] for x in ins2], [x[1] for x in ins2], [x[2] for x in ins2], stats=True)
    for i in range(len(spec)):
        assert torch.equal(ins1[i][0], ins2[i][0]), (i, G)
        np.testing.assert_array_equal(st1[i]["pcg_iters"], st2[i]["pcg_iters"])
        np.testing.assert_array_equal(m1.class_solver(i).get_state()[0], m2.class_solver(i).get_state()[0])
        np.testing.assert_array_equal(m1.class_solver(i).get_state()[1], m2.class_solver(i).get_state()[1])


def test_launch_shapes_agree_to_rounding(gpu):
    from gato_amd.solver import Solver

    wl = W.joint_sinusoid("iiwa14", N=8, B=4, seed=0)
    kw = {k: wl.params[k] for k in SCALARS}
    kw["max_sqp_iters"] = 1
    out = {}
    for G in (8, 4):  # one and two knots per group
        s = Solver("iiwa14", 8, 4, mode="graph", **kw)
        s.set_joint_cost(wl.Qd, wl.Rd, wl.xref)
        s.set_launch_shape(G)
        assert s.launch_shape() == (G, 16 * G)
        out[G] = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)
    d = np.linalg.norm(out[8]["XU"] - out[4]["XU"], axis=1) / np.linalg.norm(out[8]["XU"], axis=1)
    assert d.max() < 1e-5
    assert np.all(np.abs(out[8]["pcg_iters"] - out[4]["pcg_iters"]) <= 1)
    with pytest.raises(Exception):
        s.set_launch_shape(12)  # ceil(8 / 12) = 1 but 12 exceeds the default of 8 groups


def test_autotune_keeps_state_and_inputs(gpu):
    from gato_amd.solver import Solver

    wl = W.drop_in_figure8("iiwa14", N=32, B=8, seed=33)
    kw = {k: wl.params[k] for k in SCALARS}
    s = Solver("iiwa14", 32, 8, mode="graph", **kw)
    lam0, rho0, drho0 = s.get_state()
    G = s.autotune(wl.xu, wl.xs, wl.ref)
    assert G in (8, 16, 32) and s.launch_shape()[0] == G
    lam1, rho1, drho1 = s.get_state()
    np.testing.assert_array_equal(lam0, lam1)
    np.testing.assert_array_equal(rho0, rho1)
    np.testing.assert_array_equal(drho0, drho1)


def _oracle_first_iteration(O, wl, idx):
    p = wl.params
    cfg = O.make_cfg(wl.plant, N=wl.N, dt=p["dt"], max_sqp_iters=1, kkt_tol=p["kkt_tol"],
                     max_pcg_iters=p["max_pcg_iters"], solve_ratio=p["solve_ratio"], q_cost=p["q_cost"],
                     qd_cost=p["qd_cost"], u_cost=p["u_cost"], N_cost=p["N_cost"], q_lim_cost=p["q_lim_cost"],
                     vel_lim_cost=p["vel_lim_cost"], ctrl_lim_cost=p["ctrl_lim_cost"])
    f = lambda a: a[idx].astype(np.float32).astype(np.float64)
    o = O.solve(cfg, f(wl.xu), f(wl.xs), f(wl.ref), mu=p["mu"], pcg_tol=p["pcg_tol"], rho0=p["rho"])
    return cfg, o


@pytest.mark.parametrize("mode", ["graph", "fused"])
def test_uneven_batch_workload_full_size_and_first_iteration_against_oracle(gpu, O, mode):
    # The uneven-batch configuration as bench.py runs it: 1024 figure-8 problems, half iiwa14 and half indy7, N drawn
    # from {8, 32, 128}, radius ~U(0.1, 0.25) m and phase ~U(0, 2 pi), one graph launch with launch shapes tuned on
    # the first solve, <= 10 SQP iterations. At full size: finite trajectories, merit non-increasing over the
    # recorded iterations, steps in the step-length set, batch-uniform SQP counts within each class (the reference
    # rule), PCG counts within the cap. Then a one-iteration mixed solve of the same inputs, and for three problems of
    # every (plant, N) class the oracle's first iteration: initial merit and PCG count, and the line search's
    # consistency as in the single-class figure-8 test (the accepted merit is the oracle's merit of the GPU's new
    # iterate; a rejected step leaves the trajectory unchanged bit for bit). Merit tolerance 1e-5 up to N = 32; at
    # N = 128 the fp32 sum over 127 defects of 12-14 components (mu |.|_1 of differences of nearly equal states)
    # carries ~4e-5 of the merit, so 1e-4 there. mode "fused" runs the same through the persistent kernel, where
    # every problem stops on its own (SQP counts within the cap rather than batch-uniform).
    from gato_amd.mixed import MixedSolver

    wls = W.mixed_figure8(B=1024, seed=2)
    keys = sorted(wls)
    assert sum(wls[k].B for k in keys) == 1024 and len(keys) == 6
    kw = {k: wls[keys[0]].params[k] for k in SCALARS}
    allowed = {-1.0} | {1.0 / (1 << j) for j in range(8)}
    m = MixedSolver([(k[0], k[1], wls[k].B) for k in keys], mode=mode, **kw)
    dev = m.device
    ins = [_dev(wls[k], dev) for k in keys]
    st = m.solve_device([x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins], stats=True)
    for i, k in enumerate(keys):
        r = st[i]
        n = r["n_iters"]
        assert torch.isfinite(ins[i][0]).all(), k
        assert 1 <= n <= kw["max_sqp_iters"] and np.all((r["sqp_iters"] >= 1) & (r["sqp_iters"] <= n)), k
        if mode == "graph":
            assert np.all(r["sqp_iters"] == r["sqp_iters"][0]), k
        steps = r["ls_step_size"][:n]
        assert set(np.unique(steps)).issubset(allowed | {0.0}), k
        lsm = np.concatenate([r["initial_merit"][None], r["ls_min_merit"][:n]])
        assert np.all(np.diff(lsm, axis=0) <= 0), k
        assert np.all(r["final_merit"] <= r["initial_merit"]), k
        assert np.all((r["pcg_iters"][:n] >= 0) & (r["pcg_iters"][:n] <= kw["max_pcg_iters"])), k
    kw1 = dict(kw, max_sqp_iters=1)
    m1 = MixedSolver([(k[0], k[1], wls[k].B) for k in keys], mode=mode, **kw1)
    ins1 = [_dev(wls[k], dev) for k in keys]
    st1 = m1.solve_device([x[0] for x in ins1], [x[1] for x in ins1], [x[2] for x in ins1], stats=True)
    for i, k in enumerate(keys):
        wl = wls[k]
        tol = 1e-5 if wl.N <= 32 else 1e-4
        idx = np.array([0, wl.B // 2, wl.B - 1])
        cfg, o = _oracle_first_iteration(O, wl, idx)
        r = st1[i]
        XU = ins1[i][0].cpu().numpy()[idx]
        xu0 = wl.xu[idx].astype(np.float32)
        g_init = r["initial_merit"][idx].astype(np.float64)
        assert np.abs(g_init - o["initial_merit"]).max() <= tol * np.abs(o["initial_merit"]).max(), k
        assert np.all(np.abs(r["pcg_iters"][0][idx] - o["pcg_iters"][0]) <= 2), k
        f = lambda a: a.astype(np.float32).astype(np.float64)
        for j, b in enumerate(idx):
            step = r["ls_step_size"][0][b]
            assert step in allowed, (k, b)
            m_new = O.merit(cfg, XU[j].astype(np.float64), np.zeros(wl.xu.shape[1]), 0.0, f(wl.xs[b]), f(wl.ref[b]),
                            wl.params["mu"])
            assert abs(r["ls_min_merit"][0][b] - m_new) <= tol * abs(m_new), (k, b)
            if step < 0:
                np.testing.assert_array_equal(XU[j], xu0[j])
            else:
                assert m_new < r["initial_merit"][b], (k, b)
