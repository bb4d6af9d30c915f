"""GPU backward pass (adjoint of the SQP step at the solution) against the FP64 oracle.

The oracle's step_grad is pinned by central differences of the oracle's step map and by the implicit
derivative at a fixed point (tests/test_oracle_adjoint.py). Here the HIP backward, called through the
C-ABI, must reproduce it at the GPU's own solution, rho and weights: 1e-3 relative to each gradient's
largest entry (the tolerance set for adjoint gradients against finite differences).
"""
import numpy as np
import pytest
import torch

from gato_amd import workloads as W

pytestmark = pytest.mark.gpu


def _problem(plant, N, B, seed, hess_mode=0, terminal_mode=0):
    from gato_amd.solver import Solver

    wl = W.drop_in_figure8(plant, N=N, B=B, seed=seed)
    rng = np.random.default_rng(seed)
    nq = W.NQ[plant]
    nx, nu = 2 * nq, nq
    Qd = np.exp(rng.uniform(np.log(1e-2), np.log(1e1), (B, nx))).astype(np.float32)
    Rd = np.exp(rng.uniform(np.log(1e-3), np.log(1e0), (B, nu))).astype(np.float32)
    kw = dict(dt=0.01, max_sqp_iters=3, max_pcg_iters=200, pcg_tol=1e-6, mu=10.0, q_cost=10.0, N_cost=20.0,
              qd_cost=1e-3, u_cost=1e-3, q_lim_cost=0.0, vel_lim_cost=0.0, ctrl_lim_cost=0.0, rho=1e-2)
    s = Solver(plant, N, B, mode="graph", hess_mode=hess_mode, terminal_mode=terminal_mode, **kw)
    s.set_joint_cost(Qd, Rd)
    return s, wl, Qd, Rd, kw


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("plant,hess,term,N,groups", [("iiwa14", 0, 0, 16, None), ("iiwa14", 1, 1, 16, None),
                                                       ("indy7", 0, 0, 16, None), ("iiwa14", 0, 0, 32, None),
                                                       ("iiwa14", 0, 0, 128, None), ("iiwa14", 0, 0, 5, None),
                                                       ("indy7", 0, 0, 3, None), ("iiwa14", 0, 0, 16, 4)])
def test_backward_matches_oracle_step_gradient(gpu, O, plant, hess, term, N, groups):
    # N = 32: the differentiation configuration's horizon; N = 128: the scale configuration's (streamed PCG);
    # N = 5 and 3: odd horizons on small work-groups (eight and four knot groups), where the preconditioner's
    # four sweep sides share waves and the factorization's middle knot sits off centre; N = 16 on four knot groups:
    # the streamed PCG with both sweep sides in the work-group's one wave
    B = 4 if N <= 32 else 2
    s, wl, Qd, Rd, kw = _problem(plant, N, B, seed=21, hess_mode=hess, terminal_mode=term)
    if groups is not None:
        s.set_launch_shape(groups)
    dev = s.device
    xu = torch.tensor(wl.xu, dtype=torch.float32, device=dev)
    xs = torch.tensor(wl.xs, dtype=torch.float32, device=dev)
    ref = torch.tensor(wl.ref, dtype=torch.float32, device=dev)
    s.solve_device(xu, xs, ref)
    rng = np.random.default_rng(22)
    gxu = rng.normal(0, 1, wl.xu.shape).astype(np.float32)
    g = s.backward_device(xu, xs, ref, torch.tensor(gxu, device=dev))
    assert np.all(g["adj_pcg_iters"] > 0)
    _, rho, _ = s.get_state()
    cfg = O.make_cfg(plant, N=N, dt=kw["dt"], q_cost=kw["q_cost"], N_cost=kw["N_cost"], qd_cost=kw["qd_cost"],
                     u_cost=kw["u_cost"], q_lim_cost=0.0, hess_mode=hess, terminal_mode=term)
    X = xu.cpu().numpy().astype(np.float64)
    f = lambda a: a.astype(np.float32).astype(np.float64)
    for b in range(B):
        o = O.step_grad(cfg, X[b], f(wl.xs[b]), f(wl.ref[b]), float(rho[b]), gxu[b].astype(np.float64),
                        Qd=Qd[b].astype(np.float64), Rd=Rd[b].astype(np.float64))
        assert _rel(g["xs"][b].cpu().numpy(), o["xs"]) < 1e-3, b
        assert _rel(g["ref"][b].cpu().numpy().reshape(N, 6), o["ref"]) < 1e-3, b
        assert _rel(g["Qd"][b].cpu().numpy(), o["Qd"]) < 1e-3, b
        assert _rel(g["Rd"][b].cpu().numpy(), o["Rd"]) < 1e-3, b


def test_backward_refinement_rounds_default_by_horizon(gpu):
    # two rounds reach the fp32 floor up to N = 64; at N = 128 the default is four (each round cuts the error by
    # about 8x there, and two leave about 1e-2); -1 restores the horizon's default after an explicit setting
    from gato_amd.solver import Solver

    for N, rounds in ((32, 2), (64, 2), (128, 4)):
        s = Solver("iiwa14", N, 2)
        assert s.backward_params()["refine"] == rounds
        s.set_backward_params(refine=6)
        assert s.backward_params()["refine"] == 6
        s.set_backward_params()
        assert s.backward_params() == {"max_pcg_iters": -1, "pcg_tol": pytest.approx(1e-6), "refine": rounds}


def test_backward_leaves_the_solver_state_alone(gpu):
    s, wl, Qd, Rd, _ = _problem("iiwa14", 8, 2, seed=23)
    dev = s.device
    xu = torch.tensor(wl.xu, dtype=torch.float32, device=dev)
    xs = torch.tensor(wl.xs, dtype=torch.float32, device=dev)
    ref = torch.tensor(wl.ref, dtype=torch.float32, device=dev)
    s.solve_device(xu, xs, ref)
    lam, rho, drho = s.get_state()
    x_before = xu.clone()
    s.backward_device(xu, xs, ref, torch.ones_like(xu))
    lam2, rho2, drho2 = s.get_state()
    np.testing.assert_array_equal(lam, lam2)
    np.testing.assert_array_equal(rho, rho2)
    assert torch.equal(xu, x_before)


def test_autograd_function_routes_through_the_backward(gpu):
    from gato_amd.autograd import solve

    s, wl, Qd, Rd, _ = _problem("iiwa14", 8, 2, seed=24)
    dev = s.device
    xs = torch.tensor(wl.xs, dtype=torch.float32, device=dev, requires_grad=True)
    ref = torch.tensor(wl.ref, dtype=torch.float32, device=dev, requires_grad=True)
    Qt = torch.tensor(Qd, device=dev, requires_grad=True)
    Rt = torch.tensor(Rd, device=dev, requires_grad=True)
    xu0 = torch.tensor(wl.xu, dtype=torch.float32, device=dev)
    XU = solve(s, xs, ref, Qt, Rt, xu0)
    nx = W.NQ["iiwa14"] * 2
    loss = (XU[:, nx:nx + 7] ** 2).sum() + (XU ** 2).sum() * 1e-3  # u_0 energy + a small trajectory term
    loss.backward()
    gx = 2e-3 * XU.detach().clone()
    gx[:, nx:nx + 7] += 2 * XU.detach()[:, nx:nx + 7]
    g = s.backward_device(XU.detach().contiguous(), xs.detach(), ref.detach(), gx.contiguous())
    for t, k in ((xs, "xs"), (ref, "ref"), (Qt, "Qd"), (Rt, "Rd")):
        assert t.grad is not None and torch.isfinite(t.grad).all()
        torch.testing.assert_close(t.grad, g[k], rtol=1e-5, atol=1e-6)


def test_autograd_backward_uses_its_own_forward_state(gpu):
    # two forward solves on one solver, then the backward of the first: it must see the first forward's weights,
    # rho and multipliers (saved when the second forward changed them), give the gradients of a backward run right
    # after the first forward, and leave the solver in the second forward's state
    from gato_amd.autograd import solve

    s, wl, Qd, Rd, _ = _problem("iiwa14", 8, 2, seed=25)
    dev = s.device
    t = lambda a, g=False: torch.tensor(a, dtype=torch.float32, device=dev, requires_grad=g)
    xu0 = t(wl.xu)

    def fwd(scale):
        xs, ref, Qt, Rt = t(wl.xs, True), t(wl.ref, True), t(Qd * scale, True), t(Rd * scale, True)
        XU = solve(s, xs, ref, Qt, Rt, xu0)
        return XU, (xs, ref, Qt, Rt)

    s.reset_dual()
    s.reset_rho()
    XU1, in1 = fwd(1.0)
    (XU1 ** 2).sum().backward()
    expect = [x.grad.clone() for x in in1]
    s.reset_dual()
    s.reset_rho()
    XU1b, in1b = fwd(1.0)
    torch.testing.assert_close(XU1b, XU1, rtol=0, atol=0)
    XU2, in2 = fwd(7.0)  # different weights, warm-started multipliers, adapted rho
    after2 = s.snapshot()
    (XU1b ** 2).sum().backward()
    for x, e in zip(in1b, expect):
        torch.testing.assert_close(x.grad, e, rtol=0, atol=0)
    now = s.snapshot()
    for k in ("lam", "f_ext", "Qd", "Rd"):
        assert torch.equal(now[k], after2[k]), k
    np.testing.assert_array_equal(now["rho"], after2["rho"])
    (XU2 ** 2).sum().backward()
    assert all(torch.isfinite(x.grad).all() for x in in2)


def _config4_problems(O, n, N=32, seed=3):
    """Problems of the differentiation configuration (iiwa14 figure-8, weights log-uniform on [1e-3, 1e2], the
    consistent terminal block) whose step has a fixed point from the figure-8 start in FP64, with that solution."""
    wl = W.drop_in_figure8("iiwa14", N=N, B=64, seed=seed)
    rng = np.random.default_rng(seed)
    Qd = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), (64, 14)))
    Rd = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), (64, 7)))
    cfg = O.make_cfg("iiwa14", N=N, dt=0.01, q_cost=10.0, N_cost=10.0, qd_cost=1e-3, u_cost=1e-3, q_lim_cost=0.0,
                     terminal_mode=1, hess_mode=0)
    out = []
    for b in range(64):
        z, lam, res = O.fixed_point(cfg, wl.xu[b], wl.xs[b], wl.ref[b], Qd=Qd[b], Rd=Rd[b], iters=40)
        if res <= 1e-12:
            out.append((wl.xs[b], wl.ref[b], Qd[b], Rd[b], z))
        if len(out) == n:
            break
    assert len(out) == n
    return cfg, out


def _loss4(O, z, ref, N=32, nx=14, nu=7):
    st = nx + nu
    g = np.zeros_like(z)
    g[nx:nx + nu] = 2 * z[nx:nx + nu]
    L = z[nx:nx + nu] @ z[nx:nx + nu]
    r = ref.reshape(N, 6)
    for k in range(N):
        p, J = O.fk("iiwa14", z[k * st:k * st + 7])
        d = p - r[k, :3]
        L += d @ d
        g[k * st:k * st + 7] += 2 * J.T @ d
    return L, g


def test_implicit_backward_is_the_gradient_of_the_solution(gpu, O):
    # The differentiation configuration at the problems whose step has a fixed point: the GPU solve settles at the
    # FP64 solution (started there), then the backward with implicit rounds must give the oracle's derivative of the
    # solution (exact KKT Jacobian, pinned by central differences of re-solves on the CPU) at the GPU's own point, to
    # 1e-3 of each gradient's largest entry. Config 4's own check runs too: central differences with eps = 1e-3 of the
    # FP64 re-solved loss for x_s and the control log-weights of one problem. The one-step gradient misses both by
    # more than 5%.
    from gato_amd.solver import Solver

    N, B = 32, 3
    cfg, probs = _config4_problems(O, B, N)
    xs = np.array([p[0] for p in probs])
    ref = np.array([p[1] for p in probs])
    Qd = np.array([p[2] for p in probs])
    Rd = np.array([p[3] for p in probs])
    z0 = np.array([p[4] for p in probs])
    s = Solver("iiwa14", N, B, mode="graph", terminal_mode=1, dt=0.01, max_sqp_iters=2, max_pcg_iters=200,
               pcg_tol=1e-7, mu=1e3, q_cost=10.0, N_cost=10.0, qd_cost=1e-3, u_cost=1e-3, q_lim_cost=0.0, rho=1e-3)
    s.set_joint_cost(Qd.astype(np.float32), Rd.astype(np.float32))
    dev = s.device
    t = lambda a: torch.tensor(a, dtype=torch.float32, device=dev).contiguous()
    xu, xs_d, ref_d = t(z0), t(xs), t(ref)
    s.solve_device(xu, xs_d, ref_d)
    Z = xu.cpu().numpy().astype(np.float64)
    dL = np.array([_loss4(O, Z[b], ref[b])[1] for b in range(B)])
    s.set_backward_params(max_pcg_iters=-1, pcg_tol=1e-6, refine=2)
    g1 = s.backward_device(xu, xs_d, ref_d, t(dL))
    s.set_backward_implicit(rounds=8)
    gi = s.backward_device(xu, xs_d, ref_d, t(dL))
    keys = ("xs", "ref", "Qd", "Rd")
    worst_one = 0.0
    for b in range(B):
        zb = Z[b]
        _, lam, _ = O.step_map(cfg, zb, xs[b], ref[b], 1e-3, Qd=Qd[b], Rd=Rd[b])
        o = O.solution_grad(cfg, zb, lam, xs[b], ref[b], dL[b], Qd=Qd[b], Rd=Rd[b])
        assert o["F_inf"] < 1e-2, (b, o["F_inf"])  # the fp32 solution is a solution to its rounding
        for k in keys:
            ob = o[k][:, :3].reshape(-1) if k == "ref" else o[k]
            gb = gi[k][b].cpu().numpy().reshape(N, 6)[:, :3].reshape(-1) if k == "ref" else gi[k][b].cpu().numpy()
            g1b = g1[k][b].cpu().numpy().reshape(N, 6)[:, :3].reshape(-1) if k == "ref" else g1[k][b].cpu().numpy()
            assert _rel(gb, ob) < 1e-3, (b, k, _rel(gb, ob))
            worst_one = max(worst_one, _rel(g1b, ob))
    assert worst_one > 0.05
    # central differences of the re-solved FP64 loss, x_s and log diag R of problem 0
    eps = 1e-3

    def resolved(xs_b, Rd_b):
        zz, _, rr = O.fixed_point(cfg, Z[0], xs_b, ref[0], Qd=Qd[0], Rd=Rd_b, iters=40, tol=1e-11)
        assert rr <= 1e-11
        return _loss4(O, zz, ref[0])[0]

    fd_xs = np.zeros(14)
    for i in range(14):
        e = np.zeros(14)
        e[i] = eps
        fd_xs[i] = (resolved(xs[0] + e, Rd[0]) - resolved(xs[0] - e, Rd[0])) / (2 * eps)
    fd_R = np.zeros(7)
    for i in range(7):
        f = np.ones(7)
        f[i] = np.exp(eps)
        fd_R[i] = (resolved(xs[0], Rd[0] * f) - resolved(xs[0], Rd[0] / f)) / (2 * eps)
    assert _rel(gi["xs"][0].cpu().numpy(), fd_xs) < 1e-3
    assert _rel(gi["Rd"][0].cpu().numpy() * Rd[0], fd_R) < 1e-3
    assert _rel(g1["xs"][0].cpu().numpy(), fd_xs) > 0.05


def test_differentiation_configuration_full_size(gpu, O):
    # The differentiation configuration at full size: iiwa14, N = 32, B = 256, per-problem diagonal weights
    # log-uniform on [1e-3, 1e2], loss sum_b |u_0|^2 + sum_k |p_k - ref_k|^2 through the autograd functions (the
    # solve and the end-effector positions), gradients w.r.t. x_s, the reference, diag Q and diag R. Full size: every
    # gradient finite, the end-effector op against the oracle's forward kinematics. On three sampled problems the
    # one-step gradient (the backward's default) against the oracle's step gradient at the GPU's own solution, rho and
    # weights, with the loss's explicit dependence on the reference added, to 1e-3.
    from gato_amd.autograd import ee_positions, solve
    from gato_amd.solver import Solver

    wl = W.differentiation_figure8(N=32, B=256, seed=3)
    p = wl.params
    kw = {k: p[k] for k in ("dt", "max_sqp_iters", "max_pcg_iters", "pcg_tol", "mu", "q_cost", "qd_cost", "u_cost",
                            "N_cost", "q_lim_cost", "vel_lim_cost", "ctrl_lim_cost", "rho")}
    s = Solver("iiwa14", 32, 256, mode="graph", terminal_mode=1, **kw)
    dev = s.device
    t = lambda a, g=False: torch.tensor(a, dtype=torch.float32, device=dev, requires_grad=g)
    xs, ref, Qt, Rt = t(wl.xs, True), t(wl.ref, True), t(wl.Qd, True), t(wl.Rd, True)
    XU = solve(s, xs, ref, Qt, Rt, t(wl.xu))
    pos = ee_positions(s, XU)
    loss = (XU[:, 14:21] ** 2).sum() + ((pos - ref.view(256, 32, 6)[:, :, :3]) ** 2).sum()
    loss.backward()
    for x in (xs, ref, Qt, Rt):
        assert x.grad is not None and torch.isfinite(x.grad).all()
    Z = XU.detach().cpu().numpy().astype(np.float64)
    P = pos.detach().cpu().numpy()
    for b in (0, 100, 255):
        for k in (0, 17, 31):
            pk, _ = O.fk("iiwa14", Z[b, k * 21:k * 21 + 7])
            assert np.abs(P[b, k] - pk).max() < 1e-5
    _, rho, _ = s.get_state()
    cfg = O.make_cfg("iiwa14", N=32, dt=p["dt"], q_cost=p["q_cost"], N_cost=p["N_cost"], qd_cost=p["qd_cost"],
                     u_cost=p["u_cost"], q_lim_cost=0.0, terminal_mode=1)
    f = lambda a: a.astype(np.float32).astype(np.float64)
    for b in (0, 128, 255):
        L, dL = _loss4(O, Z[b], f(wl.ref[b]))
        o = O.step_grad(cfg, Z[b], f(wl.xs[b]), f(wl.ref[b]), float(rho[b]), dL, Qd=f(wl.Qd[b]), Rd=f(wl.Rd[b]))
        expl = np.zeros((32, 6))
        for k in range(32):
            pk, _ = O.fk("iiwa14", Z[b, k * 21:k * 21 + 7])
            expl[k, :3] = -2 * (pk - f(wl.ref[b]).reshape(32, 6)[k, :3])
        assert _rel(xs.grad[b].cpu().numpy(), o["xs"]) < 1e-3, b
        assert _rel(ref.grad[b].cpu().numpy().reshape(32, 6), o["ref"] + expl) < 1e-3, b
        assert _rel(Qt.grad[b].cpu().numpy(), o["Qd"]) < 1e-3, b
        assert _rel(Rt.grad[b].cpu().numpy(), o["Rd"]) < 1e-3, b
