"""FP64 mode (the reference's USE_DOUBLES build, BSQP_{B}_double) against the FP64 oracle.

The device runs the same SQP iteration as the oracle in the same precision; the two differ only in the order of
their floating-point sums (the device's reductions are trees over knot groups and lanes, the oracle's are loops), so
whole multi-iteration solves are compared, not single iterations: identical SQP iteration counts, line-search steps
and PCG counts (within one where a PCG stop is within rounding of its threshold), trajectories and merits to 1e-8
relative. At fp32 the same comparisons need per-iteration lockstep and 1e-3 bounds (tests/test_gpu_solve.py).
"""
import numpy as np
import pytest

from conftest import rel_err
from gato_amd import workloads as W

pytestmark = pytest.mark.gpu

SCALARS = ("dt", "max_sqp_iters", "kkt_tol", "max_pcg_iters", "pcg_tol", "solve_ratio", "mu", "q_cost", "qd_cost",
           "u_cost", "N_cost", "q_lim_cost", "vel_lim_cost", "ctrl_lim_cost", "rho")


def f32(x):
    """the value the float-typed solver settings carry"""
    return float(np.float32(x))


def make64(wl, **over):
    from gato_amd.solver import Solver64

    kw = {k: wl.params[k] for k in SCALARS}
    kw.update(over)
    return Solver64(wl.plant, wl.N, wl.B, **kw)


def oracle(O, wl, joint=False, **over):
    p = dict(wl.params)
    p.update(over)
    cfg = O.make_cfg(wl.plant, N=wl.N, dt=p["dt"], max_sqp_iters=p["max_sqp_iters"], kkt_tol=f32(p["kkt_tol"]),
                     max_pcg_iters=p["max_pcg_iters"], solve_ratio=f32(p["solve_ratio"]), q_cost=f32(p["q_cost"]),
                     qd_cost=f32(p["qd_cost"]), u_cost=f32(p["u_cost"]), N_cost=f32(p["N_cost"]),
                     q_lim_cost=f32(p["q_lim_cost"]), vel_lim_cost=f32(p["vel_lim_cost"]),
                     ctrl_lim_cost=f32(p["ctrl_lim_cost"]))
    kw = dict(Qd=wl.Qd, Rd=wl.Rd, xref=wl.xref) if joint else {}
    return O.solve(cfg, wl.xu, wl.xs, wl.ref, mu=f32(p["mu"]), pcg_tol=f32(p["pcg_tol"]), rho0=f32(p["rho"]), **kw)


def xu_rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def compare(r, o, tol=1e-8):
    np.testing.assert_array_equal(r["sqp_iters"], o["sqp_iters"])
    n = r["n_iters"]
    assert n == o["n_iters"]
    np.testing.assert_array_equal(r["ls_step_size"], o["ls_step"][: r["ls_num_iters"]].astype(np.float32))
    assert np.abs(r["all_pcg_iters"] - o["pcg_iters"][:n]).max() <= 1
    assert xu_rel(r["XU"], o["XU"]).max() < tol
    assert rel_err(r["initial_merit"], o["initial_merit"]) < tol
    assert rel_err(r["final_merit"], o["final_merit"]) < tol


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_joint_space_solve_matches_oracle(gpu, O, plant):
    # the plain SQP configuration (joint sinusoid, Q = I, R = 1e-3 I, <= 5 SQP iterations), whole solves
    wl = W.joint_sinusoid(plant, N=8, B=4, seed=0)
    s = make64(wl)
    s.set_joint_cost(wl.Qd, wl.Rd, wl.xref)
    r = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)
    o = oracle(O, wl, joint=True)
    compare(r, o)
    lam, rho, _ = s.get_state()
    assert rel_err(lam, o["lam"]) < 1e-6
    np.testing.assert_allclose(rho, o["rho"], rtol=1e-12)


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_figure8_drop_in_solve_matches_oracle(gpu, O, plant):
    # The drop-in figure-8 configuration, ten SQP iterations with the PCG at its cap: chaotic in fp32 against FP64,
    # but the FP64 device solve follows the FP64 oracle through all ten iterations.
    wl = W.drop_in_figure8(plant, N=32, B=16, seed=1)
    r = make64(wl).solve(wl.xu, wl.dt, wl.xs, wl.ref)
    o = oracle(O, wl)
    compare(r, o, tol=1e-6)


@pytest.mark.parametrize("plant,N", [("iiwa14", 3), ("indy7", 5), ("iiwa14", 64), ("indy7", 128)])
def test_short_and_long_horizons_match_oracle(gpu, O, plant, N):
    wl = W.joint_sinusoid(plant, N=N, B=3, seed=N)
    wl.params["max_sqp_iters"] = 3
    s = make64(wl)
    s.set_joint_cost(wl.Qd, wl.Rd, wl.xref)
    r = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)
    compare(r, oracle(O, wl, joint=True), tol=1e-7)


def test_setters_warm_start_and_sim_forward(gpu, O):
    wl = W.drop_in_figure8("iiwa14", N=16, B=4, seed=14)
    wl.params["max_sqp_iters"] = 2
    s = make64(wl)
    mu = np.array([0.0, 1.0, 10.0, 100.0])
    s.set_mu_batch(mu)
    rho = np.array([1e-3, 1e-2, 1e-1, 1.0])
    s.set_rho_penalty_batch(rho)
    fe = np.tile(np.array([0.0, 0.0, 0.0, 5.0, 0.0, 0.0]), (4, 1))
    s.set_f_ext_batch(fe)
    r1 = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)
    lam1, rho1, _ = s.get_state()
    r2 = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)  # warm-started from lambda and rho of the first solve
    p = wl.params
    cfg = O.make_cfg("iiwa14", N=16, dt=p["dt"], max_sqp_iters=2, q_cost=f32(p["q_cost"]), qd_cost=f32(p["qd_cost"]),
                     u_cost=f32(p["u_cost"]), N_cost=f32(p["N_cost"]), q_lim_cost=0.0, max_pcg_iters=p["max_pcg_iters"])
    o1 = O.solve(cfg, wl.xu, wl.xs, wl.ref, fext=fe, mu=mu, pcg_tol=f32(p["pcg_tol"]), rho=rho)
    compare(r1, o1)
    o2 = O.solve(cfg, wl.xu, wl.xs, wl.ref, fext=fe, mu=mu, pcg_tol=f32(p["pcg_tol"]), lam=o1["lam"], rho=o1["rho"])
    compare(r2, o2)
    s.reset_dual()
    s.reset_rho()
    lam, rho_now, _ = s.get_state()
    assert np.all(lam == 0) and np.allclose(rho_now, rho)
    x = np.random.default_rng(3).uniform(-1, 1, 14)
    u = np.random.default_rng(4).normal(0, 5, 7)
    out = s.sim_forward(x, u, 0.01)
    np.testing.assert_allclose(out, O.sim_forward("iiwa14", x, u, fe, 0.01), rtol=1e-12, atol=1e-13)


def test_drop_in_double_classes(gpu):
    import importlib

    mod = importlib.import_module("gato_amd.bsqp.bsqpN8_indy7")
    wl = W.drop_in_figure8("indy7", N=8, B=2, seed=2)
    cls = getattr(mod, "BSQP_2_double")
    s = cls(*[wl.params[k] for k in SCALARS])
    r = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)
    assert r["XU"].dtype == np.float64 and r["XU"].shape == (2, s.traj)
    assert not hasattr(mod, "BSQP_256_double")  # USE_DOUBLES registers B <= 128 (bindings.cu:244-252)


def test_drop_in_configuration_at_the_largest_double_batch(gpu, O):
    # The drop-in figure-8 configuration (iiwa14, N = 32, <= 10 SQP iterations) at B = 128, the largest batch the
    # reference registers in double: the north-star parity bar on whole solves - identical SQP iteration counts on at
    # least 99% of the problems and trajectories within 1e-4 relative - plus identical line-search steps.
    wl = W.drop_in_figure8("iiwa14", N=32, B=128, seed=1)
    r = make64(wl).solve(wl.xu, wl.dt, wl.xs, wl.ref)
    o = oracle(O, wl)
    assert (r["sqp_iters"] == o["sqp_iters"]).mean() >= 0.99
    e = xu_rel(r["XU"], o["XU"])
    assert e.max() < 1e-4, e.max()
    same_steps = (r["ls_step_size"] == o["ls_step"][: r["ls_num_iters"]].astype(np.float32)).all(0)
    assert same_steps.mean() >= 0.99
    assert rel_err(r["final_merit"], o["final_merit"]) < 1e-6
