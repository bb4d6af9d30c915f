"""GPU parity of the MPC-caller helpers: the figure-8 path, the wrench estimator and its frame transform.

Compared with the numpy restatements (gato_amd.workloads.reference_figure8_path, oracle/mpc_oracle.py) on the
same inputs. Tolerances: the path is evaluated in double on both sides and rounded to float32 once (at most one
float32 ulp apart, from the libm/device sin); the estimator's float32 recurrences agree to float32 rounding
(rtol 1e-5); radius and confidence are double on both sides (rtol 1e-12).
"""
import numpy as np
import pytest
import torch

from conftest import rel_err
from gato_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_figure8_path_matches_numpy(gpu):
    from gato_amd.mpc import figure8_path

    for kw in ({}, dict(A_x=0.3, A_z=0.2, offset=(0.1, 0.4, 0.5), period=4, cycles=3, theta=0.3)):
        dev = figure8_path(0.01, **kw).cpu().numpy()
        ref = W.reference_figure8_path(0.01, **kw).astype(np.float32)
        assert dev.shape == ref.shape
        np.testing.assert_allclose(dev, ref, rtol=0, atol=1.2e-7)


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_force_estimator_sequence_matches_restatement(gpu, O, plant):
    from oracle import mpc_oracle as M
    from gato_amd.mpc import ForceEstimator
    from gato_amd.solver import Solver

    B, nq = 16, O.nq(plant)
    s = Solver(plant, 8, B, mode="graph")
    cfg = dict(initial_radius=5.0, min_radius=2.0, max_radius=20.0, smoothing_factor=0.5)
    est = ForceEstimator(s, seed=7, **cfg)
    ref = M.EstimatorRef(B, **cfg)
    rng = np.random.default_rng(3)
    winners = [1, 4, 0, 9, 2, 15, 3, 3, 0, 7]
    for step, win in enumerate(winners):
        q = rng.uniform(-1.2, 1.2, nq).astype(np.float32)
        world = est.generate_batch(q).cpu().numpy()
        wref = ref.batch()
        scale = max(1.0, np.abs(wref).max())
        np.testing.assert_allclose(world, wref, rtol=0, atol=1e-5 * scale, err_msg=f"world batch, step {step}")
        fext = s.debug_array("f_ext").reshape(B, 6)
        fref = np.stack([M.wrench_to_solver_frame(plant, q.astype(np.float64), wref[b]) for b in range(B)])
        np.testing.assert_allclose(fext, fref, rtol=0, atol=1e-5 * scale, err_msg=f"solver wrenches, step {step}")
        errors = rng.uniform(0.5, 2.0, B).astype(np.float32)
        errors[win] = 0.1 + 0.01 * step  # the winner is the first minimum
        u = tuple(rng.uniform(0, 1, 3))
        est.update(torch.from_numpy(errors).to(s.device), alpha=0.6, beta=0.5, best=win, uniforms=u)
        ref.update(win, errors.astype(np.float64), 0.6, 0.5, u)
        st = est.state()
        for k, v in (("estimate", ref.est), ("momentum", ref.mom), ("smoothed_estimate", ref.smooth)):
            np.testing.assert_allclose(st[k], v, rtol=1e-5, atol=1e-5 * scale, err_msg=f"{k}, step {step}")
        assert st["radius"] == pytest.approx(ref.radius, rel=1e-12), step
        assert st["confidence"] == pytest.approx(ref.confidence, rel=1e-12, abs=1e-15), step
        assert st["updates"] == step + 1
        np.testing.assert_allclose(st["rotation"], ref.R, atol=1e-6)
    est.reset()
    st = est.state()
    assert st["radius"] == 5.0 and st["updates"] == 0 and not st["estimate"].any()


def test_force_estimator_follows_the_solver_selection(gpu, O):
    from oracle import mpc_oracle as M
    from gato_amd.mpc import ForceEstimator
    from gato_amd.solver import Solver

    plant, B, N = "iiwa14", 12, 8
    s = Solver(plant, N, B, mode="graph")
    est = ForceEstimator(s, initial_radius=5.0, min_radius=2.0, max_radius=20.0, smoothing_factor=0.5, seed=1)
    ref = M.EstimatorRef(B, initial_radius=5.0, min_radius=2.0, max_radius=20.0, smoothing_factor=0.5)
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-0.5, 0.5, 7), np.zeros(7)]).astype(np.float32)
    u = rng.normal(0, 1, 7).astype(np.float32)
    xu = torch.zeros(B, s.traj, dtype=torch.float32, device=s.device)
    for step in range(3):
        est.generate_batch(x[:7])
        ref_batch = ref.batch()
        # the "measured" state: hypothesis 5's prediction
        x_next = s.sim_forward(x, u, 0.01).reshape(B, -1)[5].astype(np.float32)
        best, err = s.select_hypothesis(x, u, x_next, 0.01, xu)
        uu = tuple(rng.uniform(0, 1, 3))
        est.update(err, alpha=0.6, beta=0.5, uniforms=uu)  # best taken from the solver's selection
        ref.update(int(np.argmin(err)), err.astype(np.float64), 0.6, 0.5, uu)
        assert best == int(np.argmin(err))
        st = est.state()
        scale = max(1.0, np.abs(ref_batch).max())
        np.testing.assert_allclose(st["estimate"], ref.est, rtol=1e-5, atol=1e-5 * scale)
        assert st["radius"] == pytest.approx(ref.radius, rel=1e-12)
        x = x_next


@pytest.mark.parametrize("model", ["iiwa14", "indy7", "indy7_urdf"])
def test_plant_simulation_follows_the_harness(gpu, O, model):
    # Sub-steps 7 .. 31 of 1 ms under per-knot controls (knot min(int(i / 10), N - 1), mpc_controller.py:197-201)
    # against RK4 (common.py:49-91) on the oracle's forward dynamics of the same plant model, with and without
    # wrenches; the simulated indy7 (URDF centres of mass) must differ from the solver's model.
    from gato_amd.mpc import PlantSimulator

    n, N, dt, h = 3, 8, 0.01, 1e-3
    nq = 7 if model == "iiwa14" else 6
    nx, st = 2 * nq, 3 * nq
    rng = np.random.default_rng(41)
    x0 = rng.uniform(-1, 1, (n, nx)).astype(np.float32)
    xu = rng.normal(0, 5, (n, st * N - nq)).astype(np.float32)
    fe = rng.normal(0, 2, (n, 6)).astype(np.float32)
    sim = PlantSimulator(model, N, dt, h)
    t = lambda a: torch.tensor(a, dtype=torch.float32, device="cuda:0").contiguous()
    first, steps = 7, 25
    outs = {}
    for wrench in (None, fe):
        x = t(x0)
        sim.substeps(x, t(xu), first, steps, None if wrench is None else t(wrench))
        outs[wrench is None] = x.cpu().numpy()
        f = lambda a: a.astype(np.float64)
        for b in range(n):
            q, v = f(x0[b, :nq]), f(x0[b, nq:])
            w = None if wrench is None else f(wrench[b])
            for i in range(first, first + steps):
                k = min(int(i / (dt / h)), N - 1)
                u = f(xu[b, k * st + nx:k * st + nx + nq])
                acc = lambda qq, vv: O.fd(model, qq, vv, u, w)
                k1 = acc(q, v)
                v2 = v + 0.5 * h * k1
                k2 = acc(q + 0.5 * h * v, v2)
                v3 = v + 0.5 * h * k2
                k3 = acc(q + 0.5 * h * v2, v3)
                v4 = v + h * k3
                k4 = acc(q + h * v3, v4)
                q, v = q + h / 6 * (v + 2 * v2 + 2 * v3 + v4), v + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            assert rel_err(outs[wrench is None][b], np.concatenate([q, v])) < 2e-5, (b, wrench is None)
    if model == "indy7_urdf":
        x = t(x0)
        PlantSimulator("indy7", N, dt, h).substeps(x, t(xu), first, steps)
        assert rel_err(x.cpu().numpy(), outs[True]) > 1e-3
    # the harness's clock: int(seconds / h) sub-steps, the remainders accumulated into one more sub-step
    x = t(x0)
    assert [sim.advance(x, t(xu), 0.0025) for _ in range(2)] == [2, 3]
