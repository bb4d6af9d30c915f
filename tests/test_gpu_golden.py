"""GPU results against the committed FP64 oracle fixtures (no oracle call at run time)."""
import json

import numpy as np
import pytest

from conftest import ROOT, rel_err

pytestmark = pytest.mark.gpu
FIX = json.loads((ROOT / "tests" / "golden" / "oracle_fixtures.json").read_text())
SCALARS = ("dt", "max_sqp_iters", "kkt_tol", "max_pcg_iters", "pcg_tol", "solve_ratio", "mu", "q_cost", "qd_cost",
           "u_cost", "N_cost", "q_lim_cost", "vel_lim_cost", "ctrl_lim_cost", "rho")


def _joint_solver(fx, mode):
    from gato_amd.solver import Solver

    kw = {k: fx["params"][k] for k in SCALARS}
    kw["max_sqp_iters"] = fx["iters"]
    s = Solver("iiwa14", 8, 1, mode=mode, **kw)
    s.set_joint_cost(np.array(fx["Qd"], np.float32), np.array(fx["Rd"], np.float32), np.array(fx["xref"], np.float32))
    return s


@pytest.mark.parametrize("mode", ["graph", "plain", "fused"])
def test_first_iteration_matches_fixture(gpu, mode):
    fx = FIX["joint_iter1"]
    s = _joint_solver(fx, mode)
    r = s.solve(np.array(fx["xu"], np.float32), fx["params"]["dt"], np.array(fx["xs"], np.float32),
                np.array(fx["ref"], np.float32))
    X = np.array(fx["XU"])
    assert np.linalg.norm(r["XU"] - X) / np.linalg.norm(X) < 1e-4
    np.testing.assert_array_equal(r["ls_step_size"], np.array(fx["ls_step"]))
    assert np.all(np.abs(r["pcg_iters"] - np.array(fx["pcg_iters"])) <= 2)
    assert rel_err(r["initial_merit"], np.array(fx["initial_merit"])) < 1e-5
    assert rel_err(r["final_merit"], np.array(fx["final_merit"])) < 1e-4


def test_full_solve_matches_fixture(gpu):
    # five iterations: the iteration count is exact; the trajectory lies in the envelope that the PCG
    # tolerance itself leaves (two FP64 runs at pcg_tol 1e-5 and 1e-6 differ by up to 1.1e-2 here)
    fx = FIX["joint_full"]
    s = _joint_solver(fx, "graph")
    r = s.solve(np.array(fx["xu"], np.float32), fx["params"]["dt"], np.array(fx["xs"], np.float32),
                np.array(fx["ref"], np.float32))
    np.testing.assert_array_equal(r["sqp_iters"], np.array(fx["sqp_iters"]))
    X = np.array(fx["XU"])
    assert np.linalg.norm(r["XU"] - X) / np.linalg.norm(X) < 2e-2
    assert rel_err(r["final_merit"], np.array(fx["final_merit"])) < 2e-2


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_kkt_blocks_and_dynamics_match_fixture(gpu, plant):
    from gato_amd.solver import Solver

    fx = FIX["kkt"][plant]
    kw = {k: fx["params"][k] for k in SCALARS}
    s = Solver(plant, 8, 1, mode="plain", **kw)
    nq = 7 if plant == "iiwa14" else 6
    nx, nu = 2 * nq, nq
    s.debug_phase(0, np.array(fx["xu"], np.float32)[None], np.array(fx["xs"], np.float32)[None],
                  np.array(fx["ref"], np.float32)[None])
    A = s.debug_array("A").reshape(8, nx, nx)
    Bm = s.debug_array("B").reshape(8, nx, nu)
    q = s.debug_array("q").reshape(8, nx)
    Qp = s.debug_array("Qp").reshape(8, nq, nq)
    assert rel_err(A[:7], np.array(fx["A"])[:7]) < 1e-5
    assert rel_err(Bm[:7], np.array(fx["B"])[:7]) < 1e-5
    assert rel_err(q, np.array(fx["q"])) < 1e-5
    assert rel_err(Qp, np.array(fx["Q"])[:, :nq, :nq]) < 1e-5
    # dynamics: one trapezoidal step with the fixture's forward dynamics (no external wrench)
    for d in FIX["dynamics"][plant]:
        x = np.array(d["q"] + d["qd"], np.float32)
        out = s.sim_forward(x, np.array(d["u"], np.float32), 0.01)[0]
        qdd = np.array(d["qdd"])
        expect = np.concatenate([np.array(d["q"]) + 0.01 * np.array(d["qd"]) + 0.5e-4 * qdd,
                                 np.array(d["qd"]) + 0.01 * qdd])
        assert rel_err(out, expect) < 1e-5
