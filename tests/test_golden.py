"""The committed oracle fixtures (tests/golden/oracle_fixtures.json, written by make_oracle_fixtures.py)
are reproduced by the current oracle: a regression pin on the restatement the GPU parity tests rely on."""
import json

import numpy as np
import pytest

from conftest import ROOT

FIX = json.loads((ROOT / "tests" / "golden" / "oracle_fixtures.json").read_text())


@pytest.mark.parametrize("name", ["joint_iter1", "joint_full"])
def test_oracle_reproduces_solve_fixtures(O, name):
    fx = FIX[name]
    p = fx["params"]
    cfg = O.make_cfg("iiwa14", N=8, dt=p["dt"], max_sqp_iters=fx["iters"], max_pcg_iters=p["max_pcg_iters"],
                     q_cost=p["q_cost"], qd_cost=p["qd_cost"], u_cost=p["u_cost"], N_cost=p["N_cost"],
                     q_lim_cost=p["q_lim_cost"])
    a = lambda k: np.array(fx[k])
    r = O.solve(cfg, a("xu"), a("xs"), a("ref"), Qd=a("Qd"), Rd=a("Rd"), xref=a("xref"), mu=p["mu"],
                pcg_tol=p["pcg_tol"], rho0=p["rho"])
    np.testing.assert_allclose(r["XU"], a("XU"), rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(r["pcg_iters"], a("pcg_iters"))
    np.testing.assert_array_equal(r["ls_step"], a("ls_step"))
    np.testing.assert_allclose(r["final_merit"], a("final_merit"), rtol=1e-12)


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_oracle_reproduces_kkt_and_dynamics_fixtures(O, plant):
    fx = FIX["kkt"][plant]
    p = fx["params"]
    cfg = O.make_cfg(plant, N=8, dt=p["dt"], q_cost=p["q_cost"], qd_cost=p["qd_cost"], u_cost=p["u_cost"],
                     N_cost=p["N_cost"], q_lim_cost=p["q_lim_cost"])
    k = O.setup_kkt(cfg, np.array(fx["xu"]), np.array(fx["xs"]), np.array(fx["ref"]))
    for key in ("A", "B", "c", "Q", "q", "R", "r"):
        np.testing.assert_allclose(k[key], np.array(fx[key]), rtol=1e-12, atol=1e-14)
    for d in FIX["dynamics"][plant]:
        q = np.array(d["q"])
        pp, J = O.fk(plant, q)
        np.testing.assert_allclose(pp, d["p"], rtol=1e-12)
        np.testing.assert_allclose(O.mass(plant, q), d["M"], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(O.fd(plant, q, np.array(d["qd"]), np.array(d["u"])), d["qdd"], rtol=1e-10)
