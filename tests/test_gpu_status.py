"""Per-problem failure status (gato_stats.status) against the statistics it summarises.

The reference reports failures only through its per-iteration arrays (step -1 for a failed line search, PCG
counts at the cap) and leaves PCGStats::converged unset (gato/types.cuh:28). The status bits must agree with
those arrays problem by problem, and a problem with a NaN in its data must be flagged non-finite without
changing any other problem's result.
"""
import numpy as np
import pytest
import torch

from gato_amd import _lib
from gato_amd import workloads as W
from gato_amd.solver import Solver

pytestmark = pytest.mark.gpu

SCALARS = ("dt", "max_sqp_iters", "kkt_tol", "max_pcg_iters", "pcg_tol", "solve_ratio", "mu", "q_cost", "qd_cost",
           "u_cost", "N_cost", "q_lim_cost", "vel_lim_cost", "ctrl_lim_cost", "rho")


def expected_status(r, max_sqp, max_pcg):
    """the bits from the returned per-problem and per-iteration statistics"""
    n = np.minimum(r["sqp_iters"], max_sqp)
    B = len(n)
    rows = np.arange(max_sqp)[:, None] < n[None, :]
    pcg = np.zeros(B, bool)
    if r["all_pcg_iters"].size:
        pcg = ((r["all_pcg_iters"] >= max_pcg) & rows[: r["all_pcg_iters"].shape[0]]).any(0)
    ls = np.zeros(B, bool)
    if r["ls_step_size"].size:
        ls = ((r["ls_step_size"] < 0) & rows[: r["ls_step_size"].shape[0]]).any(0)
    conv = r["kkt_converged"] != 0
    bad = ~np.isfinite(r["XU"]).all(1) | ~np.isfinite(r["initial_merit"]) | ~np.isfinite(r["final_merit"])
    return (conv * _lib.STATUS_KKT_CONVERGED + (~conv & (n >= max_sqp)) * _lib.STATUS_ITER_CAP
            + pcg * _lib.STATUS_PCG_AT_CAP + ls * _lib.STATUS_LS_FAILED + bad * _lib.STATUS_NONFINITE).astype(np.int32)


@pytest.mark.parametrize("mode", ["plain", "graph", "fused"])
def test_status_summarises_the_statistics(gpu, mode):
    wl = W.drop_in_figure8("iiwa14", N=32, B=64, seed=5)
    cfg = {k: wl.params[k] for k in SCALARS}
    s = Solver("iiwa14", 32, 64, mode=mode, **cfg)
    r = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)
    want = expected_status(r, cfg["max_sqp_iters"], cfg["max_pcg_iters"])
    np.testing.assert_array_equal(r["status"], want)
    # the drop-in figure-8 configuration runs to its iteration cap with rejected steps and capped PCG solves
    assert (r["status"] & _lib.STATUS_ITER_CAP).any()
    assert (r["status"] & _lib.STATUS_LS_FAILED).any()
    assert (r["status"] & _lib.STATUS_PCG_AT_CAP).any()
    assert not (r["status"] & _lib.STATUS_NONFINITE).any()


def test_converged_problems_are_flagged(gpu):
    # joint-space tracking under the KKT stopping rule with a loose tolerance: problems that meet it carry bit 1
    # (the residual of this workload levels off near 0.047 after three iterations in the FP64 oracle)
    wl = W.joint_sinusoid("indy7", N=8, B=32, seed=2)
    cfg = {k: wl.params[k] for k in SCALARS}
    cfg["kkt_tol"] = 0.05
    cfg["max_sqp_iters"] = 10
    s = Solver("indy7", 8, 32, mode="graph", exit_rule="kkt", **cfg)
    s.set_joint_cost(wl.Qd, wl.Rd, wl.xref)
    r = s.solve(wl.xu, wl.dt, wl.xs, wl.ref)
    np.testing.assert_array_equal(r["status"], expected_status(r, 10, cfg["max_pcg_iters"]))
    conv = r["kkt_converged"] != 0
    assert conv.any()
    np.testing.assert_array_equal((r["status"] & _lib.STATUS_KKT_CONVERGED) != 0, conv)
    assert not (r["status"][conv] & _lib.STATUS_ITER_CAP).any()


@pytest.mark.parametrize("mode", ["plain", "graph"])
def test_a_nan_problem_is_flagged_and_leaves_the_others_unchanged(gpu, mode):
    # Every index in the kernels is integer arithmetic and every loop has a fixed bound (the PCG stops at its
    # cap), so a NaN in one problem's data only changes that problem's values.
    wl = W.drop_in_figure8("indy7", N=16, B=16, seed=8)
    cfg = {k: wl.params[k] for k in SCALARS}
    cfg["max_sqp_iters"] = 4
    clean = Solver("indy7", 16, 16, mode=mode, **cfg).solve(wl.xu, wl.dt, wl.xs, wl.ref)
    xs = wl.xs.copy()
    xs[5, 2] = np.nan
    bad = Solver("indy7", 16, 16, mode=mode, **cfg).solve(wl.xu, wl.dt, xs, wl.ref)
    assert bad["status"][5] & _lib.STATUS_NONFINITE
    np.testing.assert_array_equal(bad["status"], expected_status(bad, 4, cfg["max_pcg_iters"]))
    keep = np.arange(16) != 5
    assert not (bad["status"][keep] & _lib.STATUS_NONFINITE).any()
    np.testing.assert_array_equal(bad["XU"][keep], clean["XU"][keep])
    np.testing.assert_array_equal(bad["status"][keep], clean["status"][keep])


def test_fp64_status(gpu):
    from gato_amd.solver import Solver64

    wl = W.drop_in_figure8("iiwa14", N=16, B=8, seed=9)
    cfg = {k: wl.params[k] for k in SCALARS}
    cfg["max_sqp_iters"] = 3
    r = Solver64("iiwa14", 16, 8, **cfg).solve(wl.xu, wl.dt, wl.xs, wl.ref)
    np.testing.assert_array_equal(r["status"], expected_status(r, 3, cfg["max_pcg_iters"]))
