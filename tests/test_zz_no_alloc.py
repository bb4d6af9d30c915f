"""The allocation-free solve (gato_solve_noalloc) against gato_solve on a freshly created solver.

gato_solve_noalloc runs the same kernels in the same order as gato_solve in graph mode, on a workspace carved
from the library's static per-GPU buffer instead of a hipMalloc'd pool, starting from a fresh solver's state.
Its trajectories and statistics must therefore equal a fresh solver's bit for bit. Two calls in a row, the
second on a smaller batch, check that the buffer's leftovers from one call do not leak into the next; a call
whose workspace does not fit is refused with an error.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from gato_amd import _lib
from gato_amd import workloads as W
from gato_amd.solver import Solver, _params

pytestmark = pytest.mark.gpu

SCALARS = ("dt", "max_sqp_iters", "kkt_tol", "max_pcg_iters", "pcg_tol", "solve_ratio", "mu", "q_cost", "qd_cost",
           "u_cost", "N_cost", "q_lim_cost", "vel_lim_cost", "ctrl_lim_cost", "rho")
STAT_KEYS = ("sqp_iters", "kkt_converged", "initial_merit", "final_merit", "pcg_iters", "ls_min_merit",
             "ls_step_size", "kkt_residual", "status")


def _stats(B, iters):
    st = dict(sqp_iters=np.zeros(B, np.int32), kkt_converged=np.zeros(B, np.int32),
              initial_merit=np.zeros(B, np.float32), final_merit=np.zeros(B, np.float32),
              pcg_iters=np.zeros((iters, B), np.int32), ls_min_merit=np.zeros((iters, B), np.float32),
              ls_step_size=np.zeros((iters, B), np.float32), kkt_residual=np.zeros((iters, B), np.float32),
              status=np.zeros(B, np.int32))
    ptrs = [st[k].ctypes.data_as(C.POINTER(C.c_int32 if st[k].dtype == np.int32 else C.c_float)) for k in STAT_KEYS]
    return st, _lib.GatoStats(0.0, 0, 0, *ptrs)


def _noalloc(wl, cfg, xu):
    lib = _lib.load()
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(a, dtype=torch.float32, device=dev).contiguous()
    xs, ref = t(wl.xs), t(wl.ref)
    st, gs = _stats(wl.B, int(cfg["max_sqp_iters"]))
    prm = _params(cfg, "graph", "pcg0", 0, 0)
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    rc = lib.gato_solve_noalloc(_lib.PLANT_IDS[wl.plant], wl.N, wl.B, C.byref(prm), C.c_void_p(xu.data_ptr()),
                                C.c_void_p(xs.data_ptr()), C.c_void_p(ref.data_ptr()), float(cfg["dt"]), stream,
                                C.byref(gs))
    return rc, st, gs


def _fresh(wl, cfg):
    s = Solver(wl.plant, wl.N, wl.B, mode="graph", **cfg)
    dev = s.device
    t = lambda a: torch.tensor(a, dtype=torch.float32, device=dev).contiguous()
    xu = t(wl.xu)
    st = s.solve_device(xu, t(wl.xs), t(wl.ref), stats=True)
    return xu, st


@pytest.mark.parametrize("plant,N,Bs", [("iiwa14", 32, (128, 24)), ("indy7", 8, (96, 5))])
def test_noalloc_solve_equals_a_fresh_solver(gpu, plant, N, Bs):
    for B in Bs:  # the second call is smaller and follows the first on the same static buffer
        wl = W.drop_in_figure8(plant, N=N, B=B, seed=40 + B)
        cfg = {k: wl.params[k] for k in SCALARS}
        cfg["max_sqp_iters"] = 4
        print(f"noalloc {plant} N={N} B={B}", flush=True)
        xu = torch.tensor(wl.xu, dtype=torch.float32, device="cuda:0").contiguous()
        rc, st, gs = _noalloc(wl, cfg, xu)
        assert rc == 0, _lib.load().gato_last_error().decode()
        xu_ref, st_ref = _fresh(wl, cfg)
        assert torch.equal(xu, xu_ref), (plant, N, B)
        for k in STAT_KEYS:
            np.testing.assert_array_equal(st[k], st_ref[k], err_msg=f"{k} B={B}")
        assert gs.n_iters == st_ref["n_iters"]
        assert np.isfinite(xu.cpu().numpy()).all()


def test_noalloc_refuses_a_workspace_beyond_its_buffer(gpu):
    wl = W.drop_in_figure8("iiwa14", N=128, B=512, seed=1)
    cfg = {k: wl.params[k] for k in SCALARS}
    xu = torch.tensor(wl.xu, dtype=torch.float32, device="cuda:0").contiguous()
    before = xu.clone()
    rc, _, _ = _noalloc(wl, cfg, xu)
    assert rc == 1001  # GATO_E_ARG
    assert "static" in _lib.load().gato_last_error().decode()
    assert torch.equal(xu, before)
