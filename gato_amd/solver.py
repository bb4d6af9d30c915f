"""Batched SQP solver over PyTorch-ROCm device memory, calling libgato_mi355x.so through its C-ABI.

``Solver`` mirrors the reference's pybind11 class ``PyBSQP<T, B>`` (python/bindings.cu:10-209) and
the ``BSQP<T, B>`` solver it wraps (gato/bsqp/bsqp.cuh:20-353): same constructor scalars and defaults,
same ``solve`` result dictionary, setters and ``sim_forward``. Differences, all additive:
  * the knot count N and batch size B are run-time values;
  * inputs may be numpy arrays or torch tensors (device tensors are used in place);
  * inputs are checked for dtype and size (the reference reads raw buffers unchecked);
  * keyword-only extensions: execution ``mode`` ("fused" persistent kernel, "plain" one kernel
    per phase with a host loop, or "graph" - the plain kernels for every iteration gated on a device
    flag and replayed as one hipGraph), ``exit_rule`` ("pcg0" reference rule or "kkt"), joint-space costs, state get/set.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import GatoParams, GatoStats, check

# reference C++ constructor defaults (bsqp.cuh:25-27)
DEFAULTS = dict(dt=0.01, max_sqp_iters=5, kkt_tol=1e-4, max_pcg_iters=100, pcg_tol=1e-5, solve_ratio=1.0, mu=10.0,
                q_cost=1.0, qd_cost=1e-3, u_cost=1e-6, N_cost=50.0, q_lim_cost=1e-3, vel_lim_cost=0.0,
                ctrl_lim_cost=0.0, rho=1e-3)
CTOR_ORDER = ("dt", "max_sqp_iters", "kkt_tol", "max_pcg_iters", "pcg_tol", "solve_ratio", "mu", "q_cost",
              "qd_cost", "u_cost", "N_cost", "q_lim_cost", "vel_lim_cost", "ctrl_lim_cost", "rho")
MODES = {"plain": _lib.MODE_PLAIN, "fused": _lib.MODE_FUSED, "graph": _lib.MODE_GRAPH}
EXITS = {"pcg0": _lib.EXIT_PCG0, "kkt": _lib.EXIT_KKT}


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _settings(args, kw):
    if len(args) not in (0, len(CTOR_ORDER)):
        raise TypeError(f"expected 0 or {len(CTOR_ORDER)} positional solver scalars, got {len(args)}")
    cfg = dict(DEFAULTS)
    cfg.update(zip(CTOR_ORDER, args))
    unknown = set(kw) - set(DEFAULTS)
    if unknown:
        raise TypeError(f"unknown solver settings {sorted(unknown)}")
    cfg.update(kw)
    return cfg


def _params(cfg, mode, exit_rule, terminal_mode, hess_mode):
    return GatoParams(cfg["dt"], int(cfg["max_sqp_iters"]), cfg["kkt_tol"], int(cfg["max_pcg_iters"]), cfg["pcg_tol"],
                      cfg["solve_ratio"], cfg["mu"], cfg["q_cost"], cfg["qd_cost"], cfg["u_cost"], cfg["N_cost"],
                      cfg["q_lim_cost"], cfg["vel_lim_cost"], cfg["ctrl_lim_cost"], cfg["rho"], MODES[mode],
                      EXITS[exit_rule], int(terminal_mode), int(hess_mode))


class Solver:
    def __init__(self, plant, N, B, *args, mode="graph", exit_rule="pcg0", terminal_mode=0, hess_mode=0,
                 device=None, _handle=None, **kw):
        cfg = _settings(args, kw)
        if plant not in _lib.PLANT_IDS:
            raise ValueError(f"unknown plant {plant!r}")
        self.plant, self.N, self.B = plant, int(N), int(B)
        self.nq, self.nx = _lib.plant_dims(plant)
        self.nu = self.nq
        self.traj = (self.nx + self.nu) * self.N - self.nu
        self.cfg = cfg
        self.device = torch.device(device or "cuda:0")
        if self.device.type != "cuda":
            raise ValueError("the solver runs on a HIP device (torch 'cuda' device on ROCm)")
        self.lib = _lib.load()
        self._owned = _handle is None
        if _handle is None:
            p = _params(cfg, mode, exit_rule, terminal_mode, hess_mode)
            self._h = C.c_void_p()
            with torch.cuda.device(self.device):
                check(self.lib.gato_create(C.byref(self._h), _lib.PLANT_IDS[plant], self.N, self.B, C.byref(p)),
                      "gato_create")
        else:
            self._h = _handle  # a class of a MixedSolver: owned by it
        self.mode, self.exit_rule = mode, exit_rule
        self._pending = []  # backward passes waiting for the state their forward solve left (see _touch)
        f = dict(dtype=torch.float32, device=self.device)
        self._xu = torch.zeros(self.B, self.traj, **f)
        self._xs = torch.zeros(self.B, self.nx, **f)
        self._ref = torch.zeros(self.B, 6 * self.N, **f)
        self._x = torch.zeros(self.nx, **f)
        self._u = torch.zeros(self.nu, **f)
        self._xn = torch.zeros(self.B, self.nx, **f)
        it = int(cfg["max_sqp_iters"])
        self._st = dict(sqp_iters=np.zeros(self.B, np.int32), kkt_converged=np.zeros(self.B, np.int32),
                        initial_merit=np.zeros(self.B, np.float32), final_merit=np.zeros(self.B, np.float32),
                        pcg_iters=np.zeros((it, self.B), np.int32), ls_min_merit=np.zeros((it, self.B), np.float32),
                        ls_step_size=np.zeros((it, self.B), np.float32),
                        kkt_residual=np.zeros((it, self.B), np.float32), status=np.zeros(self.B, np.int32))

    def _stats_struct(self):
        st = self._st
        return GatoStats(0.0, 0, 0, *[st[k].ctypes.data_as(C.POINTER(C.c_int32 if st[k].dtype == np.int32 else C.c_float))
                                      for k in ("sqp_iters", "kkt_converged", "initial_merit", "final_merit",
                                                "pcg_iters", "ls_min_merit", "ls_step_size", "kkt_residual",
                                                "status")])

    def __del__(self):
        h = getattr(self, "_h", None)
        if getattr(self, "_owned", False) and h is not None and h.value:
            try:
                self.lib.gato_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---------------------------------------------------------------- helpers
    def _touch(self):
        """Called before anything changes the solver's state (multipliers, rho, wrenches, weights): a backward
        pass still waiting for the state its forward solve left receives a snapshot of it first."""
        waiting = [p for p in self._pending if p.snap is None and not p.done]
        self._pending = []
        if waiting:
            snap = self.snapshot()
            for p in waiting:
                p.snap = snap

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _to_dev(self, x, dst, name):
        n = dst.numel()
        if isinstance(x, torch.Tensor):
            if x.numel() != n:
                raise ValueError(f"{name}: expected {n} values, got {x.numel()}")
            dst.view(-1).copy_(x.reshape(-1).to(dtype=torch.float32), non_blocking=False)
        else:
            a = np.asarray(x, dtype=np.float32)
            if a.size != n:
                raise ValueError(f"{name}: expected {n} values, got {a.size}")
            dst.view(-1).copy_(torch.from_numpy(np.ascontiguousarray(a).reshape(-1)))
        return dst

    def _host_f32(self, x, n, name):
        a = np.ascontiguousarray(np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x,
                                            dtype=np.float32).reshape(-1))
        if a.size != n:
            raise ValueError(f"{name}: expected {n} values, got {a.size}")
        return a

    # ---------------------------------------------------------------- reference methods
    def solve(self, xu_traj_batch, timestep, x_s_batch, reference_traj_batch, return_torch=False):
        """BSQP::solve on B problems (bindings.cu:68-148). Returns the reference's result dict
        (XU, sqp_time_us, sqp_iters, kkt_converged, final_merit, initial_merit, ls_num_iters,
        pcg_times_us, pcg_iters, ls_min_merit, ls_step_size) plus kkt_residual, n_iters and status (per-problem
        GATO_STATUS_* bits: 1 converged, 2 iteration cap, 4 PCG at its cap, 8 line-search failure, 16 non-finite)."""
        xu = self._to_dev(xu_traj_batch, self._xu, "xu_traj_batch")
        xs = self._to_dev(x_s_batch, self._xs, "x_s_batch")
        ref = self._to_dev(reference_traj_batch, self._ref, "reference_traj_batch")
        st = self._st
        gs = GatoStats(0.0, 0, 0, *[st[k].ctypes.data_as(C.POINTER(C.c_int32 if st[k].dtype == np.int32 else C.c_float))
                                    for k in ("sqp_iters", "kkt_converged", "initial_merit", "final_merit", "pcg_iters",
                                              "ls_min_merit", "ls_step_size", "kkt_residual", "status")])
        with torch.cuda.device(self.device):
            self._touch()
            check(self.lib.gato_solve(self._h, _ptr(xu), _ptr(xs), _ptr(ref), float(timestep), self._stream(),
                                      C.byref(gs)), "gato_solve")
        n_ls = gs.ls_num_iters
        out = {
            "XU": xu.clone() if return_torch else xu.cpu().numpy(),
            "sqp_time_us": float(gs.solve_time_us),
            "sqp_iters": st["sqp_iters"].copy(),
            "kkt_converged": st["kkt_converged"].copy(),
            "status": st["status"].copy(),
            "final_merit": st["final_merit"].copy(),
            "initial_merit": st["initial_merit"].copy(),
            "ls_num_iters": int(n_ls),
            "pcg_times_us": np.zeros(n_ls, np.float32),
            "pcg_iters": st["pcg_iters"][:n_ls].copy(),
            "ls_min_merit": st["ls_min_merit"][:n_ls].copy(),
            "ls_step_size": st["ls_step_size"][:n_ls].copy(),
            "kkt_residual": st["kkt_residual"][: gs.n_iters].copy(),
            "n_iters": int(gs.n_iters),
            "all_pcg_iters": st["pcg_iters"][: gs.n_iters].copy(),
        }
        return out

    def solve_device(self, xu, xs, ref, dt=None, stats=False):
        """Solve in place on device tensors (B x traj, B x nx, B x 6N float32, contiguous) on the
        current stream, without host copies of the trajectory. Returns the stats dict if asked."""
        for t, n, name in ((xu, self.B * self.traj, "xu"), (xs, self.B * self.nx, "xs"), (ref, self.B * 6 * self.N, "ref")):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                    and t.numel() == n):
                raise ValueError(f"{name}: expected a contiguous float32 device tensor with {n} elements")
        dt = self.cfg["dt"] if dt is None else dt
        st = self._st
        gs = None
        if stats:
            gs = GatoStats(0.0, 0, 0, *[st[k].ctypes.data_as(C.POINTER(C.c_int32 if st[k].dtype == np.int32 else C.c_float))
                                        for k in ("sqp_iters", "kkt_converged", "initial_merit", "final_merit",
                                                  "pcg_iters", "ls_min_merit", "ls_step_size", "kkt_residual",
                                                  "status")])
        self._touch()
        check(self.lib.gato_solve(self._h, _ptr(xu), _ptr(xs), _ptr(ref), float(dt), self._stream(),
                                  C.byref(gs) if gs is not None else None), "gato_solve")
        if stats:
            return {k: v.copy() for k, v in st.items()} | {"n_iters": gs.n_iters, "sqp_time_us": gs.solve_time_us}
        return None

    def set_f_ext_batch(self, f_ext_batch):
        f = torch.zeros(self.B * 6, dtype=torch.float32, device=self.device)
        self._to_dev(f_ext_batch, f, "f_ext_batch")
        self._touch()
        check(self.lib.gato_set_f_ext_batch(self._h, _ptr(f), self._stream()), "gato_set_f_ext_batch")

    def set_rho_penalty_batch(self, rho_batch, set_as_reset_default=True):
        a = self._host_f32(rho_batch, self.B, "rho_batch")
        self._touch()
        check(self.lib.gato_set_rho_penalty_batch(self._h, _fptr(a), int(bool(set_as_reset_default))),
              "gato_set_rho_penalty_batch")

    def set_drho_batch(self, drho_batch, set_as_reset_default=True):
        a = self._host_f32(drho_batch, self.B, "drho_batch")
        self._touch()
        check(self.lib.gato_set_drho_batch(self._h, _fptr(a), int(bool(set_as_reset_default))), "gato_set_drho_batch")

    def set_mu_batch(self, mu_batch):
        a = self._host_f32(mu_batch, self.B, "mu_batch")
        self._touch()
        check(self.lib.gato_set_mu_batch(self._h, _fptr(a)), "gato_set_mu_batch")

    def set_pcg_tol_batch(self, eps_batch):
        a = self._host_f32(eps_batch, self.B, "eps_batch")
        self._touch()
        check(self.lib.gato_set_pcg_tol_batch(self._h, _fptr(a)), "gato_set_pcg_tol_batch")

    def sim_forward(self, xk, uk, dt):
        """One trapezoidal step of (xk, uk) under each of the B wrenches (bindings.cu:180-194)."""
        self._to_dev(xk, self._x, "xk")
        self._to_dev(uk, self._u, "uk")
        check(self.lib.gato_sim_forward(self._h, _ptr(self._xn), _ptr(self._x), _ptr(self._u), float(dt),
                                        self._stream()), "gato_sim_forward")
        return self._xn.cpu().numpy()

    def reset_dual(self):
        self._touch()
        check(self.lib.gato_reset_dual(self._h), "gato_reset_dual")

    def reset_rho(self):
        self._touch()
        check(self.lib.gato_reset_rho(self._h), "gato_reset_rho")

    def set_rho_adaptation(self, enabled):
        self._touch()
        check(self.lib.gato_set_rho_adaptation(self._h, int(bool(enabled))), "gato_set_rho_adaptation")

    # ---------------------------------------------------------------- extensions
    def set_mode(self, mode=None, exit_rule=None):
        self.mode = mode or self.mode
        self.exit_rule = exit_rule or self.exit_rule
        self._touch()
        check(self.lib.gato_set_mode(self._h, MODES[self.mode], EXITS[self.exit_rule]), "gato_set_mode")

    def set_joint_cost(self, Qd=None, Rd=None, xref=None):
        """Diagonal joint-space weights (B x nx, B x nu) and joint reference (B x N*nx) replacing
        qd_cost / u_cost; all None restores the reference cost."""
        bufs = []
        for x, n, name in ((Qd, self.B * self.nx, "Qd"), (Rd, self.B * self.nu, "Rd"), (xref, self.B * self.N * self.nx, "xref")):
            if x is None:
                bufs.append(None)
            else:
                t = torch.zeros(n, dtype=torch.float32, device=self.device)
                bufs.append(self._to_dev(x, t, name))
        self._touch()
        check(self.lib.gato_set_joint_cost(self._h, *[None if b is None else _ptr(b) for b in bufs], self._stream()),
              "gato_set_joint_cost")

    def backward_device(self, xu, xs, ref, grad_xu, dt=None):
        """Gradients of a loss through the solve (the adjoint of the SQP step at the solution xu):
        returns device tensors xs [B, nx], ref [B, 6N], Qd [B, nx], Rd [B, nu] and the adjoint PCG
        iteration counts. xu is the trajectory a solve returned; grad_xu = dL/dxu [B, traj]."""
        for t, n, name in ((xu, self.B * self.traj, "xu"), (xs, self.B * self.nx, "xs"), (ref, self.B * 6 * self.N, "ref"),
                           (grad_xu, self.B * self.traj, "grad_xu")):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                    and t.numel() == n):
                raise ValueError(f"{name}: expected a contiguous float32 device tensor with {n} elements")
        dt = self.cfg["dt"] if dt is None else dt
        f = dict(dtype=torch.float32, device=self.device)
        g = dict(xs=torch.empty(self.B, self.nx, **f), ref=torch.empty(self.B, 6 * self.N, **f),
                 Qd=torch.empty(self.B, self.nx, **f), Rd=torch.empty(self.B, self.nu, **f))
        it = np.zeros(self.B, np.int32)
        check(self.lib.gato_backward(self._h, _ptr(xu), _ptr(xs), _ptr(ref), float(dt), _ptr(grad_xu), _ptr(g["xs"]),
                                     _ptr(g["ref"]), _ptr(g["Qd"]), _ptr(g["Rd"]), self._stream(),
                                     it.ctypes.data_as(C.POINTER(C.c_int32))), "gato_backward")
        g["adj_pcg_iters"] = it
        return g

    def ee_positions(self, xu, jacobian=False):
        """End-effector positions [B, N, 3] of the device trajectories xu [B, traj] (and their Jacobians dp/dq
        [B, N, 3, nq] when asked)."""
        if not (isinstance(xu, torch.Tensor) and xu.is_cuda and xu.dtype == torch.float32 and xu.is_contiguous()
                and xu.numel() == self.B * self.traj):
            raise ValueError(f"xu: expected a contiguous float32 device tensor with {self.B * self.traj} elements")
        pos = torch.empty(self.B, self.N, 3, dtype=torch.float32, device=self.device)
        jac = torch.empty(self.B, self.N, 3, self.nq, dtype=torch.float32, device=self.device) if jacobian else None
        check(self.lib.gato_ee_positions(self._h, _ptr(xu), _ptr(pos), None if jac is None else _ptr(jac),
                                         self._stream()), "gato_ee_positions")
        return (pos, jac) if jacobian else pos

    def set_backward_implicit(self, rounds=8, fd_step=1e-4):
        """Backward returns the derivative of the solution (rounds > 0; see gato_set_backward_implicit) instead of
        the one-step derivative (rounds = 0)."""
        check(self.lib.gato_set_backward_implicit(self._h, int(rounds), float(fd_step)), "gato_set_backward_implicit")

    def set_backward_params(self, max_pcg_iters=-1, pcg_tol=1e-6, refine=-1):
        """Backward solves: PCG iteration cap per round (-1: one, with the factorization preconditioner), relative
        tolerance, refinement rounds (-1: the default for this horizon, 2 up to N = 64 and 4 beyond)."""
        check(self.lib.gato_set_backward_params(self._h, int(max_pcg_iters), float(pcg_tol), int(refine)),
              "gato_set_backward_params")

    def backward_params(self):
        it, tol, rounds = C.c_int32(), C.c_float(), C.c_int32()
        check(self.lib.gato_get_backward_params(self._h, C.byref(it), C.byref(tol), C.byref(rounds)),
              "gato_get_backward_params")
        return {"max_pcg_iters": it.value, "pcg_tol": tol.value, "refine": rounds.value}

    def select_hypothesis(self, x_last, u_last, x_curr, dt, xu):
        """MPC step helper: simulate every wrench hypothesis from (x_last, u_last), pick the one closest to the
        measured x_curr, and make its trajectory every problem's warm start in the device tensor xu [B, traj].
        Returns (best index, per-hypothesis errors as numpy)."""
        if not (isinstance(xu, torch.Tensor) and xu.is_cuda and xu.dtype == torch.float32 and xu.is_contiguous()
                and xu.numel() == self.B * self.traj):
            raise ValueError(f"xu: expected a contiguous float32 device tensor with {self.B * self.traj} elements")
        f = dict(dtype=torch.float32, device=self.device)
        xl = self._to_dev(x_last, torch.zeros(self.nx, **f), "x_last")
        ul = self._to_dev(u_last, torch.zeros(self.nu, **f), "u_last")
        xc = self._to_dev(x_curr, torch.zeros(self.nx, **f), "x_curr")
        err = torch.zeros(self.B, **f)
        best = C.c_int32()
        check(self.lib.gato_select_hypothesis(self._h, _ptr(xl), _ptr(ul), _ptr(xc), float(dt), _ptr(xu), _ptr(err),
                                              C.byref(best), self._stream()), "gato_select_hypothesis")
        self.last_selection_errors = err  # device copy, for gato_amd.mpc.ForceEstimator.update
        return best.value, err.cpu().numpy()

    def simulate_rk4(self, x, u, f_ext, h, steps):
        """RK4 plant simulation of every problem's state x [B, nx] (device tensor, in place) under controls
        u [B, nu] and wrenches f_ext [B, 6] (device tensors), `steps` sub-steps of length h."""
        for t, n, name in ((x, self.B * self.nx, "x"), (u, self.B * self.nu, "u"), (f_ext, self.B * 6, "f_ext")):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                    and t.numel() == n):
                raise ValueError(f"{name}: expected a contiguous float32 device tensor with {n} elements")
        check(self.lib.gato_simulate_rk4(self._h, _ptr(x), _ptr(u), _ptr(f_ext), float(h), int(steps), self._stream()),
              "gato_simulate_rk4")

    def window_reference(self, ref, path, offset):
        """ref [B, 6N] (device tensor) <- the N points of the device path [npoints, 6] from `offset` on."""
        if not (isinstance(path, torch.Tensor) and path.is_cuda and path.dtype == torch.float32 and path.is_contiguous()
                and path.numel() % 6 == 0):
            raise ValueError("path: expected a contiguous float32 device tensor of 6-float points")
        if not (isinstance(ref, torch.Tensor) and ref.is_cuda and ref.numel() == self.B * 6 * self.N):
            raise ValueError(f"ref: expected a float32 device tensor with {self.B * 6 * self.N} elements")
        check(self.lib.gato_window_reference(self._h, _ptr(ref), _ptr(path), path.numel() // 6, int(offset),
                                             self._stream()), "gato_window_reference")

    def set_launch_shape(self, groups):
        check(self.lib.gato_set_launch_shape(self._h, int(groups)), "gato_set_launch_shape")

    def launch_shape(self):
        g, t = C.c_int32(), C.c_int32()
        check(self.lib.gato_get_launch_shape(self._h, C.byref(g), C.byref(t)), "gato_get_launch_shape")
        return g.value, t.value

    def autotune(self, xu, xs, ref, dt=None):
        """Pick the fastest launch shape by timing whole solves on these inputs (state restored)."""
        xu = self._to_dev(xu, self._xu, "xu")
        xs = self._to_dev(xs, self._xs, "xs")
        ref = self._to_dev(ref, self._ref, "ref")
        g = C.c_int32()
        self._touch()
        check(self.lib.gato_autotune(self._h, _ptr(xu), _ptr(xs), _ptr(ref), float(self.cfg["dt"] if dt is None else dt),
                                     self._stream(), C.byref(g)), "gato_autotune")
        return g.value

    def snapshot(self):
        """Everything the backward pass reads from the solver besides its arguments: multipliers (device copy),
        rho and drho, the per-problem wrenches and the diagonal joint weights (device copies)."""
        lam = torch.empty(self.B, self.N * self.nx, dtype=torch.float32, device=self.device)
        rho = np.zeros(self.B, np.float32)
        drho = np.zeros(self.B, np.float32)
        check(self.lib.gato_get_state(self._h, _ptr(lam), _fptr(rho), _fptr(drho)), "gato_get_state")
        snap = dict(lam=lam, rho=rho, drho=drho)
        for name, n in (("f_ext", 6), ("Qd", self.nx), ("Rd", self.nu)):
            t = torch.empty(self.B * n, dtype=torch.float32, device=self.device)
            cnt = C.c_int64()
            check(self.lib.gato_debug_copy(self._h, name.encode(), _ptr(t), t.numel(), C.byref(cnt)),
                  "gato_debug_copy")
            snap[name] = t
        return snap

    def restore(self, snap):
        """Put back a snapshot() (the joint reference and cost mode are left as they are)."""
        self._touch()
        check(self.lib.gato_set_state(self._h, _ptr(snap["lam"]), _fptr(snap["rho"]), _fptr(snap["drho"])),
              "gato_set_state")
        check(self.lib.gato_set_f_ext_batch(self._h, _ptr(snap["f_ext"]), self._stream()), "gato_set_f_ext_batch")
        check(self.lib.gato_set_joint_cost(self._h, _ptr(snap["Qd"]), _ptr(snap["Rd"]), None, self._stream()),
              "gato_set_joint_cost")

    def get_state(self):
        lam = torch.zeros(self.B, self.N * self.nx, dtype=torch.float32, device=self.device)
        rho = np.zeros(self.B, np.float32)
        drho = np.zeros(self.B, np.float32)
        check(self.lib.gato_get_state(self._h, _ptr(lam), _fptr(rho), _fptr(drho)), "gato_get_state")
        return lam.cpu().numpy(), rho, drho

    def set_state(self, lam=None, rho=None, drho=None):
        lt = None
        if lam is not None:
            lt = torch.zeros(self.B * self.N * self.nx, dtype=torch.float32, device=self.device)
            self._to_dev(lam, lt, "lambda")
        r = self._host_f32(rho, self.B, "rho") if rho is not None else None
        d = self._host_f32(drho, self.B, "drho") if drho is not None else None
        self._touch()
        check(self.lib.gato_set_state(self._h, None if lt is None else _ptr(lt), None if r is None else _fptr(r),
                                      None if d is None else _fptr(d)), "gato_set_state")

    def debug_phase(self, phase, xu, xs, ref, dt=None):
        """Run one plain-path phase (0 KKT, 1 Schur, 2 PCG, 3 dz, 4 merit terms) on the workspace; 5 is the graph
        path's PCG for 64 < N <= 128 (held on the CU, preconditioner in factored form)."""
        self._to_dev(xu, self._xu, "xu")
        self._to_dev(xs, self._xs, "xs")
        self._to_dev(ref, self._ref, "ref")
        dt = self.cfg["dt"] if dt is None else dt
        self._touch()
        check(self.lib.gato_debug_phase(self._h, int(phase), _ptr(self._xu), _ptr(self._xs), _ptr(self._ref),
                                        float(dt), self._stream()), "gato_debug_phase")
        return self._xu

    def debug_array(self, name):
        cap = self.B * self.N * self.nx * self.nx + self.B * (self.traj + self.nu) + 1024
        dst = torch.zeros(cap, dtype=torch.float32, device=self.device)
        n = C.c_int64()
        check(self.lib.gato_debug_copy(self._h, name.encode(), _ptr(dst), cap, C.byref(n)), "gato_debug_copy")
        out = dst[: n.value].cpu().numpy()
        return out.view(np.int32) if name == "pcg_iters" else out

    def last_kernel_time(self):
        us = C.c_double()
        n = C.c_int32()
        check(self.lib.gato_last_kernel_time(self._h, C.byref(us), C.byref(n)), "gato_last_kernel_time")
        return us.value, n.value


class Solver64:
    """The reference's BSQP<double, B> (USE_DOUBLES, gato/settings.h:7-11; BSQP_{B}_double in
    python/bindings.cu:244-252): the same constructor, solve result and setters as ``Solver``, in float64 (numpy
    float64 or float64 device tensors). Runs the one-kernel-per-phase path at double precision."""

    SET_RHO, SET_DRHO, SET_MU, SET_PCG_TOL = 0, 1, 2, 3
    RESET_DUAL, RESET_RHO = 0, 1

    def __init__(self, plant, N, B, *args, exit_rule="pcg0", terminal_mode=0, hess_mode=0, device=None, **kw):
        cfg = _settings(args, kw)
        if plant not in _lib.PLANT_IDS:
            raise ValueError(f"unknown plant {plant!r}")
        self.plant, self.N, self.B = plant, int(N), int(B)
        self.nq, self.nx = _lib.plant_dims(plant)
        self.nu = self.nq
        self.traj = (self.nx + self.nu) * self.N - self.nu
        self.cfg = cfg
        self.device = torch.device(device or "cuda:0")
        self.lib = _lib.load()
        p = _params(cfg, "plain", exit_rule, terminal_mode, hess_mode)
        self._h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.gato_create_f64(C.byref(self._h), _lib.PLANT_IDS[plant], self.N, self.B, C.byref(p)),
                  "gato_create_f64")
        f = dict(dtype=torch.float64, device=self.device)
        self._xu = torch.zeros(self.B, self.traj, **f)
        self._xs = torch.zeros(self.B, self.nx, **f)
        self._ref = torch.zeros(self.B, 6 * self.N, **f)
        it = int(cfg["max_sqp_iters"])
        self._st = dict(sqp_iters=np.zeros(self.B, np.int32), kkt_converged=np.zeros(self.B, np.int32),
                        initial_merit=np.zeros(self.B, np.float32), final_merit=np.zeros(self.B, np.float32),
                        pcg_iters=np.zeros((it, self.B), np.int32), ls_min_merit=np.zeros((it, self.B), np.float32),
                        ls_step_size=np.zeros((it, self.B), np.float32),
                        kkt_residual=np.zeros((it, self.B), np.float32), status=np.zeros(self.B, np.int32))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.gato_destroy_f64(h)
            except Exception:
                pass
            self._h = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _to_dev(self, x, dst, name):
        n = dst.numel()
        if isinstance(x, torch.Tensor):
            if x.numel() != n:
                raise ValueError(f"{name}: expected {n} values, got {x.numel()}")
            dst.view(-1).copy_(x.reshape(-1).to(dtype=torch.float64))
        else:
            a = np.asarray(x, dtype=np.float64)
            if a.size != n:
                raise ValueError(f"{name}: expected {n} values, got {a.size}")
            dst.view(-1).copy_(torch.from_numpy(np.ascontiguousarray(a).reshape(-1)))
        return dst

    def _host_f64(self, x, name):
        a = np.ascontiguousarray(np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x,
                                            dtype=np.float64).reshape(-1))
        if a.size != self.B:
            raise ValueError(f"{name}: expected {self.B} values, got {a.size}")
        return a

    def solve(self, xu_traj_batch, timestep, x_s_batch, reference_traj_batch, return_torch=False):
        """BSQP<double, B>::solve: the result dict of ``Solver.solve`` with XU in float64 and the merits in double."""
        xu = self._to_dev(xu_traj_batch, self._xu, "xu_traj_batch")
        xs = self._to_dev(x_s_batch, self._xs, "x_s_batch")
        ref = self._to_dev(reference_traj_batch, self._ref, "reference_traj_batch")
        st = self._st
        gs = GatoStats(0.0, 0, 0, *[st[k].ctypes.data_as(C.POINTER(C.c_int32 if st[k].dtype == np.int32 else C.c_float))
                                    for k in ("sqp_iters", "kkt_converged", "initial_merit", "final_merit", "pcg_iters",
                                              "ls_min_merit", "ls_step_size", "kkt_residual", "status")])
        with torch.cuda.device(self.device):
            check(self.lib.gato_solve_f64(self._h, _ptr(xu), _ptr(xs), _ptr(ref), float(timestep), self._stream(),
                                          C.byref(gs)), "gato_solve_f64")
        init, final = np.zeros(self.B), np.zeros(self.B)
        dp = C.POINTER(C.c_double)
        check(self.lib.gato_last_merits_f64(self._h, init.ctypes.data_as(dp), final.ctypes.data_as(dp)),
              "gato_last_merits_f64")
        n_ls = gs.ls_num_iters
        return {
            "XU": xu.clone() if return_torch else xu.cpu().numpy(),
            "sqp_time_us": float(gs.solve_time_us),
            "sqp_iters": st["sqp_iters"].copy(),
            "kkt_converged": st["kkt_converged"].copy(),
            "status": st["status"].copy(),
            "final_merit": final,
            "initial_merit": init,
            "ls_num_iters": int(n_ls),
            "pcg_times_us": np.zeros(n_ls, np.float32),
            "pcg_iters": st["pcg_iters"][:n_ls].copy(),
            "ls_min_merit": st["ls_min_merit"][:n_ls].copy(),
            "ls_step_size": st["ls_step_size"][:n_ls].copy(),
            "kkt_residual": st["kkt_residual"][: gs.n_iters].copy(),
            "n_iters": int(gs.n_iters),
            "all_pcg_iters": st["pcg_iters"][: gs.n_iters].copy(),
        }

    def set_joint_cost(self, Qd=None, Rd=None, xref=None):
        bufs = []
        for x, n, name in ((Qd, self.B * self.nx, "Qd"), (Rd, self.B * self.nu, "Rd"), (xref, self.B * self.N * self.nx, "xref")):
            bufs.append(None if x is None else self._to_dev(x, torch.zeros(n, dtype=torch.float64, device=self.device), name))
        check(self.lib.gato_set_joint_cost_f64(self._h, *[None if b is None else _ptr(b) for b in bufs], self._stream()),
              "gato_set_joint_cost_f64")

    def set_f_ext_batch(self, f_ext_batch):
        f = self._to_dev(f_ext_batch, torch.zeros(self.B * 6, dtype=torch.float64, device=self.device), "f_ext_batch")
        check(self.lib.gato_set_f_ext_batch_f64(self._h, _ptr(f), self._stream()), "gato_set_f_ext_batch_f64")

    def _set(self, what, values, name, reset_default=1):
        a = self._host_f64(values, name)
        check(self.lib.gato_set_batch_f64(self._h, what, a.ctypes.data_as(C.POINTER(C.c_double)), int(reset_default)),
              "gato_set_batch_f64")

    def set_rho_penalty_batch(self, rho_batch, set_as_reset_default=True):
        self._set(self.SET_RHO, rho_batch, "rho_batch", bool(set_as_reset_default))

    def set_drho_batch(self, drho_batch, set_as_reset_default=True):
        self._set(self.SET_DRHO, drho_batch, "drho_batch", bool(set_as_reset_default))

    def set_mu_batch(self, mu_batch):
        self._set(self.SET_MU, mu_batch, "mu_batch")

    def set_pcg_tol_batch(self, eps_batch):
        self._set(self.SET_PCG_TOL, eps_batch, "eps_batch")

    def reset_dual(self):
        check(self.lib.gato_reset_f64(self._h, self.RESET_DUAL), "gato_reset_f64")

    def reset_rho(self):
        check(self.lib.gato_reset_f64(self._h, self.RESET_RHO), "gato_reset_f64")

    def set_rho_adaptation(self, enabled):
        check(self.lib.gato_set_rho_adaptation_f64(self._h, int(bool(enabled))), "gato_set_rho_adaptation_f64")

    def get_state(self):
        lam = torch.zeros(self.B, self.N * self.nx, dtype=torch.float64, device=self.device)
        rho, drho = np.zeros(self.B), np.zeros(self.B)
        dp = C.POINTER(C.c_double)
        check(self.lib.gato_get_state_f64(self._h, _ptr(lam), rho.ctypes.data_as(dp), drho.ctypes.data_as(dp)),
              "gato_get_state_f64")
        return lam.cpu().numpy(), rho, drho

    def sim_forward(self, xk, uk, dt):
        f = dict(dtype=torch.float64, device=self.device)
        x = self._to_dev(xk, torch.zeros(self.nx, **f), "xk")
        u = self._to_dev(uk, torch.zeros(self.nu, **f), "uk")
        out = torch.zeros(self.B, self.nx, **f)
        check(self.lib.gato_sim_forward_f64(self._h, _ptr(out), _ptr(x), _ptr(u), float(dt), self._stream()),
              "gato_sim_forward_f64")
        return out.cpu().numpy()
