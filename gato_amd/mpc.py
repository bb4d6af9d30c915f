"""Device-side helpers for the MPC caller of the solve (the loop of python/bsqp/mpc_controller.py:136-262).

The reference's controller keeps its wrench estimator (examples/force_estimator.py), the figure-8 path
(python/bsqp/common.py:10-46) and the frame transform of every hypothesis (mpc_controller.py:279-337) on the
host with numpy and pinocchio, and moves them across PCIe each control step. Here they run on the device next to
the solver: ``figure8_path`` builds the path, ``ForceEstimator`` generates the hypotheses straight into the
solver's per-problem wrenches and updates itself from ``Solver.select_hypothesis``'s result, with no host copy.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dev_f32(t, n, name, device):
    if isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous():
        if t.numel() != n:
            raise ValueError(f"{name}: expected {n} values, got {t.numel()}")
        return t
    a = np.asarray(t.detach().cpu() if isinstance(t, torch.Tensor) else t, dtype=np.float32).reshape(-1)
    if a.size != n:
        raise ValueError(f"{name}: expected {n} values, got {a.size}")
    return torch.from_numpy(a.copy()).to(device)


def figure8_path(dt, A_x=0.4, A_z=0.4, offset=(0.0, 0.5, 0.6), period=6, cycles=5, theta=np.pi / 4,
                 device="cuda:0"):
    """The figure-8 end-effector path of common.py:10-46 (same arguments and defaults) as a float32 device
    tensor of int(period / dt) * cycles * 6 values: xyz of the rotated curve and zero rpy per point."""
    lib = _lib.load()
    m, cyc = int(period / dt), int(cycles)
    if m < 1 or cyc < 1:
        raise ValueError("period / dt and cycles must give at least one point")
    path = torch.empty(m * cyc * 6, dtype=torch.float32, device=device)
    off = (C.c_double * 3)(*[float(v) for v in offset])
    check(lib.gato_figure8_path(_ptr(path), m, cyc, float(A_x), float(A_z), off, float(theta),
                                C.c_void_p(torch.cuda.current_stream(device).cuda_stream)), "gato_figure8_path")
    return path


class PlantSimulator:
    """The MPC harness's plant advance (mpc_controller.py:187-213, rk4 of common.py:49-91) on the device.

    ``model``: "iiwa14", "indy7" (the solver's GRiD model) or "indy7_urdf" (the plant the harness simulates with
    pinocchio: indy7 with its links' centres of mass from the URDF). ``advance(x, xu, seconds)`` integrates the
    states x [n, nx] (device tensor, in place) over the time the harness gives one control step: int(seconds / h)
    RK4 sub-steps of h, sub-step i applying the control of knot min(int(i / (dt / h)), N - 1) of the trajectories xu
    [n, traj], then, as the harness does with the remainder, one more sub-step each time the accumulated remainders
    reach h. Returns the number of sub-steps run (the harness adds h per sub-step to its clock).
    """

    def __init__(self, model, knots, dt, h=1e-3, device="cuda:0"):
        if model not in _lib.SIM_MODELS:
            raise ValueError(f"unknown plant model {model!r}")
        self.model, self.knots, self.dt, self.h = model, int(knots), float(dt), float(h)
        nq = 7 if model == "iiwa14" else 6
        self.nx, self.traj = 2 * nq, 3 * nq * int(knots) - nq
        self.device = device
        self.accumulated = 0.0
        self.lib = _lib.load()

    def substeps(self, x, xu, first, steps, f_ext=None):
        """Sub-steps first .. first + steps - 1 (gato_simulate_plant); x [n, nx] in place."""
        n = x.numel() // self.nx
        x = _dev_f32(x, n * self.nx, "x", self.device)
        xu = _dev_f32(xu, n * self.traj, "xu", self.device)
        fe = None if f_ext is None else _dev_f32(f_ext, n * 6, "f_ext", self.device)
        st = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        check(self.lib.gato_simulate_plant(_lib.SIM_MODELS[self.model], n, _ptr(x), _ptr(xu), self.knots,
                                           None if fe is None else _ptr(fe), self.h, int(first), int(steps), self.dt,
                                           st), "gato_simulate_plant")
        return x

    def advance(self, x, xu, seconds, f_ext=None):
        nsteps = int(seconds / self.h)
        if nsteps > 0:
            self.substeps(x, xu, 0, nsteps, f_ext)
        run = nsteps
        rem = seconds % self.h
        if rem > 1e-5:
            self.accumulated += rem
            if self.accumulated >= self.h:
                self.accumulated = 0.0
                self.substeps(x, xu, nsteps, 1, f_ext)
                run += 1
        return run


class ForceEstimator:
    """The wrench estimator of examples/force_estimator.py with its state on the device, bound to a solver
    whose batch holds one wrench hypothesis per problem.

    ``generate_batch(q)`` is update_force_batch (mpc_controller.py:279-292): it writes every hypothesis, mapped
    to the solver's wrench frame at joint positions q, into the solver's per-problem wrenches and returns the
    world-frame batch [B, 6]. ``update(prediction_errors, alpha, beta)`` is ForceEstimator.update for the
    hypothesis the solver's last ``select_hypothesis`` chose (or ``best`` when given). ``uniforms`` fixes the
    three random numbers of the next exploration rotation (numpy.random.rand in the reference); by default the
    estimator draws its own from ``seed``.
    """

    def __init__(self, solver, initial_radius=10.0, min_radius=1.0, max_radius=100.0, smoothing_factor=0.3,
                 seed=0):
        self.lib = _lib.load()
        self.solver = solver
        self.batch_size = solver.B
        self.device = solver.device
        self._initial_radius = float(initial_radius)
        h = C.c_void_p()
        check(self.lib.gato_force_est_create(C.byref(h), int(solver.B), float(initial_radius), float(min_radius),
                                             float(max_radius), float(smoothing_factor), int(seed) & (2**64 - 1)),
              "gato_force_est_create")
        self._h = h
        self._world = torch.zeros(solver.B * 6, dtype=torch.float32, device=self.device)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.gato_force_est_destroy(h)
            except Exception:
                pass
            self._h = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def generate_batch(self, q):
        """Hypotheses at joint positions q: solver wrenches set on the device; returns the world batch [B, 6]."""
        qd = _dev_f32(q, self.solver.nq, "q", self.device)
        self.solver._touch()  # the solver's wrenches change
        check(self.lib.gato_force_est_generate(self._h, self.solver._h, _ptr(qd), _ptr(self._world), self._stream()),
              "gato_force_est_generate")
        return self._world.view(self.batch_size, 6)

    def update(self, prediction_errors=None, alpha=0.5, beta=0.8, best=None, uniforms=None):
        """Move the estimate toward the winning hypothesis (force_estimator.py:99-136). prediction_errors: the
        per-hypothesis errors [B] (default: those of the solver's last select_hypothesis, still on the device);
        best: the winning index (device int32 tensor or int; default the solver's last select_hypothesis
        choice)."""
        if prediction_errors is None:
            prediction_errors = getattr(self.solver, "last_selection_errors", None)
            if prediction_errors is None:
                raise ValueError("no prediction errors: pass them or run solver.select_hypothesis first")
        err = _dev_f32(prediction_errors, self.batch_size, "prediction_errors", self.device)
        bptr = None
        if best is not None:
            if isinstance(best, torch.Tensor) and best.is_cuda and best.dtype == torch.int32:
                bt = best
            else:
                bt = torch.tensor([int(best)], dtype=torch.int32, device=self.device)
            if not 0 <= int(bt.reshape(-1)[0]) < self.batch_size:
                raise ValueError("best outside the batch")
            bptr = _ptr(bt)
        u = None
        if uniforms is not None:
            uu = np.asarray(uniforms, dtype=np.float64).reshape(-1)
            if uu.size != 3:
                raise ValueError("uniforms: expected three values")
            u = (C.c_double * 3)(*uu.tolist())
        check(self.lib.gato_force_est_update(self._h, self.solver._h, bptr, _ptr(err), float(alpha), float(beta), u,
                                             self._stream()), "gato_force_est_update")

    def reset(self, radius=None):
        """Initial state (force_estimator.py:150-155); the radius restarts at `radius` (default: the initial)."""
        r = self._initial_radius if radius is None else float(radius)
        check(self.lib.gato_force_est_reset(self._h, r, self._stream()), "gato_force_est_reset")

    def state(self):
        """Dictionary of the estimator state (host copies)."""
        buf = np.zeros(_lib.FORCE_EST_STATE_LEN, dtype=np.float64)
        check(self.lib.gato_force_est_state(self._h, buf.ctypes.data_as(C.POINTER(C.c_double)), self._stream()),
              "gato_force_est_state")
        return {"estimate": buf[0:6].copy(), "momentum": buf[6:12].copy(), "smoothed_estimate": buf[12:18].copy(),
                "rotation": buf[18:27].reshape(3, 3).copy(), "radius": float(buf[27]), "confidence": float(buf[28]),
                "updates": int(buf[29])}
