"""Uneven batching: problems of several horizons and both plants solved by one launch.

``MixedSolver(classes, *scalars, **settings)`` takes classes ``[(plant, N, B), ...]`` that share the solver
scalars (the reference constructor's 15, ``solver.DEFAULTS``). ``solve_device`` solves every class in place
from per-class device tensors. With ``mode="graph"`` (the default) each class is one branch of a single captured
graph, so all horizons and plants run concurrently on the GPU, and the exit rule (solve_ratio) applies per class.
With ``mode="fused"`` one persistent kernel solves every problem of every class, work-groups of ``groups`` knot
groups taking problems from a work list, larger classes first; each problem runs to its own exit, as
``Solver(mode="fused")`` solves a class. ``class_solver(i)`` returns a ``Solver`` view of class i for the per-class
setters. No reference counterpart: the reference builds one module per (N, plant) and solves one batch per call.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check
from .solver import EXITS, MODES, Solver, _params, _settings


class MixedSolver:
    def __init__(self, classes, *args, exit_rule="pcg0", terminal_mode=0, hess_mode=0, autotune=True, device=None,
                 mode="graph", groups=0, **kw):
        self.cfg = _settings(args, kw)
        self.classes = [(str(p), int(n), int(b)) for p, n, b in classes]
        if not self.classes:
            raise ValueError("at least one class is needed")
        for p, _, _ in self.classes:
            if p not in _lib.PLANT_IDS:
                raise ValueError(f"unknown plant {p!r}")
        self.device = torch.device(device or "cuda:0")
        self.lib = _lib.load()
        n = len(self.classes)
        plants = (C.c_int32 * n)(*[_lib.PLANT_IDS[p] for p, _, _ in self.classes])
        knots = (C.c_int32 * n)(*[k for _, k, _ in self.classes])
        batches = (C.c_int32 * n)(*[b for _, _, b in self.classes])
        prm = _params(self.cfg, "graph", exit_rule, terminal_mode, hess_mode)
        self._h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.gato_mixed_create(C.byref(self._h), n, plants, knots, batches, C.byref(prm)),
                  "gato_mixed_create")
        check(self.lib.gato_mixed_set_autotune(self._h, int(bool(autotune))), "gato_mixed_set_autotune")
        self.set_mode(mode, groups)
        self._views = []
        for i, (p, k, b) in enumerate(self.classes):
            h = C.c_void_p()
            check(self.lib.gato_mixed_class(self._h, i, C.byref(h)), "gato_mixed_class")
            self._views.append(Solver(p, k, b, mode="graph", exit_rule=exit_rule, terminal_mode=terminal_mode,
                                      hess_mode=hess_mode, device=self.device, _handle=h, **self.cfg))
        self._stats = [v._stats_struct() for v in self._views]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._views = []
                self.lib.gato_mixed_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_mode(self, mode, groups=0):
        """``"graph"``: one hipGraph with a branch per class; ``"fused"``: one persistent kernel whose work-groups
        have ``groups`` knot groups (0: tuned on the first solve among 8, 16 and 32 when autotuning is on, else
        32)."""
        if mode not in ("graph", "fused"):
            raise ValueError(f"mode must be 'graph' or 'fused', not {mode!r}")
        check(self.lib.gato_mixed_set_mode(self._h, MODES[mode], int(groups)), "gato_mixed_set_mode")
        self.mode = mode

    def fused_groups(self):
        """Knot groups per work-group of the persistent kernel (after tuning, the chosen count)."""
        mode, g = C.c_int32(), C.c_int32()
        check(self.lib.gato_mixed_get_mode(self._h, C.byref(mode), C.byref(g)), "gato_mixed_get_mode")
        return g.value

    def class_solver(self, i):
        return self._views[i]

    def solve_device(self, xu, xs, ref, dt=None, stats=False):
        """Solve every class in place: xu[i], xs[i], ref[i] are contiguous float32 device tensors of class i."""
        n = len(self.classes)
        if not (len(xu) == len(xs) == len(ref) == n):
            raise ValueError(f"expected {n} tensors per argument")
        for i, v in enumerate(self._views):
            for t, m, name in ((xu[i], v.B * v.traj, "xu"), (xs[i], v.B * v.nx, "xs"), (ref[i], v.B * 6 * v.N, "ref")):
                if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
                        and t.numel() == m):
                    raise ValueError(f"class {i} {name}: expected a contiguous float32 device tensor with {m} elements")
        P = C.c_void_p * n
        pxu = P(*[t.data_ptr() for t in xu])
        pxs = P(*[t.data_ptr() for t in xs])
        pref = P(*[t.data_ptr() for t in ref])
        st = (_lib.GatoStats * n)(*self._stats) if stats else None
        dt = self.cfg["dt"] if dt is None else dt
        for v in self._views:
            v._touch()
        check(self.lib.gato_mixed_solve(self._h, pxu, pxs, pref, float(dt),
                                        C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream), st),
              "gato_mixed_solve")
        if not stats:
            return None
        return [{k: a.copy() for k, a in v._st.items()} | {"n_iters": st[i].n_iters} for i, v in enumerate(self._views)]

    def last_time_us(self):
        us = C.c_double()
        check(self.lib.gato_mixed_last_time(self._h, C.byref(us)), "gato_mixed_last_time")
        return us.value

    def launch_shapes(self):
        return [v.launch_shape() for v in self._views]
