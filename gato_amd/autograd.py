"""Differentiable batched solve: torch.autograd.Function over the HIP solve and its adjoint.

    XU = solve(solver, xs, ref, Qd, Rd, xu0)

runs the batched SQP solve from the initial trajectory xu0 with initial states xs [B, nx], end-effector
references ref [B, 6N] and per-problem diagonal joint weights Qd [B, nx], Rd [B, nu] (the cost's
joint-space terms 1/2 Qd (x - 0)^2 + 1/2 Rd u^2 next to the end-effector tracking term), and returns the
solution XU [B, traj]. Backward gives dL/dxs, dL/dref, dL/dQd, dL/dRd through the adjoint of the SQP step
at the returned trajectory (gato_backward in the C-ABI). No reference counterpart: the reference has no
backward.

What the gradient is: the derivative of one more SQP step z + dz(z; theta) at the returned z. Where the solve
has converged (dz = 0, KKT residual at its floor) that is the implicit derivative of the solution under the
solver's Hessian model, i.e. the gradient of the solution; where it has not, it is the derivative of that step
only. ``solve(..., return_stats=True)`` also returns the solve's statistics so a caller can check convergence
(``kkt_residual`` of the last iteration) before trusting the gradient as the solution's.

The backward re-linearises with the multipliers, rho, wrenches and joint weights the forward solve left in the
solver. If the solver is used again before backward runs (another forward, a setter), that state is saved just
before it changes and restored around the backward, so each backward sees its own forward's state.
"""
import torch


class _Pending:
    """The solver state a forward solve left, taken only if the solver changes before the backward runs."""

    __slots__ = ("snap", "done")

    def __init__(self):
        self.snap = None
        self.done = False


class _SQPSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, xs, ref, Qd, Rd, xu0, box):
        xs_c = xs.detach().contiguous().float()
        ref_c = ref.detach().contiguous().float()
        solver.set_joint_cost(Qd.detach().contiguous().float(), Rd.detach().contiguous().float())
        xu = xu0.detach().clone().contiguous().float()
        st = solver.solve_device(xu, xs_c, ref_c, stats=box is not None)
        if box is not None:
            box["stats"] = st
        ctx.solver = solver
        ctx.pending = _Pending()
        solver._pending.append(ctx.pending)
        ctx.save_for_backward(xu, xs_c, ref_c)
        return xu.clone()

    @staticmethod
    def backward(ctx, grad_xu):
        xu, xs, ref = ctx.saved_tensors
        s, pend = ctx.solver, ctx.pending
        if pend.snap is None:  # nothing has changed the solver since the forward solve
            pend.done = True
            g = s.backward_device(xu, xs, ref, grad_xu.contiguous().float())
        else:
            now = s.snapshot()
            pend.done = True
            s.restore(pend.snap)
            try:
                g = s.backward_device(xu, xs, ref, grad_xu.contiguous().float())
            finally:
                s.restore(now)
        return None, g["xs"], g["ref"], g["Qd"], g["Rd"], None, None


def solve(solver, xs, ref, Qd, Rd, xu0, return_stats=False):
    """Differentiable solve (see the module docstring). With return_stats, returns (XU, stats dict)."""
    box = {} if return_stats else None
    out = _SQPSolve.apply(solver, xs, ref, Qd, Rd, xu0, box)
    return (out, box["stats"]) if return_stats else out


class _EEPositions(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, xu):
        pos, jac = solver.ee_positions(xu.detach().contiguous().float(), jacobian=True)
        ctx.solver, ctx.jac = solver, jac
        return pos

    @staticmethod
    def backward(ctx, grad_pos):
        s = ctx.solver
        g = (grad_pos.float().unsqueeze(-1) * ctx.jac).sum(2)  # [B, N, nq] = sum_i dL/dp_i dp_i/dq
        out = torch.zeros(s.B, s.N * (s.nx + s.nu), dtype=torch.float32, device=grad_pos.device)
        out.view(s.B, s.N, s.nx + s.nu)[:, :, :s.nq] = g
        return None, out[:, :s.traj]


def ee_positions(solver, xu):
    """Differentiable end-effector positions [B, N, 3] of trajectories xu [B, traj] (gato_ee_positions)."""
    return _EEPositions.apply(solver, xu)
