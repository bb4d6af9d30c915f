"""ctypes front-end of the CPU FP64 oracle (``oracle/gato_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg, never by the product package.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libgato_oracle.so"

PLANTS = {"iiwa14": 0, "indy7": 1, "indy7_urdf": 2}  # indy7_urdf: the simulated plant (URDF inertial origins kept)
EXIT_PCG0, EXIT_KKT = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Cfg(C.Structure):
    _fields_ = [("plant", C.c_int), ("N", C.c_int), ("dt", C.c_double), ("max_sqp_iters", C.c_int),
                ("kkt_tol", C.c_double), ("max_pcg_iters", C.c_int), ("solve_ratio", C.c_double),
                ("q_cost", C.c_double), ("qd_cost", C.c_double), ("u_cost", C.c_double), ("N_cost", C.c_double),
                ("q_lim_cost", C.c_double), ("vel_lim_cost", C.c_double), ("ctrl_lim_cost", C.c_double),
                ("adapt_rho", C.c_int), ("exit_rule", C.c_int), ("terminal_mode", C.c_int),
                ("hess_mode", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("sqp_iters", _ip), ("kkt_converged", _ip), ("initial_merit", _dp), ("final_merit", _dp),
                ("pcg_iters", _ip), ("ls_min_merit", _dp), ("ls_step", _dp), ("kkt_res", _dp),
                ("n_iters", C.c_int), ("n_ls_iters", C.c_int)]


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def use_native():
    """Build (make native: -march=native for this host) and load the oracle tuned for the machine it runs on, for
    the CPU baseline's timings; returns the library path."""
    import subprocess
    global _lib, LIB_PATH
    subprocess.run(["make", "-s", "-C", str(HERE), "native"], check=True, capture_output=True)
    LIB_PATH = HERE / "_build_native" / "libgato_oracle.so"
    _lib = None
    lib()
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _lib.orc_merit.restype = C.c_double
        _lib.orc_pcg.restype = C.c_int
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _ip_(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def nq(plant):
    return lib().orc_nq(PLANTS[plant])


def limits(plant):
    n = nq(plant)
    arrs = [np.zeros(n) for _ in range(6)]
    lib().orc_limits(PLANTS[plant], *[_p(a) for a in arrs])
    return arrs


def fk(plant, q):
    q = _f64(q)
    p = np.zeros(3)
    J = np.zeros((nq(plant), 3))  # row j = dp/dq_j
    lib().orc_fk(PLANTS[plant], _p(q), _p(p), _p(J))
    return p, J.T.copy()


def frames(plant, q):
    """World orientation (nq, 3, 3) and origin (nq, 3) of every joint frame."""
    n = nq(plant)
    R = np.zeros((n, 3, 3))
    o = np.zeros((n, 3))
    lib().orc_frames(PLANTS[plant], _p(_f64(q)), _p(R), _p(o))
    return R, o


def rnea(plant, q, qd, qdd, fext=None, gravity=True):
    tau = np.zeros(nq(plant))
    lib().orc_rnea(PLANTS[plant], _p(_f64(q)), _p(_f64(qd)), _p(_f64(qdd)), _p(_f64(fext)), int(gravity), _p(tau))
    return tau


def mass(plant, q):
    n = nq(plant)
    M = np.zeros((n, n))  # column-major fill of a symmetric matrix
    lib().orc_mass(PLANTS[plant], _p(_f64(q)), _p(M))
    return M.T.copy()


def fd(plant, q, qd, u, fext=None):
    qdd = np.zeros(nq(plant))
    lib().orc_fd(PLANTS[plant], _p(_f64(q)), _p(_f64(qd)), _p(_f64(u)), _p(_f64(fext)), _p(qdd))
    return qdd


def fd_grad(plant, q, qd, u, fext=None):
    """Returns qdd, dqdd (nq x 2nq: [d/dq | d/dqd]), Minv."""
    n = nq(plant)
    qdd = np.zeros(n)
    dq = np.zeros((2 * n, n))
    Mi = np.zeros((n, n))
    lib().orc_fd_grad(PLANTS[plant], _p(_f64(q)), _p(_f64(qd)), _p(_f64(u)), _p(_f64(fext)), _p(qdd), _p(dq), _p(Mi))
    return qdd, dq.T.copy(), Mi.T.copy()


def make_cfg(plant="iiwa14", N=32, dt=0.01, max_sqp_iters=5, kkt_tol=1e-4, max_pcg_iters=100, solve_ratio=1.0,
             q_cost=1.0, qd_cost=1e-3, u_cost=1e-6, N_cost=50.0, q_lim_cost=1e-3, vel_lim_cost=0.0,
             ctrl_lim_cost=0.0, adapt_rho=True, exit_rule=EXIT_PCG0, terminal_mode=0, hess_mode=0):
    """Defaults are the C++ ctor defaults of the reference (bsqp.cuh:25-27)."""
    return Cfg(PLANTS[plant], N, dt, max_sqp_iters, kkt_tol, max_pcg_iters, solve_ratio, q_cost, qd_cost, u_cost,
               N_cost, q_lim_cost, vel_lim_cost, ctrl_lim_cost, int(adapt_rho), exit_rule, terminal_mode, hess_mode)


def dims(cfg):
    n = lib().orc_nq(cfg.plant)
    return n, 2 * n, n  # nq, nx, nu


def setup_kkt(cfg, xu, xs, ref, fext=None, Qd=None, Rd=None, xref=None):
    _, nx, nu = dims(cfg)
    N = cfg.N
    out = dict(A=np.zeros((N, nx, nx)), B=np.zeros((N, nu, nx)), c=np.zeros((N, nx)), Q=np.zeros((N, nx, nx)),
               R=np.zeros((N, nu, nu)), q=np.zeros((N, nx)), r=np.zeros((N, nu)))
    lib().orc_setup_kkt(C.byref(cfg), _p(_f64(xu)), _p(_f64(xs)), _p(_f64(ref)), _p(_f64(fext)), _p(_f64(Qd)),
                        _p(_f64(Rd)), _p(_f64(xref)), *[_p(out[k]) for k in "A B c Q R q r".split()])
    # column-major blocks -> row-major matrices
    out["A"] = out["A"].transpose(0, 2, 1).copy()
    out["B"] = out["B"].transpose(0, 2, 1).copy()
    out["Q"] = out["Q"].transpose(0, 2, 1).copy()
    out["R"] = out["R"].transpose(0, 2, 1).copy()
    return out


def schur(cfg, rho, kkt):
    """Returns Qi, Ri, Sd, So, Pd, Po, gamma as row-major numpy blocks."""
    _, nx, nu = dims(cfg)
    N = cfg.N
    cm = lambda a: np.ascontiguousarray(a.transpose(0, 2, 1))
    ins = [cm(kkt["A"]), cm(kkt["B"]), cm(kkt["Q"]), cm(kkt["R"]), _f64(kkt["q"]), _f64(kkt["r"]), _f64(kkt["c"])]
    outs = [np.zeros((N, nx, nx)), np.zeros((N, nu, nu))] + [np.zeros((N, nx, nx)) for _ in range(4)] + [np.zeros((N, nx))]
    lib().orc_schur(C.byref(cfg), C.c_double(rho), *[_p(a) for a in ins], *[_p(a) for a in outs])
    names = "Qi Ri Sd So Pd Po gamma".split()
    res = {}
    for n_, a in zip(names, outs):
        res[n_] = a if n_ == "gamma" else a.transpose(0, 2, 1).copy()
    return res


def pcg(nx, N, sch, lam, eps, max_iters):
    cm = lambda a: np.ascontiguousarray(a.transpose(0, 2, 1))
    lam = _f64(lam).copy()
    it = lib().orc_pcg(nx, N, _p(cm(sch["Sd"])), _p(cm(sch["So"])), _p(cm(sch["Pd"])), _p(cm(sch["Po"])),
                       _p(_f64(sch["gamma"])), _p(lam), C.c_double(eps), max_iters)
    return lam, it


def dz(cfg, sch, kkt, lam):
    _, nx, nu = dims(cfg)
    N = cfg.N
    cm = lambda a: np.ascontiguousarray(a.transpose(0, 2, 1))
    d = np.zeros((nx + nu) * N)
    qres = np.zeros((N, nx))
    rres = np.zeros((N, nu))
    lib().orc_dz(nx, nu, N, _p(cm(sch["Qi"])), _p(cm(sch["Ri"])), _p(cm(kkt["A"])), _p(cm(kkt["B"])),
                 _p(_f64(kkt["q"])), _p(_f64(kkt["r"])), _p(_f64(lam)), _p(d), _p(qres), _p(rres))
    return d[: (nx + nu) * N - nu], qres, rres


def merit(cfg, xu, dz_, alpha, xs, ref, mu, fext=None, Qd=None, Rd=None, xref=None):
    _, nx, nu = dims(cfg)
    d = np.zeros((nx + nu) * cfg.N)
    d[: len(dz_)] = dz_
    return lib().orc_merit(C.byref(cfg), _p(_f64(xu)), _p(d), C.c_double(alpha), _p(_f64(xs)), _p(_f64(ref)),
                           _p(_f64(fext)), _p(_f64(Qd)), _p(_f64(Rd)), _p(_f64(xref)), C.c_double(mu))


def solve(cfg, xu, xs, ref, fext=None, Qd=None, Rd=None, xref=None, lam=None, rho=None, drho=None,
          drho_init=None, mu=10.0, pcg_tol=1e-5, rho0=1e-3, nthreads=0):
    """Batched SQP solve (BSQP::solve). Arrays are (B, ...) float64; returns a dict like the
    reference's result plus the updated persistent state (lambda, rho, drho)."""
    _, nx, nu = dims(cfg)
    N = cfg.N
    xu = np.array(xu, dtype=np.float64, copy=True, order="C")
    B = xu.shape[0]
    full = lambda v, w=1: np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (B,) if w == 1 else (B, w)))
    lam = np.zeros((B, N * nx)) if lam is None else _f64(lam).copy()
    rho = full(rho0) .copy() if rho is None else _f64(rho).copy()
    drho = np.ones(B) if drho is None else _f64(drho).copy()
    drho_init = np.ones(B) if drho_init is None else _f64(drho_init)
    mu = full(mu)
    tol = full(pcg_tol)
    it = cfg.max_sqp_iters
    out = dict(sqp_iters=np.zeros(B, np.int32), kkt_converged=np.zeros(B, np.int32), initial_merit=np.zeros(B),
               final_merit=np.zeros(B), pcg_iters=np.zeros((it, B), np.int32), ls_min_merit=np.zeros((it, B)),
               ls_step=np.zeros((it, B)), kkt_res=np.zeros((it, B)))
    st = Stats(_ip_(out["sqp_iters"]), _ip_(out["kkt_converged"]), _p(out["initial_merit"]), _p(out["final_merit"]),
               _ip_(out["pcg_iters"]), _p(out["ls_min_merit"]), _p(out["ls_step"]), _p(out["kkt_res"]), 0, 0)
    lib().orc_solve(C.byref(cfg), B, _p(xu), _p(_f64(xs)), _p(_f64(ref)), _p(_f64(fext)), _p(_f64(Qd)), _p(_f64(Rd)),
                    _p(_f64(xref)), _p(lam), _p(rho), _p(drho), _p(drho_init), _p(mu), _p(tol), C.byref(st),
                    int(nthreads))
    out.update(XU=xu, lam=lam, rho=rho, drho=drho, n_iters=st.n_iters, ls_num_iters=st.n_ls_iters)
    return out


def sim_forward(plant, x, u, fext, dt):
    fext = _f64(fext)
    B = fext.shape[0]
    n = nq(plant)
    out = np.zeros((B, 2 * n))
    lib().orc_sim_forward(PLANTS[plant], B, _p(_f64(x)), _p(_f64(u)), _p(fext), C.c_double(dt), _p(out))
    return out


# ----------------------------------------------------------------------------- adjoint of the SQP step
# The backward pass differentiates the SQP step map z+ = z + dz(z; theta) at a fixed linearisation point z,
# where [dz; lambda] solves the step's KKT system K [dz; lambda] = [-g; c] (G~ = blkdiag(Q~_k, R_k) with the
# half-diagonal rho of linalg.cuh:85-96, C from the integrator Jacobians), for theta = (x_s, reference, diag Q,
# diag R). At a converged solution dz = 0 and this is the implicit derivative of the solution under the
# solver's own Hessian model. (No counterpart in the reference: the adjoint for differentiable MPC.)


def dense_kkt(kkt, rho, nq):
    """G~ (with rho on the position diagonal), g, C and c of one problem's SQP step, dense."""
    N, nx, nu = kkt["q"].shape[0], kkt["q"].shape[1], kkt["r"].shape[1]
    nz = (nx + nu) * N - nu
    off = lambda k: k * (nx + nu)
    G = np.zeros((nz, nz))
    g = np.zeros(nz)
    for k in range(N):
        Qt = kkt["Q"][k].copy()
        Qt[np.arange(nq), np.arange(nq)] += rho
        G[off(k):off(k) + nx, off(k):off(k) + nx] = Qt
        g[off(k):off(k) + nx] = kkt["q"][k]
        if k < N - 1:
            G[off(k) + nx:off(k) + nx + nu, off(k) + nx:off(k) + nx + nu] = kkt["R"][k]
            g[off(k) + nx:off(k) + nx + nu] = kkt["r"][k]
    Cm = np.zeros((N * nx, nz))
    Cm[:nx, :nx] = -np.eye(nx)
    for k in range(N - 1):
        r0 = (k + 1) * nx
        Cm[r0:r0 + nx, off(k):off(k) + nx] = kkt["A"][k]
        Cm[r0:r0 + nx, off(k) + nx:off(k) + nx + nu] = kkt["B"][k]
        Cm[r0:r0 + nx, off(k + 1):off(k + 1) + nx] = -np.eye(nx)
    return G, g, Cm, kkt["c"].reshape(-1)


def _plant_name(code):
    return {v: k for k, v in PLANTS.items()}[code]


def _joint_weights(cfg, Qd, Rd):
    nq_, nx, nu = dims(cfg)
    Qw = np.r_[np.zeros(nq_), np.full(nq_, cfg.qd_cost)] if Qd is None else _f64(Qd)
    Rw = np.full(nu, cfg.u_cost) if Rd is None else _f64(Rd)
    return Qw, Rw


class _KKTSolver:
    """Solves with one KKT matrix: LAPACK dense LU for small systems, a sparse LU (the matrix is block-banded) from
    2000 unknowns on, factorised once for all right-hand sides. Both are direct FP64 solves."""

    def __init__(self, K):
        self.K = K
        self.lu = None
        if K.shape[0] >= 2000:
            from scipy.sparse import csc_matrix
            from scipy.sparse.linalg import splu
            self.lu = splu(csc_matrix(K))

    def solve(self, rhs):
        return self.lu.solve(rhs) if self.lu is not None else np.linalg.solve(self.K, rhs)


def step_map(cfg, xu, xs, ref, rho, fext=None, Qd=None, Rd=None, xref=None, _solver=None):
    """dz and lambda of one SQP step at xu by a direct FP64 solve of the step's KKT system."""
    nq_, nx, nu = dims(cfg)
    Qw, Rw = _joint_weights(cfg, Qd, Rd)
    kkt = setup_kkt(cfg, xu, xs, ref, fext, Qw, Rw, xref)
    G, g, Cm, c = dense_kkt(kkt, rho, nq_)
    nz, nc = G.shape[0], Cm.shape[0]
    K = np.block([[G, Cm.T], [Cm, np.zeros((nc, nc))]])
    ks = _KKTSolver(K)
    if _solver is not None:
        _solver.append(ks)
    w = ks.solve(np.concatenate([-g, c]))
    return w[:nz], w[nz:], K


def step_grad(cfg, xu, xs, ref, rho, dL_dz, fext=None, Qd=None, Rd=None, xref=None):
    """Gradient of L(xu + dz(theta)) w.r.t. x_s [nx], reference [N, 6], diag Q [nx] and diag R [nu], given
    dL/dz [traj], for one problem. Adjoint: K v = [dL/dz; 0]; dL/dtheta = v^T (d[-g; c]/dtheta - dK/dtheta w)."""
    nq_, nx, nu = dims(cfg)
    N = cfg.N
    Qw, Rw = _joint_weights(cfg, Qd, Rd)
    xu = _f64(xu)
    ks = []
    dz_, lam, K = step_map(cfg, xu, xs, ref, rho, fext, Qw, Rw, xref, _solver=ks)
    nz = dz_.shape[0]
    v = ks[0].solve(np.concatenate([_f64(dL_dz), np.zeros(N * nx)]))
    vz, vl = v[:nz], v[nz:]
    xr = np.zeros((N, nx)) if xref is None else _f64(xref).reshape(N, nx)
    ref = _f64(ref).reshape(N, 6)
    st = nx + nu
    gxs = -vl[:nx]
    gQ = np.zeros(nx)
    gR = np.zeros(nu)
    gref = np.zeros((N, 6))
    for k in range(N):
        term = k == N - 1
        src = (N - 1 if cfg.terminal_mode else N - 2) if term else k
        xsrc = xu[src * st:src * st + nx]
        vx, dx = vz[k * st:k * st + nx], dz_[k * st:k * st + nx]
        gQ += -vx * (xsrc - xr[k]) - vx * dx
        if not term:
            u = xu[k * st + nx:k * st + nx + nu]
            vu, du = vz[k * st + nx:k * st + nx + nu], dz_[k * st + nx:k * st + nx + nu]
            gR += -vu * u - vu * du
        w = (cfg.N_cost if cfg.terminal_mode else cfg.q_cost) if term else cfg.q_cost
        p, J = fk(_plant_name(cfg.plant), xsrc[:nq_])
        gvec = J.T @ (p - ref[k, :3])
        vq, dq = vx[:nq_], dx[:nq_]
        for j in range(3):
            a = J[j]
            val = w * (vq @ a)
            if cfg.hess_mode == 0:
                val += w * ((vq @ a) * (gvec @ dq) + (vq @ gvec) * (a @ dq))
            gref[k, j] = val
    return dict(xs=gxs, ref=gref, Qd=gQ, Rd=gR, dz=dz_, lam=lam, v=vz, v_lam=vl)


# ----------------------------------------------------------------------------- derivative of the solution
# At a solution the step vanishes: F(z, lambda; theta) = [g(z) + C(z)^T lambda; c(z)] = 0, with g the cost gradient
# of the KKT set-up (including its terminal-block convention) and C the constraint Jacobian. The solution's
# derivative is then the implicit one, d[z; lambda]/dtheta = -J^-1 dF/dtheta with J = dF/d(z, lambda), the exact
# Jacobian: its upper-left block is the Hessian of the Lagrangian (cost curvature and the multipliers times the
# dynamics' second derivatives), not the SQP model G~ the step uses. (No counterpart in the reference.)


def kkt_function(cfg, xu, lam, xs, ref, fext=None, Qd=None, Rd=None, xref=None):
    """F(z, lambda) = [g + C^T lambda; c] of one problem (dense FP64), and C."""
    nq_, nx, nu = dims(cfg)
    Qw, Rw = _joint_weights(cfg, Qd, Rd)
    kkt = setup_kkt(cfg, _f64(xu), xs, ref, fext, Qw, Rw, xref)
    _, g, Cm, c = dense_kkt(kkt, 0.0, nq_)
    return np.concatenate([g + Cm.T @ _f64(lam), c]), Cm


def fixed_point(cfg, xu, xs, ref, rho=1e-3, fext=None, Qd=None, Rd=None, xref=None, iters=60, tol=1e-12):
    """Full SQP steps z <- z + dz from xu (no line search) until |dz|_inf <= tol; returns (z, lambda, |dz|_inf)."""
    z = _f64(xu).copy()
    for _ in range(iters):
        dz_, lam, _ = step_map(cfg, z, xs, ref, rho, fext, Qd, Rd, xref)
        z += dz_
        if np.abs(dz_).max() <= tol:
            break
    dz_, lam, _ = step_map(cfg, z, xs, ref, rho, fext, Qd, Rd, xref)
    return z, lam, float(np.abs(dz_).max())


def solution_grad(cfg, xu, lam, xs, ref, dL_dz, fext=None, Qd=None, Rd=None, xref=None, h=1e-6):
    """Gradient of L(z*(theta)) w.r.t. x_s [nx], reference [N, 6], diag Q [nx] and diag R [nu] at a solution (xu,
    lam) of one problem, given dL/dz [traj]: w = J^-1 [dL/dz; 0] with J built column by column from central
    differences (step h) of the analytic F along z and exactly in lambda; then dL/dtheta = -w^T dF/dtheta, the
    same parameter terms as step_grad with the step dz = 0."""
    nq_, nx, nu = dims(cfg)
    N = cfg.N
    Qw, Rw = _joint_weights(cfg, Qd, Rd)
    xu = _f64(xu)
    nz = xu.shape[0]
    F0, Cm = kkt_function(cfg, xu, lam, xs, ref, fext, Qw, Rw, xref)
    nc = Cm.shape[0]
    J = np.zeros((nz + nc, nz + nc))
    for j in range(nz):
        e = np.zeros(nz)
        e[j] = h
        Fp, _ = kkt_function(cfg, xu + e, lam, xs, ref, fext, Qw, Rw, xref)
        Fm, _ = kkt_function(cfg, xu - e, lam, xs, ref, fext, Qw, Rw, xref)
        J[:nz, j] = (Fp[:nz] - Fm[:nz]) / (2 * h)
    J[:nz, nz:] = Cm.T
    J[nz:, :nz] = Cm
    w = np.linalg.solve(J, np.concatenate([_f64(dL_dz), np.zeros(nc)]))
    vz, vl = w[:nz], w[nz:]
    xr = np.zeros((N, nx)) if xref is None else _f64(xref).reshape(N, nx)
    st = nx + nu
    gQ, gR, gref = np.zeros(nx), np.zeros(nu), np.zeros((N, 6))
    for k in range(N):
        term = k == N - 1
        src = (N - 1 if cfg.terminal_mode else N - 2) if term else k
        xsrc = xu[src * st:src * st + nx]
        gQ += -vz[k * st:k * st + nx] * (xsrc - xr[k])
        if not term:
            gR += -vz[k * st + nx:k * st + nx + nu] * xu[k * st + nx:k * st + nx + nu]
        w_k = (cfg.N_cost if cfg.terminal_mode else cfg.q_cost) if term else cfg.q_cost
        _, Jp = fk(_plant_name(cfg.plant), xsrc[:nq_])
        gref[k, :3] = w_k * (Jp @ vz[k * st:k * st + nq_])
    return dict(xs=-vl[:nx], ref=gref, Qd=gQ, Rd=gR, v=vz, v_lam=vl, F_inf=float(np.abs(F0).max()))
