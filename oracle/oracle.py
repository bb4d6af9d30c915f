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

PLANTS = {"iiwa14": 0, "indy7": 1}
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
