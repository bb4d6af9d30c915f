"""Seeded synthetic workloads for the batched SQP solve (figure-8 end-effector tracking).

Each generator returns float64 numpy arrays; callers cast to float32 for the GPU path.
The initial trajectory guess follows the reference MPC harness: x_k = x_s for every knot,
u_k = 0 (python/bsqp/mpc_controller.py:170-174), and x_0 = x_s (python/bsqp/interface.py:130).
"""
from dataclasses import dataclass, field

import numpy as np

NQ = {"iiwa14": 7, "indy7": 6}

# Figure-8 centres: end-effector position (origin of the last joint frame) at the nominal
# configurations below; the unit tests recompute them with the oracle's forward kinematics.
NOMINAL_Q = {
    "iiwa14": np.array([0.0, 0.6, 0.0, -1.2, 0.0, 0.9, 0.0]),
    "indy7": np.array([0.0, -0.5, 1.2, 0.0, 0.8, 0.0]),
}
CENTRE = {
    "iiwa14": np.array([6.61306655e-01, 3.68356634e-08, 5.42530285e-01]),
    "indy7": np.array([-0.17731386, -0.1865, 0.97399077]),
}


def figure8(centre, radius, period, t, phase=0.0):
    """p(t) = c + r [sin(w t + ph), 0, 0.5 sin(2 (w t + ph))], w = 2 pi / period."""
    w = 2 * np.pi / period
    a = w * np.asarray(t) + phase
    return np.stack([centre[0] + radius * np.sin(a), centre[1] + 0.0 * a, centre[2] + 0.5 * radius * np.sin(2 * a)], -1)


@dataclass
class Workload:
    plant: str
    N: int
    dt: float
    xs: np.ndarray            # (B, nx)
    ref: np.ndarray           # (B, 6N)
    xu: np.ndarray            # (B, traj)
    params: dict = field(default_factory=dict)
    xref: np.ndarray = None   # (B, N*nx) joint-space reference or None
    Qd: np.ndarray = None     # (B, nx) diagonal state weights or None
    Rd: np.ndarray = None     # (B, nu) diagonal control weights or None

    @property
    def B(self):
        return self.xs.shape[0]


def initial_guess(xs, N, nu):
    B, nx = xs.shape
    xu = np.zeros((B, (nx + nu) * N - nu))
    for k in range(N):
        xu[:, k * (nx + nu): k * (nx + nu) + nx] = xs
    return xu


def drop_in_figure8(plant="iiwa14", N=32, B=512, seed=1, dt=0.01, radius=0.2, period=4.0):
    """Drop-in plain path workload: every problem tracks the same figure-8 (radius 0.2 m, period 4 s,
    centred at CENTRE[plant]) from x_s = (q_nom + N(0, 0.05^2), N(0, 0.05^2))."""
    rng = np.random.default_rng(seed)
    nq = NQ[plant]
    q = NOMINAL_Q[plant][None, :] + rng.normal(0.0, 0.05, (B, nq))
    qd = rng.normal(0.0, 0.05, (B, nq))
    xs = np.concatenate([q, qd], 1)
    t = np.arange(N) * dt
    p = figure8(CENTRE[plant], radius, period, t)
    ref = np.zeros((B, N, 6))
    ref[:, :, :3] = p[None]
    params = dict(dt=dt, max_sqp_iters=10, kkt_tol=1e-4, max_pcg_iters=100, pcg_tol=1e-5, solve_ratio=1.0, mu=10.0,
                  q_cost=100.0, qd_cost=1e-3, u_cost=1e-3, N_cost=100.0, q_lim_cost=0.0, vel_lim_cost=0.0,
                  ctrl_lim_cost=0.0, rho=1e-3)
    return Workload(plant, N, dt, xs, ref.reshape(B, 6 * N), initial_guess(xs, N, nq), params)


def joint_sinusoid(plant="iiwa14", N=8, B=1, seed=0, dt=0.01, amp=0.3, freq=1.0):
    """Plain SQP case: x0 = (q ~ U(-1, 1), qd = 0); joint reference q_ref(t) = q0 + amp sin(2 pi f t),
    qd_ref its derivative; cost 1/2 |x - x_ref|_Q^2 + 1/2 |u|_R^2 with Q = I, R = 1e-3 I (the EE term off)."""
    rng = np.random.default_rng(seed)
    nq = NQ[plant]
    nx = 2 * nq
    q0 = rng.uniform(-1.0, 1.0, (B, nq))
    xs = np.concatenate([q0, np.zeros((B, nq))], 1)
    t = np.arange(N) * dt
    w = 2 * np.pi * freq
    qr = q0[:, None, :] + amp * np.sin(w * t)[None, :, None]
    qdr = np.broadcast_to(amp * w * np.cos(w * t)[None, :, None], qr.shape)
    xref = np.concatenate([qr, qdr], 2).reshape(B, N * nx)
    ref = np.zeros((B, 6 * N))
    params = dict(dt=dt, max_sqp_iters=5, kkt_tol=1e-4, max_pcg_iters=100, pcg_tol=1e-5, solve_ratio=1.0, mu=10.0,
                  q_cost=0.0, qd_cost=0.0, u_cost=1e-3, N_cost=0.0, q_lim_cost=0.0, vel_lim_cost=0.0,
                  ctrl_lim_cost=0.0, rho=1e-3)
    Qd = np.ones((B, nx))
    Rd = np.full((B, nq), 1e-3)
    return Workload(plant, N, dt, xs, ref, initial_guess(xs, N, nq), params, xref=xref, Qd=Qd, Rd=Rd)


def mixed_figure8(B=1024, seed=2, dt=0.01, period=4.0, knots=(8, 32, 128)):
    """Uneven batch: B problems, the first half iiwa14 and the second indy7, each with N drawn uniformly
    from `knots`; problem b tracks a figure-8 of radius ~U(0.1, 0.25) m and phase ~U(0, 2 pi) around the
    plant's centre. Returns {(plant, N): Workload} with the drop-in solver settings."""
    rng = np.random.default_rng(seed)
    out = {}
    plants = ["iiwa14"] * (B // 2) + ["indy7"] * (B - B // 2)
    Ns = rng.choice(np.asarray(knots), size=B)
    radius = rng.uniform(0.1, 0.25, B)
    phase = rng.uniform(0.0, 2 * np.pi, B)
    q_noise = rng.normal(0.0, 0.05, (B, 7))
    qd_noise = rng.normal(0.0, 0.05, (B, 7))
    for plant in ("iiwa14", "indy7"):
        nq = NQ[plant]
        for N in knots:
            idx = [b for b in range(B) if plants[b] == plant and Ns[b] == N]
            if not idx:
                continue
            q = NOMINAL_Q[plant][None, :] + q_noise[idx, :nq]
            qd = qd_noise[idx, :nq]
            xs = np.concatenate([q, qd], 1)
            t = np.arange(N) * dt
            ref = np.zeros((len(idx), N, 6))
            for j, b in enumerate(idx):
                ref[j, :, :3] = figure8(CENTRE[plant], radius[b], period, t, phase[b])
            params = dict(dt=dt, max_sqp_iters=10, kkt_tol=1e-4, max_pcg_iters=100, pcg_tol=1e-5, solve_ratio=1.0,
                          mu=10.0, q_cost=100.0, qd_cost=1e-3, u_cost=1e-3, N_cost=100.0, q_lim_cost=0.0,
                          vel_lim_cost=0.0, ctrl_lim_cost=0.0, rho=1e-3)
            out[(plant, int(N))] = Workload(plant, int(N), dt, xs, ref.reshape(len(idx), 6 * N),
                                            initial_guess(xs, int(N), nq), params)
    return out


def differentiation_figure8(N=32, B=256, seed=3, dt=0.01, plant="iiwa14"):
    """Differentiation workload: the drop-in figure-8 problems with per-problem diagonal joint weights Qd [B, nx] and
    Rd [B, nu] drawn log-uniformly from [1e-3, 1e2] (the joint terms 1/2 Qd x^2 + 1/2 Rd u^2 beside the end-effector
    term of weight q_cost = N_cost = 10). The loss of the configuration is sum_b |u_0|^2 + sum_k |p_k - ref_k|^2."""
    wl = drop_in_figure8(plant, N=N, B=B, seed=seed, dt=dt)
    rng = np.random.default_rng(seed)
    nq = NQ[plant]
    wl.Qd = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), (B, 2 * nq)))
    wl.Rd = np.exp(rng.uniform(np.log(1e-3), np.log(1e2), (B, nq)))
    wl.params.update(q_cost=10.0, N_cost=10.0)
    return wl


def scale_figure8(N=128, B=4096, seed=4, dt=0.01, period=4.0, plant="iiwa14"):
    """Scale workload: B problems of horizon N, problem b tracking a figure-8 of seeded radius ~U(0.1, 0.25) m
    and phase ~U(0, 2 pi); weights are set by the caller (the learning loop owns diag Q and diag R)."""
    rng = np.random.default_rng(seed)
    nq = NQ[plant]
    radius = rng.uniform(0.1, 0.25, B)
    phase = rng.uniform(0.0, 2 * np.pi, B)
    q = NOMINAL_Q[plant][None, :] + rng.normal(0.0, 0.05, (B, nq))
    qd = rng.normal(0.0, 0.05, (B, nq))
    xs = np.concatenate([q, qd], 1)
    t = np.arange(N) * dt
    ref = np.zeros((B, N, 6))
    for b in range(B):
        ref[b, :, :3] = figure8(CENTRE[plant], radius[b], period, t, phase[b])
    params = dict(dt=dt, max_sqp_iters=10, kkt_tol=1e-4, max_pcg_iters=100, pcg_tol=1e-5, solve_ratio=1.0, mu=10.0,
                  q_cost=10.0, qd_cost=1e-3, u_cost=1e-3, N_cost=10.0, q_lim_cost=0.0, vel_lim_cost=0.0,
                  ctrl_lim_cost=0.0, rho=1e-2)
    return Workload(plant, N, dt, xs, ref.reshape(B, 6 * N), initial_guess(xs, N, nq), params)


# The reference's published solve-time heatmap: indy7 figure-8 MPC with its default MPC solver settings
# (python/bsqp/config.py:35-50) from the "ready" configuration (config.py:22-26).
REFERENCE_MPC_PARAMS = dict(max_sqp_iters=1, kkt_tol=0.001, max_pcg_iters=200, pcg_tol=1e-4, solve_ratio=1.0,
                            mu=10.0, q_cost=2.0, qd_cost=1e-2, u_cost=2e-6, N_cost=50.0, q_lim_cost=0.01,
                            vel_lim_cost=0.0, ctrl_lim_cost=0.0, rho=0.01)
INDY7_READY = np.array([-1.096711, -0.09903229, 0.83125766, -0.10907673, 0.49704404, 0.01499449])


def reference_figure8_path(dt, A_x=0.4, A_z=0.4, offset=(0.0, 0.5, 0.6), period=6.0, cycles=5, theta=np.pi / 4):
    """End-effector path of the reference's figure-8 benchmark (python/bsqp/common.py:10-46): the curve
    (A_x sin t, 0, A_z sin(2t)/2 + A_z/2) + offset rotated by theta about z, sampled at period/dt points of
    t in [0, 2 pi] and repeated `cycles` times; 6 floats per point (xyz, zero rpy)."""
    t = np.linspace(0.0, 2 * np.pi, int(period / dt))
    p = np.stack([offset[0] + A_x * np.sin(t), np.full_like(t, offset[1]),
                  offset[2] + A_z * np.sin(2 * t) / 2 + A_z / 2], 1)
    c, s_ = np.cos(theta), np.sin(theta)
    R = np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])
    pts = np.zeros((len(t), 6))
    pts[:, :3] = p @ R.T
    return np.tile(pts.reshape(-1), int(cycles))
