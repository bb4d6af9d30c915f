"""Writes tests/golden/oracle_fixtures.json: seeded inputs and FP64 oracle outputs for the small
configurations the parity tests use (run from the repository root after building the oracle):

    python tests/golden/make_oracle_fixtures.py

Contents: the plain SQP case (iiwa14, N=8, B=1, joint sinusoid, Q = I, R = 1e-3 I) after one and after five
SQP iterations; a figure-8 KKT/merit point per plant (N=8); dynamics values (FK, mass matrix, forward
dynamics) at seeded states of both plants. Inputs are rounded to fp32 first, as the GPU sees them.
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from gato_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def joint_case(iters):
    wl = W.joint_sinusoid("iiwa14", N=8, B=1, seed=0)
    p = wl.params
    cfg = O.make_cfg("iiwa14", N=8, dt=p["dt"], max_sqp_iters=iters, max_pcg_iters=p["max_pcg_iters"],
                     q_cost=p["q_cost"], qd_cost=p["qd_cost"], u_cost=p["u_cost"], N_cost=p["N_cost"],
                     q_lim_cost=p["q_lim_cost"])
    r = O.solve(cfg, f32(wl.xu), f32(wl.xs), f32(wl.ref), Qd=f32(wl.Qd), Rd=f32(wl.Rd), xref=f32(wl.xref),
                mu=p["mu"], pcg_tol=p["pcg_tol"], rho0=p["rho"])
    return {"iters": iters, "xu": f32(wl.xu).tolist(), "xs": f32(wl.xs).tolist(), "ref": f32(wl.ref).tolist(),
            "Qd": f32(wl.Qd).tolist(), "Rd": f32(wl.Rd).tolist(), "xref": f32(wl.xref).tolist(),
            "params": p, "XU": r["XU"].tolist(), "sqp_iters": r["sqp_iters"].tolist(),
            "pcg_iters": r["pcg_iters"].tolist(), "ls_step": r["ls_step"].tolist(),
            "initial_merit": r["initial_merit"].tolist(), "final_merit": r["final_merit"].tolist(),
            "lam": r["lam"].tolist()}


def kkt_case(plant):
    wl = W.drop_in_figure8(plant, N=8, B=1, seed=40)
    rng = np.random.default_rng(41)
    xu = f32(wl.xu[0] + rng.normal(0, 0.05, wl.xu.shape[1]))
    p = wl.params
    cfg = O.make_cfg(plant, N=8, dt=p["dt"], q_cost=p["q_cost"], qd_cost=p["qd_cost"], u_cost=p["u_cost"],
                     N_cost=p["N_cost"], q_lim_cost=p["q_lim_cost"])
    k = O.setup_kkt(cfg, xu, f32(wl.xs[0]), f32(wl.ref[0]))
    m = O.merit(cfg, xu, np.zeros_like(xu), 0.0, f32(wl.xs[0]), f32(wl.ref[0]), p["mu"])
    return {"plant": plant, "xu": xu.tolist(), "xs": f32(wl.xs[0]).tolist(), "ref": f32(wl.ref[0]).tolist(),
            "params": p, "A": k["A"].tolist(), "B": k["B"].tolist(), "c": k["c"].tolist(), "Q": k["Q"].tolist(),
            "q": k["q"].tolist(), "R": k["R"].tolist(), "r": k["r"].tolist(), "merit": m}


def dynamics_case(plant):
    rng = np.random.default_rng(42 if plant == "iiwa14" else 43)
    n = W.NQ[plant]
    out = []
    for _ in range(4):
        q, qd, u = f32(rng.uniform(-1, 1, n)), f32(rng.uniform(-1, 1, n)), f32(rng.normal(0, 5, n))
        p, J = O.fk(plant, q)
        out.append({"q": q.tolist(), "qd": qd.tolist(), "u": u.tolist(), "p": p.tolist(), "J": J.tolist(),
                    "M": O.mass(plant, q).tolist(), "qdd": O.fd(plant, q, qd, u).tolist()})
    return out


def main():
    data = {"joint_iter1": joint_case(1), "joint_full": joint_case(5),
            "kkt": {pl: kkt_case(pl) for pl in ("iiwa14", "indy7")},
            "dynamics": {pl: dynamics_case(pl) for pl in ("iiwa14", "indy7")}}
    path = Path(__file__).resolve().parent / "oracle_fixtures.json"
    path.write_text(json.dumps(data))
    print(f"wrote {path} ({path.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
