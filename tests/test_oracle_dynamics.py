"""Pins of the oracle's rigid-body dynamics (oracle/gato_oracle.c) by closed forms, identities,
finite differences and the reference's own generated model tables.

Each test names the mistake it is built to catch.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, rel_err

PLANTS = ["iiwa14", "indy7", "indy7_urdf"]  # indy7_urdf: the simulated plant (inertial origins kept)
GOLDEN = Path(__file__).resolve().parent / "golden"


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.2, 1.2, n), rng.normal(0, 1.0, n), rng.normal(0, 3.0, n), rng.normal(0, 2.0, 6)


def table(plant):
    """(m, h, Ibar) per joint from the generated header, parsed as text."""
    import re

    src = (ROOT / "gato_amd/csrc/model_tables.h").read_text()
    block = re.search(r"#define GATO_%s_JOINTS_INIT \{ \\\n(.*?)\n\}" % plant.upper(), src, re.S).group(1)
    rows = []
    for r in re.findall(r"\{([^}]*)\}", block):
        v = [float(x) for x in r.split(",")]
        rows.append(v)
    return rows


def test_fk_zero_configuration_closed_form(O):
    # iiwa14 at q = 0 stands upright: the end effector (origin of the last joint frame, no tool
    # offset) sits at the sum of the URDF joint offsets along z (iiwa14.urdf:71-288):
    # 0.1575 + 0.2025 + 0.2045 + 0.2155 + 0.1845 + 0.2155 + 0.081 = 1.261, the +-0.0607 y offsets cancel.
    p, _ = O.fk("iiwa14", np.zeros(7))
    np.testing.assert_allclose(p, [0.0, 0.0, 1.261], atol=1e-12)
    _, o = O.frames("iiwa14", np.zeros(7))
    np.testing.assert_allclose(o[:, 2], np.cumsum([0.1575, 0.2025, 0.2045, 0.2155, 0.1845, 0.2155, 0.081]), atol=1e-12)


@pytest.mark.parametrize("plant", PLANTS)
def test_base_joint_rotates_about_world_z(O, plant):
    # Joint 0 of both arms is a world-z axis through the origin: turning it by th rotates the EE
    # position by Rz(th) (catches a transposed joint rotation or a wrong rpy convention).
    n = O.nq(plant)
    q = rand_state(n, 1)[0]
    p0, _ = O.fk(plant, q)
    th = 0.7
    q2 = q.copy()
    q2[0] += th
    p1, _ = O.fk(plant, q2)
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    np.testing.assert_allclose(p1, Rz @ p0, atol=1e-12)


@pytest.mark.parametrize("plant", PLANTS)
def test_frames_are_rotations(O, plant):
    R, _ = O.frames(plant, rand_state(O.nq(plant), 2)[0])
    for Ri in R:
        np.testing.assert_allclose(Ri @ Ri.T, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(Ri) - 1) < 1e-12


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_inertias_equal_reference_generated_tables(plant):
    # GRiD's spatial inertias (the reference's model, iiwa14_grid.cuh h_XImats[252..503],
    # indy7_grid.cuh h_XImats[216..431]) vs the URDF-derived tables: catches a wrong parallel-axis
    # term, an unmerged fixed link (iiwa link 7 + contact = 0.41132 kg) or a sign in h x.
    grid = json.loads((GOLDEN / "grid_inertia.json").read_text())[plant]
    for i, row in enumerate(table(plant)):
        m, h, I = row[12], row[13:16], row[16:22]
        Ib = np.array([[I[0], I[3], I[4]], [I[3], I[1], I[5]], [I[4], I[5], I[2]]])
        hx = np.array([[0, -h[2], h[1]], [h[2], 0, -h[0]], [-h[1], h[0], 0]])
        mine = np.block([[Ib, hx], [hx.T, m * np.eye(3)]])
        ref = np.array(grid[i]).reshape(6, 6).T
        np.testing.assert_allclose(mine, ref, atol=1e-12, err_msg=f"{plant} link {i}")


def test_urdf_plant_is_the_solver_model_with_the_link_centres_of_mass():
    # The simulated indy7 (pinocchio reads the URDF's <inertial><origin>, common.py:49-91) differs from the solver's
    # GRiD model, which places every centre of mass at its joint origin, only in the first moments and the
    # parallel-axis terms: h = m c and Ibar = Ibar_solver + m (|c|^2 1 - c c^T) with c the URDF centre of mass in the
    # joint frame. The COMs are read from indy7.urdf as data (links 1-6, rpy 0 on every inertial origin, each link
    # frame is its joint frame): catches an inertial origin dropped, taken in the wrong frame or with a sign error.
    com = [(-0.00023749, -0.04310313, 0.13245396), (-0.29616699, 2.254e-05, 0.04483069),
           (-0.16804016, 0.00021421, -0.07000383), (-0.00026847, -0.0709844, 0.07649128),
           (-0.09796232, -0.00023114, 0.06445892), (8.147e-05, -0.00046556, 0.03079097)]
    solver, urdf = table("indy7"), table("indy7_urdf")
    assert len(solver) == len(urdf) == 6
    for i, (a, b) in enumerate(zip(solver, urdf)):
        np.testing.assert_array_equal(a[:13], b[:13])  # joint transforms and masses unchanged
        np.testing.assert_array_equal(a[22:], b[22:])  # limits unchanged
        m, c = a[12], np.array(com[i])
        assert np.all(np.array(a[13:16]) == 0)
        np.testing.assert_allclose(b[13:16], m * c, rtol=1e-12, atol=1e-15, err_msg=f"h, link {i + 1}")
        pa = m * (c @ c * np.eye(3) - np.outer(c, c))
        want = np.array(a[16:22]) + np.array([pa[0, 0], pa[1, 1], pa[2, 2], pa[0, 1], pa[0, 2], pa[1, 2]])
        np.testing.assert_allclose(b[16:22], want, rtol=1e-12, atol=1e-15, err_msg=f"Ibar, link {i + 1}")


@pytest.mark.parametrize("plant", PLANTS)
def test_mass_matrix_spd_and_linear_in_qdd(O, plant):
    n = O.nq(plant)
    q, qd, qdd, _ = rand_state(n, 3)
    M = O.mass(plant, q)
    np.testing.assert_allclose(M, M.T, atol=1e-14)
    assert np.linalg.eigvalsh(M).min() > 0
    # ID(q, qd, qdd) - ID(q, qd, 0) = M qdd (catches a velocity-dependent term leaking into M)
    d = O.rnea(plant, q, qd, qdd) - O.rnea(plant, q, qd, np.zeros(n))
    np.testing.assert_allclose(d, M @ qdd, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("plant", PLANTS)
def test_forward_dynamics_inverts_inverse_dynamics(O, plant):
    n = O.nq(plant)
    q, qd, u, f = rand_state(n, 4)
    qdd = O.fd(plant, q, qd, u, f)
    np.testing.assert_allclose(O.rnea(plant, q, qd, qdd, f), u, atol=1e-10)


@pytest.mark.parametrize("plant", PLANTS)
def test_power_balance(O, plant):
    # qd^T (tau - g) = qd^T M qdd + 1/2 qd^T Mdot qd (skew-symmetry of Mdot - 2C); catches wrong
    # Coriolis/centrifugal terms (a sign or a missing v x* Iv) that the identities above miss.
    n = O.nq(plant)
    q, qd, qdd, _ = rand_state(n, 5)
    tau = O.rnea(plant, q, qd, qdd)
    g = O.rnea(plant, q, np.zeros(n), np.zeros(n))
    h = 1e-6
    Mdot = (O.mass(plant, q + h * qd) - O.mass(plant, q - h * qd)) / (2 * h)
    lhs = qd @ (tau - g)
    rhs = qd @ O.mass(plant, q) @ qdd + 0.5 * qd @ Mdot @ qd
    assert abs(lhs - rhs) < 1e-7 * max(1.0, abs(lhs))


@pytest.mark.parametrize("plant", PLANTS)
def test_gravity_is_gradient_of_potential(O, plant):
    # g(q) = dV/dq with V = sum_i m_i g0 z(com_i); catches a wrong gravity sign/axis or frame error.
    rows = table(plant)
    n = len(rows)

    def V(q):
        R, o = O.frames(plant, q)
        v = 0.0
        for i, r in enumerate(rows):
            m, h = r[12], np.array(r[13:16])
            com = o[i] + R[i] @ (h / m)
            v += m * 9.81 * com[2]
        return v

    q = rand_state(n, 6)[0]
    g = O.rnea(plant, q, np.zeros(n), np.zeros(n))
    eps = 1e-6
    dV = np.array([(V(q + eps * e) - V(q - eps * e)) / (2 * eps) for e in np.eye(n)])
    np.testing.assert_allclose(g, dV, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("plant", PLANTS)
def test_fd_gradient_vs_central_differences(O, plant):
    n = O.nq(plant)
    for seed in range(3):
        q, qd, u, f = rand_state(n, 10 + seed)
        qdd, dq, Mi = O.fd_grad(plant, q, qd, u, f)
        h = 1e-6
        num = np.zeros((n, 2 * n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = h
            num[:, j] = (O.fd(plant, q + e, qd, u, f) - O.fd(plant, q - e, qd, u, f)) / (2 * h)
            num[:, n + j] = (O.fd(plant, q, qd + e, u, f) - O.fd(plant, q, qd - e, u, f)) / (2 * h)
        assert rel_err(dq, num) < 1e-6
        np.testing.assert_allclose(Mi, np.linalg.inv(O.mass(plant, q)), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("plant", PLANTS)
def test_ee_jacobian_vs_central_differences(O, plant):
    n = O.nq(plant)
    q = rand_state(n, 7)[0]
    _, J = O.fk(plant, q)
    h = 1e-6
    num = np.stack([(O.fk(plant, q + h * e)[0] - O.fk(plant, q - h * e)[0]) / (2 * h) for e in np.eye(n)], 1)
    np.testing.assert_allclose(J, num, atol=1e-9)


@pytest.mark.parametrize("plant", PLANTS)
def test_external_wrench_maps_through_jacobian_transpose(O, plant):
    # f_ext is a spatial force [n; f] on the last link in its own frame, subtracted from that link's
    # force (iiwa14_fext.cuh:164-171): tau(f_ext) - tau(0) = -(J_p^T R f + J_w^T R n).
    n = O.nq(plant)
    q, qd, qdd, f = rand_state(n, 8)
    d = O.rnea(plant, q, qd, qdd, f) - O.rnea(plant, q, qd, qdd)
    R, o = O.frames(plant, q)
    _, Jp = O.fk(plant, q)
    Jw = R[:, :, 2].T  # joint axes in world
    Rn = R[-1]
    expect = -(Jp.T @ (Rn @ f[3:]) + Jw.T @ (Rn @ f[:3]))
    np.testing.assert_allclose(d, expect, atol=1e-10)


@pytest.mark.parametrize("plant", PLANTS)
def test_sim_forward_is_trapezoidal_step(O, plant):
    n = O.nq(plant)
    q, qd, u, f = rand_state(n, 9)
    x = np.concatenate([q, qd])
    fb = np.stack([f, np.zeros(6)])
    out = O.sim_forward(plant, x, u, fb, 0.01)
    for b in range(2):
        a = O.fd(plant, q, qd, u, fb[b])
        np.testing.assert_allclose(out[b, n:], qd + 0.01 * a, atol=1e-13)
        np.testing.assert_allclose(out[b, :n], q + 0.01 * qd + 0.5 * 0.01 ** 2 * a, atol=1e-13)
