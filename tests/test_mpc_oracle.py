"""Pins of the MPC-caller restatements (oracle/mpc_oracle.py, gato_amd.workloads.reference_figure8_path).

Each check comes from the mathematics or from values the reference's definitions fix by hand, not from the
restated formulas themselves: power pairing and the group law of rigid-body actions for the wrench transform,
closed-form steps of the estimator's update, and exact points of the figure-8 path.
"""
import numpy as np
import pytest

from gato_amd import workloads as W


def _hom(R, p):
    H = np.eye(4)
    H[:3, :3], H[:3, 3] = R, p
    return H


def _twist_local(H, v, w):
    """Twist (linear v at the origin, angular w) seen in the frame of placement H: H^-1 [w]x..H (matrix route)."""
    X = np.zeros((4, 4))
    X[:3, :3] = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    X[:3, 3] = v
    Y = np.linalg.inv(H) @ X @ H
    return Y[:3, 3], np.array([Y[2, 1], Y[0, 2], Y[1, 0]])


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_frames_end_at_the_forward_kinematics_point(O, plant):
    rng = np.random.default_rng(0)
    for _ in range(5):
        q = rng.uniform(-1, 1, O.nq(plant))
        R, o = O.frames(plant, q)
        np.testing.assert_allclose(o[-1], O.fk(plant, q)[0], atol=1e-12)
        for Ri in R:
            np.testing.assert_allclose(Ri @ Ri.T, np.eye(3), atol=1e-12)


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_wrench_transform_pairs_with_twists_and_composes(O, plant):
    from oracle import mpc_oracle as M

    rng = np.random.default_rng(1)
    for _ in range(10):
        q = rng.uniform(-1.5, 1.5, O.nq(plant))
        w = rng.normal(size=6)
        out = M.wrench_to_solver_frame(plant, q, w)
        R, o = O.frames(plant, q)
        H1, H0 = _hom(R[-1], o[-1]), _hom(R[-2], o[-2])
        Hrel = np.linalg.inv(H0) @ H1
        # virtual power is frame independent: the twist follows the same two placements
        v, om = rng.normal(size=3), rng.normal(size=3)
        v1, om1 = _twist_local(H1, v, om)
        v2, om2 = _twist_local(Hrel, v1, om1)
        assert abs(out[:3] @ v2 + out[3:] @ om2 - (w[:3] @ v + w[3:] @ om)) < 1e-10
        np.testing.assert_allclose(np.linalg.norm(out[:3]), np.linalg.norm(w[:3]), rtol=1e-12)
        # group law: two inverse actions = one inverse action of the composed placement H1 Hrel
        Hc = H1 @ Hrel
        Rc, pc = Hc[:3, :3], Hc[:3, 3]
        np.testing.assert_allclose(out[:3], Rc.T @ w[:3], atol=1e-12)
        np.testing.assert_allclose(out[3:], Rc.T @ (w[3:] - np.cross(pc, w[:3])), atol=1e-12)


def test_estimator_initial_batch_and_directions():
    from oracle import mpc_oracle as M

    est = M.EstimatorRef(16, initial_radius=5.0, min_radius=2.0, max_radius=20.0, smoothing_factor=0.5)
    d = est.dirs
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, rtol=1e-6)
    np.testing.assert_allclose(d[0], [0, 1, 0], atol=1e-7)   # spiral starts at the north pole
    np.testing.assert_allclose(d[-1], [0, -1, 0], atol=1e-6)  # and ends at the south pole
    b = est.batch()
    assert not b[:3].any()
    np.testing.assert_allclose(b[3:, :3], 5.0 * d, rtol=1e-6)
    assert not b[3:, 3:].any()


def test_estimator_update_closed_forms():
    from oracle import mpc_oracle as M

    r0, a, be = 5.0, 0.6, 0.5
    est = M.EstimatorRef(8, initial_radius=r0, min_radius=0.1, max_radius=20.0, smoothing_factor=0.5)
    u = (0.0, 0.0, 0.0)
    # the zero hypothesis wins from the zero state: the estimate stays at zero, the radius shrinks
    est.update(1, np.ones(8), a, be, u)
    assert not est.est.any() and not est.mom.any()
    assert est.radius == pytest.approx(r0 * 0.95) and est.confidence == pytest.approx(0.05)
    # u = 0 gives the half turn about y
    np.testing.assert_allclose(est.R, np.diag([-1.0, 1.0, -1.0]), atol=1e-7)
    # an exploration hypothesis wins: momentum (1-beta) f, estimate 0.2 (alpha f + 0.5 momentum)
    f = est.batch()[3].astype(np.float64)
    est.update(3, np.ones(8), a, be, u)
    np.testing.assert_allclose(est.mom, (1 - be) * f, rtol=1e-6)
    np.testing.assert_allclose(est.est, 0.2 * (a * f + 0.5 * (1 - be) * f), rtol=1e-6)
    np.testing.assert_allclose(est.smooth, 0.5 * est.est, rtol=1e-6)
    assert est.radius == pytest.approx(r0 * 0.95 * 1.05) and est.confidence == 0.0


def test_estimator_history_rules():
    from oracle import mpc_oracle as M

    u = (0.3, 0.2, 0.7)
    est = M.EstimatorRef(8, initial_radius=10.0, min_radius=0.1, max_radius=50.0)
    for _ in range(6):  # flat history: the sixth update also applies the 0.9 factor
        est.update(0, np.full(8, 2.0), 0.5, 0.8, u)
    assert est.radius == pytest.approx(10.0 * 0.95 ** 6 * 0.9)
    est2 = M.EstimatorRef(8, initial_radius=10.0, min_radius=0.1, max_radius=50.0)
    errs = [1.0, 1.2, 0.9, 1.1, 1.0, 3.0]  # a spike: 3 > 1.5 * mean of the four before it
    for e in errs:
        est2.update(0, np.full(8, e), 0.5, 0.8, u)
    assert est2.radius == pytest.approx(10.0 * 0.95 ** 6 * 1.3)
    assert est2.confidence == pytest.approx(0.3 * 0.5)
    R = est2.R.astype(np.float64)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
    assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-6)


def test_figure8_path_points():
    ax, az, off, th = 0.4, 0.3, (0.1, 0.5, 0.6), np.pi / 4
    p = W.reference_figure8_path(1.0, A_x=ax, A_z=az, offset=off, period=5, cycles=2, theta=th).reshape(-1, 6)
    assert p.shape == (10, 6) and not p[:, 3:].any()
    c, s = np.cos(th), np.sin(th)
    rot = lambda x, y, z: np.array([c * x - s * y, s * x + c * y, z])
    # five points per cycle at t = 0, pi/2, pi, 3pi/2, 2pi
    np.testing.assert_allclose(p[0, :3], rot(off[0], off[1], off[2] + az / 2), atol=1e-12)
    np.testing.assert_allclose(p[1, :3], rot(off[0] + ax, off[1], off[2] + az / 2), atol=1e-12)
    np.testing.assert_allclose(p[3, :3], rot(off[0] - ax, off[1], off[2] + az / 2), atol=1e-12)
    np.testing.assert_allclose(p[4], p[0], atol=1e-12)
    np.testing.assert_array_equal(p[5:], p[:5])
