"""The kinematic model both sides share (gato_amd/csrc/model_tables.h, read by the oracle and the HIP kernels)
against the reference's own generated tables and its URDFs, at q = 0 and four seeded configurations.

Golden data (tests/golden/grid_kinematics.json, written by tests/golden/make_grid_kinematics_fixture.py):
  * GRiD's 6x6 spatial transforms X[i](q) (gato/dynamics/*/*_grid.cuh, load_update_XImats_helpers) and 4x4
    homogeneous transforms Xhom[i](q) (load_update_XmatsHom_helpers), evaluated at those configurations;
  * the URDF origin (xyz, rpy) of every revolute joint in chain order (examples/*_description/*.urdf).
The oracle's joint frames (orc_frames) and end-effector point (orc_fk, the last joint frame's origin, as the
reference's end_effector_pose/positions use it) must compose to GRiD's chain; its parent-to-child transforms must
equal GRiD's X[i]; and both must equal a chain composed here from the URDF numbers alone. A wrong joint offset,
axis sign, rotation order or a transposed table entry fails one of these.
"""
import json

import numpy as np
import pytest

from conftest import ROOT

FIX = json.loads((ROOT / "tests" / "golden" / "grid_kinematics.json").read_text())
# The URDFs give the quarter turns to 9-10 digits (1.570796327, 1.5707963705062866) and iiwa14 carries a -4.4e-8 rad
# pitch, so the URDF chain agrees with the oracle to ~1e-7. GRiD's homogeneous tables agree with the oracle to 1e-9;
# its iiwa14 spatial tables keep those inexact turns in the transformed offsets (-0.0607000096 for 0.0607, ~1e-8).
TOL_GRID = 2e-8
TOL_URDF = 2e-7


def _cm(a, n):
    """column-major n x n -> row-major matrix"""
    return np.asarray(a, dtype=np.float64).reshape(n, n).T


def _rpy(r, p, y):
    cr, sr, cp, sp, cy, sy = np.cos(r), np.sin(r), np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx  # URDF: fixed-axis roll, pitch, yaw


def _urdf_chain(joints, q):
    """World pose of every joint frame: T_i = T_{i-1} Trans(xyz_i) RPY(rpy_i) Rot(axis_i, q_i)."""
    T = np.eye(4)
    out = []
    for j, qi in zip(joints, q):
        A = np.eye(4)
        A[:3, :3] = _rpy(*j["rpy"])
        A[:3, 3] = j["xyz"]
        assert j["axis"] == [0.0, 0.0, 1.0]
        Rq = np.eye(4)
        Rq[:2, :2] = [[np.cos(qi), -np.sin(qi)], [np.sin(qi), np.cos(qi)]]
        T = T @ A @ Rq
        out.append(T.copy())
    return out


def _skew(r):
    return np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_joint_frames_compose_to_grid_homogeneous_chain(O, plant):
    d = FIX[plant]
    n = O.nq(plant)
    for q, H in zip(d["q"], d["Xhom"]):
        R, o = O.frames(plant, np.array(q))
        T = np.eye(4)
        for i in range(n):
            T = T @ _cm(H[i], 4)
            assert np.abs(R[i] - T[:3, :3]).max() < TOL_GRID, (plant, i)
            assert np.abs(o[i] - T[:3, 3]).max() < TOL_GRID, (plant, i)
        p, _ = O.fk(plant, np.array(q))
        assert np.abs(p - T[:3, 3]).max() < TOL_GRID


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_parent_to_child_transforms_equal_grid_spatial_transforms(O, plant):
    # X_i = [[E, 0], [-E r^x, E]] with E the child's orientation seen from the parent, transposed, and r the child's
    # origin in parent coordinates (Featherstone's motion transform, the form GRiD's X[i] tables take)
    d = FIX[plant]
    n = O.nq(plant)
    for q, Xs in zip(d["q"], d["X"]):
        R, o = O.frames(plant, np.array(q))
        Rp, op = np.eye(3), np.zeros(3)
        for i in range(n):
            E = (Rp.T @ R[i]).T
            r = Rp.T @ (o[i] - op)
            X = np.zeros((6, 6))
            X[:3, :3] = E
            X[3:, 3:] = E
            X[3:, :3] = -E @ _skew(r)
            assert np.abs(X - _cm(Xs[i], 6)).max() < TOL_GRID, (plant, i)
            Rp, op = R[i], o[i]


@pytest.mark.parametrize("plant", ["iiwa14", "indy7"])
def test_frames_equal_a_chain_composed_from_the_urdf(O, plant):
    d = FIX[plant]
    for q in d["q"]:
        R, o = O.frames(plant, np.array(q))
        for i, T in enumerate(_urdf_chain(d["urdf_joints"], q)):
            assert np.abs(R[i] - T[:3, :3]).max() < TOL_URDF, (plant, i)
            assert np.abs(o[i] - T[:3, 3]).max() < TOL_URDF, (plant, i)


def test_end_effector_at_zero_configuration_in_closed_form(O):
    # At q = 0 the chain is a product of the constant joint placements, summed here by hand from the URDFs.
    # iiwa14 (iiwa14.urdf A1-A7): the link offsets all end up along world z: 0.1575 + 0.2025 + 0.2045 + 0.2155 + 0.1845
    # + 0.2155 + 0.081 = 1.261, with the y offsets -0.0607 (A6) and +0.0607 (A7) cancelling.
    # indy7 (indy7.urdf:208-247): frame 1 = RPY(pi/2, pi/2, 0) maps (x, y, z) to (y, -z, -x); frames 2 and 4 share it,
    # frame 3 is the identity (RPY(-pi/2, 0, pi/2) undoes it). Offsets in world axes: joint1 (0, -0.109, 0.222),
    # joint2 (-0.45, 0, -0.0305) -> (0, 0.0305, 0.45), joint3 (-0.267, 0, -0.075) -> (0, 0.075, 0.267), joint4
    # (0, -0.114, 0.083) unchanged, joint5 (-0.168, 0, 0.069) -> (0, -0.069, 0.168); with joint0's 0.0775:
    # y = -0.109 + 0.0305 + 0.075 - 0.114 - 0.069 = -0.1865, z = 0.0775 + 0.222 + 0.45 + 0.267 + 0.083 + 0.168 = 1.2675.
    p, _ = O.fk("iiwa14", np.zeros(7))
    np.testing.assert_allclose(p, [0.0, 0.0, 1.261], atol=1e-9)
    p, _ = O.fk("indy7", np.zeros(6))
    np.testing.assert_allclose(p, [0.0, -0.1865, 1.2675], atol=1e-9)
