#!/usr/bin/env python3
"""Generate ``model_tables.h``: rigid-body constants of the iiwa14 and indy7 arms.

Build step (run by hand when a URDF changes; the generated header is committed):

    python gato_amd/csrc/gen_model_tables.py [/root/reference]

Per revolute joint i (serial chain, parent = i-1) it emits
  * ``R_tree`` (3x3, row-major): orientation of joint frame i in frame i-1 at q_i = 0,
    composed from every URDF ``<origin rpy xyz>`` between the two revolute joints
    (URDF rpy convention R = Rz(yaw) Ry(pitch) Rx(roll));
  * ``p_tree`` (3): origin of joint frame i in frame i-1;
  * the link inertia about the joint-frame origin, in joint-frame coordinates, with all
    rigidly attached (fixed-joint) child links merged in:
    mass m, first moment h = m c, rotational inertia Ibar = I_c + m (|c|^2 1 - c c^T)
    stored as (xx, yy, zz, xy, xz, yz).
The joint axis is the local +z for every joint of both robots (checked below).

The joint/velocity/torque limits are the plant tables of the reference
(gato/dynamics/iiwa14/iiwa14_plant.cuh:36-70, gato/dynamics/indy7/indy7_plant.cuh:36-61):
they differ slightly from the URDF values for indy7, so they are listed here as data,
before the -0.1 margin that the solver applies.

The spatial inertias these produce equal GRiD's generated tables
(gato/dynamics/iiwa14/iiwa14_grid.cuh:1211-1730, h_XImats[252..503]); the unit tests
check that against a committed fixture.
"""
import math
import sys
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

PLANTS = {
    # name: (urdf path relative to the reference root, limits, keep the inertial origins)
    "iiwa14": (
        "examples/iiwa_description/iiwa14.urdf",
        dict(
            pos=[2.96706, 2.09440, 2.96706, 2.09440, 2.96706, 2.09440, 3.05433],
            vel=[1.48353, 1.48353, 1.74533, 1.30900, 2.26893, 2.35619, 2.35619],
            ctrl=[320.0, 320.0, 176.0, 176.0, 110.0, 40.0, 40.0],
        ),
        True,
    ),
    # The reference's generated indy7 model (gato/dynamics/indy7/indy7_grid.cuh, h_XImats[216..431])
    # places every link's centre of mass at its joint origin: the whitespace-padded <inertial><origin xyz>
    # strings of indy7.urdf were read as zero. The solver's dynamics are that model, so the same
    # convention is reproduced here (first moment 0, inertia tensor about the joint origin = URDF I_com).
    "indy7": (
        "examples/indy7_description/indy7.urdf",
        dict(
            pos=[3.0543, 3.0543, 3.0543, 3.0543, 3.0543, 3.7520],
            vel=[2.61, 2.61, 2.61, 3.14, 3.14, 3.14],
            ctrl=[431.97, 431.97, 197.23, 79.79, 79.79, 79.79],
        ),
        False,
    ),
    # The plant the reference's MPC examples simulate: pinocchio's model of the same URDF (common.py:49-91), which
    # reads the inertial origins, so indy7's links keep their centres of mass. Used by the plant simulator only; for
    # iiwa14 the solver's model already is the URDF's.
    "indy7_urdf": (
        "examples/indy7_description/indy7.urdf",
        dict(
            pos=[3.0543, 3.0543, 3.0543, 3.0543, 3.0543, 3.7520],
            vel=[2.61, 2.61, 2.61, 3.14, 3.14, 3.14],
            ctrl=[431.97, 431.97, 197.23, 79.79, 79.79, 79.79],
        ),
        True,
    ),
}


def rpy_to_R(r, p, y):
    cr, sr = math.cos(r), math.sin(r)
    cp, sp = math.cos(p), math.sin(p)
    cy, sy = math.cos(y), math.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def parse_origin(el):
    o = el.find("origin")
    if o is None:
        return np.eye(3), np.zeros(3)
    xyz = [float(v) for v in o.get("xyz", "0 0 0").split()]
    rpy = [float(v) for v in o.get("rpy", "0 0 0").split()]
    return rpy_to_R(*rpy), np.array(xyz)


def link_inertia(link, keep_origin=True):
    """(m, c, Ic) of a URDF link in its own frame (Ic about the COM)."""
    inert = link.find("inertial")
    if inert is None:
        return 0.0, np.zeros(3), np.zeros((3, 3))
    R, c = parse_origin(inert)
    if not keep_origin:
        R, c = np.eye(3), np.zeros(3)
    m = float(inert.find("mass").get("value"))
    i = inert.find("inertia")
    g = lambda k: float(i.get(k, "0"))
    Ic = np.array([[g("ixx"), g("ixy"), g("ixz")],
                   [g("ixy"), g("iyy"), g("iyz")],
                   [g("ixz"), g("iyz"), g("izz")]])
    return m, c, R @ Ic @ R.T


def parse_chain(urdf_path, keep_origin=True):
    root = ET.parse(urdf_path).getroot()
    links = {l.get("name"): l for l in root.findall("link")}
    joints = root.findall("joint")
    by_parent = {}
    for j in joints:
        by_parent.setdefault(j.find("parent").get("link"), []).append(j)
    children = {j.find("child").get("link") for j in joints}
    roots = [n for n in links if n not in children]
    assert len(roots) == 1, roots
    # walk the serial chain; accumulate fixed transforms into the next revolute joint
    out = []  # list of dict(R_tree, p_tree, bodies=[(m, c, Ic) in joint frame])
    R_acc, p_acc = np.eye(3), np.zeros(3)
    name = roots[0]
    current = None  # index into out of the joint frame that owns `name`
    R_in, p_in = np.eye(3), np.zeros(3)  # pose of link `name` in the owning joint frame
    while True:
        if current is not None:
            m, c, Ic = link_inertia(links[name], keep_origin)
            out[current]["bodies"].append((m, p_in + R_in @ c, R_in @ Ic @ R_in.T))
        kids = by_parent.get(name, [])
        if not kids:
            break
        assert len(kids) == 1, f"not a serial chain at {name}"
        j = kids[0]
        R_o, p_o = parse_origin(j)
        jt = j.get("type")
        if jt == "fixed":
            if current is None:
                # still in the base: accumulate into the first joint's tree transform
                p_acc = p_acc + R_acc @ p_o
                R_acc = R_acc @ R_o
            else:
                p_in = p_in + R_in @ p_o
                R_in = R_in @ R_o
        elif jt in ("revolute", "continuous"):
            axis = [float(v) for v in j.find("axis").get("xyz").split()]
            assert np.allclose(axis, [0, 0, 1]), f"joint {j.get('name')} axis {axis}"
            if current is None:
                R_t, p_t = R_acc @ R_o, p_acc + R_acc @ p_o
            else:
                # tree transform from the current joint frame: through the fixed offset
                R_t, p_t = R_in @ R_o, p_in + R_in @ p_o
            out.append(dict(name=j.get("name"), R_tree=R_t, p_tree=p_t, bodies=[]))
            current = len(out) - 1
            R_in, p_in = np.eye(3), np.zeros(3)
        else:
            raise ValueError(jt)
        name = j.find("child").get("link")
    for jd in out:
        m = sum(b[0] for b in jd["bodies"])
        h = sum((b[0] * b[1] for b in jd["bodies"]), np.zeros(3))
        Ibar = np.zeros((3, 3))
        for bm, bc, bI in jd["bodies"]:
            Ibar += bI + bm * (bc @ bc * np.eye(3) - np.outer(bc, bc))
        jd.update(m=m, h=h, Ibar=Ibar)
    return out


def snap(R):
    """Round rotation entries within 1e-6 of -1, 0 or 1 to those values (the URDF rpy strings
    are float32 renderings of quarter turns; GRiD's generated transforms hold the exact values)."""
    R = R.copy()
    for t in (-1.0, 0.0, 1.0):
        R[np.abs(R - t) < 1e-6] = t
    return R


def fmt(v):
    return repr(float(v)) if v != 0 else "0.0"


def emit(ref_root):
    lines = [
        "/* Generated by gen_model_tables.py from the robots' URDF files. Do not edit. */",
        "#ifndef GATO_MODEL_TABLES_H",
        "#define GATO_MODEL_TABLES_H",
        "",
        "/* Per joint: R_tree[9] (row-major), p_tree[3], mass, h[3], Ibar[6] (xx yy zz xy xz yz),",
        " * then position/velocity/torque limit magnitudes (before the 0.1 margin). */",
        "#define GATO_JOINT_REC 25",
        "",
    ]
    for plant, (rel, lim, keep_origin) in PLANTS.items():
        chain = parse_chain(Path(ref_root) / rel, keep_origin)
        nq = len(chain)
        assert nq == len(lim["pos"]) == len(lim["vel"]) == len(lim["ctrl"])
        lines.append(f"#define GATO_{plant.upper()}_NQ {nq}")
        lines.append(f"#define GATO_{plant.upper()}_JOINTS_INIT {{ \\")
        for i, jd in enumerate(chain):
            I = jd["Ibar"]
            rec = (list(snap(jd["R_tree"]).reshape(-1)) + list(jd["p_tree"]) + [jd["m"]] + list(jd["h"])
                   + [I[0, 0], I[1, 1], I[2, 2], I[0, 1], I[0, 2], I[1, 2]]
                   + [lim["pos"][i], lim["vel"][i], lim["ctrl"][i]])
            lines.append(f"  /* {jd['name']} */ {{" + ", ".join(fmt(v) for v in rec) + "}, \\")
        lines.append("}")
        lines.append(f"static const double gato_{plant}_joints[{nq}][GATO_JOINT_REC] = GATO_{plant.upper()}_JOINTS_INIT;")
        lines.append("")
    lines += [
        "#ifdef __cplusplus",
        "/* compile-time copies for device code: indices fold to constants after unrolling */",
        "namespace gato_model {",
        "template <int PLANT> struct Joints;",
        f"template <> struct Joints<0> {{ static constexpr int NQ = GATO_IIWA14_NQ; "
        f"static constexpr double T[GATO_IIWA14_NQ][GATO_JOINT_REC] = GATO_IIWA14_JOINTS_INIT; }};",
        f"template <> struct Joints<1> {{ static constexpr int NQ = GATO_INDY7_NQ; "
        f"static constexpr double T[GATO_INDY7_NQ][GATO_JOINT_REC] = GATO_INDY7_JOINTS_INIT; }};",
        f"template <> struct Joints<2> {{ static constexpr int NQ = GATO_INDY7_URDF_NQ; "
        f"static constexpr double T[GATO_INDY7_URDF_NQ][GATO_JOINT_REC] = GATO_INDY7_URDF_JOINTS_INIT; }};",
        "}  // namespace gato_model",
        "#endif",
        "",
        "#endif",
    ]
    (HERE / "model_tables.h").write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    emit(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
