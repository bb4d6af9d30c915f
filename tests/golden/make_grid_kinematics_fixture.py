#!/usr/bin/env python3
"""Write tests/golden/grid_kinematics.json: golden joint transforms of the reference's generated dynamics (GRiD),
evaluated at q = 0 and at seeded configurations, plus the URDF joint origins, as numbers.

GRiD keeps each joint's 6x6 spatial transform X[i] (column-major, h_XImats[36 i .. 36 i + 35]) and 4x4 homogeneous
transform Xhom[i] (column-major, h_XImats[72 nq + 16 i ..]) as a constant table (init_XImats) whose q-dependent
entries are overwritten by c * sin(q_i) or c * cos(q_i), then (for X) the upper-left 3x3 block is copied to the
lower-right one:

    gato/dynamics/iiwa14/iiwa14_grid.cuh  load_update_XImats_helpers, load_update_XmatsHom_helpers
    gato/dynamics/indy7/indy7_grid.cuh    (same functions)

This script reads those tables and update lines as text and evaluates them; the fixture holds only the inputs q and
the resulting matrices. The URDF joint origins (xyz, rpy of each revolute joint, in chain order) come from

    examples/iiwa_description/iiwa14.urdf, examples/indy7_description/indy7.urdf

Run once in the container that holds the reference:  python tests/golden/make_grid_kinematics_fixture.py
"""
import json
import re
import sys
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "grid_kinematics.json"

SOURCES = {
    "iiwa14": ("gato/dynamics/iiwa14/iiwa14_grid.cuh", "examples/iiwa_description/iiwa14.urdf", 7),
    "indy7": ("gato/dynamics/indy7/indy7_grid.cuh", "examples/indy7_description/indy7.urdf", 6),
}
INIT = re.compile(r"h_XImats\[(\d+)\]\s*=\s*static_cast<T>\(([-0-9.eE+]+)\)")
UPD = re.compile(r"s_(?:XImats|XmatsHom)\[(\d+)\]\s*=\s*static_cast<T>\(\s*(-?)\s*(?:([0-9.eE+-]+)\s*\*\s*)?s_temp\[(\d+)\]\s*\)")


def body(text, name):
    """Text of the first definition of `name` taking (T *s_..., const T *s_q, const robotModel<T> *, T *s_temp)."""
    m = re.search(r"void\s+" + name + r"\s*\(\s*T\s*\*\s*\w+\s*,\s*const T\s*\*\s*s_q", text)
    assert m, name
    i = text.index("{", m.end())
    depth = 0
    for j in range(i, len(text)):
        depth += text[j] == "{"
        depth -= text[j] == "}"
        if depth == 0:
            return text[i:j + 1]
    raise ValueError(name)


def updates(src):
    out = []
    for m in UPD.finditer(src):
        idx, neg, coef, k = int(m.group(1)), m.group(2) == "-", m.group(3), int(m.group(4))
        c = float(coef) if coef else 1.0
        out.append((idx, -c if neg else c, k))
    return out


def evaluate(init, upd, q, nq, base, count, copy_block):
    vals = np.array([init.get(base + i, 0.0) for i in range(count)])
    s = np.concatenate([np.sin(q), np.cos(q)])
    for idx, c, k in upd:
        vals[idx] = c * s[k]
    if copy_block:  # lower-right 3x3 of every 6x6 block = its upper-left 3x3 (column-major)
        for i in range(nq):
            for col in range(3):
                for row in range(3):
                    vals[36 * i + col * 6 + row + 21] = vals[36 * i + col * 6 + row]
    return vals


def urdf_joints(path):
    root = ET.parse(path).getroot()
    joints = []
    for j in root.findall("joint"):
        if j.get("type") != "revolute":
            continue
        o = j.find("origin")
        joints.append(dict(name=j.get("name"), xyz=[float(v) for v in o.get("xyz").split()],
                           rpy=[float(v) for v in o.get("rpy").split()],
                           axis=[float(v) for v in j.find("axis").get("xyz").split()]))
    return joints


def main():
    out = {}
    rng = np.random.default_rng(0)
    for plant, (rel, urdf, nq) in SOURCES.items():
        text = (REF / rel).read_text()
        init = {}
        for m in INIT.finditer(text):
            init.setdefault(int(m.group(1)), float(m.group(2)))  # first assignment = init_XImats
        ux = updates(body(text, "load_update_XImats_helpers"))
        uh = updates(body(text, "load_update_XmatsHom_helpers"))
        assert ux and uh and all(i < 36 * nq for i, _, _ in ux) and all(i < 16 * nq for i, _, _ in uh)
        qs = [np.zeros(nq)] + [rng.uniform(-2.5, 2.5, nq) for _ in range(4)]
        X, H = [], []
        for q in qs:
            X.append(evaluate(init, ux, q, nq, 0, 36 * nq, True).reshape(nq, 36).tolist())
            H.append(evaluate(init, uh, q, nq, 72 * nq, 16 * nq, False).reshape(nq, 16).tolist())
        out[plant] = dict(q=[q.tolist() for q in qs], X=X, Xhom=H, urdf_joints=urdf_joints(REF / urdf))
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
