#!/usr/bin/env python3
"""Write tests/golden/grid_inertia.json: the spatial inertia matrices that the reference's
generated dynamics use (GRiD's h_XImats tables, model data), read as numbers from

    gato/dynamics/iiwa14/iiwa14_grid.cuh  init_XImats(): I[0..6] = h_XImats[252..503]
    gato/dynamics/indy7/indy7_grid.cuh    init_XImats(): I[0..5] = h_XImats[216..431]

Each I[i] is a 6x6 column-major matrix [[Ibar, h x], [(h x)^T, m 1]] about the joint origin.
Run once in the container that holds the reference:  python tests/golden/make_grid_inertia_fixture.py
"""
import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "grid_inertia.json"

SOURCES = {
    "iiwa14": ("gato/dynamics/iiwa14/iiwa14_grid.cuh", 7),
    "indy7": ("gato/dynamics/indy7/indy7_grid.cuh", 6),
}
PAT = re.compile(r"h_XImats\[(\d+)\]\s*=\s*static_cast<T>\(([-0-9.eE+]+)\)")


def main():
    out = {}
    for plant, (rel, nq) in SOURCES.items():
        vals = {}
        for m in PAT.finditer((REF / rel).read_text()):
            vals.setdefault(int(m.group(1)), float(m.group(2)))  # first assignment = init_XImats
        base = 36 * nq
        out[plant] = [[vals.get(base + 36 * i + k, 0.0) for k in range(36)] for i in range(nq)]
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
