"""Oracle: gmsh MSH 2.2 ASCII reader (restates src/mesh.cpp:30-206).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
from __future__ import annotations

import numpy as np

_NODE_COUNT = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 8: 3, 9: 6, 11: 10, 15: 1}


def load_msh(path):
    """Returns (nodes (N,3), tet10 (E,10), {name: sorted node ids})."""
    phys = {}
    nodes, g2l, tets, sets = [], {}, [], {}
    with open(path) as f:
        lines = [ln.rstrip("\r\n") for ln in f]
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln == "$MeshFormat":
            ver = float(lines[i + 1].split()[0])
            if ver < 2.0 or ver >= 3.0:
                raise RuntimeError("Unsupported Gmsh format version")
            i += 2
        elif ln == "$PhysicalNames":
            n = int(lines[i + 1])
            for q in range(n):
                dim, tag, name = lines[i + 2 + q].split(maxsplit=2)
                phys[int(tag)] = name.strip().strip('"')
            i += n + 2
        elif ln == "$Nodes":
            n = int(lines[i + 1])
            for q in range(n):
                p = lines[i + 2 + q].split()
                g2l[int(p[0])] = len(nodes)
                nodes.append([float(p[1]), float(p[2]), float(p[3])])
            i += n + 2
        elif ln == "$Elements":
            n = int(lines[i + 1])
            X = np.asarray(nodes)
            for q in range(n):
                p = [int(t) for t in lines[i + 2 + q].split()]
                et, nt = p[1], p[2]
                tags = p[3:3 + nt]
                nn = _NODE_COUNT.get(et, -1)
                if nn < 0:
                    raise RuntimeError(f"Unknown Gmsh element type: {et}")
                en = [g2l[g] for g in p[3 + nt:3 + nt + nn]]
                if et == 11:
                    m13 = 0.5 * (X[en[1]] + X[en[3]])
                    m23 = 0.5 * (X[en[2]] + X[en[3]])
                    if np.sum((X[en[8]] - m23) ** 2) < np.sum((X[en[8]] - m13) ** 2):
                        en[8], en[9] = en[9], en[8]
                    tets.append(en)
                if tags and tags[0] > 0 and tags[0] in phys:
                    sets.setdefault(phys[tags[0]], set()).update(en)
            i += n + 2
        else:
            i += 1
    return (np.asarray(nodes, float), np.asarray(tets, np.int64),
            {k: np.array(sorted(v), np.int64) for k, v in sets.items()})
