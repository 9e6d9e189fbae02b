"""Tetrahedral mesh files: Gmsh (.msh: format 1, 2.2 and 4.1 ASCII) and legacy VTK (.vtk ASCII
unstructured grids), as SOFA's MeshGmshLoader / MeshVTKLoader read them
(Sofa/Component/IO/Mesh/src/sofa/component/io/mesh/MeshGmshLoader.cpp, MeshVTKLoader.cpp): the
points in file order (Gmsh node numbers mapped to their position in the file), the 4-node
tetrahedra in file order (Gmsh element type 4, VTK cell type 10); other element types skipped.
"""
import numpy as np


def _tokens(lines, i):
    return lines[i].split()


def read_gmsh(path):
    """(positions (N,3) float64, tetrahedra (E,4) int32) of a Gmsh file."""
    with open(path) as fh:
        lines = [l.strip() for l in fh if l.strip()]
    pos, ids, tets = [], {}, []
    i = 0
    fmt = 1.0
    while i < len(lines):
        tag = lines[i]
        if tag == "$MeshFormat":
            fmt = float(lines[i + 1].split()[0])
            i += 2
        elif tag in ("$NOD", "$Nodes") and fmt < 4.0:
            n = int(lines[i + 1])
            for k in range(n):
                t = lines[i + 2 + k].split()
                ids[int(t[0])] = len(pos)
                pos.append([float(t[1]), float(t[2]), float(t[3])])
            i += 2 + n
        elif tag == "$Nodes":                       # 4.x: entity blocks, tags then coordinates
            nb = int(lines[i + 1].split()[0])
            j = i + 2
            for _ in range(nb):
                cnt = int(lines[j].split()[3])
                tags = [int(lines[j + 1 + k]) for k in range(cnt)]
                for k in range(cnt):
                    t = lines[j + 1 + cnt + k].split()
                    ids[tags[k]] = len(pos)
                    pos.append([float(t[0]), float(t[1]), float(t[2])])
                j += 1 + 2 * cnt
            i = j
        elif tag in ("$ELM", "$Elements") and fmt < 4.0:
            m = int(lines[i + 1])
            for k in range(m):
                t = [int(v) for v in lines[i + 2 + k].split()]
                if t[1] != 4:
                    continue
                nodes = t[5:9] if fmt < 2.0 else t[3 + t[2]:3 + t[2] + 4]
                tets.append([ids[v] for v in nodes])
            i += 2 + m
        elif tag == "$Elements":                    # 4.x: entity blocks of one element type
            nb = int(lines[i + 1].split()[0])
            j = i + 2
            for _ in range(nb):
                h = lines[j].split()
                etype, cnt = int(h[2]), int(h[3])
                if etype == 4:
                    for k in range(cnt):
                        t = [int(v) for v in lines[j + 1 + k].split()]
                        tets.append([ids[v] for v in t[1:5]])
                j += 1 + cnt
            i = j
        else:
            i += 1
    return np.asarray(pos, dtype=np.float64).reshape(-1, 3), np.asarray(tets, dtype=np.int32).reshape(-1, 4)


def read_vtk(path):
    """(positions (N,3) float64, tetrahedra (E,4) int32) of a legacy ASCII VTK unstructured grid."""
    with open(path) as fh:
        tok = fh.read().split()
    i = 0
    pos = cells = types = None
    while i < len(tok):
        w = tok[i].upper()
        if w == "POINTS":
            n = int(tok[i + 1])
            pos = np.asarray([float(v) for v in tok[i + 3:i + 3 + 3 * n]]).reshape(n, 3)
            i += 3 + 3 * n
        elif w == "CELLS":
            m, size = int(tok[i + 1]), int(tok[i + 2])
            vals = [int(v) for v in tok[i + 3:i + 3 + size]]
            cells, j = [], 0
            for _ in range(m):
                k = vals[j]
                cells.append(vals[j + 1:j + 1 + k])
                j += 1 + k
            i += 3 + size
        elif w == "CELL_TYPES":
            m = int(tok[i + 1])
            types = [int(v) for v in tok[i + 2:i + 2 + m]]
            i += 2 + m
        else:
            i += 1
    if pos is None or cells is None or types is None:
        raise ValueError(f"{path}: not a legacy unstructured grid")
    tets = [c for c, t in zip(cells, types) if t == 10 and len(c) == 4]
    return pos, np.asarray(tets, dtype=np.int32).reshape(-1, 4)


def read_mesh(path):
    """By extension: .msh (Gmsh) or .vtk (legacy VTK)."""
    low = path.lower()
    if low.endswith(".msh"):
        return read_gmsh(path)
    if low.endswith(".vtk"):
        return read_vtk(path)
    raise ValueError(f"{path}: unknown mesh format")
