"""Scene inputs for the FEM step: regular grids, hexahedron-to-tetrahedron splits, BoxROI.

Everything here is index/geometry generation that happens once per job on the host; none of
it is on the per-step path.

* ``regular_grid`` follows RegularGridTopology / GridTopology:
  node (i, j, k) has index ``nx*(ny*k + j) + i`` (GridTopology.cpp:352-356) and position
  ``p0 + dx*i + dy*j + dz*k`` with ``dx = (xmax-xmin)/(nx-1)`` (RegularGridTopology.cpp:127-160);
  hexahedra are enumerated x fastest, then y, then z, with corners
  (x,y,z),(x+1,y,z),(x+1,y+1,z),(x,y+1,z),(x,y,z+1),(x+1,y,z+1),(x+1,y+1,z+1),(x,y+1,z+1)
  (GridTopology.cpp:197-212).
* ``split_hexa_5`` is the 5-tetrahedra pattern of SofaCUDA's benchmark helper
  (applications/plugins/SofaCUDA/examples/ElementFEMForcefield/benchmarks/utilities.py:56-83),
  applied to every hexahedron without alternation.
* ``split_hexa_6`` is Hexa2TetraTopologicalMapping with its default ``swapping=false``
  (Hexa2TetraTopologicalMapping.cpp:103-197): the non-swapped 6-tetrahedra pattern.
* ``box_roi`` is BoxROI::isPointInAlignedBox (BoxROI.inl:189-200): inclusive bounds.
"""
from __future__ import annotations

import numpy as np


def regular_grid(n, bmin, bmax):
    """Positions (N,3) float64 and hexahedra (H,8) int32 of a RegularGridTopology."""
    nx, ny, nz = (int(v) for v in n)
    bmin = [float(v) for v in bmin]
    bmax = [float(v) for v in bmax]
    step = []
    for d, nd in enumerate((nx, ny, nz)):
        step.append((bmax[d] - bmin[d]) / (nd - 1) if nd > 1 else (bmax[d] - bmin[d]))
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    # p0 + dx*i + dy*j + dz*k; the other axes contribute exact zeros
    xs = bmin[0] + step[0] * i
    ys = bmin[1] + step[1] * j
    zs = bmin[2] + step[2] * k
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float64)

    def pid(x, y, z):
        return nx * (ny * z + y) + x

    hz, hy, hx = np.meshgrid(np.arange(nz - 1), np.arange(ny - 1), np.arange(nx - 1), indexing="ij")
    hx = hx.ravel(); hy = hy.ravel(); hz = hz.ravel()
    hexa = np.stack([
        pid(hx, hy, hz), pid(hx + 1, hy, hz), pid(hx + 1, hy + 1, hz), pid(hx, hy + 1, hz),
        pid(hx, hy, hz + 1), pid(hx + 1, hy, hz + 1), pid(hx + 1, hy + 1, hz + 1), pid(hx, hy + 1, hz + 1),
    ], axis=1).astype(np.int32)
    return pos, hexa


_SPLIT5 = np.array([[0, 1, 3, 4], [1, 2, 3, 6], [1, 4, 5, 6], [3, 4, 6, 7], [1, 3, 4, 6]], dtype=np.int64)
_SPLIT6 = np.array([[0, 5, 1, 6], [0, 1, 3, 6], [1, 3, 6, 2], [6, 3, 0, 7], [6, 7, 0, 5], [7, 5, 4, 0]], dtype=np.int64)


def split_hexa_5(hexa):
    """(H,8) -> (5H,4) tetrahedra, the SofaCUDA benchmark pattern (non-conforming)."""
    return np.ascontiguousarray(hexa[:, _SPLIT5].reshape(-1, 4).astype(np.int32))


def split_hexa_6(hexa):
    """(H,8) -> (6H,4) tetrahedra, Hexa2TetraTopologicalMapping with swapping=false."""
    return np.ascontiguousarray(hexa[:, _SPLIT6].reshape(-1, 4).astype(np.int32))


_XEDGES = ((0, 1), (4, 5), (3, 2), (7, 6))
_YEDGES = ((4, 7), (5, 6), (1, 2), (0, 3))
_ZEDGES = ((4, 0), (5, 1), (6, 2), (7, 3))


def split_hexa_fem(hexa, nx=2, ny=1):
    """(H,8) -> (6H,4): TetrahedronFEMForceField::init's own tessellation of a hexahedral
    topology (TetrahedronFEMForceField.inl:1318-1370). Cube i's corners are mirrored across x
    when (i mod nx) is even, across y when floor(i/nx) mod ny is odd, across z when
    floor(i/(nx ny)) is odd (edge pairs of sofa/geometry/Hexahedron.h:67-88), then split with
    the 6-tetrahedra pattern. nx, ny: cubes along x and y (2 and 1 when the topology is not a
    grid, as in the reference)."""
    h = np.array(hexa, dtype=np.int64, copy=True)
    i = np.arange(len(h))
    for mask, edges in ((((i % nx) & 1) == 0, _XEDGES), ((((i // nx) % ny) & 1) == 1, _YEDGES),
                        (((i // (nx * ny)) & 1) == 1, _ZEDGES)):
        for a, b in edges:
            ca, cb = h[mask, a].copy(), h[mask, b].copy()
            h[mask, a], h[mask, b] = cb, ca
    return np.ascontiguousarray(h[:, _SPLIT6].reshape(-1, 4).astype(np.int32))


def box_roi(pos, box):
    """Indices of the points inside the box (xmin,ymin,zmin,xmax,ymax,zmax), bounds inclusive."""
    box = np.asarray(box, dtype=np.float64)
    inside = np.all((pos >= box[:3]) & (pos <= box[3:]), axis=1)
    return np.nonzero(inside)[0].astype(np.int32)


def beam(n, length_xyz, origin=(0.0, 0.0, 0.0), split=5):
    """Cantilever beam on a regular grid: positions, tetrahedra and the z=0 face indices."""
    bmin = np.asarray(origin, dtype=np.float64)
    bmax = bmin + np.asarray(length_xyz, dtype=np.float64)
    pos, hexa = regular_grid(n, bmin, bmax)
    tets = split_hexa_5(hexa) if split == 5 else split_hexa_6(hexa)
    fixed = np.nonzero(pos[:, 2] == bmin[2])[0].astype(np.int32)
    return pos, tets, fixed


def incidence_csr(n_nodes, tets):
    """Node -> (element*4 + local vertex) incidence, ascending element id per node.

    This is the order in which TetrahedronFEMForceField accumulates into a node (element
    loop in ascending order, TetrahedronFEMForceField.inl:1630-1634), and the order of
    TetrahedronSetTopologyContainer::createTetrahedraAroundVertexArray (:338-357).
    """
    tets = np.asarray(tets, dtype=np.int64)
    slots = np.arange(tets.size, dtype=np.int64)          # e*4 + j
    nodes = tets.ravel()
    order = np.lexsort((slots, nodes))                    # by node, then by slot (= element)
    inc = slots[order].astype(np.int32)
    counts = np.bincount(nodes, minlength=n_nodes)
    ptr = np.zeros(n_nodes + 1, dtype=np.int32)
    np.cumsum(counts, out=ptr[1:])
    return ptr, inc
