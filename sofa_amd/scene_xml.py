"""SOFA scene files (.scn, XML) for the FEM step: the subset of components on this path.

A scene is read into one spec per simulated body -- a Node holding an EulerImplicitSolver, a
MechanicalObject and a TetrahedronFEMForceField -- in the format of ``workloads.add_to_engine``,
so that the reference's own scenes (e.g. SofaCUDA's TetrahedronFEMForceField benchmarks) run
unchanged. Read, with their Data names and defaults:

* ``Node``: ``gravity`` (default 0 -9.81 0), ``dt`` (default 0.01), children;
* ``RegularGridTopology``: ``n``/``min``/``max`` or ``nx ny nz``/``xmin .. zmax`` (grid points
  x fastest, hexahedra of RegularGridTopology.cpp / GridTopology.cpp:352-398);
* ``TetrahedronSetTopologyContainer``: ``tetrahedra``/``position`` inline or ``src`` = a link to
  another container; ``Hexa2TetraTopologicalMapping`` (``input`` grid, ``output`` container,
  ``swapping`` false: the 6-tetrahedra split of Hexa2TetraTopologicalMapping.cpp:103-197);
* ``MechanicalObject``: ``position`` inline, a link such as ``@../Beam/grid.position``, or
  ``src`` = a mesh loader;
* ``MeshGmshLoader`` / ``MeshVTKLoader``: ``filename`` (Gmsh 1 / 2.2 / 4.1 or legacy VTK ASCII,
  ``sofa_amd.mesh_io``), ``scale3d``, ``translation``;
* ``TetrahedronFEMForceField``: ``youngModulus``, ``poissonRatio`` (lists), ``method`` (only
  ``large``), ``computeVonMisesStress`` (kept as the spec's ``von_mises`` mode, 1 or 2, for
  ``Engine.von_mises``); a topology with hexahedra only is tessellated as its init does;
* ``UniformMass`` (``vertexMass`` or ``totalMass``), ``DiagonalMass`` (``massDensity`` or
  ``totalMass``), ``MeshMatrixMass`` (``massDensity``, ``lumping``);
* ``BoxROI`` (``box``: one or more 6-tuples, bounds inclusive), ``FixedProjectiveConstraint``
  (``indices`` inline or ``@ROI.indices``, ``fixAll``, ``activate_projectVelocity``),
  ``OscillatorProjectiveConstraint`` (``oscillators``: index, mean, amplitude, pulsation, phase),
  ``LinearMovementProjectiveConstraint`` (``indices`` inline or ``@ROI.indices``, ``keyTimes``,
  ``movements``, ``relativeMovements``; defaults indices {0}, keyTimes {0}, movements {0 0 0},
  relative true);
* ``EulerImplicitSolver`` (``rayleighStiffness``, ``rayleighMass``, ``firstOrder``,
  ``trapezoidalScheme``, ``vdamping``), ``CGLinearSolver`` (``iterations``, ``tolerance``,
  ``threshold``);
* ``PlaneForceField`` (``normal``, ``d``, ``stiffness``, ``damping``, ``maxForce``, ``bilateral``).

``TetrahedronFEMForceField`` ``localStiffnessFactor`` is applied as the reference's
computeMaterialStiffness does (element i's Young modulus times factor[i M / N], M factors for N
elements, TetrahedronFEMForceField.inl:256-261), by handing the engine the products per element.

Every other component (visual models, collision pipeline, plugins, surface or visual mappings)
is listed in ``ignored`` and plays no part in the bodies.  A component on the path with a Data
value this build does not run raises ``ValueError`` rather than running other physics:
``method`` other than large, ``plasticMaxThreshold`` > 0, ``updateStiffnessMatrix``,
``updateStiffness`` or ``computeGlobalMatrix`` set, a force field's or mass's own Rayleigh
coefficient (``rayleighStiffness``, ``rayleighMass``) other than 0, PlaneForceField
``localRange``, ``solveConstraint``, a DiagonalMass given per-vertex masses
only, a UniformMass given no mass at all (the reference marks it invalid, UniformMass.inl
:253-266), and a solver that governs more than one body (SOFA solves such bodies as one system,
with one CG over all of them).
"""
import xml.etree.ElementTree as ET

import numpy as np

from . import meshes



def _floats(s):
    return [float(v) for v in s.replace(",", " ").split()]


def _bool(s):
    return s.strip().lower() in ("1", "true", "yes", "on")


class _Obj:
    def __init__(self, elem, node):
        self.tag = elem.tag
        self.attr = dict(elem.attrib)
        self.node = node
        self.name = self.attr.get("name", "")
        self.data = {}


class _Node:
    def __init__(self, elem, parent):
        self.name = elem.attrib.get("name", "")
        self.parent = parent
        self.attr = dict(elem.attrib)
        self.objects = []
        self.children = []

    def find(self, tag):
        return [o for o in self.objects if o.tag == tag]

    def first(self, *tags):
        for o in self.objects:
            if o.tag in tags:
                return o
        return None


HANDLED = {
    "RegularGridTopology", "GridTopology", "TetrahedronSetTopologyContainer", "Hexa2TetraTopologicalMapping",
    "MechanicalObject", "TetrahedronFEMForceField", "UniformMass", "DiagonalMass", "MeshMatrixMass", "BoxROI",
    "FixedProjectiveConstraint", "FixedConstraint", "OscillatorProjectiveConstraint", "EulerImplicitSolver",
    "LinearMovementProjectiveConstraint", "LinearMovementConstraint",
    "CGLinearSolver", "PlaneForceField", "MeshGmshLoader", "MeshVTKLoader",
}


def _walk(elem, parent, ignored):
    node = _Node(elem, parent)
    for child in elem:
        if child.tag == "Node":
            node.children.append(_walk(child, node, ignored))
        else:
            if child.tag not in HANDLED:
                ignored.append(child.tag)
            node.objects.append(_Obj(child, node))
    return node


def _resolve(node, link):
    """'@../Beam/grid.position' -> (object, 'position'); '@ROI1.indices' -> (object, 'indices')."""
    path = link[1:]
    field = None
    last = path.rsplit("/", 1)[-1]
    if "." in last and last not in (".", ".."):
        head, field = path.rsplit(".", 1)
    else:
        head = path
    parts = head.split("/")
    cur = node
    for p in parts[:-1]:
        if p == "..":
            cur = cur.parent
        elif p in ("", "."):
            continue
        else:
            cur = next(c for c in cur.children if c.name == p)
    target = parts[-1]
    if target == "..":
        return cur.parent, field
    for o in cur.objects:
        if o.name == target:
            return o, field
    raise ValueError(f"link {link} does not resolve")


def _grid(obj):
    a = obj.attr
    if "n" in a:
        n = [int(v) for v in _floats(a["n"])]
        lo = _floats(a.get("min", "0 0 0"))
        hi = _floats(a.get("max", "1 1 1"))
    else:
        n = [int(float(a.get(k, "2"))) for k in ("nx", "ny", "nz")]
        lo = [float(a.get(k, "0")) for k in ("xmin", "ymin", "zmin")]
        hi = [float(a.get(k, "1")) for k in ("xmax", "ymax", "zmax")]
    pos, hexa = meshes.regular_grid(n, lo, hi)
    obj.data.update(position=pos, hexahedra=hexa, n=n)
    return obj


def _load_mesh(obj, search):
    """MeshGmshLoader / MeshVTKLoader: points and tetrahedra of `filename` (looked up beside the
    scene, then in the search paths), then MeshLoader's scale3d and translation (TRSx,
    MeshLoader.cpp:206-234; a rotation is not supported)."""
    import os
    from . import mesh_io
    fn = obj.attr["filename"]
    cands = [fn] if os.path.isabs(fn) else [os.path.join(d, fn) for d in search]
    path = next((c for c in cands if os.path.exists(c)), None)
    if path is None:
        raise ValueError(f"{obj.tag}: {fn} not found in {search}")
    pos, tets = mesh_io.read_mesh(path)
    if any(v != 0.0 for v in _floats(obj.attr.get("rotation", "0 0 0"))):
        raise ValueError(f"{obj.tag}: rotation is not supported")
    sc = np.asarray(_floats(obj.attr.get("scale3d", "1 1 1")))
    tr = np.asarray(_floats(obj.attr.get("translation", "0 0 0")))
    pos = pos * sc + tr
    obj.data.update(position=pos, tetrahedra=tets)


def _evaluate(root, search=(".",)):
    """Fills obj.data for topology objects and mappings (in document order, links resolved)."""
    order = []

    def visit(node):
        for o in node.objects:
            order.append(o)
        for c in node.children:
            visit(c)
    visit(root)
    for o in order:
        if o.tag in ("RegularGridTopology", "GridTopology"):
            _grid(o)
        elif o.tag in ("MeshGmshLoader", "MeshVTKLoader"):
            _load_mesh(o, search)
    for _ in range(3):                      # containers fed by mappings or links to other containers
        for o in order:
            if o.tag == "Hexa2TetraTopologicalMapping":
                if _bool(o.attr.get("swapping", "false")):
                    raise ValueError("Hexa2TetraTopologicalMapping swapping=true is not supported")
                src, _ = _resolve(o.node, o.attr["input"])
                dst, _ = _resolve(o.node, o.attr["output"])
                if "hexahedra" in src.data:
                    dst.data["tetrahedra"] = meshes.split_hexa_6(src.data["hexahedra"])
                    dst.data.setdefault("position", src.data["position"])
            elif o.tag == "TetrahedronSetTopologyContainer":
                if "src" in o.attr:
                    src, _ = _resolve(o.node, o.attr["src"])
                    for k in ("tetrahedra", "position"):
                        if k in src.data:
                            o.data[k] = src.data[k]
                if "tetrahedra" in o.attr:
                    o.data["tetrahedra"] = np.asarray(_floats(o.attr["tetrahedra"]), dtype=np.int32).reshape(-1, 4)
                for k in ("position", "points"):
                    if k in o.attr:
                        o.data["position"] = np.asarray(_floats(o.attr[k])).reshape(-1, 3)
    return order


def _vec(o, key):
    v = o.attr.get(key)
    if v is None:
        return None
    if v.startswith("@"):
        src, field = _resolve(o.node, v)
        return src.data[field or key]
    return v


def load_scene(path_or_text, search_paths=()):
    """Parses a scene (a path, or the XML text itself); returns (specs, info) where specs is a
    list of body specs for ``workloads.add_to_engine`` and info names the bodies and the
    components that were ignored. Mesh files are looked up beside the scene, then in
    `search_paths` (SOFA's data directories, e.g. a checkout's ``share``)."""
    import os
    text = path_or_text
    search = [os.getcwd()]
    if not path_or_text.lstrip().startswith("<"):
        with open(path_or_text) as fh:
            text = fh.read()
        search = [os.path.dirname(os.path.abspath(path_or_text))]
    search += list(search_paths)
    ignored = []
    root = _walk(ET.fromstring(text), None, ignored)
    order = _evaluate(root, search)
    g = _floats(root.attr.get("gravity", "0 -9.81 0"))
    dt = float(root.attr.get("dt", "0.01"))
    # BoxROI indices (on the rest positions of their node's mechanical object or topology)
    bodies, names = [], []

    def node_positions(node):
        mo = node.first("MechanicalObject")
        if mo is not None:
            p = _vec(mo, "position")
            if p is None and "src" in mo.attr:
                p = _resolve(mo.node, mo.attr["src"])[0].data.get("position")
            if p is not None:
                return np.asarray(_floats(p) if isinstance(p, str) else p, dtype=np.float64).reshape(-1, 3)
        for o in node.objects:
            if "position" in o.data:
                return o.data["position"]
        return None

    for o in order:
        if o.tag == "BoxROI":
            p = node_positions(o.node)
            if p is None:
                raise ValueError("BoxROI without positions")
            o.data["indices"] = box_indices(p, o.attr.get("box", "0 0 0 1 1 1"))

    # solvers govern their node's subtree (a context lookup searches up the graph): the nearest
    # EulerImplicitSolver / CGLinearSolver at or above a body's node is its solver
    governed = {}

    def visit(node, solver_up, cg_up):
        solver = node.first("EulerImplicitSolver") or solver_up
        cg = node.first("CGLinearSolver") or cg_up
        fem = node.first("TetrahedronFEMForceField")
        mo = node.first("MechanicalObject")
        if solver is not None and fem is not None and mo is not None:
            # one solver over several mechanical objects is one linear system in SOFA (one CG,
            # its dot products summed over the objects): not what separate scenes compute
            if id(solver) in governed:
                raise ValueError(f"EulerImplicitSolver of node {solver.node.name!r} governs bodies "
                                 f"{governed[id(solver)]!r} and {node.name!r}: one system over several "
                                 "mechanical objects is not on this path (give each body its own solver)")
            governed[id(solver)] = node.name
            bodies.append(_body(node, solver, cg, fem, mo, g, dt, node_positions))
            names.append(node.name)
        for c in node.children:
            visit(c, solver, cg)
    visit(root, None, None)
    return bodies, {"bodies": names, "ignored": sorted(set(ignored))}


def _refuse(obj, key, allowed, what):
    """ValueError when Data `key` of `obj` is set to a value other than the one(s) this path runs."""
    v = obj.attr.get(key)
    if v is None:
        return
    ok = allowed(v)
    if not ok:
        raise ValueError(f"{obj.tag} {key}={v!r}: {what} is not on this path")


def _body(node, solver, cg, fem, mo, g, dt, node_positions):
    pos = node_positions(node)
    if pos is None:
        raise ValueError(f"node {node.name}: no positions")
    # topology: a tetrahedral container of the node (or one its mechanical object links to),
    # else hexahedra of a grid in the node, tessellated as TetrahedronFEMForceField::init does
    tets = None
    for o in node.objects:
        if "tetrahedra" in o.data:
            tets = o.data["tetrahedra"]
            break
    if tets is None:
        grid = next((o for o in node.objects if "hexahedra" in o.data), None)
        if grid is None:
            raise ValueError(f"node {node.name}: no tetrahedral or hexahedral topology")
        nx, ny = grid.data["n"][0] - 1, grid.data["n"][1] - 1
        tets = meshes.split_hexa_fem(grid.data["hexahedra"], nx, ny)
    if fem.attr.get("method", "large") != "large":
        raise ValueError("TetrahedronFEMForceField: only method=large is on this path")
    zero = lambda v: all(x == 0.0 for x in _floats(v))
    false = lambda v: not _bool(v)
    _refuse(fem, "plasticMaxThreshold", zero, "plasticity (TetrahedronFEMForceField.inl:357-367)")
    _refuse(fem, "updateStiffnessMatrix", false, "J recomputed from the deformed shape")
    _refuse(fem, "updateStiffness", false, "the stiffness rebuilt from its Data every step")
    for key in ("computeGlobalMatrix", "assembling"):
        _refuse(fem, key, false, "the assembled global matrix")
    _refuse(fem, "rayleighStiffness", zero, "the force field's own Rayleigh damping")
    E = _floats(fem.attr.get("youngModulus", "5000"))
    nu = _floats(fem.attr.get("poissonRatio", "0.45"))
    lsf = _floats(fem.attr.get("localStiffnessFactor", ""))
    if lsf:
        # computeMaterialStiffness (:256-261): youngModulus = localStiffnessFactor[i M / N] *
        # getYoungModulusInElement(i), the latter E[i] or E[0] beyond the list
        # (BaseLinearElasticityFEMForceField.inl:113-129); the products are the same doubles
        ne, m = len(tets), len(lsf)
        Ei = np.array([E[i] if i < len(E) else E[0] for i in range(ne)], dtype=np.float64)
        fac = np.asarray(lsf, dtype=np.float64)[(np.arange(ne, dtype=np.int64) * m) // ne]
        E = list(fac * Ei)
    n = len(pos)
    spec = dict(pos=pos, tets=np.ascontiguousarray(tets, dtype=np.int32), E=np.array(E), nu=np.array(nu),
                g=tuple(g), dt=dt)
    vm = int(fem.attr.get("computeVonMisesStress", "0"))
    if vm in (1, 2):                      # TetrahedronFEMForceField.inl:57, 0 = no computation
        spec["von_mises"] = vm
    # mass
    um, dm, mm = node.first("UniformMass"), node.first("DiagonalMass"), node.first("MeshMatrixMass")
    for m in (um, dm, mm):
        if m is not None:
            _refuse(m, "rayleighMass", zero, "the mass's own Rayleigh damping")
    if um is not None:
        # UniformMass.inl:233-266: totalMass has priority over vertexMass; neither is invalid
        if "totalMass" in um.attr:
            spec["mass"] = ("uniform", float(um.attr["totalMass"]) / float(n))   # UniformMass.inl:329-345
        elif "vertexMass" in um.attr:
            spec["mass"] = ("uniform", float(um.attr["vertexMass"]))
        else:
            raise ValueError(f"node {node.name}: UniformMass without vertexMass or totalMass "
                             "(the reference marks the component invalid)")
    elif dm is not None:
        # DiagonalMass::massInitialization (DiagonalMass.inl:880-935): totalMass, then
        # massDensity, then vertexMass; none: totalMass = 1
        if "totalMass" in dm.attr:
            spec["mass"] = ("diagonal_total", float(dm.attr["totalMass"]))
        elif "massDensity" in dm.attr:
            spec["mass"] = ("diagonal", float(dm.attr["massDensity"]))
        elif "vertexMass" in dm.attr:
            raise ValueError(f"node {node.name}: DiagonalMass given per-vertex masses (vertexMass) is not on "
                             "this path; give massDensity or totalMass")
        else:
            spec["mass"] = ("diagonal_total", 1.0)
    elif mm is not None:
        rho = _floats(mm.attr.get("massDensity", "1"))
        kind = "meshmatrix_lumped" if _bool(mm.attr.get("lumping", "false")) else "meshmatrix"
        spec["mass"] = (kind, rho[0] if len(rho) == 1 else np.array(rho))
    else:
        raise ValueError(f"node {node.name}: no mass")
    # projective constraints
    fixed = np.zeros(0, np.int32)
    fc = node.first("FixedProjectiveConstraint", "FixedConstraint")
    if fc is not None:
        idx = _vec(fc, "indices")
        if idx is None:
            fixed = np.array([0], np.int32)                  # Data default {0}
        elif isinstance(idx, str):
            fixed = np.asarray(_floats(idx), dtype=np.int32)
        else:
            fixed = np.asarray(idx, dtype=np.int32)
        spec["fix_all"] = _bool(fc.attr.get("fixAll", "false"))
        spec["project_velocity"] = _bool(fc.attr.get("activate_projectVelocity", "false"))
    spec["fixed"] = fixed
    osc = node.first("OscillatorProjectiveConstraint")
    if osc is not None:
        vals = _floats(osc.attr.get("oscillators", ""))
        rec = np.asarray(vals).reshape(-1, 9)
        spec["oscillators"] = dict(index=rec[:, 0].astype(np.int32), mean=rec[:, 1:4], amplitude=rec[:, 4:7],
                                   pulsation=rec[:, 7], phase=rec[:, 8])
    lm = node.first("LinearMovementProjectiveConstraint", "LinearMovementConstraint")
    if lm is not None:
        idx = _vec(lm, "indices")
        if idx is None:
            idx = np.array([0], np.int32)                    # Data default {0}
        elif isinstance(idx, str):
            idx = np.asarray(_floats(idx), dtype=np.int32)
        keys = _floats(lm.attr.get("keyTimes", "0"))
        moves = np.asarray(_floats(lm.attr.get("movements", "0 0 0")), dtype=np.float64).reshape(-1, 3)
        if len(moves) != len(keys):
            raise ValueError("LinearMovementProjectiveConstraint: one movement per key time")
        spec["linear_movement"] = dict(indices=np.asarray(idx, dtype=np.int32), key_times=np.asarray(keys),
                                       movements=moves,
                                       relative=_bool(lm.attr.get("relativeMovements", "true")))
    # solvers
    a = solver.attr
    if _bool(a.get("solveConstraint", "false")):
        raise ValueError("EulerImplicitSolver solveConstraint is not on this path")
    spec["rs"] = float(a.get("rayleighStiffness", "0"))
    spec["rm"] = float(a.get("rayleighMass", "0"))
    spec["integrator"] = dict(first_order=_bool(a.get("firstOrder", "false")),
                              trapezoidal=_bool(a.get("trapezoidalScheme", "false")),
                              vdamping=float(a.get("vdamping", "0")))
    c = cg.attr if cg is not None else {}
    if _bool(c.get("warmStart", "false")):
        spec["warm_start"] = True          # CGLinearSolver.inl:94-105 (FP64 engine)
    spec["iters"] = int(float(c.get("iterations", "25")))
    spec["tol"] = float(c.get("tolerance", "1e-5"))
    spec["thr"] = float(c.get("threshold", "1e-5"))
    pf = node.first("PlaneForceField")
    if pf is not None:
        _refuse(pf, "rayleighStiffness", zero, "the force field's own Rayleigh damping")
        if "localRange" in pf.attr and _floats(pf.attr["localRange"]) != [-1.0, -1.0]:
            raise ValueError(f"PlaneForceField localRange={pf.attr['localRange']!r}: a node range is not on this path")
        p = pf.attr
        spec["plane"] = dict(normal=tuple(_floats(p.get("normal", "0 1 0"))), d=float(p.get("d", "0")),
                             stiffness=float(p.get("stiffness", "500")), damping=float(p.get("damping", "5")),
                             max_force=float(p.get("maxForce", "0")), bilateral=_bool(p.get("bilateral", "false")))
    return spec


def box_indices(node_obj_positions, box_text):
    """BoxROI (BoxROI.inl:189-200): indices of the points inside any of the boxes."""
    b = np.asarray(_floats(box_text)).reshape(-1, 6)
    idx = set()
    for box in b:
        idx.update(meshes.box_roi(node_obj_positions, box).tolist())
    return np.array(sorted(idx), dtype=np.int32)
