"""Scene-spec helpers shared by the GPU parity tests (not a test module)."""
import numpy as np


def oracle_from_spec(spec, dot_mode):
    from oracle import OracleScene
    o = OracleScene(spec["pos"])
    o.set_gravity(spec["g"])
    o.set_fem(spec["tets"], list(np.atleast_1d(spec["E"])), list(np.atleast_1d(spec["nu"])), "large")
    kind, val = spec["mass"]
    if kind == "diagonal_total":
        o.set_diagonal_total_mass(val)
    elif kind == "diagonal":
        o.set_diagonal_mass(val)
    elif kind == "uniform":
        o.set_uniform_mass(vertex_mass=val)
    else:
        o.set_meshmatrix_mass(val, lumping=kind == "meshmatrix_lumped")
    o.set_fixed(spec["fixed"])
    for k in range(len(spec.get("oscillators", {}).get("index", []))):
        osc = spec["oscillators"]
        o.add_oscillator(int(osc["index"][k]), osc["mean"][k], osc["amplitude"][k], float(osc["pulsation"][k]),
                         float(osc["phase"][k]))
    if "linear_movement" in spec:
        o.add_linear_movement(**spec["linear_movement"])
    if spec.get("preconditioner"):
        o.set_preconditioner(spec["preconditioner"])
    if spec.get("warm_start"):
        o.set_warm_start(True)
    o.set_solver(spec["rs"], spec["rm"], spec["iters"], spec["tol"], spec["thr"])
    if "plane" in spec:
        o.add_plane(**spec["plane"])
    o.set_dot_mode(dot_mode)
    return o
