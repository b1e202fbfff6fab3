"""Rigid IPC's scene and result files: the JSON scene format read by
SimState / read_rb_scene (src/SimState.cpp:122-262,
src/io/read_rb_scene.cpp:58-339, OBJ meshes as src/io/read_obj.cpp) and the
sim.json written by SimState::save_simulation (src/SimState.cpp:420-457),
so that the reference's fixtures run through this build's stepper and the
reference's tools read its results.

Supported: 3D scenes of dynamic, static and kinematic bodies (a body whose
six dofs are all fixed is static, as in RigidBody's constructor,
src/physics/rigid_body.cpp:78-83; kinematic bodies with their scripted
poses and kinematic_max_time), constant external forces and torques,
meshes from OBJ files or inline arrays,
scale / dimensions, xyz Euler rotations and angular velocities in degrees,
densities and group ids (with the reference's id assignment), and the
solver settings the stepper uses. The CCD runs with the scene's
trajectory_type, "rigid" or "piecewise_linear" (the reference's default,
src/SimState.cpp:179); "linear" and "redon" raise NotImplementedError. The
broad phase of the CCD is the scene's detection_method, "bvh" (the
reference's default, src/SimState.cpp:178) or "hash_grid"; the constraint
set comes out the same from either. The coefficient of restitution is read and, as in the
reference's distance-barrier problem, not used. Anything else (2D,
partially fixed dofs, split components) raises NotImplementedError.
"""
import copy
import json
import math
import os
import sys
import time

import numpy as np

from . import scenes as S

# detection_method names (src/ccd/ccd.hpp:20-28 via NLOHMANN_JSON_SERIALIZE_ENUM)
# -> rigs_params.detection_method (RIGC_DETECTION_*)
DETECTION_METHODS = {"hash_grid": 0, "bvh": 1}
# trajectory_type names (src/ccd/ccd.hpp:30-38 via NLOHMANN_JSON_SERIALIZE_ENUM)
# -> rigs_params.trajectory_type
TRAJECTORY_TYPES = {"rigid": 0, "piecewise_linear": 1}

# SimState::init defaults (src/SimState.cpp:126-205) for the keys used here,
# with the constants filled in there (src/constants.hpp).
DEFAULT_ARGS = {
    "max_iterations": -1,
    "max_time": -1,
    "timestep": 0.01,
    "scene_type": "distance_barrier_rb_problem",
    "solver": "ipc_solver",
    "rigid_body_problem": {
        "rigid_bodies": [],
        "coefficient_restitution": 0.0,
        "coefficient_friction": 0.0,
        "gravity": [0.0, 0.0, 0.0],
        "collision_eps": 0.0,
        "time_stepper": "default",
        "do_intersection_check": False,
    },
    "newton_solver": {
        "max_iterations": 3000,
        "convergence_criteria": "velocity",
        "energy_conv_tol": 1e-5,
        "velocity_conv_tol": 1e-2,
        "is_velocity_conv_tol_abs": False,
        "line_search_lower_bound": 1e-12,
    },
    "ipc_solver": {"dhat_epsilon": 1e-9, "min_barrier_stiffness_scale": 1e11},
    "distance_barrier_constraint": {
        "detection_method": "bvh",
        "trajectory_type": "piecewise_linear",
        "initial_barrier_activation_distance": 1e-3,
        "minimum_separation_distance": 0,
        "barrier_type": "ipc",
    },
    "friction_constraints": {"static_friction_speed_bound": 1e-3, "iterations": 1},
}

BODY_DEFAULTS = {
    "mesh": "", "vertices": [], "edges": [], "faces": [], "density": 1000.0,
    "is_dof_fixed": [False] * 6, "oriented": False, "group_id": -1,
    "position": [0.0, 0.0, 0.0], "rotation": [0.0, 0.0, 0.0], "scale": [1.0, 1.0, 1.0],
    "linear_velocity": [0.0, 0.0, 0.0], "angular_velocity": [0.0, 0.0, 0.0],
    "force": [0.0, 0.0, 0.0], "torque": [0.0, 0.0, 0.0], "enabled": True, "type": "dynamic",
    "kinematic_max_time": -1, "kinematic_poses": [], "split_components": False,
}


def merge_patch(target, patch):
    """RFC 7386 merge patch (nlohmann::json::merge_patch)."""
    if not isinstance(patch, dict):
        return copy.deepcopy(patch)
    out = copy.deepcopy(target) if isinstance(target, dict) else {}
    for k, v in patch.items():
        if v is None:
            out.pop(k, None)
        else:
            out[k] = merge_patch(out.get(k), v)
    return out


def read_obj(path):
    """Vertices, triangles and polyline edges of an OBJ file (1-based indices,
    '/'-separated attribute indices ignored, polygons fanned into triangles,
    'l' lines as edges, as src/io/read_obj.cpp collects them)."""
    V, F, L = [], [], []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            if tok[0] == "v":
                V.append([float(x) for x in tok[1:4]])
            elif tok[0] == "f":
                ids = [int(t.split("/")[0]) for t in tok[1:]]
                ids = [i - 1 if i > 0 else len(V) + i for i in ids]
                for k in range(1, len(ids) - 1):
                    F.append([ids[0], ids[k], ids[k + 1]])
            elif tok[0] == "l":
                ids = [int(t) - 1 for t in tok[1:]]
                for k in range(len(ids) - 1):
                    L.append([ids[k], ids[k + 1]])
    return np.array(V, float).reshape(-1, 3), np.array(F, np.int64).reshape(-1, 3), np.array(L, np.int64).reshape(-1, 2)


def euler_xyz_degrees(rot):
    """AngleAxis(z) * AngleAxis(y) * AngleAxis(x) of xyz angles in degrees
    (read_rb_scene.cpp:183-192)."""
    x, y, z = (float(a) * math.pi / 180.0 for a in rot)
    cx, sx, cy, sy, cz, sz = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def assign_group_ids(types, groups):
    """read_rb_scene.cpp:292-336: static bodies, and bodies sharing a group id
    with a static body, form one group; other bodies without an id get the
    next free ids in order."""
    groups = list(groups)
    static_ids = {g for t, g in zip(types, groups) if t != "dynamic" and g >= 0}
    groups = [(-1 if t != "dynamic" else g) for t, g in zip(types, groups)]
    taken = sorted(g for g in groups if g >= 0)
    ti, nid, static_group = 0, 0, -1
    out = []
    for t, g in zip(types, groups):
        in_static = t != "dynamic" or g in static_ids
        if static_group >= 0 and in_static:
            g = static_group
        elif g < 0:
            while ti < len(taken) and nid >= taken[ti]:
                if nid == taken[ti]:
                    nid += 1
                ti += 1
            g = nid
            nid += 1
        if static_group < 0 and in_static:
            static_group = g
        out.append(g)
    return out


def load(scene, meshes_dir=None, base_dir=None):
    """A scene JSON (path or dict) -> (scene arrays for rigidipc_amd, merged
    args). Mesh paths are tried as given, relative to the scene file, then
    under meshes_dir."""
    if isinstance(scene, (str, os.PathLike)):
        base_dir = base_dir or os.path.dirname(os.path.abspath(scene))
        with open(scene) as f:
            scene = json.load(f)
    args = merge_patch(DEFAULT_ARGS, scene)
    # the IPC solver shares the Newton solver's settings (SimState.cpp:196-201, 236-239)
    args["ipc_solver"] = merge_patch(args["newton_solver"], args["ipc_solver"])
    rbp = args["rigid_body_problem"]
    if args.get("scene_type", "distance_barrier_rb_problem") != "distance_barrier_rb_problem":
        raise NotImplementedError(f"scene_type {args['scene_type']!r}: only distance_barrier_rb_problem")
    bodies, types, groups = [], [], []
    for jrb in rbp.get("rigid_bodies", []):
        b = merge_patch(BODY_DEFAULTS, jrb)
        if not b["enabled"]:
            continue
        if b["mesh"]:
            path = b["mesh"]
            cands = [path]
            if base_dir:
                cands.append(os.path.join(base_dir, path))
            if meshes_dir:
                cands.append(os.path.join(meshes_dir, path))
            found = next((c for c in cands if os.path.exists(c)), None)
            if found is None:
                raise FileNotFoundError(f"mesh {path!r} not found (tried {cands})")
            V, F, _ = read_obj(found)
        else:
            V = np.array(b["vertices"], float).reshape(-1, 3)
            F = np.array(b["faces"], np.int64).reshape(-1, 3)
        if V.shape[1] != 3:
            raise NotImplementedError("2D bodies")
        if b["split_components"]:
            raise NotImplementedError("split_components")
        kinematic = b["type"] == "kinematic"
        # scripted poses: degrees of freedom as given (position, rotation
        # vector), read_rb_scene.cpp:233-246
        kposes = [list(np.array(kp.get("position", [0, 0, 0]), float)[:3])
                  + list(np.array(kp.get("rotation", [0, 0, 0]), float)[:3]) for kp in b["kinematic_poses"]]
        force = list(np.array(b["force"], float)[:3]) + list(np.array(b["torque"], float)[:3])
        fixed = b["is_dof_fixed"]
        fixed = [bool(fixed)] * 6 if isinstance(fixed, bool) else [bool(x) for x in fixed[:6]]
        static = b["type"] == "static" or all(fixed)
        if any(fixed) and not static:
            raise NotImplementedError("partially fixed degrees of freedom")
        # scale (or dimensions) about the model origin, then the Euler rotation
        if "dimensions" in jrb:
            dims = np.abs(V.max(0) - V.min(0))
            dims[dims == 0] = 1.0
            scale = np.array(jrb["dimensions"], float)[:3] / dims
        elif isinstance(b["scale"], (int, float)):
            scale = np.full(3, float(b["scale"]))
        else:
            scale = np.array(b["scale"], float)[:3]
        V = (V * scale) @ euler_xyz_degrees(b["rotation"]).T
        # angular velocity in the mesh's own frame, as the reference takes it
        # (rigid_body.cpp:132-140: pose rotation R_i R0, omega = R0^T omega_0),
        # so its world value is R_i omega_0
        w = euler_xyz_degrees(b["rotation"]) @ (np.array(b["angular_velocity"], float)[:3] * math.pi / 180.0)
        v = np.array(b["linear_velocity"], float)[:3]
        p = np.array(b["position"], float)[:3]
        bodies.append(S.Body(V, F.astype(np.int64), np.eye(3), p, v, w, density=float(b["density"]), static=static,
                             kinematic=kinematic and not static,
                             kinematic_max_time=float(b["kinematic_max_time"]), kinematic_poses=kposes,
                             force=None if static or not any(x != 0.0 for x in force) else force))
        types.append("static" if static else ("kinematic" if kinematic else "dynamic"))
        groups.append(int(b["group_id"]))
    for body, g in zip(bodies, assign_group_ids(types, groups)):
        body.group = g
    sc = S.assemble(bodies)
    dbc = args["distance_barrier_constraint"]
    fc = args["friction_constraints"]
    ipc = args["ipc_solver"]
    sc.dt = float(args["timestep"])
    sc.dhat = float(dbc["initial_barrier_activation_distance"])
    sc.minimum_separation_distance = float(dbc["minimum_separation_distance"])
    sc.mu = float(rbp.get("coefficient_friction", 0.0))
    sc.gravity = np.array(rbp.get("gravity", [0, 0, 0]), float)[:3]
    steps = int(args["max_iterations"])
    if float(args["max_time"]) >= 0:
        steps = int(math.ceil(float(args["max_time"]) / sc.dt))
    sc.steps = steps
    sc.friction_iterations = int(fc["iterations"])
    sc.static_friction_speed_bound = float(fc["static_friction_speed_bound"])
    sc.newton = ipc
    sc.do_intersection_check = bool(rbp.get("do_intersection_check", False))
    traj = str(dbc.get("trajectory_type", "piecewise_linear"))
    if traj not in TRAJECTORY_TYPES:
        raise NotImplementedError(f"trajectory_type {traj!r}: only 'rigid' and 'piecewise_linear'")
    sc.trajectory_type = traj
    det = str(dbc.get("detection_method", "bvh"))
    if det not in DETECTION_METHODS:
        raise NotImplementedError(f"detection_method {det!r}: only 'hash_grid' and 'bvh'")
    sc.detection_method = det
    sc.name = str(scene.get("name", "scene"))
    return sc, args


def params_for(sc):
    """rigs_params for a loaded scene (the IPC solver's Newton settings)."""
    from . import sim as M
    p = M.default_params(sc)
    p.minimum_separation_distance = sc.minimum_separation_distance
    p.friction_iterations = sc.friction_iterations
    p.static_friction_speed_bound = sc.static_friction_speed_bound
    ipc = sc.newton
    p.max_iterations = int(ipc.get("max_iterations", p.max_iterations))
    p.line_search_lower_bound = float(ipc.get("line_search_lower_bound", p.line_search_lower_bound))
    p.dhat_epsilon = float(ipc.get("dhat_epsilon", p.dhat_epsilon))
    p.min_barrier_stiffness_scale = float(ipc.get("min_barrier_stiffness_scale", p.min_barrier_stiffness_scale))
    tol = float(ipc.get("velocity_conv_tol", p.velocity_conv_tol))
    if ipc.get("is_velocity_conv_tol_abs", False):
        # the stepper takes the tolerance relative to the initial world bbox
        # diagonal (the reference's default form)
        W = []
        for b in sc.bodies:
            R = S.rotvec_to_matrix(np.asarray(b.pose[3:]))
            W.append(b.V @ R.T + b.pose[:3])
        W = np.concatenate(W)
        tol = tol / float(np.linalg.norm(W.max(0) - W.min(0)))
    p.velocity_conv_tol = tol
    p.do_intersection_check = int(bool(sc.do_intersection_check))
    p.trajectory_type = TRAJECTORY_TYPES[getattr(sc, "trajectory_type", "rigid")]
    p.detection_method = DETECTION_METHODS[getattr(sc, "detection_method", "hash_grid")]
    return p


def run(scene, out_path=None, steps=None, meshes_dir=None, device=0, backend=None):
    """Simulate a scene file and write the reference's sim.json (args,
    animation.state_sequence, stats). Returns the result dict."""
    from . import sim as M
    sc, args = load(scene, meshes_dir=meshes_dir)
    n = sc.steps if steps is None else int(steps)
    args = copy.deepcopy(args)
    args["max_iterations"], args["max_time"] = n, -1
    s = M.Simulation(sc, params=params_for(sc), device=device, backend=backend)
    states, timings, iters, contacts, mind = [], [], [], [], []
    # the reference's log lines for intersections (rigid_body_problem.cpp:100,
    # SimState.cpp:389-393); sim.json has no field for them
    if s.stats()["initial_intersections"] > 0:
        print("The initial state contains intersections!", file=sys.stderr)

    def state():
        poses, vel = s.state()
        return {"rigid_bodies": [{"position": poses[i, :3].tolist(), "rotation": poses[i, 3:].tolist(),
                                  "linear_velocity": vel[i, :3].tolist(), "angular_velocity": vel[i, 3:].tolist()}
                                 for i in range(len(poses))]}

    states.append(state())
    prev = s.stats()
    for _ in range(n):
        t0 = time.perf_counter()
        s.step(1)
        timings.append(time.perf_counter() - t0)
        st = s.stats()
        if st["intersecting_steps"] > prev["intersecting_steps"]:
            print(f"sim_state action=simulation_step sim_it={len(timings)} status=has_intersections",
                  file=sys.stderr)
        iters.append(int(st["newton_iterations"] - prev["newton_iterations"]))
        contacts.append(int(st["num_constraints"]))
        mind.append(float(st["min_distance"]) if st["min_distance"] >= 0 else None)
        prev = st
        states.append(state())
        states[-1]["min_distance"] = mind[-1]
    s.close()
    result = {
        "args": args,
        "animation": {"state_sequence": states},
        "stats": {
            "dim": 3, "num_bodies": int(len(sc.group_ids)), "num_vertices": int(len(sc.vertices)),
            "num_edges": int(len(sc.edges)), "num_faces": int(len(sc.faces)), "num_timesteps": n,
            "step_timings": timings, "solver_iterations": iters, "num_contacts": contacts,
            "step_minimum_distances": mind,
        },
    }
    if out_path:
        with open(out_path, "w") as f:
            json.dump(result, f)
    return result


def main(argv=None):
    """python -m rigidipc_amd.scene_io SCENE.json OUTPUT_DIR [--steps N]
    (the reference's rigid_ipc_sim writes OUTPUT_DIR/sim.json the same way)."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("scene")
    ap.add_argument("output_dir")
    ap.add_argument("--steps", type=int, default=None, help="override max_time / max_iterations")
    ap.add_argument("--meshes", default=None, help="folder searched for mesh files")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    os.makedirs(a.output_dir, exist_ok=True)
    r = run(a.scene, out_path=os.path.join(a.output_dir, "sim.json"), steps=a.steps, meshes_dir=a.meshes,
            device=a.device)
    st = r["stats"]
    print(f"{st['num_timesteps']} steps, {sum(st['solver_iterations'])} Newton iterations, "
          f"{sum(st['step_timings']):.3f} s")


if __name__ == "__main__":
    main()
