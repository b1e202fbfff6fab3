"""The reference's scene format (rigidipc_amd.scene_io): JSON defaults and
merge patches, OBJ meshes, xyz Euler rotations in degrees, group-id
assignment (src/io/read_rb_scene.cpp:58-339) and the sim.json layout that
the reference's tools/benchmark.py reads (src/SimState.cpp:420-457),
checked on fixtures copied from the reference (tests/golden/fixtures, made
by tests/golden/make_fixtures.py)."""
import glob
import json
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from gpu_util import PKG  # noqa: F401  (puts the package on sys.path)
from rigidipc_amd import scene_io as IO
from rigidipc_amd import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "fixtures")


def test_merge_patch():
    a = {"x": 1, "n": {"p": 1, "q": 2}, "l": [1, 2]}
    b = {"n": {"q": None, "r": 3}, "l": [5], "y": 2}
    assert IO.merge_patch(a, b) == {"x": 1, "n": {"p": 1, "r": 3}, "l": [5], "y": 2}


@pytest.mark.parametrize("angles", [(90, 0, 0), (0, 90, 0), (0, 0, 90), (30, 45, 60), (-20, 170, 5)])
def test_euler_rotation_is_z_y_x(angles):
    # AngleAxis(z) * AngleAxis(y) * AngleAxis(x): intrinsic Z-Y-X
    want = Rotation.from_euler("ZYX", [angles[2], angles[1], angles[0]], degrees=True).as_matrix()
    np.testing.assert_allclose(IO.euler_xyz_degrees(angles), want, atol=1e-15)


def test_euler_quarter_turn_about_z_maps_x_to_y():
    np.testing.assert_allclose(IO.euler_xyz_degrees((0, 0, 90)) @ [1, 0, 0], [0, 1, 0], atol=1e-15)


@pytest.mark.parametrize("types,groups,want", [
    (["dynamic", "dynamic", "static", "static"], [-1, -1, -1, -1], [0, 1, 2, 2]),
    (["dynamic", "dynamic", "dynamic", "static"], [0, -1, -1, -1], [0, 1, 2, 3]),
    (["dynamic", "dynamic", "dynamic"], [-1, 1, -1], [0, 1, 2]),
    (["static", "dynamic", "dynamic"], [5, 5, -1], [0, 0, 1]),
    (["dynamic", "static", "dynamic", "static"], [3, -1, -1, -1], [3, 0, 1, 0]),
])
def test_group_ids_follow_the_reference_rule(types, groups, want):
    assert IO.assign_group_ids(types, groups) == want


def test_obj_reader(tmp_path):
    p = tmp_path / "quad.obj"
    p.write_text("# quad\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1/1/1 2/2/2 3/3/3 4/4/4\nl 1 3\n")
    V, F, L = IO.read_obj(str(p))
    assert V.shape == (4, 3)
    np.testing.assert_array_equal(F, [[0, 1, 2], [0, 2, 3]])
    np.testing.assert_array_equal(L, [[0, 2]])


def test_five_cubes_fixture():
    sc, args = IO.load(os.path.join(FIX, "unit-tests-5-cubes.json"))
    assert len(sc.group_ids) == 6 and len(sc.vertices) == 5 * 8 + 4
    np.testing.assert_array_equal(sc.is_static, [0, 0, 0, 0, 0, 1])
    # unit cubes centred on their mesh origin: the pose is the given position
    np.testing.assert_allclose(sc.poses[:5, :3], [[0, 1 + 2 * k, 0] for k in range(5)], atol=1e-12)
    np.testing.assert_allclose(sc.mass[:5], 1000.0, rtol=1e-12)
    assert sc.steps == 1000 and sc.dt == 0.01 and sc.dhat == 1e-3 and sc.mu == 0.0
    np.testing.assert_allclose(sc.gravity, [0, -9.8, 0])
    assert IO.params_for(sc).velocity_conv_tol == 1e-5


def test_tet_corner_fixture_scale_and_statics():
    sc, args = IO.load(os.path.join(FIX, "unit-tests-tet-corner.json"))
    np.testing.assert_array_equal(sc.is_static, [1, 1, 1, 1, 0])
    assert len(set(sc.group_ids[:4].tolist())) == 1 and sc.group_ids[4] != sc.group_ids[0]
    assert sc.dhat == 1e-4 and sc.steps == 300
    V, F, _ = IO.read_obj(os.path.join(FIX, "meshes", "tet-corner.obj"))
    ext = sc.vertices[sc.body_vertex_start[4]:sc.body_vertex_start[5]]
    # scale 0.999 about the mesh origin: every vertex-to-vertex distance
    # shrinks with it (the body frame is rotated, so compare distances)
    def diameter(X):
        return np.linalg.norm(X[:, None, :] - X[None, :, :], axis=-1).max()
    assert abs(diameter(ext) / diameter(V) - 0.999) < 1e-12


def test_rotation_and_angular_velocity_in_degrees():
    Vc, Fc = S.cube(1.0)
    scene = {"rigid_body_problem": {"rigid_bodies": [
        {"vertices": Vc.tolist(), "faces": Fc.tolist(), "rotation": [0, 0, 90], "angular_velocity": [0, 0, 180],
         "position": [1, 2, 3]}]}}
    sc, _ = IO.load(scene)
    # world angular velocity pi rad/s about z; stored in the body frame
    b = sc.bodies[0]
    Rw = S.rotvec_to_matrix(b.pose[3:])
    np.testing.assert_allclose(Rw @ sc.velocities[0, 3:], [0, 0, np.pi], atol=1e-12)
    np.testing.assert_allclose(sc.poses[0, :3], [1, 2, 3], atol=1e-12)


def test_angular_velocity_is_in_the_mesh_frame():
    """The reference stores omega = R0^T omega_0 with the pose rotation
    R_i R0 (rigid_body.cpp:132-140): the JSON angular velocity is in the
    unrotated mesh's frame, so a body rotated 90 degrees about x spinning
    about its own z turns about world -y."""
    Vc, Fc = S.cube(1.0)
    scene = {"rigid_body_problem": {"rigid_bodies": [
        {"vertices": Vc.tolist(), "faces": Fc.tolist(), "rotation": [90, 0, 0], "angular_velocity": [0, 0, 180]}]}}
    sc, _ = IO.load(scene)
    Rw = S.rotvec_to_matrix(sc.bodies[0].pose[3:])
    np.testing.assert_allclose(Rw @ sc.velocities[0, 3:], [0, -np.pi, 0], atol=1e-12)


def test_kinematic_bodies_and_forces_are_read():
    """type "kinematic" with its scripted poses (position, rotation vector as
    given) and max time, and force / torque (read_rb_scene.cpp:85-104,
    225-246); the stepper's handling is tested in test_kinematic.py."""
    Vc, Fc = S.cube(1.0)
    kin = {"vertices": Vc.tolist(), "faces": Fc.tolist(), "type": "kinematic", "kinematic_max_time": 0.5,
           "kinematic_poses": [{"position": [0, 1, 0]}, {"position": [0, 2, 0], "rotation": [0.1, 0, 0]}]}
    dyn = {"vertices": Vc.tolist(), "faces": Fc.tolist(), "position": [3, 0, 0], "force": [0, 1, 0],
           "torque": [0, 0, 2]}
    sc, _ = IO.load({"rigid_body_problem": {"rigid_bodies": [kin, dyn]}})
    assert sc.kinematic.tolist() == [1, 0]
    assert sc.kinematic_max_time.tolist() == [0.5, -1.0]
    assert sc.kinematic_pose_start.tolist() == [0, 2, 2]
    np.testing.assert_array_equal(sc.kinematic_poses, [[0, 1, 0, 0, 0, 0], [0, 2, 0, 0.1, 0, 0]])
    np.testing.assert_array_equal(sc.force, [[0] * 6, [0, 1, 0, 0, 0, 2]])
    # kinematic bodies join the static bodies' group (read_rb_scene.cpp:292-336)
    assert sc.group_ids[0] != sc.group_ids[1]


@pytest.mark.parametrize("body,msg", [
    ({"is_dof_fixed": [True, False, False, False, False, False]}, "partially fixed"),
    ({"split_components": True}, "split_components"),
])
def test_unsupported_features_raise(body, msg):
    Vc, Fc = S.cube(1.0)
    b = {"vertices": Vc.tolist(), "faces": Fc.tolist()}
    b.update(body)
    with pytest.raises(NotImplementedError, match=msg):
        IO.load({"rigid_body_problem": {"rigid_bodies": [b]}})


def test_all_copied_fixtures_load():
    names = sorted(os.path.basename(f) for f in glob.glob(os.path.join(FIX, "*.json")))
    assert len(names) >= 6
    for f in names:
        sc, _ = IO.load(os.path.join(FIX, f))
        assert len(sc.group_ids) >= 2 and sc.steps > 0


def test_sim_json_has_the_fields_the_reference_tools_read(tmp_path):
    from stepper_util import oracle_backend
    out = tmp_path / "sim.json"
    r = IO.run(os.path.join(FIX, "unit-tests-edge-edge.json"), out_path=str(out), steps=3, backend=oracle_backend())
    j = json.loads(out.read_text())
    assert j == json.loads(json.dumps(r))
    for k in ("dim", "num_bodies", "num_vertices", "num_edges", "num_faces", "num_timesteps", "step_timings",
              "solver_iterations", "num_contacts"):
        assert k in j["stats"], k
    a = j["args"]
    for path in (("timestep",), ("distance_barrier_constraint", "initial_barrier_activation_distance"),
                 ("rigid_body_problem", "coefficient_friction"), ("friction_constraints", "static_friction_speed_bound"),
                 ("friction_constraints", "iterations"), ("rigid_body_problem", "coefficient_restitution"),
                 ("ipc_solver", "velocity_conv_tol")):
        x = a
        for p in path:
            x = x[p]
    # the fixture leaves trajectory_type at the reference's default, and the
    # run uses it
    assert a["distance_barrier_constraint"]["trajectory_type"] == "piecewise_linear"
    seq = j["animation"]["state_sequence"]
    assert len(seq) == 4 and len(seq[0]["rigid_bodies"]) == 2
    assert j["stats"]["num_timesteps"] == 3 and len(j["stats"]["solver_iterations"]) == 3


def test_intersection_check_flag_reaches_the_stepper(tmp_path):
    base = os.path.join(FIX, "unit-tests-5-cubes.json")
    scene = json.load(open(base))
    sc, _ = IO.load(base)
    assert IO.params_for(sc).do_intersection_check == 0
    scene.setdefault("rigid_body_problem", {})["do_intersection_check"] = True
    p = tmp_path / "checked.json"
    p.write_text(json.dumps(scene))
    sc, _ = IO.load(str(p), base_dir=os.path.dirname(base))
    assert IO.params_for(sc).do_intersection_check == 1


def test_runner_reports_an_intersecting_initial_state(tmp_path, capsys):
    from stepper_util import oracle_backend
    base = os.path.join(FIX, "unit-tests-5-cubes.json")
    scene = json.load(open(base))
    scene["rigid_body_problem"]["rigid_bodies"][1]["position"] = [0.25, 1.5, 0.125]  # into the first cube
    p = tmp_path / "overlap.json"
    p.write_text(json.dumps(scene))
    IO.run(str(p), steps=0, meshes_dir=FIX, backend=oracle_backend())
    assert "The initial state contains intersections!" in capsys.readouterr().err
    IO.run(base, steps=0, backend=oracle_backend())
    assert "intersections" not in capsys.readouterr().err


def test_trajectory_type_reaches_the_stepper(tmp_path):
    base = os.path.join(FIX, "unit-tests-edge-edge.json")
    sc, args = IO.load(base)
    assert sc.trajectory_type == "piecewise_linear"  # SimState.cpp:179 default
    assert IO.params_for(sc).trajectory_type == 1
    scene = json.load(open(base))
    for name, want in (("rigid", 0), ("piecewise_linear", 1)):
        scene.setdefault("distance_barrier_constraint", {})["trajectory_type"] = name
        p = tmp_path / f"{name}.json"
        p.write_text(json.dumps(scene))
        sc, _ = IO.load(str(p), base_dir=os.path.dirname(base))
        assert IO.params_for(sc).trajectory_type == want
    for name in ("linear", "redon"):
        scene["distance_barrier_constraint"]["trajectory_type"] = name
        with pytest.raises(NotImplementedError, match="trajectory_type"):
            IO.load(scene, base_dir=os.path.dirname(base))


def test_piecewise_linear_steps_through_the_oracle_stepper():
    """Both trajectory types step the edge-edge fixture over the oracle
    backend. The piecewise-linear CCD is conservative with respect to the
    rigid trajectory, so its line search never takes a longer step than the
    rigid CCD allows: both runs end without intersections, and both reach
    contact (constraints active)."""
    from stepper_util import oracle_backend
    from rigidipc_amd import sim as M
    base = os.path.join(FIX, "unit-tests-edge-edge.json")
    out = {}
    for name in ("rigid", "piecewise_linear"):
        sc, _ = IO.load(base)
        sc.trajectory_type = name
        p = IO.params_for(sc)
        p.do_intersection_check = 1
        s = M.Simulation(sc, params=p, backend=oracle_backend())
        s.step(40)
        out[name] = s.stats()
        s.close()
    for st in out.values():
        assert st["failed_solves"] == 0 and st["intersecting_steps"] == 0 and st["ccd_calls"] > 0


def test_detection_method_reaches_the_stepper(tmp_path):
    """A scene's distance_barrier_constraint.detection_method (default "bvh",
    SimState.cpp:178) sets the stepper's broad phase; a run over the oracle
    backend with the BVH takes the same steps as with the hash grid (same
    constraint sets; same earliest TOIs on this fixture)."""
    from stepper_util import oracle_backend
    from rigidipc_amd import sim as M
    import oracle_lib as O
    base = os.path.join(FIX, "unit-tests-5-cubes.json")
    sc, _ = IO.load(base)
    assert sc.detection_method == "bvh"
    assert IO.params_for(sc).detection_method == 1
    scene = json.load(open(base))
    scene.setdefault("distance_barrier_constraint", {})["detection_method"] = "hash_grid"
    p = tmp_path / "grid.json"
    p.write_text(json.dumps(scene))
    sc, _ = IO.load(str(p), base_dir=os.path.dirname(base))
    assert IO.params_for(sc).detection_method == 0
    scene["distance_barrier_constraint"]["detection_method"] = "brute_force"
    with pytest.raises(NotImplementedError, match="detection_method"):
        IO.load(scene, base_dir=os.path.dirname(base))
    out = {}
    try:
        for name in ("hash_grid", "bvh"):
            sc, _ = IO.load(base)
            sc.detection_method = name
            sc.trajectory_type = "rigid"
            s = M.Simulation(sc, params=IO.params_for(sc), backend=oracle_backend())
            s.step(5)
            out[name] = (s.state(), s.stats())
            s.close()
    finally:
        O.set_detection_method("hash_grid")
    np.testing.assert_array_equal(out["bvh"][0][0], out["hash_grid"][0][0])
    assert out["bvh"][1]["newton_iterations"] == out["hash_grid"][1]["newton_iterations"]
    assert out["bvh"][1]["ccd_calls"] > 0
