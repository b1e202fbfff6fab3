"""The reference's own 3D fixtures (copied to tests/golden/fixtures), loaded
with rigidipc_amd.scene_io and stepped by the host stepper over the device
path and over the oracle: the trajectories must agree as in
test_gpu_stepper (same Newton and CCD counts, poses within 1e-6, velocities
within 1e-4), covering stacks (5 cubes), a tet dropped into a static corner
(dhat 1e-4), edge-edge and face-vertex first contacts, sliding with
friction, a concave body on a static corner, and a kinematic plane moving
at its velocity into a cube pushed by a constant force, each with the
scene's
trajectory type (piecewise-linear CCD) and with the rigid CCD. The runs
check every step's end poses for intersections (do_intersection_check):
the counts of both backends agree, and no step ends intersecting.

The reference's kinematically-scripted.json is not among them: its
kinematic cube moves a metre per step into a falling cube, and the
augmented Lagrangian's thresholds (progress >= 0.999 / < 0.99) then flip
on rounding differences between the device's and the oracle's barrier
sums (one Newton iteration apart within three steps). test_kinematic.py
pins that scene on the CPU by its scripted poses."""
import os

import numpy as np
import pytest

from stepper_util import oracle_backend
from rigidipc_amd import scene_io as IO
from rigidipc_amd import sim as M

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")
CASES = [
    ("unit-tests-5-cubes.json", 80),
    ("unit-tests-tet-corner.json", 60),
    ("unit-tests-edge-edge.json", 80),
    ("unit-tests-face-vertex.json", 80),
    ("friction-sliding.json", 60),
    ("examples-concave.json", 80),
    ("examples-kinematic.json", 40),
]


def run(path, steps, backend, trajectory):
    sc, _ = IO.load(path)
    if trajectory is not None:
        sc.trajectory_type = trajectory
    params = IO.params_for(sc)
    params.do_intersection_check = 1
    s = M.Simulation(sc, params=params, backend=backend)
    s.step(steps)
    poses, vel = s.state()
    st = s.stats()
    s.close()
    return poses, vel, st


@pytest.mark.parametrize("trajectory", [None, "rigid"])
@pytest.mark.parametrize("name,steps", CASES)
def test_fixture_device_matches_oracle(name, steps, trajectory):
    """trajectory None: the scene's own (these fixtures leave it at the
    reference's default, piecewise_linear); "rigid": the rigid CCD."""
    path = os.path.join(FIX, name)
    print(f"fixture {name}: {steps} steps, trajectory {trajectory or 'scene'}", flush=True)
    pg, vg, sg = run(path, steps, None, trajectory)
    po, vo, so = run(path, steps, oracle_backend(), trajectory)
    print(f"newton {sg['newton_iterations']} / {so['newton_iterations']}, "
          f"max pose diff {np.abs(pg - po).max():.3e}", flush=True)
    assert sg["failed_solves"] == so["failed_solves"]
    assert sg["newton_iterations"] == so["newton_iterations"]
    assert sg["ccd_calls"] == so["ccd_calls"]
    assert np.abs(pg - po).max() <= 1e-6
    assert np.abs(vg - vo).max() <= 1e-4
    for k in ("initial_intersections", "intersections", "intersecting_steps"):
        assert sg[k] == so[k] == 0, (k, sg[k], so[k])
