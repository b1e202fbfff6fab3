"""The host stepper over the device backend against the same stepper over
the oracle backend: identical Newton loop, contact terms from the device
(bit-exact CCD and constraint sets, derivatives within 1e-10 / 1e-9), so
the trajectories agree to rounding (1e-8 m / 1e-8 rad after several
steps with contact, friction and CCD-limited line searches)."""
import numpy as np
import pytest

from stepper_util import min_vertex_height, oracle_backend, resting_scene
from rigidipc_amd import scenes as S
from rigidipc_amd import sim as M
from test_oracle_barrier import contact_scene

pytestmark = pytest.mark.gpu


def run(sc, steps, backend, tol=None):
    params = M.default_params(sc)
    if tol is not None:
        params.velocity_conv_tol = tol
    sim = M.Simulation(sc, params=params, backend=backend)
    sim.step(steps)
    poses, vel = sim.state()
    st = sim.stats()
    sim.close()
    return poses, vel, st


def drop_scene():
    """The cube / sphere / torus contact scene under gravity, mu = 0.3 and
    dhat = 0.12, so that every constraint kind is active from the start."""
    sc = contact_scene(0)
    sc.dt, sc.dhat, sc.mu, sc.gravity = 0.01, 0.12, 0.3, np.array([0, -9.81, 0])
    return sc


CASES = {
    "config1": (lambda: S.config1(seed=0), 10, None),
    "sliding_cube": (lambda: resting_scene(mu=0.5, v=(1.0, 0.0, 0.0)), 6, 1e-7),
    "drop": (drop_scene, 8, None),
    # the differentiation scene's forward rollout: 64 spheres dropped on
    # the plane, mu 0.5, its 20 steps
    "config4": (lambda: S.config4(0), 20, None),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_stepper_matches_oracle_stepper(name):
    make, steps, tol = CASES[name]
    print(f"stepper: {name}, {steps} steps", flush=True)
    pg, vg, sg = run(make(), steps, None, tol)
    po, vo, so = run(make(), steps, oracle_backend(), tol)
    assert sg["failed_solves"] == so["failed_solves"] == 0
    assert np.abs(pg - po).max() <= 1e-6
    assert np.abs(vg - vo).max() <= 1e-4
    assert sg["newton_iterations"] == so["newton_iterations"]
    assert sg["ccd_calls"] == so["ccd_calls"]
    sc = make()
    assert min_vertex_height(sc, pg) > 0


def test_config2_steps_without_penetration():
    """A few steps of the 1,000-body scene on the device: no failed solve,
    CCD keeps every body above the plane."""
    sc = S.config2(seed=0)
    poses, _, st = run(sc, 3, None)
    assert st["failed_solves"] == 0
    assert min_vertex_height(sc, poses) > 0
    assert st["ccd_candidates"] > 0


def test_intersection_checks_match_the_oracle_stepper():
    """The stepper's intersection checks through the device backend: the
    initial state's count and, with do_intersection_check, every step's."""
    from test_host_stepper import intersecting_start_scene
    counts = []
    for backend in (None, oracle_backend()):
        sim = M.Simulation(intersecting_start_scene(), backend=backend)
        first = sim.stats()["initial_intersections"]
        sim.close()
        sc = S.config1(seed=0)
        params = M.default_params(sc)
        params.do_intersection_check = 1
        sim = M.Simulation(sc, params=params, backend=backend)
        sim.step(5)
        st = sim.stats()
        sim.close()
        counts.append((first, st["initial_intersections"], st["intersections"], st["intersecting_steps"]))
    print("intersection counts (device, oracle):", counts, flush=True)
    assert counts[0] == counts[1]
    assert counts[0][0] > 0 and counts[0][1:] == (0, 0, 0)
