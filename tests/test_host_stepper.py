"""The host time stepper (librigc_sim.so) over the oracle contact backend:
the sparse block LDL^T against dense solves, closed-form implicit-Euler
motion of a free body, and contact invariants of a resting / sliding body
(no penetration, the barrier holds the body in its activation band,
friction decelerates a sliding body). These pin the Newton loop that the
device backend is compared against in tests/test_gpu_stepper.py."""
import numpy as np
import pytest

from stepper_util import free_body_scene, min_vertex_height, oracle_backend, resting_scene
from rigidipc_amd import scenes as S
from rigidipc_amd import sim as M


def random_block_system(rng, n, m, spd=True):
    nbk = 6 * n
    A = np.zeros((nbk, nbk))
    pairs = set()
    while len(pairs) < m:
        i, j = rng.integers(0, n, 2)
        if i != j:
            pairs.add((min(i, j), max(i, j)))
    pairs = sorted(pairs)
    off = []
    for i, j in pairs:
        B = rng.normal(size=(6, 6))
        off.append(B)
        A[6 * i:6 * i + 6, 6 * j:6 * j + 6] += B
        A[6 * j:6 * j + 6, 6 * i:6 * i + 6] += B.T
    diag = []
    for i in range(n):
        D = rng.normal(size=(6, 6))
        D = D @ D.T + (20.0 if spd else -2.0) * np.eye(6)
        diag.append(D)
        A[6 * i:6 * i + 6, 6 * i:6 * i + 6] += D
    return A, np.array(diag), np.array(pairs, np.int32).reshape(-1, 2), np.array(off).reshape(-1, 6, 6)


@pytest.mark.parametrize("n,m,spd", [(1, 0, True), (5, 4, True), (40, 60, True), (30, 40, False)])
def test_block_ldlt_matches_dense_solve(n, m, spd):
    rng = np.random.default_rng(n * 100 + m)
    A, diag, pairs, off = random_block_system(rng, n, m, spd)
    b = rng.normal(size=6 * n)
    x = M.block_ldlt_solve(diag, pairs, off, b)
    ref = np.linalg.solve(A, b)
    assert np.abs(x - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


def test_free_body_follows_implicit_euler():
    """Without contact the linear Newton step is exact:
    v+ = v + h g, x+ = x + h v+ (compute_body_energy, IMPLICIT_EULER)."""
    sc = free_body_scene()
    sim = M.Simulation(sc, backend=oracle_backend())
    x = sc.poses[0, :3].copy()
    v = sc.velocities[0, :3].copy()
    g = np.array([0, -9.81, 0])
    for _ in range(12):
        sim.step(1)
        v = v + sc.dt * g
        x = x + sc.dt * v
    poses, vel = sim.state()
    assert np.allclose(poses[0, :3], x, rtol=0, atol=1e-9)
    assert np.allclose(vel[0, :3], v, rtol=0, atol=1e-7)
    assert np.allclose(poses[0, 3:], sc.poses[0, 3:], atol=1e-12)
    assert np.array_equal(poses[1], sc.poses[1])  # static plane untouched


def test_free_rotation_keeps_angular_speed():
    """A spinning free cube: |omega| is conserved to first order by the
    rotational implicit-Euler energy (no contact, no gravity torque)."""
    sc = free_body_scene(w=(0.0, 3.0, 0.0), v=(0.0, 0.0, 0.0))
    sim = M.Simulation(sc, backend=oracle_backend())
    w0 = np.linalg.norm(sc.velocities[0, 3:])
    sim.step(20)
    _, vel = sim.state()
    assert abs(np.linalg.norm(vel[0, 3:]) - w0) <= 0.05 * w0
    assert sim.stats()["failed_solves"] == 0


def test_resting_body_stays_in_barrier_band():
    sc = resting_scene()
    sim = M.Simulation(sc, backend=oracle_backend())
    for _ in range(15):
        sim.step(1)
        poses, _ = sim.state()
        h = min_vertex_height(sc, poses)
        assert 0 < h < sc.dhat
    st = sim.stats()
    assert st["failed_solves"] == 0
    assert st["num_constraints"] > 0


def test_friction_decelerates_sliding_body():
    """A cube sliding at 1 m/s on the plane: without friction it keeps its
    speed; with mu = 0.5 it loses about mu g h per step once the lagged
    normal force has settled. The Newton velocity tolerance is tightened
    (the reference default, 1e-2 of the 28 m scene diagonal, accepts
    iterates 0.28 m/s away from the minimiser)."""
    v0 = np.array([1.0, 0.0, 0.0])
    speeds = {}
    for mu in (0.0, 0.5):
        sc = resting_scene(mu=mu, v=v0)
        params = M.default_params(sc)
        params.velocity_conv_tol = 1e-7
        sim = M.Simulation(sc, params=params, backend=oracle_backend())
        vx = []
        for _ in range(8):
            sim.step(1)
            poses, vel = sim.state()
            assert min_vertex_height(sc, poses) > 0
            vx.append(vel[0, 0])
        assert sim.stats()["failed_solves"] == 0
        speeds[mu] = np.array(vx)
    assert np.abs(speeds[0.0] - 1.0).max() <= 1e-6
    decel = -np.diff(speeds[0.5][3:]).mean()
    assert decel == pytest.approx(0.5 * 9.81 * 0.01, rel=0.3)


def intersecting_start_scene():
    """A cube sunk 0.2 into the static plane: its side edges cross the
    plane's triangles (an intersecting initial state)."""
    Vc, Fc = S.cube(1.0)
    body = S.Body(Vc, Fc, np.eye(3), np.array([4.0, 0.3, 1.0]), np.zeros(3), np.zeros(3))
    sc = S.assemble([body, S._ground()])
    sc.dt, sc.dhat, sc.mu, sc.gravity = 0.01, 1e-3, 0.0, np.array([0, -9.81, 0])
    return sc


def test_initial_state_is_checked_for_intersections():
    import oracle_lib as O
    sc = intersecting_start_scene()
    O.set_scene(sc)
    want = len(O.detect_intersections(sc.poses))
    sim = M.Simulation(sc, backend=oracle_backend())
    st = sim.stats()
    sim.close()
    assert want > 0 and st["initial_intersections"] == want
    assert st["intersections"] == -1 and st["intersecting_steps"] == 0
    sim = M.Simulation(resting_scene(), backend=oracle_backend())
    assert sim.stats()["initial_intersections"] == 0
    sim.close()


def test_intersection_check_per_step_leaves_the_motion_alone():
    runs = {}
    for check in (0, 1):
        sc = resting_scene(mu=0.5, v=(1.0, 0.0, 0.0))
        params = M.default_params(sc)
        params.do_intersection_check = check
        sim = M.Simulation(sc, params=params, backend=oracle_backend())
        sim.step(4)
        runs[check] = (sim.state()[0], sim.stats())
        sim.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    assert runs[0][1]["intersections"] == -1
    assert runs[1][1]["intersections"] == 0 and runs[1][1]["intersecting_steps"] == 0
