"""Pin the CPU oracle's interval arithmetic, root finder and rigid TOI against
the reference's own known-answer tests, restated here with their inputs and
expectations:

  tests/interval/test_interval.cpp:12-97          interval ops, cos/sin shifts
  tests/interval/test_interval_root_finder.cpp:9-77 quadratic / trig roots
  tests/utils/test_sinc.cpp:14-122                 double + interval sinc
  tests/ccd/test_rigid_body_time_of_impact.cpp:164-303 EE / FV analytic TOIs

plus checks the mathematics fixes (closed-form rotation TOIs, enclosure of
the true cosine/sine, the interval rotation matrix containing the exact one).
"""
import math

import numpy as np
import pytest

import oracle_lib as O

PI = math.pi
CATCH_EPS = 100 * float(np.finfo(np.float32).eps)


def catch_approx(value, expected, margin=0.0):
    """Catch2 v2 Approx(expected).margin(margin): an absolute margin OR a
    relative error of 100*FLT_EPSILON of the expected value."""
    d = abs(value - expected)
    return d <= margin or d <= CATCH_EPS * abs(expected)


# ---------------------------------------------------------------- intervals
def test_simple_interval_arithmetic():
    i, j = (0.0, 1.0), (4.0, 5.0)
    for op in "+-*/":
        r = O.binop(op, i, j)
        assert r[0] <= r[1]
    r = O.binop("+", i, (10.0, 10.0))
    assert abs(r[0] - 10) <= 1e-12 and abs(r[1] - 11) <= 1e-12


def test_interval_ops_enclose_and_widen_one_ulp():
    # The filib rule: round to nearest, then step one ulp outward; exact
    # zero results are special-cased (filib_rounding.cpp:12-50).
    rng = np.random.default_rng(0)
    for _ in range(2000):
        a, b = np.sort(rng.normal(size=2) * 10.0 ** rng.integers(-3, 4))
        c, d = np.sort(rng.normal(size=2) * 10.0 ** rng.integers(-3, 4))
        for op, f in (("+", lambda x, y: x + y), ("-", lambda x, y: x - y), ("*", lambda x, y: x * y)):
            r = O.binop(op, (a, b), (c, d))
            vals = [f(x, y) for x in (a, b) for y in (c, d)]
            assert r[0] == np.nextafter(min(vals), -np.inf)
            assert r[1] == np.nextafter(max(vals), np.inf)
    # zero special cases
    r = O.binop("-", (1.5, 1.5), (1.5, 1.5))
    assert r[0] == 0.0 and r[1] == 0.0
    r = O.binop("*", (0.0, 0.0), (-3.0, 7.0))
    assert r[0] == 0.0 and r[1] == 0.0


@pytest.mark.parametrize("shift", [0.0, 2 * PI, -2 * PI, 100 * PI, -100 * PI])
def test_cosine_interval(shift):
    r = O.interval_fn("oracle_interval_cos", *O.binop("+", (-1.0, 7.0), (shift, shift)))
    assert r[0] == -1.0 and r[1] == 1.0
    r = O.interval_fn("oracle_interval_cos", *O.binop("+", (2.0, 4.0), (shift, shift)))
    assert r[1] < 0
    r = O.interval_fn("oracle_interval_cos", *O.binop("+", (0.0, 1.0), (shift, shift)))
    assert r[0] > 0
    r = O.interval_fn("oracle_interval_cos", *O.binop("+", (1.0, 2.0), (shift, shift)))
    assert r[0] < 0 and r[1] > 0


@pytest.mark.parametrize("shift", [0.0, 2 * PI, -2 * PI, 100 * PI, -100 * PI])
def test_sine_interval(shift):
    shift = shift + PI / 2
    r = O.interval_fn("oracle_interval_sin", *O.binop("+", (-1.0, 7.0), (shift, shift)))
    assert r[0] == -1.0 and r[1] == 1.0
    r = O.interval_fn("oracle_interval_sin", *O.binop("+", (2.0, 4.0), (shift, shift)))
    assert r[1] < 0
    r = O.interval_fn("oracle_interval_sin", *O.binop("+", (0.0, 1.0), (shift, shift)))
    assert r[0] > 0
    r = O.interval_fn("oracle_interval_sin", *O.binop("+", (1.0, 2.0), (shift, shift)))
    assert r[0] < 0 and r[1] > 0


def test_cos_sin_enclose_true_values():
    # Enclosure of the mathematical cos/sin (checked in float64 with a
    # margin far below the 1e-15 widening, and exactly at sample points).
    rng = np.random.default_rng(1)
    for _ in range(3000):
        c = rng.uniform(-60, 60)
        w = 10.0 ** rng.uniform(-12, 0.5)
        lo, hi = c - w, c + w
        xs = np.linspace(lo, hi, 9)
        rc = O.interval_fn("oracle_interval_cos", lo, hi)
        rs = O.interval_fn("oracle_interval_sin", lo, hi)
        assert rc[0] <= rc[1] and rs[0] <= rs[1]
        assert np.all(np.cos(xs) >= rc[0] - 1e-16) and np.all(np.cos(xs) <= rc[1] + 1e-16)
        assert np.all(np.sin(xs) >= rs[0] - 1e-16) and np.all(np.sin(xs) <= rs[1] + 1e-16)
        # and are not grossly wide
        assert rc[1] - rc[0] <= (np.ptp(np.cos(xs)) + 4 * w + 1e-14) or rc[1] - rc[0] <= 2


def test_restated_cos_sin_accuracy():
    xs = np.linspace(0, 4, 20001)
    err = max(abs(O.lib().oracle_cos_rn(float(x)) - math.cos(x)) for x in xs)
    assert err < 4e-16
    xs = np.concatenate([np.linspace(-200, 200, 20001), [1e-9, -3e-5, 1e3]])
    err = max(abs(O.lib().oracle_sin_rn(float(x)) - math.sin(x)) / max(1.0, abs(math.sin(x))) for x in xs)
    assert err < 1e-15


# -------------------------------------------------------------------- sinc
def test_double_sinc():
    s = O.lib().oracle_sinc
    assert s(0.0) == pytest.approx(1)
    assert s(1e-8) == pytest.approx(1)
    assert s(PI / 2) == pytest.approx(2 / PI)
    assert abs(s(PI)) <= 1e-16


@pytest.mark.parametrize(
    "x,expected",
    [((0, 0), (1, 1)), ((0, 1e-8), (1, 1)), ((-1e-8, 0), (1, 1)), ((-1e-8, 1e-8), (1, 1)),
     ((0, PI), (0, 1)), ((-PI, 0), (0, 1)), ((-PI, PI), (0, 1)),
     ((0, PI / 2), (2 / PI, 1)), ((-PI / 2, 0), (2 / PI, 1)), ((-PI / 2, PI / 2), (2 / PI, 1))],
)
def test_interval_sinc(x, expected):
    y = O.interval_fn("oracle_interval_sinc", float(x[0]), float(x[1]))
    assert catch_approx(y[0], expected[0], 1e-8)
    assert catch_approx(y[1], expected[1], 1e-8)


def _sinc(x):
    return math.sin(x) / x if x != 0 else 1.0


@pytest.mark.parametrize("x,expected", [((-5, 5), (-0.217233628211, 1)), ((-1, 5), (-0.217233628211, 1)),
                                        ((0, 5), (-0.217233628211, 1)), ((5, 7), None), ((21, 22), None)])
def test_interval_sinc_looser_bounds(x, expected):
    if expected is None:
        a, b = _sinc(x[0]), _sinc(x[1])
        expected = (min(a, b), max(a, b))
    y = O.interval_fn("oracle_interval_sinc", float(x[0]), float(x[1]))
    assert y[0] <= expected[0] and expected[1] <= y[1]
    assert -0.23 <= y[0] and y[1] <= 1
    if x[0] > 0:
        assert -1 / x[0] <= y[0] and y[1] <= 1 / x[0]


def test_interval_sinc_encloses_samples():
    rng = np.random.default_rng(2)
    for _ in range(2000):
        lo = rng.uniform(0, 12)
        hi = lo + 10.0 ** rng.uniform(-10, 0.7)
        y = O.interval_fn("oracle_interval_sinc", lo, hi)
        for x in np.linspace(lo, hi, 7):
            assert y[0] - 1e-15 <= _sinc(x) <= y[1] + 1e-15


# ------------------------------------------------------------- root finder
@pytest.mark.parametrize("yshift", [-1.1, -1.0 - 10 * 1e-8, -1.0, -0.5, -1e-4, 0.5, 1.0])
def test_root_of_simple_function(yshift):
    a, b, c = 4.0, -4.0, 1 + yshift
    out = np.zeros(2)
    found = O.lib().oracle_root_quadratic(a, b, c, 1e-8, 0, 0.0, O.dptr(out))
    assert bool(found) == (-1.0 <= yshift <= 0.0)
    if found:
        actual = (-b - math.sqrt(b * b - 4 * a * c)) / (2 * a)
        assert catch_approx(0.5 * (out[0] + out[1]), actual, 1e-8)


def test_root_of_trig_function():
    out = np.zeros(2)
    assert O.lib().oracle_root_trig(2 * PI, 0.0, 1e-8, O.dptr(out))
    assert catch_approx(0.5 * (out[0] + out[1]), 0.25, 1e-8)


@pytest.mark.parametrize("yshift", [-1.1, -1.0 - 1e-7, -1.0, -0.5, 0.0, 0.5, 1.0])
def test_root_with_constraint(yshift):
    a, b, c = 4.0, -4.0, 1 + yshift
    out = np.zeros(2)
    found = O.lib().oracle_root_quadratic(a, b, c, 1e-8, 1, 0.5, O.dptr(out))
    assert bool(found) == (-1.0 <= yshift <= 0.0)
    if found:
        actual = (-b + math.sqrt(b * b - 4 * a * c)) / (2 * a)
        assert catch_approx(0.5 * (out[0] + out[1]), actual, 1e-8)


# --------------------------------------------------------- rotation matrix
def _rodrigues(r):
    th = np.linalg.norm(r)
    K = np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])
    if th == 0:
        return np.eye(3)
    return np.eye(3) + math.sin(th) / th * K + (1 - math.cos(th)) / th ** 2 * K @ K


def test_interval_rotation_contains_exact():
    rng = np.random.default_rng(3)
    for _ in range(500):
        r = rng.normal(size=3) * rng.choice([1e-6, 0.1, 1.0, 4.0])
        w = 10.0 ** rng.uniform(-9, -1)
        box = np.stack([r - w, r + w], axis=1).ravel()
        out = np.zeros(18)
        O.lib().oracle_rotation_interval(O.dptr(box), O.dptr(out))
        lo, hi = out[0::2].reshape(3, 3), out[1::2].reshape(3, 3)
        for _ in range(4):
            rr = r + rng.uniform(-w, w, size=3)
            R = _rodrigues(rr)
            assert np.all(R >= lo - 1e-14) and np.all(R <= hi + 1e-14)
        Rd = np.zeros(9)
        O.lib().oracle_rotation_double(O.dptr(np.ascontiguousarray(r)), O.dptr(Rd))
        assert np.allclose(Rd.reshape(3, 3), _rodrigues(r), atol=1e-14)


# ------------------------------------------------------------- rigid TOI
def ee_query(a0, a1, b0, b1, pa0, pa1, pb0, pb1):
    return np.concatenate([a0, a1, b0, b1, pa0, pa1, pb0, pb1]).astype(np.float64)


Z6 = np.zeros(6)


@pytest.mark.parametrize("z", [1000.0, 10.0, 1.0 + 1e-8, 1.0, 1.0 - 1e-8, 0.5, 0.0])
@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("y_t1", [2.0, 1.0, 0.1, 1e-6, 0.0, -1e-6, -1, -2, -10, -1000])
def test_rigid_edge_edge_toi(z, sign, y_t1):
    # test_rigid_body_time_of_impact.cpp:164-234
    z_t0 = sign * z
    pa0 = np.array([0, 1.0, z_t0, 0, 0, 0])
    pa1 = np.array([0, y_t1, z_t0, 0, 0, 0])
    q = ee_query([-1, 0, 0], [1, 0, 0], [0, 0, -1], [0, 0, 1], pa0, pa1, Z6, Z6)
    expected = 1 / (1 - y_t1) if y_t1 != 1 else math.inf  # C++ 1/0 = inf
    expect_hit = y_t1 <= 0.0 and -1 <= z_t0 <= 1
    hit, toi, _ = O.ee_toi(q, 1.0, 1e-6)
    assert hit == expect_hit
    if hit:
        assert catch_approx(toi, expected, 1e-4)
        assert toi <= expected


@pytest.mark.parametrize("y", [1.0 + 1e-8, 1.0, 1.0 - 1e-8, 0.5, 1e-8, 0.0])
@pytest.mark.parametrize("z_t1", [1, 1e-8, 0.0, -1e-8, -2.0, -10])
def test_rigid_face_vertex_toi(y, z_t1):
    # test_rigid_body_time_of_impact.cpp:236-303: the vertex body B moves
    # its vertex (0, y, 1) to z = z_t1 through a static triangle at z = 0.
    pb1 = np.array([0, 0, z_t1 - 1.0, 0, 0, 0])
    q = np.concatenate([[0, y, 1], [-1, 0, 0], [1, 0, 0], [0, 1, 0], Z6, pb1, Z6, Z6]).astype(np.float64)
    expected = 1 / (-z_t1 + 1) if z_t1 != 1 else math.inf
    expect_hit = z_t1 <= 0.0 and 0 <= y <= 1.0
    hit, toi, _ = O.fv_toi(q, 1.0, 1e-6)
    assert hit == expect_hit
    if expect_hit:
        assert catch_approx(toi, expected, 1e-4)
        assert toi <= expected


@pytest.mark.parametrize("phi", [0.3, 1.0, 2.0, 3.0])
@pytest.mark.parametrize("theta1", [3.5, 6.0, 20.0])
def test_rotating_edge_hits_edge_closed_form(phi, theta1):
    # Edge A = [0.5, 2] along x spins about z by theta1*t; edge B is a
    # vertical segment through (cos phi, sin phi). First contact at
    # t = phi / theta1 (theta1 > phi), exercising interval sin/cos/sinc.
    pa0 = np.zeros(6)
    pa1 = np.array([0, 0, 0, 0, 0, theta1])
    b = [math.cos(phi), math.sin(phi)]
    q = ee_query([0.5, 0, 0], [2.0, 0, 0], [b[0], b[1], -1], [b[0], b[1], 1], pa0, pa1, Z6, Z6)
    hit, toi, _ = O.ee_toi(q)
    expected = phi / theta1
    assert hit
    assert toi <= expected
    # time tolerance 1e-4, plus the length tolerance mapped to time
    assert expected - toi <= 1e-4 + 2e-4 / theta1


def test_rotating_edge_misses_when_angle_too_small():
    pa1 = np.array([0, 0, 0, 0, 0, 0.5])
    q = ee_query([0.5, 0, 0], [2.0, 0, 0], [math.cos(1.0), math.sin(1.0), -1],
                 [math.cos(1.0), math.sin(1.0), 1], Z6, pa1, Z6, Z6)
    hit, toi, _ = O.ee_toi(q)
    assert not hit and toi == math.inf
