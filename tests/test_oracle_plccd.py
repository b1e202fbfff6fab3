"""The oracle's piecewise-linear CCD (and its Tight-Inclusion restatement)
against what the mathematics fixes.

* Closed forms: the reference's translation scenes
  (tests/ccd/test_rigid_body_time_of_impact.cpp:170-303). The linear CCD
  finds an impact iff the elements meet, with the TOI within 1e-6 of the
  exact time and never above it; the piecewise-linear method reports every
  impact no later than the exact time.
* Conservativeness on rotating queries: an impact is reported no later than
  the first crossing of the exact rigid trajectories.
* Linear CCD alone: a vertex falling through a triangle and crossing edges
  with known linear-trajectory TOIs, and the minimum-separation shift.
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import oracle_lib as O  # noqa: E402
import plccd_cases as C  # noqa: E402
import scenes  # noqa: E402


def _linear_vertices(q, kind):
    X0 = C.world(q, kind, [0.0])[0]
    X1 = C.world(q, kind, [1.0])[0]
    return np.concatenate([X0, X1])


@pytest.mark.parametrize("kind", [0, 1])
def test_translation_closed_forms_linear(kind):
    """Tight-Inclusion alone on the translation scenes (linear trajectories,
    so the closed form is exact): impact iff the elements meet, TOI within
    1e-6 of the exact time and never above it."""
    cases = C.ee_translation_cases() if kind == 0 else C.fv_translation_cases()
    for q, want_hit, want in cases:
        hit, toi, _, _ = O.ti_linear(_linear_vertices(q, kind), kind == 1)
        assert hit == want_hit, (q[12:24], toi, want)
        if hit:
            assert toi <= want and toi == pytest.approx(want, abs=1e-6), (q[12:24], toi, want)


@pytest.mark.parametrize("kind", [0, 1])
def test_translation_closed_forms_piecewise(kind):
    """The piecewise-linear method on the same scenes. Its interval bound of
    a piece's deviation from a straight line does not see that a pure
    translation has none (the position interval appears on both sides), so
    it cuts the step into pieces, and it reports the start of a piece as an
    impact once the distance at that piece's start has fallen below a fifth
    of the initial one (time_of_impact.cpp:341-347, :529-541). So: every
    expected impact is reported, never later than the exact time, and a
    reported impact without contact is followed by a distance below a fifth
    of the initial one."""
    cases = C.ee_translation_cases() if kind == 0 else C.fv_translation_cases()
    fn = O.pl_ee_toi if kind == 0 else O.pl_fv_toi
    for q, want_hit, want in cases:
        hit, toi, _, _ = fn(q)
        if want_hit:
            assert hit and toi <= want, (q[12:24], toi, want)
        elif hit:
            # the distance is measured at a piece's start
            d0 = math.sqrt(O.element_distance_sq(kind, C.world(q, kind, [0.0])[0]))
            X = C.world(q, kind, np.linspace(toi, 1.0, 257))
            dmin = min(math.sqrt(O.element_distance_sq(kind, x)) for x in X)
            assert dmin < 0.2 * d0, (q[12:24], toi, d0, dmin)


def test_face_vertex_exit_measures_the_piece_start():
    """A vertex approaches a triangle that spins in its own plane and stops
    short of it at a fraction (1-19 %) of the initial distance, so the
    elements never touch. The spin axis passes through the vertex's foot on
    the triangle. The decreasing-distance exit reports the start t0
    of the current piece once the distance AT t0 is below a fifth of the
    initial one (time_of_impact.cpp:529-541: the face-vertex distance is
    taken at pose_ti0, as for edges). The distance along these trajectories
    is exact (the vertex translates along the normal of a plane that stays
    put), so every reported impact must lie where the true distance is
    already below a fifth of the initial one. Measuring at the piece's end
    instead reports the start of the piece in which the threshold is
    crossed, where the distance is still above it: that fails every case
    here."""
    rng = np.random.default_rng(0)
    hits = 0
    for _ in range(200):
        x, y, h0 = rng.uniform(-0.3, 0.3), rng.uniform(0.1, 0.4), rng.uniform(0.5, 2.0)
        frac = rng.uniform(0.01, 0.19)
        zero = [0.0] * 6
        # the triangle's body origin sits under the vertex, inside the
        # triangle, so the spin keeps the vertex's foot inside it
        verts = [[0, 0, h0], [-1 - x, -y, 0], [1 - x, -y, 0], [-x, 1 - y, 0]]
        pa1 = [0, 0, -(1 - frac) * h0, 0, 0, 0]
        pb1 = [0, 0, 0, 0, 0, rng.uniform(-0.5, 0.5)]
        q = np.array(sum(verts, []) + zero + pa1 + zero + pb1, np.float64)
        hit, toi, _, _ = O.pl_fv_toi(q)
        # exact distance: the vertex height above the plane z = 0
        d0, d_toi = h0, h0 * (1 - (1 - frac) * toi) if hit else None
        if hit:
            hits += 1
            assert 0 < toi <= 1 and d_toi < 0.2 * d0, (x, y, h0, frac, toi, d_toi / d0)
    assert hits >= 100


def test_resting_pair_at_the_separation_edge_terminates():
    """A pair resting msd * (1 + 1e-12) apart: the linear CCD reports an
    impact at the start of every piece, which the reference answers by
    halving the first piece again, without end. The oracle stops after 64
    halvings (65 pieces) with an impact at t = 0, the conservative answer.
    A pair 1.5 msd apart has no impact and needs one piece."""
    for kind, q, f in C.resting_cases():
        hit, toi, pieces, boxes = (O.pl_ee_toi if kind == 0 else O.pl_fv_toi)(q, 1.0, 1e-3)
        if f == 1.5:
            assert not hit and pieces == 1, (kind, f)
        else:
            assert hit and toi == 0.0 and pieces == 65, (kind, f, hit, toi, pieces)


@pytest.mark.parametrize("seed,rot", [(1, 0.3), (2, 1.0), (3, 2.5)])
def test_conservative_against_exact_crossings(seed, rot):
    kinds, qs = scenes.random_queries(150, seed, rot_scale=rot)
    crossings = 0
    for kind, q in zip(kinds, qs):
        tc = C.first_crossing(q, kind)
        hit, toi, pieces, boxes = (O.pl_ee_toi if kind == 0 else O.pl_fv_toi)(q)
        assert pieces >= 1
        if math.isfinite(tc):
            crossings += 1
            assert hit and toi <= tc + 1e-12, (kind, tc, hit, toi)
        if hit:
            assert 0 <= toi <= 1
    assert crossings >= 5


def test_earliest_toi_bounds_the_search():
    """Queries searched over [0, T] report only impacts at or before T."""
    kinds, qs = scenes.random_queries(30, 11, rot_scale=1.0)
    for kind, q in zip(kinds, qs):
        fn = O.pl_ee_toi if kind == 0 else O.pl_fv_toi
        hit1, toi1, _, _ = fn(q, 1.0)
        hit5, toi5, _, _ = fn(q, 0.5)
        if hit5:
            assert toi5 <= 0.5
        if hit1 and toi1 < 0.25:
            assert hit5


def test_linear_ccd_known_tois():
    # vertex falling from z = 1 to z = -1 through the triangle's interior: toi 1/2
    V = [[0.2, 0.2, 1.0], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.2, -1.0], [0, 0, 0], [1, 0, 0], [0, 1, 0]]
    hit, toi, _, _ = O.ti_linear(V, True)
    assert hit and 0.5 - 1e-6 <= toi <= 0.5
    # with separation 0.1 the vertex is reported once it is 0.1 away (toi
    # 0.45). The separation box around the origin is wide, so the breadth-first
    # search stops at its refinement cap (1e6 boxes) and reports the first box
    # of its last level together with that box's width as the precision
    # reached (output_tolerance)
    hit, toi, out_tol, boxes = O.ti_linear(V, True, ms=0.1)
    assert hit and boxes == 1000001 and 0.45 - out_tol <= toi <= 0.45 and out_tol < 1e-3
    # same fall beside the triangle: no impact
    W = [list(v) for v in V]
    W[0][0] = W[4][0] = 1.5
    assert not O.ti_linear(W, True)[0]
    # edges crossing at t = 1/4: edge a moves along y from 1 to -3 across edge b (the z axis)
    E = [[-1, 1, 0], [1, 1, 0], [0, 0, -1], [0, 0, 1], [-1, -3, 0], [1, -3, 0], [0, 0, -1], [0, 0, 1]]
    hit, toi, _, _ = O.ti_linear(E, False)
    assert hit and 0.25 - 1e-6 <= toi <= 0.25
    # parallel edges that never meet
    P = [[-1, 1, 0], [1, 1, 0], [-1, 0, 0], [1, 0, 0], [-1, 1, 1], [1, 1, 1], [-1, 0, 0], [1, 0, 0]]
    assert not O.ti_linear(P, False)[0]


def test_element_distance_closed_forms():
    X = np.array([[0, 0, 1.0], [-1, -1, 0], [3, -1, 0], [-1, 3, 0]])
    assert O.element_distance_sq(1, X) == 1.0
    X = np.array([[-1, 0, 0], [1, 0, 0], [0, 2, -1], [0, 2, 1]])
    assert O.element_distance_sq(0, X) == 4.0
