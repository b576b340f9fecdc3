"""Pins of the oracle's GJK, EPA and box-box arithmetic (CPU only).

Three kinds of check, each chosen so that a plausible transcription mistake
in the oracle fails at least one of them (tests/plant_oracle.py plants such
mistakes one at a time and runs these tests against each):

* the reference's own helper-level and adversarial tests, restated
  (gjk.rs:1233-1537 and :1703-2093, epa.rs:480-738);
* exact answers from geometry alone: the penetration depth and the distance
  of the core Minkowski difference, computed in float64 from its convex hull
  (scipy's Qhull over every core-point difference), and closed-form
  triangle distances;
* constructed ties whose outcome the reference documents: find_best keeps
  the first minimum (epa.rs:212-224), horizon edges cancel by
  Vec::swap_remove and stitch in horizon order (epa.rs:333-386), cull_points
  keeps the last deepest point and sorts stably (box_box.rs:126-174), the
  clip filter keeps dep >= 0 (box_box.rs:501), EPA seeds with the six axis
  supports in the order X, -X, Y, -Y, Z, -Z (epa.rs:150-172).
"""
import math

import numpy as np
import pytest

I = (0.0, 0.0, 0.0, 1.0)
AXES = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)


def finite(v):
    return bool(np.all(np.isfinite(np.asarray(v, np.float64))))


def dist(p, q):
    """|p - q| with both sides taken as f32 vectors, as the reference's
    Vector3::distance compares them."""
    return float(np.linalg.norm(np.asarray(p, np.float32).astype(np.float64) - np.asarray(q, np.float32)))


# ------------------------------------------------------------------ gjk.rs --
# closest_origin helpers (gjk.rs:1233-1429)
def test_closest_segment_degenerate_lengths_return_endpoints(oracle):  # gjk.rs:1233-1261
    a = (1.0, 2.0, 3.0)
    q, keep = oracle.closest_origin([a, a])
    assert finite(q) and keep.shape[0] == 1 and dist(keep[0], a) < 1e-9
    tiny = 1e-30  # 1 + 1e-30 rounds to 1 in f32: the exact-zero guard
    q, _ = oracle.closest_origin([(1, 0, 0), (1.0 + tiny, 0, 0)])
    assert finite(q) and dist(q, (1, 0, 0)) < 1e-9
    q, _ = oracle.closest_origin([(-1, 0, 0), (-1.0 + tiny, 0, 0)])
    assert finite(q) and dist(q, np.float32(-1.0 + tiny) * np.array([1, 0, 0])) < 1e-9
    q, _ = oracle.closest_origin([(1, 0, 0), (1.0 + 1e-20, 0, 0)])
    assert finite(q) and dist(q, (1, 0, 0)) < 1e-9


def test_closest_segment_well_conditioned_preserved(oracle):  # gjk.rs:1263-1278
    q, keep = oracle.closest_origin([(2, 0, 0), (3, 0, 0)])
    assert dist(q, (2, 0, 0)) < 1e-9 and keep.shape[0] == 1
    q, _ = oracle.closest_origin([(-3, 0, 0), (-2, 0, 0)])
    assert dist(q, (-2, 0, 0)) < 1e-9
    q, keep = oracle.closest_origin([(1, 1, 0), (2, -1, 0)])
    assert dist(q, (1.2, 0.6, 0.0)) < 1e-6 and keep.shape[0] == 2
    # keep order is (b, a) (gjk.rs:904)
    assert np.array_equal(keep, np.array([[2, -1, 0], [1, 1, 0]], np.float32))


def test_closest_triangle_duplicate_vertices_stay_finite(oracle):  # gjk.rs:1280-1312
    shared = np.array([0.5, 0.3, 0.0])
    q, keep = oracle.closest_origin([shared, shared, (-2.0, 0.3, 0.0)])
    assert finite(q) and keep.shape[0] == 1 and dist(q, shared) < 1e-9
    q, keep = oracle.closest_origin([(0.5, 0, 0), (-0.5, 0, 0), (0.5, 0, 0)])
    assert finite(q) and np.linalg.norm(q) < 1e-7 and keep.shape[0] == 2
    q, _ = oracle.closest_origin([(0, 0.5, 0), (0.3, -0.5, 0), (0.3, -0.5, 0)])
    assert finite(q) and dist(q, (0.1376, 0.0413, 0.0)) < 1e-3
    # closed form of the b == c case: the foot of the origin on the line a-b
    a, b = np.array([0, 0.5, 0]), np.array([0.3, -0.5, 0])
    t = -a @ (b - a) / ((b - a) @ (b - a))
    assert np.linalg.norm(q - (a + t * (b - a))) < 1e-7
    p = (0.7, -0.2, 1.0)
    q, keep = oracle.closest_origin([p, p, p])
    assert finite(q) and keep.shape[0] == 1 and dist(q, p) < 1e-9


def test_closest_triangle_collinear_and_sliver_stay_finite(oracle):  # gjk.rs:1314-1349
    q, _ = oracle.closest_origin([(3, 1, 0), (0.5, 1, 0), (-3, 1, 0)])
    assert finite(q) and dist(q, (0, 1, 0)) < 1e-7
    q, _ = oracle.closest_origin([(1, 1, 0), (2, 1, 0), (3, 1, 0)])
    assert finite(q) and dist(q, (1, 1, 0)) < 1e-7
    q, _ = oracle.closest_origin([(3, 1, 0), (0.5, 1, 1e-7), (-3, 1, 0)])
    assert finite(q) and dist(q, (0, 1, 0)) < 1e-5
    q, _ = oracle.closest_origin([(-1, 1e-6, 0), (1, 1e-6, 0), (0, -1e-6, 0)])
    assert finite(q) and np.linalg.norm(q) < 2e-6


def test_closest_triangle_well_conditioned_preserved(oracle):  # gjk.rs:1351-1371
    q, keep = oracle.closest_origin([(2, 2, 0), (-2, 2, 0), (0, -2, 0)])
    assert np.linalg.norm(q) < 1e-7 and keep.shape[0] == 3
    q, keep = oracle.closest_origin([(1, 0, 0), (1, 2, 0), (3, 1, 0)])
    assert dist(q, (1, 0, 0)) < 1e-7 and keep.shape[0] == 1
    q, keep = oracle.closest_origin([(1, -1, 0), (1, 1, 0), (3, 0, 0)])
    assert dist(q, (1, 0, 0)) < 1e-7 and keep.shape[0] == 2


def test_closest_tetrahedron_degenerate_and_analytic_cases(oracle):  # gjk.rs:1373-1404
    a, b, c, d = (1, 0, 1), (-1, 0, 1), (0, 1, 1), (0, -1, 1)
    q, keep = oracle.closest_origin([a, b, c, d])
    assert finite(q) and dist(q, (0, 0, 1)) < 1e-7 and keep.shape[0] >= 1
    q, keep = oracle.closest_origin([(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)])
    assert float(q @ q) == 0.0 and keep.shape[0] == 4
    q, keep = oracle.closest_origin([a, b, c, (0, -1, 2)])
    assert finite(q) and dist(q, (0, 0, 1)) < 1e-7 and keep.shape[0] == 2


def test_reduce_degenerate_support_buffers_stay_finite(oracle):  # gjk.rs:1406-1478
    # buf[0] is the newest vertex (gjk.rs:352-364 reads a = buf[len - 1])
    inside, v, pb, n = oracle.reduce_support([(-0.5, 0, 0), (0.5, 0, 0), (0.5, 0, 0)],
                                             [(-0.2, 0, 0), (0.2, 0, 0), (0.2, 0, 0)])
    assert not inside and finite(v) and n == 1
    assert dist(v, (0.5, 0, 0)) < 1e-7 and dist(pb, (0.2, 0, 0)) < 1e-7
    inside, v, pb, n = oracle.reduce_support([(-2, 1, 0), (0, 1, 0), (2, 1, 0)], [(0, 0, 0), (1, 1, 1), (0, 0, 0)])
    assert not inside and n == 1 and dist(v, (0, 1, 0)) < 1e-7 and dist(pb, (1, 1, 1)) < 1e-7
    inside, v, pb, n = oracle.reduce_support([(1, 0, 1), (-1, 0, 1), (0, 1, 1), (0, -1, 1)], np.zeros((4, 3)))
    assert not inside and n <= 3 and dist(v, (0, 0, 1)) < 1e-7 and finite(pb)
    inside, _, _, _ = oracle.reduce_support([(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)], np.zeros((4, 3)))
    assert inside


def test_reduce_segment_interpolates_the_witness(oracle):  # gjk.rs:366-393 (closed form)
    # a = buf[1], b = buf[0]; t = -a.(b - a) / |b - a|^2
    a, b = np.array([1.0, -1.0, 0.5]), np.array([1.0, 3.0, 0.5])
    pa_, pb_ = np.array([0.0, 0.0, 0.0]), np.array([4.0, 8.0, 0.0])
    inside, v, pb, n = oracle.reduce_support([b, a], [pb_, pa_])
    t = -a @ (b - a) / ((b - a) @ (b - a))
    assert not inside and n == 2
    assert np.linalg.norm(v - (a + t * (b - a))) < 1e-6 and np.linalg.norm(pb - (pa_ + t * (pb_ - pa_))) < 1e-6


def test_gjk_intersect_coincident_shapes_report_overlap(oracle):  # gjk.rs:1480-1509
    for s in (oracle.sphere(0.5), oracle.box(0.5)):
        n, simp = oracle.gjk_intersect(s, (0, 0, 0), I, s, (0, 0, 0), I)
        assert n == 4 and finite(simp)


def test_gjk_contact_coincident_point_cores_terminate_finitely(oracle):  # gjk.rs:1511-1536
    for s in (oracle.sphere(0.5), oracle.capsule(0.5, 0.5)):
        k, dist, nrm, _, _ = oracle.gjk_contact(s, (0, 0, 0), I, s, (0, 0, 0), I, 0.0173)
        if k == 1:
            assert math.isfinite(dist) and finite(nrm)


CUBE8 = np.array([[x, y, z] for z in (-1, 1) for y in (-1, 1) for x in (-1, 1)], np.float32)


@pytest.mark.parametrize("scale", [1e-3, 1e-2, 1.0, 1e2])
def test_gjk_duplicate_support_gate_across_scales(oracle, scale):  # gjk.rs:1703-1733
    e = np.float32(0.5 * scale)
    hull = oracle.hull(CUBE8 * e)
    k, dist, nrm, _, _ = oracle.gjk_contact(hull, (0, 0, 0), I, oracle.sphere(0.5 * scale),
                                            (1.01 * scale, 0, 0), I, 0.02 * scale)
    assert k == 1
    assert abs(dist - 0.01 * scale) < max(1e-4 * scale, 1e-9)
    assert finite(nrm) and abs(nrm[0]) > 0.999


SEEDS = [None, (1, 0, 0), (-1, 0, 0), (0, 1, 0), tuple(np.array([1.0, 7.0, 3.0]) / np.linalg.norm([1.0, 7.0, 3.0]))]


def test_gjk_seeded_vs_unseeded_stability(oracle):  # gjk.rs:1735-1808
    a = oracle.sphere(0.5)
    ref = None
    for seed in SEEDS:
        k, dist, nrm, _, _ = oracle.gjk_contact(a, (0, 0, 0), I, a, (1.01, 0, 0), I, 0.0173, seed=seed)
        assert k == 1 and abs(nrm[0] - 1.0) < 1e-6
        ref = dist if ref is None else ref
        assert abs(dist - ref) < 1e-9
    bx = oracle.box(0.5)
    ref = None
    for seed in SEEDS:
        k, dist, nrm, _, _ = oracle.gjk_contact(bx, (0, 0, 0), I, bx, (1.02, 0, 0), I, 0.05, seed=seed)
        assert k == 1 and abs(nrm[0]) > 0.999
        ref = dist if ref is None else ref
        assert abs(dist - ref) < 1e-5
    for seed in SEEDS:
        k, _, _, _, simp = oracle.gjk_contact(bx, (0, 0, 0), I, bx, (0.6, 0, 0), I, 0.02, seed=seed)
        assert k == 2 and finite(simp)
        k, _, _, _, _ = oracle.gjk_contact(bx, (0, 0, 0), I, bx, (3.0, 0, 0), I, 0.02, seed=seed)
        assert k == 0


def test_gjk_degenerate_segment_hulls_terminate_finitely(oracle):  # gjk.rs:1810-1867
    seg_x = oracle.hull(np.array([[-1, 0, 0], [1, 0, 0]], np.float32))
    seg_y = oracle.hull(np.array([[0, -1, 0], [0, 1, 0]], np.float32))
    n, _ = oracle.gjk_intersect(seg_x, (0, 0, 0), I, seg_y, (0, 0, 0), I)
    assert n > 0
    k, dist, nrm, _, simp = oracle.gjk_contact(seg_x, (0, 0, 0), I, seg_y, (0, 0, 0), I, 0.0173)
    assert k != 0
    if k == 2:
        assert finite(simp)
    else:
        assert abs(dist) < 1e-3 and finite(nrm)
    n, _ = oracle.gjk_intersect(seg_x, (0, 0, 0), I, seg_x, (0, 2, 0), I)
    assert n == 0
    assert oracle.gjk_contact(seg_x, (0, 0, 0), I, seg_x, (0, 2, 0), I, 0.0173)[0] == 0
    oracle.gjk_intersect(seg_x, (0, 0, 0), I, seg_x, (0.5, 0, 0), I)
    k, dist, nrm, _, _ = oracle.gjk_contact(seg_x, (0, 0, 0), I, seg_x, (0.5, 0, 0), I, 0.0173)
    if k == 1:
        assert math.isfinite(dist) and finite(nrm)


def test_gjk_elongated_rod_vs_box_stays_finite(oracle):  # gjk.rs:1869-1924
    e = np.float32(1e-3)
    rod = oracle.hull(np.array([[sx * e, sy, sz * e] for sz in (-1, 1) for sy in (-1, 1) for sx in (-1, 1)],
                               np.float32))
    b = oracle.box(0.5)
    k, dist, nrm, _, _ = oracle.gjk_contact(rod, (0, 0, 0), I, b, (0, 1.51, 0), I, 0.0173)
    assert k == 1 and abs(dist - 0.01) < 1e-4 and finite(nrm) and abs(nrm[1]) > 0.999
    k, dist, _, _, _ = oracle.gjk_contact(rod, (0, 0, 0), I, b, (0, 1.50, 0), I, 0.0173)
    assert k == 1 and abs(dist) < 1e-5
    k, _, _, _, simp = oracle.gjk_contact(rod, (0, 0, 0), I, b, (0, 1.45, 0), I, 0.0173)
    assert k == 2 and finite(simp)
    assert oracle.gjk_contact(rod, (0, 0, 0), I, b, (0, 2.5, 0), I, 0.0173)[0] == 0


@pytest.mark.parametrize("v", [(0.5, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (2, 1, 1),
                               (1e-4, 0, 0), (0, 3, 0), (-0.2, -0.2, -0.2), (0, 0, 0)])
def test_perpendicular_always_finite_and_orthogonal(oracle, v):  # gjk.rs:1936-1965
    p = oracle.perpendicular(v).astype(np.float64)
    assert finite(p) and 0.5 < p @ p < 1.5
    if np.dot(v, v) > 0:
        assert abs(p @ np.asarray(v, np.float64)) < 1e-5


def test_perpendicular_of_nan_is_finite(oracle):  # gjk.rs:1962-1964
    assert finite(oracle.perpendicular((float("nan"), 0, 0)))


def test_solve_reductions_produce_finite_directions(oracle):  # gjk.rs:1967-1987
    enclosed, d, _ = oracle.solve([(2, 0, 0), (2, 0, 0)])
    assert not enclosed and finite(d)
    enclosed, d, _ = oracle.solve([(3, 0, 0), (1, 0, 0), (-2, 0, 0)])
    assert not enclosed and finite(d)


def test_solve_segment_points_toward_the_origin(oracle):  # gjk.rs:624-646 (closed form)
    # a = older point, b = newer; origin beside the segment -> dir = (ab x ao) x ab
    enclosed, d, n = oracle.solve([(1, -1, 0), (1, 1, 0)])
    assert not enclosed and n == 2
    a, b = np.array([1, -1, 0.0]), np.array([1, 1, 0.0])
    ab, ao = b - a, -a
    assert np.allclose(d, np.cross(np.cross(ab, ao), ab), atol=1e-6)
    # origin behind a: the simplex shrinks to a and dir = -a
    enclosed, d, n = oracle.solve([(2, 0, 0), (3, 0, 0)])
    assert n == 1 and np.allclose(d, (-2, 0, 0))


def test_enclosing_walk_never_queries_with_zero_direction(oracle):  # gjk.rs:1989-2021
    found, q = oracle.enclosing_walk_scripted([(2, 0, 0), (0, 0, 0)], (1, 0, 0))
    assert not found
    assert q.shape[0] >= 2
    assert np.all(np.einsum("ij,ij->i", q, q) > 0) and finite(q)


@pytest.mark.parametrize("offset,hit", [(0.98, True), (1.0, False), (1.05, False)])
def test_strict_walk_duplicate_supports_stay_classified(oracle, offset, hit):  # gjk.rs:2023-2093
    b = oracle.box(0.5)
    n, simp = oracle.gjk_intersect(b, (0, 0, 0), I, b, (offset, 0, 0), I)
    assert (n > 0) == hit
    if n:
        assert finite(simp)
    hull = oracle.hull(CUBE8 * np.float32(0.5))
    n, _ = oracle.gjk_intersect(hull, (0, 0, 0), I, hull, (offset, 0, 0), I)
    assert (n > 0) == hit


# ------------------------------------------------------------------ epa.rs --
def check_finite(normal, depth, scale):  # epa.rs:505-512
    assert finite(normal) and math.isfinite(depth)
    assert -1e-3 * scale <= depth <= 1e3 * scale


def test_epa_collinear_simplex_fails_clean(oracle):  # epa.rs:535-550
    simplex = np.array([[1, 1, 1], [0.5, 0.5, 0.5], [-1, -1, -1], [-2, -2, -2]], np.float32)
    st, n, d = oracle.epa_points(simplex, simplex)
    assert st == 1 and np.all(n == 0) and d == 0.0


def test_epa_near_degenerate_seed_stays_clean(oracle):  # epa.rs:552-566
    a, c, d = np.array([1, 0, 0.2]), np.array([-0.5, 1, -0.3]), np.array([-0.4, -0.9, 0.4])
    st, n, dep = oracle.epa_points([a, a, c, d], [a, c, d, a * 1.5, c * -1.2])
    check_finite(n, dep, 2.0)


def test_epa_coplanar_simplex_stays_clean(oracle):  # epa.rs:568-582
    simplex = np.array([[1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]], np.float32)
    st, n, d = oracle.epa_points(simplex, simplex)
    check_finite(n, d, 2.0)


def test_epa_support_never_offers_new_point_converges(oracle):  # epa.rs:584-601
    simplex = np.array([[0.6, 0.6, 0.6], [-0.6, -0.6, 0.6], [-0.6, 0.6, -0.6], [0.6, -0.6, -0.6]], np.float32)
    st, n, d = oracle.epa_points(simplex, simplex)
    assert st == 0
    check_finite(n, d, 2.0)
    assert d >= -1e-3
    # closed form: the regular tetrahedron's faces all lie at 0.6/sqrt(3) from the origin
    assert abs(d - 0.6 / math.sqrt(3)) < 1e-6


def _full_epa(oracle, a, pa, b, pb):
    n, simp = oracle.gjk_intersect(a, pa, I, b, pb, I)
    if not n:
        return None
    return oracle.epa_pair(simp, a, pa, I, b, pb, I, core=False)


def test_epa_thin_plate_horizon_stays_clean(oracle):  # epa.rs:603-628
    a = oracle.box(0.5, 0.0005, 0.5)
    r = _full_epa(oracle, a, (0, 0, 0), a, (0.05, 0.0004, 0.03))
    if r is not None:
        st, n, d = r
        check_finite(n, d, 1.0)
        # exact: the plates overlap by 2 * 0.0005 - 0.0004 along y
        assert abs(d - 0.0006) < 1e-6 and abs(abs(n[1]) - 1.0) < 1e-6


def test_epa_tiny_penetration_is_finite_and_tiny(oracle):  # epa.rs:630-660
    s = oracle.sphere(1.0)
    r = _full_epa(oracle, s, (0, 0, 0), s, (2.0 - 1e-4, 0, 0))
    if r is not None:
        st, n, d = r
        check_finite(n, d, 1.0)
        assert -1e-3 <= d < 0.1
        assert abs(d - (2.0 - np.float32(2.0 - 1e-4))) < 1e-6


def test_epa_elongated_convex_stays_clean(oracle):  # epa.rs:662-692
    a, b = oracle.box(1.0, 0.005, 0.005), oracle.sphere(0.02)
    r = _full_epa(oracle, a, (0, 0, 0), b, (0.3, 0.01, 0.0))
    assert r is not None
    st, n, d = r
    check_finite(n, d, 1.0)
    n2 = float(n.astype(np.float64) @ n.astype(np.float64))
    assert abs(n2 - 1.0) < 1e-2 or n2 < 1e-8
    # exact: the rounded sliver (box + sphere) holds the sphere centre 0.005 +
    # 0.02 - 0.01 = 0.015 inside its +y face; EPA stops within its 1e-4 exit
    assert st == 0 and 0.015 - 1e-4 <= d <= 0.015 + 1e-6 and n[1] > 0.999


@pytest.mark.parametrize("scale", [1e-3, 1e-2, 1.0, 1e2, 1e3])
def test_epa_scale_sweep_spheres(oracle, scale):  # epa.rs:694-737
    r = np.float32(scale)
    s = oracle.sphere(r)
    pb = np.array([r * np.float32(1.6), r * np.float32(0.3), r * np.float32(0.2)], np.float32)
    expected = 2.0 * float(r) - float(np.linalg.norm(pb.astype(np.float64)))
    res = _full_epa(oracle, s, (0, 0, 0), s, pb)
    if res is None:
        return
    st, n, d = res
    check_finite(n, d, max(float(r), 1.0))
    assert -abs(expected) * 1e-2 < d < expected * 4.0 + float(r) * 1e-2
    if expected > 1e-3 * float(r):
        assert float(n @ (pb / np.linalg.norm(pb))) > 0.0
    # the polytope lies inside the Minkowski sphere: depth never exceeds it
    assert d <= expected + 1e-6 * max(1.0, float(r))


# ---------------------------------------------------- EPA tie contracts --
def test_epa_seeds_query_the_six_axes_in_order(oracle):  # epa.rs:150-172
    simplex = np.array([[0.6, 0.6, 0.6], [-0.6, -0.6, 0.6], [-0.6, 0.6, -0.6], [0.6, -0.6, -0.6]], np.float32)
    pts = np.vstack([simplex, CUBE8 * np.float32(0.9)])
    st, n, d, q, nv, it = oracle.epa_points_trace(simplex, pts)
    assert q.shape[0] >= 7
    assert np.array_equal(q[:6], AXES)
    # every seed point is pushed, accepted or not: 4 + 6 vertices before the loop
    # (epa.rs:162-163); each later expansion adds one (epa.rs:192-193)
    assert st == 0 and nv == 10 + (it - 1)
    # the seventh query is the best face's normal after seeding
    assert abs(np.linalg.norm(q[6].astype(np.float64)) - 1.0) < 1e-6


def test_epa_find_best_keeps_the_first_minimum(oracle):  # epa.rs:212-224
    assert oracle.epa_find_best([0.5, 0.3, -0.3, 0.3, 0.7]) == 1
    assert oracle.epa_find_best([0.2, 0.2]) == 0
    assert oracle.epa_find_best([0.9, -0.2, 0.2]) == 1
    # NaN compares Equal under partial_cmp: min_by keeps what it holds
    nan = float("nan")
    assert oracle.epa_find_best([nan, 1.0, 0.5]) == 0
    assert oracle.epa_find_best([1.0, nan, 0.5]) == 2


OCTA = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
# outward-wound octahedron faces, the four x > 0 ones interleaved with the rest
OCTA_FACES = [[0, 2, 4], [1, 4, 2], [0, 5, 2], [1, 2, 5], [0, 4, 3], [1, 3, 4], [0, 3, 5], [1, 5, 3]]


def test_epa_expand_horizon_order_follows_swap_remove(oracle):  # epa.rs:309-388
    """Hand trace of expand() toward (2, 0.1, 0.1), which sees the four x > 0
    faces.  Their edges, in face order, build the horizon list:
      (0,2,4): push 0-2, 2-4, 4-0            -> [0-2, 2-4, 4-0]
      (0,5,2): push 0-5, 5-2; 2-0 cancels 0-2 at index 0, swap_remove moves
               the last entry there          -> [5-2, 2-4, 4-0, 0-5]
      (0,4,3): 0-4 cancels 4-0 at index 2    -> [5-2, 2-4, 0-5]; push 4-3, 3-0
      (0,3,5): 0-3 cancels 3-0 (last); push 3-5; 5-0 cancels 0-5 at index 2
                                             -> [5-2, 2-4, 3-5, 4-3]
    retain keeps the hidden faces in order, then each horizon edge (u, v)
    stitches the face (u, v, new) in horizon order.  An order-keeping erase
    would give [2-4, 5-2, 4-3, 3-5] instead."""
    verts = np.vstack([OCTA, [[2.0, 0.1, 0.1]]]).astype(np.float32)
    ok, faces = oracle.epa_expand(OCTA_FACES, verts)
    assert ok
    expect = [[1, 4, 2], [1, 2, 5], [1, 3, 4], [1, 5, 3], [5, 2, 6], [2, 4, 6], [3, 5, 6], [4, 3, 6]]
    assert faces.tolist() == expect


def test_epa_expand_inside_point_is_a_no_op(oracle):  # epa.rs:347-350
    verts = np.vstack([OCTA, [[0.1, 0.1, 0.1]]]).astype(np.float32)
    ok, faces = oracle.epa_expand(OCTA_FACES, verts)
    assert ok and faces.tolist() == OCTA_FACES


def test_epa_expand_coplanar_point_is_not_visible(oracle):  # epa.rs:395-398 (PLANE_EPS = 1e-5)
    # over the centre of face (0,2,4), 5e-6 outside it: within the tolerance
    c = np.array([1, 1, 1], np.float64) / 3.0
    n = np.array([1, 1, 1], np.float64) / math.sqrt(3)
    verts = np.vstack([OCTA, [c + 5e-6 * n]]).astype(np.float32)
    ok, faces = oracle.epa_expand(OCTA_FACES, verts)
    assert ok and faces.tolist() == OCTA_FACES
    # 5e-5 outside: beyond the tolerance, so visible -- one face replaced by three
    verts = np.vstack([OCTA, [c + 5e-5 * n]]).astype(np.float32)
    ok, faces = oracle.epa_expand(OCTA_FACES, verts)
    assert ok and len(faces) == 10 and [0, 2, 4] not in faces.tolist()


def test_epa_expand_rejects_a_degenerate_horizon_atomically(oracle):  # epa.rs:352-360
    """Face (0,1,2) has a 1e-13 m edge 0-1 and is seen from P; its neighbour
    across that edge is not, so 0-1 is a horizon edge, and |(v1 - v0) x
    (P - v0)| ~ 1e-13 < 1e-12 rejects the expansion with the faces intact.
    With a 1e-3 m edge the same expansion is accepted: the hidden face stays
    and the three horizon edges stitch in face-edge order."""
    V = np.array([[0, 0, 0], [1e-13, 0, 0], [0, 1, 0], [0, -1, -5], [0.3, 0.3, 1]], np.float32)
    ok, faces = oracle.epa_expand([[0, 1, 2], [1, 0, 3]], V)
    assert not ok and faces.tolist() == [[0, 1, 2], [1, 0, 3]]
    V[1] = (1e-3, 0, 0)
    ok, faces = oracle.epa_expand([[0, 1, 2], [1, 0, 3]], V)
    assert ok and faces.tolist() == [[1, 0, 3], [0, 1, 4], [1, 2, 4], [2, 0, 4]]


def test_epa_face_distance_closed_forms(oracle):  # epa.rs:262-303
    # origin projects inside the triangle: the plane distance
    assert oracle.epa_face_distance((1, -1, 2), (1, 1, 2), (-1, 0, 2)) == pytest.approx(2.0, abs=1e-7)
    # past edge a-b, beyond vertex a (a.ba > 0): |a|
    assert oracle.epa_face_distance((1, 1, 1), (2, 1, 1), (1, 2, 1)) == pytest.approx(math.sqrt(3), abs=1e-6)
    # past edge a-b, beyond vertex b (b.ba < 0): |b|
    assert oracle.epa_face_distance((-2, 1, 1), (-1, 1, 1), (-1, 2, 1)) == pytest.approx(math.sqrt(3), abs=1e-6)
    # past edge a-b, foot of the origin on it: distance to the line a-b
    a, b, c = np.array([-1.0, 1, 1]), np.array([1.0, 1, 1]), np.array([0.0, 2, 1])
    foot = a + (-(a @ (b - a)) / ((b - a) @ (b - a))) * (b - a)
    assert oracle.epa_face_distance(a, b, c) == pytest.approx(float(np.linalg.norm(foot)), abs=1e-6)


def test_epa_face_distance_is_the_true_triangle_distance(oracle):  # epa.rs:5-14, 259-261
    """Where the origin's projection leaves the triangle through exactly one
    edge region, the face distance is the true distance from the origin to
    the triangle (the closest point then lies on that edge)."""
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(400):
        t = rng.normal(size=(3, 3)).astype(np.float32)
        a, b, c = t.astype(np.float64)
        n = np.cross(b - a, c - a)
        n /= np.linalg.norm(n)
        o = -(a @ n) * n  # projection of the origin onto the plane
        outside = [np.cross(q - p, n) @ (o - p) > 1e-6 for p, q in ((a, b), (b, c), (c, a))]
        if sum(outside) != 1:
            continue
        # dense closed form: closest point of each edge segment
        best = np.inf
        for p, q in ((a, b), (b, c), (c, a)):
            s = np.clip(-(p @ (q - p)) / ((q - p) @ (q - p)), 0.0, 1.0)
            best = min(best, float(np.linalg.norm(p + s * (q - p))))
        # the reference's edge formula sqrt((|a|^2 |b|^2 - (a.b)^2) / |b - a|^2)
        # cancels in f32: allow 8 ulps of the products, carried through the root
        p, q = [(a, b), (b, c), (c, a)][outside.index(True)]
        tol = 8 * 6e-8 * (p @ p) * (q @ q) / (2 * ((q - p) @ (q - p)) * max(best, 1e-3)) + 1e-6
        assert abs(oracle.epa_face_distance(*t) - best) <= tol
        checked += 1
    assert checked > 50


# ------------------------------------------------------- box_box tie rules --
def test_cull_points_keeps_the_last_deepest_and_sorts_stably(oracle):  # box_box.rs:126-174
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [2, 2, 0]], np.float32)
    deps = np.array([0.2, 0.5, 0.1, 0.5, 0.3], np.float32)
    # deepest: the last of the two 0.5 depths (index 3).  Squared distances
    # from it: p0 2, p1 1, p2 1, p4 2; stable descending: p0, p4, p1, p2.
    op, od = oracle.cull_points(pts, deps, 4)
    assert op.tolist() == pts[[3, 0, 4, 1]].tolist()
    assert od.tolist() == deps[[3, 0, 4, 1]].tolist()


def test_cull_points_ties_keep_construction_order(oracle):  # box_box.rs:159-168
    # all five points equidistant from the deepest one: the sort keeps index order
    ang = np.arange(4) * (np.pi / 2)
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(4)], 1).astype(np.float32) * 2
    ring = np.round(ring)  # exact (2,0,0), (0,2,0), (-2,0,0), (0,-2,0)
    pts = np.vstack([ring[:2], [[0, 0, 0]], ring[2:]]).astype(np.float32)
    deps = np.array([0.1, 0.1, 0.4, 0.1, 0.1], np.float32)
    op, od = oracle.cull_points(pts, deps, 4)
    assert op.tolist() == pts[[2, 0, 1, 3]].tolist()


def test_cull_points_passes_short_lists_through(oracle):  # box_box.rs:131-135
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    op, od = oracle.cull_points(pts, [0.1, 0.3, 0.2], 4)
    assert op.tolist() == pts.tolist() and od.tolist() == pytest.approx([0.1, 0.3, 0.2])


def test_box_box_exact_touch_keeps_zero_depth_points(oracle):  # box_box.rs:497-505
    """Face-to-face at zero depth: B's bottom face lies exactly on A's top
    face, so every clipped point has dep == 0 and the reference's dep >= 0
    filter keeps all four corners of B's face."""
    m, _ = oracle.collide(oracle.box(1.0), (0, 0, 0), I, oracle.box(0.5, 1.0, 0.5), (0, 2, 0), I)
    assert m.shape[0] == 4
    assert np.all(m[:, 6] == 0.0)
    assert np.all(m[:, 3:6] == np.array([0, 1, 0], np.float32))
    got = sorted(map(tuple, m[:, :3].tolist()))
    assert got == sorted((x, 1.0, z) for x in (-0.5, 0.5) for z in (-0.5, 0.5))


def test_box_box_face_contact_depths_closed_form(oracle):  # box_box.rs:178-528
    # B rests 0.1 into A's top face, offset so that two of its corners clip
    m, _ = oracle.collide(oracle.box(1.0), (0, 0, 0), I, oracle.box(0.5), (0.75, 1.4, -0.25), I)
    assert m.shape[0] == 4
    assert np.allclose(m[:, 6], 0.1, atol=1e-6)
    got = sorted(map(tuple, np.round(m[:, [0, 2]].astype(np.float64), 6).tolist()))
    assert got == sorted((x, z) for x in (0.25, 1.0) for z in (-0.75, 0.25))


# ------------------------------------------- exact Minkowski geometry --
def _qrand(rng):
    u1, u2, u3 = rng.random(3)
    return np.array([math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                     math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3)],
                    np.float32)


def _qrot(q, v):
    x, y, z, w = (float(t) for t in q)
    b = np.array([x, y, z])
    v = np.asarray(v, np.float64)
    return v * (w * w - b @ b) + np.outer(v @ b, b) * 2 + np.cross(b, v) * (2 * w)


def _shape(oracle, rng, kind):
    """(oracle shape, local core points, margin).  Cores: a hull is its own
    core, a box shrinks by its 0.04 margin, a capsule is its segment and a
    sphere its centre (shapes/*: core_support_local)."""
    if kind == "hull":
        ax = rng.uniform(0.3, 1.2, 3)
        ax[rng.integers(3)] /= rng.choice([1, 3, 10, 30, 100])
        g = rng.normal(size=(int(rng.integers(8, 64)), 3))
        p = (g / np.linalg.norm(g, axis=1, keepdims=True) * ax).astype(np.float32)
        return oracle.hull(p), p.astype(np.float64), 0.0
    if kind == "box":
        h = rng.uniform(0.05, 1.0, 3).astype(np.float32)
        if rng.random() < 0.5:
            h[rng.integers(3)] = np.float32(rng.uniform(0.01, 0.05))
        c = np.maximum(h.astype(np.float64) - np.float64(np.float32(0.04)), 0.0)
        return oracle.box(*h), np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * c, \
            float(np.float32(0.04))
    if kind == "capsule":
        r, hh = np.float32(rng.uniform(0.1, 0.5)), np.float32(rng.uniform(0.2, 1.0))
        return oracle.capsule(r, hh), np.array([[0, hh, 0], [0, -hh, 0]], np.float64), float(r)
    r = np.float32(rng.uniform(0.1, 0.6))
    return oracle.sphere(r), np.zeros((1, 3)), float(r)


def _pt_tri_dist(a, b, c):
    """Distance from the origin to each triangle (Ericson 5.1.5), float64."""
    ab, ac = b - a, c - a
    d1, d2 = np.einsum("ij,ij->i", ab, -a), np.einsum("ij,ij->i", ac, -a)
    d3, d4 = np.einsum("ij,ij->i", ab, -b), np.einsum("ij,ij->i", ac, -b)
    d5, d6 = np.einsum("ij,ij->i", ab, -c), np.einsum("ij,ij->i", ac, -c)
    va, vb, vc = d3 * d6 - d5 * d4, d5 * d2 - d1 * d6, d1 * d4 - d3 * d2
    out = np.full(a.shape, np.nan)
    todo = np.ones(len(a), bool)

    def put(mask, val):
        m = mask & todo
        out[m] = val[m]
        todo[m] = False

    with np.errstate(all="ignore"):
        put((d1 <= 0) & (d2 <= 0), a)
        put((d3 >= 0) & (d4 <= d3), b)
        put((d6 >= 0) & (d5 <= d6), c)
        put((vc <= 0) & (d1 >= 0) & (d3 <= 0), a + ab * (d1 / (d1 - d3))[:, None])
        put((vb <= 0) & (d2 >= 0) & (d6 <= 0), a + ac * (d2 / (d2 - d6))[:, None])
        put((va <= 0) & (d4 - d3 >= 0) & (d5 - d6 >= 0), b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))[:, None])
        den = 1.0 / (va + vb + vc)
        put(np.ones(len(a), bool), a + ab * (vb * den)[:, None] + ac * (vc * den)[:, None])
    return np.linalg.norm(out, axis=1)


KINDS = [("hull", "hull"), ("box", "hull"), ("capsule", "hull"), ("hull", "box"), ("sphere", "hull"),
         ("hull", "capsule")]


def test_epa_depth_and_normal_match_the_minkowski_hull(oracle):
    """EPA (through convex_convex) against the exact penetration depth of the
    core Minkowski difference, margins added back: min over Qhull facets of
    the facet's distance from the origin.  The polytope EPA grows lies inside
    the difference, so its depth is at most the exact one; its exit
    (epa.rs:179-183) puts the support along the reported normal within 1e-4
    of the face, so the depth is at least exact - 1e-4, and the normal is one
    along which the difference reaches no further than depth + 1e-4.
    Hulls are elongated up to 100:1 and boxes include 0.01-0.05 m slabs."""
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(11)
    n_epa = n_skipped = 0
    for k in range(1800):
        ka, kb = KINDS[k % len(KINDS)]
        A, ca, ma = _shape(oracle, rng, ka)
        B, cb, mb = _shape(oracle, rng, kb)
        qa, qb = _qrand(rng), _qrand(rng)
        d = rng.normal(size=3)
        pb = (d / np.linalg.norm(d) * rng.uniform(0.0, 1.0)).astype(np.float32)
        m, path = oracle.collide(A, (0, 0, 0), qa, B, pb, qb)
        if path != 2 or m.shape[0] == 0:
            continue
        M = (_qrot(qa, ca)[:, None, :] - (_qrot(qb, cb) + pb.astype(np.float64))[None]).reshape(-1, 3)
        eq = ConvexHull(M).equations
        if not np.all(eq[:, 3] < 0):
            # GJK reported overlapping cores on a nearly flat tetrahedron (the
            # reference's f32 region tests, gjk.rs:546-558) while the exact
            # difference misses the origin: classified, not compared
            n_skipped += 1
            continue
        exact = float(np.min(-eq[:, 3])) + ma + mb
        depth = float(m[0, 6])
        nrm = m[0, 3:6].astype(np.float64)
        reach = float(np.max(M @ nrm)) + ma + mb
        assert exact - 1e-4 - 2e-6 <= depth <= exact + 2e-6, (k, ka, kb, depth, exact)
        assert reach - depth <= 1e-4 + 2e-6, (k, ka, kb, reach, depth)
        n_epa += 1
    assert n_epa > 900 and n_skipped <= 3


def test_gjk_band_distance_matches_the_minkowski_hull(oracle):
    """The margin-core walk's Band distance (gjk.rs:171-288) against the exact
    distance from the origin to the core Minkowski difference (min over its
    Qhull boundary triangles), margins subtracted.  The walk stops when
    |v|^2 - v.w <= 1e-6 |v|^2 (gjk.rs:292), so its distance is within
    1e-6 |v| of the exact one; f32 rounding adds ~1e-7."""
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(12)
    kinds = [("hull", "hull"), ("box", "hull"), ("capsule", "hull"), ("hull", "box"), ("sphere", "hull"),
             ("box", "capsule")]
    n_band = 0
    for k in range(2400):
        ka, kb = kinds[k % len(kinds)]
        A, ca, ma = _shape(oracle, rng, ka)
        B, cb, mb = _shape(oracle, rng, kb)
        qa, qb = _qrand(rng), _qrand(rng)
        wa, wb = _qrot(qa, ca), _qrot(qb, cb)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pb = (d * (np.max(wa @ d) - np.min(wb @ d) + rng.uniform(0.001, 0.3))).astype(np.float32)
        kind, dist, _, _, _ = oracle.gjk_contact(A, (0, 0, 0), qa, B, pb, qb, 1e30)
        if kind != 1:
            continue
        M = (wa[:, None, :] - (wb + pb.astype(np.float64))[None]).reshape(-1, 3)
        T = M[ConvexHull(M).simplices]
        exact = float(np.min(_pt_tri_dist(T[:, 0], T[:, 1], T[:, 2])))
        assert abs((float(dist) + ma + mb) - exact) <= 1e-6, (k, ka, kb, float(dist) + ma + mb, exact)
        n_band += 1
    assert n_band > 2300


def test_gjk_band_distance_with_near_duplicate_hull_vertices(oracle):
    """As above, for hulls whose every vertex has a twin 2e-5 - 8e-5 m away.
    The walk's duplicate gate (gjk.rs:253-256) only stops on supports within
    1e-6 of a simplex vertex; a twin is a genuine new support, and stopping
    on it would leave the distance up to its offset short."""
    from scipy.spatial import ConvexHull
    rng = np.random.default_rng(5)
    n_band = 0
    for k in range(600):
        _, ca, _ = _shape(oracle, rng, "hull")
        pts = ca.astype(np.float32)
        j = rng.normal(size=pts.shape)
        j /= np.linalg.norm(j, axis=1, keepdims=True)
        pts = np.vstack([pts, pts + (j * rng.uniform(2e-5, 8e-5, (len(pts), 1))).astype(np.float32)])
        A, ca = oracle.hull(pts.astype(np.float32)), pts.astype(np.float64)
        B, cb, mb = _shape(oracle, rng, ["sphere", "box", "hull"][k % 3])
        qa, qb = _qrand(rng), _qrand(rng)
        wa, wb = _qrot(qa, ca), _qrot(qb, cb)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pb = (d * (np.max(wa @ d) - np.min(wb @ d) + rng.uniform(0.001, 0.3))).astype(np.float32)
        kind, dist, _, _, _ = oracle.gjk_contact(A, (0, 0, 0), qa, B, pb, qb, 1e30)
        if kind != 1:
            continue
        M = (wa[:, None, :] - (wb + pb.astype(np.float64))[None]).reshape(-1, 3)
        T = M[ConvexHull(M).simplices]
        exact = float(np.min(_pt_tri_dist(T[:, 0], T[:, 1], T[:, 2])))
        assert abs((float(dist) + mb) - exact) <= 1e-6, (k, float(dist) + mb, exact)
        n_band += 1
    assert n_band > 580


# ------------------------------------------------ limits and tie rules --
@pytest.mark.parametrize("scale", [1.0, 1e2, 1e3])
def test_epa_exits_at_its_documented_limits(oracle, scale):  # epa.rs:28-46, :184-190
    """Every seed point is pushed (4 + 6), each expansion adds one vertex, and
    the run stops with VertexLimit once 128 vertices are held.  Two spheres
    of radius 100 or 1000 need more than that to meet the 1e-4 exit."""
    r = np.float32(scale)
    s = oracle.sphere(r)
    pb = np.array([r * np.float32(1.6), r * np.float32(0.3), r * np.float32(0.2)], np.float32)
    n, simp = oracle.gjk_intersect(s, (0, 0, 0), I, s, pb, I)
    assert n == 4
    st, nrm, d, nv, it = oracle.epa_pair_trace(simp, s, (0, 0, 0), I, s, pb, I)
    assert st in (0, 3)
    if st == 0:
        assert nv == 10 + (it - 1) and nv < 128
    else:
        assert nv == 128 and it == 128 - 10
    assert scale < 10 or st == 3


def test_gjk_touch_tolerance_scale(oracle):  # gjk.rs:194-198, :232-245
    """The walk treats |v|^2 < 1e-12 max(|w0|^2, 1) as touching and hands the
    pair to the strict walk, which cannot enclose the origin of two cubes
    face to face (gjk.rs:2032-2043): no contact.  Gaps of a few microns
    stay on the distance path and report their gap."""
    h = oracle.hull(CUBE8 * np.float32(0.5))
    for gap in (2e-6, 5e-6, 2e-5):
        k, dist, nrm, _, _ = oracle.gjk_contact(h, (0, 0, 0), I, h, (1.0 + gap, 0, 0), I, 0.0173)
        exact = float(np.float32(1.0 + gap)) - 1.0
        assert k == 1 and abs(dist - exact) < 1e-7 and abs(nrm[0] - 1.0) < 1e-6
    k, _, _, _, _ = oracle.gjk_contact(h, (0, 0, 0), I, h, (1.0 + 5e-7, 0, 0), I, 0.0173)
    assert k == 0


def test_hull_support_keeps_the_first_maximum(oracle):  # convex_hull.rs:55-66
    h = oracle.hull(CUBE8 * np.float32(0.5))
    # four corners tie along +x; CUBE8 lists (+0.5, -0.5, -0.5) first among them
    assert oracle.support_world(h, (0, 0, 0), I, (1, 0, 0)).tolist() == [0.5, -0.5, -0.5]
    assert oracle.support_world(h, (0, 0, 0), I, (0, 0, -1)).tolist() == [-0.5, -0.5, -0.5]


def test_capsule_core_tip_takes_plus_y_on_zero(oracle):  # capsule.rs:59-69
    c = oracle.capsule(0.3, 0.7)
    assert oracle.support_world(c, (0, 0, 0), I, (1, 0, 0), core=True).tolist() == [0.0, np.float32(0.7), 0.0]
    assert oracle.support_world(c, (0, 0, 0), I, (0, -1, 0), core=True).tolist() == [0.0, -np.float32(0.7), 0.0]


def test_sphere_inside_box_face_ties_go_x_first(oracle):  # sphere.rs:57-75
    # centre inside: 0.5 to the x and y faces, 2 to z; the x face wins the tie
    m, _ = oracle.collide(oracle.box(1, 1, 2), (0, 0, 0), I, oracle.sphere(0.3), (0.5, 0.5, 0), I)
    assert m.shape[0] == 1
    assert m[0, 3:6].tolist() == [1.0, 0.0, 0.0]
    assert m[0, 6] == pytest.approx(0.8, abs=1e-7)
    m, _ = oracle.collide(oracle.box(1, 2, 1), (0, 0, 0), I, oracle.sphere(0.3), (0, 1.5, -0.5), I)
    assert m[0, 3:6].tolist() == [0.0, 1.0, 0.0]


def test_reduce_thin_enclosing_tetrahedron_is_inside(oracle):  # gjk.rs:495-558
    """A tetrahedron 2e-10 thick around the origin: its triple product 8e-10
    passes the 1e-12 coplanar gate, and the region tests find the origin
    inside."""
    h = np.float32(1e-10)
    inside, _, _, _ = oracle.reduce_support([(0, 0, h), (0, 1, -h), (-1, -1, -h), (1, -1, -h)], np.zeros((4, 3)))
    assert inside
    # one more decade down it is coplanar: the closest face, not Inside
    h = np.float32(1e-13)
    inside, v, _, n = oracle.reduce_support([(0, 0, h), (0, 1, -h), (-1, -1, -h), (1, -1, -h)], np.zeros((4, 3)))
    assert not inside and n <= 3 and float(np.abs(v).max()) <= 2e-13


def test_epa_edge_distance_overflow_follows_f32_max(oracle):  # epa.rs:294-298
    """|a|^2 |b|^2 and (a.b)^2 both overflow to +inf for 2e10 m coordinates;
    inf - inf is NaN, and Rust's f32::max(0.0) returns the non-NaN operand,
    so the edge distance is sqrt(0) = 0."""
    a, b, c = (-1e10, 2e10, 1.0), (1e10, 2e10, 1.0), (0.0, 3e10, 1.0)
    assert oracle.epa_face_distance(a, b, c) == 0.0
