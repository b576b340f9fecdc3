"""The oracle against the reference's own known-answer tests (CPU only).

Each test restates one reference test: same shapes, poses and expected
values, same tolerances (file:line cited).  These pin the restated
closed forms, the GJK walk's band/separation routing and EPA depths.
"""
import math

import numpy as np
import pytest

I = (0.0, 0.0, 0.0, 1.0)


def rot(axis, angle):
    """glam Quat::from_rotation_{x,y,z}: (sin, cos) of angle/2 in f32."""
    h = np.float32(angle) * np.float32(0.5)
    s, c = float(np.sin(h)), float(np.cos(h))
    q = [0.0, 0.0, 0.0, c]
    q["xyz".index(axis)] = s
    return tuple(q)


OCTA = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
T_SPHERE_HALF = 0.017320508  # sphere.rs:89, plane.rs:112: 0.5 * sqrt(3) * 0.02


# ----------------------------------------------------------- sphere.rs --
def test_spheres_overlap(oracle):  # sphere.rs:92-100
    m, _ = oracle.collide(oracle.sphere(1.0), (0, 0, 0), I, oracle.sphere(1.0), (1.5, 0, 0), I)
    assert m.shape[0] == 1
    assert m[0, 6] == pytest.approx(0.5, abs=1e-6)
    assert m[0, 3] == pytest.approx(1.0, abs=1e-6)


def test_spheres_separated(oracle):  # sphere.rs:102-107
    m, _ = oracle.collide(oracle.sphere(1.0), (0, 0, 0), I, oracle.sphere(1.0), (3, 0, 0), I)
    assert m.shape[0] == 0


def test_sphere_box_contact(oracle):  # sphere.rs:109-117 (box is A, normal box -> sphere)
    m, _ = oracle.collide(oracle.box(1.0), (0, 0, 0), I, oracle.sphere(0.5), (1.4, 0, 0), I)
    assert m[0, 6] == pytest.approx(0.1, abs=1e-6)
    assert m[0, 3] == pytest.approx(1.0, abs=1e-6)
    # sphere as A: the normal is flipped (narrowphase/mod.rs:258-264)
    m, _ = oracle.collide(oracle.sphere(0.5), (1.4, 0, 0), I, oracle.box(1.0), (0, 0, 0), I)
    assert m[0, 3] == pytest.approx(-1.0, abs=1e-6)


def test_sphere_box_band_contact(oracle):  # sphere.rs:121-141
    assert oracle.threshold(oracle.sphere(0.5)) == pytest.approx(T_SPHERE_HALF, rel=1e-6)
    m, _ = oracle.collide(oracle.box(1.0), (0, 0, 0), I, oracle.sphere(0.5), (1.51, 0, 0), I)
    assert m[0, 6] == pytest.approx(-0.01, abs=1e-6)
    assert m[0, 3] == pytest.approx(1.0, abs=1e-6)
    m, _ = oracle.collide(oracle.box(1.0), (0, 0, 0), I, oracle.sphere(0.5), (1.6, 0, 0), I)
    assert m.shape[0] == 0


# ------------------------------------------------------------ plane.rs --
Y_PLANE = ((0, 1, 0), 0.0)


def test_sphere_resting_on_plane(oracle):  # plane.rs:122-142
    m, _ = oracle.collide(oracle.plane(*Y_PLANE), (0, 0, 0), I, oracle.sphere(0.5), (0, 0.4, 0), I)
    assert m[0, 4] == pytest.approx(1.0, abs=1e-6)
    assert m[0, 6] == pytest.approx(0.1, abs=1e-6)
    assert m[0, 1] == pytest.approx(-0.1, abs=1e-6)


def test_plane_separated_and_band(oracle):  # plane.rs:144-190
    P = oracle.plane(*Y_PLANE)
    assert oracle.collide(P, (0, 0, 0), I, oracle.sphere(0.5), (0, 2, 0), I)[0].shape[0] == 0
    m, _ = oracle.collide(P, (0, 0, 0), I, oracle.sphere(0.5), (0, 0.51, 0), I)
    assert m[0, 4] == pytest.approx(1.0) and m[0, 6] == pytest.approx(-0.01, abs=1e-6)
    assert oracle.collide(P, (0, 0, 0), I, oracle.sphere(0.5), (0, 0.6, 0), I)[0].shape[0] == 0


def test_sphere_below_plane_flips_the_normal(oracle):  # plane.rs:192-208
    m, _ = oracle.collide(oracle.plane(*Y_PLANE), (0, 0, 0), I, oracle.sphere(0.5), (0, -0.3, 0), I)
    assert m[0, 4] == pytest.approx(-1.0, abs=1e-6)
    assert m[0, 6] == pytest.approx(0.2, abs=1e-6)


def test_rotated_plane_uses_the_world_normal(oracle):  # plane.rs:210-227
    q = rot("z", -math.pi / 2)
    m, _ = oracle.collide(oracle.plane(*Y_PLANE), (0, 0, 0), q, oracle.sphere(0.5), (0.4, 0, 0), I)
    assert m[0, 3] == pytest.approx(1.0, abs=1e-6)
    assert m[0, 6] == pytest.approx(0.1, abs=1e-6)


def test_convex_first_plane_second_flips(oracle):  # narrowphase/mod.rs:245-253
    m, _ = oracle.collide(oracle.sphere(0.5), (0, 0.4, 0), I, oracle.plane(*Y_PLANE), (0, 0, 0), I)
    assert m[0, 4] == pytest.approx(-1.0) and m[0, 6] == pytest.approx(0.1, abs=1e-6)


# -------------------------------------------------------- narrowphase/mod.rs --
def test_plane_vs_box_through_collide(oracle):  # mod.rs:370-380
    m, _ = oracle.collide(oracle.plane(*Y_PLANE), (0, 0, 0), I, oracle.box(1.0), (0, 0.5, 0), I)
    assert m[0, 4] == pytest.approx(1.0) and m[0, 6] == pytest.approx(0.5, abs=1e-6)


def test_rotated_capsule_on_plane(oracle):  # mod.rs:382-400
    q = rot("z", math.pi / 2)
    m, _ = oracle.collide(oracle.plane(*Y_PLANE), (0, 0, 0), I, oracle.capsule(0.5, 1.0), (0, 0.4, 0), q)
    assert m[0, 4] == pytest.approx(1.0) and m[0, 6] == pytest.approx(0.1, abs=1e-6)


def test_plane_plane_is_none(oracle):  # mod.rs:104
    P = oracle.plane(*Y_PLANE)
    assert oracle.collide(P, (0, 0, 0), I, P, (0, 1, 0), I)[0].shape[0] == 0


def test_seeded_collide_matches_unseeded(oracle):  # mod.rs:515-564 (capsule pair; the cylinder/cone rows are out of scope)
    a = b = oracle.capsule(0.5, 1.0)
    q = rot("y", 0.6)
    off = (0.7, 0.2, 0.1)
    un, _ = oracle.collide(a, (0, 0, 0), q, b, off, I)
    n = un[0, 3:6]
    s3 = np.array([0.3, -0.5, 0.8], np.float32)
    for seed in (n, -n, s3 / np.linalg.norm(s3)):
        sd, _ = oracle.collide(a, (0, 0, 0), q, b, off, I, seed=seed)
        assert sd.shape[0] == un.shape[0]
        assert sd[0, 6] == pytest.approx(un[0, 6], abs=2e-4)
        assert np.linalg.norm(sd[0, 3:6] - n) < 1e-4 or np.dot(sd[0, 3:6], n) > 0.9999


# ------------------------------------------------------------ box_box.rs --
def test_box_box_face_face_contact_along_x(oracle):  # box_box.rs:536-546
    m, _ = oracle.collide(oracle.box(1.0), (0, 0, 0), I, oracle.box(1.0), (1.9, 0, 0), I)
    assert 1 <= m.shape[0] <= 4
    for p in m:
        assert p[3] == pytest.approx(1.0, abs=1e-5)
        assert p[6] == pytest.approx(0.1, abs=1e-4)


def test_box_box_separated(oracle):  # box_box.rs:548-554
    assert oracle.collide(oracle.box(1.0), (0, 0, 0), I, oracle.box(1.0), (3, 0, 0), I)[0].shape[0] == 0


# ------------------------------------------------------- convex_convex.rs --
def cc(oracle, a, pa, b, pb, qa=I, qb=I):
    m, path = oracle.convex_convex(a, pa, qa, b, pb, qb)
    return m, path


def test_capsules_along_x_overlap(oracle):  # convex_convex.rs:215-222
    m, _ = cc(oracle, oracle.capsule(0.5, 1.0), (0, 0, 0), oracle.capsule(0.5, 1.0), (0.7, 0, 0))
    assert m[0, 3] == pytest.approx(1.0, abs=1e-4)
    assert m[0, 6] == pytest.approx(0.3, abs=1e-3)


def test_separated_capsules_do_not_touch(oracle):  # convex_convex.rs:224-229
    m, _ = cc(oracle, oracle.capsule(0.5, 1.0), (0, 0, 0), oracle.capsule(0.5, 1.0), (2.0, 0, 0))
    assert m.shape[0] == 0


def test_hull_sphere_overlap(oracle):  # convex_convex.rs:254-270
    m, _ = cc(oracle, oracle.hull(OCTA), (0, 0, 0), oracle.sphere(0.5), (1.25, 0, 0))
    assert m[0, 3] == pytest.approx(1.0, abs=1e-3)
    assert m[0, 6] == pytest.approx(0.25, abs=1e-3)


def test_margin_shell_overlap_resolved_by_walk(oracle):  # convex_convex.rs:276-300
    m, path = cc(oracle, oracle.box(1.0), (0, 0, 0), oracle.sphere(0.5), (0, 1.3, 0))
    assert path == 1  # band walk, no EPA
    assert m[0, 4] == pytest.approx(1.0, abs=1e-3)
    assert m[0, 6] == pytest.approx(0.2, abs=1e-3)
    assert m[0, 1] == pytest.approx(0.8, abs=1e-3)


def test_deep_overlap_runs_epa_with_margins(oracle):  # convex_convex.rs:304-329
    m, path = cc(oracle, oracle.box(1.0), (0, 0, 0), oracle.sphere(0.5), (0, 0.6, 0))
    assert path == 2
    assert m[0, 4] == pytest.approx(1.0, abs=1e-3)
    assert m[0, 6] == pytest.approx(0.9, abs=1e-3)
    assert m[0, 1] == pytest.approx(0.1, abs=1e-3)


def test_margin_shell_separation_reports_band(oracle):  # convex_convex.rs:333-355
    m, _ = cc(oracle, oracle.capsule(0.5, 1.0), (0, 0, 0), oracle.box(1.0), (1.51, 0, 0))
    assert m[0, 3] == pytest.approx(1.0, abs=1e-3)
    assert m[0, 6] == pytest.approx(-0.01, abs=1e-3)


def _support_plane_invariant(oracle, m, b, pb, qb):
    p = m[0]
    n = p[3:6]
    sb = oracle.support_world(b, pb, qb, -n)
    assert np.all(np.isfinite(p))
    assert float(np.dot(n, p[:3])) == pytest.approx(float(np.dot(n, sb)), abs=2e-2)


def test_witness_asymmetric_hull_vs_sphere(oracle):  # convex_convex.rs:454-481
    pb = (1.25, 0.3, 0.0)
    m, _ = cc(oracle, oracle.hull(OCTA), (0, 0, 0), oracle.sphere(0.5), pb)
    assert m[0, 6] > 0
    assert np.dot(m[0, 3:6], np.array([0.25, 0.3, 0]) / np.linalg.norm([0.25, 0.3, 0])) > 0
    _support_plane_invariant(oracle, m, oracle.sphere(0.5), pb, I)


def test_witness_rotated_box_vs_box(oracle):  # convex_convex.rs:485-498
    a = oracle.box(1.0, 0.4, 0.6)
    b = oracle.box(0.5)
    qa, qb, pb = rot("z", 0.6), rot("z", -0.35), (1.3, 0.25, 0.1)
    m, _ = cc(oracle, a, (0, 0, 0), b, pb, qa, qb)
    assert m[0, 6] > 0
    _support_plane_invariant(oracle, m, b, pb, qb)


def test_witness_elongated_hull_vs_hull(oracle):  # convex_convex.rs:503-516
    a = oracle.box(1.5, 0.05, 0.05)
    b = oracle.box(0.3)
    qa, pb = rot("z", 0.25), (1.7, 0.2, 0.0)
    m, _ = cc(oracle, a, (0, 0, 0), b, pb, qa, I)
    assert m[0, 6] > 0
    _support_plane_invariant(oracle, m, b, pb, I)


def test_witness_unequal_margins_deep_overlap(oracle):  # convex_convex.rs:520-543
    m, _ = cc(oracle, oracle.box(1.0), (0, 0, 0), oracle.sphere(0.25), (0, 0.9, 0))
    assert m[0, 6] == pytest.approx(0.35, abs=2e-3)
    assert m[0, 1] == pytest.approx(0.65, abs=2e-3)
    assert m[0, 4] == pytest.approx(1.0, abs=1e-3)


def test_band_to_epa_transition_stays_continuous(oracle):  # convex_convex.rs:549-596
    step = 0.01
    prev = None
    for i in range(78):
        py = np.float32(1.32) - np.float32(step) * np.float32(i)
        m, _ = cc(oracle, oracle.box(1.0), (0, 0, 0), oracle.sphere(0.5), (0, py, 0))
        if m.shape[0] == 0:
            prev = None
            continue
        p = m[0]
        assert abs(p[4] - 1.0) <= 5e-2
        if prev is not None:
            jump = 2.5 * step + 2e-3
            assert abs(p[6] - prev[1]) <= jump
            assert abs(p[1] - prev[2]) <= jump
            assert p[4] * prev[0] > 0.5
        prev = (p[4], p[6], p[1])


def test_exact_touching_and_tiny_gaps(oracle):  # convex_convex.rs:600-652
    s = oracle.sphere(0.5)
    m, _ = cc(oracle, s, (0, 0, 0), s, (1.0, 0, 0))
    if m.shape[0]:
        assert abs(m[0, 6]) <= 1e-3 and m[0, 3] == pytest.approx(1.0, abs=1e-2)
    m, _ = cc(oracle, s, (0, 0, 0), s, (1.0 - 1e-4, 0, 0))
    if m.shape[0]:
        assert abs(m[0, 6]) <= 0.1
    m, _ = cc(oracle, s, (0, 0, 0), s, (1.0 + 1e-4, 0, 0))
    if m.shape[0]:
        assert m[0, 6] <= 0.0


# ---------------------------------------------------------------- gjk.rs --
def test_gjk_intersect_overlap_and_separation(oracle):  # gjk.rs:1054-1084
    S1, B1 = oracle.sphere(1.0), oracle.box(1.0)
    assert oracle.gjk_intersect(S1, (0, 0, 0), I, S1, (1.5, 0, 0), I)[0] == 4
    assert oracle.gjk_intersect(S1, (0, 0, 0), I, S1, (3, 0, 0), I)[0] == 0
    assert oracle.gjk_intersect(B1, (0, 0, 0), I, B1, (1.5, 0, 0), I)[0] == 4
    assert oracle.gjk_intersect(B1, (0, 0, 0), I, B1, (3, 0, 0), I)[0] == 0
    S5 = oracle.sphere(0.5)
    assert oracle.gjk_intersect(S5, (1.4, 0, 0), I, B1, (0, 0, 0), I)[0] == 4
    assert oracle.gjk_intersect(S5, (2.0, 0, 0), I, B1, (0, 0, 0), I)[0] == 0


def test_contact_walk_reports_band_separation(oracle):  # gjk.rs:1106-1127
    s = oracle.sphere(0.5)
    k, d, n, pb, _ = oracle.gjk_contact(s, (0, 0, 0), I, s, (1.01, 0, 0), I, 0.0173)
    assert k == 1
    assert abs(d - 0.01) < 1e-4
    assert np.linalg.norm(n - [1, 0, 0]) < 1e-4
    assert np.linalg.norm(pb - [0.51, 0, 0]) < 1e-4
    assert oracle.gjk_contact(s, (0, 0, 0), I, s, (1.5, 0, 0), I, 0.0173)[0] == 0


def test_contact_walk_routes_overlap_and_rotated_band(oracle):  # gjk.rs:1134-1192
    s = oracle.sphere(0.5)
    k, d, *_ = oracle.gjk_contact(s, (0, 0, 0), I, s, (0.7, 0, 0), I, 0.0173)
    assert k == 1 and abs(d + 0.3) < 1e-4
    cap = oracle.capsule(0.5, 1.0)
    k, d, *_ = oracle.gjk_contact(cap, (0, 0, 0), I, cap, (1.01, 0, 0), I, 0.5)
    assert k == 1 and abs(d - 0.01) < 1e-4
    q = rot("z", 0.5)
    pb = (1.6, 0.2, 0.1)
    k, ref, *_ = oracle.gjk_contact(cap, (0, 0, 0), q, cap, pb, I, 0.5)
    assert k == 1
    s3 = np.array([0.3, -0.5, 0.8], np.float32)
    for seed in ((1, 0, 0), (0, 1, 0), s3 / np.linalg.norm(s3)):
        k, d, *_ = oracle.gjk_contact(cap, (0, 0, 0), q, cap, pb, I, 0.5, seed=seed)
        assert k == 1 and abs(d - ref) < 2e-3


def test_gjk_touch_gap_overlap_boundary(oracle):  # gjk.rs:1538-1573
    s = oracle.sphere(0.5)
    T = 0.0173
    k, d, n, pb, _ = oracle.gjk_contact(s, (0, 0, 0), I, s, (1.0, 0, 0), I, T)
    assert k == 1 and abs(d) < 1e-6
    assert np.linalg.norm(n - [1, 0, 0]) < 1e-6 and np.linalg.norm(pb - [0.5, 0, 0]) < 1e-5
    k, d, *_ = oracle.gjk_contact(s, (0, 0, 0), I, s, (1.0 + 1e-5, 0, 0), I, T)
    assert k == 1 and abs(d - 1e-5) < 1e-6
    k, d, *_ = oracle.gjk_contact(s, (0, 0, 0), I, s, (1.0 - 1e-5, 0, 0), I, T)
    assert k == 1 and abs(d + 1e-5) < 1e-6
    assert oracle.gjk_contact(s, (0, 0, 0), I, s, (1.05, 0, 0), I, T)[0] == 0


def test_gjk_band_threshold_routing(oracle):  # gjk.rs:1576-1594
    s = oracle.sphere(0.5)
    k, d, *_ = oracle.gjk_contact(s, (0, 0, 0), I, s, (1.0199, 0, 0), I, 0.02)
    assert k == 1 and abs(d - 0.0199) < 1e-5
    assert oracle.gjk_contact(s, (0, 0, 0), I, s, (1.0201, 0, 0), I, 0.02)[0] == 0


@pytest.mark.parametrize("scale", [1e-3, 1e-2, 1.0, 1e2, 1e3])
def test_gjk_scale_sweep_sphere_pairs(oracle, scale):  # gjk.rs:1597-1639
    s = oracle.sphere(0.5 * scale)
    for centers, expect_band, expected in ((3.0, False, 2.0), (1.0, True, 0.0), (0.9, True, -0.1),
                                           (0.5, True, -0.5)):
        k, d, n, *_ = oracle.gjk_contact(s, (0, 0, 0), I, s, (centers * scale, 0, 0), I, 0.02 * scale)
        assert k != 2
        assert (k == 1) == expect_band
        if k == 1:
            assert abs(d / scale - expected) < 1e-4
            assert abs(np.linalg.norm(n) - 1.0) < 1e-5 and abs(n[0]) > 0.999


@pytest.mark.parametrize("scale", [1e-3, 1e-2, 1.0, 1e2, 1e3])
def test_gjk_scale_sweep_box_pairs(oracle, scale):  # gjk.rs:1642-1700
    margin = 0.04
    b = oracle.box(0.5 * scale)
    core_half = max(np.float32(0.5 * scale) - np.float32(margin), 0.0)
    thr = 0.02 * scale
    for centers in (3.0 * scale, 1.0 * scale):
        expected = centers - 2.0 * core_half - 2.0 * margin
        k, d, *_ = oracle.gjk_contact(b, (0, 0, 0), I, b, (centers, 0, 0), I, thr)
        assert k != 2
        assert (k == 1) == (expected < thr)
        if k == 1:
            assert abs(d - expected) < max(1e-4 * scale, 1e-6)
    k, *_ = oracle.gjk_contact(b, (0, 0, 0), I, b, (0.3 * scale, 0, 0), I, thr)
    assert k == (2 if core_half > 0 else 1)


# ---------------------------------------------------------------- epa.rs --
def test_epa_penetration_spheres(oracle):  # epa.rs:408-444
    S = oracle.sphere(1.0)
    pb = np.array([1.0, 0.5, 0.3], np.float32)
    n, simplex = oracle.gjk_intersect(S, (0, 0, 0), I, S, pb, I)
    assert n == 4
    st, normal, depth = oracle.epa_pair(simplex, S, (0, 0, 0), I, S, pb, I, core=False)
    assert depth == pytest.approx(2.0 - np.linalg.norm(pb), abs=1e-2)
    assert np.dot(normal, pb / np.linalg.norm(pb)) > 0.9


def test_epa_axis_aligned_overlap_is_exact(oracle):  # epa.rs:446-478
    S = oracle.sphere(1.0)
    n, simplex = oracle.gjk_intersect(S, (0, 0, 0), I, S, (1.5, 0, 0), I)
    st, normal, depth = oracle.epa_pair(simplex, S, (0, 0, 0), I, S, (1.5, 0, 0), I, core=False)
    assert depth == pytest.approx(0.5, abs=1e-3)
    assert normal[0] == pytest.approx(1.0, abs=1e-4)


def test_epa_degenerate_seed_returns_zero(oracle):  # epa.rs:514-530
    v = [0.3, -0.2, 0.5]
    st, normal, depth = oracle.epa_points(np.array([v] * 4, np.float32), np.array([v] * 4, np.float32))
    assert st == 1 and depth == 0.0 and np.all(normal == 0)


def test_epa_point_set_cube(oracle):
    """EPA on a cube point set [-1,1]^3 seeded by an enclosing tetrahedron:
    the penetration depth of a centred cube is its half-extent 1 along an axis."""
    cube = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float32)
    st, normal, depth = oracle.epa_points(tet, cube)
    assert st == 0
    assert depth == pytest.approx(1.0, abs=1e-4)
    assert np.max(np.abs(normal)) == pytest.approx(1.0, abs=1e-4)


def test_raycast_candidates_filters_by_ray_path(oracle):  # axis_sweep3.rs:734-748
    mn = np.array([[1, 1, 1], [10, 10, 10]], np.float32)
    mx = np.array([[2, 2, 2], [11, 11, 11]], np.float32)
    assert oracle.raycast_candidates(mn, mx, [0, 0, 0], [1.5, 1.5, 1.5]).tolist() == [0]
    inside = oracle.raycast_candidates(mn, mx, [1.5, 1.5, 1.5], [4, 4, 4]).tolist()
    assert 0 in inside and 1 not in inside
    assert oracle.raycast_candidates(mn, mx, [3, 3, 3], [4, 4, 4]).tolist() == []


# ------------------------------------------------------ persistent manifold --
# The reference's own PersistentManifold tests (crates/bullet-collision/src/
# manifold.rs:388-868) against the oracle's restatement.
_ID = [0.0, 0.0, 0.0, 1.0]
_Y = [0.0, 1.0, 0.0]
_BT = np.float32(0.02)  # CONTACT_BREAKING_THRESHOLD


def _resting(pen):
    """manifold.rs:397-407: box A at (0, 0.5 - pen, 0) on plane B at the origin."""
    pa = np.array([0.0, 0.5 - pen, 0.0], np.float32)
    pb = np.zeros(3, np.float32)
    return np.array([[0, 0, 0, 0, 1, 0, pen]], np.float32), pa, pb


def test_pm_refresh_reprojects_and_keeps_band_points(oracle):
    """manifold.rs:409-441"""
    m, pa, pb = _resting(0.01)
    e = oracle.pm_new()
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    out = oracle.pm_refresh(e, pa - np.float32(0.015) * np.array(_Y, np.float32), _ID, pb, _ID)
    assert out.shape[0] == 1
    assert abs(out[0, 6] - (-0.005)) < 1e-6
    assert np.array_equal(out[0, 0:3], [0, 0, 0]) and np.array_equal(out[0, 3:6], _Y)
    assert abs(oracle.pm_points(e)[0, 3] - (-0.005)) < 1e-6


def test_pm_refresh_drops_separated_and_drifted_points(oracle):
    """manifold.rs:443-484"""
    m, pa, pb = _resting(0.01)
    e = oracle.pm_new()
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    gone = oracle.pm_refresh(e, pa - np.float32(0.05) * np.array(_Y, np.float32), _ID, pb, _ID)
    assert gone.shape[0] == 0 and oracle.pm_count(e) == 0
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    drifted = oracle.pm_refresh(e, pa + np.array([0.03, 0, 0], np.float32), _ID, pb, _ID)
    assert drifted.shape[0] == 0 and oracle.pm_count(e) == 0


def test_pm_capture_anchors_round_trip(oracle):
    """manifold.rs:486-508"""
    m, pa, pb = _resting(0.01)
    e = oracle.pm_new()
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    p = oracle.pm_points(e)[0]
    wa = pa + p[0:3]
    wb = pb + p[4:7]
    assert np.array_equal(wb, [0, 0, 0])
    assert abs(wa[0]) < 1e-6 and abs(wa[1] - 0.01) < 1e-6 and abs(wa[2]) < 1e-6


def test_pm_capture_merges_and_inherits_the_impulse(oracle):
    """manifold.rs:510-558"""
    m, pa, pb = _resting(0.01)
    e = oracle.pm_new()
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    e[7] = 5.0  # point_mut(0).normal_impulse
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    assert oracle.pm_points(e)[0, 7] == 5.0
    flipped = np.array([[0, 0, 0, 0, -1, 0, 0.01]], np.float32)
    oracle.pm_capture(e, flipped, pa, _ID, pb, _ID, _BT)
    pts = oracle.pm_points(e)
    assert pts.shape[0] == 2 and pts[0, 7] == 5.0 and pts[1, 7] == 0.0


def test_pm_refresh_increments_lifetime_and_capture_resets_it(oracle):
    """manifold.rs:560-622"""
    m, pa, pb = _resting(0.01)
    e = oracle.pm_new()
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    assert oracle.pm_points(e)[0].view(np.uint32)[11] == 0
    for step in range(1, 6):
        out = oracle.pm_refresh(e, pa, _ID, pb, _ID)
        assert out.shape[0] == 1
        assert oracle.pm_points(e)[0].view(np.uint32)[11] == step
        assert out[0].view(np.uint32)[8] == step
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    assert oracle.pm_points(e)[0].view(np.uint32)[11] == 0


def test_pm_redetect_triggers_on_relative_drift_only(oracle):
    """manifold.rs:624-662 (from_rotation_z(0.02) = (0, 0, sin 0.01, cos 0.01))"""
    m, pa, pb = _resting(0.01)
    e = oracle.pm_new()
    assert oracle.pm_needs_redetect(e, pa, _ID, pb, _ID)
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    y = np.array(_Y, np.float32)
    assert not oracle.pm_needs_redetect(e, pa + np.float32(0.004) * y, _ID, pb, _ID)
    assert oracle.pm_needs_redetect(e, pa + np.float32(0.006) * y, _ID, pb, _ID)
    q = [0.0, 0.0, float(np.sin(np.float32(0.01))), float(np.cos(np.float32(0.01)))]
    assert oracle.pm_needs_redetect(e, pa, q, pb, _ID)
    x = np.array([1, 0, 0], np.float32)
    assert not oracle.pm_needs_redetect(e, pa + x, _ID, pb + x, _ID)


def test_pm_capture_overflow_keeps_deepest_four(oracle):
    """manifold.rs:778-826: depth truncation keeps the deep point and the
    first three shallow points in input order (the clustered ones)."""
    pt = lambda x, y, z, pen: [x, y, z, 0, 1, 0, pen]
    batch = np.array([pt(0, 0, 0, 0.5), pt(0.03, 0, 0, 0.1), pt(0, 0, 0.03, 0.1), pt(0, 0.03, 0, 0.1),
                      pt(1, 0, 0, 0.1), pt(0, 1, 0, 0.1)], np.float32)
    e = oracle.pm_new()
    o = np.zeros(3, np.float32)
    oracle.pm_capture(e, batch, o, _ID, o, _ID, _BT)
    pts = oracle.pm_points(e)
    assert pts.shape[0] == 4
    anchors = pts[:, 0:3]
    assert abs(anchors[0, 1] - 0.5) < 1e-6
    shallow = anchors[1:4]
    spread = max(np.linalg.norm(a - b) for a in shallow for b in shallow)
    assert spread < 0.1
    assert np.all(anchors.max(axis=1) < 0.6)


def test_pm_capture_overflow_on_spread_inputs(oracle):
    """manifold.rs:828-868: a shallow near-duplicate outside the matching
    radius is dropped by depth truncation."""
    base = [([0, 0, 0], 0.4), ([1, 0, 0], 0.3), ([0, 0, 1], 0.2), ([1, 0, 1], 0.1)]
    e = oracle.pm_new()
    o = np.zeros(3, np.float32)
    for p, pen in base:
        oracle.pm_capture(e, np.array([p + _Y + [pen]], np.float32), o, _ID, o, _ID, _BT)
    dup = np.array([[0.015, 0, 0.015] + _Y + [0.05]], np.float32)
    oracle.pm_capture(e, dup, o, _ID, o, _ID, _BT)
    anchors = oracle.pm_points(e)[:, 0:3]
    assert anchors.shape[0] == 4
    assert np.any(anchors[:, 1] > 0.3)
    assert not np.any((np.abs(anchors[:, 0] - 0.015) < 1e-6) & (np.abs(anchors[:, 2] - 0.015) < 1e-6))


def test_pm_seed_is_the_first_normal(oracle):
    """manifold.rs:378-385"""
    m, pa, pb = _resting(0.01)
    e = oracle.pm_new()
    assert oracle.pm_seed(e) is None
    oracle.pm_capture(e, m, pa, _ID, pb, _ID, _BT)
    assert np.array_equal(oracle.pm_seed(e), _Y)
