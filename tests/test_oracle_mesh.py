"""The oracle's convex-vs-triangle-mesh path (oracle.hpp: TriMesh, the BVH,
sphere_triangle_contact, reduce_manifold, convex_mesh) against the
reference's own tests (crates/bullet-collision/src/narrowphase/
convex_mesh.rs:274-581, shapes/triangle_mesh.rs:255-300, manifold.rs:
680-707) and against what the mathematics fixes (the BVH query returns
exactly the triangles whose AABB a brute force finds among its leaves'
boxes; reduce keeps the deepest point and a subset).
"""
import numpy as np
import pytest

ID = [0.0, 0.0, 0.0, 1.0]


def _flat_ground(oracle):
    """triangle_mesh.rs / convex_mesh.rs flat_ground(): the unit square at y = 0."""
    return oracle.Mesh([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], [[0, 1, 2], [0, 2, 3]])


def _grid(oracle, n=5, spacing=1.0, height=None):
    """A flat n x n vertex grid centred on the origin (the triangulated
    heightfield of convex_mesh.rs:555-580)."""
    h = (n - 1) * spacing / 2.0
    xs = np.linspace(-h, h, n)
    v = np.array([[x, 0.0 if height is None else height(x, z), z] for z in xs for x in xs], np.float32)
    t = []
    for r in range(n - 1):
        for c in range(n - 1):
            a, b, d, e = r * n + c, r * n + c + 1, (r + 1) * n + c, (r + 1) * n + c + 1
            t += [[a, d, b], [b, d, e]]
    return oracle.Mesh(v, np.array(t, np.uint32))


def _sphere(oracle, r):
    return oracle.Shape(oracle.SPHERE, [r])


def test_local_aabb_and_query_of_flat_ground(oracle):
    """triangle_mesh.rs:258-281: both triangles share the unit box, so the
    query over the second one reports both."""
    m = _flat_ground(oracle)
    hits = m.query([0.8, -1.0, 0.8], [1.2, 1.0, 1.2])
    assert 1 in hits.tolist() and len(hits) == 2


def test_sphere_resting_on_mesh(oracle):
    """convex_mesh.rs:283-307"""
    m = _flat_ground(oracle)
    pts, n = m.collide(_sphere(oracle, 0.5), [0.5, 0.4, 0.5], ID, [0, 0, 0], ID)
    assert n >= 1
    assert abs(pts[0, 4] - (-1.0)) < 1e-3 and abs(pts[0, 6] - 0.1) < 1e-3


def test_sphere_fast_path_matches_gjk(oracle):
    """convex_mesh.rs:309-374: the analytic sphere path against GJK+EPA on the
    triangle as a three-vertex hull."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0.3, 0.2, 0.9]], np.float32)
    m = oracle.Mesh(tri, [[0, 1, 2]])
    s = _sphere(oracle, 0.5)
    th = float(oracle.lib().oracle_threshold_f32(s.c()))
    hull = oracle.Shape(oracle.HULL, hull=tri)
    for c in ([0.4, 0.3, 0.4], [0.95, 0.15, 0.1], [-0.3, 0.3, 0.2], [0.3, 0.55, 0.3], [0.6, 0.05, 0.4]):
        fast, nf = m.collide(s, c, ID, [0, 0, 0], ID)
        slow, path = oracle.convex_convex(s, c, ID, hull, [0, 0, 0], ID, threshold=th)
        assert (nf > 0) == (slow.shape[0] > 0), f"presence at {c}"
        if nf:
            assert abs(fast[0, 6] - slow[0, 6]) <= 5e-3
            assert float(np.dot(fast[0, 3:6], slow[0, 3:6])) > 0.999
            assert np.all(np.abs(fast[0, 0:3] - slow[0, 0:3]) <= 1e-2)


def test_sphere_triangle_edge_vertex_band_backface(oracle):
    """convex_mesh.rs:376-438"""
    m = oracle.Mesh([[0, 0, 0], [1, 0, 0], [0, 0, 1]], [[0, 1, 2]])
    s = _sphere(oracle, 0.5)
    p, _ = m.collide(s, [0.0, 0.4, 0.0], ID, [0, 0, 0], ID)
    assert abs(p[0, 6] - 0.1) < 1e-4 and abs(p[0, 4] + 1.0) < 1e-3 and float(np.dot(p[0, :3], p[0, :3])) < 1e-6
    p, _ = m.collide(s, [0.5, 0.4, 0.0], ID, [0, 0, 0], ID)
    assert abs(p[0, 6] - 0.1) < 1e-4 and abs(p[0, 4] + 1.0) < 1e-3
    assert abs(p[0, 0] - 0.5) < 1e-6 and abs(p[0, 1]) < 1e-6
    th = float(oracle.lib().oracle_threshold_f32(s.c()))
    a, b, c = [0, 0, 0], [1, 0, 0], [0, 0, 1]
    point, normal, pen = oracle.sphere_triangle([0.3, 0.51, 0.3], 0.5, th, a, b, c)
    assert abs(pen + 0.01) < 1e-4 and abs(normal[1] - 1.0) < 1e-3 and abs(point[1]) < 1e-6
    assert oracle.sphere_triangle([0.3, 0.6, 0.3], 0.5, th, a, b, c) is None
    p, _ = m.collide(s, [0.3, -0.4, 0.3], ID, [0, 0, 0], ID)
    assert abs(p[0, 6] - 0.1) < 1e-4 and abs(p[0, 4] - 1.0) < 1e-3 and abs(p[0, 1]) < 1e-6


def test_separated_sphere_does_not_touch_mesh(oracle):
    """convex_mesh.rs:439-456"""
    _, n = _flat_ground(oracle).collide(_sphere(oracle, 0.5), [0.5, 5.0, 0.5], ID, [0, 0, 0], ID)
    assert n == 0


def test_box_over_many_triangles_reduces_contacts(oracle):
    """convex_mesh.rs:554-580: a unit box over the four cells around the
    origin of a flat 5 x 5 grid: at most four points, the deepest first
    with the -Y normal and penetration 0.1."""
    m = _grid(oracle, 5)
    box = oracle.Shape(oracle.BOX, [0.5, 0.5, 0.5])
    pts, n = m.collide(box, [0.0, 0.4, 0.0], ID, [0, 0, 0], ID)
    assert 1 <= n <= 4
    assert abs(pts[0, 4] + 1.0) < 1e-3 and abs(pts[0, 6] - 0.1) < 1e-3


def test_mesh_first_flips_the_normal(oracle):
    """narrowphase/mod.rs:177-190: mesh as shape A -> the manifold flipped."""
    m = _flat_ground(oracle)
    s = _sphere(oracle, 0.5)
    a, _ = m.collide(s, [0.5, 0.4, 0.5], ID, [0, 0, 0], ID)
    b, _ = m.collide(s, [0.5, 0.4, 0.5], ID, [0, 0, 0], ID, mesh_first=True)
    assert np.array_equal(a[:, 3:6], -b[:, 3:6]) and np.array_equal(a[:, [0, 1, 2, 6]], b[:, [0, 1, 2, 6]])


def test_bvh_query_is_the_brute_force_leaf_set(oracle):
    """The query returns each triangle at most once, and a triangle whose own
    AABB overlaps the query box is always among them (the leaves' boxes
    contain their triangles)."""
    rng = np.random.default_rng(5)
    m = _grid(oracle, 17, 0.5, height=lambda x, z: 0.3 * np.sin(x) * np.cos(z))
    tri = m.v[m.i]
    tmin, tmax = tri.min(axis=1), tri.max(axis=1)
    for _ in range(200):
        c = rng.uniform(-4.5, 4.5, 3).astype(np.float32)
        e = rng.uniform(0.05, 1.5, 3).astype(np.float32)
        qmin, qmax = c - e, c + e
        hits = m.query(qmin, qmax)
        assert len(set(hits.tolist())) == len(hits)
        must = np.nonzero(np.all((tmin <= qmax) & (qmin <= tmax), axis=1))[0]
        assert set(must.tolist()) <= set(hits.tolist())


def test_reduce_keeps_the_deepest_and_spreads_the_rest(oracle):
    """manifold.rs:680-700"""
    pt = lambda x, y, z, pen: [x, y, z, 0, 1, 0, pen]
    pts = np.array([pt(0, 0, 0, 0.5), pt(0.01, 0, 0, 0.4), pt(1, 0, 0, 0.1), pt(0, 1, 0, 0.1), pt(0, 0, 1, 0.1),
                    pt(1, 1, 1, 0.05)], np.float32)
    r = oracle.reduce(pts, 4)
    assert r.shape[0] == 4
    assert np.array_equal(r[0, :3], [0, 0, 0]) and r[0, 6] == np.float32(0.5)
    assert np.all(np.sum(r[1:, :3] ** 2, axis=1) > 0.5)


def test_reduce_leaves_small_manifolds_alone(oracle):
    """manifold.rs:702-707"""
    pts = np.array([[0, 0, 0, 0, 1, 0, 0.2], [1, 0, 0, 0, 1, 0, 0.1], [0, 1, 0, 0, 1, 0, 0.1]], np.float32)
    assert np.array_equal(oracle.reduce(pts, 4), pts)


# ------------------------------------------------------------- CCD sweeps --
# The reference's own sweep tests (crates/bullet-collision/src/narrowphase/
# convex_sweep.rs:526-668) against the oracle.
def test_sweep_head_on_box_hit(oracle):
    """convex_sweep.rs:535-552"""
    b = oracle.Shape(oracle.BOX, [1, 1, 1])
    f, n, p = oracle.sweep_sphere([5, 0, 0], [0, 0, 0], 0.5, b)
    assert abs(f - 3.5 / 5.0) < 1e-4 and abs(n[0] - 1.0) < 1e-3
    assert abs(p[0] - 1.0) < 1e-3 and abs(p[1]) < 1e-3 and abs(p[2]) < 1e-3


def test_sweep_offset_cast_misses_the_box(oracle):
    """convex_sweep.rs:554-559"""
    assert oracle.sweep_sphere([5, 3, 0], [-5, 3, 0], 0.5, oracle.Shape(oracle.BOX, [1, 1, 1])) is None


def test_sweep_plane_cases(oracle):
    """convex_sweep.rs:561-586: analytic entry, moving away, already touching."""
    pl = oracle.Shape(oracle.PLANE, [0, 1, 0, 0])
    f, n, p = oracle.sweep_sphere([0, 5, 0], [0, -5, 0], 0.5, pl)
    assert abs(f - 0.45) < 1e-5 and abs(n[1] - 1.0) < 1e-5 and abs(p[1]) < 1e-5
    assert oracle.sweep_sphere([0, 2, 0], [0, 5, 0], 0.5, pl) is None
    f, n, _ = oracle.sweep_sphere([0, 0.3, 0], [0, -3, 0], 0.5, pl)
    assert f == 0.0 and n[1] == 1.0


def test_sweep_sphere_against_a_sloped_triangle(oracle):
    """convex_sweep.rs:588-610"""
    m = oracle.Mesh([[0, 0, 0], [2, 2, 0], [0, 0, 2]], [[0, 1, 2]])
    f, n, _ = oracle.sweep_sphere_mesh([1, 3, 0.6], [1, -1, 0.6], 0.2, m)
    assert abs(f - (3.0 - (1.0 + 0.4 * np.sqrt(0.5))) / 4.0) < 1e-3
    assert n[1] > 0.6 and n[0] < -0.1


def test_convex_cast_cases(oracle):
    """convex_sweep.rs:616-650: box head on, box offset, capsule on a plane."""
    mv = oracle.Shape(oracle.BOX, [0.5, 0.5, 0.5])
    tg = oracle.Shape(oracle.BOX, [1, 1, 1])
    f, n, p = oracle.convex_cast(mv, [5, 0, 0], [0, 0, 0, 1], [0, 0, 0], tg)
    assert abs(f - 3.5 / 5.0) < 1e-4 and abs(n[0] - 1.0) < 1e-3 and abs(p[0] - 1.0) < 1e-3
    assert oracle.convex_cast(mv, [5, 3, 0], [0, 0, 0, 1], [-5, 3, 0], tg) is None
    cap = oracle.Shape(oracle.CAPSULE, [0.4, 0.8])
    f, n, p = oracle.convex_cast(cap, [0, 3, 0], [0, 0, 0, 1], [0, -1, 0], oracle.Shape(oracle.PLANE, [0, 1, 0, 0]))
    assert abs(f - (3.0 - 1.2) / 4.0) < 1e-4 and abs(n[1] - 1.0) < 1e-4 and abs(p[1]) < 1e-4


def test_sweep_matches_a_dense_scan(oracle):
    """Closed-form check of the search: for a segment that ends inside the
    inflated box, the entry fraction is where the centre first comes within
    the radius of the box, which a dense scan of the segment against the
    exact box distance brackets; a segment that stays clear reports no hit.
    (A segment that passes through and leaves again is not used: the
    reference's ternary search settles at the far edge of the overlap and
    can then report no hit, convex_sweep.rs:161-176, which the restatement
    reproduces.  The hit normal is not checked here: it is the direction of
    the GJK distance walk at a probe point, and the walk's tetrahedron step
    keeps the first face whose side test fails rather than the nearest one,
    gjk.rs:1024-1039, so near edges it can be far from the exact normal.)"""
    rng = np.random.default_rng(4)
    h = np.array([0.7, 0.4, 0.9])
    b = oracle.Shape(oracle.BOX, list(h))
    for k in range(60):
        frm = rng.uniform(-4, 4, 3)
        r = rng.uniform(0.1, 0.6)
        to = rng.uniform(-h, h) if k % 3 else frm + rng.normal(0, 0.5, 3)
        s = np.linspace(0, 1, 20001)
        pts = frm[None, :] + (to - frm)[None, :] * s[:, None]
        q = np.clip(pts, -h, h)
        d = np.linalg.norm(pts - q, axis=1)
        inside = d <= r
        hit = oracle.sweep_sphere(frm, to, r, b)
        if k % 3 == 0:
            if d.min() > r + 1e-3:
                assert hit is None
            continue
        assert inside[-1] and hit is not None
        first = s[np.argmax(inside)]
        assert abs(hit[0] - first) <= 1e-4


def test_box_cast_against_a_mesh_slope(oracle):
    """convex_sweep.rs:651-668"""
    m = oracle.Mesh([[0, 0, 0], [2, 2, 0], [0, 0, 2]], [[0, 1, 2]])
    sc = {"position": np.zeros((1, 3), np.float32), "orientation": np.array([[0, 0, 0, 1]], np.float32),
          "shape_type": np.array([oracle.BOX], np.uint8), "shape_params": np.array([[0.4, 0.4, 0.4, 0]], np.float32),
          "hull_first": np.zeros(1, np.uint32), "hull_count": np.zeros(1, np.uint32),
          "hull_vertices": np.zeros((0, 3), np.float32)}
    hit, out = oracle.sweep_mesh_scene(sc, [m], [0], [[0, 0, 0, 0, 0, 0, 1]], [[1, 3, 0.6]], [[1, -1, 0.6]],
                                       caster=[0])
    assert hit[0] == 1
    assert out[0, 2] > 0.6 and 0.0 < out[0, 0] < 1.0


def test_box_cast_matches_the_slab_entry(oracle):
    """Closed-form check of convex_cast's search (convex_sweep.rs:387-429):
    an axis-aligned box cast against an axis-aligned box enters their
    Minkowski sum, a box of the summed half extents about the target, at the
    slab-test entry parameter.  The search stops where the GJK distance
    first drops to 1e-4 m (d2 <= 1e-8), which is 1e-4 / |dir . n| before the
    face with normal n, so the bound is written in those terms."""
    rng = np.random.default_rng(12)
    for _ in range(60):
        h1, h2 = rng.uniform(0.2, 1.0, 3), rng.uniform(0.2, 1.0, 3)
        op = rng.uniform(-2, 2, 3)
        H = h1 + h2
        to = op + rng.uniform(-0.9, 0.9, 3) * H
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        frm = op + d * rng.uniform(1.5, 3.0) * np.linalg.norm(H)
        hit = oracle.convex_cast(oracle.Shape(oracle.BOX, list(h1)), frm, ID, to,
                                 oracle.Shape(oracle.BOX, list(h2)), op)
        assert hit is not None
        seg = to - frm
        length = np.linalg.norm(seg)
        t_in, axis = 0.0, -1
        for i in range(3):
            lo, hi = op[i] - H[i], op[i] + H[i]
            if lo <= frm[i] <= hi:
                continue
            t = ((lo if frm[i] < lo else hi) - frm[i]) / seg[i]
            if t > t_in:
                t_in, axis = t, i
        assert axis >= 0
        slope = abs(seg[axis]) / length
        assert t_in - hit[0] >= -2e-6
        assert (t_in - hit[0]) * length * slope <= 1.5e-4


def test_sphere_cast_agrees_with_the_sphere_sweep(oracle):
    """convex_cast with a sphere of radius r and swept_sphere_cast of radius r
    search the same Minkowski set (the target grown by r) from two sides:
    the cast walks target (+) -sphere about the origin
    (convex_sweep.rs:387-429), the sweep walks the grown target about the
    moving centre (:105-125).  Their entry fractions agree to the search's
    resolution (1e-4 m of distance, converted along the hit normal) on most
    boxes, capsules and hulls at random poses.  Not on all: on curved sets
    the distance walk (convex_sweep.rs:59-103) has no progress stop, keeps
    adding nearly coplanar points, and a tetrahedron of volume ~1e-8 (above
    closest_origin_tetrahedron's absolute 1e-12 flatness gate, gjk.rs:
    1008-1022) can pass all four side tests with the origin outside; the
    walk then reports zero distance and the search an early entry.  Both
    restated walks do this as the reference's does, at different points, so
    the check asks for agreement on three in four hits and a small median.
    A sign or operand mistake in either support sum fails nearly every
    case."""
    rng = np.random.default_rng(13)
    pts = rng.normal(size=(24, 3)) * np.array([0.8, 0.5, 0.6])
    targets = [oracle.Shape(oracle.BOX, [0.7, 0.4, 0.9]), oracle.Shape(oracle.CAPSULE, [0.3, 0.6]),
               oracle.Shape(oracle.HULL, hull=pts)]
    diffs = []
    for k in range(90):
        tg = targets[k % 3]
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        op = rng.uniform(-1, 1, 3)
        r = rng.uniform(0.1, 0.5)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        frm = op + d * rng.uniform(2.0, 4.0)
        to = op + rng.normal(0, 0.2, 3)
        a = oracle.sweep_sphere(frm, to, r, tg, op, q)
        b = oracle.convex_cast(oracle.Shape(oracle.SPHERE, [r]), frm, ID, to, tg, op, q)
        assert (a is None) == (b is None)
        if a is None:
            continue
        seg = to - frm
        slope = abs(np.dot(seg, a[1])) / np.linalg.norm(seg)
        diffs.append(abs(a[0] - b[0]) * np.linalg.norm(seg) * slope)
    diffs = np.array(diffs)
    assert diffs.shape[0] > 60
    assert np.median(diffs) <= 1e-4
    assert np.mean(diffs <= 2e-4) >= 0.75
