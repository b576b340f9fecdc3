"""The oracle against what the mathematics fixes (CPU only).

Closed forms, brute force and invariants chosen so that a dropped term, a
wrong sign/index or a transposed operand in the restatement fails a check:
  * broadphase: the sweep and the DbvtBroadPhase restatement (dbvt.rs, the
    CPU baseline's broadphase) equal brute force (the definition,
    brute_force.rs:54-67) on the reference's own scattered-box inputs
    (axis_sweep3.rs:570-593) and on random/degenerate sets;
  * world AABBs against float64 corner/segment/vertex sweeps;
  * hull support = brute-force argmax (first maximum);
  * sphere-sphere / sphere-plane depths in closed form;
  * GJK band distance for capsule/sphere cores = segment distance - radii;
  * box-box SAT depth/normal vs the independent GJK/EPA route on the same boxes;
  * contact_breaking_threshold = 0.02 (|extent| + |centre|) of the local AABB.
"""
import numpy as np
import pytest

I = (0.0, 0.0, 0.0, 1.0)


def quat_to_mat(q):
    x, y, z, w = [float(v) for v in q]
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rand_quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(np.float32)


def test_sweep_equals_brute_on_reference_scattered_boxes(oracle):
    """matches_brute_force_on_scattered_boxes inputs (axis_sweep3.rs:570-593)."""
    A, Cc, M = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    seed = 0x9E3779B97F4A7C15
    mn, mx = [], []
    for _ in range(64):
        seed = (seed * A + Cc) & M
        base = np.array([((seed >> 32) % 1000) / 10.0, ((seed >> 16) % 1000) / 10.0, (seed % 1000) / 10.0],
                        np.float32)
        seed = (seed * A + Cc) & M
        ext = np.array([((seed >> 32) % 100) / 100.0 + 0.1, ((seed >> 16) % 100) / 100.0 + 0.1,
                        (seed % 100) / 100.0 + 0.1], np.float32)
        mn.append(base)
        mx.append(base + ext)
    mn, mx = np.array(mn), np.array(mx)
    bp = oracle.brute_pairs(mn, mx)
    assert np.array_equal(oracle.sweep_pairs(mn, mx), bp)
    assert np.array_equal(oracle.dbvt_pairs(mn, mx), bp)
    # brute force against numpy's reading of util.rs:28-38
    ov = np.all((mn[:, None] <= mx[None]) & (mn[None] <= mx[:, None]), axis=2)
    ii, jj = np.nonzero(np.triu(ov, 1))
    assert np.array_equal(bp, np.stack([ii, jj], 1).astype(np.uint32))


@pytest.mark.parametrize("seed", range(4))
def test_sweep_equals_brute_random_and_degenerate(oracle, seed):
    rng = np.random.default_rng(seed)
    n = 1500
    mn = rng.uniform(-20, 20, (n, 3)).astype(np.float32)
    mx = mn + rng.uniform(0.0, 3.0, (n, 3)).astype(np.float32)
    mn[:50] = np.round(mn[:50])          # touching faces / shared coordinates
    mx[:50] = mn[:50] + 1
    mx[50:60] = mn[50:60]                # degenerate points
    mn[60], mx[60] = -1e30, 1e30         # plane-like
    mn[61:64] = np.nan                   # dead proxies
    mx[61:64] = np.nan
    bp = oracle.brute_pairs(mn, mx)
    assert np.array_equal(oracle.sweep_pairs(mn, mx), bp)
    assert np.array_equal(oracle.dbvt_pairs(mn, mx), bp)


def test_world_aabbs_match_float64_sweeps(oracle, scenes):
    sc = scenes.mixed_convex(3000, 4)
    mn, mx = oracle.scene_aabbs(sc, 0.0)
    for i in range(sc["shape_type"].shape[0]):
        t = sc["shape_type"][i]
        R = quat_to_mat(sc["orientation"][i])
        p = sc["position"][i].astype(np.float64)
        a = sc["shape_params"][i].astype(np.float64)
        if t == 1:
            c = np.array([[sx * a[0], sy * a[1], sz * a[2]] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
            w = c @ R.T + p
            lo, hi = w.min(0), w.max(0)
        elif t == 0:
            lo, hi = p - a[0], p + a[0]
        elif t == 2:
            e = np.array([[0, a[1], 0], [0, -a[1], 0]]) @ R.T + p
            lo, hi = e.min(0) - a[0], e.max(0) + a[0]
        else:
            f, k = sc["hull_first"][i], sc["hull_count"][i]
            w = sc["hull_vertices"][f:f + k].astype(np.float64) @ R.T + p
            lo, hi = w.min(0), w.max(0)
        assert np.allclose(mn[i], lo, atol=2e-5), i
        assert np.allclose(mx[i], hi, atol=2e-5), i


def test_margin_pads_every_side(oracle, scenes):
    sc = scenes.mixed_convex(500, 4)
    a0, b0 = oracle.scene_aabbs(sc, 0.0)
    a1, b1 = oracle.scene_aabbs(sc, 0.04)
    assert np.array_equal(a1, a0 - np.float32(0.04)) and np.array_equal(b1, b0 + np.float32(0.04))


def test_hull_support_is_first_argmax(oracle):
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(40, 3)).astype(np.float32)
    pts[7] = pts[3]  # duplicate vertex: the first one wins (strict >, convex_hull.rs:58-63)
    h = oracle.hull(pts)
    for _ in range(50):
        d = rng.normal(size=3).astype(np.float32)
        s = oracle.support_world(h, (0, 0, 0), I, d)
        dots = (pts * d).sum(1, dtype=np.float32)
        assert np.array_equal(s, pts[int(np.argmax(dots))])


def test_box_and_capsule_core_supports(oracle):
    b = oracle.box(0.5, 0.03, 1.0)
    s = oracle.support_world(b, (0, 0, 0), I, (1, -1, 1), core=True)
    assert np.allclose(s, [0.46, 0.0, -0.0 + 0.96], atol=1e-6)   # h - 0.04 clamped at 0
    c = oracle.capsule(0.3, 0.7)
    assert np.allclose(oracle.support_world(c, (0, 0, 0), I, (0.2, -3, 1), core=True), [0, -0.7, 0])
    assert np.allclose(oracle.support_world(c, (0, 0, 0), I, (0, 0, 0)), [0.3, 0.7, 0])


def test_contact_breaking_threshold_closed_form(oracle):
    assert oracle.threshold(oracle.box(1.0, 2.0, 3.0)) == pytest.approx(0.02 * np.sqrt(14), rel=1e-6)
    assert oracle.threshold(oracle.capsule(0.5, 1.0)) == pytest.approx(0.02 * np.sqrt(0.5 + 2.25), rel=1e-6)
    off = np.array([[2, 0, 0], [3, 1, 0], [2.5, -1, 1]], np.float32)   # off-centre hull adds |centre|
    lo, hi = off.min(0), off.max(0)
    exp = 0.02 * (np.linalg.norm(hi - lo) * 0.5 + np.linalg.norm((lo + hi) * 0.5))
    assert oracle.threshold(oracle.hull(off)) == pytest.approx(exp, rel=1e-5)
    assert np.isinf(oracle.threshold(oracle.plane((0, 1, 0), 0)))


def test_sphere_pairs_closed_form(oracle):
    rng = np.random.default_rng(6)
    for _ in range(200):
        ra, rb = rng.uniform(0.1, 1.0, 2)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dist = rng.uniform(0.05, (ra + rb) * 1.1)
        pa = rng.uniform(-5, 5, 3).astype(np.float32)
        pb = (pa + d * dist).astype(np.float32)
        m, _ = oracle.collide(oracle.sphere(ra), pa, I, oracle.sphere(rb), pb, I)
        true = np.linalg.norm(pb.astype(np.float64) - pa)
        if true < ra + rb - 1e-5:
            assert m.shape[0] == 1
            assert m[0, 6] == pytest.approx(ra + rb - true, abs=2e-6)
            n = (pb - pa) / true
            assert np.allclose(m[0, 3:6], n, atol=1e-6)
            assert np.allclose(m[0, :3], pa + n * (ra - m[0, 6] / 2), atol=1e-5)
        elif true > ra + rb + 1e-5:
            assert m.shape[0] == 0


def test_plane_depth_closed_form(oracle):
    """Plane-convex depth = c - n.support(-n) for a rotated box (plane.rs:19-56)."""
    rng = np.random.default_rng(7)
    nrm = np.array([0.3, 1.0, -0.2])
    nrm /= np.linalg.norm(nrm)
    for _ in range(100):
        h = rng.uniform(0.2, 1.0, 3)
        q = rand_quat(rng)
        R = quat_to_mat(q)
        pos = (nrm * rng.uniform(-0.5, 1.5) + rng.normal(size=3) * 0.3).astype(np.float32)
        corners = np.array([[sx * h[0], sy * h[1], sz * h[2]] for sx in (-1, 1) for sy in (-1, 1)
                            for sz in (-1, 1)]) @ R.T + pos
        side = 1.0 if np.dot(nrm, pos) >= 0 else -1.0
        depth = -(corners @ (side * nrm)).min()   # c = 0 at the origin
        m, _ = oracle.collide(oracle.plane(nrm, 0.0), (0, 0, 0), I, oracle.box(*h), pos, q)
        thr = oracle.threshold(oracle.box(*h))
        if depth >= -thr + 1e-4:
            assert m.shape[0] == 1 and m[0, 6] == pytest.approx(depth, abs=2e-5)
            assert np.allclose(m[0, 3:6], side * nrm, atol=1e-5)
        elif depth < -thr - 1e-4:
            assert m.shape[0] == 0


def seg_seg_distance(p1, q1, p2, q2, samples=4001):
    """Dense-sample minimum distance between two segments (float64)."""
    t = np.linspace(0, 1, samples)
    a = p1[None] + t[:, None] * (q1 - p1)[None]
    best = np.inf
    d2 = q2 - p2
    for pt in a:
        s = np.clip(np.dot(pt - p2, d2) / max(np.dot(d2, d2), 1e-300), 0, 1)
        best = min(best, np.linalg.norm(pt - (p2 + s * d2)))
    return best


def test_gjk_capsule_band_distance_closed_form(oracle):
    """Capsule cores are segments, so the walk's signed distance is exactly
    segment distance - (ra + rb) (gjk.rs:255-279); threshold = +inf."""
    rng = np.random.default_rng(8)
    for _ in range(40):
        ra, rb = rng.uniform(0.1, 0.5, 2)
        ha, hb = rng.uniform(0.2, 1.0, 2)
        qa, qb = rand_quat(rng), rand_quat(rng)
        pa = np.zeros(3, np.float32)
        pb = (rng.normal(size=3) * 1.5).astype(np.float32)
        Ra, Rb = quat_to_mat(qa), quat_to_mat(qb)
        sa = Ra @ [0, ha, 0]
        sb = Rb @ [0, hb, 0]
        dseg = seg_seg_distance(pa - sa, pa + sa, pb - sb, pb + sb)
        k, d, n, _, _ = oracle.gjk_contact(oracle.capsule(ra, ha), pa, qa, oracle.capsule(rb, hb), pb, qb,
                                           np.inf)
        if dseg > 1e-3:
            assert k == 1
            assert d == pytest.approx(dseg - ra - rb, abs=2e-4)
            assert abs(np.linalg.norm(n) - 1) < 1e-5


def _min_overlap(ha, hb, Rb, pb):
    """Minimum projected overlap of two boxes (A at the origin, unrotated)
    over the 15 SAT axes, float64: the exact penetration depth of boxes."""
    axes = [np.eye(3)[:, i] for i in range(3)] + [Rb[:, i] for i in range(3)]
    axes += [np.cross(np.eye(3)[:, i], Rb[:, j]) for i in range(3) for j in range(3)]
    best = np.inf
    for a in axes:
        l = np.linalg.norm(a)
        if l < 1e-9:
            continue
        a = a / l
        best = min(best, np.sum(ha * np.abs(a)) + np.sum(hb * np.abs(Rb.T @ a)) - abs(np.dot(pb, a)))
    return best


def test_box_box_routes_match_their_closed_forms(oracle):
    """Box-box penetration through two routes, each against its closed form:
    ODE SAT (box_box.rs) on the sharp boxes, and the margin-core GJK/EPA walk
    (convex_convex.rs) on the rounded boxes (core h - 0.04 plus a 0.04 ball,
    box_box core: box_shape.rs:42-53), whose depth is the cores' minimum
    overlap plus both margins."""
    rng = np.random.default_rng(9)
    checked = cap_exits = full = 0
    for _ in range(300):
        ha = rng.uniform(0.5, 1.0, 3)
        hb = rng.uniform(0.5, 1.0, 3)
        ang = rng.normal(size=3) * 0.08
        qb = np.array([ang[0] / 2, ang[1] / 2, ang[2] / 2, 1.0])
        qb = (qb / np.linalg.norm(qb)).astype(np.float32)
        axis = rng.integers(3)
        pb = np.zeros(3, np.float32)
        pb[axis] = (ha[axis] + hb[axis]) * rng.choice([-1, 1]) - rng.uniform(0.05, 0.15) * np.sign(rng.normal())
        pb += (rng.normal(size=3) * 0.1).astype(np.float32) * (np.arange(3) != axis)
        ha32, hb32 = ha.astype(np.float32).astype(np.float64), hb.astype(np.float32).astype(np.float64)
        A, B = oracle.box(*ha), oracle.box(*hb)
        sat, _ = oracle.collide(A, (0, 0, 0), I, B, pb, qb)
        gjk, _ = oracle.convex_convex(A, (0, 0, 0), I, B, pb, qb)
        Rb = quat_to_mat(qb)
        sharp = _min_overlap(ha32, hb32, Rb, pb.astype(np.float64))
        rounded = _min_overlap(ha32 - 0.04, hb32 - 0.04, Rb, pb.astype(np.float64)) + 0.08
        if sharp > 1e-3:
            assert sat.shape[0] >= 1 and np.all(sat[:, 6] >= 0) and sat.shape[0] <= 4
            deepest = sat[np.argmax(sat[:, 6])]
            n = deepest[3:6].astype(np.float64)
            assert abs(np.linalg.norm(n) - 1) < 1e-5
            # overlap of the boxes along the reported normal
            ov_n = np.sum(ha32 * np.abs(n)) + np.sum(hb32 * np.abs(Rb.T @ n)) - abs(np.dot(pb, n))
            assert np.dot(n, pb) > 0                       # A -> B
            face = max(abs(np.dot(n, a)) for a in list(np.eye(3)) + list(Rb.T)) > 0.99999
            if face:
                assert deepest[6] <= ov_n + 1e-4           # clipping only removes depth
            else:
                # edge-edge: the SAT overlap carries the 1e-5 fudge on every
                # |R| term, divided by |ua x ub| (box_box.rs:270-283, 371-379)
                crosses = [np.cross(np.eye(3)[:, i], Rb[:, j]) for i in range(3) for j in range(3)]
                l = max(crosses, key=lambda c: abs(np.dot(c, n)) / max(np.linalg.norm(c), 1e-30))
                slack = 1e-5 * 2 * (ha32.sum() + hb32.sum()) / np.linalg.norm(l)
                assert deepest[6] <= ov_n + 1e-4 + slack
            assert ov_n <= 1.05 * sharp + 1e-4             # FUDGE 1.05 favours face axes (box_box.rs:8)
            full += abs(deepest[6] - ov_n) < 1e-4
        if rounded > 1e-3:
            if gjk.shape[0] == 0:
                # The reference walk can cycle between tied supports of two
                # near-parallel core faces (delta misses the 1 - 1e-6
                # relative exit by ~1e-8) and leave at the 64-iteration cap
                # with Separated (gjk.rs:211, 285-287): counted, bounded.
                cap_exits += 1
                continue
            assert gjk[0, 6] == pytest.approx(rounded, abs=2e-4)
            checked += 1
    assert checked > 100
    assert cap_exits <= 0.1 * checked
    assert full > 100   # most deepest vertices lie inside the reference face


def test_raycast_slab_test_against_dense_sampling(oracle):
    """segment_hits_aabb (broadphase/mod.rs:153-183) against an independent
    check: a segment that has a sampled point inside a box must be reported;
    a reported hit whose overlap interval is long must have one.  Includes
    axis-parallel and zero-length segments."""
    rng = np.random.default_rng(5)
    n = 300
    lo = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    mn, mx = lo, (lo + rng.uniform(0.1, 2.0, (n, 3))).astype(np.float32)
    ts = np.linspace(0.0, 1.0, 4001)
    for k in range(40):
        f = rng.uniform(-6, 6, 3).astype(np.float32)
        t = rng.uniform(-6, 6, 3).astype(np.float32)
        if k % 5 == 1:
            t[k % 3] = f[k % 3]          # parallel to a slab
        if k % 10 == 7:
            t = f.copy()                 # zero length
        got = set(oracle.raycast_candidates(mn, mx, f, t).tolist())
        pts = f.astype(np.float64)[None, :] + ts[:, None] * (t.astype(np.float64) - f)[None, :]
        eps = 1e-4  # sampled points must be clearly inside (f32 rounding at the faces)
        inside = ((pts[None, :, :] >= mn[:, None, :] + eps) & (pts[None, :, :] <= mx[:, None, :] - eps)).all(axis=2)
        sampled = set(np.nonzero(inside.any(axis=1))[0].tolist())
        assert sampled <= got, f"segment {k}: sampled hits missing {sorted(sampled - got)}"
        deep = set(np.nonzero(inside.sum(axis=1) >= 8)[0].tolist())
        assert deep <= got
        for i in got - sampled:
            # a reported hit that no sample found must graze: within 1e-3 of the box
            d = np.maximum(np.maximum(mn[i] - pts, pts - mx[i]), 0).max(axis=1).min()
            assert d < 1e-3 * (1 + np.abs(pts).max()), f"segment {k}: spurious hit on box {i} (gap {d})"


def test_signed_distance_closed_forms(oracle):
    """signed_distance (the walk with threshold = +inf, then EPA): sphere
    pairs (distance - radii) at any gap, aligned boxes (gap or overlap
    along the axis), witness identity pa = pb - n sd."""
    I = [0, 0, 0, 1]
    rng = np.random.default_rng(6)
    for _ in range(50):
        ra, rb = rng.uniform(0.2, 1.0, 2).astype(np.float32).astype(np.float64)  # shape params are f32
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dist = rng.uniform(0.05, 5.0) + ra + rb
        out, st = oracle.signed_distance(oracle.sphere(ra), [0, 0, 0], I, oracle.sphere(rb), d * dist, I, f64=True)
        assert st == oracle.SD_WALK
        assert abs(out[0] - (dist - ra - rb)) < 1e-12
        assert np.allclose(out[1:4], d, atol=1e-9)
        assert np.allclose(out[4:7], out[7:10] - out[1:4] * out[0], atol=1e-12)
    for gap in (0.3, 2.0, -0.1, -0.05):
        b = oracle.box(0.5)
        out, st = oracle.signed_distance(b, [0, 0, 0], I, b, [1.0 + gap, 0.1, -0.2], I, f64=True)
        assert abs(out[0] - gap) < 1e-9, (gap, out[0])
        assert np.allclose(out[1:4], [1, 0, 0], atol=1e-9)
        assert st == (oracle.SD_WALK if gap > -0.08 else oracle.SD_EPA) or gap < 0


def test_hull_pair_scene_targets(oracle, scenes):
    """The differentiation scene puts separated pairs at their target signed
    distance and overlaps the rest."""
    sc = scenes.hull_pairs(512, 4)
    out, st = oracle.sd_pairs(sc, sc["pairs"], None, f64=True)
    sep = np.arange(512) % 2 == 0
    assert np.abs(out[sep, 0] - sc["target"][sep]).max() < 1e-5
    assert (out[~sep, 0] < 0).all()
    assert (out[~sep, 0] >= sc["target"][~sep] - 1e-5).all()


def test_signed_distance_reference_adjoint_matches_f64_differences(oracle, scenes):
    """The oracle's f64 adjoint of signed_distance (the reference gradient for
    the GPU adjoint) against f64 central differences at h = 1e-4: on every
    pair where the function is smooth within +-h (second differences halve
    with h), within 1e-3 relative; the median far below."""
    from sd_helpers import fd
    sc = scenes.hull_pairs(96, 6, spacing=0.0)
    P = sc["pairs"].shape[0]
    g10 = np.random.default_rng(3).normal(size=(P, 10))
    g10[:, 4:] = 0.0  # sd and normal (the witnesses jump at tied supports)
    ref, fails = oracle.sd_vjp_pairs(sc, sc["pairs"], g10)
    assert fails == 0
    vjp, j1, _, _ = fd(oracle, sc, g10, 1e-4)
    _, j2, _, _ = fd(oracle, sc, g10, 5e-5)
    scale = np.maximum(np.linalg.norm(vjp, axis=1), 1e-2)
    err = np.linalg.norm(ref - vjp, axis=1) / scale
    ratio = j2 / np.maximum(j1, 1e-300)
    smooth = np.all((j1 <= 1e-3 * scale[:, None]) | ((ratio >= 0.3) & (ratio <= 0.7)), axis=1)
    assert smooth.sum() >= 0.9 * P
    assert err[smooth].max() < 1e-3
    assert np.median(err[smooth]) < 1e-5


def test_capped_walk_reports_its_closest_iterate(oracle, scenes):
    """An f32 walk that misses the 1 - 1e-6 exit by rounding cycles through
    worse simplices until the 64-iteration cap.  The signed distance then
    reports the closest iterate, which meets the f64 walk (converged on the
    same pairs) to 1e-4 m on at least 90 % of them and to 1 cm on all (a walk
    that starts cycling before it nears the minimum keeps a coarser best);
    the last iterate of the cycle was up to 9 cm off (box-hull pairs of the
    de-penetration worlds)."""
    sc = scenes.batched_boxes_hulls(16, 5)
    pairs = oracle.world_pairs(sc, 0.04)
    o32, st32 = oracle.sd_pairs(sc, pairs)
    o64, st64 = oracle.sd_pairs(sc, pairs, None, f64=True)
    capped = st32 == oracle.SD_CAPPED
    both = capped & (st64 == oracle.SD_WALK)
    assert capped.sum() >= 20 and both.sum() >= 0.9 * capped.sum()
    err = np.abs(o32[both, 0] - o64[both, 0])
    close = err < 1e-4
    assert close.mean() >= 0.9 and err.max() < 1e-2
    assert np.abs(o32[both, 1:4] - o64[both, 1:4])[close].max() < 1e-3


@pytest.mark.parametrize("shift", [0.0, 20000.0])
def test_walk_that_proved_a_gap_never_reports_an_overlap(oracle, scenes, shift):
    """The signed-distance walk runs with threshold +inf, so it goes on past
    the point where the reference's band exit (gjk.rs:241-244) would have
    stopped a separated pair.  An iterate whose support plane lies beyond the
    origin (delta = v . w > 0) bounds the core distance from below by
    delta / |v|; once that bound exceeds 1e-4 m the cores are apart, and a
    later simplex that encloses or touches the origin is rounding (a flat
    simplex, or large coordinates).  Without the rule such pairs went on to
    EPA and reported penetrations of 0.2-0.6 m for cores 0.1 m apart (the
    de-penetration worlds, x up to 21 km).  Held, on 16 worlds of boxes and
    hulls (at the origin, and moved to x = 20 km where an f32 ulp is 2 mm):
    no pair whose f64 walk ends more than 1 mm apart goes to EPA in f32; the
    pairs that take the new exit (status SD_SEPARATED) meet the f64 distance
    to 1e-5 m at the origin and within a few ulps of their coordinates at
    20 km; and the exit is taken (so this test exercises it)."""
    sc = scenes.batched_boxes_hulls(16, 5)
    sc["position"] = (sc["position"].astype(np.float64) + np.array([shift, 0.0, 0.0])).astype(np.float32)
    pairs = oracle.world_pairs(sc, 0.04)
    o32, st32 = oracle.sd_pairs(sc, pairs)
    o64, st64 = oracle.sd_pairs(sc, pairs, None, f64=True)
    false_epa = (st32 == oracle.SD_EPA) & (o64[:, 0] > 1e-3)
    assert not false_epa.any(), f"{int(false_epa.sum())} separated pairs reported as overlapping"
    sep = st32 == oracle.SD_SEPARATED
    assert sep.sum() >= 10
    err = np.abs(o32[sep, 0] - o64[sep, 0])
    assert err.max() <= (1e-5 if shift == 0.0 else 4 * np.spacing(np.float32(shift)))


def test_signed_distance_rounding_exits_on_recorded_pairs(oracle):
    """Pairs recorded from the penalty descent (tests/golden, written by
    make_sd_rounding_exits.py on the GPU) on which the f32 walk takes a
    rounding exit: status 5 (an enclosure after the walk proved a gap) or 6
    (EPA's core depth beyond the walk's own bound, the least support-plane
    distance -delta / |v| of its iterates, by more than 1e-4 m).  The oracle
    takes the same exits and lands bit for bit on the GPU's recorded signed
    distance; the reported closest iterate meets the f64 signed distance to
    2e-4 m, where EPA on the same simplex had reported depths of 0.02-0.18 m
    (caught by the planted mistakes sd_gap_proof_off and sd_epa_bound_off)."""
    from sd_helpers import rounding_exit_scene
    sc, cases = rounding_exit_scene()
    o32, st32 = oracle.sd_pairs(sc, sc["pairs"])
    want = np.array([c["status"] for c in cases])
    assert np.array_equal(st32, want)
    assert (want == oracle.SD_SEPARATED).sum() >= 3 and (want == oracle.SD_EPA_CONTRADICTED).sum() >= 3
    assert np.array_equal(o32[:, 0], np.float32([c["sd_f32_gpu"] for c in cases]))
    o64, _ = oracle.sd_pairs(sc, sc["pairs"], None, f64=True)
    assert np.allclose(o64[:, 0], [c["sd_f64"] for c in cases], rtol=0, atol=1e-9)
    assert np.abs(o32[:, 0] - o64[:, 0]).max() <= 2e-4


def test_touching_cores_report_the_contact_normal_and_its_gradient(oracle):
    """Cores that touch within the walk's touch gate (gjk.rs:226-233) with no
    enclosing tetrahedron: the reference reports Separated there; the signed
    distance reports the core distance 0, sd = -(margins), with a normal and
    witnesses (status SD_TOUCH).  The normal is the walk's last support
    direction with its components along the touching simplex removed.  Held,
    on 131 box-hull pairs bisected to core touch: sd = -0.04 exactly (box
    margin, hull margin 0); the normal within 1e-3 rad of the f64 walk's
    normal with B moved 1 mm apart (the raw search direction was up to 0.84
    rad off: planted mistake sd_touch_raw_direction); pa = pb - n sd; and the
    f64 reference adjoint, which holds n fixed and moves the touching
    simplex's witnesses with their bodies, within 1e-3 of f64 central
    differences on the pairs smooth within +-h (median below 1e-6).  Without
    a normal the gradient was zero and a de-penetration descent kept such a
    pair 0.04 m deep (planted mistake sd_touch_no_normal)."""
    from sd_helpers import fd, pose14, touch_scene
    sc = touch_scene(oracle)
    P = sc["pairs"].shape[0]
    o32, st32 = oracle.sd_pairs(sc, sc["pairs"])
    assert P >= 100 and (st32 == oracle.SD_TOUCH).all()
    assert (o32[:, 0] == np.float32(-0.04)).all()
    n = o32[:, 1:4].astype(np.float64)
    assert np.abs(np.linalg.norm(n, axis=1) - 1.0).max() < 1e-6
    assert np.abs(o32[:, 4:7] - (o32[:, 7:10] - o32[:, 1:4] * o32[:, :1])).max() == 0.0
    apart = pose14(sc)
    apart[:, 7:10] += 1e-3 * sc["direction"]
    o_sep, st_sep = oracle.sd_pairs(sc, sc["pairs"], apart, f64=True)
    assert (st_sep == oracle.SD_WALK).all()
    ang = np.arccos(np.clip(np.sum(n * o_sep[:, 1:4], axis=1), -1.0, 1.0))
    assert ang.max() < 1e-3
    _, st64 = oracle.sd_pairs(sc, sc["pairs"], None, f64=True)
    assert (st64 == oracle.SD_TOUCH).sum() >= 50
    g10 = np.random.default_rng(3).normal(size=(P, 10))
    g10[:, 1:] = 0.0
    ref, fails = oracle.sd_vjp_pairs(sc, sc["pairs"], g10)
    assert fails == 0
    vjp, j1, _, _ = fd(oracle, sc, g10, 1e-4)
    _, j2, _, _ = fd(oracle, sc, g10, 5e-5)
    scale = np.maximum(np.linalg.norm(vjp, axis=1), 1e-2)
    err = np.linalg.norm(ref - vjp, axis=1) / scale
    ratio = j2 / np.maximum(j1, 1e-300)
    smooth = np.all((j1 <= 1e-3 * scale[:, None]) | ((ratio >= 0.3) & (ratio <= 0.7)), axis=1)
    assert smooth.sum() >= 0.85 * P
    assert err[smooth].max() < 1e-3 and np.median(err[smooth]) < 1e-6


def test_exact_penetration_places_overlapping_pairs_at_their_depth(oracle, scenes):
    """scenes.exact_penetration bisects B along the pair's direction until
    the signed distance (here the oracle's f32 one, as the library's is on
    the GPU) meets the target -U[0, 0.2] m.  The f64 signed distance of the
    placed pairs then meets the target to 1e-4 m (EPA's exit tolerance), the
    separated half is untouched, and at most 2 % of the overlapping pairs
    keep a lowered target because their line never reaches the drawn
    depth."""
    sc = scenes.hull_pairs(2000, 4)

    def sd32(pos):
        s2 = dict(sc)
        s2["position"] = pos
        return oracle.sd_pairs(s2, sc["pairs"])[0][:, 0]

    ex = scenes.exact_penetration(sc, sd32)
    pen = sc["target"] < 0
    assert np.array_equal(ex["overlap"], pen)
    assert np.array_equal(ex["position"][0::2], sc["position"][0::2])
    assert np.array_equal(ex["position"][1::2][~pen], sc["position"][1::2][~pen])
    assert ex["unreached"] + ex["retargeted"] <= 0.02 * pen.sum()
    out, st = oracle.sd_pairs(ex, ex["pairs"], None, f64=True)
    assert np.abs(out[:, 0] - ex["target"]).max() < 1e-4
    assert (st[pen & (ex["target"] < 0)] == oracle.SD_EPA).all()
    assert (ex["target"][pen] == sc["target"][pen]).sum() >= 0.98 * pen.sum()
