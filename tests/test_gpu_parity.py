"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Pair sets must be bit-exact (values and order).  Contacts must match in
count exactly; positions and penetrations within TOL_M metres, normals
within TOL_RAD radians (f32 tolerance of the collision path).  The number of
bit-identical contacts is printed as well.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_M = 1e-5
TOL_RAD = 1e-4
MARGIN = 0.04


def _flush(msg):
    print(msg, flush=True)


def _oracle_pass(oracle, scene, margin, pairs_fn="brute"):
    mn, mx = oracle.scene_aabbs(scene, margin)
    pairs = oracle.brute_pairs(mn, mx) if pairs_fn == "brute" else oracle.sweep_pairs(mn, mx)
    cnt, pts, path = oracle.collide_pairs(scene, pairs)
    return pairs, cnt, pts, path


def _dense(cnt, pts):
    """Oracle per-pair fixed slots -> CSR like the library's output."""
    offs = np.concatenate(([0], np.cumsum(cnt.astype(np.int64))))
    dense = np.concatenate([pts[k, :cnt[k]] for k in range(len(cnt))]) if offs[-1] else np.zeros((0, 7), np.float32)
    return offs, dense


def compare_contacts(ref_cnt, ref_pts, offs, cont, path=None, label=""):
    """Assert count equality and the f32 tolerance; return the stats."""
    got_cnt = np.diff(offs.astype(np.int64))
    bad_cnt = np.nonzero(got_cnt != ref_cnt)[0]
    assert bad_cnt.size == 0, f"{label}: contact counts differ at pairs {bad_cnt[:10]}"
    ro, rd = _dense(ref_cnt, ref_pts)
    assert rd.shape == cont.shape
    if rd.shape[0] == 0:
        return dict(n=0, exact=0)
    dp = np.linalg.norm(rd[:, :3].astype(np.float64) - cont[:, :3], axis=1)
    dd = np.abs(rd[:, 6].astype(np.float64) - cont[:, 6])
    a = rd[:, 3:6].astype(np.float64)
    b = cont[:, 3:6].astype(np.float64)
    ang = np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.sum(a * b, axis=1))
    exact = int(np.sum(np.all(rd == cont, axis=1)))
    bad = np.nonzero((dp > TOL_M) | (dd > TOL_M) | (ang > TOL_RAD))[0]
    _flush(f"{label}: {rd.shape[0]} contacts, {exact} bit-identical, max |dp| {dp.max():.3g} m, "
           f"max |dd| {dd.max():.3g} m, max angle {ang.max():.3g} rad, {bad.size} beyond tolerance")
    assert bad.size == 0, f"{label}: {bad.size} contacts beyond tolerance"
    return dict(n=rd.shape[0], exact=exact)


def test_boxes_spheres_plane_pass(bpa, scenes, oracle, ctx):
    """2,048 boxes/spheres over a ground plane (seed 1), margin 0.04: full pass."""
    _flush("case: boxes_spheres_on_plane n=2048 seed=1")
    sc = scenes.boxes_spheres_on_plane(2048, 1)
    ref_pairs, ref_cnt, ref_pts, _ = _oracle_pass(oracle, sc, MARGIN)
    w = bpa.CollisionWorld(sc, ctx)
    P, Cn = w.collide(MARGIN)
    pairs, offs, cont = w.results()
    assert P == ref_pairs.shape[0]
    assert np.array_equal(pairs, ref_pairs), "pair set differs from brute force"
    compare_contacts(ref_cnt, ref_pts, offs, cont, label="boxes_spheres_plane")


@pytest.mark.parametrize("margin", [0.02, 0.04])
def test_margins(bpa, scenes, oracle, ctx, margin):
    _flush(f"case: boxes_spheres_on_plane margin={margin}")
    sc = scenes.boxes_spheres_on_plane(2048, 1)
    ref_pairs, ref_cnt, ref_pts, _ = _oracle_pass(oracle, sc, margin)
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(margin)
    pairs, offs, cont = w.results()
    assert np.array_equal(pairs, ref_pairs)
    compare_contacts(ref_cnt, ref_pts, offs, cont, label=f"margin {margin}")


def test_mixed_convex_pass(bpa, scenes, oracle, ctx):
    """Boxes/spheres/capsules/hulls (40/30/20/10), 20k bodies: GJK + EPA paths."""
    _flush("case: mixed_convex n=20000 seed=2")
    sc = scenes.mixed_convex(20000, 2)
    ref_pairs, ref_cnt, ref_pts, path = _oracle_pass(oracle, sc, MARGIN, "sweep")
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    pairs, offs, cont = w.results()
    assert np.array_equal(pairs, ref_pairs)
    _flush(f"paths: {np.bincount(path, minlength=4)} (gjk-sep, band, epa, closed)")
    compare_contacts(ref_cnt, ref_pts, offs, cont, label="mixed_convex")
    # Dispatch route per pair: the shape-type class, and EPA exactly where the
    # oracle's convex_convex took its EPA arm (path 2).
    routes = w.routes()
    t = sc["shape_type"]
    ti, tj = t[pairs[:, 0]], t[pairs[:, 1]]
    SPH, BOX, HULL, PLANE = 0, 1, 3, 4
    cls = np.full(pairs.shape[0], 5, np.uint8)
    cls[(ti == HULL) | (tj == HULL)] = 6
    cls[(ti == BOX) & (tj == BOX)] = 4
    cls[((ti == SPH) & (tj == BOX)) | ((ti == BOX) & (tj == SPH))] = 2
    cls[(ti == SPH) & (tj == SPH)] = 1
    cls[(ti == PLANE) | (tj == PLANE)] = 3
    cls[(ti == PLANE) & (tj == PLANE)] = 0
    assert np.array_equal(routes & 7, cls), "shape-class route differs"
    assert np.array_equal((routes & 8) != 0, path == 2), "EPA route differs from the oracle's convex_convex"


def test_large_hull_pass_defers_epa_to_the_one_pair_kernel(bpa, scenes, oracle, ctx):
    """Hulls of 65-160 vertices: their EPA pairs leave the two-pair kernel's
    fast path (hulls of at most 64 vertices) and rerun in the one-pair
    kernel, hulls past 128 vertices through its memory-resident support;
    every contact must still match the oracle."""
    _flush("case: mixed_convex n=4000 seed=21, hulls of 65-160 points")
    sc = scenes.mixed_convex(4000, 21, mix=(0.25, 0.15, 0.1, 0.5), hull_pts=(65, 160))
    ref_pairs, ref_cnt, ref_pts, path = _oracle_pass(oracle, sc, MARGIN, "sweep")
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    pairs, offs, cont = w.results()
    assert np.array_equal(pairs, ref_pairs)
    routes = w.routes()
    hc = sc["hull_count"].astype(np.int64)
    big = (hc[pairs[:, 0]] > 64) | (hc[pairs[:, 1]] > 64)
    huge = (hc[pairs[:, 0]] > 128) | (hc[pairs[:, 1]] > 128)
    epa = (routes & 8) != 0
    _flush(f"EPA pairs {int(epa.sum())}, with a hull past 64 vertices {int((epa & big).sum())}, "
           f"past 128 {int((epa & huge).sum())}")
    assert int((epa & big).sum()) > 20 and int((epa & huge).sum()) > 5
    assert np.array_equal(epa, path == 2)
    compare_contacts(ref_cnt, ref_pts, offs, cont, label="large_hulls")


def test_epa_store_overflow_reruns_with_the_big_store(bpa, scenes, oracle, ctx):
    """A polytope that outgrows the one-pair kernel's store goes to the
    2048-face kernel instead of failing the pass.  No natural input is known
    to do that (a convex polytope of <= 128 vertices has <= 252 faces), so the
    one-pair store is cut to 16 faces here: nearly every deferred entry then
    overflows, and the pass must equal the default one bit for bit, and the
    oracle."""
    _flush("case: EPA overflow route, hulls of 65-160 points, one-pair store cut to 16 faces")
    sc = scenes.mixed_convex(4000, 21, mix=(0.25, 0.15, 0.1, 0.5), hull_pts=(65, 160))
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    ref = w.results()
    n_q, n_def, n_over = w.epa_routes()
    assert n_def > 20 and n_over == 0
    try:
        ctx.set_tuning(epa_wave_faces=16)
        w.collide(MARGIN)
        got = w.results()
        routes = w.epa_routes()
    finally:
        ctx.set_tuning(epa_wave_faces=256)
    _flush(f"EPA entries {n_q}, to the one-pair kernel {n_def}; with 16 faces: {routes[2]} through the 2048-face "
           f"store")
    assert routes[:2] == (n_q, n_def) and routes[2] >= 0.9 * n_def
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    ref_pairs, ref_cnt, ref_pts, _ = _oracle_pass(oracle, sc, MARGIN, "sweep")
    compare_contacts(ref_cnt, ref_pts, got[1], got[2], label="overflow route")


def test_collide_batch_matches_oracle(bpa, scenes, oracle, ctx):
    """bch_collide_batch on caller-given pairs (unsorted, with repeats)."""
    _flush("case: collide_batch mixed 5000")
    sc = scenes.mixed_convex(5000, 7)
    mn, mx = oracle.scene_aabbs(sc, MARGIN)
    pairs = oracle.sweep_pairs(mn, mx)
    rng = np.random.default_rng(0)
    sel = pairs[rng.permutation(pairs.shape[0])]
    sel = np.concatenate([sel, sel[:50]])
    cnt, pts = bpa.collide_batch(sc, sel, ctx=ctx)
    rc, rp, _ = oracle.collide_pairs(sc, sel)
    assert np.array_equal(cnt, rc)
    offs = np.concatenate(([0], np.cumsum(cnt.astype(np.int64))))
    _, dense = _dense(cnt, pts)
    compare_contacts(rc, rp, offs, dense, label="collide_batch")


def test_seeded_collide_matches_oracle(bpa, scenes, oracle, ctx):
    """GJK warm-start seeds change the walk, not the contract (mod.rs:515-564)."""
    _flush("case: seeded collide")
    sc = scenes.mixed_convex(3000, 11)
    mn, mx = oracle.scene_aabbs(sc, MARGIN)
    pairs = oracle.sweep_pairs(mn, mx)
    rng = np.random.default_rng(1)
    seeds = rng.normal(size=(pairs.shape[0], 3)).astype(np.float32)
    cnt, pts = bpa.collide_batch(sc, pairs, seeds, ctx=ctx)
    rc, rp, _ = oracle.collide_pairs(sc, pairs, seeds)
    offs = np.concatenate(([0], np.cumsum(cnt.astype(np.int64))))
    _, dense = _dense(cnt, pts)
    compare_contacts(rc, rp, offs, dense, label="seeded")


def _ref_random_aabb(lcg_state):
    """random_aabb of axis_sweep3.rs:855-859 with the reference LCG."""
    A, Cc, M = 6364136223846793005, 1442695040888963407, (1 << 64) - 1

    def unit():
        lcg_state[0] = (lcg_state[0] * A + Cc) & M
        return np.float32(float(lcg_state[0] >> 11)) / np.float32(2.0 ** 53)

    def rng(lo, hi):
        return np.float32(lo) + (np.float32(hi) - np.float32(lo)) * unit()

    base = np.array([rng(-20, 20), rng(-20, 20), rng(-20, 20)], np.float32)
    ext = np.array([rng(0.1, 2.0), rng(0.1, 2.0), rng(0.1, 2.0)], np.float32)
    return base, base + ext


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_proxy_state_machine_matches_brute_force(bpa, oracle, ctx, seed):
    """Randomized create/move/destroy, compared after every few ops with the
    brute-force oracle (the reference's state-machine test pattern,
    axis_sweep3.rs:937-957)."""
    _flush(f"case: proxy state machine seed={seed}")
    bp = bpa.BroadPhase(ctx=bpa._Context(0))
    live = {}
    boxes = {}
    st = [seed]
    A, Cc, M = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    for op in range(300):
        st[0] = (st[0] * A + Cc) & M
        r = st[0] % 4
        if r <= 1 or not live:
            mn, mx = _ref_random_aabb(st)
            i = bp.create_proxy(mn, mx)
            free = [k for k in range(len(boxes) + 1) if k not in live]
            assert i == min(free), "create_proxy must reuse the lowest free slot"
            live[i] = True
            boxes[i] = (mn, mx)
        elif r == 2:
            k = sorted(live)[int(st[0] >> 33) % len(live)]
            mn, mx = _ref_random_aabb(st)
            bp.set_aabb(k, mn, mx)
            boxes[k] = (mn, mx)
        else:
            k = sorted(live)[int(st[0] >> 33) % len(live)]
            bp.destroy_proxy(k)
            del live[k]
        if op % 25 == 24:
            n = max(boxes) + 1 if boxes else 0
            mn = np.full((n, 3), np.nan, np.float32)
            mx = np.full((n, 3), np.nan, np.float32)
            for k in live:
                mn[k], mx[k] = boxes[k]
            ref = oracle.brute_pairs(mn, mx)
            got = bp.compute_pairs()
            assert np.array_equal(got, ref), f"op {op}"
    bp.ctx.close()


def test_broadphase_edge_cases(bpa, oracle, ctx):
    """Touching boxes (closed intervals), identical boxes, degenerate points,
    a huge box, empty and single-proxy sets."""
    _flush("case: broadphase edge cases")
    bp = bpa.BroadPhase(ctx=bpa._Context(0))
    assert bp.compute_pairs().shape == (0, 2)
    bp.create_proxy([0, 0, 0], [1, 1, 1])
    assert bp.compute_pairs().shape == (0, 2)
    bp.create_proxy([1, 0, 0], [2, 1, 1])          # touching on x
    bp.create_proxy([0, 0, 0], [1, 1, 1])          # identical
    bp.create_proxy([0.5, 0.5, 0.5], [0.5, 0.5, 0.5])  # point
    bp.create_proxy([-1e30, -1e30, -1e30], [1e30, 1e30, 1e30])  # plane-like
    bp.create_proxy([0.5, 0.5, 5.0], [1.5, 1.5, 6.0])  # separated on z
    mn = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 0], [.5, .5, .5], [-1e30] * 3, [.5, .5, 5]], np.float32)
    mx = np.array([[1, 1, 1], [2, 1, 1], [1, 1, 1], [.5, .5, .5], [1e30] * 3, [1.5, 1.5, 6]], np.float32)
    assert np.array_equal(bp.compute_pairs(), oracle.brute_pairs(mn, mx))
    assert bp.proxy_aabb(3) is not None
    bp.destroy_proxy(3)
    assert bp.proxy_aabb(3) is None
    bp.set_aabb(3, [0, 0, 0], [9, 9, 9])  # dead slot: no-op
    assert bp.proxy_aabb(3) is None
    mn[3] = np.nan
    mx[3] = np.nan
    assert np.array_equal(bp.compute_pairs(), oracle.brute_pairs(mn, mx))
    cand = bp.aabb_candidates([0.9, 0.9, 0.9], [1.1, 1.1, 1.1])
    assert list(cand) == [0, 1, 2, 4]
    bp.ctx.close()


def test_nan_and_degenerate_poses(bpa, scenes, oracle, ctx):
    """NaN poses never pair and never hang; coincident bodies stay finite."""
    _flush("case: nan / degenerate poses")
    sc = scenes.mixed_convex(2000, 5)
    sc["position"][10] = np.nan
    sc["orientation"][20] = np.nan
    sc["position"][30] = sc["position"][31]
    sc["orientation"][30] = sc["orientation"][31]
    ref_pairs, ref_cnt, ref_pts, _ = _oracle_pass(oracle, sc, MARGIN, "sweep")
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    pairs, offs, cont = w.results()
    assert np.array_equal(pairs, ref_pairs)
    assert not np.any(pairs == 10)
    compare_contacts(ref_cnt, ref_pts, offs, cont, label="degenerate")


def test_empty_and_tiny_worlds(bpa, scenes, ctx):
    _flush("case: empty / tiny worlds")
    sc = scenes.boxes_spheres_on_plane(0, 1, with_plane=False)
    w = bpa.CollisionWorld(sc, ctx)
    assert w.collide(MARGIN) == (0, 0)
    sc = scenes.boxes_spheres_on_plane(1, 1)
    w = bpa.CollisionWorld(sc, ctx)
    P, _ = w.collide(MARGIN)
    assert P in (0, 1)


def _box_columns(cols=16, height=4, sink=0.01):
    """cols x cols columns of `height` unit cubes resting face to face on the
    plane y = 0, each sunk `sink` m into the one below: three face-face box
    pairs of four points per column, more contacts than pairs."""
    n = cols * cols * height + 1
    sc = {"position": np.zeros((n, 3), np.float32),
          "orientation": np.tile(np.float32([0, 0, 0, 1]), (n, 1)),
          "shape_type": np.zeros(n, np.uint8), "shape_params": np.zeros((n, 4), np.float32),
          "hull_first": np.zeros(n, np.uint32), "hull_count": np.zeros(n, np.uint32),
          "hull_vertices": np.zeros((0, 3), np.float32)}
    sc["shape_type"][0] = 4  # plane y = 0, normal +Y
    sc["shape_params"][0] = (0.0, 1.0, 0.0, 0.0)
    k = 1
    for x in range(cols):
        for z in range(cols):
            for h in range(height):
                sc["shape_type"][k] = 1  # box
                sc["shape_params"][k, :3] = 0.5
                sc["position"][k] = (1.5 * x, 0.5 - sink + h * (1.0 - sink), 1.5 * z)
                k += 1
    return sc


def test_contact_buffer_regrows_when_contacts_outnumber_pairs(bpa, oracle):
    """The eager pass gathers contacts into a buffer sized by earlier passes
    (at first one contact per pair, with headroom) and gathers again once it
    is short.  Face-resting boxes give about 1.9 contacts per pair, so on a
    fresh context the first pass takes the regrow path and the second the
    sized buffer; both must equal the oracle."""
    _flush("case: resting box columns, contacts > pairs, fresh context")
    sc = _box_columns()
    ref_pairs, ref_cnt, ref_pts, _ = _oracle_pass(oracle, sc, MARGIN)
    assert ref_cnt.sum() > 1.5 * ref_pairs.shape[0]
    fresh = bpa._Context(0)
    try:
        w = bpa.CollisionWorld(sc, fresh)
        for rep in range(2):
            P, Cn = w.collide(MARGIN)
            pairs, offs, cont = w.results()
            assert np.array_equal(pairs, ref_pairs)
            assert Cn == int(ref_cnt.sum())
            compare_contacts(ref_cnt, ref_pts, offs, cont, label=f"box columns, pass {rep}")
    finally:
        fresh.close()


@pytest.mark.parametrize("worlds", ["64 of 16-512", "8 of 16-4096"])
def test_batched_worlds_equal_plain_per_world(bpa, scenes, oracle, ctx, worlds):
    """Independent worlds in one pass == each world collided alone (plain
    path).  Worlds of at most 2048 bodies have their Morton keys sorted one
    block per world; a batch with a larger world sorts them all at once."""
    _flush(f"case: batched {worlds} bodies")
    sc = scenes.batched_worlds(64, 3, 16, 512) if worlds.startswith("64") else scenes.batched_worlds(8, 7, 16, 4096)
    if worlds.startswith("8"):
        assert np.diff(sc["world_offsets"]).max() > 2048
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    pairs, offs, cont = w.results()
    wo = sc["world_offsets"]
    world_of = np.searchsorted(wo, pairs[:, 0], side="right") - 1
    assert np.all(pairs[:, 1] < wo[world_of + 1]), "pair crosses worlds"
    plain = bpa.CollisionWorld(scenes.world_slice(sc, 0), bpa._Context(0))
    for wi in range(len(wo) - 1):
        sub = scenes.world_slice(sc, wi)
        pw = bpa.CollisionWorld(sub, plain.ctx)
        pw.collide(MARGIN)
        pp, po, pc = pw.results()
        m = world_of == wi
        got = pairs[m] - wo[wi]
        assert np.array_equal(got, pp), f"world {wi} pairs differ"
        k = np.nonzero(m)[0]
        seg = cont[offs[k[0]]:offs[k[-1] + 1]] if k.size else np.zeros((0, 7), np.float32)
        assert np.array_equal(seg, pc), f"world {wi} contacts differ"
        if wi < 4:
            rp = oracle.brute_pairs(*oracle.scene_aabbs(sub, MARGIN))
            assert np.array_equal(pp, rp)
    plain.ctx.close()


def test_large_scene_properties(bpa, scenes, oracle, ctx):
    """200k mixed bodies: sorted unique pairs; every pair overlaps; sampled
    bodies' partner sets equal brute force over all bodies."""
    _flush("case: mixed_convex n=200000 properties")
    sc = scenes.mixed_convex(200_000, 2)
    w = bpa.CollisionWorld(sc, ctx)
    P, Cn = w.collide(MARGIN)
    pairs, offs, cont = w.results()
    key = pairs[:, 0].astype(np.uint64) << np.uint64(32) | pairs[:, 1].astype(np.uint64)
    assert np.all(pairs[:, 0] < pairs[:, 1])
    assert np.all(np.diff(key.astype(np.int64)) > 0)
    mn, mx = oracle.scene_aabbs(sc, MARGIN)
    a, b = pairs[:, 0], pairs[:, 1]
    assert np.all((mn[a] <= mx[b]) & (mn[b] <= mx[a]))
    rng = np.random.default_rng(3)
    for i in rng.choice(sc["shape_type"].shape[0], 64, replace=False):
        ov = np.nonzero(np.all((mn[i] <= mx) & (mn <= mx[i]), axis=1))[0]
        ov = ov[ov != i]
        got = np.concatenate([pairs[pairs[:, 0] == i, 1], pairs[pairs[:, 1] == i, 0]])
        assert np.array_equal(np.sort(got), ov), f"body {i}"
    assert np.all(np.isfinite(cont))
    assert offs[-1] == Cn


@pytest.mark.gpu
def test_raycast_candidates_match_oracle(bpa, oracle, ctx):
    """raycast_candidates through the C ABI: the reference KAT
    (axis_sweep3.rs:734-748) and random segments against the oracle."""
    _flush("case: raycast candidates")
    bp = bpa.BroadPhase(ctx=bpa._Context(0))
    hit = bp.create_proxy([1, 1, 1], [2, 2, 2])
    miss = bp.create_proxy([10, 10, 10], [11, 11, 11])
    assert bp.raycast_candidates([0, 0, 0], [1.5, 1.5, 1.5]).tolist() == [hit]
    inside = bp.raycast_candidates([1.5, 1.5, 1.5], [4, 4, 4]).tolist()
    assert hit in inside and miss not in inside
    assert bp.raycast_candidates([3, 3, 3], [4, 4, 4]).tolist() == []
    rng = np.random.default_rng(11)
    bp2 = bpa.BroadPhase(ctx=bpa._Context(0))
    lo = rng.uniform(-20, 20, (500, 3)).astype(np.float32)
    hi = (lo + rng.uniform(0.1, 3.0, (500, 3))).astype(np.float32)
    for i in range(500):
        bp2.create_proxy(lo[i], hi[i])
    for k in range(50):
        f = rng.uniform(-25, 25, 3).astype(np.float32)
        t = rng.uniform(-25, 25, 3).astype(np.float32)
        if k % 7 == 3:
            t[1] = f[1]
        assert np.array_equal(bp2.raycast_candidates(f, t), oracle.raycast_candidates(lo, hi, f, t))


def test_autotune_keeps_results(bpa, scenes, ctx):
    """The autotuner changes launch shapes only: the pass after tuning gives
    the same pairs and bit-identical contacts."""
    _flush("case: autotune")
    sc = scenes.mixed_convex(30000, 12)
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    before = w.results()
    tuned = w.autotune(MARGIN, trials=2)
    _flush(f"autotune: {tuned}")
    w.collide(MARGIN)
    after = w.results()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert tuned["hull_lanes"] in (8, 16) and tuned["pass_ms"] > 0


def test_predicted_poses_and_sleeping_aabbs(bpa, scenes, oracle, ctx):
    """The step's AABB update (step.rs:340-363): active bodies at position +
    velocity * dt, inactive bodies keep the previous pass's AABB even though
    their pose moved; pairs bit-exact and contacts against the oracle at the
    predicted poses."""
    _flush("case: predicted poses, 20k mixed, 10% inactive")
    sc = scenes.mixed_convex(20000, 2)
    n = sc["shape_type"].shape[0]
    rng = np.random.default_rng(21)
    w = bpa.CollisionWorld(sc, bpa._Context(0))
    w.collide(MARGIN)  # previous pass: AABBs at the scene poses
    prev = oracle.scene_aabbs(sc, MARGIN)
    flags = np.full(n, 2, np.uint8)
    flags[rng.random(n) < 0.1] = 0
    w.set_body_flags(flags)
    pos2 = (sc["position"] + rng.normal(0, 0.02, (n, 3))).astype(np.float32)
    vel = rng.normal(0, 1.0, (n, 3)).astype(np.float32)
    dt = np.float32(1.0 / 60.0)
    moved = dict(sc)
    moved["position"] = pos2
    ppos, mn, mx = oracle.scene_aabbs_predicted(moved, vel, dt, MARGIN, flags=flags, prev=prev)
    assert np.array_equal(mn[flags == 0], prev[0][flags == 0])
    w.set_poses(pos2, sc["orientation"], vel, float(dt))
    P, Cn = w.collide(MARGIN)
    pairs, offs, cont = w.results()
    ref_pairs = oracle.sweep_pairs(mn, mx)
    assert np.array_equal(pairs, ref_pairs), "pairs at predicted poses differ"
    pred = dict(sc)
    pred["position"] = ppos
    cnt, pts, _ = oracle.collide_pairs(pred, ref_pairs)
    compare_contacts(cnt, pts, offs, cont, label="predicted poses")
    # the same pair set differs from one without retention (the test bites)
    _, mn2, mx2 = oracle.scene_aabbs_predicted(moved, vel, dt, MARGIN, flags=flags, prev=None)
    assert not np.array_equal(oracle.sweep_pairs(mn2, mx2), ref_pairs)
    w.ctx.close()


def _facing_quat(v, u):
    """Unit quaternion (x, y, z, w) turning direction v onto direction u."""
    v = v / np.linalg.norm(v)
    u = u / np.linalg.norm(u)
    c = np.cross(v, u)
    q = np.array([c[0], c[1], c[2], 1.0 + float(v @ u)])
    return q / np.linalg.norm(q)


def test_walk_pairs_at_the_bounding_ball_edge(bpa, oracle, ctx):
    """Walk-class pairs (hull, box and capsule combinations) turned so that
    each body's farthest point faces the other, placed at gaps of -0.5 to 3
    times the pair's contact_breaking_threshold and at the split's
    bounding-ball cut itself (the radii plus both thresholds, +-0.2 %): the
    pairs the split skips as farther apart than the band, and the ones it
    keeps, give the oracle's pair set and contacts (a wrongly skipped pair
    would lose its band contact)."""
    _flush("case: walk-class pairs at the bounding-ball edge")
    rng = np.random.default_rng(11)
    kinds = [(3, 3), (1, 3), (3, 1), (2, 3), (3, 2), (1, 2), (2, 2)]
    fracs = [-0.5, 0.25, 0.5, 0.9, 0.999, 1.001, 1.5, 3.0]
    bodies = []
    for k in range(224):
        ta, tb = kinds[k % len(kinds)]
        pair = []
        for t in (ta, tb):
            if t == 3:
                d = rng.normal(size=(24, 3))
                pts = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.3, 1.2, 3)).astype(np.float32)
                far = pts[np.argmax(np.linalg.norm(pts.astype(np.float64), axis=1))].astype(np.float64)
                pair.append(dict(t=3, pts=pts, far=far, R=float(np.linalg.norm(far)), s=oracle.hull(pts)))
            elif t == 1:
                h = rng.uniform(0.25, 1.0, 3).astype(np.float32)
                pair.append(dict(t=1, h=h, far=h.astype(np.float64), R=float(np.linalg.norm(h)), s=oracle.box(*h)))
            else:
                r, hh = np.float32(rng.uniform(0.1, 0.5)), np.float32(rng.uniform(0.2, 1.0))
                pair.append(dict(t=2, r=r, hh=hh, far=np.array([0.0, 1.0, 0.0]), R=float(hh + r),
                                 s=oracle.capsule(r, hh)))
        A, B = pair
        ta_thr, tb_thr = oracle.threshold(A["s"]), oracle.threshold(B["s"])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pa = 6.0 * np.array([k % 8, (k // 8) % 8, k // 64]) - 21.0  # near the origin: a small rounding allowance
        if k % 4 == 3:  # at the cut: radii plus both thresholds, +-0.2 %
            d = (A["R"] + B["R"] + ta_thr + tb_thr) * (1.0 + rng.uniform(-2e-3, 2e-3))
        else:
            d = A["R"] + B["R"] + fracs[(k // 4) % len(fracs)] * min(ta_thr, tb_thr)
        A["p"], A["q"] = pa, _facing_quat(A["far"], u)
        B["p"], B["q"] = pa + d * u, _facing_quat(B["far"], -u)
        bodies += [A, B]
    n = len(bodies)
    sc = dict(position=np.zeros((n, 3), np.float32), orientation=np.zeros((n, 4), np.float32),
              shape_type=np.zeros(n, np.uint8), shape_params=np.zeros((n, 4), np.float32),
              hull_first=np.zeros(n, np.uint32), hull_count=np.zeros(n, np.uint32))
    verts = []
    for i, b in enumerate(bodies):
        sc["position"][i] = b["p"]
        sc["orientation"][i] = b["q"]
        sc["shape_type"][i] = b["t"]
        if b["t"] == 3:
            sc["hull_first"][i] = 24 * len(verts)
            sc["hull_count"][i] = 24
            verts.append(b["pts"])
        elif b["t"] == 1:
            sc["shape_params"][i, :3] = b["h"]
        else:
            sc["shape_params"][i, :2] = (b["r"], b["hh"])
    sc["hull_vertices"] = np.concatenate(verts).astype(np.float32)
    ref_pairs, ref_cnt, ref_pts, _ = _oracle_pass(oracle, sc, MARGIN)
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    pairs, offs, cont = w.results()
    assert np.array_equal(pairs, ref_pairs)
    assert ref_pairs.shape[0] >= 150 and (ref_cnt > 0).sum() >= 60 and (ref_cnt == 0).sum() >= 40
    compare_contacts(ref_cnt, ref_pts, offs, cont, label="bounding-ball edge")


def test_sort_tile_keeps_results(bpa, scenes, ctx):
    """Every sort of the pass (Morton codes, pair keys, shape classes) with
    2048- and 4096-item tiles (the sort_tile knob): identical pairs, offsets
    and contacts, on one world and on a batch of worlds."""
    _flush("case: sort tile 2048 vs 4096")
    keep = ctx.tuning()["sort_tile"]
    try:
        for name, sc in (("mixed", scenes.mixed_convex(30000, 13)), ("batched", scenes.batched_worlds(48, 5))):
            w = bpa.CollisionWorld(sc, ctx)
            out = []
            for tile in (4096, 2048):
                ctx.set_tuning(sort_tile=tile)
                w.collide(MARGIN)
                out.append(w.results())
            _flush(f"  {name}: {out[0][0].shape[0]} pairs, {out[0][2].shape[0]} contacts")
            for a, b in zip(*out):
                assert np.array_equal(a, b)
    finally:
        ctx.set_tuning(sort_tile=keep)
