"""GPU parity of the pair pass with persistent manifolds
(bch_world_manifold_pass) against the oracle (oracle.hpp::pm_pair_pass,
restating step.rs:660-854 over manifold.rs:124-386), frame by frame:
routes, emitted points and every cache entry bit for bit, with the
solver's impulse writeback between frames.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MARGIN = 0.04
SLOP = 0.005
WARM = 0.85


def _flush(msg):
    print(msg, flush=True)


def _frames(sc, seed, n_frames):
    """Pose sequence: a third of the bodies drift a few millimetres per frame
    (some past the 5 mm redetect bound), a tenth turn by ~0.01 rad, the rest
    stay; planes never move."""
    rng = np.random.default_rng(seed)
    n = sc["shape_type"].shape[0]
    pos = sc["position"].astype(np.float32).copy()
    quat = sc["orientation"].astype(np.float32).copy()
    movable = sc["shape_type"] != 4
    out = [(pos.copy(), quat.copy())]
    for _ in range(n_frames - 1):
        drift = (rng.random(n) < 0.33) & movable
        pos[drift] += rng.normal(0.0, 0.004, (int(drift.sum()), 3)).astype(np.float32)
        turn = (rng.random(n) < 0.1) & movable
        k = int(turn.sum())
        axis = rng.normal(size=(k, 3))
        axis /= np.linalg.norm(axis, axis=1, keepdims=True)
        ang = rng.normal(0.0, 0.01, k)
        dq = np.concatenate([axis * np.sin(ang / 2)[:, None], np.cos(ang / 2)[:, None]], 1)
        q = quat[turn].astype(np.float64)
        x1, y1, z1, w1 = dq.T
        x2, y2, z2, w2 = q.T
        nq = np.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                       w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
        quat[turn] = (nq / np.linalg.norm(nq, axis=1, keepdims=True)).astype(np.float32)
        out.append((pos.copy(), quat.copy()))
    return out


def _dense_to_slots(offs, pts, P):
    slots = np.zeros((P, 4, 12), np.float32)
    cnt = np.diff(offs.astype(np.int64))
    for p in np.nonzero(cnt)[0]:
        slots[p, :cnt[p]] = pts[offs[p]:offs[p + 1]]
    return cnt, slots


def _run(bpa, oracle, ctx, sc, label, n_frames=5, seed=7, flags=None, gm=None):
    w = bpa.CollisionWorld(sc, ctx)
    w.set_body_flags(flags, gm)
    cache = None
    rng = np.random.default_rng(seed + 100)
    tot_routes = np.zeros(6, np.int64)
    prev = None  # with flags, inactive bodies keep the previous frame's AABB (step.rs:348-353)
    for f, (pos, quat) in enumerate(_frames(sc, seed, n_frames)):
        _flush(f"case: {label} frame {f}")
        w.set_poses(pos, quat)
        P, Np = w.manifold_pass(MARGIN, SLOP, WARM)
        pairs, route, offs, pts = w.manifold_results()
        cnt, slots = _dense_to_slots(offs, pts, P)
        valid, ent = w.manifold_cache()
        osc = dict(sc)
        osc["position"], osc["orientation"] = pos, quat
        if flags is None:
            mn, mx = oracle.scene_aabbs(osc, MARGIN)
        else:
            _, mn, mx = oracle.scene_aabbs_predicted(osc, np.zeros_like(pos), 0.0, MARGIN, flags=flags, prev=prev)
            prev = (mn, mx)
        ref_pairs = oracle.sweep_pairs(mn, mx)
        assert np.array_equal(pairs, ref_pairs), f"{label} frame {f}: pair sets differ"
        ref = oracle.pm_pass(osc, ref_pairs, cache, flags=flags, group_mask=gm, slop=SLOP, warm=WARM)
        bad = np.nonzero(route != ref["route"])[0]
        assert bad.size == 0, f"{label} frame {f}: routes differ at {bad[:8]} ({route[bad[:8]]} vs {ref['route'][bad[:8]]})"
        assert np.array_equal(cnt, ref["count"].astype(np.int64)), f"{label} frame {f}: point counts differ"
        same_pts = np.all(slots.view(np.uint32) == ref["points"].view(np.uint32), axis=(1, 2))
        assert same_pts.all(), f"{label} frame {f}: points differ at pairs {np.nonzero(~same_pts)[0][:8]}"
        assert np.array_equal(valid, ref["cache"][2]), f"{label} frame {f}: cache validity differs"
        v = valid.astype(bool)
        same_ent = np.all(ent[v].view(np.uint32) == ref["cache"][1][v].view(np.uint32), axis=1)
        assert same_ent.all(), f"{label} frame {f}: cache entries differ at {np.nonzero(v)[0][~same_ent][:8]}"
        rc = np.bincount(route, minlength=6)
        tot_routes += rc
        _flush(f"{label} frame {f}: {P} pairs, {Np} points; routes {dict(zip(bpa.CollisionWorld.ROUTES, rc))}")
        # the solver's impulses written back on both sides
        imp = rng.uniform(0.0, 2.0, Np).astype(np.float32)
        w.manifold_impulses(imp)
        keys, rent, rval = ref["cache"]
        rent = rent.copy()
        for p in np.nonzero(cnt)[0]:
            for k in range(cnt[p]):
                rent[p, 12 * k + 7] = imp[offs[p] + k]
        cache = (keys, rent, rval)
        valid2, ent2 = w.manifold_cache()
        assert np.array_equal(ent2[v].view(np.uint32), rent[v].view(np.uint32)), "impulse writeback differs"
    return tot_routes


def test_manifold_pass_matches_oracle_mixed(bpa, scenes, oracle, ctx):
    """20,000 mixed convex bodies (GJK, EPA and closed forms), five frames."""
    sc = scenes.mixed_convex(20000, 2)
    tot = _run(bpa, oracle, ctx, sc, "mixed 20k")
    assert tot[3] > 0 and tot[4] > 0  # both the refresh and the detection paths ran


def test_manifold_pass_matches_oracle_with_flags(bpa, scenes, oracle, ctx):
    """Boxes and spheres on a plane with static, sleeping and fast bodies
    and collision groups, five frames."""
    sc = scenes.boxes_spheres_on_plane(4096, 5, lo=(-12.0, 0.0, -12.0), hi=(12.0, 6.0, 12.0))
    n = sc["shape_type"].shape[0]
    rng = np.random.default_rng(3)
    flags = np.full(n, 2, np.uint8)                      # active
    flags[rng.random(n) < 0.05] = 1                      # static, inactive
    flags[rng.random(n) < 0.05] = 3                      # kinematic (static, active)
    flags[rng.random(n) < 0.05] = 0                      # sleeping
    flags[rng.random(n) < 0.15] |= 4                     # fast
    flags[0] = 3                                         # the plane is static
    gm = np.full(n, 1 | (0xffff << 16), np.uint32)       # group 1, collides with all
    grp = rng.random(n) < 0.05
    gm[grp] = 2 | (2 << 16)                              # group 2, collides with group 2 only
    tot = _run(bpa, oracle, ctx, sc, "plane 4k flags", flags=flags, gm=gm)
    assert all(tot[r] > 0 for r in range(6)), f"every route taken: {tot}"
