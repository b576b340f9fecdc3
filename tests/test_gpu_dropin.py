"""The drop-in entry points on the GPU: the device-resident proxy arena
(bch_bp_*), caller-given pair batches with pose-only uploads
(bch_collide_batch_poses), and the separation of the batch state from the
resident world on one context.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MARGIN = 0.04


def _flush(msg):
    print(msg, flush=True)


def _boxes(rng, n, side):
    lo = rng.uniform(0.0, side, (n, 3)).astype(np.float32)
    return lo, (lo + rng.uniform(0.1, 2.0, (n, 3))).astype(np.float32)


def _mirror_pairs(oracle, mn, mx, live):
    a = np.where(live[:, None], mn, np.float32(np.nan)).astype(np.float32)
    b = np.where(live[:, None], mx, np.float32(np.nan)).astype(np.float32)
    return oracle.sweep_pairs(a, b)


def test_proxy_arena_updates_match_brute_force(bpa, oracle):
    """20,000 proxies: scattered updates of 5 % (the dirty-slot scatter), a
    move of 60 % (the whole-arena upload), destroys and creates (lowest free
    slot first, then growth), and an unchanged call (the kept result), each
    compared with the sweep oracle over the same boxes."""
    _flush("case: proxy arena, 20000 proxies")
    rng = np.random.default_rng(3)
    n, side = 20000, 60.0
    bp = bpa.BroadPhase(ctx=bpa._Context(0))
    mn, mx = _boxes(rng, n, side)
    for i in range(n):
        assert bp.create_proxy(mn[i], mx[i]) == i
    live = np.ones(n, bool)
    assert np.array_equal(bp.compute_pairs(), _mirror_pairs(oracle, mn, mx, live))
    for frac, label in ((0.05, "scatter"), (0.6, "whole"), (0.01, "scatter")):
        k = rng.choice(n, int(frac * n), replace=False).astype(np.uint32)
        nmn, nmx = _boxes(rng, k.shape[0], side)
        bp.set_aabbs(k, nmn, nmx)
        mn[k], mx[k] = nmn, nmx
        got = bp.compute_pairs()
        assert np.array_equal(got, _mirror_pairs(oracle, mn, mx, live)), label
        _flush(f"  {label}: {k.shape[0]} moved, {got.shape[0]} pairs")
    dead = rng.choice(n, 600, replace=False)
    for i in dead:
        bp.destroy_proxy(int(i))
    live[dead] = False
    assert np.array_equal(bp.compute_pairs(), _mirror_pairs(oracle, mn, mx, live))
    # a dead slot ignores set_aabb; creates fill the lowest free slots first
    bp.set_aabb(int(dead[0]), [0, 0, 0], [side, side, side])
    free = sorted(int(i) for i in dead)
    amn, amx = _boxes(rng, 800, side)
    mn = np.vstack([mn, np.zeros((200, 3), np.float32)])
    mx = np.vstack([mx, np.zeros((200, 3), np.float32)])
    live = np.concatenate([live, np.zeros(200, bool)])
    for t in range(800):
        slot = bp.create_proxy(amn[t], amx[t])
        assert slot == (free[t] if t < len(free) else n + t - len(free))
        mn[slot], mx[slot], live[slot] = amn[t], amx[t], True
    ref = _mirror_pairs(oracle, mn, mx, live)
    assert np.array_equal(bp.compute_pairs(), ref)
    assert np.array_equal(bp.compute_pairs(), ref)  # unchanged: the kept result
    bp.ctx.close()


def test_compute_pairs_overflow_protocol(bpa, oracle):
    """OVERFLOW with the needed count, then the re-call with that capacity."""
    import ctypes as C
    _flush("case: compute_pairs overflow")
    rng = np.random.default_rng(4)
    bp = bpa.BroadPhase(ctx=bpa._Context(0))
    mn, mx = _boxes(rng, 3000, 20.0)
    for i in range(3000):
        bp.create_proxy(mn[i], mx[i])
    ref = oracle.sweep_pairs(mn, mx)
    out = np.zeros((16, 2), np.uint32)
    cnt = C.c_uint64()
    rc = bp.L.bch_bp_compute_pairs(bp.ctx.h, out.ctypes.data, 16, C.byref(cnt))
    assert rc == bpa._lib.BCH_OVERFLOW and cnt.value == ref.shape[0] > 16
    out = np.zeros((cnt.value, 2), np.uint32)
    rc = bp.L.bch_bp_compute_pairs(bp.ctx.h, out.ctypes.data, cnt.value, C.byref(cnt))
    assert rc == bpa._lib.BCH_OK and np.array_equal(out, ref)
    bp.ctx.close()


def test_collide_batch_poses_matches_a_full_upload(bpa, scenes, oracle):
    """Pose-only calls on the shapes of the last collide_batch equal
    collide_batch with every array uploaded, bit for bit, and the oracle."""
    _flush("case: collide_batch_poses, mixed 4000")
    ctx = bpa._Context(0)
    sc = scenes.mixed_convex(4000, 21)
    pairs = oracle.sweep_pairs(*oracle.scene_aabbs(sc, MARGIN))
    bpa.collide_batch(sc, pairs, ctx=ctx)
    rng = np.random.default_rng(2)
    for step in range(3):
        moved = dict(sc)
        moved["position"] = (sc["position"] + rng.normal(0, 0.05, sc["position"].shape)).astype(np.float32)
        q = sc["orientation"] + rng.normal(0, 0.02, sc["orientation"].shape)
        moved["orientation"] = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
        cp, pp = bpa.collide_batch_poses(moved["position"], moved["orientation"], pairs, ctx=ctx)
        full_ctx = bpa._Context(0)
        cf, pf = bpa.collide_batch(moved, pairs, ctx=full_ctx)
        full_ctx.close()
        assert np.array_equal(cp, cf) and np.array_equal(pp, pf), f"step {step}"
        rc, rp, _ = oracle.collide_pairs(moved, pairs)
        assert np.array_equal(cp, rc)
        for k in np.nonzero(cp)[0][:2000]:
            assert np.array_equal(pp[k, :cp[k]], rp[k, :cp[k]]), f"step {step} pair {k}"
    ctx.close()


def test_collide_batch_poses_needs_shapes(bpa):
    ctx = bpa._Context(0)
    with pytest.raises(bpa.BchError):
        bpa.collide_batch_poses(np.zeros((1, 3)), np.array([[0, 0, 0, 1.0]]), np.zeros((0, 2)), ctx=ctx)
    ctx.close()


def test_batch_calls_leave_the_world_untouched(bpa, scenes, oracle):
    """One context serves a batched world with planes and every batch entry
    point: collide_batch, a pose-only batch, a sphere sweep batch, a mesh
    batch, a mesh sweep and the proxy broadphase run between world passes,
    and the world pass, its captured replay and its routes stay bit-identical."""
    _flush("case: batch calls between world passes, 32 batched worlds")
    ctx = bpa._Context(0)
    sc = scenes.batched_worlds(32, 3, 16, 512)
    w = bpa.CollisionWorld(sc, ctx)
    w.collide(MARGIN)
    ref = w.results()
    ref_routes = w.routes()
    w.capture(MARGIN)
    w.replay()
    for a, b in zip(w.results(), ref):
        assert np.array_equal(a, b)

    other = scenes.mixed_convex(3000, 9)
    opairs = oracle.sweep_pairs(*oracle.scene_aabbs(other, MARGIN))
    bpa.collide_batch(other, opairs, ctx=ctx)
    bpa.collide_batch_poses(other["position"], other["orientation"], opairs, ctx=ctx)
    rng = np.random.default_rng(1)
    tgt = rng.integers(0, 3000, 500).astype(np.uint32)
    frm = (other["position"][tgt] + 2.0).astype(np.float32)
    bpa.sweep_batch(other, tgt, frm, other["position"][tgt], radius=np.full(500, 0.2, np.float32), ctx=ctx)
    v = np.array([[x, 0.0, z] for z in (-30, 30) for x in (-30, 30)], np.float32)
    mesh = bpa.TriangleMesh(v, np.array([[0, 2, 1], [1, 2, 3]], np.uint32), ctx=ctx)
    mp = np.tile(np.array([0, 0, 0, 0, 0, 0, 1], np.float32), (100, 1))
    bpa.collide_mesh_batch(other, np.stack([np.arange(100), np.full(100, mesh.id)], 1), mp, ctx=ctx)
    bpa.sweep_mesh_batch(np.full(50, mesh.id, np.uint32), mp[:50], frm[:50], other["position"][tgt[:50]],
                         radius=np.full(50, 0.2, np.float32), ctx=ctx)
    bp = bpa.BroadPhase(ctx=ctx)
    for i in range(200):
        bp.create_proxy(*_boxes(rng, 1, 10.0))
    bp.compute_pairs()

    w.replay()
    for a, b in zip(w.results(), ref):
        assert np.array_equal(a, b), "replay after batch calls"
    w.collide(MARGIN)
    for a, b in zip(w.results(), ref):
        assert np.array_equal(a, b), "eager pass after batch calls"
    assert np.array_equal(w.routes(), ref_routes)
    # and the world's pairs still never cross worlds
    pairs = ref[0]
    wo = sc["world_offsets"]
    world_of = np.searchsorted(wo, pairs[:, 0], side="right") - 1
    assert np.all(pairs[:, 1] < wo[world_of + 1])
    ctx.close()
