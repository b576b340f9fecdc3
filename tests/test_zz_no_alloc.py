"""bch_collide_batch_static, the entry point that allocates nothing, against
bch_collide_batch on the same inputs: counts equal and every contact point
bit-identical (the same kernels run on both), on two calls in a row with
the second one smaller, so nothing left in the shared static array from the
first call can leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MARGIN = 0.04


def _flush(msg):
    print(msg, flush=True)


def _device_bodies(bpa, torch, sc):
    """BchBodies whose arrays are device copies of the scene's (kept alive in
    the returned list)."""
    b = bpa.Bodies(sc)
    keep = []

    def dev(a):
        if a.size == 0:
            return None
        t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        keep.append(t)
        return t.data_ptr()

    d = bpa._lib.BchBodies()
    d.n = b.c.n
    d.position = dev(b.pos)
    d.orientation = dev(b.quat)
    d.shape_type = dev(b.type)
    d.shape_params = dev(b.params)
    d.hull_first = dev(b.hull_first)
    d.hull_count = dev(b.hull_count)
    d.hull_vertices = dev(b.hull_vertices)
    d.n_hull_vertices = b.hull_vertices.shape[0]
    return d, keep


def _static(bpa, torch, ctx, sc, pairs, seeds=None):
    d, keep = _device_bodies(bpa, torch, sc)
    P = pairs.shape[0]
    dp = torch.from_numpy(np.ascontiguousarray(pairs, np.uint32).view(np.int32)).to("cuda")
    ds = None if seeds is None else torch.from_numpy(np.ascontiguousarray(seeds, np.float32)).to("cuda")
    cnt = torch.full((max(P, 1),), 255, dtype=torch.uint8, device="cuda")
    pts = torch.full((max(P, 1) * 28,), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = ctx.L.bch_collide_batch_static(ctx.h, C.byref(d), dp.data_ptr(), P,
                                        None if ds is None else ds.data_ptr(), cnt.data_ptr(), pts.data_ptr())
    if rc != bpa._lib.BCH_OK:
        raise bpa.BchError(rc, ctx.L.bch_last_error(ctx.h).decode())
    assert ctx.L.bch_synchronize(ctx.h) == 0
    del keep
    return cnt.cpu().numpy()[:P], pts.cpu().numpy()[:28 * P].reshape(P, 4, 7)


def _check_same(label, got, ref):
    gc, gp = got
    rc, rp = ref
    assert np.array_equal(gc, rc), f"{label}: counts differ at {np.nonzero(gc != rc)[0][:10]}"
    assert np.array_equal(gp, rp), f"{label}: points differ"
    _flush(f"{label}: {gc.shape[0]} pairs, {int(gc.sum())} contacts, all bit-identical")


def test_static_batch_matches_collide_batch_on_two_calls(bpa, scenes, oracle):
    import torch
    ctx = bpa._Context(0)
    ref_ctx = bpa._Context(0)
    rng = np.random.default_rng(17)

    _flush("case: static batch call 1, mixed 7000 (hulls, capsules, boxes, spheres), pairs shuffled with repeats")
    sc = scenes.mixed_convex(7000, 31)
    pairs = oracle.sweep_pairs(*oracle.scene_aabbs(sc, MARGIN))
    pairs = pairs[rng.permutation(pairs.shape[0])][:30000]
    pairs = np.concatenate([pairs, pairs[:500]])
    assert pairs.shape[0] <= 32768
    got = _static(bpa, torch, ctx, sc, pairs)
    _check_same("static call 1", got, bpa.collide_batch(sc, pairs, ctx=ref_ctx))
    assert int(got[0].sum()) > 10000

    _flush("case: static batch call 2, 700 boxes and spheres on a plane, seeds with per-pair None")
    sc2 = scenes.boxes_spheres_on_plane(700, 5)
    pairs2 = oracle.sweep_pairs(*oracle.scene_aabbs(sc2, MARGIN))
    seeds = rng.normal(size=(pairs2.shape[0], 3)).astype(np.float32)
    seeds[::3, 0] = np.nan
    assert pairs2.shape[0] < pairs.shape[0] // 4
    got2 = _static(bpa, torch, ctx, sc2, pairs2, seeds)
    _check_same("static call 2", got2, bpa.collide_batch(sc2, pairs2, seeds, ctx=ref_ctx))
    rc, rp, _ = oracle.collide_pairs(sc2, pairs2, seeds)
    assert np.array_equal(got2[0], rc)

    _flush("case: static batch call 3, empty")
    got3 = _static(bpa, torch, ctx, sc2, np.zeros((0, 2), np.uint32))
    assert got3[0].shape == (0,)
    ref_ctx.close()
    ctx.close()


def test_static_batch_rejects_what_it_cannot_hold(bpa, scenes):
    import torch
    ctx = bpa._Context(0)
    sc = scenes.mixed_convex(500, 3)
    _flush("case: static batch limits and bad inputs")
    with pytest.raises(bpa.BchError) as e:  # more pairs than the static array holds: refused on the host
        d, keep = _device_bodies(bpa, torch, sc)
        rc = ctx.L.bch_collide_batch_static(ctx.h, C.byref(d), 0, 32769, None, 0, 0)
        raise bpa.BchError(rc, ctx.L.bch_last_error(ctx.h).decode())
    assert e.value.code == bpa._lib.BCH_INVALID_ARG
    bad = np.array([[0, 1], [2, 500]], np.uint32)  # index 500 is out of range: the device check refuses it
    with pytest.raises(bpa.BchError) as e:
        _static(bpa, torch, ctx, sc, bad)
    assert e.value.code == bpa._lib.BCH_INVALID_ARG
    hulls = np.nonzero(sc["shape_type"] == 3)[0]
    broken = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sc.items()}
    broken["hull_count"][hulls[0]] = 10 ** 6  # hull range past the vertex array
    with pytest.raises(bpa.BchError) as e:
        _static(bpa, torch, ctx, broken, np.array([[0, 1]], np.uint32))
    assert e.value.code == bpa._lib.BCH_INVALID_ARG
    # a good call after the refusals still matches
    pairs = np.array([[i, i + 1] for i in range(0, 400, 2)], np.uint32)
    _check_same("static after refusals", _static(bpa, torch, ctx, sc, pairs), bpa.collide_batch(sc, pairs, ctx=ctx))
    ctx.close()
