"""GPU parity of the CCD sweep batches (bch_sweep_batch, bch_sweep_mesh_batch)
against the oracle's restatement of narrowphase/convex_sweep.rs (oracle.hpp,
pinned by the reference's sweep tests and a dense-scan entry check in
tests/test_oracle_mesh.py): hit flags exact; entry distance and point within
1e-5 m, normal within 1e-4 rad (bit-identical expected) for sphere sweeps
and convex casts against spheres, boxes, capsules, hulls and planes, and
against a terrain mesh at a moved pose.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_M = 1e-5
TOL_RAD = 1e-4


def _flush(msg):
    print(msg, flush=True)


def _compare(tag, frm, to, hit, out, rh, ro):
    bad = np.nonzero(hit != rh)[0]
    assert bad.size == 0, f"{tag}: hit flags differ at {bad[:8]}"
    m = hit == 1
    a, b = out[m].astype(np.float64), ro[m].astype(np.float64)
    length = np.linalg.norm((to - frm)[m].astype(np.float64), axis=1)
    exact = int(np.sum(np.all(out[m] == ro[m], axis=1)))
    df = np.abs(a[:, 0] - b[:, 0]) * length
    dp = np.linalg.norm(a[:, 4:7] - b[:, 4:7], axis=1)
    na, nb = a[:, 1:4], b[:, 1:4]
    ang = np.arctan2(np.linalg.norm(np.cross(na, nb), axis=1), np.sum(na * nb, axis=1))
    _flush(f"{tag}: {hit.shape[0]} queries, {int(m.sum())} hits, {exact} bit-identical, "
           f"max |df|*len {df.max(initial=0):.3g} m, |dp| {dp.max(initial=0):.3g}, angle {ang.max(initial=0):.3g}")
    assert np.all(df <= TOL_M) and np.all(dp <= TOL_M) and np.all(ang <= TOL_RAD)
    assert np.all(out[~m] == 0.0)
    return int(m.sum())


def _with_planes(scenes, sc, rng, every=40):
    """Every `every`-th box or sphere becomes a plane through its centre c
    with a random world normal nw (local normal = nw in the body frame).  The
    sphere sweep reads the plane as nw . x = constant and the convex cast as
    nw . x = constant + nw . c (convex_sweep.rs:258-292, :431-478), so
    alternate planes carry constant nw . c (sphere sweeps reach them) and 0
    (convex casts reach them)."""
    sc = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in sc.items()}
    idx = np.nonzero(np.isin(sc["shape_type"], [0, 1]))[0][::every]
    nw = rng.normal(size=(idx.shape[0], 3))
    nw /= np.linalg.norm(nw, axis=1, keepdims=True)
    q = sc["orientation"][idx].astype(np.float64)
    qc = q * np.array([-1, -1, -1, 1.0])
    sc["shape_type"][idx] = 4
    sc["shape_params"][idx, :3] = scenes._qrot_np(qc, nw)
    const = np.sum(nw * sc["position"][idx], axis=1)
    const[1::2] = 0.0
    sc["shape_params"][idx, 3] = const
    return sc


def _queries(rng, sc, n):
    """Segments that start 1-3 m away from a target body and end near its
    centre (mostly hits), or wander past it (mostly misses)."""
    nb = sc["shape_type"].shape[0]
    target = rng.integers(0, nb, n).astype(np.uint32)
    c = sc["position"][target].astype(np.float64)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    frm = c + d * rng.uniform(1.0, 3.0, (n, 1))
    to = c + rng.normal(0, 0.3, (n, 3))
    miss = rng.random(n) < 0.3
    to[miss] = frm[miss] + rng.normal(0, 0.8, (int(miss.sum()), 3))
    return target, frm.astype(np.float32), to.astype(np.float32)


def test_sphere_sweeps_and_convex_casts_match_oracle(bpa, scenes, oracle, ctx):
    _flush("case: sphere sweeps and convex casts against mixed bodies and planes, 6000 queries each")
    rng = np.random.default_rng(5)
    sc = _with_planes(scenes, scenes.mixed_convex(2000, 11), rng)
    target, frm, to = _queries(rng, sc, 6000)
    radius = rng.uniform(0.05, 0.5, 6000).astype(np.float32)
    hit, out = bpa.sweep_batch(sc, target, frm, to, radius=radius, ctx=ctx)
    rh, ro = oracle.sweep_scene(sc, target, frm, to, radius=radius)
    n_hit = _compare("sphere sweeps", frm, to, hit, out, rh, ro)
    assert n_hit > 1000
    planes = sc["shape_type"][target] == 4
    assert int(hit[planes].sum()) > 0

    convex = np.nonzero(sc["shape_type"] != 4)[0]
    caster = convex[rng.integers(0, convex.shape[0], 6000)].astype(np.uint32)
    target, frm, to = _queries(rng, sc, 6000)
    hit, out = bpa.sweep_batch(sc, target, frm, to, caster=caster, ctx=ctx)
    rh, ro = oracle.sweep_scene(sc, target, frm, to, caster=caster)
    n_hit = _compare("convex casts", frm, to, hit, out, rh, ro)
    assert n_hit > 1000
    assert int(hit[sc["shape_type"][target] == 4].sum()) > 0
    kinds = set(int(k) for k in sc["shape_type"][caster[hit == 1]])
    assert kinds == {0, 1, 2, 3}


def test_mesh_sweeps_match_oracle(bpa, scenes, oracle, ctx):
    _flush("case: sphere sweeps and convex casts against a 33x33 terrain mesh at a moved pose")
    n = 33
    xs = np.linspace(-8, 8, n)
    v = np.array([[x, 0.5 * np.sin(0.6 * x) * np.cos(0.4 * z), z] for z in xs for x in xs], np.float32)
    t = []
    for r in range(n - 1):
        for c in range(n - 1):
            a, b, d, e = r * n + c, r * n + c + 1, (r + 1) * n + c, (r + 1) * n + c + 1
            t += [[a, d, b], [b, d, e]]
    t = np.array(t, np.uint32)
    mg = bpa.TriangleMesh(v, t, ctx=ctx)
    mr = oracle.Mesh(v, t)
    rng = np.random.default_rng(8)
    q = 3000
    ang = 0.15
    pose = np.array([0.3, -0.2, 0.1, 0.0, np.sin(ang / 2), 0.0, np.cos(ang / 2)], np.float32)
    poses = np.repeat(pose[None, :], q, axis=0)
    mesh_of = np.full(q, mg.id, np.uint32)
    frm = np.stack([rng.uniform(-7, 7, q), rng.uniform(1.0, 3.0, q), rng.uniform(-7, 7, q)], 1).astype(np.float32)
    to = (frm + np.stack([rng.normal(0, 1.0, q), -rng.uniform(1.5, 4.5, q), rng.normal(0, 1.0, q)], 1)).astype(
        np.float32)
    radius = rng.uniform(0.05, 0.4, q).astype(np.float32)
    hit, out = bpa.sweep_mesh_batch(mesh_of, poses, frm, to, radius=radius, ctx=ctx)
    rh, ro = oracle.sweep_mesh_scene(None, [mr] * (mg.id + 1), mesh_of, poses, frm, to, radius=radius)
    assert _compare("sphere vs mesh", frm, to, hit, out, rh, ro) > 500

    # convex casts: every swept triangle runs a full search (the oracle takes
    # ~20 ms per query here), so a smaller batch
    k = 400
    sc = scenes.mixed_convex(400, 12)
    convex = np.nonzero(sc["shape_type"] != 4)[0]
    caster = convex[rng.integers(0, convex.shape[0], k)].astype(np.uint32)
    hit, out = bpa.sweep_mesh_batch(mesh_of[:k], poses[:k], frm[:k], to[:k], caster=caster, scene=sc, ctx=ctx)
    rh, ro = oracle.sweep_mesh_scene(sc, [mr] * (mg.id + 1), mesh_of[:k], poses[:k], frm[:k], to[:k], caster=caster)
    assert _compare("convex vs mesh", frm[:k], to[:k], hit, out, rh, ro) > 200


def test_sweep_edge_cases(bpa, scenes, oracle, ctx):
    _flush("case: sweep edge cases (empty batch, zero-length segment, bad indices, plane caster)")
    sc = _with_planes(scenes, scenes.mixed_convex(50, 3), np.random.default_rng(1), every=5)
    hit, out = bpa.sweep_batch(sc, np.zeros(0, np.uint32), np.zeros((0, 3)), np.zeros((0, 3)),
                               radius=np.zeros(0), ctx=ctx)
    assert hit.shape == (0,) and out.shape == (0, 7)
    p = sc["position"][:1]
    hit, out = bpa.sweep_batch(sc, [0], p, p, radius=[0.3], ctx=ctx)
    assert hit[0] == 0 and np.all(out == 0)
    with pytest.raises(RuntimeError, match="out of range"):
        bpa.sweep_batch(sc, [50], p, p + 1, radius=[0.3], ctx=ctx)
    plane = int(np.nonzero(sc["shape_type"] == 4)[0][0])
    with pytest.raises(RuntimeError, match="convex"):
        bpa.sweep_batch(sc, [0], p, p + 1, caster=[plane], ctx=ctx)
    with pytest.raises(RuntimeError, match="mesh index"):
        bpa.sweep_mesh_batch([1000], [[0, 0, 0, 0, 0, 0, 1]], p, p + 1, radius=[0.3], ctx=ctx)
