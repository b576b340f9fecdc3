"""GPU parity of the convex-vs-triangle-mesh batch (bch_collide_mesh_batch)
against the oracle's convex_mesh restatement (oracle.hpp, pinned by the
reference's tests in tests/test_oracle_mesh.py): counts exact, points
within the contact tolerance (bit-identical expected), on a wavy terrain and
a flat grid (flat contacts give equal depths, which exercise the tie order
of ContactManifold::reduce), boxes, spheres, capsules and hulls, seeds with
per-pair None and mesh-first pairs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_M = 1e-5
TOL_RAD = 1e-4


def _flush(msg):
    print(msg, flush=True)


def _grid(n, spacing, height):
    h = (n - 1) * spacing / 2.0
    xs = np.linspace(-h, h, n)
    v = np.array([[x, height(x, z), z] for z in xs for x in xs], np.float32)
    t = []
    for r in range(n - 1):
        for c in range(n - 1):
            a, b, d, e = r * n + c, r * n + c + 1, (r + 1) * n + c, (r + 1) * n + c + 1
            t += [[a, d, b], [b, d, e]]
    return v, np.array(t, np.uint32)


def _rot_y(a):
    return np.array([0.0, np.sin(a / 2), 0.0, np.cos(a / 2)], np.float32)


def test_convex_mesh_batch_matches_oracle(bpa, scenes, oracle, ctx):
    _flush("case: convex vs mesh batch, terrain 65x65 + flat grid 21x21")
    rng = np.random.default_rng(9)
    terrain = _grid(65, 0.5, lambda x, z: 0.6 * np.sin(0.7 * x) * np.cos(0.5 * z))
    flat = _grid(21, 1.0, lambda x, z: 0.0)
    meshes_gpu = [bpa.TriangleMesh(*terrain, ctx=ctx), bpa.TriangleMesh(*flat, ctx=ctx)]
    meshes_ref = [oracle.Mesh(*terrain), oracle.Mesh(*flat)]
    assert [m.id for m in meshes_gpu] == [0, 1]
    sc = scenes.mixed_convex(3000, 7)
    n = sc["shape_type"].shape[0]
    keep = sc["shape_type"] != 4
    idx = np.nonzero(keep)[0]
    # place the bodies over the meshes: half over the terrain, half resting
    # over the flat grid (boxes axis-aligned and just touching: equal depths)
    pos = sc["position"].copy()
    quat = sc["orientation"].copy()
    pairs, poses, first = [], [], []
    for k, b in enumerate(idx):
        m = k % 2
        x, z = rng.uniform(-14, 14, 2)
        if m == 0:
            pos[b] = [x, 0.6 * np.sin(0.7 * x) * np.cos(0.5 * z) + rng.uniform(-0.3, 0.6), z]
            mp = np.concatenate([rng.normal(0, 0.01, 3), _rot_y(rng.normal(0, 0.02))])
        else:
            pos[b] = [x, rng.uniform(0.2, 0.9), z]
            if sc["shape_type"][b] == 1 and k % 4 == 1:  # axis-aligned resting box
                quat[b] = [0, 0, 0, 1]
                pos[b, 1] = sc["shape_params"][b, 1] - 0.01
            mp = np.array([0, 0, 0, 0, 0, 0, 1], np.float32)
        pairs.append([b, m])
        poses.append(mp)
        first.append(1 if k % 3 == 0 else 0)
    sc = dict(sc)
    sc["position"], sc["orientation"] = pos.astype(np.float32), quat.astype(np.float32)
    pairs = np.array(pairs, np.uint32)
    poses = np.array(poses, np.float32)
    first = np.array(first, np.uint8)
    seeds = rng.normal(size=(pairs.shape[0], 3)).astype(np.float32)
    seeds /= np.linalg.norm(seeds, axis=1, keepdims=True)
    seeds[::2, 0] = np.nan  # None for every other pair
    cnt, pts = bpa.collide_mesh_batch(sc, pairs, poses, first, seeds, ctx=ctx)
    rc, rp = oracle.mesh_collide_scene(sc, meshes_ref, pairs, poses, seeds, first)
    _flush(f"{pairs.shape[0]} pairs, {int((cnt > 0).sum())} touching, {int(cnt.sum())} points "
           f"({int((cnt == 4).sum())} reduced to four)")
    bad = np.nonzero(cnt != rc)[0]
    assert bad.size == 0, f"counts differ at {bad[:8]}: {cnt[bad[:8]]} vs {rc[bad[:8]]}"
    m = np.arange(4)[None, :] < cnt[:, None]
    a, b = pts[m], rp[m]
    exact = int(np.sum(np.all(a == b, axis=1)))
    dp = np.linalg.norm(a[:, :3].astype(np.float64) - b[:, :3], axis=1)
    dd = np.abs(a[:, 6].astype(np.float64) - b[:, 6])
    na, nb = a[:, 3:6].astype(np.float64), b[:, 3:6].astype(np.float64)
    ang = np.arctan2(np.linalg.norm(np.cross(na, nb), axis=1), np.sum(na * nb, axis=1))
    _flush(f"points: {a.shape[0]}, {exact} bit-identical, max |dp| {dp.max():.3g}, |dd| {dd.max():.3g}, "
           f"angle {ang.max():.3g}")
    assert np.all(dp <= TOL_M) and np.all(dd <= TOL_M) and np.all(ang <= TOL_RAD)
    assert int((cnt == 4).sum()) > 0 and int((cnt > 0).sum()) > 100
