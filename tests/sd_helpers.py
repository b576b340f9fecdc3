"""Finite-difference helpers for the signed-distance adjoint tests (f64,
through the oracle)."""
import numpy as np


def pose14(sc):
    pos = sc["position"].astype(np.float64)
    q = sc["orientation"].astype(np.float64)
    p = sc["pairs"].astype(np.int64)
    return np.concatenate([pos[p[:, 0]], q[p[:, 0]], pos[p[:, 1]], q[p[:, 1]]], axis=1)


def fd(oracle, sc, g10, h):
    """f64 central differences of the oracle's signed distance: VJP g10^T J
    [P, 14] and the second differences |g.(f+ - 2 f0 + f-)| / h per pose
    coordinate [P, 14] (proportional to h where the function is smooth)."""
    base = pose14(sc)
    f0, st0 = oracle.sd_pairs(sc, sc["pairs"], base, f64=True)
    P = base.shape[0]
    vjp = np.zeros((P, 14))
    jump = np.zeros((P, 14))
    for c in range(14):
        xp = base.copy()
        xm = base.copy()
        xp[:, c] += h
        xm[:, c] -= h
        fp, _ = oracle.sd_pairs(sc, sc["pairs"], xp, f64=True)
        fm, _ = oracle.sd_pairs(sc, sc["pairs"], xm, f64=True)
        vjp[:, c] = np.sum(g10 * (fp - fm), axis=1) / (2 * h)
        jump[:, c] = np.abs(np.sum(g10 * (fp - 2 * f0 + fm), axis=1)) / h
    return vjp, jump, f0, st0


def witness_margin(sc, n):
    """How far A's support along n is from a tie (projection gap between the
    best local support point and the next one it could switch to)."""
    p = sc["pairs"].astype(np.int64)
    ia = p[:, 0]
    q = sc["orientation"][ia].astype(np.float64)
    b, w = q[:, :3], q[:, 3:4]
    nl = (n * (w * w - np.sum(b * b, 1, keepdims=True)) + (-b) * (2 * np.sum(n * -b, 1, keepdims=True))
          + np.cross(-b, n) * (2 * w))  # n in A's frame (rotation by the conjugate)
    t = sc["shape_type"][ia]
    par = sc["shape_params"][ia].astype(np.float64)
    out = np.full(p.shape[0], np.inf)
    for k in range(p.shape[0]):
        if t[k] == 3:
            v = sc["hull_vertices"][sc["hull_first"][ia[k]]: sc["hull_first"][ia[k]] + sc["hull_count"][ia[k]]]
            pr = np.sort(v.astype(np.float64) @ nl[k])[::-1]
            out[k] = pr[0] - pr[1]
        elif t[k] == 1:
            out[k] = 2 * np.min(np.abs(nl[k]) * par[k, :3])
        elif t[k] == 2:
            out[k] = 2 * abs(nl[k, 1]) * par[k, 1]
    return out


def rounding_exit_scene():
    """The recorded rounding-exit pairs (tests/golden/sd_rounding_exits.json)
    as a scene of 2k bodies with pairs (2k, 2k + 1), plus the case list."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sd_rounding_exits.json")
    with open(path) as f:
        cases = json.load(f)["cases"]
    n = 2 * len(cases)
    sc = dict(position=np.zeros((n, 3), np.float32), orientation=np.zeros((n, 4), np.float32),
              shape_type=np.zeros(n, np.uint8), shape_params=np.zeros((n, 4), np.float32),
              hull_first=np.zeros(n, np.uint32), hull_count=np.zeros(n, np.uint32))
    verts = []
    for k, c in enumerate(cases):
        for side, body in (("a", 2 * k), ("b", 2 * k + 1)):
            s = c["shape_" + side]
            sc["position"][body] = np.float32(c["pos_" + side])
            sc["orientation"][body] = np.float32(c["quat_" + side])
            if s["type"] == "box":
                sc["shape_type"][body] = 1
                sc["shape_params"][body, :3] = np.float32(s["half_extents"])
            else:
                sc["shape_type"][body] = 3
                sc["hull_first"][body] = sum(len(v) for v in verts)
                sc["hull_count"][body] = len(s["points"])
                verts.append(np.float32(s["points"]))
    sc["hull_vertices"] = np.concatenate(verts).reshape(-1, 3) if verts else np.zeros((0, 3), np.float32)
    sc["pairs"] = np.arange(n, dtype=np.uint32).reshape(-1, 2)
    return sc, cases


def touch_scene(oracle, trials=300, seed=1):
    """Box-hull pairs whose cores touch: a box (half-extents U[0.25, 1] m,
    margin 0.04) at the origin and a hull of 24 points on an ellipsoid
    (semi-axes U[0.3, 1.2] m, margin 0) moved along a random direction u to
    where the f64 core distance crosses zero (80 bisection steps), both at
    random orientations.  Kept: the pairs whose f32 walk ends on the touch
    gate with no enclosing tetrahedron (status SD_TOUCH).  Returns a scene
    with pairs (2k, 2k + 1) and sc["direction"] = u."""
    rng = np.random.default_rng(seed)
    keep = []
    for _ in range(trials):
        he = rng.uniform(0.25, 1.0, 3).astype(np.float32)
        d = rng.normal(size=(24, 3))
        pts = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0.3, 1.2, 3)).astype(np.float32)
        qa = rng.normal(size=4)
        qb = rng.normal(size=4)
        qa = (qa / np.linalg.norm(qa)).astype(np.float32)
        qb = (qb / np.linalg.norm(qb)).astype(np.float32)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        a, b = oracle.box(*he), oracle.hull(pts)
        lo, hi = 0.0, 4.0
        for _ in range(80):
            m = 0.5 * (lo + hi)
            out, _ = oracle.signed_distance(a, np.zeros(3), qa, b, m * u, qb, f64=True)
            if out[0] + 0.04 < 0.0:
                lo = m
            else:
                hi = m
        pb = (hi * u).astype(np.float32)
        _, st = oracle.signed_distance(a, np.zeros(3), qa, b, pb, qb)
        if st == oracle.SD_TOUCH:
            keep.append((he, pts, qa, qb, pb, u))
    n = 2 * len(keep)
    sc = dict(position=np.zeros((n, 3), np.float32), orientation=np.zeros((n, 4), np.float32),
              shape_type=np.zeros(n, np.uint8), shape_params=np.zeros((n, 4), np.float32),
              hull_first=np.zeros(n, np.uint32), hull_count=np.zeros(n, np.uint32))
    for k, (he, pts, qa, qb, pb, u) in enumerate(keep):
        sc["orientation"][2 * k] = qa
        sc["orientation"][2 * k + 1] = qb
        sc["position"][2 * k + 1] = pb
        sc["shape_type"][2 * k] = 1
        sc["shape_params"][2 * k, :3] = he
        sc["shape_type"][2 * k + 1] = 3
        sc["hull_first"][2 * k + 1] = 24 * k
        sc["hull_count"][2 * k + 1] = 24
    sc["hull_vertices"] = (np.concatenate([k[1] for k in keep]) if keep else np.zeros((0, 3))).astype(np.float32)
    sc["pairs"] = np.arange(n, dtype=np.uint32).reshape(-1, 2)
    sc["direction"] = np.array([k[5] for k in keep]).reshape(-1, 3)
    return sc
