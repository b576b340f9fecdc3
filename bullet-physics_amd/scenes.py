"""Seeded synthetic scenes for the collision pass.

All randomness comes from the reference's test LCG
(crates/bullet-collision/src/broadphase/axis_sweep3.rs:823-852):
``s = s * 6364136223846793005 + 1442695040888963407 (mod 2^64)`` and
``unit() = (s >> 11) / 2^53`` rounded to f32.  The stream is generated
vectorised by jumping ahead (``s_k = A_k s_0 + C_k`` with ``A_k = a^k``,
``C_k = c (1 + a + ... + a^(k-1))``, all mod 2^64).

A scene is a dict of numpy arrays in the library's body layout:
``position f32[n,3]``, ``orientation f32[n,4]`` (x, y, z, w),
``shape_type u8[n]`` (0 sphere, 1 box, 2 capsule, 3 hull, 4 plane),
``shape_params f32[n,4]`` (sphere r | box hx hy hz | capsule r hh |
plane nx ny nz c), ``hull_first u32[n]``, ``hull_count u32[n]``,
``hull_vertices f32[V,3]``, plus ``world_offsets u32[W+1]`` for batches.
"""
from __future__ import annotations

import math

import numpy as np

SPHERE, BOX, CAPSULE, HULL, PLANE = 0, 1, 2, 3, 4
_A = np.uint64(6364136223846793005)
_C = np.uint64(1442695040888963407)


class Lcg:
    """The reference's test LCG, drawn in vectorised blocks."""

    def __init__(self, seed: int):
        self.s = np.uint64(seed)

    def next_u64(self, k: int) -> np.ndarray:
        if k <= 0:
            return np.zeros(0, np.uint64)
        with np.errstate(over="ignore"):
            a = np.full(k, _A, dtype=np.uint64)
            ak = np.cumprod(a, dtype=np.uint64)              # a^1 .. a^k
            geo = np.concatenate(([np.uint64(1)], ak[:-1]))  # a^0 .. a^(k-1)
            ck = np.cumsum(geo, dtype=np.uint64) * _C        # c (1 + .. + a^(k-1))
            out = ak * self.s + ck
        self.s = out[-1]
        return out

    def unit(self, k: int) -> np.ndarray:
        """(next >> 11) as f32 / 2^53 -- the f32 build of ``Lcg::unit``."""
        x = (self.next_u64(k) >> np.uint64(11)).astype(np.float64)  # exact (< 2^53)
        return x.astype(np.float32) / np.float32(2.0 ** 53)

    def range(self, lo: float, hi: float, k: int) -> np.ndarray:
        u = self.unit(k)
        return np.float32(lo) + (np.float32(hi) - np.float32(lo)) * u


def random_quaternions(rng: Lcg, k: int) -> np.ndarray:
    """Uniform unit quaternions (Shoemake's u1, u2, u3 method), (x, y, z, w)."""
    u = rng.unit(3 * k).astype(np.float64).reshape(k, 3)
    u1, u2, u3 = u[:, 0], u[:, 1], u[:, 2]
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * math.pi * u2), a * np.cos(2 * math.pi * u2),
                  b * np.sin(2 * math.pi * u3), b * np.cos(2 * math.pi * u3)], axis=1)
    return q.astype(np.float32)


def _gaussian(rng: Lcg, k: int) -> np.ndarray:
    """Box-Muller normals from LCG uniforms."""
    u = rng.unit(2 * k).astype(np.float64).reshape(k, 2)
    u1 = np.maximum(u[:, 0], 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * math.pi * u[:, 1])


def ellipsoid_hull(rng: Lcg, npts: int, semi_lo: float, semi_hi: float) -> np.ndarray:
    """``npts`` points on a random ellipsoid (every point is extreme)."""
    axes = rng.range(semi_lo, semi_hi, 3).astype(np.float64)
    g = _gaussian(rng, 3 * npts).reshape(npts, 3)
    d = g / np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-30)
    return (d * axes).astype(np.float32)


def _empty(n: int) -> dict:
    return dict(position=np.zeros((n, 3), np.float32), orientation=np.tile(np.float32([0, 0, 0, 1]), (n, 1)),
                shape_type=np.zeros(n, np.uint8), shape_params=np.zeros((n, 4), np.float32),
                hull_first=np.zeros(n, np.uint32), hull_count=np.zeros(n, np.uint32),
                hull_vertices=np.zeros((0, 3), np.float32))


def boxes_spheres_on_plane(n: int = 2048, seed: int = 1, lo=(-20.0, 0.0, -20.0), hi=(20.0, 40.0, 20.0),
                           with_plane: bool = True) -> dict:
    """``n`` boxes/spheres (50/50 by LCG draw) above a static ground plane.

    Boxes: half-extents U[0.25, 1.0] per axis; spheres: radius U[0.25, 1.0];
    centres uniform in [lo, hi]; orientations uniform unit quaternions.
    The plane ``y = 0`` (normal +Y) takes slot 0 when ``with_plane``.
    """
    rng = Lcg(seed)
    m = n + (1 if with_plane else 0)
    sc = _empty(m)
    o = 1 if with_plane else 0
    if with_plane:
        sc["shape_type"][0] = PLANE
        sc["shape_params"][0] = (0.0, 1.0, 0.0, 0.0)
    kind = rng.unit(n) < np.float32(0.5)
    ext = np.stack([rng.range(0.25, 1.0, n) for _ in range(3)], axis=1)
    pos = np.stack([rng.range(lo[k], hi[k], n) for k in range(3)], axis=1)
    quat = random_quaternions(rng, n)
    sc["shape_type"][o:] = np.where(kind, BOX, SPHERE)
    sc["shape_params"][o:, :3] = np.where(kind[:, None], ext, np.stack([ext[:, 0], 0 * ext[:, 0], 0 * ext[:, 0]], 1))
    sc["position"][o:] = pos
    sc["orientation"][o:] = quat
    return sc


def _aabb_widths(sc: dict) -> np.ndarray:
    """Axis-aligned full widths (no margin) of each body, from numpy."""
    t = sc["shape_type"]
    q = sc["orientation"].astype(np.float64)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], 1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], 1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)], 1)
    p = sc["shape_params"].astype(np.float64)
    wd = np.zeros((t.shape[0], 3))
    bm = t == BOX
    wd[bm] = 2 * np.einsum("nij,nj->ni", np.abs(R[bm]), p[bm, :3])
    sm = t == SPHERE
    wd[sm] = 2 * p[sm, :1]
    cm = t == CAPSULE
    wd[cm] = 2 * (np.abs(R[cm][:, :, 1]) * p[cm, 1:2] + p[cm, 0:1])
    hm = np.nonzero(t == HULL)[0]
    hv = sc["hull_vertices"].astype(np.float64)
    for i in hm:
        v = hv[sc["hull_first"][i]: sc["hull_first"][i] + sc["hull_count"][i]] @ R[i].T
        wd[i] = v.max(0) - v.min(0)
    return wd


def mixed_convex(n: int = 1_000_000, seed: int = 2, pairs_per_body: float = 4.0, margin: float = 0.04,
                 mix=(0.4, 0.3, 0.2, 0.1), hull_pts=(16, 64), side: float | None = None) -> dict:
    """``n`` convex bodies: boxes / spheres / capsules / hulls in ``mix``.

    Boxes half-extents U[0.25, 1], spheres r U[0.25, 1], capsules r U[0.1, 0.5]
    and half-height U[0.2, 1], hulls of 16-64 points on ellipsoids with
    semi-axes U[0.3, 1.2]; random orientations; centres uniform in a cube whose
    side is chosen (unless given) so that the expected number of overlapping
    AABB pairs is ``pairs_per_body * n``.
    """
    rng = Lcg(seed)
    sc = _empty(n)
    u = rng.unit(n)
    c = np.cumsum(mix)
    t = np.where(u < c[0], BOX, np.where(u < c[1], SPHERE, np.where(u < c[2], CAPSULE, HULL))).astype(np.uint8)
    sc["shape_type"][:] = t
    ext = np.stack([rng.range(0.25, 1.0, n) for _ in range(3)], axis=1)
    cap_r = rng.range(0.1, 0.5, n)
    cap_h = rng.range(0.2, 1.0, n)
    params = np.zeros((n, 4), np.float32)
    params[t == BOX, :3] = ext[t == BOX]
    params[t == SPHERE, 0] = ext[t == SPHERE, 0]
    params[t == CAPSULE, 0] = cap_r[t == CAPSULE]
    params[t == CAPSULE, 1] = cap_h[t == CAPSULE]
    sc["shape_params"] = params
    hull_ids = np.nonzero(t == HULL)[0]
    counts = (hull_pts[0] + (rng.unit(hull_ids.shape[0]) * (hull_pts[1] - hull_pts[0] + 1)).astype(np.int64)
              ).clip(hull_pts[0], hull_pts[1])
    first = np.concatenate(([0], np.cumsum(counts)[:-1])).astype(np.uint32)
    total = int(counts.sum())
    # Per-hull semi-axes and unit directions, drawn in bulk.
    axes = np.stack([rng.range(0.3, 1.2, hull_ids.shape[0]) for _ in range(3)], 1).astype(np.float64)
    g = np.stack([_gaussian(rng, total) for _ in range(3)], 1)
    d = g / np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-30)
    owner = np.repeat(np.arange(hull_ids.shape[0]), counts)
    sc["hull_vertices"] = (d * axes[owner]).astype(np.float32)
    sc["hull_first"][hull_ids] = first
    sc["hull_count"][hull_ids] = counts.astype(np.uint32)
    sc["orientation"] = random_quaternions(rng, n)
    if side is None:
        wd = _aabb_widths(sc) + 2 * margin
        k = min(n, 200_000)
        ia = (rng.unit(k) * n).astype(np.int64).clip(0, n - 1)
        ib = (rng.unit(k) * n).astype(np.int64).clip(0, n - 1)
        e = np.mean(np.prod(wd[ia] + wd[ib], axis=1))
        side = float((n * e / (2.0 * pairs_per_body)) ** (1.0 / 3.0))
    pos = np.stack([rng.range(0.0, side, n) for _ in range(3)], axis=1)
    sc["position"] = pos.astype(np.float32)
    sc["side"] = side
    return sc


def batched_worlds(n_worlds: int = 4096, seed: int = 3, lo_count: int = 16, hi_count: int = 2048,
                   cube: float = 20.0) -> dict:
    """``n_worlds`` independent worlds concatenated body-major.

    World w has ``floor(exp(U[ln lo, ln hi]))`` boxes/spheres (the mix of
    ``boxes_spheres_on_plane``) in its own ``cube``-metre cube above its own
    ground plane (its slot 0).  ``world_offsets[w]`` is its first body.
    """
    rng = Lcg(seed)
    u = rng.unit(n_worlds).astype(np.float64)
    counts = np.floor(np.exp(math.log(lo_count) + u * (math.log(hi_count) - math.log(lo_count)))).astype(np.int64)
    counts = counts.clip(lo_count, hi_count)
    parts = []
    for w in range(n_worlds):
        parts.append(boxes_spheres_on_plane(int(counts[w]), seed=(seed << 32) + w, lo=(0.0, 0.0, 0.0),
                                            hi=(cube, cube, cube)))
    sc = concat(parts)
    return sc


def concat(parts) -> dict:
    """Concatenate scenes into one batch with ``world_offsets``."""
    out = {}
    sizes = [p["shape_type"].shape[0] for p in parts]
    for k in ("position", "orientation", "shape_type", "shape_params", "hull_count"):
        out[k] = np.concatenate([p[k] for p in parts])
    vbase = np.cumsum([0] + [p["hull_vertices"].shape[0] for p in parts])
    out["hull_first"] = np.concatenate([p["hull_first"] + np.uint32(vbase[i]) for i, p in enumerate(parts)])
    out["hull_vertices"] = np.concatenate([p["hull_vertices"] for p in parts]).reshape(-1, 3)
    out["world_offsets"] = np.concatenate(([0], np.cumsum(sizes))).astype(np.uint32)
    return out


def world_slice(sc: dict, w: int) -> dict:
    """World ``w`` of a batch as a stand-alone scene (hull indices rebased)."""
    a, b = int(sc["world_offsets"][w]), int(sc["world_offsets"][w + 1])
    out = {k: sc[k][a:b].copy() for k in ("position", "orientation", "shape_type", "shape_params", "hull_count")}
    hf = sc["hull_first"][a:b]
    hm = out["shape_type"] == HULL
    if hm.any():
        v0 = int(hf[hm].min())
        v1 = int((hf[hm] + out["hull_count"][hm]).max())
        out["hull_first"] = np.where(hm, hf - np.uint32(v0), 0).astype(np.uint32)
        out["hull_vertices"] = sc["hull_vertices"][v0:v1].copy()
    else:
        out["hull_first"] = np.zeros(b - a, np.uint32)
        out["hull_vertices"] = np.zeros((0, 3), np.float32)
    return out


def _qrot_np(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Rotate v[..., 3] by unit quaternions q[..., 4] (x, y, z, w), float64."""
    b, w = q[..., :3], q[..., 3:4]
    return v * (w * w - np.sum(b * b, -1, keepdims=True)) + b * (2.0 * np.sum(v * b, -1, keepdims=True)) \
        + np.cross(b, v) * (2.0 * w)


def hull_pairs(n_pairs: int = 262_144, seed: int = 4, npts: int = 32, semi_lo: float = 0.3, semi_hi: float = 1.0,
               gap: float = 0.2, spacing: float = 6.0, shape_mix: bool = False) -> dict:
    """``n_pairs`` independent convex pairs (bodies 2k, 2k+1) for the signed
    distance and its adjoint.

    Every body is a hull of ``npts`` points on a random ellipsoid (semi-axes
    U[semi_lo, semi_hi]) with a uniform random orientation.  Pair k sits at
    its own grid cell; B is placed along a random unit direction u at the
    position where B's support point along -u lies at A's support point along
    u plus d u: d ~ +U[0, gap] for even k (then the signed distance is d:
    the supports face each other across the plane normal to u) and -U[0, gap]
    for odd k (overlapping, penetration at most |d|).  ``shape_mix`` makes every fourth A a box, sphere or
    capsule instead (for the rounded-shape paths).  ``pairs`` holds the
    (2k, 2k+1) index pairs."""
    rng = Lcg(seed)
    nb = 2 * n_pairs
    axes = rng.range(semi_lo, semi_hi, 3 * nb).astype(np.float64).reshape(nb, 1, 3)
    g = _gaussian(rng, 3 * npts * nb).reshape(nb, npts, 3)
    d = g / np.maximum(np.linalg.norm(g, axis=2, keepdims=True), 1e-30)
    verts = (d * axes).astype(np.float32)
    quat = random_quaternions(rng, nb)
    u = _gaussian(rng, 3 * n_pairs).reshape(n_pairs, 3)
    u /= np.maximum(np.linalg.norm(u, axis=1, keepdims=True), 1e-30)
    target = rng.unit(n_pairs).astype(np.float64) * gap
    target = np.where(np.arange(n_pairs) % 2 == 0, target, -target)
    sc = _empty(nb)
    sc["shape_type"][:] = HULL
    sc["hull_count"][:] = npts
    sc["hull_first"][:] = (np.arange(nb) * npts).astype(np.uint32)
    sc["hull_vertices"] = verts.reshape(-1, 3)
    sc["orientation"] = quat
    if shape_mix:
        a_idx = np.arange(0, nb, 2)
        pick = a_idx[(np.arange(n_pairs) % 4) == 3]
        kind = rng.unit(pick.shape[0])
        k3 = np.where(kind < 1 / 3, BOX, np.where(kind < 2 / 3, SPHERE, CAPSULE)).astype(np.uint8)
        sc["shape_type"][pick] = k3
        sc["hull_count"][pick] = 0
        sc["hull_first"][pick] = 0
        p = rng.range(0.25, 0.8, 3 * pick.shape[0]).reshape(-1, 3)
        sc["shape_params"][pick, :3] = p
    # world offsets (from the centre) of A's support point along u and B's
    # along -u; B is placed so that they sit d apart along u
    qa = quat[0::2].astype(np.float64)
    qb = quat[1::2].astype(np.float64)
    wa = _qrot_np(qa[:, None, :], verts[0::2].astype(np.float64))
    wb = _qrot_np(qb[:, None, :], verts[1::2].astype(np.float64))
    rows = np.arange(n_pairs)
    sa = wa[rows, np.argmax(np.einsum("pvk,pk->pv", wa, u), axis=1)]
    sb = wb[rows, np.argmax(np.einsum("pvk,pk->pv", wb, -u), axis=1)]
    if shape_mix:
        m = (np.arange(n_pairs) % 4) == 3
        ta = sc["shape_type"][0::2]
        par = sc["shape_params"][0::2].astype(np.float64)
        qconj = qa * np.array([-1, -1, -1, 1.0])
        ul = _qrot_np(qconj, u)  # u in A's frame
        box = _qrot_np(qa, np.where(ul >= 0, 1.0, -1.0) * par[:, :3])
        sph = u * par[:, :1]
        tip = np.zeros((n_pairs, 3))
        tip[:, 1] = np.where(ul[:, 1] >= 0, par[:, 1], -par[:, 1])
        cap = _qrot_np(qa, tip) + u * par[:, :1]
        sa = np.where((m & (ta == BOX))[:, None], box,
                      np.where((m & (ta == SPHERE))[:, None], sph, np.where((m & (ta == CAPSULE))[:, None], cap, sa)))
    k = np.arange(n_pairs)
    side = max(1, int(round(n_pairs ** (1 / 3))) + 1)
    cell = np.stack([k % side, (k // side) % side, k // (side * side)], axis=1).astype(np.float64) * spacing
    # (spacing = 0 puts every pair at the origin: independent pairs need no
    # room, and small coordinates keep f32 differences of nearby points exact)
    pa = cell
    pb = pa + sa + u * target[:, None] - sb
    pos = np.empty((nb, 3), np.float32)
    pos[0::2] = pa
    pos[1::2] = pb
    sc["position"] = pos
    sc["pairs"] = np.stack([2 * k, 2 * k + 1], axis=1).astype(np.uint32)
    sc["target"] = target.astype(np.float32)
    # B at s along u is b_base + s u (s = target above); exact_penetration
    # moves the overlapping half along this line
    sc["direction"] = u
    sc["b_base"] = pa + sa - sb
    return sc


def exact_penetration(sc: dict, signed_distance, iters: int = 40, reach: float = 2.5, tol: float = 1e-4) -> dict:
    """Place B of every overlapping pair of a :func:`hull_pairs` scene so that
    its signed distance equals its target -U[0, gap], by bisection along the
    pair's direction u.

    ``signed_distance(position)`` returns the signed distance of every pair
    of ``sc["pairs"]`` with the bodies at ``position`` [2P, 3] (float32).
    Along the line B = b_base + s u the signed distance is convex in s (the
    signed distance to the convex Minkowski difference) and equals s for
    s >= 0, so for a target t < 0 the set {s : sd(s) <= t} is an interval
    whose right end is the one crossing in [-reach, 0].  A pair whose line
    never reaches its target depth keeps the deepest point found on a
    coarse scan instead, and its target becomes the signed distance there;
    their number is returned in ``sc["unreached"]``.  A pair whose computed
    signed distance jumps across the crossing (f32 walks near the iteration
    cap) ends more than ``tol`` from its target and takes the signed
    distance it ends at as its target too (``sc["retargeted"]``).
    ``sc["overlap"]`` marks the pairs of the overlapping half.  Separated
    pairs are left as built (exact by construction)."""
    tgt = sc["target"].astype(np.float64)
    pen = tgt < 0
    u = sc["direction"]
    base = sc["b_base"]
    pos = sc["position"].copy()

    b_rows = 2 * np.nonzero(pen)[0] + 1

    def place(s):
        pos[b_rows] = (base[pen] + u[pen] * s[pen, None]).astype(np.float32)

    def sd_at(s):
        place(s)
        return np.asarray(signed_distance(pos), np.float64)

    # a coarse scan for a start at or below the target (the deepest sample
    # where none is)
    n = tgt.shape[0]
    lo = np.full(n, -reach)
    best_s, best_sd = np.zeros(n), np.zeros(n)
    found = ~pen
    for s_try in np.linspace(-reach / 10, -reach, 10):
        sv = np.full(n, s_try)
        sd = sd_at(sv)
        deeper = pen & (sd < best_sd)
        best_s = np.where(deeper, s_try, best_s)
        best_sd = np.where(deeper, sd, best_sd)
        hit = pen & ~found & (sd <= tgt)
        lo = np.where(hit, s_try, lo)
        found |= hit
    unreached = pen & ~found
    tgt = np.where(unreached, best_sd, tgt)
    lo = np.where(unreached, best_s, lo)
    hi = np.zeros(n)
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        below = sd_at(mid) <= tgt
        lo = np.where(pen & below, mid, lo)
        hi = np.where(pen & ~below, mid, hi)
    sd = sd_at(np.where(unreached, best_s, 0.5 * (lo + hi)))
    off = pen & (np.abs(sd - tgt) > tol)
    tgt = np.where(off | unreached, sd, tgt)
    out = dict(sc)
    out["position"] = pos
    out["target"] = tgt.astype(np.float32)
    out["overlap"] = pen
    out["unreached"] = int(unreached.sum())
    out["retargeted"] = int((off & ~unreached).sum())
    return out


def batched_boxes_hulls(n_worlds: int = 1024, seed: int = 5, lo_count: int = 64, hi_count: int = 1024,
                        fill: float = 0.35, gap: float | None = None) -> dict:
    """``n_worlds`` independent worlds of boxes and hulls (50/50), body count
    log-uniform in [lo_count, hi_count], for the differentiable batch.

    Boxes: half-extents U[0.25, 1.0] per axis; hulls: 16-64 points on
    ellipsoids with semi-axes U[0.3, 1.2]; orientations uniform.  World w
    fills its own cube [0, side_w]^3 (side chosen so that the bodies'
    bounding volume is ``fill`` of it, which leaves many pairs
    interpenetrating), in its own coordinates: the cubes of all worlds share
    the origin, as independent worlds do (the batched broadphase keeps them
    apart by world id).  With ``gap`` the cubes instead sit side by side
    along x, ``gap`` metres apart; 1,024 worlds then reach x = 21 km, where
    an f32 ulp is 2 mm and the f32 walks of the reference's arithmetic report
    false overlaps hundreds of millimetres deep."""
    rng = Lcg(seed)
    u = rng.unit(n_worlds).astype(np.float64)
    counts = np.floor(np.exp(math.log(lo_count) + u * (math.log(hi_count) - math.log(lo_count)))).astype(np.int64)
    counts = counts.clip(lo_count, hi_count)
    N = int(counts.sum())
    sc = _empty(N)
    kind = rng.unit(N) < np.float32(0.5)
    ext = np.stack([rng.range(0.25, 1.0, N) for _ in range(3)], axis=1)
    sc["shape_type"][:] = np.where(kind, BOX, HULL)
    sc["shape_params"][:, :3] = np.where(kind[:, None], ext, 0.0)
    hidx = np.nonzero(~kind)[0]
    nh = hidx.shape[0]
    npts = np.floor(rng.range(16.0, 65.0, nh)).astype(np.int64).clip(16, 64)
    # every hull draws 64 ellipsoid points and keeps its first npts
    axes = rng.range(0.3, 1.2, 3 * nh).astype(np.float64).reshape(nh, 1, 3)
    g = _gaussian(rng, 3 * 64 * nh).reshape(nh, 64, 3)
    d = g / np.maximum(np.linalg.norm(g, axis=2, keepdims=True), 1e-30)
    pts = (d * axes).astype(np.float32)
    keep = np.arange(64)[None, :] < npts[:, None]
    sc["hull_vertices"] = pts[keep].reshape(-1, 3)
    first = np.concatenate(([0], np.cumsum(npts)[:-1]))
    sc["hull_first"][hidx] = first.astype(np.uint32)
    sc["hull_count"][hidx] = npts.astype(np.uint32)
    sc["orientation"] = random_quaternions(rng, N)
    offs = np.concatenate(([0], np.cumsum(counts))).astype(np.uint32)
    side = (counts * 8 * 0.6 ** 3 / fill) ** (1.0 / 3.0)
    x0 = np.concatenate(([0.0], np.cumsum(side + gap)[:-1])) if gap is not None else np.zeros(n_worlds)
    wid = np.repeat(np.arange(n_worlds), counts)
    u3 = rng.unit(3 * N).astype(np.float64).reshape(N, 3)
    pos = u3 * side[wid, None]
    pos[:, 0] += x0[wid]
    sc["position"] = pos.astype(np.float32)
    sc["world_offsets"] = offs
    return sc
