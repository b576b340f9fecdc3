"""MI355X-native collision hot path of bullet-physics (broadphase + convex narrowphase).

Host-side mirror of the reference's interfaces, over the C ABI in
include/bullet_collision_hip.h:

* :class:`BroadPhase` -- the ``BroadPhase`` trait
  (crates/bullet-collision/src/broadphase/mod.rs:94-130): proxies with
  lowest-free-slot reuse, ``compute_pairs`` returning the sorted ``(i, j)``
  overlap set.
* :func:`collide` / :func:`collide_seeded` / :func:`collide_batch` --
  ``narrowphase::collide`` / ``collide_seeded`` (narrowphase/mod.rs:25-286):
  ``None`` for no contact, else a :class:`ContactManifold` of at most four
  :class:`ContactPoint` (normal A -> B, position = witness on B,
  penetration > 0 overlapping).
* :class:`CollisionWorld` -- the world's per-step detection pass
  (world/step.rs:349-416, 445-854) with bodies resident on the GPU.

Everything runs through the HIP library; importing works without a GPU, but
any call needs one (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, scenes  # noqa: F401
from ._lib import BchBodies, BchError

SPHERE, BOX, CAPSULE, HULL, PLANE = 0, 1, 2, 3, 4


def _check(ctx, rc):
    if rc != _lib.BCH_OK:
        msg = _lib.load().bch_last_error(ctx).decode() if ctx else ""
        raise BchError(rc, msg)


class _Context:
    """One HIP context (device + stream); one per host thread."""

    def __init__(self, device: int = 0):
        self.L = _lib.load()
        h = C.c_void_p()
        rc = self.L.bch_create(device, C.byref(h))
        if rc != _lib.BCH_OK:
            raise BchError(rc, "bch_create failed (no HIP device?)")
        self.h = h

    _KNOBS = {"hull_lanes": _lib.TUNE_HULL_LANES,
              "traverse_block": _lib.TUNE_TRAVERSE_BLOCK, "epa_grid": _lib.TUNE_EPA_GRID, "np_grid": _lib.TUNE_NP_GRID,
              "epa_wave_faces": _lib.TUNE_EPA_WAVE_FACES, "sort_tile": _lib.TUNE_SORT_TILE}

    def set_tuning(self, **knobs):
        """Launch-shape knobs (hull_lanes, traverse_block, epa_grid,
        np_grid, epa_wave_faces, sort_tile); they never change results."""
        for k, v in knobs.items():
            if k not in self._KNOBS:
                raise KeyError(f"unknown tuning knob {k!r}")
            _check(self.h, self.L.bch_set_tuning(self.h, self._KNOBS[k], int(v)))

    def tuning(self) -> dict:
        out = {}
        for k, knob in self._KNOBS.items():
            v = C.c_int()
            _check(self.h, self.L.bch_get_tuning(self.h, knob, C.byref(v)))
            out[k] = v.value
        return out

    def set_stream(self, stream: int | None) -> None:
        """Order this context's work on a caller stream (hipStream_t as an
        int; 0 is the device's null stream); None restores the context's
        own stream."""
        if stream is None:
            _check(self.h, self.L.bch_reset_stream(self.h))
        else:
            _check(self.h, self.L.bch_set_stream(self.h, C.c_void_p(stream)))

    def close(self):
        if self.h:
            self.L.bch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> _Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = _Context(0)
    return _default_ctx


# ---------------------------------------------------------------- contacts --
@dataclass
class ContactPoint:
    """manifold.rs:12-39 (position, normal A->B, penetration)."""
    position: np.ndarray
    normal: np.ndarray
    penetration: float


@dataclass
class ContactManifold:
    points: list

    def flipped(self) -> "ContactManifold":
        return ContactManifold([ContactPoint(p.position, -p.normal, p.penetration) for p in self.points])


class Bodies:
    """Body arrays kept alive for a C call (see scenes.py for the layout)."""

    def __init__(self, scene: dict):
        self.pos = np.ascontiguousarray(scene["position"], np.float32).reshape(-1, 3)
        n = self.pos.shape[0]
        self.quat = np.ascontiguousarray(scene["orientation"], np.float32).reshape(n, 4)
        self.type = np.ascontiguousarray(scene["shape_type"], np.uint8).reshape(n)
        self.params = np.ascontiguousarray(scene["shape_params"], np.float32).reshape(n, 4)
        self.hull_first = np.ascontiguousarray(scene.get("hull_first", np.zeros(n)), np.uint32).reshape(n)
        self.hull_count = np.ascontiguousarray(scene.get("hull_count", np.zeros(n)), np.uint32).reshape(n)
        self.hull_vertices = np.ascontiguousarray(scene.get("hull_vertices", np.zeros((0, 3))),
                                                  np.float32).reshape(-1, 3)
        b = BchBodies()
        b.n = n
        b.position = self.pos.ctypes.data
        b.orientation = self.quat.ctypes.data
        b.shape_type = self.type.ctypes.data
        b.shape_params = self.params.ctypes.data
        b.hull_first = self.hull_first.ctypes.data
        b.hull_count = self.hull_count.ctypes.data
        b.hull_vertices = self.hull_vertices.ctypes.data if self.hull_vertices.shape[0] else None
        b.n_hull_vertices = self.hull_vertices.shape[0]
        self.c = b


# -------------------------------------------------------------- broadphase --
class BroadPhase:
    """The BroadPhase trait (broadphase/mod.rs:94-130) on the GPU."""

    def __init__(self, ctx: _Context | None = None):
        self.ctx = ctx or _Context(0)
        self.L = self.ctx.L

    def create_proxy(self, aabb_min, aabb_max) -> int:
        mn = np.ascontiguousarray(aabb_min, np.float32)
        mx = np.ascontiguousarray(aabb_max, np.float32)
        out = C.c_uint32()
        _check(self.ctx.h, self.L.bch_bp_create_proxy(self.ctx.h, mn.ctypes.data, mx.ctypes.data, C.byref(out)))
        return out.value

    def set_aabb(self, proxy: int, aabb_min, aabb_max) -> None:
        mn = np.ascontiguousarray(aabb_min, np.float32)
        mx = np.ascontiguousarray(aabb_max, np.float32)
        _check(self.ctx.h, self.L.bch_bp_set_aabb(self.ctx.h, proxy, mn.ctypes.data, mx.ctypes.data))

    def set_aabbs(self, proxies, mins, maxs) -> None:
        p = np.ascontiguousarray(proxies, np.uint32)
        mn = np.ascontiguousarray(mins, np.float32)
        mx = np.ascontiguousarray(maxs, np.float32)
        _check(self.ctx.h, self.L.bch_bp_set_aabbs(self.ctx.h, p.shape[0], p.ctypes.data, mn.ctypes.data,
                                                    mx.ctypes.data))

    def destroy_proxy(self, proxy: int) -> None:
        _check(self.ctx.h, self.L.bch_bp_destroy_proxy(self.ctx.h, proxy))

    def proxy_aabb(self, proxy: int):
        mn = np.zeros(3, np.float32)
        mx = np.zeros(3, np.float32)
        live = C.c_int()
        _check(self.ctx.h, self.L.bch_bp_proxy_aabb(self.ctx.h, proxy, mn.ctypes.data, mx.ctypes.data,
                                                     C.byref(live)))
        return (mn, mx) if live.value else None

    def compute_pairs(self) -> np.ndarray:
        """Sorted (i, j), i < j, overlapping live proxies: uint32[P, 2]."""
        cnt = C.c_uint64()
        rc = self.L.bch_bp_compute_pairs(self.ctx.h, None, 0, C.byref(cnt))
        if rc not in (_lib.BCH_OK, _lib.BCH_OVERFLOW):
            _check(self.ctx.h, rc)
        out = np.zeros((max(cnt.value, 1), 2), np.uint32)
        _check(self.ctx.h, self.L.bch_bp_compute_pairs(self.ctx.h, out.ctypes.data, out.shape[0], C.byref(cnt)))
        return out[:cnt.value]

    def aabb_candidates(self, aabb_min, aabb_max) -> np.ndarray:
        mn = np.ascontiguousarray(aabb_min, np.float32)
        mx = np.ascontiguousarray(aabb_max, np.float32)
        cnt = C.c_uint64()
        rc = self.L.bch_bp_aabb_candidates(self.ctx.h, mn.ctypes.data, mx.ctypes.data, None, 0, C.byref(cnt))
        if rc not in (_lib.BCH_OK, _lib.BCH_OVERFLOW):
            _check(self.ctx.h, rc)
        out = np.zeros(max(cnt.value, 1), np.uint32)
        _check(self.ctx.h, self.L.bch_bp_aabb_candidates(self.ctx.h, mn.ctypes.data, mx.ctypes.data,
                                                          out.ctypes.data, out.shape[0], C.byref(cnt)))
        return out[:cnt.value]

    def raycast_candidates(self, ray_from, ray_to) -> np.ndarray:
        """Live proxies whose AABB the segment crosses (broadphase/mod.rs:119-124)."""
        f = np.ascontiguousarray(ray_from, np.float32)
        t = np.ascontiguousarray(ray_to, np.float32)
        cnt = C.c_uint64()
        rc = self.L.bch_bp_raycast_candidates(self.ctx.h, f.ctypes.data, t.ctypes.data, None, 0, C.byref(cnt))
        if rc not in (_lib.BCH_OK, _lib.BCH_OVERFLOW):
            _check(self.ctx.h, rc)
        out = np.zeros(max(cnt.value, 1), np.uint32)
        _check(self.ctx.h, self.L.bch_bp_raycast_candidates(self.ctx.h, f.ctypes.data, t.ctypes.data,
                                                             out.ctypes.data, out.shape[0], C.byref(cnt)))
        return out[:cnt.value]


# ------------------------------------------------------------- narrowphase --
def collide_batch(scene: dict, pairs, seeds=None, ctx: _Context | None = None):
    """collide_seeded for each (i, j): returns (count u8[P], points f32[P, 4, 7])."""
    ctx = ctx or default_context()
    b = Bodies(scene)
    pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    P = pr.shape[0]
    cnt = np.zeros(max(P, 1), np.uint8)
    pts = np.zeros((max(P, 1), 4, 7), np.float32)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.float32).reshape(P, 3)
    _check(ctx.h, ctx.L.bch_collide_batch(ctx.h, C.byref(b.c), pr.ctypes.data, P,
                                          None if sd is None else sd.ctypes.data, cnt.ctypes.data, pts.ctypes.data))
    return cnt[:P], pts[:P]


def collide_batch_poses(position, orientation, pairs, seeds=None, ctx: _Context | None = None):
    """collide_batch on the shapes of the last collide_batch call of ctx at new
    poses (only the poses are uploaded): (count u8[P], points f32[P, 4, 7])."""
    ctx = ctx or default_context()
    pos = np.ascontiguousarray(position, np.float32).reshape(-1, 3)
    quat = np.ascontiguousarray(orientation, np.float32).reshape(-1, 4)
    pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    P = pr.shape[0]
    cnt = np.zeros(max(P, 1), np.uint8)
    pts = np.zeros((max(P, 1), 4, 7), np.float32)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.float32).reshape(P, 3)
    _check(ctx.h, ctx.L.bch_collide_batch_poses(ctx.h, pos.ctypes.data, quat.ctypes.data, pr.ctypes.data, P,
                                                None if sd is None else sd.ctypes.data, cnt.ctypes.data,
                                                pts.ctypes.data))
    return cnt[:P], pts[:P]


def _shape_scene(shapes_poses):
    """Tiny scene from [(type, params, hull_pts|None, pos, quat), ...]."""
    n = len(shapes_poses)
    sc = scenes._empty(n)
    verts = []
    off = 0
    for i, (t, prm, hp, pos, quat) in enumerate(shapes_poses):
        sc["shape_type"][i] = t
        prm = list(prm) + [0.0] * (4 - len(prm))
        sc["shape_params"][i] = prm[:4]
        sc["position"][i] = pos
        sc["orientation"][i] = quat
        if t == HULL:
            hp = np.asarray(hp, np.float32).reshape(-1, 3)
            sc["hull_first"][i] = off
            sc["hull_count"][i] = hp.shape[0]
            verts.append(hp)
            off += hp.shape[0]
    if verts:
        sc["hull_vertices"] = np.concatenate(verts)
    return sc


def collide_seeded(shape_a, pa, qa, shape_b, pb, qb, seed=None, ctx=None):
    """narrowphase::collide_seeded for one pair; shapes are (type, params[, hull_points])."""
    def unpack(s):
        return (s[0], s[1], s[2] if len(s) > 2 else None)
    ta, pra, ha = unpack(shape_a)
    tb, prb, hb = unpack(shape_b)
    sc = _shape_scene([(ta, pra, ha, pa, qa), (tb, prb, hb, pb, qb)])
    cnt, pts = collide_batch(sc, np.array([[0, 1]], np.uint32),
                             None if seed is None else np.asarray(seed, np.float32).reshape(1, 3), ctx)
    if cnt[0] == 0:
        return None
    return ContactManifold([ContactPoint(pts[0, k, :3].copy(), pts[0, k, 3:6].copy(), float(pts[0, k, 6]))
                            for k in range(cnt[0])])


def collide(shape_a, pa, qa, shape_b, pb, qb, ctx=None):
    """narrowphase::collide (mod.rs:25-47): collide_seeded with seed None."""
    return collide_seeded(shape_a, pa, qa, shape_b, pb, qb, None, ctx)


# ------------------------------------------------------------------- world --
class TriangleMesh:
    """A static triangle mesh on a context (TriangleMesh::new,
    shapes/triangle_mesh.rs:84-123); `id` names it in collide_mesh_batch."""

    def __init__(self, vertices, indices, ctx: _Context | None = None):
        self.ctx = ctx or _Context(0)
        self.v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.i = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        out = C.c_uint32()
        _check(self.ctx.h, self.ctx.L.bch_mesh_add(self.ctx.h, self.v.ctypes.data, self.v.shape[0],
                                                   self.i.ctypes.data, self.i.shape[0], C.byref(out)))
        self.id = out.value


def collide_mesh_batch(scene: dict, pairs, mesh_poses, mesh_first=None, seeds=None, ctx: _Context | None = None):
    """collide_seeded for (convex body, mesh) pairs (narrowphase/mod.rs:177-215):
    pairs [P, 2] = (body of `scene`, TriangleMesh.id), mesh_poses [P, 7];
    returns (count u8[P], points f32[P, 4, 7])."""
    ctx = ctx or _Context(0)
    b = Bodies(scene)
    pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    P = pr.shape[0]
    mp = np.ascontiguousarray(mesh_poses, np.float32).reshape(-1, 7)
    mf = None if mesh_first is None else np.ascontiguousarray(mesh_first, np.uint8)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.float32)
    cnt = np.zeros(max(P, 1), np.uint8)
    pts = np.zeros((max(P, 1), 4, 7), np.float32)
    _check(ctx.h, ctx.L.bch_collide_mesh_batch(ctx.h, C.byref(b.c), pr.ctypes.data, mp.ctypes.data,
                                               None if mf is None else mf.ctypes.data, P,
                                               None if sd is None else sd.ctypes.data, cnt.ctypes.data,
                                               pts.ctypes.data))
    return cnt[:P], pts[:P]


def sweep_batch(scene: dict, target, frm, to, radius=None, caster=None, ctx: _Context | None = None):
    """Sweeps against scene bodies (narrowphase/convex_sweep.rs): query k moves
    from frm[k] to to[k] against body target[k]; a sphere of radius[k]
    (swept_sphere_cast) or, with caster, body caster[k]'s shape at its
    orientation (convex_cast).  Returns (hit u8[n], out f32[n, 7] = fraction,
    normal, point)."""
    ctx = ctx or _Context(0)
    b = Bodies(scene)
    t = np.ascontiguousarray(target, np.uint32).reshape(-1)
    n = t.shape[0]
    f = np.ascontiguousarray(frm, np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(to, np.float32).reshape(-1, 3)
    if f.shape[0] != n or e.shape[0] != n:
        raise ValueError("from/to need one point per query")
    rad = None if radius is None else np.ascontiguousarray(radius, np.float32).reshape(-1)
    cs = None if caster is None else np.ascontiguousarray(caster, np.uint32).reshape(-1)
    if (rad if cs is None else cs) is None or (rad if cs is None else cs).shape[0] != n:
        raise ValueError("need one radius (sphere sweeps) or one caster (convex casts) per query")
    hit = np.zeros(max(n, 1), np.uint8)
    out = np.zeros((max(n, 1), 7), np.float32)
    _check(ctx.h, ctx.L.bch_sweep_batch(ctx.h, C.byref(b.c), t.ctypes.data, None if cs is None else cs.ctypes.data,
                                        f.ctypes.data, e.ctypes.data, None if rad is None else rad.ctypes.data, n,
                                        hit.ctypes.data, out.ctypes.data))
    return hit[:n], out[:n]


def sweep_mesh_batch(mesh_of, mesh_poses, frm, to, radius=None, caster=None, scene: dict | None = None,
                     ctx: _Context | None = None):
    """The mesh arms of the sweeps (sweep_mesh / sweep_convex_mesh): query k
    against TriangleMesh id mesh_of[k] at mesh_poses[k]; a sphere of radius[k]
    or body caster[k] of `scene`.  Returns (hit u8[n], out f32[n, 7])."""
    ctx = ctx or _Context(0)
    mo = np.ascontiguousarray(mesh_of, np.uint32).reshape(-1)
    n = mo.shape[0]
    mp = np.ascontiguousarray(mesh_poses, np.float32).reshape(-1, 7)
    f = np.ascontiguousarray(frm, np.float32).reshape(-1, 3)
    e = np.ascontiguousarray(to, np.float32).reshape(-1, 3)
    if mp.shape[0] != n or f.shape[0] != n or e.shape[0] != n:
        raise ValueError("mesh poses and from/to need one row per query")
    rad = None if radius is None else np.ascontiguousarray(radius, np.float32).reshape(-1)
    cs = None if caster is None else np.ascontiguousarray(caster, np.uint32).reshape(-1)
    if (rad if cs is None else cs) is None or (rad if cs is None else cs).shape[0] != n:
        raise ValueError("need one radius (sphere sweeps) or one caster (convex casts) per query")
    if cs is not None and scene is None:
        raise ValueError("convex casts need the scene of the caster bodies")
    b = None if scene is None else Bodies(scene)
    hit = np.zeros(max(n, 1), np.uint8)
    out = np.zeros((max(n, 1), 7), np.float32)
    _check(ctx.h, ctx.L.bch_sweep_mesh_batch(ctx.h, None if b is None else C.byref(b.c), mo.ctypes.data,
                                             mp.ctypes.data, None if cs is None else cs.ctypes.data, f.ctypes.data,
                                             e.ctypes.data, None if rad is None else rad.ctypes.data, n,
                                             hit.ctypes.data, out.ctypes.data))
    return hit[:n], out[:n]


class CollisionWorld:
    """Bodies resident in HBM; one detection pass per :meth:`collide`."""

    def __init__(self, scene: dict, ctx: _Context | None = None, batched: bool | None = None):
        self.ctx = ctx or _Context(0)
        self.L = self.ctx.L
        self._bodies = Bodies(scene)
        self.n = self._bodies.pos.shape[0]
        wo = scene.get("world_offsets") if batched in (None, True) else None
        if wo is not None:
            self._wo = np.ascontiguousarray(wo, np.uint32)
            _check(self.ctx.h, self.L.bch_world_set_bodies(self.ctx.h, C.byref(self._bodies.c),
                                                            self._wo.shape[0] - 1, self._wo.ctypes.data))
        else:
            _check(self.ctx.h, self.L.bch_world_set_bodies(self.ctx.h, C.byref(self._bodies.c), 1, None))
        self.n_pairs = 0
        self.n_contacts = 0
        self.found = (0, 0)  # counts the last captured-pass replay found
        self.n_mpairs = 0
        self.n_mpoints = 0

    def set_poses(self, position, orientation, linear_velocity=None, dt: float = 0.0) -> None:
        """New poses; with velocities, active bodies go to their predicted
        position + linear_velocity * dt (step.rs:340-343)."""
        p = np.ascontiguousarray(position, np.float32)
        q = np.ascontiguousarray(orientation, np.float32)
        if linear_velocity is None:
            _check(self.ctx.h, self.L.bch_world_set_poses(self.ctx.h, p.ctypes.data, q.ctypes.data))
            return
        v = np.ascontiguousarray(linear_velocity, np.float32)
        _check(self.ctx.h, self.L.bch_world_set_poses_predicted(self.ctx.h, p.ctypes.data, q.ctypes.data,
                                                                v.ctypes.data, dt))

    def set_poses_device(self, pos_ptr: int, quat_ptr: int) -> None:
        _check(self.ctx.h, self.L.bch_world_set_poses_device(self.ctx.h, pos_ptr, quat_ptr))

    def collide(self, aabb_margin: float = 0.02):
        """Run one pass; returns (n_pairs, n_contacts)."""
        P = C.c_uint64()
        Cc = C.c_uint64()
        _check(self.ctx.h, self.L.bch_world_collide(self.ctx.h, aabb_margin, C.byref(P), C.byref(Cc)))
        self.n_pairs, self.n_contacts = P.value, Cc.value
        return self.n_pairs, self.n_contacts

    def results(self):
        """(pairs u32[P, 2], contact_offsets u32[P+1], contacts f32[C, 7]) of the last pass."""
        pairs = np.zeros((max(self.n_pairs, 1), 2), np.uint32)
        offs = np.zeros(self.n_pairs + 1, np.uint32)
        cont = np.zeros((max(self.n_contacts, 1), 7), np.float32)
        _check(self.ctx.h, self.L.bch_world_results(self.ctx.h, pairs.ctypes.data, offs.ctypes.data,
                                                     cont.ctypes.data))
        return pairs[:self.n_pairs], offs, cont[:self.n_contacts]

    def autotune(self, aabb_margin: float = 0.02, trials: int = 3) -> dict:
        """Pick the launch-shape knobs for these bodies (see bch_autotune);
        returns the chosen knobs and the best pass time in ms."""
        ms = C.c_float()
        _check(self.ctx.h, self.L.bch_autotune(self.ctx.h, aabb_margin, trials, C.byref(ms)))
        out = self.ctx.tuning()
        out["pass_ms"] = ms.value
        return out

    # ------------------------------------------ persistent manifolds --
    # body flags and pair routes of the manifold pass
    BODY_STATIC, BODY_ACTIVE, BODY_FAST = 1, 2, 4
    ROUTES = ("filtered", "sleeping", "static", "served", "detected", "removed")

    def set_body_flags(self, flags=None, group_mask=None) -> None:
        """Per-body flags (BODY_STATIC | BODY_ACTIVE | BODY_FAST) and
        collision group | mask << 16; None = active, dynamic, slow bodies
        that collide with everything."""
        self._bflags = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        self._bgm = None if group_mask is None else np.ascontiguousarray(group_mask, np.uint32)
        _check(self.ctx.h, self.L.bch_world_set_body_flags(
            self.ctx.h, None if self._bflags is None else self._bflags.ctypes.data,
            None if self._bgm is None else self._bgm.ctypes.data))

    def manifold_pass(self, aabb_margin: float = 0.02, linear_slop: float = 0.005, warm_start: float = 0.85):
        """The step's pair pass with persistent manifolds (step.rs:660-854)
        at the current poses; returns (n_pairs, n_points)."""
        P, Np = C.c_uint64(), C.c_uint64()
        _check(self.ctx.h, self.L.bch_world_manifold_pass(self.ctx.h, aabb_margin, linear_slop, warm_start,
                                                          C.byref(P), C.byref(Np)))
        self.n_mpairs, self.n_mpoints = P.value, Np.value
        return self.n_mpairs, self.n_mpoints

    def manifold_results(self):
        """(pairs u32[P, 2], route u8[P], point_offsets u32[P+1], points
        f32[Np, 12]) of the last manifold pass; a point is position on B,
        normal A->B, penetration, warm-started impulse, lifetime (u32 bits),
        A-side anchor."""
        P, Np = self.n_mpairs, self.n_mpoints
        pairs = np.zeros((max(P, 1), 2), np.uint32)
        route = np.zeros(max(P, 1), np.uint8)
        offs = np.zeros(P + 1, np.uint32)
        pts = np.zeros((max(Np, 1), 12), np.float32)
        _check(self.ctx.h, self.L.bch_world_manifold_results(self.ctx.h, pairs.ctypes.data, route.ctypes.data,
                                                             offs.ctypes.data, pts.ctypes.data))
        return pairs[:P], route[:P], offs, pts[:Np]

    def manifold_impulses(self, impulses) -> None:
        """Post-solve writeback of the last pass's point impulses."""
        imp = np.ascontiguousarray(impulses, np.float32)
        if imp.shape[0] != self.n_mpoints:
            raise ValueError(f"{imp.shape[0]} impulses for {self.n_mpoints} points")
        _check(self.ctx.h, self.L.bch_world_manifold_impulses(self.ctx.h, imp.ctypes.data))

    def manifold_cache(self):
        """(valid u8[P], entries f32[P, 64]) after the last manifold pass."""
        P = self.n_mpairs
        valid = np.zeros(max(P, 1), np.uint8)
        ent = np.zeros((max(P, 1), 64), np.float32)
        _check(self.ctx.h, self.L.bch_world_manifold_cache(self.ctx.h, valid.ctypes.data, ent.ctypes.data))
        return valid[:P], ent[:P]

    def manifold_reset(self) -> None:
        _check(self.ctx.h, self.L.bch_world_manifold_reset(self.ctx.h))

    # ------------------------------------------------ captured pass --
    def capture(self, aabb_margin: float = 0.02, pair_capacity: int | None = None,
                contact_capacity: int | None = None, headroom: float = 1.5, attempts: int = 4):
        """Capture the world pass in a HIP graph (bch_world_capture).  Missing
        capacities come from an eager pass at the current poses times
        `headroom`; when the capture's own pass overflows, the capacities
        grow to fit and the capture is repeated.  Returns (pair_capacity,
        contact_capacity) and leaves the capture's counts in n_pairs /
        n_contacts."""
        if pair_capacity is None or contact_capacity is None:
            P0, C0 = self.collide(aabb_margin)
            pair_capacity = pair_capacity or max(1024, int(P0 * headroom) + 1024)
            contact_capacity = contact_capacity or max(1024, int(C0 * headroom) + 1024)
        for _ in range(attempts):
            rc = self.L.bch_world_capture(self.ctx.h, aabb_margin, int(pair_capacity), int(contact_capacity))
            if rc == _lib.BCH_OVERFLOW:
                try:
                    self._read_counts()
                except BchError:
                    pass
                P, Cn = self.found
                pair_capacity = max(int(pair_capacity), int(P * headroom) + 1024)
                contact_capacity = max(int(contact_capacity), int(Cn * headroom) + 1024)
                continue
            _check(self.ctx.h, rc)
            self._caps = (int(pair_capacity), int(contact_capacity))
            self._read_counts()
            return self._caps
        raise BchError(_lib.BCH_OVERFLOW, "capture kept overflowing")

    def replay(self) -> None:
        """Enqueue the captured pass (one graph launch, no synchronisation)."""
        _check(self.ctx.h, self.L.bch_world_replay(self.ctx.h))

    def _read_counts(self):
        P, Cn, FP, FC = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        rc = self.L.bch_world_replay_counts(self.ctx.h, C.byref(P), C.byref(Cn), C.byref(FP), C.byref(FC))
        # held counts size results(); set them before any error is raised
        self.n_pairs, self.n_contacts = P.value, Cn.value
        self.found = (FP.value, FC.value)
        _check(self.ctx.h, rc)
        return self.n_pairs, self.n_contacts

    def replay_counts(self):
        """Wait for the last replay; (n_pairs, n_contacts) held.  Raises
        BchError with code BCH_OVERFLOW when a capacity was exceeded
        (n_pairs / n_contacts then hold what the buffers kept, `found` what
        the pass found)."""
        return self._read_counts()

    def release_graph(self) -> None:
        _check(self.ctx.h, self.L.bch_world_release_graph(self.ctx.h))

    def results_device(self):
        """Device pointers (pairs u32[2P], contact_offsets u32[P+1],
        contacts f32[7C]) of the last pass, valid until the next pass."""
        a, b, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(self.ctx.h, self.L.bch_world_results_device(self.ctx.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value or 0, b.value or 0, c.value or 0

    def sd_forward(self, d_pairs: int, n_pairs: int, d_out: int, d_status: int | None = None) -> None:
        """Signed distance of device pairs [2P] into device out [10P] (sd,
        normal A->B, witness on A, witness on B) at the current poses."""
        _check(self.ctx.h, self.L.bch_sd_forward(self.ctx.h, d_pairs, n_pairs, d_out, d_status))

    def sd_backward(self, d_pairs: int, n_pairs: int, d_grad_out: int, d_grad_pose: int,
                    d_status: int | None = None) -> None:
        """Adjoint: dL/d(out) [10P] -> dL/d(pose A, pose B) [14P] per pair."""
        _check(self.ctx.h, self.L.bch_sd_backward(self.ctx.h, d_pairs, n_pairs, d_grad_out, d_grad_pose, d_status))

    # stage names of timings(), in order
    STAGES = ("aabb", "tree_build", "tree_traverse", "pair_list", "narrow_fork", "narrow_gjk", "epa",
              "contact_compaction")

    def routes(self):
        """Dispatch route of every pair of the last pass (u8[P]): 0 none,
        1 sphere-sphere, 2 sphere-box, 3 plane-convex, 4 box-box, 5 GJK
        without a hull, 6 GJK with a hull; +8 when EPA ran."""
        out = np.zeros(max(self.n_pairs, 1), np.uint8)
        _check(self.ctx.h, self.L.bch_world_routes(self.ctx.h, out.ctypes.data))
        return out[:self.n_pairs]

    def timings(self):
        """Device milliseconds of the last pass per stage (see STAGES)."""
        return self.timings_all()[:8]

    def timings_all(self):
        """The eight stage times, then the four-pair EPA kernel alone (part
        of "epa") and the hull GJK kernel alone (part of "narrow_gjk")."""
        out = np.zeros(10, np.float32)
        _check(self.ctx.h, self.L.bch_last_timings(self.ctx.h, out.ctypes.data))
        return out

    def epa_routes(self):
        """EPA entries of the last pass: (queued for the four-pair kernel,
        handed to the one-pair kernel, rerun with the 2048-face store)."""
        out = np.zeros(3, np.uint32)
        _check(self.ctx.h, self.L.bch_last_epa_routes(self.ctx.h, out.ctypes.data))
        return tuple(int(x) for x in out)

    def synchronize(self):
        _check(self.ctx.h, self.L.bch_synchronize(self.ctx.h))
