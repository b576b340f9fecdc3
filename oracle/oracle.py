"""ctypes bridge to the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is a C++ restatement of the reference collision path
(crates/bullet-collision, see oracle.hpp for the per-function citations).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load
it; the shipped library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BCH_ORACLE_LIB names another build of the same sources (tests/plant_oracle.py
# points it at copies with one planted mistake each).
_LIB_PATH = os.environ.get("BCH_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

SPHERE, BOX, CAPSULE, HULL, PLANE = 0, 1, 2, 3, 4


class OrShape(C.Structure):
    _fields_ = [("type", C.c_int32), ("p", C.c_float * 4),
                ("hull", C.POINTER(C.c_float)), ("nhull", C.c_int32)]


class OrScene(C.Structure):
    _fields_ = [("n", C.c_int64), ("pos", C.c_void_p), ("quat", C.c_void_p),
                ("type", C.c_void_p), ("params", C.c_void_p),
                ("hull_first", C.c_void_p), ("hull_count", C.c_void_p),
                ("hull_verts", C.c_void_p), ("n_hull_verts", C.c_int64)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        f32p, f64p = C.POINTER(C.c_float), C.POINTER(C.c_double)
        sp = C.POINTER(OrShape)
        L.oracle_threshold_f32.restype = C.c_float
        L.oracle_threshold_f32.argtypes = [sp]
        L.oracle_world_aabb_f32.argtypes = [sp, f32p, f32p, f32p]
        L.oracle_support_world_f32.argtypes = [sp, f32p, f32p, f32p, C.c_int, f32p]
        L.oracle_collide_f32.argtypes = [sp, f32p, f32p, sp, f32p, f32p, f32p, f32p, C.POINTER(C.c_int)]
        L.oracle_collide_f64.argtypes = [sp, f64p, f64p, sp, f64p, f64p, f64p, f64p, C.POINTER(C.c_int)]
        L.oracle_convex_convex_f32.argtypes = [sp, f32p, f32p, sp, f32p, f32p, f32p, C.c_float, f32p,
                                               C.POINTER(C.c_int)]
        L.oracle_gjk_contact_f32.argtypes = [sp, f32p, f32p, sp, f32p, f32p, f32p, C.c_float, f32p, f32p]
        L.oracle_gjk_contact_f64.argtypes = [sp, f64p, f64p, sp, f64p, f64p, f64p, C.c_double, f64p]
        L.oracle_gjk_intersect_f32.argtypes = [sp, f32p, f32p, sp, f32p, f32p, f32p]
        L.oracle_epa_pair_f32.argtypes = [f32p, sp, f32p, f32p, sp, f32p, f32p, C.c_int, f32p]
        L.oracle_epa_points_f32.argtypes = [f32p, f32p, C.c_int, f32p]
        i32p = C.POINTER(C.c_int)
        L.oracle_closest_origin_f32.argtypes = [f32p, C.c_int, f32p, f32p]
        L.oracle_reduce_f32.argtypes = [f32p, f32p, C.c_int, f32p, f32p, i32p]
        L.oracle_perpendicular_f32.argtypes = [f32p, f32p]
        L.oracle_solve_f32.argtypes = [f32p, C.c_int, f32p, i32p]
        L.oracle_enclosing_walk_scripted_f32.argtypes = [f32p, C.c_int, f32p, f32p, C.c_int, i32p]
        L.oracle_epa_find_best_f32.argtypes = [f32p, C.c_int]
        L.oracle_epa_expand_f32.argtypes = [i32p, C.c_int, f32p, C.c_int, i32p, i32p]
        L.oracle_cull_points_f32.argtypes = [f32p, f32p, C.c_int, C.c_int, f32p, f32p]
        L.oracle_epa_points_trace_f32.argtypes = [f32p, f32p, C.c_int, f32p, f32p, C.c_int, i32p, i32p, i32p]
        L.oracle_epa_face_distance_f32.restype = C.c_float
        L.oracle_epa_pair_trace_f32.argtypes = [f32p, sp, f32p, f32p, sp, f32p, f32p, C.c_int, f32p, i32p, i32p]
        L.oracle_epa_face_distance_f32.argtypes = [f32p, f32p, f32p]
        scp = C.POINTER(OrScene)
        L.oracle_scene_aabbs_f32.argtypes = [scp, C.c_float, C.c_void_p, C.c_void_p]
        L.oracle_scene_thresholds_f32.argtypes = [scp, C.c_void_p]
        L.oracle_scene_aabbs_predicted_f32.argtypes = [scp, C.c_void_p, C.c_float] + [C.c_void_p] * 3 + [
            C.c_float] + [C.c_void_p] * 3
        L.oracle_sd_pairs_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.oracle_sd_pairs_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_sd_f64.argtypes = [C.c_void_p] * 7
        L.oracle_sd_vjp_pairs_f64.restype = C.c_int64
        L.oracle_sd_vjp_pairs_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.oracle_sd_f32.argtypes = [C.c_void_p] * 7
        L.oracle_raycast_candidates.restype = C.c_int64
        L.oracle_raycast_candidates.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_int64]
        L.oracle_brute_pairs.restype = C.c_int64
        L.oracle_brute_pairs.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_sweep_pairs.restype = C.c_int64
        L.oracle_sweep_pairs.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_gjk_exit_f32.argtypes = [sp, f32p, f32p, sp, f32p, f32p, C.c_float, C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]
        L.oracle_dbvt_pairs.restype = C.c_int64
        L.oracle_dbvt_pairs.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_collide_pairs_f32.argtypes = [scp, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p]
        L.oracle_collide_pairs_f64.argtypes = [scp, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
        L.oracle_pm_capture.argtypes = [C.c_void_p, C.c_void_p, C.c_int] + [C.c_void_p] * 4 + [C.c_float]
        L.oracle_pm_refresh.restype = C.c_int
        L.oracle_pm_refresh.argtypes = [C.c_void_p] * 6
        L.oracle_pm_needs_redetect.restype = C.c_int
        L.oracle_pm_needs_redetect.argtypes = [C.c_void_p] * 5
        L.oracle_pm_seed.restype = C.c_int
        L.oracle_pm_seed.argtypes = [C.c_void_p] * 2
        L.oracle_pm_pass.argtypes = ([scp, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_float, C.c_float]
                                     + [C.c_void_p] * 3 + [C.c_int64] + [C.c_void_p] * 5)
        L.oracle_sphere_triangle.restype = C.c_int
        L.oracle_sphere_triangle.argtypes = [C.c_void_p, C.c_float, C.c_float] + [C.c_void_p] * 4
        L.oracle_reduce.restype = C.c_int
        L.oracle_reduce.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oracle_sweep_sphere.restype = C.c_int
        L.oracle_sweep_sphere.argtypes = [C.c_void_p] * 5 + [C.c_float, C.c_void_p]
        L.oracle_sweep_sphere_mesh.restype = C.c_int
        L.oracle_sweep_sphere_mesh.argtypes = [C.c_void_p] * 5 + [C.c_float, C.c_void_p]
        L.oracle_convex_cast.restype = C.c_int
        L.oracle_convex_cast.argtypes = [C.c_void_p] * 8
        L.oracle_sweep_scene.argtypes = [scp] + [C.c_void_p] * 5 + [C.c_int64, C.c_void_p, C.c_void_p]
        L.oracle_sweep_mesh_scene.argtypes = [C.c_void_p] * 8 + [C.c_int64, C.c_void_p, C.c_void_p]
        L.oracle_mesh_create.restype = C.c_void_p
        L.oracle_mesh_create.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32]
        L.oracle_mesh_destroy.argtypes = [C.c_void_p]
        L.oracle_mesh_threshold.restype = C.c_float
        L.oracle_mesh_threshold.argtypes = [C.c_void_p]
        L.oracle_mesh_query.restype = C.c_int64
        L.oracle_mesh_query.argtypes = [C.c_void_p] * 4 + [C.c_int64]
        L.oracle_mesh_collide.restype = C.c_int
        L.oracle_mesh_collide.argtypes = [C.c_void_p] * 7 + [C.c_int, C.c_void_p]
        L.oracle_mesh_collide_scene.argtypes = [scp, C.c_void_p, C.c_void_p, C.c_int64] + [C.c_void_p] * 5
        L.oracle_pass_f32.restype = C.c_int64
        L.oracle_pass_f32.argtypes = [scp, C.c_float, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.oracle_pass_f32_mt.restype = C.c_int64
        L.oracle_pass_f32_mt.argtypes = [scp, C.c_float, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        _lib = L
    return _lib


# ------------------------------------------------------------------ shapes --
class Shape:
    """Single-shape descriptor for pair-level oracle calls."""

    def __init__(self, type_: int, params=(0, 0, 0, 0), hull=None):
        self.type = type_
        self.params = list(params) + [0.0] * (4 - len(params))
        self.hull = None if hull is None else np.ascontiguousarray(hull, dtype=np.float32).reshape(-1, 3)

    def c(self) -> OrShape:
        s = OrShape()
        s.type = self.type
        for i in range(4):
            s.p[i] = float(self.params[i])
        if self.hull is not None:
            s.hull = self.hull.ctypes.data_as(C.POINTER(C.c_float))
            s.nhull = self.hull.shape[0]
        else:
            s.nhull = 0
        return s


def sphere(r):
    return Shape(SPHERE, (r,))


def box(hx, hy=None, hz=None):
    if hy is None:
        hy = hz = hx
    return Shape(BOX, (hx, hy, hz))


def capsule(r, hh):
    return Shape(CAPSULE, (r, hh))


def plane(n, c):
    return Shape(PLANE, (n[0], n[1], n[2], c))


def hull(points):
    return Shape(HULL, (), points)


IDENT = (0.0, 0.0, 0.0, 1.0)


def _f32(v, n):
    a = (C.c_float * n)()
    for i in range(n):
        a[i] = float(v[i])
    return a


def _f64(v, n):
    a = (C.c_double * n)()
    for i in range(n):
        a[i] = float(v[i])
    return a


def threshold(s: Shape) -> float:
    sc = s.c()
    return lib().oracle_threshold_f32(C.byref(sc))


def world_aabb(s: Shape, pos, quat=IDENT):
    sc = s.c()
    out = (C.c_float * 6)()
    lib().oracle_world_aabb_f32(C.byref(sc), _f32(pos, 3), _f32(quat, 4), out)
    return np.array(out[:3], np.float32), np.array(out[3:], np.float32)


def support_world(s: Shape, pos, quat, d, core=False):
    sc = s.c()
    out = (C.c_float * 3)()
    lib().oracle_support_world_f32(C.byref(sc), _f32(pos, 3), _f32(quat, 4), _f32(d, 3), int(core), out)
    return np.array(out[:], np.float32)


def collide(a: Shape, pa, qa, b: Shape, pb, qb, seed=None, f64=False):
    """collide_seeded (narrowphase/mod.rs:63-286) -> (points[k,7], path)."""
    ac, bc = a.c(), b.c()
    path = C.c_int(0)
    if f64:
        out = (C.c_double * 28)()
        n = lib().oracle_collide_f64(C.byref(ac), _f64(pa, 3), _f64(qa, 4), C.byref(bc), _f64(pb, 3),
                                     _f64(qb, 4), None if seed is None else _f64(seed, 3), out, C.byref(path))
        return np.array(out[:], np.float64).reshape(4, 7)[:n], path.value
    out = (C.c_float * 28)()
    n = lib().oracle_collide_f32(C.byref(ac), _f32(pa, 3), _f32(qa, 4), C.byref(bc), _f32(pb, 3), _f32(qb, 4),
                                 None if seed is None else _f32(seed, 3), out, C.byref(path))
    return np.array(out[:], np.float32).reshape(4, 7)[:n], path.value


def convex_convex(a: Shape, pa, qa, b: Shape, pb, qb, seed=None, threshold=-1.0):
    """convex_convex (convex_convex.rs:30-155) -> (points[k,7], path)."""
    ac, bc = a.c(), b.c()
    out = (C.c_float * 28)()
    path = C.c_int(0)
    n = lib().oracle_convex_convex_f32(C.byref(ac), _f32(pa, 3), _f32(qa, 4), C.byref(bc), _f32(pb, 3),
                                       _f32(qb, 4), None if seed is None else _f32(seed, 3), threshold, out,
                                       C.byref(path))
    return np.array(out[:], np.float32).reshape(4, 7)[:n], path.value


def gjk_contact(a: Shape, pa, qa, b: Shape, pb, qb, threshold, seed=None, f64=False):
    """gjk_contact (gjk.rs:171-288) -> (kind, distance, normal, point_on_b, simplex)."""
    ac, bc = a.c(), b.c()
    if f64:
        out = (C.c_double * 7)()
        k = lib().oracle_gjk_contact_f64(C.byref(ac), _f64(pa, 3), _f64(qa, 4), C.byref(bc), _f64(pb, 3),
                                         _f64(qb, 4), None if seed is None else _f64(seed, 3), threshold, out)
        o = np.array(out[:], np.float64)
        return k, o[0], o[1:4], o[4:7], None
    out = (C.c_float * 7)()
    simp = (C.c_float * 12)()
    k = lib().oracle_gjk_contact_f32(C.byref(ac), _f32(pa, 3), _f32(qa, 4), C.byref(bc), _f32(pb, 3), _f32(qb, 4),
                                     None if seed is None else _f32(seed, 3), threshold, out, simp)
    o = np.array(out[:], np.float32)
    return k, o[0], o[1:4], o[4:7], np.array(simp[:], np.float32).reshape(4, 3)


def gjk_exit(a: Shape, pa, qa, b: Shape, pb, qb, threshold):
    """How the f32 gjk_contact walk ended -> (kind, iterations, capped)."""
    ac, bc = a.c(), b.c()
    it, cap = C.c_int(0), C.c_int(0)
    k = lib().oracle_gjk_exit_f32(C.byref(ac), _f32(pa, 3), _f32(qa, 4), C.byref(bc), _f32(pb, 3), _f32(qb, 4),
                                  threshold, C.byref(it), C.byref(cap))
    return k, it.value, bool(cap.value)


def gjk_intersect(a: Shape, pa, qa, b: Shape, pb, qb):
    ac, bc = a.c(), b.c()
    simp = (C.c_float * 12)()
    n = lib().oracle_gjk_intersect_f32(C.byref(ac), _f32(pa, 3), _f32(qa, 4), C.byref(bc), _f32(pb, 3),
                                       _f32(qb, 4), simp)
    return n, np.array(simp[:], np.float32).reshape(4, 3)


def epa_pair(simplex, a: Shape, pa, qa, b: Shape, pb, qb, core=False):
    ac, bc = a.c(), b.c()
    out = (C.c_float * 4)()
    st = lib().oracle_epa_pair_f32(_f32(np.asarray(simplex).ravel(), 12), C.byref(ac), _f32(pa, 3), _f32(qa, 4),
                                   C.byref(bc), _f32(pb, 3), _f32(qb, 4), int(core), out)
    return st, np.array(out[:3], np.float32), float(out[3])


def epa_pair_trace(simplex, a: Shape, pa, qa, b: Shape, pb, qb, core=False):
    """epa_pair -> (status, normal, depth, vertices, iterations)."""
    ac, bc = a.c(), b.c()
    out = (C.c_float * 4)()
    nv, it = C.c_int(0), C.c_int(0)
    st = lib().oracle_epa_pair_trace_f32(_f32(np.asarray(simplex).ravel(), 12), C.byref(ac), _f32(pa, 3),
                                         _f32(qa, 4), C.byref(bc), _f32(pb, 3), _f32(qb, 4), int(core), out,
                                         C.byref(nv), C.byref(it))
    return st, np.array(out[:3], np.float32), float(out[3]), nv.value, it.value


def epa_points(simplex, points):
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = (C.c_float * 4)()
    st = lib().oracle_epa_points_f32(_f32(np.asarray(simplex).ravel(), 12),
                                     pts.ctypes.data_as(C.POINTER(C.c_float)), pts.shape[0], out)
    return st, np.array(out[:3], np.float32), float(out[3])


# ------------------------------------------------------------- helpers --
def _fa(a, n=None):
    a = np.ascontiguousarray(a, np.float32).ravel()
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def closest_origin(points):
    """closest_origin (gjk.rs:864-1043) -> (q, kept points in keep order)."""
    pts, pp = _fa(points)
    q = (C.c_float * 3)()
    keep = (C.c_float * 12)()
    k = lib().oracle_closest_origin_f32(pp, pts.size // 3, q, keep)
    return np.array(q[:], np.float32), np.array(keep[:3 * k], np.float32).reshape(k, 3)


def reduce_support(w, p_b):
    """The contact walk's reduce (gjk.rs:352-583) -> (inside, v, p_b, len)."""
    wa, wp = _fa(w)
    ba, bp = _fa(p_b)
    v, pb = (C.c_float * 3)(), (C.c_float * 3)()
    n = C.c_int(0)
    inside = lib().oracle_reduce_f32(wp, bp, wa.size // 3, v, pb, C.byref(n))
    return bool(inside), np.array(v[:], np.float32), np.array(pb[:], np.float32), n.value


def perpendicular(v):
    out = (C.c_float * 3)()
    lib().oracle_perpendicular_f32(_f32(v, 3), out)
    return np.array(out[:], np.float32)


def solve(points):
    """solve (gjk.rs:609-620) on Simplex::new(points[0]) + add(points[1:]) -> (enclosed, dir, len)."""
    pa, pp = _fa(points)
    d = (C.c_float * 3)()
    n = C.c_int(0)
    r = lib().oracle_solve_f32(pp, pa.size // 3, d, C.byref(n))
    return bool(r), np.array(d[:], np.float32), n.value


def enclosing_walk_scripted(script, fallback, max_q=512):
    """enclosing_walk (gjk.rs:299-342) whose k-th support call returns
    script[min(k, len - 1)] -> (found, queried directions)."""
    sa, sp = _fa(script)
    q = (C.c_float * (3 * max_q))()
    nq = C.c_int(0)
    ok = lib().oracle_enclosing_walk_scripted_f32(sp, sa.size // 3, _f32(fallback, 3), q, max_q, C.byref(nq))
    n = min(nq.value, max_q)
    return bool(ok), np.array(q[:3 * n], np.float32).reshape(n, 3)


def epa_points_trace(simplex, points, max_q=1024):
    """EPA against a point set's support, recording the support queries ->
    (status, normal, depth, queried directions, vertices, iterations)."""
    pa, pp = _fa(points)
    out = (C.c_float * 4)()
    q = (C.c_float * (3 * max_q))()
    nq, nv, it = C.c_int(0), C.c_int(0), C.c_int(0)
    st = lib().oracle_epa_points_trace_f32(_f32(np.asarray(simplex).ravel(), 12), pp, pa.size // 3, out, q, max_q,
                                           C.byref(nq), C.byref(nv), C.byref(it))
    n = min(nq.value, max_q)
    return (st, np.array(out[:3], np.float32), float(out[3]), np.array(q[:3 * n], np.float32).reshape(n, 3),
            nv.value, it.value)


def epa_face_distance(a, b, c):
    return lib().oracle_epa_face_distance_f32(_f32(a, 3), _f32(b, 3), _f32(c, 3))


def epa_find_best(distances):
    da, dp = _fa(distances)
    return lib().oracle_epa_find_best_f32(dp, da.size)


def epa_expand(faces, verts):
    """expand (epa.rs:309-388) toward verts[-1] -> (accepted, faces after)."""
    f = np.ascontiguousarray(faces, np.int32).ravel()
    va, vp = _fa(verts)
    out = (C.c_int * (3 * (f.size // 3 + 64)))()
    nf = C.c_int(0)
    ok = lib().oracle_epa_expand_f32(f.ctypes.data_as(C.POINTER(C.c_int)), f.size // 3, vp, va.size // 3, out,
                                     C.byref(nf))
    return bool(ok), np.array(out[:3 * nf.value], np.int32).reshape(-1, 3)


def cull_points(points, depths, maxc=4):
    """cull_points (box_box.rs:131-174) -> (points, depths) kept."""
    pa, pp = _fa(points)
    da, dp = _fa(depths)
    op, od = (C.c_float * 12)(), (C.c_float * 4)()
    m = lib().oracle_cull_points_f32(pp, dp, da.size, maxc, op, od)
    return np.array(op[:3 * m], np.float32).reshape(m, 3), np.array(od[:m], np.float32)


# ------------------------------------------------------------------ scenes --
class _SceneRef:
    """Keeps the numpy arrays of a scene alive behind an OrScene struct."""

    def __init__(self, scene):
        self.arrs = dict(
            pos=np.ascontiguousarray(scene["position"], np.float32),
            quat=np.ascontiguousarray(scene["orientation"], np.float32),
            type=np.ascontiguousarray(scene["shape_type"], np.uint8),
            params=np.ascontiguousarray(scene["shape_params"], np.float32),
            hull_first=np.ascontiguousarray(scene["hull_first"], np.uint32),
            hull_count=np.ascontiguousarray(scene["hull_count"], np.uint32),
            hull_verts=np.ascontiguousarray(scene["hull_vertices"], np.float32).reshape(-1, 3),
        )
        s = OrScene()
        s.n = self.arrs["type"].shape[0]
        for k in ("pos", "quat", "type", "params", "hull_first", "hull_count", "hull_verts"):
            setattr(s, k, self.arrs[k].ctypes.data)
        s.n_hull_verts = self.arrs["hull_verts"].shape[0]
        self.c = s


def scene_aabbs(scene, margin):
    r = _SceneRef(scene)
    n = r.c.n
    mn = np.empty((n, 3), np.float32)
    mx = np.empty((n, 3), np.float32)
    lib().oracle_scene_aabbs_f32(C.byref(r.c), margin, mn.ctypes.data, mx.ctypes.data)
    return mn, mx


def scene_aabbs_predicted(scene, vel, dt, margin, flags=None, prev=None):
    """The step's AABB update (step.rs:340-363) -> (predicted positions, min, max)."""
    r = _SceneRef(scene)
    n = r.c.n
    v = np.ascontiguousarray(vel, np.float32)
    fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    pm = px = None
    if prev is not None:
        pm, px = (np.ascontiguousarray(prev[0], np.float32), np.ascontiguousarray(prev[1], np.float32))
    pos = np.empty((n, 3), np.float32)
    mn = np.empty((n, 3), np.float32)
    mx = np.empty((n, 3), np.float32)
    lib().oracle_scene_aabbs_predicted_f32(C.byref(r.c), v.ctypes.data, dt, None if fl is None else fl.ctypes.data,
                                           None if pm is None else pm.ctypes.data,
                                           None if px is None else px.ctypes.data, margin, pos.ctypes.data,
                                           mn.ctypes.data, mx.ctypes.data)
    return pos, mn, mx


def scene_thresholds(scene):
    r = _SceneRef(scene)
    out = np.empty(r.c.n, np.float32)
    lib().oracle_scene_thresholds_f32(C.byref(r.c), out.ctypes.data)
    return out


def brute_pairs(mn, mx):
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    n = mn.shape[0]
    cnt = lib().oracle_brute_pairs(n, mn.ctypes.data, mx.ctypes.data, None, 0)
    out = np.empty((max(cnt, 1), 2), np.uint32)
    lib().oracle_brute_pairs(n, mn.ctypes.data, mx.ctypes.data, out.ctypes.data, cnt)
    return out[:cnt]


def raycast_candidates(mn, mx, ray_from, ray_to):
    """aabb_raycast_candidates (broadphase/mod.rs:185-200): ascending ids."""
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    f = np.ascontiguousarray(ray_from, np.float32)
    t = np.ascontiguousarray(ray_to, np.float32)
    n = mn.shape[0]
    out = np.zeros(max(n, 1), np.uint32)
    cnt = lib().oracle_raycast_candidates(n, mn.ctypes.data, mx.ctypes.data, f.ctypes.data, t.ctypes.data,
                                          out.ctypes.data, out.shape[0])
    return out[:cnt]


def sweep_pairs(mn, mx):
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    n = mn.shape[0]
    cnt = lib().oracle_sweep_pairs(n, mn.ctypes.data, mx.ctypes.data, None, 0)
    out = np.empty((max(cnt, 1), 2), np.uint32)
    lib().oracle_sweep_pairs(n, mn.ctypes.data, mx.ctypes.data, out.ctypes.data, cnt)
    return out[:cnt]


def world_pairs(scene, margin):
    """Sorted pair set of a batch of independent worlds (``world_offsets``),
    world by world, in the batch's body indices: worlds may share space."""
    wo = np.asarray(scene["world_offsets"], np.int64)
    mn, mx = scene_aabbs(scene, margin)
    out = [sweep_pairs(mn[wo[w]:wo[w + 1]], mx[wo[w]:wo[w + 1]]).astype(np.int64) + wo[w]
           for w in range(wo.shape[0] - 1)]
    return np.concatenate(out).astype(np.uint32) if out else np.zeros((0, 2), np.uint32)


def dbvt_pairs(mn, mx):
    """DbvtBroadPhase's pair set on a fresh scene (dbvt.rs), sorted."""
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    n = mn.shape[0]
    cnt = lib().oracle_dbvt_pairs(n, mn.ctypes.data, mx.ctypes.data, None, 0)
    out = np.empty((max(cnt, 1), 2), np.uint32)
    lib().oracle_dbvt_pairs(n, mn.ctypes.data, mx.ctypes.data, out.ctypes.data, cnt)
    return out[:cnt]


def collide_pairs(scene, pairs, seeds=None, f64=False):
    """collide_seeded(body i, body j) per pair -> (count[P], points[P,4,7], path[P])."""
    r = _SceneRef(scene)
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    p = pairs.shape[0]
    cnt = np.zeros(p, np.uint8)
    if f64:
        pts = np.zeros((p, 4, 7), np.float64)
        lib().oracle_collide_pairs_f64(C.byref(r.c), pairs.ctypes.data, p, cnt.ctypes.data, pts.ctypes.data)
        return cnt, pts, None
    pts = np.zeros((p, 4, 7), np.float32)
    path = np.zeros(p, np.int8)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.float32)
    lib().oracle_collide_pairs_f32(C.byref(r.c), pairs.ctypes.data, p, None if sd is None else sd.ctypes.data,
                                   cnt.ctypes.data, pts.ctypes.data, path.ctypes.data)
    return cnt, pts, path


SD_WALK, SD_EPA, SD_CAPPED, SD_TOUCH, SD_EPA_UNCONVERGED, SD_SEPARATED, SD_EPA_CONTRADICTED = 0, 1, 2, 3, 4, 5, 6


def signed_distance(a: Shape, pa, qa, b: Shape, pb, qb, f64=False):
    """signed_distance (oracle.hpp) of one pair -> (out[10] = sd, n3, pa3, pb3; status)."""
    ac, bc = a.c(), b.c()
    if f64:
        out = (C.c_double * 10)()
        st = lib().oracle_sd_f64(C.byref(ac), _f64(pa, 3), _f64(qa, 4), C.byref(bc), _f64(pb, 3), _f64(qb, 4), out)
        return np.array(out[:], np.float64), st
    out = (C.c_float * 10)()
    st = lib().oracle_sd_f32(C.byref(ac), _f32(pa, 3), _f32(qa, 4), C.byref(bc), _f32(pb, 3), _f32(qb, 4), out)
    return np.array(out[:], np.float32), st


def sd_pairs(scene, pairs, pose14=None, f64=False):
    """signed_distance over scene pairs -> (out[P, 10], status[P]).  pose14
    (float64 [P, 14]: pos_a, quat_a, pos_b, quat_b) overrides the scene
    poses (f64 only)."""
    r = _SceneRef(scene)
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    p = pairs.shape[0]
    st = np.zeros(p, np.int8)
    if f64:
        out = np.zeros((p, 10), np.float64)
        ps = None if pose14 is None else np.ascontiguousarray(pose14, np.float64)
        lib().oracle_sd_pairs_f64(C.byref(r.c), pairs.ctypes.data, p, None if ps is None else ps.ctypes.data,
                                  out.ctypes.data, st.ctypes.data)
        return out, st
    assert pose14 is None
    out = np.zeros((p, 10), np.float32)
    lib().oracle_sd_pairs_f32(C.byref(r.c), pairs.ctypes.data, p, out.ctypes.data, st.ctypes.data)
    return out, st


def sd_vjp_pairs(scene, pairs, g10):
    """f64 reference adjoint of sd_pairs at the scene poses: g10 [P, 10] ->
    (g14 [P, 14], number of EPA faces not split)."""
    r = _SceneRef(scene)
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    g = np.ascontiguousarray(g10, np.float64)
    out = np.zeros((pairs.shape[0], 14), np.float64)
    fails = lib().oracle_sd_vjp_pairs_f64(C.byref(r.c), pairs.ctypes.data, pairs.shape[0], g.ctypes.data,
                                          out.ctypes.data)
    return out, fails


def collision_pass(scene, margin):
    """Timed CPU pass (AABBs + exact pairs + narrowphase): (pairs, contacts, seconds)."""
    r = _SceneRef(scene)
    nc = C.c_int64(0)
    sec = C.c_double(0)
    npairs = lib().oracle_pass_f32(C.byref(r.c), margin, C.byref(nc), C.byref(sec))
    return int(npairs), int(nc.value), float(sec.value)


def collision_pass_mt(scene, margin, threads):
    """The same pass with AABBs and narrowphase on `threads` OpenMP threads:
    (pairs, contacts, seconds)."""
    r = _SceneRef(scene)
    nc = C.c_int64(0)
    sec = C.c_double(0)
    npairs = lib().oracle_pass_f32_mt(C.byref(r.c), margin, int(threads), C.byref(nc), C.byref(sec))
    return int(npairs), int(nc.value), float(sec.value)


# ----------------------------------------------------- persistent manifold --
# One cache entry is 64 float32 (layout in oracle.hpp::pm_store); an emitted
# point is 12 float32: position, normal, penetration, impulse, lifetime
# (uint32 bits), A-side anchor.
PM_FILTERED, PM_SLEEPING, PM_STATIC_PAIR, PM_SERVED, PM_DETECTED, PM_REMOVED = 0, 1, 2, 3, 4, 5
PM_STATIC, PM_ACTIVE, PM_FAST = 1, 2, 4


def pm_new():
    """An empty PersistentManifold (manifold.rs:188-199)."""
    e = np.zeros(64, np.float32)
    e[51] = np.float32(0.02)
    e[55] = e[63] = np.float32(1.0)
    return e


def pm_count(e):
    return int(np.asarray(e, np.float32).view(np.uint32)[59] & 0xff)


def pm_points(e):
    """[n, 12] rows of the cached points (local_a, pen, local_b, impulse, normal, lifetime bits)."""
    return np.asarray(e, np.float32)[:12 * pm_count(e)].reshape(-1, 12)


def pm_capture(e, points7, pa, qa, pb, qb, threshold=0.02):
    """capture (manifold.rs:320-376) in place; points7 [k, 7] = position, normal, penetration."""
    pts = np.ascontiguousarray(points7, np.float32).reshape(-1, 7)
    lib().oracle_pm_capture(e.ctypes.data, pts.ctypes.data, pts.shape[0], _f32(pa, 3), _f32(qa, 4), _f32(pb, 3),
                            _f32(qb, 4), threshold)


def pm_refresh(e, pa, qa, pb, qb):
    """refresh (manifold.rs:252-298) in place -> emitted points [n, 12]."""
    out = np.zeros((4, 12), np.float32)
    n = lib().oracle_pm_refresh(e.ctypes.data, _f32(pa, 3), _f32(qa, 4), _f32(pb, 3), _f32(qb, 4), out.ctypes.data)
    return out[:n]


def pm_needs_redetect(e, pa, qa, pb, qb):
    return bool(lib().oracle_pm_needs_redetect(e.ctypes.data, _f32(pa, 3), _f32(qa, 4), _f32(pb, 3), _f32(qb, 4)))


def pm_seed(e):
    out = np.zeros(3, np.float32)
    return out if lib().oracle_pm_seed(e.ctypes.data, out.ctypes.data) else None


def pm_pass(scene, pairs, cache, flags=None, group_mask=None, slop=0.005, warm=0.85):
    """One step of the pair pass with persistent manifolds (step.rs:660-854).
    cache = (keys u64[M], entries f32[M, 64], valid u8[M]) of the previous
    step (or None).  Returns dict(route, count, points [P, 4, 12], cache=(keys,
    entries, valid) of this step)."""
    r = _SceneRef(scene)
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    P = pairs.shape[0]
    if cache is None:
        keys = np.zeros(1, np.uint64)
        ent = np.zeros((1, 64), np.float32)
        val = np.zeros(1, np.uint8)
        M = 0
    else:
        keys, ent, val = (np.ascontiguousarray(cache[0], np.uint64), np.ascontiguousarray(cache[1], np.float32),
                          np.ascontiguousarray(cache[2], np.uint8))
        M = keys.shape[0]
    fl = None if flags is None else np.ascontiguousarray(flags, np.uint8)
    gm = None if group_mask is None else np.ascontiguousarray(group_mask, np.uint32)
    route = np.zeros(P, np.uint8)
    valid = np.zeros(P, np.uint8)
    ent_out = np.zeros((max(P, 1), 64), np.float32)
    cnt = np.zeros(P, np.uint8)
    pts = np.zeros((max(P, 1), 4, 12), np.float32)
    lib().oracle_pm_pass(C.byref(r.c), pairs.ctypes.data, P, None if fl is None else fl.ctypes.data,
                         None if gm is None else gm.ctypes.data, slop, warm, keys.ctypes.data, ent.ctypes.data,
                         val.ctypes.data, M, route.ctypes.data, valid.ctypes.data, ent_out.ctypes.data,
                         cnt.ctypes.data, pts.ctypes.data)
    new_keys = (pairs[:, 0].astype(np.uint64) << np.uint64(32)) | pairs[:, 1].astype(np.uint64)
    return dict(route=route, count=cnt, points=pts[:P], cache=(new_keys, ent_out[:P], valid))


# ----------------------------------------------------------- triangle meshes --
class Mesh:
    """TriangleMesh (triangle_mesh.rs:84-123) with its BVH, in the oracle."""

    def __init__(self, vertices, indices):
        self.v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        self.i = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        self.h = lib().oracle_mesh_create(self.v.ctypes.data, self.v.shape[0], self.i.ctypes.data, self.i.shape[0])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.oracle_mesh_destroy(self.h)
            self.h = None

    @property
    def threshold(self):
        return float(lib().oracle_mesh_threshold(self.h))

    def query(self, qmin, qmax):
        """triangles_in_aabb in the BVH's order."""
        out = np.zeros(max(self.i.shape[0], 1), np.uint32)
        n = lib().oracle_mesh_query(self.h, _f32(qmin, 3), _f32(qmax, 3), out.ctypes.data, out.shape[0])
        return out[:n]

    def collide(self, convex: Shape, pc, qc, pm, qm, seed=None, mesh_first=False):
        """collide_seeded(convex, mesh) (or (mesh, convex) with mesh_first) -> points [k, 7]."""
        out = (C.c_float * 28)()
        n = lib().oracle_mesh_collide(C.byref(convex.c()), _f32(pc, 3), _f32(qc, 4), self.h, _f32(pm, 3),
                                      _f32(qm, 4), None if seed is None else _f32(seed, 3), int(mesh_first), out)
        return np.array(out[:], np.float32).reshape(4, 7)[:min(n, 4)], n


def mesh_collide_scene(scene, meshes, pairs, mesh_pose, seeds=None, mesh_first=None):
    """Batch of (body, mesh) pairs -> (count [P], points [P, 4, 7])."""
    r = _SceneRef(scene)
    pairs = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
    P = pairs.shape[0]
    hs = (C.c_void_p * max(len(meshes), 1))(*[m.h for m in meshes])
    mp = np.ascontiguousarray(mesh_pose, np.float32).reshape(-1, 7)
    sd = None if seeds is None else np.ascontiguousarray(seeds, np.float32)
    mf = None if mesh_first is None else np.ascontiguousarray(mesh_first, np.uint8)
    cnt = np.zeros(P, np.uint8)
    pts = np.zeros((max(P, 1), 4, 7), np.float32)
    lib().oracle_mesh_collide_scene(C.byref(r.c), hs, pairs.ctypes.data, P, mp.ctypes.data,
                                    None if sd is None else sd.ctypes.data, None if mf is None else mf.ctypes.data,
                                    cnt.ctypes.data, pts.ctypes.data)
    return cnt, pts[:P]


def sphere_triangle(center, radius, threshold, a, b, c):
    """sphere_triangle_contact (convex_mesh.rs:168-235) -> (point, normal, penetration) or None."""
    out = (C.c_float * 7)()
    if not lib().oracle_sphere_triangle(_f32(center, 3), radius, threshold, _f32(a, 3), _f32(b, 3), _f32(c, 3), out):
        return None
    o = np.array(out[:], np.float32)
    return o[0:3], o[3:6], float(o[6])


def reduce(points7, max_points=4):
    """ContactManifold::reduce (manifold.rs:76-110) -> points [k, 7]."""
    p = np.ascontiguousarray(points7, np.float32).reshape(-1, 7).copy()
    n = lib().oracle_reduce(p.ctypes.data, p.shape[0], max_points)
    return p[:n]


# ------------------------------------------------------------------ sweeps --
def _hit(ok, out):
    o = np.array(out[:], np.float32)
    return (float(o[0]), o[1:4], o[4:7]) if ok else None


def sweep_sphere(frm, to, radius, other: Shape, op=(0, 0, 0), oq=(0, 0, 0, 1)):
    """swept_sphere_cast (convex_sweep.rs:35-57) -> (fraction, normal, point) or None."""
    oc = other.c()
    out = (C.c_float * 7)()
    ok = lib().oracle_sweep_sphere(C.byref(oc), _f32(op, 3), _f32(oq, 4), _f32(frm, 3), _f32(to, 3), radius, out)
    return _hit(ok, out)


def sweep_sphere_mesh(frm, to, radius, mesh, pos=(0, 0, 0), ori=(0, 0, 0, 1)):
    out = (C.c_float * 7)()
    ok = lib().oracle_sweep_sphere_mesh(mesh.h, _f32(pos, 3), _f32(ori, 4), _f32(frm, 3), _f32(to, 3), radius, out)
    return _hit(ok, out)


def convex_cast(shape: Shape, frm, ori, to, other: Shape, op=(0, 0, 0), oq=(0, 0, 0, 1)):
    """convex_cast (convex_sweep.rs:357-385) -> (fraction, normal, point) or None."""
    a, o = shape.c(), other.c()
    out = (C.c_float * 7)()
    ok = lib().oracle_convex_cast(C.byref(a), _f32(frm, 3), _f32(ori, 4), _f32(to, 3), C.byref(o), _f32(op, 3),
                                  _f32(oq, 4), out)
    return _hit(ok, out)


def sweep_scene(scene, target, frm, to, radius=None, caster=None):
    """Batch over scene bodies -> (hit u8[n], out [n, 7] = fraction, normal, point)."""
    r = _SceneRef(scene)
    t = np.ascontiguousarray(target, np.uint32)
    n = t.shape[0]
    f = np.ascontiguousarray(frm, np.float32)
    e = np.ascontiguousarray(to, np.float32)
    rad = None if radius is None else np.ascontiguousarray(radius, np.float32)
    cs = None if caster is None else np.ascontiguousarray(caster, np.uint32)
    hit = np.zeros(n, np.uint8)
    out = np.zeros((max(n, 1), 7), np.float32)
    lib().oracle_sweep_scene(C.byref(r.c), t.ctypes.data, f.ctypes.data, e.ctypes.data,
                             None if rad is None else rad.ctypes.data, None if cs is None else cs.ctypes.data, n,
                             hit.ctypes.data, out.ctypes.data)
    return hit, out[:n]


def sweep_mesh_scene(scene, meshes, mesh_of, mesh_pose, frm, to, radius=None, caster=None):
    """Sphere sweeps (radius) or convex casts (caster bodies of scene) against
    meshes -> (hit u8[n], out [n, 7])."""
    r = None if scene is None else _SceneRef(scene)
    mo = np.ascontiguousarray(mesh_of, np.uint32)
    n = mo.shape[0]
    hs = (C.c_void_p * max(len(meshes), 1))(*[m.h for m in meshes])
    mp = np.ascontiguousarray(mesh_pose, np.float32)
    f = np.ascontiguousarray(frm, np.float32)
    e = np.ascontiguousarray(to, np.float32)
    rad = None if radius is None else np.ascontiguousarray(radius, np.float32)
    cs = None if caster is None else np.ascontiguousarray(caster, np.uint32)
    hit = np.zeros(n, np.uint8)
    out = np.zeros((max(n, 1), 7), np.float32)
    lib().oracle_sweep_mesh_scene(None if r is None else C.byref(r.c), hs, mo.ctypes.data, mp.ctypes.data,
                                  f.ctypes.data, e.ctypes.data, None if rad is None else rad.ctypes.data,
                                  None if cs is None else cs.ctypes.data, n, hit.ctypes.data, out.ctypes.data)
    return hit, out[:n]
