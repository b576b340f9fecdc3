// bch_sweep.hpp -- CCD sweeps (crates/bullet-collision/src/narrowphase/
// convex_sweep.rs): a sphere swept along a segment (swept_sphere_cast,
// :35-57) or a convex shape cast along one (convex_cast, :357-385), against
// convex shapes, planes and triangle meshes; one thread per query.
//
// The convex targets use the point-to-shape GJK distance walk
// (point_convex_distance, :59-103) over the simplex reduction closest_origin
// of gjk.rs:864-1043; the sweep core is the reference's ternary search for
// the closest approach and binary search for the entry (:150-206).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bch {

struct CoSimplex {
  V3 p[4];
  int n;
};

// gjk.rs:887-906
__device__ __forceinline__ V3 co_segment(V3 a, V3 b, CoSimplex& k) {
  V3 ab = b - a;
  if (len2(ab) == 0.0f) {
    k.p[0] = a;
    k.n = 1;
    return a;
  }
  float t = -dot(a, ab) / len2(ab);
  if (t <= 0.0f) {
    k.p[0] = a;
    k.n = 1;
    return a;
  }
  if (t >= 1.0f) {
    k.p[0] = b;
    k.n = 1;
    return b;
  }
  k.p[0] = b;
  k.p[1] = a;
  k.n = 2;
  return a + ab * t;
}

// gjk.rs:908-985
__device__ __forceinline__ V3 co_triangle(V3 a, V3 b, V3 c, CoSimplex& k) {
  V3 ab = b - a, ac = c - a, ap = -a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) { k.p[0] = a; k.n = 1; return a; }
  V3 bp = -b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) { k.p[0] = b; k.n = 1; return b; }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    if (d1 == d3) { k.p[0] = a; k.n = 1; return a; }
    float v = d1 / (d1 - d3);
    k.p[0] = b; k.p[1] = a; k.n = 2;
    return a + ab * v;
  }
  V3 cp = -c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) { k.p[0] = c; k.n = 1; return c; }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    if (d2 == d6) { k.p[0] = a; k.n = 1; return a; }
    float w = d2 / (d2 - d6);
    k.p[0] = c; k.p[1] = a; k.n = 2;
    return a + ac * w;
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    if ((d4 - d3) + (d5 - d6) == 0.0f) { k.p[0] = b; k.n = 1; return b; }
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    k.p[0] = c; k.p[1] = b; k.n = 2;
    return b + (c - b) * w;
  }
  float sum = (va + vb) + vc;
  if (sum <= 0.0f || is_nan(sum)) {  // degenerate: the closest of the three edges
    CoSimplex kab, kbc, kca;
    V3 qab = co_segment(a, b, kab), qbc = co_segment(b, c, kbc), qca = co_segment(c, a, kca);
    if (len2(qab) <= len2(qbc) && len2(qab) <= len2(qca)) { k = kab; return qab; }
    if (len2(qbc) <= len2(qca)) { k = kbc; return qbc; }
    k = kca;
    return qca;
  }
  float denom = 1.0f / sum;
  float v = vb * denom, w = vc * denom;
  k.p[0] = c; k.p[1] = b; k.p[2] = a; k.n = 3;
  return (a + ab * v) + ac * w;
}

// gjk.rs:987-1043: the first face whose side test puts the origin outside.
__device__ __forceinline__ V3 co_tetra(V3 a, V3 b, V3 c, V3 d, CoSimplex& k) {
  V3 ao = -a, ab = b - a, ac = c - a, ad = d - a;
  if (fabsf(dot(cross(ab, ac), ad)) < 1e-12f) {  // flat: the nearest face, first of equals
    CoSimplex kk;
    V3 best = co_triangle(a, b, c, k);
    V3 q = co_triangle(a, b, d, kk);
    if (len2(q) < len2(best)) { best = q; k = kk; }
    q = co_triangle(a, c, d, kk);
    if (len2(q) < len2(best)) { best = q; k = kk; }
    q = co_triangle(b, c, d, kk);
    if (len2(q) < len2(best)) { best = q; k = kk; }
    return best;
  }
  V3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
  V3 bo = -b;
  V3 bcd = cross(c - b, d - b);
  if (dot(abc, ao) * dot(abc, ad) < 0.0f) return co_triangle(a, b, c, k);
  if (dot(acd, ao) * dot(acd, ab) < 0.0f) return co_triangle(a, c, d, k);
  if (dot(adb, ao) * dot(adb, ac) < 0.0f) return co_triangle(a, d, b, k);
  if (dot(bcd, bo) * dot(bcd, a - b) < 0.0f) return co_triangle(b, c, d, k);
  k.p[0] = a; k.p[1] = b; k.p[2] = c; k.p[3] = d; k.n = 4;
  return splat(0.0f);
}

__device__ __forceinline__ V3 closest_origin(CoSimplex& s) {
  CoSimplex k;
  V3 v;
  if (s.n <= 0) return splat(0.0f);
  if (s.n == 1) return s.p[0];
  if (s.n == 2) v = co_segment(s.p[0], s.p[1], k);
  else if (s.n == 3) v = co_triangle(s.p[0], s.p[1], s.p[2], k);
  else v = co_tetra(s.p[0], s.p[1], s.p[2], s.p[3], k);
  s = k;
  return v;
}

// point_convex_distance (convex_sweep.rs:59-103)
template <class F>
__device__ __forceinline__ void point_convex_distance(const F& support, V3 point, float& d2, V3& dir) {
  V3 v = point;
  if (len2(v) < 1e-12f) v = V3{1.0f, 0.0f, 0.0f};
  CoSimplex s;
  s.p[0] = point - support(-v);
  s.n = 1;
  v = s.p[0];
  for (int it = 0; it < 64; ++it) {
    if (len2(v) < 1e-10f) {
      d2 = 0.0f;
      dir = splat(0.0f);
      return;
    }
    V3 w = point - support(v);  // mink(-v)
    bool dup = false;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < s.n && len2(s.p[i] - w) < 1e-12f) dup = true;
    if (dup) {
      d2 = len2(v);
      dir = v / len(v);
      return;
    }
    // append with constant indices only, so the simplex stays in registers
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i == s.n) s.p[i] = w;
    s.n++;
    v = closest_origin(s);
  }
  if (len2(v) < 1e-10f) {
    d2 = 0.0f;
    dir = splat(0.0f);
    return;
  }
  d2 = len2(v);
  dir = v / len(v);
}

struct SweepHit {
  float fraction;
  V3 normal, point;
};

// sweep_with_dist (convex_sweep.rs:150-206)
template <class D>
__device__ __forceinline__ bool sweep_with_dist(V3 from, V3 dir, float length, float radius, const D& dist,
                                                SweepHit& h) {
  float lo = 0.0f, hi = 1.0f;
  for (int i = 0; i < 40; ++i) {
    float m1 = lo + (hi - lo) / 3.0f;
    float m2 = hi - (hi - lo) / 3.0f;
    float a2, b2;
    V3 t;
    dist(from + dir * (m1 * length), a2, t);
    dist(from + dir * (m2 * length), b2, t);
    if (a2 < b2) hi = m2;
    else lo = m1;
  }
  float s_min = (lo + hi) * 0.5f;
  float dm;
  V3 t;
  dist(from + dir * (s_min * length), dm, t);
  if (dm > 1e-8f) return false;
  float low = 0.0f, high = s_min;
  for (int i = 0; i < 60; ++i) {
    float mid = (low + high) * 0.5f;
    float d2;
    dist(from + dir * (mid * length), d2, t);
    if (d2 <= 1e-8f) high = mid;
    else low = mid;
  }
  float fraction = (low + high) * 0.5f;
  float pr = (fraction - 2e-4f) * length;
  V3 probe = from + dir * (0.0f > pr ? 0.0f : pr);
  float d2;
  V3 v;
  dist(probe, d2, v);
  V3 normal = len2(v) > 1e-12f ? v / len(v) : -dir;
  V3 entry = from + dir * (fraction * length);
  h.fraction = fraction;
  h.normal = normal;
  h.point = entry - normal * radius;
  return true;
}

// point_triangle_distance (convex_sweep.rs:208-256)
__device__ __forceinline__ float point_tri_dist(V3 p, V3 a, V3 b, V3 c, V3& dv) {
  V3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) { dv = p - a; return len2(p - a); }
  V3 bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) { dv = p - b; return len2(p - b); }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / (d1 - d3);
    V3 q = a + ab * v;
    dv = p - q;
    return len2(p - q);
  }
  V3 cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) { dv = p - c; return len2(p - c); }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    V3 q = a + ac * w;
    dv = p - q;
    return len2(p - q);
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    V3 q = b + (c - b) * w;
    dv = p - q;
    return len2(p - q);
  }
  float denom = 1.0f / ((va + vb) + vc);
  float v = vb * denom, w = vc * denom;
  V3 q = (a + ab * v) + ac * w;
  dv = p - q;
  return len2(p - q);
}

// glam Quat * Quat (scalar order)
__device__ __forceinline__ Q4 qmul(Q4 a, Q4 b) {
  return Q4{((a.w * b.x + a.x * b.w) + a.y * b.z) - a.z * b.y, ((a.w * b.y - a.x * b.z) + a.y * b.w) + a.z * b.x,
            ((a.w * b.z + a.x * b.y) - a.y * b.x) + a.z * b.w, ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z};
}

// swept_sphere_cast against a convex or plane target (convex_sweep.rs:35-57,
// sweep_convex :105-125, sweep_plane :258-292).
__device__ __forceinline__ bool sweep_sphere(V3 from, V3 to, float radius, const ShapeRef& other, V3 op, Q4 oq,
                                             SweepHit& h) {
  V3 delta = to - from;
  float length = len(delta);
  if (length <= 1.1920929e-7f) return false;
  V3 dir = delta / length;
  if (other.type == PLANE) {
    V3 n = qrot(oq, V3{other.a0, other.a1, other.a2});
    float dn0 = dot(n, from) - other.a3;
    if (fabsf(dn0) <= radius) {
      h.fraction = 0.0f;
      h.normal = dn0 >= 0.0f ? n : -n;
      h.point = from - n * dn0;
      return true;
    }
    float dd = dot(n, dir * length);
    if (dd >= 0.0f) return false;
    float sign = dn0 > 0.0f ? 1.0f : -1.0f;
    float fraction = (sign * radius - dn0) / dd;
    if (!(fraction >= 0.0f && fraction <= 1.0f)) return false;
    V3 normal = sign > 0.0f ? n : -n;
    V3 entry = from + (dir * length) * fraction;
    h.fraction = fraction;
    h.normal = normal;
    h.point = entry - normal * radius;
    return true;
  }
  auto support = [&](V3 d) {
    V3 sph = len2(d) > 1e-12f ? (d / len(d)) * radius : splat(0.0f);
    return support_world(other, op, oq, d) + sph;
  };
  auto dist = [&](V3 p, float& d2, V3& v) { point_convex_distance(support, p, d2, v); };
  return sweep_with_dist(from, dir, length, radius, dist, h);
}

// convex_cast against a convex or plane target (sweep_convex_shape :387-429,
// sweep_convex_plane :431-478).
__device__ __forceinline__ bool convex_cast(const ShapeRef& shape, V3 from, Q4 ori, V3 to, const ShapeRef& other,
                                            V3 op, Q4 oq, SweepHit& h) {
  V3 delta = to - from;
  float length = len(delta);
  if (length <= 1.1920929e-7f) return false;
  if (other.type == PLANE) {
    V3 n = qrot(oq, V3{other.a0, other.a1, other.a2});
    float c = other.a3 + dot(n, op);
    V3 normal = dot(n, from) >= c ? n : -n;
    V3 dfrom = support_world(shape, from, ori, -normal);
    float da = dot(normal, dfrom) - c;
    float db = dot(normal, support_world(shape, to, ori, -normal)) - c;
    if (db >= da) return false;
    if (da <= 0.0f) {
      h.fraction = 0.0f;
      h.normal = normal;
      h.point = dfrom;
      return true;
    }
    float fraction = da / (da - db);
    if (!(fraction >= 0.0f && fraction <= 1.0f)) return false;
    V3 entry = from + delta * fraction;
    h.fraction = fraction;
    h.normal = normal;
    h.point = support_world(shape, entry, ori, -normal);
    return true;
  }
  V3 dir = delta / length;
  auto dist = [&](V3 p, float& d2, V3& v) {
    auto support = [&](V3 d) { return support_world(other, op, oq, d) - support_world(shape, p, ori, -d); };
    point_convex_distance(support, splat(0.0f), d2, v);
  };
  if (!sweep_with_dist(from, dir, length, 0.0f, dist, h)) return false;
  h.point = support_world(other, op, oq, h.normal);
  return true;
}

// Shape class of a body-target query (target type x caster type, 0 for a
// sphere sweep): the sort key that groups like queries into wavefronts, so
// a wavefront runs one support map instead of every branch.
__global__ void k_sweep_keys(uint32_t n, const uint32_t* __restrict__ target, const uint32_t* __restrict__ caster,
                             const BodyRec* __restrict__ bodies, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ idx) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  uint32_t tt = bodies[target[k]].type & 0xffu;
  uint32_t ct = caster ? (bodies[caster[k]].type & 0xffu) + 1u : 0u;
  keys[k] = tt * 8u + ct;
  idx[k] = k;
}

// Body-target sweeps: query k against body target[k]; a sphere of radius[k]
// when caster is NULL, else body caster[k]'s shape at its orientation.
// Thread i runs query order[i] (the class-sorted order) when order is set.
// Four waves per SIMD (128 VGPRs, a short spill) hold 262,144 queries in one
// round on 1024 SIMDs; at three (158 VGPRs) that batch took 139 ms, now 120.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_sweep(uint32_t n, const uint32_t* __restrict__ order,
                                               const uint32_t* __restrict__ target,
                                               const uint32_t* __restrict__ caster, const float* __restrict__ from,
                                               const float* __restrict__ to, const float* __restrict__ radius,
                                               const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                                               uint8_t* __restrict__ hit, float* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = order ? order[i] : i;
  BodyRec T = bodies[target[k]];
  ShapeRef ts = shape_of(T, hull);
  V3 f{from[3 * (uint64_t)k], from[3 * (uint64_t)k + 1], from[3 * (uint64_t)k + 2]};
  V3 e{to[3 * (uint64_t)k], to[3 * (uint64_t)k + 1], to[3 * (uint64_t)k + 2]};
  SweepHit h;
  bool ok;
  if (caster) {
    BodyRec Cb = bodies[caster[k]];
    ok = convex_cast(shape_of(Cb, hull), f, quat_of(Cb), e, ts, pos_of(T), quat_of(T), h);
  } else {
    ok = sweep_sphere(f, e, radius[k], ts, pos_of(T), quat_of(T), h);
  }
  float* o = out + 7 * (uint64_t)k;
  hit[k] = ok ? 1 : 0;
  o[0] = ok ? h.fraction : 0.0f;
  o[1] = ok ? h.normal.x : 0.0f;
  o[2] = ok ? h.normal.y : 0.0f;
  o[3] = ok ? h.normal.z : 0.0f;
  o[4] = ok ? h.point.x : 0.0f;
  o[5] = ok ? h.point.y : 0.0f;
  o[6] = ok ? h.point.z : 0.0f;
}

// Mesh sweeps (sweep_mesh :294-344, sweep_convex_mesh :480-524): the BVH walk
// (no triangle-AABB filter), per-triangle casts in mesh space, the smallest
// fraction (first on ties), mapped back to world space.
__global__ __launch_bounds__(256) void k_sweep_mesh(uint32_t n, const uint32_t* __restrict__ mesh_of,
                                                    const float* __restrict__ mesh_pose,
                             const uint32_t* __restrict__ caster, const float* __restrict__ from,
                             const float* __restrict__ to, const float* __restrict__ radius,
                             const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                             const MeshRec* __restrict__ meshes, const float4* __restrict__ nodes,
                             const uint32_t* __restrict__ order, const float4* __restrict__ tris,
                             uint8_t* __restrict__ hit, float* __restrict__ out, Counters* cnt) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  MeshRec M = meshes[mesh_of[k]];
  const float* mp = mesh_pose + 7 * (uint64_t)k;
  V3 pos{mp[0], mp[1], mp[2]};
  Q4 ori{mp[3], mp[4], mp[5], mp[6]};
  Q4 inv = qconj(ori);
  V3 f{from[3 * (uint64_t)k], from[3 * (uint64_t)k + 1], from[3 * (uint64_t)k + 2]};
  V3 e{to[3 * (uint64_t)k], to[3 * (uint64_t)k + 1], to[3 * (uint64_t)k + 2]};
  V3 lf = qrot(inv, f - pos), lt = qrot(inv, e - pos);
  bool any = false;
  SweepHit best;
  best.fraction = 0.0f;
  best.normal = best.point = splat(0.0f);
  bool walked = true;
  if (!caster) {
    float r = radius[k];
    V3 delta = lt - lf;
    float length = len(delta);
    if (length > 1.1920929e-7f) {
      V3 dir = delta / length;
      V3 rr = splat(r);
      Box6 q{vmin(lf, lt) - rr, vmax(lf, lt) + rr};
      walked = mesh_walk(M, nodes, order, nullptr, q, [&](uint32_t t) {
        const float4* tv = tris + 3 * (uint64_t)(M.tri_off + t);
        V3 a = hull_vertex(tv, 0), b = hull_vertex(tv, 1), c = hull_vertex(tv, 2);
        auto dist = [&](V3 p, float& d2, V3& v) {
          V3 dv;
          float q2 = point_tri_dist(p, a, b, c, dv);
          if (q2 <= r * r) {
            d2 = 0.0f;
            v = splat(0.0f);
          } else {
            float d = sqrtf(q2);
            d2 = (d - r) * (d - r);
            v = dv / d;
          }
        };
        SweepHit hh;
        if (sweep_with_dist(lf, dir, length, r, dist, hh) && (!any || hh.fraction < best.fraction)) {
          best = hh;
          any = true;
        }
      });
    }
  } else {
    BodyRec Cb = bodies[caster[k]];
    ShapeRef cs = shape_of(Cb, hull);
    Q4 ol = qmul(inv, quat_of(Cb));
    Box6 a = world_aabb(cs, lf, ol), b = world_aabb(cs, lt, ol);
    Box6 q{vmin(a.lo, b.lo), vmax(a.hi, b.hi)};
    walked = mesh_walk(M, nodes, order, nullptr, q, [&](uint32_t t) {
      ShapeRef tri;
      tri.type = HULL;
      tri.a0 = tri.a1 = tri.a2 = tri.a3 = 0.0f;
      tri.hull = tris + 3 * (uint64_t)(M.tri_off + t);
      tri.nhull = 3;
      tri.margin = 0.0f;
      SweepHit hh;
      if (convex_cast(cs, lf, ol, lt, tri, splat(0.0f), Q4{0.0f, 0.0f, 0.0f, 1.0f}, hh) &&
          (!any || hh.fraction < best.fraction)) {
        best = hh;
        any = true;
      }
    });
  }
  if (!walked) atomicOr(&cnt->errors, 32u);
  float* o = out + 7 * (uint64_t)k;
  hit[k] = any ? 1 : 0;
  V3 nw = qrot(ori, best.normal), pw = qrot(ori, best.point) + pos;
  o[0] = any ? best.fraction : 0.0f;
  o[1] = any ? nw.x : 0.0f;
  o[2] = any ? nw.y : 0.0f;
  o[3] = any ? nw.z : 0.0f;
  o[4] = any ? pw.x : 0.0f;
  o[5] = any ? pw.y : 0.0f;
  o[6] = any ? pw.z : 0.0f;
}

}  // namespace bch
