// bch_shapes.hpp -- device body record and the convex support maps.
//
// Shape semantics follow crates/bullet-collision/src/shapes:
//   sphere   support  r * d/|d| (x-axis when d = 0), core = centre, margin r
//   box      support  sign(d) * h (d >= 0 -> +h), core = max(h - 0.04, 0), margin 0.04
//   capsule  support  (0, +-hh, 0) + r d/|d|, core = the segment tip, margin r
//   hull     support  first vertex maximising p.d (strict >), core = hull, margin 0
//   plane    n.x = c, AABB +-1e30, never a support query (plane-convex is a closed form)
// (sphere.rs:17-48, box_shape.rs:17-72, capsule.rs:24-79,
//  convex_hull.rs:43-72, plane.rs:25-40, shapes/mod.rs:273-469).
#pragma once
#include "bch_math.hpp"

namespace bch {

enum : int { SPHERE = 0, BOX = 1, CAPSULE = 2, HULL = 3, PLANE = 4 };

// 64-byte body record kept resident in HBM (one cache-line half per body):
//   r0 = position.xyz, type   r1 = orientation (x, y, z, w)
//   r2 = shape params (hull: bits(first vertex), bits(vertex count))
//   r3 = contact_breaking_threshold, margin, bits(world), bound
// bound = bound_radius (a ball about the position holding the shape and its
// margin-grown core) + contact_breaking_threshold.
struct alignas(16) BodyRec {
  float px, py, pz;
  uint32_t type;
  float qx, qy, qz, qw;
  float a0, a1, a2, a3;
  float threshold, margin;
  uint32_t world;
  float bound;
};

struct ShapeRef {
  int type;
  float a0, a1, a2, a3;
  const float4* hull;
  int nhull;
  float margin;
};

BCH_HD uint32_t f2u(float f) {
  union {
    float f;
    uint32_t u;
  } c;
  c.f = f;
  return c.u;
}
BCH_HD float u2f(uint32_t u) {
  union {
    float f;
    uint32_t u;
  } c;
  c.u = u;
  return c.f;
}

BCH_HD ShapeRef shape_of(const BodyRec& b, const float4* hull_verts) {
  ShapeRef s;
  s.type = (int)(b.type & 0xffu);
  s.a0 = b.a0;
  s.a1 = b.a1;
  s.a2 = b.a2;
  s.a3 = b.a3;
  s.margin = b.margin;
  if (s.type == HULL) {
    s.hull = hull_verts + f2u(b.a0);
    s.nhull = (int)f2u(b.a1);
  } else {
    s.hull = nullptr;
    s.nhull = 0;
  }
  return s;
}
BCH_HD V3 pos_of(const BodyRec& b) { return V3{b.px, b.py, b.pz}; }
BCH_HD Q4 quat_of(const BodyRec& b) { return Q4{b.qx, b.qy, b.qz, b.qw}; }

BCH_HD float shape_margin(int type, float a0) {
  return type == SPHERE ? a0 : (type == BOX ? 0.04f : (type == CAPSULE ? a0 : 0.0f));
}

BCH_HD V3 hull_vertex(const float4* h, int i) {
  float4 v = h[i];
  return V3{v.x, v.y, v.z};
}

BCH_HD V3 hull_support(const float4* h, int n, V3 d) {
  V3 best = splat(0.0f);
  float best_dot = -INFINITY;
  for (int i = 0; i < n; ++i) {
    V3 p = hull_vertex(h, i);
    float dd = dot(p, d);
    if (dd > best_dot) {
      best_dot = dd;
      best = p;
    }
  }
  return best;
}

BCH_HD V3 support_local(const ShapeRef& s, V3 d) {
  switch (s.type) {
    case SPHERE: {
      float l = len(d);
      if (l > 0.0f) return (d / l) * s.a0;
      return V3{s.a0, 0.0f, 0.0f};
    }
    case BOX:
      return V3{d.x >= 0.0f ? s.a0 : -s.a0, d.y >= 0.0f ? s.a1 : -s.a1, d.z >= 0.0f ? s.a2 : -s.a2};
    case CAPSULE: {
      float l = len(d);
      if (l > 0.0f) {
        V3 tip{0.0f, d.y >= 0.0f ? s.a1 : -s.a1, 0.0f};
        return tip + (d / l) * s.a0;
      }
      return V3{s.a0, s.a1, 0.0f};
    }
    case HULL: return hull_support(s.hull, s.nhull, d);
    default: return splat(0.0f);
  }
}

BCH_HD V3 core_support_local(const ShapeRef& s, V3 d) {
  switch (s.type) {
    case SPHERE: return splat(0.0f);
    case BOX: {
      float hx = mx(s.a0 - 0.04f, 0.0f), hy = mx(s.a1 - 0.04f, 0.0f), hz = mx(s.a2 - 0.04f, 0.0f);
      return V3{d.x >= 0.0f ? hx : -hx, d.y >= 0.0f ? hy : -hy, d.z >= 0.0f ? hz : -hz};
    }
    case CAPSULE: return V3{0.0f, d.y >= 0.0f ? s.a1 : -s.a1, 0.0f};
    default: return support_local(s, d);
  }
}

BCH_HD V3 support_world(const ShapeRef& s, V3 p, Q4 q, V3 d) {
  V3 ld = qrot(qconj(q), d);
  return qrot(q, support_local(s, ld)) + p;
}
BCH_HD V3 core_support_world(const ShapeRef& s, V3 p, Q4 q, V3 d) {
  V3 ld = qrot(qconj(q), d);
  return qrot(q, core_support_local(s, ld)) + p;
}

// World AABB (shapes/mod.rs:350-375), not yet padded by the broadphase margin.
BCH_HD Box6 world_aabb(const ShapeRef& s, V3 p, Q4 q) {
  switch (s.type) {
    case SPHERE: {
      V3 r = splat(s.a0);
      return Box6{p - r, p + r};
    }
    case BOX: {
      V3 lo = splat(INFINITY), hi = splat(-INFINITY);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        V3 corner{(i & 1) == 0 ? s.a0 : -s.a0, (i & 2) == 0 ? s.a1 : -s.a1, (i & 4) == 0 ? s.a2 : -s.a2};
        V3 w = qrot(q, corner) + p;
        lo = vmin(lo, w);
        hi = vmax(hi, w);
      }
      return Box6{lo, hi};
    }
    case CAPSULE: {
      V3 top = qrot(q, V3{0.0f, s.a1, 0.0f}) + p;
      V3 bottom = qrot(q, V3{0.0f, -s.a1, 0.0f}) + p;
      V3 r = splat(s.a0);
      return Box6{vmin(top, bottom) - r, vmax(top, bottom) + r};
    }
    case HULL: {
      Q4 qi = qconj(q);
      V3 ux = qrot(qi, V3{1.0f, 0.0f, 0.0f});
      V3 uy = qrot(qi, V3{0.0f, 1.0f, 0.0f});
      V3 uz = qrot(qi, V3{0.0f, 0.0f, 1.0f});
      // Six support projections fused into one pass over the vertices; each
      // keeps the first maximiser like six separate scans would.
      float bx = -INFINITY, by = -INFINITY, bz = -INFINITY, cx = -INFINITY, cy = -INFINITY, cz = -INFINITY;
      V3 px = splat(0.0f), py = px, pz = px, nx = px, ny = px, nz = px;
      V3 mux = -ux, muy = -uy, muz = -uz;
      for (int i = 0; i < s.nhull; ++i) {
        V3 v = hull_vertex(s.hull, i);
        float d;
        d = dot(v, ux); if (d > bx) { bx = d; px = v; }
        d = dot(v, uy); if (d > by) { by = d; py = v; }
        d = dot(v, uz); if (d > bz) { bz = d; pz = v; }
        d = dot(v, mux); if (d > cx) { cx = d; nx = v; }
        d = dot(v, muy); if (d > cy) { cy = d; ny = v; }
        d = dot(v, muz); if (d > cz) { cz = d; nz = v; }
      }
      Box6 r;
      r.lo = V3{p.x + dot(nx, ux), p.y + dot(ny, uy), p.z + dot(nz, uz)};
      r.hi = V3{p.x + dot(px, ux), p.y + dot(py, uy), p.z + dot(pz, uz)};
      return r;
    }
    default: {
      V3 b = splat(1e30f);
      return Box6{-b, b};
    }
  }
}

// Radius of a ball about the body's position that holds both the full shape
// and the core grown by the margin (the rounded shape the margin-core walk
// measures): sphere r, capsule hh + r, box max(|h|, |core h| + 0.04), hull
// the farthest vertex (margin 0), plane +inf.  Not a reference quantity: the
// pair split uses it to skip walks that cannot end within the band.  NaN
// when a parameter is not finite (no pair of the body is skipped then).
BCH_HD float bound_radius(const ShapeRef& s) {
  float r;
  if (s.type == SPHERE) {
    r = s.a0;
  } else if (s.type == CAPSULE) {
    r = s.a1 + s.a0;
  } else if (s.type == BOX) {
    V3 h{s.a0, s.a1, s.a2};
    V3 hc{fmaxf(s.a0 - 0.04f, 0.0f), fmaxf(s.a1 - 0.04f, 0.0f), fmaxf(s.a2 - 0.04f, 0.0f)};
    float rh = sqrtf(len2(h)), rc = sqrtf(len2(hc)) + 0.04f;
    r = rh > rc ? rh : rc;
  } else if (s.type == HULL) {
    float r2 = 0.0f;
    bool finite = true;
    for (int v = 0; v < s.nhull; ++v) {
      float q = len2(V3{s.hull[v].x, s.hull[v].y, s.hull[v].z});
      finite = finite && q <= 3.0e38f;
      r2 = q > r2 ? q : r2;
    }
    r = finite ? sqrtf(r2) : NAN;
  } else {
    return INFINITY;
  }
  return r >= 0.0f && r <= 3.0e38f ? r : NAN;
}

// Local AABB and contact_breaking_threshold (shapes/mod.rs:254-270, 388-392).
BCH_HD Box6 local_aabb(const ShapeRef& s) {
  switch (s.type) {
    case SPHERE: {
      V3 r = splat(s.a0);
      return Box6{-r, r};
    }
    case BOX: {
      V3 h{s.a0, s.a1, s.a2};
      return Box6{-h, h};
    }
    case CAPSULE: {
      V3 r{s.a0, s.a1 + s.a0, s.a0};
      return Box6{-r, r};
    }
    case HULL: {
      V3 lo = splat(INFINITY), hi = splat(-INFINITY);
      for (int i = 0; i < s.nhull; ++i) {
        V3 v = hull_vertex(s.hull, i);
        lo = vmin(lo, v);
        hi = vmax(hi, v);
      }
      if (s.nhull == 0) {
        lo = splat(0.0f);
        hi = splat(0.0f);
      }
      return Box6{lo, hi};
    }
    default: {
      V3 b = splat(1e30f);
      return Box6{-b, b};
    }
  }
}

BCH_HD float breaking_threshold(const ShapeRef& s) {
  Box6 a = local_aabb(s);
  V3 c = (a.lo + a.hi) * 0.5f;
  float disc = len(a.hi - a.lo) * 0.5f + len(c);
  return disc * 0.02f;
}

}  // namespace bch
