// oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the bullet-collision crate's broadphase + convex
// narrowphase (CrazyboyQCD/bullet-physics), templated on float / double so the
// same text gives the f32 build (the parity oracle) and an f64 build (for the
// error-vs-f64 figures).  Every function cites the reference file:line it
// follows; statement order, operand order and branch structure are kept so an
// f32 build with -ffp-contract=off reproduces the reference arithmetic except
// where glam's internal op order is unknown (quaternion rotate, matrix-vector
// product: the order used is the one written in `qrot` / `mmul` below).
//
// Only tests/, the smoke check and bench.py's CPU-baseline leg may use this
// code, and only as the checker or the timed CPU baseline; the shipped library
// never links it.
//
// Parity pinning (which checks fix which function) is listed in the test
// files under tests/ (test_oracle_*.py); tests/plant_oracle.py plants single
// mistakes here and checks that some test fails.  Without a closed-form,
// known-answer or brute-force check: the quaternion op order (glam boundary),
// and so the last f32 bits of GJK / EPA on generic pairs (geometry pins them
// to the walks' own tolerances) -- "parity unpinned" beyond the restatement.  The CCD sweeps' entry fractions are pinned (the
// reference's sweep tests, a dense scan, the slab entry of box casts, the
// sphere cast against the sphere sweep); their hit normal is pinned only on
// the reference's axis-aligned cases and is otherwise "parity unpinned": it
// is the distance walk's direction at a probe point, restated step by step.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

namespace oracle {

// A literal spelled as the reference spells it, rounded once into T (an f32
// build parses the decimal straight to f32 like rustc; no double rounding).
#define ORL(x) (std::is_same<T, float>::value ? (T)(x##f) : (T)(x))

// ---------------------------------------------------------------- vectors --
// glam Vec3A / DVec3 semantics (crates/bullet-math/src/lib.rs:30-54).
template <class T>
struct V3 {
  T x, y, z;
};
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class T> inline T len2(V3<T> a) { return dot(a, a); }
template <class T> inline T len(V3<T> a) { return std::sqrt(dot(a, a)); }
// glam normalize: self * (1 / length)
template <class T> inline V3<T> normalize(V3<T> a) { return a * (T(1) / len(a)); }
template <class T> inline T fmin_(T a, T b) { return b < a ? b : a; }
template <class T> inline T fmax_(T a, T b) { return b > a ? b : a; }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
template <class T> inline V3<T> vabs(V3<T> a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }
template <class T> inline V3<T> splat(T s) { return {s, s, s}; }
template <class T> inline T comp(V3<T> v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
template <class T> inline bool abs_diff_eq(V3<T> a, V3<T> b, T eps) {
  return std::fabs(a.x - b.x) <= eps && std::fabs(a.y - b.y) <= eps && std::fabs(a.z - b.z) <= eps;
}
// Rust f32::signum for non-NaN input.
template <class T> inline T signum(T x) { return std::isnan(x) ? x : std::copysign(T(1), x); }

template <class T>
struct Quat {
  T x, y, z, w;
};
// Quaternion rotation v' = v (w^2 - b.b) + b (2 v.b) + (b x v)(2 w), b = (x,y,z).
template <class T> inline V3<T> qrot(Quat<T> q, V3<T> v) {
  V3<T> b{q.x, q.y, q.z};
  T b2 = dot(b, b);
  V3<T> t1 = v * (q.w * q.w - b2);
  V3<T> t2 = b * (dot(v, b) * T(2));
  V3<T> t3 = cross(b, v) * (q.w * T(2));
  return (t1 + t2) + t3;
}
template <class T> inline Quat<T> qconj(Quat<T> q) { return {-q.x, -q.y, -q.z, q.w}; }

// Column-major 3x3 (glam Mat3A).
template <class T>
struct M3 {
  V3<T> c0, c1, c2;
};
template <class T> inline V3<T> col(const M3<T>& m, int i) { return i == 0 ? m.c0 : (i == 1 ? m.c1 : m.c2); }
template <class T> inline M3<T> from_quat(Quat<T> q) {
  T x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  T xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  T yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  T wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  M3<T> m;
  m.c0 = {T(1) - (yy + zz), xy + wz, xz - wy};
  m.c1 = {xy - wz, T(1) - (xx + zz), yz + wx};
  m.c2 = {xz + wy, yz - wx, T(1) - (xx + yy)};
  return m;
}
template <class T> inline V3<T> mmul(const M3<T>& m, V3<T> v) { return (m.c0 * v.x + m.c1 * v.y) + m.c2 * v.z; }
template <class T> inline M3<T> transpose(const M3<T>& m) {
  M3<T> t;
  t.c0 = {m.c0.x, m.c1.x, m.c2.x};
  t.c1 = {m.c0.y, m.c1.y, m.c2.y};
  t.c2 = {m.c0.z, m.c1.z, m.c2.z};
  return t;
}
template <class T> inline M3<T> mmulm(const M3<T>& a, const M3<T>& b) {
  return M3<T>{mmul(a, b.c0), mmul(a, b.c1), mmul(a, b.c2)};
}

// ------------------------------------------------------------------- Aabb --
// crates/bullet-math/src/util.rs:5-125
template <class T>
struct Aabb {
  V3<T> min, max;
  V3<T> center() const { return (min + max) * T(0.5); }
  // util.rs:28-38 (closed intervals; touching counts)
  bool overlaps(const Aabb& o) const {
    return min.x <= o.max.x && min.y <= o.max.y && min.z <= o.max.z && o.min.x <= max.x &&
           o.min.y <= max.y && o.min.z <= max.z;
  }
  // util.rs:52-55
  Aabb expanded(T m) const {
    V3<T> pad = splat(m);
    return Aabb{min - pad, max + pad};
  }
};

// broadphase/mod.rs:153-183: slab test of the segment from -> to (t in
// [0, 1]); f32::max / f32::min keep the non-NaN operand like fmax / fmin.
template <class T>
inline bool segment_hits_aabb(const Aabb<T>& a, V3<T> from, V3<T> to) {
  V3<T> dir = to - from;
  T t_min = T(0), t_max = T(1);
  for (int k = 0; k < 3; ++k) {
    T d = k == 0 ? dir.x : (k == 1 ? dir.y : dir.z);
    T f = k == 0 ? from.x : (k == 1 ? from.y : from.z);
    T amin = k == 0 ? a.min.x : (k == 1 ? a.min.y : a.min.z);
    T amax = k == 0 ? a.max.x : (k == 1 ? a.max.y : a.max.z);
    T lo = d >= T(0) ? amin : amax;
    T hi = d >= T(0) ? amax : amin;
    if (d == T(0)) {
      if (f < lo || f > hi) return false;
      continue;
    }
    T inv = T(1) / d;
    t_min = std::fmax(t_min, (lo - f) * inv);
    t_max = std::fmin(t_max, (hi - f) * inv);
    if (t_min > t_max) return false;
  }
  return true;
}

// ----------------------------------------------------------------- shapes --
enum Kind : uint8_t { SPHERE = 0, BOX = 1, CAPSULE = 2, HULL = 3, PLANE = 4 };

// CONTACT_BREAKING_THRESHOLD, crates/bullet-collision/src/manifold.rs:8
template <class T> inline T contact_breaking_threshold_const() { return ORL(0.02); }

template <class T>
struct Shape {
  Kind kind = SPHERE;
  T radius = 0, half_height = 0;    // sphere, capsule
  V3<T> half{0, 0, 0};              // box
  V3<T> pnormal{0, 1, 0};           // plane (normalized)
  T pconst = 0;                     // plane
  const V3<T>* pts = nullptr;       // hull
  int npts = 0;
  Aabb<T> hull_bounds{{0, 0, 0}, {0, 0, 0}};

  static Shape sphere(T r) { Shape s; s.kind = SPHERE; s.radius = r; return s; }
  static Shape box(V3<T> h) { Shape s; s.kind = BOX; s.half = h; return s; }
  static Shape capsule(T r, T hh) { Shape s; s.kind = CAPSULE; s.radius = r; s.half_height = hh; return s; }
  // shapes/plane.rs:18-23: the normal is normalized on construction.
  static Shape plane(V3<T> n, T c) { Shape s; s.kind = PLANE; s.pnormal = normalize(n); s.pconst = c; return s; }
  // shapes/convex_hull.rs:21-36: cached local bounds, zero when empty.
  static Shape hull(const V3<T>* p, int n) {
    Shape s;
    s.kind = HULL;
    s.pts = p;
    s.npts = n;
    V3<T> mn = splat(std::numeric_limits<T>::infinity());
    V3<T> mx = splat(-std::numeric_limits<T>::infinity());
    for (int i = 0; i < n; ++i) {
      mn = vmin(mn, p[i]);
      mx = vmax(mx, p[i]);
    }
    if (n == 0) {
      mn = splat(T(0));
      mx = splat(T(0));
    }
    s.hull_bounds = Aabb<T>{mn, mx};
    return s;
  }

  // shapes/mod.rs:236-251 + per-shape margin()
  T margin() const {
    switch (kind) {
      case SPHERE: return radius;           // sphere.rs:17-20
      case BOX: return ORL(0.04);           // box_shape.rs:17-21
      case CAPSULE: return radius;          // capsule.rs:24-28
      case HULL: return T(0);               // convex_hull.rs:43-45
      case PLANE: return T(0);              // plane.rs:25-28
    }
    return T(0);
  }
  // shapes/mod.rs:254-270 + per-shape local_aabb()
  Aabb<T> local_aabb() const {
    switch (kind) {
      case SPHERE: { V3<T> r = splat(radius); return Aabb<T>{-r, r}; }
      case BOX: return Aabb<T>{-half, half};
      case CAPSULE: { V3<T> r = mk(radius, half_height + radius, radius); return Aabb<T>{-r, r}; }
      case HULL: return hull_bounds;
      case PLANE: { V3<T> b = splat(ORL(1e30)); return Aabb<T>{-b, b}; }
    }
    return Aabb<T>{};
  }
  // shapes/mod.rs:273-290
  V3<T> support_local(V3<T> dir) const {
    switch (kind) {
      case SPHERE: {  // sphere.rs:35-42
        T l = len(dir);
        if (l > T(0)) return (dir / l) * radius;
        return mk(radius, T(0), T(0));
      }
      case BOX:  // box_shape.rs:30-37
        return mk(dir.x >= T(0) ? half.x : -half.x, dir.y >= T(0) ? half.y : -half.y,
                  dir.z >= T(0) ? half.z : -half.z);
      case CAPSULE: {  // capsule.rs:38-56
        T l = len(dir);
        if (l > T(0)) {
          V3<T> tip = mk(T(0), dir.y >= T(0) ? half_height : -half_height, T(0));
          return tip + (dir / l) * radius;
        }
        return mk(radius, half_height, T(0));
      }
      case HULL: {  // convex_hull.rs:55-66 (strict > keeps the first maximum)
        V3<T> best = splat(T(0));
        T best_dot = -std::numeric_limits<T>::infinity();
        for (int i = 0; i < npts; ++i) {
          T d = dot(pts[i], dir);
          if (d > best_dot) {
            best_dot = d;
            best = pts[i];
          }
        }
        return best;
      }
      case PLANE: return splat(T(0));  // mod.rs:286-288
    }
    return splat(T(0));
  }
  // shapes/mod.rs:305-319 + per-shape core support
  V3<T> core_support_local(V3<T> dir) const {
    switch (kind) {
      case SPHERE: return splat(T(0));  // sphere.rs:23-26
      case BOX: {                       // box_shape.rs:42-53
        V3<T> h = mk(fmax_(half.x - margin(), T(0)), fmax_(half.y - margin(), T(0)),
                     fmax_(half.z - margin(), T(0)));
        return mk(dir.x >= T(0) ? h.x : -h.x, dir.y >= T(0) ? h.y : -h.y, dir.z >= T(0) ? h.z : -h.z);
      }
      case CAPSULE:  // capsule.rs:59-69
        return mk(T(0), dir.y >= T(0) ? half_height : -half_height, T(0));
      default: return support_local(dir);  // hull: core == full shape
    }
  }
  // shapes/mod.rs:323-332
  V3<T> support_world(V3<T> p, Quat<T> q, V3<T> dir) const {
    V3<T> ld = qrot(qconj(q), dir);
    V3<T> l = support_local(ld);
    return qrot(q, l) + p;
  }
  // shapes/mod.rs:337-346
  V3<T> core_support_world(V3<T> p, Quat<T> q, V3<T> dir) const {
    V3<T> ld = qrot(qconj(q), dir);
    V3<T> l = core_support_local(ld);
    return qrot(q, l) + p;
  }
  // shapes/mod.rs:350-375 + per-shape world_aabb()
  Aabb<T> world_aabb(V3<T> p, Quat<T> q) const {
    switch (kind) {
      case SPHERE: { V3<T> r = splat(radius); return Aabb<T>{p - r, p + r}; }  // sphere.rs:45-48
      case BOX: {  // box_shape.rs:57-72
        V3<T> mn = splat(std::numeric_limits<T>::infinity());
        V3<T> mx = splat(-std::numeric_limits<T>::infinity());
        for (int i = 0; i < 8; ++i) {
          V3<T> corner = mk((i & 1) == 0 ? half.x : -half.x, (i & 2) == 0 ? half.y : -half.y,
                            (i & 4) == 0 ? half.z : -half.z);
          V3<T> w = qrot(q, corner) + p;
          mn = vmin(mn, w);
          mx = vmax(mx, w);
        }
        return Aabb<T>{mn, mx};
      }
      case CAPSULE: {  // capsule.rs:74-79
        V3<T> top = qrot(q, mk(T(0), half_height, T(0))) + p;
        V3<T> bottom = qrot(q, mk(T(0), -half_height, T(0))) + p;
        V3<T> r = splat(radius);
        return Aabb<T>{vmin(top, bottom) - r, vmax(top, bottom) + r};
      }
      case HULL: {  // convex_hull.rs:70-72 -> shapes/mod.rs:448-469
        Quat<T> qi = qconj(q);
        V3<T> ux = qrot(qi, mk(T(1), T(0), T(0)));
        V3<T> uy = qrot(qi, mk(T(0), T(1), T(0)));
        V3<T> uz = qrot(qi, mk(T(0), T(0), T(1)));
        V3<T> mn = mk(p.x + dot(support_local(-ux), ux), p.y + dot(support_local(-uy), uy),
                      p.z + dot(support_local(-uz), uz));
        V3<T> mx = mk(p.x + dot(support_local(ux), ux), p.y + dot(support_local(uy), uy),
                      p.z + dot(support_local(uz), uz));
        return Aabb<T>{mn, mx};
      }
      case PLANE: return local_aabb();  // plane.rs:37-40
    }
    return Aabb<T>{};
  }
  // shapes/mod.rs:388-392
  T contact_breaking_threshold() const {
    Aabb<T> a = local_aabb();
    T disc = len(a.max - a.min) * T(0.5) + len(a.center());
    return disc * contact_breaking_threshold_const<T>();
  }
};

// ---------------------------------------------------------------- contacts --
// manifold.rs:12-112
template <class T>
struct ContactPoint {
  V3<T> position, normal;
  T penetration;
};
template <class T>
struct Manifold {
  int n = 0;
  ContactPoint<T> pts[8];
  void push(V3<T> p, V3<T> nrm, T d) { pts[n++] = ContactPoint<T>{p, nrm, d}; }
  // manifold.rs:68-74
  void flip() {
    for (int i = 0; i < n; ++i) pts[i].normal = -pts[i].normal;
  }
};

// manifold.rs:81-112 (ContactManifold::reduce), kept for completeness: the
// convex pairs restated here never produce more than four points.
template <class T>
void reduce_manifold(std::vector<ContactPoint<T>>& points, size_t max) {
  if (points.size() <= max) return;
  std::vector<ContactPoint<T>> cand = points;
  // sort_by(b.pen.total_cmp(a.pen)): stable, descending by penetration.
  std::stable_sort(cand.begin(), cand.end(), [](const ContactPoint<T>& a, const ContactPoint<T>& b) {
    return a.penetration > b.penetration;
  });
  std::vector<ContactPoint<T>> kept;
  kept.push_back(cand[0]);
  std::vector<bool> used(cand.size(), false);
  used[0] = true;
  while (kept.size() < max) {
    size_t best = 0;
    T best_sep = -std::numeric_limits<T>::infinity();
    for (size_t i = 0; i < cand.size(); ++i) {
      if (used[i]) continue;
      T sep = std::numeric_limits<T>::infinity();
      for (auto& q : kept) sep = fmin_(sep, len2(cand[i].position - q.position));
      if (sep > best_sep) {
        best_sep = sep;
        best = i;
      }
    }
    kept.push_back(cand[best]);
    used[best] = true;
  }
  points = kept;
}

// ------------------------------------------------------------ sphere.rs --
// narrowphase/sphere.rs:10-32
template <class T>
bool sphere_sphere(const Shape<T>& sa, V3<T> pa, const Shape<T>& sb, V3<T> pb, Manifold<T>& m) {
  V3<T> delta = pb - pa;
  T dist_sq = len2(delta);
  T radius_sum = sa.radius + sb.radius;
  if (dist_sq >= radius_sum * radius_sum) return false;
  T dist = std::sqrt(dist_sq);
  V3<T> normal = dist > ORL(1e-12) ? delta / dist : mk(T(0), T(1), T(0));
  T pen = radius_sum - dist;
  V3<T> position = pa + normal * (sa.radius - pen * T(0.5));
  m.push(position, normal, pen);
  return true;
}

// narrowphase/sphere.rs:38-82 (normal from the box to the sphere)
template <class T>
bool sphere_box(const Shape<T>& s, V3<T> ps, const Shape<T>& b, V3<T> pb, Quat<T> ob, T threshold,
                Manifold<T>& m) {
  V3<T> local = qrot(qconj(ob), ps - pb);
  V3<T> h = b.half;
  V3<T> clamped = vmin(vmax(local, -h), h);
  V3<T> delta = local - clamped;
  T dist_sq = len2(delta);
  T contact_dist = s.radius + threshold;
  if (dist_sq >= contact_dist * contact_dist) return false;
  T dist = std::sqrt(dist_sq);
  V3<T> normal_local;
  T pen;
  if (dist > ORL(1e-12)) {
    normal_local = delta / dist;
    pen = s.radius - dist;
  } else {
    T dx = h.x - std::fabs(local.x);
    T dy = h.y - std::fabs(local.y);
    T dz = h.z - std::fabs(local.z);
    V3<T> n;
    T d;
    if (dx <= dy && dx <= dz) {
      n = mk(signum(local.x), T(0), T(0));
      d = dx;
    } else if (dy <= dz) {
      n = mk(T(0), signum(local.y), T(0));
      d = dy;
    } else {
      n = mk(T(0), T(0), signum(local.z));
      d = dz;
    }
    normal_local = n;
    pen = s.radius + d;
  }
  V3<T> normal = qrot(ob, normal_local);
  V3<T> position = qrot(ob, clamped) + pb;
  m.push(position, normal, pen);
  return true;
}

// ------------------------------------------------------------- plane.rs --
// narrowphase/plane.rs:19-56
template <class T>
bool plane_convex(const Shape<T>& plane, V3<T> pp, Quat<T> pq, const Shape<T>& convex, V3<T> cp,
                  Quat<T> cq, T threshold, Manifold<T>& m) {
  V3<T> n = qrot(pq, plane.pnormal);
  T c = plane.pconst + dot(n, pp);
  V3<T> normal = dot(n, cp - pp) >= T(0) ? n : -n;
  if (!(normal.x == n.x && normal.y == n.y && normal.z == n.z)) c = -c;
  V3<T> deepest = convex.support_world(cp, cq, -normal);
  T depth = c - dot(normal, deepest);
  if (depth < -threshold) return false;
  m.push(deepest, normal, depth);
  return true;
}

// ----------------------------------------------------------- box_box.rs --
// narrowphase/box_box.rs:31-46
template <class T>
inline void line_closest_approach(V3<T> pa, V3<T> ua, V3<T> pb, V3<T> ub, T& alpha, T& beta) {
  V3<T> p = pb - pa;
  T uaub = dot(ua, ub);
  T q1 = dot(ua, p);
  T q2 = -dot(ub, p);
  T d = T(1) - uaub * uaub;
  if (d <= ORL(1e-4)) {
    alpha = T(0);
    beta = T(0);
  } else {
    d = T(1) / d;
    alpha = (q1 + uaub * q2) * d;
    beta = (uaub * q1 + q2) * d;
  }
}

template <class T>
struct ClipPoly {
  T p[8][2];
  int n = 0;
  bool overflow = false;
  void push(T a, T b) {
    if (n >= 8) { overflow = true; return; }  // reference: debug_assert + index panic
    p[n][0] = a;
    p[n][1] = b;
    ++n;
  }
};

// box_box.rs:92-114
template <class T>
void clip_poly(const ClipPoly<T>& src, int axis, T sign, T h, ClipPoly<T>& out) {
  int n = src.n;
  out.n = 0;
  out.overflow = src.overflow;
  for (int i = 0; i < n; ++i) {
    const T* cur = src.p[i];
    const T* next = src.p[(i + 1) % n];
    bool cur_in = sign * cur[axis] < h;
    bool next_in = sign * next[axis] < h;
    if (cur_in) out.push(cur[0], cur[1]);
    if (cur_in != next_in) {
      T t = (sign * h - cur[axis]) / (next[axis] - cur[axis]);
      T q[2];
      q[axis] = sign * h;
      q[1 - axis] = cur[1 - axis] + (next[1 - axis] - cur[1 - axis]) * t;
      out.push(q[0], q[1]);
    }
  }
}

// box_box.rs:131-174
template <class T>
int cull_points(const V3<T>* pts, const T* deps, int n, int maxc, V3<T>* opts, T* odeps) {
  if (n <= maxc) {
    for (int i = 0; i < n; ++i) { opts[i] = pts[i]; odeps[i] = deps[i]; }
    return n;
  }
  int i0 = 0;
  T deepest_depth = deps[0];
  for (int i = 1; i < n; ++i) {
    if (deps[i] >= deepest_depth) {
      i0 = i;
      deepest_depth = deps[i];
    }
  }
  int ri[8];
  T rd[8];
  int rn = 0;
  for (int i = 0; i < n; ++i) {
    if (i != i0) {
      ri[rn] = i;
      rd[rn] = len2(pts[i] - pts[i0]);
      ++rn;
    }
  }
  for (int j = 1; j < rn; ++j) {
    int ki = ri[j];
    T kd = rd[j];
    int k = j;
    while (k > 0 && rd[k - 1] < kd) {
      ri[k] = ri[k - 1];
      rd[k] = rd[k - 1];
      --k;
    }
    ri[k] = ki;
    rd[k] = kd;
  }
  opts[0] = pts[i0];
  odeps[0] = deps[i0];
  int m = 1;
  for (int j = 0; j < maxc - 1 && j < rn; ++j) {
    opts[m] = pts[ri[j]];
    odeps[m] = deps[ri[j]];
    ++m;
  }
  return m;
}

// box_box.rs:178-528 (ODE dBoxBox2 SAT, 15 axes, face clipping, <= 4 points)
template <class T>
bool box_box(const Shape<T>& a, V3<T> p1, Quat<T> oa, const Shape<T>& b, V3<T> p2, Quat<T> ob,
             Manifold<T>& out) {
  const T EPS = ORL(1e-6), FUDGE = ORL(1.05);
  M3<T> r1 = from_quat(oa), r2 = from_quat(ob);
  V3<T> ha = a.half, hb = b.half;
  V3<T> p = p2 - p1;
  M3<T> r1t = transpose(r1);
  V3<T> pp = mmul(r1t, p);
  M3<T> r = mmulm(r1t, r2);
  T r11 = r.c0.x, r12 = r.c1.x, r13 = r.c2.x;
  T r21 = r.c0.y, r22 = r.c1.y, r23 = r.c2.y;
  T r31 = r.c0.z, r32 = r.c1.z, r33 = r.c2.z;
  T q11 = std::fabs(r11), q12 = std::fabs(r12), q13 = std::fabs(r13);
  T q21 = std::fabs(r21), q22 = std::fabs(r22), q23 = std::fabs(r23);
  T q31 = std::fabs(r31), q32 = std::fabs(r32), q33 = std::fabs(r33);

  T s = -std::numeric_limits<T>::infinity();
  bool invert_normal = false;
  int code = 0;
  bool normal_from_r2 = false;
  int normal_axis = 0;
  bool normal_is_cross = false;
  V3<T> normal_c = splat(T(0));

#define OR_TST_FACE(e1, e2, fr2, ax, cc)          \
  {                                               \
    T ex1 = (e1);                                 \
    T s2 = std::fabs(ex1) - (e2);                 \
    if (s2 > T(0)) return false;                  \
    if (s2 > s) {                                 \
      s = s2;                                     \
      normal_is_cross = false;                    \
      normal_from_r2 = fr2;                       \
      normal_axis = ax;                           \
      invert_normal = ex1 < T(0);                 \
      code = cc;                                  \
    }                                             \
  }
#define OR_TST_CROSS(e1, e2, n1, n2, n3, cc)                   \
  {                                                            \
    T ex1 = (e1);                                              \
    T s2 = std::fabs(ex1) - (e2);                              \
    if (s2 > EPS) return false;                                \
    T N1 = (n1), N2 = (n2), N3 = (n3);                         \
    T l = std::sqrt(N1 * N1 + N2 * N2 + N3 * N3);              \
    if (l > EPS) {                                             \
      s2 = s2 / l;                                             \
      if (s2 * FUDGE > s) {                                    \
        s = s2;                                                \
        normal_is_cross = true;                                \
        normal_c = mk(N1 / l, N2 / l, N3 / l);                 \
        invert_normal = ex1 < T(0);                            \
        code = cc;                                             \
      }                                                        \
    }                                                          \
  }

  OR_TST_FACE(pp.x, ha.x + hb.x * q11 + hb.y * q12 + hb.z * q13, false, 0, 1);
  OR_TST_FACE(pp.y, ha.y + hb.x * q21 + hb.y * q22 + hb.z * q23, false, 1, 2);
  OR_TST_FACE(pp.z, ha.z + hb.x * q31 + hb.y * q32 + hb.z * q33, false, 2, 3);
  OR_TST_FACE(dot(r2.c0, p), ha.x * q11 + ha.y * q21 + ha.z * q31 + hb.x, true, 0, 4);
  OR_TST_FACE(dot(r2.c1, p), ha.x * q12 + ha.y * q22 + ha.z * q32 + hb.y, true, 1, 5);
  OR_TST_FACE(dot(r2.c2, p), ha.x * q13 + ha.y * q23 + ha.z * q33 + hb.z, true, 2, 6);

  const T fudge2 = ORL(1e-5);
  q11 += fudge2; q12 += fudge2; q13 += fudge2;
  q21 += fudge2; q22 += fudge2; q23 += fudge2;
  q31 += fudge2; q32 += fudge2; q33 += fudge2;

  OR_TST_CROSS(pp.z * r21 - pp.y * r31, ha.y * q31 + ha.z * q21 + hb.y * q13 + hb.z * q12, T(0), -r31, r21, 7);
  OR_TST_CROSS(pp.z * r22 - pp.y * r32, ha.y * q32 + ha.z * q22 + hb.x * q13 + hb.z * q11, T(0), -r32, r22, 8);
  OR_TST_CROSS(pp.z * r23 - pp.y * r33, ha.y * q33 + ha.z * q23 + hb.x * q12 + hb.y * q11, T(0), -r33, r23, 9);
  OR_TST_CROSS(pp.x * r31 - pp.z * r11, ha.x * q31 + ha.z * q11 + hb.y * q23 + hb.z * q22, r31, T(0), -r11, 10);
  OR_TST_CROSS(pp.x * r32 - pp.z * r12, ha.x * q32 + ha.z * q12 + hb.x * q23 + hb.z * q21, r32, T(0), -r12, 11);
  OR_TST_CROSS(pp.x * r33 - pp.z * r13, ha.x * q33 + ha.z * q13 + hb.x * q22 + hb.y * q21, r33, T(0), -r13, 12);
  OR_TST_CROSS(pp.y * r11 - pp.x * r21, ha.x * q21 + ha.y * q11 + hb.y * q33 + hb.z * q32, -r21, r11, T(0), 13);
  OR_TST_CROSS(pp.y * r12 - pp.x * r22, ha.x * q22 + ha.y * q12 + hb.x * q33 + hb.z * q31, -r22, r12, T(0), 14);
  OR_TST_CROSS(pp.y * r13 - pp.x * r23, ha.x * q23 + ha.y * q13 + hb.x * q32 + hb.y * q31, -r23, r13, T(0), 15);
#undef OR_TST_FACE
#undef OR_TST_CROSS

  V3<T> normal_sat = normal_is_cross ? mmul(r1, normal_c) : (normal_from_r2 ? col(r2, normal_axis) : col(r1, normal_axis));
  if (invert_normal) normal_sat = -normal_sat;
  T depth = -s;

  if (code > 6) {  // edge-edge: box_box.rs:402-423
    V3<T> pa = p1;
    for (int j = 0; j < 3; ++j) {
      V3<T> c = col(r1, j);
      T sign = dot(normal_sat, c) > T(0) ? T(1) : T(-1);
      pa = pa + (sign * comp(ha, j)) * c;
    }
    V3<T> pb = p2;
    for (int j = 0; j < 3; ++j) {
      V3<T> c = col(r2, j);
      T sign = dot(normal_sat, c) > T(0) ? T(-1) : T(1);
      pb = pb + (sign * comp(hb, j)) * c;
    }
    V3<T> ua = col(r1, (code - 7) / 3);
    V3<T> ub = col(r2, (code - 7) % 3);
    T alpha, beta;
    line_closest_approach(pa, ua, pb, ub, alpha, beta);
    pa = pa + ua * alpha;
    pb = pb + ub * beta;
    out.push(pb, normal_sat, depth);
    return true;
  }

  // Face-face: box_box.rs:425-527
  const M3<T>& ra = code <= 3 ? r1 : r2;
  const M3<T>& rb = code <= 3 ? r2 : r1;
  V3<T> pa_ref = code <= 3 ? p1 : p2;
  V3<T> pb_inc = code <= 3 ? p2 : p1;
  V3<T> sa = code <= 3 ? ha : hb;
  V3<T> sb = code <= 3 ? hb : ha;
  V3<T> normal2 = code <= 3 ? normal_sat : -normal_sat;
  V3<T> nr = mmul(transpose(rb), normal2);
  V3<T> anr = vabs(nr);
  int lanr, a1, a2;
  if (anr.y > anr.x) {
    if (anr.y > anr.z) { lanr = 1; a1 = 0; a2 = 2; } else { lanr = 2; a1 = 0; a2 = 1; }
  } else if (anr.x > anr.z) {
    lanr = 0; a1 = 1; a2 = 2;
  } else {
    lanr = 2; a1 = 0; a2 = 1;
  }
  V3<T> center = comp(nr, lanr) < T(0) ? (pb_inc - pa_ref) + comp(sb, lanr) * col(rb, lanr)
                                          : (pb_inc - pa_ref) - comp(sb, lanr) * col(rb, lanr);
  int code_n = code <= 3 ? code - 1 : code - 4;
  int code1, code2;
  if (code_n == 0) { code1 = 1; code2 = 2; } else if (code_n == 1) { code1 = 0; code2 = 2; } else { code1 = 0; code2 = 1; }
  T c1 = dot(center, col(ra, code1));
  T c2 = dot(center, col(ra, code2));
  T m11 = dot(col(ra, code1), col(rb, a1));
  T m12 = dot(col(ra, code1), col(rb, a2));
  T m21 = dot(col(ra, code2), col(rb, a1));
  T m22 = dot(col(ra, code2), col(rb, a2));
  T k1 = m11 * comp(sb, a1);
  T k2 = m21 * comp(sb, a1);
  T k3 = m12 * comp(sb, a2);
  T k4 = m22 * comp(sb, a2);
  ClipPoly<T> buf, tmp;
  buf.n = 4;
  buf.p[0][0] = c1 - k1 - k3; buf.p[0][1] = c2 - k2 - k4;
  buf.p[1][0] = c1 - k1 + k3; buf.p[1][1] = c2 - k2 + k4;
  buf.p[2][0] = c1 + k1 + k3; buf.p[2][1] = c2 + k2 + k4;
  buf.p[3][0] = c1 + k1 - k3; buf.p[3][1] = c2 + k2 - k4;
  T h0 = comp(sa, code1), h1 = comp(sa, code2);
  clip_poly(buf, 0, T(-1), h0, tmp);
  clip_poly(tmp, 0, T(1), h0, buf);
  clip_poly(buf, 1, T(-1), h1, tmp);
  clip_poly(tmp, 1, T(1), h1, buf);
  if (buf.n == 0) return false;

  T det1 = T(1) / (m11 * m22 - m12 * m21);
  T n11 = m11 * det1, n12 = m12 * det1, n21 = m21 * det1, n22 = m22 * det1;
  V3<T> pen_pts[8];
  T pen_d[8];
  int np = 0;
  for (int i = 0; i < buf.n; ++i) {
    T rx = buf.p[i][0], ry = buf.p[i][1];
    T kk1 = n22 * (rx - c1) - n12 * (ry - c2);
    T kk2 = -n21 * (rx - c1) + n11 * (ry - c2);
    V3<T> point = (center + kk1 * col(rb, a1)) + kk2 * col(rb, a2);
    T dep = comp(sa, code_n) - dot(normal2, point);
    if (dep >= T(0)) {
      pen_pts[np] = point;
      pen_d[np] = dep;
      ++np;
    }
  }
  if (np == 0) return false;
  V3<T> cp[4];
  T cd[4];
  int nc = cull_points(pen_pts, pen_d, np, 4, cp, cd);
  for (int i = 0; i < nc; ++i) {
    V3<T> pos = cp[i] + pa_ref;
    if (code > 3) pos = pos + normal_sat * cd[i];
    out.push(pos, normal_sat, cd[i]);
  }
  return true;
}

// ---------------------------------------------------------------- gjk.rs --
template <class T>
struct SupportVertex {
  V3<T> w, p_b;
};
template <class T>
struct GjkOut {
  int kind = 0;  // 0 = Separated, 1 = Band, 2 = Overlap
  T distance = 0;
  V3<T> normal{0, 0, 0}, point_on_b{0, 0, 0};
  V3<T> simplex[4];
  int iterations = 0;
  // Not in the reference (signed-distance callers only): when the walk
  // stopped at the 64-iteration cap, its closest iterate v (the least |v| of
  // all iterations; an f32 walk that misses the 1 - 1e-6 exit by rounding
  // cycles through worse simplices afterwards) and B witness.
  bool capped = false;
  // Also not in the reference: the walk had proven a core gap (an iterate
  // with delta / |v| > SD_GAP_PROOF) and later ended on a simplex that
  // enclosed or touched the origin, which rounding alone can produce (a
  // flat simplex, or coordinates far from the origin).  The closest iterate
  // is reported as for the cap, with capped and separated both set.
  bool separated = false;
  // Signed-distance walks: the least support-plane bound -delta / |v| over
  // the iterations.  Each iterate's support plane along -v lies that far
  // beyond the origin, so the core penetration depth is at most this.
  T depth_bound = std::numeric_limits<T>::infinity();
  V3<T> last_v{0, 0, 0}, last_pb{0, 0, 0};
  // the active simplex (w, b) at a Band exit, or of that iterate at the cap
  SupportVertex<T> active[4];
  int n_active = 0;
  // Also not in the reference (signed-distance walks): the cores touched
  // (|v|^2 below the touch gate) and the enclosing walk found no
  // tetrahedron, where the reference reports Separated.  The touching
  // simplex is kept in active[] with its B witness in last_pb, and
  // touch_dir is the walk's last support direction (the initial direction
  // on the first iteration), the normal reported for the touch.
  bool touch = false;
  V3<T> touch_dir{0, 0, 0};
};
template <class T>
struct Reduce {
  bool inside;
  V3<T> v, p_b;
};

// gjk.rs:398-464
template <class T>
void closest_triangle_bary(const SupportVertex<T>& a, const SupportVertex<T>& b, const SupportVertex<T>& c,
                           V3<T>& q, T wts[3]) {
  V3<T> ab = b.w - a.w, ac = c.w - a.w, ap = -a.w;
  T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= T(0) && d2 <= T(0)) { q = a.w; wts[0] = 1; wts[1] = 0; wts[2] = 0; return; }
  V3<T> bp = -b.w;
  T d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= T(0) && d4 <= d3) { q = b.w; wts[0] = 0; wts[1] = 1; wts[2] = 0; return; }
  T vc = d1 * d4 - d3 * d2;
  if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
    if (d1 == d3) { q = a.w; wts[0] = 1; wts[1] = 0; wts[2] = 0; return; }
    T t = d1 / (d1 - d3);
    q = a.w + ab * t; wts[0] = T(1) - t; wts[1] = t; wts[2] = 0; return;
  }
  V3<T> cp = -c.w;
  T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= T(0) && d5 <= d6) { q = c.w; wts[0] = 0; wts[1] = 0; wts[2] = 1; return; }
  T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    if (d2 == d6) { q = a.w; wts[0] = 1; wts[1] = 0; wts[2] = 0; return; }
    T t = d2 / (d2 - d6);
    q = a.w + ac * t; wts[0] = T(1) - t; wts[1] = 0; wts[2] = t; return;
  }
  T va = d3 * d6 - d5 * d4;
  if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
    if ((d4 - d3) + (d5 - d6) == T(0)) { q = b.w; wts[0] = 0; wts[1] = 1; wts[2] = 0; return; }
    T t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    q = b.w + (c.w - b.w) * t; wts[0] = 0; wts[1] = T(1) - t; wts[2] = t; return;
  }
  T sum = va + vb + vc;
  if (sum <= T(0) || std::isnan(sum)) {
    if (len2(a.w) <= len2(b.w) && len2(a.w) <= len2(c.w)) { q = a.w; wts[0] = 1; wts[1] = 0; wts[2] = 0; }
    else if (len2(b.w) <= len2(c.w)) { q = b.w; wts[0] = 0; wts[1] = 1; wts[2] = 0; }
    else { q = c.w; wts[0] = 0; wts[1] = 0; wts[2] = 1; }
    return;
  }
  T denom = T(1) / sum;
  T v = vb * denom, w = vc * denom;
  q = (a.w + ab * v) + ac * w;
  wts[0] = T(1) - v - w; wts[1] = v; wts[2] = w;
}

// gjk.rs:366-393
template <class T>
Reduce<T> reduce_segment(SupportVertex<T>* buf, int& n) {
  SupportVertex<T> a = buf[1], b = buf[0];
  V3<T> ab = b.w - a.w;
  T denom = len2(ab);
  if (denom < ORL(1e-24)) { n = 1; buf[0] = a; return {false, a.w, a.p_b}; }
  T t = -dot(a.w, ab) / denom;
  if (t <= T(0)) { n = 1; buf[0] = a; return {false, a.w, a.p_b}; }
  if (t >= T(1)) { n = 1; buf[0] = b; return {false, b.w, b.p_b}; }
  n = 2;
  buf[0] = a;
  buf[1] = b;
  return {false, a.w + ab * t, a.p_b + (b.p_b - a.p_b) * t};
}

// gjk.rs:466-486
template <class T>
Reduce<T> reduce_triangle(SupportVertex<T>* buf, int& n) {
  SupportVertex<T> a = buf[2], b = buf[1], c = buf[0];
  V3<T> v;
  T wt[3];
  closest_triangle_bary(a, b, c, v, wt);
  SupportVertex<T> kept[4] = {a, a, a, a};
  SupportVertex<T> src[3] = {a, b, c};
  int k = 0;
  for (int i = 0; i < 3; ++i)
    if (wt[i] > T(0)) kept[k++] = src[i];
  if (k == 0) { k = 1; kept[0] = a; }
  for (int i = 0; i < k; ++i) buf[i] = kept[i];
  n = k;
  V3<T> p_b = (a.p_b * wt[0] + b.p_b * wt[1]) + c.p_b * wt[2];
  return {false, v, p_b};
}

// gjk.rs:569-583
template <class T>
Reduce<T> triangle_step(SupportVertex<T>* buf, int& n, SupportVertex<T> a, SupportVertex<T> b, SupportVertex<T> c) {
  buf[0] = c;
  buf[1] = b;
  buf[2] = a;
  n = 3;
  return reduce_triangle(buf, n);
}

// gjk.rs:488-567
template <class T>
Reduce<T> reduce_tetrahedron(SupportVertex<T>* buf, int& n) {
  SupportVertex<T> a = buf[3], b = buf[2], c = buf[1], d = buf[0];
  V3<T> ab = b.w - a.w, ac = c.w - a.w, ad = d.w - a.w;
  if (std::fabs(dot(cross(ab, ac), ad)) < ORL(1e-12)) {
    SupportVertex<T> faces[4][3] = {{a, b, c}, {a, b, d}, {a, c, d}, {b, c, d}};
    int best_face = 0;
    T best_v2 = std::numeric_limits<T>::infinity();
    V3<T> best_q;
    T best_w[3];
    closest_triangle_bary(faces[0][0], faces[0][1], faces[0][2], best_q, best_w);
    for (int f = 1; f < 4; ++f) {
      V3<T> q;
      T w[3];
      closest_triangle_bary(faces[f][0], faces[f][1], faces[f][2], q, w);
      T v2 = len2(q);
      if (v2 < best_v2) {  // best_v2 starts at +inf (the face-0 result is never scored)
        best_v2 = v2;
        best_q = q;
        best_w[0] = w[0]; best_w[1] = w[1]; best_w[2] = w[2];
        best_face = f;
      }
    }
    SupportVertex<T> kept[4] = {buf[0], buf[0], buf[0], buf[0]};
    int k = 0;
    for (int i = 0; i < 3; ++i)
      if (best_w[i] > T(0)) kept[k++] = faces[best_face][i];
    if (k == 0) { k = 1; kept[0] = faces[best_face][0]; }
    for (int i = 0; i < k; ++i) buf[i] = kept[i];
    n = k;
    V3<T> p_b = (faces[best_face][0].p_b * best_w[0] + faces[best_face][1].p_b * best_w[1]) +
                faces[best_face][2].p_b * best_w[2];
    return {false, best_q, p_b};
  }
  V3<T> abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
  V3<T> ao = -a.w, bo = -b.w;
  V3<T> bcd = cross(c.w - b.w, d.w - b.w);
  if (dot(abc, ao) * dot(abc, ad) < T(0)) return triangle_step(buf, n, a, b, c);
  if (dot(acd, ao) * dot(acd, ab) < T(0)) return triangle_step(buf, n, a, c, d);
  if (dot(adb, ao) * dot(adb, ac) < T(0)) return triangle_step(buf, n, a, d, b);
  if (dot(bcd, bo) * dot(bcd, a.w - b.w) < T(0)) return triangle_step(buf, n, b, c, d);
  return {true, splat(T(0)), splat(T(0))};
}

// gjk.rs:352-364
template <class T>
Reduce<T> reduce(SupportVertex<T>* buf, int& n) {
  switch (n) {
    case 1: return {false, buf[0].w, buf[0].p_b};
    case 2: return reduce_segment(buf, n);
    case 3: return reduce_triangle(buf, n);
    default: return reduce_tetrahedron(buf, n);
  }
}

// gjk.rs:668-680
template <class T>
V3<T> perpendicular(V3<T> v) {
  V3<T> X = mk(T(1), T(0), T(0)), Y = mk(T(0), T(1), T(0));
  V3<T> helper = std::fabs(v.x) < ORL(0.9) ? X : Y;
  V3<T> perp = cross(v, helper);
  if (len2(perp) > T(0)) return normalize(perp);
  V3<T> alt = std::fabs(v.x) < ORL(0.9) ? Y : X;
  perp = cross(v, alt);
  if (len2(perp) > T(0)) return normalize(perp);
  return mk(T(0), T(0), T(1));
}

// Strict-walk simplex, gjk.rs:17-51 (pts[0] is the most recent point).
template <class T>
struct Simplex {
  V3<T> pts[4];
  int n = 0;
  void init(V3<T> p) { pts[0] = pts[1] = pts[2] = pts[3] = p; n = 1; }
  void add(V3<T> v) {
    for (int i = n; i > 0; --i) pts[i] = pts[i - 1];
    pts[0] = v;
    ++n;
  }
  void set2(V3<T> a, V3<T> b) { pts[0] = a; pts[1] = b; n = 2; }
  void set3(V3<T> a, V3<T> b, V3<T> c) { pts[0] = a; pts[1] = b; pts[2] = c; n = 3; }
  void set1(V3<T> a) { pts[0] = a; n = 1; }
};

// gjk.rs:785-807
template <class T>
bool shrink_or_normal(Simplex<T>& s, V3<T>& dir, V3<T> closest, V3<T> normal, V3<T> n0, V3<T> n1) {
  T span2 = T(0);
  for (int i = 0; i < s.n; ++i) span2 = std::fmax(span2, len2(s.pts[i]));
  if (len2(closest) < span2 * ORL(1e-12)) {
    dir = normal;
  } else {
    s.set2(n0, n1);
    dir = -closest;
  }
  return false;
}

// gjk.rs:624-646
template <class T>
bool solve_segment(Simplex<T>& s, V3<T>& dir) {
  V3<T> a = s.pts[1], b = s.pts[0];
  V3<T> ab = b - a, ao = -a;
  if (dot(ab, ao) > T(0)) {
    V3<T> cr = cross(ab, ao);
    if (len2(cr) > len2(ab) * len2(ao) * ORL(1e-10)) dir = cross(cr, ab);
    else dir = perpendicular(ab);
  } else {
    s.set1(a);
    dir = ao;
  }
  return false;
}

// gjk.rs:687-773
template <class T>
bool solve_triangle(Simplex<T>& s, V3<T>& dir) {
  V3<T> a = s.pts[2], b = s.pts[1], c = s.pts[0];
  V3<T> ab = b - a, ac = c - a;
  V3<T> normal = cross(ab, ac);
  V3<T> ap = -a;
  T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= T(0) && d2 <= T(0)) { s.set1(a); dir = -a; return false; }
  V3<T> bp = -b;
  T d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= T(0) && d4 <= d3) { s.set1(b); dir = -b; return false; }
  T vc = d1 * d4 - d3 * d2;
  if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
    T v = d1 == d3 ? T(0) : d1 / (d1 - d3);
    V3<T> closest = a + ab * v;
    return shrink_or_normal(s, dir, closest, normal, b, a);
  }
  V3<T> cp = -c;
  T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= T(0) && d5 <= d6) { s.set1(c); dir = -c; return false; }
  T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    T w = d2 == d6 ? T(0) : d2 / (d2 - d6);
    V3<T> closest = a + ac * w;
    return shrink_or_normal(s, dir, closest, normal, c, a);
  }
  T va = d3 * d6 - d5 * d4;
  if (va <= T(0) && (d4 - d3) >= T(0) && (d5 - d6) >= T(0)) {
    T w = (d4 - d3) + (d5 - d6) == T(0) ? T(0) : (d4 - d3) / ((d4 - d3) + (d5 - d6));
    V3<T> closest = b + (c - b) * w;
    return shrink_or_normal(s, dir, closest, normal, c, b);
  }
  T sum = va + vb + vc;
  if (sum <= T(0) || std::isnan(sum)) { dir = normal; return false; }
  T denom = T(1) / sum;
  T v = vb * denom, w = vc * denom;
  V3<T> closest = (a + ab * v) + ac * w;
  T span2 = std::fmax(std::fmax(len2(a), len2(b)), len2(c));
  if (len2(closest) < span2 * ORL(1e-12)) dir = normal;
  else dir = -closest;
  return false;
}

// gjk.rs:849-858
template <class T> inline V3<T> toward_origin(V3<T> n, V3<T> t) { return dot(n, t) < T(0) ? -n : n; }

// gjk.rs:811-846
template <class T>
bool solve_tetrahedron(Simplex<T>& s, V3<T>& dir) {
  V3<T> a = s.pts[3], b = s.pts[2], c = s.pts[1], d = s.pts[0];
  V3<T> ab = b - a, ac = c - a, ad = d - a, ao = -a, bo = -b;
  V3<T> abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
  V3<T> bcd = cross(b - c, d - c);
  if (dot(abc, ao) * dot(abc, ad) < T(0)) { s.set3(c, b, a); dir = toward_origin(abc, ao); return false; }
  if (dot(acd, ao) * dot(acd, ab) < T(0)) { s.set3(d, c, a); dir = toward_origin(acd, ao); return false; }
  if (dot(adb, ao) * dot(adb, ac) < T(0)) { s.set3(d, b, a); dir = toward_origin(adb, ao); return false; }
  if (dot(bcd, bo) * dot(bcd, a - b) < T(0)) { s.set3(d, c, b); dir = toward_origin(bcd, bo); return false; }
  return true;
}

// gjk.rs:609-620
template <class T>
bool solve(Simplex<T>& s, V3<T>& dir) {
  switch (s.n) {
    case 2: return solve_segment(s, dir);
    case 3: return solve_triangle(s, dir);
    case 4: return solve_tetrahedron(s, dir);
    default: dir = -s.pts[0]; return false;
  }
}

// gjk.rs:299-342 (strict origin-enclosing walk).
template <class T, class F>
bool enclosing_walk(F support, V3<T> fallback, const V3<T>* seed, Simplex<T>& simplex) {
  V3<T> initial = (seed && len2(*seed) > ORL(1e-12)) ? *seed : fallback;
  simplex.init(support(initial));
  V3<T> dir = -simplex.pts[0];
  for (int it = 0; it < 64; ++it) {
    T dir2 = len2(dir);
    if (dir2 <= T(0) || std::isnan(dir2)) dir = perpendicular(simplex.pts[0]);
    V3<T> a = support(dir);
    if (dot(a, dir) < T(0)) return false;
    simplex.add(a);
    if (solve(simplex, dir)) return true;
    V3<T> reference = simplex.n >= 2 ? simplex.pts[0] - simplex.pts[simplex.n - 1] : simplex.pts[0];
    if (len2(dir) < len2(reference) * ORL(1e-12)) dir = perpendicular(reference);
  }
  return false;
}

// gjk.rs:76-89 core Minkowski support.
template <class T>
struct PairPose {
  const Shape<T>* a;
  V3<T> pa;
  Quat<T> qa;
  const Shape<T>* b;
  V3<T> pb;
  Quat<T> qb;
  SupportVertex<T> core_support(V3<T> dir) const {
    V3<T> p_a = a->core_support_world(pa, qa, dir);
    V3<T> p_b = b->core_support_world(pb, qb, -dir);
    return SupportVertex<T>{p_a - p_b, p_b};
  }
  V3<T> full_support(V3<T> dir) const {
    return a->support_world(pa, qa, dir) - b->support_world(pb, qb, -dir);
  }
};

// Signed-distance walks only: a core gap larger than this (metres) proven
// by one iterate (support plane delta / |v|) outranks a later enclosure.
constexpr double SD_GAP_PROOF = 1e-4;
// Signed-distance EPA only: a core depth more than this (EPA's own exit
// tolerance, epa.rs:35) beyond the walk's depth bound is contradicted.
constexpr double SD_EPA_SLACK = 1e-4;

// gjk.rs:171-288 -- the margin-core distance walk.  sd_walk (not in the
// reference; signed-distance callers, threshold = +inf) keeps the closest
// iterate and reports it when the walk hits the cap or ends enclosing the
// origin after it proved a gap.
template <class T>
GjkOut<T> gjk_contact(const PairPose<T>& pr, const V3<T>* seed, T threshold, bool sd_walk = false) {
  GjkOut<T> out;
  T margin_a = pr.a->margin(), margin_b = pr.b->margin();
  T margin_sum = margin_a + margin_b;
  V3<T> center_dir = pr.pb - pr.pa;
  V3<T> fallback = len2(center_dir) > ORL(1e-12) ? center_dir : mk(T(1), T(0), T(0));
  V3<T> initial = (seed && len2(*seed) > ORL(1e-12)) ? *seed : fallback;
  SupportVertex<T> first = pr.core_support(initial);
  SupportVertex<T> buf[4] = {first, first, first, first};
  int n = 1;
  V3<T> dir = -first.w;
  T touch2 = ORL(1e-12) * std::fmax(len2(first.w), T(1));
  V3<T> search = initial;  // the direction of the latest support query
  const T REL_ERROR2 = ORL(1e-6);
  T best_v2 = std::numeric_limits<T>::infinity();
  const T gap2 = T(SD_GAP_PROOF) * T(SD_GAP_PROOF);
  bool proven_gap = false;
  for (int it = 0; it < 64; ++it) {
    out.iterations = it + 1;
    Reduce<T> r = reduce(buf, n);
    V3<T> v = r.v;
    T v2 = len2(v);
    if (sd_walk && proven_gap && (r.inside || v2 < touch2)) {
      out.kind = 0;
      out.capped = true;
      out.separated = true;
      return out;
    }
    if (r.inside) {
      out.kind = 2;
      for (int i = 0; i < 4; ++i) out.simplex[i] = buf[i].w;
      return out;
    }
    if (v2 < touch2) {
      Simplex<T> s;
      auto sup = [&](V3<T> d) { return pr.core_support(d).w; };
      bool ok = enclosing_walk<T>(sup, fallback, &dir, s);
      if (ok && s.n == 4) {
        out.kind = 2;
        for (int i = 0; i < 4; ++i) out.simplex[i] = s.pts[i];
      } else {
        out.kind = 0;
        if (sd_walk) {
          out.touch = true;
          out.touch_dir = search;
          for (int i = 0; i < n; ++i) out.active[i] = buf[i];
          out.n_active = n;
          out.last_pb = r.p_b;
        }
      }
      return out;
    }
    dir = -v;
    search = dir;
    SupportVertex<T> w = pr.core_support(dir);
    T delta = dot(v, w.w);
    T band = margin_sum + threshold;
    if (delta > T(0) && delta * delta > v2 * band * band) { out.kind = 0; return out; }
    if (delta > T(0) && delta * delta > v2 * gap2) proven_gap = true;
    if (sd_walk) {
      T bound = -delta / std::sqrt(v2);
      if (bound < out.depth_bound) out.depth_bound = bound;
    }
    bool duplicate = false;
    for (int i = 0; i < n; ++i)
      if (abs_diff_eq(w.w, buf[i].w, ORL(1e-6))) duplicate = true;
    if (duplicate || delta >= v2 * (T(1) - REL_ERROR2)) {
      T core_distance = std::sqrt(v2);
      T distance = core_distance - margin_sum;
      if (distance < threshold) {
        V3<T> normal = dir / core_distance;
        out.kind = 1;
        out.distance = distance;
        out.normal = normal;
        out.point_on_b = r.p_b - normal * margin_b;
        for (int i = 0; i < n; ++i) out.active[i] = buf[i];
        out.n_active = n;
        return out;
      }
      out.kind = 0;
      return out;
    }
    if (!(v2 >= best_v2)) {  // the closest iterate so far: the first of equals, and NaN passes
      best_v2 = v2;
      for (int i = 0; i < n; ++i) out.active[i] = buf[i];
      out.n_active = n;
      out.last_v = v;
      out.last_pb = r.p_b;
    }
    buf[n] = w;
    ++n;
  }
  out.kind = 0;
  out.capped = true;
  return out;
}

// gjk.rs:590-607 (full-shape strict walk).
template <class T>
bool gjk_intersect(const PairPose<T>& pr, const V3<T>* seed, Simplex<T>& s) {
  V3<T> center_dir = pr.pb - pr.pa;
  V3<T> fallback = len2(center_dir) > ORL(1e-12) ? center_dir : mk(T(1), T(0), T(0));
  auto sup = [&](V3<T> d) { return pr.full_support(d); };
  return enclosing_walk<T>(sup, fallback, seed, s);
}

// ---------------------------------------------------------------- epa.rs --
enum EpaStatus { EPA_CONVERGED = 0, EPA_DEGENERATE = 1, EPA_ITERATION_LIMIT = 2, EPA_VERTEX_LIMIT = 3 };

template <class T>
struct Face {
  int a, b, c;
  V3<T> normal;
  T distance, plane_distance;
};

// epa.rs:283-303
template <class T>
bool edge_distance(V3<T> pa, V3<T> pb, V3<T> normal, T& out) {
  V3<T> ba = pb - pa;
  V3<T> n_ab = cross(ba, normal);
  if (dot(pa, n_ab) < T(0)) {
    T ba_l2 = len2(ba);
    T a_dot_ba = dot(pa, ba);
    T b_dot_ba = dot(pb, ba);
    if (a_dot_ba > T(0)) out = len(pa);
    else if (b_dot_ba < T(0)) out = len(pb);
    else {
      T a_dot_b = dot(pa, pb);
      // Rust's f32::max(0.0) returns the non-NaN operand: an overflowed
      // (inf - inf) quotient gives 0, not NaN
      T q = (len2(pa) * len2(pb) - a_dot_b * a_dot_b) / ba_l2;
      out = std::sqrt(q > T(0) ? q : T(0));
    }
    return true;
  }
  return false;
}

// epa.rs:262-278
template <class T>
T face_distance(V3<T> normal, V3<T> pa, V3<T> pb, V3<T> pc) {
  T d;
  if (edge_distance(pa, pb, normal, d)) return d;
  if (edge_distance(pb, pc, normal, d)) return d;
  if (edge_distance(pc, pa, normal, d)) return d;
  return dot(pa, normal);
}

// epa.rs:395-398
template <class T> inline bool face_visible(const Face<T>& f, V3<T> p) {
  return dot(f.normal, p) - f.plane_distance > ORL(1e-5);
}
template <> inline bool face_visible<double>(const Face<double>& f, V3<double> p) {
  return dot(f.normal, p) - f.plane_distance > 1e-14;
}

// epa.rs:309-388 (two-phase, transactional expansion)
template <class T>
bool expand(std::vector<Face<T>>& faces, const std::vector<V3<T>>& verts, std::vector<std::pair<int, int>>& boundary,
            V3<T> new_point, int new_idx) {
  boundary.clear();
  bool any_visible = false;
  for (const auto& f : faces) {
    if (!face_visible(f, new_point)) continue;
    any_visible = true;
    int e[3][2] = {{f.a, f.b}, {f.b, f.c}, {f.c, f.a}};
    for (auto& uv : e) {
      int u = uv[0], v = uv[1];
      int found = -1;
      for (size_t i = 0; i < boundary.size(); ++i)
        if (boundary[i].first == v && boundary[i].second == u) { found = (int)i; break; }
      if (found >= 0) {
        boundary[found] = boundary.back();  // swap_remove
        boundary.pop_back();
      } else {
        boundary.push_back({u, v});
      }
    }
  }
  if (!any_visible) return true;
  for (auto& uv : boundary) {
    V3<T> cr = cross(verts[uv.second] - verts[uv.first], new_point - verts[uv.first]);
    if (len(cr) < ORL(1e-12)) return false;
  }
  std::vector<Face<T>> kept;
  kept.reserve(faces.size());
  for (const auto& f : faces)
    if (!face_visible(f, new_point)) kept.push_back(f);
  faces.swap(kept);
  for (auto& uv : boundary) {
    int u = uv.first, v = uv.second;
    V3<T> cr = cross(verts[v] - verts[u], new_point - verts[u]);
    T l = len(cr);
    V3<T> normal = cr / l;
    T distance = face_distance(normal, verts[u], verts[v], new_point);
    faces.push_back(Face<T>{u, v, new_idx, normal, distance, dot(normal, verts[u])});
  }
  boundary.clear();
  return true;
}

// epa.rs:212-224 (first minimum of |distance|)
template <class T>
int find_best(const std::vector<Face<T>>& faces) {
  int best = 0;
  for (size_t i = 1; i < faces.size(); ++i)
    if (std::fabs(faces[best].distance) > std::fabs(faces[i].distance)) best = (int)i;
  return best;
}

// epa.rs:226-257
template <class T>
void build_initial_faces(const V3<T>* s, std::vector<Face<T>>& out) {
  const int quads[4][4] = {{0, 1, 2, 3}, {0, 3, 1, 2}, {0, 2, 3, 1}, {1, 3, 2, 0}};
  for (auto& q : quads) {
    int a = q[0], b = q[1], c = q[2], opp = q[3];
    V3<T> normal = cross(s[b] - s[a], s[c] - s[a]);
    if (dot(normal, s[opp] - s[a]) > T(0)) {
      normal = -normal;
      std::swap(b, c);
    }
    T l = len(normal);
    if (l < ORL(1e-12)) continue;
    normal = normal / l;
    out.push_back(Face<T>{a, b, c, normal, face_distance(normal, s[a], s[b], s[c]), dot(normal, s[a])});
  }
}

template <class T>
struct EpaOut {
  int status;
  V3<T> normal;
  T depth;
  int iterations;
  int vertices;
  V3<T> face[3];  // vertices of the reported face (a, b, c order)
};

// epa.rs:127-205
template <class T, class F>
EpaOut<T> epa_penetration(const V3<T>* simplex, F support) {
  const T EPSILON = std::is_same<T, float>::value ? (T)1e-4f : (T)1e-12;
  std::vector<V3<T>> verts(simplex, simplex + 4);
  std::vector<Face<T>> faces;
  std::vector<std::pair<int, int>> boundary;
  build_initial_faces(simplex, faces);
  if (faces.empty()) return EpaOut<T>{EPA_DEGENERATE, splat(T(0)), T(0), 0, 4, {}};
  int status = EPA_ITERATION_LIMIT;
  V3<T> dirs[6] = {mk(T(1), T(0), T(0)), mk(T(-1), T(0), T(0)), mk(T(0), T(1), T(0)),
                   mk(T(0), T(-1), T(0)), mk(T(0), T(0), T(1)), mk(T(0), T(0), T(-1))};
  bool seeded_ok = true;
  for (auto& d : dirs) {
    int idx = (int)verts.size();
    V3<T> p = support(d);
    verts.push_back(p);
    if (!expand(faces, verts, boundary, p, idx)) {
      status = EPA_DEGENERATE;
      seeded_ok = false;
      break;
    }
  }
  (void)seeded_ok;
  int it = 0;
  for (; it < 255; ++it) {
    int bi = find_best(faces);
    V3<T> bn = faces[bi].normal;
    T bd = faces[bi].distance;
    V3<T> np = support(bn);
    T wdist = dot(bn, np) - bd;
    if (wdist <= EPSILON)
      return EpaOut<T>{EPA_CONVERGED, bn, bd, it + 1, (int)verts.size(),
                       {verts[faces[bi].a], verts[faces[bi].b], verts[faces[bi].c]}};
    if ((int)verts.size() >= 128) { status = EPA_VERTEX_LIMIT; break; }
    int nidx = (int)verts.size();
    verts.push_back(np);
    if (!expand(faces, verts, boundary, np, nidx)) { status = EPA_DEGENERATE; break; }
  }
  int bi = find_best(faces);
  return EpaOut<T>{status, faces[bi].normal, faces[bi].distance, it, (int)verts.size(),
                   {verts[faces[bi].a], verts[faces[bi].b], verts[faces[bi].c]}};
}

// ------------------------------------------------------ convex_convex.rs --
// narrowphase/convex_convex.rs:30-155
template <class T>
bool convex_convex(const PairPose<T>& pr, const V3<T>* seed, T threshold, Manifold<T>& m, int* path = nullptr) {
  GjkOut<T> g = gjk_contact(pr, seed, threshold);
  if (g.kind == 0) { if (path) *path = 0; return false; }
  if (g.kind == 1) {
    if (path) *path = 1;
    m.push(g.point_on_b, g.normal, -g.distance);
    return true;
  }
  if (path) *path = 2;
  auto sup = [&](V3<T> d) { return pr.core_support(d).w; };
  EpaOut<T> e = epa_penetration<T>(g.simplex, sup);
  if (e.depth <= T(0)) return false;
  T depth = (e.depth + pr.a->margin()) + pr.b->margin();
  V3<T> point_a = pr.a->support_world(pr.pa, pr.qa, e.normal);
  V3<T> position = point_a - e.normal * depth;
  m.push(position, e.normal, depth);
  return true;
}

// ------------------------------------------------------ signed distance --
// Not a reference function: the signed distance between two convex shapes
// built from the reference's own walks.  The margin-core walk (gjk.rs:171-288)
// runs with threshold = +inf, so a separated pair ends in its Band arm at any
// distance; overlapping cores go to EPA as convex_convex.rs:117-154 does.
// Three exits report the walk's closest iterate the way the Band arm would:
// the 64-iteration cap; an enclosure after the walk proved a core gap
// (SD_GAP_PROOF); and an EPA depth beyond the walk's own depth bound
// (SD_EPA_SLACK), which only a polytope corrupted by rounding can give (the
// reference's f32 EPA, started from the nearly flat simplex of just-touching
// hulls, has converged on faces 0.02-0.18 m deep).  Outputs: sd (> 0
// separated), unit normal A -> B, witness points on A and on B with
// pa = pb - n sd.
enum SdStatus {
  SD_WALK = 0,
  SD_EPA = 1,
  SD_CAPPED = 2,
  SD_TOUCH = 3,            // cores touch, no enclosing tetrahedron: sd = -(margins), normal = last search direction
  SD_EPA_UNCONVERGED = 4,
  SD_SEPARATED = 5,        // the walk proved a core gap, then enclosed the origin by rounding
  SD_EPA_CONTRADICTED = 6  // EPA's depth exceeds the walk's depth bound
};
template <class T>
struct SdOut {
  int status = SD_TOUCH;
  T sd = 0;
  V3<T> n{0, 0, 0}, pa{0, 0, 0}, pb{0, 0, 0};
};
// The normal reported for touching cores: the walk's last support direction
// with its components along the touching simplex removed (a triangle's face
// normal, a segment's perpendicular; a single point keeps the direction).
template <class T>
V3<T> sd_touch_normal(const GjkOut<T>& g) {
  V3<T> s = g.touch_dir, n = s;
  if (g.n_active == 3) {
    V3<T> c = cross(g.active[1].w - g.active[0].w, g.active[2].w - g.active[0].w);
    T c2 = len2(c);
    if (c2 > T(0)) n = c * (dot(c, s) / c2);
  } else if (g.n_active == 2) {
    V3<T> e = g.active[1].w - g.active[0].w;
    T e2 = len2(e);
    if (e2 > T(0)) n = s - e * (dot(s, e) / e2);
  }
  T n2 = len2(n);
  if (!(n2 > T(0))) n = s, n2 = len2(s);
  return n / std::sqrt(n2);
}
template <class T>
SdOut<T> signed_distance(const PairPose<T>& pr) {
  SdOut<T> o;
  T ma = pr.a->margin(), mb = pr.b->margin();
  GjkOut<T> g = gjk_contact(pr, (const V3<T>*)nullptr, std::numeric_limits<T>::infinity(), true);
  EpaOut<T> e{};
  bool contradicted = false;
  if (g.kind == 2) {
    auto sup = [&](V3<T> d) { return pr.core_support(d).w; };
    e = epa_penetration<T>(g.simplex, sup);
    contradicted = e.depth > g.depth_bound + T(SD_EPA_SLACK);
  }
  if (g.kind == 1 || (g.kind == 0 && g.capped) || contradicted) {
    if (g.kind == 1) {
      o.status = SD_WALK;
      o.sd = g.distance;
      o.n = g.normal;
      o.pb = g.point_on_b;
    } else {
      T core_distance = std::sqrt(len2(g.last_v));
      o.status = contradicted ? SD_EPA_CONTRADICTED : (g.separated ? SD_SEPARATED : SD_CAPPED);
      o.sd = core_distance - (ma + mb);
      o.n = (-g.last_v) / core_distance;
      o.pb = g.last_pb - o.n * mb;
    }
    o.pa = o.pb - o.n * o.sd;
    return o;
  }
  if (g.kind == 2) {
    T depth = (e.depth + ma) + mb;
    o.status = e.status == EPA_CONVERGED ? SD_EPA : SD_EPA_UNCONVERGED;
    o.sd = -depth;
    o.n = e.normal;
    o.pa = pr.a->support_world(pr.pa, pr.qa, e.normal);
    o.pb = o.pa - e.normal * depth;
    return o;
  }
  o.status = SD_TOUCH;
  o.sd = -(ma + mb);
  if (g.touch) {
    // the cores touch: core distance 0, normal along the walk's last support
    // direction, witnesses from the touching simplex
    o.n = sd_touch_normal(g);
    o.pb = g.last_pb - o.n * mb;
    o.pa = o.pb - o.n * o.sd;
  }
  return o;
}

// ------------------------------------------------ signed distance adjoint --
// Not a reference function: the vector-Jacobian product of signed_distance
// with respect to both poses (position xyz, quaternion xyzw; 14 values), for
// an f64 reference gradient.  The Band arm differentiates the affine
// least-norm point of the active simplex; the EPA arm differentiates the
// reported face's plane distance and normal; support points are fixed local
// points placed by qrot(q, x) + p.  Pieces where a support feature changes
// are not smooth; this is the derivative of the piece the pose is on.
template <class T>
void qrot_vjp(Quat<T> q, V3<T> x, V3<T> g, T gq[4]) {
  V3<T> b = mk(q.x, q.y, q.z);
  T w = q.w;
  V3<T> gb = ((g * dot(x, b) + x * dot(b, g)) - b * dot(x, g) + cross(x, g) * w) * T(2);
  gq[0] += gb.x;
  gq[1] += gb.y;
  gq[2] += gb.z;
  gq[3] += T(2) * (w * dot(x, g) + dot(cross(b, x), g));
}
template <class T>
V3<T> sd_to_local(Quat<T> q, V3<T> p, V3<T> world) {
  T q2 = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
  return qrot(qconj(q), world - p) * (T(1) / (q2 * q2));
}
// the finite set a core support point is drawn from (local coordinates)
template <class T>
int core_set_size(const Shape<T>& s) {
  return s.kind == HULL ? s.npts : (s.kind == BOX ? 8 : (s.kind == CAPSULE ? 2 : 1));
}
template <class T>
V3<T> core_set_point(const Shape<T>& s, int i) {
  if (s.kind == HULL) return s.pts[i];
  if (s.kind == BOX) {
    V3<T> h = mk(fmax_(s.half.x - s.margin(), T(0)), fmax_(s.half.y - s.margin(), T(0)),
                 fmax_(s.half.z - s.margin(), T(0)));
    return mk((i & 1) ? -h.x : h.x, (i & 2) ? -h.y : h.y, (i & 4) ? -h.z : h.z);
  }
  if (s.kind == CAPSULE) return mk(T(0), i ? -s.half_height : s.half_height, T(0));
  return splat(T(0));
}
// split a Minkowski vertex into local support points (exact recomputation)
template <class T>
bool split_vertex(const PairPose<T>& pr, V3<T> w, V3<T>& xa, V3<T>& xb) {
  for (int i = 0; i < core_set_size(*pr.a); ++i) {
    V3<T> la = core_set_point(*pr.a, i);
    V3<T> wa = qrot(pr.qa, la) + pr.pa;
    for (int j = 0; j < core_set_size(*pr.b); ++j) {
      V3<T> lb = core_set_point(*pr.b, j);
      V3<T> d = wa - (qrot(pr.qb, lb) + pr.pb);
      if (d.x == w.x && d.y == w.y && d.z == w.z) {
        xa = la;
        xb = lb;
        return true;
      }
    }
  }
  return false;
}
// returns false when a face vertex could not be split (gradient left 0)
template <class T>
bool sd_vjp(const PairPose<T>& pr, const T* g10, T* g14) {
  for (int k = 0; k < 14; ++k) g14[k] = T(0);
  T ma = pr.a->margin(), mb = pr.b->margin();
  V3<T> gpa_pos = splat(T(0)), gpb_pos = splat(T(0));
  T gqa[4] = {0, 0, 0, 0}, gqb[4] = {0, 0, 0, 0};
  V3<T> gn = mk(g10[1], g10[2], g10[3]), gpa = mk(g10[4], g10[5], g10[6]), gpb = mk(g10[7], g10[8], g10[9]);
  GjkOut<T> g = gjk_contact(pr, (const V3<T>*)nullptr, std::numeric_limits<T>::infinity(), true);
  bool ok = true;
  EpaOut<T> e{};
  bool contradicted = false;
  if (g.kind == 2) {
    auto sup = [&](V3<T> d) { return pr.core_support(d).w; };
    e = epa_penetration<T>(g.simplex, sup);
    contradicted = e.depth > g.depth_bound + T(SD_EPA_SLACK);
  }
  // touching cores (SD_TOUCH): the witnesses move with their bodies and the
  // reported normal is held fixed, d sd = n . (d b - d a) over the touching
  // simplex's weights
  bool touch = g.kind == 0 && g.touch;
  if (g.kind == 1 || (g.kind == 0 && g.capped) || contradicted || touch) {
    int k = g.n_active;
    V3<T> w[3], b[3];
    for (int i = 0; i < k && i < 3; ++i) {
      w[i] = g.active[i].w;
      b[i] = g.active[i].p_b;
    }
    T lam[3] = {1, 0, 0}, mu0 = 0, mu1 = 0, i00 = 0, i01 = 0, i11 = 0;
    V3<T> D0 = splat(T(0)), D1 = splat(T(0));
    if (k == 2) {
      D0 = w[1] - w[0];
      i00 = T(1) / len2(D0);
      mu0 = -dot(D0, w[0]) * i00;
      lam[0] = T(1) - mu0;
      lam[1] = mu0;
    } else if (k == 3) {
      D0 = w[1] - w[0];
      D1 = w[2] - w[0];
      T m00 = len2(D0), m01 = dot(D0, D1), m11 = len2(D1);
      T det = m00 * m11 - m01 * m01;
      i00 = m11 / det;
      i01 = -m01 / det;
      i11 = m00 / det;
      T r0 = -dot(D0, w[0]), r1 = -dot(D1, w[0]);
      mu0 = i00 * r0 + i01 * r1;
      mu1 = i01 * r0 + i11 * r1;
      lam[0] = (T(1) - mu0) - mu1;
      lam[1] = mu0;
      lam[2] = mu1;
    }
    V3<T> v = splat(T(0));
    for (int i = 0; i < k; ++i) v = v + w[i] * lam[i];
    T core = std::sqrt(len2(v));
    V3<T> n = touch ? sd_touch_normal(g) : (-v) / core;
    T sd = touch ? -(ma + mb) : core - (ma + mb);
    V3<T> g_pbt = gpb + gpa;
    V3<T> g_nt = (gn - gpa * sd) - g_pbt * mb;
    T g_sdt = g10[0] - dot(n, gpa);
    V3<T> g_v = touch ? n * (-g_sdt) : (v / core) * g_sdt - (g_nt - n * dot(n, g_nt)) / core;
    V3<T> gw[3], gb[3];
    T glam[3];
    for (int i = 0; i < 3; ++i) {
      gw[i] = g_v * lam[i];
      gb[i] = g_pbt * lam[i];
      glam[i] = i < k ? dot(w[i], g_v) + dot(b[i], g_pbt) : T(0);
    }
    if (k >= 2 && !touch) {
      T gm0 = glam[1] - glam[0], gm1 = k == 3 ? glam[2] - glam[0] : T(0);
      T u0 = k == 3 ? i00 * gm0 + i01 * gm1 : i00 * gm0;
      T u1 = k == 3 ? i01 * gm0 + i11 * gm1 : T(0);
      V3<T> Du = D0 * u0 + D1 * u1;
      V3<T> gD0 = -(v * u0 + Du * mu0), gD1 = -(v * u1 + Du * mu1);
      gw[0] = ((gw[0] - Du) - gD0) - (k == 3 ? gD1 : splat(T(0)));
      gw[1] = gw[1] + gD0;
      if (k == 3) gw[2] = gw[2] + gD1;
    }
    for (int i = 0; i < k; ++i) {
      V3<T> a_i = w[i] + b[i];
      gpa_pos = gpa_pos + gw[i];
      qrot_vjp(pr.qa, sd_to_local(pr.qa, pr.pa, a_i), gw[i], gqa);
      V3<T> gbi = gb[i] - gw[i];
      gpb_pos = gpb_pos + gbi;
      qrot_vjp(pr.qb, sd_to_local(pr.qb, pr.pb, b[i]), gbi, gqb);
    }
  } else if (g.kind == 2) {
    T depth = (e.depth + ma) + mb;
    V3<T> n = e.normal;
    V3<T> g_pat = gpa + gpb;
    V3<T> g_nt = gn - gpb * depth;
    T g_depth = -g10[0] - dot(n, gpb);
    // witness on A: supportA(n), with the rounded shapes' n-dependence
    gpa_pos = gpa_pos + g_pat;
    V3<T> ld = qrot(qconj(pr.qa), n);
    if (pr.a->kind == SPHERE || pr.a->kind == CAPSULE) {
      V3<T> tip = pr.a->kind == CAPSULE ? mk(T(0), ld.y >= T(0) ? pr.a->half_height : -pr.a->half_height, T(0))
                                        : splat(T(0));
      qrot_vjp(pr.qa, tip, g_pat, gqa);
      T r = pr.a->radius, ng = dot(n, g_pat);
      gqa[0] += T(2) * r * ng * pr.qa.x;
      gqa[1] += T(2) * r * ng * pr.qa.y;
      gqa[2] += T(2) * r * ng * pr.qa.z;
      gqa[3] += T(2) * r * ng * pr.qa.w;
      g_nt = g_nt + (g_pat - n * ng) * r;
    } else {
      qrot_vjp(pr.qa, pr.a->support_local(ld), g_pat, gqa);
    }
    V3<T> xa[3], xb[3];
    for (int i = 0; i < 3; ++i) ok = split_vertex(pr, e.face[i], xa[i], xb[i]) && ok;
    if (ok) {
      V3<T> w0 = e.face[0], w1 = e.face[1], w2 = e.face[2];
      g_nt = g_nt + w0 * g_depth;
      V3<T> e1 = w1 - w0, e2 = w2 - w0;
      V3<T> cc = cross(e1, e2);
      T cl = std::sqrt(len2(cc));
      V3<T> nn = cc / cl;
      V3<T> g_c = (g_nt - nn * dot(nn, g_nt)) / cl;
      V3<T> g_e1 = cross(e2, g_c), g_e2 = cross(g_c, e1);
      V3<T> gw[3] = {(n * g_depth - g_e1) - g_e2, g_e1, g_e2};
      for (int i = 0; i < 3; ++i) {
        gpa_pos = gpa_pos + gw[i];
        gpb_pos = gpb_pos - gw[i];
        qrot_vjp(pr.qa, xa[i], gw[i], gqa);
        qrot_vjp(pr.qb, xb[i], -gw[i], gqb);
      }
    }
  }
  g14[0] = gpa_pos.x; g14[1] = gpa_pos.y; g14[2] = gpa_pos.z;
  for (int k = 0; k < 4; ++k) g14[3 + k] = gqa[k];
  g14[7] = gpb_pos.x; g14[8] = gpb_pos.y; g14[9] = gpb_pos.z;
  for (int k = 0; k < 4; ++k) g14[10 + k] = gqb[k];
  return ok;
}

// ---------------------------------------------------------- dispatch ----
// narrowphase/mod.rs:52-57
template <class T>
T pair_threshold(const Shape<T>& a, const Shape<T>& b) {
  return fmin_(a.contact_breaking_threshold(), b.contact_breaking_threshold());
}

// narrowphase/mod.rs:63-286, restricted to the convex primitives and planes
// (compound and concave arms are outside this path).  `threshold_override`
// (NaN = unused) replaces the pair band, which lets the signed-distance
// callers pass +inf to the same walk.
template <class T>
bool collide_seeded(const Shape<T>& sa, V3<T> pa, Quat<T> qa, const Shape<T>& sb, V3<T> pb, Quat<T> qb,
                    const V3<T>* seed, Manifold<T>& m, int* path = nullptr) {
  T threshold = pair_threshold(sa, sb);
  m.n = 0;
  if (path) *path = 3;
  if (sa.kind == PLANE && sb.kind == PLANE) return false;
  if (sa.kind == PLANE) return plane_convex(sa, pa, qa, sb, pb, qb, threshold, m);
  if (sb.kind == PLANE) {
    bool r = plane_convex(sb, pb, qb, sa, pa, qa, threshold, m);
    if (r) m.flip();
    return r;
  }
  if (sa.kind == SPHERE && sb.kind == SPHERE) return sphere_sphere(sa, pa, sb, pb, m);
  if (sa.kind == SPHERE && sb.kind == BOX) {
    bool r = sphere_box(sa, pa, sb, pb, qb, threshold, m);
    if (r) m.flip();
    return r;
  }
  if (sa.kind == BOX && sb.kind == SPHERE) return sphere_box(sb, pb, sa, pa, qa, threshold, m);
  if (sa.kind == BOX && sb.kind == BOX) return box_box(sa, pa, qa, sb, pb, qb, m);
  PairPose<T> pr{&sa, pa, qa, &sb, pb, qb};
  return convex_convex(pr, seed, threshold, m, path);
}

// --------------------------------------------------- persistent manifold --
// PersistentManifold (crates/bullet-collision/src/manifold.rs:124-386) and
// the rigid-pair arm of the world step's pair pass
// (crates/bullet-dynamics/src/world/step.rs:660-854), in f32 like the
// reference.  Quaternion products follow glam's scalar Quat * Quat order.

// manifold.rs:146-152, world.rs:62
constexpr float PM_REDRIFT_T = 0.005f;
constexpr float PM_REDRIFT_T2 = PM_REDRIFT_T * PM_REDRIFT_T;
constexpr float PM_REDRIFT_ANGLE = 0.01f;
constexpr int PM_CACHE = 4;

// glam Quat * Quat (Hamilton product, scalar path).
inline Quat<float> qmul(Quat<float> a, Quat<float> b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

// manifold.rs:154-158: 2 acos(clamp(w, -1, 1))
inline float rotation_angle(Quat<float> q) {
  float w = q.w < -1.0f ? -1.0f : (q.w > 1.0f ? 1.0f : q.w);
  return 2.0f * std::acos(w);
}

// manifold.rs:124-144
struct PPoint {
  V3<float> local_a, local_b, normal;
  float penetration = 0.0f, normal_impulse = 0.0f;
  uint32_t lifetime = 0;
};

// manifold.rs:160-184 (points may exceed four inside capture only)
struct PMan {
  PPoint pts[PM_CACHE + 8];
  int n = 0;
  V3<float> dpa{0.0f, 0.0f, 0.0f}, dpb{0.0f, 0.0f, 0.0f};
  Quat<float> dqa{0.0f, 0.0f, 0.0f, 1.0f}, dqb{0.0f, 0.0f, 0.0f, 1.0f};
  bool initialized = false;
  float threshold = 0.02f;  // CONTACT_BREAKING_THRESHOLD, manifold.rs:8
};

// A refreshed point as the step emits it (ContactPoint, manifold.rs:12-41).
struct PContact {
  V3<float> position, normal;
  float penetration, normal_impulse;
  uint32_t lifetime;
};

// refresh (manifold.rs:252-298): re-project, drop separated or sideways-
// drifted points, emit the survivors in cache order.
inline int pm_refresh(PMan& m, V3<float> pa, Quat<float> qa, V3<float> pb, Quat<float> qb, PContact* out) {
  float thr = m.threshold;
  float thr2 = thr * thr;
  int w = 0;
  for (int r = 0; r < m.n; ++r) {
    PPoint p = m.pts[r];
    V3<float> wa = pa + qrot(qa, p.local_a);
    V3<float> wb = pb + qrot(qb, p.local_b);
    V3<float> d = wa - wb;
    float pen = dot(d, p.normal);
    if (pen < -thr) continue;
    V3<float> lat = d - p.normal * pen;
    if (len2(lat) > thr2) continue;
    p.penetration = pen;
    p.lifetime += 1;
    if (out) out[w] = PContact{wb, p.normal, pen, p.normal_impulse, p.lifetime};
    m.pts[w++] = p;
  }
  m.n = w;
  return w;
}

// needs_redetect (manifold.rs:300-318)
inline bool pm_needs_redetect(const PMan& m, V3<float> pa, Quat<float> qa, V3<float> pb, Quat<float> qb) {
  if (!m.initialized) return true;
  V3<float> drift = (pb - m.dpb) - (pa - m.dpa);
  if (len2(drift) > PM_REDRIFT_T2) return true;
  Quat<float> ra = qmul(qa, qconj(m.dqa));
  Quat<float> rb = qmul(qb, qconj(m.dqb));
  return rotation_angle(ra) + rotation_angle(rb) > PM_REDRIFT_ANGLE;
}

// f32::total_cmp key (larger key = larger value)
inline int64_t total_key(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  int32_t s = (int32_t)u;
  return s < 0 ? -(int64_t)(u & 0x7fffffffu) - 1 : (int64_t)u;
}

// capture (manifold.rs:320-376): merge by A-side anchor (< threshold) and
// normal (dot >= 0.9), inherit the impulse, append the rest; deepest four
// on overflow (stable sort, total order).
inline void pm_capture(PMan& m, const ContactPoint<float>* cps, int ncp, V3<float> pa, Quat<float> qa, V3<float> pb,
                       Quat<float> qb, float threshold) {
  m.threshold = threshold;
  Quat<float> inv_a = qconj(qa), inv_b = qconj(qb);
  float thr2 = threshold * threshold;
  for (int k = 0; k < ncp; ++k) {
    const ContactPoint<float>& cp = cps[k];
    V3<float> wb = cp.position;
    V3<float> wa = cp.position + cp.normal * cp.penetration;
    PPoint pt;
    pt.local_a = qrot(inv_a, wa - pa);
    pt.local_b = qrot(inv_b, wb - pb);
    pt.normal = cp.normal;
    pt.penetration = cp.penetration;
    int slot = -1;
    for (int o = 0; o < m.n; ++o) {
      if (len2(m.pts[o].local_a - pt.local_a) < thr2 && dot(m.pts[o].normal, pt.normal) >= 0.9f) {
        slot = o;
        break;
      }
    }
    if (slot >= 0) {
      pt.normal_impulse = m.pts[slot].normal_impulse;
      m.pts[slot] = pt;
    } else {
      m.pts[m.n++] = pt;
    }
  }
  if (m.n > PM_CACHE) {
    std::stable_sort(m.pts, m.pts + m.n,
                     [](const PPoint& a, const PPoint& b) { return total_key(a.penetration) > total_key(b.penetration); });
    m.n = PM_CACHE;
  }
  m.dpa = pa;
  m.dqa = qa;
  m.dpb = pb;
  m.dqb = qb;
  m.initialized = true;
}

// seed (manifold.rs:378-385)
inline bool pm_seed(const PMan& m, V3<float>* out) {
  if (m.n == 0 || !(len2(m.pts[0].normal) > 0.5f)) return false;
  *out = m.pts[0].normal;
  return true;
}

// Per-body flags of the pair pass (the body state step.rs reads).
enum : uint8_t { PM_STATIC = 1, PM_ACTIVE = 2, PM_FAST = 4 };
// Routes of a pair through the pass.
enum : uint8_t { PM_FILTERED = 0, PM_SLEEPING = 1, PM_STATIC_PAIR = 2, PM_SERVED = 3, PM_DETECTED = 4, PM_REMOVED = 5 };

struct PmPassOut {
  uint8_t route = PM_FILTERED;
  int n_out = 0;
  PContact out[PM_CACHE];
  V3<float> anchor_a[PM_CACHE];  // pa + qa * cached local_a of each emitted point
  bool keep = false;              // the pair has a cache entry after the pass
};

// The rigid-pair arm of the pair pass for one pair (step.rs:660-854):
// group/mask filter (step.rs:660-664), sleeping and static pairs keep their
// entry untouched (step.rs:683-704, 711-713), the refresh path unless the
// pair is fast and a refreshed point is deeper than the slop
// (step.rs:715-776), otherwise collide_seeded with the cache's seed and
// capture + refresh (step.rs:777-847) or removal (step.rs:848-852).
// Emitted impulses are scaled by the warm-start factor (step.rs:751, 821).
inline PmPassOut pm_pair_pass(bool has_entry, PMan& m, const Shape<float>& sa, V3<float> pa, Quat<float> qa,
                              const Shape<float>& sb, V3<float> pb, Quat<float> qb, uint8_t fa, uint8_t fb,
                              uint32_t gma, uint32_t gmb, float slop, float warm) {
  PmPassOut r;
  uint32_t ga = gma & 0xffffu, ma = gma >> 16, gb = gmb & 0xffffu, mb = gmb >> 16;
  if ((ga & mb) == 0 || (gb & ma) == 0) return r;
  if (!(fa & PM_ACTIVE) && !(fb & PM_ACTIVE)) {
    r.route = PM_SLEEPING;
    r.keep = has_entry;
    return r;
  }
  if ((fa & PM_STATIC) && (fb & PM_STATIC)) {
    r.route = PM_STATIC_PAIR;
    r.keep = has_entry;
    return r;
  }
  bool fast = (fa & PM_FAST) || (fb & PM_FAST);
  auto emit = [&](int n) {
    r.n_out = n;
    for (int k = 0; k < n; ++k) {
      r.out[k].normal_impulse = r.out[k].normal_impulse * warm;
      r.anchor_a[k] = pa + qrot(qa, m.pts[k].local_a);
    }
  };
  if (has_entry && !pm_needs_redetect(m, pa, qa, pb, qb)) {
    int n = pm_refresh(m, pa, qa, pb, qb, r.out);
    bool deep = false;
    for (int k = 0; k < n; ++k) deep = deep || r.out[k].penetration > slop;
    if (!fast || !deep) {
      emit(n);
      r.route = PM_SERVED;
      r.keep = true;
      return r;
    }
  }
  V3<float> seed;
  bool has_seed = has_entry && pm_seed(m, &seed);
  Manifold<float> man;
  collide_seeded(sa, pa, qa, sb, pb, qb, has_seed ? &seed : nullptr, man);
  if (man.n > 0) {
    if (!has_entry) m = PMan();
    pm_capture(m, man.pts, man.n, pa, qa, pb, qb, pair_threshold(sa, sb));
    int n = pm_refresh(m, pa, qa, pb, qb, r.out);
    emit(n);
    r.route = PM_DETECTED;
    r.keep = true;
  } else {
    r.route = PM_REMOVED;
    r.keep = false;
  }
  return r;
}

// Entry serialisation shared with the device cache: 64 floats (256 B).
//   point k (k < 4): [12k .. 12k+2] local_a, [12k+3] penetration,
//                    [12k+4 .. +6] local_b, [12k+7] normal_impulse,
//                    [12k+8 .. +10] normal, [12k+11] lifetime (u32 bits)
//   [48..50] detect position A, [51] threshold, [52..55] detect orientation A,
//   [56..58] detect position B, [59] n | initialized << 8 (u32 bits),
//   [60..63] detect orientation B
inline void pm_store(const PMan& m, float* e) {
  std::memset(e, 0, 64 * sizeof(float));
  for (int k = 0; k < m.n && k < PM_CACHE; ++k) {
    const PPoint& p = m.pts[k];
    float* q = e + 12 * k;
    q[0] = p.local_a.x; q[1] = p.local_a.y; q[2] = p.local_a.z; q[3] = p.penetration;
    q[4] = p.local_b.x; q[5] = p.local_b.y; q[6] = p.local_b.z; q[7] = p.normal_impulse;
    q[8] = p.normal.x; q[9] = p.normal.y; q[10] = p.normal.z;
    std::memcpy(q + 11, &p.lifetime, 4);
  }
  e[48] = m.dpa.x; e[49] = m.dpa.y; e[50] = m.dpa.z; e[51] = m.threshold;
  e[52] = m.dqa.x; e[53] = m.dqa.y; e[54] = m.dqa.z; e[55] = m.dqa.w;
  e[56] = m.dpb.x; e[57] = m.dpb.y; e[58] = m.dpb.z;
  uint32_t nb = (uint32_t)m.n | (m.initialized ? 0x100u : 0u);
  std::memcpy(e + 59, &nb, 4);
  e[60] = m.dqb.x; e[61] = m.dqb.y; e[62] = m.dqb.z; e[63] = m.dqb.w;
}

inline void pm_load(const float* e, PMan& m) {
  uint32_t nb;
  std::memcpy(&nb, e + 59, 4);
  m = PMan();
  m.n = (int)(nb & 0xffu);
  m.initialized = (nb & 0x100u) != 0;
  for (int k = 0; k < m.n; ++k) {
    const float* q = e + 12 * k;
    PPoint& p = m.pts[k];
    p.local_a = mk(q[0], q[1], q[2]);
    p.penetration = q[3];
    p.local_b = mk(q[4], q[5], q[6]);
    p.normal_impulse = q[7];
    p.normal = mk(q[8], q[9], q[10]);
    std::memcpy(&p.lifetime, q + 11, 4);
  }
  m.dpa = mk(e[48], e[49], e[50]);
  m.threshold = e[51];
  m.dqa = Quat<float>{e[52], e[53], e[54], e[55]};
  m.dpb = mk(e[56], e[57], e[58]);
  m.dqb = Quat<float>{e[60], e[61], e[62], e[63]};
}

// ------------------------------------------------------- triangle meshes --
// TriangleMesh (crates/bullet-collision/src/shapes/triangle_mesh.rs:84-154),
// its Bvh (shapes/bvh.rs:40-269) and convex_mesh
// (narrowphase/convex_mesh.rs:24-272), f32.
struct TriMesh {
  std::vector<V3<float>> v;
  std::vector<uint32_t> idx;  // 3 per triangle
  std::vector<Aabb<float>> tri_aabb;
  std::vector<V3<float>> tri_pts;  // 3 per triangle: the three-vertex hulls
  Aabb<float> bounds{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
  struct Node {
    Aabb<float> box;
    bool leaf;
    uint32_t left, right, start, count;
  };
  std::vector<Node> nodes;
  std::vector<uint32_t> order;
  size_t ntri() const { return idx.size() / 3; }
  V3<float> vert(uint32_t t, int k) const { return v[idx[3 * t + k]]; }

  // bvh.rs:252-255
  V3<float> center(uint32_t t) const { return ((vert(t, 0) + vert(t, 1)) + vert(t, 2)) * (1.0f / 3.0f); }
  // bvh.rs:204-222
  Aabb<float> slice_aabb(uint32_t start, uint32_t count) const {
    bool any = false;
    Aabb<float> a{{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    for (uint32_t k = start; k < start + count; ++k)
      for (int c = 0; c < 3; ++c) {
        V3<float> p = vert(order[k], c);
        if (!any) {
          a = Aabb<float>{p, p};
          any = true;
        } else {
          a = Aabb<float>{vmin(a.min, p), vmax(a.max, p)};  // util.rs:41-46
        }
      }
    return a;
  }
  // bvh.rs:170-202 (longest axis bvh.rs:257-269; stable sort on total order)
  uint32_t build_node(uint32_t start, uint32_t count) {
    Aabb<float> box = slice_aabb(start, count);
    uint32_t index = (uint32_t)nodes.size();
    nodes.push_back(Node{box, true, 0, 0, start, count});
    if (count <= 4) return index;
    V3<float> e = (box.max - box.min) * 0.5f;
    int axis = 0;
    float mxv = e.x;
    if (e.y > mxv) {
      axis = 1;
      mxv = e.y;
    }
    if (e.z > mxv) axis = 2;
    std::stable_sort(order.begin() + start, order.begin() + start + count, [&](uint32_t a, uint32_t b) {
      return total_key(comp(center(a), axis)) < total_key(comp(center(b), axis));
    });
    uint32_t mid = count / 2;
    uint32_t l = build_node(start, mid);
    uint32_t r = build_node(start + mid, count - mid);
    nodes[index].leaf = false;
    nodes[index].left = l;
    nodes[index].right = r;
    return index;
  }
  // triangle_mesh.rs:84-123 (one part)
  void build(const float* verts, uint32_t nv, const uint32_t* ind, uint32_t nt) {
    v.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) v[i] = mk(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    idx.assign(ind, ind + 3 * (size_t)nt);
    order.resize(nt);
    for (uint32_t t = 0; t < nt; ++t) order[t] = t;
    nodes.clear();
    build_node(0, nt);
    tri_aabb.resize(nt);
    tri_pts.resize(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t) {
      V3<float> a = vert(t, 0), b = vert(t, 1), c = vert(t, 2);
      tri_pts[3 * t] = a;
      tri_pts[3 * t + 1] = b;
      tri_pts[3 * t + 2] = c;
      tri_aabb[t] = Aabb<float>{vmin(a, vmin(b, c)), vmax(a, vmax(b, c))};
      bounds = t == 0 ? tri_aabb[0] : Aabb<float>{vmin(bounds.min, tri_aabb[t].min), vmax(bounds.max, tri_aabb[t].max)};
    }
  }
  // bvh.rs:58-79: depth first, right child popped first
  void query(const Aabb<float>& q, std::vector<uint32_t>& out) const {
    out.clear();
    if (nodes.empty()) return;
    std::vector<uint32_t> stack{0u};
    while (!stack.empty()) {
      uint32_t k = stack.back();
      stack.pop_back();
      const Node& n = nodes[k];
      if (!n.box.overlaps(q)) continue;
      if (!n.leaf) {
        stack.push_back(n.left);
        stack.push_back(n.right);
      } else {
        for (uint32_t i = n.start; i < n.start + n.count; ++i) out.push_back(order[i]);
      }
    }
  }
  // shapes/mod.rs:388-392 over the local bounds
  float threshold() const {
    float disc = len((bounds.max - bounds.min)) * 0.5f + len(bounds.center());
    return disc * 0.02f;
  }
};

// convex_mesh.rs:237-255
inline float seg_closest_sq(V3<float> from, V3<float> to, V3<float> p, V3<float>& nearest) {
  V3<float> v = to - from;
  float t = dot(v, p - from);
  if (t > 0.0f) {
    float dvv = dot(v, v);
    if (t < dvv) t /= dvv;
    else t = 1.0f;
  } else {
    t = 0.0f;
  }
  nearest = from + v * t;
  return len2(p - nearest);
}

// convex_mesh.rs:257-272
inline bool point_in_triangle(V3<float> p, V3<float> v1, V3<float> v2, V3<float> v3, V3<float> normal) {
  V3<float> e1 = v2 - v1, e2 = v3 - v2, e3 = v1 - v3;
  float r1 = dot(cross(e1, normal), p - v1);
  float r2 = dot(cross(e2, normal), p - v2);
  float r3 = dot(cross(e3, normal), p - v3);
  return (r1 > 0.0f && r2 > 0.0f && r3 > 0.0f) || (r1 <= 0.0f && r2 <= 0.0f && r3 <= 0.0f);
}

// convex_mesh.rs:168-235: point (mesh frame), centre-pointing normal, penetration
inline bool sphere_triangle_contact(V3<float> center, float radius, float threshold, V3<float> a, V3<float> b,
                                    V3<float> c, V3<float>& point, V3<float>& normal_out, float& pen) {
  const float EPS = std::numeric_limits<float>::epsilon();
  float rt = radius + threshold;
  V3<float> normal = cross(b - a, c - a);
  float l2 = len2(normal);
  if (l2 < EPS * EPS) return false;
  normal = normal * (1.0f / std::sqrt(l2));
  float dist = dot(normal, center - a);
  if (dist < 0.0f) {
    dist = -dist;
    normal = -normal;
  }
  if (dist >= rt) return false;
  V3<float> cp;
  if (point_in_triangle(center, a, b, c, normal)) {
    cp = center - normal * dist;
  } else {
    float min_d2 = rt * rt;
    bool found = false;
    V3<float> ends[3][2] = {{a, b}, {b, c}, {c, a}};
    for (int k = 0; k < 3; ++k) {
      V3<float> on;
      float d2 = seg_closest_sq(ends[k][0], ends[k][1], center, on);
      if (d2 < min_d2) {
        min_d2 = d2;
        cp = on;
        found = true;
      }
    }
    if (!found) return false;
  }
  V3<float> to_c = center - cp;
  float d2 = len2(to_c);
  if (d2 >= rt * rt) return false;
  point = cp;
  if (d2 > EPS) {
    float d = std::sqrt(d2);
    normal_out = to_c * (1.0f / d);
    pen = radius - d;
  } else {
    normal_out = normal;
    pen = radius;
  }
  return true;
}

// ContactManifold::reduce (manifold.rs:76-110): deepest first, then greedily
// the candidate farthest (min squared distance) from those kept.
inline void reduce_manifold(std::vector<ContactPoint<float>>& pts, size_t max) {
  if (pts.size() <= max) return;
  std::vector<ContactPoint<float>> cand = pts;
  std::stable_sort(cand.begin(), cand.end(), [](const ContactPoint<float>& a, const ContactPoint<float>& b) {
    return total_key(a.penetration) > total_key(b.penetration);
  });
  std::vector<ContactPoint<float>> kept{cand[0]};
  std::vector<bool> used(cand.size(), false);
  used[0] = true;
  while (kept.size() < max) {
    size_t best = 0;
    float best_sep = -std::numeric_limits<float>::infinity();
    for (size_t i = 0; i < cand.size(); ++i) {
      if (used[i]) continue;
      float sep = std::numeric_limits<float>::infinity();
      for (const auto& q : kept) sep = std::fmin(sep, len2(cand[i].position - q.position));
      if (sep > best_sep) {
        best_sep = sep;
        best = i;
      }
    }
    kept.push_back(cand[best]);
    used[best] = true;
  }
  pts = kept;
}

// convex_mesh (convex_mesh.rs:24-145) without internal-edge info: the convex
// is A, the mesh B, the normal points from the convex to the mesh.
inline bool convex_mesh(const Shape<float>& convex, V3<float> pc, Quat<float> qc, const TriMesh& mesh, V3<float> pm,
                        Quat<float> qm, const V3<float>* seed, float threshold,
                        std::vector<ContactPoint<float>>& out) {
  out.clear();
  Aabb<float> cw = convex.world_aabb(pc, qc);
  Quat<float> inv_m = qconj(qm);
  V3<float> lmin = splat(std::numeric_limits<float>::infinity());
  V3<float> lmax = splat(-std::numeric_limits<float>::infinity());
  V3<float> c = cw.center(), e = (cw.max - cw.min) * 0.5f;
  for (int i = 0; i < 8; ++i) {
    V3<float> corner = mk((i & 1) == 0 ? c.x + e.x : c.x - e.x, (i & 2) == 0 ? c.y + e.y : c.y - e.y,
                          (i & 4) == 0 ? c.z + e.z : c.z - e.z);
    V3<float> p = qrot(inv_m, corner - pm);
    lmin = vmin(lmin, p);
    lmax = vmax(lmax, p);
  }
  Aabb<float> query{lmin, lmax};
  std::vector<uint32_t> cand;
  mesh.query(query, cand);
  if (convex.kind == SPHERE) {
    V3<float> center = qrot(inv_m, pc - pm);
    float r = convex.radius;
    for (uint32_t t : cand) {
      if (!mesh.tri_aabb[t].overlaps(query)) continue;
      V3<float> pl, nl;
      float pen;
      if (sphere_triangle_contact(center, r, threshold, mesh.vert(t, 0), mesh.vert(t, 1), mesh.vert(t, 2), pl, nl, pen))
        out.push_back(ContactPoint<float>{pm + qrot(qm, pl), qrot(qm, -nl), pen});
    }
  } else {
    for (uint32_t t : cand) {
      if (!mesh.tri_aabb[t].overlaps(query)) continue;
      Shape<float> tri = Shape<float>::hull(&mesh.tri_pts[3 * (size_t)t], 3);
      PairPose<float> pr{&convex, pc, qc, &tri, pm, qm};
      Manifold<float> m;
      if (convex_convex(pr, seed, threshold, m))
        for (int k = 0; k < m.n; ++k) out.push_back(m.pts[k]);
    }
  }
  if (out.empty()) return false;
  reduce_manifold(out, 4);
  return true;
}

// ------------------------------------------------------------ CCD sweeps --
// convex_sweep.rs (swept_sphere_cast :35-57, convex_cast :357-385) over the
// simplex reduction closest_origin of gjk.rs:864-1043, f32.
struct CoSimplex {
  V3<float> p[4];
  int n = 0;
};

// gjk.rs:887-906
inline V3<float> co_segment(V3<float> a, V3<float> b, CoSimplex& k) {
  V3<float> ab = b - a;
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
inline V3<float> co_triangle(V3<float> a, V3<float> b, V3<float> c, CoSimplex& k) {
  V3<float> ab = b - a, ac = c - a, ap = -a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) { k.p[0] = a; k.n = 1; return a; }
  V3<float> bp = -b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) { k.p[0] = b; k.n = 1; return b; }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    if (d1 == d3) { k.p[0] = a; k.n = 1; return a; }
    float v = d1 / (d1 - d3);
    k.p[0] = b; k.p[1] = a; k.n = 2;
    return a + ab * v;
  }
  V3<float> cp = -c;
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
  if (sum <= 0.0f || std::isnan(sum)) {
    CoSimplex kab, kbc, kca;
    V3<float> qab = co_segment(a, b, kab), qbc = co_segment(b, c, kbc), qca = co_segment(c, a, kca);
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

// gjk.rs:987-1043
inline V3<float> co_tetra(V3<float> a, V3<float> b, V3<float> c, V3<float> d, CoSimplex& k) {
  V3<float> ao = -a, ab = b - a, ac = c - a, ad = d - a;
  if (std::fabs(dot(cross(ab, ac), ad)) < 1e-12f) {
    V3<float> f[4][3] = {{a, b, c}, {a, b, d}, {a, c, d}, {b, c, d}};
    V3<float> best;
    CoSimplex bk;
    for (int i = 0; i < 4; ++i) {  // min_by keeps the first of equal minima
      CoSimplex kk;
      V3<float> q = co_triangle(f[i][0], f[i][1], f[i][2], kk);
      if (i == 0 || len2(q) < len2(best)) { best = q; bk = kk; }
    }
    k = bk;
    return best;
  }
  V3<float> abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
  V3<float> bo = -b;
  V3<float> bcd = cross(c - b, d - b);
  if (dot(abc, ao) * dot(abc, ad) < 0.0f) return co_triangle(a, b, c, k);
  if (dot(acd, ao) * dot(acd, ab) < 0.0f) return co_triangle(a, c, d, k);
  if (dot(adb, ao) * dot(adb, ac) < 0.0f) return co_triangle(a, d, b, k);
  if (dot(bcd, bo) * dot(bcd, a - b) < 0.0f) return co_triangle(b, c, d, k);
  k.p[0] = a; k.p[1] = b; k.p[2] = c; k.p[3] = d; k.n = 4;
  return splat(0.0f);
}

inline V3<float> closest_origin(CoSimplex& s) {
  CoSimplex k;
  V3<float> v;
  switch (s.n) {
    case 0: return splat(0.0f);
    case 1: return s.p[0];
    case 2: v = co_segment(s.p[0], s.p[1], k); break;
    case 3: v = co_triangle(s.p[0], s.p[1], s.p[2], k); break;
    default: v = co_tetra(s.p[0], s.p[1], s.p[2], s.p[3], k); break;
  }
  s = k;
  return v;
}

// point_convex_distance (convex_sweep.rs:59-103): squared distance and the
// unit direction from the shape toward `point`.
template <class F>
inline void point_convex_distance(const F& support, V3<float> point, float& d2, V3<float>& dir) {
  auto mink = [&](V3<float> v) { return point - support(-v); };
  V3<float> v = point;
  if (len2(v) < 1e-12f) v = mk(1.0f, 0.0f, 0.0f);
  CoSimplex s;
  s.p[0] = mink(v);
  s.n = 1;
  v = s.p[0];
  for (int it = 0; it < 64; ++it) {
    if (len2(v) < 1e-10f) { d2 = 0.0f; dir = splat(0.0f); return; }
    V3<float> w = mink(-v);
    for (int i = 0; i < s.n; ++i)
      if (len2(s.p[i] - w) < 1e-12f) { d2 = len2(v); dir = v / len(v); return; }
    s.p[s.n++] = w;
    v = closest_origin(s);
  }
  if (len2(v) < 1e-10f) { d2 = 0.0f; dir = splat(0.0f); return; }
  d2 = len2(v);
  dir = v / len(v);
}

struct SweepHit {
  float fraction;
  V3<float> normal, point;
};

// sweep_with_dist (convex_sweep.rs:150-206)
template <class D>
inline bool sweep_with_dist(V3<float> from, V3<float> dir, float length, float radius, const D& dist, SweepHit& h) {
  auto dist2 = [&](V3<float> p) {
    float d2;
    V3<float> v;
    dist(p, d2, v);
    return d2;
  };
  float lo = 0.0f, hi = 1.0f;
  for (int i = 0; i < 40; ++i) {
    float m1 = lo + (hi - lo) / 3.0f;
    float m2 = hi - (hi - lo) / 3.0f;
    if (dist2(from + dir * (m1 * length)) < dist2(from + dir * (m2 * length))) hi = m2;
    else lo = m1;
  }
  float s_min = (lo + hi) * 0.5f;
  if (dist2(from + dir * (s_min * length)) > 1e-8f) return false;
  float low = 0.0f, high = s_min;
  for (int i = 0; i < 60; ++i) {
    float mid = (low + high) * 0.5f;
    if (dist2(from + dir * (mid * length)) <= 1e-8f) high = mid;
    else low = mid;
  }
  float fraction = (low + high) * 0.5f;
  V3<float> probe = from + dir * fmax_((fraction - 2e-4f) * length, 0.0f);
  float d2;
  V3<float> v;
  dist(probe, d2, v);
  V3<float> normal = len2(v) > 1e-12f ? v / len(v) : -dir;
  V3<float> entry = from + dir * (fraction * length);
  h = SweepHit{fraction, normal, entry - normal * radius};
  return true;
}

// point_triangle_distance (convex_sweep.rs:208-256)
inline float point_triangle_distance(V3<float> p, V3<float> a, V3<float> b, V3<float> c, V3<float>& dv) {
  V3<float> ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) { dv = p - a; return len2(p - a); }
  V3<float> bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) { dv = p - b; return len2(p - b); }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / (d1 - d3);
    V3<float> q = a + ab * v;
    dv = p - q;
    return len2(p - q);
  }
  V3<float> cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) { dv = p - c; return len2(p - c); }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    V3<float> q = a + ac * w;
    dv = p - q;
    return len2(p - q);
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    V3<float> q = b + (c - b) * w;
    dv = p - q;
    return len2(p - q);
  }
  float denom = 1.0f / ((va + vb) + vc);
  float v = vb * denom, w = vc * denom;
  V3<float> q = (a + ab * v) + ac * w;
  dv = p - q;
  return len2(p - q);
}

// swept_sphere_cast (convex_sweep.rs:35-57) against a convex or plane target.
inline bool sweep_sphere(V3<float> from, V3<float> to, float radius, const Shape<float>& other, V3<float> op,
                         Quat<float> oq, SweepHit& h) {
  V3<float> delta = to - from;
  float length = len(delta);
  if (length <= std::numeric_limits<float>::epsilon()) return false;
  V3<float> dir = delta / length;
  if (other.kind == PLANE) {  // sweep_plane (convex_sweep.rs:258-292)
    V3<float> n = qrot(oq, other.pnormal);
    float dn0 = dot(n, from) - other.pconst;
    if (std::fabs(dn0) <= radius) {
      h = SweepHit{0.0f, dn0 >= 0.0f ? n : -n, from - n * dn0};
      return true;
    }
    float dd = dot(n, dir * length);
    if (dd >= 0.0f) return false;
    float sign = dn0 > 0.0f ? 1.0f : -1.0f;
    float fraction = (sign * radius - dn0) / dd;
    if (!(fraction >= 0.0f && fraction <= 1.0f)) return false;
    V3<float> normal = sign > 0.0f ? n : -n;
    V3<float> entry = from + (dir * length) * fraction;
    h = SweepHit{fraction, normal, entry - normal * radius};
    return true;
  }
  // sweep_convex (convex_sweep.rs:105-125)
  auto support = [&](V3<float> d) {
    V3<float> sph = len2(d) > 1e-12f ? (d / len(d)) * radius : splat(0.0f);
    return other.support_world(op, oq, d) + sph;
  };
  auto dist = [&](V3<float> p, float& d2, V3<float>& v) { point_convex_distance(support, p, d2, v); };
  return sweep_with_dist(from, dir, length, radius, dist, h);
}

// sweep_mesh (convex_sweep.rs:294-344): per-triangle casts in mesh space, the
// smallest fraction wins (the first on ties).
inline bool sweep_sphere_mesh(V3<float> from, V3<float> to, float radius, const TriMesh& mesh, V3<float> pos,
                              Quat<float> ori, SweepHit& out) {
  Quat<float> inv = qconj(ori);
  V3<float> lf = qrot(inv, from - pos), lt = qrot(inv, to - pos);
  V3<float> delta = lt - lf;
  float length = len(delta);
  if (length <= std::numeric_limits<float>::epsilon()) return false;
  V3<float> dir = delta / length;
  V3<float> r = splat(radius);
  Aabb<float> box{vmin(lf, lt) - r, vmax(lf, lt) + r};
  std::vector<uint32_t> cand;
  mesh.query(box, cand);
  bool any = false;
  SweepHit best{};
  for (uint32_t t : cand) {
    V3<float> a = mesh.vert(t, 0), b = mesh.vert(t, 1), c = mesh.vert(t, 2);
    auto dist = [&](V3<float> p, float& d2, V3<float>& v) {
      V3<float> dv;
      float q2 = point_triangle_distance(p, a, b, c, dv);
      if (q2 <= radius * radius) {
        d2 = 0.0f;
        v = splat(0.0f);
      } else {
        float d = std::sqrt(q2);
        d2 = (d - radius) * (d - radius);
        v = dv / d;
      }
    };
    SweepHit hh;
    if (sweep_with_dist(lf, dir, length, radius, dist, hh) && (!any || hh.fraction < best.fraction)) {
      best = hh;
      any = true;
    }
  }
  if (!any) return false;
  out = SweepHit{best.fraction, qrot(ori, best.normal), qrot(ori, best.point) + pos};
  return true;
}

// convex_cast (convex_sweep.rs:357-385) of `shape` (orientation `ori`) against
// a convex or plane target.
inline bool convex_cast(const Shape<float>& shape, V3<float> from, Quat<float> ori, V3<float> to,
                        const Shape<float>& other, V3<float> op, Quat<float> oq, SweepHit& h) {
  V3<float> delta = to - from;
  float length = len(delta);
  if (length <= std::numeric_limits<float>::epsilon()) return false;
  if (other.kind == PLANE) {  // sweep_convex_plane (convex_sweep.rs:431-478)
    V3<float> n = qrot(oq, other.pnormal);
    float c = other.pconst + dot(n, op);
    V3<float> normal = dot(n, from) >= c ? n : -n;
    auto deepest = [&](V3<float> p) { return shape.support_world(p, ori, -normal); };
    float da = dot(normal, deepest(from)) - c;
    float db = dot(normal, deepest(to)) - c;
    if (db >= da) return false;
    if (da <= 0.0f) {
      h = SweepHit{0.0f, normal, deepest(from)};
      return true;
    }
    float fraction = da / (da - db);
    if (!(fraction >= 0.0f && fraction <= 1.0f)) return false;
    V3<float> entry = from + delta * fraction;
    h = SweepHit{fraction, normal, deepest(entry)};
    return true;
  }
  // sweep_convex_shape (convex_sweep.rs:387-429)
  V3<float> dir = delta / length;
  auto dist = [&](V3<float> p, float& d2, V3<float>& v) {
    auto support = [&](V3<float> d) { return other.support_world(op, oq, d) - shape.support_world(p, ori, -d); };
    point_convex_distance(support, splat(0.0f), d2, v);
  };
  if (!sweep_with_dist(from, dir, length, 0.0f, dist, h)) return false;
  h.point = other.support_world(op, oq, h.normal);
  return true;
}

// sweep_convex_mesh (convex_sweep.rs:480-524): BVH-pruned per-triangle convex
// casts in mesh space (no triangle-AABB filter), the smallest fraction wins.
inline bool convex_cast_mesh(const Shape<float>& shape, V3<float> from, Quat<float> ori, V3<float> to,
                             const TriMesh& mesh, V3<float> op, Quat<float> oq, SweepHit& out) {
  Quat<float> inv = qconj(oq);
  V3<float> fl = qrot(inv, from - op), tl = qrot(inv, to - op);
  Quat<float> ol = qmul(inv, ori);
  Aabb<float> a = shape.world_aabb(fl, ol), b = shape.world_aabb(tl, ol);
  Aabb<float> box{vmin(a.min, b.min), vmax(a.max, b.max)};
  std::vector<uint32_t> cand;
  mesh.query(box, cand);
  bool any = false;
  SweepHit best{};
  for (uint32_t t : cand) {
    Shape<float> tri = Shape<float>::hull(&mesh.tri_pts[3 * (size_t)t], 3);
    SweepHit hh;
    if (convex_cast(shape, fl, ol, tl, tri, splat(0.0f), Quat<float>{0.0f, 0.0f, 0.0f, 1.0f}, hh) &&
        (!any || hh.fraction < best.fraction)) {
      best = hh;
      any = true;
    }
  }
  if (!any) return false;
  out = SweepHit{best.fraction, qrot(oq, best.normal), qrot(oq, best.point) + op};
  return true;
}

}  // namespace oracle
