// bch_narrow.hpp -- per-pair convex narrowphase for the collision kernels.
//
// One lane evaluates one pair.  The closed forms (sphere-sphere, sphere-box,
// plane-convex, box-box SAT with face clipping) and the GJK margin-core walk
// run entirely in registers; the walk hands deep overlaps to EPA, which keeps
// its polytope in a caller-provided store (LDS or global scratch).
//
// Behaviour follows the reference narrowphase (file:line per routine) with
// the same statement order, branch structure and tie rules, so results match
// the reference's f32 path to the last bit wherever the shared operation order
// is used (see bch_math.hpp) and the f32 division/square root are correctly
// rounded.
//
// Conventions (narrowphase/mod.rs:21-23): the normal points from A to B,
// `pos` is the witness point on B, depth > 0 means penetration, < 0 a
// separation inside the contact band.
#pragma once
#include "bch_shapes.hpp"

namespace bch {

struct Contact {
  V3 pos, n;
  float depth;
};

struct Manifold4 {
  int n;
  Contact c[4];
  // Constant-index stores keep the manifold in registers on the device.
  BCH_HD void push(V3 p, V3 nrm, float d) {
    Contact ct{p, nrm, d};
    if (n == 0) c[0] = ct;
    else if (n == 1) c[1] = ct;
    else if (n == 2) c[2] = ct;
    else if (n == 3) c[3] = ct;
    ++n;
  }
  BCH_HD void flip() {
    for (int i = 0; i < 4; ++i)
      if (i < n) c[i].n = -c[i].n;
  }
};

// --------------------------------------------------------- closed forms --
// narrowphase/sphere.rs:10-32
BCH_HD bool sphere_sphere(float ra, V3 pa, float rb, V3 pb, Manifold4& m) {
  V3 delta = pb - pa;
  float dist_sq = len2(delta);
  float radius_sum = ra + rb;
  if (dist_sq >= radius_sum * radius_sum) return false;
  float dist = fsqrt(dist_sq);
  V3 normal = sel(dist > 1e-12f, delta / dist, V3{0.0f, 1.0f, 0.0f});
  float pen = radius_sum - dist;
  m.push(pa + normal * (ra - pen * 0.5f), normal, pen);
  return true;
}

// narrowphase/sphere.rs:38-82 (normal from the box to the sphere)
BCH_HD bool sphere_box(float r, V3 ps, V3 h, V3 pb, Q4 ob, float threshold, Manifold4& m) {
  V3 local = qrot(qconj(ob), ps - pb);
  V3 clamped = vmin(vmax(local, -h), h);
  V3 delta = local - clamped;
  float dist_sq = len2(delta);
  float contact_dist = r + threshold;
  if (dist_sq >= contact_dist * contact_dist) return false;
  float dist = fsqrt(dist_sq);
  V3 nl;
  float pen;
  if (dist > 1e-12f) {
    nl = delta / dist;
    pen = r - dist;
  } else {
    float dx = h.x - fabsf(local.x);
    float dy = h.y - fabsf(local.y);
    float dz = h.z - fabsf(local.z);
    float d;
    if (dx <= dy && dx <= dz) {
      nl = V3{signum(local.x), 0.0f, 0.0f};
      d = dx;
    } else if (dy <= dz) {
      nl = V3{0.0f, signum(local.y), 0.0f};
      d = dy;
    } else {
      nl = V3{0.0f, 0.0f, signum(local.z)};
      d = dz;
    }
    pen = r + d;
  }
  m.push(qrot(ob, clamped) + pb, qrot(ob, nl), pen);
  return true;
}

// narrowphase/plane.rs:19-56 (plane params: local normal, constant)
BCH_HD bool plane_convex(V3 pn, float pc, V3 pp, Q4 pq, const ShapeRef& cv, V3 cp, Q4 cq, float threshold,
                         Manifold4& m) {
  V3 n = qrot(pq, pn);
  float c = pc + dot(n, pp);
  V3 normal = sel(dot(n, cp - pp) >= 0.0f, n, -n);
  if (!(normal.x == n.x && normal.y == n.y && normal.z == n.z)) c = -c;
  V3 deepest = support_world(cv, cp, cq, -normal);
  float depth = c - dot(normal, deepest);
  if (depth < -threshold) return false;
  m.push(deepest, normal, depth);
  return true;
}

// ---------------------------------------------------------------- box-box --
// Every array of the clip and cull below is indexed by compile-time
// constants only (unrolled loops; a store at a run-time position is a select
// over the eight slots), so the arrays stay in registers instead of scratch.
struct ClipPoly {
  float x[8], y[8];
  int n;
  bool overflow;
};

BCH_HD void clip_push(ClipPoly& o, float a, float b) {
  // past eight vertices: flagged, never written past
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    bool at = k == o.n;
    o.x[k] = at ? a : o.x[k];
    o.y[k] = at ? b : o.y[k];
  }
  o.overflow = o.overflow || o.n >= 8;
  o.n = o.n < 8 ? o.n + 1 : o.n;
}

// box_box.rs:92-114: keep `sign * coord[axis] < h`.
BCH_HD void clip_poly(const ClipPoly& s, int axis, float sign, float h, ClipPoly& o) {
  o.n = 0;
  o.overflow = s.overflow;
#pragma unroll
  for (int k = 0; k < 8; ++k) o.x[k] = o.y[k] = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (i >= s.n) break;
    // the next vertex wraps to 0 after the last (i = 7 only when n = 8)
    bool wrap = i + 1 == s.n;
    float cx = s.x[i], cy = s.y[i];
    float nx = wrap ? s.x[0] : s.x[i < 7 ? i + 1 : 0], ny = wrap ? s.y[0] : s.y[i < 7 ? i + 1 : 0];
    float ca = axis == 0 ? cx : cy, na = axis == 0 ? nx : ny;
    float cb = axis == 0 ? cy : cx, nb = axis == 0 ? ny : nx;
    bool cur_in = sign * ca < h;
    bool next_in = sign * na < h;
    if (cur_in) clip_push(o, cx, cy);
    if (cur_in != next_in) {
      float t = (sign * h - ca) / (na - ca);
      float pa = sign * h;
      float pb = cb + (nb - cb) * t;
      if (axis == 0) clip_push(o, pa, pb);
      else clip_push(o, pb, pa);
    }
  }
}

// box_box.rs:178-528 (ODE dBoxBox2: 15 SAT axes, incident-face clipping,
// deepest + farthest culling to <= 4 points).
BCH_HD bool box_box(V3 ha, V3 p1, Q4 oa, V3 hb, V3 p2, Q4 ob, Manifold4& out) {
  const float EPS = 1e-6f, FUDGE = 1.05f;
  M3 r1 = from_quat(oa), r2 = from_quat(ob);
  V3 p = p2 - p1;
  M3 r1t = transpose(r1);
  V3 pp = mmul(r1t, p);
  M3 r = mmulm(r1t, r2);
  float r11 = r.c0.x, r12 = r.c1.x, r13 = r.c2.x;
  float r21 = r.c0.y, r22 = r.c1.y, r23 = r.c2.y;
  float r31 = r.c0.z, r32 = r.c1.z, r33 = r.c2.z;
  float q11 = fabsf(r11), q12 = fabsf(r12), q13 = fabsf(r13);
  float q21 = fabsf(r21), q22 = fabsf(r22), q23 = fabsf(r23);
  float q31 = fabsf(r31), q32 = fabsf(r32), q33 = fabsf(r33);

  float s = -INFINITY;
  bool invert = false;
  int code = 0;
  bool from_r2 = false;
  int axis = 0;
  bool is_cross = false;
  V3 nc = splat(0.0f);

#define BCH_FACE(E1, E2, FR2, AX, CC)   \
  {                                     \
    float e1 = (E1);                    \
    float s2 = fabsf(e1) - (E2);        \
    if (s2 > 0.0f) return false;        \
    if (s2 > s) {                       \
      s = s2;                           \
      is_cross = false;                 \
      from_r2 = FR2;                    \
      axis = AX;                        \
      invert = e1 < 0.0f;               \
      code = CC;                        \
    }                                   \
  }
#define BCH_CROSS(E1, E2, N1, N2, N3, CC)          \
  {                                                \
    float e1 = (E1);                               \
    float s2 = fabsf(e1) - (E2);                   \
    if (s2 > EPS) return false;                    \
    float n1 = (N1), n2 = (N2), n3 = (N3);         \
    float l = fsqrt(n1 * n1 + n2 * n2 + n3 * n3);  \
    if (l > EPS) {                                 \
      s2 = s2 / l;                                 \
      if (s2 * FUDGE > s) {                        \
        s = s2;                                    \
        is_cross = true;                           \
        nc = V3{n1 / l, n2 / l, n3 / l};           \
        invert = e1 < 0.0f;                        \
        code = CC;                                 \
      }                                            \
    }                                              \
  }
  BCH_FACE(pp.x, ha.x + hb.x * q11 + hb.y * q12 + hb.z * q13, false, 0, 1);
  BCH_FACE(pp.y, ha.y + hb.x * q21 + hb.y * q22 + hb.z * q23, false, 1, 2);
  BCH_FACE(pp.z, ha.z + hb.x * q31 + hb.y * q32 + hb.z * q33, false, 2, 3);
  BCH_FACE(dot(r2.c0, p), ha.x * q11 + ha.y * q21 + ha.z * q31 + hb.x, true, 0, 4);
  BCH_FACE(dot(r2.c1, p), ha.x * q12 + ha.y * q22 + ha.z * q32 + hb.y, true, 1, 5);
  BCH_FACE(dot(r2.c2, p), ha.x * q13 + ha.y * q23 + ha.z * q33 + hb.z, true, 2, 6);
  const float f2 = 1e-5f;
  q11 += f2; q12 += f2; q13 += f2;
  q21 += f2; q22 += f2; q23 += f2;
  q31 += f2; q32 += f2; q33 += f2;
  BCH_CROSS(pp.z * r21 - pp.y * r31, ha.y * q31 + ha.z * q21 + hb.y * q13 + hb.z * q12, 0.0f, -r31, r21, 7);
  BCH_CROSS(pp.z * r22 - pp.y * r32, ha.y * q32 + ha.z * q22 + hb.x * q13 + hb.z * q11, 0.0f, -r32, r22, 8);
  BCH_CROSS(pp.z * r23 - pp.y * r33, ha.y * q33 + ha.z * q23 + hb.x * q12 + hb.y * q11, 0.0f, -r33, r23, 9);
  BCH_CROSS(pp.x * r31 - pp.z * r11, ha.x * q31 + ha.z * q11 + hb.y * q23 + hb.z * q22, r31, 0.0f, -r11, 10);
  BCH_CROSS(pp.x * r32 - pp.z * r12, ha.x * q32 + ha.z * q12 + hb.x * q23 + hb.z * q21, r32, 0.0f, -r12, 11);
  BCH_CROSS(pp.x * r33 - pp.z * r13, ha.x * q33 + ha.z * q13 + hb.x * q22 + hb.y * q21, r33, 0.0f, -r13, 12);
  BCH_CROSS(pp.y * r11 - pp.x * r21, ha.x * q21 + ha.y * q11 + hb.y * q33 + hb.z * q32, -r21, r11, 0.0f, 13);
  BCH_CROSS(pp.y * r12 - pp.x * r22, ha.x * q22 + ha.y * q12 + hb.x * q33 + hb.z * q31, -r22, r12, 0.0f, 14);
  BCH_CROSS(pp.y * r13 - pp.x * r23, ha.x * q23 + ha.y * q13 + hb.x * q32 + hb.y * q31, -r23, r13, 0.0f, 15);
#undef BCH_FACE
#undef BCH_CROSS

  V3 normal = sel(is_cross, mmul(r1, nc), sel(from_r2, col(r2, axis), col(r1, axis)));
  if (invert) normal = -normal;
  float depth = -s;

  if (code > 6) {  // edge-edge (box_box.rs:402-423)
    V3 pa = p1;
    for (int j = 0; j < 3; ++j) {
      V3 c = col(r1, j);
      float sg = dot(normal, c) > 0.0f ? 1.0f : -1.0f;
      pa = pa + (sg * comp(ha, j)) * c;
    }
    V3 pb = p2;
    for (int j = 0; j < 3; ++j) {
      V3 c = col(r2, j);
      float sg = dot(normal, c) > 0.0f ? -1.0f : 1.0f;
      pb = pb + (sg * comp(hb, j)) * c;
    }
    V3 ua = col(r1, (code - 7) / 3);
    V3 ub = col(r2, (code - 7) % 3);
    // line_closest_approach (box_box.rs:31-46); only the B-side point is used.
    V3 pd = pb - pa;
    float uaub = dot(ua, ub);
    float q1 = dot(ua, pd);
    float q2 = -dot(ub, pd);
    float d = 1.0f - uaub * uaub;
    float beta = 0.0f;
    if (!(d <= 1e-4f)) {
      d = 1.0f / d;
      beta = (uaub * q1 + q2) * d;
    }
    pb = pb + ub * beta;
    out.push(pb, normal, depth);
    return true;
  }

  bool a_ref = code <= 3;
  M3 ra = sel(a_ref, r1, r2);
  M3 rb = sel(a_ref, r2, r1);
  V3 pa_ref = sel(a_ref, p1, p2);
  V3 pb_inc = sel(a_ref, p2, p1);
  V3 sa = sel(a_ref, ha, hb);
  V3 sb = sel(a_ref, hb, ha);
  V3 normal2 = sel(a_ref, normal, -normal);
  V3 nr = mmul(transpose(rb), normal2);
  V3 anr = vabs(nr);
  int lanr, a1, a2;
  if (anr.y > anr.x) {
    if (anr.y > anr.z) { lanr = 1; a1 = 0; a2 = 2; } else { lanr = 2; a1 = 0; a2 = 1; }
  } else if (anr.x > anr.z) {
    lanr = 0; a1 = 1; a2 = 2;
  } else {
    lanr = 2; a1 = 0; a2 = 1;
  }
  V3 center = sel(comp(nr, lanr) < 0.0f, (pb_inc - pa_ref) + comp(sb, lanr) * col(rb, lanr),
                  (pb_inc - pa_ref) - comp(sb, lanr) * col(rb, lanr));
  int code_n = a_ref ? code - 1 : code - 4;
  int code1 = code_n == 0 ? 1 : 0;
  int code2 = code_n == 2 ? 1 : 2;
  V3 ra1 = col(ra, code1), ra2 = col(ra, code2), rba1 = col(rb, a1), rba2 = col(rb, a2);
  float c1 = dot(center, ra1), c2 = dot(center, ra2);
  float m11 = dot(ra1, rba1), m12 = dot(ra1, rba2), m21 = dot(ra2, rba1), m22 = dot(ra2, rba2);
  float k1 = m11 * comp(sb, a1), k2 = m21 * comp(sb, a1), k3 = m12 * comp(sb, a2), k4 = m22 * comp(sb, a2);
  ClipPoly buf, tmp;
  buf.n = 4;
  buf.overflow = false;
  buf.x[0] = c1 - k1 - k3; buf.y[0] = c2 - k2 - k4;
  buf.x[1] = c1 - k1 + k3; buf.y[1] = c2 - k2 + k4;
  buf.x[2] = c1 + k1 + k3; buf.y[2] = c2 + k2 + k4;
  buf.x[3] = c1 + k1 - k3; buf.y[3] = c2 + k2 - k4;
  float h0 = comp(sa, code1), h1 = comp(sa, code2);
  clip_poly(buf, 0, -1.0f, h0, tmp);
  clip_poly(tmp, 0, 1.0f, h0, buf);
  clip_poly(buf, 1, -1.0f, h1, tmp);
  clip_poly(tmp, 1, 1.0f, h1, buf);
  if (buf.n == 0) return false;
  float det1 = 1.0f / (m11 * m22 - m12 * m21);
  float n11 = m11 * det1, n12 = m12 * det1, n21 = m21 * det1, n22 = m22 * det1;
  V3 pts[8];
  float deps[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    pts[k] = V3{0.0f, 0.0f, 0.0f};
    deps[k] = 0.0f;
  }
  int np = 0;
  float sdep = comp(sa, code_n);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (i >= buf.n) break;
    float rx = buf.x[i], ry = buf.y[i];
    float kk1 = n22 * (rx - c1) - n12 * (ry - c2);
    float kk2 = -n21 * (rx - c1) + n11 * (ry - c2);
    V3 point = (center + kk1 * rba1) + kk2 * rba2;
    float dep = sdep - dot(normal2, point);
    if (dep >= 0.0f) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        bool at = k == np;
        pts[k] = V3{at ? point.x : pts[k].x, at ? point.y : pts[k].y, at ? point.z : pts[k].z};
        deps[k] = at ? dep : deps[k];
      }
      ++np;
    }
  }
  if (np == 0) return false;
  // cull_points (box_box.rs:131-174): deepest (last maximum), then the rest by
  // descending squared distance to it (stable insertion sort), keep 4.
  int sel[4] = {0, 1, 2, 3};
  int nsel = np;
  if (np > 4) {
    int i0 = 0;
    float dd = deps[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) {
      if (i >= np) break;
      if (deps[i] >= dd) {
        i0 = i;
        dd = deps[i];
      }
    }
    V3 p0 = pts[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      bool at = k == i0;
      p0 = V3{at ? pts[k].x : p0.x, at ? pts[k].y : p0.y, at ? pts[k].z : p0.z};
    }
    int ri[8];
    float rd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      ri[k] = 0;
      rd[k] = 0.0f;
    }
    int rn = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if (i >= np) break;
      if (i != i0) {
        float d2 = len2(pts[i] - p0);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          bool at = k == rn;
          ri[k] = at ? i : ri[k];
          rd[k] = at ? d2 : rd[k];
        }
        ++rn;
      }
    }
    // The insertion sort step by step: item j moves down while the item
    // before it compares `<` (the same comparisons in the same order, so
    // equal and NaN keys land where the loop of the reference puts them).
#pragma unroll
    for (int j = 1; j < 8; ++j) {
      if (j >= rn) break;
      int ki = ri[j];
      float kd = rd[j];
      bool moving = true;
#pragma unroll
      for (int k = j; k > 0; --k) {
        bool shift = moving && rd[k - 1] < kd;
        bool place = moving && !shift;
        ri[k] = shift ? ri[k - 1] : (place ? ki : ri[k]);
        rd[k] = shift ? rd[k - 1] : (place ? kd : rd[k]);
        moving = shift;
      }
      ri[0] = moving ? ki : ri[0];
      rd[0] = moving ? kd : rd[0];
    }
    sel[0] = i0;
    sel[1] = ri[0];
    sel[2] = ri[1];
    sel[3] = ri[2];
    nsel = 4;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= nsel) break;
    V3 pt = pts[0];
    float dp = deps[0];
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      bool at = k == sel[i];
      pt = V3{at ? pts[k].x : pt.x, at ? pts[k].y : pt.y, at ? pts[k].z : pt.z};
      dp = at ? deps[k] : dp;
    }
    V3 pos = pt + pa_ref;
    if (!a_ref) pos = pos + normal * dp;
    out.push(pos, normal, dp);
  }
  if (buf.overflow) out.n = -1;  // more than 8 clip vertices: report, never write past
  return true;
}

// -------------------------------------------------------------------- GJK --
struct Pair {
  ShapeRef a, b;
  V3 pa, pb;
  Q4 qa, qb;
};

struct SV {
  V3 w, p_b;
  // local support points the vertex came from (index into the shape's core
  // support set); only the signed-distance walk fills and reads them
  int ia = 0, ib = 0;
};

BCH_HD SV core_support(const Pair& pr, V3 d) {
  V3 a = core_support_world(pr.a, pr.pa, pr.qa, d);
  V3 b = core_support_world(pr.b, pr.pb, pr.qb, -d);
  return SV{a - b, b};
}

struct Reduce {
  bool inside;
  V3 v, p_b;
};

// gjk.rs:398-464: closest point of triangle (a, b, c) to the origin with
// barycentric weights.
struct Tri {
  V3 q;
  float w0, w1, w2;
};

BCH_HD Tri closest_tri(V3 a, V3 b, V3 c) {
  V3 ab = b - a, ac = c - a, ap = -a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return Tri{a, 1, 0, 0};
  V3 bp = -b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return Tri{b, 0, 1, 0};
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    if (d1 == d3) return Tri{a, 1, 0, 0};
    float t = d1 / (d1 - d3);
    return Tri{a + ab * t, 1.0f - t, t, 0};
  }
  V3 cp = -c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return Tri{c, 0, 0, 1};
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    if (d2 == d6) return Tri{a, 1, 0, 0};
    float t = d2 / (d2 - d6);
    return Tri{a + ac * t, 1.0f - t, 0, t};
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    if ((d4 - d3) + (d5 - d6) == 0.0f) return Tri{b, 0, 1, 0};
    float t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return Tri{b + (c - b) * t, 0, 1.0f - t, t};
  }
  float sum = va + vb + vc;
  if (sum <= 0.0f || is_nan(sum)) {
    if (len2(a) <= len2(b) && len2(a) <= len2(c)) return Tri{a, 1, 0, 0};
    if (len2(b) <= len2(c)) return Tri{b, 0, 1, 0};
    return Tri{c, 0, 0, 1};
  }
  float denom = 1.0f / sum;
  float v = vb * denom, w = vc * denom;
  return Tri{(a + ab * v) + ac * w, 1.0f - v - w, v, w};
}

// The walk's simplex: s[0..n), newest appended last by the walk; the
// reductions rewrite it exactly like gjk.rs:366-583.
struct Simp {
  SV s0, s1, s2, s3;
  int n;
};


BCH_HD SV sel(bool c, const SV& a, const SV& b) {
  return SV{sel(c, a.w, b.w), sel(c, a.p_b, b.p_b), c ? a.ia : b.ia, c ? a.ib : b.ib};
}

// Append at position n (1..3): selects on every slot, so the simplex stays in
// registers.
BCH_HD void simp_append(Simp& S, SV v) {
  S.s1 = sel(S.n == 1, v, S.s1);
  S.s2 = sel(S.n == 2, v, S.s2);
  S.s3 = sel(S.n == 3, v, S.s3);
  S.n += 1;
}

// gjk.rs:466-486 on (a, b, c) = (buf[2], buf[1], buf[0]): keep the vertices
// with positive weight, in order a, b, c (a alone when none).
BCH_HD Reduce reduce_tri_abc(Simp& S, SV a, SV b, SV c) {
  Tri T = closest_tri(a.w, b.w, c.w);
  bool p0 = T.w0 > 0.0f, p1 = T.w1 > 0.0f, p2 = T.w2 > 0.0f;
  int k = (int)p0 + (int)p1 + (int)p2;
  SV k0 = sel(p0, a, sel(p1, b, sel(p2, c, a)));
  SV k1 = sel(p0 && p1, b, c);
  S.s0 = k0;
  S.s1 = sel(k >= 2, k1, S.s1);
  S.s2 = sel(k >= 3, c, S.s2);
  S.n = k > 0 ? k : 1;
  V3 p_b = (a.p_b * T.w0 + b.p_b * T.w1) + c.p_b * T.w2;
  return Reduce{false, T.q, p_b};
}

// gjk.rs:352-567
BCH_HD Reduce reduce(Simp& S) {
  if (S.n == 1) return Reduce{false, S.s0.w, S.s0.p_b};
  if (S.n == 2) {  // gjk.rs:366-393
    SV a = S.s1, b = S.s0;
    V3 ab = b.w - a.w;
    float denom = len2(ab);
    if (denom < 1e-24f) { S.n = 1; S.s0 = a; return Reduce{false, a.w, a.p_b}; }
    float t = -dot(a.w, ab) / denom;
    if (t <= 0.0f) { S.n = 1; S.s0 = a; return Reduce{false, a.w, a.p_b}; }
    if (t >= 1.0f) { S.n = 1; S.s0 = b; return Reduce{false, b.w, b.p_b}; }
    S.n = 2;
    S.s0 = a;
    S.s1 = b;
    return Reduce{false, a.w + ab * t, a.p_b + (b.p_b - a.p_b) * t};
  }
  if (S.n == 3) return reduce_tri_abc(S, S.s2, S.s1, S.s0);
  // Tetrahedron (gjk.rs:488-567).
  SV a = S.s3, b = S.s2, c = S.s1, d = S.s0;
  V3 ab = b.w - a.w, ac = c.w - a.w, ad = d.w - a.w;
  if (fabsf(dot(cross(ab, ac), ad)) < 1e-12f) {
    // Closest of faces (a,b,d), (a,c,d), (b,c,d); face (a,b,c) is only kept
    // when every other face scores NaN (its score starts at +inf).
    SV fa = a, fb = b, fc = c;
    Tri T0 = closest_tri(a.w, b.w, c.w);
    V3 bq = T0.q;
    float b0 = T0.w0, b1 = T0.w1, b2 = T0.w2;
    float best_v2 = INFINITY;
    for (int f = 1; f < 4; ++f) {
      SV x = sel(f == 3, b, a);
      SV y = sel(f == 1, b, c);
      SV z = d;
      Tri T = closest_tri(x.w, y.w, z.w);
      float v2 = len2(T.q);
      bool better = v2 < best_v2;
      best_v2 = better ? v2 : best_v2;
      bq = sel(better, T.q, bq);
      b0 = better ? T.w0 : b0;
      b1 = better ? T.w1 : b1;
      b2 = better ? T.w2 : b2;
      fa = sel(better, x, fa);
      fb = sel(better, y, fb);
      fc = sel(better, z, fc);
    }
    bool p0 = b0 > 0.0f, p1 = b1 > 0.0f, p2 = b2 > 0.0f;
    int k = (int)p0 + (int)p1 + (int)p2;
    SV k0 = sel(p0, fa, sel(p1, fb, sel(p2, fc, fa)));
    SV k1 = sel(p0 && p1, fb, fc);
    S.s0 = k0;
    S.s1 = sel(k >= 2, k1, S.s1);
    S.s2 = sel(k >= 3, fc, S.s2);
    S.n = k > 0 ? k : 1;
    V3 p_b = (fa.p_b * b0 + fb.p_b * b1) + fc.p_b * b2;
    return Reduce{false, bq, p_b};
  }
  V3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
  V3 ao = -a.w, bo = -b.w;
  V3 bcd = cross(c.w - b.w, d.w - b.w);
  if (dot(abc, ao) * dot(abc, ad) < 0.0f) return reduce_tri_abc(S, a, b, c);
  if (dot(acd, ao) * dot(acd, ab) < 0.0f) return reduce_tri_abc(S, a, c, d);
  if (dot(adb, ao) * dot(adb, ac) < 0.0f) return reduce_tri_abc(S, a, d, b);
  if (dot(bcd, bo) * dot(bcd, a.w - b.w) < 0.0f) return reduce_tri_abc(S, b, c, d);
  return Reduce{true, splat(0.0f), splat(0.0f)};
}

// gjk.rs:668-680
BCH_HD V3 perpendicular(V3 v) {
  bool small = fabsf(v.x) < 0.9f;
  V3 helper = sel(small, V3{1.0f, 0.0f, 0.0f}, V3{0.0f, 1.0f, 0.0f});
  V3 perp = cross(v, helper);
  if (len2(perp) > 0.0f) return normalize(perp);
  V3 alt = sel(small, V3{0.0f, 1.0f, 0.0f}, V3{1.0f, 0.0f, 0.0f});
  perp = cross(v, alt);
  if (len2(perp) > 0.0f) return normalize(perp);
  return V3{0.0f, 0.0f, 1.0f};
}

// Strict enclosing walk (gjk.rs:299-342 with solve_* at 609-858) over the
// core Minkowski support.  P[0] is the newest point.
struct WSimp {
  V3 p0, p1, p2, p3;
  int n;
};

BCH_HD bool shrink_or_normal(WSimp& S, V3& dir, V3 closest, V3 normal, V3 n0, V3 n1) {
  float span2 = 0.0f;
  span2 = fmax_nan(span2, len2(S.p0));
  if (S.n > 1) span2 = fmax_nan(span2, len2(S.p1));
  if (S.n > 2) span2 = fmax_nan(span2, len2(S.p2));
  if (S.n > 3) span2 = fmax_nan(span2, len2(S.p3));
  if (len2(closest) < span2 * 1e-12f) {
    dir = normal;
  } else {
    S.p0 = n0;
    S.p1 = n1;
    S.n = 2;
    dir = -closest;
  }
  return false;
}

BCH_HD V3 toward_origin(V3 n, V3 t) { return sel(dot(n, t) < 0.0f, -n, n); }

BCH_HD bool solve_simplex(WSimp& S, V3& dir) {
  if (S.n == 2) {  // gjk.rs:624-646
    V3 a = S.p1, b = S.p0;
    V3 ab = b - a, ao = -a;
    if (dot(ab, ao) > 0.0f) {
      V3 cr = cross(ab, ao);
      if (len2(cr) > len2(ab) * len2(ao) * 1e-10f) dir = cross(cr, ab);
      else dir = perpendicular(ab);
    } else {
      S.p0 = a;
      S.n = 1;
      dir = ao;
    }
    return false;
  }
  if (S.n == 3) {  // gjk.rs:687-773
    V3 a = S.p2, b = S.p1, c = S.p0;
    V3 ab = b - a, ac = c - a;
    V3 normal = cross(ab, ac);
    V3 ap = -a;
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) { S.p0 = a; S.n = 1; dir = -a; return false; }
    V3 bp = -b;
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) { S.p0 = b; S.n = 1; dir = -b; return false; }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
      float v = d1 == d3 ? 0.0f : d1 / (d1 - d3);
      return shrink_or_normal(S, dir, a + ab * v, normal, b, a);
    }
    V3 cp = -c;
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) { S.p0 = c; S.n = 1; dir = -c; return false; }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
      float w = d2 == d6 ? 0.0f : d2 / (d2 - d6);
      return shrink_or_normal(S, dir, a + ac * w, normal, c, a);
    }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
      float w = (d4 - d3) + (d5 - d6) == 0.0f ? 0.0f : (d4 - d3) / ((d4 - d3) + (d5 - d6));
      return shrink_or_normal(S, dir, b + (c - b) * w, normal, c, b);
    }
    float sum = va + vb + vc;
    if (sum <= 0.0f || is_nan(sum)) { dir = normal; return false; }
    float denom = 1.0f / sum;
    float v = vb * denom, w = vc * denom;
    V3 closest = (a + ab * v) + ac * w;
    float span2 = fmax_nan(fmax_nan(len2(a), len2(b)), len2(c));
    if (len2(closest) < span2 * 1e-12f) dir = normal;
    else dir = -closest;
    return false;
  }
  if (S.n == 4) {  // gjk.rs:811-846
    V3 a = S.p3, b = S.p2, c = S.p1, d = S.p0;
    V3 ab = b - a, ac = c - a, ad = d - a, ao = -a, bo = -b;
    V3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab);
    V3 bcd = cross(b - c, d - c);
    if (dot(abc, ao) * dot(abc, ad) < 0.0f) { S.p0 = c; S.p1 = b; S.p2 = a; S.n = 3; dir = toward_origin(abc, ao); return false; }
    if (dot(acd, ao) * dot(acd, ab) < 0.0f) { S.p0 = d; S.p1 = c; S.p2 = a; S.n = 3; dir = toward_origin(acd, ao); return false; }
    if (dot(adb, ao) * dot(adb, ac) < 0.0f) { S.p0 = d; S.p1 = b; S.p2 = a; S.n = 3; dir = toward_origin(adb, ao); return false; }
    if (dot(bcd, bo) * dot(bcd, a - b) < 0.0f) { S.p0 = d; S.p1 = c; S.p2 = b; S.n = 3; dir = toward_origin(bcd, bo); return false; }
    return true;
  }
  dir = -S.p0;
  return false;
}

template <class SupF>
BCH_HD bool enclosing_walk_t(SupF sup, V3 fallback, V3 seed, WSimp& S) {
  V3 initial = sel(len2(seed) > 1e-12f, seed, fallback);
  S.p0 = sup(initial).w;
  S.p1 = S.p2 = S.p3 = S.p0;
  S.n = 1;
  V3 dir = -S.p0;
  for (int it = 0; it < 64; ++it) {
    float dir2 = len2(dir);
    if (dir2 <= 0.0f || is_nan(dir2)) dir = perpendicular(S.p0);
    V3 a = sup(dir).w;
    if (dot(a, dir) < 0.0f) return false;
    S.p3 = S.p2;
    S.p2 = S.p1;
    S.p1 = S.p0;
    S.p0 = a;
    S.n += 1;
    if (solve_simplex(S, dir)) return true;
    V3 last = sel(S.n == 1, S.p0, sel(S.n == 2, S.p1, sel(S.n == 3, S.p2, S.p3)));
    V3 reference = sel(S.n >= 2, S.p0 - last, S.p0);
    if (len2(dir) < len2(reference) * 1e-12f) dir = perpendicular(reference);
  }
  return false;
}

struct GjkResult {
  int kind;  // 0 separated, 1 band, 2 overlap (tet in t0..t3)
  float distance;
  V3 normal, point_on_b;
  V3 t0, t1, t2, t3;
  // Signed-distance walks only (SD = true): the active simplex at a Band
  // exit; at the iteration cap, the closest iterate (least |v| of all
  // iterations: an f32 walk that misses the 1 - 1e-6 exit by rounding cycles
  // through worse simplices afterwards) with its active simplex.
  bool capped;
  // Set with capped: the walk had proven a core gap (an iterate with
  // delta / |v| > 1e-4 m) and later ended on a simplex enclosing or touching
  // the origin, which only rounding can produce (a flat simplex, or
  // coordinates far from the origin); the closest iterate stands.
  bool separated;
  // The least support-plane bound -delta / |v| over the iterations: the
  // core penetration depth is at most this (EPA beyond it is contradicted).
  float depth_bound;
  V3 last_v, last_pb;
  Simp act;
  // Signed-distance walks: the cores touched (|v|^2 below the touch gate)
  // and the enclosing walk found no tetrahedron, where the reference reports
  // Separated.  act then holds the touching simplex, last_pb its B witness
  // and last_v the walk's last support direction (the initial direction on
  // the first iteration), from which the reported normal is taken.
  bool touch;
};

// gjk.rs:171-288 -- the margin-core distance walk.
// The support is a functor SV(V3 dir): one lane's scan (core_support) or a
// lane group's cooperative scan of a hull; every lane of a group runs the
// walk on identical values.
template <class SupF, bool SD = false>
BCH_HD GjkResult gjk_contact_t(const Pair& pr, bool has_seed, V3 seed, float threshold, SupF sup) {
  GjkResult R;
  R.kind = 0;
  if (SD) {
    R.capped = R.separated = R.touch = false;
    R.depth_bound = INFINITY;
    R.last_v = R.last_pb = V3{0.0f, 0.0f, 0.0f};
    R.act.n = 0;
  }
  float margin_a = pr.a.margin, margin_b = pr.b.margin;
  float margin_sum = margin_a + margin_b;
  V3 center_dir = pr.pb - pr.pa;
  V3 fallback = sel(len2(center_dir) > 1e-12f, center_dir, V3{1.0f, 0.0f, 0.0f});
  V3 initial = sel(has_seed && len2(seed) > 1e-12f, seed, fallback);
  SV first = sup(initial);
  Simp S;
  S.s0 = S.s1 = S.s2 = S.s3 = first;
  S.n = 1;
  V3 dir = -first.w;
  float touch2 = 1e-12f * fmax_nan(len2(first.w), 1.0f);
  float best_v2 = INFINITY;
  bool proven_gap = false;
  V3 search = initial;  // the direction of the latest support query (SD only)
  for (int it = 0; it < 64; ++it) {
    Reduce r = reduce(S);
    V3 v = r.v;
    float v2 = len2(v);
    if (SD && proven_gap && (r.inside || v2 < touch2)) {
      R.capped = R.separated = true;
      return R;
    }
    if (r.inside) {
      R.kind = 2;
      R.t0 = S.s0.w;
      R.t1 = S.s1.w;
      R.t2 = S.s2.w;
      R.t3 = S.s3.w;
      return R;
    }
    if (v2 < touch2) {
      WSimp W;
      bool ok = enclosing_walk_t(sup, fallback, dir, W);
      if (ok && W.n == 4) {
        R.kind = 2;
        R.t0 = W.p0;
        R.t1 = W.p1;
        R.t2 = W.p2;
        R.t3 = W.p3;
      } else if (SD) {
        R.touch = true;
        R.act = S;
        R.last_v = search;
        R.last_pb = r.p_b;
      }
      return R;
    }
    dir = -v;
    if (SD) search = dir;
    SV w = sup(dir);
    float delta = dot(v, w.w);
    float band = margin_sum + threshold;
    if (delta > 0.0f && delta * delta > v2 * band * band) return R;
    if (SD && delta > 0.0f && delta * delta > v2 * (1e-4f * 1e-4f)) proven_gap = true;
    if (SD) {
      float bound = -delta / fsqrt(v2);
      if (bound < R.depth_bound) R.depth_bound = bound;
    }
    bool dup = near_eq(w.w, S.s0.w, 1e-6f);
    if (S.n > 1) dup = dup || near_eq(w.w, S.s1.w, 1e-6f);
    if (S.n > 2) dup = dup || near_eq(w.w, S.s2.w, 1e-6f);
    if (S.n > 3) dup = dup || near_eq(w.w, S.s3.w, 1e-6f);
    if (dup || delta >= v2 * (1.0f - 1e-6f)) {
      float core_distance = fsqrt(v2);
      float distance = core_distance - margin_sum;
      if (distance < threshold) {
        V3 normal = dir / core_distance;
        R.kind = 1;
        R.distance = distance;
        R.normal = normal;
        R.point_on_b = r.p_b - normal * margin_b;
        if (SD) R.act = S;
      }
      return R;
    }
    if (SD && !(v2 >= best_v2)) {  // the first of equal iterates; NaN passes
      best_v2 = v2;
      R.act = S;
      R.last_v = v;
      R.last_pb = r.p_b;
    }
    simp_append(S, w);
  }
  if (SD) R.capped = true;
  return R;
}

struct LaneSupport {
  const Pair* pr;
  BCH_HD SV operator()(V3 d) const { return core_support(*pr, d); }
};

BCH_HD GjkResult gjk_contact(const Pair& pr, bool has_seed, V3 seed, float threshold) {
  return gjk_contact_t(pr, has_seed, seed, threshold, LaneSupport{&pr});
}

// -------------------------------------------------------------------- EPA --
// The polytope store (epa.rs:87-103): up to 128 vertices, EPA_MAX_FACES faces,
// horizon edges; the caller points these at LDS or global scratch.
#define BCH_EPA_MAX_VERTS 128
#define BCH_EPA_MAX_FACES 256
#define BCH_EPA_MAX_EDGES 384

struct EpaFace {
  V3 n;
  float dist, pd;
  int abc;  // a | b << 8 | c << 16
};

struct EpaStore {
  float4* verts;     // [BCH_EPA_MAX_VERTS]
  EpaFace* faces;    // [BCH_EPA_MAX_FACES]
  uint32_t* edges;   // [BCH_EPA_MAX_EDGES], u | v << 8
};

struct EpaResult {
  int status;  // 0 converged, 1 degenerate, 2 iteration limit, 3 vertex limit, 4 store overflow
  V3 n;
  float depth;
  int face;    // index of the reported face (wavefront EPA)
};

BCH_HD V3 ev(const EpaStore& st, int i) {
  float4 v = st.verts[i];
  return V3{v.x, v.y, v.z};
}

// epa.rs:283-303
BCH_HD bool edge_dist(V3 pa, V3 pb, V3 n, float& out) {
  V3 ba = pb - pa;
  V3 n_ab = cross(ba, n);
  if (dot(pa, n_ab) < 0.0f) {
    float ba_l2 = len2(ba);
    float a_dot_ba = dot(pa, ba);
    float b_dot_ba = dot(pb, ba);
    if (a_dot_ba > 0.0f) out = len(pa);
    else if (b_dot_ba < 0.0f) out = len(pb);
    else {
      float a_dot_b = dot(pa, pb);
      // f32::max(0.0) keeps the non-NaN operand: an overflowed quotient gives 0
      float q = (len2(pa) * len2(pb) - a_dot_b * a_dot_b) / ba_l2;
      out = fsqrt(q > 0.0f ? q : 0.0f);
    }
    return true;
  }
  return false;
}

// epa.rs:262-278
BCH_HD float face_dist(V3 n, V3 pa, V3 pb, V3 pc) {
  float d;
  if (edge_dist(pa, pb, n, d)) return d;
  if (edge_dist(pb, pc, n, d)) return d;
  if (edge_dist(pc, pa, n, d)) return d;
  return dot(pa, n);
}

BCH_HD bool face_vis(const EpaFace& f, V3 p) { return dot(f.n, p) - f.pd > 1e-5f; }

// epa.rs:309-388.  Returns 1 ok, 0 degenerate (hull untouched), -1 store overflow.
BCH_HD int epa_expand(const EpaStore& st, int& nfaces, V3 np, int nidx) {
  int nedges = 0;
  bool any = false;
  for (int fi = 0; fi < nfaces; ++fi) {
    EpaFace f = st.faces[fi];
    if (!face_vis(f, np)) continue;
    any = true;
    int fa = f.abc & 0xff, fb = (f.abc >> 8) & 0xff, fc = (f.abc >> 16) & 0xff;
    for (int e = 0; e < 3; ++e) {
      int u = e == 0 ? fa : (e == 1 ? fb : fc);
      int v = e == 0 ? fb : (e == 1 ? fc : fa);
      uint32_t rev = (uint32_t)v | ((uint32_t)u << 8);
      int found = -1;
      for (int k = 0; k < nedges; ++k)
        if (st.edges[k] == rev) { found = k; break; }
      if (found >= 0) {
        st.edges[found] = st.edges[nedges - 1];
        --nedges;
      } else {
        if (nedges >= BCH_EPA_MAX_EDGES) return -1;
        st.edges[nedges++] = (uint32_t)u | ((uint32_t)v << 8);
      }
    }
  }
  if (!any) return 1;
  for (int k = 0; k < nedges; ++k) {
    int u = st.edges[k] & 0xff, v = (st.edges[k] >> 8) & 0xff;
    V3 vu = ev(st, u);
    if (len(cross(ev(st, v) - vu, np - vu)) < 1e-12f) return 0;
  }
  // Commit: stable removal of the visible faces, then the horizon fan.
  int w = 0;
  for (int fi = 0; fi < nfaces; ++fi) {
    EpaFace f = st.faces[fi];
    if (!face_vis(f, np)) st.faces[w++] = f;
  }
  if (w + nedges > BCH_EPA_MAX_FACES) return -1;
  for (int k = 0; k < nedges; ++k) {
    int u = st.edges[k] & 0xff, v = (st.edges[k] >> 8) & 0xff;
    V3 vu = ev(st, u), vv = ev(st, v);
    V3 cr = cross(vv - vu, np - vu);
    float l = len(cr);
    V3 n = cr / l;
    EpaFace f;
    f.n = n;
    f.dist = face_dist(n, vu, vv, np);
    f.pd = dot(n, vu);
    f.abc = u | (v << 8) | (nidx << 16);
    st.faces[w++] = f;
  }
  nfaces = w;
  return 1;
}

BCH_HD int epa_best(const EpaStore& st, int nfaces) {
  int best = 0;
  float bd = fabsf(st.faces[0].dist);
  for (int i = 1; i < nfaces; ++i) {
    float d = fabsf(st.faces[i].dist);
    if (bd > d) {
      best = i;
      bd = d;
    }
  }
  return best;
}

// epa.rs:127-205 over the core Minkowski support of `pr`.
BCH_HD EpaResult epa_solve(const Pair& pr, V3 t0, V3 t1, V3 t2, V3 t3, const EpaStore& st) {
  EpaResult R;
  V3 s[4] = {t0, t1, t2, t3};
  for (int i = 0; i < 4; ++i) st.verts[i] = float4{s[i].x, s[i].y, s[i].z, 0.0f};
  int nverts = 4, nfaces = 0;
  const int quads[4][4] = {{0, 1, 2, 3}, {0, 3, 1, 2}, {0, 2, 3, 1}, {1, 3, 2, 0}};
  for (int qi = 0; qi < 4; ++qi) {  // epa.rs:226-257
    int a = quads[qi][0], b = quads[qi][1], c = quads[qi][2], opp = quads[qi][3];
    V3 sa = s[a];
    V3 n = cross(s[b] - sa, s[c] - sa);
    if (dot(n, s[opp] - sa) > 0.0f) {
      n = -n;
      int t = b;
      b = c;
      c = t;
    }
    float l = len(n);
    if (l < 1e-12f) continue;
    n = n / l;
    EpaFace f;
    f.n = n;
    f.dist = face_dist(n, sa, s[b], s[c]);
    f.pd = dot(n, sa);
    f.abc = a | (b << 8) | (c << 16);
    st.faces[nfaces++] = f;
  }
  if (nfaces == 0) {
    R.status = 1;
    R.n = splat(0.0f);
    R.depth = 0.0f;
    return R;
  }
  // Seed with the six axis supports; a degenerate seeding expansion ends the
  // seeding only -- the main loop still runs on the intact hull (epa.rs:148-168).
  int status = 2;
  bool stop = false;
  for (int k = 0; k < 6; ++k) {
    V3 d = k == 0 ? V3{1.0f, 0.0f, 0.0f} : k == 1 ? V3{-1.0f, 0.0f, 0.0f} : k == 2 ? V3{0.0f, 1.0f, 0.0f}
         : k == 3 ? V3{0.0f, -1.0f, 0.0f} : k == 4 ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 0.0f, -1.0f};
    V3 p = core_support(pr, d).w;
    int idx = nverts;
    st.verts[nverts++] = float4{p.x, p.y, p.z, 0.0f};
    int e = epa_expand(st, nfaces, p, idx);
    if (e <= 0) {
      status = e == 0 ? 1 : 4;
      stop = e < 0;
      break;
    }
  }
  if (!stop) {
    for (int it = 0; it < 255; ++it) {
      int bi = epa_best(st, nfaces);
      EpaFace bf = st.faces[bi];
      V3 np = core_support(pr, bf.n).w;
      float wdist = dot(bf.n, np) - bf.dist;
      if (wdist <= 1e-4f) {
        R.status = 0;
        R.n = bf.n;
        R.depth = bf.dist;
        return R;
      }
      if (nverts >= BCH_EPA_MAX_VERTS) {
        status = 3;
        break;
      }
      int nidx = nverts;
      st.verts[nverts++] = float4{np.x, np.y, np.z, 0.0f};
      int e = epa_expand(st, nfaces, np, nidx);
      if (e <= 0) {
        status = e == 0 ? 1 : 4;
        break;
      }
    }
  }
  int bi = epa_best(st, nfaces);
  R.status = status;
  R.n = st.faces[bi].n;
  R.depth = st.faces[bi].dist;
  return R;
}

// convex_convex.rs:117-154: margins added back, witness = supportA(n) - n depth.
BCH_HD bool epa_contact(const Pair& pr, const EpaResult& e, Manifold4& m) {
  if (e.depth <= 0.0f) return false;
  float depth = (e.depth + pr.a.margin) + pr.b.margin;
  V3 point_a = support_world(pr.a, pr.pa, pr.qa, e.n);
  m.push(point_a - e.n * depth, e.n, depth);
  return true;
}

// ---------------------------------------------------------- pair classes --
// Which routine collide_seeded (narrowphase/mod.rs:99-285) sends a pair to.
// CLS_GJK_HULL: the same GJK/EPA route as CLS_GJK for pairs with a hull,
// evaluated by lane groups (the hull support scan is split over the group).
enum : int {
  CLS_NONE = 0,
  CLS_SPHERE_SPHERE = 1,
  CLS_SPHERE_BOX = 2,
  CLS_PLANE = 3,
  CLS_BOX_BOX = 4,
  CLS_GJK = 5,
  CLS_GJK_HULL = 6
};

BCH_HD int pair_class(int ta, int tb) {
  if (ta == PLANE && tb == PLANE) return CLS_NONE;
  if (ta == PLANE || tb == PLANE) return CLS_PLANE;
  if (ta == SPHERE && tb == SPHERE) return CLS_SPHERE_SPHERE;
  if ((ta == SPHERE && tb == BOX) || (ta == BOX && tb == SPHERE)) return CLS_SPHERE_BOX;
  if (ta == BOX && tb == BOX) return CLS_BOX_BOX;
  if (ta == HULL || tb == HULL) return CLS_GJK_HULL;
  return CLS_GJK;
}

// collide_seeded for every class but the EPA continuation.  Returns:
//   0 -> done (m holds 0..4 points), 1 -> needs EPA (g holds the tetrahedron).
BCH_HD int collide_pair(const Pair& pr, float threshold, bool has_seed, V3 seed, Manifold4& m, GjkResult& g,
                        int cls) {
  m.n = 0;
  switch (cls) {
    case CLS_NONE: return 0;
    case CLS_PLANE:
      if (pr.a.type == PLANE) {
        plane_convex(V3{pr.a.a0, pr.a.a1, pr.a.a2}, pr.a.a3, pr.pa, pr.qa, pr.b, pr.pb, pr.qb, threshold, m);
      } else {
        if (plane_convex(V3{pr.b.a0, pr.b.a1, pr.b.a2}, pr.b.a3, pr.pb, pr.qb, pr.a, pr.pa, pr.qa, threshold, m))
          m.flip();
      }
      return 0;
    case CLS_SPHERE_SPHERE: sphere_sphere(pr.a.a0, pr.pa, pr.b.a0, pr.pb, m); return 0;
    case CLS_SPHERE_BOX:
      if (pr.a.type == SPHERE) {
        if (sphere_box(pr.a.a0, pr.pa, V3{pr.b.a0, pr.b.a1, pr.b.a2}, pr.pb, pr.qb, threshold, m)) m.flip();
      } else {
        sphere_box(pr.b.a0, pr.pb, V3{pr.a.a0, pr.a.a1, pr.a.a2}, pr.pa, pr.qa, threshold, m);
      }
      return 0;
    case CLS_BOX_BOX:
      box_box(V3{pr.a.a0, pr.a.a1, pr.a.a2}, pr.pa, pr.qa, V3{pr.b.a0, pr.b.a1, pr.b.a2}, pr.pb, pr.qb, m);
      return 0;
    default: {
      g = gjk_contact(pr, has_seed, seed, threshold);
      if (g.kind == 1) {
        m.push(g.point_on_b, g.normal, -g.distance);
        return 0;
      }
      return g.kind == 2 ? 1 : 0;
    }
  }
}

}  // namespace bch
