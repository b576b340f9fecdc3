// bch_adjoint.hpp -- signed distance between convex pairs and its adjoint
// with respect to both bodies' poses (position xyz, quaternion xyzw).
//
// Forward (per pair): the reference's margin-core walk (gjk.rs:171-288) with
// an unbounded band (threshold = +inf), so a separated pair always ends in
// its Band arm; overlapping cores go to EPA (epa.rs:127-205) and the contact
// of convex_convex.rs:117-154.  A walk stopped by its 64-iteration cap
// reports its last iterate as the Band arm would.  Outputs per pair, 10
// floats: sd (> 0 separated), unit normal A -> B, witness on A, witness on B
// (pa = pb - n sd).
//
// Backward: recomputation, no stored simplices.  The walk runs again; the
// Band arm's closest point v = sum l_i w_i of the final active simplex (w_i
// = a_i - b_i, support points of the two cores) is differentiated through
// the affine least-norm problem that fixes the weights l; EPA runs again and
// its reported face (vertices w_a, w_b, w_c) is differentiated through the
// plane distance n.w_a with n = normalize((w_b - w_a) x (w_c - w_a)).  Each
// support point is a fixed local point x of its shape (hull vertex, core box
// corner, capsule tip, sphere centre) placed by qrot(q, x) + p, so pose
// gradients come from the derivative of the quaternion-rotate formula.  The
// face vertices' split into (a, b) is found by recomputing every candidate
// difference of the two support sets and matching it exactly (the same
// arithmetic produced them).
#pragma once
#include "bch_epa_quad.hpp"
#include "bch_epa_wave.hpp"

namespace bch {

enum : uint32_t {
  SD_WALK = 0,             // Band arm of the walk (separated, or inside the margins)
  SD_EPA = 1,              // EPA converged
  SD_CAPPED = 2,           // walk stopped at its iteration cap (closest iterate)
  SD_TOUCH = 3,            // cores touch and no enclosing tetrahedron: sd = -(margins), normal from the touching simplex
  SD_EPA_UNCONVERGED = 4,  // EPA ended without converging (reported face)
  SD_SEPARATED = 5,        // walk proved a core gap, then enclosed the origin by rounding (closest iterate)
  SD_EPA_CONTRADICTED = 6, // EPA depth beyond the walk's depth bound (closest iterate)
  SD_FLAG_EDGE = 8,        // EPA face distance came from an edge or vertex, not the plane
  SD_FLAG_SPLIT = 16       // a face vertex could not be split into support points
};

// float4 per queued signed-distance EPA entry: the walk's enclosing
// simplex with the pair index, then its depth bound and closest iterate.
constexpr uint32_t kSdEpaEntry = 8;

// Pose gradient accumulator for one pair: position A, quaternion A,
// position B, quaternion B.
struct PoseGrad {
  V3 pa, pb;
  float qa[4], qb[4];
  BCH_HD void zero() {
    pa = pb = splat(0.0f);
    for (int k = 0; k < 4; ++k) qa[k] = qb[k] = 0.0f;
  }
};

// Band arm (or capped walk): outputs from the active simplex and the VJP.
// act holds k = act.n vertices (w, b); v is the walk's closest point.
struct SdOut {
  float sd;
  V3 n, pa, pb;
};

BCH_HD SdOut band_outputs(const Pair& pr, V3 v, V3 p_b) {
  float core = fsqrt(len2(v));
  SdOut o;
  o.sd = core - (pr.a.margin + pr.b.margin);
  o.n = (-v) / core;
  o.pb = p_b - o.n * pr.b.margin;
  o.pa = o.pb - o.n * o.sd;
  return o;
}

// The normal reported for touching cores: the walk's last support direction
// s with its components along the touching simplex removed (a triangle's
// face normal, a segment's perpendicular; a single point keeps s).
BCH_HD V3 sd_touch_normal(const Simp& S, V3 s) {
  V3 n = s;
  if (S.n == 3) {
    V3 c = cross(S.s1.w - S.s0.w, S.s2.w - S.s0.w);
    float c2 = len2(c);
    if (c2 > 0.0f) n = c * (dot(c, s) / c2);
  } else if (S.n == 2) {
    V3 e = S.s1.w - S.s0.w;
    float e2 = len2(e);
    if (e2 > 0.0f) n = s - e * (dot(s, e) / e2);
  }
  float n2 = len2(n);
  if (!(n2 > 0.0f)) {
    n = s;
    n2 = len2(s);
  }
  return n / fsqrt(n2);
}

// Core support point `idx` of a shape in local coordinates (the finite set
// the core supports are drawn from) and the set's size.
BCH_HD int core_set_size(const ShapeRef& s) {
  return s.type == HULL ? s.nhull : (s.type == BOX ? 8 : (s.type == CAPSULE ? 2 : 1));
}
BCH_HD V3 core_set_point(const ShapeRef& s, int idx) {
  if (s.type == HULL) return hull_vertex(s.hull, idx);
  if (s.type == BOX) {
    float hx = mx(s.a0 - 0.04f, 0.0f), hy = mx(s.a1 - 0.04f, 0.0f), hz = mx(s.a2 - 0.04f, 0.0f);
    return V3{(idx & 1) ? -hx : hx, (idx & 2) ? -hy : hy, (idx & 4) ? -hz : hz};
  }
  if (s.type == CAPSULE) return V3{0.0f, idx ? -s.a1 : s.a1, 0.0f};
  return splat(0.0f);
}

// ---- double-precision Band arm.  Near-touching cores make v = a - b a
// difference of nearly equal points: its f32 rounding, amplified by the
// 1/|v| of the normal's derivative, would dominate the gradient.  The walk
// records which local point every support came from (hull vertex, core box
// corner, capsule tip, sphere centre); the adjoint rebuilds a, b from them in
// double and differentiates the least-norm point there.
struct D3 {
  double x, y, z;
};
BCH_HD D3 d3(V3 v) { return D3{(double)v.x, (double)v.y, (double)v.z}; }
BCH_HD D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BCH_HD D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
BCH_HD D3 operator-(D3 a) { return D3{-a.x, -a.y, -a.z}; }
BCH_HD D3 operator*(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
BCH_HD double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BCH_HD D3 crossd(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
BCH_HD D3 qrotd(Q4 q, D3 v) {
  D3 b{(double)q.x, (double)q.y, (double)q.z};
  double w = q.w;
  return (v * (w * w - dotd(b, b)) + b * (2.0 * dotd(v, b))) + crossd(b, v) * (2.0 * w);
}
// J^T g of qrot(q, x) = x (w^2 - b.b) + b (2 x.b) + (b x x)(2w) with respect
// to q = (b, w), added to gq.
BCH_HD void qrot_vjpd(Q4 q, D3 x, D3 g, double gq[4]) {
  D3 b{(double)q.x, (double)q.y, (double)q.z};
  double w = q.w;
  D3 gb = ((g * dotd(x, b) + x * dotd(b, g)) - b * dotd(x, g) + crossd(x, g) * w) * 2.0;
  gq[0] += gb.x;
  gq[1] += gb.y;
  gq[2] += gb.z;
  gq[3] += 2.0 * (w * dotd(x, g) + dotd(crossd(b, x), g));
}

struct AffineD {
  double lam[3], mu0, mu1, i00, i01, i11;
  D3 D0, D1;
};

BCH_HD AffineD affine_weights_d(const D3* w, int k) {
  AffineD A;
  A.lam[0] = 1.0;
  A.lam[1] = A.lam[2] = 0.0;
  A.mu0 = A.mu1 = A.i00 = A.i01 = A.i11 = 0.0;
  A.D0 = A.D1 = D3{0.0, 0.0, 0.0};
  if (k == 2) {
    A.D0 = w[1] - w[0];
    A.i00 = 1.0 / dotd(A.D0, A.D0);
    A.mu0 = -dotd(A.D0, w[0]) * A.i00;
    A.lam[0] = 1.0 - A.mu0;
    A.lam[1] = A.mu0;
  } else if (k == 3) {
    A.D0 = w[1] - w[0];
    A.D1 = w[2] - w[0];
    double m00 = dotd(A.D0, A.D0), m01 = dotd(A.D0, A.D1), m11 = dotd(A.D1, A.D1);
    double det = m00 * m11 - m01 * m01;
    A.i00 = m11 / det;
    A.i01 = -m01 / det;
    A.i11 = m00 / det;
    double r0 = -dotd(A.D0, w[0]), r1 = -dotd(A.D1, w[0]);
    A.mu0 = A.i00 * r0 + A.i01 * r1;
    A.mu1 = A.i01 * r0 + A.i11 * r1;
    A.lam[0] = (1.0 - A.mu0) - A.mu1;
    A.lam[1] = A.mu0;
    A.lam[2] = A.mu1;
  }
  return A;
}

// touch_dir (touching cores, SD_TOUCH): the witnesses move with their bodies
// and the reported normal n is held fixed, d sd = n . (d b - d a) over the
// touching simplex's weights; n is rebuilt in double from the simplex and
// the walk's last support direction as the forward builds it in f32.
BCH_HD void band_vjp(const Pair& pr, const Simp& act, const float* g10, PoseGrad& G,
                     const V3* touch_dir = nullptr) {
  int k = act.n;
  const SV* sv[3] = {&act.s0, &act.s1, &act.s2};
  D3 xa[3], xb[3], w[3], b[3];
  for (int i = 0; i < 3; ++i) {
    xa[i] = d3(core_set_point(pr.a, sv[i]->ia));
    xb[i] = d3(core_set_point(pr.b, sv[i]->ib));
    D3 ai = qrotd(pr.qa, xa[i]) + d3(pr.pa);
    b[i] = qrotd(pr.qb, xb[i]) + d3(pr.pb);
    w[i] = ai - b[i];
  }
  // the least-norm point's feature: drop vertices whose exact weight is not
  // positive (the f32 walk can keep one at rounding level)
  AffineD A = affine_weights_d(w, k);
  for (int pass = 0; pass < 2 && k > 1 && !touch_dir; ++pass) {
    int drop = -1;
    double low = 0.0;
    for (int i = 0; i < 3; ++i)
      if (i < k && A.lam[i] <= low) {
        low = A.lam[i];
        drop = i;
      }
    if (drop < 0) break;
    for (int i = drop; i < 2; ++i) {
      w[i] = w[i + 1];
      b[i] = b[i + 1];
      xa[i] = xa[i + 1];
      xb[i] = xb[i + 1];
    }
    --k;
    A = affine_weights_d(w, k);
  }
  D3 v{0.0, 0.0, 0.0};
  for (int i = 0; i < 3; ++i)
    if (i < k) v = v + w[i] * A.lam[i];
  double ma = pr.a.margin, mb = pr.b.margin;
  double core = sqrt(dotd(v, v));
  D3 nv = v * (1.0 / core);
  D3 n = -nv;
  double sd = core - (ma + mb);
  if (touch_dir) {
    D3 s = d3(*touch_dir);
    n = s;
    if (k == 3) {
      D3 c = crossd(w[1] - w[0], w[2] - w[0]);
      double c2 = dotd(c, c);
      if (c2 > 0.0) n = c * (dotd(c, s) / c2);
    } else if (k == 2) {
      D3 e = w[1] - w[0];
      double e2 = dotd(e, e);
      if (e2 > 0.0) n = s - e * (dotd(s, e) / e2);
    }
    double n2 = dotd(n, n);
    if (!(n2 > 0.0)) {
      n = s;
      n2 = dotd(s, s);
    }
    n = n * (1.0 / sqrt(n2));
    sd = -(ma + mb);
  }
  D3 gn{g10[1], g10[2], g10[3]}, gpa{g10[4], g10[5], g10[6]}, gpb{g10[7], g10[8], g10[9]};
  D3 g_pbt = gpb + gpa;
  D3 g_nt = (gn - gpa * sd) - g_pbt * mb;
  double g_sdt = (double)g10[0] - dotd(n, gpa);
  D3 g_v = touch_dir ? n * (-g_sdt) : nv * g_sdt - (g_nt - n * dotd(n, g_nt)) * (1.0 / core);
  D3 gw[3], gb[3];
  double glam[3];
  for (int i = 0; i < 3; ++i) {
    gw[i] = g_v * A.lam[i];
    gb[i] = g_pbt * A.lam[i];
    glam[i] = dotd(w[i], g_v) + dotd(b[i], g_pbt);
  }
  if (k >= 2 && !touch_dir) {
    // g_mu_j = g_lam_j - g_lam_0; u = M^-1 g_mu;
    // g_D = -(v u^T + (D u) mu^T), g_w0 -= D u
    double gm0 = glam[1] - glam[0], gm1 = k == 3 ? glam[2] - glam[0] : 0.0;
    double u0 = k == 3 ? A.i00 * gm0 + A.i01 * gm1 : A.i00 * gm0;
    double u1 = k == 3 ? A.i01 * gm0 + A.i11 * gm1 : 0.0;
    D3 Du = A.D0 * u0 + A.D1 * u1;
    D3 gD0 = -(v * u0 + Du * A.mu0);
    D3 gD1 = -(v * u1 + Du * A.mu1);
    gw[0] = (gw[0] - Du) - gD0;
    gw[1] = gw[1] + gD0;
    if (k == 3) {
      gw[2] = gw[2] + gD1;
      gw[0] = gw[0] - gD1;
    }
  }
  D3 gpA{0.0, 0.0, 0.0}, gpB{0.0, 0.0, 0.0};
  double gqa[4] = {0, 0, 0, 0}, gqb[4] = {0, 0, 0, 0};
  for (int i = 0; i < 3; ++i) {
    if (i >= k) break;
    gpA = gpA + gw[i];
    qrot_vjpd(pr.qa, xa[i], gw[i], gqa);
    D3 gbi = gb[i] - gw[i];
    gpB = gpB + gbi;
    qrot_vjpd(pr.qb, xb[i], gbi, gqb);
  }
  G.pa = G.pa + V3{(float)gpA.x, (float)gpA.y, (float)gpA.z};
  G.pb = G.pb + V3{(float)gpB.x, (float)gpB.y, (float)gpB.z};
  for (int c = 0; c < 4; ++c) {
    G.qa[c] += (float)gqa[c];
    G.qb[c] += (float)gqb[c];
  }
}

// EPA arm in double: the reported face rebuilt from its exact local support
// points, depth = n . w_a + margins with n = normalize((w_b - w_a) x
// (w_c - w_a)), witness pa = supportA(n) (its support point chosen by the
// forward normal n_fwd), pb = pa - n depth.
BCH_HD void epa_vjp(const Pair& pr, V3 n_fwd, const V3* xa_f, const V3* xb_f, const float* g10, PoseGrad& G) {
  D3 xa[3], xb[3], w[3];
  for (int i = 0; i < 3; ++i) {
    xa[i] = d3(xa_f[i]);
    xb[i] = d3(xb_f[i]);
    w[i] = (qrotd(pr.qa, xa[i]) + d3(pr.pa)) - (qrotd(pr.qb, xb[i]) + d3(pr.pb));
  }
  D3 e1 = w[1] - w[0], e2 = w[2] - w[0];
  D3 cc = crossd(e1, e2);
  double cl = sqrt(dotd(cc, cc));
  D3 n = cc * (1.0 / cl);
  double depth = (dotd(n, w[0]) + (double)pr.a.margin) + (double)pr.b.margin;
  D3 gn{g10[1], g10[2], g10[3]}, gpa{g10[4], g10[5], g10[6]}, gpb{g10[7], g10[8], g10[9]};
  D3 g_pat = gpa + gpb;
  D3 g_nt = gn - gpb * depth;
  double g_depth = -(double)g10[0] - dotd(n, gpb);
  D3 gpA = g_pat, gpB{0.0, 0.0, 0.0};
  double gqa[4] = {0, 0, 0, 0}, gqb[4] = {0, 0, 0, 0};
  // witness on A
  const ShapeRef& s = pr.a;
  V3 ld = qrot(qconj(pr.qa), n_fwd);
  if (s.type == SPHERE || s.type == CAPSULE) {
    V3 tip = s.type == CAPSULE ? V3{0.0f, ld.y >= 0.0f ? s.a1 : -s.a1, 0.0f} : splat(0.0f);
    qrot_vjpd(pr.qa, d3(tip), g_pat, gqa);
    double r = s.a0, ng = dotd(n, g_pat);
    // qrot(q, r ld/|ld|) = r |q|^2 n: d/dq = 2 r (n.g) q, d/dn = r (I - n n^T)
    gqa[0] += 2.0 * r * ng * pr.qa.x;
    gqa[1] += 2.0 * r * ng * pr.qa.y;
    gqa[2] += 2.0 * r * ng * pr.qa.z;
    gqa[3] += 2.0 * r * ng * pr.qa.w;
    g_nt = g_nt + (g_pat - n * ng) * r;
  } else {
    qrot_vjpd(pr.qa, d3(support_local(s, ld)), g_pat, gqa);
  }
  // plane distance and face normal
  g_nt = g_nt + w[0] * g_depth;
  D3 g_c = (g_nt - n * dotd(n, g_nt)) * (1.0 / cl);
  D3 g_e1 = crossd(e2, g_c), g_e2 = crossd(g_c, e1);
  D3 gw[3] = {(n * g_depth - g_e1) - g_e2, g_e1, g_e2};
  for (int i = 0; i < 3; ++i) {
    gpA = gpA + gw[i];
    gpB = gpB - gw[i];
    qrot_vjpd(pr.qa, xa[i], gw[i], gqa);
    qrot_vjpd(pr.qb, xb[i], -gw[i], gqb);
  }
  G.pa = G.pa + V3{(float)gpA.x, (float)gpA.y, (float)gpA.z};
  G.pb = G.pb + V3{(float)gpB.x, (float)gpB.y, (float)gpB.z};
  for (int c = 0; c < 4; ++c) {
    G.qa[c] += (float)gqa[c];
    G.qb[c] += (float)gqb[c];
  }
}

BCH_HD void store_grad(float* g14, const PoseGrad& G) {
  g14[0] = G.pa.x; g14[1] = G.pa.y; g14[2] = G.pa.z;
  for (int k = 0; k < 4; ++k) g14[3 + k] = G.qa[k];
  g14[7] = G.pb.x; g14[8] = G.pb.y; g14[9] = G.pb.z;
  for (int k = 0; k < 4; ++k) g14[10 + k] = G.qb[k];
}

BCH_HD void store_out(float* o10, const SdOut& o) {
  o10[0] = o.sd;
  o10[1] = o.n.x; o10[2] = o.n.y; o10[3] = o.n.z;
  o10[4] = o.pa.x; o10[5] = o.pa.y; o10[6] = o.pa.z;
  o10[7] = o.pb.x; o10[8] = o.pb.y; o10[9] = o.pb.z;
}

// The walk for every pair, G lanes per pair (hull support scans split over
// the group).  BWD = false writes out[10P]; BWD = true reads gin[10P] and
// writes grad[14P].  Overlapping cores are queued for the EPA kernel.  P_dev
// (when given) caps P with a count computed on the device.
// BWD = true leaves the Band and touch arms' VJP to k_sd_band_vjp: the walk
// writes, per pair, the final active simplex's local support indices (and
// the touch direction) to brec (kSdBandRec float4, kind 0 = nothing there:
// an EPA pair or a zero gradient already stored), so the double-precision
// adjoint runs one lane per pair instead of on one lane of each group.
constexpr uint32_t kSdBandRec = 3;
// Three waves per SIMD (168 VGPRs, a few spills; 224 uncapped, two waves):
// the 1024-world de-penetration step runs 9.4 -> 8.1 ms, 262,144 hull pairs
// about as before (forward 1.85 -> 1.87 ms, backward 2.87 -> 2.81 ms); at four
// waves (128 VGPRs) the spills cost more than the occupancy gains.
#ifndef BCH_SD_WALK_WAVES
#define BCH_SD_WALK_WAVES 3
#endif
template <int G, bool BWD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BCH_SD_WALK_WAVES))) void k_sd_walk(uint32_t P, const uint32_t* __restrict__ P_dev, uint32_t nbodies,
                                                 const uint32_t* __restrict__ pairs,
                                                 const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                                                 const float* __restrict__ gin, float* __restrict__ out,
                                                 uint8_t* __restrict__ status, float4* __restrict__ epa_q,
                                                 Counters* cnt, float4* __restrict__ brec = nullptr) {
  P = dev_count(P, P_dev);
  __shared__ float4 hstage[256 / G][64];
  const uint32_t per_block = blockDim.x / G;
  int grp = threadIdx.x / G, sub = threadIdx.x % G;
  for (uint32_t base = blockIdx.x * per_block; base < P; base += gridDim.x * per_block) {
    uint32_t t = base + grp;
    bool need_epa = false;
    GjkResult g;
    uint32_t i = 0, j = 0;
    if (t < P) {
      i = pairs[2 * t];
      j = pairs[2 * t + 1];
    }
    bool bad = t < P && (i >= nbodies || j >= nbodies);
    if (bad && sub == 0) {
      // a pair index outside the bodies: NaN outputs, error status
      if (status) status[t] = 0xffu;
      for (int k = 0; k < (BWD ? 14 : 10); ++k) out[(BWD ? 14 : 10) * (uint64_t)t + k] = NAN;
      if (BWD) brec[kSdBandRec * (uint64_t)t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (t < P && !bad) {
      Pair pr = make_pair(bodies[i], bodies[j], hull);
      // one 64-vertex slot per group: both hulls when they fit together,
      // else one hull of at most 64 vertices (A first), the other read
      // through the cache
      const int na = pr.a.type == HULL ? (int)pr.a.nhull : 65, nb = pr.b.type == HULL ? (int)pr.b.nhull : 65;
      int used = 0;
      if (na <= 64) {
        for (int v = sub; v < na; v += G) hstage[grp][v] = pr.a.hull[v];
        pr.a.hull = hstage[grp];
        used = na;
      }
      if (nb <= 64 - used) {
        for (int v = sub; v < nb; v += G) hstage[grp][used + v] = pr.b.hull[v];
        pr.b.hull = hstage[grp] + used;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      g = gjk_contact_t<GroupSupportIdx<G>, true>(pr, false, splat(0.0f), INFINITY, GroupSupportIdx<G>{&pr, sub});
      need_epa = g.kind == 2;
      if (sub == 0 && !need_epa) {
        bool band = g.kind == 1 || g.capped;
        uint32_t st = band ? (g.kind == 1 ? SD_WALK : (g.separated ? SD_SEPARATED : SD_CAPPED)) : SD_TOUCH;
        SdOut o;
        if (band) {
          if (g.kind == 1) {
            o.sd = g.distance;
            o.n = g.normal;
            o.pb = g.point_on_b;
            o.pa = o.pb - o.n * o.sd;
          } else {
            o = band_outputs(pr, g.last_v, g.last_pb);
          }
        } else {
          o.sd = -(pr.a.margin + pr.b.margin);
          o.n = o.pa = o.pb = splat(0.0f);
          if (g.touch) {
            o.n = sd_touch_normal(g.act, g.last_v);
            o.pb = g.last_pb - o.n * pr.b.margin;
            o.pa = o.pb - o.n * o.sd;
          }
        }
        if (status) status[t] = (uint8_t)st;
        if (!BWD) {
          store_out(out + 10 * (uint64_t)t, o);
        } else if (band || g.touch) {
          float4* r = brec + kSdBandRec * (uint64_t)t;
          r[0] = make_float4(u2f((band ? 1u : 2u) | ((uint32_t)g.act.n << 8)), u2f((uint32_t)g.act.s0.ia),
                             u2f((uint32_t)g.act.s1.ia), u2f((uint32_t)g.act.s2.ia));
          r[1] = make_float4(u2f((uint32_t)g.act.s0.ib), u2f((uint32_t)g.act.s1.ib), u2f((uint32_t)g.act.s2.ib), 0.0f);
          r[2] = make_float4(g.last_v.x, g.last_v.y, g.last_v.z, 0.0f);
        } else {
          PoseGrad Gr;
          Gr.zero();
          store_grad(out + 14 * (uint64_t)t, Gr);
          brec[kSdBandRec * (uint64_t)t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
      }
      if (BWD && sub == 0 && need_epa) brec[kSdBandRec * (uint64_t)t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    bool enq = need_epa && sub == 0;
    uint64_t mask = __ballot(enq);
    if (mask) {
      uint32_t lane = threadIdx.x & 63;
      uint32_t leader = __ffsll((unsigned long long)mask) - 1;
      uint32_t qb = 0;
      if (lane == leader) qb = atomicAdd(&cnt->n_epa, (uint32_t)__popcll(mask));
      qb = __shfl(qb, leader);
      if (enq) {
        // entry of kSdEpaEntry float4: the enclosing simplex and the pair,
        // then the walk's depth bound and closest iterate (for an EPA
        // result the bound contradicts)
        uint64_t e = (uint64_t)kSdEpaEntry * (qb + __popcll(mask & ((1ull << lane) - 1ull)));
        epa_q[e] = make_float4(g.t0.x, g.t0.y, g.t0.z, u2f(t));
        epa_q[e + 1] = make_float4(g.t1.x, g.t1.y, g.t1.z, 0.0f);
        epa_q[e + 2] = make_float4(g.t2.x, g.t2.y, g.t2.z, 0.0f);
        epa_q[e + 3] = make_float4(g.t3.x, g.t3.y, g.t3.z, 0.0f);
        epa_q[e + 4] = make_float4(g.last_v.x, g.last_v.y, g.last_v.z, g.depth_bound);
        epa_q[e + 5] = make_float4(g.last_pb.x, g.last_pb.y, g.last_pb.z, u2f((uint32_t)g.act.n));
        epa_q[e + 6] = make_float4(u2f((uint32_t)g.act.s0.ia), u2f((uint32_t)g.act.s1.ia),
                                   u2f((uint32_t)g.act.s2.ia), 0.0f);
        epa_q[e + 7] = make_float4(u2f((uint32_t)g.act.s0.ib), u2f((uint32_t)g.act.s1.ib),
                                   u2f((uint32_t)g.act.s2.ib), 0.0f);
      }
    }
  }
}

// The Band and touch arms' VJP of the pairs k_sd_walk<G, true> recorded, one
// lane per pair (the same band_vjp, on the recorded simplex indices).
__global__ __launch_bounds__(256) void k_sd_band_vjp(uint32_t P, const uint32_t* __restrict__ P_dev,
                                                     const float4* __restrict__ brec,
                                                     const uint32_t* __restrict__ pairs,
                                                     const BodyRec* __restrict__ bodies,
                                                     const float4* __restrict__ hull, const float* __restrict__ gin,
                                                     float* __restrict__ out) {
  P = dev_count(P, P_dev);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < P; t += gridDim.x * blockDim.x) {
    const float4* r = brec + kSdBandRec * (uint64_t)t;
    const float4 r0 = r[0];
    const uint32_t h = f2u(r0.x);
    const uint32_t kind = h & 3u;
    if (kind == 0u) continue;
    const float4 r1 = r[1], r2 = r[2];
    Simp act;
    act.n = (int)(h >> 8);
    act.s0.ia = (int)f2u(r0.y);
    act.s1.ia = (int)f2u(r0.z);
    act.s2.ia = (int)f2u(r0.w);
    act.s0.ib = (int)f2u(r1.x);
    act.s1.ib = (int)f2u(r1.y);
    act.s2.ib = (int)f2u(r1.z);
    const Pair pr = make_pair(bodies[pairs[2 * (uint64_t)t]], bodies[pairs[2 * (uint64_t)t + 1]], hull);
    const V3 dir{r2.x, r2.y, r2.z};
    PoseGrad Gr;
    Gr.zero();
    band_vjp(pr, act, gin + 10 * (uint64_t)t, Gr, kind == 2u ? &dir : nullptr);
    store_grad(out + 14 * (uint64_t)t, Gr);
  }
}

// Split Minkowski vertex w into a support point of A's core (local xa) and
// one of B's (local xb) with (qrot(qa, xa) + pa) - (qrot(qb, xb) + pb) == w.
// Lane l tries A candidates l, l + 64, ...; B candidates are looped.
__device__ __forceinline__ bool wave_split(const Pair& pr, V3 w, V3& xa, V3& xb, int lane) {
  int na = core_set_size(pr.a), nb = core_set_size(pr.b);
  for (int ca = 0; ca < na; ca += 64) {
    int ia = ca + lane;
    bool va = ia < na;
    V3 la = va ? core_set_point(pr.a, ia) : splat(0.0f);
    V3 wa = qrot(pr.qa, la) + pr.pa;
    for (int ib = 0; ib < nb; ++ib) {
      V3 lb = core_set_point(pr.b, ib);
      V3 wb = qrot(pr.qb, lb) + pr.pb;
      V3 d = wa - wb;
      uint64_t hit = __ballot(va && d.x == w.x && d.y == w.y && d.z == w.z);
      if (hit) {
        int l = ctz64(hit);
        xa = V3{rdl_f(la.x, l), rdl_f(la.y, l), rdl_f(la.z, l)};
        xb = lb;
        return true;
      }
    }
  }
  return false;
}

// wave_split for the three vertices of a face at once, with the same result
// per vertex (first B candidate in index order, then the lowest A lane,
// A candidates in chunks of 64): B's world core points are computed once,
// one per lane, and read back by index (the same arithmetic as wave_split's
// per-candidate recomputation, so the same bits), instead of once per A
// chunk, B candidate and vertex.  B sets of more than 64 points take
// wave_split.
__device__ __forceinline__ bool wave_split3(const Pair& pr, const V3* w, V3* xa, V3* xb, int lane) {
  const int na = core_set_size(pr.a), nb = core_set_size(pr.b);
  if (nb > 64) {
    bool ok = wave_split(pr, w[0], xa[0], xb[0], lane);
    ok = wave_split(pr, w[1], xa[1], xb[1], lane) && ok;
    return wave_split(pr, w[2], xa[2], xb[2], lane) && ok;
  }
  const V3 lbl = lane < nb ? core_set_point(pr.b, lane) : splat(0.0f);
  const V3 wbl = qrot(pr.qb, lbl) + pr.pb;
  bool f0 = false, f1 = false, f2 = false;
  for (int ca = 0; ca < na && !(f0 && f1 && f2); ca += 64) {
    const int ia = ca + lane;
    const bool va = ia < na;
    const V3 la = va ? core_set_point(pr.a, ia) : splat(0.0f);
    const V3 wa = qrot(pr.qa, la) + pr.pa;
    for (int ib = 0; ib < nb && !(f0 && f1 && f2); ++ib) {
      const V3 d = wa - V3{rdl_f(wbl.x, ib), rdl_f(wbl.y, ib), rdl_f(wbl.z, ib)};
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const bool done = k == 0 ? f0 : (k == 1 ? f1 : f2);
        if (done) continue;
        const uint64_t hit = __ballot(va && d.x == w[k].x && d.y == w[k].y && d.z == w[k].z);
        if (hit) {
          const int l = ctz64(hit);
          xa[k] = V3{rdl_f(la.x, l), rdl_f(la.y, l), rdl_f(la.z, l)};
          xb[k] = V3{rdl_f(lbl.x, ib), rdl_f(lbl.y, ib), rdl_f(lbl.z, ib)};
          if (k == 0) f0 = true;
          else if (k == 1) f1 = true;
          else f2 = true;
        }
      }
    }
  }
  return f0 && f1 && f2;
}

// The EPA arm of the VJP for one queue entry `ent` of pair t, after EPA
// reported (n, depth, status) and, with has_face, the face (w0, w1, w2) of
// Minkowski vertices; the whole wavefront takes part (the vertex split),
// lane 0 stores.  A depth beyond the walk's own bound takes the Band arm at
// the walk's closest iterate instead, as the forward did.
__device__ __forceinline__ void sd_epa_vjp(const Pair& pr, const float4* ent, uint32_t t, V3 en, float edepth,
                                           int estatus, bool has_face, V3 w0, V3 w1, V3 w2,
                                           const float* __restrict__ gin, float* __restrict__ out,
                                           uint8_t* __restrict__ status, int lane) {
  uint32_t st = estatus == 0 ? SD_EPA : SD_EPA_UNCONVERGED;
  const float4 it_v = ent[4], it_pb = ent[5];
  PoseGrad Gr;
  Gr.zero();
  if (edepth > it_v.w + 1e-4f) {
    st = SD_EPA_CONTRADICTED;
    float4 ia = ent[6], ib = ent[7];
    Simp act;
    act.n = (int)f2u(it_pb.w);
    act.s0.ia = (int)f2u(ia.x); act.s1.ia = (int)f2u(ia.y); act.s2.ia = (int)f2u(ia.z);
    act.s0.ib = (int)f2u(ib.x); act.s1.ib = (int)f2u(ib.y); act.s2.ib = (int)f2u(ib.z);
    band_vjp(pr, act, gin + 10 * (uint64_t)t, Gr);
  } else if (has_face) {
    float fd_plane = dot(w0, en);
    if (!(fd_plane == edepth)) st |= SD_FLAG_EDGE;
    V3 xa[3], xb[3];
    const V3 wf[3] = {w0, w1, w2};
    bool ok = wave_split3(pr, wf, xa, xb, lane);
    if (!ok) st |= SD_FLAG_SPLIT;
    if (ok) epa_vjp(pr, en, xa, xb, gin + 10 * (uint64_t)t, Gr);
  }
  if (lane == 0) {
    store_grad(out + 14 * (uint64_t)t, Gr);
    if (status) status[t] = (uint8_t)st;
  }
}

// EPA for the queued pairs (one wavefront each) and the EPA arm of the
// outputs (BWD = false) or of the VJP (BWD = true).
template <bool BWD>
__global__ __launch_bounds__(64) void k_sd_epa(const uint32_t* __restrict__ n_queue, const float4* __restrict__ epa_q,
                                              const uint32_t* __restrict__ pairs, const BodyRec* __restrict__ bodies,
                                              const float4* __restrict__ hull, const float* __restrict__ gin,
                                              float* __restrict__ out, uint8_t* __restrict__ status) {
  __shared__ EpaLds S;
  const int lane = threadIdx.x & 63;
  const uint32_t nq = *n_queue;
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    const float4* ent = epa_q + (uint64_t)kSdEpaEntry * q;
    float4 a = ent[0], b = ent[1], c = ent[2], d = ent[3];
    uint32_t t = f2u(a.w);
    uint32_t i = pairs[2 * t], j = pairs[2 * t + 1];
    Pair pr = make_pair(bodies[i], bodies[j], hull);
    V3 t0{a.x, a.y, a.z}, t1{b.x, b.y, b.z}, t2{c.x, c.y, c.z}, t3{d.x, d.y, d.z};
    EpaResult e;
    V3 wa;
    if (SplitPair::fits(pr)) {
      SplitPair sp;
      sp.load(pr, lane);
      e = wave_epa(sp, t0, t1, t2, t3, S, lane);
      wa = sp.support_a(e.n, lane);
    } else {
      EpaPair ep;
      ep.pr = pr;
      ep.ha.load(pr.a, lane);
      ep.hb.load(pr.b, lane);
      SeqPair sq{&ep};
      e = wave_epa(sq, t0, t1, t2, t3, S, lane);
      wa = sq.support_a(e.n, lane);
    }
    if (!BWD) {
      float depth = (e.depth + pr.a.margin) + pr.b.margin;
      SdOut o;
      o.sd = -depth;
      o.n = e.n;
      o.pa = wa;
      o.pb = o.pa - e.n * depth;
      uint32_t st = e.status == 0 ? SD_EPA : SD_EPA_UNCONVERGED;
      // A core depth beyond the walk's own bound (plus EPA's exit tolerance)
      // comes from a polytope corrupted by rounding: the closest iterate of
      // the walk stands, as at the iteration cap.
      float4 it_v = ent[4], it_pb = ent[5];
      if (e.depth > it_v.w + 1e-4f) {
        st = SD_EPA_CONTRADICTED;
        o = band_outputs(pr, V3{it_v.x, it_v.y, it_v.z}, V3{it_pb.x, it_pb.y, it_pb.z});
      }
      if (lane == 0) {
        store_out(out + 10 * (uint64_t)t, o);
        if (status) status[t] = (uint8_t)st;
      }
    } else {
      V3 w0 = splat(0.0f), w1 = w0, w2 = w0;
      if (e.face >= 0) {
        uint32_t abc = S.abc[e.face];
        w0 = lv(S, abc & 0xffu);
        w1 = lv(S, (abc >> 8) & 0xffu);
        w2 = lv(S, (abc >> 16) & 0xffu);
      }
      sd_epa_vjp(pr, ent, t, e.n, e.depth, e.status, e.face >= 0, w0, w1, w2, gin, out, status, lane);
    }
    wave_sync();
  }
}

// The signed distance's EPA, four pairs per wavefront: the world pass's
// four-pair EPA (same algorithm, statement order and tie rules, same budget)
// on the signed-distance queue.  An entry outside the budget goes whole to
// defer_q (the same kSdEpaEntry layout), which k_sd_epa then runs from the
// start, so every result is the one-pair kernel's.  BWD = false writes the
// outputs; BWD = true writes, per queue entry, the reported normal, depth,
// status and face (rec, 4 float4) for k_sd_epa_vjp.
constexpr uint32_t kSdRecFace = 0x100u, kSdRecDeferred = 0x200u;
template <bool BWD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(BCH_EPA_QUAD_WAVES))) void k_sd_epa_quad(
    const uint32_t* __restrict__ n_queue, const float4* __restrict__ epa_q, const uint32_t* __restrict__ pairs,
    const BodyRec* __restrict__ bodies, const float4* __restrict__ hull, float* __restrict__ out,
    uint8_t* __restrict__ status, float4* __restrict__ rec, float4* __restrict__ defer_q, uint32_t* n_defer) {
  __shared__ EpaQuarter S[4];
  const int lane = threadIdx.x & 63;
  const int qi = lane >> 4;
  const int ql = lane & 15;
  const uint32_t nq = *n_queue;
  const WorldEpaSrc src{pairs, bodies, hull};
  for (uint32_t base = 4u * blockIdx.x; base < nq; base += 4u * gridDim.x) {
    const uint32_t q = base + (uint32_t)qi;
    const bool has = q < nq;
    const uint32_t qq = has ? q : base;  // an empty quarter reads the first entry, unused
    const float4* ent = epa_q + (uint64_t)kSdEpaEntry * qq;
    const float4 a = ent[0], b = ent[1], c = ent[2], d = ent[3];
    const uint32_t t = f2u(a.w);
    const Pair pr = src.pair(t);
    const bool fits = QuadShapes::fits(pr);
    QuadShapes qs;
    qs.load(pr, lane);
    EpaResult e;
    bool deferred;
    quad_epa(qs, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, V3{d.x, d.y, d.z}, S[qi], lane, qi,
             has && fits, e, deferred);
    deferred = has && (deferred || !fits);
    const V3 wa = qs.support_a(e.n, lane);
    if (ql == 0 && has) {
      if (deferred) {
        uint32_t slot = atomicAdd(n_defer, 1u);
        for (uint32_t k = 0; k < kSdEpaEntry; ++k) defer_q[(uint64_t)kSdEpaEntry * slot + k] = ent[k];
        if (BWD) rec[4 * (uint64_t)q + 1] = make_float4(0.0f, 0.0f, 0.0f, u2f(kSdRecDeferred));
      } else if (!BWD) {
        const float depth = (e.depth + pr.a.margin) + pr.b.margin;
        SdOut o;
        o.sd = -depth;
        o.n = e.n;
        o.pa = wa;
        o.pb = o.pa - e.n * depth;
        uint32_t st = e.status == 0 ? SD_EPA : SD_EPA_UNCONVERGED;
        // as in k_sd_epa: a core depth beyond the walk's own bound stands
        // down for the walk's closest iterate
        const float4 it_v = ent[4], it_pb = ent[5];
        if (e.depth > it_v.w + 1e-4f) {
          st = SD_EPA_CONTRADICTED;
          o = band_outputs(pr, V3{it_v.x, it_v.y, it_v.z}, V3{it_pb.x, it_pb.y, it_pb.z});
        }
        store_out(out + 10 * (uint64_t)t, o);
        if (status) status[t] = (uint8_t)st;
      } else {
        V3 w0 = splat(0.0f), w1 = w0, w2 = w0;
        const bool face = e.face >= 0;
        if (face) {
          uint32_t abc = S[qi].abc[e.face];
          w0 = qv(S[qi], abc & 0xffu);
          w1 = qv(S[qi], (abc >> 8) & 0xffu);
          w2 = qv(S[qi], (abc >> 16) & 0xffu);
        }
        float4* r = rec + 4 * (uint64_t)q;
        r[0] = make_float4(e.n.x, e.n.y, e.n.z, e.depth);
        r[1] = make_float4(w0.x, w0.y, w0.z, u2f((uint32_t)(e.status & 0xff) | (face ? kSdRecFace : 0u)));
        r[2] = make_float4(w1.x, w1.y, w1.z, 0.0f);
        r[3] = make_float4(w2.x, w2.y, w2.z, 0.0f);
      }
    }
    wave_sync();
  }
}

// The EPA arm of the VJP for the entries k_sd_epa_quad<true> finished, one
// wavefront per entry (the face's vertex split runs over its 64 lanes).
__global__ __launch_bounds__(64) void k_sd_epa_vjp(const uint32_t* __restrict__ n_queue,
                                                   const float4* __restrict__ epa_q, const float4* __restrict__ rec,
                                                   const uint32_t* __restrict__ pairs,
                                                   const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                                                   const float* __restrict__ gin, float* __restrict__ out,
                                                   uint8_t* __restrict__ status) {
  const int lane = threadIdx.x & 63;
  const uint32_t nq = *n_queue;
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    const float4* r = rec + 4 * (uint64_t)q;
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
    const uint32_t flags = f2u(r1.w);
    if (flags & kSdRecDeferred) continue;  // wavefront-uniform
    const float4* ent = epa_q + (uint64_t)kSdEpaEntry * q;
    const uint32_t t = f2u(ent[0].w);
    const Pair pr = make_pair(bodies[pairs[2 * t]], bodies[pairs[2 * t + 1]], hull);
    sd_epa_vjp(pr, ent, t, V3{r0.x, r0.y, r0.z}, r0.w, (int)(flags & 0xffu), (flags & kSdRecFace) != 0,
               V3{r1.x, r1.y, r1.z}, V3{r2.x, r2.y, r2.z}, V3{r3.x, r3.y, r3.z}, gin, out, status, lane);
    wave_sync();
  }
}

}  // namespace bch
