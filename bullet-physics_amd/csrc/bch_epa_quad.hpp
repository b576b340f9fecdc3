// bch_epa_quad.hpp -- EPA for four pairs per wavefront, one per 16-lane
// quarter (one DPP row).
//
// Same algorithm, statement order and tie rules as k_epa_wave and the mesh
// batch's two-pair kernel (narrowphase/epa.rs:127-398), on a smaller budget per pair: faces l and
// l + 16 of a quarter's polytope in lane l (32 faces), at most 5 visible
// faces per expansion (15 edge events, one per lane), at most 64 vertices,
// the shapes' supports on 8 lanes each with up to 8 hull vertices per lane
// in registers.  Over the EPA pairs of a 200k-body mixed scene 97.8 % end
// with at most 32 faces and never see more than 5 visible faces (oracle
// statistics).  Every wavefront-wide instruction then serves four pairs
// instead of two, the cross-lane steps of a quarter stay inside its DPP row,
// and the supports of A and B are exchanged by one row mirror.
//
// A pair outside that budget (a hull of more than 64 vertices, more than 32
// faces, more than 5 visible faces, more than 64 vertices, a directed edge
// shared by two visible faces) is not finished here: its entry goes to a
// second queue, which k_epa_wave runs from the start, so every result is the
// one-pair kernel's.  (Routing these through the two-pair kernel first made
// the EPA stage 0.1 ms longer on the 1M-body scene: they are the largest
// polytopes, and two per wavefront ran them more slowly than one.)
//
// As in the dual kernel, every cross-lane operation runs with all 64 lanes
// active; a quarter's state lives in quarter-uniform registers and its
// updates are selected, not branched around.
#pragma once
#include "bch_epa_dual.hpp"

namespace bch {

constexpr int kQuadF = 32;  // face slots per quarter (two per lane)
constexpr int kQuadV = 64;  // vertex slots per quarter

struct EpaQuarter {
  float4 v[kQuadV];
  float nx[kQuadF], ny[kQuadF], nz[kQuadF], dist[kQuadF], pd[kQuadF];
  uint32_t abc[kQuadF];
  // expansion scratch: the visible faces' vertex triples by rank, and the
  // horizon list of the swap-remove replay (list[p] = event at position p,
  // pos[e] = position of event e)
  uint32_t vis_abc[8];
  int list[16], pos[16];
};

// --------------------------------------------------------- per-quarter ops --
// Ballot of the own quarter (bit l = lane 16 qi + l).
__device__ __forceinline__ uint32_t qballot(bool p, int qi) {
  return (uint32_t)(__ballot(p) >> (16 * qi)) & 0xffffu;
}
__device__ __forceinline__ int qsel4(int r0, int r1, int r2, int r3, int qi) {
  return qi < 2 ? (qi == 0 ? r0 : r1) : (qi == 2 ? r2 : r3);
}
// x at lane k of the own quarter, k uniform over the wavefront.
__device__ __forceinline__ uint32_t qget_u(uint32_t x, int k, int qi) {
  return (uint32_t)qsel4(rdl_i((int)x, k), rdl_i((int)x, 16 + k), rdl_i((int)x, 32 + k), rdl_i((int)x, 48 + k), qi);
}
__device__ __forceinline__ int qget_i(int x, int k, int qi) {
  return qsel4(rdl_i(x, k), rdl_i(x, 16 + k), rdl_i(x, 32 + k), rdl_i(x, 48 + k), qi);
}
// the largest of the four quarters' values of a quarter-uniform int
__device__ __forceinline__ int qmax_i(int x) {
  int a = rdl_i(x, 0), b = rdl_i(x, 16), c = rdl_i(x, 32), d = rdl_i(x, 48);
  int m0 = a > b ? a : b, m1 = c > d ? c : d;
  return m0 > m1 ? m0 : m1;
}
// Row min / max of values without NaN: the four DPP row steps leave it in
// every lane of the row, i.e. of the quarter.
template <bool MAX>
__device__ __forceinline__ float quarter_ext_f(float v) {
  float o;
  o = dpp_f<0xB1>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = dpp_f<0x4E>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = dpp_f<0x141>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = dpp_f<0x140>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  return v;
}

// -------------------------------------------------------------- supports --
// The two shapes of a quarter's pair: lanes 0-7 of the quarter hold shape A,
// lanes 8-15 shape B (the two halves of the DPP row).  A hull of at most 64
// vertices sits in registers: vertex sub + 8 k in slot k of lane sub; a
// box, capsule or sphere keeps its parameters in slot 0 instead.  The
// support's reduction carries the winning vertex along, so no hull is read
// from memory after the load.
struct QuadShapes {
  int tn;  // type | vertex count << 8
  V3 p;
  Q4 q;
  float x[8], y[8], z[8];

  __device__ __forceinline__ static bool fits(const Pair& pr) {
    return (pr.a.type != HULL || pr.a.nhull <= 64) && (pr.b.type != HULL || pr.b.nhull <= 64);
  }
  __device__ __forceinline__ int type() const { return tn & 0xff; }
  __device__ __forceinline__ int count() const { return tn >> 8; }
  __device__ __forceinline__ void load(const Pair& pr, int lane) {
    const bool sb = (lane & 8) != 0;
    const int t = sb ? pr.b.type : pr.a.type;
    const int n = t == HULL ? (sb ? pr.b.nhull : pr.a.nhull) : 0;
    tn = t | (n << 8);
    p = sel(sb, pr.pb, pr.pa);
    q = Q4{sb ? pr.qb.x : pr.qa.x, sb ? pr.qb.y : pr.qa.y, sb ? pr.qb.z : pr.qa.z, sb ? pr.qb.w : pr.qa.w};
    const float4* h = sb ? pr.b.hull : pr.a.hull;
    const int sub = lane & 7;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (sub + 8 * k < n) u = h[sub + 8 * k];
      x[k] = u.x;
      y[k] = u.y;
      z[k] = u.z;
    }
    if (t != HULL) {
      x[0] = sb ? pr.b.a0 : pr.a.a0;
      y[0] = sb ? pr.b.a1 : pr.a.a1;
      z[0] = sb ? pr.b.a2 : pr.a.a2;
    }
  }
  // convex_hull.rs:55-66 over the half-row: per-lane first maximiser (its
  // vertices in index order), then largest value / lowest index, with the
  // vertex carried through the reduction.
  __device__ __forceinline__ V3 hull_support(V3 d, int lane) const {
    const int sub = lane & 7;
    const int n = count();
    float bv = -INFINITY;
    int bi = -1;
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      float dd = dot(V3{x[k], y[k], z[k]}, d);
      if (sub + 8 * k < n && dd > bv) {
        bv = dd;
        bi = sub + 8 * k;
        bx = x[k];
        by = y[k];
        bz = z[k];
      }
    }
    amax_step<0xB1>(bv, bi, bx, by, bz);
    amax_step<0x4E>(bv, bi, bx, by, bz);
    amax_step<0x141>(bv, bi, bx, by, bz);
    return bi >= 0 ? V3{bx, by, bz} : splat(0.0f);
  }
  template <int CTRL>
  __device__ __forceinline__ static void amax_step(float& v, int& i, float& x0, float& y0, float& z0) {
    float ov = dpp_f<CTRL>(v), ox = dpp_f<CTRL>(x0), oy = dpp_f<CTRL>(y0), oz = dpp_f<CTRL>(z0);
    int oi = dpp_i<CTRL>(i);
    if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) {
      v = ov;
      i = oi;
      x0 = ox;
      y0 = oy;
      z0 = oz;
    }
  }
  // support_world / core_support_world of this lane's shape at d
  __device__ __forceinline__ V3 world(V3 d, int lane, bool core) const {
    ShapeRef s;
    s.type = type();
    s.a0 = x[0];
    s.a1 = y[0];
    s.a2 = z[0];
    s.a3 = 0.0f;
    s.hull = nullptr;
    s.nhull = 0;
    s.margin = 0.0f;
    V3 ld = qrot(qconj(q), d);
    V3 l;
    if (s.type == HULL) l = hull_support(ld, lane);
    else l = core ? core_support_local(s, ld) : support_local(s, ld);
    return qrot(q, l) + p;
  }
  // core supportA(d) - core supportB(-d) of the own quarter's pair, d
  // uniform within each quarter.  Each half-row holds its shape's support
  // in every lane, so the row mirror hands every lane the other half's.
  __device__ __forceinline__ V3 mink(V3 d, int lane) const {
    const bool sb = (lane & 8) != 0;
    V3 w = world(sel(sb, -d, d), lane, true);
    V3 o{dpp_f<0x140>(w.x), dpp_f<0x140>(w.y), dpp_f<0x140>(w.z)};
    return sel(sb, o - w, w - o);
  }
  // full supportA(d) of the own quarter's pair (uniform within each quarter)
  __device__ __forceinline__ V3 support_a(V3 d, int lane) const {
    const bool sb = (lane & 8) != 0;
    V3 w = world(d, lane, false);
    V3 o{dpp_f<0x140>(w.x), dpp_f<0x140>(w.y), dpp_f<0x140>(w.z)};
    return sel(sb, o, w);
  }
};

__device__ __forceinline__ V3 qv(const EpaQuarter& S, int i) {
  float4 v = S.v[i];
  return V3{v.x, v.y, v.z};
}

// epa.rs:212-224 over the (<= 32) faces of the quarter, faces ql and ql + 16
// in lane ql: first face with the smallest |distance|, NaN distances
// skipped, a NaN first face kept.
template <bool TWO>
__device__ __forceinline__ int quad_best_t(const EpaQuarter& S, int nfaces, int ql, int qi) {
  const bool nan0 = is_nan(S.dist[0]);
  float v0 = INFINITY, v1 = INFINITY;
  if (ql < nfaces) {
    float d = fabsf(S.dist[ql]);
    v0 = is_nan(d) ? INFINITY : d;
  }
  if constexpr (TWO) {
    if (ql + 16 < nfaces) {
      float d = fabsf(S.dist[ql + 16]);
      v1 = is_nan(d) ? INFINITY : d;
    }
  }
  float m = quarter_ext_f<false>(v1 < v0 ? v1 : v0);
  uint32_t b0 = qballot(v0 == m, qi);
  uint32_t b1 = TWO ? qballot(v1 == m, qi) : 0u;
  if (nan0 || !(m < INFINITY)) return 0;
  return b0 ? ctz64((uint64_t)b0) : (b1 ? 16 + ctz64((uint64_t)b1) : 0);
}
// The second face slot is only scanned while a quarter holds more than 16 faces.
__device__ __forceinline__ int quad_best(const EpaQuarter& S, int nfaces, int ql, int qi) {
  return __ballot(nfaces > 16) ? quad_best_t<true>(S, nfaces, ql, qi) : quad_best_t<false>(S, nfaces, ql, qi);
}

// The fan face of horizon edge (u, w) and the new vertex (epa.rs:372-386).
__device__ __forceinline__ void quad_fan_face(EpaQuarter& S, int d, uint32_t u, uint32_t w, V3 vu, V3 vw, V3 np,
                                              int nidx) {
  V3 cr = cross(vw - vu, np - vu);
  float l = len(cr);
  V3 n = cr / l;
  S.nx[d] = n.x;
  S.ny[d] = n.y;
  S.nz[d] = n.z;
  S.dist[d] = face_dist(n, vu, vw, np);
  S.pd[d] = dot(n, vu);
  S.abc[d] = u | (w << 8) | ((uint32_t)nidx << 16);
}

// One face slot of a lane: whether face f of the quarter's store is live
// and visible from the new point.  The face itself stays in LDS until the
// commit moves it (fewer registers live across the expansion).
struct QuadFace {
  bool live, vis;
  __device__ __forceinline__ void load(const EpaQuarter& S, int f, bool l, V3 np) {
    live = l;
    vis = false;
    if (l) vis = dot(V3{S.nx[f], S.ny[f], S.nz[f]}, np) - S.pd[f] > 1e-5f;
  }
};
// A face record moved by the commit's compaction.
struct QuadFaceRec {
  float nx, ny, nz, d, pd;
  uint32_t abc;
  __device__ __forceinline__ void read(const EpaQuarter& S, int f) {
    nx = S.nx[f];
    ny = S.ny[f];
    nz = S.nz[f];
    d = S.dist[f];
    pd = S.pd[f];
    abc = S.abc[f];
  }
  __device__ __forceinline__ void store(EpaQuarter& S, int f) const {
    S.nx[f] = nx;
    S.ny[f] = ny;
    S.nz[f] = nz;
    S.dist[f] = d;
    S.pd[f] = pd;
    S.abc[f] = abc;
  }
};

// epa.rs:309-388 for the quarters with `act` set, faces ql and ql + 16 in
// lane ql.  Per-quarter result: 1 done (or nothing visible), 0 degenerate
// horizon (polytope untouched), -2 the pair leaves the budget (polytope
// untouched).
template <bool TWO>
__device__ __forceinline__ int quad_expand_t(EpaQuarter& S, int& nfaces, V3 np, int nidx, int ql, int qi, bool act) {
  QuadFace f0, f1;
  f0.load(S, ql, act && ql < nfaces, np);
  f1.load(S, ql + 16, TWO && act && ql + 16 < nfaces, np);
  const uint32_t vm0 = qballot(f0.vis, qi), vm1 = TWO ? qballot(f1.vis, qi) : 0u;
  int res = 1;
  bool go = act && (vm0 | vm1) != 0u;
  const int nv0 = __popc(vm0);
  const int nv = nv0 + __popc(vm1);
  if (go && nv > 5) {
    res = -2;
    go = false;
  }
  if (!__ballot(go)) return res;
  // Edge events: lane l < 3 nv holds edge l % 3 of the visible face of rank
  // l / 3, in the reference's visiting order (face index order).
  const int nev = go ? 3 * nv : 0;
  const int rank = ql / 3;
  const int e = ql - 3 * rank;
  const uint32_t below = (uint32_t)lanes_below(ql);
  if (go && f0.vis) S.vis_abc[__popc(vm0 & below)] = S.abc[ql];
  if (go && f1.vis) S.vis_abc[nv0 + __popc(vm1 & below)] = S.abc[ql + 16];
  wave_sync();
  const uint32_t eabc = (go && ql < nev) ? S.vis_abc[rank] : 0u;
  const uint32_t fa = eabc & 0xffu, fb = (eabc >> 8) & 0xffu, fc = (eabc >> 16) & 0xffu;
  const uint32_t u = e == 0 ? fa : (e == 1 ? fb : fc);
  const uint32_t w = e == 0 ? fb : (e == 1 ? fc : fa);
  const uint32_t edge = u | (w << 8), rev = w | (u << 8);
  const bool is_ev = go && ql < nev;
  // The earlier event that is this edge's reverse (it is then still on the
  // list, so this event cancels it); a repeated directed edge leaves the
  // budget.
  int match = -1;
  bool dup = false;
  const int nevmax = qmax_i(nev);
  for (int k = 0; k < nevmax; ++k) {
    uint32_t ek = qget_u(edge, k, qi);
    bool kin = k < nev;
    if (kin && ek == rev && k < ql) match = k;
    if (kin && ek == edge && k != ql) dup = true;
  }
  if (go && qballot(is_ev && dup, qi)) {
    res = -2;
    go = false;
  }
  if (!__ballot(go)) return res;
  // Replay of the pushes and swap-removes on the quarter's LDS lists
  // (list[p] = event at position p, pos[v] = position of event v): every
  // lane of the quarter reads the same address, one lane writes.
  const uint32_t cancel = qballot(is_ev && match >= 0, qi);
  int ne = 0;
  for (int l = 0; l < nevmax; ++l) {
    const bool lin = go && l < nev;
    const bool c = ((cancel >> l) & 1u) != 0u;
    const int ml = qget_i(match, l, qi);
    if (lin && c) {
      const int p = S.pos[ml];
      const int last = S.list[ne - 1];
      if (ql == 0) {
        S.list[p] = last;
        S.pos[last] = p;
      }
      --ne;
    } else if (lin) {
      if (ql == 0) {
        S.list[ne] = l;
        S.pos[l] = ne;
      }
      ++ne;
    }
  }
  wave_sync();
  // Horizon entry k (lane k of the quarter): the edge of the event at list
  // position k.
  const bool is_h = go && ql < ne;
  const int list = is_h ? S.list[ql] : 0;
  const uint32_t hedge = (uint32_t)__shfl((int)edge, 16 * qi + list, 64);
  const uint32_t hu = hedge & 0xffu, hw = (hedge >> 8) & 0xffu;
  V3 vu = splat(0.0f), vw = splat(0.0f);
  bool bad = false;
  if (is_h) {
    vu = qv(S, hu);
    vw = qv(S, hw);
    bad = len(cross(vw - vu, np - vu)) < 1e-12f;
  }
  if (go && qballot(bad, qi)) {
    res = 0;
    go = false;
  }
  // Commit: stable compaction of the faces that stay, then the fan.  A
  // polytope that would pass the store leaves the budget instead.
  const bool k0 = go && f0.live && !f0.vis, k1 = TWO && go && f1.live && !f1.vis;
  const uint32_t km0 = qballot(k0, qi), km1 = TWO ? qballot(k1, qi) : 0u;
  const int kept0 = __popc(km0);
  const int kept = kept0 + __popc(km1);
  if (go && kept + ne > kQuadF) {
    res = -2;
    go = false;
  }
  QuadFaceRec r0, r1;
  if (go && k0) r0.read(S, ql);
  if (TWO && go && k1) r1.read(S, ql + 16);
  wave_sync();
  if (go && k0) r0.store(S, __popc(km0 & below));
  if (TWO && go && k1) r1.store(S, kept0 + __popc(km1 & below));
  wave_sync();
  if (go && is_h) quad_fan_face(S, kept + ql, hu, hw, vu, vw, np, nidx);
  if (go) nfaces = kept + ne;
  wave_sync();
  return res;
}
// The second face slot is only visited while a quarter holds more than 16 faces.
__device__ __forceinline__ int quad_expand(EpaQuarter& S, int& nfaces, V3 np, int nidx, int ql, int qi, bool act) {
  return __ballot(act && nfaces > 16) ? quad_expand_t<true>(S, nfaces, np, nidx, ql, qi, act)
                                      : quad_expand_t<false>(S, nfaces, np, nidx, ql, qi, act);
}

// epa.rs:127-205 for the pair of each quarter.  `act` false: the quarter has
// no pair.  Returns per quarter: the result in R, `deferred` when the pair
// left the budget (R is then meaningless).
__device__ __forceinline__ void quad_epa(const QuadShapes& qs, V3 t0, V3 t1, V3 t2, V3 t3, EpaQuarter& S, int lane,
                                         int qi, bool act, EpaResult& R, bool& deferred) {
  const int ql = lane & 15;
  deferred = false;
  R.status = 1;
  R.n = splat(0.0f);
  R.depth = 0.0f;
  R.face = -1;
  if (act && ql < 4) {
    V3 w = sel(ql == 0, t0, sel(ql == 1, t1, sel(ql == 2, t2, t3)));
    S.v[ql] = make_float4(w.x, w.y, w.z, 0.0f);
  }
  // initial faces (epa.rs:226-257): lane k builds quad k; survivors keep
  // quad order.  Quads (a, b, c, opposite): (0,1,2,3) (0,3,1,2) (0,2,3,1)
  // (1,3,2,0).
  bool valid = false;
  V3 qn = splat(0.0f);
  float qd = 0.0f, qpd = 0.0f;
  uint32_t qabc = 0;
  if (act && ql < 4) {
    int a = ql == 3 ? 1 : 0;
    int b = ql == 0 ? 1 : (ql == 2 ? 2 : 3);
    int c = ql == 0 ? 2 : (ql == 1 ? 1 : (ql == 2 ? 3 : 2));
    int opp = 3 - ql;
    V3 sa = sel(a == 0, t0, t1);
    V3 sb = sel(b == 1, t1, sel(b == 2, t2, t3));
    V3 sc = sel(c == 1, t1, sel(c == 2, t2, t3));
    V3 so = sel(opp == 0, t0, sel(opp == 1, t1, sel(opp == 2, t2, t3)));
    V3 n = cross(sb - sa, sc - sa);
    if (dot(n, so - sa) > 0.0f) {
      n = -n;
      V3 t = sb;
      sb = sc;
      sc = t;
      int ti = b;
      b = c;
      c = ti;
    }
    float l = len(n);
    valid = !(l < 1e-12f);
    if (valid) {
      n = n / l;
      qn = n;
      qd = face_dist(n, sa, sb, sc);
      qpd = dot(n, sa);
      qabc = (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16);
    }
  }
  const uint32_t vm = qballot(valid, qi);
  int nfaces = __popc(vm);
  if (valid) {
    int d = __popc(vm & (uint32_t)lanes_below(ql));
    S.nx[d] = qn.x;
    S.ny[d] = qn.y;
    S.nz[d] = qn.z;
    S.dist[d] = qd;
    S.pd[d] = qpd;
    S.abc[d] = qabc;
  }
  wave_sync();
  bool alive = act && nfaces > 0;  // the quarter still has a pair in work
  int nverts = 4;
  int status = 2;
  bool stop = false;
  bool seeding = alive;
  for (int k = 0; k < 6; ++k) {  // seeding; a degenerate expansion only ends the seeding
    V3 d = k == 0 ? V3{1.0f, 0.0f, 0.0f} : k == 1 ? V3{-1.0f, 0.0f, 0.0f} : k == 2 ? V3{0.0f, 1.0f, 0.0f}
         : k == 3 ? V3{0.0f, -1.0f, 0.0f} : k == 4 ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 0.0f, -1.0f};
    V3 p = qs.mink(d, lane);
    int idx = nverts;
    if (seeding) {
      if (ql == 0) S.v[idx] = make_float4(p.x, p.y, p.z, 0.0f);
      ++nverts;
    }
    wave_sync();
    int e = quad_expand(S, nfaces, p, idx, ql, qi, seeding);
    if (seeding && e == -2) {
      deferred = true;
      alive = false;
      seeding = false;
    } else if (seeding && e <= 0) {
      status = e == 0 ? 1 : 4;
      stop = e < 0;
      seeding = false;
    }
  }
  bool run = alive && !stop;
  bool conv = false;
  for (int it = 0; it < 255; ++it) {
    if (!__ballot(run)) break;
    int bi = quad_best(S, nfaces, ql, qi);
    V3 bn{S.nx[bi], S.ny[bi], S.nz[bi]};
    float bd = S.dist[bi];
    V3 np = qs.mink(bn, lane);
    float wdist = dot(bn, np) - bd;
    if (run && wdist <= 1e-4f) {
      R.status = 0;
      R.n = bn;
      R.depth = bd;
      R.face = bi;
      conv = true;
      run = false;
    }
    // the reference's vertex limit is 128 (status 3); a quarter holds 64
    // vertices, so a pair that needs more is deferred instead
    if (run && nverts >= kQuadV) {
      deferred = true;
      alive = false;
      run = false;
    }
    int nidx = nverts;
    if (run) {
      if (ql == 0) S.v[nidx] = make_float4(np.x, np.y, np.z, 0.0f);
      ++nverts;
    }
    wave_sync();
    int e = quad_expand(S, nfaces, np, nidx, ql, qi, run);
    if (run && e == -2) {
      deferred = true;
      alive = false;
      run = false;
    } else if (run && e <= 0) {
      status = e == 0 ? 1 : 4;
      run = false;
    }
  }
  // the closest face of an unconverged polytope
  int bi = quad_best(S, nfaces, ql, qi);
  if (alive && !conv) {
    R.status = status;
    R.n = V3{S.nx[bi], S.ny[bi], S.nz[bi]};
    R.depth = S.dist[bi];
    R.face = bi;
  }
}

// Grid-stride over an EPA queue, four entries per wavefront (Src as in
// epa_queue_run).  Outputs as epa_queue_run; an entry that leaves the
// budget is appended to defer_q (count in *n_defer) for the two-pair kernel.
template <class Src>
__device__ __forceinline__ void epa_quad_run(const uint32_t* __restrict__ n_queue, const float4* __restrict__ epa_q,
                                             const Src& src, uint32_t* __restrict__ ccount,
                                             float4* __restrict__ cslots, uint32_t stride,
                                             uint32_t* __restrict__ route, float4* __restrict__ defer_q,
                                             uint32_t* n_defer, Counters* cnt, EpaQuarter* S) {
  const int lane = threadIdx.x & 63;
  const int qi = lane >> 4;
  const int ql = lane & 15;
  const uint32_t nq = *n_queue;
  for (uint32_t base = 4u * blockIdx.x; base < nq; base += 4u * gridDim.x) {
    const uint32_t q = base + (uint32_t)qi;
    const bool has = q < nq;
    const uint32_t qq = has ? q : base;  // an empty quarter reads the first entry, unused
    float4 a = epa_q[4 * (uint64_t)qq], b = epa_q[4 * (uint64_t)qq + 1], c = epa_q[4 * (uint64_t)qq + 2],
           d = epa_q[4 * (uint64_t)qq + 3];
    const uint32_t p = f2u(a.w);
    Pair pr = src.pair(p);
    const bool fits = QuadShapes::fits(pr);
    QuadShapes qs;
    qs.load(pr, lane);
    V3 t0{a.x, a.y, a.z}, t1{b.x, b.y, b.z}, t2{c.x, c.y, c.z}, t3{d.x, d.y, d.z};
    EpaResult e;
    bool deferred;
    quad_epa(qs, t0, t1, t2, t3, S[qi], lane, qi, has && fits, e, deferred);
    deferred = has && (deferred || !fits);
    const bool hit = has && !deferred && !(e.depth <= 0.0f);
    V3 point_a = qs.support_a(e.n, lane);
    float depth = 0.0f;
    if (hit) depth = (e.depth + pr.a.margin) + pr.b.margin;
    if (ql == 0 && has) {
      if (deferred) {
        // the entry is read again rather than kept in registers through the
        // EPA (epa_q is not written while this kernel runs)
        uint32_t slot = atomicAdd(n_defer, 1u);
        for (int k = 0; k < 4; ++k) defer_q[4 * (uint64_t)slot + k] = epa_q[4 * (uint64_t)qq + k];
      } else {
        if (e.status == 4) atomicOr(&cnt->errors, 2u);
        ccount[p] = hit ? 1u : 0u;
        if (route) route[p] |= 8u;
        if (hit) {
          V3 pos = point_a - e.n * depth;
          cslots[(uint64_t)p * stride] = make_float4(pos.x, pos.y, pos.z, depth);
          cslots[(uint64_t)p * stride + 1] = make_float4(e.n.x, e.n.y, e.n.z, 0.0f);
        }
      }
    }
    wave_sync();
  }
}

// Waves per SIMD the four-pair EPA is compiled for (its register budget).
#ifndef BCH_EPA_QUAD_WAVES
#define BCH_EPA_QUAD_WAVES 4
#endif

// The world pass's EPA queue, four entries per wavefront; the entries that
// leave the budget go to defer_q for k_epa_wave.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(BCH_EPA_QUAD_WAVES))) void k_epa_quad(const uint32_t* __restrict__ n_queue,
                                                 const float4* __restrict__ epa_q, const uint32_t* __restrict__ pairs,
                                                 const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                                                 uint32_t* __restrict__ ccount, float4* __restrict__ cslots,
                                                 uint32_t* __restrict__ route, float4* __restrict__ defer_q,
                                                 uint32_t* n_defer, Counters* cnt) {
  __shared__ EpaQuarter S[4];
  epa_quad_run(n_queue, epa_q, WorldEpaSrc{pairs, bodies, hull}, ccount, cslots, 8u, route, defer_q, n_defer, cnt, S);
}

}  // namespace bch
