// bch_epa_dual.hpp -- EPA for two pairs per wavefront, one per 32-lane half
// (the mesh batch's sub-pairs; the world pass runs four pairs per wavefront,
// bch_epa_quad.hpp, whose helpers build on these).
//
// The polytopes of the world pass are small: the closest-face search and
// the expansions see at most 64 faces for all but a few pairs in 10^5, and
// the visible set rarely exceeds a handful of faces.  One wavefront per pair leaves most
// lanes idle while every instruction still costs the whole wavefront, so
// here each half runs its own pair with the same algorithm, statement order
// and tie rules as k_epa_wave (narrowphase/epa.rs:127-398): faces l and
// l + 32 in lane l, the reference's swap-remove horizon replayed per half,
// the shapes' supports on 16 lanes each.
//
// A pair that leaves that fast path (a hull of more than 64 vertices, more
// than 64 faces, more than 10 visible faces, a directed edge shared by two
// visible faces) is not finished here: its queue entry goes to a second
// queue, which k_epa_wave runs from the start afterwards, so every result is
// the one-pair kernel's.
//
// Every cross-lane operation below runs with all 64 lanes active; a half's
// state (face count, done, deferred) lives in registers that are uniform
// within the half, and its updates are selected, not branched around.
#pragma once
#include "bch_epa_wave.hpp"

namespace bch {

constexpr int kDualF = 64;  // face slots per half (two per lane)

struct EpaHalf {
  float4 v[BCH_WEPA_V];
  float nx[kDualF], ny[kDualF], nz[kDualF], dist[kDualF], pd[kDualF];
  uint32_t abc[kDualF];
  // expansion scratch: the visible faces' vertex triples by rank, and the
  // horizon list of the swap-remove replay (list[p] = event at position p,
  // pos[e] = position of event e)
  uint32_t vis_abc[16];
  int list[32], pos[32];
};

// ---------------------------------------------------------- per-half ops --
// Ballot of the own half (bit l = lane 32 h + l).
__device__ __forceinline__ uint32_t hballot(bool p, int h) {
  uint64_t b = __ballot(p);
  return h ? (uint32_t)(b >> 32) : (uint32_t)b;
}
// x at lane k of the own half, k uniform over the wavefront.
__device__ __forceinline__ uint32_t hget_u(uint32_t x, int k, int h) {
  uint32_t r0 = rdl_u(x, k), r1 = rdl_u(x, 32 + k);
  return h ? r1 : r0;
}
__device__ __forceinline__ int hget_i(int x, int k, int h) {
  int r0 = rdl_i(x, k), r1 = rdl_i(x, 32 + k);
  return h ? r1 : r0;
}
// x at lane idx of the own half, idx uniform within each half.
__device__ __forceinline__ int hgetv_i(int x, int idx, int h) {
  int i0 = rdl_i(idx, 0) & 31, i1 = rdl_i(idx, 32) & 31;
  int r0 = rdl_i(x, i0), r1 = rdl_i(x, 32 + i1);
  return h ? r1 : r0;
}
// the larger of the two halves' values of a half-uniform int
__device__ __forceinline__ int hmax_i(int x) {
  int a = rdl_i(x, 0), b = rdl_i(x, 32);
  return a > b ? a : b;
}
// Half min / max of values without NaN (uniform within each half): rows
// by DPP, then the two rows of each half.
template <bool MAX>
__device__ __forceinline__ float half_ext_f(float v, int h) {
  float o;
  o = dpp_f<0xB1>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = dpp_f<0x4E>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = dpp_f<0x141>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  o = dpp_f<0x140>(v);
  v = MAX ? (o > v ? o : v) : (o < v ? o : v);
  float a = rdl_f(v, 0), b = rdl_f(v, 16), c = rdl_f(v, 32), d = rdl_f(v, 48);
  float m0 = MAX ? (b > a ? b : a) : (b < a ? b : a);
  float m1 = MAX ? (d > c ? d : c) : (d < c ? d : c);
  return h ? m1 : m0;
}

// ------------------------------------------------------------- supports --
// The two shapes of a half's pair: lanes 0-15 of the half hold shape A,
// lanes 16-31 shape B (one DPP row each).  A hull of at most 64 vertices
// sits in registers: vertex sub + 16 k in slot k of lane sub.
struct DualShapes {
  int type;
  float a0, a1, a2;
  V3 p;
  Q4 q;
  const float4* h;
  int n;
  float x[4], y[4], z[4];

  __device__ __forceinline__ static bool fits(const Pair& pr) {
    return (pr.a.type != HULL || pr.a.nhull <= 64) && (pr.b.type != HULL || pr.b.nhull <= 64);
  }
  __device__ __forceinline__ void load(const Pair& pr, int lane) {
    const bool sb = (lane & 16) != 0;
    type = sb ? pr.b.type : pr.a.type;
    a0 = sb ? pr.b.a0 : pr.a.a0;
    a1 = sb ? pr.b.a1 : pr.a.a1;
    a2 = sb ? pr.b.a2 : pr.a.a2;
    p = sel(sb, pr.pb, pr.pa);
    q = Q4{sb ? pr.qb.x : pr.qa.x, sb ? pr.qb.y : pr.qa.y, sb ? pr.qb.z : pr.qa.z, sb ? pr.qb.w : pr.qa.w};
    h = sb ? pr.b.hull : pr.a.hull;
    n = type == HULL ? (sb ? pr.b.nhull : pr.a.nhull) : 0;
    const int sub = lane & 15;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (sub + 16 * k < n) u = h[sub + 16 * k];
      x[k] = u.x;
      y[k] = u.y;
      z[k] = u.z;
    }
  }
  // convex_hull.rs:55-66 over the row: per-lane first maximiser (its
  // vertices in index order), then largest value / lowest index.
  __device__ __forceinline__ V3 hull_support(V3 d, int lane) const {
    const int sub = lane & 15;
    float bv = -INFINITY;
    int bi = -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float dd = dot(V3{x[k], y[k], z[k]}, d);
      if (sub + 16 * k < n && dd > bv) {
        bv = dd;
        bi = sub + 16 * k;
      }
    }
    row_reduce<true, 16>(bv, bi);
    return bi >= 0 ? hull_vertex(h, bi) : splat(0.0f);
  }
  // support_world / core_support_world of this lane's shape at d
  __device__ __forceinline__ V3 world(V3 d, int lane, bool core) const {
    ShapeRef s;
    s.type = type;
    s.a0 = a0;
    s.a1 = a1;
    s.a2 = a2;
    s.a3 = 0.0f;
    s.hull = nullptr;
    s.nhull = n;
    s.margin = 0.0f;
    V3 ld = qrot(qconj(q), d);
    V3 l;
    if (type == HULL) l = hull_support(ld, lane);
    else l = core ? core_support_local(s, ld) : support_local(s, ld);
    return qrot(q, l) + p;
  }
  // core supportA(d) - core supportB(-d) of the own half's pair, with d
  // uniform within each half; the result is uniform within each half.
  __device__ __forceinline__ V3 mink(V3 d, int lane, int hh) const {
    const bool sb = (lane & 16) != 0;
    V3 w = world(sel(sb, -d, d), lane, true);
    V3 a0v{rdl_f(w.x, 0), rdl_f(w.y, 0), rdl_f(w.z, 0)};
    V3 b0v{rdl_f(w.x, 16), rdl_f(w.y, 16), rdl_f(w.z, 16)};
    V3 a1v{rdl_f(w.x, 32), rdl_f(w.y, 32), rdl_f(w.z, 32)};
    V3 b1v{rdl_f(w.x, 48), rdl_f(w.y, 48), rdl_f(w.z, 48)};
    return sel(hh != 0, a1v - b1v, a0v - b0v);
  }
  // full supportA(d) of the own half's pair (uniform within each half)
  __device__ __forceinline__ V3 support_a(V3 d, int lane, int hh) const {
    V3 w = world(d, lane, false);
    V3 a0v{rdl_f(w.x, 0), rdl_f(w.y, 0), rdl_f(w.z, 0)};
    V3 a1v{rdl_f(w.x, 32), rdl_f(w.y, 32), rdl_f(w.z, 32)};
    return sel(hh != 0, a1v, a0v);
  }
};

__device__ __forceinline__ V3 hv(const EpaHalf& S, int i) {
  float4 v = S.v[i];
  return V3{v.x, v.y, v.z};
}

// epa.rs:212-224 over the (<= 64) faces of the half, faces hl and hl + 32
// in lane hl: first face with the smallest |distance|, NaN distances
// skipped, a NaN first face kept.
template <bool TWO>
__device__ __forceinline__ int dual_best_t(const EpaHalf& S, int nfaces, int hl, int h) {
  const bool nan0 = is_nan(S.dist[0]);
  float v0 = INFINITY, v1 = INFINITY;
  if (hl < nfaces) {
    float d = fabsf(S.dist[hl]);
    v0 = is_nan(d) ? INFINITY : d;
  }
  if constexpr (TWO) {
    if (hl + 32 < nfaces) {
      float d = fabsf(S.dist[hl + 32]);
      v1 = is_nan(d) ? INFINITY : d;
    }
  }
  float m = half_ext_f<false>(v1 < v0 ? v1 : v0, h);
  uint32_t b0 = hballot(v0 == m, h);
  uint32_t b1 = TWO ? hballot(v1 == m, h) : 0u;
  if (nan0 || !(m < INFINITY)) return 0;
  return b0 ? ctz64((uint64_t)b0) : (b1 ? 32 + ctz64((uint64_t)b1) : 0);
}
// The second face slot is only scanned while a half holds more than 32 faces.
__device__ __forceinline__ int dual_best(const EpaHalf& S, int nfaces, int hl, int h) {
  return __ballot(nfaces > 32) ? dual_best_t<true>(S, nfaces, hl, h) : dual_best_t<false>(S, nfaces, hl, h);
}

// The fan face of horizon edge (u, w) and the new vertex (epa.rs:372-386).
__device__ __forceinline__ void dual_fan_face(EpaHalf& S, int d, uint32_t u, uint32_t w, V3 vu, V3 vw, V3 np,
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

// One face slot of a lane: face f of the half's store, with its
// visibility against the new point.
struct DualFace {
  float nx, ny, nz, d, pd;
  uint32_t abc;
  bool live, vis;
  __device__ __forceinline__ void load(const EpaHalf& S, int f, bool l, V3 np) {
    live = l;
    nx = ny = nz = d = pd = 0.0f;
    abc = 0u;
    if (l) {
      nx = S.nx[f];
      ny = S.ny[f];
      nz = S.nz[f];
      d = S.dist[f];
      pd = S.pd[f];
      abc = S.abc[f];
    }
    vis = l && dot(V3{nx, ny, nz}, np) - pd > 1e-5f;
  }
  __device__ __forceinline__ void store(EpaHalf& S, int f) const {
    S.nx[f] = nx;
    S.ny[f] = ny;
    S.nz[f] = nz;
    S.dist[f] = d;
    S.pd[f] = pd;
    S.abc[f] = abc;
  }
};

// epa.rs:309-388 for the halves with `act` set, faces hl and hl + 32 in lane
// hl.  Per-half result: 1 done (or nothing visible), 0 degenerate horizon
// (polytope untouched), -2 the pair leaves the fast path (polytope
// untouched).
template <bool TWO>
__device__ __forceinline__ int dual_expand_t(EpaHalf& S, int& nfaces, V3 np, int nidx, int hl, int h, bool act) {
  DualFace f0, f1;
  f0.load(S, hl, act && hl < nfaces, np);
  f1.load(S, hl + 32, TWO && act && hl + 32 < nfaces, np);
  const uint32_t vm0 = hballot(f0.vis, h), vm1 = TWO ? hballot(f1.vis, h) : 0u;
  int res = 1;
  bool go = act && (vm0 | vm1) != 0u;
  const int nv0 = __popc(vm0);
  const int nv = nv0 + __popc(vm1);
  if (go && nv > 10) {
    res = -2;
    go = false;
  }
  if (!__ballot(go)) return res;
  // Edge events: lane l < 3 nv holds edge l % 3 of the visible face of rank
  // l / 3, in the reference's visiting order (face index order).
  const int nev = go ? 3 * nv : 0;
  const int rank = hl / 3;
  const int e = hl - 3 * rank;
  const uint32_t below = (uint32_t)lanes_below(hl);
  if (go && f0.vis) S.vis_abc[__popc(vm0 & below)] = f0.abc;
  if (go && f1.vis) S.vis_abc[nv0 + __popc(vm1 & below)] = f1.abc;
  wave_sync();
  const uint32_t eabc = (go && hl < nev) ? S.vis_abc[rank] : 0u;
  const uint32_t fa = eabc & 0xffu, fb = (eabc >> 8) & 0xffu, fc = (eabc >> 16) & 0xffu;
  const uint32_t u = e == 0 ? fa : (e == 1 ? fb : fc);
  const uint32_t w = e == 0 ? fb : (e == 1 ? fc : fa);
  const uint32_t edge = u | (w << 8), rev = w | (u << 8);
  const bool is_ev = go && hl < nev;
  // The earlier event that is this edge's reverse (it is then still on the
  // list, so this event cancels it); a repeated directed edge leaves the
  // fast path.
  int match = -1;
  bool dup = false;
  const int nevmax = hmax_i(nev);
  for (int k = 0; k < nevmax; ++k) {
    uint32_t ek = hget_u(edge, k, h);
    bool kin = k < nev;
    if (kin && ek == rev && k < hl) match = k;
    if (kin && ek == edge && k != hl) dup = true;
  }
  if (go && hballot(is_ev && dup, h)) {
    res = -2;
    go = false;
  }
  if (!__ballot(go)) return res;
  // Replay of the pushes and swap-removes on the half's LDS lists (list[p]
  // = event at position p, pos[v] = position of event v): every lane of
  // the half reads the same address, one lane writes.
  const uint32_t cancel = hballot(is_ev && match >= 0, h);
  int ne = 0;
  for (int l = 0; l < nevmax; ++l) {
    const bool lin = go && l < nev;
    const bool c = ((cancel >> l) & 1u) != 0u;
    const int ml = hget_i(match, l, h);
    if (lin && c) {
      const int p = S.pos[ml];
      const int last = S.list[ne - 1];
      if (hl == 0) {
        S.list[p] = last;
        S.pos[last] = p;
      }
      --ne;
    } else if (lin) {
      if (hl == 0) {
        S.list[ne] = l;
        S.pos[l] = ne;
      }
      ++ne;
    }
  }
  wave_sync();
  // Horizon entry k (lane k of the half): the edge of the event at list
  // position k.
  const bool is_h = go && hl < ne;
  const int list = is_h ? S.list[hl] : 0;
  const uint32_t hedge = (uint32_t)__shfl((int)edge, 32 * h + list, 64);
  const uint32_t hu = hedge & 0xffu, hw = (hedge >> 8) & 0xffu;
  V3 vu = splat(0.0f), vw = splat(0.0f);
  bool bad = false;
  if (is_h) {
    vu = hv(S, hu);
    vw = hv(S, hw);
    bad = len(cross(vw - vu, np - vu)) < 1e-12f;
  }
  if (go && hballot(bad, h)) {
    res = 0;
    go = false;
  }
  // Commit: stable compaction of the faces that stay, then the fan.  A
  // polytope that would pass the store leaves the fast path instead.
  const bool k0 = go && f0.live && !f0.vis, k1 = TWO && go && f1.live && !f1.vis;
  const uint32_t km0 = hballot(k0, h), km1 = TWO ? hballot(k1, h) : 0u;
  const int kept0 = __popc(km0);
  const int kept = kept0 + __popc(km1);
  if (go && kept + ne > kDualF) {
    res = -2;
    go = false;
  }
  wave_sync();
  if (go && k0) f0.store(S, __popc(km0 & below));
  if (TWO && go && k1) f1.store(S, kept0 + __popc(km1 & below));
  wave_sync();
  if (go && is_h) dual_fan_face(S, kept + hl, hu, hw, vu, vw, np, nidx);
  if (go) nfaces = kept + ne;
  wave_sync();
  return res;
}
// The second face slot is only visited while a half holds more than 32 faces.
__device__ __forceinline__ int dual_expand(EpaHalf& S, int& nfaces, V3 np, int nidx, int hl, int h, bool act) {
  return __ballot(act && nfaces > 32) ? dual_expand_t<true>(S, nfaces, np, nidx, hl, h, act)
                                      : dual_expand_t<false>(S, nfaces, np, nidx, hl, h, act);
}

// epa.rs:127-205 for the pair of each half.  `act` false: the half has no
// pair.  Returns per half: the result in R, `deferred` when the pair left
// the fast path (R is then meaningless).
__device__ __forceinline__ void dual_epa(const DualShapes& ds, V3 t0, V3 t1, V3 t2, V3 t3, EpaHalf& S, int lane,
                                         int h, bool act, EpaResult& R, bool& deferred) {
  const int hl = lane & 31;
  deferred = false;
  R.status = 1;
  R.n = splat(0.0f);
  R.depth = 0.0f;
  R.face = -1;
  if (act && hl < 4) {
    V3 w = sel(hl == 0, t0, sel(hl == 1, t1, sel(hl == 2, t2, t3)));
    S.v[hl] = make_float4(w.x, w.y, w.z, 0.0f);
  }
  // initial faces (epa.rs:226-257): lane q builds quad q; survivors keep
  // quad order.  Quads (a, b, c, opposite): (0,1,2,3) (0,3,1,2) (0,2,3,1)
  // (1,3,2,0).
  bool valid = false;
  V3 qn = splat(0.0f);
  float qd = 0.0f, qpd = 0.0f;
  uint32_t qabc = 0;
  if (act && hl < 4) {
    int a = hl == 3 ? 1 : 0;
    int b = hl == 0 ? 1 : (hl == 2 ? 2 : 3);
    int c = hl == 0 ? 2 : (hl == 1 ? 1 : (hl == 2 ? 3 : 2));
    int opp = 3 - hl;
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
  const uint32_t vm = hballot(valid, h);
  int nfaces = __popc(vm);
  if (valid) {
    int d = __popc(vm & (uint32_t)lanes_below(hl));
    S.nx[d] = qn.x;
    S.ny[d] = qn.y;
    S.nz[d] = qn.z;
    S.dist[d] = qd;
    S.pd[d] = qpd;
    S.abc[d] = qabc;
  }
  wave_sync();
  bool alive = act && nfaces > 0;  // the half still has a pair in work
  int nverts = 4;
  int status = 2;
  bool stop = false;
  bool seeding = alive;
  for (int k = 0; k < 6; ++k) {  // seeding; a degenerate expansion only ends the seeding
    V3 d = k == 0 ? V3{1.0f, 0.0f, 0.0f} : k == 1 ? V3{-1.0f, 0.0f, 0.0f} : k == 2 ? V3{0.0f, 1.0f, 0.0f}
         : k == 3 ? V3{0.0f, -1.0f, 0.0f} : k == 4 ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 0.0f, -1.0f};
    V3 p = ds.mink(d, lane, h);
    int idx = nverts;
    if (seeding) {
      if (hl == 0) S.v[idx] = make_float4(p.x, p.y, p.z, 0.0f);
      ++nverts;
    }
    wave_sync();
    int e = dual_expand(S, nfaces, p, idx, hl, h, seeding);
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
    int bi = dual_best(S, nfaces, hl, h);
    V3 bn{S.nx[bi], S.ny[bi], S.nz[bi]};
    float bd = S.dist[bi];
    V3 np = ds.mink(bn, lane, h);
    float wdist = dot(bn, np) - bd;
    if (run && wdist <= 1e-4f) {
      R.status = 0;
      R.n = bn;
      R.depth = bd;
      R.face = bi;
      conv = true;
      run = false;
    }
    if (run && nverts >= BCH_WEPA_V) {
      status = 3;
      run = false;
    }
    int nidx = nverts;
    if (run) {
      if (hl == 0) S.v[nidx] = make_float4(np.x, np.y, np.z, 0.0f);
      ++nverts;
    }
    wave_sync();
    int e = dual_expand(S, nfaces, np, nidx, hl, h, run);
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
  int bi = dual_best(S, nfaces, hl, h);
  if (alive && !conv) {
    R.status = status;
    R.n = V3{S.nx[bi], S.ny[bi], S.nz[bi]};
    R.depth = S.dist[bi];
    R.face = bi;
  }
}

// Grid-stride over an EPA queue, two entries per wavefront (Src as in
// epa_queue_run).  Outputs as epa_queue_run; an entry that leaves the fast
// path is appended to defer_q (count in *n_defer) for the one-pair kernel.
template <class Src>
__device__ __forceinline__ void epa_dual_run(const uint32_t* __restrict__ n_queue, const float4* __restrict__ epa_q,
                                             const Src& src, uint32_t* __restrict__ ccount,
                                             float4* __restrict__ cslots, uint32_t stride,
                                             uint32_t* __restrict__ route, float4* __restrict__ defer_q,
                                             uint32_t* n_defer, Counters* cnt, EpaHalf* S) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;
  const int hl = lane & 31;
  const uint32_t nq = *n_queue;
  for (uint32_t base = 2u * blockIdx.x; base < nq; base += 2u * gridDim.x) {
    const uint32_t q = base + (uint32_t)h;
    const bool has = q < nq;
    const uint32_t qq = has ? q : base;  // an empty half reads its partner's entry, unused
    float4 a = epa_q[4 * (uint64_t)qq], b = epa_q[4 * (uint64_t)qq + 1], c = epa_q[4 * (uint64_t)qq + 2],
           d = epa_q[4 * (uint64_t)qq + 3];
    const uint32_t p = f2u(a.w);
    Pair pr = src.pair(p);
    const bool fits = DualShapes::fits(pr);
    DualShapes ds;
    ds.load(pr, lane);
    V3 t0{a.x, a.y, a.z}, t1{b.x, b.y, b.z}, t2{c.x, c.y, c.z}, t3{d.x, d.y, d.z};
    EpaResult e;
    bool deferred;
    dual_epa(ds, t0, t1, t2, t3, S[h], lane, h, has && fits, e, deferred);
    deferred = has && (deferred || !fits);
    const bool hit = has && !deferred && !(e.depth <= 0.0f);
    V3 point_a = ds.support_a(e.n, lane, h);
    float depth = 0.0f;
    if (hit) depth = (e.depth + pr.a.margin) + pr.b.margin;
    if (hl == 0 && has) {
      if (deferred) {
        uint32_t slot = atomicAdd(n_defer, 1u);
        defer_q[4 * (uint64_t)slot] = a;
        defer_q[4 * (uint64_t)slot + 1] = b;
        defer_q[4 * (uint64_t)slot + 2] = c;
        defer_q[4 * (uint64_t)slot + 3] = d;
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

}  // namespace bch
