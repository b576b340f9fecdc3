// bch_epa_wave.hpp -- EPA with one wavefront per pair, polytope in LDS.
//
// Same algorithm, statement order and tie rules as the reference
// (narrowphase/epa.rs:127-398): tetrahedron seed + six axis supports,
// closest face by edge-aware distance (first minimum), visibility
// n.p - d > 1e-5, transactional two-phase expansion, <= 255 iterations,
// <= 128 vertices.
//
// The 64 lanes share the work that has no order:
//  * hull supports: lane l holds hull vertices l and l + 64 in registers;
//    a DPP max over the projections, then a ballot of the lanes that reach
//    it, gives the lowest index with the largest projection = the
//    sequential first maximum (convex_hull.rs:55-66);
//  * closest face: the same with a min over |distance| (epa.rs:212-224);
//  * visibility, compaction of the surviving faces (ballot prefix counts
//    keep their order) and the faces of the horizon fan (one per lane).
// The horizon is the reference's swap-remove edge list, replayed exactly:
// every edge of the visible faces first finds, in parallel, the earlier
// edge that is its reverse; a scalar replay of the pushes and swap-removes
// over a lane-indexed list (v_readlane / lane selects, a few
// instructions per edge) then yields the reference's horizon order.  A
// polytope outside that fast path (more than 64 faces, more than 21 visible
// faces, or a directed edge shared by two visible faces) runs the general
// stitch, which searches the list with one ballot per 64 entries.
//
// The polytope store is a template: the one-pair kernels keep 256 faces in
// LDS (8 KB per wavefront).  A convex polytope of <= 128 vertices has at most
// 2 * 128 - 4 = 252 faces, so only a numerically non-convex expansion (a
// visible set that is not one disk) can outgrow it; such an entry goes to an
// overflow queue and k_epa_big reruns it from the start with a 2048-face
// store (51 KB of LDS), so one pathological pair never fails a pass.
#pragma once
#include "bch_kernels.hpp"

namespace bch {

#define BCH_WEPA_V 128
#define BCH_WEPA_F 256
#define BCH_WEPA_BIG_F 2048

// F faces; the general stitch holds up to F horizon edges, F / 64 registers
// per lane.
template <int F>
struct WaveEpaStore {
  static constexpr int kF = F;
  float4 v[BCH_WEPA_V];
  float nx[F], ny[F], nz[F], dist[F], pd[F];
  uint32_t abc[F];
};
using EpaLds = WaveEpaStore<BCH_WEPA_F>;
using EpaBig = WaveEpaStore<BCH_WEPA_BIG_F>;

// Order LDS traffic between lanes of one wavefront.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int rdl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint32_t rdl_u(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ __forceinline__ float rdl_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
// lane l of `old` replaced by the uniform value `val` (v_cmp + v_cndmask)
__device__ __forceinline__ int wrl_i(int val, int l, int old, int lane) { return lane == l ? val : old; }
__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane ? (~0ull >> (64 - lane)) : 0ull; }
__device__ __forceinline__ int ctz64(uint64_t m) { return (int)__builtin_ctzll(m); }

// Wavefront max / min of values without NaN (uniform result).  The DPP
// row steps leave each row's extremum in every lane of the row; the four
// rows are combined from lanes 0, 16, 32 and 48.
template <bool MAX>
__device__ __forceinline__ float wave_ext_f(float v) {
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
  a = MAX ? (b > a ? b : a) : (b < a ? b : a);
  c = MAX ? (d > c ? d : c) : (d < c ? d : c);
  return MAX ? (c > a ? c : a) : (c < a ? c : a);
}

__device__ __forceinline__ int wave_min_i(int v) {
  int o;
  o = dpp_i<0xB1>(v);
  v = o < v ? o : v;
  o = dpp_i<0x4E>(v);
  v = o < v ? o : v;
  o = dpp_i<0x141>(v);
  v = o < v ? o : v;
  o = dpp_i<0x140>(v);
  v = o < v ? o : v;
  int a = rdl_i(v, 0), b = rdl_i(v, 16), c = rdl_i(v, 32), d = rdl_i(v, 48);
  a = b < a ? b : a;
  c = d < c ? d : c;
  return c < a ? c : a;
}

// convex_hull.rs:55-66 over a hull in global memory (more vertices than the
// registers hold): per-lane first maximum, then the lowest index among the
// lanes that reach the wavefront maximum.
__device__ __forceinline__ V3 wave_hull_support_mem(const float4* h, int n, V3 d, int lane) {
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  for (int v = lane; v < n; v += 64) {
    float dd = dot(hull_vertex(h, v), d);
    if (dd > bv) {
      bv = dd;
      bi = v;
    }
  }
  float m = wave_ext_f<true>(bv);
  if (!(m > -INFINITY)) return splat(0.0f);
  int idx = wave_min_i(bv == m ? bi : 0x7fffffff);
  return hull_vertex(h, idx);
}

// Hull vertices held in registers: lane l keeps vertices l and l + 64.
struct RegHull {
  float x0, y0, z0, x1, y1, z1;
  const float4* h;
  int n;
  bool inreg;
  __device__ __forceinline__ void load(const ShapeRef& s, int lane) {
    n = s.type == HULL ? s.nhull : 0;
    h = s.hull;
    inreg = n <= 128;
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
    if (inreg && lane < n) a = h[lane];
    if (inreg && lane + 64 < n) b = h[lane + 64];
    x0 = a.x;
    y0 = a.y;
    z0 = a.z;
    x1 = b.x;
    y1 = b.y;
    z1 = b.z;
  }
  // convex_hull.rs:55-66: first maximiser of p.d (NaN projections never win).
  __device__ __forceinline__ V3 support(V3 d, int lane) const {
    if (!inreg) return wave_hull_support_mem(h, n, d, lane);
    float a = dot(V3{x0, y0, z0}, d), b = dot(V3{x1, y1, z1}, d);
    a = (lane < n && a > -INFINITY) ? a : -INFINITY;
    b = (lane + 64 < n && b > -INFINITY) ? b : -INFINITY;
    float m = wave_ext_f<true>(b > a ? b : a);
    if (!(m > -INFINITY)) return splat(0.0f);
    uint64_t ba = __ballot(a == m);
    if (ba) {
      int l = ctz64(ba);
      return V3{rdl_f(x0, l), rdl_f(y0, l), rdl_f(z0, l)};
    }
    int l = ctz64(__ballot(b == m));
    return V3{rdl_f(x1, l), rdl_f(y1, l), rdl_f(z1, l)};
  }
};

struct EpaPair {
  Pair pr;
  RegHull ha, hb;
};

__device__ __forceinline__ V3 wave_support_world(const ShapeRef& s, const RegHull& rh, V3 p, Q4 q, V3 d, int lane,
                                                 bool core) {
  V3 ld = qrot(qconj(q), d);
  V3 l = s.type == HULL ? rh.support(ld, lane) : (core ? core_support_local(s, ld) : support_local(s, ld));
  return qrot(q, l) + p;
}

__device__ __forceinline__ V3 wave_minkowski(const EpaPair& ep, V3 d, int lane) {
  V3 a = wave_support_world(ep.pr.a, ep.ha, ep.pr.pa, ep.pr.qa, d, lane, true);
  V3 b = wave_support_world(ep.pr.b, ep.hb, ep.pr.pb, ep.pr.qb, -d, lane, true);
  return a - b;
}

// Both core supports of one Minkowski point in the two halves of the
// wavefront at once: lanes 0-31 evaluate shape A at d, lanes 32-63 shape B
// at -d, each with the same arithmetic as the one-side form above; the two
// world points are then read from lanes 0 and 32.  Hulls of at most 64
// vertices: lane l holds vertices l % 32 and l % 32 + 32 of its side.
struct SplitPair {
  int type;
  float a0, a1, a2;
  V3 p;
  Q4 q;
  float x0, y0, z0, x1, y1, z1;
  int n;
  // pairs whose hulls fit: at most 64 vertices each
  __device__ __forceinline__ static bool fits(const Pair& pr) {
    return (pr.a.type != HULL || pr.a.nhull <= 64) && (pr.b.type != HULL || pr.b.nhull <= 64);
  }
  __device__ __forceinline__ void load(const Pair& pr, int lane) {
    bool hi = lane >= 32;
    const ShapeRef& sa = pr.a;
    const ShapeRef& sb = pr.b;
    type = hi ? sb.type : sa.type;
    a0 = hi ? sb.a0 : sa.a0;
    a1 = hi ? sb.a1 : sa.a1;
    a2 = hi ? sb.a2 : sa.a2;
    p = sel(hi, pr.pb, pr.pa);
    q = Q4{hi ? pr.qb.x : pr.qa.x, hi ? pr.qb.y : pr.qa.y, hi ? pr.qb.z : pr.qa.z, hi ? pr.qb.w : pr.qa.w};
    n = type == HULL ? (hi ? sb.nhull : sa.nhull) : 0;
    const float4* h = hi ? sb.hull : sa.hull;
    int hl = lane & 31;
    float4 u = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v = u;
    if (hl < n) u = h[hl];
    if (hl + 32 < n) v = h[hl + 32];
    x0 = u.x; y0 = u.y; z0 = u.z;
    x1 = v.x; y1 = v.y; z1 = v.z;
  }
  // convex_hull.rs:55-66 per half: first maximiser of p.d.
  __device__ __forceinline__ V3 hull_support(V3 d, int lane) const {
    const bool hi = lane >= 32;
    const int hl = lane & 31;
    float a = dot(V3{x0, y0, z0}, d), b = dot(V3{x1, y1, z1}, d);
    a = (hl < n && a > -INFINITY) ? a : -INFINITY;
    b = (hl + 32 < n && b > -INFINITY) ? b : -INFINITY;
    float v = b > a ? b : a;
    float o;
    o = dpp_f<0xB1>(v);
    v = o > v ? o : v;
    o = dpp_f<0x4E>(v);
    v = o > v ? o : v;
    o = dpp_f<0x141>(v);
    v = o > v ? o : v;
    o = dpp_f<0x140>(v);
    v = o > v ? o : v;
    float r0 = rdl_f(v, 0), r1 = rdl_f(v, 16), r2 = rdl_f(v, 32), r3 = rdl_f(v, 48);
    float mA = r1 > r0 ? r1 : r0, mB = r3 > r2 ? r3 : r2;
    float m = hi ? mB : mA;
    uint64_t ba = __ballot(a == m), bb = __ballot(b == m);
    uint32_t aA = (uint32_t)ba, aB = (uint32_t)(ba >> 32), bA = (uint32_t)bb, bB = (uint32_t)(bb >> 32);
    int lA = aA ? ctz64(aA) : (bA ? ctz64(bA) : 0);
    int lB = 32 + (aB ? ctz64(aB) : (bB ? ctz64(bB) : 0));
    V3 vA = aA ? V3{rdl_f(x0, lA), rdl_f(y0, lA), rdl_f(z0, lA)} : V3{rdl_f(x1, lA), rdl_f(y1, lA), rdl_f(z1, lA)};
    V3 vB = aB ? V3{rdl_f(x0, lB), rdl_f(y0, lB), rdl_f(z0, lB)} : V3{rdl_f(x1, lB), rdl_f(y1, lB), rdl_f(z1, lB)};
    V3 r = sel(hi, vB, vA);
    return m > -INFINITY ? r : splat(0.0f);
  }
  // support_world / core_support_world of this lane's side at its direction
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
  // the Minkowski point supportA(d) - supportB(-d) (cores), uniform
  __device__ __forceinline__ V3 operator()(V3 d, int lane) const {
    const bool hi = lane >= 32;
    V3 w = world(sel(hi, -d, d), lane, true);
    V3 a{rdl_f(w.x, 0), rdl_f(w.y, 0), rdl_f(w.z, 0)};
    V3 b{rdl_f(w.x, 32), rdl_f(w.y, 32), rdl_f(w.z, 32)};
    return a - b;
  }
  // supportA(d) (full shape), uniform
  __device__ __forceinline__ V3 support_a(V3 d, int lane) const {
    V3 w = world(d, lane, false);
    return V3{rdl_f(w.x, 0), rdl_f(w.y, 0), rdl_f(w.z, 0)};
  }
};

// The one-side form as the same functor interface (hulls of any size).
struct SeqPair {
  const EpaPair* ep;
  __device__ __forceinline__ V3 operator()(V3 d, int lane) const { return wave_minkowski(*ep, d, lane); }
  __device__ __forceinline__ V3 support_a(V3 d, int lane) const {
    return wave_support_world(ep->pr.a, ep->ha, ep->pr.pa, ep->pr.qa, d, lane, false);
  }
};

template <class St>
__device__ __forceinline__ V3 lv(const St& S, int i) {
  float4 v = S.v[i];
  return V3{v.x, v.y, v.z};
}

// epa.rs:212-224: index of the first face with the smallest |distance|
// (NaN distances are skipped; a NaN first face is never replaced).
template <class St>
__device__ __forceinline__ int wave_best(const St& S, int nfaces, int lane) {
  if (is_nan(S.dist[0])) return 0;
  if (nfaces <= 64) {
    float v = INFINITY;
    if (lane < nfaces) {
      float d = fabsf(S.dist[lane]);
      v = is_nan(d) ? INFINITY : d;
    }
    float m = wave_ext_f<false>(v);
    if (!(m < INFINITY)) return 0;
    return ctz64(__ballot(v == m));
  }
  float bv = INFINITY;
  int bi = 0x7fffffff;
  for (int f = lane; f < nfaces; f += 64) {
    float d = fabsf(S.dist[f]);
    if (d < bv) {
      bv = d;
      bi = f;
    }
  }
  float m = wave_ext_f<false>(bv);
  if (!(m < INFINITY)) return 0;
  return wave_min_i(bv == m ? bi : 0x7fffffff);
}

template <class St>
__device__ __forceinline__ bool face_vis_lds(const St& S, int f, V3 np) {
  return dot(V3{S.nx[f], S.ny[f], S.nz[f]}, np) - S.pd[f] > 1e-5f;
}

// The fan face of horizon edge (u, w) and the new vertex (epa.rs:372-386).
template <class St>
__device__ __forceinline__ void fan_face(St& S, int d, uint32_t u, uint32_t w, V3 vu, V3 vw, V3 np, int nidx) {
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

// epa.rs:309-388, general form (any face count).  Returns 1 ok (or nothing
// visible), 0 degenerate horizon (hull untouched), -1 store overflow: more
// than `flimit` (<= St::kF) faces or St::kF horizon edges (hull untouched).
template <class St>
__device__ __noinline__ int wave_expand_general(St& S, int& nfaces, V3 np, int nidx, int lane, int flimit) {
  constexpr int R = St::kF / 64;
  const int nchunks = (nfaces + 63) / 64;
  // Horizon in visible-face order with swap-remove cancellation; entry k of
  // the edge list is held by lane k % 64 in register bl[k / 64].
  uint32_t bl[R];
#pragma unroll
  for (int k = 0; k < R; ++k) bl[k] = 0xffffffffu;
  int ne = 0;
  bool over = false, any = false;
  for (int c = 0; c < nchunks; ++c) {
    int f = lane + 64 * c;
    bool v = false;
    uint32_t ab = 0u;
    if (f < nfaces) {
      v = face_vis_lds(S, f, np);
      ab = S.abc[f];
    }
    uint64_t m = __ballot(v);
    any = any || m != 0ull;
    while (m) {
      int b = ctz64(m);
      m &= m - 1ull;
      uint32_t t = rdl_u(ab, b);
      uint32_t fa = t & 0xffu, fb = (t >> 8) & 0xffu, fc = (t >> 16) & 0xffu;
      for (int e = 0; e < 3; ++e) {
        uint32_t u = e == 0 ? fa : (e == 1 ? fb : fc);
        uint32_t w = e == 0 ? fb : (e == 1 ? fc : fa);
        uint32_t rev = w | (u << 8);
        int found = -1;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if (found < 0 && 64 * k < ne) {
            uint64_t h = __ballot(lane + 64 * k < ne && bl[k] == rev);
            if (h) found = 64 * k + ctz64(h);
          }
        }
        if (found >= 0) {
          int last = ne - 1;
          uint32_t lastv = 0;
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (last / 64 == k) lastv = rdl_u(bl[k], last % 64);
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (found / 64 == k && lane == found % 64) bl[k] = lastv;
          --ne;
        } else if (ne < 64 * R) {
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (ne / 64 == k && lane == ne % 64) bl[k] = u | (w << 8);
          ++ne;
        } else {
          over = true;
        }
      }
    }
  }
  if (!any) return 1;
  if (over) return -1;
  // Validate every horizon edge before touching the hull.
  bool bad = false;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    if (lane + 64 * k < ne) {
      uint32_t e = bl[k];
      V3 vu = lv(S, e & 0xffu);
      if (len(cross(lv(S, (e >> 8) & 0xffu) - vu, np - vu)) < 1e-12f) bad = true;
    }
  }
  if (__ballot(bad)) return 0;
  // Count the survivors first: a hull that would outgrow the store is left
  // untouched.
  int kept = 0;
  for (int c = 0; c < nchunks; ++c) {
    int f = lane + 64 * c;
    bool live = f < nfaces;
    kept += __popcll(__ballot(live && !face_vis_lds(S, live ? f : 0, np)));
  }
  if (kept + ne > flimit) return -1;
  // Commit: stable compaction of the faces that stay (visibility re-evaluated
  // on the same values, as the reference does) ...
  int outp = 0;
  for (int c = 0; c < nchunks; ++c) {
    int f = lane + 64 * c;
    bool live = f < nfaces;
    uint64_t keepm = __ballot(live && !face_vis_lds(S, live ? f : 0, np));
    bool keep = (keepm >> lane) & 1ull;
    int d = outp + __popcll(keepm & lanes_below(lane));
    outp += __popcll(keepm);
    float fnx = 0, fny = 0, fnz = 0, fd = 0, fpd = 0;
    uint32_t fabc = 0;
    if (keep) {
      fnx = S.nx[f];
      fny = S.ny[f];
      fnz = S.nz[f];
      fd = S.dist[f];
      fpd = S.pd[f];
      fabc = S.abc[f];
    }
    wave_sync();
    if (keep) {
      S.nx[d] = fnx;
      S.ny[d] = fny;
      S.nz[d] = fnz;
      S.dist[d] = fd;
      S.pd[d] = fpd;
      S.abc[d] = fabc;
    }
    wave_sync();
  }
  // ... then the horizon fan, in horizon order (entry k built by its lane).
#pragma unroll
  for (int kk = 0; kk < R; ++kk) {
    int k = lane + 64 * kk;
    if (k >= ne) continue;
    uint32_t e = bl[kk];
    uint32_t u = e & 0xffu, w = (e >> 8) & 0xffu;
    fan_face(S, kept + k, u, w, lv(S, u), lv(S, w), np, nidx);
  }
  nfaces = kept + ne;
  wave_sync();
  return 1;
}

// epa.rs:309-388.  Same contract as wave_expand_general.
template <class St>
__device__ __forceinline__ int wave_expand(St& S, int& nfaces, V3 np, int nidx, int lane, int flimit) {
  if (nfaces > 64) return wave_expand_general(S, nfaces, np, nidx, lane, flimit);
  const bool live = lane < nfaces;
  float fnx = 0.0f, fny = 0.0f, fnz = 0.0f, fd = 0.0f, fpd = 0.0f;
  uint32_t fabc = 0u;
  if (live) {
    fnx = S.nx[lane];
    fny = S.ny[lane];
    fnz = S.nz[lane];
    fd = S.dist[lane];
    fpd = S.pd[lane];
    fabc = S.abc[lane];
  }
  const bool vis = live && dot(V3{fnx, fny, fnz}, np) - fpd > 1e-5f;
  const uint64_t vm = __ballot(vis);
  if (!vm) return 1;
  const int nv = __popcll(vm);
  if (nv > 21) return wave_expand_general(S, nfaces, np, nidx, lane, flimit);
  // Edge events: lane l < 3 nv holds edge l % 3 of the visible face of rank
  // l / 3, in the reference's visiting order.
  const int nev = 3 * nv;
  const int rank = lane / 3;
  const int e = lane - 3 * rank;
  uint32_t eabc = 0u;
  {
    uint64_t mm = vm;
    for (int r = 0; r < nv; ++r) {
      int b = ctz64(mm);
      mm &= mm - 1ull;
      uint32_t t = rdl_u(fabc, b);
      if (rank == r) eabc = t;
    }
  }
  const uint32_t fa = eabc & 0xffu, fb = (eabc >> 8) & 0xffu, fc = (eabc >> 16) & 0xffu;
  const uint32_t u = e == 0 ? fa : (e == 1 ? fb : fc);
  const uint32_t w = e == 0 ? fb : (e == 1 ? fc : fa);
  const uint32_t edge = u | (w << 8), rev = w | (u << 8);
  const bool is_ev = lane < nev;
  // The earlier event that is this edge's reverse (it is then still on the
  // list, so this event cancels it); any repeated directed edge leaves the
  // fast path.
  int match = -1;
  bool dup = false;
  for (int k = 0; k < nev; ++k) {
    uint32_t ek = rdl_u(edge, k);
    if (ek == rev && k < lane) match = k;
    if (ek == edge && k != lane) dup = true;
  }
  if (__ballot(is_ev && dup)) return wave_expand_general(S, nfaces, np, nidx, lane, flimit);
  // Replay of the pushes and swap-removes: lane p of `list` holds the event
  // at list position p, lane v of `pos` the position of event v.
  const uint64_t cancel = __ballot(is_ev && match >= 0);
  int list = 0, pos = 0, ne = 0;
  for (int l = 0; l < nev; ++l) {
    if ((cancel >> l) & 1ull) {
      int p = rdl_i(pos, rdl_i(match, l));
      int last = rdl_i(list, ne - 1);
      list = wrl_i(last, p, list, lane);
      pos = wrl_i(p, last, pos, lane);
      --ne;
    } else {
      list = wrl_i(l, ne, list, lane);
      pos = wrl_i(ne, l, pos, lane);
      ++ne;
    }
  }
  // Horizon entry k (lane k): the edge of the event at list position k.
  const bool is_h = lane < ne;
  const uint32_t hedge = (uint32_t)__shfl((int)edge, is_h ? list : 0, 64);
  const uint32_t hu = hedge & 0xffu, hw = (hedge >> 8) & 0xffu;
  V3 vu = splat(0.0f), vw = splat(0.0f);
  bool bad = false;
  if (is_h) {
    vu = lv(S, hu);
    vw = lv(S, hw);
    bad = len(cross(vw - vu, np - vu)) < 1e-12f;
  }
  if (__ballot(bad)) return 0;
  // Commit: stable compaction of the faces that stay, then the fan (at most
  // 64 + 63 faces, within the store unless the caller set a lower limit).
  const uint64_t keepm = __ballot(live && !vis);
  const int kept = __popcll(keepm);
  if (kept + ne > flimit) return -1;
  wave_sync();
  if (live && !vis) {
    int d = __popcll(keepm & lanes_below(lane));
    S.nx[d] = fnx;
    S.ny[d] = fny;
    S.nz[d] = fnz;
    S.dist[d] = fd;
    S.pd[d] = fpd;
    S.abc[d] = fabc;
  }
  wave_sync();
  if (is_h) fan_face(S, kept + lane, hu, hw, vu, vw, np, nidx);
  nfaces = kept + ne;
  wave_sync();
  return 1;
}

// epa.rs:127-205 for one pair; every lane returns the result.  `mink(d,
// lane)` is the wavefront-uniform Minkowski support.
template <class Mink, class St>
__device__ __forceinline__ EpaResult wave_epa(const Mink& mink, V3 t0, V3 t1, V3 t2, V3 t3, St& S, int lane,
                                              int flimit = St::kF) {
  EpaResult R;
  if (lane < 4) {
    V3 w = sel(lane == 0, t0, sel(lane == 1, t1, sel(lane == 2, t2, t3)));
    S.v[lane] = make_float4(w.x, w.y, w.z, 0.0f);
  }
  // initial faces (epa.rs:226-257): lane q builds quad q; survivors keep
  // quad order.  Quads (a, b, c, opposite): (0,1,2,3) (0,3,1,2) (0,2,3,1)
  // (1,3,2,0).
  bool valid = false;
  V3 qn = splat(0.0f);
  float qd = 0.0f, qpd = 0.0f;
  uint32_t qabc = 0;
  if (lane < 4) {
    int a = lane == 3 ? 1 : 0;
    int b = lane == 0 ? 1 : (lane == 2 ? 2 : 3);
    int c = lane == 0 ? 2 : (lane == 1 ? 1 : (lane == 2 ? 3 : 2));
    int opp = 3 - lane;
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
  uint64_t vm = __ballot(valid);
  int nfaces = __popcll(vm);
  if (valid) {
    int d = __popcll(vm & lanes_below(lane));
    S.nx[d] = qn.x;
    S.ny[d] = qn.y;
    S.nz[d] = qn.z;
    S.dist[d] = qd;
    S.pd[d] = qpd;
    S.abc[d] = qabc;
  }
  wave_sync();
  if (nfaces == 0) {
    R.status = 1;
    R.n = splat(0.0f);
    R.depth = 0.0f;
    R.face = -1;
    return R;
  }
  int nverts = 4;
  int status = 2;
  bool stop = false;
  for (int k = 0; k < 6; ++k) {  // seeding; a degenerate expansion only ends the seeding
    V3 d = k == 0 ? V3{1.0f, 0.0f, 0.0f} : k == 1 ? V3{-1.0f, 0.0f, 0.0f} : k == 2 ? V3{0.0f, 1.0f, 0.0f}
         : k == 3 ? V3{0.0f, -1.0f, 0.0f} : k == 4 ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 0.0f, -1.0f};
    V3 p = mink(d, lane);
    int idx = nverts;
    if (lane == 0) S.v[idx] = make_float4(p.x, p.y, p.z, 0.0f);
    ++nverts;
    wave_sync();
    int e = wave_expand(S, nfaces, p, idx, lane, flimit);
    if (e <= 0) {
      status = e == 0 ? 1 : 4;
      stop = e < 0;
      break;
    }
  }
  if (!stop) {
    for (int it = 0; it < 255; ++it) {
      int bi = wave_best(S, nfaces, lane);
      V3 bn{S.nx[bi], S.ny[bi], S.nz[bi]};
      float bd = S.dist[bi];
      V3 np = mink(bn, lane);
      float wdist = dot(bn, np) - bd;
      if (wdist <= 1e-4f) {
        R.status = 0;
        R.n = bn;
        R.depth = bd;
        R.face = bi;
        return R;
      }
      if (nverts >= BCH_WEPA_V) {
        status = 3;
        break;
      }
      int nidx = nverts;
      if (lane == 0) S.v[nidx] = make_float4(np.x, np.y, np.z, 0.0f);
      ++nverts;
      wave_sync();
      int e = wave_expand(S, nfaces, np, nidx, lane, flimit);
      if (e <= 0) {
        status = e == 0 ? 1 : 4;
        break;
      }
    }
  }
  int bi = wave_best(S, nfaces, lane);
  R.status = status;
  R.n = V3{S.nx[bi], S.ny[bi], S.nz[bi]};
  R.depth = S.dist[bi];
  R.face = bi;
  return R;
}

// The world pass's pairs as an EPA source: entry p is pair p of the list.
struct WorldEpaSrc {
  const uint32_t* pairs;
  const BodyRec* bodies;
  const float4* hull;
  __device__ __forceinline__ Pair pair(uint32_t p) const {
    uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    BodyRec A = bodies[i], B = bodies[j];
    Pair pr;
    pr.a = shape_of(A, hull);
    pr.b = shape_of(B, hull);
    pr.pa = pos_of(A);
    pr.pb = pos_of(B);
    pr.qa = quat_of(A);
    pr.qb = quat_of(B);
    return pr;
  }
};

// Grid-stride over the EPA queue, one wavefront per entry and one LDS
// polytope per wavefront; the contact of entry p goes to ccount[p] and
// cslots[stride * p ..] (convex_convex.rs:117-154: margins added back,
// witness = supportA(n) - n depth).  `route` (may be NULL) gets bit 3.  An
// entry whose polytope outgrows `flimit` faces is appended to `over_q`
// (counted in cnt->n_epa_over, its outputs left to the rerun) when that
// queue is given, and flags errors bit 1 otherwise.
template <class Src, class St>
__device__ __forceinline__ void epa_queue_run(const uint32_t* __restrict__ n_queue, const float4* __restrict__ epa_q,
                                              const Src& src, uint32_t* __restrict__ ccount,
                                              float4* __restrict__ cslots, uint32_t stride,
                                              uint32_t* __restrict__ route, Counters* cnt, St& S,
                                              float4* __restrict__ over_q = nullptr, int flimit = St::kF) {
  const int lane = threadIdx.x & 63;
  const uint32_t nq = *n_queue;
  for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
    float4 a = epa_q[4 * (uint64_t)q], b = epa_q[4 * (uint64_t)q + 1], c = epa_q[4 * (uint64_t)q + 2],
           d = epa_q[4 * (uint64_t)q + 3];
    uint32_t p = f2u(a.w);
    Pair pr = src.pair(p);
    V3 t0{a.x, a.y, a.z}, t1{b.x, b.y, b.z}, t2{c.x, c.y, c.z}, t3{d.x, d.y, d.z};
    EpaResult e;
    V3 point_a = splat(0.0f);
    if (SplitPair::fits(pr)) {
      SplitPair sp;
      sp.load(pr, lane);
      e = wave_epa(sp, t0, t1, t2, t3, S, lane, flimit);
      if (!(e.depth <= 0.0f)) point_a = sp.support_a(e.n, lane);
    } else {
      EpaPair ep;
      ep.pr = pr;
      ep.ha.load(pr.a, lane);
      ep.hb.load(pr.b, lane);
      SeqPair sq{&ep};
      e = wave_epa(sq, t0, t1, t2, t3, S, lane, flimit);
      if (!(e.depth <= 0.0f)) point_a = sq.support_a(e.n, lane);
    }
    bool hit = !(e.depth <= 0.0f);
    float depth = 0.0f;
    if (hit) depth = (e.depth + pr.a.margin) + pr.b.margin;
    if (e.status == 4 && over_q) {
      if (lane == 0) {
        uint32_t k = atomicAdd(&cnt->n_epa_over, 1u);
        over_q[4 * (uint64_t)k] = a;
        over_q[4 * (uint64_t)k + 1] = b;
        over_q[4 * (uint64_t)k + 2] = c;
        over_q[4 * (uint64_t)k + 3] = d;
      }
      wave_sync();
      continue;
    }
    if (lane == 0) {
      if (e.status == 4) atomicOr(&cnt->errors, 2u);
      ccount[p] = hit ? 1u : 0u;
      if (route) route[p] |= 8u;
      if (hit) {
        V3 pos = point_a - e.n * depth;
        cslots[(uint64_t)p * stride] = make_float4(pos.x, pos.y, pos.z, depth);
        cslots[(uint64_t)p * stride + 1] = make_float4(e.n.x, e.n.y, e.n.z, 0.0f);
      }
    }
    wave_sync();
  }
}

// `over_q` receives the entries that outgrow `flimit` faces (256 unless a
// test lowers it) for k_epa_big.
__global__ __launch_bounds__(64) void k_epa_wave(const uint32_t* __restrict__ n_queue,
                                                 const float4* __restrict__ epa_q, const uint32_t* __restrict__ pairs,
                                                 const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                                                 uint32_t* __restrict__ ccount, float4* __restrict__ cslots,
                                                 uint32_t* __restrict__ route, Counters* cnt,
                                                 float4* __restrict__ over_q, int flimit) {
  __shared__ EpaLds S;
  epa_queue_run(n_queue, epa_q, WorldEpaSrc{pairs, bodies, hull}, ccount, cslots, 8u, route, cnt, S, over_q, flimit);
}

// The overflow queue of k_epa_wave, rerun from the start with the 2048-face
// store; an entry that outgrows even that flags errors bit 1.
__global__ __launch_bounds__(64) void k_epa_big(const uint32_t* __restrict__ n_over,
                                                const float4* __restrict__ over_q, const uint32_t* __restrict__ pairs,
                                                const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                                                uint32_t* __restrict__ ccount, float4* __restrict__ cslots,
                                                uint32_t* __restrict__ route, Counters* cnt) {
  __shared__ EpaBig S;
  epa_queue_run(n_over, over_q, WorldEpaSrc{pairs, bodies, hull}, ccount, cslots, 8u, route, cnt, S);
}

}  // namespace bch
