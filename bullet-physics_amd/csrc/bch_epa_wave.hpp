// bch_epa_wave.hpp -- EPA with a group of G lanes per pair (G = 16, 32 or
// 64), polytope in LDS.
//
// Same algorithm, statement order and tie rules as the reference
// (narrowphase/epa.rs:127-398): tetrahedron seed + six axis supports,
// closest face by edge-aware distance (first minimum), visibility
// n.p - d > 1e-5, transactional two-phase expansion, <= 255 iterations,
// <= 128 vertices.  The G lanes of a group share the work that has no
// order: face scans (closest face, visibility), hull support scans
// (largest projection, lowest index = the sequential first maximum), face
// compaction (ballot prefix counts keep the survivors' order) and the new
// faces of the horizon fan.  The horizon is stitched in visible-face order
// with the reference's swap-remove edge cancellation on a register-resident
// edge list, so faces are created in exactly the reference order.
//
// A wavefront runs 64 / G pairs side by side: every instruction is issued
// for the whole wavefront, so small groups cut the issue cost per pair.
#pragma once
#include "bch_kernels.hpp"

namespace bch {

#define BCH_WEPA_V 128
#define BCH_WEPA_F 256

struct EpaLds {
  float4 v[BCH_WEPA_V];
  float nx[BCH_WEPA_F], ny[BCH_WEPA_F], nz[BCH_WEPA_F], dist[BCH_WEPA_F], pd[BCH_WEPA_F];
  uint32_t abc[BCH_WEPA_F];
};

// Order LDS traffic between lanes of one wavefront (no cross-wave barrier:
// groups run their own pairs).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Lane-group view of a wavefront: G lanes per pair, group `grp` owns lanes
// [grp * G, grp * G + G).
template <int G>
struct Grp {
  int sub;   // lane within the group
  int base;  // first wavefront lane of the group
  __device__ __forceinline__ uint64_t ballot(bool p) const {
    uint64_t m = __ballot(p);
    return G == 64 ? m : (m >> base) & ((1ull << G) - 1ull);
  }
  __device__ __forceinline__ uint64_t below() const { return sub ? ((1ull << sub) - 1ull) : 0ull; }
  // value of `v` on group lane k (k uniform within the group); a whole
  // wavefront reads it with v_readlane, smaller groups with a permute.
  __device__ __forceinline__ uint32_t bcast(uint32_t v, int k) const {
    if (G == 64) return (uint32_t)__builtin_amdgcn_readlane((int)v, __builtin_amdgcn_readfirstlane(k));
    return (uint32_t)__shfl((int)v, base + k, 64);
  }
  __device__ __forceinline__ float bcastf(float v, int k) const { return __uint_as_float(bcast(__float_as_uint(v), k)); }
};

// (max value, lowest index) or (min value, lowest index) over the group.
template <bool MAX, int G>
__device__ __forceinline__ void group_reduce(float& v, int& i) {
  if (G == 64) {
    wave_reduce<MAX>(v, i);
    return;
  }
  row_reduce<MAX, 16>(v, i);
  if (G == 32) {
    float ov = __shfl_xor(v, 16, 64);
    int oi = __shfl_xor(i, 16, 64);
    if (MAX) amax_merge(v, i, ov, oi); else amin_merge(v, i, ov, oi);
  }
}

template <int G>
__device__ __forceinline__ V3 grp_hull_support(const float4* h, int n, V3 d, const Grp<G>& g) {
  float bv = -INFINITY;
  int bi = -1;
  for (int v = g.sub; v < n; v += G) {
    float dd = dot(hull_vertex(h, v), d);
    if (dd > bv) {
      bv = dd;
      bi = v;
    }
  }
  group_reduce<true, G>(bv, bi);
  return bi >= 0 ? hull_vertex(h, bi) : splat(0.0f);
}

// Hull vertices of one shape held in registers: group lane `sub` keeps
// vertices sub, sub + G, ... (NR of them).  Hulls with more vertices than
// the group holds scan global memory instead.
template <int G>
struct RegHull {
  static constexpr int NR = 128 / G;
  float x[NR], y[NR], z[NR];
  const float4* h;
  int n;
  bool inreg;
  __device__ __forceinline__ void load(const ShapeRef& s, const Grp<G>& g) {
    n = s.type == HULL ? s.nhull : 0;
    h = s.hull;
    inreg = n <= G * NR;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      int v = g.sub + G * r;
      float4 t = (inreg && v < n) ? h[v] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      x[r] = t.x;
      y[r] = t.y;
      z[r] = t.z;
    }
  }
  // convex_hull.rs:55-66: first maximiser of p.d over the vertices.
  __device__ __forceinline__ V3 support(V3 d, const Grp<G>& g) const {
    if (!inreg) return grp_hull_support<G>(h, n, d, g);
    float bv = -INFINITY;
    int bi = -1;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      int v = g.sub + G * r;
      if (v < n) {
        float dd = dot(V3{x[r], y[r], z[r]}, d);
        if (dd > bv) {
          bv = dd;
          bi = v;
        }
      }
    }
    group_reduce<true, G>(bv, bi);
    if (bi < 0) return splat(0.0f);
    int rr = bi / G, ll = bi % G;
    float sx = x[0], sy = y[0], sz = z[0];
#pragma unroll
    for (int r = 1; r < NR; ++r) {
      if (rr == r) {
        sx = x[r];
        sy = y[r];
        sz = z[r];
      }
    }
    return V3{g.bcastf(sx, ll), g.bcastf(sy, ll), g.bcastf(sz, ll)};
  }
};

template <int G>
__device__ __forceinline__ V3 grp_support_world(const ShapeRef& s, const RegHull<G>& rh, V3 p, Q4 q, V3 d,
                                                const Grp<G>& g, bool core) {
  V3 ld = qrot(qconj(q), d);
  V3 l = s.type == HULL ? rh.support(ld, g) : (core ? core_support_local(s, ld) : support_local(s, ld));
  return qrot(q, l) + p;
}

template <int G>
struct EpaPair {
  Pair pr;
  RegHull<G> ha, hb;
};

template <int G>
__device__ __forceinline__ V3 grp_minkowski(const EpaPair<G>& ep, V3 d, const Grp<G>& g) {
  V3 a = grp_support_world<G>(ep.pr.a, ep.ha, ep.pr.pa, ep.pr.qa, d, g, true);
  V3 b = grp_support_world<G>(ep.pr.b, ep.hb, ep.pr.pb, ep.pr.qb, -d, g, true);
  return a - b;
}

__device__ __forceinline__ V3 lv(const EpaLds& S, int i) {
  float4 v = S.v[i];
  return V3{v.x, v.y, v.z};
}

// epa.rs:212-224: index of the first face with the smallest |distance|
// (a NaN first face is never replaced).
template <int G>
__device__ __forceinline__ int grp_best(const EpaLds& S, int nfaces, const Grp<G>& g) {
  float bv = INFINITY;
  int bi = -1;
  for (int f = g.sub; f < nfaces; f += G) {
    float v = fabsf(S.dist[f]);
    if (!is_nan(v) && (bi < 0 || v < bv)) {
      bv = v;
      bi = f;
    }
  }
  group_reduce<false, G>(bv, bi);
  if (bi < 0 || is_nan(S.dist[0])) bi = 0;
  return bi;
}

// Horizon edge list capacity in registers: R entries per lane (G * R edges).
template <int G>
struct HorizonCap {
  static constexpr int R = G >= 64 ? 4 : (G >= 32 ? 8 : 8);
};

__device__ __forceinline__ bool face_vis_lds(const EpaLds& S, int f, V3 np) {
  return dot(V3{S.nx[f], S.ny[f], S.nz[f]}, np) - S.pd[f] > 1e-5f;
}

// epa.rs:309-388.  Returns 1 ok (or nothing visible), 0 degenerate horizon
// (hull untouched), -1 capacity exceeded (hull untouched).
template <int G>
__device__ __forceinline__ int grp_expand(EpaLds& S, int& nfaces, V3 np, int nidx, const Grp<G>& g) {
  constexpr int R = HorizonCap<G>::R;
  const int nchunks = (nfaces + G - 1) / G;
  // Horizon in visible-face order with swap-remove cancellation; entry k of
  // the edge list is held by group lane k % G in register bl[k / G].
  uint32_t bl[R];
#pragma unroll
  for (int k = 0; k < R; ++k) bl[k] = 0xffffffffu;
  int ne = 0;
  bool over = false, any = false;
  for (int c = 0; c < nchunks; ++c) {
    int f = g.sub + G * c;
    bool v = false;
    uint32_t ab = 0u;
    if (f < nfaces) {
      v = face_vis_lds(S, f, np);
      ab = S.abc[f];
    }
    uint64_t m = g.ballot(v);
    any = any || m != 0ull;
    while (m) {
      int b = __builtin_ctzll(m);
      m &= m - 1ull;
      uint32_t t = g.bcast(ab, b);
      uint32_t fa = t & 0xffu, fb = (t >> 8) & 0xffu, fc = (t >> 16) & 0xffu;
      for (int e = 0; e < 3; ++e) {
        uint32_t u = e == 0 ? fa : (e == 1 ? fb : fc);
        uint32_t w = e == 0 ? fb : (e == 1 ? fc : fa);
        uint32_t rev = w | (u << 8);
        int found = -1;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if (found < 0 && G * k < ne) {
            uint64_t h = g.ballot(g.sub + G * k < ne && bl[k] == rev);
            if (h) found = G * k + __builtin_ctzll(h);
          }
        }
        if (found >= 0) {
          int last = ne - 1;
          uint32_t lastv = 0;
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (last / G == k) lastv = g.bcast(bl[k], last % G);
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (found / G == k && g.sub == found % G) bl[k] = lastv;
          --ne;
        } else if (ne < G * R) {
#pragma unroll
          for (int k = 0; k < R; ++k)
            if (ne / G == k && g.sub == ne % G) bl[k] = u | (w << 8);
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
    if (g.sub + G * k < ne) {
      uint32_t e = bl[k];
      V3 vu = lv(S, e & 0xffu);
      if (len(cross(lv(S, (e >> 8) & 0xffu) - vu, np - vu)) < 1e-12f) bad = true;
    }
  }
  if (g.ballot(bad)) return 0;
  // Count the survivors first: a hull that would outgrow the store is left
  // untouched.
  int kept = 0;
  for (int c = 0; c < nchunks; ++c) {
    int f = g.sub + G * c;
    bool live = f < nfaces;
    kept += __popcll(g.ballot(live && !face_vis_lds(S, live ? f : 0, np)));
  }
  if (kept + ne > BCH_WEPA_F) return -1;
  // Commit: stable compaction of the faces that stay (visibility re-evaluated
  // on the same values, as the reference does) ...
  int outp = 0;
  for (int c = 0; c < nchunks; ++c) {
    int f = g.sub + G * c;
    bool live = f < nfaces;
    uint64_t keepm = g.ballot(live && !face_vis_lds(S, live ? f : 0, np));
    bool keep = (keepm >> g.sub) & 1ull;
    int d = outp + __popcll(keepm & g.below());
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
    int k = g.sub + G * kk;
    if (k >= ne) continue;
    uint32_t e = bl[kk];
    uint32_t u = e & 0xffu, w = (e >> 8) & 0xffu;
    V3 vu = lv(S, u), vw = lv(S, w);
    V3 cr = cross(vw - vu, np - vu);
    float l = len(cr);
    V3 n = cr / l;
    int d = kept + k;
    S.nx[d] = n.x;
    S.ny[d] = n.y;
    S.nz[d] = n.z;
    S.dist[d] = face_dist(n, vu, vw, np);
    S.pd[d] = dot(n, vu);
    S.abc[d] = u | (w << 8) | ((uint32_t)nidx << 16);
  }
  nfaces = kept + ne;
  wave_sync();
  return 1;
}

// epa.rs:127-205 for one pair; every lane of the group returns the result.
template <int G>
__device__ __forceinline__ EpaResult grp_epa(const EpaPair<G>& pr, V3 t0, V3 t1, V3 t2, V3 t3, EpaLds& S, const Grp<G>& g) {
  EpaResult R;
  int sub = g.sub;
  if (sub < 4) {
    V3 w = sel(sub == 0, t0, sel(sub == 1, t1, sel(sub == 2, t2, t3)));
    S.v[sub] = make_float4(w.x, w.y, w.z, 0.0f);
  }
  // initial faces (epa.rs:226-257): group lane q builds quad q; survivors
  // keep quad order.  Quads (a, b, c, opposite): (0,1,2,3) (0,3,1,2)
  // (0,2,3,1) (1,3,2,0).
  bool valid = false;
  V3 qn = splat(0.0f);
  float qd = 0.0f, qpd = 0.0f;
  uint32_t qabc = 0;
  if (sub < 4) {
    int a = sub == 3 ? 1 : 0;
    int b = sub == 0 ? 1 : (sub == 2 ? 2 : 3);
    int c = sub == 0 ? 2 : (sub == 1 ? 1 : (sub == 2 ? 3 : 2));
    int opp = 3 - sub;
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
  uint64_t vm = g.ballot(valid);
  int nfaces = __popcll(vm);
  if (valid) {
    int d = __popcll(vm & g.below());
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
    return R;
  }
  int nverts = 4;
  int status = 2;
  bool stop = false;
  for (int k = 0; k < 6; ++k) {  // seeding; a degenerate expansion only ends the seeding
    V3 d = k == 0 ? V3{1.0f, 0.0f, 0.0f} : k == 1 ? V3{-1.0f, 0.0f, 0.0f} : k == 2 ? V3{0.0f, 1.0f, 0.0f}
         : k == 3 ? V3{0.0f, -1.0f, 0.0f} : k == 4 ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 0.0f, -1.0f};
    V3 p = grp_minkowski<G>(pr, d, g);
    int idx = nverts;
    if (sub == 0) S.v[idx] = make_float4(p.x, p.y, p.z, 0.0f);
    ++nverts;
    wave_sync();
    int e = grp_expand<G>(S, nfaces, p, idx, g);
    if (e <= 0) {
      status = e == 0 ? 1 : 4;
      stop = e < 0;
      break;
    }
  }
  if (!stop) {
    for (int it = 0; it < 255; ++it) {
      int bi = grp_best<G>(S, nfaces, g);
      V3 bn{S.nx[bi], S.ny[bi], S.nz[bi]};
      float bd = S.dist[bi];
      V3 np = grp_minkowski<G>(pr, bn, g);
      float wdist = dot(bn, np) - bd;
      if (wdist <= 1e-4f) {
        R.status = 0;
        R.n = bn;
        R.depth = bd;
        return R;
      }
      if (nverts >= BCH_WEPA_V) {
        status = 3;
        break;
      }
      int nidx = nverts;
      if (sub == 0) S.v[nidx] = make_float4(np.x, np.y, np.z, 0.0f);
      ++nverts;
      wave_sync();
      int e = grp_expand<G>(S, nfaces, np, nidx, g);
      if (e <= 0) {
        status = e == 0 ? 1 : 4;
        break;
      }
    }
  }
  int bi = grp_best<G>(S, nfaces, g);
  R.status = status;
  R.n = V3{S.nx[bi], S.ny[bi], S.nz[bi]};
  R.depth = S.dist[bi];
  return R;
}

// Grid-stride over the EPA queue: one wavefront per block, 64 / G pairs per
// wavefront, one LDS polytope per group.
template <int G>
__global__ __launch_bounds__(64) void k_epa_grp(const uint32_t* __restrict__ n_queue, const float4* __restrict__ epa_q,
                                               const uint32_t* __restrict__ pairs, const BodyRec* __restrict__ bodies,
                                               const float4* __restrict__ hull, uint32_t* __restrict__ ccount,
                                               float4* __restrict__ cslots, Counters* cnt) {
  constexpr int NG = 64 / G;
  __shared__ EpaLds lds[NG];
  int lane = threadIdx.x & 63;
  int grp = lane / G;
  Grp<G> g{lane % G, grp * G};
  EpaLds& S = lds[grp];
  uint32_t nq = *n_queue;
  for (uint32_t q0 = blockIdx.x * NG; q0 < nq; q0 += gridDim.x * NG) {
    uint32_t q = q0 + grp;
    if (q < nq) {
      float4 a = epa_q[4 * (uint64_t)q], b = epa_q[4 * (uint64_t)q + 1], c = epa_q[4 * (uint64_t)q + 2],
             d = epa_q[4 * (uint64_t)q + 3];
      uint32_t p = f2u(a.w);
      uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
      BodyRec A = bodies[i], B = bodies[j];
      EpaPair<G> ep;
      Pair& pr = ep.pr;
      pr.a = shape_of(A, hull);
      pr.b = shape_of(B, hull);
      pr.pa = pos_of(A);
      pr.pb = pos_of(B);
      pr.qa = quat_of(A);
      pr.qb = quat_of(B);
      ep.ha.load(pr.a, g);
      ep.hb.load(pr.b, g);
      EpaResult e = grp_epa<G>(ep, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, V3{d.x, d.y, d.z}, S, g);
      {
        // convex_convex.rs:117-154
        bool hit = !(e.depth <= 0.0f);
        V3 point_a = splat(0.0f);
        float depth = 0.0f;
        if (hit) {
          depth = (e.depth + pr.a.margin) + pr.b.margin;
          point_a = grp_support_world<G>(pr.a, ep.ha, pr.pa, pr.qa, e.n, g, false);
        }
        if (g.sub == 0) {
          if (e.status == 4) atomicOr(&cnt->errors, 2u);
          ccount[p] = hit ? 1u : 0u;
          if (hit) {
            V3 pos = point_a - e.n * depth;
            cslots[(uint64_t)p * 8] = make_float4(pos.x, pos.y, pos.z, depth);
            cslots[(uint64_t)p * 8 + 1] = make_float4(e.n.x, e.n.y, e.n.z, 0.0f);
          }
        }
      }
    }
    wave_sync();
  }
}

}  // namespace bch
