// bch_epa_wave.hpp -- EPA with one wavefront per pair, polytope in LDS.
//
// Same algorithm, statement order and tie rules as the reference
// (narrowphase/epa.rs:127-398): tetrahedron seed + six axis supports,
// closest face by edge-aware distance (first minimum), visibility
// n.p - d > 1e-5, transactional two-phase expansion, <= 255 iterations,
// <= 128 vertices.  The 64 lanes share the work that has no order:
// face scans (closest face, visibility), hull support scans (largest
// projection, lowest index = the sequential first maximum), face
// compaction (ballot prefix counts keep the survivors' order) and the new
// faces of the horizon fan.  The horizon itself is stitched by lane 0 in
// visible-face order with the reference's swap-remove edge cancellation,
// so faces are created in exactly the reference order.
#pragma once
#include "bch_narrow.hpp"

namespace bch {

#define BCH_WEPA_V 128
#define BCH_WEPA_F 256
#define BCH_WEPA_E 384

struct EpaLds {
  float4 v[BCH_WEPA_V];
  float nx[BCH_WEPA_F], ny[BCH_WEPA_F], nz[BCH_WEPA_F], dist[BCH_WEPA_F], pd[BCH_WEPA_F];
  uint32_t abc[BCH_WEPA_F];
  uint32_t edges[BCH_WEPA_E];
  int misc[4];
};

// Order LDS traffic between lanes of one wavefront (no cross-wave barrier:
// each wave of a block runs its own pair).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane ? ((1ull << lane) - 1ull) : 0ull; }

// First maximiser of p.d over a hull's vertices, across the wavefront.
__device__ __forceinline__ V3 wave_hull_support(const float4* h, int n, V3 d, int lane) {
  float bv = -INFINITY;
  int bi = -1;
  for (int v = lane; v < n; v += 64) {
    float dd = dot(hull_vertex(h, v), d);
    if (dd > bv) {
      bv = dd;
      bi = v;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    float ov = __shfl_xor(bv, off);
    int oi = __shfl_xor(bi, off);
    if (oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi))) {
      bv = ov;
      bi = oi;
    }
  }
  return bi >= 0 ? hull_vertex(h, bi) : splat(0.0f);
}

__device__ __forceinline__ V3 wave_core_support_world(const ShapeRef& s, V3 p, Q4 q, V3 d, int lane) {
  V3 ld = qrot(qconj(q), d);
  V3 l = s.type == HULL ? wave_hull_support(s.hull, s.nhull, ld, lane) : core_support_local(s, ld);
  return qrot(q, l) + p;
}

__device__ __forceinline__ V3 wave_support_world(const ShapeRef& s, V3 p, Q4 q, V3 d, int lane) {
  V3 ld = qrot(qconj(q), d);
  V3 l = s.type == HULL ? wave_hull_support(s.hull, s.nhull, ld, lane) : support_local(s, ld);
  return qrot(q, l) + p;
}

__device__ __forceinline__ V3 wave_minkowski(const Pair& pr, V3 d, int lane) {
  V3 a = wave_core_support_world(pr.a, pr.pa, pr.qa, d, lane);
  V3 b = wave_core_support_world(pr.b, pr.pb, pr.qb, -d, lane);
  return a - b;
}

__device__ __forceinline__ V3 lv(const EpaLds& S, int i) {
  float4 v = S.v[i];
  return V3{v.x, v.y, v.z};
}

// epa.rs:212-224: index of the first face with the smallest |distance|
// (a NaN first face is never replaced).
__device__ __forceinline__ int wave_best(const EpaLds& S, int nfaces, int lane) {
  float bv = INFINITY;
  int bi = -1;
  for (int f = lane; f < nfaces; f += 64) {
    float v = fabsf(S.dist[f]);
    if (!is_nan(v) && (bi < 0 || v < bv)) {
      bv = v;
      bi = f;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    float ov = __shfl_xor(bv, off);
    int oi = __shfl_xor(bi, off);
    if (oi >= 0 && (bi < 0 || ov < bv || (ov == bv && oi < bi))) {
      bv = ov;
      bi = oi;
    }
  }
  if (bi < 0 || is_nan(S.dist[0])) bi = 0;
  return bi;
}

// epa.rs:309-388.  Returns 1 ok (or nothing visible), 0 degenerate horizon
// (hull untouched), -1 LDS capacity exceeded.
__device__ int wave_expand(EpaLds& S, int& nfaces, V3 np, int nidx, int lane) {
  uint64_t vis[4];
  bool any = false;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int f = lane + 64 * c;
    bool v = false;
    if (f < nfaces) v = dot(V3{S.nx[f], S.ny[f], S.nz[f]}, np) - S.pd[f] > 1e-5f;
    vis[c] = __ballot(v);
    any = any || vis[c] != 0ull;
  }
  if (!any) return 1;
  // Horizon, lane 0, in visible-face order with swap-remove cancellation.
  if (lane == 0) {
    int ne = 0;
    bool over = false;
    for (int c = 0; c < 4; ++c) {
      uint64_t m = vis[c];
      while (m) {
        int f = 64 * c + __builtin_ctzll(m);
        m &= m - 1ull;
        uint32_t t = S.abc[f];
        uint32_t fa = t & 0xffu, fb = (t >> 8) & 0xffu, fc = (t >> 16) & 0xffu;
        for (int e = 0; e < 3; ++e) {
          uint32_t u = e == 0 ? fa : (e == 1 ? fb : fc);
          uint32_t v = e == 0 ? fb : (e == 1 ? fc : fa);
          uint32_t rev = v | (u << 8);
          int found = -1;
          for (int k = 0; k < ne; ++k)
            if (S.edges[k] == rev) {
              found = k;
              break;
            }
          if (found >= 0) {
            S.edges[found] = S.edges[ne - 1];
            --ne;
          } else if (ne < BCH_WEPA_E) {
            S.edges[ne++] = u | (v << 8);
          } else {
            over = true;
          }
        }
      }
    }
    S.misc[0] = over ? -1 : ne;
  }
  wave_sync();
  int ne = S.misc[0];
  if (ne < 0) return -1;
  // Validate every horizon edge before touching the hull.
  bool bad = false;
  for (int k = lane; k < ne; k += 64) {
    uint32_t e = S.edges[k];
    V3 vu = lv(S, e & 0xffu);
    if (len(cross(lv(S, (e >> 8) & 0xffu) - vu, np - vu)) < 1e-12f) bad = true;
  }
  if (__ballot(bad)) return 0;
  // Commit: stable compaction of the faces that stay ...
  float fnx[4], fny[4], fnz[4], fd[4], fpd[4];
  uint32_t fabc[4];
  int dst[4];
  int kept = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int f = lane + 64 * c;
    bool live = f < nfaces;
    uint64_t keepm = __ballot(live && !((vis[c] >> lane) & 1ull));
    bool keep = (keepm >> lane) & 1ull;
    dst[c] = keep ? kept + __popcll(keepm & lanes_below(lane)) : -1;
    kept += __popcll(keepm);
    if (keep) {
      fnx[c] = S.nx[f];
      fny[c] = S.ny[f];
      fnz[c] = S.nz[f];
      fd[c] = S.dist[f];
      fpd[c] = S.pd[f];
      fabc[c] = S.abc[f];
    }
  }
  if (kept + ne > BCH_WEPA_F) return -1;
  wave_sync();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    int d = dst[c];
    if (d >= 0) {
      S.nx[d] = fnx[c];
      S.ny[d] = fny[c];
      S.nz[d] = fnz[c];
      S.dist[d] = fd[c];
      S.pd[d] = fpd[c];
      S.abc[d] = fabc[c];
    }
  }
  // ... then the horizon fan, in horizon order.
  for (int k = lane; k < ne; k += 64) {
    uint32_t e = S.edges[k];
    uint32_t u = e & 0xffu, v = (e >> 8) & 0xffu;
    V3 vu = lv(S, u), vv = lv(S, v);
    V3 cr = cross(vv - vu, np - vu);
    float l = len(cr);
    V3 n = cr / l;
    int d = kept + k;
    S.nx[d] = n.x;
    S.ny[d] = n.y;
    S.nz[d] = n.z;
    S.dist[d] = face_dist(n, vu, vv, np);
    S.pd[d] = dot(n, vu);
    S.abc[d] = u | (v << 8) | ((uint32_t)nidx << 16);
  }
  nfaces = kept + ne;
  wave_sync();
  return 1;
}

// epa.rs:127-205 for one pair; every lane returns the same result.
__device__ EpaResult wave_epa(const Pair& pr, V3 t0, V3 t1, V3 t2, V3 t3, EpaLds& S, int lane) {
  EpaResult R;
  V3 s[4] = {t0, t1, t2, t3};
  if (lane < 4) {
    V3 w = lane == 0 ? t0 : (lane == 1 ? t1 : (lane == 2 ? t2 : t3));
    S.v[lane] = make_float4(w.x, w.y, w.z, 0.0f);
  }
  // initial faces (epa.rs:226-257): lane q builds quad q; survivors keep quad order
  bool valid = false;
  V3 qn = splat(0.0f);
  float qd = 0.0f, qpd = 0.0f;
  uint32_t qabc = 0;
  if (lane < 4) {
    const int A[4] = {0, 0, 0, 1}, B[4] = {1, 3, 2, 3}, Cc[4] = {2, 1, 3, 2}, O[4] = {3, 2, 1, 0};
    int a = A[lane], b = B[lane], c = Cc[lane], opp = O[lane];
    V3 sa = s[a], sb = s[b], sc = s[c], so = s[opp];
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
    return R;
  }
  int nverts = 4;
  int status = 2;
  bool stop = false;
  for (int k = 0; k < 6; ++k) {  // seeding; a degenerate expansion only ends the seeding
    V3 d = k == 0 ? V3{1.0f, 0.0f, 0.0f} : k == 1 ? V3{-1.0f, 0.0f, 0.0f} : k == 2 ? V3{0.0f, 1.0f, 0.0f}
         : k == 3 ? V3{0.0f, -1.0f, 0.0f} : k == 4 ? V3{0.0f, 0.0f, 1.0f} : V3{0.0f, 0.0f, -1.0f};
    V3 p = wave_minkowski(pr, d, lane);
    int idx = nverts;
    if (lane == 0) S.v[idx] = make_float4(p.x, p.y, p.z, 0.0f);
    ++nverts;
    wave_sync();
    int e = wave_expand(S, nfaces, p, idx, lane);
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
      V3 np = wave_minkowski(pr, bn, lane);
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
      if (lane == 0) S.v[nidx] = make_float4(np.x, np.y, np.z, 0.0f);
      ++nverts;
      wave_sync();
      int e = wave_expand(S, nfaces, np, nidx, lane);
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
  return R;
}

// Grid-stride over the EPA queue, one pair per wavefront.
__global__ __launch_bounds__(256) void k_epa_wave(const Counters* __restrict__ cnt_ro, const float4* __restrict__ epa_q,
                                                  const uint32_t* __restrict__ pairs, const BodyRec* __restrict__ bodies,
                                                  const float4* __restrict__ hull, uint32_t* __restrict__ ccount,
                                                  float4* __restrict__ cslots, Counters* cnt) {
  __shared__ EpaLds lds[4];
  int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  EpaLds& S = lds[wid];
  uint32_t nq = cnt_ro->n_epa;
  for (uint32_t q = blockIdx.x * 4 + wid; q < nq; q += gridDim.x * 4) {
    float4 a = epa_q[4 * (uint64_t)q], b = epa_q[4 * (uint64_t)q + 1], c = epa_q[4 * (uint64_t)q + 2],
           d = epa_q[4 * (uint64_t)q + 3];
    uint32_t p = f2u(a.w);
    uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    BodyRec A = bodies[i], B = bodies[j];
    Pair pr;
    pr.a = shape_of(A, hull);
    pr.b = shape_of(B, hull);
    pr.pa = pos_of(A);
    pr.pb = pos_of(B);
    pr.qa = quat_of(A);
    pr.qb = quat_of(B);
    EpaResult e = wave_epa(pr, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, V3{d.x, d.y, d.z}, S, lane);
    // convex_convex.rs:117-154
    bool hit = !(e.depth <= 0.0f);
    V3 point_a = splat(0.0f);
    float depth = 0.0f;
    if (hit) {
      depth = (e.depth + pr.a.margin) + pr.b.margin;
      point_a = wave_support_world(pr.a, pr.pa, pr.qa, e.n, lane);
    }
    if (lane == 0) {
      if (e.status == 4) atomicOr(&cnt->errors, 2u);
      ccount[p] = hit ? 1u : 0u;
      if (hit) {
        V3 pos = point_a - e.n * depth;
        cslots[(uint64_t)p * 8] = make_float4(pos.x, pos.y, pos.z, depth);
        cslots[(uint64_t)p * 8 + 1] = make_float4(e.n.x, e.n.y, e.n.z, 0.0f);
      }
    }
    wave_sync();
  }
}

}  // namespace bch
