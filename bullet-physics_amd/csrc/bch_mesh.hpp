// bch_mesh.hpp -- convex shapes against static triangle meshes
// (convex_mesh, crates/bullet-collision/src/narrowphase/convex_mesh.rs:24-272,
// over the TriangleMesh BVH of shapes/bvh.rs and shapes/triangle_mesh.rs).
//
// Per (convex body, mesh) pair:
//   k_mesh_count   the convex's world AABB corners mapped into the mesh frame
//                  give the query box; a depth-first walk of the mesh BVH
//                  (right child first, as the reference's stack pops it)
//                  counts the candidate triangles whose own AABB overlaps
//   scan + k_mesh_emit   the (pair, triangle) sub-pairs in candidate order
//   k_mesh_detect  per sub-pair: the analytic sphere-triangle contact for a
//                  sphere, else the margin-core GJK walk against the triangle
//                  as a three-vertex hull (EPA queued for overlapping cores)
//   k_epa_mesh     EPA of the queued sub-pairs (bch_epa_wave.hpp)
//   k_mesh_reduce  the pair's points in candidate order, reduced to four as
//                  ContactManifold::reduce does (manifold.rs:76-110)
//
// Mesh buffers: nodes (lo, hi) with lo.w = left child or first slot, hi.w =
// right child or slot count | kMeshLeaf; the BVH's triangle order; the
// triangles as three float4 vertices; their AABBs (lo, hi).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bch {

constexpr uint32_t kMeshLeaf = 0x80000000u;
constexpr int kMeshStack = 64;

struct MeshRec {
  uint32_t node_off, order_off, tri_off, ntri;
  float threshold;  // shapes/mod.rs:388-392 over the mesh's local bounds
  uint32_t pad[3];
};

// Query box of a convex body in the mesh frame (convex_mesh.rs:37-55).
__device__ __forceinline__ Box6 mesh_query_box(const BodyRec& B, const float4* hull, V3 pm, Q4 qm) {
  ShapeRef s = shape_of(B, hull);
  Box6 w = world_aabb(s, pos_of(B), quat_of(B));
  Q4 inv = qconj(qm);
  V3 c = (w.lo + w.hi) * 0.5f, e = (w.hi - w.lo) * 0.5f;
  V3 lo = splat(INFINITY), hi = splat(-INFINITY);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    V3 corner{(i & 1) == 0 ? c.x + e.x : c.x - e.x, (i & 2) == 0 ? c.y + e.y : c.y - e.y,
              (i & 4) == 0 ? c.z + e.z : c.z - e.z};
    V3 p = qrot(inv, corner - pm);
    lo = vmin(lo, p);
    hi = vmax(hi, p);
  }
  return Box6{lo, hi};
}

// triangles_in_aabb (bvh.rs:58-79) filtered by the triangle's own AABB
// (convex_mesh.rs:96-104) unless tri_aabb is null: f(t) for each candidate,
// in the reference's order. Returns false when the stack would overflow (the walk stops).
template <class F>
__device__ __forceinline__ bool mesh_walk(const MeshRec& M, const float4* __restrict__ nodes,
                                          const uint32_t* __restrict__ order, const float4* __restrict__ tri_aabb,
                                          const Box6& q, F f) {
  uint32_t stack[kMeshStack];
  int sp = 0;
  stack[sp++] = 0u;
  while (sp > 0) {
    uint32_t k = stack[--sp];
    float4 lo = nodes[2 * (uint64_t)(M.node_off + k)], hi = nodes[2 * (uint64_t)(M.node_off + k) + 1];
    Box6 nb{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}};
    if (!overlaps(nb, q)) continue;
    uint32_t a = f2u(lo.w), b = f2u(hi.w);
    if (b & kMeshLeaf) {
      uint32_t cnt = b & ~kMeshLeaf;
      for (uint32_t i = a; i < a + cnt; ++i) {
        uint32_t t = order[M.order_off + i];
        if (!tri_aabb) {  // the sweeps walk without the triangle filter
          f(t);
          continue;
        }
        float4 tl = tri_aabb[2 * (uint64_t)(M.tri_off + t)], th = tri_aabb[2 * (uint64_t)(M.tri_off + t) + 1];
        Box6 tb{V3{tl.x, tl.y, tl.z}, V3{th.x, th.y, th.z}};
        if (overlaps(tb, q)) f(t);
      }
    } else {
      if (sp + 2 > kMeshStack) return false;
      stack[sp++] = a;
      stack[sp++] = b;
    }
  }
  return true;
}

__global__ void k_mesh_count(uint32_t P, const uint2* __restrict__ pairs, const float* __restrict__ poses,
                             const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                             const MeshRec* __restrict__ meshes, const float4* __restrict__ nodes,
                             const uint32_t* __restrict__ order, const float4* __restrict__ tri_aabb,
                             uint32_t* __restrict__ counts, Counters* cnt) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint2 pr = pairs[p];
  MeshRec M = meshes[pr.y];
  const float* ps = poses + 7 * (uint64_t)p;
  Box6 q = mesh_query_box(bodies[pr.x], hull, V3{ps[0], ps[1], ps[2]}, Q4{ps[3], ps[4], ps[5], ps[6]});
  uint32_t c = 0;
  if (!mesh_walk(M, nodes, order, tri_aabb, q, [&](uint32_t) { ++c; })) atomicOr(&cnt->errors, 32u);
  counts[p] = c;
}

__global__ void k_mesh_emit(uint32_t P, const uint2* __restrict__ pairs, const float* __restrict__ poses,
                            const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                            const MeshRec* __restrict__ meshes, const float4* __restrict__ nodes,
                            const uint32_t* __restrict__ order, const float4* __restrict__ tri_aabb,
                            const uint32_t* __restrict__ offsets, uint2* __restrict__ sub) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint2 pr = pairs[p];
  MeshRec M = meshes[pr.y];
  const float* ps = poses + 7 * (uint64_t)p;
  Box6 q = mesh_query_box(bodies[pr.x], hull, V3{ps[0], ps[1], ps[2]}, Q4{ps[3], ps[4], ps[5], ps[6]});
  uint32_t o = offsets[p], end = offsets[p + 1];
  (void)mesh_walk(M, nodes, order, tri_aabb, q, [&](uint32_t t) {
    if (o < end) sub[o++] = make_uint2(p, t);
  });
}

// segment_closest_point_sq (convex_mesh.rs:237-255)
__device__ __forceinline__ float seg_closest_sq(V3 from, V3 to, V3 p, V3& nearest) {
  V3 v = to - from;
  float t = dot(v, p - from);
  if (t > 0.0f) {
    float dvv = dot(v, v);
    if (t < dvv) t = t / dvv;
    else t = 1.0f;
  } else {
    t = 0.0f;
  }
  nearest = from + v * t;
  return len2(p - nearest);
}

// point_in_triangle (convex_mesh.rs:257-272)
__device__ __forceinline__ bool point_in_tri(V3 p, V3 v1, V3 v2, V3 v3, V3 n) {
  float r1 = dot(cross(v2 - v1, n), p - v1);
  float r2 = dot(cross(v3 - v2, n), p - v2);
  float r3 = dot(cross(v1 - v3, n), p - v3);
  return (r1 > 0.0f && r2 > 0.0f && r3 > 0.0f) || (r1 <= 0.0f && r2 <= 0.0f && r3 <= 0.0f);
}

// sphere_triangle_contact (convex_mesh.rs:168-235), mesh frame.
__device__ __forceinline__ bool sphere_tri(V3 center, float radius, float threshold, V3 a, V3 b, V3 c, V3& point,
                                           V3& nout, float& pen) {
  const float eps = 1.1920929e-7f;  // f32::EPSILON
  float rt = radius + threshold;
  V3 normal = cross(b - a, c - a);
  float l2 = len2(normal);
  if (l2 < eps * eps) return false;
  normal = normal * (1.0f / sqrtf(l2));
  float dist = dot(normal, center - a);
  if (dist < 0.0f) {
    dist = -dist;
    normal = -normal;
  }
  if (dist >= rt) return false;
  V3 cp = splat(0.0f);
  if (point_in_tri(center, a, b, c, normal)) {
    cp = center - normal * dist;
  } else {
    float min_d2 = rt * rt;
    bool found = false;
    V3 on;
    float d2 = seg_closest_sq(a, b, center, on);
    if (d2 < min_d2) { min_d2 = d2; cp = on; found = true; }
    d2 = seg_closest_sq(b, c, center, on);
    if (d2 < min_d2) { min_d2 = d2; cp = on; found = true; }
    d2 = seg_closest_sq(c, a, center, on);
    if (d2 < min_d2) { min_d2 = d2; cp = on; found = true; }
    if (!found) return false;
  }
  V3 tc = center - cp;
  float d2 = len2(tc);
  if (d2 >= rt * rt) return false;
  point = cp;
  if (d2 > eps) {
    float d = sqrtf(d2);
    nout = tc * (1.0f / d);
    pen = radius - d;
  } else {
    nout = normal;
    pen = radius;
  }
  return true;
}

// The pair of sub-pair (p, t): the convex as A, the triangle hull as B.
__device__ __forceinline__ Pair mesh_sub_pair(const BodyRec& Bc, const float4* hull, const MeshRec& M,
                                              const float4* tris, uint32_t t, const float* ps) {
  Pair pr;
  pr.a = shape_of(Bc, hull);
  pr.pa = pos_of(Bc);
  pr.qa = quat_of(Bc);
  ShapeRef tri;
  tri.type = HULL;
  tri.a0 = tri.a1 = tri.a2 = tri.a3 = 0.0f;
  tri.hull = tris + 3 * (uint64_t)(M.tri_off + t);
  tri.nhull = 3;
  tri.margin = 0.0f;  // shapes/mod.rs:243 (convex_hull.rs:43-45)
  pr.b = tri;
  pr.pb = V3{ps[0], ps[1], ps[2]};
  pr.qb = Q4{ps[3], ps[4], ps[5], ps[6]};
  return pr;
}

// One sub-pair: the analytic sphere path (convex_mesh.rs:62-95) or
// convex_convex (convex_mesh.rs:96-137) with the pair's seed and band.
__global__ void k_mesh_detect(uint32_t S, const uint2* __restrict__ sub, const uint2* __restrict__ pairs,
                              const float* __restrict__ poses, const float* __restrict__ seeds,
                              const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                              const MeshRec* __restrict__ meshes, const float4* __restrict__ tris,
                              uint32_t* __restrict__ sub_cnt, float4* __restrict__ sub_slots,
                              float4* __restrict__ epa_q, Counters* cnt) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  bool need_epa = false;
  GjkResult g;
  if (k < S) {
    uint2 st = sub[k];
    uint32_t p = st.x, t = st.y;
    uint2 pr_ = pairs[p];
    BodyRec Bc = bodies[pr_.x];
    MeshRec M = meshes[pr_.y];
    const float* ps = poses + 7 * (uint64_t)p;
    V3 pm{ps[0], ps[1], ps[2]};
    Q4 qm{ps[3], ps[4], ps[5], ps[6]};
    float thr = mn(Bc.threshold, M.threshold);  // narrowphase/mod.rs:52-57
    uint32_t n = 0;
    float4 s0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), s1 = s0;
    if ((Bc.type & 0xffu) == SPHERE) {
      Q4 inv = qconj(qm);
      V3 center = qrot(inv, pos_of(Bc) - pm);
      const float4* tv = tris + 3 * (uint64_t)(M.tri_off + t);
      V3 a = hull_vertex(tv, 0), b = hull_vertex(tv, 1), c = hull_vertex(tv, 2);
      V3 pl, nl;
      float pen;
      if (sphere_tri(center, Bc.a0, thr, a, b, c, pl, nl, pen)) {
        V3 pw = pm + qrot(qm, pl);
        V3 nw = qrot(qm, -nl);
        s0 = make_float4(pw.x, pw.y, pw.z, pen);
        s1 = make_float4(nw.x, nw.y, nw.z, 0.0f);
        n = 1;
      }
    } else {
      Pair pr = mesh_sub_pair(Bc, hull, M, tris, t, ps);
      bool has_seed = seeds != nullptr && !is_nan(seeds[3 * (uint64_t)p]);
      V3 sd = has_seed ? V3{seeds[3 * (uint64_t)p], seeds[3 * (uint64_t)p + 1], seeds[3 * (uint64_t)p + 2]}
                       : splat(0.0f);
      Manifold4 m;
      need_epa = collide_pair(pr, thr, has_seed, sd, m, g, CLS_GJK) == 1;
      if (m.n > 0) {
        s0 = make_float4(m.c[0].pos.x, m.c[0].pos.y, m.c[0].pos.z, m.c[0].depth);
        s1 = make_float4(m.c[0].n.x, m.c[0].n.y, m.c[0].n.z, 0.0f);
        n = 1;
      }
    }
    sub_cnt[k] = n;
    sub_slots[2 * (uint64_t)k] = s0;
    sub_slots[2 * (uint64_t)k + 1] = s1;
  }
  uint64_t mask = __ballot(need_epa);
  if (mask) {
    uint32_t lane = threadIdx.x & 63;
    uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t qb = 0;
    if (lane == leader) qb = atomicAdd(&cnt->n_epa, (uint32_t)__popcll(mask));
    qb = __shfl(qb, leader);
    if (need_epa) {
      uint32_t q = qb + __popcll(mask & ((1ull << lane) - 1ull));
      epa_q[4 * (uint64_t)q] = make_float4(g.t0.x, g.t0.y, g.t0.z, u2f(k));
      epa_q[4 * (uint64_t)q + 1] = make_float4(g.t1.x, g.t1.y, g.t1.z, 0.0f);
      epa_q[4 * (uint64_t)q + 2] = make_float4(g.t2.x, g.t2.y, g.t2.z, 0.0f);
      epa_q[4 * (uint64_t)q + 3] = make_float4(g.t3.x, g.t3.y, g.t3.z, 0.0f);
    }
  }
}

// EPA pair source for mesh sub-pairs (entry index = sub-pair index).
struct MeshEpaSrc {
  const uint2* sub;
  const uint2* pairs;
  const float* poses;
  const BodyRec* bodies;
  const float4* hull;
  const MeshRec* meshes;
  const float4* tris;
  __device__ __forceinline__ Pair pair(uint32_t k) const {
    uint2 st = sub[k];
    uint2 pr = pairs[st.x];
    return mesh_sub_pair(bodies[pr.x], hull, meshes[pr.y], tris, st.y, poses + 7 * (uint64_t)st.x);
  }
};

__global__ __launch_bounds__(64) void k_epa_mesh(const uint32_t* __restrict__ n_queue,
                                                 const float4* __restrict__ epa_q, MeshEpaSrc src,
                                                 uint32_t* __restrict__ sub_cnt, float4* __restrict__ sub_slots,
                                                 Counters* cnt, float4* __restrict__ over_q, int flimit) {
  __shared__ EpaLds S;
  epa_queue_run(n_queue, epa_q, src, sub_cnt, sub_slots, 2u, (uint32_t*)nullptr, cnt, S, over_q, flimit);
}

// k_epa_mesh's overflow queue with the 2048-face store (as k_epa_big).
__global__ __launch_bounds__(64) void k_epa_mesh_big(const uint32_t* __restrict__ n_over,
                                                     const float4* __restrict__ over_q, MeshEpaSrc src,
                                                     uint32_t* __restrict__ sub_cnt, float4* __restrict__ sub_slots,
                                                     Counters* cnt) {
  __shared__ EpaBig S;
  epa_queue_run(n_over, over_q, src, sub_cnt, sub_slots, 2u, (uint32_t*)nullptr, cnt, S);
}

// The same queue two sub-pairs per wavefront (bch_epa_dual.hpp); entries
// outside its fast path go to defer_q for k_epa_mesh.
__global__ __launch_bounds__(64) void k_epa_mesh_dual(const uint32_t* __restrict__ n_queue,
                                                      const float4* __restrict__ epa_q, MeshEpaSrc src,
                                                      uint32_t* __restrict__ sub_cnt, float4* __restrict__ sub_slots,
                                                      float4* __restrict__ defer_q, Counters* cnt) {
  __shared__ EpaHalf S[2];
  epa_dual_run(n_queue, epa_q, src, sub_cnt, sub_slots, 2u, (uint32_t*)nullptr, defer_q, &cnt->n_epa_defer, cnt, S);
}

// ContactManifold::reduce (manifold.rs:76-110) over the pair's sub-pair
// points in candidate order, without sorting: the stable sort by depth makes
// ties fall to the earlier candidate, which the comparisons below encode.
__device__ __forceinline__ int64_t pen_key(float f) {
  uint32_t u = f2u(f);
  return (u & 0x80000000u) ? -(int64_t)(u & 0x7fffffffu) - 1 : (int64_t)u;
}

__global__ void k_mesh_reduce(uint32_t P, const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ sub_cnt,
                              const float4* __restrict__ sub_slots, const uint8_t* __restrict__ mesh_first,
                              uint8_t* __restrict__ out_count, float* __restrict__ out_pts) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint32_t b = offsets[p], e = offsets[p + 1];
  uint32_t m = 0;
  for (uint32_t k = b; k < e; ++k) m += sub_cnt[k];
  uint32_t pick[4];
  int np = 0;
  if (m <= 4) {  // reduce leaves small manifolds alone
    for (uint32_t k = b; k < e && np < 4; ++k)
      if (sub_cnt[k]) pick[np++] = k;
  } else {
    // deepest (first in the stable descending sort)
    uint32_t best = 0xffffffffu;
    int64_t bk = 0;
    for (uint32_t k = b; k < e; ++k) {
      if (!sub_cnt[k]) continue;
      int64_t kk = pen_key(sub_slots[2 * (uint64_t)k].w);
      if (best == 0xffffffffu || kk > bk) {
        best = k;
        bk = kk;
      }
    }
    pick[np++] = best;
    while (np < 4) {
      uint32_t bi = 0xffffffffu;
      float bsep = -INFINITY;
      int64_t bpk = 0;
      for (uint32_t k = b; k < e; ++k) {
        if (!sub_cnt[k]) continue;
        bool used = false;
        for (int j = 0; j < np; ++j) used = used || pick[j] == k;
        if (used) continue;
        float4 pk = sub_slots[2 * (uint64_t)k];
        float sep = INFINITY;
        for (int j = 0; j < np; ++j) {
          float4 q = sub_slots[2 * (uint64_t)pick[j]];
          V3 d{pk.x - q.x, pk.y - q.y, pk.z - q.z};
          sep = fminf(sep, len2(d));
        }
        int64_t kk = pen_key(pk.w);
        // strict > in sorted order: a tie goes to the deeper, then the earlier candidate
        bool better = bi == 0xffffffffu ? sep > bsep : (sep > bsep || (sep == bsep && kk > bpk));
        if (better) {
          bi = k;
          bsep = sep;
          bpk = kk;
        }
      }
      if (bi == 0xffffffffu) break;
      pick[np++] = bi;
    }
  }
  bool flip = mesh_first && mesh_first[p];
  float* o = out_pts + 28 * (uint64_t)p;
  for (int j = 0; j < 4; ++j) {
    if (j < np) {
      float4 a = sub_slots[2 * (uint64_t)pick[j]], n = sub_slots[2 * (uint64_t)pick[j] + 1];
      o[7 * j] = a.x;
      o[7 * j + 1] = a.y;
      o[7 * j + 2] = a.z;
      o[7 * j + 3] = flip ? -n.x : n.x;
      o[7 * j + 4] = flip ? -n.y : n.y;
      o[7 * j + 5] = flip ? -n.z : n.z;
      o[7 * j + 6] = a.w;
    } else {
      for (int c = 0; c < 7; ++c) o[7 * j + c] = 0.0f;
    }
  }
  out_count[p] = (uint8_t)np;
}

}  // namespace bch
