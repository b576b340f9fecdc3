// bch_kernels.hpp -- HIP kernels of the collision pass (gfx950).
//
// Pass structure (one stream, no host work between stages except one
// count read-back before the pair sort):
//   k_aabb        body record -> world AABB + margin (fused), scene bounds,
//                 big / NaN classification
//   k_morton      30-bit Morton code of the AABB centre (world id above it)
//   radix sort    (rocPRIM) codes -> leaf order
//   k_leaves      leaf nodes: box, body id, rope
//   k_build       Karras (2012) internal nodes: left child, rope, parent
//   k_refit       bottom-up boxes (second arriving child computes the union)
//   k_traverse    one lane per leaf walks the rope-linked tree over the
//                 leaves to its right: exact pair set without duplicates
//   k_big         bodies with unbounded AABBs (planes) against their world
//   radix sort    (rocPRIM) pair keys i << 32 | j -> lexicographic order
//   k_narrow      one lane per pair (pairs grouped by shape class)
//   k_epa         one lane per deep-overlap pair, polytope in scratch
//   k_gather      per-pair contact slots -> dense CSR contacts
#pragma once
#include "bch_narrow.hpp"

namespace bch {

// Traversal node (32 B = two 16-byte loads): box, left child (>= 0) or
// ~body (< 0) for a leaf, rope = next node in depth-first order after this
// subtree (-1 = end).  The right child of an internal node is its left
// child's rope.
struct alignas(16) Node {
  float lx, ly, lz;
  int left;
  float hx, hy, hz;
  int rope;
};

struct Counters {
  uint32_t n_reg;       // bodies in the tree
  uint32_t n_big;       // bodies with unbounded AABBs
  uint32_t bounds[6];   // orderable-encoded min xyz, max xyz of regular centres
  uint32_t max_leaf;    // largest per-leaf pair count seen
  uint32_t n_big_pairs;
  uint32_t n_epa;
  uint32_t errors;      // bit 0 clip overflow, bit 1 epa store overflow, bit 2 big-pair overflow
  uint32_t pad[4];
};

__device__ __forceinline__ uint32_t ord_enc(float f) {
  uint32_t u = f2u(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord_dec(uint32_t u) {
  return u2f((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Body flags.
enum : uint8_t { FLAG_REG = 0, FLAG_BIG = 1, FLAG_NAN = 2 };
// An AABB this wide on any axis (or non-finite) is tested brute force
// against its world instead of entering the tree (planes: +-1e30).
#define BCH_BIG_EXTENT 1.0e6f

__global__ void k_reset(Counters* c) {
  if (threadIdx.x == 0) {
    c->n_reg = 0;
    c->n_big = 0;
    c->bounds[0] = c->bounds[1] = c->bounds[2] = 0xffffffffu;
    c->bounds[3] = c->bounds[4] = c->bounds[5] = 0u;
    c->max_leaf = 0;
    c->n_big_pairs = 0;
    c->n_epa = 0;
    c->errors = 0;
  }
}

// Pack caller SoA arrays into body records; shape constants (margin,
// contact_breaking_threshold, normalised plane normal) are computed here.
__global__ void k_pack_bodies(uint32_t n, const float* __restrict__ pos, const float* __restrict__ quat,
                              const uint8_t* __restrict__ type, const float* __restrict__ params,
                              const uint32_t* __restrict__ hull_first, const uint32_t* __restrict__ hull_count,
                              const float4* __restrict__ hull, const uint32_t* __restrict__ world_of,
                              BodyRec* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  BodyRec b;
  b.px = pos[3 * i];
  b.py = pos[3 * i + 1];
  b.pz = pos[3 * i + 2];
  b.type = type[i];
  b.qx = quat[4 * i];
  b.qy = quat[4 * i + 1];
  b.qz = quat[4 * i + 2];
  b.qw = quat[4 * i + 3];
  b.a0 = params[4 * i];
  b.a1 = params[4 * i + 1];
  b.a2 = params[4 * i + 2];
  b.a3 = params[4 * i + 3];
  if (b.type == PLANE) {  // shapes/plane.rs:18-23
    V3 nn = normalize(V3{b.a0, b.a1, b.a2});
    b.a0 = nn.x;
    b.a1 = nn.y;
    b.a2 = nn.z;
  }
  if (b.type == HULL) {
    b.a0 = u2f(hull_first[i]);
    b.a1 = u2f(hull_count[i]);
  }
  b.margin = shape_margin((int)b.type, b.a0);
  ShapeRef s = shape_of(b, hull);
  b.threshold = breaking_threshold(s);
  b.world = world_of ? world_of[i] : 0u;
  b.pad = 0u;
  out[i] = b;
}

__global__ void k_pack_poses(uint32_t n, const float* __restrict__ pos, const float* __restrict__ quat,
                             BodyRec* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i].px = pos[3 * i];
  out[i].py = pos[3 * i + 1];
  out[i].pz = pos[3 * i + 2];
  out[i].qx = quat[4 * i];
  out[i].qy = quat[4 * i + 1];
  out[i].qz = quat[4 * i + 2];
  out[i].qw = quat[4 * i + 3];
}

__global__ void k_xyz_to_float4(uint32_t n, const float* __restrict__ in, float4* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float4(in[3 * i], in[3 * i + 1], in[3 * i + 2], 0.0f);
}

// Classify an AABB and fold its centre into the scene bounds.
__device__ __forceinline__ void classify_box(uint32_t i, const Box6& bx, uint8_t* flags, Counters* cnt,
                                             uint32_t* big_ids, float* cmin, float* cmax) {
  bool nan = is_nan(bx.lo.x) || is_nan(bx.lo.y) || is_nan(bx.lo.z) || is_nan(bx.hi.x) || is_nan(bx.hi.y) ||
             is_nan(bx.hi.z);
  float wx = bx.hi.x - bx.lo.x, wy = bx.hi.y - bx.lo.y, wz = bx.hi.z - bx.lo.z;
  bool big = !nan && !(wx <= BCH_BIG_EXTENT && wy <= BCH_BIG_EXTENT && wz <= BCH_BIG_EXTENT);
  uint8_t f = nan ? FLAG_NAN : (big ? FLAG_BIG : FLAG_REG);
  flags[i] = f;
  if (big) big_ids[atomicAdd(&cnt->n_big, 1u)] = i;
  if (f == FLAG_REG) {
    V3 c = (bx.lo + bx.hi) * 0.5f;
    cmin[0] = mn(cmin[0], c.x); cmin[1] = mn(cmin[1], c.y); cmin[2] = mn(cmin[2], c.z);
    cmax[0] = mx(cmax[0], c.x); cmax[1] = mx(cmax[1], c.y); cmax[2] = mx(cmax[2], c.z);
  }
}

__device__ __forceinline__ void reduce_bounds(float* cmin, float* cmax, uint32_t nreg, Counters* cnt) {
  // wave reduction then one atomic per wave
  for (int k = 0; k < 3; ++k) {
    float a = cmin[k], b = cmax[k];
    for (int off = 32; off > 0; off >>= 1) {
      a = mn(a, __shfl_xor(a, off));
      b = mx(b, __shfl_xor(b, off));
    }
    cmin[k] = a;
    cmax[k] = b;
  }
  uint32_t r = nreg;
  for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off);
  if ((threadIdx.x & 63) == 0) {
    if (r) {
      atomicAdd(&cnt->n_reg, r);
      for (int k = 0; k < 3; ++k) {
        atomicMin(&cnt->bounds[k], ord_enc(cmin[k]));
        atomicMax(&cnt->bounds[3 + k], ord_enc(cmax[k]));
      }
    }
  }
}

// World AABB + margin per body (world/step.rs:349-363), fused with
// classification and the scene-bounds reduction.
__global__ void k_aabb(uint32_t n, const BodyRec* __restrict__ bodies, const float4* __restrict__ hull, float margin,
                       float4* __restrict__ aabb, uint8_t* __restrict__ flags, uint32_t* __restrict__ big_ids,
                       Counters* cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t reg = 0;
  if (i < n) {
    BodyRec b = bodies[i];
    ShapeRef s = shape_of(b, hull);
    Box6 bx = world_aabb(s, pos_of(b), quat_of(b));
    V3 pad = splat(margin);  // util.rs:52-55
    bx.lo = bx.lo - pad;
    bx.hi = bx.hi + pad;
    aabb[2 * i] = make_float4(bx.lo.x, bx.lo.y, bx.lo.z, 0.0f);
    aabb[2 * i + 1] = make_float4(bx.hi.x, bx.hi.y, bx.hi.z, 0.0f);
    classify_box(i, bx, flags, cnt, big_ids, cmin, cmax);
    reg = flags[i] == FLAG_REG;
  }
  reduce_bounds(cmin, cmax, reg, cnt);
}

// Same classification for caller-given AABBs (the proxy API); dead proxies
// arrive as NaN boxes.
__global__ void k_aabb_given(uint32_t n, float4* __restrict__ aabb, uint8_t* __restrict__ flags,
                             uint32_t* __restrict__ big_ids, Counters* cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t reg = 0;
  if (i < n) {
    float4 lo = aabb[2 * i], hi = aabb[2 * i + 1];
    Box6 bx{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}};
    classify_box(i, bx, flags, cnt, big_ids, cmin, cmax);
    reg = flags[i] == FLAG_REG;
  }
  reduce_bounds(cmin, cmax, reg, cnt);
}

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

// Sort key: world << 30 | Morton(centre); bodies outside the tree get
// all-ones and sort last.
__global__ void k_morton(uint32_t n, const float4* __restrict__ aabb, const uint8_t* __restrict__ flags,
                         const uint32_t* __restrict__ world_of_body, const Counters* __restrict__ cnt,
                         uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t key = ~0ull;
  if (flags[i] == FLAG_REG) {
    float lo[3], ext[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = ord_dec(cnt->bounds[k]);
      ext[k] = ord_dec(cnt->bounds[3 + k]) - lo[k];
    }
    float4 a = aabb[2 * i], b = aabb[2 * i + 1];
    float c[3] = {(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      float t = ext[k] > 0.0f ? (c[k] - lo[k]) / ext[k] : 0.0f;
      t = fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
      q[k] = (uint32_t)t;
    }
    uint32_t m = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    uint64_t w = world_of_body ? (uint64_t)world_of_body[i] : 0ull;
    key = (w << 30) | (uint64_t)m;
  }
  keys[i] = key;
  vals[i] = i;
}

// Common-prefix length of sorted keys i, j with the index as tie-break
// (Karras 2012); -1 outside [0, n).
__device__ __forceinline__ int kdelta(const uint64_t* __restrict__ keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint64_t a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll(a ^ b);
}

__device__ __forceinline__ int kdir(const uint64_t* __restrict__ keys, int n, int i) {
  return (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
}

// Rope of a subtree whose leaf range ends at `last`: the subtree starting at
// last + 1 -- internal node last + 1 when that node's range starts there,
// else leaf last + 1 (the right child of the nearest ancestor that holds
// this subtree on its left).
__device__ __forceinline__ int rope_after(const uint64_t* __restrict__ keys, int n, int last) {
  int k = last + 1;
  if (k >= n) return -1;
  if (k <= n - 2 && kdir(keys, n, k) == 1) return k;
  return (n - 1) + k;
}

__global__ void k_leaves(uint32_t n_total, const Counters* __restrict__ cnt, const uint64_t* __restrict__ keys,
                         const uint32_t* __restrict__ order, const float4* __restrict__ aabb, Node* __restrict__ nodes,
                         uint32_t* __restrict__ node_first) {
  int n = (int)cnt->n_reg;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t body = order[s];
  float4 lo = aabb[2 * body], hi = aabb[2 * body + 1];
  Node L;
  L.lx = lo.x; L.ly = lo.y; L.lz = lo.z;
  L.hx = hi.x; L.hy = hi.y; L.hz = hi.z;
  L.left = -(int)body - 1;
  L.rope = rope_after(keys, n, s);
  nodes[(n - 1) + s] = L;
  if (node_first) node_first[(n - 1) + s] = (uint32_t)s;
}

__global__ void k_build(const Counters* __restrict__ cnt, const uint64_t* __restrict__ keys, Node* __restrict__ nodes,
                        int* __restrict__ parent, uint32_t* __restrict__ refit_flag, uint32_t* __restrict__ node_first) {
  int n = (int)cnt->n_reg;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = kdir(keys, n, i);
  int dmin = kdelta(keys, n, i, i - d);
  int lmax = 2;
  while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = kdelta(keys, n, i, j);
  int s = 0;
  int t = l;
  while (true) {
    t = (t + 1) >> 1;
    if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int first = i < j ? i : j, last = i < j ? j : i;
  int left = (first == gamma) ? (n - 1) + gamma : gamma;
  int right = (last == gamma + 1) ? (n - 1) + gamma + 1 : gamma + 1;
  nodes[i].left = left;
  nodes[i].rope = rope_after(keys, n, last);
  parent[left] = i;
  parent[right] = i;
  refit_flag[i] = 0u;
  if (node_first) node_first[i] = (uint32_t)first;
  if (i == 0) parent[0] = -1;
}

// Bottom-up boxes: each leaf climbs; the first child to arrive at a parent
// stops, the second computes the union (both children are then final).
__global__ void k_refit(const Counters* __restrict__ cnt, Node* nodes, const int* __restrict__ parent,
                        uint32_t* refit_flag) {
  int n = (int)cnt->n_reg;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || n < 2) return;
  int node = (n - 1) + s;
  while (true) {
    int p = parent[node];
    if (p < 0) break;
    __threadfence();
    if (atomicAdd(&refit_flag[p], 1u) == 0u) break;
    __threadfence();
    int l = __atomic_load_n(&nodes[p].left, __ATOMIC_RELAXED);
    int r = __atomic_load_n(&nodes[l].rope, __ATOMIC_RELAXED);
    const volatile Node* A = &nodes[l];
    const volatile Node* B = &nodes[r];
    Node P;
    P.lx = mn(A->lx, B->lx); P.ly = mn(A->ly, B->ly); P.lz = mn(A->lz, B->lz);
    P.hx = mx(A->hx, B->hx); P.hy = mx(A->hy, B->hy); P.hz = mx(A->hz, B->hz);
    volatile Node* D = &nodes[p];
    D->lx = P.lx; D->ly = P.ly; D->lz = P.lz;
    D->hx = P.hx; D->hy = P.hy; D->hz = P.hz;
    node = p;
  }
}

__device__ __forceinline__ void load_node(const Node* __restrict__ nodes, int i, Box6& b, int& left, int& rope) {
  const float4* p = reinterpret_cast<const float4*>(nodes + i);
  float4 a = p[0], c = p[1];
  b.lo = V3{a.x, a.y, a.z};
  b.hi = V3{c.x, c.y, c.z};
  left = (int)f2u(a.w);
  rope = (int)f2u(c.w);
}

// One lane per leaf (sorted position s): walk the subtrees to the right of
// s in depth-first order via ropes; every overlapping leaf is a pair found
// exactly once.  Pair keys go to the leaf's K slots; counts beyond K are
// recorded so the host can rerun with more slots.  With `world_end`
// (batches), the walk stops at the first node that starts past the query's
// world.
__global__ void k_traverse(const Counters* __restrict__ cnt, const Node* __restrict__ nodes,
                           const uint32_t* __restrict__ node_first, const uint32_t* __restrict__ world_end_of_leaf,
                           uint32_t K, uint64_t* __restrict__ slots, uint32_t* __restrict__ counts,
                           Counters* cnt_rw) {
  int n = (int)cnt->n_reg;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  Box6 q;
  int qleft, node;
  load_node(nodes, (n - 1) + s, q, qleft, node);
  uint32_t body = (uint32_t)(-qleft - 1);
  uint32_t wend = world_end_of_leaf ? world_end_of_leaf[s] : 0xffffffffu;
  uint32_t c = 0;
  uint64_t* my = slots + (uint64_t)s * K;
  while (node >= 0) {
    if (node_first && node_first[node] >= wend) break;
    Box6 b;
    int left, rope;
    load_node(nodes, node, b, left, rope);
    if (overlaps(q, b)) {
      if (left < 0) {
        uint32_t other = (uint32_t)(-left - 1);
        uint32_t lo = body < other ? body : other, hi = body < other ? other : body;
        if (c < K) my[c] = ((uint64_t)lo << 32) | hi;
        ++c;
        node = rope;
      } else {
        node = left;
      }
    } else {
      node = rope;
    }
  }
  counts[s] = c < K ? c : K;
  if (c > K) atomicMax(&cnt_rw->max_leaf, c);
}

__global__ void k_compact_slots(const Counters* __restrict__ cnt, uint32_t K, const uint64_t* __restrict__ slots,
                                const uint32_t* __restrict__ counts, const uint32_t* __restrict__ offsets,
                                uint64_t* __restrict__ keys) {
  int n = (int)cnt->n_reg;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint32_t c = counts[s], o = offsets[s];
  for (uint32_t k = 0; k < c; ++k) keys[o + k] = slots[(uint64_t)s * K + k];
}

// Bodies with unbounded AABBs: tested against every other live body of their
// world (big-big pairs once, from the lower id).  Wave-aggregated append.
__global__ void k_big(const Counters* __restrict__ cnt, const uint32_t* __restrict__ big_ids,
                      const float4* __restrict__ aabb, const uint8_t* __restrict__ flags,
                      const BodyRec* __restrict__ bodies, const uint32_t* __restrict__ world_offsets, uint32_t n,
                      uint64_t* __restrict__ out, uint32_t cap, Counters* cnt_rw) {
  uint32_t nbig = cnt->n_big;
  for (uint32_t bi = blockIdx.y; bi < nbig; bi += gridDim.y) {
    uint32_t b = big_ids[bi];
    uint32_t w0 = 0, w1 = n;
    if (world_offsets && bodies) {
      uint32_t w = bodies[b].world;
      w0 = world_offsets[w];
      w1 = world_offsets[w + 1];
    }
    float4 blo = aabb[2 * b], bhi = aabb[2 * b + 1];
    Box6 bb{V3{blo.x, blo.y, blo.z}, V3{bhi.x, bhi.y, bhi.z}};
    for (uint32_t base = w0 + blockIdx.x * blockDim.x; base < w1; base += gridDim.x * blockDim.x) {
      uint32_t j = base + threadIdx.x;
      bool emit = false;
      if (j < w1 && j != b) {
        uint8_t f = flags[j];
        if (f == FLAG_REG || (f == FLAG_BIG && j > b)) {
          float4 lo = aabb[2 * j], hi = aabb[2 * j + 1];
          Box6 o{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}};
          emit = overlaps(bb, o);
        }
      }
      uint64_t mask = __ballot(emit);
      if (mask) {
        uint32_t lane = threadIdx.x & 63;
        uint32_t leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base_idx = 0;
        if (lane == leader) base_idx = atomicAdd(&cnt_rw->n_big_pairs, (uint32_t)__popcll(mask));
        base_idx = __shfl(base_idx, leader);
        if (emit) {
          uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
          uint32_t idx = base_idx + rank;
          uint32_t lo = b < j ? b : j, hi = b < j ? j : b;
          if (idx < cap) out[idx] = ((uint64_t)lo << 32) | hi;
          else atomicOr(&cnt_rw->errors, 4u);
        }
      }
    }
  }
}

// Sorted keys -> (i, j) pairs and the shape-class sort keys.
__global__ void k_split_pairs(uint32_t P, const uint64_t* __restrict__ keys, const BodyRec* __restrict__ bodies,
                              uint32_t* __restrict__ pairs, uint32_t* __restrict__ cls, uint32_t* __restrict__ iota) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint64_t k = keys[p];
  uint32_t i = (uint32_t)(k >> 32), j = (uint32_t)k;
  pairs[2 * p] = i;
  pairs[2 * p + 1] = j;
  cls[p] = (uint32_t)pair_class((int)(bodies[i].type & 0xffu), (int)(bodies[j].type & 0xffu));
  iota[p] = p;
}

__device__ __forceinline__ void store_contacts(float4* __restrict__ cslots, uint32_t p, const Manifold4& m) {
  for (int c = 0; c < 4; ++c) {
    if (c >= m.n) break;
    cslots[(uint64_t)p * 8 + 2 * c] = make_float4(m.c[c].pos.x, m.c[c].pos.y, m.c[c].pos.z, m.c[c].depth);
    cslots[(uint64_t)p * 8 + 2 * c + 1] = make_float4(m.c[c].n.x, m.c[c].n.y, m.c[c].n.z, 0.0f);
  }
}

__device__ __forceinline__ Pair make_pair(const BodyRec& A, const BodyRec& B, const float4* hull) {
  Pair pr;
  pr.a = shape_of(A, hull);
  pr.b = shape_of(B, hull);
  pr.pa = pos_of(A);
  pr.pb = pos_of(B);
  pr.qa = quat_of(A);
  pr.qb = quat_of(B);
  return pr;
}

// One lane per pair in shape-class order.  Closed forms and GJK finish here;
// overlaps past the margin shell are queued for EPA with their tetrahedron.
__global__ void k_narrow(uint32_t P, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ pairs,
                         const BodyRec* __restrict__ bodies, const float4* __restrict__ hull,
                         const float* __restrict__ seeds, uint32_t* __restrict__ ccount, float4* __restrict__ cslots,
                         float4* __restrict__ epa_q, Counters* cnt) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  bool need_epa = false;
  uint32_t p = 0;
  GjkResult g;
  if (t < P) {
    p = perm[t];
    uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    BodyRec A = bodies[i], B = bodies[j];
    Pair pr = make_pair(A, B, hull);
    float thr = mn(A.threshold, B.threshold);  // narrowphase/mod.rs:52-57
    int cls = pair_class(pr.a.type, pr.b.type);
    Manifold4 m;
    V3 sd;
    const V3* sp = nullptr;
    if (seeds) {
      sd = V3{seeds[3 * p], seeds[3 * p + 1], seeds[3 * p + 2]};
      sp = &sd;
    }
    int r = collide_pair(pr, thr, sp, m, g, cls);
    if (m.n < 0) {
      atomicOr(&cnt->errors, 1u);
      m.n = 0;
    }
    need_epa = r == 1;
    ccount[p] = need_epa ? 0u : (uint32_t)m.n;
    if (!need_epa) store_contacts(cslots, p, m);
  }
  uint64_t mask = __ballot(need_epa);
  if (mask) {
    uint32_t lane = threadIdx.x & 63;
    uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&cnt->n_epa, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (need_epa) {
      uint32_t q = base + __popcll(mask & ((1ull << lane) - 1ull));
      epa_q[4 * (uint64_t)q] = make_float4(g.t0.x, g.t0.y, g.t0.z, u2f(p));
      epa_q[4 * (uint64_t)q + 1] = make_float4(g.t1.x, g.t1.y, g.t1.z, 0.0f);
      epa_q[4 * (uint64_t)q + 2] = make_float4(g.t2.x, g.t2.y, g.t2.z, 0.0f);
      epa_q[4 * (uint64_t)q + 3] = make_float4(g.t3.x, g.t3.y, g.t3.z, 0.0f);
    }
  }
}

// EPA, one lane per queued pair; lane-private polytope in global scratch
// (grid-stride over the queue with a fixed grid, one store per lane).
__global__ void k_epa(const Counters* __restrict__ cnt_ro, const float4* __restrict__ epa_q,
                      const uint32_t* __restrict__ pairs, const BodyRec* __restrict__ bodies,
                      const float4* __restrict__ hull, uint32_t* __restrict__ ccount, float4* __restrict__ cslots,
                      float4* __restrict__ sv, EpaFace* __restrict__ sf, uint32_t* __restrict__ se, Counters* cnt) {
  uint32_t nq = cnt_ro->n_epa;
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  EpaStore st{sv + (uint64_t)gid * BCH_EPA_MAX_VERTS, sf + (uint64_t)gid * BCH_EPA_MAX_FACES,
              se + (uint64_t)gid * BCH_EPA_MAX_EDGES};
  for (uint32_t q = gid; q < nq; q += gridDim.x * blockDim.x) {
    float4 a = epa_q[4 * (uint64_t)q], b = epa_q[4 * (uint64_t)q + 1], c = epa_q[4 * (uint64_t)q + 2],
           d = epa_q[4 * (uint64_t)q + 3];
    uint32_t p = f2u(a.w);
    uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    Pair pr = make_pair(bodies[i], bodies[j], hull);
    EpaResult e = epa_solve(pr, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, V3{d.x, d.y, d.z}, st);
    if (e.status == 4) atomicOr(&cnt->errors, 2u);
    Manifold4 m;
    m.n = 0;
    epa_contact(pr, e, m);
    ccount[p] = (uint32_t)m.n;
    store_contacts(cslots, p, m);
  }
}

// Dense contacts: contact c of pair p -> contacts[7 * (offsets[p] + c)].
__global__ void k_gather(uint32_t P, const uint32_t* __restrict__ ccount, const uint32_t* __restrict__ offsets,
                         const float4* __restrict__ cslots, float* __restrict__ contacts) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint32_t c = ccount[p], o = offsets[p];
  for (uint32_t k = 0; k < c; ++k) {
    float4 a = cslots[(uint64_t)p * 8 + 2 * k], b = cslots[(uint64_t)p * 8 + 2 * k + 1];
    float* dst = contacts + 7 * (uint64_t)(o + k);
    dst[0] = a.x; dst[1] = a.y; dst[2] = a.z;
    dst[3] = b.x; dst[4] = b.y; dst[5] = b.z;
    dst[6] = a.w;
  }
}

// World end (exclusive leaf position) per sorted leaf, for batches: leaves are
// sorted world-major, so world w spans a contiguous run.
__global__ void k_world_end(const Counters* __restrict__ cnt, const uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ wend) {
  int n = (int)cnt->n_reg;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  uint64_t w = keys[s] >> 30;
  // binary search for the first leaf of a later world
  int lo = s + 1, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if ((keys[mid] >> 30) > w) hi = mid;
    else lo = mid + 1;
  }
  wend[s] = (uint32_t)lo;
}

__global__ void k_reset_big(Counters* c) {
  if (threadIdx.x == 0) c->n_big_pairs = 0;
}
__global__ void k_reset_epa(Counters* c) {
  if (threadIdx.x == 0) {
    c->n_epa = 0;
    c->errors &= ~2u;
  }
}

__global__ void k_fill_u32(uint32_t n, uint32_t* p, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace bch
