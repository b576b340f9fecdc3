// bch_kernels.hpp -- HIP kernels of the collision pass (gfx950).
//
// Pass structure (one stream, forked once so the one-lane narrowphase
// classes run beside k_np_hull; no host work between stages except one count
// read-back before the pair sort):
//   k_aabb        body record -> world AABB + margin (fused), scene bounds,
//                 big / NaN classification
//   k_morton      30-bit Morton code of the AABB centre (world id above it)
//   radix sort    (bch_sort.hpp) codes -> leaf order
//   k_tree_*      implicit complete binary tree over the sorted leaves;
//                 boxes unioned level by level in LDS
//   k_traverse    each leaf walks the tree over the leaves to its right
//                 (one leaf per lane, XCD-contiguous leaf ranges): exact pair
//                 set without duplicates
//   k_big         bodies with unbounded AABBs (planes) against their world
//   radix sort    (bch_sort.hpp) pair keys i << 32 | j -> lexicographic order
//   k_np<CLS>     one lane per pair, one kernel per shape class
//   k_np_hull     a group of lanes per pair with a hull
//   k_epa_quad    EPA of the deep-overlap pairs, four per wavefront, polytopes
//                 in LDS (bch_epa_quad.hpp); the few outside its budget
//                 rerun one per wavefront in k_epa_wave (bch_epa_wave.hpp)
//   k_gather      per-pair contact slots -> dense CSR contacts
#pragma once
#include "bch_narrow.hpp"

namespace bch {

struct Counters {
  uint32_t n_reg;       // bodies in the tree
  uint32_t n_big;       // bodies with unbounded AABBs
  uint32_t bounds[6];   // orderable-encoded min xyz, max xyz of regular centres
  uint32_t wave_over;   // largest tree-walk wavefront pair count past its region (0: none)
  uint32_t n_big_pairs;
  uint32_t n_epa;
  uint32_t errors;      // bit 0 clip overflow, bit 1 epa store overflow, bit 3 pair capacity overflow,
                        // bit 4 contact capacity overflow, bit 5 mesh BVH stack, bit 6 bad static-batch
                        // input, bit 7 a tree-walk wavefront's pair region overflowed
  uint32_t cls_start[8];  // pair range of each shape class after the class sort
  uint32_t cls_end[8];
  // capacity-mode pass (counts stay on the device)
  uint32_t n_pairs;        // pairs held (<= capacity)
  uint32_t n_pairs_total;  // pairs found
  uint32_t n_contacts;     // contacts found
  uint32_t n_pen;          // penetrating pairs (de-penetration step)
  uint32_t n_epa_defer;    // EPA entries handed to the one-pair kernel (world pass: by the four-pair
                           // kernel; mesh batch: by the two-pair kernel)
  uint32_t n_epa_over;     // entries whose polytope outgrew the one-pair kernel's store (rerun by k_epa_big)
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
    c->wave_over = 0;
    c->n_big_pairs = 0;
    c->n_epa = 0;
    c->errors = 0;
    c->n_pairs = 0;
    c->n_pairs_total = 0;
    c->n_contacts = 0;
    c->n_pen = 0;
    c->n_epa_defer = 0;
    c->n_epa_over = 0;
  }
}

// Pack caller SoA arrays into body records; shape constants (margin,
// contact_breaking_threshold, normalised plane normal) are computed here.
__global__ void k_pack_bodies(uint32_t n, const float* __restrict__ pos, const float* __restrict__ quat,
                              const uint8_t* __restrict__ type, const float* __restrict__ params,
                              const uint32_t* __restrict__ hull_first, const uint32_t* __restrict__ hull_count,
                              const float4* __restrict__ hull, const uint32_t* __restrict__ world_of,
                              BodyRec* __restrict__ out, uint8_t* __restrict__ type_out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // the pair split's type byte: shape type, bit 4 set for a hull of more
  // than 32 vertices (its supports take twice the scan)
  type_out[i] = (uint8_t)(type[i] | (type[i] == HULL && hull_count[i] > 32u ? 0x10u : 0u));
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
  b.bound = bound_radius(s) + b.threshold;
  out[i] = b;
}

// Poses into the body records.  With velocities, an active body (flag bit 1,
// every body without flags) is placed at its predicted position
// position + linear_velocity * dt (body.rs:1175-1181, step.rs:340-343).
__global__ void k_pack_poses(uint32_t n, const float* __restrict__ pos, const float* __restrict__ quat,
                             BodyRec* __restrict__ out, const float* __restrict__ vel = nullptr, float dt = 0.0f,
                             const uint8_t* __restrict__ bflags = nullptr) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  if (vel && (!bflags || (bflags[i] & 2u))) {
    px = px + vel[3 * i] * dt;
    py = py + vel[3 * i + 1] * dt;
    pz = pz + vel[3 * i + 2] * dt;
  }
  out[i].px = px;
  out[i].py = py;
  out[i].pz = pz;
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

// Block-level reduction of the regular-body centre bounds and count into a
// per-block partial (8 floats); k_bounds_final folds the partials, so no two
// blocks ever touch the same address.
__device__ __forceinline__ void reduce_bounds(float* cmin, float* cmax, uint32_t nreg, float* partial) {
  __shared__ float sred[7][16];
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
  int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    for (int k = 0; k < 3; ++k) {
      sred[k][w] = cmin[k];
      sred[3 + k][w] = cmax[k];
    }
    sred[6][w] = u2f(r);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float out[6];
    uint32_t tot = 0;
    for (int k = 0; k < 3; ++k) {
      out[k] = INFINITY;
      out[3 + k] = -INFINITY;
    }
    for (int i = 0; i < nw; ++i) {
      for (int k = 0; k < 3; ++k) {
        out[k] = mn(out[k], sred[k][i]);
        out[3 + k] = mx(out[3 + k], sred[3 + k][i]);
      }
      tot += f2u(sred[6][i]);
    }
    float* dst = partial + 8 * blockIdx.x;
    for (int k = 0; k < 6; ++k) dst[k] = out[k];
    dst[6] = u2f(tot);
  }
}

// Fold the per-block partials of up to two kernels into the counters.
// One block of 1024 threads reads the partials (8 floats each: min xyz,
// max xyz, count) with two float4 loads per partial, then wave shuffles
// and one value per wave in LDS.  min / max of finite centres and an integer
// count are exact in any order.
__global__ __launch_bounds__(1024) void k_bounds_final(const float* __restrict__ pa, uint32_t na,
                                                       const float* __restrict__ pb, uint32_t nb, Counters* cnt) {
  __shared__ float sm[6][16];
  __shared__ uint32_t sc[16];
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t tot = 0;
  for (uint32_t i = threadIdx.x; i < na + nb; i += blockDim.x) {
    const float4* p = reinterpret_cast<const float4*>(i < na ? pa + 8 * (uint64_t)i : pb + 8 * (uint64_t)(i - na));
    float4 a = p[0], b = p[1];
    lo[0] = mn(lo[0], a.x);
    lo[1] = mn(lo[1], a.y);
    lo[2] = mn(lo[2], a.z);
    hi[0] = mx(hi[0], a.w);
    hi[1] = mx(hi[1], b.x);
    hi[2] = mx(hi[2], b.y);
    tot += f2u(b.z);
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = mn(lo[k], __shfl_xor(lo[k], off));
      hi[k] = mx(hi[k], __shfl_xor(hi[k], off));
    }
    tot += __shfl_xor(tot, off);
  }
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    for (int k = 0; k < 3; ++k) {
      sm[k][w] = lo[k];
      sm[3 + k][w] = hi[k];
    }
    sc[w] = tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < nw; ++i) {
      for (int k = 0; k < 3; ++k) {
        lo[k] = mn(lo[k], sm[k][i]);
        hi[k] = mx(hi[k], sm[3 + k][i]);
      }
      tot += sc[i];
    }
    cnt->n_reg = tot;
    for (int k = 0; k < 3; ++k) {
      cnt->bounds[k] = ord_enc(lo[k]);
      cnt->bounds[3 + k] = ord_enc(hi[k]);
    }
  }
}

// A body whose AABB stays as the previous pass left it: inactive bodies
// (flag bit 1 clear) when `keep` is set (step.rs:348-353: sleeping bodies
// keep their broadphase AABB).
__device__ __forceinline__ bool keeps_aabb(const uint8_t* bflags, uint32_t i, int keep) {
  return keep && bflags && !(bflags[i] & 2u);
}

// World AABB + margin per body (world/step.rs:349-363), fused with
// classification and the scene-bounds reduction.
__global__ void k_aabb(uint32_t n, const BodyRec* __restrict__ bodies, const float4* __restrict__ hull, float margin,
                       float4* __restrict__ aabb, uint8_t* __restrict__ flags, uint32_t* __restrict__ big_ids,
                       Counters* cnt, float* __restrict__ partial, const uint8_t* __restrict__ bflags = nullptr,
                       int keep = 0, float4* __restrict__ bsph = nullptr) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t reg = 0;
  BodyRec b;
  if (i < n) b = bodies[i];
  if (i < n && bsph) {
    // the body's bounding ball at its current pose (every body, also one
    // that keeps its previous AABB); NaN radius for a non-finite pose
    float s = ((b.px + b.py) + b.pz) + (((b.qx + b.qy) + b.qz) + b.qw);
    bool finite = s - s == 0.0f;
    bsph[i] = make_float4(b.px, b.py, b.pz, finite ? b.bound : NAN);
  }
  if (i < n && (b.type & 0xffu) != HULL) {  // hull boxes: k_hull_aabb
    Box6 bx;
    if (keeps_aabb(bflags, i, keep)) {
      float4 lo = aabb[2 * i], hi = aabb[2 * i + 1];
      bx = Box6{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}};
    } else {
      ShapeRef s = shape_of(b, hull);
      bx = world_aabb(s, pos_of(b), quat_of(b));
      V3 pad = splat(margin);  // util.rs:52-55
      bx.lo = bx.lo - pad;
      bx.hi = bx.hi + pad;
      aabb[2 * i] = make_float4(bx.lo.x, bx.lo.y, bx.lo.z, 0.0f);
      aabb[2 * i + 1] = make_float4(bx.hi.x, bx.hi.y, bx.hi.z, 0.0f);
    }
    classify_box(i, bx, flags, cnt, big_ids, cmin, cmax);
    reg = flags[i] == FLAG_REG;
  }
  reduce_bounds(cmin, cmax, reg, partial);
}

// Same classification for caller-given AABBs (the proxy API); dead proxies
// arrive as NaN boxes.
__global__ void k_aabb_given(uint32_t n, float4* __restrict__ aabb, uint8_t* __restrict__ flags,
                             uint32_t* __restrict__ big_ids, Counters* cnt, float* __restrict__ partial) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t reg = 0;
  if (i < n) {
    float4 lo = aabb[2 * i], hi = aabb[2 * i + 1];
    Box6 bx{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}};
    classify_box(i, bx, flags, cnt, big_ids, cmin, cmax);
    reg = flags[i] == FLAG_REG;
  }
  reduce_bounds(cmin, cmax, reg, partial);
}

// Hull world AABBs (shapes/mod.rs:448-469): 16 lanes per hull scan the
// vertices for the six axis supports; each lane keeps its first maximiser
// and the group reduces (largest value, then lowest index), which is the
// first maximiser of the sequential scan (convex_hull.rs:55-66).
__device__ __forceinline__ void argmax_merge(float& v, int& i, float ov, int oi) {
  if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) {
    v = ov;
    i = oi;
  }
}

__global__ void k_hull_aabb(uint32_t n_hulls, const uint32_t* __restrict__ hull_ids,
                            const BodyRec* __restrict__ bodies, const float4* __restrict__ hull, float margin,
                            float4* __restrict__ aabb, uint8_t* __restrict__ flags, uint32_t* __restrict__ big_ids,
                            Counters* cnt, float* __restrict__ partial, const uint8_t* __restrict__ bflags = nullptr,
                            int keep = 0) {
  uint32_t h = blockIdx.x * (blockDim.x / 16) + threadIdx.x / 16;
  int lane = threadIdx.x & 15;
  float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
  uint32_t reg = 0;
  bool active = h < n_hulls;
  uint32_t body = active ? hull_ids[h] : 0u;
  BodyRec b = bodies[body];
  const float4* hv = hull + f2u(b.a0);
  int nv = active ? (int)f2u(b.a1) : 0;
  Q4 qi = qconj(quat_of(b));
  V3 u[3] = {qrot(qi, V3{1.0f, 0.0f, 0.0f}), qrot(qi, V3{0.0f, 1.0f, 0.0f}), qrot(qi, V3{0.0f, 0.0f, 1.0f})};
  float bv[6];
  int bi[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    bv[k] = -INFINITY;
    bi[k] = -1;
  }
  for (int v = lane; v < nv; v += 16) {
    V3 p = hull_vertex(hv, v);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float d = dot(p, u[k]);
      if (d > bv[k]) { bv[k] = d; bi[k] = v; }
      float e = dot(p, -u[k]);
      if (e > bv[3 + k]) { bv[3 + k] = e; bi[3 + k] = v; }
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    for (int off = 8; off > 0; off >>= 1) {
      float ov = __shfl_xor(bv[k], off, 16);
      int oi = __shfl_xor(bi[k], off, 16);
      argmax_merge(bv[k], bi[k], ov, oi);
    }
  }
  if (active && lane == 0) {
    Box6 bx;
    if (keeps_aabb(bflags, body, keep)) {
      float4 lo = aabb[2 * body], hi = aabb[2 * body + 1];
      bx = Box6{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}};
    } else {
      V3 p = pos_of(b);
      V3 sp[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) sp[k] = bi[k] >= 0 ? hull_vertex(hv, bi[k]) : splat(0.0f);
      bx.lo = V3{p.x + dot(sp[3], u[0]), p.y + dot(sp[4], u[1]), p.z + dot(sp[5], u[2])};
      bx.hi = V3{p.x + dot(sp[0], u[0]), p.y + dot(sp[1], u[1]), p.z + dot(sp[2], u[2])};
      V3 pad = splat(margin);
      bx.lo = bx.lo - pad;
      bx.hi = bx.hi + pad;
      aabb[2 * body] = make_float4(bx.lo.x, bx.lo.y, bx.lo.z, 0.0f);
      aabb[2 * body + 1] = make_float4(bx.hi.x, bx.hi.y, bx.hi.z, 0.0f);
    }
    classify_box(body, bx, flags, cnt, big_ids, cmin, cmax);
    reg = flags[body] == FLAG_REG;
  }
  reduce_bounds(cmin, cmax, reg, partial);
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

// ------------------------------------------------------------------ tree --
// Implicit complete binary tree over N = 2^L leaf slots, 1-based heap order:
// node k has children 2k and 2k + 1, leaves are nodes N .. 2N - 1 and hold the
// boxes in Morton order (slots past the regular bodies hold an empty box).
// Each node is two float4 (lo, hi); a leaf's lo.w carries its body id.
// Depth-first order needs no pointers: the next subtree after node k (its
// rope) is found by stripping k's trailing one bits and stepping right.
__device__ __forceinline__ uint32_t tree_rope(uint32_t k) {
  k >>= __builtin_ctz(~k);
  return k ? k + 1u : 0u;
}

__global__ void k_tree_leaves(uint32_t N, const Counters* __restrict__ cnt, const uint32_t* __restrict__ order,
                              const float4* __restrict__ aabb, float4* __restrict__ nodes) {
  uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
  if (s < cnt->n_reg) {
    uint32_t b = order[s];
    lo = aabb[2 * b];
    hi = aabb[2 * b + 1];
    lo.w = u2f(b);
  }
  nodes[2 * (uint64_t)(N + s)] = lo;
  nodes[2 * (uint64_t)(N + s) + 1] = hi;
}

// Boxes of `m` levels above the child level starting at node `first`
// (= 2^level): each block unions 2^m consecutive children in LDS.
__global__ void k_tree_up(float4* __restrict__ nodes, uint32_t first, int m) {
  __shared__ float4 slo[512], shi[512];
  uint32_t per = 1u << m;
  uint64_t base = (uint64_t)first + (uint64_t)blockIdx.x * per;
  for (uint32_t i = threadIdx.x; i < per; i += blockDim.x) {
    slo[i] = nodes[2 * (base + i)];
    shi[i] = nodes[2 * (base + i) + 1];
  }
  __syncthreads();
  uint32_t cnt = per;
  for (int t = 0; t < m; ++t) {
    cnt >>= 1;
    base >>= 1;
    float4 l, h;
    uint32_t i = threadIdx.x;
    if (i < cnt) {
      float4 a = slo[2 * i], b = slo[2 * i + 1], c = shi[2 * i], d = shi[2 * i + 1];
      l = make_float4(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z), 0.0f);
      h = make_float4(mx(c.x, d.x), mx(c.y, d.y), mx(c.z, d.z), 0.0f);
    }
    __syncthreads();
    if (i < cnt) {
      slo[i] = l;
      shi[i] = h;
      nodes[2 * (base + i)] = l;
      nodes[2 * (base + i) + 1] = h;
    }
    __syncthreads();
  }
}

// Leaf s walks the subtrees right of s in depth-first order; every
// overlapping leaf is a pair found exactly once, as the key i << 32 | j
// (i < j).  With `wend` (batches: leaves sorted world-major), the walk stops
// at the first node whose first leaf lies past the query's world.
struct LeafWalk {
  float4 qlo, qhi;
  uint32_t body, wend, k;
};
__device__ __forceinline__ void walk_begin(LeafWalk& w, uint32_t s, const float4* __restrict__ nodes, uint32_t N,
                                           const uint32_t* __restrict__ wend_of_leaf, bool valid) {
  w.k = 0;
  if (!valid) return;
  w.qlo = nodes[2 * (uint64_t)(N + s)];
  w.qhi = nodes[2 * (uint64_t)(N + s) + 1];
  w.body = f2u(w.qlo.w);
  w.wend = wend_of_leaf ? wend_of_leaf[s] : 0xffffffffu;
  w.k = tree_rope(N + s);
}
// One node of the walk; returns whether it completed a pair (*key).
__device__ __forceinline__ bool walk_step(LeafWalk& w, float4 lo, float4 hi, uint32_t N, uint64_t* key) {
  bool ov = w.qlo.x <= hi.x && w.qlo.y <= hi.y && w.qlo.z <= hi.z && lo.x <= w.qhi.x && lo.y <= w.qhi.y &&
            lo.z <= w.qhi.z;
  bool pair = false;
  if (ov && w.k < N) {
    w.k <<= 1;
  } else {
    if (ov) {
      uint32_t other = f2u(lo.w);
      uint32_t a = w.body < other ? w.body : other, b = w.body < other ? other : w.body;
      *key = ((uint64_t)a << 32) | b;
      pair = true;
    }
    w.k = tree_rope(w.k);
  }
  return pair;
}
__device__ __forceinline__ void walk_world_stop(LeafWalk& w, const uint32_t* __restrict__ wend_of_leaf, uint32_t N,
                                                int L) {
  if (wend_of_leaf && w.k) {
    int d = 31 - __clz(w.k);
    uint32_t first = (w.k << (L - d)) - N;
    if (first >= w.wend) w.k = 0;
  }
}
// Workgroups go to the eight XCDs round-robin (block b to XCD b mod 8), and
// each XCD has its own L2.  This bijection of [0, nb) gives XCD x the
// contiguous logical blocks x q + min(x, r) ... (q = nb / 8, r = nb mod 8),
// so neighbouring leaves, which walk the same tree nodes, share one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

// Per-wavefront pair compaction.  Each step's pairs are ranked by a ballot
// and appended to the wavefront's LDS list; every 64 of them go out as one
// coalesced 512-B store into the wavefront's own region of R slots (no
// atomics).  Pairs past R are counted but not written.
struct PairStage {
  uint64_t* st;       // this wavefront's 128 LDS slots
  uint64_t* region;   // its R output slots
  uint32_t fill, lane, written, R;
};
__device__ __forceinline__ void stage_flush(PairStage& ps, uint32_t m) {
  if (ps.lane < m && ps.written + ps.lane < ps.R) ps.region[ps.written + ps.lane] = ps.st[ps.lane];
  ps.written += m;
}
__device__ __forceinline__ void stage_push(PairStage& ps, bool emit, uint64_t key) {
  uint64_t m = __ballot(emit);
  if (!m) return;
  uint32_t r = __popcll(m & ((1ull << ps.lane) - 1ull));
  if (emit) ps.st[ps.fill + r] = key;
  ps.fill += (uint32_t)__popcll(m);
  if (ps.fill >= 64) {
    stage_flush(ps, 64);
    uint32_t rest = ps.fill - 64;
    uint64_t v = ps.lane < rest ? ps.st[64 + ps.lane] : 0ull;
    if (ps.lane < rest) ps.st[ps.lane] = v;
    ps.fill = rest;
  }
}

// The tree walk of every leaf (one leaf per lane, lanes of a wavefront kept
// together until the last walk ends).  Wavefront w (logical, after the XCD
// remap: leaves 64 w .. 64 w + 63) writes its pairs to regions[w R ..] and
// its count to wave_counts[w]; a count past R raises wave_over to it.
// Blocks are remapped with xcd_block so each XCD walks a contiguous run of
// leaves.
template <int B>
__global__ __launch_bounds__(B) void k_traverse(const Counters* __restrict__ cnt, const float4* __restrict__ nodes,
                                                uint32_t N, int L, const uint32_t* __restrict__ wend_of_leaf,
                                                uint64_t* __restrict__ regions, uint32_t R,
                                                uint32_t* __restrict__ wave_counts, Counters* cnt_rw) {
  __shared__ uint64_t stage[B / 64][128];
  const uint32_t n = cnt->n_reg;
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x) * B + threadIdx.x;
  const uint32_t wave = t >> 6;
  PairStage ps{stage[threadIdx.x >> 6], regions + (uint64_t)wave * R, 0u, threadIdx.x & 63u, 0u, R};
  LeafWalk w;
  walk_begin(w, t, nodes, N, wend_of_leaf, t < n);
  walk_world_stop(w, wend_of_leaf, N, L);
  while (__any(w.k != 0)) {
    bool emit = false;
    uint64_t key = 0;
    if (w.k) {
      float4 lo = nodes[2 * (uint64_t)w.k], hi = nodes[2 * (uint64_t)w.k + 1];
      emit = walk_step(w, lo, hi, N, &key);
      walk_world_stop(w, wend_of_leaf, N, L);
    }
    stage_push(ps, emit, key);
  }
  if (ps.fill) stage_flush(ps, ps.fill);
  if (ps.lane == 0) {
    wave_counts[wave] = ps.written;
    if (ps.written > R) atomicMax(&cnt_rw->wave_over, ps.written);
  }
}

// The wavefronts' regions -> one dense list at their scanned offsets (one
// block per wavefront), up to `cap`.
__global__ __launch_bounds__(64) void k_concat_regions(const uint64_t* __restrict__ regions, uint32_t R,
                                                       const uint32_t* __restrict__ wave_counts,
                                                       const uint32_t* __restrict__ wave_offs, uint64_t cap,
                                                       uint64_t* __restrict__ out) {
  uint32_t w = blockIdx.x;
  uint32_t c = wave_counts[w];
  c = c < R ? c : R;
  uint64_t o = wave_offs[w];
  const uint64_t* src = regions + (uint64_t)w * R;
  for (uint32_t k = threadIdx.x; k < c; k += 64)
    if (o + k < cap) out[o + k] = src[k];
}

// Capacity mode: the held and found pair counts (tree pairs at
// *tree_total, then the big bodies' pairs) and the overflow flags: bit 3
// pair capacity, bit 7 a wavefront region.
__global__ void k_pair_counts(Counters* cnt, const uint32_t* __restrict__ tree_total, uint64_t cap, uint32_t R) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint64_t total = (uint64_t)*tree_total + cnt->n_big_pairs;
    cnt->n_pairs = (uint32_t)(total < cap ? total : cap);
    cnt->n_pairs_total = (uint32_t)(total < 0xffffffffull ? total : 0xffffffffull);
    if (total > cap) atomicOr(&cnt->errors, 8u);
    if (cnt->wave_over > R) atomicOr(&cnt->errors, 128u);
  }
}

// Bodies with unbounded boxes (planes) against every body of their world,
// brute force; pairs appended after the tree walk's *tree_total pairs (the
// count, Counters::n_big_pairs, still grows past `cap`).  One block takes
// a big body and a chunk of kBigChunk bodies of its world: its pairs are
// ranked by ballots within the wavefronts and by wavefront counts in LDS,
// staged there, and claimed with one atomic per block (one counter shared
// by every block: an atomic per wavefront serialised 27k same-address
// atomics over 4096 planes).
constexpr uint32_t kBigChunk = 2048;
__global__ __launch_bounds__(256) void k_big(const Counters* __restrict__ cnt, const uint32_t* __restrict__ big_ids,
                                             const float4* __restrict__ aabb, const uint8_t* __restrict__ flags,
                                             const BodyRec* __restrict__ bodies,
                                             const uint32_t* __restrict__ world_offsets, uint32_t n,
                                             uint64_t* __restrict__ out, uint64_t cap,
                                             const uint32_t* __restrict__ tree_total, Counters* cnt_rw) {
  __shared__ uint64_t stage[kBigChunk];
  __shared__ uint32_t wcnt[4];
  __shared__ uint32_t base_s;
  const uint64_t t0 = *tree_total;
  const uint32_t nbig = cnt->n_big;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (uint32_t bi = blockIdx.y; bi < nbig; bi += gridDim.y) {
    const uint32_t b = big_ids[bi];
    uint32_t w0 = 0, w1 = n;
    if (world_offsets && bodies) {
      uint32_t w = bodies[b].world;
      w0 = world_offsets[w];
      w1 = world_offsets[w + 1];
    }
    const float4 blo = aabb[2 * b], bhi = aabb[2 * b + 1];
    const Box6 bb{V3{blo.x, blo.y, blo.z}, V3{bhi.x, bhi.y, bhi.z}};
    for (uint32_t c0 = w0 + blockIdx.x * kBigChunk; c0 < w1; c0 += gridDim.x * kBigChunk) {
      const uint32_t c1 = c0 + kBigChunk < w1 ? c0 + kBigChunk : w1;
      uint32_t fill = 0;  // block-uniform
      for (uint32_t base = c0; base < c1; base += blockDim.x) {
        const uint32_t j = base + threadIdx.x;
        bool emit = false;
        if (j < c1 && j != b) {
          const uint8_t f = flags[j];
          if (f == FLAG_REG || (f == FLAG_BIG && j > b)) {
            const float4 lo = aabb[2 * j], hi = aabb[2 * j + 1];
            emit = overlaps(bb, Box6{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}});
          }
        }
        const uint64_t mask = __ballot(emit);
        if (lane == 0) wcnt[wv] = (uint32_t)__popcll(mask);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t k = 0; k < blockDim.x / 64; ++k) {
          before += k < wv ? wcnt[k] : 0u;
          total += wcnt[k];
        }
        if (emit) {
          const uint32_t lo_id = b < j ? b : j, hi_id = b < j ? j : b;
          stage[fill + before + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] =
              ((uint64_t)lo_id << 32) | hi_id;
        }
        fill += total;
        __syncthreads();
      }
      if (threadIdx.x == 0) base_s = fill ? atomicAdd(&cnt_rw->n_big_pairs, fill) : 0u;
      __syncthreads();
      const uint64_t at = t0 + base_s;
      for (uint32_t k = threadIdx.x; k < fill; k += blockDim.x)
        if (at + k < cap) out[at + k] = stage[k];
      __syncthreads();
    }
  }
}

// Sorted keys -> (i, j) pairs and the shape-class sort keys.
// P_dev (when given) caps P with a count the device computed.
__device__ __forceinline__ uint32_t dev_count(uint32_t n, const uint32_t* n_dev) {
  if (!n_dev) return n;
  uint32_t m = *n_dev;
  return m < n ? m : n;
}

// Per-pair class word: bits 0-2 the shape class (the route readback's low
// byte; EPA sets bit 3), bits 8-15 the sort digit class << 5 | sub, which
// groups the pairs by class and, within a class, by `sub`.
__device__ __forceinline__ uint32_t cls_key(uint32_t c, uint32_t sub) { return c | (((c << 5) | sub) << 8); }
__device__ __forceinline__ uint32_t cls_of_key(uint32_t k) { return (k >> 13) & 7u; }

// The sort class of walk-class pairs whose bounding balls are apart by more
// than the band: no kernel runs for them (their count is written here).
constexpr uint32_t kClsNoKernel = 7;

// With bsph (the bodies' bounding balls of this pass), a walk-class pair
// (GJK, GJK_HULL) whose centres are farther apart than the two radii
// (each grown by its contact_breaking_threshold) plus a rounding allowance
// cannot end in the reference's Band or Overlap arms: the walk's closest
// point v is a point of the core Minkowski difference, so |v| exceeds the
// margins plus the band (gjk.rs:236-244, 261-275).  Its result is None
// (count 0) and it is sorted into kClsNoKernel; its route keeps its class.
// The rest of a walk class is grouped by shape types, hull size and whether
// the centres lie within 0.85 of the radii (deeper pairs walk longer), so the
// lanes of a wavefront take the same support code and similar walk lengths.
// One pair of the split: writes pairs[2p..], cls[p] (and ccount[p] = 0 for
// a skipped walk pair); returns the class word.
__device__ __forceinline__ uint32_t split_pair(uint32_t p, const uint64_t* __restrict__ keys,
                                               const uint8_t* __restrict__ btype, uint32_t* __restrict__ pairs,
                                               uint32_t* __restrict__ cls, const float4* __restrict__ bsph,
                                               uint32_t* __restrict__ ccount) {
  uint64_t k = keys[p];
  uint32_t i = (uint32_t)(k >> 32), j = (uint32_t)k;
  pairs[2 * p] = i;
  pairs[2 * p + 1] = j;
  const uint32_t ta = btype[i], tb = btype[j];
  const uint32_t c = (uint32_t)pair_class((int)(ta & 7u), (int)(tb & 7u));
  uint32_t sort_c = c, sub = 0u;
  if (c == CLS_GJK_HULL || c == CLS_GJK) {
    bool deep = false;
    if (bsph) {
      float4 A = bsph[i], B = bsph[j];
      float dx = A.x - B.x, dy = A.y - B.y, dz = A.z - B.z;
      float d2 = (dx * dx + dy * dy) + dz * dz;
      float lim = A.w + B.w;
      float mag = fmaxf(fmaxf(fmaxf(fabsf(A.x), fabsf(A.y)), fmaxf(fabsf(A.z), fabsf(B.x))),
                        fmaxf(fabsf(B.y), fabsf(B.z)));
      float far = lim * (1.0f + 1e-5f) + 1e-5f * (1.0f + mag);
      if (d2 > far * far) {  // NaN never passes
        sort_c = kClsNoKernel;
        if (ccount) ccount[p] = 0u;
      }
      deep = d2 < (0.85f * lim) * (0.85f * lim);
    }
    const uint32_t ta3 = ta & 3u, tb3 = tb & 3u;
    if (c == CLS_GJK_HULL) {
      // 7 type combinations (one shape is a hull), the large-hull bit, deep
      const uint32_t combo = ta3 == (uint32_t)HULL ? tb3 : 4u + ta3;
      sub = (combo << 2) | (((ta | tb) & 0x10u) >> 3) | (deep ? 1u : 0u);
    } else {
      sub = (ta3 << 2) | tb3 | (deep ? 0x10u : 0u);
    }
  }
  const uint32_t w = c | (((sort_c << 5) | sub) << 8);
  cls[p] = w;
  return w;
}

__global__ void k_split_pairs(uint32_t P, const uint32_t* __restrict__ P_dev, const uint64_t* __restrict__ keys,
                              const uint8_t* __restrict__ btype, uint32_t* __restrict__ pairs,
                              uint32_t* __restrict__ cls, uint32_t* __restrict__ iota,
                              const float4* __restrict__ bsph = nullptr, uint32_t* __restrict__ ccount = nullptr) {
  P = dev_count(P, P_dev);
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  (void)split_pair(p, keys, btype, pairs, cls, bsph, ccount);
  iota[p] = p;
}

// The split of one class-sort tile of ROUNDS x 256 pairs per block, with the
// tile's histogram of the class digit (bits 8-15 of the class word) written
// as k_rs_hist writes it ([256][tiles]), so the class sort skips its own
// histogram pass.  The sort takes the pair index itself as the value.
template <int ROUNDS>
__global__ __launch_bounds__(256) void k_split_pairs_hist(uint32_t P, const uint32_t* __restrict__ P_dev,
                                                         const uint64_t* __restrict__ keys,
                                                         const uint8_t* __restrict__ btype,
                                                         uint32_t* __restrict__ pairs, uint32_t* __restrict__ cls,
                                                         const float4* __restrict__ bsph,
                                                         uint32_t* __restrict__ ccount, uint32_t tiles,
                                                         uint32_t* __restrict__ hist) {
  P = dev_count(P, P_dev);
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * (256 * ROUNDS);
  for (int r = 0; r < ROUNDS; ++r) {
    const uint64_t p = base + (uint64_t)r * 256 + threadIdx.x;
    if (p < P) atomicAdd(&h[(split_pair((uint32_t)p, keys, btype, pairs, cls, bsph, ccount) >> 8) & 0xffu], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Input check of the static batch: shape types, hull vertex ranges and pair
// indices; a bad input sets error bit 6 and nothing else runs.
__global__ void k_static_check(uint32_t n, const uint8_t* __restrict__ type, const uint32_t* __restrict__ hull_first,
                               const uint32_t* __restrict__ hull_count, uint32_t nv, uint32_t P,
                               const uint32_t* __restrict__ pairs, Counters* cnt) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < n) {
    uint8_t t = type[i];
    if (t > PLANE) bad = true;
    else if (t == HULL) bad = !hull_first || !hull_count || (uint64_t)hull_first[i] + hull_count[i] > nv;
  }
  if (i < P) bad = bad || pairs[2 * (uint64_t)i] >= n || pairs[2 * (uint64_t)i + 1] >= n;
  if (bad) atomicOr(&cnt->errors, 64u);
}

// Proxy arena update: slot idx[k] takes the box (lo, hi) = upd[2k], upd[2k + 1].
__global__ void k_arena_scatter(uint32_t m, const uint32_t* __restrict__ idx, const float4* __restrict__ upd,
                                float4* __restrict__ arena) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  uint32_t i = idx[k];
  arena[2 * (uint64_t)i] = upd[2 * (uint64_t)k];
  arena[2 * (uint64_t)i + 1] = upd[2 * (uint64_t)k + 1];
}

// Sorted pair keys i << 32 | j -> (i, j) as u32 pairs.
__global__ void k_keys_to_pairs(uint32_t P, const uint64_t* __restrict__ keys, uint2* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint64_t k = keys[p];
  out[p] = make_uint2((uint32_t)(k >> 32), (uint32_t)k);
}

// Caller pairs (i, j) -> keys i << 32 | j, kept in caller order.
__global__ void k_pairs_to_keys(uint32_t P, const uint2* __restrict__ pairs, uint64_t* __restrict__ keys) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint2 q = pairs[p];
  keys[p] = ((uint64_t)q.x << 32) | q.y;
}

// Per-pair contact slots -> the batch layout: a u8 count and four points of
// (position xyz, normal xyz, depth), zero past the count.
__global__ void k_batch_out(uint32_t P, const uint32_t* __restrict__ ccount, const float4* __restrict__ cslots,
                            uint8_t* __restrict__ out_count, float* __restrict__ out_pts) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint32_t n = ccount[p];
  out_count[p] = (uint8_t)n;
  float* o = out_pts + 28 * (uint64_t)p;
#pragma unroll
  for (uint32_t q = 0; q < 4; ++q) {
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = a;
    if (q < n) {
      a = cslots[8 * (uint64_t)p + 2 * q];
      b = cslots[8 * (uint64_t)p + 2 * q + 1];
    }
    o[7 * q] = a.x;
    o[7 * q + 1] = a.y;
    o[7 * q + 2] = a.z;
    o[7 * q + 3] = b.x;
    o[7 * q + 4] = b.y;
    o[7 * q + 5] = b.z;
    o[7 * q + 6] = a.w;
  }
}

__device__ __forceinline__ void store_contacts(float4* __restrict__ cslots, uint32_t p, const Manifold4& m) {
#pragma unroll
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

// Pair ranges of each shape class in the class-sorted order.
__global__ void k_class_bounds(uint32_t P, const uint32_t* __restrict__ P_dev, const uint32_t* __restrict__ cls_sorted,
                               Counters* cnt) {
  P = dev_count(P, P_dev);
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P) return;
  uint32_t c = cls_of_key(cls_sorted[t]);
  if (t == 0 || cls_of_key(cls_sorted[t - 1]) != c) cnt->cls_start[c] = t;
  if (t == P - 1 || cls_of_key(cls_sorted[t + 1]) != c) cnt->cls_end[c] = t + 1;
}

// One lane per pair of shape class CLS (grid-stride over the class range):
// each class gets its own kernel and register budget.  Closed forms and the
// GJK walk finish here; overlaps past the margin shell are queued for EPA
// with their tetrahedron.
template <int CLS>
__global__ __launch_bounds__(256) void k_np(const Counters* __restrict__ cnt_ro, const uint32_t* __restrict__ perm,
                     const uint32_t* __restrict__ pairs, const BodyRec* __restrict__ bodies,
                     const float4* __restrict__ hull, const float* __restrict__ seeds, uint32_t* __restrict__ ccount,
                     float4* __restrict__ cslots, float4* __restrict__ epa_q, Counters* cnt) {
  uint32_t beg = cnt_ro->cls_start[CLS], end = cnt_ro->cls_end[CLS];
  for (uint32_t base = beg + blockIdx.x * blockDim.x; base < end; base += gridDim.x * blockDim.x) {
    uint32_t t = base + threadIdx.x;
    bool need_epa = false;
    uint32_t p = 0;
    GjkResult g;
    if (t < end) {
      p = perm[t];
      uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
      BodyRec A = bodies[i], B = bodies[j];
      Pair pr = make_pair(A, B, hull);
      float thr = mn(A.threshold, B.threshold);  // narrowphase/mod.rs:52-57
      Manifold4 m;
      V3 sd{0.0f, 0.0f, 0.0f};
      // a seed whose x is NaN is None for that pair
      bool has_seed = CLS == CLS_GJK && seeds != nullptr && !is_nan(seeds[3 * p]);
      if (has_seed) sd = V3{seeds[3 * p], seeds[3 * p + 1], seeds[3 * p + 2]};
      int r = collide_pair(pr, thr, has_seed, sd, m, g, CLS);
      if (m.n < 0) {
        atomicOr(&cnt->errors, 1u);
        m.n = 0;
      }
      need_epa = r == 1;
      ccount[p] = need_epa ? 0u : (uint32_t)m.n;
      if (!need_epa) store_contacts(cslots, p, m);
    }
    if (CLS == CLS_GJK) {
      uint64_t mask = __ballot(need_epa);
      if (mask) {
        uint32_t lane = threadIdx.x & 63;
        uint32_t leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t qb = 0;
        if (lane == leader) qb = atomicAdd(&cnt->n_epa, (uint32_t)__popcll(mask));
        qb = __shfl(qb, leader);
        if (need_epa) {
          uint32_t q = qb + __popcll(mask & ((1ull << lane) - 1ull));
          epa_q[4 * (uint64_t)q] = make_float4(g.t0.x, g.t0.y, g.t0.z, u2f(p));
          epa_q[4 * (uint64_t)q + 1] = make_float4(g.t1.x, g.t1.y, g.t1.z, 0.0f);
          epa_q[4 * (uint64_t)q + 2] = make_float4(g.t2.x, g.t2.y, g.t2.z, 0.0f);
          epa_q[4 * (uint64_t)q + 3] = make_float4(g.t3.x, g.t3.y, g.t3.z, 0.0f);
        }
      }
    }
  }
}

// ------------------------------------------------------ DPP reductions --
// Cross-lane moves within 16-lane rows (gfx9 DPP controls): quad_perm xor 1
// (0xB1), quad_perm xor 2 (0x4E), row half-mirror (0x141), row mirror
// (0x140).  Combining a lane with these partners in turn leaves every lane
// of the row with the row's reduction (the combine is commutative and
// associative).  A few cycles each, against ~100 for ds_bpermute.
// A partner lane that is switched off yields the lane's own value (old =
// src, bound_ctrl off), which the combine absorbs.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}

// (value, index) argmax: larger value wins, ties and invalid (-1) by index.
__device__ __forceinline__ void amax_merge(float& v, int& i, float ov, int oi) {
  if (oi >= 0 && (i < 0 || ov > v || (ov == v && oi < i))) {
    v = ov;
    i = oi;
  }
}
__device__ __forceinline__ void amin_merge(float& v, int& i, float ov, int oi) {
  if (oi >= 0 && (i < 0 || ov < v || (ov == v && oi < i))) {
    v = ov;
    i = oi;
  }
}

template <bool MAX, int WIDTH>
__device__ __forceinline__ void row_reduce(float& v, int& i) {
  // WIDTH in {4, 8, 16}: number of lanes of the group, aligned in the row
  {
    float ov = dpp_f<0xB1>(v);
    int oi = dpp_i<0xB1>(i);
    if (MAX) amax_merge(v, i, ov, oi); else amin_merge(v, i, ov, oi);
  }
  {
    float ov = dpp_f<0x4E>(v);
    int oi = dpp_i<0x4E>(i);
    if (MAX) amax_merge(v, i, ov, oi); else amin_merge(v, i, ov, oi);
  }
  if (WIDTH >= 8) {
    float ov = dpp_f<0x141>(v);
    int oi = dpp_i<0x141>(i);
    if (MAX) amax_merge(v, i, ov, oi); else amin_merge(v, i, ov, oi);
  }
  if (WIDTH >= 16) {
    float ov = dpp_f<0x140>(v);
    int oi = dpp_i<0x140>(i);
    if (MAX) amax_merge(v, i, ov, oi); else amin_merge(v, i, ov, oi);
  }
}

// Whole-wavefront reduction (all 64 lanes active): rows by DPP, then the
// four row results combined from lanes 0, 16, 32, 48 (uniform result).
template <bool MAX>
__device__ __forceinline__ void wave_reduce(float& v, int& i) {
  row_reduce<MAX, 16>(v, i);
  float rv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  int ri = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    float ov = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16 * r));
    int oi = __builtin_amdgcn_readlane(i, 16 * r);
    if (MAX) amax_merge(rv, ri, ov, oi); else amin_merge(rv, ri, ov, oi);
  }
  v = rv;
  i = ri;
}

// ---------------------------------------------------- hull pairs by group --
// Support of a hull split over the G lanes of a group: each lane keeps the
// first maximiser of its vertex subset, the group reduces (largest value,
// then lowest index) = the sequential first maximum (convex_hull.rs:55-66).
template <int G>
__device__ __forceinline__ V3 group_hull_support(const float4* h, int n, V3 d, int sub) {
  float bv = -INFINITY;
  int bi = -1;
  for (int v = sub; v < n; v += G) {
    float dd = dot(hull_vertex(h, v), d);
    if (dd > bv) {
      bv = dd;
      bi = v;
    }
  }
  static_assert(G == 4 || G == 8 || G == 16, "group of 4, 8 or 16 lanes within a DPP row");
  row_reduce<true, G>(bv, bi);
  return bi >= 0 ? hull_vertex(h, bi) : splat(0.0f);
}

template <int G>
__device__ __forceinline__ V3 group_core_support_world(const ShapeRef& s, V3 p, Q4 q, V3 d, int sub) {
  V3 ld = qrot(qconj(q), d);
  V3 l = s.type == HULL ? group_hull_support<G>(s.hull, s.nhull, ld, sub) : core_support_local(s, ld);
  return qrot(q, l) + p;
}

// The same support with the index of the local support point (hull vertex,
// core box corner by sign bits, capsule tip, sphere centre), for walks whose
// adjoint rebuilds the support points.
template <int G>
__device__ __forceinline__ V3 group_core_support_world_idx(const ShapeRef& s, V3 p, Q4 q, V3 d, int sub, int& idx) {
  V3 ld = qrot(qconj(q), d);
  V3 l;
  if (s.type == HULL) {
    float bv = -INFINITY;
    int bi = -1;
    for (int v = sub; v < s.nhull; v += G) {
      float dd = dot(hull_vertex(s.hull, v), ld);
      if (dd > bv) {
        bv = dd;
        bi = v;
      }
    }
    row_reduce<true, G>(bv, bi);
    l = bi >= 0 ? hull_vertex(s.hull, bi) : splat(0.0f);
    idx = bi >= 0 ? bi : 0;
  } else {
    l = core_support_local(s, ld);
    idx = s.type == BOX ? ((ld.x >= 0.0f ? 0 : 1) | (ld.y >= 0.0f ? 0 : 2) | (ld.z >= 0.0f ? 0 : 4))
                        : (s.type == CAPSULE ? (ld.y >= 0.0f ? 0 : 1) : 0);
  }
  return qrot(q, l) + p;
}

template <int G>
struct GroupSupportIdx {
  const Pair* pr;
  int sub;
  __device__ SV operator()(V3 d) const {
    int ia, ib;
    V3 a = group_core_support_world_idx<G>(pr->a, pr->pa, pr->qa, d, sub, ia);
    V3 b = group_core_support_world_idx<G>(pr->b, pr->pb, pr->qb, -d, sub, ib);
    return SV{a - b, b, ia, ib};
  }
};

template <int G>
struct GroupSupport {
  const Pair* pr;
  int sub;
  __device__ SV operator()(V3 d) const {
    V3 a = group_core_support_world<G>(pr->a, pr->pa, pr->qa, d, sub);
    V3 b = group_core_support_world<G>(pr->b, pr->pb, pr->qb, -d, sub);
    return SV{a - b, b};
  }
};

// Pairs with a hull (class CLS_GJK_HULL), G lanes per pair: every lane of a
// group runs the same walk; the hull support scans are split over the group.
// The walk is latency-bound, so occupancy pays: capped at 128 VGPRs, G >= 8
// runs four waves per SIMD (three uncapped, at 158 VGPRs), with a few
// spills: 678,321 hull pairs of the 1M-body mixed scene 1.03 -> 0.92 ms (at
// five waves, 96 VGPRs, the spills make it 1.13 ms).  G = 4 stages 64 KB per
// block, two blocks per CU, so it stays uncapped.
#ifndef BCH_NP_HULL_WAVES
#define BCH_NP_HULL_WAVES 4
#endif
template <int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(G >= 8 ? BCH_NP_HULL_WAVES : 1))) void k_np_hull(const Counters* __restrict__ cnt_ro, const uint32_t* __restrict__ perm,
                                                 const uint32_t* __restrict__ pairs, const BodyRec* __restrict__ bodies,
                                                 const float4* __restrict__ hull, const float* __restrict__ seeds,
                                                 uint32_t* __restrict__ ccount, float4* __restrict__ cslots,
                                                 float4* __restrict__ epa_q, Counters* cnt) {
  // One hull of the pair (A if it is a hull of <= 64 vertices, else B) is
  // staged in LDS, one slot per group; a second hull is read through the
  // cache.  32 KB per block at G = 8 keeps three blocks per CU, where
  // staging both (64 KB) allowed two.
  __shared__ float4 hstage[256 / G][64];
  uint32_t beg = cnt_ro->cls_start[CLS_GJK_HULL], end = cnt_ro->cls_end[CLS_GJK_HULL];
  const uint32_t per_block = blockDim.x / G;
  int grp = threadIdx.x / G, sub = threadIdx.x % G;
  for (uint32_t base = beg + blockIdx.x * per_block; base < end; base += gridDim.x * per_block) {
    uint32_t t = base + grp;
    bool need_epa = false;
    uint32_t p = 0;
    GjkResult g;
    if (t < end) {
      p = perm[t];
      uint32_t i = pairs[2 * p], j = pairs[2 * p + 1];
      BodyRec A = bodies[i], B = bodies[j];
      Pair pr = make_pair(A, B, hull);
      if (pr.a.type == HULL && pr.a.nhull <= 64) {
        for (int v = sub; v < pr.a.nhull; v += G) hstage[grp][v] = pr.a.hull[v];
        pr.a.hull = hstage[grp];
      } else if (pr.b.type == HULL && pr.b.nhull <= 64) {
        for (int v = sub; v < pr.b.nhull; v += G) hstage[grp][v] = pr.b.hull[v];
        pr.b.hull = hstage[grp];
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      float thr = mn(A.threshold, B.threshold);
      V3 sd{0.0f, 0.0f, 0.0f};
      bool has_seed = seeds != nullptr && !is_nan(seeds[3 * p]);  // NaN x: None
      if (has_seed) sd = V3{seeds[3 * p], seeds[3 * p + 1], seeds[3 * p + 2]};
      g = gjk_contact_t(pr, has_seed, sd, thr, GroupSupport<G>{&pr, sub});
      need_epa = g.kind == 2;
      if (sub == 0) {
        Manifold4 m;
        m.n = 0;
        if (g.kind == 1) m.push(g.point_on_b, g.normal, -g.distance);
        ccount[p] = (uint32_t)m.n;
        if (!need_epa) store_contacts(cslots, p, m);
      }
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
        uint32_t q = qb + __popcll(mask & ((1ull << lane) - 1ull));
        epa_q[4 * (uint64_t)q] = make_float4(g.t0.x, g.t0.y, g.t0.z, u2f(p));
        epa_q[4 * (uint64_t)q + 1] = make_float4(g.t1.x, g.t1.y, g.t1.z, 0.0f);
        epa_q[4 * (uint64_t)q + 2] = make_float4(g.t2.x, g.t2.y, g.t2.z, 0.0f);
        epa_q[4 * (uint64_t)q + 3] = make_float4(g.t3.x, g.t3.y, g.t3.z, 0.0f);
      }
    }
  }
}

// Dense contacts: contact c of pair p -> contacts[7 * (offsets[p] + c)];
// contacts past `cap` are dropped and flagged.
__global__ void k_gather(uint32_t P, const uint32_t* __restrict__ P_dev, const uint32_t* __restrict__ ccount,
                         const uint32_t* __restrict__ offsets, const float4* __restrict__ cslots, uint64_t cap,
                         float* __restrict__ contacts, Counters* cnt) {
  P = dev_count(P, P_dev);
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint32_t c = ccount[p], o = offsets[p];
  if ((uint64_t)o + c > cap) {
    atomicOr(&cnt->errors, 16u);
    c = (uint64_t)o < cap ? (uint32_t)(cap - o) : 0u;
  }
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

__global__ void k_reset_tree(Counters* c) {
  if (threadIdx.x == 0) {
    c->wave_over = 0;
    c->n_big_pairs = 0;
  }
}
__global__ void k_reset_epa(Counters* c) {
  if (threadIdx.x == 0) {
    c->n_epa = 0;
    c->n_epa_defer = 0;
    c->n_epa_over = 0;
    c->errors &= ~2u;
    for (int k = 0; k < 8; ++k) c->cls_start[k] = c->cls_end[k] = 0u;
  }
}

__global__ void k_fill_u32(uint32_t n, uint32_t* p, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace bch
