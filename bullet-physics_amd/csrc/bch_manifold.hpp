// bch_manifold.hpp -- persistent contact manifolds on the device: the
// PersistentManifold of crates/bullet-collision/src/manifold.rs:124-386 and
// the rigid-pair arm of the world step's pair pass
// (crates/bullet-dynamics/src/world/step.rs:660-854), one thread per pair.
//
// The cache is kept per pair of the step that wrote it: the step's sorted
// pair keys, a valid flag and one 256-B entry per pair.  The next step finds
// a pair's entry by binary search in the previous keys; entries of pairs
// that left the pair list are not carried over (the prune of
// step.rs:1006-1010).
//
//   k_mf_pre   lookup, filters, sleeping / static pairs, the refresh path;
//              pairs that need detection keep their shape class and get the
//              cache's seed (NaN x = None), all others get class 7, which no
//              narrowphase kernel runs
//   (narrowphase over the detection pairs)
//   k_mf_post  capture + refresh of the detected manifold, or removal
//   scan + k_mf_gather   dense emitted points
//
// Entry layout (16 float4): point k < 4 at [3k] (local_a, penetration),
// [3k + 1] (local_b, normal impulse), [3k + 2] (normal, lifetime bits);
// [12] (detect position A, threshold), [13] detect orientation A,
// [14] (detect position B, n | initialized << 8 bits), [15] detect
// orientation B.
// Emitted point (3 float4): (position, normal.x), (normal.yz, penetration,
// impulse x warm-start factor), (lifetime bits, A-side anchor).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bch {

constexpr float kRedriftT = 0.005f;                       // manifold.rs:146
constexpr float kRedriftT2 = kRedriftT * kRedriftT;       // manifold.rs:147
constexpr float kRedriftAngle = 0.01f;                    // manifold.rs:148
constexpr int kMfCache = 4;                               // manifold.rs:151
constexpr uint32_t kMfNoDetect = 7;                       // shape class no kernel runs

enum : uint8_t { MF_STATIC = 1, MF_ACTIVE = 2, MF_FAST = 4 };
// Routes; the two pending ones exist between the pre and post kernels
// (MR_PENDING_REFRESHED: the entry was refreshed before detection).
enum : uint8_t { MR_FILTERED = 0, MR_SLEEPING = 1, MR_STATIC = 2, MR_SERVED = 3, MR_DETECTED = 4, MR_REMOVED = 5,
                 MR_PENDING = 6, MR_PENDING_REFRESHED = 7 };
constexpr uint32_t kMfNone = 0xffffffffu;

struct MfPoint {
  V3 la, lb, n;
  float pen, imp;
  uint32_t life;
};

// The cache's point array is indexed by compile-time constants only
// (unrolled loops; a store or load at a run-time position is a select over
// the slots), so a manifold stays in registers instead of scratch.
struct MfMan {
  MfPoint pt[kMfCache + 4];  // capture appends up to four before truncating
  int n;
  V3 dpa, dpb;
  Q4 dqa, dqb;
  uint32_t init;
  float thr;
};

// glam Quat * Quat (scalar order); only w is read by the redetect test.
BCH_HD float qmul_w(Q4 a, Q4 b) { return ((a.w * b.w - a.x * b.x) - a.y * b.y) - a.z * b.z; }

__device__ __forceinline__ float rotation_angle_w(float w) {  // manifold.rs:154-158
  w = w < -1.0f ? -1.0f : (w > 1.0f ? 1.0f : w);
  return 2.0f * acosf(w);
}

// dst = v where c (component-wise: no struct select)
__device__ __forceinline__ void mf_put(MfPoint& dst, bool c, const MfPoint& v) {
  dst.la = sel(c, v.la, dst.la);
  dst.lb = sel(c, v.lb, dst.lb);
  dst.n = sel(c, v.n, dst.n);
  dst.pen = c ? v.pen : dst.pen;
  dst.imp = c ? v.imp : dst.imp;
  dst.life = c ? v.life : dst.life;
}

__device__ __forceinline__ void mf_clear_points(MfMan& m) {
  MfPoint z;
  z.la = z.lb = z.n = splat(0.0f);
  z.pen = z.imp = 0.0f;
  z.life = 0u;
#pragma unroll
  for (int k = 0; k < kMfCache + 4; ++k) m.pt[k] = z;
}

__device__ __forceinline__ void mf_empty(MfMan& m) {
  mf_clear_points(m);
  m.n = 0;
  m.dpa = m.dpb = splat(0.0f);
  m.dqa = m.dqb = Q4{0.0f, 0.0f, 0.0f, 1.0f};
  m.init = 0;
  m.thr = 0.02f;  // CONTACT_BREAKING_THRESHOLD, manifold.rs:8
}

__device__ __forceinline__ void mf_load(const float4* __restrict__ e, MfMan& m) {
  float4 a = e[12], b = e[13], c = e[14], d = e[15];
  uint32_t nb = f2u(c.w);
  m.n = (int)(nb & 0xffu);
  m.init = (nb >> 8) & 1u;
  m.dpa = V3{a.x, a.y, a.z};
  m.thr = a.w;
  m.dqa = Q4{b.x, b.y, b.z, b.w};
  m.dpb = V3{c.x, c.y, c.z};
  m.dqb = Q4{d.x, d.y, d.z, d.w};
  mf_clear_points(m);
#pragma unroll
  for (int k = 0; k < kMfCache; ++k) {
    if (k >= m.n) break;
    float4 p0 = e[3 * k], p1 = e[3 * k + 1], p2 = e[3 * k + 2];
    m.pt[k].la = V3{p0.x, p0.y, p0.z};
    m.pt[k].pen = p0.w;
    m.pt[k].lb = V3{p1.x, p1.y, p1.z};
    m.pt[k].imp = p1.w;
    m.pt[k].n = V3{p2.x, p2.y, p2.z};
    m.pt[k].life = f2u(p2.w);
  }
}

__device__ __forceinline__ void mf_store(const MfMan& m, float4* __restrict__ e) {
#pragma unroll
  for (int k = 0; k < kMfCache; ++k) {
    if (k < m.n) {
      const MfPoint& p = m.pt[k];
      e[3 * k] = make_float4(p.la.x, p.la.y, p.la.z, p.pen);
      e[3 * k + 1] = make_float4(p.lb.x, p.lb.y, p.lb.z, p.imp);
      e[3 * k + 2] = make_float4(p.n.x, p.n.y, p.n.z, u2f(p.life));
    } else {
      e[3 * k] = e[3 * k + 1] = e[3 * k + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  e[12] = make_float4(m.dpa.x, m.dpa.y, m.dpa.z, m.thr);
  e[13] = make_float4(m.dqa.x, m.dqa.y, m.dqa.z, m.dqa.w);
  e[14] = make_float4(m.dpb.x, m.dpb.y, m.dpb.z, u2f((uint32_t)m.n | (m.init << 8)));
  e[15] = make_float4(m.dqb.x, m.dqb.y, m.dqb.z, m.dqb.w);
}

// refresh (manifold.rs:252-298); the survivors keep their order.
__device__ __forceinline__ int mf_refresh(MfMan& m, V3 pa, Q4 qa, V3 pb, Q4 qb) {
  float thr = m.thr;
  float thr2 = thr * thr;
  int w = 0;
#pragma unroll
  for (int r = 0; r < kMfCache; ++r) {
    if (r >= m.n) break;
    MfPoint p = m.pt[r];
    V3 wa = pa + qrot(qa, p.la);
    V3 wb = pb + qrot(qb, p.lb);
    V3 d = wa - wb;
    float pen = dot(d, p.n);
    V3 lat = d - p.n * pen;
    bool keep = !(pen < -thr) && !(len2(lat) > thr2);
    p.pen = pen;
    p.life += 1u;
    // survivor r goes to slot w <= r (slots above r are read later, never written)
#pragma unroll
    for (int k = 0; k <= r; ++k) mf_put(m.pt[k], keep && k == w, p);
    w += keep ? 1 : 0;
  }
  m.n = w;
  return w;
}

// needs_redetect (manifold.rs:300-318)
__device__ __forceinline__ bool mf_needs_redetect(const MfMan& m, V3 pa, Q4 qa, V3 pb, Q4 qb) {
  if (!m.init) return true;
  V3 drift = (pb - m.dpb) - (pa - m.dpa);
  if (len2(drift) > kRedriftT2) return true;
  float wa = qmul_w(qa, qconj(m.dqa));
  float wb = qmul_w(qb, qconj(m.dqb));
  return rotation_angle_w(wa) + rotation_angle_w(wb) > kRedriftAngle;
}

// f32::total_cmp as a signed key
__device__ __forceinline__ int64_t total_key(float f) {
  uint32_t u = f2u(f);
  return (u & 0x80000000u) ? -(int64_t)(u & 0x7fffffffu) - 1 : (int64_t)u;
}

// capture (manifold.rs:320-376) of n contacts from the per-pair slots
// (position, depth), (normal, -).
__device__ __forceinline__ void mf_capture(MfMan& m, const float4* __restrict__ slots, int n, V3 pa, Q4 qa, V3 pb,
                                           Q4 qb, float threshold) {
  m.thr = threshold;
  Q4 ia = qconj(qa), ib = qconj(qb);
  float thr2 = threshold * threshold;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= n) break;
    float4 s0 = slots[2 * k], s1 = slots[2 * k + 1];
    V3 wb{s0.x, s0.y, s0.z};
    V3 nrm{s1.x, s1.y, s1.z};
    float depth = s0.w;
    V3 wa = wb + nrm * depth;
    MfPoint pt;
    pt.la = qrot(ia, wa - pa);
    pt.lb = qrot(ib, wb - pb);
    pt.n = nrm;
    pt.pen = depth;
    pt.imp = 0.0f;
    pt.life = 0u;
    // first cached point within the threshold with a like normal (its
    // impulse is inherited), else a new slot at the end
    int slot = -1;
    float imp = 0.0f;
#pragma unroll
    for (int o = 0; o < kMfCache + 4; ++o) {
      bool hit = slot < 0 && o < m.n && len2(m.pt[o].la - pt.la) < thr2 && dot(m.pt[o].n, pt.n) >= 0.9f;
      imp = hit ? m.pt[o].imp : imp;
      slot = hit ? o : slot;
    }
    pt.imp = imp;
    int target = slot >= 0 ? slot : m.n;
#pragma unroll
    for (int o = 0; o < kMfCache + 4; ++o) mf_put(m.pt[o], o == target, pt);
    m.n += slot >= 0 ? 0 : 1;
  }
  if (m.n > kMfCache) {
    // stable insertion sort, deepest first (total order), step by step:
    // item a moves down while the item before it has the smaller key
#pragma unroll
    for (int a = 1; a < kMfCache + 4; ++a) {
      if (a >= m.n) break;
      MfPoint x = m.pt[a];
      int64_t kx = total_key(x.pen);
      bool moving = true;
#pragma unroll
      for (int b = a - 1; b >= 0; --b) {
        bool shift = moving && total_key(m.pt[b].pen) < kx;
        bool place = moving && !shift;
        MfPoint y = m.pt[b];
        mf_put(m.pt[b + 1], shift, y);
        mf_put(m.pt[b + 1], place, x);
        moving = shift;
      }
      mf_put(m.pt[0], moving, x);
    }
    m.n = kMfCache;
  }
  m.dpa = pa;
  m.dqa = qa;
  m.dpb = pb;
  m.dqb = qb;
  m.init = 1u;
}

// Emit the refreshed points (step.rs:748-760, 818-830).
__device__ __forceinline__ void mf_emit(const MfMan& m, V3 pa, Q4 qa, V3 pb, Q4 qb, float warm,
                                        float4* __restrict__ out) {
#pragma unroll
  for (int k = 0; k < kMfCache; ++k) {
    if (k >= m.n) break;
    const MfPoint& p = m.pt[k];
    V3 wb = pb + qrot(qb, p.lb);
    V3 anchor = pa + qrot(qa, p.la);
    out[3 * k] = make_float4(wb.x, wb.y, wb.z, p.n.x);
    out[3 * k + 1] = make_float4(p.n.y, p.n.z, p.pen, p.imp * warm);
    out[3 * k + 2] = make_float4(u2f(p.life), anchor.x, anchor.y, anchor.z);
  }
}

__device__ __forceinline__ int64_t mf_lookup(const uint64_t* __restrict__ keys, uint32_t M, uint64_t key) {
  uint32_t lo = 0, hi = M;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < M && keys[lo] == key) ? (int64_t)lo : -1;
}

// Per pair, before the narrowphase.  src[p] = the previous step's entry
// index of the pair, kMfNone when it had none.
__global__ __launch_bounds__(256) void k_mf_pre(uint32_t P, const uint64_t* __restrict__ keys, const BodyRec* __restrict__ bodies,
                         const uint8_t* __restrict__ bflags, const uint32_t* __restrict__ group_mask,
                         const uint64_t* __restrict__ okeys, const uint8_t* __restrict__ ovalid,
                         const float4* __restrict__ oent, uint32_t M, float slop, float warm,
                         uint32_t* __restrict__ cls, uint32_t* __restrict__ ccount, float* __restrict__ seeds,
                         uint8_t* __restrict__ route, uint32_t* __restrict__ src, uint8_t* __restrict__ nvalid,
                         float4* __restrict__ nent, uint32_t* __restrict__ ocnt, float4* __restrict__ oslots) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint64_t key = keys[p];
  uint32_t i = (uint32_t)(key >> 32), j = (uint32_t)key;
  ocnt[p] = 0;
  src[p] = kMfNone;
  nvalid[p] = 0;
  uint8_t r = MR_PENDING;
  if (group_mask) {  // step.rs:660-664
    uint32_t ga = group_mask[i], gb = group_mask[j];
    if (((ga & 0xffffu) & (gb >> 16)) == 0u || ((gb & 0xffffu) & (ga >> 16)) == 0u) r = MR_FILTERED;
  }
  int64_t e = -1;
  if (r == MR_PENDING) {
    e = mf_lookup(okeys, M, key);
    if (e >= 0 && !ovalid[e]) e = -1;
    uint8_t fa = bflags ? bflags[i] : (uint8_t)MF_ACTIVE, fb = bflags ? bflags[j] : (uint8_t)MF_ACTIVE;
    if (!(fa & MF_ACTIVE) && !(fb & MF_ACTIVE)) r = MR_SLEEPING;          // step.rs:683-704
    else if ((fa & MF_STATIC) && (fb & MF_STATIC)) r = MR_STATIC;         // step.rs:711-713
    if (r != MR_PENDING) {
      if (e >= 0) {
        for (int k = 0; k < 16; ++k) nent[16 * (uint64_t)p + k] = oent[16 * (uint64_t)e + k];
        nvalid[p] = 1;
      }
    } else {
      BodyRec A = bodies[i], B = bodies[j];
      V3 pa = pos_of(A), pb = pos_of(B);
      Q4 qa = quat_of(A), qb = quat_of(B);
      bool fast = ((fa | fb) & MF_FAST) != 0;
      MfMan m;
      bool refreshed = false;
      if (e >= 0) {
        mf_load(oent + 16 * (uint64_t)e, m);
        if (!mf_needs_redetect(m, pa, qa, pb, qb)) {  // step.rs:731-776
          int n = mf_refresh(m, pa, qa, pb, qb);
          bool deep = false;
#pragma unroll
          for (int k = 0; k < kMfCache; ++k)
            if (k < n) deep = deep || m.pt[k].pen > slop;
          if (!fast || !deep) {
            mf_emit(m, pa, qa, pb, qb, warm, oslots + 12 * (uint64_t)p);
            ocnt[p] = (uint32_t)n;
            mf_store(m, nent + 16 * (uint64_t)p);
            nvalid[p] = 1;
            r = MR_SERVED;
          } else {
            refreshed = true;
          }
        }
      }
      if (r == MR_PENDING) {  // detection with the cache's seed (step.rs:777-786)
        float sx = NAN, sy = 0.0f, sz = 0.0f;
        if (e >= 0 && m.n > 0 && len2(m.pt[0].n) > 0.5f) {  // seed, manifold.rs:378-385
          sx = m.pt[0].n.x;
          sy = m.pt[0].n.y;
          sz = m.pt[0].n.z;
        }
        seeds[3 * (uint64_t)p] = sx;
        seeds[3 * (uint64_t)p + 1] = sy;
        seeds[3 * (uint64_t)p + 2] = sz;
        src[p] = e >= 0 ? (uint32_t)e : kMfNone;
        if (refreshed) r = MR_PENDING_REFRESHED;
      }
    }
  }
  route[p] = r;
  if (r != MR_PENDING && r != MR_PENDING_REFRESHED) {
    cls[p] = cls_key(kMfNoDetect, 0u);
    ccount[p] = 0u;
  }
}

// Per pair after the narrowphase: capture + refresh, or removal
// (step.rs:787-852).
__global__ __launch_bounds__(256) void k_mf_post(uint32_t P, const uint64_t* __restrict__ keys, const BodyRec* __restrict__ bodies,
                          const float4* __restrict__ oent, const uint32_t* __restrict__ ccount,
                          const float4* __restrict__ cslots, float warm, uint8_t* __restrict__ route,
                          const uint32_t* __restrict__ src, uint8_t* __restrict__ nvalid, float4* __restrict__ nent,
                          uint32_t* __restrict__ ocnt, float4* __restrict__ oslots) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint8_t r = route[p];
  if (r != MR_PENDING && r != MR_PENDING_REFRESHED) return;
  uint64_t key = keys[p];
  uint32_t i = (uint32_t)(key >> 32), j = (uint32_t)key;
  uint32_t nc = ccount[p];
  if (nc == 0u) {  // collide_seeded returned None: drop the entry
    route[p] = MR_REMOVED;
    nvalid[p] = 0;
    ocnt[p] = 0;
    return;
  }
  BodyRec A = bodies[i], B = bodies[j];
  V3 pa = pos_of(A), pb = pos_of(B);
  Q4 qa = quat_of(A), qb = quat_of(B);
  MfMan m;
  uint32_t e = src[p];
  if (e == kMfNone) {
    mf_empty(m);
  } else {
    mf_load(oent + 16 * (uint64_t)e, m);
    if (r == MR_PENDING_REFRESHED) (void)mf_refresh(m, pa, qa, pb, qb);  // step.rs:737-745 ran first
  }
  mf_capture(m, cslots + 8 * (uint64_t)p, (int)nc, pa, qa, pb, qb, mn(A.threshold, B.threshold));
  int n = mf_refresh(m, pa, qa, pb, qb);
  mf_emit(m, pa, qa, pb, qb, warm, oslots + 12 * (uint64_t)p);
  ocnt[p] = (uint32_t)n;
  mf_store(m, nent + 16 * (uint64_t)p);
  nvalid[p] = 1;
  route[p] = MR_DETECTED;
}

// Dense emitted points: point k of pair p -> out[3 * (offsets[p] + k)].
__global__ void k_mf_gather(uint32_t P, const uint32_t* __restrict__ ocnt, const uint32_t* __restrict__ offsets,
                            const float4* __restrict__ oslots, float4* __restrict__ out) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  uint32_t c = ocnt[p], o = offsets[p];
  for (uint32_t k = 0; k < c; ++k)
    for (int q = 0; q < 3; ++q) out[3 * (uint64_t)(o + k) + q] = oslots[12 * (uint64_t)p + 3 * k + q];
}

// Post-solve impulse writeback (step.rs:978-1005): impulse of emitted point
// offsets[p] + k into cached point k of pair p.
__global__ void k_mf_impulses(uint32_t P, const uint32_t* __restrict__ ocnt, const uint32_t* __restrict__ offsets,
                              const uint8_t* __restrict__ nvalid, const float* __restrict__ impulses,
                              float4* __restrict__ nent) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P || !nvalid[p]) return;
  uint32_t c = ocnt[p], o = offsets[p];
  for (uint32_t k = 0; k < c; ++k) nent[16 * (uint64_t)p + 3 * k + 1].w = impulses[o + k];
}

// per-u8 fill (route / count buffers are bytes)
__global__ void k_fill_u8(uint64_t n, uint8_t* p, uint8_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace bch
