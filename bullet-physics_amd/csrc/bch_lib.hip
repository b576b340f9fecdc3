// bch_lib.hip -- context, pass orchestration and the C ABI
// (include/bullet_collision_hip.h).
#include <hip/hip_runtime.h>

#include <cstring>


#include <algorithm>
#include <cstdio>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/bullet_collision_hip.h"
#include "bch_kernels.hpp"
#include "bch_epa_wave.hpp"
#include "bch_epa_dual.hpp"
#include "bch_epa_quad.hpp"
#include "bch_adjoint.hpp"
#include "bch_sort.hpp"
#include "bch_depen.hpp"
#include "bch_manifold.hpp"
#include "bch_mesh.hpp"
#include "bch_sweep.hpp"

using namespace bch;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool borrowed = false;  // a view into memory this buffer does not own: never grown or freed
  bool ensure(size_t need) {
    if (need <= bytes) return true;
    if (borrowed) return false;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t nb = std::max<size_t>(need + need / 4, 256);
    if (hipMalloc(&p, nb) != hipSuccess) {
      p = nullptr;
      return false;
    }
    bytes = nb;
    return true;
  }
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    borrowed = false;
  }
  void view(void* q, size_t nb) {
    release();
    p = q;
    bytes = nb;
    borrowed = true;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

// Page-locked host memory that grows by reallocation, keeping its contents.
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool ensure(size_t need) {
    if (need <= bytes) return true;
    size_t nb = std::max<size_t>(need + need / 2, 4096);
    void* q = nullptr;
    if (hipHostMalloc(&q, nb, hipHostMallocDefault) != hipSuccess) return false;
    if (p) {
      memcpy(q, p, bytes);
      (void)hipHostFree(p);
    }
    p = q;
    bytes = nb;
    return true;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

constexpr int kBlock = 256;
constexpr uint32_t kEpaDeferGrid = 16384;  // one-pair EPA blocks for the deferred entries (~4% of the queue)
constexpr uint32_t kEpaBigGrid = 256;      // blocks for the entries that outgrow the one-pair store
inline uint32_t blocks(uint64_t n, int b = kBlock) { return (uint32_t)((n + b - 1) / b); }

// Working memory of bch_collide_batch_static, the entry point that allocates
// nothing: one static device array per GPU, carved into the buffers the
// narrowphase over caller pairs uses.
constexpr size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kStP = BCH_STATIC_MAX_PAIRS, kStN = BCH_STATIC_MAX_BODIES, kStV = BCH_STATIC_MAX_HULL_VERTICES;
constexpr size_t kStTiles = (kStP + kRsTile - 1) / kRsTile;
constexpr size_t kStBodies = 0;
constexpr size_t kStBtype = kStBodies + al256(kStN * sizeof(BodyRec));
constexpr size_t kStHull = kStBtype + al256(kStN);
constexpr size_t kStKeys = kStHull + al256(kStV * sizeof(float4));
constexpr size_t kStPairs = kStKeys + al256(kStP * sizeof(uint64_t));
constexpr size_t kStU32 = kStPairs + al256(kStP * 2 * sizeof(uint32_t));     // cls, cls_alt, perm, perm_alt
constexpr size_t kStCount = kStU32 + 4 * al256(kStP * sizeof(uint32_t));    // ccount, coffsets
constexpr size_t kStSlots = kStCount + 2 * al256((kStP + 1) * sizeof(uint32_t));
constexpr size_t kStEpa = kStSlots + al256(kStP * 8 * sizeof(float4));       // two EPA queues
constexpr size_t kStCounters = kStEpa + 2 * al256(kStP * 4 * sizeof(float4));
constexpr size_t kStHist = kStCounters + al256(sizeof(Counters));
constexpr size_t kStOffs = kStHist + al256(256 * kStTiles * sizeof(uint32_t));
constexpr size_t kStPart = kStOffs + al256((256 * kStTiles + 256) * sizeof(uint32_t));  // scan partials
constexpr size_t kStParts = (256 * kStTiles + kScanTile - 1) / kScanTile;
constexpr size_t kStBytes = kStPart + al256(kStParts * sizeof(uint32_t));
static_assert(kStBytes == BCH_STATIC_BYTES, "BCH_STATIC_BYTES must equal the carved layout");

}  // namespace

// The array itself.  In host code `scratch` names a host-side stand-in, not
// the device address: the address comes from hipGetSymbolAddress, once per
// device (scratch_init), and only that address is ever handed to kernels.
static __device__ __align__(256) unsigned char scratch[kStBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Resolve the device address of `scratch` on `device` (current); once per device.
int scratch_init(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return BCH_NO_DEVICE;
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BCH_HIP_ERROR;
    g_scratch[device] = p;
  }
  *out = g_scratch[device];
  return BCH_OK;
}
}  // namespace

struct bch_context {
  int device = 0;
  hipStream_t stream = nullptr;      // the stream every call is ordered on
  hipStream_t own_stream = nullptr;  // the context's own stream
  // Second stream (default priority) for the closed-form and non-hull GJK
  // launches, which run beside the hull GJK: forked from and joined back to
  // `stream` with the two events.
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  std::string err;
  hipEvent_t ev[11] = {};  // stage boundaries 0-8; 9 ends the four-pair EPA kernel, 10 the hull GJK kernel
  float timings[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

  // world-resident bodies
  uint32_t n = 0, n_worlds = 1, max_world = 0;
  bool batched = false;
  DevBuf bodies, hull, world_offsets, world_of, staging_a, staging_b, staging_c, staging_d, staging_e, staging_f;
  DevBuf btype;  // u8 shape type per body (the pair split reads it, not the 64-B records)

  // pass buffers
  DevBuf aabb, flags, big_ids, counters, partials, keys, keys_alt, vals, vals_alt, nodes, wend, hull_ids;
  DevBuf bsph;  // per body: position and bound (bounding ball), written by k_aabb each pass
  DevBuf sd_defer;  // signed-distance EPA entries the four-pair kernel hands on
  DevBuf sd_rec;    // the adjoint's EPA faces, one record (4 float4) per queue entry
  DevBuf sd_brec;   // the adjoint walk's Band / touch simplex per pair (kSdBandRec float4)
  uint32_t n_hulls = 0;
  DevBuf world_reg;  // per world: tree bodies, then their exclusive scan (the per-world Morton sort)
  DevBuf walk_regions, wave_counts, wave_offs, pair_keys, pair_keys_alt, pairs, cls, cls_alt, perm,
      perm_alt, ccount, coffsets, cslots, contacts, epa_q, epa_q2, scan_part, rs_hist, rs_offs;
  int hull_lanes = 16;      // lanes per hull GJK pair (4, 8 or 16)
  uint32_t epa_grid = 32768; // EPA blocks (one wavefront each)
  uint32_t np_grid = 2048;  // narrowphase blocks per shape class
  int epa_wave_faces = 256;  // face store the one-pair EPA may fill before the overflow route
  uint32_t trav_block = 256;  // threads per tree-walk block (64, 128, 256, 512)
  uint32_t sort_tile = 2048;  // items per radix-sort tile (2048 or 4096; the static batch keeps 4096)
  bool cls_hist_ready = false;  // the split has written the class sort's tile histograms
  uint64_t tree_cap = 0;      // pair buffer of the eager pass, sized by earlier passes
  // capacities of the next eager pass's device-count run (0: none yet)
  uint64_t eager_P_cap = 0, eager_C_cap = 0;
  uint32_t walk_R = 1024;     // pair slots of one tree-walk wavefront's region
  uint32_t walk_nw = 0;       // wavefronts of the last walk
  bool walk_grow = false;     // a replayed walk overflowed a region: the next capture doubles walk_R
  uint64_t P = 0, C = 0;
  Counters host_cnt{};
  uint32_t n_planes = 0;

  // capacity-mode pass: fixed buffers, counts and overflow flags stay on the
  // device, no host round trip, so the pass can be captured in a HIP graph
  bool cap_mode = false, capturing = false;
  uint64_t P_cap = 0, C_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t realloc_gen = 0, graph_gen = 0;  // a reallocation invalidates the graph
  uint64_t graph_P_cap = 0, graph_C_cap = 0;

  // captured de-penetration step
  hipGraph_t dgraph = nullptr;
  hipGraphExec_t dgraph_exec = nullptr;
  uint64_t dgraph_gen = 0;
  uint32_t dhist_cap = 0;
  DevBuf sd_out, sd_status, pen_pairs, pen_g, g14, body_grad, pen_part, dhist, dstep;

  // persistent manifolds: the cache written by the last pair pass, one entry
  // per pair of that pass (keys, valid flags, 256-B entries), double-buffered
  DevBuf mkeys[2], mvalid[2], ment[2];
  int mcur = 0;     // the cache the next pass reads
  uint64_t mM = 0;  // its pair count
  DevBuf bflags, bgm;
  bool has_bflags = false, has_bgm = false;
  DevBuf mroute, msrc, mcnt, mslots, moff, mpts, mseeds, mimp;
  uint64_t mNP = 0;  // points emitted by the last pass
  bool aabb_valid = false;  // c->aabb holds the previous pass's boxes (kept for inactive bodies)
  DevBuf vel;

  // triangle meshes (host copies; uploaded when changed) and the buffers of
  // the convex-vs-mesh batch
  std::vector<MeshRec> mesh_recs;
  std::vector<float4> mesh_nodes, mesh_tris, mesh_tri_aabb;
  std::vector<uint32_t> mesh_order;
  bool mesh_dirty = false;
  DevBuf d_mesh_recs, d_mesh_nodes, d_mesh_tris, d_mesh_tri_aabb, d_mesh_order;
  DevBuf mx_pairs, mx_poses, mx_seeds, mx_first, mx_count, mx_offs, mx_sub, mx_sub_cnt, mx_sub_slots, mx_epa_q,
      mx_epa_q2, mx_out_count, mx_out_pts;
  DevBuf sw_target, sw_caster, sw_from, sw_to, sw_radius, sw_pose, sw_hit, sw_out, sw_keys, sw_order;

  // Proxy broadphase (held by the batch context below): the arena of slot
  // boxes lives on the device; the host keeps a pinned mirror (lo, hi as two
  // float4 per slot, NaN for a free slot, which overlaps nothing), the live
  // flags, the free slots (lowest first) and the slots changed since the
  // last upload.
  PinnedBuf proxy_box;  // 2 float4 per slot
  uint32_t n_proxy = 0;
  std::vector<uint8_t> proxy_live, proxy_dirty;
  std::set<uint32_t> proxy_free;
  std::vector<uint32_t> dirty_list;
  uint32_t arena_n = 0;  // slots the device arena holds
  bool arena_whole = false;  // so many slots changed that the next sync uploads the whole arena
  DevBuf arena, arena_upd, bp_pairs;
  PinnedBuf arena_stage;
  bool bp_pairs_valid = false;  // the last compute_pairs result still holds (no change since)
  uint64_t bp_count = 0;

  // collide_batch: the body count and hull-vertex count of the shapes
  // uploaded by the last bch_collide_batch (for bch_collide_batch_poses)
  bool batch_shapes = false;
  uint32_t batch_n = 0;
  DevBuf batch_pairs, batch_seeds, batch_out_count, batch_out_pts;

  // The batch entry points (caller-given pairs, mesh and sweep batches, the
  // proxy broadphase) run on this second context of the same device and
  // stream, so they never touch the world-resident state above.
  bch_context* aux = nullptr;
  bool is_aux = false;

  // bch_collide_batch_static runs on this context: its buffers are views into
  // the GPU's static array, its streams and events are borrowed from the
  // owner, and a buffer that would need to grow is an error.
  bch_context* stat = nullptr;
  bool no_alloc = false;
};

namespace {

int fail(bch_context* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                          \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(ctx, BCH_HIP_ERROR, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

#define ENSURE(ctx, buf, want)                                                               \
  do {                                                                                       \
    size_t need_ = (want);                                                                   \
    if (need_ > (buf).bytes) {                                                               \
      if ((ctx)->no_alloc) return fail(ctx, BCH_INVALID_ARG, "input exceeds the static scratch"); \
      if ((ctx)->capturing) return fail(ctx, BCH_HIP_ERROR, "buffer growth during capture"); \
      ++(ctx)->realloc_gen;                                                                  \
    }                                                                                        \
    if (!(buf).ensure(need_)) return fail(ctx, BCH_HIP_ERROR, "device allocation failed");  \
  } while (0)

// Stage times of the last eager pass from its boundary events; the ninth is
// the two-pair EPA kernel alone (stage 6 minus the one-pair launch after it).
void stage_times(bch_context* c) {
  for (int k = 0; k < 8; ++k) (void)hipEventElapsedTime(&c->timings[k], c->ev[k], c->ev[k + 1]);
  (void)hipEventElapsedTime(&c->timings[8], c->ev[6], c->ev[9]);
  (void)hipEventElapsedTime(&c->timings[9], c->ev[5], c->ev[10]);
}

// Stage boundary event (not while capturing: the graph is timed as a whole).
void mark(bch_context* c, int k) {
  if (!c->capturing && !c->no_alloc) (void)hipEventRecord(c->ev[k], c->stream);
}

// Exclusive scan of n u32 values: out[0..n) offsets, out[n] the total.  With
// n_dev the count is min(n, *n_dev), read on the device; total_copy, when
// given, receives the total as well.
int dev_scan(bch_context* c, const uint32_t* in, uint32_t* out, uint64_t n, DevBuf& part, hipStream_t s,
             const uint32_t* n_dev = nullptr, uint32_t* total_copy = nullptr) {
  if (n > 0xffffffffull) return fail(c, BCH_INVALID_ARG, "scan longer than 2^32 - 1");
  uint32_t nb = (uint32_t)std::max<uint64_t>(1, (n + kScanTile - 1) / kScanTile);
  ENSURE(c, part, (size_t)nb * sizeof(uint32_t));
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, s, in, (uint32_t)n, n_dev, part.as<uint32_t>());
  const int prefixed = nb > kScanSelfTiles ? 1 : 0;
  if (prefixed)
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanThreads), 0, s, part.as<uint32_t>(), nb, out, (uint32_t)n, n_dev,
                       total_copy);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanThreads), 0, s, in, (uint32_t)n, n_dev, part.as<uint32_t>(),
                     out, total_copy, prefixed);
  return BCH_OK;
}

// Stable LSD radix sort of n keys (with values when v0 != nullptr) over the
// 8-bit digits starting at the given bit shifts.  Ping-pongs between (k0, v0)
// and (k1, v1); *in_second tells which pair holds the result.
template <class K, int ROUNDS>
int radix_pass(bch_context* c, K* kin, K* kout, uint32_t* vin, uint32_t* vout, bool vals, uint64_t n,
               const uint32_t* n_dev, int shift, int pack, uint32_t tiles, hipStream_t s, bool hist_ready = false) {
  if (!hist_ready)
    hipLaunchKernelGGL((k_rs_hist<K, ROUNDS>), dim3(tiles), dim3(kRsThreads), 0, s, kin, (uint32_t)n, n_dev, shift,
                       pack, tiles, c->rs_hist.as<uint32_t>());
  uint32_t* offs = c->rs_offs.as<uint32_t>();
  uint32_t* rowtot = offs + (size_t)256 * tiles;
  hipLaunchKernelGGL(k_rs_row_scan, dim3(256), dim3(kScanThreads), 0, s, c->rs_hist.as<uint32_t>(), tiles, offs,
                     rowtot);
  if (vals)
    hipLaunchKernelGGL((k_rs_scatter<K, true, ROUNDS>), dim3(tiles), dim3(kRsThreads), 0, s, kin, vin, (uint32_t)n,
                       n_dev, shift, pack, tiles, c->rs_hist.as<uint32_t>(), offs, rowtot, kout, vout);
  else
    hipLaunchKernelGGL((k_rs_scatter<K, false, ROUNDS>), dim3(tiles), dim3(kRsThreads), 0, s, kin, vin, (uint32_t)n,
                       n_dev, shift, pack, tiles, c->rs_hist.as<uint32_t>(), offs, rowtot, kout, vout);
  return BCH_OK;
}

// Stable LSD radix sort of n keys (with values when v0 != nullptr) over the
// 8-bit digits starting at the given bit shifts.  Ping-pongs between (k0, v0)
// and (k1, v1); *in_second tells which pair holds the result.  The tile
// (items per histogram and scatter block) is the context's sort_tile knob.
// With index_vals the values start as the items' indices (v0 is not read);
// with hist_ready the first pass's tile histograms are already in rs_hist.
template <class K>
int dev_radix_sort(bch_context* c, K* k0, K* k1, uint32_t* v0, uint32_t* v1, uint64_t n, const int* shifts,
                   int nshift, hipStream_t s, bool* in_second, const uint32_t* n_dev = nullptr, int pack = 0,
                   bool index_vals = false, bool hist_ready = false) {
  *in_second = false;
  if (n <= 1 || nshift == 0) return BCH_OK;
  if (n > 0xffffffffull) return fail(c, BCH_INVALID_ARG, "sort longer than 2^32 - 1");
  const uint32_t tile = c->sort_tile == 2048 ? 2048u : (uint32_t)kRsTile;
  uint32_t tiles = (uint32_t)((n + tile - 1) / tile);
  ENSURE(c, c->rs_hist, (size_t)256 * tiles * sizeof(uint32_t));
  ENSURE(c, c->rs_offs, ((size_t)256 * tiles + 256) * sizeof(uint32_t));  // offsets, then the digit totals
  K* kin = k0;
  K* kout = k1;
  uint32_t* vin = index_vals ? nullptr : v0;
  uint32_t* vout = v1;
  const bool vals = v0 != nullptr || index_vals;
  for (int p = 0; p < nshift; ++p) {
    const bool hr = hist_ready && p == 0;
    int rc = tile == 2048
                 ? radix_pass<K, 8>(c, kin, kout, vin, vout, vals, n, n_dev, shifts[p], pack, tiles, s, hr)
                 : radix_pass<K, kRsRounds>(c, kin, kout, vin, vout, vals, n, n_dev, shifts[p], pack, tiles, s, hr);
    if (rc) return rc;
    std::swap(kin, kout);
    vin = vout;
    vout = vin == v1 ? v0 : v1;
  }
  *in_second = (nshift & 1) != 0;
  return BCH_OK;
}

// digit shifts covering bits [lo, hi)
int digit_shifts(int lo, int hi, int* out) {
  int n = 0;
  for (int b = lo; b < hi; b += 8) out[n++] = b;
  return n;
}

int bits_for(uint64_t v) {
  int b = 1;
  while ((1ull << b) <= v) ++b;
  return b;
}

// The pair sort's digits: keys are i << 32 | j with i, j < n < 2^b, sorted
// as the packed 2b-bit number i << b | j (rs_digit's pack = b).  In a batch
// every pair lies inside one world, so j - i < max_world <= 2^w, and the
// key is sorted as the (b + w)-bit number i << w | (j - i) when that is
// shorter (pack = w | kPackDelta): 4 digit passes instead of 6 for 4096
// worlds of at most 2048 bodies.
int pair_key_shifts(const bch_context* c, uint32_t n, int* ps, int* pack) {
  int bn = bits_for(n);
  if (c->batched && c->max_world > 0) {
    int bw = bits_for(c->max_world - 1);
    if (bw < bn) {
      *pack = bw | kPackDelta;
      return digit_shifts(0, bn + bw, ps);
    }
  }
  *pack = bn;
  return digit_shifts(0, 2 * bn, ps);
}

// The tree walk of every leaf into its wavefront's region (stage event 3
// follows it), the regions scanned and concatenated into c->pair_keys, then
// the big bodies' brute force appended after them, all up to `cap` keys.
// The tree total lands at c->wave_offs[nw], the big count in
// Counters::n_big_pairs, and a region overflow in Counters::wave_over.
int walk_pairs(bch_context* c, uint32_t n, const uint32_t* d_world_offsets, const float4* nodes, uint32_t N, int L,
               const uint32_t* wend, uint64_t cap) {
  hipStream_t s = c->stream;
  Counters* cnt = c->counters.as<Counters>();
  const uint32_t B = c->trav_block;
  const uint32_t nb = blocks(std::max<uint32_t>(n, 1), B);
  const uint32_t nw = nb * (B / 64);
  ENSURE(c, c->walk_regions, (size_t)nw * c->walk_R * sizeof(uint64_t));
  ENSURE(c, c->wave_counts, ((size_t)nw + 1) * sizeof(uint32_t));
  ENSURE(c, c->wave_offs, ((size_t)nw + 1) * sizeof(uint32_t));
  hipLaunchKernelGGL(k_reset_tree, dim3(1), dim3(64), 0, s, cnt);
  uint64_t* reg = c->walk_regions.as<uint64_t>();
  uint32_t* wc = c->wave_counts.as<uint32_t>();
  const uint32_t R = c->walk_R;
#define BCH_WALK(BS)                                                                                          \
  hipLaunchKernelGGL(k_traverse<BS>, dim3(nb), dim3(BS), 0, s, cnt, nodes, N, L, wend, reg, R, wc, cnt)
  switch (B) {
    case 64: BCH_WALK(64); break;
    case 128: BCH_WALK(128); break;
    case 512: BCH_WALK(512); break;
    default: BCH_WALK(256);
  }
#undef BCH_WALK
  mark(c, 3);
  int rc = dev_scan(c, wc, c->wave_offs.as<uint32_t>(), nw, c->scan_part, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_concat_regions, dim3(nw), dim3(64), 0, s, reg, R, wc, c->wave_offs.as<uint32_t>(), cap,
                     c->pair_keys.as<uint64_t>());
  // k_big reads the big-body count on the device; an empty list exits at once
  uint32_t span = c->batched ? c->max_world : n;
  uint32_t gx = std::max<uint32_t>(1u, std::min<uint32_t>((span + kBigChunk - 1) / kBigChunk, 1024u));
  uint32_t gy = std::max<uint32_t>(1u, std::min<uint32_t>(c->n_planes, 65535u));
  hipLaunchKernelGGL(k_big, dim3(gx, gy), dim3(kBlock), 0, s, cnt, c->big_ids.as<uint32_t>(), c->aabb.as<float4>(),
                     c->flags.as<uint8_t>(), c->bodies.as<BodyRec>(), d_world_offsets, n, c->pair_keys.as<uint64_t>(),
                     cap, c->wave_offs.as<uint32_t>() + nw, cnt);
  c->walk_nw = nw;
  return BCH_OK;
}

// Capacity-mode tail of the broadphase: the walk into the fixed pair buffer,
// the held / found counts on the device, and the sort over the held count.
// Nothing is read back.
int broadphase_capacity(bch_context* c, uint32_t n, const uint32_t* d_world_offsets, const float4* nodes, uint32_t N,
                        int L, const uint32_t* wend) {
  hipStream_t s = c->stream;
  Counters* cnt = c->counters.as<Counters>();
  uint64_t cap = c->P_cap;
  ENSURE(c, c->pair_keys, cap * sizeof(uint64_t));
  ENSURE(c, c->pair_keys_alt, cap * sizeof(uint64_t));
  int rc = walk_pairs(c, n, d_world_offsets, nodes, N, L, wend, cap);
  if (rc) return rc;
  hipLaunchKernelGGL(k_pair_counts, dim3(1), dim3(64), 0, s, cnt, c->wave_offs.as<uint32_t>() + c->walk_nw, cap,
                     c->walk_R);
  c->P = cap;
  int ps[16];
  int pack = 0;
  int np = pair_key_shifts(c, n, ps, &pack);
  bool palt = false;
  rc = dev_radix_sort<uint64_t>(c, c->pair_keys.as<uint64_t>(), c->pair_keys_alt.as<uint64_t>(), nullptr, nullptr, cap,
                                ps, np, s, &palt, &cnt->n_pairs, pack);
  if (rc) return rc;
  if (palt) std::swap(c->pair_keys, c->pair_keys_alt);
  return BCH_OK;
}

// Broadphase over the AABBs already in c->aabb (n boxes).  Leaves the sorted
// pair keys in c->pair_keys and c->P set (in capacity mode c->P is the
// capacity and the held count is Counters::n_pairs on the device).
int run_broadphase(bch_context* c, uint32_t n, const uint32_t* d_world_of, const uint32_t* d_world_offsets,
                   bool given_boxes) {
  hipStream_t s = c->stream;
  Counters* cnt = c->counters.as<Counters>();
  if (given_boxes) {
    hipLaunchKernelGGL(k_aabb_given, dim3(blocks(n)), dim3(kBlock), 0, s, n, c->aabb.as<float4>(),
                       c->flags.as<uint8_t>(), c->big_ids.as<uint32_t>(), cnt, c->partials.as<float>());
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(1024), 0, s, c->partials.as<float>(), blocks(n),
                       (const float*)nullptr, 0u, cnt);
  }
  hipLaunchKernelGGL(k_morton, dim3(blocks(n)), dim3(kBlock), 0, s, n, c->aabb.as<float4>(), c->flags.as<uint8_t>(),
                     d_world_of, cnt, c->keys.as<uint64_t>(), c->vals.as<uint32_t>());
  int wbits = d_world_of ? bits_for(c->n_worlds) : 0;
  int end_bit = std::min(64, 31 + wbits);
  // all-ones keys of excluded bodies still sort last within end_bit bits.
  // The lowest code bits are not sorted: they leave 2^(30 - drop) >= 8 m
  // sorted cells for the m bodies of the largest world (drop = 7 at 1M
  // bodies in one world, 16 for worlds of at most 2048 bodies), so a cell
  // holds ~0.1 bodies of a world whose bodies spread over the scene bounds,
  // the leaf order (and the tree) barely changes, and the pair set is exact
  // for any leaf order.  At 1M bodies that is 3 digit passes instead of 4,
  // and 4 instead of 5 for 4096 worlds.
  uint32_t m = d_world_of ? std::max<uint32_t>(c->max_world, 1u) : std::max<uint32_t>(n, 1u);
  int drop = 30 - bits_for(8ull * m - 1);
  drop = std::max(0, std::min(drop, 24));
  int shifts[8];
  int ns = digit_shifts(drop, end_bit, shifts);
  bool alt = false, seg = false;
  int rc = BCH_OK;
  if (d_world_of && d_world_offsets && c->max_world <= kSegMax) {
    // worlds of at most kSegMax bodies: one block sorts each world (same order)
    const uint32_t nw = c->n_worlds;
    ENSURE(c, c->world_reg, (size_t)2 * (nw + 1) * sizeof(uint32_t));
    uint32_t* cntw = c->world_reg.as<uint32_t>();
    hipLaunchKernelGGL(k_seg_count, dim3(nw), dim3(256), 0, s, c->keys.as<uint64_t>(), d_world_offsets, cntw);
    rc = dev_scan(c, cntw, cntw + nw + 1, nw, c->scan_part, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_seg_sort, dim3(nw), dim3(256), 0, s, c->keys.as<uint64_t>(), d_world_offsets, cntw + nw + 1,
                       nw, drop, c->keys_alt.as<uint64_t>(), c->vals_alt.as<uint32_t>(), c->wend.as<uint32_t>());
    alt = true;
    seg = true;
  } else {
    rc = dev_radix_sort<uint64_t>(c, c->keys.as<uint64_t>(), c->keys_alt.as<uint64_t>(), c->vals.as<uint32_t>(),
                                  c->vals_alt.as<uint32_t>(), n, shifts, ns, s, &alt);
    if (rc) return rc;
  }
  const uint64_t* skeys = alt ? c->keys_alt.as<uint64_t>() : c->keys.as<uint64_t>();
  const uint32_t* order = alt ? c->vals_alt.as<uint32_t>() : c->vals.as<uint32_t>();
  int L = 0;
  while ((1u << L) < n) ++L;
  uint32_t N = 1u << L;
  float4* nodes = c->nodes.as<float4>();
  hipLaunchKernelGGL(k_tree_leaves, dim3(blocks(N)), dim3(kBlock), 0, s, N, cnt, order, c->aabb.as<float4>(), nodes);
  for (int lc = L; lc > 0;) {
    int m = std::min(9, lc);
    uint32_t per = 1u << m;
    hipLaunchKernelGGL(k_tree_up, dim3((1u << lc) / per), dim3(std::max<uint32_t>(per / 2, 1u)), 0, s, nodes,
                       1u << lc, m);
    lc -= m;
  }
  uint32_t* wend = nullptr;
  if (d_world_of) {
    wend = c->wend.as<uint32_t>();
    if (!seg) hipLaunchKernelGGL(k_world_end, dim3(blocks(n)), dim3(kBlock), 0, s, cnt, skeys, wend);
  }
  mark(c, 2);
  if (c->cap_mode) return broadphase_capacity(c, n, d_world_offsets, nodes, N, L, wend);
  // The pair buffer and the wavefronts' regions are sized by earlier passes
  // (5 pairs per body, 1024 per wavefront at first); a walk whose
  // wavefront overflowed its region runs again with a region that holds
  // it, and a list longer than the buffer is concatenated again.
  uint64_t cap = std::max<uint64_t>(c->tree_cap, std::max<uint64_t>(5ull * n, 1024));
  uint32_t tree_total = 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    ENSURE(c, c->pair_keys, cap * sizeof(uint64_t));
    ENSURE(c, c->pair_keys_alt, cap * sizeof(uint64_t));
    int rc2 = walk_pairs(c, n, d_world_offsets, nodes, N, L, wend, cap);
    if (rc2) return rc2;
    HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipMemcpyAsync(&tree_total, c->wave_offs.as<uint32_t>() + c->walk_nw, sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    uint64_t total = (uint64_t)tree_total + c->host_cnt.n_big_pairs;
    bool region_short = c->host_cnt.wave_over > c->walk_R;
    if (!region_short && total <= cap) break;
    if (region_short) {
      uint32_t r = c->walk_R;
      while (r < c->host_cnt.wave_over) r *= 2;
      c->walk_R = r;
    }
    if (total > cap) cap = total + total / 4;
  }
  uint64_t total = (uint64_t)tree_total + c->host_cnt.n_big_pairs;
  if (c->host_cnt.wave_over > c->walk_R || total > cap) return fail(c, BCH_HIP_ERROR, "pair buffers did not converge");
  c->tree_cap = cap;
  c->P = total;
  if (c->P > 1) {
    int ps[16];
    int pack = 0;
    int np = pair_key_shifts(c, n, ps, &pack);
    bool palt = false;
    rc = dev_radix_sort<uint64_t>(c, c->pair_keys.as<uint64_t>(), c->pair_keys_alt.as<uint64_t>(), nullptr, nullptr,
                                  c->P, ps, np, s, &palt, nullptr, pack);
    if (rc) return rc;
    if (palt) std::swap(c->pair_keys, c->pair_keys_alt);
  }
  return BCH_OK;
}

int ensure_pass_buffers(bch_context* c, uint32_t n) {
  ENSURE(c, c->aabb, (size_t)n * 2 * sizeof(float4));
  ENSURE(c, c->flags, (size_t)n);
  ENSURE(c, c->big_ids, (size_t)n * sizeof(uint32_t));
  ENSURE(c, c->counters, sizeof(Counters));
  ENSURE(c, c->partials, ((size_t)blocks(n) + (size_t)blocks(n, kBlock / 16) + 2) * 8 * sizeof(float));
  ENSURE(c, c->keys, (size_t)n * sizeof(uint64_t));
  ENSURE(c, c->keys_alt, (size_t)n * sizeof(uint64_t));
  ENSURE(c, c->vals, (size_t)n * sizeof(uint32_t));
  ENSURE(c, c->vals_alt, (size_t)n * sizeof(uint32_t));
  uint64_t N = 1;
  while (N < n) N <<= 1;
  ENSURE(c, c->nodes, (size_t)N * 4 * sizeof(float4));
  ENSURE(c, c->wend, (size_t)(n + 1) * sizeof(uint32_t));
  ENSURE(c, c->bsph, (size_t)n * sizeof(float4));
  return BCH_OK;
}

// First part of the narrowphase: buffers, and the pair split of
// c->pair_keys into (i, j), shape classes and the identity permutation.
int narrow_split(bch_context* c, const float4* bsph = nullptr) {
  hipStream_t s = c->stream;
  Counters* cnt = c->counters.as<Counters>();
  uint64_t P = c->P;
  size_t Pm = std::max<uint64_t>(P, 1);
  ENSURE(c, c->pairs, Pm * 2 * sizeof(uint32_t));
  ENSURE(c, c->cls, Pm * sizeof(uint32_t));
  ENSURE(c, c->cls_alt, Pm * sizeof(uint32_t));
  ENSURE(c, c->perm, Pm * sizeof(uint32_t));
  ENSURE(c, c->perm_alt, Pm * sizeof(uint32_t));
  ENSURE(c, c->ccount, (Pm + 1) * sizeof(uint32_t));
  ENSURE(c, c->coffsets, (Pm + 1) * sizeof(uint32_t));
  ENSURE(c, c->cslots, Pm * 8 * sizeof(float4));
  ENSURE(c, c->epa_q, Pm * 4 * sizeof(float4));
  ENSURE(c, c->epa_q2, Pm * 4 * sizeof(float4));
  // capacity mode: P is the capacity, the device holds the count
  const uint32_t* P_dev = c->cap_mode ? &cnt->n_pairs : nullptr;
  hipLaunchKernelGGL(k_reset_epa, dim3(1), dim3(64), 0, s, cnt);
  c->cls_hist_ready = false;
  if (P > 1) {
    // the split also counts the class sort's first (only) digit per tile
    const uint32_t tile = c->sort_tile == 2048 ? 2048u : (uint32_t)kRsTile;
    const uint32_t tiles = (uint32_t)((P + tile - 1) / tile);
    ENSURE(c, c->rs_hist, (size_t)256 * tiles * sizeof(uint32_t));
    if (tile == 2048)
      hipLaunchKernelGGL(k_split_pairs_hist<8>, dim3(tiles), dim3(256), 0, s, (uint32_t)P, P_dev,
                         c->pair_keys.as<uint64_t>(), c->btype.as<uint8_t>(), c->pairs.as<uint32_t>(),
                         c->cls.as<uint32_t>(), bsph, c->ccount.as<uint32_t>(), tiles, c->rs_hist.as<uint32_t>());
    else
      hipLaunchKernelGGL(k_split_pairs_hist<kRsRounds>, dim3(tiles), dim3(256), 0, s, (uint32_t)P, P_dev,
                         c->pair_keys.as<uint64_t>(), c->btype.as<uint8_t>(), c->pairs.as<uint32_t>(),
                         c->cls.as<uint32_t>(), bsph, c->ccount.as<uint32_t>(), tiles, c->rs_hist.as<uint32_t>());
    c->cls_hist_ready = true;
  } else if (P) {
    hipLaunchKernelGGL(k_split_pairs, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, P_dev,
                       c->pair_keys.as<uint64_t>(), c->btype.as<uint8_t>(), c->pairs.as<uint32_t>(),
                       c->cls.as<uint32_t>(), c->perm.as<uint32_t>(), bsph, c->ccount.as<uint32_t>());
  }
  return BCH_OK;
}

// Rest of the narrowphase: class sort, the per-class kernels (with per-pair
// GJK seeds when d_seeds is given; NaN x = none), EPA, dense contacts.
// Fills c->coffsets (P+1), c->contacts (7C), c->C; without `dense` (eager
// pass only) it stops at the per-pair slots (c->ccount, c->cslots).
int narrow_finish(bch_context* c, const float* d_seeds, bool dense = true) {
  hipStream_t s = c->stream;
  Counters* cnt = c->counters.as<Counters>();
  uint64_t P = c->P;
  const uint32_t* P_dev = c->cap_mode ? &cnt->n_pairs : nullptr;
  if (P) {
    // one 8-bit digit (bits 8-15) holds the class and its sub-group; one pass
    // leaves the result in the alt pair
    int cs[1] = {8};
    bool calt = false;
    // the values are the pair indices (the split writes no identity
    // permutation), and the split has counted the digit unless something
    // rewrote the classes since
    int rc = dev_radix_sort<uint32_t>(c, c->cls.as<uint32_t>(), c->cls_alt.as<uint32_t>(), c->perm.as<uint32_t>(),
                                      c->perm_alt.as<uint32_t>(), P, cs, 1, s, &calt, P_dev, 0, true,
                                      c->cls_hist_ready);
    c->cls_hist_ready = false;
    if (rc) return rc;
    if (!calt) {  // P <= 1 (eager pass only): nothing moved
      HIP_TRY(c, hipMemcpyAsync(c->cls_alt.p, c->cls.p, P * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
      HIP_TRY(c, hipMemcpyAsync(c->perm_alt.p, c->perm.p, P * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL(k_class_bounds, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, P_dev,
                       c->cls_alt.as<uint32_t>(), cnt);
    mark(c, 4);
    // The one-lane-per-pair classes run on the side stream beside k_np_hull
    // (disjoint pairs; both append to the EPA queue through one atomic
    // counter): each kernel fills the other's idle issue slots and tail.
    // k_np_hull, the longer branch, is submitted first and the side classes
    // longest first.  Stage 4 -> 5 is then only the fork; 5 -> 6 covers both
    // streams.
    HIP_TRY(c, hipEventRecord(c->fork_ev, s));
    HIP_TRY(c, hipStreamWaitEvent(c->side, c->fork_ev, 0));
    mark(c, 5);
    const uint32_t* perm = c->perm_alt.as<uint32_t>();
    uint32_t g = std::max<uint32_t>(1u, std::min<uint32_t>(blocks(P), c->np_grid));
#define BCH_NP_HULL(G)                                                                                  \
  hipLaunchKernelGGL(k_np_hull<G>, dim3(g), dim3(kBlock), 0, s, cnt, perm, c->pairs.as<uint32_t>(),        \
                     c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_seeds, c->ccount.as<uint32_t>(),     \
                     c->cslots.as<float4>(), c->epa_q.as<float4>(), cnt)
    if (c->hull_lanes == 4)
      BCH_NP_HULL(4);
    else if (c->hull_lanes == 8)
      BCH_NP_HULL(8);
    else
      BCH_NP_HULL(16);
#undef BCH_NP_HULL
    mark(c, 10);
#define BCH_NP(CLS)                                                                                     \
  hipLaunchKernelGGL(k_np<CLS>, dim3(g), dim3(kBlock), 0, c->side, cnt, perm, c->pairs.as<uint32_t>(),   \
                     c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_seeds, c->ccount.as<uint32_t>(),     \
                     c->cslots.as<float4>(), c->epa_q.as<float4>(), cnt)
    BCH_NP(CLS_GJK);
    BCH_NP(CLS_BOX_BOX);
    BCH_NP(CLS_SPHERE_BOX);
    BCH_NP(CLS_SPHERE_SPHERE);
    BCH_NP(CLS_PLANE);
    BCH_NP(CLS_NONE);
#undef BCH_NP
    HIP_TRY(c, hipEventRecord(c->join_ev, c->side));
    HIP_TRY(c, hipStreamWaitEvent(s, c->join_ev, 0));
    mark(c, 6);
    // four pairs per wavefront; an entry outside its budget reruns from the
    // start one pair per wavefront (epa_q2, then epa_q: each queue is
    // consumed by the kernel before the one that refills it), and an entry
    // that outgrows the one-pair store once more with the 2048-face store
    hipLaunchKernelGGL(k_epa_quad, dim3(c->epa_grid), dim3(64), 0, s, &cnt->n_epa, c->epa_q.as<float4>(),
                       c->pairs.as<uint32_t>(), c->bodies.as<BodyRec>(), c->hull.as<float4>(),
                       c->ccount.as<uint32_t>(), c->cslots.as<float4>(), c->cls.as<uint32_t>(),
                       c->epa_q2.as<float4>(), &cnt->n_epa_defer, cnt);
    mark(c, 9);
    hipLaunchKernelGGL(k_epa_wave, dim3(kEpaDeferGrid), dim3(64), 0, s, &cnt->n_epa_defer, c->epa_q2.as<float4>(),
                       c->pairs.as<uint32_t>(), c->bodies.as<BodyRec>(), c->hull.as<float4>(),
                       c->ccount.as<uint32_t>(), c->cslots.as<float4>(), c->cls.as<uint32_t>(), cnt,
                       c->epa_q.as<float4>(), c->epa_wave_faces);
    hipLaunchKernelGGL(k_epa_big, dim3(kEpaBigGrid), dim3(64), 0, s, &cnt->n_epa_over, c->epa_q.as<float4>(),
                       c->pairs.as<uint32_t>(), c->bodies.as<BodyRec>(), c->hull.as<float4>(),
                       c->ccount.as<uint32_t>(), c->cslots.as<float4>(), c->cls.as<uint32_t>(), cnt);
    mark(c, 7);
  } else {
    for (int k = 4; k <= 7; ++k) mark(c, k);
    mark(c, 9);
    mark(c, 10);
  }
  if (!dense && !c->cap_mode) {
    HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    c->C = 0;
    if (c->host_cnt.errors & 2u) return fail(c, BCH_HIP_ERROR, "EPA polytope store overflow");
    return BCH_OK;
  }
  if (c->cap_mode) {
    int rc = dev_scan(c, c->ccount.as<uint32_t>(), c->coffsets.as<uint32_t>(), P, c->scan_part, s, P_dev,
                      &cnt->n_contacts);
    if (rc) return rc;
    c->C = c->C_cap;
    ENSURE(c, c->contacts, c->C_cap * 7 * sizeof(float));
    if (P)
      hipLaunchKernelGGL(k_gather, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, P_dev, c->ccount.as<uint32_t>(),
                         c->coffsets.as<uint32_t>(), c->cslots.as<float4>(), c->C_cap, c->contacts.as<float>(), cnt);
    mark(c, 8);
    return BCH_OK;
  }
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, 1u, c->ccount.as<uint32_t>() + P, 0u);
  int rc = dev_scan(c, c->ccount.as<uint32_t>(), c->coffsets.as<uint32_t>(), P, c->scan_part, s);
  if (rc) return rc;
  // Gather into the contact buffer as earlier passes sized it (at first one
  // contact per pair), read the total with the counters, and gather again
  // only when this pass has more contacts than that buffer holds.
  if (c->contacts.bytes == 0) ENSURE(c, c->contacts, std::max<uint64_t>(P, 1) * 7 * sizeof(float));
  uint64_t ccap = c->contacts.bytes / (7 * sizeof(float));
  if (P)
    hipLaunchKernelGGL(k_gather, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, (const uint32_t*)nullptr,
                       c->ccount.as<uint32_t>(), c->coffsets.as<uint32_t>(), c->cslots.as<float4>(), ccap,
                       c->contacts.as<float>(), cnt);
  mark(c, 8);
  uint32_t total = 0;
  HIP_TRY(c, hipMemcpyAsync(&total, c->coffsets.as<uint32_t>() + P, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  c->C = total;
  if (total > ccap) {
    ENSURE(c, c->contacts, (uint64_t)total * 7 * sizeof(float));
    hipLaunchKernelGGL(k_gather, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, (const uint32_t*)nullptr,
                       c->ccount.as<uint32_t>(), c->coffsets.as<uint32_t>(), c->cslots.as<float4>(), (uint64_t)total,
                       c->contacts.as<float>(), cnt);
    HIP_TRY(c, hipStreamSynchronize(s));
  }
  if (c->host_cnt.errors & 2u) return fail(c, BCH_HIP_ERROR, "EPA polytope store overflow");
  return BCH_OK;
}

// Narrowphase over c->pair_keys (sorted, c->P pairs) with the world bodies.
// Fills c->pairs (2P), c->coffsets (P+1), c->contacts (7C), c->C.
int run_narrowphase(bch_context* c, const float* d_seeds) {
  int rc = narrow_split(c, c->bsph.as<float4>());
  if (rc) return rc;
  return narrow_finish(c, d_seeds);
}

int upload_bodies(bch_context* c, const bch_bodies* b, const uint32_t* world_of_host) {
  hipStream_t s = c->stream;
  uint32_t n = b->n;
  ENSURE(c, c->bodies, (size_t)std::max<uint32_t>(n, 1) * sizeof(BodyRec));
  ENSURE(c, c->btype, (size_t)std::max<uint32_t>(n, 1));
  ENSURE(c, c->staging_a, (size_t)std::max<uint32_t>(n, 1) * 3 * sizeof(float));
  ENSURE(c, c->staging_b, (size_t)std::max<uint32_t>(n, 1) * 4 * sizeof(float));
  ENSURE(c, c->staging_c, (size_t)std::max<uint32_t>(n, 1));
  ENSURE(c, c->staging_d, (size_t)std::max<uint32_t>(n, 1) * 4 * sizeof(float));
  ENSURE(c, c->staging_e, (size_t)std::max<uint32_t>(n, 1) * 2 * sizeof(uint32_t));
  uint32_t nv = b->hull_vertices ? b->n_hull_vertices : 0;
  ENSURE(c, c->hull, (size_t)std::max<uint32_t>(nv, 1) * sizeof(float4));
  ENSURE(c, c->staging_f, (size_t)std::max<uint32_t>(nv, 1) * 3 * sizeof(float));
  if (n) {
    HIP_TRY(c, hipMemcpyAsync(c->staging_a.p, b->position, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->staging_b.p, b->orientation, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->staging_c.p, b->shape_type, (size_t)n, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->staging_d.p, b->shape_params, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, s));
  }
  uint32_t* hf = nullptr;
  uint32_t* hc = nullptr;
  if (nv && n) {
    if (!b->hull_first || !b->hull_count) return fail(c, BCH_INVALID_ARG, "hull vertices without hull_first/count");
    hf = c->staging_e.as<uint32_t>();
    hc = hf + n;
    HIP_TRY(c, hipMemcpyAsync(hf, b->hull_first, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(hc, b->hull_count, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->staging_f.p, b->hull_vertices, (size_t)nv * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_xyz_to_float4, dim3(blocks(nv)), dim3(kBlock), 0, s, nv, c->staging_f.as<float>(),
                       c->hull.as<float4>());
  }
  uint32_t* wo = nullptr;
  if (world_of_host) {
    ENSURE(c, c->world_of, (size_t)std::max<uint32_t>(n, 1) * sizeof(uint32_t));
    HIP_TRY(c, hipMemcpyAsync(c->world_of.p, world_of_host, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    wo = c->world_of.as<uint32_t>();
  }
  {
    std::vector<uint32_t> hid;
    c->n_planes = 0;
    for (uint32_t i = 0; i < n; ++i) {
      if (b->shape_type[i] == BCH_HULL) hid.push_back(i);
      c->n_planes += b->shape_type[i] == BCH_PLANE;
    }
    c->n_hulls = (uint32_t)hid.size();
    ENSURE(c, c->hull_ids, std::max<size_t>(hid.size(), 1) * sizeof(uint32_t));
    if (!hid.empty())
      HIP_TRY(c, hipMemcpyAsync(c->hull_ids.p, hid.data(), hid.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipStreamSynchronize(s));
  }
  if (n)
    hipLaunchKernelGGL(k_pack_bodies, dim3(blocks(n)), dim3(kBlock), 0, s, n, c->staging_a.as<float>(),
                       c->staging_b.as<float>(), c->staging_c.as<uint8_t>(), c->staging_d.as<float>(), hf, hc,
                       c->hull.as<float4>(), wo, c->bodies.as<BodyRec>(), c->btype.as<uint8_t>());
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int validate_bodies(bch_context* c, const bch_bodies* b) {
  if (!b) return fail(c, BCH_INVALID_ARG, "bodies is NULL");
  if (b->n && (!b->position || !b->orientation || !b->shape_type || !b->shape_params))
    return fail(c, BCH_INVALID_ARG, "body arrays must be non-NULL");
  if (b->n > 0x7fffffffu) return fail(c, BCH_INVALID_ARG, "too many bodies");
  for (uint32_t i = 0; i < b->n; ++i) {
    uint8_t t = b->shape_type[i];
    if (t > BCH_PLANE) return fail(c, BCH_INVALID_ARG, "unknown shape type at body " + std::to_string(i));
    if (t == BCH_HULL) {
      if (!b->hull_first || !b->hull_count || !b->hull_vertices)
        return fail(c, BCH_INVALID_ARG, "hull body without hull arrays");
      uint64_t end = (uint64_t)b->hull_first[i] + b->hull_count[i];
      if (end > b->n_hull_vertices) return fail(c, BCH_INVALID_ARG, "hull range out of bounds at body " + std::to_string(i));
    }
  }
  return BCH_OK;
}

// World AABBs (padded by the margin), body classes and scene bounds.
void run_aabbs(bch_context* c, float aabb_margin) {
  uint32_t n = c->n;
  hipStream_t s = c->stream;
  // with body flags, inactive bodies keep the previous pass's box
  const uint8_t* bf = c->has_bflags ? c->bflags.as<uint8_t>() : nullptr;
  int keep = c->aabb_valid ? 1 : 0;
  hipLaunchKernelGGL(k_aabb, dim3(blocks(n)), dim3(kBlock), 0, s, n, c->bodies.as<BodyRec>(), c->hull.as<float4>(),
                     aabb_margin, c->aabb.as<float4>(), c->flags.as<uint8_t>(), c->big_ids.as<uint32_t>(),
                     c->counters.as<Counters>(), c->partials.as<float>(), bf, keep, c->bsph.as<float4>());
  uint32_t hb = c->n_hulls ? blocks(c->n_hulls, kBlock / 16) : 0;
  float* hpart = c->partials.as<float>() + 8 * (size_t)blocks(n);
  if (c->n_hulls)
    hipLaunchKernelGGL(k_hull_aabb, dim3(hb), dim3(kBlock), 0, s, c->n_hulls, c->hull_ids.as<uint32_t>(),
                       c->bodies.as<BodyRec>(), c->hull.as<float4>(), aabb_margin, c->aabb.as<float4>(),
                       c->flags.as<uint8_t>(), c->big_ids.as<uint32_t>(), c->counters.as<Counters>(), hpart, bf, keep);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(1024), 0, s, c->partials.as<float>(), blocks(n), hpart, hb,
                     c->counters.as<Counters>());
  c->aabb_valid = true;
}

// ----------------------------------------------------- mesh BVH (host) --
// The TriangleMesh BVH (shapes/bvh.rs:40-56, 170-269): a node's box covers
// its triangles' vertices; a node of more than four triangles splits its
// slice at the middle after a stable sort of the triangle centres
// ((a + b) + c) / 3 along the longest axis of the box (ties x, y, z).
struct MeshBuilder {
  const float* v;
  const uint32_t* idx;
  std::vector<uint32_t> order;
  std::vector<float4> nodes;  // two per node
  float coord(uint32_t t, int k, int axis) const { return v[3 * idx[3 * t + k] + axis]; }
  float centre(uint32_t t, int axis) const {
    return ((coord(t, 0, axis) + coord(t, 1, axis)) + coord(t, 2, axis)) * (1.0f / 3.0f);
  }
  static int64_t key(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? -(int64_t)(u & 0x7fffffffu) - 1 : (int64_t)u;
  }
  uint32_t build(uint32_t start, uint32_t count) {
    float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
    bool any = false;
    for (uint32_t k = start; k < start + count; ++k)
      for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
          float x = coord(order[k], c, a);
          if (!any) {
            lo[a] = hi[a] = x;
          } else {
            lo[a] = x < lo[a] ? x : lo[a];
            hi[a] = x > hi[a] ? x : hi[a];
          }
          if (a == 2) any = true;
        }
    uint32_t index = (uint32_t)(nodes.size() / 2);
    nodes.push_back(make_float4(lo[0], lo[1], lo[2], u2f(start)));
    nodes.push_back(make_float4(hi[0], hi[1], hi[2], u2f(count | kMeshLeaf)));
    if (count <= 4) return index;
    float e[3] = {(hi[0] - lo[0]) * 0.5f, (hi[1] - lo[1]) * 0.5f, (hi[2] - lo[2]) * 0.5f};
    int axis = 0;
    float m = e[0];
    if (e[1] > m) {
      axis = 1;
      m = e[1];
    }
    if (e[2] > m) axis = 2;
    std::stable_sort(order.begin() + start, order.begin() + start + count,
                     [&](uint32_t x, uint32_t y) { return key(centre(x, axis)) < key(centre(y, axis)); });
    uint32_t mid = count / 2;
    uint32_t l = build(start, mid);
    uint32_t r = build(start + mid, count - mid);
    nodes[2 * index].w = u2f(l);
    nodes[2 * index + 1].w = u2f(r);
    return index;
  }
};

// One world pass (AABBs, broadphase, narrowphase) on c->stream.  Eager mode
// reads counts back as it goes; capacity mode (c->cap_mode) keeps them on the
// device.
int run_pass(bch_context* c, float aabb_margin) {
  uint32_t n = c->n;
  hipStream_t s = c->stream;
  if (!c->cap_mode) {
    c->P = 0;
    c->C = 0;
  }
  int rc = ensure_pass_buffers(c, std::max<uint32_t>(n, 1));
  if (rc) return rc;
  mark(c, 0);
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, c->counters.as<Counters>());
  if (n) {
    run_aabbs(c, aabb_margin);
    mark(c, 1);
    rc = run_broadphase(c, n, c->batched ? c->world_of.as<uint32_t>() : nullptr,
                        c->batched ? c->world_offsets.as<uint32_t>() : nullptr, false);
    if (rc) return rc;
  } else {
    for (int k = 1; k <= 3; ++k) mark(c, k);
    if (c->cap_mode) c->P = c->P_cap;
  }
  rc = run_narrowphase(c, nullptr);
  if (rc) return rc;
  HIP_TRY(c, hipGetLastError());
  return BCH_OK;
}

void release_graph(bch_context* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

void release_depen_graph(bch_context* c) {
  if (c->dgraph_exec) (void)hipGraphExecDestroy(c->dgraph_exec);
  if (c->dgraph) (void)hipGraphDestroy(c->dgraph);
  c->dgraph_exec = nullptr;
  c->dgraph = nullptr;
}

// Capture whatever `body` enqueues on the context's own stream into a graph.
template <class F>
int capture_graph(bch_context* c, F body, hipGraph_t* out_g, hipGraphExec_t* out_x) {
  *out_g = nullptr;
  *out_x = nullptr;
  hipStream_t caller = c->stream;
  c->stream = c->own_stream;
  c->capturing = true;
  int rc = BCH_OK;
  hipError_t e = hipStreamBeginCapture(c->own_stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) rc = body();
  hipGraph_t g = nullptr;
  hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(c->own_stream, &g) : e;
  c->capturing = false;
  c->stream = caller;
  if (rc || e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    return fail(c, BCH_HIP_ERROR, std::string("stream capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(c, BCH_HIP_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  *out_g = g;
  *out_x = x;
  return BCH_OK;
}

// The forward signed distance's EPA over the walk's queue (c->epa_q, count
// Counters::n_epa): four pairs per wavefront, then the entries it hands on
// one pair per wavefront (c->sd_defer, count Counters::n_epa_defer).  The
// queues hold one entry per pair at most, sized by the caller.
void sd_forward_epa(bch_context* c, const uint32_t* d_pairs, float* d_out, uint8_t* d_status, hipStream_t s) {
  Counters* cnt = c->counters.as<Counters>();
  hipLaunchKernelGGL(k_sd_epa_quad<false>, dim3(c->epa_grid), dim3(64), 0, s, &cnt->n_epa, c->epa_q.as<float4>(),
                     d_pairs, c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_out, d_status, (float4*)nullptr,
                     c->sd_defer.as<float4>(), &cnt->n_epa_defer);
  hipLaunchKernelGGL(k_sd_epa<false>, dim3(kEpaDeferGrid), dim3(64), 0, s, &cnt->n_epa_defer,
                     c->sd_defer.as<float4>(), d_pairs, c->bodies.as<BodyRec>(), c->hull.as<float4>(),
                     (const float*)nullptr, d_out, d_status);
}

// The adjoint's EPA arm over the backward walk's queue: EPA again, four
// pairs per wavefront, recording each entry's face (c->sd_rec); the VJP of
// those entries one wavefront each; the entries handed on, whole, one pair
// per wavefront.  Needs c->sd_rec and c->sd_defer for the queue's length.
void sd_backward_epa(bch_context* c, const uint32_t* d_pairs, const float* d_gin, float* d_grad, uint8_t* d_status,
                     hipStream_t s) {
  Counters* cnt = c->counters.as<Counters>();
  hipLaunchKernelGGL(k_sd_epa_quad<true>, dim3(c->epa_grid), dim3(64), 0, s, &cnt->n_epa, c->epa_q.as<float4>(),
                     d_pairs, c->bodies.as<BodyRec>(), c->hull.as<float4>(), (float*)nullptr, (uint8_t*)nullptr,
                     c->sd_rec.as<float4>(), c->sd_defer.as<float4>(), &cnt->n_epa_defer);
  hipLaunchKernelGGL(k_sd_epa_vjp, dim3(c->epa_grid), dim3(64), 0, s, &cnt->n_epa, c->epa_q.as<float4>(),
                     c->sd_rec.as<float4>(), d_pairs, c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_gin, d_grad,
                     d_status);
  hipLaunchKernelGGL(k_sd_epa<true>, dim3(kEpaDeferGrid), dim3(64), 0, s, &cnt->n_epa_defer,
                     c->sd_defer.as<float4>(), d_pairs, c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_gin, d_grad,
                     d_status);
}

// One de-penetration step in capacity mode (c->P_cap pairs): poses packed
// from the caller's arrays, AABBs, broadphase, signed distance of every pair,
// loss record, then (update) the adjoint of the penetrating pairs, the
// per-body gradient and the pose update written back to the caller's arrays.
int run_depen_step(bch_context* c, float aabb_margin, float lr, float* d_pos, float* d_quat, bool update) {
  uint32_t n = c->n;
  hipStream_t s = c->stream;
  Counters* cnt = c->counters.as<Counters>();
  uint64_t cap = c->P_cap;
  int rc = ensure_pass_buffers(c, n);
  if (rc) return rc;
  ENSURE(c, c->pairs, cap * 2 * sizeof(uint32_t));
  ENSURE(c, c->cls, cap * sizeof(uint32_t));
  ENSURE(c, c->perm, cap * sizeof(uint32_t));
  ENSURE(c, c->epa_q, cap * kSdEpaEntry * sizeof(float4));
  ENSURE(c, c->sd_defer, cap * kSdEpaEntry * sizeof(float4));
  ENSURE(c, c->sd_rec, cap * 4 * sizeof(float4));
  ENSURE(c, c->sd_brec, cap * kSdBandRec * sizeof(float4));
  ENSURE(c, c->sd_out, cap * 10 * sizeof(float));
  ENSURE(c, c->sd_status, cap);
  ENSURE(c, c->pen_pairs, cap * 2 * sizeof(uint32_t));
  ENSURE(c, c->pen_g, cap * 10 * sizeof(float));
  ENSURE(c, c->g14, cap * 14 * sizeof(float));
  ENSURE(c, c->body_grad, (size_t)n * 7 * sizeof(float));
  uint32_t nb = blocks(cap, kDepenBlock);
  ENSURE(c, c->pen_part, (size_t)nb * sizeof(float2));
  hipLaunchKernelGGL(k_pack_poses, dim3(blocks(n)), dim3(kBlock), 0, s, n, (const float*)d_pos,
                     (const float*)d_quat, c->bodies.as<BodyRec>());
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, cnt);
  run_aabbs(c, aabb_margin);
  rc = run_broadphase(c, n, c->batched ? c->world_of.as<uint32_t>() : nullptr,
                      c->batched ? c->world_offsets.as<uint32_t>() : nullptr, false);
  if (rc) return rc;
  const uint32_t* P_dev = &cnt->n_pairs;
  hipLaunchKernelGGL(k_split_pairs, dim3(blocks(cap)), dim3(kBlock), 0, s, (uint32_t)cap, P_dev,
                     c->pair_keys.as<uint64_t>(), c->btype.as<uint8_t>(), c->pairs.as<uint32_t>(),
                     c->cls.as<uint32_t>(), c->perm.as<uint32_t>());
  constexpr int G = 8;
  uint32_t g = std::max<uint32_t>(1u, std::min<uint32_t>(blocks(cap, kBlock / G), 8192u));
  hipLaunchKernelGGL(k_reset_epa, dim3(1), dim3(64), 0, s, cnt);
  hipLaunchKernelGGL((k_sd_walk<G, false>), dim3(g), dim3(kBlock), 0, s, (uint32_t)cap, P_dev, n,
                     c->pairs.as<uint32_t>(), c->bodies.as<BodyRec>(), c->hull.as<float4>(), (const float*)nullptr,
                     c->sd_out.as<float>(), c->sd_status.as<uint8_t>(), c->epa_q.as<float4>(), cnt);
  sd_forward_epa(c, c->pairs.as<uint32_t>(), c->sd_out.as<float>(), c->sd_status.as<uint8_t>(), s);
  hipLaunchKernelGGL(k_depen_terms, dim3(nb), dim3(kDepenBlock), 0, s, (uint32_t)cap, P_dev, c->sd_out.as<float>(),
                     c->pairs.as<uint32_t>(), c->pen_pairs.as<uint32_t>(), c->pen_g.as<float>(), &cnt->n_pen,
                     c->pen_part.as<float2>());
  hipLaunchKernelGGL(k_depen_record, dim3(1), dim3(kDepenBlock), 0, s, c->pen_part.as<float2>(), nb, cnt,
                     c->dhist.as<DepenRecord>(), c->dhist_cap, c->dstep.as<uint32_t>());
  if (!update) return BCH_OK;
  const uint32_t* pen_dev = &cnt->n_pen;
  hipLaunchKernelGGL(k_reset_epa, dim3(1), dim3(64), 0, s, cnt);
  hipLaunchKernelGGL((k_sd_walk<G, true>), dim3(g), dim3(kBlock), 0, s, (uint32_t)cap, pen_dev, n,
                     c->pen_pairs.as<uint32_t>(), c->bodies.as<BodyRec>(), c->hull.as<float4>(),
                     c->pen_g.as<float>(), c->g14.as<float>(), (uint8_t*)nullptr, c->epa_q.as<float4>(), cnt,
                     c->sd_brec.as<float4>());
  hipLaunchKernelGGL(k_sd_band_vjp, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(blocks(cap), 8192u))),
                     dim3(kBlock), 0, s, (uint32_t)cap, pen_dev, c->sd_brec.as<float4>(), c->pen_pairs.as<uint32_t>(),
                     c->bodies.as<BodyRec>(), c->hull.as<float4>(), c->pen_g.as<float>(), c->g14.as<float>());
  sd_backward_epa(c, c->pen_pairs.as<uint32_t>(), c->pen_g.as<float>(), c->g14.as<float>(), (uint8_t*)nullptr, s);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks((uint64_t)n * 7)), dim3(kBlock), 0, s, n * 7,
                     c->body_grad.as<uint32_t>(), 0u);
  float* gpos = c->body_grad.as<float>();
  float* gquat = gpos + (size_t)n * 3;
  hipLaunchKernelGGL(k_depen_scatter, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(blocks(cap), 4096u))),
                     dim3(kBlock), 0, s, pen_dev, c->pen_pairs.as<uint32_t>(), c->g14.as<float>(), gpos, gquat);
  hipLaunchKernelGGL(k_depen_update, dim3(blocks(n)), dim3(kBlock), 0, s, n, lr, gpos, gquat, d_pos, d_quat);
  return BCH_OK;
}

// Counts of the last capacity-mode pass (waits for the stream): held (what
// the buffers hold, also left in c->P and c->C for the result getters) and
// found.
// The uncaptured sizing pass of a capture: run `pass` (capacity mode) and
// read its counters; while a tree-walk wavefront overflowed its region,
// give the regions twice that wavefront's count and run it again.  A region
// overflow reported by an earlier replay doubles the regions first.
template <class F>
int size_walk(bch_context* c, F pass) {
  if (c->walk_grow) {
    c->walk_R *= 2;
    c->walk_grow = false;
  }
  for (int attempt = 0; attempt < 4; ++attempt) {
    int rc = pass();
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->host_cnt.wave_over <= c->walk_R) return BCH_OK;
    uint32_t r = c->walk_R;
    while (r < 2 * c->host_cnt.wave_over) r *= 2;
    c->walk_R = r;
  }
  return fail(c, BCH_HIP_ERROR, "tree-walk regions did not converge");
}

int read_capacity_counts(bch_context* c, uint64_t* n_pairs, uint64_t* n_contacts, uint64_t* found_pairs,
                         uint64_t* found_contacts) {
  HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  const Counters& h = c->host_cnt;
  c->P = std::min<uint64_t>(h.n_pairs, c->graph_P_cap);
  c->C = std::min<uint64_t>(h.n_contacts, c->graph_C_cap);
  if (n_pairs) *n_pairs = c->P;
  if (n_contacts) *n_contacts = c->C;
  if (found_pairs) *found_pairs = h.n_pairs_total;
  if (found_contacts) *found_contacts = h.n_contacts;
  if (h.errors & 1u) return fail(c, BCH_HIP_ERROR, "box-box clip polygon overflow");
  if (h.errors & 2u) return fail(c, BCH_HIP_ERROR, "EPA polytope store overflow");
  if (h.errors & 128u) {
    c->walk_grow = true;
    return fail(c, BCH_OVERFLOW, "a tree-walk wavefront's pair region overflowed: capture again");
  }
  if (h.errors & 8u) return fail(c, BCH_OVERFLOW, "pair capacity exceeded");
  if (h.errors & 16u) return fail(c, BCH_OVERFLOW, "contact capacity exceeded");
  return BCH_OK;
}

}  // namespace

extern "C" {

int bch_create(int device, bch_context** out) {
  if (!out) return BCH_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return BCH_NO_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return BCH_NO_DEVICE;
  bch_context* c = new bch_context();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return BCH_HIP_ERROR;
  }
  c->stream = c->own_stream;
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess) {
    if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
    if (c->side) (void)hipStreamDestroy(c->side);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return BCH_HIP_ERROR;
  }
  for (auto& e : c->ev) (void)hipEventCreate(&e);
  // the static scratch of this device (bch_collide_batch_static): its
  // address, and the views its context works on
  void* sp = nullptr;
  if (scratch_init(device, &sp) != BCH_OK) {
    bch_destroy(c);
    return BCH_HIP_ERROR;
  }
  bch_context* st = new bch_context();
  st->device = device;
  st->no_alloc = true;
  st->sort_tile = (uint32_t)kRsTile;  // the carved layout holds kStTiles tiles
  char* b = static_cast<char*>(sp);
  st->bodies.view(b + kStBodies, kStN * sizeof(BodyRec));
  st->btype.view(b + kStBtype, kStN);
  st->hull.view(b + kStHull, kStV * sizeof(float4));
  st->pair_keys.view(b + kStKeys, kStP * sizeof(uint64_t));
  st->pairs.view(b + kStPairs, kStP * 2 * sizeof(uint32_t));
  DevBuf* u32s[4] = {&st->cls, &st->cls_alt, &st->perm, &st->perm_alt};
  for (int k = 0; k < 4; ++k) u32s[k]->view(b + kStU32 + k * al256(kStP * sizeof(uint32_t)), kStP * sizeof(uint32_t));
  st->ccount.view(b + kStCount, (kStP + 1) * sizeof(uint32_t));
  st->coffsets.view(b + kStCount + al256((kStP + 1) * sizeof(uint32_t)), (kStP + 1) * sizeof(uint32_t));
  st->cslots.view(b + kStSlots, kStP * 8 * sizeof(float4));
  st->epa_q.view(b + kStEpa, kStP * 4 * sizeof(float4));
  st->epa_q2.view(b + kStEpa + al256(kStP * 4 * sizeof(float4)), kStP * 4 * sizeof(float4));
  st->counters.view(b + kStCounters, sizeof(Counters));
  st->rs_hist.view(b + kStHist, 256 * kStTiles * sizeof(uint32_t));
  st->rs_offs.view(b + kStOffs, (256 * kStTiles + 256) * sizeof(uint32_t));
  st->scan_part.view(b + kStPart, kStParts * sizeof(uint32_t));
  c->stat = st;
  *out = c;
  return BCH_OK;
}

void bch_destroy(bch_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  release_graph(c);
  release_depen_graph(c);
  DevBuf* all[] = {&c->bodies, &c->hull, &c->world_offsets, &c->world_of, &c->staging_a, &c->staging_b,
                   &c->staging_c, &c->staging_d, &c->staging_e, &c->staging_f, &c->aabb, &c->flags, &c->big_ids,
                   &c->counters, &c->partials, &c->keys, &c->keys_alt, &c->vals, &c->vals_alt, &c->nodes, &c->hull_ids,
                   &c->btype, &c->wend, &c->world_reg, &c->walk_regions,
                   &c->wave_counts, &c->wave_offs, &c->pair_keys, &c->pair_keys_alt, &c->pairs, &c->cls, &c->cls_alt, &c->perm, &c->perm_alt,
                   &c->ccount, &c->coffsets, &c->cslots, &c->contacts, &c->epa_q, &c->epa_q2,
                   &c->scan_part, &c->rs_hist, &c->rs_offs, &c->sd_out, &c->sd_status, &c->pen_pairs,
                   &c->pen_g, &c->g14, &c->body_grad, &c->pen_part, &c->dhist, &c->dstep, &c->mkeys[0],
                   &c->mkeys[1], &c->mvalid[0], &c->mvalid[1], &c->ment[0], &c->ment[1], &c->bflags, &c->bgm,
                   &c->mroute, &c->msrc, &c->mcnt, &c->mslots, &c->moff, &c->mpts, &c->mseeds, &c->mimp,
                   &c->vel, &c->d_mesh_recs, &c->d_mesh_nodes, &c->d_mesh_tris, &c->d_mesh_tri_aabb,
                   &c->d_mesh_order, &c->mx_pairs, &c->mx_poses, &c->mx_seeds, &c->mx_first, &c->mx_count,
                   &c->mx_offs, &c->mx_sub, &c->mx_sub_cnt, &c->mx_sub_slots, &c->mx_epa_q, &c->mx_epa_q2,
                   &c->mx_out_count, &c->mx_out_pts, &c->sw_target, &c->sw_caster, &c->sw_from, &c->sw_to,
                   &c->sw_radius, &c->sw_pose, &c->sw_hit, &c->sw_out, &c->sw_keys, &c->sw_order, &c->arena,
                   &c->arena_upd, &c->bp_pairs, &c->batch_pairs, &c->batch_seeds, &c->batch_out_count,
                   &c->batch_out_pts};
  for (DevBuf* b : all) b->release();
  c->proxy_box.release();
  c->arena_stage.release();
  if (c->aux) bch_destroy(c->aux);
  delete c->stat;  // views only: nothing of the static array is freed
  for (auto& e : c->ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(c->fork_ev);
  (void)hipEventDestroy(c->join_ev);
  (void)hipStreamDestroy(c->side);
  (void)hipStreamDestroy(c->own_stream);
  delete c;
}

const char* bch_last_error(const bch_context* c) { return c ? c->err.c_str() : "no context"; }

int bch_synchronize(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return BCH_OK;
}

int bch_last_timings(bch_context* c, float* out_ms) {
  if (!c || !out_ms) return BCH_INVALID_ARG;
  for (int i = 0; i < 10; ++i) out_ms[i] = c->timings[i];
  return BCH_OK;
}

int bch_last_epa_routes(bch_context* c, uint32_t* out3) {
  if (!c || !out3) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  if (!c->counters.p) {
    out3[0] = out3[1] = out3[2] = 0;
    return BCH_OK;
  }
  Counters h;
  HIP_TRY(c, hipMemcpyAsync(&h, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  out3[0] = h.n_epa;
  out3[1] = h.n_epa_defer;
  out3[2] = h.n_epa_over;
  return BCH_OK;
}

int bch_set_stream(bch_context* c, void* stream) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->stream = (hipStream_t)stream;  // NULL is the device's null stream
  return BCH_OK;
}

int bch_reset_stream(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->stream = c->own_stream;
  return BCH_OK;
}

int bch_set_tuning(bch_context* c, int knob, int value) {
  if (!c) return BCH_INVALID_ARG;
  switch (knob) {
    case BCH_TUNE_HULL_LANES:
      if (value != 4 && value != 8 && value != 16) return fail(c, BCH_INVALID_ARG, "hull lanes must be 4, 8 or 16");
      c->hull_lanes = value;
      return BCH_OK;
    case BCH_TUNE_TRAVERSE_BLOCK:
      if (value != 64 && value != 128 && value != 256 && value != 512)
        return fail(c, BCH_INVALID_ARG, "tree-walk block must be 64, 128, 256 or 512 threads");
      c->trav_block = (uint32_t)value;
      return BCH_OK;
    case BCH_TUNE_EPA_GRID:
      if (value < 1 || value > (1 << 20)) return fail(c, BCH_INVALID_ARG, "EPA grid must be in [1, 2^20]");
      c->epa_grid = (uint32_t)value;
      return BCH_OK;
    case BCH_TUNE_NP_GRID:
      if (value < 1 || value > (1 << 20)) return fail(c, BCH_INVALID_ARG, "narrowphase grid must be in [1, 2^20]");
      c->np_grid = (uint32_t)value;
      return BCH_OK;
    case BCH_TUNE_EPA_WAVE_FACES:
      if (value < 16 || value > 256) return fail(c, BCH_INVALID_ARG, "one-pair EPA face store must be in [16, 256]");
      c->epa_wave_faces = value;
      return BCH_OK;
    case BCH_TUNE_SORT_TILE:
      if (value != 2048 && value != 4096) return fail(c, BCH_INVALID_ARG, "sort tile must be 2048 or 4096 items");
      c->sort_tile = (uint32_t)value;
      return BCH_OK;
    default:
      return fail(c, BCH_INVALID_ARG, "unknown tuning knob");
  }
}

// Autotuner: time the world pass (device time from its own events, median
// of `trials`) over the launch-shape knobs and keep the fastest.  Results
// do not depend on the knobs; only the time does.
static int time_pass(bch_context* c, float margin, int trials, float* ms) {
  std::vector<float> t;
  for (int k = 0; k < trials; ++k) {
    uint64_t P = 0, Cn = 0;
    int rc = bch_world_collide(c, margin, &P, &Cn);
    if (rc) return rc;
    float s = 0.0f;
    for (int i = 0; i < 8; ++i) s += c->timings[i];
    t.push_back(s);
  }
  std::sort(t.begin(), t.end());
  *ms = t[t.size() / 2];
  return BCH_OK;
}

int bch_autotune(bch_context* c, float aabb_margin, int trials, float* out_ms) {
  if (!c || trials < 1) return BCH_INVALID_ARG;
  if (!c->n) return fail(c, BCH_INVALID_ARG, "no bodies: call bch_world_set_bodies first");
  float best = 0.0f;
  int rc = time_pass(c, aabb_margin, trials, &best);  // also sizes the pair buffer
  if (rc) return rc;
  // one knob at a time, in the order of their weight in the pass
  const int hull_opts[] = {8, 16};
  const uint32_t epa_opts[] = {16384, 32768, 65536, 131072};
  const uint32_t np_opts[] = {2048, 4096, 8192};
  {
    int keep = c->hull_lanes;
    for (int v : hull_opts) {
      c->hull_lanes = v;
      float ms;
      if ((rc = time_pass(c, aabb_margin, trials, &ms))) return rc;
      if (ms < best) { best = ms; keep = v; }
    }
    c->hull_lanes = keep;
  }
  {
    uint32_t keep = c->epa_grid;
    for (uint32_t v : epa_opts) {
      c->epa_grid = v;
      float ms;
      if ((rc = time_pass(c, aabb_margin, trials, &ms))) return rc;
      if (ms < best) { best = ms; keep = v; }
    }
    c->epa_grid = keep;
  }
  {
    uint32_t keep = c->np_grid;
    for (uint32_t v : np_opts) {
      c->np_grid = v;
      float ms;
      if ((rc = time_pass(c, aabb_margin, trials, &ms))) return rc;
      if (ms < best) { best = ms; keep = v; }
    }
    c->np_grid = keep;
  }
  {
    // threads per tree-walk block (the walk's tile: leaves per block, and so
    // the run of leaves each XCD's blocks cover)
    uint32_t keep = c->trav_block;
    for (uint32_t v : {128u, 256u, 512u}) {
      c->trav_block = v;
      float ms;
      if ((rc = time_pass(c, aabb_margin, trials, &ms))) return rc;
      if (ms < best) { best = ms; keep = v; }
    }
    c->trav_block = keep;
  }
  {
    // items per radix-sort tile (the histogram and scatter blocks of every
    // sort in the pass: Morton codes, pair keys, classes)
    uint32_t keep = c->sort_tile;
    for (uint32_t v : {2048u, 4096u}) {
      c->sort_tile = v;
      float ms;
      if ((rc = time_pass(c, aabb_margin, trials, &ms))) return rc;
      if (ms < best) { best = ms; keep = v; }
    }
    c->sort_tile = keep;
  }
  if (out_ms) *out_ms = best;
  return BCH_OK;
}

int bch_get_tuning(bch_context* c, int knob, int* value) {
  if (!c || !value) return BCH_INVALID_ARG;
  switch (knob) {
    case BCH_TUNE_HULL_LANES: *value = c->hull_lanes; return BCH_OK;
    case BCH_TUNE_TRAVERSE_BLOCK: *value = (int)c->trav_block; return BCH_OK;
    case BCH_TUNE_EPA_GRID: *value = (int)c->epa_grid; return BCH_OK;
    case BCH_TUNE_NP_GRID: *value = (int)c->np_grid; return BCH_OK;
    case BCH_TUNE_EPA_WAVE_FACES: *value = c->epa_wave_faces; return BCH_OK;
    case BCH_TUNE_SORT_TILE: *value = (int)c->sort_tile; return BCH_OK;
    default: return fail(c, BCH_INVALID_ARG, "unknown tuning knob");
  }
}

// ------------------------------------------------------------ proxies --
// The proxy broadphase and the batch entry points run on the batch context
// (bch_context::aux): created on first use, bound to the caller's stream and
// launch knobs on every call.
static int batch_ctx(bch_context* c, bch_context** out) {
  if (!c) return BCH_INVALID_ARG;
  if (!c->aux) {
    bch_context* a = nullptr;
    int rc = bch_create(c->device, &a);
    if (rc) return fail(c, rc, "cannot create the batch context");
    a->is_aux = true;
    c->aux = a;
  }
  bch_context* a = c->aux;
  a->stream = c->stream;
  a->hull_lanes = c->hull_lanes;
  a->epa_grid = c->epa_grid;
  a->np_grid = c->np_grid;
  a->epa_wave_faces = c->epa_wave_faces;
  a->sort_tile = c->sort_tile;
  *out = a;
  return BCH_OK;
}

static int from_batch(bch_context* c, const bch_context* a, int rc) {
  if (rc) c->err = a->err;
  return rc;
}

#define BCH_ON_BATCH(c, a)              \
  bch_context* a = nullptr;             \
  do {                                  \
    int rc_ = batch_ctx(c, &a);         \
    if (rc_) return rc_;                \
  } while (0)

static void proxy_put(bch_context* a, uint32_t slot, const float* mn, const float* mx) {
  float4* b = a->proxy_box.as<float4>() + 2 * (size_t)slot;
  b[0] = mn ? make_float4(mn[0], mn[1], mn[2], 0.0f) : make_float4(NAN, NAN, NAN, 0.0f);
  b[1] = mx ? make_float4(mx[0], mx[1], mx[2], 0.0f) : make_float4(NAN, NAN, NAN, 0.0f);
  if (!a->arena_whole && slot < a->arena_n && !a->proxy_dirty[slot]) {
    a->proxy_dirty[slot] = 1;
    a->dirty_list.push_back(slot);
    if (a->dirty_list.size() > (size_t)a->n_proxy / 4) a->arena_whole = true;
  }
  a->bp_pairs_valid = false;
}

// broadphase/mod.rs:94-130.  The lowest free slot is reused
// (brute_force.rs:26-34, axis_sweep3.rs:398-421, dbvt.rs:679-697).
int bch_bp_create_proxy(bch_context* c, const float* mn, const float* mx, uint32_t* out) {
  if (!c || !mn || !mx || !out) return fail(c, BCH_INVALID_ARG, "NULL argument");
  BCH_ON_BATCH(c, a);
  uint32_t slot;
  if (!a->proxy_free.empty()) {
    slot = *a->proxy_free.begin();
    a->proxy_free.erase(a->proxy_free.begin());
  } else {
    if (a->n_proxy == 0x7fffffffu) return fail(c, BCH_INVALID_ARG, "too many proxies");
    if (!a->proxy_box.ensure((size_t)(a->n_proxy + 1) * 2 * sizeof(float4)))
      return fail(c, BCH_HIP_ERROR, "pinned host allocation failed");
    slot = a->n_proxy++;
    a->proxy_live.push_back(0);
    a->proxy_dirty.push_back(0);
  }
  a->proxy_live[slot] = 1;
  proxy_put(a, slot, mn, mx);
  *out = slot;
  return BCH_OK;
}

// A dead slot is a no-op (axis_sweep3.rs:424-426, dbvt.rs:700-702).
int bch_bp_set_aabb(bch_context* c, uint32_t proxy, const float* mn, const float* mx) {
  if (!c || !mn || !mx) return fail(c, BCH_INVALID_ARG, "NULL argument");
  BCH_ON_BATCH(c, a);
  if (proxy >= a->n_proxy || !a->proxy_live[proxy]) return BCH_OK;
  proxy_put(a, proxy, mn, mx);
  return BCH_OK;
}

int bch_bp_set_aabbs(bch_context* c, uint32_t count, const uint32_t* proxies, const float* mins, const float* maxs) {
  if (!c || (count && (!proxies || !mins || !maxs))) return fail(c, BCH_INVALID_ARG, "NULL argument");
  BCH_ON_BATCH(c, a);
  if (count > a->n_proxy / 4) a->arena_whole = true;  // the next sync uploads everything: skip the bookkeeping
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t p = proxies[k];
    if (p < a->n_proxy && a->proxy_live[p]) proxy_put(a, p, mins + 3 * (size_t)k, maxs + 3 * (size_t)k);
  }
  return BCH_OK;
}

int bch_bp_destroy_proxy(bch_context* c, uint32_t proxy) {
  if (!c) return BCH_INVALID_ARG;
  BCH_ON_BATCH(c, a);
  if (proxy < a->n_proxy && a->proxy_live[proxy]) {
    a->proxy_live[proxy] = 0;
    proxy_put(a, proxy, nullptr, nullptr);  // NaN boxes overlap nothing
    a->proxy_free.insert(proxy);
  }
  return BCH_OK;
}

int bch_bp_proxy_aabb(bch_context* c, uint32_t proxy, float* mn, float* mx, int* live) {
  if (!c || !live) return fail(c, BCH_INVALID_ARG, "NULL argument");
  BCH_ON_BATCH(c, a);
  *live = proxy < a->n_proxy && a->proxy_live[proxy];
  if (*live) {
    const float4* b = a->proxy_box.as<float4>() + 2 * (size_t)proxy;
    if (mn) { mn[0] = b[0].x; mn[1] = b[0].y; mn[2] = b[0].z; }
    if (mx) { mx[0] = b[1].x; mx[1] = b[1].y; mx[2] = b[1].z; }
  }
  return BCH_OK;
}

// Bring the device arena up to the host mirror: a whole upload when it grew
// past its buffer or most slots changed, else the new tail slots plus a
// scatter of the changed ones.
static int arena_sync(bch_context* a) {
  hipStream_t s = a->stream;
  uint32_t n = a->n_proxy;
  const float4* mirror = a->proxy_box.as<float4>();
  bool whole = a->arena.bytes < (size_t)n * 2 * sizeof(float4) || a->arena_n == 0 || a->arena_whole ||
               a->dirty_list.size() > (size_t)n / 4;
  if (whole) {
    if (a->arena.bytes < (size_t)n * 2 * sizeof(float4)) ENSURE(a, a->arena, (size_t)n * 2 * sizeof(float4));
    HIP_TRY(a, hipMemcpyAsync(a->arena.p, mirror, (size_t)n * 2 * sizeof(float4), hipMemcpyHostToDevice, s));
  } else {
    if (n > a->arena_n)
      HIP_TRY(a, hipMemcpyAsync(a->arena.as<float4>() + 2 * (size_t)a->arena_n, mirror + 2 * (size_t)a->arena_n,
                                (size_t)(n - a->arena_n) * 2 * sizeof(float4), hipMemcpyHostToDevice, s));
    uint32_t m = (uint32_t)a->dirty_list.size();
    if (m) {
      size_t ib = ((size_t)m * sizeof(uint32_t) + 15) & ~(size_t)15;
      size_t need = ib + (size_t)m * 2 * sizeof(float4);
      if (!a->arena_stage.ensure(need)) return fail(a, BCH_HIP_ERROR, "pinned host allocation failed");
      ENSURE(a, a->arena_upd, need);
      uint32_t* idx = a->arena_stage.as<uint32_t>();
      float4* box = reinterpret_cast<float4*>(a->arena_stage.as<char>() + ib);
      for (uint32_t k = 0; k < m; ++k) {
        uint32_t i = a->dirty_list[k];
        idx[k] = i;
        box[2 * k] = mirror[2 * (size_t)i];
        box[2 * k + 1] = mirror[2 * (size_t)i + 1];
      }
      HIP_TRY(a, hipMemcpyAsync(a->arena_upd.p, a->arena_stage.p, need, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_arena_scatter, dim3(blocks(m)), dim3(kBlock), 0, s, m, a->arena_upd.as<uint32_t>(),
                         reinterpret_cast<const float4*>(a->arena_upd.as<char>() + ib), a->arena.as<float4>());
    }
  }
  for (uint32_t i : a->dirty_list) a->proxy_dirty[i] = 0;
  a->dirty_list.clear();
  a->arena_whole = false;
  a->arena_n = n;
  return BCH_OK;
}

// compute_pairs_into (broadphase/mod.rs:103-118): every live closed overlap
// (i < j), lexicographically sorted.  The result is kept until a proxy
// changes, so the OVERFLOW re-call with a larger capacity only copies it.
int bch_bp_compute_pairs(bch_context* c, uint32_t* out_pairs, uint64_t capacity, uint64_t* out_count) {
  if (!c || !out_count) return fail(c, BCH_INVALID_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  BCH_ON_BATCH(c, a);
  *out_count = 0;
  uint32_t n = a->n_proxy;
  hipStream_t s = a->stream;
  if (!a->bp_pairs_valid) {
    a->bp_count = 0;
    if (n) {
      int rc = ensure_pass_buffers(a, n);
      if (!rc) rc = arena_sync(a);
      if (rc) return from_batch(c, a, rc);
      hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, a->counters.as<Counters>());
      HIP_TRY(c, hipMemcpyAsync(a->aabb.p, a->arena.p, (size_t)n * 2 * sizeof(float4), hipMemcpyDeviceToDevice, s));
      a->batched = false;
      a->cap_mode = false;
      rc = run_broadphase(a, n, nullptr, nullptr, true);
      if (rc) return from_batch(c, a, rc);
      a->bp_count = a->P;
      if (a->P) {
        ENSURE(a, a->bp_pairs, a->P * sizeof(uint2));
        hipLaunchKernelGGL(k_keys_to_pairs, dim3(blocks(a->P)), dim3(kBlock), 0, s, (uint32_t)a->P,
                           a->pair_keys.as<uint64_t>(), a->bp_pairs.as<uint2>());
      }
    }
    a->bp_pairs_valid = true;
  }
  *out_count = a->bp_count;
  if (a->bp_count > capacity || (!out_pairs && a->bp_count)) return BCH_OVERFLOW;
  if (a->bp_count)
    HIP_TRY(c, hipMemcpyAsync(out_pairs, a->bp_pairs.p, a->bp_count * sizeof(uint2), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int bch_bp_aabb_candidates(bch_context* c, const float* mn, const float* mx, uint32_t* out, uint64_t capacity,
                           uint64_t* out_count) {
  if (!c || !mn || !mx || !out_count) return fail(c, BCH_INVALID_ARG, "NULL argument");
  BCH_ON_BATCH(c, a);
  // Host scan over the proxy arena: the exact query of
  // broadphase/mod.rs:135-146 (one box against every live proxy).
  uint64_t k = 0;
  Box6 q{V3{mn[0], mn[1], mn[2]}, V3{mx[0], mx[1], mx[2]}};
  const float4* bx = a->proxy_box.as<float4>();
  for (uint32_t i = 0; i < a->n_proxy; ++i) {
    if (!a->proxy_live[i]) continue;
    float4 lo = bx[2 * (size_t)i], hi = bx[2 * (size_t)i + 1];
    Box6 o{V3{lo.x, lo.y, lo.z}, V3{hi.x, hi.y, hi.z}};
    if (overlaps(o, q)) {
      if (out && k < capacity) out[k] = i;
      ++k;
    }
  }
  *out_count = k;
  return k > capacity ? BCH_OVERFLOW : BCH_OK;
}

int bch_bp_raycast_candidates(bch_context* c, const float* from, const float* to, uint32_t* out,
                              uint64_t capacity, uint64_t* out_count) {
  if (!c || !from || !to || !out_count) return fail(c, BCH_INVALID_ARG, "NULL argument");
  BCH_ON_BATCH(c, a);
  // Host scan over the proxy arena with the slab test of
  // broadphase/mod.rs:153-183 (t in [0, 1]; fmaxf / fminf keep the non-NaN
  // operand, as f32::max / f32::min do).
  uint64_t k = 0;
  const float4* bx = a->proxy_box.as<float4>();
  for (uint32_t i = 0; i < a->n_proxy; ++i) {
    if (!a->proxy_live[i]) continue;
    float4 lo4 = bx[2 * (size_t)i], hi4 = bx[2 * (size_t)i + 1];
    const float b[6] = {lo4.x, lo4.y, lo4.z, hi4.x, hi4.y, hi4.z};
    float t_min = 0.0f, t_max = 1.0f;
    bool hit = true;
    for (int ax = 0; ax < 3 && hit; ++ax) {
      float d = to[ax] - from[ax];
      float lo = d >= 0.0f ? b[ax] : b[3 + ax];
      float hi = d >= 0.0f ? b[3 + ax] : b[ax];
      if (d == 0.0f) {
        if (from[ax] < lo || from[ax] > hi) hit = false;
        continue;
      }
      float inv = 1.0f / d;
      t_min = fmaxf(t_min, (lo - from[ax]) * inv);
      t_max = fminf(t_max, (hi - from[ax]) * inv);
      if (t_min > t_max) hit = false;
    }
    if (hit) {
      if (out && k < capacity) out[k] = i;
      ++k;
    }
  }
  *out_count = k;
  return k > capacity ? BCH_OVERFLOW : BCH_OK;
}

// ------------------------------------------------------- collide_batch --
// The narrowphase over caller-given pairs on the batch context's bodies:
// pairs to keys on the device, the per-pair contact slots back in the batch
// layout (u8 count, four 7-float points).
static int collide_pairs_run(bch_context* a, const uint32_t* pairs, uint64_t n_pairs, const float* seeds,
                             uint8_t* out_count, float* out_points) {
  if (!n_pairs) return BCH_OK;
  hipStream_t s = a->stream;
  ENSURE(a, a->counters, sizeof(Counters));
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, a->counters.as<Counters>());
  a->P = n_pairs;
  a->cap_mode = false;
  ENSURE(a, a->pair_keys, n_pairs * sizeof(uint64_t));
  ENSURE(a, a->batch_pairs, n_pairs * sizeof(uint2));
  HIP_TRY(a, hipMemcpyAsync(a->batch_pairs.p, pairs, n_pairs * sizeof(uint2), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_pairs_to_keys, dim3(blocks(n_pairs)), dim3(kBlock), 0, s, (uint32_t)n_pairs,
                     a->batch_pairs.as<uint2>(), a->pair_keys.as<uint64_t>());
  float* d_seeds = nullptr;
  if (seeds) {
    ENSURE(a, a->batch_seeds, n_pairs * 3 * sizeof(float));
    d_seeds = a->batch_seeds.as<float>();
    HIP_TRY(a, hipMemcpyAsync(d_seeds, seeds, n_pairs * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  }
  int rc = narrow_split(a);
  if (!rc) rc = narrow_finish(a, d_seeds, false);
  if (rc) return rc;
  if (a->host_cnt.errors & 1u) return fail(a, BCH_HIP_ERROR, "box-box clip polygon overflow");
  ENSURE(a, a->batch_out_count, n_pairs);
  ENSURE(a, a->batch_out_pts, n_pairs * 28 * sizeof(float));
  hipLaunchKernelGGL(k_batch_out, dim3(blocks(n_pairs)), dim3(kBlock), 0, s, (uint32_t)n_pairs,
                     a->ccount.as<uint32_t>(), a->cslots.as<float4>(), a->batch_out_count.as<uint8_t>(),
                     a->batch_out_pts.as<float>());
  HIP_TRY(a, hipMemcpyAsync(out_count, a->batch_out_count.p, n_pairs, hipMemcpyDeviceToHost, s));
  HIP_TRY(a, hipMemcpyAsync(out_points, a->batch_out_pts.p, n_pairs * 28 * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(a, hipStreamSynchronize(s));
  return BCH_OK;
}

static int check_pairs(bch_context* c, const uint32_t* pairs, uint64_t n_pairs, uint32_t n) {
  if (n_pairs > 0x7fffffffull) return fail(c, BCH_INVALID_ARG, "too many pairs");
  for (uint64_t k = 0; k < 2 * n_pairs; ++k)
    if (pairs[k] >= n) return fail(c, BCH_INVALID_ARG, "pair index out of range");
  return BCH_OK;
}

int bch_collide_batch(bch_context* c, const bch_bodies* bodies, const uint32_t* pairs, uint64_t n_pairs,
                      const float* seeds, uint8_t* out_count, float* out_points) {
  if (!c) return BCH_INVALID_ARG;
  int rc = validate_bodies(c, bodies);
  if (rc) return rc;
  if (n_pairs && (!pairs || !out_count || !out_points)) return fail(c, BCH_INVALID_ARG, "NULL argument");
  rc = check_pairs(c, pairs, n_pairs, bodies->n);
  if (rc) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  BCH_ON_BATCH(c, a);
  a->batch_shapes = false;
  rc = upload_bodies(a, bodies, nullptr);
  if (rc) return from_batch(c, a, rc);
  a->n = bodies->n;
  a->batched = false;
  a->batch_shapes = true;
  a->batch_n = bodies->n;
  return from_batch(c, a, collide_pairs_run(a, pairs, n_pairs, seeds, out_count, out_points));
}

int bch_collide_batch_poses(bch_context* c, const float* position, const float* orientation, const uint32_t* pairs,
                            uint64_t n_pairs, const float* seeds, uint8_t* out_count, float* out_points) {
  if (!c) return BCH_INVALID_ARG;
  BCH_ON_BATCH(c, a);
  if (!a->batch_shapes) return fail(c, BCH_INVALID_ARG, "no shapes: call bch_collide_batch first");
  uint32_t n = a->batch_n;
  if (n && (!position || !orientation)) return fail(c, BCH_INVALID_ARG, "NULL pose arrays");
  if (n_pairs && (!pairs || !out_count || !out_points)) return fail(c, BCH_INVALID_ARG, "NULL argument");
  int rc = check_pairs(c, pairs, n_pairs, n);
  if (rc) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = a->stream;
  if (n) {
    HIP_TRY(c, hipMemcpyAsync(a->staging_a.p, position, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(a->staging_b.p, orientation, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_pack_poses, dim3(blocks(n)), dim3(kBlock), 0, s, n, a->staging_a.as<float>(),
                       a->staging_b.as<float>(), a->bodies.as<BodyRec>(), (const float*)nullptr, 0.0f,
                       (const uint8_t*)nullptr);
  }
  return from_batch(c, a, collide_pairs_run(a, pairs, n_pairs, seeds, out_count, out_points));
}

// ------------------------------------------------------ world pipeline --
int bch_world_set_bodies(bch_context* c, const bch_bodies* bodies, uint32_t n_worlds, const uint32_t* world_offsets) {
  if (!c) return BCH_INVALID_ARG;
  int rc = validate_bodies(c, bodies);
  if (rc) return rc;
  HIP_TRY(c, hipSetDevice(c->device));
  std::vector<uint32_t> world_of;
  uint32_t max_world = bodies->n;
  if (world_offsets) {
    max_world = 0;
    if (n_worlds == 0 || world_offsets[0] != 0 || world_offsets[n_worlds] != bodies->n)
      return fail(c, BCH_INVALID_ARG, "world_offsets must start at 0 and end at n");
    world_of.resize(bodies->n);
    for (uint32_t w = 0; w < n_worlds; ++w) {
      if (world_offsets[w + 1] < world_offsets[w]) return fail(c, BCH_INVALID_ARG, "world_offsets not monotone");
      max_world = std::max<uint32_t>(max_world, world_offsets[w + 1] - world_offsets[w]);
      for (uint32_t i = world_offsets[w]; i < world_offsets[w + 1]; ++i) world_of[i] = w;
    }
  } else if (n_worlds > 1) {
    return fail(c, BCH_INVALID_ARG, "n_worlds > 1 needs world_offsets");
  }
  release_graph(c);  // the captured graphs are shaped for the old bodies
  c->eager_P_cap = c->eager_C_cap = 0;
  release_depen_graph(c);
  c->mM = 0;  // cached manifolds name bodies by index
  c->has_bflags = c->has_bgm = false;
  c->aabb_valid = false;
  rc = upload_bodies(c, bodies, world_offsets ? world_of.data() : nullptr);
  if (rc) return rc;
  c->n = bodies->n;
  c->n_worlds = world_offsets ? n_worlds : 1;
  c->batched = world_offsets != nullptr;
  c->max_world = max_world;
  if (world_offsets) {
    ENSURE(c, c->world_offsets, (size_t)(n_worlds + 1) * sizeof(uint32_t));
    HIP_TRY(c, hipMemcpy(c->world_offsets.p, world_offsets, (size_t)(n_worlds + 1) * sizeof(uint32_t),
                         hipMemcpyHostToDevice));
  }
  return ensure_pass_buffers(c, std::max<uint32_t>(c->n, 1));
}

int bch_world_set_poses(bch_context* c, const float* position, const float* orientation) {
  if (!c || (c->n && (!position || !orientation))) return fail(c, BCH_INVALID_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if (!c->n) return BCH_OK;
  HIP_TRY(c, hipMemcpyAsync(c->staging_a.p, position, (size_t)c->n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(c->staging_b.p, orientation, (size_t)c->n * 4 * sizeof(float), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_pack_poses, dim3(blocks(c->n)), dim3(kBlock), 0, s, c->n, c->staging_a.as<float>(),
                     c->staging_b.as<float>(), c->bodies.as<BodyRec>());
  HIP_TRY(c, hipGetLastError());
  return BCH_OK;
}

int bch_world_set_poses_predicted(bch_context* c, const float* position, const float* orientation,
                                  const float* linear_velocity, float dt) {
  if (!c || (c->n && (!position || !orientation || !linear_velocity)))
    return fail(c, BCH_INVALID_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if (!c->n) return BCH_OK;
  ENSURE(c, c->vel, (size_t)c->n * 3 * sizeof(float));
  HIP_TRY(c, hipMemcpyAsync(c->staging_a.p, position, (size_t)c->n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(c->staging_b.p, orientation, (size_t)c->n * 4 * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(c->vel.p, linear_velocity, (size_t)c->n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_pack_poses, dim3(blocks(c->n)), dim3(kBlock), 0, s, c->n, c->staging_a.as<float>(),
                     c->staging_b.as<float>(), c->bodies.as<BodyRec>(), (const float*)c->vel.as<float>(), dt,
                     c->has_bflags ? (const uint8_t*)c->bflags.as<uint8_t>() : nullptr);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int bch_world_set_poses_predicted_device(bch_context* c, const float* d_position, const float* d_orientation,
                                         const float* d_linear_velocity, float dt) {
  if (!c || (c->n && (!d_position || !d_orientation || !d_linear_velocity)))
    return fail(c, BCH_INVALID_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  if (!c->n) return BCH_OK;
  hipLaunchKernelGGL(k_pack_poses, dim3(blocks(c->n)), dim3(kBlock), 0, c->stream, c->n, d_position, d_orientation,
                     c->bodies.as<BodyRec>(), d_linear_velocity, dt,
                     c->has_bflags ? (const uint8_t*)c->bflags.as<uint8_t>() : nullptr);
  HIP_TRY(c, hipGetLastError());
  return BCH_OK;
}

int bch_world_set_poses_device(bch_context* c, const float* d_position, const float* d_orientation) {
  if (!c || (c->n && (!d_position || !d_orientation))) return fail(c, BCH_INVALID_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  if (!c->n) return BCH_OK;
  hipLaunchKernelGGL(k_pack_poses, dim3(blocks(c->n)), dim3(kBlock), 0, c->stream, c->n, d_position, d_orientation,
                     c->bodies.as<BodyRec>());
  HIP_TRY(c, hipGetLastError());
  return BCH_OK;
}

int bch_world_collide(bch_context* c, float aabb_margin, uint64_t* out_pairs, uint64_t* out_contacts) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  // Once a pass has counted this world's pairs and contacts, the next one
  // runs the captured pass's kernels (counts kept on the device, no host
  // round trip inside the pass) against 1.25x those counts and reads the
  // counts back at its end.  A run that overflows a capacity or a walk
  // region goes again as the first pass did, with its round trips.
  bool done = false;
  if (c->eager_P_cap && c->n) {
    const uint64_t pc = c->P_cap, cc = c->C_cap;
    c->P_cap = c->eager_P_cap;
    c->C_cap = c->eager_C_cap;
    c->cap_mode = true;
    int rc = run_pass(c, aabb_margin);
    c->cap_mode = false;
    c->P_cap = pc;
    c->C_cap = cc;
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const Counters& h = c->host_cnt;
    if (!(h.errors & (8u | 16u | 128u))) {
      c->P = h.n_pairs;
      c->C = h.n_contacts;
      done = true;
    } else if (h.errors & 128u) {
      uint32_t r = c->walk_R;
      while (r < 2 * h.wave_over) r *= 2;
      c->walk_R = r;
    }
  }
  if (!done) {
    int rc = run_pass(c, aabb_margin);
    if (rc) return rc;
  }
  c->eager_P_cap = std::max<uint64_t>(c->P + c->P / 4, 1024);
  c->eager_C_cap = std::max<uint64_t>(c->C + c->C / 4, 1024);
  stage_times(c);
  if (c->host_cnt.errors & 1u) return fail(c, BCH_HIP_ERROR, "box-box clip polygon overflow");
  if (c->host_cnt.errors & 2u) return fail(c, BCH_HIP_ERROR, "EPA polytope store overflow");
  if (out_pairs) *out_pairs = c->P;
  if (out_contacts) *out_contacts = c->C;
  return BCH_OK;
}

// ---------------------------------------------------- captured world pass --
int bch_world_capture(bch_context* c, float aabb_margin, uint64_t pair_capacity, uint64_t contact_capacity) {
  if (!c) return BCH_INVALID_ARG;
  if (pair_capacity < 2 || pair_capacity > 0xffffffffull || contact_capacity < 1 ||
      contact_capacity > 0xffffffffull)
    return fail(c, BCH_INVALID_ARG, "capacities must be in [2, 2^32 - 1] pairs and [1, 2^32 - 1] contacts");
  HIP_TRY(c, hipSetDevice(c->device));
  release_graph(c);
  c->P_cap = pair_capacity;
  c->C_cap = contact_capacity;
  // One uncaptured pass in capacity mode sizes every buffer for these
  // poses; its counts are what the capture reports.
  c->cap_mode = true;
  int rc = size_walk(c, [&] { return run_pass(c, aabb_margin); });
  if (rc) {
    c->cap_mode = false;
    return rc;
  }
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  rc = capture_graph(c, [&] { return run_pass(c, aabb_margin); }, &g, &x);
  c->cap_mode = false;
  if (rc) return rc;
  c->graph = g;
  c->graph_exec = x;
  c->graph_gen = c->realloc_gen;
  c->graph_P_cap = c->P_cap;
  c->graph_C_cap = c->C_cap;
  return read_capacity_counts(c, nullptr, nullptr, nullptr, nullptr);
}

int bch_world_replay(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  if (!c->graph_exec) return fail(c, BCH_INVALID_ARG, "no captured pass: call bch_world_capture first");
  if (c->graph_gen != c->realloc_gen)
    return fail(c, BCH_INVALID_ARG, "a pass buffer was reallocated since the capture: capture again");
  HIP_TRY(c, hipSetDevice(c->device));
  HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream));
  return BCH_OK;
}

int bch_world_replay_counts(bch_context* c, uint64_t* n_pairs, uint64_t* n_contacts, uint64_t* found_pairs,
                            uint64_t* found_contacts) {
  if (!c) return BCH_INVALID_ARG;
  if (!c->graph_exec) return fail(c, BCH_INVALID_ARG, "no captured pass: call bch_world_capture first");
  HIP_TRY(c, hipSetDevice(c->device));
  return read_capacity_counts(c, n_pairs, n_contacts, found_pairs, found_contacts);
}

// --------------------------------------------- persistent contact manifolds --
int bch_world_set_body_flags(bch_context* c, const uint8_t* flags, const uint32_t* group_mask) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  size_t n = std::max<uint32_t>(c->n, 1);
  c->has_bflags = flags != nullptr && c->n;
  c->has_bgm = group_mask != nullptr && c->n;
  if (c->has_bflags) {
    ENSURE(c, c->bflags, n);
    HIP_TRY(c, hipMemcpyAsync(c->bflags.p, flags, c->n, hipMemcpyHostToDevice, s));
  }
  if (c->has_bgm) {
    ENSURE(c, c->bgm, n * sizeof(uint32_t));
    HIP_TRY(c, hipMemcpyAsync(c->bgm.p, group_mask, c->n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  }
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int bch_world_manifold_pass(bch_context* c, float aabb_margin, float linear_slop, float warm_start,
                            uint64_t* out_pairs, uint64_t* out_points) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  uint32_t n = c->n;
  hipStream_t s = c->stream;
  c->P = 0;
  c->C = 0;
  int rc = ensure_pass_buffers(c, std::max<uint32_t>(n, 1));
  if (rc) return rc;
  mark(c, 0);
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, c->counters.as<Counters>());
  if (n) {
    run_aabbs(c, aabb_margin);
    mark(c, 1);
    rc = run_broadphase(c, n, c->batched ? c->world_of.as<uint32_t>() : nullptr,
                        c->batched ? c->world_offsets.as<uint32_t>() : nullptr, false);
    if (rc) return rc;
  } else {
    for (int k = 1; k <= 3; ++k) mark(c, k);
  }
  uint64_t P = c->P;
  size_t Pm = std::max<uint64_t>(P, 1);
  int nw = c->mcur ^ 1;
  ENSURE(c, c->mkeys[nw], Pm * sizeof(uint64_t));
  ENSURE(c, c->mvalid[nw], Pm);
  ENSURE(c, c->ment[nw], Pm * 16 * sizeof(float4));
  ENSURE(c, c->mroute, Pm);
  ENSURE(c, c->msrc, Pm * sizeof(uint32_t));
  ENSURE(c, c->mcnt, (Pm + 1) * sizeof(uint32_t));
  ENSURE(c, c->moff, (Pm + 1) * sizeof(uint32_t));
  ENSURE(c, c->mslots, Pm * 12 * sizeof(float4));
  ENSURE(c, c->mseeds, Pm * 3 * sizeof(float));
  rc = narrow_split(c, c->bsph.as<float4>());
  if (rc) return rc;
  int cur = c->mcur;
  if (P)
    hipLaunchKernelGGL(k_mf_pre, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, c->pair_keys.as<uint64_t>(),
                       c->bodies.as<BodyRec>(), c->has_bflags ? c->bflags.as<uint8_t>() : nullptr,
                       c->has_bgm ? c->bgm.as<uint32_t>() : nullptr, c->mkeys[cur].as<uint64_t>(),
                       c->mvalid[cur].as<uint8_t>(), c->ment[cur].as<float4>(), (uint32_t)c->mM, linear_slop,
                       warm_start, c->cls.as<uint32_t>(), c->ccount.as<uint32_t>(), c->mseeds.as<float>(),
                       c->mroute.as<uint8_t>(), c->msrc.as<uint32_t>(), c->mvalid[nw].as<uint8_t>(),
                       c->ment[nw].as<float4>(), c->mcnt.as<uint32_t>(), c->mslots.as<float4>());
  c->cls_hist_ready = false;  // k_mf_pre rewrote classes: the class sort counts them itself
  rc = narrow_finish(c, c->mseeds.as<float>());
  if (rc) return rc;
  if (P)
    hipLaunchKernelGGL(k_mf_post, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, c->pair_keys.as<uint64_t>(),
                       c->bodies.as<BodyRec>(), c->ment[cur].as<float4>(), c->ccount.as<uint32_t>(),
                       c->cslots.as<float4>(), warm_start, c->mroute.as<uint8_t>(), c->msrc.as<uint32_t>(),
                       c->mvalid[nw].as<uint8_t>(), c->ment[nw].as<float4>(), c->mcnt.as<uint32_t>(),
                       c->mslots.as<float4>());
  rc = dev_scan(c, c->mcnt.as<uint32_t>(), c->moff.as<uint32_t>(), P, c->scan_part, s);
  if (rc) return rc;
  uint32_t total = 0;
  HIP_TRY(c, hipMemcpyAsync(&total, c->moff.as<uint32_t>() + P, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  ENSURE(c, c->mpts, (size_t)std::max<uint32_t>(total, 1) * 3 * sizeof(float4));
  if (P) {
    hipLaunchKernelGGL(k_mf_gather, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, c->mcnt.as<uint32_t>(),
                       c->moff.as<uint32_t>(), c->mslots.as<float4>(), c->mpts.as<float4>());
    HIP_TRY(c, hipMemcpyAsync(c->mkeys[nw].p, c->pair_keys.p, P * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
  }
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(s));
  c->mcur = nw;
  c->mM = P;
  c->mNP = total;
  stage_times(c);
  if (c->host_cnt.errors & 1u) return fail(c, BCH_HIP_ERROR, "box-box clip polygon overflow");
  if (out_pairs) *out_pairs = P;
  if (out_points) *out_points = total;
  return BCH_OK;
}

int bch_world_manifold_results(bch_context* c, uint32_t* pairs, uint8_t* route, uint32_t* point_offsets,
                               float* points) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  uint64_t P = c->mM;
  if (pairs && P) HIP_TRY(c, hipMemcpyAsync(pairs, c->pairs.p, P * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (route && P) HIP_TRY(c, hipMemcpyAsync(route, c->mroute.p, P, hipMemcpyDeviceToHost, s));
  if (point_offsets) {
    if (P)
      HIP_TRY(c, hipMemcpyAsync(point_offsets, c->moff.p, (P + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    else
      point_offsets[0] = 0;
  }
  if (points && c->mNP)
    HIP_TRY(c, hipMemcpyAsync(points, c->mpts.p, c->mNP * 12 * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int bch_world_manifold_impulses(bch_context* c, const float* impulses) {
  if (!c) return BCH_INVALID_ARG;
  if (!impulses && c->mNP) return fail(c, BCH_INVALID_ARG, "NULL impulses");
  HIP_TRY(c, hipSetDevice(c->device));
  if (!c->mM || !c->mNP) return BCH_OK;
  hipStream_t s = c->stream;
  ENSURE(c, c->mimp, c->mNP * sizeof(float));
  HIP_TRY(c, hipMemcpyAsync(c->mimp.p, impulses, c->mNP * sizeof(float), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mf_impulses, dim3(blocks(c->mM)), dim3(kBlock), 0, s, (uint32_t)c->mM, c->mcnt.as<uint32_t>(),
                     c->moff.as<uint32_t>(), c->mvalid[c->mcur].as<uint8_t>(), c->mimp.as<float>(),
                     c->ment[c->mcur].as<float4>());
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int bch_world_manifold_cache(bch_context* c, uint8_t* valid, float* entries) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  uint64_t M = c->mM;
  if (valid && M) HIP_TRY(c, hipMemcpyAsync(valid, c->mvalid[c->mcur].p, M, hipMemcpyDeviceToHost, s));
  if (entries && M)
    HIP_TRY(c, hipMemcpyAsync(entries, c->ment[c->mcur].p, M * 64 * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int bch_world_manifold_reset(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  c->mM = 0;
  c->mNP = 0;
  return BCH_OK;
}

// ------------------------------------------------------ triangle meshes --
static int mesh_add_on(bch_context* c, const float* vertices, uint32_t n_vertices, const uint32_t* indices,
                 uint32_t n_triangles, uint32_t* out_mesh) {
  if (!c || !out_mesh || (n_vertices && !vertices) || (n_triangles && !indices))
    return fail(c, BCH_INVALID_ARG, "NULL argument");
  for (uint64_t k = 0; k < 3ull * n_triangles; ++k)
    if (indices[k] >= n_vertices) return fail(c, BCH_INVALID_ARG, "triangle index out of range");
  MeshRec r{};
  r.node_off = (uint32_t)(c->mesh_nodes.size() / 2);
  r.order_off = (uint32_t)c->mesh_order.size();
  r.tri_off = (uint32_t)(c->mesh_tris.size() / 3);
  r.ntri = n_triangles;
  // triangles, their AABBs and the local bounds (triangle_mesh.rs:99-113)
  float blo[3] = {0.0f, 0.0f, 0.0f}, bhi[3] = {0.0f, 0.0f, 0.0f};
  for (uint32_t t = 0; t < n_triangles; ++t) {
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      const float* p = vertices + 3 * (size_t)indices[3 * (size_t)t + k];
      c->mesh_tris.push_back(make_float4(p[0], p[1], p[2], 0.0f));
    }
    const float* a = vertices + 3 * (size_t)indices[3 * (size_t)t];
    const float* b = vertices + 3 * (size_t)indices[3 * (size_t)t + 1];
    const float* d = vertices + 3 * (size_t)indices[3 * (size_t)t + 2];
    for (int k = 0; k < 3; ++k) {  // a.min(b.min(c)), a.max(b.max(c))
      float mbc = d[k] < b[k] ? d[k] : b[k], xbc = d[k] > b[k] ? d[k] : b[k];
      lo[k] = mbc < a[k] ? mbc : a[k];
      hi[k] = xbc > a[k] ? xbc : a[k];
    }
    c->mesh_tri_aabb.push_back(make_float4(lo[0], lo[1], lo[2], 0.0f));
    c->mesh_tri_aabb.push_back(make_float4(hi[0], hi[1], hi[2], 0.0f));
    for (int k = 0; k < 3; ++k) {
      blo[k] = t == 0 ? lo[k] : (lo[k] < blo[k] ? lo[k] : blo[k]);
      bhi[k] = t == 0 ? hi[k] : (hi[k] > bhi[k] ? hi[k] : bhi[k]);
    }
  }
  // contact_breaking_threshold over the local bounds (shapes/mod.rs:388-392)
  float dx = bhi[0] - blo[0], dy = bhi[1] - blo[1], dz = bhi[2] - blo[2];
  float cx = (blo[0] + bhi[0]) * 0.5f, cy = (blo[1] + bhi[1]) * 0.5f, cz = (blo[2] + bhi[2]) * 0.5f;
  float disc = sqrtf((dx * dx + dy * dy) + dz * dz) * 0.5f + sqrtf((cx * cx + cy * cy) + cz * cz);
  r.threshold = disc * 0.02f;
  MeshBuilder mb{vertices, indices, {}, {}};
  mb.order.resize(n_triangles);
  for (uint32_t t = 0; t < n_triangles; ++t) mb.order[t] = t;
  mb.build(0, n_triangles);
  c->mesh_nodes.insert(c->mesh_nodes.end(), mb.nodes.begin(), mb.nodes.end());
  c->mesh_order.insert(c->mesh_order.end(), mb.order.begin(), mb.order.end());
  *out_mesh = (uint32_t)c->mesh_recs.size();
  c->mesh_recs.push_back(r);
  c->mesh_dirty = true;
  return BCH_OK;
}

int bch_mesh_add(bch_context* c, const float* vertices, uint32_t n_vertices, const uint32_t* indices,
                 uint32_t n_triangles, uint32_t* out_mesh) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  BCH_ON_BATCH(c, a);
  return from_batch(c, a, mesh_add_on(a, vertices, n_vertices, indices, n_triangles, out_mesh));
}

static int mesh_clear_on(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  c->mesh_recs.clear();
  c->mesh_nodes.clear();
  c->mesh_tris.clear();
  c->mesh_tri_aabb.clear();
  c->mesh_order.clear();
  c->mesh_dirty = true;
  return BCH_OK;
}

int bch_mesh_clear(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  BCH_ON_BATCH(c, a);
  return from_batch(c, a, mesh_clear_on(a));
}

// Upload the host-side mesh tables when a mesh was added or cleared.
static int sync_meshes(bch_context* c, hipStream_t s) {
  if (!c->mesh_dirty) return BCH_OK;
  size_t nm = std::max<size_t>(c->mesh_recs.size(), 1);
  ENSURE(c, c->d_mesh_recs, nm * sizeof(MeshRec));
  ENSURE(c, c->d_mesh_nodes, std::max<size_t>(c->mesh_nodes.size(), 1) * sizeof(float4));
  ENSURE(c, c->d_mesh_tris, std::max<size_t>(c->mesh_tris.size(), 1) * sizeof(float4));
  ENSURE(c, c->d_mesh_tri_aabb, std::max<size_t>(c->mesh_tri_aabb.size(), 1) * sizeof(float4));
  ENSURE(c, c->d_mesh_order, std::max<size_t>(c->mesh_order.size(), 1) * sizeof(uint32_t));
  if (!c->mesh_recs.empty())
    HIP_TRY(c, hipMemcpyAsync(c->d_mesh_recs.p, c->mesh_recs.data(), c->mesh_recs.size() * sizeof(MeshRec),
                              hipMemcpyHostToDevice, s));
  if (!c->mesh_nodes.empty())
    HIP_TRY(c, hipMemcpyAsync(c->d_mesh_nodes.p, c->mesh_nodes.data(), c->mesh_nodes.size() * sizeof(float4),
                              hipMemcpyHostToDevice, s));
  if (!c->mesh_tris.empty()) {
    HIP_TRY(c, hipMemcpyAsync(c->d_mesh_tris.p, c->mesh_tris.data(), c->mesh_tris.size() * sizeof(float4),
                              hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemcpyAsync(c->d_mesh_tri_aabb.p, c->mesh_tri_aabb.data(),
                              c->mesh_tri_aabb.size() * sizeof(float4), hipMemcpyHostToDevice, s));
  }
  if (!c->mesh_order.empty())
    HIP_TRY(c, hipMemcpyAsync(c->d_mesh_order.p, c->mesh_order.data(), c->mesh_order.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, s));
  c->mesh_dirty = false;
  return BCH_OK;
}

static int collide_mesh_batch_on(bch_context* c, const bch_bodies* bodies, const uint32_t* pairs, const float* mesh_poses,
                           const uint8_t* mesh_first, uint64_t n_pairs, const float* seeds, uint8_t* out_count,
                           float* out_points) {
  if (!c) return BCH_INVALID_ARG;
  if (n_pairs && (!pairs || !mesh_poses || !out_count || !out_points)) return fail(c, BCH_INVALID_ARG, "NULL argument");
  if (n_pairs > 0x7fffffffull) return fail(c, BCH_INVALID_ARG, "too many pairs");
  int rc = validate_bodies(c, bodies);
  if (rc) return rc;
  for (uint64_t k = 0; k < n_pairs; ++k) {
    uint32_t b = pairs[2 * k], m = pairs[2 * k + 1];
    if (b >= bodies->n || m >= c->mesh_recs.size()) return fail(c, BCH_INVALID_ARG, "pair index out of range");
    if (bodies->shape_type[b] == BCH_PLANE)
      return fail(c, BCH_INVALID_ARG, "plane-vs-mesh pairs are not part of this path");
  }
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  rc = upload_bodies(c, bodies, nullptr);
  if (rc) return rc;
  rc = sync_meshes(c, s);
  if (rc) return rc;
  if (!n_pairs) return BCH_OK;
  uint64_t P = n_pairs;
  ENSURE(c, c->counters, sizeof(Counters));
  ENSURE(c, c->mx_pairs, P * sizeof(uint2));
  ENSURE(c, c->mx_poses, P * 7 * sizeof(float));
  ENSURE(c, c->mx_seeds, P * 3 * sizeof(float));
  ENSURE(c, c->mx_first, P);
  ENSURE(c, c->mx_count, (P + 1) * sizeof(uint32_t));
  ENSURE(c, c->mx_offs, (P + 1) * sizeof(uint32_t));
  ENSURE(c, c->mx_out_count, P);
  ENSURE(c, c->mx_out_pts, P * 28 * sizeof(float));
  HIP_TRY(c, hipMemcpyAsync(c->mx_pairs.p, pairs, P * sizeof(uint2), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(c->mx_poses.p, mesh_poses, P * 7 * sizeof(float), hipMemcpyHostToDevice, s));
  if (seeds) HIP_TRY(c, hipMemcpyAsync(c->mx_seeds.p, seeds, P * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  if (mesh_first) HIP_TRY(c, hipMemcpyAsync(c->mx_first.p, mesh_first, P, hipMemcpyHostToDevice, s));
  Counters* cnt = c->counters.as<Counters>();
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, cnt);
  const MeshRec* recs = c->d_mesh_recs.as<MeshRec>();
  hipLaunchKernelGGL(k_mesh_count, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, c->mx_pairs.as<uint2>(),
                     (const float*)c->mx_poses.as<float>(), c->bodies.as<BodyRec>(), c->hull.as<float4>(), recs,
                     c->d_mesh_nodes.as<float4>(), c->d_mesh_order.as<uint32_t>(), c->d_mesh_tri_aabb.as<float4>(),
                     c->mx_count.as<uint32_t>(), cnt);
  rc = dev_scan(c, c->mx_count.as<uint32_t>(), c->mx_offs.as<uint32_t>(), P, c->scan_part, s);
  if (rc) return rc;
  uint32_t S = 0;
  HIP_TRY(c, hipMemcpyAsync(&S, c->mx_offs.as<uint32_t>() + P, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  size_t Sm = std::max<uint32_t>(S, 1);
  ENSURE(c, c->mx_sub, Sm * sizeof(uint2));
  ENSURE(c, c->mx_sub_cnt, Sm * sizeof(uint32_t));
  ENSURE(c, c->mx_sub_slots, Sm * 2 * sizeof(float4));
  ENSURE(c, c->mx_epa_q, Sm * 4 * sizeof(float4));
  ENSURE(c, c->mx_epa_q2, Sm * 4 * sizeof(float4));
  hipLaunchKernelGGL(k_mesh_emit, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, c->mx_pairs.as<uint2>(),
                     (const float*)c->mx_poses.as<float>(), c->bodies.as<BodyRec>(), c->hull.as<float4>(), recs,
                     c->d_mesh_nodes.as<float4>(), c->d_mesh_order.as<uint32_t>(), c->d_mesh_tri_aabb.as<float4>(),
                     c->mx_offs.as<uint32_t>(), c->mx_sub.as<uint2>());
  if (S) {
    hipLaunchKernelGGL(k_mesh_detect, dim3(blocks(S)), dim3(kBlock), 0, s, S, c->mx_sub.as<uint2>(),
                       c->mx_pairs.as<uint2>(), (const float*)c->mx_poses.as<float>(),
                       seeds ? (const float*)c->mx_seeds.as<float>() : nullptr, c->bodies.as<BodyRec>(),
                       c->hull.as<float4>(), recs, c->d_mesh_tris.as<float4>(), c->mx_sub_cnt.as<uint32_t>(),
                       c->mx_sub_slots.as<float4>(), c->mx_epa_q.as<float4>(), cnt);
    MeshEpaSrc src{c->mx_sub.as<uint2>(), c->mx_pairs.as<uint2>(), c->mx_poses.as<float>(), c->bodies.as<BodyRec>(),
                   c->hull.as<float4>(), recs, c->d_mesh_tris.as<float4>()};
    hipLaunchKernelGGL(k_epa_mesh_dual, dim3(c->epa_grid), dim3(64), 0, s, &cnt->n_epa, c->mx_epa_q.as<float4>(),
                       src, c->mx_sub_cnt.as<uint32_t>(), c->mx_sub_slots.as<float4>(), c->mx_epa_q2.as<float4>(),
                       cnt);
    hipLaunchKernelGGL(k_epa_mesh, dim3(kEpaDeferGrid), dim3(64), 0, s, &cnt->n_epa_defer,
                       c->mx_epa_q2.as<float4>(), src, c->mx_sub_cnt.as<uint32_t>(), c->mx_sub_slots.as<float4>(),
                       cnt, c->mx_epa_q.as<float4>(), c->epa_wave_faces);
    hipLaunchKernelGGL(k_epa_mesh_big, dim3(kEpaBigGrid), dim3(64), 0, s, &cnt->n_epa_over,
                       c->mx_epa_q.as<float4>(), src, c->mx_sub_cnt.as<uint32_t>(), c->mx_sub_slots.as<float4>(),
                       cnt);
  }
  hipLaunchKernelGGL(k_mesh_reduce, dim3(blocks(P)), dim3(kBlock), 0, s, (uint32_t)P, c->mx_offs.as<uint32_t>(),
                     c->mx_sub_cnt.as<uint32_t>(), c->mx_sub_slots.as<float4>(),
                     mesh_first ? (const uint8_t*)c->mx_first.as<uint8_t>() : nullptr, c->mx_out_count.as<uint8_t>(),
                     c->mx_out_pts.as<float>());
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(out_count, c->mx_out_count.p, P, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(out_points, c->mx_out_pts.p, P * 28 * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  if (c->host_cnt.errors & 32u) return fail(c, BCH_HIP_ERROR, "mesh BVH deeper than the traversal stack");
  if (c->host_cnt.errors & 2u) return fail(c, BCH_HIP_ERROR, "EPA polytope store overflow");
  return BCH_OK;
}

int bch_collide_mesh_batch(bch_context* c, const bch_bodies* bodies, const uint32_t* pairs, const float* mesh_poses,
                           const uint8_t* mesh_first, uint64_t n_pairs, const float* seeds, uint8_t* out_count,
                           float* out_points) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  BCH_ON_BATCH(c, a);
  return from_batch(c, a, collide_mesh_batch_on(a, bodies, pairs, mesh_poses, mesh_first, n_pairs, seeds, out_count, out_points));
}

// --------------------------------------------------------- CCD sweeps --
// Shared checks and uploads of the sweep batches: query arrays to the device.
static int sweep_upload(bch_context* c, uint64_t n, const uint32_t* idx, const uint32_t* caster, const float* from,
                        const float* to, const float* radius, const float* pose, hipStream_t s) {
  ENSURE(c, c->counters, sizeof(Counters));
  ENSURE(c, c->sw_target, n * sizeof(uint32_t));
  ENSURE(c, c->sw_from, n * 3 * sizeof(float));
  ENSURE(c, c->sw_to, n * 3 * sizeof(float));
  ENSURE(c, c->sw_hit, n);
  ENSURE(c, c->sw_out, n * 7 * sizeof(float));
  HIP_TRY(c, hipMemcpyAsync(c->sw_target.p, idx, n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(c->sw_from.p, from, n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipMemcpyAsync(c->sw_to.p, to, n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  if (caster) {
    ENSURE(c, c->sw_caster, n * sizeof(uint32_t));
    HIP_TRY(c, hipMemcpyAsync(c->sw_caster.p, caster, n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  } else {
    ENSURE(c, c->sw_radius, n * sizeof(float));
    HIP_TRY(c, hipMemcpyAsync(c->sw_radius.p, radius, n * sizeof(float), hipMemcpyHostToDevice, s));
  }
  if (pose) {
    ENSURE(c, c->sw_pose, n * 7 * sizeof(float));
    HIP_TRY(c, hipMemcpyAsync(c->sw_pose.p, pose, n * 7 * sizeof(float), hipMemcpyHostToDevice, s));
  }
  Counters* cnt = c->counters.as<Counters>();
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, cnt);
  return BCH_OK;
}

static int sweep_download(bch_context* c, uint64_t n, uint8_t* out_hit, float* out_sweep, hipStream_t s) {
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipMemcpyAsync(out_hit, c->sw_hit.p, n, hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(out_sweep, c->sw_out.p, n * 7 * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipMemcpyAsync(&c->host_cnt, c->counters.p, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  if (c->host_cnt.errors & 32u) return fail(c, BCH_HIP_ERROR, "mesh BVH deeper than the traversal stack");
  return BCH_OK;
}

static int sweep_batch_on(bch_context* c, const bch_bodies* bodies, const uint32_t* target, const uint32_t* caster,
                    const float* from, const float* to, const float* radius, uint64_t n, uint8_t* out_hit,
                    float* out_sweep) {
  if (!c) return BCH_INVALID_ARG;
  if (n && (!target || !from || !to || !out_hit || !out_sweep || (!caster && !radius)))
    return fail(c, BCH_INVALID_ARG, "NULL argument");
  if (n > 0x7fffffffull) return fail(c, BCH_INVALID_ARG, "too many queries");
  int rc = validate_bodies(c, bodies);
  if (rc) return rc;
  for (uint64_t k = 0; k < n; ++k) {
    if (target[k] >= bodies->n) return fail(c, BCH_INVALID_ARG, "target index out of range");
    if (caster) {
      if (caster[k] >= bodies->n) return fail(c, BCH_INVALID_ARG, "caster index out of range");
      if (bodies->shape_type[caster[k]] == BCH_PLANE) return fail(c, BCH_INVALID_ARG, "a cast shape must be convex");
    }
  }
  if (!n) return BCH_OK;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  rc = upload_bodies(c, bodies, nullptr);
  if (rc) return rc;
  rc = sweep_upload(c, n, target, caster, from, to, radius, nullptr, s);
  if (rc) return rc;
  // group the queries by shape class (one 8-bit radix pass, stable)
  ENSURE(c, c->sw_keys, n * 2 * sizeof(uint32_t));
  ENSURE(c, c->sw_order, n * 2 * sizeof(uint32_t));
  uint32_t* keys = c->sw_keys.as<uint32_t>();
  uint32_t* ord = c->sw_order.as<uint32_t>();
  hipLaunchKernelGGL(k_sweep_keys, dim3(blocks(n)), dim3(kBlock), 0, s, (uint32_t)n, c->sw_target.as<uint32_t>(),
                     caster ? (const uint32_t*)c->sw_caster.as<uint32_t>() : nullptr, c->bodies.as<BodyRec>(), keys,
                     ord);
  const int shift0 = 0;
  bool second = false;
  rc = dev_radix_sort<uint32_t>(c, keys, keys + n, ord, ord + n, n, &shift0, 1, s, &second);
  if (rc) return rc;
  const uint32_t* order = second ? ord + n : ord;
  hipLaunchKernelGGL(k_sweep, dim3(blocks(n)), dim3(kBlock), 0, s, (uint32_t)n, order, c->sw_target.as<uint32_t>(),
                     caster ? (const uint32_t*)c->sw_caster.as<uint32_t>() : nullptr, c->sw_from.as<float>(),
                     c->sw_to.as<float>(), caster ? nullptr : (const float*)c->sw_radius.as<float>(),
                     c->bodies.as<BodyRec>(), c->hull.as<float4>(), c->sw_hit.as<uint8_t>(), c->sw_out.as<float>());
  return sweep_download(c, n, out_hit, out_sweep, s);
}

int bch_sweep_batch(bch_context* c, const bch_bodies* bodies, const uint32_t* target, const uint32_t* caster,
                    const float* from, const float* to, const float* radius, uint64_t n, uint8_t* out_hit,
                    float* out_sweep) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  BCH_ON_BATCH(c, a);
  return from_batch(c, a, sweep_batch_on(a, bodies, target, caster, from, to, radius, n, out_hit, out_sweep));
}

static int sweep_mesh_batch_on(bch_context* c, const bch_bodies* bodies, const uint32_t* mesh_of, const float* mesh_poses,
                         const uint32_t* caster, const float* from, const float* to, const float* radius, uint64_t n,
                         uint8_t* out_hit, float* out_sweep) {
  if (!c) return BCH_INVALID_ARG;
  if (n && (!mesh_of || !mesh_poses || !from || !to || !out_hit || !out_sweep || (!caster && !radius)))
    return fail(c, BCH_INVALID_ARG, "NULL argument");
  if (caster && !bodies) return fail(c, BCH_INVALID_ARG, "convex casts need the bodies");
  if (n > 0x7fffffffull) return fail(c, BCH_INVALID_ARG, "too many queries");
  int rc;
  if (caster) {
    rc = validate_bodies(c, bodies);
    if (rc) return rc;
  }
  for (uint64_t k = 0; k < n; ++k) {
    if (mesh_of[k] >= c->mesh_recs.size()) return fail(c, BCH_INVALID_ARG, "mesh index out of range");
    if (caster) {
      if (caster[k] >= bodies->n) return fail(c, BCH_INVALID_ARG, "caster index out of range");
      if (bodies->shape_type[caster[k]] == BCH_PLANE) return fail(c, BCH_INVALID_ARG, "a cast shape must be convex");
    }
  }
  if (!n) return BCH_OK;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if (caster) {
    rc = upload_bodies(c, bodies, nullptr);
    if (rc) return rc;
  }
  rc = sync_meshes(c, s);
  if (rc) return rc;
  rc = sweep_upload(c, n, mesh_of, caster, from, to, radius, mesh_poses, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sweep_mesh, dim3(blocks(n)), dim3(kBlock), 0, s, (uint32_t)n, c->sw_target.as<uint32_t>(),
                     (const float*)c->sw_pose.as<float>(), caster ? (const uint32_t*)c->sw_caster.as<uint32_t>() : nullptr,
                     c->sw_from.as<float>(), c->sw_to.as<float>(),
                     caster ? nullptr : (const float*)c->sw_radius.as<float>(),
                     caster ? (const BodyRec*)c->bodies.as<BodyRec>() : nullptr,
                     caster ? (const float4*)c->hull.as<float4>() : nullptr, c->d_mesh_recs.as<MeshRec>(),
                     c->d_mesh_nodes.as<float4>(), c->d_mesh_order.as<uint32_t>(), c->d_mesh_tris.as<float4>(),
                     c->sw_hit.as<uint8_t>(), c->sw_out.as<float>(), c->counters.as<Counters>());
  return sweep_download(c, n, out_hit, out_sweep, s);
}

int bch_sweep_mesh_batch(bch_context* c, const bch_bodies* bodies, const uint32_t* mesh_of, const float* mesh_poses,
                         const uint32_t* caster, const float* from, const float* to, const float* radius, uint64_t n,
                         uint8_t* out_hit, float* out_sweep) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  BCH_ON_BATCH(c, a);
  return from_batch(c, a, sweep_mesh_batch_on(a, bodies, mesh_of, mesh_poses, caster, from, to, radius, n, out_hit, out_sweep));
}

// ---------------------------------------------- captured de-penetration --
int bch_depen_capture(bch_context* c, float aabb_margin, float lr, float* d_position, float* d_orientation,
                      uint64_t pair_capacity, uint32_t history_capacity, uint64_t* found_pairs) {
  if (!c) return BCH_INVALID_ARG;
  if (found_pairs) *found_pairs = 0;
  if (!c->n) return fail(c, BCH_INVALID_ARG, "no bodies: call bch_world_set_bodies first");
  if (!d_position || !d_orientation) return fail(c, BCH_INVALID_ARG, "NULL pose arrays");
  if (pair_capacity < 2 || pair_capacity > 0xffffffffull)
    return fail(c, BCH_INVALID_ARG, "pair capacity must be in [2, 2^32 - 1]");
  HIP_TRY(c, hipSetDevice(c->device));
  release_depen_graph(c);
  c->P_cap = pair_capacity;
  c->dhist_cap = std::max<uint32_t>(history_capacity, 1u);
  ENSURE(c, c->dhist, (size_t)c->dhist_cap * sizeof(DepenRecord));
  ENSURE(c, c->dstep, sizeof(uint32_t));
  // An uncaptured step without the update sizes the buffers at the current
  // poses (which it leaves unchanged).
  c->cap_mode = true;
  int rc = size_walk(c, [&] { return run_depen_step(c, aabb_margin, lr, d_position, d_orientation, false); });
  if (!rc && found_pairs) *found_pairs = c->host_cnt.n_pairs_total;
  if (!rc && c->host_cnt.n_pairs_total > pair_capacity)
    rc = fail(c, BCH_OVERFLOW, "pair capacity exceeded at the current poses");
  if (rc) {
    c->cap_mode = false;
    return rc;
  }
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  rc = capture_graph(c, [&] { return run_depen_step(c, aabb_margin, lr, d_position, d_orientation, true); }, &g, &x);
  c->cap_mode = false;
  if (rc) return rc;
  c->dgraph = g;
  c->dgraph_exec = x;
  c->dgraph_gen = c->realloc_gen;
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, 1u, c->dstep.as<uint32_t>(), 0u);
  HIP_TRY(c, hipGetLastError());
  return BCH_OK;
}

int bch_depen_replay(bch_context* c, uint32_t steps) {
  if (!c) return BCH_INVALID_ARG;
  if (!c->dgraph_exec) return fail(c, BCH_INVALID_ARG, "no captured step: call bch_depen_capture first");
  if (c->dgraph_gen != c->realloc_gen)
    return fail(c, BCH_INVALID_ARG, "a pass buffer was reallocated since the capture: capture again");
  HIP_TRY(c, hipSetDevice(c->device));
  for (uint32_t k = 0; k < steps; ++k) HIP_TRY(c, hipGraphLaunch(c->dgraph_exec, c->stream));
  return BCH_OK;
}

int bch_depen_history(bch_context* c, float* loss, float* max_pen, uint32_t* n_pairs, uint32_t* n_pen,
                      uint32_t capacity, uint32_t* count, int reset) {
  if (!c || !count) return fail(c, BCH_INVALID_ARG, "NULL argument");
  if (!c->dgraph_exec) return fail(c, BCH_INVALID_ARG, "no captured step: call bch_depen_capture first");
  HIP_TRY(c, hipSetDevice(c->device));
  uint32_t steps = 0;
  HIP_TRY(c, hipMemcpyAsync(&steps, c->dstep.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  uint32_t held = std::min(steps, c->dhist_cap);
  std::vector<DepenRecord> rec(held);
  if (held)
    HIP_TRY(c, hipMemcpyAsync(rec.data(), c->dhist.p, held * sizeof(DepenRecord), hipMemcpyDeviceToHost, c->stream));
  if (reset) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->stream, 1u, c->dstep.as<uint32_t>(), 0u);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  *count = steps;
  bool over = false;
  for (uint32_t k = 0; k < held; ++k) {
    if (k < capacity) {
      if (loss) loss[k] = rec[k].loss;
      if (max_pen) max_pen[k] = rec[k].max_pen;
      if (n_pairs) n_pairs[k] = rec[k].n_pairs;
      if (n_pen) n_pen[k] = rec[k].n_pen;
    }
    over = over || (rec[k].flags & (8u | 128u)) != 0;
    if (rec[k].flags & 128u) c->walk_grow = true;
  }
  if (over) return fail(c, BCH_OVERFLOW, "a step's pairs exceeded the capacity: capture again with a larger one");
  if (steps > c->dhist_cap || steps > capacity) return fail(c, BCH_OVERFLOW, "more steps than history capacity");
  return BCH_OK;
}

int bch_world_release_graph(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  (void)hipSetDevice(c->device);
  release_graph(c);
  return BCH_OK;
}

int bch_depen_release(bch_context* c) {
  if (!c) return BCH_INVALID_ARG;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  release_depen_graph(c);
  return BCH_OK;
}

static int run_sd(bch_context* c, const uint32_t* d_pairs, uint64_t P, const float* d_gin, float* d_out,
                  uint8_t* d_status, bool bwd) {
  if (!c) return BCH_INVALID_ARG;
  if (P && (!d_pairs || !d_out || (bwd && !d_gin))) return fail(c, BCH_INVALID_ARG, "NULL argument");
  if (P > 0xffffffffull) return fail(c, BCH_INVALID_ARG, "more than 2^32 - 1 pairs");
  if (P && !c->n) return fail(c, BCH_INVALID_ARG, "no bodies: call bch_world_set_bodies first");
  HIP_TRY(c, hipSetDevice(c->device));
  ENSURE(c, c->epa_q, std::max<uint64_t>(P, 1) * kSdEpaEntry * sizeof(float4));
  ENSURE(c, c->sd_defer, std::max<uint64_t>(P, 1) * kSdEpaEntry * sizeof(float4));
  if (bwd) ENSURE(c, c->sd_rec, std::max<uint64_t>(P, 1) * 4 * sizeof(float4));
  if (bwd) ENSURE(c, c->sd_brec, std::max<uint64_t>(P, 1) * kSdBandRec * sizeof(float4));
  hipStream_t s = c->stream;
  Counters* cnt = c->counters.as<Counters>();
  hipLaunchKernelGGL(k_reset_epa, dim3(1), dim3(64), 0, s, cnt);
  if (P) {
    constexpr int G = 8;
    uint32_t g = std::max<uint32_t>(1u, std::min<uint32_t>(blocks(P, kBlock / G), 8192u));
    if (bwd) {
      hipLaunchKernelGGL((k_sd_walk<G, true>), dim3(g), dim3(kBlock), 0, s, (uint32_t)P, (const uint32_t*)nullptr,
                         c->n, d_pairs,
                         c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_gin, d_out, d_status,
                         c->epa_q.as<float4>(), cnt, c->sd_brec.as<float4>());
      hipLaunchKernelGGL(k_sd_band_vjp, dim3(std::max<uint32_t>(1u, std::min<uint32_t>(blocks(P), 8192u))),
                         dim3(kBlock), 0, s, (uint32_t)P, (const uint32_t*)nullptr, c->sd_brec.as<float4>(), d_pairs,
                         c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_gin, d_out);
      sd_backward_epa(c, d_pairs, d_gin, d_out, d_status, s);
    } else {
      hipLaunchKernelGGL((k_sd_walk<G, false>), dim3(g), dim3(kBlock), 0, s, (uint32_t)P, (const uint32_t*)nullptr,
                         c->n, d_pairs,
                         c->bodies.as<BodyRec>(), c->hull.as<float4>(), d_gin, d_out, d_status,
                         c->epa_q.as<float4>(), cnt);
      sd_forward_epa(c, d_pairs, d_out, d_status, s);
    }
  }
  HIP_TRY(c, hipGetLastError());
  return BCH_OK;
}

int bch_sd_forward(bch_context* c, const uint32_t* d_pairs, uint64_t n_pairs, float* d_out, uint8_t* d_status) {
  return run_sd(c, d_pairs, n_pairs, nullptr, d_out, d_status, false);
}

int bch_sd_backward(bch_context* c, const uint32_t* d_pairs, uint64_t n_pairs, const float* d_grad_out,
                    float* d_grad_pose, uint8_t* d_status) {
  return run_sd(c, d_pairs, n_pairs, d_grad_out, d_grad_pose, d_status, true);
}

int bch_world_routes(bch_context* c, uint8_t* routes) {
  if (!c || (!routes && c->P)) return fail(c, BCH_INVALID_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  if (!c->P) return BCH_OK;
  std::vector<uint32_t> cls(c->P);
  HIP_TRY(c, hipMemcpyAsync(cls.data(), c->cls.p, c->P * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (uint64_t k = 0; k < c->P; ++k) routes[k] = (uint8_t)cls[k];
  return BCH_OK;
}

int bch_world_results(bch_context* c, uint32_t* pairs, uint32_t* offsets, float* contacts) {
  if (!c) return BCH_INVALID_ARG;
  HIP_TRY(c, hipSetDevice(c->device));
  hipStream_t s = c->stream;
  if (pairs && c->P)
    HIP_TRY(c, hipMemcpyAsync(pairs, c->pairs.p, c->P * 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (offsets)
    HIP_TRY(c, hipMemcpyAsync(offsets, c->coffsets.p, (c->P + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  if (contacts && c->C)
    HIP_TRY(c, hipMemcpyAsync(contacts, c->contacts.p, c->C * 7 * sizeof(float), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return BCH_OK;
}

int bch_world_results_device(bch_context* c, const uint32_t** d_pairs, const uint32_t** d_offsets,
                             const float** d_contacts) {
  if (!c) return BCH_INVALID_ARG;
  if (d_pairs) *d_pairs = c->pairs.as<uint32_t>();
  if (d_offsets) *d_offsets = c->coffsets.as<uint32_t>();
  if (d_contacts) *d_contacts = c->contacts.as<float>();
  return BCH_OK;
}

// ------------------------------------------------ no allocation (static) --
// bch_collide_batch over device arrays with the static scratch as working
// memory: no hipMalloc, hipFree or allocating library call.  The inputs are
// checked on the device first (shape types, hull ranges, pair indices), so
// no kernel reads past a caller array.
int bch_collide_batch_static(bch_context* c, const bch_bodies* d_bodies, const uint32_t* d_pairs, uint64_t n_pairs,
                             const float* d_seeds, uint8_t* d_out_count, float* d_out_points) {
  if (!c) return BCH_INVALID_ARG;
  if (!d_bodies) return fail(c, BCH_INVALID_ARG, "bodies is NULL");
  const bch_bodies* b = d_bodies;
  uint32_t n = b->n;
  uint32_t nv = b->hull_vertices ? b->n_hull_vertices : 0;
  if (n > BCH_STATIC_MAX_BODIES) return fail(c, BCH_INVALID_ARG, "more bodies than BCH_STATIC_MAX_BODIES");
  if (nv > BCH_STATIC_MAX_HULL_VERTICES)
    return fail(c, BCH_INVALID_ARG, "more hull vertices than BCH_STATIC_MAX_HULL_VERTICES");
  if (n_pairs > BCH_STATIC_MAX_PAIRS) return fail(c, BCH_INVALID_ARG, "more pairs than BCH_STATIC_MAX_PAIRS");
  if (n && (!b->position || !b->orientation || !b->shape_type || !b->shape_params))
    return fail(c, BCH_INVALID_ARG, "body arrays must be non-NULL");
  if (n_pairs && (!d_pairs || !d_out_count || !d_out_points)) return fail(c, BCH_INVALID_ARG, "NULL argument");
  HIP_TRY(c, hipSetDevice(c->device));
  bch_context* a = c->stat;
  a->stream = c->stream;
  a->side = c->side;
  a->fork_ev = c->fork_ev;
  a->join_ev = c->join_ev;
  a->hull_lanes = c->hull_lanes;
  a->epa_grid = c->epa_grid;
  a->np_grid = c->np_grid;
  a->epa_wave_faces = c->epa_wave_faces;  // sort_tile stays the carved layout's
  a->cap_mode = false;
  hipStream_t s = a->stream;
  Counters* cnt = a->counters.as<Counters>();
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, cnt);
  uint64_t span = std::max<uint64_t>(std::max<uint64_t>(n, n_pairs), 1);
  hipLaunchKernelGGL(k_static_check, dim3(blocks(span)), dim3(kBlock), 0, s, n, b->shape_type, b->hull_first,
                     b->hull_count, nv, (uint32_t)n_pairs, d_pairs, cnt);
  HIP_TRY(c, hipMemcpyAsync(&a->host_cnt, cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  if (a->host_cnt.errors & 64u)
    return fail(c, BCH_INVALID_ARG, "shape type, hull vertex range or pair index out of range");
  if (!n_pairs) return BCH_OK;
  if (nv)
    hipLaunchKernelGGL(k_xyz_to_float4, dim3(blocks(nv)), dim3(kBlock), 0, s, nv, b->hull_vertices,
                       a->hull.as<float4>());
  if (n)
    hipLaunchKernelGGL(k_pack_bodies, dim3(blocks(n)), dim3(kBlock), 0, s, n, b->position, b->orientation,
                       b->shape_type, b->shape_params, b->hull_first, b->hull_count, a->hull.as<float4>(),
                       (const uint32_t*)nullptr, a->bodies.as<BodyRec>(), a->btype.as<uint8_t>());
  a->P = n_pairs;
  hipLaunchKernelGGL(k_pairs_to_keys, dim3(blocks(n_pairs)), dim3(kBlock), 0, s, (uint32_t)n_pairs,
                     reinterpret_cast<const uint2*>(d_pairs), a->pair_keys.as<uint64_t>());
  int rc = narrow_split(a);
  if (!rc) rc = narrow_finish(a, d_seeds, false);
  if (rc) return from_batch(c, a, rc);
  if (a->host_cnt.errors & 1u) return fail(c, BCH_HIP_ERROR, "box-box clip polygon overflow");
  hipLaunchKernelGGL(k_batch_out, dim3(blocks(n_pairs)), dim3(kBlock), 0, s, (uint32_t)n_pairs,
                     a->ccount.as<uint32_t>(), a->cslots.as<float4>(), d_out_count, d_out_points);
  HIP_TRY(c, hipGetLastError());
  return BCH_OK;
}

}  // extern "C"
