// bch_sort.hpp -- exclusive scan and LSD radix sort, with no memsets: every
// buffer a launch reads was fully written by an earlier launch of the same
// sequence, so the sequence can be captured in a HIP graph and replayed.
//
// Scan (u32, n items -> n + 1 offsets, the last one the total):
//   k_scan_reduce   per-block sums of 2048-item tiles
//   k_scan_apply    per-tile exclusive scan plus the tile's offset (the sum
//                   of the earlier tiles' sums, which every block adds up
//                   itself); the last tile writes the total at out[n]
//   k_scan_top      for more than kScanSelfTiles tiles only: one block scans
//                   the tile sums in place first (and writes the total), and
//                   k_scan_apply reads its offset from there
// An optional device count n_dev caps n (a count an earlier kernel of the
// same sequence wrote), so a fixed launch shape serves any count up to n.
// Radix sort (u64 keys, optional u32 values, 8-bit digits, stable):
//   k_rs_hist       per-tile digit histograms, digit-major [256][tiles]
//   k_rs_row_scan   one block per digit scans its row of tile counts (the
//                   tile's offset among the digit's items) and writes the
//                   digit's total
//   k_rs_scatter    adds the digit's start (the exclusive scan of the 256
//                   totals, done by every block) to its row offsets; then
//                   per tile, in item order: each wavefront ranks its items
//                   among equal digits with eight ballots, wavefront counts
//                   meet in LDS, a running per-digit base keeps the order
//                   across the tile's rounds (stable, as LSD needs)
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bch {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048
constexpr int kRsThreads = 256;
constexpr int kRsRounds = 16;
constexpr int kRsTile = kRsThreads * kRsRounds;  // 4096

// Exclusive scan of one value per thread over a block of kScanThreads
// (= 4 wavefronts); returns the block total in `total`.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / 64; ++k) {
    uint32_t s = wsum[k];
    base += k < w ? s : 0u;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + x - v;
}

__device__ __forceinline__ uint32_t capped_count(uint32_t n, const uint32_t* n_dev) {
  if (!n_dev) return n;
  uint32_t m = *n_dev;
  return m < n ? m : n;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                              const uint32_t* __restrict__ n_dev,
                                                              uint32_t* __restrict__ partials) {
  n = capped_count(n, n_dev);
  uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) s += in[base + k];
  uint32_t tot;
  (void)block_exclusive_scan(s, tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Exclusive scan in place of nb partials by one block; the sum goes to
// out[n] (n capped by n_dev) and, when given, to total_copy.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(uint32_t* __restrict__ partials, uint32_t nb,
                                                           uint32_t* __restrict__ out, uint32_t n,
                                                           const uint32_t* __restrict__ n_dev,
                                                           uint32_t* __restrict__ total_copy) {
  uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
  uint32_t b0 = threadIdx.x * per;
  uint32_t s = 0;
  for (uint32_t k = 0; k < per; ++k)
    if (b0 + k < nb) s += partials[b0 + k];
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, tot);
  for (uint32_t k = 0; k < per; ++k) {
    if (b0 + k < nb) {
      uint32_t v = partials[b0 + k];
      partials[b0 + k] = run;
      run += v;
    }
  }
  if (threadIdx.x == 0) {
    out[capped_count(n, n_dev)] = tot;
    if (total_copy) total_copy[0] = tot;
  }
}

// Launched with the same grid as k_scan_reduce (nb blocks).  With
// `prefixed` the partials already hold the tiles' offsets (k_scan_top ran).
constexpr uint32_t kScanSelfTiles = 4096;
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n,
                                                             const uint32_t* __restrict__ n_dev,
                                                             const uint32_t* __restrict__ partials,
                                                             uint32_t* __restrict__ out,
                                                             uint32_t* __restrict__ total_copy, int prefixed) {
  n = capped_count(n, n_dev);
  // this tile's offset: the sum of the earlier tiles' sums
  uint32_t off = 0;
  if (prefixed) {
    off = partials[blockIdx.x];
  } else {
    uint32_t before = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += kScanThreads) before += partials[b];
    (void)block_exclusive_scan(before, off);
  }
  uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, tot) + off;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (!prefixed && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    out[n] = off + tot;
    if (total_copy) total_copy[0] = off + tot;
  }
}

// With pack = b > 0 the key i << 32 | j (i, j < 2^b) is read as the 2b-bit
// number i << b | j, so a pair sort needs ceil(2b / 8) digits, not
// 2 ceil(b / 8).  With pack = b | kPackDelta the key is read as
// i << b | (j - i), which orders the keys the same way when every j - i is
// below 2^b (pairs inside worlds of at most 2^b bodies).
constexpr int kPackDelta = 0x100;
template <class K>
__device__ __forceinline__ uint32_t rs_digit(K key, int shift, int pack) {
  uint64_t k = (uint64_t)key;
  if (pack & kPackDelta) {
    const int b = pack & 0xff;
    k = ((k >> 32) << b) | (((k & 0xffffffffull) - (k >> 32)) & ((1ull << b) - 1ull));
  } else if (pack) {
    k = (k & ((1ull << pack) - 1ull)) | ((k >> 32) << pack);
  }
  return (uint32_t)(k >> shift) & 0xffu;
}

// ROUNDS rounds of kRsThreads items per tile (the tile size is a tuning
// knob: 8 or 16 rounds, 2048 or 4096 items).
template <class K, int ROUNDS = kRsRounds>
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const K* __restrict__ keys, uint32_t n,
                                                        const uint32_t* __restrict__ n_dev, int shift, int pack,
                                                        uint32_t tiles,
                                                        uint32_t* __restrict__ hist) {
  n = capped_count(n, n_dev);
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * (kRsThreads * ROUNDS);
  for (int r = 0; r < ROUNDS; ++r) {
    uint64_t i = base + (uint64_t)r * kRsThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[rs_digit(keys[i], shift, pack)], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Row d of counts ([256][tiles]) scanned into offs (exclusive, per digit:
// how many items of digit d the earlier tiles hold); rowtot[d] = the row's
// sum.  One block per digit.
__global__ __launch_bounds__(kScanThreads) void k_rs_row_scan(const uint32_t* __restrict__ counts, uint32_t tiles,
                                                              uint32_t* __restrict__ offs,
                                                              uint32_t* __restrict__ rowtot) {
  const uint64_t row = (uint64_t)blockIdx.x * tiles;
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < tiles; b0 += kScanTile) {
    const uint32_t base = b0 + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < tiles ? counts[row + base + k] : 0u;
      s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, tot) + carry;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < tiles) offs[row + base + k] = run;
      run += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) rowtot[blockIdx.x] = carry;
}

// counts = the per-tile digit counts k_rs_hist wrote ([256][tiles]), offs and
// rowtot = what k_rs_row_scan made of them: a tile's first item of digit d
// goes to (rowtot[0] + ... + rowtot[d - 1]) + offs[d][tile].
// The tile's items are first placed in LDS in digit order (stable), then
// written out from there: consecutive LDS slots of one digit go to
// consecutive global positions, so the stores come in runs (a wavefront's
// direct stores hit up to 64 separate lines otherwise).
template <class K, bool VALS, int ROUNDS = kRsRounds>
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                           uint32_t n, const uint32_t* __restrict__ n_dev, int shift,
                                                           int pack, uint32_t tiles, const uint32_t* __restrict__ counts,
                                                           const uint32_t* __restrict__ offs,
                                                           const uint32_t* __restrict__ rowtot, K* __restrict__ kout,
                                                           uint32_t* __restrict__ vout) {
  n = capped_count(n, n_dev);
  __shared__ K skey[kRsThreads * ROUNDS];
  __shared__ uint32_t sval[VALS ? kRsThreads * ROUNDS : 1];
  __shared__ uint32_t running[256], lstart[256], gstart[256];
  __shared__ uint32_t wcount[kRsThreads / 64][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t ti = (uint64_t)threadIdx.x * tiles + blockIdx.x;
  uint32_t all;
  const uint32_t dstart = block_exclusive_scan(rowtot[threadIdx.x], all);
  uint32_t total;
  const uint32_t ls = block_exclusive_scan(counts[ti], total);
  lstart[threadIdx.x] = ls;
  running[threadIdx.x] = ls;
  gstart[threadIdx.x] = dstart + offs[ti];
#pragma unroll
  for (int k = 0; k < kRsThreads / 64; ++k) wcount[k][threadIdx.x] = 0u;
  __syncthreads();
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t base = (uint64_t)blockIdx.x * (kRsThreads * ROUNDS);
  for (int r = 0; r < ROUNDS; ++r) {
    uint64_t i = base + (uint64_t)r * kRsThreads + threadIdx.x;
    bool valid = i < n;
    K key = valid ? kin[i] : (K)0;
    uint32_t val = 0u;
    if (VALS && valid) val = vin ? vin[i] : (uint32_t)i;  // no vin: the item's index
    uint32_t d = rs_digit(key, shift, pack);
    // lanes of this wavefront holding the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    uint32_t cnt = (uint32_t)__popcll(peers);
    if (valid && rank + 1 == cnt) wcount[w][d] = cnt;
    __syncthreads();
    if (valid) {
      uint32_t pos = running[d] + rank;
      for (int k = 0; k < kRsThreads / 64; ++k)
        if (k < w) pos += wcount[k][d];
      skey[pos] = key;
      if (VALS) sval[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0u;
#pragma unroll
    for (int k = 0; k < kRsThreads / 64; ++k) {
      add += wcount[k][threadIdx.x];
      wcount[k][threadIdx.x] = 0u;
    }
    running[threadIdx.x] += add;
    __syncthreads();
  }
  // the tile's `total` items, in digit order
  for (uint32_t k = threadIdx.x; k < total; k += kRsThreads) {
    K key = skey[k];
    uint32_t d = rs_digit(key, shift, pack);
    uint32_t pos = gstart[d] + (k - lstart[d]);
    kout[pos] = key;
    if (VALS) vout[pos] = sval[k];
  }
}

// ------------------------------------------------ per-world Morton sort --
// The Morton keys of a batch of worlds, where world w's bodies are the index
// range [off[w], off[w + 1]) of at most kSegMax bodies and a body outside
// the tree carries the key ~0: the same order as the stable radix sort of
// all keys over bits [drop, 31 + world bits), i.e. the tree's bodies world
// by world in (code >> drop) order, ties in body order, then every excluded
// body in body order, with one block per world and no pass over the whole
// array per digit.  k_seg_count counts each world's tree bodies; their
// exclusive scan gives each world's first output slot.
constexpr uint32_t kSegMax = 2048;
__global__ __launch_bounds__(256) void k_seg_count(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ off,
                                                   uint32_t* __restrict__ count) {
  const uint32_t w = blockIdx.x, o0 = off[w], o1 = off[w + 1];
  uint32_t c = 0;
  for (uint32_t i = o0 + threadIdx.x; i < o1; i += 256) c += keys[i] != ~0ull ? 1u : 0u;
  uint32_t tot;
  (void)block_exclusive_scan(c, tot);
  if (threadIdx.x == 0) count[w] = tot;
}

// One block per world: the world's keys as (sort key << 11 | local index),
// sort key = code bits [drop, 30) below a top bit set for an excluded body,
// LSD-sorted in LDS by 8-bit digits (ballot ranking, as k_rs_scatter), then
// written to their global slots: tree body k of world w to first[w] + k,
// excluded body k to n_tree + (off[w] - first[w]) + k, where first holds
// the scan of k_seg_count's counts and first[n_worlds] = n_tree.  wend[s]
// of a tree slot s = the first slot of a later world (first[w + 1]).
__global__ __launch_bounds__(256) void k_seg_sort(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ off,
                                                  const uint32_t* __restrict__ first, uint32_t n_worlds, int drop,
                                                  uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                  uint32_t* __restrict__ wend) {
  __shared__ uint32_t buf[2][kSegMax];
  __shared__ uint32_t running[256];
  __shared__ uint32_t wcount[4][256];
  const uint32_t w = blockIdx.x, o0 = off[w], m = off[w + 1] - o0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int cb = 30 - drop;  // code bits kept
  for (uint32_t k = threadIdx.x; k < m; k += 256) {
    const uint64_t key = keys[o0 + k];
    const uint32_t sk = key == ~0ull ? (1u << cb) : (uint32_t)((key >> drop) & ((1ull << cb) - 1ull));
    buf[0][k] = (sk << 11) | k;
  }
  int cur = 0;
  for (int sh = 11; sh < 11 + cb + 1; sh += 8) {
    running[threadIdx.x] = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) wcount[k][threadIdx.x] = 0u;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < m; k += 256) atomicAdd(&running[(buf[cur][k] >> sh) & 0xffu], 1u);
    __syncthreads();
    uint32_t all;
    const uint32_t base = block_exclusive_scan(running[threadIdx.x], all);
    running[threadIdx.x] = base;
    __syncthreads();
    for (uint32_t r0 = 0; r0 < m; r0 += 256) {
      const uint32_t k = r0 + threadIdx.x;
      const bool valid = k < m;
      const uint32_t item = valid ? buf[cur][k] : 0u;
      const uint32_t d = (item >> sh) & 0xffu;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint64_t mb = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? mb : ~mb;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt), cnt = (uint32_t)__popcll(peers);
      if (valid && rank + 1 == cnt) wcount[wv][d] = cnt;
      __syncthreads();
      if (valid) {
        uint32_t pos = running[d] + rank;
        for (int k2 = 0; k2 < 4; ++k2)
          if (k2 < wv) pos += wcount[k2][d];
        buf[cur ^ 1][pos] = item;
      }
      __syncthreads();
      uint32_t add = 0u;
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) {
        add += wcount[k2][threadIdx.x];
        wcount[k2][threadIdx.x] = 0u;
      }
      running[threadIdx.x] += add;
      __syncthreads();
    }
    cur ^= 1;
  }
  const uint32_t f = first[w], nt = first[n_worlds], rw = first[w + 1] - f;
  for (uint32_t k = threadIdx.x; k < m; k += 256) {
    const uint32_t body = o0 + (buf[cur][k] & 0x7ffu);
    const uint32_t dest = k < rw ? f + k : nt + (o0 - f) + (k - rw);
    kout[dest] = keys[body];
    vout[dest] = body;
    if (k < rw) wend[dest] = f + rw;
  }
}

}  // namespace bch
