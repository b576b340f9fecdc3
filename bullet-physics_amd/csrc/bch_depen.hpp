// bch_depen.hpp -- one descent step on the penetration penalty
//   L = sum over broadphase pairs of max(0, -sd)^2
// at the bodies' poses, on the device: the loss terms and the list of
// penetrating pairs (only they carry a gradient, dL/dsd = -2 max(0, -sd)),
// the per-body gradient sum and the pose update
//   position -= lr * dL/dposition
//   orientation = normalize(orientation - lr * dL/dorientation).
// The signed distance and its adjoint are the k_sd_walk / k_sd_epa kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace bch {

// One step's record (32 bytes).
struct DepenRecord {
  float loss;         // sum of max(0, -sd)^2
  float max_pen;      // largest max(0, -sd)
  uint32_t n_pairs;   // broadphase pairs found
  uint32_t n_pen;     // penetrating pairs
  uint32_t flags;     // Counters::errors bits (8 = pair capacity exceeded)
  uint32_t pad[3];
};

constexpr int kDepenBlock = 256;

// Per pair: the loss term, and (for sd < 0) the pair and its output
// gradient appended to the penetrating list.  Per block: the sum of the
// loss terms and the largest penetration, reduced in a fixed order.
__global__ __launch_bounds__(kDepenBlock) void k_depen_terms(uint32_t P, const uint32_t* __restrict__ P_dev,
                                                             const float* __restrict__ sd_out,
                                                             const uint32_t* __restrict__ pairs,
                                                             uint32_t* __restrict__ pen_pairs,
                                                             float* __restrict__ pen_g, uint32_t* __restrict__ n_pen,
                                                             float2* __restrict__ partials) {
  __shared__ float ss[kDepenBlock], sm[kDepenBlock];
  P = dev_count(P, P_dev);
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  float pen = 0.0f;
  if (t < P) {
    float sd = sd_out[10 * (uint64_t)t];
    pen = sd < 0.0f ? -sd : 0.0f;
  }
  bool emit = pen > 0.0f;
  uint64_t mask = __ballot(emit);
  if (mask) {
    uint32_t lane = threadIdx.x & 63;
    uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(n_pen, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (emit) {
      uint64_t k = base + __popcll(mask & ((1ull << lane) - 1ull));
      pen_pairs[2 * k] = pairs[2 * (uint64_t)t];
      pen_pairs[2 * k + 1] = pairs[2 * (uint64_t)t + 1];
      float* g = pen_g + 10 * k;
      g[0] = -2.0f * pen;
      for (int c = 1; c < 10; ++c) g[c] = 0.0f;
    }
  }
  ss[threadIdx.x] = pen * pen;
  sm[threadIdx.x] = pen;
  __syncthreads();
  for (int h = kDepenBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      ss[threadIdx.x] = ss[threadIdx.x] + ss[threadIdx.x + h];
      sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = make_float2(ss[0], sm[0]);
}

// One block: the step's loss and record; hist[*step] when it fits, then
// ++*step.
__global__ __launch_bounds__(kDepenBlock) void k_depen_record(const float2* __restrict__ partials, uint32_t nb,
                                                              const Counters* __restrict__ cnt,
                                                              DepenRecord* __restrict__ hist, uint32_t hist_cap,
                                                              uint32_t* __restrict__ step) {
  __shared__ float ss[kDepenBlock], sm[kDepenBlock];
  float s = 0.0f, m = 0.0f;
  for (uint32_t i = threadIdx.x; i < nb; i += kDepenBlock) {
    float2 p = partials[i];
    s = s + p.x;
    m = fmaxf(m, p.y);
  }
  ss[threadIdx.x] = s;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int h = kDepenBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      ss[threadIdx.x] = ss[threadIdx.x] + ss[threadIdx.x + h];
      sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint32_t k = *step;
    if (k < hist_cap) {
      DepenRecord r;
      r.loss = ss[0];
      r.max_pen = sm[0];
      r.n_pairs = cnt->n_pairs_total;
      r.n_pen = cnt->n_pen;
      r.flags = cnt->errors;
      r.pad[0] = r.pad[1] = r.pad[2] = 0u;
      hist[k] = r;
    }
    *step = k + 1;
  }
}

// Per-body gradient: sum over the penetrating pairs' pose gradients
// [14] = (position A, orientation A, position B, orientation B).
__global__ void k_depen_scatter(const uint32_t* __restrict__ n_pen, const uint32_t* __restrict__ pen_pairs,
                                const float* __restrict__ g14, float* __restrict__ gpos, float* __restrict__ gquat) {
  uint32_t n = *n_pen;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    uint32_t i = pen_pairs[2 * (uint64_t)t], j = pen_pairs[2 * (uint64_t)t + 1];
    const float* g = g14 + 14 * (uint64_t)t;
    for (int k = 0; k < 3; ++k) {
      atomicAdd(&gpos[3 * (uint64_t)i + k], g[k]);
      atomicAdd(&gpos[3 * (uint64_t)j + k], g[7 + k]);
    }
    for (int k = 0; k < 4; ++k) {
      atomicAdd(&gquat[4 * (uint64_t)i + k], g[3 + k]);
      atomicAdd(&gquat[4 * (uint64_t)j + k], g[10 + k]);
    }
  }
}

__global__ void k_depen_update(uint32_t n, float lr, const float* __restrict__ gpos, const float* __restrict__ gquat,
                               float* __restrict__ pos, float* __restrict__ quat) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; ++k) pos[3 * (uint64_t)i + k] = pos[3 * (uint64_t)i + k] - lr * gpos[3 * (uint64_t)i + k];
  float q[4];
  for (int k = 0; k < 4; ++k) q[k] = quat[4 * (uint64_t)i + k] - lr * gquat[4 * (uint64_t)i + k];
  float nrm = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) quat[4 * (uint64_t)i + k] = q[k] / nrm;
}

}  // namespace bch
