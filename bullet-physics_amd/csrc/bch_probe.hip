// bch_probe.hip -- measurement probes for the roofline figures (not part of
// the collision path).  k_valu_probe issues independent f32 FMA chains on
// every lane at full occupancy, so its VALU wave-instruction count (from the
// SQ counters) over its duration is the issue rate the hardware actually
// sustains: the ceiling the collision kernels' VALU rates are set against.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// Eight independent FMA chains per lane; each inner round is 8 v_fma_f32 with
// no dependency between the chains, so issue, not latency, bounds it.
__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ sink, int rounds) {
  float a0 = threadIdx.x * 1e-3f, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
  float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
  const float m = 0.999f, c = 1e-3f;
  for (int r = 0; r < rounds; ++r) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      a0 = fmaf(a0, m, c);
      a1 = fmaf(a1, m, c);
      a2 = fmaf(a2, m, c);
      a3 = fmaf(a3, m, c);
      a4 = fmaf(a4, m, c);
      a5 = fmaf(a5, m, c);
      a6 = fmaf(a6, m, c);
      a7 = fmaf(a7, m, c);
    }
  }
  float s = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
  if (s == 1234.5f) sink[blockIdx.x * blockDim.x + threadIdx.x] = s;  // keeps the chains live
}

}  // namespace

extern "C" {

// Launch `blocks` blocks of 256 threads, `rounds` rounds of 128 FMAs per
// lane, `reps` times on the null stream; *ms = the mean device time of one
// launch.  `sink` is a device buffer of blocks * 256 floats.
int bch_probe_valu(int blocks, int rounds, int reps, float* sink, float* ms) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 3;
  hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, 0, sink, rounds);
  (void)hipEventRecord(a, 0);
  for (int k = 0; k < reps; ++k) hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, 0, sink, rounds);
  (void)hipEventRecord(b, 0);
  int rc = hipEventSynchronize(b) == hipSuccess ? 0 : 3;
  float t = 0.0f;
  (void)hipEventElapsedTime(&t, a, b);
  *ms = t / (float)(reps > 0 ? reps : 1);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // extern "C"
