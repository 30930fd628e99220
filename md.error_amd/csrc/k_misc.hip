// k_misc.hip -- Gaussian position perturbation used between force
// evaluations by the benchmark (a stand-in for the integrator step that sits
// outside the force path, H/src/Integrator.cu:273-350).
//
// Counter-based: the normal deviates of (seed, step, atom) come from
// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014) followed by the Box-Muller transform, so the
// perturbation is independent of launch shape.
#include "sre_internal.h"

namespace sre {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline float u01(uint64_t x) {
  // (0, 1]: never zero, so log() is finite
  return ((float)(x >> 40) + 1.0f) * (1.0f / 16777216.0f);
}

// pos = (anchor ? anchor : pos) + sigma * N(0,1) per component; the normal
// deviates come from Box-Muller on counter-based (seed, step, atom) hashes.
__global__ __launch_bounds__(256) void k_perturb(int n, float *pos,
                                                 const float *anchor, int stride,
                                                 float sigma, uint64_t seed,
                                                 uint64_t step,
                                                 const uint64_t *step_ctr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (step_ctr) step = *step_ctr;
  const uint64_t base = splitmix64(seed ^ splitmix64(step * 0x632BE59BD9B4E019ull ^ (uint64_t)i));
  const float u1 = u01(splitmix64(base + 1)), u2 = u01(splitmix64(base + 2));
  const float u3 = u01(splitmix64(base + 3)), u4 = u01(splitmix64(base + 4));
  const float r1 = sqrtf(-2.f * logf(u1)), r2 = sqrtf(-2.f * logf(u3));
  const float t1 = 6.2831853f * u2, t2 = 6.2831853f * u4;
  float *p = pos + (size_t)i * stride;
  const float *o = anchor ? anchor + (size_t)i * stride : p;
  const float x = o[0], y = o[1], z = o[2];
  p[0] = x + sigma * r1 * cosf(t1);
  p[1] = y + sigma * r1 * sinf(t1);
  p[2] = z + sigma * r2 * cosf(t2);
}

void launch_perturb(int n, float *pos, const float *anchor, int stride,
                    float sigma, uint64_t seed, uint64_t step,
                    const uint64_t *step_ctr, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_perturb, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     pos, anchor, stride, sigma, seed, step, step_ctr);
}

}  // namespace sre
