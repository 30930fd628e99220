// k_misc.hip -- Gaussian position perturbation used between force
// evaluations by the benchmark (a stand-in for the integrator step that sits
// outside the force path, H/src/Integrator.cu:273-350).
//
// The normal deviates are the counter-based gauss3() of (seed, step, atom),
// the same the position refresh applies inside graph replays.
#include "sre_internal.h"

namespace sre {

// pos += sigma * N(0,1) per component, atoms in their original order
__global__ __launch_bounds__(256) void k_perturb(int n, float *pos, int stride,
                                                 float sigma, uint64_t seed,
                                                 uint64_t step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float3 g = gauss3(seed, step, (uint64_t)i);
  float *p = pos + (size_t)i * stride;
  const float x = p[0], y = p[1], z = p[2];
  p[0] = x + sigma * g.x;
  p[1] = y + sigma * g.y;
  p[2] = z + sigma * g.z;
}

void launch_perturb(int n, float *pos, int stride, float sigma, uint64_t seed,
                    uint64_t step, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_perturb, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     pos, stride, sigma, seed, step);
}

}  // namespace sre
