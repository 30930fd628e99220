// k_misc.hip -- Gaussian position perturbation used between force
// evaluations by the benchmark (a stand-in for the integrator step that sits
// outside the force path, H/src/Integrator.cu:273-350), and the per-atom add
// of a grid force correction.
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

// AssignForceCorr::assign (FC/src/AssignForceCorr.cu:116-145): one-image wrap,
// ix = int((x * nx) / Lx) in FP32 (clamped here), then the force of that cell
// is added to the atom's force (FC/src/InteractionEngine_interface.cu:1118).
__global__ __launch_bounds__(256) void k_add_force_grid(
    int n, const float *__restrict__ pos, int stride,
    const uint4 *__restrict__ backup, const uint32_t *__restrict__ inv_perm,
    const RepInfo *__restrict__ reps, const float4 *__restrict__ grid,
    int64_t *fx, int64_t *fy, int64_t *fz, double *fxd, double *fyd,
    double *fzd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const RepInfo &R = reps[backup[i].w >> 8];
  if (R.fc_off < 0) return;
  int c[3];
  for (int d = 0; d < 3; ++d)
    c[d] = rc_bin(pos[(size_t)i * stride + d], R.L[d], R.fc_n[d]);
  const float4 f = grid[R.fc_off + c[2] + R.fc_n[2] * (c[1] + R.fc_n[1] * c[0])];
  const uint32_t s = inv_perm[i];
  if (fx) {
    fx[s] += to_fix((double)f.x);
    fy[s] += to_fix((double)f.y);
    fz[s] += to_fix((double)f.z);
  } else {
    fxd[s] += (double)f.x;
    fyd[s] += (double)f.y;
    fzd[s] += (double)f.z;
  }
}

void launch_add_force_grid(int n, const float *pos, int stride, const uint4 *backup,
                           const uint32_t *inv_perm, const RepInfo *reps,
                           const float4 *grid, int64_t *fx, int64_t *fy, int64_t *fz,
                           double *fxd, double *fyd, double *fzd, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add_force_grid, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     pos, stride, backup, inv_perm, reps, grid, fx, fy, fz, fxd, fyd,
                     fzd);
}

void launch_perturb(int n, float *pos, int stride, float sigma, uint64_t seed,
                    uint64_t step, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_perturb, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     pos, stride, sigma, seed, step);
}

}  // namespace sre
