// Phase unwrapping on the device: the residue map of 2x2 loops and the raster
// integration of wrapped differences with 2 pi step maps, in single precision
// as the reference's goldstein_GPU kernels compute them
// (Codes/sub_code/goldstein_GPU/get_residue_cuda.cu:35-66,
//  unwrapp_phase_goldstein_cuda.cu:60-113).  Arrays are column-major
// (n1 fastest, then n2, then the angle index n3).
#include <hip/hip_runtime.h>
#include <cstdint>

#include "unwrap_kernels.hpp"

namespace cbs {

__device__ __forceinline__ float wrap_mod(float a, float b) { return fmodf(fmodf(a, b) + b, b); }

constexpr float kPiF = 3.141592654f;

// residue[d1, d2, z] over the (n1 - 1) x (n2 - 1) loops of each angle
__global__ void phase_residues(const float* __restrict__ p, float* __restrict__ out, int n1, int n2, int n3) {
  const long long m1 = n1 - 1, m2 = n2 - 1;
  const long long total = m1 * m2 * n3;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
    const long long d12 = i % (m1 * m2);
    const int z = (int)(i / (m1 * m2));
    const int d2 = (int)(d12 / m1), d1 = (int)(d12 % m1);
    const long long b = (long long)z * n1 * n2;
    const float v11 = p[b + (long long)d2 * n1 + d1];
    const float v21 = p[b + (long long)d2 * n1 + d1 + 1];
    const float v12 = p[b + (long long)(d2 + 1) * n1 + d1];
    const float v22 = p[b + (long long)(d2 + 1) * n1 + d1 + 1];
    const float tp = 2.0f * kPiF;
    const float a = wrap_mod(v11 - v12 + kPiF, tp);   // along n2 at row d1
    const float c = wrap_mod(v21 - v22 + kPiF, tp);   // along n2 at row d1 + 1
    const float e = wrap_mod(v11 - v21 + kPiF, tp);   // along n1 at column d2
    const float f = wrap_mod(v12 - v22 + kPiF, tp);   // along n1 at column d2 + 1
    out[i] = roundf((a + f - c - e) / tp);
  }
}

// first column (d1 = 0) of every angle, integrated along n2
__global__ void phase_integrate_column(float* __restrict__ p, const int32_t* __restrict__ step2, int n1, int n2,
                                       int n3) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n3) return;
  const float tp = 2.0f * kPiF;
  const long long b = (long long)z * n1 * n2;
  float last = p[b];
  for (int i = 1; i < n2; ++i) {
    float cur = p[b + (long long)i * n1];
    const float d = wrap_mod(cur - last + kPiF, tp) - kPiF;
    cur = last + d - (float)step2[(long long)z * n2 + i] * tp;
    p[b + (long long)i * n1] = cur;
    last = cur;
  }
}

// every row (fixed d2, angle), integrated along n1 from its already unwrapped first element
__global__ void phase_integrate_rows(float* __restrict__ p, const int32_t* __restrict__ step1, int n1, int n2, int n3) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long long)n2 * n3) return;
  const float tp = 2.0f * kPiF;
  const long long b = r * n1;   // (d2, z) row start: z * n1 * n2 + d2 * n1
  float last = p[b];
  for (int i = 1; i < n1; ++i) {
    float cur = p[b + i];
    const float d = wrap_mod(cur - last + kPiF, tp) - kPiF;
    cur = last + d - (float)step1[b + i] * tp;
    p[b + i] = cur;
    last = cur;
  }
}

}  // namespace cbs

