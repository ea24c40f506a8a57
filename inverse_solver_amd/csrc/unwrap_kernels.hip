// Phase unwrapping on the device: the residue map of 2x2 loops and the raster
// integration of wrapped differences with 2 pi step maps, in single precision
// as the reference's goldstein_GPU kernels compute them
// (Codes/sub_code/goldstein_GPU/get_residue_cuda.cu:35-66,
//  unwrapp_phase_goldstein_cuda.cu:60-113).  Arrays are column-major
// (n1 fastest, then n2, then the angle index n3).
#include <hip/hip_runtime.h>
#include <cstdint>

#include "goldstein.hpp"
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

// The reference's update last + d - step * 2 * pi is one fused multiply-add under nvcc's
// default contraction (--fmad=true); it is written as fmaf here so that the rounding does
// not depend on the compiler's choice.

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
    cur = fmaf(-(float)step2[(long long)z * n2 + i], tp, last + d);
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
    cur = fmaf(-(float)step1[b + i], tp, last + d);
    p[b + i] = cur;
    last = cur;
  }
}

// remove_dipoles_cuda.cu:50-106, one launch per direction (0..3 = d1 - 1, d2 - 1, d1 + 1,
// d2 + 1).  pos: 1-based flat indices of the positive residues, set to -1 once cancelled.
// Within one direction every negative neighbour belongs to one positive residue, so no
// element is touched by two threads.  A dipole along d2 writes +-1 (directions 1 / 3)
// into step1 at (d1 + 1, min(d2, d2') + 1).
__global__ void phase_dipoles(float* __restrict__ res, int32_t* __restrict__ pos, int32_t* __restrict__ step1,
                              int npos, int m1, int m2, int n3, int direction) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npos || pos[k] <= 0) return;
  const long long rid = (long long)pos[k] - 1;
  const long long plane = (long long)m1 * m2;
  const int z = (int)(rid / plane);
  const int d12 = (int)(rid % plane);
  const int d2 = d12 / m1, d1 = d12 % m1;
  if (z >= n3) return;
  const int p1 = d1 - (direction == 0 ? 1 : 0) + (direction == 2 ? 1 : 0);
  const int p2 = d2 - (direction == 1 ? 1 : 0) + (direction == 3 ? 1 : 0);
  if (p1 < 0 || p2 < 0 || p1 >= m1 || p2 >= m2) return;
  const long long pid = (long long)z * plane + (long long)p2 * m1 + p1;
  float pv = res[pid];
  if (!(pv < 0.0f)) return;
  float v = res[rid];
  const float rem = fminf(v, -pv);
  v = roundf(v - rem);
  pv = roundf(pv + rem);
  res[rid] = v;
  res[pid] = pv;
  if (v == 0.0f) pos[k] = -1;
  if (direction == 1 || direction == 3) {
    const long long n1 = m1 + 1, n2 = m2 + 1;
    const long long c1 = p1 + 1, c2 = min(d2, p2) + 1;
    step1[(long long)z * n1 * n2 + c2 * n1 + c1] = (int32_t)rintf(direction == 1 ? rem : -rem);
  }
}

// get_shortest_sort_cuda.cu:64-87: the gs::kNearest nearest residues of every listed
// residue, out[(i * kNearest + k) * 2 + {0, 1}] = {id, distance}.
__global__ void phase_nearest(const int32_t* __restrict__ rows, int n, const int32_t* __restrict__ table,
                              int table_size, const int32_t* __restrict__ zstart, const int32_t* __restrict__ square,
                              int s1, int s2, int32_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cur[2 * gs::kNearest];
  gs::nearest<gs::kNearest>(nullptr, rows, n, table, table_size, zstart, square, s1, s2, i, cur);
  for (int k = 0; k < 2 * gs::kNearest; ++k) out[(long long)i * 2 * gs::kNearest + k] = cur[k];
}

// raster_branches_cuda.cu:65-135: add each branch's jump to the step maps.  A branch to
// the border (p1 = -1) ends at the nearer border of the (n1, n2) map; on the d1 = -1
// side it adds to step2 of column p2 d2 + 1.  Otherwise each row path between the two
// ends is crossed at d1 = p2 d1 + trunc(i / steps * (p1 d1 - p2 d1)) + 1, taken as a flat
// index of step1 (so a crossing one past the last row wraps to the next column, as there);
// a flat index outside the map is dropped.
__global__ void phase_raster(const int32_t* __restrict__ br, int nb, int32_t* __restrict__ step1,
                             int32_t* __restrict__ step2, int n1, int n2, int n3) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int32_t* r = br + 6 * (long long)b;
  int p1x = r[0], p1y = r[1], p2x = r[2], p2y = r[3];
  const int z = r[4];
  int jump = r[5];
  if (p1x == -1) {
    const bool xs = p2x < n1 - p2x;
    const bool ys = p2y < n2 - p2y;
    const int mdx = xs ? p2x : n1 - p2x;
    const int mdy = ys ? p2y : n2 - p2y;
    p1x = mdx > mdy ? p2x : (xs ? -1 : n1 - 1);
    p1y = mdx < mdy ? p2y : (ys ? -1 : n2 - 1);
    if (p1x == -1) {
      const long long k = (long long)n2 * z + p2y + 1;
      if (k >= 0 && k < (long long)n2 * n3) atomicAdd(&step2[k], jump);
      return;
    }
  }
  if (p1y == p2y) return;
  if (p2y > p1y) {
    int t = p1y;
    p1y = p2y;
    p2y = t;
    t = p1x;
    p1x = p2x;
    p2x = t;
    jump = -jump;
  }
  const int steps = p1y - p2y;
  const long long total = (long long)n1 * n2 * n3;
  const float dxf = (float)(p1x - p2x);
  for (int i = 0; i < steps; ++i) {
    const int cy = p2y + i + 1;
    const float frac = (float)((double)i / (double)steps);   // the single-precision quotient
    const int cx = p2x + (int)(frac * dxf) + 1;
    const long long k = (long long)n1 * n2 * z + (long long)n1 * cy + cx;
    if (k >= 0 && k < total) atomicAdd(&step1[k], jump);
  }
}

}  // namespace cbs

