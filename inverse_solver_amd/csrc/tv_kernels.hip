// Isotropic 3D TV proximal step (FISTA on the dual), real FP32, on the device.
// Codes/sub_code/TV_FISTA_inner_v2.m:1-123 for a real 3D input (kappamax = 0):
//   y   = clamp(in - lam * div(R))                     (:38, project_non_neg + TV_L)
//   P'  = R + grad(y) / (12 lam);  P' /= max(|P'|, 1)  (:38-42, TV_L_trans, project_TV)
//   R   = P' + (t - 1)/t' (P' - P)                     (:44-47)
//   out = clamp(in - lam * div(P))                     (:51)
// with periodic forward differences grad_k(y) = y - y(+e_k) and
// div(P) = sum_k P_k - P_k(-e_k).
#include "tv_kernels.hpp"

namespace cbs {

#define GRID_STRIDE_TV(i, total) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (total); i += (long long)gridDim.x * blockDim.x)

__device__ __forceinline__ float tv_project(float v, TvArgs a, int x, int y, int z) {
  if (a.non_neg) v = fminf(fmaxf(v, a.vmin), a.vmax);
  if (a.dirichlet && (x == 0 || y == 0 || z == 0 || x == a.nx - 1 || y == a.ny - 1 || z == a.nz - 1)) v = 0.f;
  return v;
}

// y = project(in - lam * div(Q)), Q = R (iterations) or P (final)
__global__ void tv_div_project(const float* __restrict__ in, const float* __restrict__ Q, float* __restrict__ y,
                               TvArgs a) {
  const long long n = (long long)a.nx * a.ny * a.nz;
  GRID_STRIDE_TV(i, n) {
    const int x = (int)(i % a.nx);
    const long long yz = i / a.nx;
    const int yy = (int)(yz % a.ny), z = (int)(yz / a.ny);
    const long long ix = x == 0 ? i + (a.nx - 1) : i - 1;
    const long long iy = yy == 0 ? i + (long long)(a.ny - 1) * a.nx : i - a.nx;
    const long long iz = z == 0 ? i + (long long)(a.nz - 1) * a.nx * a.ny : i - (long long)a.nx * a.ny;
    const float div = Q[i] + Q[n + i] + Q[2 * n + i] - Q[ix] - Q[n + iy] - Q[2 * n + iz];
    y[i] = tv_project(in[i] - a.lam * div, a, x, yy, z);
  }
}

// P' = proj(R + grad(y)/(12 lam)); R = P' + theta (P' - P); P = P'
__global__ void tv_dual_update(const float* __restrict__ y, float* __restrict__ P, float* __restrict__ R, TvArgs a,
                               float theta) {
  const long long n = (long long)a.nx * a.ny * a.nz;
  const float c = 1.0f / (12.0f * a.lam);
  GRID_STRIDE_TV(i, n) {
    const int x = (int)(i % a.nx);
    const long long yz = i / a.nx;
    const int yy = (int)(yz % a.ny), z = (int)(yz / a.ny);
    const long long jx = x == a.nx - 1 ? i - (a.nx - 1) : i + 1;
    const long long jy = yy == a.ny - 1 ? i - (long long)(a.ny - 1) * a.nx : i + a.nx;
    const long long jz = z == a.nz - 1 ? i - (long long)(a.nz - 1) * a.nx * a.ny : i + (long long)a.nx * a.ny;
    const float yi = y[i];
    float p0 = R[i] + c * (yi - y[jx]);
    float p1 = R[n + i] + c * (yi - y[jy]);
    float p2 = R[2 * n + i] + c * (yi - y[jz]);
    const float nrm = sqrtf(fmaxf(p0 * p0 + p1 * p1 + p2 * p2, 1.0f));
    p0 /= nrm;
    p1 /= nrm;
    p2 /= nrm;
    const float o0 = P[i], o1 = P[n + i], o2 = P[2 * n + i];
    P[i] = p0;
    P[n + i] = p1;
    P[2 * n + i] = p2;
    R[i] = p0 + theta * (p0 - o0);
    R[n + i] = p1 + theta * (p1 - o1);
    R[2 * n + i] = p2 + theta * (p2 - o2);
  }
}

__global__ void tv_zero(float* __restrict__ buf, long long n) {
  GRID_STRIDE_TV(i, n) buf[i] = 0.f;
}

}  // namespace cbs
