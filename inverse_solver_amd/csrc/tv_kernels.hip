// Isotropic 3D TV proximal step (FISTA on the dual) on the device, FP32, for a
// real potential (kappamax = 0) or a complex one.
// Codes/sub_code/TV_FISTA_inner_v2.m:1-123 for 3D input:
//   y   = project(in - lam * div(R))                   (:38, project_non_neg + TV_L)
//   P'  = R + grad(y) / (12 lam);  P' /= max(|P'|, 1)  (:38-42, TV_L_trans, project_TV)
//   R   = P' + (t - 1)/t' (P' - P)                     (:44-47)
//   out = project(in - lam * div(P))                   (:51)
// with periodic forward differences grad_k(y) = y - y(+e_k) and
// div(P) = sum_k P_k - P_k(-e_k).  |P'| is the isotropic norm over the three
// components, each |.|^2 of a complex value for complex input (project_TV :79-91).
// project (project_non_neg :57-75): with non_neg, real part clamped to [vmin, vmax];
// for complex input a negative imaginary part is set to 0 and, when vimag_max > 0,
// the imaginary part is capped at vimag_max; dirichlet zeroes the six faces.
#include "tv_kernels.hpp"

namespace cbs {

#define GRID_STRIDE_TV(i, total) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (total); i += (long long)gridDim.x * blockDim.x)

__device__ __forceinline__ float tv_abs2(float v) { return v * v; }
__device__ __forceinline__ float tv_abs2(float2 v) { return v.x * v.x + v.y * v.y; }
__device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 operator*(float s, float2 a) { return make_float2(s * a.x, s * a.y); }
__device__ __forceinline__ float2 operator/(float2 a, float s) { return make_float2(a.x / s, a.y / s); }
__device__ __forceinline__ void tv_zero_val(float& v) { v = 0.f; }
__device__ __forceinline__ void tv_zero_val(float2& v) { v = make_float2(0.f, 0.f); }

__device__ __forceinline__ float tv_bounds(float v, const TvArgs& a) {
  if (a.non_neg) v = fminf(fmaxf(v, a.vmin), a.vmax);
  return v;
}
__device__ __forceinline__ float2 tv_bounds(float2 v, const TvArgs& a) {
  if (a.non_neg) v.x = fminf(fmaxf(v.x, a.vmin), a.vmax);
  if (v.y < 0.f) v.y = 0.f;
  if (a.vimag_max > 0.f && v.y > a.vimag_max) v.y = a.vimag_max;
  return v;
}

template <class T>
__device__ __forceinline__ void tv_div_project_body(const T* __restrict__ in, const T* __restrict__ Q,
                                                    T* __restrict__ y, const TvArgs& a) {
  const long long n = (long long)a.nx * a.ny * a.nz;
  GRID_STRIDE_TV(i, n) {
    const int x = (int)(i % a.nx);
    const long long yz = i / a.nx;
    const int yy = (int)(yz % a.ny), z = (int)(yz / a.ny);
    const long long ix = x == 0 ? i + (a.nx - 1) : i - 1;
    const long long iy = yy == 0 ? i + (long long)(a.ny - 1) * a.nx : i - a.nx;
    const long long iz = z == 0 ? i + (long long)(a.nz - 1) * a.nx * a.ny : i - (long long)a.nx * a.ny;
    const T div = ((Q[i] + Q[n + i]) + Q[2 * n + i]) - Q[ix] - Q[n + iy] - Q[2 * n + iz];
    T v = tv_bounds(in[i] - a.lam * div, a);
    if (a.dirichlet && (x == 0 || yy == 0 || z == 0 || x == a.nx - 1 || yy == a.ny - 1 || z == a.nz - 1))
      tv_zero_val(v);
    y[i] = v;
  }
}

template <class T>
__device__ __forceinline__ void tv_dual_update_body(const T* __restrict__ y, T* __restrict__ P, T* __restrict__ R,
                                                    const TvArgs& a, float theta) {
  const long long n = (long long)a.nx * a.ny * a.nz;
  const float c = 1.0f / (12.0f * a.lam);
  GRID_STRIDE_TV(i, n) {
    const int x = (int)(i % a.nx);
    const long long yz = i / a.nx;
    const int yy = (int)(yz % a.ny), z = (int)(yz / a.ny);
    const long long jx = x == a.nx - 1 ? i - (a.nx - 1) : i + 1;
    const long long jy = yy == a.ny - 1 ? i - (long long)(a.ny - 1) * a.nx : i + a.nx;
    const long long jz = z == a.nz - 1 ? i - (long long)(a.nz - 1) * a.nx * a.ny : i + (long long)a.nx * a.ny;
    const T yi = y[i];
    T p0 = R[i] + c * (yi - y[jx]);
    T p1 = R[n + i] + c * (yi - y[jy]);
    T p2 = R[2 * n + i] + c * (yi - y[jz]);
    const float nrm = sqrtf(fmaxf(tv_abs2(p0) + tv_abs2(p1) + tv_abs2(p2), 1.0f));
    p0 = p0 / nrm;
    p1 = p1 / nrm;
    p2 = p2 / nrm;
    const T o0 = P[i], o1 = P[n + i], o2 = P[2 * n + i];
    P[i] = p0;
    P[n + i] = p1;
    P[2 * n + i] = p2;
    R[i] = p0 + theta * (p0 - o0);
    R[n + i] = p1 + theta * (p1 - o1);
    R[2 * n + i] = p2 + theta * (p2 - o2);
  }
}

__global__ void tv_div_project(const float* __restrict__ in, const float* __restrict__ Q, float* __restrict__ y,
                               TvArgs a) {
  tv_div_project_body(in, Q, y, a);
}
__global__ void tv_div_project_c(const float2* __restrict__ in, const float2* __restrict__ Q, float2* __restrict__ y,
                                 TvArgs a) {
  tv_div_project_body(in, Q, y, a);
}
__global__ void tv_dual_update(const float* __restrict__ y, float* __restrict__ P, float* __restrict__ R, TvArgs a,
                               float theta) {
  tv_dual_update_body(y, P, R, a, theta);
}
__global__ void tv_dual_update_c(const float2* __restrict__ y, float2* __restrict__ P, float2* __restrict__ R,
                                 TvArgs a, float theta) {
  tv_dual_update_body(y, P, R, a, theta);
}

__global__ void tv_zero(float* __restrict__ buf, long long n) {
  GRID_STRIDE_TV(i, n) buf[i] = 0.f;
}

// grad[i] = -(i / n_rhs) * sum_r conj(u_r[i]) * conj(w_r[flip_z(i)]), FP64 sums.
// Codes/BACKWARD_SOLVER_MULTI.m:50-53 with w = flip(conj(w_adjoint), 3).
__global__ void field_gradient(const float2* __restrict__ u, const float2* __restrict__ w, int nx, int ny, int nz,
                               int n_rhs, double2* __restrict__ grad) {
  const long long nxy = (long long)nx * ny;
  const long long n = nxy * nz;
  GRID_STRIDE_TV(i, n) {
    const long long z = i / nxy;
    const long long j = i + (nz - 1 - 2 * z) * nxy;   // same (x, y), z -> nz - 1 - z
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < n_rhs; ++r) {
      const float2 a = u[(long long)r * n + i];
      const float2 b = w[(long long)r * n + j];
      // conj(a) * conj(b) = (a.x b.x - a.y b.y) - i (a.x b.y + a.y b.x)
      sr += (double)a.x * b.x - (double)a.y * b.y;
      si -= (double)a.x * b.y + (double)a.y * b.x;
    }
    const double inv = 1.0 / n_rhs;
    grad[i] = make_double2(si * inv, -sr * inv);   // -i (sr + i si) / n_rhs
  }
}

// FISTA of BACKWARD_SOLVER_MULTI (Codes/BACKWARD_SOLVER_MULTI.m:152-160) on the device,
// the iterates complex128 and x fastest.  Proximal input u - inv_alpha grad, rounded to
// complex64 (cplx) or its real part to float32, as the host loop forms it before the TV
// prox of that type.
__global__ void fista_prox_input(const double2* __restrict__ u, const double2* __restrict__ grad, double inv_alpha,
                                 long long n, int cplx, void* __restrict__ out) {
  GRID_STRIDE_TV(i, n) {
    const double2 g = grad[i], v = u[i];
    const double re = v.x - inv_alpha * g.x, im = v.y - inv_alpha * g.y;
    if (cplx) static_cast<float2*>(out)[i] = make_float2((float)re, (float)im);
    else static_cast<float*>(out)[i] = (float)re;
  }
}

// Principal square root of a + ib (a branch cut on the negative real axis, the sign of
// b kept for a <= 0): t = sqrt((|z| + |a|) / 2), then (t, b / 2t) for a >= 0 and
// (|b| / 2t, copysign(t, b)) for a < 0, so that neither part suffers cancellation.
__device__ __forceinline__ double2 csqrt_principal(double a, double b) {
  if (a == 0.0 && b == 0.0) return make_double2(0.0, b);
  const double t = sqrt(0.5 * (hypot(a, b) + fabs(a)));
  if (a >= 0.0) return make_double2(t, b / (2.0 * t));
  return make_double2(fabs(b) / (2.0 * t), copysign(t, b));
}

// After the proximal step s (complex64 when cplx, else float32, widened to FP64):
// u = s + c (s - x), x = s, ri = n_m sqrt(u + 1).
__global__ void fista_momentum(const void* __restrict__ s, int cplx, double2* __restrict__ x, double2* __restrict__ u,
                               double2* __restrict__ ri, long long n, double c, double n_m) {
  GRID_STRIDE_TV(i, n) {
    double2 sv;
    if (cplx) {
      const float2 f = static_cast<const float2*>(s)[i];
      sv = make_double2((double)f.x, (double)f.y);
    } else {
      sv = make_double2((double)static_cast<const float*>(s)[i], 0.0);
    }
    const double2 xv = x[i];
    const double2 un = make_double2(sv.x + c * (sv.x - xv.x), sv.y + c * (sv.y - xv.y));
    u[i] = un;
    x[i] = sv;
    const double2 r = csqrt_principal(un.x + 1.0, un.y);
    ri[i] = make_double2(n_m * r.x, n_m * r.y);
  }
}

}  // namespace cbs
