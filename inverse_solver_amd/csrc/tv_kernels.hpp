#pragma once
#include <hip/hip_runtime.h>

namespace cbs {

struct TvArgs {
  int nx, ny, nz;
  float lam, vmin, vmax, vimag_max;
  int non_neg, dirichlet;
};

// Real potential (kappamax = 0): one float per voxel; the dual P, R hold 3 * n floats.
__global__ void tv_div_project(const float* __restrict__ in, const float* __restrict__ Q, float* __restrict__ y,
                               TvArgs a);
__global__ void tv_dual_update(const float* __restrict__ y, float* __restrict__ P, float* __restrict__ R, TvArgs a,
                               float theta);
// Complex potential: complex64 per voxel; the dual holds 3 * n complex values.
__global__ void tv_div_project_c(const float2* __restrict__ in, const float2* __restrict__ Q, float2* __restrict__ y,
                                 TvArgs a);
__global__ void tv_dual_update_c(const float2* __restrict__ y, float2* __restrict__ P, float2* __restrict__ R,
                                 TvArgs a, float theta);
__global__ void tv_zero(float* __restrict__ buf, long long n);

// Misfit gradient from the forward and adjoint 3D fields (n_rhs blocks of nx*ny*nz complex64).
__global__ void field_gradient(const float2* __restrict__ u, const float2* __restrict__ w, int nx, int ny, int nz,
                               int n_rhs, double2* __restrict__ grad);

// FISTA iterates on the device (complex128, x fastest): the proximal input and the
// momentum / RI update of BACKWARD_SOLVER_MULTI.m:152-160.
__global__ void fista_prox_input(const double2* __restrict__ u, const double2* __restrict__ grad, double inv_alpha,
                                 long long n, int cplx, void* __restrict__ out);
__global__ void fista_momentum(const void* __restrict__ s, int cplx, double2* __restrict__ x, double2* __restrict__ u,
                               double2* __restrict__ ri, long long n, double c, double n_m);

}  // namespace cbs
