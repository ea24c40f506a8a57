#pragma once
#include <hip/hip_runtime.h>

namespace cbs {

struct TvArgs {
  int nx, ny, nz;
  float lam, vmin, vmax;
  int non_neg, dirichlet;
};

__global__ void tv_div_project(const float* __restrict__ in, const float* __restrict__ Q, float* __restrict__ y,
                               TvArgs a);
__global__ void tv_dual_update(const float* __restrict__ y, float* __restrict__ P, float* __restrict__ R, TvArgs a,
                               float theta);
__global__ void tv_zero(float* __restrict__ buf, long long n);

}  // namespace cbs
