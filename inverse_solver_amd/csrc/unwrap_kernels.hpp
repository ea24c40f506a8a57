#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cbs {

// Residues of the 2x2 loops of n3 column-major n1 x n2 phase maps: out is (n1-1) x (n2-1) x n3.
__global__ void phase_residues(const float* __restrict__ p, float* __restrict__ out, int n1, int n2, int n3);
// Raster unwrap in place: first column along n2 (step2: n2 x n3), then every row along n1 (step1: n1 x n2 x n3).
__global__ void phase_integrate_column(float* __restrict__ p, const int32_t* __restrict__ step2, int n1, int n2, int n3);
__global__ void phase_integrate_rows(float* __restrict__ p, const int32_t* __restrict__ step1, int n1, int n2, int n3);

}  // namespace cbs
