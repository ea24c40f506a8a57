#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace cbs {

// Residues of the 2x2 loops of n3 column-major n1 x n2 phase maps: out is (n1-1) x (n2-1) x n3.
__global__ void phase_residues(const float* __restrict__ p, float* __restrict__ out, int n1, int n2, int n3);
// Raster unwrap in place: first column along n2 (step2: n2 x n3), then every row along n1 (step1: n1 x n2 x n3).
__global__ void phase_integrate_column(float* __restrict__ p, const int32_t* __restrict__ step2, int n1, int n2, int n3);
__global__ void phase_integrate_rows(float* __restrict__ p, const int32_t* __restrict__ step1, int n1, int n2, int n3);

// Goldstein branch cuts (see goldstein.hpp): dipole removal in one direction, the nearest
// lists of the listed residues, and rastering of the branches into the step maps.
__global__ void phase_dipoles(float* __restrict__ res, int32_t* __restrict__ pos, int32_t* __restrict__ step1,
                              int npos, int m1, int m2, int n3, int direction);
__global__ void phase_nearest(const int32_t* __restrict__ rows, int n, const int32_t* __restrict__ table,
                              int table_size, const int32_t* __restrict__ zstart, const int32_t* __restrict__ square,
                              int s1, int s2, int32_t* __restrict__ out);
__global__ void phase_raster(const int32_t* __restrict__ br, int nb, int32_t* __restrict__ step1,
                             int32_t* __restrict__ step2, int n1, int n2, int n3);

}  // namespace cbs
