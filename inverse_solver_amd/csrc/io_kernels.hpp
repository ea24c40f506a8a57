// 2D input/output projections of the CBS solve.
//
// Input  (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:543-554): the filtered 2D field
//   is zero-padded to the convolution grid, transformed, multiplied by the
//   refocusing kernel for every z of the RI grid, transformed back and cropped
//   to the RI grid -> the incident 3D field u_in.
// Output (:444-511): the emitter u * Vt * dV on the RI grid is padded to the
//   convolution grid and transformed per z; the transmitted / reflected 2D
//   spectra are the z-sums of emitter * kernel (accumulated in FP64); they are
//   transformed back and cropped to the field grid.
// The centred transforms fftshift(fft2(ifftshift(.))) are carried out as
// index maps in the pad / multiply / crop kernels around plain 2D FFTs (whose
// output is in LineFft position order, mapped to centred frequencies by kcx/kcy).
#pragma once
#include "cbs_kernels.hpp"

namespace cbs {

struct IoGeom {
  int fx, fy;          // field grid
  int cx, cy;          // convolution grid
  int rx, ry, rz;      // RI grid (rx = fx, ry = fy)
  int pad_x, pad_y;    // start of the field grid inside the conv grid (padd_field2conv)
  int ri_x, ri_y;      // start of the RI grid inside the conv grid (padd_RI2conv / crop_conv2RI)
};

__global__ void io_pad_in(const float2* __restrict__ src, float2* __restrict__ buf, IoGeom g, int nb);
__global__ void io_refocus(const float2* __restrict__ spec2, const float2* __restrict__ refocus,
                           float2* __restrict__ buf3, const int* __restrict__ kcx, const int* __restrict__ kcy,
                           IoGeom g, int nb, float scale);
__global__ void io_crop_in(const float2* __restrict__ buf3, float2* __restrict__ u_in, IoGeom g, int nb);
__global__ void io_emitter(const float2* __restrict__ field, const float2* __restrict__ u_in,
                           const float2* __restrict__ V, float2* __restrict__ buf3, IoGeom g, Consts c, int ox, int oy,
                           int oz, long long rhs_stride, int nb, float dV);
__global__ void io_zsum(const float2* __restrict__ buf3, const float2* __restrict__ ktrans,
                        const float2* __restrict__ kref, float2* __restrict__ out_t, float2* __restrict__ out_r,
                        const int* __restrict__ kcx, const int* __restrict__ kcy, IoGeom g, int nb);
__global__ void io_crop_out(const float2* __restrict__ buf2, float2* __restrict__ out, IoGeom g, int nb, float scale);
// centred pupil filter of the field grid (field_pupil in solver.hip): out = a [+ b], then the
// field-grid FFTs around the pupil product
__global__ void field_sum(const float2* a, const float2* __restrict__ b, float2* out,   // out may be a
                          long long n);
__global__ void field_pupil_mul(float2* __restrict__ buf, const float* __restrict__ pupil, int nx, int ny, long long nf);

struct IoState {
  bool ready = false;
  bool have_ref = false;
  unsigned long long footprint(int nb, long long cx, long long cy, long long rz) const {
    if (!ready) return 0ull;
    // conv-grid 3D work buffer + 2 conv 2D spectra + u_in
    return (unsigned long long)nb * (cx * cy * rz + 2 * cx * cy) * 8ull;
  }
};

}  // namespace cbs
