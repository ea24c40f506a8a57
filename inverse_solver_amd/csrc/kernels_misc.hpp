#pragma once
#include "cbs_kernels.hpp"

namespace cbs {

struct RoiGeom {
  int ox, oy, oz;        // 0-based start of the RI grid inside the simulation grid
  int nx, ny, nz;        // RI grid size
  long long nvox;
};

__global__ void green_plain(float2* __restrict__ spec, Tables tb, Consts c, const uint8_t* __restrict__ parity, int jj,
                            ConvArgs cv);
__global__ void green_plain_vec(float2* __restrict__ spec, long long comp_stride, Tables tb, Consts c,
                                const uint8_t* __restrict__ parity, int jj, ConvArgs cv);
__global__ void update_plain(StateArgs st, Tables tb, Consts c, int upd, int fwd, int jj_inv, int jj_fwd, ConvArgs cv);
__global__ void fill_parity(uint8_t* __restrict__ parity, int nslots, int P, int flip_state, int rhs0, int M);
__global__ void zero_words(uint32_t* __restrict__ p, long long n);
__global__ void seed_source(const float2* __restrict__ u_in, const float2* __restrict__ V, float2* __restrict__ fstate,
                            Consts c, RoiGeom rg, long long rhs_stride, int nrhs);
__global__ void finalize_field(const float2* __restrict__ field, const float2* __restrict__ u_in, float2* __restrict__ u_out,
                               Consts c, RoiGeom rg, long long rhs_stride, int nrhs);
__global__ void tile_rows(const float2* __restrict__ V, float2* __restrict__ Vt, Consts c);
__global__ void tile_rows_real(const float2* __restrict__ V, float* __restrict__ Vr, Consts c);
__global__ void sum_slots(const double* __restrict__ src, double* __restrict__ dst, long long n);

// RI grid given as FP64 values, x fastest: real (cplx = 0) or interleaved complex.
struct RiGrid {
  const double* ri;
  int cplx;
  int nx, ny, nz;        // RI grid size
  double kk;             // k^2 with k = 2 pi n_m / lambda
  double inv_n2;         // 1 / n_m^2
};
__global__ void potential_vmax(RiGrid g, double* __restrict__ block_max);
__global__ void potential_fill(RiGrid g, int flip_conj, RoiGeom rg, int snx, int sny, long long snvox, double eps,
                               float2* __restrict__ V);

// ---- registry of templated pass kernels (passes_inst.hip)
using PassRFn = void (*)(struct PassRArgs);
using PassKFn = void (*)(struct PassKArgs);
using RowsFn = void (*)(float2*, long long, const float2*, int);
using StridedFn = void (*)(float2*, int, long long, long long, int, const float2*, int, int, long long);

// fn takes every update kind; fn_steady only UPD_DEFER / UPD_PAIR with FWD_GAMMA
// fn_vr, fn_steady_vr: the same passes reading a real potential (st.Vr; fused form only)
struct PassRInfo { PassRFn fn, fn_steady, fn_vr, fn_steady_vr; int nx, a, nt, lds, rows; };
struct PassKInfo { PassKFn fn; int b, nz, t, nt, lds, pole; };
struct RowsInfo { RowsFn fn; int n, rows, nt; };
struct StridedInfo { StridedFn fn; int n, t, nt; };

const PassRInfo* find_pass_real(int nx, int a);
const PassKInfo* find_pass_kspace(int b, int nz, int t, int pole);
const RowsInfo* find_fft_rows(int n);
const StridedInfo* find_fft_strided(int n, int t);
// per-axis form, one field component, compiled line lengths (lines_fused_inst.hip):
// x inverse + pointwise update + x forward, and z forward + Green's multiply + z inverse
using RowsUpdateFn = void (*)(StateArgs, Tables, Consts, const float2*, int, int, int, int, ConvArgs);
using StridedGreenFn = void (*)(float2*, Tables, Consts, const float2*, const uint8_t*, int, ConvArgs);
struct RowsUpdateInfo { RowsUpdateFn fn; int n, rows, nt; };
struct StridedGreenInfo { StridedGreenFn fn; int n, t, nt; };
const RowsUpdateInfo* find_rows_update(int n);
const StridedGreenInfo* find_strided_green(int n);
// enumerate for planning
int list_pass_real(const PassRInfo** out);
int list_pass_kspace(const PassKInfo** out);

}  // namespace cbs
