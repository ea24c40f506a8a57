// Line FFTs of any length, planned at run time (the per-axis form and the 2D
// convolution-grid transforms of grids whose line lengths have no compiled kernel).
//
// A length n = product of radices from {2, 3, 4, 5, 7, 8, 9, 11, 13, 16} runs as a
// mixed-radix Stockham transform (natural order in and out) in LDS, one stage per
// radix, ping-ponging between two LDS buffers.  Any other n (a prime factor above 13)
// runs as Bluestein's chirp-z transform: x_n conj(c_n), c_n = exp(i pi n^2 / n), is
// convolved with c through a Stockham transform of a length L >= 2n - 1 whose factors
// are 2, 3, 5 and 7, and the result is multiplied by conj(c_k).  Unnormalised in both
// directions, like the compiled line FFTs.
#pragma once
#include "cbs_kernels.hpp"

namespace cbs {

constexpr int kRtMaxStages = 24;
constexpr int kRtLdsElems = 8192;   // complex64 per block (64 KiB): both ping-pong buffers of every line

struct RtFft {
  int n;                    // line length
  int L;                    // transform length: n, or the Bluestein convolution length
  int nst;                  // Stockham stages
  int blue;                 // 1: Bluestein
  int radix[kRtMaxStages];  // stage radices, product L
  const float2* tw;         // [L] exp(-2 pi i e / L)
  const float2* chirp;      // [n] c_k = exp(i pi k^2 / n)             (Bluestein)
  const float2* bspec;      // [L] DFT_L of the chirp filter, over L   (Bluestein)
};

// Lines per block that fit the LDS budget (at least 1, at most cap).
__host__ __device__ inline int rt_lines(int L, int cap) {
  const int k = kRtLdsElems / (2 * L);
  return k < 1 ? 1 : (k > cap ? cap : k);
}

// n contiguous elements per row, nrows rows; lines_per_block = rt_lines(p.L, 64);
// dynamic LDS = 2 * lines_per_block * p.L * sizeof(float2).
__global__ void fft_rows_rt(float2* data, long long nrows, RtFft p, int lines_per_block, int inverse);

// Strided lines: element e of line t in the block's tile at base + t + e * es, the tile
// being T consecutive x positions (x tile xt of ceil(nx / T)) of outer set o < nouter:
// base = (o / inner) * os2 + (o % inner) * os + xt * T.  T = rt_lines(p.L, 16); dynamic
// LDS = 2 * T * p.L * 8 bytes.
__global__ void fft_strided_rt(float2* data, int nx, long long es, long long os, int nouter, RtFft p, int T,
                               int inverse, int inner, long long os2);

// Bluestein's transform on a compiled convolution length L >= 2n - 1 (fft_blue.hip):
// the chirp products are fused into the tile load and store, and the convolution is
// one compiled forward FFT, multiply and inverse FFT of length L (line_fft_mul_ifft),
// so no stage runs on run-time radices.  The inverse is conj(DFT(conj x)).
struct BlueArgs {
  int n;                    // line length
  const float2* chirp;      // [n] c_k = exp(i pi k^2 / n)
  const float2* bpos;       // [L] DFT_L of the chirp filter over L, in the forward output order of L
  const float2* tw;         // [L] exp(-2 pi i e / L)
};
// rows: NL contiguous lines of n per block; strided: as fft_strided_rt with T = t lines
using BlueRowsFn = void (*)(float2*, long long, BlueArgs, int);
using BlueStridedFn = void (*)(float2*, int, long long, long long, int, BlueArgs, int, int, long long);
// per-axis form, one field component: x inverse + pointwise update + x forward, and z
// forward + Green's multiply + z inverse (fft_blue.hip)
using BlueRowsUpdateFn = void (*)(StateArgs, Tables, Consts, BlueArgs, int, int, int, int, ConvArgs);
using BlueStridedGreenFn = void (*)(float2*, Tables, Consts, BlueArgs, const uint8_t*, int, ConvArgs);
struct BlueInfo {
  BlueRowsFn rows;
  BlueStridedFn strided;
  BlueRowsUpdateFn rows_update;
  BlueStridedGreenFn strided_green;
  int L, nl, t, nt;
  int r2;                   // forward output position p holds frequency p / r2 + (L / r2) (p % r2)
};
// the compiled length L >= m with the fewest points (null when m is above every one)
const BlueInfo* find_blue(int m);

}  // namespace cbs
