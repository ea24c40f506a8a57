// Fused two-pass Born iteration and the unfused (plain) per-axis form.
// Templates are instantiated per supported line length in passes_*.hip.
#pragma once
#include "cbs_kernels.hpp"

namespace cbs {

// LDS row pitch: odd (in 8-byte words) so that lanes walking down a column of
// rows hit distinct banks.
__host__ __device__ constexpr int row_pitch(int n) { return (n % 2 == 0) ? n + 1 : n; }

// ------------------------------------------------------------------ pass_real
// One block per (rhs, y2, z).  Tile = a rows (y1) of nx points.
struct PassRArgs {
  Tables tb;
  Consts c;
  StateArgs st;
  int upd, fwd, jj_inv, jj_fwd;
};

template <int NX, int A, int NT>
__global__ __launch_bounds__(NT) void pass_real(PassRArgs args) {
  constexpr int PR = row_pitch(NX);
  __shared__ float2 lds[A * PR];
  const Consts& c = args.c;
  const Tables& tb = args.tb;
  const int tid = threadIdx.x;
  const int bid = blockIdx.x;
  const int r = bid % c.nrhs;
  const int tile = bid / c.nrhs;
  const int y2 = tile % c.b;
  const int z = tile / c.b;
  const long long roff = (long long)r * args.st.rhs_stride;
  float2* __restrict__ spec = args.st.spec + roff + (long long)(z * c.b + y2) * (A * NX);
  float2* __restrict__ F = args.st.field + roff;
  float2* __restrict__ f = args.st.fstate + roff;
  const int flip_inv = args.jj_inv ? flip_of(args.jj_inv, args.st.parity, r) : 0;
  const int flip_fwd = args.jj_fwd ? flip_of(args.jj_fwd, args.st.parity, r) : 0;

  auto row = [](int l, int e) { return l * PR + e; };
  auto col = [](int l, int e) { return e * PR + l; };

  if (args.upd != UPD_NONE) {
    const float2* __restrict__ cr = tb.cross + y2 * A;
    for (int i = tid; i < A * NX; i += NT) {
      const int p1 = i / NX, px = i - p1 * NX;
      float2 v = spec[i];
      if (c.b > 1) v = cmulc(v, cr[p1]);
      lds[p1 * PR + px] = v;
    }
    __syncthreads();
    if constexpr (A > 1) line_fft<A, true, NX, NT>(lds, col, tb.tw_a, tid);
    line_fft<NX, true, A, NT>(lds, row, tb.tw_x, tid);
  }
  double nrm = 0.0;
  for (int i = tid; i < A * NX; i += NT) {
    const int y1 = i / NX, x = i - y1 * NX;
    const int y = y1 * c.b + y2;
    const long long g = ((long long)z * c.ny + y) * NX + x;
    const float2 q = (args.upd != UPD_NONE) ? lds[y1 * PR + x] : make_float2(0.f, 0.f);
    const float2 p = born_point(args.upd, args.fwd, q, g, x, y, z, tb, c, F, f, args.st.V, flip_inv, flip_fwd, nrm);
    if (args.fwd != FWD_NONE) lds[y1 * PR + x] = p;
  }
  if (args.st.norms) flush_norm(nrm, args.st.norms + r);
  if (args.fwd == FWD_NONE) return;
  __syncthreads();
  line_fft<NX, false, A, NT>(lds, row, tb.tw_x, tid);
  if constexpr (A > 1) line_fft<A, false, NX, NT>(lds, col, tb.tw_a, tid);
  {
    const float2* __restrict__ cr = tb.cross + y2 * A;
    for (int i = tid; i < A * NX; i += NT) {
      const int p1 = i / NX, px = i - p1 * NX;
      float2 v = lds[p1 * PR + px];
      if (c.b > 1) v = cmul(v, cr[p1]);
      spec[i] = v;
    }
  }
}

// ------------------------------------------------------------------ pass_kspace
// One block per (rhs, p1, x tile of T).  Tile = T x-positions * B (y2) * NZ.
struct PassKArgs {
  Tables tb;
  Consts c;
  float2* spec;
  const uint8_t* parity;
  long long rhs_stride;
  int jj;
};

template <int B, int NZ, int T, int NT>
__global__ __launch_bounds__(NT) void pass_kspace(PassKArgs args) {
  constexpr int LW = B * T;   // lines of the z-FFT
  __shared__ float2 lds[NZ * LW];
  const Consts& c = args.c;
  const Tables& tb = args.tb;
  const int tid = threadIdx.x;
  const int bid = blockIdx.x;
  const int r = bid % c.nrhs;
  const int t2 = bid / c.nrhs;
  const int p1 = t2 % c.a;
  const int xt = t2 / c.a;
  const int x0 = xt * T;
  const long long rowstride = (long long)c.a * c.nx;     // between consecutive y2
  float2* __restrict__ spec = args.spec + (long long)r * args.rhs_stride + (long long)p1 * c.nx + x0;
  const int flip = flip_of(args.jj, args.parity, r);

  const int tvalid = c.nx - x0 < T ? c.nx - x0 : T;
  for (int task = tid; task < NZ * T; task += NT) {
    const int t = task % T, z = task / T;
    float2 v[B];
    const float2* src = spec + (long long)z * B * rowstride + t;
#pragma unroll
    for (int y2 = 0; y2 < B; ++y2) v[y2] = t < tvalid ? src[y2 * rowstride] : make_float2(0.f, 0.f);
    Dft<B, false>::run(v);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) lds[z * LW + k2 * T + t] = v[k2];
  }
  __syncthreads();
  auto zl = [](int l, int e) { return e * LW + l; };
  line_fft<NZ, false, LW, NT>(lds, zl, tb.tw_z, tid);
  {
    const float* gx = tb.green_x + flip * c.nx + x0;
    const float* gy = tb.green_y + flip * c.ny + p1;
    const float* gz = tb.green_z + flip * c.nz;
    for (int i = tid; i < NZ * LW; i += NT) {
      const int pz = i / LW, l = i - pz * LW;
      const int k2 = l / T, t = l - k2 * T;
      const float d = gx[t < tvalid ? t : 0] + gy[k2 * c.a] + gz[pz] - c.km2;
      // (1/N) / (d - i eps) = (d + i eps) / (N (d^2 + eps^2))
      const float s = c.inv_n / fmaf(d, d, c.eps * c.eps);
      const float2 g = make_float2(d * s, c.eps * s);
      lds[i] = cmul(lds[i], g);
    }
  }
  __syncthreads();
  line_fft<NZ, true, LW, NT>(lds, zl, tb.tw_z, tid);
  for (int task = tid; task < NZ * T; task += NT) {
    const int t = task % T, z = task / T;
    if (t >= tvalid) continue;
    float2 v[B];
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[k2] = lds[z * LW + k2 * T + t];
    Dft<B, true>::run(v);
    float2* dst = spec + (long long)z * B * rowstride + t;
#pragma unroll
    for (int y2 = 0; y2 < B; ++y2) dst[y2 * rowstride] = v[y2];
  }
}

// ------------------------------------------------------------------ plain form
// Unfused single-rhs iteration: per-axis FFT launches on a natural-layout
// spectrum, a Green's multiply launch and a pointwise update launch.

// FFT along x for R rows per block (rows are contiguous nx-vectors).
template <int NX, int ROWS, int NT>
__global__ __launch_bounds__(NT) void fft_rows(float2* data, long long nrows, const float2* tw, int inverse) {
  constexpr int PR = row_pitch(NX);
  __shared__ float2 lds[ROWS * PR];
  const long long r0 = (long long)blockIdx.x * ROWS;
  float2* base = data + r0 * NX;
  const int nr = (int)((nrows - r0) < ROWS ? (nrows - r0) : ROWS);
  for (int i = threadIdx.x; i < ROWS * NX; i += NT) {
    const int rr = i / NX, e = i - rr * NX;
    lds[rr * PR + e] = rr < nr ? base[i] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  auto row = [](int l, int e) { return l * PR + e; };
  if (inverse) line_fft<NX, true, ROWS, NT>(lds, row, tw, threadIdx.x);
  else line_fft<NX, false, ROWS, NT>(lds, row, tw, threadIdx.x);
  for (int i = threadIdx.x; i < ROWS * NX; i += NT) {
    const int rr = i / NX, e = i - rr * NX;
    if (rr < nr) base[i] = lds[rr * PR + e];
  }
}

// FFT along a strided axis of length N (stride `es` elements).  A block takes
// T consecutive x positions of one line set: element e of line t at
// base + t + e*es.  Blocks enumerate (x tile, outer index) with
// base = outer * os + xt * T.
template <int N, int T, int NT>
__global__ __launch_bounds__(NT) void fft_strided(float2* data, int nx, long long es, long long os, int nouter,
                                                  const float2* tw, int inverse) {
  __shared__ float2 lds[N * T];
  const int ntile = (nx + T - 1) / T;
  const int xt = blockIdx.x % ntile;
  const int o = blockIdx.x / ntile;
  const int tvalid = nx - xt * T < T ? nx - xt * T : T;
  float2* base = data + (long long)o * os + (long long)xt * T;
  for (int i = threadIdx.x; i < N * T; i += NT) {
    const int e = i / T, t = i - e * T;
    lds[i] = t < tvalid ? base[e * es + t] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  auto ad = [](int l, int e) { return e * T + l; };
  if (inverse) line_fft<N, true, T, NT>(lds, ad, tw, threadIdx.x);
  else line_fft<N, false, T, NT>(lds, ad, tw, threadIdx.x);
  for (int i = threadIdx.x; i < N * T; i += NT) {
    const int e = i / T, t = i - e * T;
    if (t < tvalid) base[e * es + t] = lds[i];
  }
}

}  // namespace cbs
