// Bluestein line transforms on compiled convolution lengths (see fft_rt.hpp).
//
// For a line x of length n with a prime factor above 13,
//   X_k = conj(c_k) sum_m x_m conj(c_m) c_{k-m},   c_k = exp(i pi k^2 / n),
// the sum being a circular convolution of length L >= 2n - 1 with the chirp filter
// (c_m at m and at L - m).  A block loads a tile of lines, multiplies by conj(c) on the
// way into LDS (zero padded to L), runs the compiled forward FFT, the multiply by the
// filter's spectrum and the inverse FFT (line_fft_mul_ifft), and multiplies by conj(c)
// on the way out.  The filter spectrum carries 1/L, so the result is the plain DFT.
#include "fft_rt.hpp"

namespace cbs {

// LDS index of element e of line l: rows of NL + 1 words, so lanes that walk the
// lines of one element and lanes that walk the elements of one line both spread
// over the banks.
template <int NL>
struct BlueTile {
  __device__ __forceinline__ int operator()(int l, int e) const { return e * (NL + 1) + l; }
};

template <int L, int NL, int NT>
__device__ __forceinline__ void blue_convolve(float2* lds, const float2* s_tw, const float2* s_b) {
  const BlueTile<NL> ad;
  auto mul = [&](int, int p0, int p1, bool hi, float2 v) { return cmul(v, s_b[hi ? p1 : p0]); };
  line_fft_mul_ifft<L, NL, NT>(lds, ad, s_tw, (int)threadIdx.x, mul);
}

// Tables of the convolution (twiddles of L, filter spectrum) into LDS
template <int L, int NT>
__device__ __forceinline__ void blue_tables(const BlueArgs& p, float2* s_tw, float2* s_b) {
  constexpr int K = (L + NT - 1) / NT;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    if (i < L) {
      s_tw[i] = p.tw[i];
      s_b[i] = p.bpos[i];
    }
  }
}

// NL contiguous lines of p.n elements per block (line r at data + r n).  Every global
// load of a thread (data and chirp) is issued before the first one is used: the element
// loop runs over the compile-time bound n <= (L + 1) / 2.
template <int L, int NL, int NT>
__global__ __launch_bounds__(NT) void blue_rows(float2* data, long long nrows, BlueArgs p, int inverse) {
  constexpr int NMAX = (L + 1) / 2;                    // L >= 2n - 1
  constexpr int K = (NL * NMAX + NT - 1) / NT;
  constexpr int KZ = (NL * (L - NMAX) + NT - 1) / NT;
  __shared__ float2 lds[L * (NL + 1)];
  __shared__ float2 s_tw[L];
  __shared__ float2 s_b[L];
  const int n = p.n;
  const long long r0 = (long long)blockIdx.x * NL;
  const int nr = nrows - r0 < NL ? (int)(nrows - r0) : NL;
  float2* base = data + r0 * n;
  const BlueTile<NL> ad;
  float2 xv[K], cv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / NMAX, e = i - l * NMAX;
    xv[k] = cv[k] = make_float2(0.f, 0.f);
    if (i < NL * NMAX && e < n && l < nr) {
      xv[k] = base[(long long)l * n + e];
      cv[k] = p.chirp[e];
    }
  }
  blue_tables<L, NT>(p, s_tw, s_b);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / NMAX, e = i - l * NMAX;
    if (i < NL * NMAX) lds[ad(l, e)] = cmulc(inverse ? cconj(xv[k]) : xv[k], cv[k]);
  }
#pragma unroll
  for (int k = 0; k < KZ; ++k) {                       // zero padding above NMAX
    const int i = threadIdx.x + k * NT;
    const int l = i / (L - NMAX), e = NMAX + i - l * (L - NMAX);
    if (i < NL * (L - NMAX)) lds[ad(l, e)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  blue_convolve<L, NL, NT>(lds, s_tw, s_b);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / NMAX, e = i - l * NMAX;
    if (i < NL * NMAX && e < n && l < nr) cv[k] = p.chirp[e];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / NMAX, e = i - l * NMAX;
    if (i < NL * NMAX && e < n && l < nr) {
      const float2 v = cmulc(lds[ad(l, e)], cv[k]);
      base[(long long)l * n + e] = inverse ? cconj(v) : v;
    }
  }
}

// Strided lines: element e of line t at base + t + e es, the tile being T consecutive x
// positions (tile xt of ceil(nx / T)) of outer set o < nouter,
// base = (o / inner) os2 + (o % inner) os + xt T.  Loads issued as in blue_rows.
template <int L, int T, int NT>
__global__ __launch_bounds__(NT) void blue_strided(float2* data, int nx, long long es, long long os, int nouter,
                                                   BlueArgs p, int inverse, int inner, long long os2) {
  constexpr int NMAX = (L + 1) / 2;
  constexpr int K = (NMAX * T + NT - 1) / NT;
  constexpr int KZ = ((L - NMAX) * T + NT - 1) / NT;
  __shared__ float2 lds[L * (T + 1)];
  __shared__ float2 s_tw[L];
  __shared__ float2 s_b[L];
  const int n = p.n;
  const int ntile = (nx + T - 1) / T;
  const int xt = blockIdx.x % ntile;
  const int o = blockIdx.x / ntile;
  if (o >= nouter) return;
  const int tv = nx - xt * T < T ? nx - xt * T : T;
  float2* base = data + (long long)(o / inner) * os2 + (long long)(o % inner) * os + (long long)xt * T;
  const BlueTile<T> ad;
  float2 xv[K], cv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    xv[k] = cv[k] = make_float2(0.f, 0.f);
    if (i < NMAX * T && e < n && t < tv) {
      xv[k] = base[e * es + t];
      cv[k] = p.chirp[e];
    }
  }
  blue_tables<L, NT>(p, s_tw, s_b);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    if (i < NMAX * T) lds[ad(t, e)] = cmulc(inverse ? cconj(xv[k]) : xv[k], cv[k]);
  }
#pragma unroll
  for (int k = 0; k < KZ; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = NMAX + i / T, t = i % T;
    if (i < (L - NMAX) * T) lds[ad(t, e)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  blue_convolve<L, T, NT>(lds, s_tw, s_b);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    if (i < NMAX * T && e < n && t < tv) cv[k] = p.chirp[e];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    if (i < NMAX * T && e < n && t < tv) {
      const float2 v = cmulc(lds[ad(t, e)], cv[k]);
      base[e * es + t] = inverse ? cconj(v) : v;
    }
  }
}

// (L, lines per rows block, lines per strided tile, threads): 5-smooth lengths whose
// two-step plan has both factors <= 36
#define BLUE_LIST(X)                                                                                           \
  X(48, 16, 16, 256) X(64, 16, 16, 256) X(80, 16, 16, 256) X(96, 16, 16, 256) X(128, 16, 16, 256)              \
  X(160, 16, 16, 256) X(192, 16, 16, 256) X(256, 16, 16, 256) X(320, 16, 16, 256) X(384, 16, 16, 256)          \
  X(512, 16, 16, 256) X(640, 8, 8, 256) X(768, 8, 8, 256) X(1024, 8, 8, 256) X(1152, 8, 8, 256)                \
  X(1296, 8, 8, 256)

#define BLUE_ENTRY(L, NL, T, NT) {&blue_rows<L, NL, NT>, &blue_strided<L, T, NT>, L, NL, T, NT, LineFft<L>::R2},
static const BlueInfo kBlue[] = {BLUE_LIST(BLUE_ENTRY)};

const BlueInfo* find_blue(int m) {
  const BlueInfo* best = nullptr;
  for (const auto& e : kBlue)
    if (e.L >= m && (!best || e.L < best->L)) best = &e;
  return best;
}

}  // namespace cbs
