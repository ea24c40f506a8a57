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

// l = i / n, e = i % n for a block-uniform n, by a float reciprocal corrected by one step
// (the quotient is off by at most one for i < 2^24)
__device__ __forceinline__ void split_line(int i, int n, float inv_n, int& l, int& e) {
  l = (int)((float)i * inv_n);
  e = i - l * n;
  if (e < 0) {
    --l;
    e += n;
  } else if (e >= n) {
    ++l;
    e -= n;
  }
}

// NL contiguous lines of p.n elements per block (line r at data + r n)
template <int L, int NL, int NT>
__global__ __launch_bounds__(NT) void blue_rows(float2* data, long long nrows, BlueArgs p, int inverse) {
  __shared__ float2 lds[L * (NL + 1)];
  __shared__ float2 s_tw[L];
  __shared__ float2 s_b[L];
  for (int i = threadIdx.x; i < L; i += NT) {
    s_tw[i] = p.tw[i];
    s_b[i] = p.bpos[i];
  }
  const int n = p.n;
  const long long r0 = (long long)blockIdx.x * NL;
  const int nr = nrows - r0 < NL ? (int)(nrows - r0) : NL;
  float2* base = data + r0 * n;
  const BlueTile<NL> ad;
  for (int i = threadIdx.x; i < NL * L; i += NT) {
    const int l = i / L, e = i - l * L;
    float2 v = make_float2(0.f, 0.f);
    if (e < n && l < nr) {
      v = base[(long long)l * n + e];
      if (inverse) v = cconj(v);
      v = cmulc(v, p.chirp[e]);
    }
    lds[ad(l, e)] = v;
  }
  __syncthreads();
  blue_convolve<L, NL, NT>(lds, s_tw, s_b);
  const float inv_n = 1.0f / (float)n;
  for (int i = threadIdx.x; i < nr * n; i += NT) {
    int l, e;
    split_line(i, n, inv_n, l, e);
    const float2 v = cmulc(lds[ad(l, e)], p.chirp[e]);
    base[i] = inverse ? cconj(v) : v;
  }
}

// Strided lines: element e of line t at base + t + e es, the tile being T consecutive x
// positions (tile xt of ceil(nx / T)) of outer set o < nouter,
// base = (o / inner) os2 + (o % inner) os + xt T.
template <int L, int T, int NT>
__global__ __launch_bounds__(NT) void blue_strided(float2* data, int nx, long long es, long long os, int nouter,
                                                   BlueArgs p, int inverse, int inner, long long os2) {
  __shared__ float2 lds[L * (T + 1)];
  __shared__ float2 s_tw[L];
  __shared__ float2 s_b[L];
  for (int i = threadIdx.x; i < L; i += NT) {
    s_tw[i] = p.tw[i];
    s_b[i] = p.bpos[i];
  }
  const int n = p.n;
  const int ntile = (nx + T - 1) / T;
  const int xt = blockIdx.x % ntile;
  const int o = blockIdx.x / ntile;
  if (o >= nouter) return;
  const int tv = nx - xt * T < T ? nx - xt * T : T;
  float2* base = data + (long long)(o / inner) * os2 + (long long)(o % inner) * os + (long long)xt * T;
  const BlueTile<T> ad;
  for (int i = threadIdx.x; i < L * T; i += NT) {
    const int e = i / T, t = i - e * T;
    float2 v = make_float2(0.f, 0.f);
    if (e < n && t < tv) {
      v = base[e * es + t];
      if (inverse) v = cconj(v);
      v = cmulc(v, p.chirp[e]);
    }
    lds[ad(t, e)] = v;
  }
  __syncthreads();
  blue_convolve<L, T, NT>(lds, s_tw, s_b);
  for (int i = threadIdx.x; i < n * T; i += NT) {
    const int e = i / T, t = i - e * T;
    if (t < tv) {
      const float2 v = cmulc(lds[ad(t, e)], p.chirp[e]);
      base[e * es + t] = inverse ? cconj(v) : v;
    }
  }
}

// (L, lines per rows block, lines per strided tile): 5-smooth lengths whose two-step
// plan has both factors <= 36, lines per block sized to keep two blocks per CU up to
// L = 512 (16 lines: 128-byte tile rows in the strided form)
#define BLUE_LIST(X)                                                                                           \
  X(48, 16, 16) X(64, 16, 16) X(80, 16, 16) X(96, 16, 16) X(128, 16, 16) X(160, 16, 16) X(192, 16, 16)         \
  X(256, 16, 16) X(320, 16, 16) X(384, 16, 16) X(512, 16, 16) X(640, 8, 8) X(768, 8, 8) X(1024, 8, 8)          \
  X(1152, 8, 8) X(1296, 8, 8)

#define BLUE_ENTRY(L, NL, T) {&blue_rows<L, NL, 256>, &blue_strided<L, T, 256>, L, NL, T, 256, LineFft<L>::R2},
static const BlueInfo kBlue[] = {BLUE_LIST(BLUE_ENTRY)};

const BlueInfo* find_blue(int m) {
  const BlueInfo* best = nullptr;
  for (const auto& e : kBlue)
    if (e.L >= m && (!best || e.L < best->L)) best = &e;
  return best;
}

}  // namespace cbs
