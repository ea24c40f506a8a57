// Run-time-planned line FFTs (see fft_rt.hpp).
#include "fft_rt.hpp"

namespace cbs {

// One Stockham stage of radix R on nl lines of length L (line l at l * L in src and dst):
// butterfly j reads src[j + q L / R], scales by W_L^{(j mod Ns) q L / (Ns R)}, transforms
// and writes dst[(j / Ns) Ns R + (j mod Ns) + q Ns]; Ns is the product of the earlier radices.
template <int R, bool INV>
__device__ void rt_stage(const float2* __restrict__ src, float2* __restrict__ dst, int L, int Ns, int nl,
                         const float2* __restrict__ tw) {
  const int nb = L / R;
  const int step = L / (Ns * R);
  for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
    const int l = t / nb, j = t - l * nb;
    const float2* s = src + l * L;
    float2* d = dst + l * L;
    const int jm = j % Ns;
    float2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = s[j + q * nb];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      const float2 w = tw[jm * q * step];
      v[q] = INV ? cmulc(v[q], w) : cmul(v[q], w);
    }
    Dft<R, INV>::run(v);
    const int o = (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int q = 0; q < R; ++q) d[o + q * Ns] = v[q];
  }
}

// The Stockham transform of length p.L on nl lines in a; returns the buffer (a or b)
// that holds the result.  Called by every thread of the block.
template <bool INV>
__device__ float2* rt_stockham(float2* a, float2* b, const RtFft& p, int nl) {
  int Ns = 1;
  for (int s = 0; s < p.nst; ++s) {
    const int R = p.radix[s];
    switch (R) {
      case 2: rt_stage<2, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 3: rt_stage<3, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 4: rt_stage<4, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 5: rt_stage<5, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 7: rt_stage<7, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 8: rt_stage<8, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 9: rt_stage<9, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 11: rt_stage<11, INV>(a, b, p.L, Ns, nl, p.tw); break;
      case 13: rt_stage<13, INV>(a, b, p.L, Ns, nl, p.tw); break;
      default: rt_stage<16, INV>(a, b, p.L, Ns, nl, p.tw); break;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// The line transform of length p.n on nl lines whose first p.n elements sit in a (line
// l at l * p.L); returns the buffer holding the result in the same positions.
template <bool INV>
__device__ float2* rt_line(float2* a, float2* b, const RtFft& p, int nl) {
  if (!p.blue) return rt_stockham<INV>(a, b, p, nl);
  const int n = p.n, L = p.L;
  // x_n conj(c_n), zero padded to L (the inverse transform is conj(DFT(conj x)))
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) {
    const int m = t % L;
    float2 v = make_float2(0.f, 0.f);
    if (m < n) {
      v = a[t];
      if (INV) v = cconj(v);
      v = cmulc(v, p.chirp[m]);
    }
    a[t] = v;
  }
  __syncthreads();
  float2* r = rt_stockham<false>(a, b, p, nl);
  float2* o = r == a ? b : a;
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) r[t] = cmul(r[t], p.bspec[t % L]);
  __syncthreads();
  r = rt_stockham<true>(r, o, p, nl);
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) {
    const int m = t % L;
    if (m < n) {
      float2 v = cmulc(r[t], p.chirp[m]);
      r[t] = INV ? cconj(v) : v;
    }
  }
  __syncthreads();
  return r;
}

__global__ void fft_rows_rt(float2* data, long long nrows, RtFft p, int lines_per_block, int inverse) {
  extern __shared__ float2 rt_lds[];
  const int L = p.L, n = p.n;
  float2* a = rt_lds;
  float2* b = rt_lds + (long long)lines_per_block * L;
  const long long r0 = (long long)blockIdx.x * lines_per_block;
  const int nl = (int)((nrows - r0) < lines_per_block ? (nrows - r0) : lines_per_block);
  float2* base = data + r0 * n;
  for (int i = threadIdx.x; i < nl * n; i += blockDim.x) {
    const int l = i / n, e = i - l * n;
    a[l * L + e] = base[i];
  }
  __syncthreads();
  float2* r = inverse ? rt_line<true>(a, b, p, nl) : rt_line<false>(a, b, p, nl);
  for (int i = threadIdx.x; i < nl * n; i += blockDim.x) {
    const int l = i / n, e = i - l * n;
    base[i] = r[l * L + e];
  }
}

__global__ void fft_strided_rt(float2* data, int nx, long long es, long long os, int nouter, RtFft p, int T,
                               int inverse, int inner, long long os2) {
  extern __shared__ float2 rt_lds[];
  const int L = p.L, n = p.n;
  float2* a = rt_lds;
  float2* b = rt_lds + (long long)T * L;
  const int ntile = (nx + T - 1) / T;
  const int xt = blockIdx.x % ntile;
  const int o = blockIdx.x / ntile;
  if (o >= nouter) return;
  const int tv = nx - xt * T < T ? nx - xt * T : T;
  float2* base = data + (long long)(o / inner) * os2 + (long long)(o % inner) * os + (long long)xt * T;
  for (int i = threadIdx.x; i < n * tv; i += blockDim.x) {
    const int e = i / tv, t = i - e * tv;
    a[t * L + e] = base[e * es + t];
  }
  __syncthreads();
  float2* r = inverse ? rt_line<true>(a, b, p, tv) : rt_line<false>(a, b, p, tv);
  for (int i = threadIdx.x; i < n * tv; i += blockDim.x) {
    const int e = i / tv, t = i - e * tv;
    base[e * es + t] = r[t * L + e];
  }
}

}  // namespace cbs
