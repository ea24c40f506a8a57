// Run-time-planned line FFTs (see fft_rt.hpp).
#include "fft_rt.hpp"

namespace cbs {

// i / n and i % n for a block-uniform n > 0 and 0 <= i < 2^22, from a float reciprocal
// corrected by one step (the float quotient is off by at most one there)
__device__ __forceinline__ int qdiv(int i, int n, float inv_n, int& rem) {
  int q = (int)((float)i * inv_n);
  int r = i - q * n;
  if (r < 0) {
    --q;
    r += n;
  } else if (r >= n) {
    ++q;
    r -= n;
  }
  rem = r;
  return q;
}

// One Stockham stage of radix R on nl lines of length L (line l at l * L in src and dst):
// butterfly j reads src[j + q L / R], scales by W_L^{(j mod Ns) q L / (Ns R)}, transforms
// and writes dst[(j / Ns) Ns R + (j mod Ns) + q Ns]; Ns is the product of the earlier radices.
template <int R, bool INV>
__device__ void rt_stage(const float2* __restrict__ src, float2* __restrict__ dst, int L, int Ns, int nl,
                         const float2* __restrict__ tw) {
  const int nb = L / R;
  const int step = L / (Ns * R);
  const float inv_nb = 1.0f / (float)nb, inv_ns = 1.0f / (float)Ns;
#pragma unroll 2
  for (int t = threadIdx.x; t < nl * nb; t += blockDim.x) {
    int j, jm;
    const int l = qdiv(t, nb, inv_nb, j);
    const int jq = qdiv(j, Ns, inv_ns, jm);
    const float2* s = src + l * L;
    float2* d = dst + l * L;
    float2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = s[j + q * nb];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      const float2 w = tw[jm * q * step];
      v[q] = INV ? cmulc(v[q], w) : cmul(v[q], w);
    }
    Dft<R, INV>::run(v);
    const int o = jq * Ns * R + jm;
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
  const float inv_L = 1.0f / (float)L;
#pragma unroll 2
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) {
    int m;
    (void)qdiv(t, L, inv_L, m);
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
#pragma unroll 2
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) {
    int m;
    (void)qdiv(t, L, inv_L, m);
    r[t] = cmul(r[t], p.bspec[m]);
  }
  __syncthreads();
  r = rt_stockham<true>(r, o, p, nl);
#pragma unroll 2
  for (int t = threadIdx.x; t < nl * L; t += blockDim.x) {
    int m;
    (void)qdiv(t, L, inv_L, m);
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
  const float inv_n = 1.0f / (float)n;
#pragma unroll 4
  for (int i = threadIdx.x; i < nl * n; i += blockDim.x) {
    int e;
    const int l = qdiv(i, n, inv_n, e);
    a[l * L + e] = base[i];
  }
  __syncthreads();
  float2* r = inverse ? rt_line<true>(a, b, p, nl) : rt_line<false>(a, b, p, nl);
#pragma unroll 4
  for (int i = threadIdx.x; i < nl * n; i += blockDim.x) {
    int e;
    const int l = qdiv(i, n, inv_n, e);
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
  const float inv_tv = 1.0f / (float)tv;
#pragma unroll 4
  for (int i = threadIdx.x; i < n * tv; i += blockDim.x) {
    int t;
    const int e = qdiv(i, tv, inv_tv, t);
    a[t * L + e] = base[e * es + t];
  }
  __syncthreads();
  float2* r = inverse ? rt_line<true>(a, b, p, tv) : rt_line<false>(a, b, p, tv);
#pragma unroll 4
  for (int i = threadIdx.x; i < n * tv; i += blockDim.x) {
    int t;
    const int e = qdiv(i, tv, inv_tv, t);
    base[e * es + t] = r[t * L + e];
  }
}

}  // namespace cbs
