// Register-resident small DFTs and LDS-staged line FFTs for gfx950 (CDNA4).
//
// Building blocks for the convergent-Born-series passes:
//   * Dft<R, INV>::run(v)   in-place R-point DFT on R complex values held in
//     registers, natural order in and out; twiddles are compile-time constants.
//   * LineFft<N>            two-step (four-step) plan N = R1 * R2 for a line FFT
//     whose data sits in LDS.  The forward transform leaves X[k1 + R1*k2] at
//     position R2*k1 + k2 (a transposed order); the inverse transform consumes
//     exactly that order and returns natural order, so a forward/inverse pair
//     never needs a reorder pass.  pos2freq() maps a position to its frequency.
//
// Sign convention: forward = exp(-2*pi*i*n*k/N), inverse = exp(+2*pi*i*n*k/N),
// both unnormalised (the 1/N of the inverse is folded into the Green's multiply).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cbs {

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward quarter turn) / +i
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ __forceinline__ float2 cmul_pi(float2 a) { return make_float2(-a.y, a.x); }

// ---------------------------------------------------------------- compile-time trig
constexpr double kPi = 3.14159265358979323846264338327950288;

// sin/cos of 2*pi*j/n by Taylor series on an argument reduced to [-pi, pi];
// accurate to ~1e-15, far below the float rounding applied afterwards.
constexpr double ct_sin_red(double x) {
  double term = x, sum = x, x2 = x * x;
  for (int n = 1; n < 30; ++n) { term *= -x2 / double((2 * n) * (2 * n + 1)); sum += term; }
  return sum;
}
constexpr double ct_cos_red(double x) {
  double term = 1.0, sum = 1.0, x2 = x * x;
  for (int n = 1; n < 30; ++n) { term *= -x2 / double((2 * n - 1) * (2 * n)); sum += term; }
  return sum;
}
constexpr double ct_arg(long j, long n) {
  long m = ((j % n) + n) % n;          // 0..n-1
  if (2 * m > n) m -= n;                // -n/2..n/2
  return 2.0 * kPi * double(m) / double(n);
}
constexpr double ct_cos2pi(long j, long n) { return ct_cos_red(ct_arg(j, n)); }
constexpr double ct_sin2pi(long j, long n) { return ct_sin_red(ct_arg(j, n)); }

template <int R>
struct RootTable {
  float c[R > 0 ? R : 1];
  float s[R > 0 ? R : 1];
  constexpr RootTable() : c(), s() {
    for (int e = 0; e < R; ++e) {
      c[e] = float(ct_cos2pi(e, R));
      s[e] = float(ct_sin2pi(e, R));
    }
  }
};
template <int R>
__device__ constexpr RootTable<R> kRoots = RootTable<R>();

// v * W_R^e with W_R = exp(-2 pi i / R) (forward) or its conjugate (inverse).
// e is a compile-time constant after unrolling; trivial rotations fold away.
template <int R, bool INV>
__device__ __forceinline__ float2 twiddle_ct(float2 v, int e) {
  e %= R;
  if (e == 0) return v;
  if (4 * e == R) return INV ? cmul_pi(v) : cmul_mi(v);
  if (2 * e == R) return make_float2(-v.x, -v.y);
  if (4 * e == 3 * R) return INV ? cmul_mi(v) : cmul_pi(v);
  const float c = kRoots<R>.c[e];
  const float s = INV ? kRoots<R>.s[e] : -kRoots<R>.s[e];
  return make_float2(fmaf(v.x, c, -v.y * s), fmaf(v.x, s, v.y * c));
}

// ---------------------------------------------------------------- small DFTs
constexpr int first_factor(int r) {
  return (r % 4 == 0 && r != 4) ? 4 : (r % 2 == 0 && r != 2) ? 2 : (r % 3 == 0 && r != 3) ? 3
       : (r % 5 == 0 && r != 5) ? 5 : (r % 7 == 0 && r != 7) ? 7 : (r % 11 == 0 && r != 11) ? 11
       : (r % 13 == 0 && r != 13) ? 13 : r;
}
constexpr bool is_prime_above5(int r) {
  if (r <= 5) return false;
  for (int d = 2; d * d <= r; ++d)
    if (r % d == 0) return false;
  return true;
}

template <int R, bool INV>
struct Dft;

template <bool INV>
struct Dft<1, INV> { __device__ __forceinline__ static void run(float2*) {} };

template <bool INV>
struct Dft<2, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <bool INV>
struct Dft<3, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float c1 = -0.5f;
    constexpr float s1 = INV ? 0.86602540378443864676f : -0.86602540378443864676f;
    float2 a = v[0], b = v[1], c = v[2];
    float2 t = cadd(b, c), d = csub(b, c);
    float2 m = make_float2(fmaf(t.x, c1, a.x), fmaf(t.y, c1, a.y));
    v[0] = cadd(a, t);
    // (b - c) * i*s1
    float2 r = make_float2(-d.y * s1, d.x * s1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <bool INV>
struct Dft<4, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    float2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
    b1 = INV ? cmul_pi(b1) : cmul_mi(b1);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  }
};

template <bool INV>
struct Dft<5, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    // Winograd-style 5-point DFT
    constexpr float c1 = 0.30901699437494742410f;   // cos(2pi/5)
    constexpr float c2 = -0.80901699437494742410f;  // cos(4pi/5)
    constexpr float s1 = INV ? 0.95105651629515357212f : -0.95105651629515357212f;
    constexpr float s2 = INV ? 0.58778525229247312917f : -0.58778525229247312917f;
    float2 x0 = v[0];
    float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    float2 m1 = make_float2(x0.x + c1 * t1.x + c2 * t2.x, x0.y + c1 * t1.y + c2 * t2.y);
    float2 m2 = make_float2(x0.x + c2 * t1.x + c1 * t2.x, x0.y + c2 * t1.y + c1 * t2.y);
    // i*(s1*t3 + s2*t4) and i*(s2*t3 - s1*t4)
    float2 n1 = make_float2(-(s1 * t3.y + s2 * t4.y), s1 * t3.x + s2 * t4.x);
    float2 n2 = make_float2(-(s2 * t3.y - s1 * t4.y), s2 * t3.x - s1 * t4.x);
    v[0] = cadd(x0, cadd(t1, t2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
};

// Odd prime R > 5 (7, 11, 13 for the run-time plans): with a_j = x_j + x_{R-j} and
// b_j = x_j - x_{R-j}, j = 1..(R-1)/2, and t = 2 pi j k / R,
//   X_k = x_0 + sum_j a_j cos t - i sum_j b_j sin t,  X_{R-k} = the same with + i,
// and the signs of i swapped for the inverse.
template <int R, bool INV>
struct DftPrime {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr int H = (R - 1) / 2;
    float2 a[H], b[H];
    const float2 x0 = v[0];
    float2 sum = x0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      a[j - 1] = cadd(v[j], v[R - j]);
      b[j - 1] = csub(v[j], v[R - j]);
      sum = cadd(sum, a[j - 1]);
    }
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      float2 A = x0, B = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const int e = (j * k) % R;
        const float c = kRoots<R>.c[e], sn = kRoots<R>.s[e];
        A = make_float2(fmaf(a[j - 1].x, c, A.x), fmaf(a[j - 1].y, c, A.y));
        B = make_float2(fmaf(b[j - 1].x, sn, B.x), fmaf(b[j - 1].y, sn, B.y));
      }
      const float2 iB = make_float2(-B.y, B.x);
      v[k] = INV ? cadd(A, iB) : csub(A, iB);
      v[R - k] = INV ? csub(A, iB) : cadd(A, iB);
    }
    v[0] = sum;
  }
};

// Composite R = F * M, four-step in registers:
//   n = M*n1 + n2, k = k1 + F*k2;  X[k] = sum_n2 W_R^{n2 k1} W_M^{n2 k2} sum_n1 x[M n1 + n2] W_F^{n1 k1}
template <int R, bool INV>
struct Dft {
  static constexpr int F = is_prime_above5(R) ? 1 : first_factor(R);
  static constexpr int M = is_prime_above5(R) ? R : R / F;
  static_assert(is_prime_above5(R) || (F > 1 && F < R), "radix not supported (factors 2, 3, 5, 7, 11, 13)");
  __device__ __forceinline__ static void run(float2* v) {
    if constexpr (is_prime_above5(R)) {
      DftPrime<R, INV>::run(v);
    } else {
      float2 w[R];
#pragma unroll
      for (int n2 = 0; n2 < M; ++n2) {
        float2 u[F];
#pragma unroll
        for (int n1 = 0; n1 < F; ++n1) u[n1] = v[M * n1 + n2];
        Dft<F, INV>::run(u);
#pragma unroll
        for (int k1 = 0; k1 < F; ++k1) w[k1 * M + n2] = twiddle_ct<R, INV>(u[k1], n2 * k1);
      }
#pragma unroll
      for (int k1 = 0; k1 < F; ++k1) {
        float2 u[M];
#pragma unroll
        for (int n2 = 0; n2 < M; ++n2) u[n2] = w[k1 * M + n2];
        Dft<M, INV>::run(u);
#pragma unroll
        for (int k2 = 0; k2 < M; ++k2) v[k1 + F * k2] = u[k2];
      }
    }
  }
};

// ---------------------------------------------------------------- line FFT plans
constexpr bool is_smooth(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}
// R2 = largest divisor of n not above sqrt(n) (single step when n <= 16)
constexpr int plan_r2(int n) {
  if (n <= 16) return 1;
  int best = 1;
  for (int d = 1; d * d <= n; ++d)
    if (n % d == 0) best = d;
  return best;
}

template <int N>
struct LineFft {
  static_assert(is_smooth(N), "line length must factor into 2, 3 and 5");
  static constexpr int R2 = plan_r2(N);
  static constexpr int R1 = N / R2;
  static_assert(R1 <= 36 && R2 <= 36, "line too long for a two-step plan");
  // frequency held at position p after the forward transform
  __host__ __device__ static constexpr int pos2freq(int p) {
    return R2 == 1 ? p : (p / R2) + R1 * (p % R2);
  }
  __host__ __device__ static constexpr int freq2pos(int k) {
    return R2 == 1 ? k : (k % R1) * R2 + (k / R1);
  }
};

// Forward (INV=false) or inverse line FFTs on NL lines held in LDS.
// addr(l, e) gives the LDS index of element e of line l.  tw points at the
// forward twiddle table W_N^e = exp(-2 pi i e / N), e in [0, N), in any memory.
// Must be called by all NT threads of the block; contains __syncthreads().
template <int N, bool INV, int NL, int NT, class Addr>
__device__ __forceinline__ void line_fft(float2* __restrict__ lds, const Addr& addr,
                                         const float2* __restrict__ tw, int tid) {
  using P = LineFft<N>;
  constexpr int R1 = P::R1, R2 = P::R2;
  if constexpr (R2 == 1) {
    for (int l = tid; l < NL; l += NT) {
      float2 v[N];
#pragma unroll
      for (int e = 0; e < N; ++e) v[e] = lds[addr(l, e)];
      Dft<N, INV>::run(v);
#pragma unroll
      for (int e = 0; e < N; ++e) lds[addr(l, e)] = v[e];
    }
    __syncthreads();
  } else if constexpr (!INV) {
    for (int t = tid; t < NL * R2; t += NT) {
      const int l = t % NL, n2 = t / NL;
      float2 v[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) v[n1] = lds[addr(l, R2 * n1 + n2)];
      Dft<R1, false>::run(v);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], tw[n2 * k1]);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[addr(l, R2 * k1 + n2)] = v[k1];
    }
    __syncthreads();
    for (int t = tid; t < NL * R1; t += NT) {
      const int l = t % NL, k1 = t / NL;
      float2 v[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) v[n2] = lds[addr(l, R2 * k1 + n2)];
      Dft<R2, false>::run(v);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) lds[addr(l, R2 * k1 + k2)] = v[k2];
    }
    __syncthreads();
  } else {
    for (int t = tid; t < NL * R1; t += NT) {
      const int l = t % NL, k1 = t / NL;
      float2 v[R2];
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k2] = lds[addr(l, R2 * k1 + k2)];
      Dft<R2, true>::run(v);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) lds[addr(l, R2 * k1 + n2)] = v[n2];
    }
    __syncthreads();
    for (int t = tid; t < NL * R2; t += NT) {
      const int l = t % NL, n2 = t / NL;
      float2 v[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[k1] = lds[addr(l, R2 * k1 + n2)];
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmulc(v[k1], tw[n2 * k1]);
      Dft<R1, true>::run(v);
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) lds[addr(l, R2 * n1 + n2)] = v[n1];
    }
    __syncthreads();
  }
}

// Line-group index t / NL of task t.  When NL and NT are multiples of the wave
// size, every lane of a wave has the same quotient; saying so with readfirstlane
// lets the compiler fetch twiddles and per-position factors that depend only on it
// with scalar loads instead of one vector load per lane.
template <int NL, int NT>
__device__ __forceinline__ int group_of(int t) {
  if constexpr (NL % 64 == 0 && NT % 64 == 0) return __builtin_amdgcn_readfirstlane(t / NL);
  else return t / NL;
}

// Line groups of 32 lines in 64-lane waves (NT a multiple of 64): a wave holds group g0
// in lanes 0-31 and g0 + 1 in lanes 32-63, g0 wave-uniform, so a per-group constant comes
// from two scalar loads and a select instead of one vector load per lane.
template <int NL, int NT>
constexpr bool half_wave_groups() { return NL == 32 && NT % 64 == 0; }

// Group of task t: g0 (wave-uniform when the groups allow) plus hi.
template <int NL, int NT>
__device__ __forceinline__ int group_pair(int t, bool& hi) {
  if constexpr (half_wave_groups<NL, NT>()) {
    hi = (t & 32) != 0;
    return __builtin_amdgcn_readfirstlane(t >> 6) * 2;
  } else {
    hi = false;
    return group_of<NL, NT>(t);
  }
}

// a[hi ? i1 : i0] with both indices loaded (uniform indices become scalar loads)
template <class T>
__device__ __forceinline__ T pick2(const T* __restrict__ a, int i0, int i1, bool hi) {
  const T x = a[i0];
  const T y = a[i1];
  return hi ? y : x;
}

// Forward FFT, pointwise multiply, inverse FFT on NL lines in LDS, with the
// last forward step, the multiply and the first inverse step fused in
// registers (they act on the same R2 values of one task).  mul(l, p0, p1, hi, v)
// returns v times the factor for line l at position p = hi ? p1 : p0 (forward output
// order; p0 and p1 are wave-uniform when the line groups allow).
// Result: the circular convolution, unnormalised (N times the usual inverse).
template <int N, int NL, int NT, class Addr, class Mul>
__device__ __forceinline__ void line_fft_mul_ifft(float2* __restrict__ lds, const Addr& addr,
                                                  const float2* __restrict__ tw, int tid, const Mul& mul) {
  using P = LineFft<N>;
  constexpr int R1 = P::R1, R2 = P::R2;
  if constexpr (R2 == 1) {
    for (int l = tid; l < NL; l += NT) {
      float2 v[N];
#pragma unroll
      for (int e = 0; e < N; ++e) v[e] = lds[addr(l, e)];
      Dft<N, false>::run(v);
#pragma unroll
      for (int e = 0; e < N; ++e) v[e] = mul(l, e, e, false, v[e]);
      Dft<N, true>::run(v);
#pragma unroll
      for (int e = 0; e < N; ++e) lds[addr(l, e)] = v[e];
    }
    __syncthreads();
  } else {
    for (int t = tid; t < NL * R2; t += NT) {           // forward step A
      bool hi;
      const int l = t % NL, g0 = group_pair<NL, NT>(t, hi), n2 = g0 + hi;
      float2 v[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) v[n1] = lds[addr(l, R2 * n1 + n2)];
      Dft<R1, false>::run(v);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], pick2(tw, g0 * k1, (g0 + 1) * k1, hi));
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[addr(l, R2 * k1 + n2)] = v[k1];
    }
    __syncthreads();
    for (int t = tid; t < NL * R1; t += NT) {           // step B, multiply, inverse step B
      bool hi;
      const int l = t % NL, g0 = group_pair<NL, NT>(t, hi), k1 = g0 + hi;
      float2 v[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) v[n2] = lds[addr(l, R2 * k1 + n2)];
      Dft<R2, false>::run(v);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[k2] = mul(l, R2 * g0 + k2, R2 * (g0 + 1) + k2, hi, v[k2]);
      Dft<R2, true>::run(v);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) lds[addr(l, R2 * k1 + n2)] = v[n2];
    }
    __syncthreads();
    for (int t = tid; t < NL * R2; t += NT) {           // inverse step A
      bool hi;
      const int l = t % NL, g0 = group_pair<NL, NT>(t, hi), n2 = g0 + hi;
      float2 v[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[k1] = lds[addr(l, R2 * k1 + n2)];
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmulc(v[k1], pick2(tw, g0 * k1, (g0 + 1) * k1, hi));
      Dft<R1, true>::run(v);
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) lds[addr(l, R2 * n1 + n2)] = v[n1];
    }
    __syncthreads();
  }
}

// Three-component variant for the dyadic Green's multiply: NL lines per
// component, component c's line l at index c * NL + l of addr().  The forward
// step A and inverse step A run over all 3 NL lines; step B, the multiply and
// the inverse step B are fused per (l, k1) so that mul3(l, p, v0, v1, v2) sees
// the three components of one frequency together (in place on v0..v2).
template <int N, int NL, int NT, class Addr, class Mul3>
__device__ __forceinline__ void line_fft_mul3_ifft(float2* __restrict__ lds, const Addr& addr,
                                                   const float2* __restrict__ tw, int tid, const Mul3& mul3) {
  using P = LineFft<N>;
  constexpr int R1 = P::R1, R2 = P::R2;
  if constexpr (R2 == 1) {
    for (int l = tid; l < NL; l += NT) {
      float2 v0[N], v1[N], v2[N];
#pragma unroll
      for (int e = 0; e < N; ++e) {
        v0[e] = lds[addr(l, e)];
        v1[e] = lds[addr(NL + l, e)];
        v2[e] = lds[addr(2 * NL + l, e)];
      }
      Dft<N, false>::run(v0);
      Dft<N, false>::run(v1);
      Dft<N, false>::run(v2);
#pragma unroll
      for (int e = 0; e < N; ++e) mul3(l, e, v0[e], v1[e], v2[e]);
      Dft<N, true>::run(v0);
      Dft<N, true>::run(v1);
      Dft<N, true>::run(v2);
#pragma unroll
      for (int e = 0; e < N; ++e) {
        lds[addr(l, e)] = v0[e];
        lds[addr(NL + l, e)] = v1[e];
        lds[addr(2 * NL + l, e)] = v2[e];
      }
    }
    __syncthreads();
  } else {
    for (int t = tid; t < 3 * NL * R2; t += NT) {       // forward step A, all components
      const int l = t % (3 * NL), n2 = group_of<3 * NL, NT>(t);
      float2 v[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) v[n1] = lds[addr(l, R2 * n1 + n2)];
      Dft<R1, false>::run(v);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], tw[n2 * k1]);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[addr(l, R2 * k1 + n2)] = v[k1];
    }
    __syncthreads();
    for (int t = tid; t < NL * R1; t += NT) {           // step B, 3x3 multiply, inverse step B
      const int l = t % NL, k1 = group_of<NL, NT>(t);
      float2 v0[R2], v1[R2], v2[R2];
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        v0[n2] = lds[addr(l, R2 * k1 + n2)];
        v1[n2] = lds[addr(NL + l, R2 * k1 + n2)];
        v2[n2] = lds[addr(2 * NL + l, R2 * k1 + n2)];
      }
      Dft<R2, false>::run(v0);
      Dft<R2, false>::run(v1);
      Dft<R2, false>::run(v2);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) mul3(l, R2 * k1 + k2, v0[k2], v1[k2], v2[k2]);
      Dft<R2, true>::run(v0);
      Dft<R2, true>::run(v1);
      Dft<R2, true>::run(v2);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) {
        lds[addr(l, R2 * k1 + n2)] = v0[n2];
        lds[addr(NL + l, R2 * k1 + n2)] = v1[n2];
        lds[addr(2 * NL + l, R2 * k1 + n2)] = v2[n2];
      }
    }
    __syncthreads();
    for (int t = tid; t < 3 * NL * R2; t += NT) {       // inverse step A, all components
      const int l = t % (3 * NL), n2 = group_of<3 * NL, NT>(t);
      float2 v[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[k1] = lds[addr(l, R2 * k1 + n2)];
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmulc(v[k1], tw[n2 * k1]);
      Dft<R1, true>::run(v);
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) lds[addr(l, R2 * n1 + n2)] = v[n1];
    }
    __syncthreads();
  }
}

}  // namespace cbs
