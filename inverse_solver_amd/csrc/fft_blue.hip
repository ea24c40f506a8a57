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
#include "cbs_kernels.hpp"

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

// ------------------------------------------------------------------ per-axis form, fused
// Per-axis form of a single-component run whose x lines are Bluestein-transformed: the
// inverse x transform of the spectrum rows, the pointwise update of update_plain
// (born_step) and the forward x transform of the next p in one launch.  Block (bx, slot)
// takes rows bx NL .. of component slot `slot` (ny nz rows of nx per slot, contiguous).
// upd = UPD_NONE skips the inverse transform (the first launch: p from the source term);
// fwd = FWD_NONE skips the forward one (the last).  A slot that the convergence test has
// stopped leaves Field, f and the spectrum alone (after the flush of a deferred
// increment), as update_plain does.
template <int L, int NL, int NT>
__global__ __launch_bounds__(NT) void blue_rows_update(StateArgs st, Tables tb, Consts c, BlueArgs p, int upd,
                                                       int fwd, int jj_inv, int jj_fwd, ConvArgs cv) {
  constexpr int NMAX = (L + 1) / 2;
  constexpr int K = (NL * NMAX + NT - 1) / NT;
  constexpr int KZ = (NL * (L - NMAX) + NT - 1) / NT;
  __shared__ float2 lds[L * (NL + 1)];
  __shared__ float2 s_tw[L];
  __shared__ float2 s_b[L];
  const int n = p.n;
  const int slot = blockIdx.y;
  const int rows = c.ny * c.nz;                          // < 2^31: a slot's voxels fit a launch
  const int r0 = blockIdx.x * NL;
  const int nr = rows - r0 < NL ? rows - r0 : NL;
  const long long ro = (long long)slot * st.rhs_stride + (long long)r0 * n;
  float2* __restrict__ spec = st.spec + ro;
  float2* __restrict__ F = st.field + ro;
  float2* __restrict__ f = st.fstate + ro;
  const float2* __restrict__ V = st.V + (long long)r0 * n;
  if (cv.norms && jj_inv > 0) {
    const int rr = slot / cv.P;
    if (conv_stopped(cv, jj_inv - 1, rr)) {
      if (!conv_stopped(cv, jj_inv - 2, rr) && upd_kind(jj_inv - 1, cv.M) == UPD_DEFER)
        for (int i = threadIdx.x; i < nr * n; i += NT) F[i] = cadd(F[i], f[i]);
      return;
    }
  }
  const BlueTile<NL> ad;
  const bool rq = upd != UPD_NONE;
  const bool rf = reads_f(upd, fwd), rF = reads_F(upd);
  // every global operand of the block's elements is requested before the first is used:
  // the spectrum, the chirp, V, f and Field (element k of a thread: l = i / NMAX, e = i % NMAX)
  float2 xq[K], ch[K], vv[K], xo[K], Fo[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / NMAX, e = i - l * NMAX;
    xq[k] = ch[k] = vv[k] = xo[k] = Fo[k] = make_float2(0.f, 0.f);
    if (i < NL * NMAX && e < n && l < nr) {
      const int g = l * n + e;
      ch[k] = p.chirp[e];
      if (rq) xq[k] = spec[g];
      vv[k] = V[g];
      if (rf) xo[k] = f[g];
      if (rF) Fo[k] = F[g];
    }
  }
  blue_tables<L, NT>(p, s_tw, s_b);
  if (rq) {
    // inverse x transform: conj(DFT(conj q))
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = threadIdx.x + k * NT;
      const int l = i / NMAX, e = i - l * NMAX;
      if (i < NL * NMAX) lds[ad(l, e)] = cmulc(cconj(xq[k]), ch[k]);
    }
#pragma unroll
    for (int k = 0; k < KZ; ++k) {
      const int i = threadIdx.x + k * NT;
      const int l = i / (L - NMAX), e = NMAX + i - l * (L - NMAX);
      if (i < NL * (L - NMAX)) lds[ad(l, e)] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    blue_convolve<L, NL, NT>(lds, s_tw, s_b);
  } else {
    __syncthreads();   // the tables
  }
  const bool wF = writes_F(upd), wf = fwd == FWD_GAMMA;
  const bool want_norm = st.norms != nullptr;
  const bool want_fref = st.fref != nullptr && upd == UPD_VQ_ADD;
  const float sg_inv = sign_inv(jj_inv ? flip_of(jj_inv, st.parity, slot) : 0);
  const float sg_fwd = sign_fwd(jj_fwd ? flip_of(jj_fwd, st.parity, slot) : 0);
  double nrm = 0.0, nrmF = 0.0;
  // pointwise: element (l, e) is voxel x = e of row r0 + l; the inverse result is
  // c_e conj(w_e), the forward input p conj(c_e) (zero outside the lines)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / NMAX, e = i - l * NMAX;
    if (i >= NL * NMAX) continue;
    if (e >= n || l >= nr) {
      lds[ad(l, e)] = make_float2(0.f, 0.f);
      continue;
    }
    const int g = l * n + e;
    const int row = r0 + l;
    const int z = row / c.ny, y = row - z * c.ny;
    const float2 r = cmul(cmul(tb.ramp_x[e], tb.ramp_y[y]), tb.ramp_z[z]);
    const float m = tb.mask_x[e] * tb.mask_y[y] * tb.mask_z[z];
    const float2 q = rq ? cmul(ch[k], cconj(lds[ad(l, e)])) : make_float2(0.f, 0.f);
    float2 Fn, xn;
    const float2 pn = born_step(upd, fwd, q, r, m, vv[k], xo[k], Fo[k], c.inv_eps, sg_inv, sg_fwd, Fn, xn,
                                want_norm, nrm);
    if (wF) F[g] = Fn;
    if (want_fref) nrmF += (double)Fn.x * Fn.x + (double)Fn.y * Fn.y;
    if (wf) f[g] = xn;
    lds[ad(l, e)] = cmulc(pn, ch[k]);
  }
  if (want_norm) flush_norm(nrm, st.norms, slot);
  if (want_fref) flush_norm(nrmF, st.fref, slot);
  if (fwd == FWD_NONE) return;
#pragma unroll
  for (int k = 0; k < KZ; ++k) {                         // zero padding above NMAX
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
    if (i < NL * NMAX && e < n && l < nr) spec[l * n + e] = cmulc(lds[ad(l, e)], ch[k]);
  }
}

// Per-axis form of a single-component run whose z lines are Bluestein-transformed: the
// forward z transform, the Green's multiply of green_plain and the inverse z transform in
// one launch.  Tiles as blue_strided with es = nx ny and outer set o = y + ny slot.
template <int L, int T, int NT>
__global__ __launch_bounds__(NT) void blue_strided_green(float2* data, Tables tb, Consts c, BlueArgs p,
                                                         const uint8_t* __restrict__ parity, int jj, ConvArgs cv) {
  constexpr int NMAX = (L + 1) / 2;
  constexpr int K = (NMAX * T + NT - 1) / NT;
  constexpr int KZ = ((L - NMAX) * T + NT - 1) / NT;
  __shared__ float2 lds[L * (T + 1)];
  __shared__ float2 s_tw[L];
  __shared__ float2 s_b[L];
  const int n = p.n;
  const int ntile = (c.nx + T - 1) / T;
  const int xt = blockIdx.x % ntile;
  const int o = blockIdx.x / ntile;
  const int y = o % c.ny, slot = o / c.ny;      // the grid holds ceil(nx / T) ny slots blocks
  if (cv.norms && conv_stopped(cv, jj - 1, slot)) return;
  const int flip = flip_of(jj, parity, slot);
  const int x0 = xt * T;
  const int tv = c.nx - x0 < T ? c.nx - x0 : T;
  const long long es = (long long)c.nx * c.ny;
  float2* base = data + (long long)slot * c.nvox + (long long)y * c.nx + x0;
  const BlueTile<T> ad;
  float2 xv[K], chv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    xv[k] = chv[k] = make_float2(0.f, 0.f);
    if (i < NMAX * T && e < n && t < tv) {
      xv[k] = base[e * es + t];
      chv[k] = p.chirp[e];
    }
  }
  blue_tables<L, NT>(p, s_tw, s_b);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    if (i < NMAX * T) lds[ad(t, e)] = cmulc(xv[k], chv[k]);
  }
#pragma unroll
  for (int k = 0; k < KZ; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = NMAX + i / T, t = i % T;
    if (i < (L - NMAX) * T) lds[ad(t, e)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  blue_convolve<L, T, NT>(lds, s_tw, s_b);
  // X_k = conj(c_k) w_k; Y_k = G_k X_k; the inverse transform's input conj(Y_k) conj(c_k)
  const float gy = tb.green_y[flip * c.ny + y] - c.km2;
  const float eps2 = c.eps * c.eps;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    if (i >= NMAX * T) continue;
    float2 v = make_float2(0.f, 0.f);
    if (e < n && t < tv) {
      const float d = tb.green_x[flip * c.nx + x0 + t] + gy + tb.green_z[flip * c.nz + e];
      const float s = c.inv_n * __builtin_amdgcn_rcpf(fmaf(d, d, eps2));   // hardware reciprocal (1 ulp)
      const float2 X = cmulc(lds[ad(t, e)], chv[k]);
      const float2 Y = cmul(X, make_float2(d * s, c.eps * s));
      v = cmulc(cconj(Y), chv[k]);
    }
    lds[ad(t, e)] = v;
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
    if (i < NMAX * T && e < n && t < tv) base[e * es + t] = cmul(chv[k], cconj(lds[ad(t, e)]));
  }
}

// (L, lines per rows block, lines per strided tile, threads): 5-smooth lengths whose
// two-step plan has both factors <= 36
#define BLUE_LIST(X)                                                                                           \
  X(48, 16, 16, 256) X(64, 16, 16, 256) X(80, 16, 16, 256) X(96, 16, 16, 256) X(128, 16, 16, 256)              \
  X(160, 16, 16, 256) X(192, 16, 16, 256) X(256, 16, 16, 256) X(320, 16, 16, 256) X(384, 16, 16, 256)          \
  X(512, 16, 16, 256) X(640, 8, 8, 256) X(768, 8, 8, 256) X(1024, 8, 8, 256) X(1152, 8, 8, 256)                \
  X(1296, 8, 8, 256)

#define BLUE_ENTRY(L, NL, T, NT) \
  {&blue_rows<L, NL, NT>, &blue_strided<L, T, NT>, &blue_rows_update<L, NL, NT>, &blue_strided_green<L, T, NT>, \
   L, NL, T, NT, LineFft<L>::R2},
static const BlueInfo kBlue[] = {BLUE_LIST(BLUE_ENTRY)};

const BlueInfo* find_blue(int m) {
  const BlueInfo* best = nullptr;
  for (const auto& e : kBlue)
    if (e.L >= m && (!best || e.L < best->L)) best = &e;
  return best;
}

}  // namespace cbs
