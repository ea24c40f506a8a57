// Fused two-pass Born iteration and the unfused (plain) per-axis form.
// Templates are instantiated per supported line length in passes_*.hip.
#pragma once
#include "cbs_kernels.hpp"

namespace cbs {

// LDS row pitch: odd (in 8-byte words) so that lanes walking down a column of
// rows hit distinct banks.
__host__ __device__ constexpr int row_pitch(int n) { return (n % 2 == 0) ? n + 1 : n; }

// ------------------------------------------------------------------ pass_real
// One block per (rhs, y2, z).  Tile = A rows (y1) of NX points.
// Global traffic is issued in bulk: each thread loads its share of the tile as
// 16-byte pairs before touching LDS, and the pointwise step loads V, f and Field
// for CH pairs at a time.  Twiddles and the y/z ramp and mask factors live in LDS;
// the x ramp and mask pairs are read through the cache, which keeps the 40 x 160
// tile at three blocks per CU.
struct PassRArgs {
  Tables tb;
  Consts c;
  StateArgs st;
  int upd, fwd, jj_inv, jj_fwd;
  int order;   // block order: 0 = rhs fastest (V rows reused by consecutive blocks),
               // 1 = y2 fastest, then rhs (the b blocks of one (rhs, z) slab run together),
               // 2 = y2 fastest, then z, then rhs (the tiles of one rhs in storage order, so
               // the blocks in flight stream one contiguous window of each state array)
  int nt;      // 1: nontemporal access to the state arrays (they exceed the caches)
  ConvArgs cv; // convergence test (cv.norms = null: off)
};

// ROWS = 0: the tile is the y coset y = y1 * b + y2 and the pass does the a-point
// part of the y transform (pass_kspace applies the cross twiddles); V, Field and f store
// each z plane's rows in tile order (row y at y2 * A + y1, state_row), so a tile is one
// contiguous run of every array.  ROWS = 1 (four-pass form for grids without a two-pass
// plan): the tile is A contiguous rows y = y2 * A + y1 of the natural layout, only the x
// transforms are done here, and the spectrum stays in the natural [z][y][x] layout for
// the separate y and z passes.  Either way pair i of the tile sits at 2 i past the
// tile's first element in every array.
//
// STEADY = 1: the launch is known to be a steady-state one (update UPD_DEFER or
// UPD_PAIR, next p = gamma x), so the pointwise step compiles to straight-line code;
// STEADY = 0 takes every update kind at run time.
// Waves per SIMD to hold pass_real's register budget to (1: no bound).  The 160 x 40
// tile's 53.6 KB of LDS lets three 512-thread blocks share a CU (6 waves per SIMD, at most
// 80 VGPRs); its steady pass fits 80 without spilling, and left alone the compiler took
// 81 and with them a third of the occupancy.  Other shapes either reach their LDS limit
// unbounded or would spill under the bound (96 x 48 and 160 x 32 at 512 threads).
constexpr int pass_real_min_waves(int nx, int a, int nt) { return nx == 160 && a == 40 && nt == 512 ? 6 : 1; }

//
// VR = 1 (fused form, real RI): V is read as its real part from st.Vr, 4 bytes a voxel,
// with the imaginary part -eps that V holds at every voxel for a real RI.
template <int NX, int A, int NT, int ROWS, int STEADY, int VR>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(pass_real_min_waves(NX, A, NT))))
void pass_real(PassRArgs args) {
  static_assert(NX % 2 == 0, "pass_real processes x in pairs");
  constexpr int PR = row_pitch(NX);
  constexpr int E2 = A * NX / 2;              // complex pairs in the tile
  constexpr int L = (E2 + NT - 1) / NT;       // pairs per thread
  constexpr int CH = 1;                       // pointwise chunk (pairs): one keeps the VGPRs at three blocks per CU
  __shared__ float2 lds[A * PR];
  __shared__ float2 s_twx[NX];
  __shared__ float2 s_twa[A];
  __shared__ float2 s_ryz[A];
  __shared__ float s_myz[A];

  const Consts& c = args.c;
  const Tables& tb = args.tb;
  const int tid = threadIdx.x;
  const int bid = blockIdx.x;
  int r, y2, z;
  if (args.order == 0) {
    r = bid % c.nrhs;
    const int tile = bid / c.nrhs;
    y2 = tile % c.b;
    z = tile / c.b;
  } else if (args.order == 1) {
    y2 = bid % c.b;
    const int t = bid / c.b;
    r = t % c.nrhs;
    z = t / c.nrhs;
  } else {
    y2 = bid % c.b;
    const int t = bid / c.b;
    z = t % c.nz;
    r = t / c.nz;
  }
  const long long roff = (long long)r * args.st.rhs_stride;
  float4* __restrict__ spec4 = reinterpret_cast<float4*>(args.st.spec + roff + (long long)(z * c.b + y2) * (A * NX));
  const int upd = STEADY ? (args.upd == UPD_PAIR ? UPD_PAIR : UPD_DEFER) : args.upd;
  const int fwd = STEADY ? FWD_GAMMA : args.fwd;
  const float sg_inv = sign_inv(args.jj_inv ? flip_of(args.jj_inv, args.st.parity, r) : 0);
  const float sg_fwd = sign_fwd(args.jj_fwd ? flip_of(args.jj_fwd, args.st.parity, r) : 0);
  // block-uniform base of this z plane; per-lane offsets inside a plane fit in 32 bits
  const long long zoff = (long long)z * c.ny * NX;
  // byte offset inside the z plane of pair i of the tile (the tile's rows are stored together)
  const unsigned tile_off = (unsigned)y2 * (unsigned)(A * NX) * 8u;
  auto pair_off = [&](int i) { return tile_off + 16u * (unsigned)i; };

  if (args.cv.norms && args.jj_inv > 0) {
    // convergence test: this slot's rhs stopped after an earlier iteration
    const int rr = r / args.cv.P;
    if (conv_stopped(args.cv, args.jj_inv - 1, rr)) {
      if (!conv_stopped(args.cv, args.jj_inv - 2, rr) && upd_kind(args.jj_inv - 1, args.cv.M) == UPD_DEFER) {
        // the last iteration run deferred its increment: Field += f
        float2* __restrict__ F = args.st.field + roff + zoff;
        const float2* __restrict__ f = args.st.fstate + roff + zoff;
        for (int i = tid; i < E2; i += NT) {
          const unsigned g = pair_off(i);
          const float4 a = *byte_at<const float4>(F, g), b = *byte_at<const float4>(f, g);
          *byte_at<float4>(F, g) = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
        }
      }
      return;
    }
  }

  // bulk load of the spectrum tile (all loads in flight before LDS writes)
  float4 reg[L];
  if (upd != UPD_NONE) {
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int i = tid + k * NT;
      if (i < E2) reg[k] = load_stream(spec4 + i, args.nt);
    }
  }
  const bool rf = reads_f(upd, fwd), rF = reads_F(upd);
  const float2* __restrict__ V = args.st.V + zoff;
  const float* __restrict__ Vr = args.st.Vr + zoff;
  // the potential at the pair whose state byte offset is g
  auto load_v = [&](unsigned g) {
    if constexpr (VR) {
      const float2 q = *byte_at<const float2>(Vr, g >> 1);
      return make_float4(q.x, -c.eps, q.y, -c.eps);
    } else {
      return *byte_at<const float4>(V, g);
    }
  };
  float2* __restrict__ F = args.st.field + roff + zoff;
  float2* __restrict__ f = args.st.fstate + roff + zoff;
  // small tiles (few blocks, e.g. a single rhs): the pointwise operands are requested with
  // the tile, so their latency overlaps the inverse transforms instead of following them
  constexpr bool PF = L <= 5;
  constexpr int LP = PF ? L : 1;
  float4 pv[LP], px[LP], pF[LP];
  if constexpr (PF) {
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int i = tid + k * NT;
      if (i < E2) {
        const unsigned g = pair_off(i);
        pv[k] = load_v(g);
        if (rf) px[k] = load_stream(byte_at<const float4>(f, g), args.nt);
        if (rF) pF[k] = load_stream(byte_at<const float4>(F, g), args.nt);
      }
    }
  }
  for (int i = tid; i < NX; i += NT) {
    s_twx[i] = tb.tw_x[i];
  }
  for (int i = tid; i < A; i += NT) {
    s_twa[i] = tb.tw_a[i];
    const int y = ROWS ? y2 * A + i : i * c.b + y2;
    s_ryz[i] = cmul(tb.ramp_y[y], tb.ramp_z[z]);
    s_myz[i] = tb.mask_y[y] * tb.mask_z[z];
  }
  __syncthreads();

  auto row = [](int l, int e) { return l * PR + e; };
  auto col = [](int l, int e) { return e * PR + l; };

  if (upd != UPD_NONE) {
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int i = tid + k * NT;
      if (i < E2) {
        const int e = 2 * i;
        const int p1 = e / NX, px = e - p1 * NX;
        lds[p1 * PR + px] = lo2(reg[k]);
        lds[p1 * PR + px + 1] = hi2(reg[k]);
      }
    }
    __syncthreads();
    if constexpr (A > 1 && !ROWS) line_fft<A, true, NX, NT>(lds, col, s_twa, tid);
    line_fft<NX, true, A, NT>(lds, row, s_twx, tid);
  }

  // pointwise update, CH pairs per thread at a time
  const bool wF = writes_F(upd), wf = fwd == FWD_GAMMA;
  const bool want_norm = args.st.norms != nullptr;
  // |F_1 + F_2|^2 for the convergence test, formed by the jj = 2 launch
  const bool want_fref = !STEADY && args.st.fref != nullptr && upd == UPD_VQ_ADD;
  double nrm = 0.0, nrmF = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < L; k0 += CH) {
    float4 vv[CH], xx[CH], FF[CH], rx[CH];
    float2 mx[CH];
    unsigned gi[CH];     // byte offsets inside the z plane (< 2^32): base + 32-bit offset addressing
    bool ok[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      const int i = tid + (k0 + k) * NT;
      ok[k] = k0 + k < L && i < E2;
      gi[k] = 0;
      if (ok[k]) {
        const unsigned e = 2u * (unsigned)i;
        const unsigned x = e - (e / NX) * NX;
        const unsigned g = pair_off(i);
        gi[k] = g;
        // per-axis x ramp and mask pairs (read through the cache: the LDS holds three tiles per CU)
        rx[k] = *byte_at<const float4>(tb.ramp_x, x * 8u);
        mx[k] = *byte_at<const float2>(tb.mask_x, x * 4u);
        if constexpr (PF) {
          vv[k] = pv[k0 + k];
          if (rf) xx[k] = px[k0 + k];
          if (rF) FF[k] = pF[k0 + k];
        } else {
          vv[k] = load_v(g);
          if (rf) xx[k] = load_stream(byte_at<const float4>(f, g), args.nt);
          if (rF) FF[k] = load_stream(byte_at<const float4>(F, g), args.nt);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < CH; ++k) {
      if (!ok[k]) continue;
      const int i = tid + (k0 + k) * NT;
      const int e = 2 * i;
      const int y1 = e / NX, x = e - y1 * NX;
      const float2 ryz = s_ryz[y1];
      const float myz = s_myz[y1];
      const float4 xo4 = rf ? xx[k] : make_float4(0.f, 0.f, 0.f, 0.f);
      const float4 Fo4 = rF ? FF[k] : make_float4(0.f, 0.f, 0.f, 0.f);
      float2 Fn0, Fn1, xn0, xn1;
      const float2 q0 = upd != UPD_NONE ? lds[y1 * PR + x] : make_float2(0.f, 0.f);
      const float2 q1 = upd != UPD_NONE ? lds[y1 * PR + x + 1] : make_float2(0.f, 0.f);
      const float2 p0 = born_step(upd, fwd, q0, cmul(lo2(rx[k]), ryz), mx[k].x * myz, lo2(vv[k]), lo2(xo4), lo2(Fo4),
                                  c.inv_eps, sg_inv, sg_fwd, Fn0, xn0, want_norm, nrm);
      const float2 p1 = born_step(upd, fwd, q1, cmul(hi2(rx[k]), ryz), mx[k].y * myz, hi2(vv[k]), hi2(xo4),
                                  hi2(Fo4), c.inv_eps, sg_inv, sg_fwd, Fn1, xn1, want_norm, nrm);
      if (want_fref)
        nrmF += (double)Fn0.x * Fn0.x + (double)Fn0.y * Fn0.y + (double)Fn1.x * Fn1.x + (double)Fn1.y * Fn1.y;
      if (wF) store_stream(byte_at<float4>(F, gi[k]), pack2(Fn0, Fn1), args.nt);
      if (wf) store_stream(byte_at<float4>(f, gi[k]), pack2(xn0, xn1), args.nt);
      if (fwd != FWD_NONE) {
        lds[y1 * PR + x] = p0;
        lds[y1 * PR + x + 1] = p1;
      }
    }
  }
  if (want_norm) flush_norm(nrm, args.st.norms, r);
  if (want_fref) flush_norm(nrmF, args.st.fref, r);
  if (fwd == FWD_NONE) return;
  __syncthreads();
  line_fft<NX, false, A, NT>(lds, row, s_twx, tid);
  if constexpr (A > 1 && !ROWS) line_fft<A, false, NX, NT>(lds, col, s_twa, tid);
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const int i = tid + k * NT;
    if (i < E2) {
      const int e = 2 * i;
      const int p1 = e / NX, px = e - p1 * NX;
      store_stream(spec4 + i, pack2(lds[p1 * PR + px], lds[p1 * PR + px + 1]), args.nt);
    }
  }
}

// ------------------------------------------------------------------ pass_kspace
// One block per (rhs, p1, x tile of T).  Tile = T x-positions * B (y2) * NZ.
// Loads are 16-byte pairs of x positions, all issued before the b-DFTs.
struct PassKArgs {
  Tables tb;
  Consts c;
  float2* spec;
  const uint8_t* parity;
  long long rhs_stride;
  int jj;
  int order;   // block order: 0 = rhs fastest, then a-index, then x tile;
               // 1 = x tile fastest, then a-index, then rhs (whole rows read together)
  int nt;      // 1: nontemporal stores of the spectrum (it exceeds the caches)
  int ntl;     // 1: nontemporal loads of the spectrum as well
  ConvArgs cv; // convergence test (cv.norms = null: off): a stopped rhs is skipped
};

// P = 1: scalar Green's function.  P = 3: vector run; the block holds the three
// field components of one rhs (stored as consecutive rhs slots 3r, 3r+1, 3r+2)
// and applies the dyadic Green's function G (I - c r r^T) of
// Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:607-615 at every frequency.
template <int B, int NZ, int T, int NT, int P>
__global__ __launch_bounds__(NT) void pass_kspace(PassKArgs args) {
  static_assert(T % 2 == 0, "x tile must hold pairs");
  static_assert(P == 1 || P == 3, "one or three field components");
  constexpr int LW = B * T;                   // lines of the z-FFT per component
  constexpr int TP = T / 2;                   // pairs per z row
  constexpr int NTASK = NZ * TP;              // (z, pair) tasks
  constexpr int L = (NTASK + NT - 1) / NT;
  constexpr int PLW = P * LW;                 // LDS row width (all components)
  // scalar runs: the tile alone fills the LDS budget of two blocks per CU, so the
  // small per-axis tables are read through the caches.  Vector runs: the tile is
  // three times larger and a CU holds one block anyway, so the z twiddles and the
  // z Green's and direction factors of this block's flip state sit after it.
  constexpr int ZTAB = P == 3 ? 2 * NZ : 0;
  __shared__ float2 lds[NZ * PLW + ZTAB];
  const Consts& c = args.c;
  const Tables& tb = args.tb;
  const int tid = threadIdx.x;
  const int bid = blockIdx.x;
  int r, p1, xt;
  if (args.order == 0) {
    r = bid % c.nrhs;
    const int t2 = bid / c.nrhs;
    p1 = t2 % c.a;
    xt = t2 / c.a;
  } else {
    const int ntile = (c.nx + T - 1) / T;
    xt = bid % ntile;
    const int t2 = bid / ntile;
    p1 = t2 % c.a;
    r = t2 / c.a;
  }
  if (args.cv.norms && conv_stopped(args.cv, args.jj - 1, r)) return;
  const int x0 = xt * T;
  const long long rowstride = (long long)c.a * c.nx;     // between consecutive y2
  float2* __restrict__ spec = args.spec + (long long)(P * r) * args.rhs_stride + (long long)p1 * c.nx + x0;
  const int flip = flip_of(args.jj, args.parity, P * r);
  const int tvalid = c.nx - x0 < T ? c.nx - x0 : T;      // even: nx and x0 are even

  if constexpr (P == 3) {
    for (int i = tid; i < NZ; i += NT) {
      lds[NZ * PLW + i] = tb.tw_z[i];
      lds[NZ * PLW + NZ + i] = make_float2(tb.green_z[flip * c.nz + i], tb.rad_z[flip * c.nz + i]);
    }
  }
  // cross twiddles W_ny^(y2 k1) of the four-step y transform for this block's a-position
  // (y2 = 0 needs none); applied before the forward b-DFT and, conjugated, after the inverse
  float2 wcr[B];
#pragma unroll
  for (int y2 = 0; y2 < B; ++y2) wcr[y2] = y2 ? tb.cross[y2 * c.a + p1] : make_float2(1.f, 0.f);

  float4 reg[L][P][B];
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const int task = tid + k * NT;
    if (task < NTASK) {
      const int tp = task % TP, z = task / TP;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        const float2* src = spec + (long long)q * args.rhs_stride + (long long)z * B * rowstride + 2 * tp;
#pragma unroll
        for (int y2 = 0; y2 < B; ++y2)
          reg[k][q][y2] = (2 * tp < tvalid) ? load_stream(reinterpret_cast<const float4*>(src + y2 * rowstride),
                                                          args.ntl)
                                            : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const int task = tid + k * NT;
    if (task < NTASK) {
      const int tp = task % TP, z = task / TP;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        float2 u0[B], u1[B];
#pragma unroll
        for (int y2 = 0; y2 < B; ++y2) {
          u0[y2] = lo2(reg[k][q][y2]);
          u1[y2] = hi2(reg[k][q][y2]);
          if (y2) {
            u0[y2] = cmul(u0[y2], wcr[y2]);
            u1[y2] = cmul(u1[y2], wcr[y2]);
          }
        }
        Dft<B, false>::run(u0);
        Dft<B, false>::run(u1);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
          lds[z * PLW + q * LW + k2 * T + 2 * tp] = u0[k2];
          lds[z * PLW + q * LW + k2 * T + 2 * tp + 1] = u1[k2];
        }
      }
    }
  }
  __syncthreads();
  auto zl = [](int l, int e) { return e * PLW + l; };
  {
    // (1/N) / (d - i eps) = (d + i eps) / (N (d^2 + eps^2)),  d = 4 pi^2 |k_s|^2 - k_m^2,
    // applied between the forward and inverse z transforms of each line
    const float* gz = tb.green_z + flip * c.nz;
    const float* gx = tb.green_x + flip * c.nx + x0;
    const float* gy = tb.green_y + flip * c.ny + p1;
    const float km2 = c.km2, eps2 = c.eps * c.eps, inv_n = c.inv_n, eps = c.eps;
    if constexpr (P == 1) {
      auto green = [&](int l, int pz0, int pz1, bool hi, float2 v) {
        const int k2 = l / T, t = l - k2 * T;
        const float d = gx[t < tvalid ? t : 0] + gy[k2 * c.a] + pick2(gz, pz0, pz1, hi) - km2;
        const float s = inv_n * __builtin_amdgcn_rcpf(fmaf(d, d, eps2));   // hardware reciprocal (1 ulp)
        return cmul(v, make_float2(d * s, eps * s));
      };
      line_fft_mul_ifft<NZ, LW, NT>(lds, zl, tb.tw_z, tid, green);
    } else {
      const float2* gr_z = lds + NZ * PLW + NZ;                // (green_z, rad_z) of this flip state
      const float* rx = tb.rad_x + flip * c.nx + x0;
      const float* ry = tb.rad_y + flip * c.ny + p1;
      const float2 cc = make_float2(c.c_re, c.c_im);
      // PSI_j = G (v_j - c r_j (r . v))  ==  sum_i G (delta_ji - c r_j r_i) v_i
      auto dyadic = [&](int l, int pz, float2& v0, float2& v1, float2& v2) {
        const int k2 = l / T, t = l - k2 * T;
        const int tx = t < tvalid ? t : 0;
        const float2 grz = gr_z[pz];
        const float d = gx[tx] + gy[k2 * c.a] + grz.x - km2;
        const float s = inv_n * __builtin_amdgcn_rcpf(fmaf(d, d, eps2));   // hardware reciprocal (1 ulp)
        const float2 g = make_float2(d * s, eps * s);
        const float r0 = rx[tx], r1 = ry[k2 * c.a], r2 = grz.y;
        const float2 dot = make_float2(r0 * v0.x + r1 * v1.x + r2 * v2.x, r0 * v0.y + r1 * v1.y + r2 * v2.y);
        const float2 cd = cmul(cc, dot);
        v0 = cmul(g, make_float2(v0.x - r0 * cd.x, v0.y - r0 * cd.y));
        v1 = cmul(g, make_float2(v1.x - r1 * cd.x, v1.y - r1 * cd.y));
        v2 = cmul(g, make_float2(v2.x - r2 * cd.x, v2.y - r2 * cd.y));
      };
      line_fft_mul3_ifft<NZ, LW, NT>(lds, zl, lds + NZ * PLW, tid, dyadic);
    }
  }
#pragma unroll
  for (int k = 0; k < L; ++k) {
    const int task = tid + k * NT;
    if (task < NTASK) {
      const int tp = task % TP, z = task / TP;
      if (2 * tp >= tvalid) continue;
#pragma unroll
      for (int q = 0; q < P; ++q) {
        float2 u0[B], u1[B];
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
          u0[k2] = lds[z * PLW + q * LW + k2 * T + 2 * tp];
          u1[k2] = lds[z * PLW + q * LW + k2 * T + 2 * tp + 1];
        }
        Dft<B, true>::run(u0);
        Dft<B, true>::run(u1);
        float2* dst = spec + (long long)q * args.rhs_stride + (long long)z * B * rowstride + 2 * tp;
#pragma unroll
        for (int y2 = 0; y2 < B; ++y2) {
          if (y2) {
            u0[y2] = cmulc(u0[y2], wcr[y2]);
            u1[y2] = cmulc(u1[y2], wcr[y2]);
          }
          store_stream(reinterpret_cast<float4*>(dst + y2 * rowstride), pack2(u0[y2], u1[y2]), args.nt);
        }
      }
    }
  }
}

// ------------------------------------------------------------------ pass_kspace, vector runs, one component at a time
// The same work as pass_kspace<B, NZ, T, NT, 3>, with one component's T x B x NZ tile in LDS
// at a time: each component is loaded, b-transformed and taken through the forward z
// steps A and B; its step-B output stays in registers, so after the third component every
// thread holds the three components of its frequencies and applies the dyadic factor
// there.  The inverse steps then run component by component and store.  The LDS is a third
// of the all-component tile, so an x tile of 16 (128-byte rows) fits two blocks per CU.
// Every value goes through the same operations in the same order as in pass_kspace.
template <int B, int NZ, int T, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(B * T * LineFft<NZ>::R1 == NT ? 4 : 1)))
void pass_kspace_seq3(PassKArgs args) {
  using LF = LineFft<NZ>;
  constexpr int R1 = LF::R1, R2 = LF::R2;
  static_assert(R2 > 1, "two-step z transform");
  static_assert(T % 2 == 0, "x tile must hold pairs");
  constexpr int LW = B * T;                   // z lines of one component
  constexpr int TP = T / 2;
  constexpr int NTASK = NZ * TP;
  constexpr int L = (NTASK + NT - 1) / NT;
  static_assert((LW * R1) % NT == 0, "step-B groups split evenly over the block");
  constexpr int G = LW * R1 / NT;             // step-B groups per thread
  __shared__ float2 lds[NZ * LW];
  const Consts& c = args.c;
  const Tables& tb = args.tb;
  const int tid = threadIdx.x;
  const int bid = blockIdx.x;
  int r, p1, xt;
  if (args.order == 0) {
    r = bid % c.nrhs;
    const int t2 = bid / c.nrhs;
    p1 = t2 % c.a;
    xt = t2 / c.a;
  } else {
    const int ntile = (c.nx + T - 1) / T;
    xt = bid % ntile;
    const int t2 = bid / ntile;
    p1 = t2 % c.a;
    r = t2 / c.a;
  }
  if (args.cv.norms && conv_stopped(args.cv, args.jj - 1, r)) return;
  const int x0 = xt * T;
  const long long rowstride = (long long)c.a * c.nx;
  float2* __restrict__ spec = args.spec + (long long)(3 * r) * args.rhs_stride + (long long)p1 * c.nx + x0;
  const int flip = flip_of(args.jj, args.parity, 3 * r);
  const int tvalid = c.nx - x0 < T ? c.nx - x0 : T;
  float2 wcr[B];
#pragma unroll
  for (int y2 = 0; y2 < B; ++y2) wcr[y2] = y2 ? tb.cross[y2 * c.a + p1] : make_float2(1.f, 0.f);
  auto zl = [](int l, int e) { return e * LW + l; };
  const float2* __restrict__ tw = tb.tw_z;
  float2 acc[G][3][R2];

#pragma unroll
  for (int q = 0; q < 3; ++q) {
    float4 reg[L][B];
    const float2* srcq = spec + (long long)q * args.rhs_stride;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int task = tid + k * NT;
      if (task < NTASK) {
        const int tp = task % TP, z = task / TP;
        const float2* src = srcq + (long long)z * B * rowstride + 2 * tp;
#pragma unroll
        for (int y2 = 0; y2 < B; ++y2)
          reg[k][y2] = (2 * tp < tvalid) ? load_stream(reinterpret_cast<const float4*>(src + y2 * rowstride), args.ntl)
                                         : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int task = tid + k * NT;
      if (task < NTASK) {
        const int tp = task % TP, z = task / TP;
        float2 u0[B], u1[B];
#pragma unroll
        for (int y2 = 0; y2 < B; ++y2) {
          u0[y2] = lo2(reg[k][y2]);
          u1[y2] = hi2(reg[k][y2]);
          if (y2) {
            u0[y2] = cmul(u0[y2], wcr[y2]);
            u1[y2] = cmul(u1[y2], wcr[y2]);
          }
        }
        Dft<B, false>::run(u0);
        Dft<B, false>::run(u1);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
          lds[z * LW + k2 * T + 2 * tp] = u0[k2];
          lds[z * LW + k2 * T + 2 * tp + 1] = u1[k2];
        }
      }
    }
    __syncthreads();
    for (int t = tid; t < LW * R2; t += NT) {            // forward step A
      const int l = t % LW, n2 = group_of<LW, NT>(t);
      float2 v[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) v[n1] = lds[zl(l, R2 * n1 + n2)];
      Dft<R1, false>::run(v);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], tw[n2 * k1]);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) lds[zl(l, R2 * k1 + n2)] = v[k1];
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g) {                         // forward step B, kept in registers
      const int t = tid + g * NT;
      const int l = t % LW, k1 = group_of<LW, NT>(t);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) acc[g][q][n2] = lds[zl(l, R2 * k1 + n2)];
      Dft<R2, false>::run(acc[g][q]);
    }
    __syncthreads();
  }
  {
    // G (v - c r (r . v)) per frequency, as in pass_kspace
    const float* gz = tb.green_z + flip * c.nz;
    const float* rz = tb.rad_z + flip * c.nz;
    const float* gx = tb.green_x + flip * c.nx + x0;
    const float* gy = tb.green_y + flip * c.ny + p1;
    const float* rx = tb.rad_x + flip * c.nx + x0;
    const float* ry = tb.rad_y + flip * c.ny + p1;
    const float km2 = c.km2, eps2 = c.eps * c.eps, inv_n = c.inv_n, eps = c.eps;
    const float2 cc = make_float2(c.c_re, c.c_im);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int t = tid + g * NT;
      const int l = t % LW, k1 = group_of<LW, NT>(t);
      const int k2y = l / T, tt = l - k2y * T;
      const int tx = tt < tvalid ? tt : 0;
      const float r0 = rx[tx], r1 = ry[k2y * c.a];
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        const int pz = R2 * k1 + k2;
        float2& v0 = acc[g][0][k2];
        float2& v1 = acc[g][1][k2];
        float2& v2 = acc[g][2][k2];
        const float d = gx[tx] + gy[k2y * c.a] + gz[pz] - km2;
        const float s = inv_n * __builtin_amdgcn_rcpf(fmaf(d, d, eps2));   // hardware reciprocal (1 ulp)
        const float2 gg = make_float2(d * s, eps * s);
        const float r2 = rz[pz];
        const float2 dot = make_float2(r0 * v0.x + r1 * v1.x + r2 * v2.x, r0 * v0.y + r1 * v1.y + r2 * v2.y);
        const float2 cd = cmul(cc, dot);
        v0 = cmul(gg, make_float2(v0.x - r0 * cd.x, v0.y - r0 * cd.y));
        v1 = cmul(gg, make_float2(v1.x - r1 * cd.x, v1.y - r1 * cd.y));
        v2 = cmul(gg, make_float2(v2.x - r2 * cd.x, v2.y - r2 * cd.y));
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
#pragma unroll
    for (int g = 0; g < G; ++g) {                         // inverse step B
      const int t = tid + g * NT;
      const int l = t % LW, k1 = group_of<LW, NT>(t);
      Dft<R2, true>::run(acc[g][q]);
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) lds[zl(l, R2 * k1 + n2)] = acc[g][q][n2];
    }
    __syncthreads();
    for (int t = tid; t < LW * R2; t += NT) {            // inverse step A
      const int l = t % LW, n2 = group_of<LW, NT>(t);
      float2 v[R1];
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) v[k1] = lds[zl(l, R2 * k1 + n2)];
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmulc(v[k1], tw[n2 * k1]);
      Dft<R1, true>::run(v);
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) lds[zl(l, R2 * n1 + n2)] = v[n1];
    }
    __syncthreads();
    float2* dstq = spec + (long long)q * args.rhs_stride;
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int task = tid + k * NT;
      if (task < NTASK) {
        const int tp = task % TP, z = task / TP;
        if (2 * tp < tvalid) {
          float2 u0[B], u1[B];
#pragma unroll
          for (int k2 = 0; k2 < B; ++k2) {
            u0[k2] = lds[z * LW + k2 * T + 2 * tp];
            u1[k2] = lds[z * LW + k2 * T + 2 * tp + 1];
          }
          Dft<B, true>::run(u0);
          Dft<B, true>::run(u1);
          float2* dst = dstq + (long long)z * B * rowstride + 2 * tp;
#pragma unroll
          for (int y2 = 0; y2 < B; ++y2) {
            if (y2) {
              u0[y2] = cmulc(u0[y2], wcr[y2]);
              u1[y2] = cmulc(u1[y2], wcr[y2]);
            }
            store_stream(reinterpret_cast<float4*>(dst + y2 * rowstride), pack2(u0[y2], u1[y2]), args.nt);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ plain form
// Unfused single-rhs iteration: per-axis FFT launches on a natural-layout
// spectrum, a Green's multiply launch and a pointwise update launch.

// Per-axis form of a single-component run on a compiled x length N: the inverse x
// transform of the spectrum rows, the pointwise update of update_plain (born_step) and the
// forward x transform of the next p in one launch.  Block (bx, slot) takes rows bx NL ..
// of slot `slot` (ny nz rows of N per slot, contiguous).  upd = UPD_NONE skips the inverse
// transform, fwd = FWD_NONE the forward one; a slot the convergence test has stopped
// leaves Field, f and the spectrum alone after the flush of a deferred increment.  Every
// global operand of a thread's elements is requested before the first transform.
template <int N, int NL, int NT>
__global__ __launch_bounds__(NT) void line_rows_update(StateArgs st, Tables tb, Consts c, const float2* tw, int upd,
                                                       int fwd, int jj_inv, int jj_fwd, ConvArgs cv) {
  constexpr int PR = row_pitch(N);
  constexpr int K = (NL * N + NT - 1) / NT;
  __shared__ float2 lds[NL * PR];
  const int slot = blockIdx.y;
  const int rows = c.ny * c.nz;                          // < 2^31: a slot's voxels fit a launch
  const int r0 = blockIdx.x * NL;
  const int nr = rows - r0 < NL ? rows - r0 : NL;
  const long long ro = (long long)slot * st.rhs_stride + (long long)r0 * N;
  float2* __restrict__ spec = st.spec + ro;
  float2* __restrict__ F = st.field + ro;
  float2* __restrict__ f = st.fstate + ro;
  const float2* __restrict__ V = st.V + (long long)r0 * N;
  if (cv.norms && jj_inv > 0) {
    const int rr = slot / cv.P;
    if (conv_stopped(cv, jj_inv - 1, rr)) {
      if (!conv_stopped(cv, jj_inv - 2, rr) && upd_kind(jj_inv - 1, cv.M) == UPD_DEFER)
        for (int i = threadIdx.x; i < nr * N; i += NT) F[i] = cadd(F[i], f[i]);
      return;
    }
  }
  auto row = [](int l, int e) { return l * PR + e; };
  const bool rq = upd != UPD_NONE;
  const bool rf = reads_f(upd, fwd), rF = reads_F(upd);
  float2 xq[K], vv[K], xo[K], Fo[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    xq[k] = vv[k] = xo[k] = Fo[k] = make_float2(0.f, 0.f);
    if (i < nr * N) {
      if (rq) xq[k] = spec[i];
      vv[k] = V[i];
      if (rf) xo[k] = f[i];
      if (rF) Fo[k] = F[i];
    }
  }
  if (rq) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int i = threadIdx.x + k * NT;
      const int l = i / N, e = i - l * N;
      if (i < NL * N) lds[row(l, e)] = xq[k];
    }
    __syncthreads();
    line_fft<N, true, NL, NT>(lds, row, tw, threadIdx.x);
  }
  const bool wF = writes_F(upd), wf = fwd == FWD_GAMMA;
  const bool want_norm = st.norms != nullptr;
  const bool want_fref = st.fref != nullptr && upd == UPD_VQ_ADD;
  const float sg_inv = sign_inv(jj_inv ? flip_of(jj_inv, st.parity, slot) : 0);
  const float sg_fwd = sign_fwd(jj_fwd ? flip_of(jj_fwd, st.parity, slot) : 0);
  double nrm = 0.0, nrmF = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / N, e = i - l * N;
    if (i >= NL * N) continue;
    if (l >= nr) {
      lds[row(l, e)] = make_float2(0.f, 0.f);
      continue;
    }
    const int rw = r0 + l;
    const int z = rw / c.ny, y = rw - z * c.ny;
    const float2 r = cmul(cmul(tb.ramp_x[e], tb.ramp_y[y]), tb.ramp_z[z]);
    const float m = tb.mask_x[e] * tb.mask_y[y] * tb.mask_z[z];
    const float2 q = rq ? lds[row(l, e)] : make_float2(0.f, 0.f);
    float2 Fn, xn;
    const float2 pn = born_step(upd, fwd, q, r, m, vv[k], xo[k], Fo[k], c.inv_eps, sg_inv, sg_fwd, Fn, xn,
                                want_norm, nrm);
    if (wF) F[i] = Fn;
    if (want_fref) nrmF += (double)Fn.x * Fn.x + (double)Fn.y * Fn.y;
    if (wf) f[i] = xn;
    lds[row(l, e)] = pn;
  }
  if (want_norm) flush_norm(nrm, st.norms, slot);
  if (want_fref) flush_norm(nrmF, st.fref, slot);
  if (fwd == FWD_NONE) return;
  __syncthreads();
  line_fft<N, false, NL, NT>(lds, row, tw, threadIdx.x);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int l = i / N, e = i - l * N;
    if (i < nr * N) spec[i] = lds[row(l, e)];
  }
}

// Per-axis form of a single-component run on a compiled z length N: the forward z
// transform, the Green's multiply of green_plain (at the transform's output positions, the
// order the per-axis tables are built in) and the inverse z transform in one launch.
// Tiles as fft_strided: T consecutive x positions of z lines, outer set o = y + ny slot.
template <int N, int T, int NT>
__global__ __launch_bounds__(NT) void line_strided_green(float2* data, Tables tb, Consts c, const float2* tw,
                                                         const uint8_t* __restrict__ parity, int jj, ConvArgs cv) {
  constexpr int K = (N * T + NT - 1) / NT;
  __shared__ float2 lds[N * T];
  __shared__ float2 s_tw[N];
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
  auto ad = [](int l, int e) { return e * T + l; };
  float2 xv[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    xv[k] = make_float2(0.f, 0.f);
    if (i < N * T && t < tv) xv[k] = base[e * es + t];
  }
  for (int i = threadIdx.x; i < N; i += NT) s_tw[i] = tw[i];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    if (i < N * T) lds[i] = xv[k];              // e T + t == i
  }
  __syncthreads();
  const float* gx = tb.green_x + flip * c.nx + x0;
  const float* gz = tb.green_z + flip * c.nz;
  const float gy = tb.green_y[flip * c.ny + y] - c.km2;
  const float eps2 = c.eps * c.eps, inv_n = c.inv_n, eps = c.eps;
  auto green = [&](int l, int pz0, int pz1, bool hi, float2 v) {
    const float d = gx[l < tv ? l : 0] + gy + pick2(gz, pz0, pz1, hi);
    const float s = inv_n * __builtin_amdgcn_rcpf(fmaf(d, d, eps2));   // hardware reciprocal (1 ulp)
    return cmul(v, make_float2(d * s, eps * s));
  };
  line_fft_mul_ifft<N, T, NT>(lds, ad, s_tw, threadIdx.x, green);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = threadIdx.x + k * NT;
    const int e = i / T, t = i - e * T;
    if (i < N * T && t < tv) base[e * es + t] = lds[i];
  }
}

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
// base + t + e*es.  Blocks enumerate (x tile, outer index o < nouter) with
// base = (o / inner) * os2 + (o % inner) * os + xt * T (two outer levels: e.g. the
// y positions of a z transform and the right-hand sides).
template <int N, int T, int NT>
__global__ __launch_bounds__(NT) void fft_strided(float2* data, int nx, long long es, long long os, int nouter,
                                                  const float2* tw, int inverse, int inner, long long os2) {
  __shared__ float2 lds[N * T];
  __shared__ float2 s_tw[N];   // twiddles: a task's R1 - 1 of them differ from lane to lane
  for (int i = threadIdx.x; i < N; i += NT) s_tw[i] = tw[i];
  const int ntile = (nx + T - 1) / T;
  const int xt = blockIdx.x % ntile;
  const int o = blockIdx.x / ntile;
  const int tvalid = nx - xt * T < T ? nx - xt * T : T;
  float2* base = data + (long long)(o / inner) * os2 + (long long)(o % inner) * os + (long long)xt * T;
  auto ad = [](int l, int e) { return e * T + l; };
  // even x extent (and so even strides and tile origins): 16-byte pair accesses,
  // all of a thread's loads issued before the LDS writes
  const bool pairs = (T % 2 == 0) && (nx % 2 == 0) && (es % 2 == 0) && (os % 2 == 0) && (os2 % 2 == 0);
  constexpr int TP = T / 2;
  constexpr int LP = (N * TP + NT - 1) / NT;
  if (pairs) {
    float4 reg[LP];
#pragma unroll
    for (int k = 0; k < LP; ++k) {
      const int i = threadIdx.x + k * NT;
      if (i < N * TP) {
        const int e = i / TP, tp = i - e * TP;
        reg[k] = 2 * tp < tvalid ? *reinterpret_cast<const float4*>(base + e * es + 2 * tp)
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
#pragma unroll
    for (int k = 0; k < LP; ++k) {
      const int i = threadIdx.x + k * NT;
      if (i < N * TP) {
        const int e = i / TP, tp = i - e * TP;
        lds[e * T + 2 * tp] = lo2(reg[k]);
        lds[e * T + 2 * tp + 1] = hi2(reg[k]);
      }
    }
  } else {
    for (int i = threadIdx.x; i < N * T; i += NT) {
      const int e = i / T, t = i - e * T;
      lds[i] = t < tvalid ? base[e * es + t] : make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  if (inverse) line_fft<N, true, T, NT>(lds, ad, s_tw, threadIdx.x);
  else line_fft<N, false, T, NT>(lds, ad, s_tw, threadIdx.x);
  if (pairs) {
#pragma unroll
    for (int k = 0; k < LP; ++k) {
      const int i = threadIdx.x + k * NT;
      if (i < N * TP) {
        const int e = i / TP, tp = i - e * TP;
        if (2 * tp < tvalid)
          *reinterpret_cast<float4*>(base + e * es + 2 * tp) = pack2(lds[e * T + 2 * tp], lds[e * T + 2 * tp + 1]);
      }
    }
  } else {
    for (int i = threadIdx.x; i < N * T; i += NT) {
      const int e = i / T, t = i - e * T;
      if (t < tvalid) base[e * es + t] = lds[i];
    }
  }
}

}  // namespace cbs
