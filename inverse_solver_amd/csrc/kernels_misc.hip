// Pointwise kernels: plain-form Green's multiply and update, source seeding and
// the final Field[ROI] + u_in assembly.  Grid-stride loops (the largest grids
// exceed 2^32 work items when launched one thread per element).
#include "kernels_misc.hpp"

namespace cbs {

#define GRID_STRIDE(i, total) \
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (total); i += (long long)gridDim.x * blockDim.x)

// Plain form: spectrum in natural [z][y][x] layout (each axis in LineFft
// position order).  Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:609-613 (Greenp .* fftn)
// with the 1/N of the following ifftn folded in.  blockIdx.y is the rhs; its flip
// state of iteration jj follows from its parity; a stopped rhs (cv) is skipped.
__global__ void green_plain(float2* __restrict__ spec, Tables tb, Consts c, const uint8_t* __restrict__ parity, int jj,
                            ConvArgs cv) {
  const int r = blockIdx.y;
  if (cv.norms && conv_stopped(cv, jj - 1, r)) return;
  const int flip = flip_of(jj, parity, r);
  spec += (long long)r * c.nvox;
  GRID_STRIDE(i, c.nvox) {
    const int x = (int)(i % c.nx);
    const long long yz = i / c.nx;
    const int y = (int)(yz % c.ny);
    const int z = (int)(yz / c.ny);
    const float d = tb.green_x[flip * c.nx + x] + tb.green_y[flip * c.ny + y] + tb.green_z[flip * c.nz + z] - c.km2;
    const float s = c.inv_n * __builtin_amdgcn_rcpf(fmaf(d, d, c.eps * c.eps));   // hardware reciprocal (1 ulp)
    spec[i] = cmul(spec[i], make_float2(d * s, c.eps * s));
  }
}

// Plain form, vector runs: the dyadic multiply G (I - c r r^T) on the three
// component spectra of one rhs (comp_stride elements apart; rhs blockIdx.y); :607-615.
__global__ void green_plain_vec(float2* __restrict__ spec, long long comp_stride, Tables tb, Consts c,
                                const uint8_t* __restrict__ parity, int jj, ConvArgs cv) {
  const int rr = blockIdx.y;
  if (cv.norms && conv_stopped(cv, jj - 1, rr)) return;
  const int flip = flip_of(jj, parity, 3 * rr);
  spec += (long long)3 * rr * comp_stride;
  const float2 cc = make_float2(c.c_re, c.c_im);
  GRID_STRIDE(i, c.nvox) {
    const int x = (int)(i % c.nx);
    const long long yz = i / c.nx;
    const int y = (int)(yz % c.ny);
    const int z = (int)(yz / c.ny);
    const float d = tb.green_x[flip * c.nx + x] + tb.green_y[flip * c.ny + y] + tb.green_z[flip * c.nz + z] - c.km2;
    const float s = c.inv_n * __builtin_amdgcn_rcpf(fmaf(d, d, c.eps * c.eps));   // hardware reciprocal (1 ulp)
    const float2 g = make_float2(d * s, c.eps * s);
    const float r0 = tb.rad_x[flip * c.nx + x], r1 = tb.rad_y[flip * c.ny + y], r2 = tb.rad_z[flip * c.nz + z];
    const float2 v0 = spec[i], v1 = spec[comp_stride + i], v2 = spec[2 * comp_stride + i];
    const float2 dot = make_float2(r0 * v0.x + r1 * v1.x + r2 * v2.x, r0 * v0.y + r1 * v1.y + r2 * v2.y);
    const float2 cd = cmul(cc, dot);
    spec[i] = cmul(g, make_float2(v0.x - r0 * cd.x, v0.y - r0 * cd.y));
    spec[comp_stride + i] = cmul(g, make_float2(v1.x - r1 * cd.x, v1.y - r1 * cd.y));
    spec[2 * comp_stride + i] = cmul(g, make_float2(v2.x - r2 * cd.x, v2.y - r2 * cd.y));
  }
}

// Plain form: the pointwise part of one iteration on natural-layout arrays
// (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:595-642).
// Component slot blockIdx.y.  cv: the convergence test of pass_real (a stopped rhs is
// skipped; the first skipped iteration after a deferred one adds the deferred increment
// to Field).
__global__ void update_plain(StateArgs st, Tables tb, Consts c, int upd, int fwd, int jj_inv, int jj_fwd, ConvArgs cv) {
  const int rhs = blockIdx.y;
  double nrm = 0.0, nrmF = 0.0;
  const long long ro = (long long)rhs * st.rhs_stride;
  float2* spec = st.spec + ro;
  float2* F = st.field + ro;
  float2* f = st.fstate + ro;
  if (cv.norms && jj_inv > 0) {
    const int rr = rhs / cv.P;
    if (conv_stopped(cv, jj_inv - 1, rr)) {
      if (!conv_stopped(cv, jj_inv - 2, rr) && upd_kind(jj_inv - 1, cv.M) == UPD_DEFER) {
        GRID_STRIDE(i, c.nvox) F[i] = cadd(F[i], f[i]);
      }
      return;
    }
  }
  const bool want_fref = st.fref != nullptr && upd == UPD_VQ_ADD;
  const float sg_inv = sign_inv(jj_inv ? flip_of(jj_inv, st.parity, rhs) : 0);
  const float sg_fwd = sign_fwd(jj_fwd ? flip_of(jj_fwd, st.parity, rhs) : 0);
  GRID_STRIDE(i, c.nvox) {
    const int x = (int)(i % c.nx);
    const long long yz = i / c.nx;
    const int y = (int)(yz % c.ny);
    const int z = (int)(yz / c.ny);
    const float2 r = cmul(cmul(tb.ramp_x[x], tb.ramp_y[y]), tb.ramp_z[z]);
    const float m = tb.mask_x[x] * tb.mask_y[y] * tb.mask_z[z];
    const float2 zero = make_float2(0.f, 0.f);
    const float2 q = (upd != UPD_NONE) ? spec[i] : zero;
    const float2 xo = reads_f(upd, fwd) ? f[i] : zero;
    const float2 Fo = reads_F(upd) ? F[i] : zero;
    float2 Fn, xn;
    const float2 p = born_step(upd, fwd, q, r, m, st.V[i], xo, Fo, c.inv_eps, sg_inv, sg_fwd, Fn, xn, st.norms != nullptr, nrm);
    if (writes_F(upd)) F[i] = Fn;
    if (want_fref) nrmF += (double)Fn.x * Fn.x + (double)Fn.y * Fn.y;
    if (fwd == FWD_GAMMA) f[i] = xn;
    if (fwd != FWD_NONE) spec[i] = p;
  }
  if (st.norms) flush_norm(nrm, st.norms, rhs);
  if (want_fref) flush_norm(nrmF, st.fref, rhs);
}

// Flip parity of each component slot: (flip_state + (rhs0 + v / P) * M) & 1, the
// persistent alternating-shift state of the reference's handle after the earlier
// right-hand sides (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:579-585).
__global__ void fill_parity(uint8_t* __restrict__ parity, int nslots, int P, int flip_state, int rhs0, int M) {
  GRID_STRIDE(v, nslots) parity[v] = (uint8_t)((flip_state + (long long)(rhs0 + v / P) * M) & 1);
}

// Zero fill of n 32-bit words (accumulators, counters), in place of a memset node.
__global__ void zero_words(uint32_t* __restrict__ p, long long n) {
  GRID_STRIDE(i, n) p[i] = 0u;
}

// p0 = (i/eps) * S / 2 with S = (V + i eps) .* u_in on the ROI, zero elsewhere
// (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:566,596).
__global__ void seed_source(const float2* __restrict__ u_in, const float2* __restrict__ V, float2* __restrict__ fstate,
                            Consts c, RoiGeom rg, long long rhs_stride, int nrhs) {
  GRID_STRIDE(i, c.nvox * nrhs) {
    const int r = (int)(i / c.nvox);
    const long long g = i - (long long)r * c.nvox;
    const int x = (int)(g % c.nx);
    const long long yz = g / c.nx;
    const int y = (int)(yz % c.ny);
    const int z = (int)(yz / c.ny);
    const int xr = x - rg.ox, yr = y - rg.oy, zr = z - rg.oz;
    float2 p = make_float2(0.f, 0.f);
    if (xr >= 0 && xr < rg.nx && yr >= 0 && yr < rg.ny && zr >= 0 && zr < rg.nz) {
      const long long gi = ((long long)zr * rg.ny + yr) * rg.nx + xr;
      const float2 u = u_in[(long long)r * rg.nvox + gi];
      const float2 v = V[g];
      const float2 s = cmul(make_float2(v.x, v.y + c.eps), u);   // (V + i eps) u = Vt u
      const float h = 0.5f * c.inv_eps;
      p = make_float2(-s.y * h, s.x * h);                         // (i / eps) s / 2
    }
    fstate[(long long)r * rhs_stride + ((long long)z * c.ny + state_row(c, y)) * c.nx + x] = p;
  }
}

// dst[i] = sum of the kNormSlots partial sums of accumulator i (src), n accumulators.
__global__ void sum_slots(const double* __restrict__ src, double* __restrict__ dst, long long n) {
  GRID_STRIDE(i, n) {
    double s = 0.0;
    for (int k = 0; k < kNormSlots; ++k) s += src[i * kNormSlots + k];
    dst[i] = s;
  }
}

// The potential in the state arrays' row order (V natural -> Vt, c.tiled = 1).
__global__ void tile_rows(const float2* __restrict__ V, float2* __restrict__ Vt, Consts c) {
  GRID_STRIDE(i, c.nvox) {
    const int x = (int)(i % c.nx);
    const long long yz = i / c.nx;
    const int y = (int)(yz % c.ny);
    const int z = (int)(yz / c.ny);
    Vt[((long long)z * c.ny + state_row(c, y)) * c.nx + x] = V[i];
  }
}

// Re V in the state arrays' row order (V natural -> Vr, c.tiled = 1), for a potential
// whose imaginary part is -eps everywhere (a real RI).
__global__ void tile_rows_real(const float2* __restrict__ V, float* __restrict__ Vr, Consts c) {
  GRID_STRIDE(i, c.nvox) {
    const int x = (int)(i % c.nx);
    const long long yz = i / c.nx;
    const int y = (int)(yz % c.ny);
    const int z = (int)(yz / c.ny);
    Vr[((long long)z * c.ny + state_row(c, y)) * c.nx + x] = V[i].x;
  }
}

// u = Field[ROI] + u_in (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:646-647).
__global__ void finalize_field(const float2* __restrict__ field, const float2* __restrict__ u_in, float2* __restrict__ u_out,
                               Consts c, RoiGeom rg, long long rhs_stride, int nrhs) {
  GRID_STRIDE(i, rg.nvox * nrhs) {
    const int r = (int)(i / rg.nvox);
    const long long gi = i - (long long)r * rg.nvox;
    const int xr = (int)(gi % rg.nx);
    const long long yz = gi / rg.nx;
    const int yr = (int)(yz % rg.ny);
    const int zr = (int)(yz / rg.ny);
    const long long g = ((long long)(zr + rg.oz) * c.ny + state_row(c, yr + rg.oy)) * c.nx + (xr + rg.ox);
    u_out[i] = cadd(field[(long long)r * rhs_stride + g], u_in[i]);
  }
}

// Vt = k^2 (RI^2 / n_m^2 - 1) (Codes/sub_code/RI2potential.m:1-4) in FP64, with the
// operations in the order and rounding of the host restatement in complex128 NumPy
// (no FMA contraction; the division of a complex by the real n_m^2 is a product with
// its reciprocal there), so that both give the same potential bit for bit.
// conj: use conj(RI).
__device__ __forceinline__ double2 ri_potential(const RiGrid& g, long long i, int conj) {
#pragma clang fp contract(off)
  double re, im;
  if (g.cplx) {
    re = g.ri[2 * i];
    im = conj ? -g.ri[2 * i + 1] : g.ri[2 * i + 1];
  } else {
    re = g.ri[i];
    im = 0.0;
  }
  const double sr = re * re - im * im, si = re * im + im * re;
  return make_double2(g.kk * (sr * g.inv_n2 - 1.0), g.kk * (si * g.inv_n2));
}

// Per-block max |Vt| over the RI grid (the max of eps = 1.01 max|Vt|,
// Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:205); blockDim.x = 256.
__global__ void potential_vmax(RiGrid g, double* __restrict__ block_max) {
  __shared__ double s_max[4];
  const long long n = (long long)g.nx * g.ny * g.nz;
  double m = 0.0;
  // a non-finite RI value makes the maximum NaN (as the host max does), so the caller rejects it
  auto vmax = [](double a, double b) { return (a != a || b != b) ? __builtin_nan("") : fmax(a, b); };
  GRID_STRIDE(i, n) {
    const double2 v = ri_potential(g, i, 0);
    m = vmax(m, hypot(v.x, v.y));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = vmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) block_max[blockIdx.x] = vmax(vmax(s_max[0], s_max[1]), vmax(s_max[2], s_max[3]));
}

// V = Vt - i eps on the simulation grid: Vt of the RI grid inside the ROI, 0 in the
// boundary (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:212-215,238-241), rounded to complex64.
// flip_conj: the RI taken as conj(RI) reversed along z, as the adjoint solve of
// Codes/BACKWARD_SOLVER_MULTI.m:48 uses it.
__global__ void potential_fill(RiGrid g, int flip_conj, RoiGeom rg, int snx, int sny, long long snvox, double eps,
                               float2* __restrict__ V) {
#pragma clang fp contract(off)
  GRID_STRIDE(i, snvox) {
    const int x = (int)(i % snx);
    const long long yz = i / snx;
    const int y = (int)(yz % sny);
    const int z = (int)(yz / sny);
    const int xr = x - rg.ox, yr = y - rg.oy, zr = z - rg.oz;
    double2 vt = make_double2(0.0, 0.0);
    if (xr >= 0 && xr < rg.nx && yr >= 0 && yr < rg.ny && zr >= 0 && zr < rg.nz) {
      const int zs = flip_conj ? rg.nz - 1 - zr : zr;
      vt = ri_potential(g, ((long long)zs * rg.ny + yr) * rg.nx + xr, flip_conj);
    }
    V[i] = make_float2((float)(vt.x - 0.0), (float)(vt.y - eps));
  }
}

}  // namespace cbs
