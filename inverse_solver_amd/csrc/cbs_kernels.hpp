// Convergent Born series kernels for gfx950.
//
// Memory layout (all complex64, interleaved re/im, x fastest like MATLAB):
//   V      [nz][ny][nx]            potential V = Vt - i*eps on the simulation grid
//   Field  [rhs][nz][ny][nx]       accumulated series
//   fstate [rhs][nz][ny][nx]       current increment f (and the source term p0 at start)
//   spec   [rhs][nz][b][a][nx]     intermediate spectrum between the two passes
//                                  (x in transposed k order, y split ny = a*b with
//                                  the a-part in k space and the b-part in real space)
//
// One Born iteration (Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:577-643) is two launches:
//   pass_real  : inverse a-DFT (y) + inverse x-FFT, the pointwise update with
//                ramp/mask/potential, Field accumulation, gamma*f with the next
//                ramp, forward x-FFT + a-DFT
//   pass_kspace: cross twiddle + b-DFT (y) + z-FFT, Green's multiply with the
//                +-1/4-bin shifted grid and 1/N, inverse z-FFT + inverse b-DFT +
//                conjugate cross twiddle
#pragma once
#include <type_traits>

#include "fft_core.hpp"

namespace cbs {

// Update kinds of the pointwise step (the two-chain recurrence of
// Codes/FORWARD_SOLVER_CONVERGENT_BORN.m:587-642).
enum UpdKind : int {
  UPD_NONE = 0,      // no inverse input: p comes from the source term
  UPD_VQ_SET = 1,    // jj = 1 : f1 = A V q;            Field  = f1
  UPD_VQ_ADD = 2,    // jj = 2 : f2 = A V q;            Field += f2; x = f2
  UPD_FULL3 = 3,     // jj = 3 : f3 = A (x - gx + V q); Field += f3; x = f3 + F1 (F1 = Field - x)
  UPD_FULL = 4,      // jj > 3 : fn = A (x - gx + V q); Field += fn; x = fn
  // Field accumulation in pairs (jj >= 4): an even jj below the last leaves Field
  // alone, and the odd jj after it adds both increments, the earlier one being the
  // x it reads.  Field then crosses HBM every other iteration instead of every one.
  UPD_DEFER = 5,     // fn = A (x - gx + V q); x = fn                 (Field untouched)
  UPD_PAIR = 6,      // fn = A (x - gx + V q); Field += x + fn; x = fn
};
enum FwdKind : int {
  FWD_NONE = 0,      // last iteration: nothing to transform
  FWD_P0 = 1,        // p = source term stored in fstate (jj = 1, 2)
  FWD_GAMMA = 2,     // p = gamma * x = (i/eps) V x
};

struct Tables {
  const float2* tw_x;     // W_nx^e, e < nx
  const float2* tw_a;     // W_a^e
  const float2* tw_z;     // W_nz^e
  const float2* cross;    // [b][a]: W_ny^(y2 * k1(p1))
  const float2* ramp_x;   // sigma=+1 ramp per axis (1 where the axis has no ramp)
  const float2* ramp_y;
  const float2* ramp_z;
  const float* mask_x;    // attenuation mask factors per axis
  const float* mask_y;
  const float* mask_z;
  const float* green_x;   // [2][nx]  4 pi^2 (k + sh dk/4)^2 at x position p (index 0: sigma=+1, 1: flipped)
  const float* green_y;   // [2][b][a] at (k2, p1)
  const float* green_z;   // [2][nz]  at z position p
  const float* rad_x;     // vector runs: (k + sh dk/4) / k0 per axis, same layout as green_* (else null)
  const float* rad_y;
  const float* rad_z;
};

struct Consts {
  int nx, ny, nz, a, b;
  int tiled;          // 1 (fused form): the rows of each z plane of V, Field and f are stored
                      // in the real-space pass's tile order, row y at (y % b) a + y / b
  int nrhs;
  long long nvox;
  float inv_eps;
  float eps;
  float km2;
  float inv_n;
  float c_re, c_im;   // vector runs: k^2 / (k^2 + i eps), the dyadic absorption correction
};

// storage row of simulation row y in a z plane of the state arrays
__host__ __device__ __forceinline__ int state_row(const Consts& c, int y) {
  return c.tiled ? (y % c.b) * c.a + y / c.b : y;
}

struct StateArgs {
  const float2* V;
  const float* Vr;         // fused form, real RI: Re V in the state arrays' row order (Im V = -eps
                           // everywhere, so the real-space pass reads 4 bytes a voxel); null: use V
  float2* field;
  float2* fstate;
  float2* spec;
  const uint8_t* parity;   // per-rhs flip parity (0/1)
  long long rhs_stride;    // elements between consecutive rhs in field/fstate/spec
  double* norms;           // optional [nrhs][kNormSlots] FP64 |f_n|^2 partial sums (may be null)
  double* fref;            // optional [nrhs][kNormSlots] FP64 |Field|^2 partial sums, filled by the jj = 2 launch
};

// FP64 norm accumulators: every (iteration, component slot) sum is held in kNormSlots
// partial sums (flush_norm), added up by their readers.
constexpr int kNormSlots = 64;

// FP64 convergence test (an option; the reference runs a fixed Bornmax).  Rhs r has
// stopped after iteration n >= 3 when |f_n|^2 <= tol^2 |F_1 + F_2|^2, both summed in
// FP64 over the grid and the rhs's P component slots: the increment norm relative to
// the first-order (jj = 2) Field.  Iterations after the stop leave Field and f alone;
// an iteration that skips, the increment it adds is zero, so the test stays true and
// every later launch skips as well.  norms = null: no test.
struct ConvArgs {
  const double* norms;     // [M][nslots][kNormSlots] |f_n|^2 partial sums (row n - 1: iteration n)
  const double* fref;      // [nslots][kNormSlots] |F_1 + F_2|^2 partial sums
  double tol2;
  int P;                   // component slots per rhs
  int nslots;              // slots per norms row
  int M;                   // Bornmax
};

__device__ __forceinline__ bool conv_stopped(const ConvArgs& cv, int n, int r) {
  if (n < 3) return false;
  double s = 0.0, ref = 0.0;
  for (int q = 0; q < cv.P; ++q) {
    const long long slot = (long long)cv.P * r + q;
    for (int k = 0; k < kNormSlots; ++k) {
      s += cv.norms[((long long)(n - 1) * cv.nslots + slot) * kNormSlots + k];
      ref += cv.fref[slot * kNormSlots + k];
    }
  }
  return s <= cv.tol2 * ref;
}

// flip state of iteration jj for rhs r: 1 = Greenp flipped (sigma = -1), 0 = sigma = +1
__device__ __forceinline__ int flip_of(int jj, const uint8_t* parity, int r) { return (jj + parity[r]) & 1; }

// Which state arrays a pointwise step reads.
__host__ __device__ constexpr bool reads_f(int upd, int fwd) { return upd >= UPD_FULL3 || fwd == FWD_P0; }
__host__ __device__ constexpr bool reads_F(int upd) {
  return upd == UPD_VQ_ADD || upd == UPD_FULL3 || upd == UPD_FULL || upd == UPD_PAIR;
}
__host__ __device__ constexpr bool writes_F(int upd) { return upd != UPD_NONE && upd != UPD_DEFER; }

// Update kind of iteration jj of a series of M iterations (see UpdKind).
__host__ __device__ constexpr int upd_kind(int jj, int M) {
  return jj == 1 ? UPD_VQ_SET : jj == 2 ? UPD_VQ_ADD : jj == 3 ? UPD_FULL3
       : (jj % 2 == 0 && jj < M) ? UPD_DEFER : (jj % 2 == 1 && jj >= 5) ? UPD_PAIR : UPD_FULL;
}

// Pointwise step on one voxel.
//   q   : G_s applied to the previous p, before the conj ramp (unused for UPD_NONE)
//   r   : sigma=+1 ramp at the voxel, m: attenuation mask, v: potential
//   sg_inv, sg_fwd: ramp orientation of the inverse and forward sides, +1 or -1:
//         q is multiplied by (r.x, sg_inv r.y) (= conj(ramp_s) for the inverse flip
//         state) and the next p by (r.x, sg_fwd r.y) (= ramp_s of the forward one)
//   xo  : fstate value (p0 for FWD_P0, x for the full update), Fo: Field value
// Outputs: Fn (new Field, when writes_F(upd)), xn (new fstate, when fwd == GAMMA),
// |fn|^2 added to nrm in FP64 when want_norm; returns the next p times the forward
// ramp (0 when fwd == NONE).
__device__ __forceinline__ float2 born_step(int upd, int fwd, float2 q, float2 r, float m, float2 v, float2 xo,
                                            float2 Fo, float inv_eps, float sg_inv, float sg_fwd, float2& Fn,
                                            float2& xn, bool want_norm, double& nrm) {
  xn = make_float2(0.f, 0.f);
  if (upd != UPD_NONE) {
    const float2 qq = cmul(q, make_float2(r.x, sg_inv * r.y));
    float2 fn;
    if (upd == UPD_VQ_SET || upd == UPD_VQ_ADD) {
      fn = cscale(cmul(v, qq), m);
      Fn = (upd == UPD_VQ_SET) ? fn : cadd(Fo, fn);
      xn = fn;
    } else {
      // A (x - gamma x + V q) with gamma x = (i/eps) V x, as A (x + V (q - (i/eps) x))
      const float2 t = make_float2(fmaf(xo.y, inv_eps, qq.x), fmaf(-xo.x, inv_eps, qq.y));
      fn = cscale(cadd(xo, cmul(v, t)), m);
      Fn = upd == UPD_PAIR ? cadd(Fo, cadd(xo, fn)) : cadd(Fo, fn);
      xn = (upd == UPD_FULL3) ? cadd(fn, csub(Fo, xo)) : fn;
    }
    if (want_norm) nrm += (double)fn.x * fn.x + (double)fn.y * fn.y;
  }
  if (fwd == FWD_NONE) return make_float2(0.f, 0.f);
  float2 p;
  if (fwd == FWD_P0) {
    p = xo;
  } else {
    const float2 vx = cmul(v, xn);
    p = make_float2(-vx.y * inv_eps, vx.x * inv_eps);
  }
  return cmul(p, make_float2(r.x, sg_fwd * r.y));
}

// Ramp orientation signs for born_step from the flip states (1 = Greenp flipped).
__device__ __forceinline__ float sign_inv(int flip_inv) { return flip_inv ? 1.f : -1.f; }
__device__ __forceinline__ float sign_fwd(int flip_fwd) { return flip_fwd ? -1.f : 1.f; }

// Wave-level reduction of a double then one atomic per wave.
// Adds a wave's partial sum to the accumulator of component slot `slot`, which is
// spread over kNormSlots sub-slots picked by block and wave: blocks in flight that all
// work on one slot (storage-order numbering) then do not queue on one address.  The
// readers sum the sub-slots (conv_stopped, sum_slots).
__device__ __forceinline__ void flush_norm(double v, double* base, long long slot) {
  if (!base) return;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int sub = (int)((blockIdx.x * 8u + (threadIdx.x >> 6)) & (unsigned)(kNormSlots - 1));
  if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(base + slot * kNormSlots + sub, v);
}

// 16-byte store of a state array.  With nt set (state larger than the caches: the
// array is next read a whole launch later, after every other tile) it is marked
// nontemporal; a state that fits the 256 MB Infinity Cache keeps ordinary stores.
__device__ __forceinline__ void store_stream(float4* dst, float4 v, bool nt) {
  if (nt) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<v4f*>(dst));
  } else {
    *dst = v;
  }
}

// 16-byte load of a state array read once per launch (nontemporal when nt, as above).
__device__ __forceinline__ float4 load_stream(const float4* src, bool nt) {
  if (nt) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    const v4f t = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(src));
    return make_float4(t.x, t.y, t.z, t.w);
  }
  return *src;
}

// T* at a 32-bit byte offset from a (block-uniform) base pointer
template <class T, class B>
__device__ __forceinline__ T* byte_at(B* base, unsigned byte_off) {
  using C = typename std::conditional<std::is_const<B>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<C*>(base) + byte_off);
}

__device__ __forceinline__ float2 lo2(float4 v) { return make_float2(v.x, v.y); }
__device__ __forceinline__ float2 hi2(float4 v) { return make_float2(v.z, v.w); }
__device__ __forceinline__ float4 pack2(float2 a, float2 b) { return make_float4(a.x, a.y, b.x, b.y); }

}  // namespace cbs
