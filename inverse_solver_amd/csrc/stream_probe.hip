// Streaming-bandwidth probe (libcbs_probe.so): the practical HBM ceiling that the
// solver's passes are compared against.  Not part of the solver's C ABI; bench.py
// loads it to measure the ceiling on the same device in the same run.
//
// Each thread moves four 16-byte vectors per array with nontemporal loads and stores
// (the access form of the solver's state streams), 256 threads per block, one pass
// over the arrays per launch:
//   kind 0: c = a                 (1 read + 1 write stream)
//   kind 1: b = a + b, c = c/2 + a (3 read + 2 write streams, the mix of a Field-pair launch)
// and cbs_probe_pattern moves the fused passes' own access patterns (below).
#include <hip/hip_runtime.h>

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kU = 4;
constexpr int kNT = 256;

__device__ __forceinline__ v4f ld(const v4f* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st(v4f* p, v4f v) { __builtin_nontemporal_store(v, p); }

__global__ __launch_bounds__(kNT) void stream_copy(const v4f* __restrict__ a, v4f* __restrict__ c, long long n) {
  const long long i0 = (long long)blockIdx.x * kNT * kU + threadIdx.x;
  v4f r[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const long long i = i0 + u * kNT;
    r[u] = i < n ? ld(a + i) : v4f{0.f, 0.f, 0.f, 0.f};
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const long long i = i0 + u * kNT;
    if (i < n) st(c + i, r[u]);
  }
}

__global__ __launch_bounds__(kNT) void stream_3r2w(const v4f* __restrict__ a, v4f* __restrict__ b,
                                                   v4f* __restrict__ c, long long n) {
  const long long i0 = (long long)blockIdx.x * kNT * kU + threadIdx.x;
  v4f ra[kU], rb[kU], rc[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const long long i = i0 + u * kNT;
    if (i < n) {
      ra[u] = ld(a + i);
      rb[u] = ld(b + i);
      rc[u] = ld(c + i);
    }
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const long long i = i0 + u * kNT;
    if (i < n) {
      st(b + i, ra[u] + rb[u]);
      st(c + i, rc[u] * 0.5f + ra[u]);
    }
  }
}

// The fused passes' own traffic with no arithmetic besides a scale, so that each pass can
// be compared with what its access pattern moves on this device.  Arrays: nrhs blocks of
// nx*ny*nz complex64 (V one block); pairs of complex64 per 16-byte access.
//
// k-space pattern (pass_kspace): block (rhs, p1, x tile of 16) reads and writes back, in
// place, nz*b rows of 16 complex64 that lie a*nx elements apart (the spectrum layout
// [z][y2][p1][x]); blocks numbered rhs fastest; cached loads, nontemporal stores.
constexpr int kPT = 512;    // threads per block of the pattern probes
constexpr int kPK = 8;      // accesses per thread in flight

__global__ __launch_bounds__(kPT) void pattern_kspace(v4f* spec, int nx, int a, int b, int nz, int nrhs) {
  constexpr int T = 16, TP = T / 2;
  int bid = blockIdx.x;
  const int r = bid % nrhs;
  bid /= nrhs;
  const int p1 = bid % a, xt = bid / a;
  const long long rs = (long long)a * nx / 2;                // pairs between rows
  v4f* base = spec + (long long)r * nz * b * rs + ((long long)p1 * nx + xt * T) / 2;
  const int ntask = nz * b * TP;
  for (int t0 = 0; t0 < ntask; t0 += kPK * kPT) {
    v4f v[kPK];
#pragma unroll
    for (int k = 0; k < kPK; ++k) {
      const int t = t0 + threadIdx.x + k * kPT;
      if (t < ntask) v[k] = base[(t / TP) * rs + t % TP];
    }
#pragma unroll
    for (int k = 0; k < kPK; ++k) {
      const int t = t0 + threadIdx.x + k * kPT;
      if (t < ntask) st(base + (t / TP) * rs + t % TP, v[k] * 1.0001f);
    }
  }
}

// Real-space pattern (pass_real, storage-order numbering): block (y2, z, rhs), y2 fastest,
// moves one tile of a*nx contiguous complex64 in each array: reads spec, V, f (and Field
// when ODD), writes spec and f (and Field); nontemporal except V, which is shared by the
// rhs and read through the caches.  E2C > 0: the tile's pair count known at compile time,
// as it is in pass_real (only the last access of a thread needs a bounds test); 0: e2
// taken at run time.
// VF: V read as one float per voxel (a real RI's potential, as the real-space pass reads it).
template <int ODD, int E2C, int VF>
__global__ __launch_bounds__(kPT) void pattern_real(v4f* spec, const v4f* V, v4f* f, v4f* F, int nx, int ny, int nz,
                                                     int a, int b) {
  const int e2 = E2C > 0 ? E2C : a * nx / 2;                  // pairs per tile
  constexpr int L = E2C > 0 ? (E2C + kPT - 1) / kPT : kPK;    // accesses per thread in flight
  int bid = blockIdx.x;
  const int y2 = bid % b;
  bid /= b;
  const int z = bid % nz, r = bid / nz;
  const long long toff = (long long)(z * b + y2) * e2;
  const long long roff = (long long)r * nx * ny * nz / 2;
  for (int i0 = 0; i0 < e2; i0 += L * kPT) {
    v4f q[L], v[L], x[L], w[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int i = i0 + threadIdx.x + k * kPT;
      if (i < e2) {
        q[k] = ld(spec + roff + toff + i);
        if (VF) {
          const float2 h = reinterpret_cast<const float2*>(V)[toff + i];
          v[k] = v4f{h.x, -1.f, h.y, -1.f};
        } else {
          v[k] = V[toff + i];
        }
        x[k] = ld(f + roff + toff + i);
        if (ODD) w[k] = ld(F + roff + toff + i);
      }
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
      const int i = i0 + threadIdx.x + k * kPT;
      if (i < e2) {
        st(spec + roff + toff + i, q[k] * v[k] + x[k]);
        st(f + roff + toff + i, x[k] * 0.5f + q[k]);
        if (ODD) st(F + roff + toff + i, w[k] + x[k]);
      }
    }
  }
}

// the tile shapes of the compiled 160-point real-space passes (a = 20, 32, 40, 80)
using PatternFn = void (*)(v4f*, const v4f*, v4f*, v4f*, int, int, int, int, int);
template <int ODD, int VF>
PatternFn pattern_real_for(int e2) {
  switch (e2) {
    case 1600: return pattern_real<ODD, 1600, VF>;
    case 2560: return pattern_real<ODD, 2560, VF>;
    case 3200: return pattern_real<ODD, 3200, VF>;
    case 6400: return pattern_real<ODD, 6400, VF>;
    default: return pattern_real<ODD, 0, VF>;
  }
}

}  // namespace

extern "C" {

// Runs `reps` launches of probe `kind` over arrays of `bytes` bytes each (a multiple of
// 16; b is unused by kind 0) on `stream` (null: the null stream) after one untimed
// launch, and writes the average launch time in ms, from HIP events on that stream.
// Returns 0 or a hipError_t code (1 for bad arguments).
int cbs_probe_stream(int kind, const void* a, void* b, void* c, long long bytes, int reps, void* stream,
                     float* ms_per_launch) {
  if ((kind != 0 && kind != 1) || !a || !c || (kind == 1 && !b) || bytes <= 0 || bytes % 16 != 0 || reps < 1 ||
      !ms_per_launch)
    return 1;
  const long long n = bytes / 16;
  const long long nblk = (n + kNT * kU - 1) / (kNT * kU);
  if (nblk >= (1ll << 31)) return 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto launch = [&] {
    if (kind == 0)
      hipLaunchKernelGGL(stream_copy, dim3((unsigned)nblk), dim3(kNT), 0, s, static_cast<const v4f*>(a),
                         static_cast<v4f*>(c), n);
    else
      hipLaunchKernelGGL(stream_3r2w, dim3((unsigned)nblk), dim3(kNT), 0, s, static_cast<const v4f*>(a),
                         static_cast<v4f*>(b), static_cast<v4f*>(c), n);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  if (err == hipSuccess) {
    launch();
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e0, s);
  if (err == hipSuccess) {
    for (int r = 0; r < reps; ++r) launch();
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e1, s);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (err != hipSuccess) return (int)err;
  *ms_per_launch = ms / (float)reps;
  return 0;
}

// Runs `reps` launches of the pass pattern `kind` after one untimed launch and writes the
// average launch time in ms (HIP events on `stream`): kind 0 the k-space pattern on spec;
// 1 and 2 the real-space pattern of an even (spec, V, f) and an odd (with Field) launch;
// 3 and 4 the same with V read as one float per voxel (V then holds nx*ny*nz floats).
// dims = {nx, ny, nz, a, b, nrhs}: nx and a*nx even, ny = a*b, nx a multiple of 16 for
// kind 0; spec, f and F hold nrhs*nx*ny*nz complex64, V nx*ny*nz.  Returns 0, 1 for bad
// arguments, or a hipError_t code.
int cbs_probe_pattern(int kind, void* spec, const void* V, void* f, void* F, const int* dims, int reps, void* stream,
                      float* ms_per_launch) {
  if (kind < 0 || kind > 4 || !spec || !dims || reps < 1 || !ms_per_launch) return 1;
  const int nx = dims[0], ny = dims[1], nz = dims[2], a = dims[3], b = dims[4], nrhs = dims[5];
  if (nx < 2 || nx % 2 || a < 1 || b < 1 || ny != a * b || nz < 1 || nrhs < 1) return 1;
  if (kind == 0 && nx % 16) return 1;
  if (kind > 0 && (!V || !f || ((kind == 2 || kind == 4) && !F))) return 1;
  const long long nblk = kind == 0 ? (long long)nrhs * a * (nx / 16) : (long long)nrhs * b * nz;
  if (nblk >= (1ll << 31)) return 1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto launch = [&] {
    if (kind == 0)
      hipLaunchKernelGGL(pattern_kspace, dim3((unsigned)nblk), dim3(kPT), 0, s, static_cast<v4f*>(spec), nx, a, b, nz,
                         nrhs);
    else {
      const int e2 = a * nx / 2;
      const PatternFn fn = kind == 1 ? pattern_real_for<0, 0>(e2) : kind == 2 ? pattern_real_for<1, 0>(e2)
                         : kind == 3 ? pattern_real_for<0, 1>(e2) : pattern_real_for<1, 1>(e2);
      hipLaunchKernelGGL(fn, dim3((unsigned)nblk), dim3(kPT), 0, s, static_cast<v4f*>(spec), static_cast<const v4f*>(V),
                         static_cast<v4f*>(f), static_cast<v4f*>(F), nx, ny, nz, a, b);
    }
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  if (err == hipSuccess) {
    launch();
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e0, s);
  if (err == hipSuccess) {
    for (int r = 0; r < reps; ++r) launch();
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(e1, s);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (err != hipSuccess) return (int)err;
  *ms_per_launch = ms / (float)reps;
  return 0;
}

}  // extern "C"
