// Streaming-bandwidth probe (libcbs_probe.so): the practical HBM ceiling that the
// solver's passes are compared against.  Not part of the solver's C ABI; bench.py
// loads it to measure the ceiling on the same device in the same run.
//
// Each thread moves four 16-byte vectors per array with nontemporal loads and stores
// (the access form of the solver's state streams), 256 threads per block, one pass
// over the arrays per launch:
//   kind 0: c = a                 (1 read + 1 write stream)
//   kind 1: b = a + b, c = c/2 + a (3 read + 2 write streams, the mix of a Field-pair launch)
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

}  // extern "C"
