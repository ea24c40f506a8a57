// Host side of the CBS solver: handle registry, kernel planning, FP64 table
// construction, the iteration schedule and the C ABI (include/cbs_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cbs_hip.h"
#include "cbs_passes.hpp"
#include "fft_rt.hpp"
#include "goldstein.hpp"
#include "io_kernels.hpp"
#include "kernels_misc.hpp"
#include "tv_kernels.hpp"
#include "unwrap_kernels.hpp"

namespace cbs {

// ------------------------------------------------------------------ errors
struct CbsError : std::runtime_error {
  int code;
  CbsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
static thread_local std::string g_last_error;

#define HIP_CHECK(x)                                                                                    \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess) throw CbsError(CBS_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

static void require(bool cond, int code, const std::string& msg) {
  if (!cond) throw CbsError(code, msg);
}

// ------------------------------------------------------------------ device buffers
// Operations that synchronise the whole device (frees, allocations, stream and graph
// teardown) and graph capture exclude each other across threads: a device-wide
// synchronisation waits on a stream that another handle is capturing, which invalidates
// that capture.
static std::recursive_mutex& device_mutex() {
  static std::recursive_mutex m;
  return m;
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      std::lock_guard<std::recursive_mutex> lk(device_mutex());
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    release();
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw CbsError(CBS_ERR_MEMORY, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed: " +
                                         hipGetErrorString(e));
    }
    n = count;
  }
  void upload(const std::vector<T>& v) {
    ensure(v.size());
    HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
};

// ------------------------------------------------------------------ FFT position maps
static int pos2freq(int n, int p) {
  const int r2 = plan_r2(n);
  if (r2 == 1) return p;
  const int r1 = n / r2;
  return (p / r2) + r1 * (p % r2);
}

// ------------------------------------------------------------------ table set
struct TableSet {
  int a = 0, b = 0;
  DevBuf<float2> tw_x, tw_a, tw_y, tw_z, cross, ramp_x, ramp_y, ramp_z;
  DevBuf<float> mask_x, mask_y, mask_z, green_x, green_y, green_z, rad_x, rad_y, rad_z;
  Tables view() const {
    Tables t;
    t.tw_x = tw_x.p; t.tw_a = tw_a.p; t.tw_z = tw_z.p; t.cross = cross.p;
    t.ramp_x = ramp_x.p; t.ramp_y = ramp_y.p; t.ramp_z = ramp_z.p;
    t.mask_x = mask_x.p; t.mask_y = mask_y.p; t.mask_z = mask_z.p;
    t.green_x = green_x.p; t.green_y = green_y.p; t.green_z = green_z.p;
    t.rad_x = rad_x.p; t.rad_y = rad_y.p; t.rad_z = rad_z.p;
    return t;
  }
};

static std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n);
  for (int e = 0; e < n; ++e) {
    const double ang = -2.0 * M_PI * (double)e / (double)n;
    t[e] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  return t;
}

struct Plan {
  int form = CBS_FORM_FUSED;
  int a = 0, b = 0, t = 0;
  const PassRInfo* pr = nullptr;
  const PassKInfo* pk = nullptr;
  const RowsInfo* rw = nullptr;
  const StridedInfo* sy = nullptr;
  const StridedInfo* sz = nullptr;
  bool rt[3] = {false, false, false};   // per-axis form: axis j runs the run-time-planned FFT
};

// ------------------------------------------------------------------ run-time-planned line FFTs
struct RtPlan {
  RtFft k{};
  DevBuf<float2> tw, chirp, bspec;
  const BlueInfo* bi = nullptr;   // Bluestein on a compiled convolution length (else k)
  BlueArgs ba{};
  int pref = 0;                   // built with prefer_blue
};

// Stage radices of a Stockham transform of length L (16, 8, 4, 2 for the power of two,
// then 9, 3, 5, 7, 11, 13); false when a larger prime factor is left.
static bool rt_factor(int L, std::vector<int>& rs) {
  rs.clear();
  int m = L;
  while (m % 16 == 0) { rs.push_back(16); m /= 16; }
  for (int r : {8, 4, 2})
    if (m % r == 0) { rs.push_back(r); m /= r; }
  for (int r : {9, 3, 5, 7, 11, 13})
    while (m % r == 0) { rs.push_back(r); m /= r; }
  return m == 1;
}

// smallest m' >= m whose prime factors are 2, 3, 5 and 7
static int smooth_at_least(int m) {
  for (;; ++m) {
    int k = m;
    for (int q : {2, 3, 5, 7})
      while (k % q == 0) k /= q;
    if (k == 1) return m;
  }
}

// Plan of the length-n line transform: Stockham when n factors into the radices,
// else Bluestein with the chirp c_k = exp(i pi k^2 / n) and the DFT of the chirp
// filter over a length L >= 2n - 1 (FP64 on the host, rounded to complex64) -- on the
// shortest compiled convolution length when one is long enough (fft_blue.hip), else
// on a run-time Stockham transform.
// prefer_blue: take Bluestein's transform on a compiled convolution length even when n
// factors into run-time radices (the per-axis form fuses those transforms with the update
// and the Green's multiply; run-time stages have no fused form)
static void build_rt(RtPlan& P, int n, bool prefer_blue = false) {
  require(n >= 1, CBS_ERR_ARG, "line length must be >= 1");
  std::vector<int> rs;
  int L = n, blue = 0;
  P.bi = nullptr;
  P.pref = prefer_blue ? 1 : 0;
  if (!rt_factor(n, rs) || (prefer_blue && find_blue(2 * n - 1))) {
    blue = 1;
    P.bi = find_blue(2 * n - 1);
    L = P.bi ? P.bi->L : smooth_at_least(2 * n - 1);
    rt_factor(L, rs);
  }
  require(2 * L <= kRtLdsElems && (int)rs.size() <= kRtMaxStages, CBS_ERR_UNSUPPORTED,
          "line length " + std::to_string(n) + " is beyond the run-time FFT (transform length " + std::to_string(L) +
              " > " + std::to_string(kRtLdsElems / 2) + ")");
  P.k = RtFft{};
  P.k.n = n;
  P.k.L = L;
  P.k.blue = blue;
  P.k.nst = (int)rs.size();
  for (size_t i = 0; i < rs.size(); ++i) P.k.radix[i] = rs[i];
  std::vector<float2> tw(L);
  std::vector<std::complex<double>> w(L);
  for (int e = 0; e < L; ++e) {
    const double ang = -2.0 * M_PI * (double)e / (double)L;
    w[e] = std::complex<double>(std::cos(ang), std::sin(ang));
    tw[e] = make_float2((float)w[e].real(), (float)w[e].imag());
  }
  P.tw.upload(tw);
  P.k.tw = P.tw.p;
  P.k.chirp = P.k.bspec = nullptr;
  if (blue) {
    std::vector<std::complex<double>> c(n), bv(L, 0.0);
    std::vector<float2> cf(n), bf(L);
    for (int k = 0; k < n; ++k) {
      const long long q = ((long long)k * k) % (2ll * n);   // k^2 mod 2n: the exact phase
      const double ang = M_PI * (double)q / (double)n;
      c[k] = std::complex<double>(std::cos(ang), std::sin(ang));
      cf[k] = make_float2((float)c[k].real(), (float)c[k].imag());
    }
    for (int m = 0; m < n; ++m) bv[m] = c[m];
    for (int m = 1; m < n; ++m) bv[L - m] = c[m];
    for (int k = 0; k < L; ++k) {
      std::complex<double> acc = 0.0;
      for (int m = 0; m < L; ++m)
        if (bv[m] != 0.0) acc += bv[m] * w[(size_t)(((long long)k * m) % L)];
      acc /= (double)L;
      bf[k] = make_float2((float)acc.real(), (float)acc.imag());
    }
    if (P.bi) {
      // the compiled convolution multiplies position p of its forward output, which
      // holds frequency p / r2 + (L / r2)(p % r2)
      const int r2 = P.bi->r2, r1 = L / r2;
      std::vector<float2> bp(L);
      for (int q = 0; q < L; ++q) bp[q] = bf[r2 == 1 ? q : q / r2 + r1 * (q % r2)];
      bf.swap(bp);
    }
    P.chirp.upload(cf);
    P.bspec.upload(bf);
    P.k.chirp = P.chirp.p;
    P.k.bspec = P.bspec.p;
    P.ba = BlueArgs{n, P.chirp.p, P.bspec.p, P.tw.p};
  }
}

// ------------------------------------------------------------------ no-allocation workspace
// Device memory of cbs_solve_3d_noalloc: the state arrays (Field, f, the spectrum), the
// per-slot flip parities and the FP64 norm accumulators of one batch.  The module holds
// one copy on every GPU it runs on, for the life of the process; every such call on a
// GPU shares that GPU's copy.  In host code the array's name is not a device address:
// scratch_base() gives the address of the current device's copy.
constexpr size_t kScratchBytes = CBS_NOALLOC_WORKSPACE_BYTES;
__attribute__((used)) static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on device dev, which must be the current device; resolved once
// per device (hipGetSymbolAddress) and kept.
static unsigned char* scratch_base(int dev) {
  require(dev >= 0 && dev < kMaxDevices, CBS_ERR_ARG, "device ordinal out of range for the no-allocation workspace");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      (void)hipGetLastError();
      throw CbsError(CBS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

// ------------------------------------------------------------------ solver
struct Solver {
  int device = 0;
  int P = 1;                      // field components: 1 scalar, 3 vector (dyadic Green's function)
  int nx = 0, ny = 0, nz = 0;
  int rx = 0, ry = 0, rz = 0;     // RI grid
  int fx = 0, fy = 0;             // field grid
  int cx = 0, cy = 0;             // convolution grid
  long long nvox = 0, nri = 0;
  hipStream_t stream = nullptr;
  cbs_options opt{};
  bool have_potential = false;
  cbs_potential_params prm{};
  DevBuf<float2> V;
  // the fused form's copy of V in the state arrays' tile row order (state_row), made when
  // the potential or the plan changes, never inside a solve: Vt, or for a real RI (V's
  // imaginary part -eps everywhere, as potential_fill forms it) only its real part, Vr
  DevBuf<float2> Vt;
  DevBuf<float> Vr;
  bool v_real = false;
  long long v_version = 0;
  std::vector<long long> vt_key;
  Plan plan;
  TableSet tabs;        // tables of the active plan
  int flip_state = 0;   // persistent alternating-shift parity (reference handle semantics)
  // workspace: the handle's own buffers; ws points at them (or at memory a caller provides)
  DevBuf<float2> field, fstate, spec, uin, uout;
  DevBuf<uint8_t> parity;
  DevBuf<double> norms, fref;
  struct Ws {
    float2* field = nullptr;
    float2* fstate = nullptr;
    float2* spec = nullptr;
    uint8_t* parity = nullptr;
    double* norms = nullptr;   // [M][slots][kNormSlots] partial sums, then [M][slots] totals, when tracked
    double* fref = nullptr;    // [slots][kNormSlots] partial sums, then [slots] totals, with the test on
  } ws;
  std::vector<double> last_norms;
  int last_norm_rhs = 0, last_norm_iter = 0;
  // FP64 convergence test (0 = off: every solve runs Bornmax iterations)
  double conv_tol = 0.0;
  std::vector<uint32_t> last_iters;   // iterations each rhs of the last solve ran
  bool tracking() const { return opt.track_norms || conv_tol > 0.0; }
  // kernel timing
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_used;   // (kind, index of start event)
  double ktime_ms[3] = {0, 0, 0};
  uint64_t kcount[3] = {0, 0, 0};
  // graph cache: two captured loops (a pipelined host solve alternates between its chunk
  // width and a shorter last chunk)
  struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<long long> key;
  };
  GraphSlot gslot[2];
  int gslot_next = 0;
  // overlapped host I/O: a copy stream and its events (created on first use)
  hipStream_t cstream = nullptr;
  hipEvent_t ev_up[2] = {nullptr, nullptr}, ev_done[2] = {nullptr, nullptr}, ev_down[2] = {nullptr, nullptr};
  // I/O projection state
  IoState io;
  int rc_x = 0, rc_y = 0;
  IoGeom iog{};
  DevBuf<float2> io_refocus_k, io_ktrans, io_kref, io_tw_cx, io_tw_cy;
  DevBuf<int> io_kcx, io_kcy;
  DevBuf<float2> io_buf3, io_spec2, io_out_t, io_out_r, io_uin, io_src, io_u3d, io_tr, io_rf;
  const RowsInfo* io_rows = nullptr;
  const StridedInfo* io_cols = nullptr;
  // run-time-planned line FFTs: the per-axis form's axes without a compiled kernel, and
  // the convolution grid's x / y transforms (used when io_rows / io_cols are null)
  RtPlan rtp[3], io_rtx, io_rty;

  void rt_rows(const RtPlan& P, float2* data, long long nrows, int inverse) {
    if (P.bi) {
      const long long nblk = (nrows + P.bi->nl - 1) / P.bi->nl;
      require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "run-time FFT batch too large for one launch");
      timed(2, [&] { hipLaunchKernelGGL(P.bi->rows, dim3((unsigned)nblk), dim3(P.bi->nt), 0, stream, data, nrows,
                                        P.ba, inverse); });
      return;
    }
    const int lines = rt_lines(P.k.L, 64);
    const long long nblk = (nrows + lines - 1) / lines;
    require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "run-time FFT batch too large for one launch");
    const size_t lds = 2ull * lines * P.k.L * sizeof(float2);
    timed(2, [&] { hipLaunchKernelGGL(fft_rows_rt, dim3((unsigned)nblk), dim3(256), lds, stream, data, nrows, P.k,
                                      lines, inverse); });
  }
  // outer set o < nouter at (o / inner) * os2 + (o % inner) * os (inner = 0: one level)
  void rt_strided(const RtPlan& P, float2* data, int ncols, long long es, long long os, long long nouter,
                  int inverse, long long inner = 0, long long os2 = 0) {
    const int in = (int)(inner > 0 ? inner : nouter);
    if (P.bi) {
      const int T = P.bi->t;
      const long long nblk = (long long)((ncols + T - 1) / T) * nouter;
      require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "run-time FFT batch too large for one launch");
      timed(2, [&] { hipLaunchKernelGGL(P.bi->strided, dim3((unsigned)nblk), dim3(P.bi->nt), 0, stream, data, ncols,
                                        es, os, (int)nouter, P.ba, inverse, in, os2); });
      return;
    }
    const int T = rt_lines(P.k.L, 16);
    const long long nblk = (long long)((ncols + T - 1) / T) * nouter;
    require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "run-time FFT batch too large for one launch");
    const size_t lds = 2ull * T * P.k.L * sizeof(float2);
    timed(2, [&] { hipLaunchKernelGGL(fft_strided_rt, dim3((unsigned)nblk), dim3(256), lds, stream, data, ncols,
                                      es, os, (int)nouter, P.k, T, inverse, in, os2); });
  }

  void drop_graph() {
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    for (auto& g : gslot) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      if (g.graph) (void)hipGraphDestroy(g.graph);
      g = GraphSlot();
    }
    gslot_next = 0;
  }

  ~Solver() {
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    drop_graph();
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; ++k)
      for (hipEvent_t e : {ev_up[k], ev_done[k], ev_down[k]})
        if (e) (void)hipEventDestroy(e);
    if (cstream) (void)hipStreamDestroy(cstream);
    if (stream) (void)hipStreamDestroy(stream);
    (void)hipSetDevice(prev);
  }

  Consts consts(int nb) const {
    Consts c;
    c.nx = nx; c.ny = ny; c.nz = nz; c.a = plan.a; c.b = plan.b; c.nrhs = nb; c.nvox = nvox;
    c.tiled = plan.form == CBS_FORM_FUSED ? 1 : 0;
    c.eps = (float)prm.eps_imag;
    c.inv_eps = (float)(1.0 / prm.eps_imag);
    const double km = 2.0 * M_PI * prm.k0_nm;
    c.km2 = (float)(km * km);
    c.inv_n = (float)(1.0 / (double)nvox);
    // k^2 / (k^2 + i eps) = k^2 (k^2 - i eps) / (k^4 + eps^2)
    const double k2 = km * km, e = prm.eps_imag, den = k2 * k2 + e * e;
    c.c_re = (float)(k2 * k2 / den);
    c.c_im = (float)(-k2 * e / den);
    return c;
  }
  RoiGeom roi() const {
    RoiGeom g;
    g.ox = (int)prm.ROI[0] - 1; g.oy = (int)prm.ROI[2] - 1; g.oz = (int)prm.ROI[4] - 1;
    g.nx = rx; g.ny = ry; g.nz = rz; g.nvox = nri;
    return g;
  }

  // ---- planning
  void make_plan() {
    Plan p;
    p.form = opt.form;
    if (p.form == CBS_FORM_ROWS) {
      if (make_rows_plan(p)) return;
      throw CbsError(CBS_ERR_UNSUPPORTED, "no four-pass (row block) kernels for grid " + shape_str());
    }
    if (p.form == CBS_FORM_FUSED) {
      const PassRInfo* prs = nullptr;
      const int npr = list_pass_real(&prs);
      const PassKInfo* pks = nullptr;
      const int npk = list_pass_kspace(&pks);
      // default preference: the largest a whose LDS tile keeps two blocks per CU
      // (<= 80 KiB), then the widest x tile, then 256 threads; cbs_autotune
      // replaces this choice by measurement.
      // (measured on MI355X at 160^3 x 49 rhs: 512-thread blocks beat 256 for
      // tiles of 40 KiB and more)
      auto score_r = [&](const PassRInfo& r) {
        return (r.lds <= 81920 ? 1000000 : 0) + r.a * 10 + ((r.nt == 512) == (r.lds >= 40960) ? 1 : 0);
      };
      auto score_k = [&](const PassKInfo& q) {
        return (q.lds <= 81920 ? 1000 : 0) + (q.t >= 16 ? 100 : 0) - q.t + ((q.nt == 512) == (q.lds >= 40960) ? 1 : 0);
      };
      long best = -1;
      for (int i = 0; i < npr; ++i) {
        const PassRInfo& r = prs[i];
        if (!fused_ok_r(r)) continue;
        const int b = ny / r.a;
        const PassKInfo* k = nullptr;
        for (int j = 0; j < npk; ++j) {
          const PassKInfo& q = pks[j];
          if (!fused_ok_k(q, b)) continue;
          if (!k || score_k(q) > score_k(*k)) k = &q;
        }
        if (!k) continue;
        const long sc = (long)score_r(r) * 10000 + score_k(*k);
        if (sc > best) {
          best = sc;
          p.pr = &r;
          p.pk = k;
        }
      }
      if (p.pr) {
        p.a = p.pr->a;
        p.b = ny / p.a;
        p.t = p.pk->t;
        plan = p;
        return;
      }
      if (opt.split_a > 0 || opt.tile_t > 0 || opt.threads_real > 0 || opt.threads_kspace > 0)
        throw CbsError(CBS_ERR_UNSUPPORTED,
                       "no fused kernel plan for the forced split/tile/threads on grid " + shape_str());
      if (make_rows_plan(p)) return;   // then the four-pass form
      p.form = CBS_FORM_PLAIN;         // grids without either run the per-axis form
    }
    // per-axis form: compiled line FFTs where the length has one, run-time-planned
    // (Stockham, or Bluestein for prime factors above 13) for every other length
    p.rw = find_fft_rows(nx);
    p.sy = find_fft_strided(ny, 0);
    p.sz = find_fft_strided(nz, 0);
    p.rt[0] = !p.rw;
    p.rt[1] = !p.sy;
    p.rt[2] = !p.sz;
    const int dims[3] = {nx, ny, nz};
    // one-component runs fuse the x transforms with the update and the z transforms with
    // the Green's multiply when they are compiled (line FFT or Bluestein on a compiled
    // convolution length), so those two axes take Bluestein's transform over run-time
    // stages (12-rhs solves on 169^2 x 75, 98^2 x 90, 126^2 x 78 and 147^2 x 77 grids: 22.0,
    // 22.4, 17.1 and 17.5 -> 25.0, 29.3, 26.8 and 22.4 Gvoxel-iter/s); y, which no pass
    // fuses, keeps the run-time stages
    const bool fuse = P == 1 && p.form == CBS_FORM_PLAIN;
    for (int j = 0; j < 3; ++j) {
      const int pref = fuse && j != 1 ? 1 : 0;
      if (p.rt[j] && (rtp[j].k.n != dims[j] || rtp[j].pref != pref)) build_rt(rtp[j], dims[j], pref);
    }
    p.a = ny;
    p.b = 1;
    p.t = p.sy ? p.sy->t : 0;
    plan = p;
  }
  // four-pass form: the real-space pass on blocks of contiguous rows, strided y
  // transforms, the k-space pass with b = 1 on the natural layout
  bool make_rows_plan(Plan& p) {
    const PassRInfo* prs = nullptr;
    const int npr = list_pass_real(&prs);
    const PassKInfo* pks = nullptr;
    const int npk = list_pass_kspace(&pks);
    const StridedInfo* sy = find_fft_strided(ny, 0);
    if (!sy || nx % 2 != 0) return false;
    const PassRInfo* br = nullptr;
    const PassKInfo* bk = nullptr;
    for (int i = 0; i < npr; ++i) {
      const PassRInfo& r = prs[i];
      if (!r.rows || r.nx != nx || ny % r.a != 0) continue;
      if (opt.threads_real > 0 && r.nt != opt.threads_real) continue;
      if (!br || (r.lds <= 81920 && (br->lds > 81920 || r.a > br->a))) br = &r;
    }
    for (int j = 0; j < npk; ++j) {
      const PassKInfo& q = pks[j];
      if (q.b != 1 || q.nz != nz || q.pole != P) continue;
      if (opt.tile_t > 0 && q.t != opt.tile_t) continue;
      if (opt.threads_kspace > 0 && q.nt != opt.threads_kspace) continue;
      if (!bk || (q.lds <= 81920 && (bk->lds > 81920 || q.t > bk->t))) bk = &q;
    }
    if (!br || !bk) return false;
    p.form = CBS_FORM_ROWS;
    p.pr = br;
    p.pk = bk;
    p.sy = sy;
    p.a = br->a;          // rows per real-space block
    p.b = ny / br->a;     // row blocks per z plane
    p.t = bk->t;
    plan = p;
    return true;
  }
  bool fused_ok_r(const PassRInfo& r) const {
    if (r.rows) return false;
    if (r.nx != nx || ny % r.a != 0 || nx % 2 != 0) return false;   // pairs of x per 16-byte access
    if (opt.split_a > 0 && r.a != opt.split_a) return false;
    if (opt.threads_real > 0 && r.nt != opt.threads_real) return false;
    return true;
  }
  bool fused_ok_k(const PassKInfo& q, int b) const {
    if (q.b != b || q.nz != nz || q.pole != P) return false;
    if (opt.tile_t > 0 && q.t != opt.tile_t) return false;
    if (opt.threads_kspace > 0 && q.nt != opt.threads_kspace) return false;
    return true;
  }
  std::string shape_str() const {
    return std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz);
  }

  // ---- FP64 tables for the active plan
  // what build_tables() reads: the plan and the potential's geometry scalars
  std::vector<double> tab_key;
  std::vector<double> table_key() const {
    std::vector<double> k = {(double)plan.form, (double)plan.a, (double)plan.b, (double)plan.t,
                             (double)(uintptr_t)plan.pr, (double)(uintptr_t)plan.pk, (double)(uintptr_t)plan.sy,
                             (double)plan.rt[0], (double)plan.rt[1], (double)plan.rt[2], prm.k0_nm,
                             prm.fourier_res[0], prm.fourier_res[1], prm.fourier_res[2], prm.boundary_sharpness,
                             (double)prm.acyclic, (double)prm.shift_axes, (double)prm.ramp_axes};
    for (int j = 0; j < 6; ++j) k.push_back((double)prm.ROI[j]);
    return k;
  }
  void build_tables() {
    drop_graph();   // captured loops hold the table addresses, which may move here
    tab_key.clear();
    const int a = plan.a, b = plan.b;
    TableSet& T = tabs;
    T.a = a; T.b = b;
    T.tw_x.upload(twiddles(nx));
    T.tw_a.upload(twiddles(a));
    if (plan.form == CBS_FORM_ROWS) T.tw_y.upload(twiddles(ny));
    T.tw_z.upload(twiddles(nz));
    {
      std::vector<float2> cr((size_t)a * b);
      for (int y2 = 0; y2 < b; ++y2)
        for (int p1 = 0; p1 < a; ++p1) {
          const double ang = -2.0 * M_PI * (double)y2 * (double)pos2freq(a, p1) / (double)ny;
          cr[(size_t)y2 * a + p1] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
      T.cross.upload(cr);
    }
    const int dims[3] = {nx, ny, nz};
    DevBuf<float2>* ramps[3] = {&T.ramp_x, &T.ramp_y, &T.ramp_z};
    DevBuf<float>* masks[3] = {&T.mask_x, &T.mask_y, &T.mask_z};
    for (int j = 0; j < 3; ++j) {
      const int n = dims[j];
      // phase ramp exp(-i pi (d - floor(n/2)) / (2 n))  (FORWARD_SOLVER_CONVERGENT_BORN.m:355-377)
      std::vector<float2> rp(n);
      const bool on = prm.acyclic && (prm.ramp_axes >> j & 1u);
      for (int d = 0; d < n; ++d) {
        const double ang = -M_PI * (double)(d - n / 2) / (2.0 * n);
        rp[d] = on ? make_float2((float)std::cos(ang), (float)std::sin(ang)) : make_float2(1.f, 0.f);
      }
      ramps[j]->upload(rp);
      // attenuation mask (create_boundary_RI :251-271)
      const int bpx = (int)prm.ROI[2 * j] - 1;
      std::vector<float> mk(n);
      const double sh = prm.boundary_sharpness;
      for (int d = 0; d < n; ++d) {
        // x(i) = |i - (floor(n/2)+1) + 0.5| - 0.5 taken after circshift(x, -floor(n/2)),
        // i.e. the distance of d to the nearest edge of the axis
        const int dd = (d + n / 2) % n;
        double xv = std::fabs((double)(dd + 1) - std::floor(n / 2.0 + 1.0) + 0.5) - 0.5;
        double val0;
        if (bpx == 0) {
          val0 = 0.0;
        } else {
          xv = xv / ((double)bpx - 0.5);
          val0 = std::fabs(xv) >= 1.0 ? 1.0 : std::fabs(xv);
          val0 = 1.0 - val0;
          if (val0 > sh) val0 = sh;
          val0 = val0 / sh;
        }
        mk[d] = (float)(1.0 - val0);
      }
      masks[j]->upload(mk);
    }
    // Green's function per-axis terms 4 pi^2 (k + sh dk/4)^2 (init :299-351), position ordered,
    // index 1 = flipped grid (fft_flip(...,[1 1 1],false) :580): G1(d) = G0((n - d) mod n).
    auto g0 = [&](int j, int n, int d) {
      const int m = ((d + n / 2) % n) - n / 2;
      const double sh = (prm.acyclic && (prm.shift_axes >> j & 1u)) ? 0.25 : 0.0;
      const double k = ((double)m + sh) * prm.fourier_res[j];
      return 4.0 * M_PI * M_PI * k * k;
    };
    auto gfl = [&](int j, int n, int flip, int d) { return flip ? g0(j, n, (n - d) % n) : g0(j, n, d); };
    // vector runs: rads = (k + sh dk/4) / k0_nm (init :318-321), flipped like Greenp (:582)
    auto r0 = [&](int j, int n, int d) {
      const int m = ((d + n / 2) % n) - n / 2;
      const double sh = (prm.acyclic && (prm.shift_axes >> j & 1u)) ? 0.25 : 0.0;
      return ((double)m + sh) * prm.fourier_res[j] / prm.k0_nm;
    };
    auto rfl = [&](int j, int n, int flip, int d) { return flip ? r0(j, n, (n - d) % n) : r0(j, n, d); };
    // frequency at spectrum position p of axis j: the compiled line FFTs leave a transposed
    // order, the run-time-planned ones the natural order
    auto p2f = [&](int j, int n, int p) { return plan.rt[j] ? p : pos2freq(n, p); };
    {
      std::vector<float> gx(2 * (size_t)nx + 64, 0.f);
      for (int f = 0; f < 2; ++f)
        for (int p = 0; p < nx; ++p) gx[(size_t)f * nx + p] = (float)gfl(0, nx, f, p2f(0, nx, p));
      T.green_x.upload(gx);
      std::vector<float> gy(2 * (size_t)ny, 0.f);
      for (int f = 0; f < 2; ++f)
        for (int k2 = 0; k2 < b; ++k2)
          for (int p1 = 0; p1 < a; ++p1) {
            const int idx = k2 * a + p1;   // plain and four-pass forms: position of the strided y transform
            const int ky = (plan.form == CBS_FORM_FUSED) ? pos2freq(a, p1) + a * k2 : p2f(1, ny, idx);
            gy[(size_t)f * ny + idx] = (float)gfl(1, ny, f, ky);
          }
      T.green_y.upload(gy);
      std::vector<float> gz(2 * (size_t)nz, 0.f);
      for (int f = 0; f < 2; ++f)
        for (int p = 0; p < nz; ++p) gz[(size_t)f * nz + p] = (float)gfl(2, nz, f, p2f(2, nz, p));
      T.green_z.upload(gz);
    }
    if (P == 3) {
      std::vector<float> qx(2 * (size_t)nx + 64, 0.f), qy(2 * (size_t)ny, 0.f), qz(2 * (size_t)nz, 0.f);
      for (int f = 0; f < 2; ++f) {
        for (int p = 0; p < nx; ++p) qx[(size_t)f * nx + p] = (float)rfl(0, nx, f, p2f(0, nx, p));
        for (int k2 = 0; k2 < b; ++k2)
          for (int p1 = 0; p1 < a; ++p1) {
            const int idx = k2 * a + p1;
            const int ky = (plan.form == CBS_FORM_FUSED) ? pos2freq(a, p1) + a * k2 : p2f(1, ny, idx);
            qy[(size_t)f * ny + idx] = (float)rfl(1, ny, f, ky);
          }
        for (int p = 0; p < nz; ++p) qz[(size_t)f * nz + p] = (float)rfl(2, nz, f, p2f(2, nz, p));
      }
      T.rad_x.upload(qx);
      T.rad_y.upload(qy);
      T.rad_z.upload(qz);
    }
    tab_key = table_key();
    sync_vtile();
  }

  // ---- kernel timing
  hipEvent_t event_at(size_t i) {
    while (ev_pool.size() <= i) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[i];
  }
  template <class F>
  void timed(int kind, F&& launch) {
    if (!opt.time_kernels) {
      launch();
      return;
    }
    const size_t i = 2 * ev_used.size();
    hipEvent_t e0 = event_at(i), e1 = event_at(i + 1);
    HIP_CHECK(hipEventRecord(e0, stream));
    launch();
    HIP_CHECK(hipEventRecord(e1, stream));
    ev_used.emplace_back(kind, (int)i);
  }
  void resolve_times() {
    for (int k = 0; k < 3; ++k) { ktime_ms[k] = 0; kcount[k] = 0; }
    for (auto& u : ev_used) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, ev_pool[u.second], ev_pool[u.second + 1]));
      ktime_ms[u.first] += ms;
      kcount[u.first] += 1;
    }
    ev_used.clear();
  }

  // ---- launches
  StateArgs state_args(int nb) {
    (void)nb;
    StateArgs s;
    const bool fused = plan.form == CBS_FORM_FUSED;
    s.V = fused && !v_real ? Vt.p : V.p;
    s.Vr = fused && v_real ? Vr.p : nullptr;
    s.field = ws.field; s.fstate = ws.fstate; s.spec = ws.spec; s.parity = ws.parity;
    s.rhs_stride = nvox;
    s.norms = nullptr;
    s.fref = nullptr;
    return s;
  }
  // convergence test arguments for a batch of nv component slots (norms = null: off)
  ConvArgs conv_args(int nv, int M) const {
    ConvArgs cv;
    cv.norms = conv_tol > 0.0 ? ws.norms : nullptr;
    cv.fref = ws.fref;
    cv.tol2 = conv_tol * conv_tol;
    cv.P = P;
    cv.nslots = nv;
    cv.M = M;
    return cv;
  }
  // norm accumulators of the launch that completes iteration jj_inv (null: not tracked)
  double* norm_row(int jj_inv, int nv) const {
    return (tracking() && jj_inv > 0) ? ws.norms + (size_t)(jj_inv - 1) * nv * kNormSlots : nullptr;
  }

  static int upd_of(int jj, int M) { return upd_kind(jj, M); }
  // nontemporal state access only when Field, f and the spectrum of nv slots plus V
  // exceed what the 256 MB Infinity Cache can keep between launches
  int stream_hint(int nv) const { return (3.0 * nv + 1.0) * (double)nvox * sizeof(float2) > 192.0e6 ? 1 : 0; }
  int fwd_of(int jj, int M) const { return jj == M ? FWD_NONE : (jj == 1 ? FWD_P0 : FWD_GAMMA); }

  // fused schedule for one batch already seeded in fstate
  void run_fused(int nb, int M) {
    const PassRInfo* pr = plan.pr;
    const PassKInfo* pk = plan.pk;
    const int nv = nb * P;                 // component slots: the real-space pass treats each as an rhs
    const Consts c = consts(nv);
    const Consts ck = consts(nb);
    const Tables tb = tabs.view();
    const ConvArgs cv = conv_args(nv, M);
    auto pass_r = [&](int upd, int fwd, int jj_inv, int jj_fwd) {
      PassRArgs args;
      args.tb = tb; args.c = c; args.st = state_args(nb);
      args.st.norms = norm_row(jj_inv, nv);
      if (conv_tol > 0.0 && upd == UPD_VQ_ADD) args.st.fref = ws.fref;
      args.upd = upd; args.fwd = fwd; args.jj_inv = jj_inv; args.jj_fwd = jj_fwd;
      args.order = (opt.block_order & 4) ? 2 : (opt.block_order & 1);
      args.nt = stream_hint(nv);
      args.cv = cv;
      const dim3 grid((unsigned)((long long)nv * plan.b * nz));
      const bool steady = (upd == UPD_DEFER || upd == UPD_PAIR) && fwd == FWD_GAMMA;
      const bool vr = args.st.Vr != nullptr;   // real potential: its 4-byte form
      const PassRFn fn = steady ? (vr ? pr->fn_steady_vr : pr->fn_steady) : (vr ? pr->fn_vr : pr->fn);
      require(fn != nullptr, CBS_ERR_UNSUPPORTED, "no real-potential form of the real-space pass");
      timed(0, [&] { hipLaunchKernelGGL(fn, grid, dim3(pr->nt), 0, stream, args); });
    };
    auto pass_k = [&](int jj) {
      PassKArgs args;
      args.tb = tb; args.c = ck; args.spec = ws.spec; args.parity = ws.parity; args.rhs_stride = nvox; args.jj = jj;
      args.order = (opt.block_order >> 1) & 1;
      args.nt = stream_hint(nv);
      args.ntl = args.nt && (opt.block_order & 8) ? 1 : 0;
      args.cv = cv;
      const int ntile = (nx + pk->t - 1) / pk->t;
      const dim3 grid((unsigned)((long long)nb * plan.a * ntile));
      timed(1, [&] { hipLaunchKernelGGL(pk->fn, grid, dim3(pk->nt), 0, stream, args); });
    };
    pass_r(UPD_NONE, FWD_P0, 0, 1);
    for (int jj = 1; jj <= M; ++jj) {
      pass_k(jj);
      const int fwd = fwd_of(jj, M);
      pass_r(upd_of(jj, M), fwd, jj, fwd == FWD_NONE ? 0 : jj + 1);
    }
    HIP_CHECK(hipGetLastError());
  }

  // four-pass schedule for one batch already seeded in fstate: the real-space pass on
  // row blocks (x transforms + pointwise), the forward y transform, the k-space pass
  // (z transforms around the Green's multiply, b = 1), the inverse y transform
  void run_rows(int nb, int M) {
    const PassRInfo* pr = plan.pr;
    const PassKInfo* pk = plan.pk;
    const StridedInfo* sy = plan.sy;
    const int nv = nb * P;
    const Consts c = consts(nv);
    Consts ck = consts(nb);
    ck.a = ny;   // k-space pass: every y is its own row, natural layout
    ck.b = 1;
    const Tables tb = tabs.view();
    const ConvArgs cv = conv_args(nv, M);
    auto pass_r = [&](int upd, int fwd, int jj_inv, int jj_fwd) {
      PassRArgs args;
      args.tb = tb; args.c = c; args.st = state_args(nb);
      args.st.norms = norm_row(jj_inv, nv);
      if (conv_tol > 0.0 && upd == UPD_VQ_ADD) args.st.fref = ws.fref;
      args.upd = upd; args.fwd = fwd; args.jj_inv = jj_inv; args.jj_fwd = jj_fwd;
      args.order = 0;
      args.nt = stream_hint(nv);
      args.cv = cv;
      const dim3 grid((unsigned)((long long)nv * plan.b * nz));
      const bool steady = (upd == UPD_DEFER || upd == UPD_PAIR) && fwd == FWD_GAMMA;
      const PassRFn fn = steady ? pr->fn_steady : pr->fn;
      timed(0, [&] { hipLaunchKernelGGL(fn, grid, dim3(pr->nt), 0, stream, args); });
    };
    auto pass_k = [&](int jj) {
      PassKArgs args;
      args.tb = tb; args.c = ck; args.spec = ws.spec; args.parity = ws.parity; args.rhs_stride = nvox; args.jj = jj;
      args.order = 0;
      args.nt = stream_hint(nv);
      args.ntl = 0;
      args.cv = cv;
      const int ntile = (nx + pk->t - 1) / pk->t;
      const dim3 grid((unsigned)((long long)nb * ny * ntile));
      timed(1, [&] { hipLaunchKernelGGL(pk->fn, grid, dim3(pk->nt), 0, stream, args); });
    };
    auto fft_y = [&](int inverse) {
      // all z planes of all component slots: consecutive planes are nx * ny apart
      const int nt = (nx + sy->t - 1) / sy->t;
      const long long nouter = (long long)nz * nv;
      const dim3 g((unsigned)(nt * nouter));
      timed(2, [&] { hipLaunchKernelGGL(sy->fn, g, dim3(sy->nt), 0, stream, ws.spec, nx, (long long)nx,
                                        (long long)nx * ny, (int)nouter, tabs.tw_y.p, inverse, (int)nouter, 0ll); });
    };
    pass_r(UPD_NONE, FWD_P0, 0, 1);
    for (int jj = 1; jj <= M; ++jj) {
      fft_y(0);
      pass_k(jj);
      fft_y(1);
      const int fwd = fwd_of(jj, M);
      pass_r(upd_of(jj, M), fwd, jj, fwd == FWD_NONE ? 0 : jj + 1);
    }
    HIP_CHECK(hipGetLastError());
  }

  // per-axis schedule for a batch of nb rhs (nv = nb P component slots): every launch covers
  // the whole batch (the rhs or slot on the grid's y dimension, or as an outer index of the
  // line transforms)
  void run_plain(int nb, int M) {
    const int nv = nb * P;
    const Consts c = consts(1);
    const Tables tb = tabs.view();
    const unsigned pw_grid = std::min<unsigned>(grid_for(nvox), 65535u);
    const ConvArgs cv = conv_args(nv, M);
    auto update = [&](int upd, int fwd, int jj_inv, int jj_fwd) {
      StateArgs s = state_args(1);
      s.norms = norm_row(jj_inv, nv);
      if (conv_tol > 0.0 && upd == UPD_VQ_ADD) s.fref = ws.fref;
      timed(2, [&] { hipLaunchKernelGGL(update_plain, dim3(pw_grid, (unsigned)nv), dim3(256), 0, stream, s, tb, c, upd,
                                        fwd, jj_inv, jj_fwd, cv); });
    };
    // the line transforms of one axis (0 x, 1 y, 2 z) over every slot
    auto fft1 = [&](int axis, int inverse) {
      float2* sp = ws.spec;
      const RowsInfo* rw = plan.rw;
      const StridedInfo* sy = plan.sy;
      const StridedInfo* sz = plan.sz;
      auto fx = [&] {
        const long long nrows = (long long)ny * nz * nv;
        if (plan.rt[0]) return rt_rows(rtp[0], sp, nrows, inverse);
        const long long nblk = (nrows + rw->rows - 1) / rw->rows;
        require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "x transform batch too large for one launch");
        timed(2, [&] { hipLaunchKernelGGL(rw->fn, dim3((unsigned)nblk), dim3(rw->nt), 0, stream, sp, nrows,
                                          tabs.tw_x.p, inverse); });
      };
      auto fy = [&] {   // outer: the z planes of every slot, nx * ny apart
        const long long nouter = (long long)nz * nv;
        if (plan.rt[1]) return rt_strided(rtp[1], sp, nx, (long long)nx, (long long)nx * ny, nouter, inverse);
        const long long nblk = (long long)((nx + sy->t - 1) / sy->t) * nouter;
        require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "y transform batch too large for one launch");
        timed(2, [&] { hipLaunchKernelGGL(sy->fn, dim3((unsigned)nblk), dim3(sy->nt), 0, stream, sp, nx,
                                          (long long)nx, (long long)nx * ny, (int)nouter, tabs.tw_a.p, inverse,
                                          (int)nouter, 0ll); });
      };
      auto fz = [&] {   // outer: y (nx apart) within a slot, then the slots (nvox apart)
        const long long nouter = (long long)ny * nv;
        if (plan.rt[2])
          return rt_strided(rtp[2], sp, nx, (long long)nx * ny, (long long)nx, nouter, inverse, ny, nvox);
        const long long nblk = (long long)((nx + sz->t - 1) / sz->t) * nouter;
        require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "z transform batch too large for one launch");
        timed(2, [&] { hipLaunchKernelGGL(sz->fn, dim3((unsigned)nblk), dim3(sz->nt), 0, stream, sp, nx,
                                          (long long)nx * ny, (long long)nx, (int)nouter, tabs.tw_z.p, inverse, ny,
                                          (long long)nvox); });
      };
      if (axis == 0) fx();
      else if (axis == 1) fy();
      else fz();
    };
    // single-component runs whose x (z) lines are Bluestein-transformed on a compiled
    // length: the x inverse, the update and the x forward (the z forward, the Green's
    // multiply and the z inverse) run as one launch each
    // (on a compiled line length, the compiled fused passes of lines_fused_inst.hip)
    const bool fuse = P == 1 && plan.form == CBS_FORM_PLAIN;   // CBS_FORM_AXES: every pass on its own
    const BlueInfo* bx = (fuse && plan.rt[0]) ? rtp[0].bi : nullptr;
    const BlueInfo* bz = (fuse && plan.rt[2]) ? rtp[2].bi : nullptr;
    const RowsUpdateInfo* cx = (fuse && !plan.rt[0]) ? find_rows_update(nx) : nullptr;
    const StridedGreenInfo* cz = (fuse && !plan.rt[2]) ? find_strided_green(nz) : nullptr;
    const long long rows = (long long)ny * nz;
    require(rows < (1ll << 31), CBS_ERR_UNSUPPORTED, "grid too large for the per-axis passes");
    auto xpass = [&](int upd, int fwd, int jj_inv, int jj_fwd) {   // (x inverse,) update(, x forward)
      if (!bx && !cx) {
        if (upd != UPD_NONE) fft1(0, 1);
        update(upd, fwd, jj_inv, jj_fwd);
        if (fwd != FWD_NONE) fft1(0, 0);
        return;
      }
      StateArgs s = state_args(1);
      s.norms = norm_row(jj_inv, nv);
      if (conv_tol > 0.0 && upd == UPD_VQ_ADD) s.fref = ws.fref;
      const int nl = bx ? bx->nl : cx->rows;
      const dim3 g((unsigned)((rows + nl - 1) / nl), (unsigned)nv);
      if (bx)
        timed(2, [&] { hipLaunchKernelGGL(bx->rows_update, g, dim3(bx->nt), 0, stream, s, tb, c, rtp[0].ba, upd, fwd,
                                          jj_inv, jj_fwd, cv); });
      else
        timed(2, [&] { hipLaunchKernelGGL(cx->fn, g, dim3(cx->nt), 0, stream, s, tb, c, tabs.tw_x.p, upd, fwd,
                                          jj_inv, jj_fwd, cv); });
    };
    auto zpass = [&](int jj) {   // z forward, Green's multiply, z inverse
      if (cz) {
        const long long nblk = (long long)((nx + cz->t - 1) / cz->t) * ny * nv;
        require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "z transform batch too large for one launch");
        timed(2, [&] { hipLaunchKernelGGL(cz->fn, dim3((unsigned)nblk), dim3(cz->nt), 0, stream, ws.spec, tb, c,
                                          tabs.tw_z.p, ws.parity, jj, cv); });
        return;
      }
      if (!bz) {
        fft1(2, 0);
        if (P == 1)
          timed(2, [&] { hipLaunchKernelGGL(green_plain, dim3(pw_grid, (unsigned)nb), dim3(256), 0, stream, ws.spec,
                                            tb, c, ws.parity, jj, cv); });
        else
          timed(2, [&] { hipLaunchKernelGGL(green_plain_vec, dim3(pw_grid, (unsigned)nb), dim3(256), 0, stream,
                                            ws.spec, nvox, tb, c, ws.parity, jj, cv); });
        fft1(2, 1);
        return;
      }
      const long long nblk = (long long)((nx + bz->t - 1) / bz->t) * ny * nv;
      require(nblk < (1ll << 31), CBS_ERR_UNSUPPORTED, "z transform batch too large for one launch");
      timed(2, [&] { hipLaunchKernelGGL(bz->strided_green, dim3((unsigned)nblk), dim3(bz->nt), 0, stream, ws.spec,
                                        tb, c, rtp[2].ba, ws.parity, jj, cv); });
    };
    xpass(UPD_NONE, FWD_P0, 0, 1);
    for (int jj = 1; jj <= M; ++jj) {
      fft1(1, 0);
      zpass(jj);
      fft1(1, 1);
      const int fwd = fwd_of(jj, M);
      xpass(upd_of(jj, M), fwd, jj, fwd == FWD_NONE ? 0 : jj + 1);
    }
    HIP_CHECK(hipGetLastError());
  }

  // the handle's own state buffers for batches of nb rhs, and ws pointing at them
  void ensure_workspace(int nb, bool host_io, int host_bufs = 1) {
    const size_t nv = (size_t)nb * P;
    const size_t n = nv * nvox;
    field.ensure(n);
    fstate.ensure(n);
    spec.ensure(n);
    parity.ensure(nv);
    if (tracking()) norms.ensure(std::max<size_t>(1, (size_t)prm.Bornmax) * nv * (kNormSlots + 1));
    if (conv_tol > 0.0) fref.ensure(nv * (kNormSlots + 1));
    if (host_io) {
      uin.ensure((size_t)host_bufs * nv * nri);
      uout.ensure((size_t)host_bufs * nv * nri);
    }
    ws.field = field.p;
    ws.fstate = fstate.p;
    ws.spec = spec.p;
    ws.parity = parity.p;
    ws.norms = tracking() ? norms.p : nullptr;
    ws.fref = conv_tol > 0.0 ? fref.p : nullptr;
  }

  uint64_t footprint(uint32_t n_rhs) const {
    const int nv = batch_width(n_rhs) * P;
    uint64_t b = 3ull * nv * nvox * sizeof(float2);   // field, fstate, spec
    b += 2ull * nv * nri * sizeof(float2);             // staging for host I/O
    b += (uint64_t)nvox * sizeof(float2);   // V
    if (plan.form == CBS_FORM_FUSED) b += (uint64_t)nvox * (v_real ? sizeof(float) : sizeof(float2));   // its tile-order copy
    b += io.footprint(nv, cx, cy, rz) + (io.ready ? (unsigned long long)nv * nri * 16ull : 0ull);
    return b;
  }
  int batch_width(uint32_t n_rhs) const {
    // the per-axis form's launches put the component slots on grid y (at most 65535)
    long long nb = std::min<long long>((long long)n_rhs, 65535 / P);
    if (opt.max_batch > 0 && opt.max_batch < nb) nb = opt.max_batch;
    return (int)nb;
  }

  // Born loop for a batch whose incident fields sit at d_uin (device, RI grid).
  // Born loop for a batch whose incident fields sit at d_uin (device, RI grid), on the
  // state buffers ws points at.  no_alloc: no graph capture and no kernel timing (both
  // create runtime objects), for callers that must not allocate.
  void born_batch(const float2* d_uin, float2* d_uout, int nb, int rhs_index0, bool no_alloc = false) {
    const int M = (int)prm.Bornmax;
    const int nv = nb * P;
    if (last_iters.size() < (size_t)(rhs_index0 + nb)) last_iters.resize((size_t)(rhs_index0 + nb), 0u);
    if (tracking() && last_norms.size() < (size_t)(rhs_index0 + nb) * M)
      last_norms.resize((size_t)(rhs_index0 + nb) * M, 0.0);
    // the device copy is formed in stream order, after the previous batch's kernels
    hipLaunchKernelGGL(fill_parity, dim3(grid_for(nv)), dim3(256), 0, stream, ws.parity, nv, P, flip_state,
                       rhs_index0, M);
    const Consts c = consts(nv);
    const RoiGeom g = roi();
    if (tracking() && M > 0) {
      const long long words = 2ll * M * nv * kNormSlots;   // FP64 partial sums as 32-bit words
      hipLaunchKernelGGL(zero_words, dim3(grid_for(words)), dim3(256), 0, stream,
                         reinterpret_cast<uint32_t*>(ws.norms), words);
      if (conv_tol > 0.0)
        hipLaunchKernelGGL(zero_words, dim3(grid_for(2ll * nv * kNormSlots)), dim3(256), 0, stream,
                           reinterpret_cast<uint32_t*>(ws.fref), 2ll * nv * kNormSlots);
    }
    {
      const long long tot = (long long)nv * nvox;
      timed(2, [&] { hipLaunchKernelGGL(seed_source, dim3(grid_for(tot)), dim3(256), 0, stream, d_uin,
                                        V.p, ws.fstate, c, g, nvox, nv); });
    }
    if (M == 0) {
      const long long words = 2ll * nv * nvox;
      hipLaunchKernelGGL(zero_words, dim3(grid_for(words)), dim3(256), 0, stream,
                         reinterpret_cast<uint32_t*>(ws.field), words);
    } else {
      if (opt.use_graph && !opt.time_kernels && !no_alloc) {
        // replay a captured iteration loop (the kernel arguments, including the
        // potential's scalars, are baked in; the cache is dropped whenever the
        // potential, the options or the plan change)
        const std::vector<long long> key = {nb, M, (long long)ws.field, (long long)ws.fstate, (long long)ws.spec,
                                            (long long)ws.parity, (long long)ws.norms, (long long)ws.fref,
                                            (long long)(conv_tol > 0.0), (long long)plan.form,
                                            (long long)plan.pr, (long long)plan.pk, (long long)opt.block_order};
        GraphSlot* gs = nullptr;
        for (auto& g : gslot)
          if (g.exec && g.key == key) gs = &g;
        if (!gs) {
          gs = &gslot[gslot_next];
          gslot_next ^= 1;
          std::lock_guard<std::recursive_mutex> lk(device_mutex());   // no device-wide sync while capturing
          if (gs->exec) (void)hipGraphExecDestroy(gs->exec);
          if (gs->graph) (void)hipGraphDestroy(gs->graph);
          *gs = GraphSlot();
          HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
          run_loop(nb, M);
          hipGraph_t captured = nullptr;
          hipError_t e = hipStreamEndCapture(stream, &captured);
          if (e != hipSuccess) throw CbsError(CBS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
          gs->graph = captured;
          HIP_CHECK(hipGraphInstantiate(&gs->exec, gs->graph, nullptr, nullptr, 0));
          gs->key = key;
        }
        HIP_CHECK(hipGraphLaunch(gs->exec, stream));
      } else {
        run_loop(nb, M);
      }
    }
    {
      const long long tot = (long long)nv * nri;
      timed(2, [&] { hipLaunchKernelGGL(finalize_field, dim3(grid_for(tot)), dim3(256), 0, stream,
                                        ws.field, d_uin, d_uout, c, g, nvox, nv); });
    }
    HIP_CHECK(hipGetLastError());
    if (tracking() && M > 0) {
      // the totals go after the partial sums
      std::vector<double> tmp((size_t)M * nv), ref((size_t)nv, 0.0);
      double* tot = ws.norms + (size_t)M * nv * kNormSlots;
      hipLaunchKernelGGL(sum_slots, dim3(grid_for((long long)M * nv)), dim3(256), 0, stream, ws.norms, tot,
                         (long long)M * nv);
      HIP_CHECK(hipMemcpyAsync(tmp.data(), tot, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      if (conv_tol > 0.0) {
        double* rtot = ws.fref + (size_t)nv * kNormSlots;
        hipLaunchKernelGGL(sum_slots, dim3(grid_for(nv)), dim3(256), 0, stream, ws.fref, rtot, (long long)nv);
        HIP_CHECK(hipMemcpyAsync(ref.data(), rtot, ref.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
      }
      HIP_CHECK(hipStreamSynchronize(stream));
      for (int r = 0; r < nb; ++r) {
        double fr = 0.0;
        for (int q = 0; q < P; ++q) fr += ref[(size_t)P * r + q];
        uint32_t stop = (uint32_t)M;
        for (int j = 0; j < M; ++j) {
          double acc = 0.0;   // |f_n|^2 summed over the field components
          for (int q = 0; q < P; ++q) acc += tmp[(size_t)j * nv + (size_t)P * r + q];
          last_norms[(size_t)(rhs_index0 + r) * M + j] = acc;
          // the device's test (conv_stopped) on the same FP64 values
          if (conv_tol > 0.0 && stop == (uint32_t)M && j + 1 >= 3 && acc <= conv_tol * conv_tol * fr)
            stop = (uint32_t)(j + 1);
        }
        last_iters[(size_t)(rhs_index0 + r)] = stop;
      }
    } else {
      for (int r = 0; r < nb; ++r) last_iters[(size_t)(rhs_index0 + r)] = (uint32_t)M;
    }
  }

  void ensure_copy_stream() {
    if (cstream) return;
    std::lock_guard<std::recursive_mutex> lk(device_mutex());
    HIP_CHECK(hipStreamCreateWithFlags(&cstream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
      HIP_CHECK(hipEventCreateWithFlags(&ev_up[k], hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&ev_done[k], hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&ev_down[k], hipEventDisableTiming));
    }
  }

  // Host buffers, K >= 2 chunks of nbc right-hand sides: the upload of chunk k+1 and the
  // download of chunk k-1 run on the copy stream while chunk k computes, through two
  // device staging buffers each way.
  void solve_3d_host_pipelined(const float2* u_in, uint32_t n_rhs, float2* u_out, int nbc) {
    ensure_copy_stream();
    const size_t per = (size_t)P * nri;            // elements per rhs ([rhs][component][z][y][x])
    const size_t half = (size_t)nbc * per;         // one staging buffer
    const int K = (int)((n_rhs + (uint32_t)nbc - 1) / (uint32_t)nbc);
    auto rhs_of = [&](int k) { return (int)std::min<uint32_t>((uint32_t)nbc, n_rhs - (uint32_t)k * (uint32_t)nbc); };
    HIP_CHECK(hipMemcpyAsync(uin.p, u_in, (size_t)rhs_of(0) * per * sizeof(float2), hipMemcpyHostToDevice, cstream));
    HIP_CHECK(hipEventRecord(ev_up[0], cstream));
    for (int k = 0; k < K; ++k) {
      const int b = k & 1, o = b ^ 1;
      HIP_CHECK(hipStreamWaitEvent(stream, ev_up[b], 0));                   // chunk k is on the device
      if (k >= 2) HIP_CHECK(hipStreamWaitEvent(stream, ev_down[b], 0));     // chunk k-2 left uout[b]
      born_batch(uin.p + b * half, uout.p + b * half, rhs_of(k), k * nbc);
      HIP_CHECK(hipEventRecord(ev_done[b], stream));
      if (k + 1 < K) {                                                      // chunk k-1 no longer reads uin[o]
        if (k >= 1) HIP_CHECK(hipStreamWaitEvent(cstream, ev_done[o], 0));
        HIP_CHECK(hipMemcpyAsync(uin.p + o * half, u_in + (size_t)(k + 1) * half,
                                 (size_t)rhs_of(k + 1) * per * sizeof(float2), hipMemcpyHostToDevice, cstream));
        HIP_CHECK(hipEventRecord(ev_up[o], cstream));
      }
      if (k >= 1) {
        HIP_CHECK(hipStreamWaitEvent(cstream, ev_done[o], 0));
        HIP_CHECK(hipMemcpyAsync(u_out + (size_t)(k - 1) * half, uout.p + o * half,
                                 (size_t)rhs_of(k - 1) * per * sizeof(float2), hipMemcpyDeviceToHost, cstream));
        HIP_CHECK(hipEventRecord(ev_down[o], cstream));
      }
    }
    const int bl = (K - 1) & 1;
    HIP_CHECK(hipStreamWaitEvent(cstream, ev_done[bl], 0));
    HIP_CHECK(hipMemcpyAsync(u_out + (size_t)(K - 1) * half, uout.p + bl * half,
                             (size_t)rhs_of(K - 1) * per * sizeof(float2), hipMemcpyDeviceToHost, cstream));
    HIP_CHECK(hipStreamSynchronize(cstream));
  }

  void solve_3d(const float2* u_in, uint32_t n_rhs, float2* u_out, int mem) {
    require(have_potential, CBS_ERR_STATE, "cbs_set_potential must be called before solving");
    require(n_rhs >= 1, CBS_ERR_ARG, "n_rhs must be >= 1");
    require(u_in && u_out, CBS_ERR_ARG, "null field pointer");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    int nbw = batch_width(n_rhs);
    // host buffers: at least two chunks of >= 8 rhs (about six for large batches) so that
    // the PCIe copies overlap the solve
    bool pipe = false;
    if (mem == CBS_MEM_HOST) {
      const int nbc = std::min(nbw, std::max((int)((n_rhs + 5) / 6), std::min((int)n_rhs, 8)));
      if ((int)n_rhs > nbc) {
        pipe = true;
        nbw = nbc;
      }
    }
    ensure_workspace(nbw, mem == CBS_MEM_HOST, pipe ? 2 : 1);
    const int M = (int)prm.Bornmax;
    begin_solve(n_rhs, M);
    ev_used.clear();
    if (pipe) {
      solve_3d_host_pipelined(u_in, n_rhs, u_out, nbw);
      HIP_CHECK(hipStreamSynchronize(stream));
      if (opt.time_kernels) resolve_times();
      if (prm.acyclic) flip_state = (int)((flip_state + (long long)n_rhs * M) & 1);
      return;
    }
    for (uint32_t r0 = 0; r0 < n_rhs; r0 += nbw) {
      const int nb = (int)std::min<uint32_t>(nbw, n_rhs - r0);
      const float2* din;
      float2* dout;
      const size_t off = (size_t)r0 * P * nri, cnt = (size_t)nb * P * nri;   // [rhs][component][z][y][x]
      if (mem == CBS_MEM_HOST) {
        HIP_CHECK(hipMemcpyAsync(uin.p, u_in + off, cnt * sizeof(float2), hipMemcpyHostToDevice, stream));
        din = uin.p;
        dout = uout.p;
      } else {
        din = u_in + off;
        dout = u_out + off;
      }
      born_batch(din, dout, nb, (int)r0);
      if (mem == CBS_MEM_HOST) {
        HIP_CHECK(hipMemcpyAsync(u_out + off, uout.p, cnt * sizeof(float2), hipMemcpyDeviceToHost, stream));
      }
    }
    HIP_CHECK(hipStreamSynchronize(stream));
    if (opt.time_kernels) resolve_times();
    if (prm.acyclic) flip_state = (int)((flip_state + (long long)n_rhs * M) & 1);
  }

  // cbs_solve_3d on device buffers with the state in the no-allocation workspace scr
  // (kScratchBytes): no device allocation, no graph capture, no timing events.
  void solve_3d_noalloc(const float2* u_in, uint32_t n_rhs, float2* u_out, unsigned char* scr) {
    require(have_potential, CBS_ERR_STATE, "cbs_set_potential must be called before solving");
    require(n_rhs >= 1, CBS_ERR_ARG, "n_rhs must be >= 1");
    require(u_in && u_out, CBS_ERR_ARG, "null field pointer");
    const int M = (int)prm.Bornmax;
    const int nbw = batch_width(n_rhs);
    const size_t nv = (size_t)nbw * P;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_par = off;
    off += up(nv);
    const size_t o_nrm = off;
    if (tracking()) off += up((size_t)std::max(M, 1) * nv * (kNormSlots + 1) * sizeof(double));
    const size_t o_ref = off;
    if (conv_tol > 0.0) off += up(nv * (kNormSlots + 1) * sizeof(double));
    const size_t state = up(nv * (size_t)nvox * sizeof(float2));
    const size_t o_state = off;
    off += 3 * state;
    if (off > kScratchBytes) {
      const size_t per_rhs = 3 * (size_t)P * (size_t)nvox * sizeof(float2);
      throw CbsError(CBS_ERR_MEMORY, "cbs_solve_3d_noalloc: a batch of " + std::to_string(nbw) + " rhs on the " +
                                         shape_str() + " grid needs " + std::to_string(off) + " bytes, over the " +
                                         std::to_string(kScratchBytes) + "-byte no-allocation workspace (about " +
                                         std::to_string(kScratchBytes / per_rhs) +
                                         " rhs of this grid fit; set max_batch or call cbs_solve_3d)");
    }
    ws.parity = scr + o_par;
    ws.norms = tracking() ? reinterpret_cast<double*>(scr + o_nrm) : nullptr;
    ws.fref = conv_tol > 0.0 ? reinterpret_cast<double*>(scr + o_ref) : nullptr;
    ws.field = reinterpret_cast<float2*>(scr + o_state);
    ws.fstate = reinterpret_cast<float2*>(scr + o_state + state);
    ws.spec = reinterpret_cast<float2*>(scr + o_state + 2 * state);
    const int saved_time = opt.time_kernels;
    opt.time_kernels = 0;
    try {
      begin_solve(n_rhs, M);
      for (uint32_t r0 = 0; r0 < n_rhs; r0 += (uint32_t)nbw) {
        const int nb = (int)std::min<uint32_t>((uint32_t)nbw, n_rhs - r0);
        const size_t o = (size_t)r0 * P * nri;
        born_batch(u_in + o, u_out + o, nb, (int)r0, /*no_alloc=*/true);
      }
      HIP_CHECK(hipStreamSynchronize(stream));
    } catch (...) {
      opt.time_kernels = saved_time;
      ws = Ws();
      throw;
    }
    opt.time_kernels = saved_time;
    ws = Ws();     // the workspace belongs to the GPU, not to this handle
    if (prm.acyclic) flip_state = (int)((flip_state + (long long)n_rhs * M) & 1);
  }

  void run_loop(int nb, int M) {
    if (plan.form == CBS_FORM_ROWS) run_rows(nb, M);
    else if (plan.form == CBS_FORM_PLAIN || plan.form == CBS_FORM_AXES) run_plain(nb, M);
    else run_fused(nb, M);
  }

  // per-solve records: norms (when tracked) and the iterations each rhs ran
  void begin_solve(uint32_t n_rhs, int M) {
    last_iters.assign((size_t)n_rhs, (uint32_t)M);
    last_norms.clear();
    if (tracking()) last_norms.assign((size_t)n_rhs * M, 0.0);
    last_norm_rhs = tracking() ? (int)n_rhs : 0;
    last_norm_iter = M;
  }

  static unsigned grid_for(long long n) {
    const long long b = (n + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(b, 1 << 20));
  }

  // ---- 2D input/output projections
  void set_io_kernels(const float2* refocus, const float2* ktrans, const float2* kref, int mem) {
    require(refocus != nullptr && ktrans != nullptr, CBS_ERR_ARG, "refocus and kernel_trans are required");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    require(fx == rx && fy == ry, CBS_ERR_ARG,
            "2D projections need the field xy size equal to the RI xy size (as the reference's crop_field2RI assumes)");
    require(cx >= fx && cy >= fy, CBS_ERR_ARG, "conv_size must be at least field_size");
    // FORWARD_SOLVER_CONVERGENT_BORN.m:173-179 crops parameters.size out of an array of that same size,
    // offset by RI_center, so every 2D solve with RI_center other than [0 0] stops there (:554)
    require(rc_x == 0 && rc_y == 0, CBS_ERR_ARG,
            "RI_center other than [0 0]: the reference's crop_field2RI indexes outside the field grid");
    io_rows = find_fft_rows(cx);
    io_cols = find_fft_strided(cy, 0);
    if (!io_rows) build_rt(io_rtx, cx);    // lengths without a compiled kernel: run-time plans
    if (!io_cols) build_rt(io_rty, cy);
    const size_t n = (size_t)cx * cy * rz;
    const hipMemcpyKind kind = mem == CBS_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    io_refocus_k.ensure(n);
    HIP_CHECK(hipMemcpy(io_refocus_k.p, refocus, n * sizeof(float2), kind));
    io_ktrans.ensure(n);
    HIP_CHECK(hipMemcpy(io_ktrans.p, ktrans, n * sizeof(float2), kind));
    io.have_ref = kref != nullptr;
    if (kref) {
      io_kref.ensure(n);
      HIP_CHECK(hipMemcpy(io_kref.p, kref, n * sizeof(float2), kind));
    }
    std::vector<int> kx(cx), ky(cy);
    for (int p = 0; p < cx; ++p) kx[p] = ((io_rows ? pos2freq(cx, p) : p) + cx / 2) % cx;
    for (int p = 0; p < cy; ++p) ky[p] = ((io_cols ? pos2freq(cy, p) : p) + cy / 2) % cy;
    io_kcx.upload(kx);
    io_kcy.upload(ky);
    io_tw_cx.upload(twiddles(cx));
    io_tw_cy.upload(twiddles(cy));
    iog.fx = fx; iog.fy = fy; iog.cx = cx; iog.cy = cy; iog.rx = rx; iog.ry = ry; iog.rz = rz;
    iog.pad_x = cx / 2 - fx / 2; iog.pad_y = cy / 2 - fy / 2;
    iog.ri_x = iog.pad_x + rc_x; iog.ri_y = iog.pad_y + rc_y;
    require(iog.ri_x >= 0 && iog.ri_y >= 0 && iog.ri_x + rx <= cx && iog.ri_y + ry <= cy, CBS_ERR_ARG,
            "RI_center moves the RI grid outside the convolution grid");
    io.ready = true;
  }

  // First-octant kernels of CudaBorn action 0 (FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m:28-35:
  // refocusing_util and kernel_trans over floor(conv/2)+1 x floor(conv/2)+1 x floor(RI_z/2)+1,
  // the centred (kx, ky) quarter from the most negative frequency to zero, z from the first
  // plane to the middle one).  Expanded to the whole (conv_x, conv_y, RI_z) kernels with the
  // symmetry the reference's field_in/field_out kernels read them with (execution2.cu:195-275):
  // centred index c takes octant index conv/2 - |c - conv/2| on each axis; z takes
  // RI_z/2 - |z - RI_z/2|, conjugated for the refocusing and negated and conjugated for the
  // transmission from the middle plane on; the reflection kernel is the conjugated
  // transmission kernel, as field_out_kernel forms it.  (The y index uses conv_y; the
  // reference's conv_x there, quirk Q1, is the same for the square grids it supports.)
  void set_io_kernels_octant(const float2* r_oct, const float2* t_oct, int mem) {
    require(r_oct != nullptr && t_oct != nullptr, CBS_ERR_ARG, "refocus and kernel_trans octants are required");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    const int gx = cx / 2 + 1, gy = cy / 2 + 1, gz = rz / 2 + 1;
    const size_t no = (size_t)gx * gy * gz;
    std::vector<float2> ro(no), to(no);
    const hipMemcpyKind kind = mem == CBS_MEM_HOST ? hipMemcpyHostToHost : hipMemcpyDeviceToHost;
    HIP_CHECK(hipMemcpy(ro.data(), r_oct, no * sizeof(float2), kind));
    HIP_CHECK(hipMemcpy(to.data(), t_oct, no * sizeof(float2), kind));
    const size_t n = (size_t)cx * cy * rz;
    std::vector<float2> R(n), T(n), K(n);
    for (int z = 0; z < rz; ++z) {
      const int g3 = rz / 2 - std::abs(z - rz / 2);
      const bool upper = z >= rz / 2;
      for (int c2 = 0; c2 < cy; ++c2) {
        const int g2 = cy / 2 - std::abs(c2 - cy / 2);
        for (int c1 = 0; c1 < cx; ++c1) {
          const int g1 = cx / 2 - std::abs(c1 - cx / 2);
          const size_t o = (size_t)g1 + (size_t)gx * ((size_t)g2 + (size_t)gy * g3);
          const size_t i = (size_t)c1 + (size_t)cx * ((size_t)c2 + (size_t)cy * z);
          float2 r = ro[o], t = to[o];
          if (upper) {
            r = make_float2(r.x, -r.y);
            t = make_float2(-t.x, t.y);
          }
          R[i] = r;
          T[i] = t;
          K[i] = make_float2(t.x, -t.y);
        }
      }
    }
    set_io_kernels(R.data(), T.data(), K.data(), CBS_MEM_HOST);
  }

  // ---- centred pupil transforms of the field grid (cbs_field_pupil)
  struct FieldFft {
    int nx = 0, ny = 0;
    const RowsInfo* rw = nullptr;
    const StridedInfo* st = nullptr;
    RtPlan rtx, rty;
    DevBuf<float2> twx, twy;
  } ffp;
  DevBuf<float2> fp_in, fp_add, io_fsrc;
  // the last two pupils (a solve alternates the input and the output one): host copy and
  // position-ordered device table
  struct PupilSlot { std::vector<float> host; DevBuf<float> tab; int nx = 0, ny = 0; long long used = 0; } fp_slot[2];
  long long fp_clock = 0;

  // the device table of a host pupil; the least recently used slot is replaced, so two
  // tables looked up one after the other both stay valid
  const float* pupil_table(const float* pupil) {
    const size_t n2 = (size_t)fx * fy;
    for (auto& sl : fp_slot)
      if (sl.nx == fx && sl.ny == fy && std::memcmp(sl.host.data(), pupil, n2 * sizeof(float)) == 0) {
        sl.used = ++fp_clock;
        return sl.tab.p;
      }
    PupilSlot& sl = fp_slot[0].used <= fp_slot[1].used ? fp_slot[0] : fp_slot[1];
    sl.used = ++fp_clock;
    // the pupil at the frequency each spectrum position holds (compiled transforms leave
    // a transposed order, run-time ones the natural order), over fx fy
    const bool cx_ = find_fft_rows(fx) != nullptr, cy_ = find_fft_strided(fy, 0) != nullptr;
    std::vector<int> kx(fx);
    for (int px = 0; px < fx; ++px) kx[px] = ((cx_ ? pos2freq(fx, px) : px) + fx / 2) % fx;
    std::vector<float> tab(n2);
    const double inv_n = 1.0 / (double)n2;
    for (int py = 0; py < fy; ++py) {
      const int cyi = ((cy_ ? pos2freq(fy, py) : py) + fy / 2) % fy;
      for (int px = 0; px < fx; ++px)
        tab[(size_t)py * fx + px] = (float)((double)pupil[(size_t)cyi * fx + kx[px]] * inv_n);
    }
    sl.tab.upload(tab);
    sl.host.assign(pupil, pupil + n2);
    sl.nx = fx;
    sl.ny = fy;
    return sl.tab.p;
  }

  void field_fft2d(float2* buf, long long nf, int inverse) {
    if (ffp.nx != fx || ffp.ny != fy) {
      ffp.rw = find_fft_rows(fx);
      ffp.st = find_fft_strided(fy, 0);
      if (!ffp.rw) build_rt(ffp.rtx, fx);
      if (!ffp.st) build_rt(ffp.rty, fy);
      ffp.twx.upload(twiddles(fx));
      ffp.twy.upload(twiddles(fy));
      ffp.nx = fx;
      ffp.ny = fy;
    }
    auto rows = [&] {
      const long long nrows = nf * fy;
      if (!ffp.rw) return rt_rows(ffp.rtx, buf, nrows, inverse);
      const dim3 g((unsigned)((nrows + ffp.rw->rows - 1) / ffp.rw->rows));
      timed(2, [&] { hipLaunchKernelGGL(ffp.rw->fn, g, dim3(ffp.rw->nt), 0, stream, buf, nrows, ffp.twx.p, inverse); });
    };
    auto cols = [&] {
      if (!ffp.st) return rt_strided(ffp.rty, buf, fx, (long long)fx, (long long)fx * fy, nf, inverse);
      const int nt = (fx + ffp.st->t - 1) / ffp.st->t;
      require((long long)nt * nf < (1ll << 31), CBS_ERR_UNSUPPORTED, "field transform batch too large for one launch");
      timed(2, [&] { hipLaunchKernelGGL(ffp.st->fn, dim3((unsigned)(nt * nf)), dim3(ffp.st->nt), 0, stream, buf, fx,
                                        (long long)fx, (long long)fx * fy, (int)nf, ffp.twy.p, inverse, (int)nf, 0ll); });
    };
    if (!inverse) { rows(); cols(); }
    else { cols(); rows(); }
  }

  // out = centred_ifft2(P .* centred_fft2(in + add)) over nf fields of the field grid, on
  // device pointers (add may be null, out may be in); tab from pupil_table.  The centring
  // shifts cancel (a shift by n/2 before the forward transform multiplies the spectrum by
  // exp(2 pi i k (n/2) / n), the one after the inverse by its conjugate), so this is
  // ifft2(ifftshift(P) .* fft2(in + add)), in place in out.
  void pupil_apply(const float2* in, const float2* add, long long nf, const float* tab, float2* out) {
    const long long total = (long long)fx * fy * nf;
    const unsigned g = grid_for(total);
    if (add || in != out)
      timed(2, [&] { hipLaunchKernelGGL(field_sum, dim3(g), dim3(256), 0, stream, in, add, out, total); });
    field_fft2d(out, nf, 0);
    timed(2, [&] { hipLaunchKernelGGL(field_pupil_mul, dim3(g), dim3(256), 0, stream, out, tab, fx, fy, nf); });
    field_fft2d(out, nf, 1);
  }

  // P: fx x fy floats, x fastest, in centred (fftshifted) frequency order, on the host
  void field_pupil(const void* in, const void* add, uint32_t nf, const float* pupil, void* out, int mem) {
    require(in && out && pupil, CBS_ERR_ARG, "null field argument");
    require(nf >= 1, CBS_ERR_ARG, "n_fields must be >= 1");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    const long long total = (long long)fx * fy * (long long)nf;
    require(total < (1ll << 40), CBS_ERR_ARG, "field batch too large");
    const float* tab = pupil_table(pupil);
    const float2* din = static_cast<const float2*>(in);
    const float2* dadd = static_cast<const float2*>(add);
    float2* dout = static_cast<float2*>(out);
    if (mem == CBS_MEM_HOST) {
      fp_in.ensure((size_t)total);
      HIP_CHECK(hipMemcpyAsync(fp_in.p, in, (size_t)total * sizeof(float2), hipMemcpyHostToDevice, stream));
      din = fp_in.p;
      if (add) {
        fp_add.ensure((size_t)total);
        HIP_CHECK(hipMemcpyAsync(fp_add.p, add, (size_t)total * sizeof(float2), hipMemcpyHostToDevice, stream));
        dadd = fp_add.p;
      }
      dout = fp_in.p;
    }
    pupil_apply(din, dadd, nf, tab, dout);
    HIP_CHECK(hipGetLastError());
    if (mem == CBS_MEM_HOST)
      HIP_CHECK(hipMemcpyAsync(out, fp_in.p, (size_t)total * sizeof(float2), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  // batched 2D FFT over nslices (cx x cy) planes
  void fft2d(float2* buf, long long nslices, int inverse) {
    const RowsInfo* rw = io_rows;
    const StridedInfo* st = io_cols;
    auto fx_ = [&] {
      const long long nrows = nslices * cy;
      if (!rw) return rt_rows(io_rtx, buf, nrows, inverse);
      const dim3 g((unsigned)((nrows + rw->rows - 1) / rw->rows));
      timed(2, [&] { hipLaunchKernelGGL(rw->fn, g, dim3(rw->nt), 0, stream, buf, nrows, io_tw_cx.p, inverse); });
    };
    auto fy_ = [&] {
      if (!st) return rt_strided(io_rty, buf, cx, (long long)cx, (long long)cx * cy, nslices, inverse);
      const int nt = (cx + st->t - 1) / st->t;
      require((long long)nt * nslices < (1ll << 31), CBS_ERR_UNSUPPORTED, "2D FFT batch too large for one launch");
      const dim3 g((unsigned)(nt * nslices));
      timed(2, [&] { hipLaunchKernelGGL(st->fn, g, dim3(st->nt), 0, stream, buf, cx, (long long)cx,
                                        (long long)cx * cy, (int)nslices, io_tw_cy.p, inverse, (int)nslices, 0ll); });
    };
    if (!inverse) { fx_(); fy_(); }
    else { fy_(); fx_(); }
    HIP_CHECK(hipGetLastError());
  }

  // pupil_in / pupil_out (host, may be null): the input field is filtered into the source,
  // and the outputs are formed as the centred pupil transforms of cbs_solve_field
  void solve_2d(const float2* src, uint32_t n_rhs, float2* trans, float2* refl, float2* f3d, int mem,
                const float* pupil_in = nullptr, const float* pupil_out = nullptr) {
    require(have_potential, CBS_ERR_STATE, "cbs_set_potential must be called before solving");
    require(io.ready, CBS_ERR_STATE, "cbs_set_io_kernels must be called before cbs_solve_2d");
    require(n_rhs >= 1 && src != nullptr, CBS_ERR_ARG, "null source or n_rhs = 0");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    require(refl == nullptr || io.have_ref, CBS_ERR_STATE, "reflection requested but no kernel_ref was uploaded");
    const int nbw = batch_width(n_rhs);
    ensure_workspace(nbw, false);
    const long long n2 = (long long)cx * cy, nf = (long long)fx * fy;
    const size_t nvw = (size_t)nbw * P;   // component slots per batch; every 2D step is per component
    io_buf3.ensure(nvw * n2 * rz);
    io_spec2.ensure(nvw * n2);
    if (trans) io_out_t.ensure(nvw * n2);
    if (refl) io_out_r.ensure(nvw * n2);
    io_uin.ensure(nvw * nri);
    if (mem == CBS_MEM_HOST || !f3d) io_u3d.ensure(nvw * nri);
    if (mem == CBS_MEM_HOST) io_src.ensure(nvw * nf);
    if (mem == CBS_MEM_HOST || pupil_out) {
      if (trans) io_tr.ensure(nvw * nf);
      if (refl) io_rf.ensure(nvw * nf);
    }
    const float* tab_in = pupil_in ? pupil_table(pupil_in) : nullptr;
    const float* tab_out = pupil_out ? pupil_table(pupil_out) : nullptr;
    if (pupil_in) io_fsrc.ensure(nvw * nf);
    const int M = (int)prm.Bornmax;
    // host 3D fields: their download runs on the copy stream while the 2D projections of
    // the same batch compute (it is issued after them: a copy to pageable memory can hold
    // the host thread until it completes)
    const bool f3_async = mem == CBS_MEM_HOST && f3d != nullptr;
    if (f3_async) ensure_copy_stream();
    begin_solve(n_rhs, M);
    ev_used.clear();
    const float sc = (float)(1.0 / (double)n2);
    const Consts c = consts(1);
    const RoiGeom rg = roi();
    // downloads still running write into the caller's buffers: they are waited for before
    // an error leaves this call
    try {
      for (uint32_t r0 = 0; r0 < n_rhs; r0 += nbw) {
        const int nb = (int)std::min<uint32_t>(nbw, n_rhs - r0);
        const int nv = nb * P;
        const size_t v0 = (size_t)r0 * P;
        const float2* s_d = src + v0 * nf;
        if (mem == CBS_MEM_HOST) {
          HIP_CHECK(hipMemcpyAsync(io_src.p, s_d, (size_t)nv * nf * sizeof(float2), hipMemcpyHostToDevice, stream));
          s_d = io_src.p;
        }
        if (tab_in) {
          pupil_apply(s_d, nullptr, nv, tab_in, io_fsrc.p);
          s_d = io_fsrc.p;
        }
        // incident field: pad -> fft2 -> x refocus(z) -> ifft2 -> crop  (:543-554)
        timed(2, [&] { hipLaunchKernelGGL(io_pad_in, dim3(grid_for(nv * n2)), dim3(256), 0, stream, s_d, io_spec2.p, iog, nv); });
        fft2d(io_spec2.p, nv, 0);
        timed(2, [&] { hipLaunchKernelGGL(io_refocus, dim3(grid_for(nv * n2 * rz)), dim3(256), 0, stream, io_spec2.p,
                                          io_refocus_k.p, io_buf3.p, io_kcx.p, io_kcy.p, iog, nv, sc); });
        fft2d(io_buf3.p, (long long)nv * rz, 1);
        timed(2, [&] { hipLaunchKernelGGL(io_crop_in, dim3(grid_for(nv * nri)), dim3(256), 0, stream, io_buf3.p,
                                          io_uin.p, iog, nv); });
        float2* u3 = (mem == CBS_MEM_DEVICE && f3d) ? f3d + v0 * nri : io_u3d.p;
        if (f3_async && r0 > 0) HIP_CHECK(hipStreamWaitEvent(stream, ev_down[0], 0));   // io_u3d downloaded
        born_batch(io_uin.p, u3, nb, (int)r0);
        if (f3_async) HIP_CHECK(hipEventRecord(ev_done[0], stream));
        if (trans || refl) {
          // emitter -> fft2 per z -> sum_z E K (FP64) -> ifft2 -> crop  (:444-511)
          timed(2, [&] { hipLaunchKernelGGL(io_emitter, dim3(grid_for(nv * n2 * rz)), dim3(256), 0, stream, ws.field,
                                            io_uin.p, V.p, io_buf3.p, iog, c, rg.ox, rg.oy, rg.oz, nvox, nv,
                                            (float)prm.dV); });
          fft2d(io_buf3.p, (long long)nv * rz, 0);
          timed(2, [&] { hipLaunchKernelGGL(io_zsum, dim3(grid_for(nv * n2)), dim3(256), 0, stream, io_buf3.p,
                                            io_ktrans.p, io.have_ref ? io_kref.p : nullptr, trans ? io_out_t.p : nullptr,
                                            refl ? io_out_r.p : nullptr, io_kcx.p, io_kcy.p, iog, nv); });
          if (trans) {
            fft2d(io_out_t.p, nv, 1);
            float2* dst = mem == CBS_MEM_HOST ? io_tr.p : trans + v0 * nf;
            float2* crop = tab_out ? io_tr.p : dst;
            timed(2, [&] { hipLaunchKernelGGL(io_crop_out, dim3(grid_for(nv * nf)), dim3(256), 0, stream, io_out_t.p,
                                              crop, iog, nv, sc); });
            // fft2(t) + U = fft2(t + source): the scattered field plus the filtered incident one
            if (tab_out) pupil_apply(io_tr.p, s_d, nv, tab_out, dst);
          }
          if (refl) {
            fft2d(io_out_r.p, nv, 1);
            float2* dst = mem == CBS_MEM_HOST ? io_rf.p : refl + v0 * nf;
            float2* crop = tab_out ? io_rf.p : dst;
            timed(2, [&] { hipLaunchKernelGGL(io_crop_out, dim3(grid_for(nv * nf)), dim3(256), 0, stream, io_out_r.p,
                                              crop, iog, nv, sc); });
            if (tab_out) pupil_apply(io_rf.p, nullptr, nv, tab_out, dst);
          }
        }
        HIP_CHECK(hipGetLastError());
        if (mem == CBS_MEM_HOST) {
          if (f3_async) {
            HIP_CHECK(hipStreamWaitEvent(cstream, ev_done[0], 0));
            HIP_CHECK(hipMemcpyAsync(f3d + v0 * nri, io_u3d.p, (size_t)nv * nri * sizeof(float2),
                                     hipMemcpyDeviceToHost, cstream));
            HIP_CHECK(hipEventRecord(ev_down[0], cstream));
          }
          if (trans)
            HIP_CHECK(hipMemcpyAsync(trans + v0 * nf, io_tr.p, (size_t)nv * nf * sizeof(float2),
                                     hipMemcpyDeviceToHost, stream));
          if (refl)
            HIP_CHECK(hipMemcpyAsync(refl + v0 * nf, io_rf.p, (size_t)nv * nf * sizeof(float2),
                                     hipMemcpyDeviceToHost, stream));
        }
      }
    } catch (...) {
      if (f3_async) (void)hipStreamSynchronize(cstream);
      (void)hipStreamSynchronize(stream);
      throw;
    }
    if (f3_async) HIP_CHECK(hipStreamSynchronize(cstream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (opt.time_kernels) resolve_times();
    if (prm.acyclic) flip_state = (int)((flip_state + (long long)n_rhs * M) & 1);
  }

  // Times every fused plan for the grid and keeps the fastest.
  double autotune(uint32_t n_rhs, uint32_t iters) {
    require(have_potential, CBS_ERR_STATE, "cbs_set_potential must be called before cbs_autotune");
    require(n_rhs >= 1 && iters >= 3, CBS_ERR_ARG, "autotune needs n_rhs >= 1 and iterations >= 3");
    const PassRInfo* prs = nullptr;
    const int npr = list_pass_real(&prs);
    const PassKInfo* pks = nullptr;
    const int npk = list_pass_kspace(&pks);
    drop_graph();
    const cbs_options saved = opt;
    opt.split_a = opt.tile_t = opt.threads_real = opt.threads_kspace = 0;
    opt.form = CBS_FORM_FUSED;
    std::vector<std::pair<const PassRInfo*, const PassKInfo*>> cands;
    for (int i = 0; i < npr; ++i) {
      if (!fused_ok_r(prs[i])) continue;
      for (int j = 0; j < npk; ++j)
        if (fused_ok_k(pks[j], ny / prs[i].a)) cands.emplace_back(&prs[i], &pks[j]);
    }
    if (cands.empty()) {
      opt = saved;
      throw CbsError(CBS_ERR_UNSUPPORTED, "no fused kernel plan to tune for grid " + shape_str());
    }
    const int nb = batch_width(n_rhs);
    ensure_workspace(nb, false);
    DevBuf<float2> src, dst;
    src.ensure((size_t)nb * P * nri);
    dst.ensure((size_t)nb * P * nri);
    {
      std::vector<float2> h((size_t)nb * P * nri);
      for (size_t i = 0; i < h.size(); ++i) h[i] = make_float2(std::cos(0.37 * (double)i), std::sin(0.21 * (double)i));
      HIP_CHECK(hipMemcpy(src.p, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    const uint32_t saved_iters = prm.Bornmax;
    const int saved_flip = flip_state;
    const int saved_time = opt.time_kernels, saved_norm = opt.track_norms;
    const double saved_tol = conv_tol;
    opt.time_kernels = 0;
    opt.track_norms = 0;
    opt.use_graph = 0;      // every trial plan runs eagerly (its tables are rebuilt per trial)
    conv_tol = 0.0;
    prm.Bornmax = iters;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    double best = 1e30;
    Plan best_plan;
    int best_order = 4;
    int best_w = nb;
    opt.block_order = 4;
    // the faster of two timed runs after a warm-up (single runs differ by a few per cent)
    auto time_plan = [&](int w) {
      born_batch(src.p, dst.p, w, 0);
      double t = 1e30;
      for (int k = 0; k < 2; ++k) {
        HIP_CHECK(hipEventRecord(e0, stream));
        born_batch(src.p, dst.p, w, 0);
        HIP_CHECK(hipEventRecord(e1, stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        t = std::min(t, (double)ms);
      }
      return t;
    };
    try {
      // every kernel pair in block order 4, then the block orders of the three fastest
      std::vector<std::pair<double, Plan>> timed_plans;
      for (auto& cd : cands) {
        Plan p;
        p.form = CBS_FORM_FUSED;
        p.pr = cd.first;
        p.pk = cd.second;
        p.a = cd.first->a;
        p.b = ny / p.a;
        p.t = cd.second->t;
        plan = p;
        build_tables();
        timed_plans.emplace_back(time_plan(nb), p);
      }
      std::sort(timed_plans.begin(), timed_plans.end(),
                [](const std::pair<double, Plan>& x, const std::pair<double, Plan>& y) { return x.first < y.first; });
      const size_t ntop = std::min<size_t>(3, timed_plans.size());
      for (size_t i = 0; i < ntop; ++i) {
        plan = timed_plans[i].second;
        build_tables();
        // real-space tiles in storage order (4), rhs first (0) or rows of one rhs first (1);
        // k-space rhs or x tile first (2), nontemporal k-space loads (8: 2 % faster with
        // 512-thread k-space blocks at 160^3 x 49, 13-25 % slower with 1024-thread ones)
        for (int ord : {4, 0, 1, 2, 3, 6, 12}) {
          opt.block_order = ord;
          const double ms = ord == 4 ? timed_plans[i].first : time_plan(nb);
          if (ms < best) {
            best = ms;
            best_plan = plan;
            best_order = ord;
          }
        }
        opt.block_order = 4;
      }
      // batch width: the chosen plan with nb / 2, nb / 4, nb / 8, 2 and 1 rhs per launch
      // (narrower batches keep more of the state in the caches between launches but give
      // the k-space pass fewer blocks); a narrower width is kept when its predicted time
      // for nb rhs is at least 3 % below the full batch's
      plan = best_plan;
      build_tables();
      opt.block_order = best_order;
      for (int w : {(nb + 1) / 2, (nb + 3) / 4, (nb + 7) / 8, 2, 1}) {
        if (w >= best_w || w < 1) continue;
        const double total = time_plan(w) * (double)((nb + w - 1) / w);
        if (total < 0.97 * best) {
          best = total;
          best_w = w;
        }
      }
    } catch (...) {
      prm.Bornmax = saved_iters;
      flip_state = saved_flip;
      conv_tol = saved_tol;
      opt = saved;
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      make_plan();
      build_tables();
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    prm.Bornmax = saved_iters;
    flip_state = saved_flip;
    opt = saved;
    opt.form = CBS_FORM_FUSED;
    opt.split_a = best_plan.a;
    opt.tile_t = best_plan.t;
    opt.threads_real = best_plan.pr->nt;
    opt.threads_kspace = best_plan.pk->nt;
    opt.block_order = best_order;
    if (best_w < nb) opt.max_batch = best_w;
    opt.time_kernels = saved_time;
    opt.track_norms = saved_norm;
    conv_tol = saved_tol;
    plan = best_plan;
    build_tables();
    HIP_CHECK(hipStreamSynchronize(stream));
    return best;
  }

  DevBuf<float2> tv_in, tv_y, tv_P, tv_R;   // used as float arrays for real input
  DevBuf<float2> gr_u, gr_w;
  DevBuf<double2> gr_out;

  // Both element types share the buffers: a real problem of n voxels uses the first
  // n (or 3n) floats of buffers sized for n complex values.
  void tv_prox(const void* in, void* out, const uint32_t* dims, double lam, double vmin, double vmax,
               double vimag_max, int non_neg, int dirichlet, uint32_t inner, int mem, bool cplx) {
    require(in && out && dims, CBS_ERR_ARG, "null TV argument");
    require(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, CBS_ERR_ARG, "TV dims must be >= 1");
    require(lam > 0, CBS_ERR_ARG, "TV lambda must be > 0");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    const size_t esz = cplx ? sizeof(float2) : sizeof(float);
    const long long nf = cplx ? 2 * n : n;   // floats per volume
    tv_y.ensure((size_t)n);
    tv_P.ensure((size_t)(3 * n));
    tv_R.ensure((size_t)(3 * n));
    const void* din = in;
    if (mem == CBS_MEM_HOST) {
      tv_in.ensure((size_t)n);
      HIP_CHECK(hipMemcpyAsync(tv_in.p, in, (size_t)n * esz, hipMemcpyHostToDevice, stream));
      din = tv_in.p;
    }
    TvArgs a;
    a.nx = (int)dims[0]; a.ny = (int)dims[1]; a.nz = (int)dims[2];
    a.lam = (float)lam; a.vmin = (float)vmin; a.vmax = (float)vmax;
    a.vimag_max = (vimag_max > 0) ? (float)vimag_max : 0.f;   // NaN or <= 0: no cap (the reference tests > 0)
    a.non_neg = non_neg; a.dirichlet = dirichlet;
    float* P = reinterpret_cast<float*>(tv_P.p);
    float* R = reinterpret_cast<float*>(tv_R.p);
    float* Y = reinterpret_cast<float*>(tv_y.p);
    hipLaunchKernelGGL(tv_zero, dim3(grid_for(3 * nf)), dim3(256), 0, stream, P, 3 * nf);
    hipLaunchKernelGGL(tv_zero, dim3(grid_for(3 * nf)), dim3(256), 0, stream, R, 3 * nf);
    const unsigned g = grid_for(n);
    auto div_project = [&](const float* Q, void* dst) {
      if (cplx)
        hipLaunchKernelGGL(tv_div_project_c, dim3(g), dim3(256), 0, stream, (const float2*)din, (const float2*)Q,
                           (float2*)dst, a);
      else
        hipLaunchKernelGGL(tv_div_project, dim3(g), dim3(256), 0, stream, (const float*)din, Q, (float*)dst, a);
    };
    double t = 1.0;
    for (uint32_t it = 0; it < inner; ++it) {
      const double tn = (1.0 + std::sqrt(1.0 + 4.0 * t * t)) / 2.0;
      const float theta = (float)((t - 1.0) / tn);
      div_project(R, Y);
      if (cplx)
        hipLaunchKernelGGL(tv_dual_update_c, dim3(g), dim3(256), 0, stream, (const float2*)Y, (float2*)P,
                           (float2*)R, a, theta);
      else
        hipLaunchKernelGGL(tv_dual_update, dim3(g), dim3(256), 0, stream, (const float*)Y, P, R, a, theta);
      t = tn;
    }
    div_project(P, mem == CBS_MEM_HOST ? (void*)Y : out);
    HIP_CHECK(hipGetLastError());
    if (mem == CBS_MEM_HOST) HIP_CHECK(hipMemcpyAsync(out, Y, (size_t)n * esz, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  void field_grad(const void* u, const void* w, uint32_t nr, const uint32_t* dims, void* grad, int mem) {
    require(u && w && grad && dims, CBS_ERR_ARG, "null gradient argument");
    require(nr >= 1, CBS_ERR_ARG, "n_rhs must be >= 1");
    require(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, CBS_ERR_ARG, "gradient dims must be >= 1");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    const float2* du = (const float2*)u;
    const float2* dw = (const float2*)w;
    double2* dg = (double2*)grad;
    if (mem == CBS_MEM_HOST) {
      gr_u.ensure((size_t)(n * nr));
      gr_w.ensure((size_t)(n * nr));
      gr_out.ensure((size_t)n);
      HIP_CHECK(hipMemcpyAsync(gr_u.p, u, (size_t)(n * nr) * sizeof(float2), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipMemcpyAsync(gr_w.p, w, (size_t)(n * nr) * sizeof(float2), hipMemcpyHostToDevice, stream));
      du = gr_u.p; dw = gr_w.p; dg = gr_out.p;
    }
    hipLaunchKernelGGL(field_gradient, dim3(grid_for(n)), dim3(256), 0, stream, du, dw, (int)dims[0], (int)dims[1],
                       (int)dims[2], (int)nr, dg);
    HIP_CHECK(hipGetLastError());
    if (mem == CBS_MEM_HOST)
      HIP_CHECK(hipMemcpyAsync(grad, gr_out.p, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  void set_potential(const cbs_potential_params* p, const float2* v, int mem) {
    require(p != nullptr && v != nullptr, CBS_ERR_ARG, "null potential argument");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    check_params(p);
    prm = *p;
    drop_graph();
    v_real = false;   // a caller's V is taken as complex
    V.ensure((size_t)nvox);
    HIP_CHECK(hipMemcpyAsync(V.p, v, (size_t)nvox * sizeof(float2),
                             mem == CBS_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, stream));
    finish_potential();
  }

  // ---- the RI -> V conversion of set_RI on the device
  DevBuf<double> ri_stage, vmax_part;
  RiGrid ri_grid(const void* ri, int cplx, double wavelength, double n_m, int mem) {
    require(ri != nullptr, CBS_ERR_ARG, "null RI pointer");
    require(cplx == 0 || cplx == 1, CBS_ERR_ARG, "ri_complex must be 0 or 1");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    require(wavelength > 0 && n_m > 0 && std::isfinite(wavelength) && std::isfinite(n_m), CBS_ERR_ARG,
            "wavelength and RI_bg must be > 0");
    RiGrid g;
    const size_t n = (size_t)nri * (cplx ? 2 : 1);
    if (mem == CBS_MEM_DEVICE) {
      g.ri = reinterpret_cast<const double*>(ri);
    } else {
      ri_stage.ensure(n);
      HIP_CHECK(hipMemcpyAsync(ri_stage.p, ri, n * sizeof(double), hipMemcpyHostToDevice, stream));
      g.ri = ri_stage.p;
    }
    g.cplx = cplx;
    g.nx = rx; g.ny = ry; g.nz = rz;
    const double k = 2.0 * M_PI * n_m / wavelength;   // the host restatement's k, then k * k
    g.kk = k * k;
    g.inv_n2 = 1.0 / (n_m * n_m);
    return g;
  }
  double ri_vmax(const void* ri, int cplx, double wavelength, double n_m, int mem) {
    const RiGrid g = ri_grid(ri, cplx, wavelength, n_m, mem);
    const int nb = (int)std::min<long long>(1024, (nri + 255) / 256);
    vmax_part.ensure((size_t)nb);
    hipLaunchKernelGGL(potential_vmax, dim3(nb), dim3(256), 0, stream, g, vmax_part.p);
    HIP_CHECK(hipGetLastError());
    std::vector<double> part((size_t)nb);
    HIP_CHECK(hipMemcpyAsync(part.data(), vmax_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    double m = 0.0;
    for (double v : part) m = (std::isnan(v) || std::isnan(m)) ? v + m : std::max(m, v);
    return m;
  }
  void set_potential_ri(const cbs_potential_params* p, const void* ri, int cplx, int flip_conj, double wavelength,
                        double n_m, int mem) {
    require(p != nullptr, CBS_ERR_ARG, "null potential argument");
    require(flip_conj == 0 || flip_conj == 1, CBS_ERR_ARG, "flip_conj must be 0 or 1");
    check_params(p);
    const RiGrid g = ri_grid(ri, cplx, wavelength, n_m, mem);
    prm = *p;
    drop_graph();
    v_real = cplx == 0;   // Im V = 0 - eps at every voxel, inside the RI grid and in the boundary
    V.ensure((size_t)nvox);
    const RoiGeom rg = roi();
    hipLaunchKernelGGL(potential_fill, dim3(grid_for(nvox)), dim3(256), 0, stream, g, flip_conj, rg, nx, ny,
                       (long long)nvox, p->eps_imag, V.p);
    HIP_CHECK(hipGetLastError());
    finish_potential();
  }

  // Vt for the fused plan (its a and b) from the current V
  void sync_vtile() {
    if (plan.form != CBS_FORM_FUSED || !V.p) return;
    const std::vector<long long> key = {v_version, (long long)plan.a, (long long)plan.b, nvox, (long long)v_real};
    if (key == vt_key) return;
    if (v_real) {
      Vr.ensure((size_t)nvox);
      hipLaunchKernelGGL(tile_rows_real, dim3(grid_for(nvox)), dim3(256), 0, stream, V.p, Vr.p, consts(1));
    } else {
      Vt.ensure((size_t)nvox);
      hipLaunchKernelGGL(tile_rows, dim3(grid_for(nvox)), dim3(256), 0, stream, V.p, Vt.p, consts(1));
    }
    HIP_CHECK(hipGetLastError());
    vt_key = key;
  }

  void finish_potential() {
    ++v_version;
    make_plan();
    // the tables depend on the plan and on the geometry scalars, not on V or eps: a new
    // potential on the same geometry (every set_RI of a reconstruction or a serving loop)
    // keeps them
    if (table_key() != tab_key) build_tables();
    sync_vtile();
    HIP_CHECK(hipStreamSynchronize(stream));
    flip_state = 0;
    have_potential = true;
  }

  void check_params(const cbs_potential_params* p) const {
    require(p->eps_imag > 0 && std::isfinite(p->eps_imag), CBS_ERR_ARG,
            "eps_imag must be > 0 (a homogeneous RI has eps_imag = 0 and is not solvable by CBS)");
    require(p->k0_nm > 0, CBS_ERR_ARG, "k0_nm must be > 0");
    require(p->boundary_sharpness > 0, CBS_ERR_ARG, "boundary_sharpness must be > 0");
    const int dims[3] = {nx, ny, nz};
    const int rdims[3] = {rx, ry, rz};
    for (int j = 0; j < 3; ++j) {
      require(p->fourier_res[j] > 0, CBS_ERR_ARG, "fourier_res must be > 0");
      const long lo = (long)p->ROI[2 * j], hi = (long)p->ROI[2 * j + 1];
      require(lo >= 1 && hi <= dims[j] && hi - lo + 1 == rdims[j], CBS_ERR_ARG,
              "ROI does not match RI_size inside simulation_size on axis " + std::to_string(j));
      require(lo - 1 == dims[j] - hi, CBS_ERR_ARG, "ROI must be centred (equal boundary on both sides) on axis " +
                                                       std::to_string(j));
    }
  }
};

// ------------------------------------------------------------------ registry
static std::mutex g_mu;
static std::set<Solver*> g_live;

static Solver* lookup(cbs_handle h) {
  std::lock_guard<std::mutex> lk(g_mu);
  Solver* s = reinterpret_cast<Solver*>(h);
  if (!s || !g_live.count(s)) throw CbsError(CBS_ERR_HANDLE, "unknown or destroyed cbs handle");
  return s;
}

struct DeviceGuard {
  int prev = 0;
  explicit DeviceGuard(int dev) {
    (void)hipGetDevice(&prev);
    HIP_CHECK(hipSetDevice(dev));
  }
  ~DeviceGuard() { (void)hipSetDevice(prev); }
};

template <class F>
static int guarded(F&& f) {
  try {
    f();
    return CBS_OK;
  } catch (const CbsError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return CBS_ERR_HIP;
  } catch (...) {
    g_last_error = "unknown error";
    return CBS_ERR_HIP;
  }
}

}  // namespace cbs

using namespace cbs;

// ------------------------------------------------------------------ Goldstein branch cuts (host part)
namespace {

// remove_dipoles + goldstein_branch_cut on device-resident data: dres (m1 x m2 x n3,
// consumed), dstep1 / dstep2 (n1 x n2 x n3, n2 x n3) overwritten with the step maps.
void gs_step_maps(float* dres, int32_t* dstep1, int32_t* dstep2, int n1, int n2, int n3) {
  const int m1 = n1 - 1, m2 = n2 - 1;
  const long long m = (long long)m1 * m2 * n3;
  HIP_CHECK(hipMemset(dstep1, 0, (size_t)n1 * n2 * n3 * sizeof(int32_t)));
  HIP_CHECK(hipMemset(dstep2, 0, (size_t)n2 * n3 * sizeof(int32_t)));
  std::vector<float> res((size_t)m);
  HIP_CHECK(hipMemcpy(res.data(), dres, m * sizeof(float), hipMemcpyDeviceToHost));
  std::vector<int32_t> pos;
  for (long long i = 0; i < m; ++i)
    if (res[i] > 0.0f) pos.push_back((int32_t)(i + 1));
  if (!pos.empty()) {
    DevBuf<int32_t> dpos;
    dpos.upload(pos);
    const unsigned grid = (unsigned)((pos.size() + 255) / 256);
    for (int dir = 0; dir < 4; ++dir)
      hipLaunchKernelGGL(phase_dipoles, dim3(grid), dim3(256), 0, 0, dres, dpos.p, dstep1, (int)pos.size(), m1, m2, n3,
                         dir);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(res.data(), dres, m * sizeof(float), hipMemcpyDeviceToHost));
  }
  const gs::GsList L = gs::gs_list(res, m1, m2, n3);
  if (L.n == 0) return;
  DevBuf<int32_t> drows, dtab, dzs, dsq, dsl;
  drows.upload(L.rows);
  dtab.upload(L.table);
  dzs.upload(L.zstart);
  dsq.upload(L.square);
  dsl.ensure((size_t)L.n * 2 * gs::kNearest);
  hipLaunchKernelGGL(phase_nearest, dim3((L.n + 127) / 128), dim3(128), 0, 0, drows.p, L.n, dtab.p, (int)L.table.size(),
                     dzs.p, dsq.p, m1, m2, dsl.p);
  HIP_CHECK(hipGetLastError());
  std::vector<int32_t> sl((size_t)L.n * 2 * gs::kNearest);
  HIP_CHECK(hipMemcpy(sl.data(), dsl.p, sl.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  const std::vector<int32_t> br = gs::gs_branches(L, sl, m1, m2);
  const int nb = (int)(br.size() / 6);
  if (nb > 0) {
    DevBuf<int32_t> dbr;
    dbr.upload(br);
    hipLaunchKernelGGL(phase_raster, dim3((nb + 127) / 128), dim3(128), 0, 0, dbr.p, nb, dstep1, dstep2, n1, n2, n3);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(nullptr));   // this call's work, on the null stream
  }
}

void check_phase_dims(const uint32_t dims[3]) {
  require(dims[0] >= 2 && dims[1] >= 2 && dims[2] >= 1, CBS_ERR_ARG, "phase maps need at least 2 x 2 pixels");
  require((unsigned long long)dims[0] * dims[1] * dims[2] < (1ull << 31), CBS_ERR_UNSUPPORTED,
          "phase maps of 2^31 pixels or more");
}
}  // namespace

extern "C" {

// ------------------------------------------------------------------ phase unwrapping (handle-free)
namespace {
struct DevSel {
  int prev = 0;
  explicit DevSel(int d) {
    HIP_CHECK(hipGetDevice(&prev));
    HIP_CHECK(hipSetDevice(d));
  }
  ~DevSel() { (void)hipSetDevice(prev); }
};
}  // namespace

int cbs_phase_residues(int device, const float* wrapped, float* residues, const uint32_t dims[3], int mem) {
  return guarded([&] {
    require(wrapped && residues && dims, CBS_ERR_ARG, "null argument");
    require(dims[0] >= 2 && dims[1] >= 2 && dims[2] >= 1, CBS_ERR_ARG, "phase maps need at least 2 x 2 pixels");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    DevSel ds(device);
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    const long long m = (long long)(dims[0] - 1) * (dims[1] - 1) * dims[2];
    DevBuf<float> bp, br;
    const float* dp = wrapped;
    float* dr = residues;
    if (mem == CBS_MEM_HOST) {
      bp.ensure((size_t)n);
      br.ensure((size_t)m);
      HIP_CHECK(hipMemcpy(bp.p, wrapped, n * sizeof(float), hipMemcpyHostToDevice));
      dp = bp.p;
      dr = br.p;
    }
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>((m + 255) / 256, 1 << 20));
    hipLaunchKernelGGL(phase_residues, dim3(grid), dim3(256), 0, 0, dp, dr, (int)dims[0], (int)dims[1], (int)dims[2]);
    HIP_CHECK(hipGetLastError());
    if (mem == CBS_MEM_HOST) HIP_CHECK(hipMemcpy(residues, br.p, m * sizeof(float), hipMemcpyDeviceToHost));
    HIP_CHECK(hipStreamSynchronize(nullptr));   // this call's work, on the null stream
  });
}

int cbs_phase_integrate(int device, const float* wrapped, const int32_t* step1, const int32_t* step2, float* unwrapped,
                        const uint32_t dims[3], int mem) {
  return guarded([&] {
    require(wrapped && step1 && step2 && unwrapped && dims, CBS_ERR_ARG, "null argument");
    require(dims[0] >= 1 && dims[1] >= 1 && dims[2] >= 1, CBS_ERR_ARG, "empty phase map");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    DevSel ds(device);
    const long long n = (long long)dims[0] * dims[1] * dims[2];
    const long long n23 = (long long)dims[1] * dims[2];
    DevBuf<float> bp;
    DevBuf<int32_t> b1, b2;
    float* dp = unwrapped;
    const int32_t* d1 = step1;
    const int32_t* d2 = step2;
    if (mem == CBS_MEM_HOST) {
      bp.ensure((size_t)n);
      b1.ensure((size_t)n);
      b2.ensure((size_t)n23);
      HIP_CHECK(hipMemcpy(bp.p, wrapped, n * sizeof(float), hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(b1.p, step1, n * sizeof(int32_t), hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(b2.p, step2, n23 * sizeof(int32_t), hipMemcpyHostToDevice));
      dp = bp.p;
      d1 = b1.p;
      d2 = b2.p;
    } else if (unwrapped != wrapped) {
      HIP_CHECK(hipMemcpy(unwrapped, wrapped, n * sizeof(float), hipMemcpyDeviceToDevice));
    }
    hipLaunchKernelGGL(phase_integrate_column, dim3((dims[2] + 63) / 64), dim3(64), 0, 0, dp, d2, (int)dims[0],
                       (int)dims[1], (int)dims[2]);
    hipLaunchKernelGGL(phase_integrate_rows, dim3((unsigned)((n23 + 127) / 128)), dim3(128), 0, 0, dp, d1,
                       (int)dims[0], (int)dims[1], (int)dims[2]);
    HIP_CHECK(hipGetLastError());
    if (mem == CBS_MEM_HOST) HIP_CHECK(hipMemcpy(unwrapped, bp.p, n * sizeof(float), hipMemcpyDeviceToHost));
    HIP_CHECK(hipStreamSynchronize(nullptr));   // this call's work, on the null stream
  });
}

int cbs_phase_branch_cuts(int device, const float* residues, int32_t* step1, int32_t* step2, const uint32_t dims[3],
                          int mem) {
  return guarded([&] {
    require(residues && step1 && step2 && dims, CBS_ERR_ARG, "null argument");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    check_phase_dims(dims);
    DevSel ds(device);
    const int n1 = (int)dims[0], n2 = (int)dims[1], n3 = (int)dims[2];
    const long long m = (long long)(n1 - 1) * (n2 - 1) * n3;
    DevBuf<float> dres;
    DevBuf<int32_t> b1, b2;
    dres.ensure((size_t)m);
    HIP_CHECK(hipMemcpy(dres.p, residues, m * sizeof(float),
                        mem == CBS_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice));
    int32_t* d1 = step1;
    int32_t* d2 = step2;
    if (mem == CBS_MEM_HOST) {
      b1.ensure((size_t)n1 * n2 * n3);
      b2.ensure((size_t)n2 * n3);
      d1 = b1.p;
      d2 = b2.p;
    }
    gs_step_maps(dres.p, d1, d2, n1, n2, n3);
    if (mem == CBS_MEM_HOST) {
      HIP_CHECK(hipMemcpy(step1, b1.p, (size_t)n1 * n2 * n3 * sizeof(int32_t), hipMemcpyDeviceToHost));
      HIP_CHECK(hipMemcpy(step2, b2.p, (size_t)n2 * n3 * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    HIP_CHECK(hipStreamSynchronize(nullptr));   // this call's work, on the null stream
  });
}

int cbs_phase_unwrap(int device, const float* wrapped, float* unwrapped, const uint32_t dims[3], int mem) {
  return guarded([&] {
    require(wrapped && unwrapped && dims, CBS_ERR_ARG, "null argument");
    require(mem == CBS_MEM_HOST || mem == CBS_MEM_DEVICE, CBS_ERR_ARG, "mem must be CBS_MEM_HOST or CBS_MEM_DEVICE");
    check_phase_dims(dims);
    DevSel ds(device);
    const int n1 = (int)dims[0], n2 = (int)dims[1], n3 = (int)dims[2];
    const long long n = (long long)n1 * n2 * n3;
    const long long m = (long long)(n1 - 1) * (n2 - 1) * n3;
    DevBuf<float> bp, dres;
    DevBuf<int32_t> d1, d2;
    float* dp = unwrapped;
    if (mem == CBS_MEM_HOST) {
      bp.ensure((size_t)n);
      dp = bp.p;
      HIP_CHECK(hipMemcpy(dp, wrapped, n * sizeof(float), hipMemcpyHostToDevice));
    } else if (unwrapped != wrapped) {
      HIP_CHECK(hipMemcpy(dp, wrapped, n * sizeof(float), hipMemcpyDeviceToDevice));
    }
    dres.ensure((size_t)m);
    d1.ensure((size_t)n);
    d2.ensure((size_t)n2 * n3);
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>((m + 255) / 256, 1 << 20));
    hipLaunchKernelGGL(phase_residues, dim3(grid), dim3(256), 0, 0, dp, dres.p, n1, n2, n3);
    HIP_CHECK(hipGetLastError());
    gs_step_maps(dres.p, d1.p, d2.p, n1, n2, n3);
    hipLaunchKernelGGL(phase_integrate_column, dim3((n3 + 63) / 64), dim3(64), 0, 0, dp, d2.p, n1, n2, n3);
    hipLaunchKernelGGL(phase_integrate_rows, dim3((unsigned)(((long long)n2 * n3 + 127) / 128)), dim3(128), 0, 0, dp,
                       d1.p, n1, n2, n3);
    HIP_CHECK(hipGetLastError());
    if (mem == CBS_MEM_HOST) HIP_CHECK(hipMemcpy(unwrapped, dp, n * sizeof(float), hipMemcpyDeviceToHost));
    HIP_CHECK(hipStreamSynchronize(nullptr));   // this call's work, on the null stream
  });
}

int cbs_abi_version(void) { return CBS_ABI_VERSION; }

const char* cbs_last_error(void) { return g_last_error.c_str(); }

int cbs_create(const cbs_geometry* geo, cbs_handle* out) {
  return guarded([&] {
    require(geo && out, CBS_ERR_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    require(geo->device >= 0 && geo->device < ndev, CBS_ERR_ARG, "device ordinal out of range");
    for (int j = 0; j < 3; ++j) {
      require(geo->simulation_size[j] >= 2 && geo->RI_size[j] >= 1 && geo->RI_size[j] <= geo->simulation_size[j],
              CBS_ERR_ARG, "invalid simulation/RI size on axis " + std::to_string(j));
    }
    auto s = std::make_unique<Solver>();
    s->device = geo->device;
    DeviceGuard dg(s->device);
    s->nx = (int)geo->simulation_size[0]; s->ny = (int)geo->simulation_size[1]; s->nz = (int)geo->simulation_size[2];
    s->rx = (int)geo->RI_size[0]; s->ry = (int)geo->RI_size[1]; s->rz = (int)geo->RI_size[2];
    s->fx = (int)geo->field_size[0]; s->fy = (int)geo->field_size[1];
    s->cx = (int)geo->conv_size[0]; s->cy = (int)geo->conv_size[1];
    s->rc_x = geo->RI_center[0]; s->rc_y = geo->RI_center[1];
    require(geo->pole_num == 0 || geo->pole_num == 1 || geo->pole_num == 3, CBS_ERR_ARG,
            "pole_num must be 1 (scalar) or 3 (vector)");
    s->P = geo->pole_num == 3 ? 3 : 1;
    s->nvox = (long long)s->nx * s->ny * s->nz;
    s->nri = (long long)s->rx * s->ry * s->rz;
    s->opt.form = CBS_FORM_FUSED;
    s->opt.use_graph = 1;   // the fused and four-pass loops replay a captured graph unless kernels are timed
    s->opt.block_order = 4; // real-space tiles in storage order: 2.5 % faster than rhs first at 160^3 x 49
    {
      std::lock_guard<std::recursive_mutex> lk(device_mutex());
      HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    }
    // set-up of the no-allocation workspace on this device (a failure is reported by
    // cbs_solve_3d_noalloc, which resolves it again)
    try {
      (void)scratch_base(s->device);
    } catch (const CbsError&) {
    }
    Solver* raw = s.release();
    std::lock_guard<std::mutex> lk(g_mu);
    g_live.insert(raw);
    *out = reinterpret_cast<cbs_handle>(raw);
  });
}

int cbs_destroy(cbs_handle h) {
  return guarded([&] {
    Solver* s;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      s = reinterpret_cast<Solver*>(h);
      if (!s || !g_live.count(s)) throw CbsError(CBS_ERR_HANDLE, "unknown or destroyed cbs handle");
      g_live.erase(s);
    }
    {
      DeviceGuard dg(s->device);
      (void)hipStreamSynchronize(s->stream);
    }
    delete s;
  });
}

int cbs_set_options(cbs_handle h, const cbs_options* opt) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(opt != nullptr, CBS_ERR_ARG, "null options");
    require(opt->form == CBS_FORM_FUSED || opt->form == CBS_FORM_PLAIN || opt->form == CBS_FORM_ROWS ||
                opt->form == CBS_FORM_AXES, CBS_ERR_ARG,
            "unknown form");
    require(opt->max_batch >= 0 && opt->split_a >= 0 && opt->tile_t >= 0, CBS_ERR_ARG, "negative option");
    require(opt->block_order >= 0 && opt->block_order <= 15, CBS_ERR_ARG, "block_order must be 0..15");
    DeviceGuard dg(s->device);
    s->opt = *opt;
    s->drop_graph();
    if (s->have_potential) {
      s->make_plan();
      s->build_tables();
    }
  });
}

int cbs_get_options(cbs_handle h, cbs_options* opt) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(opt != nullptr, CBS_ERR_ARG, "null options");
    *opt = s->opt;
  });
}

int cbs_set_potential(cbs_handle h, const cbs_potential_params* prm, const void* V, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->set_potential(prm, reinterpret_cast<const float2*>(V), mem);
  });
}

int cbs_potential_max(cbs_handle h, const void* ri, int ri_complex, double wavelength, double RI_bg, int mem,
                      double* vmax) {
  return guarded([&] {
    require(vmax != nullptr, CBS_ERR_ARG, "null vmax pointer");
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    *vmax = s->ri_vmax(ri, ri_complex, wavelength, RI_bg, mem);
  });
}

int cbs_set_potential_ri(cbs_handle h, const cbs_potential_params* prm, const void* ri, int ri_complex, int flip_conj,
                         double wavelength, double RI_bg, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->set_potential_ri(prm, ri, ri_complex, flip_conj, wavelength, RI_bg, mem);
  });
}

int cbs_solve_3d(cbs_handle h, const void* u_in, uint32_t n_rhs, void* u_out, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->solve_3d(reinterpret_cast<const float2*>(u_in), n_rhs, reinterpret_cast<float2*>(u_out), mem);
  });
}

int cbs_solve_3d_noalloc(cbs_handle h, const void* u_in, uint32_t n_rhs, void* u_out) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    unsigned char* scr = scratch_base(s->device);
    s->solve_3d_noalloc(reinterpret_cast<const float2*>(u_in), n_rhs, reinterpret_cast<float2*>(u_out), scr);
  });
}

int cbs_tv_prox(cbs_handle h, const float* in, float* out, const uint32_t dims[3], double lambda, double vmin,
                double vmax, int non_neg, int dirichlet, uint32_t inner_itt, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->tv_prox(in, out, dims, lambda, vmin, vmax, 0.0, non_neg, dirichlet, inner_itt, mem, false);
  });
}

int cbs_tv_prox_complex(cbs_handle h, const void* in, void* out, const uint32_t dims[3], double lambda, double vmin,
                        double vmax, double vimag_max, int non_neg, int dirichlet, uint32_t inner_itt, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->tv_prox(in, out, dims, lambda, vmin, vmax, vimag_max, non_neg, dirichlet, inner_itt, mem, true);
  });
}

int cbs_field_gradient(cbs_handle h, const void* u, const void* w, uint32_t n_rhs, const uint32_t dims[3], void* grad,
                       int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->field_grad(u, w, n_rhs, dims, grad, mem);
  });
}

int cbs_solve_field(cbs_handle h, const void* in_field, uint32_t n_rhs, const float* pupil_in, const float* pupil_out,
                    void* trans, void* refl, void* field3d, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    require(pupil_in && pupil_out, CBS_ERR_ARG, "null pupil");
    require(s->P == 1, CBS_ERR_UNSUPPORTED, "cbs_solve_field takes scalar fields (pole_num 1)");
    s->solve_2d(static_cast<const float2*>(in_field), n_rhs, static_cast<float2*>(trans), static_cast<float2*>(refl),
                static_cast<float2*>(field3d), mem, pupil_in, pupil_out);
  });
}

int cbs_field_pupil(cbs_handle h, const void* in, const void* add, uint32_t n_fields, const float* pupil, void* out,
                    int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->field_pupil(in, add, n_fields, pupil, out, mem);
  });
}

int cbs_fista_prox_input(cbs_handle h, const void* u, const void* grad, double inv_alpha, uint64_t n, int cplx,
                         void* out) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(u && grad && out, CBS_ERR_ARG, "null FISTA argument");
    require(n >= 1 && n < (1ull << 40), CBS_ERR_ARG, "FISTA length out of range");
    require(cplx == 0 || cplx == 1, CBS_ERR_ARG, "cplx must be 0 or 1");
    require(std::isfinite(inv_alpha), CBS_ERR_ARG, "inv_alpha must be finite");
    DeviceGuard dg(s->device);
    hipLaunchKernelGGL(fista_prox_input, dim3(Solver::grid_for((long long)n)), dim3(256), 0, s->stream,
                       static_cast<const double2*>(u), static_cast<const double2*>(grad), inv_alpha, (long long)n, cplx,
                       out);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s->stream));
  });
}

int cbs_fista_momentum(cbs_handle h, const void* sp, int cplx, void* x, void* u, void* ri, uint64_t n, double c,
                       double n_m) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(sp && x && u && ri, CBS_ERR_ARG, "null FISTA argument");
    require(n >= 1 && n < (1ull << 40), CBS_ERR_ARG, "FISTA length out of range");
    require(cplx == 0 || cplx == 1, CBS_ERR_ARG, "cplx must be 0 or 1");
    require(std::isfinite(c) && std::isfinite(n_m) && n_m > 0, CBS_ERR_ARG, "momentum and RI_bg must be finite");
    DeviceGuard dg(s->device);
    hipLaunchKernelGGL(fista_momentum, dim3(Solver::grid_for((long long)n)), dim3(256), 0, s->stream, sp, cplx,
                       static_cast<double2*>(x), static_cast<double2*>(u), static_cast<double2*>(ri), (long long)n, c,
                       n_m);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s->stream));
  });
}

int cbs_set_convergence(cbs_handle h, double tol) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(tol >= 0.0 && std::isfinite(tol), CBS_ERR_ARG, "convergence tolerance must be finite and >= 0");
    DeviceGuard dg(s->device);
    HIP_CHECK(hipStreamSynchronize(s->stream));
    s->conv_tol = tol;
    s->drop_graph();
  });
}

int cbs_get_iterations(cbs_handle h, uint32_t* out, uint32_t n_rhs) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(out != nullptr, CBS_ERR_ARG, "null output");
    require(n_rhs == s->last_iters.size(), CBS_ERR_ARG, "n_rhs does not match the last solve");
    std::memcpy(out, s->last_iters.data(), s->last_iters.size() * sizeof(uint32_t));
  });
}

int cbs_get_norms(cbs_handle h, double* out, uint32_t n_rhs, uint32_t n_iter) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(out != nullptr, CBS_ERR_ARG, "null output");
    require(s->tracking(), CBS_ERR_STATE, "track_norms is off (and no convergence test is set)");
    require((int)n_rhs == s->last_norm_rhs && (int)n_iter == s->last_norm_iter, CBS_ERR_ARG,
            "n_rhs/n_iter do not match the last solve");
    std::memcpy(out, s->last_norms.data(), s->last_norms.size() * sizeof(double));
  });
}

int cbs_get_kernel_times(cbs_handle h, double* out_ms, uint64_t* out_count, uint32_t n) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(out_ms && out_count, CBS_ERR_ARG, "null output");
    for (uint32_t k = 0; k < n && k < 3; ++k) {
      out_ms[k] = s->ktime_ms[k];
      out_count[k] = s->kcount[k];
    }
  });
}

int cbs_get_stream(cbs_handle h, void** stream) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(stream != nullptr, CBS_ERR_ARG, "null output");
    *stream = (void*)s->stream;
  });
}

int cbs_footprint(cbs_handle h, uint32_t n_rhs, uint64_t* bytes) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(bytes != nullptr, CBS_ERR_ARG, "null output");
    *bytes = s->footprint(n_rhs);
  });
}

int cbs_get_plan(cbs_handle h, int32_t* form, int32_t* a, int32_t* b, int32_t* tile_t) {
  return guarded([&] {
    Solver* s = lookup(h);
    require(s->have_potential, CBS_ERR_STATE, "plan is chosen at cbs_set_potential");
    if (form) *form = s->plan.form;
    if (a) *a = s->plan.a;
    if (b) *b = s->plan.b;
    if (tile_t) *tile_t = s->plan.t;
  });
}

int cbs_autotune(cbs_handle h, uint32_t n_rhs, uint32_t iterations, double* out_ms) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    const double ms = s->autotune(n_rhs, iterations);
    if (out_ms) *out_ms = ms;
  });
}

int cbs_set_io_kernels(cbs_handle h, const void* refocus, const void* kernel_trans, const void* kernel_ref, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->set_io_kernels(reinterpret_cast<const float2*>(refocus), reinterpret_cast<const float2*>(kernel_trans),
                      reinterpret_cast<const float2*>(kernel_ref), mem);
  });
}

int cbs_set_io_kernels_octant(cbs_handle h, const void* refocus_octant, const void* trans_octant, int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->set_io_kernels_octant(reinterpret_cast<const float2*>(refocus_octant),
                             reinterpret_cast<const float2*>(trans_octant), mem);
  });
}

int cbs_solve_2d(cbs_handle h, const void* source_2d, uint32_t n_rhs, void* trans_2d, void* ref_2d, void* field_3d,
                 int mem) {
  return guarded([&] {
    Solver* s = lookup(h);
    DeviceGuard dg(s->device);
    s->solve_2d(reinterpret_cast<const float2*>(source_2d), n_rhs, reinterpret_cast<float2*>(trans_2d),
                reinterpret_cast<float2*>(ref_2d), reinterpret_cast<float2*>(field_3d), mem);
  });
}

}  // extern "C"
