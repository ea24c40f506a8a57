// Explicit instantiations of the templated passes and a lookup registry.
// The shape lists cover the simulation grids of the supported workloads
// (sim 96^3, 160^3, 128x128x135, 128^3, 192^2x128, 288^2x160, 576^2x320) and the small grids
// used by the tests; any other grid runs the per-axis form on the line FFTs of
// lines_inst.hip (every 5-smooth length up to 1296) or the run-time-planned ones
// (fft_rt.hip, fft_blue.hip).
#include "cbs_passes.hpp"
#include "kernels_misc.hpp"

namespace cbs {

// (NX, A, NT): x length, rows of the y split, threads per block
#define PR_LIST(X) \
  X(24, 24, 256) X(24, 12, 256) X(24, 6, 256) X(32, 32, 256) X(32, 16, 256) X(40, 40, 256) X(40, 20, 256) \
  X(96, 96, 256) X(96, 96, 512) X(96, 48, 256) X(96, 48, 512) X(96, 32, 256) X(96, 24, 256) \
  X(128, 64, 256) X(128, 64, 512) X(128, 32, 256) \
  X(160, 40, 256) X(160, 40, 512) X(160, 32, 256) X(160, 32, 512) X(160, 20, 256) X(160, 80, 512) \
  X(192, 48, 256) X(192, 48, 512) X(192, 32, 256) \
  X(288, 32, 256) X(288, 32, 512)

// four-pass form (contiguous row blocks, y and z transforms in their own passes): (NX, rows, NT)
#define PRR_LIST(X) X(24, 8, 256) X(24, 6, 256) X(576, 16, 512) X(576, 8, 256) X(512, 16, 512) X(512, 8, 256)

// (B, NZ, T, NT): b-part of the y split, z length, x tile, threads per block
#define PK_LIST(X) \
  X(1, 18, 8, 256) X(1, 24, 8, 256) X(1, 24, 16, 256) X(2, 24, 8, 256) X(4, 24, 8, 256) X(1, 20, 8, 256) \
  X(1, 32, 16, 256) X(2, 32, 16, 256) X(1, 40, 16, 256) X(2, 40, 16, 256) X(1, 30, 8, 256) X(1, 36, 8, 256) \
  X(1, 96, 32, 256) X(1, 96, 32, 512) X(1, 96, 16, 256) X(2, 96, 16, 256) X(2, 96, 32, 512) X(3, 96, 16, 256) \
  X(2, 96, 8, 256) X(3, 96, 8, 256) X(4, 96, 8, 256) X(4, 96, 16, 256) \
  X(2, 135, 32, 256) X(2, 135, 32, 512) X(2, 135, 16, 256) X(4, 135, 8, 256) \
  X(2, 128, 32, 256) X(4, 128, 16, 256) X(6, 128, 8, 256) X(4, 128, 8, 256) \
  X(4, 160, 16, 256) X(4, 160, 16, 512) X(4, 160, 16, 1024) X(5, 160, 8, 256) X(8, 160, 8, 256) X(4, 160, 8, 256) \
  X(2, 160, 32, 512) X(2, 160, 32, 256) \
  X(9, 160, 4, 256) \
  X(1, 320, 16, 256) X(1, 320, 32, 512) X(1, 256, 16, 256) X(1, 256, 32, 512)

// vector runs (three field components per block): (B, NZ, T, NT)
#define PKV_LIST(X) \
  X(1, 18, 8, 256) X(1, 20, 8, 256) X(1, 24, 8, 256) X(2, 24, 8, 256) X(4, 24, 8, 256) X(1, 30, 8, 256) \
  X(1, 32, 16, 256) X(2, 32, 16, 256) X(1, 36, 8, 256) X(1, 40, 16, 256) X(2, 40, 16, 256) \
  X(1, 96, 16, 256) X(2, 96, 16, 512) X(3, 96, 8, 256) \
  X(2, 128, 16, 512) X(4, 128, 8, 512) X(2, 128, 8, 256) \
  X(2, 135, 16, 512) X(2, 135, 8, 256) X(4, 135, 8, 512) \
  X(2, 160, 8, 512) X(4, 160, 8, 1024) X(2, 160, 16, 1024) \
  X(1, 320, 8, 256) X(1, 320, 16, 512) X(1, 256, 16, 512)

// vector runs, one component at a time through LDS (pass_kspace_seq3): (B, NZ, T, NT)
#define PKS_LIST(X) X(4, 160, 16, 1024) X(4, 160, 8, 512)

// LDS of pass_real: the tile plus its per-axis tables (x: twiddle; rows: twiddle, ramp, mask)
constexpr int pr_lds(int nx, int a) { return a * row_pitch(nx) * 8 + nx * 8 + a * 20; }
#define PR_ENTRY(NX, A, NT)                                                                                   \
  {&pass_real<NX, A, NT, 0, 0, 0>, &pass_real<NX, A, NT, 0, 1, 0>, &pass_real<NX, A, NT, 0, 0, 1>,                  \
   &pass_real<NX, A, NT, 0, 1, 1>, NX, A, NT, pr_lds(NX, A), 0},
#define PRR_ENTRY(NX, A, NT) \
  {&pass_real<NX, A, NT, 1, 0, 0>, &pass_real<NX, A, NT, 1, 1, 0>, nullptr, nullptr, NX, A, NT, pr_lds(NX, A), 1},
#define PK_ENTRY(B, NZ, T, NT) {&pass_kspace<B, NZ, T, NT, 1>, B, NZ, T, NT, (int)(NZ * B * T * sizeof(float2)), 1},
#define PKV_ENTRY(B, NZ, T, NT) \
  {&pass_kspace<B, NZ, T, NT, 3>, B, NZ, T, NT, (int)((3 * NZ * B * T + 2 * NZ) * sizeof(float2)), 3},
#define PKS_ENTRY(B, NZ, T, NT) \
  {&pass_kspace_seq3<B, NZ, T, NT>, B, NZ, T, NT, (int)(NZ * B * T * sizeof(float2)), 3},

static const PassRInfo kPR[] = {PR_LIST(PR_ENTRY) PRR_LIST(PRR_ENTRY)};
static const PassKInfo kPK[] = {PK_LIST(PK_ENTRY) PKV_LIST(PKV_ENTRY) PKS_LIST(PKS_ENTRY)};

const PassRInfo* find_pass_real(int nx, int a) {
  for (const auto& e : kPR)
    if (e.nx == nx && e.a == a && !e.rows) return &e;
  return nullptr;
}
const PassKInfo* find_pass_kspace(int b, int nz, int t, int pole) {
  for (const auto& e : kPK)
    if (e.b == b && e.nz == nz && e.t == t && e.pole == pole) return &e;
  return nullptr;
}
int list_pass_real(const PassRInfo** out) {
  *out = kPR;
  return (int)(sizeof(kPR) / sizeof(kPR[0]));
}
int list_pass_kspace(const PassKInfo** out) {
  *out = kPK;
  return (int)(sizeof(kPK) / sizeof(kPK[0]));
}

}  // namespace cbs
