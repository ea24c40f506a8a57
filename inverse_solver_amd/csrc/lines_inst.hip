// Explicit instantiations of the per-axis line FFTs (fft_rows, fft_strided) and their
// registry: every 5-smooth length from 16 to 1296 whose two-step plan has both factors
// <= 36 (all but 1000, 1125, 1200, 1215, 1250 and 1280).  Used by the per-axis form and
// the 2D convolution-grid transforms; lengths absent here run the run-time-planned
// transforms (fft_rt.hip, fft_blue.hip).
#include "cbs_passes.hpp"
#include "kernels_misc.hpp"

namespace cbs {

constexpr int kNT = 256;

#define LINE_LIST(X)                                                                                           \
  X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) X(45) X(48) X(50) X(54) X(60) X(64) X(72) X(75)  \
  X(80) X(81) X(90) X(96) X(100) X(108) X(120) X(125) X(128) X(135) X(144) X(150) X(160) X(162) X(180)         \
  X(192) X(200) X(216) X(225) X(240) X(243) X(250) X(256) X(270) X(288) X(300) X(320) X(324) X(360) X(375)     \
  X(384) X(400) X(405) X(432) X(450) X(480) X(486) X(500) X(512) X(540) X(576) X(600) X(625) X(640) X(648)     \
  X(675) X(720) X(729) X(750) X(768) X(800) X(810) X(864) X(900) X(960) X(972) X(1024) X(1080) X(1152) X(1296)

constexpr int rows_for(int n) { return n <= 64 ? 16 : (n <= 256 ? 8 : (n <= 576 ? 4 : 2)); }
// 16-wide column tiles (128-byte rows) up to n = 576 (78 KiB with the twiddles: two blocks
// per CU); at 576 this took the y transforms of the 576^2 x 320 grid from 12.1 to 9.8 ms
constexpr int strided_t(int n) { return n <= 576 ? 16 : 8; }

#define RW_ENTRY(N) {&fft_rows<N, rows_for(N), kNT>, N, rows_for(N), kNT},
#define ST_ENTRY(N) {&fft_strided<N, strided_t(N), kNT>, N, strided_t(N), kNT},

static const RowsInfo kRW[] = {LINE_LIST(RW_ENTRY)};
static const StridedInfo kST[] = {LINE_LIST(ST_ENTRY)};

const RowsInfo* find_fft_rows(int n) {
  for (const auto& e : kRW)
    if (e.n == n) return &e;
  return nullptr;
}
const StridedInfo* find_fft_strided(int n, int t) {
  for (const auto& e : kST)
    if (e.n == n && (t <= 0 || e.t == t)) return &e;
  return nullptr;
}

}  // namespace cbs
