// Explicit instantiations of the per-axis form's fused passes on compiled line lengths
// (line_rows_update, line_strided_green) and their registry: the lengths of
// lines_inst.hip.
#include "cbs_passes.hpp"
#include "kernels_misc.hpp"

namespace cbs {

#define FUSED_LIST(X)                                                                                          \
  X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) X(45) X(48) X(50) X(54) X(60) X(64) X(72) X(75)  \
  X(80) X(81) X(90) X(96) X(100) X(108) X(120) X(125) X(128) X(135) X(144) X(150) X(160) X(162) X(180)         \
  X(192) X(200) X(216) X(225) X(240) X(243) X(250) X(256) X(270) X(288) X(300) X(320) X(324) X(360) X(375)     \
  X(384) X(400) X(405) X(432) X(450) X(480) X(486) X(500) X(512) X(540) X(576) X(600) X(625) X(640) X(648)     \
  X(675) X(720) X(729) X(750) X(768) X(800) X(810) X(864) X(900) X(960) X(972) X(1024) X(1080) X(1152) X(1296)

// rows per block and x tile as the plain line FFTs (lines_inst.hip)
constexpr int fused_rows(int n) { return n <= 64 ? 16 : (n <= 256 ? 8 : (n <= 576 ? 4 : 2)); }
constexpr int fused_t(int n) { return n <= 576 ? 16 : 8; }

#define RU_ENTRY(N) {&line_rows_update<N, fused_rows(N), 256>, N, fused_rows(N), 256},
#define SG_ENTRY(N) {&line_strided_green<N, fused_t(N), 256>, N, fused_t(N), 256},

static const RowsUpdateInfo kRU[] = {FUSED_LIST(RU_ENTRY)};
static const StridedGreenInfo kSG[] = {FUSED_LIST(SG_ENTRY)};

const RowsUpdateInfo* find_rows_update(int n) {
  for (const auto& e : kRU)
    if (e.n == n) return &e;
  return nullptr;
}
const StridedGreenInfo* find_strided_green(int n) {
  for (const auto& e : kSG)
    if (e.n == n) return &e;
  return nullptr;
}

}  // namespace cbs
