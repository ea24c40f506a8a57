// 2D input/output projections (refocus of the incident field, transmitted and
// reflected field accumulation).
#pragma once
#include "cbs_kernels.hpp"

namespace cbs {

struct IoState {
  bool ready = false;
  unsigned long long footprint(int nb) const { (void)nb; return 0ull; }
};

}  // namespace cbs
