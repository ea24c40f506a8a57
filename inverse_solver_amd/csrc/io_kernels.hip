// 2D input/output projection kernels.
#include "io_kernels.hpp"
