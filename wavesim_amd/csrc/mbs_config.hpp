// Launch configuration shared by the host runtime and the kernel variants.
#pragma once
#include "mbs_kernels.hpp"

namespace wsim {

constexpr int kThreads = 256;  // threads per workgroup of every axis pass
// registers (complex values) per thread: 32 for scalar fields, 24 for the
// 3-component vector field
constexpr int ept_for(int C) { return C == 1 ? 32 : 24; }

using KernelFn = void (*)(const PassArgs*);

// the pass variants the solver launches, as numbers (contig, mode, prologue,
// epilogue) with mode 0 fwd / 1 inv / 2 mid, prologue 1 = source + ramps,
// epilogue 1 = mix, 2 = ramps only
#define WSIM_FOR_EACH_VARIANT(X) \
  X(1, 0, 1, 0)                  \
  X(1, 1, 0, 1)                  \
  X(1, 1, 0, 2)                  \
  X(1, 2, 1, 1)                  \
  X(1, 2, 1, 2)                  \
  X(0, 0, 0, 0)                  \
  X(0, 1, 0, 0)                  \
  X(0, 2, 0, 0)

#define WSIM_DECLARE_KPTR(ct, m, p, e)              \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_1(); \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_3();
WSIM_FOR_EACH_VARIANT(WSIM_DECLARE_KPTR)

}  // namespace wsim
