// Launch configuration shared by the host runtime and the kernel variants.
#pragma once
#include "mbs_kernels.hpp"

namespace wsim {

// workgroup size: every variant is compiled for up to kMaxThreads threads and
// launched with 256 or 512 (a per-axis plan parameter).  Strided passes hold
// wide tiles (LDS-limited to a few workgroups per CU) and default to 512 to
// keep more waves resident; contiguous passes default to 256 (more, smaller
// workgroups for small grids).
constexpr int kMaxThreads = 512;
constexpr int default_threads(int contig) { return contig ? 256 : 512; }
// kernel families: 0 complex64, 1 complex64 compact, 2 complex128
constexpr int kFamilyF64 = 2;
// registers (complex values) per thread and stage: 32 for scalar fields, 24
// for the 3-component vector field; the compact family halves them, which
// lowers the kernel's register count (set by its largest radix path) and
// raises occupancy for tiles that fit.  complex128 values take twice the
// registers: that family holds the compact counts (the same register budget
// as the complex64 family)
constexpr int ept_for(int C, int family) { return C == 1 ? (family ? 16 : 32) : (family ? 12 : 24); }

using KernelFn = void (*)(const PassArgs*);

// the pass variants the solver launches, as numbers (contig, mode, prologue,
// epilogue) with mode 0 fwd / 1 inv / 2 mid, prologue 1 = source + ramps,
// epilogue 1 = mix, 2 = ramps only, 3 = mix + the next iteration's first pass,
// 4 = the PSTD time step
#define WSIM_FOR_EACH_VARIANT(X) \
  X(1, 0, 1, 0)                  \
  X(1, 1, 0, 1)                  \
  X(1, 1, 0, 3)                  \
  X(1, 1, 0, 2)                  \
  X(1, 2, 1, 1)                  \
  X(1, 2, 1, 2)                  \
  X(1, 1, 0, 4)                  \
  X(1, 1, 0, 5)                  \
  X(1, 2, 1, 4)                  \
  X(0, 0, 0, 0)                  \
  X(0, 1, 0, 0)                  \
  X(0, 2, 0, 0)

// the variants built in complex128 as well (the MBS passes; PSTD runs in
// complex64 only)
#define WSIM_FOR_EACH_VARIANT_F64(X) \
  X(1, 0, 1, 0)                      \
  X(1, 1, 0, 1)                      \
  X(1, 1, 0, 3)                      \
  X(1, 1, 0, 2)                      \
  X(1, 2, 1, 1)                      \
  X(1, 2, 1, 2)                      \
  X(0, 0, 0, 0)                      \
  X(0, 1, 0, 0)                      \
  X(0, 2, 0, 0)

#define WSIM_DECLARE_KPTR(ct, m, p, e)                \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_1_0(); \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_1_1(); \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_3_0(); \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_3_1();
WSIM_FOR_EACH_VARIANT(WSIM_DECLARE_KPTR)
#define WSIM_DECLARE_KPTR_F64(ct, m, p, e)            \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_1_2(); \
  KernelFn wsim_kptr_##ct##_##m##_##p##_##e##_3_2();
WSIM_FOR_EACH_VARIANT_F64(WSIM_DECLARE_KPTR_F64)

}  // namespace wsim

namespace wsim {
// A pass kernel compiled with its FFT plan as compile-time constants
// (FixedPlan): the host launches it in place of the run-time-plan kernel for
// launches of exactly this shape.  The table (generated at build time from
// the shipped tuned plans) ends with an entry whose `get` is null.
struct FixedKernel {
  int contig, mode, pro, epi, C, family, n, ncol, threads, nst;
  int rad[MAXST];
  KernelFn (*get)();
};
const FixedKernel* fixed_kernel_table();
}  // namespace wsim
