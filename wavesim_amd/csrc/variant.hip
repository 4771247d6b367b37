// One instantiation of axis_pass_kernel, selected by -D flags at build time
// (WSIM_CONTIG, WSIM_MODE, WSIM_PRO, WSIM_EPI, WSIM_C, WSIM_COMPACT), so
// that the variants compile in parallel.  The host runtime gets the kernel
// through wsim_kptr_<contig>_<mode>_<pro>_<epi>_<C>_<family>(); family
// (WSIM_COMPACT) 0 and 1 are the complex64 kernels, 2 the complex128 ones.
#include <type_traits>

#include "mbs_kernels.hpp"
#include "mbs_config.hpp"

#define WSIM_CAT7(a, b, c, d, e, f, g) a##_##b##_##c##_##d##_##e##_##f##_##g
#define WSIM_NAME(a, b, c, d, e, f, g) WSIM_CAT7(a, b, c, d, e, f, g)

namespace wsim {
KernelFn WSIM_NAME(wsim_kptr, WSIM_CONTIG, WSIM_MODE, WSIM_PRO, WSIM_EPI, WSIM_C, WSIM_COMPACT)() {
  using F = std::conditional_t<WSIM_COMPACT == kFamilyF64, double, float>;
  return axis_pass_kernel<(WSIM_CONTIG != 0), WSIM_MODE, WSIM_PRO, WSIM_EPI, WSIM_C, kMaxThreads,
                          ept_for(WSIM_C, WSIM_COMPACT), F>;
}
}  // namespace wsim
