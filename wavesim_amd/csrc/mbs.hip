// Host runtime and C ABI of the MBS solver (see include/wavesim_mbs.h).
//
// One wsim_solver owns, on one GPU:
//   gamma [nvox]            shared by every rhs (WaveSimBase.m:73)
//   E', dE', T, S           [nrhs][C][nvox] complex64 each
//   per-wiggle 1-D tables   pe/ramp per axis (WaveSimBase.m:300-316)
//   twiddle tables          per axis, forward and inverse
//   DevState                the iteration counter and per-rhs stop state
// and a list of kernel launches that make one MBS iteration
// (WaveSimBase.m:165-189): 2D-1 axis passes plus the control kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <complex>
#include <cstring>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "mbs_config.hpp"
#include "mbs_kernels.hpp"
#include "wavesim_mbs.h"
#include "wavesim_pstd.h"

namespace wsim {

static thread_local std::string g_err;

// the message wsim_last_error() returns, for the entry points of other
// translation units of the library
void set_last_error(const std::string& msg) { g_err = msg; }

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess)                                                                  \
      throw Error(e_ == hipErrorOutOfMemory ? WSIM_ERR_NOMEM : WSIM_ERR_HIP,               \
                  std::string(#x) + ": " + hipGetErrorString(e_) + " (line " +             \
                      std::to_string(__LINE__) + ")");                                     \
  } while (0)

static void require(bool ok, const std::string& msg, int code = WSIM_ERR_ARG) {
  if (!ok) throw Error(code, msg);
}

// a float2 / double2 on the host
template <class V>
static V mkv_host(real_of<V> x, real_of<V> y) {
  V r;
  r.x = x;
  r.y = y;
  return r;
}

// ----------------------------------------------------------------------------
// index arithmetic
// ----------------------------------------------------------------------------
static FastDiv make_fastdiv(uint32_t d) {
  FastDiv f{1u, 0u, 0u};
  if (d <= 1) return f;
  uint32_t s = 0;
  while ((1ull << s) < d) ++s;  // ceil(log2 d) >= 1
  const uint64_t m = ((1ull << (31 + s)) + d - 1) / d;
  f.d = d;
  f.m = (uint32_t)m;
  f.sh = s - 1;
  return f;
}

static uint32_t host_fdiv(uint32_t x, const FastDiv& f) {
  if (f.d == 1) return x;
  return (uint32_t)(((uint64_t)x * f.m) >> 32) >> f.sh;
}

// radices for a transform of length n: the power-of-two part split as evenly
// as possible into radix 2..16 stages, then the odd primes 3, 5, 7, 11
static std::vector<int> default_radices(int n, int C = 1) {
  std::vector<int> r;
  {
    // experiments/autotuning: WSIM_RADICES_<n>=r0,r1,... overrides the plan
    const std::string key = "WSIM_RADICES_" + std::to_string(n);
    if (const char* e = std::getenv(key.c_str())) {
      long long p = 1;
      for (const char* q = e; *q;) {
        const int x = std::atoi(q);
        if (x > 1) {
          r.push_back(x);
          p *= x;
        }
        while (*q && *q != ',') ++q;
        if (*q == ',') ++q;
      }
      if (p == n) return r;
      r.clear();
    }
  }
  int p2 = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++p2;
  }
  if (p2 > 0) {
    const int maxbits = radix_fits(16, C) ? 4 : 3;
    const int k = (p2 + maxbits - 1) / maxbits;
    for (int i = 0; i < k; ++i) {
      const int bits = p2 / k + (i < p2 % k ? 1 : 0);
      r.push_back(1 << bits);
    }
  }
  for (int p : {11, 7, 5, 3}) {
    while (n % p == 0) {
      n /= p;
      r.push_back(p);
    }
  }
  std::sort(r.begin(), r.end(), [](int a, int b) {
    const bool pa = (a & (a - 1)) == 0, pb = (b & (b - 1)) == 0;
    if (pa != pb) return pa;  // powers of two first
    return a > b;
  });
  if (n != 1) r.clear();  // a prime factor above 11: unsupported
  return r;
}

struct AxisTables {
  AxisPlan plan;
  std::vector<float2> twf, twi;      // complex64 solvers
  std::vector<double2> twf_d, twi_d; // complex128 solvers (same angles)
};

static AxisTables build_axis(int n, const std::vector<int>& radices) {
  AxisTables t{};
  AxisPlan& p = t.plan;
  p.n = n;
  p.nst = (int)radices.size();
  require(p.nst <= MAXST, "too many FFT stages", WSIM_ERR_UNSUPPORTED);
  auto fill = [&](Stage* st, std::vector<float2>& tw, std::vector<double2>& twd, const std::vector<int>& rad,
                  int sign) {
    int Ns = 1;
    for (size_t s = 0; s < rad.size(); ++s) {
      const int R = rad[s];
      st[s].R = R;
      st[s].Ns = Ns;
      st[s].nb = n / R;
      st[s].tw = (int)tw.size();
      st[s].fNs = make_fastdiv((uint32_t)Ns);
      st[s].fnb = make_fastdiv((uint32_t)(n / R));
      // [r-1][k] order: the butterflies of one wave have consecutive k, so
      // their twiddle reads from LDS hit consecutive banks
      for (int r = 1; r < R; ++r)
        for (int k = 0; k < Ns; ++k) {
          const double ang = sign * 2.0 * M_PI * (double)r * (double)k / ((double)Ns * R);
          tw.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
          twd.push_back(make_double2(std::cos(ang), std::sin(ang)));
        }
      Ns *= R;
    }
  };
  std::vector<int> rev(radices.rbegin(), radices.rend());
  fill(p.f, t.twf, t.twf_d, radices, -1);
  fill(p.i, t.twi, t.twi_d, rev, +1);
  if (t.twf.empty()) t.twf.push_back(make_float2(1.f, 0.f));
  if (t.twi.empty()) t.twi.push_back(make_float2(1.f, 0.f));
  if (t.twf_d.empty()) t.twf_d.push_back(make_double2(1.0, 0.0));
  if (t.twi_d.empty()) t.twi_d.push_back(make_double2(1.0, 0.0));
  return t;
}

// One block: reduce the energy partials of every rhs, record diff_energy the
// way Simulation.next does (Simulation.m:193-217) and advance the iteration.
constexpr int kControlThreads = 1024;
constexpr int kXcdAutoBlocks = 2048;  // automatic XCD tile ranges up to this many workgroups per rhs
// largest grid: byte offsets 8 * (nvox - 1) of one component fit 32 bits
// (16 * (nvox - 1) for complex128 fields)
constexpr long long kMaxVoxels = (1ll << 29);
constexpr long long kMaxVoxelsF64 = (1ll << 28);

// FFT plans tuned on MI355X for given problem shapes, applied when a solver of
// that shape is created (the rows are generated from tuned_plans.json by the
// build; a build without them uses the default plans)
struct ShippedPlan {
  int n0, n1, n2, c, nrhs, nw, f64;  // f64: the plan of complex128 solvers of that shape
  const char* plan;
};
static const ShippedPlan kShippedPlans[] = {
#if __has_include("shipped_plans.inc")
#include "shipped_plans.inc"
#endif
    {0, 0, 0, 0, 0, 0, 0, nullptr}};
// Decisions every thread takes (return, skip a stopped rhs) come from one
// thread's read of the state, broadcast through LDS: the state is written by
// thread 0 of this workgroup, possibly in the previous step of the same launch
// (control_cycle_kernel), and all threads must agree on the barriers they
// reach.
__device__ void control_one(DevState* st, const double* partials, int nblk) {
  constexpr int NW = kControlThreads / 64;
  __shared__ double red[NW];
  __shared__ int flag;
  if (threadIdx.x == 0) flag = st->all_stopped;
  __syncthreads();
  if (flag) return;
  const int it = st->it;  // used by thread 0 only
  const int nrhs = st->nrhs;
  int all = 1;
  for (int r = 0; r < nrhs; ++r) {
    __syncthreads();  // every thread has read the previous rhs's flag
    if (threadIdx.x == 0) flag = st->stopped[r];
    __syncthreads();
    if (flag) continue;
    // four independent accumulators keep four loads in flight per thread;
    // the summation order is fixed, so the result is deterministic
    const double* pr = partials + (size_t)r * nblk;
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0;
    int i = threadIdx.x;
    for (; i + 3 * kControlThreads < nblk; i += 4 * kControlThreads) {
      e0 += pr[i];
      e1 += pr[i + kControlThreads];
      e2 += pr[i + 2 * kControlThreads];
      e3 += pr[i + 3 * kControlThreads];
    }
    for (; i < nblk; i += kControlThreads) e0 += pr[i];
    double e = (e0 + e1) + (e2 + e3);
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
      double en = 0.0;
      for (int k = 0; k < NW; ++k) en += red[k];
      const size_t h = (size_t)r * st->hist_cap + (size_t)(it - 1);
      st->energy_all[h] = en;
      if (st->calc_energy) st->last_energy[r] = en;
      const double de = st->last_energy[r];
      st->diff[h] = de;
      const bool can = (it % st->nw) == 0;
      const double ref = st->ref_energy[r];
      const double d1 = ref > 0.0 ? ref : st->diff[(size_t)r * st->hist_cap];
      // failure exit (no counterpart in the reference, which runs such a
      // solve to its cap): a non-finite increment energy, or one that grew
      // over each of the last two wiggle cycles to above 100 times the first
      // iteration's (or the fixed reference) energy -- a diverging series,
      // e.g. epsilon forced below the convergence bound (WaveSimBase.m:63)
      bool diverged = !isfinite(en);
      if (can && it > 2 * st->nw) {
        const double* ea = st->energy_all + (size_t)r * st->hist_cap;
        const double e1 = ea[it - 1 - st->nw], e2 = ea[it - 1 - 2 * st->nw];
        diverged = diverged || (en > e1 && e1 > e2 && en > 100.0 * fmax(ea[0], ref));
      }
      if (diverged) {
        st->stopped[r] = 1;
        st->converged[r] = -1;
        st->n_iter[r] = it;
      } else if (can && de / d1 < st->threshold) {
        st->stopped[r] = 1;
        st->converged[r] = 1;
        st->n_iter[r] = it;
      } else if ((can && it >= st->max_iter) || it >= st->hist_cap) {
        st->stopped[r] = 1;
        st->converged[r] = 0;
        st->n_iter[r] = it;
      }
    }
    __syncthreads();
    all &= st->stopped[r];
  }
  if (threadIdx.x == 0) {
    st->calc_energy = ((it - 1) % st->interval) == 0;
    st->it = it + 1;
    st->all_stopped = all;
    if (st->host_stopped != nullptr && all) *st->host_stopped = 1;
  }
}

__global__ void __launch_bounds__(kControlThreads) control_kernel(DevState* st, const double* partials, int nblk) {
  control_one(st, partials, nblk);
}

// The control steps of L iterations at once, in order (iteration j's energy
// partials in slot j, `slot` doubles apart).  A solve runs one per wiggle
// cycle: the stop test applies at multiples of Nw only (Simulation.m:202), so
// no rhs stops inside a cycle except on the failure exit (whose iteration is
// still the one recorded)
//
// A single rhs with at most one partial per thread and at most kCycleFast
// iterations takes a short path: all L energies are reduced at once (the same
// order of additions as control_one, so the energies are bit-identical), and
// thread 0 then runs the L stop decisions with the state in registers, so the
// steps do not wait on each other's global-memory round trips.
constexpr int kCycleFast = 8;
__global__ void __launch_bounds__(kControlThreads) control_cycle_kernel(DevState* st, const double* partials, int nblk,
                                                                        int L, long long slot) {
  if (st->nrhs != 1 || nblk > kControlThreads || L > kCycleFast) {
    for (int j = 0; j < L; ++j) {
      control_one(st, partials + (size_t)j * slot, nblk);
      __syncthreads();
    }
    return;
  }
  constexpr int NW = kControlThreads / 64;
  __shared__ double red[kCycleFast][NW];
  double v[kCycleFast];
#pragma unroll
  for (int j = 0; j < kCycleFast; ++j) {
    v[j] = 0.0;
    if (j < L && (int)threadIdx.x < nblk) v[j] += partials[(size_t)j * slot + threadIdx.x];
  }
#pragma unroll
  for (int j = 0; j < kCycleFast; ++j) {
    if (j < L) {
      double e = v[j];
      for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
      if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = e;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0 || st->all_stopped) return;
  // the decisions of control_one for rhs 0, state in registers
  int it = st->it, calc = st->calc_energy, stopped = st->stopped[0], all_stopped = 0;
  double last = st->last_energy[0], diff0 = st->diff[0];
  const double ref = st->ref_energy[0], thr = st->threshold;
  const int nw = st->nw, interval = st->interval, hist_cap = st->hist_cap;
  const long long max_iter = st->max_iter;
  double* ea = st->energy_all;
  for (int j = 0; j < L && !all_stopped; ++j) {
    int all = 1;
    if (!stopped) {
      double en = 0.0;
      for (int k = 0; k < NW; ++k) en += red[j][k];
      ea[it - 1] = en;
      if (calc) last = en;
      const double de = last;
      st->diff[it - 1] = de;
      if (it == 1) diff0 = de;
      const bool can = (it % nw) == 0;
      const double d1 = ref > 0.0 ? ref : diff0;
      bool diverged = !isfinite(en);
      if (can && it > 2 * nw) {
        const double e1 = ea[it - 1 - nw], e2 = ea[it - 1 - 2 * nw];
        diverged = diverged || (en > e1 && e1 > e2 && en > 100.0 * fmax(ea[0], ref));
      }
      int conv = 2;  // 2: running
      if (diverged)
        conv = -1;
      else if (can && de / d1 < thr)
        conv = 1;
      else if ((can && it >= max_iter) || it >= hist_cap)
        conv = 0;
      if (conv != 2) {
        stopped = 1;
        st->stopped[0] = 1;
        st->converged[0] = conv;
        st->n_iter[0] = it;
      }
      all = stopped;
    }
    calc = ((it - 1) % interval) == 0;
    ++it;
    all_stopped = all;
  }
  st->last_energy[0] = last;
  st->calc_energy = calc;
  st->it = it;
  st->all_stopped = all_stopped;
  if (st->host_stopped != nullptr && all_stopped) *st->host_stopped = 1;
}

// grouped launches: workgroup e runs the control step of problem e and
// publishes its all-stopped flag into one array the host reads at once
struct CtlEntry {
  DevState* st;
  const double* partials;
  int nblk;
  int* done;
};
__global__ void __launch_bounds__(kControlThreads) control_group_kernel(const CtlEntry* __restrict__ ce) {
  const CtlEntry& c = ce[blockIdx.x];
  control_one(c.st, c.partials, c.nblk);
  if (threadIdx.x == 0) *c.done = c.st->all_stopped;
}

// PSTD source amplitude PSTD.default_source (PSTD.m:130-137): a CW phasor
// exp(-2 pi i N) switched on by a raised cosine over 100 optical cycles
__host__ __device__ inline void pstd_amplitude(double ncycle, double& re, double& im) {
  const double env = 0.5 - 0.5 * cos(M_PI * fmin(ncycle / 100.0, 1.0));
  const double ph = -2.0 * M_PI * ncycle;
  re = cos(ph) * env;
  im = sin(ph) * env;
}

// One block, after each PSTD step: reduce the energy partials, record the
// step energy relative to |A|^2 (PSTD.m:113-116) the way Simulation.next does
// (Simulation.m:193-217) with PSTD's can_terminate = real(A) > 0.9 (PSTD.m:120),
// then set up the next step's amplitude and expected phase advance.
__global__ void __launch_bounds__(kControlThreads) control_pstd_kernel(DevState* st, const double* partials,
                                                                       int nblk) {
  constexpr int NW = kControlThreads / 64;
  __shared__ double red[NW];
  if (st->all_stopped) return;
  double e = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kControlThreads) e += partials[i];
  for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double en = 0.0;
  for (int k = 0; k < NW; ++k) en += red[k];
  const int it = st->it;
  en /= st->amp_re * st->amp_re + st->amp_im * st->amp_im;
  st->energy_all[it - 1] = en;
  if (st->calc_energy) st->last_energy[0] = en;
  const double de = st->last_energy[0];
  st->diff[it - 1] = de;
  const bool can = st->amp_re > 0.9;
  if (!isfinite(en)) {  // failure exit: a non-finite step energy
    st->stopped[0] = 1;
    st->converged[0] = -1;
    st->n_iter[0] = it;
  } else if (can && de / st->diff[0] < st->threshold) {
    st->stopped[0] = 1;
    st->converged[0] = 1;
    st->n_iter[0] = it;
  } else if ((can && it >= st->max_iter) || it >= st->hist_cap) {
    st->stopped[0] = 1;
    st->converged[0] = 0;
    st->n_iter[0] = it;
  }
  st->all_stopped = st->stopped[0];
  st->calc_energy = ((it - 1) % st->interval) == 0;
  st->it = it + 1;
  // the next step's amplitude (PSTD.m:94-96) and phase advance (PSTD.m:114)
  const double pre = st->amp_re, pim = st->amp_im;
  double re, im;
  pstd_amplitude((double)(it + 1) / st->ipc, re, im);
  st->amp_re = re;
  st->amp_im = im;
  st->amp = make_float2((float)re, (float)im);
  const double dph = atan2(im, re) - atan2(pim, pre);
  st->phase = make_float2((float)cos(dph), (float)sin(dph));
}

// x[c][v] *= gamma[v] for the C components of one field (warm-start seeding)
template <class V>
__global__ void __launch_bounds__(256) gamma_mul_kernel(V* x, const V* g, long long nvox, int C) {
  const long long n = nvox * C;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    x[i] = cmul(x[i], g[i % nvox]);
}

// sparse source upload: the entries of a source that touch few voxels,
// unique indices, written into a field that is zero elsewhere
template <class V>
__global__ void __launch_bounds__(256) scatter_kernel(V* S, const unsigned long long* idx, const V* val, long long m) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < m; i += (long long)gridDim.x * 256)
    S[idx[i]] = val[i];
}

// The mix epilogue adds the increments to E' two at a time, on even
// iterations; a rhs whose last iteration was odd still lacks its last dE'.
// One block row per rhs; blocks of a rhs that needs nothing return at once.
template <class V>
__global__ void __launch_bounds__(256) finish_e_kernel(const DevState* st, V* E, const V* dE, long long fe) {
  const int r = blockIdx.y;
  const int last = st->stopped[r] ? st->n_iter[r] : st->it - 1;
  if ((last & 1) == 0) return;
  V* e = E + (size_t)r * fe;
  const V* d = dE + (size_t)r * fe;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < fe; i += (long long)gridDim.x * 256)
    e[i] = cadd(e[i], d[i]);
}

// ----------------------------------------------------------------------------
// kernel table
// ----------------------------------------------------------------------------
constexpr size_t LDS_CAP = 80 * 1024;  // two workgroups per 160 KB CU

// family: 0 complex64, 1 complex64 compact, kFamilyF64 complex128
static KernelFn select_kernel(bool contig, int mode, int pro, int epi, int C, int family) {
  if (family == kFamilyF64) {
#define WSIM_SELECT_F64(ct, m, p, e)                                   \
  if ((contig ? 1 : 0) == ct && mode == m && pro == p && epi == e)     \
    return C == 1 ? wsim_kptr_##ct##_##m##_##p##_##e##_1_2() : wsim_kptr_##ct##_##m##_##p##_##e##_3_2();
    WSIM_FOR_EACH_VARIANT_F64(WSIM_SELECT_F64)
#undef WSIM_SELECT_F64
    throw Error(WSIM_ERR_UNSUPPORTED, "no complex128 kernel for this pass variant");
  }
  const bool compact = family == 1;
#define WSIM_SELECT(ct, m, p, e)                                                                       \
  if ((contig ? 1 : 0) == ct && mode == m && pro == p && epi == e)                                     \
    return C == 1 ? (compact ? wsim_kptr_##ct##_##m##_##p##_##e##_1_1() : wsim_kptr_##ct##_##m##_##p##_##e##_1_0()) \
                  : (compact ? wsim_kptr_##ct##_##m##_##p##_##e##_3_1() : wsim_kptr_##ct##_##m##_##p##_##e##_3_0());
  WSIM_FOR_EACH_VARIANT(WSIM_SELECT)
#undef WSIM_SELECT
  throw Error(WSIM_ERR_UNSUPPORTED, "no kernel for this pass variant");
}

// LDS in front of the transform data: column table and the per-row ramp and
// coordinate tables (layout of axis_pass_kernel)
static size_t lds_header(int ncol, int n, int ntw, bool fuse, bool f64) {
  auto r16 = [](size_t x) { return (x + 15) & ~(size_t)15; };
  const size_t ci = f64 ? sizeof(ColInfoT<double>) : sizeof(ColInfo);
  const size_t cz = f64 ? sizeof(double2) : sizeof(float2), rz = f64 ? sizeof(double) : sizeof(float);
  return r16(ncol * ci) + (fuse ? 2 : 1) * r16((size_t)n * cz) + r16((size_t)n * rz) + (size_t)ntw * cz;
}

// contig LDS line pitch: one padding slot every 8 rows, plus one
static int contig_pitch(int n) { return n + (n >> WSIM_PAD_SHIFT) + 1; }

static int its_for(int C, int R, int family = 0) { return its_per_stage(ept_for(C, family), C, R); }

// radices of the run-time-plan kernels, and the composite register radices
// (R1 x R2 in registers) that only plan-specialised kernels run (scalar fields)
static bool radix_generic(int R) {
  return R == 2 || R == 3 || R == 4 || R == 5 || R == 7 || R == 8 || R == 11 || R == 16;
}
static bool radix_specialised(int R) { return R == 9 || R == 12 || R == 20 || R == 24 || R == 32 || R == 36; }

// every stage of the plan covers ncol lines of length n within the
// butterflies-per-thread budget of the kernel family
static bool tile_fits(const AxisPlan& pl, int n, int ncol, int C, int family, int NT) {
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.f[s].R;
    const long long need = ((long long)ncol * (n / R) + NT - 1) / NT;
    if (need > its_for(C, R, family)) return false;
  }
  return true;
}

struct Launch {
  KernelFn fn;
  KernelFn fn_generic;  // the run-time-plan kernel of the same pass (fn, unless fn is plan-specialised)
  dim3 grid;
  int threads;
  size_t lds;
  PassArgs args;
  const PassArgs* dargs;  // device copy of args
  const char* name;
  double bytes;  // algorithmic HBM bytes of one launch
};

// ----------------------------------------------------------------------------
// the solver
// ----------------------------------------------------------------------------
struct Solver {
  int N[3] = {1, 1, 1};
  int C = 1, nrhs = 1, nw = 1, interval = 8;
  long long nvox = 0;
  float epsilon = 1.f;
  float2 k02e{0.f, 0.f};
  // complex128 solver (Simulation.m:221-252 with single_precision = false):
  // every field, table and twiddle buffer below holds double2 / double
  // elements (the pointers keep their complex64 types; cz() is the element
  // size), and epsilon and k02e are kept in double
  bool f64 = false;
  double epsilon_d = 1.0;
  double2 k02e_d{0.0, 0.0};
  int roi_lo[4], roi_hi[4];
  std::vector<int> axes;  // non-trivial axes in [y x z] order
  int dev = 0;
  hipStream_t stream = nullptr;
  float2 *gamma = nullptr, *E = nullptr, *dE = nullptr, *T = nullptr, *S = nullptr;
  float2* probe = nullptr;  // scratch for wsim_solver_propagate
  std::vector<char> has_source;
  double* partials = nullptr;
  int nblk = 1;
  DevState* dst = nullptr;
  DevState* dprobe = nullptr;
  int *d_stopped = nullptr, *d_conv = nullptr, *d_niter = nullptr;
  double *d_last = nullptr, *d_diff = nullptr, *d_eall = nullptr;
  double* d_ref = nullptr;              // [nrhs] stop-test denominators
  float2* W0 = nullptr;                 // [nrhs][C][nvox] gamma * E0 of warm starts
  std::vector<char> has_warm;           // per rhs
  bool warm_mode = false;
  int *p_stopped = nullptr;  // probe state arrays
  double* p_dummy = nullptr;
  int hist_cap = 0;
  float *d_pe[3] = {nullptr, nullptr, nullptr};
  float2 *d_ramp[3] = {nullptr, nullptr, nullptr};
  AxisTables tab[3];
  float2 *d_twf[3] = {nullptr, nullptr, nullptr}, *d_twi[3] = {nullptr, nullptr, nullptr};
  std::vector<int> plan_rad[3];    // per-axis radix plan (empty: default)
  int plan_ncol[3] = {0, 0, 0};    // per-axis workgroup columns cap (0: default)
  int plan_threads[3] = {0, 0, 0}; // per-axis workgroup size, a multiple of 64 up to 512 (0: default)
  // per-axis: each XCD takes one contiguous range of tiles (1), every 8th
  // tile (0), or -1: the former when a launch has at most kXcdAutoBlocks
  // workgroups per rhs (measured: 1152^2 x pass, 72 workgroups, 34 -> 27 us;
  // 320^3 passes, >= 3200 workgroups, 2-4% slower)
  int plan_xcd[3] = {-1, -1, -1};
  std::vector<Launch> iter;  // one MBS iteration
  Launch pre{};                // first forward pass of iteration 1 when fused (fn == nullptr: none)
  PassArgs* d_pre = nullptr;
  PassArgs* d_args = nullptr;  // device copies of iter[k].args, then those of the other iterations of a cycle
  int part_cap = 0;            // energy partials allocated per iteration slot (of nw slots)
  std::vector<Launch> probe_iter;
  PassArgs* d_probe_args = nullptr;
  size_t lds_limit = 64 * 1024;    // per-workgroup LDS the tile search aims at
  size_t lds_device = 64 * 1024;   // the device's per-workgroup maximum (long axes)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_iters = 0;
  DevState host_state{};
  int* h_stopped = nullptr;  // pinned host flag the control step sets when every rhs has stopped
  hipEvent_t batch_ev[2] = {nullptr, nullptr};  // ends of the two batches in flight
  long long last_iters = 0;
  // Allocation-free mode (wsim_run_algorithm_noalloc): every device buffer is
  // carved from a caller-provided region (arena), gamma and E' are the
  // caller's device arrays, the caller's stream is used and no graph is
  // captured, so a solve calls no hipMalloc, hipFree, hipStreamCreate or
  // hipGraphInstantiate.
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  // PSTD mode (PSTD.m:79-127): E and dE hold the two time levels, S the
  // unscaled source, coef = [c1 | c2 | c3], srcline the source lines of the
  // first axis; pe tables hold the unscaled p ranges, ramps are 1
  bool pstd = false;
  float* coef = nullptr;
  unsigned char* srcline = nullptr;
  double pstd_ipc = 1.0;
  bool external_stream = false;
  bool ext_gamma = false, ext_E = false;

  ~Solver() { release(); }

  void release() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    if (h_stopped) (void)hipHostFree(h_stopped);
    h_stopped = nullptr;
    for (hipEvent_t& e : batch_ev)
      if (e) (void)hipEventDestroy(e);
    void* ptrs[] = {dE, T, S, probe, partials, dst, dprobe, d_stopped, d_conv, d_niter,
                    d_last, d_diff, d_eall, p_stopped, p_dummy, d_args, d_probe_args, d_ref, W0, d_pre};
    for (void* p : ptrs) dfree(p);
    if (!ext_gamma) dfree(gamma);
    if (!ext_E) dfree(E);
    dfree(coef);
    dfree(srcline);
    for (int a = 0; a < 3; ++a) {
      dfree(d_pe[a]);
      dfree(d_ramp[a]);
      dfree(d_twf[a]);
      dfree(d_twi[a]);
    }
    if (stream && !external_stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  long long field_elems() const { return (long long)C * nvox; }
  // bytes of one complex / real element of the fields and tables
  size_t cz() const { return f64 ? sizeof(double2) : sizeof(float2); }
  size_t rz() const { return f64 ? sizeof(double) : sizeof(float); }
  // element k of a complex buffer
  float2* at(float2* p, size_t k) const { return reinterpret_cast<float2*>(reinterpret_cast<char*>(p) + k * cz()); }
  int family_base() const { return f64 ? kFamilyF64 : 0; }

  template <class T_>
  void dalloc(T_** p, size_t n) {
    dalloc_bytes(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T_));
  }
  // complex and real buffers of n elements of the solver's precision
  void dalloc_c(float2** p, size_t n) { dalloc_bytes(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * cz()); }
  void dalloc_r(float** p, size_t n) { dalloc_bytes(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * rz()); }
  void dalloc_bytes(void** p, size_t bytes) {
    if (arena) {
      const size_t b = (bytes + 255) & ~(size_t)255;
      require(arena_used + b <= arena_cap,
              "problem needs more than the " + std::to_string(arena_cap >> 20) +
                  " MiB scratch of the allocation-free entry point",
              WSIM_ERR_UNSUPPORTED);
      *p = arena + arena_used;
      arena_used += b;
      return;
    }
    HIPCHK(hipMalloc(p, bytes));
  }
  // frees what dalloc allocated (arena memory is the caller's: nothing to do)
  void dfree(void* p) {
    if (p && !arena) (void)hipFree(p);
  }

  // pass configuration: columns per workgroup and LDS size
  int threads_of(int axis, bool contig) const {
    return plan_threads[axis] > 0 ? plan_threads[axis] : default_threads(contig ? 1 : 0);
  }

  void configure(PassArgs& a, bool contig, int mode, int axis, int& blocks, size_t& lds, bool fuse = false) {
    const AxisPlan& pl = tab[axis].plan;
    const int nthr = threads_of(axis, contig);
    const int n = N[axis];
    a.ntwf = (int)tab[axis].twf.size();
    a.ntwi = (int)tab[axis].twi.size();
    const int ntw = ((mode == MODE_INV && !fuse) ? 0 : a.ntwf) + (mode == MODE_FWD ? 0 : a.ntwi);
    a.axis = axis;
    a.n = n;
    a.inner = 1;
    for (int b = 0; b < axis; ++b) a.inner *= N[b];
    a.rstride = contig ? 1 : a.inner;
    a.plan = pl;
    size_t lds_cap = lds_limit;
    auto fits = [&](int ncol) {
      if (!tile_fits(pl, n, ncol, C, family_base(), nthr)) return false;
      size_t bytes;
      if (contig)
        bytes = (size_t)C * ncol * contig_pitch(n) * cz();
      else
        bytes = (size_t)n * (C * ncol + 1) * cz();
      bytes += lds_header(ncol, n, ntw, fuse, f64);
      return bytes <= lds_cap;
    };
    // long axes whose single-column tile exceeds the target take up to the
    // device's per-workgroup LDS (one workgroup per CU)
    if (!fits(1)) lds_cap = lds_device;
    int ncol = 0;
    const char* env_c = std::getenv(contig ? "WSIM_NCOL_CONTIG" : "WSIM_NCOL_STRIDED");
    int cap = env_c ? std::max(1, std::atoi(env_c)) : (contig ? 16 : 32);
    if (plan_ncol[axis] > 0) cap = plan_ncol[axis];
    if (contig) {
      const long long nlines = nvox / n;
      for (int c = cap; c >= 1; --c)
        if (fits(c)) {
          ncol = c;
          break;
        }
      require(ncol > 0, "transform length " + std::to_string(n) + " does not fit one workgroup",
              WSIM_ERR_UNSUPPORTED);
      ncol = (int)std::min<long long>(ncol, nlines);
      a.ncol = ncol;
      a.lds_pitch = contig_pitch(n);
      a.nlines = (int)nlines;
      blocks = (int)((nlines + ncol - 1) / ncol);
      a.ntile_inner = 1;
      a.ftile = make_fastdiv(1);
    } else {
      int c0 = 1;
      while (c0 * 2 <= cap) c0 *= 2;
      for (int c = c0; c >= 1; c >>= 1)
        if (fits(c)) {
          ncol = c;
          break;
        }
      require(ncol > 0, "transform length " + std::to_string(n) + " does not fit one workgroup",
              WSIM_ERR_UNSUPPORTED);
      while (ncol > 1 && ncol / 2 >= a.inner) ncol >>= 1;
      a.ncol = ncol;
      a.lds_pitch = C * ncol + 1;
      for (int q = 0; q < a.plan.nst; ++q)
        for (Stage* st : {&a.plan.f[q], &a.plan.i[q]}) {
          st->nbp = st->nb * a.lds_pitch;
          st->nsp = st->Ns * a.lds_pitch;
        }
      a.ntile_inner = (a.inner + ncol - 1) / ncol;
      a.ftile = make_fastdiv((uint32_t)a.ntile_inner);
      const long long outer = nvox / ((long long)a.inner * n);
      a.nlines = 0;
      blocks = (int)(a.ntile_inner * outer);
    }
    if (contig)
      lds = (size_t)C * a.ncol * contig_pitch(n) * cz();
    else
      lds = (size_t)n * (C * a.ncol + 1) * cz();
    lds += lds_header(a.ncol, n, ntw, fuse, f64);
  }

  PassArgs base_args() const {
    PassArgs a{};
    for (int d = 0; d < 3; ++d) a.N[d] = N[d];
    a.nvox = nvox;
    a.rhs_stride = field_elems();
    for (int d = 0; d < 3; ++d) {
      a.pe[d] = d_pe[d];
      a.ramp[d] = d_ramp[d];
    }
    a.k02e = k02e;
    const double kr = k02e.x, ki = k02e.y, kk = kr * kr + ki * ki;
    a.inv_k02e = make_float2((float)(kr / kk), (float)(-ki / kk));
    a.invN = (float)(1.0 / (double)nvox);
    {
      const double dr = k02e_d.x, di = k02e_d.y, dd = dr * dr + di * di;
      a.k02e_d = k02e_d;
      a.inv_k02e_d = make_double2(dr / dd, -di / dd);
      a.invN_d = 1.0 / (double)nvox;
    }
    for (int d = 0; d < 4; ++d) {
      a.roi_lo[d] = roi_lo[d];
      a.roi_hi[d] = roi_hi[d];
    }
    a.st = dst;
    return a;
  }

  Launch make_launch(bool contig, int mode, int pro, int epi, int axis, const char* name) {
    Launch L{};
    L.args = base_args();
    L.args.twf = d_twf[axis];
    L.args.twi = d_twi[axis];
    int blocks = 0;
    configure(L.args, contig, mode, axis, blocks, L.lds, epi == EPI_MIX_FWD || epi == EPI_PSTD_FWD);
    // the compact register family whenever the tile fits it (WSIM_COMPACT=0 disables)
    const char* env_k = std::getenv("WSIM_COMPACT");
    const bool compact = !f64 && !(env_k && std::atoi(env_k) == 0) &&
                         tile_fits(tab[axis].plan, N[axis], L.args.ncol, C, 1, threads_of(axis, contig));
    L.fn = select_kernel(contig, mode, pro, epi, C, f64 ? kFamilyF64 : (compact ? 1 : 0));
    L.fn_generic = L.fn;
    L.fn = fixed_kernel(L.fn, contig, mode, pro, epi, axis, L.args.ncol, f64 ? kFamilyF64 : (compact ? 1 : 0));
    // the composite radices are built into plan-specialised kernels only
    for (int q = 0; q < tab[axis].plan.nst; ++q)
      require(radix_generic(tab[axis].plan.f[q].R) || L.fn != L.fn_generic,
              "radix " + std::to_string(tab[axis].plan.f[q].R) +
                  " runs in plan-specialised kernels only; this launch shape has none",
              WSIM_ERR_UNSUPPORTED);
    // a strided tile keeps each thread on one column: the workgroup size must
    // be a multiple of the tile width
    require(contig || threads_of(axis, contig) % L.args.ncol == 0,
            "workgroup size " + std::to_string(threads_of(axis, contig)) + " is not a multiple of the tile width " +
                std::to_string(L.args.ncol),
            WSIM_ERR_UNSUPPORTED);
    L.grid = dim3((unsigned)blocks, (unsigned)nrhs, 1);
    L.threads = threads_of(axis, contig);
    L.args.blocks = blocks;
    L.args.xcd = plan_xcd[axis] >= 0 ? plan_xcd[axis] : (blocks <= kXcdAutoBlocks ? 1 : 0);
    L.args.nrhs = nrhs;
    L.args.grp_n = 1;
    L.args.grp_end = nullptr;
    // several rhs: the mix passes (the ones that read gamma) run a tile's rhs
    // back to back on one XCD (WSIM_RHS_INNER=0: rhs along y)
    const char* env_ri = std::getenv("WSIM_RHS_INNER");
    if (nrhs > 1 && (epi == EPI_MIX || epi == EPI_MIX_FWD) && !(env_ri && std::atoi(env_ri) == 0)) {
      L.args.rhs_inner = 1;
      L.grid = dim3((unsigned)(8ll * ((blocks + 7) / 8) * nrhs), 1, 1);
      require(L.grid.x < (1u << 31), "too many workgroups for one launch");
    }
    L.name = name;
    // algorithmic bytes per voxel, averaged over an even and an odd
    // iteration (E' is read and written on even iterations only): the mix
    // passes read T, gamma, dE' (+ E') and write dE' (+ E') (+ the next T)
    // (the PSTD update reads T, E, E_prev and c1..c3 and writes E_next (+ the
    // next T))
    const double perv = epi == EPI_MIX ? 32.0 * C + 8.0
                        : epi == EPI_MIX_FWD ? 40.0 * C + 8.0
                        : epi == EPI_PSTD ? 32.0 * C + 12.0
                        : epi == EPI_PSTD_FWD ? 40.0 * C + 12.0
                                              : 16.0 * C;
    L.bytes = perv * (double)nvox * nrhs * (f64 ? 2.0 : 1.0);
    if (L.lds > 48 * 1024)
      HIPCHK(hipFuncSetAttribute((const void*)L.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)L.lds));
    return L;
  }

  // the plan-specialised kernel of a launch of exactly this shape, if one
  // was compiled (shipped tuned plans), else `generic`; WSIM_FIXED=0 keeps
  // the run-time-plan kernels
  KernelFn fixed_kernel(KernelFn generic, bool contig, int mode, int pro, int epi, int axis, int ncol,
                        int family) const {
    const char* env = std::getenv("WSIM_FIXED");
    if (env && std::atoi(env) == 0) return generic;
    const AxisPlan& pl = tab[axis].plan;
    const int thr = threads_of(axis, contig);
    for (const FixedKernel* fk = fixed_kernel_table(); fk->get; ++fk) {
      if (fk->contig != (contig ? 1 : 0) || fk->mode != mode || fk->pro != pro || fk->epi != epi || fk->C != C ||
          fk->family != family || fk->n != N[axis] || fk->ncol != ncol || fk->threads != thr || fk->nst != pl.nst)
        continue;
      bool same = true;
      for (int q = 0; q < pl.nst; ++q) same = same && fk->rad[q] == pl.f[q].R;
      if (same) return fk->get();
    }
    return generic;
  }

  // launches of one iteration that run plan-specialised kernels
  int fixed_launches() const {
    int k = 0;
    for (const Launch& L : iter) k += L.fn_generic != nullptr && L.fn != L.fn_generic;
    return k;
  }

  // the last inverse pass also runs the next iteration's first forward pass
  // (same axis, same lines) when that axis has at least two FFT stages;
  // WSIM_FUSE=0 keeps them apart
  bool fuse_first_pass() const {
    const char* e = std::getenv("WSIM_FUSE");
    if (e && std::atoi(e) == 0) return false;
    return !pstd && axes.size() >= 2 && tab[axes[0]].plan.nst >= 2;
  }

  // the launches of one iteration (in = dE', out = T, T in place, T -> E', dE')
  // PSTD step (PSTD.m:108-111): the forward passes on the current time
  // level, the Laplacian in the middle pass, the inverse passes, and the
  // update in the last pass
  void build_pstd_step(std::vector<Launch>& out) {
    const int D = (int)axes.size();
    // the update also runs the next step's first forward pass (same axis,
    // same lines; that axis with at least two FFT stages), as the MBS
    // iteration does, and the first pass runs once, before step 1 (1152^2:
    // 0.163 -> 0.128 s for 2006 steps; 320^3 on its shipped plan 0.701 ->
    // 0.632 s for 928 steps, though 0.746 -> 0.773 s on the default plan,
    // whose fused update holds more registers).  WSIM_FUSE=0 keeps them apart.
    const char* env_f = std::getenv("WSIM_FUSE");
    const bool fuse = D >= 2 && tab[axes[0]].plan.nst >= 2 && !(env_f && std::atoi(env_f) == 0);
    if (D == 1) {
      out.push_back(make_launch(true, MODE_MID, PRO_SRC_RAMP, EPI_PSTD, axes[0], "pstd_step_1d"));
    } else {
      Launch L0 = make_launch(true, MODE_FWD, PRO_SRC_RAMP, EPI_NONE, axes[0], "pstd_fwd");
      L0.args.out = T;
      out.push_back(L0);
      for (int k = 1; k < D - 1; ++k) {
        Launch L = make_launch(false, MODE_FWD, PRO_NONE, EPI_NONE, axes[k], "pass_fwd");
        L.args.in = T;
        L.args.out = T;
        out.push_back(L);
      }
      Launch LM = make_launch(false, MODE_MID, PRO_NONE, EPI_NONE, axes[D - 1], "pstd_mid_laplacian");
      LM.args.in = T;
      LM.args.out = T;
      out.push_back(LM);
      for (int k = D - 2; k >= 1; --k) {
        Launch L = make_launch(false, MODE_INV, PRO_NONE, EPI_NONE, axes[k], "pass_inv");
        L.args.in = T;
        L.args.out = T;
        out.push_back(L);
      }
      Launch LE = make_launch(true, MODE_INV, PRO_NONE, fuse ? EPI_PSTD_FWD : EPI_PSTD, axes[0],
                              fuse ? "pstd_inv_update_fwd" : "pstd_inv_update");
      LE.args.in = T;
      LE.args.out = T;
      out.push_back(LE);
    }
    for (auto& L : out) {
      L.args.pstd = 1;
      L.args.fld[0] = E;
      L.args.fld[1] = dE;
      L.args.coef = coef;
      L.args.Sp = S;
      L.args.srcline = srcline;
      L.args.partials = partials;
    }
    if (fuse) {
      pre = out.front();
      out.erase(out.begin());
    }
  }

  void build_iteration(std::vector<Launch>& out, bool probe_mode) {
    out.clear();
    if (!probe_mode) pre = Launch{};
    const int D = (int)axes.size();
    const int epi = probe_mode ? EPI_RAMP : EPI_MIX;
    const float2* in0 = probe_mode ? probe : dE;
    if (pstd && !probe_mode) {
      build_pstd_step(out);
    } else if (D == 1) {
      Launch L = make_launch(true, MODE_MID, PRO_SRC_RAMP, epi, axes[0], "pass_mid_pro_epi");
      L.args.in = in0;
      L.args.S = probe_mode ? nullptr : S;
      L.args.W0 = probe_mode ? nullptr : W0;
      L.args.out = probe_mode ? probe : T;
      L.args.dE = dE;
      L.args.E = E;
      L.args.gamma = gamma;
      L.args.partials = partials;
      out.push_back(L);
    } else {
      Launch L0 = make_launch(true, MODE_FWD, PRO_SRC_RAMP, EPI_NONE, axes[0], "pass_fwd_pro");
      L0.args.in = in0;
      L0.args.S = probe_mode ? nullptr : S;
      L0.args.W0 = probe_mode ? nullptr : W0;
      L0.args.out = T;
      out.push_back(L0);
      for (int k = 1; k < D - 1; ++k) {
        Launch L = make_launch(false, MODE_FWD, PRO_NONE, EPI_NONE, axes[k], "pass_fwd");
        L.args.in = T;
        L.args.out = T;
        out.push_back(L);
      }
      Launch LM = make_launch(false, MODE_MID, PRO_NONE, EPI_NONE, axes[D - 1], "pass_mid_green");
      LM.args.in = T;
      LM.args.out = T;
      out.push_back(LM);
      for (int k = D - 2; k >= 1; --k) {
        Launch L = make_launch(false, MODE_INV, PRO_NONE, EPI_NONE, axes[k], "pass_inv");
        L.args.in = T;
        L.args.out = T;
        out.push_back(L);
      }
      const bool fuse = !probe_mode && fuse_first_pass();
      Launch LE = make_launch(true, MODE_INV, PRO_NONE, fuse ? EPI_MIX_FWD : epi, axes[0],
                              fuse ? "pass_inv_mix_fwd" : "pass_inv_mix");
      LE.args.in = T;
      LE.args.out = probe_mode ? probe : T;
      LE.args.dE = dE;
      LE.args.E = E;
      LE.args.gamma = gamma;
      LE.args.partials = partials;
      LE.args.W0 = probe_mode ? nullptr : W0;
      LE.args.S = fuse ? S : nullptr;
      out.push_back(LE);
      if (fuse) {
        // the first forward pass runs once per solve (before iteration 1);
        // every later one is the tail of the previous iteration's last pass
        pre = out.front();
        out.erase(out.begin());
      }
    }
    // tile order: the last two passes of the iteration walk memory from its
    // end, the others from its start.  Measured at 320^3 (fields larger than
    // the 256 MB memory-side cache) over all 16 orders of the 4 passes: every
    // order with the last (mix) pass backwards ran it 5-7 us faster, this one
    // 0.2547 s against 0.2566 s for alternating directions from the first
    // pass; grids that fit the cache showed no difference.
    // WSIM_REVERSE=<mask> overrides (bit k: pass k backwards).
    const char* rv = std::getenv("WSIM_REVERSE");
    for (size_t k = 0; k < out.size(); ++k)
      out[k].args.reverse = rv ? (std::atoi(rv) >> k) & 1 : (k + 2 >= out.size()) ? 1 : 0;
    if (probe_mode)
      for (auto& L : out) {
        L.args.st = dprobe;
        L.grid.y = 1;
        if (L.args.rhs_inner) {  // the probe runs one rhs
          L.args.rhs_inner = 0;
          L.grid.x = (unsigned)L.args.blocks;
        }
      }
  }

  void launch(const Launch& L) {
    hipLaunchKernelGGL(L.fn, L.grid, dim3(L.threads), L.lds, stream, L.dargs);
    HIPCHK(hipGetLastError());
  }

  // per-wiggle coordinate and ramp tables (WaveSimBase.m:300-316):
  // wsim_wiggle (complex64 solvers) or wsim_wiggle_f64 (complex128)
  void upload_wiggles(const void* wig) {
    if (f64)
      upload_wiggles_t<double, double2>(static_cast<const wsim_wiggle_f64*>(wig));
    else
      upload_wiggles_t<float, float2>(static_cast<const wsim_wiggle*>(wig));
  }
  template <class R, class V, class W>
  void upload_wiggles_t(const W* wig) {
    for (int a = 0; a < 3; ++a) {
      std::vector<R> pe((size_t)nw * N[a]);
      std::vector<V> rp((size_t)nw * N[a]);
      for (int w = 0; w < nw; ++w) {
        const auto* src = a == 0 ? wig[w].pye : (a == 1 ? wig[w].pxe : wig[w].pze);
        const auto* gr = a == 0 ? wig[w].gy : (a == 1 ? wig[w].gx : wig[w].gz);
        require(src != nullptr && gr != nullptr, "null wiggle table");
        for (int i = 0; i < N[a]; ++i) {
          pe[(size_t)w * N[a] + i] = src[i];
          rp[(size_t)w * N[a] + i] = mkv_host<V>(gr[i].re, gr[i].im);
        }
      }
      HIPCHK(hipMemcpy(d_pe[a], pe.data(), pe.size() * sizeof(R), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(d_ramp[a], rp.data(), rp.size() * sizeof(V), hipMemcpyHostToDevice));
    }
  }

  // a new medium on the same grid: gamma, epsilon, k02e and wiggle tables
  // (sources are scaled by i/eps/Nw when set: set them again afterwards)
  void set_medium(const wsim_cfloat* g, float eps, wsim_cfloat k, const wsim_wiggle* wig) {
    require(g != nullptr && wig != nullptr, "null argument");
    require(!f64, "a complex128 solver takes no new medium: create another one", WSIM_ERR_UNSUPPORTED);
    upload_wiggles(wig);
    HIPCHK(hipMemcpy(gamma, g, nvox * sizeof(float2), hipMemcpyHostToDevice));
    epsilon = eps;
    k02e = make_float2(k.re, k.im);
    rebuild_launches();
    std::fill(has_source.begin(), has_source.end(), 0);
    const size_t fe = (size_t)field_elems() * nrhs;
    HIPCHK(hipMemset(S, 0, fe * sizeof(float2)));
    HIPCHK(hipDeviceSynchronize());  // uploads on the null stream are done before the solver stream runs
  }

  // twiddle tables of axis a for its current radix plan
  void build_tables(int a) {
    std::vector<int> rad = plan_rad[a].empty() ? default_radices(N[a], C) : plan_rad[a];
    if (N[a] == 1) rad = {};
    tab[a] = build_axis(N[a], rad);
    dfree(d_twf[a]);
    dfree(d_twi[a]);
    d_twf[a] = d_twi[a] = nullptr;
    dalloc_c(&d_twf[a], tab[a].twf.size());
    dalloc_c(&d_twi[a], tab[a].twi.size());
    const void* tf = f64 ? (const void*)tab[a].twf_d.data() : (const void*)tab[a].twf.data();
    const void* ti = f64 ? (const void*)tab[a].twi_d.data() : (const void*)tab[a].twi.data();
    HIPCHK(hipMemcpy(d_twf[a], tf, tab[a].twf.size() * cz(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_twi[a], ti, tab[a].twi.size() * cz(), hipMemcpyHostToDevice));
  }

  // (re)build the launches of one iteration; partials are sized for the
  // energy pass; cached probe launches and the captured graph are dropped
  void rebuild_launches() {
    std::vector<Launch> it2;
    build_iteration(it2, false);
    const int nb = it2.back().args.blocks;  // the mix pass's tiles (its grid also spans the rhs when rhs_inner)
    if (!partials || nb > part_cap) {
      dfree(partials);
      partials = nullptr;
      dalloc(&partials, (size_t)nb * nrhs * nw);  // one slot per iteration of a wiggle cycle
      part_cap = nb;
    }
    nblk = nb;
    for (auto& L : it2) L.args.partials = partials;
    dfree(d_args);
    d_args = nullptr;
    iter.swap(it2);
    upload_cycle_args();
    dfree(d_pre);
    d_pre = nullptr;
    if (pre.fn) {
      std::vector<Launch> one{pre};
      upload_args(one, &d_pre);
      pre.dargs = d_pre;
    }
    probe_iter.clear();
    dfree(d_probe_args);
    d_probe_args = nullptr;
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    graph_iters = 0;
  }

  // mean time (ms) of one steady-state iteration (every rhs forced to run)
  // ms per iteration: the fastest of three timed runs of `reps` iterations
  // (the minimum filters clock and scheduling noise out of the tuner's choices)
  double time_iteration(int reps) {
    ensure_history(3 * reps + nw + 2);
    reset_state(1ll << 40, -1.0, true);
    enqueue_iterations(nw);
    hipEvent_t ev[4];
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(ev[0], stream));
    for (int k = 1; k < 4; ++k) {
      enqueue_iterations(reps);
      HIPCHK(hipEventRecord(ev[k], stream));
    }
    HIPCHK(hipEventSynchronize(ev[3]));
    double best = 1e30;
    for (int k = 1; k < 4; ++k) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, ev[k - 1], ev[k]));
      best = std::min(best, (double)ms / reps);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return best;
  }

  // plan text: "y:8,8,5/16/512;x:4,16,5/16;z:4,16,5/32/0/xcd" (axis:radices/columns
  // [/threads[/xcd|noxcd]]; columns 0 = widest tile that fits, threads 0 or omitted =
  // the default for the axis; xcd = each XCD takes one contiguous range of the
  // tiles, noxcd = every 8th tile, omitted = by the launch's workgroup count)
  std::string plan_string() const {
    static const char* nm = "yxz";
    std::string out;
    for (int a = 0; a < 3; ++a) {
      if (N[a] == 1) continue;
      if (!out.empty()) out += ";";
      out += nm[a];
      out += ":";
      const AxisPlan& p = tab[a].plan;
      for (int s2 = 0; s2 < p.nst; ++s2) out += (s2 ? "," : "") + std::to_string(p.f[s2].R);
      out += "/" + std::to_string(plan_ncol[a]);
      if (plan_threads[a] > 0 || plan_xcd[a] >= 0) out += "/" + std::to_string(plan_threads[a]);
      if (plan_xcd[a] >= 0) out += plan_xcd[a] ? "/xcd" : "/noxcd";
    }
    return out;
  }

  void reset_plan() {
    for (int a = 0; a < 3; ++a) {
      plan_rad[a].clear();
      plan_ncol[a] = 0;
      plan_threads[a] = 0;
      plan_xcd[a] = -1;
      build_tables(a);
    }
    rebuild_launches();
  }

  // A plan the launches cannot be built for (a tile that does not fit, a
  // radix without a kernel) leaves the previous plan in force: its tables and
  // launches are rebuilt before the error is passed on, so no launch keeps a
  // pointer to a freed table.
  void set_plan(const std::string& txt) {
    std::vector<int> old_rad[3];
    int old_ncol[3], old_thr[3], old_xcd[3];
    for (int a = 0; a < 3; ++a) {
      old_rad[a] = plan_rad[a];
      old_ncol[a] = plan_ncol[a];
      old_thr[a] = plan_threads[a];
      old_xcd[a] = plan_xcd[a];
    }
    try {
      apply_plan(txt);
    } catch (...) {
      for (int a = 0; a < 3; ++a) {
        plan_rad[a] = old_rad[a];
        plan_ncol[a] = old_ncol[a];
        plan_threads[a] = old_thr[a];
        plan_xcd[a] = old_xcd[a];
        build_tables(a);
      }
      rebuild_launches();
      throw;
    }
  }

  void apply_plan(const std::string& txt) {
    size_t pos = 0;
    while (pos < txt.size()) {
      size_t end = txt.find(';', pos);
      if (end == std::string::npos) end = txt.size();
      const std::string item = txt.substr(pos, end - pos);
      pos = end + 1;
      if (item.size() < 3 || item[1] != ':') throw Error(WSIM_ERR_ARG, "bad plan item '" + item + "'");
      const int a = item[0] == 'y' ? 0 : item[0] == 'x' ? 1 : item[0] == 'z' ? 2 : -1;
      require(a >= 0, "bad plan axis in '" + item + "'");
      const size_t slash = item.find('/');
      const std::string rs = item.substr(2, slash == std::string::npos ? std::string::npos : slash - 2);
      std::vector<int> rad;
      long long prod = 1;
      for (size_t q = 0; q < rs.size();) {
        const int x = std::atoi(rs.c_str() + q);
        require(radix_generic(x) || radix_specialised(x), "unsupported radix in plan '" + item + "'");
        require(radix_specialised(x) ? C == 1 : radix_fits(x, C),
                "radix " + std::to_string(x) + " is not built for " + std::to_string(C) +
                    "-component fields in plan '" + item + "'");
        rad.push_back(x);
        prod *= x;
        while (q < rs.size() && rs[q] != ',') ++q;
        ++q;
      }
      require(prod == N[a], "plan radices do not multiply to the axis length in '" + item + "'");
      plan_rad[a] = rad;
      plan_ncol[a] = slash == std::string::npos ? 0 : std::max(0, std::atoi(item.c_str() + slash + 1));
      const size_t slash2 = slash == std::string::npos ? std::string::npos : item.find('/', slash + 1);
      const int thr = slash2 == std::string::npos ? 0 : std::atoi(item.c_str() + slash2 + 1);
      require(thr == 0 || (thr >= 64 && thr <= kMaxThreads && thr % 64 == 0),
              "plan workgroup size must be a multiple of 64 up to " + std::to_string(kMaxThreads) + " in '" + item + "'");
      plan_threads[a] = thr;
      const size_t slash3 = slash2 == std::string::npos ? std::string::npos : item.find('/', slash2 + 1);
      const std::string opt = slash3 == std::string::npos ? "" : item.substr(slash3 + 1);
      require(slash3 == std::string::npos || opt == "xcd" || opt == "noxcd",
              "unknown plan option '" + opt + "' in '" + item + "'");
      plan_xcd[a] = slash3 == std::string::npos ? -1 : opt == "xcd" ? 1 : 0;
      build_tables(a);
    }
    rebuild_launches();
  }

  // every factorisation of n into supported radices with at most `extra`
  // stages more than the fewest, in every distinct order (capped)
  static std::vector<std::vector<int>> radix_candidates(int n, int extra, size_t cap, int C) {
    static const int R[] = {16, 8, 4, 2, 11, 7, 5, 3};
    std::vector<std::vector<int>> sets;
    std::vector<int> cur;
    std::function<void(int, int)> rec = [&](int rem, int start) {
      if (rem == 1) {
        sets.push_back(cur);
        return;
      }
      if (cur.size() >= (size_t)MAXST) return;
      for (int i = start; i < 8; ++i)
        if (rem % R[i] == 0 && radix_fits(R[i], C)) {
          cur.push_back(R[i]);
          rec(rem / R[i], i);
          cur.pop_back();
        }
    };
    rec(n, 0);
    size_t best = 99;
    for (auto& v : sets) best = std::min(best, v.size());
    std::vector<std::vector<int>> out;
    for (auto& v : sets) {
      if (v.size() > best + extra) continue;
      std::sort(v.begin(), v.end());
      size_t k = 0;
      do {
        out.push_back(v);
      } while (++k < 24 && std::next_permutation(v.begin(), v.end()));
    }
    if (out.size() > cap) out.resize(cap);
    return out;
  }

  // coordinate search over (radix order, workgroup columns, workgroup size,
  // XCD tile ranges) per axis, timing whole iterations; keeps the fastest plan
  std::string tune(int reps) {
    double best = time_iteration(reps);
    for (int a : axes) {
      if (N[a] == 1) continue;
      const bool contig = (a == axes[0]);
      auto try_cfg = [&](const std::vector<int>& rad, int ncol, int thr, int xcd) {
        const std::vector<int> old_r = plan_rad[a];
        const int old_c = plan_ncol[a], old_t = plan_threads[a], old_x = plan_xcd[a];
        plan_rad[a] = rad;
        plan_ncol[a] = ncol;
        plan_threads[a] = thr;
        plan_xcd[a] = xcd;
        double t = 1e30;
        try {
          build_tables(a);
          rebuild_launches();
          t = time_iteration(reps);
        } catch (const Error&) {
          t = 1e30;  // the plan does not fit a workgroup
        }
        if (t < best * 0.997) {  // a plan change must beat the noise
          best = t;
        } else {
          plan_rad[a] = old_r;
          plan_ncol[a] = old_c;
          plan_threads[a] = old_t;
          plan_xcd[a] = old_x;
        }
      };
      auto radices = [&] { return plan_rad[a].empty() ? default_radices(N[a], C) : plan_rad[a]; };
      // two rounds: the register family a tile can use depends on the radix
      // order, the tile width and the workgroup size together
      // (narrow tiles give small grids enough workgroups to fill the CUs)
      const std::vector<int> cand = contig ? std::vector<int>{1, 2, 4, 6, 8, 12, 16, 24, 32, 48}
                                           : std::vector<int>{1, 2, 4, 8, 16, 32};
      for (int round = 0; round < 2; ++round) {
        for (const auto& rad : radix_candidates(N[a], 1, 48, C))
          try_cfg(rad, plan_ncol[a], plan_threads[a], plan_xcd[a]);
        for (int thr : {256, 512})
          for (int c : cand) try_cfg(radices(), c, thr, plan_xcd[a]);
        for (int x : {0, 1}) try_cfg(radices(), plan_ncol[a], plan_threads[a], x);
      }
      build_tables(a);
    }
    rebuild_launches();
    return plan_string();
  }

  // the launches of the nw iterations of a wiggle cycle: iteration j's copy of
  // every pass carries it_off = j and energy-partials slot j; iter[k].dargs is
  // iteration 0's (per-iteration control, groups, profiling)
  void upload_cycle_args() {
    const size_t n = iter.size();
    std::vector<PassArgs> h(n * nw);
    for (int j = 0; j < nw; ++j)
      for (size_t k = 0; k < n; ++k) {
        PassArgs a = iter[k].args;
        a.it_off = j;
        if (a.partials) a.partials = partials + (size_t)j * nrhs * nblk;
        h[j * n + k] = a;
      }
    dalloc(&d_args, h.size());
    HIPCHK(hipMemcpy(d_args, h.data(), h.size() * sizeof(PassArgs), hipMemcpyHostToDevice));
    for (size_t k = 0; k < n; ++k) iter[k].dargs = d_args + k;
  }

  void upload_args(std::vector<Launch>& ls, PassArgs** dptr) {
    std::vector<PassArgs> h(ls.size());
    for (size_t k = 0; k < ls.size(); ++k) h[k] = ls[k].args;
    dalloc(dptr, h.size());
    HIPCHK(hipMemcpy(*dptr, h.data(), h.size() * sizeof(PassArgs), hipMemcpyHostToDevice));
    for (size_t k = 0; k < ls.size(); ++k) ls[k].dargs = *dptr + k;
  }

  void launch_control() {
    if (pstd)
      hipLaunchKernelGGL(control_pstd_kernel, dim3(1), dim3(kControlThreads), 0, stream, dst, partials, nblk);
    else
      hipLaunchKernelGGL(control_kernel, dim3(1), dim3(kControlThreads), 0, stream, dst, partials, nblk);
    HIPCHK(hipGetLastError());
  }

  // k iterations; whole wiggle cycles run their control steps once per cycle
  // (WSIM_CYCLE_CONTROL=0: every iteration)
  void enqueue_iterations(int k) {
    const char* e = std::getenv("WSIM_CYCLE_CONTROL");
    const bool cyc = !pstd && nw > 1 && k % nw == 0 && !(e && std::atoi(e) == 0);
    if (!cyc) {
      for (int i = 0; i < k; ++i) {
        for (const auto& L : iter) launch(L);
        launch_control();
      }
      return;
    }
    const size_t n = iter.size();
    for (int c = 0; c < k / nw; ++c) {
      for (int j = 0; j < nw; ++j)
        for (size_t q = 0; q < n; ++q) {
          const Launch& L = iter[q];
          hipLaunchKernelGGL(L.fn, L.grid, dim3(L.threads), L.lds, stream, (const PassArgs*)(d_args + j * n + q));
          HIPCHK(hipGetLastError());
        }
      hipLaunchKernelGGL(control_cycle_kernel, dim3(1), dim3(kControlThreads), 0, stream, dst, partials, nblk, nw,
                         (long long)nrhs * nblk);
      HIPCHK(hipGetLastError());
    }
  }

  // g: nvox complex values of the solver's precision (wsim_cfloat or
  // wsim_cdouble), wig: wsim_wiggle or wsim_wiggle_f64 tables
  void create(const wsim_solver_desc* d, const void* g, const void* wig) {
    require(d != nullptr && (g != nullptr || pstd) && wig != nullptr, "null argument");
    for (int a = 0; a < 3; ++a) {
      require(d->N[a] >= 1 && d->N[a] < (1 << 30), "grid size out of range");
      N[a] = (int)d->N[a];
    }
    require(d->N[3] == 1 || d->N[3] == 3, "N(4) must be 1 (scalar) or 3 (vector)");
    C = (int)d->N[3];
    require(d->nrhs >= 1 && d->nrhs <= 65535, "nrhs out of range");
    nrhs = d->nrhs;
    require(d->nwiggles >= 1 && d->nwiggles <= 8, "Nwiggles must be 1..8");
    nw = d->nwiggles;
    interval = std::max(1, d->energy_interval);
    epsilon = d->epsilon;
    k02e = make_float2(d->k02e.re, d->k02e.im);
    nvox = (long long)N[0] * N[1] * N[2];
    // the passes address a component's field through a 32-bit byte offset
    // from a per-workgroup scalar base (global_load/store with a 32-bit lane
    // offset): an element offset reaches nvox - 1 in the strided passes, so
    // one component's field must stay below 4 GiB
    const long long maxv = f64 ? kMaxVoxelsF64 : kMaxVoxels;
    require(nvox <= maxv, "grid of " + std::to_string(nvox) + " voxels: at most " + std::to_string(maxv) +
                              " (one component below 4 GiB)",
            WSIM_ERR_UNSUPPORTED);
    for (int c = 0; c < 4; ++c) {
      const long long lim = c < 3 ? N[c] : C;
      long long lo = d->roi[0][c], hi = d->roi[1][c];
      require(lo >= 1 && hi <= lim && lo <= hi, "roi outside the grid");
      roi_lo[c] = (int)(lo - 1);
      roi_hi[c] = (int)(hi - 1);
    }
    axes.clear();
    for (int a = 0; a < 3; ++a)
      if (N[a] > 1) axes.push_back(a);
    if (axes.empty()) axes.push_back(0);
    require(!(C == 3 && axes.size() == 1), "vector fields need at least two non-trivial axes",
            WSIM_ERR_UNSUPPORTED);
    HIPCHK(hipGetDevice(&dev));
    {
      int lim = 0;
      HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
      const char* env_l = std::getenv("WSIM_LDS_CAP_KB");
      lds_limit = std::min<size_t>((size_t)lim, env_l ? (size_t)std::atoi(env_l) * 1024 : LDS_CAP);
      lds_device = std::max<size_t>((size_t)lim, lds_limit);
    }
    if (!external_stream) HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // tables
    for (int a = 0; a < 3; ++a) {
      require(N[a] == 1 || !default_radices(N[a], C).empty(),
              "axis length " + std::to_string(N[a]) + " has a prime factor above 11",
              WSIM_ERR_UNSUPPORTED);
      build_tables(a);
      dalloc_r(&d_pe[a], (size_t)nw * N[a]);
      dalloc_c(&d_ramp[a], (size_t)nw * N[a]);
    }
    upload_wiggles(wig);
    // fields
    const size_t fe = (size_t)field_elems() * nrhs;
    if (ext_gamma) {
      gamma = reinterpret_cast<float2*>(const_cast<void*>(g));  // the caller's device array
    } else if (pstd) {
      gamma = nullptr;  // PSTD has no gamma
    } else {
      dalloc_c(&gamma, nvox);
      HIPCHK(hipMemcpy(gamma, g, nvox * cz(), hipMemcpyHostToDevice));
    }
    if (!ext_E) dalloc_c(&E, fe);
    dalloc_c(&dE, fe);
    dalloc_c(&T, fe);
    dalloc_c(&S, fe);
    HIPCHK(hipMemset(S, 0, fe * cz()));
    has_source.assign(nrhs, 0);
    if (pstd) {
      dalloc(&coef, 3 * (size_t)nvox);
      const size_t nl = (size_t)(nvox / N[axes[0]]);
      dalloc(&srcline, nl);
      HIPCHK(hipMemset(srcline, 0, nl));
    }
    // state
    dalloc(&dst, 1);
    dalloc(&dprobe, 1);
    dalloc(&d_stopped, nrhs);
    dalloc(&d_conv, nrhs);
    dalloc(&d_niter, nrhs);
    dalloc(&d_last, nrhs);
    dalloc(&d_ref, nrhs);
    // a result read before the first solve reports 0 iterations, not converged
    for (int* p : {d_stopped, d_conv, d_niter}) HIPCHK(hipMemset(p, 0, nrhs * sizeof(int)));
    {
      std::vector<double> z(nrhs, 0.0);
      HIPCHK(hipMemcpy(d_ref, z.data(), nrhs * sizeof(double), hipMemcpyHostToDevice));
    }
    has_warm.assign(nrhs, 0);
    dalloc(&p_stopped, 1);
    HIPCHK(hipMemset(p_stopped, 0, sizeof(int)));
    dalloc(&p_dummy, 1);
    rebuild_launches();
    // the shipped plan of this shape and precision; a complex128 solver
    // without one takes the complex64 plan, and a PSTD solver (one wiggle)
    // the MBS plan of its grid
    const ShippedPlan* shipped = nullptr;
    for (int want = f64 ? 1 : 0; want >= 0 && !shipped; --want)
      for (const ShippedPlan* sp = kShippedPlans; sp->plan; ++sp)
        if (sp->n0 == N[0] && sp->n1 == N[1] && sp->n2 == N[2] && sp->c == C && sp->nrhs == nrhs &&
            (sp->nw == nw || pstd) && sp->f64 == want) {
          shipped = sp;
          break;
        }
    if (shipped) {
      try {
        set_plan(shipped->plan);
      } catch (const Error&) {
        reset_plan();  // keep the default plan
      }
    }
    // the probe state: one rhs, never stopped
    DevState ps{};
    ps.it = nw + 1;
    ps.nw = nw;
    ps.nrhs = 1;
    ps.stopped = p_stopped;
    HIPCHK(hipMemcpy(dprobe, &ps, sizeof(ps), hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
  }

  void set_source(int rhs, int np, const wsim_source_patch* patches) {
    require(!f64, "this solver computes in complex128: set its sources with wsim_solver_set_source_f64");
    set_source_t<float2>(rhs, np, patches);
  }
  void set_source_f64(int rhs, int np, const wsim_source_patch_f64* patches) {
    require(f64, "this solver computes in complex64: set its sources with wsim_solver_set_source");
    set_source_t<double2>(rhs, np, patches);
  }

  // V = float2 (complex64 patches) or double2 (complex128 patches)
  template <class V, class Patch>
  void set_source_t(int rhs, int np, const Patch* patches) {
    using R = real_of<V>;
    require(rhs >= 0 && rhs < nrhs, "rhs out of range");
    require(np >= 0 && (np == 0 || patches != nullptr), "null patches");
    // source pre-multiplied by A = i/eps/Nw in the field's precision, the patches
    // added in order into a zero field (WaveSimBase.m:171-172; Source.m:163-169)
    const size_t fe = (size_t)field_elems();
    // (PSTD: unscaled, the amplitude of each step multiplies it in the update)
    const R eps = f64 ? (R)epsilon_d : (R)epsilon;
    const R ar = pstd ? R(1) : R(0), ai = pstd ? R(0) : (R(1) / eps) / (R)nw;  // (1i / eps) / Nw
    bool any = false;
    struct Entry {
      size_t g;
      V v;
    };
    std::vector<Entry> ent;
    size_t total = 0;
    for (int p = 0; p < np; ++p) {
      const Patch& sp = patches[p];
      long long pos[4], sz[4];
      for (int c = 0; c < 4; ++c) {
        pos[c] = (long long)std::llround(sp.pos[c]) - 1;
        sz[c] = sp.size[c];
        const long long lim = c < 3 ? N[c] : C;
        require(sz[c] >= 0 && pos[c] >= 0 && pos[c] + sz[c] <= lim, "source patch outside the grid");
      }
      require(sp.values != nullptr || sz[0] * sz[1] * sz[2] * sz[3] == 0, "null patch values");
      total += (size_t)(sz[0] * sz[1] * sz[2] * sz[3]);
    }
    // few voxels: upload only those (a point source at 320^3 would otherwise
    // move a 262 MB dense field through host memory and PCIe)
    const bool sparse = total * 4 < fe;
    std::vector<V> h(sparse ? 0 : fe, mkv_host<V>(R(0), R(0)));
    if (sparse) ent.reserve(total);
    for (int p = 0; p < np; ++p) {
      const Patch& sp = patches[p];
      long long pos[4], sz[4];
      for (int c = 0; c < 4; ++c) {
        pos[c] = (long long)std::llround(sp.pos[c]) - 1;
        sz[c] = sp.size[c];
      }
      size_t k = 0;
      for (long long q = 0; q < sz[3]; ++q)
        for (long long z = 0; z < sz[2]; ++z)
          for (long long x = 0; x < sz[1]; ++x)
            for (long long y = 0; y < sz[0]; ++y, ++k) {
              const auto v = sp.values[k];
              const size_t g = (size_t)(pos[3] + q) * nvox +
                               (size_t)((pos[2] + z) * N[1] + (pos[1] + x)) * N[0] + (size_t)(pos[0] + y);
              // A * val with A = (ar + i ai)
              const R re = ar * v.re - ai * v.im;
              const R im = ar * v.im + ai * v.re;
              if (sparse) {
                ent.push_back({g, mkv_host<V>(re, im)});
              } else {
                h[g].x += re;
                h[g].y += im;
              }
              if (v.re != 0 || v.im != 0) any = true;
            }
    }
    V* Sr = reinterpret_cast<V*>(at(S, (size_t)rhs * fe));
    if (!sparse) {
      HIPCHK(hipMemcpy(Sr, h.data(), fe * sizeof(V), hipMemcpyHostToDevice));
    } else {
      // overlapping patches add in patch order, from zero, as in the dense field
      std::stable_sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) { return a.g < b.g; });
      std::vector<unsigned long long> hi;
      std::vector<V> hv;
      for (size_t i = 0; i < ent.size();) {
        V acc = mkv_host<V>(R(0), R(0));
        const size_t g = ent[i].g;
        for (; i < ent.size() && ent[i].g == g; ++i) {
          acc.x += ent[i].v.x;
          acc.y += ent[i].v.y;
        }
        hi.push_back(g);
        hv.push_back(acc);
      }
      HIPCHK(hipMemset(Sr, 0, fe * sizeof(V)));
      if (!hi.empty()) {
        struct Tmp {  // device staging freed on every path out
          Solver* s;
          unsigned long long* i = nullptr;
          V* v = nullptr;
          ~Tmp() {
            s->dfree(i);
            s->dfree(v);
          }
        } t{this};
        dalloc(&t.i, hi.size());
        dalloc(&t.v, hv.size());
        HIPCHK(hipMemcpy(t.i, hi.data(), hi.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(t.v, hv.data(), hv.size() * sizeof(V), hipMemcpyHostToDevice));
        const long long m = (long long)hi.size();
        const unsigned blocks = (unsigned)std::min<long long>((m + 255) / 256, 4096);
        hipLaunchKernelGGL(scatter_kernel<V>, dim3(blocks), dim3(256), 0, 0, Sr, t.i, t.v, m);
        HIPCHK(hipGetLastError());
        HIPCHK(hipDeviceSynchronize());
      }
    }
    if (pstd) {
      // lines of the first axis that carry source entries (the update reads
      // S only there)
      const size_t nl = (size_t)(nvox / N[axes[0]]);
      std::vector<unsigned char> fl(nl, 0);
      for (int p = 0; p < np; ++p) {
        const Patch& sp = patches[p];
        long long pos[4], sz[4];
        for (int c = 0; c < 4; ++c) {
          pos[c] = (long long)std::llround(sp.pos[c]) - 1;
          sz[c] = sp.size[c];
        }
        for (long long z = 0; z < sz[2]; ++z)
          for (long long x = 0; x < sz[1]; ++x)
            for (long long y = 0; y < sz[0]; ++y) {
              const size_t g = (size_t)((pos[2] + z) * N[1] + (pos[1] + x)) * N[0] + (size_t)(pos[0] + y);
              fl[g / (size_t)N[axes[0]]] = 1;
            }
      }
      HIPCHK(hipMemcpy(srcline, fl.data(), nl, hipMemcpyHostToDevice));
    }
    HIPCHK(hipDeviceSynchronize());
    has_source[rhs] = any ? 1 : 0;
  }

  // warm start of one rhs: W0 = gamma * E0 (E0 unprimed, [pol][z][x][y])
  void set_warm(int rhs, const wsim_cfloat* E0, bool on_device) {
    require(rhs >= 0 && rhs < nrhs && E0 != nullptr, "bad argument");
    require(!f64, "warm starts are built for complex64 solvers", WSIM_ERR_UNSUPPORTED);
    const size_t f1 = (size_t)field_elems();
    if (!W0) {
      dalloc(&W0, f1 * nrhs);
      HIPCHK(hipMemset(W0, 0, f1 * nrhs * sizeof(float2)));
      rebuild_launches();  // the passes take the W0 pointer
    }
    float2* dst = W0 + rhs * f1;
    // E0 may come from any stream: wait for all device work, then copy on
    // the solver's stream (a device-to-device hipMemcpy would return before
    // the copy is done and is not ordered with that stream)
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpyAsync(dst, E0, f1 * sizeof(float2), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          stream));
    const unsigned blocks = (unsigned)std::min<size_t>((f1 + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(gamma_mul_kernel<float2>, dim3(blocks), dim3(256), 0, stream, dst, gamma, (long long)nvox, C);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    has_warm[rhs] = 1;
    warm_mode = true;
  }

  void clear_warm() {
    warm_mode = false;
    std::fill(has_warm.begin(), has_warm.end(), 0);
    if (W0) HIPCHK(hipMemset(W0, 0, (size_t)field_elems() * nrhs * sizeof(float2)));
    HIPCHK(hipDeviceSynchronize());
  }

  void set_reference_energy(int rhs, double e) {
    require(rhs >= 0 && rhs < nrhs, "rhs out of range");
    HIPCHK(hipMemcpy(d_ref + rhs, &e, sizeof(double), hipMemcpyHostToDevice));
  }

  // history length for a solve capped at max_iter: the cap applies at the
  // first multiple of Nw at or after it (Simulation.m:202), so the history
  // (whose end also stops the loop) must reach that iteration
  int history_for(long long max_iter) const {
    const long long cap = std::min<long long>(max_iter, 2000000);
    if (pstd) {
      // PSTD applies the cap on the first step at or after it whose source
      // amplitude has real part above 0.9: within about one optical cycle
      // once the ramp has passed 0.9 (from about 80 cycles on)
      const long long ramp = (long long)std::ceil(101.0 * pstd_ipc);
      return (int)std::min<long long>(std::max(cap, ramp) + (long long)std::ceil(2.0 * pstd_ipc) + 2, 2000001);
    }
    return (int)(((cap + nw - 1) / nw) * nw + 1);
  }

  void ensure_history(int cap) {
    if (cap <= hist_cap) return;
    dfree(d_diff);
    dfree(d_eall);
    d_diff = d_eall = nullptr;
    dalloc(&d_diff, (size_t)cap * nrhs);
    dalloc(&d_eall, (size_t)cap * nrhs);
    hist_cap = cap;
  }

  // force_run: every rhs iterates even without a source (kernel timing)
  void reset_state(long long max_iter, double thr, bool force_run = false) {
    DevState h{};
    h.it = 1;
    h.calc_energy = 1;
    h.all_stopped = 0;
    h.nw = nw;
    h.nrhs = nrhs;
    h.interval = interval;
    h.hist_cap = hist_cap;
    h.nblk = nblk;
    h.max_iter = max_iter;
    h.threshold = thr;
    h.stopped = d_stopped;
    h.converged = d_conv;
    h.n_iter = d_niter;
    h.last_energy = d_last;
    h.diff = d_diff;
    h.energy_all = d_eall;
    h.ref_energy = d_ref;
    h.warm = warm_mode ? 1 : 0;
    if (pstd) {
      // step 1: A_1 = default_source(1 / ipc), previous amplitude 1 (PSTD.m:83, 94-96)
      h.pstd = 1;
      h.ipc = pstd_ipc;
      double re, im;
      pstd_amplitude(1.0 / pstd_ipc, re, im);
      h.amp_re = re;
      h.amp_im = im;
      h.amp = make_float2((float)re, (float)im);
      const double dph = std::atan2(im, re);
      h.phase = make_float2((float)std::cos(dph), (float)std::sin(dph));
    }
    std::vector<int> st(nrhs), zero(nrhs, 0);
    int all = 1;
    for (int r = 0; r < nrhs; ++r) {
      st[r] = (has_source[r] || force_run) ? 0 : 1;  // zero source: nothing to do (Simulation.m:157-161)
      all &= st[r];
    }
    h.all_stopped = all;
    std::vector<double> inf(nrhs, INFINITY);
    if (arena == nullptr && !pstd) {
      if (h_stopped == nullptr)
        HIPCHK(hipHostMalloc((void**)&h_stopped, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
      *h_stopped = all;
      int* dev_flag = nullptr;
      HIPCHK(hipHostGetDevicePointer((void**)&dev_flag, h_stopped, 0));
      h.host_stopped = dev_flag;
    }
    HIPCHK(hipMemcpyAsync(dst, &h, sizeof(h), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_stopped, st.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_conv, zero.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_niter, zero.data(), nrhs * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_last, inf.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, stream));
    const size_t fe = (size_t)field_elems() * nrhs;
    HIPCHK(hipMemsetAsync(dE, 0, fe * cz(), stream));
    if (warm_mode) {
      // E' starts at gamma E0 (W0 is zero for rhs without a warm start)
      HIPCHK(hipMemcpyAsync(E, W0, fe * sizeof(float2), hipMemcpyDeviceToDevice, stream));
    } else {
      HIPCHK(hipMemsetAsync(E, 0, fe * cz(), stream));
    }
    // fused iterations: the first forward pass of iteration 1 runs here
    if (pre.fn) launch(pre);
    HIPCHK(hipStreamSynchronize(stream));
  }

  void capture_graph(int k) {
    if (gexec && graph_iters == k) return;
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    enqueue_iterations(k);
    HIPCHK(hipStreamEndCapture(stream, &graph));
    HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    graph_iters = k;
  }

  int solve(long long max_iter, double thr, long long batch, const char* cbtype, double cbint,
            wsim_callback cb, void* user) {
    require(max_iter >= 1, "max_iter must be >= 1");
    ensure_history(history_for(max_iter));
    reset_state(max_iter, thr);
    if (cb != nullptr && cbint >= 1.0) {
      // the reference calls the callback at every multiple of
      // callback_interval and at the stop (Simulation.m:213): batches end on
      // those iterations (the largest divisor of the interval up to 256
      // iterations, so a long interval does not make one huge graph)
      const long long ci = std::max<long long>(1, std::llround(cbint));
      long long d = std::min<long long>(ci, 256);
      while (ci % d != 0) --d;
      batch = d;
    }
    const bool use_graph = std::getenv("WSIM_NO_GRAPH") == nullptr && arena == nullptr;
    // Without a callback two batches are kept in flight: the host polls the
    // stop flag the control step writes into pinned host memory while the
    // next batch runs, so the GPU does not wait for the host between batches,
    // and a batch needs only cover the host's turn-around (about 150 us of
    // work, a multiple of Nw).  At most two batches of no-op launches follow
    // the stop.  With a callback (or without graphs) a batch ends in a host
    // synchronisation; it then holds about 2 ms of work.
    const bool pipelined = use_graph && cb == nullptr && h_stopped != nullptr;
    if (batch <= 0) {
      const double it_us = std::max(112.0 * (double)nvox * C * nrhs / 5e6, 20.0);  // ~5 TB/s, >= 20 us
      long long k = (long long)std::ceil((pipelined ? 150.0 : 2000.0) / it_us);
      k = std::max<long long>(nw, std::min<long long>(k, 256));
      batch = ((k + nw - 1) / nw) * nw;
    }
    if (use_graph) capture_graph((int)batch);
    if (pipelined) {
      for (hipEvent_t& e : batch_ev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      if (!*h_stopped) {
        HIPCHK(hipGraphLaunch(gexec, stream));
        HIPCHK(hipEventRecord(batch_ev[0], stream));
        for (long long b = 1;; ++b) {
          HIPCHK(hipGraphLaunch(gexec, stream));
          HIPCHK(hipEventRecord(batch_ev[b & 1], stream));
          HIPCHK(hipEventSynchronize(batch_ev[(b - 1) & 1]));
          if (__atomic_load_n(h_stopped, __ATOMIC_ACQUIRE)) break;
        }
      }
      HIPCHK(hipStreamSynchronize(stream));
    }
    long long done = 0;
    double next_cb = cbint > 0 ? cbint : INFINITY;
    std::vector<double> hist;
    int rc = WSIM_OK;
    for (; !pipelined;) {
      HIPCHK(hipMemcpy(&host_state, dst, sizeof(DevState), hipMemcpyDeviceToHost));
      if (host_state.all_stopped) break;
      if (use_graph)
        HIPCHK(hipGraphLaunch(gexec, stream));
      else
        enqueue_iterations((int)batch);
      done += batch;
      HIPCHK(hipStreamSynchronize(stream));
      if (cb != nullptr) {
        HIPCHK(hipMemcpy(&host_state, dst, sizeof(DevState), hipMemcpyDeviceToHost));
        const long long it_now = host_state.it - 1;
        if (it_now >= next_cb || host_state.all_stopped) {
          hist.resize((size_t)it_now);
          HIPCHK(hipMemcpy(hist.data(), d_diff, hist.size() * sizeof(double), hipMemcpyDeviceToHost));
          const int stop = cb(cbtype ? cbtype : "NoCallback", it_now, hist.data(), (long long)hist.size(),
                              thr, user);
          while (next_cb <= it_now) next_cb += cbint;
          if (stop) {
            rc = WSIM_ERR_ABORTED;
            break;
          }
        }
      }
    }
    if (!pstd) finish_e();
    HIPCHK(hipMemcpy(&host_state, dst, sizeof(DevState), hipMemcpyDeviceToHost));
    last_iters = host_state.it - 1;
    if (rc == WSIM_OK && any_diverged()) {
      g_err = "the iteration diverged (non-finite or growing increment energy)";
      rc = WSIM_ERR_DIVERGED;
    }
    return rc;
  }

  // some rhs stopped on the failure exit of the stop rule
  bool any_diverged() {
    std::vector<int> cv(nrhs);
    HIPCHK(hipMemcpy(cv.data(), d_conv, nrhs * sizeof(int), hipMemcpyDeviceToHost));
    for (int c : cv)
      if (c < 0) return true;
    return false;
  }

  // E' of every rhs complete (finish_e_kernel), ordered on the solver's stream
  void finish_e() {
    const long long fe = field_elems();
    const unsigned bx = (unsigned)std::min<long long>((fe + 255) / 256, 4096);
    if (f64)
      hipLaunchKernelGGL(finish_e_kernel<double2>, dim3(bx, (unsigned)nrhs), dim3(256), 0, stream,
                         (const DevState*)dst, reinterpret_cast<double2*>(E), reinterpret_cast<const double2*>(dE), fe);
    else
      hipLaunchKernelGGL(finish_e_kernel<float2>, dim3(bx, (unsigned)nrhs), dim3(256), 0, stream,
                         (const DevState*)dst, E, (const float2*)dE, fe);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
  }

  void result(int rhs, int64_t* n_iter, int* conv, double* diff, double* eall, int64_t cap) {
    require(rhs >= 0 && rhs < nrhs, "rhs out of range");
    int ni = 0, cv = 0, stp = 0;
    HIPCHK(hipMemcpy(&ni, d_niter + rhs, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&cv, d_conv + rhs, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&stp, d_stopped + rhs, sizeof(int), hipMemcpyDeviceToHost));
    long long n = stp ? ni : last_iters;  // aborted solves report the iterations run
    if (n_iter) *n_iter = n;
    if (conv) *conv = cv;
    const long long m = std::min<long long>(n, cap);
    if (m > 0 && diff)
      HIPCHK(hipMemcpy(diff, d_diff + (size_t)rhs * hist_cap, m * sizeof(double), hipMemcpyDeviceToHost));
    if (m > 0 && eall)
      HIPCHK(hipMemcpy(eall, d_eall + (size_t)rhs * hist_cap, m * sizeof(double), hipMemcpyDeviceToHost));
  }

  // out / in: complex values of the solver's precision
  void get_field(int rhs, void* out) {
    require(rhs >= 0 && rhs < nrhs && out != nullptr, "bad argument");
    const size_t fe = (size_t)field_elems();
    HIPCHK(hipMemcpy(out, at(E, (size_t)rhs * fe), fe * cz(), hipMemcpyDeviceToHost));
  }

  void propagate(int w, const void* in, void* out) {
    require(w >= 0 && w < nw && in && out, "bad argument");
    const size_t fe = (size_t)field_elems();
    if (!probe) dalloc_c(&probe, fe);
    HIPCHK(hipMemcpy(probe, in, fe * cz(), hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
    DevState ps{};
    ps.it = w + 1 + nw;  // wiggle w, past the source iterations
    ps.nw = nw;
    ps.nrhs = 1;
    ps.stopped = p_stopped;
    HIPCHK(hipMemcpy(dprobe, &ps, sizeof(ps), hipMemcpyHostToDevice));
    if (probe_iter.empty()) {
      build_iteration(probe_iter, true);
      upload_args(probe_iter, &d_probe_args);
    }
    for (const auto& L : probe_iter) launch(L);
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(hipMemcpy(out, probe, fe * cz(), hipMemcpyDeviceToHost));
  }

  int profile(int reps, int maxk, double* ms, double* bytes, const char** names, int* nk) {
    require(reps >= 1, "reps must be >= 1");
    const int K = (int)iter.size() + 1;
    if (nk) *nk = K;
    std::vector<hipEvent_t> ev(2 * K);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    std::vector<double> acc(K, 0.0);
    // run on a fresh state that never stops
    ensure_history(reps + nw + 2);
    reset_state(1ll << 40, -1.0, true);
    enqueue_iterations(nw);  // past the source iterations: steady-state kernels
    HIPCHK(hipStreamSynchronize(stream));
    for (int r = 0; r < reps; ++r) {
      for (int k = 0; k < K; ++k) {
        HIPCHK(hipEventRecord(ev[2 * k], stream));
        if (k < K - 1)
          launch(iter[k]);
        else
          launch_control();
        HIPCHK(hipEventRecord(ev[2 * k + 1], stream));
      }
      HIPCHK(hipStreamSynchronize(stream));
      for (int k = 0; k < K; ++k) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
        acc[k] += t;
      }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < K && k < maxk; ++k) {
      if (ms) ms[k] = acc[k] / reps;
      if (bytes) bytes[k] = k < K - 1 ? iter[k].bytes : 0.0;
      if (names) names[k] = k < K - 1 ? iter[k].name : "control";
    }
    return WSIM_OK;
  }
};

template <class F>
static int guard(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return WSIM_ERR_HIP;
  } catch (...) {
    g_err = "unknown error";
    return WSIM_ERR_HIP;
  }
}


// ----------------------------------------------------------------------------
// groups: several independent problems (any grid sizes) advanced together.
// Problems whose iterations launch the same kernel sequence share one
// launch per pass (up to 64 problems per launch, each owning a range of
// workgroups); one control launch serves every problem.  A replayed graph
// of B iterations therefore costs (passes + 1) launches per iteration
// whatever the number of problems.
// ----------------------------------------------------------------------------
constexpr int kGroupMax = 64;

template <class T_>
static void hip_alloc(T_** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T_)));
}

struct Group {
  std::vector<Solver*> ps;
  hipStream_t stream = nullptr;
  struct Sub {
    std::vector<int> members;
    std::vector<Launch> launches;
    PassArgs* d_tab = nullptr;  // [passes][members]
    int* d_end = nullptr;       // [passes][members]
  };
  std::vector<Sub> subs;
  CtlEntry* d_ctl = nullptr;
  int* d_done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;

  ~Group() { release(); }

  void release_tables() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    for (auto& sb : subs) {
      if (sb.d_tab) (void)hipFree(sb.d_tab);
      if (sb.d_end) (void)hipFree(sb.d_end);
    }
    subs.clear();
    if (d_ctl) (void)hipFree(d_ctl);
    if (d_done) (void)hipFree(d_done);
    d_ctl = nullptr;
    d_done = nullptr;
  }

  void release() {
    release_tables();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  void create(Solver* const* s, int n) {
    require(n >= 1, "a group needs at least one solver");
    for (int i = 0; i < n; ++i) {
      require(s[i] != nullptr, "null solver in group");
      for (int j = 0; j < i; ++j) require(s[j] != s[i], "a solver appears twice in the group");
      ps.push_back(s[i]);
    }
    HIPCHK(hipSetDevice(ps[0]->dev));
    for (Solver* p : ps) require(p->dev == ps[0]->dev, "every solver of a group must be on one device");
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  }

  // grouped launch tables from the solvers' current launches (sources,
  // plans and media may have changed since the last solve)
  void build() {
    release_tables();
    std::vector<int> todo(ps.size());
    for (size_t i = 0; i < ps.size(); ++i) todo[i] = (int)i;
    auto same_seq = [&](int a, int b) {
      const auto &A = ps[a]->iter, &B = ps[b]->iter;
      if (A.size() != B.size()) return false;
      for (size_t k = 0; k < A.size(); ++k)
        // the same kernels: plan-specialised kernels are shared by problems of
        // one shape only (32 ragged 2-D problems in six sizes: 0.1291 s
        // grouped by kernel against 0.1331 s all in one launch on the
        // run-time-plan kernels)
        if (A[k].fn != B[k].fn || A[k].threads != B[k].threads) return false;
      return true;
    };
    while (!todo.empty()) {
      Sub sb;
      std::vector<int> rest;
      for (int i : todo) {
        if ((sb.members.empty() || same_seq(sb.members[0], i)) && (int)sb.members.size() < kGroupMax)
          sb.members.push_back(i);
        else
          rest.push_back(i);
      }
      todo.swap(rest);
      const size_t K = ps[sb.members[0]]->iter.size(), M = sb.members.size();
      std::vector<PassArgs> tab(K * M);
      std::vector<int> ends(K * M);
      hip_alloc(&sb.d_tab, K * M);
      hip_alloc(&sb.d_end, K * M);
      for (size_t k = 0; k < K; ++k) {
        Launch G = ps[sb.members[0]]->iter[k];
        long long x = 0;
        unsigned y = 1;
        size_t lds = 0;
        for (size_t m = 0; m < M; ++m) {
          const Launch& L = ps[sb.members[m]]->iter[k];
          x += L.grid.x;
          require(x < (1ll << 31), "group too large for one launch");
          y = std::max(y, L.grid.y);
          lds = std::max(lds, L.lds);
          ends[k * M + m] = (int)x;
          PassArgs a = L.args;
          a.grp_n = (int)M;
          a.grp_end = sb.d_end + k * M;
          tab[k * M + m] = a;
        }
        G.grid = dim3((unsigned)x, y, 1);
        G.lds = lds;
        G.dargs = sb.d_tab + k * M;
        if (G.lds > 48 * 1024)
          HIPCHK(hipFuncSetAttribute((const void*)G.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)G.lds));
        sb.launches.push_back(G);
      }
      HIPCHK(hipMemcpy(sb.d_tab, tab.data(), tab.size() * sizeof(PassArgs), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(sb.d_end, ends.data(), ends.size() * sizeof(int), hipMemcpyHostToDevice));
      subs.push_back(std::move(sb));
    }
    hip_alloc(&d_done, ps.size());
    hip_alloc(&d_ctl, ps.size());
    std::vector<CtlEntry> ce(ps.size());
    for (size_t i = 0; i < ps.size(); ++i)
      ce[i] = CtlEntry{ps[i]->dst, ps[i]->partials, ps[i]->nblk, d_done + i};
    HIPCHK(hipMemcpy(d_ctl, ce.data(), ce.size() * sizeof(CtlEntry), hipMemcpyHostToDevice));
  }

  void enqueue_iterations(int k) {
    for (int i = 0; i < k; ++i) {
      for (const auto& sb : subs)
        for (const auto& L : sb.launches) {
          hipLaunchKernelGGL(L.fn, L.grid, dim3(L.threads), L.lds, stream, L.dargs);
          HIPCHK(hipGetLastError());
        }
      hipLaunchKernelGGL(control_group_kernel, dim3((unsigned)ps.size()), dim3(kControlThreads), 0, stream,
                         (const CtlEntry*)d_ctl);
      HIPCHK(hipGetLastError());
    }
  }

  // every problem from E' = dE' = 0 (or its warm start) to its own stop
  int solve(const int64_t* max_iter, const double* thr, long long batch) {
    int nw = 1;
    double it_us = 0.0;
    for (size_t i = 0; i < ps.size(); ++i) {
      Solver* p = ps[i];
      require(max_iter[i] >= 1, "max_iter must be >= 1");
      if (p->iter.empty() || !p->d_args) p->rebuild_launches();
      p->ensure_history(p->history_for(max_iter[i]));
      p->reset_state(max_iter[i], thr[i]);
      nw = std::max(nw, p->nw);
      it_us += 112.0 * (double)p->nvox * p->C * p->nrhs / 5e6;
    }
    build();
    if (batch <= 0) {
      long long k = (long long)std::ceil(2000.0 / std::max(it_us, 1.0));
      k = std::max<long long>(nw, std::min<long long>(k, 256));
      batch = ((k + nw - 1) / nw) * nw;
    }
    std::vector<int> done(ps.size());
    for (size_t i = 0; i < ps.size(); ++i) {
      HIPCHK(hipMemcpy(&ps[i]->host_state, ps[i]->dst, sizeof(DevState), hipMemcpyDeviceToHost));
      done[i] = ps[i]->host_state.all_stopped;
    }
    HIPCHK(hipMemcpy(d_done, done.data(), done.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    enqueue_iterations((int)batch);
    HIPCHK(hipStreamEndCapture(stream, &graph));
    HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    for (;;) {
      bool all = true;
      for (int d : done) all = all && d;
      if (all) break;
      HIPCHK(hipGraphLaunch(gexec, stream));
      HIPCHK(hipStreamSynchronize(stream));
      HIPCHK(hipMemcpy(done.data(), d_done, done.size() * sizeof(int), hipMemcpyDeviceToHost));
    }
    for (Solver* p : ps) {
      p->finish_e();
      HIPCHK(hipMemcpy(&p->host_state, p->dst, sizeof(DevState), hipMemcpyDeviceToHost));
      p->last_iters = p->host_state.it - 1;
    }
    for (Solver* p : ps)
      if (p->any_diverged()) {
        g_err = "a problem of the group diverged (non-finite or growing increment energy)";
        return WSIM_ERR_DIVERGED;
      }
    return WSIM_OK;
  }
};

// ----------------------------------------------------------------------------
// scratch of the allocation-free entry point: one static device array per GPU
// (each device that loads the library holds its own copy until the process
// ends).  Host code never uses the name `scratch` as an address: its device
// address comes from hipGetSymbolAddress, once per device (scratch_ptr).
// ----------------------------------------------------------------------------
constexpr size_t kScratchBytes = 64ull << 20;
constexpr long long kNoAllocMaxIter = 100000;  // history length held in the scratch
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;  // calls on one device run one at a time
static void* g_scratch_ptr[kMaxDevices] = {};

// device address of the scratch array on the current device (caller holds g_scratch_mu)
static char* scratch_ptr(int dev) {
  require(dev >= 0 && dev < kMaxDevices, "device index out of range", WSIM_ERR_HIP);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, "no device address for the scratch array", WSIM_ERR_HIP);
    g_scratch_ptr[dev] = p;
  }
  return static_cast<char*>(g_scratch_ptr[dev]);
}

}  // namespace wsim

using namespace wsim;

struct wsim_solver {
  Solver s;
};

extern "C" {

const char* wsim_last_error(void) { return g_err.c_str(); }
int wsim_abi_version(void) { return WSIM_ABI_VERSION; }

int wsim_selftest(void) {
  return guard([&]() -> int {
    const uint32_t ds[] = {2, 3, 5, 7, 11, 16, 40, 64, 320, 1000, 1152, 2187, 102400, 32768000};
    for (uint32_t d : ds) {
      FastDiv f = make_fastdiv(d);
      for (uint64_t x = 0; x < (1ull << 31); x += 1 + x / 7)
        if (host_fdiv((uint32_t)x, f) != (uint32_t)x / d) {
          g_err = "fastdiv mismatch d=" + std::to_string(d) + " x=" + std::to_string(x);
          return WSIM_ERR_HIP;
        }
      for (uint32_t x = (1u << 31) - 4096; x < (1u << 31); ++x)
        if (host_fdiv(x, f) != x / d) {
          g_err = "fastdiv mismatch near 2^31 d=" + std::to_string(d);
          return WSIM_ERR_HIP;
        }
    }
    const int ns[] = {1, 2, 8, 12, 64, 192, 256, 320, 1100, 1152, 2187, 840, 77};
    for (int n : ns) {
      auto r = default_radices(n);
      long long p = 1;
      for (int x : r) p *= x;
      if (n > 1 && p != n) {
        g_err = "radix plan mismatch for n=" + std::to_string(n);
        return WSIM_ERR_HIP;
      }
    }
    return WSIM_OK;
  });
}

int wsim_solver_create(wsim_solver** out, const wsim_solver_desc* desc, const wsim_cfloat* gamma,
                       const wsim_wiggle* wiggles) {
  return guard([&]() -> int {
    require(out != nullptr, "null output handle");
    *out = nullptr;
    auto* h = new wsim_solver();
    try {
      h->s.create(desc, gamma, wiggles);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return WSIM_OK;
  });
}

int wsim_solver_create_f64(wsim_solver** out, const wsim_solver_desc_f64* desc, const wsim_cdouble* gamma,
                           const wsim_wiggle_f64* wiggles) {
  return guard([&]() -> int {
    require(out != nullptr, "null output handle");
    *out = nullptr;
    require(desc != nullptr && gamma != nullptr && wiggles != nullptr, "null argument");
    wsim_solver_desc d{};
    for (int c = 0; c < 4; ++c) {
      d.N[c] = desc->N[c];
      d.roi[0][c] = desc->roi[0][c];
      d.roi[1][c] = desc->roi[1][c];
    }
    d.nrhs = desc->nrhs;
    d.nwiggles = desc->nwiggles;
    d.epsilon = (float)desc->epsilon;
    d.k02e = wsim_cfloat{(float)desc->k02e.re, (float)desc->k02e.im};
    d.energy_interval = desc->energy_interval;
    auto* h = new wsim_solver();
    try {
      h->s.f64 = true;
      h->s.epsilon_d = desc->epsilon;
      h->s.k02e_d = make_double2(desc->k02e.re, desc->k02e.im);
      h->s.create(&d, gamma, wiggles);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return WSIM_OK;
  });
}

int wsim_solver_set_source_f64(wsim_solver* s, int rhs, int np, const wsim_source_patch_f64* p) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    s->s.set_source_f64(rhs, np, p);
    return WSIM_OK;
  });
}

int wsim_solver_get_field_f64(wsim_solver* s, int rhs, wsim_cdouble* E_out) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    require(s->s.f64, "this solver computes in complex64: read its field with wsim_solver_get_field");
    s->s.get_field(rhs, E_out);
    return WSIM_OK;
  });
}

int wsim_solver_propagate_f64(wsim_solver* s, int w, const wsim_cdouble* in, wsim_cdouble* out) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    require(s->s.f64, "this solver computes in complex64: use wsim_solver_propagate");
    s->s.propagate(w, in, out);
    return WSIM_OK;
  });
}

int wsim_solver_fixed_launches(wsim_solver* s) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    return s->s.fixed_launches();
  });
}

int wsim_solver_element_size(wsim_solver* s) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    return (int)s->s.cz();
  });
}

int wsim_solver_set_source(wsim_solver* s, int rhs, int np, const wsim_source_patch* p) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    s->s.set_source(rhs, np, p);
    return WSIM_OK;
  });
}

int wsim_solver_solve(wsim_solver* s, int64_t max_iter, double thr, int64_t batch, const char* cbtype,
                      double cbint, wsim_callback cb, void* user) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    return s->s.solve(max_iter, thr, batch, cbtype, cbint, cb, user);
  });
}

int wsim_solver_result(wsim_solver* s, int rhs, int64_t* n_iter, int* conv, double* diff, double* eall,
                       int64_t cap) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    s->s.result(rhs, n_iter, conv, diff, eall, cap);
    return WSIM_OK;
  });
}

int wsim_solver_get_field(wsim_solver* s, int rhs, wsim_cfloat* E_out) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    require(!s->s.f64, "this solver computes in complex128: read its field with wsim_solver_get_field_f64");
    s->s.get_field(rhs, E_out);
    return WSIM_OK;
  });
}

int wsim_solver_device_ptr(wsim_solver* s, int which, int rhs, void** dptr) {
  return guard([&]() -> int {
    require(s != nullptr && dptr != nullptr, "null argument");
    Solver& v = s->s;
    require(rhs >= 0 && rhs < v.nrhs, "rhs out of range");
    const size_t off = (size_t)rhs * v.field_elems();
    switch (which) {
      case 0: *dptr = v.at(v.E, off); break;
      case 1: *dptr = v.at(v.dE, off); break;
      case 2: *dptr = v.gamma; break;
      case 3: *dptr = v.at(v.T, off); break;
      case 4: *dptr = v.at(v.S, off); break;
      default: throw Error(WSIM_ERR_ARG, "which must be 0..4");
    }
    return WSIM_OK;
  });
}

int wsim_solver_propagate(wsim_solver* s, int w, const wsim_cfloat* in, wsim_cfloat* out) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    require(!s->s.f64, "this solver computes in complex128: use wsim_solver_propagate_f64");
    s->s.propagate(w, in, out);
    return WSIM_OK;
  });
}

int wsim_solver_profile(wsim_solver* s, int reps, int maxk, double* ms, double* bytes, const char** names,
                        int* nk) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    return s->s.profile(reps, maxk, ms, bytes, names, nk);
  });
}

int wsim_solver_tune(wsim_solver* s, int reps, char* plan_out, int plan_cap, double* ms_per_iteration) {
  return guard([&]() -> int {
    require(s != nullptr && reps >= 1, "bad argument");
    const std::string p = s->s.tune(reps);
    if (ms_per_iteration) *ms_per_iteration = s->s.time_iteration(reps);
    if (plan_out && plan_cap > 0) {
      std::strncpy(plan_out, p.c_str(), (size_t)plan_cap - 1);
      plan_out[plan_cap - 1] = 0;
    }
    return WSIM_OK;
  });
}

int wsim_solver_get_plan(wsim_solver* s, char* plan_out, int plan_cap) {
  return guard([&]() -> int {
    require(s != nullptr && plan_out != nullptr && plan_cap > 0, "bad argument");
    const std::string p = s->s.plan_string();
    std::strncpy(plan_out, p.c_str(), (size_t)plan_cap - 1);
    plan_out[plan_cap - 1] = 0;
    return WSIM_OK;
  });
}

int wsim_solver_set_plan(wsim_solver* s, const char* plan) {
  return guard([&]() -> int {
    require(s != nullptr && plan != nullptr, "bad argument");
    s->s.set_plan(plan);
    return WSIM_OK;
  });
}

int wsim_solver_set_medium(wsim_solver* s, const wsim_cfloat* gamma, float epsilon, wsim_cfloat k02e,
                           const wsim_wiggle* wiggles) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    s->s.set_medium(gamma, epsilon, k02e, wiggles);
    return WSIM_OK;
  });
}

int wsim_solver_set_warm_start(wsim_solver* s, int rhs, const wsim_cfloat* E0, int on_device) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    s->s.set_warm(rhs, E0, on_device != 0);
    return WSIM_OK;
  });
}

int wsim_solver_clear_warm_start(wsim_solver* s) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    s->s.clear_warm();
    return WSIM_OK;
  });
}

int wsim_solver_set_reference_energy(wsim_solver* s, int rhs, double energy) {
  return guard([&]() -> int {
    require(s != nullptr, "null solver");
    s->s.set_reference_energy(rhs, energy);
    return WSIM_OK;
  });
}

struct wsim_group {
  Group g;
};

int wsim_group_create(wsim_group** out, wsim_solver* const* solvers, int n) {
  return guard([&]() -> int {
    require(out != nullptr && solvers != nullptr, "null argument");
    require(n >= 1 && n <= 4096, "group size out of range");
    std::vector<Solver*> v(n);
    for (int i = 0; i < n; ++i) {
      require(solvers[i] != nullptr, "null solver in group");
      v[i] = &solvers[i]->s;
    }
    auto* g = new wsim_group();
    try {
      g->g.create(v.data(), n);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
    return WSIM_OK;
  });
}

int wsim_group_solve(wsim_group* g, const int64_t* max_iter, const double* energy_threshold, int64_t batch) {
  return guard([&]() -> int {
    require(g != nullptr && max_iter != nullptr && energy_threshold != nullptr, "null argument");
    HIPCHK(hipSetDevice(g->g.ps[0]->dev));
    return g->g.solve(max_iter, energy_threshold, batch);
  });
}

int wsim_group_destroy(wsim_group* g) {
  return guard([&]() -> int {
    delete g;
    return WSIM_OK;
  });
}

int wsim_solver_destroy(wsim_solver* s) {
  return guard([&]() -> int {
    delete s;
    return WSIM_OK;
  });
}

int wsim_run_algorithm(const double N[4], const wsim_cfloat* gamma, int nwiggles, const wsim_wiggle* wiggles,
                       int npatches, const wsim_source_patch* patches, float epsilon, wsim_cfloat k02e,
                       const double roi[8], double max_iter, double energy_threshold, const char* callback_type,
                       double callback_interval, wsim_callback cb, void* user, wsim_cfloat* E_out,
                       double* diff_energy_out, int64_t diff_energy_cap, int64_t* n_diff_energy) {
  return guard([&]() -> int {
    require(N != nullptr && roi != nullptr && E_out != nullptr, "null argument");
    wsim_solver_desc d{};
    for (int c = 0; c < 4; ++c) d.N[c] = (int64_t)std::llround(N[c]);
    d.nrhs = 1;
    d.nwiggles = nwiggles;
    d.epsilon = epsilon;
    d.k02e = k02e;
    for (int c = 0; c < 4; ++c) {
      d.roi[0][c] = (int64_t)std::llround(roi[2 * c]);
      d.roi[1][c] = (int64_t)std::llround(roi[2 * c + 1]);
    }
    // energy_calculation_interval = min(8, callback_interval) (Simulation.m:61, 135)
    d.energy_interval = (int)std::max(1.0, std::min(8.0, callback_interval > 0 ? callback_interval : 8.0));
    // maxIter: max_iterations or 100000 (WaveSimBase.m:127-131)
    // (clamped before the conversion: a double beyond the int64 range has no
    // defined conversion, and the energy history stops at 2e6 iterations anyway)
    long long mi = (std::isfinite(max_iter) && max_iter > 0) ? (long long)std::min(max_iter, 1e15) : 100000;
    Solver sv;
    sv.create(&d, gamma, wiggles);
    sv.set_source(0, npatches, patches);
    const int rc = sv.solve(mi, energy_threshold, 0, callback_type, callback_interval, cb, user);
    int64_t n = 0;
    int conv = 0;
    sv.result(0, &n, &conv, diff_energy_out, nullptr, diff_energy_out ? diff_energy_cap : 0);
    if (n_diff_energy) *n_diff_energy = n;
    sv.get_field(0, E_out);
    return rc;
  });
}

// ---- PSTD (PSTD.m) ----------------------------------------------------------
struct wsim_pstd {
  Solver s;
};

int wsim_pstd_create(wsim_pstd** out, const wsim_pstd_desc* desc, const float* c1, const float* c2,
                     const float* c3, const float* py, const float* px, const float* pz) {
  return guard([&]() -> int {
    require(out != nullptr && desc != nullptr && c1 && c2 && c3 && py && px && pz, "null argument");
    *out = nullptr;
    require(desc->iterations_per_cycle > 0 && std::isfinite(desc->iterations_per_cycle),
            "iterations_per_cycle must be positive");
    wsim_solver_desc d{};
    for (int a = 0; a < 3; ++a) d.N[a] = desc->N[a];
    d.N[3] = 1;
    d.nrhs = 1;
    d.nwiggles = 1;
    d.epsilon = 1.f;
    for (int c = 0; c < 4; ++c) {
      d.roi[0][c] = desc->roi[0][c];
      d.roi[1][c] = desc->roi[1][c];
    }
    d.energy_interval = desc->energy_interval;
    // the wiggle tables carry the unscaled coordinates and unit ramps
    const int64_t nmax = std::max(d.N[0], std::max(d.N[1], d.N[2]));
    require(nmax >= 1 && nmax < (1 << 30), "grid size out of range");
    std::vector<wsim_cfloat> ones((size_t)nmax, wsim_cfloat{1.f, 0.f});
    const wsim_wiggle wg{px, py, pz, ones.data(), ones.data(), ones.data()};
    auto* h = new wsim_pstd();
    try {
      h->s.pstd = true;
      h->s.pstd_ipc = desc->iterations_per_cycle;
      h->s.create(&d, nullptr, &wg);
      const size_t nv = (size_t)h->s.nvox;
      for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpy(h->s.coef + k * nv, k == 0 ? c1 : k == 1 ? c2 : c3, nv * sizeof(float),
                         hipMemcpyHostToDevice));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return WSIM_OK;
  });
}

int wsim_pstd_set_source(wsim_pstd* p, int np, const wsim_source_patch* patches) {
  return guard([&]() -> int {
    require(p != nullptr, "null handle");
    p->s.set_source(0, np, patches);
    return WSIM_OK;
  });
}

int wsim_pstd_run(wsim_pstd* p, double max_iterations, double thr, int64_t* n_iter, int* converged) {
  return guard([&]() -> int {
    require(p != nullptr, "null handle");
    const long long mi = (std::isfinite(max_iterations) && max_iterations > 0)
                             ? (long long)std::min(max_iterations, 1e15)
                             : 2000000;
    const int rc = p->s.solve(mi, thr, 0, nullptr, 0.0, nullptr, nullptr);
    int64_t n = 0;
    int cv = 0;
    p->s.result(0, &n, &cv, nullptr, nullptr, 0);
    if (n_iter) *n_iter = n;
    if (converged) *converged = cv;
    return rc;
  });
}

int wsim_pstd_result(wsim_pstd* p, wsim_cfloat* E_out, double* diff_energy, int64_t cap) {
  return guard([&]() -> int {
    require(p != nullptr && E_out != nullptr, "null argument");
    Solver& v = p->s;
    int64_t n = 0;
    int cv = 0;
    v.result(0, &n, &cv, diff_energy, nullptr, diff_energy ? cap : 0);
    // step k writes fld[k & 1] (E before any step: fld[0], zero)
    const float2* F = (n & 1) ? v.dE : v.E;
    HIPCHK(hipMemcpy(E_out, F, (size_t)v.nvox * sizeof(float2), hipMemcpyDeviceToHost));
    return WSIM_OK;
  });
}

int wsim_pstd_fixed_launches(wsim_pstd* p) {
  return guard([&]() -> int {
    require(p != nullptr, "null handle");
    return p->s.fixed_launches();
  });
}

int wsim_pstd_destroy(wsim_pstd* p) {
  return guard([&]() -> int {
    delete p;
    return WSIM_OK;
  });
}

int wsim_run_algorithm_noalloc(const double N[4], const wsim_cfloat* gamma_dev, int nwiggles,
                               const wsim_wiggle* wiggles, int npatches, const wsim_source_patch* patches,
                               float epsilon, wsim_cfloat k02e, const double roi[8], double max_iter,
                               double energy_threshold, const char* callback_type, double callback_interval,
                               wsim_callback cb, void* user, wsim_cfloat* E_out_dev, double* diff_energy_out,
                               int64_t diff_energy_cap, int64_t* n_diff_energy, void* stream) {
  return guard([&]() -> int {
    require(N != nullptr && roi != nullptr && gamma_dev != nullptr && E_out_dev != nullptr, "null argument");
    wsim_solver_desc d{};
    for (int c = 0; c < 4; ++c) d.N[c] = (int64_t)std::llround(N[c]);
    d.nrhs = 1;
    d.nwiggles = nwiggles;
    d.epsilon = epsilon;
    d.k02e = k02e;
    for (int c = 0; c < 4; ++c) {
      d.roi[0][c] = (int64_t)std::llround(roi[2 * c]);
      d.roi[1][c] = (int64_t)std::llround(roi[2 * c + 1]);
    }
    d.energy_interval = (int)std::max(1.0, std::min(8.0, callback_interval > 0 ? callback_interval : 8.0));
    const long long mi = (std::isfinite(max_iter) && max_iter > 0) ? (long long)std::min(max_iter, 1e15) : 100000;
    require(mi <= kNoAllocMaxIter, "the allocation-free entry point runs at most " +
                                       std::to_string(kNoAllocMaxIter) + " iterations", WSIM_ERR_UNSUPPORTED);
    // the three work fields dE', T and S dominate; the arena check below is exact
    long long nv = 1;
    for (int c = 0; c < 3; ++c) {
      require(d.N[c] >= 1 && d.N[c] < (1 << 30), "grid size out of range");
      nv *= d.N[c];
    }
    require(d.N[3] == 1 || d.N[3] == 3, "N(4) must be 1 (scalar) or 3 (vector)");
    const double need = 3.0 * 8.0 * (double)d.N[3] * (double)nv;
    require(need <= (double)kScratchBytes,
            "the allocation-free entry point holds problems of at most " + std::to_string(kScratchBytes >> 20) +
                " MiB of work fields (3 x Npol x voxels x 8 bytes); this one needs " +
                std::to_string((long long)(need / (1 << 20))) + " MiB",
            WSIM_ERR_UNSUPPORTED);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    char* base = scratch_ptr(dev);
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIPCHK(hipStreamSynchronize(st));  // gamma and E' may come from the caller's pending work
    Solver sv;
    sv.arena = base;
    sv.arena_cap = kScratchBytes;
    sv.stream = st;
    sv.external_stream = true;
    sv.ext_gamma = sv.ext_E = true;
    sv.E = reinterpret_cast<float2*>(E_out_dev);
    sv.create(&d, gamma_dev, wiggles);
    sv.set_source(0, npatches, patches);
    const int rc = sv.solve(mi, energy_threshold, 0, callback_type, callback_interval, cb, user);
    int64_t n = 0;
    int conv = 0;
    sv.result(0, &n, &conv, diff_energy_out, nullptr, diff_energy_out ? diff_energy_cap : 0);
    if (n_diff_energy) *n_diff_energy = n;
    HIPCHK(hipStreamSynchronize(st));
    return rc;
  });
}

}  // extern "C"
