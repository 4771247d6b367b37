// Device side of the MBS solver: one templated "axis pass" kernel that runs a
// batch of 1-D Stockham FFTs along one grid axis through LDS, with the
// elementwise work of the MBS iteration fused into its first load / last store:
//
//   PRO_SRC_RAMP  (first forward pass): load dE' (+ i/eps/Nw * S during the
//                 first Nw iterations), multiply by the wiggle ramps gx gy gz
//                 (Source.m:139-172, WaveSim.m:38-48, private/f_wiggle.m:5)
//   MODE_MID      (middle pass): forward FFT along the last axis, multiply by
//                 the Green's function (private/f_g0_scalar.m:6 or
//                 private/f_g0_vector.m:4-8) and 1/N, inverse FFT along the
//                 same axis -- the spectrum never leaves registers/LDS
//   EPI_MIX       (last inverse pass): undo the ramps (WaveSim.m:49-59), mix
//                 and accumulate (private/f_mix.m:3-4, WaveSimBase.m:179-180),
//                 and the per-workgroup ROI energy partial sums
//                 (Simulation.m:265-273)
//   EPI_MIX_FWD   as EPI_MIX, then the next iteration's first forward pass on
//                 the same lines: dE' (+ the next step's source / warm-start
//                 portion) times the next wiggle's ramps goes to LDS and
//                 through the forward FFT stages to the work field.  The
//                 first forward pass then runs only once, before iteration 1.
//   EPI_PSTD_FWD  the PSTD update (EPI_PSTD below), then the next step's first
//                 forward pass on the new time level, fused the same way.
//
// Layout: a field is [rhs][comp][z][x][y] complex64 (MATLAB column-major
// [y,x,z,pol], y contiguous).  A pass along axis `a` sees each component as
// [outer][n][inner]: the first non-trivial axis has inner == 1 (CONTIG kernel,
// a workgroup owns `ncol` whole lines), the others are strided (a workgroup
// owns a tile of `ncol` consecutive inner indices of one outer index, so each
// row of the tile is one contiguous 8*ncol-byte segment).
//
// Stockham autosort (Govindaraju et al. 2008): stage s with radix R and
// Ns = product of earlier radices reads x[j + r n/R], twiddles by
// exp(-+2 pi i r (j mod Ns)/(Ns R)), does a radix-R DFT and writes
// y[(j - j mod Ns) R + j mod Ns + r Ns].  The first stage reads straight from
// HBM and the last stage writes straight to HBM; only the stages in between go
// through LDS.  In the middle pass the inverse transform uses the forward
// radices in reverse order, so the forward transform's last-stage outputs are
// exactly the inverse transform's first-stage inputs held by the same thread:
// the Green's-function multiply happens in registers.
#pragma once
#include <stdint.h>

#include <type_traits>

#include "fft_radix.hpp"

namespace wsim {

constexpr int MAXST = 10;  // 3^7 = 2187 needs 7 stages

struct FastDiv {
  uint32_t d, m, sh;  // q = umulhi(x, m) >> sh  for x < 2^31, d >= 2
};

__device__ __forceinline__ uint32_t fdiv(uint32_t x, const FastDiv& f) {
  return f.d == 1 ? x : (__umulhi(x, f.m) >> f.sh);
}

struct Stage {
  int R;   // radix
  int Ns;  // product of the radices of the earlier stages
  int nb;  // n / R butterflies per line
  int tw;  // offset of this stage's twiddles: tw[(r-1)*Ns + k], k < Ns, 1 <= r < R
  FastDiv fNs, fnb;
  int nbp, nsp;  // strided tiles: nb and Ns rows in LDS words of float2 (times the pitch)
};

struct AxisPlan {
  int n, nst;
  Stage f[MAXST];  // forward stages (sign -1)
  Stage i[MAXST];  // inverse stages (sign +1), radices in reverse order
};

// A stage whose numbers are compile-time constants (kernels specialised for
// one plan): the same fields as Stage, so that divisions by Ns and nb become
// multiply-shifts by constants and LDS offsets become instruction immediates.
template <int R_, int NS, int NB, int TW, int PITCH>
struct CStage {
  static constexpr int R = R_, Ns = NS, nb = NB, tw = TW;
  static constexpr int nbp = NB * PITCH, nsp = NS * PITCH;
};
__device__ __forceinline__ uint32_t div_nb(const Stage& s, uint32_t x) { return fdiv(x, s.fnb); }
__device__ __forceinline__ uint32_t div_Ns(const Stage& s, uint32_t x) { return fdiv(x, s.fNs); }
template <int R, int NS, int NB, int TW, int P>
__device__ __forceinline__ uint32_t div_nb(const CStage<R, NS, NB, TW, P>&, uint32_t x) { return x / (uint32_t)NB; }
template <int R, int NS, int NB, int TW, int P>
__device__ __forceinline__ uint32_t div_Ns(const CStage<R, NS, NB, TW, P>&, uint32_t x) { return x / (uint32_t)NS; }

// The stage sequence of a pass: read from PassArgs at run time (DynPlan, one
// kernel for every plan), or fixed at compile time (FixedPlan: transform
// length N, NCOL lines per workgroup, NTH threads, radices Rs in forward
// order; the host launches it only for launches of exactly that shape).
// kWaves: the waves per SIMD the kernel's registers must allow (the
// workgroups its LDS lets a CU hold), a launch bound for the compiler.
// kEarly: issue the first stage's HBM loads before the LDS tables are built
// (pays where few workgroups share a CU and cannot hide each other's set-up).
struct DynPlan {
  static constexpr bool kFixed = false, kEarly = false;
  static constexpr int kN = 0, kNcol = 0, kNT = 0, kWaves = 1;
};
template <int N, int NCOL, int NTH, int WAVES, bool EARLY, int... Rs>
struct FixedPlan {
  static constexpr bool kFixed = true, kEarly = EARLY;
  static constexpr int kN = N, kNcol = NCOL, kNT = NTH, kWaves = WAVES;
  static constexpr int S = (int)sizeof...(Rs);
  static constexpr int r[S] = {Rs...};
  // radix of stage i of the forward (INV false) or inverse (reversed) transform
  static constexpr int rad(int i, bool inv) { return inv ? r[S - 1 - i] : r[i]; }
  static constexpr int span(int i, bool inv) {
    int p = 1;
    for (int k = 0; k < i; ++k) p *= rad(k, inv);
    return p;
  }
  // offset of stage i's twiddles in the axis table (the host's layout:
  // (R-1)*Ns entries per stage, stages in order)
  static constexpr int twoff(int i, bool inv) {
    int t = 0;
    for (int k = 0; k < i; ++k) t += (rad(k, inv) - 1) * span(k, inv);
    return t;
  }
  template <int I, bool INV, int PITCH>
  using St = CStage<rad(I, INV), span(I, INV), N / rad(I, INV), twoff(I, INV), PITCH>;
};

template <int B, int E, class Fn>
__device__ __forceinline__ void static_for(Fn&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// Iteration state shared by all kernels of the loop; only control_kernel
// writes it (WaveSimBase.m:165-189, Simulation.m:193-217).
struct DevState {
  int it;           // current iteration, 1-based
  int calc_energy;  // refresh last_energy at this iteration
  int all_stopped;
  int nw;           // number of wiggles
  int nrhs;
  int interval;     // energy_calculation_interval
  int hist_cap;     // capacity of the per-rhs histories
  int nblk;         // energy partials per rhs
  long long max_iter;
  double threshold;
  int warm;           // warm start: W0 = gamma E0 is spread over the first Nw iterations
  int* stopped;        // [nrhs]
  int* converged;      // [nrhs]
  int* n_iter;         // [nrhs]
  double* last_energy; // [nrhs]
  double* diff;        // [nrhs][hist_cap]  diff_energy as the reference records it
  double* energy_all;  // [nrhs][hist_cap]  |crop dE'|^2 of every iteration
  double* ref_energy;  // [nrhs] stop-test denominator; <= 0: diff_energy(1) (Simulation.m:198)
  // PSTD time stepping (PSTD.m:79-127): the source amplitude A of the
  // current step in double and single precision, the expected phase advance
  // exp(i(angle A - angle A_prev)) of the step, and steps per optical cycle
  int pstd;
  double amp_re, amp_im;
  float2 amp, phase;
  double ipc;
  // host-visible copy of all_stopped (pinned, mapped host memory; null: none),
  // written by the control step so that the host can poll it while the next
  // batch of iterations is already queued
  int* host_stopped;
};

enum { MODE_FWD = 0, MODE_INV = 1, MODE_MID = 2 };
enum { PRO_NONE = 0, PRO_SRC_RAMP = 1 };
enum { EPI_NONE = 0, EPI_MIX = 1, EPI_RAMP = 2, EPI_MIX_FWD = 3, EPI_PSTD = 4, EPI_PSTD_FWD = 5 };

struct PassArgs {
  int N[3];  // Ny Nx Nz
  long long nvox;
  long long rhs_stride;  // C * nvox
  int axis;
  int n;
  int inner;      // product of N over the axes before `axis`
  int rstride;    // element stride along the axis (1 or inner)
  int ncol;       // columns per component in one workgroup
  int lds_pitch;  // contig: n + pad ; strided: C*ncol + pad  (in float2)
  int ntile_inner;
  FastDiv ftile;  // division by ntile_inner
  int nlines;     // contig: lines per component (nvox / n)
  int ntwf, ntwi; // twiddle table lengths (float2)
  AxisPlan plan;
  const float2* twf;
  const float2* twi;
  const float2* in;
  float2* out;
  const float2* S;  // premultiplied dense source (i/eps/Nw * S), or null
  const float2* W0; // warm start gamma*E0 per rhs, or null
  float2* dE;
  float2* E;
  const float2* gamma;
  double* partials;  // [rhs][blocks]
  const DevState* st;
  // iteration of this launch relative to the state's counter: solves advance
  // the counter once per wiggle cycle (one control step for Nw iterations),
  // and the launches of iteration j of a cycle carry it_off = j and their own
  // energy-partials slot
  int it_off;
  const float* pe[3];     // [nw][N[a]] scaled Fourier coordinates
  const float2* ramp[3];  // [nw][N[a]] wiggle phase ramps
  float2 k02e;
  float2 inv_k02e;
  float invN;
  // complex128 solves (Simulation.m:221-252 with single_precision = false):
  // the same constants in double; every field, table and twiddle pointer
  // above then addresses double2 / double data (typed float2 / float here)
  double2 k02e_d;
  double2 inv_k02e_d;
  double invN_d;
  int roi_lo[4];  // 0-based inclusive, [y x z pol]
  int roi_hi[4];
  // one launch may carry several problems (entries of an array of PassArgs):
  // entry e owns workgroups [grp_end[e-1], grp_end[e]) along x.  Read from
  // entry 0 only.
  int grp_n;           // entries in the launch (1: a single problem), <= 64
  const int* grp_end;  // [grp_n] exclusive prefix sums of the entries' widths
  int blocks;          // this entry's workgroups along x (row length of partials)
  int nrhs;            // this entry's rhs count; workgroups with y beyond it idle
  int reverse;         // walk the tiles from the end of memory
  int xcd;             // hand each XCD one contiguous range of tiles
  // several rhs, mix passes: the rhs index varies fastest along x (grid
  // 8 * ceil(blocks / 8) * nrhs by 1), so the workgroups of one tile's rhs
  // run on one XCD one after another and read the shared gamma tile from its
  // L2 instead of HBM
  int rhs_inner;
  // PSTD (PSTD.m:79-127): the two time levels E, E_prev alternate between
  // fld[0] and fld[1] (step k reads fld[(k-1)&1], writes fld[k&1]); the
  // middle pass multiplies by the Laplacian -|p|^2 (pe = p, unscaled)
  // instead of the Green's function; the last pass applies the update
  // c2 E + c1 E_prev + c3 (lap E + A S) with coef = [c1 | c2 | c3] per voxel,
  // S the unscaled source and srcline[l] != 0 for lines that carry source
  int pstd;
  float2* fld[2];
  const float* coef;
  const float2* Sp;
  const unsigned char* srcline;
};

// radix R is compiled for C-component fields when the R*C values of one
// butterfly fit the register budget
__host__ __device__ constexpr bool radix_fits(int R, int C) { return R * C <= 33; }

// butterflies per thread per stage: EPT registers' worth, at most kMaxIts
// (more unrolled iterations of the small radices only add register pressure)
constexpr int kMaxIts = 6;
__host__ __device__ constexpr int its_per_stage(int ept, int C, int R) {
  return ept / (C * R) < 1 ? 1 : (ept / (C * R) < kMaxIts ? ept / (C * R) : kMaxIts);
}

// F = float (complex64 fields) or double (complex128 fields)
template <class F>
struct ColInfoT {
  uint32_t off;    // element offset of row 0 from the workgroup's base (same for every component)
  vec_of<F> rc;    // product of the other axes' ramps
  vec_of<F> rc2;   // the same for the next iteration's wiggle (EPI_MIX_FWD)
  F p[3];          // other axes' coordinates (axis entry unused)
  F q;             // sum of squares of the other axes' coordinates
  int valid;
  int roi_ok;      // other axes inside the ROI
  int src;         // PSTD: the line carries source entries
};
using ColInfo = ColInfoT<float>;
// the host sizes the LDS column table with the same layout
static_assert(sizeof(ColInfoT<float>) == 56 && sizeof(ColInfoT<double>) == 96, "column table layout");

// the fields and tables of PassArgs as the kernel's element type
template <class T>
__device__ __forceinline__ const T* as(const void* p) { return reinterpret_cast<const T*>(p); }
template <class T>
__device__ __forceinline__ T* as_mut(void* p) { return reinterpret_cast<T*>(p); }

// 32-bit element offsets from a workgroup-uniform base pointer, so the
// compiler can use SGPR-base + 32-bit VGPR-offset addressing
// The field pointers are read from the argument block in device memory, so
// the compiler cannot infer that they address global memory and would emit
// flat instructions -- which also count against the LDS wait counter, making
// every LDS wait of a stage wait for the outstanding HBM traffic too.  Casting
// to the global address space gives global_load / global_store with a
// scalar base.
using gptr_c = const __attribute__((address_space(1))) char*;
using gptr = __attribute__((address_space(1))) char*;
using v2f = float __attribute__((ext_vector_type(2)));  // plain vector type: no constructors in address space 1
__device__ __forceinline__ float2 ldg(const float2* base, uint32_t i) {
  const v2f x = *reinterpret_cast<const __attribute__((address_space(1))) v2f*>((gptr_c)base + (size_t)(i * 8u));
  return make_float2(x.x, x.y);
}
__device__ __forceinline__ void stg(float2* base, uint32_t i, float2 v) {
  const v2f x = {v.x, v.y};
  *reinterpret_cast<__attribute__((address_space(1))) v2f*>((gptr)base + (size_t)(i * 8u)) = x;
}
// complex128: 16-byte elements (i * 16 stays below 2^32: a component holds
// at most 2^28 voxels in double)
using v2d = double __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ldg(const double2* base, uint32_t i) {
  const v2d x = *reinterpret_cast<const __attribute__((address_space(1))) v2d*>((gptr_c)base + (size_t)(i * 16u));
  return make_double2(x.x, x.y);
}
__device__ __forceinline__ void stg(double2* base, uint32_t i, double2 v) {
  const v2d x = {v.x, v.y};
  *reinterpret_cast<__attribute__((address_space(1))) v2d*>((gptr)base + (size_t)(i * 16u)) = x;
}
// A pointer the whole workgroup shares, moved to scalar registers.  Values
// computed with 64-bit multiplies land in vector registers even when every
// lane holds the same one, and a vector base costs each access a 64-bit
// address add; a scalar base lets global_load / global_store take it with
// the 32-bit lane offset directly.
template <typename T>
__device__ __forceinline__ T* wave_uniform(T* p) {
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (T*)(((uint64_t)hi << 32) | lo);
}
// one element of a table in global memory
__device__ __forceinline__ float ldt(const float* p) {
  return *reinterpret_cast<const __attribute__((address_space(1))) float*>((gptr_c)p);
}
__device__ __forceinline__ float2 ldt(const float2* p) { return ldg(p, 0); }
__device__ __forceinline__ double ldt(const double* p) {
  return *reinterpret_cast<const __attribute__((address_space(1))) double*>((gptr_c)p);
}
__device__ __forceinline__ double2 ldt(const double2* p) { return ldg(p, 0); }
__device__ __forceinline__ float ldf(const float* base, uint32_t i) {
  return *reinterpret_cast<const __attribute__((address_space(1))) float*>((gptr_c)base + (size_t)(i * 4u));
}

// LDS pointers carry their address space, so the index arithmetic of every
// LDS access is 32-bit (as a generic pointer it was widened to 64 bits, with a
// quarter-rate 64-bit multiply-add per strided row offset)
// (complex values are held as a plain two-float vector there: float2's
// constructors and assignments exist for generic pointers only)
#define WSIM_LDS __attribute__((address_space(3)))
using lf2 = WSIM_LDS v2f*;
using lcf2 = const WSIM_LDS v2f*;
using lcf = const WSIM_LDS float*;
__device__ __forceinline__ float2 ldl(lcf2 p, int i) {
  const v2f x = p[i];
  return make_float2(x.x, x.y);
}
__device__ __forceinline__ void stl(lf2 p, int i, float2 v) {
  const v2f x = {v.x, v.y};
  p[i] = x;
}
using ld2 = WSIM_LDS v2d*;
using lcd2 = const WSIM_LDS v2d*;
using lcd = const WSIM_LDS double*;
__device__ __forceinline__ double2 ldl(lcd2 p, int i) {
  const v2d x = p[i];
  return make_double2(x.x, x.y);
}
__device__ __forceinline__ void stl(ld2 p, int i, double2 v) {
  const v2d x = {v.x, v.y};
  p[i] = x;
}
// LDS pointer types by scalar type
template <class F> struct LdsOf;
template <> struct LdsOf<float> {
  using vec = lf2;
  using cvec = lcf2;
  using cscal = lcf;
  using scal = WSIM_LDS float*;
};
template <> struct LdsOf<double> {
  using vec = ld2;
  using cvec = lcd2;
  using cscal = lcd;
  using scal = WSIM_LDS double*;
};
// reciprocal of the Green's-function denominator: the hardware approximation
// in single precision (1 ulp), a correctly rounded division in double
__device__ __forceinline__ float rcp_green(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_green(double x) { return 1.0 / x; }

#ifndef WSIM_PAD_SHIFT  // contiguous lines in LDS: one padding slot every 2^WSIM_PAD_SHIFT rows
#define WSIM_PAD_SHIFT 3
#endif

#ifndef WSIM_EPI_CHUNK  // rows of one column whose mix operands are loaded together
#define WSIM_EPI_CHUNK 4
#endif
#ifndef WSIM_FUSED_CHUNK  // the same when the next forward pass is fused into the mix
#define WSIM_FUSED_CHUNK 4
#endif
#ifndef WSIM_FIXED_FUSED_CHUNK  // the same in plan-specialised kernels (4 rows spill there)
#define WSIM_FIXED_FUSED_CHUNK 2
#endif

// F: the scalar type of the field (float: complex64, double: complex128)
template <bool CONTIG, int MODE, int PRO, int EPI, int C, int NT, int EPT, class F = float, class K = DynPlan>
struct Pass {
  using V = vec_of<F>;
  using CI = ColInfoT<F>;
  using L2 = typename LdsOf<F>::vec;
  using LC2 = typename LdsOf<F>::cvec;
  using LCS = typename LdsOf<F>::cscal;
  static constexpr bool F64 = sizeof(F) == 8;
  const PassArgs& a;
  L2 lds;
  const CI* cols;
  int tid;
  bool add_src;
  bool warm_step;  // one of the first Nw steps of a warm start
  F inv_nw;
  uint32_t lg_ncol;
  LC2 ramp_a;  // this wiggle's ramp along the axis
  LC2 ramp2_a; // the next iteration's (EPI_MIX_FWD)
  bool add_src2;         // the next iteration adds the source (EPI_MIX_FWD)
  bool warm_step2;       // the next iteration adds the warm-start portion
  LCS pe_a;
  const V* in_b[C];  // workgroup bases, per component
  V* out_b[C];
  const V* S_b[C];
  const V* W0_b[C];
  V* dE_b[C];
  V* E_b[C];
  const V* gam_b;
  // PSTD epilogue: current and previous time level, coefficients, source
  const V* fcur_b[C];
  V* fprv_b[C];
  const V* Sp_b[C];
  const float* c1_b;
  const float* c2_b;
  const float* c3_b;
  V amp, phase;
  uint32_t nt;   // threads of the workgroup (NT is the compile-time maximum)
  CI cfix;  // strided passes: the thread's column never changes
  double eacc;
  // E' += dE' is carried out on even iterations only, as E' += dE'_{k-1} +
  // dE'_k: the mix reads dE'_{k-1} and forms dE'_k anyway, so E' is read and
  // written every other iteration.  (E' + dE'_{k-1}) + dE'_k rounds exactly
  // as the two separate updates; a solve that ends on an odd iteration adds
  // its last dE' afterwards (finish_e_kernel).
  bool acc_e;

  // tile geometry: compile-time constants of a fixed plan, else the launch's
  // values (lines per workgroup, LDS pitch, threads)
  static constexpr int kPitch =
      K::kFixed ? (CONTIG ? K::kN + (K::kN >> WSIM_PAD_SHIFT) + 1 : C * K::kNcol + 1) : 0;
  __device__ __forceinline__ int ncol() const {
    if constexpr (K::kFixed) return K::kNcol; else return a.ncol;
  }
  __device__ __forceinline__ int nlen() const {
    if constexpr (K::kFixed) return K::kN; else return a.n;
  }
  __device__ __forceinline__ int pitch() const {
    if constexpr (K::kFixed) return kPitch; else return a.lds_pitch;
  }
  __device__ __forceinline__ uint32_t nthr() const {
    if constexpr (K::kFixed) return (uint32_t)K::kNT; else return nt;
  }
  __device__ __forceinline__ uint32_t lgncol() const {
    if constexpr (K::kFixed) return (uint32_t)(31 - __builtin_clz((unsigned)K::kNcol)); else return lg_ncol;
  }
  // a fixed plan knows at compile time that every thread has a butterfly in
  // step `it` of a stage
  template <class ST>
  static constexpr bool full(int it) {
    if constexpr (K::kFixed) return (it + 1) * K::kNT <= K::kNcol * ST::nb; else return false;
  }

  // contig lines get one padding slot every 8 rows: the Stockham stores of
  // the early stages (Ns < 8) write rows R apart, which would otherwise land
  // on the same LDS banks
  // (24-bit products: LDS indices stay far below 2^24, and a 24-bit multiply
  // issues at full rate where a 32-bit one takes four cycles)
  __device__ __forceinline__ int lidx(int t, int c, int row) const {
    return CONTIG ? (int)__umul24((uint32_t)(c * ncol() + t), (uint32_t)pitch()) + row + (row >> WSIM_PAD_SHIFT)
                  : (int)__umul24((uint32_t)row, (uint32_t)pitch()) + c * ncol() + t;
  }

  // HBM element index of row `row` of a column whose row 0 is at `off`
  // (contiguous lines: unit stride)
  __device__ __forceinline__ uint32_t gidx(uint32_t off, int row) const {
    return CONTIG ? off + (uint32_t)row : off + (uint32_t)row * (uint32_t)a.rstride;
  }
  // the step of gidx for rows d apart (uniform)
  __device__ __forceinline__ uint32_t gstep(int d) const {
    return CONTIG ? (uint32_t)d : (uint32_t)d * (uint32_t)a.rstride;
  }
  // LDS index of rows row0 + r*d: contiguous lines pad every 8 rows, so their
  // index is computed per row; strided tiles are row-major, so the index of
  // row0 is computed once and each further row adds d * pitch (the row
  // products were the bulk of the integer work of a stage)
  // (dp = d * pitch, precomputed on the host: the compiler would otherwise
  // fold the products back into one multiply per row)
  __device__ __forceinline__ int lrow(int t, int c, int row0, int d, int dp, int r) const {
    return CONTIG ? lidx(t, c, row0 + r * d) : lidx(t, c, row0) + r * dp;
  }

  // butterfly b -> (column t, butterfly index j)
  template <class ST>
  __device__ __forceinline__ void decode(uint32_t b, const ST& s, int& t, int& j) const {
    if (CONTIG) {
      const uint32_t q = div_nb(s, b);
      j = (int)(b - q * (uint32_t)s.nb);
      t = (int)q;
    } else {
      t = (int)(b & (uint32_t)(ncol() - 1));
      j = (int)(b >> lgncol());
    }
  }

  __device__ __forceinline__ const CI& col(int t) const { return CONTIG ? cols[t] : cfix; }

  // the problem's constants in the field's precision
  __device__ __forceinline__ V k02e() const {
    if constexpr (F64) return a.k02e_d; else return a.k02e;
  }
  __device__ __forceinline__ V inv_k02e() const {
    if constexpr (F64) return a.inv_k02e_d; else return a.inv_k02e;
  }
  __device__ __forceinline__ F invN() const {
    if constexpr (F64) return a.invN_d; else return a.invN;
  }

  // The butterflies of a stage that this thread takes are b = tid + it*nt,
  // it = 0, 1, ...  Rather than dividing every b, a cursor starts from one
  // division and steps by nt = dt*nb + dj (contiguous lines) or by dj rows
  // (strided tiles: nt is a multiple of ncol, so the column never changes).
  // k = j mod Ns steps by dk = dj mod Ns (nb is a multiple of Ns, so the
  // wrap of j past nb leaves k alone); g is the HBM index of row j (strided).
  struct Cursor {
    int t, j, k;
    uint32_t g;
  };
  struct Walk {
    int dt, dj, dk;
    uint32_t dg;
  };
  template <class ST>
  __device__ __forceinline__ Walk walk(const ST& s) const {
    Walk w;
    if (CONTIG) {
      w.dt = (int)div_nb(s, nthr());
      w.dj = (int)nthr() - w.dt * s.nb;
    } else {
      w.dt = 0;
      w.dj = (int)(nthr() >> lgncol());
    }
    w.dk = w.dj - (int)div_Ns(s, (uint32_t)w.dj) * s.Ns;
    w.dg = (uint32_t)w.dj * (uint32_t)a.rstride;
    return w;
  }
  template <class ST>
  __device__ __forceinline__ Cursor start(const ST& s) const {
    Cursor c;
    decode((uint32_t)tid, s, c.t, c.j);
    c.k = c.j - (int)div_Ns(s, (uint32_t)c.j) * s.Ns;
    c.g = CONTIG ? 0u : cfix.off + (uint32_t)c.j * (uint32_t)a.rstride;
    return c;
  }
  // b < ncol * nb (b = tid + it*nt)
  template <class ST>
  __device__ __forceinline__ bool live(const Cursor& c, const ST& s, int it) const {
    if (full<ST>(it)) return true;
    return CONTIG ? c.t < ncol() : c.j < s.nb;
  }
  template <class ST>
  __device__ __forceinline__ void step(Cursor& c, const Walk& w, const ST& s) const {
    c.t += w.dt;
    c.j += w.dj;
    c.k += w.dk;
    c.g += w.dg;
    if (CONTIG && c.j >= s.nb) {
      c.j -= s.nb;
      ++c.t;
    }
    if (c.k >= s.Ns) c.k -= s.Ns;
  }
  // HBM index of row j of the cursor's column
  __device__ __forceinline__ uint32_t gbase(const Cursor& c, const CI& ci) const {
    return CONTIG ? ci.off + (uint32_t)c.j : c.g;
  }

  // ---- element transforms --------------------------------------------------
  // first-stage operand before the ramps: dE' (+ W0/Nw in the warm steps,
  // + i/eps/Nw * S in the source steps)
  __device__ __forceinline__ V load_raw(uint32_t i, int c) const {
    V v = ldg(in_b[c], i);
    if (PRO == PRO_SRC_RAMP) {
      if (warm_step) v = cadd(v, cscale(ldg(W0_b[c], i), inv_nw));
      if (add_src) v = cadd(v, ldg(S_b[c], i));
    }
    return v;
  }

  // scalar Green's function times 1/N (f_g0_scalar.m:6)
  __device__ __forceinline__ V green_scalar(const CI& ci, int row, V v) const {
    const F pa = pe_a[row];
    const F dr = ci.q + (pa * pa - k02e().x);
    const F di = -k02e().y;
    const F s = invN() * rcp_green(dr * dr + di * di);
    return cmul(v, mkv<V>(dr * s, -di * s));
  }

  // butterflies per thread in a stage of radix R: the register budget of the
  // kernel family, or exactly what a fixed plan's tile needs
  template <int R, class ST = Stage>
  static constexpr int its() {
    if constexpr (K::kFixed) return (K::kNcol * ST::nb + K::kNT - 1) / K::kNT;
    else return its_per_stage(EPT, C, R);
  }

  // ---- stage pieces (C components, radix R) --------------------------------
  // first stage from HBM.  Only loads here: a use of a loaded value inside
  // the guarded block would make every iteration wait for its own loads.
  template <int R, class ST>
  __device__ __forceinline__ void stage_load_global(V* v, const ST& s) const {
    constexpr int ITS = its<R, ST>();
    const Walk w = walk(s);
    Cursor cu = start(s);
    const uint32_t gs = gstep(s.nb);
#pragma unroll
    for (int it = 0; it < ITS; ++it) {
      const bool lv = live(cu, s, it);
      const CI& ci = col(lv ? cu.t : 0);
      const bool ok = lv && ci.valid;
      const uint32_t g0 = gbase(cu, ci);
#pragma unroll
      for (int c = 0; c < C; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[(c * ITS + it) * R + r] = ok ? load_raw(g0 + (uint32_t)r * gs, c) : mkv<V>(F(0), F(0));
      step(cu, w, s);
    }
    if (PRO == PRO_SRC_RAMP) {
      // wiggle ramps of the first forward pass (WaveSim.m:38-47)
      cu = start(s);
#pragma unroll
      for (int it = 0; it < ITS; ++it) {
        if (live(cu, s, it)) {
          const V rc = col(cu.t).rc;
#pragma unroll
          for (int c = 0; c < C; ++c)
#pragma unroll
            for (int r = 0; r < R; ++r)
              v[(c * ITS + it) * R + r] = cmul(v[(c * ITS + it) * R + r], cmul(ldl(ramp_a, cu.j + r * s.nb), rc));
        }
        step(cu, w, s);
      }
    }
  }

  // The first stage's loads issued before the column table is staged (fixed
  // plans without a prologue), so that they are in flight while the tables
  // are built: a line's offset and validity come from the tile geometry
  // (line0: the workgroup's first line; col_ok: the strided thread's column
  // lies inside the grid).  The same loads as stage_load_global.
  template <int R, class ST>
  __device__ __forceinline__ void stage_load_early(V* v, const ST& s, int line0, bool col_ok) const {
    constexpr int ITS = its<R, ST>();
    const Walk w = walk(s);
    Cursor cu = start(s);
    const uint32_t gs = gstep(s.nb);
#pragma unroll
    for (int it = 0; it < ITS; ++it) {
      const bool lv = live(cu, s, it);
      const bool ok = lv && (CONTIG ? line0 + cu.t < a.nlines : col_ok);
      const uint32_t g0 = CONTIG ? (uint32_t)cu.t * (uint32_t)nlen() + (uint32_t)cu.j : cu.g;
#pragma unroll
      for (int c = 0; c < C; ++c)
#pragma unroll
        for (int r = 0; r < R; ++r)
          v[(c * ITS + it) * R + r] = ok ? load_raw(g0 + (uint32_t)r * gs, c) : mkv<V>(F(0), F(0));
      step(cu, w, s);
    }
  }

  // A wave with at least one live butterfly reads without branching per lane:
  // lanes past the tile (the idle lanes of a partly filled last step) read
  // butterfly 0's operands, whose results are never stored, so the reads of
  // a butterfly issue together.  Waves with no live butterfly (a stage with
  // fewer butterflies than threads) skip the reads.
  template <int R, bool FIRST = false, class ST>
  __device__ __forceinline__ void stage_load_lds(V* v, const ST& s, LC2 tw) const {
    constexpr int ITS = its<R, ST>();
    const Walk w = walk(s);
    Cursor cu = start(s);
#pragma unroll
    for (int it = 0; it < ITS; ++it) {
      const bool lv = live(cu, s, it);
      if (full<ST>(it) || __builtin_amdgcn_ballot_w64(lv) != 0) {
        const int t = lv ? cu.t : 0, j = lv ? cu.j : 0, k = lv ? cu.k : 0;
        LC2 twk = tw + s.tw + k;  // table [r-1][k]: lanes with consecutive k read consecutive words
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const V x = ldl(lds, lrow(t, c, j, s.nb, s.nbp, r));
            v[(c * ITS + it) * R + r] = (r == 0 || FIRST) ? x : cmul(x, ldl(twk, (r - 1) * s.Ns));
          }
      } else {
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
          for (int r = 0; r < R; ++r) v[(c * ITS + it) * R + r] = mkv<V>(F(0), F(0));
      }
      step(cu, w, s);
    }
  }

  template <int R, int DIR, class ST = Stage>
  __device__ __forceinline__ void stage_compute(V* v) const {
    constexpr int ITS = its<R, ST>();
#pragma unroll
    for (int k = 0; k < C * ITS; ++k) dft<R, DIR>(v + k * R);
  }

  template <int R, class ST>
  __device__ __forceinline__ void stage_store_lds(const V* v, const ST& s) const {
    constexpr int ITS = its<R, ST>();
    const Walk w = walk(s);
    Cursor cu = start(s);
#pragma unroll
    for (int it = 0; it < ITS; ++it) {
      if (live(cu, s, it)) {
        const int base = (cu.j - cu.k) * R + cu.k;
#pragma unroll
        for (int c = 0; c < C; ++c)
#pragma unroll
          for (int r = 0; r < R; ++r) stl(lds, lrow(cu.t, c, base, s.Ns, s.nsp, r), v[(c * ITS + it) * R + r]);
      }
      step(cu, w, s);
    }
  }

  // last stage: Ns = n / R, outputs land on rows j + r*nb; E2 = the
  // epilogue applied on the way out (EPI, or EPI_NONE for the forward tail of
  // EPI_MIX_FWD)
  template <int R, int E2 = EPI, class ST>
  __device__ __forceinline__ void stage_store_global(const V* v, const ST& s) {
    constexpr int ITS = its<R, ST>();
    const Walk w = walk(s);
    Cursor cu = start(s);
    const uint32_t gs = gstep(s.nb);
#pragma unroll
    for (int it = 0; it < ITS; ++it, step(cu, w, s)) {
      if (live(cu, s, it)) {
        const int t = cu.t, j = cu.j;
        const CI& ci = col(t);
        if (!ci.valid) continue;
        if (E2 == EPI_MIX || E2 == EPI_MIX_FWD) {
          mix_rows<R, ST>(v, it, t, j, s.nb, ci);
          // a fixed plan's steps are branch-free: keep the scheduler from
          // overlapping one step's mix with the next step's operand loads,
          // which would hold more registers than the occupancy allows
          if constexpr (K::kFixed) __builtin_amdgcn_sched_barrier(0);
          continue;
        }
        if constexpr (E2 == EPI_PSTD || E2 == EPI_PSTD_FWD) {
          pstd_rows<R, ST>(v, it, t, j, s.nb, ci);
          continue;
        }
        const uint32_t g0 = gbase(cu, ci);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = j + r * s.nb;
          const uint32_t i = g0 + (uint32_t)r * gs;
          if (E2 == EPI_NONE) {
#pragma unroll
            for (int c = 0; c < C; ++c) stg(out_b[c], i, v[(c * ITS + it) * R + r]);
          } else {  // EPI_RAMP: undo the ramps only (propagation probe)
            const V ramp = cconj(cmul(ldl(ramp_a, row), ci.rc));
#pragma unroll
            for (int c = 0; c < C; ++c) stg(out_b[c], i, cmul(v[(c * ITS + it) * R + r], ramp));
          }
        }
      }
    }
  }

  // mix + accumulate of the rows j + r*nb of one column (f_mix.m:3-4 with
  // alpha = 0.7: dE' <- (1 - a g) dE' - (0.7 i g^2) Eprop; E' <- E' + dE').
  // The operands of a chunk of rows are loaded before any of its stores:
  // a load after a store to a possibly aliasing address would wait for the
  // previous row's whole load-compute-store chain.
  template <int R, class ST>
  __device__ __forceinline__ void mix_rows(const V* v, int it, int t, int j, int nb, const CI& ci) {
    constexpr int ITS = its<R, ST>();
    // rows per chunk: 4 for scalar fields (118 VGPRs in the fused variant,
    // four waves per SIMD; 2 rows ran the pass 4% slower), 1 per component
    // for vector fields
    constexpr int CH0 =
        C == 1 ? (EPI == EPI_MIX_FWD ? (K::kFixed ? WSIM_FIXED_FUSED_CHUNK : WSIM_FUSED_CHUNK) : WSIM_EPI_CHUNK) : 1;
    constexpr int CH = CH0 < R ? CH0 : R;
    const int rlo = a.roi_lo[a.axis], rhi = a.roi_hi[a.axis];
#pragma unroll
    for (int r0 = 0; r0 < R; r0 += CH) {
      V gm[CH], d0[CH][C], e0[CH][C], w0[CH][C];
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        if (r0 + q >= R) break;
        const uint32_t i = gidx(ci.off, j + (r0 + q) * nb);
        gm[q] = ldg(gam_b, i);
#pragma unroll
        for (int c = 0; c < C; ++c) {
          d0[q][c] = ldg(dE_b[c], i);
          e0[q][c] = acc_e ? ldg(E_b[c], i) : mkv<V>(F(0), F(0));
          w0[q][c] = warm_step ? cscale(ldg(W0_b[c], i), inv_nw) : mkv<V>(F(0), F(0));
        }
      }
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        if (r0 + q >= R) break;
        const int r = r0 + q;
        const int row = j + r * nb;
        const uint32_t i = gidx(ci.off, row);
        const V ramp = cconj(cmul(ldl(ramp_a, row), ci.rc));
        const V c1 = mkv<V>(F(1) - F(0.7) * gm[q].x, F(-0.7) * gm[q].y);
        const V g2 = cmul(gm[q], gm[q]);
        const V c2 = mkv<V>(F(-0.7) * g2.y, F(0.7) * g2.x);
        const bool in_roi = ci.roi_ok && row >= rlo && row <= rhi;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const V ep = cmul(v[(c * ITS + it) * R + r], ramp);
          // warm start (first Nw steps): the increment of the correction
          // W0/Nw, i.e. (M_w - 1) W0/Nw, joins the recurrence
          const V w = w0[q][c];
          const V dn = csub(csub(cmul(c1, cadd(d0[q][c], w)), cmul(c2, ep)), w);
          stg(dE_b[c], i, dn);
          if (acc_e) stg(E_b[c], i, cadd(cadd(e0[q][c], d0[q][c]), dn));
          if (in_roi && c >= a.roi_lo[3] && c <= a.roi_hi[3])
            eacc += (double)dn.x * (double)dn.x + (double)dn.y * (double)dn.y;
          if (EPI == EPI_MIX_FWD) {
            // the next iteration's first-pass operand, as load_raw and the
            // ramp step of the first forward pass would form it
            // (the source is read here, not with the other operands: it is
            // needed in the first Nw iterations only and would hold registers)
            V x = dn;
            if (warm_step2) x = cadd(x, w0[q][c]);
            if (add_src2) x = cadd(x, ldg(S_b[c], i));
            stl(lds, lidx(t, c, row), cmul(x, cmul(ldl(ramp2_a, row), ci.rc2)));
          }
        }
      }
    }
  }

  // PSTD update of the rows j + r*nb of one column (PSTD.m:108-116):
  // Etmp = lap E (+ A S on source lines); E_next = c2 E + c1 E_prev + c3 Etmp
  // into E_prev's buffer; energy partials of E_next - E exp(i dphi) in the ROI.
  // EPI_PSTD_FWD: E_next times the next step's ramps also goes to LDS, the
  // operand of the next step's first forward pass (as stage_load_global of
  // that pass would form it).
  // Operands of a chunk of rows are loaded before any store (as mix_rows).
  template <int R, class ST>
  __device__ __forceinline__ void pstd_rows(const V* v, int it, int t, int j, int nb, const CI& ci) {
    static_assert(!F64, "the PSTD step is built for complex64 fields only");
    constexpr int ITS = its<R, ST>();
    constexpr int CH = 4 < R ? 4 : R;
    const int rlo = a.roi_lo[a.axis], rhi = a.roi_hi[a.axis];
#pragma unroll
    for (int r0 = 0; r0 < R; r0 += CH) {
      float k1[CH], k2[CH], k3[CH];
      V ec[CH][C], ep[CH][C], sv[CH][C];
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        if (r0 + q >= R) break;
        const uint32_t i = gidx(ci.off, j + (r0 + q) * nb);
        k1[q] = ldf(c1_b, i);
        k2[q] = ldf(c2_b, i);
        k3[q] = ldf(c3_b, i);
#pragma unroll
        for (int c = 0; c < C; ++c) {
          ec[q][c] = ldg(fcur_b[c], i);
          ep[q][c] = ldg(fprv_b[c], i);
          sv[q][c] = ci.src ? ldg(Sp_b[c], i) : mkv<V>(F(0), F(0));
        }
      }
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        if (r0 + q >= R) break;
        const int row = j + (r0 + q) * nb;
        const uint32_t i = gidx(ci.off, row);
        const bool in_roi = ci.roi_ok && row >= rlo && row <= rhi;
#pragma unroll
        for (int c = 0; c < C; ++c) {
          V et = v[(c * ITS + it) * R + r0 + q];
          if (ci.src) et = cadd(et, cmul(amp, sv[q][c]));
          const V e = ec[q][c], em = ep[q][c];
          const V en = mkv<V>((k2[q] * e.x + k1[q] * em.x) + k3[q] * et.x,
                                        (k2[q] * e.y + k1[q] * em.y) + k3[q] * et.y);
          stg(fprv_b[c], i, en);
          if (EPI == EPI_PSTD_FWD) stl(lds, lidx(t, c, row), cmul(en, cmul(ldl(ramp2_a, row), ci.rc2)));
          if (in_roi && c >= a.roi_lo[3] && c <= a.roi_hi[3]) {
            const V d = csub(en, cmul(e, phase));
            eacc += (double)d.x * (double)d.x + (double)d.y * (double)d.y;
          }
        }
      }
    }
  }

  // Green's function on the forward last-stage outputs (rows j + r*nb)
  template <int R, class ST>
  __device__ __forceinline__ void stage_green(V* v, const ST& s) const {
    constexpr int ITS = its<R, ST>();
    const Walk w = walk(s);
    Cursor cu = start(s);
#pragma unroll
    for (int it = 0; it < ITS; ++it, step(cu, w, s)) {
      if (live(cu, s, it)) {
        const int j = cu.j;
        const CI& ci = col(cu.t);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = j + r * s.nb;
          if (C == 1) {
            if (a.pstd) {
              // PSTD Laplacian -|p|^2 times 1/N (PSTD.m:68, 108)
              const F pa = pe_a[row];
              v[it * R + r] = cscale(v[it * R + r], -(ci.q + pa * pa) * invN());
            } else {
              v[it * R + r] = green_scalar(ci, row, v[it * R + r]);
            }
          } else {
            // f_g0_vector.m:4-8; component 0 <-> x <-> array axis 1
            const F pr = pe_a[row];
            const F py = a.axis == 0 ? pr : ci.p[0];
            const F px = a.axis == 1 ? pr : ci.p[1];
            const F pz = a.axis == 2 ? pr : ci.p[2];
            const V fx = v[(0 * ITS + it) * R + r];
            const V fy = v[(1 * ITS + it) * R + r];
            const V fz = v[(2 * ITS + it) * R + r];
            V dv = mkv<V>(px * fx.x + py * fy.x + pz * fz.x, px * fx.y + py * fy.y + pz * fz.y);
            dv = cmul(dv, inv_k02e());
            const F dr = px * px + py * py + pz * pz - k02e().x;
            const F di = -k02e().y;
            const F sc = invN() * rcp_green(dr * dr + di * di);
            const V g0 = mkv<V>(dr * sc, -di * sc);
            v[(0 * ITS + it) * R + r] = cmul(g0, csub(fx, cscale(dv, px)));
            v[(1 * ITS + it) * R + r] = cmul(g0, csub(fy, cscale(dv, py)));
            v[(2 * ITS + it) * R + r] = cmul(g0, csub(fz, cscale(dv, pz)));
          }
        }
      }
    }
  }
};

#ifdef WSIM_ONLY_RADIX  // experiments: compile a single radix path
#define WSIM_RADIX_SWITCH(RV, ...) \
  { constexpr int R = WSIM_ONLY_RADIX; __VA_ARGS__ }
#else
// radices whose register tile a C-component field can hold (16 is left out
// for 3-component fields: 48 values per thread spill)
#define WSIM_RADIX_CASE(RR, ...) \
  case RR: { constexpr int R = RR; if constexpr (radix_fits(RR, C)) { __VA_ARGS__ } } break;
#define WSIM_RADIX_SWITCH(RV, ...)          \
  switch (RV) {                             \
    WSIM_RADIX_CASE(2, __VA_ARGS__)         \
    WSIM_RADIX_CASE(3, __VA_ARGS__)         \
    WSIM_RADIX_CASE(4, __VA_ARGS__)         \
    WSIM_RADIX_CASE(5, __VA_ARGS__)         \
    WSIM_RADIX_CASE(7, __VA_ARGS__)         \
    WSIM_RADIX_CASE(8, __VA_ARGS__)         \
    WSIM_RADIX_CASE(11, __VA_ARGS__)        \
    WSIM_RADIX_CASE(16, __VA_ARGS__)        \
    default: break;                         \
  }
#endif

template <bool CONTIG, int MODE, int PRO, int EPI, int C, int NT, int EPT, class F = float, class K = DynPlan>
__global__ void __launch_bounds__(NT, K::kWaves) axis_pass_kernel(const PassArgs* __restrict__ ap) {
  using P = Pass<CONTIG, MODE, PRO, EPI, C, NT, EPT, F, K>;
  using V = typename P::V;
  using CI = typename P::CI;
  using L2 = typename P::L2;
  using LC2 = typename P::LC2;
  // the arguments live in device memory (read through the scalar cache):
  // indexing the stage table of a by-value kernel argument would copy the
  // whole struct to scratch
  int bx = (int)blockIdx.x;
  if (ap->grp_n > 1) {
    // grouped launch: one lane per entry compares the block index with the
    // entry's end; the count of passed ends is this workgroup's entry
    const int lane = (int)(threadIdx.x & 63);
    const int end = lane < ap->grp_n ? ap->grp_end[lane] : 0x7fffffff;
    const int e = (int)__popcll(__ballot(bx >= end));
    if (e > 0) bx -= ap->grp_end[e - 1];
    ap += e;
  }
  const PassArgs& a = *ap;
  if constexpr (K::kFixed) {
    // the host launches a plan-specialised kernel only for its own shape
    if (a.n != K::kN || a.ncol != K::kNcol || (int)blockDim.x != K::kNT || a.plan.nst != K::S) return;
  }
  int rhs = blockIdx.y;
  if (a.rhs_inner) {
    // XCD x runs workgroups x, x+8, x+16, ... in order: its k-th takes rhs
    // k mod nrhs of its (k / nrhs)-th tile
    if (blockIdx.y != 0) return;
    const int x = bx & 7, k = bx >> 3, j = k / a.nrhs;
    rhs = k - j * a.nrhs;
    if (a.xcd) {
      const int q = a.blocks >> 3, r = a.blocks & 7;
      if (j >= q + (x < r ? 1 : 0)) return;
      bx = x * q + (x < r ? x : r) + j;
    } else {
      bx = j * 8 + x;
      if (bx >= a.blocks) return;
    }
  } else if (a.xcd) {
    // workgroups go to the 8 XCDs round-robin by index: give XCD x the
    // contiguous block range [x*q + min(x, r), +q + (x < r)) (a bijection)
    const int q = a.blocks >> 3, r = a.blocks & 7, x = bx & 7, j = bx >> 3;
    bx = x * q + (x < r ? x : r) + j;
  }
  if (a.reverse) bx = a.blocks - 1 - bx;
  extern __shared__ __align__(16) unsigned char smem[];
  if (rhs >= a.nrhs) return;
  const DevState* st = a.st;
  if (st->stopped[rhs]) return;

  // LDS: [CI x ncol][ramp_a x n][(ramp2 x n)][pe_a x n][twiddles][transform data]
  constexpr bool FUSE = EPI == EPI_MIX_FWD || EPI == EPI_PSTD_FWD;
  CI* cols = reinterpret_cast<CI*>(smem);
  const size_t rn = ((size_t)a.n * sizeof(V) + 15) & ~(size_t)15;
  const size_t off_ramp = (a.ncol * sizeof(CI) + 15) & ~(size_t)15;
  const size_t off_ramp2 = off_ramp + rn;
  const size_t off_pe = off_ramp + (FUSE ? 2 : 1) * rn;
  const int ntf = (MODE == MODE_INV && !FUSE) ? 0 : a.ntwf;  // forward twiddles used by this pass
  const int nti = MODE == MODE_FWD ? 0 : a.ntwi;  // inverse twiddles used by this pass
  const size_t off_twf = off_pe + (((size_t)a.n * sizeof(F) + 15) & ~(size_t)15);
  const size_t off_twi = off_twf + (size_t)ntf * sizeof(V);
  const size_t off_data = off_twi + (size_t)nti * sizeof(V);
  L2 ramp_l = (L2)(smem + off_ramp);
  L2 ramp2_l = (L2)(smem + off_ramp2);
  typename LdsOf<F>::scal pe_l = (typename LdsOf<F>::scal)(smem + off_pe);
  L2 twf_l = (L2)(smem + off_twf);
  L2 twi_l = (L2)(smem + off_twi);
  L2 lds = (L2)(smem + off_data);
  const int nthr = (int)blockDim.x;  // 256 or 512 (<= NT), chosen per pass by the plan
  const int it_cur = st->it + a.it_off;
  const int nw = st->nw;
  const int w = (it_cur - 1) % nw;
  const int w2 = it_cur % nw;  // the next iteration's wiggle (EPI_MIX_FWD)

  // workgroup geometry: voxel base and the decomposition of block index
  long long vblk;
  int o = 0, itile = 0;
  if (CONTIG) {
    vblk = (long long)bx * a.ncol * a.n;
  } else {
    o = (int)fdiv((uint32_t)bx, a.ftile);
    itile = bx - o * a.ntile_inner;
    vblk = (long long)o * a.n * a.inner + (long long)itile * a.ncol;
  }

  P p{a, lds, cols, (int)threadIdx.x, a.S != nullptr && it_cur <= nw,
      st->warm != 0 && a.W0 != nullptr && it_cur <= nw, F(1) / F(nw), 0u, ramp_l, ramp2_l,
      FUSE && a.S != nullptr && it_cur + 1 <= nw, FUSE && st->warm != 0 && a.W0 != nullptr && it_cur + 1 <= nw,
      pe_l};
  p.eacc = 0.0;
  p.acc_e = (it_cur & 1) == 0;
  p.nt = (uint32_t)nthr;
  if (!CONTIG) p.lg_ncol = 31 - __builtin_clz((unsigned)a.ncol);
  const long long rbase = (long long)rhs * a.rhs_stride + vblk;
  // PSTD: the first pass transforms the current time level
  const V* in_sel = as<V>((PRO == PRO_SRC_RAMP && a.pstd) ? a.fld[(it_cur - 1) & 1] : a.in);
  constexpr bool PSTD_STEP = EPI == EPI_PSTD || EPI == EPI_PSTD_FWD;
  if constexpr (PSTD_STEP) {
    p.c1_b = wave_uniform(a.coef + vblk);
    p.c2_b = wave_uniform(a.coef + a.nvox + vblk);
    p.c3_b = wave_uniform(a.coef + 2 * a.nvox + vblk);
    p.amp = st->amp;
    p.phase = st->phase;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const long long cb = rbase + (long long)c * a.nvox;
    if constexpr (PSTD_STEP) {
      p.fcur_b[c] = wave_uniform(a.fld[(it_cur - 1) & 1] + cb);
      p.fprv_b[c] = wave_uniform(a.fld[it_cur & 1] + cb);
      p.Sp_b[c] = wave_uniform(a.Sp + cb);
    }
    p.in_b[c] = wave_uniform(in_sel + cb);
    p.out_b[c] = wave_uniform(as_mut<V>(a.out) + cb);
    p.S_b[c] = wave_uniform(as<V>(a.S) + cb);
    p.W0_b[c] = wave_uniform(as<V>(a.W0) + cb);
    p.dE_b[c] = wave_uniform(as_mut<V>(a.dE) + cb);
    p.E_b[c] = wave_uniform(as_mut<V>(a.E) + cb);
  }
  p.gam_b = wave_uniform(as<V>(a.gamma) + vblk);

  // the LDS tables: twiddles, this axis's ramps and coordinates, and the
  // per-column coordinates (axes other than a.axis)
  auto setup_tables = [&]() {
    for (int i = threadIdx.x; i < ntf; i += nthr) stl(twf_l, i, ldt(as<V>(a.twf) + i));
    for (int i = threadIdx.x; i < nti; i += nthr) stl(twi_l, i, ldt(as<V>(a.twi) + i));
    for (int r = threadIdx.x; r < a.n; r += nthr) {
      stl(ramp_l, r, ldt(as<V>(a.ramp[a.axis]) + (size_t)w * a.n + r));
      if (FUSE) stl(ramp2_l, r, ldt(as<V>(a.ramp[a.axis]) + (size_t)w2 * a.n + r));
      pe_l[r] = ldt(as<F>(a.pe[a.axis]) + (size_t)w * a.n + r);
    }
    // per-column coordinates (axes other than a.axis)
    for (int t = threadIdx.x; t < a.ncol; t += nthr) {
      CI ci;
      int crd[3] = {0, 0, 0};
      int valid;
      if (CONTIG) {
        const int line = bx * a.ncol + t;  // < nvox < 2^31
        valid = line < a.nlines;
        ci.off = (uint32_t)(t * a.n);
        int rem = valid ? line : 0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
          if (b > a.axis) { crd[b] = rem % a.N[b]; rem /= a.N[b]; }
      } else {
        const int i = itile * a.ncol + t;
        valid = i < a.inner;
        ci.off = (uint32_t)t;
        int rem = valid ? i : 0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
          if (b < a.axis) { crd[b] = rem % a.N[b]; rem /= a.N[b]; }
        rem = o;
#pragma unroll
        for (int b = 0; b < 3; ++b)
          if (b > a.axis) { crd[b] = rem % a.N[b]; rem /= a.N[b]; }
      }
      ci.valid = valid;
      ci.src = 0;
      if (CONTIG && a.srcline != nullptr && valid) ci.src = a.srcline[bx * a.ncol + t];
      V rc = mkv<V>(F(1), F(0)), rc2 = mkv<V>(F(1), F(0));
      F q = F(0);
      int roi_ok = 1;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        ci.p[b] = F(0);
        if (b == a.axis) continue;
        const F pb = ldt(as<F>(a.pe[b]) + (size_t)w * a.N[b] + crd[b]);
        ci.p[b] = pb;
        q += pb * pb;
        rc = cmul(rc, ldt(as<V>(a.ramp[b]) + (size_t)w * a.N[b] + crd[b]));
        if (FUSE) rc2 = cmul(rc2, ldt(as<V>(a.ramp[b]) + (size_t)w2 * a.N[b] + crd[b]));
        roi_ok &= (crd[b] >= a.roi_lo[b]) & (crd[b] <= a.roi_hi[b]);
      }
      ci.rc = rc;
      ci.rc2 = rc2;
      ci.q = q;
      ci.roi_ok = roi_ok;
      cols[t] = ci;
    }
  };

  constexpr bool MID = MODE == MODE_MID;
  constexpr int D0 = (MODE == MODE_INV) ? 1 : -1;
  if constexpr (K::kFixed) {
    // the plan's stages as compile-time constants: the same stage pieces
    // and order as the run-time sequence below
    constexpr int S = K::S;
    constexpr bool FI = MODE == MODE_INV;  // the first transform runs the inverse stages
    constexpr int PT = P::kPitch;
    LC2 tw = FI ? twi_l : twf_l;
    using S0 = typename K::template St<0, FI, PT>;
    V v0[P::template its<S0::R, S0>() * C * S0::R];
    constexpr bool EARLY = PRO == PRO_NONE && K::kEarly;
    if constexpr (EARLY) {
      const int t0 = (int)(threadIdx.x & (uint32_t)(K::kNcol - 1));
      if (!CONTIG) p.cfix.off = (uint32_t)t0;
      p.template stage_load_early<S0::R>(v0, S0{}, CONTIG ? bx * K::kNcol : 0,
                                         !CONTIG && itile * K::kNcol + t0 < a.inner);
    }
    setup_tables();
    __syncthreads();
    if (!CONTIG) p.cfix = cols[threadIdx.x & (uint32_t)(K::kNcol - 1)];
    if constexpr (!EARLY) p.template stage_load_global<S0::R>(v0, S0{});
    if constexpr (S == 1) {
      constexpr int R = S0::R;
      V* v = v0;
      p.template stage_compute<R, D0, S0>(v);
      if constexpr (MID) {
        using I0 = typename K::template St<0, true, PT>;
        p.template stage_green<R>(v, S0{});
        p.template stage_compute<R, 1, S0>(v);
        p.template stage_store_global<R>(v, I0{});
      } else {
        p.template stage_store_global<R>(v, S0{});
      }
    } else {
      {
        constexpr int R = S0::R;
        V* v = v0;
        p.template stage_compute<R, D0, S0>(v);
        p.template stage_store_lds<R>(v, S0{});
      }
      __syncthreads();
      static_for<1, S - 1>([&](auto I) {
        using SI = typename K::template St<decltype(I)::value, FI, PT>;
        constexpr int R = SI::R;
        V v[P::template its<R, SI>() * C * R];
        p.template stage_load_lds<R>(v, SI{}, tw);
        __syncthreads();
        p.template stage_compute<R, D0, SI>(v);
        p.template stage_store_lds<R>(v, SI{});
        __syncthreads();
      });
      using SL = typename K::template St<S - 1, FI, PT>;
      if constexpr (!MID) {
        {
          constexpr int R = SL::R;
          V v[P::template its<R, SL>() * C * R];
          p.template stage_load_lds<R>(v, SL{}, tw);
          if (FUSE) __syncthreads();  // the epilogue writes the next pass's operand to LDS
          p.template stage_compute<R, D0, SL>(v);
          p.template stage_store_global<R>(v, SL{});
        }
        if constexpr (FUSE) {
          __syncthreads();
          {
            using F0 = typename K::template St<0, false, PT>;
            constexpr int R = F0::R;
            V v[P::template its<R, F0>() * C * R];
            p.template stage_load_lds<R, true>(v, F0{}, twf_l);
            __syncthreads();
            p.template stage_compute<R, -1, F0>(v);
            p.template stage_store_lds<R>(v, F0{});
          }
          __syncthreads();
          static_for<1, S - 1>([&](auto I) {
            using FI_ = typename K::template St<decltype(I)::value, false, PT>;
            constexpr int R = FI_::R;
            V v[P::template its<R, FI_>() * C * R];
            p.template stage_load_lds<R>(v, FI_{}, twf_l);
            __syncthreads();
            p.template stage_compute<R, -1, FI_>(v);
            p.template stage_store_lds<R>(v, FI_{});
            __syncthreads();
          });
          using FL = typename K::template St<S - 1, false, PT>;
          constexpr int R = FL::R;
          V v[P::template its<R, FL>() * C * R];
          p.template stage_load_lds<R>(v, FL{}, twf_l);
          p.template stage_compute<R, -1, FL>(v);
          p.template stage_store_global<R, EPI_NONE>(v, FL{});
        }
      } else {
        {
          using I0 = typename K::template St<0, true, PT>;
          constexpr int R = SL::R;
          V v[P::template its<R, SL>() * C * R];
          p.template stage_load_lds<R>(v, SL{}, tw);
          __syncthreads();
          p.template stage_compute<R, -1, SL>(v);
          p.template stage_green<R>(v, SL{});
          p.template stage_compute<R, 1, SL>(v);
          p.template stage_store_lds<R>(v, I0{});
        }
        __syncthreads();
        static_for<1, S - 1>([&](auto I) {
          using II = typename K::template St<decltype(I)::value, true, PT>;
          constexpr int R = II::R;
          V v[P::template its<R, II>() * C * R];
          p.template stage_load_lds<R>(v, II{}, twi_l);
          __syncthreads();
          p.template stage_compute<R, 1, II>(v);
          p.template stage_store_lds<R>(v, II{});
          __syncthreads();
        });
        using IL = typename K::template St<S - 1, true, PT>;
        constexpr int R = IL::R;
        V v[P::template its<R, IL>() * C * R];
        p.template stage_load_lds<R>(v, IL{}, twi_l);
        p.template stage_compute<R, 1, IL>(v);
        p.template stage_store_global<R>(v, IL{});
      }
    }
  } else {
    setup_tables();
    __syncthreads();
    if (!CONTIG) p.cfix = cols[threadIdx.x & (a.ncol - 1)];
    const AxisPlan& pl = a.plan;
    const int S = pl.nst;
    const Stage* fs = (MODE == MODE_INV) ? pl.i : pl.f;
    LC2 tw = (MODE == MODE_INV) ? twi_l : twf_l;

    // Each region below is its own radix switch, so the register allocator
    // sees one short straight-line path per radix.
    if (S == 1) {
      // single stage: HBM -> registers -> HBM (the middle pass applies the
      // Green's function and the inverse DFT in registers)
      WSIM_RADIX_SWITCH(fs[0].R,
        V v[P::template its<R>() * C * R];
        p.template stage_load_global<R>(v, fs[0]);
        p.template stage_compute<R, D0>(v);
        if (MID) {
          p.template stage_green<R>(v, fs[0]);
          p.template stage_compute<R, 1>(v);
          p.template stage_store_global<R>(v, pl.i[0]);
        } else {
          p.template stage_store_global<R>(v, fs[0]);
        }
      )
    } else {
      // first stage: HBM -> registers -> LDS
      WSIM_RADIX_SWITCH(fs[0].R,
        V v[P::template its<R>() * C * R];
        p.template stage_load_global<R>(v, fs[0]);
        p.template stage_compute<R, D0>(v);
        p.template stage_store_lds<R>(v, fs[0]);
      )
      __syncthreads();
      // middle stages: LDS -> registers -> LDS
      for (int s = 1; s < S - 1; ++s) {
        const Stage st_ = fs[s];
        WSIM_RADIX_SWITCH(st_.R,
          V v[P::template its<R>() * C * R];
          p.template stage_load_lds<R>(v, st_, tw);
          __syncthreads();
          p.template stage_compute<R, D0>(v);
          p.template stage_store_lds<R>(v, st_);
        )
        __syncthreads();
      }
      const Stage sl = fs[S - 1];
      if (!MID) {
        // last stage: LDS -> registers -> HBM
        WSIM_RADIX_SWITCH(sl.R,
          V v[P::template its<R>() * C * R];
          p.template stage_load_lds<R>(v, sl, tw);
          if (FUSE) __syncthreads();  // the epilogue writes the next pass's operand to LDS
          p.template stage_compute<R, D0>(v);
          p.template stage_store_global<R>(v, sl);
        )
        if (FUSE) {
          // the next iteration's forward transform of the same lines: LDS ->
          // forward stages -> work field (the plan has at least two stages)
          __syncthreads();
          const Stage* ff = pl.f;
          WSIM_RADIX_SWITCH(ff[0].R,
            V v[P::template its<R>() * C * R];
            p.template stage_load_lds<R, true>(v, ff[0], twf_l);
            __syncthreads();
            p.template stage_compute<R, -1>(v);
            p.template stage_store_lds<R>(v, ff[0]);
          )
          __syncthreads();
          for (int s = 1; s < S - 1; ++s) {
            const Stage st_ = ff[s];
            WSIM_RADIX_SWITCH(st_.R,
              V v[P::template its<R>() * C * R];
              p.template stage_load_lds<R>(v, st_, twf_l);
              __syncthreads();
              p.template stage_compute<R, -1>(v);
              p.template stage_store_lds<R>(v, st_);
            )
            __syncthreads();
          }
          const Stage fl = ff[S - 1];
          WSIM_RADIX_SWITCH(fl.R,
            V v[P::template its<R>() * C * R];
            p.template stage_load_lds<R>(v, fl, twf_l);
            p.template stage_compute<R, -1>(v);
            p.template stage_store_global<R, EPI_NONE>(v, fl);
          )
        }
      } else {
        // forward last stage, Green's function, inverse first stage (same
        // radix, Ns = 1, no twiddles), back to LDS
        WSIM_RADIX_SWITCH(sl.R,
          V v[P::template its<R>() * C * R];
          p.template stage_load_lds<R>(v, sl, tw);
          __syncthreads();
          p.template stage_compute<R, -1>(v);
          p.template stage_green<R>(v, sl);
          p.template stage_compute<R, 1>(v);
          p.template stage_store_lds<R>(v, pl.i[0]);
        )
        __syncthreads();
        for (int s = 1; s < S - 1; ++s) {
          const Stage st_ = pl.i[s];
          WSIM_RADIX_SWITCH(st_.R,
            V v[P::template its<R>() * C * R];
            p.template stage_load_lds<R>(v, st_, twi_l);
            __syncthreads();
            p.template stage_compute<R, 1>(v);
            p.template stage_store_lds<R>(v, st_);
          )
          __syncthreads();
        }
        const Stage il = pl.i[S - 1];
        WSIM_RADIX_SWITCH(il.R,
          V v[P::template its<R>() * C * R];
          p.template stage_load_lds<R>(v, il, twi_l);
          p.template stage_compute<R, 1>(v);
          p.template stage_store_global<R>(v, il);
        )
      }
    }
  }

  if (EPI == EPI_MIX || EPI == EPI_MIX_FWD || EPI == EPI_PSTD || EPI == EPI_PSTD_FWD) {
    // block reduction of the ROI energy (fixed order: deterministic)
    __shared__ double red[NT / 64];
    double e = p.eacc;
    for (int off = 32; off > 0; off >>= 1) e += __shfl_down(e, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = 0.0;
      for (int i = 0; i < nthr / 64; ++i) sum += red[i];
      a.partials[(size_t)rhs * a.blocks + bx] = sum;
    }
  }
}

}  // namespace wsim
