// Host-side checks run under AddressSanitizer and UndefinedBehaviorSanitizer
// (built by tests/mex_stub/build_shim.py, run by tests/test_host_sanitized.py):
// the problem set-up (wsim_problem_query / wsim_problem_build) on several
// shapes and on invalid input, and the MEX shim's marshalling through the fake
// MATLAB API with an in-process stand-in for wsim_run_algorithm (no GPU):
// output shapes and values, callback forwarding, and every argument check.
// Exit code 0 when every check holds; a sanitizer report aborts.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"
#include "wavesim_mbs.h"

namespace wsim {
static std::string g_msg;
void set_last_error(const std::string& m) { g_msg = m; }
}  // namespace wsim

extern "C" {
const char* wsim_last_error(void) { return wsim::g_msg.c_str(); }

// stand-in for the GPU loop: E' = pattern, diff_energy(k) = 1/k for 23
// iterations (or the capacity), the callback after every multiple of the
// interval and after the last iteration, as the library calls it
int wsim_run_algorithm(const double N[4], const wsim_cfloat*, int, const wsim_wiggle*, int,
                       const wsim_source_patch* patches, float, wsim_cfloat, const double*, double, double thr,
                       const char* cbtype, double cbint, wsim_callback cb, void* user, wsim_cfloat* E_out,
                       double* de, int64_t cap, int64_t* n) {
  const size_t nv = (size_t)(N[0] * N[1] * N[2] * N[3]);
  for (size_t i = 0; i < nv; ++i) E_out[i] = wsim_cfloat{(float)i * 1e-3f, -(float)i * 2e-3f};
  // the first source value shows that the patches arrived
  if (patches && patches[0].values) E_out[0] = patches[0].values[0];
  const int64_t it = std::min<int64_t>(23, cap);
  for (int64_t k = 0; k < it; ++k) de[k] = 1.0 / (double)(k + 1);
  *n = it;
  if (cb) {
    const int64_t ci = cbint >= 1 ? (int64_t)cbint : it;
    for (int64_t k = ci; k <= it; k += ci)
      if (cb(cbtype, k, de, k, thr, user)) return WSIM_ERR_ABORTED;
    if (it % ci != 0 && cb(cbtype, it, de, it, thr, user)) return WSIM_ERR_ABORTED;
  }
  return WSIM_OK;
}

mxArray* mxf_numeric(int cls, int cplx, int ndim, const int64_t* dims, const void* data);
mxArray* mxf_string(const char* s);
mxArray* mxf_cell(int64_t n);
void mxf_cell_set(mxArray* c, int64_t i, mxArray* v);
mxArray* mxf_struct(void);
mxArray* mxf_object(void);
void mxf_set(mxArray* s, const char* name, mxArray* v);
void mxf_destroy(mxArray* a);
int64_t mxf_numel(const mxArray* a);
int mxf_ndims(const mxArray* a);
int64_t mxf_dim(const mxArray* a, int k);
const void* mxf_data(const mxArray* a);
int64_t mxf_live_temporaries(void);
int mxf_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs, char* err, int errcap);
void mxf_log_clear(void);
int mxf_log_count(void);
int mxf_log_entry(int i, char* name, int ncap, char* type, int tcap, double* de, int64_t cap, int64_t* n_de,
                  int64_t* n_energy, double* thr, int* nargs);
void mxf_fail_callback_at(int k);
}

static int g_fail = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      ++g_fail;                                                    \
    }                                                              \
  } while (0)

static mxArray* dbl(std::vector<int64_t> dims, std::vector<double> v) {
  return mxf_numeric(0, 0, (int)dims.size(), dims.data(), v.data());
}
static mxArray* sgl(std::vector<int64_t> dims, const std::vector<float>& v, bool cplx) {
  return mxf_numeric(1, cplx ? 1 : 0, (int)dims.size(), dims.data(), v.data());
}

// ---------------------------------------------------------------- problem set-up
static void check_problem_setup() {
  struct Case {
    int64_t shape[3];
    int64_t bw[3];
    int vector;
    int roi;
  } cases[] = {{{24, 20, 1}, {6, 6, 0}, 0, 0}, {{10, 12, 8}, {3, 3, 3}, 1, 0}, {{16, 18, 1}, {4, 0, 0}, 0, 0},
               {{9, 7, 5}, {2, 3, 0}, 0, 1}, {{40, 1, 1}, {8, 0, 0}, 0, 0}};
  for (const Case& c : cases) {
    wsim_problem_desc d{};
    for (int a = 0; a < 3; ++a) {
      d.shape[a] = c.shape[a];
      d.boundary_widths[a] = c.bw[a];
      d.acc[a] = -1;
    }
    d.pixel_size = 0.25;
    d.vector = c.vector;
    if (c.roi) {
      d.has_roi = 1;
      for (int a = 0; a < 4; ++a) {
        d.roi[0][a] = 1;
        d.roi[1][a] = a < 3 ? std::max<int64_t>(1, c.shape[a] - 1) : 1;
      }
    }
    wsim_problem_info info{};
    CHECK(wsim_problem_query(&d, &info) == WSIM_OK);
    const size_t nv = (size_t)(info.N[0] * info.N[1] * info.N[2]);
    std::vector<wsim_cdouble> n((size_t)(c.shape[0] * c.shape[1] * c.shape[2]));
    for (size_t i = 0; i < n.size(); ++i) n[i] = wsim_cdouble{1.0 + 0.4 * std::sin(0.37 * (double)i), 0.01};
    std::vector<wsim_cfloat> gamma(nv);
    std::vector<std::vector<float>> pe(3 * (size_t)info.nwiggles);
    std::vector<std::vector<wsim_cfloat>> ramp(3 * (size_t)info.nwiggles);
    std::vector<wsim_wiggle_out> wo((size_t)info.nwiggles);
    const int64_t len[3] = {info.N[1], info.N[0], info.N[2]};  // x, y, z
    for (int w = 0; w < info.nwiggles; ++w) {
      for (int k = 0; k < 3; ++k) {
        pe[3 * w + k].resize((size_t)len[k]);
        ramp[3 * w + k].resize((size_t)len[k]);
      }
      wo[w] = wsim_wiggle_out{pe[3 * w].data(), pe[3 * w + 1].data(), pe[3 * w + 2].data(), ramp[3 * w].data(),
                              ramp[3 * w + 1].data(), ramp[3 * w + 2].data()};
    }
    CHECK(wsim_problem_build(&d, n.data(), &info, gamma.data(), wo.data()) == WSIM_OK);
    bool finite = true;
    for (const auto& g : gamma) finite = finite && std::isfinite(g.re) && std::isfinite(g.im);
    CHECK(finite);
    CHECK(info.epsilon >= 3.0f && std::isfinite(info.k02e.re));
    for (const auto& r : ramp)
      for (const auto& v : r) CHECK(std::fabs(std::hypot(v.re, v.im) - 1.0f) < 1e-5f);
  }
  // invalid input: gain, a zero-size map, an ROI outside the map, null pointers
  wsim_problem_desc d{};
  d.shape[0] = 8;
  d.shape[1] = 8;
  d.shape[2] = 1;
  d.boundary_widths[0] = d.boundary_widths[1] = 2;
  d.pixel_size = 0.25;
  d.acc[0] = d.acc[1] = d.acc[2] = -1;
  wsim_problem_info info{};
  std::vector<wsim_cdouble> n(64, wsim_cdouble{1.0, -0.1});
  CHECK(wsim_problem_query(&d, &info) == WSIM_OK);
  std::vector<wsim_cfloat> g((size_t)(info.N[0] * info.N[1]));
  CHECK(wsim_problem_build(&d, n.data(), &info, g.data(), nullptr) != WSIM_OK);
  d.has_roi = 1;
  d.roi[0][0] = 1;
  d.roi[1][0] = 99;
  CHECK(wsim_problem_query(&d, &info) != WSIM_OK);
  d.has_roi = 0;
  d.shape[0] = 0;
  CHECK(wsim_problem_query(&d, &info) != WSIM_OK);
  CHECK(wsim_problem_query(nullptr, &info) != WSIM_OK);
}

// ---------------------------------------------------------------- MEX shim
struct MexInputs {
  mxArray* a[12];
  ~MexInputs() {
    for (auto* x : a)
      if (x) mxf_destroy(x);
  }
};

// a valid 8 x 6 scalar problem with Nw = 4 (values are placeholders: the
// stand-in loop does not read them)
static void make_inputs(MexInputs& in, const char* cb, double cbint) {
  const int64_t Ny = 8, Nx = 6, nw = 4;
  in.a[0] = dbl({1, 4}, {(double)Ny, (double)Nx, 1, 1});
  in.a[1] = sgl({Ny, Nx}, std::vector<float>(2 * Ny * Nx, 0.5f), true);
  in.a[2] = dbl({1, 1}, {(double)nw});
  in.a[3] = mxf_cell(nw);
  for (int w = 0; w < nw; ++w) {
    mxArray* s = mxf_struct();
    mxf_set(s, "pxe", sgl({1, Nx}, std::vector<float>(Nx, 1.f), false));
    mxf_set(s, "pye", sgl({Ny, 1}, std::vector<float>(Ny, 1.f), false));
    mxf_set(s, "pze", sgl({1, 1, 1}, std::vector<float>(1, 0.f), false));
    mxf_set(s, "gx", sgl({1, Nx}, std::vector<float>(2 * Nx, 1.f), true));
    mxf_set(s, "gy", sgl({Ny, 1}, std::vector<float>(2 * Ny, 1.f), true));
    mxf_set(s, "gz", sgl({1, 1, 1}, std::vector<float>(2, 1.f), true));
    mxf_cell_set(in.a[3], w, s);
  }
  mxArray* src = mxf_object();
  mxArray* pos = mxf_cell(2);
  mxArray* val = mxf_cell(2);
  mxf_cell_set(pos, 0, dbl({1, 4}, {4, 3, 1, 1}));
  mxf_cell_set(val, 0, sgl({1, 1}, {2.5f, -1.5f}, true));
  mxf_cell_set(pos, 1, dbl({1, 4}, {1, 1, 1, 1}));
  mxf_cell_set(val, 1, sgl({Ny, 1}, std::vector<float>(Ny, 1.f), false));
  mxf_set(src, "positions", pos);
  mxf_set(src, "values", val);
  in.a[4] = src;
  in.a[5] = sgl({1, 1}, {3.0f}, false);
  in.a[6] = sgl({1, 1}, {1.2f, 1.0f}, true);
  in.a[7] = dbl({2, 4}, {2, 7, 2, 5, 1, 1, 1, 1});
  in.a[8] = dbl({1, 1}, {1e5});
  in.a[9] = dbl({1, 1}, {1e-6});
  in.a[10] = mxf_string(cb);
  in.a[11] = dbl({1, 1}, {cbint});
}

static std::string call(MexInputs& in, mxArray** out, int nrhs = 12) {
  char err[512] = {0};
  const int rc = mxf_call(2, out, nrhs, in.a, err, sizeof(err));
  return rc == 0 ? std::string() : std::string(err);
}

static void check_shim() {
  {
    MexInputs in;
    make_inputs(in, "NoCallback", 50);
    mxArray* out[2];
    mxf_log_clear();
    CHECK(call(in, out).empty());
    CHECK(out[0] && out[1]);
    if (out[0] && out[1]) {
      CHECK(mxf_ndims(out[0]) == 4 && mxf_dim(out[0], 0) == 8 && mxf_dim(out[0], 1) == 6);
      const float* e = static_cast<const float*>(mxf_data(out[0]));
      CHECK(e[0] == 2.5f && e[1] == -1.5f);  // first source value
      CHECK(std::fabs(e[2 * 5] - 5e-3f) < 1e-9f);
      CHECK(mxf_dim(out[1], 0) == 1 && mxf_numel(out[1]) == 23);
      const double* de = static_cast<const double*>(mxf_data(out[1]));
      CHECK(de[0] == 1.0 && std::fabs(de[22] - 1.0 / 23) < 1e-15);
      mxf_destroy(out[0]);
      mxf_destroy(out[1]);
    }
    CHECK(mxf_log_count() == 0);
    CHECK(mxf_live_temporaries() == 0);
  }
  {
    MexInputs in;
    make_inputs(in, "EnergyAddedDisp", 5);
    mxArray* out[2];
    mxf_log_clear();
    CHECK(call(in, out).empty());
    CHECK(mxf_log_count() == 5);  // after 5, 10, 15, 20 and the last (23)
    const int64_t want[5] = {5, 10, 15, 20, 23};
    for (int i = 0; i < mxf_log_count() && i < 5; ++i) {
      char name[64], type[64];
      double de[64], thr;
      int64_t nde, nen;
      int nargs;
      CHECK(mxf_log_entry(i, name, 64, type, 64, de, 64, &nde, &nen, &thr, &nargs) == 0);
      CHECK(std::strcmp(name, "SimulationCallback") == 0 && std::strcmp(type, "EnergyAddedDisp") == 0);
      CHECK(nde == want[i] && nen == 0 && thr == 1e-6 && nargs == 4);
    }
    if (out[0]) mxf_destroy(out[0]);
    if (out[1]) mxf_destroy(out[1]);
    // a failing callback stops the solve and raises after the library returned
    mxf_fail_callback_at(2);
    mxf_log_clear();
    const std::string e = call(in, out);
    CHECK(e.find("wsim:callback") != std::string::npos && mxf_log_count() == 2);
    mxf_fail_callback_at(0);
    CHECK(mxf_live_temporaries() == 0);
  }
  // argument checks: each raises an error naming the argument
  struct Bad {
    int arg;
    const char* expect;
  } bad[] = {{-1, "12 inputs"}, {0, "N"}, {1, "gamma"}, {3, "wiggles"}, {4, "single"}, {5, "epsilon"},
             {6, "k02e"}, {7, "roi"}, {10, "callbackString"}, {31, "pxe"}, {32, "wrong length"}, {41, "position"}};
  for (const Bad& b : bad) {
    MexInputs in;
    make_inputs(in, "NoCallback", 50);
    int nrhs = 12;
    switch (b.arg) {
      case -1: nrhs = 11; break;
      case 0: mxf_destroy(in.a[0]); in.a[0] = sgl({1, 4}, {8, 6, 1, 1}, false); break;
      case 1: mxf_destroy(in.a[1]); in.a[1] = sgl({8, 5}, std::vector<float>(80, 0.f), true); break;
      case 3: mxf_destroy(in.a[3]); in.a[3] = mxf_cell(3); break;
      case 4: {  // double-precision source values (single_precision = false)
        mxf_destroy(in.a[4]);
        mxArray* src = mxf_object();
        mxArray* pos = mxf_cell(1);
        mxArray* val = mxf_cell(1);
        mxf_cell_set(pos, 0, dbl({1, 4}, {1, 1, 1, 1}));
        mxf_cell_set(val, 0, dbl({1, 1}, {1.0}));
        mxf_set(src, "positions", pos);
        mxf_set(src, "values", val);
        in.a[4] = src;
        break;
      }
      case 5: mxf_destroy(in.a[5]); in.a[5] = mxf_string("x"); break;
      case 6: mxf_destroy(in.a[6]); in.a[6] = sgl({1, 1}, {1.0f}, false); break;
      case 7: mxf_destroy(in.a[7]); in.a[7] = dbl({1, 4}, {1, 1, 1, 1}); break;
      case 10: mxf_destroy(in.a[10]); in.a[10] = dbl({1, 1}, {0}); break;
      case 31: {  // a wiggle struct without pxe
        mxArray* s = mxf_struct();
        mxf_set(s, "pye", sgl({8, 1}, std::vector<float>(8, 1.f), false));
        mxArray* c = mxf_cell(4);
        for (int w = 0; w < 4; ++w) mxf_cell_set(c, w, w == 0 ? s : mxf_struct());
        mxf_destroy(in.a[3]);
        in.a[3] = c;
        break;
      }
      case 32: {  // pxe of the wrong length
        mxArray* s = mxf_struct();
        mxf_set(s, "pxe", sgl({1, 5}, std::vector<float>(5, 1.f), false));
        mxf_set(s, "pye", sgl({8, 1}, std::vector<float>(8, 1.f), false));
        mxf_set(s, "pze", sgl({1, 1}, {0.f}, false));
        mxf_set(s, "gx", sgl({1, 5}, std::vector<float>(10, 1.f), true));
        mxf_set(s, "gy", sgl({8, 1}, std::vector<float>(16, 1.f), true));
        mxf_set(s, "gz", sgl({1, 1}, {1.f, 0.f}, true));
        mxArray* c = mxf_cell(4);
        for (int w = 0; w < 4; ++w) mxf_cell_set(c, w, w == 0 ? s : mxf_struct());
        mxf_destroy(in.a[3]);
        in.a[3] = c;
        break;
      }
      case 41: {  // a 1x3 position
        mxf_destroy(in.a[4]);
        mxArray* src = mxf_object();
        mxArray* pos = mxf_cell(1);
        mxArray* val = mxf_cell(1);
        mxf_cell_set(pos, 0, dbl({1, 3}, {1, 1, 1}));
        mxf_cell_set(val, 0, sgl({1, 1}, {1.f}, false));
        mxf_set(src, "positions", pos);
        mxf_set(src, "values", val);
        in.a[4] = src;
        break;
      }
    }
    mxArray* out[2];
    const std::string e = call(in, out, nrhs);
    if (e.find(b.expect) == std::string::npos) {
      std::fprintf(stderr, "argument check %d: expected '%s', got '%s'\n", b.arg, b.expect, e.c_str());
      ++g_fail;
    }
    CHECK(mxf_live_temporaries() == 0);
  }
}

int main() {
  check_problem_setup();
  check_shim();
  if (g_fail) {
    std::fprintf(stderr, "%d host checks failed\n", g_fail);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
