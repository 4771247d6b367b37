// MEX shim: RunWaveSimAlgorithmMex on top of libwavesim_mbs (include/wavesim_mbs.h).
//
// Same MATLAB signature as the call in WaveSimBase.m:145-146:
//   [E, diff_energy] = RunWaveSimAlgorithmMex(N, gamma, Nwiggles, wiggles, source, ...
//        epsilon, k02e, roi, maxIter, energy_threshold, callbackString, callback_interval)
// Inputs are host mxArrays (R2018a+ interleaved complex, build with -R2018a); gamma and
// the wiggle tables are single precision as WaveSimBase stores them.  Every error
// becomes mexErrMsgIdAndTxt after the library call returned (no exception crosses
// the C ABI).  Build (on a machine with MATLAB):
//   mex -R2018a RunWaveSimAlgorithmMex.cpp -I<repo>/include -L<repo>/wavesim_amd -lwavesim_mbs
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"
#include "wavesim_mbs.h"

namespace {

const mxArray* field(const mxArray* s, mwIndex i, const char* name) {
  const mxArray* f = mxGetField(s, i, name);
  if (!f) mexErrMsgIdAndTxt("wsim:arg", "wiggle descriptor lacks field %s", name);
  return f;
}

const float* real_single(const mxArray* a, const char* what) {
  if (!mxIsSingle(a) || mxIsComplex(a)) mexErrMsgIdAndTxt("wsim:arg", "%s must be real single", what);
  return mxGetSingles(a);
}

const wsim_cfloat* complex_single(const mxArray* a, const char* what) {
  if (!mxIsSingle(a) || !mxIsComplex(a)) mexErrMsgIdAndTxt("wsim:arg", "%s must be complex single", what);
  return reinterpret_cast<const wsim_cfloat*>(mxGetComplexSingles(a));
}

// a real double array of at least n elements (N, roi, source positions)
const double* real_double(const mxArray* a, size_t n, const char* what) {
  if (!a || !mxIsDouble(a) || mxIsComplex(a) || mxGetNumberOfElements(a) < n)
    mexErrMsgIdAndTxt("wsim:arg", "%s must be a real double array of %d elements", what, (int)n);
  return mxGetDoubles(a);
}

// a numeric scalar (any real class: mxGetScalar converts it)
double scalar(const mxArray* a, const char* what) {
  if (!a || !(mxIsDouble(a) || mxIsSingle(a)) || mxGetNumberOfElements(a) < 1)
    mexErrMsgIdAndTxt("wsim:arg", "%s must be a numeric scalar", what);
  return mxGetScalar(a);
}

// forwards the library's progress callback to SimulationCallback.m
// (SimulationCallback.m:1, arguments callbackType, diff_energy, energy,
// energy_threshold).  A MATLAB error in the callback is trapped (it must not
// unwind through the library) and stops the solve; mexFunction raises it
// after the library returned.
int forward_callback(const char* type, int64_t, const double* de, int64_t n, double thr, void* user) {
  if (std::strcmp(type, "NoCallback") == 0) return 0;
  mxArray* args[4];
  args[0] = mxCreateString(type);
  args[1] = mxCreateDoubleMatrix(1, (mwSize)n, mxREAL);
  std::memcpy(mxGetDoubles(args[1]), de, (size_t)n * sizeof(double));
  args[2] = mxCreateDoubleMatrix(0, 0, mxREAL);
  args[3] = mxCreateDoubleScalar(thr);
  mxArray* err = mexCallMATLABWithTrap(0, nullptr, 4, args, "SimulationCallback");
  for (auto* a : args) mxDestroyArray(a);
  if (err != nullptr) {
    *static_cast<bool*>(user) = true;
    return 1;
  }
  return 0;
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs != 12) mexErrMsgIdAndTxt("wsim:arg", "RunWaveSimAlgorithmMex takes 12 inputs");
  if (nlhs > 2) mexErrMsgIdAndTxt("wsim:arg", "RunWaveSimAlgorithmMex returns at most 2 outputs");
  const double* N = real_double(prhs[0], 4, "N");
  for (int c = 0; c < 4; ++c)
    if (!(N[c] >= 1 && N[c] == std::floor(N[c]))) mexErrMsgIdAndTxt("wsim:arg", "N must hold positive integers");
  const wsim_cfloat* gamma = complex_single(prhs[1], "gamma");
  if (mxGetNumberOfElements(prhs[1]) != (size_t)(N[0] * N[1] * N[2]))
    mexErrMsgIdAndTxt("wsim:arg", "gamma must have N(1)*N(2)*N(3) elements");
  const int nw = (int)scalar(prhs[2], "Nwiggles");
  if (nw < 1 || !mxIsCell(prhs[3]) || mxGetNumberOfElements(prhs[3]) != (size_t)nw)
    mexErrMsgIdAndTxt("wsim:arg", "wiggles must be a cell array of Nwiggles descriptors");
  std::vector<wsim_wiggle> wig(nw);
  for (int w = 0; w < nw; ++w) {
    const mxArray* d = mxGetCell(prhs[3], w);
    if (!d || !mxIsStruct(d)) mexErrMsgIdAndTxt("wsim:arg", "wiggles{%d} must be a struct", w + 1);
    wig[w].pxe = real_single(field(d, 0, "pxe"), "pxe");
    wig[w].pye = real_single(field(d, 0, "pye"), "pye");
    wig[w].pze = real_single(field(d, 0, "pze"), "pze");
    wig[w].gx = complex_single(field(d, 0, "gx"), "gx");
    wig[w].gy = complex_single(field(d, 0, "gy"), "gy");
    wig[w].gz = complex_single(field(d, 0, "gz"), "gz");
    const size_t len[3] = {(size_t)N[1], (size_t)N[0], (size_t)N[2]};
    const char* names[6] = {"pxe", "pye", "pze", "gx", "gy", "gz"};
    for (int k = 0; k < 6; ++k)
      if (mxGetNumberOfElements(field(d, 0, names[k])) != len[k % 3])
        mexErrMsgIdAndTxt("wsim:arg", "wiggles{%d}.%s has the wrong length", w + 1, names[k]);
  }
  // Source object (Source.m:17-21): cells of positions (1x4 double) and values
  // (single: data_array cast them, Simulation.m:154, 230-248)
  const mxArray* pos = mxGetProperty(prhs[4], 0, "positions");
  const mxArray* val = mxGetProperty(prhs[4], 0, "values");
  if (!pos || !val || !mxIsCell(pos) || !mxIsCell(val) ||
      mxGetNumberOfElements(pos) != mxGetNumberOfElements(val))
    mexErrMsgIdAndTxt("wsim:arg", "source must be a Source object with matching positions and values cells");
  const mwSize np = mxGetNumberOfElements(pos);
  std::vector<wsim_source_patch> patches(np);
  std::vector<std::vector<wsim_cfloat>> keep(np);
  for (mwSize p = 0; p < np; ++p) {
    const double* pp = real_double(mxGetCell(pos, p), 4, "source position");
    const mxArray* v = mxGetCell(val, p);
    if (!v || !mxIsSingle(v))
      mexErrMsgIdAndTxt("wsim:arg", "source values must be single (the solver computes in single precision)");
    const mwSize nd = mxGetNumberOfDimensions(v);
    if (nd > 4) mexErrMsgIdAndTxt("wsim:arg", "source values have more than 4 dimensions");
    const mwSize* dims = mxGetDimensions(v);
    size_t count = 1;
    for (int c = 0; c < 4; ++c) {
      patches[p].pos[c] = pp[c];
      patches[p].size[c] = c < (int)nd ? (int64_t)dims[c] : 1;
      count *= (size_t)patches[p].size[c];
    }
    keep[p].resize(count);
    if (mxIsComplex(v)) {
      const mxComplexSingle* s = mxGetComplexSingles(v);
      for (size_t k = 0; k < count; ++k) keep[p][k] = {s[k].real, s[k].imag};
    } else {
      const float* s = mxGetSingles(v);
      for (size_t k = 0; k < count; ++k) keep[p][k] = {s[k], 0.f};
    }
    patches[p].values = keep[p].data();
  }
  const float eps = (float)scalar(prhs[5], "epsilon");
  const wsim_cfloat k = complex_single(prhs[6], "k02e")[0];
  const double* roi = real_double(prhs[7], 8, "roi");
  const double max_iter = scalar(prhs[8], "maxIter");
  const double thr = scalar(prhs[9], "energy_threshold");
  char cb[64];
  if (!mxIsChar(prhs[10]) || mxGetString(prhs[10], cb, sizeof(cb)) != 0)
    mexErrMsgIdAndTxt("wsim:arg", "callbackString must be a char array");
  const double cb_int = scalar(prhs[11], "callback_interval");

  const mwSize dims[4] = {(mwSize)N[0], (mwSize)N[1], (mwSize)N[2], (mwSize)N[3]};
  plhs[0] = mxCreateNumericArray(4, dims, mxSINGLE_CLASS, mxCOMPLEX);
  // diff_energy has one entry per iteration run: the cap (max_iter, or
  // 100000 when it is not a positive number, WaveSimBase.m:127-131) takes
  // effect at the first multiple of Nwiggles at or after it (Simulation.m:202)
  const double mi = (std::isfinite(max_iter) && max_iter > 0) ? std::floor(max_iter) : 100000.0;
  const int64_t cap = (int64_t)std::min(std::ceil(mi / std::max(nw, 1)) * std::max(nw, 1), 1e7);
  std::vector<double> de((size_t)cap);
  int64_t n = 0;
  bool cb_failed = false;
  const int rc = wsim_run_algorithm(N, gamma, nw, wig.data(), (int)np, patches.data(), eps,
                                    k, roi, max_iter, thr, cb, cb_int,
                                    std::strcmp(cb, "NoCallback") == 0 ? nullptr : forward_callback, &cb_failed,
                                    reinterpret_cast<wsim_cfloat*>(mxGetComplexSingles(plhs[0])),
                                    de.data(), cap, &n);
  if (cb_failed) mexErrMsgIdAndTxt("wsim:callback", "SimulationCallback failed; the solve was stopped");
  if (rc == WSIM_ERR_DIVERGED)
    mexErrMsgIdAndTxt("wsim:diverged", "the iteration diverged after %d iterations: %s", (int)n, wsim_last_error());
  if (rc != WSIM_OK && rc != WSIM_ERR_ABORTED)
    mexErrMsgIdAndTxt("wsim:run", "wsim_run_algorithm failed (%d): %s", rc, wsim_last_error());
  n = std::min(n, cap);
  plhs[1] = mxCreateDoubleMatrix(1, (mwSize)n, mxREAL);
  std::memcpy(mxGetDoubles(plhs[1]), de.data(), (size_t)n * sizeof(double));
}
