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

// forwards the library's progress callback to SimulationCallback.m (SimulationCallback.m:1)
int forward_callback(const char* type, int64_t, const double* de, int64_t n, double thr, void*) {
  if (std::strcmp(type, "NoCallback") == 0) return 0;
  mxArray* args[4];
  args[0] = mxCreateString(type);
  args[1] = mxCreateDoubleMatrix(1, (mwSize)n, mxREAL);
  std::memcpy(mxGetDoubles(args[1]), de, (size_t)n * sizeof(double));
  args[2] = mxCreateDoubleMatrix(0, 0, mxREAL);
  args[3] = mxCreateDoubleScalar(thr);
  mexCallMATLAB(0, nullptr, 4, args, "SimulationCallback");
  for (auto* a : args) mxDestroyArray(a);
  return 0;
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs != 12) mexErrMsgIdAndTxt("wsim:arg", "RunWaveSimAlgorithmMex takes 12 inputs");
  const double* N = mxGetDoubles(prhs[0]);
  const wsim_cfloat* gamma = complex_single(prhs[1], "gamma");
  const int nw = (int)mxGetScalar(prhs[2]);
  std::vector<wsim_wiggle> wig(nw);
  for (int w = 0; w < nw; ++w) {
    const mxArray* d = mxGetCell(prhs[3], w);
    wig[w].pxe = real_single(field(d, 0, "pxe"), "pxe");
    wig[w].pye = real_single(field(d, 0, "pye"), "pye");
    wig[w].pze = real_single(field(d, 0, "pze"), "pze");
    wig[w].gx = complex_single(field(d, 0, "gx"), "gx");
    wig[w].gy = complex_single(field(d, 0, "gy"), "gy");
    wig[w].gz = complex_single(field(d, 0, "gz"), "gz");
  }
  // Source object (Source.m:17-21): cells of positions (1x4 double) and values
  const mxArray* pos = mxGetProperty(prhs[4], 0, "positions");
  const mxArray* val = mxGetProperty(prhs[4], 0, "values");
  const mwSize np = mxGetNumberOfElements(pos);
  std::vector<wsim_source_patch> patches(np);
  std::vector<std::vector<wsim_cfloat>> keep(np);
  for (mwSize p = 0; p < np; ++p) {
    const double* pp = mxGetDoubles(mxGetCell(pos, p));
    const mxArray* v = mxGetCell(val, p);
    const mwSize nd = mxGetNumberOfDimensions(v);
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
  const float eps = (float)mxGetScalar(prhs[5]);
  const mxComplexSingle k = mxGetComplexSingles(prhs[6])[0];
  const double* roi = mxGetDoubles(prhs[7]);
  const double max_iter = mxGetScalar(prhs[8]);
  const double thr = mxGetScalar(prhs[9]);
  char cb[64];
  mxGetString(prhs[10], cb, sizeof(cb));
  const double cb_int = mxGetScalar(prhs[11]);

  const mwSize dims[4] = {(mwSize)N[0], (mwSize)N[1], (mwSize)N[2], (mwSize)N[3]};
  plhs[0] = mxCreateNumericArray(4, dims, mxSINGLE_CLASS, mxCOMPLEX);
  // diff_energy has one entry per iteration run: the cap (max_iter, or
  // 100000 when it is not a positive number, WaveSimBase.m:127-131) takes
  // effect at the first multiple of Nwiggles at or after it (Simulation.m:202)
  const double mi = (std::isfinite(max_iter) && max_iter > 0) ? std::floor(max_iter) : 100000.0;
  const int64_t cap = (int64_t)std::min(std::ceil(mi / std::max(nw, 1)) * std::max(nw, 1), 1e7);
  std::vector<double> de((size_t)cap);
  int64_t n = 0;
  const int rc = wsim_run_algorithm(N, gamma, nw, wig.data(), (int)np, patches.data(), eps,
                                    wsim_cfloat{k.real, k.imag}, roi, max_iter, thr, cb, cb_int,
                                    forward_callback, nullptr,
                                    reinterpret_cast<wsim_cfloat*>(mxGetComplexSingles(plhs[0])),
                                    de.data(), cap, &n);
  if (rc != WSIM_OK && rc != WSIM_ERR_ABORTED)
    mexErrMsgIdAndTxt("wsim:run", "wsim_run_algorithm failed (%d): %s", rc, wsim_last_error());
  n = std::min(n, cap);
  plhs[1] = mxCreateDoubleMatrix(1, (mwSize)n, mxREAL);
  std::memcpy(mxGetDoubles(plhs[1]), de.data(), (size_t)n * sizeof(double));
}
