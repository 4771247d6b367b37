// A CPU fake of the part of MATLAB's C API that the MEX shim uses, to run the
// shim's marshalling without MATLAB: numeric, char, cell, struct and object
// arrays; mexErrMsgIdAndTxt throws (the harness entry point catches it, as
// MATLAB's error handling would end the MEX call); mexCallMATLAB records each
// call (the SimulationCallback arguments) instead of running MATLAB code.
// Arrays created during a call and not returned are freed when the call ends,
// as MATLAB frees a MEX function's temporaries.  Test infrastructure only.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
  mxClassID cls = mxUNKNOWN_CLASS;
  bool complex = false;
  std::vector<mwSize> dims;
  std::vector<unsigned char> data;  // numeric (complex interleaved) and char (UTF-8 bytes)
  std::vector<mxArray*> cells;
  std::vector<std::pair<std::string, mxArray*>> fields;  // 1x1 struct fields or object properties
};

namespace {

struct MexError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

std::set<mxArray*> g_temps;  // created during the current call
bool g_in_call = false;

struct Call {
  std::string name, type;
  std::vector<double> de;
  size_t energy_numel = 0;
  double thr = 0;
  int nargs = 0;
};
std::vector<Call> g_log;
int g_fail_callback = 0;  // make mexCallMATLABWithTrap report an error

size_t numel_of(const mxArray* a) {
  size_t n = 1;
  for (mwSize d : a->dims) n *= d;
  return a->dims.empty() ? 0 : n;
}

size_t elem_bytes(const mxArray* a) {
  const size_t b = a->cls == mxDOUBLE_CLASS ? 8 : a->cls == mxSINGLE_CLASS ? 4 : a->cls == mxCHAR_CLASS ? 1 : 0;
  return a->complex ? 2 * b : b;
}

mxArray* make(mxClassID cls, bool cplx, size_t ndim, const mwSize* dims) {
  auto* a = new mxArray_tag();
  a->cls = cls;
  a->complex = cplx;
  a->dims.assign(dims, dims + ndim);
  while (a->dims.size() < 2) a->dims.push_back(1);
  a->data.assign(numel_of(a) * elem_bytes(a), 0);
  if (g_in_call) g_temps.insert(a);
  return a;
}

void destroy(mxArray* a) {
  if (!a) return;
  g_temps.erase(a);
  for (auto* c : a->cells) destroy(c);
  for (auto& f : a->fields) destroy(f.second);
  delete a;
}

mxArray* deep_copy(const mxArray* a) {
  auto* b = new mxArray_tag(*a);
  for (auto*& c : b->cells) c = c ? deep_copy(c) : nullptr;
  for (auto& f : b->fields) f.second = f.second ? deep_copy(f.second) : nullptr;
  return b;
}

}  // namespace

extern "C" {

// ---- the MATLAB C API subset --------------------------------------------------
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw MexError(std::string(id) + ": " + buf);
}

static void record(int nrhs, mxArray** prhs, const char* name) {
  Call c;
  c.name = name;
  c.nargs = nrhs;
  if (nrhs > 0 && prhs[0]->cls == mxCHAR_CLASS) c.type.assign(prhs[0]->data.begin(), prhs[0]->data.end());
  if (nrhs > 1 && prhs[1]->cls == mxDOUBLE_CLASS) {
    const double* d = reinterpret_cast<const double*>(prhs[1]->data.data());
    c.de.assign(d, d + numel_of(prhs[1]));
  }
  if (nrhs > 2) c.energy_numel = numel_of(prhs[2]);
  if (nrhs > 3 && prhs[3]->cls == mxDOUBLE_CLASS) c.thr = *reinterpret_cast<const double*>(prhs[3]->data.data());
  g_log.push_back(c);
}

int mexCallMATLAB(int, mxArray**, int nrhs, mxArray** prhs, const char* name) {
  record(nrhs, prhs, name);
  return 0;
}

mxArray* mexCallMATLABWithTrap(int, mxArray**, int nrhs, mxArray** prhs, const char* name) {
  record(nrhs, prhs, name);
  if (g_fail_callback > 0 && (int)g_log.size() >= g_fail_callback) {
    const mwSize one[2] = {1, 1};
    return make(mxDOUBLE_CLASS, false, 2, one);  // stands for the MException object
  }
  return nullptr;
}

mxArray* mxGetCell(const mxArray* a, mwIndex i) {
  if (!a || a->cls != mxCELL_CLASS || i >= a->cells.size()) return nullptr;
  return a->cells[i];
}

mxComplexSingle* mxGetComplexSingles(const mxArray* a) {
  if (!a || a->cls != mxSINGLE_CLASS || !a->complex) return nullptr;
  return reinterpret_cast<mxComplexSingle*>(const_cast<unsigned char*>(a->data.data()));
}

mxSingle* mxGetSingles(const mxArray* a) {
  if (!a || a->cls != mxSINGLE_CLASS || a->complex) return nullptr;
  return reinterpret_cast<mxSingle*>(const_cast<unsigned char*>(a->data.data()));
}

mxDouble* mxGetDoubles(const mxArray* a) {
  if (!a || a->cls != mxDOUBLE_CLASS || a->complex) return nullptr;
  return reinterpret_cast<mxDouble*>(const_cast<unsigned char*>(a->data.data()));
}

const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
size_t mxGetNumberOfElements(const mxArray* a) {
  if (a->cls == mxSTRUCT_CLASS || a->cls == mxOBJECT_CLASS) return 1;
  return a->cls == mxCELL_CLASS ? a->cells.size() : numel_of(a);
}

mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name) {
  if (!a || a->cls != mxSTRUCT_CLASS || i != 0) return nullptr;
  for (auto& f : a->fields)
    if (f.first == name) return f.second;
  return nullptr;
}

// MATLAB returns a copy of the property (freed with the call's temporaries)
mxArray* mxGetProperty(const mxArray* a, mwIndex i, const char* name) {
  if (!a || a->cls != mxOBJECT_CLASS || i != 0) return nullptr;
  for (auto& f : a->fields)
    if (f.first == name) {
      mxArray* c = deep_copy(f.second);
      if (g_in_call) g_temps.insert(c);
      return c;
    }
  return nullptr;
}

double mxGetScalar(const mxArray* a) {
  if (!a || numel_of(a) == 0) return 0.0;
  if (a->cls == mxDOUBLE_CLASS) return *reinterpret_cast<const double*>(a->data.data());
  if (a->cls == mxSINGLE_CLASS) return *reinterpret_cast<const float*>(a->data.data());
  if (a->cls == mxCHAR_CLASS) return (double)a->data[0];
  return 0.0;
}

int mxGetString(const mxArray* a, char* buf, mwSize len) {
  if (!a || a->cls != mxCHAR_CLASS || len == 0) return 1;
  const size_t n = a->data.size();
  const size_t m = n < len - 1 ? n : len - 1;
  std::memcpy(buf, a->data.data(), m);
  buf[m] = 0;
  return m < n ? 1 : 0;
}

bool mxIsCell(const mxArray* a) { return a && a->cls == mxCELL_CLASS; }
bool mxIsChar(const mxArray* a) { return a && a->cls == mxCHAR_CLASS; }
bool mxIsComplex(const mxArray* a) { return a && a->complex; }
bool mxIsDouble(const mxArray* a) { return a && a->cls == mxDOUBLE_CLASS; }
bool mxIsSingle(const mxArray* a) { return a && a->cls == mxSINGLE_CLASS; }
bool mxIsStruct(const mxArray* a) { return a && a->cls == mxSTRUCT_CLASS; }

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
  const mwSize d[2] = {m, n};
  return make(mxDOUBLE_CLASS, c == mxCOMPLEX, 2, d);
}

mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  *reinterpret_cast<double*>(a->data.data()) = v;
  return a;
}

mxArray* mxCreateNumericArray(mwSize nd, const mwSize* dims, mxClassID cls, mxComplexity c) {
  return make(cls, c == mxCOMPLEX, nd, dims);
}

mxArray* mxCreateString(const char* s) {
  const mwSize d[2] = {1, std::strlen(s)};
  mxArray* a = make(mxCHAR_CLASS, false, 2, d);
  std::memcpy(a->data.data(), s, std::strlen(s));
  return a;
}

void mxDestroyArray(mxArray* a) { destroy(a); }

// ---- harness entry points (Python ctypes side) ----------------------------------
mxArray* mxf_numeric(int cls, int cplx, int ndim, const int64_t* dims, const void* data) {
  std::vector<mwSize> d(dims, dims + ndim);
  mxArray* a = make(cls == 0 ? mxDOUBLE_CLASS : mxSINGLE_CLASS, cplx != 0, (size_t)ndim, d.data());
  if (data && !a->data.empty()) std::memcpy(a->data.data(), data, a->data.size());
  return a;
}
mxArray* mxf_string(const char* s) { return mxCreateString(s); }
mxArray* mxf_cell(int64_t n) {
  const mwSize d[2] = {(mwSize)n, 1};
  mxArray* a = make(mxCELL_CLASS, false, 2, d);
  a->cells.assign((size_t)n, nullptr);
  return a;
}
void mxf_cell_set(mxArray* c, int64_t i, mxArray* v) { c->cells.at((size_t)i) = v; }
mxArray* mxf_struct(void) {
  const mwSize d[2] = {1, 1};
  return make(mxSTRUCT_CLASS, false, 2, d);
}
mxArray* mxf_object(void) {
  const mwSize d[2] = {1, 1};
  return make(mxOBJECT_CLASS, false, 2, d);
}
void mxf_set(mxArray* s, const char* name, mxArray* v) { s->fields.emplace_back(name, v); }
void mxf_destroy(mxArray* a) { destroy(a); }
int64_t mxf_numel(const mxArray* a) { return (int64_t)numel_of(a); }
int mxf_ndims(const mxArray* a) { return (int)a->dims.size(); }
int64_t mxf_dim(const mxArray* a, int k) { return k < (int)a->dims.size() ? (int64_t)a->dims[k] : 1; }
int mxf_class(const mxArray* a) { return a->cls == mxDOUBLE_CLASS ? 0 : a->cls == mxSINGLE_CLASS ? 1 : -1; }
int mxf_is_complex(const mxArray* a) { return a->complex ? 1 : 0; }
const void* mxf_data(const mxArray* a) { return a->data.data(); }
int64_t mxf_live_temporaries(void) { return (int64_t)g_temps.size(); }

// runs mexFunction; 0 on success, 1 when it raised an error (message in err)
int mxf_call(int nlhs, mxArray** plhs, int nrhs, mxArray** prhs, char* err, int errcap) {
  g_in_call = true;
  g_temps.clear();
  for (int i = 0; i < nlhs; ++i) plhs[i] = nullptr;
  int rc = 0;
  try {
    mexFunction(nlhs, plhs, nrhs, const_cast<const mxArray**>(prhs));
  } catch (const MexError& e) {
    std::snprintf(err, (size_t)errcap, "%s", e.what());
    rc = 1;
  }
  g_in_call = false;
  // outputs survive; the other temporaries are freed as MATLAB frees them
  std::set<mxArray*> keep;
  if (rc == 0)
    for (int i = 0; i < nlhs; ++i) keep.insert(plhs[i]);
  std::vector<mxArray*> drop;
  for (mxArray* t : g_temps)
    if (!keep.count(t)) drop.push_back(t);
  for (mxArray* t : drop) destroy(t);
  g_temps.clear();
  if (rc != 0)
    for (int i = 0; i < nlhs; ++i) plhs[i] = nullptr;
  return rc;
}

void mxf_log_clear(void) { g_log.clear(); }
int mxf_log_count(void) { return (int)g_log.size(); }
// call i: function name, callback type, length of diff_energy (and its
// values into de up to cap), number of elements of the energy argument,
// threshold, argument count
int mxf_log_entry(int i, char* name, int ncap, char* type, int tcap, double* de, int64_t cap, int64_t* n_de,
                  int64_t* n_energy, double* thr, int* nargs) {
  if (i < 0 || i >= (int)g_log.size()) return 1;
  const Call& c = g_log[(size_t)i];
  std::snprintf(name, (size_t)ncap, "%s", c.name.c_str());
  std::snprintf(type, (size_t)tcap, "%s", c.type.c_str());
  *n_de = (int64_t)c.de.size();
  for (size_t k = 0; k < c.de.size() && (int64_t)k < cap; ++k) de[k] = c.de[k];
  *n_energy = (int64_t)c.energy_numel;
  *thr = c.thr;
  *nargs = c.nargs;
  return 0;
}
void mxf_fail_callback_at(int k) { g_fail_callback = k; }

}  // extern "C"
