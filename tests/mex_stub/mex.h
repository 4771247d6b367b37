// Declarations of the MATLAB C API functions the MEX shim calls, for building
// the shim where MATLAB is not installed: against these declarations for a
// syntax check, and against the CPU fake of tests/mex_stub/mxfake.cpp to run
// it (not MATLAB's header; names and signatures as MATLAB's C API documents).
#pragma once
#include <cstddef>
#include <cstdint>
typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;
typedef enum { mxUNKNOWN_CLASS, mxCELL_CLASS, mxSTRUCT_CLASS, mxCHAR_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS,
               mxOBJECT_CLASS } mxClassID;
typedef struct { float real, imag; } mxComplexSingle;
typedef float mxSingle;
typedef double mxDouble;
extern "C" {
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
void mexErrMsgIdAndTxt(const char*, const char*, ...);
int mexCallMATLAB(int, mxArray**, int, mxArray**, const char*);
mxArray* mexCallMATLABWithTrap(int, mxArray**, int, mxArray**, const char*);
mxArray* mxGetCell(const mxArray*, mwIndex);
mxComplexSingle* mxGetComplexSingles(const mxArray*);
const mwSize* mxGetDimensions(const mxArray*);
mxDouble* mxGetDoubles(const mxArray*);
mxArray* mxGetField(const mxArray*, mwIndex, const char*);
mwSize mxGetNumberOfDimensions(const mxArray*);
size_t mxGetNumberOfElements(const mxArray*);
mxArray* mxGetProperty(const mxArray*, mwIndex, const char*);
double mxGetScalar(const mxArray*);
mxSingle* mxGetSingles(const mxArray*);
int mxGetString(const mxArray*, char*, mwSize);
bool mxIsCell(const mxArray*);
bool mxIsChar(const mxArray*);
bool mxIsComplex(const mxArray*);
bool mxIsDouble(const mxArray*);
bool mxIsSingle(const mxArray*);
bool mxIsStruct(const mxArray*);
mxArray* mxCreateDoubleMatrix(mwSize, mwSize, mxComplexity);
mxArray* mxCreateDoubleScalar(double);
mxArray* mxCreateNumericArray(mwSize, const mwSize*, mxClassID, mxComplexity);
mxArray* mxCreateString(const char*);
void mxDestroyArray(mxArray*);
}
