// Declarations of the MATLAB C API functions the MEX shim calls, for a
// syntax check of the shim where MATLAB is not installed (not MATLAB's header).
#pragma once
#include <cstddef>
#include <cstdint>
typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;
typedef enum { mxDOUBLE_CLASS, mxSINGLE_CLASS } mxClassID;
#define mxSINGLE mxSINGLE_CLASS
typedef struct { float real, imag; } mxComplexSingle;
typedef float mxSingle;
typedef double mxDouble;
extern "C" {
void mexErrMsgIdAndTxt(const char*, const char*, ...);
int mexCallMATLAB(int, mxArray**, int, mxArray**, const char*);
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
bool mxIsComplex(const mxArray*);
bool mxIsSingle(const mxArray*);
mxArray* mxCreateDoubleMatrix(mwSize, mwSize, mxComplexity);
mxArray* mxCreateDoubleScalar(double);
mxArray* mxCreateNumericArray(mwSize, const mwSize*, mxClassID, mxComplexity);
mxArray* mxCreateString(const char*);
void mxDestroyArray(mxArray*);
}
