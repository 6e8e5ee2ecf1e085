/* Test double of the MATLAB MEX API -- TEST INFRASTRUCTURE ONLY.
 *
 * Declares the part of the documented MEX C API (interleaved complex,
 * R2018a) that comp_ui_toolbox_amd/mex/combined_quick_gpu.cpp calls, so the
 * shim can be compiled and exercised without MATLAB.  mock_mex.cpp
 * implements it over a small in-memory object model. */
#ifndef MEX_MOCK_H
#define MEX_MOCK_H
#include <stddef.h>

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct { double real, imag; } mxComplexDouble;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

#ifdef __cplusplus
extern "C" {
#endif
int mexAtExit(void (*fn)(void));
#ifdef __cplusplus
[[noreturn]]
#endif
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
mxArray* mxGetProperty(const mxArray* obj, mwIndex index, const char* name);
mxArray* mxGetCell(const mxArray* cell, mwIndex index);
bool mxIsDouble(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsNumeric(const mxArray* a);
bool mxIsClass(const mxArray* a, const char* name);
bool mxIsChar(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize buflen);
size_t mxGetNumberOfElements(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
double mxGetScalar(const mxArray* a);
double* mxGetPr(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
#ifdef __cplusplus
}
#endif

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
#endif
