/* Declarations of the MATLAB MEX C API (R2018a interleaved-complex form) that
 * null--space-free-6d-projection-method_amd/mex/fame6d_mex.c uses, so that the gateway can be
 * compiled for syntax and type checking where MATLAB is not installed.  Signatures follow the
 * public MATLAB External API documentation; nothing here is linked or run. */
#ifndef FAME6D_TEST_MEX_H
#define FAME6D_TEST_MEX_H
#include <stddef.h>
#include <stdint.h>

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxUINT64_CLASS = 15 } mxClassID;
typedef struct { double real, imag; } mxComplexDouble;
typedef double mxDouble;
typedef uint64_t mxUint64;

void mexErrMsgIdAndTxt(const char* errorid, const char* errormsg, ...);
void* mxCalloc(mwSize n, mwSize size);
void mxFree(void* ptr);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity complexity);
mxArray* mxCreateDoubleScalar(double value);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity complexity);
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** fieldnames);
void mxSetField(mxArray* pm, mwSize index, const char* fieldname, mxArray* pvalue);
mxComplexDouble* mxGetComplexDoubles(const mxArray* pa);
mxDouble* mxGetDoubles(const mxArray* pa);
mxUint64* mxGetUint64s(const mxArray* pa);
size_t mxGetM(const mxArray* pm);
size_t mxGetN(const mxArray* pm);
size_t mxGetNumberOfElements(const mxArray* pm);
double mxGetScalar(const mxArray* pm);
int mxGetString(const mxArray* pm, char* str, mwSize strlen);
int mxIsComplex(const mxArray* pm);
int mxIsUint64(const mxArray* pm);
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#endif
