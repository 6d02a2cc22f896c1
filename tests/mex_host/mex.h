/* Test host for the MATLAB gateway multem_amd/mex/ilc_multem_hip.cpp: the
   small part of the MATLAB C matrix API (R2017b separate-complex storage)
   that the gateway calls, implemented in mex_host.cpp so that the gateway
   can be compiled and driven from Python without MATLAB.  Not part of the
   library. */
#ifndef MEX_HOST_MEX_H
#define MEX_HOST_MEX_H

#include <stddef.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

#ifdef __cplusplus
extern "C" {
#endif

mxArray *mxGetField(const mxArray *s, mwIndex idx, const char *name);
void mxSetField(mxArray *s, mwIndex idx, const char *name, mxArray *v);
bool mxIsEmpty(const mxArray *a);
bool mxIsStruct(const mxArray *a);
bool mxIsComplex(const mxArray *a);
double mxGetScalar(const mxArray *a);
double *mxGetPr(const mxArray *a);
double *mxGetPi(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
size_t mxGetM(const mxArray *a);
size_t mxGetN(const mxArray *a);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char **names);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);

#ifdef __cplusplus
}
#endif

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#endif
