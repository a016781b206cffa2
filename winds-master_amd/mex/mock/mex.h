// mex.h -- the subset of MATLAB's MEX API that BiotSavartMex.cpp uses, for
// building the gateway against mock_mex.cpp in tests (MATLAB and Octave are
// not available here).  A real build uses MATLAB's or Octave's own mex.h.
#pragma once

#include <cstddef>

typedef size_t mwSize;
typedef struct mxArray_tag mxArray;
typedef enum { mxDOUBLE_CLASS = 6, mxCHAR_CLASS = 4 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

extern "C" {
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
double *mxGetPr(const mxArray *a);
bool mxIsChar(const mxArray *a);
bool mxIsDouble(const mxArray *a);
bool mxIsComplex(const mxArray *a);
int mxGetString(const mxArray *a, char *buf, mwSize buflen);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID c, mxComplexity f);
void mxDestroyArray(mxArray *a);
void mexErrMsgTxt(const char *msg);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
