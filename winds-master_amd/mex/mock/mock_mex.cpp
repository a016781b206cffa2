// mock_mex.cpp -- in-process stand-in for MATLAB's MEX runtime, so tests can
// call the gateway (BiotSavartMex.cpp) from Python through ctypes.
// mexErrMsgTxt throws; mock_call catches it and returns the message.
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls;
    std::vector<mwSize> dims;
    std::vector<double> data;
    std::string str;
};

namespace {
struct MexError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
}  // namespace

extern "C" {

mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->dims.size(); }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims.data(); }
size_t mxGetNumberOfElements(const mxArray *a) {
    size_t n = 1;
    for (mwSize d : a->dims) n *= d;
    return n;
}
double *mxGetPr(const mxArray *a) { return const_cast<double *>(a->data.data()); }
bool mxIsChar(const mxArray *a) { return a->cls == mxCHAR_CLASS; }
bool mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsComplex(const mxArray *) { return false; }
int mxGetString(const mxArray *a, char *buf, mwSize buflen) {
    if (a->cls != mxCHAR_CLASS || buflen == 0) return 1;
    std::strncpy(buf, a->str.c_str(), buflen - 1);
    buf[buflen - 1] = 0;
    return a->str.size() + 1 > buflen ? 1 : 0;
}
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID c, mxComplexity) {
    mxArray *a = new mxArray_tag();
    a->cls = c;
    a->dims.assign(dims, dims + ndim);
    a->data.assign(mxGetNumberOfElements(a), 0.0);
    return a;
}
void mxDestroyArray(mxArray *a) { delete a; }
void mexErrMsgTxt(const char *msg) { throw MexError(msg); }

// ---- test driver entry points
mxArray *mock_double(int ndim, const int64_t *dims, const double *data) {
    std::vector<mwSize> d(dims, dims + ndim);
    mxArray *a = mxCreateNumericArray(ndim, d.data(), mxDOUBLE_CLASS, mxREAL);
    if (data) std::memcpy(a->data.data(), data, a->data.size() * sizeof(double));
    return a;
}
mxArray *mock_string(const char *s) {
    mxArray *a = new mxArray_tag();
    a->cls = mxCHAR_CLASS;
    a->str = s;
    a->dims = {1, std::strlen(s)};
    return a;
}
int mock_ndims(const mxArray *a) { return (int)a->dims.size(); }
void mock_dims(const mxArray *a, int64_t *out) {
    for (size_t i = 0; i < a->dims.size(); ++i) out[i] = (int64_t)a->dims[i];
}
const double *mock_data(const mxArray *a) { return a->data.data(); }
void mock_free(mxArray *a) { delete a; }

// 0 on success; else 1 with the error message in msg
int mock_call(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *msg, int len) {
    try {
        mexFunction(nlhs, plhs, nrhs, prhs);
        return 0;
    } catch (const std::exception &e) {
        std::strncpy(msg, e.what(), len - 1);
        msg[len - 1] = 0;
        return 1;
    }
}

}  // extern "C"
