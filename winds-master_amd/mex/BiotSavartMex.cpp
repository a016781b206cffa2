// BiotSavartMex.cpp -- MATLAB/Octave gateway of the engine.
//
// Same call as the reference MEX (cuda/source/BiotSavartMex.cu:688-807):
//
//   [uind, L] = BiotSavartMex(F1, F2, P, gamma, rc, d, cmod, co, type, gpu, gpuhw)
//
// with exactly 11 inputs and 2 outputs (:720-723), P padded to 4-D (:745-764),
// uind of P's padded dims, and the work done by winds_bs_eval (include/
// winds_bs.h), which evaluates on the GPU.  Differences, each deliberate:
// L has the dims of F ([ns,1,nt,nb], as BiotSavart.m:71-76 returns it) rather
// than P's; F1 and F2 of fewer than 4 dims are padded instead of being read
// past their dimension array; a character d (a 'viscN' string that reached
// the MEX unparsed) is an error instead of being read as doubles; failures
// free everything before mexErrMsgTxt.
//
// Build: mex -I<repo>/include BiotSavartMex.cpp -L<repo>/winds-master_amd/lib -lwinds_bs
//        (Octave: mkoctfile --mex with the same flags).

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"
#include "winds_bs.h"

namespace {

// 2-D and 3-D arrays are padded with trailing unit dims
bool dims4(const mxArray *a, int64_t out[4]) {
    const mwSize nd = mxGetNumberOfDimensions(a);
    const mwSize *d = mxGetDimensions(a);
    if (nd < 2 || nd > 4) return false;
    for (int i = 0; i < 4; ++i) out[i] = i < (int)nd ? (int64_t)d[i] : 1;
    return true;
}

bool get_string(const mxArray *a, std::string &s) {
    if (!mxIsChar(a)) return false;
    std::vector<char> buf(mxGetNumberOfElements(a) + 1, 0);
    if (mxGetString(a, buf.data(), buf.size())) return false;
    s.assign(buf.data());
    return true;
}

bool is_real_double(const mxArray *a) { return mxIsDouble(a) && !mxIsComplex(a); }

// Returns "" on success, else the message for mexErrMsgTxt.
std::string run(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs != 11) return "11 inputs expected.";
    if (nlhs != 2) return "2 outputs expected.";
    const mxArray *F1 = prhs[0], *F2 = prhs[1], *P = prhs[2], *G = prhs[3], *RC = prhs[4];
    const mxArray *D = prhs[5], *CM = prhs[6], *CO = prhs[7], *TY = prhs[8], *GPU = prhs[9],
                  *HW = prhs[10];
    std::string cmod, type, gpu;
    if (!get_string(CM, cmod)) return "Couldn't make Core Model string.";
    if (!get_string(TY, type)) return "Couldn't make TYPE string.";
    if (!get_string(GPU, gpu)) return "Couldn't make GPU string.";
    if (mxIsChar(D))
        return "d must be numeric: split 'viscN' into cmod 'visc' and d = N (WInDS.m:84-91).";
    for (const mxArray *a : {F1, F2, P, G, RC, D, CO, HW})
        if (!is_real_double(a)) return "numeric inputs must be real double arrays.";
    if (mxGetNumberOfElements(D) < 1 || mxGetNumberOfElements(CO) < 1)
        return "d and co must be scalars.";
    if (mxGetNumberOfDimensions(P) == 1) return "Need more dimensions in P.";
    int64_t pd[4], fd[4];
    if (!dims4(P, pd)) return "Wrong dimensions in P.";
    if (!dims4(F1, fd)) return "Wrong dimensions in F1.";
    int64_t f2d[4];
    if (!dims4(F2, f2d) || std::memcmp(fd, f2d, sizeof(fd)) != 0)
        return "F1 and F2 must have the same dims.";
    // the engine reads ns*nt*nb values of gamma and of rc (one per segment)
    const size_t nf = (size_t)(fd[0] * fd[2] * fd[3]);
    if (mxGetNumberOfElements(G) != nf) return "gamma must hold one value per segment.";
    const size_t nrc = mxGetNumberOfElements(RC);
    if (nrc != 0 && nrc != nf) return "rc must be empty or hold one value per segment.";
    double hw[3] = {0.0, 0.0, 0.0};
    const size_t nhw = mxGetNumberOfElements(HW);
    for (size_t i = 0; i < 3 && i < nhw; ++i) hw[i] = mxGetPr(HW)[i];

    mwSize udims[4] = {(mwSize)pd[0], (mwSize)pd[1], (mwSize)pd[2], (mwSize)pd[3]};
    mwSize ldims[4] = {(mwSize)fd[0], 1, (mwSize)fd[2], (mwSize)fd[3]};
    mxArray *U = mxCreateNumericArray(4, udims, mxDOUBLE_CLASS, mxREAL);
    mxArray *L = mxCreateNumericArray(4, ldims, mxDOUBLE_CLASS, mxREAL);
    if (!U || !L) {
        if (U) mxDestroyArray(U);
        if (L) mxDestroyArray(L);
        return "out of host memory";
    }
    const int rc = winds_bs_eval(mxGetPr(F1), mxGetPr(F2), fd, mxGetPr(P), pd, mxGetPr(G),
                                 nrc ? mxGetPr(RC) : nullptr, mxGetPr(D)[0], cmod.c_str(),
                                 mxGetPr(CO)[0],
                                 type.c_str(), gpu.c_str(), hw, mxGetPr(U), mxGetPr(L));
    if (rc != WINDS_OK) {
        mxDestroyArray(U);
        mxDestroyArray(L);
        return std::string("BiotSavartMex: ") + winds_last_error();
    }
    plhs[0] = U;
    plhs[1] = L;
    return std::string();
}

}  // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    static char msg[512];
    {
        const std::string m = run(nlhs, plhs, nrhs, prhs);
        if (m.empty()) return;
        std::strncpy(msg, m.c_str(), sizeof(msg) - 1);
        msg[sizeof(msg) - 1] = 0;
    }  // every local is gone before the error unwinds
    mexErrMsgTxt(msg);
}
