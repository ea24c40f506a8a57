// MATLAB MEX gateway for phase unwrapping onto libcbs_hip.so (cbs_phase_unwrap).
//
// Drop-in for unwrapp2_gpu (Codes/sub_code/goldstein_GPU/unwrapp2_gpu.m:1-5), which
// BACKWARD_SOLVER_RYTOV.m:49 calls on the wrapped phase of every angle:
//   unwrapped = unwrapp2_mi355x(single(gather(wrapped_phase)));
// takes a real single (n1, n2[, n3]) host array and returns the unwrapped phases, same
// size and class: residues, the reference's Goldstein branch cuts (dipoles, nearest lists,
// branch search, raster) and the raster integration, on GPU 0 (the device index may be
// given as a second argument).
//
// Build (with MATLAB R2018a+):  mex -R2018a unwrapp2_mi355x.cpp -I../../include -L.. -lcbs_hip
#include <cstdint>

#include "cbs_hip.h"
#include "mex.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 1 || nrhs > 2) mexErrMsgIdAndTxt("cbs:arg", "unwrapp2_mi355x(wrapped_phase[, device])");
  if (nlhs > 1) mexErrMsgIdAndTxt("cbs:arg", "one output");
  const mxArray* w = prhs[0];
  if (!mxIsSingle(w) || mxIsComplex(w)) mexErrMsgIdAndTxt("cbs:arg", "the wrapped phase must be real single");
  const mwSize nd = mxGetNumberOfDimensions(w);
  const mwSize* d = mxGetDimensions(w);
  if (nd > 3) mexErrMsgIdAndTxt("cbs:arg", "phase maps are (n1, n2[, n3])");
  const uint32_t dims[3] = {(uint32_t)d[0], (uint32_t)d[1], nd == 3 ? (uint32_t)d[2] : 1u};
  int device = 0;
  if (nrhs == 2) device = (int)mxGetScalar(prhs[1]);
  plhs[0] = mxCreateNumericArray(nd, d, mxSINGLE_CLASS, mxREAL);
  if (cbs_phase_unwrap(device, mxGetSingles(w), mxGetSingles(plhs[0]), dims, CBS_MEM_HOST) != CBS_OK)
    mexErrMsgIdAndTxt("cbs:lib", "%s", cbs_last_error());
}
