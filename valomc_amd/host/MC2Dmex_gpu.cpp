// MC2Dmex on the MI355X engine: the reference gateway's 19-argument
// signature, GaussianSigma 17th, and outputs (cpp/2d/MC2Dmex.cpp:98-252);
// see vmc_mex.hpp.
//   mex -R2018a MC2Dmex_gpu.cpp -I<repo>/include -I<repo>/valomc_amd/host -L<repo>/valomc_amd -lvalomc
#include "vmc_mex.hpp"

void mexFunction(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
  vmc_mex::gateway(2, nlhs, plhs, nrhs, prhs);
}
