// MC3Dmex on the MI355X engine: the reference gateway's 18-argument
// signature and outputs (cpp/3d/MC3Dmex.cpp:99-254); see vmc_mex.hpp.
//   mex -R2018a MC3Dmex_gpu.cpp -I<repo>/include -I<repo>/valomc_amd/host -L<repo>/valomc_amd -lvalomc
#include "vmc_mex.hpp"

void mexFunction(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
  vmc_mex::gateway(3, nlhs, plhs, nrhs, prhs);
}
