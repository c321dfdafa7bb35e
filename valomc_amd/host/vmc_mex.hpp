// MEX gateway body shared by MC3Dmex_gpu.cpp and MC2Dmex_gpu.cpp.
//
// Keeps the reference gateways' argument lists and outputs
// (cpp/3d/MC3Dmex.cpp:99-254, cpp/2d/MC2Dmex.cpp:98-252):
//
//   [vsol, bsol, ebsol, simulationtime, rnseed, [HN]] =
//       MC3Dmex(H, HN, BH, r, BCType, BCIntensity, BCLightDirectionType,
//               BCLNormal, BCn, mua, mus, g, n, f, phase0, Nphoton,
//               disablepbar, rnseed)                               (18 inputs)
//       MC2Dmex(... Nphoton, GaussianSigma, disablepbar, rndseed)  (19 inputs)
//
// with the classes ValoMC.m passes (ValoMC.m:128-136, 262-265, 338): int64
// H/HN/BH (0-based, HN may be empty), double r and parameters, int8 or char
// BCType and BCLightDirectionType, int64 Nphoton and disablepbar, uint64
// rnseed = [seed, use-seed flag].  vsol/bsol/ebsol are complex columns
// (ER + i EI, EBR + i EBI, DEBR + i DEBI), simulationtime is in seconds,
// rnseed returns the seed used, HN the neighbour table.
//
// The solve runs on the GPU through the C ABI: vmc_monte_carlo when the
// progress bar is disabled, vmc_problem_create + vmc_problem_run with a
// progress callback that drives the waitbar and its Cancel button otherwise
// (an aborted run is still normalised by the planned Nphoton, as in the
// reference).  Transport is FP64, the reference's arithmetic; the
// environment variable VALOMC_PRECISION=fp32 selects FP32 transport.
//
// Built on a MATLAB host with (one line)
//   mex -R2018a MC3Dmex_gpu.cpp -I<repo>/include -I<repo>/valomc_amd/host
//       -L<repo>/valomc_amd -lvalomc
#ifndef VMC_HOST_VMC_MEX_HPP
#define VMC_HOST_VMC_MEX_HPP

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"
#include "valomc.h"

namespace vmc_mex {

inline const char *syntax(int dim) {
  return dim == 3
             ? "Syntax:\n [vsol, bsol, ebsol, simulationtime, rnseed, [HN]] = MC3Dmex(H, HN, BH, r, BCType, "
               "BCIntensity, BCLightDirectionType, BCLNormal, BCn, mua, mus, g, n, f, phase0, Nphoton, disablepbar, "
               "rnseed)\n"
             : "Syntax:\n [vsol, bsol, ebsol, simulationtime, seed, [HN]] = MC2Dmex(H, HN, BH, r, BCType, "
               "BCIntensity, BCLightDirectionType, BCLNormal, BCn, mua, mus, g, n, f, phase0, Nphoton, "
               "GaussianSigma, disablepbar, rndseed)\n";
}

// an input of class `id` with `cols` columns (0: any) and `rows` rows (-1:
// any); empty inputs are accepted where `optional`.  Raises a MATLAB error
// naming the argument otherwise.
inline const mxArray *arg(const mxArray **prhs, int k, const char *name, mxClassID id, int64_t rows, int64_t cols,
                          bool optional) {
  const mxArray *a = prhs[k];
  if (mxIsEmpty(a)) {
    if (optional) return nullptr;
    mexErrMsgTxt((std::string(name) + " must not be empty\n").c_str());
  }
  const mxClassID got = mxGetClassID(a);
  const bool chars_ok = id == mxINT8_CLASS && (got == mxCHAR_CLASS || got == mxUINT8_CLASS);
  if (got != id && !chars_ok) mexErrMsgTxt((std::string("mxArray type error! Incorrect type received for ") + name + "\n").c_str());
  if (mxGetNumberOfDimensions(a) != 2) mexErrMsgTxt("Only 2D-mxArrays are supported!\n");
  if ((rows >= 0 && (int64_t)mxGetM(a) != rows) || (cols > 0 && (int64_t)mxGetN(a) != cols))
    mexErrMsgTxt((std::string(name) + " has the wrong size\n").c_str());
  return a;
}

// characters of an int8 / uint8 / char array (MATLAB chars are 16-bit)
inline std::vector<char> chars(const mxArray *a) {
  std::vector<char> out;
  if (!a) return out;
  const size_t n = mxGetNumberOfElements(a);
  out.resize(n);
  if (mxIsChar(a)) {
    const mxChar *c = mxGetChars(a);
    for (size_t i = 0; i < n; ++i) out[i] = (char)c[i];
  } else {
    std::memcpy(out.data(), mxGetData(a), n);
  }
  return out;
}

inline const double *doubles(const mxArray *a) { return a ? mxGetDoubles(a) : nullptr; }
inline const int64_t *int64s(const mxArray *a) { return a ? reinterpret_cast<const int64_t *>(mxGetInt64s(a)) : nullptr; }

// complex column re + i im
inline mxArray *complex_column(const std::vector<double> &re, const std::vector<double> &im) {
  mxArray *a = mxCreateDoubleMatrix(re.size(), 1, mxCOMPLEX);
  mxComplexDouble *z = mxGetComplexDoubles(a);
  for (size_t i = 0; i < re.size(); ++i) {
    z[i].real = re[i];
    z[i].imag = im[i];
  }
  return a;
}

// Progress with the waitbar and its Cancel button (MC3Dmex.cpp:48-91):
// the button sets abort_photonMC in the base workspace
struct Bar {
  std::chrono::steady_clock::time_point t0;
  static int call(double pct, void *user) {
    Bar *b = static_cast<Bar *>(user);
    mxArray *flag = mexGetVariable("base", "abort_photonMC");
    if (flag) {
      const bool stop = mxIsLogicalScalarTrue(flag);
      mxDestroyArray(flag);
      if (stop) return 0;
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - b->t0).count();
    char cmd[256];
    if (el > 0 && pct > 0) {
      const double left = (100.0 - pct) / (pct / el);
      const int h = (int)(left / 3600), m = (int)((left - 3600 * h) / 60), s = (int)std::ceil(left - 3600 * h - 60 * m);
      std::snprintf(cmd, sizeof cmd, "waitbar(%f,mcwaitbar,'%i hours %i minutes and %i seconds left');", pct / 100.0, h,
                    m, s);
    } else {
      std::snprintf(cmd, sizeof cmd, "waitbar(0, mcwaitbar,'Estimating the time left');");
    }
    mexEvalString(cmd);
    return 1;
  }
};

inline void gateway(int dim, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
  const int nin = dim == 3 ? 18 : 19;
  if (nrhs != nin || (nlhs != 5 && nlhs != 6)) {
    mexPrintf("nrhs %i nlhs %i", nrhs, nlhs);
    mexErrMsgTxt(syntax(dim));
  }
  const int nv = dim + 1;
  const mxArray *H = arg(prhs, 0, "H", mxINT64_CLASS, -1, nv, false);
  const int64_t Ne = (int64_t)mxGetM(H);
  const mxArray *HN = arg(prhs, 1, "HN", mxINT64_CLASS, Ne, nv, true);
  const mxArray *BH = arg(prhs, 2, "BH", mxINT64_CLASS, -1, dim, false);
  const int64_t Nb = (int64_t)mxGetM(BH);
  const mxArray *r = arg(prhs, 3, "r", mxDOUBLE_CLASS, -1, dim, false);
  const int64_t Nr = (int64_t)mxGetM(r);
  const std::vector<char> bct = chars(arg(prhs, 4, "BCType", mxINT8_CLASS, -1, 0, false));
  const mxArray *BCIntensity = arg(prhs, 5, "BCIntensity", mxDOUBLE_CLASS, -1, 0, true);
  const std::vector<char> ldt = chars(arg(prhs, 6, "BCLightDirectionType", mxINT8_CLASS, -1, 0, true));
  const mxArray *BCLNormal = arg(prhs, 7, "BCLNormal", mxDOUBLE_CLASS, Nb, dim, true);
  const mxArray *BCn = arg(prhs, 8, "BCn", mxDOUBLE_CLASS, -1, 0, false);
  const mxArray *mua = arg(prhs, 9, "mua", mxDOUBLE_CLASS, -1, 0, false);
  const mxArray *mus = arg(prhs, 10, "mus", mxDOUBLE_CLASS, -1, 0, false);
  const mxArray *g = arg(prhs, 11, "g", mxDOUBLE_CLASS, -1, 0, false);
  const mxArray *n = arg(prhs, 12, "n", mxDOUBLE_CLASS, -1, 0, false);
  const double f = mxGetScalar(arg(prhs, 13, "f", mxDOUBLE_CLASS, -1, 0, false));
  const double phase0 = mxGetScalar(arg(prhs, 14, "phase0", mxDOUBLE_CLASS, -1, 0, false));
  const int64_t Nphoton = int64s(arg(prhs, 15, "Nphoton", mxINT64_CLASS, -1, 0, false))[0];
  const mxArray *GaussianSigma = dim == 2 ? arg(prhs, 16, "GaussianSigma", mxDOUBLE_CLASS, -1, 0, true) : nullptr;
  const int k0 = dim == 3 ? 16 : 17;
  const int64_t disable_pbar = int64s(arg(prhs, k0, "disablepbar", mxINT64_CLASS, -1, 0, false))[0];
  const mxArray *seedarg = arg(prhs, k0 + 1, "rnseed", mxUINT64_CLASS, -1, 0, false);
  if (mxGetNumberOfElements(seedarg) < 2) mexErrMsgTxt("rnseed must hold [seed, use-seed flag]\n");
  // every per-element and per-face array must match the mesh (the sizes
  // MC3D::ErrorChecks compares, MC3D.hpp:460-571)
  struct Sized {
    const mxArray *a;
    int64_t want;
    int err;   // status of the reference's mcerror for it
  };
  const Sized sized[] = {{g, Ne, 1 + 1}, {mua, Ne, 1 + 2}, {mus, Ne, 1 + 3}, {n, Ne, 1 + 4}, {BCn, Nb, 1 + 5},
                         {BCIntensity, Nb, VMC_ERR_INVALID_ARGUMENT}, {GaussianSigma, Nb, VMC_ERR_INVALID_ARGUMENT}};
  for (const Sized &s : sized)
    if (s.a && (int64_t)mxGetNumberOfElements(s.a) != s.want)
      mexErrMsgTxt((std::string("Error in initializing MC") + (dim == 3 ? "3D" : "2D") + ": " +
                    vmc_status_string(s.err) + "\n")
                       .c_str());
  if ((int64_t)bct.size() != Nb) mexErrMsgTxt(vmc_status_string(1 + 7));
  if (!ldt.empty() && (int64_t)ldt.size() != Nb) mexErrMsgTxt(vmc_status_string(1 + 9));

  const uint64_t *rs = reinterpret_cast<const uint64_t *>(mxGetUint64s(seedarg));
  const uint64_t seed =
      rs[1] ? rs[0] : (uint64_t)std::chrono::high_resolution_clock::now().time_since_epoch().count();
  const char *pe = std::getenv("VALOMC_PRECISION");
  const int precision = (pe && std::strcmp(pe, "fp32") == 0) ? VMC_FP32 : VMC_FP64;

  std::vector<double> ER(Ne), EI(Ne), EBR(Nb), EBI(Nb), DEBR(Nb), DEBI(Nb);
  std::vector<int64_t> hn((size_t)Ne * nv);
  int64_t loss = 0;
  const auto t0 = std::chrono::steady_clock::now();
  int st;
  mexPrintf("Computing... \n");
  if (disable_pbar != 0) {
    st = vmc_monte_carlo(dim, Ne, Nb, Nr, int64s(H), int64s(HN), int64s(BH), doubles(r), bct.data(),
                         doubles(BCIntensity), ldt.empty() ? nullptr : ldt.data(), doubles(BCLNormal), doubles(BCn),
                         doubles(mua), doubles(mus), doubles(g), doubles(n), f, phase0, Nphoton,
                         doubles(GaussianSigma), seed, precision, ER.data(), EI.data(), EBR.data(), EBI.data(),
                         DEBR.data(), DEBI.data(), hn.data(), &loss);
  } else {
    vmc_problem_desc d = {};
    d.dim = dim;
    d.Ne = Ne;
    d.Nb = Nb;
    d.Nr = Nr;
    d.H = int64s(H);
    d.HN = int64s(HN);
    d.BH = int64s(BH);
    d.r = doubles(r);
    d.mua = doubles(mua);
    d.mus = doubles(mus);
    d.g = doubles(g);
    d.n = doubles(n);
    d.BCType = bct.data();
    d.BCn = doubles(BCn);
    d.BCLNormal = doubles(BCLNormal);
    d.BCLightDirectionType = ldt.empty() ? nullptr : ldt.data();
    d.BCIntensity = doubles(BCIntensity);
    d.GaussianSigma = doubles(GaussianSigma);
    vmc_problem *prob = nullptr;
    st = vmc_problem_create(&d, &prob);
    if (st == VMC_OK) {
      vmc_run_desc run = {};
      run.f = f;
      run.phase0 = phase0;
      run.Nphoton = Nphoton;
      run.seed = seed;
      run.precision = precision;
      vmc_output out = {ER.data(), EI.data(), EBR.data(), EBI.data(), DEBR.data(), DEBI.data()};
      vmc_stats stats = {};
      mexEvalString("assignin('base','abort_photonMC', false);");
      mexEvalString("mcwaitbar = waitbar(0,'Please wait..', 'name', 'Running simulation', "
                    "'CreateCancelBtn','abort_photonMC=true;');");
      Bar bar{t0};
      st = vmc_problem_run(prob, &run, &out, &stats, &Bar::call, &bar);
      loss = stats.loss;
      if (st == VMC_OK || st == VMC_ERR_ABORTED) (void)vmc_problem_get_hn(prob, hn.data());
      vmc_problem_destroy(prob);
      mexEvalString("delete(mcwaitbar);");
      if (st == VMC_ERR_ABORTED) {
        mexPrintf("WARNING: RUN WAS ABORTED OR PARALLEL COMPUTING ENVIRONMENT IS NOT WORKING CORRECTLY. \n");
        st = VMC_OK;
      }
    }
  }
  if (st != VMC_OK) {
    std::string msg = std::string("Error in initializing MC") + (dim == 3 ? "3D" : "2D") + ": " +
                      vmc_status_string(st);
    const char *detail = vmc_last_error();
    if (st >= 100 && detail && *detail) msg += std::string(" (") + detail + ")";
    mexErrMsgTxt((msg + "\n").c_str());
  }
  mexPrintf("...done\n");
  if (loss) mexPrintf(" %lld photons lost during computation!\n", (long long)loss);
  plhs[0] = complex_column(ER, EI);
  plhs[1] = complex_column(EBR, EBI);
  plhs[2] = complex_column(DEBR, DEBI);
  plhs[3] = mxCreateDoubleScalar(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  plhs[4] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  *reinterpret_cast<uint64_t *>(mxGetData(plhs[4])) = seed;
  if (nlhs == 6) {
    plhs[5] = mxCreateNumericMatrix((size_t)Ne, (size_t)nv, mxINT64_CLASS, mxREAL);
    std::memcpy(mxGetData(plhs[5]), hn.data(), hn.size() * sizeof(int64_t));
  }
  mexPrintf("Done\n");
}

} // namespace vmc_mex

#endif
