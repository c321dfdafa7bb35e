/* Test harness only: the subset of MATLAB's MEX API (R2018a interleaved
   complex) that valomc_amd/host/vmc_mex.hpp uses, so the gateways compile
   and run without MATLAB.  mexErrMsgTxt throws instead of unwinding into
   MATLAB; the base workspace holds one logical, abort_photonMC, whose value
   the test program (driver.cpp) can script. */
#ifndef VMC_TEST_MEX_MOCK_H
#define VMC_TEST_MEX_MOCK_H

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

typedef size_t mwSize;
typedef char16_t mxChar;
typedef double mxDouble;
typedef int64_t mxInt64;
typedef uint64_t mxUint64;
typedef struct { double real, imag; } mxComplexDouble;
typedef enum {
  mxUNKNOWN_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS,
  mxINT16_CLASS, mxUINT16_CLASS, mxINT32_CLASS, mxUINT32_CLASS, mxINT64_CLASS, mxUINT64_CLASS
} mxClassID;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;

struct mxArray {
  mxClassID id = mxDOUBLE_CLASS;
  size_t m = 0, n = 0;
  bool complex = false;
  std::vector<unsigned char> data;   /* column-major, interleaved when complex */
};

size_t mxElementSize(mxClassID id);
mxArray *mxCreateArray(mxClassID id, size_t m, size_t n, bool complex);

inline size_t mxGetM(const mxArray *a) { return a->m; }
inline size_t mxGetN(const mxArray *a) { return a->n; }
inline size_t mxGetNumberOfElements(const mxArray *a) { return a->m * a->n; }
inline mwSize mxGetNumberOfDimensions(const mxArray *) { return 2; }
inline mxClassID mxGetClassID(const mxArray *a) { return a->id; }
inline bool mxIsEmpty(const mxArray *a) { return a->m * a->n == 0; }
inline bool mxIsChar(const mxArray *a) { return a->id == mxCHAR_CLASS; }
inline void *mxGetData(const mxArray *a) { return const_cast<unsigned char *>(a->data.data()); }
inline mxChar *mxGetChars(const mxArray *a) { return static_cast<mxChar *>(mxGetData(a)); }
inline mxDouble *mxGetDoubles(const mxArray *a) { return static_cast<mxDouble *>(mxGetData(a)); }
inline mxInt64 *mxGetInt64s(const mxArray *a) { return static_cast<mxInt64 *>(mxGetData(a)); }
inline mxUint64 *mxGetUint64s(const mxArray *a) { return static_cast<mxUint64 *>(mxGetData(a)); }
inline mxComplexDouble *mxGetComplexDoubles(const mxArray *a) { return static_cast<mxComplexDouble *>(mxGetData(a)); }
double mxGetScalar(const mxArray *a);
inline mxArray *mxCreateDoubleMatrix(size_t m, size_t n, mxComplexity c) {
  return mxCreateArray(mxDOUBLE_CLASS, m, n, c == mxCOMPLEX);
}
inline mxArray *mxCreateNumericMatrix(size_t m, size_t n, mxClassID id, mxComplexity c) {
  return mxCreateArray(id, m, n, c == mxCOMPLEX);
}
mxArray *mxCreateDoubleScalar(double v);
void mxDestroyArray(mxArray *a);
bool mxIsLogicalScalarTrue(const mxArray *a);

void mexErrMsgTxt(const char *msg);
int mexPrintf(const char *fmt, ...);
int mexEvalString(const char *cmd);
mxArray *mexGetVariable(const char *workspace, const char *name);

#endif
