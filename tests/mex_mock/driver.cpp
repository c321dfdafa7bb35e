// Test harness only: runs a MEX gateway's mexFunction on arrays read from
// files, against the mock MEX API of mex.h.
//   driver <manifest> <outdir> <nlhs>
// The manifest lists the right-hand sides in order, one per line:
//   <class> <rows> <cols> <file>      class: double int64 uint64 int8 char;
//                                     an empty array is "<class> 0 0 -"
// Outputs go to <outdir>/out<k>.bin (complex: re then im), with a line
//   out <k> <class> <rows> <cols> <complex>
// per output on stdout, the waitbar commands as "eval <cmd>" lines, and
// "error <message>" (on one line) plus exit status 3 when the gateway
// raises an error.
// VMC_MOCK_CANCEL_AFTER=k makes the Cancel button be pressed after k
// progress calls.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

void mexFunction(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs);

size_t mxElementSize(mxClassID id) {
  switch (id) {
  case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
  case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
  case mxCHAR_CLASS: case mxINT16_CLASS: case mxUINT16_CLASS: return 2;
  default: return 1;
  }
}
mxArray *mxCreateArray(mxClassID id, size_t m, size_t n, bool complex) {
  mxArray *a = new mxArray;
  a->id = id;
  a->m = m;
  a->n = n;
  a->complex = complex;
  a->data.assign(m * n * mxElementSize(id) * (complex ? 2 : 1), 0);
  return a;
}
double mxGetScalar(const mxArray *a) {
  if (mxIsEmpty(a)) return 0.0;
  switch (a->id) {
  case mxDOUBLE_CLASS: return mxGetDoubles(a)[0];
  case mxINT64_CLASS: return (double)mxGetInt64s(a)[0];
  case mxUINT64_CLASS: return (double)mxGetUint64s(a)[0];
  case mxINT8_CLASS: return (double)((const signed char *)mxGetData(a))[0];
  default: return 0.0;
  }
}
mxArray *mxCreateDoubleScalar(double v) {
  mxArray *a = mxCreateArray(mxDOUBLE_CLASS, 1, 1, false);
  mxGetDoubles(a)[0] = v;
  return a;
}
void mxDestroyArray(mxArray *a) { delete a; }
bool mxIsLogicalScalarTrue(const mxArray *a) {
  return a->id == mxLOGICAL_CLASS && a->m * a->n == 1 && a->data[0] != 0;
}

struct GatewayError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
void mexErrMsgTxt(const char *msg) { throw GatewayError(msg); }
int mexPrintf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const int r = vfprintf(stderr, fmt, ap);
  va_end(ap);
  return r;
}

static bool g_abort = false;
static int g_polls = 0;
int mexEvalString(const char *cmd) {
  std::cout << "eval " << cmd << "\n";
  if (std::strstr(cmd, "assignin('base','abort_photonMC', false)")) g_abort = false;
  return 0;
}
mxArray *mexGetVariable(const char *workspace, const char *name) {
  if (std::strcmp(workspace, "base") != 0 || std::strcmp(name, "abort_photonMC") != 0) return nullptr;
  ++g_polls;
  const char *after = std::getenv("VMC_MOCK_CANCEL_AFTER");
  if (after && g_polls > std::atoi(after)) g_abort = true;
  mxArray *a = mxCreateArray(mxLOGICAL_CLASS, 1, 1, false);
  a->data[0] = g_abort ? 1 : 0;
  return a;
}

static mxClassID class_of(const std::string &s) {
  if (s == "double") return mxDOUBLE_CLASS;
  if (s == "int64") return mxINT64_CLASS;
  if (s == "uint64") return mxUINT64_CLASS;
  if (s == "int8") return mxINT8_CLASS;
  if (s == "char") return mxCHAR_CLASS;
  throw std::runtime_error("unknown class " + s);
}
static const char *class_name(mxClassID id) {
  switch (id) {
  case mxDOUBLE_CLASS: return "double";
  case mxINT64_CLASS: return "int64";
  case mxUINT64_CLASS: return "uint64";
  default: return "other";
  }
}

int main(int argc, char **argv) {
  if (argc != 4) {
    std::cerr << "usage: driver <manifest> <outdir> <nlhs>\n";
    return 2;
  }
  std::ifstream man(argv[1]);
  std::vector<mxArray *> rhs;
  std::string line;
  while (std::getline(man, line)) {
    if (line.empty()) continue;
    std::istringstream in(line);
    std::string cls, path;
    size_t m = 0, n = 0;
    in >> cls >> m >> n >> path;
    mxArray *a = mxCreateArray(class_of(cls), m, n, false);
    if (m * n != 0) {
      std::ifstream f(path, std::ios::binary);
      f.read(reinterpret_cast<char *>(a->data.data()), (std::streamsize)a->data.size());
      if (!f) {
        std::cerr << "short file " << path << "\n";
        return 2;
      }
    }
    rhs.push_back(a);
  }
  const int nlhs = std::atoi(argv[3]);
  std::vector<mxArray *> lhs(6, nullptr);
  try {
    mexFunction(nlhs, lhs.data(), (int)rhs.size(), const_cast<const mxArray **>(rhs.data()));
  } catch (const GatewayError &e) {
    std::string msg = e.what();   // one line: the message's line breaks become spaces
    for (char &c : msg)
      if (c == '\n') c = ' ';
    std::cout << "error " << msg << "\n";
    return 3;
  }
  for (int k = 0; k < nlhs; ++k) {
    const mxArray *a = lhs[k];
    if (!a) {
      std::cout << "missing " << k << "\n";
      return 4;
    }
    const std::string path = std::string(argv[2]) + "/out" + std::to_string(k) + ".bin";
    std::ofstream f(path, std::ios::binary);
    if (a->complex) {   // interleaved in memory: write the real parts, then the imaginary parts
      const size_t cnt = a->m * a->n;
      const mxComplexDouble *z = mxGetComplexDoubles(a);
      std::vector<double> re(cnt), im(cnt);
      for (size_t i = 0; i < cnt; ++i) {
        re[i] = z[i].real;
        im[i] = z[i].imag;
      }
      f.write(reinterpret_cast<const char *>(re.data()), (std::streamsize)(cnt * 8));
      f.write(reinterpret_cast<const char *>(im.data()), (std::streamsize)(cnt * 8));
    } else {
      f.write(reinterpret_cast<const char *>(a->data.data()), (std::streamsize)a->data.size());
    }
    std::cout << "out " << k << " " << class_name(a->id) << " " << a->m << " " << a->n << " " << (a->complex ? 1 : 0)
              << "\n";
  }
  for (mxArray *a : rhs) mxDestroyArray(a);
  for (mxArray *a : lhs) if (a) mxDestroyArray(a);
  return 0;
}
