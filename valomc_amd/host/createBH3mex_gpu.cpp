// createBH3mex over the C ABI: the boundary triangles of a tetrahedral
// topology, i.e. the element faces that no other element shares
// (cpp/3d/createBH3mex.cpp:15-78).
//
//   [BH] = createBH3mex(H)
//
// H: int64 [Ne x 4], 1-based node ids.  BH: int64 [Nb x 3], 1-based, one
// row per unshared face in element-major order, faces in the reference's
// order and vertex order: f0 = (v0, v1, v2), f1 = (v0, v1, v3),
// f2 = (v0, v2, v3), f3 = (v1, v2, v3).  The neighbour search is
// vmc_build_neighbours (BuildNeighbourhoods alone, on the host); unlike the
// reference gateway, which decrements the caller's H in place, H is left
// as given.
//
// Built on a MATLAB host with (one line)
//   mex -R2018a createBH3mex_gpu.cpp -I<repo>/include -L<repo>/valomc_amd -lvalomc
#include <cstdint>
#include <string>
#include <vector>

#include "mex.h"
#include "valomc.h"

void mexFunction(int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs) {
  if (nrhs != 1 || nlhs != 1) mexErrMsgTxt("Syntax:\n [BH] = createBH3mex(H)\n");
  const mxArray *h = prhs[0];
  if (mxGetClassID(h) != mxINT64_CLASS || mxGetN(h) != 4 || mxGetM(h) == 0)
    mexErrMsgTxt("createBH3mex: H must be a non-empty int64 matrix with 4 columns");
  const int64_t ne = (int64_t)mxGetM(h);
  const mxInt64 *h1 = mxGetInt64s(h);
  std::vector<int64_t> h0((size_t)ne * 4);
  int64_t maxi = 0;   // the node count, as the reference sizes r from max(H)
  for (size_t i = 0; i < h0.size(); ++i) {
    if (h1[i] < 1) mexErrMsgTxt("createBH3mex: node ids in H are 1-based");
    h0[i] = h1[i] - 1;
    if (h1[i] > maxi) maxi = h1[i];
  }
  std::vector<int64_t> hn((size_t)ne * 4);
  const int st = vmc_build_neighbours(3, ne, 0, maxi, h0.data(), nullptr, hn.data());
  if (st != VMC_OK) mexErrMsgTxt((std::string("createBH3mex: ") + vmc_status_string(st)).c_str());
  static const int face[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  std::vector<int64_t> rows;   // triples, element-major
  for (int64_t e = 0; e < ne; ++e)
    for (int f = 0; f < 4; ++f)
      if (hn[(size_t)(e + ne * f)] == INT64_MAX)
        for (int k = 0; k < 3; ++k) rows.push_back(h0[(size_t)(e + ne * face[f][k])] + 1);
  const size_t nb = rows.size() / 3;
  plhs[0] = mxCreateNumericMatrix(nb, 3, mxINT64_CLASS, mxREAL);
  mxInt64 *out = mxGetInt64s(plhs[0]);
  for (size_t i = 0; i < nb; ++i)
    for (int k = 0; k < 3; ++k) out[i + nb * k] = rows[3 * i + k];
}
