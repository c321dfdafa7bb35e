// Host side of the C ABI (include/valomc.h).
//
// vmc_problem_create is the reference's ErrorChecks() + Init()
// (cpp/3d/MC3D.hpp:460-738, cpp/2d/MC2D.hpp:390-643): same checks in the same
// order with the same error codes, the same neighbour table, the same light
// source list and CDF, the same BCLNormal preprocessing.  It then packs the
// element records the kernels walk on and keeps them resident on the device.
// vmc_problem_run is MonteCarlo() (MC3D.hpp:1641-1822): it launches the
// persistent transport kernel, sums the tally replicas on the device and
// normalises on the host with the reference's own expressions.
#include "../../include/valomc.h"
#include "vmc_types.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace vmc {
hipError_t walk_occupancy(int dim, bool f64, bool fd, bool fixed, bool batch, int block, size_t lds_bytes,
                          int *blocks_per_cu);
hipError_t launch_walk(int dim, bool f64, bool fd, bool fixed, bool batch, int grid, int block, size_t lds_bytes,
                       hipStream_t s, const WalkParams &P);
hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_ranges(unsigned long long *p, size_t n, int nrep, size_t stride, unsigned long long v,
                              hipStream_t s);
hipError_t launch_reduce(bool fixed, const void *rep, double *out, size_t n, int nrep, size_t stride,
                         double inv_scale, hipStream_t s);
} // namespace vmc

using namespace vmc;

namespace {

thread_local std::string g_last_error;

// mcerror codes (cpp/Errors.hpp:4-21); the ABI reports 1 + code
enum McErr {
  NO_LIGHTSOURCE, SIZE_MISMATCH_G, SIZE_MISMATCH_MUA, SIZE_MISMATCH_MUS, SIZE_MISMATCH_N,
  SIZE_MISMATCH_BCN, MISSING_BCN, SIZE_MISMATCH_BCTYPE, SIZE_MISMATCH_LIGHT_DIRECTION,
  SIZE_MISMATCH_LIGHT_DIRECTION_TYPE, INCONSISTENT_H, INCONSISTENT_BH, INCONSISTENT_MESH,
  INCONSISTENT_MESH_DUPLICATE_NEIGHBORS, NO_GAUSSIAN_SIGMA, MISSING_BOUNDARY
};
inline int mc_status(int e) { return 1 + e; }

const int kFace3[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}}; // MC3D.hpp:1013
const int kEdge2[3][2] = {{0, 1}, {1, 2}, {2, 0}};                     // MC2D.hpp:921
const int64_t kUnset = INT64_MAX;                                      // MC3D.hpp:29-31
const double kC0 = 2.99792458e11;                                      // MC3D.hpp:178

struct HipFail {
  hipError_t e;
  const char *what;
};
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HipFail{e, what};
}

template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    hip_check(hipMalloc((void **)&p, count * sizeof(T)), "hipMalloc");
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Tally arena and launch resources of one stream: nrep replicas of
// [ne][W] element words and [nb][4] boundary words, the launch's item
// records, the weight ledger, the counters and a per-item loss count.
struct Arena {
  int64_t ne = 0, nb = 0;
  int W = 1, nrep = 0;
  size_t nitems = 0;
  DevBuf<unsigned long long> etally, btally, counters, item_loss;
  DevBuf<double> ledger, raw_e, raw_b;
  DevBuf<ItemRec> items;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void open() {
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&ev0), "hipEventCreate");
    hip_check(hipEventCreate(&ev1), "hipEventCreate");
  }
  void close() {
    etally.release();
    btally.release();
    counters.release();
    item_loss.release();
    ledger.release();
    raw_e.release();
    raw_b.release();
    items.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
    ev0 = ev1 = nullptr;
    stream = nullptr;
  }
};

} // namespace

// Mesh geometry and what is derived from it alone: shared by a problem and
// its variants, immutable once the problem is created.
struct MeshData {
  std::vector<int64_t> H, HN, BH;
  std::vector<double> r;
  std::vector<double> vol, area;    // element volumes (2D areas), boundary areas (2D lengths)
  std::vector<i4> nbcode;           // per element (caller's numbering), neighbours as device ids
  std::vector<int32_t> perm, inv;   // device order: perm[device id] = element, inv[element] = device id
};

// Optical parameters and the device element records packed from them (with
// the geometry): shared by variants that keep the medium, copied when one of
// them changes it.
struct MediumData {
  std::vector<double> mua, mus, g, n;
  DevBuf<unsigned char> elems32, elems64;
  bool dirty32 = true, dirty64 = true;
  ~MediumData() {
    elems32.release();
    elems64.release();
  }
};

struct vmc_problem {
  std::shared_ptr<MeshData> md;
  std::shared_ptr<MediumData> med;
  int dim = 3, nv = 4, nbv = 3;
  int64_t Ne = 0, Nb = 0, Nr = 0;
  std::vector<int64_t> &H, &HN, &BH;
  std::vector<double> &r;
  std::vector<char> BCType, BCLDirType;
  std::vector<double> BCn, BCLNormal, BCIntensity, GaussianSigma;
  std::vector<double> BCLNormal_in;  // BCLNormal as given, before its preprocessing
  double shift[3] = {0, 0, 0};
  std::vector<double> &vol, &area;
  std::vector<int64_t> lsFace, lsMother;
  std::vector<double> lsCDF;
  std::vector<SrcRec> src;
  std::vector<BndRec> bnd;
  std::vector<i4> &nbcode;
  std::vector<int32_t> &perm, &inv;

  int device = 0, num_cus = 0;
  int64_t lds_per_cu = 65536;
  DevBuf<double> cdf;
  DevBuf<SrcRec> dsrc;
  DevBuf<BndRec> dbnd;
  Arena arena;   // tallies and stream of this problem's own runs

  explicit vmc_problem(std::shared_ptr<MeshData> mesh = std::make_shared<MeshData>(),
                       std::shared_ptr<MediumData> medium = std::make_shared<MediumData>())
      : md(std::move(mesh)), med(std::move(medium)), H(md->H), HN(md->HN), BH(md->BH), r(md->r),
        vol(md->vol), area(md->area), nbcode(md->nbcode), perm(md->perm), inv(md->inv) {}

  int64_t h(int64_t el, int c) const { return H[el + Ne * c]; }
  int64_t hn(int64_t el, int c) const { return HN[el + Ne * c]; }
  int64_t bh(int64_t ib, int c) const { return BH[ib + Nb * c]; }
  double x(int64_t node, int d) const { return r[node + Nr * d]; }
  double bcl(int64_t ib, int d) const { return BCLNormal[ib + Nb * d]; }
  double &bcl_ref(int64_t ib, int d) { return BCLNormal[ib + Nb * d]; }
};

namespace {

struct McError {
  int code;
};

// ---------------------------------------------------------------------------
// Geometry with the reference's expressions (MC3D.hpp:306-457, MC2D.hpp:303-387)
// ---------------------------------------------------------------------------

double volume3(const vmc_problem &m, int64_t el) {
  double ax = m.x(m.h(el, 0), 0), ay = m.x(m.h(el, 0), 1), az = m.x(m.h(el, 0), 2);
  double bx = m.x(m.h(el, 1), 0), by = m.x(m.h(el, 1), 1), bz = m.x(m.h(el, 1), 2);
  double cx = m.x(m.h(el, 2), 0), cy = m.x(m.h(el, 2), 1), cz = m.x(m.h(el, 2), 2);
  double dx = m.x(m.h(el, 3), 0), dy = m.x(m.h(el, 3), 1), dz = m.x(m.h(el, 3), 2);
  double v = (ax - dx) * ((by - dy) * (cz - dz) - (bz - dz) * (cy - dy)) -
             (ay - dy) * ((bx - dx) * (cz - dz) - (bz - dz) * (cx - dx)) +
             (az - dz) * ((bx - dx) * (cy - dy) - (by - dy) * (cx - dx));
  return fabs(v) / 6.0;
}

double heron3(const vmc_problem &m, int64_t i0, int64_t i1, int64_t i2) {
  double a = sqrt(pow(m.x(i1, 0) - m.x(i0, 0), 2) + pow(m.x(i1, 1) - m.x(i0, 1), 2) + pow(m.x(i1, 2) - m.x(i0, 2), 2));
  double b = sqrt(pow(m.x(i2, 0) - m.x(i0, 0), 2) + pow(m.x(i2, 1) - m.x(i0, 1), 2) + pow(m.x(i2, 2) - m.x(i0, 2), 2));
  double c = sqrt(pow(m.x(i1, 0) - m.x(i2, 0), 2) + pow(m.x(i1, 1) - m.x(i2, 1), 2) + pow(m.x(i1, 2) - m.x(i2, 2), 2));
  return sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c)) / 4.0;
}

double area2(const vmc_problem &m, int64_t el) {
  double ax = m.x(m.h(el, 1), 0) - m.x(m.h(el, 0), 0);
  double ay = m.x(m.h(el, 1), 1) - m.x(m.h(el, 0), 1);
  double bx = m.x(m.h(el, 2), 0) - m.x(m.h(el, 0), 0);
  double by = m.x(m.h(el, 2), 1) - m.x(m.h(el, 0), 1);
  double c = ax * by - ay * bx;
  return 0.5 * (c >= 0.0 ? c : -c);
}

double len2(const vmc_problem &m, int64_t i0, int64_t i1) {
  return sqrt(pow(m.x(i1, 0) - m.x(i0, 0), 2) + pow(m.x(i1, 1) - m.x(i0, 1), 2));
}

void bnd_normal(const vmc_problem &m, int64_t ib, double *nrm) {
  if (m.dim == 3) {
    int64_t i0 = m.bh(ib, 0), i1 = m.bh(ib, 1), i2 = m.bh(ib, 2);
    double x1 = m.x(i1, 0) - m.x(i0, 0), y1 = m.x(i1, 1) - m.x(i0, 1), z1 = m.x(i1, 2) - m.x(i0, 2);
    double x2 = m.x(i2, 0) - m.x(i0, 0), y2 = m.x(i2, 1) - m.x(i0, 1), z2 = m.x(i2, 2) - m.x(i0, 2);
    double nx = y1 * z2 - z1 * y2, ny = z1 * x2 - x1 * z2, nz = x1 * y2 - y1 * x2;
    double len = sqrt(nx * nx + ny * ny + nz * nz);
    nrm[0] = nx / len; nrm[1] = ny / len; nrm[2] = nz / len;
  } else {
    int64_t i0 = m.bh(ib, 0), i1 = m.bh(ib, 1);
    double dx = m.x(i1, 0) - m.x(i0, 0), dy = m.x(i1, 1) - m.x(i0, 1);
    double len = sqrt(dx * dx + dy * dy);
    nrm[0] = dy / len; nrm[1] = -dx / len; nrm[2] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// ErrorChecks (MC3D.hpp:460-571, MC2D.hpp:390-513)
// ---------------------------------------------------------------------------

void error_checks(const vmc_problem &m) {
  if (m.BCn.empty()) throw McError{MISSING_BCN};
  bool has_ls = false, has_gauss = false;
  for (char c : m.BCType) {
    if (c != 'a') has_ls = true;
    if (c == 'g') has_gauss = true;
  }
  if (!has_ls) throw McError{NO_LIGHTSOURCE};
  if (m.dim == 2 && has_gauss && m.GaussianSigma.empty()) throw McError{NO_GAUSSIAN_SIGMA};
  for (int64_t v : m.H)
    if (v < 0 || v >= m.Nr) throw McError{INCONSISTENT_H};
  for (int64_t v : m.BH)
    if (v < 0 || v >= m.Nr) throw McError{INCONSISTENT_BH};
}

// ---------------------------------------------------------------------------
// Neighbour table.  Same result as the reference's vertex-incidence search
// (MC3D.hpp:740-871, MC2D.hpp:645-773): face f of element e gets the other
// element, or boundary entry -1-ib, holding all of its vertices; two
// different matches are an inconsistent mesh.  Found here by sorting face
// keys, O(F log F) instead of O(sum of vertex degree^2).
// ---------------------------------------------------------------------------

struct FaceKey {
  int64_t a, b, c;
  int64_t owner;   // (el << 2 | f) for element faces, -1 - ib for boundary entries
  bool operator<(const FaceKey &o) const {
    if (a != o.a) return a < o.a;
    if (b != o.b) return b < o.b;
    if (c != o.c) return c < o.c;
    return owner < o.owner;
  }
  bool same(const FaceKey &o) const { return a == o.a && b == o.b && c == o.c; }
};

inline void sort3(int64_t &a, int64_t &b, int64_t &c) {
  if (a > b) std::swap(a, b);
  if (b > c) std::swap(b, c);
  if (a > b) std::swap(a, b);
}

void build_neighbours(vmc_problem &m) {
  if ((int64_t)m.HN.size() == m.Ne * m.nv) return;   // given (MC3D.hpp:839)
  m.HN.assign(m.Ne * m.nv, kUnset);
  std::vector<FaceKey> keys;
  keys.reserve(m.Ne * m.nv + m.Nb);
  for (int64_t e = 0; e < m.Ne; ++e)
    for (int f = 0; f < m.nv; ++f) {
      FaceKey k;
      if (m.dim == 3) {
        k.a = m.h(e, kFace3[f][0]); k.b = m.h(e, kFace3[f][1]); k.c = m.h(e, kFace3[f][2]);
        sort3(k.a, k.b, k.c);
      } else {
        k.a = std::min(m.h(e, kEdge2[f][0]), m.h(e, kEdge2[f][1]));
        k.b = std::max(m.h(e, kEdge2[f][0]), m.h(e, kEdge2[f][1]));
        k.c = -1;
      }
      k.owner = (e << 2) | f;
      keys.push_back(k);
    }
  for (int64_t b = 0; b < m.Nb; ++b) {
    FaceKey k;
    if (m.dim == 3) {
      k.a = m.bh(b, 0); k.b = m.bh(b, 1); k.c = m.bh(b, 2);
      sort3(k.a, k.b, k.c);
    } else {
      k.a = std::min(m.bh(b, 0), m.bh(b, 1));
      k.b = std::max(m.bh(b, 0), m.bh(b, 1));
      k.c = -1;
    }
    k.owner = -1 - b;
    keys.push_back(k);
  }
  std::sort(keys.begin(), keys.end());
  const int dup = m.dim == 3 ? INCONSISTENT_MESH_DUPLICATE_NEIGHBORS : INCONSISTENT_MESH;
  size_t i = 0;
  while (i < keys.size()) {
    size_t j = i + 1;
    while (j < keys.size() && keys[j].same(keys[i])) ++j;
    // every element face in the group takes the one other entry
    for (size_t q = i; q < j; ++q) {
      if (keys[q].owner < 0) continue;
      const int64_t e = keys[q].owner >> 2;
      const int f = (int)(keys[q].owner & 3);
      int64_t val = kUnset;
      for (size_t o = i; o < j; ++o) {
        if (o == q) continue;
        const int64_t cand = keys[o].owner < 0 ? keys[o].owner : (keys[o].owner >> 2);
        if (cand == e) continue;
        if (val != kUnset && val != cand) throw McError{dup};
        val = cand;
      }
      m.HN[e + m.Ne * f] = val;
    }
    i = j;
  }
  if (m.dim == 2)
    for (int64_t v : m.HN)
      if (v == kUnset) throw McError{MISSING_BOUNDARY};   // MC2D.hpp:768-771
}

bool is_source(int dim, char t) {
  if (dim == 3) return t == 'l' || t == 'L' || t == 'i' || t == 'I' || t == 'c' || t == 'C' || t == 'p';
  return t == 'l' || t == 'i' || t == 'c' || t == 'g';
}

// BuildLightSource (MC3D.hpp:941-1002, MC2D.hpp:851-911): element-major order
void build_light_sources(vmc_problem &m) {
  m.lsFace.clear(); m.lsMother.clear(); m.lsCDF.clear();
  for (int64_t e = 0; e < m.Ne; ++e)
    for (int f = 0; f < m.nv; ++f) {
      int64_t v = m.hn(e, f);
      if (v <= -1) {
        int64_t ib = -1 - v;
        if (ib < m.Nb && is_source(m.dim, m.BCType[ib])) {
          m.lsFace.push_back(ib);
          m.lsMother.push_back(e);
          m.lsCDF.push_back(m.dim == 3 ? heron3(m, m.bh(ib, 0), m.bh(ib, 1), m.bh(ib, 2))
                                       : len2(m, m.bh(ib, 0), m.bh(ib, 1)));
        }
      }
    }
  const size_t ns = m.lsCDF.size();
  if (ns == 0) throw McError{NO_LIGHTSOURCE};
  for (size_t i = 0; i < ns; ++i) m.lsCDF[i] *= m.BCIntensity[m.lsFace[i]];
  for (size_t i = 1; i < ns; ++i) m.lsCDF[i] += m.lsCDF[i - 1];
  for (size_t i = 0; i < ns; ++i) m.lsCDF[i] /= m.lsCDF[ns - 1];
}

// BCLNormal preprocessing of Init (MC3D.hpp:670-735, MC2D.hpp:611-640)
void prepare_directions(vmc_problem &m) {
  if (m.BCLNormal.empty()) return;
  if (m.dim == 3) {
    for (int64_t i = 0; i < m.Nb; ++i) {
      double len = sqrt(pow(m.bcl(i, 0), 2) + pow(m.bcl(i, 1), 2) + pow(m.bcl(i, 2), 2));
      if (m.BCLDirType[i] != 'n') continue;   // only 'n' rows are normalised
      for (int d = 0; d < 3; ++d) m.bcl_ref(i, d) /= len;
    }
    for (int64_t i = 0; i < m.Nb; ++i) {
      if (m.BCLDirType[i] != 'r') continue;
      double nn[3], e1[3], e2[3], len, dpn, dp1;
      bnd_normal(m, i, nn);
      for (int d = 0; d < 3; ++d) {
        e1[d] = m.x(m.bh(i, 1), d) - m.x(m.bh(i, 0), d);
        e2[d] = m.x(m.bh(i, 2), d) - m.x(m.bh(i, 0), d);
      }
      dpn = nn[0] * e1[0] + nn[1] * e1[1] + nn[2] * e1[2];
      for (int d = 0; d < 3; ++d) e1[d] -= dpn * nn[d];
      len = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
      for (int d = 0; d < 3; ++d) e1[d] /= len;
      dpn = nn[0] * e2[0] + nn[1] * e2[1] + nn[2] * e2[2];
      dp1 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
      for (int d = 0; d < 3; ++d) e2[d] -= dpn * nn[d] + dp1 * e1[d];
      len = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
      for (int d = 0; d < 3; ++d) e2[d] /= len;
      double ox = m.bcl(i, 0), oy = m.bcl(i, 1), oz = m.bcl(i, 2);
      for (int d = 0; d < 3; ++d) m.bcl_ref(i, d) = ox * e1[d] + oy * e2[d] + oz * nn[d];
    }
  } else {
    for (int64_t i = 0; i < m.Nb; ++i) {
      double len = sqrt(pow(m.bcl(i, 0), 2) + pow(m.bcl(i, 1), 2));
      m.bcl_ref(i, 0) /= len;
      m.bcl_ref(i, 1) /= len;
    }
    for (int64_t i = 0; i < m.Nb; ++i) {
      if (m.BCLDirType[i] != 'r') continue;
      double nn[3], t[2], len;
      bnd_normal(m, i, nn);
      t[0] = m.x(m.bh(i, 1), 0) - m.x(m.bh(i, 0), 0);
      t[1] = m.x(m.bh(i, 1), 1) - m.x(m.bh(i, 0), 1);
      len = sqrt(t[0] * t[0] + t[1] * t[1]);
      t[0] /= len; t[1] /= len;
      double ox = m.bcl(i, 0), oy = m.bcl(i, 1);
      m.bcl_ref(i, 0) = ox * t[0] + oy * nn[0];
      m.bcl_ref(i, 1) = ox * t[1] + oy * nn[1];
    }
  }
}

// Device order of the elements: sorted along a Morton (Z-order) curve of
// their centroids, so elements that share faces get nearby ids.  A photon's
// successive element records and tally words then share cache lines and
// pages; meshes numbered by construction (a grid generator emits each of the
// six tetrahedra of a cube in a separate block) otherwise put face
// neighbours far apart.  Results are mapped back to the caller's numbering.
uint64_t spread_bits3(uint64_t v) {   // 21 bits -> every third bit
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}
uint64_t spread_bits2(uint64_t v) {   // 32 bits -> every second bit
  v &= 0xffffffffull;
  v = (v | v << 16) & 0x0000ffff0000ffffull;
  v = (v | v << 8) & 0x00ff00ff00ff00ffull;
  v = (v | v << 4) & 0x0f0f0f0f0f0f0f0full;
  v = (v | v << 2) & 0x3333333333333333ull;
  v = (v | v << 1) & 0x5555555555555555ull;
  return v;
}

void build_order(vmc_problem &m) {
  const int bits = m.dim == 3 ? 21 : 32;
  double lo[3] = {0, 0, 0}, maxspan = 0;
  for (int d = 0; d < m.dim; ++d) {
    double a = m.x(0, d), b = a;
    for (int64_t i = 1; i < m.Nr; ++i) {
      a = std::min(a, m.x(i, d));
      b = std::max(b, m.x(i, d));
    }
    lo[d] = a;
    maxspan = std::max(maxspan, b - a);
  }
  // Morton cells about the size of the hexahedron six tetrahedra (or the
  // quadrilateral two triangles) of a structured mesh fill, so the elements
  // of one cell stay consecutive
  double mean = 0;
  for (int64_t e = 0; e < m.Ne; ++e) mean += std::fabs(m.dim == 3 ? volume3(m, e) : area2(m, e));
  mean = m.Ne > 0 ? mean / (double)m.Ne : 0.0;
  const double cell = m.dim == 3 ? std::cbrt(6.0 * mean) : std::sqrt(2.0 * mean);
  const double maxcells = (double)((1ull << bits) - 1);
  const double inv_cell = (cell > 0 && maxspan / cell < maxcells) ? 1.0 / cell
                                                                  : (maxspan > 0 ? maxcells / maxspan : 1.0);
  std::vector<uint64_t> key(m.Ne);
  for (int64_t e = 0; e < m.Ne; ++e) {
    uint64_t q[3] = {0, 0, 0};
    for (int d = 0; d < m.dim; ++d) {
      double c = 0;
      for (int k = 0; k < m.nv; ++k) c += m.x(m.h(e, k), d);
      c /= m.nv;
      q[d] = (uint64_t)std::min(std::max((c - lo[d]) * inv_cell, 0.0), maxcells);
    }
    key[e] = m.dim == 3 ? (spread_bits3(q[0]) | spread_bits3(q[1]) << 1 | spread_bits3(q[2]) << 2)
                        : (spread_bits2(q[0]) | spread_bits2(q[1]) << 1);
  }
  m.perm.resize(m.Ne);
  for (int64_t e = 0; e < m.Ne; ++e) m.perm[e] = (int32_t)e;
  std::sort(m.perm.begin(), m.perm.end(), [&](int32_t a, int32_t b) {
    return key[a] != key[b] ? key[a] < key[b] : a < b;
  });
  m.inv.resize(m.Ne);
  for (int64_t d = 0; d < m.Ne; ++d) m.inv[m.perm[d]] = (int32_t)d;
}

// Per-source launch records (CreatePhoton, MC3D.hpp:1060-1294 / MC2D.hpp:958-1132)
void build_source_records(vmc_problem &m) {
  m.src.assign(m.lsFace.size(), SrcRec{});
  for (size_t i = 0; i < m.lsFace.size(); ++i) {
    SrcRec &s = m.src[i];
    const int64_t ib = m.lsFace[i], el = m.lsMother[i];
    s.ib = (int32_t)ib;
    s.mother = m.inv[el];
    s.face = -1;
    for (int f = 0; f < m.nv; ++f)
      if (-1 - m.hn(el, f) == ib) { s.face = f; break; }
    const int nb = m.nbv;
    double *vv[3] = {s.v0, s.v1, s.v2};
    for (int c = 0; c < nb; ++c)
      for (int d = 0; d < m.dim; ++d) vv[c][d] = m.x(m.bh(ib, c), d) - m.shift[d];
    // inward normal: orient towards the element centroid (MC3D.hpp:1092-1107)
    double nn[3];
    bnd_normal(m, ib, nn);
    double t[3] = {0, 0, 0};
    for (int d = 0; d < m.dim; ++d) {
      double cen = 0, fc = 0;
      for (int c = 0; c < m.nv; ++c) cen += m.x(m.h(el, c), d);
      for (int c = 0; c < nb; ++c) fc += m.x(m.bh(ib, c), d);
      t[d] = cen / m.nv - fc / nb;
    }
    double dot = 0;
    for (int d = 0; d < m.dim; ++d) dot += nn[d] * t[d];
    if (dot < 0.0)
      for (int d = 0; d < 3; ++d) nn[d] = -nn[d];
    for (int d = 0; d < 3; ++d) s.nin[d] = nn[d];
    const char bt = m.BCType[ib];
    const bool given = m.BCLDirType[ib] != 'n';
    if (m.dim == 3) {
      if (!given && (bt == 'L' || bt == 'I' || bt == 'p')) throw McError{-1 + VMC_ERR_UNDEFINED_DIRECTION};
      if (bt == 'l' || bt == 'L') {
        s.kind = given ? kSrcGiven : kSrcNormal;
        if (given)
          for (int d = 0; d < 3; ++d) s.dir[d] = m.bcl(ib, d);
      } else if (bt == 'p') {
        s.kind = kSrcPencil;
        for (int d = 0; d < 3; ++d) s.dir[d] = m.bcl(ib, d) - m.shift[d];
      } else if (bt == 'i' || bt == 'I') {
        s.kind = kSrcIso3;
      } else {   // 'c', 'C'
        s.kind = kSrcCos3;
        double e1[3], e2[3], len, dpn, dp1;
        for (int d = 0; d < 3; ++d) {
          e1[d] = m.x(m.bh(ib, 1), d) - m.x(m.bh(ib, 0), d);
          e2[d] = m.x(m.bh(ib, 2), d) - m.x(m.bh(ib, 0), d);
        }
        if (!given) {
          len = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
          for (int d = 0; d < 3; ++d) e1[d] /= len;
          dpn = nn[0] * e2[0] + nn[1] * e2[1] + nn[2] * e2[2];
          dp1 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
          for (int d = 0; d < 3; ++d) e2[d] -= dpn * nn[d] + dp1 * e1[d];
        } else {
          // frame orthogonal to BCLNormal, first axis still the inward
          // normal (MC3D.hpp:1259-1285)
          double b[3] = {m.bcl(ib, 0), m.bcl(ib, 1), m.bcl(ib, 2)};
          dpn = b[0] * e1[0] + b[1] * e1[1] + b[2] * e1[2];
          for (int d = 0; d < 3; ++d) e1[d] -= dpn * b[d];
          len = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
          for (int d = 0; d < 3; ++d) e1[d] /= len;
          dpn = b[0] * e2[0] + b[1] * e2[1] + b[2] * e2[2];
          dp1 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
          for (int d = 0; d < 3; ++d) e2[d] -= dpn * b[d] + dp1 * e1[d];
        }
        len = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        for (int d = 0; d < 3; ++d) {
          e2[d] /= len;
          s.a1[d] = e1[d];
          s.a2[d] = e2[d];
        }
      }
    } else {
      s.phi0 = given ? atan2(m.bcl(ib, 1), m.bcl(ib, 0)) : atan2(nn[1], nn[0]);
      if (bt == 'l') {
        s.kind = given ? kSrcGiven : kSrcNormal;
        if (given) { s.dir[0] = m.bcl(ib, 0); s.dir[1] = m.bcl(ib, 1); }
      } else if (bt == 'i') {
        s.kind = kSrcIso2;
      } else if (bt == 'c') {
        s.kind = given ? kSrcCos2Rej : kSrcCos2;
      } else {   // 'g'
        s.kind = given ? kSrcGauss2Rej : kSrcGauss2;
        s.gsig = m.GaussianSigma[ib];
        s.gh = 0.5 - 0.5 * erf(M_PI / s.gsig / 2.0 / sqrt(2.0));
      }
    }
  }
}

// Element records: outward face planes, neighbour codes, optical parameters.
template <int NF, class R>
void pack_elements(const vmc_problem &m, std::vector<Elem<NF, R>> &out) {
  out.assign(m.Ne, Elem<NF, R>{});
  for (int64_t dev = 0; dev < m.Ne; ++dev) {
    const int64_t e = m.perm[dev];
    double v[4][3] = {{0}};
    for (int c = 0; c < m.nv; ++c)
      for (int d = 0; d < m.dim; ++d) v[c][d] = m.x(m.h(e, c), d) - m.shift[d];
    Elem<NF, R> &E = out[dev];
    for (int f = 0; f < NF; ++f) {
      double nrm[3] = {0, 0, 0}, off = 0;
      if (m.dim == 3) {
        const double *a = v[kFace3[f][0]], *b = v[kFace3[f][1]], *c = v[kFace3[f][2]];
        const double *o = v[6 - kFace3[f][0] - kFace3[f][1] - kFace3[f][2]];
        double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        double w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        nrm[0] = u[1] * w[2] - u[2] * w[1];
        nrm[1] = u[2] * w[0] - u[0] * w[2];
        nrm[2] = u[0] * w[1] - u[1] * w[0];
        double len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int d = 0; d < 3; ++d) nrm[d] /= len;
        if (nrm[0] * (o[0] - a[0]) + nrm[1] * (o[1] - a[1]) + nrm[2] * (o[2] - a[2]) > 0)
          for (int d = 0; d < 3; ++d) nrm[d] = -nrm[d];
        for (int d = 0; d < 3; ++d) off += nrm[d] * (a[d] + b[d] + c[d]) / 3.0;
      } else {
        const double *a = v[kEdge2[f][0]], *b = v[kEdge2[f][1]];
        const double *o = v[3 - kEdge2[f][0] - kEdge2[f][1]];
        double dx = b[0] - a[0], dy = b[1] - a[1];
        double len = sqrt(dx * dx + dy * dy);
        nrm[0] = dy / len;
        nrm[1] = -dx / len;
        if (nrm[0] * (o[0] - a[0]) + nrm[1] * (o[1] - a[1]) > 0) { nrm[0] = -nrm[0]; nrm[1] = -nrm[1]; }
        off = nrm[0] * (a[0] + b[0]) / 2.0 + nrm[1] * (a[1] + b[1]) / 2.0;
      }
      E.pl[f].x = (R)nrm[0];
      E.pl[f].y = (R)nrm[1];
      E.pl[f].z = (R)nrm[2];
      E.pl[f].w = (R)off;
    }
    E.opt.x = (R)m.med->mua[e];
    E.opt.y = (R)m.med->mus[e];
    E.opt.z = (R)m.med->n[e];
    E.opt.w = (R)m.med->g[e];
    E.nb = m.nbcode[e];
    for (int f = 0; f < NF; ++f) {
      const int code = m.nbcode[e][f];
      const int64_t nx = code >= 0 ? (int64_t)m.perm[code >> 3] : -1;
      E.nbo[f].x = nx >= 0 ? (R)m.med->mus[nx] : R(0);
      E.nbo[f].y = nx >= 0 ? (R)m.med->n[nx] : R(0);
    }
  }
}

void build_nb_codes(vmc_problem &m) {
  m.nbcode.assign(m.Ne, i4{kNbNone, kNbNone, kNbNone, kNbNone});
  for (int64_t e = 0; e < m.Ne; ++e) {
    int code[4] = {kNbNone, kNbNone, kNbNone, kNbNone};
    for (int f = 0; f < m.nv; ++f) {
      const int64_t v = m.hn(e, f);
      if (v == kUnset) continue;
      if (v < 0) {
        if (-1 - v < m.Nb) code[f] = (int)v;
        continue;
      }
      if (v >= m.Ne) continue;
      int entry = 0, back = 0;
      for (int q = 0; q < m.nv; ++q)
        if (m.hn(v, q) == e) { entry = q; back = 1; break; }   // MC3D.hpp:1616-1628
      code[f] = (int)(((int64_t)m.inv[v] << 3) | (back << 2) | entry);
    }
    m.nbcode[e] = i4{code[0], code[1], code[2], code[3]};
  }
}

void build_boundary(vmc_problem &m) {
  m.bnd.assign(m.Nb, BndRec{});
  for (int64_t i = 0; i < m.Nb; ++i) {
    const char t = m.BCType[i];
    BndRec &b = m.bnd[i];
    b.bcn = m.BCn[i];
    if (t == 'm' || t == 'L' || t == 'I' || t == 'C') b.kind = kBndMirror;   // MC3D.hpp:1541
    else b.kind = m.BCn[i] > 0.0 ? kBndFresnel : kBndAbsorb;                  // MC3D.hpp:1552
  }
}

void upload_params(vmc_problem &m, bool f64) {
  if (f64 && !m.med->dirty64) return;
  if (!f64 && !m.med->dirty32) return;
  if (m.dim == 3) {
    if (f64) {
      std::vector<Elem<4, double>> v;
      pack_elements<4, double>(m, v);
      m.med->elems64.ensure(v.size() * sizeof(v[0]));
      hip_check(hipMemcpy(m.med->elems64.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice), "upload");
    } else {
      std::vector<Elem<4, float>> v;
      pack_elements<4, float>(m, v);
      m.med->elems32.ensure(v.size() * sizeof(v[0]));
      hip_check(hipMemcpy(m.med->elems32.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice), "upload");
    }
  } else {
    if (f64) {
      std::vector<Elem<3, double>> v;
      pack_elements<3, double>(m, v);
      m.med->elems64.ensure(v.size() * sizeof(v[0]));
      hip_check(hipMemcpy(m.med->elems64.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice), "upload");
    } else {
      std::vector<Elem<3, float>> v;
      pack_elements<3, float>(m, v);
      m.med->elems32.ensure(v.size() * sizeof(v[0]));
      hip_check(hipMemcpy(m.med->elems32.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice), "upload");
    }
  }
  (f64 ? m.med->dirty64 : m.med->dirty32) = false;
}

// Normalisation of MonteCarlo (MC3D.hpp:1793-1818, MC2D.hpp:1673-1703)
void normalise(const vmc_problem &m, double omega, const double *kfac_n, int64_t N, double *ER, double *EI,
               double *EBR, double *EBI, double *DEBR, double *DEBI) {
  const bool d3 = m.dim == 3;
  if (omega <= 0.0) {
    for (int64_t i = 0; i < m.Ne; ++i) {
      if (m.med->mua[i] > 0.0) ER[i] /= m.med->mua[i] * m.vol[i] * (double)N;
      else ER[i] /= m.vol[i] * (double)N;
    }
    for (int64_t i = 0; i < m.Nb; ++i) EBR[i] /= (double)N * m.area[i];
    if (!d3)
      for (int64_t i = 0; i < m.Nb; ++i) DEBR[i] /= (double)N * m.area[i];
  } else {
    for (int64_t i = 0; i < m.Ne; ++i) {
      const double a = ER[i], b = EI[i], k = kfac_n[i];
      ER[i] = (b * k + a * m.med->mua[i]) / (pow(k, 2) + pow(m.med->mua[i], 2)) / (double)N / m.vol[i];
      EI[i] = -(a * k - b * m.med->mua[i]) / (pow(k, 2) + pow(m.med->mua[i], 2)) / (double)N / m.vol[i];
    }
    for (int64_t i = 0; i < m.Nb; ++i) {
      EBR[i] /= (double)N * m.area[i];
      EBI[i] /= (double)N * m.area[i];
      if (!d3) {
        DEBR[i] /= (double)N * m.area[i];
        DEBI[i] /= (double)N * m.area[i];
      }
    }
  }
}

int fail_hip(const HipFail &f) {
  g_last_error = std::string(f.what) + ": " + hipGetErrorString(f.e);
  return f.e == hipErrorOutOfMemory ? VMC_ERR_OUT_OF_MEMORY : VMC_ERR_HIP;
}

template <class T> void copy_in(std::vector<T> &dst, const T *src, int64_t n) {
  if (src) dst.assign(src, src + n);
  else dst.clear();
}

} // namespace

namespace {

// ErrorChecks() and the host part of Init(); returns a status.
int host_prepare(const vmc_problem_desc *d, vmc_problem *m) {
  if ((d->dim != 2 && d->dim != 3) || d->Ne <= 0 || d->Nb <= 0 || d->Nr <= 0 || !d->H || !d->BH || !d->r ||
      !d->mua || !d->mus || !d->g || !d->n || !d->BCType)
    return VMC_ERR_INVALID_ARGUMENT;
  if (d->Ne >= kMaxElements || d->Nb >= INT32_MAX) return VMC_ERR_INVALID_ARGUMENT;
  try {
    m->dim = d->dim;
    m->nv = d->dim + 1;
    m->nbv = d->dim;
    m->Ne = d->Ne; m->Nb = d->Nb; m->Nr = d->Nr;
    copy_in(m->H, d->H, d->Ne * m->nv);
    copy_in(m->HN, d->HN, d->Ne * m->nv);
    copy_in(m->BH, d->BH, d->Nb * m->nbv);
    copy_in(m->r, d->r, d->Nr * m->dim);
    copy_in(m->med->mua, d->mua, d->Ne);
    copy_in(m->med->mus, d->mus, d->Ne);
    copy_in(m->med->g, d->g, d->Ne);
    copy_in(m->med->n, d->n, d->Ne);
    copy_in(m->BCType, d->BCType, d->Nb);
    copy_in(m->BCn, d->BCn, d->Nb);
    copy_in(m->BCLNormal, d->BCLNormal, d->Nb * m->dim);
    copy_in(m->BCLDirType, d->BCLightDirectionType, d->Nb);
    copy_in(m->BCIntensity, d->BCIntensity, d->Nb);
    if (m->dim == 2) copy_in(m->GaussianSigma, d->GaussianSigma, d->Nb);

    error_checks(*m);
    build_neighbours(*m);
    if (m->BCIntensity.empty()) m->BCIntensity.assign(m->Nb, 1.0);
    if (m->BCLDirType.empty()) m->BCLDirType.assign(m->Nb, 'n');
    build_light_sources(*m);
    m->BCLNormal_in = m->BCLNormal;
    prepare_directions(*m);

    for (int dd = 0; dd < m->dim; ++dd) {
      double lo = m->x(0, dd), hi = lo;
      for (int64_t i = 1; i < m->Nr; ++i) {
        lo = std::min(lo, m->x(i, dd));
        hi = std::max(hi, m->x(i, dd));
      }
      m->shift[dd] = 0.5 * (lo + hi);
    }
    build_order(*m);
    build_source_records(*m);
    build_nb_codes(*m);
    build_boundary(*m);
    m->vol.resize(m->Ne);
    m->area.resize(m->Nb);
    for (int64_t i = 0; i < m->Ne; ++i) m->vol[i] = m->dim == 3 ? volume3(*m, i) : area2(*m, i);
    for (int64_t i = 0; i < m->Nb; ++i)
      m->area[i] = m->dim == 3 ? heron3(*m, m->bh(i, 0), m->bh(i, 1), m->bh(i, 2)) : len2(*m, m->bh(i, 0), m->bh(i, 1));
  } catch (McError e) {
    return e.code + 1;
  }
  return VMC_OK;
}

} // namespace

namespace {

// ---------------------------------------------------------------------------
// Launch engine: persistent walk launches over a list of items (problem
// instances), shared by single runs and batches
// ---------------------------------------------------------------------------

const uint64_t kChunk = 128;   // photon ids a wave takes from the dispenser at once

struct LaunchCfg {
  int dim = 3;
  bool f64 = false, fd = false, fixed = false, xcc = true;
  int W = 1, block = 256, slots = 0, refill = 6, nrep = 8;
  bool direct = false;
  size_t lds = 0;
};

// Validates a run's precision, tally and tuning fields and resolves the
// launch shape on a device with `num_cus` CUs of `lds_per_cu` bytes of LDS.
int resolve_cfg(int dim, int num_cus, int64_t lds_per_cu, bool fd, const vmc_run_desc *run, LaunchCfg *c) {
  (void)num_cus;
  if (run->precision != VMC_FP64 && run->precision != VMC_FP32) return VMC_ERR_INVALID_ARGUMENT;
  if (run->tally != VMC_TALLY_FP64 && run->tally != VMC_TALLY_FIXED64) return VMC_ERR_INVALID_ARGUMENT;
  c->dim = dim;
  c->f64 = run->precision == VMC_FP64;
  c->fixed = run->tally == VMC_TALLY_FIXED64;
  c->fd = fd;
  c->W = fd ? 2 : 1;
  // FP32 transport runs as one 1024-lane workgroup per CU (16 waves sharing
  // one sector cache), FP64 (occupancy 2 waves/SIMD) as 256-lane
  // workgroups, two per CU
  c->block = run->block_size > 0 ? run->block_size : (c->f64 ? 256 : 1024);
  if (c->block % 64 != 0 || c->block > (c->f64 ? 256 : 1024)) return VMC_ERR_INVALID_ARGUMENT;
  // replicas == 0: one per XCD, added with L2 atomics (choose_replicas may
  // make it a single one for the problem's size); > 0: that many, picked by
  // workgroup index and added with device-scope atomics
  c->xcc = run->replicas <= 0;
  c->nrep = c->xcc ? 8 : run->replicas;
  if (c->nrep > 4096) return VMC_ERR_INVALID_ARGUMENT;
  if (run->direct_flush != 0 && run->direct_flush != 1) return VMC_ERR_INVALID_ARGUMENT;
  c->direct = run->direct_flush == 1;
  const int rbytes = c->f64 ? 8 : 4;
  // sector cache: 0 = fill the CU's LDS left over by the lane buffers of
  // the workgroups that fit by registers, -1 = off
  c->slots = run->lds_slots < 0 ? 0 : run->lds_slots;
  if (run->lds_slots == 0) {
    const int waves_per_cu = c->f64 ? 8 : 16;
    const int bpc_target = std::max(1, waves_per_cu / (c->block / 64));
    const size_t per_block = (size_t)lds_per_cu / bpc_target;
    const size_t fixed_part = WalkLds(c->W, rbytes, 0, c->block).total + 1024;   // + ledger, alignment
    const size_t entry = (size_t)kSectorElems * c->W * 8 + 4;
    c->slots = per_block > fixed_part ? (int)((per_block - fixed_part) / entry) : 0;
  }
  if (c->slots > (1 << 20)) return VMC_ERR_INVALID_ARGUMENT;
  c->lds = WalkLds(c->W, rbytes, c->slots, c->block).total;
  // creating photons for several dead lanes at once keeps the creation
  // code's lanes busy; 6 was best on the 384k-tet box (4..16 within 1%)
  c->refill = run->refill_min > 0 ? std::min(run->refill_min, 64) : 6;
  return VMC_OK;
}

// replicas == 0: the tally copies follow the size of one copy.  Up to an
// XCD's L2 (4 MiB) one copy per XCD, added with L2 atomics, costs nothing
// extra; a copy that exceeds the L2 but fits the Infinity Cache comfortably
// (64 MiB) is better kept single, device-scope atomics meeting in the cache
// instead of eight copies streaming through it (the 2M-tet FD box: +18%);
// larger ones (many batch instances) go back to one per XCD, whose L2
// atomics beat device-scope atomics that reach HBM.
void choose_replicas(LaunchCfg &c, const vmc_run_desc *run, int64_t ne, int64_t nb) {
  if (run->replicas > 0) return;
  const double bytes = ((double)ne * c.W + (double)nb * 4) * 8;
  const bool single = bytes > 4.0 * (1 << 20) && bytes <= 64.0 * (1 << 20);
  c.xcc = !single;
  c.nrep = single ? 1 : 8;
}

// Sizes the arena for ne element and nb boundary slots and clears its
// tallies, ledger, counters and per-item loss counts.
void arena_reset(Arena &a, const LaunchCfg &c, int64_t ne, int64_t nb, size_t nitems) {
  a.ne = ne;
  a.nb = nb;
  a.W = c.W;
  a.nrep = c.nrep;
  a.nitems = nitems;
  const size_t ew = (size_t)ne * c.W * c.nrep, bw = (size_t)nb * 4 * c.nrep;
  a.etally.ensure(ew);
  a.btally.ensure(bw);
  a.raw_e.ensure((size_t)ne * c.W);
  a.raw_b.ensure((size_t)nb * 4);
  a.ledger.ensure(8);
  a.counters.ensure(8);
  a.item_loss.ensure(nitems);
  a.items.ensure(nitems);
  hip_check(launch_fill_u64(a.etally.p, ew, 0ull, a.stream), "fill");
  hip_check(launch_fill_u64(a.btally.p, bw, 0ull, a.stream), "fill");
  hip_check(launch_fill_u64(reinterpret_cast<unsigned long long *>(a.ledger.p), 8, 0ull, a.stream), "fill");
  hip_check(launch_fill_u64(a.counters.p, 8, 0ull, a.stream), "fill");
  hip_check(launch_fill_u64(a.item_loss.p, nitems, 0ull, a.stream), "fill");
}

// One persistent launch over `items` (vstart set here, chunk-aligned and in
// order).  Returns the kernel time from HIP events on the arena's stream and
// the grid size.
float launch_items(Arena &a, int num_cus, const LaunchCfg &c, std::vector<ItemRec> &items, int *grid_out) {
  uint64_t v = 0;
  for (ItemRec &it : items) {
    it.vstart = v;
    v += (it.count + kChunk - 1) / kChunk * kChunk;
  }
  if (v == 0) return 0.f;
  const bool batch = items.size() > 1;
  int bpc = 0;
  hip_check(walk_occupancy(c.dim, c.f64, c.fd, c.fixed, batch, c.block, c.lds, &bpc), "occupancy");
  if (bpc < 1) throw HipFail{hipErrorInvalidConfiguration, "workgroup does not fit on a CU (block size / LDS tally cache)"};
  const int grid = bpc * num_cus;
  hip_check(hipMemcpyAsync(a.items.p, items.data(), items.size() * sizeof(ItemRec), hipMemcpyHostToDevice, a.stream),
            "upload");
  hip_check(launch_fill_u64(a.counters.p, 1, 0ull, a.stream), "fill");   // dispenser
  WalkParams P{};
  P.items = a.items.p;
  P.nitems = (int32_t)items.size();
  P.vtotal = v;
  P.etally = a.etally.p;
  P.btally = a.btally.p;
  P.ledger = a.ledger.p;
  P.counters = a.counters.p;
  P.item_loss = a.item_loss.p;
  P.ne = a.ne;
  P.nb = a.nb;
  P.nrep = c.nrep;
  P.step_cap = 1u << 22;   // a photon still walking after 4M steps is counted as lost
  P.refill_min = c.refill;
  P.chunk = (int32_t)kChunk;
  P.xcc_replicas = c.xcc ? 1 : 0;
  P.lds_slots = c.slots;
  P.direct_flush = c.direct ? 1 : 0;
  P.weight0 = 0.001;   // MC3D.hpp:182
  P.chance = 0.1;      // MC3D.hpp:183
  P.weight0_f = (float)P.weight0;
  P.chance_f = (float)P.chance;
  P.fix_scale = 4294967296.0;
  hip_check(hipEventRecord(a.ev0, a.stream), "event");
  hip_check(launch_walk(c.dim, c.f64, c.fd, c.fixed, batch, grid, c.block, c.lds, a.stream, P), "walk");
  hip_check(hipEventRecord(a.ev1, a.stream), "event");
  hip_check(hipEventSynchronize(a.ev1), "sync");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a.ev0, a.ev1), "event");
  if (grid_out) *grid_out = grid;
  return ms;
}

// Sums the replicas and copies raw element words [Ne_pad*W], boundary
// words [nb*4], the ledger, the counters and the per-item loss counts to
// the host.
void arena_collect(Arena &a, const LaunchCfg &c, std::vector<double> &raw_e, std::vector<double> &raw_b,
                   double led[5], unsigned long long cnt[4], std::vector<unsigned long long> &loss) {
  const double inv = c.fixed ? 1.0 / 4294967296.0 : 1.0;
  const size_t ew = (size_t)a.ne * c.W, bw = (size_t)a.nb * 4;
  raw_e.resize(ew);
  raw_b.resize(bw);
  loss.resize(a.nitems);
  hip_check(launch_reduce(c.fixed, a.etally.p, a.raw_e.p, ew, c.nrep, ew, inv, a.stream), "reduce");
  hip_check(launch_reduce(c.fixed, a.btally.p, a.raw_b.p, bw, c.nrep, bw, inv, a.stream), "reduce");
  hip_check(hipMemcpyAsync(raw_e.data(), a.raw_e.p, ew * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipMemcpyAsync(raw_b.data(), a.raw_b.p, bw * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipMemcpyAsync(led, a.ledger.p, 5 * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipMemcpyAsync(cnt, a.counters.p, 4 * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  if (a.nitems)
    hip_check(hipMemcpyAsync(loss.data(), a.item_loss.p, a.nitems * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipStreamSynchronize(a.stream), "sync");
}

// Item record of a problem run: its device arrays, stream key and phase.
ItemRec make_item(const vmc_problem &m, const vmc_run_desc *run, bool f64, uint64_t count, uint64_t id0,
                  int64_t tbase, int64_t bbase) {
  ItemRec it{};
  it.count = count;
  it.id0 = id0;
  it.elems = f64 ? (const void *)m.med->elems64.p : (const void *)m.med->elems32.p;
  it.cdf = m.cdf.p;
  it.src = m.dsrc.p;
  it.bnd = m.dbnd.p;
  it.tbase = (int32_t)tbase;
  it.bbase = (int32_t)bbase;
  it.nsrc = (int32_t)m.src.size();
  it.key0 = (uint32_t)run->seed;
  it.key1 = (uint32_t)(run->seed >> 32);
  it.tag = run->stream_tag;
  double phase0 = run->phase0;
  if (phase0 < 0) phase0 += 2 * M_PI * ceil(-phase0 / (2 * M_PI));   // MC3D.cpp:124-126
  it.kfac = 2.0 * M_PI * run->f / kC0;
  it.cph0 = cos(phase0);
  it.sph0 = sin(phase0);
  return it;
}

// Splits a problem's raw element words (device order, from slot tbase) and
// boundary words (from slot bbase) and writes normalised outputs
// (MC3D.hpp:1793-1818).
void finish_outputs(const vmc_problem &m, double f, int64_t N, int W, const double *raw_e, const double *raw_b,
                    vmc_output *out) {
  const double omega = 2.0 * M_PI * f;
  const bool fd = omega > 0.0;
  std::vector<double> ER(m.Ne), EI(m.Ne, 0.0), EBR(m.Nb), EBI(m.Nb), DEBR(m.Nb), DEBI(m.Nb), kn(m.Ne);
  for (int64_t i = 0; i < m.Ne; ++i) {
    const int64_t dv = m.inv[i];
    ER[i] = raw_e[W * dv];
    if (fd) EI[i] = raw_e[W * dv + 1];
    kn[i] = omega / kC0 * m.med->n[i];   // k = omega / c0 * n (MC3D.hpp:660-662)
  }
  for (int64_t i = 0; i < m.Nb; ++i) {
    EBR[i] = raw_b[4 * i];
    EBI[i] = raw_b[4 * i + 1];
    DEBR[i] = raw_b[4 * i + 2];
    DEBI[i] = raw_b[4 * i + 3];
  }
  normalise(m, omega, kn.data(), N, ER.data(), EI.data(), EBR.data(), EBI.data(), DEBR.data(), DEBI.data());
  if (!out) return;
  if (out->ER) std::memcpy(out->ER, ER.data(), 8 * m.Ne);
  if (out->EI) std::memcpy(out->EI, EI.data(), 8 * m.Ne);
  if (out->EBR) std::memcpy(out->EBR, EBR.data(), 8 * m.Nb);
  if (out->EBI) std::memcpy(out->EBI, EBI.data(), 8 * m.Nb);
  if (out->DEBR) std::memcpy(out->DEBR, DEBR.data(), 8 * m.Nb);
  if (out->DEBI) std::memcpy(out->DEBI, DEBI.data(), 8 * m.Nb);
}

int64_t sector_round(int64_t n) { return (n + kSectorElems - 1) / kSectorElems * kSectorElems; }

Arena g_batch_arena;   // tallies of vmc_batch_run (one batch at a time)
std::mutex g_batch_mutex;

} // namespace

extern "C" {

int vmc_problem_check(const vmc_problem_desc *d, int64_t *HN_out, int64_t *nsources) {
  if (!d) return VMC_ERR_INVALID_ARGUMENT;
  vmc_problem m;
  const int st = host_prepare(d, &m);
  if (st != VMC_OK) return st;
  if (HN_out) std::memcpy(HN_out, m.HN.data(), sizeof(int64_t) * m.HN.size());
  if (nsources) *nsources = (int64_t)m.src.size();
  return VMC_OK;
}

int vmc_problem_create(const vmc_problem_desc *d, vmc_problem **out) {
  if (!d || !out) return VMC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  vmc_problem *m = new vmc_problem();
  const int st = host_prepare(d, m);
  if (st != VMC_OK) {
    delete m;
    return st;
  }
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      delete m;
      g_last_error = "no HIP device";
      return VMC_ERR_NO_DEVICE;
    }
    hip_check(hipGetDevice(&m->device), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, m->device), "hipGetDeviceProperties");
    m->num_cus = prop.multiProcessorCount;
    m->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor > 0 ? (int64_t)prop.maxSharedMemoryPerMultiProcessor
                                                             : (int64_t)prop.sharedMemPerBlock;
    m->arena.open();
    m->cdf.ensure(m->lsCDF.size());
    hip_check(hipMemcpy(m->cdf.p, m->lsCDF.data(), m->lsCDF.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
    m->dsrc.ensure(m->src.size());
    hip_check(hipMemcpy(m->dsrc.p, m->src.data(), m->src.size() * sizeof(SrcRec), hipMemcpyHostToDevice), "upload");
    m->dbnd.ensure(m->bnd.size());
    hip_check(hipMemcpy(m->dbnd.p, m->bnd.data(), m->bnd.size() * sizeof(BndRec), hipMemcpyHostToDevice), "upload");
  } catch (HipFail f) {
    int st = fail_hip(f);
    vmc_problem_destroy(m);
    return st;
  }
  *out = m;
  return VMC_OK;
}

int vmc_problem_create_variant(const vmc_problem *base, const vmc_variant_desc *d, vmc_problem **out) {
  if (!base || !d || !out) return VMC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::shared_ptr<MediumData> med = base->med;
  if (d->mua || d->mus || d->g || d->n) {
    med = std::make_shared<MediumData>();
    med->mua = d->mua ? std::vector<double>(d->mua, d->mua + base->Ne) : base->med->mua;
    med->mus = d->mus ? std::vector<double>(d->mus, d->mus + base->Ne) : base->med->mus;
    med->g = d->g ? std::vector<double>(d->g, d->g + base->Ne) : base->med->g;
    med->n = d->n ? std::vector<double>(d->n, d->n + base->Ne) : base->med->n;
  }
  vmc_problem *m = new vmc_problem(base->md, med);
  m->dim = base->dim;
  m->nv = base->nv;
  m->nbv = base->nbv;
  m->Ne = base->Ne;
  m->Nb = base->Nb;
  m->Nr = base->Nr;
  for (int k = 0; k < 3; ++k) m->shift[k] = base->shift[k];
  m->device = base->device;
  m->num_cus = base->num_cus;
  m->lds_per_cu = base->lds_per_cu;
  const int64_t Nb = base->Nb;
  m->BCn = base->BCn;
  m->BCType = d->BCType ? std::vector<char>(d->BCType, d->BCType + Nb) : base->BCType;
  m->BCLDirType = d->BCLightDirectionType ? std::vector<char>(d->BCLightDirectionType, d->BCLightDirectionType + Nb)
                                          : base->BCLDirType;
  m->BCIntensity = d->BCIntensity ? std::vector<double>(d->BCIntensity, d->BCIntensity + Nb) : base->BCIntensity;
  m->GaussianSigma = d->GaussianSigma ? std::vector<double>(d->GaussianSigma, d->GaussianSigma + Nb)
                                      : base->GaussianSigma;
  m->BCLNormal = d->BCLNormal ? std::vector<double>(d->BCLNormal, d->BCLNormal + Nb * base->dim) : base->BCLNormal_in;
  m->BCLNormal_in = m->BCLNormal;
  try {
    // the checks of ErrorChecks() that a source or medium change can fail
    bool has_ls = false, has_gauss = false;
    for (char c : m->BCType) {
      if (c != 'a') has_ls = true;
      if (c == 'g') has_gauss = true;
    }
    if (!has_ls) throw McError{NO_LIGHTSOURCE};
    if (m->dim == 2 && has_gauss && m->GaussianSigma.empty()) throw McError{NO_GAUSSIAN_SIGMA};
    build_light_sources(*m);
    prepare_directions(*m);
    build_source_records(*m);
    build_boundary(*m);
  } catch (McError e) {
    delete m;
    return e.code + 1;
  }
  try {
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    m->arena.open();
    m->cdf.ensure(m->lsCDF.size());
    hip_check(hipMemcpy(m->cdf.p, m->lsCDF.data(), m->lsCDF.size() * sizeof(double), hipMemcpyHostToDevice), "upload");
    m->dsrc.ensure(m->src.size());
    hip_check(hipMemcpy(m->dsrc.p, m->src.data(), m->src.size() * sizeof(SrcRec), hipMemcpyHostToDevice), "upload");
    m->dbnd.ensure(m->bnd.size());
    hip_check(hipMemcpy(m->dbnd.p, m->bnd.data(), m->bnd.size() * sizeof(BndRec), hipMemcpyHostToDevice), "upload");
  } catch (HipFail f) {
    int st = fail_hip(f);
    vmc_problem_destroy(m);
    return st;
  }
  *out = m;
  return VMC_OK;
}

int vmc_problem_set_medium(vmc_problem *m, const double *mua, const double *mus, const double *g,
                           const double *n) {
  if (!m) return VMC_ERR_INVALID_ARGUMENT;
  if (m->med.use_count() > 1) {   // shared with a variant: copy on write
    auto own = std::make_shared<MediumData>();
    own->mua = m->med->mua;
    own->mus = m->med->mus;
    own->g = m->med->g;
    own->n = m->med->n;
    m->med = own;
  }
  if (mua) m->med->mua.assign(mua, mua + m->Ne);
  if (mus) m->med->mus.assign(mus, mus + m->Ne);
  if (g) m->med->g.assign(g, g + m->Ne);
  if (n) m->med->n.assign(n, n + m->Ne);
  m->med->dirty32 = m->med->dirty64 = true;
  return VMC_OK;
}

int vmc_problem_get_hn(const vmc_problem *m, int64_t *HN) {
  if (!m || !HN) return VMC_ERR_INVALID_ARGUMENT;
  std::memcpy(HN, m->HN.data(), sizeof(int64_t) * m->HN.size());
  return VMC_OK;
}

int vmc_problem_run(vmc_problem *m, const vmc_run_desc *run, vmc_output *out, vmc_stats *stats,
                    vmc_progress_fn progress, void *user) {
  if (!m || !run || run->Nphoton < 0) return VMC_ERR_INVALID_ARGUMENT;
  const auto t_start = std::chrono::steady_clock::now();
  const bool fd = 2.0 * M_PI * run->f > 0.0;
  LaunchCfg c;
  const int st = resolve_cfg(m->dim, m->num_cus, m->lds_per_cu, fd, run, &c);
  if (st != VMC_OK) return st;
  const int64_t N = run->Nphoton;
  int aborted = 0, grid = 0;
  float kernel_ms = 0.f;
  double led[5] = {0, 0, 0, 0, 0};
  unsigned long long cnt[4] = {0, 0, 0, 0};
  std::vector<double> raw_e, raw_b;
  std::vector<unsigned long long> loss;
  try {
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    upload_params(*m, c.f64);
    choose_replicas(c, run, sector_round(m->Ne), m->Nb);
    arena_reset(m->arena, c, sector_round(m->Ne), m->Nb, 1);
    // photon ranges: one launch, or ~20 when a progress callback wants ticks
    const int nseg = progress ? 20 : 1;
    int64_t done = 0;
    for (int sgm = 0; sgm < nseg && done < N; ++sgm) {
      const int64_t end = (sgm == nseg - 1) ? N : (N * (sgm + 1)) / nseg;
      if (end <= done) continue;
      std::vector<ItemRec> items{make_item(*m, run, c.f64, (uint64_t)(end - done), (uint64_t)done, 0, 0)};
      kernel_ms += launch_items(m->arena, m->num_cus, c, items, &grid);
      done = end;
      if (progress && !progress(100.0 * (double)done / (double)(N > 0 ? N : 1), user)) {
        aborted = 1;
        break;
      }
    }
    arena_collect(m->arena, c, raw_e, raw_b, led, cnt, loss);
  } catch (HipFail f) {
    return fail_hip(f);
  }
  // like the reference, an aborted run still divides by the planned N
  finish_outputs(*m, run->f, N, c.W, raw_e.data(), raw_b.data(), out);
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->photons = N;
    stats->steps = (int64_t)cnt[1];
    stats->loss = (int64_t)cnt[2];
    stats->rechecks = (int64_t)cnt[3];
    stats->kernel_ms = kernel_ms;
    for (int q = 0; q < 5; ++q) stats->ledger[q] = led[q];
    stats->grid_blocks = grid;
    stats->block_size = c.block;
    stats->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  }
  return aborted ? VMC_ERR_ABORTED : VMC_OK;
}

int vmc_batch_run(vmc_batch_item *items, int32_t n, vmc_stats *stats) {
  if (!items || n <= 0) return VMC_ERR_INVALID_ARGUMENT;
  const auto t_start = std::chrono::steady_clock::now();
  vmc_problem *m0 = items[0].prob;
  if (!m0) return VMC_ERR_INVALID_ARGUMENT;
  // one launch per (dimension, CW/FD) group; precision, tally mode and
  // tuning come from the first item and must agree across the batch
  for (int32_t i = 0; i < n; ++i) {
    const vmc_batch_item &b = items[i];
    if (!b.prob || b.run.Nphoton < 0 || b.prob->device != m0->device) return VMC_ERR_INVALID_ARGUMENT;
    if (b.run.precision != items[0].run.precision || b.run.tally != items[0].run.tally)
      return VMC_ERR_INVALID_ARGUMENT;
    items[i].loss = 0;
    items[i].status = VMC_OK;
  }
  vmc_stats tot{};
  float kernel_ms = 0.f;
  try {
    hip_check(hipSetDevice(m0->device), "hipSetDevice");
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    if (!g_batch_arena.stream) g_batch_arena.open();
    for (int dim = 2; dim <= 3; ++dim) {
      for (int fdi = 0; fdi < 2; ++fdi) {
        std::vector<int32_t> group;
        for (int32_t i = 0; i < n; ++i)
          if (items[i].prob->dim == dim && ((2.0 * M_PI * items[i].run.f > 0.0) == (fdi == 1))) group.push_back(i);
        if (group.empty()) continue;
        LaunchCfg c;
        int st = resolve_cfg(dim, m0->num_cus, m0->lds_per_cu, fdi == 1, &items[0].run, &c);
        if (st != VMC_OK) return st;
        int64_t ne = 0, nb = 0;
        std::vector<int64_t> tb(group.size()), bb(group.size());
        for (size_t g = 0; g < group.size(); ++g) {
          vmc_problem &m = *items[group[g]].prob;
          upload_params(m, c.f64);
          tb[g] = ne;
          bb[g] = nb;
          ne += sector_round(m.Ne);
          nb += m.Nb;
        }
        if (ne > ((int64_t)1 << 30) / c.W || nb > ((int64_t)1 << 29)) return VMC_ERR_INVALID_ARGUMENT;
        choose_replicas(c, &items[0].run, ne, nb);
        arena_reset(g_batch_arena, c, ne, nb, group.size());
        std::vector<ItemRec> recs;
        for (size_t g = 0; g < group.size(); ++g) {
          const vmc_batch_item &b = items[group[g]];
          recs.push_back(make_item(*b.prob, &b.run, c.f64, (uint64_t)b.run.Nphoton, 0, tb[g], bb[g]));
        }
        int grid = 0;
        kernel_ms += launch_items(g_batch_arena, m0->num_cus, c, recs, &grid);
        std::vector<double> raw_e, raw_b;
        std::vector<unsigned long long> loss;
        double led[5];
        unsigned long long cnt[4];
        arena_collect(g_batch_arena, c, raw_e, raw_b, led, cnt, loss);
        for (size_t g = 0; g < group.size(); ++g) {
          vmc_batch_item &b = items[group[g]];
          finish_outputs(*b.prob, b.run.f, b.run.Nphoton, c.W, raw_e.data() + tb[g] * c.W, raw_b.data() + bb[g] * 4,
                         &b.out);
          b.loss = (int64_t)loss[g];
          tot.photons += b.run.Nphoton;
        }
        tot.steps += (int64_t)cnt[1];
        tot.loss += (int64_t)cnt[2];
        tot.rechecks += (int64_t)cnt[3];
        for (int q = 0; q < 5; ++q) tot.ledger[q] += led[q];
        tot.grid_blocks = grid;
        tot.block_size = c.block;
      }
    }
  } catch (HipFail f) {
    return fail_hip(f);
  }
  if (stats) {
    *stats = tot;
    stats->kernel_ms = kernel_ms;
    stats->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  }
  return VMC_OK;
}

void vmc_problem_destroy(vmc_problem *m) {
  if (!m) return;
  m->cdf.release();
  m->dsrc.release();
  m->dbnd.release();
  m->arena.close();
  delete m;
}

int vmc_monte_carlo(int32_t dim, int64_t Ne, int64_t Nb, int64_t Nr, const int64_t *H, const int64_t *HN,
                    const int64_t *BH, const double *r, const char *BCType, const double *BCIntensity,
                    const char *BCLightDirectionType, const double *BCLNormal, const double *BCn,
                    const double *mua, const double *mus, const double *g, const double *n, double f,
                    double phase0, int64_t Nphoton, const double *GaussianSigma, uint64_t seed,
                    int32_t precision, double *ER, double *EI, double *EBR, double *EBI, double *DEBR,
                    double *DEBI, int64_t *HN_out, int64_t *loss) {
  vmc_problem_desc d{};
  d.dim = dim;
  d.Ne = Ne; d.Nb = Nb; d.Nr = Nr;
  d.H = H; d.HN = HN; d.BH = BH; d.r = r;
  d.mua = mua; d.mus = mus; d.g = g; d.n = n;
  d.BCType = BCType; d.BCn = BCn; d.BCLNormal = BCLNormal;
  d.BCLightDirectionType = BCLightDirectionType;
  d.BCIntensity = BCIntensity;
  d.GaussianSigma = GaussianSigma;
  vmc_problem *p = nullptr;
  int st = vmc_problem_create(&d, &p);
  if (st != VMC_OK) return st;
  vmc_run_desc rd{};
  rd.f = f;
  rd.phase0 = phase0;
  rd.Nphoton = Nphoton;
  rd.seed = seed;
  rd.precision = precision;
  rd.tally = VMC_TALLY_FP64;
  vmc_output o{ER, EI, EBR, EBI, DEBR, DEBI};
  vmc_stats s{};
  st = vmc_problem_run(p, &rd, &o, &s, nullptr, nullptr);
  if (st == VMC_OK && HN_out) vmc_problem_get_hn(p, HN_out);
  if (loss) *loss = s.loss;
  vmc_problem_destroy(p);
  return st;
}

const char *vmc_status_string(int st) {
  static const char *mc[] = {
      // cpp/Errors.hpp:24-59 wording
      "No lightsource",
      "Scattering anisotropy array is not equal in size to the number of elements",
      "Absorption array is not equal in size to the number of elements",
      "Scattering coefficient array is not equal in size to the number of elements",
      "Refractive index array is not equal in size to the number of elements",
      "External refractive index array is not equal in size to the number of boundary elements",
      "No external refractive index array given",
      "The boundary condition array is not equal in size to the number of boundary elements",
      "The light direction array is not equal in size to the number of boundary elements",
      "The light direction type array is not equal in size to the number of boundary elements",
      "H matrix contain indices that are not in the coordinate matrix",
      "BH matrix contain indices that are not in the coordinate matrix",
      "Ill defined mesh - cannot find nodes. Check Mesh.",
      "Ill defined mesh - some boundary elements are trapped between two elements. A boundary element must belong to a single element. Check Mesh.",
      "A gaussian light source exists but no sigma parameter was provided",
      "Ill defined mesh - a complete boundary cannot be constructed. Check Mesh.",
  };
  if (st == VMC_OK) return "Success";
  if (st >= 1 && st <= 16) return mc[st - 1];
  switch (st) {
  case VMC_ERR_UNDEFINED_DIRECTION:
    return "A light source of type L, I or p needs a light direction (type a, r or p)";
  case VMC_ERR_INVALID_ARGUMENT: return "Invalid argument";
  case VMC_ERR_HIP: return "HIP runtime error";
  case VMC_ERR_OUT_OF_MEMORY: return "Out of device memory";
  case VMC_ERR_NO_DEVICE: return "No GPU device";
  case VMC_ERR_ABORTED: return "Computation aborted by the progress callback";
  default: return "Undefined error";
  }
}

const char *vmc_last_error(void) { return g_last_error.c_str(); }

int vmc_device_info(char *buf, int32_t buflen, int32_t *num_cus) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return VMC_ERR_NO_DEVICE;
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return VMC_ERR_HIP;
  if (buf && buflen > 0)
    snprintf(buf, (size_t)buflen, "%s (%s), %d CUs, %.1f GiB", prop.name, prop.gcnArchName,
             prop.multiProcessorCount, (double)prop.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
  if (num_cus) *num_cus = prop.multiProcessorCount;
  return VMC_OK;
}

const char *vmc_version(void) { return "valomc-mi355x 0.1 (gfx950)"; }

} // extern "C"

// ---------------------------------------------------------------------------
// Serving loop: continuous batching of forward solves
//
// A worker thread owns one stream and a persistent tally arena.  Requests
// wait in FIFO order until their tallies fit in the arena (admission by
// device-memory footprint); each launch then takes one (dimension, CW/FD)
// group of the admitted requests and advances every request in it by a
// share of the launch's photon budget, so a request that arrives while
// others run joins at the next launch and launch times stay bounded.  A
// request's photons keep their stream ids (seed, tag, photon index) however
// they are sliced, so its result does not depend on what it shared the GPU
// with: bit for bit in fixed-point mode.
// ---------------------------------------------------------------------------

namespace {

struct Request {
  uint64_t ticket = 0;
  vmc_problem *prob = nullptr;
  vmc_run_desc run{};
  vmc_output out{};
  bool fd = false;
  int64_t launched = 0;          // photons handed to launches so far
  int64_t eoff = -1, ewords = 0; // element words in each arena replica
  int64_t boff = -1;             // first boundary slot
  int64_t loss = 0;
  double kernel_ms = 0;
  int status = VMC_OK;
  bool finished = false;
  std::chrono::steady_clock::time_point t_submit, t_done;
};

// first fit over [0, cap) in whole units of `unit`
struct RangeAlloc {
  std::map<int64_t, int64_t> free_;   // start -> length
  void reset(int64_t cap) {
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  int64_t alloc(int64_t len) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < len) continue;
      const int64_t start = it->first, rest = it->second - len;
      free_.erase(it);
      if (rest > 0) free_[start + len] = rest;
      return start;
    }
    return -1;
  }
  void release(int64_t start, int64_t len) {
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + len) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        prev->second += len;
        return;
      }
    }
    free_[start] = len;
  }
};

} // namespace

struct vmc_server {
  vmc_server_desc desc{};
  int device = 0, num_cus = 0;
  int64_t lds_per_cu = 65536;
  int nrep = 8;
  int64_t cap_words = 0, cap_b = 0;   // per replica
  Arena arena;
  RangeAlloc ealloc, balloc;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<Request *> waiting;
  std::vector<Request *> active;
  std::map<uint64_t, std::unique_ptr<Request>> all;
  uint64_t next_ticket = 1;
  bool stop = false;
  int turn = 0;
  int64_t launches = 0, photons = 0;
  double busy_ms = 0;
  uint64_t epoch = 0;   // bumped by every submit and every release of arena ranges
  std::thread worker;
  // finisher: reduces, copies and normalises completed requests on its own
  // stream while the worker keeps launching, then returns their ranges
  std::deque<Request *> to_finish;
  std::condition_variable fin_cv;
  bool fin_stop = false;
  hipStream_t fstream = nullptr;
  DevBuf<double> fraw_e, fraw_b;
  std::thread finisher;
};

namespace {

const int64_t kWordUnit = 16;   // arena element words are handed out in whole FD sectors

void server_finish(vmc_server *s, Request *r) {
  std::vector<double> raw_e((size_t)r->ewords), raw_b((size_t)r->prob->Nb * 4);
  const bool fixed = s->desc.tally == VMC_TALLY_FIXED64;
  const double inv = fixed ? 1.0 / 4294967296.0 : 1.0;
  Arena &a = s->arena;
  s->fraw_e.ensure((size_t)r->ewords);
  s->fraw_b.ensure((size_t)r->prob->Nb * 4);
  hip_check(launch_reduce(fixed, a.etally.p + r->eoff, s->fraw_e.p, (size_t)r->ewords, s->nrep, (size_t)s->cap_words,
                          inv, s->fstream), "reduce");
  hip_check(launch_reduce(fixed, a.btally.p + r->boff * 4, s->fraw_b.p, (size_t)r->prob->Nb * 4, s->nrep,
                          (size_t)s->cap_b * 4, inv, s->fstream), "reduce");
  hip_check(hipMemcpyAsync(raw_e.data(), s->fraw_e.p, raw_e.size() * 8, hipMemcpyDeviceToHost, s->fstream), "d2h");
  hip_check(hipMemcpyAsync(raw_b.data(), s->fraw_b.p, raw_b.size() * 8, hipMemcpyDeviceToHost, s->fstream), "d2h");
  hip_check(hipStreamSynchronize(s->fstream), "sync");
  finish_outputs(*r->prob, r->run.f, r->run.Nphoton, r->fd ? 2 : 1, raw_e.data(), raw_b.data(), &r->out);
}

void finisher_loop(vmc_server *s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    Request *r = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->fin_cv.wait(lk, [&] { return s->fin_stop || !s->to_finish.empty(); });
      if (s->to_finish.empty()) return;
      r = s->to_finish.front();
      s->to_finish.pop_front();
    }
    if (r->status == VMC_OK) {
      try {
        server_finish(s, r);
      } catch (HipFail f) {
        r->status = fail_hip(f);
      }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t ew = (r->ewords + kWordUnit - 1) / kWordUnit * kWordUnit;
    s->ealloc.release(r->eoff, ew);
    s->balloc.release(r->boff, r->prob->Nb);
    s->epoch++;
    r->finished = true;
    r->t_done = std::chrono::steady_clock::now();
    s->done_cv.notify_all();
    s->cv.notify_all();
  }
}

void server_loop(vmc_server *s) {
  (void)hipSetDevice(s->device);
  uint64_t seen = ~0ull;   // epoch of the last admission attempt that found no room
  for (;;) {
    std::vector<Request *> admitted, group;
    int gdim = 3;
    bool gfd = false;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] {
        return (s->stop && s->waiting.empty()) || !s->active.empty() || (!s->waiting.empty() && s->epoch != seen);
      });
      if (s->stop && s->waiting.empty() && s->active.empty()) return;
      // admission in arrival order while the tallies fit
      while (!s->waiting.empty()) {
        Request *r = s->waiting.front();
        const int64_t ew = (r->ewords + kWordUnit - 1) / kWordUnit * kWordUnit;
        const int64_t eo = s->ealloc.alloc(ew);
        if (eo < 0) break;
        const int64_t bo = s->balloc.alloc(r->prob->Nb);
        if (bo < 0) {
          s->ealloc.release(eo, ew);
          break;
        }
        r->eoff = eo;
        r->boff = bo;
        s->waiting.pop_front();
        s->active.push_back(r);
        admitted.push_back(r);
      }
      if (s->active.empty()) {
        seen = s->epoch;   // nothing fits until the finisher returns ranges
        continue;
      }
      // the next group in turn among the (dimension, CW/FD) groups with work
      for (int k = 0; k < 4; ++k) {
        const int g = (s->turn + k) % 4;
        const int dim = g < 2 ? 3 : 2;
        const bool fd = (g & 1) != 0;
        for (Request *r : s->active)
          if (r->prob->dim == dim && r->fd == fd && r->launched < r->run.Nphoton) group.push_back(r);
        if (!group.empty()) {
          gdim = dim;
          gfd = fd;
          s->turn = (g + 1) % 4;
          break;
        }
      }
    }
    LaunchCfg c;
    vmc_run_desc tmpl{};
    tmpl.precision = s->desc.precision;
    tmpl.tally = s->desc.tally;
    std::vector<Request *> done_now;
    try {
      Arena &a = s->arena;
      for (Request *r : admitted) {
        hip_check(launch_fill_ranges(a.etally.p + r->eoff, (size_t)r->ewords, s->nrep, (size_t)s->cap_words, 0ull,
                                     a.stream), "fill");
        hip_check(launch_fill_ranges(a.btally.p + r->boff * 4, (size_t)r->prob->Nb * 4, s->nrep,
                                     (size_t)s->cap_b * 4, 0ull, a.stream), "fill");
      }
      if (!group.empty()) {
        const int st = resolve_cfg(gdim, s->num_cus, s->lds_per_cu, gfd, &tmpl, &c);
        if (st != VMC_OK) throw HipFail{hipErrorInvalidValue, "server launch configuration"};
        // fair shares of the launch budget
        const int64_t share = std::max(s->desc.min_slice, s->desc.launch_photons / (int64_t)group.size());
        std::vector<ItemRec> items;
        std::vector<int64_t> slice;
        for (Request *r : group) {
          const int64_t n = std::min(share, r->run.Nphoton - r->launched);
          upload_params(*r->prob, c.f64);
          items.push_back(make_item(*r->prob, &r->run, c.f64, (uint64_t)n, (uint64_t)r->launched, r->eoff / c.W,
                                    r->boff));
          slice.push_back(n);
        }
        a.ne = s->cap_words / c.W;
        a.nb = s->cap_b;
        a.nitems = items.size();
        a.items.ensure(items.size());
        a.item_loss.ensure(items.size());
        hip_check(launch_fill_u64(a.item_loss.p, items.size(), 0ull, a.stream), "fill");
        int grid = 0;
        const float ms = launch_items(a, s->num_cus, c, items, &grid);
        std::vector<unsigned long long> loss(items.size());
        hip_check(hipMemcpyAsync(loss.data(), a.item_loss.p, loss.size() * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
        hip_check(hipStreamSynchronize(a.stream), "sync");
        std::lock_guard<std::mutex> lk(s->mu);
        s->launches++;
        s->busy_ms += ms;
        for (size_t k = 0; k < group.size(); ++k) {
          Request *r = group[k];
          r->launched += slice[k];
          r->loss += (int64_t)loss[k];
          r->kernel_ms += ms;
          s->photons += slice[k];
          if (r->launched >= r->run.Nphoton) done_now.push_back(r);
        }
      }
      // requests with nothing to launch (zero photons) finish right away
      {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Request *r : s->active)
          if (r->run.Nphoton == 0 && std::find(done_now.begin(), done_now.end(), r) == done_now.end())
            done_now.push_back(r);
      }
    } catch (HipFail f) {
      const int st = fail_hip(f);
      std::lock_guard<std::mutex> lk(s->mu);
      for (Request *r : group) {
        r->status = st;
        if (std::find(done_now.begin(), done_now.end(), r) == done_now.end()) done_now.push_back(r);
      }
    }
    // hand completed requests to the finisher; their last launch has
    // completed, so their tallies are final
    std::lock_guard<std::mutex> lk(s->mu);
    for (Request *r : done_now) {
      s->active.erase(std::find(s->active.begin(), s->active.end(), r));
      s->to_finish.push_back(r);
    }
    if (!done_now.empty()) s->fin_cv.notify_all();
  }
}

} // namespace

extern "C" {

int vmc_server_create(const vmc_server_desc *d, vmc_server **out) {
  if (!d || !out) return VMC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (d->precision != VMC_FP64 && d->precision != VMC_FP32) return VMC_ERR_INVALID_ARGUMENT;
  if (d->tally != VMC_TALLY_FP64 && d->tally != VMC_TALLY_FIXED64) return VMC_ERR_INVALID_ARGUMENT;
  std::unique_ptr<vmc_server> s(new vmc_server());
  s->desc = *d;
  if (s->desc.arena_bytes <= 0) s->desc.arena_bytes = (int64_t)4 << 30;
  if (s->desc.launch_photons <= 0) s->desc.launch_photons = 20000000;
  if (s->desc.min_slice <= 0) s->desc.min_slice = 262144;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
      g_last_error = "no HIP device";
      return VMC_ERR_NO_DEVICE;
    }
    hip_check(hipGetDevice(&s->device), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, s->device), "hipGetDeviceProperties");
    s->num_cus = prop.multiProcessorCount;
    s->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor > 0 ? (int64_t)prop.maxSharedMemoryPerMultiProcessor
                                                             : (int64_t)prop.sharedMemPerBlock;
    // 7/8 of the budget for element words, 1/8 for boundary words, over
    // one replica per XCD
    s->nrep = 8;
    const int64_t per_rep = s->desc.arena_bytes / 8 / s->nrep;
    s->cap_words = std::min<int64_t>(per_rep * 7 / 8, (int64_t)1 << 30) / kWordUnit * kWordUnit;
    s->cap_b = std::min<int64_t>(per_rep / 8 / 4, (int64_t)1 << 28);
    if (s->cap_words <= 0 || s->cap_b <= 0) return VMC_ERR_INVALID_ARGUMENT;
    s->arena.open();
    hip_check(hipStreamCreateWithFlags(&s->fstream, hipStreamNonBlocking), "hipStreamCreate");
    s->arena.W = 1;
    s->arena.nrep = s->nrep;
    s->arena.etally.ensure((size_t)s->cap_words * s->nrep);
    s->arena.btally.ensure((size_t)s->cap_b * 4 * s->nrep);
    s->arena.ledger.ensure(8);
    s->arena.counters.ensure(8);
    hip_check(launch_fill_u64(reinterpret_cast<unsigned long long *>(s->arena.ledger.p), 8, 0ull, s->arena.stream),
              "fill");
    hip_check(launch_fill_u64(s->arena.counters.p, 8, 0ull, s->arena.stream), "fill");
    hip_check(hipStreamSynchronize(s->arena.stream), "sync");
  } catch (HipFail f) {
    const int st = fail_hip(f);
    s->arena.close();
    if (s->fstream) (void)hipStreamDestroy(s->fstream);
    return st;
  }
  s->ealloc.reset(s->cap_words);
  s->balloc.reset(s->cap_b);
  vmc_server *raw = s.release();
  raw->worker = std::thread(server_loop, raw);
  raw->finisher = std::thread(finisher_loop, raw);
  *out = raw;
  return VMC_OK;
}

int vmc_server_submit(vmc_server *s, vmc_problem *prob, const vmc_run_desc *run, const vmc_output *out,
                      uint64_t *ticket) {
  if (!s || !prob || !run || !ticket || run->Nphoton < 0 || prob->device != s->device) return VMC_ERR_INVALID_ARGUMENT;
  std::unique_ptr<Request> r(new Request());
  r->prob = prob;
  r->run = *run;
  r->run.precision = s->desc.precision;
  r->run.tally = s->desc.tally;
  if (out) r->out = *out;
  r->fd = 2.0 * M_PI * run->f > 0.0;
  r->ewords = sector_round(prob->Ne) * (r->fd ? 2 : 1);
  if (r->ewords > s->cap_words || prob->Nb > s->cap_b) {
    g_last_error = "request larger than the server's tally arena";
    return VMC_ERR_OUT_OF_MEMORY;
  }
  r->t_submit = std::chrono::steady_clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  r->ticket = s->next_ticket++;
  *ticket = r->ticket;
  s->epoch++;
  s->waiting.push_back(r.get());
  s->all[r->ticket] = std::move(r);
  s->cv.notify_all();
  return VMC_OK;
}

int vmc_server_poll(vmc_server *s, uint64_t ticket, int32_t *done) {
  if (!s || !done) return VMC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return VMC_ERR_INVALID_ARGUMENT;
  *done = it->second->finished ? 1 : 0;
  return VMC_OK;
}

int vmc_server_wait(vmc_server *s, uint64_t ticket, vmc_stats *stats) {
  if (!s) return VMC_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return VMC_ERR_INVALID_ARGUMENT;
  Request *r = it->second.get();
  s->done_cv.wait(lk, [&] { return r->finished; });
  const int st = r->status;
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->photons = r->run.Nphoton;
    stats->loss = r->loss;
    stats->kernel_ms = r->kernel_ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  }
  s->all.erase(it);
  return st;
}

int vmc_server_counters(vmc_server *s, int64_t *launches, int64_t *photons, double *busy_ms, int32_t *waiting,
                        int32_t *active) {
  if (!s) return VMC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  if (launches) *launches = s->launches;
  if (photons) *photons = s->photons;
  if (busy_ms) *busy_ms = s->busy_ms;
  if (waiting) *waiting = (int32_t)s->waiting.size();
  if (active) *active = (int32_t)s->active.size();
  return VMC_OK;
}

void vmc_server_destroy(vmc_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->fin_stop = true;
    s->fin_cv.notify_all();
  }
  if (s->finisher.joinable()) s->finisher.join();
  s->fraw_e.release();
  s->fraw_b.release();
  if (s->fstream) (void)hipStreamDestroy(s->fstream);
  s->arena.close();
  delete s;
}

} // extern "C"
