// ErrorChecks() + the host part of Init() of the reference
// (cpp/3d/MC3D.hpp:460-738, cpp/2d/MC2D.hpp:390-643): same checks in the
// same order with the same error codes, the same neighbour table, the same
// light-source list and CDF, the same BCLNormal preprocessing; then the
// device element order and the element, source and boundary records the
// kernels walk on.
#include "vmc_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <system_error>
#include <thread>

namespace vmc {

// Ranges [0, n) is split into by parallel_ranges: up to 16 host threads,
// at least 65536 items each.
inline int64_t range_count(int64_t n) {
  const int64_t hw = (int64_t)std::max(1u, std::thread::hardware_concurrency());
  return std::min<int64_t>(std::min<int64_t>(hw, 16), std::max<int64_t>(1, n / 65536));
}

// Runs fn(lo, hi) over the range_count(n) contiguous ranges of [0, n), each
// on a host thread (ranges whose thread cannot be started run on the
// caller's); the first exception a range throws is rethrown here after all
// ranges end.
template <class Fn>
void parallel_ranges(int64_t n, Fn fn) {
  const int64_t nt = range_count(n);
  std::vector<std::exception_ptr> err((size_t)nt);
  auto range = [&](int64_t t) {
    try {
      fn(n * t / nt, n * (t + 1) / nt);
    } catch (...) {
      err[(size_t)t] = std::current_exception();
    }
  };
  std::vector<std::thread> pool;
  for (int64_t t = 1; t < nt; ++t) {
    try {
      pool.emplace_back(range, t);
    } catch (const std::system_error &) {
      range(t);
    }
  }
  range(0);
  for (auto &th : pool) th.join();
  for (auto &e : err)
    if (e) std::rethrow_exception(e);
}

thread_local std::string g_last_error;


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
// different matches are an inconsistent mesh.  Found here by bucketing face
// keys by their smallest vertex (a counting sort, O(F)) and sorting each
// small bucket, instead of the reference's O(sum of vertex degree^2).
// ---------------------------------------------------------------------------

struct FaceKey {
  int64_t b, c;    // the face's other vertices, ascending (c = -1 for 2D edges)
  int64_t owner;   // (el << 2 | f) for element faces, -1 - ib for boundary entries
  bool operator<(const FaceKey &o) const {
    if (b != o.b) return b < o.b;
    if (c != o.c) return c < o.c;
    return owner < o.owner;
  }
  bool same(const FaceKey &o) const { return b == o.b && c == o.c; }
};

inline void sort3(int64_t &a, int64_t &b, int64_t &c) {
  if (a > b) std::swap(a, b);
  if (b > c) std::swap(b, c);
  if (a > b) std::swap(a, b);
}

void build_neighbours(vmc_problem &m) {
  if ((int64_t)m.HN.size() == m.Ne * m.nv) return;   // given (MC3D.hpp:839)
  m.HN.assign(m.Ne * m.nv, kUnset);
  // key k of the element faces, then of the boundary entries: smallest
  // vertex a, the others b <= c (vertex indices are checked against Nr)
  auto key = [&](int64_t k, int64_t &a, FaceKey &fk) {
    if (k < m.Ne * m.nv) {
      const int64_t e = k / m.nv;
      const int f = (int)(k % m.nv);
      if (m.dim == 3) {
        a = m.h(e, kFace3[f][0]); fk.b = m.h(e, kFace3[f][1]); fk.c = m.h(e, kFace3[f][2]);
        sort3(a, fk.b, fk.c);
      } else {
        a = std::min(m.h(e, kEdge2[f][0]), m.h(e, kEdge2[f][1]));
        fk.b = std::max(m.h(e, kEdge2[f][0]), m.h(e, kEdge2[f][1]));
        fk.c = -1;
      }
      fk.owner = (e << 2) | f;
    } else {
      const int64_t b = k - m.Ne * m.nv;
      if (m.dim == 3) {
        a = m.bh(b, 0); fk.b = m.bh(b, 1); fk.c = m.bh(b, 2);
        sort3(a, fk.b, fk.c);
      } else {
        a = std::min(m.bh(b, 0), m.bh(b, 1));
        fk.b = std::max(m.bh(b, 0), m.bh(b, 1));
        fk.c = -1;
      }
      fk.owner = -1 - b;
    }
  };
  // counting sort by a, on several threads: the order inside a bucket
  // depends on the threads, but each bucket is fully sorted below
  const int64_t nk = m.Ne * m.nv + m.Nb;
  std::vector<int64_t> start(m.Nr + 1, 0);
  parallel_ranges(nk, [&](int64_t k0, int64_t k1) {
    FaceKey fk;
    int64_t a;
    for (int64_t k = k0; k < k1; ++k) {
      key(k, a, fk);
      __atomic_fetch_add(&start[a + 1], 1, __ATOMIC_RELAXED);
    }
  });
  for (int64_t v = 0; v < m.Nr; ++v) start[v + 1] += start[v];
  std::unique_ptr<FaceKey[]> keys(new FaceKey[nk]);   // every entry written below
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    parallel_ranges(nk, [&](int64_t k0, int64_t k1) {
      FaceKey fk;
      int64_t a;
      for (int64_t k = k0; k < k1; ++k) {
        key(k, a, fk);
        keys[__atomic_fetch_add(&fill[a], 1, __ATOMIC_RELAXED)] = fk;
      }
    });
  }
  const int dup = m.dim == 3 ? INCONSISTENT_MESH_DUPLICATE_NEIGHBORS : INCONSISTENT_MESH;
  // buckets are independent: ranges of vertices on separate threads, each
  // face written by the one bucket that holds its key
  parallel_ranges(m.Nr, [&](int64_t v0, int64_t v1) {
  for (int64_t v = v0; v < v1; ++v) {
    const size_t lo = (size_t)start[v], hi = (size_t)start[v + 1];
    std::sort(keys.get() + lo, keys.get() + hi);
    size_t i = lo;
    while (i < hi) {
      size_t j = i + 1;
      while (j < hi && keys[j].same(keys[i])) ++j;
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
  }
  });
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

// Tally sectors are runs of kSectorElems device slots (one 64-byte line of
// CW tally words, two in the frequency domain).  The elements of one Morton
// cell -- a hexahedron's six tetrahedra or a quadrilateral's two triangles
// in a structured mesh -- are packed into one sector: cells go into sectors
// in curve order, next fit (a cell that does not fit in what is left of a
// sector starts a new one; a cell larger than a sector fills whole sectors
// first), and the slots left empty are padding that no neighbour code or
// source refers to.  A photon then changes sector only when it leaves a
// cell, about one run in six less often than with sectors cut every eight
// elements across cell boundaries.  Inside a cell the elements follow a
// face-adjacent path, and a sector of five or more elements puts the first
// half of that path in its first four slots and the rest in the last four,
// so a run through neighbouring elements mostly touches one of the two
// frequency-domain lines.  Meshes whose padded order would exceed the
// device id range keep the dense order.
static void pack_sectors(vmc_problem &m, const std::vector<std::pair<uint64_t, int32_t>> &order) {
  const int S = kSectorElems;
  std::vector<int32_t> perm;
  perm.reserve((size_t)(m.Ne + m.Ne / 2 + S));
  auto emit = [&](const int32_t *els, int n) {
    const size_t base = perm.size();
    perm.resize(base + S, -1);
    const int first = n <= S / 2 ? n : (n + 1) / 2;
    for (int i = 0; i < n; ++i) perm[base + (i < first ? i : S / 2 + (i - first))] = els[i];
  };
  std::vector<int32_t> mark(m.Ne, -1);
  std::vector<char> seen(m.Ne, 0);
  std::vector<int32_t> path, cur;
  int32_t stamp = 0;
  for (int64_t b = 0; b < m.Ne; ++stamp) {
    int64_t e = b + 1;
    while (e < m.Ne && order[e].first == order[b].first) ++e;
    // a face-adjacent path through the cell, from an element with the
    // fewest neighbours inside it; a dead end continues at the first
    // element not yet on the path
    for (int64_t i = b; i < e; ++i) mark[order[i].second] = stamp;
    auto inside = [&](int64_t v) { return v >= 0 && v < m.Ne && mark[v] == stamp && !seen[v]; };
    int32_t x = order[b].second;
    int fewest = m.nv + 1;
    for (int64_t i = b; i < e && e - b > 1; ++i) {
      const int32_t el = order[i].second;
      int c = 0;
      for (int f = 0; f < m.nv; ++f) c += inside(m.hn(el, f)) ? 1 : 0;
      if (c < fewest) { fewest = c; x = el; }
    }
    path.clear();
    int64_t scan = b;
    for (;;) {
      seen[x] = 1;
      path.push_back(x);
      if ((int64_t)path.size() == e - b) break;
      int32_t nx = -1;
      for (int f = 0; f < m.nv && nx < 0; ++f)
        if (inside(m.hn(x, f))) nx = (int32_t)m.hn(x, f);
      if (nx < 0) {
        while (seen[order[scan].second]) ++scan;
        nx = order[scan].second;
      }
      x = nx;
    }
    const int n = (int)path.size();
    if (n <= S) {
      if ((int)cur.size() + n > S) {
        emit(cur.data(), (int)cur.size());
        cur.clear();
      }
      cur.insert(cur.end(), path.begin(), path.end());
    } else {
      if (!cur.empty()) emit(cur.data(), (int)cur.size());
      int i = 0;
      for (; i + S <= n; i += S) emit(path.data() + i, S);
      cur.assign(path.begin() + i, path.end());
    }
    b = e;
  }
  if (!cur.empty()) emit(cur.data(), (int)cur.size());
  if ((int64_t)perm.size() >= kMaxElements) {   // dense: order as sorted, sectors of 8
    perm.assign((size_t)((m.Ne + S - 1) / S * S), -1);
    for (int64_t d = 0; d < m.Ne; ++d) perm[d] = order[d].second;
  }
  m.md->slots = (int64_t)perm.size();
  m.perm.swap(perm);
  m.inv.assign(m.Ne, -1);
  for (int64_t d = 0; d < m.md->slots; ++d)
    if (m.perm[d] >= 0) m.inv[m.perm[d]] = (int32_t)d;
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
  parallel_ranges(m.Ne, [&](int64_t e0, int64_t e1) {
  for (int64_t e = e0; e < e1; ++e) {
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
  });
  // (key, element) pairs sort by key, ties by element: ranges sorted on
  // separate threads, then merged pairwise (the pairs are distinct, so the
  // order is the one a single sort gives)
  std::vector<std::pair<uint64_t, int32_t>> order(m.Ne);
  for (int64_t e = 0; e < m.Ne; ++e) order[e] = {key[e], (int32_t)e};
  parallel_ranges(m.Ne, [&](int64_t e0, int64_t e1) { std::sort(order.begin() + e0, order.begin() + e1); });
  std::vector<int64_t> bounds{0};   // the ranges parallel_ranges sorted
  const int64_t nt = range_count(m.Ne);
  for (int64_t t = 1; t <= nt; ++t) bounds.push_back(m.Ne * t / nt);
  while (bounds.size() > 2) {
    std::vector<int64_t> next{0};
    std::vector<std::thread> pool;
    for (size_t k = 0; k + 2 < bounds.size(); k += 2) {
      const int64_t b0 = bounds[k], b1 = bounds[k + 1], b2 = bounds[k + 2];
      auto merge = [&order, b0, b1, b2] {
        std::inplace_merge(order.begin() + b0, order.begin() + b1, order.begin() + b2);
      };
      try {
        pool.emplace_back(merge);
      } catch (const std::system_error &) {
        merge();
      }
      next.push_back(b2);
    }
    if (bounds.size() % 2 == 0) next.push_back(bounds.back());   // an odd range left over
    for (auto &th : pool) th.join();
    bounds.swap(next);
  }
  pack_sectors(m, order);
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
  out.assign((size_t)m.md->slots, Elem<NF, R>{});   // padding slots stay zero and are never reached
  parallel_ranges(m.md->slots, [&](int64_t d0, int64_t d1) {
  for (int64_t dev = d0; dev < d1; ++dev) {
    const int64_t e = m.perm[dev];
    if (e < 0) continue;
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
  });
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

// Normalisation of MonteCarlo (MC3D.hpp:1793-1818, MC2D.hpp:1673-1703),
// fused with the gather from device element order: raw element words
// (W per element, device order) and boundary words (EBR, EBI, DEBR, DEBI
// per face) become the caller's outputs in one pass, with the reference's
// expressions.  Words the reference leaves unnormalised (EBI in CW, DEBR
// and DEBI in 3D) are copied as they are; NULL outputs are skipped.
void normalise_into(const vmc_problem &m, double omega, double c0, int64_t N, int W, const double *raw_e,
                    const double *raw_b, vmc_output *out) {
  const bool d3 = m.dim == 3, fd = omega > 0.0;
  if (out->ER || out->EI)
    parallel_ranges(m.Ne, [&](int64_t i0, int64_t i1) {
      for (int64_t i = i0; i < i1; ++i) {
        const int64_t dv = m.inv[i];
        const double a = raw_e[W * dv], mua = m.med->mua[i];
        double er, ei = 0.0;
        if (!fd) {
          if (mua > 0.0) er = a / (mua * m.vol[i] * (double)N);
          else er = a / (m.vol[i] * (double)N);
        } else {
          const double b = raw_e[W * dv + 1];
          const double k = omega / c0 * m.med->n[i];   // k = omega / c0 * n (MC3D.hpp:660-662)
          er = (b * k + a * mua) / (pow(k, 2) + pow(mua, 2)) / (double)N / m.vol[i];
          ei = -(a * k - b * mua) / (pow(k, 2) + pow(mua, 2)) / (double)N / m.vol[i];
        }
        if (out->ER) out->ER[i] = er;
        if (out->EI) out->EI[i] = ei;
      }
    });
  for (int64_t i = 0; i < m.Nb; ++i) {
    const double na = (double)N * m.area[i];
    const double *w = raw_b + 4 * i;
    if (out->EBR) out->EBR[i] = w[0] / na;
    if (out->EBI) out->EBI[i] = fd ? w[1] / na : w[1];
    if (out->DEBR) out->DEBR[i] = d3 ? w[2] : w[2] / na;
    if (out->DEBI) out->DEBI[i] = (d3 || !fd) ? w[3] : w[3] / na;
  }
}

int fail_hip(const HipFail &f) {
  g_last_error = std::string(f.what) + ": " + hipGetErrorString(f.e);
  return f.e == hipErrorOutOfMemory ? VMC_ERR_OUT_OF_MEMORY : VMC_ERR_HIP;
}



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

// BuildNeighbourhoods() alone (MC3D.hpp:835-871, MC2D.hpp:724-773): the
// neighbour table of H and the optional boundary entries BH, nothing else
// of ErrorChecks() or Init().
int neighbours_only(int32_t dim, int64_t Ne, int64_t Nb, int64_t Nr, const int64_t *H, const int64_t *BH,
                    int64_t *HN_out) {
  if ((dim != 2 && dim != 3) || Ne <= 0 || Nb < 0 || Nr <= 0 || !H || !HN_out || (Nb > 0 && !BH))
    return VMC_ERR_INVALID_ARGUMENT;
  if (Ne >= kMaxElements || Nb >= INT32_MAX) return VMC_ERR_INVALID_ARGUMENT;
  vmc_problem m;
  m.dim = dim;
  m.nv = dim + 1;
  m.nbv = dim;
  m.Ne = Ne;
  m.Nb = Nb;
  m.Nr = Nr;
  copy_in(m.H, H, Ne * m.nv);
  if (Nb > 0) copy_in(m.BH, BH, Nb * m.nbv);
  try {
    // node ids index the reference's vertex incidence lists (sized by the
    // node count), so they must lie in [0, Nr)
    for (int64_t v : m.H)
      if (v < 0 || v >= Nr) throw McError{INCONSISTENT_H};
    for (int64_t v : m.BH)
      if (v < 0 || v >= Nr) throw McError{INCONSISTENT_BH};
    build_neighbours(m);
  } catch (McError e) {
    return e.code + 1;
  }
  std::copy(m.HN.begin(), m.HN.end(), HN_out);
  return VMC_OK;
}


} // namespace vmc
