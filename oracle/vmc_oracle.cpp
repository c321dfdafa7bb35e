/*
 * oracle/vmc_oracle.cpp -- TEST INFRASTRUCTURE ONLY (see vmc_oracle.h).
 *
 * CPU restatement of ValoMC's photon-packet transport (reference
 * cpp/3d/MC3D.hpp and cpp/2d/MC2D.hpp).  It exists to check the HIP product
 * path and to serve as a timed CPU baseline; nothing under valomc_amd/ links
 * it.
 *
 * Two stream modes:
 *   ORACLE_RNG_MT19937 -- the reference's own streams: std::mt19937 seeded
 *     with seed+t for OpenMP thread t, N/T photons per thread with the
 *     remainder on the last thread, tallies reduced in thread order
 *     (MC3D.hpp:1648-1748).  Pinned bit-for-bit against the reference engine.
 *   ORACLE_RNG_PHILOX -- photon i draws from its own xoshiro128** stream
 *     started from a Philox4x32-10 block keyed by (seed, i)
 *     (PhiloxStream below).  This is the stream definition
 *     the HIP kernels use, so their FP64 path can be compared with this
 *     file photon-for-photon.
 *
 * Floating-point expressions that feed the tallies are written in the same
 * order as the reference so that the MT19937 mode reproduces its bits with
 * the same compiler flags (-O3, no -ffast-math, no FMA contraction).
 */
#include "vmc_oracle.h"

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kEps = std::numeric_limits<double>::epsilon(); /* MC3D.hpp:36 */
const int64_t kUnset = INT_FAST64_MAX;                       /* MC3D.hpp:29-31 */

/* mcerror codes, cpp/Errors.hpp:4-21 */
enum OErr {
  E_NO_LIGHTSOURCE, E_SIZE_MISMATCH_G, E_SIZE_MISMATCH_MUA, E_SIZE_MISMATCH_MUS,
  E_SIZE_MISMATCH_N, E_SIZE_MISMATCH_BCN, E_MISSING_BCN, E_SIZE_MISMATCH_BCTYPE,
  E_SIZE_MISMATCH_LIGHT_DIRECTION, E_SIZE_MISMATCH_LIGHT_DIRECTION_TYPE,
  E_INCONSISTENT_H, E_INCONSISTENT_BH, E_INCONSISTENT_MESH,
  E_INCONSISTENT_MESH_DUPLICATE_NEIGHBORS, E_NO_GAUSSIAN_SIGMA, E_MISSING_BOUNDARY,
  /* not a reference code: a source whose direction the reference leaves
     undefined; reported as status 17 */
  E_UNDEFINED_DIRECTION
};
struct OracleError { int code; };

/* ------------------------------------------------------------------------ */
/* Random streams                                                            */
/* ------------------------------------------------------------------------ */

/* glibc's sin/cos and sincos can disagree in the last bit, and g++ -O3
   fuses sin(x)/cos(x) into sincos(x) at some sites of the reference build
   and not at others: CreatePhoton (both dims) and the 2D ScatterPhoton fuse
   every pair; the 3D ScatterPhoton fuses sincos(phi) with separate
   sin/cos(theta) when |uz| > 0.999 and sincos(theta) with separate
   sin/cos(phi) otherwise; the frequency-domain tallies fuse none.  Calls
   go through these opaque pointers so each site makes the same libm call
   the reference makes. */
double (*volatile p_sin)(double) = ::sin;
double (*volatile p_cos)(double) = ::cos;
void (*volatile p_sincos)(double, double *, double *) = ::sincos;
inline double lsin(double x) { return p_sin(x); }
inline double lcos(double x) { return p_cos(x); }
inline void lsincos(double x, double *s, double *c) { p_sincos(x, s, c); }

/* uniform maps of cpp/3d/mt_rng.hpp:186-204 applied to a 32-bit draw */
inline double u_closed(uint32_t u) { return (double)u * (1.0 / 4294967295.0); }
inline double u_open(uint32_t u) { return ((double)u + 0.5) * (1.0 / 4294967296.0); }
inline double u_halfup(uint32_t u) { return (double)u * (1.0 / 4294967296.0); }

/* mt_rng (mt_rng.hpp:119-184) is MT19937 with the standard init_genrand
   seeding of the low 32 bits of the seed, i.e. std::mt19937. */
struct MtStream {
  std::mt19937 g;
  explicit MtStream(uint64_t s) : g((uint32_t)(s & 0xffffffffu)) {}
  void begin_photon(uint64_t) {}
  uint32_t next() { return (uint32_t)g(); }
};

void philox_block(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  /* Philox4x32 with 10 rounds: Salmon, Moraes, Dror, Shaw, "Parallel random
     numbers: as easy as 1, 2, 3", SC'11.  Multipliers and Weyl increments
     are the published ones. */
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int round = 0; round < 10; ++round) {
    if (round) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One stream per photon: xoshiro128** (Blackman and Vigna, "Scrambled
   linear pseudorandom number generators", 2018: output rotl(s1 * 5, 7) * 9,
   state update by xor/shift/rotate) started from the four words of the
   Philox4x32-10 block with counter (0x80000000, tag, photon lo, photon hi)
   and key (seed lo, seed hi); an all-zero start is replaced by s0 = 1.
   Creation and walk draw from it in the reference's order.  This is the
   HIP kernels' definition. */
struct PhiloxStream {
  uint32_t key[2];
  uint32_t tag;
  uint32_t xs[4];
  explicit PhiloxStream(uint64_t seed, uint32_t tag_ = 0) : tag(tag_) {
    key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32);
  }
  void begin_photon(uint64_t id) {
    uint32_t ctr[4] = {0x80000000u, tag, (uint32_t)id, (uint32_t)(id >> 32)};
    philox_block(ctr, key, xs);
    if ((xs[0] | xs[1] | xs[2] | xs[3]) == 0u) xs[0] = 1u;
  }
  static uint32_t rotl(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
  uint32_t next() {
    const uint32_t r = rotl(xs[1] * 5u, 7) * 9u;
    const uint32_t t = xs[1] << 9;
    xs[2] ^= xs[0];
    xs[3] ^= xs[1];
    xs[1] ^= xs[2];
    xs[0] ^= xs[3];
    xs[2] ^= t;
    xs[3] = rotl(xs[3], 11);
    return r;
  }
};

/* ------------------------------------------------------------------------ */
/* Problem model shared by all threads (read-only during transport)          */
/* ------------------------------------------------------------------------ */

struct Model {
  int dim;                   /* 2 or 3                                        */
  int nv;                    /* vertices per element: dim+1                   */
  int nbv;                   /* vertices per boundary element: dim            */
  int64_t Ne, Nb, Nr;
  std::vector<int64_t> H, HN, BH;
  std::vector<double> r, mua, mus, g, n, k, g2;
  std::vector<char> BCType, BCLDirType;
  std::vector<double> BCn, BCLNormal, BCIntensity, GaussianSigma;
  double f = 0, omega = 0, phase0 = 0;
  double c0 = 2.99792458e11;           /* MC3D.hpp:178, the caller's member */
  double weight0 = 0.001, chance = 0.1; /* MC3D.hpp:182-183 */
  std::vector<int64_t> lsFace, lsMother;
  std::vector<double> lsCDF;

  int64_t h(int64_t el, int c) const { return H[el + Ne * c]; }
  int64_t hn(int64_t el, int c) const { return HN[el + Ne * c]; }
  int64_t &hn_ref(int64_t el, int c) { return HN[el + Ne * c]; }
  int64_t bh(int64_t ib, int c) const { return BH[ib + Nb * c]; }
  double x(int64_t node, int d) const { return r[node + Nr * d]; }
  double bcl(int64_t ib, int d) const { return BCLNormal[ib + Nb * d]; }
  double &bcl_ref(int64_t ib, int d) { return BCLNormal[ib + Nb * d]; }
};

void load_model(Model &m, const oracle_problem *p, int dim) {
  m.dim = dim; m.nv = dim + 1; m.nbv = dim;
  m.Ne = p->Ne; m.Nb = p->Nb; m.Nr = p->Nr;
  m.H.assign(p->H, p->H + p->Ne * m.nv);
  if (p->HN) m.HN.assign(p->HN, p->HN + p->Ne * m.nv);
  m.BH.assign(p->BH, p->BH + p->Nb * m.nbv);
  m.r.assign(p->r, p->r + p->Nr * dim);
  m.mua.assign(p->mua, p->mua + p->Ne);
  m.mus.assign(p->mus, p->mus + p->Ne);
  m.g.assign(p->g, p->g + p->Ne);
  m.n.assign(p->n, p->n + p->Ne);
  m.BCType.assign(p->BCType, p->BCType + p->Nb);
  if (p->BCn) m.BCn.assign(p->BCn, p->BCn + p->Nb);
  if (p->BCLNormal) m.BCLNormal.assign(p->BCLNormal, p->BCLNormal + p->Nb * dim);
  if (p->BCLDirType) m.BCLDirType.assign(p->BCLDirType, p->BCLDirType + p->Nb);
  if (p->BCIntensity) m.BCIntensity.assign(p->BCIntensity, p->BCIntensity + p->Nb);
  if (p->GaussianSigma) m.GaussianSigma.assign(p->GaussianSigma, p->GaussianSigma + p->Nb);
  m.f = p->f;
  m.phase0 = p->phase0;
  m.weight0 = p->weight0;
  m.chance = p->chance;
  m.c0 = p->c0;
}

/* Input checks in the order of MC3D::ErrorChecks (MC3D.hpp:460-571) and
   MC2D::ErrorChecks (MC2D.hpp:390-513).  Array sizes are implied by the
   struct, so only the content checks can fire here. */
void error_checks(const Model &m) {
  if (m.BCn.empty()) throw OracleError{E_MISSING_BCN};
  bool has_ls = false, has_gauss = false;
  for (char c : m.BCType) {
    if (c != 'a') has_ls = true;
    if (c == 'g') has_gauss = true;
  }
  if (!has_ls) throw OracleError{E_NO_LIGHTSOURCE};
  if (m.dim == 2 && has_gauss && (int64_t)m.GaussianSigma.size() != m.Nb)
    throw OracleError{E_NO_GAUSSIAN_SIGMA};
  for (int64_t v : m.H)
    if (v < 0 || v >= m.Nr) throw OracleError{E_INCONSISTENT_H};
  for (int64_t v : m.BH)
    if (v < 0 || v >= m.Nr) throw OracleError{E_INCONSISTENT_BH};
}

/* ---- geometry, 3D (MC3D.hpp:306-457) ---------------------------------- */

const int kFace3[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}}; /* MC3D.hpp:1013 */
const int kEdge2[3][2] = {{0, 1}, {1, 2}, {2, 0}};                     /* MC2D.hpp:921 */

double volume3(const Model &m, int64_t el) {
  double ax = m.x(m.h(el, 0), 0), ay = m.x(m.h(el, 0), 1), az = m.x(m.h(el, 0), 2);
  double bx = m.x(m.h(el, 1), 0), by = m.x(m.h(el, 1), 1), bz = m.x(m.h(el, 1), 2);
  double cx = m.x(m.h(el, 2), 0), cy = m.x(m.h(el, 2), 1), cz = m.x(m.h(el, 2), 2);
  double dx = m.x(m.h(el, 3), 0), dy = m.x(m.h(el, 3), 1), dz = m.x(m.h(el, 3), 2);
  double v = (ax - dx) * ((by - dy) * (cz - dz) - (bz - dz) * (cy - dy)) -
             (ay - dy) * ((bx - dx) * (cz - dz) - (bz - dz) * (cx - dx)) +
             (az - dz) * ((bx - dx) * (cy - dy) - (by - dy) * (cx - dx));
  return fabs(v) / 6.0;
}

/* Heron's formula on the three node ids, MC3D.hpp:335-343 */
double heron3(const Model &m, int64_t i0, int64_t i1, int64_t i2) {
  double a = sqrt(pow(m.x(i1, 0) - m.x(i0, 0), 2) + pow(m.x(i1, 1) - m.x(i0, 1), 2) + pow(m.x(i1, 2) - m.x(i0, 2), 2));
  double b = sqrt(pow(m.x(i2, 0) - m.x(i0, 0), 2) + pow(m.x(i2, 1) - m.x(i0, 1), 2) + pow(m.x(i2, 2) - m.x(i0, 2), 2));
  double c = sqrt(pow(m.x(i1, 0) - m.x(i2, 0), 2) + pow(m.x(i1, 1) - m.x(i2, 1), 2) + pow(m.x(i1, 2) - m.x(i2, 2), 2));
  return sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c)) / 4.0;
}
double bnd_area3(const Model &m, int64_t ib) { return heron3(m, m.bh(ib, 0), m.bh(ib, 1), m.bh(ib, 2)); }

/* unit normal of the triangle (i0, i1, i2): (v1-v0) x (v2-v0), MC3D.hpp:386-457 */
void tri_normal(const Model &m, int64_t i0, int64_t i1, int64_t i2, double *nrm) {
  double x1 = m.x(i1, 0) - m.x(i0, 0), y1 = m.x(i1, 1) - m.x(i0, 1), z1 = m.x(i1, 2) - m.x(i0, 2);
  double x2 = m.x(i2, 0) - m.x(i0, 0), y2 = m.x(i2, 1) - m.x(i0, 1), z2 = m.x(i2, 2) - m.x(i0, 2);
  double nx = y1 * z2 - z1 * y2, ny = z1 * x2 - x1 * z2, nz = x1 * y2 - y1 * x2;
  double len = sqrt(nx * nx + ny * ny + nz * nz);
  nrm[0] = nx / len; nrm[1] = ny / len; nrm[2] = nz / len;
}
void bnd_normal3(const Model &m, int64_t ib, double *nrm) { tri_normal(m, m.bh(ib, 0), m.bh(ib, 1), m.bh(ib, 2), nrm); }
void face_normal3(const Model &m, int64_t el, int64_t f, double *nrm) {
  if (f < 0 || f > 3) { nrm[0] = nrm[1] = nrm[2] = 0.0; return; }
  tri_normal(m, m.h(el, kFace3[f][0]), m.h(el, kFace3[f][1]), m.h(el, kFace3[f][2]), nrm);
}

/* ---- geometry, 2D (MC2D.hpp:303-387) ---------------------------------- */

double area2(const Model &m, int64_t el) {
  double ax = m.x(m.h(el, 1), 0) - m.x(m.h(el, 0), 0);
  double ay = m.x(m.h(el, 1), 1) - m.x(m.h(el, 0), 1);
  double bx = m.x(m.h(el, 2), 0) - m.x(m.h(el, 0), 0);
  double by = m.x(m.h(el, 2), 1) - m.x(m.h(el, 0), 1);
  double c = ax * by - ay * bx;
  return 0.5 * (c >= 0.0 ? c : -c);
}
double seg_len2(const Model &m, int64_t i0, int64_t i1) {
  return sqrt(pow(m.x(i1, 0) - m.x(i0, 0), 2) + pow(m.x(i1, 1) - m.x(i0, 1), 2));
}
double bnd_len2(const Model &m, int64_t ib) { return seg_len2(m, m.bh(ib, 0), m.bh(ib, 1)); }
void seg_normal2(const Model &m, int64_t i0, int64_t i1, double *nrm) {
  double dx = m.x(i1, 0) - m.x(i0, 0), dy = m.x(i1, 1) - m.x(i0, 1);
  double len = sqrt(dx * dx + dy * dy);
  nrm[0] = dy / len; nrm[1] = -dx / len;
}
void bnd_normal2(const Model &m, int64_t ib, double *nrm) { seg_normal2(m, m.bh(ib, 0), m.bh(ib, 1), nrm); }
void edge_normal2(const Model &m, int64_t el, int64_t f, double *nrm) {
  if (f < 0 || f > 2) { nrm[0] = nrm[1] = 0.0; return; }
  seg_normal2(m, m.h(el, kEdge2[f][0]), m.h(el, kEdge2[f][1]), nrm);
}

/* ---- neighbour table ---------------------------------------------------- */

/* Vertex-incidence search of MC3D::BuildNeighbourhoods / search_neighbor
   (MC3D.hpp:740-871) and MC2D (MC2D.hpp:645-773): face f of element e gets
   the other element, or boundary entry -1-ib, that contains all of its
   vertices; a second, different match is an inconsistent mesh. */
void build_neighbours(Model &m) {
  if ((int64_t)m.HN.size() == m.Ne * m.nv) return;
  m.HN.assign(m.Ne * m.nv, kUnset);
  std::vector<std::vector<int64_t>> parent(m.Nr);
  for (int64_t e = 0; e < m.Ne; ++e)
    for (int c = 0; c < m.nv; ++c) parent[m.h(e, c)].push_back(e);
  for (int64_t b = 0; b < m.Nb; ++b)
    for (int c = 0; c < m.nbv; ++c) parent[m.bh(b, c)].push_back(-1 - b);
  const int dup_code = (m.dim == 3) ? E_INCONSISTENT_MESH_DUPLICATE_NEIGHBORS : E_INCONSISTENT_MESH;
  const int nf = m.nv;
  for (int64_t e = 0; e < m.Ne; ++e) {
    for (int c = 0; c < m.nv; ++c) {
      for (int64_t cand : parent[m.h(e, c)]) {
        if (cand == e) continue;
        for (int f = 0; f < nf; ++f) {
          bool all = true;
          for (int q = 0; q < m.dim && all; ++q) {
            int64_t v = m.h(e, m.dim == 3 ? kFace3[f][q] : kEdge2[f][q]);
            bool found = false;
            if (cand >= 0) {
              for (int w = 0; w < m.nv; ++w) found |= (m.h(cand, w) == v);
            } else {
              for (int w = 0; w < m.nbv; ++w) found |= (m.bh(-1 - cand, w) == v);
            }
            all = found;
          }
          if (!all) continue;
          int64_t &slot = m.hn_ref(e, f);
          if (slot != kUnset && slot != cand) throw OracleError{dup_code};
          slot = cand;
        }
      }
    }
    if (m.dim == 2)
      for (int f = 0; f < 3; ++f)
        if (m.hn(e, f) == kUnset) throw OracleError{E_MISSING_BOUNDARY}; /* MC2D.hpp:768-771 */
  }
}

bool is_source_type(const Model &m, char t) {
  if (m.dim == 3)
    return t == 'l' || t == 'L' || t == 'i' || t == 'I' || t == 'c' || t == 'C' || t == 'p'; /* MC3D.hpp:957-959 */
  return t == 'l' || t == 'i' || t == 'c' || t == 'g';                                     /* MC2D.hpp:865-868 */
}

/* Light sources in element-major (el, face) order, CDF of area*intensity
   (MC3D.hpp:941-1002, MC2D.hpp:851-911). */
void build_light_sources(Model &m) {
  for (int64_t e = 0; e < m.Ne; ++e)
    for (int f = 0; f < m.nv; ++f) {
      int64_t v = m.hn(e, f);
      if (v <= -1) {
        int64_t ib = -1 - v;
        if (is_source_type(m, m.BCType[ib])) {
          m.lsFace.push_back(ib);
          m.lsMother.push_back(e);
          m.lsCDF.push_back(m.dim == 3 ? bnd_area3(m, ib) : bnd_len2(m, ib));
        }
      }
    }
  size_t ns = m.lsCDF.size();
  /* source-type faces that no element references (an inconsistent mesh)
     leave the list empty; the reference would divide by an empty CDF */
  if (ns == 0) throw OracleError{E_NO_LIGHTSOURCE};
  for (size_t i = 0; i < ns; ++i) m.lsCDF[i] *= m.BCIntensity[m.lsFace[i]];
  for (size_t i = 1; i < ns; ++i) m.lsCDF[i] += m.lsCDF[i - 1];
  for (size_t i = 0; i < ns; ++i) m.lsCDF[i] /= m.lsCDF[ns - 1];
}

/* MC3D::Init (MC3D.hpp:576-738) / MC2D::Init (MC2D.hpp:518-643) */
void init_model(Model &m) {
  m.omega = 2.0 * M_PI * m.f;
  build_neighbours(m);
  if (m.BCIntensity.empty()) m.BCIntensity.assign(m.Nb, 1.0);
  if (m.BCLDirType.empty()) m.BCLDirType.assign(m.Nb, 'n');
  build_light_sources(m);
  m.k.resize(m.Ne);
  for (int64_t i = 0; i < m.Ne; ++i) m.k[i] = m.omega / m.c0 * m.n[i];
  m.g2.resize(m.Ne);
  for (int64_t i = 0; i < m.Ne; ++i) m.g2[i] = pow(m.g[i], 2);

  if (m.dim == 3) {
    if (!m.BCLNormal.empty()) {
      /* only rows of direction type 'n' are normalised (MC3D.hpp:671-688) */
      for (int64_t i = 0; i < m.Nb; ++i) {
        double len = sqrt(pow(m.bcl(i, 0), 2) + pow(m.bcl(i, 1), 2) + pow(m.bcl(i, 2), 2));
        if (m.BCLDirType[i] != 'n') continue;
        m.bcl_ref(i, 0) /= len; m.bcl_ref(i, 1) /= len; m.bcl_ref(i, 2) /= len;
      }
    }
    for (int64_t i = 0; i < m.Nb; ++i) {
      if (m.BCLDirType[i] != 'r') continue;
      /* surface frame (e1, e2, n) by Gram-Schmidt, MC3D.hpp:696-733 */
      double nn[3], e1[3], e2[3], len, dpn, dp1;
      bnd_normal3(m, i, nn);
      for (int d = 0; d < 3; ++d) {
        e1[d] = m.x(m.bh(i, 1), d) - m.x(m.bh(i, 0), d);
        e2[d] = m.x(m.bh(i, 2), d) - m.x(m.bh(i, 0), d);
      }
      dpn = nn[0] * e1[0] + nn[1] * e1[1] + nn[2] * e1[2];
      e1[0] -= dpn * nn[0]; e1[1] -= dpn * nn[1]; e1[2] -= dpn * nn[2];
      len = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
      e1[0] /= len; e1[1] /= len; e1[2] /= len;
      dpn = nn[0] * e2[0] + nn[1] * e2[1] + nn[2] * e2[2];
      dp1 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
      e2[0] -= dpn * nn[0] + dp1 * e1[0];
      e2[1] -= dpn * nn[1] + dp1 * e1[1];
      e2[2] -= dpn * nn[2] + dp1 * e1[2];
      len = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
      e2[0] /= len; e2[1] /= len; e2[2] /= len;
      double ox = m.bcl(i, 0), oy = m.bcl(i, 1), oz = m.bcl(i, 2);
      for (int d = 0; d < 3; ++d) m.bcl_ref(i, d) = ox * e1[d] + oy * e2[d] + oz * nn[d];
    }
  } else {
    if (!m.BCLNormal.empty()) {
      /* every row is normalised in 2D (MC2D.hpp:612-622) */
      for (int64_t i = 0; i < m.Nb; ++i) {
        double len = sqrt(pow(m.bcl(i, 0), 2) + pow(m.bcl(i, 1), 2));
        m.bcl_ref(i, 0) /= len; m.bcl_ref(i, 1) /= len;
      }
    }
    for (int64_t i = 0; i < m.Nb; ++i) {
      if (m.BCLDirType[i] != 'r') continue;
      double nn[2], t[2], len;
      bnd_normal2(m, i, nn);
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

/* ------------------------------------------------------------------------ */
/* Photon transport                                                          */
/* ------------------------------------------------------------------------ */

struct Packet {              /* MC3D.hpp:42-48 */
  double pos[3], dir[3];
  int64_t curel, nextel, curface, nextface;
  double weight, phase;
};

/* Moeller-Trumbore ray/triangle test, MC3D.hpp:1825-1869 */
int ray_tri(const double O[3], const double D[3], const double V0[3], const double V1[3],
            const double V2[3], double *t) {
  double e1[3] = {V1[0] - V0[0], V1[1] - V0[1], V1[2] - V0[2]};
  double e2[3] = {V2[0] - V0[0], V2[1] - V0[1], V2[2] - V0[2]};
  double pv[3] = {D[1] * e2[2] - D[2] * e2[1], D[2] * e2[0] - D[0] * e2[2], D[0] * e2[1] - D[1] * e2[0]};
  double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
  if ((-kEps < det) && (det < kEps)) return 0;
  double inv = 1.0 / det;
  double tv[3] = {O[0] - V0[0], O[1] - V0[1], O[2] - V0[2]};
  double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
  if ((u < 0.0) || (u > 1.0)) return 0;
  double qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
  double v = (D[0] * qv[0] + D[1] * qv[1] + D[2] * qv[2]) * inv;
  if ((v < 0.0) || (u + v > 1.0)) return 0;
  *t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
  return 1;
}

/* Ray/segment test, MC2D.hpp:1710-1722 */
int ray_seg(const double R[2], const double D[2], const double U[2], const double V[2], double *t) {
  double det = D[0] * (U[1] - V[1]) - D[1] * (U[0] - V[0]);
  if ((-kEps < det) && (det < kEps)) return 0;
  double s = (D[1] * (R[0] - U[0]) - D[0] * (R[1] - U[1])) / det;
  if ((s < 0.0) || (s > 1.0)) return 0;
  *t = ((R[1] - U[1]) * (U[0] - V[0]) - (R[0] - U[0]) * (U[1] - V[1])) / det;
  return 1;
}

template <class Rng>
struct Walker {
  const Model &m;
  Rng rng;
  std::vector<double> ER, EI, EBR, EBI, DEBR, DEBI; /* raw per-thread tallies */
  int64_t loss = 0;
  double led[5] = {0, 0, 0, 0, 0};

  Walker(const Model &model, Rng r) : m(model), rng(r) {
    ER.assign(m.Ne, 0.0); EI.assign(m.Ne, 0.0);
    EBR.assign(m.Nb, 0.0); EBI.assign(m.Nb, 0.0);
    DEBR.assign(m.Nb, 0.0); DEBI.assign(m.Nb, 0.0);
  }
  double closed() { return u_closed(rng.next()); }
  double open() { return u_open(rng.next()); }
  double halfup() { return u_halfup(rng.next()); }

  /* pick the source entry: first CDF value above xi; clamped as in
     MC2D.hpp:969 (MC3D.hpp:1068-1070 can run one past the end when xi==1) */
  int64_t pick_source() {
    double xi = closed();
    int64_t ns = (int64_t)m.lsCDF.size(), i;
    for (i = 0; i < ns; ++i)
      if (xi < m.lsCDF[i]) break;
    return i > ns - 1 ? ns - 1 : i;
  }

  /* ---------------- 3D ---------------- */

  /* MC3D::WhichFace, MC3D.hpp:1006-1057 */
  int which_face3(Packet &ph, double *dist) {
    double V[4][3];
    for (int c = 0; c < 4; ++c)
      for (int d = 0; d < 3; ++d) V[c][d] = m.x(m.h(ph.curel, c), d);
    for (int f = 0; f < 4; ++f) {
      if (ph.curface == f) continue;
      if (ray_tri(ph.pos, ph.dir, V[kFace3[f][0]], V[kFace3[f][1]], V[kFace3[f][2]], dist) && *dist > 0.0) {
        ph.nextface = f;
        ph.nextel = m.hn(ph.curel, f);
        return f;
      }
    }
    return -1;
  }

  /* MC3D::CreatePhoton, MC3D.hpp:1060-1294 */
  void create3(Packet &ph) {
    int64_t is = pick_source();
    ph.curel = m.lsMother[is];
    int64_t ib = m.lsFace[is];
    ph.curface = -1;
    for (int f = 0; f < 4; ++f)
      if (-1 - m.hn(ph.curel, f) == ib) { ph.curface = f; break; }

    double r1 = open(), r2 = open();
    for (int d = 0; d < 3; ++d)
      ph.pos[d] = (1.0 - sqrt(r1)) * m.x(m.bh(ib, 0), d) + sqrt(r1) * (1.0 - r2) * m.x(m.bh(ib, 1), d) +
                  sqrt(r1) * r2 * m.x(m.bh(ib, 2), d);
    double nn[3], t[3], len;
    bnd_normal3(m, ib, nn);
    for (int d = 0; d < 3; ++d)
      t[d] = (m.x(m.h(ph.curel, 0), d) + m.x(m.h(ph.curel, 1), d) + m.x(m.h(ph.curel, 2), d) +
              m.x(m.h(ph.curel, 3), d)) / 4.0 - ph.pos[d];
    len = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    t[0] /= len; t[1] /= len; t[2] /= len;
    if (nn[0] * t[0] + nn[1] * t[1] + nn[2] * t[2] < 0.0) { nn[0] = -nn[0]; nn[1] = -nn[1]; nn[2] = -nn[2]; }

    const char bt = m.BCType[ib];
    const bool given = m.BCLDirType[ib] != 'n';
    if (!given ? bt == 'l' : (bt == 'l' || bt == 'L')) {
      for (int d = 0; d < 3; ++d) ph.dir[d] = given ? m.bcl(ib, d) : nn[d];
    } else if (given && bt == 'p') {
      /* pencil: inward normal, position from BCLNormal (MC3D.hpp:1204-1218) */
      for (int d = 0; d < 3; ++d) { ph.dir[d] = nn[d]; ph.pos[d] = m.bcl(ib, d); }
    } else if (!given ? bt == 'i' : (bt == 'i' || bt == 'I')) {
      /* uniform on the inward hemisphere by rejection (MC3D.hpp:1124-1141) */
      double dot, rr[3], th, u;
      do {
        th = 2.0 * M_PI * halfup();
        u = 2.0 * closed() - 1.0;
        double st, ct;
        lsincos(th, &st, &ct);
        rr[0] = sqrt(1.0 - pow(u, 2)) * ct;
        rr[1] = sqrt(1.0 - pow(u, 2)) * st;
        rr[2] = u;
        dot = nn[0] * rr[0] + nn[1] * rr[1] + nn[2] * rr[2];
      } while (dot <= 0.0);
      for (int d = 0; d < 3; ++d) ph.dir[d] = rr[d];
    } else if (bt == 'c' || bt == 'C') {
      /* "cosinic": phi = asin(2U-1) about the first frame axis
         (MC3D.hpp:1142-1191, 1237-1286) */
      double e1[3], e2[3], fv[3], ph_, th, len2, dpn, dp1;
      for (int d = 0; d < 3; ++d) {
        e1[d] = m.x(m.bh(ib, 1), d) - m.x(m.bh(ib, 0), d);
        e2[d] = m.x(m.bh(ib, 2), d) - m.x(m.bh(ib, 0), d);
      }
      ph_ = asin(2.0 * open() - 1.0);
      th = 2.0 * M_PI * closed();
      double sp, cp, st, ct;
      lsincos(ph_, &sp, &cp);
      lsincos(th, &st, &ct);
      fv[0] = cp; fv[1] = ct * sp; fv[2] = st * sp;
      if (!given) {
        len2 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        e1[0] /= len2; e1[1] /= len2; e1[2] /= len2;
        dpn = nn[0] * e2[0] + nn[1] * e2[1] + nn[2] * e2[2];
        dp1 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
        e2[0] -= dpn * nn[0] + dp1 * e1[0];
        e2[1] -= dpn * nn[1] + dp1 * e1[1];
        e2[2] -= dpn * nn[2] + dp1 * e1[2];
      } else {
        double b0 = m.bcl(ib, 0), b1 = m.bcl(ib, 1), b2 = m.bcl(ib, 2);
        dpn = b0 * e1[0] + b1 * e1[1] + b2 * e1[2];
        e1[0] -= dpn * b0; e1[1] -= dpn * b1; e1[2] -= dpn * b2;
        len2 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        e1[0] /= len2; e1[1] /= len2; e1[2] /= len2;
        dpn = b0 * e2[0] + b1 * e2[1] + b2 * e2[2];
        dp1 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
        e2[0] -= dpn * b0 + dp1 * e1[0];
        e2[1] -= dpn * b1 + dp1 * e1[1];
        e2[2] -= dpn * b2 + dp1 * e1[2];
      }
      len2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
      e2[0] /= len2; e2[1] /= len2; e2[2] /= len2;
      /* the first axis is the inward normal in both branches (MC3D.hpp:1283-1285) */
      for (int d = 0; d < 3; ++d) ph.dir[d] = fv[0] * nn[d] + fv[1] * e1[d] + fv[2] * e2[d];
    }
    /* other combinations leave dir undefined in the reference; the
       harness rejects them before transport */
    ph.nextel = -1; ph.nextface = -1;
    ph.weight = 1.0;
    ph.phase = m.phase0;
  }

  /* MC3D::ScatterPhoton, MC3D.hpp:1297-1340 */
  void scatter3(Packet &ph) {
    double xi, theta, phi, dxn, dyn, dzn;
    const int64_t el = ph.curel;
    if (m.g[el] != 0.0) {
      xi = closed();
      if ((0.0 < xi) && (xi < 1.0))
        theta = acos((1.0 + m.g2[el] - pow((1.0 - m.g2[el]) / (1.0 - m.g[el] * (1.0 - 2.0 * xi)), 2)) / (2.0 * m.g[el]));
      else
        theta = (1.0 - xi) * M_PI;
    } else {
      theta = acos(2.0 * closed() - 1.0);
    }
    phi = 2.0 * M_PI * closed();
    double sth, cth, sph, cph;
    if (fabs(ph.dir[2]) > 0.999) {
      lsincos(phi, &sph, &cph);
      sth = lsin(theta);
      cth = lcos(theta);
      dxn = sth * cph;
      dyn = sth * sph;
      dzn = ph.dir[2] * cth / fabs(ph.dir[2]);
    } else {
      lsincos(theta, &sth, &cth);
      cph = lcos(phi);
      sph = lsin(phi);
      dxn = sth * (ph.dir[0] * ph.dir[2] * cph - ph.dir[1] * sph) / sqrt(1.0 - ph.dir[2] * ph.dir[2]) + ph.dir[0] * cth;
      dyn = sth * (ph.dir[1] * ph.dir[2] * cph + ph.dir[0] * sph) / sqrt(1.0 - ph.dir[2] * ph.dir[2]) + ph.dir[1] * cth;
      dzn = -sth * cph * sqrt(1.0 - ph.dir[2] * ph.dir[2]) + ph.dir[2] * cth;
    }
    double len = sqrt(dxn * dxn + dyn * dyn + dzn * dzn);
    ph.dir[0] = dxn / len; ph.dir[1] = dyn / len; ph.dir[2] = dzn / len;
    ph.curface = -1;
  }

  /* specular reflection about a unit normal (MC3D.hpp:1343-1362) */
  static void mirror(Packet &ph, const double *nrm, int dim) {
    double cd = 0;
    for (int d = 0; d < dim; ++d) cd += nrm[d] * ph.dir[d];
    /* same association as n0*d0 + n1*d1 (+ n2*d2) */
    for (int d = 0; d < dim; ++d) ph.dir[d] -= 2.0 * cd * nrm[d];
  }

  /* Reflection probability shared by MC3D::FresnelPhoton (MC3D.hpp:1397-1420)
     and MC2D::FresnelPhoton (MC2D.hpp:1210-1234).  Returns 1 on total
     internal reflection, otherwise sets costht and R. */
  static int fresnel_R(double nipnt, double costhi, double *costht, double *R) {
    if (1.0 - pow(nipnt, 2) * (1.0 - pow(costhi, 2)) <= 0.0) return 1;
    *costht = sqrt(1.0 - pow(nipnt, 2) * (1.0 - pow(costhi, 2)));
    double thi = costhi > 0.0 ? acos(costhi) : acos(-costhi);
    double tht = acos(*costht);
    if (!(sin(thi + tht) > kEps))
      *R = pow((nipnt - 1.0) / (nipnt + 1.0), 2);
    else
      *R = 0.5 * (pow(sin(thi - tht) / sin(thi + tht), 2) + pow(tan(thi - tht) / tan(thi + tht), 2));
    return 0;
  }

  /* MC3D::FresnelPhoton, MC3D.hpp:1365-1447 */
  int fresnel3(Packet &ph) {
    double nor[3];
    face_normal3(m, ph.curel, ph.nextface, nor);
    double nipnt = ph.nextel < 0 ? m.n[ph.curel] / m.BCn[-1 - ph.nextel] : m.n[ph.curel] / m.n[ph.nextel];
    double costhi = -(ph.dir[0] * nor[0] + ph.dir[1] * nor[1] + ph.dir[2] * nor[2]);
    double costht = 0, R = 0;
    bool reflect = fresnel_R(nipnt, costhi, &costht, &R) != 0;
    if (!reflect) reflect = closed() <= R;
    if (reflect) {
      for (int d = 0; d < 3; ++d) ph.dir[d] += 2.0 * costhi * nor[d];
      ph.curface = ph.nextface;
      return 1;
    }
    const double s = costhi > 0.0 ? (nipnt * costhi - costht) : (nipnt * costhi + costht);
    for (int d = 0; d < 3; ++d) ph.dir[d] = nipnt * ph.dir[d] + s * nor[d];
    return 0;
  }

  /* ---------------- 2D ---------------- */

  /* MC2D::WhichFace, MC2D.hpp:914-955 */
  int which_face2(Packet &ph, double *dist) {
    double V[3][2];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < 2; ++d) V[c][d] = m.x(m.h(ph.curel, c), d);
    for (int f = 0; f < 3; ++f) {
      if (ph.curface == f) continue;
      if (ray_seg(ph.pos, ph.dir, V[kEdge2[f][0]], V[kEdge2[f][1]], dist) && *dist > 0.0) {
        ph.nextface = f;
        ph.nextel = m.hn(ph.curel, f);
        return f;
      }
    }
    return -1;
  }

  /* 16-step bisection of the truncated-normal angle CDF (MC2D.hpp:1044-1056) */
  double gaussian_angle(int64_t ib) {
    double sg = m.GaussianSigma[ib];
    double hh = 0.5 - 0.5 * erf(M_PI / sg / 2.0 / sqrt(2.0));
    double xi = open();
    double a = -M_PI / 2, b = M_PI / 2, phi = 0, cdf;
    for (int it = 0; it < 16; ++it) {
      phi = (a + b) / 2.0;
      cdf = (0.5 + 0.5 * erf(phi / sg / sqrt(2.0)) - hh) / (1 - 2.0 * hh);
      if (cdf - xi < 0.0) a = phi; else b = phi;
    }
    return phi;
  }

  /* MC2D::CreatePhoton, MC2D.hpp:958-1132 */
  void create2(Packet &ph) {
    int64_t is = pick_source();
    ph.curel = m.lsMother[is];
    int64_t ib = m.lsFace[is];
    ph.curface = -1;
    for (int f = 0; f < 3; ++f)
      if (-1 - m.hn(ph.curel, f) == ib) { ph.curface = f; break; }
    double w = open();
    ph.pos[0] = m.x(m.bh(ib, 0), 0) + w * (m.x(m.bh(ib, 1), 0) - m.x(m.bh(ib, 0), 0));
    ph.pos[1] = m.x(m.bh(ib, 0), 1) + w * (m.x(m.bh(ib, 1), 1) - m.x(m.bh(ib, 0), 1));
    ph.pos[2] = 0.0;
    double nn[2], t[2], len;
    bnd_normal2(m, ib, nn);
    t[0] = (m.x(m.h(ph.curel, 0), 0) + m.x(m.h(ph.curel, 1), 0) + m.x(m.h(ph.curel, 2), 0)) / 3.0 - ph.pos[0];
    t[1] = (m.x(m.h(ph.curel, 0), 1) + m.x(m.h(ph.curel, 1), 1) + m.x(m.h(ph.curel, 2), 1)) / 3.0 - ph.pos[1];
    len = sqrt(t[0] * t[0] + t[1] * t[1]);
    t[0] /= len; t[1] /= len;
    if (nn[0] * t[0] + nn[1] * t[1] < 0.0) { nn[0] = -nn[0]; nn[1] = -nn[1]; }

    const char bt = m.BCType[ib];
    if (m.BCLDirType[ib] == 'n') {
      double phi0 = atan2(nn[1], nn[0]), phi;
      if (bt == 'l') {
        ph.dir[0] = nn[0]; ph.dir[1] = nn[1];
      } else if (bt == 'i') {
        phi = M_PI * (open() - 0.5);
        lsincos(phi0 + phi, &ph.dir[1], &ph.dir[0]);
      } else if (bt == 'c') {
        phi = asin(2.0 * open() - 1.0);
        lsincos(phi0 + phi, &ph.dir[1], &ph.dir[0]);
      } else if (bt == 'g') {
        phi = gaussian_angle(ib);
        lsincos(phi0 + phi, &ph.dir[1], &ph.dir[0]);
      }
    } else {
      double phi0 = atan2(m.bcl(ib, 1), m.bcl(ib, 0)), phi;
      if (bt == 'l') {
        ph.dir[0] = m.bcl(ib, 0); ph.dir[1] = m.bcl(ib, 1);
      } else if (bt == 'i') {
        phi = M_PI * (open() - 0.5);
        lsincos(phi0 + phi, &ph.dir[1], &ph.dir[0]);
      } else if (bt == 'c') {
        do {
          phi = asin(2.0 * open() - 1.0);
          lsincos(phi0 + phi, &ph.dir[1], &ph.dir[0]);
        } while (nn[0] * ph.dir[0] + nn[1] * ph.dir[1] <= 0.0);
      } else if (bt == 'g') {
        do {
          phi = gaussian_angle(ib);
          lsincos(phi0 + phi, &ph.dir[1], &ph.dir[0]);
        } while (nn[0] * ph.dir[0] + nn[1] * ph.dir[1] <= 0.0);
      }
    }
    ph.dir[2] = 0.0;
    ph.nextel = -1; ph.nextface = -1;
    ph.weight = 1.0;
    ph.phase = m.phase0;
  }

  /* MC2D::ScatterPhoton, MC2D.hpp:1135-1155 */
  void scatter2(Packet &ph) {
    const int64_t el = ph.curel;
    if (m.g[el] != 0.0) {
      double xi = halfup();
      double tt = sqrt(pow(m.g2[el] + 1.0, 2) - 4.0 * m.g2[el]);
      double th = 2.0 * atan(tan(-xi * tt * M_PI / (m.g2[el] - 1.0)) * (m.g2[el] - 2 * m.g[el] + 1.0) / tt);
      double th0 = atan2(ph.dir[1], ph.dir[0]);
      lsincos(th0 + th, &ph.dir[1], &ph.dir[0]);
    } else {
      double th = 2.0 * M_PI * halfup();
      lsincos(th, &ph.dir[1], &ph.dir[0]);
    }
    ph.curface = -1;
  }

  /* MC2D::FresnelPhoton, MC2D.hpp:1178-1304: no refraction of dir when
     the photon leaves through the outer boundary */
  int fresnel2(Packet &ph) {
    double nor[3];
    edge_normal2(m, ph.curel, ph.nextface, nor);
    const bool outer = ph.nextel < 0;
    double nipnt = outer ? m.n[ph.curel] / m.BCn[-1 - ph.nextel] : m.n[ph.curel] / m.n[ph.nextel];
    double costhi = -(ph.dir[0] * nor[0] + ph.dir[1] * nor[1]);
    double costht = 0, R = 0;
    bool reflect = fresnel_R(nipnt, costhi, &costht, &R) != 0;
    if (!reflect) reflect = closed() <= R;
    if (reflect) {
      ph.dir[0] += 2.0 * costhi * nor[0];
      ph.dir[1] += 2.0 * costhi * nor[1];
      ph.curface = ph.nextface;
      return 1;
    }
    if (outer) return 0;
    const double s = costhi > 0.0 ? (nipnt * costhi - costht) : (nipnt * costhi + costht);
    ph.dir[0] = nipnt * ph.dir[0] + s * nor[0];
    ph.dir[1] = nipnt * ph.dir[1] + s * nor[1];
    return 0;
  }

  /* ---------------- the walk (MC3D.hpp:1450-1638, MC2D.hpp:1307-1511) ---------------- */

  bool is_mirror(char t) const { return t == 'm' || t == 'L' || t == 'I' || t == 'C'; }

  void propagate(Packet &ph) {
    const bool d3 = m.dim == 3;
    double prop, dist = 0, ds;
    for (;;) {
      prop = -log(open()) / m.mus[ph.curel];
      for (;;) {
        int wf = d3 ? which_face3(ph, &dist) : which_face2(ph, &dist);
        if (wf == -1) { ++loss; led[4] += ph.weight; return; }
        ds = fmin(prop, dist);
        ph.pos[0] += ph.dir[0] * ds;
        ph.pos[1] += ph.dir[1] * ds;
        if (d3) ph.pos[2] += ph.dir[2] * ds;
        const int64_t el = ph.curel;
        const double att = exp(-m.mua[el] * ds);
        if (m.omega <= 0.0) {
          if (m.mua[el] > 0.0)
            ER[el] += (1.0 - att) * ph.weight;
          else
            ER[el] += ph.weight * ds;
        } else {
          ER[el] += ph.weight * (lcos(ph.phase) - lcos(-ph.phase - m.k[el] * ds) * att);
          EI[el] += ph.weight * (-lsin(ph.phase) + lsin(ph.phase + m.k[el] * ds) * att);
          ph.phase += m.k[el] * ds;
        }
        led[0] += ph.weight - ph.weight * att;
        ph.weight *= att;
        prop -= ds;
        if (prop <= 0.0) break;

        if (ph.nextel < 0) {
          const int64_t ib = -1 - ph.nextel;
          if (is_mirror(m.BCType[ib])) {
            double nrm[2 + 1];
            if (d3) bnd_normal3(m, ib, nrm); else bnd_normal2(m, ib, nrm);
            mirror(ph, nrm, m.dim);
            ph.curface = ph.nextface;
            continue;
          }
          if (m.BCn[ib] > 0.0)
            if (d3 ? fresnel3(ph) : fresnel2(ph)) continue;
          if (m.omega <= 0.0) {
            EBR[ib] += ph.weight;
          } else {
            EBR[ib] += ph.weight * lcos(ph.phase);
            EBI[ib] -= ph.weight * lsin(ph.phase);
          }
          if (!d3) {
            /* exitance weighted by 1/(dir . outward normal), MC2D.hpp:1410-1437 */
            double nn[3], t[2];
            edge_normal2(m, ph.curel, ph.nextface, nn);
            t[0] = (m.x(m.h(ph.curel, 0), 0) + m.x(m.h(ph.curel, 1), 0) + m.x(m.h(ph.curel, 2), 0)) / 3.0 - ph.pos[0];
            t[1] = (m.x(m.h(ph.curel, 0), 1) + m.x(m.h(ph.curel, 1), 1) + m.x(m.h(ph.curel, 2), 1)) / 3.0 - ph.pos[1];
            if (nn[0] * t[0] + nn[1] * t[1] > 0.0) { nn[0] = -nn[0]; nn[1] = -nn[1]; }
            if (m.omega <= 0.0) {
              DEBR[ib] += ph.weight / (ph.dir[0] * nn[0] + ph.dir[1] * nn[1]);
            } else {
              DEBR[ib] += ph.weight * lcos(ph.phase) / (ph.dir[0] * nn[0] + ph.dir[1] * nn[1]);
              DEBI[ib] -= ph.weight * lsin(ph.phase) / (ph.dir[0] * nn[0] + ph.dir[1] * nn[1]);
            }
          }
          led[1] += ph.weight;
          return;
        }

        const int64_t nx = ph.nextel;
        /* a face the neighbour search left unmatched (INT_FAST64_MAX in 3D);
           the reference would index past the arrays here */
        if (nx >= m.Ne) { ++loss; led[4] += ph.weight; return; }
        if ((m.mus[ph.curel] <= 0.0) && (m.mus[nx] > 0.0)) {
          /* vacuum -> scattering medium: fresh free path, MC3D.hpp:1573-1592 */
          prop = -log(open()) / m.mus[nx];
          if (ph.weight < m.weight0) {
            if (closed() > m.chance) { led[2] += ph.weight; return; }
            led[3] += ph.weight / m.chance - ph.weight;
            ph.weight /= m.chance;
          }
          if (m.n[ph.curel] != m.n[nx])
            if (d3 ? fresnel3(ph) : fresnel2(ph)) continue;
        } else {
          if (ph.weight < m.weight0) {
            if (closed() > m.chance) { led[2] += ph.weight; return; }
            led[3] += ph.weight / m.chance - ph.weight;
            ph.weight /= m.chance;
          }
          if (m.n[ph.curel] != m.n[nx])
            if (d3 ? fresnel3(ph) : fresnel2(ph)) continue;
          prop *= m.mus[ph.curel] / m.mus[nx];
        }
        /* entry face in the next element, MC3D.hpp:1616-1628 */
        int64_t entry = -1;
        for (int f = 0; f < m.nv; ++f)
          if (m.hn(nx, f) == ph.curel) { entry = f; break; }
        if (entry < 0) { ++loss; led[4] += ph.weight; return; }
        ph.curface = entry;
        ph.curel = nx;
      }
      if (m.mus[ph.curel] > 0.0) {
        if (d3) scatter3(ph); else scatter2(ph);
      }
    }
  }

  void run_photon(uint64_t pid) {
    rng.begin_photon(pid);
    Packet ph;
    if (m.dim == 3) create3(ph); else create2(ph);
    propagate(ph);
  }
};

/* directions the reference leaves undefined (MC3D.hpp:1112-1192 has no
   branch for 'L', 'I' or 'p' when the direction type is 'n') */
void check_defined_directions(const Model &m) {
  if (m.dim != 3) return;
  for (size_t i = 0; i < m.lsFace.size(); ++i) {
    int64_t ib = m.lsFace[i];
    char t = m.BCType[ib];
    if (m.BCLDirType[ib] == 'n' && (t == 'L' || t == 'I' || t == 'p'))
      throw OracleError{E_UNDEFINED_DIRECTION};
  }
}

/* normalisation of MC3D::MonteCarlo (MC3D.hpp:1793-1818) and
   MC2D::MonteCarlo (MC2D.hpp:1673-1703) */
void normalise(const Model &m, int64_t N, std::vector<double> &ER, std::vector<double> &EI,
               std::vector<double> &EBR, std::vector<double> &EBI, std::vector<double> &DEBR,
               std::vector<double> &DEBI) {
  const bool d3 = m.dim == 3;
  if (m.omega <= 0.0) {
    for (int64_t i = 0; i < m.Ne; ++i) {
      double v = d3 ? volume3(m, i) : area2(m, i);
      if (m.mua[i] > 0.0) ER[i] /= m.mua[i] * v * (double)N;
      else ER[i] /= v * (double)N;
    }
    for (int64_t i = 0; i < m.Nb; ++i) EBR[i] /= (double)N * (d3 ? bnd_area3(m, i) : bnd_len2(m, i));
    if (!d3)
      for (int64_t i = 0; i < m.Nb; ++i) DEBR[i] /= (double)N * bnd_len2(m, i);
  } else {
    for (int64_t i = 0; i < m.Ne; ++i) {
      double a = ER[i], b = EI[i];
      double v = d3 ? volume3(m, i) : area2(m, i);
      ER[i] = (b * m.k[i] + a * m.mua[i]) / (pow(m.k[i], 2) + pow(m.mua[i], 2)) / (double)N / v;
      EI[i] = -(a * m.k[i] - b * m.mua[i]) / (pow(m.k[i], 2) + pow(m.mua[i], 2)) / (double)N / v;
    }
    for (int64_t i = 0; i < m.Nb; ++i) {
      double ar = d3 ? bnd_area3(m, i) : bnd_len2(m, i);
      EBR[i] /= (double)N * ar;
      EBI[i] /= (double)N * ar;
      if (!d3) {
        DEBR[i] /= (double)N * ar;
        DEBI[i] /= (double)N * ar;
      }
    }
  }
}

int run(const oracle_problem *p, int nthreads, int rng_mode, oracle_result *out, int dim) {
  Model m;
  try {
    load_model(m, p, dim);
    error_checks(m);
    init_model(m);
    check_defined_directions(m);
  } catch (OracleError e) {
    return 1 + e.code;
  }
  if (m.phase0 < 0) m.phase0 += 2 * M_PI * ceil(-m.phase0 / (2 * M_PI));
  if (nthreads < 1) nthreads = 1;
  const int64_t N = p->Nphoton;
  std::vector<int64_t> count(nthreads, N / nthreads), first(nthreads, 0);
  int64_t acc = 0;
  for (int t = 0; t < nthreads; ++t) acc += count[t];
  count[nthreads - 1] += N - acc;
  for (int t = 1; t < nthreads; ++t) first[t] = first[t - 1] + count[t - 1];

  std::vector<std::vector<double>> tER(nthreads), tEI(nthreads), tEBR(nthreads), tEBI(nthreads),
      tDEBR(nthreads), tDEBI(nthreads);
  std::vector<int64_t> tloss(nthreads, 0);
  std::vector<std::vector<double>> tled(nthreads, std::vector<double>(5, 0.0));

  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nthreads)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    auto finish = [&](auto &w) {
      tER[t].swap(w.ER); tEI[t].swap(w.EI); tEBR[t].swap(w.EBR); tEBI[t].swap(w.EBI);
      tDEBR[t].swap(w.DEBR); tDEBI[t].swap(w.DEBI);
      tloss[t] = w.loss;
      for (int q = 0; q < 5; ++q) tled[t][q] = w.led[q];
    };
    if (rng_mode == ORACLE_RNG_PHILOX) {
      Walker<PhiloxStream> w(m, PhiloxStream(p->seed, p->stream_tag));
      for (int64_t i = 0; i < count[t]; ++i) w.run_photon(p->first_photon + (uint64_t)(first[t] + i));
      finish(w);
    } else {
      /* seed*totalthreads + t with totalthreads == 1 (MC3D.hpp:1658) */
      Walker<MtStream> w(m, MtStream((uint64_t)(p->seed + (uint64_t)t)));
      for (int64_t i = 0; i < count[t]; ++i) w.run_photon(0);
      finish(w);
    }
  }
  auto t1 = std::chrono::steady_clock::now();

  /* ordered reduction, MC3D.hpp:1714-1748 */
  std::vector<double> ER(m.Ne, 0.0), EI(m.Ne, 0.0), EBR(m.Nb, 0.0), EBI(m.Nb, 0.0), DEBR(m.Nb, 0.0),
      DEBI(m.Nb, 0.0);
  int64_t loss = 0;
  double led[5] = {0, 0, 0, 0, 0};
  for (int t = 0; t < nthreads; ++t) {
    loss += tloss[t];
    for (int q = 0; q < 5; ++q) led[q] += tled[t][q];
  }
  for (int64_t i = 0; i < m.Ne; ++i)
    for (int t = 0; t < nthreads; ++t) { ER[i] += tER[t][i]; EI[i] += tEI[t][i]; }
  for (int64_t i = 0; i < m.Nb; ++i)
    for (int t = 0; t < nthreads; ++t) { EBR[i] += tEBR[t][i]; EBI[i] += tEBI[t][i]; }
  for (int64_t i = 0; i < m.Nb; ++i)
    for (int t = 0; t < nthreads; ++t) { DEBR[i] += tDEBR[t][i]; DEBI[i] += tDEBI[t][i]; }

  normalise(m, N, ER, EI, EBR, EBI, DEBR, DEBI);

  if (out->ER) std::memcpy(out->ER, ER.data(), sizeof(double) * m.Ne);
  if (out->EI) std::memcpy(out->EI, EI.data(), sizeof(double) * m.Ne);
  if (out->EBR) std::memcpy(out->EBR, EBR.data(), sizeof(double) * m.Nb);
  if (out->EBI) std::memcpy(out->EBI, EBI.data(), sizeof(double) * m.Nb);
  if (out->DEBR) std::memcpy(out->DEBR, DEBR.data(), sizeof(double) * m.Nb);
  if (out->DEBI) std::memcpy(out->DEBI, DEBI.data(), sizeof(double) * m.Nb);
  if (out->HN) std::memcpy(out->HN, m.HN.data(), sizeof(int64_t) * m.Ne * m.nv);
  out->loss = loss;
  for (int q = 0; q < 5; ++q) out->ledger[q] = led[q];
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  return 0;
}

} // namespace

extern "C" {

int oracle_run3d(const oracle_problem *p, int nthreads, int rng_mode, oracle_result *out) {
  return run(p, nthreads, rng_mode, out, 3);
}

int oracle_run2d(const oracle_problem *p, int nthreads, int rng_mode, oracle_result *out) {
  return run(p, nthreads, rng_mode, out, 2);
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox_block(ctr, key, out);
}

void oracle_xoshiro128ss(const uint32_t state[4], int n, uint32_t *out) {
  PhiloxStream st(0);
  for (int k = 0; k < 4; ++k) st.xs[k] = state[k];
  for (int i = 0; i < n; ++i) out[i] = st.next();
}

} // extern "C"
