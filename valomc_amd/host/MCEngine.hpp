// Drop-in engine classes over the C ABI.
//
// MC3D and MC2D keep the reference classes' public members and their
// meaning (cpp/3d/MC3D.hpp:107-172, cpp/2d/MC2D.hpp:112-180): fill the
// arrays, call ErrorChecks(), Init() and MonteCarlo(progress, finalchecks),
// read the normalised ER/EI/EBR/EBI/DEBR/DEBI and loss.  The photon-level
// methods of the reference (CreatePhoton, PropagatePhoton, WhichFace, ...)
// run on the GPU inside MonteCarlo() and have no host counterpart here.
// Two members are added: `precision` (VMC_FP64 by default, the reference's
// arithmetic, or VMC_FP32) and `tally` (VMC_TALLY_FP64 or
// VMC_TALLY_FIXED64), plus `stats` with the device counters of the last run.
#ifndef VMC_HOST_MCENGINE_HPP
#define VMC_HOST_MCENGINE_HPP

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "Array.hpp"
#include "Errors.hpp"
#include "valomc.h"

template <int DIM> class MCEngine {
public:
  MCEngine()
      : f(0.0), omega(0.0), phase0(0.0), Nphoton(1), c0(2.99792458e11), weight0(0.001), chance(0.1),
        loss(0), seed(5489UL), threadcount(1), rank(0), nodecount(1), totalthreads(1),
        precision(VMC_FP64), tally(VMC_TALLY_FP64) {}
  virtual ~MCEngine() {}

  // Size checks of ErrorChecks (MC3D.hpp:460-571, MC2D.hpp:390-513); the
  // content checks run in Init() through the C ABI, in the same order.
  void ErrorChecks() {
    if (g.Nx != H.Nx) throw SIZE_MISMATCH_G;
    if (mua.Nx != H.Nx) throw SIZE_MISMATCH_MUA;
    if (mus.Nx != H.Nx) throw SIZE_MISMATCH_MUS;
    if (n.Nx != H.Nx) throw SIZE_MISMATCH_N;
    if (!BCn.N) throw MISSING_BCN;
    if (BCn.Nx != BH.Nx) throw SIZE_MISMATCH_BCN;
    if (BCType.Nx != BH.Nx) throw SIZE_MISMATCH_BCTYPE;
    bool has_ls = false, has_gauss = false;
    for (int_fast64_t i = 0; i < BCType.N; ++i) {
      if (BCType[i] != 'a') has_ls = true;
      if (BCType[i] == 'g') has_gauss = true;
    }
    if (!has_ls) throw NO_LIGHTSOURCE;
    if (DIM == 2 && has_gauss && GaussianSigma.N != BH.Nx) throw NO_GAUSSIAN_SIGMA;
    if (BCLNormal.N && BCLNormal.Nx != BH.Nx) throw SIZE_MISMATCH_LIGHT_DIRECTION;
    if (BCLightDirectionType.N && BCLightDirectionType.Nx != BH.Nx) throw SIZE_MISMATCH_LIGHT_DIRECTION_TYPE;
    for (int_fast64_t i = 0; i < H.N; ++i)
      if (H[i] < 0 || H[i] >= r.Nx) throw INCONSISTENT_H;
    for (int_fast64_t i = 0; i < BH.N; ++i)
      if (BH[i] < 0 || BH[i] >= r.Nx) throw INCONSISTENT_BH;
  }

  // Init (MC3D.hpp:576-738): neighbour table, defaults, k and g^2, zeroed
  // outputs.  Light sources and direction frames are prepared by the
  // library from the same inputs.
  void Init() {
    omega = 2.0 * M_PI * f;
    vmc_problem_desc d = desc();
    std::vector<int64_t> hn((size_t)H.Nx * (DIM + 1));
    vmc_throw_status(vmc_problem_check(&d, hn.data(), nullptr));
    if (HN.N != H.N) {
      HN.resize(H.Nx, DIM + 1);
      for (size_t i = 0; i < hn.size(); ++i) HN[i] = hn[i];
    }
    ER.resize(H.Nx); EI.resize(H.Nx);
    EBR.resize(BH.Nx); EBI.resize(BH.Nx);
    DEBR.resize(BH.Nx); DEBI.resize(BH.Nx);
    if (!BCIntensity.N) {
      BCIntensity.resize(BCType.N);
      for (int_fast64_t i = 0; i < BCIntensity.N; ++i) BCIntensity[i] = 1.0;
    }
    if (!BCLightDirectionType.N) {
      BCLightDirectionType.resize(BCType.N);
      for (int_fast64_t i = 0; i < BCLightDirectionType.N; ++i) BCLightDirectionType[i] = 'n';
    }
    k.resize(n.N);
    g2.resize(g.N);
    for (int_fast64_t i = 0; i < n.N; ++i) k[i] = omega / c0 * n[i];
    for (int_fast64_t i = 0; i < g.N; ++i) g2[i] = pow(g[i], 2);
  }

  void BuildNeighbourhoods() {
    vmc_problem_desc d = desc();
    std::vector<int64_t> hn((size_t)H.Nx * (DIM + 1));
    vmc_throw_status(vmc_problem_check(&d, hn.data(), nullptr));
    if (HN.N != H.N) {
      HN.resize(H.Nx, DIM + 1);
      for (size_t i = 0; i < hn.size(); ++i) HN[i] = hn[i];
    }
  }

  // MonteCarlo (MC3D.hpp:1641-1822): transport on the GPU, normalised
  // outputs in ER/EI/EBR/EBI/DEBR/DEBI.  progress(percent) is called
  // between photon chunks; returning false stops the run, which is still
  // normalised by the planned Nphoton as in the reference.
  void MonteCarlo(bool (*progress)(double) = NULL, void (*finalchecks)(int, int) = NULL) {
    if (ER.N != H.Nx) Init();
    vmc_problem_desc d = desc();
    vmc_problem *prob = nullptr;
    vmc_throw_status(vmc_problem_create(&d, &prob));
    vmc_run_desc rd = {};
    rd.f = f;
    rd.phase0 = phase0;
    rd.Nphoton = Nphoton;
    rd.seed = (uint64_t)seed;
    rd.precision = precision;
    rd.tally = tally;
    rd.model = 1;   // the members, as the reference's roulette and k read them
    rd.weight0 = weight0;
    rd.chance = chance;
    rd.c0 = c0;
    vmc_output o = {ER.data, EI.data, EBR.data, EBI.data, DEBR.data, DEBI.data};
    Progress pg{progress, 0};
    const int st = vmc_problem_run(prob, &rd, &o, &stats, progress ? &Progress::call : nullptr, &pg);
    vmc_problem_destroy(prob);
    if (st != VMC_OK && st != VMC_ERR_ABORTED) vmc_throw_status(st);
    loss = (long)stats.loss;
    if (finalchecks) finalchecks(st == VMC_OK ? (int)Nphoton : (int)(pg.last * Nphoton / 100.0), (int)Nphoton);
    if (progress) progress(100);
  }

  double ElementVolume(int_fast64_t el);    // 3D: tet volume, 2D: triangle area
  double ElementArea(int_fast64_t ib);      // 3D: triangle area, 2D: edge length

public:
  Array<int_fast64_t> H, HN, BH;
  Array<double> r;
  Array<double> mua, mus, g, n;
  Array<double> k, g2;
  Array<char> BCType;
  Array<double> BCLNormal;
  Array<double> BCn;
  Array<double> BCIntensity;
  Array<char> BCLightDirectionType;
  Array<double> GaussianSigma;   // 2D 'g' sources
  double f, omega, phase0;
  int_fast64_t Nphoton;
  double c0;
  Array<double> ER, EI;
  Array<double> EBR, EBI;
  Array<double> DEBR, DEBI;
  double weight0, chance;
  long loss;
  unsigned long seed;
  int threadcount;
  int rank, nodecount;
  int totalthreads;
  int precision;
  int tally;
  vmc_stats stats{};

private:
  struct Progress {
    bool (*fn)(double);
    double last;
    static int call(double pct, void *user) {
      Progress *p = static_cast<Progress *>(user);
      p->last = pct;
      return p->fn(pct) ? 1 : 0;
    }
  };

  vmc_problem_desc desc() {
    vmc_problem_desc d = {};
    d.dim = DIM;
    d.Ne = H.Nx;
    d.Nb = BH.Nx;
    d.Nr = r.Nx;
    d.H = reinterpret_cast<const int64_t *>(H.data);
    d.HN = HN.N == H.N ? reinterpret_cast<const int64_t *>(HN.data) : nullptr;
    d.BH = reinterpret_cast<const int64_t *>(BH.data);
    d.r = r.data;
    d.mua = mua.data; d.mus = mus.data; d.g = g.data; d.n = n.data;
    d.BCType = BCType.data;
    d.BCn = BCn.N ? BCn.data : nullptr;
    d.BCLNormal = BCLNormal.N ? BCLNormal.data : nullptr;
    d.BCLightDirectionType = BCLightDirectionType.N ? BCLightDirectionType.data : nullptr;
    d.BCIntensity = BCIntensity.N ? BCIntensity.data : nullptr;
    d.GaussianSigma = (DIM == 2 && GaussianSigma.N) ? GaussianSigma.data : nullptr;
    return d;
  }
};

// geometry with the reference's expressions (MC3D.hpp:306-343, MC2D.hpp:303-318)
template <> inline double MCEngine<3>::ElementVolume(int_fast64_t el) {
  double ax = r(H(el, 0), 0), ay = r(H(el, 0), 1), az = r(H(el, 0), 2);
  double bx = r(H(el, 1), 0), by = r(H(el, 1), 1), bz = r(H(el, 1), 2);
  double cx = r(H(el, 2), 0), cy = r(H(el, 2), 1), cz = r(H(el, 2), 2);
  double dx = r(H(el, 3), 0), dy = r(H(el, 3), 1), dz = r(H(el, 3), 2);
  double v = (ax - dx) * ((by - dy) * (cz - dz) - (bz - dz) * (cy - dy)) -
             (ay - dy) * ((bx - dx) * (cz - dz) - (bz - dz) * (cx - dx)) +
             (az - dz) * ((bx - dx) * (cy - dy) - (by - dy) * (cx - dx));
  return fabs(v) / 6.0;
}
template <> inline double MCEngine<3>::ElementArea(int_fast64_t ib) {
  int_fast64_t i0 = BH(ib, 0), i1 = BH(ib, 1), i2 = BH(ib, 2);
  double a = sqrt(pow(r(i1, 0) - r(i0, 0), 2) + pow(r(i1, 1) - r(i0, 1), 2) + pow(r(i1, 2) - r(i0, 2), 2));
  double b = sqrt(pow(r(i2, 0) - r(i0, 0), 2) + pow(r(i2, 1) - r(i0, 1), 2) + pow(r(i2, 2) - r(i0, 2), 2));
  double c = sqrt(pow(r(i1, 0) - r(i2, 0), 2) + pow(r(i1, 1) - r(i2, 1), 2) + pow(r(i1, 2) - r(i2, 2), 2));
  return sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c)) / 4.0;
}
template <> inline double MCEngine<2>::ElementVolume(int_fast64_t el) {
  double ax = r(H(el, 1), 0) - r(H(el, 0), 0), ay = r(H(el, 1), 1) - r(H(el, 0), 1);
  double bx = r(H(el, 2), 0) - r(H(el, 0), 0), by = r(H(el, 2), 1) - r(H(el, 0), 1);
  double c = ax * by - ay * bx;
  return 0.5 * (c >= 0.0 ? c : -c);
}
template <> inline double MCEngine<2>::ElementArea(int_fast64_t ib) {
  return sqrt(pow(r(BH(ib, 1), 0) - r(BH(ib, 0), 0), 2) + pow(r(BH(ib, 1), 1) - r(BH(ib, 0), 1), 2));
}

#endif
