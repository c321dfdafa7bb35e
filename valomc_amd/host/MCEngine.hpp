// Drop-in engine classes over the C ABI.
//
// MC3D and MC2D keep the reference classes' public members and their
// meaning (cpp/3d/MC3D.hpp:107-172, cpp/2d/MC2D.hpp:112-180): fill the
// arrays, call ErrorChecks(), Init() and MonteCarlo(progress, finalchecks),
// read the normalised ER/EI/EBR/EBI/DEBR/DEBI and loss.  The photon-level
// methods of the reference (CreatePhoton, PropagatePhoton, WhichFace, ...)
// run on the GPU inside MonteCarlo() and have no host counterpart here.
// The host-side helpers keep their reference meaning: the light-source list
// (BuildLightSource, filled by Init), the random stream `rng` with
// InitRand/Unif*, the geometry queries of MC3D/MC2D (in MC3D.hpp, MC2D.hpp)
// and the sharing assignment operator.
// Two members are added: `precision` (VMC_FP64 by default, the reference's
// arithmetic, or VMC_FP32) and `tally` (VMC_TALLY_FP64 or
// VMC_TALLY_FIXED64), plus `stats` with the device counters of the last run.
#ifndef VMC_HOST_MCENGINE_HPP
#define VMC_HOST_MCENGINE_HPP

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <type_traits>
#include <vector>

#include "Array.hpp"
#include "Errors.hpp"
#include "mt_rng.hpp"
#include "valomc.h"

template <int DIM> class MCEngine {
public:
  // light-source indices: int in MC3D (MC3D.hpp:155-156), int_fast64_t in
  // MC2D (MC2D.hpp:161-162)
  typedef typename std::conditional<DIM == 3, int, int_fast64_t>::type SourceIndex;

  MCEngine()
      : f(0.0), omega(0.0), phase0(0.0), Nphoton(1), c0(2.99792458e11), weight0(0.001), chance(0.1),
        loss(0), seed(5489UL), threadcount(1), rank(0), nodecount(1), totalthreads(1),
        precision(VMC_FP64), tally(VMC_TALLY_FP64) {}
  virtual ~MCEngine() {}

  // Assignment shares the problem (geometry, medium, boundary, light
  // sources: Array copies are views) and gives the left-hand side its own
  // zeroed outputs and a random stream reseeded from seed, as the
  // reference's per-thread clones (MC3D.hpp:200-273).
  MCEngine &operator=(const MCEngine &ref) {
    if (this == &ref) return *this;
    H = ref.H; HN = ref.HN; BH = ref.BH; r = ref.r;
    mua = ref.mua; mus = ref.mus; g = ref.g; n = ref.n; k = ref.k; g2 = ref.g2;
    BCType = ref.BCType; BCLNormal = ref.BCLNormal; BCIntensity = ref.BCIntensity;
    BCLightDirectionType = ref.BCLightDirectionType; BCn = ref.BCn; GaussianSigma = ref.GaussianSigma;
    f = ref.f; omega = ref.omega; phase0 = ref.phase0; Nphoton = ref.Nphoton; c0 = ref.c0;
    ER.resize(ref.ER.N); EI.resize(ref.EI.N);
    EBR.resize(ref.EBR.N); EBI.resize(ref.EBI.N);
    DEBR.resize(ref.DEBR.N); DEBI.resize(ref.DEBI.N);   // resize() zero-fills
    if (!BCIntensity.N) {
      BCIntensity.resize(BCType.N);
      for (int_fast64_t i = 0; i < BCIntensity.N; ++i) BCIntensity[i] = 1.0;
    }
    LightSources = ref.LightSources;
    LightSourcesMother = ref.LightSourcesMother;
    LightSourcesCDF = ref.LightSourcesCDF;
    weight0 = ref.weight0; chance = ref.chance; loss = ref.loss;
    threadcount = ref.threadcount; rank = ref.rank; nodecount = ref.nodecount; totalthreads = ref.totalthreads;
    precision = ref.precision; tally = ref.tally;
    seed = ref.seed;
    InitRand();
    return *this;
  }

  // random stream of host code (MC3D.hpp:276-303); the transport itself
  // draws from per-photon streams on the GPU
  void InitRand() { rng.Seed(seed); }
  double UnifClosed() { return rng.drand_closed(); }      // [0, 1]
  double UnifOpen() { return rng.drand_open(); }          // ]0, 1[
  double UnifHalfDown() { return rng.drand_open_down(); } // ]0, 1]
  double UnifHalfUp() { return rng.drand_open_up(); }     // [0, 1[

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
    BuildLightSource();
  }

  // Light sources of the boundary conditions (MC3D.hpp:941-1002): the
  // emitting boundary faces in element-major order, their elements and the
  // normalised CDF of area (2D: length) times BCIntensity.  Kept once built,
  // as in the reference.
  void BuildLightSource() {
    if (LightSources.N != 0) return;
    vmc_problem_desc d = desc();
    int64_t ns = 0;
    vmc_throw_status(vmc_problem_sources(&d, nullptr, nullptr, nullptr, 0, &ns));
    std::vector<int64_t> face((size_t)ns), mother((size_t)ns);
    std::vector<double> cdf((size_t)ns);
    vmc_throw_status(vmc_problem_sources(&d, face.data(), mother.data(), cdf.data(), ns, &ns));
    LightSources.resize(ns);
    LightSourcesMother.resize(ns);
    LightSourcesCDF.resize(ns);
    for (int64_t i = 0; i < ns; ++i) {
      LightSources[i] = (SourceIndex)face[i];
      LightSourcesMother[i] = (SourceIndex)mother[i];
      LightSourcesCDF[i] = cdf[i];
    }
  }

  // Neighbour table of H (and BH, if set) alone (MC3D.hpp:835-871): faces
  // that match nothing stay INT_FAST64_MAX, as createBH3mex reads them.
  // Needs only H, BH and the node count r.Nx; a given HN is kept.
  void BuildNeighbourhoods() {
    if (HN.N == H.N) return;   // given (MC3D.hpp:839)
    std::vector<int64_t> hn((size_t)H.Nx * (DIM + 1));
    vmc_throw_status(vmc_build_neighbours(DIM, H.Nx, BH.Nx, r.Nx, H.data, BH.N ? BH.data : nullptr, hn.data()));
    HN.resize(H.Nx, DIM + 1);
    for (size_t i = 0; i < hn.size(); ++i) HN[i] = hn[i];
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
  Array<SourceIndex> LightSources;
  Array<SourceIndex> LightSourcesMother;
  Array<double> LightSourcesCDF;
  double weight0, chance;
  long loss;
  mt_rng rng;
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

protected:
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

#endif
