// Test program for the sanitizer build: the host part of the library (the
// ErrorChecks/Init pre-pass on up to 16 threads: neighbour matching, Morton
// order, light sources, records; and the device-free entry points) and the
// oracle, compiled with AddressSanitizer and UndefinedBehaviorSanitizer, on
// a problem read from files; also the device element order the pre-pass
// builds.  Prints "key value" lines; a sanitizer report
// ends the program with an error.
//   sanitize_driver <dir> <dim> <Ne> <Nb> <Nr> <nphoton>   (nphoton < 0: no oracle run)
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "valomc.h"
#include "vmc_host.h"
#include "vmc_oracle.h"

template <class T> static std::vector<T> load(const std::string &path, size_t n) {
  std::vector<T> v(n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T)));
  if (!f) {
    fprintf(stderr, "short file %s\n", path.c_str());
    exit(2);
  }
  return v;
}

int main(int argc, char **argv) {
  if (argc != 7) return 2;
  setvbuf(stdout, nullptr, _IOLBF, 0);
  const std::string d = argv[1];
  const int dim = atoi(argv[2]);
  const int64_t Ne = atoll(argv[3]), Nb = atoll(argv[4]), Nr = atoll(argv[5]), N = atoll(argv[6]);
  const auto H = load<int64_t>(d + "/H", Ne * (dim + 1));
  const auto BH = load<int64_t>(d + "/BH", Nb * dim);
  const auto r = load<double>(d + "/r", Nr * dim);
  const auto mua = load<double>(d + "/mua", Ne), mus = load<double>(d + "/mus", Ne);
  const auto g = load<double>(d + "/g", Ne), n = load<double>(d + "/n", Ne);
  const auto bct = load<char>(d + "/BCType", Nb);
  const auto bcn = load<double>(d + "/BCn", Nb), inten = load<double>(d + "/BCIntensity", Nb);

  vmc_problem_desc pd = {};
  pd.dim = dim;
  pd.Ne = Ne;
  pd.Nb = Nb;
  pd.Nr = Nr;
  pd.H = H.data();
  pd.BH = BH.data();
  pd.r = r.data();
  pd.mua = mua.data();
  pd.mus = mus.data();
  pd.g = g.data();
  pd.n = n.data();
  pd.BCType = bct.data();
  pd.BCn = bcn.data();
  pd.BCIntensity = inten.data();
  std::vector<int64_t> hn(Ne * (dim + 1));
  int64_t ns = 0;
  printf("check %d\n", vmc_problem_check(&pd, hn.data(), &ns));
  int64_t boundary = 0;
  for (int64_t v : hn) boundary += v < 0;
  printf("nsources %lld boundary_faces %lld\n", (long long)ns, (long long)boundary);
  std::vector<int64_t> faces(ns), mothers(ns);
  std::vector<double> cdf(ns);
  int64_t ns2 = 0;
  printf("sources %d %lld\n", vmc_problem_sources(&pd, faces.data(), mothers.data(), cdf.data(), ns, &ns2),
         (long long)ns2);
  // without a device: the entry points that need one say so
  vmc_problem *prob = nullptr;
  printf("create %d\n", vmc_problem_create(&pd, &prob));
  // an index outside the nodes: the reference's INCONSISTENT_H
  std::vector<int64_t> badH = H;
  badH[0] = Nr + 5;
  pd.H = badH.data();
  printf("bad_h %d\n", vmc_problem_check(&pd, nullptr, nullptr));
  pd.H = H.data();
  {
    // BuildNeighbourhoods alone, with and without the boundary entries:
    // without them the boundary faces match nothing (3D: INT64_MAX)
    std::vector<int64_t> hb(hn.size()), ht(hn.size());
    const int sb = vmc_build_neighbours(dim, Ne, Nb, Nr, H.data(), pd.BH, hb.data());
    const int st = vmc_build_neighbours(dim, Ne, 0, Nr, H.data(), nullptr, ht.data());
    int64_t same = 0, unmatched = 0;
    for (size_t i = 0; i < hn.size(); ++i) {
      same += hb[i] == hn[i];
      unmatched += ht[i] == INT64_MAX;
    }
    printf("neighbours_only %d %lld %lld %d %lld\n", sb, (long long)same, (long long)hn.size(), st,
           (long long)unmatched);
  }
  {
    // the device order the pre-pass builds: slots of whole sectors, every
    // element in one slot and back, no neighbour code or light source on a
    // padding slot; and, for a grid mesh whose element e lies in cell
    // e % (Ne / k) (k = 6 tetrahedra or 2 triangles per cell), how many
    // cells are split over two sectors
    vmc_problem m;
    const int st = vmc::host_prepare(&pd, &m);
    const int64_t slots = m.md->slots;
    int64_t used = 0, mismatch = 0, to_padding = 0, split = 0;
    for (int32_t e : m.perm) used += e >= 0;
    for (int64_t e = 0; e < Ne; ++e)
      if (m.inv[e] < 0 || m.inv[e] >= slots || m.perm[m.inv[e]] != e) ++mismatch;
    if (st == 0) {
      for (int64_t e = 0; e < Ne; ++e)
        for (int f = 0; f <= dim; ++f) {
          const int c = m.nbcode[e][f];
          if (c >= 0 && m.perm[c >> 3] < 0) ++to_padding;
        }
      for (const auto &src : m.src)
        if (m.perm[src.mother] < 0) ++to_padding;
    }
    const int k = dim == 3 ? 6 : 2;
    if (Ne % k == 0) {
      const int64_t nc = Ne / k;
      for (int64_t c = 0; c < nc; ++c) {
        bool two = false;
        for (int t = 1; t < k; ++t) two = two || m.inv[c + t * nc] / vmc::kSectorElems != m.inv[c] / vmc::kSectorElems;
        split += two ? 1 : 0;
      }
    }
    printf("order %d %lld %lld %lld %lld %lld\n", st, (long long)slots, (long long)used, (long long)mismatch,
           (long long)to_padding, (long long)split);
  }

  if (N < 0) return 0;   // host pre-pass only
  oracle_problem op = {};
  op.Ne = Ne;
  op.Nb = Nb;
  op.Nr = Nr;
  op.H = H.data();
  op.BH = BH.data();
  op.r = r.data();
  op.mua = mua.data();
  op.mus = mus.data();
  op.g = g.data();
  op.n = n.data();
  op.BCType = bct.data();
  op.BCn = bcn.data();
  op.BCIntensity = inten.data();
  op.Nphoton = N;
  op.seed = 3;
  op.weight0 = 0.001;
  op.chance = 0.1;
  op.c0 = 2.99792458e11;
  std::vector<double> ER(Ne), EI(Ne), EBR(Nb), EBI(Nb), DEBR(Nb), DEBI(Nb);
  oracle_result res = {ER.data(), EI.data(), EBR.data(), EBI.data(), DEBR.data(), DEBI.data(), nullptr, 0, {0}, 0};
  for (int mode : {ORACLE_RNG_MT19937, ORACLE_RNG_PHILOX}) {
    const int st = dim == 3 ? oracle_run3d(&op, 4, mode, &res) : oracle_run2d(&op, 4, mode, &res);
    double bal = res.ledger[0] + res.ledger[1] + res.ledger[2] - res.ledger[3] + res.ledger[4] - (double)N;
    printf("oracle %d %d loss %lld balance %.3e\n", mode, st, (long long)res.loss, bal / (double)N);
  }
  return 0;
}
