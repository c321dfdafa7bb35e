// Test program: the drop-in classes' host-side members on a problem read
// from files (the arrays tests/test_host_classes.py writes).  Prints
// "key value..." lines.  Needs no GPU: Init() runs the host pre-pass only.
#include <cstdio>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "MC2D.hpp"
#include "MC3D.hpp"

template <class T> static void load(Array<T> &a, const std::string &path, int64_t rows, int64_t cols) {
  std::ifstream f(path, std::ios::binary);
  if (cols == 1) a.resize(rows); else a.resize(rows, cols);
  f.read(reinterpret_cast<char *>(a.data), (std::streamsize)(sizeof(T) * rows * cols));
}

template <class E> static void fill(E &mc, const char *dir, int dim, int64_t Ne, int64_t Nb, int64_t Nr) {
  const std::string d(dir);
  load(mc.H, d + "/H", Ne, dim + 1);
  load(mc.BH, d + "/BH", Nb, dim);
  load(mc.r, d + "/r", Nr, dim);
  load(mc.mua, d + "/mua", Ne, 1);
  load(mc.mus, d + "/mus", Ne, 1);
  load(mc.g, d + "/g", Ne, 1);
  load(mc.n, d + "/n", Ne, 1);
  load(mc.BCType, d + "/BCType", Nb, 1);
  load(mc.BCn, d + "/BCn", Nb, 1);
  load(mc.BCIntensity, d + "/BCIntensity", Nb, 1);
}

template <class E> static void common(E &mc) {
  mc.ErrorChecks();
  mc.Init();
  printf("nsources %lld\n", (long long)mc.LightSources.N);
  printf("sources");
  for (int64_t i = 0; i < mc.LightSources.N; ++i) printf(" %lld", (long long)mc.LightSources[i]);
  printf("\nmothers");
  for (int64_t i = 0; i < mc.LightSources.N; ++i) printf(" %lld", (long long)mc.LightSourcesMother[i]);
  printf("\ncdf");
  for (int64_t i = 0; i < mc.LightSources.N; ++i) printf(" %.17g", mc.LightSourcesCDF[i]);
  printf("\n");
  // the host random stream: seed 77, the reference's four maps
  mc.seed = 77;
  mc.InitRand();
  std::mt19937 ref(77);
  int bad = 0;
  for (int i = 0; i < 1000; ++i) {
    const double u = ref();
    bad += mc.UnifClosed() != u * (1.0 / 4294967295.0);
    bad += mc.UnifOpen() != ((double)ref() + 0.5) * (1.0 / 4294967296.0);
    bad += mc.UnifHalfUp() != (double)ref() * (1.0 / 4294967296.0);
    bad += mc.UnifHalfDown() != ((double)ref() + 1.0) * (1.0 / 4294967296.0);
  }
  printf("rng_mismatch %d\n", bad);
  // assignment: the problem is shared, the outputs are the copy's own and zero
  mc.ER[0] = 5.0;
  E cl;
  cl = mc;
  printf("shared %d %d %d\n", cl.H.data == mc.H.data, cl.mua.data == mc.mua.data,
         cl.LightSourcesCDF.data == mc.LightSourcesCDF.data);
  printf("own_outputs %d %g %lld\n", cl.ER.data != mc.ER.data, cl.ER[0], (long long)cl.ER.N);
  printf("copied_seed %lu %d\n", cl.seed, cl.UnifClosed() == (double)std::mt19937(77)() * (1.0 / 4294967295.0));
}

// BuildNeighbourhoods() on topology alone, as createBH3mex calls it (H and
// a node array of the right length, nothing else; with_bh: BH too)
template <class E> static void topology_only(const E &mc, int64_t Nr, bool with_bh, const char *key) {
  E t;
  t.H = mc.H;
  if (with_bh) t.BH = mc.BH;
  t.r.resize(Nr);
  t.BuildNeighbourhoods();
  printf("%s", key);
  for (int64_t i = 0; i < t.HN.N; ++i) printf(" %lld", (long long)t.HN[i]);
  printf("\n");
}

int main(int argc, char **argv) {
  if (argc != 6) return 2;
  const int dim = atoi(argv[2]);
  const int64_t Ne = atoll(argv[3]), Nb = atoll(argv[4]), Nr = atoll(argv[5]);
  if (dim == 3) {
    MC3D mc;
    fill(mc, argv[1], 3, Ne, Nb, Nr);
    common(mc);
    topology_only(mc, Nr, false, "hn_topology");
    topology_only(mc, Nr, true, "hn_with_bh");
    printf("volume");
    for (int64_t e = 0; e < Ne; ++e) printf(" %.17g", mc.ElementVolume(e));
    printf("\nbarea");
    for (int64_t b = 0; b < Nb; ++b) printf(" %.17g", mc.ElementArea(b));
    printf("\nbnormal");
    for (int64_t b = 0; b < Nb; ++b) {
      double nn[3];
      mc.Normal(b, nn);
      printf(" %.17g %.17g %.17g", nn[0], nn[1], nn[2]);
    }
    printf("\nfarea");
    for (int64_t e = 0; e < Ne; ++e)
      for (long f = 0; f < 5; ++f) printf(" %.17g", mc.ElementArea(e, f));
    printf("\nfnormal");
    for (int64_t e = 0; e < Ne; ++e)
      for (long f = 0; f < 5; ++f) {
        double nn[3];
        mc.Normal(e, f, nn);
        printf(" %.17g %.17g %.17g", nn[0], nn[1], nn[2]);
      }
    printf("\n");
  } else {
    MC2D mc;
    fill(mc, argv[1], 2, Ne, Nb, Nr);
    common(mc);
    topology_only(mc, Nr, true, "hn_with_bh");
    printf("volume");
    for (int64_t e = 0; e < Ne; ++e) printf(" %.17g", mc.ElementArea(e));
    printf("\nbarea");
    for (int64_t b = 0; b < Nb; ++b) printf(" %.17g", mc.ElementLength(b));
    printf("\nbnormal");
    for (int64_t b = 0; b < Nb; ++b) {
      double nn[2];
      mc.Normal(b, nn);
      printf(" %.17g %.17g", nn[0], nn[1]);
    }
    printf("\nfarea");
    for (int64_t e = 0; e < Ne; ++e)
      for (int64_t f = 0; f < 4; ++f) printf(" %.17g", mc.ElementLength(e, f));
    printf("\nfnormal");
    for (int64_t e = 0; e < Ne; ++e)
      for (int64_t f = 0; f < 4; ++f) {
        double nn[2];
        mc.Normal(e, f, nn);
        printf(" %.17g %.17g", nn[0], nn[1]);
      }
    printf("\n");
  }
  return 0;
}
