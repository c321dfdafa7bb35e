// Command-line drop-in for the reference's external executables
// (cpp/3d/MC3D.cpp, cpp/2d/MC2D.cpp): same syntax, same text problem
// format in, same solution format out.
//
// Input (written by ValoMC.m:200-258 / 293-336, read by MC3D.cpp:97-179):
//   Ne Nb Nr Nphoton
//   f phase0 seed useseed
//   <one line skipped: nx ny [nz]>
//   H / BH / r / "mua mus g n" / BCType sections, then the optional
//   BCn, BCLightDirection, BCLightDirectionType (3D reader: BCLDirType),
//   BCIntensity (MATLAB exporter: BCLIntesity) and, in 2D, GaussianSigma.
// Each section is a tag line followed by one row per element.  A tag that
// does not match leaves the stream where it was (fileio.hpp:12-48); the
// optional sections are accepted in any order and under both spellings.
// Output: "<seconds> 0", then Ne rows "ER EI" and Nb rows "EBR EBI", with
// %18.10e in 3D and %e in 2D (MC3D.cpp:186-201, MC2D.cpp:212-226).
#ifndef VMC_HOST_CLI_HPP
#define VMC_HOST_CLI_HPP

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <chrono>
#include <sstream>
#include <string>
#include <vector>

#include "MCEngine.hpp"

namespace vmc_cli {

struct LineReader {
  FILE *fp;
  bool line(std::string &out) {
    out.clear();
    char buf[8192];
    bool any = false;
    while (fgets(buf, sizeof(buf), fp)) {
      any = true;
      out += buf;
      if (!out.empty() && out.back() == '\n') break;
    }
    return any;
  }
};

static std::vector<std::string> words(const std::string &s) {
  std::istringstream is(s);
  std::vector<std::string> w;
  std::string t;
  while (is >> t) w.push_back(t);
  return w;
}

// Read `rows` lines of `cols` values per array after a matched tag.
template <class T>
static bool read_rows(LineReader &lr, int64_t rows, int cols, std::vector<Array<T> *> arrays) {
  for (Array<T> *a : arrays) {
    if (cols == 1) a->resize(rows); else a->resize(rows, cols);
  }
  std::string s;
  for (int64_t i = 0; i < rows; ++i) {
    if (!lr.line(s)) return false;
    std::istringstream is(s);
    for (Array<T> *a : arrays)
      for (int c = 0; c < cols; ++c) {
        T v;
        if (!(is >> v)) return false;
        (*a)(i, c) = v;
      }
  }
  for (Array<T> *a : arrays) printf("%12s   (%ldx%ld)\n", "", (long)a->Nx, (long)a->Ny);
  return true;
}

// Try one tagged section at the current position; rewind when the tag
// line is not one of `tags`.
template <class T>
static int try_section(LineReader &lr, const std::vector<std::string> &tags, int64_t rows, int cols,
                       std::vector<Array<T> *> arrays) {
  fpos_t pos;
  fgetpos(lr.fp, &pos);
  std::string s;
  if (!lr.line(s)) return 0;
  std::vector<std::string> w = words(s);
  std::string tag;
  for (size_t i = 0; i < w.size(); ++i) tag += (i ? " " : "") + w[i];
  for (const std::string &t : tags)
    if (tag == t) {
      printf("%12s\n", t.c_str());
      return read_rows<T>(lr, rows, cols, arrays) ? 1 : -1;
    }
  fsetpos(lr.fp, &pos);
  return 0;
}

template <int DIM>
static int load_problem(const char *fin, MCEngine<DIM> &MC) {
  FILE *fp = fopen(fin, "r");
  if (!fp) return 1;
  LineReader lr{fp};
  long Ne = 0, Nb = 0, Nr = 0, nph = 0, sd1 = 0, sd2 = 0;
  std::string s;
  if (!lr.line(s) || sscanf(s.c_str(), "%ld %ld %ld %ld", &Ne, &Nb, &Nr, &nph) != 4) { fclose(fp); return 1; }
  if (!lr.line(s) || sscanf(s.c_str(), "%lf %lf %ld %ld", &MC.f, &MC.phase0, &sd1, &sd2) < 2) { fclose(fp); return 1; }
  lr.line(s);   // nx ny [nz]
  MC.Nphoton = nph;
  if (sd2) MC.seed = (unsigned long)sd1;
  else MC.seed = (unsigned long)std::chrono::high_resolution_clock::now().time_since_epoch().count();
  if (MC.phase0 < 0) MC.phase0 += 2 * M_PI * ceil(-MC.phase0 / (2 * M_PI));   // MC3D.cpp:124-126
  printf("Constants:\n  %10s   (%e)\n  %10s   (%e)\n  %10s   (%li)\n  %10s   (%li)\n  %10s   (%li)\n"
         "  %10s   (%li)\n", "f", MC.f, "phase0", MC.phase0, "Ne", Ne, "Nb", Nb, "Nr", Nr, "Nphoton", nph);
  printf("seed %lu\nArrays:\n", MC.seed);
  bool ok = try_section<int_fast64_t>(lr, {"H"}, Ne, DIM + 1, {&MC.H}) == 1 &&
            try_section<int_fast64_t>(lr, {"BH"}, Nb, DIM, {&MC.BH}) == 1 &&
            try_section<double>(lr, {"r"}, Nr, DIM, {&MC.r}) == 1 &&
            try_section<double>(lr, {"mua mus g n"}, Ne, 1, {&MC.mua, &MC.mus, &MC.g, &MC.n}) == 1 &&
            try_section<char>(lr, {"BCType"}, Nb, 1, {&MC.BCType}) == 1;
  if (!ok) {
    printf("Reading error!\n");
    fclose(fp);
    return 1;
  }
  for (;;) {
    int got = try_section<double>(lr, {"BCn"}, Nb, 1, {&MC.BCn});
    if (!got) got = try_section<double>(lr, {"BCLightDirection"}, Nb, DIM, {&MC.BCLNormal});
    if (!got) got = try_section<char>(lr, {"BCLDirType", "BCLightDirectionType"}, Nb, 1, {&MC.BCLightDirectionType});
    if (!got) got = try_section<double>(lr, {"BCIntensity", "BCLIntesity"}, Nb, 1, {&MC.BCIntensity});
    if (!got && DIM == 2) got = try_section<double>(lr, {"GaussianSigma"}, Nb, 1, {&MC.GaussianSigma});
    if (got <= 0) break;
  }
  fclose(fp);
  return 0;
}

template <int DIM>
static int save_problem(const char *fout, MCEngine<DIM> &MC, int seconds, bool precise) {
  FILE *fp = fopen(fout, "w");
  if (!fp) return 1;
  fprintf(fp, "%i 0\n", seconds);
  const char *fmt = precise ? "%.17e %.17e\n" : (DIM == 3 ? "%18.10e %18.10e\n" : "%e %e\n");
  for (int_fast64_t i = 0; i < MC.ER.Nx; ++i) fprintf(fp, fmt, MC.ER[i], MC.EI[i]);
  for (int_fast64_t i = 0; i < MC.EBR.Nx; ++i) fprintf(fp, fmt, MC.EBR[i], MC.EBI[i]);
  fclose(fp);
  return 0;
}

static bool g_quiet = false;
static bool progress_print(double perc) {
  if (!g_quiet) {
    printf("  %f %%\r", perc);
    fflush(stdout);
  }
  return true;
}
static void final_checks(int, int) {}

template <int DIM>
static int cli_main(int argc, char **argv) {
  const char *name = DIM == 3 ? "MC3D" : "MC2D";
  printf("                 ValoMC-%dD (%s)\n", DIM, vmc_version());
  if (argc < 3) {
    printf("Use syntax:\n %s inputfile outputfile [-s] [--fp32] [--fixed] [--precise] [--check]\n", name);
    printf("  -s        suppress the progress output\n");
    printf("  --fp32    single-precision transport (FP64 tallies)\n");
    printf("  --fixed   64-bit fixed-point tallies (bit-reproducible)\n");
    printf("  --precise write the solution with 17 significant digits\n");
    printf("  --check   parse the file and run ErrorChecks() + Init() only (no GPU)\n");
    return 0;
  }
  MCEngine<DIM> MC;
  bool precise = false, check_only = false;
  for (int i = 3; i < argc; ++i) {
    if (!strcmp(argv[i], "-s")) g_quiet = true;
    else if (!strcmp(argv[i], "--fp32")) MC.precision = VMC_FP32;
    else if (!strcmp(argv[i], "--fixed")) MC.tally = VMC_TALLY_FIXED64;
    else if (!strcmp(argv[i], "--precise")) precise = true;
    else if (!strcmp(argv[i], "--check")) check_only = true;
  }
  printf("Loading problem %s\n", argv[1]);
  if (load_problem<DIM>(argv[1], MC)) {
    printf("Error while loading!\n");
    return 1;
  }
  // optional sections the file held (Init() fills in defaults)
  const int had_bcn = MC.BCn.N > 0, had_bcl = MC.BCLNormal.N > 0, had_int = MC.BCIntensity.N > 0,
            had_sigma = MC.GaussianSigma.N > 0;
  printf("Initializing %s\n", name);
  try {
    MC.ErrorChecks();
    MC.Init();
  } catch (mcerror e) {
    printf("MC Initialization failed!: Reason %s\n", errorstring(e));
    return 1;
  } catch (vmc_device_error e) {
    printf("MC Initialization failed!: Reason %s\n", e.what());
    return 1;
  }
  if (check_only) {
    // parse + ErrorChecks + Init only: a summary line (sizes, constants, a
    // checksum of the neighbour table, the optional sections read), no
    // transport
    long long hsum = 0;
    for (int_fast64_t i = 0; i < MC.HN.N; ++i) hsum += (long long)(MC.HN[i] % 1000003);
    int nls = 0, ndir = 0;
    for (int_fast64_t i = 0; i < MC.BCType.N; ++i) nls += MC.BCType[i] != 'a';
    for (int_fast64_t i = 0; i < MC.BCLightDirectionType.N; ++i) ndir += MC.BCLightDirectionType[i] != 'n';
    printf("CHECK dim=%d Ne=%ld Nb=%ld Nr=%ld Nphoton=%ld seed=%lu f=%.10g phase0=%.10g hnsum=%lld "
           "nonabsorbing_bc=%d directed=%d bcn=%d bcl=%d intensity=%d sigma=%d\n",
           DIM, (long)MC.H.Nx, (long)MC.BH.Nx, (long)MC.r.Nx, (long)MC.Nphoton, MC.seed, MC.f, MC.phase0,
           hsum, nls, ndir, had_bcn, had_bcl, had_int, had_sigma);
    return 0;
  }
  printf("Computing...\n");
  const int tstart = (int)time(NULL);
  try {
    MC.MonteCarlo(progress_print, final_checks);
  } catch (mcerror e) {
    printf("MC failed!: Reason %s\n", errorstring(e));
    return 1;
  } catch (vmc_device_error e) {
    printf("MC failed!: Reason %s (%s)\n", e.what(), vmc_last_error());
    return 1;
  }
  const int tend = (int)time(NULL);
  printf("Saving problem\n");
  if (save_problem<DIM>(argv[2], MC, tend - tstart, precise)) {
    printf("Error while saving!\n");
    return 1;
  }
  if (MC.loss) printf(" %ld photons lost during computation!\n", MC.loss);
  printf("Computation took %i seconds (transport kernels %.3f s, %lld element steps)\n", tend - tstart,
         MC.stats.kernel_ms / 1e3, (long long)MC.stats.steps);
  return 0;
}

} // namespace vmc_cli

#endif
