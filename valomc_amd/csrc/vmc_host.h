// Internal interface of the host library (not part of the C ABI in
// include/valomc.h): device buffers, the tally arena, the resident problem,
// the preparation steps of the reference's ErrorChecks()/Init() and the
// launch engine shared by single runs, batches and the serving loop.
//   vmc_prepare.cpp  ErrorChecks + Init: checks, neighbours, light sources,
//                    device order, element / source / boundary records
//   vmc_engine.cpp   launch engine and the problem / batch C entry points
//   vmc_server.cpp   the serving loop (vmc_server_*)
#pragma once

#include "../../include/valomc.h"
#include "vmc_types.h"

#include <climits>
#include <cstdint>
#include <memory>
#include <new>
#include <string>
#include <algorithm>
#include <vector>

namespace vmc {

// kernel launchers (vmc_kernels.hip)
// (small: walk_small_kernel, the item's element records in LDS; one item)
hipError_t walk_occupancy(int dim, bool f64, bool fd, bool fixed, bool batch, bool small, int block,
                          size_t lds_bytes, int *blocks_per_cu);
hipError_t launch_walk(int dim, bool f64, bool fd, bool fixed, bool batch, bool small, int grid, int block,
                       size_t lds_bytes, hipStream_t s, const WalkParams &P);
hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_ranges(unsigned long long *p, size_t n, int nrep, size_t stride, unsigned long long v,
                              hipStream_t s);
hipError_t launch_reduce(bool fixed, const void *rep, double *out, size_t n, int nrep, size_t stride,
                         double inv_scale, hipStream_t s);
// device address of the allocation-free entry point's scratch (kScratchBytes)
// on `device`, the current device
hipError_t scratch_address(int device, void **ptr);


// last error text of the calling thread (vmc_last_error)
extern thread_local std::string g_last_error;

// mcerror codes (cpp/Errors.hpp:4-21); the ABI reports 1 + code
enum McErr {
  NO_LIGHTSOURCE, SIZE_MISMATCH_G, SIZE_MISMATCH_MUA, SIZE_MISMATCH_MUS, SIZE_MISMATCH_N,
  SIZE_MISMATCH_BCN, MISSING_BCN, SIZE_MISMATCH_BCTYPE, SIZE_MISMATCH_LIGHT_DIRECTION,
  SIZE_MISMATCH_LIGHT_DIRECTION_TYPE, INCONSISTENT_H, INCONSISTENT_BH, INCONSISTENT_MESH,
  INCONSISTENT_MESH_DUPLICATE_NEIGHBORS, NO_GAUSSIAN_SIGMA, MISSING_BOUNDARY
};
inline int mc_status(int e) { return 1 + e; }

constexpr int kFace3[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}}; // MC3D.hpp:1013
constexpr int kEdge2[3][2] = {{0, 1}, {1, 2}, {2, 0}};                     // MC2D.hpp:921
constexpr int64_t kUnset = INT64_MAX;                                      // MC3D.hpp:29-31
constexpr double kC0 = 2.99792458e11;                                      // MC3D.hpp:178
constexpr double kWeight0 = 0.001, kChance = 0.1;                          // MC3D.hpp:182-183

// Model constants of a run (vmc_run_desc.model; the reference's members
// weight0, chance and c0)
struct RunModel {
  double weight0 = kWeight0, chance = kChance, c0 = kC0;
};
// the run's constants, validated; a status
int run_model(const vmc_run_desc *run, RunModel *m);

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


// an mcerror thrown by the preparation steps; the ABI reports 1 + code
struct McError {
  int code;
};

// Mesh geometry and what is derived from it alone: shared by a problem and
// its variants, immutable once the problem is created.
struct MeshData {
  std::vector<int64_t> H, HN, BH;
  std::vector<double> r;
  std::vector<double> vol, area;    // element volumes (2D areas), boundary areas (2D lengths)
  std::vector<i4> nbcode;           // per element (caller's numbering), neighbours as device ids
  std::vector<int32_t> perm, inv;   // device order: perm[device slot] = element (-1: padding), inv[element] = slot
  int64_t slots = 0;                // device slots: element records and tally words per replica, whole sectors
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

} // namespace vmc

struct vmc_problem {
  std::shared_ptr<vmc::MeshData> md;
  std::shared_ptr<vmc::MediumData> med;
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
  std::vector<vmc::SrcRec> src;
  std::vector<vmc::BndRec> bnd;
  std::vector<vmc::i4> &nbcode;
  std::vector<int32_t> &perm, &inv;

  int device = 0, num_cus = 0;
  int64_t lds_per_cu = 65536;
  vmc::DevBuf<double> cdf;
  vmc::DevBuf<vmc::SrcRec> dsrc;
  vmc::DevBuf<vmc::BndRec> dbnd;
  vmc::Arena arena;   // tallies and stream of this problem's own runs

  explicit vmc_problem(std::shared_ptr<vmc::MeshData> mesh = std::make_shared<vmc::MeshData>(),
                       std::shared_ptr<vmc::MediumData> medium = std::make_shared<vmc::MediumData>())
      : md(std::move(mesh)), med(std::move(medium)), H(md->H), HN(md->HN), BH(md->BH), r(md->r),
        vol(md->vol), area(md->area), nbcode(md->nbcode), perm(md->perm), inv(md->inv) {}

  int64_t h(int64_t el, int c) const { return H[el + Ne * c]; }
  int64_t hn(int64_t el, int c) const { return HN[el + Ne * c]; }
  int64_t bh(int64_t ib, int c) const { return BH[ib + Nb * c]; }
  double x(int64_t node, int d) const { return r[node + Nr * d]; }
  double bcl(int64_t ib, int d) const { return BCLNormal[ib + Nb * d]; }
  double &bcl_ref(int64_t ib, int d) { return BCLNormal[ib + Nb * d]; }
};

namespace vmc {

// ---- preparation (vmc_prepare.cpp) ----

// ErrorChecks() and the host part of Init() (MC3D.hpp:460-738); a status
int host_prepare(const vmc_problem_desc *d, vmc_problem *m);
// steps a variant repeats when its sources or boundary change
void build_light_sources(vmc_problem &m);
void prepare_directions(vmc_problem &m);
void build_source_records(vmc_problem &m);
void build_boundary(vmc_problem &m);
// element records of the requested precision on the device (when stale)
void upload_params(vmc_problem &m, bool f64);
// normalisation of MonteCarlo (MC3D.hpp:1793-1818, MC2D.hpp:1673-1703),
// from raw device-order words straight into the caller's outputs
void normalise_into(const vmc_problem &m, double omega, double c0, int64_t N, int W, const double *raw_e,
                    const double *raw_b, vmc_output *out);
// a HIP failure as a status, with its text in g_last_error
int fail_hip(const HipFail &f);
// a failed host allocation as a status (the C entry points catch
// std::bad_alloc, so no exception leaves the library)
inline int host_out_of_memory() noexcept {
  try {
    g_last_error = "out of host memory";
  } catch (...) {
  }
  return VMC_ERR_OUT_OF_MEMORY;
}
template <class T> void copy_in(std::vector<T> &dst, const T *src, int64_t n) {
  if (src) dst.assign(src, src + n);
  else dst.clear();
}

// ---- launch engine (vmc_engine.cpp) ----

constexpr uint64_t kChunk = 128;   // default photon ids a wave takes from the dispenser at once
// default dead lanes a wave collects before it creates photons: measured on
// the 384k-tet box (CW: 10, within 0.5 % of 6..16) and the 2M-tet FD box
// (16, 1.6 % faster than 6)
inline int default_refill(bool fd) { return fd ? 16 : 10; }
// device counters of a launch: [0] photon dispenser, [1] element steps,
// [2] lost photons, [3] FP32 exit faces re-decided in FP64, [4] fixed-point
// range flag
constexpr int kCounters = 8;
// fixed-point tallies: runs are limited to fewer photons than this, so the
// int64 sums of exitance and absorbed weight (at most about one unit per
// photon and word) stay below 2^31 * 2^32
constexpr int64_t kFixedMaxPhotons = (int64_t)1 << 31;
// VMC_ERR_TALLY_RANGE with its text
inline int range_error() {
  g_last_error = "a fixed-point tally deposit reached 2^30 (or was not finite): the int64 sums would not be exact";
  return VMC_ERR_TALLY_RANGE;
}
// throws when a launch's dispenser ended below its photon count
struct WalkLaunch;
void check_dispenser(unsigned long long dispensed, const WalkLaunch &w);

struct LaunchCfg {
  int dim = 3;
  bool f64 = false, fd = false, fixed = false, xcc = true;
  int W = 1, block = 256, slots = 0, refill = 6, nrep = 8;
  bool direct = false;
  uint64_t chunk = kChunk;
  int epoch = 0;   // -1: never; 0: by launch (batches 1024); > 0: passes
  int bpc = 0;     // workgroups per CU: 0 = as many as fit, else at most this many
  size_t lds = 0;
  double weight0 = kWeight0, chance = kChance;   // roulette constants of the launch
  // small mesh (choose_small): the single item's small_elems element records
  // are staged in each workgroup's LDS, with one cache slot per sector
  bool small = false;
  int32_t small_elems = 0;
};

int neighbours_only(int32_t dim, int64_t Ne, int64_t Nb, int64_t Nr, const int64_t *H, const int64_t *BH,
                    int64_t *HN_out);   // vmc_build_neighbours
int resolve_cfg(int dim, int num_cus, int64_t lds_per_cu, bool fd, const vmc_run_desc *run, LaunchCfg *c,
                bool batch);
void choose_replicas(LaunchCfg &c, const vmc_run_desc *run, int64_t ne, int64_t nb);
// A single problem of ne elements walks from LDS-resident records when they
// and a cache slot per sector fit the workgroup's LDS share and the cache
// knobs are the defaults; sets c.small, c.slots and c.lds.
void choose_small(LaunchCfg &c, const vmc_run_desc *run, int64_t ne, int64_t lds_per_cu);
// The arena's buffers for ne element and nb boundary slots and nitems
// items (no device work; nothing is allocated under stream capture).
void arena_alloc(Arena &a, const LaunchCfg &c, int64_t ne, int64_t nb, size_t nitems);
// Enqueues the clearing of its tallies, ledger, counters and loss counts
// (fill kernels, so a captured graph records no memset).
void arena_fill(Arena &a, const LaunchCfg &c);
void arena_reset(Arena &a, const LaunchCfg &c, int64_t ne, int64_t nb, size_t nitems);   // alloc + fill

// A persistent walk launch over a list of items: prepare_walk sets the
// items' virtual photon ranges (chunk-aligned, in order), uploads the item
// records and fills in the kernel parameters (grid 0: no photons);
// enqueue_walk enqueues the dispenser reset and the kernel.
struct WalkLaunch {
  WalkParams P{};
  int grid = 0;
  bool batch = false;
};
WalkLaunch prepare_walk(Arena &a, int num_cus, const LaunchCfg &c, std::vector<ItemRec> &items);
void enqueue_walk(Arena &a, const LaunchCfg &c, const WalkLaunch &w);
// prepare + enqueue + wait; returns the kernel time (HIP events on the arena's stream)
float launch_items(Arena &a, int num_cus, const LaunchCfg &c, std::vector<ItemRec> &items, int *grid_out);

// Enqueues the replica sums into raw element words [ne*W] and boundary
// words [nb*4]; arena_copy_out copies them, the item ledgers (and their sum
// in led), the counters and the per-item loss counts to the host and waits.
void arena_reduce(Arena &a, const LaunchCfg &c);
void arena_copy_out(Arena &a, const LaunchCfg &c, std::vector<double> &raw_e, std::vector<double> &raw_b,
                    double led[5], unsigned long long cnt[kCounters], std::vector<unsigned long long> &loss,
                    std::vector<double> &item_led);
void arena_collect(Arena &a, const LaunchCfg &c, std::vector<double> &raw_e, std::vector<double> &raw_b,
                   double led[5], unsigned long long cnt[kCounters], std::vector<unsigned long long> &loss,
                   std::vector<double> &item_led);   // reduce + copy out
// Ledger words of a launch of `nitems` items: five per item (absorbed,
// exited, killed, gained, lost; one item: the launch's own), at least 8.
inline size_t ledger_words(size_t nitems) { return std::max<size_t>(8, 5 * nitems); }
// led = the sum of the first nitems item ledgers
void sum_ledgers(const std::vector<double> &item_led, size_t nitems, double led[5]);
ItemRec make_item(const vmc_problem &m, const vmc_run_desc *run, bool f64, uint64_t count, uint64_t id0,
                  int64_t tbase, int64_t bbase);
void finish_outputs(const vmc_problem &m, const vmc_run_desc &run, int64_t N, int W, const double *raw_e,
                    const double *raw_b, vmc_output *out);
int64_t sector_round(int64_t n);

} // namespace vmc
