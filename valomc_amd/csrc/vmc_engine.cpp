// Launch engine and the problem / batch entry points of the C ABI.
// vmc_problem_run is the reference's MonteCarlo() (MC3D.hpp:1641-1822): it
// launches the persistent transport kernel, sums the tally replicas on the
// device and normalises on the host with the reference's own expressions.
#include "vmc_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>

namespace vmc {

// ---------------------------------------------------------------------------
// Launch engine: persistent walk launches over a list of items (problem
// instances), shared by single runs and batches
// ---------------------------------------------------------------------------


int run_model(const vmc_run_desc *run, RunModel *m) {
  *m = RunModel{};
  if (run->model == 0) return VMC_OK;
  if (run->model != 1) return VMC_ERR_INVALID_ARGUMENT;
  if (!(std::isfinite(run->weight0) && run->weight0 >= 0.0) || !(run->chance > 0.0 && run->chance <= 1.0) ||
      !(std::isfinite(run->c0) && run->c0 > 0.0))
    return VMC_ERR_INVALID_ARGUMENT;
  m->weight0 = run->weight0;
  m->chance = run->chance;
  m->c0 = run->c0;
  return VMC_OK;
}

// Validates a run's precision, tally, model and tuning fields and resolves
// the launch shape on a device with `num_cus` CUs of `lds_per_cu` bytes of
// LDS; `batch`: the launch may walk several items (their kernels keep four
// ledger slots per wave in static LDS).
// LDS of one workgroup when a CU holds the waves its registers allow (FP32
// 16, FP64 8), and the part of it the lane buffers and the kernel's static
// LDS take (16 waves' ledgers: 40 bytes a wave for one item, 44 a slot + 8
// for a batch; plus alignment).
static void lds_budget(const LaunchCfg &c, int64_t lds_per_cu, bool batch, size_t *per_block, size_t *fixed_part) {
  const int waves_per_cu = c.f64 ? 8 : 16;
  const int bpc_target = std::max(1, waves_per_cu / (c.block / 64));
  *per_block = (size_t)lds_per_cu / bpc_target;
  const size_t wave_ledger = batch ? ((size_t)kBatchLedgerSlots * 44 + 8 + 7) / 8 * 8 : 40;
  *fixed_part = WalkLds(c.W, c.f64 ? 8 : 4, 0, c.block).total + 16 * wave_ledger + 384;
}

int resolve_cfg(int dim, int num_cus, int64_t lds_per_cu, bool fd, const vmc_run_desc *run, LaunchCfg *c,
                bool batch) {
  (void)num_cus;
  if (run->precision != VMC_FP64 && run->precision != VMC_FP32) return VMC_ERR_INVALID_ARGUMENT;
  if (run->tally != VMC_TALLY_FP64 && run->tally != VMC_TALLY_FIXED64) return VMC_ERR_INVALID_ARGUMENT;
  RunModel mdl;
  if (run_model(run, &mdl) != VMC_OK) return VMC_ERR_INVALID_ARGUMENT;
  c->weight0 = mdl.weight0;
  c->chance = mdl.chance;
  c->dim = dim;
  c->f64 = run->precision == VMC_FP64;
  c->fixed = run->tally == VMC_TALLY_FIXED64;
  c->fd = fd;
  c->W = fd ? 2 : 1;
  // FP32 transport runs as one 1024-lane workgroup per CU (16 waves sharing
  // one sector cache), FP64 (occupancy 2 waves/SIMD) as 256-lane
  // workgroups, two per CU.  A launch of several items runs FP32 as two
  // 512-lane workgroups per CU: its cache epochs are barriers, at which
  // eight waves idle less than sixteen (measured: 2-3 % faster on batches
  // and serving launches of 2D and 3D, CW and FD problems)
  c->block = run->block_size > 0 ? run->block_size : (c->f64 ? 256 : (batch ? 512 : 1024));
  if (c->block % 64 != 0 || c->block > (c->f64 ? 256 : 1024)) return VMC_ERR_INVALID_ARGUMENT;
  // replicas == 0: one per XCD, workgroup-scope atomics (choose_replicas may
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
    size_t per_block, fixed_part;
    lds_budget(*c, lds_per_cu, batch, &per_block, &fixed_part);
    const size_t entry = (size_t)kSectorElems * c->W * 8 + 4;
    c->slots = per_block > fixed_part ? (int)((per_block - fixed_part) / entry) : 0;
  }
  if (c->slots > (1 << 20)) return VMC_ERR_INVALID_ARGUMENT;
  c->lds = WalkLds(c->W, rbytes, c->slots, c->block).total;
  // creating photons for several dead lanes at once keeps the creation
  // code's lanes busy
  c->refill = run->refill_min > 0 ? std::min(run->refill_min, 64) : default_refill(fd);
  // photon ids per dispenser fetch: a power of 2 (chunks never straddle
  // items), at least a wave's 64 lanes (one fetch refills all dead lanes)
  if (run->chunk != 0 && (run->chunk < 64 || run->chunk > 4096 || (run->chunk & (run->chunk - 1))))
    return VMC_ERR_INVALID_ARGUMENT;
  c->chunk = run->chunk ? (uint64_t)run->chunk : kChunk;
  if (run->cache_epoch < -1) return VMC_ERR_INVALID_ARGUMENT;
  c->epoch = run->cache_epoch;
  if (run->blocks_per_cu < 0) return VMC_ERR_INVALID_ARGUMENT;
  c->bpc = run->blocks_per_cu;
  return VMC_OK;
}

// replicas == 0: the tally copies follow the size of one copy.  The L2s
// pass these FP64/u64 atomics on to the memory side either way (the TCC
// counters count every atomic request as an EA atomic); what the copies
// change is how many lines the requests spread over and whether they stay
// in the Infinity Cache.  Up to an XCD's L2 (4 MiB) one copy per XCD,
// added with workgroup-scope atomics, spreads them at no cost; a copy that
// exceeds that but fits the Infinity Cache comfortably (64 MiB) is better
// kept single, its device-scope atomics meeting in the cache instead of
// eight copies streaming through it (the 2M-tet FD box: +18%); larger ones
// (many batch instances) go back to one per XCD.
void choose_replicas(LaunchCfg &c, const vmc_run_desc *run, int64_t ne, int64_t nb) {
  if (run->replicas > 0) return;
  const double bytes = ((double)ne * c.W + (double)nb * 4) * 8;
  const bool single = bytes > 4.0 * (1 << 20) && bytes <= 64.0 * (1 << 20);
  c.xcc = !single;
  c.nrep = single ? 1 : 8;
}

static size_t elem_bytes(int dim, bool f64) {
  if (dim == 3) return f64 ? sizeof(Elem<4, double>) : sizeof(Elem<4, float>);
  return f64 ? sizeof(Elem<3, double>) : sizeof(Elem<3, float>);
}

// ne: the problem's device element slots (whole sectors).  With the records
// in LDS the walk reads no record from memory, and with a slot per sector
// every deposit stays in LDS until the workgroup's final flush.  Explicit
// cache knobs keep the general path.
void choose_small(LaunchCfg &c, const vmc_run_desc *run, int64_t ne, int64_t lds_per_cu) {
  c.small = false;
  if (run->lds_slots != 0 || run->cache_epoch > 0 || ne <= 0 || ne % kSectorElems != 0) return;
  size_t per_block, fixed_part;
  lds_budget(c, lds_per_cu, false, &per_block, &fixed_part);
  const size_t rec = elem_bytes(c.dim, c.f64);
  if ((size_t)ne * rec >= per_block) return;
  const int nsect = (int)(ne / kSectorElems);
  const size_t lay = WalkLds(c.W, c.f64 ? 8 : 4, nsect, c.block).total;
  const size_t statics = fixed_part - WalkLds(c.W, c.f64 ? 8 : 4, 0, c.block).total;
  if (lay + (size_t)ne * rec + statics > per_block) return;
  c.small = true;
  c.small_elems = (int32_t)ne;
  c.slots = nsect;
  c.lds = lay + (size_t)ne * rec;
}

// Sizes the arena for ne element and nb boundary slots and clears its
// tallies, ledger, counters and per-item loss counts.
void arena_alloc(Arena &a, const LaunchCfg &c, int64_t ne, int64_t nb, size_t nitems) {
  a.ne = ne;
  a.nb = nb;
  a.W = c.W;
  a.nrep = c.nrep;
  a.nitems = nitems;
  a.etally.ensure((size_t)ne * c.W * c.nrep);
  a.btally.ensure((size_t)nb * 4 * c.nrep);
  a.raw_e.ensure((size_t)ne * c.W);
  a.raw_b.ensure((size_t)nb * 4);
  a.ledger.ensure(ledger_words(nitems));
  a.counters.ensure(8);
  a.item_loss.ensure(nitems);
  a.items.ensure(nitems);
}

void arena_fill(Arena &a, const LaunchCfg &c) {
  const size_t ew = (size_t)a.ne * c.W * c.nrep, bw = (size_t)a.nb * 4 * c.nrep;
  hip_check(launch_fill_u64(a.etally.p, ew, 0ull, a.stream), "fill");
  hip_check(launch_fill_u64(a.btally.p, bw, 0ull, a.stream), "fill");
  hip_check(launch_fill_u64(reinterpret_cast<unsigned long long *>(a.ledger.p), ledger_words(a.nitems), 0ull, a.stream),
            "fill");
  hip_check(launch_fill_u64(a.counters.p, 8, 0ull, a.stream), "fill");
  hip_check(launch_fill_u64(a.item_loss.p, a.nitems, 0ull, a.stream), "fill");
}

void arena_reset(Arena &a, const LaunchCfg &c, int64_t ne, int64_t nb, size_t nitems) {
  arena_alloc(a, c, ne, nb, nitems);
  arena_fill(a, c);
}

WalkLaunch prepare_walk(Arena &a, int num_cus, const LaunchCfg &c, std::vector<ItemRec> &items) {
  WalkLaunch w;
  uint64_t v = 0;
  for (ItemRec &it : items) {
    it.vstart = v;
    v += (it.count + c.chunk - 1) / c.chunk * c.chunk;
  }
  if (v == 0) return w;
  w.batch = items.size() > 1;
  int bpc = 0;
  if (c.small && w.batch) throw HipFail{hipErrorInvalidValue, "a small-mesh launch walks one item"};
  hip_check(walk_occupancy(c.dim, c.f64, c.fd, c.fixed, w.batch, c.small, c.block, c.lds, &bpc), "occupancy");
  if (bpc < 1) throw HipFail{hipErrorInvalidConfiguration, "workgroup does not fit on a CU (block size / LDS tally cache)"};
  w.grid = (c.bpc > 0 ? std::min(bpc, c.bpc) : bpc) * num_cus;
  hip_check(hipMemcpyAsync(a.items.p, items.data(), items.size() * sizeof(ItemRec), hipMemcpyHostToDevice, a.stream),
            "upload");
  WalkParams &P = w.P;
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
  P.chunk = (int32_t)c.chunk;
  P.xcc_replicas = c.xcc ? 1 : 0;
  P.lds_slots = c.slots;
  P.direct_flush = c.direct ? 1 : 0;
  // a batch's items follow one another through the launch, so the cache
  // slots are handed back every 1024 passes for the items walked next
  P.cache_epoch = c.epoch > 0 ? c.epoch : (c.epoch == 0 && w.batch && c.slots > 0 ? 1024 : 0);
  P.weight0 = c.weight0;   // MC3D::weight0, chance (MC3D.hpp:160, defaults :182-183)
  P.chance = c.chance;
  P.weight0_f = (float)P.weight0;
  P.chance_f = (float)P.chance;
  P.fix_scale = 4294967296.0;
  P.mesh_elems = c.small ? c.small_elems : 0;
  return w;
}

void enqueue_walk(Arena &a, const LaunchCfg &c, const WalkLaunch &w) {
  if (w.grid == 0) return;
  hip_check(launch_fill_u64(a.counters.p, 1, 0ull, a.stream), "fill");   // dispenser
  hip_check(launch_walk(c.dim, c.f64, c.fd, c.fixed, w.batch, c.small, w.grid, c.block, c.lds, a.stream, w.P),
            "walk");
}

void check_dispenser(unsigned long long dispensed, const WalkLaunch &w) {
  // every wave leaves the walk only after drawing an id at or past vtotal,
  // so a dispenser below it means photons were left unwalked
  if (w.grid > 0 && dispensed < w.P.vtotal)
    throw HipFail{hipErrorLaunchFailure, "walk ended before its photon dispenser reached the launch's photon count"};
}

float launch_items(Arena &a, int num_cus, const LaunchCfg &c, std::vector<ItemRec> &items, int *grid_out) {
  const WalkLaunch w = prepare_walk(a, num_cus, c, items);
  if (w.grid == 0) return 0.f;
  hip_check(hipEventRecord(a.ev0, a.stream), "event");
  enqueue_walk(a, c, w);
  hip_check(hipEventRecord(a.ev1, a.stream), "event");
  unsigned long long dispensed = 0;
  hip_check(hipMemcpyAsync(&dispensed, a.counters.p, 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipStreamSynchronize(a.stream), "sync");
  check_dispenser(dispensed, w);
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a.ev0, a.ev1), "event");
  if (grid_out) *grid_out = w.grid;
  return ms;
}

void arena_reduce(Arena &a, const LaunchCfg &c) {
  const double inv = c.fixed ? 1.0 / 4294967296.0 : 1.0;
  const size_t ew = (size_t)a.ne * c.W, bw = (size_t)a.nb * 4;
  hip_check(launch_reduce(c.fixed, a.etally.p, a.raw_e.p, ew, c.nrep, ew, inv, a.stream), "reduce");
  hip_check(launch_reduce(c.fixed, a.btally.p, a.raw_b.p, bw, c.nrep, bw, inv, a.stream), "reduce");
}

void arena_copy_out(Arena &a, const LaunchCfg &c, std::vector<double> &raw_e, std::vector<double> &raw_b,
                    double led[5], unsigned long long cnt[kCounters], std::vector<unsigned long long> &loss,
                    std::vector<double> &item_led) {
  const size_t ew = (size_t)a.ne * c.W, bw = (size_t)a.nb * 4;
  const size_t ni = std::max<size_t>(1, a.nitems);
  raw_e.resize(ew);
  raw_b.resize(bw);
  loss.resize(a.nitems);
  item_led.resize(5 * ni);
  hip_check(hipMemcpyAsync(raw_e.data(), a.raw_e.p, ew * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipMemcpyAsync(raw_b.data(), a.raw_b.p, bw * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipMemcpyAsync(item_led.data(), a.ledger.p, item_led.size() * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipMemcpyAsync(cnt, a.counters.p, kCounters * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  if (a.nitems)
    hip_check(hipMemcpyAsync(loss.data(), a.item_loss.p, a.nitems * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
  hip_check(hipStreamSynchronize(a.stream), "sync");
  sum_ledgers(item_led, ni, led);
}

void arena_collect(Arena &a, const LaunchCfg &c, std::vector<double> &raw_e, std::vector<double> &raw_b,
                   double led[5], unsigned long long cnt[kCounters], std::vector<unsigned long long> &loss,
                   std::vector<double> &item_led) {
  arena_reduce(a, c);
  arena_copy_out(a, c, raw_e, raw_b, led, cnt, loss, item_led);
}

void sum_ledgers(const std::vector<double> &item_led, size_t nitems, double led[5]) {
  for (int q = 0; q < 5; ++q) led[q] = 0.0;
  for (size_t k = 0; k < nitems; ++k)
    for (int q = 0; q < 5; ++q) led[q] += item_led[5 * k + q];
}

// Item record of a problem run: its device arrays, stream key and phase.
ItemRec make_item(const vmc_problem &m, const vmc_run_desc *run, bool f64, uint64_t count, uint64_t id0,
                  int64_t tbase, int64_t bbase) {
  ItemRec it{};
  it.count = count;
  it.id0 = run->first_photon + id0;
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
  RunModel mdl;
  (void)run_model(run, &mdl);   // validated by resolve_cfg
  it.kfac = 2.0 * M_PI * run->f / mdl.c0;
  it.cph0 = cos(phase0);
  it.sph0 = sin(phase0);
  return it;
}

// Splits a problem's raw element words (device order, from slot tbase) and
// boundary words (from slot bbase) and writes normalised outputs
// (MC3D.hpp:1793-1818).
void finish_outputs(const vmc_problem &m, const vmc_run_desc &run, int64_t N, int W, const double *raw_e,
                    const double *raw_b, vmc_output *out) {
  if (!out || (!out->ER && !out->EI && !out->EBR && !out->EBI && !out->DEBR && !out->DEBI)) return;
  RunModel mdl;
  (void)run_model(&run, &mdl);
  normalise_into(m, 2.0 * M_PI * run.f, mdl.c0, N, W, raw_e, raw_b, out);
}

int64_t sector_round(int64_t n) { return (n + kSectorElems - 1) / kSectorElems * kSectorElems; }

namespace {
Arena g_batch_arena;   // tallies of vmc_batch_run (one batch at a time)
std::mutex g_batch_mutex;
} // namespace

// Coordinate descent over the launch knobs from `base`: each knob's values
// are timed with the others fixed (`time_of` returns kernel milliseconds,
// +inf when the device refuses), and a knob moves only for a gain above 2 %.
// The first timing of the base configuration only warms up (arena
// allocation, code loading).
// (several: a batch of several items, whose FP32 default block is 512)
template <class TimeOf>
int autotune_knobs(const vmc_run_desc &base, bool several, vmc_run_desc *tuned, int32_t *trials, TimeOf time_of) {
  int count = 0;
  auto timed = [&](const vmc_run_desc &c) {
    ++count;
    return time_of(c);
  };
  vmc_run_desc best = base;
  const bool f64 = base.precision == VMC_FP64;
  // defaults spelled out, so a knob's default value is not timed twice
  if (best.block_size == 0) best.block_size = f64 ? 256 : (several ? 512 : 1024);
  if (best.refill_min == 0) best.refill_min = default_refill(2.0 * M_PI * base.f > 0.0);
  if (best.chunk == 0) best.chunk = (int32_t)kChunk;
  if (!std::isfinite(timed(best))) return VMC_ERR_INVALID_ARGUMENT;
  double tb = timed(best);
  struct Knob {
    int32_t vmc_run_desc::*field;
    std::vector<int32_t> values;
  };
  const std::vector<Knob> knobs = {
      {&vmc_run_desc::block_size, f64 ? std::vector<int32_t>{256, 128, 64} : std::vector<int32_t>{1024, 512, 256}},
      {&vmc_run_desc::refill_min, {6, 10, 16, 24, 32}},
      {&vmc_run_desc::chunk, {128, 64, 256, 512}},
      {&vmc_run_desc::replicas, {0, 8, 1}},
      {&vmc_run_desc::lds_slots, {0, -1}},
      {&vmc_run_desc::cache_epoch, {0, -1, 256, 4096}},
  };
  for (const Knob &k : knobs) {
    for (int32_t v : k.values) {
      if (best.*(k.field) == v) continue;
      vmc_run_desc cand = best;
      cand.*(k.field) = v;
      const double t = timed(cand);
      if (t < 0.98 * tb) {
        best = cand;
        tb = t;
      }
    }
  }
  *tuned = best;
  if (trials) *trials = count;
  return VMC_OK;
}

// One (dimension, CW/FD) group of a batch: its launch settings, the batch
// positions of its items and their element / boundary offsets in the arena.
struct BatchGroup {
  LaunchCfg c;
  std::vector<int32_t> idx;
  std::vector<int64_t> tb, bb;
  int64_t ne = 0, nb = 0;
};

// Checks a batch (problems on one device, one precision and tally mode),
// splits it into groups in launch order -- one per (dimension, CW/FD,
// roulette constants) -- with the settings of items[0].run, and uploads the
// problems' element records; a status.
int plan_batch(const vmc_batch_item *items, int32_t n, std::vector<BatchGroup> &groups) {
  if (!items || n <= 0 || !items[0].prob) return VMC_ERR_INVALID_ARGUMENT;
  const vmc_problem *m0 = items[0].prob;
  std::vector<RunModel> model((size_t)n);
  for (int32_t i = 0; i < n; ++i) {
    const vmc_batch_item &b = items[i];
    if (!b.prob || b.run.Nphoton < 0 || b.prob->device != m0->device) return VMC_ERR_INVALID_ARGUMENT;
    if (b.run.precision != items[0].run.precision || b.run.tally != items[0].run.tally)
      return VMC_ERR_INVALID_ARGUMENT;
    if (b.run.tally == VMC_TALLY_FIXED64 && b.run.Nphoton >= kFixedMaxPhotons) return VMC_ERR_INVALID_ARGUMENT;
    if (run_model(&b.run, &model[i]) != VMC_OK) return VMC_ERR_INVALID_ARGUMENT;
  }
  try {
    hip_check(hipSetDevice(m0->device), "hipSetDevice");
    for (int dim = 2; dim <= 3; ++dim) {
      for (int fdi = 0; fdi < 2; ++fdi) {
        auto in_class = [&](int32_t i) {
          return items[i].prob->dim == dim && ((2.0 * M_PI * items[i].run.f > 0.0) == (fdi == 1));
        };
        // the roulette constants are launch-wide: one group per pair, in
        // order of first appearance
        std::vector<std::pair<double, double>> pairs;
        for (int32_t i = 0; i < n; ++i) {
          const std::pair<double, double> pr(model[i].weight0, model[i].chance);
          if (in_class(i) && std::find(pairs.begin(), pairs.end(), pr) == pairs.end()) pairs.push_back(pr);
        }
        for (const auto &pr : pairs) {
          BatchGroup g;
          for (int32_t i = 0; i < n; ++i)
            if (in_class(i) && model[i].weight0 == pr.first && model[i].chance == pr.second) g.idx.push_back(i);
          const int st = resolve_cfg(dim, m0->num_cus, m0->lds_per_cu, fdi == 1, &items[0].run, &g.c, g.idx.size() > 1);
          if (st != VMC_OK) return st;
          g.c.weight0 = pr.first;
          g.c.chance = pr.second;
          for (int32_t i : g.idx) {
            vmc_problem &m = *items[i].prob;
            upload_params(m, g.c.f64);
            g.tb.push_back(g.ne);
            g.bb.push_back(g.nb);
            g.ne += m.md->slots;
            g.nb += m.Nb;
          }
          if (g.ne > ((int64_t)1 << 30) / g.c.W || g.nb > ((int64_t)1 << 29)) return VMC_ERR_INVALID_ARGUMENT;
          choose_replicas(g.c, &items[0].run, g.ne, g.nb);
          groups.push_back(std::move(g));
        }
      }
    }
  } catch (HipFail f) {
    return fail_hip(f);
  }
  return VMC_OK;
}

// Normalised outputs and loss counts of a group's items from its raw words;
// adds the group's counters and ledger to *tot.
void finish_group(vmc_batch_item *items, const BatchGroup &g, const std::vector<double> &raw_e,
                  const std::vector<double> &raw_b, const double led[5], const unsigned long long cnt[kCounters],
                  const std::vector<unsigned long long> &loss, const std::vector<double> &item_led, vmc_stats *tot) {
  for (size_t k = 0; k < g.idx.size(); ++k) {
    vmc_batch_item &b = items[g.idx[k]];
    finish_outputs(*b.prob, b.run, b.run.Nphoton, g.c.W, raw_e.data() + g.tb[k] * g.c.W,
                   raw_b.data() + g.bb[k] * 4, &b.out);
    b.loss = (int64_t)loss[k];
    for (int q = 0; q < 5; ++q) b.ledger[q] = item_led[5 * k + q];
    tot->photons += b.run.Nphoton;
  }
  tot->steps += (int64_t)cnt[1];
  tot->loss += (int64_t)cnt[2];
  tot->rechecks += (int64_t)cnt[3];
  for (int q = 0; q < 5; ++q) tot->ledger[q] += led[q];
  tot->block_size = g.c.block;
}


// ---------------------------------------------------------------------------
// One problem's MonteCarlo() (vmc_problem_run, vmc_problem_run_noalloc)
// ---------------------------------------------------------------------------

namespace {
// the scratch of a GPU is shared by every allocation-free run on it: the
// runs of this process take turns
std::mutex g_scratch_mu[kMaxDevices];
} // namespace

// Lays the arena of a run over ne element and nb boundary slots out in the
// device's scratch (no allocation): the replicas, the reduced words in
// place of replica 0 (the replica sums read every replica of a word before
// writing it), ledger, counters, the item's loss count and its record.
// One replica per XCD falls back to a single one when the copies do not
// fit; the borrowed stream and events are the problem's.  A status.
int scratch_arena(const vmc_problem &m, const vmc_run_desc *run, LaunchCfg &c, int64_t ne, int64_t nb, Arena &a) {
  void *base = nullptr;
  hip_check(scratch_address(m.device, &base), "hipGetSymbolAddress(scratch)");
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  auto need = [&](int nrep) {
    return up((size_t)ne * c.W * nrep * 8) + up((size_t)nb * 4 * nrep * 8) + 3 * 256 + up(sizeof(ItemRec));
  };
  if (need(c.nrep) > kScratchBytes && run->replicas <= 0 && c.nrep > 1) {
    c.nrep = 1;
    c.xcc = false;
  }
  if (need(c.nrep) > kScratchBytes) {
    g_last_error = "run needs " + std::to_string(need(c.nrep)) + " bytes of tallies; the allocation-free scratch holds " +
                   std::to_string(kScratchBytes);
    return VMC_ERR_OUT_OF_MEMORY;
  }
  unsigned char *p = static_cast<unsigned char *>(base);
  auto carve = [&](size_t bytes) {
    unsigned char *q = p;
    p += up(bytes);
    return q;
  };
  a.etally.p = reinterpret_cast<unsigned long long *>(carve((size_t)ne * c.W * c.nrep * 8));
  a.etally.n = (size_t)ne * c.W * c.nrep;
  a.btally.p = reinterpret_cast<unsigned long long *>(carve((size_t)nb * 4 * c.nrep * 8));
  a.btally.n = (size_t)nb * 4 * c.nrep;
  a.raw_e.p = reinterpret_cast<double *>(a.etally.p);
  a.raw_e.n = (size_t)ne * c.W;
  a.raw_b.p = reinterpret_cast<double *>(a.btally.p);
  a.raw_b.n = (size_t)nb * 4;
  a.ledger.p = reinterpret_cast<double *>(carve(8 * 8));
  a.ledger.n = 8;
  a.counters.p = reinterpret_cast<unsigned long long *>(carve(kCounters * 8));
  a.counters.n = kCounters;
  a.item_loss.p = reinterpret_cast<unsigned long long *>(carve(8));
  a.item_loss.n = 1;
  a.items.p = reinterpret_cast<ItemRec *>(carve(sizeof(ItemRec)));
  a.items.n = 1;
  a.stream = m.arena.stream;
  a.ev0 = m.arena.ev0;
  a.ev1 = m.arena.ev1;
  return VMC_OK;
}

int problem_run(vmc_problem *m, const vmc_run_desc *run, vmc_output *out, vmc_stats *stats,
                vmc_progress_fn progress, void *user, bool use_scratch) {
  if (!m || !run || run->Nphoton < 0) return VMC_ERR_INVALID_ARGUMENT;
  const auto t_start = std::chrono::steady_clock::now();
  const bool fd = 2.0 * M_PI * run->f > 0.0;
  LaunchCfg c;
  const int st = resolve_cfg(m->dim, m->num_cus, m->lds_per_cu, fd, run, &c, false);
  if (st != VMC_OK) return st;
  const int64_t N = run->Nphoton;
  if (c.fixed && N >= kFixedMaxPhotons) return VMC_ERR_INVALID_ARGUMENT;
  if (use_scratch && (c.f64 ? m->med->elems64.p : m->med->elems32.p) == nullptr) {
    // the records are allocated by the first run of a precision; refreshing
    // them later (a new medium) reuses that allocation
    g_last_error = "element records of this precision are not on the device yet: run vmc_problem_run once first";
    return VMC_ERR_INVALID_ARGUMENT;
  }
  int aborted = 0, grid = 0;
  float kernel_ms = 0.f;
  double led[5] = {0, 0, 0, 0, 0};
  unsigned long long cnt[kCounters] = {};
  // host staging of the raw words: kept per thread on the allocation-free
  // path, so repeated runs reuse it
  std::vector<double> own_e, own_b;
  std::vector<unsigned long long> own_loss;
  thread_local std::vector<double> keep_e, keep_b;
  thread_local std::vector<unsigned long long> keep_loss;
  std::vector<double> &raw_e = use_scratch ? keep_e : own_e, &raw_b = use_scratch ? keep_b : own_b;
  std::vector<unsigned long long> &loss = use_scratch ? keep_loss : own_loss;
  thread_local std::vector<double> keep_led;
  std::vector<double> own_led;
  std::vector<double> &item_led = use_scratch ? keep_led : own_led;
  const int64_t ne = m->md->slots, nb = m->Nb;
  try {
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    upload_params(*m, c.f64);
    choose_replicas(c, run, ne, nb);
    choose_small(c, run, ne, m->lds_per_cu);
    Arena view;
    std::unique_lock<std::mutex> turn;
    if (use_scratch) {
      if (m->device < 0 || m->device >= kMaxDevices) return VMC_ERR_INVALID_ARGUMENT;
      turn = std::unique_lock<std::mutex>(g_scratch_mu[m->device]);
      const int sst = scratch_arena(*m, run, c, ne, nb, view);
      if (sst != VMC_OK) return sst;
    }
    Arena &a = use_scratch ? view : m->arena;
    arena_reset(a, c, ne, nb, 1);
    // photon ranges: one launch, or ~20 when a progress callback wants ticks
    const int nseg = progress ? 20 : 1;
    int64_t done = 0;
    for (int sgm = 0; sgm < nseg && done < N; ++sgm) {
      const int64_t end = (sgm == nseg - 1) ? N : (N * (sgm + 1)) / nseg;
      if (end <= done) continue;
      std::vector<ItemRec> items{make_item(*m, run, c.f64, (uint64_t)(end - done), (uint64_t)done, 0, 0)};
      kernel_ms += launch_items(a, m->num_cus, c, items, &grid);
      done = end;
      if (progress && !progress(100.0 * (double)done / (double)(N > 0 ? N : 1), user)) {
        aborted = 1;
        break;
      }
    }
    arena_collect(a, c, raw_e, raw_b, led, cnt, loss, item_led);
  } catch (HipFail f) {
    return fail_hip(f);
  }
  // like the reference, an aborted run still divides by the planned N
  finish_outputs(*m, *run, N, c.W, raw_e.data(), raw_b.data(), out);
  const bool out_of_range = c.fixed && cnt[4] != 0;
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
    stats->lds_mesh = c.small ? 1 : 0;
    stats->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  }
  if (out_of_range) return range_error();
  return aborted ? VMC_ERR_ABORTED : VMC_OK;
}

} // namespace vmc

// A batch recorded as a HIP graph (vmc_batch_plan_*): per group an arena on
// the plan's stream, the items' records in device order and the launch.
struct vmc_batch_plan {
  struct Group {
    vmc::BatchGroup g;
    vmc::Arena arena;
    std::vector<vmc::ItemRec> recs;
    vmc::WalkLaunch w;
  };
  int device = 0, num_cus = 0;
  int32_t precision = 0, tally = 0;
  std::vector<vmc_problem *> probs;
  std::vector<int64_t> counts;
  std::vector<bool> fd;
  std::vector<vmc::RunModel> model;
  std::vector<std::unique_ptr<Group>> groups;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~vmc_batch_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &pg : groups) {
      pg->arena.stream = nullptr;   // the plan's, destroyed below
      pg->arena.close();
    }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

using namespace vmc;

extern "C" {

int vmc_problem_check(const vmc_problem_desc *d, int64_t *HN_out, int64_t *nsources) try {
  if (!d) return VMC_ERR_INVALID_ARGUMENT;
  vmc_problem m;
  const int st = host_prepare(d, &m);
  if (st != VMC_OK) return st;
  if (HN_out) std::memcpy(HN_out, m.HN.data(), sizeof(int64_t) * m.HN.size());
  if (nsources) *nsources = (int64_t)m.src.size();
  return VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_build_neighbours(int32_t dim, int64_t Ne, int64_t Nb, int64_t Nr, const int64_t *H, const int64_t *BH,
                         int64_t *HN_out) try {
  return neighbours_only(dim, Ne, Nb, Nr, H, BH, HN_out);
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_sources(const vmc_problem_desc *d, int64_t *faces, int64_t *mothers, double *cdf,
                        int64_t capacity, int64_t *nsources) try {
  if (!d || !nsources || capacity < 0) return VMC_ERR_INVALID_ARGUMENT;
  vmc_problem m;
  const int st = host_prepare(d, &m);
  if (st != VMC_OK) return st;
  const int64_t ns = (int64_t)m.lsFace.size(), k = std::min(ns, capacity);
  if (faces) std::copy(m.lsFace.begin(), m.lsFace.begin() + k, faces);
  if (mothers) std::copy(m.lsMother.begin(), m.lsMother.begin() + k, mothers);
  if (cdf) std::copy(m.lsCDF.begin(), m.lsCDF.begin() + k, cdf);
  *nsources = ns;
  return VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_create(const vmc_problem_desc *d, vmc_problem **out) try {
  if (!d || !out) return VMC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  // owned here until it is handed out, so a host allocation failure (which
  // leaves through the catch below) frees it
  std::unique_ptr<vmc_problem> own(new vmc_problem());
  vmc_problem *m = own.get();
  const int st = host_prepare(d, m);
  if (st != VMC_OK) return st;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
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
    const int err = fail_hip(f);
    vmc_problem_destroy(own.release());
    return err;
  }
  *out = own.release();
  return VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_create_variant(const vmc_problem *base, const vmc_variant_desc *d, vmc_problem **out) try {
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
  std::unique_ptr<vmc_problem> own(new vmc_problem(base->md, med));   // freed on any early return
  vmc_problem *m = own.get();
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
    vmc_problem_destroy(own.release());
    return st;
  }
  *out = own.release();
  return VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_set_medium(vmc_problem *m, const double *mua, const double *mus, const double *g,
                           const double *n) try {
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
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_get_hn(const vmc_problem *m, int64_t *HN) try {
  if (!m || !HN) return VMC_ERR_INVALID_ARGUMENT;
  std::memcpy(HN, m->HN.data(), sizeof(int64_t) * m->HN.size());
  return VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_slots(const vmc_problem *m, int64_t *slots) {
  if (!m || !slots) return VMC_ERR_INVALID_ARGUMENT;
  *slots = m->md->slots;
  return VMC_OK;
}

int vmc_problem_run(vmc_problem *m, const vmc_run_desc *run, vmc_output *out, vmc_stats *stats,
                    vmc_progress_fn progress, void *user) try {
  return problem_run(m, run, out, stats, progress, user, false);
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_run_noalloc(vmc_problem *m, const vmc_run_desc *run, vmc_output *out, vmc_stats *stats,
                            vmc_progress_fn progress, void *user) try {
  return problem_run(m, run, out, stats, progress, user, true);
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_problem_autotune(vmc_problem *m, const vmc_run_desc *run, vmc_run_desc *tuned, int64_t trial_photons,
                         int32_t *trials) try {
  if (!m || !run || !tuned || run->Nphoton < 0 || trial_photons < 0) return VMC_ERR_INVALID_ARGUMENT;
  {
    LaunchCfg c;
    const int st = resolve_cfg(m->dim, m->num_cus, m->lds_per_cu, 2.0 * M_PI * run->f > 0.0, run, &c, false);
    if (st != VMC_OK) return st;
  }
  const int64_t n = trial_photons > 0 ? trial_photons
                                      : std::min<int64_t>(10000000, std::max<int64_t>(100000, run->Nphoton / 50));
  vmc_output none{};
  const int st = autotune_knobs(*run, false, tuned, trials, [&](vmc_run_desc cand) {
    cand.Nphoton = n;
    vmc_stats s{};
    return vmc_problem_run(m, &cand, &none, &s, nullptr, nullptr) == VMC_OK ? s.kernel_ms
                                                                           : std::numeric_limits<double>::infinity();
  });
  tuned->Nphoton = run->Nphoton;
  return st;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_batch_autotune(const vmc_batch_item *items, int32_t n, double trial_fraction, vmc_run_desc *tuned,
                       int32_t *trials) try {
  if (!items || n <= 0 || !tuned || !(trial_fraction >= 0.0 && trial_fraction <= 1.0)) return VMC_ERR_INVALID_ARGUMENT;
  const double frac = trial_fraction > 0.0 ? trial_fraction : 0.02;
  // trial batches: the same items with a fraction of their photons and no
  // outputs
  std::vector<vmc_batch_item> trial(items, items + n);
  for (vmc_batch_item &b : trial) {
    if (!b.prob) return VMC_ERR_INVALID_ARGUMENT;
    if (b.run.Nphoton > 0) b.run.Nphoton = std::max<int64_t>(1, (int64_t)(frac * (double)b.run.Nphoton));
    b.out = vmc_output{};
  }
  return autotune_knobs(items[0].run, n > 1, tuned, trials, [&](vmc_run_desc cand) {
    // the candidate's launch knobs on every item; each item keeps its own
    // photons, stream, phase and model constants
    for (vmc_batch_item &b : trial) {
      b.run.block_size = cand.block_size;
      b.run.blocks_per_cu = cand.blocks_per_cu;
      b.run.replicas = cand.replicas;
      b.run.refill_min = cand.refill_min;
      b.run.lds_slots = cand.lds_slots;
      b.run.direct_flush = cand.direct_flush;
      b.run.chunk = cand.chunk;
      b.run.cache_epoch = cand.cache_epoch;
    }
    vmc_stats s{};
    return vmc_batch_run(trial.data(), n, &s) == VMC_OK ? s.kernel_ms : std::numeric_limits<double>::infinity();
  });
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_batch_run(vmc_batch_item *items, int32_t n, vmc_stats *stats) try {
  const auto t_start = std::chrono::steady_clock::now();
  std::vector<BatchGroup> groups;
  {
    const int st = plan_batch(items, n, groups);
    if (st != VMC_OK) return st;
  }
  for (int32_t i = 0; i < n; ++i) {
    items[i].loss = 0;
    items[i].status = VMC_OK;
  }
  vmc_problem *m0 = items[0].prob;
  vmc_stats tot{};
  float kernel_ms = 0.f;
  bool range_bad = false;
  try {
    std::lock_guard<std::mutex> lock(g_batch_mutex);
    if (!g_batch_arena.stream) g_batch_arena.open();
    for (const BatchGroup &g : groups) {
      arena_reset(g_batch_arena, g.c, g.ne, g.nb, g.idx.size());
      std::vector<ItemRec> recs;
      for (size_t k = 0; k < g.idx.size(); ++k) {
        const vmc_batch_item &b = items[g.idx[k]];
        recs.push_back(make_item(*b.prob, &b.run, g.c.f64, (uint64_t)b.run.Nphoton, 0, g.tb[k], g.bb[k]));
      }
      int grid = 0;
      kernel_ms += launch_items(g_batch_arena, m0->num_cus, g.c, recs, &grid);
      std::vector<double> raw_e, raw_b;
      std::vector<unsigned long long> loss;
      std::vector<double> item_led;
      double led[5];
      unsigned long long cnt[kCounters];
      arena_collect(g_batch_arena, g.c, raw_e, raw_b, led, cnt, loss, item_led);
      finish_group(items, g, raw_e, raw_b, led, cnt, loss, item_led, &tot);
      tot.grid_blocks = grid;
      if (g.c.fixed && cnt[4]) {   // the launch's items left the fixed-point range
        range_bad = true;
        for (int32_t i : g.idx) items[i].status = VMC_ERR_TALLY_RANGE;
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
  return range_bad ? range_error() : VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_batch_plan_create(const vmc_batch_item *items, int32_t n, vmc_batch_plan **out) try {
  if (!out) return VMC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  std::vector<BatchGroup> groups;
  {
    const int st = plan_batch(items, n, groups);
    if (st != VMC_OK) return st;
  }
  std::unique_ptr<vmc_batch_plan> pl(new vmc_batch_plan);
  pl->device = items[0].prob->device;
  pl->num_cus = items[0].prob->num_cus;
  pl->precision = items[0].run.precision;
  pl->tally = items[0].run.tally;
  for (int32_t i = 0; i < n; ++i) {
    pl->probs.push_back(items[i].prob);
    pl->counts.push_back(items[i].run.Nphoton);
    pl->fd.push_back(2.0 * M_PI * items[i].run.f > 0.0);
    RunModel m;
    (void)run_model(&items[i].run, &m);   // validated by plan_batch
    pl->model.push_back(m);
  }
  try {
    hip_check(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&pl->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&pl->ev1), "hipEventCreate");
    for (BatchGroup &g : groups) {
      std::unique_ptr<vmc_batch_plan::Group> pg(new vmc_batch_plan::Group);
      pg->g = g;
      pg->arena.stream = pl->stream;   // the plan's stream, not the arena's own
      arena_alloc(pg->arena, g.c, g.ne, g.nb, g.idx.size());
      for (size_t k = 0; k < g.idx.size(); ++k) {
        const vmc_batch_item &b = items[g.idx[k]];
        pg->recs.push_back(make_item(*b.prob, &b.run, g.c.f64, (uint64_t)b.run.Nphoton, 0, g.tb[k], g.bb[k]));
      }
      pg->w = prepare_walk(pg->arena, pl->num_cus, g.c, pg->recs);
      pl->groups.push_back(std::move(pg));
    }
    hip_check(hipStreamSynchronize(pl->stream), "sync");
    // the device work of a run, recorded once: fill kernels (no memset
    // nodes), the walk launches and the replica sums
    hip_check(hipStreamBeginCapture(pl->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      for (auto &pg : pl->groups) {
        arena_fill(pg->arena, pg->g.c);
        enqueue_walk(pg->arena, pg->g.c, pg->w);
        arena_reduce(pg->arena, pg->g.c);
      }
    } catch (HipFail) {
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(pl->stream, &partial);
      if (partial) (void)hipGraphDestroy(partial);
      throw;
    }
    hip_check(hipStreamEndCapture(pl->stream, &pl->graph), "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  } catch (HipFail f) {
    return fail_hip(f);
  }
  *out = pl.release();
  return VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_batch_plan_run(vmc_batch_plan *pl, vmc_batch_item *items, int32_t n, vmc_stats *stats) try {
  if (!pl || !items || n != (int32_t)pl->probs.size()) return VMC_ERR_INVALID_ARGUMENT;
  for (int32_t i = 0; i < n; ++i) {
    const vmc_run_desc &r = items[i].run;
    RunModel m;
    if (items[i].prob != pl->probs[i] || r.Nphoton != pl->counts[i] || r.precision != pl->precision ||
        r.tally != pl->tally || (2.0 * M_PI * r.f > 0.0) != pl->fd[i] || run_model(&r, &m) != VMC_OK ||
        m.weight0 != pl->model[i].weight0 || m.chance != pl->model[i].chance)
      return VMC_ERR_INVALID_ARGUMENT;
    items[i].loss = 0;
    items[i].status = VMC_OK;
  }
  const auto t_start = std::chrono::steady_clock::now();
  vmc_stats tot{};
  float kernel_ms = 0.f;
  bool range_bad = false;
  try {
    hip_check(hipSetDevice(pl->device), "hipSetDevice");
    // current records: seeds, tags, phases and the problems' current media
    for (auto &pg : pl->groups) {
      const BatchGroup &g = pg->g;
      for (size_t k = 0; k < g.idx.size(); ++k) {
        const vmc_batch_item &b = items[g.idx[k]];
        upload_params(*b.prob, g.c.f64);
        ItemRec it = make_item(*b.prob, &b.run, g.c.f64, (uint64_t)b.run.Nphoton, 0, g.tb[k], g.bb[k]);
        it.vstart = pg->recs[k].vstart;
        pg->recs[k] = it;
      }
      hip_check(hipMemcpyAsync(pg->arena.items.p, pg->recs.data(), pg->recs.size() * sizeof(ItemRec),
                               hipMemcpyHostToDevice, pl->stream),
                "upload");
    }
    hip_check(hipEventRecord(pl->ev0, pl->stream), "event");
    hip_check(hipGraphLaunch(pl->exec, pl->stream), "hipGraphLaunch");
    hip_check(hipEventRecord(pl->ev1, pl->stream), "event");
    hip_check(hipEventSynchronize(pl->ev1), "sync");
    hip_check(hipEventElapsedTime(&kernel_ms, pl->ev0, pl->ev1), "event");
    for (auto &pg : pl->groups) {
      std::vector<double> raw_e, raw_b;
      std::vector<unsigned long long> loss;
      std::vector<double> item_led;
      double led[5];
      unsigned long long cnt[kCounters];
      arena_copy_out(pg->arena, pg->g.c, raw_e, raw_b, led, cnt, loss, item_led);
      check_dispenser(cnt[0], pg->w);
      finish_group(items, pg->g, raw_e, raw_b, led, cnt, loss, item_led, &tot);
      tot.grid_blocks = pg->w.grid;
      if (pg->g.c.fixed && cnt[4]) {
        range_bad = true;
        for (int32_t i : pg->g.idx) items[i].status = VMC_ERR_TALLY_RANGE;
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
  return range_bad ? range_error() : VMC_OK;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

void vmc_batch_plan_destroy(vmc_batch_plan *pl) { delete pl; }

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
                    double *DEBI, int64_t *HN_out, int64_t *loss) try {
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
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
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
  case VMC_ERR_OUT_OF_MEMORY: return "Out of memory (device, or host: see vmc_last_error)";
  case VMC_ERR_NO_DEVICE: return "No GPU device";
  case VMC_ERR_ABORTED: return "Computation aborted by the progress callback";
  case VMC_ERR_TALLY_RANGE: return "A fixed-point tally word left its exact range";
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
