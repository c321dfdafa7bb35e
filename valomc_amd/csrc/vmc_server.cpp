// Serving loop of the C ABI (vmc_server_*).
#include "vmc_host.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

using namespace vmc;

// ---------------------------------------------------------------------------
// Serving loop: continuous batching of forward solves
//
// A worker thread owns one stream and a persistent tally arena.  Requests
// wait in FIFO order until their tallies fit in the arena (admission by
// device-memory footprint); each launch then takes one group (dimension,
// CW/FD, roulette constants) of the admitted requests and advances them by parts of the
// launch's photon budget, so a request that arrives while others run joins
// at the next launch and launch times stay bounded.  Fair policy: groups in
// turn, equal shares.  Shortest policy: the group of the request with the
// least estimated remaining device time, whose requests take the budget in
// order of remaining time (each problem's device time per photon is
// measured from the launches it took part in).  A
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
  RunModel model;                // roulette constants and c0 of the run
  int64_t launched = 0;          // photons handed to launches so far
  int64_t eoff = -1, ewords = 0; // element words in each arena replica
  int64_t boff = -1;             // first boundary slot
  int64_t loss = 0;
  double ledger[5] = {0, 0, 0, 0, 0};   // absorbed, exited, killed, gained, lost over its launches
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
  // device ms per photon of each problem, from the launches it took part in
  // (shortest policy; a scheduling estimate only)
  std::map<const vmc_problem *, double> cost;
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
  finish_outputs(*r->prob, r->run, r->run.Nphoton, r->fd ? 2 : 1, raw_e.data(), raw_b.data(), &r->out);
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
      } catch (const std::bad_alloc &) {
        r->status = host_out_of_memory();
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

// Requests that can share a launch: same dimension, CW/FD mode and
// roulette constants (launch-wide kernel parameters).
bool same_group(const Request *a, const Request *b) {
  return a->prob->dim == b->prob->dim && a->fd == b->fd && a->model.weight0 == b->model.weight0 &&
         a->model.chance == b->model.chance;
}

// Shortest policy: the active requests with photons left, of the group of
// the one with the least estimated remaining device time, in order of that
// estimate (ties: earlier ticket).  A problem never launched yet is
// estimated at the least known cost per photon of its (dimension, CW/FD)
// class, or 0, so a new request is measured in its first launch.  Called
// under the server's lock.
void pick_shortest(vmc_server *s, std::vector<Request *> &group) {
  struct Est {
    double t;
    Request *r;
  };
  std::vector<Est> all;
  double gmin[4] = {-1, -1, -1, -1};
  auto gid = [](const Request *r) { return (r->prob->dim == 3 ? 0 : 2) + (r->fd ? 1 : 0); };
  for (Request *r : s->active) {
    auto it = s->cost.find(r->prob);
    if (it == s->cost.end()) continue;
    double &g = gmin[gid(r)];
    g = g < 0 ? it->second : std::min(g, it->second);
  }
  for (Request *r : s->active) {
    const int64_t left = r->run.Nphoton - r->launched;
    if (left <= 0) continue;
    auto it = s->cost.find(r->prob);
    const double c = it != s->cost.end() ? it->second : std::max(0.0, gmin[gid(r)]);
    all.push_back({c * (double)left, r});
  }
  if (all.empty()) return;
  std::sort(all.begin(), all.end(), [](const Est &a, const Est &b) {
    return a.t != b.t ? a.t < b.t : a.r->ticket < b.r->ticket;
  });
  for (const Est &e : all)
    if (same_group(e.r, all[0].r)) group.push_back(e.r);
}

// Fair policy: the groups with photons left take turns, in order of their
// oldest request; all of the chosen group's requests share the launch.
// Called under the server's lock.
void pick_fair(vmc_server *s, std::vector<Request *> &group) {
  std::vector<Request *> heads;   // one request per group, oldest first
  for (Request *r : s->active) {
    if (r->launched >= r->run.Nphoton) continue;
    bool seen = false;
    for (Request *h : heads) seen = seen || same_group(h, r);
    if (!seen) heads.push_back(r);
  }
  if (heads.empty()) return;
  Request *h = heads[(size_t)s->turn % heads.size()];
  s->turn = (int)(((size_t)s->turn + 1) % heads.size());
  for (Request *r : s->active)
    if (same_group(r, h) && r->launched < r->run.Nphoton) group.push_back(r);
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
      if (s->desc.policy == VMC_SERVE_SHORTEST) pick_shortest(s, group);
      else pick_fair(s, group);
      if (!group.empty()) {
        gdim = group[0]->prob->dim;
        gfd = group[0]->fd;
      }
    }
    LaunchCfg c;
    vmc_run_desc tmpl{};
    tmpl.precision = s->desc.precision;
    tmpl.tally = s->desc.tally;
    std::vector<Request *> done_now;
    // a HIP failure or a failed host allocation ends the launch's requests
    auto fail_group = [&](int st) {
      std::lock_guard<std::mutex> lk(s->mu);
      for (Request *r : group) {
        r->status = st;
        if (std::find(done_now.begin(), done_now.end(), r) == done_now.end()) done_now.push_back(r);
      }
    };
    try {
      Arena &a = s->arena;
      for (Request *r : admitted) {
        hip_check(launch_fill_ranges(a.etally.p + r->eoff, (size_t)r->ewords, s->nrep, (size_t)s->cap_words, 0ull,
                                     a.stream), "fill");
        hip_check(launch_fill_ranges(a.btally.p + r->boff * 4, (size_t)r->prob->Nb * 4, s->nrep,
                                     (size_t)s->cap_b * 4, 0ull, a.stream), "fill");
      }
      if (!group.empty()) {
        const int st = resolve_cfg(gdim, s->num_cus, s->lds_per_cu, gfd, &tmpl, &c, true);
        if (st != VMC_OK) throw HipFail{hipErrorInvalidValue, "server launch configuration"};
        c.weight0 = group[0]->model.weight0;
        c.chance = group[0]->model.chance;
        // fair: equal shares of the launch budget; shortest: the budget in
        // the group's order (least remaining time first), at least
        // min_slice for the first
        const int64_t share = std::max(s->desc.min_slice, s->desc.launch_photons / (int64_t)group.size());
        int64_t budget = std::max(s->desc.min_slice, s->desc.launch_photons);
        std::vector<ItemRec> items;
        std::vector<int64_t> slice;
        for (Request *r : group) {
          const int64_t left = r->run.Nphoton - r->launched;
          int64_t n = std::min(share, left);
          if (s->desc.policy == VMC_SERVE_SHORTEST) {
            if (budget <= 0) break;
            n = std::min(budget, left);
            budget -= n;
          }
          upload_params(*r->prob, c.f64);
          items.push_back(make_item(*r->prob, &r->run, c.f64, (uint64_t)n, (uint64_t)r->launched, r->eoff / c.W,
                                    r->boff));
          slice.push_back(n);
        }
        group.resize(items.size());   // the shortest policy may leave requests out
        if (items.size() == 1) {
          // one request walks alone: the single-item kernel's launch shape
          if (resolve_cfg(gdim, s->num_cus, s->lds_per_cu, gfd, &tmpl, &c, false) != VMC_OK)
            throw HipFail{hipErrorInvalidValue, "server launch configuration"};
          c.weight0 = group[0]->model.weight0;
          c.chance = group[0]->model.chance;
        }
        a.ne = s->cap_words / c.W;
        a.nb = s->cap_b;
        a.nitems = items.size();
        a.items.ensure(items.size());
        a.item_loss.ensure(items.size());
        a.ledger.ensure(ledger_words(items.size()));
        hip_check(launch_fill_u64(a.item_loss.p, items.size(), 0ull, a.stream), "fill");
        hip_check(launch_fill_u64(reinterpret_cast<unsigned long long *>(a.ledger.p), ledger_words(items.size()), 0ull,
                                  a.stream), "fill");
        int grid = 0;
        const float ms = launch_items(a, s->num_cus, c, items, &grid);
        std::vector<unsigned long long> loss(items.size());
        std::vector<double> item_led(5 * items.size());
        unsigned long long range_flag = 0;
        hip_check(hipMemcpyAsync(loss.data(), a.item_loss.p, loss.size() * 8, hipMemcpyDeviceToHost, a.stream), "d2h");
        hip_check(hipMemcpyAsync(item_led.data(), a.ledger.p, item_led.size() * 8, hipMemcpyDeviceToHost, a.stream),
                  "d2h");
        hip_check(hipMemcpyAsync(&range_flag, a.counters.p + 4, 8, hipMemcpyDeviceToHost, a.stream), "d2h");
        hip_check(hipStreamSynchronize(a.stream), "sync");
        if (c.fixed && range_flag) {
          // this launch's requests are out of the fixed-point range; the
          // flag is cleared for the launches that follow
          hip_check(launch_fill_u64(a.counters.p + 4, 1, 0ull, a.stream), "fill");
          (void)range_error();
          std::lock_guard<std::mutex> lk(s->mu);
          for (Request *r : group) r->status = VMC_ERR_TALLY_RANGE;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        s->launches++;
        s->busy_ms += ms;
        int64_t nl = 0;
        for (int64_t n : slice) nl += n;
        if (s->cost.size() > 4096) s->cost.clear();   // estimates of problems long gone
        if (nl > 0)
          for (Request *r : group) {
            double &est = s->cost[r->prob];
            est = est > 0 ? 0.5 * est + 0.5 * ms / (double)nl : ms / (double)nl;
          }
        for (size_t k = 0; k < group.size(); ++k) {
          Request *r = group[k];
          r->launched += slice[k];
          r->loss += (int64_t)loss[k];
          for (int q = 0; q < 5; ++q) r->ledger[q] += item_led[5 * k + q];
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
      fail_group(fail_hip(f));
    } catch (const std::bad_alloc &) {
      fail_group(host_out_of_memory());
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

int vmc_server_create(const vmc_server_desc *d, vmc_server **out) try {
  if (!d || !out) return VMC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (d->precision != VMC_FP64 && d->precision != VMC_FP32) return VMC_ERR_INVALID_ARGUMENT;
  if (d->tally != VMC_TALLY_FP64 && d->tally != VMC_TALLY_FIXED64) return VMC_ERR_INVALID_ARGUMENT;
  if (d->policy != VMC_SERVE_FAIR && d->policy != VMC_SERVE_SHORTEST) return VMC_ERR_INVALID_ARGUMENT;
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
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_server_submit(vmc_server *s, vmc_problem *prob, const vmc_run_desc *run, const vmc_output *out,
                      uint64_t *ticket) try {
  if (!s || !prob || !run || !ticket || run->Nphoton < 0 || prob->device != s->device) return VMC_ERR_INVALID_ARGUMENT;
  RunModel model;
  if (run_model(run, &model) != VMC_OK) return VMC_ERR_INVALID_ARGUMENT;
  if (s->desc.tally == VMC_TALLY_FIXED64 && run->Nphoton >= kFixedMaxPhotons) return VMC_ERR_INVALID_ARGUMENT;
  std::unique_ptr<Request> r(new Request());
  r->model = model;
  r->prob = prob;
  r->run = *run;
  r->run.precision = s->desc.precision;
  r->run.tally = s->desc.tally;
  if (out) r->out = *out;
  r->fd = 2.0 * M_PI * run->f > 0.0;
  r->ewords = prob->md->slots * (r->fd ? 2 : 1);
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
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
}

int vmc_server_poll(vmc_server *s, uint64_t ticket, int32_t *done) {
  if (!s || !done) return VMC_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(ticket);
  if (it == s->all.end()) return VMC_ERR_INVALID_ARGUMENT;
  *done = it->second->finished ? 1 : 0;
  return VMC_OK;
}

int vmc_server_wait(vmc_server *s, uint64_t ticket, vmc_stats *stats) try {
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
    for (int q = 0; q < 5; ++q) stats->ledger[q] = r->ledger[q];
    stats->kernel_ms = r->kernel_ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  }
  s->all.erase(it);
  return st;
} catch (const std::bad_alloc &) {
  return host_out_of_memory();
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
