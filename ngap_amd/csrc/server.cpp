// Request server: continuous batching of scan requests on resident groups.
//
// A request is (group, bytes), scanned as one chunk (start-of-data states at
// its first byte).  Requests wait in a queue; whenever one of `max_inflight`
// slots is free and the HBM budget allows, every waiting request that fits is
// packed into one persistent launch on that slot's stream (work items of 4 KiB
// windows, each request its own chunk bounds), so launches overlap on the
// device while new requests keep arriving.  Finished launches are retired by
// polling their events: reports are split back per request by offset range.
// There is no counterpart in the reference (it runs one input per process).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <unordered_map>

#include "host_pool.h"
#include "c_guard.h"
#include "runtime.h"

using namespace ngapamd;

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Pinned host buffers for request bytes, reused by size class (a power of
// two): the host-to-device copy of a launch is an asynchronous DMA from them.
struct PinnedPool {
  std::multimap<size_t, uint8_t *> free_;
  std::vector<uint8_t *> all;
  static size_t cls(size_t n) {
    size_t c = 1 << 16;
    while (c < n) c <<= 1;
    return c;
  }
  uint8_t *get(size_t n, size_t *cap) {
    const size_t c = cls(n);
    *cap = c;
    auto it = free_.find(c);
    if (it != free_.end()) {
      uint8_t *p = it->second;
      free_.erase(it);
      return p;
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess) return nullptr;
    all.push_back((uint8_t *)p);
    return (uint8_t *)p;
  }
  void put(uint8_t *p, size_t cap) {
    if (p) free_.emplace(cap, p);
  }
  ~PinnedPool() {
    for (uint8_t *p : all) (void)hipHostFree(p);
  }
};

struct Request {
  int32_t group = 0;
  uint8_t *data = nullptr;      // pinned, `cap` bytes, `len` used
  uint64_t len = 0, cap = 0;
  double t_submit = 0, t_done = 0;
  int state = 0;  // 0 queued, 1 in flight, 2 done, 3 failed
  std::vector<uint64_t> reports;
};

struct Slot {
  RunCtx ctx;
  DevBuf input;
  DevBuf last_map;              // bit per input byte: the last byte of a request
  Task *task_host = nullptr;    // pinned staging of the work items (async upload)
  size_t task_cap = 0;
  uint32_t *map_host = nullptr; // pinned staging of last_map
  size_t map_cap = 0;
  ~Slot() {
    if (task_host) (void)hipHostFree(task_host);
    if (map_host) (void)hipHostFree(map_host);
  }
  bool busy = false;
  std::vector<uint64_t> ids;
  std::vector<uint32_t> base;   // request start offsets in the slot input
  std::vector<Task> tasks;
  uint64_t bytes = 0;           // HBM held while in flight
  KParams P;
  uint32_t blocks = 0;
  // device bytes this slot's launch buffers hold (reserved, in flight or not)
  uint64_t held() const {
    return input.n + last_map.n + ctx.reports.n + ctx.spill.n + ctx.tasks.n + ctx.overflow.n +
           ctx.task_items.n + ctx.fb_tasks.n + ctx.fb_spill.n;
  }
  // gives the launch buffers back (an idle slot, when the budget needs room)
  void release() {
    input.release();
    last_map.release();
    ctx.reports.release();
    ctx.spill.release();
    ctx.tasks.release();
    ctx.overflow.release();
    ctx.task_items.release();
    ctx.fb_tasks.release();
    ctx.fb_spill.release();
    ctx.tasks_resident = false;
  }
};

constexpr uint32_t kMaxServeWindow = 4096;  // work-item positions, at most
constexpr uint32_t kServeSpill = 16384;     // worklist spill items per wave
constexpr uint64_t kMinLaunchBytes = 1 << 20;
// While launches are in flight, a free slot waits for this many queued bytes
// (or for the launches to finish) before it launches.
constexpr uint64_t kBusyLaunchBytes = 4 << 20;
constexpr uint64_t kMinReportCap = 1 << 16;  // report entries of a launch, at least

// Work-item window of a launch over `total` bytes: about total / (2 x waves)
// positions, a power of two in 256 .. 4096.  A short request spreads over
// many waves, which is what bounds its latency.
uint32_t serve_window(uint64_t total, uint64_t waves) {
  uint32_t win = 256;
  while (win < kMaxServeWindow && (uint64_t)win * 2 * waves < total) win *= 2;
  return win;
}

// Device bytes of a launch's buffers: input, last-byte map, reports (an
// overflow re-runs the launch with room for all), work items and their
// flags, and the worklist spill of its workgroups (kServeSpill items per
// wave).  A launch takes at most one workgroup per waves_per_block work
// items, and fewer when the HBM budget leaves room for fewer.
struct LaunchSize {
  uint64_t input, map, reports, tasks, overflow;
  uint64_t report_cap;
  uint32_t blocks;        // workgroups (spill areas)
  uint64_t per_block;     // spill bytes per workgroup
  uint64_t fixed() const { return input + map + reports + tasks + overflow; }
  uint64_t spill() const { return (uint64_t)blocks * per_block; }
  uint64_t total() const { return fixed() + spill(); }
};

LaunchSize launch_size(uint64_t bytes, uint64_t ntasks, uint32_t max_blocks,
                       uint32_t waves_per_block) {
  LaunchSize z;
  z.input = bytes + 16;
  z.map = (bytes / 32 + 2) * 4;
  z.report_cap = std::max<uint64_t>(kMinReportCap, bytes / 2);
  z.reports = z.report_cap * 8;
  z.tasks = ntasks * sizeof(Task);
  z.overflow = ntasks * 4 + 4;
  z.blocks = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(max_blocks, (ntasks + waves_per_block - 1) / waves_per_block));
  z.per_block = (uint64_t)waves_per_block * kServeSpill * 8;
  return z;
}

}  // namespace

struct ngap_server {
  ngap_engine *eng = nullptr;
  uint64_t budget = 0;
  ngap_run_opts opts;
  Mapping mp;
  std::vector<std::unique_ptr<Slot>> slots;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_id = 1;
  uint64_t inflight_bytes = 0;
  uint64_t queued_bytes = 0;
  uint32_t blocks_per_launch = 0;
  // statistics (ngap_server_stats)
  uint64_t peak_held = 0, peak_inflight = 0, launches = 0, rejected = 0, redone = 0;
  PinnedPool pinned;  // freed by ngap_server_destroy after every launch completed
  std::unique_ptr<HostPool> copier;  // staging copies of large requests
  uint64_t held() const {
    uint64_t h = 0;
    for (auto &x : slots) h += x->held();
    return h;
  }
};

namespace {

void fail_request(ngap_server *s, Request &r) {
  r.state = 3;
  r.t_done = now_ms();
  s->pinned.put(r.data, r.cap);
  r.data = nullptr;
}

// Device bytes the other slots hold.
uint64_t held_by_others(const ngap_server *s, const Slot &sl) {
  uint64_t h = 0;
  for (auto &x : s->slots)
    if (x.get() != &sl) h += x->held();
  return h;
}

// Whether a launch of size z fits the HBM budget on slot sl.  Slot buffers
// only grow, so first with what sl already holds; otherwise the buffers of
// idle slots (sl's own included) are given back, and the launch may take
// fewer workgroups (z.blocks lowered) so that its spill areas fit.
bool fits_budget(ngap_server *s, Slot &sl, LaunchSize &z) {
  const uint64_t keep = std::max<uint64_t>(sl.input.n, z.input) +
                        std::max<uint64_t>(sl.last_map.n, z.map) +
                        std::max<uint64_t>(sl.ctx.reports.n, z.reports) +
                        std::max<uint64_t>(sl.ctx.tasks.n, z.tasks) +
                        std::max<uint64_t>(sl.ctx.overflow.n, z.overflow) +
                        std::max<uint64_t>(sl.ctx.spill.n, z.spill()) + sl.ctx.task_items.n +
                        sl.ctx.fb_tasks.n + sl.ctx.fb_spill.n;
  if (held_by_others(s, sl) + keep <= s->budget) return true;
  sl.release();
  for (auto &x : s->slots)
    if (!x->busy) x->release();
  const uint64_t used = held_by_others(s, sl) + z.fixed();
  if (used + z.per_block > s->budget) return false;
  z.blocks = (uint32_t)std::min<uint64_t>(z.blocks, (s->budget - used) / z.per_block);
  return true;
}

int launch_slot(ngap_server *s, Slot &sl) {
  sl.ids.clear();
  sl.base.clear();
  sl.tasks.clear();
  const uint32_t wpb = kBlockThreads / kWave;
  const uint64_t waves = (uint64_t)s->blocks_per_launch * wpb;
  // continuous batching: a launch takes its share of what is queued (split
  // over the idle slots), so later arrivals find a slot soon
  uint32_t idle = 0;
  for (auto &x : s->slots) idle += !x->busy;
  auto share_now = [&] {
    return std::max<uint64_t>(kMinLaunchBytes, s->queued_bytes / std::max(1u, idle));
  };
  uint64_t share = share_now();
  // admission: the queue's head requests while the launch's buffers fit the
  // HBM budget next to everything the other slots hold.  A request that does
  // not fit waits for launches in flight to finish; one that cannot fit even
  // with nothing else held fails.
  uint64_t total = 0, ntasks_ub = 0;
  uint32_t adm_blocks = 0;  // workgroups the admitted requests' budget check allowed
  while (!s->queue.empty()) {
    const uint64_t id = s->queue.front();
    Request &r = s->reqs[id];
    const uint64_t need = total + r.len;
    if (!sl.ids.empty() && (need >= (1ull << 31) || total >= share)) break;
    // the window only grows with the launch's bytes: 256-position work items
    // bound the work-item count from above
    const uint64_t nt = ntasks_ub + (r.len + 255) / 256;
    LaunchSize z = launch_size(need, nt, s->blocks_per_launch, wpb);
    if (!fits_budget(s, sl, z)) {
      if (!sl.ids.empty()) break;
      bool others = false;
      for (auto &x : s->slots) others |= x->busy;
      if (others) break;  // wait for memory to free
      s->queue.pop_front();
      s->queued_bytes -= r.len;
      fail_request(s, r);
      s->rejected++;
      share = share_now();  // the share of what is left
      continue;
    }
    sl.ids.push_back(id);
    sl.base.push_back((uint32_t)total);
    total = need;
    ntasks_ub = nt;
    adm_blocks = z.blocks;
    s->queue.pop_front();
    s->queued_bytes -= r.len;
  }
  if (sl.ids.empty()) return NGAP_OK;
  const uint32_t win = serve_window(total, waves);
  // last-byte map: bit p set when p is the last byte of its request (the
  // streaming kernel's chunk-end test for items of any request)
  const size_t map_words = total / 32 + 2;
  if (sl.map_cap < map_words) {
    if (sl.map_host) HIP_TRY(hipHostFree(sl.map_host));
    sl.map_host = nullptr;
    sl.map_cap = 0;
    const size_t cap = std::max<size_t>(4096, map_words * 2);
    HIP_TRY(hipHostMalloc((void **)&sl.map_host, cap * 4, hipHostMallocDefault));
    sl.map_cap = cap;
  }
  memset(sl.map_host, 0, map_words * 4);
  for (size_t i = 0; i < sl.ids.size(); i++) {
    Request &r = s->reqs[sl.ids[i]];
    const uint32_t cs = sl.base[i], ce = cs + (uint32_t)r.len;
    sl.map_host[(ce - 1) >> 5] |= 1u << ((ce - 1) & 31);
    for (uint32_t a = cs; a < ce; a += win) {
      const uint32_t b = std::min<uint32_t>(a + win, ce);
      sl.tasks.push_back({(uint32_t)r.group, cs, ce, a, b, a == cs ? kTaskSeedStart : 0u});
    }
  }
  // the actual work items are no more than the admission bound: the launch
  // needs no more than what was admitted
  LaunchSize z = launch_size(total, sl.tasks.size(), s->blocks_per_launch, wpb);
  z.blocks = std::min(z.blocks, adm_blocks);
  HIP_TRY(sl.input.reserve(z.input));
  HIP_TRY(sl.last_map.reserve(z.map));
  for (size_t i = 0; i < sl.ids.size(); i++) {
    Request &r = s->reqs[sl.ids[i]];
    HIP_TRY(hipMemcpyAsync(sl.input.as<uint8_t>() + sl.base[i], r.data, r.len,
                           hipMemcpyHostToDevice, sl.ctx.stream));
    r.state = 1;
  }
  HIP_TRY(hipMemcpyAsync(sl.last_map.p, sl.map_host, map_words * 4, hipMemcpyHostToDevice,
                         sl.ctx.stream));
  // longest requests first inside the launch
  std::stable_sort(sl.tasks.begin(), sl.tasks.end(), [](const Task &x, const Task &y) {
    return (x.cend - x.cstart) > (y.cend - y.cstart);
  });
  if (sl.task_cap < sl.tasks.size()) {
    if (sl.task_host) HIP_TRY(hipHostFree(sl.task_host));
    sl.task_host = nullptr;
    sl.task_cap = 0;
    const size_t cap = std::max<size_t>(4096, sl.tasks.size() * 2);
    HIP_TRY(hipHostMalloc((void **)&sl.task_host, cap * sizeof(Task), hipHostMallocDefault));
    sl.task_cap = cap;
  }
  std::copy(sl.tasks.begin(), sl.tasks.end(), sl.task_host);
  HIP_TRY(sl.ctx.tasks.reserve(z.tasks));
  HIP_TRY(hipMemcpyAsync(sl.ctx.tasks.p, sl.task_host, sl.tasks.size() * sizeof(Task),
                         hipMemcpyHostToDevice, sl.ctx.stream));
  sl.ctx.tasks_resident = false;
  HIP_TRY(sl.ctx.overflow.reserve(z.overflow));
  HIP_TRY(sl.ctx.reports.reserve(z.reports));
  sl.ctx.report_cap = z.report_cap;
  HIP_TRY(sl.ctx.spill.reserve(z.spill()));
  KParams P = base_params(s->eng, s->mp, s->opts);
  P.input = sl.input.as<uint8_t>();
  P.chunk_len = 0;
  P.num_seg = 0;
  P.last_map = sl.last_map.as<uint32_t>();
  P.tasks = sl.ctx.tasks.as<Task>();
  P.num_tasks = (uint32_t)sl.tasks.size();
  P.overflow = sl.ctx.overflow.as<uint32_t>();
  P.spill = sl.ctx.spill.as<uint64_t>();
  P.spill_cap = kServeSpill;
  attach_ctx(P, sl.ctx);
  sl.P = P;
  // a small launch takes only the workgroups its work items can feed, so
  // short requests do not pay for a full persistent grid
  sl.blocks = z.blocks;
  HIP_TRY(launch_reset(sl.ctx.stream, P.report_count, P.items, P.task_head, P.overflow,
                       P.num_tasks, P.overflow_count, 0, nullptr));
  HIP_TRY(launch_batched(sl.ctx.stream, P, sl.blocks, lds_bytes(P.ring_cap)));
  // the counters follow the launch into pinned memory: retiring reads them
  // without another copy
  HIP_TRY(hipMemcpyAsync(sl.ctx.host_cnt, sl.ctx.counters.p, 32, hipMemcpyDeviceToHost,
                         sl.ctx.stream));
  HIP_TRY(hipEventRecord(sl.ctx.ev1, sl.ctx.stream));
  sl.bytes = z.total();
  s->inflight_bytes += sl.bytes;
  s->peak_inflight = std::max(s->peak_inflight, s->inflight_bytes);
  s->peak_held = std::max(s->peak_held, s->held());
  s->launches++;
  sl.busy = true;
  return NGAP_OK;
}

int retire_slot(ngap_server *s, Slot &sl) {
  HIP_TRY(hipEventSynchronize(sl.ctx.ev1));
  unsigned long long c4[4] = {0, 0, 0, 0};
  memcpy(c4, sl.ctx.host_cnt, 32);
  const unsigned long long cnt = c4[0];
  // overflowed work items: the low word of the fourth counter
  const bool redo = cnt > sl.ctx.report_cap || (uint32_t)c4[3] != 0;
  if (redo) {
    // overflow: re-run the whole launch synchronously with room for every
    // report and the engine's overflow fallback for the worklists (on the
    // launch's own grid).  This re-run's buffers are not held to the budget:
    // the launch already holds its share, and the fallback grows only the
    // worklists of the work items that overflowed.
    KParams P = sl.P;
    ngap_run_opts o = s->opts;
    o.fallback = 1;
    o.persistent_blocks = sl.blocks;
    o.spill_capacity = kServeSpill;
    o.root_pass = 0;
    o.task_traces = 0;
    if (cnt > sl.ctx.report_cap) {
      HIP_TRY(sl.ctx.reports.reserve(cnt * 8 + 8));
      sl.ctx.report_cap = cnt + 1;
    }
    s->redone++;
    int rc = run_tasks(s->eng, sl.ctx, P, sl.tasks, 0, o, nullptr);
    s->peak_held = std::max(s->peak_held, s->held());
    if (rc) {
      for (uint64_t id : sl.ids) fail_request(s, s->reqs[id]);
      sl.busy = false;
      s->inflight_bytes -= sl.bytes;
      return rc;
    }
  } else {
    sl.ctx.last_count = cnt;
  }
  std::vector<uint64_t> v;
  int rc = sorted_reports(sl.ctx, v);
  if (rc) return rc;
  // split by request: device reports are (offset << 32) | id, offsets global
  // in the slot input; requests are contiguous and sorted by base
  std::vector<size_t> order(sl.ids.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return sl.base[a] < sl.base[b]; });
  const double t = now_ms();
  size_t k = 0;
  for (size_t oi = 0; oi < order.size(); oi++) {
    const size_t i = order[oi];
    Request &r = s->reqs[sl.ids[i]];
    const uint64_t lo = sl.base[i], hi = lo + r.len;
    r.reports.clear();
    while (k < v.size() && (v[k] >> 32) < hi) {
      const uint64_t off = (v[k] >> 32) - lo;
      r.reports.push_back(((v[k] & 0xffffffffull) << 32) | off);
      k++;
    }
    r.state = 2;
    r.t_done = t;
    s->pinned.put(r.data, r.cap);
    r.data = nullptr;
  }
  sl.busy = false;
  s->inflight_bytes -= sl.bytes;
  return NGAP_OK;
}

}  // namespace

extern "C" {

int ngap_server_create(ngap_engine *eng, uint64_t budget, int32_t max_inflight,
                       const ngap_run_opts *opts, ngap_server **out) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !out || !opts) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (!eng->loaded) return ngap_fail(NGAP_ERR_STATE, "no groups loaded");
    if (max_inflight < 1 || max_inflight > 16)
      return ngap_fail(NGAP_ERR_INVALID, "max_inflight must be in [1, 16]");
    if (opts->report_off)
      return ngap_fail(NGAP_ERR_INVALID,
                       "report_off is not supported by the server: requests return their reports");
    std::unique_ptr<ngap_server> s(new ngap_server());
    s->eng = eng;
    s->budget = budget ? budget : (200ull << 30);
    s->opts = *opts;
    std::string err;
    if (!map_algorithm(opts->algorithm, opts->precompute_depth, eng->img.memo_depth, s->mp, err))
      return ngap_fail(NGAP_ERR_INVALID, err);
    HIP_TRY(hipSetDevice(eng->device));
    uint32_t blocks = 0;
    if (!ring_of(*opts))
      return ngap_fail(NGAP_ERR_INVALID, "ring_capacity must be a power of two in [64, 2048]");
    // request launches run the streaming kernel (chunk ends from the last-byte map)
    int rc = persistent_blocks(eng, opts->persistent_blocks, &blocks, ring_of(*opts), 3);
    if (rc) return rc;
    // every launch gets the whole persistent grid: launches in flight share the
    // CUs, and a lone launch is not held to a fraction of the GPU
    s->blocks_per_launch = std::max<uint32_t>(1, blocks);
    s->copier.reset(new HostPool());  // the copy of a 4 MiB request in up to 8 parts
    for (int i = 0; i < max_inflight; i++) {
      s->slots.emplace_back(new Slot());
      rc = s->slots.back()->ctx.init();
      if (rc) return rc;
    }
    *out = s.release();
    return NGAP_OK;
  });
}

int ngap_server_destroy(ngap_server *s) {
  return ngapamd::c_guard([&]() -> int {
    if (!s) return NGAP_OK;
    for (auto &sl : s->slots)
      if (sl->busy) (void)hipEventSynchronize(sl->ctx.ev1);
    delete s;
    return NGAP_OK;
  });
}

int ngap_server_submit(ngap_server *s, int32_t group, const uint8_t *data, uint64_t len,
                       uint64_t *req_id) {
  return ngapamd::c_guard([&]() -> int {
    if (!s || !req_id || (!data && len)) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (group < 0 || group >= s->eng->img.num_groups)
      return ngap_fail(NGAP_ERR_INVALID, "group out of range");
    if (len == 0 || len >= (1ull << 31)) return ngap_fail(NGAP_ERR_INVALID, "bad request length");
    const uint64_t id = s->next_id++;
    uint64_t cap = 0;
    uint8_t *buf = s->pinned.get(len, &cap);
    if (!buf) return ngap_fail(NGAP_ERR_OOM, "pinned host allocation failed");
    s->copier->copy(buf, data, len);
    Request &r = s->reqs[id];
    r.group = group;
    r.data = buf;
    r.len = len;
    r.cap = cap;
    r.t_submit = now_ms();
    s->queue.push_back(id);
    s->queued_bytes += len;
    *req_id = id;
    return NGAP_OK;
  });
}

int ngap_server_poll(ngap_server *s, int32_t block) {
  return ngapamd::c_guard([&]() -> int {
    if (!s) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(s->eng->device));
    bool retired = false;
    for (auto &sl : s->slots) {
      if (!sl->busy) continue;
      hipError_t q = hipEventQuery(sl->ctx.ev1);
      if (q == hipSuccess) {
        int rc = retire_slot(s, *sl);
        if (rc) return rc;
        retired = true;
      } else if (q != hipErrorNotReady) {
        HIP_TRY(q);
      }
    }
    // an idle GPU takes the queue at once; while launches are in flight a free
    // slot waits until kBusyLaunchBytes are queued or they finish, so that busy
    // periods pack requests into fewer, larger launches (each launch costs host
    // work and a kernel's fixed start-up, whatever its size)
    bool busy_any = false;
    for (auto &sl : s->slots) busy_any |= sl->busy;
    for (auto &sl : s->slots) {
      if (s->queue.empty()) break;
      if (sl->busy) continue;
      if (busy_any && s->queued_bytes < kBusyLaunchBytes) break;
      int rc = launch_slot(s, *sl);
      if (rc) return rc;
      busy_any = true;
    }
    if (block && !retired) {
      // wait for the oldest launch in flight
      for (auto &sl : s->slots)
        if (sl->busy) {
          int rc = retire_slot(s, *sl);
          if (rc) return rc;
          break;
        }
    }
    return NGAP_OK;
  });
}

int ngap_server_take(ngap_server *s, uint64_t req_id, int32_t *state, uint64_t *out,
                     uint64_t cap, uint64_t *n_out, double *latency_ms) {
  return ngapamd::c_guard([&]() -> int {
    if (!s || !state || !n_out) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return ngap_fail(NGAP_ERR_INVALID, "unknown request");
    Request &r = it->second;
    *state = r.state;
    *n_out = r.reports.size();
    if (latency_ms) *latency_ms = r.state >= 2 ? r.t_done - r.t_submit : -1.0;
    if (r.state == 2 && out) {
      const uint64_t n = std::min<uint64_t>(cap, r.reports.size());
      std::copy(r.reports.begin(), r.reports.begin() + n, out);
      if (n == r.reports.size()) s->reqs.erase(it);  // delivered
    }
    return NGAP_OK;
  });
}

int ngap_server_pending(ngap_server *s, uint64_t *queued, uint64_t *inflight) {
  return ngapamd::c_guard([&]() -> int {
    if (!s || !queued || !inflight) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    *queued = s->queue.size();
    uint64_t n = 0;
    for (auto &sl : s->slots)
      if (sl->busy) n += sl->ids.size();
    *inflight = n;
    return NGAP_OK;
  });
}

int ngap_server_stats(ngap_server *s, ngap_server_info *info) {
  return ngapamd::c_guard([&]() -> int {
    if (!s || !info) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    memset(info, 0, sizeof *info);
    info->budget_bytes = s->budget;
    info->queued = s->queue.size();
    for (auto &sl : s->slots)
      if (sl->busy) info->inflight_requests += sl->ids.size();
    info->inflight_bytes = s->inflight_bytes;
    info->held_bytes = s->held();
    info->peak_inflight_bytes = s->peak_inflight;
    info->peak_held_bytes = s->peak_held;
    info->launches = s->launches;
    info->rejected = s->rejected;
    info->redone = s->redone;
    return NGAP_OK;
  });
}

}  // extern "C"
