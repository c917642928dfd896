// Host runtime behind the C ABI: device image upload, input upload, launch of
// the plain and batched schedules (optionally replayed from a hipGraph),
// overflow fallback, report retrieval.
//
// Reference counterparts: ngap::launch_non_blocking_*_groups
// (code/src/ngap/ngap.cu:284-1990), NonBlockingBuffer::init_nfagroups
// (code/src/ngap/ngap_buffer.cu:475-748) and the result pack/sort/unique
// (code/src/ngap/ngap.cu:1899-1933).  Differences by design: capacities are
// runtime values, start states are seeded on the device, device buffers live
// with the engine (resident across runs) instead of per launch, and errors are
// status codes instead of exit().
#include "runtime.h"
#include "c_guard.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>

// Schedule ids are unique across the engines of a process, so an id handed
// back through ngap_run_opts.tasks_id never matches another engine's schedule.
static uint64_t next_schedule_id() {
  static std::atomic<uint64_t> next{1};
  return next.fetch_add(1);
}

int ngap_fail(int code, const std::string &msg) {
  return ngapamd::c_guard([&]() -> int {
    ngap_set_last_error_internal(msg);
    return code;
  });
}

namespace ngapamd {

hipError_t DevBuf::reserve(size_t bytes) {
  if (bytes <= n && p) return hipSuccess;
  release();
  size_t want = bytes ? bytes : 8;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipSuccess) n = want;
  else p = nullptr;
  return e;
}

void DevBuf::release() {
  if (p && owned) (void)hipFree(p);
  p = nullptr;
  n = 0;
  owned = true;
}

RunCtx::~RunCtx() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
  if (host_cnt) (void)hipHostFree(host_cnt);
}

int RunCtx::init() {
  HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&ev0));
  HIP_TRY(hipEventCreate(&ev1));
  HIP_TRY(counters.reserve(kCounterBytes));
  // the reset kernel clears the counters a run uses; the rest stays zero
  HIP_TRY(hipMemset(counters.p, 0, counters.n));
  HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&host_cnt), kCounterBytes, hipHostMallocDefault));
  return NGAP_OK;
}

int RunCtx::finish(float *ms, unsigned long long cnt[4], bool record_end) {
  if (record_end) HIP_TRY(hipEventRecord(ev1, stream));
  HIP_TRY(hipMemcpyAsync(host_cnt, counters.p, kCounterBytes, hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  HIP_TRY(hipEventElapsedTime(ms, ev0, ev1));
  memcpy(cnt, host_cnt, 32);
  cnt[1] = items_sum(host_cnt);
  return NGAP_OK;
}

void RunCtx::drop_graph() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
  gkey.clear();
}

uint32_t lds_bytes(uint32_t ring) {
  return 16 + (kBlockThreads / kWave) * wave_lds_bytes(ring, kRbuf, kDc);
}

// ring_capacity option -> LDS ring entries (0 = default); 0 when invalid
uint32_t ring_of(const ngap_run_opts &o) {
  const uint32_t r = o.ring_capacity ? o.ring_capacity : kRing;
  if (r < kMinRing || r > kMaxRing || (r & (r - 1))) return 0;
  return r;
}

bool map_algorithm(int algorithm, int depth_req, int loaded_depth, Mapping &m,
                   std::string &err) {
  // --algorithm -> engine features: each name keeps the optimisations the
  // reference adds at that step (main.cu:114-130, README.md:141-150).
  // NAPgroups injects the all-input roots of one position at a time
  // (kernel_bap.cu:98-105, a fake item of capacity 1); the later algorithms
  // take a batch of positions per step (kernel_ngap_O1.cu:84-112), a wave's
  // width here.
  const int batch = (int)kWave;
  switch (algorithm) {
    case NGAP_ALGO_NAP:
      m = {1, 0, 1};
      break;
    case NGAP_ALGO_NONBLOCKING:
      m = {1, 0, batch};
      break;
    case NGAP_ALGO_R1:
      m = {1, 1, batch};
      break;
    case NGAP_ALGO_R2:
      m = {1, 2, batch};
      break;
    case NGAP_ALGO_PC:
      m = {depth_req < 1 ? 1 : depth_req, 0, batch};
      break;
    case NGAP_ALGO_ALL:
      m = {depth_req < 1 ? 1 : depth_req, -1, batch};
      break;
    case NGAP_ALGO_BLOCKING:
      m = {1, 0, batch};
      break;
    default:
      err = "unknown algorithm";
      return false;
  }
  if (m.memo_depth > kMaxMemoDepth) {
    err = "precompute depth must be <= 3";
    return false;
  }
  if (m.memo_depth > loaded_depth) {
    err = "precompute depth exceeds the depth the groups were loaded with";
    return false;
  }
  return true;
}

KParams base_params(const ngap_engine *eng, const Mapping &mp, const ngap_run_opts &o) {
  KParams P;
  memset(&P, 0, sizeof P);
  P.img.meta = eng->d_meta.as<GroupMeta>();
  P.img.succ4 = eng->d_succ4.as<int32_t>();
  P.img.cls = eng->d_cls.as<uint32_t>();
  P.img.sod = eng->d_sod.as<int32_t>();
  for (int d = 0; d < kMaxMemoDepth; d++) {
    P.img.beg[d] = eng->d_beg[d].as<uint32_t>();
    P.img.nrep[d] = eng->d_nrep[d].as<uint32_t>();
    P.img.list[d] = eng->d_list[d].as<int32_t>();
    P.img.sel[d] = eng->d_sel[d].as<uint32_t>();
  }
  P.img.mask1 = eng->d_mask1.as<uint32_t>();
  P.img.xr1 = eng->d_xr1.as<uint32_t>();
  P.img.xe1 = eng->d_xe1.as<uint32_t>();
  P.img.xmask1 = eng->d_xmask1.as<uint32_t>();
  P.img.xr2 = eng->d_xr2.as<uint32_t>();
  P.img.xe2 = eng->d_xe2.as<uint32_t>();
  P.img.xsod = eng->d_xsod.as<uint32_t>();
  P.img.xs3 = eng->d_xs3.as<uint32_t>();
  P.img.xp3 = eng->d_xp3.as<uint32_t>();
  P.img.x2rep = eng->d_x2rep.as<uint32_t>();
  P.img.n_xsod = (uint32_t)eng->img.n_xsod;
  P.num_groups = (uint32_t)eng->img.num_groups;
  P.memo_depth = mp.memo_depth;
  P.is_dup = o.is_dup;
  P.dedup = o.unique;
  P.priv_limit = mp.priv_limit;
  P.active_threshold = o.active_threshold;
  P.report_off = o.report_off;
  P.ring_cap = o.schedule == NGAP_SCHED_BATCHED && ring_of(o) ? ring_of(o) : kRing;
  P.rbuf_cap = kRbuf;
  P.dc_cap = kDc;
  P.root_batch = (uint32_t)mp.root_batch;
  P.waves_per_block = kBlockThreads / kWave;
  return P;
}

void attach_ctx(KParams &P, RunCtx &ctx) {
  unsigned long long *counters = ctx.counters.as<unsigned long long>();
  P.report_count = counters + 0;
  P.items = counters + kItemsWord;
  P.task_head = reinterpret_cast<uint32_t *>(counters + 16);  // byte 128: the queue counters
  P.overflow_count = reinterpret_cast<uint32_t *>(counters + 3);
  P.reports = ctx.reports.as<uint64_t>();
  P.report_cap = ctx.report_cap;
}

// Kernel kind an engine run of o launches: 0 k_batched (traces), 1
// k_batched_stream, 2 k_expand (after the root pass).
int engine_kind(const ngap_run_opts &o) { return o.task_traces ? 0 : o.root_pass ? 2 : 1; }

int persistent_blocks(const ngap_engine *eng, uint32_t requested, uint32_t *blocks,
                      uint32_t ring, int kind) {
  const int slot = kBatchedKinds * (31 - __builtin_clz(ring)) + kind;  // ring in [64, 2048]
  int per_cu = eng->batched_per_cu[slot];
  if (per_cu == 0) {
    HIP_TRY(batched_occupancy(&per_cu, lds_bytes(ring), kind));
    const_cast<ngap_engine *>(eng)->batched_per_cu[slot] = per_cu;
  }
  if (per_cu < 1) per_cu = 1;
  *blocks = requested ? requested : (uint32_t)(per_cu * eng->num_cus);
  return NGAP_OK;
}

uint32_t job_blocks(uint32_t full_blocks, uint32_t waves_per_block, uint64_t work) {
  const uint64_t per_block = kMinWavePositions * std::max<uint32_t>(1, waves_per_block);
  const uint64_t want = (work + per_block - 1) / per_block;
  return (uint32_t)std::min<uint64_t>(full_blocks, std::max<uint64_t>(64, want));
}

// Root pass set-up for the resident work items: each work item's area of
// root items (factor items per position, the start-of-data seeds of a seeding
// work item and 64 spare) and its pieces of at most kRootPiece positions.
// Returns false when the areas would exceed kRootBudget bytes (no root pass).
constexpr uint32_t kRootPiece = 512;
constexpr uint64_t kRootBudget = 16ull << 30;

static int root_areas(const ngap_engine *eng, RunCtx &ctx, const std::vector<Task> &tasks,
                      uint32_t factor, bool same_tasks, bool *ok) {
  *ok = false;
  if (same_tasks && ctx.root_factor == factor && ctx.root_off.p) {
    *ok = true;
    return NGAP_OK;
  }
  ctx.root_factor = 0;
  const uint64_t n_xsod = eng->img.n_xsod;
  std::vector<uint64_t> off(tasks.size() + 1, 0);
  std::vector<RootPiece> pieces;
  for (size_t t = 0; t < tasks.size(); t++) {
    const Task &T = tasks[t];
    const bool seed = (T.flags & kTaskSeedStart) != 0;
    uint64_t sods = 0;
    if (seed)
      sods = T.group == kAllGroups ? n_xsod
                                   : eng->img.meta[T.group].sod_end - eng->img.meta[T.group].sod_begin;
    off[t + 1] = off[t] + (uint64_t)(T.end - T.begin) * factor + sods + kWave;
    if (T.begin == T.end && seed) pieces.push_back({(uint32_t)t, T.begin, T.end, T.flags});
    for (uint32_t a = T.begin; a < T.end; a += kRootPiece) {
      const uint32_t b = T.end - a > kRootPiece ? a + kRootPiece : T.end;
      pieces.push_back({(uint32_t)t, a, b, a == T.begin ? T.flags : 0u});
    }
  }
  if (off.back() * 8 > kRootBudget || pieces.size() >= (1u << 31)) return NGAP_OK;
  HIP_TRY(ctx.root_off.upload(off));
  HIP_TRY(ctx.pieces.upload(pieces));
  HIP_TRY(ctx.root_items.reserve(std::max<uint64_t>(off.back(), 1) * 8));
  HIP_TRY(ctx.root_count.reserve(std::max<size_t>(tasks.size(), 1) * 8));
  ctx.num_pieces = (uint32_t)pieces.size();
  ctx.root_factor = factor;
  *ok = true;
  return NGAP_OK;
}

// One batched launch: clears the run's counters (k_reset, a kernel), then
// launches the matching between the events `begin` and `end`: the root pass
// (when on) and the persistent kernel, replayed from a captured graph when
// use_graph is set and the parameters match the capture.  The graph holds
// the matching kernels only; the events are recorded on the stream around
// its launch (event records captured into a graph are not re-recorded on
// replay), so begin..end is the matching plus the graph's launch latency.
static int launch_batched_ctx(RunCtx &ctx, const KParams &P, uint32_t blocks, uint32_t n_flags,
                              bool keep_reports, bool use_graph, hipEvent_t begin,
                              hipEvent_t end) {
  const uint32_t lds = lds_bytes(P.ring_cap);
  HIP_TRY(launch_reset(ctx.stream, P.report_count, P.items, P.task_head, P.overflow, n_flags,
                       P.overflow_count, keep_reports, P.root_count));
  if (use_graph) {
    std::vector<unsigned char> key(sizeof P + 4);
    memcpy(key.data(), &P, sizeof P);
    memcpy(key.data() + sizeof P, &blocks, 4);
    if (!ctx.gexec || key != ctx.gkey) {
      ctx.drop_graph();
      HIP_TRY(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal));
      hipError_t e4 = P.root_items ? launch_roots(ctx.stream, P, lds) : hipSuccess;
      hipError_t e2 = launch_batched(ctx.stream, P, blocks, lds);
      hipGraph_t g = nullptr;
      hipError_t e3 = hipStreamEndCapture(ctx.stream, &g);
      ctx.graph = g;
      HIP_TRY(e4);
      HIP_TRY(e2);
      HIP_TRY(e3);
      HIP_TRY(hipGraphInstantiate(&ctx.gexec, ctx.graph, nullptr, nullptr, 0));
      ctx.gkey = key;
    }
    HIP_TRY(hipEventRecord(begin, ctx.stream));
    HIP_TRY(hipGraphLaunch(ctx.gexec, ctx.stream));
    HIP_TRY(hipEventRecord(end, ctx.stream));
    return NGAP_OK;
  }
  HIP_TRY(hipEventRecord(begin, ctx.stream));
  if (P.root_items) HIP_TRY(launch_roots(ctx.stream, P, lds));
  HIP_TRY(launch_batched(ctx.stream, P, blocks, lds));
  HIP_TRY(hipEventRecord(end, ctx.stream));
  return NGAP_OK;
}

int run_tasks(ngap_engine *eng, RunCtx &ctx, KParams P, const std::vector<Task> &tasks,
              uint64_t sched_id, const ngap_run_opts &o, ngap_run_stats *st) {
  if (tasks.size() >= (1u << 31)) return ngap_fail(NGAP_ERR_INVALID, "too many work items");
  const uint32_t T = (uint32_t)tasks.size();
  // the work items stay on the device between runs of the same schedule
  const bool same = ctx.tasks_resident &&
                    ((sched_id && sched_id == ctx.resident_sched) ||
                     (tasks.size() == ctx.last_tasks.size() &&
                      (T == 0 || memcmp(tasks.data(), ctx.last_tasks.data(),
                                        T * sizeof(Task)) == 0)));
  if (!same) {
    ctx.tasks_resident = false;
    HIP_TRY(ctx.tasks.upload(tasks));
    ctx.last_tasks = tasks;
    ctx.tasks_resident = true;
  }
  ctx.resident_sched = sched_id;
  HIP_TRY(ctx.task_items.reserve((size_t)std::max<uint32_t>(T, 1) * 8));
  HIP_TRY(ctx.overflow.reserve((size_t)std::max<uint32_t>(T, 1) * 4));
  P.tasks = ctx.tasks.as<Task>();
  P.task_order = nullptr;
  P.num_tasks = T;
  P.task_items = o.task_traces ? ctx.task_items.as<unsigned long long>() : nullptr;
  ctx.last_traces = o.task_traces != 0;
  P.overflow = ctx.overflow.as<uint32_t>();
  // root pass: engine runs on the streaming kernel only
  int rc = NGAP_OK;
  if (o.root_pass && P.chunk_len && !P.task_items && T) {
    bool ok = false;
    if (ctx.root_gen != eng->generation) {
      ctx.root_gen = eng->generation;
      ctx.root_boost = 1;
    }
    const uint32_t factor = (o.root_capacity ? o.root_capacity : 4) * ctx.root_boost;
    rc = root_areas(eng, ctx, tasks, factor, same, &ok);
    if (rc) return rc;
    if (ok) {
      P.pieces = ctx.pieces.as<RootPiece>();
      P.num_pieces = ctx.num_pieces;
      P.root_off = ctx.root_off.as<uint64_t>();
      P.root_items = ctx.root_items.as<uint64_t>();
      P.root_count = ctx.root_count.as<unsigned long long>();
    }
  }
  uint32_t blocks = 0;
  // engine runs without traces launch the streaming kernel (k_expand after a
  // root pass), the others the per-work-item kernel (launch_batched)
  rc = persistent_blocks(eng, o.persistent_blocks, &blocks, P.ring_cap, batched_kind(P));
  if (rc) return rc;
  P.spill_cap = o.spill_capacity ? o.spill_capacity : 65536;
  HIP_TRY(ctx.spill.reserve((size_t)blocks * P.waves_per_block * P.spill_cap * 8));
  P.spill = ctx.spill.as<uint64_t>();
  attach_ctx(P, ctx);

  rc = launch_batched_ctx(ctx, P, blocks, T, false, o.use_graph != 0, ctx.ev0, ctx.ev1);
  if (rc) return rc;
  float ms = 0;
  unsigned long long cnt[4] = {0, 0, 0, 0};
  rc = ctx.finish(&ms, cnt, false);
  if (rc) return rc;
  std::vector<uint32_t> bad;
  bool spill_over = false;  // some work item lost items (not only its root area)
  if ((uint32_t)cnt[3] && T) {  // per-item flags only when some work item overflowed
    std::vector<uint32_t> flags(T);
    HIP_TRY(hipMemcpy(flags.data(), P.overflow, T * 4ull, hipMemcpyDeviceToHost));
    bool root_over = false;
    for (uint32_t i = 0; i < T; i++) {
      if (flags[i]) bad.push_back(i);
      spill_over |= (flags[i] & kSpillOverflow) != 0;
      root_over |= (flags[i] & kRootOverflow) != 0;
    }
    // later runs reserve larger root areas
    if (root_over) ctx.root_boost = std::min<uint32_t>(ctx.root_boost * 2, 64);
  }
  if (st) {
    st->kernel_ms = ms;
    st->items = cnt[1];
    st->overflow_units = bad.size();
    st->num_tasks = T;
  }
  // Overflow fallback: re-run only the overflowed work items, on a small grid
  // with a worklist 16x larger each round; their earlier partial reports stay
  // (the set union is exact).  Work items whose root-pass area filled up are
  // re-run (without a root pass) even when the fallback is off, at the same
  // worklist capacity then.
  uint32_t spill = P.spill_cap;
  double fb_ms = 0;
  uint32_t fb_tasks = 0;
  for (int round = 0; !bad.empty() && round < 3 && (o.fallback || (round == 0 && !spill_over));
       round++) {
    std::vector<Task> redo;
    for (uint32_t i : bad) redo.push_back(tasks[i]);
    fb_tasks += (uint32_t)redo.size();
    if (o.fallback) spill = (uint32_t)std::min<uint64_t>((uint64_t)spill * 16, 1u << 24);
    uint32_t fb_blocks = std::min<uint32_t>(blocks, ((uint32_t)redo.size() + 3) / 4);
    const size_t per_block = (size_t)P.waves_per_block * spill * 8;
    if ((size_t)fb_blocks * per_block > (32ull << 30))
      fb_blocks = (uint32_t)std::max<size_t>(1, (32ull << 30) / per_block);
    KParams F = P;
    HIP_TRY(ctx.fb_tasks.upload(redo));
    HIP_TRY(ctx.fb_spill.reserve((size_t)fb_blocks * per_block));
    F.tasks = ctx.fb_tasks.as<Task>();
    F.num_tasks = (uint32_t)redo.size();
    F.task_items = nullptr;
    F.spill = ctx.fb_spill.as<uint64_t>();
    F.spill_cap = spill;
    // the re-run computes its roots in the persistent kernel
    F.pieces = nullptr;
    F.num_pieces = 0;
    F.root_off = nullptr;
    F.root_items = nullptr;
    F.root_count = nullptr;
    rc = launch_batched_ctx(ctx, F, fb_blocks, F.num_tasks, true, false, ctx.ev0, ctx.ev1);
    if (rc) return rc;
    HIP_TRY(hipEventSynchronize(ctx.ev1));
    HIP_TRY(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1));
    fb_ms += ms;
    std::vector<uint32_t> f2(F.num_tasks);
    HIP_TRY(hipMemcpy(f2.data(), F.overflow, F.num_tasks * 4ull, hipMemcpyDeviceToHost));
    std::vector<uint32_t> still;
    for (uint32_t i = 0; i < F.num_tasks; i++)
      if (f2[i]) still.push_back(bad[i]);
    bad.swap(still);
  }
  if (fb_tasks) HIP_TRY(hipMemcpy(cnt, ctx.counters.p, 16, hipMemcpyDeviceToHost));
  ctx.last_count = cnt[0];
  ctx.last_report_off = o.report_off;
  ctx.have_run = true;
  if (st) {
    st->reports_raw = cnt[0];
    st->fallback_tasks = fb_tasks;
    st->fallback_ms = fb_ms;
  }
  if (!bad.empty())
    return ngap_fail(NGAP_ERR_OVERFLOW, "worklist spill capacity exceeded in " +
                                            std::to_string(bad.size()) + " work items");
  if (!o.report_off && cnt[0] > ctx.report_cap)
    return ngap_fail(NGAP_ERR_OVERFLOW, "report capacity exceeded: " + std::to_string(cnt[0]) +
                                            " reports > --result-capacity " +
                                            std::to_string(ctx.report_cap));
  return NGAP_OK;
}

int sorted_reports(RunCtx &ctx, std::vector<uint64_t> &v) {
  v.resize(std::min<uint64_t>(ctx.last_count, ctx.report_cap));
  if (!v.empty())
    HIP_TRY(hipMemcpy(v.data(), ctx.reports.p, v.size() * 8, hipMemcpyDeviceToHost));
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  return NGAP_OK;
}

namespace {

// Static schedule (no host work items): a group whose roots are rare under the
// input's byte histogram takes whole chunks per work item; others take
// `window` positions.  Work items run longest-expected-first.
// Work-item length near `bytes` / (2 x waves), rounded to a power of two
// (256 .. 65536): about two work items per wave.
uint32_t auto_window(uint64_t bytes, uint32_t waves) {
  const double per = (double)bytes / std::max<uint64_t>(1, 2ull * waves);
  uint32_t win = 256;
  while (win < 65536 && (double)win * 1.41421356 < per) win *= 2;
  return win;
}

std::vector<Task> static_schedule(const ngap_engine *eng, uint32_t window, bool is_dup,
                                  uint32_t waves) {
  const Image &im = eng->img;
  const uint32_t G = (uint32_t)im.num_groups, S = eng->num_seg, L = (uint32_t)eng->chunk_len;
  uint32_t win = window ? window : auto_window((uint64_t)G * S * L, waves);
  if (win > L) win = L;
  std::vector<double> root_p(G, 0.0);
  for (uint32_t g = 0; g < G; g++)
    for (int b = 0; b < 256; b++)
      if ((im.mask1[g * 16 + b / 32] >> (b % 32)) & 1u) root_p[g] += eng->byte_freq[b];
  std::vector<Task> tasks;
  std::vector<double> cost;
  for (uint32_t c = 0; c < S; c++)
    for (uint32_t g = 0; g < G; g++) {
      uint32_t wg = root_p[g] < 1.0 / 16 ? L : win;
      for (uint32_t a = 0; a < L; a += wg) {
        uint32_t b = (uint32_t)std::min<uint64_t>((uint64_t)a + wg, L);
        const uint32_t cs = c * L;
        const uint32_t seed = (a == 0 && (c == 0 || is_dup)) ? kTaskSeedStart : 0u;
        tasks.push_back({g, cs, cs + L, cs + a, cs + b, seed});
        cost.push_back((b - a) * (0.02 + root_p[g]));
      }
    }
  std::vector<uint32_t> idx(tasks.size());
  for (uint32_t i = 0; i < idx.size(); i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
  std::vector<Task> out(tasks.size());
  for (size_t i = 0; i < idx.size(); i++) out[i] = tasks[idx[i]];
  return out;
}

// Static schedule of cross-group work items: every chunk cut into windows
// that scan all groups at once; by default about four work items per wave.
std::vector<Task> cross_schedule(const ngap_engine *eng, uint32_t window, bool is_dup,
                                 uint32_t waves) {
  const uint32_t S = eng->num_seg, L = (uint32_t)eng->chunk_len;
  uint32_t win = window ? window : auto_window((uint64_t)S * L, waves);
  if (win > L) win = L;
  std::vector<Task> tasks;
  for (uint32_t c = 0; c < S; c++)
    for (uint32_t a = 0; a < L; a += win) {
      const uint32_t b = (uint32_t)std::min<uint64_t>((uint64_t)a + win, L);
      const uint32_t cs = c * L;
      const uint32_t seed = (a == 0 && (c == 0 || is_dup)) ? kTaskSeedStart : 0u;
      tasks.push_back({kAllGroups, cs, cs + L, cs + a, cs + b, seed});
    }
  return tasks;
}

// Host work items -> device tasks, checking that each (group, chunk) unit is
// tiled exactly once (one sort over the work items, O(T log T)).
int host_tasks(const ngap_engine *eng, const ngap_run_opts &o, int memo_depth,
               std::vector<Task> &out) {
  const uint32_t G = (uint32_t)eng->img.num_groups, S = eng->num_seg, L = (uint32_t)eng->chunk_len;
  // (chunk, group or G for all groups, begin, end)
  struct Range {
    uint32_t chunk, group, begin, end;
  };
  std::vector<Range> rs;
  rs.reserve(o.num_tasks);
  out.clear();
  out.reserve(o.num_tasks);
  auto bad_tiling = [] {
    return ngap_fail(NGAP_ERR_INVALID, "work items must tile every (group, chunk) exactly once");
  };
  for (uint32_t i = 0; i < o.num_tasks; i++) {
    const ngap_task &t = o.tasks[i];
    const bool every = t.group == NGAP_ALL_GROUPS;
    if ((!every && t.group >= G) || t.chunk >= S || t.begin >= t.end || t.end > L)
      return ngap_fail(NGAP_ERR_INVALID, "work item " + std::to_string(i) + " out of range");
    if (every && memo_depth < 2)
      return ngap_fail(NGAP_ERR_INVALID, "all-group work items need precompute depth >= 2");
    rs.push_back({t.chunk, every ? G : t.group, t.begin, t.end});
    const uint32_t cs = t.chunk * L;
    const uint32_t seed = (t.begin == 0 && (t.chunk == 0 || o.is_dup)) ? kTaskSeedStart : 0u;
    out.push_back({every ? kAllGroups : t.group, cs, cs + L, cs + t.begin, cs + t.end, seed});
  }
  std::sort(rs.begin(), rs.end(), [](const Range &x, const Range &y) {
    if (x.chunk != y.chunk) return x.chunk < y.chunk;
    if (x.group != y.group) return x.group < y.group;
    return x.begin < y.begin;
  });
  // walk (chunk, group) runs: each must tile [0, L); a chunk covers all G
  // groups either with per-group runs or with one all-group run
  size_t i = 0;
  uint64_t covered = 0;
  while (i < rs.size()) {
    const uint32_t c = rs[i].chunk, g = rs[i].group;
    uint32_t at = 0;
    for (; i < rs.size() && rs[i].chunk == c && rs[i].group == g; i++) {
      if (rs[i].begin != at) return bad_tiling();
      at = rs[i].end;
    }
    if (at != L) return bad_tiling();
    covered += g == G ? G : 1;
  }
  // per-group and all-group runs in one chunk: detect by counting per chunk
  {
    size_t j = 0;
    while (j < rs.size()) {
      const uint32_t c = rs[j].chunk;
      bool per = false, all = false;
      for (; j < rs.size() && rs[j].chunk == c; j++) (rs[j].group == G ? all : per) = true;
      if (per && all)
        return ngap_fail(NGAP_ERR_INVALID,
                         "a chunk with all-group work items takes no per-group work items");
    }
  }
  if (covered != (uint64_t)G * S) return bad_tiling();
  return NGAP_OK;
}

// One workgroup per (group, chunk) unit: the non-blocking engine (k_plain) or
// the lock-step baseline (k_blocking, blockinggroups).
int run_plain(ngap_engine *eng, RunCtx &ctx, KParams P, const ngap_run_opts &o,
              ngap_run_stats *st) {
  const uint32_t G = (uint32_t)eng->img.num_groups, S = eng->num_seg;
  const uint32_t units = G * S;
  const bool blocking = o.algorithm == NGAP_ALGO_BLOCKING;
  P.ring_cap = kRing;  // these launches size their LDS for the default ring
  uint32_t spill = o.spill_capacity;
  if (!spill) {
    // default budget: 8 GiB of spill (or frontier) for the whole grid, at
    // most 64K entries per wave (per frontier half)
    uint64_t per = (8ull << 30) / 8 / ((uint64_t)units * P.waves_per_block);
    spill = (uint32_t)std::min<uint64_t>(65536, std::max<uint64_t>(1024, per));
  }
  P.spill_cap = spill;
  // k_plain: per-wave spill of u64 items; k_blocking: two u32 frontiers per unit
  HIP_TRY(ctx.spill.reserve(blocking ? (size_t)units * 2 * spill * 4
                                     : (size_t)units * P.waves_per_block * spill * 8));
  P.spill = ctx.spill.as<uint64_t>();
  HIP_TRY(ctx.overflow.reserve((size_t)units * 4));
  P.overflow = ctx.overflow.as<uint32_t>();
  attach_ctx(P, ctx);
  HIP_TRY(launch_reset(ctx.stream, P.report_count, P.items, P.task_head, P.overflow, units,
                       P.overflow_count, 0, nullptr));
  HIP_TRY(hipEventRecord(ctx.ev0, ctx.stream));  // times the matching kernel alone
  if (blocking) HIP_TRY(launch_blocking(ctx.stream, P, units, lds_bytes()));
  else HIP_TRY(launch_plain(ctx.stream, P, units, lds_bytes()));
  float ms = 0;
  unsigned long long cnt[4] = {0, 0, 0, 0};
  const int frc = ctx.finish(&ms, cnt, true);
  if (frc) return frc;
  const uint64_t overflowed = (uint32_t)cnt[3];
  ctx.last_count = cnt[0];
  ctx.last_report_off = o.report_off;
  ctx.have_run = true;
  ctx.last_tasks.clear();
  ctx.last_traces = false;
  ctx.tasks_resident = false;
  ctx.resident_sched = 0;
  if (st) {
    st->kernel_ms = ms;
    st->reports_raw = cnt[0];
    st->overflow_units = overflowed;
    st->items = cnt[1];
  }
  if (overflowed)
    return ngap_fail(NGAP_ERR_OVERFLOW, "worklist spill capacity exceeded in " +
                                            std::to_string(overflowed) + " work units");
  if (!o.report_off && cnt[0] > ctx.report_cap)
    return ngap_fail(NGAP_ERR_OVERFLOW, "report capacity exceeded: " + std::to_string(cnt[0]) +
                                            " reports > --result-capacity " +
                                            std::to_string(ctx.report_cap));
  return NGAP_OK;
}

}  // namespace
}  // namespace ngapamd

using namespace ngapamd;

extern "C" {

void ngap_default_opts(ngap_run_opts *o) {
  // defaults of the reference options (ngap_option.h:52-68, common_func.h:50-58)
  memset(o, 0, sizeof *o);
  o->algorithm = NGAP_ALGO_ALL;
  o->schedule = NGAP_SCHED_PLAIN;
  o->precompute_depth = 0;
  o->active_threshold = 20;
  o->unique = 0;
  o->data_buffer_fetch_size = 128;
  o->add_aas_start = 0;
  o->add_aas_interval = 1;
  o->result_capacity = 54619400ull;
  o->cross_groups = 1;
}

int ngap_create(int32_t device, ngap_engine **out) {
  return ngapamd::c_guard([&]() -> int {
    if (!out) return ngap_fail(NGAP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      return ngap_fail(NGAP_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= count) return ngap_fail(NGAP_ERR_INVALID, "bad device ordinal");
    std::unique_ptr<ngap_engine> e(new ngap_engine());
    e->device = device;
    HIP_TRY(hipSetDevice(device));
    int rc = e->ctx.init();
    if (rc) return rc;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    e->num_cus = prop.multiProcessorCount;
    HIP_TRY(set_lds_limit(lds_bytes(kMaxRing)));
    rc = noalloc_attach(device, &e->scratch);  // this device's scratch array (noalloc.cpp)
    if (rc) return rc;
    *out = e.release();
    return NGAP_OK;
  });
}

int ngap_destroy(ngap_engine *eng) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng) return NGAP_OK;
    (void)hipSetDevice(eng->device);
    if (eng->ctx.stream) (void)hipStreamSynchronize(eng->ctx.stream);
    delete eng;
    return NGAP_OK;
  });
}

namespace {

// Moves the image's arrays into one allocation (the arena): arrays of 2 MiB
// or more each start on a 2 MiB boundary, the small ones follow packed at
// 256 B.  The matching kernels gather from these arrays at random; with the
// arrays built among the builder's temporaries they ran up to 5 % slower
// than from one fresh allocation.  If the arena cannot be allocated next to
// the arrays (a very large image), they stay where they are.
int pack_image(ngap_engine *eng) {
  using ngapamd::DevBuf;
  DevBuf *arr[] = {&eng->d_succ4,  &eng->d_cls,     &eng->d_xs3,     &eng->d_xp3,
                   &eng->d_list[0], &eng->d_list[1], &eng->d_list[2], &eng->d_sel[0],
                   &eng->d_sel[1], &eng->d_sel[2],  &eng->d_beg[0],  &eng->d_beg[1],
                   &eng->d_beg[2], &eng->d_nrep[0], &eng->d_nrep[1], &eng->d_nrep[2],
                   &eng->d_xr1,    &eng->d_xe1,     &eng->d_xmask1,  &eng->d_xr2,
                   &eng->d_xe2,    &eng->d_xsod,    &eng->d_x2rep,   &eng->d_meta,
                   &eng->d_sod,    &eng->d_mask1};
  constexpr size_t kBig = 2u << 20, kSmall = 256;
  constexpr int kN = sizeof(arr) / sizeof(arr[0]);
  size_t off[kN], total = 0;
  for (int pass = 0; pass < 2; pass++)
    for (int i = 0; i < kN; i++) {
      if (!arr[i]->p || (arr[i]->n >= kBig) != (pass == 0)) continue;
      const size_t a = pass == 0 ? kBig : kSmall;
      total = (total + a - 1) / a * a;
      off[i] = total;
      total += arr[i]->n;
    }
  DevBuf arena;
  if (total == 0) return NGAP_OK;
  if (arena.reserve(total) != hipSuccess) {
    (void)hipGetLastError();
    return NGAP_OK;
  }
  char *base = arena.as<char>();
  for (int i = 0; i < kN; i++)
    if (arr[i]->p)
      HIP_TRY(hipMemcpyAsync(base + off[i], arr[i]->p, arr[i]->n, hipMemcpyDeviceToDevice,
                             eng->ctx.stream));
  HIP_TRY(hipStreamSynchronize(eng->ctx.stream));
  for (int i = 0; i < kN; i++) {
    if (!arr[i]->p) continue;
    const size_t n = arr[i]->n;
    arr[i]->release();
    arr[i]->p = base + off[i];
    arr[i]->n = n;
    arr[i]->owned = false;
  }
  std::swap(eng->d_arena.p, arena.p);  // the previous arena is freed on return
  std::swap(eng->d_arena.n, arena.n);
  return NGAP_OK;
}

}  // namespace

int ngap_load_groups_ex(ngap_engine *eng, const ngap_group_desc *groups, int32_t num_groups,
                        int32_t max_depth, int32_t build) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !groups) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (build != NGAP_BUILD_DEVICE && build != NGAP_BUILD_HOST)
      return ngap_fail(NGAP_ERR_INVALID, "build must be NGAP_BUILD_DEVICE or NGAP_BUILD_HOST");
    HIP_TRY(hipSetDevice(eng->device));
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<HostGroup> hg(num_groups > 0 ? num_groups : 0);
    for (int i = 0; i < num_groups; i++) {
      hg[i].num_states = groups[i].num_states;
      hg[i].row = groups[i].row_offsets;
      hg[i].col = groups[i].col_indices;
      hg[i].sets = groups[i].symbol_sets;
      hg[i].attrs = groups[i].node_attrs;
      hg[i].num_aas = groups[i].num_always_active;
      hg[i].aas = groups[i].always_active;
      hg[i].num_sod = groups[i].num_start;
      hg[i].sod = groups[i].start_states;
    }
    std::string err;
    eng->loaded = false;
    eng->generation++;
    eng->ctx.have_run = false;
    eng->ctx.drop_graph();
    const int depth = max_depth < 1 ? 1 : max_depth;
    if (build == NGAP_BUILD_DEVICE) {
      const int rc = build_image_device(eng, hg.data(), num_groups, depth, eng->ctx.stream,
                                        eng->img, eng->img_sizes, err);
      if (rc) return err.empty() ? rc : ngap_fail(rc, err);
    } else {
      if (!build_image(hg.data(), num_groups, depth, eng->img, err))
        return ngap_fail(NGAP_ERR_INVALID, err);
      Image &im = eng->img;
      HIP_TRY(eng->d_meta.upload(im.meta));
      HIP_TRY(eng->d_succ4.upload(im.succ4));
      HIP_TRY(eng->d_cls.upload(im.cls));
      HIP_TRY(eng->d_sod.upload(im.sod));
      HIP_TRY(eng->d_mask1.upload(im.mask1));
      HIP_TRY(eng->d_xr1.upload(im.xr1));
      HIP_TRY(eng->d_xe1.upload(im.xe1));
      HIP_TRY(eng->d_xmask1.upload(im.xmask1));
      HIP_TRY(eng->d_xr2.upload(im.xr2));
      HIP_TRY(eng->d_xe2.upload(im.xe2));
      HIP_TRY(eng->d_xsod.upload(im.xsod));
      HIP_TRY(eng->d_xs3.upload(im.xs3));
      HIP_TRY(eng->d_xp3.upload(im.xp3));
      HIP_TRY(eng->d_x2rep.upload(im.x2rep));
      for (int d = 0; d < kMaxMemoDepth; d++) {
        HIP_TRY(eng->d_beg[d].upload(im.memo[d].beg));
        HIP_TRY(eng->d_nrep[d].upload(im.memo[d].nrep));
        HIP_TRY(eng->d_list[d].upload(im.memo[d].list));
        HIP_TRY(eng->d_sel[d].upload(im.memo[d].sel));
      }
      eng->img_sizes = sizes_of(im);
      // the host keeps what the runtime reads (group metadata, root masks)
      Image kept;
      kept.num_groups = im.num_groups;
      kept.memo_depth = im.memo_depth;
      kept.meta.swap(im.meta);
      kept.mask1.swap(im.mask1);
      kept.n_xsod = im.n_xsod;
      kept.total_states = im.total_states;
      im = std::move(kept);
    }
    if (const int rc = pack_image(eng)) return rc;
    eng->built_on_device = build == NGAP_BUILD_DEVICE;
    eng->build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    eng->loaded = true;
    eng->generation++;
    return NGAP_OK;
  });
}

int ngap_load_groups(ngap_engine *eng, const ngap_group_desc *groups, int32_t num_groups,
                     int32_t max_depth) {
  return ngapamd::c_guard([&]() -> int {
    return ngap_load_groups_ex(eng, groups, num_groups, max_depth, NGAP_BUILD_DEVICE);
  });
}

int ngap_image_info(ngap_engine *eng, ngap_image_summary *info) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !info) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (!eng->loaded) return ngap_fail(NGAP_ERR_STATE, "no groups loaded");
    memset(info, 0, sizeof *info);
    info->built_on_device = eng->built_on_device ? 1 : 0;
    info->memo_depth = eng->img.memo_depth;
    info->num_groups = eng->img.num_groups;
    info->num_states = eng->img.total_states;
    info->bytes = eng->img_sizes.total();
    info->build_ms = eng->build_ms;
    return NGAP_OK;
  });
}

int ngap_image_array(ngap_engine *eng, const char *name, void *out, uint64_t cap,
                     uint64_t *n_bytes) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !name || !n_bytes) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (!eng->loaded) return ngap_fail(NGAP_ERR_STATE, "no groups loaded");
    const ImageSizes &z = eng->img_sizes;
    struct Entry {
      const char *name;
      const DevBuf *buf;
      uint64_t bytes;
    };
    const Entry table[] = {
        {"meta", &eng->d_meta, z.meta},       {"succ4", &eng->d_succ4, z.succ4},
        {"cls", &eng->d_cls, z.cls},          {"sod", &eng->d_sod, z.sod},
        {"beg1", &eng->d_beg[0], z.beg[0]},   {"beg2", &eng->d_beg[1], z.beg[1]},
        {"beg3", &eng->d_beg[2], z.beg[2]},   {"nrep1", &eng->d_nrep[0], z.nrep[0]},
        {"nrep2", &eng->d_nrep[1], z.nrep[1]}, {"nrep3", &eng->d_nrep[2], z.nrep[2]},
        {"list1", &eng->d_list[0], z.list[0]}, {"list2", &eng->d_list[1], z.list[1]},
        {"list3", &eng->d_list[2], z.list[2]}, {"sel2", &eng->d_sel[1], z.sel[1]},
        {"sel3", &eng->d_sel[2], z.sel[2]},   {"mask1", &eng->d_mask1, z.mask1},
        {"xr1", &eng->d_xr1, z.xr1},          {"xe1", &eng->d_xe1, z.xe1},
        {"xmask1", &eng->d_xmask1, z.xmask1}, {"xr2", &eng->d_xr2, z.xr2},
        {"xe2", &eng->d_xe2, z.xe2},          {"xsod", &eng->d_xsod, z.xsod},
        {"xs3", &eng->d_xs3, z.xs3},          {"xp3", &eng->d_xp3, z.xp3},
        {"x2rep", &eng->d_x2rep, z.x2rep}};
    for (const Entry &e : table) {
      if (strcmp(e.name, name) != 0) continue;
      *n_bytes = e.bytes;
      if (out && e.bytes) {
        HIP_TRY(hipSetDevice(eng->device));
        HIP_TRY(hipMemcpy(out, e.buf->p, std::min<uint64_t>(cap, e.bytes), hipMemcpyDeviceToHost));
      }
      return NGAP_OK;
    }
    return ngap_fail(NGAP_ERR_INVALID, std::string("unknown image array ") + name);
  });
}

int ngap_set_input(ngap_engine *eng, const uint8_t *data, uint64_t chunk_len,
                   uint32_t num_seg) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || (!data && chunk_len * num_seg)) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (chunk_len == 0 || num_seg == 0) return ngap_fail(NGAP_ERR_INVALID, "empty input");
    // offsets are 32-bit; keep headroom for the position cursors
    if (chunk_len * num_seg >= (1ull << 32) - (1ull << 20))
      return ngap_fail(NGAP_ERR_INVALID,
                       "input offsets are 32-bit: total input must be < 4 GiB - 1 MiB");
    HIP_TRY(hipSetDevice(eng->device));
    size_t n = (size_t)chunk_len * num_seg;
    HIP_TRY(eng->d_input.reserve(n + 16));
    // The input goes through two pinned blocks: host threads copy a block in
    // (counting its bytes for the static per-group schedule on the way) while
    // the other block's copy to the device runs.
    InputStage &stg = eng->stage;
    HIP_TRY(stg.init());
    HostPool &pool = *stg.pool;
    std::vector<std::array<uint64_t, 256>> hist(pool.width());
    for (auto &h : hist) h.fill(0);
    for (size_t off = 0, k = 0; off < n; off += InputStage::kBlock, k++) {
      const size_t len = std::min(InputStage::kBlock, n - off);
      uint8_t *st = stg.buf[k & 1];
      if (k >= 2) HIP_TRY(hipEventSynchronize(stg.done[k & 1]));  // its last copy is done
      const uint8_t *src = data + off;
      pool.run_parts(len, [&](size_t part, size_t a, size_t b) {
        std::copy(src + a, src + b, st + a);
        // four partial counts: no store-to-load dependence between neighbours
        uint32_t h[4][256] = {{0}};
        size_t i = a;
        for (; i + 4 <= b; i += 4) {
          h[0][src[i]]++;
          h[1][src[i + 1]]++;
          h[2][src[i + 2]]++;
          h[3][src[i + 3]]++;
        }
        for (; i < b; i++) h[0][src[i]]++;
        for (int x = 0; x < 256; x++) hist[part][x] += (uint64_t)h[0][x] + h[1][x] + h[2][x] + h[3][x];
      });
      HIP_TRY(hipMemcpyAsync(eng->d_input.as<uint8_t>() + off, st, len, hipMemcpyHostToDevice,
                             eng->ctx.stream));
      HIP_TRY(hipEventRecord(stg.done[k & 1], eng->ctx.stream));
    }
    HIP_TRY(hipStreamSynchronize(eng->ctx.stream));
    for (int b = 0; b < 256; b++) {
      uint64_t c = 0;
      for (auto &h : hist) c += h[b];
      eng->byte_freq[b] = (double)c / (double)n;
    }
    eng->chunk_len = chunk_len;
    eng->num_seg = num_seg;
    eng->have_input = true;
    eng->generation++;
    eng->ctx.have_run = false;
    // a captured graph stays: its launch parameters (input address, chunk
    // sizes) are compared with the next run's before it is replayed
    return NGAP_OK;
  });
}

int ngap_run(ngap_engine *eng, const ngap_run_opts *opts, ngap_run_stats *st) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !opts) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (st) memset(st, 0, sizeof *st);
    ngap_run_opts o = *opts;  // a copy: the default grid is resolved below
    if (!eng->loaded) return ngap_fail(NGAP_ERR_STATE, "no groups loaded");
    if (!eng->have_input) return ngap_fail(NGAP_ERR_STATE, "no input set");
    HIP_TRY(hipSetDevice(eng->device));
    Mapping mp;
    std::string err;
    if (!map_algorithm(o.algorithm, o.precompute_depth, eng->img.memo_depth, mp, err))
      return ngap_fail(NGAP_ERR_INVALID, err);
    if (!ring_of(o))
      return ngap_fail(NGAP_ERR_INVALID, "ring_capacity must be a power of two in [64, 2048]");
    const uint64_t units = (uint64_t)eng->img.num_groups * eng->num_seg;
    if (units >= (1u << 31)) return ngap_fail(NGAP_ERR_INVALID, "too many (group, chunk) units");
    RunCtx &ctx = eng->ctx;
    const uint64_t cap = o.result_capacity ? o.result_capacity : 54619400ull;
    if (!o.report_off) HIP_TRY(ctx.reports.reserve(cap * 8));
    ctx.report_cap = o.report_off ? 0 : cap;
    KParams P = base_params(eng, mp, o);
    P.input = eng->d_input.as<uint8_t>();
    P.chunk_len = (uint32_t)eng->chunk_len;
    P.chunk_magic = ~0ull / P.chunk_len + 1;  // wraps to 0 for 1-byte chunks, as the test needs
    P.num_seg = eng->num_seg;
    if (st) st->num_units = (uint32_t)units;
    if (o.schedule == NGAP_SCHED_PLAIN || o.algorithm == NGAP_ALGO_BLOCKING)
      return run_plain(eng, ctx, P, o, st);
    // the schedule: host work items, all-group windows or per-group windows,
    // rebuilt only when its inputs change
    ngap_engine::SchedCache &sc = eng->sched;
    // all-group windows wherever the cross-group index exists (depth >= 2),
    // one group included: a 3-gram slot per position replaces the group's
    // rank-table lookups (regex20k: 0.185 ms against 0.217)
    const int kind = o.tasks ? 0 : (mp.memo_depth >= 2 && o.cross_groups != 0) ? 1 : 2;
    // the default grid, sized to the job
    if (o.persistent_blocks == 0) {
      uint32_t full = 0;
      int rc = persistent_blocks(eng, 0, &full, P.ring_cap, engine_kind(o));
      if (rc) return rc;
      uint64_t work = 0;
      if (kind == 0) {
        for (uint32_t i = 0; i < o.num_tasks; i++)
          work += o.tasks[i].end > o.tasks[i].begin ? o.tasks[i].end - o.tasks[i].begin : 0;
      } else {
        work = (uint64_t)eng->num_seg * eng->chunk_len * (kind == 2 ? eng->img.num_groups : 1);
      }
      o.persistent_blocks = job_blocks(full, P.waves_per_block, work);
    }
    uint32_t waves = 0;
    if (kind != 0) {
      uint32_t blocks = 0;
      int rc = persistent_blocks(eng, o.persistent_blocks, &blocks, P.ring_cap, engine_kind(o));
      if (rc) return rc;
      waves = blocks * P.waves_per_block;
    }
    bool hit = sc.id && sc.generation == eng->generation && sc.kind == kind &&
               sc.is_dup == (o.is_dup != 0);
    if (hit && kind == 0)
      hit = sc.memo_depth == mp.memo_depth && sc.host.size() == o.num_tasks &&
            (o.num_tasks == 0 || (o.tasks_id && o.tasks_id == sc.id) ||
             memcmp(sc.host.data(), o.tasks, o.num_tasks * sizeof(ngap_task)) == 0);
    else if (hit)
      hit = sc.window == o.window && sc.waves == waves;
    if (!hit) {
      sc.id = 0;
      if (kind == 0) {
        int rc = host_tasks(eng, o, mp.memo_depth, sc.tasks);
        if (rc) return rc;
        sc.host.assign(o.tasks, o.tasks + o.num_tasks);
      } else if (kind == 1) {
        sc.tasks = cross_schedule(eng, o.window, o.is_dup != 0, waves);
      } else {
        sc.tasks = static_schedule(eng, o.window, o.is_dup != 0, waves);
      }
      sc.generation = eng->generation;
      sc.kind = kind;
      sc.is_dup = o.is_dup != 0;
      sc.memo_depth = mp.memo_depth;
      sc.window = o.window;
      sc.waves = waves;
      sc.id = next_schedule_id();
    }
    const int rc = run_tasks(eng, ctx, P, sc.tasks, sc.id, o, st);
    if (st) st->tasks_id = sc.id;
    return rc;
  });
}

int ngap_autotune(ngap_engine *eng, const ngap_run_opts *base, int32_t max_runs,
                  ngap_run_opts *best, double *best_ms) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !base || !best) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    if (!eng->loaded || !eng->have_input) return ngap_fail(NGAP_ERR_STATE, "no groups or input");
    HIP_TRY(hipSetDevice(eng->device));
    ngap_run_opts o = *base;
    o.schedule = NGAP_SCHED_BATCHED;
    o.fallback = 1;
    o.tasks = nullptr;
    o.num_tasks = 0;
    const uint32_t L = (uint32_t)eng->chunk_len;
    int runs = 0;
    double best_t = 1e300;
    ngap_run_opts cur = o;
    // time one setting: the faster of two runs (kernel plus fallback time)
    auto timeit = [&](const ngap_run_opts &c, double &t) -> int {
      t = 1e300;
      for (int r = 0; r < 2; r++) {
        ngap_run_stats st;
        int rc = ngap_run(eng, &c, &st);
        if (rc) return rc;
        t = std::min(t, st.kernel_ms + st.fallback_ms);
      }
      runs++;
      return NGAP_OK;
    };
    int rc = timeit(cur, best_t);
    if (rc) return rc;
    const std::vector<uint32_t> windows = {256, 512, 1024, 2048, 4096, 8192, 16384};
    const std::vector<uint32_t> rings = {128, 512, 1024};
    const std::vector<uint32_t> spills = {8192, 16384, 262144};
    // the reference's tuning values of --active-threshold (ngap.cu:1999)
    const std::vector<uint32_t> thresholds = {0, 8, 16, 24, 32};
    for (int axis = 0; axis < 6; axis++) {
      // the grid axis scales the full grid of the ring capacity picked so far
      uint32_t full = 0;
      if (axis == 2) {
        ngap_run_opts f = cur;
        f.persistent_blocks = 0;
        int rc2 = persistent_blocks(eng, 0, &full, ring_of(f), engine_kind(f));
        if (rc2) return rc2;
      }
      const std::vector<uint32_t> grids = {full / 4, full / 2, (full * 3) / 4};
      const std::vector<uint32_t> passes = {cur.root_pass ? 0u : 1u};
      const std::vector<uint32_t> &vals = axis == 0   ? windows
                                          : axis == 1 ? rings
                                          : axis == 2 ? grids
                                          : axis == 3 ? spills
                                          : axis == 4 ? passes
                                                      : thresholds;
      ngap_run_opts pick = cur;
      for (uint32_t v : vals) {
        if (runs >= max_runs) break;
        ngap_run_opts c = cur;
        if (axis == 0) {
          if (v > L) continue;
          c.window = v;
        } else if (axis == 1) {
          c.ring_capacity = v;
        } else if (axis == 2) {
          if (v == 0) continue;
          c.persistent_blocks = v;
        } else if (axis == 3) {
          c.spill_capacity = v;
        } else if (axis == 4) {
          c.root_pass = (int32_t)v;
        } else {
          if ((int32_t)v == cur.active_threshold) continue;
          c.active_threshold = (int32_t)v;
        }
        double t = 0;
        rc = timeit(c, t);
        if (rc) return rc;
        if (t < best_t) {
          best_t = t;
          pick = c;
        }
      }
      cur = pick;
    }
    *best = cur;
    if (best_ms) *best_ms = best_t;
    return NGAP_OK;
  });
}

int ngap_batched_grid(ngap_engine *eng, uint32_t ring_capacity, uint32_t *blocks,
                      uint32_t *waves_per_block) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !blocks || !waves_per_block) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    ngap_run_opts o;
    ngap_default_opts(&o);
    o.ring_capacity = ring_capacity;
    const uint32_t ring = ring_of(o);
    if (!ring) return ngap_fail(NGAP_ERR_INVALID, "ring_capacity must be a power of two in [64, 2048]");
    HIP_TRY(hipSetDevice(eng->device));
    int rc = persistent_blocks(eng, 0, blocks, ring, engine_kind(o));
    if (rc) return rc;
    *waves_per_block = kBlockThreads / kWave;
    return NGAP_OK;
  });
}

int ngap_get_reports(ngap_engine *eng, uint64_t *out, uint64_t cap, uint64_t *n_out) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !n_out) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    RunCtx &ctx = eng->ctx;
    if (!ctx.have_run) return ngap_fail(NGAP_ERR_STATE, "no run yet");
    if (ctx.last_report_off) return ngap_fail(NGAP_ERR_STATE, "last run had --report-off");
    if (ctx.last_count > ctx.report_cap)
      return ngap_fail(NGAP_ERR_OVERFLOW, "reports exceeded the capacity of the last run");
    HIP_TRY(hipSetDevice(eng->device));
    std::vector<uint64_t> v;
    int rc = sorted_reports(ctx, v);
    if (rc) return rc;
    *n_out = v.size();
    if (out) {
      uint64_t n = std::min<uint64_t>(cap, v.size());
      for (uint64_t i = 0; i < n; i++) out[i] = (v[i] << 32) | (v[i] >> 32);
    }
    return NGAP_OK;
  });
}

int ngap_get_task_items(ngap_engine *eng, uint64_t *out, uint64_t cap, uint64_t *n_out) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !n_out) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    RunCtx &ctx = eng->ctx;
    if (!ctx.have_run) return ngap_fail(NGAP_ERR_STATE, "no run yet");
    if (!ctx.last_traces)
      return ngap_fail(NGAP_ERR_STATE, "the last run did not record work-item traces (task_traces)");
    *n_out = ctx.last_tasks.size();
    if (out && !ctx.last_tasks.empty()) {
      uint64_t n = std::min<uint64_t>(cap, ctx.last_tasks.size());
      HIP_TRY(hipSetDevice(eng->device));
      HIP_TRY(hipMemcpy(out, ctx.task_items.p, n * 8, hipMemcpyDeviceToHost));
    }
    return NGAP_OK;
  });
}

int ngap_get_tasks(ngap_engine *eng, ngap_task *out, uint64_t cap, uint64_t *n_out) {
  return ngapamd::c_guard([&]() -> int {
    if (!eng || !n_out) return ngap_fail(NGAP_ERR_INVALID, "NULL argument");
    RunCtx &ctx = eng->ctx;
    if (!ctx.have_run) return ngap_fail(NGAP_ERR_STATE, "no run yet");
    *n_out = ctx.last_tasks.size();
    const uint32_t L = (uint32_t)eng->chunk_len;
    if (out)
      for (uint64_t i = 0; i < std::min<uint64_t>(cap, ctx.last_tasks.size()); i++) {
        const Task &t = ctx.last_tasks[i];
        out[i] = {t.group, L ? t.cstart / L : 0, t.begin - t.cstart, t.end - t.cstart};
      }
    return NGAP_OK;
  });
}

}  // extern "C"
