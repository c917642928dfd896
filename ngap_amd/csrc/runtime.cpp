// Host runtime behind the C ABI: device image upload, input upload, launch of
// the plain and batched schedules, overflow fallback, report retrieval.
//
// Reference counterparts: ngap::launch_non_blocking_*_groups
// (code/src/ngap/ngap.cu:284-1990), NonBlockingBuffer::init_nfagroups
// (code/src/ngap/ngap_buffer.cu:475-748) and the result pack/sort/unique
// (code/src/ngap/ngap.cu:1899-1933).  Differences by design: capacities are
// runtime values, start states are seeded on the device, device buffers live
// with the engine (resident across runs) instead of per launch, and errors are
// status codes instead of exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ngap_amd.h"
#include "kernels.h"
#include "layout.h"

namespace ngapamd {
hipError_t launch_reset(hipStream_t s, unsigned long long *rc, unsigned long long *items,
                        uint32_t *head, uint32_t *ovf, uint32_t n);
hipError_t launch_plain(hipStream_t s, const KParams &P, uint32_t units, uint32_t lds);
hipError_t launch_batched(hipStream_t s, const KParams &P, uint32_t blocks, uint32_t lds);
hipError_t batched_occupancy(int *blocks_per_cu, uint32_t lds);
hipError_t set_lds_limit(uint32_t lds);
}  // namespace ngapamd

using namespace ngapamd;

void ngap_set_last_error_internal(const std::string &m);  // errors.cpp

static int fail(int code, const std::string &msg) {
  ngap_set_last_error_internal(msg);
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return fail(_e == hipErrorOutOfMemory ? NGAP_ERR_OOM : NGAP_ERR_DEVICE,      \
                  std::string(#expr) + ": " + hipGetErrorString(_e));              \
  } while (0)

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Grows to at least `bytes` (contents not preserved).
  hipError_t reserve(size_t bytes) {
    if (bytes <= n && p) return hipSuccess;
    release();
    size_t want = bytes ? bytes : 8;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) n = want;
    else p = nullptr;
    return e;
  }
  template <class T>
  hipError_t upload(const std::vector<T> &v) {
    hipError_t e = reserve(v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    if (!v.empty()) e = hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    return e;
  }
  template <class T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

// LDS capacities per wave (entries)
constexpr uint32_t kRing = 512, kRbuf = 128, kDc = 256;

}  // namespace

struct ngap_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Image img;
  bool loaded = false;
  DevBuf d_meta, d_succ4, d_cls, d_sod, d_mask1, d_mask2;
  DevBuf d_beg[kMaxMemoDepth], d_nrep[kMaxMemoDepth], d_list[kMaxMemoDepth],
      d_key[kMaxMemoDepth], d_dir[kMaxMemoDepth];
  DevBuf d_input;
  uint64_t chunk_len = 0;
  uint32_t num_seg = 0;
  bool have_input = false;
  double byte_freq[256] = {0};   // input byte histogram (normalised)
  DevBuf d_reports, d_counters, d_overflow, d_spill, d_tasks, d_order, d_task_items;
  uint64_t report_cap = 0;
  uint64_t last_count = 0;
  bool last_report_off = false;
  bool have_run = false;
  uint32_t last_num_tasks = 0;
  int num_cus = 0;
  // captured graph of the last batched configuration
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> graph_key;
  ~ngap_engine() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

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
}

int ngap_create(int32_t device, ngap_engine **out) {
  if (!out) return fail(NGAP_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(NGAP_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(NGAP_ERR_INVALID, "bad device ordinal");
  std::unique_ptr<ngap_engine> e(new ngap_engine());
  e->device = device;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&e->ev0));
  HIP_TRY(hipEventCreate(&e->ev1));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  e->num_cus = prop.multiProcessorCount;
  uint32_t lds = 16 + (kBlockThreads / kWave) * wave_lds_bytes(kRing, kRbuf, kDc);
  HIP_TRY(set_lds_limit(lds));
  HIP_TRY(e->d_counters.reserve(64));
  *out = e.release();
  return NGAP_OK;
}

int ngap_destroy(ngap_engine *eng) {
  if (!eng) return NGAP_OK;
  (void)hipSetDevice(eng->device);
  if (eng->stream) (void)hipStreamSynchronize(eng->stream);
  delete eng;
  return NGAP_OK;
}

int ngap_load_groups(ngap_engine *eng, const ngap_group_desc *groups, int32_t num_groups,
                     int32_t max_depth) {
  if (!eng || !groups) return fail(NGAP_ERR_INVALID, "NULL argument");
  HIP_TRY(hipSetDevice(eng->device));
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
  if (!build_image(hg.data(), num_groups, max_depth < 1 ? 1 : max_depth, eng->img, err))
    return fail(NGAP_ERR_INVALID, err);
  const Image &im = eng->img;
  HIP_TRY(eng->d_meta.upload(im.meta));
  HIP_TRY(eng->d_succ4.upload(im.succ4));
  HIP_TRY(eng->d_cls.upload(im.cls));
  HIP_TRY(eng->d_sod.upload(im.sod));
  HIP_TRY(eng->d_mask1.upload(im.mask1));
  HIP_TRY(eng->d_mask2.upload(im.mask2));
  for (int d = 0; d < kMaxMemoDepth; d++) {
    HIP_TRY(eng->d_beg[d].upload(im.memo[d].beg));
    HIP_TRY(eng->d_nrep[d].upload(im.memo[d].nrep));
    HIP_TRY(eng->d_list[d].upload(im.memo[d].list));
    HIP_TRY(eng->d_key[d].upload(im.memo[d].key));
    HIP_TRY(eng->d_dir[d].upload(im.memo[d].dir));
  }
  eng->loaded = true;
  eng->have_run = false;
  if (eng->gexec) {
    (void)hipGraphExecDestroy(eng->gexec);
    eng->gexec = nullptr;
  }
  return NGAP_OK;
}

int ngap_set_input(ngap_engine *eng, const uint8_t *data, uint64_t chunk_len,
                   uint32_t num_seg) {
  if (!eng || (!data && chunk_len * num_seg)) return fail(NGAP_ERR_INVALID, "NULL argument");
  if (chunk_len == 0 || num_seg == 0) return fail(NGAP_ERR_INVALID, "empty input");
  // offsets are 32-bit; keep headroom for the position cursors
  if (chunk_len * num_seg >= (1ull << 32) - (1ull << 20))
    return fail(NGAP_ERR_INVALID, "input offsets are 32-bit: total input must be < 4 GiB - 1 MiB");
  HIP_TRY(hipSetDevice(eng->device));
  size_t n = (size_t)chunk_len * num_seg;
  HIP_TRY(eng->d_input.reserve(n + 16));
  HIP_TRY(hipMemcpy(eng->d_input.p, data, n, hipMemcpyHostToDevice));
  uint64_t hist[256] = {0};
  for (size_t i = 0; i < n; i++) hist[data[i]]++;
  for (int b = 0; b < 256; b++) eng->byte_freq[b] = (double)hist[b] / (double)n;
  eng->chunk_len = chunk_len;
  eng->num_seg = num_seg;
  eng->have_input = true;
  eng->have_run = false;
  if (eng->gexec) {
    (void)hipGraphExecDestroy(eng->gexec);
    eng->gexec = nullptr;
  }
  return NGAP_OK;
}

}  // extern "C"

namespace {

struct Mapping {
  int memo_depth;
  int priv_limit;
};

// --algorithm -> engine features (each name keeps the optimisations the
// reference adds at that step, main.cu:114-130 / README.md:141-150).
bool map_algorithm(const ngap_run_opts &o, int loaded_depth, Mapping &m, std::string &err) {
  int depth_req = o.precompute_depth;
  switch (o.algorithm) {
    case NGAP_ALGO_NAP:
    case NGAP_ALGO_NONBLOCKING:
      m = {1, 0};
      break;
    case NGAP_ALGO_R1:
      m = {1, 1};
      break;
    case NGAP_ALGO_R2:
      m = {1, 2};
      break;
    case NGAP_ALGO_PC:
      m = {depth_req < 1 ? 1 : depth_req, 0};
      break;
    case NGAP_ALGO_ALL:
      m = {depth_req < 1 ? 1 : depth_req, -1};
      break;
    case NGAP_ALGO_BLOCKING:
      err = "blockinggroups is not supported by this engine";
      return false;
    default:
      err = "unknown algorithm";
      return false;
  }
  if (m.memo_depth > loaded_depth) {
    err = "precompute depth exceeds the depth the groups were loaded with";
    return false;
  }
  if (m.memo_depth > kMaxMemoDepth) {
    err = "precompute depth must be <= 3";
    return false;
  }
  return true;
}

}  // namespace

static int run_impl(ngap_engine *eng, const ngap_run_opts &o, ngap_run_stats *st) {
  if (!eng->loaded) return fail(NGAP_ERR_STATE, "no groups loaded");
  if (!eng->have_input) return fail(NGAP_ERR_STATE, "no input set");
  Mapping mp;
  std::string err;
  if (!map_algorithm(o, eng->img.memo_depth, mp, err)) return fail(NGAP_ERR_INVALID, err);
  const Image &im = eng->img;
  const uint32_t G = (uint32_t)im.num_groups, S = eng->num_seg;
  const uint64_t units64 = (uint64_t)G * S;
  if (units64 >= (1u << 31)) return fail(NGAP_ERR_INVALID, "too many (group, chunk) units");
  const uint32_t units = (uint32_t)units64;
  const uint64_t cap = o.result_capacity ? o.result_capacity : 54619400ull;
  if (!o.report_off) HIP_TRY(eng->d_reports.reserve(cap * 8));
  eng->report_cap = o.report_off ? 0 : cap;

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
    P.img.key[d] = eng->d_key[d].as<uint32_t>();
    P.img.dir[d] = eng->d_dir[d].as<uint32_t>();
  }
  P.img.mask1 = eng->d_mask1.as<uint32_t>();
  P.img.mask2 = eng->d_mask2.as<uint32_t>();
  P.input = eng->d_input.as<uint8_t>();
  P.chunk_len = (uint32_t)eng->chunk_len;
  P.num_seg = S;
  P.num_groups = G;
  P.memo_depth = mp.memo_depth;
  P.is_dup = o.is_dup;
  P.dedup = o.unique;
  P.priv_limit = mp.priv_limit;
  P.report_off = o.report_off;
  P.ring_cap = kRing;
  P.rbuf_cap = kRbuf;
  P.dc_cap = kDc;
  P.root_batch = kWave;
  unsigned long long *counters = eng->d_counters.as<unsigned long long>();
  P.report_count = counters + 0;
  P.items = counters + 1;
  P.task_head = reinterpret_cast<uint32_t *>(counters + 2);
  P.reports = eng->d_reports.as<uint64_t>();
  P.report_cap = eng->report_cap;
  P.waves_per_block = kBlockThreads / kWave;
  const uint32_t lds = 16 + P.waves_per_block * wave_lds_bytes(kRing, kRbuf, kDc);

  uint32_t n_flags = 0;
  uint32_t blocks = 0;
  std::vector<Task> tasks;
  if (o.schedule == NGAP_SCHED_PLAIN) {
    n_flags = units;
    uint32_t spill = o.spill_capacity;
    if (!spill) {
      // default budget: 8 GiB of spill for the whole grid, at most 64K items
      uint64_t per = (8ull << 30) / 8 / ((uint64_t)units * P.waves_per_block);
      spill = (uint32_t)std::min<uint64_t>(65536, std::max<uint64_t>(1024, per));
    }
    P.spill_cap = spill;
    HIP_TRY(eng->d_spill.reserve((size_t)units * P.waves_per_block * spill * 8));
  } else {
    // Static schedule: a group whose roots are rare under the input's byte
    // histogram takes whole chunks per work item; others take `window`
    // positions.  Work items run longest-expected-first.
    uint32_t win = o.window ? o.window : 4096;
    if (win > P.chunk_len) win = P.chunk_len;
    std::vector<double> root_p(G, 0.0);
    for (uint32_t g = 0; g < G; g++)
      for (int b = 0; b < 256; b++)
        if ((im.mask1[g * 8 + b / 32] >> (b % 32)) & 1u) root_p[g] += eng->byte_freq[b];
    std::vector<double> cost;
    for (uint32_t c = 0; c < S; c++)
      for (uint32_t g = 0; g < G; g++) {
        uint32_t wg = root_p[g] < 1.0 / 16 ? P.chunk_len : win;
        for (uint32_t a = 0; a < P.chunk_len; a += wg) {
          uint32_t b = std::min<uint64_t>((uint64_t)a + wg, P.chunk_len);
          tasks.push_back({g, c, c * P.chunk_len + a, c * P.chunk_len + b});
          cost.push_back((b - a) * (0.02 + root_p[g]));
        }
      }
    std::vector<uint32_t> lpt(tasks.size());
    for (uint32_t i = 0; i < lpt.size(); i++) lpt[i] = i;
    std::stable_sort(lpt.begin(), lpt.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
    if (tasks.size() >= (1u << 31)) return fail(NGAP_ERR_INVALID, "too many work items");
    n_flags = (uint32_t)tasks.size();
    HIP_TRY(eng->d_tasks.upload(tasks));
    P.tasks = eng->d_tasks.as<Task>();
    P.num_tasks = (uint32_t)tasks.size();
    std::vector<uint32_t> ord = lpt;
    if (o.task_order) {
      if (o.task_order_len != tasks.size())
        return fail(NGAP_ERR_INVALID, "task_order length does not match the work items");
      ord.assign(o.task_order, o.task_order + o.task_order_len);
      std::vector<char> seen(tasks.size(), 0);
      for (uint32_t x : ord) {
        if (x >= tasks.size() || seen[x])
          return fail(NGAP_ERR_INVALID, "task_order must be a permutation of the work items");
        seen[x] = 1;
      }
    }
    HIP_TRY(eng->d_order.upload(ord));
    P.task_order = eng->d_order.as<uint32_t>();
    HIP_TRY(eng->d_task_items.reserve(tasks.size() * 8));
    P.task_items = eng->d_task_items.as<unsigned long long>();
    int per_cu = 0;
    HIP_TRY(batched_occupancy(&per_cu, lds));
    if (per_cu < 1) per_cu = 1;
    blocks = o.persistent_blocks ? o.persistent_blocks : (uint32_t)(per_cu * eng->num_cus);
    uint32_t spill = o.spill_capacity ? o.spill_capacity : 65536;
    P.spill_cap = spill;
    HIP_TRY(eng->d_spill.reserve((size_t)blocks * P.waves_per_block * spill * 8));
  }
  P.spill = eng->d_spill.as<uint64_t>();
  HIP_TRY(eng->d_overflow.reserve((size_t)std::max<uint32_t>(n_flags, 1) * 4));
  P.overflow = eng->d_overflow.as<uint32_t>();

  HIP_TRY(hipSetDevice(eng->device));
  HIP_TRY(hipEventRecord(eng->ev0, eng->stream));
  HIP_TRY(launch_reset(eng->stream, P.report_count, P.items, P.task_head, P.overflow, n_flags));
  if (o.schedule == NGAP_SCHED_PLAIN) HIP_TRY(launch_plain(eng->stream, P, units, lds));
  else HIP_TRY(launch_batched(eng->stream, P, blocks, lds));
  HIP_TRY(hipEventRecord(eng->ev1, eng->stream));
  HIP_TRY(hipEventSynchronize(eng->ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, eng->ev0, eng->ev1));

  std::vector<uint32_t> flags(n_flags);
  HIP_TRY(hipMemcpy(flags.data(), P.overflow, n_flags * 4ull, hipMemcpyDeviceToHost));
  uint64_t overflowed = 0;
  for (uint32_t f : flags) overflowed += f != 0;
  unsigned long long cnt[2] = {0, 0};
  HIP_TRY(hipMemcpy(cnt, counters, 16, hipMemcpyDeviceToHost));
  eng->last_count = cnt[0];
  eng->last_report_off = o.report_off;
  eng->have_run = true;
  eng->last_num_tasks = (uint32_t)tasks.size();
  if (st) {
    st->kernel_ms = ms;
    st->reports_raw = cnt[0];
    st->overflow_units = overflowed;
    st->items = cnt[1];
    st->num_units = units;
    st->num_tasks = (uint32_t)tasks.size();
  }
  if (overflowed)
    return fail(NGAP_ERR_OVERFLOW, "worklist spill capacity exceeded in " +
                                       std::to_string(overflowed) + " work units");
  if (!o.report_off && cnt[0] > cap)
    return fail(NGAP_ERR_OVERFLOW, "report capacity exceeded: " + std::to_string(cnt[0]) +
                                       " reports > --result-capacity " + std::to_string(cap));
  return NGAP_OK;
}

extern "C" {

int ngap_run(ngap_engine *eng, const ngap_run_opts *opts, ngap_run_stats *stats) {
  if (!eng || !opts) return fail(NGAP_ERR_INVALID, "NULL argument");
  if (stats) memset(stats, 0, sizeof *stats);
  return run_impl(eng, *opts, stats);
}

int ngap_get_reports(ngap_engine *eng, uint64_t *out, uint64_t cap, uint64_t *n_out) {
  if (!eng || !n_out) return fail(NGAP_ERR_INVALID, "NULL argument");
  if (!eng->have_run) return fail(NGAP_ERR_STATE, "no run yet");
  if (eng->last_report_off) return fail(NGAP_ERR_STATE, "last run had --report-off");
  if (eng->last_count > eng->report_cap)
    return fail(NGAP_ERR_OVERFLOW, "reports exceeded the capacity of the last run");
  HIP_TRY(hipSetDevice(eng->device));
  std::vector<uint64_t> v(eng->last_count);
  if (!v.empty())
    HIP_TRY(hipMemcpy(v.data(), eng->d_reports.p, v.size() * 8, hipMemcpyDeviceToHost));
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  *n_out = v.size();
  if (out) {
    uint64_t n = std::min<uint64_t>(cap, v.size());
    for (uint64_t i = 0; i < n; i++) out[i] = (v[i] << 32) | (v[i] >> 32);
  }
  return NGAP_OK;
}

int ngap_get_task_items(ngap_engine *eng, uint64_t *out, uint64_t cap, uint64_t *n_out) {
  if (!eng || !n_out) return fail(NGAP_ERR_INVALID, "NULL argument");
  if (!eng->have_run) return fail(NGAP_ERR_STATE, "no run yet");
  *n_out = eng->last_num_tasks;
  if (out && eng->last_num_tasks) {
    uint64_t n = std::min<uint64_t>(cap, eng->last_num_tasks);
    HIP_TRY(hipSetDevice(eng->device));
    HIP_TRY(hipMemcpy(out, eng->d_task_items.p, n * 8, hipMemcpyDeviceToHost));
  }
  return NGAP_OK;
}

}  // extern "C"
