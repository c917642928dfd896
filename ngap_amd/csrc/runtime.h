// Host-side engine state shared by the C ABI entry points (runtime.cpp) and
// the request server (server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include <memory>

#include "../../include/ngap_amd.h"
#include "host_pool.h"
#include "kernels.h"
#include "layout.h"

void ngap_set_last_error_internal(const std::string &m) noexcept;  // errors.cpp
int ngap_fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return ngap_fail(_e == hipErrorOutOfMemory ? NGAP_ERR_OOM : NGAP_ERR_DEVICE,    \
                       std::string(#expr) + ": " + hipGetErrorString(_e));            \
  } while (0)

namespace ngapamd {

// LDS capacities per wave (entries)
constexpr uint32_t kRing = 256, kRbuf = 128, kDc = 256;
constexpr uint32_t kMinRing = 64, kMaxRing = 2048;  // batched ring_capacity range

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  bool owned = true;  // false: a view into another buffer (the image arena)
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release();
  hipError_t reserve(size_t bytes);  // grows; contents not preserved
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

// Pinned staging of ngap_set_input: two blocks, so that host threads fill one
// while the other's copy to the device runs.
struct InputStage {
  static constexpr size_t kBlock = 8u << 20;
  uint8_t *buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};  // the block's copy to the device
  std::unique_ptr<HostPool> pool;
  InputStage() = default;
  InputStage(const InputStage &) = delete;
  InputStage &operator=(const InputStage &) = delete;
  ~InputStage() {
    for (int i = 0; i < 2; i++) {
      if (done[i]) (void)hipEventSynchronize(done[i]);
      if (buf[i]) (void)hipHostFree(buf[i]);
      if (done[i]) (void)hipEventDestroy(done[i]);
    }
  }
  // allocates what an earlier, failed call left missing
  hipError_t init() {
    for (int i = 0; i < 2; i++) {
      if (!buf[i]) {
        const hipError_t e =
            hipHostMalloc(reinterpret_cast<void **>(&buf[i]), kBlock, hipHostMallocDefault);
        if (e != hipSuccess) {
          buf[i] = nullptr;
          return e;
        }
      }
      if (!done[i]) {
        const hipError_t e = hipEventCreateWithFlags(&done[i], hipEventDisableTiming);
        if (e != hipSuccess) {
          done[i] = nullptr;
          return e;
        }
      }
    }
    if (!pool) pool.reset(new HostPool());
    return hipSuccess;
  }
};

// One stream with its own per-run buffers and captured graph.
struct RunCtx {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  unsigned long long *host_cnt = nullptr;  // pinned: the run's counters after finish()
  DevBuf reports, counters, overflow, spill, tasks, task_items, fb_tasks, fb_spill;
  // root pass: pieces of the resident work items, each work item's area
  // offsets (items), the areas and the per-work-item counts
  DevBuf pieces, root_off, root_items, root_count;
  uint32_t num_pieces = 0;
  uint32_t root_factor = 0;     // items per position the resident areas were sized for
  uint32_t root_boost = 1;      // doubles (up to 64x) after a root area overflowed,
  uint64_t root_gen = 0;        // for this engine generation (groups and input)
  uint64_t report_cap = 0, last_count = 0;
  bool last_report_off = false, have_run = false;
  bool tasks_resident = false;  // `tasks` holds last_tasks
  bool last_traces = false;     // the last run recorded items per work item
  uint64_t resident_sched = 0;  // schedule id of last_tasks (0 = unknown)
  std::vector<Task> last_tasks;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;
  RunCtx() = default;
  RunCtx(const RunCtx &) = delete;
  ~RunCtx();
  int init();
  void drop_graph();
  // Ends a timed launch begun with ev0: records ev1 (unless the launch
  // recorded it), queues the copy of the run counters into host_cnt behind
  // it and waits for the stream once.  *ms = ev0..ev1; cnt = reports, items
  // (the lanes summed), -, overflowed units.
  int finish(float *ms, unsigned long long cnt[4], bool record_end);
};

struct Mapping {
  int memo_depth;
  int priv_limit;
  int root_batch;  // positions per root step of per-group work
};

bool map_algorithm(int algorithm, int depth_req, int loaded_depth, Mapping &m, std::string &err);
uint32_t lds_bytes(uint32_t ring = kRing);

// Kernel launchers (kernels.hip).  launch_reset clears the run's counters
// and n per-work-item flags (and root counts, when given); launch_roots runs
// the root pass (k_roots) over P.pieces.
hipError_t launch_reset(hipStream_t s, unsigned long long *rc, unsigned long long *items,
                        uint32_t *head, uint32_t *ovf, uint32_t n, uint32_t *ovf_count,
                        int keep_reports, unsigned long long *root_count);
hipError_t launch_plain(hipStream_t s, const KParams &P, uint32_t units, uint32_t lds);
hipError_t launch_blocking(hipStream_t s, const KParams &P, uint32_t units, uint32_t lds);
hipError_t launch_roots(hipStream_t s, const KParams &P, uint32_t lds);
hipError_t launch_batched(hipStream_t s, const KParams &P, uint32_t blocks, uint32_t lds);
hipError_t batched_occupancy(int *blocks_per_cu, uint32_t lds, int kind);  // kind: batched_kind()
hipError_t set_lds_limit(uint32_t lds);

}  // namespace ngapamd

struct ngap_engine {
  int device = 0;
  int num_cus = 0;
  // blocks per CU by (log2(ring), batched kernel kind), 0 = not queried
  int batched_per_cu[12 * ngapamd::kBatchedKinds] = {0};
  ngapamd::Image img;
  bool loaded = false;
  ngapamd::DevBuf d_meta, d_succ4, d_cls, d_sod, d_mask1;
  ngapamd::DevBuf d_beg[ngapamd::kMaxMemoDepth], d_nrep[ngapamd::kMaxMemoDepth],
      d_list[ngapamd::kMaxMemoDepth], d_sel[ngapamd::kMaxMemoDepth];
  ngapamd::DevBuf d_xr1, d_xe1, d_xmask1, d_xr2, d_xe2, d_xsod, d_xs3, d_xp3, d_x2rep;
  // one allocation holding every image array above (they are views into it)
  ngapamd::DevBuf d_arena;
  ngapamd::DevBuf d_input;
  ngapamd::InputStage stage;    // pinned staging of ngap_set_input
  uint64_t chunk_len = 0;
  uint32_t num_seg = 0;
  bool have_input = false;
  double byte_freq[256] = {0};  // input byte histogram (normalised)
  ngapamd::RunCtx ctx;          // the context ngap_run uses
  // last schedule built by ngap_run, reused while its inputs are unchanged
  struct SchedCache {
    uint64_t id = 0;            // 0 = empty; ids are never reused
    uint64_t generation = 0;    // engine generation it was built for
    int kind = -1;              // 0 host work items, 1 all-group, 2 per-group
    uint32_t window = 0, waves = 0;
    int is_dup = 0, memo_depth = 0;
    std::vector<ngap_task> host;
    std::vector<ngapamd::Task> tasks;
  } sched;
  uint64_t generation = 1;      // bumped by ngap_load_groups / ngap_set_input
  void *scratch = nullptr;      // device address of this device's scratch array
                                // (ngap_run_noalloc)
  ngapamd::ImageSizes img_sizes;  // bytes of each device image array
  bool built_on_device = false;   // the last ngap_load_groups built on the device
  double build_ms = 0;            // its wall time
};

namespace ngapamd {
KParams base_params(const ngap_engine *eng, const Mapping &mp, const ngap_run_opts &o);
void attach_ctx(KParams &P, RunCtx &ctx);
int engine_kind(const ngap_run_opts &o);
// Persistent grid of the batched kernel `kind` (batched_occupancy) for ring.
int persistent_blocks(const ngap_engine *eng, uint32_t requested, uint32_t *blocks,
                      uint32_t ring = kRing, int kind = 1);
// The default persistent grid for a job of `work` positions (summed over the
// groups each work item scans): the full grid, or fewer workgroups when that
// would leave a wave fewer than kMinWavePositions positions.  A small job on
// the full grid spends its time on the waves' own set-up and claims
// (literal64, 1 MiB: 0.079 ms on 384 workgroups against 0.132 on 1,280).
constexpr uint64_t kMinWavePositions = 640;
uint32_t job_blocks(uint32_t full_blocks, uint32_t waves_per_block, uint64_t work);
uint32_t ring_of(const ngap_run_opts &o);
// Runs tasks on ctx (batched kernel), with the overflow fallback of o.
int run_tasks(ngap_engine *eng, RunCtx &ctx, KParams P, const std::vector<Task> &tasks,
              uint64_t sched_id,
              const ngap_run_opts &o, ngap_run_stats *st);
// Device reports of ctx's last run, sorted and unique (device encoding).
int sorted_reports(RunCtx &ctx, std::vector<uint64_t> &out);
// Builds the image of the groups on the device into eng's image buffers
// (build.hip): the arrays build_image makes, byte for byte.  img gets the
// parts the host reads (group metadata, root masks, counts), sz the arrays'
// sizes.  Returns a status; err holds the message.
int build_image_device(ngap_engine *eng, const HostGroup *groups, int32_t num_groups,
                       int32_t depth, hipStream_t s, Image &img, ImageSizes &sz,
                       std::string &err);
// With `device` current: the device address of that device's copy of the
// scratch array of ngap_run_noalloc (looked up once per device).
int noalloc_attach(int device, void **out);
}  // namespace ngapamd
