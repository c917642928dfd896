// Device-side structures shared by the HIP kernels and the host runtime.
#pragma once

#include <cstdint>

#include "layout.h"

#if defined(__HIPCC__)
#define NGAP_HD __host__ __device__
#else
#define NGAP_HD
#endif

namespace ngapamd {

struct DevImage {
  const GroupMeta *meta;
  const int32_t *succ4;      // kRecWords per record (ids, then tests; two int4 loads)
  const uint32_t *cls;       // 8 per state
  const int32_t *sod;
  const uint32_t *beg[kMaxMemoDepth];
  const uint32_t *nrep[kMaxMemoDepth];
  const int32_t *list[kMaxMemoDepth];
  const uint32_t *sel[kMaxMemoDepth];  // rank tables of depths 2 and 3
  const uint32_t *mask1;     // 16 words per group (layout.h)
  // cross-group root index (layout.h)
  const uint32_t *xr1, *xe1, *xmask1, *xr2, *xe2, *xsod, *xs3, *xp3, *x2rep;
  uint32_t n_xsod;
};

// One batched work item: roots at positions [begin, end) of the chunk
// [cstart, cend) scanned by group `group`; descendants run to cend.  Bit 0 of
// flags seeds the group's start-of-data states at cstart.
struct Task {
  uint32_t group, cstart, cend, begin, end, flags;
};
constexpr uint32_t kTaskSeedStart = 1;
// Work-item overflow flags: the worklist spill area filled up (the item's
// results are incomplete: fallback or error), or its root-pass area did (it
// is re-run without a root pass).
constexpr uint32_t kSpillOverflow = 1, kRootOverflow = 2;
// Task.group of a work item that scans its window for every group at once
// (roots from the cross-group index, memo depth >= 2).
constexpr uint32_t kAllGroups = 0xFFFFFFFFu;

// A piece of a work item for the root pass (k_roots): roots at positions
// [begin, end) of work item `task`; flags as Task.flags (the first piece of a
// work item carries its start-of-data seeding).
struct RootPiece {
  uint32_t task, begin, end, flags;
};

struct KParams {
  DevImage img;
  const uint8_t *input;
  uint32_t chunk_len, num_seg, num_groups;
  uint64_t chunk_magic;    // ceil(2^64 / chunk_len): divisibility test (chunk_len > 0)
  // Request runs (chunk_len 0): bit p set when position p is the last of its
  // request, so an item of any request finds its chunk end (nullable: then a
  // wave's current work item gives it)
  const uint32_t *last_map;
  int32_t memo_depth;      // 1..3, d-gram roots
  int32_t is_dup;
  int32_t dedup;
  int32_t priv_limit;      // in-lane continuation steps, -1 = unlimited
  int32_t active_threshold;  // continue in-lane only while more lanes than
                             // this are busy (--active-threshold)
  int32_t report_off;
  uint32_t ring_cap;       // LDS ring entries per wave (power of two)
  uint32_t rbuf_cap;       // LDS report buffer entries per wave
  uint32_t dc_cap;         // LDS dedup cache entries per wave (power of two)
  uint32_t root_batch;     // positions whose roots one step takes (per-group
                           // roots; 1 = NAPgroups' one position at a time)
  unsigned long long *report_count;
  uint64_t *reports;
  uint64_t report_cap;
  uint64_t *spill;
  uint32_t spill_cap;      // items per wave
  uint32_t *overflow;      // plain: per unit; batched: per task
  uint32_t *overflow_count;  // units / tasks that set their overflow flag
  unsigned long long *items;   // expanded items (statistic): kItemLanes lanes
  const Task *tasks;        // nullable: work items generated from their index
  // Generated work items (tasks == nullptr): chunks cut into windows of
  // gen_window positions; all-group work items (gen_cross) ordered by
  // (chunk, window), else one group per work item ordered by (chunk, group,
  // window).
  uint32_t gen_window;
  int32_t gen_cross;
  const uint32_t *task_order;  // nullable
  uint32_t num_tasks;
  uint32_t *task_head;     // kQueues claim counters, kQueueStride words apart
  unsigned long long *task_items;  // nullable, per task
  uint32_t waves_per_block;
  // Root pass (nullable root_items = off): k_roots writes the root items of
  // work item t to root_items[root_off[t] .. root_off[t + 1]), counting them
  // in root_count[t]; the streaming kernel then expands them instead of
  // computing roots.
  const RootPiece *pieces;
  uint32_t num_pieces;
  const uint64_t *root_off;
  uint64_t *root_items;
  unsigned long long *root_count;
};

constexpr uint32_t kWave = 64;
constexpr uint32_t kMaxOpen = 8;  // work items a wave overlaps (k_batched)
// The work queue is split by XCD: workgroup b runs on XCD b % 8, and queue q
// holds work items q, q + 8, q + 16, ...; each claim counter has its own
// 128-byte line.  A wave whose queue is empty takes from the next ones.
constexpr uint32_t kQueues = 8, kQueueStride = 32;
constexpr uint32_t kBlockThreads = 256;
// Counters of a run (one buffer per run context, cleared by k_reset): bytes
// 0-31 hold the report count and the count of overflowed work items (word
// 3), byte 128 on the queue claim counters, then one mask per XCD of the
// queues its waves found empty (one line each), then kItemLanes lines of
// expanded-item counts, one 64-bit word at the start of each line:
// the waves of workgroup b add theirs at lane b % kItemLanes and the host
// sums the lanes.  A single count that every wave adds to as it ends costs
// 5-8 % of a launch of the streaming kernel: atomics on one address
// serialise.
constexpr uint32_t kItemLanes = 32, kItemStride = 16;  // stride in 64-bit words
constexpr uint32_t kEmptyMasks = kQueues * kQueueStride;  // task_head word of XCD 0's mask
constexpr uint32_t kItemsWord = (128 + 2 * kQueues * kQueueStride * 4) / 8;
constexpr uint32_t kCounterBytes = (kItemsWord + kItemLanes * kItemStride) * 8;
NGAP_HD inline unsigned long long items_sum(const unsigned long long *counters) {
  unsigned long long n = 0;
  for (uint32_t k = 0; k < kItemLanes; k++) n += counters[kItemsWord + k * kItemStride];
  return n;
}

// Work item t of a generated schedule (KParams::gen_window).
NGAP_HD inline Task gen_task(const KParams &P, uint32_t t) {
  const uint32_t L = P.chunk_len, win = P.gen_window;
  const uint32_t per_chunk = (L + win - 1) / win;
  const uint32_t per_unit = P.gen_cross ? per_chunk : per_chunk * P.num_groups;
  const uint32_t c = t / per_unit, r = t % per_unit;
  const uint32_t g = P.gen_cross ? kAllGroups : r / per_chunk, w = r % per_chunk;
  Task T;
  T.group = g;
  T.cstart = c * L;
  T.cend = T.cstart + L;
  T.begin = T.cstart + w * win;
  T.end = T.begin + win < T.cend ? T.begin + win : T.cend;
  T.flags = (w == 0 && (c == 0 || P.is_dup)) ? kTaskSeedStart : 0u;
  return T;
}

// Batched kernel a launch of P runs: 0 k_batched (one open work item per
// wave: per-work-item traces, or request runs without a last-byte map), 1
// k_batched_stream of engine runs, 2 k_expand (after the root pass), 3
// k_batched_stream of request runs (last-byte map), 4 k_batched_stream of
// generated work items.
constexpr int kBatchedKinds = 5;
NGAP_HD inline int batched_kind(const KParams &P) {
  if (P.task_items || (!P.chunk_len && !P.last_map)) return 0;
  if (P.root_items) return 2;
  if (!P.chunk_len) return 3;
  return P.tasks ? 1 : 4;
}

// LDS bytes per wave for the given capacities
NGAP_HD inline uint32_t wave_lds_bytes(uint32_t ring, uint32_t rbuf, uint32_t dc) {
  return ring * 8 + rbuf * 8 + dc * 8 + (65 + 64 + 64 + 64 + 16 + kMaxOpen) * 4 + 4;
}

}  // namespace ngapamd
