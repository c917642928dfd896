// Device image builder: the arrays of build_image (layout.cpp, which stays
// the reference for every array: the GPU tests compare the two byte for byte)
// built by HIP kernels from the groups' CSR arrays uploaded once.
//
// Reference counterpart: the host-side prefix-memoization tables of
// ngap::getPrecomputeResultsForKGroupsInCsr / recursivePrecomputeForK
// (code/src/ngap/ngap.cu:2358-2513) and PrecTable::setVertices/setResults
// (code/src/commons/precompute_table.cu:25-395), which the reference builds
// on the host before its kernels run.
//
// The build is sort-based instead of recursive:
//  * successor lists: every non-all-input edge becomes a 64-bit key
//    (state << 22 | successor), sorted and de-duplicated: each state's
//    distinct successors in ascending order, whatever the CSR order;
//  * memo level d: every (group, d-gram, state) entry is a 64-bit key
//      group << 47 | d-gram << 23 | not-reporting << 22 | state
//    (level 1: every all-input state under every byte of its class; level
//    d+1: every entry of level d expanded by its successors and the bytes of
//    their classes), sorted and de-duplicated, so the keys come out in the
//    image's list order (group, d-gram, reporting first, state id) and a
//    list is a run of equal (group, d-gram);
//  * rank tables: presence bits set per list, ranks by an exclusive scan of
//    the words' popcounts;
//  * cross-group index: lists sorted by (d-gram, list index), which is group
//    order inside one d-gram; the 3-gram slots from runs of equal 3-grams.
// Sorts and scans are rocPRIM's device primitives; the rest are the kernels
// below (one thread per state, edge, entry, list or slot).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <cstring>
#include <thread>

#include "runtime.h"

namespace ngapamd {

namespace {

constexpr uint32_t kThreads = 256;
constexpr uint64_t kKeySentinel = ~0ull;
// entry keys: group << 47 | d-gram << 23 | not-reporting << 22 | state
constexpr int kKeyGramShift = 23, kKeyGroupShift = 47;
constexpr uint64_t kKeyStateMask = (1ull << 22) - 1;

// device error flags (bits)
constexpr uint32_t kErrRowOrder = 1, kErrColRange = 2, kErrAasRange = 4, kErrSodRange = 8;

inline uint32_t grid_for(uint64_t n) {
  uint64_t b = (n + kThreads - 1) / kThreads;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

#define GRID_LOOP(i, n) \
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (uint64_t)gridDim.x * blockDim.x)

__device__ __forceinline__ uint32_t class_popc(const uint32_t *cls, uint64_t s) {
  uint32_t c = 0;
#pragma unroll
  for (int w = 0; w < 8; w++) c += __popc(cls[s * 8 + w]);
  return c;
}

// ---- inputs: groups of states, per-state group ids, global row offsets ----

// state -> group (one block per group)
__global__ void k_state_group(const uint32_t *sb, uint16_t *sgrp, uint32_t G) {
  const uint32_t g = blockIdx.x;
  if (g >= G) return;
  for (uint32_t s = sb[g] + threadIdx.x; s < sb[g + 1]; s += blockDim.x) sgrp[s] = (uint16_t)g;
}

// Global row offsets from the per-group local ones (each group's n + 1
// entries follow each other in row_l); checks monotone rows, column range.
__global__ void k_rows(const int32_t *row_l, const uint32_t *sb, const uint64_t *eb,
                       const uint16_t *sgrp, uint64_t N, uint64_t *grow, uint32_t *err) {
  GRID_LOOP(s, N) {
    const uint32_t g = sgrp[s];
    const uint64_t at = s + g;  // group g's local rows start at sb[g] + g
    const int32_t r0 = row_l[at], r1 = row_l[at + 1];
    if (r0 < 0 || r0 > r1) atomicOr(err, kErrRowOrder);
    grow[s] = eb[g] + (uint64_t)(r0 < 0 ? 0 : r0);
  }
}

// End of state s's edges: the next state's start, or the group's edge end
// (row[n] of the group) for its last state.
__device__ __forceinline__ uint64_t edge_end(const uint64_t *grow, const uint32_t *sb,
                                             const uint64_t *eb, uint32_t g, uint64_t s) {
  return s + 1 == sb[g + 1] ? eb[g + 1] : grow[s + 1];
}

__global__ void k_mark_aas(const int32_t *aas_l, const uint64_t *ab, const uint32_t *sb,
                           uint32_t G, uint64_t A, uint8_t *is_aas, uint32_t *err) {
  GRID_LOOP(k, A) {
    uint32_t lo = 0, hi = G;  // group g with ab[g] <= k < ab[g + 1]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ab[mid] <= k) lo = mid;
      else hi = mid;
    }
    const int32_t v = aas_l[k];
    const uint32_t n = sb[lo + 1] - sb[lo];
    if (v < 0 || (uint32_t)v >= n) atomicOr(err, kErrAasRange);
    else is_aas[sb[lo] + v] = 1;
  }
}

// One key per edge: (state << 22 | successor) for non-all-input successors,
// the sentinel otherwise (sorted to the end).
__global__ void k_edge_keys(const uint64_t *grow, const int32_t *col, const uint32_t *sb,
                            const uint64_t *eb, const uint16_t *sgrp, const uint8_t *is_aas,
                            uint64_t N, uint64_t *keys, uint32_t *err) {
  GRID_LOOP(s, N) {
    const uint32_t g = sgrp[s];
    const uint32_t n = sb[g + 1] - sb[g];
    for (uint64_t e = grow[s], end = edge_end(grow, sb, eb, g, s); e < end; e++) {
      const int32_t u = col[e];
      uint64_t key = kKeySentinel;
      if (u < 0 || (uint32_t)u >= n) atomicOr(err, kErrColRange);
      else if (!is_aas[sb[g] + u]) key = (s << 22) | (uint64_t)u;
      keys[e] = key;
    }
  }
}

// Distinct successors: counts per state and in-degrees (global ids).
__global__ void k_succ_count(const uint64_t *dkeys, uint64_t D, const uint32_t *sb,
                             const uint16_t *sgrp, uint32_t *nsucc, uint32_t *indeg) {
  GRID_LOOP(j, D) {
    const uint64_t k = dkeys[j];
    const uint64_t s = k >> 22;
    atomicAdd(&nsucc[s], 1u);
    atomicAdd(&indeg[sb[sgrp[s]] + (k & kKeyStateMask)], 1u);
  }
}

// Tagged ids, successor tests and continuation-record counts (layout.h).
__global__ void k_tags(const uint32_t *sb, const uint16_t *sgrp, const uint8_t *attrs,
                       const uint32_t *cls, const uint32_t *nsucc, const uint32_t *indeg,
                       uint64_t N, int32_t *tag, uint32_t *test, uint32_t *ext) {
  GRID_LOOP(s, N) {
    const uint32_t v = (uint32_t)(s - sb[sgrp[s]]);
    int32_t t = (int32_t)v;
    if (attrs[s] & 2) t |= kFlagReport;
    if (nsucc[s] == 0) t |= kFlagDeadEnd;
    if (indeg[s] >= 2) t |= kFlagDedup;
    tag[s] = t;
    uint32_t cnt = 0, last = 0;
    for (int w = 0; w < 8; w++) {
      const uint32_t x = cls[s * 8 + w];
      if (x) last = w * 32 + 31 - __clz(x);
      cnt += __popc(x);
    }
    test[s] = cnt == 1 ? (kTestLiteral | last) : cnt == 256 ? kTestAll : 0u;
    const uint32_t d = nsucc[s];
    ext[s] = d <= 4 ? 0u : (d - 4 + 2) / 3;
  }
}

// Group record bases: group g's records start at sb[g] + (continuation
// records of the groups before it).
__global__ void k_rec_base(const uint32_t *sb, const uint64_t *ext_scan, uint32_t G,
                           uint64_t *rec_base) {
  GRID_LOOP(g, (uint64_t)G + 1) rec_base[g] = sb[g] + ext_scan[sb[g]];
}

// Successor records: primary record of state v at rec_base + v, its
// continuation records after the group's n primary records, in state order.
__global__ void k_records(const uint32_t *sb, const uint16_t *sgrp, const uint64_t *rec_base,
                          const uint64_t *ext_scan, const uint64_t *succ_off,
                          const uint64_t *dkeys, const int32_t *tag, const uint32_t *test,
                          uint64_t N, int32_t *succ4) {
  GRID_LOOP(s, N) {
    const uint32_t g = sgrp[s];
    const uint32_t v = (uint32_t)(s - sb[g]);
    const uint32_t n = sb[g + 1] - sb[g];
    const uint64_t cont0 = rec_base[g] + n + (ext_scan[s] - ext_scan[sb[g]]);
    const uint64_t d = succ_off[s + 1] - succ_off[s];
    uint64_t rec = rec_base[g] + v, next = cont0, k = 0;
    for (;;) {
      int32_t *R = succ4 + rec * kRecWords;
      const uint64_t left = d - k;
      const uint64_t here = left <= 4 ? left : 3;
      for (int i = 0; i < 4; i++) {
        int32_t id = -1;
        uint32_t t = 0;
        if ((uint64_t)i < here) {
          const uint64_t u = sb[g] + (dkeys[succ_off[s] + k + i] & kKeyStateMask);
          id = tag[u];
          t = test[u];
        }
        R[i] = id;
        R[4 + i] = (int32_t)t;
      }
      if (left <= 4) break;
      R[3] = -(int32_t)next - 2;
      k += 3;
      rec = next++;
    }
  }
}

__global__ void k_sod(const int32_t *sod_l, const uint64_t *qb, const uint32_t *sb,
                      const int32_t *tag, uint32_t G, uint64_t Q, int32_t *sod,
                      uint32_t *xsod, uint32_t *err) {
  GRID_LOOP(k, Q) {
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (qb[mid] <= k) lo = mid;
      else hi = mid;
    }
    const int32_t v = sod_l[k];
    const uint32_t n = sb[lo + 1] - sb[lo];
    int32_t t = 0;
    if (v < 0 || (uint32_t)v >= n) atomicOr(err, kErrSodRange);
    else t = tag[sb[lo] + v];
    sod[k] = t;
    xsod[2 * k] = (uint32_t)t;
    xsod[2 * k + 1] = lo;
  }
}

// ---- memo levels ----

__device__ __forceinline__ uint64_t entry_key(uint32_t g, uint32_t gram, bool rep, uint32_t v) {
  return ((uint64_t)g << kKeyGroupShift) | ((uint64_t)gram << kKeyGramShift) |
         ((uint64_t)(rep ? 0 : 1) << 22) | v;
}

// level 1: all-input entries k -> their class sizes (count pass) or keys
__global__ void k_l1(const int32_t *aas_l, const uint64_t *ab, const uint32_t *sb,
                     const uint8_t *attrs, const uint32_t *cls, uint32_t G, uint64_t A,
                     const uint64_t *off, uint32_t *cnt, uint64_t *keys) {
  GRID_LOOP(k, A) {
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ab[mid] <= k) lo = mid;
      else hi = mid;
    }
    const uint32_t v = (uint32_t)aas_l[k];
    const uint64_t s = sb[lo] + v;
    if (!keys) {
      cnt[k] = class_popc(cls, s);
      continue;
    }
    uint64_t o = off[k];
    const bool rep = attrs[s] & 2;
    for (int w = 0; w < 8; w++)
      for (uint32_t x = cls[s * 8 + w]; x; x &= x - 1)
        keys[o++] = entry_key(lo, w * 32 + __ffs(x) - 1, rep, v);
  }
}

// level d+1 from the entries of level d: each entry's successors under the
// bytes of their classes (count pass when keys is null)
__global__ void k_expand_level(const uint64_t *ent, uint64_t n, const uint32_t *sb,
                               const uint64_t *succ_off, const uint64_t *dkeys,
                               const uint8_t *attrs, const uint32_t *cls, const uint64_t *off,
                               uint32_t *cnt, uint64_t *keys) {
  GRID_LOOP(j, n) {
    const uint64_t e = ent[j];
    const uint32_t g = (uint32_t)(e >> kKeyGroupShift);
    const uint32_t gram = (uint32_t)(e >> kKeyGramShift) & 0xFFFFFFu;
    const uint64_t s = sb[g] + (e & kKeyStateMask);
    uint32_t c = 0;
    uint64_t o = keys ? off[j] : 0;
    for (uint64_t q = succ_off[s]; q < succ_off[s + 1]; q++) {
      const uint32_t u = (uint32_t)(dkeys[q] & kKeyStateMask);
      const uint64_t su = sb[g] + u;
      if (!keys) {
        c += class_popc(cls, su);
        continue;
      }
      const bool rep = attrs[su] & 2;
      for (int w = 0; w < 8; w++)
        for (uint32_t x = cls[su * 8 + w]; x; x &= x - 1)
          keys[o++] = entry_key(g, (gram << 8) | (w * 32 + __ffs(x) - 1), rep, u);
    }
    if (!keys) cnt[j] = c;
  }
}

// 1 where entry j starts a list (a new (group, d-gram))
__global__ void k_run_starts(const uint64_t *ent, uint64_t n, uint32_t *flag) {
  GRID_LOOP(j, n)
    flag[j] = (j == 0 || (ent[j] >> kKeyGramShift) != (ent[j - 1] >> kKeyGramShift)) ? 1u : 0u;
}

// Lists of a level from its sorted entries: list[j] = tagged state, list
// starts (beg), reporting entries (nrep), group and d-gram of each list.
__global__ void k_lists(const uint64_t *ent, uint64_t n, const uint32_t *run_scan,
                        const uint32_t *sb, const int32_t *tag, int32_t *list, uint32_t *beg,
                        uint32_t *nrep, uint32_t *lgrp, uint32_t *lgram, uint64_t nlists) {
  GRID_LOOP(j, n) {
    const uint64_t e = ent[j];
    const uint32_t g = (uint32_t)(e >> kKeyGroupShift);
    list[j] = tag[sb[g] + (e & kKeyStateMask)];
    const uint32_t li = run_scan[j];  // inclusive scan of the start flags, minus 1 below
    const bool start = j == 0 || (e >> kKeyGramShift) != (ent[j - 1] >> kKeyGramShift);
    if (start) {
      beg[li - 1] = (uint32_t)j;
      lgrp[li - 1] = g;
      lgram[li - 1] = (uint32_t)(e >> kKeyGramShift) & 0xFFFFFFu;
    }
    if (!((e >> 22) & 1u)) atomicAdd(&nrep[li - 1], 1u);
    if (j + 1 == n) beg[nlists] = (uint32_t)n;
  }
}

// Level 1 keeps all 256 lists of every group: counts per (group, byte).
__global__ void k_l1_counts(const uint64_t *ent, uint64_t n, uint32_t *cnt, uint32_t *rep) {
  GRID_LOOP(j, n) {
    const uint64_t e = ent[j];
    const uint32_t key = (uint32_t)(e >> kKeyGroupShift) * 256u +
                         ((uint32_t)(e >> kKeyGramShift) & 255u);
    atomicAdd(&cnt[key], 1u);
    if (!((e >> 22) & 1u)) atomicAdd(&rep[key], 1u);
  }
}

__global__ void k_l1_list(const uint64_t *ent, uint64_t n, const uint32_t *sb, const int32_t *tag,
                          int32_t *list) {
  GRID_LOOP(j, n) {
    const uint64_t e = ent[j];
    list[j] = tag[sb[(uint32_t)(e >> kKeyGroupShift)] + (e & kKeyStateMask)];
  }
}

// root masks: per group 16 words (bytes with a list, bytes whose list reports)
__global__ void k_mask1(const uint32_t *cnt, const uint32_t *rep, uint32_t G, uint32_t *mask1) {
  GRID_LOOP(k, (uint64_t)G * 256) {
    const uint32_t g = (uint32_t)(k >> 8), b = (uint32_t)(k & 255);
    if (cnt[k]) atomicOr(&mask1[g * 16 + b / 32], 1u << (b % 32));
    if (rep[k]) atomicOr(&mask1[g * 16 + 8 + b / 32], 1u << (b % 32));
  }
}

// rank tables: presence bits of each list's key in its pair
// (level 2: pair g * 2048 + 2-gram / 32; level 3: pair li2 * 8 + byte / 32)
__global__ void k_sel_bits(const uint32_t *lgrp, const uint32_t *lgram, uint64_t nlists, int level,
                           const uint32_t *sel2, uint32_t *words) {
  GRID_LOOP(li, nlists) {
    const uint32_t gram = lgram[li];
    uint64_t pair;
    if (level == 2) {
      pair = (uint64_t)lgrp[li] * 2048 + (gram >> 5);
    } else {
      const uint32_t k2 = gram >> 8;
      const uint2 w = reinterpret_cast<const uint2 *>(sel2)[(uint64_t)lgrp[li] * 2048 + (k2 >> 5)];
      const uint32_t li2 = w.y + __popc(w.x & ((1u << (k2 & 31)) - 1u));
      pair = (uint64_t)li2 * 8 + ((gram & 255) >> 5);
    }
    atomicOr(&words[pair], 1u << (gram & 31));
  }
}

__global__ void k_popc(const uint32_t *words, uint64_t n, uint32_t *out) {
  GRID_LOOP(i, n) out[i] = __popc(words[i]);
}

__global__ void k_sel_pack(const uint32_t *words, const uint32_t *ranks, uint64_t n,
                           uint32_t *sel) {
  GRID_LOOP(i, n) {
    sel[2 * i] = words[i];
    sel[2 * i + 1] = ranks[i];
  }
}

// ---- cross-group index ----

// x1: per byte, the depth-1 lists with reporting entries (g * 256 + b, g order)
__global__ void k_x1(const uint32_t *rep1, uint32_t G, const uint32_t *xr1, uint32_t *cnt,
                     uint32_t *xe1, uint32_t *xmask1) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= 256) return;
  uint32_t c = 0;
  for (uint32_t g = 0; g < G; g++)
    if (rep1[g * 256 + b]) {
      if (xe1) xe1[xr1[b] + c] = g * 256 + b;
      c++;
    }
  if (!xe1) {
    cnt[b] = c;
    if (c) atomicOr(&xmask1[b / 32], 1u << (b % 32));
  }
}

__global__ void k_sort_keys_of_lists(const uint32_t *lgram, uint64_t nlists, uint64_t *keys) {
  GRID_LOOP(li, nlists) keys[li] = ((uint64_t)lgram[li] << 32) | li;
}

__global__ void k_count_by_gram(const uint32_t *lgram, uint64_t nlists, uint32_t *cnt) {
  GRID_LOOP(li, nlists) atomicAdd(&cnt[lgram[li] + 1], 1u);
}

// x2: (list index, begin, end, reporting entries | group << 22) in 2-gram
// order, group order inside a 2-gram
__global__ void k_xe2(const uint64_t *sorted, uint64_t nlists, const uint32_t *beg,
                      const uint32_t *nrep, const uint32_t *lgrp, uint32_t *xe2) {
  GRID_LOOP(j, nlists) {
    const uint32_t li = (uint32_t)sorted[j];
    xe2[4 * j] = li;
    xe2[4 * j + 1] = beg[li];
    xe2[4 * j + 2] = beg[li + 1];
    xe2[4 * j + 3] = nrep[li] | (lgrp[li] << 22);
  }
}

__global__ void k_x2rep(const uint32_t *lgram, const uint32_t *nrep, uint64_t nlists,
                        uint32_t *x2rep) {
  GRID_LOOP(li, nlists)
    if (nrep[li]) atomicOr(&x2rep[lgram[li] >> 5], 1u << (lgram[li] & 31));
}

// x3: lists sorted by (3-gram, list index).  Run starts of equal 3-grams and
// each sorted list's pairs in xp3 (0 for a 3-gram held inline).
__global__ void k_x3_runs(const uint64_t *sorted, uint64_t n, const uint32_t *beg,
                          uint32_t *start, uint64_t *contrib) {
  GRID_LOOP(j, n) {
    const uint32_t k = (uint32_t)(sorted[j] >> 32);
    const bool first = j == 0 || (uint32_t)(sorted[j - 1] >> 32) != k;
    const bool last = j + 1 == n || (uint32_t)(sorted[j + 1] >> 32) != k;
    start[j] = first;
    const uint32_t li = (uint32_t)sorted[j];
    const uint32_t len = beg[li + 1] - beg[li];
    const bool inl = first && last && len == 1;
    contrib[j] = inl ? 0 : len;
  }
}

__global__ void k_x3_fill(const uint64_t *sorted, uint64_t n, const uint32_t *beg,
                          const int32_t *list3, const uint32_t *lgrp, const uint64_t *xpos,
                          uint32_t *xs3, uint32_t *xp3, uint32_t *err_big) {
  GRID_LOOP(j, n) {
    const uint32_t k = (uint32_t)(sorted[j] >> 32);
    const uint32_t li = (uint32_t)sorted[j];
    const bool first = j == 0 || (uint32_t)(sorted[j - 1] >> 32) != k;
    const bool last = j + 1 == n || (uint32_t)(sorted[j + 1] >> 32) != k;
    const uint32_t b0 = beg[li], len = beg[li + 1] - b0;
    if (first && last && len == 1) {
      xs3[2 * (uint64_t)k] = (uint32_t)list3[b0];
      xs3[2 * (uint64_t)k + 1] = kSlotInline | lgrp[li];
      continue;
    }
    const uint64_t p = xpos[j];
    for (uint32_t e = 0; e < len; e++) {
      xp3[2 * (p + e)] = (uint32_t)list3[b0 + e];
      xp3[2 * (p + e) + 1] = lgrp[li];
    }
    if (first) {
      // the run's pairs: from this list's first to the run end's last
      uint64_t end_j = j;
      while (end_j + 1 < n && (uint32_t)(sorted[end_j + 1] >> 32) == k) end_j++;
      const uint64_t total = xpos[end_j] + (beg[(uint32_t)sorted[end_j] + 1] -
                                            beg[(uint32_t)sorted[end_j]]) - p;
      if (p >= (1ull << 31) || total >= kSlotPrefixReports) atomicOr(err_big, 1u);
      xs3[2 * (uint64_t)k] = (uint32_t)p;
      xs3[2 * (uint64_t)k + 1] = (uint32_t)total;
    }
  }
}

__global__ void k_x3_flags(const uint32_t *xmask1, const uint32_t *x2rep, uint32_t *xs3) {
  GRID_LOOP(k, 1ull << 24) {
    const uint32_t b0 = (uint32_t)(k >> 16), k2 = (uint32_t)(k >> 8);
    if (((xmask1[b0 >> 5] >> (b0 & 31)) & 1u) || ((x2rep[k2 >> 5] >> (k2 & 31)) & 1u))
      xs3[2 * k + 1] |= kSlotPrefixReports;
  }
}

// ---- rocPRIM wrappers (temporary storage kept across calls) ----

struct Prim {
  hipStream_t s;
  DevBuf tmp;
  hipError_t sort(const uint64_t *in, uint64_t *out, uint64_t n, int bits) {
    if (n == 0) return hipSuccess;
    size_t bytes = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, bytes, in, out, n, 0, bits, s);
    if (e != hipSuccess) return e;
    if ((e = tmp.reserve(bytes)) != hipSuccess) return e;
    return rocprim::radix_sort_keys(tmp.p, bytes, in, out, n, 0, bits, s);
  }
  // out[i] = in[0] + ... + in[i-1], n + 1 entries (out[n] = total)
  template <class In, class Out>
  hipError_t exscan(const In *in, Out *out, uint64_t n) {
    size_t bytes = 0;
    hipError_t e =
        rocprim::exclusive_scan(nullptr, bytes, in, out, Out(0), n, rocprim::plus<Out>(), s);
    if (e != hipSuccess) return e;
    if ((e = tmp.reserve(bytes)) != hipSuccess) return e;
    if (n && (e = rocprim::exclusive_scan(tmp.p, bytes, in, out, Out(0), n,
                                          rocprim::plus<Out>(), s)) != hipSuccess)
      return e;
    return total(in, out, n);
  }
  // out[n] = out[n-1] + in[n-1]
  template <class In, class Out>
  hipError_t total(const In *in, Out *out, uint64_t n) {
    if (n == 0) return hipMemsetAsync(out, 0, sizeof(Out), s);
    Out a = 0;
    In b = 0;
    hipError_t e = hipMemcpyAsync(&a, out + n - 1, sizeof(Out), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&b, in + n - 1, sizeof(In), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    a += (Out)b;
    e = hipMemcpyAsync(out + n, &a, sizeof(Out), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
  }
  template <class In, class Out>
  hipError_t inscan(const In *in, Out *out, uint64_t n) {
    size_t bytes = 0;
    hipError_t e = rocprim::inclusive_scan(nullptr, bytes, in, out, n, rocprim::plus<Out>(), s);
    if (e != hipSuccess) return e;
    if ((e = tmp.reserve(bytes)) != hipSuccess) return e;
    return n ? rocprim::inclusive_scan(tmp.p, bytes, in, out, n, rocprim::plus<Out>(), s)
             : hipSuccess;
  }
  // unique of sorted keys into out; *count = kept
  hipError_t unique(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *count,
                    DevBuf &cnt_buf) {
    hipError_t e = cnt_buf.reserve(8);
    if (e != hipSuccess) return e;
    size_t bytes = 0;
    e = rocprim::unique(nullptr, bytes, in, out, cnt_buf.as<uint64_t>(), n,
                        rocprim::equal_to<uint64_t>(), s);
    if (e != hipSuccess) return e;
    if ((e = tmp.reserve(bytes)) != hipSuccess) return e;
    if (n == 0) {
      *count = 0;
      return hipSuccess;
    }
    e = rocprim::unique(tmp.p, bytes, in, out, cnt_buf.as<uint64_t>(), n,
                        rocprim::equal_to<uint64_t>(), s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(count, cnt_buf.p, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e;
  }
};

template <class T>
T read1(hipStream_t s, const T *p) {
  T v{};
  (void)hipMemcpyAsync(&v, p, sizeof(T), hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  return v;
}

}  // namespace

// Builds the device image of `groups` into eng's image buffers on stream s.
// Fills the host-side parts the runtime reads (meta, mask1, counts) into img.
int build_image_device(ngap_engine *eng, const HostGroup *hg, int32_t G, int32_t depth,
                       hipStream_t s, Image &img, ImageSizes &sz, std::string &err) {
  img = Image();
  sz = ImageSizes();
  if (G <= 0 || G > 1024) {
    err = "number of groups must be in [1, 1024] (10 group bits in report ids)";
    return NGAP_ERR_INVALID;
  }
  if (depth < 1) depth = 1;
  if (depth > kMaxMemoDepth) {
    err = "precompute depth must be <= 3";
    return NGAP_ERR_INVALID;
  }
  img.num_groups = G;
  img.memo_depth = depth;
  // group offsets (states, edges, all-input, start-of-data)
  std::vector<uint32_t> sb(G + 1, 0);
  std::vector<uint64_t> eb(G + 1, 0), ab(G + 1, 0), qb(G + 1, 0);
  for (int32_t g = 0; g < G; g++) {
    const int n = hg[g].num_states;
    if (n < 0 || n > kIdMask + 1) {
      err = "group " + std::to_string(g) + " has " + std::to_string(n) +
            " states; state ids are limited to 22 bits";
      return NGAP_ERR_INVALID;
    }
    if (n > 0 && (!hg[g].row || !hg[g].sets || !hg[g].attrs)) {
      err = "group arrays missing";
      return NGAP_ERR_INVALID;
    }
    const int E = n > 0 ? hg[g].row[n] : 0;
    if (E < 0) {
      err = "CSR row offsets are not monotone";
      return NGAP_ERR_INVALID;
    }
    if ((E > 0 && !hg[g].col) || (hg[g].num_aas > 0 && !hg[g].aas) ||
        (hg[g].num_sod > 0 && !hg[g].sod)) {
      err = "group arrays missing";
      return NGAP_ERR_INVALID;
    }
    if ((uint64_t)sb[g] + n >= (1ull << 31)) {
      err = "automata image too large";
      return NGAP_ERR_INVALID;
    }
    sb[g + 1] = sb[g] + n;
    eb[g + 1] = eb[g] + E;
    ab[g + 1] = ab[g] + std::max(0, hg[g].num_aas);
    qb[g + 1] = qb[g] + std::max(0, hg[g].num_sod);
  }
  const uint64_t N = sb[G], E = eb[G], A = ab[G], Q = qb[G];
  img.total_states = N;
  // upload: one pinned staging buffer, copied in by host threads
  const uint64_t o_row = 0, o_col = o_row + (N + G) * 4, o_sets = o_col + E * 4,
                 o_attr = o_sets + N * 32, o_aas = o_attr + ((N + 3) & ~3ull),
                 o_sod = o_aas + A * 4, o_end = o_sod + Q * 4;
  DevBuf d_in;
  HIP_TRY(d_in.reserve(o_end + 16));
  {
    uint8_t *pin = nullptr;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&pin), o_end + 16, hipHostMallocDefault));
    std::vector<std::thread> pool;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (unsigned t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        for (int32_t g = (int32_t)t; g < G; g += (int32_t)nt) {
          const HostGroup &h = hg[g];
          const int n = h.num_states;
          if (n > 0) {
            memcpy(pin + o_row + ((uint64_t)sb[g] + g) * 4, h.row, (size_t)(n + 1) * 4);
            memcpy(pin + o_sets + (uint64_t)sb[g] * 32, h.sets, (size_t)n * 32);
            memcpy(pin + o_attr + sb[g], h.attrs, (size_t)n);
          } else {
            const int32_t zero = 0;
            memcpy(pin + o_row + ((uint64_t)sb[g] + g) * 4, &zero, 4);
          }
          if (eb[g + 1] > eb[g])
            memcpy(pin + o_col + eb[g] * 4, h.col, (size_t)(eb[g + 1] - eb[g]) * 4);
          if (ab[g + 1] > ab[g])
            memcpy(pin + o_aas + ab[g] * 4, h.aas, (size_t)(ab[g + 1] - ab[g]) * 4);
          if (qb[g + 1] > qb[g])
            memcpy(pin + o_sod + qb[g] * 4, h.sod, (size_t)(qb[g + 1] - qb[g]) * 4);
        }
      });
    for (auto &t : pool) t.join();
    hipError_t e = hipMemcpyAsync(d_in.p, pin, o_end, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipHostFree(pin);
    HIP_TRY(e);
  }
  uint8_t *in = d_in.as<uint8_t>();
  const int32_t *row_l = reinterpret_cast<const int32_t *>(in + o_row);
  const int32_t *col = reinterpret_cast<const int32_t *>(in + o_col);
  const uint32_t *sets = reinterpret_cast<const uint32_t *>(in + o_sets);
  const uint8_t *attrs = in + o_attr;
  const int32_t *aas_l = reinterpret_cast<const int32_t *>(in + o_aas);
  const int32_t *sod_l = reinterpret_cast<const int32_t *>(in + o_sod);

  Prim prim{s, {}};
  DevBuf d_sb, d_eb, d_ab, d_qb, d_sgrp, d_grow, d_isaas, d_err, d_cnt8;
  HIP_TRY(d_sb.upload(sb));
  HIP_TRY(d_eb.upload(eb));
  HIP_TRY(d_ab.upload(ab));
  HIP_TRY(d_qb.upload(qb));
  HIP_TRY(d_sgrp.reserve(N * 2 + 2));
  HIP_TRY(d_grow.reserve(N * 8 + 8));
  HIP_TRY(d_isaas.reserve(N + 1));
  HIP_TRY(d_err.reserve(4));
  HIP_TRY(hipMemsetAsync(d_isaas.p, 0, N + 1, s));
  HIP_TRY(hipMemsetAsync(d_err.p, 0, 4, s));
  uint32_t *derr = d_err.as<uint32_t>();
  hipLaunchKernelGGL(k_state_group, dim3(G), dim3(kThreads), 0, s, d_sb.as<uint32_t>(),
                     d_sgrp.as<uint16_t>(), (uint32_t)G);
  hipLaunchKernelGGL(k_rows, dim3(grid_for(N)), dim3(kThreads), 0, s, row_l, d_sb.as<uint32_t>(),
                     d_eb.as<uint64_t>(), d_sgrp.as<uint16_t>(), N, d_grow.as<uint64_t>(), derr);
  hipLaunchKernelGGL(k_mark_aas, dim3(grid_for(A)), dim3(kThreads), 0, s, aas_l,
                     d_ab.as<uint64_t>(), d_sb.as<uint32_t>(), (uint32_t)G, A,
                     d_isaas.as<uint8_t>(), derr);
  HIP_TRY(hipGetLastError());
  {
    const uint32_t e0 = read1(s, derr);
    if (e0 & kErrRowOrder) err = "CSR row offsets are not monotone";
    else if (e0 & kErrAasRange) err = "all-input id out of range";
    if (e0) return NGAP_ERR_INVALID;
  }
  // distinct successors
  DevBuf d_ekeys, d_esorted, d_dkeys, d_nsucc, d_indeg, d_soff;
  HIP_TRY(d_ekeys.reserve(E * 8 + 8));
  HIP_TRY(d_esorted.reserve(E * 8 + 8));
  hipLaunchKernelGGL(k_edge_keys, dim3(grid_for(N)), dim3(kThreads), 0, s, d_grow.as<uint64_t>(),
                     col, d_sb.as<uint32_t>(), d_eb.as<uint64_t>(), d_sgrp.as<uint16_t>(),
                     d_isaas.as<uint8_t>(), N,
                     d_ekeys.as<uint64_t>(), derr);
  HIP_TRY(hipGetLastError());
  if (read1(s, derr) & kErrColRange) {
    err = "CSR column index out of range";
    return NGAP_ERR_INVALID;
  }
  HIP_TRY(prim.sort(d_ekeys.as<uint64_t>(), d_esorted.as<uint64_t>(), E, 64));
  uint64_t D = 0;
  HIP_TRY(d_dkeys.reserve(E * 8 + 8));
  HIP_TRY(prim.unique(d_esorted.as<uint64_t>(), d_dkeys.as<uint64_t>(), E, &D, d_cnt8));
  if (D && read1(s, d_dkeys.as<uint64_t>() + D - 1) == kKeySentinel) D--;
  d_ekeys.release();
  d_esorted.release();
  HIP_TRY(d_nsucc.reserve(N * 4 + 4));
  HIP_TRY(d_indeg.reserve(N * 4 + 4));
  HIP_TRY(hipMemsetAsync(d_nsucc.p, 0, N * 4 + 4, s));
  HIP_TRY(hipMemsetAsync(d_indeg.p, 0, N * 4 + 4, s));
  hipLaunchKernelGGL(k_succ_count, dim3(grid_for(D)), dim3(kThreads), 0, s,
                     d_dkeys.as<uint64_t>(), D, d_sb.as<uint32_t>(), d_sgrp.as<uint16_t>(),
                     d_nsucc.as<uint32_t>(), d_indeg.as<uint32_t>());
  HIP_TRY(d_soff.reserve((N + 1) * 8));
  HIP_TRY(prim.exscan(d_nsucc.as<uint32_t>(), d_soff.as<uint64_t>(), N));
  // tags, tests, records
  DevBuf d_tag, d_test, d_ext, d_extscan, d_recbase;
  HIP_TRY(d_tag.reserve(N * 4 + 4));
  HIP_TRY(d_test.reserve(N * 4 + 4));
  HIP_TRY(d_ext.reserve(N * 4 + 4));
  HIP_TRY(d_extscan.reserve((N + 1) * 8));
  HIP_TRY(d_recbase.reserve((G + 1) * 8));
  hipLaunchKernelGGL(k_tags, dim3(grid_for(N)), dim3(kThreads), 0, s, d_sb.as<uint32_t>(),
                     d_sgrp.as<uint16_t>(), attrs, sets, d_nsucc.as<uint32_t>(),
                     d_indeg.as<uint32_t>(), N, d_tag.as<int32_t>(), d_test.as<uint32_t>(),
                     d_ext.as<uint32_t>());
  HIP_TRY(prim.exscan(d_ext.as<uint32_t>(), d_extscan.as<uint64_t>(), N));
  hipLaunchKernelGGL(k_rec_base, dim3(grid_for(G + 1)), dim3(kThreads), 0, s,
                     d_sb.as<uint32_t>(), d_extscan.as<uint64_t>(), (uint32_t)G,
                     d_recbase.as<uint64_t>());
  std::vector<uint64_t> rec_base(G + 1);
  HIP_TRY(hipMemcpyAsync(rec_base.data(), d_recbase.p, (G + 1) * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const uint64_t R = rec_base[G];
  if (R >= (uint64_t)INT32_MAX / 2) {
    err = "automata image too large";
    return NGAP_ERR_INVALID;
  }
  HIP_TRY(eng->d_succ4.reserve(R * kRecWords * 4 + 4));
  hipLaunchKernelGGL(k_records, dim3(grid_for(N)), dim3(kThreads), 0, s, d_sb.as<uint32_t>(),
                     d_sgrp.as<uint16_t>(), d_recbase.as<uint64_t>(), d_extscan.as<uint64_t>(),
                     d_soff.as<uint64_t>(), d_dkeys.as<uint64_t>(), d_tag.as<int32_t>(),
                     d_test.as<uint32_t>(), N, eng->d_succ4.as<int32_t>());
  HIP_TRY(hipGetLastError());
  sz.succ4 = R * kRecWords * 4;
  d_ext.release();
  d_extscan.release();
  // class words: the uploaded sets, in state order
  HIP_TRY(eng->d_cls.reserve(N * 32 + 4));
  HIP_TRY(hipMemcpyAsync(eng->d_cls.p, sets, N * 32, hipMemcpyDeviceToDevice, s));
  sz.cls = N * 32;
  // start-of-data
  HIP_TRY(eng->d_sod.reserve(Q * 4 + 4));
  HIP_TRY(eng->d_xsod.reserve(Q * 8 + 8));
  hipLaunchKernelGGL(k_sod, dim3(grid_for(Q)), dim3(kThreads), 0, s, sod_l, d_qb.as<uint64_t>(),
                     d_sb.as<uint32_t>(), d_tag.as<int32_t>(), (uint32_t)G, Q,
                     eng->d_sod.as<int32_t>(), eng->d_xsod.as<uint32_t>(), derr);
  HIP_TRY(hipGetLastError());
  if (read1(s, derr) & kErrSodRange) {
    err = "start-of-data id out of range";
    return NGAP_ERR_INVALID;
  }
  sz.sod = Q * 4;
  sz.xsod = Q * 8;
  img.n_xsod = Q;
  // group metadata (host)
  img.meta.resize(G);
  for (int32_t g = 0; g < G; g++) {
    GroupMeta &m = img.meta[g];
    memset(&m, 0, sizeof m);
    m.rec_base = (uint32_t)rec_base[g];
    m.cls_base = sb[g];
    m.sod_begin = (uint32_t)qb[g];
    m.sod_end = (uint32_t)qb[g + 1];
    m.num_states = (uint32_t)hg[g].num_states;
    m.num_aas = (uint32_t)std::max(0, hg[g].num_aas);
  }
  HIP_TRY(eng->d_meta.upload(img.meta));
  sz.meta = (uint64_t)G * sizeof(GroupMeta);

  // memo level 1
  DevBuf d_cnt, d_off, d_keys, d_sorted, d_ent[kMaxMemoDepth];
  uint64_t nent[kMaxMemoDepth] = {0, 0, 0};
  HIP_TRY(d_cnt.reserve(A * 4 + 4));
  HIP_TRY(d_off.reserve((A + 1) * 8));
  hipLaunchKernelGGL(k_l1, dim3(grid_for(A)), dim3(kThreads), 0, s, aas_l, d_ab.as<uint64_t>(),
                     d_sb.as<uint32_t>(), attrs, sets, (uint32_t)G, A, nullptr,
                     d_cnt.as<uint32_t>(), nullptr);
  HIP_TRY(prim.exscan(d_cnt.as<uint32_t>(), d_off.as<uint64_t>(), A));
  uint64_t T = read1(s, d_off.as<uint64_t>() + A);
  HIP_TRY(d_keys.reserve(T * 8 + 8));
  HIP_TRY(d_sorted.reserve(T * 8 + 8));
  hipLaunchKernelGGL(k_l1, dim3(grid_for(A)), dim3(kThreads), 0, s, aas_l, d_ab.as<uint64_t>(),
                     d_sb.as<uint32_t>(), attrs, sets, (uint32_t)G, A, d_off.as<uint64_t>(),
                     nullptr, d_keys.as<uint64_t>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(prim.sort(d_keys.as<uint64_t>(), d_sorted.as<uint64_t>(), T, 57));
  HIP_TRY(d_ent[0].reserve(T * 8 + 8));
  HIP_TRY(prim.unique(d_sorted.as<uint64_t>(), d_ent[0].as<uint64_t>(), T, &nent[0], d_cnt8));
  {
    const uint64_t K1 = (uint64_t)G * 256;
    DevBuf d_c1, d_r1;
    HIP_TRY(d_c1.reserve(K1 * 4));
    HIP_TRY(d_r1.reserve(K1 * 4));
    HIP_TRY(hipMemsetAsync(d_c1.p, 0, K1 * 4, s));
    HIP_TRY(hipMemsetAsync(d_r1.p, 0, K1 * 4, s));
    hipLaunchKernelGGL(k_l1_counts, dim3(grid_for(nent[0])), dim3(kThreads), 0, s,
                       d_ent[0].as<uint64_t>(), nent[0], d_c1.as<uint32_t>(),
                       d_r1.as<uint32_t>());
    HIP_TRY(eng->d_beg[0].reserve((K1 + 1) * 4));
    HIP_TRY(prim.exscan(d_c1.as<uint32_t>(), eng->d_beg[0].as<uint32_t>(), K1));
    HIP_TRY(eng->d_nrep[0].reserve(K1 * 4));
    HIP_TRY(hipMemcpyAsync(eng->d_nrep[0].p, d_r1.p, K1 * 4, hipMemcpyDeviceToDevice, s));
    HIP_TRY(eng->d_list[0].reserve(nent[0] * 4 + 4));
    hipLaunchKernelGGL(k_l1_list, dim3(grid_for(nent[0])), dim3(kThreads), 0, s,
                       d_ent[0].as<uint64_t>(), nent[0], d_sb.as<uint32_t>(),
                       d_tag.as<int32_t>(), eng->d_list[0].as<int32_t>());
    HIP_TRY(eng->d_mask1.reserve((uint64_t)G * 64));
    HIP_TRY(hipMemsetAsync(eng->d_mask1.p, 0, (uint64_t)G * 64, s));
    hipLaunchKernelGGL(k_mask1, dim3(grid_for(K1)), dim3(kThreads), 0, s, d_c1.as<uint32_t>(),
                       d_r1.as<uint32_t>(), (uint32_t)G, eng->d_mask1.as<uint32_t>());
    // cross-group byte index
    DevBuf d_xc;
    HIP_TRY(d_xc.reserve(256 * 4));
    HIP_TRY(eng->d_xr1.reserve(257 * 4));
    HIP_TRY(eng->d_xmask1.reserve(8 * 4));
    HIP_TRY(hipMemsetAsync(eng->d_xmask1.p, 0, 32, s));
    hipLaunchKernelGGL(k_x1, dim3(1), dim3(256), 0, s, d_r1.as<uint32_t>(), (uint32_t)G,
                       nullptr, d_xc.as<uint32_t>(), nullptr, eng->d_xmask1.as<uint32_t>());
    HIP_TRY(prim.exscan(d_xc.as<uint32_t>(), eng->d_xr1.as<uint32_t>(), 256));
    const uint32_t nx1 = read1(s, eng->d_xr1.as<uint32_t>() + 256);
    HIP_TRY(eng->d_xe1.reserve((uint64_t)nx1 * 4 + 4));
    hipLaunchKernelGGL(k_x1, dim3(1), dim3(256), 0, s, d_r1.as<uint32_t>(), (uint32_t)G,
                       eng->d_xr1.as<uint32_t>(), nullptr, eng->d_xe1.as<uint32_t>(), nullptr);
    HIP_TRY(hipGetLastError());
    sz.beg[0] = (K1 + 1) * 4;
    sz.nrep[0] = K1 * 4;
    sz.list[0] = nent[0] * 4;
    sz.sel[0] = 0;
    sz.mask1 = (uint64_t)G * 64;
    sz.xr1 = 257 * 4;
    sz.xe1 = (uint64_t)nx1 * 4;
    sz.xmask1 = 32;
    img.mask1.resize((size_t)G * 16);
    HIP_TRY(hipMemcpyAsync(img.mask1.data(), eng->d_mask1.p, (uint64_t)G * 64,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  // levels 2 and 3
  DevBuf d_lgrp[kMaxMemoDepth], d_lgram[kMaxMemoDepth];
  uint64_t nlists[kMaxMemoDepth] = {(uint64_t)G * 256, 0, 0};
  for (int d = 2; d <= depth; d++) {
    const uint64_t n = nent[d - 2];
    const uint64_t *prev = d_ent[d - 2].as<uint64_t>();
    HIP_TRY(d_cnt.reserve(n * 4 + 4));
    HIP_TRY(d_off.reserve((n + 1) * 8));
    hipLaunchKernelGGL(k_expand_level, dim3(grid_for(n)), dim3(kThreads), 0, s, prev, n,
                       d_sb.as<uint32_t>(), d_soff.as<uint64_t>(), d_dkeys.as<uint64_t>(), attrs,
                       sets, nullptr, d_cnt.as<uint32_t>(), nullptr);
    HIP_TRY(prim.exscan(d_cnt.as<uint32_t>(), d_off.as<uint64_t>(), n));
    T = read1(s, d_off.as<uint64_t>() + n);
    HIP_TRY(d_keys.reserve(T * 8 + 8));
    HIP_TRY(d_sorted.reserve(T * 8 + 8));
    hipLaunchKernelGGL(k_expand_level, dim3(grid_for(n)), dim3(kThreads), 0, s, prev, n,
                       d_sb.as<uint32_t>(), d_soff.as<uint64_t>(), d_dkeys.as<uint64_t>(), attrs,
                       sets, d_off.as<uint64_t>(), nullptr, d_keys.as<uint64_t>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(prim.sort(d_keys.as<uint64_t>(), d_sorted.as<uint64_t>(), T, 57));
    HIP_TRY(d_ent[d - 1].reserve(T * 8 + 8));
    HIP_TRY(prim.unique(d_sorted.as<uint64_t>(), d_ent[d - 1].as<uint64_t>(), T, &nent[d - 1],
                        d_cnt8));
    const uint64_t m = nent[d - 1];
    if (m >= (1ull << 31)) {
      err = "memo tables too large (2^31 entries); use a smaller precompute depth";
      return NGAP_ERR_INVALID;
    }
    const uint64_t *ent = d_ent[d - 1].as<uint64_t>();
    // lists: runs of equal (group, d-gram)
    DevBuf d_flag, d_runscan;
    HIP_TRY(d_flag.reserve(m * 4 + 4));
    HIP_TRY(d_runscan.reserve(m * 4 + 4));
    hipLaunchKernelGGL(k_run_starts, dim3(grid_for(m)), dim3(kThreads), 0, s, ent, m,
                       d_flag.as<uint32_t>());
    HIP_TRY(prim.inscan(d_flag.as<uint32_t>(), d_runscan.as<uint32_t>(), m));
    const uint64_t L = m ? read1(s, d_runscan.as<uint32_t>() + m - 1) : 0;
    nlists[d - 1] = L;
    HIP_TRY(eng->d_list[d - 1].reserve(m * 4 + 4));
    HIP_TRY(eng->d_beg[d - 1].reserve((L + 1) * 4));
    HIP_TRY(eng->d_nrep[d - 1].reserve(L * 4 + 4));
    HIP_TRY(d_lgrp[d - 1].reserve(L * 4 + 4));
    HIP_TRY(d_lgram[d - 1].reserve(L * 4 + 4));
    HIP_TRY(hipMemsetAsync(eng->d_nrep[d - 1].p, 0, L * 4 + 4, s));
    HIP_TRY(hipMemsetAsync(eng->d_beg[d - 1].p, 0, 4, s));
    hipLaunchKernelGGL(k_lists, dim3(grid_for(m)), dim3(kThreads), 0, s, ent, m,
                       d_runscan.as<uint32_t>(), d_sb.as<uint32_t>(), d_tag.as<int32_t>(),
                       eng->d_list[d - 1].as<int32_t>(), eng->d_beg[d - 1].as<uint32_t>(),
                       eng->d_nrep[d - 1].as<uint32_t>(), d_lgrp[d - 1].as<uint32_t>(),
                       d_lgram[d - 1].as<uint32_t>(), L);
    HIP_TRY(hipGetLastError());
    // rank table
    const uint64_t pairs = d == 2 ? (uint64_t)G * 2048 : nlists[1] * 8;
    DevBuf d_words, d_pc, d_ranks;
    HIP_TRY(d_words.reserve(pairs * 4 + 4));
    HIP_TRY(d_pc.reserve(pairs * 4 + 4));
    HIP_TRY(d_ranks.reserve((pairs + 1) * 4));
    HIP_TRY(hipMemsetAsync(d_words.p, 0, pairs * 4 + 4, s));
    hipLaunchKernelGGL(k_sel_bits, dim3(grid_for(L)), dim3(kThreads), 0, s,
                       d_lgrp[d - 1].as<uint32_t>(), d_lgram[d - 1].as<uint32_t>(), L, d,
                       eng->d_sel[1].as<uint32_t>(), d_words.as<uint32_t>());
    hipLaunchKernelGGL(k_popc, dim3(grid_for(pairs)), dim3(kThreads), 0, s,
                       d_words.as<uint32_t>(), pairs, d_pc.as<uint32_t>());
    HIP_TRY(prim.exscan(d_pc.as<uint32_t>(), d_ranks.as<uint32_t>(), pairs));
    HIP_TRY(eng->d_sel[d - 1].reserve(pairs * 8 + 8));
    hipLaunchKernelGGL(k_sel_pack, dim3(grid_for(pairs)), dim3(kThreads), 0, s,
                       d_words.as<uint32_t>(), d_ranks.as<uint32_t>(), pairs,
                       eng->d_sel[d - 1].as<uint32_t>());
    HIP_TRY(hipGetLastError());
    sz.beg[d - 1] = (L + 1) * 4;
    sz.nrep[d - 1] = L * 4;
    sz.list[d - 1] = m * 4;
    sz.sel[d - 1] = pairs * 8;
  }
  // levels above the depth: one list offset (0), as the host image has
  for (int d = depth + 1; d <= kMaxMemoDepth; d++) {
    HIP_TRY(eng->d_beg[d - 1].reserve(4));
    HIP_TRY(hipMemsetAsync(eng->d_beg[d - 1].p, 0, 4, s));
    sz.beg[d - 1] = 4;
  }
  // the entry keys of the deepest level are no longer needed
  for (auto &b : d_ent) b.release();
  d_keys.release();
  d_sorted.release();
  if (depth >= 2) {
    const uint64_t L2 = nlists[1];
    DevBuf d_k, d_o, d_c;
    HIP_TRY(d_k.reserve(L2 * 8 + 8));
    HIP_TRY(d_o.reserve(L2 * 8 + 8));
    HIP_TRY(d_c.reserve(65537 * 4));
    HIP_TRY(hipMemsetAsync(d_c.p, 0, 65537 * 4, s));
    hipLaunchKernelGGL(k_count_by_gram, dim3(grid_for(L2)), dim3(kThreads), 0, s,
                       d_lgram[1].as<uint32_t>(), L2, d_c.as<uint32_t>());
    HIP_TRY(eng->d_xr2.reserve(65537 * 4));
    HIP_TRY(prim.inscan(d_c.as<uint32_t>(), eng->d_xr2.as<uint32_t>(), 65537));
    hipLaunchKernelGGL(k_sort_keys_of_lists, dim3(grid_for(L2)), dim3(kThreads), 0, s,
                       d_lgram[1].as<uint32_t>(), L2, d_k.as<uint64_t>());
    HIP_TRY(prim.sort(d_k.as<uint64_t>(), d_o.as<uint64_t>(), L2, 48));
    HIP_TRY(eng->d_xe2.reserve(L2 * 16 + 16));
    hipLaunchKernelGGL(k_xe2, dim3(grid_for(L2)), dim3(kThreads), 0, s, d_o.as<uint64_t>(), L2,
                       eng->d_beg[1].as<uint32_t>(), eng->d_nrep[1].as<uint32_t>(),
                       d_lgrp[1].as<uint32_t>(), eng->d_xe2.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    sz.xr2 = 65537 * 4;
    sz.xe2 = L2 * 16;
  }
  if (depth >= 3) {
    const uint64_t L2 = nlists[1], L3 = nlists[2];
    HIP_TRY(eng->d_x2rep.reserve(2048 * 4));
    HIP_TRY(hipMemsetAsync(eng->d_x2rep.p, 0, 2048 * 4, s));
    hipLaunchKernelGGL(k_x2rep, dim3(grid_for(L2)), dim3(kThreads), 0, s,
                       d_lgram[1].as<uint32_t>(), eng->d_nrep[1].as<uint32_t>(), L2,
                       eng->d_x2rep.as<uint32_t>());
    DevBuf d_k, d_o, d_start, d_contrib, d_xpos, d_big;
    HIP_TRY(d_k.reserve(L3 * 8 + 8));
    HIP_TRY(d_o.reserve(L3 * 8 + 8));
    HIP_TRY(d_start.reserve(L3 * 4 + 4));
    HIP_TRY(d_contrib.reserve(L3 * 8 + 8));
    HIP_TRY(d_xpos.reserve((L3 + 1) * 8));
    HIP_TRY(d_big.reserve(4));
    HIP_TRY(hipMemsetAsync(d_big.p, 0, 4, s));
    hipLaunchKernelGGL(k_sort_keys_of_lists, dim3(grid_for(L3)), dim3(kThreads), 0, s,
                       d_lgram[2].as<uint32_t>(), L3, d_k.as<uint64_t>());
    HIP_TRY(prim.sort(d_k.as<uint64_t>(), d_o.as<uint64_t>(), L3, 56));
    hipLaunchKernelGGL(k_x3_runs, dim3(grid_for(L3)), dim3(kThreads), 0, s, d_o.as<uint64_t>(),
                       L3, eng->d_beg[2].as<uint32_t>(), d_start.as<uint32_t>(),
                       d_contrib.as<uint64_t>());
    HIP_TRY(prim.exscan(d_contrib.as<uint64_t>(), d_xpos.as<uint64_t>(), L3));
    const uint64_t P3 = read1(s, d_xpos.as<uint64_t>() + L3);
    if (P3 >= (1ull << 31)) {
      err = "3-gram index too large (2^31 pairs); use a smaller precompute depth";
      return NGAP_ERR_INVALID;
    }
    HIP_TRY(eng->d_xs3.reserve((2ull << 24) * 4));
    HIP_TRY(hipMemsetAsync(eng->d_xs3.p, 0, (2ull << 24) * 4, s));
    HIP_TRY(eng->d_xp3.reserve(P3 * 8 + 8));
    hipLaunchKernelGGL(k_x3_fill, dim3(grid_for(L3)), dim3(kThreads), 0, s, d_o.as<uint64_t>(),
                       L3, eng->d_beg[2].as<uint32_t>(), eng->d_list[2].as<int32_t>(),
                       d_lgrp[2].as<uint32_t>(), d_xpos.as<uint64_t>(),
                       eng->d_xs3.as<uint32_t>(), eng->d_xp3.as<uint32_t>(),
                       d_big.as<uint32_t>());
    hipLaunchKernelGGL(k_x3_flags, dim3(grid_for(1ull << 24)), dim3(kThreads), 0, s,
                       eng->d_xmask1.as<uint32_t>(), eng->d_x2rep.as<uint32_t>(),
                       eng->d_xs3.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    if (read1(s, d_big.as<uint32_t>())) {
      err = "a 3-gram has 2^30 depth-3 states or more; use a smaller precompute depth";
      return NGAP_ERR_INVALID;
    }
    sz.x2rep = 2048 * 4;
    sz.xs3 = (2ull << 24) * 4;
    sz.xp3 = P3 * 8;
  }
  HIP_TRY(hipStreamSynchronize(s));
  return NGAP_OK;
}

}  // namespace ngapamd
