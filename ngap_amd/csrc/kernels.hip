// HIP kernels for gfx950 (wave64): the non-blocking (state, offset) worklist
// engine.
//
// Reference behaviour (what is computed): the worklist kernels
// advanceAndFilterNonBlocking{NAP,,R1,R2,Prec,All}Groups
// (code/src/ngap/kernel_bap.cu:10-185, kernel_ngap_O1.cu:10-185,
// kernel_ngap_O4.cu:9-548, kernel_ngap_O3.cu:10-295, kernel_ngap_OA.cu:10-374):
// an item (v, p) means state v matched input[p]; expanding it tests every
// successor u of v against input[p+1], reports u when it reports, and enables
// (u, p+1).  All-input states are enabled at every position; start-of-data
// states at the first position of chunk 0 (or of every chunk when the input
// was duplicated).  The report set is the reference's; item order is free.
//
// How it is computed here (MI355X-first, not the reference's structure):
//  * Each wave owns a ring of items in LDS.  Pushes are compacted with
//    ballot + mbcnt (one ring slot computation per wave, no atomics at all:
//    head/tail are wave-uniform registers).  Items beyond the ring go to a
//    per-wave HBM spill area and come back when the ring runs low.
//  * Idle lanes refill from the ring every step, so a wave stays converged
//    and busy without block barriers.
//  * A lane keeps the first surviving successor of its item in registers and
//    continues with it (work privatization generalised to whole chains; the
//    reference expands one more level, kernel_ngap_OA.cu:96-138).
//  * Roots come from the memo tables: one lookup per position gives the
//    states matched d-1 positions later (prefix memoization,
//    kernel_ngap_OA.cu:243-310), flattened across the wave with a prefix sum.
//  * Reports are compacted into a per-wave LDS buffer and flushed with one
//    global atomic per flush (the reference: one atomic per report,
//    kernel_helper.h:78-104).
//  * (state, offset) items of states with two or more incoming edges pass a
//    lossy LDS cache (64-bit atomic exchange): an exact duplicate seen
//    recently is dropped.  A duplicate is never needed (set semantics), and a
//    first occurrence is never dropped.
//  * The batched engine is a persistent grid whose waves pull work items
//    (a window of one chunk, for one group or for all groups) from eight
//    XCD-local queues; an all-group work item takes the roots of every group
//    from the cross-group 2-/3-gram index, one lookup per position.
//  * The loads of a step (input bytes, index offsets, a record and its class
//    words) are issued unconditionally at valid stand-in addresses and
//    selected afterwards, so they go out together: the engine is bound by the
//    latency of dependent gathers.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ngapamd {

#define NGAP_DEV __device__ __forceinline__

NGAP_DEV uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
NGAP_DEV uint32_t mask_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
NGAP_DEV uint64_t wballot(bool p) { return __ballot(p); }
NGAP_DEV void wsync() { __builtin_amdgcn_wave_barrier(); }

NGAP_DEV uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  return x;
}

// A wave's expanded items, added at its workgroup's lane of the count.
NGAP_DEV void add_items(const KParams &P, unsigned long long n) {
  atomicAdd(P.items + (blockIdx.x % kItemLanes) * kItemStride, n);
}

struct WaveState {
  uint64_t *ring;
  uint64_t *rbuf;
  uint64_t *dc;
  uint32_t *sh_pref;  // 65
  uint32_t *sh_beg;   // 64
  uint32_t *sh_pos;   // 64
  uint32_t *sh_tag;   // 64
  uint32_t *mask1;    // 16: root bytes, then reporting depth-1 bytes (layout.h)
  uint32_t *open;     // kMaxOpen: work items whose items may still be queued
  uint32_t n_open;
  uint64_t *spill;
  uint32_t head, tail, rn, spill_n;
  uint32_t lane;
  bool overflow;
  unsigned long long items;
  // root pass: the work item's area of root_items and its item count
  uint64_t *out;
  unsigned long long *out_count;
  uint64_t out_cap;
  // current (group, chunk)
  uint32_t g, gtag, rec_base, cls_base, cstart, cend;
};

NGAP_DEV bool class_has(const KParams &P, uint32_t cls_base, uint32_t u, uint32_t b) {
  return (P.img.cls[(size_t)(cls_base + u) * 8 + (b >> 5)] >> (b & 31)) & 1u;
}

// An item (and a report) is (offset << 32) | (group << 22) | state.
NGAP_DEV uint64_t item_of(uint32_t tagged, uint32_t pos) {
  return ((uint64_t)pos << 32) | tagged;
}

// A 32-byte successor record: four tagged successor ids and their tests.
struct Rec {
  int4 ids;
  uint4 tests;
};

NGAP_DEV Rec load_rec(const KParams &P, uint32_t r) {
  const int4 *R = reinterpret_cast<const int4 *>(P.img.succ4) + (size_t)r * 2;
  Rec x;
  x.ids = R[0];
  const int4 t = R[1];
  x.tests = make_uint4((uint32_t)t.x, (uint32_t)t.y, (uint32_t)t.z, (uint32_t)t.w);
  return x;
}

NGAP_DEV Rec empty_rec() {
  Rec x;
  x.ids = make_int4(-1, -1, -1, -1);
  x.tests = make_uint4(0, 0, 0, 0);
  return x;
}

// Class word a successor test needs (0 when the test word decides alone).
NGAP_DEV uint32_t test_word(const KParams &P, uint32_t cls_base, bool live, int32_t e,
                            uint32_t t, uint32_t b) {
  if (!live || e < 0 || (t & (kTestLiteral | kTestAll))) return 0u;
  return P.img.cls[(size_t)(cls_base + ((uint32_t)e & (uint32_t)kIdMask)) * 8 + (b >> 5)];
}

NGAP_DEV bool test_hit(bool live, int32_t e, uint32_t t, uint32_t word, uint32_t b) {
  if (!live || e < 0) return false;
  if (t & kTestLiteral) return (t & 255u) == b;
  if (t & kTestAll) return true;
  return (word >> (b & 31)) & 1u;
}

// Root pass: the wave's ring is a staging buffer (filled from 0), appended
// to the work item's area of root_items when it fills and at the end of a
// piece.  Items beyond the area's capacity are dropped and the work item is
// marked for the fallback.
NGAP_DEV void flush_stage(const KParams &P, WaveState &W) {
  if (W.tail == 0) return;
  wsync();
  unsigned long long base = 0;
  if (W.lane == 0) base = atomicAdd(W.out_count, (unsigned long long)W.tail);
  base = __shfl(base, 0, 64);
  for (uint32_t i = W.lane; i < W.tail; i += kWave)
    if (base + i < W.out_cap) W.out[base + i] = W.ring[i];
  if (base + W.tail > W.out_cap) W.overflow = true;
  W.tail = 0;
  wsync();
}

template <bool Stage = false>
NGAP_DEV void push(const KParams &P, WaveState &W, bool pred, uint64_t item) {
  uint64_t m = wballot(pred);
  if (m == 0) return;
  uint32_t cnt = __popcll(m);
  uint32_t r = mask_rank(m);
  if (Stage) {
    if (W.tail + cnt > P.ring_cap) flush_stage(P, W);
    if (pred) W.ring[W.tail + r] = item;
    W.tail += cnt;
    wsync();
    return;
  }
  if (W.tail - W.head + cnt <= P.ring_cap) {
    if (pred) W.ring[(W.tail + r) & (P.ring_cap - 1)] = item;
    W.tail += cnt;
  } else if (W.spill_n + cnt <= P.spill_cap) {
    if (pred) W.spill[W.spill_n + r] = item;
    W.spill_n += cnt;
  } else {
    W.overflow = true;
  }
  wsync();
}

NGAP_DEV void flush_reports(const KParams &P, WaveState &W) {
  if (W.rn == 0) return;
  wsync();
  unsigned long long base = 0;
  if (W.lane == 0) base = atomicAdd(P.report_count, (unsigned long long)W.rn);
  base = __shfl(base, 0, 64);
  if (!P.report_off) {
    for (uint32_t i = W.lane; i < W.rn; i += kWave) {
      unsigned long long idx = base + i;
      if (idx < P.report_cap) P.reports[idx] = W.rbuf[i];
    }
  }
  W.rn = 0;
  wsync();
}

// Report encoding on the device: (offset << 32) | (state | group << 22), so a
// plain numeric sort gives the reference's (offset, id) order.
NGAP_DEV void emit(const KParams &P, WaveState &W, bool pred, uint64_t val) {
  uint64_t m = wballot(pred);
  if (m == 0) return;
  uint32_t cnt = __popcll(m);
  if (W.rn + cnt > P.rbuf_cap) flush_reports(P, W);
  if (pred) W.rbuf[W.rn + mask_rank(m)] = val;
  W.rn += cnt;
  wsync();
}

NGAP_DEV bool dedup_fresh(const KParams &P, WaveState &W, uint64_t key) {
  uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & (P.dc_cap - 1);
  unsigned long long old = atomicExch((unsigned long long *)&W.dc[h], (unsigned long long)key);
  return old != key;
}

// Spilled items return to the ring when it runs low.
NGAP_DEV void refill_from_spill(const KParams &P, WaveState &W) {
  uint32_t room = P.ring_cap - (W.tail - W.head);
  uint32_t k = W.spill_n < room ? W.spill_n : room;
  uint32_t from = W.spill_n - k;
  for (uint32_t i = W.lane; i < k; i += kWave)
    W.ring[(W.tail + i) & (P.ring_cap - 1)] = W.spill[from + i];
  W.tail += k;
  W.spill_n = from;
  wsync();
}

// Processes a variable number of memo-list entries per lane as one flat index
// space (prefix sum over the wave), 64 entries per step.  Every entry that
// reports is emitted; entries are also pushed as items unless the lane's
// list is report-only (bit 31 of beg), the state has no successors, or the
// position is the chunk's last.
template <bool Stage = false>
NGAP_DEV void flatten(const KParams &P, WaveState &W, uint32_t cnt, uint32_t beg,
                      uint32_t pos, uint32_t gtag, const int32_t *list) {
  if (wballot(cnt > 1u) == 0) {
    // at most one entry per lane (the common case): no prefix sum
    const bool ok = cnt == 1u;
    if (wballot(ok) == 0) return;
    const int32_t e = ok ? list[beg & 0x7fffffffu] : 0;
    const uint64_t it = item_of(((uint32_t)e & (uint32_t)kIdMask) | gtag, pos);
    emit(P, W, ok && (e & kFlagReport), it);
    bool keep = ok && !(beg >> 31) && !(e & kFlagDeadEnd) && (pos + 1 < W.cend);
    if (keep && P.dedup && (e & kFlagDedup)) keep = dedup_fresh(P, W, it);
    push<Stage>(P, W, keep, it);
    return;
  }
  uint32_t incl = wave_incl_scan(cnt, W.lane);
  uint32_t total = __shfl(incl, 63, 64);
  if (total == 0) return;
  W.sh_pref[W.lane] = incl - cnt;
  W.sh_beg[W.lane] = beg;
  W.sh_pos[W.lane] = pos;
  W.sh_tag[W.lane] = gtag;
  if (W.lane == 63) W.sh_pref[64] = total;
  wsync();
  for (uint32_t base = 0; base < total; base += kWave) {
    uint32_t j = base + W.lane;
    bool ok = j < total;
    int32_t e = 0;
    uint32_t ppos = 0, tg = 0;
    bool ronly = true;
    if (ok) {
      int lo = 0, hi = 64;  // largest o with pref[o] <= j (pref[64] = total > j)
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (W.sh_pref[mid] <= j) lo = mid;
        else hi = mid - 1;
      }
      const uint32_t bg = W.sh_beg[lo];
      ronly = bg >> 31;
      e = list[(bg & 0x7fffffffu) + (j - W.sh_pref[lo])];
      ppos = W.sh_pos[lo];
      tg = W.sh_tag[lo];
    }
    const uint32_t u = ((uint32_t)e & (uint32_t)kIdMask) | tg;
    emit(P, W, ok && (e & kFlagReport), item_of(u, ppos));
    bool keep = ok && !ronly && !(e & kFlagDeadEnd) && (ppos + 1 < W.cend);
    if (keep && P.dedup && (e & kFlagDedup)) keep = dedup_fresh(P, W, item_of(u, ppos));
    push<Stage>(P, W, keep, item_of(u, ppos));
  }
  wsync();
}

// Rank-table lookup: pair `at` covers 32 consecutive keys; returns the
// key's list index, or -1 when bit `bit` of the pair's word is clear.
NGAP_DEV int32_t memo_rank(const uint32_t *sel, size_t at, uint32_t bit) {
  const uint2 w = reinterpret_cast<const uint2 *>(sel)[at];
  if (!((w.x >> bit) & 1u)) return -1;
  return (int32_t)(w.y + __popc(w.x & ((1u << bit) - 1u)));
}

// Roots for positions [p0, pend): one lane per position.  A position whose
// byte starts no root costs one load and one bit test.  A lane uses the
// deepest memo level that fits before the chunk end (kernel_ngap_OA.cu:243-
// 244); shallower levels contribute their reports only.  With depth >= 2 the
// independent loads of all levels are issued together (input bytes, then the
// rank pairs, then the list ranges) and the depth-1 lists are read only for
// bytes whose list reports.
// Loads in the root and expansion steps are issued unconditionally, at a
// valid address when a lane does not need the value (index 0, or the chunk's
// last byte), and the value is selected afterwards: the loads of a step then
// go out together instead of one branch (and one wait) at a time.
NGAP_DEV uint32_t min_u(uint32_t a, uint32_t b) { return a < b ? a : b; }

template <bool Stage = false>
NGAP_DEV void roots(const KParams &P, WaveState &W, uint32_t p0, uint32_t pend) {
  const uint8_t *in = P.input;
  const uint32_t p = p0 + W.lane;
  const bool act = p < pend;
  const uint32_t left = act ? W.cend - p : 0u;
  const uint32_t k = left < (uint32_t)P.memo_depth ? left : (uint32_t)P.memo_depth;
  const uint32_t pc = act ? p : p0, last = W.cend - 1;
  const uint32_t b0 = in[pc];
  const uint32_t b1 = in[min_u(pc + 1, last)];
  const uint32_t b2 = in[min_u(pc + 2, last)];
  const bool has = act && ((W.mask1[b0 >> 5] >> (b0 & 31)) & 1u);
  if (wballot(has) == 0) return;
  const uint32_t kReportOnly = 1u << 31;
  if (P.memo_depth < 2) {  // depth 1: all-input states matching b0 become items
    const uint32_t li = W.g * 256u + b0;
    const uint32_t beg = P.img.beg[0][li], end = P.img.beg[0][li + 1];
    flatten<Stage>(P, W, has ? end - beg : 0u, beg, p, W.gtag, P.img.list[0]);
    return;
  }
  // rank pair of the 2-gram (and the depth-1 reports of rare bytes)
  const uint32_t key = (b0 << 8) | b1;
  const uint2 w2 = reinterpret_cast<const uint2 *>(P.img.sel[1])[(size_t)W.g * 2048 + (key >> 5)];
  const uint32_t bit2 = key & 31;
  const int32_t li2 = has && k >= 2 && ((w2.x >> bit2) & 1u)
                          ? (int32_t)(w2.y + __popc(w2.x & ((1u << bit2) - 1u)))
                          : -1;
  const bool rep1 = has && ((W.mask1[8 + (b0 >> 5)] >> (b0 & 31)) & 1u);
  if (wballot(rep1)) {  // depth 1: reporting entries only (the last position included)
    uint32_t beg = 0, cnt = 0;
    if (rep1) {
      const uint32_t li = W.g * 256u + b0;
      beg = P.img.beg[0][li] | kReportOnly;
      cnt = P.img.nrep[0][li];
    }
    flatten<Stage>(P, W, cnt, beg, p, W.gtag, P.img.list[0]);
  }
  const bool has2 = li2 >= 0;
  if (wballot(has2) == 0) return;
  // depth-2 list range and the depth-3 rank pair (some lane has a depth-2
  // list, so index 0 and 1 are valid stand-ins)
  const uint32_t l2 = has2 ? (uint32_t)li2 : 0u;
  const uint32_t beg2r = P.img.beg[1][l2], end2 = P.img.beg[1][l2 + 1];
  const uint32_t nrep2 = P.img.nrep[1][l2];
  int32_t li3 = -1;
  if (P.memo_depth >= 3) {
    const uint2 w3 = reinterpret_cast<const uint2 *>(P.img.sel[2])[(size_t)l2 * 8 + (b2 >> 5)];
    const uint32_t bit3 = b2 & 31;
    li3 = has2 && k >= 3 && ((w3.x >> bit3) & 1u)
              ? (int32_t)(w3.y + __popc(w3.x & ((1u << bit3) - 1u)))
              : -1;
  }
  uint32_t cnt2 = has2 ? (k == 2 ? end2 - beg2r : nrep2) : 0u;
  const uint32_t beg2 = k == 2 ? beg2r : (beg2r | kReportOnly);
  uint32_t beg3 = 0, cnt3 = 0;
  if (wballot(li3 >= 0)) {
    const uint32_t l3 = li3 >= 0 ? (uint32_t)li3 : 0u;
    const uint32_t b3 = P.img.beg[2][l3], e3 = P.img.beg[2][l3 + 1];
    beg3 = b3;
    cnt3 = li3 >= 0 ? e3 - b3 : 0u;
  }
  flatten<Stage>(P, W, cnt2, beg2, p + 1, W.gtag, P.img.list[1]);
  if (wballot(cnt3 != 0)) flatten<Stage>(P, W, cnt3, beg3, p + 2, W.gtag, P.img.list[2]);
}

// Distributes per-lane index ranges [e0, e0 + ecnt) over the wave, 64
// entries per round, and calls f(ok, entry index, owner lane) each round (a
// wave-uniform loop; the owner is found by binary search over the inclusive
// prefix sums held in the lanes).
template <class F>
NGAP_DEV void for_entries(const WaveState &W, uint32_t e0, uint32_t ecnt, F &&f) {
  if (wballot(ecnt > 1u) == 0) {  // at most one entry per lane: its own
    f(ecnt == 1u, e0, W.lane);
    return;
  }
  const uint32_t incl = wave_incl_scan(ecnt, W.lane);
  const uint32_t total = __shfl(incl, 63, 64);
  for (uint32_t base = 0; base < total; base += kWave) {
    const uint32_t j = base + W.lane;
    const bool ok = j < total;
    uint32_t lo = 0, hi = 63;  // smallest lane o with incl[o] > j
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const uint32_t mid = (lo + hi) >> 1;
      if (__shfl(incl, (int)mid, 64) > j) hi = mid;
      else lo = mid + 1;
    }
    const uint32_t oi = __shfl(incl, (int)lo, 64);
    const uint32_t oc = __shfl(ecnt, (int)lo, 64);
    const uint32_t oe = __shfl(e0, (int)lo, 64);
    f(ok, oe + j - (oi - oc), lo);
  }
}

// for_entries over two ranges per lane, [a0, a0 + ac) then [b0, b0 + bc)
// (the two halves of a 128-position batch): f(ok, entry index, owner lane,
// range 0 or 1).
template <class F>
NGAP_DEV void for_entries2(const WaveState &W, uint32_t a0, uint32_t ac, uint32_t b0,
                           uint32_t bc, F &&f) {
  const uint32_t n = ac + bc;
  const uint32_t incl = wave_incl_scan(n, W.lane);
  const uint32_t total = __shfl(incl, 63, 64);
  for (uint32_t base = 0; base < total; base += kWave) {
    const uint32_t j = base + W.lane;
    const bool ok = j < total;
    uint32_t lo = 0, hi = 63;  // smallest lane o with incl[o] > j
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const uint32_t mid = (lo + hi) >> 1;
      if (__shfl(incl, (int)mid, 64) > j) hi = mid;
      else lo = mid + 1;
    }
    const uint32_t off = j - (__shfl(incl, (int)lo, 64) - __shfl(n, (int)lo, 64));
    const uint32_t oac = __shfl(ac, (int)lo, 64);
    // both bases are read by every lane before the select: a shuffle inside
    // a divergent branch reads nothing from lanes outside that branch
    const uint32_t oa0 = __shfl(a0, (int)lo, 64), ob0 = __shfl(b0, (int)lo, 64);
    const bool second = off >= oac;
    f(ok, second ? ob0 + (off - oac) : oa0 + off, lo, second ? 1u : 0u);
  }
}

// Per-lane state of one position in roots_cross.
struct CrossLane {
  uint32_t p, k, b0, e0, ec, t0, tc, r0, rc;
};

// The loads of one position's lookups, in stages shared by both halves of a
// 128-position batch: input bytes, then the index offsets.
NGAP_DEV void cross_lookup(const KParams &P, const WaveState &W, uint32_t p0, uint32_t pend,
                           CrossLane (&L)[2]) {
  const uint8_t *in = P.input;
  const uint32_t last = W.cend - 1;
  uint32_t b1[2], b2[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t p = p0 + h * kWave + W.lane;
    const bool act = p < pend;
    const uint32_t left = act ? W.cend - p : 0u;
    const uint32_t pc = act ? p : p0;
    L[h].p = p;
    L[h].k = left < (uint32_t)P.memo_depth ? left : (uint32_t)P.memo_depth;
    L[h].b0 = in[pc];
    b1[h] = in[min_u(pc + 1, last)];
    b2[h] = in[min_u(pc + 2, last)];
  }
  // Depth 3: the 3-gram slots first.  A slot flags the 3-grams whose first
  // byte or first two bytes begin reporting lists of depth 1 or 2 in some
  // group; the byte and 2-gram indexes are read only when some lane has such
  // a 3-gram or fewer than three bytes left in its chunk (for every other
  // lane their ranges are empty).
  bool full = P.memo_depth < 3;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    L[h].t0 = L[h].tc = L[h].e0 = L[h].ec = L[h].r0 = L[h].rc = 0;
    if (P.memo_depth >= 3) {
      const uint32_t k = L[h].k;
      const uint32_t key3 = (L[h].b0 << 16) | (b1[h] << 8) | b2[h];
      const uint2 slot = reinterpret_cast<const uint2 *>(P.img.xs3)[key3];
      L[h].t0 = slot.x;
      L[h].tc = k >= 3 ? (slot.y & ~kSlotPrefixReports) : 0u;
      full = full || (k >= 1 && k < 3) || (k >= 3 && (slot.y & kSlotPrefixReports));
    }
  }
  if (wballot(full) == 0) return;
  uint32_t x2w[2], x2a[2], x2b[2], x1a[2], x1b[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t key2 = (L[h].b0 << 8) | b1[h];
    x2a[h] = P.img.xr2[key2];
    x2b[h] = P.img.xr2[key2 + 1];
    x1a[h] = P.img.xr1[L[h].b0];
    x1b[h] = P.img.xr1[L[h].b0 + 1];
    x2w[h] = P.memo_depth >= 3 ? P.img.x2rep[key2 >> 5] : 0u;
  }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t k = L[h].k, b0 = L[h].b0, key2 = (b0 << 8) | b1[h];
    // depth-2 lists: the deepest level at depth 2; at depth 3 only where two
    // bytes are left or the 2-gram has reporting depth-2 lists
    const bool need2 = k >= 2 && (P.memo_depth == 2 || k == 2 || ((x2w[h] >> (key2 & 31)) & 1u));
    L[h].e0 = x2a[h];
    L[h].ec = need2 ? x2b[h] - x2a[h] : 0u;
    L[h].r0 = x1a[h];
    L[h].rc = k >= 1 && ((W.mask1[8 + (b0 >> 5)] >> (b0 & 31)) & 1u) ? x1b[h] - x1a[h] : 0u;
  }
}

// One entry per lane at most: the entry and its single list item are loaded
// for both halves before either is emitted.
template <bool Stage = false>
NGAP_DEV void emit_one(const KParams &P, WaveState &W, bool ok, int32_t e, bool ronly,
                       uint32_t pos, uint32_t tag) {
  const uint64_t it = item_of(((uint32_t)e & (uint32_t)kIdMask) | tag, pos);
  emit(P, W, ok && (e & kFlagReport), it);
  bool keep = ok && !ronly && !(e & kFlagDeadEnd) && (pos + 1 < W.cend);
  if (keep && P.dedup && (e & kFlagDedup)) keep = dedup_fresh(P, W, it);
  push<Stage>(P, W, keep, it);
}

// Roots of every group for positions [p0, pend), pend - p0 <= 128 (cross-
// group work items, memo depth >= 2): one lane per position and half; the
// position's byte selects the groups whose all-input states report on it,
// its 2-gram the depth-2 lists and its 3-gram the depth-3 lists of every
// group.  The same lists as roots() visits group by group, so the same items.
template <bool Stage = false>
NGAP_DEV void roots_cross(const KParams &P, WaveState &W, uint32_t p0, uint32_t pend) {
  const uint32_t kReportOnly = 1u << 31;
  CrossLane L[2];
  cross_lookup(P, W, p0, pend, L);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t p = L[h].p;
    // depth 1: reporting entries only (positions that continue use depth >= 2)
    if (wballot(L[h].rc != 0)) {
      for_entries(W, L[h].r0, L[h].rc, [&](bool ok, uint32_t idx, uint32_t o) {
        const uint32_t pos = __shfl(p, (int)o, 64);
        uint32_t beg = 0, cnt = 0, tag = 0;
        if (ok) {
          const uint32_t li = P.img.xe1[idx];
          beg = P.img.beg[0][li] | kReportOnly;
          cnt = P.img.nrep[0][li];
          tag = (li >> 8) << 22;
        }
        flatten<Stage>(P, W, cnt, beg, pos, tag, P.img.list[0]);
      });
    }
    // depth 2: full lists where it is the deepest level, reports otherwise
    if (wballot(L[h].ec != 0)) {
      const uint32_t k = L[h].k;
      for_entries(W, L[h].e0, L[h].ec, [&](bool ok, uint32_t idx, uint32_t o) {
        const uint32_t pos = __shfl(p, (int)o, 64);
        const uint32_t ko = __shfl(k, (int)o, 64);
        uint32_t beg = 0, cnt = 0, tag = 0;
        if (ok) {
          // (list index, begin, end, reporting entries | group << 22)
          const uint4 e = reinterpret_cast<const uint4 *>(P.img.xe2)[idx];
          tag = (e.w >> 22) << 22;
          beg = e.y;
          cnt = ko == 2 || P.memo_depth == 2 ? e.z - e.y : (e.w & (uint32_t)kIdMask);
          if (!(ko == 2 || P.memo_depth == 2)) beg |= kReportOnly;
        }
        flatten<Stage>(P, W, cnt, beg, pos + 1, tag, P.img.list[1]);
      });
    }
  }
  if (P.memo_depth < 3) return;
  // depth 3 from the 3-gram slots: an inline state is emitted at once; the
  // other slots' (tagged state, group) pairs of both halves are spread over
  // the wave together, one 8-byte load per pair
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const bool inl = (L[h].tc & kSlotInline) != 0;
    emit_one<Stage>(P, W, inl, (int32_t)L[h].t0, false, L[h].p + 2, (L[h].tc & 1023u) << 22);
    if (inl) L[h].tc = 0;
  }
  if (wballot(L[0].tc != 0 || L[1].tc != 0) == 0) return;
  // some lane has pairs, so pair 0 is a valid stand-in address
  for_entries2(W, L[0].t0, L[0].tc, L[1].t0, L[1].tc,
               [&](bool ok, uint32_t idx, uint32_t o, uint32_t h) {
                 const uint2 e = reinterpret_cast<const uint2 *>(P.img.xp3)[ok ? idx : 0u];
                 emit_one<Stage>(P, W, ok, (int32_t)e.x, false, p0 + h * kWave + o + 2,
                                 e.y << 22);
               });
}

template <bool Stage = false>
NGAP_DEV void seed_start_of_data(const KParams &P, WaveState &W, uint32_t sod_begin,
                                 uint32_t sod_end) {
  uint32_t b = P.input[W.cstart];
  for (uint32_t i = sod_begin; i < sod_end; i += kWave) {
    uint32_t idx = i + W.lane;
    bool ok = idx < sod_end;
    int32_t e = ok ? P.img.sod[idx] : 0;
    uint32_t u = (uint32_t)e & (uint32_t)kIdMask;
    bool hit = ok && class_has(P, W.cls_base, u, b);
    const uint64_t it = item_of(u | W.gtag, W.cstart);
    emit(P, W, hit && (e & kFlagReport), it);
    bool keep = hit && !(e & kFlagDeadEnd) && (W.cstart + 1 < W.cend);
    if (keep && P.dedup && (e & kFlagDedup)) keep = dedup_fresh(P, W, it);
    push<Stage>(P, W, keep, it);
  }
}

// The same for every group (cross-group work items): (tagged id, group)
// pairs, each group's class base from its metadata.
template <bool Stage = false>
NGAP_DEV void seed_start_all(const KParams &P, WaveState &W) {
  const uint32_t b = P.input[W.cstart];
  for (uint32_t i = 0; i < P.img.n_xsod; i += kWave) {
    const uint32_t idx = i + W.lane;
    const bool ok = idx < P.img.n_xsod;
    int32_t e = 0;
    uint32_t g = 0;
    bool hit = false;
    if (ok) {
      e = (int32_t)P.img.xsod[2 * idx];
      g = P.img.xsod[2 * idx + 1];
      hit = class_has(P, P.img.meta[g].cls_base, (uint32_t)e & (uint32_t)kIdMask, b);
    }
    const uint64_t it = item_of(((uint32_t)e & (uint32_t)kIdMask) | (g << 22), W.cstart);
    emit(P, W, hit && (e & kFlagReport), it);
    bool keep = hit && !(e & kFlagDeadEnd) && (W.cstart + 1 < W.cend);
    if (keep && P.dedup && (e & kFlagDedup)) keep = dedup_fresh(P, W, it);
    push<Stage>(P, W, keep, it);
  }
}

// One step of every valid lane's item: test its successors on the next byte.
// Lane context of the item being expanded: group tag and the group's record
// and class bases.
struct LaneGroup {
  uint32_t gtag, rec_base, cls_base;
};

// Where expand() finds a position's chunk end: engine runs (chunks of
// chunk_len bytes, a divisibility test valid for items of any work item),
// request runs of the streaming kernel (the last-byte map), or decided at run
// time (chunk_len, else the wave's one open work item's bounds).
enum : int { kEndChunkLen = 0, kEndMap = 1, kEndRuntime = 2 };

template <int End = kEndRuntime>
NGAP_DEV void expand(const KParams &P, WaveState &W, bool &valid, uint32_t &v,
                     uint32_t &p, int &depth, const LaneGroup &lg) {
  const uint32_t p1 = p + 1;
  // unconditional loads (index 0 stands in for idle lanes; some lane is valid,
  // so the image has a record)
  const uint32_t b = P.input[valid ? p1 : 0u];
  Rec rec = load_rec(P, valid ? lg.rec_base + v : 0u);
  if (!valid) rec = empty_rec();
  const uint32_t busy = __popcll(wballot(valid));
  W.items += busy;
  // is p1 the last position of its chunk (no successors enabled after it)?
  bool last;
  if (End == kEndMap) {
    const uint32_t lw = P.last_map[valid ? p1 >> 5 : 0u];
    last = ((lw >> (p1 & 31)) & 1u) != 0u;
  } else if (End == kEndChunkLen || P.chunk_len) {
    last = (uint64_t)(p1 + 1) * P.chunk_magic <= P.chunk_magic - 1;
  } else {
    last = p1 + 1 >= W.cend;
  }
  bool has_next = false;
  uint32_t nv = 0;
  // work privatization (kernel_ngap_O4.cu:115-122): a matched successor is
  // expanded by the same lane while more than active_threshold lanes are
  // busy, and otherwise queued for the idle lanes
  const bool can_priv = (P.priv_limit < 0 || depth < P.priv_limit) &&
                        (int32_t)busy > P.active_threshold;
  bool more = valid;
  while (wballot(more)) {
    int32_t link = -1;
    const int32_t es[4] = {rec.ids.x, rec.ids.y, rec.ids.z, rec.ids.w};
    const uint32_t ts[4] = {rec.tests.x, rec.tests.y, rec.tests.z, rec.tests.w};
    // all loads of this record first: the class words its tests need and the
    // continuation record, then the wave-wide emits and pushes
    uint32_t cw[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
      if (more && es[s] <= -2) link = -(es[s] + 2);
      const bool need = more && es[s] >= 0 && !(ts[s] & (kTestLiteral | kTestAll));
      const size_t at =
          need ? (size_t)(lg.cls_base + ((uint32_t)es[s] & (uint32_t)kIdMask)) * 8 + (b >> 5) : 0;
      cw[s] = P.img.cls[at];
    }
    Rec next = load_rec(P, more && link >= 0 ? (uint32_t)link : 0u);
    if (!(more && link >= 0)) next = empty_rec();
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const int32_t e = es[s];
      // successors fill the slots from 0 (a continuation link needs slots
      // 0-2 full): once no lane has one in this slot, none has one later
      if (s < 3 && wballot(more && e >= 0) == 0) break;
      const uint32_t u = (uint32_t)e & (uint32_t)kIdMask;
      const bool hit = test_hit(more, e, ts[s], cw[s], b);
      const uint64_t it = item_of(u | lg.gtag, p1);
      emit(P, W, hit && (e & kFlagReport), it);
      bool keep = hit && !(e & kFlagDeadEnd) && !last;
      if (keep && P.dedup && (e & kFlagDedup)) keep = dedup_fresh(P, W, it);
      const bool take = keep && !has_next && can_priv;
      if (take) {
        has_next = true;
        nv = u;
      }
      push(P, W, keep && !take, it);
    }
    more = more && link >= 0;
    rec = next;
  }
  if (has_next) {
    v = nv;
    p = p1;
    depth++;
  } else {
    valid = false;
  }
}

NGAP_DEV void clear_dedup(const KParams &P, WaveState &W) {
  for (uint32_t i = W.lane; i < P.dc_cap; i += kWave) W.dc[i] = ~0ull;
  wsync();
}

// The current (group, chunk): group bases, tag and root masks in LDS.
NGAP_DEV void set_unit(const KParams &P, WaveState &W, uint32_t g, uint32_t cstart,
                       uint32_t cend) {
  const GroupMeta gm = P.img.meta[g];
  W.g = g;
  W.gtag = g << 22;
  W.rec_base = gm.rec_base;
  W.cls_base = gm.cls_base;
  W.cstart = cstart;
  W.cend = cend;
  if (W.lane < 16) W.mask1[W.lane] = P.img.mask1[g * 16 + W.lane];
  wsync();
}

// A cross-group work item: every group's roots in one window of a chunk.
NGAP_DEV void set_cross(const KParams &P, WaveState &W, uint32_t cstart, uint32_t cend) {
  W.g = kAllGroups;
  W.gtag = 0;
  W.cstart = cstart;
  W.cend = cend;
  if (W.lane < 8) W.mask1[8 + W.lane] = P.img.xmask1[W.lane];  // bytes with depth-1 reports
  wsync();
}

NGAP_DEV void begin_unit(const KParams &P, WaveState &W, uint32_t g, uint32_t cstart,
                         uint32_t cend) {
  clear_dedup(P, W);
  set_unit(P, W, g, cstart, cend);
}

NGAP_DEV void begin_cross(const KParams &P, WaveState &W, uint32_t cstart, uint32_t cend) {
  clear_dedup(P, W);
  set_cross(P, W, cstart, cend);
}

NGAP_DEV void setup_wave(const KParams &P, WaveState &W, unsigned char *base,
                         uint64_t *spill) {
  W.ring = reinterpret_cast<uint64_t *>(base);
  W.rbuf = W.ring + P.ring_cap;
  W.dc = W.rbuf + P.rbuf_cap;
  W.sh_pref = reinterpret_cast<uint32_t *>(W.dc + P.dc_cap);
  W.sh_beg = W.sh_pref + 65;
  W.sh_pos = W.sh_beg + 64;
  W.sh_tag = W.sh_pos + 64;
  W.mask1 = W.sh_tag + 64;
  W.open = W.mask1 + 16;
  W.n_open = 0;
  W.spill = spill;
  W.out = nullptr;
  W.out_count = nullptr;
  W.out_cap = 0;
  W.head = W.tail = W.rn = W.spill_n = 0;
  W.lane = lane_id();
  W.overflow = false;
  W.items = 0;
}

// Runs items to completion, pulling root batches from `src` whenever the
// ring runs low.  Every loop iteration is wave-uniform.
template <bool Cross, class Src>
NGAP_DEV void drain(const KParams &P, WaveState &W, Src &src) {
  bool valid = false;
  uint32_t v = 0, p = 0;
  int depth = 0;
  LaneGroup lg{W.gtag, W.rec_base, W.cls_base};
  bool roots_left = true;
  for (;;) {
    uint32_t qn = W.tail - W.head;
    if (qn < kWave && W.spill_n > 0) {
      refill_from_spill(P, W);
      qn = W.tail - W.head;
    }
    // the next root batch runs while half the ring is still queued, so the
    // expansion steps after it find items for every idle lane
    if (qn < P.ring_cap / 2 && W.spill_n == 0 && roots_left) {
      uint32_t a, b;
      if (!src.next(W, a, b)) roots_left = false;
      else if (Cross) roots_cross(P, W, a, b);
      else roots(P, W, a, b);
      continue;
    }
    const uint64_t idle = wballot(!valid);
    uint32_t avail = W.tail - W.head;
    uint32_t nidle = __popcll(idle);
    uint32_t take = nidle < avail ? nidle : avail;
    uint32_t r = mask_rank(idle);
    if (!valid && r < take) {
      uint64_t it = W.ring[(W.head + r) & (P.ring_cap - 1)];
      v = (uint32_t)it & (uint32_t)kIdMask;
      p = (uint32_t)(it >> 32);
      valid = true;
      depth = 0;
      if (Cross) {
        const uint32_t g = ((uint32_t)it >> 22) & 1023u;
        const uint2 base = *reinterpret_cast<const uint2 *>(&P.img.meta[g]);
        lg.gtag = g << 22;
        lg.rec_base = base.x;
        lg.cls_base = base.y;
      }
    }
    W.head += take;
    wsync();
    if (wballot(valid) == 0) {
      if (W.tail == W.head && W.spill_n == 0 && !roots_left) break;
      continue;
    }
    expand(P, W, valid, v, p, depth, lg);
  }
}

// Claims the next work item for this wave: from the queue of its XCD first,
// then from the others.  Returns P.num_tasks when all are empty.  `q` is the
// wave's current queue (advanced when one runs dry).  A wave that finds a
// queue empty adds it to its XCD's mask of empty queues and gets back what
// the XCD's other waves found, so it claims only from queues not known to be
// empty (a queue found empty stays empty): at the end of a run a wave learns
// that every queue is empty from two atomics, not from a failed claim on each
// of the eight counters, which thousands of waves would hit at once.
NGAP_DEV uint32_t claim_task(const KParams &P, const WaveState &W, uint32_t &q) {
  uint32_t *empty = P.task_head + kEmptyMasks + (blockIdx.x % kQueues) * kQueueStride;
  uint32_t known = 0;
  for (uint32_t tries = 0; tries < kQueues; tries++) {
    uint32_t j = 0;
    if (W.lane == 0) j = atomicAdd(P.task_head + q * kQueueStride, 1u);
    j = __shfl(j, 0, 64);
    const uint64_t t = (uint64_t)j * kQueues + q;
    if (t < P.num_tasks) return (uint32_t)t;
    uint32_t m = 0;
    if (W.lane == 0) m = atomicOr(empty, 1u << q);
    known |= __shfl(m, 0, 64) | (1u << q);
    if (known == (1u << kQueues) - 1) return P.num_tasks;
    uint32_t d = 1;
    while ((known >> ((q + d) % kQueues)) & 1u) d++;
    q = (q + d) % kQueues;
  }
  return P.num_tasks;
}

// Batched schedule, engine runs: the wave's stream of work items.  A wave
// starts the next work item as soon as the current one has no positions left
// and its queue runs low, so the tail of one work item's items overlaps the
// roots of the next; at most kMaxOpen work items are open (items possibly
// queued) at a time.  next(): 0 = a root batch [a, b), 1 = wait for the open
// work items to finish, 2 = the queue is empty.
struct TaskStream {
  uint32_t cur = 0, end = 0;
  uint64_t rbase = 0;       // root pass: the work item's first root item
  bool cross = false;
  bool cross_mask = false;  // LDS holds the cross-group report mask
  uint32_t q = 0;           // the wave's queue
  NGAP_DEV void start(const KParams &P, const WaveState &W) { q = blockIdx.x % kQueues; }
  template <bool RootItems, bool Gen>
  NGAP_DEV int next(const KParams &P, WaveState &W, uint32_t &a, uint32_t &b) {
    if (cur >= end) {
      if (W.n_open >= kMaxOpen) return 1;
      const uint32_t t = claim_task(P, W, q);
      if (t >= P.num_tasks) return 2;
      const uint32_t ti = P.task_order ? P.task_order[t] : t;
      const Task T = Gen ? gen_task(P, ti) : P.tasks[ti];
      if (W.lane == 0) W.open[W.n_open] = ti;
      W.n_open++;
      if (RootItems) {  // the root pass wrote this work item's roots
        rbase = P.root_off[ti];
        const uint64_t cap = P.root_off[ti + 1] - rbase;
        const unsigned long long n = P.root_count[ti];
        cur = 0;
        end = (uint32_t)(n < cap ? n : cap);
        a = cur;
        b = end < kWave ? end : kWave;
        cur = b;
        return 0;
      }
      cross = T.group == kAllGroups;
      if (cross) {
        if (!cross_mask) set_cross(P, W, T.cstart, T.cend);
        W.g = kAllGroups;
        W.gtag = 0;
        W.cstart = T.cstart;
        W.cend = T.cend;
        cross_mask = true;
        if (T.flags & kTaskSeedStart) seed_start_all(P, W);
      } else {
        set_unit(P, W, T.group, T.cstart, T.cend);
        cross_mask = false;
        if (T.flags & kTaskSeedStart) {
          const GroupMeta gm = P.img.meta[T.group];
          seed_start_of_data(P, W, gm.sod_begin, gm.sod_end);
        }
      }
      cur = T.begin;
      end = T.end;
    }
    a = cur;
    // roots_cross takes two halves; the root pass's items go 64 at a time,
    // per-group roots root_batch positions at a time
    const uint32_t step = RootItems ? kWave : cross ? 2 * kWave : P.root_batch;
    b = cur + step < end ? cur + step : end;
    cur = b;
    return 0;
  }
};

// All queued items are done: the open work items are complete, and each is
// marked for the overflow fallback when the wave dropped an item since they
// opened.
NGAP_DEV void close_open(const KParams &P, WaveState &W) {
  if (W.n_open == 0) return;
  wsync();
  if (W.overflow) {
    if (W.lane < W.n_open) atomicOr(&P.overflow[W.open[W.lane]], kSpillOverflow);
    if (W.lane == 0) atomicAdd(P.overflow_count, W.n_open);
  }
  W.n_open = 0;
  W.overflow = false;
  wsync();
}

// Queues root items [from, from + n) of the root pass (n <= 64).
NGAP_DEV void take_roots(const KParams &P, WaveState &W, uint64_t from, uint32_t n) {
  if (n == 0) return;
  const bool ok = W.lane < n;
  const uint64_t it = P.root_items[ok ? from + W.lane : from];
  push(P, W, ok, it);
}

// drain() over a TaskStream: items of several work items (and groups) share
// the queue, so a lane looks up its item's group bases when it takes one.
template <bool RootItems, int End, bool Gen>
NGAP_DEV void drain_stream(const KParams &P, WaveState &W) {
  TaskStream src;
  src.start(P, W);
  bool valid = false;
  uint32_t v = 0, p = 0;
  int depth = 0;
  LaneGroup lg{0u, 0u, 0u};
  bool more = true;
  for (;;) {
    uint32_t qn = W.tail - W.head;
    if (qn < kWave && W.spill_n > 0) {
      refill_from_spill(P, W);
      qn = W.tail - W.head;
    }
    if (qn < P.ring_cap / 2 && W.spill_n == 0 && more) {  // as in drain()
      uint32_t a, b;
      const int st = src.template next<RootItems, Gen>(P, W, a, b);
      if (st == 0) {
        if (RootItems) take_roots(P, W, src.rbase + a, b - a);
        else if (src.cross) roots_cross(P, W, a, b);
        else roots(P, W, a, b);
        continue;
      }
      if (st == 2) more = false;
    }
    const uint64_t idle = wballot(!valid);
    const uint32_t avail = W.tail - W.head;
    const uint32_t nidle = __popcll(idle);
    const uint32_t take = nidle < avail ? nidle : avail;
    const uint32_t r = mask_rank(idle);
    if (!valid && r < take) {
      const uint64_t it = W.ring[(W.head + r) & (P.ring_cap - 1)];
      v = (uint32_t)it & (uint32_t)kIdMask;
      p = (uint32_t)(it >> 32);
      valid = true;
      depth = 0;
      const uint32_t g = ((uint32_t)it >> 22) & 1023u;
      const uint2 base = *reinterpret_cast<const uint2 *>(&P.img.meta[g]);
      lg.gtag = g << 22;
      lg.rec_base = base.x;
      lg.cls_base = base.y;
    }
    W.head += take;
    wsync();
    if (wballot(valid) == 0) {
      if (W.tail == W.head && W.spill_n == 0) {
        close_open(P, W);
        if (!more) break;
      }
      continue;
    }
    expand<End>(P, W, valid, v, p, depth, lg);
  }
}

// Plain schedule: the cursor over a (group, chunk) unit shared by the block's
// waves (the reference's one-block-per-unit decomposition, ngap.cu:1507).
struct SharedCursor {
  uint32_t *cursor;
  uint32_t end, batch;
  NGAP_DEV bool next(const WaveState &W, uint32_t &a, uint32_t &b) {
    uint32_t x = 0;
    if (W.lane == 0) x = atomicAdd(cursor, batch);
    x = __shfl(x, 0, 64);
    if (x >= end) return false;
    a = x;
    b = x + batch < end ? x + batch : end;
    return true;
  }
};

// Batched schedule: a wave walks its own work item in 64-position steps.
struct LocalCursor {
  uint32_t cur, end;
  NGAP_DEV bool next(const WaveState &W, uint32_t &a, uint32_t &b) {
    if (cur >= end) return false;
    a = cur;
    b = cur + kWave < end ? cur + kWave : end;
    cur = b;
    return true;
  }
};

__global__ void __launch_bounds__(kBlockThreads) k_plain(KParams P) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t unit = blockIdx.x;
  const uint32_t g = unit % P.num_groups, c = unit / P.num_groups;
  uint32_t *cursor = reinterpret_cast<uint32_t *>(lds);
  const uint32_t wbytes = wave_lds_bytes(P.ring_cap, P.rbuf_cap, P.dc_cap);
  WaveState W;
  setup_wave(P, W, lds + 16 + wave * wbytes,
             P.spill + ((size_t)unit * P.waves_per_block + wave) * P.spill_cap);
  begin_unit(P, W, g, c * P.chunk_len, (c + 1) * P.chunk_len);
  if (threadIdx.x == 0) *cursor = W.cstart;
  __syncthreads();
  if (wave == 0 && (c == 0 || P.is_dup)) {
    const GroupMeta gm = P.img.meta[g];
    seed_start_of_data(P, W, gm.sod_begin, gm.sod_end);
  }
  SharedCursor src{cursor, W.cend, P.root_batch};
  drain<false>(P, W, src);
  flush_reports(P, W);
  if (W.lane == 0) {
    if (W.overflow && atomicOr(&P.overflow[unit], 1u) == 0u) atomicAdd(P.overflow_count, 1u);
    add_items(P, W.items);
  }
}

__global__ void __launch_bounds__(kBlockThreads) k_batched(KParams P) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t wbytes = wave_lds_bytes(P.ring_cap, P.rbuf_cap, P.dc_cap);
  WaveState W;
  setup_wave(P, W, lds + 16 + wave * wbytes,
             P.spill + ((size_t)blockIdx.x * P.waves_per_block + wave) * P.spill_cap);
  uint32_t q = blockIdx.x % kQueues;
  for (;;) {
    const uint32_t t = claim_task(P, W, q);
    if (t >= P.num_tasks) break;
    const uint32_t ti = P.task_order ? P.task_order[t] : t;
    const Task T = P.tasks[ti];
    W.overflow = false;
    const unsigned long long items0 = W.items;
    LocalCursor src{T.begin, T.end};
    if (T.group == kAllGroups) {
      begin_cross(P, W, T.cstart, T.cend);
      if (T.flags & kTaskSeedStart) seed_start_all(P, W);
      drain<true>(P, W, src);
    } else {
      begin_unit(P, W, T.group, T.cstart, T.cend);
      if (T.flags & kTaskSeedStart) {
        const GroupMeta gm = P.img.meta[T.group];
        seed_start_of_data(P, W, gm.sod_begin, gm.sod_end);
      }
      drain<false>(P, W, src);
    }
    if (W.lane == 0) {
      if (W.overflow) {
        P.overflow[ti] = 1u;
        atomicAdd(P.overflow_count, 1u);
      }
      if (P.task_items) P.task_items[ti] = W.items - items0;
    }
  }
  flush_reports(P, W);
  if (W.lane == 0) add_items(P, W.items);
}

// The batched engine without per-work-item traces: each wave overlaps its
// work items (drain_stream).  Kernels of their own, so their registers are
// not sized for the per-item path: engine runs (End = kEndChunkLen, work
// items from P.tasks), request runs (kEndMap) and engine runs whose work
// items are generated from their index (Gen).
template <int End, bool Gen>
__global__ void __launch_bounds__(kBlockThreads) k_batched_stream(KParams P) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t wbytes = wave_lds_bytes(P.ring_cap, P.rbuf_cap, P.dc_cap);
  WaveState W;
  setup_wave(P, W, lds + 16 + wave * wbytes,
             P.spill + ((size_t)blockIdx.x * P.waves_per_block + wave) * P.spill_cap);
  clear_dedup(P, W);
  drain_stream<false, End, Gen>(P, W);
  flush_reports(P, W);
  if (W.lane == 0) add_items(P, W.items);
}

// The streaming kernel after a root pass (k_roots): its work items' roots
// are read from root_items, so it holds only the expansion's registers.
__global__ void __launch_bounds__(kBlockThreads) k_expand(KParams P) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t wbytes = wave_lds_bytes(P.ring_cap, P.rbuf_cap, P.dc_cap);
  WaveState W;
  setup_wave(P, W, lds + 16 + wave * wbytes,
             P.spill + ((size_t)blockIdx.x * P.waves_per_block + wave) * P.spill_cap);
  clear_dedup(P, W);
  drain_stream<true, kEndChunkLen, false>(P, W);
  flush_reports(P, W);
  if (W.lane == 0) add_items(P, W.items);
}

// Root pass of an engine run: one wave per piece of a work item (a non-
// persistent grid, so the gathers of many waves overlap).  The piece's roots
// (and start-of-data seeds) are staged in the wave's LDS ring and appended
// to the work item's area of root_items; their reports go out at once.
// k_batched_stream then expands the items.
__global__ void __launch_bounds__(kBlockThreads) k_roots(KParams P) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t piece = blockIdx.x * P.waves_per_block + wave;
  if (piece >= P.num_pieces) return;  // wave-uniform; no block barrier follows
  const uint32_t wbytes = wave_lds_bytes(P.ring_cap, P.rbuf_cap, P.dc_cap);
  WaveState W;
  setup_wave(P, W, lds + 16 + wave * wbytes, nullptr);
  const RootPiece R = P.pieces[piece];
  const Task T = P.tasks[R.task];
  const uint64_t o0 = P.root_off[R.task];
  W.out = P.root_items + o0;
  W.out_cap = P.root_off[R.task + 1] - o0;
  W.out_count = P.root_count + R.task;
  if (T.group == kAllGroups) {
    begin_cross(P, W, T.cstart, T.cend);
    if (R.flags & kTaskSeedStart) seed_start_all<true>(P, W);
    for (uint32_t a = R.begin; a < R.end; a += 2 * kWave)
      roots_cross<true>(P, W, a, min_u(a + 2 * kWave, R.end));
  } else {
    begin_unit(P, W, T.group, T.cstart, T.cend);
    if (R.flags & kTaskSeedStart) {
      const GroupMeta gm = P.img.meta[T.group];
      seed_start_of_data<true>(P, W, gm.sod_begin, gm.sod_end);
    }
    for (uint32_t a = R.begin; a < R.end; a += kWave)
      roots<true>(P, W, a, min_u(a + kWave, R.end));
  }
  flush_stage(P, W);
  flush_reports(P, W);
  if (W.lane == 0 && W.overflow && atomicOr(&P.overflow[R.task], kRootOverflow) == 0u)
    atomicAdd(P.overflow_count, 1u);
}

// blockinggroups (the reference's BAP baseline, kernel_ngap_O0.cu:7-126): one
// workgroup per (group, chunk) unit advances the set of states matched at
// position p to those matched at p + 1 in lock step, with a block barrier per
// input byte.  The frontier lives in a per-unit HBM buffer (two halves of
// P.spill_cap states); appends are compacted per wave.
NGAP_DEV void block_append(const KParams &P, WaveState &W, uint32_t *count, uint32_t *buf,
                           bool pred, uint32_t u) {
  const uint64_t m = wballot(pred);
  if (m == 0) return;
  uint32_t base = 0;
  if (W.lane == 0) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, 0, 64);
  const uint32_t slot = base + mask_rank(m);
  if (pred) {
    if (slot < P.spill_cap) buf[slot] = u;
    else W.overflow = true;
  }
}

__global__ void __launch_bounds__(kBlockThreads) k_blocking(KParams P) {
  extern __shared__ __align__(16) unsigned char lds[];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t unit = blockIdx.x;
  const uint32_t g = unit % P.num_groups, c = unit / P.num_groups;
  uint32_t *count = reinterpret_cast<uint32_t *>(lds);  // [0] current, [1] next
  const uint32_t wbytes = wave_lds_bytes(P.ring_cap, P.rbuf_cap, P.dc_cap);
  WaveState W;
  setup_wave(P, W, lds + 16 + wave * wbytes, nullptr);
  begin_unit(P, W, g, c * P.chunk_len, (c + 1) * P.chunk_len);
  uint32_t *front = reinterpret_cast<uint32_t *>(P.spill) + (size_t)unit * 2 * P.spill_cap;
  const GroupMeta gm = P.img.meta[g];
  const bool seed = c == 0 || P.is_dup;
  if (threadIdx.x == 0) count[0] = count[1] = 0;
  __syncthreads();
  // position cstart: matched start-of-data and all-input states
  {
    uint32_t *cur = front;
    const uint32_t b = P.input[W.cstart];
    for (uint32_t i = gm.sod_begin + wave * kWave; seed && i < gm.sod_end; i += kBlockThreads) {
      const uint32_t idx = i + W.lane;
      const bool ok = idx < gm.sod_end;
      const int32_t e = ok ? P.img.sod[idx] : 0;
      const uint32_t u = (uint32_t)e & (uint32_t)kIdMask;
      const bool hit = ok && class_has(P, W.cls_base, u, b);
      emit(P, W, hit && (e & kFlagReport), ((uint64_t)W.cstart << 32) | u | W.gtag);
      block_append(P, W, &count[0], cur, hit && !(e & kFlagDeadEnd), u);
    }
    const uint32_t li = g * 256u + b;
    const uint32_t lb = P.img.beg[0][li], le = P.img.beg[0][li + 1];
    for (uint32_t i = lb + wave * kWave; i < le; i += kBlockThreads) {
      const uint32_t idx = i + W.lane;
      const bool ok = idx < le;
      const int32_t e = ok ? P.img.list[0][idx] : 0;
      const uint32_t u = (uint32_t)e & (uint32_t)kIdMask;
      emit(P, W, ok && (e & kFlagReport), ((uint64_t)W.cstart << 32) | u | W.gtag);
      block_append(P, W, &count[0], cur, ok && !(e & kFlagDeadEnd), u);
    }
  }
  __syncthreads();
  for (uint32_t p = W.cstart; p + 1 < W.cend; p++) {
    const uint32_t sel = (p - W.cstart) & 1u;
    uint32_t *cur = front + sel * P.spill_cap;
    uint32_t *nxt = front + (sel ^ 1u) * P.spill_cap;
    uint32_t *ncount = &count[sel ^ 1u];
    const uint32_t n = min(count[sel], P.spill_cap);
    const uint32_t p1 = p + 1;
    const uint32_t b = P.input[p1];
    // successors of the matched states, tested on input[p + 1]
    for (uint32_t i = wave * kWave; i < n; i += kBlockThreads) {
      const uint32_t idx = i + W.lane;
      bool more = idx < n;
      Rec rec = empty_rec();
      if (more) rec = load_rec(P, W.rec_base + cur[idx]);
      while (wballot(more)) {
        int32_t link = -1;
        const int32_t es[4] = {rec.ids.x, rec.ids.y, rec.ids.z, rec.ids.w};
        const uint32_t ts[4] = {rec.tests.x, rec.tests.y, rec.tests.z, rec.tests.w};
#pragma unroll
        for (int s2 = 0; s2 < 4; s2++) {
          const int32_t e = es[s2];
          if (more && e <= -2) link = -(e + 2);
          const uint32_t u = (uint32_t)e & (uint32_t)kIdMask;
          const bool hit =
              test_hit(more, e, ts[s2], test_word(P, W.cls_base, more, e, ts[s2], b), b);
          emit(P, W, hit && (e & kFlagReport), item_of(u | W.gtag, p1));
          bool keep = hit && !(e & kFlagDeadEnd);
          if (keep && P.dedup && (e & kFlagDedup)) keep = dedup_fresh(P, W, item_of(u | W.gtag, p1));
          block_append(P, W, ncount, nxt, keep, u);
        }
        more = more && link >= 0;
        if (more) rec = load_rec(P, (uint32_t)link);
      }
    }
    // all-input states matched at p + 1
    const uint32_t li = g * 256u + b;
    const uint32_t lb = P.img.beg[0][li], le = P.img.beg[0][li + 1];
    for (uint32_t i = lb + wave * kWave; i < le; i += kBlockThreads) {
      const uint32_t idx = i + W.lane;
      const bool ok = idx < le;
      const int32_t e = ok ? P.img.list[0][idx] : 0;
      const uint32_t u = (uint32_t)e & (uint32_t)kIdMask;
      emit(P, W, ok && (e & kFlagReport), ((uint64_t)p1 << 32) | u | W.gtag);
      block_append(P, W, ncount, nxt, ok && !(e & kFlagDeadEnd), u);
    }
    if (wave == 0) W.items += n;
    __syncthreads();
    if (threadIdx.x == 0) count[sel] = 0;
    __syncthreads();
  }
  flush_reports(P, W);
  if (W.lane == 0) {
    if (W.overflow && atomicOr(&P.overflow[unit], 1u) == 0u) atomicAdd(P.overflow_count, 1u);
    add_items(P, W.items);
  }
}

// Clears the per-run counters with a kernel (graph-capture friendly).  A
// re-run of overflowed work items keeps the report count (keep_reports).
__global__ void k_reset(unsigned long long *report_count, unsigned long long *items,
                        uint32_t *task_head, uint32_t *overflow, uint32_t n_overflow,
                        uint32_t *overflow_count, int keep_reports,
                        unsigned long long *root_count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    if (!keep_reports) *report_count = 0;
    for (uint32_t k = 0; k < kItemLanes; k++) items[k * kItemStride] = 0;
    for (uint32_t k = 0; k < 2 * kQueues; k++) task_head[k * kQueueStride] = 0;
    *overflow_count = 0;
  }
  for (; i < n_overflow; i += gridDim.x * blockDim.x) {
    overflow[i] = 0;
    if (root_count) root_count[i] = 0;
  }
}

// Host-side launch wrappers (this translation unit owns the kernels).
hipError_t launch_reset(hipStream_t s, unsigned long long *rc, unsigned long long *items,
                        uint32_t *head, uint32_t *ovf, uint32_t n, uint32_t *ovf_count,
                        int keep_reports, unsigned long long *root_count) {
  uint32_t blocks = (n + 255) / 256;
  if (blocks == 0) blocks = 1;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_reset, dim3(blocks), dim3(256), 0, s, rc, items, head, ovf, n,
                     ovf_count, keep_reports, root_count);
  return hipGetLastError();
}

hipError_t launch_plain(hipStream_t s, const KParams &P, uint32_t units, uint32_t lds) {
  hipLaunchKernelGGL(k_plain, dim3(units), dim3(kBlockThreads), lds, s, P);
  return hipGetLastError();
}

hipError_t launch_blocking(hipStream_t s, const KParams &P, uint32_t units, uint32_t lds) {
  hipLaunchKernelGGL(k_blocking, dim3(units), dim3(kBlockThreads), lds, s, P);
  return hipGetLastError();
}

hipError_t launch_roots(hipStream_t s, const KParams &P, uint32_t lds) {
  const uint32_t blocks = (P.num_pieces + P.waves_per_block - 1) / P.waves_per_block;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_roots, dim3(blocks), dim3(kBlockThreads), lds, s, P);
  return hipGetLastError();
}

// The kernel of each batched kind (batched_kind).
static const void *batched_kernel(int kind) {
  switch (kind) {
    case 1: return reinterpret_cast<const void *>(k_batched_stream<kEndChunkLen, false>);
    case 2: return reinterpret_cast<const void *>(k_expand);
    case 3: return reinterpret_cast<const void *>(k_batched_stream<kEndMap, false>);
    case 4: return reinterpret_cast<const void *>(k_batched_stream<kEndChunkLen, true>);
    default: return reinterpret_cast<const void *>(k_batched);
  }
}

hipError_t launch_batched(hipStream_t s, const KParams &P, uint32_t blocks, uint32_t lds) {
  const int kind = batched_kind(P);
  switch (kind) {
    case 1:
      hipLaunchKernelGGL((k_batched_stream<kEndChunkLen, false>), dim3(blocks),
                         dim3(kBlockThreads), lds, s, P);
      break;
    case 2:
      hipLaunchKernelGGL(k_expand, dim3(blocks), dim3(kBlockThreads), lds, s, P);
      break;
    case 3:
      hipLaunchKernelGGL((k_batched_stream<kEndMap, false>), dim3(blocks), dim3(kBlockThreads),
                         lds, s, P);
      break;
    case 4:
      hipLaunchKernelGGL((k_batched_stream<kEndChunkLen, true>), dim3(blocks),
                         dim3(kBlockThreads), lds, s, P);
      break;
    default:
      hipLaunchKernelGGL(k_batched, dim3(blocks), dim3(kBlockThreads), lds, s, P);
  }
  return hipGetLastError();
}

hipError_t batched_occupancy(int *blocks_per_cu, uint32_t lds, int kind) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batched_kernel(kind),
                                                      kBlockThreads, lds);
}

hipError_t set_lds_limit(uint32_t lds) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_plain),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_blocking),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if (e != hipSuccess) return e;
  for (int kind = 1; kind <= kBatchedKinds - 1; kind++) {
    e = hipFuncSetAttribute(batched_kernel(kind), hipFuncAttributeMaxDynamicSharedMemorySize,
                            lds);
    if (e != hipSuccess) return e;
  }
  e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_roots),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_expand),
                          hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(k_batched),
                             hipFuncAttributeMaxDynamicSharedMemorySize, lds);
}

}  // namespace ngapamd
