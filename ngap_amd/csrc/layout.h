// Device image of a set of NFA groups: the layout the HIP kernels read.
//
// Replaces the reference's per-group device structs (Csr, Matchset,
// GroupNodeAttrs, GroupAAS, PrecTable; code/include/commons/graph.h:13-30,
// 128-248, group_graph.h:6-109, precompute_table.h:8-100) with flat arrays,
// one allocation per kind for all groups:
//
//  * succ4  int32[8] per 32-byte record.  Words 0-3 of record v (< n) hold
//    the successors of state v that are NOT all-input states, each tagged
//    with the successor's flags; -1 pads, a value <= -2 continues at record
//    -(value+2) (appended after the group's n records) when a state has more
//    than 4 such successors.  Words 4-7 hold each successor's test: a class
//    of one byte b is kTestLiteral | b (compared, no class word read), the
//    class of all 256 bytes is kTestAll, any other class is 0 (read its class
//    word).  One 32-byte record (half a cache line) yields a state's
//    successor list and, for literal successors, their tests.
//    All-input successors are dropped: an all-input state is enabled at every
//    position anyway and enters as a root there, so the reference's re-adding
//    it is a duplicate (kernel_ngap_OA.cu re-tests it; the oracle's set
//    semantics merge it).
//  * cls    uint32[8] per state: 256-bit symbol class, bit b in word b/32
//    (the reference's AoS Matchset layout).
//  * memo tables by depth d = 1..3 over the d-gram input[p..p+d-1]:
//      d = 1: all-input states matching input[p]            (matched at p)
//      d = 2: non-all-input successors of d=1 entries that match input[p+1]
//      d = 3: ... of d=2 entries that match input[p+2]
//    Each list is a set (no duplicates) with reporting states first.  d = 1
//    is indexed directly by byte; d = 2, 3 through rank tables (one 8-byte
//    load and a popcount per lookup, MemoLevel below).  This is the
//    reference's prefix memoization (ngap.cu:2358-2513,
//    precompute_table.cu:25-395) with set lists and rank-table indexes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ngapamd {

// successor-entry flags (bits above the 22-bit state id)
constexpr int32_t kIdMask = (1 << 22) - 1;
constexpr int32_t kFlagReport = 1 << 22;    // state reports when matched
constexpr int32_t kFlagDeadEnd = 1 << 23;   // no non-all-input successors
constexpr int32_t kFlagDedup = 1 << 24;     // reachable along >= 2 edges
constexpr int kMaxMemoDepth = 3;
constexpr int kRecWords = 8;                  // int32 words per successor record
constexpr uint32_t kTestLiteral = 1u << 31;   // successor test: one byte (bits 0-7)
constexpr uint32_t kTestAll = 1u << 30;       // successor test: every byte
constexpr uint32_t kSlotInline = 1u << 31;    // 3-gram slot holds its one state
// 3-gram slot flag: the first byte begins a reporting depth-1 list, or the
// first two bytes a reporting depth-2 list, in some group
constexpr uint32_t kSlotPrefixReports = 1u << 30;

struct HostGroup {  // what the C ABI hands in for one group
  int32_t num_states = 0;
  const int32_t *row = nullptr;
  const int32_t *col = nullptr;
  const uint32_t *sets = nullptr;
  const uint8_t *attrs = nullptr;
  int32_t num_aas = 0;
  const int32_t *aas = nullptr;
  int32_t num_sod = 0;
  const int32_t *sod = nullptr;
};

struct GroupMeta {            // per group, read by every kernel (32 B)
  uint32_t rec_base;          // first succ4 record
  uint32_t cls_base;          // first state in cls
  uint32_t sod_begin, sod_end;// start-of-data entries (tagged ids) in sod[]
  uint32_t num_states;
  uint32_t num_aas;
  uint32_t pad[2];
};

// Memo level d holds one list per d-gram key that has one.  Keys of a level
// are numbered in (group, key) order; list k is list[beg[k], beg[k+1]) with
// its nrep[k] reporting entries first.  Depth 1 is indexed by g*256 + byte
// (all 256 keys present).  Depths 2 and 3 are found through rank tables of
// (presence word, rank) pairs: the 32 keys x..x+31 share one pair whose word
// has bit i set when key x+i has a list and whose rank is the index of the
// first such key, so key k's index is rank + popcount(word & ((1<<i)-1)).
//   depth 2: pair g*2048 + (2-gram >> 5)
//   depth 3: pair k2*8 + (third byte >> 5), k2 = index of the first two bytes'
//            depth-2 key (8 pairs per depth-2 key)
struct MemoLevel {            // one depth for all groups
  std::vector<uint32_t> beg;   // [nkeys + 1] list offsets (global)
  std::vector<uint32_t> nrep;  // reporting entries at the head of each list
  std::vector<uint32_t> sel;   // d >= 2: rank table, (word, rank) pairs
  std::vector<int32_t> list;   // tagged state ids
  uint64_t keys() const { return nrep.size(); }
};

struct Image {
  int32_t num_groups = 0;
  int32_t memo_depth = 0;
  std::vector<GroupMeta> meta;
  std::vector<int32_t> succ4;   // kRecWords per record
  std::vector<uint32_t> cls;    // 8 per state
  std::vector<int32_t> sod;     // tagged SOD ids
  MemoLevel memo[kMaxMemoDepth];// memo[0] = depth 1 (always built)
  // root filters: per group 16 words; bit b of words 0-7 is set when byte b
  // has a depth-1 list, of words 8-15 when that list has reporting entries
  std::vector<uint32_t> mask1;
  // Cross-group root index (memo depth >= 2), for work items that scan one
  // window for every group at once:
  //   x1: byte b -> depth-1 lists with reporting entries, xe1[xr1[b]..xr1[b+1])
  //       holding list indices g*256 + b; xmask1 marks bytes with any;
  //   x2: 2-gram k -> depth-2 lists of every group: entries xr2[k] ..
  //       xr2[k+1]-1 of xe2, each 4 words (list index, list begin, list end,
  //       reporting entries | group << 22), one 16-byte load per entry;
  //   xsod: (tagged id, group) pairs of every group's start-of-data states.
  //   x3 (depth 3): 3-gram k -> every group's depth-3 list, through the
  //       8-byte slot xs3[2k], xs3[2k+1]: (0, 0) when no group has one;
  //       (tagged state, kSlotInline | group) when exactly one group has one
  //       and it holds one state (the common case: one load, no list);
  //       otherwise (first pair, pair count) into xp3, whose 8-byte pairs
  //       (tagged state, group) are the states of every group's list, group
  //       by group in list order: a lookup reads the slot, then its pairs,
  //       with no list indirection.  Bit 30 of the second word
  //       (kSlotPrefixReports) flags the 3-grams whose first byte has a
  //       reporting depth-1 list or whose first two bytes have a reporting
  //       depth-2 list in some group (the count is the word without it).
  //       x2rep marks the 2-grams that have a reporting depth-2 list (2048
  //       words).
  std::vector<uint32_t> xr1, xe1, xmask1, xr2, xe2, xsod, xs3, xp3, x2rep;
  uint64_t n_xsod = 0;          // (state, group) pairs in xsod
  uint64_t total_states = 0;
};

// Bytes of each image array as built (the device builder keeps the arrays on
// the device only; the host builder's vectors give the same numbers).
struct ImageSizes {
  uint64_t meta = 0, succ4 = 0, cls = 0, sod = 0;
  uint64_t beg[kMaxMemoDepth] = {0, 0, 0}, nrep[kMaxMemoDepth] = {0, 0, 0},
           list[kMaxMemoDepth] = {0, 0, 0}, sel[kMaxMemoDepth] = {0, 0, 0};
  uint64_t mask1 = 0, xr1 = 0, xe1 = 0, xmask1 = 0, xr2 = 0, xe2 = 0, xsod = 0, xs3 = 0,
           xp3 = 0, x2rep = 0;
  uint64_t total() const {
    uint64_t t = meta + succ4 + cls + sod + mask1 + xr1 + xe1 + xmask1 + xr2 + xe2 + xsod + xs3 +
                 xp3 + x2rep;
    for (int d = 0; d < kMaxMemoDepth; d++) t += beg[d] + nrep[d] + list[d] + sel[d];
    return t;
  }
};
ImageSizes sizes_of(const Image &img);

// Validates the groups (ids < 2^22, <= 1024 groups, CSR bounds) and builds the
// image.  Returns false with err on bad input.
bool build_image(const HostGroup *groups, int32_t num_groups, int32_t memo_depth,
                 Image &img, std::string &err);

}  // namespace ngapamd
