// Device image of a set of NFA groups: the layout the HIP kernels read.
//
// Replaces the reference's per-group device structs (Csr, Matchset,
// GroupNodeAttrs, GroupAAS, PrecTable; code/include/commons/graph.h:13-30,
// 128-248, group_graph.h:6-109, precompute_table.h:8-100) with flat arrays,
// one allocation per kind for all groups:
//
//  * succ4  int32[4] per record.  Record v (< n) holds the successors of
//    state v that are NOT all-input states, each tagged with the successor's
//    flags; -1 pads, a value <= -2 continues at record -(value+2) (appended
//    after the group's n records) when a state has more than 4 such
//    successors.  One 16-byte load yields a state's whole successor list.
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
//    is indexed directly by byte; d = 2, 3 keep sorted non-empty keys with a
//    per-first-byte directory and a binary search inside the bucket.  This is
//    the reference's prefix memoization (ngap.cu:2358-2513,
//    precompute_table.cu:25-395) with set lists and a bucketed key index.
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
  uint32_t memo_key_base[2];  // first key index for depth 2 and 3
};

struct MemoLevel {            // one depth for all groups
  std::vector<uint32_t> key;   // sorted keys per group (d >= 2)
  std::vector<uint32_t> beg;   // [nkeys + 1] list offsets (global)
  std::vector<uint32_t> nrep;  // reporting entries at the head of each list
  std::vector<uint32_t> dir;   // d >= 2: [groups*257] key-index ranges by first byte
  std::vector<int32_t> list;   // tagged state ids
};

struct Image {
  int32_t num_groups = 0;
  int32_t memo_depth = 0;
  std::vector<GroupMeta> meta;
  std::vector<int32_t> succ4;   // 4 per record
  std::vector<uint32_t> cls;    // 8 per state
  std::vector<int32_t> sod;     // tagged SOD ids
  MemoLevel memo[kMaxMemoDepth];// memo[0] = depth 1 (always built)
  // root filters: per group 8 words (byte b has a depth-1 list) and, when
  // depth >= 2 is built, 2048 words (2-gram key has a depth-2 list)
  std::vector<uint32_t> mask1;
  std::vector<uint32_t> mask2;
  // statistics (for the learned scheduler's features and reports)
  std::vector<uint32_t> stat_edges, stat_aas, stat_sod, stat_report,
      stat_wide, stat_selfloop;
  uint64_t total_states = 0;
  uint64_t bytes() const;
};

// Validates the groups (ids < 2^22, <= 1024 groups, CSR bounds) and builds the
// image.  Returns false with err on bad input.
bool build_image(const HostGroup *groups, int32_t num_groups, int32_t memo_depth,
                 Image &img, std::string &err);

}  // namespace ngapamd
