// Device image builder (host side).  Layout rationale in layout.h.
#include "layout.h"

#include <algorithm>
#include <atomic>
#include <bitset>
#include <sstream>
#include <thread>

namespace ngapamd {

ImageSizes sizes_of(const Image &img) {
  ImageSizes z;
  z.meta = img.meta.size() * sizeof(GroupMeta);
  z.succ4 = img.succ4.size() * 4;
  z.cls = img.cls.size() * 4;
  z.sod = img.sod.size() * 4;
  for (int d = 0; d < kMaxMemoDepth; d++) {
    z.beg[d] = img.memo[d].beg.size() * 4;
    z.nrep[d] = img.memo[d].nrep.size() * 4;
    z.list[d] = img.memo[d].list.size() * 4;
    z.sel[d] = img.memo[d].sel.size() * 4;
  }
  z.mask1 = img.mask1.size() * 4;
  z.xr1 = img.xr1.size() * 4;
  z.xe1 = img.xe1.size() * 4;
  z.xmask1 = img.xmask1.size() * 4;
  z.xr2 = img.xr2.size() * 4;
  z.xe2 = img.xe2.size() * 4;
  z.xsod = img.xsod.size() * 4;
  z.xs3 = img.xs3.size() * 4;
  z.xp3 = img.xp3.size() * 4;
  z.x2rep = img.x2rep.size() * 4;
  return z;
}

namespace {

// Calls f(b) for every byte b of state v's class, in increasing order.
template <class F>
inline void for_each_byte(const uint32_t *sets, int v, F &&f) {
  for (int w = 0; w < 8; w++)
    for (uint32_t x = sets[(size_t)v * 8 + w]; x; x &= x - 1) f(w * 32 + __builtin_ctz(x));
}

inline bool has_byte(const uint32_t *sets, int v, int b) {
  return (sets[(size_t)v * 8 + (b >> 5)] >> (b & 31)) & 1u;
}

// Appends one list (reporting entries first, each part in id order).
void push_list(MemoLevel &m, std::vector<int32_t> &ids, const std::vector<int32_t> &tag) {
  std::sort(ids.begin(), ids.end(), [&](int32_t a, int32_t b) {
    bool ra = tag[a] & kFlagReport, rb = tag[b] & kFlagReport;
    if (ra != rb) return ra;
    return a < b;
  });
  uint32_t nrep = 0;
  for (int32_t v : ids) {
    if (tag[v] & kFlagReport) nrep++;
    m.list.push_back(tag[v]);
  }
  m.nrep.push_back(nrep);
  m.beg.push_back((uint32_t)m.list.size());
}

// One group's share of the image, with indices local to the group (record
// links, list offsets, key ranks); build_image rebases them when it
// concatenates the groups.  Groups are independent, so parts build in
// parallel.
struct GroupPart {
  bool ok = true;
  std::string err;
  GroupMeta gm{};
  std::vector<int32_t> succ4;        // records; links -(local index + 2)
  std::vector<int32_t> sod;          // tagged start-of-data ids
  uint32_t mask1[16] = {0};
  MemoLevel memo[kMaxMemoDepth];     // beg from 0, sel ranks local
  std::vector<uint32_t> key2, key3;  // keys of the depth-2 / depth-3 lists
};

bool fail_part(GroupPart &P, const std::string &m) {
  P.ok = false;
  P.err = m;
  return false;
}

bool build_part(const HostGroup &hg, int32_t g, int memo_depth, GroupPart &P) {
  const int n = hg.num_states;
  if (n < 0 || n > kIdMask + 1) {
    std::ostringstream o;
    o << "group " << g << " has " << n << " states; state ids are limited to 22 bits";
    return fail_part(P, o.str());
  }
  if (n > 0 && (!hg.row || !hg.sets || !hg.attrs)) return fail_part(P, "group arrays missing");
  const int E = n > 0 ? hg.row[n] : 0;
  if (E < 0) return fail_part(P, "CSR row offsets are not monotone");
  if ((E > 0 && !hg.col) || (hg.num_aas > 0 && !hg.aas) || (hg.num_sod > 0 && !hg.sod))
    return fail_part(P, "group arrays missing");
  for (int v = 0; v < n; v++)
    if (hg.row[v] > hg.row[v + 1] || hg.row[v] < 0)
      return fail_part(P, "CSR row offsets are not monotone");
  for (int e = 0; e < E; e++)
    if (hg.col[e] < 0 || hg.col[e] >= n) return fail_part(P, "CSR column index out of range");
  std::vector<char> is_aas(n, 0);
  for (int i = 0; i < hg.num_aas; i++) {
    if (hg.aas[i] < 0 || hg.aas[i] >= n) return fail_part(P, "all-input id out of range");
    is_aas[hg.aas[i]] = 1;
  }
  for (int i = 0; i < hg.num_sod; i++)
    if (hg.sod[i] < 0 || hg.sod[i] >= n) return fail_part(P, "start-of-data id out of range");

  // successor lists without all-input successors (distinct), in-degrees
  std::vector<std::vector<int32_t>> succ(n);
  std::vector<int32_t> indeg(n, 0);
  for (int v = 0; v < n; v++) {
    for (int e = hg.row[v]; e < hg.row[v + 1]; e++) {
      int u = hg.col[e];
      if (is_aas[u]) continue;
      if (std::find(succ[v].begin(), succ[v].end(), u) == succ[v].end()) succ[v].push_back(u);
    }
    std::sort(succ[v].begin(), succ[v].end());
    for (int u : succ[v]) indeg[u]++;
  }
  std::vector<int32_t> tag(n);
  for (int v = 0; v < n; v++) {
    int32_t t = v;
    if (hg.attrs[v] & 2) t |= kFlagReport;
    if (succ[v].empty()) t |= kFlagDeadEnd;
    if (indeg[v] >= 2) t |= kFlagDedup;
    tag[v] = t;
  }

  // successor test words: a one-byte class is tested by comparison, an
  // all-byte class always matches, any other class reads its class word
  std::vector<int32_t> test(n, 0);
  for (int v = 0; v < n; v++) {
    int cnt = 0, last = 0;
    for (int w = 0; w < 8; w++) {
      const uint32_t x = hg.sets[(size_t)v * 8 + w];
      if (x) last = w * 32 + 31 - __builtin_clz(x);
      cnt += __builtin_popcount(x);
    }
    if (cnt == 1) test[v] = (int32_t)(kTestLiteral | (uint32_t)last);
    else if (cnt == 256) test[v] = (int32_t)kTestAll;
  }

  P.gm.num_states = n;
  P.gm.num_aas = hg.num_aas;
  // records: n primary records, continuation records appended
  auto new_records = [&](size_t count) {
    const size_t at = P.succ4.size();
    P.succ4.resize(at + count * kRecWords, -1);
    for (size_t r = 0; r < count; r++)
      for (int i = 4; i < 8; i++) P.succ4[at + r * kRecWords + i] = 0;
    return at;
  };
  new_records((size_t)n);
  for (int v = 0; v < n; v++) {
    const auto &s = succ[v];
    size_t rec = (size_t)v * kRecWords;
    size_t k = 0;
    for (;;) {
      size_t left = s.size() - k;
      const size_t here = left <= 4 ? left : 3;
      for (size_t i = 0; i < here; i++) {
        P.succ4[rec + i] = tag[s[k + i]];
        P.succ4[rec + 4 + i] = test[s[k + i]];
      }
      if (left <= 4) break;
      k += 3;
      const size_t next = new_records(1);
      P.succ4[rec + 3] = -(int32_t)(next / kRecWords) - 2;
      rec = next;
    }
  }
  for (int i = 0; i < hg.num_sod; i++) P.sod.push_back(tag[hg.sod[i]]);

  for (int d = 0; d < kMaxMemoDepth; d++) P.memo[d].beg.push_back(0);
  // memo depth 1: all-input states by byte (all 256 keys present)
  std::vector<std::vector<int32_t>> l1(256);
  for (int i = 0; i < hg.num_aas; i++)
    for_each_byte(hg.sets, hg.aas[i], [&](int b) { l1[b].push_back(hg.aas[i]); });
  for (int b = 0; b < 256; b++) {
    std::sort(l1[b].begin(), l1[b].end());
    l1[b].erase(std::unique(l1[b].begin(), l1[b].end()), l1[b].end());
    if (!l1[b].empty()) P.mask1[b / 32] |= 1u << (b % 32);
    push_list(P.memo[0], l1[b], tag);
    if (P.memo[0].nrep.back()) P.mask1[8 + b / 32] |= 1u << (b % 32);
  }
  // deeper levels, each list's successors split by the next byte
  std::vector<char> mark(n, 0);
  std::vector<std::vector<int32_t>> bucket(256);
  std::vector<std::vector<int32_t>> prev_lists(l1.begin(), l1.end());
  for (int d = 2; d <= memo_depth; d++) {
    MemoLevel &m = P.memo[d - 1];
    // depth 2: 2048 pairs (256 first bytes x 8); depth 3: 8 pairs per depth-2 list
    m.sel.assign(prev_lists.size() * 16, 0u);
    std::vector<std::vector<int32_t>> cur_lists;
    for (size_t li = 0; li < prev_lists.size(); li++) {
      // key bits: depth 2 uses (first byte li, second byte b) inside the
      // group's 65536-key block; depth 3 uses the 256 keys below list li
      uint32_t *pairs = &m.sel[li * 16];
      std::vector<int32_t> S;
      for (int32_t v : prev_lists[li])
        for (int32_t u : succ[v])
          if (!mark[u]) { mark[u] = 1; S.push_back(u); }
      for (int32_t u : S) mark[u] = 0;
      for (auto &bk : bucket) bk.clear();
      for (int32_t u : S) for_each_byte(hg.sets, u, [&](int b) { bucket[b].push_back(u); });
      for (int w = 0; w < 8; w++) {
        pairs[2 * w + 1] = (uint32_t)m.keys();  // local rank, rebased on merge
        for (int b = w * 32; b < w * 32 + 32; b++) {
          if (bucket[b].empty()) continue;
          pairs[2 * w] |= 1u << (b & 31);
          if (d == 2) P.key2.push_back(((uint32_t)li << 8) | (uint32_t)b);
          else P.key3.push_back((P.key2[li] << 8) | (uint32_t)b);
          cur_lists.push_back(bucket[b]);
          push_list(m, bucket[b], tag);
        }
      }
    }
    prev_lists.swap(cur_lists);
  }
  return true;
}

}  // namespace

bool build_image(const HostGroup *groups, int32_t num_groups, int32_t memo_depth,
                 Image &img, std::string &err) {
  img = Image();
  if (num_groups <= 0 || num_groups > 1024) {
    err = "number of groups must be in [1, 1024] (10 group bits in report ids)";
    return false;
  }
  if (memo_depth < 1) memo_depth = 1;
  if (memo_depth > kMaxMemoDepth) {
    err = "precompute depth must be <= 3";
    return false;
  }
  img.num_groups = num_groups;
  img.memo_depth = memo_depth;
  for (int d = 0; d < kMaxMemoDepth; d++) img.memo[d].beg.push_back(0);
  std::vector<uint32_t> key2, grp2;  // 2-gram and group of every depth-2 list
  std::vector<uint32_t> key3, grp3;  // 3-gram and group of every depth-3 list

  // the groups' parts, built in parallel
  std::vector<GroupPart> parts(num_groups);
  {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nthreads = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)num_groups);
    std::atomic<int32_t> next(0);
    auto worker = [&]() {
      for (int32_t g; (g = next.fetch_add(1)) < num_groups;) build_part(groups[g], g, memo_depth, parts[g]);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; t++) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
  }
  for (int32_t g = 0; g < num_groups; g++)
    if (!parts[g].ok) {
      err = parts[g].err;
      return false;
    }
  // concatenation, rebasing the part-local indices
  for (int32_t g = 0; g < num_groups; g++) {
    GroupPart &P = parts[g];
    const HostGroup &hg = groups[g];
    const int n = hg.num_states;
    GroupMeta gm = P.gm;
    const uint32_t rec_base = (uint32_t)(img.succ4.size() / kRecWords);
    gm.rec_base = rec_base;
    gm.cls_base = (uint32_t)(img.cls.size() / 8);
    for (size_t r = 0; r < P.succ4.size(); r += kRecWords)
      if (P.succ4[r + 3] <= -2) P.succ4[r + 3] = -(int32_t)(rec_base - P.succ4[r + 3] - 2) - 2;
    img.succ4.insert(img.succ4.end(), P.succ4.begin(), P.succ4.end());
    img.cls.insert(img.cls.end(), hg.sets, hg.sets + (size_t)n * 8);
    gm.sod_begin = (uint32_t)img.sod.size();
    img.sod.insert(img.sod.end(), P.sod.begin(), P.sod.end());
    gm.sod_end = (uint32_t)img.sod.size();
    for (int32_t t : P.sod) {
      img.xsod.push_back((uint32_t)t);
      img.xsod.push_back((uint32_t)g);
    }
    img.mask1.insert(img.mask1.end(), P.mask1, P.mask1 + 16);
    for (int d = 0; d < kMaxMemoDepth; d++) {
      MemoLevel &m = img.memo[d];
      const MemoLevel &pm = P.memo[d];
      const uint32_t list_base = (uint32_t)m.list.size(), key_base = (uint32_t)m.keys();
      for (size_t i = 1; i < pm.beg.size(); i++) m.beg.push_back(list_base + pm.beg[i]);
      m.nrep.insert(m.nrep.end(), pm.nrep.begin(), pm.nrep.end());
      m.list.insert(m.list.end(), pm.list.begin(), pm.list.end());
      const size_t s0 = m.sel.size();
      m.sel.insert(m.sel.end(), pm.sel.begin(), pm.sel.end());
      for (size_t i = s0 + 1; i < m.sel.size(); i += 2) m.sel[i] += key_base;
    }
    for (uint32_t k : P.key2) {
      key2.push_back(k);
      grp2.push_back((uint32_t)g);
    }
    for (uint32_t k : P.key3) {
      key3.push_back(k);
      grp3.push_back((uint32_t)g);
    }
    img.meta.push_back(gm);
    img.total_states += n;
    P = GroupPart();  // release as we go
  }
  // cross-group index: counting sort of the lists by byte / 2-gram
  img.xr1.assign(257, 0);
  img.xmask1.assign(8, 0);
  for (int32_t g = 0; g < num_groups; g++)
    for (int b = 0; b < 256; b++)
      if (img.memo[0].nrep[(size_t)g * 256 + b]) img.xr1[b + 1]++;
  for (int b = 0; b < 256; b++) {
    if (img.xr1[b + 1]) img.xmask1[b / 32] |= 1u << (b % 32);
    img.xr1[b + 1] += img.xr1[b];
  }
  img.xe1.resize(img.xr1[256]);
  {
    std::vector<uint32_t> at(img.xr1.begin(), img.xr1.end() - 1);
    for (int32_t g = 0; g < num_groups; g++)
      for (int b = 0; b < 256; b++)
        if (img.memo[0].nrep[(size_t)g * 256 + b]) img.xe1[at[b]++] = (uint32_t)g * 256 + b;
  }
  if (memo_depth >= 2) {
    img.xr2.assign(65537, 0);
    for (uint32_t k : key2) img.xr2[k + 1]++;
    for (int k = 0; k < 65536; k++) img.xr2[k + 1] += img.xr2[k];
    img.xe2.resize((size_t)img.xr2[65536] * 4);
    std::vector<uint32_t> at(img.xr2.begin(), img.xr2.end() - 1);
    const MemoLevel &m2 = img.memo[1];
    for (size_t i = 0; i < key2.size(); i++) {  // list index i (depth-2 lists in order)
      uint32_t *e = &img.xe2[4 * (size_t)at[key2[i]]++];
      e[0] = (uint32_t)i;
      e[1] = m2.beg[i];
      e[2] = m2.beg[i + 1];
      e[3] = m2.nrep[i] | (grp2[i] << 22);  // nrep < 2^22: at most the group's states
    }
  }
  if (memo_depth >= 3) {
    // 3-gram k -> every group's depth-3 list (counting sort by key)
    const MemoLevel &m3 = img.memo[2];
    std::vector<uint32_t> start((1u << 24) + 1, 0);
    for (uint32_t k : key3) start[k + 1]++;
    for (uint32_t k = 0; k < (1u << 24); k++) start[k + 1] += start[k];
    std::vector<uint32_t> order(key3.size());
    {
      std::vector<uint32_t> at(start.begin(), start.end() - 1);
      for (size_t i = 0; i < key3.size(); i++) order[at[key3[i]]++] = (uint32_t)i;
    }
    img.xs3.assign(2u << 24, 0);
    img.xp3.reserve(m3.list.size() * 2);
    for (uint32_t k = 0; k < (1u << 24); k++) {
      const uint32_t n = start[k + 1] - start[k];
      if (n == 0) continue;
      const uint32_t li0 = order[start[k]];
      if (n == 1 && m3.beg[li0 + 1] - m3.beg[li0] == 1) {
        // one group, one state: the slot holds the state itself
        img.xs3[2 * (size_t)k] = (uint32_t)m3.list[m3.beg[li0]];
        img.xs3[2 * (size_t)k + 1] = kSlotInline | grp3[li0];
        continue;
      }
      const size_t first = img.xp3.size() / 2;
      for (uint32_t j = start[k]; j < start[k + 1]; j++) {
        const uint32_t li = order[j];
        for (uint32_t i = m3.beg[li]; i < m3.beg[li + 1]; i++) {
          img.xp3.push_back((uint32_t)m3.list[i]);
          img.xp3.push_back(grp3[li]);
        }
      }
      if (img.xp3.size() / 2 >= (1ull << 31)) {
        err = "3-gram index too large (2^31 pairs); use a smaller precompute depth";
        return false;
      }
      if (img.xp3.size() / 2 - first >= kSlotPrefixReports) {
        err = "a 3-gram has 2^30 depth-3 states or more; use a smaller precompute depth";
        return false;
      }
      img.xs3[2 * (size_t)k] = (uint32_t)first;
      img.xs3[2 * (size_t)k + 1] = (uint32_t)(img.xp3.size() / 2 - first);
    }
    // 2-grams with a reporting depth-2 list (their reports come from depth 2)
    img.x2rep.assign(2048, 0);
    for (size_t i = 0; i < key2.size(); i++)
      if (img.memo[1].nrep[i]) img.x2rep[key2[i] >> 5] |= 1u << (key2[i] & 31);
    // slots of the 3-grams whose prefixes report at depth 1 or 2
    for (uint32_t k = 0; k < (1u << 24); k++) {
      const uint32_t b0 = k >> 16, k2 = k >> 8;
      if (((img.xmask1[b0 >> 5] >> (b0 & 31)) & 1u) || ((img.x2rep[k2 >> 5] >> (k2 & 31)) & 1u))
        img.xs3[2 * (size_t)k + 1] |= kSlotPrefixReports;
    }
  }
  img.n_xsod = img.xsod.size() / 2;
  for (auto &m : img.memo)
    if (m.list.size() >= (1ull << 31)) {
      err = "memo tables too large (2^31 entries); use a smaller precompute depth";
      return false;
    }
  if (img.succ4.size() / kRecWords >= (size_t)INT32_MAX / 2) {
    err = "automata image too large";
    return false;
  }
  return true;
}

}  // namespace ngapamd
