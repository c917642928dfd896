// ORACLE — test infrastructure only.  Nothing in the product links, loads or
// calls this file; only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg use it, and only as the checker.
//
// CPU restatement of ngAP's validator, ngap::automataGroupsReference
// (code/src/ngap/ngap.cu:2171-2268), with set semantics (--unique=true,
// ngap.cu:2246-2251), plus the ordering / de-duplication of
// ngap::automataValidation (ngap.cu:2270-2344, compareResult ngap.cu:19-30).
//
// For group j and chunk c covering positions [cL, (c+1)L):
//   enabled(cL)   = AAS u (SOD if c == 0 or is_dup)          ngap.cu:2204-2210
//   matched(p)    = {v in enabled(p) : input[p] in class(v)}  ngap.cu:2226-2228
//   report (v | j<<22, p) for reporting v in matched(p)       ngap.cu:2229-2232
//   enabled(p+1)  = succ(matched(p)) u AAS                    ngap.cu:2233-2256
// Output: ((state | group<<22) << 32) | offset, sorted by (offset, id), unique.
//
// Deliberate difference: the reference indexes its two frontiers by
// iter % 2 with iter starting at c*L, so for odd c*L it starts on the empty
// frontier and reports nothing for that chunk (ngap.cu:2200-2212).  The GPU
// kernels of the reference do process such chunks; this oracle follows the
// kernels (and the stated semantics above) and treats every chunk alike.
//
// It also counts the algorithmic work the roofline uses (per unit, distinct
// items): T_ne tested non-AAS enabled states, M_aas matched AAS, M matched
// states that advance (position < chunk end - 1), S their successor ids, R
// reports.  For a prefix-memoized engine of depth k it further tracks, per
// matched item, the longest chain from an all-input root (capped at k; items
// seeded from start-of-data count as k): an item whose every chain is shorter
// than k-1 is covered by the memo tables and never materialised; an item
// with depth k-1 is read from a memo list; materialised items that advance
// are expanded (one successor record, one class word per non-all-input
// successor).
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

struct oracle_group {
  int32_t num_states;
  const int32_t *row;     // [n+1]
  const int32_t *col;     // [E]
  const uint32_t *sets;   // [n*8]
  const uint8_t *attrs;   // bit0 all-input, bit1 report
  int32_t num_aas;
  const int32_t *aas;
  int32_t num_sod;
  const int32_t *sod;
};

struct oracle_stats {
  uint64_t tested_non_aas;   // T_ne
  uint64_t matched_aas;      // M_aas
  uint64_t matched_advance;  // M
  uint64_t successors;       // S
  uint64_t reports;          // R (distinct)
  uint64_t enabled_items;    // all distinct enabled (state, pos)
  uint64_t matched_items;    // all distinct matched (state, pos)
  uint64_t memo_list_items;  // materialised from a depth-k memo list
  uint64_t expanded_items;   // materialised items that advance
  uint64_t expanded_tests;   // their non-all-input successors
  uint64_t root_positions;   // positions where at least one all-input state matched
  uint64_t level2_lists;     // memo depth >= 2: positions where a state one
                             // step below an all-input root matched (a
                             // non-empty depth-2 list)
  uint64_t deepest_lists;    // positions where a state k-1 steps below an
                             // all-input root matched (a non-empty list of the
                             // deepest memo level k)
};

// Runs the oracle.  `unit_mask` (nullable, length num_groups*num_seg, index
// c*num_groups + j) selects units; `threads` <= 0 means all cores.  Reports
// are written to out (capacity cap) in reference order; *n_out gets the full
// count even when it exceeds cap.  Returns 0, or 1 when out was too small.
int oracle_run_ex(int32_t num_groups, const oracle_group *groups,
                  const uint8_t *input, uint32_t chunk_len, uint32_t num_seg,
                  int32_t is_dup, const uint8_t *unit_mask, int32_t threads,
                  int32_t memo_depth, int32_t implicit_aas, uint64_t *out, uint64_t cap,
                  uint64_t *n_out, oracle_stats *stats_out) {
  const int K = memo_depth < 1 ? 1 : memo_depth;
  const int64_t units = (int64_t)num_groups * num_seg;
  std::vector<std::vector<uint64_t>> per_unit(units);
  std::vector<oracle_stats> per_stats(units);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t u = 0; u < units; u++) {
    if (unit_mask && !unit_mask[u]) continue;
    const int c = (int)(u / num_groups), j = (int)(u % num_groups);
    const oracle_group &g = groups[j];
    oracle_stats st;
    memset(&st, 0, sizeof st);
    std::vector<uint64_t> &rep = per_unit[u];
    std::vector<char> is_aas(g.num_states, 0);
    for (int i = 0; i < g.num_aas; i++) is_aas[g.aas[i]] = 1;
    std::vector<int32_t> nonaas_deg(g.num_states, 0);
    for (int v = 0; v < g.num_states; v++)
      for (int e = g.row[v]; e < g.row[v + 1]; e++) nonaas_deg[v] += !is_aas[g.col[e]];
    // frontier entries: (state, root depth capped at K); duplicates merged
    // keeping the deepest chain
    typedef std::pair<int32_t, int32_t> SD;
    std::vector<SD> cur, nxt;
    auto merge = [](std::vector<SD> &f) {
      std::sort(f.begin(), f.end(), [](const SD &a, const SD &b) {
        return a.first != b.first ? a.first < b.first : a.second > b.second;
      });
      size_t w = 0;
      for (size_t i = 0; i < f.size(); i++)
        if (w == 0 || f[w - 1].first != f[i].first) f[w++] = f[i];
      f.resize(w);
    };
    if (c == 0 || is_dup)
      for (int i = 0; i < g.num_sod; i++) cur.emplace_back(g.sod[i], K);
    for (int i = 0; i < g.num_aas; i++) cur.emplace_back(g.aas[i], 0);
    merge(cur);
    const uint64_t begin = (uint64_t)c * chunk_len, end = begin + chunk_len;
    // implicit_aas: the same frontier with the all-input states left out of
    // it after the first position.  They are enabled at every position with
    // chain depth 0 (they are re-added with depth 0 and no edge leads to them,
    // ngap.cu:2233-2256), so each position visits the all-input states whose
    // class holds its byte (a per-byte list) instead of testing every one, and
    // counts them as enabled; everything it reports and counts is what the
    // loop below would.
    std::vector<std::vector<int32_t>> aas_by_byte;
    if (implicit_aas && g.num_aas > 0) {
      aas_by_byte.resize(256);
      std::vector<int32_t> ids(g.aas, g.aas + g.num_aas);
      std::sort(ids.begin(), ids.end());
      ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
      for (int32_t v : ids)
        for (int b = 0; b < 256; b++)
          if ((g.sets[(size_t)v * 8 + b / 32] >> (b % 32)) & 1u) aas_by_byte[b].push_back(v);
    }
    const int64_t n_aas_distinct = [&] {
      std::vector<int32_t> ids(g.aas, g.aas + g.num_aas);
      std::sort(ids.begin(), ids.end());
      return (int64_t)(std::unique(ids.begin(), ids.end()) - ids.begin());
    }();
    if (implicit_aas && g.num_aas > 0 && chunk_len > 0) {
      // first position: the general loop (start-of-data seeds may share ids
      // with all-input states); then every later position with the all-input
      // states implicit
      std::vector<SD> first;
      first.swap(cur);
      for (uint64_t p = begin; p < end; p++) {
        const uint8_t b = input[p];
        nxt.clear();
        const bool at_first = p == begin;
        const std::vector<SD> &front = at_first ? first : cur;
        st.enabled_items += front.size() + (at_first ? 0 : (uint64_t)n_aas_distinct);
        bool root_here = false, list2_here = false, deep_here = false;
        auto visit = [&](int32_t v, int32_t d) {
          st.matched_items++;
          if (is_aas[v]) {
            st.matched_aas++;
            root_here = true;
          }
          if (g.attrs[v] & 2) {
            uint32_t id = (uint32_t)v | ((uint32_t)j << 22);
            rep.push_back(((uint64_t)id << 32) | (uint32_t)p);
          }
          const bool materialised = d >= K - 1;
          if (d == K - 1) st.memo_list_items++;
          if (d == 1) list2_here = true;
          if (d == K - 1) deep_here = true;
          if (p + 1 < end) {
            st.matched_advance++;
            st.successors += (uint64_t)(g.row[v + 1] - g.row[v]);
            if (materialised && nonaas_deg[v] > 0) {
              st.expanded_items++;
              st.expanded_tests += nonaas_deg[v];
            }
          }
          const int32_t dn = d + 1 < K ? d + 1 : K;
          for (int e = g.row[v]; e < g.row[v + 1]; e++)
            if (!is_aas[g.col[e]]) nxt.emplace_back(g.col[e], dn);
        };
        for (const SD &vd : front) {
          const int32_t v = vd.first, d = vd.second;
          if (!is_aas[v]) st.tested_non_aas++;
          if (!((g.sets[(size_t)v * 8 + b / 32] >> (b % 32)) & 1u)) continue;
          visit(v, d);
        }
        if (!at_first)
          for (int32_t v : aas_by_byte[b]) visit(v, 0);
        st.root_positions += root_here;
        st.level2_lists += list2_here;
        st.deepest_lists += deep_here;
        merge(nxt);
        cur.swap(nxt);
      }
      st.reports = rep.size();
      per_stats[u] = st;
      continue;
    }
    for (uint64_t p = begin; p < end && !cur.empty(); p++) {
      const uint8_t b = input[p];
      nxt.clear();
      st.enabled_items += cur.size();
      bool root_here = false, list2_here = false, deep_here = false;
      for (const SD &vd : cur) {
        const int32_t v = vd.first, d = vd.second;
        if (!is_aas[v]) st.tested_non_aas++;
        if (!((g.sets[(size_t)v * 8 + b / 32] >> (b % 32)) & 1u)) continue;
        st.matched_items++;
        if (is_aas[v]) {
          st.matched_aas++;
          root_here = true;
        }
        if (g.attrs[v] & 2) {
          uint32_t id = (uint32_t)v | ((uint32_t)j << 22);
          rep.push_back(((uint64_t)id << 32) | (uint32_t)p);
        }
        const bool materialised = d >= K - 1;
        if (d == K - 1) st.memo_list_items++;
        if (d == 1) list2_here = true;
        if (d == K - 1) deep_here = true;
        if (p + 1 < end) {
          st.matched_advance++;
          st.successors += (uint64_t)(g.row[v + 1] - g.row[v]);
          if (materialised && nonaas_deg[v] > 0) {
            st.expanded_items++;
            st.expanded_tests += nonaas_deg[v];
          }
        }
        // an all-input successor is re-enabled as a root (depth 0) below
        const int32_t dn = d + 1 < K ? d + 1 : K;
        for (int e = g.row[v]; e < g.row[v + 1]; e++)
          if (!is_aas[g.col[e]]) nxt.emplace_back(g.col[e], dn);
      }
      st.root_positions += root_here;
      st.level2_lists += list2_here;
      st.deepest_lists += deep_here;
      for (int i = 0; i < g.num_aas; i++) nxt.emplace_back(g.aas[i], 0);
      merge(nxt);
      cur.swap(nxt);
    }
    st.reports = rep.size();
    per_stats[u] = st;
  }
  std::vector<uint64_t> all;
  oracle_stats tot;
  memset(&tot, 0, sizeof tot);
  size_t total = 0;
  for (int64_t u = 0; u < units; u++) total += per_unit[u].size();
  all.reserve(total);
  for (int64_t u = 0; u < units; u++) {
    all.insert(all.end(), per_unit[u].begin(), per_unit[u].end());
    std::vector<uint64_t>().swap(per_unit[u]);
    tot.tested_non_aas += per_stats[u].tested_non_aas;
    tot.matched_aas += per_stats[u].matched_aas;
    tot.matched_advance += per_stats[u].matched_advance;
    tot.successors += per_stats[u].successors;
    tot.reports += per_stats[u].reports;
    tot.enabled_items += per_stats[u].enabled_items;
    tot.matched_items += per_stats[u].matched_items;
    tot.memo_list_items += per_stats[u].memo_list_items;
    tot.expanded_items += per_stats[u].expanded_items;
    tot.expanded_tests += per_stats[u].expanded_tests;
    tot.root_positions += per_stats[u].root_positions;
    tot.level2_lists += per_stats[u].level2_lists;
    tot.deepest_lists += per_stats[u].deepest_lists;
  }
  // compareResult order: by offset (low word), then by id (high word)
  std::sort(all.begin(), all.end(), [](uint64_t a, uint64_t b) {
    uint32_t ia = (uint32_t)a, ib = (uint32_t)b;
    if (ia != ib) return ia < ib;
    return (a >> 32) < (b >> 32);
  });
  all.erase(std::unique(all.begin(), all.end()), all.end());
  *n_out = all.size();
  if (stats_out) *stats_out = tot;
  if (all.size() > cap) {
    if (out) std::copy(all.begin(), all.begin() + cap, out);
    return 1;
  }
  if (out) std::copy(all.begin(), all.end(), out);
  return 0;
}

// The reference loop as written (every all-input state in the frontier).
int oracle_run(int32_t num_groups, const oracle_group *groups, const uint8_t *input,
               uint32_t chunk_len, uint32_t num_seg, int32_t is_dup, const uint8_t *unit_mask,
               int32_t threads, int32_t memo_depth, uint64_t *out, uint64_t cap,
               uint64_t *n_out, oracle_stats *stats_out) {
  return oracle_run_ex(num_groups, groups, input, chunk_len, num_seg, is_dup, unit_mask,
                       threads, memo_depth, 0, out, cap, n_out, stats_out);
}

}  // extern "C"
