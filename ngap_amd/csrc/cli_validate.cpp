// CPU check behind the CLI's --validation flag (the reference binary's
// ngap::automataGroupsReference + automataValidation, code/src/ngap/
// ngap.cu:2171-2344).  It never produces the reported results: the CLI prints
// the GPU's reports and only uses this to print PASS/FAILED.
//
// Per (group, chunk) unit, a set frontier of non-all-input states advances
// position by position; all-input states are taken from a per-byte bucket
// instead of being re-tested at every position.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "frontend.h"

namespace ngapamd {

std::vector<uint64_t> cpu_reference(const std::vector<GroupGraph> &gs, const uint8_t *input,
                                    uint64_t chunk_len, uint32_t num_seg, bool is_dup) {
  const int64_t G = (int64_t)gs.size();
  const int64_t units = G * num_seg;
  std::vector<std::vector<uint64_t>> out(units);
  // per group: all-input states bucketed by byte, all-input flags
  std::vector<std::vector<std::vector<int32_t>>> bucket(G);
  std::vector<std::vector<char>> aas(G);
  for (int64_t j = 0; j < G; j++) {
    const GroupGraph &g = gs[j];
    bucket[j].resize(256);
    aas[j].assign(g.num_states, 0);
    for (int32_t a : g.aas) {
      aas[j][a] = 1;
      for (int b = 0; b < 256; b++)
        if ((g.sets[(size_t)a * 8 + b / 32] >> (b % 32)) & 1u) bucket[j][b].push_back(a);
    }
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t u = 0; u < units; u++) {
    const int64_t c = u / G, j = u % G;
    const GroupGraph &g = gs[j];
    const uint64_t begin = c * chunk_len, end = begin + chunk_len;
    std::vector<int32_t> cur, nxt;
    if (c == 0 || is_dup)
      for (int32_t s : g.sod) cur.push_back(s);
    std::sort(cur.begin(), cur.end());
    auto &rep = out[u];
    auto step = [&](int32_t v, uint64_t p) {
      if (g.attrs[v] & 2) {
        uint32_t id = (uint32_t)v | ((uint32_t)j << 22);
        rep.push_back(((uint64_t)id << 32) | (uint32_t)p);
      }
      for (int e = g.row[v]; e < g.row[v + 1]; e++)
        if (!aas[j][g.col[e]]) nxt.push_back(g.col[e]);
    };
    for (uint64_t p = begin; p < end; p++) {
      const uint8_t b = input[p];
      nxt.clear();
      for (int32_t v : cur)
        if ((g.sets[(size_t)v * 8 + b / 32] >> (b % 32)) & 1u) step(v, p);
      for (int32_t a : bucket[j][b]) step(a, p);
      std::sort(nxt.begin(), nxt.end());
      nxt.erase(std::unique(nxt.begin(), nxt.end()), nxt.end());
      cur.swap(nxt);
    }
  }
  std::vector<uint64_t> all;
  for (auto &v : out) all.insert(all.end(), v.begin(), v.end());
  std::sort(all.begin(), all.end(), [](uint64_t a, uint64_t b) {
    uint32_t ia = (uint32_t)a, ib = (uint32_t)b;
    return ia != ib ? ia < ib : (a >> 32) < (b >> 32);
  });
  all.erase(std::unique(all.begin(), all.end()), all.end());
  return all;
}

}  // namespace ngapamd
