// Host check of the device image's index tables (ngap_amd/csrc/layout.cpp):
// for random groups, every memo lookup the kernels make — depth-1 lists by
// byte, depth-2 and depth-3 rank tables, the cross-group byte and 2-gram
// index, the start-of-data pairs — is compared with the lists computed by
// brute force from the definition (states reached from all-input states by a
// d-gram, non-all-input successors only).  Prints "OK" or the first mismatch.
#include <algorithm>
#include <cstdio>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "../ngap_amd/csrc/layout.h"

using namespace ngapamd;

struct G {
  int n;
  std::vector<int32_t> row, col, aas, sod;
  std::vector<uint32_t> sets;
  std::vector<uint8_t> attrs;
};

static bool has(const G &g, int v, int b) { return (g.sets[v * 8 + b / 32] >> (b % 32)) & 1u; }

static G random_group(std::mt19937 &r, int n) {
  G g;
  g.n = n;
  g.sets.assign(n * 8, 0);
  g.attrs.assign(n, 0);
  std::vector<std::set<int>> adj(n);
  for (int v = 0; v < n; v++) {
    const int width = 1 + r() % 12;  // narrow classes over a small alphabet
    for (int i = 0; i < width; i++) {
      int b = r() % 40;
      g.sets[v * 8 + b / 32] |= 1u << (b % 32);
    }
    if (r() % 3 == 0) g.attrs[v] |= 2;
    const int deg = r() % 7;  // > 4 exercises continuation records
    for (int i = 0; i < deg; i++) adj[v].insert(r() % n);
  }
  for (int v = 0; v < n; v++)
    if (r() % 4 == 0) {
      g.attrs[v] |= 1;
      g.aas.push_back(v);
    } else if (r() % 10 == 0) {
      g.sod.push_back(v);
    }
  g.row.push_back(0);
  for (int v = 0; v < n; v++) {
    for (int u : adj[v]) g.col.push_back(u);
    g.row.push_back((int32_t)g.col.size());
  }
  return g;
}

// Brute-force memo list of a d-gram (as a set of state ids)
static std::set<int> brute(const G &g, const std::vector<int> &gram) {
  std::vector<char> aas(g.n, 0);
  for (int v : g.aas) aas[v] = 1;
  std::set<int> cur;
  for (int v : g.aas)
    if (has(g, v, gram[0])) cur.insert(v);
  for (size_t i = 1; i < gram.size(); i++) {
    std::set<int> nxt;
    for (int v : cur)
      for (int e = g.row[v]; e < g.row[v + 1]; e++) {
        int u = g.col[e];
        if (!aas[u] && has(g, u, gram[i])) nxt.insert(u);
      }
    cur.swap(nxt);
  }
  return cur;
}

static int fails = 0;
static void fail(const std::string &m) {
  if (fails++ < 5) printf("MISMATCH %s\n", m.c_str());
}

static std::set<int> list_of(const MemoLevel &m, uint32_t li, uint32_t *nrep_out) {
  std::set<int> s;
  uint32_t nrep = 0;
  bool rep_part = true;
  for (uint32_t i = m.beg[li]; i < m.beg[li + 1]; i++) {
    const int32_t e = m.list[i];
    const bool rep = e & kFlagReport;
    if (rep && !rep_part) fail("reporting entry after a non-reporting one");
    if (!rep) rep_part = false;
    nrep += rep;
    s.insert(e & kIdMask);
  }
  if (nrep != m.nrep[li]) fail("nrep");
  if (nrep_out) *nrep_out = nrep;
  return s;
}

static int rank_lookup(const std::vector<uint32_t> &sel, size_t at, uint32_t bit) {
  const uint32_t w = sel[2 * at], rank = sel[2 * at + 1];
  if (!((w >> bit) & 1u)) return -1;
  return (int)(rank + __builtin_popcount(w & ((1u << bit) - 1u)));
}

int main() {
  std::mt19937 r(12345);
  for (int trial = 0; trial < 6; trial++) {
    const int ng = 1 + trial * 3;
    std::vector<G> gs;
    for (int i = 0; i < ng; i++) gs.push_back(random_group(r, 5 + r() % 60));
    if (ng > 1) gs[1] = random_group(r, 0);  // a group without states
    std::vector<HostGroup> hg(ng);
    for (int i = 0; i < ng; i++) {
      hg[i].num_states = gs[i].n;
      hg[i].row = gs[i].row.data();
      hg[i].col = gs[i].col.data();
      hg[i].sets = gs[i].sets.data();
      hg[i].attrs = gs[i].attrs.data();
      hg[i].num_aas = (int32_t)gs[i].aas.size();
      hg[i].aas = gs[i].aas.data();
      hg[i].num_sod = (int32_t)gs[i].sod.size();
      hg[i].sod = gs[i].sod.data();
    }
    Image im;
    std::string err;
    if (!build_image(hg.data(), ng, 3, im, err)) {
      printf("build failed: %s\n", err.c_str());
      return 1;
    }
    // cross-group 2-gram entries, gathered per (group, key)
    std::vector<std::vector<int>> x2(65536 * (size_t)ng, std::vector<int>());
    for (uint32_t k = 0; k < 65536; k++)
      for (uint32_t i = im.xr2[k]; i < im.xr2[k + 1]; i++) {
        const uint32_t li = im.xe2[4 * i], g = im.xe2[4 * i + 3] >> 22;
        if (g >= (uint32_t)ng) fail("xe2 group");
        else x2[(size_t)g * 65536 + k].push_back((int)li);
        if (im.xe2[4 * i + 1] != im.memo[1].beg[li] || im.xe2[4 * i + 2] != im.memo[1].beg[li + 1] ||
            (im.xe2[4 * i + 3] & kIdMask) != im.memo[1].nrep[li])
          fail("xe2 list range");
      }
    std::vector<std::vector<int>> x1(256);
    for (int b = 0; b < 256; b++) {
      if (((im.xmask1[b / 32] >> (b % 32)) & 1u) != (im.xr1[b + 1] > im.xr1[b]))
        fail("xmask1");
      for (uint32_t i = im.xr1[b]; i < im.xr1[b + 1]; i++) x1[b].push_back((int)im.xe1[i]);
    }
    for (int gi = 0; gi < ng; gi++) {
      const G &g = gs[gi];
      for (int b0 = 0; b0 < 42; b0++) {
        const uint32_t li1 = gi * 256 + b0;
        uint32_t nrep1 = 0;
        const std::set<int> want1 = brute(g, {b0});
        if (list_of(im.memo[0], li1, &nrep1) != want1) fail("depth 1");
        const uint32_t *m1 = &im.mask1[(size_t)gi * 16];
        if (((m1[b0 / 32] >> (b0 % 32)) & 1u) != !want1.empty()) fail("mask1 roots");
        if (((m1[8 + b0 / 32] >> (b0 % 32)) & 1u) != (nrep1 > 0)) fail("mask1 reports");
        const bool in_x1 =
            std::find(x1[b0].begin(), x1[b0].end(), (int)li1) != x1[b0].end();
        if (in_x1 != (nrep1 > 0)) fail("xe1 membership");
        for (int b1 = 0; b1 < 42; b1++) {
          const uint32_t key = (b0 << 8) | b1;
          const std::set<int> want2 = brute(g, {b0, b1});
          const int li2 = rank_lookup(im.memo[1].sel, (size_t)gi * 2048 + (key >> 5), key & 31);
          const auto &xe = x2[(size_t)gi * 65536 + key];
          if (want2.empty()) {
            if (li2 >= 0) fail("depth 2 present for an empty list");
            if (!xe.empty()) fail("xe2 entry for an empty list");
            continue;
          }
          if (li2 < 0) {
            fail("depth 2 missing");
            continue;
          }
          if (list_of(im.memo[1], li2, nullptr) != want2) fail("depth 2 list");
          if (xe.size() != 1 || xe[0] != li2) fail("xe2 entry");
          uint32_t nrep2 = 0;
          list_of(im.memo[1], li2, &nrep2);
          if (nrep2 && !((im.x2rep[key >> 5] >> (key & 31)) & 1u)) fail("x2rep");
          for (int b2 = 0; b2 < 42; b2++) {
            const std::set<int> want3 = brute(g, {b0, b1, b2});
            const int li3 = rank_lookup(im.memo[2].sel, (size_t)li2 * 8 + (b2 >> 5), b2 & 31);
            if (want3.empty() != (li3 < 0)) fail("depth 3 presence");
            else if (li3 >= 0 && list_of(im.memo[2], li3, nullptr) != want3) fail("depth 3 list");
            // the 3-gram cross-group slot: this group's list, inline or
            // through its entry
            const uint32_t k3 = (key << 8) | (uint32_t)b2;
            const uint32_t s0 = im.xs3[2 * (size_t)k3], s1f = im.xs3[2 * (size_t)k3 + 1];
            const uint32_t s1 = s1f & ~kSlotPrefixReports;
            // the prefix flag: a reporting depth-1 list for b0 or depth-2
            // list for (b0, b1) in some group
            const bool pre = ((im.xmask1[b0 >> 5] >> (b0 & 31)) & 1u) ||
                             ((im.x2rep[key >> 5] >> (key & 31)) & 1u);
            if (pre != ((s1f & kSlotPrefixReports) != 0)) fail("3-gram prefix flag");
            int found = 0;
            if (s1 & kSlotInline) {
              if ((s1 & 1023u) == (uint32_t)gi) {
                found++;
                if (want3 != std::set<int>{(int)(s0 & kIdMask)}) fail("inline slot");
              }
            } else {
              // this group's pairs: one contiguous run holding its list
              std::set<int> got;
              uint32_t first = 0, last = 0, cnt = 0;
              for (uint32_t i = s0; i < s0 + s1; i++) {
                if (im.xp3[2 * (size_t)i + 1] != (uint32_t)gi) continue;
                if (cnt == 0) first = i;
                last = i;
                cnt++;
                got.insert(im.xp3[2 * (size_t)i] & kIdMask);
              }
              if (cnt) {
                found++;
                if (got != want3 || got.size() != cnt) fail("xp3 pairs");
                if (last - first + 1 != cnt) fail("xp3 run not contiguous");
              }
              if (s1 == 1) fail("a one-state slot not inline");
            }
            if (found != (want3.empty() ? 0 : 1)) fail("xp3 group");
          }
        }
      }
    }
    // successor records: each state's non-all-input successors with their
    // flags and tests, through continuation records
    for (int gi = 0; gi < ng; gi++) {
      const G &g = gs[gi];
      std::vector<char> aas(g.n, 0);
      for (int v : g.aas) aas[v] = 1;
      const uint32_t base = im.meta[gi].rec_base;
      if (im.meta[gi].cls_base + (uint32_t)g.n > im.cls.size() / 8) fail("cls base");
      for (int v = 0; v < g.n; v++) {
        std::set<int> want;
        for (int e = g.row[v]; e < g.row[v + 1]; e++)
          if (!aas[g.col[e]]) want.insert(g.col[e]);
        std::set<int> got;
        int64_t r = base + v;
        for (int hops = 0; r >= 0 && hops < 100; hops++) {
          const int32_t *rec = &im.succ4[(size_t)r * kRecWords];
          int64_t next = -1;
          for (int s = 0; s < 4; s++) {
            const int32_t e = rec[s];
            if (e <= -2) { next = -(int64_t)e - 2; continue; }
            if (e < 0) continue;
            const int u = e & kIdMask;
            got.insert(u);
            int cnt = 0;
            for (int b = 0; b < 256; b++) cnt += has(g, u, b);
            const uint32_t t = (uint32_t)rec[4 + s];
            if (cnt == 1) {
              if (!(t & kTestLiteral) || !has(g, u, (int)(t & 255))) fail("literal test");
            } else if (cnt == 256) {
              if (t != kTestAll) fail("all-byte test");
            } else if (t != 0) {
              fail("class test");
            }
            if (((e & kFlagReport) != 0) != ((g.attrs[u] & 2) != 0)) fail("report flag");
          }
          r = next;
        }
        if (got != want) fail("successor set");
      }
    }
    // start-of-data pairs: every group's entries, in group order
    size_t at = 0;
    for (int gi = 0; gi < ng; gi++)
      for (int v : gs[gi].sod) {
        if (at + 1 >= im.xsod.size() || (int)(im.xsod[at] & kIdMask) != v ||
            (int)im.xsod[at + 1] != gi)
          fail("xsod");
        at += 2;
      }
    if (at != im.xsod.size()) fail("xsod size");
  }
  if (fails) {
    printf("FAILED (%d mismatches)\n", fails);
    return 1;
  }
  printf("OK\n");
  return 0;
}
