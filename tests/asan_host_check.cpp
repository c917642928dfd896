// The host-side front-end (ANML -> groups), image builder and the CLI's
// --validation check compiled with AddressSanitizer and UBSan on the CPU:
// every ANML file named on the command line goes through the whole pipeline
// (two group counts) and the image builder at precompute depths 1-3, and the
// CLI's CPU check (cli_validate.cpp) must give the oracle's reports on seeded
// inputs over the automata's alphabet (one chunk and eight, with and without
// duplicated chunks); then malformed groups, each of which the builder must
// reject with a message and without touching memory it does not own.  Prints
// PASS or the first failures.
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../ngap_amd/csrc/frontend.h"
#include "../ngap_amd/csrc/layout.h"
// the oracle, the checker, compiled into this test alone
#include "../oracle/oracle.cpp"

namespace ngapamd {
std::vector<uint64_t> cpu_reference(const std::vector<GroupGraph> &gs, const uint8_t *input,
                                    uint64_t chunk_len, uint32_t num_seg, bool is_dup);
}

using namespace ngapamd;

static int fails = 0;
static void fail(const std::string &what) {
  std::printf("FAIL %s\n", what.c_str());
  fails++;
}

static HostGroup host_of(const GroupGraph &g) {
  HostGroup h;
  h.num_states = g.num_states;
  h.row = g.row.data();
  h.col = g.col.data();
  h.sets = g.sets.data();
  h.attrs = g.attrs.data();
  h.num_aas = (int32_t)g.aas.size();
  h.aas = g.aas.data();
  h.num_sod = (int32_t)g.sod.size();
  h.sod = g.sod.data();
  return h;
}

// The CLI's CPU check against the oracle on seeded inputs over the bytes the
// automata's classes hold (so that chains advance).
static void check_validator(const char *path, const std::vector<GroupGraph> &groups) {
  std::vector<int> alphabet;
  for (int b = 0; b < 256; b++) {
    bool any = false;
    for (const GroupGraph &g : groups)
      for (int v = 0; v < g.num_states && !any; v++)
        any = (g.sets[(size_t)v * 8 + b / 32] >> (b % 32)) & 1u;
    if (any) alphabet.push_back(b);
  }
  if (alphabet.empty()) return;
  std::mt19937 rng(1234);
  const uint32_t n = 4096;
  std::vector<uint8_t> input(n);
  for (auto &x : input)
    x = rng() % 8 == 0 ? (uint8_t)(rng() % 256) : (uint8_t)alphabet[rng() % alphabet.size()];
  std::vector<oracle_group> og;
  for (const GroupGraph &g : groups)
    og.push_back({g.num_states, g.row.data(), g.col.data(), g.sets.data(), g.attrs.data(),
                  (int32_t)g.aas.size(), g.aas.data(), (int32_t)g.sod.size(), g.sod.data()});
  for (uint32_t chunk : {n, n / 8})
    for (int dup : {0, 1}) {
      const std::vector<uint64_t> got = cpu_reference(groups, input.data(), chunk, n / chunk, dup);
      std::vector<uint64_t> want(got.size() + 4096);
      uint64_t nw = 0;
      oracle_stats st;
      if (oracle_run((int32_t)og.size(), og.data(), input.data(), chunk, n / chunk, dup, nullptr,
                     1, 1, want.data(), want.size(), &nw, &st) != 0) {
        fail(std::string(path) + ": oracle output larger than expected");
        continue;
      }
      want.resize(nw);
      if (got != want)
        fail(std::string(path) + ": CLI validator differs from the oracle (chunk " +
             std::to_string(chunk) + ", dup " + std::to_string(dup) + ": " +
             std::to_string(got.size()) + " vs " + std::to_string(want.size()) + " reports)");
    }
}

static void check_file(const char *path) {
  Nfa nfa;
  std::string err;
  if (!load_anml_file(path, nfa, err)) return fail(std::string(path) + ": " + err);
  for (int gn : {1, 10}) {
    FrontendOptions opt;
    opt.group_num = gn;
    std::vector<GroupGraph> groups;
    FrontendStats st;
    if (!build_groups(nfa, opt, groups, st, err)) return fail(std::string(path) + ": " + err);
    if (dump_groups(groups).empty()) fail(std::string(path) + ": empty dump");
    check_validator(path, groups);
    std::vector<HostGroup> hg;
    for (const GroupGraph &g : groups) hg.push_back(host_of(g));
    for (int depth = 1; depth <= 3; depth++) {
      Image img;
      if (!build_image(hg.data(), (int32_t)hg.size(), depth, img, err))
        fail(std::string(path) + " depth " + std::to_string(depth) + ": " + err);
      else if (sizes_of(img).total() == 0)
        fail(std::string(path) + ": empty image");
    }
  }
}

// a 3-state group: 0 (all-input, 'a') -> 1 ('b') -> 2 ('c', reporting)
struct Small {
  std::vector<int32_t> row{0, 1, 2, 2}, col{1, 2}, aas{0}, sod{};
  std::vector<uint32_t> sets = std::vector<uint32_t>(24, 0);
  std::vector<uint8_t> attrs{1, 0, 2};
  Small() {
    sets[0 * 8 + 'a' / 32] |= 1u << ('a' % 32);
    sets[1 * 8 + 'b' / 32] |= 1u << ('b' % 32);
    sets[2 * 8 + 'c' / 32] |= 1u << ('c' % 32);
  }
  HostGroup host() const {
    HostGroup h;
    h.num_states = 3;
    h.row = row.data();
    h.col = col.data();
    h.sets = sets.data();
    h.attrs = attrs.data();
    h.num_aas = (int32_t)aas.size();
    h.aas = aas.data();
    h.num_sod = (int32_t)sod.size();
    h.sod = sod.data();
    return h;
  }
};

static void expect_reject(const char *what, const HostGroup &h, int32_t groups = 1,
                          int32_t depth = 3) {
  std::vector<HostGroup> hg(groups > 0 && groups <= 2048 ? groups : 1, h);
  Image img;
  std::string err;
  if (build_image(hg.data(), groups, depth, img, err)) fail(std::string("accepted: ") + what);
  else if (err.empty()) fail(std::string("no message: ") + what);
}

static void check_malformed() {
  Small s;
  {
    Image img;
    std::string err;
    HostGroup h = s.host();
    if (!build_image(&h, 1, 3, img, err)) fail("the valid small group: " + err);
  }
  {
    Small t;
    t.col[1] = 3;
    expect_reject("column index out of range", t.host());
  }
  {
    Small t;
    t.row = {0, 2, 1, 2};
    expect_reject("row offsets not monotone", t.host());
  }
  {
    Small t;
    t.row = {0, 1, 2, -1};
    expect_reject("negative edge count", t.host());
  }
  {
    Small t;
    t.aas = {5};
    expect_reject("all-input id out of range", t.host());
  }
  {
    Small t;
    t.sod = {-1};
    expect_reject("start-of-data id out of range", t.host());
  }
  {
    HostGroup h = s.host();
    h.num_states = -1;
    expect_reject("negative state count", h);
  }
  {
    HostGroup h = s.host();
    h.col = nullptr;
    expect_reject("edges without a column array", h);
  }
  {
    HostGroup h = s.host();
    h.aas = nullptr;
    expect_reject("all-input states without their array", h);
  }
  {
    HostGroup h = s.host();
    h.num_sod = 2;
    h.sod = nullptr;
    expect_reject("start-of-data states without their array", h);
  }
  expect_reject("no groups", s.host(), 0);
  expect_reject("1025 groups", s.host(), 1025);
  expect_reject("precompute depth 4", s.host(), 1, 4);
}

int main(int argc, char **argv) {
  for (int i = 1; i < argc; i++) check_file(argv[i]);
  check_malformed();
  if (!fails) std::printf("PASS\n");
  return fails ? 1 : 0;
}
