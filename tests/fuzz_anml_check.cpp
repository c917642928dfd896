// Seeded mutations of ANML text and VASim symbol sets through the front-end
// (and, when a mutant still parses, the grouping and the host image builder)
// under AddressSanitizer and UBSan on the CPU: malformed input must come back
// as an error, never as a crash or undefined behaviour.  argv: iterations,
// then ANML files.  Prints PASS and the count of mutants that still parsed.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../ngap_amd/csrc/frontend.h"
#include "../ngap_amd/csrc/layout.h"

using namespace ngapamd;

static std::string mutate(std::string s, std::mt19937 &rng) {
  static const char kChars[] = "<>\"=/[]\\^-*{}()|.x0123456789 \n&;#abc\xff";
  const int edits = 1 + rng() % 4;
  for (int k = 0; k < edits && !s.empty(); k++) {
    const size_t at = rng() % s.size();
    switch (rng() % 5) {
      case 0: s[at] = kChars[rng() % (sizeof kChars - 1)]; break;       // replace
      case 1: s.erase(at, 1 + rng() % 8); break;                        // delete
      case 2: s.insert(at, 1, kChars[rng() % (sizeof kChars - 1)]); break;  // insert
      case 3: s.resize(at); break;                                      // truncate
      default: {                                                        // duplicate a piece
        const size_t len = std::min<size_t>(s.size() - at, 1 + rng() % 64);
        s.insert(rng() % s.size(), s.substr(at, len));
      }
    }
  }
  return s;
}

int main(int argc, char **argv) {
  const int iters = argc > 1 ? std::atoi(argv[1]) : 1000;
  std::mt19937 rng(20260);
  int parsed = 0, grouped = 0;
  for (int f = 2; f < argc; f++) {
    std::ifstream in(argv[f], std::ios::binary);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    for (int i = 0; i < iters; i++) {
      const std::string m = mutate(text, rng);
      Nfa nfa;
      std::string err;
      if (!load_anml_text(m, nfa, err)) continue;
      parsed++;
      FrontendOptions opt;
      opt.group_num = 1 + rng() % 8;
      std::vector<GroupGraph> groups;
      FrontendStats st;
      if (!build_groups(nfa, opt, groups, st, err) || groups.empty()) continue;
      grouped++;
      std::vector<HostGroup> hg;
      for (const GroupGraph &g : groups) {
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
        hg.push_back(h);
      }
      Image img;
      (void)build_image(hg.data(), (int32_t)hg.size(), 2, img, err);
    }
  }
  // symbol sets: mutations of well-formed ones
  const char *sets[] = {"[abc]", "[^\\x00-\\x1f]", "*", "[\\x41-\\x5a\\x61-\\x7a]", "a",
                        "[\\[\\]\\\\]", "[a-z0-9_]", "\\x2e", "[^a]"};
  for (int i = 0; i < iters * 4; i++) {
    const std::string m = mutate(sets[rng() % (sizeof sets / sizeof sets[0])], rng);
    std::bitset<256> b;
    std::string err;
    (void)parse_symbol_set(m, b, err);
  }
  std::printf("PASS parsed %d grouped %d\n", parsed, grouped);
  return 0;
}
