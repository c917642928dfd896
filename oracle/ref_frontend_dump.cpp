// Test infrastructure (never shipped, never measured): a small driver that runs
// the reference's own CUDA-free NFA front-end, compiled from the sources under
// the reference tree by oracle/Makefile, and prints the grouped automata in
// the text format of ngapamd::dump_groups.  It pins the product front-end's
// state-id namespace against the reference implementation.
//
// Steps follow code/src/ngap/main.cu:56-78 and ngap::preprocessing
// (code/src/ngap/ngap.cu:63-90).  calc_scc/topo_sort are skipped: they only set
// scc_id/topo_order/bfs_layer, which nothing on the id path reads, and
// topo_sort is quadratic in the component size.  Graph::ReadNFA and
// Csr::fromCoo live in CUDA files, so their two loops (node attributes in id
// order; edges (i, sid(adj)) sorted by (src, dst)) are written out below.
//
// usage: ref_frontend_dump <file.anml> <group_num> [max_nfa_size]
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "NFA.h"
#include "NFALoader.h"
#include "nfa_utils.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s file.anml group_num [max_nfa_size]\n", argv[0]);
    return 2;
  }
  int group_num = atoi(argv[2]);
  int max_cc = argc > 3 ? atoi(argv[3]) : -1;
  NFA *nfa = load_nfa_from_file(argv[1]);
  nfa->mark_cc_id();
  std::vector<NFA *> ccs = nfa_utils::split_nfa_by_ccs(*nfa);
  nfa_utils::limit_out_degree_on_ccs(ccs, 4, true, false);
  if (max_cc != -1) {
    std::vector<NFA *> keep;
    for (auto *c : ccs)
      if (c->size() <= max_cc) keep.push_back(c);
    ccs = keep;
  }
  auto groups = nfa_utils::group_nfas_by_num(group_num, ccs);
  printf("groups %zu\n", groups.size());
  for (size_t gi = 0; gi < groups.size(); gi++) {
    NFA *g = groups[gi];
    int n = g->size();
    std::vector<std::pair<int, int>> e;
    int naas = 0, nsod = 0;
    for (int i = 0; i < n; i++) {
      Node *node = g->get_node_by_int_id(i);
      if (node->is_start_always_enabled()) naas++;
      if (node->is_start() && !node->is_start_always_enabled()) nsod++;
      for (auto &to : g->adj[node->str_id])
        e.emplace_back(i, g->get_node_by_str_id(to)->sid);
    }
    std::sort(e.begin(), e.end());
    printf("group %zu states %d edges %zu aas %d sod %d\n", gi, n, e.size(), naas, nsod);
    size_t k = 0;
    for (int i = 0; i < n; i++) {
      Node *node = g->get_node_by_int_id(i);
      int attrs = (node->is_start_always_enabled() ? 1 : 0) | (node->is_report() ? 2 : 0);
      unsigned w[8] = {0};
      for (int b = 0; b < 256; b++)
        if (node->symbol_set.test(b)) w[b / 32] |= 1u << (b % 32);
      printf("%d %s a%d s", i, node->str_id.c_str(), attrs);
      for (int j = 7; j >= 0; j--) printf("%08x", w[j]);
      printf(" ->");
      while (k < e.size() && e[k].first == i) printf(" %d", e[k++].second);
      printf("\n");
    }
  }
  return 0;
}
