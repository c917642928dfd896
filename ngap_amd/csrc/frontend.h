// NFA front-end: ANML text -> grouped automata in the reference's state-id
// namespace.  Every report carries a state id produced by this pipeline, so the
// steps below follow the reference's own order of operations exactly:
//   ANML document order            code/src/commons/NFALoader.cpp:70-154
//   symbol-set parsing (VASim)     code/src/commons/vasim_helper.cpp:40-478
//   connected components           code/src/commons/NFA.cpp:188-226
//   split into one NFA per CC      code/src/commons/nfa_utils.cpp:54-90
//   out-degree limit 4 (dup split) code/src/commons/nfa_utils.cpp:507-676
//   optional max CC size filter    code/src/ngap/ngap.cu:77-90
//   grouping by CC count (float)   code/src/commons/nfa_utils.cpp:382-424
//   merge + Graph::ReadNFA + CSR   code/src/commons/nfa_utils.cpp:466-486,
//                                  code/src/commons/graph.cu:169-206,
//                                  code/include/commons/graph.h:184-219
#pragma once

#include <bitset>
#include <cstdint>
#include <string>
#include <vector>

namespace ngapamd {

// start attribute values (code/src/commons/NFALoader.cpp:23-34)
enum StartKind : int { kStartNone = 0, kStartOfData = 1, kStartAllInput = 2 };

struct Ste {
  std::string name;          // ANML id, plus "_dup_<p>" suffixes after splits
  std::bitset<256> cls;      // symbol class
  int start = kStartNone;
  bool report = false;
  std::string report_code;   // parsed, never used for report ids
};

// A mutable automaton whose integer ids behave like the reference NFA class:
// ids are dense [0, V); a split node's first copy reuses its id and further
// copies are appended.  adj/from keep insertion order (it decides split order).
struct Nfa {
  std::vector<Ste> node;
  std::vector<std::vector<int>> adj;   // successors, insertion order
  std::vector<std::vector<int>> from;  // predecessors, insertion order
  int size() const { return (int)node.size(); }
  int add_node(const Ste &s);          // id = size()
  void add_edge(int a, int b);         // de-duplicated, appends on both sides
  void remove_node_edges(int x);       // detach x from all neighbours
};

// One group in the form the kernels consume (Graph + Csr of the reference).
struct GroupGraph {
  int32_t num_states = 0;
  std::vector<int32_t> row;        // [n+1] CSR offsets, edges sorted (src,dst)
  std::vector<int32_t> col;        // [E]
  std::vector<uint32_t> sets;      // [n*8] 256-bit class, bit b in word b/32
  std::vector<uint8_t> attrs;      // bit0 all-input, bit1 reporting
  std::vector<int32_t> aas;        // all-input states, ascending id
  std::vector<int32_t> sod;        // start-of-data (not all-input), ascending
  std::vector<std::string> names;  // state names (for dumps / debugging)
};

struct FrontendOptions {
  int group_num = 10;        // --group-num; lowered to the CC count if larger
  int max_nfa_size = -1;     // --max-nfa-size (CC filter), -1 = off
  bool quit_degree = true;   // --quit-degree
  bool remove_degree = false;// --remove-degree
};

struct FrontendStats {
  int total_states = 0, start_states = 0, always_active = 0, report_states = 0;
  int num_cc = 0, limit_succeed = 0, limit_fail = 0, removed_by_size = 0;
  int cc_256 = 0, max_cc = 0;   // components > 256 states, largest component
  double avg_cc = 0;
};

// Parses a VASim symbol-set string.  Returns false (and fills err) where the
// reference exits (curly-brace sets, unbalanced brackets).
bool parse_symbol_set(const std::string &s, std::bitset<256> &out,
                      std::string &err);

// Loads ANML text into an Nfa in document order.  false + err on failure.
bool load_anml_text(const std::string &text, Nfa &out, std::string &err);
bool load_anml_file(const std::string &path, Nfa &out, std::string &err);

// Whole pipeline: Nfa -> groups.  `opt.group_num` may be lowered (as the
// reference does by reference argument).
bool build_groups(const Nfa &nfa, FrontendOptions &opt,
                  std::vector<GroupGraph> &groups, FrontendStats &stats,
                  std::string &err);

// Individual stages, exposed for tests.
std::vector<Nfa> split_by_cc(const Nfa &nfa);
bool limit_out_degree(Nfa &cc, int limit);
std::vector<std::vector<int>> group_ccs_by_num(int &group_num, int cc_num);
GroupGraph to_group_graph(const Nfa &merged);

// Text dump of a group list; the same format is written by the dumper built
// against the reference front-end, so the two can be compared byte for byte.
std::string dump_groups(const std::vector<GroupGraph> &groups);

}  // namespace ngapamd
