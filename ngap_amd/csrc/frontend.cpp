// NFA front-end: ANML -> grouped automata in the reference's id namespace.
// Stage-by-stage citations are in frontend.h.
#include "frontend.h"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace ngapamd {

// ---------------------------------------------------------------- Nfa basics
int Nfa::add_node(const Ste &s) {
  node.push_back(s);
  adj.emplace_back();
  from.emplace_back();
  return (int)node.size() - 1;
}

void Nfa::add_edge(int a, int b) {
  // NFA::addEdge (code/src/commons/NFA.cpp:127-135): both lists de-duplicated
  auto &sa = adj[a];
  if (std::find(sa.begin(), sa.end(), b) == sa.end()) sa.push_back(b);
  auto &fb = from[b];
  if (std::find(fb.begin(), fb.end(), a) == fb.end()) fb.push_back(a);
}

void Nfa::remove_node_edges(int x) {
  // NFA::remove_node_unsafe (code/src/commons/NFA.cpp:528-570): drop x from
  // every predecessor's successor list and every successor's predecessor list.
  for (int f : from[x]) {
    auto &l = adj[f];
    auto it = std::find(l.begin(), l.end(), x);
    if (it != l.end()) l.erase(it);
  }
  for (int t : adj[x]) {
    auto &l = from[t];
    auto it = std::find(l.begin(), l.end(), x);
    if (it != l.end()) l.erase(it);
  }
  adj[x].clear();
  from[x].clear();
}

// ------------------------------------------------------ VASim symbol sets
bool parse_symbol_set(const std::string &s, std::bitset<256> &col,
                      std::string &err) {
  col.reset();
  if (s == "*") {  // every byte
    col.set();
    return true;
  }
  if (s == ".") {  // every byte but '\n' (apcompile convention)
    col.set();
    col.reset('\n');
    return true;
  }
  if (!s.empty() && s.front() == '{' && s.back() == '}') {
    err = "CURLY BRACES NOT IMPLEMENTED";
    return false;
  }
  const uint32_t kOpenBracket = 256;
  bool escaped = false, inverting = false, range_set = false;
  int bracket = 0;
  uint32_t last = 0, range_start = 0;
  auto set_range = [&](int a, int b) {
    for (; a <= b; a++) col.set(a);
  };
  // A literal byte: set it, close a pending range, remember it.
  auto lit = [&](uint32_t c) {
    col.set(c);
    if (range_set) {
      set_range((int)range_start, (int)c);
      range_set = false;
    }
    last = c;
  };
  auto at = [&](size_t i) -> char { return i < s.size() ? s[i] : '\0'; };
  size_t i = 0;
  while (i < s.size()) {
    unsigned char c = (unsigned char)s[i];
    switch (c) {
      case '[':
        if (escaped) { lit(c); escaped = false; }
        else { last = kOpenBracket; bracket++; }
        break;
      case ']':
        if (escaped) { lit(c); escaped = false; }
        else bracket--;
        break;
      case '{':
      case '}':
        lit(c);  // braces are always literal; an escape stays pending
        break;
      case '\\':
        if (escaped) { lit(c); escaped = false; }
        else escaped = true;
        break;
      case 'n': case 'r': case 't': case 'a': case 'b': case 'f': case 'v': {
        if (escaped) {
          static const char from_c[] = "nrtabfv";
          static const char to_c[] = "\n\r\t\a\b\f\v";
          lit((unsigned char)to_c[strchr(from_c, c) - from_c]);
          escaped = false;
        } else {
          lit(c);
        }
        break;
      }
      case '\'':
      case '"':
        lit(c);
        escaped = false;
        break;
      case '-':
        if (escaped || last == kOpenBracket) {
          lit('-');
          escaped = false;
        } else {
          range_set = true;
          range_start = last;
        }
        break;
      case 's':
        if (escaped) {  // whitespace class; leaves last/range untouched
          for (int w : {0x0A, 0x09, 0x0D, 0x0B, 0x0C, 0x20}) col.set(w);
          escaped = false;
        } else {
          lit(c);
        }
        break;
      case 'd':
        if (escaped) { set_range('0', '9'); escaped = false; }
        else lit(c);
        break;
      case 'w':
        if (escaped) {
          col.set('_');
          set_range('0', '9');
          set_range('A', 'Z');
          set_range('a', 'z');
          escaped = false;
        } else {
          lit(c);
        }
        break;
      case '^':
        if (escaped) { lit(c); escaped = false; }
        else inverting = true;
        break;
      case 'x':
        if (escaped) {
          ++i;
          char hex[3] = {at(i), at(i + 1), '\0'};
          unsigned char num = (unsigned char)strtoul(hex, nullptr, 16);
          ++i;
          lit(num);
          escaped = false;
        } else {
          lit(c);
        }
        break;
      default:
        escaped = false;  // an unknown escape is an ordinary byte
        lit(c);
    }
    i++;
  }
  if (inverting) col.flip();
  if (bracket != 0) {
    err = "MALFORMED BRACKETS OR BRACES: " + s;
    return false;
  }
  return true;
}

// ------------------------------------------------------------ minimal XML
namespace {

struct XNode {
  std::string name;  // empty for text / CDATA
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<XNode> kids;
  const std::string *attr(const char *k) const {
    for (auto &a : attrs)
      if (a.first == k) return &a.second;
    return nullptr;
  }
};

class XmlParser {
 public:
  explicit XmlParser(const std::string &t) : t_(t) {}
  bool parse(XNode &doc, std::string &err) {
    if (t_.compare(0, 3, "\xEF\xBB\xBF") == 0) p_ = 3;
    if (!content(doc, "")) { err = err_; return false; }
    return true;
  }

 private:
  const std::string &t_;
  size_t p_ = 0;
  std::string err_;

  bool fail(const std::string &m) {
    std::ostringstream o;
    o << "XML parse error at byte " << p_ << ": " << m;
    err_ = o.str();
    return false;
  }
  static bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
  static bool is_name(char c) {
    return isalnum((unsigned char)c) || c == '_' || c == '-' || c == ':' || c == '.' ||
           (unsigned char)c >= 0x80;
  }
  // Entity decoding plus attribute whitespace conversion (pugixml's default
  // parse flags: escapes, wconv_attribute, eol).
  static void decode(const std::string &in, std::string &out, bool attr) {
    out.clear();
    for (size_t i = 0; i < in.size(); i++) {
      char c = in[i];
      if (c == '&') {
        size_t e = in.find(';', i);
        if (e != std::string::npos) {
          std::string ent = in.substr(i + 1, e - i - 1);
          long cp = -1;
          if (ent == "lt") cp = '<';
          else if (ent == "gt") cp = '>';
          else if (ent == "amp") cp = '&';
          else if (ent == "quot") cp = '"';
          else if (ent == "apos") cp = '\'';
          else if (ent.size() > 1 && ent[0] == '#') {
            char *endp = nullptr;
            cp = (ent[1] == 'x') ? strtol(ent.c_str() + 2, &endp, 16)
                                 : strtol(ent.c_str() + 1, &endp, 10);
            if (!endp || *endp) cp = -1;
          }
          if (cp >= 0) {
            if (cp < 0x80) out.push_back((char)cp);
            else if (cp < 0x800) {
              out.push_back((char)(0xC0 | (cp >> 6)));
              out.push_back((char)(0x80 | (cp & 0x3F)));
            } else if (cp < 0x10000) {
              out.push_back((char)(0xE0 | (cp >> 12)));
              out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
              out.push_back((char)(0x80 | (cp & 0x3F)));
            } else {
              out.push_back((char)(0xF0 | (cp >> 18)));
              out.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
              out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
              out.push_back((char)(0x80 | (cp & 0x3F)));
            }
            i = e;
            continue;
          }
        }
        out.push_back(c);
      } else if (attr && c == '\r' && i + 1 < in.size() && in[i + 1] == '\n') {
        out.push_back(' ');
        i++;
      } else if (attr && (c == '\r' || c == '\n' || c == '\t')) {
        out.push_back(' ');
      } else {
        out.push_back(c);
      }
    }
  }
  bool starts(const char *s) const { return t_.compare(p_, strlen(s), s) == 0; }
  bool skip_until(const char *s) {
    size_t e = t_.find(s, p_);
    if (e == std::string::npos) return fail(std::string("missing ") + s);
    p_ = e + strlen(s);
    return true;
  }
  // Parses children until </close> (or end of input when close is empty).
  bool content(XNode &parent, const std::string &close) {
    while (p_ < t_.size()) {
      if (t_[p_] != '<') {
        size_t e = t_.find('<', p_);
        if (e == std::string::npos) e = t_.size();
        std::string raw = t_.substr(p_, e - p_);
        p_ = e;
        bool ws = std::all_of(raw.begin(), raw.end(), is_ws);
        if (!ws && !close.empty()) {
          XNode txt;
          decode(raw, txt.name, false);
          txt.name.clear();  // text nodes carry no element name
          parent.kids.push_back(std::move(txt));
        }
        continue;
      }
      if (starts("<?")) { if (!skip_until("?>")) return false; continue; }
      if (starts("<!--")) { if (!skip_until("-->")) return false; continue; }
      if (starts("<![CDATA[")) {
        if (!skip_until("]]>")) return false;
        if (!close.empty()) parent.kids.emplace_back();
        continue;
      }
      if (starts("<!")) { if (!skip_until(">")) return false; continue; }
      if (starts("</")) {
        p_ += 2;
        size_t s = p_;
        while (p_ < t_.size() && is_name(t_[p_])) p_++;
        std::string nm = t_.substr(s, p_ - s);
        while (p_ < t_.size() && is_ws(t_[p_])) p_++;
        if (p_ >= t_.size() || t_[p_] != '>') return fail("bad end tag");
        p_++;
        if (nm != close) return fail("mismatched end tag </" + nm + ">");
        return true;
      }
      // element
      p_++;
      XNode el;
      size_t s = p_;
      while (p_ < t_.size() && is_name(t_[p_])) p_++;
      el.name = t_.substr(s, p_ - s);
      if (el.name.empty()) return fail("bad element name");
      for (;;) {
        while (p_ < t_.size() && is_ws(t_[p_])) p_++;
        if (p_ >= t_.size()) return fail("unterminated tag");
        if (t_[p_] == '/') {
          if (p_ + 1 >= t_.size() || t_[p_ + 1] != '>') return fail("bad '/'");
          p_ += 2;
          parent.kids.push_back(std::move(el));
          break;
        }
        if (t_[p_] == '>') {
          p_++;
          if (!content(el, el.name)) return false;
          parent.kids.push_back(std::move(el));
          break;
        }
        size_t a = p_;
        while (p_ < t_.size() && is_name(t_[p_])) p_++;
        std::string an = t_.substr(a, p_ - a);
        if (an.empty()) return fail("bad attribute");
        while (p_ < t_.size() && is_ws(t_[p_])) p_++;
        if (p_ >= t_.size() || t_[p_] != '=') return fail("attribute without '='");
        p_++;
        while (p_ < t_.size() && is_ws(t_[p_])) p_++;
        if (p_ >= t_.size() || (t_[p_] != '"' && t_[p_] != '\'')) return fail("unquoted attribute");
        char q = t_[p_++];
        size_t e = t_.find(q, p_);
        if (e == std::string::npos) return fail("unterminated attribute");
        std::string v;
        decode(t_.substr(p_, e - p_), v, true);
        p_ = e + 1;
        el.attrs.emplace_back(std::move(an), std::move(v));
      }
    }
    if (!close.empty()) return fail("missing </" + close + ">");
    return true;
  }
};

const XNode *child(const XNode &n, const char *name) {
  for (auto &k : n.kids)
    if (k.name == name) return &k;
  return nullptr;
}

}  // namespace

// ------------------------------------------------------------ ANML loading
bool load_anml_text(const std::string &text, Nfa &nfa, std::string &err) {
  XNode doc;
  XmlParser xp(text);
  if (!xp.parse(doc, err)) return false;
  // automata-network under <anml>, or at the top level
  // (code/src/commons/NFALoader.cpp:85-98)
  const XNode *net = nullptr;
  if (const XNode *anml = child(doc, "anml")) net = child(*anml, "automata-network");
  if (!net) net = child(doc, "automata-network");
  nfa = Nfa();
  if (!net) return true;
  std::unordered_map<std::string, int> by_name;
  std::vector<std::pair<int, std::string>> edges;  // (from id, to name)
  for (const XNode &el : net->kids) {
    if (el.name == "state-transition-element") {
      Ste s;
      if (const std::string *v = el.attr("id")) s.name = *v;
      std::string sym, start;
      if (const std::string *v = el.attr("symbol-set")) sym = *v;
      if (const std::string *v = el.attr("start")) start = *v;
      s.start = start == "start-of-data" ? kStartOfData
                : start == "all-input"   ? kStartAllInput
                                         : kStartNone;
      if (!parse_symbol_set(sym, s.cls, err)) return false;
      if (by_name.count(s.name)) {
        err = "duplicate state id " + s.name;
        return false;
      }
      for (const XNode &k : el.kids) {
        if (k.name == "report-on-match") {
          s.report = true;
          if (const std::string *rc = k.attr("reportcode")) s.report_code = *rc;
          else s.report_code.clear();
        }
      }
      int id = nfa.add_node(s);
      by_name[s.name] = id;
      for (const XNode &k : el.kids)
        if (k.name == "activate-on-match") {
          const std::string *to = k.attr("element");
          edges.emplace_back(id, to ? *to : std::string());
        }
    } else if (el.name == "description") {
      // ignored
    } else {
      err = "NODE: " + el.name + " NOT IMPLEMENTED IN PARSER...";
      return false;
    }
  }
  // Edges are applied in document order, as NFA::addEdge is called while
  // reading, so successor and predecessor lists keep the reference order.
  for (auto &e : edges) {
    auto it = by_name.find(e.second);
    if (it == by_name.end()) {
      err = "activate-on-match names unknown element " + e.second;
      return false;
    }
    nfa.add_edge(e.first, it->second);
  }
  return true;
}

bool load_anml_file(const std::string &path, Nfa &nfa, std::string &err) {
  // load_nfa_from_file accepts only the .anml suffix
  // (code/src/commons/NFALoader.cpp:158-171)
  if (path.size() < 5 || path.compare(path.size() - 5, 5, ".anml") != 0) {
    err = "Unsupported NFA file type: " + path;
    return false;
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "Could not load .xml file: " + path;
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return load_anml_text(ss.str(), nfa, err);
}

// ------------------------------------------------------ CC split and limit
std::vector<Nfa> split_by_cc(const Nfa &nfa) {
  // Components of the undirected graph, numbered by their smallest id; local
  // ids follow global id order (NFA::mark_cc_id, split_nfa_by_ccs).
  int V = nfa.size();
  std::vector<int> cc(V, -1), local(V, 0);
  std::vector<std::vector<int>> und(V);
  for (int i = 0; i < V; i++)
    for (int j : nfa.adj[i]) {
      und[i].push_back(j);
      und[j].push_back(i);
    }
  int ncc = 0;
  std::vector<int> stack;
  for (int i = 0; i < V; i++) {
    if (cc[i] >= 0) continue;
    cc[i] = ncc;
    stack.assign(1, i);
    while (!stack.empty()) {
      int x = stack.back();
      stack.pop_back();
      for (int y : und[x])
        if (cc[y] < 0) {
          cc[y] = ncc;
          stack.push_back(y);
        }
    }
    ncc++;
  }
  std::vector<Nfa> out(ncc);
  for (int i = 0; i < V; i++) local[i] = out[cc[i]].add_node(nfa.node[i]);
  for (int i = 0; i < V; i++)
    for (int j : nfa.adj[i]) out[cc[i]].add_edge(local[i], local[j]);
  return out;
}

bool limit_out_degree(Nfa &g, int limit) {
  // nfa_utils::limit_out_degree_on_nfa (code/src/commons/nfa_utils.cpp:507-634).
  // Each pass splits one node with out-degree > limit: scanning ids from high
  // to low, the first with the smallest out-degree (the reference's
  // get_indegree_of_node returns out-degree, NFA.cpp:505-508).  Copy 0 keeps
  // the id, further copies are appended; every copy gets the node's
  // predecessors and up to `limit` successors (limit-1 plus itself when the
  // node loops on itself).
  int iteration = 0;
  const int original_size = g.size();
  for (;;) {
    iteration++;
    int best = -1;
    size_t best_deg = SIZE_MAX;
    for (int i = g.size() - 1; i >= 0; i--) {
      size_t d = g.adj[i].size();
      if (d > (size_t)limit && d < best_deg) {
        best_deg = d;
        best = i;
      }
    }
    if (best < 0) break;
    std::vector<int> from_nodes = g.from[best];
    std::vector<int> to_nodes = g.adj[best];
    Ste deleted = g.node[best];
    g.remove_node_edges(best);
    auto s1 = std::find(to_nodes.begin(), to_nodes.end(), best);
    auto s2 = std::find(from_nodes.begin(), from_nodes.end(), best);
    bool self_loop = s1 != to_nodes.end();
    if (self_loop && s2 != from_nodes.end()) {
      to_nodes.erase(s1);
      from_nodes.erase(s2);
    }
    size_t t = 0;
    int p = 0;
    while (t < to_nodes.size()) {
      Ste copy = deleted;
      copy.name += "_dup_" + std::to_string(p);
      int id;
      if (p == 0) {
        id = best;
        g.node[id] = copy;
      } else {
        id = g.add_node(copy);
      }
      int L = limit;
      if (self_loop) {
        L = limit - 1;
        g.add_edge(id, id);
      }
      for (int kk = 0; kk < L && t + kk < to_nodes.size(); kk++)
        g.add_edge(id, to_nodes[t + kk]);
      for (int f : from_nodes) g.add_edge(f, id);
      p++;
      t += L;
    }
    if (iteration >= original_size * 2) return false;
  }
  return true;
}

std::vector<std::vector<int>> group_ccs_by_num(int &group_num, int cc_num) {
  // nfa_utils::group_nfas_by_num (code/src/commons/nfa_utils.cpp:382-424):
  // contiguous CC ranges cut against a single-precision running bound.
  std::vector<std::vector<int>> res;
  if (group_num > cc_num) group_num = cc_num;
  float per = (float)(1.0 * cc_num / group_num);
  float bound = per;
  std::vector<int> cur;
  for (int i = 0; i < cc_num; i++) {
    if (!cur.empty() && (float)(i + 1) > bound && (int)res.size() + 1 < group_num) {
      bound += per;
      res.push_back(cur);
      cur.clear();
    }
    cur.push_back(i);
  }
  if (!cur.empty()) res.push_back(cur);
  return res;
}

GroupGraph to_group_graph(const Nfa &g) {
  // Graph::ReadNFA + Csr::fromCoo: edges sorted by (src, dst).
  GroupGraph gg;
  int n = g.size();
  gg.num_states = n;
  gg.sets.assign((size_t)n * 8, 0u);
  gg.attrs.assign(n, 0);
  gg.names.resize(n);
  std::vector<std::pair<int, int>> e;
  for (int i = 0; i < n; i++) {
    const Ste &s = g.node[i];
    gg.names[i] = s.name;
    for (int b = 0; b < 256; b++)
      if (s.cls.test(b)) gg.sets[(size_t)i * 8 + b / 32] |= 1u << (b % 32);
    if (s.start == kStartAllInput) {
      gg.attrs[i] |= 1;
      gg.aas.push_back(i);
    }
    if (s.start == kStartOfData) gg.sod.push_back(i);
    if (s.report) gg.attrs[i] |= 2;
    for (int j : g.adj[i]) e.emplace_back(i, j);
  }
  std::sort(e.begin(), e.end());
  gg.row.assign(n + 1, 0);
  gg.col.resize(e.size());
  for (size_t k = 0; k < e.size(); k++) {
    gg.row[e[k].first + 1]++;
    gg.col[k] = e[k].second;
  }
  for (int i = 0; i < n; i++) gg.row[i + 1] += gg.row[i];
  return gg;
}

bool build_groups(const Nfa &nfa, FrontendOptions &opt,
                  std::vector<GroupGraph> &groups, FrontendStats &st,
                  std::string &err) {
  st = FrontendStats();
  st.total_states = nfa.size();
  for (auto &s : nfa.node) {
    if (s.start == kStartOfData) st.start_states++;
    if (s.start == kStartAllInput) st.always_active++;
    if (s.report) st.report_states++;
  }
  if (nfa.size() <= 1) {
    // mark_cc_id asserts V > 0 and prints for V <= 1
    if (nfa.size() == 0) {
      err = "empty automaton";
      return false;
    }
  }
  std::vector<Nfa> ccs = split_by_cc(nfa);
  st.num_cc = (int)ccs.size();
  long long sum = 0;
  for (auto &c : ccs) {
    sum += c.size();
    st.max_cc = std::max(st.max_cc, c.size());
    if (c.size() > 256) st.cc_256++;
  }
  st.avg_cc = ccs.empty() ? 0.0 : (double)sum / ccs.size();
  // ngap::preprocessing (code/src/ngap/ngap.cu:63-90)
  std::vector<Nfa> kept;
  for (size_t i = 0; i < ccs.size(); i++) {
    if (limit_out_degree(ccs[i], 4)) {
      kept.push_back(std::move(ccs[i]));
      st.limit_succeed++;
    } else {
      st.limit_fail++;
      if (opt.remove_degree && opt.quit_degree) {
        err = "ERROR: Quit when degree > 4";
        return false;
      }
    }
  }
  if (opt.max_nfa_size != -1) {
    std::vector<Nfa> tmp;
    for (auto &c : kept) {
      if (c.size() <= opt.max_nfa_size) tmp.push_back(std::move(c));
      else st.removed_by_size++;
    }
    kept.swap(tmp);
  }
  if (kept.empty()) {
    err = "no connected component left after preprocessing";
    return false;
  }
  if (opt.group_num <= 0) {
    err = "--group-num must be positive";
    return false;
  }
  auto parts = group_ccs_by_num(opt.group_num, (int)kept.size());
  groups.clear();
  for (auto &part : parts) {
    Nfa merged;
    for (int ci : part) {
      const Nfa &c = kept[ci];
      int base = merged.size();
      for (int i = 0; i < c.size(); i++) merged.add_node(c.node[i]);
      for (int i = 0; i < c.size(); i++)
        for (int j : c.adj[i]) merged.add_edge(base + i, base + j);
    }
    groups.push_back(to_group_graph(merged));
  }
  return true;
}

std::string dump_groups(const std::vector<GroupGraph> &groups) {
  std::ostringstream o;
  o << "groups " << groups.size() << "\n";
  char buf[128];
  for (size_t gi = 0; gi < groups.size(); gi++) {
    const GroupGraph &g = groups[gi];
    o << "group " << gi << " states " << g.num_states << " edges " << g.col.size()
      << " aas " << g.aas.size() << " sod " << g.sod.size() << "\n";
    for (int i = 0; i < g.num_states; i++) {
      o << i << " " << g.names[i] << " a" << (int)g.attrs[i] << " s";
      for (int w = 7; w >= 0; w--) {
        snprintf(buf, sizeof buf, "%08x", g.sets[(size_t)i * 8 + w]);
        o << buf;
      }
      o << " ->";
      for (int e = g.row[i]; e < g.row[i + 1]; e++) o << " " << g.col[e];
      o << "\n";
    }
  }
  return o.str();
}

}  // namespace ngapamd
