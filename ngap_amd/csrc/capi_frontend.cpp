// C ABI of the front-end (ANML -> groups).  Host only; no device calls.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ngap_amd.h"
#include "c_guard.h"
#include "frontend.h"

using namespace ngapamd;

struct ngap_frontend {
  std::vector<GroupGraph> groups;
  FrontendStats stats;
  std::string dump;
};

void ngap_set_last_error_internal(const std::string &m) noexcept;  // errors.cpp

namespace {
int fe_fail(const std::string &m) {
  ngap_set_last_error_internal(m);
  return NGAP_ERR_INVALID;
}

int finish(Nfa &nfa, int32_t group_num, int32_t max_nfa_size, int32_t quit, int32_t remove,
           ngap_frontend **out) {
  std::unique_ptr<ngap_frontend> fe(new ngap_frontend());
  FrontendOptions opt;
  opt.group_num = group_num;
  opt.max_nfa_size = max_nfa_size;
  opt.quit_degree = quit != 0;
  opt.remove_degree = remove != 0;
  std::string err;
  if (!build_groups(nfa, opt, fe->groups, fe->stats, err)) return fe_fail(err);
  *out = fe.release();
  return NGAP_OK;
}
}  // namespace

extern "C" {

int ngap_frontend_load(const char *path, int32_t group_num, int32_t max_nfa_size,
                       int32_t quit_degree, int32_t remove_degree, ngap_frontend **out) {
  return ngapamd::c_guard([&]() -> int {
    if (!path || !out) return fe_fail("NULL argument");
    *out = nullptr;
    Nfa nfa;
    std::string err;
    if (!load_anml_file(path, nfa, err)) return fe_fail(err);
    return finish(nfa, group_num, max_nfa_size, quit_degree, remove_degree, out);
  });
}

int ngap_frontend_load_text(const char *text, int32_t group_num, int32_t max_nfa_size,
                            ngap_frontend **out) {
  return ngapamd::c_guard([&]() -> int {
    if (!text || !out) return fe_fail("NULL argument");
    *out = nullptr;
    Nfa nfa;
    std::string err;
    if (!load_anml_text(text, nfa, err)) return fe_fail(err);
    return finish(nfa, group_num, max_nfa_size, 1, 0, out);
  });
}

int ngap_frontend_free(ngap_frontend *fe) {
  return ngapamd::c_guard([&]() -> int {
    delete fe;
    return NGAP_OK;
  });
}

int32_t ngap_frontend_num_groups(const ngap_frontend *fe) {
  return fe ? (int32_t)fe->groups.size() : 0;
}

int ngap_frontend_group(const ngap_frontend *fe, int32_t g, ngap_group_desc *d) {
  return ngapamd::c_guard([&]() -> int {
    if (!fe || !d || g < 0 || g >= (int32_t)fe->groups.size()) return fe_fail("bad group index");
    const GroupGraph &gg = fe->groups[g];
    d->num_states = gg.num_states;
    d->row_offsets = gg.row.data();
    d->col_indices = gg.col.data();
    d->symbol_sets = gg.sets.data();
    d->node_attrs = gg.attrs.data();
    d->num_always_active = (int32_t)gg.aas.size();
    d->always_active = gg.aas.data();
    d->num_start = (int32_t)gg.sod.size();
    d->start_states = gg.sod.data();
    return NGAP_OK;
  });
}

const char *ngap_frontend_state_name(const ngap_frontend *fe, int32_t g, int32_t s) {
  if (!fe || g < 0 || g >= (int32_t)fe->groups.size()) return nullptr;
  const GroupGraph &gg = fe->groups[g];
  if (s < 0 || s >= gg.num_states) return nullptr;
  return gg.names[s].c_str();
}

const char *ngap_frontend_dump(ngap_frontend *fe) {
  return ngapamd::c_guard_or([&]() -> const char * {
    if (!fe) return nullptr;
    fe->dump = dump_groups(fe->groups);
    return fe->dump.c_str();
  }, (const char *)nullptr);
}

int ngap_parse_symbol_set(const char *s, uint32_t words[8]) {
  return ngapamd::c_guard([&]() -> int {
    if (!s || !words) return fe_fail("NULL argument");
    std::bitset<256> b;
    std::string err;
    if (!parse_symbol_set(s, b, err)) return fe_fail(err);
    for (int w = 0; w < 8; w++) words[w] = 0;
    for (int i = 0; i < 256; i++)
      if (b.test(i)) words[i / 32] |= 1u << (i % 32);
    return NGAP_OK;
  });
}

}  // extern "C"
