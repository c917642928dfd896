// Thread-local last-error message shared by every C ABI entry point.
#include <string>

#include "../../include/ngap_amd.h"

static thread_local std::string g_last_error;
// a message with static storage, used when storing the text failed or when
// the error is that host memory is exhausted (nothing is allocated for it)
static thread_local const char *g_fixed = nullptr;

void ngap_set_last_error_internal(const std::string &m) noexcept {
  try {
    g_last_error = m;
    g_fixed = nullptr;
  } catch (...) {
    g_fixed = "host memory exhausted (error message not stored)";
  }
}

void ngap_set_last_error_fixed(const char *m) noexcept { g_fixed = m; }

extern "C" const char *ngap_last_error(void) { return g_fixed ? g_fixed : g_last_error.c_str(); }
