// Thread-local last-error message shared by every C ABI entry point.
#include <string>

#include "../../include/ngap_amd.h"

static thread_local std::string g_last_error;

void ngap_set_last_error_internal(const std::string &m) { g_last_error = m; }

extern "C" const char *ngap_last_error(void) { return g_last_error.c_str(); }
