// The body of every C entry point runs under c_guard: a C++ exception thrown
// inside (host memory exhausted, a thread the system refused) becomes
// NGAP_ERR_HOST (or the entry point's own failure value, c_guard_or) and a
// message for ngap_last_error(), so nothing unwinds into a C caller.  The
// handlers allocate nothing they cannot do without: host memory exhausted is
// reported through a message with static storage.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/ngap_amd.h"

void ngap_set_last_error_internal(const std::string &m) noexcept;  // errors.cpp
void ngap_set_last_error_fixed(const char *m) noexcept;            // errors.cpp

namespace ngapamd {

template <class F, class R>
R c_guard_or(F &&body, R fail) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    ngap_set_last_error_fixed("host memory exhausted");
  } catch (const std::exception &e) {
    try {
      ngap_set_last_error_internal(std::string("host error: ") + e.what());
    } catch (...) {
      ngap_set_last_error_fixed("host error");
    }
  } catch (...) {
    ngap_set_last_error_fixed("host error");
  }
  return fail;
}

template <class F>
int c_guard(F &&body) noexcept {
  return c_guard_or(static_cast<F &&>(body), (int)NGAP_ERR_HOST);
}

}  // namespace ngapamd
