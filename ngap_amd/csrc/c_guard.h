// The body of every status-returning C entry point runs under c_guard: a C++
// exception thrown inside (host memory exhausted, a thread the system
// refused) becomes NGAP_ERR_HOST and a message for ngap_last_error(), so
// nothing unwinds into a C caller.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "../../include/ngap_amd.h"

void ngap_set_last_error_internal(const std::string &m);  // errors.cpp

namespace ngapamd {

template <class F>
int c_guard(F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    ngap_set_last_error_internal("host memory exhausted");
  } catch (const std::exception &e) {
    ngap_set_last_error_internal(std::string("host error: ") + e.what());
  } catch (...) {
    ngap_set_last_error_internal("host error");
  }
  return NGAP_ERR_HOST;
}

}  // namespace ngapamd
