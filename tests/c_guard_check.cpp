// c_guard (ngap_amd/csrc/c_guard.h) on the CPU: a body's status passes
// through; an exception thrown inside becomes NGAP_ERR_HOST with a message
// for ngap_last_error().  Prints PASS or the first failure.
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <system_error>

#include "../ngap_amd/csrc/c_guard.h"

extern "C" const char *ngap_last_error(void);

static int fails = 0;
static void expect(bool ok, const char *what) {
  if (!ok) {
    std::printf("FAIL %s (last error: %s)\n", what, ngap_last_error());
    fails++;
  }
}

int main() {
  using ngapamd::c_guard;
  expect(c_guard([] { return NGAP_OK; }) == NGAP_OK, "ok passes");
  expect(c_guard([] { return NGAP_ERR_OVERFLOW; }) == NGAP_ERR_OVERFLOW, "status passes");
  expect(c_guard([]() -> int { throw std::bad_alloc(); }) == NGAP_ERR_HOST, "bad_alloc");
  expect(std::strstr(ngap_last_error(), "memory") != nullptr, "bad_alloc message");
  expect(c_guard([]() -> int { throw std::system_error(std::make_error_code(std::errc::resource_unavailable_try_again)); }) == NGAP_ERR_HOST,
         "system_error");
  expect(std::strstr(ngap_last_error(), "host error") != nullptr, "system_error message");
  expect(c_guard([]() -> int { throw 7; }) == NGAP_ERR_HOST, "anything else");
  expect(std::strcmp(ngap_last_error(), "host error") == 0, "anything else message");
  // entry points that return a pointer: NULL on an exception
  static const char text[] = "text";
  expect(ngapamd::c_guard_or([]() -> const char * { return text; }, (const char *)nullptr) == text,
         "pointer passes");
  expect(ngapamd::c_guard_or([]() -> const char * { throw std::bad_alloc(); },
                             (const char *)nullptr) == nullptr,
         "pointer on bad_alloc");
  expect(std::strcmp(ngap_last_error(), "host memory exhausted") == 0, "pointer message");
  // a stored message replaces the static one
  ngap_set_last_error_internal("worklist spill capacity exceeded");
  expect(std::strcmp(ngap_last_error(), "worklist spill capacity exceeded") == 0, "stored message");
  if (!fails) std::printf("PASS\n");
  return fails ? 1 : 0;
}
