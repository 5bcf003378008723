// Host-side sanitizer check of the C-ABI library: libcolloc built with
// AddressSanitizer and UndefinedBehaviorSanitizer on its host code (the
// device code is not instrumented), driven through the calls that need no
// GPU: version and error text, argument and shape checks, hiprtc
// compilation of a generated problem unit and of a broken source (the
// compiler log path), and handle creation, which without a device fails in
// its first allocation and must release what it made.  Any sanitizer report
// ends the process with a non-zero status.
//
// usage: asan_driver <unit.hip> <kernel include dir> <expect_device 0|1>
#include "colloc.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

static int failures = 0;

static void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s (last error: %s)\n", what, cs_last_error());
    ++failures;
  }
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::printf("usage: %s unit.hip include_dir expect_device\n", argv[0]);
    return 2;
  }
  std::ifstream f(argv[1]);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string src = ss.str();
  const bool device = std::strcmp(argv[3], "1") == 0;

  expect(cs_version() && std::strlen(cs_version()) > 0, "version string");

  // hiprtc compilation of the generated unit, twice (the second call reuses
  // nothing of the first)
  for (int rep = 0; rep < 2; ++rep) {
    void* code = nullptr;
    size_t size = 0;
    char* log = nullptr;
    const int rc = cs_jit_compile(src.c_str(), "unit.hip", argv[2], &code, &size, &log);
    expect(rc == CS_OK && code && size > 0, "compile the unit");
    cs_free(code);
    cs_free(log);
  }
  // a source that does not compile: an error code, a message and a log
  {
    void* code = nullptr;
    size_t size = 0;
    char* log = nullptr;
    const int rc = cs_jit_compile("__global__ void k() { undefined_name(); }", "bad.hip", argv[2], &code, &size,
                                  &log);
    expect(rc == CS_E_COMPILE, "broken source is a compile error");
    expect(code == nullptr, "no code object for a broken source");
    expect(log != nullptr && std::strlen(log) > 0, "compiler log of a broken source");
    expect(std::strlen(cs_last_error()) > 0, "error text of a broken source");
    cs_free(log);
  }
  expect(cs_jit_compile(nullptr, "x.hip", argv[2], nullptr, nullptr, nullptr) == CS_E_ARG, "null source");
  cs_free(nullptr);

  // argument and shape checks before any device work
  cs_handle* h = nullptr;
  expect(cs_create(nullptr, &h) == CS_E_ARG, "null dims");
  cs_dims bad = {0, 1, 1, 5, 16, 0, 1};
  expect(cs_create(&bad, &h) == CS_E_SHAPE, "ny = 0");
  cs_dims few = {4, 1, 0, 5, 3, 0, 2};
  expect(cs_create(&few, &h) == CS_E_SHAPE, "n_max below two nodes per instance");
  cs_destroy(nullptr);
  int st = 0;
  expect(cs_status(nullptr, &st) == CS_E_ARG, "null handle");

  // a valid handle: created and destroyed with a device, a clean HIP error
  // without one
  cs_dims ok = {4, 1, 0, 5, 64, 0, 2};
  const int rc = cs_create(&ok, &h);
  if (device) {
    expect(rc == CS_OK && h, "create a handle");
    cs_destroy(h);
  } else {
    expect(rc == CS_E_HIP && h == nullptr, "create without a device fails cleanly");
  }
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
