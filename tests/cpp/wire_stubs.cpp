// Host-only stand-ins for the runtime symbols that mg_wire.cpp references, so
// the wire parser and the client functions can be built alone with the
// address and undefined-behaviour sanitizers (tests/cpp/fuzz_wire.cpp). The
// fuzz harness never reaches the request loop, so those entry points only
// report that they are absent.
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../morphogenesis_amd/csrc/mg_internal.hpp"

namespace mg {

namespace {
thread_local std::string g_err;
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

const char* last_error() { return g_err.c_str(); }

uint32_t server_row_bytes(const mg_server*) { return 0; }

}  // namespace mg

extern "C" {
const char* mg_last_error(void) { return mg::last_error(); }
int mg_server_submit(mg_server*, const uint8_t*, const size_t*, uint32_t, uint64_t*) {
  return MG_ERR_STATE;
}
int mg_server_wait(mg_server*, uint64_t, uint8_t*, uint32_t, mg_request_info*) {
  return MG_ERR_STATE;
}
int mg_server_abandon(mg_server*, uint64_t) { return MG_ERR_STATE; }
}
