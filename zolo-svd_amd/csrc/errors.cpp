// Last-error message of the C ABI (zolo_last_error), per host thread.
#include <string>

#include "common.hpp"

namespace zolo {
namespace {
thread_local std::string g_err;
}
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }
}  // namespace zolo
