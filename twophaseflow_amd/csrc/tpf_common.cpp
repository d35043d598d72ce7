// Error state and version of libtpf.
#include "tpf_common.hpp"

namespace tpf {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace tpf

extern "C" const char* tpf_last_error(void) { return tpf::g_last_error.c_str(); }
extern "C" int tpf_version(void) { return 1; }
