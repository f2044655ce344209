// kp_error.cpp -- thread-local last error (the ABI never throws).
#include "kp_error.h"

namespace {
thread_local std::string g_last_error;
}

namespace kp {
void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
} // namespace kp

extern "C" const char *kp_last_error(void) { return g_last_error.c_str(); }
