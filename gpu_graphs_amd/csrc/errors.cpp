// Thread-local message of the last failing call, returned by gb_last_error().
#include <string>

#include "common.hpp"

namespace gg {

namespace {
thread_local std::string t_last_error;
}

void set_last_error(const std::string& msg) { t_last_error = msg; }
const char* last_error() { return t_last_error.c_str(); }

}  // namespace gg
