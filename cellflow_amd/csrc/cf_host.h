// Host-only helpers shared by the engine (cf_engine.hip) and the host-only sources
// (cf_preset.cpp, cf_host.cpp).  No HIP headers: the host-only sources also build with a
// plain C++ compiler (the sanitizer build of the tests).
#pragma once

#include <string>

#include "../../include/cellflow.h"

namespace cf {

// Records the message of a failing call (read back by cf_last_error_string) and returns
// `code`.
int fail(int code, const char* fmt, ...);
// The calling thread's last error message.
std::string& error_text();
// CF_OK when every field of `t` holds an allowed value, else CF_ERR_ARG with the reason.
int check_tuning(const cf_tuning* t);

}  // namespace cf
