// errors.hpp — the message behind bfg_last_error(), shared by the library's
// translation units.
#pragma once
#include <string>

namespace bfg {

// Sets the message bfg_last_error() returns on the calling thread.
void set_last_error(const std::string &msg);

}  // namespace bfg
