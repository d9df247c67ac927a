// common.hpp — error reporting shared by the host-side translation units.
#pragma once

#include <stdexcept>
#include <string>

namespace clipamd {

struct ConfigError : std::runtime_error {
    explicit ConfigError(const std::string& m) : std::runtime_error(m) {}
};

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);

}  // namespace clipamd
