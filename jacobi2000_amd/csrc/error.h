// Host-side failure type shared by every part of libj2k.
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/j2k.h"

namespace j2k {

// A failed call carries the C-ABI status it maps to and a message that
// j2k_last_error() returns.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

}  // namespace j2k
