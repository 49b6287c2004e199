// Error plumbing shared by the C-ABI entry points.
#pragma once

#include <exception>
#include <string>

#include "error.h"

namespace j2k {

void set_last_error(const std::string& msg);

// Runs f; maps an Error to its status (J2K_PENDING is a status, not an
// error, and leaves the last error alone).
template <class F>
int guard(F&& f) {
    try {
        f();
        return J2K_OK;
    } catch (const Error& e) {
        if (e.code != J2K_PENDING) set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return J2K_ERR_POLL;
    }
}

}  // namespace j2k
