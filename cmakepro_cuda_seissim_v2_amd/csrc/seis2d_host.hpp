// seis2d_host.hpp - error reporting shared by the 2D library's translation units.
#pragma once

#include <exception>
#include <string>

#include "../../include/seis2d.h"

namespace seis2d::host {

// Message of the last failed call on this thread (seis2d_last_error).
inline thread_local std::string g_error;

struct Failure {
    int code;
    std::string message;
};
[[noreturn]] inline void fail(int code, const std::string& m) { throw Failure{code, m}; }
[[noreturn]] inline void invalid(const std::string& m) { fail(SEIS2D_EINVAL, m); }

// Runs fn, mapping a Failure to its code and any other exception to SEIS2D_EHIP.
template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return SEIS2D_OK;
    } catch (const Failure& f) {
        g_error = f.message;
        return f.code;
    } catch (const std::exception& e) {
        g_error = e.what();
        return SEIS2D_EHIP;
    }
}

}  // namespace seis2d::host
