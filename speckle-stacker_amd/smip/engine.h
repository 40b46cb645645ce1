// engine.h -- glue between the drop-in C++ surface and libsmip_mi355x.so:
// the process-wide device context and the mapping of ABI status codes back to
// the exception types the reference throws.
#pragma once

#include <cstdlib>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>

#include "smip_mi355x.h"

namespace smip::detail {

[[noreturn]] inline void throw_status(int rc)
{
    const std::string m = smip_last_error();
    switch (rc) {
    case SMIP_ERR_INVALID_ARGUMENT:
        throw std::invalid_argument(m);
    case SMIP_ERR_OUT_OF_RANGE:
        throw std::out_of_range(m);
    case SMIP_ERR_BAD_FUNCTION_CALL:
        throw std::bad_function_call();
    case SMIP_ERR_LOGIC:
        throw std::logic_error(m);
    case SMIP_ERR_DOMAIN:
        throw std::domain_error(m);
    default:
        throw std::runtime_error(m);
    }
}

inline void check(int rc)
{
    if (rc != SMIP_OK)
        throw_status(rc);
}

// Device context shared by every drop-in object (device from SMIP_DEVICE, default 0).
// It is made on first use and never destroyed: a static destructor that released
// it would call HIP while the process exits, after or beside the HIP runtime's
// own teardown and while static drop-in objects of other translation units may
// still hold it.  The runtime reclaims the context's stream at exit.
inline smip_ctx* context()
{
    static smip_ctx* const ctx = [] {
        const char* dev = std::getenv("SMIP_DEVICE");
        smip_ctx* c = nullptr;
        check(smip_ctx_create(dev ? std::atoi(dev) : 0, &c));
        return c;
    }();
    return ctx;
}

} // namespace smip::detail
