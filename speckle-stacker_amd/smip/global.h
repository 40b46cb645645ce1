// global.h -- version and export macros of the drop-in surface (the reference
// generates its global.h from config/config.h:1-52 at configure time).
#pragma once

#include <string>

#if defined(__GNUC__)
#define SMIP_PUBLIC __attribute__((visibility("default")))
#else
#define SMIP_PUBLIC
#endif

namespace smip::Version {
constexpr int major { 0 };
constexpr int minor { 3 };
constexpr int patch { 2 };
constexpr const char* additional { "mi355x" };
[[nodiscard]] inline auto string() -> std::string
{
    return std::to_string(major) + "." + std::to_string(minor) + "." + std::to_string(patch) + "-" + additional;
}
} // namespace smip::Version
