// Error types of the library; the C ABI turns each into an OLFHIP_E_* code plus the message.
#pragma once

#include <stdexcept>
#include <string>

namespace olfhip {

struct ArgumentError : std::invalid_argument {
    explicit ArgumentError(const std::string &m) : std::invalid_argument(m) {}
};
struct UnsupportedError : std::domain_error {
    explicit UnsupportedError(const std::string &m) : std::domain_error(m) {}
};
struct DeviceError : std::runtime_error {
    explicit DeviceError(const std::string &m) : std::runtime_error(m) {}
};
struct MemoryError : std::runtime_error {
    explicit MemoryError(const std::string &m) : std::runtime_error(m) {}
};

}  // namespace olfhip
