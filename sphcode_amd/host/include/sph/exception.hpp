// Errors: the reference throws SPHException from THROW_ERROR
// (include/exception.hpp:9-56); here the same exception carries the C ABI
// status of the failing call.
#pragma once

#include <stdexcept>
#include <string>

namespace sph {

class SPHException : public std::runtime_error {
    int m_code;

public:
    SPHException(const std::string &message, int code = -1) : std::runtime_error(message), m_code(code) {}
    int code() const { return m_code; }
};

// Throws SPHException for a non-zero status of a sphgpu_* call.
void check_status(int status, const char *call);

}  // namespace sph
