/**
 * goetia_amd.hh -- common pieces of the C++ facade over libgoetia_amd:
 * goetia's exception types and count/byte typedefs, and the mapping of C ABI
 * status codes to those exceptions.
 *
 * The facade (include/goetia/...) is header-only C++17 with no HIP headers,
 * so it can be compiled by any host compiler or parsed by cling; every
 * storage and sequence operation goes through the C ABI in goetia_amd.h and
 * runs on the GPU.
 */
#ifndef GOETIA_AMD_FACADE_HH
#define GOETIA_AMD_FACADE_HH

#include <cstdint>
#include <exception>
#include <stdexcept>
#include <string>

#include "goetia_amd.h"

namespace goetia {

typedef uint8_t byte_t;                       // storage.hh:79
typedef int16_t count_t;                      // storage.hh:80

// goetia.hh:140-153
class GoetiaException : public std::exception {
  public:
    explicit GoetiaException(const std::string& msg = "Generic goetia exception") : _msg(msg) {}
    const char* what() const noexcept override { return _msg.c_str(); }

  protected:
    std::string _msg;
};

class GoetiaFileException : public GoetiaException {
  public:
    using GoetiaException::GoetiaException;
};

// exceptions.hh
class InvalidSequenceException : public GoetiaException {
  public:
    using GoetiaException::GoetiaException;
};

class SequenceLengthException : public GoetiaException {
  public:
    using GoetiaException::GoetiaException;
};

class UninitializedShifterException : public GoetiaException {
  public:
    UninitializedShifterException() : GoetiaException("Shifter used without hash_base") {}
};

// alphabets.hh:41-55
class InvalidCharacterException : public GoetiaException {
  public:
    using GoetiaException::GoetiaException;
};

class DeviceException : public GoetiaException {
  public:
    using GoetiaException::GoetiaException;
};

namespace detail {

inline void check(int rc, const char* what) {
    if (rc == GT_OK) return;
    std::string msg = std::string(what) + ": " + gt_last_error();
    switch (rc) {
        case GT_ERR_SEQUENCE_LENGTH: throw SequenceLengthException(msg);
        case GT_ERR_IO: throw GoetiaFileException(msg);
        case GT_ERR_INVALID_CHARACTER: throw InvalidCharacterException(msg);
        case GT_ERR_INVALID_ARG: throw std::invalid_argument(msg);
        case GT_ERR_OUT_OF_MEMORY: throw std::bad_alloc();
        case GT_ERR_DEVICE:
        case GT_ERR_NO_DEVICE: throw DeviceException(msg);
        default: throw GoetiaException(msg);
    }
}

}  // namespace detail
}  // namespace goetia

#endif
