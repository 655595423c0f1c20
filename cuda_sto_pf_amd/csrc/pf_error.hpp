// Per-thread last-error message for the C ABI.
#pragma once
#include <string>

namespace pf {
std::string& last_error();
inline int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}
}  // namespace pf
