// Error type and status codes shared by the host layer. No HIP dependency,
// so host-only parts (the proof server's scheduler) build without ROCm.
#pragma once
#include <stdexcept>
#include <string>

namespace vx {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
enum Status : int {
  OK = 0,
  ERR_ARG = 1,      // invalid argument (shape, size, null pointer)
  ERR_HIP = 2,      // HIP runtime error
  ERR_ALLOC = 3,    // device allocation failed
  ERR_VERIFY = 4,   // proof rejected
  ERR_INTERNAL = 5,
};
inline void require(bool cond, const char* what) {
  if (!cond) throw Error(ERR_ARG, what);
}

}  // namespace vx
