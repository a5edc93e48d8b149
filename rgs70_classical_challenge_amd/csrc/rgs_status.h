// Status codes and error plumbing shared by host translation units.
#pragma once
#include <string>

namespace rgs {

// Status codes returned by every C entry point (mirrored in include/rgs_amd.h).
enum Status : int {
  kOk = 0,
  kInvalid = 1,     // bad argument (Python ValueError)
  kArith = 2,       // non-finite or failed norm check (ArithmeticError)
  kNoMemory = 3,    // device memory exhausted (MemoryError)
  kHip = 4,         // HIP runtime failure (RuntimeError)
};

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

}  // namespace rgs
