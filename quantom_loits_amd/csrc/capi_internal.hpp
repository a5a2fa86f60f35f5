// capi_internal.hpp -- error helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "loits_device.hpp"  // kMaxLds

namespace qloits {

void set_error(const std::string& msg);
int fail(int status, const std::string& msg);
int check_hip(hipError_t e, const char* what);

}  // namespace qloits
