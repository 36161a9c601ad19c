// sph_internal.h — host-side helpers shared by the C ABI translation units.
#pragma once

#include <string>

#include "sph_c_api.h"

namespace sphhost {

// Record `message` as the calling thread's last error and return `status`.
int fail(int status, const std::string& message);
void clear_error();

}  // namespace sphhost
