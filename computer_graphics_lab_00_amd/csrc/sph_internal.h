// sph_internal.h — host-side helpers shared by the C ABI translation units.
#pragma once

#include <stdint.h>

#include <string>

#include "sph_c_api.h"

namespace sphhost {

// Record `message` as the calling thread's last error and return `status`.
int fail(int status, const std::string& message);
void clear_error();

// The marching-cubes case table (256 x 16 edge indices, -1 terminated), built once
// from its rule (sph_mc_table.cpp); nullptr if the rule failed to produce one.
const int8_t* mc_table();

}  // namespace sphhost
