// noalloc.hpp -- the fixed device workspace (see noalloc.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace dmgx {

// Bytes of the fixed workspace each GPU holds: room for about 17,000 slots of
// the carbon model (rows of 152 entries) or 21,000 of the Cu model (128).
constexpr size_t kFixedWorkspaceBytes = size_t(128) << 20;

// Device address and size of the current device's copy of the fixed
// workspace; the address is looked up on the first call on each device.
hipError_t fixed_workspace(void** ptr, size_t* bytes);

}  // namespace dmgx
