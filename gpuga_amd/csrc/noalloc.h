// The fixed device workspace of the allocation-free entry point (noalloc.hip).
#pragma once

#include <cstddef>

namespace gpuga {

// Size of the workspace: per GPU, held until the process ends.
constexpr size_t kStaticWorkspaceBytes = (size_t)64 << 20;

// Device address of the current device's workspace.  The first call on a
// device looks it up (hipGetSymbolAddress, with that device current); later
// calls return the kept address.  Allocates nothing.
void* static_workspace();

} // namespace gpuga
