// Workspace for the entry points that allocate nothing: one static device array per GPU,
// reserved when the library's code object is loaded and held until the process ends.
#pragma once
#include <cstdint>

#include "vector_ops.h"

namespace ExaDG
{
// largest operator (DoF) the no-allocation CG accepts
constexpr int64_t kNoAllocMaxDofs = EXADG_NOALLOC_MAX_DOFS;
// doubles in the workspace: the CG vectors r, z, p, v (kNoAllocMaxDofs each, in that order)
// followed by the device memory of one Reducer
constexpr int64_t kNoAllocDoubles = 4 * kNoAllocMaxDofs + Reducer::kDeviceDoubles;

// Device address of the current GPU's workspace: looked up once per device
// (hipGetSymbolAddress with that device current) and cached; allocates nothing.
double * noalloc_workspace();
}  // namespace ExaDG
