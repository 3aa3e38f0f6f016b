// Cell kernel instantiations for the fast tier (own translation unit so the
// device code of the three tiers compiles in parallel).
#include <hip/hip_runtime.h>

#include "vgpu_cells.h"

namespace vgpu {
VGPU_DEFINE_CELLS_LAUNCHER(launch_cells_fast, TierFast)
}
