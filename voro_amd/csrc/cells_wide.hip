// Cell kernel instantiations for the wide tier (own translation unit so the
// device code of the tiers compiles in parallel).
#include <hip/hip_runtime.h>

#include "vgpu_cells.h"

namespace vgpu {
int launch_cells_wide(bool poly, int occ, int waves, hipStream_t s, const RectView &v, int ncells, const int *list,
		      const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		      unsigned stamp_base, int per_wave, const CellOut &out) {
	(void)occ;
	if (poly) VGPU_CELLS_LAUNCH(TierWide, true, 2);
	else VGPU_CELLS_LAUNCH(TierWide, false, 2);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}
}
