// Cell kernel instantiations for the periodic containers
// (container_periodic, container_periodic_poly): the order-3, wide and huge
// tiers with the unit-cell start and periodic block indexing.
#include <hip/hip_runtime.h>

#include "vgpu_cells.h"

namespace vgpu {
int launch_cells_prd(int tier, bool poly, int waves, hipStream_t s, const RectView &v, int ncells, const int *list,
		     const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		     unsigned stamp_base, int per_wave, const CellOut &out) {
	if (tier == 0) {
		if (poly) VGPU_CELLS_LAUNCH_PRD(Tier3, true, 2);
		else VGPU_CELLS_LAUNCH_PRD(Tier3, false, 2);
	} else if (tier == 1) {
		if (poly) VGPU_CELLS_LAUNCH_PRD(TierWide, true, 2);
		else VGPU_CELLS_LAUNCH_PRD(TierWide, false, 2);
	} else {
		if (poly) VGPU_CELLS_LAUNCH_PRD(TierHuge, true, 2);
		else VGPU_CELLS_LAUNCH_PRD(TierHuge, false, 2);
	}
	return hipGetLastError() == hipSuccess ? 0 : 6;
}
}
