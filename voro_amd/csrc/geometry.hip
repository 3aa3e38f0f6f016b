// Polyhedron statistics pass (VGPU_F_GEOMETRY): per-cell output sizes in loop
// order, then one thread per cell computing every statistic from the cell
// tables the tiers kept (geom_dev.h).
#include <hip/hip_runtime.h>

#include "geom_dev.h"

namespace vgpu {

namespace {

// Output sizes of cell c (0 for a cell that does not exist): vertices, edge
// slots, face-vertex entries (one count per face plus its vertices); entry nc
// is 0 so the arrays scan to their totals.
__global__ void k_geo_sizes(int nc, const int *__restrict__ status, const int *__restrict__ nfaces, GeoSink g,
			    int *__restrict__ nv, int *__restrict__ ne, int *__restrict__ nfv) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c > nc) return;
	int a = 0, b = 0, f = 0;
	if (c < nc && (status[c] & ST_EXISTS) && g.np[c] > 0) {
		a = g.np[c];
		b = g.esz[c];
		f = b + nfaces[c];
	}
	nv[c] = a; ne[c] = b; nfv[c] = f;
}

__global__ void k_geo_stats(int nc, GeoSink g, GeoStats o) {
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < nc) cell_geometry_stats(g, o, c);
}

int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

int launch_geo_sizes(hipStream_t s, int nc, const int *status, const int *nfaces, const GeoSink &g, int *nv, int *ne,
		     int *nfv) {
	hipLaunchKernelGGL(k_geo_sizes, dim3(nblk(nc + 1, 256)), dim3(256), 0, s, nc, status, nfaces, g, nv, ne, nfv);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_geo_stats(hipStream_t s, int nc, const GeoSink &g, const GeoStats &o) {
	if (nc > 0) hipLaunchKernelGGL(k_geo_stats, dim3(nblk(nc, 64)), dim3(64), 0, s, nc, g, o);
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

}  // namespace vgpu
