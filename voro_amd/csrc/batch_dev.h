// Shared by the binning kernels and the batch kernels: the rectangular
// container's binning rule, and the batch work-queue types and launchers.
#pragma once
#include "vgpu_cells.h"

namespace vgpu {

// Geometry needed to bin particles of a rectangular container.
struct RectGeom {
	double ax, bx, ay, by, az, bz;
	int nx, ny, nz;
	double boxx, boxy, boxz, xsp, ysp, zsp;
	int xper, yper, zper;
};

__device__ __forceinline__ int step_int_d(double a) { return a < 0 ? int(a) - 1 : int(a); }
__device__ __forceinline__ int step_mod_d(int a, int b) { return a >= 0 ? a % b : b - 1 - (b - 1 - a) % b; }

// container_base::put_remap (src/container.cc:174-191): the block of (x,y,z)
// in the container, the position remapped into the domain in periodic
// directions; -1 if the particle is dropped.  Non-finite or huge
// coordinates would make the int conversion undefined: they lie outside any
// box, so they are dropped.
__device__ __forceinline__ int locate_rect(const RectGeom &g, double &x, double &y, double &z) {
	const double lim = 1073741824.0;
	const double ax_ = (x - g.ax) * g.xsp, ay_ = (y - g.ay) * g.ysp, az_ = (z - g.az) * g.zsp;
	if (!(fabs(ax_) < lim && fabs(ay_) < lim && fabs(az_) < lim)) return -1;
	int i = step_int_d(ax_), j, k, l;
	bool ok = true;
	if (g.xper) { l = step_mod_d(i, g.nx); x += g.boxx * (l - i); i = l; }
	else if (i < 0 || i >= g.nx) ok = false;
	j = step_int_d(ay_);
	if (g.yper) { l = step_mod_d(j, g.ny); y += g.boxy * (l - j); j = l; }
	else if (j < 0 || j >= g.ny) ok = false;
	k = step_int_d(az_);
	if (g.zper) { l = step_mod_d(k, g.nz); z += g.boxz * (l - k); k = l; }
	else if (k < 0 || k >= g.nz) ok = false;
	return ok ? i + g.nx * j + g.nx * g.ny * k : -1;
}

// A run of up to CH consecutive cells of one container.
struct BatchChunk { int cont, first, count, pad; };

// Set-up kernels (cells_batch.hip).
int launch_batch_mrad(hipStream_t s, int ncont, const RectGeom *geoms, const unsigned *wl, double *mrad);
int launch_batch_locate(hipStream_t s, int n, const double *xyz, const double *r, const RectGeom *geoms,
			const int *pcont, const int *boff, int *blk, int *slot, dbl4 *rem, int *cnt,
			unsigned long long *maxr_bits);
int launch_batch_counts(hipStream_t s, int ncont, int ch, const int *boff, const int *bstart,
			const unsigned long long *maxr_bits, int poly, RectView *views, int *cfirst, int *nchunk);
int launch_batch_chunks(hipStream_t s, int ncont, int ch, const int *cfirst, const int *coff, BatchChunk *chunks,
			int *nchunks);

// The cell kernels.  First tier over the chunk queue (order-3 cells at
// 2/3/4 waves per SIMD); retry tiers (tier 0 order-3, 1 wide, 2 huge) over a
// list.
// Block-mask stamps start at *stamp_dev, read on the device.
int launch_batch_first(bool poly, int occ, int waves, hipStream_t s, const RectView *views, const BatchChunk *chunks,
		       const int *nchunks, const int *cellblk, int *counters, char *ws, size_t slab,
		       const unsigned *stamp_dev, const CellOut &out);
// Wavefront-cell first tier over the chunk queue (cells_wave.hip), occ
// (4, 5 or 6) waves per SIMD.
int launch_batch_wfirst(bool poly, int occ, int waves, hipStream_t s, const RectView *views, const BatchChunk *chunks,
			const int *nchunks, const int *cellblk, int *counters, char *ws, size_t slab,
			const unsigned *stamp_dev, const CellOut &out);
int launch_batch_retry(int tier, bool poly, int waves, hipStream_t s, const RectView *views, const int *boff,
		       int ncont, const int *list, const int *list_count, const int *cellblk, int *counters, char *ws,
		       size_t slab, const unsigned *stamp_dev, const CellOut &out);

// The learned tier of a batch (cells_list.hip): the container of every
// block; block features; candidate lists built in one pass into a fixed
// capacity (lists that do not fit give their cells no list; *top: the
// records everything needed); the list-driven wavefront tier over the
// schedule, whose cells without a usable list continue as plain wavefront
// cells (block-mask stamps from *stamp_dev).  Counts are read on the device
// (ncp: the cells), so the sequence can be captured.
int launch_batch_bcont(hipStream_t s, int ncont, const int *boff, int *bcont);
int launch_batch_features(hipStream_t s, int nb, const int *bcont, const RectView *views, float *X);
int launch_batch_list_build(hipStream_t s, const RectView *views, const int *bcont, int n, const int *ncp,
			    const int *cellblk, const int *order, const unsigned char *depth, long long cap,
			    unsigned long long *top, int *lcnt, int *loff, ListRec *LA);
int launch_batch_wcells_list(bool poly, int occ, int waves, int chunk, hipStream_t s, const RectView *views,
			     const int *bcont, const int *ncp, const int *cellblk, const int *order, const int *lcnt,
			     const int *loff, const ListRec *LA, int *counters, char *ws, size_t slab,
			     const unsigned *stamp_dev, int *fallbacks, int *routed, const CellOut &out);

}  // namespace vgpu
