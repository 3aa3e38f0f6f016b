// The persistent cell kernel and the per-tier launchers.
#pragma once
#include "compute_dev.h"
#include "geom_dev.h"

namespace vgpu {

// Results of the cell kernels, loop order.
struct CellOut {
	double *vol;
	double *cent;        // 3 per cell
	int *nfaces;
	int *status;         // ST_* bits
	int *nbr_start;      // offset into the neighbour arena
	int *nbr;            // neighbour arena
	int *arena_top;      // bump allocator over nbr
	int arena_cap;
	int *retry;          // cells for the next tier
	int *retry_count;
	int *fail_count;     // cells the last tier could not finish
	short *st_planes;    // optional statistics
	signed char *st_exit;
	signed char *st_tier;
	unsigned char *touch; // periodic containers: 1 if the cell read an image block
	int tier;
	int want_nbrs;
	GeoSink geo;          // cell tables for the polyhedron statistics (geo.top null: not kept)
};

// Vertex order of a per-lane cell: 3 in the order-3 cell, stored in the
// general one.
template<class CELL> VG_INL int cell_order(CELL &, int) { return 3; }
template<class T_> VG_INL int cell_order(VCell<T_, 64> &c, int v) { return (int)c.NU(v); }

// Copies a finished per-lane cell's tables (coordinates, orders, edges and
// back pointers) into the geometry arena; a cell that does not fit is
// counted in top[2] (the caller grows the arena and computes again).
template<class CELL>
__device__ void geo_write_lane(CELL &c, const GeoSink &g, int cell, bool ok) {
	if (!ok) { g.vstart[cell] = 0; g.np[cell] = 0; g.esz[cell] = 0; return; }
	const int p = c.p;
	int esz = 0;
	for (int v = 0; v < p; v++) esz += cell_order(c, v);
	const int vs = atomicAdd(&g.top[0], p), es = atomicAdd(&g.top[1], esz);
	if ((long long)vs + p > g.vcap || (long long)es + esz > g.ecap) {
		g.vstart[cell] = -1; g.np[cell] = 0; g.esz[cell] = 0;
		atomicAdd(&g.top[2], 1);
		return;
	}
	g.vstart[cell] = vs; g.np[cell] = p; g.estart[cell] = es; g.esz[cell] = esz;
	int e = es;
	for (int v = 0; v < p; v++) {
		const int n = cell_order(c, v);
		g.xyz[3 * (size_t)(vs + v)] = c.X(v);
		g.xyz[3 * (size_t)(vs + v) + 1] = c.Y(v);
		g.xyz[3 * (size_t)(vs + v) + 2] = c.Z(v);
		g.nu[vs + v] = (unsigned char)n;
		g.eoff[vs + v] = e;
		for (int i = 0; i < n; i++) {
			g.ed[e + i] = (short)c.E(v, i);
			g.rel[e + i] = (unsigned char)c.R(v, i);
		}
		e += n;
	}
}

// Neighbour ids of one cell go to a scratch row of the lane's slab first,
// then to the arena once the count is known.
template<class CELL>
struct SlabSink {
	CELL *c;
	VG_INL void put(int i, int v) {
		if (i < (int)CELL::L::NBCAP) c->template A<int>(CELL::L::o_nbs, i) = v;
	}
};

// Cell type of each tier.
template<class T_> struct TierCell { typedef VCell<T_, 64> type; };
template<> struct TierCell<Tier3> { typedef VCell3 type; };

// Persistent cell kernel, one wavefront per workgroup.  Wavefronts take
// 64 consecutive cells at a time; the cell range is split into one
// contiguous share per XCD (workgroups are dispatched to the 8 XCDs round
// robin) so that neighbouring cells, which read the same particle blocks,
// share an L2; a wavefront whose share is exhausted moves to the next one.
// per_wave (1..64) is the number of cells a wavefront takes at a time: 64
// for a full tier, fewer for a short retry list of expensive cells, which
// then spreads over all wavefronts instead of a few full ones.  PRD selects
// the periodic container's start cell and block indexing.
template<class T_, bool POLY, int OCC, bool PRD = false>
__global__ void __launch_bounds__(64, OCC)
k_cells(RectView v, int ncells, const int *__restrict__ list, const int *__restrict__ list_count,
	const int *__restrict__ cellblk, int *__restrict__ counters, char *__restrict__ ws, size_t slab_bytes,
	unsigned stamp_base, int per_wave, CellOut out) {
	typedef typename TierCell<T_>::type CELL;
	const int lane = threadIdx.x;
	CELL c;
	c.setup(ws + (size_t)blockIdx.x * slab_bytes, lane, v.max_len_sq);
	const int total = list_count ? *list_count : ncells;
	const int nshare = 8;
	const int my = blockIdx.x % nshare;
	const int share = (total + nshare - 1) / nshare;
	const bool last_tier = T_::MAXV >= TierHuge::MAXV;
	for (int h = 0; h < nshare; h++) {
		const int sh = (my + h) % nshare;
		const int lo = sh * share, hi = min(total, lo + share);
		for (;;) {
			int base = 0;
			if (lane == 0) base = atomicAdd(&counters[sh], per_wave);
			base = __shfl(base, 0);
			if (lo + base >= hi) break;
			const int idx = lo + base + lane;
			if (idx < hi && lane < per_wave) {
				const int cell = list ? list[idx] : idx;
				const int b = cellblk[cell];
				const int s = cell - v.cstart[b];
				const int ci = b % v.nx, cj = (b / v.nx) % v.gy, ck = b / v.nxy;
				CellStats stt;
				bool ok = compute_cell_rect<CELL, POLY, PRD>(c, v, b, s, ci, cj, ck, stamp_base + (unsigned)cell + 1u,
									     stt, (unsigned long long)cell << 24);
				double vol = 0, cx = 0, cy = 0, cz = 0;
				int nf = 0;
				if (!c.ovf && ok) {
					SlabSink<CELL> sk{&c};
					c.outputs(vol, cx, cy, cz, nf, sk);
					if (nf > (int)CELL::L::NBCAP) c.ovf = true;
				}
				if (c.ovf) {
					if (last_tier) {
						atomicAdd(out.fail_count, 1);
						out.status[cell] = c.fatal ? ST_FATAL : ST_OVERFLOW;
					} else {
						int slot_ = atomicAdd(out.retry_count, 1);
						out.retry[slot_] = cell;
						out.status[cell] = ST_OVERFLOW;
					}
				} else {
					out.vol[cell] = vol;
					out.cent[3 * cell] = cx; out.cent[3 * cell + 1] = cy; out.cent[3 * cell + 2] = cz;
					out.nfaces[cell] = nf;
					out.status[cell] = ok ? ST_EXISTS : 0;
					if (out.want_nbrs) {
						int off = nf > 0 ? atomicAdd(out.arena_top, nf) : 0;
						if (off + nf > out.arena_cap) { out.nbr_start[cell] = -1; atomicAdd(out.fail_count, 1); }
						else {
							out.nbr_start[cell] = off;
							for (int t = 0; t < nf; t++) out.nbr[off + t] = c.template A<int>(CELL::L::o_nbs, t);
						}
					}
					if (out.st_planes) {
						out.st_planes[cell] = (short)min(stt.n_planes, 32767);
						out.st_exit[cell] = (signed char)stt.wl_exit;
						out.st_tier[cell] = (signed char)out.tier;
					}
					if (PRD && out.touch) out.touch[cell] = (unsigned char)stt.img;
					if (out.geo.top) geo_write_lane(c, out.geo, cell, ok);
				}
			}
		}
	}
}

// Launches k_cells<tier,poly,occupancy> (one translation unit per tier
// keeps the build parallel).  `occ` is the waves-per-SIMD register budget
// the kernel was compiled for (2, 3 or 4 for the fast tier).
typedef int (*cells_launcher)(bool poly, int occ, int waves, hipStream_t s, const RectView &v, int ncells,
			      const int *list, const int *list_count, const int *cellblk, int *counters, char *ws,
			      size_t slab, unsigned stamp_base, int per_wave, const CellOut &out);
int launch_cells_fast3(bool poly, int occ, int waves, hipStream_t s, const RectView &v, int ncells, const int *list,
		       const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		       unsigned stamp_base, int per_wave, const CellOut &out);
int launch_cells_wide(bool poly, int occ, int waves, hipStream_t s, const RectView &v, int ncells, const int *list,
		      const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		      unsigned stamp_base, int per_wave, const CellOut &out);
int launch_cells_huge(bool poly, int occ, int waves, hipStream_t s, const RectView &v, int ncells, const int *list,
		      const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		      unsigned stamp_base, int per_wave, const CellOut &out);

// The wavefront-cell tiers (cells_wave.hip): one cell per wavefront with k
// = 1, 2 or 4 vertices per lane (up to 64, 128, 254 vertices), over every
// cell or a device list (count at *list_count), at 2, 3, 4 or 8 waves per
// SIMD (k = 1; k = 2 and 4 run at 2); cells it cannot finish go to
// out.retry.
// find_voronoi_cell for n query points (find_cell.hip): tiers 0, 1, 2 search
// with the order-3, wide and huge cells' mask windows and send queries that
// leave the window to `retry` (none: found[q] = -1).  prd: a periodic
// container (touch[q] = 1 when query q read an image block; first visits in
// v.fv).
int launch_find_cells(int tier, bool poly, bool prd, int waves, hipStream_t s, const RectView &v, int n,
		      const double *xyz, const int *list, const int *list_count, int *found, int *pid, double *pos,
		      int *retry, int *retry_count, int *counter, char *ws, size_t slab, unsigned stamp_base,
		      unsigned char *touch);
struct RectGeom;
// compute_ghost_cell for n points (or the list of them counted at
// *list_count) (find_cell.hip), general per-lane cells: tier 0 the wide
// cell (points it cannot finish go to retry, counted at *failed), tier 1 the
// huge one (retry null: *failed counts the points it cannot finish).
// prd: a periodic container (the searches read v's block arrays and stop at
// the blocks v.img_ok excludes).
int launch_ghost_cells(int tier, bool poly, bool prd, int waves, hipStream_t s, const RectView &v,
		       const RectGeom &geo, int n, const double *xyz, const double *rad, const int *list,
		       const int *list_count, int *exists, double *vol, double *cent, int *nfaces, int *nbr, int max_nbr,
		       int *counter, int *retry, int *failed, char *ws, size_t slab, unsigned stamp_base);
// Waves per SIMD of the 128-vertex wavefront tier (k = 2): four, whose
// register spills cost less than the latency two waves leave exposed
// (periodic radical 2M, ~44k cells over 64 vertices: the tiers after the
// first take 135 instead of 154 ms per compute; 3, 5, 6 waves: 141, 135,
// 144 ms).
#ifndef VG_K2_OCC
#define VG_K2_OCC 4
#endif
int launch_cells_wave(int k, bool poly, bool prd, int occ, int waves, hipStream_t s, const RectView &v, int ncells,
		      const int *list, const int *list_count, const int *cellblk, int *counters, char *ws,
		      unsigned stamp_base, int chunk, const CellOut &out);

// Polyhedron statistics (geometry.hip): per-cell output sizes (nc + 1
// entries, the last 0) and the statistics pass.
int launch_geo_sizes(hipStream_t s, int nc, const int *status, const int *nfaces, const GeoSink &g, int *nv, int *ne,
		     int *nfv);
int launch_geo_stats(hipStream_t s, int nc, const GeoSink &g, const GeoStats &o);

// Periodic containers (cells_prd.hip): tier 0 order-3, 1 wide, 2 huge.
int launch_cells_prd(int tier, bool poly, int waves, hipStream_t s, const RectView &v, int ncells, const int *list,
		     const int *list_count, const int *cellblk, int *counters, char *ws, size_t slab,
		     unsigned stamp_base, int per_wave, const CellOut &out);

// Learned path (cells_list.hip): list gathering and the list-driven tier.
// prd: a periodic container's view (the gather needs v.fv null; the cell
// kernels record first visits through v.fv and image reads in out.touch).
// ncp (may be null): the cell count on the device, at most nc.
int launch_list_order(hipStream_t s, int nc, const int *cellblk, const unsigned char *cost, int per_cell, int *hist,
		      int *order, const int *ncp = nullptr);
int launch_list_gather(bool prd, hipStream_t s, const RectView &v, int nc, const int *cellblk, const int *order,
		       const unsigned char *depth, int per_cell, int *cnt, int *bmax);
int launch_list_fill(bool prd, hipStream_t s, const RectView &v, int nc, const int *cellblk, const int *order,
		     const unsigned char *depth, int per_cell, const int *boff, bool csr, ListRec *LA);
int launch_cells_list(bool poly, bool prd, int waves, hipStream_t s, const RectView &v, int nc, const int *cellblk,
		      const int *order, const int *cnt, const int *boff, const ListRec *LA,
		      int *counters, char *ws, int *fallbacks, int *routed, const CellOut &out);
// The list-driven tier with one cell per wavefront (4 waves per SIMD) over
// lists filled with csr, cells taken `chunk` at a time; ws holds one
// wavefront-cell slab per wavefront.  Cells without a usable list are
// computed by the plain algorithm in place (block-mask stamps from
// stamp_base + cell + 1); only overflows go to out's retry list.
int launch_cells_wlist(bool poly, bool prd, int occ, int waves, int chunk, hipStream_t s, const RectView &v, int nc,
		       const int *cellblk, const int *order, const int *cnt, const int *boff, const ListRec *LA,
		       int *counters, char *ws, unsigned stamp_base, int *fallbacks, int *routed, const CellOut &out);

#define VGPU_CELLS_LAUNCH(TIER, P, O)                                                                      \
	hipLaunchKernelGGL((k_cells<TIER, P, O>), dim3(waves), dim3(64), 0, s, v, ncells, list, list_count,    \
			   cellblk, counters, ws, slab, stamp_base, per_wave, out)
#define VGPU_CELLS_LAUNCH_PRD(TIER, P, O)                                                                  \
	hipLaunchKernelGGL((k_cells<TIER, P, O, true>), dim3(waves), dim3(64), 0, s, v, ncells, list,          \
			   list_count, cellblk, counters, ws, slab, stamp_base, per_wave, out)

}  // namespace vgpu
