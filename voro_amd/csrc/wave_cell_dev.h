// One cell computed by a wavefront (wcell_dev.h): compute_cell_rect, then
// the outputs or an entry in the retry list.  Shared by the wavefront-cell
// tiers (cells_wave.hip) and the list-driven tier (cells_list.hip), whose
// cells without a usable list continue here.
#pragma once
#include <hip/hip_runtime.h>

#include "vgpu_cells.h"
#include "wcell_dev.h"

namespace vgpu {

// Neighbour ids in face-walk order: id i in lane i & 63 of set i >> 6.
template<int K>
struct WSink {
	WCellT<K> *c;
	VG_INL void put(int i, int v) { if (i < (int)WLayoutT<K>::NBCAP) c->put_nb(i, v); }
};

// Copies a finished wavefront cell's tables into the geometry arena: each
// lane writes the vertices it holds (order 3: three edges and back pointers
// from the topology word); a cell that does not fit is counted in top[2].
template<int K>
__device__ __forceinline__ void wave_geo_write(WCellT<K> &c, const GeoSink &g, int cell, bool ok) {
	const int lane = threadIdx.x;
	if (!ok) {
		if (lane == 0) { g.vstart[cell] = 0; g.np[cell] = 0; g.esz[cell] = 0; }
		return;
	}
	const int p = c.p;
	int vs = 0, es = 0;
	if (lane == 0) { vs = atomicAdd(&g.top[0], p); es = atomicAdd(&g.top[1], 3 * p); }
	vs = __builtin_amdgcn_readfirstlane(vs);
	es = __builtin_amdgcn_readfirstlane(es);
	if ((long long)vs + p > g.vcap || (long long)es + 3 * p > g.ecap) {
		if (lane == 0) { g.vstart[cell] = -1; g.np[cell] = 0; g.esz[cell] = 0; atomicAdd(&g.top[2], 1); }
		return;
	}
	if (lane == 0) { g.vstart[cell] = vs; g.np[cell] = p; g.estart[cell] = es; g.esz[cell] = 3 * p; }
	eachk<K>([&](auto hc) __attribute__((always_inline)) {
		constexpr int h = decltype(hc)::value;
		const int v = 64 * h + lane;
		if (v < p) {
			const size_t a = (size_t)(vs + v);
			g.xyz[3 * a] = c.vx[h]; g.xyz[3 * a + 1] = c.vy[h]; g.xyz[3 * a + 2] = c.vz[h];
			g.nu[a] = 3;
			g.eoff[a] = es + 3 * v;
			const unsigned w = c.vt[h];
			for (int i = 0; i < 3; i++) {
				g.ed[es + 3 * v + i] = (short)WCellT<K>::twE(w, i);
				g.rel[es + 3 * v + i] = (unsigned char)WCellT<K>::twR(w, i);
			}
		}
	});
}

// One cell of the container with view v, block b of the view's numbering:
// compute_cell, then the outputs (lane 0 writes the scalars, every lane its
// neighbour ids) or, if the cell overflowed, an entry in out's retry list.
// resume: the cell is continued where a candidate list ended.
template<int K, bool POLY, bool PRD>
__device__ __forceinline__ void wave_cell(WCellT<K> &c, const RectView &v, int cell, int b, unsigned stamp,
					  const CellOut &out, const ListResume *resume = nullptr) {
	typedef WLayoutT<K> L;
	const int lane = threadIdx.x;
	const int s = cell - v.cstart[b];
	const int ci = b % v.nx, cj = (b / v.nx) % v.gy, ck = b / v.nxy;
	CellStats stt;
	const bool ok = compute_cell_rect<WCellT<K>, POLY, PRD>(c, v, b, s, ci, cj, ck, stamp, stt,
							  (unsigned long long)cell << 24, nullptr, 0, resume);
	double vol = 0, cx = 0, cy = 0, cz = 0;
	int nf = 0;
	if (!c.ovf && ok) {
		WSink<K> sk{&c};
		c.outputs(vol, cx, cy, cz, nf, sk);
		if (nf > (int)L::NBCAP) c.ovf = true;
	}
	if (c.ovf) {
		if (lane == 0) {
			const int slot_ = atomicAdd(out.retry_count, 1);
			out.retry[slot_] = cell;
			out.status[cell] = ST_OVERFLOW;
		}
		return;
	}
	if (lane == 0) {
		out.vol[cell] = vol;
		out.cent[3 * cell] = cx; out.cent[3 * cell + 1] = cy; out.cent[3 * cell + 2] = cz;
		out.nfaces[cell] = nf;
		out.status[cell] = ok ? ST_EXISTS : 0;
		if (out.st_planes) {
			out.st_planes[cell] = (short)min(stt.n_planes, 32767);
			out.st_exit[cell] = (signed char)stt.wl_exit;
			out.st_tier[cell] = (signed char)out.tier;
		}
		if (PRD && out.touch) out.touch[cell] = (unsigned char)stt.img;
	}
	if (out.geo.top) wave_geo_write(c, out.geo, cell, ok);
	if (out.want_nbrs) {
		int off = 0;
		if (lane == 0 && nf > 0) off = atomicAdd(out.arena_top, nf);
		off = __builtin_amdgcn_readfirstlane(off);
		if (off + nf > out.arena_cap) {
			if (lane == 0) { out.nbr_start[cell] = -1; atomicAdd(out.fail_count, 1); }
		} else {
			if (lane == 0) out.nbr_start[cell] = off;
			eachk<K>([&](auto hc) __attribute__((always_inline)) {
				constexpr int h = decltype(hc)::value;
				if (64 * h + lane < nf) out.nbr[off + 64 * h + lane] = c.nb_lane(h);
			});
		}
	}
}

}  // namespace vgpu
