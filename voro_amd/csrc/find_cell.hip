// find_voronoi_cell for many query points at once: for each point, the
// particle whose Voronoi cell contains it (the nearest particle; in power
// distance for radical containers), for rectangular containers (with or
// without periodic directions).  One query per lane; the block search of a
// query keeps its mask window and queue in the lane's slab of a per-lane
// cell type (the order-3 tier's; then the wide and the huge tier's for the
// few searches that leave the smaller windows).
#include <hip/hip_runtime.h>

#include "batch_dev.h"
#include "vgpu_cells.h"

namespace vgpu {

namespace {

// step_int / step_div of src/common.hh (rounding toward minus infinity).
VG_INL int q_step_int(double a) { return a < 0 ? int(a) - 1 : int(a); }

// put_locate_block of a periodic container (src/container_prd.cc:189-217):
// the block of (x,y,z) in the nx*oy*oz grid and the position remapped into
// the primary domain; -1 for a coordinate whose block index would not fit an
// int.
VG_INL int locate_prd_point(const RectView &v, double &x, double &y, double &z) {
	const double lim = 1073741824.0;
	const double zs = z * v.zsp;
	if (!(fabs(zs) < lim)) return -1;
	int k = q_step_int(zs);
	if (k < 0 || k >= v.nz) {
		const int ak = step_div_i(k, v.nz);
		z -= ak * v.bz; y -= ak * v.byz; x -= ak * v.bxz; k -= ak * v.nz;
	}
	const double ys = y * v.ysp;
	if (!(fabs(ys) < lim)) return -1;
	int j = q_step_int(ys);
	if (j < 0 || j >= v.ny) {
		const int aj = step_div_i(j, v.ny);
		y -= aj * v.by; x -= aj * v.bxy; j -= aj * v.ny;
	}
	const double xs = x * v.xsp;
	if (!(fabs(xs) < lim)) return -1;
	int i = q_step_int(xs);
	if (i < 0 || i >= v.nx) {
		const int ai = step_div_i(i, v.nx);
		x -= ai * v.bx; i -= ai * v.nx;
	}
	return i + v.nx * (j + v.ey + v.oy * (k + v.ez));
}

struct Found {
	int b, l;           // storage block and index of the particle, b = -1: none yet
	int di, dj, dk;     // worklist displacement of that block
};

// scan_all, src/v_compute.cc:46-57: the particles of storage block b against
// (x,y,z), keeping the first strict minimum.
template<bool POLY>
VG_INL void scan_block(const RectView &v, int b, double x, double y, double z, int di, int dj, int dk, Found &w,
		       double &mrs) {
	const int b0 = v.bstart[b], n = v.bstart[b + 1] - b0;
	bool in_block = false;
	for (int l = 0; l < n; l++) {
		const dbl4 o = v.P[b0 + l];
		const double x1 = o.x - x, y1 = o.y - y, z1 = o.z - z;
		double rs = x1 * x1 + y1 * y1 + z1 * z1;
		if (POLY) rs = rs - o.w * o.w;
		if (rs < mrs) { mrs = rs; w.l = l; in_block = true; }
	}
	if (in_block) { w.b = b; w.di = di; w.dj = dj; w.dk = dk; }
}

// compute_min_radius, src/v_compute.cc:947-961.
template<bool POLY>
VG_INL bool min_radius_skip(const RectView &v, int di, int dj, int dk, double fx, double fy, double fz, double mrs) {
	double t, crs;
	if (di > 0) { t = di * v.boxx - fx; crs = t * t; }
	else if (di < 0) { t = (di + 1) * v.boxx - fx; crs = t * t; }
	else crs = 0;
	if (dj > 0) { t = dj * v.boxy - fy; crs += t * t; }
	else if (dj < 0) { t = (dj + 1) * v.boxy - fy; crs += t * t; }
	if (dk > 0) { t = dk * v.boxz - fz; crs += t * t; }
	else if (dk < 0) { t = (dk + 1) * v.boxz - fz; crs += t * t; }
	return crs > (POLY ? mrs + v.max_radius * v.max_radius : mrs);
}

// container::find_voronoi_cell (src/container.cc:237-262, container_poly's
// src/container.cc:274-297) around voro_compute::find_voronoi_cell
// (src/v_compute.cc:71-228); with PRD, container_periodic(_poly)::
// find_voronoi_cell (src/container_prd.cc:303-360), whose searches read the
// image blocks (their first visits are recorded in v.fv with keys
// qkey | visit, and *touched is set when one is read).  Returns 1 and the
// particle (its storage index, its position shifted by whole periods) if one
// was found, 0 if none (a point outside a non-periodic direction, or no
// particle), -1 if the block search left the mask window (retry with a wider
// one).
template<class CELL, bool POLY, bool PRD>
VG_FN int find_cell(CELL &c, const RectView &v, double x, double y, double z, unsigned stamp, int &idx, double &rx,
		    double &ry, double &rz, unsigned long long qkey = 0, bool *touched = nullptr) {
	// A non-finite or huge coordinate would make the int conversion
	// undefined (on x86-64 the reference's lands outside the grid): the
	// point is reported as not found.
	const double lim = 1073741824.0;
	int ai, aj, ak, ci, cj, ck, ijk0, i, j, k, disp = 0;
	double fx, fy, fz;
	if (PRD) {
		// container_periodic_base::remap, src/container_prd.cc:264-290: z,
		// then y, then x, each shift carrying the lattice's off-diagonal terms
		const double zs = z * v.zsp;
		if (!(fabs(zs) < lim)) return 0;
		ck = q_step_int(zs);
		if (ck < 0 || ck >= v.nz) {
			ak = step_div_i(ck, v.nz);
			z -= ak * v.bz; y -= ak * v.byz; x -= ak * v.bxz; ck -= ak * v.nz;
		} else ak = 0;
		const double ys = y * v.ysp;
		if (!(fabs(ys) < lim)) return 0;
		cj = q_step_int(ys);
		if (cj < 0 || cj >= v.ny) {
			aj = step_div_i(cj, v.ny);
			y -= aj * v.by; x -= aj * v.bxy; cj -= aj * v.ny;
		} else aj = 0;
		const double xs = x * v.xsp;
		if (!(fabs(xs) < lim)) return 0;
		ci = q_step_int(xs);
		if (ci < 0 || ci >= v.nx) {
			ai = step_div_i(ci, v.nx);
			x -= ai * v.bx; ci -= ai * v.nx;
		} else ai = 0;
		cj += v.ey; ck += v.ez;
		ijk0 = ci + v.nx * (cj + v.oy * ck);
		// initialize_search and frac_pos, src/container_prd.hh:143-157
		i = v.nx; j = v.ey; k = v.ez;
		fx = x - v.boxx * ci; fy = y - v.boxy * (cj - v.ey); fz = z - v.boxz * (ck - v.ez);
	} else {
		// remap, src/container.cc:203-224
		{
			const double a = (x - v.ax) * v.xsp, b = (y - v.ay) * v.ysp, c = (z - v.az) * v.zsp;
			if (!(fabs(a) < lim && fabs(b) < lim && fabs(c) < lim)) return 0;
		}
		ci = q_step_int((x - v.ax) * v.xsp);
		if (ci < 0 || ci >= v.nx) {
			if (!v.xper) return 0;
			ai = step_div_i(ci, v.nx); x -= ai * (v.bx - v.ax); ci -= ai * v.nx;
		} else ai = 0;
		cj = q_step_int((y - v.ay) * v.ysp);
		if (cj < 0 || cj >= v.ny) {
			if (!v.yper) return 0;
			aj = step_div_i(cj, v.ny); y -= aj * (v.by - v.ay); cj -= aj * v.ny;
		} else aj = 0;
		ck = q_step_int((z - v.az) * v.zsp);
		if (ck < 0 || ck >= v.nz) {
			if (!v.zper) return 0;
			ak = step_div_i(ck, v.nz); z -= ak * (v.bz - v.az); ck -= ak * v.nz;
		} else ak = 0;
		ijk0 = ci + v.nx * cj + v.nxy * ck;
		// initialize_search, src/container.hh:214-219
		i = v.xper ? v.nx : ci; j = v.yper ? v.ny : cj; k = v.zper ? v.nz : ck;
		disp = ijk0 - i - v.nx * (j + v.ny * k);
		fx = x - v.ax - v.boxx * ci; fy = y - v.ay - v.boxy * cj; fz = z - v.az - v.boxz * ck;
	}
	// region_index of the reference's container class
	unsigned long long key = qkey;
	int img = 0;
	auto region = [&](int ei, int ej, int ek, double &qx, double &qy, double &qz) {
		if (PRD) return prd_region(v, ci, cj, ck, ei, ej, ek, qx, key++, img);
		return rect_region(v, ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
	};

	Found w{-1, 0, 0, 0, 0};
	double mrs = 1.7976931348623157e308;     // large_number, src/config.hh:87
	scan_block<POLY>(v, ijk0, x, y, z, 0, 0, 0, w, mrs);

	int di = int(fx * v.xsp * VGPU_WL_FGRID), dj = int(fy * v.ysp * VGPU_WL_FGRID), dk = int(fz * v.zsp * VGPU_WL_FGRID);
	unsigned m1, m2;
	double mxs, mys, mzs;
	if (di >= VGPU_WL_HGRID) {
		mxs = v.boxx - fx;
		m1 = 127 + (3 << 21); m2 = 1 + (1 << 21); di = VGPU_WL_FGRID - 1 - di; if (di < 0) di = 0;
	} else { m1 = m2 = 0; mxs = fx; }
	if (dj >= VGPU_WL_HGRID) {
		mys = v.boxy - fy;
		m1 |= (127 << 7) + (3 << 24); m2 |= (1 << 7) + (1 << 24); dj = VGPU_WL_FGRID - 1 - dj; if (dj < 0) dj = 0;
	} else mys = fy;
	if (dk >= VGPU_WL_HGRID) {
		mzs = v.boxz - fz;
		m1 |= (127 << 14) + (3 << 27); m2 |= (1 << 14) + (1 << 27); dk = VGPU_WL_FGRID - 1 - dk; if (dk < 0) dk = 0;
	} else mzs = fz;
	const double r2 = POLY ? v.max_radius * v.max_radius : 0.0;
	bool done = false;
	{
		const double rs = POLY ? mrs + r2 : mrs;
		if (mxs * mxs > rs && mys * mys > rs && mzs * mzs > rs) done = true;
	}
	if (!done) {
		const int wli = di + VGPU_WL_HGRID * (dj + VGPU_WL_HGRID * dk);
		const double *radp = v.mrad + wli * 64;
		const unsigned *e = v.wl + wli * 64;
		double qx = 0, qy = 0, qz = 0;
		const int f = (int)e[0];
		int g = 0;
		// worklist blocks that need no mask
		do {
			if ((POLY ? mrs + r2 : mrs) < radp[g]) { done = true; break; }
			g++;
			unsigned q = e[g];
			q ^= m1; q += m2;
			const int ddi = (int)(q & 127) - 64, ddj = (int)((q >> 7) & 127) - 64, ddk = (int)((q >> 14) & 127) - 64;
			const int ei = ddi + i, ej = ddj + j, ek = ddk + k;
			if (ei < 0 || ei >= v.hx || ej < 0 || ej >= v.hy || ek < 0 || ek >= v.hz) continue;
			if (min_radius_skip<POLY>(v, ddi, ddj, ddk, fx, fy, fz, mrs)) continue;
			const int b = region(ei, ej, ek, qx, qy, qz);
			scan_block<POLY>(v, b, x - qx, y - qy, z - qz, ddi, ddj, ddk, w, mrs);
		} while (g < f);
		BlockSearch<CELL> bs(c, stamp, i, j, k);
		const unsigned b1 = 1u << 21, b2 = 1u << 22, b3 = 1u << 24, b4 = 1u << 25, b5 = 1u << 27, b6 = 1u << 28;
		// the rest of the worklist, marking blocks and queueing their outside
		// neighbours (scan_bits_mask_add, src/v_compute.cc:249-264)
		while (!done && g < 63) {
			if ((POLY ? mrs + r2 : mrs) < radp[g]) { done = true; break; }
			g++;
			unsigned q = e[g];
			q ^= m1; q += m2;
			const int ddi = (int)(q & 127) - 64, ddj = (int)((q >> 7) & 127) - 64, ddk = (int)((q >> 14) & 127) - 64;
			const int ei = ddi + i, ej = ddj + j, ek = ddk + k;
			if (ei < 0 || ei >= v.hx || ej < 0 || ej >= v.hy || ek < 0 || ek >= v.hz) continue;
			bs.mark(ei, ej, ek);
			if (min_radius_skip<POLY>(v, ddi, ddj, ddk, fx, fy, fz, mrs)) continue;
			const int b = region(ei, ej, ek, qx, qy, qz);
			scan_block<POLY>(v, b, x - qx, y - qy, z - qz, ddi, ddj, ddk, w, mrs);
			if ((q & b2) == b2) {
				if (ei > 0) bs.mark_push(ei - 1, ej, ek);
				if ((q & b1) == 0 && ei < v.hx - 1) bs.mark_push(ei + 1, ej, ek);
			} else if ((q & b1) == b1 && ei < v.hx - 1) bs.mark_push(ei + 1, ej, ek);
			if ((q & b4) == b4) {
				if (ej > 0) bs.mark_push(ei, ej - 1, ek);
				if ((q & b3) == 0 && ej < v.hy - 1) bs.mark_push(ei, ej + 1, ek);
			} else if ((q & b3) == b3 && ej < v.hy - 1) bs.mark_push(ei, ej + 1, ek);
			if ((q & b6) == b6) {
				if (ek > 0) bs.mark_push(ei, ej, ek - 1);
				if ((q & b5) == 0 && ek < v.hz - 1) bs.mark_push(ei, ej, ek + 1);
			} else if ((q & b5) == b5 && ek < v.hz - 1) bs.mark_push(ei, ej, ek + 1);
			if (c.ovf) return -1;
		}
		if (!done && (POLY ? mrs + r2 : mrs) < radp[g]) done = true;
		// breadth-first over the queued blocks (src/v_compute.cc:211-227)
		int ei, ej, ek;
		while (!done && bs.pop(ei, ej, ek)) {
			const int ddi = ei - i, ddj = ej - j, ddk = ek - k;
			if (min_radius_skip<POLY>(v, ddi, ddj, ddk, fx, fy, fz, mrs)) continue;
			const int b = region(ei, ej, ek, qx, qy, qz);
			scan_block<POLY>(v, b, x - qx, y - qy, z - qz, ddi, ddj, ddk, w, mrs);
			// add_to_mask, src/v_compute.cc:236-243
			if (ek > 0 && !bs.marked(ei, ej, ek - 1)) bs.mark_push(ei, ej, ek - 1);
			if (ej > 0 && !bs.marked(ei, ej - 1, ek)) bs.mark_push(ei, ej - 1, ek);
			if (ei > 0 && !bs.marked(ei - 1, ej, ek)) bs.mark_push(ei - 1, ej, ek);
			if (ei < v.hx - 1 && !bs.marked(ei + 1, ej, ek)) bs.mark_push(ei + 1, ej, ek);
			if (ej < v.hy - 1 && !bs.marked(ei, ej + 1, ek)) bs.mark_push(ei, ej + 1, ek);
			if (ek < v.hz - 1 && !bs.marked(ei, ej, ek + 1)) bs.mark_push(ei, ej, ek + 1);
			if (c.ovf) return -1;
		}
		if (c.ovf) return -1;
	}
	if (touched && img) *touched = true;
	if (w.b < 0) return 0;
	if (PRD) {
		// src/container_prd.cc:316-322: the image blocks hold positions
		// shifted in y and z; the x period and the point's own image are added
		ci += w.di;
		if (ci < 0 || ci >= v.nx) ai += step_div_i(ci, v.nx);
		idx = v.bstart[w.b] + w.l;
		const dbl4 o = v.P[idx];
		rx = o.x + ak * v.bxz + aj * v.bxy + ai * v.bx;
		ry = o.y + ak * v.byz + aj * v.by;
		rz = o.z + ak * v.bz;
		return 1;
	}
	// the particle's position, shifted by whole periods
	if (v.xper) { ci += w.di; if (ci < 0 || ci >= v.nx) ai += step_div_i(ci, v.nx); }
	if (v.yper) { cj += w.dj; if (cj < 0 || cj >= v.ny) aj += step_div_i(cj, v.ny); }
	if (v.zper) { ck += w.dk; if (ck < 0 || ck >= v.nz) ak += step_div_i(ck, v.nz); }
	idx = v.bstart[w.b] + w.l;
	const dbl4 o = v.P[idx];
	rx = o.x + ai * (v.bx - v.ax);
	ry = o.y + aj * (v.by - v.ay);
	rz = o.z + ak * (v.bz - v.az);
	return 1;
}

// Persistent kernel, one query per lane, 64 queries per counter increment,
// over every query or over a device list of them (its count at *list_count).
// found[q]: 1 / 0 as find_cell returns; a query whose search left the window
// goes to the retry list (or, without one, gets -1 and is counted at
// *retry_count).
//
// Periodic containers (PRD): touch[q] is set to 1 when query q read an image
// block, and the first visit of every block is recorded in v.fv with the key
// (q << 24 | visit number), the order in which the reference's queries,
// made one after another, would create the image blocks.
template<class T_, bool POLY, bool PRD>
__global__ void __launch_bounds__(64, 2)
k_find_cells(RectView v, int n, const double *__restrict__ xyz, const int *__restrict__ list,
	     const int *__restrict__ list_count, int *__restrict__ found, int *__restrict__ pid,
	     double *__restrict__ pos, int *__restrict__ retry, int *__restrict__ retry_count, int *__restrict__ counter,
	     char *__restrict__ ws, size_t slab, unsigned stamp_base, unsigned char *__restrict__ touch) {
	typedef typename TierCell<T_>::type CELL;
	const int lane = threadIdx.x;
	CELL c;
	c.setup(ws + (size_t)blockIdx.x * slab, lane, v.max_len_sq);
	const int total = list_count ? *list_count : n;
	for (;;) {
		int base = 0;
		if (lane == 0) base = atomicAdd(counter, 64);
		base = __builtin_amdgcn_readfirstlane(base);
		if (base >= total) break;
		const int t = base + lane;
		if (t >= total) continue;
		const int qi = list ? list[t] : t;
		c.ovf = false;
		int idx = 0;
		double rx = 0, ry = 0, rz = 0;
		bool touched = false;
		const int r = find_cell<CELL, POLY, PRD>(c, v, xyz[3 * qi], xyz[3 * qi + 1], xyz[3 * qi + 2],
							 stamp_base + (unsigned)t + 1u, idx, rx, ry, rz,
							 (unsigned long long)qi << 24, &touched);
		if (PRD && touched) touch[qi] = 1;
		if (r < 0) {
			if (retry) {
				retry[atomicAdd(retry_count, 1)] = qi;
				continue;
			}
			if (retry_count) atomicAdd(retry_count, 1);   // the last tier counts failures
		}
		found[qi] = r;
		if (r > 0) {
			pid[qi] = v.pid[idx];
			pos[3 * qi] = rx; pos[3 * qi + 1] = ry; pos[3 * qi + 2] = rz;
		}
	}
}


// compute_ghost_cell (src/container.hh:488-498, container_poly's
// src/container.hh:710-721) for n points, one per lane, with the general
// per-lane cell (up to 1,024 vertices): the point is located as put() would
// (put_remap, dropped outside a non-periodic direction: exists 0), then its
// cell is computed as that of a particle stored last in its block, with the
// container's maximum radius raised to the point's own (radical).  exists:
// 1 / 0 as the reference's return value, -1 if the cell could not be
// computed here (capacity, or a search that comes back to the point's own
// block); nbr holds up to max_nbr neighbour ids per point.
//
// PRD: container_periodic(_poly)::compute_ghost_cell (src/container_prd.hh:
// 414-420,635-645): the point is located and remapped as put_locate_block
// does; the searches read the block arrays of v (the image blocks the
// reference holds before the call) and give up on the blocks v.img_ok
// excludes.
template<class T_, bool POLY, bool PRD>
__global__ void __launch_bounds__(64, 2)
k_ghost_cells(RectView v, RectGeom geo, int n, const double *__restrict__ xyz, const double *__restrict__ rad,
	      const int *__restrict__ list, const int *__restrict__ list_count, int *__restrict__ exists,
	      double *__restrict__ vol, double *__restrict__ cent, int *__restrict__ nfaces, int *__restrict__ nbr,
	      int max_nbr, int *__restrict__ counter, int *__restrict__ retry, int *__restrict__ failed,
	      char *__restrict__ ws, size_t slab, unsigned stamp_base) {
	typedef typename TierCell<T_>::type CELL;
	const int lane = threadIdx.x;
	CELL c;
	c.setup(ws + (size_t)blockIdx.x * slab, lane, v.max_len_sq);
	const int total = list_count ? *list_count : n;
	for (;;) {
		int base = 0;
		if (lane == 0) base = atomicAdd(counter, 64);
		base = __builtin_amdgcn_readfirstlane(base);
		if (base >= total) break;
		const int tt = base + lane;
		if (tt >= total) continue;
		const int t = list ? list[tt] : tt;
		double x = xyz[3 * t], y = xyz[3 * t + 1], z = xyz[3 * t + 2];
		const int ijk = PRD ? locate_prd_point(v, x, y, z) : locate_rect(geo, x, y, z);
		if (ijk < 0) {
			exists[t] = 0; nfaces[t] = 0; vol[t] = 0;
			cent[3 * t] = cent[3 * t + 1] = cent[3 * t + 2] = 0;
			continue;
		}
		const double r = POLY ? rad[t] : 0.5;
		const dbl4 gp{x, y, z, r};
		const double maxr = POLY && r > v.max_radius ? r : v.max_radius;
		const int ci = ijk % v.nx, cj = (ijk / v.nx) % v.gy, ck = ijk / v.nxy;
		const int s = v.bstart[ijk + 1] - v.bstart[ijk];
		CellStats stt;
		c.ovf = false; c.fatal = false;
		const bool ok = compute_cell_rect<CELL, POLY, PRD>(c, v, ijk, s, ci, cj, ck, stamp_base + (unsigned)t + 1u,
								   stt, 0, &gp, maxr);
		double cv = 0, cx = 0, cy = 0, cz = 0;
		int nf = 0;
		if (!c.ovf && ok) {
			SlabSink<CELL> sk{&c};
			c.outputs(cv, cx, cy, cz, nf, sk);
			if (nf > (int)CELL::L::NBCAP || nf > max_nbr) c.ovf = true;
		}
		if (c.ovf) {
			exists[t] = -1; nfaces[t] = 0; vol[t] = 0;
			cent[3 * t] = cent[3 * t + 1] = cent[3 * t + 2] = 0;
			// the next tier's list, or (last tier) a failure
			if (retry) retry[atomicAdd(failed, 1)] = t;
			else atomicAdd(failed, 1);
			continue;
		}
		exists[t] = ok ? 1 : 0;
		vol[t] = cv;
		cent[3 * t] = cx; cent[3 * t + 1] = cy; cent[3 * t + 2] = cz;
		nfaces[t] = nf;
		for (int f = 0; f < nf; f++) nbr[(size_t)t * max_nbr + f] = c.template A<int>(CELL::L::o_nbs, f);
	}
}

}  // namespace

int launch_find_cells(int tier, bool poly, bool prd, int waves, hipStream_t s, const RectView &v, int n,
		      const double *xyz, const int *list, const int *list_count, int *found, int *pid, double *pos,
		      int *retry, int *retry_count, int *counter, char *ws, size_t slab, unsigned stamp_base,
		      unsigned char *touch) {
#define VGPU_FIND(T, P, R)                                                                                         \
	hipLaunchKernelGGL((k_find_cells<T, P, R>), dim3(waves), dim3(64), 0, s, v, n, xyz, list, list_count, found,    \
			   pid, pos, retry, retry_count, counter, ws, slab, stamp_base, touch)
#define VGPU_FIND_T(T)                                                                                             \
	do {                                                                                                       \
		if (prd) { if (poly) VGPU_FIND(T, true, true); else VGPU_FIND(T, false, true); }                   \
		else { if (poly) VGPU_FIND(T, true, false); else VGPU_FIND(T, false, false); }                     \
	} while (0)
	if (tier == 0) VGPU_FIND_T(Tier3);
	else if (tier == 1) VGPU_FIND_T(TierWide);
	else VGPU_FIND_T(TierHuge);
#undef VGPU_FIND_T
#undef VGPU_FIND
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

int launch_ghost_cells(int tier, bool poly, bool prd, int waves, hipStream_t s, const RectView &v,
		       const RectGeom &geo, int n, const double *xyz, const double *rad, const int *list,
		       const int *list_count, int *exists, double *vol, double *cent, int *nfaces, int *nbr, int max_nbr,
		       int *counter, int *retry, int *failed, char *ws, size_t slab, unsigned stamp_base) {
#define VGPU_GHOST(T, P, R)                                                                                       \
	hipLaunchKernelGGL((k_ghost_cells<T, P, R>), dim3(waves), dim3(64), 0, s, v, geo, n, xyz, rad, list,          \
			   list_count, exists, vol, cent, nfaces, nbr, max_nbr, counter, retry, failed, ws, slab,      \
			   stamp_base)
#define VGPU_GHOST_T(T)                                                                                           \
	do {                                                                                                      \
		if (prd) { if (poly) VGPU_GHOST(T, true, true); else VGPU_GHOST(T, false, true); }               \
		else { if (poly) VGPU_GHOST(T, true, false); else VGPU_GHOST(T, false, false); }                 \
	} while (0)
	if (tier == 0) VGPU_GHOST_T(TierWide);
	else VGPU_GHOST_T(TierHuge);
#undef VGPU_GHOST_T
#undef VGPU_GHOST
	return hipGetLastError() == hipSuccess ? 0 : 6;
}

}  // namespace vgpu
