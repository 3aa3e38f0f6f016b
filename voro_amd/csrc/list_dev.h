// Gathering of candidate lists for the learned path (see compute_dev.h,
// "Candidate lists").  One cell's list follows the reference's candidate
// order (src/v_compute.cc:328-431): one record per worklist step 0 -> 1 ...
// G-1 -> G, then the end record at depth G.
#pragma once
#include "compute_dev.h"

namespace vgpu {

// The number of records of a list of depth G: 0 for G >= 64 (no list: the
// cell takes the plain path).
VG_INL int list_length(int G) { return G >= 64 ? 0 : G + 1; }

// Writes the list of stored particle `cell` (block ijk of view v) at depth
// G, record k at L[k * stride]; returns its length.  PRD: a periodic
// container's view (unit-cell start, image regions; v.fv must be null here,
// the list kernel records the visits).
template<bool PRD = false>
VG_FN int gather_cell_list(const RectView &v, int ijk, int cell, int G, ListRec *L, int stride) {
	if (G >= 64) return 0;
	const int b0 = v.bstart[ijk], s = cell - v.cstart[ijk];
	const dbl4 me = v.P[b0 + s];
	const double x = me.x, y = me.y, z = me.z;
	const int ci = ijk % v.nx, cj = (ijk / v.nx) % v.gy, ck = ijk / v.nxy;
	int i, j, k, disp = 0;
	if (PRD) { i = v.nx; j = v.ey; k = v.ez; }
	else {
		i = v.xper ? v.nx : ci; j = v.yper ? v.ny : cj; k = v.zper ? v.nz : ck;
		disp = ijk - i - v.nx * (j + v.ny * k);
	}
	// worklist selection (src/v_compute.cc:328-362)
	double fx, fy, fz;
	if (PRD) { fx = x - v.boxx * ci; fy = y - v.boxy * (cj - v.ey); fz = z - v.boxz * (ck - v.ez); }
	else { fx = x - v.ax - v.boxx * ci; fy = y - v.ay - v.boxy * cj; fz = z - v.az - v.boxz * ck; }
	double gxs, gys, gzs;
	unsigned m1, m2;
	int di = int(fx * v.xsp * VGPU_WL_FGRID), dj = int(fy * v.ysp * VGPU_WL_FGRID), dk = int(fz * v.zsp * VGPU_WL_FGRID);
	if (di >= VGPU_WL_HGRID) {
		gxs = fx;
		m1 = 127 + (3 << 21); m2 = 1 + (1 << 21); di = VGPU_WL_FGRID - 1 - di; if (di < 0) di = 0;
	} else { m1 = m2 = 0; gxs = v.boxx - fx; }
	if (dj >= VGPU_WL_HGRID) {
		gys = fy;
		m1 |= (127 << 7) + (3 << 24); m2 |= (1 << 7) + (1 << 24); dj = VGPU_WL_FGRID - 1 - dj; if (dj < 0) dj = 0;
	} else gys = v.boxy - fy;
	if (dk >= VGPU_WL_HGRID) {
		gzs = fz;
		m1 |= (127 << 14) + (3 << 27); m2 |= (1 << 14) + (1 << 27); dk = VGPU_WL_FGRID - 1 - dk; if (dk < 0) dk = 0;
	} else gzs = v.boxz - fz;
	gxs *= gxs; gys *= gys; gzs *= gzs;
	const int wli = di + VGPU_WL_HGRID * (dj + VGPU_WL_HGRID * dk);
	const double *radp = v.mrad + wli * 64;
	const unsigned *e = v.wl + wli * 64;
	for (int g = 0; g < G; g++) {
		unsigned q = e[g + 1];
		q ^= m1; q += m2;
		int ddi = q & 127; ddi -= 64;
		int ddj = (q >> 7) & 127; ddj -= 64;
		int ddk = (q >> 14) & 127; ddk -= 64;
		const int ei = ddi + i, ej = ddj + j, ek = ddk + k;
		if (ei < 0 || ei >= v.hx || ej < 0 || ej >= v.hy || ek < 0 || ek >= v.hz) {
			L[(size_t)g * stride] = ListRec{radp[g], 0, 0, 0, LK_NOBLOCK};
			continue;
		}
		double qx = 0, qy = 0, qz = 0;
		int img = 0;
		const int rijk = PRD ? prd_region(v, ci, cj, ck, ei, ej, ek, qx, 0, img)
				     : rect_region(v, ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
		double cn, cf;
		int form;
		min_max_radius_pair(v, ddi, ddj, ddk, fx, fy, fz, gxs, gys, gzs, cn, cf, form);
		const int info = LK_BLOCK | img << 3 | list_code(qx) << 4 | list_code(qy) << 6 | list_code(qz) << 8;
		L[(size_t)g * stride] = ListRec{radp[g], cn, cf, rijk, info};
	}
	L[(size_t)G * stride] = ListRec{radp[G], 0, 0, 0, LK_END};
	return G + 1;
}

}  // namespace vgpu
