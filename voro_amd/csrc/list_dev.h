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

// The worklist of stored particle `cell` (block ijk of view v): its
// sub-block's entries and radii, the mirror masks and the block-face
// distances (src/v_compute.cc:328-362), from which each record follows.
struct ListCtx {
	double fx, fy, fz, gxs, gys, gzs;
	unsigned m1, m2;
	int i, j, k, ci, cj, ck, disp;
	const double *radp;
	const unsigned *e;
};
template<bool PRD>
VG_FN ListCtx list_ctx(const RectView &v, int ijk, int cell) {
	ListCtx x;
	const int b0 = v.bstart[ijk], s = cell - v.cstart[ijk];
	const dbl4 me = v.P[b0 + s];
	const double px = me.x, py = me.y, pz = me.z;
	x.ci = ijk % v.nx; x.cj = (ijk / v.nx) % v.gy; x.ck = ijk / v.nxy;
	x.disp = 0;
	if (PRD) { x.i = v.nx; x.j = v.ey; x.k = v.ez; }
	else {
		x.i = v.xper ? v.nx : x.ci; x.j = v.yper ? v.ny : x.cj; x.k = v.zper ? v.nz : x.ck;
		x.disp = ijk - x.i - v.nx * (x.j + v.ny * x.k);
	}
	if (PRD) { x.fx = px - v.boxx * x.ci; x.fy = py - v.boxy * (x.cj - v.ey); x.fz = pz - v.boxz * (x.ck - v.ez); }
	else { x.fx = px - v.ax - v.boxx * x.ci; x.fy = py - v.ay - v.boxy * x.cj; x.fz = pz - v.az - v.boxz * x.ck; }
	int di = int(x.fx * v.xsp * VGPU_WL_FGRID), dj = int(x.fy * v.ysp * VGPU_WL_FGRID), dk = int(x.fz * v.zsp * VGPU_WL_FGRID);
	if (di >= VGPU_WL_HGRID) {
		x.gxs = x.fx;
		x.m1 = 127 + (3 << 21); x.m2 = 1 + (1 << 21); di = VGPU_WL_FGRID - 1 - di; if (di < 0) di = 0;
	} else { x.m1 = x.m2 = 0; x.gxs = v.boxx - x.fx; }
	if (dj >= VGPU_WL_HGRID) {
		x.gys = x.fy;
		x.m1 |= (127 << 7) + (3 << 24); x.m2 |= (1 << 7) + (1 << 24); dj = VGPU_WL_FGRID - 1 - dj; if (dj < 0) dj = 0;
	} else x.gys = v.boxy - x.fy;
	if (dk >= VGPU_WL_HGRID) {
		x.gzs = x.fz;
		x.m1 |= (127 << 14) + (3 << 27); x.m2 |= (1 << 14) + (1 << 27); dk = VGPU_WL_FGRID - 1 - dk; if (dk < 0) dk = 0;
	} else x.gzs = v.boxz - x.fz;
	x.gxs *= x.gxs; x.gys *= x.gys; x.gzs *= x.gzs;
	const int wli = di + VGPU_WL_HGRID * (dj + VGPU_WL_HGRID * dk);
	x.radp = v.mrad + wli * 64;
	x.e = v.wl + wli * 64;
	return x;
}

// Record g of the list: the step g -> g+1 (g < G), the end record (g = G).
// Records are independent of one another.
template<bool PRD>
VG_FN ListRec list_record(const RectView &v, const ListCtx &x, int g, int G) {
	if (g == G) return ListRec{x.radp[G], 0, 0, 0, LK_END};
	unsigned q = x.e[g + 1];
	q ^= x.m1; q += x.m2;
	int ddi = q & 127; ddi -= 64;
	int ddj = (q >> 7) & 127; ddj -= 64;
	int ddk = (q >> 14) & 127; ddk -= 64;
	const int ei = ddi + x.i, ej = ddj + x.j, ek = ddk + x.k;
	if (ei < 0 || ei >= v.hx || ej < 0 || ej >= v.hy || ek < 0 || ek >= v.hz) return ListRec{x.radp[g], 0, 0, 0, LK_NOBLOCK};
	double qx = 0, qy = 0, qz = 0;
	int img = 0;
	const int rijk = PRD ? prd_region(v, x.ci, x.cj, x.ck, ei, ej, ek, qx, 0, img)
			     : rect_region(v, x.ci, x.cj, x.ck, ei, ej, ek, qx, qy, qz, x.disp);
	double cn, cf;
	int form;
	min_max_radius_pair(v, ddi, ddj, ddk, x.fx, x.fy, x.fz, x.gxs, x.gys, x.gzs, cn, cf, form);
	const int info = LK_BLOCK | img << 3 | list_code(qx) << 4 | list_code(qy) << 6 | list_code(qz) << 8;
	return ListRec{x.radp[g], cn, cf, rijk, info};
}

// Writes the list of stored particle `cell` (block ijk of view v) at depth
// G, record k at L[k * stride]; returns its length.  PRD: a periodic
// container's view (unit-cell start, image regions; v.fv must be null here,
// the list kernel records the visits).
template<bool PRD = false>
VG_FN int gather_cell_list(const RectView &v, int ijk, int cell, int G, ListRec *L, int stride) {
	if (G >= 64) return 0;
	const ListCtx x = list_ctx<PRD>(v, ijk, cell);
	for (int g = 0; g <= G; g++) L[(size_t)g * stride] = list_record<PRD>(v, x, g, G);
	return G + 1;
}

}  // namespace vgpu
