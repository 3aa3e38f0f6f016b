// Host-side grid set-up shared by the C-ABI library: worklist radii.
#pragma once
#include "vgpu_common.h"
#include "worklist_table.h"

namespace vgpu {

// Squared distance from sub-region [xlo,xhi]x[ylo,yhi]x[zlo,zhi] of a block
// to block (ti,tj,tk), folded into minr (src/v_base.cc:81-94).
VG_FN void wl_block_min(double &minr, double boxx, double boxy, double boxz, double xlo, double xhi, double ylo,
			double yhi, double zlo, double zhi, int ti, int tj, int tk) {
	double radsq, t;
	if (ti > 0) { t = boxx * ti - xhi; radsq = t * t; }
	else if (ti < 0) { t = xlo - boxx * (1 + ti); radsq = t * t; }
	else radsq = 0;
	if (tj > 0) { t = boxy * tj - yhi; radsq += t * t; }
	else if (tj < 0) { t = ylo - boxy * (1 + tj); radsq += t * t; }
	if (tk > 0) { t = boxz * tk - zhi; radsq += t * t; }
	else if (tk < 0) { t = zlo - boxz * (1 + tk); radsq += t * t; }
	if (radsq < minr) minr = radsq;
}

// The 64 worklist radii of sub-region w = lx + 4 (ly + 4 lz): radp[g] is
// the squared distance from the sub-region to the nearest block not among
// worklist entries 1..g (src/v_base.cc:23-69: the shell outside the list
// first, then backwards).  The sub-region bounds are accumulated step by
// step, as the reference's loops do.
VG_FN void worklist_radii_region(double boxx, double boxy, double boxz, int w, const unsigned *wl, double *radp) {
	const unsigned b1 = 1u << 21, b2 = 1u << 22, b3 = 1u << 24, b4 = 1u << 25, b5 = 1u << 27, b6 = 1u << 28;
	const double xs = boxx / VGPU_WL_FGRID, ys = boxy / VGPU_WL_FGRID, zs = boxz / VGPU_WL_FGRID;
	const int lx = w % VGPU_WL_HGRID, ly = (w / VGPU_WL_HGRID) % VGPU_WL_HGRID, lz = w / (VGPU_WL_HGRID * VGPU_WL_HGRID);
	double xlo = 0, xhi = xs, ylo = 0, yhi = ys, zlo = 0, zhi = zs;
	for (int t = 0; t < lx; t++) { xlo = xhi; xhi += xs; }
	for (int t = 0; t < ly; t++) { ylo = yhi; yhi += ys; }
	for (int t = 0; t < lz; t++) { zlo = zhi; zhi += zs; }
	const unsigned *e = wl + 64 * w;
	double minr = 1.7976931348623157e308;   // std::numeric_limits<double>::max()
	int q;
	for (q = e[0] + 1; q < 64; q++) {
		unsigned f = e[q];
		int i = (f & 127) - 64, j = (f >> 7 & 127) - 64, k = (f >> 14 & 127) - 64;
		if ((f & b2) == b2) {
			wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i - 1, j, k);
			if ((f & b1) == 0) wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i + 1, j, k);
		} else if ((f & b1) == b1) wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i + 1, j, k);
		if ((f & b4) == b4) {
			wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i, j - 1, k);
			if ((f & b3) == 0) wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i, j + 1, k);
		} else if ((f & b3) == b3) wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i, j + 1, k);
		if ((f & b6) == b6) {
			wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k - 1);
			if ((f & b5) == 0) wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k + 1);
		} else if ((f & b5) == b5) wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k + 1);
	}
	q--;
	while (q > 0) {
		radp[q] = minr;
		unsigned f = e[q];
		int i = (f & 127) - 64, j = (f >> 7 & 127) - 64, k = (f >> 14 & 127) - 64;
		wl_block_min(minr, boxx, boxy, boxz, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k);
		q--;
	}
	radp[0] = minr;
}

// mrad[w*64+g] for all 64 sub-regions, for blocks of size boxx x boxy x boxz.
static inline void worklist_radii(double boxx, double boxy, double boxz, double *mrad) {
	for (int w = 0; w < VGPU_WL_HGRID * VGPU_WL_HGRID * VGPU_WL_HGRID; w++)
		worklist_radii_region(boxx, boxy, boxz, w, vgpu_worklist_table, mrad + 64 * w);
}

}  // namespace vgpu
