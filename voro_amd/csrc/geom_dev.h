// Cell tables kept for the polyhedron statistics (VGPU_F_GEOMETRY), and the
// statistics computed from them.
//
// A tier that finishes a cell copies its final tables -- the doubled vertex
// coordinates (pts), the vertex orders (nu), each edge's end vertex (ed[i][j])
// and back pointer (ed[i][nu[i]+j]) -- into a bump-allocated arena.  One
// thread per cell then walks the faces exactly as the reference's functions
// do (src/cell.cc: face_areas 1791-1818, surface_area 1822-1849,
// total_edge_distance 1921-1934, face_perimeters 2251-2278, face_vertices
// 2282-2305, face_orders 2309-2328, normals 2087-2160, vertex_orders
// 2191-2194, vertices 2208-2216, max_radius_squared 1907-1917,
// number_of_edges 2468-2473), with the same expressions in the same order, so
// every value is the reference's bit for bit.  Every face-based statistic
// enumerates the faces in the order the reference's loops do (vertex i = 1
// .. p-1, edge j = 0 .. nu[i]-1, the first unmarked edge opens a face), the
// order of the neighbour ids, so per-face outputs share the neighbour CSR.
#pragma once
#include "vgpu_common.h"

#include <math.h>

namespace vgpu {

struct GeoSink {
	int *top = nullptr;          // [0] vertices used, [1] edge slots used, [2] cells that did not fit
	int vcap = 0, ecap = 0;
	int *vstart = nullptr;       // per cell: first vertex, -1 if it did not fit
	int *np = nullptr;           // per cell: vertex count (0: no cell)
	int *estart = nullptr;       // per cell: first edge slot
	int *esz = nullptr;          // per cell: edge slots (the sum of the vertex orders)
	double *xyz = nullptr;       // per vertex: the cell's doubled coordinates
	unsigned char *nu = nullptr; // per vertex: order
	int *eoff = nullptr;         // per vertex: first edge slot
	short *ed = nullptr;         // per edge slot: end vertex (cell-local)
	unsigned char *rel = nullptr;// per edge slot: back pointer
};

// Loop-order outputs of the statistics pass (any array may be null).
struct GeoStats {
	const int *status = nullptr;     // ST_EXISTS bit per cell
	const int *face_off = nullptr;   // per cell: first face (= the neighbour CSR offsets)
	const int *vert_off = nullptr;   // per cell: first output vertex
	const int *edge_off = nullptr;   // per cell: first output edge slot
	const int *fv_off = nullptr;     // per cell: first face-vertex entry
	double tol = 0;
	double *vertices = nullptr;      // 3 per vertex, local coordinates (vertices())
	int *orders = nullptr;           // per vertex
	int *edges = nullptr;            // per edge slot, cell-local end vertex
	int *edge_rel = nullptr;         // per edge slot, back pointer
	double *face_area = nullptr;     // per face
	double *face_perim = nullptr;    // per face
	double *normal = nullptr;        // 3 per face
	int *face_order = nullptr;       // per face
	int *face_verts = nullptr;       // per face: count, then the vertices (face_vertices())
	double *surface = nullptr;       // per cell
	double *edge_dist = nullptr;     // per cell
	int *n_edges = nullptr;          // per cell
	double *maxr2 = nullptr;         // per cell
};

// The statistics of one cell from its tables (edges are marked in place, as
// the reference negates them, and restored at the end).
VG_FN void cell_geometry_stats(const GeoSink &g, const GeoStats &o, int cell) {
	const int vs = g.vstart[cell], p = g.np[cell];
	if (!(o.status[cell] & ST_EXISTS) || p <= 0 || vs < 0) {
		if (o.surface) o.surface[cell] = 0;
		if (o.edge_dist) o.edge_dist[cell] = 0;
		if (o.n_edges) o.n_edges[cell] = 0;
		if (o.maxr2) o.maxr2[cell] = 0;
		return;
	}
	const double *pts = g.xyz + 3 * (size_t)vs;
	const unsigned char *nu = g.nu + vs;
	const int *eo = g.eoff + vs;
	short *ed = g.ed;
	const unsigned char *rel = g.rel;
	auto X = [&](int i) { return pts[3 * i]; };
	auto Y = [&](int i) { return pts[3 * i + 1]; };
	auto Z = [&](int i) { return pts[3 * i + 2]; };
	auto cyc = [&](int a, int v) { return a == nu[v] - 1 ? 0 : a + 1; };

	// vertices, orders, edges; max_radius_squared; number_of_edges
	int esum = 0;
	{
		const int vo = o.vert_off ? o.vert_off[cell] : 0;
		const int eb = o.edge_off ? o.edge_off[cell] : 0;
		double r = X(0) * X(0) + Y(0) * Y(0) + Z(0) * Z(0), s;
		for (int i = 0; i < p; i++) {
			if (o.vertices) {
				o.vertices[3 * (size_t)(vo + i)] = X(i) * 0.5;
				o.vertices[3 * (size_t)(vo + i) + 1] = Y(i) * 0.5;
				o.vertices[3 * (size_t)(vo + i) + 2] = Z(i) * 0.5;
			}
			if (o.orders) o.orders[vo + i] = nu[i];
			for (int j = 0; j < nu[i]; j++) {
				if (o.edges) o.edges[eb + esum + j] = ed[eo[i] + j];
				if (o.edge_rel) o.edge_rel[eb + esum + j] = rel[eo[i] + j];
			}
			esum += nu[i];
			if (i > 0) {
				s = X(i) * X(i);
				s += Y(i) * Y(i);
				s += Z(i) * Z(i);
				if (s > r) r = s;
			}
		}
		if (o.maxr2) o.maxr2[cell] = r;
		if (o.n_edges) o.n_edges[cell] = esum >> 1;
	}
	// total_edge_distance
	{
		double dis = 0, dx, dy, dz;
		for (int i = 0; i < p - 1; i++)
			for (int j = 0; j < nu[i]; j++) {
				const int k = ed[eo[i] + j];
				if (k > i) {
					dx = X(k) - X(i);
					dy = Y(k) - Y(i);
					dz = Z(k) - Z(i);
					dis += sqrt(dx * dx + dy * dy + dz * dz);
				}
			}
		if (o.edge_dist) o.edge_dist[cell] = 0.5 * dis;
	}
	// one walk per face: order, vertices, area (and the surface sum),
	// perimeter, normal
	double surf = 0;
	int f = o.face_off ? o.face_off[cell] : 0;
	int fv = o.fv_off ? o.fv_off[cell] : 0;
	const double tol = o.tol;
	for (int i = 1; i < p; i++)
		for (int j = 0; j < nu[i]; j++) {
			int k = ed[eo[i] + j];
			if (k < 0) continue;
			double area = 0, perim;
			double ux, uy, uz, vx, vy, vz, wx, wy, wz;
			// first edge i -> k
			ed[eo[i] + j] = (short)(-1 - k);
			int l = cyc(rel[eo[i] + j], k);
			{
				const double dx = X(k) - X(i), dy = Y(k) - Y(i), dz = Z(k) - Z(i);
				perim = sqrt(dx * dx + dy * dy + dz * dz);
			}
			int q = 1;
			const int fv0 = fv;
			if (o.face_verts) o.face_verts[fv0 + 1] = i;
			// normal state: first edge (after the first) longer than tol, then
			// the first product with a later edge above tol
			bool have_u = false, have_n = false;
			double nux = 0, nuy = 0, nuz = 0, nx = 0, ny = 0, nz = 0;
			for (;;) {
				q++;
				if (o.face_verts) o.face_verts[fv0 + q] = k;
				const int m = ed[eo[k] + l];
				ed[eo[k] + l] = (short)(-1 - m);
				// edge k -> m
				{
					const double dx = X(m) - X(k), dy = Y(m) - Y(k), dz = Z(m) - Z(k);
					perim += sqrt(dx * dx + dy * dy + dz * dz);
					if (!have_n) {
						if (!have_u) {
							if (dx * dx + dy * dy + dz * dz > tol) { have_u = true; nux = dx; nuy = dy; nuz = dz; }
						} else {
							const double px = nuz * dy - nuy * dz, py = nux * dz - nuz * dx, pz = nuy * dx - nux * dy;
							double wmag = px * px + py * py + pz * pz;
							if (wmag > tol) {
								wmag = 1 / sqrt(wmag);
								nx = px * wmag; ny = py * wmag; nz = pz * wmag;
								have_n = true;
							}
						}
					}
				}
				// triangle (i, k, m) unless m closes the face
				if (m != i) {
					ux = X(k) - X(i);
					uy = Y(k) - Y(i);
					uz = Z(k) - Z(i);
					vx = X(m) - X(i);
					vy = Y(m) - Y(i);
					vz = Z(m) - Z(i);
					wx = uy * vz - uz * vy;
					wy = uz * vx - ux * vz;
					wz = ux * vy - uy * vx;
					const double t = sqrt(wx * wx + wy * wy + wz * wz);
					area += t;
					surf += t;
				}
				const int n = cyc(rel[eo[k] + l], m);
				k = m;
				l = n;
				if (k == i) break;
			}
			// q: the vertices of the face, i and the q - 1 walked to
			if (o.face_verts) o.face_verts[fv0] = q;
			fv += q + 1;
			if (o.face_order) o.face_order[f] = q;
			if (o.face_area) o.face_area[f] = 0.125 * area;
			if (o.face_perim) o.face_perim[f] = 0.5 * perim;
			if (o.normal) {
				o.normal[3 * (size_t)f] = nx;
				o.normal[3 * (size_t)f + 1] = ny;
				o.normal[3 * (size_t)f + 2] = nz;
			}
			f++;
		}
	if (o.surface) o.surface[cell] = 0.125 * surf;
	// reset_edges, src/cell.cc:2025-2031
	for (int i = 0; i < p; i++)
		for (int j = 0; j < nu[i]; j++) {
			const int k = ed[eo[i] + j];
			if (k < 0) ed[eo[i] + j] = (short)(-1 - k);
		}
}

}  // namespace vgpu
