// Test-infrastructure driver around the unmodified Voro++ sources.
//
// This file is compiled together with the reference library's own source
// files (taken where they lie under the reference checkout, never copied into
// this repository) into oracle/_ref/libvororef.so.  It is a checker and a CPU
// baseline only: tests, the smoke check and the benchmark's CPU leg load it;
// the product library never does.
//
// It exposes a plain C interface so that Python (ctypes) can hand it seeded
// inputs and read back, in c_loop_all order, exactly what the GPU engine
// returns: particle id, block, existence flag, volume, centroid, face count
// and neighbour IDs.  The calls it makes are the reference's public API:
//   container / container_poly            src/container.hh:288-728
//   container_periodic / _poly            src/container_prd.hh:213-651
//   put()                                 src/container.cc:90-115
//   c_loop_all / c_loop_all_periodic      src/c_loops.hh:165-209,325-393
//   compute_cell(c,vl)                    src/container.hh:463-465
//   voronoicell_neighbor outputs          src/cell.cc:1584,1861,2168,2660
//   compute_all_cells()                   src/container.cc:438-443
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "voro++.hh"

using namespace voro;

extern "C" {

struct ref_spec {
	int kind;            // 0 container, 1 container_poly, 2 container_periodic, 3 container_periodic_poly
	double geom[6];      // rectangular: ax,bx,ay,by,az,bz ; periodic: bx,bxy,by,bxz,byz,bz
	int grid[3];         // nx,ny,nz
	int periodic[3];     // rectangular containers only
	int init_mem;
	int n_spheres;       // sphere walls (rectangular containers only)
	const double *spheres;  // n_spheres x {xc,yc,zc,rc,w_id}
	int n_walls;         // further walls, added after the spheres
	const double *walls;    // n_walls x {type, w_id, constructor arguments (4 or 7)}: 0 sphere,
	                        // 1 plane, 2 cylinder, 3 cone (src/wall.hh:27-107)
};

struct ref_out {
	int *id;             // [n] particle id in loop order
	int *ijk;            // [n] block index
	int *q;              // [n] index in block
	int *exists;         // [n] compute_cell() return value
	double *vol;         // [n]
	double *cent;        // [3n] centroid relative to the particle
	int *nfaces;         // [n]
	long long *nbr_off;  // [n+1]
	int *nbr;            // [nbr_cap]
	long long nbr_cap;
	double *pos;         // [3n] stored (possibly remapped) position, or null
};

// Polyhedron statistics of every cell in loop order, each from the
// reference's own voronoicell_neighbor functions (src/cell.cc), plus the raw
// tables (pts, nu, ed with back pointers) they are computed from.  The
// *_off arrays have n+1 entries; an array is written only up to its cap,
// and the offsets always count everything (so a caller can retry).
struct ref_geo {
	long long *vert_off;   // vertices per cell (vertices(), vertex_orders())
	double *vertices;      // 3 per vertex, local coordinates
	int *orders;
	double *pts;           // 3 per vertex, the raw (doubled) coordinates
	long long vcap;
	long long *edge_off;   // edge slots per cell (sum of the orders)
	int *edges;            // ed[i][j]
	int *edge_rel;         // ed[i][nu[i]+j]
	long long ecap;
	long long *face_off;   // faces per cell
	double *face_area;
	double *face_perim;
	double *normal;        // 3 per face
	int *face_order;
	long long fcap;
	long long *fv_off;     // face_vertices() entries per cell
	int *face_verts;
	long long fvcap;
	double *surface;       // per cell
	double *edge_dist;
	int *n_edges;
	double *maxr2;
	double tol;            // the cells' tolerance (tolerance * max_len_sq)
};

}

namespace {

template<class c_class, class c_loop>
long long run_loop(c_class &con, ref_out *o, double *secs) {
	voronoicell_neighbor c(con);
	c_loop vl(con);
	std::vector<int> v;
	long long m = 0, nn = 0;
	double sink = 0;
	auto t0 = std::chrono::steady_clock::now();
	if (vl.start()) do {
		bool ok = con.compute_cell(c, vl);
		if (!o && ok) {
			// Timing mode: produce the same outputs, keep only a checksum.
			double cx, cy, cz;
			c.centroid(cx, cy, cz);
			c.neighbors(v);
			sink += c.volume() + cx + c.number_of_faces() + (double)v.size();
		}
		if (o) {
			o->id[m] = vl.pid();
			o->ijk[m] = vl.ijk;
			o->q[m] = vl.q;
			o->exists[m] = ok ? 1 : 0;
			if (o->pos) {
				double x, y, z;
				vl.pos(x, y, z);
				o->pos[3 * m] = x; o->pos[3 * m + 1] = y; o->pos[3 * m + 2] = z;
			}
			o->nbr_off[m] = nn;
			if (ok) {
				o->vol[m] = c.volume();
				double cx, cy, cz;
				c.centroid(cx, cy, cz);
				o->cent[3 * m] = cx; o->cent[3 * m + 1] = cy; o->cent[3 * m + 2] = cz;
				o->nfaces[m] = c.number_of_faces();
				c.neighbors(v);
				for (size_t t = 0; t < v.size(); t++) {
					if (nn < o->nbr_cap) o->nbr[nn] = v[t];
					nn++;
				}
			} else {
				o->vol[m] = 0; o->cent[3 * m] = o->cent[3 * m + 1] = o->cent[3 * m + 2] = 0;
				o->nfaces[m] = 0;
			}
		}
		m++;
	} while (vl.inc());
	auto t1 = std::chrono::steady_clock::now();
	if (o) o->nbr_off[m] = nn;
	if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
	if (sink == -1.0) m = -m;  // keeps the timing-mode work observable
	return m;
}

template<class c_class>
void put_all(c_class &con, int n, const int *ids, const double *xyz, const double *) {
	for (int i = 0; i < n; i++) con.put(ids[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

template<class c_class>
void put_all_r(c_class &con, int n, const int *ids, const double *xyz, const double *r) {
	for (int i = 0; i < n; i++) con.put(ids[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], r[i]);
}

template<class c_class>
void add_spheres(c_class &con, const ref_spec *s, std::vector<wall*> &keep) {
	for (int w = 0; w < s->n_spheres; w++) {
		const double *q = s->spheres + 5 * w;
		wall_sphere *ws = new wall_sphere(q[0], q[1], q[2], q[3], (int)q[4]);
		keep.push_back(ws);
		con.add_wall(*ws);
	}
	for (int w = 0; w < s->n_walls; w++) {
		const double *q = s->walls + 9 * w;
		const int type = (int)q[0], id = (int)q[1];
		const double *a = q + 2;
		wall *x;
		if (type == 0) x = new wall_sphere(a[0], a[1], a[2], a[3], id);
		else if (type == 1) x = new wall_plane(a[0], a[1], a[2], a[3], id);
		else if (type == 2) x = new wall_cylinder(a[0], a[1], a[2], a[3], a[4], a[5], a[6], id);
		else x = new wall_cone(a[0], a[1], a[2], a[3], a[4], a[5], a[6], id);
		keep.push_back(x);
		con.add_wall(*x);
	}
}

// mode 0: neighbour loop with outputs; mode 1: time compute_all_cells() only;
// mode 2: time the neighbour loop only (outputs discarded).
long long dispatch(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
		   ref_out *o, double *secs, int mode) {
	const double *g = s->geom;
	std::vector<wall*> keep;
	long long m = 0;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container, c_loop_all>(con, mode == 0 ? o : 0, secs);
	} else if (s->kind == 1) {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all_r(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container_poly, c_loop_all>(con, mode == 0 ? o : 0, secs);
	} else if (s->kind == 2) {
		container_periodic con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				       s->init_mem);
		put_all(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container_periodic, c_loop_all_periodic>(con, mode == 0 ? o : 0, secs);
	} else {
		container_periodic_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1],
					    s->grid[2], s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		if (mode == 1) {
			auto t0 = std::chrono::steady_clock::now();
			con.compute_all_cells();
			auto t1 = std::chrono::steady_clock::now();
			*secs = std::chrono::duration<double>(t1 - t0).count();
			m = n;
		} else m = run_loop<container_periodic_poly, c_loop_all_periodic>(con, mode == 0 ? o : 0, secs);
	}
	for (size_t w = 0; w < keep.size(); w++) delete keep[w];
	return m;
}

template<class c_class>
void text_of(c_class &con, int what, const char *format, const char *filename) {
	switch (what) {
	case 0: con.print_custom(format, filename); break;
	case 1: con.draw_particles(filename); break;
	case 2: con.draw_particles_pov(filename); break;
	case 3: con.draw_cells_gnuplot(filename); break;
	default: con.draw_cells_pov(filename); break;
	}
}

long long ref_text(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r, int what,
		   const char *format, const char *filename) {
	const double *g = s->geom;
	std::vector<wall*> keep;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all(con, n, ids, xyz, r);
		text_of(con, what, format, filename);
	} else if (s->kind == 1) {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all_r(con, n, ids, xyz, r);
		text_of(con, what, format, filename);
	} else if (s->kind == 2) {
		container_periodic con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2], s->init_mem);
		put_all(con, n, ids, xyz, r);
		text_of(con, what, format, filename);
	} else {
		container_periodic_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
					    s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		text_of(con, what, format, filename);
	}
	for (size_t w = 0; w < keep.size(); w++) delete keep[w];
	return 0;
}

// The periodic classes' find_voronoi_cell (src/container_prd.cc:303-360) and
// compute_ghost_cell (src/container_prd.hh:414-420,635-645) for nq points,
// made one after another on one container; pre != 0: compute_all_cells()
// first (the image blocks it creates are then the ones the queries start
// from).  Outputs as ref_find_voronoi_cells / ref_compute_ghost_cells;
// ghost = 0 finds, 1 ghost cells.  The ghost point is written into its block
// without growing it (the reference's code), so the caller's init_mem must
// exceed every block's count.
inline bool ghost_call(container_periodic &con, voronoicell_neighbor &c, double x, double y, double z, const double *) {
	return con.compute_ghost_cell(c, x, y, z);
}
inline bool ghost_call(container_periodic_poly &con, voronoicell_neighbor &c, double x, double y, double z,
		       const double *r) {
	return con.compute_ghost_cell(c, x, y, z, *r);
}

template<class c_class>
long long prd_queries_impl(c_class &con, int pre, int ghost, int nq, const double *q, const double *qr, int *found,
			   int *pid, double *pos, double *vol, double *cent, int *nfaces, int *nbr, int max_nbr,
			   bool poly) {
	if (pre) con.compute_all_cells();
	long long m = 0;
	std::vector<int> nb;
	voronoicell_neighbor c(con);
	for (int i = 0; i < nq; i++) {
		const double x = q[3 * i], y = q[3 * i + 1], z = q[3 * i + 2];
		if (!ghost) {
			found[i] = con.find_voronoi_cell(x, y, z, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], pid[i]) ? 1 : 0;
			m += found[i];
			continue;
		}
		const bool ok = ghost_call(con, c, x, y, z, poly ? qr + i : 0);
		found[i] = ok ? 1 : 0;
		vol[i] = 0; cent[3 * i] = cent[3 * i + 1] = cent[3 * i + 2] = 0; nfaces[i] = 0;
		if (!ok) continue;
		m++;
		vol[i] = c.volume();
		c.centroid(cent[3 * i], cent[3 * i + 1], cent[3 * i + 2]);
		nfaces[i] = c.number_of_faces();
		c.neighbors(nb);
		for (int f = 0; f < (int)nb.size() && f < max_nbr; f++) nbr[(size_t)i * max_nbr + f] = nb[f];
	}
	return m;
}

}  // namespace

extern "C" {

// Runs the voronoicell_neighbor loop and fills `o` in c_loop_all order.
// Returns the number of loop entries (particles stored in the container).
long long ref_compute_neighbor(const ref_spec *s, int n, const int *ids, const double *xyz,
			       const double *r, ref_out *o, double *secs) {
	return dispatch(s, n, ids, xyz, r, o, secs, 0);
}

// Times container::compute_all_cells() (steady_clock around the call only, as
// examples/timing/timing_test.cc:51-57 does).  Returns seconds.
double ref_time_compute_all_cells(const ref_spec *s, int n, const int *ids, const double *xyz,
				  const double *r) {
	double secs = 0;
	dispatch(s, n, ids, xyz, r, 0, &secs, 1);
	return secs;
}

// container(_poly)::find_voronoi_cell (src/container.cc:237-297) for nq
// query points; found[i] is its return value.  With `secs`, the time of the
// query loop alone.  Rectangular kinds only (returns -1 otherwise).
long long ref_find_voronoi_cells(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
				 int nq, const double *q, int *found, int *pid, double *pos, double *secs) {
	const double *g = s->geom;
	if (s->kind > 1) return -1;
	long long nf = 0;
	std::chrono::steady_clock::time_point t0, t1;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all(con, n, ids, xyz, r);
		t0 = std::chrono::steady_clock::now();
		for (int i = 0; i < nq; i++) {
			found[i] = con.find_voronoi_cell(q[3 * i], q[3 * i + 1], q[3 * i + 2], pos[3 * i], pos[3 * i + 1],
							 pos[3 * i + 2], pid[i]) ? 1 : 0;
			nf += found[i];
		}
		t1 = std::chrono::steady_clock::now();
	} else {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		t0 = std::chrono::steady_clock::now();
		for (int i = 0; i < nq; i++) {
			found[i] = con.find_voronoi_cell(q[3 * i], q[3 * i + 1], q[3 * i + 2], pos[3 * i], pos[3 * i + 1],
							 pos[3 * i + 2], pid[i]) ? 1 : 0;
			nf += found[i];
		}
		t1 = std::chrono::steady_clock::now();
	}
	if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
	return nf;
}

// container(_poly)::compute_ghost_cell (src/container.hh:488-498,710-721)
// for nq points: return value, volume, centroid, face count and up to
// max_nbr neighbour ids per point.  Rectangular kinds only (returns -1).
long long ref_compute_ghost_cells(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
				  int nq, const double *q, const double *qr, int *exists, double *vol, double *cent,
				  int *nfaces, int *nbr, int max_nbr) {
	const double *g = s->geom;
	if (s->kind > 1) return -1;
	std::vector<int> nb;
	long long m = 0;
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all(con, n, ids, xyz, r);
		voronoicell_neighbor c(con);
		for (int i = 0; i < nq; i++) {
			exists[i] = con.compute_ghost_cell(c, q[3 * i], q[3 * i + 1], q[3 * i + 2]) ? 1 : 0;
			vol[i] = 0; cent[3 * i] = cent[3 * i + 1] = cent[3 * i + 2] = 0; nfaces[i] = 0;
			if (!exists[i]) continue;
			m++;
			vol[i] = c.volume();
			c.centroid(cent[3 * i], cent[3 * i + 1], cent[3 * i + 2]);
			nfaces[i] = c.number_of_faces();
			c.neighbors(nb);
			for (int f = 0; f < (int)nb.size() && f < max_nbr; f++) nbr[(size_t)i * max_nbr + f] = nb[f];
		}
	} else {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		voronoicell_neighbor c(con);
		for (int i = 0; i < nq; i++) {
			exists[i] = con.compute_ghost_cell(c, q[3 * i], q[3 * i + 1], q[3 * i + 2], qr[i]) ? 1 : 0;
			vol[i] = 0; cent[3 * i] = cent[3 * i + 1] = cent[3 * i + 2] = 0; nfaces[i] = 0;
			if (!exists[i]) continue;
			m++;
			vol[i] = c.volume();
			c.centroid(cent[3 * i], cent[3 * i + 1], cent[3 * i + 2]);
			nfaces[i] = c.number_of_faces();
			c.neighbors(nb);
			for (int f = 0; f < (int)nb.size() && f < max_nbr; f++) nbr[(size_t)i * max_nbr + f] = nb[f];
		}
	}
	return m;
}

// print_custom(format, filename) (src/container.cc:401-432 and the periodic
// classes'), and the drawings: what 0 print_custom, 1 draw_particles,
// 2 draw_particles_pov, 3 draw_cells_gnuplot, 4 draw_cells_pov.
long long ref_prd_queries(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r, int pre,
			  int ghost, int nq, const double *q, const double *qr, int *found, int *pid, double *pos,
			  double *vol, double *cent, int *nfaces, int *nbr, int max_nbr) {
	const double *g = s->geom;
	if (s->kind == 2) {
		container_periodic con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				       s->init_mem);
		put_all(con, n, ids, xyz, r);
		return prd_queries_impl(con, pre, ghost, nq, q, qr, found, pid, pos, vol, cent, nfaces, nbr, max_nbr,
					false);
	}
	if (s->kind == 3) {
		container_periodic_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
					    s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		return prd_queries_impl(con, pre, ghost, nq, q, qr, found, pid, pos, vol, cent, nfaces, nbr, max_nbr,
					true);
	}
	return -1;
}

long long ref_print_custom(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
			   const char *format, const char *filename) {
	return ref_text(s, n, ids, xyz, r, 0, format, filename);
}

// Times the voronoicell_neighbor loop with volume/centroid/faces/neighbours.
double ref_time_neighbor_loop(const ref_spec *s, int n, const int *ids, const double *xyz,
			      const double *r) {
	double secs = 0;
	dispatch(s, n, ids, xyz, r, 0, &secs, 2);
	return secs;
}


}  // extern "C"

namespace {

template<class c_class, class c_loop>
long long geo_loop(c_class &con, ref_geo *g) {
	voronoicell_neighbor c(con);
	c_loop vl(con);
	std::vector<int> vi;
	std::vector<double> vd;
	long long m = 0, nv = 0, ne = 0, nf = 0, nfv = 0;
	g->tol = c.tol;
	if (vl.start()) do {
		g->vert_off[m] = nv; g->edge_off[m] = ne; g->face_off[m] = nf; g->fv_off[m] = nfv;
		if (con.compute_cell(c, vl)) {
			c.vertices(vd);
			for (int i = 0; i < c.p; i++) {
				if (nv + i < g->vcap) {
					for (int a = 0; a < 3; a++) {
						g->vertices[3 * (nv + i) + a] = vd[3 * i + a];
						g->pts[3 * (nv + i) + a] = c.pts[4 * i + a];
					}
					g->orders[nv + i] = c.nu[i];
				}
				for (int j = 0; j < c.nu[i]; j++) {
					if (ne < g->ecap) { g->edges[ne] = c.ed[i][j]; g->edge_rel[ne] = c.ed[i][c.nu[i] + j]; }
					ne++;
				}
			}
			nv += c.p;
			std::vector<double> fa, fp, nr;
			std::vector<int> fo;
			c.face_areas(fa);
			c.face_perimeters(fp);
			c.normals(nr);
			c.face_orders(fo);
			for (size_t f = 0; f < fa.size(); f++) {
				if (nf + (long long)f < g->fcap) {
					g->face_area[nf + f] = fa[f];
					g->face_perim[nf + f] = fp[f];
					g->face_order[nf + f] = fo[f];
					for (int a = 0; a < 3; a++) g->normal[3 * (nf + f) + a] = nr[3 * f + a];
				}
			}
			nf += (long long)fa.size();
			c.face_vertices(vi);
			for (size_t t = 0; t < vi.size(); t++) if (nfv + (long long)t < g->fvcap) g->face_verts[nfv + t] = vi[t];
			nfv += (long long)vi.size();
			g->surface[m] = c.surface_area();
			g->edge_dist[m] = c.total_edge_distance();
			g->n_edges[m] = c.number_of_edges();
			g->maxr2[m] = c.max_radius_squared();
		} else {
			g->surface[m] = 0; g->edge_dist[m] = 0; g->n_edges[m] = 0; g->maxr2[m] = 0;
		}
		m++;
	} while (vl.inc());
	g->vert_off[m] = nv; g->edge_off[m] = ne; g->face_off[m] = nf; g->fv_off[m] = nfv;
	return m;
}

// ---- Threaded runs: one container copy per thread, each thread computing
// the cells of its own z-slab of blocks through c_loop_subset::setup_intbox
// (src/c_loops.cc:44-50) -- the all-core harness of the CPU baseline.  The
// reference has no threading of its own; separate container objects share
// nothing (its scratch lives in each container, src/v_compute.hh:103-119).
// A cell of a rectangular container depends only on the particles, so the
// threaded outputs equal c_loop_all's, in the same order.

// z-slab boundaries [kb[t], kb[t+1]) balanced by particle count
std::vector<int> slab_bounds(const int *co, int nx, int ny, int nz, int T) {
	std::vector<long long> layer(nz, 0);
	long long tot = 0;
	for (int k = 0; k < nz; k++) {
		for (int b = 0; b < nx * ny; b++) layer[k] += co[k * nx * ny + b];
		tot += layer[k];
	}
	std::vector<int> kb(T + 1, nz);
	kb[0] = 0;
	long long acc = 0;
	int t = 1;
	for (int k = 0; k < nz && t < T; k++) {
		acc += layer[k];
		while (t < T && acc * T >= tot * t) kb[t++] = k + 1;
	}
	for (int u = 1; u <= T; u++) if (kb[u] < kb[u - 1]) kb[u] = kb[u - 1];
	return kb;
}

// Releases every thread at once after all have built their containers.
struct Gate {
	std::atomic<int> ready{0};
	std::atomic<int> go{0};
	void arrive_and_wait() {
		ready.fetch_add(1);
		while (!go.load()) std::this_thread::yield();
	}
};

template<class c_class>
void slab_outputs(c_class &con, int k0, int k1, long long m0, ref_out *o, std::vector<int> &nb,
		  std::vector<long long> &cnt, double &sink) {
	voronoicell_neighbor c(con);
	c_loop_subset vl(con);
	std::vector<int> v;
	if (k1 <= k0) return;
	vl.setup_intbox(0, con.nx - 1, 0, con.ny - 1, k0, k1 - 1);
	long long m = m0;
	if (vl.start()) do {
		const bool ok = con.compute_cell(c, vl);
		if (!o) {
			if (ok) {
				double cx, cy, cz;
				c.centroid(cx, cy, cz);
				c.neighbors(v);
				sink += c.volume() + cx + c.number_of_faces() + (double)v.size();
			}
			continue;
		}
		o->id[m] = vl.pid();
		o->ijk[m] = vl.ijk;
		o->q[m] = vl.q;
		o->exists[m] = ok ? 1 : 0;
		if (o->pos) {
			double x, y, z;
			vl.pos(x, y, z);
			o->pos[3 * m] = x; o->pos[3 * m + 1] = y; o->pos[3 * m + 2] = z;
		}
		if (ok) {
			o->vol[m] = c.volume();
			double cx, cy, cz;
			c.centroid(cx, cy, cz);
			o->cent[3 * m] = cx; o->cent[3 * m + 1] = cy; o->cent[3 * m + 2] = cz;
			o->nfaces[m] = c.number_of_faces();
			c.neighbors(v);
			nb.insert(nb.end(), v.begin(), v.end());
			cnt.push_back((long long)v.size());
		} else {
			o->vol[m] = 0; o->cent[3 * m] = o->cent[3 * m + 1] = o->cent[3 * m + 2] = 0;
			o->nfaces[m] = 0;
			cnt.push_back(0);
		}
		m++;
	} while (vl.inc());
}

// compute_all_cells() restricted to a slab (the same loop body,
// src/container.cc:438-443).
template<class c_class>
void slab_compute(c_class &con, int k0, int k1) {
	if (k1 <= k0) return;
	voronoicell c(con);
	c_loop_subset vl(con);
	vl.setup_intbox(0, con.nx - 1, 0, con.ny - 1, k0, k1 - 1);
	if (vl.start()) do con.compute_cell(c, vl);
	while (vl.inc());
}

template<class c_class>
c_class *make_rect(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
		   std::vector<wall*> &keep);
template<>
container *make_rect<container>(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
				std::vector<wall*> &keep) {
	const double *g = s->geom;
	container *con = new container(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				       s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
	add_spheres(*con, s, keep);
	put_all(*con, n, ids, xyz, r);
	return con;
}
template<>
container_poly *make_rect<container_poly>(const ref_spec *s, int n, const int *ids, const double *xyz,
					  const double *r, std::vector<wall*> &keep) {
	const double *g = s->geom;
	container_poly *con = new container_poly(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1],
						 s->grid[2], s->periodic[0] != 0, s->periodic[1] != 0,
						 s->periodic[2] != 0, s->init_mem);
	add_spheres(*con, s, keep);
	put_all_r(*con, n, ids, xyz, r);
	return con;
}

// mode 0: outputs into o; mode 1: compute_all_cells-equivalent timing;
// mode 2: neighbour-loop timing.  Returns the loop length.
template<class c_class>
long long threaded_rect(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r, ref_out *o,
			int T, double *secs, int mode) {
	std::vector<wall*> keep;
	// the block occupancy decides the slabs and each slab's first loop index
	c_class *probe = make_rect<c_class>(s, n, ids, xyz, r, keep);
	const int nx = probe->nx, ny = probe->ny, nz = probe->nz;
	std::vector<int> kb = slab_bounds(probe->co, nx, ny, nz, T);
	std::vector<long long> first(T + 1, 0);
	for (int t = 0; t < T; t++) {
		long long a = 0;
		for (int b = kb[t] * nx * ny; b < kb[t + 1] * nx * ny; b++) a += probe->co[b];
		first[t + 1] = first[t] + a;
	}
	std::vector<std::vector<int>> nbs(T);
	std::vector<std::vector<long long>> cnts(T);
	std::vector<c_class*> cons(T, nullptr);
	cons[0] = probe;
	std::vector<std::vector<wall*>> keeps(T);
	Gate gate;
	std::vector<double> sinks(T, 0);
	std::vector<std::thread> th;
	for (int t = 0; t < T; t++) {
		th.emplace_back([&, t]() {
			if (t > 0) cons[t] = make_rect<c_class>(s, n, ids, xyz, r, keeps[t]);
			gate.arrive_and_wait();
			if (mode == 1) slab_compute(*cons[t], kb[t], kb[t + 1]);
			else slab_outputs(*cons[t], kb[t], kb[t + 1], first[t], mode == 0 ? o : 0, nbs[t], cnts[t], sinks[t]);
		});
	}
	while (gate.ready.load() < T) std::this_thread::yield();
	auto t0 = std::chrono::steady_clock::now();
	gate.go.store(1);
	for (auto &x : th) x.join();
	auto t1 = std::chrono::steady_clock::now();
	if (secs) *secs = std::chrono::duration<double>(t1 - t0).count();
	const long long m = first[T];
	if (o) {
		long long nn = 0, i = 0;
		for (int t = 0; t < T; t++) {
			for (size_t c = 0; c < cnts[t].size(); c++) { o->nbr_off[i++] = nn; nn += cnts[t][c]; }
		}
		o->nbr_off[m] = nn;
		long long w = 0;
		for (int t = 0; t < T; t++)
			for (size_t c = 0; c < nbs[t].size(); c++, w++) if (w < o->nbr_cap) o->nbr[w] = nbs[t][c];
	}
	for (int t = 0; t < T; t++) delete cons[t];
	for (auto *w : keep) delete w;
	for (auto &kv : keeps) for (auto *w : kv) delete w;
	double sk = 0;
	for (double x : sinks) sk += x;
	return sk == -1.0 ? -m : m;
}

void put_prd(container_periodic &con, int n, const int *ids, const double *xyz, const double *r) {
	put_all(con, n, ids, xyz, r);
}
void put_prd(container_periodic_poly &con, int n, const int *ids, const double *xyz, const double *r) {
	put_all_r(con, n, ids, xyz, r);
}

// Periodic kinds: each thread loops c_loop_all_periodic over its own
// container and computes the cells of its z-slab of primary blocks.  Image
// blocks are created lazily per container, so only the timing is used.
template<class c_class>
double threaded_prd_time(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r, int T) {
	const double *g = s->geom;
	std::vector<c_class*> cons(T, nullptr);
	Gate gate;
	std::vector<std::thread> th;
	for (int t = 0; t < T; t++) {
		th.emplace_back([&, t]() {
			cons[t] = new c_class(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
					      s->init_mem);
			put_prd(*cons[t], n, ids, xyz, r);
			gate.arrive_and_wait();
			c_class &con = *cons[t];
			const int nz = con.nz, k0 = nz * t / T, k1 = nz * (t + 1) / T;
			voronoicell c(con);
			c_loop_all_periodic vl(con);
			if (vl.start()) do {
				const int kk = vl.k - con.ez;   // primary layers are ez .. wz-1
				if (kk >= k0 && kk < k1) con.compute_cell(c, vl);
			} while (vl.inc());
		});
	}
	while (gate.ready.load() < T) std::this_thread::yield();
	auto t0 = std::chrono::steady_clock::now();
	gate.go.store(1);
	for (auto &x : th) x.join();
	auto t1 = std::chrono::steady_clock::now();
	for (int t = 0; t < T; t++) delete cons[t];
	return std::chrono::duration<double>(t1 - t0).count();
}

}  // namespace

extern "C" {

// The voronoicell_neighbor loop of a rectangular container on T threads
// (z-slabs), filling `o` in c_loop_all order exactly as
// ref_compute_neighbor does.  Returns the loop length, -1 for periodic kinds.
long long ref_compute_neighbor_mt(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
				  ref_out *o, int threads, double *secs) {
	if (s->kind > 1) return -1;
	if (threads < 1) threads = 1;
	return s->kind == 0 ? threaded_rect<container>(s, n, ids, xyz, r, o, threads, secs, 0)
			    : threaded_rect<container_poly>(s, n, ids, xyz, r, o, threads, secs, 0);
}

// Wall time of compute_all_cells() spread over T threads (z-slabs, one
// container copy each; building the copies is not timed).  loop 1: the
// compute_all_cells body; loop 2: the voronoicell_neighbor output loop.
double ref_time_mt(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r, int threads,
		   int loop) {
	double secs = 0;
	if (threads < 1) threads = 1;
	if (s->kind == 0) threaded_rect<container>(s, n, ids, xyz, r, 0, threads, &secs, loop);
	else if (s->kind == 1) threaded_rect<container_poly>(s, n, ids, xyz, r, 0, threads, &secs, loop);
	else if (s->kind == 2) secs = threaded_prd_time<container_periodic>(s, n, ids, xyz, nullptr, threads);
	else secs = threaded_prd_time<container_periodic_poly>(s, n, ids, xyz, r, threads);
	return secs;
}


// The polyhedron statistics of every cell (loop order), rectangular and
// periodic kinds; returns the loop length.
long long ref_cell_geometry(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r,
			    ref_geo *g) {
	const double *gm = s->geom;
	std::vector<wall*> keep;
	long long m = 0;
	if (s->kind == 0) {
		container con(gm[0], gm[1], gm[2], gm[3], gm[4], gm[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all(con, n, ids, xyz, r);
		m = geo_loop<container, c_loop_all>(con, g);
	} else if (s->kind == 1) {
		container_poly con(gm[0], gm[1], gm[2], gm[3], gm[4], gm[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		add_spheres(con, s, keep);
		put_all_r(con, n, ids, xyz, r);
		m = geo_loop<container_poly, c_loop_all>(con, g);
	} else if (s->kind == 2) {
		container_periodic con(gm[0], gm[1], gm[2], gm[3], gm[4], gm[5], s->grid[0], s->grid[1], s->grid[2],
				       s->init_mem);
		put_all(con, n, ids, xyz, r);
		m = geo_loop<container_periodic, c_loop_all_periodic>(con, g);
	} else {
		container_periodic_poly con(gm[0], gm[1], gm[2], gm[3], gm[4], gm[5], s->grid[0], s->grid[1],
					    s->grid[2], s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		m = geo_loop<container_periodic_poly, c_loop_all_periodic>(con, g);
	}
	for (size_t w = 0; w < keep.size(); w++) delete keep[w];
	return m;
}


// The reference's drawings (what 1..4 as ref_text) into filename.
long long ref_draw(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r, int what,
		   const char *filename) {
	return ref_text(s, n, ids, xyz, r, what, "", filename);
}


// c_loop_subset (src/c_loops.hh:218-258) of the reference over a rectangular
// container: mode 0 setup_sphere(v0..v3, bounds_test), 1 setup_box(v0..v5,
// bounds_test), 2 setup_intbox(iv0..iv5); for each step up to cap: the
// particle id, i, j, k, ijk, q.  Returns the loop length.
long long ref_loop_subset(const ref_spec *s, int n, const int *ids, const double *xyz, const double *r, int mode,
			  const double *v, const int *iv, int bounds_test, int *out6, long long cap) {
	const double *g = s->geom;
	std::vector<wall*> keep;
	long long m = 0;
	auto run = [&](auto &con) {
		c_loop_subset vl(con);
		if (mode == 0) vl.setup_sphere(v[0], v[1], v[2], v[3], bounds_test != 0);
		else if (mode == 1) vl.setup_box(v[0], v[1], v[2], v[3], v[4], v[5], bounds_test != 0);
		else vl.setup_intbox(iv[0], iv[1], iv[2], iv[3], iv[4], iv[5]);
		if (vl.start()) do {
			if (m < cap) {
				int *o = out6 + 6 * m;
				o[0] = vl.pid(); o[1] = vl.i; o[2] = vl.j; o[3] = vl.k; o[4] = vl.ijk; o[5] = vl.q;
			}
			m++;
		} while (vl.inc());
	};
	if (s->kind == 0) {
		container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
			      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all(con, n, ids, xyz, r);
		run(con);
	} else if (s->kind == 1) {
		container_poly con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
				   s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
		put_all_r(con, n, ids, xyz, r);
		run(con);
	} else return -1;
	return m;
}


// container::point_inside (src/container.cc:484-487) of a rectangular
// container with the spec's walls, for nq points.
long long ref_point_inside(const ref_spec *s, int nq, const double *q, int *inside) {
	const double *g = s->geom;
	std::vector<wall*> keep;
	container con(g[0], g[1], g[2], g[3], g[4], g[5], s->grid[0], s->grid[1], s->grid[2],
		      s->periodic[0] != 0, s->periodic[1] != 0, s->periodic[2] != 0, s->init_mem);
	add_spheres(con, s, keep);
	long long k = 0;
	for (int i = 0; i < nq; i++) { inside[i] = con.point_inside(q[3 * i], q[3 * i + 1], q[3 * i + 2]) ? 1 : 0; k += inside[i]; }
	for (size_t w = 0; w < keep.size(); w++) delete keep[w];
	return k;
}

}
