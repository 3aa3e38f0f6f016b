// voro_amd.hh -- header-only C++ drop-in for the reference's hot path,
// over the C ABI of voro_gpu.h (link with -lvoro_gpu).
//
// Mirrors the reference classes a program uses to compute cells:
//   container / container_poly        src/container.hh:288-728
//   container_periodic / _poly        src/container_prd.hh:213-651
//   wall_sphere                       src/wall.hh:27-30
//   c_loop_all / c_loop_all_periodic  src/c_loops.hh:165-209, 325-393
//   voronoicell_neighbor outputs      src/cell.hh:432-549
// A reference program that does
//     container con(...); con.put(i,x,y,z); ...
//     c_loop_all vl(con); voronoicell_neighbor c(con);
//     if(vl.start()) do if(con.compute_cell(c,vl)) { c.volume(); c.neighbors(v); ... } while(vl.inc());
// compiles unchanged against this header (namespace voro_amd instead of
// voro).  The first compute_cell (or compute_all_cells) computes every
// cell on the GPU; later compute_cell calls read the stored cell.  Errors the
// reference reports with exit() are thrown as voro_amd::error here.
#ifndef VORO_AMD_HH
#define VORO_AMD_HH

#include <stdexcept>
#include <string>
#include <vector>

#include "voro_gpu.h"

namespace voro_amd {

struct error : std::runtime_error {
	int status;
	error(int s, const char *what)
		: std::runtime_error(std::string(what) + ": " + vgpu_status_string(s)), status(s) {}
};

inline void check(int s, const char *what) {
	if (s != VGPU_OK) throw error(s, what);
}

struct wall_sphere {
	double xc, yc, zc, rc;
	int w_id;
	wall_sphere(double xc_, double yc_, double zc_, double rc_, int w_id_ = -99)
		: xc(xc_), yc(yc_), zc(zc_), rc(rc_), w_id(w_id_) {}
};

class voronoicell_neighbor;
class c_loop_all;

class container_base_gpu {
    public:
	~container_base_gpu() { if (h) vgpu_destroy(h); }
	container_base_gpu(const container_base_gpu &) = delete;
	container_base_gpu &operator=(const container_base_gpu &) = delete;

	void add_wall(wall_sphere &w) {
		check(vgpu_add_wall_sphere(h, w.xc, w.yc, w.zc, w.rc, w.w_id), "add_wall");
		dirty = true;
	}
	void clear() { check(vgpu_clear(h), "clear"); dirty = true; }
	// Computes every cell and keeps the results (src/container.cc:438-443).
	void compute_all_cells() {
		check(vgpu_compute_all(h, VGPU_F_NEIGHBORS | VGPU_F_CSR, nullptr), "compute_all_cells");
		fetch();
		dirty = false;
	}
	double sum_cell_volumes() {
		ensure();
		double v = 0;
		check(vgpu_sum_volumes(h, &v), "sum_cell_volumes");
		return v;
	}
	template<class v_cell, class c_loop> bool compute_cell(v_cell &c, c_loop &vl);
	vgpu_container *handle() { return h; }

	// loop-order results
	std::vector<int> id, ijk, q, exists, nfaces, nbr;
	std::vector<double> pos, vol, cent;
	std::vector<long long> nbr_off;

    protected:
	container_base_gpu(int kind, const double *geom, const int *grid, const int *per) {
		check(vgpu_create(kind, geom, grid, per, 8, &h), "container");
	}
	void ensure() { if (dirty) compute_all_cells(); }
	void fetch() {
		long long n = 0;
		check(vgpu_num_cells(h, &n), "num_cells");
		id.resize(n); ijk.resize(n); q.resize(n); exists.resize(n); nfaces.resize(n);
		pos.resize(3 * n); vol.resize(n); cent.resize(3 * n); nbr_off.resize(n + 1);
		long long tot = 0;
		check(vgpu_get_results(h, id.data(), ijk.data(), q.data(), exists.data(), pos.data(), vol.data(),
				       cent.data(), nfaces.data(), nbr_off.data(), nullptr, 0, &tot), "results");
		nbr.resize(tot > 0 ? tot : 1);
		check(vgpu_get_results(h, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
				       nullptr, nbr.data(), tot, &tot), "results");
	}
	vgpu_container *h = nullptr;
	bool dirty = true;
	friend class c_loop_all;
};

// container(ax,bx,ay,by,az,bz,nx,ny,nz,xp,yp,zp,init_mem), src/container.hh:290-293
class container : public container_base_gpu {
    public:
	container(double ax, double bx, double ay, double by, double az, double bz, int nx, int ny, int nz, bool xp,
		  bool yp, bool zp, int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER, geom6(ax, bx, ay, by, az, bz).v, grid3(nx, ny, nz).v,
				     grid3(xp, yp, zp).v) { (void)init_mem; }
	void put(int n, double x, double y, double z) {
		double p[3] = {x, y, z};
		check(vgpu_put(h, 1, &n, p, nullptr), "put");
		dirty = true;
	}
	// find_voronoi_cell, src/container.cc:237-297 (one query; batches of
	// queries: vgpu_find_voronoi_cells)
	bool find_voronoi_cell(double x, double y, double z, double &rx, double &ry, double &rz, int &pid) {
		double xyz[3] = {x, y, z}, r[3];
		int found = 0, id_ = 0;
		check(vgpu_find_voronoi_cells(h, 1, xyz, &found, &id_, r, 0), "find_voronoi_cell");
		if (!found) return false;
		rx = r[0]; ry = r[1]; rz = r[2]; pid = id_;
		return true;
	}

    private:
	struct geom6 { double v[6]; geom6(double a, double b, double c, double d, double e, double f) : v{a, b, c, d, e, f} {} };
	struct grid3 { int v[3]; grid3(int a, int b, int c) : v{a, b, c} {} };
};

// container_poly, src/container.hh:510-515
class container_poly : public container_base_gpu {
    public:
	container_poly(double ax, double bx, double ay, double by, double az, double bz, int nx, int ny, int nz,
		       bool xp, bool yp, bool zp, int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER_POLY, g6(ax, bx, ay, by, az, bz).v, g3(nx, ny, nz).v, g3(xp, yp, zp).v) {
		(void)init_mem;
	}
	void put(int n, double x, double y, double z, double r) {
		double p[3] = {x, y, z};
		check(vgpu_put(h, 1, &n, p, &r), "put");
		dirty = true;
	}
	// find_voronoi_cell, src/container.cc:237-297 (one query; batches of
	// queries: vgpu_find_voronoi_cells)
	bool find_voronoi_cell(double x, double y, double z, double &rx, double &ry, double &rz, int &pid) {
		double xyz[3] = {x, y, z}, r[3];
		int found = 0, id_ = 0;
		check(vgpu_find_voronoi_cells(h, 1, xyz, &found, &id_, r, 0), "find_voronoi_cell");
		if (!found) return false;
		rx = r[0]; ry = r[1]; rz = r[2]; pid = id_;
		return true;
	}

    private:
	struct g6 { double v[6]; g6(double a, double b, double c, double d, double e, double f) : v{a, b, c, d, e, f} {} };
	struct g3 { int v[3]; g3(int a, int b, int c) : v{a, b, c} {} };
};

// container_periodic(bx,bxy,by,bxz,byz,bz,nx,ny,nz,init_mem),
// src/container_prd.hh:213-217
class container_periodic : public container_base_gpu {
    public:
	container_periodic(double bx, double bxy, double by, double bxz, double byz, double bz, int nx, int ny, int nz,
			   int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER_PERIODIC, g6(bx, bxy, by, bxz, byz, bz).v, g3(nx, ny, nz).v, nullptr) {
		(void)init_mem;
	}
	void put(int n, double x, double y, double z) {
		double p[3] = {x, y, z};
		check(vgpu_put(h, 1, &n, p, nullptr), "put");
		dirty = true;
	}

    private:
	struct g6 { double v[6]; g6(double a, double b, double c, double d, double e, double f) : v{a, b, c, d, e, f} {} };
	struct g3 { int v[3]; g3(int a, int b, int c) : v{a, b, c} {} };
};

// container_periodic_poly, src/container_prd.hh:434-441
class container_periodic_poly : public container_base_gpu {
    public:
	container_periodic_poly(double bx, double bxy, double by, double bxz, double byz, double bz, int nx, int ny,
				int nz, int init_mem = 8)
		: container_base_gpu(VGPU_CONTAINER_PERIODIC_POLY, g6(bx, bxy, by, bxz, byz, bz).v, g3(nx, ny, nz).v,
				     nullptr) {
		(void)init_mem;
	}
	void put(int n, double x, double y, double z, double r) {
		double p[3] = {x, y, z};
		check(vgpu_put(h, 1, &n, p, &r), "put");
		dirty = true;
	}

    private:
	struct g6 { double v[6]; g6(double a, double b, double c, double d, double e, double f) : v{a, b, c, d, e, f} {} };
	struct g3 { int v[3]; g3(int a, int b, int c) : v{a, b, c} {} };
};

// c_loop_all, src/c_loops.hh:165-209: blocks by ascending index, then
// storage order.
class c_loop_all {
    public:
	template<class c_class>
	explicit c_loop_all(c_class &con_) : con(con_) {}
	bool start() { con.ensure(); m = 0; return !con.id.empty(); }
	bool inc() { return ++m < (long long)con.id.size(); }
	int pid() { return con.id[m]; }
	void pos(double &x, double &y, double &z) { x = con.pos[3 * m]; y = con.pos[3 * m + 1]; z = con.pos[3 * m + 2]; }
	long long index() const { return m; }

    private:
	container_base_gpu &con;
	long long m = 0;
	friend class container_base_gpu;
};

// c_loop_all_periodic, src/c_loops.hh:325-393: the primary domain's
// blocks by ascending index, then storage order -- the order the engine
// stores a periodic container's cells in.
typedef c_loop_all c_loop_all_periodic;

// The outputs of voronoicell_neighbor that the engine produces.
class voronoicell_neighbor {
    public:
	voronoicell_neighbor() {}
	template<class c_class> explicit voronoicell_neighbor(c_class &) {}
	double volume() { return vol; }
	void centroid(double &x, double &y, double &z) { x = cx; y = cy; z = cz; }
	int number_of_faces() { return nf; }
	void neighbors(std::vector<int> &v) { v = nb; }

    private:
	double vol = 0, cx = 0, cy = 0, cz = 0;
	int nf = 0;
	std::vector<int> nb;
	friend class container_base_gpu;
};

template<class v_cell, class c_loop>
bool container_base_gpu::compute_cell(v_cell &c, c_loop &vl) {
	ensure();
	long long m = vl.index();
	c.vol = vol[m];
	c.cx = cent[3 * m]; c.cy = cent[3 * m + 1]; c.cz = cent[3 * m + 2];
	c.nf = nfaces[m];
	c.nb.assign(nbr.begin() + nbr_off[m], nbr.begin() + nbr_off[m + 1]);
	return exists[m] != 0;
}

}  // namespace voro_amd

#endif
