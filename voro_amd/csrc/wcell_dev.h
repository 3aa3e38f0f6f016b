// Wavefront cell: one Voronoi cell per wavefront, vertex v held in the
// registers of lane v.
//
// The algorithm, the order of every operation and every floating-point
// expression are those of the order-3 tier (cell3_dev.h, which restates
// src/cell.cc: nplane 718-870 with search_upward/downward 480-515 and
// 597-629, definite_max/min 522-595 and 631-708, create_facet 875-1364,
// max_radius_squared 1907-1917, plane_intersects(_guess) 2366-2410 and the
// output walks 1584-1616, 1861-1901, 2168-2187, 2660-2678).  What differs is
// where the cell lives:
//
//   per lane v:  x, y, z (doubles, twice the coordinates, as the reference
//                stores them), the topology word (three neighbouring vertex
//                indices, 0xff = the reference's -1, and the three edge
//                relations), the three neighbour ids, and the vertex's
//                value for the current cutting plane;
//   uniform:     the vertex count, the delete / visited / search marks as
//                64-bit masks (bit v = vertex v), the stacks as byte arrays
//                spread four entries per lane.
//
// The cell's control flow is the same in every lane, so it runs on the
// scalar unit; reading vertex v is a readlane, writing it a lane-masked
// select, and the plane value of every vertex is one vector expression.
// A plane that leaves every vertex below -tol cannot cut the cell: the
// reference's upward search then ends at a definite maximum without a
// change to the cell (all vertices classify as below, and its ascent and
// tie search are bounded well inside the capacities here), so that test
// replaces the search.
//
// Cells this tier does not finish set ovf and are recomputed from scratch
// by the per-lane tiers: more than 64 vertices, more than 64 faces, and the
// marginal configurations the order-3 tier leaves to the general one (a
// vertex within tol of a cutting plane, a kept vertex within big_tol of it).
//
// CPU builds hold the 64 lanes in arrays, so the same code can be checked
// against the oracle on a host.
#pragma once
#include "vgpu_common.h"

#include <type_traits>

namespace vgpu {

// Vertex indices are bytes (0xff stands for the reference's -1), so four
// vertices per lane hold up to 254.  The last wavefront tier (K = 4) also
// takes the cells whose block search leaves the others' window of +-6
// blocks: its window reaches +-31 blocks (the wide per-lane tier's), its
// ring 8,192 blocks.
template<int K>
struct WLayoutT {
	enum { MAXV = K < 4 ? 64 * K : 254, MAXS = 256, W = K < 4 ? 6 : 31, MAXQ = K < 4 ? 256 : 8192, MAXF = 64 * K };
	enum { WD = 2 * W + 1, WBITS = WD * WD * WD, WWORDS = (WBITS + 31) / 32 };
	// per-wavefront global slab: block-search window and ring only
	static constexpr unsigned o_bm = 0;                          // WWORDS
	static constexpr unsigned o_bms = o_bm + WWORDS * 4;         // WWORDS stamps
	static constexpr unsigned o_qu = o_bms + WWORDS * 4;         // MAXQ ints
	static constexpr size_t bytes = (o_qu + MAXQ * 4 + 255) & ~255u;
	enum { NBCAP = MAXF };
};
typedef WLayoutT<1> WLayout;

#if defined(__HIP_DEVICE_COMPILE__)
#define VG_WDEV 1
#else
#define VG_WDEV 0
#endif

// Value selection among the K register sets of a wavefront cell (never an
// lvalue choice: a conditional or branch choosing between members makes the
// compiler address the cell in memory).
template<int K, class T>
VG_INL T pickk(const T (&a)[K], int h) {
	if constexpr (K == 1) {
		return a[0];
	} else if constexpr (K == 2) {
		const T a0 = a[0], a1 = a[1];
		return h == 0 ? a0 : a1;
	} else {
		// an OR of masked words: a chain of selects over consecutive
		// elements would be turned into an indexed load
		typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type U;
		const U r = (h == 0 ? __builtin_bit_cast(U, a[0]) : U(0)) | (h == 1 ? __builtin_bit_cast(U, a[1]) : U(0)) |
			    (h == 2 ? __builtin_bit_cast(U, a[2]) : U(0)) | (h == 3 ? __builtin_bit_cast(U, a[3]) : U(0));
		return __builtin_bit_cast(T, r);
	}
}
// a[h] = x where c, every set updated by value arithmetic
template<int K, class T>
VG_INL void putk(T (&a)[K], int h, bool c, T x) {
	if constexpr (K == 1) {      // one set: every valid index is below 64
		(void)h;
		a[0] = c ? x : a[0];
		return;
	}
	a[0] = (c && h == 0) ? x : a[0];
	if (K > 1) a[1] = (c && h == 1) ? x : a[1];
	if (K > 2) { a[2] = (c && h == 2) ? x : a[2]; a[3] = (c && h == 3) ? x : a[3]; }
}

// f(integral_constant<h>) for h = 0 .. K-1: register sets indexed by constants.
template<int K, int H = 0, class F>
VG_INL void eachk(F &&f) {
	if constexpr (H < K) {
		f(std::integral_constant<int, H>());
		eachk<K, H + 1>(f);
	}
}

// Vertex marks, bit v = vertex v (K words of 64 bits).
template<int K>
struct WMask {
	unsigned long long w[K];
	VG_INL void clear() {
		w[0] = 0;
		if (K > 1) w[1] = 0;
		if (K > 2) { w[2] = 0; w[3] = 0; }
	}
	VG_INL bool get(int v) const { return (pickk<K>(w, v >> 6) >> (v & 63)) & 1ull; }
	VG_INL void set(int v) { set_if(true, v); }
	VG_INL void unset(int v) {
		if constexpr (K == 1) { w[0] &= ~(1ull << (v & 63)); return; }
		const unsigned long long b = 1ull << (v & 63);
		const int h = v >> 6;
		w[0] &= h == 0 ? ~b : ~0ull;
		if (K > 1) w[1] &= h == 1 ? ~b : ~0ull;
		if (K > 2) { w[2] &= h == 2 ? ~b : ~0ull; w[3] &= h == 3 ? ~b : ~0ull; }
	}
	VG_INL void toggle(int v) {
		if constexpr (K == 1) { w[0] ^= 1ull << (v & 63); return; }
		const unsigned long long b = 1ull << (v & 63);
		const int h = v >> 6;
		w[0] ^= h == 0 ? b : 0ull;
		if (K > 1) w[1] ^= h == 1 ? b : 0ull;
		if (K > 2) { w[2] ^= h == 2 ? b : 0ull; w[3] ^= h == 3 ? b : 0ull; }
	}
	// set(v) when c
	VG_INL void set_if(bool c, int v) {
		if constexpr (K == 1) { w[0] |= c ? 1ull << (v & 63) : 0ull; return; }
		const unsigned long long b = c ? 1ull << (v & 63) : 0ull;
		const int h = v >> 6;
		w[0] |= h == 0 ? b : 0ull;
		if (K > 1) w[1] |= h == 1 ? b : 0ull;
		if (K > 2) { w[2] |= h == 2 ? b : 0ull; w[3] |= h == 3 ? b : 0ull; }
	}
};

// K vertices per lane: vertex v in lane v & 63 of register set v >> 6.
template<int K>
struct WCellT {
	static_assert(K == 1 || K == 2 || K == 4, "one, two or four vertices per lane");
	typedef WLayoutT<K> L;
	enum { MAXV = L::MAXV, MAXS = L::MAXS };
	// the candidate loop may evaluate worklist steps one per lane
	static constexpr bool lane_parallel = VG_WDEV != 0;

#if VG_WDEV
	// register set h holds vertex 64 h + `lane`
	double vx[K], vy[K], vz[K];
	unsigned vt[K];          // e0 | e1 << 8 | e2 << 16 | (r0 | r1 << 2 | r2 << 4) << 24
	int vna[K], vnb[K], vnc[K];   // neighbour ids
	double vtv[K];           // value for the current plane
	unsigned vds, vdm;       // stack entries 4 * lane .. 4 * lane + 3
	int nbv[K];              // output staging: neighbour id number 64 h + `lane`
	unsigned vrec;           // one vertex per lane: the facet records of create_facet1
	unsigned vdec;           // one vertex per lane: the search decisions of decide_search
	int lane;
	__amdgpu_buffer_rsrc_t rs;
#else
	double vx[MAXV], vy[MAXV], vz[MAXV];
	unsigned vt[MAXV];
	int vn[MAXV][3];
	double vtv[MAXV];
	unsigned char vds[MAXS], vdm[MAXS];
	int nbv[MAXV];
	unsigned vrec[MAXV];
	unsigned vdec[MAXV];
	char *wsb;
#endif

	int p, up;
	double px, py, pz, prsq;
	double tol, tol_cu, big_tol;
	int nds;
	bool ovf, fatal;
	int n_cuts;
	WMask<K> delm;               // deletion marks
	WMask<K> flipm;              // tie-search marks (the reference negates back pointers)
	WMask<K> vis0, vis1, vis2;   // face-walk marks per edge slot
	// one register set: the live vertices above the plane (value > tol),
	// below it (< -tol), and below it but within big_tol (plane_values)
	unsigned long long posm, negm, nearm;

	// Value (never lvalue) selection: a conditional over members would make
	// the compiler address the cell in memory.
	template<class T> static VG_INL T sel3(int i, T a, T b, T c) { return i == 0 ? a : (i == 1 ? b : c); }

	// ---- lane primitives ----
#if VG_WDEV
	static VG_INL unsigned rlu(unsigned a, int i) { return (unsigned)__builtin_amdgcn_readlane((int)a, i); }
	static VG_INL int rli(int a, int i) { return __builtin_amdgcn_readlane(a, i); }
	static VG_INL double rld(double a, int i) {
		const unsigned long long b = __builtin_bit_cast(unsigned long long, a);
		const unsigned lo = rlu((unsigned)b, i), hi = rlu((unsigned)(b >> 32), i);
		return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
	}
	static VG_INL double rfd(double a) {
		const unsigned long long b = __builtin_bit_cast(unsigned long long, a);
		const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
		const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
		return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
	}
	VG_INL double X(int v) const { return rld(pickk<K>(vx, v >> 6), v & 63); }
	VG_INL double Y(int v) const { return rld(pickk<K>(vy, v >> 6), v & 63); }
	VG_INL double Z(int v) const { return rld(pickk<K>(vz, v >> 6), v & 63); }
	VG_INL double T(int v) const { return rld(pickk<K>(vtv, v >> 6), v & 63); }
	VG_INL unsigned TW(int v) const { return rlu(pickk<K>(vt, v >> 6), v & 63); }
	VG_INL void setXYZ(int v, double a, double b, double c) {
		const bool m = lane == (v & 63);
		putk<K>(vx, v >> 6, m, a); putk<K>(vy, v >> 6, m, b); putk<K>(vz, v >> 6, m, c);
	}
	VG_INL void setTW(int v, unsigned w) { putk<K>(vt, v >> 6, lane == (v & 63), w); }
	VG_INL int N(int v, int i) const {
		const int l = v & 63, h = v >> 6;
		return sel3(i, rli(pickk<K>(vna, h), l), rli(pickk<K>(vnb, h), l), rli(pickk<K>(vnc, h), l));
	}
	VG_INL void setN3(int v, int a, int b, int c) {
		const bool m = lane == (v & 63);
		putk<K>(vna, v >> 6, m, a); putk<K>(vnb, v >> 6, m, b); putk<K>(vnc, v >> 6, m, c);
	}
	VG_INL int DS(int i) const { return (int)((rlu(vds, i >> 2) >> (8 * (i & 3))) & 0xffu); }
	// a stack write changes one byte of lane i >> 2's word, in that lane's
	// own register (no lane read)
	VG_INL void setDS(int i, int v) {
		const unsigned sh = 8u * (unsigned)(i & 3);
		const unsigned w = (vds & ~(0xffu << sh)) | ((unsigned)v & 0xffu) << sh;
		vds = lane == (i >> 2) ? w : vds;
	}
	VG_INL int DM(int i) const { return (int)((rlu(vdm, i >> 2) >> (8 * (i & 3))) & 0xffu); }
	VG_INL void setDM(int i, int v) {
		const unsigned sh = 8u * (unsigned)(i & 3);
		const unsigned w = (vdm & ~(0xffu << sh)) | ((unsigned)v & 0xffu) << sh;
		vdm = lane == (i >> 2) ? w : vdm;
	}
	VG_INL void put_nb(int i, int v) { putk<K>(nbv, i >> 6, lane == (i & 63), v); }
	VG_INL void put_rec(int j, unsigned r) { vrec = lane == j ? r : vrec; }
	static VG_INL double bperm64(double a, int addr) {
		const unsigned long long b = __builtin_bit_cast(unsigned long long, a);
		const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)b);
		const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(unsigned)(b >> 32));
		return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
	}
	// neighbour id number h * 64 + `lane`
	VG_INL int nb_lane(int h = 0) const { return pickk<K>(nbv, h); }
	// vertex-parallel: bit l set for live vertices h * 64 + l whose predicate holds
	VG_INL unsigned long long live_ballot(bool b, int h = 0) const {
		const unsigned long long m = __ballot(b);
		const int q = p - 64 * h;
		return q >= 64 ? m : (q <= 0 ? 0ull : (m & ((1ull << q) - 1)));
	}
#else
	VG_INL double X(int v) const { return vx[v]; }
	VG_INL double Y(int v) const { return vy[v]; }
	VG_INL double Z(int v) const { return vz[v]; }
	VG_INL double T(int v) const { return vtv[v]; }
	VG_INL unsigned TW(int v) const { return vt[v]; }
	VG_INL void setXYZ(int v, double a, double b, double c) { vx[v] = a; vy[v] = b; vz[v] = c; }
	VG_INL void setTW(int v, unsigned w) { vt[v] = w; }
	VG_INL int N(int v, int i) const { return vn[v][i]; }
	VG_INL void setN3(int v, int a, int b, int c) { vn[v][0] = a; vn[v][1] = b; vn[v][2] = c; }
	VG_INL int DS(int i) const { return vds[i]; }
	VG_INL void setDS(int i, int v) { vds[i] = (unsigned char)v; }
	VG_INL int DM(int i) const { return vdm[i]; }
	VG_INL void setDM(int i, int v) { vdm[i] = (unsigned char)v; }
	VG_INL void put_nb(int i, int v) { nbv[i] = v; }
	VG_INL void put_rec(int j, unsigned r) { vrec[j] = r; }
	VG_INL int nb_lane(int h = 0) const { return nbv[64 * h]; }
#endif

	// ---- per-wavefront slab (block-search arrays) ----
	template<class T> VG_INL T ld(unsigned off) const {
#if VG_WDEV
		static_assert(sizeof(T) == 4, "slab words");
		return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
#else
		T v;
		__builtin_memcpy(&v, wsb + off, sizeof(T));
		return v;
#endif
	}
	template<class T> VG_INL void st(unsigned off, T v) const {
#if VG_WDEV
		static_assert(sizeof(T) == 4, "slab words");
		__builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, off, 0, 0);
#else
		__builtin_memcpy(wsb + off, &v, sizeof(T));
#endif
	}
	template<class T>
	struct Ref {
		const WCellT *c;
		unsigned off;
		VG_INL operator T() const { return c->template ld<T>(off); }
		VG_INL const Ref &operator=(T v) const { c->template st<T>(off, v); return *this; }
		VG_INL const Ref &operator|=(T b) const { c->template st<T>(off, (T)(c->template ld<T>(off) | b)); return *this; }
	};
	template<class T> VG_INL Ref<T> A(unsigned off, int i) const { return Ref<T>{this, off + (unsigned)i * (unsigned)sizeof(T)}; }

	// ---- topology fields ----
	static VG_INL int twE(unsigned w, int i) { int e = (int)((w >> (8 * i)) & 0xff); return e == 0xff ? -1 : e; }
	static VG_INL int twR(unsigned w, int i) { return (int)((w >> (24 + 2 * i)) & 3); }
	static VG_INL unsigned twSetE(unsigned w, int i, int e) {
		const unsigned sh = 8u * (unsigned)i;
		return (w & ~(0xffu << sh)) | (((unsigned)e & 0xffu) << sh);
	}
	static VG_INL unsigned twSetR(unsigned w, int i, int r) {
		const unsigned sh = 24u + 2u * (unsigned)i;
		return (w & ~(3u << sh)) | ((unsigned)r << sh);
	}
	VG_INL int E(int v, int i) const { return twE(TW(v), i); }
	VG_INL int R(int v, int i) const { return twR(TW(v), i); }
#if VG_WDEV
	// edge and relation updates made in the vertex's own lane (no lane read)
	VG_INL void setE(int v, int i, int e) {
		const int h = v >> 6;
		const bool me = lane == (v & 63);
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int q = decltype(hc)::value;
			vt[q] = (me && h == q) ? twSetE(vt[q], i, e) : vt[q];
		});
	}
	VG_INL void setER(int v, int i, int e, int r) {
		const int h = v >> 6;
		const bool me = lane == (v & 63);
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int q = decltype(hc)::value;
			vt[q] = (me && h == q) ? twSetR(twSetE(vt[q], i, e), i, r) : vt[q];
		});
	}
#else
	VG_INL void setE(int v, int i, int e) { setTW(v, twSetE(TW(v), i, e)); }
	VG_INL void setER(int v, int i, int e, int r) { setTW(v, twSetR(twSetE(TW(v), i, e), i, r)); }
#endif
	VG_INL bool DEL(int v) const { return delm.get(v); }
	VG_INL void setDEL(int v) { delm.set(v); }
	VG_INL bool FL(int v) const { return flipm.get(v); }
	VG_INL void flip(int v) { flipm.toggle(v); }

	static VG_INL int cyc_up(int a) { return a == 2 ? 0 : a + 1; }

	VG_FN void setup(char *slab, int lane_, double max_len_sq) {
#if VG_WDEV
		lane = lane_;
		rs = __builtin_amdgcn_make_buffer_rsrc(slab, 0, (int)L::bytes, 0x00020000);
#else
		(void)lane_;
		wsb = slab;
#endif
		set_tol(max_len_sq);
	}
	// src/cell.cc:20-27: tol = 10 eps max_len_sq, tol_cu = tol^1.5, big_tol = 20 tol.
	VG_INL void set_tol(double max_len_sq) {
		const double eps = 2.220446049250313e-16;
		tol = 10.0 * eps * max_len_sq;
		tol_cu = tol * sqrt(tol);
		big_tol = 20.0 * tol;
	}

	// Vertex `from` takes index `to` (src/cell.cc:849-861).
	VG_FN void move_vertex(int from, int to) {
		setXYZ(to, X(from), Y(from), Z(from));
		const unsigned w = TW(from);
		setTW(to, w);
		setN3(to, N(from, 0), N(from, 1), N(from, 2));
		if (DEL(from)) delm.set(to); else delm.unset(to);
		flipm.unset(to);
		for (int i = 0; i < 3; i++) setE(twE(w, i), twR(w, i), to);
	}

	VG_INL bool room_v(int v) { if (v >= MAXV) { ovf = true; return false; } return true; }
	VG_INL bool push_ds(int v) { if (nds >= MAXS) { ovf = true; return false; } setDS(nds++, v); return true; }

	// A fresh vertex: coordinates, no edges yet, no marks.
	VG_INL void new_vertex(int v, double x, double y, double z) {
		setXYZ(v, x, y, z);
		setTW(v, 0x00ffffffu);
		delm.unset(v);
		flipm.unset(v);
	}

	// Box start cell, faces -1..-6 (src/cell.cc:294-318, 2567-2580); corner v
	// has bits x,y,z = v&1, v>>1&1, v>>2&1 and relations 2,1,0.
	VG_FN void init_box(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax) {
		ovf = false; fatal = false; n_cuts = 0;
		up = 0;
		p = 8;
		delm.clear(); flipm.clear();
		xmin *= 2; xmax *= 2; ymin *= 2; ymax *= 2; zmin *= 2; zmax *= 2;
#if VG_WDEV
		if (lane < 8) box_corner(lane, xmin, xmax, ymin, ymax, zmin, zmax);
#else
		for (int v = 0; v < 8; v++) box_corner(v, xmin, xmax, ymin, ymax, zmin, zmax);
#endif
	}
	// Corner v of the box: edges to v^(1+par), v^4, v^(2-par) with relations
	// 2, 1, 0, faces z, then x and y (x first for odd parity), where par is
	// the parity of v -- the reference's corner table in closed form.
	VG_INL void box_corner(int v, double xmin, double xmax, double ymin, double ymax, double zmin, double zmax) {
		const int par = __builtin_popcount((unsigned)v) & 1;
		const unsigned e = (unsigned)(v ^ (1 + par)) | (unsigned)(v ^ 4) << 8 | (unsigned)(v ^ (2 - par)) << 16;
		const int xf = -1 - (v & 1), yf = -3 - ((v >> 1) & 1);
#if VG_WDEV
		vx[0] = (v & 1) ? xmax : xmin; vy[0] = (v & 2) ? ymax : ymin; vz[0] = (v & 4) ? zmax : zmin;
		vt[0] = e | (6u << 24);
		vna[0] = -5 - (v >> 2); vnb[0] = par ? xf : yf; vnc[0] = par ? yf : xf;
#else
		setXYZ(v, (v & 1) ? xmax : xmin, (v & 2) ? ymax : ymin, (v & 4) ? zmax : zmin);
		setTW(v, e | (6u << 24));
		setN3(v, -5 - (v >> 2), par ? xf : yf, par ? yf : xf);
#endif
	}

	// The unit Voronoi cell of a periodic container as the start cell
	// (src/container_prd.hh:124-130, src/cell.cc:71-95).  A vertex of
	// another order than 3 leaves the cell to the per-lane tiers.
	VG_FN void init_unit(const UnitVert *__restrict__ u, int np) {
		ovf = false; fatal = false; n_cuts = 0;
		up = 0;
		delm.clear(); flipm.clear();
		if (np > MAXV) { ovf = true; return; }
		p = np;
#if VG_WDEV
		bool bad = false;
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			if (64 * h + lane < np) {
				const UnitVert a = u[64 * h + lane];
				bad = bad || a.nu != 3;
				vx[h] = a.x; vy[h] = a.y; vz[h] = a.z;
				vt[h] = unit_word(a);
				vna[h] = vnb[h] = vnc[h] = 0;
			}
		});
		if (__ballot(bad)) ovf = true;
#else
		for (int v = 0; v < np; v++) {
			const UnitVert a = u[v];
			if (a.nu != 3) { ovf = true; return; }
			setXYZ(v, a.x, a.y, a.z);
			setTW(v, unit_word(a));
			setN3(v, 0, 0, 0);
		}
#endif
	}
	static VG_INL unsigned unit_word(const UnitVert &a) {
		return (unsigned)(a.ed[0] & 0xff) | (unsigned)(a.ed[1] & 0xff) << 8 | (unsigned)(a.ed[2] & 0xff) << 16 |
		       (unsigned)((a.rel[0] & 3) | (a.rel[1] & 3) << 2 | (a.rel[2] & 3) << 4) << 24;
	}

	// Plane values of every vertex, the expression of src/cell.cc:2043-2059.
	// Returns true if some live vertex is not below -tol.
	VG_FN bool plane_values() {
#if VG_WDEV
		bool any = false;
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			if (h == 0 || p > 64 * h) {
				double a = vx[h] * px;
				a += vy[h] * py;
				a += vz[h] * pz - prsq;
				vtv[h] = a;
				if constexpr (K == 1) {
					// a plane that leaves every vertex below -tol (most planes)
					// needs no other mask
					negm = live_ballot(a < -tol);
					const unsigned long long live = p >= 64 ? ~0ull : (1ull << p) - 1;
					any = (live & ~negm) != 0;
					if (any) {
						posm = live_ballot(a > tol);
						nearm = live_ballot(a < -tol && a > -big_tol);
					}
				} else any = any || live_ballot(!(a < -tol), h) != 0;
			}
		});
		return any;
#else
		bool any = false;
		posm = negm = nearm = 0;
		for (int v = 0; v < MAXV; v++) {
			double a = vx[v] * px;
			a += vy[v] * py;
			a += vz[v] * pz - prsq;
			vtv[v] = a;
			if (v < p && !(a < -tol)) any = true;
			if (K == 1 && v < p) {
				if (a > tol) posm |= 1ull << v;
				if (a < -tol) negm |= 1ull << v;
				if (a < -tol && a > -big_tol) nearm |= 1ull << v;
			}
		}
		return any;
#endif
	}

	VG_INL unsigned m_test(int n, double &ans) {
		if (n < 0 || n >= MAXV) { fatal = ovf = true; ans = 0; return 1u; }
		ans = T(n);
		return ans < -tol ? 0u : (ans > tol ? 2u : 1u);
	}
	// m_test that also hands back the vertex's coordinates, and flags a kept
	// vertex within big_tol of the plane (src/cell.cc:2072-2083).
	VG_INL unsigned m_testx3(int n, double &ans, double &x, double &y, double &z) {
		if (n < 0 || n >= MAXV) { fatal = ovf = true; ans = 0; x = y = z = 0; return 1u; }
		ans = T(n);
		x = X(n); y = Y(n); z = Z(n);
		const unsigned r = ans < -tol ? 0u : (ans > tol ? 2u : 1u);
		if (r == 0 && ans > -big_tol) ovf = true;
		return r;
	}

	// src/cell.cc:522-595 (the member `up` is pushed in the second loop,
	// src/cell.cc:546-549, as there).  ta, tb, tc: the values of lp's
	// neighbours.
	VG_FN bool definite_max(int &lp, int &ls, double &l, double &u, unsigned &uw, double ta, double tb, double tc) {
		int tp = lp, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < 3; ts++) {
			qp = E(tp, ts);
			q = sel3(ts, ta, tb, tc);
			if (q > l - big_tol) break;
		}
		if (ts == 3) return true;
		int nst = 0;
		flip(lp); flip(qp);
		setDM(nst++, qp);
		ts++;
		while (ts < 3) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q > l - big_tol) {
				if (nst >= MAXS) { ovf = true; return true; }
				setDM(nst++, up); flip(up);
			}
			ts++;
		}
		int spp = 0;
		while (spp < nst) {
			tp = DM(spp++);
			for (ts = 0; ts < 3; ts++) {
				qp = E(tp, ts);
				if (FL(qp)) continue;
				qw = m_test(qp, q);
				if (q > l) {
					flip(lp);
					lp = tp; ls = ts;
					m_test(lp, l);
					up = qp; uw = qw; u = q;
					while (nst > 0) flip(DM(--nst));
					return false;
				}
				if (q > l - big_tol) {
					if (nst >= MAXS) { ovf = true; return true; }
					setDM(nst++, qp); flip(qp);
				}
			}
		}
		flip(lp);
		while (nst > 0) flip(DM(--nst));
		return true;
	}

	// The searches' step decisions of every vertex for the current plane
	// (one register set): for each slot vs the search entered by (-1 for
	// none; bits 2(vs+1) up, 8 + 2(vs+1) down), the first other slot whose
	// neighbour's value is greater (upward) or smaller (downward) than the
	// vertex's -- the comparisons of the reference's loops in
	// src/cell.cc:480-515 and 597-629 -- or 3 for none; bit 16: a neighbour
	// index the value test rejects (m_test).
	VG_FN void decide_search() {
#if VG_WDEV
		const unsigned w = vt[0];
		const double own = vtv[0];
		double t[3];
		unsigned bad = 0;
		for (int i = 0; i < 3; i++) {
			const unsigned e = (w >> (8 * i)) & 0xffu;
			bad |= e >= (unsigned)MAXV ? 1u : 0u;
			t[i] = bperm64(vtv[0], (int)((e & 63u) << 2));
		}
		unsigned d = bad << 16;
		for (int vs = -1; vs < 3; vs++) {
			unsigned a = 3, b = 3;
			for (int i = 2; i >= 0; i--) {
				if (i == vs) continue;
				if (t[i] > own) a = (unsigned)i;
				if (own > t[i]) b = (unsigned)i;
			}
			d |= a << (2 * (vs + 1)) | b << (8 + 2 * (vs + 1));
		}
		vdec = d;
#else
		for (int v = 0; v < p; v++) {
			const unsigned w = vt[v];
			const double own = vtv[v];
			double t[3];
			unsigned bad = 0;
			for (int i = 0; i < 3; i++) {
				const unsigned e = (w >> (8 * i)) & 0xffu;
				bad |= e >= (unsigned)MAXV ? 1u : 0u;
				t[i] = vtv[e & 63u];
			}
			unsigned d = bad << 16;
			for (int vs = -1; vs < 3; vs++) {
				unsigned a = 3, b = 3;
				for (int i = 2; i >= 0; i--) {
					if (i == vs) continue;
					if (t[i] > own) a = (unsigned)i;
					if (own > t[i]) b = (unsigned)i;
				}
				d |= a << (2 * (vs + 1)) | b << (8 + 2 * (vs + 1));
			}
			vdec[v] = d;
		}
#endif
	}
#if VG_WDEV
	VG_INL unsigned DEC(int v) const { return rlu(vdec, v); }
#else
	VG_INL unsigned DEC(int v) const { return vdec[v]; }
#endif
	// m_test's class from the plane's masks (one register set)
	VG_INL unsigned cls(int n) const {
		return ((posm >> n) & 1ull) ? 2u : (((negm >> n) & 1ull) ? 0u : 1u);
	}

	// search_upward over the decisions of decide_search: a step reads the
	// vertex's edges and decision word; a vertex with no greater neighbour
	// takes the reference's tie search (definite_max) with the values.
	VG_FN bool search_upward1(unsigned &uw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		unsigned w = 0;
		lp = up;
		for (;;) {
			w = TW(lp);
			const unsigned d = DEC(lp);
			if ((d >> 16) & 1u) { fatal = ovf = true; return false; }
			ls = (int)((d >> (2 * (vs + 1))) & 3u);
			if (ls != 3) {
				up = (int)((w >> (8 * ls)) & 0xffu);
				uw = cls(up);
			} else {
				l = T(lp);
				const int e0 = twE(w, 0), e1 = twE(w, 1), e2 = twE(w, 2);
				double ta, tb, tc;
				const unsigned wa = m_test(e0, ta), wb = m_test(e1, tb), wc = m_test(e2, tc);
				const int last = vs == 2 ? 1 : 2;
				up = sel3(last, e0, e1, e2);
				uw = sel3(last, wa, wb, wc);
				u = sel3(last, ta, tb, tc);
				if (definite_max(lp, ls, l, u, uw, ta, tb, tc)) { up = lp; return false; }
				w = TW(lp);
			}
			if (uw != 0) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = twR(w, ls); lp = up;
		}
		us = twR(w, ls);
		return true;
	}

	// search_downward over the decisions of decide_search (definite_min for
	// a vertex with no smaller neighbour).
	VG_FN bool search_downward1(unsigned &lw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		unsigned w = 0;
		for (;;) {
			w = TW(up);
			const unsigned d = DEC(up);
			if ((d >> 16) & 1u) { fatal = ovf = true; return false; }
			us = (int)((d >> (8 + 2 * (vs + 1))) & 3u);
			if (us != 3) {
				lp = (int)((w >> (8 * us)) & 0xffu);
				lw = cls(lp);
			} else {
				u = T(up);
				const int e0 = twE(w, 0), e1 = twE(w, 1), e2 = twE(w, 2);
				double ta, tb, tc;
				const unsigned wa = m_test(e0, ta), wb = m_test(e1, tb), wc = m_test(e2, tc);
				const int last = vs == 2 ? 1 : 2;
				lp = sel3(last, e0, e1, e2);
				lw = sel3(last, wa, wb, wc);
				l = sel3(last, ta, tb, tc);
				if (definite_min(lp, us, l, u, lw, ta, tb, tc)) return false;
				w = TW(up);
			}
			if (lw != 2) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = twR(w, us); up = lp;
		}
		ls = twR(w, us);
		return true;
	}

	// src/cell.cc:480-515 with definite_max folded in.
	VG_FN bool search_upward(unsigned &uw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		lp = up; l = u;
		for (;;) {
			const unsigned w = TW(lp);
			const int e0 = twE(w, 0), e1 = twE(w, 1), e2 = twE(w, 2);
			double ta, tb, tc;
			const unsigned wa = m_test(e0, ta), wb = m_test(e1, tb), wc = m_test(e2, tc);
			for (ls = 0; ls < 3; ls++) {
				if (ls == vs) continue;
				up = sel3(ls, e0, e1, e2);
				uw = sel3(ls, wa, wb, wc);
				u = sel3(ls, ta, tb, tc);
				if (u > l) break;
			}
			if (ls == 3 && definite_max(lp, ls, l, u, uw, ta, tb, tc)) { up = lp; return false; }
			if (uw != 0) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = R(lp, ls); lp = up; l = u;
		}
		us = R(lp, ls);
		return true;
	}

	// src/cell.cc:631-708 (the argument `lp` is pushed in the second loop,
	// src/cell.cc:654-657, as there).  ta, tb, tc: the values of up's
	// neighbours.
	VG_FN bool definite_min(int &lp, int &us, double &l, double &u, unsigned &lw, double ta, double tb, double tc) {
		int tp = up, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < 3; ts++) {
			qp = E(tp, ts);
			q = sel3(ts, ta, tb, tc);
			if (q < u + big_tol) break;
		}
		if (ts == 3) return true;
		int nst = 0;
		flip(up); flip(qp);
		setDM(nst++, qp);
		ts++;
		while (ts < 3) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q < u + big_tol) {
				if (nst >= MAXS) { ovf = true; return true; }
				setDM(nst++, lp); flip(lp);
			}
			ts++;
		}
		int spp = 0;
		while (spp < nst) {
			tp = DM(spp++);
			for (ts = 0; ts < 3; ts++) {
				qp = E(tp, ts);
				if (FL(qp)) continue;
				qw = m_test(qp, q);
				if (q < u) {
					flip(up);
					up = tp; us = ts;
					m_test(up, u);
					lp = qp; lw = qw; l = q;
					while (nst > 0) flip(DM(--nst));
					return false;
				}
				if (q < u + big_tol) {
					if (nst >= MAXS) { ovf = true; return true; }
					setDM(nst++, qp); flip(qp);
				}
			}
		}
		flip(up);
		while (nst > 0) flip(DM(--nst));
		return true;
	}

	// src/cell.cc:597-629 with definite_min folded in.
	VG_FN bool search_downward(unsigned &lw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		for (;;) {
			const unsigned w = TW(up);
			const int e0 = twE(w, 0), e1 = twE(w, 1), e2 = twE(w, 2);
			double ta, tb, tc;
			const unsigned wa = m_test(e0, ta), wb = m_test(e1, tb), wc = m_test(e2, tc);
			for (us = 0; us < 3; us++) {
				if (us == vs) continue;
				lp = sel3(us, e0, e1, e2);
				lw = sel3(us, wa, wb, wc);
				l = sel3(us, ta, tb, tc);
				if (u > l) break;
			}
			if (us == 3 && definite_min(lp, us, l, u, lw, ta, tb, tc)) return false;
			if (lw != 2) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = R(up, us); up = lp; u = l;
		}
		ls = R(up, us);
		return true;
	}

	// src/cell.cc:875-1364 for a plane that cuts edges cleanly.  Returns
	// true if the cell was deleted (or on overflow).
	VG_FN bool create_facet(int lp, int ls, double l, int us, double u, int pid) {
		int qp, qs, cp, cs, rp;
		unsigned lw;
		double q, r;
		if (!room_v(p)) return true;
		if (!push_ds(up)) return true;
		r = u / (u - l); l = 1 - r;
		new_vertex(p, X(lp) * r + X(up) * l, Y(lp) * r + Y(up) * l, Z(lp) * r + Z(up) * l);
		setN3(p, pid, N(up, us), N(lp, ls));
		setE(up, us, -1);
		setER(lp, ls, p, 1);
		setTW(p, twSetR(twSetE(0x00ffffffu, 1, lp), 1, ls));
		cs = 2;
		qs = cyc_up(us);
		qp = up; q = u;
		double qx = X(up), qy = Y(up), qz = Z(up);
		cp = p; rp = p; p++;
		int guard = 0;
		while (qp != up || qs != us) {
			if (++guard > 16 * MAXV) { fatal = ovf = true; return true; }
			lp = E(qp, qs);
			double lx, ly, lz;
			lw = m_testx3(lp, l, lx, ly, lz);
			if (ovf) return true;
			if (lw == 2) {
				qs = cyc_up(R(qp, qs));
				qp = lp;
				q = l;
				qx = lx; qy = ly; qz = lz;
				if (!push_ds(qp)) return true;
			} else if (lw == 0) {
				if (!room_v(p)) return true;
				r = q / (q - l); l = 1 - r;
				new_vertex(p, lx * r + qx * l, ly * r + qy * l, lz * r + qz * l);
				ls = R(qp, qs);
				setN3(p, pid, N(qp, qs), N(lp, ls));
				setTW(p, twSetR(twSetR(twSetE(twSetE(0x00ffffffu, 0, cp), 1, lp), 0, cs), 1, ls));
				setER(lp, ls, p, 1);
				setER(cp, cs, p, 0);
				setE(qp, qs, -1);
				qs = cyc_up(qs);
				cp = p++;
				cs = 2;
			} else {
				ovf = true;      // a vertex on the plane: marginal case
				return true;
			}
		}
		setER(cp, cs, rp, 0);
		setER(rp, 0, cp, cs);
		return false;
	}

	// create_facet for one register set, in two passes.  The walk around
	// the cut (the reference's loop of src/cell.cc:875-1364, in the same
	// order) classifies vertices by the masks of plane_values and only
	// records, per new vertex, its cut edge (positive vertex and slot,
	// negative vertex and slot) in lane `new vertex - p` of vrec; the cut
	// slot of the positive vertex takes the new vertex's index (the
	// reference writes -1 there; both stop the deletion search, which treats
	// indices >= the old vertex count as cut).  A vector pass then builds
	// every new vertex at once -- coordinates by the reference's
	// interpolation, edges to its neighbours on the facet and to the
	// negative vertex, face ids -- and points the negative vertices' cut
	// slots at them.  The result is the state the reference's vertex-by-vertex
	// construction leaves.  A cut edge the walk did not reach (the cell is not
	// convex to working precision) sets ovf.  Returns true if the cell was
	// deleted (or on overflow).
	VG_FN bool create_facet1(int lp, int ls, int us, int pid) {
		static_assert(K == 1, "one register set");
		if (!room_v(p)) return true;
		if (!push_ds(up)) return true;
		const int p0 = p;
		unsigned w = TW(up);
		put_rec(0, (unsigned)up | (unsigned)us << 8 | (unsigned)lp << 16 | (unsigned)ls << 24);
		w = twSetE(w, us, p0);
		setTW(up, w);
		int m = 1, qp = up, qs = cyc_up(us), guard = 0;
		while (qp != up || qs != us) {
			if (++guard > 16 * MAXV) { fatal = ovf = true; return true; }
			const int n = (int)((w >> (8 * qs)) & 0xffu);
			if (n >= p0) { fatal = ovf = true; return true; }   // -1, or a slot already cut
			const unsigned long long b = 1ull << n;
			if (posm & b) {
				qs = cyc_up(twR(w, qs));
				qp = n;
				w = TW(qp);
				if (!push_ds(qp)) return true;
			} else if (negm & b) {
				if (nearm & b) { ovf = true; return true; }   // a kept vertex within big_tol
				if (!room_v(p0 + m)) return true;
				put_rec(m, (unsigned)qp | (unsigned)qs << 8 | (unsigned)n << 16 | (unsigned)twR(w, qs) << 24);
				w = twSetE(w, qs, p0 + m);
				setTW(qp, w);
				m++;
				qs = cyc_up(qs);
			} else {
				ovf = true;      // a vertex on the plane: marginal case
				return true;
			}
		}
		p = p0 + m;
		build_facet(p0, m, pid);
		if (!link_cut(p0, m)) { ovf = true; return true; }
		return false;
	}

	// The new vertices p0 .. p0+m-1 from their records (see create_facet1):
	// new vertex j lies on the cut edge of record j, between the facet's
	// vertices j-1 and j+1 (relations 2 and 0) and joined to the negative
	// vertex (its relation the negative vertex's slot).
	VG_FN void build_facet(int p0, int m, int pid) {
#if VG_WDEV
		const int j = lane - p0;
		const bool act = j >= 0 && j < m;
		const unsigned r = bperm(vrec, (act ? j : 0) << 2);
		const int qp = (int)(r & 0xffu), qs = (int)((r >> 8) & 3u), lp = (int)((r >> 16) & 0xffu),
			  ls = (int)(r >> 24);
		const int aq = qp << 2, al = lp << 2;
		const double qx = bperm64(vx[0], aq), qy = bperm64(vy[0], aq), qz = bperm64(vz[0], aq);
		const double qv = bperm64(vtv[0], aq);
		const double lx = bperm64(vx[0], al), ly = bperm64(vy[0], al), lz = bperm64(vz[0], al);
		const double lv = bperm64(vtv[0], al);
		const int na = (int)bperm((unsigned)vna[0], aq), nb = (int)bperm((unsigned)vnb[0], aq),
			  nc = (int)bperm((unsigned)vnc[0], aq);
		const int la = (int)bperm((unsigned)vna[0], al), lb = (int)bperm((unsigned)vnb[0], al),
			  lc = (int)bperm((unsigned)vnc[0], al);
		if (act) {
			const double rr = qv / (qv - lv), l1 = 1 - rr;
			vx[0] = lx * rr + qx * l1;
			vy[0] = ly * rr + qy * l1;
			vz[0] = lz * rr + qz * l1;
			const int prv = p0 + (j == 0 ? m - 1 : j - 1), nxt = p0 + (j + 1 == m ? 0 : j + 1);
			vt[0] = (unsigned)prv | (unsigned)lp << 8 | (unsigned)nxt << 16 | (2u | (unsigned)ls << 2) << 24;
			vna[0] = pid;
			vnb[0] = sel3(qs, na, nb, nc);
			vnc[0] = sel3(ls, la, lb, lc);
		}
#else
		for (int j = 0; j < m; j++) {
			const unsigned r = vrec[j];
			const int qp = (int)(r & 0xffu), qs = (int)((r >> 8) & 3u), lp = (int)((r >> 16) & 0xffu),
				  ls = (int)(r >> 24);
			const double qv = vtv[qp], lv = vtv[lp];
			const double rr = qv / (qv - lv), l1 = 1 - rr;
			const int v = p0 + j;
			vx[v] = vx[lp] * rr + vx[qp] * l1;
			vy[v] = vy[lp] * rr + vy[qp] * l1;
			vz[v] = vz[lp] * rr + vz[qp] * l1;
			const int prv = p0 + (j == 0 ? m - 1 : j - 1), nxt = p0 + (j + 1 == m ? 0 : j + 1);
			vt[v] = (unsigned)prv | (unsigned)lp << 8 | (unsigned)nxt << 16 | (2u | (unsigned)ls << 2) << 24;
			vn[v][0] = pid;
			vn[v][1] = vn[qp][qs];
			vn[v][2] = vn[lp][ls];
		}
#endif
		const unsigned long long nm = (m >= 64 ? ~0ull : (1ull << m) - 1) << p0;
		delm.w[0] &= ~nm;
		flipm.w[0] &= ~nm;
	}

	// Points every cut slot of the negative vertices (below p0) at the new
	// vertex on that edge: the positive end's slot holds it since the walk.
	// False if some cut edge was not walked (its count differs from m).
	VG_FN bool link_cut(int p0, int m) {
#if VG_WDEV
		const unsigned w = vt[0];
		const bool neg = lane < p0 && ((negm >> lane) & 1ull);
		int cnt = 0;
		unsigned nw = w;
		for (int i = 0; i < 3; i++) {
			const unsigned e = (w >> (8 * i)) & 0xffu;
			const bool cut = neg && e < (unsigned)p0 && ((posm >> e) & 1ull);
			const unsigned we = bperm(w, (int)((cut ? e : 0u) << 2));
			const unsigned back = (w >> (24 + 2 * i)) & 3u;
			const unsigned nv = (we >> (8 * back)) & 0xffu;
			if (cut) nw = (nw & ~(0xffu << (8 * i)) & ~(3u << (24 + 2 * i))) | nv << (8 * i) | 1u << (24 + 2 * i);
			cnt += __builtin_popcountll(__ballot(cut));
		}
		vt[0] = nw;
		return cnt == m;
#else
		int cnt = 0;
		unsigned nw[64];
		for (int v = 0; v < p0; v++) {
			const unsigned w = vt[v];
			nw[v] = w;
			if (!((negm >> v) & 1ull)) continue;
			for (int i = 0; i < 3; i++) {
				const unsigned e = (w >> (8 * i)) & 0xffu;
				if (!(e < (unsigned)p0 && ((posm >> e) & 1ull))) continue;
				const unsigned back = (w >> (24 + 2 * i)) & 3u;
				const unsigned nv = (vt[e] >> (8 * back)) & 0xffu;
				nw[v] = (nw[v] & ~(0xffu << (8 * i)) & ~(3u << (24 + 2 * i))) | nv << (8 * i) | 1u << (24 + 2 * i);
				cnt++;
			}
		}
		for (int v = 0; v < p0; v++) vt[v] = nw[v];
		return cnt == m;
#endif
	}

	// src/cell.cc:718-870.  Returns false if the cell was deleted or a
	// capacity was exceeded (ovf set).
	VG_FN bool nplane(double x, double y, double z, double rsq, int pid) {
		bool r = nplane_impl(x, y, z, rsq, pid);
		return ovf ? false : r;
	}

	VG_FN bool nplane_impl(double x, double y, double z, double rsq, int pid) {
		n_cuts++;
		int j, lp = up, cp, qp;
		int us = 0, ls = 0;
		unsigned uw, lw;
		double u, l = 0;
		up = 0;
		px = x; py = y; pz = z; prsq = rsq;
		if (!plane_values()) return true;   // every vertex below -tol: no cut
		if constexpr (K == 1) decide_search();
		uw = m_test(up, u);
		if (uw == 2) {
			if constexpr (K == 1) {
				if (!search_downward1(lw, lp, ls, us, l, u)) return false;
			} else {
				if (!search_downward(lw, lp, ls, us, l, u)) return false;
			}
			if (lw == 1) { ovf = true; return false; }     // vertex on the plane
		} else if (uw == 0) {
			if constexpr (K == 1) {
				if (!search_upward1(uw, lp, ls, us, l, u)) return true;
			} else {
				if (!search_upward(uw, lp, ls, us, l, u)) return true;
			}
			if (uw == 1) { ovf = true; return false; }     // vertex on the plane
		} else {
			ovf = true;                                      // vertex 0 on the plane
			return false;
		}
		if (ovf) return false;
		nds = 0;
		const int p0 = p;
		if constexpr (K == 1) {
			if (create_facet1(lp, ls, us, pid)) return false;
		} else {
			if (create_facet(lp, ls, l, us, u, pid)) return false;
		}
		// Delete the cut vertices and everything reachable from them
		// (src/cell.cc:786-863); a cut slot holds -1 or (one register set)
		// the index of a new vertex, both at least p0.
		int d = 0;
		while (d < nds) {
			j = DS(d);
			if (!DEL(j)) { setDEL(j); d++; }
			else setDS(d, DS(--nds));
		}
		bool search = true;
		if constexpr (K == 1) {
			// Every cut edge was walked (create_facet1), so a non-cut edge
			// of a vertex above the plane ends above or on it.  A vertex on
			// the plane that the walk did not meet is left to the general
			// tier; otherwise the search only has vertices above the plane
			// to reach, none once all of them are marked.
			const unsigned long long low = p0 >= 64 ? ~0ull : (1ull << p0) - 1;
			if (low & ~posm & ~negm) { ovf = true; return false; }
			search = __builtin_popcountll(delm.w[0] & low) != __builtin_popcountll(posm);
		}
		if (search)
			for (d = 0; d < nds; d++) {
				cp = DS(d);
				const unsigned w = TW(cp);
				for (int e = 0; e < 3; e++) {
					qp = twE(w, e);
					if (qp != -1 && (K > 1 || qp < p0) && !DEL(qp)) {
						if (!push_ds(qp)) return false;
						setDEL(qp);
					}
				}
			}
		return compact();
	}

	// Compaction, src/cell.cc:843-862: the deleted slots are refilled from
	// the end of the vertex list, delete-stack entries taken last first.  The
	// loop runs on the marks alone and records, per slot, the original vertex
	// it ends up holding (org) and, per moved vertex, the slot it ends in
	// (inv); one permutation then gathers the vertex records and renames
	// every edge through inv -- the state the reference's vertex-by-vertex
	// moves (each rewriting the back pointers of the moved vertex's
	// neighbours) leave behind.
	// (With four register sets the permutation's temporaries do not fit
	// the register budget: the moves are made one by one.)
	VG_FN bool compact() {
		if constexpr (K == 4) {
			up = 0;
			while (nds > 0) {
				--p;
				while (DEL(p)) {
					if (p == 0) { fatal = ovf = true; return false; }
					--p;
				}
				up = DS(--nds);
				if (up < p) move_vertex(p, up);
				else up = p++;
			}
			return true;
		}
#if VG_WDEV
		if constexpr (K == 1) return compact1();
		int org[K], inv[K];
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			org[h] = 64 * h + lane;
			inv[h] = 64 * h + lane;
		});
#else
		int org[MAXV], inv[MAXV];
		for (int v = 0; v < MAXV; v++) org[v] = inv[v] = v;
#endif
		bool moved = false;
		up = 0;
		while (nds > 0) {
			--p;
			while (DEL(p)) {
				if (p == 0) { fatal = ovf = true; return false; }
				--p;
			}
			up = DS(--nds);
			if (up < p) {
				// slot p's vertex moves to slot up
#if VG_WDEV
				const int o = rli(pickk<K>(org, p >> 6), p & 63);
				putk<K>(org, up >> 6, lane == (up & 63), o);
				putk<K>(inv, o >> 6, lane == (o & 63), up);
#else
				const int o = org[p];
				org[up] = o;
				inv[o] = up;
#endif
				delm.unset(up);
				flipm.unset(up);
				moved = true;
			} else up = p++;
		}
		if (moved) permute(org, inv);
		return true;
	}

#if VG_WDEV
	// compact() for one register set with the replay's scalar chain kept
	// short: the last live vertex below p is found with a bit scan of the
	// deletion marks, the first 16 delete-stack entries are read into scalar
	// registers at once, and a moved vertex's original index is read back
	// from `org` only when its slot was itself filled earlier in the replay
	// (otherwise it is the slot).  The moves, and so the permutation, are
	// the same as the vertex-by-vertex loop's.
	VG_FN bool compact1() {
		int org[1] = {lane}, inv[1] = {lane};
		unsigned long long dm = delm.w[0], fm = flipm.w[0], tgt = 0;
		const unsigned d0 = rlu(vds, 0), d1 = rlu(vds, 1), d2 = rlu(vds, 2), d3 = rlu(vds, 3);
		bool moved = false;
		up = 0;
		while (nds > 0) {
			const unsigned long long below = p >= 64 ? ~0ull : (1ull << p) - 1;
			const unsigned long long live = ~dm & below;
			if (live == 0) { fatal = ovf = true; return false; }
			p = 63 - __builtin_clzll(live);
			--nds;
			if (nds < 16) {
				const unsigned wd = nds < 4 ? d0 : (nds < 8 ? d1 : (nds < 12 ? d2 : d3));
				up = (int)((wd >> (8 * (nds & 3))) & 0xffu);
			} else up = DS(nds);
			if (up < p) {
				// slot p's vertex moves to slot up
				const int o = ((tgt >> p) & 1ull) ? rli(org[0], p) : p;
				org[0] = lane == up ? o : org[0];
				inv[0] = lane == o ? up : inv[0];
				dm &= ~(1ull << up);
				fm &= ~(1ull << up);
				tgt |= 1ull << up;
				moved = true;
			} else up = p++;
		}
		delm.w[0] = dm;
		flipm.w[0] = fm;
		if (moved) permute(org, inv);
		return true;
	}

	static VG_INL unsigned bperm(unsigned v, int addr) {
		return (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)v);
	}
	// value of a[] at vertex idx (per lane), 0 <= idx < 256
	template<class T>
	static VG_INL unsigned lookup(const T (&a)[K], unsigned idx) {
		const int addr = (int)((idx & 63u) << 2);
		const unsigned hs = idx >> 6;
		unsigned r = 0;
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int q = decltype(hc)::value;
			const unsigned g = bperm((unsigned)a[q], addr);
			r = hs == (unsigned)q ? g : r;
		});
		return r;
	}
	template<class T>
	static VG_INL void gather32(T (&a)[K], const int (&src)[K]) {
		T o[K];
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			o[h] = (T)lookup<T>(a, (unsigned)src[h]);
		});
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			a[h] = o[h];
		});
	}
	static VG_INL void gather64(double (&a)[K], const int (&src)[K]) {
		unsigned lo[K], hi[K];
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			const unsigned long long b = __builtin_bit_cast(unsigned long long, a[h]);
			lo[h] = (unsigned)b;
			hi[h] = (unsigned)(b >> 32);
		});
		gather32(lo, src);
		gather32(hi, src);
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			a[h] = __builtin_bit_cast(double, ((unsigned long long)hi[h] << 32) | lo[h]);
		});
	}
	VG_FN void permute(const int (&org)[K], const int (&inv)[K]) {
		gather64(vx, org); gather64(vy, org); gather64(vz, org);
		gather32(vt, org); gather32(vna, org); gather32(vnb, org); gather32(vnc, org);
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			// every lane takes part in every lookup (a lane outside the
			// exec mask would read as zero to the others)
			const unsigned w = vt[h];
			const unsigned e0 = w & 0xffu, e1 = (w >> 8) & 0xffu, e2 = (w >> 16) & 0xffu;
			const unsigned l0 = lookup<int>(inv, e0), l1 = lookup<int>(inv, e1), l2 = lookup<int>(inv, e2);
			const unsigned n0 = e0 == 0xffu ? 0xffu : l0 & 0xffu;
			const unsigned n1 = e1 == 0xffu ? 0xffu : l1 & 0xffu;
			const unsigned n2 = e2 == 0xffu ? 0xffu : l2 & 0xffu;
			vt[h] = (w & 0xff000000u) | n0 | n1 << 8 | n2 << 16;
		});
	}
#else
	VG_FN void permute(const int *org, const int *inv) {
		double ax[MAXV], ay[MAXV], az[MAXV];
		unsigned at[MAXV];
		int an[MAXV][3];
		for (int v = 0; v < MAXV; v++) {
			const int o = org[v];
			ax[v] = vx[o]; ay[v] = vy[o]; az[v] = vz[o]; at[v] = vt[o];
			an[v][0] = vn[o][0]; an[v][1] = vn[o][1]; an[v][2] = vn[o][2];
		}
		for (int v = 0; v < MAXV; v++) {
			unsigned w = at[v];
			for (int i = 0; i < 3; i++) {
				const int e = (int)((w >> (8 * i)) & 0xffu);
				if (e != 0xff && e < MAXV) w = twSetE(w, i, inv[e]);
			}
			vx[v] = ax[v]; vy[v] = ay[v]; vz[v] = az[v]; vt[v] = w;
			vn[v][0] = an[v][0]; vn[v][1] = an[v][1]; vn[v][2] = an[v][2];
		}
	}
#endif

	// src/cell.cc:1907-1917: the largest squared vertex distance (maxima are
	// exact, so the order of the comparisons does not matter).
	VG_FN double max_radius_squared() {
#if VG_WDEV
		double s = 0;
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			if (h == 0 || p > 64 * h) {
				double t = vx[h] * vx[h];
				t += vy[h] * vy[h];
				t += vz[h] * vz[h];
				if (64 * h + lane < p && t > s) s = t;
			}
		});
		for (int o = 32; o > 0; o >>= 1) {
			const double t = __shfl_xor(s, o);
			s = t > s ? t : s;
		}
		return rfd(s);
#else
		double r = X(0) * X(0) + Y(0) * Y(0) + Z(0) * Z(0), s;
		for (int v = 1; v < p; v++) {
			double a = X(v), b = Y(v), c = Z(v);
			s = a * a;
			s += b * b;
			s += c * c;
			if (s > r) r = s;
		}
		return r;
#endif
	}

	// plane_intersects(_guess), src/cell.cc:2366-2410: true if the value of
	// the start vertex is >= rsq or that of any vertex exceeds rsq (the
	// guess samples are vertices too); `up` follows the guess walk.
	VG_FN bool plane_probe(double x, double y, double z, double rsq, bool guess) {
		if (guess) up = 0;
#if VG_WDEV
		double gl[K];
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			gl[h] = (h == 0 || p > 64 * h) ? x * vx[h] + y * vy[h] + z * vz[h] : 0.0;
		});
		double g = rld(pickk<K>(gl, up >> 6), up & 63);
		if (!(g < rsq)) return true;
		if (guess) {
			int ca = 1, cc = p >> 3, mp = 1;
			while (ca < cc) {
				const double m = rld(pickk<K>(gl, mp >> 6), mp & 63);
				if (m > g) {
					if (m > rsq) return true;
					g = m; up = mp;
				}
				ca += mp++;
			}
		}
		bool any = false;
		eachk<K>([&](auto hc) __attribute__((always_inline)) {
			constexpr int h = decltype(hc)::value;
			any = any || ((h == 0 || p > 64 * h) && live_ballot(gl[h] > rsq, h) != 0);
		});
		return any;
#else
		double g = x * X(up) + y * Y(up) + z * Z(up);
		if (!(g < rsq)) return true;
		if (guess) {
			int ca = 1, cc = p >> 3, mp = 1;
			while (ca < cc) {
				double m = x * X(mp) + y * Y(mp) + z * Z(mp);
				if (m > g) {
					if (m > rsq) return true;
					g = m; up = mp;
				}
				ca += mp++;
			}
		}
		for (int tp = 0; tp < p; tp++)
			if (x * X(tp) + y * Y(tp) + z * Z(tp) > rsq) return true;
		return false;
#endif
	}

	// Volume, centroid, face count and neighbour ids in one face walk
	// (src/cell.cc:1584-1616, 1861-1901, 2168-2187, 2660-2678).  Walked
	// edges are marked in the visited masks (the reference negates them).
	VG_INL bool vis(int v, int i) const {
		const bool a = vis0.get(v), b = vis1.get(v), c = vis2.get(v);
		return sel3(i, a, b, c);
	}
	VG_INL void setvis(int v, int i) {
		vis0.set_if(i == 0, v); vis1.set_if(i == 1, v); vis2.set_if(i == 2, v);
	}
	template<class NbSink>
	VG_FN void outputs(double &vol, double &cx, double &cy, double &cz, int &nf, NbSink &nb) {
		const double fe = 1 / 48.0;
		double v = 0, tvol;
		double ux, uy, uz, vx_, vy_, vz_, wx, wy, wz;
		const double x0 = X(0), y0 = Y(0), z0 = Z(0);
		vis0.clear(); vis1.clear(); vis2.clear();
		cx = cy = cz = 0;
		nf = 0;
		for (int i = 1; i < p; i++) {
			const double xi = X(i), yi = Y(i), zi = Z(i);
			ux = x0 - xi;
			uy = y0 - yi;
			uz = z0 - zi;
			for (int j = 0; j < 3; j++) {
				if (vis(i, j)) continue;
				const unsigned wi = TW(i);
				int k = twE(wi, j);
				nb.put(nf++, N(i, j));
				setvis(i, j);
				int l = cyc_up(twR(wi, j));
				vx_ = X(k) - x0;
				vy_ = Y(k) - y0;
				vz_ = Z(k) - z0;
				unsigned wk = TW(k);
				int m = twE(wk, l);
				setvis(k, l);
				int guard = 0;
				while (m != i) {
					if (++guard > 4 * MAXV) { fatal = ovf = true; return; }
					int n = cyc_up(twR(wk, l));
					wx = X(m) - x0;
					wy = Y(m) - y0;
					wz = Z(m) - z0;
					tvol = ux * vy_ * wz + uy * vz_ * wx + uz * vx_ * wy - uz * vy_ * wx - uy * vx_ * wz - ux * vz_ * wy;
					v += tvol;
					cx += (wx + vx_ - ux) * tvol;
					cy += (wy + vy_ - uy) * tvol;
					cz += (wz + vz_ - uz) * tvol;
					k = m; l = n; vx_ = wx; vy_ = wy; vz_ = wz;
					wk = TW(k);
					m = twE(wk, l);
					setvis(k, l);
				}
			}
		}
		vol = v * fe;
		if (v > tol_cu) {
			v = 0.125 / v;
			cx = cx * v + 0.5 * x0;
			cy = cy * v + 0.5 * y0;
			cz = cz * v + 0.5 * z0;
		} else cx = cy = cz = 0;
	}
};

typedef WCellT<1> WCell;     // up to 64 vertices
typedef WCellT<2> WCell2;    // up to 128 vertices
typedef WCellT<4> WCell4;    // up to 254 vertices

}  // namespace vgpu
