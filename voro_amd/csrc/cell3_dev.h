// Fast-tier cell: the same algorithm as cell_dev.h (src/cell.cc nplane
// family and outputs), specialised to cells whose vertices all have order 3,
// which is every vertex of a cell built from points in general position.
// It also leaves to the general tier every cell that meets a marginal
// configuration -- a cutting plane through a vertex (|plane value| <= tol),
// or a kept vertex within big_tol of the plane, the cases that take the
// reference into its marginal-vertex and extra-search code -- as well as a
// vertex count above MAXV or a full stack: such a cell sets ovf and is
// recomputed from scratch by the general tier, so results never depend on
// the tier.  Points in general position never meet a marginal case.
//
// Storage is per lane and contiguous: one 32-byte record per vertex
//   +0  x, y, z (doubles, twice the coordinates, as the reference stores)
//   +24 e0 e1 e2   neighbouring vertex indices (0xff = the reference's -1)
//   +27 r0 r1 r2   relation of each edge (2 bits each); bit 6 = deletion
//                  mark (the reference's ed[v][nu[v]] = -1)
//   +28 bp         back pointer / marker (int16: v, -1-v flipped, -p, -1)
//   +30 visited    face-walk marks (3 bits; the reference negates edges)
// so that a vertex's coordinates and topology share one 32-byte sector,
// plus 3 neighbour ids per vertex, the delete stack and the marginal-search
// stack, the output staging row, and the block-search window and ring (used
// only when a cell reaches the block search).  Since a vertex of order 3 is
// all this tier holds, the reference's per-order row pools are empty apart
// from the order-3 pool, whose row order does not influence results.
#pragma once
#include "vgpu_common.h"

namespace vgpu {

struct Tier3 {
	enum { MAXV = 128, MAXS = 96, W = 6, MAXQ = 256, MAXF = 64 };
};

template<int ST_UNUSED = 0>
struct Layout3 {
	enum { MAXV = Tier3::MAXV, MAXS = Tier3::MAXS, W = Tier3::W, MAXQ = Tier3::MAXQ, MAXF = Tier3::MAXF };
	enum { WD = 2 * W + 1, WBITS = WD * WD * WD, WWORDS = (WBITS + 31) / 32 };
	static constexpr unsigned o_rec = 0;                              // MAXV x 32
	static constexpr unsigned o_ne = o_rec + MAXV * 32;               // MAXV x 3 ints
	static constexpr unsigned o_ds = o_ne + MAXV * 12;                // bytes
	static constexpr unsigned o_dm = o_ds + MAXS;
	static constexpr unsigned o_nb = (o_dm + MAXS + 15) & ~15u;       // MAXF ints (output staging)
	static constexpr unsigned o_bm = o_nb + MAXF * 4;                 // WWORDS
	static constexpr unsigned o_bms = o_bm + WWORDS * 4;              // WWORDS stamps
	static constexpr unsigned o_qu = o_bms + WWORDS * 4;              // MAXQ ints
	static constexpr unsigned lane_bytes = (o_qu + MAXQ * 4 + 127) & ~127u;
	static constexpr size_t bytes = (size_t)lane_bytes * 64;          // one wavefront slab
	static constexpr unsigned o_nbs = o_nb;
	enum { NBCAP = MAXF };
};

struct VCell3 {
	typedef Layout3<> L;
	enum { MAXV = L::MAXV, MAXS = L::MAXS };

	char *wsb;          // CPU builds: this lane's region
	unsigned lo;        // byte offset of this lane's region in the slab
#if defined(__HIP_DEVICE_COMPILE__)
	__amdgpu_buffer_rsrc_t rs;
#endif

	int p, up;
	double px, py, pz, prsq;
	double tol, tol_cu, big_tol;
	int nds;
	bool ovf, fatal;
	int n_cuts;

	// ---- raw loads / stores at a byte offset within this lane's region ----
	template<class T> VG_INL T ld(unsigned off) const {
#if defined(__HIP_DEVICE_COMPILE__)
		if constexpr (sizeof(T) == 1) return (T)__builtin_amdgcn_raw_buffer_load_b8(rs, lo + off, 0, 0);
		else if constexpr (sizeof(T) == 2) return (T)__builtin_amdgcn_raw_buffer_load_b16(rs, lo + off, 0, 0);
		else if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, lo + off, 0, 0));
		else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, lo + off, 0, 0));
#else
		T v;
		__builtin_memcpy(&v, wsb + off, sizeof(T));
		return v;
#endif
	}
	template<class T> VG_INL void st(unsigned off, T v) const {
#if defined(__HIP_DEVICE_COMPILE__)
		if constexpr (sizeof(T) == 1) __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v, rs, lo + off, 0, 0);
		else if constexpr (sizeof(T) == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, rs, lo + off, 0, 0);
		else if constexpr (sizeof(T) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, lo + off, 0, 0);
		else {
			typedef unsigned u2 __attribute__((ext_vector_type(2)));
			__builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, lo + off, 0, 0);
		}
#else
		__builtin_memcpy(wsb + off, &v, sizeof(T));
#endif
	}

	// x and y of a record in one 16-byte load (they share a 16-byte slot).
	VG_INL void ldXY(unsigned off, double &x, double &y) const {
#if defined(__HIP_DEVICE_COMPILE__)
		typedef unsigned u4 __attribute__((ext_vector_type(4)));
		u4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, lo + off, 0, 0);
		typedef unsigned u2 __attribute__((ext_vector_type(2)));
		u2 a = {v.x, v.y}, b = {v.z, v.w};
		x = __builtin_bit_cast(double, a);
		y = __builtin_bit_cast(double, b);
#else
		__builtin_memcpy(&x, wsb + off, 8);
		__builtin_memcpy(&y, wsb + off + 8, 8);
#endif
	}

	// Proxy used by the shared block-search code (BM/BS/QU arrays).
	template<class T>
	struct Ref {
		const VCell3 *c;
		unsigned off;
		VG_INL operator T() const { return c->template ld<T>(off); }
		VG_INL const Ref &operator=(T v) const { c->template st<T>(off, v); return *this; }
		VG_INL const Ref &operator|=(T b) const { c->template st<T>(off, (T)(c->template ld<T>(off) | b)); return *this; }
	};
	template<class T> VG_INL Ref<T> A(unsigned off, int i) const { return Ref<T>{this, off + (unsigned)i * (unsigned)sizeof(T)}; }

	// ---- vertex record fields ----
	VG_INL unsigned rec(int v) const { return L::o_rec + (unsigned)v * 32u; }
	VG_INL double X(int v) const { return ld<double>(rec(v)); }
	VG_INL double Y(int v) const { return ld<double>(rec(v) + 8); }
	VG_INL double Z(int v) const { return ld<double>(rec(v) + 16); }
	VG_INL void setXYZ(int v, double x, double y, double z) const {
		st<double>(rec(v), x); st<double>(rec(v) + 8, y); st<double>(rec(v) + 16, z);
	}
	VG_INL int E(int v, int i) const { int e = ld<unsigned char>(rec(v) + 24 + i); return e == 0xff ? -1 : e; }
	VG_INL void setE(int v, int i, int e) const { st<unsigned char>(rec(v) + 24 + i, (unsigned char)(e & 0xff)); }
	VG_INL int R(int v, int i) const { return (ld<unsigned char>(rec(v) + 27) >> (2 * i)) & 3; }
	VG_INL void setR(int v, int i, int r) const {
		unsigned b = ld<unsigned char>(rec(v) + 27);
		b = (b & ~(3u << (2 * i))) | ((unsigned)r << (2 * i));
		st<unsigned char>(rec(v) + 27, (unsigned char)b);
	}
	VG_INL bool DEL(int v) const { return (ld<unsigned char>(rec(v) + 27) >> 6) & 1; }
	VG_INL void setDEL(int v) const { st<unsigned char>(rec(v) + 27, (unsigned char)(ld<unsigned char>(rec(v) + 27) | 64u)); }
	VG_INL int BP(int v) const { return ld<short>(rec(v) + 28); }
	VG_INL void setBP(int v, int b) const { st<short>(rec(v) + 28, (short)b); }
	// whole topology word (edges, relations, bp, visited) as one 8-byte move
	VG_INL unsigned long long TW(int v) const { return ld<unsigned long long>(rec(v) + 24); }
	VG_INL void setTW(int v, unsigned long long w) const { st<unsigned long long>(rec(v) + 24, w); }
	VG_INL int N(int v, int i) const { return ld<int>(L::o_ne + (unsigned)(v * 3 + i) * 4u); }
	VG_INL void setN(int v, int i, int x) const { st<int>(L::o_ne + (unsigned)(v * 3 + i) * 4u, x); }
	VG_INL int DS(int i) const { return ld<unsigned char>(L::o_ds + i); }
	VG_INL int DM(int i) const { return ld<unsigned char>(L::o_dm + i); }
	VG_INL void setDS(int i, int v) const { st<unsigned char>(L::o_ds + i, (unsigned char)v); }
	VG_INL void setDM(int i, int v) const { st<unsigned char>(L::o_dm + i, (unsigned char)v); }

	// Fields of a topology word held in registers.
	static VG_INL int twE(unsigned long long w, int i) { int e = (int)((w >> (8 * i)) & 0xff); return e == 0xff ? -1 : e; }
	static VG_INL int twR(unsigned long long w, int i) { return (int)((w >> (24 + 2 * i)) & 3); }
	static VG_INL int twBP(unsigned long long w) { return (int)(short)(unsigned short)(w >> 32); }

	static VG_INL int cyc_up(int a) { return a == 2 ? 0 : a + 1; }
	static VG_INL int cyc_down(int a) { return a == 0 ? 2 : a - 1; }
	VG_INL void flip(int v) const { setBP(v, -1 - BP(v)); }

	VG_FN void setup(char *slab, int lane, double max_len_sq) {
		lo = (unsigned)lane * L::lane_bytes;
#if defined(__HIP_DEVICE_COMPILE__)
		wsb = slab;
		rs = __builtin_amdgcn_make_buffer_rsrc(slab, 0, (int)L::bytes, 0x00020000);
#else
		wsb = slab + lo;
		lo = 0;
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

	VG_INL bool room_v(int v) { if (v >= MAXV) { ovf = true; return false; } return true; }
	VG_INL bool push_ds(int v) { if (nds >= MAXS) { ovf = true; return false; } setDS(nds++, v); return true; }

	// A fresh vertex record: coordinates, no edges yet, bp = v.
	VG_INL void new_vertex(int v, double x, double y, double z) const {
		setXYZ(v, x, y, z);
		setTW(v, 0x0000000000ffffffull | ((unsigned long long)(unsigned short)(short)v << 32));
	}

	// Vertex `from` takes index `to` (src/cell.cc:849-861).
	VG_FN void move_vertex(int from, int to) {
		setXYZ(to, X(from), Y(from), Z(from));
		unsigned long long w = TW(from);
		w = (w & ~(0xffffull << 32)) | ((unsigned long long)(unsigned short)(short)to << 32);
		setTW(to, w);
		for (int i = 0; i < 3; i++) setN(to, i, N(from, i));
		for (int i = 0; i < 3; i++) setE(E(to, i), R(to, i), to);
	}

	// Box start cell, faces -1..-6 (src/cell.cc:294-318, 2567-2580).
	VG_FN void init_box(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax) {
		ovf = false; fatal = false; n_cuts = 0;
		up = 0;
		p = 8;
		xmin *= 2; xmax *= 2; ymin *= 2; ymax *= 2; zmin *= 2; zmax *= 2;
		// edges of corner v (bits x,y,z = v&1, v>>1&1, v>>2&1), relations 2,1,0
		const unsigned EDG[8] = {0x020401, 0x000503, 0x030600, 0x010702, 0x050006, 0x070104, 0x040207, 0x060305};
		const int NID[8][3] = {{-5, -3, -1}, {-5, -2, -3}, {-5, -1, -4}, {-5, -4, -2},
				       {-6, -1, -3}, {-6, -3, -2}, {-6, -4, -1}, {-6, -2, -4}};
		for (int v = 0; v < 8; v++) {
			setXYZ(v, (v & 1) ? xmax : xmin, (v & 2) ? ymax : ymin, (v & 4) ? zmax : zmin);
			// relations r0=2, r1=1, r2=0 -> 0b000110 = 6
			setTW(v, (unsigned long long)EDG[v] | (6ull << 24) | ((unsigned long long)v << 32));
			for (int i = 0; i < 3; i++) setN(v, i, NID[v][i]);
		}
	}

	// The unit Voronoi cell of a periodic container as the start cell
	// (src/container_prd.hh:124-130: c=unit_voro, which copies vertices and
	// edges, zeroes the neighbour ids and sets up=0, src/cell.cc:71-95).  A
	// vertex of another order than 3 leaves the cell to the general tier.
	VG_FN void init_unit(const UnitVert *__restrict__ u, int np) {
		ovf = false; fatal = false; n_cuts = 0;
		up = 0;
		if (np > MAXV) { ovf = true; return; }
		p = np;
		for (int v = 0; v < np; v++) {
			const UnitVert a = u[v];
			if (a.nu != 3) { ovf = true; return; }
			setXYZ(v, a.x, a.y, a.z);
			unsigned long long w = (unsigned long long)(a.ed[0] & 0xff) | (unsigned long long)(a.ed[1] & 0xff) << 8 |
					       (unsigned long long)(a.ed[2] & 0xff) << 16;
			w |= (unsigned long long)((a.rel[0] & 3) | (a.rel[1] & 3) << 2 | (a.rel[2] & 3) << 4) << 24;
			w |= (unsigned long long)v << 32;
			setTW(v, w);
			for (int i = 0; i < 3; i++) setN(v, i, 0);
		}
	}

	// src/cell.cc:2043-2059 (the reference caches the result per vertex and
	// plane; recomputing gives the same value).
	VG_INL unsigned m_test(int n, double &ans) {
		if (n < 0 || n >= MAXV) { fatal = ovf = true; ans = 0; return 1u; }
		const unsigned r = rec(n);
		double x, y;
		ldXY(r, x, y);
		const double z = ld<double>(r + 16);
		ans = x * px;
		ans += y * py;
		ans += z * pz - prsq;
		return ans < -tol ? 0u : (ans > tol ? 2u : 1u);
	}
	// src/cell.cc:2072-2083: a kept vertex within big_tol of the plane
	// would go on the extra-search stack; this tier leaves such cells to the
	// general one.
	VG_INL unsigned m_testx(int n, double &ans) {
		unsigned r = m_test(n, ans);
		if (r == 0 && ans > -big_tol) ovf = true;
		return r;
	}
	// m_testx that also hands back the vertex's coordinates.
	VG_INL unsigned m_testx3(int n, double &ans, double &x, double &y, double &z) {
		if (n < 0 || n >= MAXV) { fatal = ovf = true; ans = 0; x = y = z = 0; return 1u; }
		const unsigned rr = rec(n);
		ldXY(rr, x, y);
		z = ld<double>(rr + 16);
		ans = x * px;
		ans += y * py;
		ans += z * pz - prsq;
		const unsigned r = ans < -tol ? 0u : (ans > tol ? 2u : 1u);
		if (r == 0 && ans > -big_tol) ovf = true;
		return r;
	}

	// src/cell.cc:522-595 (the reference pushes the member `up` in the
	// second loop, src/cell.cc:546-549; kept).
	// t0: the plane values of lp's three neighbours, already computed.
	VG_FN bool definite_max(int &lp, int &ls, double &l, double &u, unsigned &uw, const double *t0) {
		int tp = lp, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < 3; ts++) {
			qp = E(tp, ts);
			q = t0[ts];
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
				if (BP(qp) < 0) continue;
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

	// src/cell.cc:480-515, folded so definite_max has one call site.
	// The three neighbours' plane values are loaded together (one memory
	// round trip instead of three dependent ones); they are then taken in
	// the reference's order with its early exit, so up/u/uw end as there.
	VG_FN bool search_upward(unsigned &uw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		lp = up; l = u;
		for (;;) {
			const unsigned long long w = TW(lp);
			const int e0 = twE(w, 0), e1 = twE(w, 1), e2 = twE(w, 2);
			double t[3];
			unsigned tw[3];
			tw[0] = m_test(e0, t[0]);
			tw[1] = m_test(e1, t[1]);
			tw[2] = m_test(e2, t[2]);
			for (ls = 0; ls < 3; ls++) {
				if (ls == vs) continue;
				up = ls == 0 ? e0 : (ls == 1 ? e1 : e2);
				uw = tw[ls];
				u = t[ls];
				if (u > l) break;
			}
			if (ls == 3 && definite_max(lp, ls, l, u, uw, t)) { up = lp; return false; }
			if (uw != 0) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = R(lp, ls); lp = up; l = u;
		}
		us = R(lp, ls);
		return true;
	}

	// src/cell.cc:631-708 (the reference pushes the argument `lp` in the
	// second loop, src/cell.cc:654-657; kept).
	// t0: the plane values of up's three neighbours, already computed.
	VG_FN bool definite_min(int &lp, int &us, double &l, double &u, unsigned &lw, const double *t0) {
		int tp = up, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < 3; ts++) {
			qp = E(tp, ts);
			q = t0[ts];
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
				if (BP(qp) < 0) continue;
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

	// src/cell.cc:597-629, folded as search_upward.
	VG_FN bool search_downward(unsigned &lw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		for (;;) {
			const unsigned long long w = TW(up);
			const int e0 = twE(w, 0), e1 = twE(w, 1), e2 = twE(w, 2);
			double t[3];
			unsigned tw[3];
			tw[0] = m_test(e0, t[0]);
			tw[1] = m_test(e1, t[1]);
			tw[2] = m_test(e2, t[2]);
			for (us = 0; us < 3; us++) {
				if (us == vs) continue;
				lp = us == 0 ? e0 : (us == 1 ? e1 : e2);
				lw = tw[us];
				l = t[us];
				if (u > l) break;
			}
			if (us == 3 && definite_min(lp, us, l, u, lw, t)) return false;
			if (lw != 2) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = R(up, us); up = lp; u = l;
		}
		ls = R(up, us);
		return true;
	}

	// src/cell.cc:875-1364, the path of a plane that cuts edges cleanly
	// (the lp != -1 start, and only regular crossings while walking the new
	// facet).  Returns true if the cell was deleted (or on overflow).
	VG_FN bool create_facet(int lp, int ls, double l, int us, double u, int pid) {
		int qp, qs, cp, cs, rp;
		unsigned lw;
		double q, r;
		if (!room_v(p)) return true;
		if (!push_ds(up)) return true;
		r = u / (u - l); l = 1 - r;
		new_vertex(p, X(lp) * r + X(up) * l, Y(lp) * r + Y(up) * l, Z(lp) * r + Z(up) * l);
		setN(p, 0, pid);
		setN(p, 1, N(up, us));
		setN(p, 2, N(lp, ls));
		setE(up, us, -1);
		setE(lp, ls, p);
		setR(lp, ls, 1);
		setE(p, 1, lp);
		setR(p, 1, ls);
		cs = 2;
		qs = cyc_up(us);
		qp = up; q = u;
		// coordinates of qp, carried along the walk
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
				setN(p, 0, pid);
				setN(p, 1, N(qp, qs));
				setN(p, 2, N(lp, ls));
				setE(p, 0, cp);
				setE(p, 1, lp);
				setR(p, 0, cs);
				setR(p, 1, ls);
				setE(lp, ls, p);
				setR(lp, ls, 1);
				setE(cp, cs, p);
				setR(cp, cs, 0);
				setE(qp, qs, -1);
				qs = cyc_up(qs);
				cp = p++;
				cs = 2;
			} else {
				ovf = true;      // a vertex on the plane: marginal case
				return true;
			}
		}
		setE(cp, cs, rp);
		setE(rp, 0, cp);
		setR(cp, cs, 0);
		setR(rp, 0, cs);
		return false;
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
		uw = m_test(up, u);
		if (uw == 2) {
			if (!search_downward(lw, lp, ls, us, l, u)) return false;
			if (lw == 1) { ovf = true; return false; }     // vertex on the plane
		} else if (uw == 0) {
			if (!search_upward(uw, lp, ls, us, l, u)) return true;
			if (uw == 1) { ovf = true; return false; }     // vertex on the plane
		} else {
			ovf = true;                                      // vertex 0 on the plane
			return false;
		}
		if (ovf) return false;
		nds = 0;
		if (create_facet(lp, ls, l, us, u, pid)) return false;
		// Delete the cut vertices and everything reachable from them
		// (src/cell.cc:786-863); without marginal vertices the auxiliary
		// and extra-search stacks are empty.
		int d = 0;
		while (d < nds) {
			j = DS(d);
			if (!DEL(j)) { setDEL(j); d++; }
			else setDS(d, DS(--nds));
		}
		for (d = 0; d < nds; d++) {
			cp = DS(d);
			for (int e = 0; e < 3; e++) {
				qp = E(cp, e);
				if (qp != -1 && !DEL(qp)) {
					if (!push_ds(qp)) return false;
					setDEL(qp);
				}
			}
		}
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

	// src/cell.cc:1907-1917
	VG_FN double max_radius_squared() {
		double r = X(0) * X(0) + Y(0) * Y(0) + Z(0) * Z(0), s;
		for (int v = 1; v < p; v++) {
			double a = X(v), b = Y(v), c = Z(v);
			s = a * a;
			s += b * b;
			s += c * c;
			if (s > r) r = s;
		}
		return r;
	}

	// plane_intersects(_guess), src/cell.cc:2366-2410
	VG_FN bool plane_probe(double x, double y, double z, double rsq, bool guess) {
		if (guess) up = 0;
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
	}

	// Volume, centroid, face count and neighbour ids in one face walk
	// (src/cell.cc:1584-1616, 1861-1901, 2168-2187, 2660-2678 share it).
	// The reference marks walked edges by negating them; here a per-record
	// visited bit does (the edges read are always unmarked ones).
	VG_INL bool vis(int v, int i) const { return (ld<unsigned char>(rec(v) + 30) >> i) & 1; }
	VG_INL void setvis(int v, int i) const {
		st<unsigned char>(rec(v) + 30, (unsigned char)(ld<unsigned char>(rec(v) + 30) | (1u << i)));
	}
	template<class NbSink>
	VG_FN void outputs(double &vol, double &cx, double &cy, double &cz, int &nf, NbSink &nb) {
		const double fe = 1 / 48.0;
		double v = 0, tvol;
		double ux, uy, uz, vx, vy, vz, wx, wy, wz;
		const double x0 = X(0), y0 = Y(0), z0 = Z(0);
		cx = cy = cz = 0;
		nf = 0;
		for (int i = 1; i < p; i++) {
			ux = x0 - X(i);
			uy = y0 - Y(i);
			uz = z0 - Z(i);
			for (int j = 0; j < 3; j++) {
				if (vis(i, j)) continue;
				int k = E(i, j);
				nb.put(nf++, N(i, j));
				setvis(i, j);
				int l = cyc_up(R(i, j));
				vx = X(k) - x0;
				vy = Y(k) - y0;
				vz = Z(k) - z0;
				int m = E(k, l);
				setvis(k, l);
				int guard = 0;
				while (m != i) {
					if (++guard > 4 * MAXV) { fatal = ovf = true; return; }
					int n = cyc_up(R(k, l));
					wx = X(m) - x0;
					wy = Y(m) - y0;
					wz = Z(m) - z0;
					tvol = ux * vy * wz + uy * vz * wx + uz * vx * wy - uz * vy * wx - uy * vx * wz - ux * vz * wy;
					v += tvol;
					cx += (wx + vx - ux) * tvol;
					cy += (wy + vy - uy) * tvol;
					cz += (wz + vz - uz) * tvol;
					k = m; l = n; vx = wx; vy = wy; vz = wz;
					m = E(k, l);
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

}  // namespace vgpu
