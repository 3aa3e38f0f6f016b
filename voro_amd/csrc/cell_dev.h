// Voronoi cell geometry for one lane: plane cutting, order-1/2 collapse and
// the face walk that yields volume, centroid, face count and neighbour IDs.
//
// The algorithm is the reference's (src/cell.cc), restated for a bounded,
// lane-interleaved workspace; the steps and floating-point operation order are
// kept identical because vertex numbering, neighbour order and volumes must
// match the reference bit for bit:
//   m_calc / m_test / m_testx        src/cell.cc:2043-2083
//   search_upward / definite_max     src/cell.cc:480-595
//   search_downward / definite_min   src/cell.cc:597-708
//   nplane                           src/cell.cc:718-870
//   create_facet                     src/cell.cc:875-1364
//   collapse_order2 / 1              src/cell.cc:1377-1464
//   delete_connection                src/cell.cc:1473-1513
//   volume / centroid / faces / nbrs src/cell.cc:1584-1616, 1861-1901, 2168-2187, 2660-2678
//   max_radius_squared               src/cell.cc:1907-1917
//   plane_intersects(_guess)         src/cell.cc:2366-2410
//
// Storage differs from the reference: a vertex owns a fixed row of MAXO edge,
// relation and neighbour slots, so changing a vertex's order moves nothing.
// The reference's per-order row pools influence results only through the
// order in which order-1 and order-2 vertices are collapsed (the last pool
// row is taken first), so only those two pools are tracked, with the same
// append / move-last-into-hole behaviour.
#pragma once
#include "vgpu_common.h"

namespace vgpu {

template<class T_, int ST>
struct VCell {
	typedef Layout<T_, ST> L;
	enum { MAXV = L::MAXV, MAXO = L::MAXO, MAXS = L::MAXS, MAXP = L::MAXP };

	// Workspace addressing.  On the GPU the slab is reached through a buffer
	// resource (wavefront-uniform, 4 SGPRs): element i of a field is at
	// voffset (i*ST + lane)*sizeof(T) plus the field's constant offset as the
	// scalar offset, so no per-field vector address is ever live, and an
	// out-of-slab index reads 0 / drops the write instead of faulting.  CPU
	// builds use plain pointers.
	char *wsb;
	int lane;
#if defined(__HIP_DEVICE_COMPILE__)
	__amdgpu_buffer_rsrc_t rs;
#endif

	template<class T>
	struct Ref {
#if defined(__HIP_DEVICE_COMPILE__)
		__amdgpu_buffer_rsrc_t rs;
		unsigned vo, so;
		VG_INL T get() const {
			if constexpr (sizeof(T) == 1) return (T)__builtin_amdgcn_raw_buffer_load_b8(rs, vo, so, 0);
			else if constexpr (sizeof(T) == 2) return (T)__builtin_amdgcn_raw_buffer_load_b16(rs, vo, so, 0);
			else if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, 0));
			else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, 0));
		}
		VG_INL void set(T v) const {
			if constexpr (sizeof(T) == 1) __builtin_amdgcn_raw_buffer_store_b8((unsigned char)v, rs, vo, so, 0);
			else if constexpr (sizeof(T) == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, rs, vo, so, 0);
			else if constexpr (sizeof(T) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, vo, so, 0);
			else {
				typedef unsigned u2 __attribute__((ext_vector_type(2)));
				__builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), rs, vo, so, 0);
			}
		}
#else
		T *ptr;
		VG_INL T get() const { return *ptr; }
		VG_INL void set(T v) const { *ptr = v; }
#endif
		VG_INL operator T() const { return get(); }
		VG_INL const Ref &operator=(T v) const { set(v); return *this; }
		VG_INL const Ref &operator=(const Ref &o) const { set(o.get()); return *this; }
		VG_INL T operator--(int) const { T v = get(); set((T)(v - 1)); return v; }
		VG_INL const Ref &operator|=(T b) const { set((T)(get() | b)); return *this; }
	};

	template<class T> VG_INL Ref<T> A(size_t off, int i) {
		Ref<T> r;
#if defined(__HIP_DEVICE_COMPILE__)
		r.rs = rs;
		r.vo = (unsigned)(i * ST + lane) * (unsigned)sizeof(T);
		r.so = (unsigned)off;
#else
		r.ptr = (T *)(wsb + off + ((size_t)i * ST + lane) * sizeof(T));
#endif
		return r;
	}
	int p, up;
	unsigned maskc;
	double px, py, pz, prsq;
	double tol, tol_cu, big_tol;
	int nds, nds2, nxse, np0, np1, np2;
	bool ovf, fatal;
	int n_cuts;

	VG_INL Ref<double> X(int v) { return A<double>(L::o_x, v); }
	VG_INL Ref<double> Y(int v) { return A<double>(L::o_y, v); }
	VG_INL Ref<double> Z(int v) { return A<double>(L::o_z, v); }
	VG_INL Ref<double> AN(int v) { return A<double>(L::o_a, v); }
	VG_INL Ref<unsigned> M(int v) { return A<unsigned>(L::o_mask, v); }
	VG_INL Ref<int> N(int v, int i) { return A<int>(L::o_ne, v * MAXO + i); }
	VG_INL Ref<short> E(int v, int i) { return A<short>(L::o_ed, v * MAXO + i); }
	VG_INL Ref<signed char> R(int v, int i) { return A<signed char>(L::o_rel, v * MAXO + i); }
	VG_INL Ref<signed char> NU(int v) { return A<signed char>(L::o_nu, v); }
	VG_INL Ref<short> BP(int v) { return A<short>(L::o_bp, v); }
	VG_INL Ref<short> SL(int v) { return A<short>(L::o_slot, v); }
	VG_INL Ref<short> DS(int i) { return A<short>(L::o_ds, i); }
	VG_INL Ref<short> DS2(int i) { return A<short>(L::o_ds2, i); }
	VG_INL Ref<short> XS(int i) { return A<short>(L::o_xse, i); }
	VG_INL Ref<short> DM(int i) { return A<short>(L::o_dm, i); }
	VG_INL Ref<short> P1(int i) { return A<short>(L::o_pool1, i); }
	VG_INL Ref<short> P2(int i) { return A<short>(L::o_pool2, i); }

	VG_INL int cycle_up(int a, int v) { return a == NU(v) - 1 ? 0 : a + 1; }
	VG_INL int cycle_down(int a, int v) { return a == 0 ? NU(v) - 1 : a - 1; }
	VG_INL void flip(int v) { BP(v) = (short)(-1 - BP(v)); }

	// src/cell.cc:20-27: tol = 10 eps max_len_sq, tol_cu = tol^1.5, big_tol = 20 tol.
	VG_FN void setup(char *slab, int lane_, double max_len_sq) {
		wsb = slab; lane = lane_;
#if defined(__HIP_DEVICE_COMPILE__)
		rs = __builtin_amdgcn_make_buffer_rsrc(slab, 0, (int)L::bytes, 0x00020000);
#endif
		set_tol(max_len_sq);
		maskc = 0;
		for (int v = 0; v < MAXV; v++) M(v) = 0;
	}
	VG_INL void set_tol(double max_len_sq) {
		const double eps = 2.220446049250313e-16;
		tol = 10.0 * eps * max_len_sq;
		tol_cu = tol * sqrt(tol);
		big_tol = 20.0 * tol;
	}

	// -- capacity-checked containers -------------------------------------
	VG_INL bool room_v(int v) { if (v >= MAXV) { ovf = true; return false; } return true; }
	VG_INL bool push_ds(int v) { if (nds >= MAXS) { ovf = true; return false; } DS(nds++) = (short)v; return true; }
	VG_INL bool push_ds2(int v) { if (nds2 >= MAXS) { ovf = true; return false; } DS2(nds2++) = (short)v; return true; }
	VG_INL bool push_xse(int v) { if (nxse >= MAXS) { ovf = true; return false; } XS(nxse++) = (short)v; return true; }
	VG_INL bool pool_add(int o, int v) {
		if (o > MAXO) { ovf = true; return false; }
		if (o == 1) { if (np1 >= MAXP) { ovf = true; return false; } SL(v) = (short)np1; P1(np1++) = (short)v; }
		else if (o == 2) { if (np2 >= MAXP) { ovf = true; return false; } SL(v) = (short)np2; P2(np2++) = (short)v; }
		else if (o == 0) np0++;
		return true;
	}
	VG_INL void pool_remove_at(int o, int s) {
		if (o == 1) { int last = P1(--np1); if (s < np1) { P1(s) = (short)last; SL(last) = (short)s; } }
		else if (o == 2) { int last = P2(--np2); if (s < np2) { P2(s) = (short)last; SL(last) = (short)s; } }
		else if (o == 0) np0--;
	}
	VG_INL void pool_remove(int o, int v) { pool_remove_at(o, (o == 1 || o == 2) ? (int)SL(v) : 0); }
	// Vertex `from` takes index `to` (src/cell.cc:849-861, 1416-1426).
	VG_FN void move_vertex(int from, int to) {
		int o = NU(from);
		for (int i = 0; i < o; i++) { E(to, i) = E(from, i); R(to, i) = R(from, i); N(to, i) = N(from, i); }
		NU(to) = (signed char)o;
		for (int i = 0; i < o; i++) E(E(to, i), R(to, i)) = (short)to;
		BP(to) = (short)to;
		if (o == 1) { P1(SL(from)) = (short)to; SL(to) = SL(from); }
		else if (o == 2) { P2(SL(from)) = (short)to; SL(to) = SL(from); }
	}
	VG_INL void copy_pts(int to, int from) { X(to) = X(from); Y(to) = Y(from); Z(to) = Z(from); }

	// Box start cell, faces -1..-6 (src/cell.cc:294-318, 2567-2580).
	VG_FN void init_box(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax) {
		np0 = np1 = np2 = 0;
		ovf = false; fatal = false; n_cuts = 0;
		up = 0;
		p = 8;
		xmin *= 2; xmax *= 2; ymin *= 2; ymax *= 2; zmin *= 2; zmax *= 2;
		// Corner v has bits (x,y,z) = (v&1, v>>1&1, v>>2&1).  Edge rows and
		// face ids of the reference box (src/cell.cc:307-317, 2569-2576).
		static const signed char E0[8][3] = {{1,4,2},{3,5,0},{0,6,3},{2,7,1},{6,0,5},{4,1,7},{7,2,4},{5,3,6}};
		static const signed char N0[8][3] = {{-5,-3,-1},{-5,-2,-3},{-5,-1,-4},{-5,-4,-2},
						     {-6,-1,-3},{-6,-3,-2},{-6,-4,-1},{-6,-2,-4}};
		for (int v = 0; v < 8; v++) {
			X(v) = (v & 1) ? xmax : xmin;
			Y(v) = (v & 2) ? ymax : ymin;
			Z(v) = (v & 4) ? zmax : zmin;
			NU(v) = 3; BP(v) = (short)v;
			for (int i = 0; i < 3; i++) { E(v, i) = E0[v][i]; R(v, i) = (signed char)(2 - i); N(v, i) = N0[v][i]; }
		}
	}

	// The unit Voronoi cell of a periodic container as the start cell
	// (src/container_prd.hh:124-130: c=unit_voro copies vertices and edges,
	// zeroes the neighbour ids and sets up=0, src/cell.cc:71-95).  The unit
	// cell holds no order-1 or order-2 vertex, so both pools start empty.
	VG_FN void init_unit(const UnitVert *__restrict__ u, int np) {
		np0 = np1 = np2 = 0;
		ovf = false; fatal = false; n_cuts = 0;
		up = 0;
		if (np > MAXV) { ovf = true; return; }
		p = np;
		for (int v = 0; v < np; v++) {
			const UnitVert a = u[v];
			if (a.nu > MAXO) { ovf = true; return; }
			X(v) = a.x; Y(v) = a.y; Z(v) = a.z;
			NU(v) = (signed char)a.nu; BP(v) = (short)v;
			for (int i = 0; i < a.nu; i++) { E(v, i) = a.ed[i]; R(v, i) = a.rel[i]; N(v, i) = 0; }
		}
	}

	// src/cell.cc:2043-2059
	VG_INL unsigned m_calc(int n, double &ans) {
		ans = X(n) * px;
		ans += Y(n) * py;
		ans += Z(n) * pz - prsq;
		AN(n) = ans;
		unsigned r = ans < -tol ? 0u : (ans > tol ? 2u : 1u);
		M(n) = maskc | r;
		return r;
	}
	// A negative index would be a corrupt edge table: abort the cell.
	VG_INL bool bad_index(int n) {
		if (n < 0 || n >= MAXV) { fatal = true; ovf = true; return true; }
		return false;
	}
	VG_INL unsigned m_test(int n, double &ans) {
		if (bad_index(n)) { ans = 0; return 1u; }
		unsigned mk = M(n);
		if (mk >= maskc) { ans = AN(n); return mk & 3; }
		return m_calc(n, ans);
	}
	// src/cell.cc:2072-2083
	VG_INL unsigned m_testx(int n, double &ans) {
		if (bad_index(n)) { ans = 0; return 1u; }
		unsigned r, mk = M(n);
		if (mk >= maskc) { ans = AN(n); r = mk & 3; }
		else r = m_calc(n, ans);
		if (r == 0 && ans > -big_tol && BP(n) != -1) {
			BP(n) = -1;
			push_xse(n);
		}
		return r;
	}

	// src/cell.cc:444-475
	VG_FN bool search_for_outside_edge(int &u_) {
		int sc2 = nds2;
		if (!push_ds2(u_)) return false;
		int j = sc2;
		double l;
		while (j < nds2) {
			u_ = DS2(j++);
			for (int i = 0; i < NU(u_); i++) {
				int lp = E(u_, i);
				unsigned lw = m_test(lp, l);
				if (lw == 0) { nds2 = sc2; return true; }
				else if (lw == 1) {
					bool seen = false;
					for (int k = sc2; k < nds2; k++) if (DS2(k) == lp) { seen = true; break; }
					if (!seen && !push_ds2(lp)) return false;
				}
			}
		}
		nds2 = sc2;
		return false;
	}

	// src/cell.cc:522-595.  The reference pushes the member `up` in the
	// second loop (src/cell.cc:546-549); that is kept.
	VG_FN bool definite_max(int &lp, int &ls, double &l, double &u, unsigned &uw) {
		int tp = lp, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < NU(tp); ts++) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q > l - big_tol) break;
		}
		if (ts == NU(tp)) return true;
		int nst = 0;
		flip(lp); flip(qp);
		DM(nst++) = (short)qp;
		ts++;
		while (ts < NU(tp)) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q > l - big_tol) {
				if (nst >= MAXS) { ovf = true; return true; }
				DM(nst++) = (short)up; flip(up);
			}
			ts++;
		}
		int spp = 0;
		while (spp < nst) {
			tp = DM(spp++);
			for (ts = 0; ts < NU(tp); ts++) {
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
					DM(nst++) = (short)qp; flip(qp);
				}
			}
		}
		flip(lp);
		while (nst > 0) flip(DM(--nst));
		return true;
	}

	// src/cell.cc:480-515.  The reference's first scan and its while loop
	// are folded into one loop (vs = -1 skips nothing on the first pass), so
	// definite_max has a single call site.
	VG_FN bool search_upward(unsigned &uw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		lp = up; l = u;
		for (;;) {
			for (ls = 0; ls < NU(lp); ls++) {
				if (ls == vs) continue;
				up = E(lp, ls);
				uw = m_test(up, u);
				if (u > l) break;
			}
			if (ls == NU(lp) && definite_max(lp, ls, l, u, uw)) { up = lp; return false; }
			if (uw != 0) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = R(lp, ls); lp = up; l = u;
		}
		us = R(lp, ls);
		return true;
	}

	// src/cell.cc:631-708.  The reference pushes the argument `lp` in the
	// second loop (src/cell.cc:654-657); that is kept.
	VG_FN bool definite_min(int &lp, int &us, double &l, double &u, unsigned &lw) {
		int tp = up, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < NU(tp); ts++) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q < u + big_tol) break;
		}
		if (ts == NU(tp)) return true;
		int nst = 0;
		flip(up); flip(qp);
		DM(nst++) = (short)qp;
		ts++;
		while (ts < NU(tp)) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q < u + big_tol) {
				if (nst >= MAXS) { ovf = true; return true; }
				DM(nst++) = (short)lp; flip(lp);
			}
			ts++;
		}
		int spp = 0;
		while (spp < nst) {
			tp = DM(spp++);
			for (ts = 0; ts < NU(tp); ts++) {
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
					DM(nst++) = (short)qp; flip(qp);
				}
			}
		}
		flip(up);
		while (nst > 0) flip(DM(--nst));
		return true;
	}

	// src/cell.cc:597-629, folded as search_upward is.
	VG_FN bool search_downward(unsigned &lw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs = -1, guard = 0;
		for (;;) {
			for (us = 0; us < NU(up); us++) {
				if (us == vs) continue;
				lp = E(up, us);
				lw = m_test(lp, l);
				if (u > l) break;
			}
			if (us == NU(up) && definite_min(lp, us, l, u, lw)) return false;
			if (lw != 2) break;
			if (ovf || ++guard > 4 * MAXV) { fatal = ovf = true; return false; }
			vs = R(up, us); up = lp; u = l;
		}
		ls = R(up, us);
		return true;
	}

	// Moves edge (up,i) of the vertex being replaced onto the new vertex
	// np at slot k and rewires its far end (src/cell.cc:962-972).
	VG_INL void take_edge(int np_, int k, int src, int i) {
		int qp = E(src, i), qs = R(src, i);
		N(np_, k) = N(src, i);
		E(np_, k) = (short)qp; R(np_, k) = (signed char)qs;
		E(qp, qs) = (short)np_; R(qp, qs) = (signed char)k;
		E(src, i) = -1;
	}

	// src/cell.cc:875-1364.  Returns true if the cell was deleted (or on
	// capacity overflow, with ovf set).
	VG_FN bool create_facet(int lp, int ls, double l, int us, double u, int pid) {
		int i, j, k, qp, qs, iqs, cp, cs, rp;
		unsigned lw, qw;
		bool new_double_edge = false, double_edge = false;
		double q, r;
		if (!room_v(p)) return true;
		if (lp == -1) {
			if (!search_for_outside_edge(up)) return true;
			copy_pts(p, up);
			i = 0;
			lp = E(up, 0);
			lw = m_testx(lp, l);
			if (lw != 0) {
				unsigned rw = lw;
				do {
					i++;
					if (i == NU(up)) return true;
					lp = E(up, i);
					lw = m_testx(lp, l);
				} while (lw != 0);
				j = i + 1;
				while (j < NU(up)) {
					lp = E(up, j);
					lw = m_testx(lp, l);
					if (lw != 0) break;
					j++;
				}
				int o = (j == NU(up) && i == 1 && rw == 1) ? (int)NU(up) : j - i + 2;
				double_edge = (j == NU(up) && i == 1 && rw == 1);
				if (!pool_add(o, p)) return true;
				NU(p) = (signed char)o;
				k = 1;
				BP(p) = (short)p;
				us = cycle_down(i, up);
				while (i < j) { take_edge(p, k, up, i); i++; k++; }
				qs = i == NU(up) ? 0 : i;
			} else {
				i = NU(up) - 1;
				lp = E(up, i);
				lw = m_testx(lp, l);
				while (lw == 0) {
					i--;
					if (i == 0) return false;
					lp = E(up, i);
					lw = m_testx(lp, l);
				}
				j = 1;
				qp = E(up, j);
				qw = m_testx(qp, q);
				while (qw == 0) {
					j++;
					qp = E(up, j);
					qw = m_testx(qp, l);   // the reference writes l here (src/cell.cc:1005)
				}
				int o = (i == j && qw == 1) ? (int)NU(up) : NU(up) - i + j + 1;
				double_edge = (i == j && qw == 1);
				if (!pool_add(o, p)) return true;
				NU(p) = (signed char)o;
				k = 1;
				BP(p) = (short)p;
				us = i++;
				while (i < NU(up)) { take_edge(p, k, up, i); i++; k++; }
				i = 0;
				while (i < j) { take_edge(p, k, up, i); i++; k++; }
				qs = j;
			}
			if (!double_edge) {
				N(p, k) = N(up, qs);
				N(p, 0) = pid;
			} else N(p, 0) = N(up, qs);
			if (!push_ds2(up)) return true;
			cs = k;
			qp = up; q = u;
			i = E(up, us);
			us = R(up, us);
			up = i;
			BP(qp) = (short)(-p);
		} else {
			if (!push_ds(up)) return true;
			r = u / (u - l); l = 1 - r;
			X(p) = X(lp) * r + X(up) * l;
			Y(p) = Y(lp) * r + Y(up) * l;
			Z(p) = Z(lp) * r + Z(up) * l;
			NU(p) = 3;
			N(p, 0) = pid;
			N(p, 1) = N(up, us);
			N(p, 2) = N(lp, ls);
			BP(p) = (short)p;
			E(up, us) = -1;
			E(lp, ls) = (short)p;
			R(lp, ls) = 1;
			E(p, 1) = (short)lp;
			R(p, 1) = (signed char)ls;
			cs = 2;
			qs = cycle_up(us, up);
			qp = up; q = u;
		}
		cp = p; rp = p; p++;
		int guard = 0;
		while (qp != up || qs != us) {
			if (++guard > 16 * MAXV) { fatal = ovf = true; return true; }
			lp = E(qp, qs);
			lw = m_testx(lp, l);
			if (ovf) return true;
			if (lw == 2) {
				qs = cycle_up(R(qp, qs), lp);
				qp = lp;
				q = l;
				if (!push_ds(qp)) return true;
			} else if (lw == 0) {
				if (!room_v(p)) return true;
				r = q / (q - l); l = 1 - r;
				X(p) = X(lp) * r + X(qp) * l;
				Y(p) = Y(lp) * r + Y(qp) * l;
				Z(p) = Z(lp) * r + Z(qp) * l;
				NU(p) = 3;
				ls = R(qp, qs);
				N(p, 0) = pid;
				N(p, 1) = N(qp, qs);
				N(p, 2) = N(lp, ls);
				E(p, 0) = (short)cp;
				E(p, 1) = (short)lp;
				R(p, 0) = (signed char)cs;
				R(p, 1) = (signed char)ls;
				BP(p) = (short)p;
				E(lp, ls) = (short)p;
				R(lp, ls) = 1;
				E(cp, cs) = (short)p;
				R(cp, cs) = 0;
				E(qp, qs) = -1;
				qs = cycle_up(qs, qp);
				cp = p++;
				cs = 2;
			} else {
				if (!room_v(p)) return true;
				k = double_edge ? 0 : 1;
				qs = R(qp, qs);
				qp = lp;
				iqs = qs;
				do {
					k++;
					qs = cycle_up(qs, qp);
					lp = E(qp, qs);
					lw = m_testx(lp, l);
				} while (lw == 0);
				if (ovf) return true;
				j = -BP(qp);
				if (qp == up && qs == us) {
					new_double_edge = false;
					if (j > 0) k += NU(j);
				} else {
					if (j > 0) {
						k += NU(j);
						if (lw == 1) {
							i = -BP(lp);
							if (i > 0) {
								if (E(i, NU(i) - 1) == j) { new_double_edge = true; k -= 1; }
								else new_double_edge = false;
							} else {
								if (j == rp && lp == up && R(qp, qs) == us) { new_double_edge = true; k -= 1; }
								else new_double_edge = false;
							}
						} else new_double_edge = false;
					} else {
						if (lw == 1) {
							i = -BP(lp);
							if (i == cp) { new_double_edge = true; k -= 1; }
							else new_double_edge = false;
						} else new_double_edge = false;
					}
				}
				if (k > MAXO) { ovf = true; return true; }
				if (j > 0) {
					if (NU(j) != k) {
						// Vertex j grows to order k (src/cell.cc:1252-1281).
						int old = NU(j), s_old = (old == 1 || old == 2) ? (int)SL(j) : 0;
						if (!pool_add(k, j)) return true;
						pool_remove_at(old, s_old);
						BP(j) = (short)j;
					}
					i = NU(j);
				} else {
					if (!pool_add(k, p)) return true;
					BP(p) = (short)p;
					if (!push_ds2(qp)) return true;
					copy_pts(p, qp);
					BP(qp) = (short)(-p);
					j = p++;
					i = 0;
				}
				NU(j) = (signed char)k;
				if (!double_edge) {
					E(j, i) = (short)cp;
					R(j, i) = (signed char)cs;
					N(j, i) = pid;
					E(cp, cs) = (short)j;
					R(cp, cs) = (signed char)i;
					i++;
				}
				qs = iqs;
				while (i < (new_double_edge ? k : k - 1)) {
					qs = cycle_up(qs, qp);
					lp = E(qp, qs); ls = R(qp, qs);
					N(j, i) = N(qp, qs);
					E(j, i) = (short)lp;
					R(j, i) = (signed char)ls;
					E(lp, ls) = (short)j;
					R(lp, ls) = (signed char)i;
					E(qp, qs) = -1;
					i++;
				}
				qs = cycle_up(qs, qp);
				cs = i;
				cp = j;
				N(j, new_double_edge ? 0 : cs) = N(qp, qs);
				double_edge = new_double_edge;
			}
		}
		E(cp, cs) = (short)rp;
		E(rp, 0) = (short)cp;
		R(cp, cs) = 0;
		R(rp, 0) = (signed char)cs;
		return false;
	}

	// src/cell.cc:1473-1513.  Removes edge k of vertex j; the neighbour id
	// removed is entry q (k itself when hand, else the next one).
	VG_FN bool delete_connection(int j, int k, bool hand) {
		int q = hand ? k : cycle_up(k, j);
		int i = NU(j) - 1, l, m;
		if (i < 1) return false;
		for (l = q; l < i; l++) N(j, l) = N(j, l + 1);
		for (l = k; l < i; l++) {
			m = E(j, l + 1);
			int kk = R(j, l + 1);
			E(j, l) = (short)m;
			R(j, l) = (signed char)kk;
			R(m, kk)--;
		}
		int old = NU(j), s_old = (old == 1 || old == 2) ? (int)SL(j) : 0;
		if (!pool_add(i, j)) return false;
		pool_remove_at(old, s_old);
		NU(j) = (signed char)i;
		return true;
	}

	// src/cell.cc:1441-1464
	VG_FN bool collapse_order1() {
		while (np1 > 0) {
			up = 0;
			int i = P1(--np1);
			int j = E(i, 0), k = R(i, 0);
			if (!delete_connection(j, k, false)) return false;
			--p;
			if (up == i) up = 0;
			if (p != i) {
				if (up == p) up = i;
				copy_pts(i, p);
				move_vertex(p, i);
			}
		}
		return true;
	}

	// src/cell.cc:1377-1431
	VG_FN bool collapse_order2() {
		if (!collapse_order1()) return false;
		while (np2 > 0) {
			int i = P2(--np2);
			int j = E(i, 0), k = E(i, 1);
			if (j == k) return false;
			int l;
			for (l = 0; l < NU(j); l++) if (E(j, l) == k) break;
			int a = R(i, 0), b = R(i, 1);
			if (l == NU(j)) {
				E(j, a) = (short)k;
				E(k, b) = (short)j;
				R(j, a) = (signed char)b;
				R(k, b) = (signed char)a;
			} else {
				if (!delete_connection(j, a, false)) return false;
				if (!delete_connection(k, b, true)) return false;
			}
			--p;
			if (up == i) up = 0;
			if (p != i) {
				if (up == p) up = i;
				copy_pts(i, p);
				move_vertex(p, i);
			}
			if (!collapse_order1()) return false;
		}
		return true;
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
		maskc += 4;
		if (maskc < 4) { for (int t = 0; t < MAXV; t++) M(t) = 0; maskc = 4; }
		uw = m_test(up, u);
		if (uw == 2) {
			if (!search_downward(lw, lp, ls, us, l, u)) return false;
			if (lw == 1) { up = lp; lp = -1; }
		} else if (uw == 0) {
			if (!search_upward(uw, lp, ls, us, l, u)) return true;
			if (uw == 1) lp = -1;
		} else {
			lp = -1;
		}
		if (ovf) return false;
		nds = nds2 = nxse = 0;
		int op = p;

		// The first facet, then one facet per marginal or cut edge met
		// while scanning the extra-search stack (src/cell.cc:746-780).  The
		// scan is written as a resumable cursor so that create_facet has a
		// single call site.
		int f_lp = lp, f_ls = ls, f_us = us;
		double f_l = l, f_u = u;
		int xk = 0, xlp = -1, xls = 0;
		for (;;) {
			if (create_facet(f_lp, f_ls, f_l, f_us, f_u, pid)) return false;
			bool have = false;
			for (;;) {
				if (xlp < 0 || xls >= NU(xlp)) {
					if (xk >= nxse) break;
					xlp = XS(xk++);
					m_test(xlp, l);
					xls = 0;
					continue;
				}
				int e = xls++;
				up = E(xlp, e);
				uw = m_test(up, u);
				if (up >= op) continue;
				if (uw == 0) {
					if (u > -big_tol && BP(up) != -1) {
						BP(up) = -1;
						if (!push_xse(up)) return false;
					}
				} else if (uw == 1) {
					f_lp = -1; f_ls = 0; f_l = 0; f_us = 0; f_u = u;
					have = true;
					break;
				} else {
					f_us = R(xlp, e);
					m_test(xlp, l);
					f_lp = xlp; f_ls = e; f_l = l; f_u = u;
					have = true;
					break;
				}
			}
			if (!have) break;
		}

		for (int t = 0; t < nxse; t++) { int v = XS(t); BP(v) = (short)v; }
		int d = 0;
		while (d < nds) {
			j = DS(d);
			if (R(j, 0) != -1) { R(j, 0) = -1; d++; }
			else DS(d) = DS(--nds);
		}
		for (int t = 0; t < nds2; t++) {
			j = DS2(t);
			BP(j) = (short)j;
			if (R(j, 0) != -1) { R(j, 0) = -1; if (!push_ds(j)) return false; }
		}
		for (d = 0; d < nds; d++) {
			cp = DS(d);
			for (int e = 0; e < NU(cp); e++) {
				qp = E(cp, e);
				if (qp != -1 && R(qp, 0) != -1) {
					if (!push_ds(qp)) return false;
					R(qp, 0) = -1;
				}
			}
		}
		up = 0;
		while (nds > 0) {
			--p;
			while (R(p, 0) == -1) {
				if (p == 0) { fatal = ovf = true; return false; }
				pool_remove(NU(p), p);
				--p;
			}
			up = DS(--nds);
			if (up < p) {
				copy_pts(up, p);
				pool_remove(NU(up), up);
				move_vertex(p, up);
			} else up = p++;
		}
		if (np0 > 0) { fatal = true; return false; }
		return collapse_order2();
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

	// plane_intersects / plane_intersects_guess, src/cell.cc:2366-2410, as
	// one routine: does any vertex lie strictly beyond x.v = rsq?  The guess
	// form restarts from vertex 0 and climbs through a sparse sample of
	// vertices first; both probe vertex `up` with >= before the full scan.
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

	// Volume, centroid, face count and neighbour ids in one face walk; the
	// four reference routines share this walk and accumulation order.
	// Neighbour ids go to nb[0..cap); the face count is returned in nf even
	// when it exceeds cap.
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
			for (int j = 0; j < NU(i); j++) {
				int k = E(i, j);
				if (k < 0) continue;
				nb.put(nf++, N(i, j));
				E(i, j) = (short)(-1 - k);
				int l = cycle_up(R(i, j), k);
				vx = X(k) - x0;
				vy = Y(k) - y0;
				vz = Z(k) - z0;
				int m = E(k, l);
				E(k, l) = (short)(-1 - m);
				while (m != i) {
					int n = cycle_up(R(k, l), m);
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
					E(k, l) = (short)(-1 - m);
				}
			}
		}
		// The walk leaves every edge flipped; the cell is discarded after
		// its outputs are taken, so the reference's reset_edges() pass
		// (src/cell.cc:2025-2031) is not needed here.
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
