// CPU restatement of the Voro++ per-particle cell computation.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle: tests, the smoke check and
// the benchmark's CPU-baseline leg may call it, as the checker.  The product
// (voro_amd/) never links, loads or calls anything in oracle/.
//
// It restates, in its own storage model, the algorithm of the reference
// (Voro++ 0.4.6 / 0.5-dev, chr1shr/voro), citing the file:line each part follows:
//   cell geometry and plane cutting      src/cell.cc:294-318, 444-870, 875-1513
//   cell outputs                         src/cell.cc:1584-1616, 1861-1917, 2168-2187, 2660-2678
//   per-particle compute                 src/v_compute.cc:236-264, 291-621, 634-961
//   block grid, worklist radii           src/v_base.cc:23-94
//   rectangular containers, binning      src/container.cc:30-191, src/container.hh:192-251
//   radical (power) option               src/rad_option.hh:24-158
//   sphere wall                          src/wall.cc:28-36
//   c_loop_all order                     src/c_loops.hh:165-209
//   periodic triclinic containers        src/container_prd.cc, src/container_prd.hh, src/unitcell.cc
//
// Storage model (differs from the reference's per-order memory pools, same
// semantics): every vertex owns a fixed-stride row of S edge slots, S
// relation slots and S neighbour-id slots; the reference's back-pointer slot
// (used for marking) is a separate per-vertex int.  The reference's per-order
// row pools matter for results only through the order in which order-1/2
// vertices are collapsed (src/cell.cc:1380-1464 pops the last row), so the
// row order of every pool is tracked as a list with the same append /
// move-last-into-hole behaviour.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "../voro_amd/csrc/worklist_table.h"

namespace orc {

// src/config.hh:80-84
static const double kTolerance = 10.0 * std::numeric_limits<double>::epsilon();
static const double kBigTolFac = 20.0;

struct Cell {
	int p = 0, up = 0;
	int S = 16;                 // row stride (maximum storable vertex order + 1)
	int cap = 0;                // vertex capacity
	std::vector<double> pts;    // 4 per vertex: 2x coordinates, cached plane value
	std::vector<int> nu, bp, ed, rel, ne, slot;
	std::vector<unsigned> mask;
	std::vector<std::vector<int> > pool;
	std::vector<int> ds, ds2, xse;
	unsigned maskc = 0;
	double px = 0, py = 0, pz = 0, prsq = 0;
	double tol, tol_cu, big_tol;
	bool fatal = false;         // the reference would have called exit()
	long long n_cuts = 0;       // nplane calls on this object (statistics)
	int max_p = 0;              // peak vertex count seen (statistics)
	int max_order = 0;          // peak vertex order seen (statistics)

	explicit Cell(double max_len_sq) {
		// src/cell.cc:20-44
		tol = kTolerance * max_len_sq;
		tol_cu = tol * std::sqrt(tol);
		big_tol = kBigTolFac * tol;
		grow_vertices(256);
		pool.resize(S);
	}

	int &E(int v, int i) { return ed[v * S + i]; }
	int &R(int v, int i) { return rel[v * S + i]; }
	int &N(int v, int i) { return ne[v * S + i]; }
	int cycle_up(int a, int v) { return a == nu[v] - 1 ? 0 : a + 1; }
	int cycle_down(int a, int v) { return a == 0 ? nu[v] - 1 : a - 1; }
	void flip(int v) { bp[v] = -1 - bp[v]; }

	void grow_vertices(int ncap) {
		pts.resize(4 * ncap, 0.0);
		nu.resize(ncap, 0); bp.resize(ncap, 0); slot.resize(ncap, 0);
		mask.resize(ncap, 0u);
		ed.resize((size_t)ncap * S, 0); rel.resize((size_t)ncap * S, 0); ne.resize((size_t)ncap * S, 0);
		cap = ncap;
	}
	void need_vertex(int v) { if (v >= cap) grow_vertices(2 * cap > v + 1 ? 2 * cap : v + 1); }
	void need_order(int k) {
		if (k > max_order) max_order = k;
		if (k < S) return;
		int nS = S;
		while (nS <= k) nS <<= 1;
		std::vector<int> e2((size_t)cap * nS, 0), r2((size_t)cap * nS, 0), n2((size_t)cap * nS, 0);
		for (int v = 0; v < cap; v++) for (int i = 0; i < S; i++) {
			e2[(size_t)v * nS + i] = ed[(size_t)v * S + i];
			r2[(size_t)v * nS + i] = rel[(size_t)v * S + i];
			n2[(size_t)v * nS + i] = ne[(size_t)v * S + i];
		}
		ed.swap(e2); rel.swap(r2); ne.swap(n2);
		S = nS;
		pool.resize(S);
	}
	void pool_add(int o, int v) {
		need_order(o);
		slot[v] = (int)pool[o].size();
		pool[o].push_back(v);
	}
	// Removes the row at position s of pool o by moving the last row into
	// it, as the reference's memory compaction does.
	void pool_remove_at(int o, int s) {
		int last = pool[o].back();
		pool[o].pop_back();
		if (s < (int)pool[o].size()) { pool[o][s] = last; slot[last] = s; }
	}
	void pool_remove(int o, int v) { pool_remove_at(o, slot[v]); }
	// Vertex `from` takes index `to` (src/cell.cc:849-861, 1416-1426).
	void move_vertex(int from, int to) {
		int o = nu[from];
		for (int i = 0; i < o; i++) {
			E(to, i) = E(from, i); R(to, i) = R(from, i); N(to, i) = N(from, i);
		}
		nu[to] = o;
		for (int i = 0; i < o; i++) E(E(to, i), R(to, i)) = to;
		bp[to] = to;
		pool[o][slot[from]] = to;
		slot[to] = slot[from];
	}
	void copy_pts(int to, int from) {
		pts[4 * to] = pts[4 * from];
		pts[4 * to + 1] = pts[4 * from + 1];
		pts[4 * to + 2] = pts[4 * from + 2];
	}

	// Box start cell with faces -1..-6 (src/cell.cc:294-318, 2567-2580).
	void init_box(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax) {
		for (size_t o = 0; o < pool.size(); o++) pool[o].clear();
		up = 0;
		p = 8;
		xmin *= 2; xmax *= 2; ymin *= 2; ymax *= 2; zmin *= 2; zmax *= 2;
		const double X[2] = {xmin, xmax}, Y[2] = {ymin, ymax}, Z[2] = {zmin, zmax};
		for (int v = 0; v < 8; v++) {
			pts[4 * v] = X[v & 1]; pts[4 * v + 1] = Y[(v >> 1) & 1]; pts[4 * v + 2] = Z[(v >> 2) & 1];
		}
		static const int E0[8][3] = {{1,4,2},{3,5,0},{0,6,3},{2,7,1},{6,0,5},{4,1,7},{7,2,4},{5,3,6}};
		static const int R0[8][3] = {{2,1,0},{2,1,0},{2,1,0},{2,1,0},{2,1,0},{2,1,0},{2,1,0},{2,1,0}};
		static const int N0[8][3] = {{-5,-3,-1},{-5,-2,-3},{-5,-1,-4},{-5,-4,-2},
					     {-6,-1,-3},{-6,-3,-2},{-6,-4,-1},{-6,-2,-4}};
		for (int v = 0; v < 8; v++) {
			nu[v] = 3; bp[v] = v;
			for (int i = 0; i < 3; i++) { E(v, i) = E0[v][i]; R(v, i) = R0[v][i]; N(v, i) = N0[v][i]; }
			pool_add(3, v);
		}
	}

	// src/cell.cc:2043-2059
	unsigned m_calc(int n, double &ans) {
		double *pp = &pts[4 * n];
		ans = pp[0] * px;
		ans += pp[1] * py;
		ans += pp[2] * pz - prsq;
		pp[3] = ans;
		unsigned r = ans < -tol ? 0u : (ans > tol ? 2u : 1u);
		mask[n] = maskc | r;
		return r;
	}
	unsigned m_test(int n, double &ans) {
		if (mask[n] >= maskc) { ans = pts[4 * n + 3]; return mask[n] & 3; }
		return m_calc(n, ans);
	}
	// src/cell.cc:2072-2083
	unsigned m_testx(int n, double &ans) {
		unsigned r;
		if (mask[n] >= maskc) { ans = pts[4 * n + 3]; r = mask[n] & 3; }
		else r = m_calc(n, ans);
		if (r == 0 && ans > -big_tol && bp[n] != -1) {
			bp[n] = -1;
			xse.push_back(n);
		}
		return r;
	}

	// src/cell.cc:444-475
	bool search_for_outside_edge(int &u_) {
		int sc2 = (int)ds2.size();
		ds2.push_back(u_);
		size_t j = sc2;
		double l;
		while (j < ds2.size()) {
			u_ = ds2[j++];
			for (int i = 0; i < nu[u_]; i++) {
				int lp = E(u_, i);
				unsigned lw = m_test(lp, l);
				if (lw == 0) { ds2.resize(sc2); return true; }
				else if (lw == 1) {
					bool seen = false;
					for (size_t k = sc2; k < ds2.size(); k++) if (ds2[k] == lp) { seen = true; break; }
					if (!seen) ds2.push_back(lp);
				}
			}
		}
		ds2.resize(sc2);
		return false;
	}

	// src/cell.cc:522-595
	bool definite_max(int &lp, int &ls, double &l, double &u, unsigned &uw) {
		int tp = lp, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < nu[tp]; ts++) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q > l - big_tol) break;
		}
		if (ts == nu[tp]) return true;
		std::vector<int> st;
		flip(lp); flip(qp);
		st.push_back(qp);
		ts++;
		while (ts < nu[tp]) {
			qp = E(tp, ts);
			m_test(qp, q);
			// The reference pushes the member `up` here (src/cell.cc:546-549).
			if (q > l - big_tol) { st.push_back(up); flip(up); }
			ts++;
		}
		size_t spp = 0;
		while (spp < st.size()) {
			tp = st[spp++];
			for (ts = 0; ts < nu[tp]; ts++) {
				qp = E(tp, ts);
				if (bp[qp] < 0) continue;
				qw = m_test(qp, q);
				if (q > l) {
					flip(lp);
					lp = tp; ls = ts;
					m_test(lp, l);
					up = qp; uw = qw; u = q;
					while (!st.empty()) { flip(st.back()); st.pop_back(); }
					return false;
				}
				if (q > l - big_tol) { st.push_back(qp); flip(qp); }
			}
		}
		flip(lp);
		while (!st.empty()) { flip(st.back()); st.pop_back(); }
		return true;
	}

	// src/cell.cc:480-515
	bool search_upward(unsigned &uw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs;
		lp = up; l = u;
		for (ls = 0; ls < nu[lp]; ls++) {
			up = E(lp, ls);
			uw = m_test(up, u);
			if (u > l) break;
		}
		if (ls == nu[lp]) if (definite_max(lp, ls, l, u, uw)) { up = lp; return false; }
		while (uw == 0) {
			vs = R(lp, ls); lp = up; l = u;
			for (ls = 0; ls < nu[lp]; ls++) {
				if (ls == vs) continue;
				up = E(lp, ls);
				uw = m_test(up, u);
				if (u > l) break;
			}
			if (ls == nu[lp] && definite_max(lp, ls, l, u, uw)) { up = lp; return false; }
		}
		us = R(lp, ls);
		return true;
	}

	// src/cell.cc:631-708
	bool definite_min(int &lp, int &us, double &l, double &u, unsigned &lw) {
		int tp = up, ts, qp = 0;
		unsigned qw;
		double q;
		for (ts = 0; ts < nu[tp]; ts++) {
			qp = E(tp, ts);
			m_test(qp, q);
			if (q < u + big_tol) break;
		}
		if (ts == nu[tp]) return true;
		std::vector<int> st;
		flip(up); flip(qp);
		st.push_back(qp);
		ts++;
		while (ts < nu[tp]) {
			qp = E(tp, ts);
			m_test(qp, q);
			// The reference pushes the argument `lp` here (src/cell.cc:654-657).
			if (q < u + big_tol) { st.push_back(lp); flip(lp); }
			ts++;
		}
		size_t spp = 0;
		while (spp < st.size()) {
			tp = st[spp++];
			for (ts = 0; ts < nu[tp]; ts++) {
				qp = E(tp, ts);
				if (bp[qp] < 0) continue;
				qw = m_test(qp, q);
				if (q < u) {
					flip(up);
					up = tp; us = ts;
					m_test(up, u);
					lp = qp; lw = qw; l = q;
					while (!st.empty()) { flip(st.back()); st.pop_back(); }
					return false;
				}
				if (q < u + big_tol) { st.push_back(qp); flip(qp); }
			}
		}
		flip(up);
		while (!st.empty()) { flip(st.back()); st.pop_back(); }
		return true;
	}

	// src/cell.cc:597-629
	bool search_downward(unsigned &lw, int &lp, int &ls, int &us, double &l, double &u) {
		int vs;
		for (us = 0; us < nu[up]; us++) {
			lp = E(up, us);
			lw = m_test(lp, l);
			if (u > l) break;
		}
		if (us == nu[up]) if (definite_min(lp, us, l, u, lw)) return false;
		while (lw == 2) {
			vs = R(up, us); up = lp; u = l;
			for (us = 0; us < nu[up]; us++) {
				if (us == vs) continue;
				lp = E(up, us);
				lw = m_test(lp, l);
				if (u > l) break;
			}
			if (us == nu[up] && definite_min(lp, us, l, u, lw)) return false;
		}
		ls = R(up, us);
		return true;
	}

	// src/cell.cc:875-1364.  Returns true if the cell was deleted.
	bool create_facet(int lp, int ls, double l, int us, double u, int pid) {
		int i, j, k, qp, qs, iqs, cp, cs, rp;
		unsigned lw, qw;
		bool new_double_edge = false, double_edge = false;
		double q, r;
		need_vertex(p);
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
					if (i == nu[up]) return true;
					lp = E(up, i);
					lw = m_testx(lp, l);
				} while (lw != 0);
				j = i + 1;
				while (j < nu[up]) {
					lp = E(up, j);
					lw = m_testx(lp, l);
					if (lw != 0) break;
					j++;
				}
				if (j == nu[up] && i == 1 && rw == 1) {
					nu[p] = nu[up];
					double_edge = true;
				} else nu[p] = j - i + 2;
				k = 1;
				pool_add(nu[p], p);
				bp[p] = p;
				us = cycle_down(i, up);
				while (i < j) {
					qp = E(up, i); qs = R(up, i);
					N(p, k) = N(up, i);
					E(p, k) = qp; R(p, k) = qs;
					E(qp, qs) = p; R(qp, qs) = k;
					E(up, i) = -1;
					i++; k++;
				}
				qs = i == nu[up] ? 0 : i;
			} else {
				i = nu[up] - 1;
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
				if (i == j && qw == 1) {
					double_edge = true;
					nu[p] = nu[up];
				} else nu[p] = nu[up] - i + j + 1;
				k = 1;
				pool_add(nu[p], p);
				bp[p] = p;
				us = i++;
				while (i < nu[up]) {
					qp = E(up, i); qs = R(up, i);
					N(p, k) = N(up, i);
					E(p, k) = qp; R(p, k) = qs;
					E(qp, qs) = p; R(qp, qs) = k;
					E(up, i) = -1;
					i++; k++;
				}
				i = 0;
				while (i < j) {
					qp = E(up, i); qs = R(up, i);
					N(p, k) = N(up, i);
					E(p, k) = qp; R(p, k) = qs;
					E(qp, qs) = p; R(qp, qs) = k;
					E(up, i) = -1;
					i++; k++;
				}
				qs = j;
			}
			if (!double_edge) {
				N(p, k) = N(up, qs);
				N(p, 0) = pid;
			} else N(p, 0) = N(up, qs);
			ds2.push_back(up);
			cs = k;
			qp = up; q = u;
			i = E(up, us);
			us = R(up, us);
			up = i;
			bp[qp] = -p;
		} else {
			ds.push_back(up);
			r = u / (u - l); l = 1 - r;
			pts[4 * p] = pts[4 * lp] * r + pts[4 * up] * l;
			pts[4 * p + 1] = pts[4 * lp + 1] * r + pts[4 * up + 1] * l;
			pts[4 * p + 2] = pts[4 * lp + 2] * r + pts[4 * up + 2] * l;
			nu[p] = 3;
			pool_add(3, p);
			N(p, 0) = pid;
			N(p, 1) = N(up, us);
			N(p, 2) = N(lp, ls);
			bp[p] = p;
			E(up, us) = -1;
			E(lp, ls) = p;
			R(lp, ls) = 1;
			E(p, 1) = lp;
			R(p, 1) = ls;
			cs = 2;
			qs = cycle_up(us, up);
			qp = up; q = u;
		}
		cp = p; rp = p; p++;
		while (qp != up || qs != us) {
			lp = E(qp, qs);
			lw = m_testx(lp, l);
			if (lw == 2) {
				qs = cycle_up(R(qp, qs), lp);
				qp = lp;
				q = l;
				ds.push_back(qp);
			} else if (lw == 0) {
				need_vertex(p);
				r = q / (q - l); l = 1 - r;
				pts[4 * p] = pts[4 * lp] * r + pts[4 * qp] * l;
				pts[4 * p + 1] = pts[4 * lp + 1] * r + pts[4 * qp + 1] * l;
				pts[4 * p + 2] = pts[4 * lp + 2] * r + pts[4 * qp + 2] * l;
				nu[p] = 3;
				pool_add(3, p);
				ls = R(qp, qs);
				N(p, 0) = pid;
				N(p, 1) = N(qp, qs);
				N(p, 2) = N(lp, ls);
				E(p, 0) = cp;
				E(p, 1) = lp;
				R(p, 0) = cs;
				R(p, 1) = ls;
				bp[p] = p;
				E(lp, ls) = p;
				R(lp, ls) = 1;
				E(cp, cs) = p;
				R(cp, cs) = 0;
				E(qp, qs) = -1;
				qs = cycle_up(qs, qp);
				cp = p++;
				cs = 2;
			} else {
				need_vertex(p);
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
				j = -bp[qp];
				if (qp == up && qs == us) {
					new_double_edge = false;
					if (j > 0) k += nu[j];
				} else {
					if (j > 0) {
						k += nu[j];
						if (lw == 1) {
							i = -bp[lp];
							if (i > 0) {
								if (E(i, nu[i] - 1) == j) {
									new_double_edge = true;
									k -= 1;
								} else new_double_edge = false;
							} else {
								if (j == rp && lp == up && R(qp, qs) == us) {
									new_double_edge = true;
									k -= 1;
								} else new_double_edge = false;
							}
						} else new_double_edge = false;
					} else {
						if (lw == 1) {
							i = -bp[lp];
							if (i == cp) {
								new_double_edge = true;
								k -= 1;
							} else new_double_edge = false;
						} else new_double_edge = false;
					}
				}
				need_order(k);
				if (j > 0) {
					if (nu[j] != k) {
						// Vertex j grows to order k: new pool row, old row
						// removed (src/cell.cc:1252-1281); edges stay in place.
						int old = nu[j], s_old = slot[j];
						pool_add(k, j);
						pool_remove_at(old, s_old);
						bp[j] = j;
					}
					i = nu[j];
				} else {
					pool_add(k, p);
					bp[p] = p;
					ds2.push_back(qp);
					copy_pts(p, qp);
					bp[qp] = -p;
					j = p++;
					i = 0;
				}
				nu[j] = k;
				if (!double_edge) {
					E(j, i) = cp;
					R(j, i) = cs;
					N(j, i) = pid;
					E(cp, cs) = j;
					R(cp, cs) = i;
					i++;
				}
				qs = iqs;
				while (i < (new_double_edge ? k : k - 1)) {
					qs = cycle_up(qs, qp);
					lp = E(qp, qs); ls = R(qp, qs);
					N(j, i) = N(qp, qs);
					E(j, i) = lp;
					R(j, i) = ls;
					E(lp, ls) = j;
					R(lp, ls) = i;
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
		E(cp, cs) = rp;
		E(rp, 0) = cp;
		R(cp, cs) = 0;
		R(rp, 0) = cs;
		return false;
	}

	// src/cell.cc:1473-1513.  Removes edge k of vertex j.
	bool delete_connection(int j, int k, bool hand) {
		int q = hand ? k : cycle_up(k, j);
		int i = nu[j] - 1, l, m;
		if (i < 1) return false;
		int te[4096], tr[4096], tn[4096];
		for (l = 0; l < q; l++) tn[l] = N(j, l);
		while (l < i) { tn[l] = N(j, l + 1); l++; }
		for (l = 0; l < k; l++) { te[l] = E(j, l); tr[l] = R(j, l); }
		while (l < i) {
			m = E(j, l + 1);
			te[l] = m;
			int kk = R(j, l + 1);
			tr[l] = kk;
			R(m, kk)--;
			l++;
		}
		int old = nu[j], s_old = slot[j];
		pool_add(i, j);
		pool_remove_at(old, s_old);
		for (l = 0; l < i; l++) { E(j, l) = te[l]; R(j, l) = tr[l]; N(j, l) = tn[l]; }
		nu[j] = i;
		return true;
	}

	// src/cell.cc:1441-1464
	bool collapse_order1() {
		while (!pool[1].empty()) {
			up = 0;
			int i = pool[1].back();
			pool[1].pop_back();
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
	bool collapse_order2() {
		if (!collapse_order1()) return false;
		while (pool.size() > 2 && !pool[2].empty()) {
			int i = pool[2].back();
			pool[2].pop_back();
			int j = E(i, 0), k = E(i, 1);
			if (j == k) return false;
			int l;
			for (l = 0; l < nu[j]; l++) if (E(j, l) == k) break;
			int a = R(i, 0), b = R(i, 1);
			if (l == nu[j]) {
				E(j, a) = k;
				E(k, b) = j;
				R(j, a) = b;
				R(k, b) = a;
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

	// src/cell.cc:718-870.  Returns false if the cell was deleted.
	bool nplane(double x, double y, double z, double rsq, int pid) {
		n_cuts++;
		int i, j, lp = up, cp, qp;
		int us = 0, ls = 0;
		unsigned uw, lw;
		double u, l = 0;
		up = 0;
		px = x; py = y; pz = z; prsq = rsq;
		maskc += 4;
		if (maskc < 4) { for (int t = 0; t < cap; t++) mask[t] = 0; maskc = 4; }
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
		ds.clear(); ds2.clear(); xse.clear();
		int op = p;
		if (create_facet(lp, ls, l, us, u, pid)) return false;
		size_t k = 0;
		while (k < xse.size()) {
			lp = xse[k++];
			uw = m_test(lp, l);
			for (ls = 0; ls < nu[lp]; ls++) {
				up = E(lp, ls);
				uw = m_test(up, u);
				if (up >= op) continue;
				if (uw == 0) {
					if (u > -big_tol && bp[up] != -1) {
						bp[up] = -1;
						xse.push_back(up);
					}
				} else if (uw == 1) {
					if (create_facet(-1, 0, 0, 0, u, pid)) return false;
				} else {
					us = R(lp, ls);
					m_test(lp, l);
					if (create_facet(lp, ls, l, us, u, pid)) return false;
				}
			}
		}
		if (p > max_p) max_p = p;
		for (size_t t = 0; t < xse.size(); t++) bp[xse[t]] = xse[t];
		size_t d = 0;
		while (d < ds.size()) {
			j = ds[d];
			if (R(j, 0) != -1) { R(j, 0) = -1; d++; }
			else { ds[d] = ds.back(); ds.pop_back(); }
		}
		for (size_t t = 0; t < ds2.size(); t++) {
			j = ds2[t];
			bp[j] = j;
			if (R(j, 0) != -1) { R(j, 0) = -1; ds.push_back(j); }
		}
		for (d = 0; d < ds.size(); d++) {
			cp = ds[d];
			for (int e = 0; e < nu[cp]; e++) {
				qp = E(cp, e);
				if (qp != -1 && R(qp, 0) != -1) { ds.push_back(qp); R(qp, 0) = -1; }
			}
		}
		up = 0;
		while (!ds.empty()) {
			--p;
			while (R(p, 0) == -1) {
				pool_remove(nu[p], p);
				--p;
			}
			up = ds.back(); ds.pop_back();
			if (up < p) {
				copy_pts(up, p);
				pool_remove(nu[up], up);
				move_vertex(p, up);
			} else up = p++;
		}
		(void)i;
		if (!pool[0].empty()) { fatal = true; return false; }
		return collapse_order2();
	}

	// src/cell.cc:1907-1917
	double max_radius_squared() {
		double r = pts[0] * pts[0] + pts[1] * pts[1] + pts[2] * pts[2], s;
		for (int v = 1; v < p; v++) {
			const double *q = &pts[4 * v];
			s = q[0] * q[0];
			s += q[1] * q[1];
			s += q[2] * q[2];
			if (s > r) r = s;
		}
		return r;
	}

	// src/cell.cc:2366-2410
	bool plane_intersects_track(double x, double y, double z, double rsq) {
		for (int tp = 0; tp < p; tp++)
			if (x * pts[4 * tp] + y * pts[4 * tp + 1] + z * pts[4 * tp + 2] > rsq) return true;
		return false;
	}
	bool plane_intersects(double x, double y, double z, double rsq) {
		double g = x * pts[4 * up] + y * pts[4 * up + 1] + z * pts[4 * up + 2];
		if (g < rsq) return plane_intersects_track(x, y, z, rsq);
		return true;
	}
	bool plane_intersects_guess(double x, double y, double z, double rsq) {
		up = 0;
		double g = x * pts[4 * up] + y * pts[4 * up + 1] + z * pts[4 * up + 2];
		if (g < rsq) {
			int ca = 1, cc = p >> 3, mp = 1;
			double m;
			while (ca < cc) {
				m = x * pts[4 * mp] + y * pts[4 * mp + 1] + z * pts[4 * mp + 2];
				if (m > g) {
					if (m > rsq) return true;
					g = m; up = mp;
				}
				ca += mp++;
			}
			return plane_intersects_track(x, y, z, rsq);
		}
		return true;
	}

	// Volume, centroid, face count and neighbours in one face walk.  The
	// walk visits faces in the order of src/cell.cc:1584-1616 (volume),
	// 1861-1901 (centroid), 2168-2187 (faces) and 2660-2678 (neighbours),
	// which all share it, and accumulates in the same floating-point order.
	void outputs(double &vol, double &cx, double &cy, double &cz, int &nf, std::vector<int> &nb) {
		const double fe = 1 / 48.0;
		double v = 0, tvol;
		double ux, uy, uz, vx, vy, vz, wx, wy, wz;
		cx = cy = cz = 0;
		nf = 0;
		nb.clear();
		for (int i = 1; i < p; i++) {
			ux = pts[0] - pts[4 * i];
			uy = pts[1] - pts[4 * i + 1];
			uz = pts[2] - pts[4 * i + 2];
			for (int j = 0; j < nu[i]; j++) {
				int k = E(i, j);
				if (k < 0) continue;
				nf++;
				nb.push_back(N(i, j));
				E(i, j) = -1 - k;
				int l = cycle_up(R(i, j), k);
				vx = pts[4 * k] - pts[0];
				vy = pts[4 * k + 1] - pts[1];
				vz = pts[4 * k + 2] - pts[2];
				int m = E(k, l);
				E(k, l) = -1 - m;
				while (m != i) {
					int n = cycle_up(R(k, l), m);
					wx = pts[4 * m] - pts[0];
					wy = pts[4 * m + 1] - pts[1];
					wz = pts[4 * m + 2] - pts[2];
					tvol = ux * vy * wz + uy * vz * wx + uz * vx * wy - uz * vy * wx - uy * vx * wz - ux * vz * wy;
					v += tvol;
					cx += (wx + vx - ux) * tvol;
					cy += (wy + vy - uy) * tvol;
					cz += (wz + vz - uz) * tvol;
					k = m; l = n; vx = wx; vy = wy; vz = wz;
					m = E(k, l);
					E(k, l) = -1 - m;
				}
			}
		}
		for (int i = 0; i < p; i++) for (int j = 0; j < nu[i]; j++) {
			if (E(i, j) >= 0) fatal = true;
			E(i, j) = -1 - E(i, j);
		}
		vol = v * fe;
		if (v > tol_cu) {
			v = 0.125 / v;
			cx = cx * v + 0.5 * pts[0];
			cy = cy * v + 0.5 * pts[1];
			cz = cz * v + 0.5 * pts[2];
		} else cx = cy = cz = 0;
	}
};

// ---------------------------------------------------------------------------
// Containers.

struct Spec {
	int kind;            // 0 container, 1 container_poly, 2 container_periodic, 3 container_periodic_poly
	double geom[6];
	int grid[3];
	int periodic[3];
	int init_mem;
	int n_spheres;
	const double *spheres;
	int n_walls;         // plane, cylinder and cone walls are not restated here
	const double *walls;
};

struct Out {
	int *id;
	int *ijk;
	int *q;
	int *exists;
	double *vol;
	double *cent;
	int *nfaces;
	long long *nbr_off;
	int *nbr;
	long long nbr_cap;
	double *pos;
};

// Per-cell statistics recorded by the restatement (candidate counts, exit
// depth in the worklist, phase reached, peak sizes).
struct Stats {
	int *n_planes;       // nplane calls for the cell
	int *wl_exit;        // worklist index g at which the security test stopped (64 = phase 3)
	int *phase;          // 1, 2 or 3
	int *max_p;          // peak vertex count
	int *max_order;      // peak vertex order
};

static inline int step_int(double a) { return a < 0 ? int(a) - 1 : int(a); }
static inline int step_mod(int a, int b) { return a >= 0 ? a % b : b - 1 - (b - 1 - a) % b; }
static inline int step_div(int a, int b) { return a >= 0 ? a / b : -1 + (a + 1) / b; }

struct RectContainer {
	// geometry (src/container.cc:30-44, src/v_base.cc:23-25)
	double ax, bx, ay, by, az, bz;
	int nx, ny, nz, nxy, nxyz;
	double boxx, boxy, boxz, xsp, ysp, zsp;
	bool xper, yper, zper;
	double max_len_sq;
	int ps;
	bool poly;
	double max_radius = 0;
	std::vector<std::vector<int> > id;
	std::vector<std::vector<double> > pp;
	std::vector<double> mrad;
	std::vector<double> walls;   // spheres: xc,yc,zc,rc,w_id
	// compute-engine state (src/v_compute.cc:22-31)
	int hx, hy, hz, hxy, hxyz;
	double bxsq;
	unsigned mv = 0;
	std::vector<unsigned> bmask;
	// radical state (src/rad_option.hh:98-156)
	double r_rad = 0, r_mul = 0, r_val = 0;

	RectContainer(const Spec &s) {
		ax = s.geom[0]; bx = s.geom[1]; ay = s.geom[2]; by = s.geom[3]; az = s.geom[4]; bz = s.geom[5];
		nx = s.grid[0]; ny = s.grid[1]; nz = s.grid[2];
		nxy = nx * ny; nxyz = nxy * nz;
		xper = s.periodic[0] != 0; yper = s.periodic[1] != 0; zper = s.periodic[2] != 0;
		boxx = (bx - ax) / nx; boxy = (by - ay) / ny; boxz = (bz - az) / nz;
		xsp = 1 / boxx; ysp = 1 / boxy; zsp = 1 / boxz;
		max_len_sq = (bx - ax) * (bx - ax) * (xper ? 0.25 : 1) + (by - ay) * (by - ay) * (yper ? 0.25 : 1)
			     + (bz - az) * (bz - az) * (zper ? 0.25 : 1);
		poly = s.kind == 1;
		ps = poly ? 4 : 3;
		id.resize(nxyz); pp.resize(nxyz);
		for (int w = 0; w < s.n_spheres; w++) for (int t = 0; t < 5; t++) walls.push_back(s.spheres[5 * w + t]);
		hx = xper ? 2 * nx + 1 : nx; hy = yper ? 2 * ny + 1 : ny; hz = zper ? 2 * nz + 1 : nz;
		hxy = hx * hy; hxyz = hxy * hz;
		bxsq = boxx * boxx + boxy * boxy + boxz * boxz;
		bmask.assign(hxyz, 0u);
		init_mrad();
	}

	// src/v_base.cc:81-94
	void compute_minimum(double &minr, double xlo, double xhi, double ylo, double yhi, double zlo, double zhi,
			     int ti, int tj, int tk) {
		double radsq, temp;
		if (ti > 0) { temp = boxx * ti - xhi; radsq = temp * temp; }
		else if (ti < 0) { temp = xlo - boxx * (1 + ti); radsq = temp * temp; }
		else radsq = 0;
		if (tj > 0) { temp = boxy * tj - yhi; radsq += temp * temp; }
		else if (tj < 0) { temp = ylo - boxy * (1 + tj); radsq += temp * temp; }
		if (tk > 0) { temp = boxz * tk - zhi; radsq += temp * temp; }
		else if (tk < 0) { temp = zlo - boxz * (1 + tk); radsq += temp * temp; }
		if (radsq < minr) minr = radsq;
	}

	// src/v_base.cc:23-69
	void init_mrad() {
		const unsigned b1 = 1u << 21, b2 = 1u << 22, b3 = 1u << 24, b4 = 1u << 25, b5 = 1u << 27, b6 = 1u << 28;
		const double xstep = boxx / VGPU_WL_FGRID, ystep = boxy / VGPU_WL_FGRID, zstep = boxz / VGPU_WL_FGRID;
		mrad.assign(64 * 64, 0.0);
		const unsigned *e = vgpu_worklist_table;
		double *radp = mrad.data();
		int lx, ly, lz, q, i, j, k;
		double xlo, ylo, zlo, xhi, yhi, zhi, minr;
		unsigned f;
		for (zlo = 0, zhi = zstep, lz = 0; lz < VGPU_WL_HGRID; zlo = zhi, zhi += zstep, lz++)
		for (ylo = 0, yhi = ystep, ly = 0; ly < VGPU_WL_HGRID; ylo = yhi, yhi += ystep, ly++)
		for (xlo = 0, xhi = xstep, lx = 0; lx < VGPU_WL_HGRID; xlo = xhi, xhi += xstep, lx++) {
			minr = std::numeric_limits<double>::max();
			for (q = e[0] + 1; q < 64; q++) {
				f = e[q];
				i = (f & 127) - 64; j = (f >> 7 & 127) - 64; k = (f >> 14 & 127) - 64;
				if ((f & b2) == b2) {
					compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i - 1, j, k);
					if ((f & b1) == 0) compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i + 1, j, k);
				} else if ((f & b1) == b1) compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i + 1, j, k);
				if ((f & b4) == b4) {
					compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i, j - 1, k);
					if ((f & b3) == 0) compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i, j + 1, k);
				} else if ((f & b3) == b3) compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i, j + 1, k);
				if ((f & b6) == b6) {
					compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k - 1);
					if ((f & b5) == 0) compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k + 1);
				} else if ((f & b5) == b5) compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k + 1);
			}
			q--;
			while (q > 0) {
				radp[q] = minr;
				f = e[q];
				i = (f & 127) - 64; j = (f >> 7 & 127) - 64; k = (f >> 14 & 127) - 64;
				compute_minimum(minr, xlo, xhi, ylo, yhi, zlo, zhi, i, j, k);
				q--;
			}
			*radp = minr;
			e += 64;
			radp += 64;
		}
	}

	// src/container.cc:90-191
	bool put(int n, double x, double y, double z, double r) {
		int ijk = step_int((x - ax) * xsp), l;
		if (xper) { l = step_mod(ijk, nx); x += boxx * (l - ijk); ijk = l; }
		else if (ijk < 0 || ijk >= nx) return false;
		int j = step_int((y - ay) * ysp);
		if (yper) { l = step_mod(j, ny); y += boxy * (l - j); j = l; }
		else if (j < 0 || j >= ny) return false;
		int k = step_int((z - az) * zsp);
		if (zper) { l = step_mod(k, nz); z += boxz * (l - k); k = l; }
		else if (k < 0 || k >= nz) return false;
		ijk += nx * j + nxy * k;
		id[ijk].push_back(n);
		pp[ijk].push_back(x); pp[ijk].push_back(y); pp[ijk].push_back(z);
		if (poly) { pp[ijk].push_back(r); if (max_radius < r) max_radius = r; }
		return true;
	}
	int co(int ijk) const { return (int)id[ijk].size(); }
	double P(int ijk, int l, int c) const { return pp[ijk][ps * l + c]; }

	// Radical / mono policies, src/rad_option.hh
	void r_init(int ijk, int s) {
		if (!poly) return;
		double r = P(ijk, s, 3);
		r_rad = r * r;
		r_mul = (r - max_radius) * (r + max_radius);
	}
	void r_prime(double rv) { if (poly) r_val = 1 + r_mul / rv; }
	bool r_ctest(double crs, double mrs) const {
		if (!poly) return crs > mrs;
		double cc = crs + r_mul;
		return cc > 0 && cc * cc > mrs * crs;
	}
	double r_cutoff(double lrs) const { return poly ? lrs * r_val : lrs; }
	double r_scale(double rs, int ijk, int q) const {
		if (!poly) return rs;
		return rs + r_rad - P(ijk, q, 3) * P(ijk, q, 3);
	}
	bool r_scale_check(double &rs, double mrs, int ijk, int q) const {
		if (!poly) return rs < mrs;
		double trs = rs;
		rs += r_rad - P(ijk, q, 3) * P(ijk, q, 3);
		return rs <= 0 || rs * rs < mrs * trs;
	}

	// src/container.hh:227-251
	int region_index(int ci, int cj, int ck, int ei, int ej, int ek, double &qx, double &qy, double &qz, int disp) {
		if (xper) { if (ci + ei < nx) { ei += nx; qx = -(bx - ax); } else if (ci + ei >= (nx << 1)) { ei -= nx; qx = bx - ax; } else qx = 0; }
		if (yper) { if (cj + ej < ny) { ej += ny; qy = -(by - ay); } else if (cj + ej >= (ny << 1)) { ej -= ny; qy = by - ay; } else qy = 0; }
		if (zper) { if (ck + ek < nz) { ek += nz; qz = -(bz - az); } else if (ck + ek >= (nz << 1)) { ek -= nz; qz = bz - az; } else qz = 0; }
		return disp + ei + nx * (ej + ny * ek);
	}

	// src/v_compute.cc:797-944
	bool compute_min_max_radius(int di, int dj, int dk, double fx, double fy, double fz, double gxs, double gys,
				    double gzs, double &crs, double mrs) {
		double xlo, ylo, zlo;
		if (di > 0) {
			xlo = di * boxx - fx;
			crs = xlo * xlo;
			if (dj > 0) {
				ylo = dj * boxy - fy;
				crs += ylo * ylo;
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (boxx * xlo + boxy * ylo + boxz * zlo);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (boxx * xlo + boxy * ylo - boxz * zlo);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += boxx * (2 * xlo + boxx) + boxy * (2 * ylo + boxy) + gzs;
				}
			} else if (dj < 0) {
				ylo = (dj + 1) * boxy - fy;
				crs += ylo * ylo;
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (boxx * xlo - boxy * ylo + boxz * zlo);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (boxx * xlo - boxy * ylo - boxz * zlo);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += boxx * (2 * xlo + boxx) + boxy * (-2 * ylo + boxy) + gzs;
				}
			} else {
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (2 * zlo + boxz);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (-2 * zlo + boxz);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += gzs;
				}
				crs += gys + boxx * (2 * xlo + boxx);
			}
		} else if (di < 0) {
			xlo = (di + 1) * boxx - fx;
			crs = xlo * xlo;
			if (dj > 0) {
				ylo = dj * boxy - fy;
				crs += ylo * ylo;
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (-boxx * xlo + boxy * ylo + boxz * zlo);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (-boxx * xlo + boxy * ylo - boxz * zlo);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += boxx * (-2 * xlo + boxx) + boxy * (2 * ylo + boxy) + gzs;
				}
			} else if (dj < 0) {
				ylo = (dj + 1) * boxy - fy;
				crs += ylo * ylo;
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (-boxx * xlo - boxy * ylo + boxz * zlo);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += bxsq + 2 * (-boxx * xlo - boxy * ylo - boxz * zlo);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += boxx * (-2 * xlo + boxx) + boxy * (-2 * ylo + boxy) + gzs;
				}
			} else {
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (2 * zlo + boxz);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (-2 * zlo + boxz);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += gzs;
				}
				crs += gys + boxx * (-2 * xlo + boxx);
			}
		} else {
			if (dj > 0) {
				ylo = dj * boxy - fy;
				crs = ylo * ylo;
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (2 * zlo + boxz);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (-2 * zlo + boxz);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += gzs;
				}
				crs += boxy * (2 * ylo + boxy);
			} else if (dj < 0) {
				ylo = (dj + 1) * boxy - fy;
				crs = ylo * ylo;
				if (dk > 0) {
					zlo = dk * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (2 * zlo + boxz);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz;
					crs += zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (-2 * zlo + boxz);
				} else {
					if (r_ctest(crs, mrs)) return true;
					crs += gzs;
				}
				crs += boxy * (-2 * ylo + boxy);
			} else {
				if (dk > 0) {
					zlo = dk * boxz - fz; crs = zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (2 * zlo + boxz);
				} else if (dk < 0) {
					zlo = (dk + 1) * boxz - fz; crs = zlo * zlo; if (r_ctest(crs, mrs)) return true;
					crs += boxz * (-2 * zlo + boxz);
				} else {
					crs = 0;
				}
				crs += gys;
			}
			crs += gxs;
		}
		return false;
	}

	// Region tests, src/v_compute.cc:634-780
	bool corner_test(Cell &c, double xl, double yl, double zl, double xh, double yh, double zh) {
		r_prime(xl * xl + yl * yl + zl * zl);
		if (c.plane_intersects_guess(xh, yl, zl, r_cutoff(xl * xh + yl * yl + zl * zl))) return false;
		if (c.plane_intersects(xh, yh, zl, r_cutoff(xl * xh + yl * yh + zl * zl))) return false;
		if (c.plane_intersects(xl, yh, zl, r_cutoff(xl * xl + yl * yh + zl * zl))) return false;
		if (c.plane_intersects(xl, yh, zh, r_cutoff(xl * xl + yl * yh + zl * zh))) return false;
		if (c.plane_intersects(xl, yl, zh, r_cutoff(xl * xl + yl * yl + zl * zh))) return false;
		if (c.plane_intersects(xh, yl, zh, r_cutoff(xl * xh + yl * yl + zl * zh))) return false;
		return true;
	}
	bool edge_x_test(Cell &c, double x0, double yl, double zl, double x1, double yh, double zh) {
		r_prime(yl * yl + zl * zl);
		if (c.plane_intersects_guess(x0, yl, zh, r_cutoff(yl * yl + zl * zh))) return false;
		if (c.plane_intersects(x1, yl, zh, r_cutoff(yl * yl + zl * zh))) return false;
		if (c.plane_intersects(x1, yl, zl, r_cutoff(yl * yl + zl * zl))) return false;
		if (c.plane_intersects(x0, yl, zl, r_cutoff(yl * yl + zl * zl))) return false;
		if (c.plane_intersects(x0, yh, zl, r_cutoff(yl * yh + zl * zl))) return false;
		if (c.plane_intersects(x1, yh, zl, r_cutoff(yl * yh + zl * zl))) return false;
		return true;
	}
	bool edge_y_test(Cell &c, double xl, double y0, double zl, double xh, double y1, double zh) {
		r_prime(xl * xl + zl * zl);
		if (c.plane_intersects_guess(xl, y0, zh, r_cutoff(xl * xl + zl * zh))) return false;
		if (c.plane_intersects(xl, y1, zh, r_cutoff(xl * xl + zl * zh))) return false;
		if (c.plane_intersects(xl, y1, zl, r_cutoff(xl * xl + zl * zl))) return false;
		if (c.plane_intersects(xl, y0, zl, r_cutoff(xl * xl + zl * zl))) return false;
		if (c.plane_intersects(xh, y0, zl, r_cutoff(xl * xh + zl * zl))) return false;
		if (c.plane_intersects(xh, y1, zl, r_cutoff(xl * xh + zl * zl))) return false;
		return true;
	}
	bool edge_z_test(Cell &c, double xl, double yl, double z0, double xh, double yh, double z1) {
		r_prime(xl * xl + yl * yl);
		if (c.plane_intersects_guess(xl, yh, z0, r_cutoff(xl * xl + yl * yh))) return false;
		if (c.plane_intersects(xl, yh, z1, r_cutoff(xl * xl + yl * yh))) return false;
		if (c.plane_intersects(xl, yl, z1, r_cutoff(xl * xl + yl * yl))) return false;
		if (c.plane_intersects(xl, yl, z0, r_cutoff(xl * xl + yl * yl))) return false;
		if (c.plane_intersects(xh, yl, z0, r_cutoff(xl * xh + yl * yl))) return false;
		if (c.plane_intersects(xh, yl, z1, r_cutoff(xl * xh + yl * yl))) return false;
		return true;
	}
	bool face_x_test(Cell &c, double xl, double y0, double z0, double y1, double z1) {
		r_prime(xl * xl);
		if (c.plane_intersects_guess(xl, y0, z0, r_cutoff(xl * xl))) return false;
		if (c.plane_intersects(xl, y0, z1, r_cutoff(xl * xl))) return false;
		if (c.plane_intersects(xl, y1, z1, r_cutoff(xl * xl))) return false;
		if (c.plane_intersects(xl, y1, z0, r_cutoff(xl * xl))) return false;
		return true;
	}
	bool face_y_test(Cell &c, double x0, double yl, double z0, double x1, double z1) {
		r_prime(yl * yl);
		if (c.plane_intersects_guess(x0, yl, z0, r_cutoff(yl * yl))) return false;
		if (c.plane_intersects(x0, yl, z1, r_cutoff(yl * yl))) return false;
		if (c.plane_intersects(x1, yl, z1, r_cutoff(yl * yl))) return false;
		if (c.plane_intersects(x1, yl, z0, r_cutoff(yl * yl))) return false;
		return true;
	}
	bool face_z_test(Cell &c, double x0, double y0, double zl, double x1, double y1) {
		r_prime(zl * zl);
		if (c.plane_intersects_guess(x0, y0, zl, r_cutoff(zl * zl))) return false;
		if (c.plane_intersects(x0, y1, zl, r_cutoff(zl * zl))) return false;
		if (c.plane_intersects(x1, y1, zl, r_cutoff(zl * zl))) return false;
		if (c.plane_intersects(x1, y0, zl, r_cutoff(zl * zl))) return false;
		return true;
	}

	// src/v_compute.cc:236-264
	void add_to_mask(int ei, int ej, int ek, std::vector<int> &qu) {
		unsigned *mijk = &bmask[ei + hx * (ej + hy * ek)];
		if (ek > 0) if (*(mijk - hxy) != mv) { *(mijk - hxy) = mv; qu.push_back(ei); qu.push_back(ej); qu.push_back(ek - 1); }
		if (ej > 0) if (*(mijk - hx) != mv) { *(mijk - hx) = mv; qu.push_back(ei); qu.push_back(ej - 1); qu.push_back(ek); }
		if (ei > 0) if (*(mijk - 1) != mv) { *(mijk - 1) = mv; qu.push_back(ei - 1); qu.push_back(ej); qu.push_back(ek); }
		if (ei < hx - 1) if (*(mijk + 1) != mv) { *(mijk + 1) = mv; qu.push_back(ei + 1); qu.push_back(ej); qu.push_back(ek); }
		if (ej < hy - 1) if (*(mijk + hx) != mv) { *(mijk + hx) = mv; qu.push_back(ei); qu.push_back(ej + 1); qu.push_back(ek); }
		if (ek < hz - 1) if (*(mijk + hxy) != mv) { *(mijk + hxy) = mv; qu.push_back(ei); qu.push_back(ej); qu.push_back(ek + 1); }
	}
	void scan_bits_mask_add(unsigned q, unsigned *mijk, int ei, int ej, int ek, std::vector<int> &qu) {
		const unsigned b1 = 1u << 21, b2 = 1u << 22, b3 = 1u << 24, b4 = 1u << 25, b5 = 1u << 27, b6 = 1u << 28;
		if ((q & b2) == b2) {
			if (ei > 0) { *(mijk - 1) = mv; qu.push_back(ei - 1); qu.push_back(ej); qu.push_back(ek); }
			if ((q & b1) == 0 && ei < hx - 1) { *(mijk + 1) = mv; qu.push_back(ei + 1); qu.push_back(ej); qu.push_back(ek); }
		} else if ((q & b1) == b1 && ei < hx - 1) { *(mijk + 1) = mv; qu.push_back(ei + 1); qu.push_back(ej); qu.push_back(ek); }
		if ((q & b4) == b4) {
			if (ej > 0) { *(mijk - hx) = mv; qu.push_back(ei); qu.push_back(ej - 1); qu.push_back(ek); }
			if ((q & b3) == 0 && ej < hy - 1) { *(mijk + hx) = mv; qu.push_back(ei); qu.push_back(ej + 1); qu.push_back(ek); }
		} else if ((q & b3) == b3 && ej < hy - 1) { *(mijk + hx) = mv; qu.push_back(ei); qu.push_back(ej + 1); qu.push_back(ek); }
		if ((q & b6) == b6) {
			if (ek > 0) { *(mijk - hxy) = mv; qu.push_back(ei); qu.push_back(ej); qu.push_back(ek - 1); }
			if ((q & b5) == 0 && ek < hz - 1) { *(mijk + hxy) = mv; qu.push_back(ei); qu.push_back(ej); qu.push_back(ek + 1); }
		} else if ((q & b5) == b5 && ek < hz - 1) { *(mijk + hxy) = mv; qu.push_back(ei); qu.push_back(ej); qu.push_back(ek + 1); }
	}

	// Candidate cut in block ijk, particle l (plain or radical).
	// src/v_compute.cc:291-621.  Returns false if the cell was deleted.
	bool compute_cell(Cell &c, int ijk, int s, int ci, int cj, int ck, int *wl_exit, int *phase) {
		static const int count_list[8] = {7, 11, 15, 19, 26, 35, 45, 59};
		double x, y, z, x1, y1, z1, qx = 0, qy = 0, qz = 0;
		double xlo, ylo, zlo, xhi, yhi, zhi, x2, y2, z2, rs;
		int i, j, k, di, dj, dk, ei, ej, ek, f, g, l, disp;
		double fx, fy, fz, gxs, gys, gzs;
		const double *radp;
		unsigned q, *mijk;
		const unsigned *e;
		*wl_exit = 0; *phase = 1;

		// container_base::initialize_voronoicell, src/container.hh:192-203
		x = P(ijk, s, 0); y = P(ijk, s, 1); z = P(ijk, s, 2);
		double xa, xb, ya, yb, za, zb;
		if (xper) { xa = -(xb = 0.5 * (bx - ax)); i = nx; } else { xa = ax - x; xb = bx - x; i = ci; }
		if (yper) { ya = -(yb = 0.5 * (by - ay)); j = ny; } else { ya = ay - y; yb = by - y; j = cj; }
		if (zper) { za = -(zb = 0.5 * (bz - az)); k = nz; } else { za = az - z; zb = bz - z; k = ck; }
		c.init_box(xa, xb, ya, yb, za, zb);
		for (size_t w = 0; w < walls.size(); w += 5) {
			// wall_sphere::cut_cell_base, src/wall.cc:28-36
			double xd = x - walls[w], yd = y - walls[w + 1], zd = z - walls[w + 2], dq = xd * xd + yd * yd + zd * zd;
			if (dq > 1e-5) {
				dq = 2 * (std::sqrt(dq) * walls[w + 3] - dq);
				if (!c.nplane(xd, yd, zd, dq, (int)walls[w + 4])) return false;
			}
		}
		disp = ijk - i - nx * (j + ny * k);
		r_init(ijk, s);

		double crs, mrs;
		int next_count = 3, cpi = 0;

		for (l = 0; l < s; l++) {
			x1 = P(ijk, l, 0) - x; y1 = P(ijk, l, 1) - y; z1 = P(ijk, l, 2) - z;
			rs = r_scale(x1 * x1 + y1 * y1 + z1 * z1, ijk, l);
			if (!c.nplane(x1, y1, z1, rs, id[ijk][l])) return false;
		}
		l++;
		while (l < co(ijk)) {
			x1 = P(ijk, l, 0) - x; y1 = P(ijk, l, 1) - y; z1 = P(ijk, l, 2) - z;
			rs = r_scale(x1 * x1 + y1 * y1 + z1 * z1, ijk, l);
			if (!c.nplane(x1, y1, z1, rs, id[ijk][l])) return false;
			l++;
		}
		mrs = c.max_radius_squared();

		unsigned m1, m2;
		// frac_pos, src/container.hh:227-232
		fx = x - ax - boxx * ci; fy = y - ay - boxy * cj; fz = z - az - boxz * ck;
		di = int(fx * xsp * VGPU_WL_FGRID); dj = int(fy * ysp * VGPU_WL_FGRID); dk = int(fz * zsp * VGPU_WL_FGRID);
		if (di >= VGPU_WL_HGRID) {
			gxs = fx;
			m1 = 127 + (3 << 21); m2 = 1 + (1 << 21); di = VGPU_WL_FGRID - 1 - di; if (di < 0) di = 0;
		} else { m1 = m2 = 0; gxs = boxx - fx; }
		if (dj >= VGPU_WL_HGRID) {
			gys = fy;
			m1 |= (127 << 7) + (3 << 24); m2 |= (1 << 7) + (1 << 24); dj = VGPU_WL_FGRID - 1 - dj; if (dj < 0) dj = 0;
		} else gys = boxy - fy;
		if (dk >= VGPU_WL_HGRID) {
			gzs = fz;
			m1 |= (127 << 14) + (3 << 27); m2 |= (1 << 14) + (1 << 27); dk = VGPU_WL_FGRID - 1 - dk; if (dk < 0) dk = 0;
		} else gzs = boxz - fz;
		gxs *= gxs; gys *= gys; gzs *= gzs;

		int wli = di + VGPU_WL_HGRID * (dj + VGPU_WL_HGRID * dk);
		radp = &mrad[wli * 64];
		e = vgpu_worklist_table + wli * 64;

		f = e[0]; g = 0;
		do {
			if (g == next_count) {
				mrs = c.max_radius_squared();
				if (cpi != 8) next_count = count_list[cpi++];
			}
			if (r_ctest(radp[g], mrs)) { *wl_exit = g; return true; }
			g++;
			q = e[g]; q ^= m1; q += m2;
			di = q & 127; di -= 64;
			dj = (q >> 7) & 127; dj -= 64;
			dk = (q >> 14) & 127; dk -= 64;
			ei = di + i; if (ei < 0 || ei >= hx) continue;
			ej = dj + j; if (ej < 0 || ej >= hy) continue;
			ek = dk + k; if (ek < 0 || ek >= hz) continue;
			if (compute_min_max_radius(di, dj, dk, fx, fy, fz, gxs, gys, gzs, crs, mrs)) continue;
			int rijk = region_index(ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
			if (co(rijk) > 0) {
				l = 0; x2 = x - qx; y2 = y - qy; z2 = z - qz;
				if (!r_ctest(crs, mrs)) {
					do {
						x1 = P(rijk, l, 0) - x2; y1 = P(rijk, l, 1) - y2; z1 = P(rijk, l, 2) - z2;
						rs = r_scale(x1 * x1 + y1 * y1 + z1 * z1, rijk, l);
						if (!c.nplane(x1, y1, z1, rs, id[rijk][l])) return false;
						l++;
					} while (l < co(rijk));
				} else {
					do {
						x1 = P(rijk, l, 0) - x2; y1 = P(rijk, l, 1) - y2; z1 = P(rijk, l, 2) - z2;
						rs = x1 * x1 + y1 * y1 + z1 * z1;
						if (r_scale_check(rs, mrs, rijk, l) && !c.nplane(x1, y1, z1, rs, id[rijk][l])) return false;
						l++;
					} while (l < co(rijk));
				}
			}
		} while (g < f);

		*phase = 2;
		mv++;
		if (mv == 0) { for (size_t t = 0; t < bmask.size(); t++) bmask[t] = 0; mv = 1; }
		std::vector<int> qu;
		size_t qs = 0;

		while (g < 63) {
			if (g == next_count) {
				mrs = c.max_radius_squared();
				if (cpi != 8) next_count = count_list[cpi++];
			}
			if (r_ctest(radp[g], mrs)) { *wl_exit = g; return true; }
			g++;
			q = e[g]; q ^= m1; q += m2;
			di = q & 127; di -= 64;
			dj = (q >> 7) & 127; dj -= 64;
			dk = (q >> 14) & 127; dk -= 64;
			ei = di + i; if (ei < 0 || ei >= hx) continue;
			ej = dj + j; if (ej < 0 || ej >= hy) continue;
			ek = dk + k; if (ek < 0 || ek >= hz) continue;
			mijk = &bmask[ei + hx * (ej + hy * ek)];
			*mijk = mv;
			if (compute_min_max_radius(di, dj, dk, fx, fy, fz, gxs, gys, gzs, crs, mrs)) continue;
			int rijk = region_index(ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
			if (co(rijk) > 0) {
				l = 0; x2 = x - qx; y2 = y - qy; z2 = z - qz;
				if (!r_ctest(crs, mrs)) {
					do {
						x1 = P(rijk, l, 0) - x2; y1 = P(rijk, l, 1) - y2; z1 = P(rijk, l, 2) - z2;
						rs = r_scale(x1 * x1 + y1 * y1 + z1 * z1, rijk, l);
						if (!c.nplane(x1, y1, z1, rs, id[rijk][l])) return false;
						l++;
					} while (l < co(rijk));
				} else {
					do {
						x1 = P(rijk, l, 0) - x2; y1 = P(rijk, l, 1) - y2; z1 = P(rijk, l, 2) - z2;
						rs = x1 * x1 + y1 * y1 + z1 * z1;
						if (r_scale_check(rs, mrs, rijk, l) && !c.nplane(x1, y1, z1, rs, id[rijk][l])) return false;
						l++;
					} while (l < co(rijk));
				}
			}
			scan_bits_mask_add(q, mijk, ei, ej, ek, qu);
		}

		if (r_ctest(radp[g], mrs)) { *wl_exit = g; return true; }
		*wl_exit = 64;
		*phase = 3;

		while (qs < qu.size()) {
			ei = qu[qs]; ej = qu[qs + 1]; ek = qu[qs + 2]; qs += 3;
			xlo = (ei - i) * boxx - fx; xhi = xlo + boxx;
			ylo = (ej - j) * boxy - fy; yhi = ylo + boxy;
			zlo = (ek - k) * boxz - fz; zhi = zlo + boxz;
			bool skip;
			if (ei > i) {
				if (ej > j) {
					if (ek > k) skip = corner_test(c, xlo, ylo, zlo, xhi, yhi, zhi);
					else if (ek < k) skip = corner_test(c, xlo, ylo, zhi, xhi, yhi, zlo);
					else skip = edge_z_test(c, xlo, ylo, zlo, xhi, yhi, zhi);
				} else if (ej < j) {
					if (ek > k) skip = corner_test(c, xlo, yhi, zlo, xhi, ylo, zhi);
					else if (ek < k) skip = corner_test(c, xlo, yhi, zhi, xhi, ylo, zlo);
					else skip = edge_z_test(c, xlo, yhi, zlo, xhi, ylo, zhi);
				} else {
					if (ek > k) skip = edge_y_test(c, xlo, ylo, zlo, xhi, yhi, zhi);
					else if (ek < k) skip = edge_y_test(c, xlo, ylo, zhi, xhi, yhi, zlo);
					else skip = face_x_test(c, xlo, ylo, zlo, yhi, zhi);
				}
			} else if (ei < i) {
				if (ej > j) {
					if (ek > k) skip = corner_test(c, xhi, ylo, zlo, xlo, yhi, zhi);
					else if (ek < k) skip = corner_test(c, xhi, ylo, zhi, xlo, yhi, zlo);
					else skip = edge_z_test(c, xhi, ylo, zlo, xlo, yhi, zhi);
				} else if (ej < j) {
					if (ek > k) skip = corner_test(c, xhi, yhi, zlo, xlo, ylo, zhi);
					else if (ek < k) skip = corner_test(c, xhi, yhi, zhi, xlo, ylo, zlo);
					else skip = edge_z_test(c, xhi, yhi, zlo, xlo, ylo, zhi);
				} else {
					if (ek > k) skip = edge_y_test(c, xhi, ylo, zlo, xlo, yhi, zhi);
					else if (ek < k) skip = edge_y_test(c, xhi, ylo, zhi, xlo, yhi, zlo);
					else skip = face_x_test(c, xhi, ylo, zlo, yhi, zhi);
				}
			} else {
				if (ej > j) {
					if (ek > k) skip = edge_x_test(c, xlo, ylo, zlo, xhi, yhi, zhi);
					else if (ek < k) skip = edge_x_test(c, xlo, ylo, zhi, xhi, yhi, zlo);
					else skip = face_y_test(c, xlo, ylo, zlo, xhi, zhi);
				} else if (ej < j) {
					if (ek > k) skip = edge_x_test(c, xlo, yhi, zlo, xhi, ylo, zhi);
					else if (ek < k) skip = edge_x_test(c, xlo, yhi, zhi, xhi, ylo, zlo);
					else skip = face_y_test(c, xlo, yhi, zlo, xhi, zhi);
				} else {
					if (ek > k) skip = face_z_test(c, xlo, ylo, zlo, xhi, yhi);
					else if (ek < k) skip = face_z_test(c, xlo, ylo, zhi, xhi, yhi);
					else { c.fatal = true; return false; }
				}
			}
			if (skip) continue;
			int rijk = region_index(ci, cj, ck, ei, ej, ek, qx, qy, qz, disp);
			if (co(rijk) > 0) {
				l = 0; x2 = x - qx; y2 = y - qy; z2 = z - qz;
				do {
					x1 = P(rijk, l, 0) - x2; y1 = P(rijk, l, 1) - y2; z1 = P(rijk, l, 2) - z2;
					rs = r_scale(x1 * x1 + y1 * y1 + z1 * z1, rijk, l);
					if (!c.nplane(x1, y1, z1, rs, id[rijk][l])) return false;
					l++;
				} while (l < co(rijk));
			}
			add_to_mask(ei, ej, ek, qu);
		}
		return true;
	}
};

}  // namespace orc

using namespace orc;

extern "C" {

// Same contract as ref_compute_neighbor in ref_driver.cc, for rectangular
// containers (kind 0/1).  `st` may be null.  Returns the number of cells, or
// -1 if the restatement hit a state the reference treats as fatal.
long long orc_compute_neighbor(const Spec *s, int n, const int *ids, const double *xyz, const double *r,
			       Out *o, Stats *st) {
	if (s->kind > 1) return -2;
	if (s->n_walls > 0) return -3;   // only sphere walls are restated
	RectContainer con(*s);
	for (int t = 0; t < n; t++) con.put(ids[t], xyz[3 * t], xyz[3 * t + 1], xyz[3 * t + 2], r ? r[t] : 0.5);
	Cell c(con.max_len_sq);
	std::vector<int> nb;
	long long m = 0, nn = 0;
	bool bad = false;
	for (int ijk = 0; ijk < con.nxyz; ijk++) {
		int ci = ijk % con.nx, cj = (ijk / con.nx) % con.ny, ck = ijk / con.nxy;
		for (int q = 0; q < con.co(ijk); q++) {
			int wl_exit, phase;
			c.n_cuts = 0; c.max_p = 0; c.max_order = 0;
			bool ok = con.compute_cell(c, ijk, q, ci, cj, ck, &wl_exit, &phase);
			if (c.fatal) bad = true;
			if (st) {
				st->n_planes[m] = (int)c.n_cuts;
				st->wl_exit[m] = wl_exit;
				st->phase[m] = phase;
				st->max_p[m] = c.max_p;
				st->max_order[m] = c.max_order;
			}
			if (o) {
				o->id[m] = con.id[ijk][q];
				o->ijk[m] = ijk;
				o->q[m] = q;
				o->exists[m] = ok ? 1 : 0;
				if (o->pos) { for (int t = 0; t < 3; t++) o->pos[3 * m + t] = con.P(ijk, q, t); }
				o->nbr_off[m] = nn;
				if (ok) {
					double v, cx, cy, cz;
					int nf;
					c.outputs(v, cx, cy, cz, nf, nb);
					o->vol[m] = v;
					o->cent[3 * m] = cx; o->cent[3 * m + 1] = cy; o->cent[3 * m + 2] = cz;
					o->nfaces[m] = nf;
					for (size_t t = 0; t < nb.size(); t++) {
						if (nn < o->nbr_cap) o->nbr[nn] = nb[t];
						nn++;
					}
				} else {
					o->vol[m] = 0; o->cent[3 * m] = o->cent[3 * m + 1] = o->cent[3 * m + 2] = 0;
					o->nfaces[m] = 0;
				}
			}
			m++;
		}
	}
	if (o) o->nbr_off[m] = nn;
	return bad ? -1 : m;
}

}
