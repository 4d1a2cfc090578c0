// tg_tree.cpp -- host-side tree, multipoles and interaction lists (see tg_tree.hpp).
//
// Every floating-point expression that decides the tree shape or a list entry is evaluated with
// the reference's operand order and precision: box midpoints and partition tests in FP64
// (src/tree.cpp:136-157, src/part_vect.cpp:905-926), centres of mass in FP64 then quantised
// (src/tree.cpp:229-247), P2M in float 16-lane sums (src/part_vect.cpp:584-617), M2M with float
// shifts (tigergrav/multipole.hpp:414-433), radii in FP64 rounded to float
// (src/tree.cpp:256-275), the MAC in FP64 with float radius sums (src/tree.cpp:428-429,452-459).
// Build with -ffp-contract=off so no multiply-add is fused where the reference rounds twice.
#include "tg_tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>

namespace tg {

int32_t double_to_pos(double x) {
	const double POS_MAX = 4294967296.0;
	const double y = x - 0.5;
	long z = std::lround(y * POS_MAX);
	if (z >= POS_MAX / 2.0) z -= (long) POS_MAX;
	return (int32_t) z;
}

void box_id_to_range(boxid_t id, double* mn, double* mx) {
	double x[3] = {0.0, 0.0, 0.0};
	double dx[3] = {1.0, 1.0, 1.0};
	while (id != 1) {
		// each level rotates the axes so the split dimension cycles x, y, z
		const double t1 = x[0], t2 = dx[0];
		x[0] = x[2];
		x[2] = x[1];
		x[1] = t1;
		dx[0] = dx[2];
		dx[2] = dx[1];
		dx[1] = t2;
		if (id & 1)
			x[0] = 0.5 * x[0] + 0.5;
		else
			x[0] = 0.5 * x[0];
		dx[0] *= 0.5;
		id >>= 1;
	}
	for (int d = 0; d < 3; d++) {
		mn[d] = x[d];
		mx[d] = x[d] + dx[d];
	}
}

static double range_max_span(const double* mn, const double* mx) {
	double s = 0.0;
	for (int d = 0; d < 3; d++) s += (mx[d] - mn[d]) * (mx[d] - mn[d]);
	return std::sqrt(s);
}

static double vabs(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

static double ewald_near_separation(const double* x) {
	double d = 0.0;
	for (int dim = 0; dim < 3; dim++) {
		const double absx = std::abs(x[dim]);
		const double this_d = std::min(absx, 1.0 - absx);
		d += this_d * this_d;
	}
	return std::sqrt(d);
}

// ---- tree build ---------------------------------------------------------------------------------
void Tree::build(const int32_t* x, const float* masses, int64_t n, const TreeParams& p, double particle_mass) {
	P = p;
	m = particle_mass;
	pos.assign(x, x + 3 * n);
	mass.resize(n);
	for (int64_t i = 0; i < n; i++) mass[i] = masses ? masses[i] : float(particle_mass);
	order.resize(n);
	for (int64_t i = 0; i < n; i++) order[i] = i;
	nodes.clear();
	nodes.reserve(size_t(4 * (n / std::max(1, p.parts_per_node)) + 16));
	Node root;
	root.pbegin = 0;
	root.pend = int32_t(n);
	root.boxid = 1;
	nodes.push_back(root);
	// breadth-first: the reference refines one level per pass; any order of splitting disjoint
	// ranges gives the same particle arrangement, so the nodes of a level split in parallel, their
	// children numbered in node order (as a sequential pass appends them)
	for (size_t lb = 0; lb < nodes.size();) {
		const size_t le = nodes.size();
		std::vector<int32_t> first(le - lb, -1);
		size_t nc = le;
		for (size_t i = lb; i < le; i++)
			if (nodes[i].pend - nodes[i].pbegin > P.parts_per_node && nodes[i].level < MAX_SPLIT_LEVEL) {
				first[i - lb] = int32_t(nc);
				nc += 2;
			}
		nodes.resize(nc);
#pragma omp parallel for schedule(dynamic, 1)
		for (int64_t i = int64_t(lb); i < int64_t(le); i++)
			if (first[size_t(i) - lb] >= 0) refine(int(i), first[size_t(i) - lb]);
		lb = le;
	}
	leaves.clear();
	max_level = 0;
	for (size_t i = 0; i < nodes.size(); i++) {
		if (nodes[i].leaf()) leaves.push_back(int32_t(i));
		max_level = std::max(max_level, nodes[i].level);
	}
}

// tree::refine (src/tree.cpp:136-172) with part_vect_sort (src/part_vect.cpp:895-933)
void Tree::refine(int ni, int li) {
	Node nd = nodes[ni];
	double bmin[3], bmax[3];
	box_id_to_range(nd.boxid, bmin, bmax);
	double max_span = 0.0;
	int max_dim = 0;
	for (int dim = 0; dim < 3; dim++) {
		const double span = bmax[dim] - bmin[dim];
		if (span > max_span) {
			max_span = span;
			max_dim = dim;
		}
	}
	const double mid = (bmax[max_dim] + bmin[max_dim]) * 0.5;
	// in-place partition: scan up from lo, on a high element search down from hi for a low one
	int64_t lo = nd.pbegin, hi = nd.pend;
	int32_t* X = pos.data();
	auto val = [&](int64_t i) { return pos_to_double(X[3 * i + max_dim]); };
	auto swap_parts = [&](int64_t a, int64_t b) {
		for (int d = 0; d < 3; d++) std::swap(X[3 * a + d], X[3 * b + d]);
		std::swap(mass[a], mass[b]);
		std::swap(order[a], order[b]);
	};
	if (lo != hi) {
		while (lo < hi) {
			if (val(lo) >= mid) {
				while (lo != hi) {
					hi--;
					if (val(hi) < mid) {
						swap_parts(lo, hi);
						break;
					}
				}
			}
			lo++;
		}
	}
	const int32_t midi = int32_t(hi);
	Node l, r;
	l.pbegin = nd.pbegin;
	l.pend = midi;
	r.pbegin = midi;
	r.pend = nd.pend;
	l.boxid = nd.boxid << 1;
	r.boxid = (nd.boxid << 1) + 1;
	l.level = r.level = nd.level + 1;
	l.parent = r.parent = ni;
	nodes[li] = l;
	nodes[li + 1] = r;
	nodes[ni].child[0] = li;
	nodes[ni].child[1] = li + 1;
}

// ---- multipoles ---------------------------------------------------------------------------------
static inline int m2s(int a, int b) {  // multipole<T>::operator()(i, j) slot (tigergrav/multipole.hpp:356-366)
	static const int map2[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
	return 1 + map2[a][b];
}
static inline int m3s(int a, int b, int c) {  // tigergrav/multipole.hpp:368-380
	static const int map3[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}}, {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
			{{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
	return 7 + map3[a][b][c];
}

// multipole<float>::operator>>= (tigergrav/multipole.hpp:414-433): shift to a new centre Y away
static void m2m_shift(float* me, const float* Y) {
	for (int p = 0; p < 3; p++)
		for (int q = p; q < 3; q++)
			for (int l = q; l < 3; l++) {
				float& t = me[m3s(p, q, l)];
				t += me[0] * Y[p] * Y[q] * Y[l];
				t += me[m2s(p, q)] * Y[l];
				t += me[m2s(q, l)] * Y[p];
				t += me[m2s(l, p)] * Y[q];
			}
	for (int p = 0; p < 3; p++)
		for (int q = p; q < 3; q++) me[m2s(p, q)] += me[0] * Y[p] * Y[q];
}

void Tree::update_moments(const int32_t* pos_tree, const float* mass_tree) {
	std::copy(pos_tree, pos_tree + pos.size(), pos.begin());
	if (mass_tree) std::copy(mass_tree, mass_tree + mass.size(), mass.begin());
	for (int ni = int(nodes.size()) - 1; ni >= 0; ni--) {
		Node& nd = nodes[ni];
		if (nd.empty()) continue;  // zero moments, as built
		if (nd.leaf()) {
			// the P2M of part_vect_multipole_info about the kept centre, 16 float lanes
			float M[17][16];
			std::memset(M, 0, sizeof(M));
			for (int64_t i = nd.pbegin; i < nd.pend; i += 16) {
				for (int k = 0; k < 16; k++) {
					// offset from the kept centre; periodic: the wrap-around difference, so a particle
					// moved across the box edge stays next to its node
					float dx[3] = {0.0f, 0.0f, 0.0f};
					float ms = 0.0f;
					if (i + k < nd.pend) {
						for (int d = 0; d < 3; d++)
							dx[d] = float(double(centre_offset_host(pos[3 * (i + k) + d], nd.x[d], P.ewald != 0)) * (1.0 / 4294967296.0));
						ms = mass[i + k];
					}
					M[0][k] += ms;
					for (int j = 0; j < 3; j++)
						for (int kk = 0; kk <= j; kk++) {
							const float Xjk = dx[j] * dx[kk];
							M[m2s(j, kk)][k] += ms * Xjk;
							for (int l = 0; l <= kk; l++) M[m3s(j, kk, l)][k] += ms * Xjk * dx[l];
						}
				}
			}
			for (int q = 0; q < 17; q++) {
				float s = 0.0f;
				for (int k = 0; k < 16; k++) s += M[q][k];
				nd.m[q] = s;
			}
			continue;
		}
		const Node& L = nodes[nd.child[0]];
		const Node& R = nodes[nd.child[1]];
		float ml[17], mr[17], yl[3], yr[3];
		std::memcpy(ml, L.m, sizeof(ml));
		std::memcpy(mr, R.m, sizeof(mr));
		for (int d = 0; d < 3; d++) {
			const double x = pos_to_double(nd.x[d]);
			yl[d] = float(pos_to_double(L.x[d]) - x);
			yr[d] = float(pos_to_double(R.x[d]) - x);
		}
		m2m_shift(ml, yl);
		m2m_shift(mr, yr);
		for (int q = 0; q < 17; q++) nd.m[q] = ml[q] + mr[q];
	}
}

void Tree::compute_multipoles() {
	// nodes are numbered level by level, children after their parents: levels from the deepest up,
	// the nodes of a level in parallel
	std::vector<size_t> level_begin;
	for (size_t i = 0; i < nodes.size(); i++)
		if (i == 0 || nodes[i].level != nodes[i - 1].level) level_begin.push_back(i);
	level_begin.push_back(nodes.size());
	for (size_t l = level_begin.size() - 1; l-- > 0;) {
#pragma omp parallel for schedule(dynamic, 64)
		for (int64_t i = int64_t(level_begin[l]); i < int64_t(level_begin[l + 1]); i++) multipoles(int(i));
	}
}

void Tree::multipoles(int ni) {
	Node& nd = nodes[ni];
	if (nd.empty()) {  // src/tree.cpp:214-227
		double bmin[3], bmax[3];
		box_id_to_range(nd.boxid, bmin, bmax);
		nd.N = 0;
		for (int q = 0; q < 17; q++) nd.m[q] = 0.0f;
		for (int d = 0; d < 3; d++) {
			nd.x[d] = double_to_pos(0.5 * (bmax[d] - bmin[d]));
			nd.rmin[d] = bmin[d];
			nd.rmax[d] = bmax[d];
		}
		nd.r = 0.0f;
		return;
	}
	if (nd.leaf()) {
		// part_vect_center_of_mass (src/part_vect.cpp:542-565)
		double com[3] = {0.0, 0.0, 0.0};
		for (int64_t i = nd.pbegin; i < nd.pend; i++)
			for (int d = 0; d < 3; d++) com[d] += pos_to_double(pos[3 * i + d]);
		nd.N = uint64_t(nd.pend - nd.pbegin);
		for (int d = 0; d < 3; d++) nd.x[d] = double_to_pos(com[d] / double(nd.N));
		// part_vect_multipole_info (src/part_vect.cpp:567-635): 16 float lanes, lane k takes
		// particles i+k of each block, summed lane 0..15 at the end
		double Xcom[3];
		for (int d = 0; d < 3; d++) Xcom[d] = pos_to_double(nd.x[d]);
		float M[17][16];
		std::memset(M, 0, sizeof(M));
		for (int64_t i = nd.pbegin; i < nd.pend; i += 16) {
			for (int k = 0; k < 16; k++) {
				double X[3];
				float ms;
				if (i + k < nd.pend) {
					for (int d = 0; d < 3; d++) X[d] = pos_to_double(pos[3 * (i + k) + d]);
					ms = mass[i + k];
				} else {
					X[0] = X[1] = X[2] = 0.0;
					ms = 0.0f;
				}
				float dx[3];
				for (int d = 0; d < 3; d++) dx[d] = float(X[d] - Xcom[d]);
				M[0][k] += ms;
				for (int j = 0; j < 3; j++)
					for (int kk = 0; kk <= j; kk++) {
						const float Xjk = dx[j] * dx[kk];
						M[m2s(j, kk)][k] += ms * Xjk;
						for (int l = 0; l <= kk; l++) M[m3s(j, kk, l)][k] += ms * Xjk * dx[l];
					}
			}
		}
		for (int q = 0; q < 17; q++) {
			float s = 0.0f;
			for (int k = 0; k < 16; k++) s += M[q][k];
			nd.m[q] = s;
		}
		nd.r = 0.0f;
		for (int64_t i = nd.pbegin; i < nd.pend; i++) {
			double v[3];
			for (int d = 0; d < 3; d++) v[d] = pos_to_double(pos[3 * i + d]) - Xcom[d];
			nd.r = std::max(nd.r, float(vabs(v)));
		}
		// part_vect_range (src/part_vect.cpp:935-968)
		for (int d = 0; d < 3; d++) {
			nd.rmax[d] = 0.0;
			nd.rmin[d] = 1.0;
		}
		for (int64_t i = nd.pbegin; i < nd.pend; i++)
			for (int d = 0; d < 3; d++) {
				const double x = pos_to_double(pos[3 * i + d]);
				nd.rmax[d] = std::max(nd.rmax[d], x);
				nd.rmin[d] = std::min(nd.rmin[d], x);
			}
		return;
	}
	// interior node (src/tree.cpp:235-288)
	const Node& L = nodes[nd.child[0]];
	const Node& R = nodes[nd.child[1]];
	nd.N = L.N + R.N;
	double lx[3], rx[3], x[3];
	for (int d = 0; d < 3; d++) {
		lx[d] = pos_to_double(L.x[d]);
		rx[d] = pos_to_double(R.x[d]);
		nd.x[d] = double_to_pos((lx[d] * double(L.N) + rx[d] * double(R.N)) / double(nd.N));
		x[d] = pos_to_double(nd.x[d]);
	}
	float ml[17], mr[17], yl[3], yr[3];
	std::memcpy(ml, L.m, sizeof(ml));
	std::memcpy(mr, R.m, sizeof(mr));
	for (int d = 0; d < 3; d++) {
		yl[d] = float(lx[d] - x[d]);
		yr[d] = float(rx[d] - x[d]);
	}
	m2m_shift(ml, yl);
	m2m_shift(mr, yr);
	for (int q = 0; q < 17; q++) nd.m[q] = ml[q] + mr[q];
	double dl[3], dr[3];
	for (int d = 0; d < 3; d++) {
		dl[d] = lx[d] - x[d];
		dr[d] = rx[d] - x[d];
	}
	if (L.N == 0)
		nd.r = R.r;
	else if (R.N == 0)
		nd.r = L.r;
	else
		nd.r = float(std::max(vabs(dl) + double(L.r), vabs(dr) + double(R.r)));
	for (int d = 0; d < 3; d++) {
		nd.rmax[d] = std::max(L.rmax[d], R.rmax[d]);
		nd.rmin[d] = std::min(L.rmin[d], R.rmin[d]);
	}
	double rmax = 0.0;
	bool first = true;
	for (int c = 0; c < 8; c++) {  // the eight corners, in the reference's order
		const double corner[3] = {(c & 1) ? nd.rmax[0] : nd.rmin[0], (c & 2) ? nd.rmax[1] : nd.rmin[1],
				(c & 4) ? nd.rmax[2] : nd.rmin[2]};
		double v[3];
		for (int d = 0; d < 3; d++) v[d] = x[d] - corner[d];
		const double a = vabs(v);
		rmax = first ? a : std::max(rmax, a);
		first = false;
	}
	nd.r = std::min(nd.r, float(rmax));
}

// ---- interaction lists: tree::kick_fmm (src/tree.cpp:378-635) -------------------------------------
// A check item is a node id; the "opened" flag (a leaf whose particles are listed directly) is the
// sign: ~id.
namespace {
struct NodeLists {
	std::vector<int32_t> cc_d, cp_d, cc_e, cp_e, pc_d, pp_d, pc_e, pp_e;
};
inline bool opened(int32_t c) { return c < 0; }
inline int32_t nid(int32_t c) { return c < 0 ? ~c : c; }
}  // namespace

struct WalkCtx {
	std::vector<NodeLists>* out;
	const std::vector<float>* rmax;  // per node: the span of its particles' bounding box
	double theta_inv;
	float r_ewald;
};

static void walk_rec(Tree& T, const WalkCtx& W, int ni, std::vector<int32_t> dlist, std::vector<int32_t> elist);

// per node, the span of its particles' bounding box (src/tree.cpp:406, part_vect_range): leaves
// from their particles, interior nodes as the union of their non-empty children's boxes (min and
// max are exact, so this equals a scan of the node's particles)
std::vector<float> Tree::sink_spans() const {
	const size_t nn = nodes.size();
	std::vector<double> mn(3 * nn, 1.0), mx(3 * nn, 0.0);
	std::vector<float> span(nn, 0.0f);
	// nodes are numbered level by level: levels from the deepest up, each level's nodes in parallel
	std::vector<size_t> level_begin;
	for (size_t i = 0; i < nn; i++)
		if (i == 0 || nodes[i].level != nodes[i - 1].level) level_begin.push_back(i);
	level_begin.push_back(nn);
	for (size_t l = level_begin.size() - 1; l-- > 0;) {
#pragma omp parallel for schedule(dynamic, 64)
		for (int64_t ki = int64_t(level_begin[l]); ki < int64_t(level_begin[l + 1]); ki++) {
			const size_t k = size_t(ki);
			const Node& nd = nodes[k];
			if (nd.empty()) continue;
			if (nd.leaf()) {
				for (int64_t i = nd.pbegin; i < nd.pend; i++)
					for (int d = 0; d < 3; d++) {
						const double x = pos_to_double(pos[3 * i + d]);
						mx[3 * k + d] = std::max(mx[3 * k + d], x);
						mn[3 * k + d] = std::min(mn[3 * k + d], x);
					}
			} else {
				for (int c = 0; c < 2; c++) {
					const size_t ch = size_t(nd.child[c]);
					if (nodes[ch].empty()) continue;
					for (int d = 0; d < 3; d++) {
						mx[3 * k + d] = std::max(mx[3 * k + d], mx[3 * ch + d]);
						mn[3 * k + d] = std::min(mn[3 * k + d], mn[3 * ch + d]);
					}
				}
			}
			span[k] = float(range_max_span(&mn[3 * k], &mx[3 * k]));
		}
	}
	return span;
}

void Tree::walk() {
	std::vector<NodeLists> per(nodes.size());
	WalkCtx W;
	W.out = &per;
	W.theta_inv = 1.0 / P.theta;
	W.r_ewald = float(2.0 * P.theta * std::pow(1.0e-3 / 8.0, 1.0 / 3.0));  // src/tree.cpp:403-404
	const std::vector<float> rmax = sink_spans();
	W.rmax = &rmax;
	std::vector<int32_t> root = {0};
#pragma omp parallel
#pragma omp single nowait
	walk_rec(*this, W, 0, root, root);
	// flatten, in node order
	Lists& Ls = lists;
	Ls = Lists();
	auto flat = [&](hvec<int64_t>& ptr, hvec<int32_t>& dst, std::vector<int32_t> NodeLists::*mem) {
		ptr.assign(nodes.size() + 1, 0);
		size_t tot = 0;
		for (size_t i = 0; i < nodes.size(); i++) tot += (per[i].*mem).size();
		dst.clear();
		dst.reserve(tot);
		for (size_t i = 0; i < nodes.size(); i++) {
			const auto& v = per[i].*mem;
			dst.insert(dst.end(), v.begin(), v.end());
			ptr[i + 1] = int64_t(dst.size());
		}
	};
	flat(Ls.cc_d_ptr, Ls.cc_d, &NodeLists::cc_d);
	flat(Ls.cp_d_ptr, Ls.cp_d, &NodeLists::cp_d);
	flat(Ls.cc_e_ptr, Ls.cc_e, &NodeLists::cc_e);
	flat(Ls.cp_e_ptr, Ls.cp_e, &NodeLists::cp_e);
	flat(Ls.pc_d_ptr, Ls.pc_d, &NodeLists::pc_d);
	flat(Ls.pp_d_ptr, Ls.pp_d, &NodeLists::pp_d);
	flat(Ls.pc_e_ptr, Ls.pc_e, &NodeLists::pc_e);
	flat(Ls.pp_e_ptr, Ls.pp_e, &NodeLists::pp_e);
	count_census();
}

// total length of the [begin, end) ranges v[b..e) (2 ints each)
static uint64_t range_total(const hvec<int32_t>& v, int64_t b, int64_t e) {
	uint64_t t = 0;
	for (int64_t k = b; k < e; k += 2) t += uint64_t(int64_t(v[size_t(k) + 1]) - int64_t(v[size_t(k)]));
	return t;
}

// census (src/tree.cpp:483-494,618-628) of the flat lists: node kinds count sources, leaf kinds
// sinks x sources
void Tree::count_census() {
	const Lists& Ls = lists;
	const int64_t nn = Ls.cc_d_ptr.empty() ? 0 : int64_t(Ls.cc_d_ptr.size()) - 1;
	uint64_t ccd = 0, cce = 0, cpd = 0, cpe = 0, pcd = 0, pce = 0, ppd = 0, ppe = 0;  // integer sums: any order
#pragma omp parallel for schedule(static) reduction(+ : ccd, cce, cpd, cpe, pcd, pce, ppd, ppe)
	for (int64_t i = 0; i < nn; i++) {
		const uint64_t ns = uint64_t(nodes[size_t(i)].pend - nodes[size_t(i)].pbegin);
		ccd += uint64_t(Ls.cc_d_ptr[size_t(i) + 1] - Ls.cc_d_ptr[size_t(i)]);
		cce += uint64_t(Ls.cc_e_ptr[size_t(i) + 1] - Ls.cc_e_ptr[size_t(i)]);
		cpd += range_total(Ls.cp_d, Ls.cp_d_ptr[size_t(i)], Ls.cp_d_ptr[size_t(i) + 1]);
		cpe += range_total(Ls.cp_e, Ls.cp_e_ptr[size_t(i)], Ls.cp_e_ptr[size_t(i) + 1]);
		pcd += ns * uint64_t(Ls.pc_d_ptr[size_t(i) + 1] - Ls.pc_d_ptr[size_t(i)]);
		pce += ns * uint64_t(Ls.pc_e_ptr[size_t(i) + 1] - Ls.pc_e_ptr[size_t(i)]);
		ppd += ns * range_total(Ls.pp_d, Ls.pp_d_ptr[size_t(i)], Ls.pp_d_ptr[size_t(i) + 1]);
		ppe += ns * range_total(Ls.pp_e, Ls.pp_e_ptr[size_t(i)], Ls.pp_e_ptr[size_t(i) + 1]);
	}
	Census C;
	C.CC_direct = ccd;
	C.CC_ewald = cce;
	C.CP_direct = cpd;
	C.CP_ewald = cpe;
	C.PC_direct = pcd;
	C.PC_ewald = pce;
	C.PP_direct = ppd;
	C.PP_ewald = ppe;
	census = C;
}

static void walk_rec(Tree& T, const WalkCtx& W, int ni, std::vector<int32_t> dlist, std::vector<int32_t> elist) {
	const Node& n = T.nodes[ni];
	if (n.empty()) return;  // src/tree.cpp:396-398
	NodeLists& out = (*W.out)[ni];
	const double h = T.P.soft_len;
	double X[3];
	for (int d = 0; d < 3; d++) X[d] = pos_to_double(n.x[d]);
	// rmax of the sink's particle range (src/tree.cpp:406)
	const float rmax = (*W.rmax)[size_t(ni)];
	auto sep = [&](const Node& c, double* dX) {
		for (int d = 0; d < 3; d++) dX[d] = X[d] - pos_to_double(c.x[d]);
	};
	auto direct_far = [&](const Node& c) {
		double dX[3];
		sep(c, dX);
		const double dx = T.P.ewald ? ewald_near_separation(dX) : vabs(dX);
		const float rsum = n.r + c.r;
		return dx > double(rsum) * W.theta_inv && dx > double(rsum) + h;
	};
	auto ewald_far = [&](const Node& c) {
		double dX[3];
		sep(c, dX);
		const double dx = std::max(0.25, ewald_near_separation(dX));
		const float rsum = n.r + c.r;
		if (dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2] > 0.0) return dx > double(rsum) * W.theta_inv && dx > double(rsum) + h;
		return rmax < W.r_ewald;
	};
	auto push_range = [](std::vector<int32_t>& v, const Node& c) {
		v.push_back(c.pbegin);
		v.push_back(c.pend);
	};
	// first pass (every node, leaves included): far items become CC / CP
	std::vector<int32_t> next_d, next_e, deferred;
	next_d.reserve(2 * dlist.size());
	for (int32_t c : dlist) {
		const Node& cn = T.nodes[nid(c)];
		if (cn.empty()) continue;
		if (direct_far(cn)) {
			if (opened(c))
				push_range(out.cp_d, cn);
			else
				out.cc_d.push_back(nid(c));
		} else {
			if (cn.leaf())
				next_d.push_back(~nid(c));
			else
				deferred.push_back(nid(c));
		}
	}
	for (int32_t c : deferred) {
		next_d.push_back(T.nodes[c].child[0]);
		next_d.push_back(T.nodes[c].child[1]);
	}
	if (T.P.ewald) {
		deferred.clear();
		next_e.reserve(2 * elist.size());
		for (int32_t c : elist) {
			const Node& cn = T.nodes[nid(c)];
			if (cn.empty()) continue;
			if (ewald_far(cn)) {
				if (opened(c))
					push_range(out.cp_e, cn);
				else
					out.cc_e.push_back(nid(c));
			} else {
				if (cn.leaf())
					next_e.push_back(~nid(c));
				else
					deferred.push_back(nid(c));
			}
		}
		for (int32_t c : deferred) {
			next_e.push_back(T.nodes[c].child[0]);
			next_e.push_back(T.nodes[c].child[1]);
		}
	}
	std::vector<int32_t>().swap(dlist);
	std::vector<int32_t>().swap(elist);
	if (!n.leaf()) {
		const int c0 = n.child[0], c1 = n.child[1];
		const bool spawn = (n.pend - n.pbegin) > 4096;
		if (spawn) {
#pragma omp task default(shared) firstprivate(c0, next_d, next_e)
			walk_rec(T, W, c0, next_d, next_e);
			walk_rec(T, W, c1, std::move(next_d), std::move(next_e));
#pragma omp taskwait
		} else {
			walk_rec(T, W, c0, next_d, next_e);
			walk_rec(T, W, c1, std::move(next_d), std::move(next_e));
		}
		return;
	}
	// leaf: drain the lists into PP / PC (src/tree.cpp:530-602)
	std::vector<int32_t> cur = std::move(next_d), nxt;
	while (!cur.empty()) {
		deferred.clear();
		nxt.clear();
		for (int32_t c : cur) {
			const Node& cn = T.nodes[nid(c)];
			if (cn.empty()) continue;
			const bool far = direct_far(cn);
			if (opened(c))
				push_range(out.pp_d, cn);
			else if (far)
				out.pc_d.push_back(nid(c));
			else if (cn.leaf())
				nxt.push_back(~nid(c));
			else
				deferred.push_back(nid(c));
		}
		for (int32_t c : deferred) {
			nxt.push_back(T.nodes[c].child[0]);
			nxt.push_back(T.nodes[c].child[1]);
		}
		std::swap(cur, nxt);
	}
	if (T.P.ewald) {
		cur = std::move(next_e);
		while (!cur.empty()) {
			deferred.clear();
			nxt.clear();
			for (int32_t c : cur) {
				const Node& cn = T.nodes[nid(c)];
				if (cn.empty()) continue;
				const bool far = ewald_far(cn);
				if (opened(c))
					push_range(out.pp_e, cn);
				else if (far)
					out.pc_e.push_back(nid(c));
				else if (cn.leaf())
					nxt.push_back(~nid(c));
				else
					deferred.push_back(nid(c));
			}
			for (int32_t c : deferred) {
				nxt.push_back(T.nodes[c].child[0]);
				nxt.push_back(T.nodes[c].child[1]);
			}
			std::swap(cur, nxt);
		}
	}
}

}  // namespace tg
