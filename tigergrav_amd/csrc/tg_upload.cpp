// tg_upload.cpp -- the host side of a solver upload: the particle and node records and the work
// the batched kernels take (leaf items, node waves, list runs, partial slots, root paths), built
// from the tree and its interaction lists.  Compiled by the host compiler with OpenMP.
#include "tg_upload.hpp"

#include <algorithm>
#include <parallel/algorithm>
#include <cmath>
#include <cstring>

#include "tg_gravity.h"
#include "tg_walk.hpp"

namespace tg {

void walk_nodes(const Tree& T, WalkNode* out) {
	const std::vector<float> span = T.sink_spans();
	const int64_t nn = int64_t(T.nodes.size());
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < nn; i++) {
		const Node& nd = T.nodes[size_t(i)];
		WalkNode& w = out[i];
		w.pbegin = nd.pbegin;
		w.pend = nd.pend;
		w.c0 = nd.child[0];
		w.c1 = nd.child[1];
		for (int d = 0; d < 3; d++) w.x[d] = nd.x[d];
		w.r = nd.r;
		w.rmax = span[size_t(i)];
	}
}

int prepare_tree_work(const Tree& T, int64_t n, HostWork& W, std::string& err) {
	const int64_t nn = int64_t(T.nodes.size());
	// particles (x, y, z, mass bits), tree order
	hvec<int4> parts(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < n; i++) {
		int w;
		std::memcpy(&w, &T.mass[i], 4);
		parts[i] = make_int4(T.pos[3 * i], T.pos[3 * i + 1], T.pos[3 * i + 2], w);
	}
	hvec<MultiSrc> multis(static_cast<size_t>(nn));
	hvec<int4> npos(static_cast<size_t>(nn));
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < nn; i++) {
		const Node& nd = T.nodes[i];
		std::memcpy(multis[i].m, nd.m, sizeof(nd.m));
		std::memcpy(multis[i].x, nd.x, sizeof(nd.x));
		npos[i] = make_int4(nd.x[0], nd.x[1], nd.x[2], 0);
	}
	// non-empty leaves in node order: (node, first particle, end, root-path offset), and the
	// root-to-leaf paths of the local-expansion pass
	hvec<int4> leaf_items;
	for (int32_t lf : T.leaves)
		if (!T.nodes[lf].empty()) leaf_items.push_back(make_int4(lf, T.nodes[lf].pbegin, T.nodes[lf].pend, 0));
	const int64_t nl = int64_t(leaf_items.size());
	hvec<int> path, level(static_cast<size_t>(nn));
	for (int64_t i = 0; i < nn; i++) level[size_t(i)] = T.nodes[size_t(i)].level;
	{
		std::vector<int64_t> poff(size_t(nl) + 1, 0);
		for (int64_t i = 0; i < nl; i++) poff[size_t(i) + 1] = poff[size_t(i)] + T.nodes[leaf_items[size_t(i)].x].level + 1;
		if (poff[size_t(nl)] > int64_t(INT32_MAX)) {
			err = "root-to-leaf paths exceed 2^31 entries";
			return TG_EINVAL;
		}
		path.resize(size_t(poff[size_t(nl)]));
#pragma omp parallel for schedule(static)
		for (int64_t i = 0; i < nl; i++) {
			leaf_items[size_t(i)].w = int(poff[size_t(i)]);
			int64_t p = poff[size_t(i) + 1];
			for (int a = leaf_items[size_t(i)].x; a >= 0; a = T.nodes[a].parent) path[size_t(--p)] = a;
		}
	}
	W.parts = std::move(parts);
	W.multis = std::move(multis);
	W.npos = std::move(npos);
	W.leaf_items = std::move(leaf_items);
	W.path = std::move(path);
	W.level = std::move(level);
	return TG_OK;
}

int prepare_work(const Tree& T, int64_t n, int ewald, double soft_len, int node_chunk, HostWork& W, std::string& err) {
	const int64_t nn = int64_t(T.nodes.size());
	{
		const int rc = prepare_tree_work(T, n, W, err);
		if (rc != TG_OK) return rc;
	}
	// node kinds: each node's list is cut into chunks of <= node_chunk entries, one partial slot per
	// chunk (a node's slots are contiguous); consecutive chunks are packed into waves of <= 64
	// entries, and a chunk longer than 64 is a wave of its own.
	// Source ranges (CP, PP) are not flattened here: each becomes a run that the device expands
	// (k_expand_runs), as do the partial-slot owners of the node kinds' entries.  Node and leaf
	// work are built in three passes: per-node (per-leaf) counts in parallel, offsets by a serial
	// prefix, then the runs, chunks and items in parallel at those offsets -- the arrays a single
	// in-order pass would produce.
	const Lists& Ls = T.lists;
	hvec<int4> items[NKINDS];
	hvec<int> lists[NKINDS];  // host-built lists (PC kinds)
	hvec<ListRun> runs[NKINDS], owner_runs[NKINDS];
	int64_t list_len[NKINDS] = {};
	hvec<int2> chunks[NKINDS];
	hvec<int> slot_node;
	hvec<int64_t> slot_ptr(static_cast<size_t>(nn + 1));
	slot_ptr[0] = 0;
	const int chunk = std::min(NODE_CHUNK_MAX, std::max(1, node_chunk));
	struct KindSrc {
		int kind;
		const hvec<int64_t>* ptr;
		const hvec<int32_t>* v;
		bool ranges;
	};
	const KindSrc nsrc[4] = {{CC_DIRECT, &Ls.cc_d_ptr, &Ls.cc_d, false}, {CP_DIRECT, &Ls.cp_d_ptr, &Ls.cp_d, true},
			{CC_EWALD, &Ls.cc_e_ptr, &Ls.cc_e, false}, {CP_EWALD, &Ls.cp_e_ptr, &Ls.cp_e, true}};
	// entries and runs of a node's (a leaf's) ranges v[b..e)
	auto range_counts = [](const hvec<int32_t>& v, int64_t b, int64_t e, int64_t& ent, int64_t& rn) {
		ent = rn = 0;
		for (int64_t k = b; k < e; k += 2) {
			const int64_t len = int64_t(v[size_t(k) + 1]) - int64_t(v[size_t(k)]);
			if (len > 0) {
				ent += len;
				rn++;
			}
		}
	};
	hvec<int64_t> nent(size_t(4 * nn)), nrun(size_t(4 * nn));
#pragma omp parallel for schedule(static)
	for (int64_t nd = 0; nd < nn; nd++)
		for (int q = 0; q < 4; q++) {
			const KindSrc& k = nsrc[q];
			int64_t ent = 0, rn = 0;
			if (k.ranges)
				range_counts(*k.v, (*k.ptr)[nd], (*k.ptr)[nd + 1], ent, rn);
			else
				ent = (*k.ptr)[nd + 1] - (*k.ptr)[nd];
			nent[size_t(4 * nd + q)] = ent;
			nrun[size_t(4 * nd + q)] = rn;
		}
	hvec<int64_t> oent(size_t(4 * nn)), orun(size_t(4 * nn)), ochk(size_t(4 * nn)), oslot(size_t(4 * nn));
	int64_t trun[4] = {}, tchk[4] = {}, slot = 0;
	{
		// each node's offsets in node order (entries, runs, chunks per kind; slots over all kinds):
		// per-block totals in parallel, their scan, then every block's running offsets from its start
		constexpr int NC = 13;  // entries[4], runs[4], chunks[4], slots
		const int64_t NB = std::min<int64_t>(nn, 1024);
		auto first = [&](int64_t b) { return nn * b / NB; };
		std::vector<int64_t> base(size_t(NB + 1) * NC, 0);
#pragma omp parallel for schedule(static)
		for (int64_t b = 0; b < NB; b++) {
			int64_t* S = base.data() + NC * (b + 1);
			for (int64_t nd = first(b); nd < first(b + 1); nd++)
				for (int q = 0; q < 4; q++) {
					const size_t i = size_t(4 * nd + q);
					const int64_t nch = (nent[i] + chunk - 1) / chunk;
					S[q] += nent[i];
					S[4 + q] += nrun[i];
					S[8 + q] += nch;
					S[12] += nch;
				}
		}
		for (int64_t b = 1; b <= NB; b++)
			for (int c = 0; c < NC; c++) base[size_t(NC * b + c)] += base[size_t(NC * (b - 1) + c)];
		int in_order = 1;
#pragma omp parallel for schedule(static) reduction(min : in_order)
		for (int64_t b = 0; b < NB; b++) {
			int64_t R[NC];
			for (int c = 0; c < NC; c++) R[c] = base[size_t(NC * b + c)];
			for (int64_t nd = first(b); nd < first(b + 1); nd++) {
				for (int q = 0; q < 4; q++) {
					const KindSrc& k = nsrc[q];
					const size_t i = size_t(4 * nd + q);
					if (!k.ranges && (*k.ptr)[nd] != R[q]) in_order = 0;
					const int64_t nch = (nent[i] + chunk - 1) / chunk;
					oent[i] = R[q];
					R[q] += nent[i];
					orun[i] = R[4 + q];
					R[4 + q] += nrun[i];
					ochk[i] = R[8 + q];
					R[8 + q] += nch;
					oslot[i] = R[12];
					R[12] += nch;
				}
				slot_ptr[nd + 1] = R[12];
			}
		}
		if (!in_order) {
			err = "node interaction list not in node order";
			return TG_EINVAL;
		}
		const int64_t* tot = base.data() + NC * NB;
		for (int q = 0; q < 4; q++) {
			list_len[nsrc[q].kind] += tot[q];
			trun[q] = tot[4 + q];
			tchk[q] = tot[8 + q];
		}
		slot = tot[12];
	}
	for (int q = 0; q < 4; q++) {
		runs[nsrc[q].kind].resize(size_t(trun[q]));
		chunks[nsrc[q].kind].resize(size_t(tchk[q]));
		owner_runs[nsrc[q].kind].resize(size_t(tchk[q]));
	}
	slot_node.resize(size_t(slot));
#pragma omp parallel for schedule(static)
	for (int64_t nd = 0; nd < nn; nd++)
		for (int q = 0; q < 4; q++) {
			const KindSrc& k = nsrc[q];
			const size_t i = size_t(4 * nd + q);
			if (k.ranges) {
				int64_t pos = oent[i], r = orun[i];
				for (int64_t j = (*k.ptr)[nd]; j < (*k.ptr)[nd + 1]; j += 2) {
					const int32_t b = (*k.v)[size_t(j)], len = (*k.v)[size_t(j) + 1] - b;
					if (len <= 0) continue;
					runs[k.kind][size_t(r++)] = ListRun{b, len, pos};
					pos += len;
				}
			}
			const int64_t lb = oent[i], le = lb + nent[i];
			int64_t c = ochk[i], sl = oslot[i];
			for (int64_t b = lb; b < le; b += chunk, c++, sl++) {
				const int64_t e = std::min<int64_t>(le, b + chunk);
				chunks[k.kind][size_t(c)] = make_int2(int(b), int(e));
				owner_runs[k.kind][size_t(c)] = ListRun{int32_t(sl), int32_t(e - b), b};
				slot_node[size_t(sl)] = int(nd);
			}
		}
	hvec<int2> waves[NKINDS];
	for (int k : {CC_DIRECT, CP_DIRECT, CC_EWALD, CP_EWALD}) {
		// greedy, restarted every WAVE_SEGMENT chunks (so the device can pack segments in parallel)
		int2 cur = make_int2(0, 0);
		bool open = false;
		for (size_t q = 0; q < chunks[k].size(); q++) {
			const int2& c = chunks[k][q];
			if (open && (q % WAVE_SEGMENT == 0 || c.y - cur.x > 64)) {
				waves[k].push_back(cur);
				open = false;
			}
			if (!open) {
				cur = c;
				open = true;
			} else {
				cur.y = c.y;
			}
		}
		if (open) waves[k].push_back(cur);
	}
	W.n_slots = slot;
	// leaf kinds: items (first sink, count <= 64, lb, le).  A PP-direct list is reordered so the
	// source ranges whose bounding box comes within the softening length of the sink leaf's box
	// (the leaf itself among them) come first; the item records that near prefix's length.
	hvec<int32_t> leaf_at(size_t(n) + 1);
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i <= n; i++) leaf_at[size_t(i)] = -1;
	for (int32_t lf : T.leaves)
		if (!T.nodes[lf].empty()) leaf_at[T.nodes[lf].pbegin] = lf;
	const double hnear = soft_len * 1.001 + 1e-9;  // margin over the kernel's single-precision test
	auto range_box = [&](int32_t b, int32_t e, double* mn, double* mx) {
		const int32_t lf = leaf_at[b];
		if (lf >= 0 && T.nodes[lf].pend == e) {
			for (int d = 0; d < 3; d++) {
				mn[d] = T.nodes[lf].rmin[d];
				mx[d] = T.nodes[lf].rmax[d];
			}
			return;
		}
		for (int d = 0; d < 3; d++) {
			mn[d] = 1.0;
			mx[d] = 0.0;
		}
		for (int32_t i = b; i < e; i++)
			for (int d = 0; d < 3; d++) {
				const double x = pos_to_double(T.pos[3 * size_t(i) + d]);
				mn[d] = std::min(mn[d], x);
				mx[d] = std::max(mx[d], x);
			}
	};
	auto may_be_near = [&](const Node& sk, const double* mn, const double* mx) {
		double d2 = 0.0;
		for (int d = 0; d < 3; d++) {
			const double h1 = 0.5 * (sk.rmax[d] - sk.rmin[d]), h2 = 0.5 * (mx[d] - mn[d]);
			if (ewald && h1 + h2 > 0.25) return true;  // minimum image of the centres is not decisive
			double dc = 0.5 * (sk.rmax[d] + sk.rmin[d]) - 0.5 * (mx[d] + mn[d]);
			if (ewald) dc -= std::nearbyint(dc);
			const double gap = std::fabs(dc) - h1 - h2;
			if (gap > 0.0) d2 += gap * gap;
		}
		return d2 <= hnear * hnear;
	};
	std::vector<int32_t> lvs;  // non-empty leaves, in the tree's leaf order
	for (int32_t lf : T.leaves)
		if (!T.nodes[lf].empty()) lvs.push_back(lf);
	const int64_t nl = int64_t(lvs.size());
	const KindSrc lsrc[4] = {{PC_DIRECT, &Ls.pc_d_ptr, &Ls.pc_d, false}, {PP_DIRECT, &Ls.pp_d_ptr, &Ls.pp_d, true},
			{PC_EWALD, &Ls.pc_e_ptr, &Ls.pc_e, false}, {PP_EWALD, &Ls.pp_e_ptr, &Ls.pp_e, true}};
	// per leaf and kind: entries, runs, near-prefix length (PP direct) or image count (PP Ewald), items
	hvec<int64_t> lent(size_t(4 * nl)), lrun(size_t(4 * nl)), lnear(size_t(4 * nl)), litem(size_t(4 * nl));
	std::vector<uint8_t> pp_near(Ls.pp_d.size() / 2, 0);
#pragma omp parallel for schedule(dynamic, 64)
	for (int64_t i = 0; i < nl; i++) {
		const int32_t lf = lvs[size_t(i)];
		const Node& nd = T.nodes[lf];
		const int64_t per_leaf = (int64_t(nd.pend - nd.pbegin) + 63) / 64;
		for (int q = 0; q < 4; q++) {
			const KindSrc& k = lsrc[q];
			const size_t o = size_t(4 * i + q);
			int64_t ent = 0, rn = 0, nearl = 0;
			if (size_t(lf) + 1 < k.ptr->size()) {
				const int64_t kb = (*k.ptr)[lf], ke = (*k.ptr)[lf + 1];
				if (k.kind == PP_DIRECT) {
					for (int64_t j = kb; j < ke; j += 2) {
						const int32_t b = (*k.v)[size_t(j)], e = (*k.v)[size_t(j) + 1];
						double mn[3], mx[3];
						range_box(b, e, mn, mx);
						const bool near = may_be_near(nd, mn, mx);
						pp_near[size_t(j / 2)] = near ? 1 : 0;
						if (e > b) {
							ent += e - b;
							rn++;
							if (near) nearl += e - b;
						}
					}
				} else if (k.kind == PP_EWALD) {
					range_counts(*k.v, kb, ke, ent, rn);
				} else {
					ent = ke - kb;
				}
			}
			lent[o] = ent;
			lrun[o] = rn;
			lnear[o] = nearl;
			litem[o] = ent > 0 ? per_leaf : 0;
		}
	}
	hvec<int64_t> olent(size_t(4 * nl)), olrun(size_t(4 * nl)), olitem(size_t(4 * nl));
	{
		int64_t lrt[4] = {}, lit[4] = {};
		for (int64_t i = 0; i < nl; i++)
			for (int q = 0; q < 4; q++) {
				const size_t o = size_t(4 * i + q);
				olent[o] = list_len[lsrc[q].kind];
				list_len[lsrc[q].kind] += lent[o];
				olrun[o] = lrt[q];
				lrt[q] += lrun[o];
				olitem[o] = lit[q];
				lit[q] += litem[o];
			}
		for (int q = 0; q < 4; q++) {
			const int kind = lsrc[q].kind;
			if (lsrc[q].ranges)
				runs[kind].resize(size_t(lrt[q]));
			else
				lists[kind].resize(size_t(list_len[kind]));
			items[kind].resize(size_t(lit[q]));
		}
	}
#pragma omp parallel for schedule(dynamic, 64)
	for (int64_t i = 0; i < nl; i++) {
		const int32_t lf = lvs[size_t(i)];
		const Node& nd = T.nodes[lf];
		for (int q = 0; q < 4; q++) {
			const KindSrc& k = lsrc[q];
			const size_t o = size_t(4 * i + q);
			const int64_t lb = olent[o], le = lb + lent[o];
			if (le == lb) continue;
			const int64_t kb = (*k.ptr)[lf], ke = (*k.ptr)[lf + 1];
			if (k.ranges) {
				int64_t r = olrun[o], pos = lb;
				// PP direct: the near ranges first, then the far ones; PP Ewald: in list order
				for (int pass = 0; pass < (k.kind == PP_DIRECT ? 2 : 1); pass++)
					for (int64_t j = kb; j < ke; j += 2) {
						const int32_t b = (*k.v)[size_t(j)], len = (*k.v)[size_t(j) + 1] - b;
						if (len <= 0) continue;
						if (k.kind == PP_DIRECT && (pp_near[size_t(j / 2)] != 0) != (pass == 0)) continue;
						runs[k.kind][size_t(r++)] = ListRun{b, len, pos};
						pos += len;
					}
			} else {
				std::copy(k.v->begin() + kb, k.v->begin() + ke, lists[k.kind].begin() + lb);
			}
			int64_t it = olitem[o];
			for (int32_t b = nd.pbegin; b < nd.pend; b += 64) {
				const int cnt = std::min<int32_t>(64, nd.pend - b);
				const int y = (k.kind == PP_DIRECT || k.kind == PP_EWALD) ? pp_item_y(cnt, lnear[o]) : cnt;
				items[k.kind][size_t(it++)] = make_int4(b, y, int(lb), int(le));
			}
		}
	}
	// heaviest cost class first, leaf order within a class (item_class): the dispatcher hands out
	// work groups in order (a stable sort: the result is the unique stable order, however the
	// threads merge)
	const int leaf_kinds[4] = {PC_DIRECT, PP_DIRECT, PC_EWALD, PP_EWALD};
	for (int q = 0; q < 4; q++) {
		const int k = leaf_kinds[q];
		__gnu_parallel::stable_sort(items[k].begin(), items[k].end(), [](const int4& a, const int4& b) {
			const int na = a.y & PP_COUNT_MASK, nb = b.y & PP_COUNT_MASK;
			return item_class(a.w - a.z, na) > item_class(b.w - b.z, nb);
		});
	}
	
	for (int k = 0; k < NKINDS; k++)
		if (list_len[k] > int64_t(INT32_MAX)) {
			err = "an interaction list exceeds 2^31 entries (32-bit list offsets)";
			return TG_EINVAL;
		}
	for (int k = 0; k < NKINDS; k++) {
		W.items[k] = std::move(items[k]);
		W.lists[k] = std::move(lists[k]);
		W.runs[k] = std::move(runs[k]);
		W.owner_runs[k] = std::move(owner_runs[k]);
		W.waves[k] = std::move(waves[k]);
		W.list_len[k] = list_len[k];
	}
	W.slot_node = std::move(slot_node);
	W.slot_ptr = std::move(slot_ptr);
	return TG_OK;
}

}  // namespace tg
