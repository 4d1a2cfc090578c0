// tg_prep.hip -- a solver's list work built on the device from the device walk's lists, which
// never leave the device: the arrays prepare_work (tg_upload.cpp) builds on the host from a host
// copy of the lists, element for element.
//
// Node kinds (CC, CP; direct and Ewald), one thread per node: the entries of each kind, cut into
// chunks of <= node_chunk entries (one partial slot each; a node's slots are contiguous, its kinds
// in the order CC_d, CP_d, CC_e, CP_e).  Exclusive scans over the nodes give every offset; then
// the chunks, the slot of every entry, the node of every slot and the CP ranges expanded to
// particle indices are written, and the chunks packed into waves of <= 64 entries (a longer chunk
// is a wave alone): greedily, restarting every WAVE_SEGMENT chunks, so the segments pack in
// parallel.  The CC lists are the
// walk's own (its node-order CSR is the list order).
// Leaf kinds (PC, PP), one wave per non-empty leaf, leaves in node order: the entries, and for PP
// direct the near prefix -- the sources of ranges whose bounding box comes within the softening
// length of the sink leaf's box, the test made in FP64 with the host's operations.  Scans over
// the leaves give the offsets; the PP ranges are expanded (PP direct: near ranges first, both
// groups in list order), the items of <= 64 sinks written, and every kind's items put in the
// order of a stable sort by descending cost class (a radix sort of (class, index) pairs).  The PC lists
// are the walk's own.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "tg_gravity.h"
#include "tg_solver.hpp"
#include "tg_devmem.hpp"

namespace tg {
namespace {

// walk kinds 0..3 are the node kinds, 4..7 the leaf kinds, in these solver kinds
constexpr int NODE_KIND[4] = {CC_DIRECT, CP_DIRECT, CC_EWALD, CP_EWALD};
constexpr int LEAF_KIND[4] = {PC_DIRECT, PP_DIRECT, PC_EWALD, PP_EWALD};

__device__ __forceinline__ double pos_d(int32_t x) {
#pragma clang fp contract(off)
	return double(x) * (1.0 / 4294967296.0) + 0.5;
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

__device__ __forceinline__ int wave_min(int v) {
	for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
	return v;
}

__device__ __forceinline__ int wave_max(int v) {
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
	return v;
}

// ---- node kinds ----------------------------------------------------------------------------------

struct NodeIn {
	const int64_t* ptr[4];  // walk kinds 0..3
	const int32_t* val[4];
	int64_t nn;
	int chunk;
};

// per node and kind: entries and chunks; per node: slots.  Arrays of nn + 1 (the last zero, so an
// exclusive scan over nn + 1 ends with the total)
__global__ __launch_bounds__(256) void k_node_count(NodeIn A, int64_t* __restrict__ nent, int64_t* __restrict__ nchk,
		int64_t* __restrict__ nslot) {
	const int64_t nd = int64_t(blockIdx.x) * 256 + threadIdx.x;
	const int64_t np = A.nn + 1;
	if (nd > A.nn) return;
	if (nd == A.nn) {
		for (int q = 0; q < 4; q++) nent[q * np + nd] = nchk[q * np + nd] = 0;
		nslot[nd] = 0;
		return;
	}
	int64_t slots = 0;
	for (int q = 0; q < 4; q++) {
		const int64_t kb = A.ptr[q][nd], ke = A.ptr[q][nd + 1];
		int64_t ent = 0;
		if (q & 1) {  // CP: [begin, end) ranges
			for (int64_t j = kb; j < ke; j += 2) {
				const int64_t len = int64_t(A.val[q][j + 1]) - int64_t(A.val[q][j]);
				if (len > 0) ent += len;
			}
		} else {
			ent = ke - kb;
		}
		const int64_t nch = (ent + A.chunk - 1) / A.chunk;
		nent[q * np + nd] = ent;
		nchk[q * np + nd] = nch;
		slots += nch;
	}
	nslot[nd] = slots;
}

struct NodeOut {
	int2* chunks[4];       // per kind: [begin, end) entries of each chunk
	int* chunk_slot[4];    // per kind: the partial slot of each chunk
	int* slot_node;
};

// the chunks of every node: records, slots and the slots' node
__global__ __launch_bounds__(256) void k_node_chunks(NodeIn A, const int64_t* __restrict__ nent, const int64_t* __restrict__ oent,
		const int64_t* __restrict__ ochk, const int64_t* __restrict__ slot_ptr, NodeOut O) {
	const int64_t nd = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (nd >= A.nn) return;
	const int64_t np = A.nn + 1;
	int64_t sl = slot_ptr[nd];
	for (int q = 0; q < 4; q++) {
		const int64_t lb = oent[q * np + nd], le = lb + nent[q * np + nd];
		int64_t c = ochk[q * np + nd];
		for (int64_t b = lb; b < le; b += A.chunk, c++, sl++) {
			const int64_t e = le < b + A.chunk ? le : b + A.chunk;
			O.chunks[q][c] = make_int2(int(b), int(e));
			O.chunk_slot[q][c] = int(sl);
			O.slot_node[sl] = int(nd);
		}
	}
}

// the partial slot of every entry: one wave per chunk
__global__ __launch_bounds__(256) void k_node_owner(const int2* __restrict__ chunks, const int* __restrict__ chunk_slot, int64_t nch,
		int* __restrict__ owner) {
	const int64_t c = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
	if (c >= nch) return;
	const int2 ch = chunks[c];
	const int sl = chunk_slot[c];
	for (int j = ch.x + int(threadIdx.x & 63); j < ch.y; j += 64) owner[j] = sl;
}

// greedy packing of chunks [c0, c1) into waves of <= 64 entries: calls emit(begin, end) per wave
template <class F>
__device__ __forceinline__ void pack_waves(const int2* __restrict__ chunks, int64_t c0, int64_t c1, F emit) {
	int2 cur = make_int2(0, 0);
	bool open = false;
	for (int64_t c = c0; c < c1; c++) {
		const int2 ch = chunks[c];
		if (open && ch.y - cur.x > 64) {
			emit(cur);
			open = false;
		}
		if (!open) {
			cur = ch;
			open = true;
		} else {
			cur.y = ch.y;
		}
	}
	if (open) emit(cur);
}

// waves per segment of WAVE_SEGMENT chunks (array of nseg + 1, the last zero)
__global__ __launch_bounds__(256) void k_wave_count(const int2* __restrict__ chunks, int64_t nch, int64_t nseg, int64_t* __restrict__ nw) {
	const int64_t sg = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (sg > nseg) return;
	if (sg == nseg) {
		nw[sg] = 0;
		return;
	}
	const int64_t c0 = sg * WAVE_SEGMENT, c1 = c0 + WAVE_SEGMENT < nch ? c0 + WAVE_SEGMENT : nch;
	int64_t k = 0;
	pack_waves(chunks, c0, c1, [&](int2) { k++; });
	nw[sg] = k;
}

__global__ __launch_bounds__(256) void k_wave_write(const int2* __restrict__ chunks, int64_t nch, int64_t nseg, const int64_t* __restrict__ ow,
		int2* __restrict__ waves) {
	const int64_t sg = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (sg >= nseg) return;
	const int64_t c0 = sg * WAVE_SEGMENT, c1 = c0 + WAVE_SEGMENT < nch ? c0 + WAVE_SEGMENT : nch;
	int64_t k = ow[sg];
	pack_waves(chunks, c0, c1, [&](int2 w) { waves[k++] = w; });
}

// ---- range expansion (CP, PP) --------------------------------------------------------------------

// One wave writes the entries of the non-empty ranges of v[kb, ke) (2 ints each) that keep()
// accepts, consecutively from out: 64 ranges per round, an inclusive scan of their lengths, then
// every lane writes entries base + lane + 64 t, finding its range by binary search of the scan.
// Returns the entries written.  The block is this one wave (the barriers are the wave's).
template <class Keep>
__device__ int64_t expand_ranges(const int32_t* __restrict__ v, int64_t kb, int64_t ke, Keep keep, int32_t* __restrict__ out) {
	__shared__ int incl[64], first[64];
	const int lane = threadIdx.x;
	int64_t pos = 0;
	for (int64_t j0 = kb; j0 < ke; j0 += 128) {
		const int64_t j = j0 + 2 * lane;
		int b = 0, len = 0;
		if (j < ke) {
			b = v[j];
			const int e = v[j + 1];
			len = e > b ? e - b : 0;
			if (len > 0 && !keep(b, e, j)) len = 0;
		}
		int s = len;
		for (int o = 1; o < 64; o <<= 1) {
			const int t = __shfl_up(s, o, 64);
			if (lane >= o) s += t;
		}
		const int tot = __shfl(s, 63, 64);
		incl[lane] = s;
		first[lane] = b;
		__syncthreads();
		for (int p = lane; p < tot; p += 64) {
			int lo = 0, hi = 63;  // the first range whose inclusive end passes p
			while (lo < hi) {
				const int mid = (lo + hi) >> 1;
				if (incl[mid] > p)
					hi = mid;
				else
					lo = mid + 1;
			}
			const int before = lo ? incl[lo - 1] : 0;
			out[pos + p] = first[lo] + (p - before);
		}
		__syncthreads();
		pos += tot;
	}
	return pos;
}

// CP ranges of every node expanded: one wave per node
__global__ __launch_bounds__(64) void k_node_expand(NodeIn A, const int64_t* __restrict__ oent, int32_t* cp_d, int32_t* cp_e) {
	const int64_t nd = blockIdx.x;
	const int64_t np = A.nn + 1;
	for (int q = 1; q < 4; q += 2) {
		const int64_t kb = A.ptr[q][nd], ke = A.ptr[q][nd + 1];
		if (ke == kb) continue;
		int32_t* out = (q == 1 ? cp_d : cp_e) + oent[q * np + nd];
		expand_ranges(A.val[q], kb, ke, [](int, int, int64_t) { return true; }, out);
	}
}

// ---- leaf kinds ----------------------------------------------------------------------------------

struct LeafIn {
	const int64_t* ptr[4];  // walk kinds 4..7: pc_d, pp_d, pc_e, pp_e
	const int32_t* val[4];
	const int4* leaf_items;  // (node, first particle, end, -)
	const int4* parts;
	const int* pleaf;        // particle -> leaf
	const WalkNode* nodes;
	const int4* box;         // per node, 2 records: lowest and highest particle coordinates (leaves)
	uint8_t* near;           // PP direct: per range (entry pair j / 2), the near test's outcome
	int64_t nl;
	int ewald;
	double hnear;
};

// the bounding box of every non-empty leaf's particles: one wave per leaf
__global__ __launch_bounds__(64) void k_leaf_boxes(const int4* __restrict__ leaf_items, const int4* __restrict__ parts, int4* __restrict__ box) {
	const int4 li = leaf_items[blockIdx.x];
	int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
	for (int p = li.y + int(threadIdx.x); p < li.z; p += 64) {
		const int4 x = parts[p];
		lo[0] = min(lo[0], x.x);
		lo[1] = min(lo[1], x.y);
		lo[2] = min(lo[2], x.z);
		hi[0] = max(hi[0], x.x);
		hi[1] = max(hi[1], x.y);
		hi[2] = max(hi[2], x.z);
	}
	for (int d = 0; d < 3; d++) {
		lo[d] = wave_min(lo[d]);
		hi[d] = wave_max(hi[d]);
	}
	if (threadIdx.x == 0) {
		box[2 * li.x] = make_int4(lo[0], lo[1], lo[2], 0);
		box[2 * li.x + 1] = make_int4(hi[0], hi[1], hi[2], 0);
	}
}

// may a source of the box [lo, hi] lie within the softening length of the sink box [slo, shi]?
// The host's test (prepare_work, tg_upload.cpp) with the same FP64 operations.
__device__ bool may_be_near(int4 slo, int4 shi, int4 lo, int4 hi, int ewald, double hnear) {
#pragma clang fp contract(off)
	const int a0[3] = {slo.x, slo.y, slo.z}, a1[3] = {shi.x, shi.y, shi.z};
	const int b0[3] = {lo.x, lo.y, lo.z}, b1[3] = {hi.x, hi.y, hi.z};
	double d2 = 0.0;
	for (int d = 0; d < 3; d++) {
		const double smn = pos_d(a0[d]), smx = pos_d(a1[d]), mn = pos_d(b0[d]), mx = pos_d(b1[d]);
		const double h1 = 0.5 * (smx - smn), h2 = 0.5 * (mx - mn);
		if (ewald && h1 + h2 > 0.25) return true;
		double dc = 0.5 * (smx + smn) - 0.5 * (mx + mn);
		if (ewald) dc -= rint(dc);
		const double gap = fabs(dc) - h1 - h2;
		if (gap > 0.0) d2 += gap * gap;
	}
	return d2 <= hnear * hnear;
}

// the near test of the source range [b, e) (non-empty) for a sink leaf of box [slo, shi]: a whole
// leaf takes that leaf's box, any other range its particles' box
__device__ bool range_near(const LeafIn& A, int4 slo, int4 shi, int b, int e) {
	const int lf = A.pleaf[b];
	const WalkNode& nd = A.nodes[lf];
	int4 lo, hi;
	if (nd.pbegin == b && nd.pend == e) {
		lo = A.box[2 * lf];
		hi = A.box[2 * lf + 1];
	} else {
		lo = make_int4(INT32_MAX, INT32_MAX, INT32_MAX, 0);
		hi = make_int4(INT32_MIN, INT32_MIN, INT32_MIN, 0);
		for (int p = b; p < e; p++) {
			const int4 x = A.parts[p];
			lo = make_int4(min(lo.x, x.x), min(lo.y, x.y), min(lo.z, x.z), 0);
			hi = make_int4(max(hi.x, x.x), max(hi.y, x.y), max(hi.z, x.z), 0);
		}
	}
	return may_be_near(slo, shi, lo, hi, A.ewald, A.hnear);
}

// per leaf and kind: entries (list ints for PC, expanded sources for PP), near-prefix length (PP
// direct) and items; per leaf the census terms (sinks x entries, sums[q]) and the raw list
// lengths (sums[4 + q], to check that the leaves hold every entry), summed afterwards by
// reductions (one atomic per leaf on a shared total would serialise at one cache line)
__global__ __launch_bounds__(64) void k_leaf_count(LeafIn A, int64_t* __restrict__ lent, int64_t* __restrict__ lnear, int64_t* __restrict__ litem,
		int64_t* __restrict__ sums) {
	const int64_t i = blockIdx.x;
	const int lane = threadIdx.x;
	const int64_t np = A.nl + 1;
	if (i == A.nl) {
		if (lane < 4) lent[lane * np + i] = litem[lane * np + i] = 0;
		if (lane < 8) sums[lane * np + i] = 0;
		if (lane == 0) lnear[i] = 0;
		return;
	}
	const int4 li = A.leaf_items[i];
	const int lf = li.x;
	const int64_t ns = li.z - li.y;
	const int4 slo = A.box[2 * lf], shi = A.box[2 * lf + 1];
	for (int q = 0; q < 4; q++) {
		const int64_t kb = A.ptr[q][lf], ke = A.ptr[q][lf + 1];
		int64_t ent = 0, nearl = 0;
		if (q & 1) {
			for (int64_t j = kb + 2 * lane; j < ke; j += 128) {
				const int b = A.val[q][j], e = A.val[q][j + 1];
				if (e <= b) continue;
				ent += e - b;
				if (q == 1) {
					const bool nr = range_near(A, slo, shi, b, e);
					A.near[j >> 1] = nr ? 1 : 0;  // for the write pass
					if (nr) nearl += e - b;
				}
			}
			ent = wave_sum(ent);
			nearl = wave_sum(nearl);
		} else {
			ent = ke - kb;
		}
		if (lane == 0) {
			lent[q * np + i] = ent;
			litem[q * np + i] = ent > 0 ? (ns + 63) / 64 : 0;
			if (q == 1) lnear[i] = nearl;
			sums[q * np + i] = ns * ent;
			sums[(4 + q) * np + i] = ke - kb;
		}
	}
}

struct LeafOut {
	int32_t* list[4];  // PP kinds: expanded lists (PC: unused)
	int4* items[4];    // unsorted items
	unsigned long long* key[4];
	int* idx[4];
};


// the expanded PP lists and every kind's items, in leaf order: one wave per leaf
__global__ __launch_bounds__(64) void k_leaf_write(LeafIn A, const int64_t* __restrict__ lent, const int64_t* __restrict__ olent,
		const int64_t* __restrict__ lnear, const int64_t* __restrict__ olitem, LeafOut O) {
	const int64_t i = blockIdx.x;
	const int lane = threadIdx.x;
	const int64_t np = A.nl + 1;
	const int4 li = A.leaf_items[i];
	const int lf = li.x;
	const int4 slo = A.box[2 * lf], shi = A.box[2 * lf + 1];
	for (int q = 0; q < 4; q++) {
		const int64_t ent = lent[q * np + i];
		if (ent == 0) continue;
		const int64_t kb = A.ptr[q][lf], ke = A.ptr[q][lf + 1];
		int64_t lb, le, nearl = 0;
		if (q & 1) {
			lb = olent[q * np + i];
			le = lb + ent;
			int32_t* out = O.list[q] + lb;
			if (q == 1) {
				nearl = lnear[i];
				const uint8_t* nr = A.near;
				const int64_t w = expand_ranges(A.val[q], kb, ke, [&](int, int, int64_t j) { return nr[j >> 1] != 0; }, out);
				expand_ranges(A.val[q], kb, ke, [&](int, int, int64_t j) { return nr[j >> 1] == 0; }, out + w);
			} else {
				expand_ranges(A.val[q], kb, ke, [](int, int, int64_t) { return true; }, out);
			}
		} else {
			lb = kb;
			le = ke;
		}
		const int64_t it0 = olitem[q * np + i];
		const int nblk = (li.z - li.y + 63) / 64;
		for (int k = lane; k < nblk; k += 64) {
			const int b = li.y + 64 * k;
			const int cnt = li.z - b < 64 ? li.z - b : 64;
			const int y = (q & 1) ? pp_item_y(cnt, nearl) : cnt;
			const int64_t it = it0 + k;
			O.items[q][it] = make_int4(b, y, int(lb), int(le));
			O.key[q][it] = (unsigned long long) item_class(le - lb, cnt);
			O.idx[q][it] = int(it);
		}
	}
}

__global__ __launch_bounds__(256) void k_gather_items(const int4* __restrict__ in, const int* __restrict__ idx, int64_t m, int4* __restrict__ out) {
	const int64_t r = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (r < m) out[r] = in[idx[r]];
}

__global__ void k_zero_u64(unsigned long long* p, int m) {
	if (int(threadIdx.x) < m) p[threadIdx.x] = 0;
}

// the last element of each array, gathered for one copy to the host
__global__ void k_last(const int64_t* const* __restrict__ arrs, const int64_t* __restrict__ at, int m, int64_t* __restrict__ out) {
	const int k = threadIdx.x;
	if (k < m) out[k] = arrs[k][at[k]];
}

unsigned blocks(int64_t n, int per) { return unsigned((n + per - 1) / per); }

// ---- the tree's arrays from a device-built tree ---------------------------------------------------

// the walk's node table: one level of nodes, deepest levels first.  sbox: per node the box of its
// particles (leaves) or the union of its non-empty children's (Tree::sink_spans); span its
// diagonal, in FP64 with the host's operations
__global__ __launch_bounds__(256) void k_tw_walk_nodes(const DNode* __restrict__ nodes, const DMulti* __restrict__ multi, int lb, int nl,
		double* __restrict__ sbox, WalkNode* __restrict__ wn) {
#pragma clang fp contract(off)
	const int j = blockIdx.x * 256 + threadIdx.x;
	if (j >= nl) return;
	const int k = lb + j;
	const DNode nd = nodes[k];
	const DMulti& m = multi[k];
	double mn[3] = {1.0, 1.0, 1.0}, mx[3] = {0.0, 0.0, 0.0};
	const bool empty = nd.pend == nd.pbegin;
	if (!empty) {
		if (nd.c0 < 0) {
			for (int d = 0; d < 3; d++) {
				mn[d] = m.rmin[d];
				mx[d] = m.rmax[d];
			}
		} else {
			const int ch[2] = {nd.c0, nd.c1};
			for (int c = 0; c < 2; c++) {
				if (nodes[ch[c]].pend == nodes[ch[c]].pbegin) continue;
				const double* b = sbox + 6 * int64_t(ch[c]);
				for (int d = 0; d < 3; d++) {
					mx[d] = mx[d] < b[3 + d] ? b[3 + d] : mx[d];
					mn[d] = b[d] < mn[d] ? b[d] : mn[d];
				}
			}
		}
	}
	double* o = sbox + 6 * int64_t(k);
	for (int d = 0; d < 3; d++) {
		o[d] = mn[d];
		o[3 + d] = mx[d];
	}
	float span = 0.0f;
	if (!empty) {
		double s2 = 0.0;
		for (int d = 0; d < 3; d++) s2 += (mx[d] - mn[d]) * (mx[d] - mn[d]);
		span = float(sqrt(s2));
	}
	WalkNode w;
	w.pbegin = nd.pbegin;
	w.pend = nd.pend;
	w.c0 = nd.c0;
	w.c1 = nd.c1;
	for (int d = 0; d < 3; d++) w.x[d] = m.x[d];
	w.r = m.r;
	w.rmax = span;
	wn[k] = w;
}

// particle records (x, y, z, mass bits)
__global__ __launch_bounds__(256) void k_tw_parts(const int32_t* __restrict__ pos, const float* __restrict__ mass, int64_t n, int4* __restrict__ parts) {
	const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (i < n) parts[i] = make_int4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], __float_as_int(mass[i]));
}

// per node: source multipole, centre, level, parent; flags of the non-empty leaves and their
// root-path lengths (arrays of nn + 1, the last zero); the leaf count (empty leaves included)
__global__ __launch_bounds__(256) void k_tw_nodes(const DNode* __restrict__ nodes, const DMulti* __restrict__ multi, int64_t nn,
		MultiSrc* __restrict__ multis, int4* __restrict__ npos, int* __restrict__ level, int* __restrict__ parent, int64_t* __restrict__ lflag,
		int64_t* __restrict__ plen, unsigned long long* __restrict__ n_leaves) {
	const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
	// the leaf count: one atomic per wave
	const bool is_leaf = i < nn && nodes[i].c0 < 0;
	const unsigned long long lv = __ballot(is_leaf);
	if ((threadIdx.x & 63) == 0 && lv) atomicAdd(n_leaves, (unsigned long long) __popcll(lv));
	if (i > nn) return;
	if (i == nn) {
		lflag[i] = plen[i] = 0;
		return;
	}
	const DNode nd = nodes[i];
	const DMulti& m = multi[i];
	MultiSrc ms;
	for (int q = 0; q < 17; q++) ms.m[q] = m.m[q];
	for (int d = 0; d < 3; d++) ms.x[d] = m.x[d];
	multis[i] = ms;
	npos[i] = make_int4(m.x[0], m.x[1], m.x[2], 0);
	level[i] = nd.level;
	parent[i] = nd.parent;
	const bool leaf = nd.c0 < 0;
	const bool use = leaf && nd.pend > nd.pbegin;
	lflag[i] = use ? 1 : 0;
	plen[i] = use ? nd.level + 1 : 0;
}

// the non-empty leaves in node order: (node, first particle, end, root-path offset) and the paths
__global__ __launch_bounds__(256) void k_tw_leaves(const DNode* __restrict__ nodes, int64_t nn, const int64_t* __restrict__ lflag,
		const int64_t* __restrict__ loff, const int64_t* __restrict__ poff, int4* __restrict__ leaf_items, int* __restrict__ path) {
	const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (i >= nn || !lflag[i]) return;
	const DNode nd = nodes[i];
	leaf_items[loff[i]] = make_int4(int(i), nd.pbegin, nd.pend, int(poff[i]));
	int64_t p = poff[i] + nd.level + 1;
	for (int a = int(i); a >= 0; a = nodes[a].parent) path[--p] = a;
}

// each particle's leaf: one wave per leaf
__global__ __launch_bounds__(64) void k_tw_pleaf(const int4* __restrict__ leaf_items, int* __restrict__ pleaf) {
	const int4 li = leaf_items[blockIdx.x];
	for (int p = li.y + int(threadIdx.x); p < li.z; p += 64) pleaf[p] = li.x;
}

}  // namespace

hipError_t walk_nodes_device(const DeviceTree& D, WalkNode** out, hipStream_t s, std::string& err) {
	*out = nullptr;
	const int64_t nn = D.nn;
	WalkNode* wn = nullptr;
	double* sbox = nullptr;
	hipError_t e = dev_malloc(&wn, size_t(std::max<int64_t>(nn, 1)) * sizeof(WalkNode));
	if (e == hipSuccess) e = dev_malloc(&sbox, size_t(std::max<int64_t>(nn, 1)) * 6 * sizeof(double));
	for (size_t l = D.level_begin.size() - 1; e == hipSuccess && l-- > 0;) {
		const int b = D.level_begin[l], nl = D.level_begin[l + 1] - b;
		if (nl <= 0) continue;
		hipLaunchKernelGGL(k_tw_walk_nodes, dim3(blocks(nl, 256)), dim3(256), 0, s, D.nodes, D.multi, b, nl, sbox, wn);
		e = hipGetLastError();
	}
	if (e == hipSuccess) e = hipStreamSynchronize(s);
	if (sbox) dev_free(sbox);
	if (e != hipSuccess) {
		if (wn) dev_free(wn);
		err = std::string("walk node table: ") + hipGetErrorString(e);
		return e;
	}
	*out = wn;
	return hipSuccess;
}

#define PCHK(call)                                                   \
	do {                                                             \
		const hipError_t e_ = (call);                                \
		if (e_ != hipSuccess) {                                      \
			err = std::string("device list work: ") + #call + ": " + hipGetErrorString(e_); \
			return e_ == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;  \
		}                                                            \
	} while (0)

int Solver::upload_tree_device() {
	hipStream_t s = stream;
	const DeviceTree& D = dtree;
	const int64_t nn = D.nn;
	if (D.level_begin.size() < 2 || D.level_begin.back() != nn) {
		err = "device tree work: the level ranges do not cover the nodes";
		return TG_EINVAL;
	}
	if (max_level + 1 > 64) {  // the local-expansion pass's limit
		err = "tree deeper than the local-expansion pass supports (64 levels)";
		return TG_EINVAL;
	}
	PCHK(d_parts.alloc(std::max<size_t>(size_t(n) * sizeof(int4), 16)));
	PCHK(d_multis.alloc(std::max<size_t>(size_t(nn) * sizeof(MultiSrc), 16)));
	PCHK(d_nodepos.alloc(std::max<size_t>(size_t(nn) * sizeof(int4), 16)));
	PCHK(d_level.alloc(std::max<size_t>(size_t(nn) * 4, 16)));
	PCHK(d_parent.alloc(std::max<size_t>(size_t(nn) * 4, 16)));
	PCHK(d_pleaf.alloc(std::max<size_t>(size_t(n) * 4, 16)));
	PCHK(d_order.alloc(std::max<size_t>(size_t(n) * 8, 16)));
	PCHK(d_Lnode.alloc(std::max<size_t>(size_t(nn) * 35 * sizeof(double), 16)));
	DeviceArray flag, loff, plen, poff, cnt, scan_tmp;
	struct Free {
		hipStream_t s;
		DeviceArray* a[6];
		~Free() {
			(void) hipStreamSynchronize(s);
			for (auto* x : a) x->release();
		}
	} free_on_return{s, {&flag, &loff, &plen, &poff, &cnt, &scan_tmp}};
	const int64_t np = nn + 1;
	PCHK(flag.alloc(np * 8));
	PCHK(loff.alloc(np * 8));
	PCHK(plen.alloc(np * 8));
	PCHK(poff.alloc(np * 8));
	PCHK(cnt.alloc(64));
	if (n > 0) {
		hipLaunchKernelGGL(k_tw_parts, dim3(blocks(n, 256)), dim3(256), 0, s, D.pos, D.mass, n, d_parts.as<int4>());
		PCHK(hipGetLastError());
		PCHK(hipMemcpyAsync(d_order.p, D.order, size_t(n) * 8, hipMemcpyDeviceToDevice, s));
	}
	order_on_device = true;
	hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, cnt.as<unsigned long long>(), 1);
	PCHK(hipGetLastError());
	hipLaunchKernelGGL(k_tw_nodes, dim3(blocks(np, 256)), dim3(256), 0, s, D.nodes, D.multi, nn, d_multis.as<MultiSrc>(), d_nodepos.as<int4>(),
			d_level.as<int>(), d_parent.as<int>(), flag.as<int64_t>(), plen.as<int64_t>(), cnt.as<unsigned long long>());
	PCHK(hipGetLastError());
	for (int k = 0; k < 2; k++) {
		const int64_t* in = (k ? plen : flag).as<const int64_t>();
		int64_t* out = (k ? poff : loff).as<int64_t>();
		size_t need = 0;
		PCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, np, s));
		PCHK(scan_tmp.alloc(std::max<size_t>(need, 16)));
		PCHK(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, need, in, out, np, s));
	}
	int64_t tot[3];
	PCHK(hipMemcpyAsync(tot, loff.as<int64_t>() + nn, 8, hipMemcpyDeviceToHost, s));
	PCHK(hipMemcpyAsync(tot + 1, poff.as<int64_t>() + nn, 8, hipMemcpyDeviceToHost, s));
	PCHK(hipMemcpyAsync(tot + 2, cnt.p, 8, hipMemcpyDeviceToHost, s));
	PCHK(hipStreamSynchronize(s));
	if (tot[1] > int64_t(INT32_MAX)) {
		err = "root-to-leaf paths exceed 2^31 entries";
		return TG_EINVAL;
	}
	n_leaf_items = int(tot[0]);
	n_leaves = tot[2];
	path_len = tot[1];
	PCHK(d_leaf_items.alloc(std::max<size_t>(size_t(tot[0]) * sizeof(int4), 16)));
	PCHK(d_path.alloc(std::max<size_t>(size_t(tot[1]) * 4, 16)));
	if (nn > 0) {
		hipLaunchKernelGGL(k_tw_leaves, dim3(blocks(nn, 256)), dim3(256), 0, s, D.nodes, nn, flag.as<const int64_t>(), loff.as<const int64_t>(),
				poff.as<const int64_t>(), d_leaf_items.as<int4>(), d_path.as<int>());
		PCHK(hipGetLastError());
	}
	if (tot[0] > 0) {
		hipLaunchKernelGGL(k_tw_pleaf, dim3(unsigned(tot[0])), dim3(64), 0, s, d_leaf_items.as<const int4>(), d_pleaf.as<int>());
		PCHK(hipGetLastError());
	}
	level_begin.assign(D.level_begin.begin(), D.level_begin.end());
	PCHK(hipStreamSynchronize(s));
	return TG_OK;
}

int Solver::upload_lists_device() {
	hipStream_t s = stream;
	const DeviceLists& L = dlists;
	const int64_t nn = n_nodes;
	const int64_t nl = n_leaf_items;
	const int chunk = std::min(NODE_CHUNK_MAX, std::max(1, node_chunk));
	if (L.nn != nn || !L.ptr_all) {
		err = "device list work: the device walk's lists do not belong to this tree";
		return TG_EINVAL;
	}
	for (int k = 0; k < 8; k++)
		if (L.len[k] > int64_t(INT32_MAX)) {
			err = "an interaction list exceeds 2^31 entries (32-bit list offsets)";
			return TG_EINVAL;
		}
	std::vector<DeviceArray> tmp;  // scratch, freed on return (after the stream drains)
	struct Drain {
		hipStream_t s;
		std::vector<DeviceArray>& t;
		~Drain() {
			(void) hipStreamSynchronize(s);
			for (auto& a : t) a.release();
		}
	} drain{s, tmp};
	tmp.reserve(64);
	auto scratch = [&](size_t bytes) -> void* {
		tmp.emplace_back();
		if (tmp.back().alloc(std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
		return tmp.back().p;
	};
	void* scan_tmp = nullptr;
	size_t scan_bytes = 0;
	auto exclusive_scan = [&](const int64_t* in, int64_t* out, int64_t m) -> hipError_t {
		size_t need = 0;
		hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, m, s);
		if (e != hipSuccess) return e;
		if (need > scan_bytes) {
			scan_bytes = std::max(need, 2 * scan_bytes);
			scan_tmp = scratch(scan_bytes);
			if (!scan_tmp) return hipErrorOutOfMemory;
		}
		size_t b = scan_bytes;
		return hipcub::DeviceScan::ExclusiveSum(scan_tmp, b, in, out, m, s);
	};
	// the last elements of up to 16 arrays to the host (one copy, one wait)
	int64_t** d_arrs = static_cast<int64_t**>(scratch(16 * sizeof(int64_t*)));
	int64_t* d_at = static_cast<int64_t*>(scratch(16 * sizeof(int64_t)));
	int64_t* d_last = static_cast<int64_t*>(scratch(16 * sizeof(int64_t)));
	if (!d_arrs || !d_at || !d_last) PCHK(hipErrorOutOfMemory);
	auto lasts = [&](const std::vector<const int64_t*>& arrs, const std::vector<int64_t>& at, int64_t* out) -> hipError_t {
		const int m = int(arrs.size());
		hipError_t e = hipMemcpyAsync(d_arrs, arrs.data(), m * sizeof(int64_t*), hipMemcpyHostToDevice, s);
		if (e == hipSuccess) e = hipMemcpyAsync(d_at, at.data(), m * sizeof(int64_t), hipMemcpyHostToDevice, s);
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(k_last, dim3(1), dim3(64), 0, s, d_arrs, d_at, m, d_last);
		if ((e = hipGetLastError()) != hipSuccess) return e;
		if ((e = hipMemcpyAsync(out, d_last, m * sizeof(int64_t), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
		return hipStreamSynchronize(s);
	};
	list_bytes = 0;
	const int64_t np = nn + 1;

	// ---- node kinds ----
	NodeIn ni{};
	for (int q = 0; q < 4; q++) {
		ni.ptr[q] = L.ptr(q);
		ni.val[q] = L.val[q];
	}
	ni.nn = nn;
	ni.chunk = chunk;
	int64_t* nent = static_cast<int64_t*>(scratch(4 * np * 8));
	int64_t* oent = static_cast<int64_t*>(scratch(4 * np * 8));
	int64_t* nchk = static_cast<int64_t*>(scratch(4 * np * 8));
	int64_t* ochk = static_cast<int64_t*>(scratch(4 * np * 8));
	int64_t* nslot = static_cast<int64_t*>(scratch(np * 8));
	if (!nent || !oent || !nchk || !ochk || !nslot) PCHK(hipErrorOutOfMemory);
	PCHK(d_node_slot_ptr.alloc(np * 8));
	int64_t* slot_ptr = d_node_slot_ptr.as<int64_t>();
	hipLaunchKernelGGL(k_node_count, dim3(blocks(np, 256)), dim3(256), 0, s, ni, nent, nchk, nslot);
	PCHK(hipGetLastError());
	for (int q = 0; q < 4; q++) {
		PCHK(exclusive_scan(nent + q * np, oent + q * np, np));
		PCHK(exclusive_scan(nchk + q * np, ochk + q * np, np));
	}
	PCHK(exclusive_scan(nslot, slot_ptr, np));
	int64_t nlen[4], nch[4], nslots;
	{
		int64_t v[9];
		std::vector<const int64_t*> arrs;
		std::vector<int64_t> at;
		for (int q = 0; q < 4; q++) {
			arrs.push_back(oent + q * np);
			at.push_back(nn);
			arrs.push_back(ochk + q * np);
			at.push_back(nn);
		}
		arrs.push_back(slot_ptr);
		at.push_back(nn);
		PCHK(lasts(arrs, at, v));
		for (int q = 0; q < 4; q++) {
			nlen[q] = v[2 * q];
			nch[q] = v[2 * q + 1];
		}
		nslots = v[8];
	}
	for (int q = 0; q < 4; q++)
		if (nlen[q] > int64_t(INT32_MAX)) {
			err = "an interaction list exceeds 2^31 entries (32-bit list offsets)";
			return TG_EINVAL;
		}
	if (nslots > int64_t(INT32_MAX)) {
		err = "partial slots exceed 2^31";
		return TG_EINVAL;
	}
	n_slots = nslots;
	PCHK(d_slot_node.alloc(std::max<size_t>(size_t(nslots) * 4, 16)));
	NodeOut no{};
	no.slot_node = d_slot_node.as<int>();
	for (int q = 0; q < 4; q++) {
		no.chunks[q] = static_cast<int2*>(scratch(size_t(nch[q]) * sizeof(int2)));
		no.chunk_slot[q] = static_cast<int*>(scratch(size_t(nch[q]) * 4));
		if (!no.chunks[q] || !no.chunk_slot[q]) PCHK(hipErrorOutOfMemory);
	}
	if (nn > 0) {
		hipLaunchKernelGGL(k_node_chunks, dim3(blocks(nn, 256)), dim3(256), 0, s, ni, nent, oent, ochk, slot_ptr, no);
		PCHK(hipGetLastError());
	}
	// lists: CC the walk's own, CP expanded; owners; waves
	for (int q = 0; q < 4; q++) {
		KindWork& w = work[NODE_KIND[q]];
		w.list.release();
		if (q & 1) {
			PCHK(w.list.alloc(std::max<size_t>(size_t(nlen[q]) * 4, 16)));
		} else {
			w.list.p = L.val[q];
			w.list.bytes = size_t(L.len[q]) * 4;
			w.list.owned = false;
		}
		w.nlist = nlen[q];
		PCHK(w.owner.alloc(std::max<size_t>(size_t(nlen[q]) * 4, 16)));
		if (nch[q] > 0) {
			hipLaunchKernelGGL(k_node_owner, dim3(blocks(nch[q], 4)), dim3(256), 0, s, no.chunks[q], no.chunk_slot[q], nch[q], w.owner.as<int>());
			PCHK(hipGetLastError());
		}
	}
	if (nn > 0 && (L.len[1] > 0 || L.len[3] > 0)) {
		hipLaunchKernelGGL(k_node_expand, dim3(unsigned(nn)), dim3(64), 0, s, ni, oent, work[CP_DIRECT].list.as<int32_t>(),
				work[CP_EWALD].list.as<int32_t>());
		PCHK(hipGetLastError());
	}
	{
		int64_t* nw[4];
		int64_t* ow[4];
		int64_t nseg[4];
		std::vector<const int64_t*> arrs;
		std::vector<int64_t> at;
		for (int q = 0; q < 4; q++) {
			nseg[q] = (nch[q] + WAVE_SEGMENT - 1) / WAVE_SEGMENT;
			nw[q] = static_cast<int64_t*>(scratch(size_t(nseg[q] + 1) * 8));
			ow[q] = static_cast<int64_t*>(scratch(size_t(nseg[q] + 1) * 8));
			if (!nw[q] || !ow[q]) PCHK(hipErrorOutOfMemory);
			hipLaunchKernelGGL(k_wave_count, dim3(blocks(nseg[q] + 1, 256)), dim3(256), 0, s, no.chunks[q], nch[q], nseg[q], nw[q]);
			PCHK(hipGetLastError());
			PCHK(exclusive_scan(nw[q], ow[q], nseg[q] + 1));
			arrs.push_back(ow[q]);
			at.push_back(nseg[q]);
		}
		int64_t nwaves[4];
		PCHK(lasts(arrs, at, nwaves));
		for (int q = 0; q < 4; q++) {
			KindWork& w = work[NODE_KIND[q]];
			PCHK(w.items.alloc(std::max<size_t>(size_t(nwaves[q]) * sizeof(int2), 16)));
			w.nitems = int(nwaves[q]);
			if (nseg[q] > 0) {
				hipLaunchKernelGGL(k_wave_write, dim3(blocks(nseg[q], 256)), dim3(256), 0, s, no.chunks[q], nch[q], nseg[q], ow[q],
						w.items.as<int2>());
				PCHK(hipGetLastError());
			}
			list_bytes += int64_t(nwaves[q]) * int64_t(sizeof(int2)) + nlen[q] * 4;
		}
	}

	// ---- leaf kinds ----
	LeafIn li{};
	for (int q = 0; q < 4; q++) {
		li.ptr[q] = L.ptr(4 + q);
		li.val[q] = L.val[4 + q];
	}
	li.leaf_items = d_leaf_items.as<int4>();
	li.parts = d_parts.as<int4>();
	li.pleaf = d_pleaf.as<int>();
	li.nodes = L.nodes;
	li.nl = nl;
	li.ewald = ewald;
	li.hnear = soft_len * 1.001 + 1e-9;  // the host's margin over the kernel's single-precision test
	int4* box = static_cast<int4*>(scratch(size_t(std::max<int64_t>(nn, 1)) * 2 * sizeof(int4)));
	const int64_t lp = nl + 1;
	int64_t* lent = static_cast<int64_t*>(scratch(4 * lp * 8));
	int64_t* olent = static_cast<int64_t*>(scratch(4 * lp * 8));
	int64_t* litem = static_cast<int64_t*>(scratch(4 * lp * 8));
	int64_t* olitem = static_cast<int64_t*>(scratch(4 * lp * 8));
	int64_t* lnear = static_cast<int64_t*>(scratch(lp * 8));
	int64_t* sums = static_cast<int64_t*>(scratch(8 * lp * 8));   // per leaf, 8 columns
	int64_t* sum_tot = static_cast<int64_t*>(scratch(8 * 8));     // their totals
	if (!box || !lent || !olent || !litem || !olitem || !lnear || !sums || !sum_tot) PCHK(hipErrorOutOfMemory);
	li.box = box;
	li.near = static_cast<uint8_t*>(scratch(size_t(std::max<int64_t>(L.len[5], 2)) / 2 + 1));
	if (!li.near) PCHK(hipErrorOutOfMemory);
	if (nl > 0) {
		hipLaunchKernelGGL(k_leaf_boxes, dim3(unsigned(nl)), dim3(64), 0, s, li.leaf_items, li.parts, box);
		PCHK(hipGetLastError());
	}
	hipLaunchKernelGGL(k_leaf_count, dim3(unsigned(lp)), dim3(64), 0, s, li, lent, lnear, litem, sums);
	PCHK(hipGetLastError());
	for (int q = 0; q < 4; q++) {
		PCHK(exclusive_scan(lent + q * lp, olent + q * lp, lp));
		PCHK(exclusive_scan(litem + q * lp, olitem + q * lp, lp));
	}
	for (int k = 0; k < 8; k++) {
		size_t need = 0;
		PCHK(hipcub::DeviceReduce::Sum(nullptr, need, sums + k * lp, sum_tot + k, lp, s));
		void* tmp_r = scratch(need);
		if (!tmp_r) PCHK(hipErrorOutOfMemory);
		PCHK(hipcub::DeviceReduce::Sum(tmp_r, need, sums + k * lp, sum_tot + k, lp, s));
	}
	int64_t llen[4], nitems[4];
	uint64_t census[8];
	{
		std::vector<const int64_t*> arrs;
		std::vector<int64_t> at;
		for (int q = 0; q < 4; q++) {
			arrs.push_back(olent + q * lp);
			at.push_back(nl);
			arrs.push_back(olitem + q * lp);
			at.push_back(nl);
		}
		for (int k = 0; k < 8; k++) {
			arrs.push_back(sum_tot);
			at.push_back(k);
		}
		int64_t v[16];
		PCHK(lasts(arrs, at, v));
		for (int q = 0; q < 4; q++) {
			llen[q] = v[2 * q];
			nitems[q] = v[2 * q + 1];
		}
		for (int k = 0; k < 8; k++) census[k] = uint64_t(v[8 + k]);
	}
	for (int q = 0; q < 4; q++) {
		if (int64_t(census[4 + q]) != L.len[4 + q]) {
			err = "device list work: leaf-kind entries outside the non-empty leaves";
			return TG_EINVAL;
		}
		if (llen[q] > int64_t(INT32_MAX) || nitems[q] > int64_t(INT32_MAX)) {
			err = "an interaction list exceeds 2^31 entries (32-bit list offsets)";
			return TG_EINVAL;
		}
	}
	LeafOut lo{};
	for (int q = 0; q < 4; q++) {
		KindWork& w = work[LEAF_KIND[q]];
		w.list.release();
		if (q & 1) {
			PCHK(w.list.alloc(std::max<size_t>(size_t(llen[q]) * 4, 16)));
			lo.list[q] = w.list.as<int32_t>();
		} else {
			w.list.p = L.val[4 + q];
			w.list.bytes = size_t(L.len[4 + q]) * 4;
			w.list.owned = false;
		}
		w.owner.release();
		w.nlist = llen[q];
		lo.items[q] = static_cast<int4*>(scratch(size_t(nitems[q]) * sizeof(int4)));
		lo.key[q] = static_cast<unsigned long long*>(scratch(size_t(nitems[q]) * 8));
		lo.idx[q] = static_cast<int*>(scratch(size_t(nitems[q]) * 4));
		if (!lo.items[q] || !lo.key[q] || !lo.idx[q]) PCHK(hipErrorOutOfMemory);
	}
	if (nl > 0) {
		hipLaunchKernelGGL(k_leaf_write, dim3(unsigned(nl)), dim3(64), 0, s, li, lent, olent, lnear, olitem, lo);
		PCHK(hipGetLastError());
	}
	// heaviest cost class first, leaf order within a class: a stable radix sort of (class, index),
	// descending
	for (int q = 0; q < 4; q++) {
		KindWork& w = work[LEAF_KIND[q]];
		const int m = int(nitems[q]);
		PCHK(w.items.alloc(std::max<size_t>(size_t(m) * sizeof(int4), 16)));
		w.nitems = m;
		list_bytes += int64_t(m) * int64_t(sizeof(int4));
		if (m == 0) continue;
		unsigned long long* key_out = static_cast<unsigned long long*>(scratch(size_t(m) * 8));
		int* idx_out = static_cast<int*>(scratch(size_t(m) * 4));
		if (!key_out || !idx_out) PCHK(hipErrorOutOfMemory);
		size_t need = 0;
		PCHK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, need, lo.key[q], key_out, lo.idx[q], idx_out, m, 0, 8, s));
		void* sort_tmp = scratch(need);
		if (!sort_tmp) PCHK(hipErrorOutOfMemory);
		PCHK(hipcub::DeviceRadixSort::SortPairsDescending(sort_tmp, need, lo.key[q], key_out, lo.idx[q], idx_out, m, 0, 8, s));
		hipLaunchKernelGGL(k_gather_items, dim3(blocks(m, 256)), dim3(256), 0, s, lo.items[q], idx_out, int64_t(m), w.items.as<int4>());
		PCHK(hipGetLastError());
	}
	// the census as the reference counts it (node kinds: sources; leaf kinds: sinks x sources)
	Census& C = tree.census;
	C.CC_direct = uint64_t(nlen[0]);
	C.CP_direct = uint64_t(nlen[1]);
	C.CC_ewald = uint64_t(nlen[2]);
	C.CP_ewald = uint64_t(nlen[3]);
	C.PC_direct = census[0];
	C.PP_direct = census[1];
	C.PC_ewald = census[2];
	C.PP_ewald = census[3];
	PCHK(hipStreamSynchronize(s));
	return TG_OK;
}

}  // namespace tg
