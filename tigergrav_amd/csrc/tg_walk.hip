// tg_walk.hip -- the dual-tree walk (tree::kick_fmm's list construction, src/tree.cpp:378-635) on
// the device, producing the same lists as the host walk (Tree::walk).
//
// The recursion is flattened into steps.  A step holds work units, each one checklist pass of
// one node:
//   FIRST  (node n, direct checklist D, Ewald checklist E): the first MAC pass every non-empty
//          node makes (src/tree.cpp:424-494): far entries become CC (node) or CP (opened leaf's
//          particle range); near leaves are passed on opened, near interior nodes as their two
//          children (after all the opened leaves, as the reference appends them).  An interior
//          node's two next lists become the checklists of both children in the next step; a
//          leaf's become its drains.
//   DRAIN  (leaf n, one checklist, direct or Ewald): one round of the leaf's drain loop
//          (src/tree.cpp:530-602): opened leaves become PP ranges, far nodes PC, the rest is
//          opened or split for the next round.
// Per step, on the device: k_walk_count (one wave per unit, 64 checklist entries at a time)
// classifies every entry and counts every unit's outputs, k_walk_sizes turns the counts
// into per-unit sizes (entries of each output, of the next checklists, next units), exclusive
// scans give the offsets, k_walk_write (one wave per unit, from the kept categories) writes and
// k_walk_units emits the next step's units; the
// host reads back only the step's totals (to size the pools).  Each node's outputs are segments
// in step order, which is the host walk's order; k_walk_place puts them into node-order CSR at
// the end.  The MAC arithmetic is the host walk's, in FP64 without
// contraction, so every decision is bit-identical.
#include "tg_walk.hpp"
#include "tg_devmem.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

namespace tg {
namespace {

enum { MODE_FIRST = 0, MODE_DRAIN_D = 1, MODE_DRAIN_E = 2 };

struct WalkUnit {
	int32_t node, mode, len_d, len_e;
	int64_t off_d, off_e;
	int32_t child, pad;  // a FIRST unit's sibling (0, 1): the two children of a node share their checklists
};

struct WalkParams {
	const WalkNode* nodes;
	const int32_t* lists;  // this step's checklists
	double theta_inv, h;
	float r_ewald;
	int ewald;
};

// the node of every unit of a step (kept for the final placement)
__global__ __launch_bounds__(64) void k_walk_nodes_of(const WalkUnit* __restrict__ units, int nu, int32_t* __restrict__ nodes_of) {
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i < nu) nodes_of[i] = units[i].node;
}

__device__ __forceinline__ double pos_d(int32_t x) {
#pragma clang fp contract(off)
	return double(x) * (1.0 / 4294967296.0) + 0.5;
}

__device__ __forceinline__ double near_sep(const double* x) {
#pragma clang fp contract(off)
	double d = 0.0;
	for (int k = 0; k < 3; k++) {
		const double ax = fabs(x[k]);
		const double o = 1.0 - ax;
		const double t = o < ax ? o : ax;
		d += t * t;
	}
	return sqrt(d);
}

// the host walk's direct_far / ewald_far (tg_tree.cpp walk_rec) for sink n (centre X) and source c
__device__ __forceinline__ bool far_test(const WalkParams& P, const WalkNode& n, const double* X, const WalkNode& c, bool ewald_kind) {
#pragma clang fp contract(off)
	double dX[3];
	for (int k = 0; k < 3; k++) dX[k] = X[k] - pos_d(c.x[k]);
	const float rsum = n.r + c.r;
	if (!ewald_kind) {
		const double dx = P.ewald ? near_sep(dX) : sqrt(dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]);
		return dx > double(rsum) * P.theta_inv && dx > double(rsum) + P.h;
	}
	const double ns = near_sep(dX);
	const double dx = 0.25 < ns ? ns : 0.25;
	if (dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2] > 0.0) return dx > double(rsum) * P.theta_inv && dx > double(rsum) + P.h;
	return n.rmax < P.r_ewald;
}

// Checklist entries are classified 64 at a time by one wave, each lane one entry, in checklist
// order.  Categories: FIRST {0 cc, 1 cp range, 2 opened, 3 deferred}; DRAIN {0 pp range, 1 pc,
// 2 opened, 3 deferred}; 4: an empty node (nothing).  The count pass keeps every entry's category
// (one byte, beside the entry) for the write pass, which then places each lane's outputs by its
// rank among the lanes of the same category: the serial pass's order.  The two children of a node
// read the same checklists, so entry i keeps a category byte per child (2 i + child).
enum { CAT_SKIP = 4 };

__device__ __forceinline__ unsigned long long lanes_below(int lane) { return lane ? ~0ull >> (64 - lane) : 0ull; }

__device__ __forceinline__ int classify(const WalkParams& P, const WalkNode& n, const double* X, int32_t e, bool ewald_kind, bool drain) {
	const bool opened = e < 0;
	const WalkNode cn = P.nodes[opened ? ~e : e];
	if (cn.pend == cn.pbegin) return CAT_SKIP;
	const bool far = far_test(P, n, X, cn, ewald_kind);
	if (drain) return opened ? 0 : far ? 1 : cn.c0 < 0 ? 2 : 3;
	return far ? (opened ? 1 : 0) : cn.c0 < 0 ? 2 : 3;
}

// one checklist pass classified and counted by the wave; c (4, the same in every lane) += counts
__device__ void count_pass(const WalkParams& P, const WalkNode& n, const double* X, const int32_t* list, uint8_t* codes, int len,
		bool ewald_kind, bool drain, int* c) {
	const int lane = threadIdx.x & 63;
	for (int b = 0; b < len; b += 64) {
		const int i = b + lane;
		int cat = CAT_SKIP;
		if (i < len) {
			cat = classify(P, n, X, list[i], ewald_kind, drain);
			codes[2 * i] = uint8_t(cat);
		}
		for (int k = 0; k < 4; k++) c[k] += __popcll(__ballot(cat == k));
	}
}

// one wave (work group) per unit
__global__ __launch_bounds__(64) void k_walk_count(WalkParams P, uint8_t* __restrict__ codes, const WalkUnit* __restrict__ units, int nunits,
		int32_t* __restrict__ counts) {
	const int i = blockIdx.x;
	if (i >= nunits) return;
	const WalkUnit u = units[i];
	const WalkNode n = P.nodes[u.node];
	const double X[3] = {pos_d(n.x[0]), pos_d(n.x[1]), pos_d(n.x[2])};
	int c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (u.mode == MODE_FIRST) {
		count_pass(P, n, X, P.lists + u.off_d, codes + 2 * u.off_d + u.child, u.len_d, false, false, c);
		if (P.ewald) count_pass(P, n, X, P.lists + u.off_e, codes + 2 * u.off_e + u.child, u.len_e, true, false, c + 4);
	} else {
		count_pass(P, n, X, P.lists + u.off_d, codes + 2 * u.off_d + u.child, u.len_d, u.mode == MODE_DRAIN_E, true, c);
	}
	if ((threadIdx.x & 63) < 8) counts[8 * int64_t(i) + (threadIdx.x & 63)] = c[threadIdx.x & 63];
}

__device__ __forceinline__ void put_at(int32_t* p, int64_t i, int64_t cap, int32_t v, int32_t* overflow) {
	if (i < cap)
		p[i] = v;
	else
		*overflow = 1;
}

// one checklist pass, written by the wave from the count pass's categories: FIRST into (a: cc, b: cp)
// and next; DRAIN into (a: pp, b: pc) and next.  Next holds the opened leaves first (n_opened of
// them), then the deferred nodes' children.
__device__ void write_pass(const WalkParams& P, const int32_t* list, const uint8_t* codes, int len, bool drain, int32_t* a, int64_t oa,
		int64_t capa, int32_t* b, int64_t ob, int64_t capb, int32_t* next, int64_t on, int64_t capn, int n_opened, int32_t* overflow) {
	const int lane = threadIdx.x & 63;
	const unsigned long long below = lanes_below(lane);
	int64_t op = on, od = on + n_opened;
	for (int base = 0; base < len; base += 64) {
		const int i = base + lane;
		const int cat = i < len ? int(codes[2 * i]) : CAT_SKIP;
		const int32_t e = i < len ? list[i] : 0;
		unsigned long long m[4];
		for (int k = 0; k < 4; k++) m[k] = __ballot(cat == k);
		if (cat != CAT_SKIP) {
			const int32_t id = e < 0 ? ~e : e;
			const int r = __popcll(m[cat] & below);
			if (cat == 2) {
				put_at(next, op + r, capn, ~id, overflow);
			} else if (cat == 3) {
				const WalkNode cn = P.nodes[id];
				put_at(next, od + 2 * r, capn, cn.c0, overflow);
				put_at(next, od + 2 * r + 1, capn, cn.c1, overflow);
			} else if ((cat == 0) == drain) {  // a range: PP (drain) or CP (first pass)
				const WalkNode cn = P.nodes[id];
				int32_t* dst = drain ? a : b;
				const int64_t o = (drain ? oa : ob) + 2 * r;
				const int64_t cap = drain ? capa : capb;
				put_at(dst, o, cap, cn.pbegin, overflow);
				put_at(dst, o + 1, cap, cn.pend, overflow);
			} else {  // a node: CC (first pass) or PC (drain)
				if (drain)
					put_at(b, ob + r, capb, id, overflow);
				else
					put_at(a, oa + r, capa, id, overflow);
			}
		}
		const int n0 = __popcll(m[0]), n1 = __popcll(m[1]);
		oa += drain ? 2 * n0 : n0;
		ob += drain ? n1 : 2 * n1;
		op += __popcll(m[2]);
		od += 2 * __popcll(m[3]);
	}
}

enum { K_CC_D, K_CP_D, K_CC_E, K_CP_E, K_PC_D, K_PP_D, K_PC_E, K_PP_E, K_N };
constexpr int COL_NEXT = K_N, COL_UNITS = K_N + 1, NCOL = K_N + 2;  // sizes: 8 outputs, next lists, next units

// per unit, the entries of each output, of its next checklists and the number of its next units
__global__ __launch_bounds__(64) void k_walk_sizes(const WalkNode* __restrict__ nodes, const WalkUnit* __restrict__ units, int nu,
		const int32_t* __restrict__ counts, int ewald, int64_t* __restrict__ sizes) {
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i >= nu) return;
	const WalkUnit u = units[i];
	const int32_t* c = counts + 8 * int64_t(i);
	int64_t z[NCOL] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
	if (u.mode == MODE_FIRST) {
		const int64_t ld = int64_t(c[2]) + 2 * int64_t(c[3]);
		const int64_t le = ewald ? int64_t(c[6]) + 2 * int64_t(c[7]) : 0;
		z[K_CC_D] = c[0];
		z[K_CP_D] = 2 * int64_t(c[1]);
		if (ewald) {
			z[K_CC_E] = c[4];
			z[K_CP_E] = 2 * int64_t(c[5]);
		}
		z[COL_NEXT] = ld + le;
		const WalkNode n = nodes[u.node];
		if (n.c0 >= 0)
			z[COL_UNITS] = int64_t(nodes[n.c0].pend > nodes[n.c0].pbegin) + int64_t(nodes[n.c1].pend > nodes[n.c1].pbegin);
		else
			z[COL_UNITS] = int64_t(ld > 0) + int64_t(le > 0);
	} else {
		const bool ew = u.mode == MODE_DRAIN_E;
		const int64_t l = int64_t(c[2]) + 2 * int64_t(c[3]);
		z[ew ? K_PP_E : K_PP_D] = 2 * int64_t(c[0]);
		z[ew ? K_PC_E : K_PC_D] = c[1];
		z[COL_NEXT] = l;
		z[COL_UNITS] = l > 0;
	}
	for (int k = 0; k < NCOL; k++) sizes[int64_t(k) * nu + i] = z[k];
}

// the sizes of all columns are scanned as one sequence: every column's start (tot[NCOL + k]) and
// total (tot[k])
__global__ void k_walk_totals(const int64_t* __restrict__ sizes, const int64_t* __restrict__ offs, int nu, int64_t* __restrict__ tot) {
	const int k = threadIdx.x;
	if (k < NCOL) {
		const int64_t c0 = offs[int64_t(k) * nu];
		tot[NCOL + k] = c0;
		tot[k] = offs[int64_t(k) * nu + nu - 1] + sizes[int64_t(k) * nu + nu - 1] - c0;
	}
}

// offsets within each column: the one-sequence scan less the column's start
__global__ __launch_bounds__(256) void k_walk_colfix(const int64_t* __restrict__ tot, int nu, int64_t* __restrict__ offs) {
	const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (i < int64_t(NCOL) * nu) offs[i] -= tot[NCOL + i / nu];
}

struct WalkPools {
	int32_t* out[K_N];   // output pools, appended step by step
	int64_t base[K_N];   // this step's start in each output pool
	int64_t cap[K_N + 1];
	int32_t* next;       // the next step's checklists
	int32_t* overflow;
};

// one wave (work group) per unit
__global__ __launch_bounds__(64) void k_walk_write(WalkParams P, const uint8_t* __restrict__ codes, const WalkUnit* __restrict__ units, int nu,
		const int32_t* __restrict__ counts, const int64_t* __restrict__ offs, WalkPools W) {
	const int i = blockIdx.x;
	if (i >= nu) return;
	const WalkUnit u = units[i];
	const int32_t* c = counts + 8 * int64_t(i);
	auto at = [&](int k) { return W.base[k] + offs[int64_t(k) * nu + i]; };
	const int64_t nb = offs[int64_t(COL_NEXT) * nu + i];
	if (u.mode == MODE_FIRST) {
		write_pass(P, P.lists + u.off_d, codes + 2 * u.off_d + u.child, u.len_d, false, W.out[K_CC_D], at(K_CC_D), W.cap[K_CC_D], W.out[K_CP_D],
				at(K_CP_D), W.cap[K_CP_D], W.next, nb, W.cap[K_N], c[2], W.overflow);
		if (P.ewald)
			write_pass(P, P.lists + u.off_e, codes + 2 * u.off_e + u.child, u.len_e, false, W.out[K_CC_E], at(K_CC_E), W.cap[K_CC_E], W.out[K_CP_E],
					at(K_CP_E), W.cap[K_CP_E], W.next, nb + int64_t(c[2]) + 2 * int64_t(c[3]), W.cap[K_N], c[6], W.overflow);
	} else if (u.mode == MODE_DRAIN_D) {
		write_pass(P, P.lists + u.off_d, codes + 2 * u.off_d + u.child, u.len_d, true, W.out[K_PP_D], at(K_PP_D), W.cap[K_PP_D], W.out[K_PC_D],
				at(K_PC_D), W.cap[K_PC_D], W.next, nb, W.cap[K_N], c[2], W.overflow);
	} else {
		write_pass(P, P.lists + u.off_d, codes + 2 * u.off_d + u.child, u.len_d, true, W.out[K_PP_E], at(K_PP_E), W.cap[K_PP_E], W.out[K_PC_E],
				at(K_PC_E), W.cap[K_PC_E], W.next, nb, W.cap[K_N], c[2], W.overflow);
	}
}

// the next step's units: an interior node's first pass hands its two next lists to its non-empty
// children, a leaf's to its drains; a drain round with a non-empty next list continues
__global__ __launch_bounds__(64) void k_walk_units(const WalkNode* __restrict__ nodes, const WalkUnit* __restrict__ units, int nu,
		const int32_t* __restrict__ counts, const int64_t* __restrict__ offs, int ewald, WalkUnit* __restrict__ next_units, int cap) {
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i >= nu) return;
	const WalkUnit u = units[i];
	const int32_t* c = counts + 8 * int64_t(i);
	int64_t q = offs[int64_t(COL_UNITS) * nu + i];
	const int64_t nb = offs[int64_t(COL_NEXT) * nu + i];
	auto emit = [&](WalkUnit w) {
		if (q < cap) next_units[q] = w;
		q++;
	};
	if (u.mode == MODE_FIRST) {
		const int32_t ld = c[2] + 2 * c[3];
		const int32_t le = ewald ? c[6] + 2 * c[7] : 0;
		const WalkNode n = nodes[u.node];
		if (n.c0 >= 0) {
			const int ch[2] = {n.c0, n.c1};
			for (int k = 0; k < 2; k++)
				if (nodes[ch[k]].pend > nodes[ch[k]].pbegin) emit(WalkUnit{ch[k], MODE_FIRST, ld, le, nb, nb + ld, k, 0});
		} else {
			if (ld > 0) emit(WalkUnit{u.node, MODE_DRAIN_D, ld, 0, nb, 0, 0, 0});
			if (le > 0) emit(WalkUnit{u.node, MODE_DRAIN_E, le, 0, nb + ld, 0, 0, 0});
		}
	} else {
		const int32_t l = c[2] + 2 * c[3];
		if (l > 0) emit(WalkUnit{u.node, u.mode, l, 0, nb, 0, 0, 0});
	}
}

// per node and output kind, the entries of its segment in this step (one unit per node and kind
// per step at most: one first pass, one direct and one Ewald drain round)
__global__ __launch_bounds__(64) void k_walk_node_len(const int32_t* __restrict__ nodes_of, int nu, const int64_t* __restrict__ sizes,
		int64_t nn, int64_t* __restrict__ node_len) {
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i >= nu) return;
	const int64_t nd = nodes_of[i];
	for (int k = 0; k < K_N; k++) {
		// only the one unit of (node, kind) with entries touches its counter (a direct and an Ewald
		// drain of one leaf run in the same step and share the node)
		const int64_t len = sizes[int64_t(k) * nu + i];
		if (len > 0) node_len[int64_t(k) * (nn + 1) + nd] += len;
	}
}

// a step's segments into the node-order lists: dst = ptr[node] + fill[node].  One wave per unit:
// lanes 0..7 read the eight kinds' lengths and places at once, then all lanes copy the unit's
// segments as one run of entries (64 at a time across the kinds), so a unit costs one round of
// metadata loads and one of copies instead of one of each per kind.  Only the one unit of (node,
// kind) with entries moves its fill counter (k_walk_node_len).
__global__ __launch_bounds__(64) void k_walk_place(const int32_t* __restrict__ nodes_of, int nu, const int64_t* __restrict__ sizes,
		const int64_t* __restrict__ offs, const int64_t* __restrict__ base, int64_t nn, const int64_t* __restrict__ ptr,
		int64_t* __restrict__ fill, WalkPools src, int32_t* const* __restrict__ dst) {
	__shared__ int64_t s_pre[K_N + 1];
	__shared__ const int32_t* s_from[K_N];
	__shared__ int32_t* s_to[K_N];
	const int i = blockIdx.x;
	const int lane = threadIdx.x;
	const int64_t nd = nodes_of[i];
	int64_t len = 0, f = 0;
	if (lane < K_N) {
		len = sizes[int64_t(lane) * nu + i];
		if (len > 0) {
			f = fill[int64_t(lane) * (nn + 1) + nd];
			s_from[lane] = src.out[lane] + base[lane] + offs[int64_t(lane) * nu + i];
			s_to[lane] = dst[lane] + ptr[int64_t(lane) * (nn + 1) + nd] + f;
		}
	}
	// the start of each kind's part of the run: an exclusive scan of the eight lengths
	int64_t pre = 0;
#pragma unroll
	for (int k = 0; k < K_N - 1; k++) {
		const int64_t lk = __shfl(len, k, 64);
		if (lane > k) pre += lk;
	}
	if (lane < K_N) s_pre[lane] = pre;
	if (lane == K_N - 1) s_pre[K_N] = pre + len;
	__syncthreads();
	const int64_t total = s_pre[K_N];
	for (int64_t t = lane; t < total; t += 64) {
		int k = 0;
#pragma unroll
		for (int q = 1; q < K_N; q++) k += (t >= s_pre[q]);  // the last kind that starts at or before t
		const int64_t o = t - s_pre[k];
		s_to[k][o] = s_from[k][o];
	}
	if (lane < K_N && len > 0) fill[int64_t(lane) * (nn + 1) + nd] = f + len;
}

// a device buffer of int32 that grows, keeping its contents
struct Pool {
	int32_t* p = nullptr;
	int64_t cap = 0;
	hipError_t reserve(int64_t n, hipStream_t s, bool keep = true) {
		if (n <= cap) return hipSuccess;
		int64_t nc = std::max<int64_t>(n, cap * 2);
		nc = std::max<int64_t>(nc, 1024);
		int32_t* q = nullptr;
		hipError_t e = dev_malloc(&q, size_t(nc) * sizeof(int32_t));
		if (e != hipSuccess) return e;
		if (p && cap > 0 && keep) {
			if ((e = hipMemcpyAsync(q, p, size_t(cap) * sizeof(int32_t), hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
			if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
		}
		if (p) dev_free(p);
		p = q;
		cap = nc;
		return hipSuccess;
	}
	void release() {
		if (p) dev_free(p);
		p = nullptr;
		cap = 0;
	}
};

// device memory handed out in 256-byte-aligned pieces from blocks of >= 64 MiB, freed together
struct Arena {
	std::vector<void*> blocks;
	char* cur = nullptr;
	size_t left = 0;
	template <class T>
	hipError_t get(size_t count, T** p) {
		const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
		if (bytes > left) {
			const size_t b = std::max<size_t>(bytes, size_t(64) << 20);
			void* q = nullptr;
			const hipError_t e = dev_malloc(&q, b);
			if (e != hipSuccess) return e;
			blocks.push_back(q);
			cur = static_cast<char*>(q);
			left = b;
		}
		*p = reinterpret_cast<T*>(cur);
		cur += bytes;
		left -= bytes;
		return hipSuccess;
	}
	void release() {
		for (void* b : blocks) dev_free(b);
		blocks.clear();
		cur = nullptr;
		left = 0;
	}
};

// one step's record for the final placement: its units' nodes, sizes and offsets (device) and
// the output bases (host)
struct StepLog {
	int32_t* nodes_of = nullptr;
	int64_t* sizes = nullptr;
	int64_t* offs = nullptr;
	int nu = 0;
	int64_t base[K_N];
};

}  // namespace

static double wall() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

void DeviceLists::release() {
	if (nodes) dev_free(nodes);
	if (ptr_all) dev_free(ptr_all);
	for (auto& v : val)
		if (v) dev_free(v);
	*this = DeviceLists();
}

hipError_t device_walk(Tree& T, int device, hipStream_t s, std::string& err, double* t_out, DeviceLists* keep, WalkNode* given,
		int64_t given_nn, WalkHint* hint) {
	static_assert(sizeof(WalkHint::out) / sizeof(int64_t) == K_N, "one hint per output kind");
	double tw[5] = {wall(), 0, 0, 0, 0};
	const int64_t nn = given ? given_nn : int64_t(T.nodes.size());
	hipError_t e = hipSuccess;
	void* dev_nodes = nullptr;
	WalkUnit* dev_units = nullptr;
	WalkUnit* dev_next_units = nullptr;
	int32_t* dev_counts = nullptr;
	int64_t* dev_tot = nullptr;
	int64_t* dev_node_len = nullptr;
	int64_t* dev_fill = nullptr;
	int64_t* dev_base = nullptr;
	int32_t** dev_dst = nullptr;
	int32_t* dev_overflow = nullptr;
	void* scan_tmp = nullptr;
	size_t scan_tmp_bytes = 0;
	int32_t overflow = 0;
	int64_t units_cap = 0, next_units_cap = 0, counts_cap = 0, codes_cap = 0, in_len = 2;
	uint8_t* dev_codes = nullptr;  // the category of every checklist entry of the step: byte 2 i + child for entry i
	Arena arena;
	Pool lists_in, lists_out, out[K_N], fin[K_N];
	std::vector<StepLog> log;
	std::vector<int64_t> base_h;  // every step's pool starts, for the placement
	int64_t max_lists = 2;        // the longest checklist set of a step, for the hint
	int64_t tot[K_N] = {};
	int nu = 0;
	WalkParams P{};
	WalkNode root_node{};
	hipError_t pending = hipSuccess;
#define WCHK(call)                                     \
	do {                                               \
		e = (call);                                    \
		if (e != hipSuccess) {                         \
			err = std::string(#call) + ": " + hipGetErrorString(e); \
			goto done;                                 \
		}                                              \
	} while (0)
#define WLAUNCH(name)                                                                       \
	do {                                                                                    \
		e = hipGetLastError();                                                              \
		if (e != hipSuccess) {                                                              \
			err = std::string("launch ") + name + ": " + hipGetErrorString(e);              \
			if (pending != hipSuccess) err += std::string(" (pending before the walk: ") + hipGetErrorString(pending) + ")"; \
			goto done;                                                                      \
		}                                                                                   \
	} while (0)
	pending = hipGetLastError();  // an error left by an earlier call is not the walk's
	dev_nodes = given;  // owned from here on
	WCHK(hipSetDevice(device));
	if (!given) {
		hvec<WalkNode> wn(size_t(std::max<int64_t>(nn, 1)));
		walk_nodes(T, wn.data());
		WCHK(dev_malloc(&dev_nodes, wn.size() * sizeof(WalkNode)));
		WCHK(hipMemcpyAsync(dev_nodes, wn.data(), wn.size() * sizeof(WalkNode), hipMemcpyHostToDevice, s));
	}
	WCHK(dev_malloc(&dev_overflow, sizeof(int32_t)));
	WCHK(hipMemsetAsync(dev_overflow, 0, sizeof(int32_t), s));
	WCHK(dev_malloc(&dev_tot, 2 * NCOL * sizeof(int64_t)));
	if (hint) {
		for (int k = 0; k < K_N; k++)
			if (hint->out[k] > 0) WCHK(out[k].reserve(hint->out[k], s, false));
		if (hint->lists > 0) {
			WCHK(lists_in.reserve(hint->lists, s, false));
			WCHK(lists_out.reserve(hint->lists, s, false));
		}
		if (hint->counts > 0) {
			counts_cap = hint->counts;
			WCHK(dev_malloc(&dev_counts, size_t(counts_cap) * sizeof(int32_t)));
		}
		if (hint->codes > 0) {
			codes_cap = hint->codes;
			WCHK(dev_malloc(&dev_codes, size_t(codes_cap)));
		}
	}
	P.nodes = static_cast<const WalkNode*>(dev_nodes);
	P.theta_inv = 1.0 / T.P.theta;
	P.h = T.P.soft_len;
	P.r_ewald = float(2.0 * T.P.theta * std::pow(1.0e-3 / 8.0, 1.0 / 3.0));  // src/tree.cpp:403-404
	P.ewald = T.P.ewald ? 1 : 0;
	if (nn > 0) {
		if (given) {
			WCHK(hipMemcpyAsync(&root_node, dev_nodes, sizeof(WalkNode), hipMemcpyDeviceToHost, s));
			WCHK(hipStreamSynchronize(s));
		} else {
			root_node.pbegin = T.nodes[0].pbegin;
			root_node.pend = T.nodes[0].pend;
		}
	}
	if (nn > 0 && root_node.pend > root_node.pbegin) {
		// the root's checklists: {root} for both kinds
		const int32_t root[2] = {0, 0};
		const WalkUnit r0{0, MODE_FIRST, 1, P.ewald ? 1 : 0, 0, 1, 0, 0};
		WCHK(lists_in.reserve(2, s));
		WCHK(hipMemcpyAsync(lists_in.p, root, sizeof(root), hipMemcpyHostToDevice, s));
		units_cap = 1024;
		WCHK(dev_malloc(&dev_units, size_t(units_cap) * sizeof(WalkUnit)));
		WCHK(hipMemcpyAsync(dev_units, &r0, sizeof(r0), hipMemcpyHostToDevice, s));
		nu = 1;
	}
	WCHK(hipStreamSynchronize(s));
	tw[1] = wall();
	while (nu > 0) {
		const unsigned g = unsigned((nu + 63) / 64);
		StepLog L;
		L.nu = nu;
		for (int k = 0; k < K_N; k++) L.base[k] = tot[k];
		WCHK(arena.get(size_t(nu), &L.nodes_of));
		WCHK(arena.get(size_t(nu) * NCOL, &L.sizes));
		WCHK(arena.get(size_t(nu) * NCOL, &L.offs));
		log.push_back(L);
		if (int64_t(nu) * 8 > counts_cap) {
			WCHK(hipStreamSynchronize(s));  // the last step's kernels read the old counts
			if (dev_counts) dev_free(dev_counts);
			dev_counts = nullptr;
			counts_cap = std::max<int64_t>(int64_t(nu) * 8, 2 * counts_cap);
			WCHK(dev_malloc(&dev_counts, size_t(counts_cap) * sizeof(int32_t)));
		}
		if (2 * in_len > codes_cap) {
			WCHK(hipStreamSynchronize(s));  // the last step's kernels read the old categories
			if (dev_codes) dev_free(dev_codes);
			dev_codes = nullptr;
			codes_cap = std::max<int64_t>(2 * in_len, 2 * codes_cap);
			WCHK(dev_malloc(&dev_codes, size_t(codes_cap)));
		}
		P.lists = lists_in.p;
		WLAUNCH("walk step set-up");
		hipLaunchKernelGGL(k_walk_count, dim3(unsigned(nu)), dim3(64), 0, s, P, dev_codes, dev_units, nu, dev_counts);
		WLAUNCH("k_walk_count");
		hipLaunchKernelGGL(k_walk_sizes, dim3(g), dim3(64), 0, s, P.nodes, dev_units, nu, dev_counts, P.ewald, L.sizes);
		WLAUNCH("k_walk_sizes");
		hipLaunchKernelGGL(k_walk_nodes_of, dim3(g), dim3(64), 0, s, dev_units, nu, L.nodes_of);
		WLAUNCH("k_walk_nodes_of");
		{
			// all columns in one scan, then each column's start taken off
			const int64_t m = int64_t(nu) * NCOL;
			size_t need = 0;
			WCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, L.sizes, L.offs, m, s));
			if (need > scan_tmp_bytes) {
				if (scan_tmp) dev_free(scan_tmp);
				scan_tmp = nullptr;
				scan_tmp_bytes = std::max(need, 2 * scan_tmp_bytes);
				WCHK(dev_malloc(&scan_tmp, scan_tmp_bytes));
			}
			size_t b = scan_tmp_bytes;
			WCHK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, b, L.sizes, L.offs, m, s));
			hipLaunchKernelGGL(k_walk_totals, dim3(1), dim3(64), 0, s, L.sizes, L.offs, nu, dev_tot);
			WLAUNCH("k_walk_totals");
			hipLaunchKernelGGL(k_walk_colfix, dim3(unsigned((m + 255) / 256)), dim3(256), 0, s, dev_tot, nu, L.offs);
			WLAUNCH("k_walk_colfix");
		}
		int64_t t[NCOL];
		WCHK(hipMemcpyAsync(t, dev_tot, sizeof(t), hipMemcpyDeviceToHost, s));
		WCHK(hipStreamSynchronize(s));
		if (t[COL_NEXT] > INT32_MAX || t[COL_UNITS] > INT32_MAX) {
			err = "walk checklists exceed 2^31 entries";
			e = hipErrorInvalidValue;
			goto done;
		}
		for (int k = 0; k < K_N; k++) {
			WCHK(out[k].reserve(std::max<int64_t>(tot[k] + t[k], 1), s));
			tot[k] += t[k];
		}
		WCHK(lists_out.reserve(std::max<int64_t>(t[COL_NEXT], 1), s, false));
		max_lists = std::max(max_lists, t[COL_NEXT]);
		if (t[COL_UNITS] > next_units_cap) {
			if (dev_next_units) dev_free(dev_next_units);
			dev_next_units = nullptr;
			next_units_cap = std::max<int64_t>(t[COL_UNITS], 2 * next_units_cap);
			WCHK(dev_malloc(&dev_next_units, size_t(next_units_cap) * sizeof(WalkUnit)));
		}
		{
			WalkPools W{};
			for (int k = 0; k < K_N; k++) {
				W.out[k] = out[k].p;
				W.base[k] = L.base[k];
				W.cap[k] = out[k].cap;
			}
			W.cap[K_N] = lists_out.cap;
			W.next = lists_out.p;
			W.overflow = dev_overflow;
			hipLaunchKernelGGL(k_walk_write, dim3(unsigned(nu)), dim3(64), 0, s, P, dev_codes, dev_units, nu, dev_counts, L.offs, W);
			WLAUNCH("k_walk_write");
			hipLaunchKernelGGL(k_walk_units, dim3(g), dim3(64), 0, s, P.nodes, dev_units, nu, dev_counts, L.offs, P.ewald, dev_next_units,
					int(next_units_cap));
			WLAUNCH("k_walk_units");
		}
		std::swap(lists_in, lists_out);
		std::swap(dev_units, dev_next_units);
		std::swap(units_cap, next_units_cap);
		nu = int(t[COL_UNITS]);
		in_len = t[COL_NEXT];
	}
	WCHK(hipMemcpyAsync(&overflow, dev_overflow, sizeof(int32_t), hipMemcpyDeviceToHost, s));
	WCHK(hipStreamSynchronize(s));
	if (overflow) {
		err = "device walk: a write pass disagreed with its count pass";
		e = hipErrorUnknown;
		goto done;
	}
	tw[2] = wall();
	if (hint) {
		for (int k = 0; k < K_N; k++) hint->out[k] = tot[k];
		hint->lists = max_lists;
		hint->counts = counts_cap;
		hint->codes = codes_cap;
	}
	{
		// node-order CSR per kind: per-node lengths over all steps, their scan, then every step's
		// segments placed in step order
		const int64_t np = nn + 1;
		WCHK(dev_malloc(&dev_node_len, size_t(K_N * np) * sizeof(int64_t)));
		WCHK(dev_malloc(&dev_fill, size_t(K_N * np) * sizeof(int64_t)));
		WCHK(hipMemsetAsync(dev_node_len, 0, size_t(K_N * np) * sizeof(int64_t), s));
		WCHK(hipMemsetAsync(dev_fill, 0, size_t(K_N * np) * sizeof(int64_t), s));
		for (const StepLog& L : log) {
			hipLaunchKernelGGL(k_walk_node_len, dim3(unsigned((L.nu + 63) / 64)), dim3(64), 0, s, L.nodes_of, L.nu, L.sizes, nn, dev_node_len);
			WLAUNCH("k_walk_node_len");
		}
		int64_t* dev_ptr = nullptr;
		WCHK(dev_malloc(&dev_ptr, size_t(K_N * np) * sizeof(int64_t)));
		for (int k = 0; k < K_N; k++) {
			size_t need = 0;
			e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, dev_node_len + k * np, dev_ptr + k * np, np, s);
			if (e == hipSuccess && need > scan_tmp_bytes) {
				if (scan_tmp) dev_free(scan_tmp);
				scan_tmp = nullptr;
				scan_tmp_bytes = need;
				e = dev_malloc(&scan_tmp, scan_tmp_bytes);
			}
			size_t b = scan_tmp_bytes;
			if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(scan_tmp, b, dev_node_len + k * np, dev_ptr + k * np, np, s);
			if (e == hipSuccess) e = fin[k].reserve(std::max<int64_t>(tot[k], 1), s, false);
			if (e != hipSuccess) {
				dev_free(dev_ptr);
				err = std::string("list assembly: ") + hipGetErrorString(e);
				goto done;
			}
		}
		int32_t* dst_h[K_N];
		for (int k = 0; k < K_N; k++) dst_h[k] = fin[k].p;
		e = dev_malloc(&dev_dst, sizeof(dst_h));
		if (e == hipSuccess) e = hipMemcpyAsync(dev_dst, dst_h, sizeof(dst_h), hipMemcpyHostToDevice, s);
		if (e == hipSuccess) e = dev_malloc(&dev_base, K_N * sizeof(int64_t) * std::max<size_t>(log.size(), 1));
		// every step's pool starts in one copy (the host array outlives the stream's wait below)
		base_h.resize(K_N * std::max<size_t>(log.size(), 1));
		for (size_t st = 0; st < log.size(); st++) std::memcpy(base_h.data() + K_N * st, log[st].base, K_N * sizeof(int64_t));
		if (e == hipSuccess) e = hipMemcpyAsync(dev_base, base_h.data(), base_h.size() * sizeof(int64_t), hipMemcpyHostToDevice, s);
		WalkPools src{};
		for (int k = 0; k < K_N; k++) src.out[k] = out[k].p;
		for (size_t st = 0; st < log.size() && e == hipSuccess; st++) {
			const StepLog& L = log[st];
			hipLaunchKernelGGL(k_walk_place, dim3(unsigned(L.nu)), dim3(64), 0, s, L.nodes_of, L.nu, L.sizes, L.offs,
					dev_base + K_N * st, nn, dev_ptr, dev_fill, src, dev_dst);
			e = hipGetLastError();
		}
		Lists& Ls = T.lists;
		Ls = Lists();
		if (keep) {
			// everything stays on the device: the caller's DeviceLists takes the arrays
			if (e == hipSuccess) e = hipStreamSynchronize(s);
			if (e != hipSuccess) {
				dev_free(dev_ptr);
				err = std::string("list assembly: ") + hipGetErrorString(e);
				goto done;
			}
			keep->release();
			keep->nn = nn;
			keep->ptr_all = dev_ptr;
			keep->nodes = static_cast<WalkNode*>(dev_nodes);
			dev_nodes = nullptr;
			for (int k = 0; k < K_N; k++) {
				keep->val[k] = fin[k].p;
				keep->len[k] = tot[k];
				fin[k].p = nullptr;
				fin[k].cap = 0;
			}
			T.census = Census();
			tw[3] = tw[4] = wall();
			if (t_out)
				for (int k = 0; k < 4; k++) t_out[k] = tw[k + 1] - tw[k];
			goto done;
		}
		hvec<int64_t>* ptrs[K_N] = {&Ls.cc_d_ptr, &Ls.cp_d_ptr, &Ls.cc_e_ptr, &Ls.cp_e_ptr,
				&Ls.pc_d_ptr, &Ls.pp_d_ptr, &Ls.pc_e_ptr, &Ls.pp_e_ptr};
		hvec<int32_t>* vals[K_N] = {&Ls.cc_d, &Ls.cp_d, &Ls.cc_e, &Ls.cp_e, &Ls.pc_d, &Ls.pp_d, &Ls.pc_e, &Ls.pp_e};
		for (int k = 0; k < K_N && e == hipSuccess; k++) {
			ptrs[k]->resize(size_t(np));
			vals[k]->resize(size_t(tot[k]));
			e = hipMemcpyAsync(ptrs[k]->data(), dev_ptr + k * np, size_t(np) * sizeof(int64_t), hipMemcpyDeviceToHost, s);
			if (e == hipSuccess && tot[k] > 0)
				e = hipMemcpyAsync(vals[k]->data(), fin[k].p, size_t(tot[k]) * sizeof(int32_t), hipMemcpyDeviceToHost, s);
		}
		if (e == hipSuccess) e = hipStreamSynchronize(s);
		dev_free(dev_ptr);
		if (e != hipSuccess) {
			err = std::string("list assembly: ") + hipGetErrorString(e);
			goto done;
		}
		tw[3] = wall();
		T.count_census();
		tw[4] = wall();
		if (t_out)
			for (int k = 0; k < 4; k++) t_out[k] = tw[k + 1] - tw[k];
	}
done:
#undef WCHK
#undef WLAUNCH
	(void) hipStreamSynchronize(s);
	if (dev_nodes) dev_free(dev_nodes);
	if (dev_units) dev_free(dev_units);
	if (dev_next_units) dev_free(dev_next_units);
	if (dev_counts) dev_free(dev_counts);
	if (dev_codes) dev_free(dev_codes);
	if (dev_tot) dev_free(dev_tot);
	if (dev_node_len) dev_free(dev_node_len);
	if (dev_fill) dev_free(dev_fill);
	if (dev_base) dev_free(dev_base);
	if (dev_dst) dev_free(dev_dst);
	if (dev_overflow) dev_free(dev_overflow);
	if (scan_tmp) dev_free(scan_tmp);
	arena.release();
	lists_in.release();
	lists_out.release();
	for (auto& p : out) p.release();
	for (auto& p : fin) p.release();
	return e;
}

}  // namespace tg
