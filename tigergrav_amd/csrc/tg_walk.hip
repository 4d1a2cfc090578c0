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
// Per step: k_walk_count counts every unit's outputs; the host turns the counts into output
// offsets and the next step's units (a few integers per unit); k_walk_write writes.  Each node's
// outputs are recorded as segments in step order, which is the host walk's order, and gathered
// into node-order CSR at the end.  The MAC arithmetic is the host walk's, in FP64 without
// contraction, so every decision is bit-identical.
#include "tg_walk.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace tg {
namespace {

struct WalkNode {
	int32_t pbegin, pend, c0, c1;
	int32_t x[3];
	float r;
	float rmax;  // span of the node's particle box (the sink rmax of the Ewald self test)
};

enum { MODE_FIRST = 0, MODE_DRAIN_D = 1, MODE_DRAIN_E = 2 };

struct WalkUnit {
	int32_t node, mode, len_d, len_e;
	int64_t off_d, off_e;
};

struct WalkParams {
	const WalkNode* nodes;
	const int32_t* lists;  // this step's checklists
	double theta_inv, h;
	float r_ewald;
	int ewald;
};

// output pointers and per-unit offsets (six per unit: FIRST cc_d, cp_d, next_d, cc_e, cp_e,
// next_e; DRAIN pp, pc, next)
struct WalkOut {
	int32_t *cc_d, *cp_d, *cc_e, *cp_e, *pp_d, *pc_d, *pp_e, *pc_e;
	int32_t* next;
	const int64_t* offs;
	const int32_t* counts;
	// capacities of the output pools (cc_d, cp_d, cc_e, cp_e, pc_d, pp_d, pc_e, pp_e) and of next: a
	// write past one is dropped and flagged, never made (the count pass makes the same decisions,
	// so none is expected)
	int64_t cap[9];
	int32_t* overflow;
};

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

// counts of one checklist pass: FIRST {cc, cp ranges, opened, deferred}; DRAIN {pp ranges, pc,
// opened, deferred}
__device__ void count_pass(const WalkParams& P, const WalkNode& n, const double* X, const int32_t* list, int len, bool ewald_kind,
		bool drain, int* c) {
	for (int i = 0; i < len; i++) {
		const int32_t e = list[i];
		const bool opened = e < 0;
		const WalkNode cn = P.nodes[opened ? ~e : e];
		if (cn.pend == cn.pbegin) continue;
		const bool far = far_test(P, n, X, cn, ewald_kind);
		if (drain) {
			if (opened)
				c[0]++;
			else if (far)
				c[1]++;
			else if (cn.c0 < 0)
				c[2]++;
			else
				c[3]++;
		} else {
			if (far) {
				if (opened)
					c[1]++;
				else
					c[0]++;
			} else if (cn.c0 < 0) {
				c[2]++;
			} else {
				c[3]++;
			}
		}
	}
}

__global__ __launch_bounds__(64) void k_walk_count(WalkParams P, const WalkUnit* __restrict__ units, int nunits, int32_t* __restrict__ counts) {
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i >= nunits) return;
	const WalkUnit u = units[i];
	const WalkNode n = P.nodes[u.node];
	const double X[3] = {pos_d(n.x[0]), pos_d(n.x[1]), pos_d(n.x[2])};
	int c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (u.mode == MODE_FIRST) {
		count_pass(P, n, X, P.lists + u.off_d, u.len_d, false, false, c);
		if (P.ewald) count_pass(P, n, X, P.lists + u.off_e, u.len_e, true, false, c + 4);
	} else {
		count_pass(P, n, X, P.lists + u.off_d, u.len_d, u.mode == MODE_DRAIN_E, true, c);
	}
	for (int k = 0; k < 8; k++) counts[8 * int64_t(i) + k] = c[k];
}

__device__ __forceinline__ void put_at(int32_t* p, int64_t i, int64_t cap, int32_t v, int32_t* overflow) {
	if (i < cap)
		p[i] = v;
	else
		*overflow = 1;
}

// one checklist pass, written: FIRST into (a: cc, b: cp) and next; DRAIN into (a: pp, b: pc) and next
__device__ void write_pass(const WalkParams& P, const WalkNode& n, const double* X, const int32_t* list, int len, bool ewald_kind,
		bool drain, int32_t* a, int64_t oa, int64_t capa, int32_t* b, int64_t ob, int64_t capb, int32_t* next, int64_t on, int64_t capn,
		int n_opened, int32_t* overflow) {
	int64_t op = on, od = on + n_opened;
	for (int i = 0; i < len; i++) {
		const int32_t e = list[i];
		const bool opened = e < 0;
		const int32_t id = opened ? ~e : e;
		const WalkNode cn = P.nodes[id];
		if (cn.pend == cn.pbegin) continue;
		const bool far = far_test(P, n, X, cn, ewald_kind);
		if (drain) {
			if (opened) {
				put_at(a, oa++, capa, cn.pbegin, overflow);
				put_at(a, oa++, capa, cn.pend, overflow);
			} else if (far) {
				put_at(b, ob++, capb, id, overflow);
			} else if (cn.c0 < 0) {
				put_at(next, op++, capn, ~id, overflow);
			} else {
				put_at(next, od++, capn, cn.c0, overflow);
				put_at(next, od++, capn, cn.c1, overflow);
			}
		} else {
			if (far) {
				if (opened) {
					put_at(b, ob++, capb, cn.pbegin, overflow);
					put_at(b, ob++, capb, cn.pend, overflow);
				} else {
					put_at(a, oa++, capa, id, overflow);
				}
			} else if (cn.c0 < 0) {
				put_at(next, op++, capn, ~id, overflow);
			} else {
				put_at(next, od++, capn, cn.c0, overflow);
				put_at(next, od++, capn, cn.c1, overflow);
			}
		}
	}
}

__global__ __launch_bounds__(64) void k_walk_write(WalkParams P, const WalkUnit* __restrict__ units, int nunits, WalkOut O) {
	const int i = blockIdx.x * 64 + threadIdx.x;
	if (i >= nunits) return;
	const WalkUnit u = units[i];
	const WalkNode n = P.nodes[u.node];
	const double X[3] = {pos_d(n.x[0]), pos_d(n.x[1]), pos_d(n.x[2])};
	const int64_t* o = O.offs + 6 * int64_t(i);
	const int32_t* c = O.counts + 8 * int64_t(i);
	if (u.mode == MODE_FIRST) {
		write_pass(P, n, X, P.lists + u.off_d, u.len_d, false, false, O.cc_d, o[0], O.cap[0], O.cp_d, o[1], O.cap[1], O.next, o[2], O.cap[8], c[2],
				O.overflow);
		if (P.ewald)
			write_pass(P, n, X, P.lists + u.off_e, u.len_e, true, false, O.cc_e, o[3], O.cap[2], O.cp_e, o[4], O.cap[3], O.next, o[5], O.cap[8],
					c[6], O.overflow);
	} else if (u.mode == MODE_DRAIN_D) {
		write_pass(P, n, X, P.lists + u.off_d, u.len_d, false, true, O.pp_d, o[0], O.cap[5], O.pc_d, o[1], O.cap[4], O.next, o[2], O.cap[8], c[2],
				O.overflow);
	} else {
		write_pass(P, n, X, P.lists + u.off_d, u.len_d, true, true, O.pp_e, o[0], O.cap[7], O.pc_e, o[1], O.cap[6], O.next, o[2], O.cap[8], c[2],
				O.overflow);
	}
}

// dst[d + l] = src[s + l] for every run (s, d, len)
struct CopyRun {
	int64_t src, dst, len;
};
__global__ __launch_bounds__(64) void k_gather_runs(const CopyRun* __restrict__ runs, int64_t nruns, const int32_t* __restrict__ src,
		int32_t* __restrict__ dst) {
	const int64_t r0 = int64_t(blockIdx.x) * 64;
	const int64_t r1 = r0 + 64 < nruns ? r0 + 64 : nruns;
	for (int64_t r = r0; r < r1; r++) {
		const CopyRun g = runs[r];
		for (int64_t l = threadIdx.x; l < g.len; l += 64) dst[g.dst + l] = src[g.src + l];
	}
}

// a device buffer of int32 that grows, keeping its contents
struct Pool {
	int32_t* p = nullptr;
	int64_t cap = 0;
	hipError_t reserve(int64_t n, hipStream_t s) {
		if (n <= cap) return hipSuccess;
		int64_t nc = std::max<int64_t>(n, cap * 2);
		nc = std::max<int64_t>(nc, 1024);
		int32_t* q = nullptr;
		hipError_t e = hipMalloc(&q, size_t(nc) * sizeof(int32_t));
		if (e != hipSuccess) return e;
		if (p && cap > 0) {
			if ((e = hipMemcpyAsync(q, p, size_t(cap) * sizeof(int32_t), hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
			if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
			(void) hipFree(p);
		}
		p = q;
		cap = nc;
		return hipSuccess;
	}
	void release() {
		if (p) (void) hipFree(p);
		p = nullptr;
		cap = 0;
	}
};

struct Seg {
	int32_t node;
	int64_t off, len;
};

enum { K_CC_D, K_CP_D, K_CC_E, K_CP_E, K_PC_D, K_PP_D, K_PC_E, K_PP_E, K_N };

}  // namespace

hipError_t device_walk(Tree& T, int device, hipStream_t s, std::string& err) {
	const int64_t nn = int64_t(T.nodes.size());
	hipError_t e = hipSuccess;
	void* dev_nodes = nullptr;
	void* dev_units = nullptr;
	void* dev_counts = nullptr;
	void* dev_offs = nullptr;
	int32_t* dev_overflow = nullptr;
	int32_t overflow = 0;
	int64_t units_cap = 0;
	Pool lists_in, lists_out, out[K_N];
	std::vector<Seg> segs[K_N];
	int64_t tot[K_N] = {};
	std::vector<WalkUnit> units, next_units;
	std::vector<int32_t> counts;
	std::vector<int64_t> offs;
	WalkParams P{};
#define WCHK(call)                                     \
	do {                                               \
		e = (call);                                    \
		if (e != hipSuccess) {                         \
			err = std::string(#call) + ": " + hipGetErrorString(e); \
			goto done;                                 \
		}                                              \
	} while (0)
	WCHK(hipSetDevice(device));
	{
		const std::vector<float> span = T.sink_spans();
		std::vector<WalkNode> wn(size_t(std::max<int64_t>(nn, 1)));
		for (int64_t i = 0; i < nn; i++) {
			const Node& nd = T.nodes[size_t(i)];
			WalkNode& w = wn[size_t(i)];
			w.pbegin = nd.pbegin;
			w.pend = nd.pend;
			w.c0 = nd.child[0];
			w.c1 = nd.child[1];
			for (int d = 0; d < 3; d++) w.x[d] = nd.x[d];
			w.r = nd.r;
			w.rmax = span[size_t(i)];
		}
		WCHK(hipMalloc(&dev_nodes, wn.size() * sizeof(WalkNode)));
		WCHK(hipMemcpyAsync(dev_nodes, wn.data(), wn.size() * sizeof(WalkNode), hipMemcpyHostToDevice, s));
	}
	WCHK(hipMalloc(&dev_overflow, sizeof(int32_t)));
	WCHK(hipMemsetAsync(dev_overflow, 0, sizeof(int32_t), s));
	P.nodes = static_cast<const WalkNode*>(dev_nodes);
	P.theta_inv = 1.0 / T.P.theta;
	P.h = T.P.soft_len;
	P.r_ewald = float(2.0 * T.P.theta * std::pow(1.0e-3 / 8.0, 1.0 / 3.0));  // src/tree.cpp:403-404
	P.ewald = T.P.ewald ? 1 : 0;
	if (nn > 0 && T.nodes[0].pend > T.nodes[0].pbegin) {
		// the root's checklists: {root} for both kinds
		const int32_t root[2] = {0, 0};
		WCHK(lists_in.reserve(2, s));
		WCHK(hipMemcpyAsync(lists_in.p, root, sizeof(root), hipMemcpyHostToDevice, s));
		units.push_back(WalkUnit{0, MODE_FIRST, 1, P.ewald ? 1 : 0, 0, 1});
	}
	while (!units.empty()) {
		const int nu = int(units.size());
		if (nu > units_cap) {
			if (dev_units) (void) hipFree(dev_units);
			if (dev_counts) (void) hipFree(dev_counts);
			if (dev_offs) (void) hipFree(dev_offs);
			dev_units = dev_counts = dev_offs = nullptr;
			units_cap = std::max<int64_t>(nu, 2 * units_cap);
			WCHK(hipMalloc(&dev_units, size_t(units_cap) * sizeof(WalkUnit)));
			WCHK(hipMalloc(&dev_counts, size_t(units_cap) * 8 * sizeof(int32_t)));
			WCHK(hipMalloc(&dev_offs, size_t(units_cap) * 6 * sizeof(int64_t)));
		}
		WCHK(hipMemcpyAsync(dev_units, units.data(), size_t(nu) * sizeof(WalkUnit), hipMemcpyHostToDevice, s));
		P.lists = lists_in.p;
		hipLaunchKernelGGL(k_walk_count, dim3(unsigned((nu + 63) / 64)), dim3(64), 0, s, P, static_cast<const WalkUnit*>(dev_units), nu,
				static_cast<int32_t*>(dev_counts));
		WCHK(hipGetLastError());
		counts.resize(size_t(nu) * 8);
		WCHK(hipMemcpyAsync(counts.data(), dev_counts, counts.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
		WCHK(hipStreamSynchronize(s));
		// offsets, segments and the next step's units
		offs.assign(size_t(nu) * 6, 0);
		next_units.clear();
		int64_t next_len = 0;
		for (int i = 0; i < nu; i++) {
			const WalkUnit& u = units[size_t(i)];
			const int32_t* c = &counts[size_t(i) * 8];
			int64_t* o = &offs[size_t(i) * 6];
			auto put = [&](int kind, int64_t len) {
				const int64_t off = tot[kind];
				if (len > 0) segs[kind].push_back(Seg{u.node, off, len});
				tot[kind] += len;
				return off;
			};
			const Node& nd = T.nodes[size_t(u.node)];
			if (u.mode == MODE_FIRST) {
				o[0] = put(K_CC_D, c[0]);
				o[1] = put(K_CP_D, 2 * int64_t(c[1]));
				const int64_t ld = int64_t(c[2]) + 2 * int64_t(c[3]);
				o[2] = next_len;
				next_len += ld;
				int64_t le = 0;
				if (P.ewald) {
					o[3] = put(K_CC_E, c[4]);
					o[4] = put(K_CP_E, 2 * int64_t(c[5]));
					le = int64_t(c[6]) + 2 * int64_t(c[7]);
					o[5] = next_len;
					next_len += le;
				}
				if (!nd.leaf()) {
					for (int k = 0; k < 2; k++) {
						const int ch = nd.child[k];
						if (T.nodes[size_t(ch)].empty()) continue;
						next_units.push_back(WalkUnit{ch, MODE_FIRST, int32_t(ld), int32_t(le), o[2], o[5]});
					}
				} else {
					if (ld > 0) next_units.push_back(WalkUnit{u.node, MODE_DRAIN_D, int32_t(ld), 0, o[2], 0});
					if (le > 0) next_units.push_back(WalkUnit{u.node, MODE_DRAIN_E, int32_t(le), 0, o[5], 0});
				}
			} else {
				const bool ew = u.mode == MODE_DRAIN_E;
				o[0] = put(ew ? K_PP_E : K_PP_D, 2 * int64_t(c[0]));
				o[1] = put(ew ? K_PC_E : K_PC_D, c[1]);
				const int64_t l = int64_t(c[2]) + 2 * int64_t(c[3]);
				o[2] = next_len;
				next_len += l;
				if (l > 0) next_units.push_back(WalkUnit{u.node, u.mode, int32_t(l), 0, o[2], 0});
			}
		}
		if (next_len > INT32_MAX) {
			err = "walk checklists exceed 2^31 entries";
			e = hipErrorInvalidValue;
			goto done;
		}
		for (int k = 0; k < K_N; k++) WCHK(out[k].reserve(std::max<int64_t>(tot[k], 1), s));
		WCHK(lists_out.reserve(std::max<int64_t>(next_len, 1), s));
		WCHK(hipMemcpyAsync(dev_offs, offs.data(), offs.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
		{
			WalkOut O{out[K_CC_D].p, out[K_CP_D].p, out[K_CC_E].p, out[K_CP_E].p, out[K_PP_D].p, out[K_PC_D].p, out[K_PP_E].p, out[K_PC_E].p,
					lists_out.p, static_cast<const int64_t*>(dev_offs), static_cast<const int32_t*>(dev_counts),
					{out[K_CC_D].cap, out[K_CP_D].cap, out[K_CC_E].cap, out[K_CP_E].cap, out[K_PC_D].cap, out[K_PP_D].cap, out[K_PC_E].cap,
							out[K_PP_E].cap, lists_out.cap},
					dev_overflow};
			hipLaunchKernelGGL(k_walk_write, dim3(unsigned((nu + 63) / 64)), dim3(64), 0, s, P, static_cast<const WalkUnit*>(dev_units), nu, O);
			WCHK(hipGetLastError());
		}
		std::swap(lists_in, lists_out);
		units.swap(next_units);
		// the units and offsets arrays are rewritten next step: wait for this step's kernels first
		WCHK(hipMemcpyAsync(&overflow, dev_overflow, sizeof(int32_t), hipMemcpyDeviceToHost, s));
		WCHK(hipStreamSynchronize(s));
		if (overflow) {
			err = "device walk: a write pass disagreed with its count pass";
			e = hipErrorUnknown;
			goto done;
		}
	}
	{
		// node-order CSR per kind: each node's segments in step order
		Lists& Ls = T.lists;
		Ls = Lists();
		std::vector<int64_t>* ptrs[K_N] = {&Ls.cc_d_ptr, &Ls.cp_d_ptr, &Ls.cc_e_ptr, &Ls.cp_e_ptr,
				&Ls.pc_d_ptr, &Ls.pp_d_ptr, &Ls.pc_e_ptr, &Ls.pp_e_ptr};
		std::vector<int32_t>* vals[K_N] = {&Ls.cc_d, &Ls.cp_d, &Ls.cc_e, &Ls.cp_e, &Ls.pc_d, &Ls.pp_d, &Ls.pc_e, &Ls.pp_e};
		std::vector<int64_t> fill(size_t(nn), 0);
		for (int k = 0; k < K_N; k++) {
			std::vector<int64_t>& ptr = *ptrs[k];
			ptr.assign(size_t(nn) + 1, 0);
			for (const Seg& g : segs[k]) ptr[size_t(g.node) + 1] += g.len;
			for (int64_t i = 0; i < nn; i++) ptr[size_t(i) + 1] += ptr[size_t(i)];
			vals[k]->assign(size_t(tot[k]), 0);
			if (tot[k] == 0) continue;
			std::fill(fill.begin(), fill.end(), 0);
			std::vector<CopyRun> runs;
			runs.reserve(segs[k].size());
			for (const Seg& g : segs[k]) {
				runs.push_back(CopyRun{g.off, ptr[size_t(g.node)] + fill[size_t(g.node)], g.len});
				fill[size_t(g.node)] += g.len;
			}
			Pool dst, rt;
			e = dst.reserve(tot[k], s);
			if (e == hipSuccess) e = rt.reserve(int64_t(runs.size() * sizeof(CopyRun) / sizeof(int32_t)) + 1, s);
			if (e == hipSuccess) e = hipMemcpyAsync(rt.p, runs.data(), runs.size() * sizeof(CopyRun), hipMemcpyHostToDevice, s);
			if (e == hipSuccess) {
				hipLaunchKernelGGL(k_gather_runs, dim3(unsigned((runs.size() + 63) / 64)), dim3(64), 0, s,
						reinterpret_cast<const CopyRun*>(rt.p), int64_t(runs.size()), out[k].p, dst.p);
				e = hipGetLastError();
			}
			if (e == hipSuccess) e = hipMemcpyAsync(vals[k]->data(), dst.p, size_t(tot[k]) * sizeof(int32_t), hipMemcpyDeviceToHost, s);
			if (e == hipSuccess) e = hipStreamSynchronize(s);
			dst.release();
			rt.release();
			if (e != hipSuccess) {
				err = std::string("list gather: ") + hipGetErrorString(e);
				goto done;
			}
		}
		T.count_census();
	}
done:
#undef WCHK
	(void) hipStreamSynchronize(s);
	if (dev_nodes) (void) hipFree(dev_nodes);
	if (dev_units) (void) hipFree(dev_units);
	if (dev_counts) (void) hipFree(dev_counts);
	if (dev_offs) (void) hipFree(dev_offs);
	if (dev_overflow) (void) hipFree(dev_overflow);
	lists_in.release();
	lists_out.release();
	for (auto& p : out) p.release();
	return e;
}

}  // namespace tg
