// tg_fof.hip -- friends-of-friends groups on the device (tree::find_groups, src/tree.cpp:637-760,
// with part_vect_find_groups, src/part_vect.cpp:84-120).
//
// The reference links particles i, j when the squared int32 wrap-around separation, taken in
// uint64, is below L^2 and not zero (L = N^(-1/3) link_len (2^32 - 1), truncated to an integer),
// and relabels every linked particle with the smallest id among its links until nothing changes,
// each particle's first id being its own tree-order index.  At convergence a linked particle
// carries the smallest tree-order index of its connected component; an unlinked one keeps
// DEFAULT_GROUP.  Here the same components come from a lock-free union-find in which the larger
// root is always hooked under the smaller, so every root is its component's smallest index:
//   * candidate leaf pairs: one thread per leaf walks the tree and keeps the leaves whose particle
//     boxes come within L of its own (minimum image), counted, then written;
//   * links: one 64-lane group per candidate pair tests every particle pair with the reference's
//     integer test and unites the linked ones;
//   * labels: the root of every linked particle, DEFAULT_GROUP for the rest.
// The candidate search only has to find a superset of the linked pairs, so its box test is in
// FP64 with a margin; the link test itself is the reference's exact integer one.
//
// Across the periodic boundary: the reference's integer test wraps, but the ranges it searches
// (range_expand / ranges_intersect, src/tree.cpp:660-670, src/range.cpp:80-116) are not
// wrapped, so a pair straddling the box edge is seen only where a kd box reaches across it, and
// then by one side only.  Here such a pair is never a friend (its coordinate difference must not
// wrap), in either boundary mode, and groups are the plain connected components; every pair away
// from the edge is treated as the reference treats it.
#include "tg_fof.hpp"

#include <hipcub/hipcub.hpp>

#include <climits>
#include <cmath>
#include <vector>

namespace tg {

namespace {

struct FofNode {
	int32_t pbegin, pend, c0, c1;
	double mn[3], mx[3];
};

// the smallest distance between two boxes on the unit torus (periodic) or in space, per axis
__device__ __forceinline__ double axis_gap(double amn, double amx, double bmn, double bmx, bool periodic) {
	double g = 0.0;
	const double d1 = bmn - amx, d2 = amn - bmx;
	g = d1 > g ? d1 : g;
	g = d2 > g ? d2 : g;
	if (periodic) {
		// the box shifted by one period either way
		double gp = 0.0, gm = 0.0;
		const double p1 = (bmn + 1.0) - amx, p2 = amn - (bmx + 1.0);
		gp = p1 > gp ? p1 : gp;
		gp = p2 > gp ? p2 : gp;
		const double m1 = (bmn - 1.0) - amx, m2 = amn - (bmx - 1.0);
		gm = m1 > gm ? m1 : gm;
		gm = m2 > gm ? m2 : gm;
		g = gp < g ? gp : g;
		g = gm < g ? gm : g;
	}
	return g;
}

__device__ __forceinline__ bool boxes_near(const FofNode& a, const FofNode& b, double L2, bool periodic) {
	double s = 0.0;
	for (int d = 0; d < 3; d++) {
		const double g = axis_gap(a.mn[d], a.mx[d], b.mn[d], b.mx[d], periodic);
		s += g * g;
	}
	return s <= L2;
}

constexpr int STACK = 256;

// pass 0 counts each leaf's candidates (leaves b >= a whose boxes come within L), pass 1 writes
// them at the leaf's offset
__global__ void k_fof_candidates(const FofNode* __restrict__ nodes, const int32_t* __restrict__ leaves, int nleaves, double L2, int periodic,
		int pass, int64_t* __restrict__ count, const int64_t* __restrict__ offset, int2* __restrict__ pairs, int* __restrict__ overflow) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nleaves) return;
	const int a = leaves[k];
	const FofNode A = nodes[a];
	int64_t c = 0;
	if (A.pend == A.pbegin) {
		if (!pass) count[k] = 0;
		return;
	}
	int64_t o = pass ? offset[k] : 0;
	int stack[STACK];
	int sp = 0;
	stack[sp++] = 0;
	while (sp > 0) {
		const int id = stack[--sp];
		const FofNode B = nodes[id];
		if (B.pend == B.pbegin || !boxes_near(A, B, L2, periodic != 0)) continue;
		if (B.c0 < 0) {
			if (id >= a) {
				if (pass) pairs[o + c] = make_int2(a, id);
				c++;
			}
		} else {
			if (sp + 2 > STACK) {
				atomicOr(overflow, 1);
				continue;
			}
			stack[sp++] = B.c1;
			stack[sp++] = B.c0;
		}
	}
	if (!pass) count[k] = c;
}

// parents only ever decrease (parent[i] <= i, equal at roots), so the walk ends even on stale
// reads; loads are device-coherent so other groups' hooks are seen
__device__ __forceinline__ int ld(const int* a) { return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ int uf_find(int* parent, int x) {
	int p = ld(parent + x);
	while (p != x) {
		const int gp = ld(parent + p);
		if (gp != p) parent[x] = gp;  // path halving: only ever lowers a parent to an ancestor
		x = p;
		p = gp;
	}
	return x;
}

__device__ void uf_union(int* parent, int a, int b) {
	for (;;) {
		a = uf_find(parent, a);
		b = uf_find(parent, b);
		if (a == b) return;
		if (a > b) {
			const int t = a;
			a = b;
			b = t;
		}
		// hook the larger root under the smaller one
		const int old = atomicCAS(&parent[b], b, a);
		if (old == b) return;
		b = old;
	}
}

// every particle pair of each candidate leaf pair, the reference's integer test
__global__ __launch_bounds__(64) void k_fof_link(const FofNode* __restrict__ nodes, const int2* __restrict__ pairs, int64_t npairs,
		const int4* __restrict__ parts, uint64_t L2, int* __restrict__ parent, int* __restrict__ linked) {
	const int64_t q = blockIdx.x;
	if (q >= npairs) return;
	const int2 pr = pairs[q];
	const FofNode A = nodes[pr.x], B = nodes[pr.y];
	const bool same = pr.x == pr.y;
	for (int i = A.pbegin + threadIdx.x; i < A.pend; i += 64) {
		const int4 xi = parts[i];
		for (int j = same ? i + 1 : B.pbegin; j < B.pend; j++) {
			const int4 xj = parts[j];
			// vect<pos_type> difference (int32 wrap), then vect<uint64_t> and its dot product; a
			// difference that wraps crosses the box edge
			const int64_t ex = int64_t(xi.x) - int64_t(xj.x), ey = int64_t(xi.y) - int64_t(xj.y), ez = int64_t(xi.z) - int64_t(xj.z);
			const bool wraps = ex != int64_t(int32_t(ex)) || ey != int64_t(int32_t(ey)) || ez != int64_t(int32_t(ez));
			const uint64_t dx = uint64_t(int64_t(int32_t(uint32_t(xi.x) - uint32_t(xj.x))));
			const uint64_t dy = uint64_t(int64_t(int32_t(uint32_t(xi.y) - uint32_t(xj.y))));
			const uint64_t dz = uint64_t(int64_t(int32_t(uint32_t(xi.z) - uint32_t(xj.z))));
			const uint64_t d2 = dx * dx + dy * dy + dz * dz;
			if (d2 < L2 && d2 != 0 && !wraps) {
				linked[i] = 1;
				linked[j] = 1;
				uf_union(parent, i, j);
			}
		}
	}
}

__global__ void k_fof_init(int* parent, int* linked, int64_t n) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) {
		parent[i] = int(i);
		linked[i] = 0;
	}
}

__global__ void k_fof_labels(int* parent, const int* __restrict__ linked, int64_t n, uint64_t* __restrict__ group) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) group[i] = linked[i] ? uint64_t(uf_find(parent, int(i))) : TG_FOF_NONE;
}

struct Buf {
	void* p = nullptr;
	~Buf() {
		if (p) (void) hipFree(p);
	}
	hipError_t alloc(size_t b) { return hipMalloc(&p, b < 256 ? 256 : b); }
	template <class T>
	T* as() const {
		return static_cast<T*>(p);
	}
};

}  // namespace

hipError_t fof_device(const Tree& T, const int4* d_parts, double link_len, int device, hipStream_t s, uint64_t* group_tree,
		int64_t* n_candidates, std::string& err) {
	const int64_t n = int64_t(T.pos.size() / 3);
	const int64_t nn = int64_t(T.nodes.size());
	hipError_t e = hipSuccess;
#define FCHK(call)                                                 \
	do {                                                           \
		e = (call);                                                \
		if (e != hipSuccess) {                                     \
			err = std::string(#call) + ": " + hipGetErrorString(e); \
			return e;                                              \
		}                                                          \
	} while (0)
	if (n == 0) return hipSuccess;
	FCHK(hipSetDevice(device));
	// the reference's integer linking length (src/part_vect.cpp:86-88)
	const uint64_t L = uint64_t((std::pow(double(n), -1.0 / 3.0) * link_len) * 4294967295.0);
	const uint64_t L2 = L * L;
	// candidate boxes within the linking length, with a margin (the exact test is per pair)
	const double Lb = 1.01 * double(L) / 4294967296.0 + 1e-9;
	std::vector<FofNode> hn(static_cast<size_t>(nn));
	for (int64_t i = 0; i < nn; i++) {
		const Node& d = T.nodes[size_t(i)];
		FofNode& o = hn[size_t(i)];
		o.pbegin = d.pbegin;
		o.pend = d.pend;
		o.c0 = d.child[0];
		o.c1 = d.child[1];
		for (int k = 0; k < 3; k++) {
			o.mn[k] = d.rmin[k];
			o.mx[k] = d.rmax[k];
		}
	}
	const int nleaves = int(T.leaves.size());
	Buf dnodes, dleaves, dcount, doffset, dpairs, dparent, dlinked, dgroup, doverflow, tmp;
	FCHK(dnodes.alloc(hn.size() * sizeof(FofNode)));
	FCHK(dleaves.alloc(size_t(nleaves) * 4));
	FCHK(dcount.alloc(size_t(nleaves + 1) * 8));
	FCHK(doffset.alloc(size_t(nleaves + 1) * 8));
	FCHK(doverflow.alloc(4));
	FCHK(hipMemcpyAsync(dnodes.p, hn.data(), hn.size() * sizeof(FofNode), hipMemcpyHostToDevice, s));
	FCHK(hipMemcpyAsync(dleaves.p, T.leaves.data(), size_t(nleaves) * 4, hipMemcpyHostToDevice, s));
	FCHK(hipMemsetAsync(doverflow.p, 0, 4, s));
	FCHK(hipMemsetAsync(dcount.p, 0, size_t(nleaves + 1) * 8, s));
	const bool periodic = false;  // pairs across the box edge are not friends (above)
	const unsigned lb = unsigned((nleaves + 127) / 128);
	hipLaunchKernelGGL(k_fof_candidates, dim3(lb), dim3(128), 0, s, dnodes.as<const FofNode>(), dleaves.as<const int32_t>(), nleaves, Lb * Lb,
			int(periodic), 0, dcount.as<int64_t>(), nullptr, nullptr, doverflow.as<int>());
	FCHK(hipGetLastError());
	size_t tb = 0;
	FCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, dcount.as<int64_t>(), doffset.as<int64_t>(), nleaves + 1, s));
	FCHK(tmp.alloc(tb));
	FCHK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, dcount.as<int64_t>(), doffset.as<int64_t>(), nleaves + 1, s));
	int64_t npairs = 0;
	int overflow = 0;
	FCHK(hipMemcpyAsync(&npairs, doffset.as<int64_t>() + nleaves, 8, hipMemcpyDeviceToHost, s));
	FCHK(hipMemcpyAsync(&overflow, doverflow.p, 4, hipMemcpyDeviceToHost, s));
	FCHK(hipStreamSynchronize(s));
	if (overflow) {
		err = "friends-of-friends: tree deeper than the candidate search's stack";
		return hipErrorInvalidValue;
	}
	if (npairs > int64_t(INT32_MAX)) {
		err = "friends-of-friends: more candidate leaf pairs than one launch takes";
		return hipErrorInvalidValue;
	}
	FCHK(dpairs.alloc(size_t(npairs) * sizeof(int2)));
	hipLaunchKernelGGL(k_fof_candidates, dim3(lb), dim3(128), 0, s, dnodes.as<const FofNode>(), dleaves.as<const int32_t>(), nleaves, Lb * Lb,
			int(periodic), 1, dcount.as<int64_t>(), doffset.as<const int64_t>(), dpairs.as<int2>(), doverflow.as<int>());
	FCHK(hipGetLastError());
	FCHK(dparent.alloc(size_t(n) * 4));
	FCHK(dlinked.alloc(size_t(n) * 4));
	FCHK(dgroup.alloc(size_t(n) * 8));
	const unsigned pb = unsigned((n + 255) / 256);
	hipLaunchKernelGGL(k_fof_init, dim3(pb), dim3(256), 0, s, dparent.as<int>(), dlinked.as<int>(), n);
	if (npairs > 0)
		hipLaunchKernelGGL(k_fof_link, dim3(unsigned(npairs)), dim3(64), 0, s, dnodes.as<const FofNode>(), dpairs.as<const int2>(), npairs,
				d_parts, L2, dparent.as<int>(), dlinked.as<int>());
	hipLaunchKernelGGL(k_fof_labels, dim3(pb), dim3(256), 0, s, dparent.as<int>(), dlinked.as<const int>(), n, dgroup.as<uint64_t>());
	FCHK(hipGetLastError());
	FCHK(hipMemcpyAsync(group_tree, dgroup.p, size_t(n) * 8, hipMemcpyDeviceToHost, s));
	FCHK(hipStreamSynchronize(s));
	if (n_candidates) *n_candidates = npairs;
	return hipSuccess;
#undef FCHK
}

}  // namespace tg
