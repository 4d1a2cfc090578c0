// tg_devtree.hip -- the kd-tree (tree::refine, src/tree.cpp:122-201) and its multipoles
// (tree::compute_multipoles, src/tree.cpp:203-296) on the device, bit for bit as the host
// restatement in tg_tree.cpp builds them.
//
// Split passes, one per tree level (the reference refines one level per pass, src/main.cpp:170-174):
//   * every node of the level with more than parts_per_node particles (below the depth limit) is
//     split at the midpoint of its box's longest side; its children are numbered after the level,
//     in node order (a scan over the level's split flags);
//   * the reference's in-place partition (src/part_vect.cpp:905-926: scan up from the low end, on
//     a high particle search down from the high end for a low one and swap) exchanges the k-th high
//     particle left of the split point with the k-th low particle right of it counting from the
//     high end, and moves nothing else.  So each particle's partner follows from two prefix counts
//     of the "low" flags over the particle array: every exchange is made at once, each pair by the
//     thread of its left particle.
// Multipole passes, levels from the deepest up: leaves by 16 threads each (the 16 float lanes of
// part_vect_multipole_info, src/part_vect.cpp:567-635, lane k taking particles i + k of each block
// of 16, the lanes summed in order at the end), interior nodes by one thread each.  Every
// floating-point expression keeps the host's operand order and roundings (no contraction).
#include "tg_devtree.hpp"
#include "tg_devmem.hpp"
#include "tg_math.hpp"

#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace tg {

namespace {

struct SplitInfo {
	double mid;
	int32_t dim, first_child, lo_end, pad;  // lo_end: the split point (first particle of the right child)
};

__device__ __forceinline__ double d_pos_to_double(int32_t x) {
#pragma clang fp contract(off)
	return double(x) * (1.0 / 4294967296.0) + 0.5;
}

__device__ __forceinline__ int32_t d_double_to_pos(double x) {
#pragma clang fp contract(off)
	const double POS_MAX = 4294967296.0;
	const double y = x - 0.5;
	long long z = llround(y * POS_MAX);
	if (double(z) >= POS_MAX / 2.0) z -= (long long) POS_MAX;
	return int32_t(z);
}

__device__ __forceinline__ void d_box_range(uint64_t id, double* mn, double* mx) {
#pragma clang fp contract(off)
	double x[3] = {0.0, 0.0, 0.0};
	double dx[3] = {1.0, 1.0, 1.0};
	while (id != 1) {
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

__device__ __forceinline__ double d_vabs(const double* v) {
#pragma clang fp contract(off)
	return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

__global__ void k_iota64(int64_t* order, int64_t n) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) order[i] = i;
}

__global__ void k_fill_i32(int32_t* a, int64_t n, int32_t v) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) a[i] = v;
}

// The level being split is [lev[0], lev[1]) in device memory: the host launches a level's kernels
// with grids sized by a bound on its node count and reads the counts back only every few levels.

// 1 for every node of the level that splits, 0 from nl up to bound (the scan's total at nl)
__global__ void k_split_flags(const DNode* __restrict__ nodes, const int32_t* __restrict__ lev, int ppn, int max_level,
		int32_t* __restrict__ flags, int bound) {
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j > bound) return;
	const int lb = lev[0], nl = lev[1] - lev[0];
	int32_t f = 0;
	if (j < nl) {
		const DNode nd = nodes[lb + j];
		f = (nd.pend - nd.pbegin > ppn && nd.level < max_level) ? 1 : 0;
	}
	flags[j] = f;
}

// the next level: [le, le + 2 nsplit); log = (lb, le, nsplit) of the level just split
__global__ void k_next_level(int32_t* __restrict__ lev, const int32_t* __restrict__ rank, int32_t* __restrict__ log) {
	if (threadIdx.x != 0 || blockIdx.x != 0) return;
	const int32_t lb = lev[0], le = lev[1];
	const int32_t ns = rank[le - lb];
	log[0] = lb;
	log[1] = le;
	log[2] = ns;
	lev[0] = le;
	lev[1] = le + 2 * ns;
}

// the split dimension and midpoint of every splitting node (src/tree.cpp:136-157), its children's
// index
__global__ void k_split_setup(const DNode* __restrict__ nodes, const int32_t* __restrict__ lev, const int32_t* __restrict__ flags,
		const int32_t* __restrict__ rank, SplitInfo* __restrict__ info) {
#pragma clang fp contract(off)
	const int lb = lev[0], le = lev[1], nl = le - lb;
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nl) return;
	SplitInfo si{};
	si.first_child = -1;
	if (flags[j]) {
		double bmin[3], bmax[3];
		d_box_range(nodes[lb + j].boxid, bmin, bmax);
		double max_span = 0.0;
		int max_dim = 0;
		for (int dim = 0; dim < 3; dim++) {
			const double span = bmax[dim] - bmin[dim];
			if (span > max_span) {
				max_span = span;
				max_dim = dim;
			}
		}
		si.mid = (bmax[max_dim] + bmin[max_dim]) * 0.5;
		si.dim = max_dim;
		si.first_child = le + 2 * rank[j];
	}
	info[j] = si;
}

// per particle: 1 if it lies in a splitting node of the level and below that node's midpoint
__global__ void k_low_flags(const int32_t* __restrict__ pos, const int32_t* __restrict__ pnode, int64_t n, const int32_t* __restrict__ lev,
		const SplitInfo* __restrict__ info, int32_t* __restrict__ low) {
	const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (p > n) return;
	const int lb = lev[0], nl = lev[1] - lev[0];
	int32_t f = 0;
	if (p < n) {
		const int j = pnode[p] - lb;
		if (j >= 0 && j < nl) {
			const SplitInfo si = info[j];
			if (si.first_child >= 0) f = d_pos_to_double(pos[3 * p + si.dim]) < si.mid ? 1 : 0;
		}
	}
	low[p] = f;
}

// the children of every splitting node: [pbegin, split) and [split, pend), the split point being
// where the node's low particles end after the partition
__global__ void k_split_nodes(DNode* __restrict__ nodes, const int32_t* __restrict__ lev, SplitInfo* __restrict__ info,
		const int32_t* __restrict__ S) {
	const int lb = lev[0], nl = lev[1] - lev[0];
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nl) return;
	SplitInfo si = info[j];
	if (si.first_child < 0) return;
	DNode nd = nodes[lb + j];
	const int32_t split = nd.pbegin + (S[nd.pend] - S[nd.pbegin]);
	si.lo_end = split;
	info[j] = si;
	DNode l{nd.pbegin, split, -1, -1, lb + j, nd.level + 1, nd.boxid << 1};
	DNode r{split, nd.pend, -1, -1, lb + j, nd.level + 1, (nd.boxid << 1) + 1};
	nodes[si.first_child] = l;
	nodes[si.first_child + 1] = r;
	nd.c0 = si.first_child;
	nd.c1 = si.first_child + 1;
	nodes[lb + j] = nd;
}

// the low particles right of the split point, by rank counted from the node's high end
__global__ void k_high_side_lows(const int32_t* __restrict__ pnode, int64_t n, const int32_t* __restrict__ lev,
		const SplitInfo* __restrict__ info, const DNode* __restrict__ nodes, const int32_t* __restrict__ low, const int32_t* __restrict__ S,
		int32_t* __restrict__ mr) {
	const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (p >= n) return;
	const int lb = lev[0], nl = lev[1] - lev[0];
	const int j = pnode[p] - lb;
	if (j < 0 || j >= nl) return;
	const SplitInfo si = info[j];
	if (si.first_child < 0 || p < si.lo_end || !low[p]) return;
	const int32_t e = nodes[lb + j].pend;
	const int32_t b = nodes[lb + j].pbegin;
	mr[b + (S[e] - S[p] - 1)] = int32_t(p);
}

// every high particle left of the split point exchanges its record with its partner; every
// particle of a splitting node moves to its child
__global__ void k_exchange(int32_t* __restrict__ pos, float* __restrict__ mass, int64_t* __restrict__ order, int32_t* __restrict__ pnode,
		int64_t n, const int32_t* __restrict__ lev, const SplitInfo* __restrict__ info, const DNode* __restrict__ nodes,
		const int32_t* __restrict__ low, const int32_t* __restrict__ S, const int32_t* __restrict__ mr) {
	const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (p >= n) return;
	const int lb = lev[0], nl = lev[1] - lev[0];
	const int j = pnode[p] - lb;
	if (j < 0 || j >= nl) return;
	const SplitInfo si = info[j];
	if (si.first_child < 0) return;
	if (p < si.lo_end && !low[p]) {
		const int32_t b = nodes[lb + j].pbegin;
		const int64_t q = mr[b + ((int32_t(p) - b) - (S[p] - S[b]))];
		for (int d = 0; d < 3; d++) {
			const int32_t t = pos[3 * p + d];
			pos[3 * p + d] = pos[3 * q + d];
			pos[3 * q + d] = t;
		}
		const float tm = mass[p];
		mass[p] = mass[q];
		mass[q] = tm;
		const int64_t to = order[p];
		order[p] = order[q];
		order[q] = to;
	}
	pnode[p] = p < si.lo_end ? si.first_child : si.first_child + 1;
}

// ---- multipoles ----------------------------------------------------------------------------------
// a node without particles (src/tree.cpp:214-227)
__device__ void empty_node(const DNode& nd, DMulti& M) {
#pragma clang fp contract(off)
	double bmin[3], bmax[3];
	d_box_range(nd.boxid, bmin, bmax);
	M.N = 0;
	for (int q = 0; q < 17; q++) M.m[q] = 0.0f;
	for (int d = 0; d < 3; d++) {
		M.x[d] = d_double_to_pos(0.5 * (bmax[d] - bmin[d]));
		M.rmin[d] = bmin[d];
		M.rmax[d] = bmax[d];
	}
	M.r = 0.0f;
}

// leaves of the level: 16 threads per leaf, four leaves per work group
__global__ __launch_bounds__(64) void k_leaf_moments(const DNode* __restrict__ nodes, int lb, int nl, const int32_t* __restrict__ pos,
		const float* __restrict__ mass, DMulti* __restrict__ out) {
#pragma clang fp contract(off)
	__shared__ float red[4][17][16];
	__shared__ float rr[4][16];
	__shared__ double bb[4][16][6];
	const int slot = threadIdx.x >> 4;
	const int k = threadIdx.x & 15;
	const int j = blockIdx.x * 4 + slot;
	const bool active = j < nl && nodes[lb + j].c0 < 0;
	DNode nd{};
	if (active) nd = nodes[lb + j];
	const bool nonempty = active && nd.pend > nd.pbegin;
	float M[17];
	for (int q = 0; q < 17; q++) M[q] = 0.0f;
	double Xcom[3] = {0.0, 0.0, 0.0};
	float r = 0.0f;
	double mn[3] = {1.0, 1.0, 1.0}, mx[3] = {0.0, 0.0, 0.0};
	int32_t cx[3] = {0, 0, 0};
	if (nonempty) {
		// part_vect_center_of_mass (src/part_vect.cpp:542-565): FP64 sum in particle order
		double com[3] = {0.0, 0.0, 0.0};
		for (int64_t i = nd.pbegin; i < nd.pend; i++)
			for (int d = 0; d < 3; d++) com[d] += d_pos_to_double(pos[3 * i + d]);
		const double cnt = double(nd.pend - nd.pbegin);
		for (int d = 0; d < 3; d++) {
			cx[d] = d_double_to_pos(com[d] / cnt);
			Xcom[d] = d_pos_to_double(cx[d]);
		}
		// lane k: particles i + k of each block of 16 (padding past the end: position 0, mass 0)
		for (int64_t i = nd.pbegin; i < nd.pend; i += 16) {
			double X[3];
			float ms;
			if (i + k < nd.pend) {
				for (int d = 0; d < 3; d++) X[d] = d_pos_to_double(pos[3 * (i + k) + d]);
				ms = mass[i + k];
			} else {
				X[0] = X[1] = X[2] = 0.0;
				ms = 0.0f;
			}
			float dx[3];
			for (int d = 0; d < 3; d++) dx[d] = float(X[d] - Xcom[d]);
			M[0] += ms;
#pragma unroll
			for (int a = 0; a < 3; a++)
#pragma unroll
				for (int b = 0; b <= a; b++) {
					const float Xab = dx[a] * dx[b];
					M[M2(a, b)] += ms * Xab;
#pragma unroll
					for (int l = 0; l <= b; l++) M[M3(a, b, l)] += ms * Xab * dx[l];
				}
		}
		// radius and particle box of this lane's particles (max / min are exact in any order)
		for (int64_t i = nd.pbegin + k; i < nd.pend; i += 16) {
			double v[3];
			for (int d = 0; d < 3; d++) {
				const double xd = d_pos_to_double(pos[3 * i + d]);
				v[d] = xd - Xcom[d];
				mx[d] = mx[d] < xd ? xd : mx[d];
				mn[d] = xd < mn[d] ? xd : mn[d];
			}
			const float a = float(d_vabs(v));
			r = r < a ? a : r;
		}
	}
	for (int q = 0; q < 17; q++) red[slot][q][k] = M[q];
	rr[slot][k] = r;
	for (int d = 0; d < 3; d++) {
		bb[slot][k][d] = mn[d];
		bb[slot][k][3 + d] = mx[d];
	}
	__syncthreads();
	if (!active) return;
	DMulti& o = out[lb + j];
	if (!nonempty) {
		if (k == 0) empty_node(nd, o);
		return;
	}
	// lanes summed in order, 0..15 (slot q by thread q, slot 16 by thread 0)
	for (int q = k; q < 17; q += 16) {
		float s = 0.0f;
		for (int l = 0; l < 16; l++) s += red[slot][q][l];
		o.m[q] = s;
	}
	if (k == 0) {
		float rm = 0.0f;
		double lo[3] = {1.0, 1.0, 1.0}, hi[3] = {0.0, 0.0, 0.0};
		for (int l = 0; l < 16; l++) {
			rm = rm < rr[slot][l] ? rr[slot][l] : rm;
			for (int d = 0; d < 3; d++) {
				lo[d] = bb[slot][l][d] < lo[d] ? bb[slot][l][d] : lo[d];
				hi[d] = hi[d] < bb[slot][l][3 + d] ? bb[slot][l][3 + d] : hi[d];
			}
		}
		o.N = uint64_t(nd.pend - nd.pbegin);
		o.r = rm;
		for (int d = 0; d < 3; d++) {
			o.x[d] = cx[d];
			o.rmin[d] = lo[d];
			o.rmax[d] = hi[d];
		}
	}
}

// multipole<float>::operator>>= (tigergrav/multipole.hpp:414-433)
__device__ __forceinline__ void d_m2m_shift(float* me, const float* Y) {
#pragma clang fp contract(off)
#pragma unroll
	for (int p = 0; p < 3; p++)
#pragma unroll
		for (int q = p; q < 3; q++)
#pragma unroll
			for (int l = q; l < 3; l++) {
				float t = me[M3(p, q, l)];
				t += me[0] * Y[p] * Y[q] * Y[l];
				t += me[M2(p, q)] * Y[l];
				t += me[M2(q, l)] * Y[p];
				t += me[M2(l, p)] * Y[q];
				me[M3(p, q, l)] = t;
			}
#pragma unroll
	for (int p = 0; p < 3; p++)
#pragma unroll
		for (int q = p; q < 3; q++) me[M2(p, q)] += me[0] * Y[p] * Y[q];
}

// interior nodes of the level (src/tree.cpp:235-288), one thread each
__global__ void k_interior_moments(const DNode* __restrict__ nodes, int lb, int nl, DMulti* __restrict__ out) {
#pragma clang fp contract(off)
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= nl) return;
	const DNode nd = nodes[lb + j];
	if (nd.c0 < 0) return;
	DMulti o{};
	if (nd.pend == nd.pbegin) {
		empty_node(nd, o);
		out[lb + j] = o;
		return;
	}
	const DMulti L = out[nd.c0];
	const DMulti R = out[nd.c1];
	o.N = L.N + R.N;
	double lx[3], rx[3], x[3];
	for (int d = 0; d < 3; d++) {
		lx[d] = d_pos_to_double(L.x[d]);
		rx[d] = d_pos_to_double(R.x[d]);
		o.x[d] = d_double_to_pos((lx[d] * double(L.N) + rx[d] * double(R.N)) / double(o.N));
		x[d] = d_pos_to_double(o.x[d]);
	}
	float ml[17], mr[17], yl[3], yr[3];
	for (int q = 0; q < 17; q++) {
		ml[q] = L.m[q];
		mr[q] = R.m[q];
	}
	for (int d = 0; d < 3; d++) {
		yl[d] = float(lx[d] - x[d]);
		yr[d] = float(rx[d] - x[d]);
	}
	d_m2m_shift(ml, yl);
	d_m2m_shift(mr, yr);
	for (int q = 0; q < 17; q++) o.m[q] = ml[q] + mr[q];
	double dl[3], dr[3];
	for (int d = 0; d < 3; d++) {
		dl[d] = lx[d] - x[d];
		dr[d] = rx[d] - x[d];
	}
	if (L.N == 0) {
		o.r = R.r;
	} else if (R.N == 0) {
		o.r = L.r;
	} else {
		const double a = d_vabs(dl) + double(L.r), b = d_vabs(dr) + double(R.r);
		o.r = float(a < b ? b : a);
	}
	for (int d = 0; d < 3; d++) {
		o.rmax[d] = L.rmax[d] < R.rmax[d] ? R.rmax[d] : L.rmax[d];
		o.rmin[d] = R.rmin[d] < L.rmin[d] ? R.rmin[d] : L.rmin[d];
	}
	double rmax = 0.0;
	for (int c = 0; c < 8; c++) {  // the eight corners, in the reference's order
		const double corner[3] = {(c & 1) ? o.rmax[0] : o.rmin[0], (c & 2) ? o.rmax[1] : o.rmin[1], (c & 4) ? o.rmax[2] : o.rmin[2]};
		double v[3];
		for (int d = 0; d < 3; d++) v[d] = x[d] - corner[d];
		const double a = d_vabs(v);
		rmax = c == 0 ? a : (rmax < a ? a : rmax);
	}
	const float rf = float(rmax);
	o.r = rf < o.r ? rf : o.r;
	out[lb + j] = o;
}

unsigned blocks(int64_t n, int b) { return unsigned((n + b - 1) / b); }

// pinned host staging for the copy back (full-rate transfers), grown on demand and kept for the
// process; one build at a time uses it
std::mutex pinned_mtx;
void* pinned_buf = nullptr;
size_t pinned_bytes = 0;

hipError_t pinned_staging(size_t bytes, void** out) {
	if (bytes > pinned_bytes) {
		if (pinned_buf) (void) hipHostFree(pinned_buf);
		pinned_buf = nullptr;
		pinned_bytes = 0;
		const size_t want = bytes + bytes / 4;
		const hipError_t e = hipHostMalloc(&pinned_buf, want, hipHostMallocDefault);
		if (e != hipSuccess) return e;
		pinned_bytes = want;
	}
	*out = pinned_buf;
	return hipSuccess;
}

double wall() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// device array freed with its owner
struct Buf {
	void* p = nullptr;
	size_t bytes = 0;
	~Buf() {
		if (p) dev_free(p);
	}
	hipError_t alloc(size_t b) {
		if (b <= bytes) return hipSuccess;
		if (p) dev_free(p);
		p = nullptr;
		bytes = 0;
		const hipError_t e = dev_malloc(&p, b < 256 ? 256 : b);
		if (e == hipSuccess) bytes = b < 256 ? 256 : b;
		return e;
	}
	template <class T>
	T* as() const {
		return static_cast<T*>(p);
	}
};

}  // namespace

void DeviceTree::release() {
	if (pos) dev_free(pos);
	if (mass) dev_free(mass);
	if (order) dev_free(order);
	if (nodes) dev_free(nodes);
	if (multi) dev_free(multi);
	*this = DeviceTree();
}

hipError_t device_build_tree(Tree& T, const int32_t* x, const float* masses, int64_t n, const TreeParams& p,
		double particle_mass, int device, hipStream_t s, std::string& err, double* t_out, bool on_device, DeviceTree* keep) {
	const double t0 = wall();
	hipError_t e = hipSuccess;
#define DTCHK(call)                                                    \
	do {                                                               \
		e = (call);                                                    \
		if (e != hipSuccess) {                                         \
			err = std::string(#call) + ": " + hipGetErrorString(e);    \
			return e;                                                  \
		}                                                              \
	} while (0)
#define DTLAUNCH(name)                                                                 \
	do {                                                                               \
		e = hipGetLastError();                                                         \
		if (e != hipSuccess) {                                                         \
			err = std::string("device tree build: launch ") + name + ": " + hipGetErrorString(e); \
			return e;                                                                  \
		}                                                                              \
	} while (0)
	(void) hipGetLastError();  // an error left by an earlier call is not the build's
	if (n > int64_t(INT32_MAX) - 1) {
		err = "device tree build: more than 2^31 - 2 particles";
		return hipErrorInvalidValue;
	}
	DTCHK(hipSetDevice(device));
	T.P = p;
	T.m = particle_mass;
	T.lists = Lists();
	T.census = Census();
	Buf dpos, dmass, dorder, dpnode, dlow, dS, dmr, dnodes, dinfo, dflags, drank, dmulti, scan_tmp;
	const int64_t nn_cap0 = 4 * (n / std::max(1, p.parts_per_node)) + 16;
	int64_t node_cap = nn_cap0;
	DTCHK(dpos.alloc(size_t(3 * n) * 4));
	DTCHK(dmass.alloc(size_t(n) * 4));
	DTCHK(dorder.alloc(size_t(n) * 8));
	DTCHK(dpnode.alloc(size_t(n) * 4));
	DTCHK(dlow.alloc(size_t(n + 1) * 4));
	DTCHK(dS.alloc(size_t(n + 1) * 4));
	DTCHK(dmr.alloc(size_t(n) * 4));
	DTCHK(dnodes.alloc(size_t(node_cap) * sizeof(DNode)));
	if (n > 0) {
		const hipMemcpyKind in = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
		DTCHK(hipMemcpyAsync(dpos.p, x, size_t(3 * n) * 4, in, s));
		if (masses) {
			DTCHK(hipMemcpyAsync(dmass.p, masses, size_t(n) * 4, in, s));
		} else {
			std::vector<float> mm(static_cast<size_t>(n), static_cast<float>(particle_mass));
			DTCHK(hipMemcpyAsync(dmass.p, mm.data(), size_t(n) * 4, hipMemcpyHostToDevice, s));
			DTCHK(hipStreamSynchronize(s));
		}
		hipLaunchKernelGGL(k_iota64, dim3(blocks(n, 256)), dim3(256), 0, s, dorder.as<int64_t>(), n);
		hipLaunchKernelGGL(k_fill_i32, dim3(blocks(n, 256)), dim3(256), 0, s, dpnode.as<int32_t>(), n, 0);
		DTLAUNCH("k_iota64/k_fill_i32");
	}
	{
		const DNode root{0, int32_t(n), -1, -1, -1, 0, 1};
		DTCHK(hipMemcpyAsync(dnodes.p, &root, sizeof(root), hipMemcpyHostToDevice, s));
	}
	// Level by level, a batch of DT_BATCH levels at a time without a host wait: each level's kernels
	// read its extent from dlev and a last kernel moves dlev on and logs (lb, le, nsplit); grids are
	// sized by bounds (a level's splitting nodes hold more than ppn particles each, so there are at
	// most split_cap of them, and a level has at most twice the previous level's nodes).  After a
	// batch the host reads the log, stops at the first level that split nothing (the levels after
	// it in the batch found no nodes and did nothing), and grows the node array for the next batch.
	constexpr int DT_BATCH = 4;
	const int64_t split_cap = n / (int64_t(std::max(1, p.parts_per_node)) + 1) + 1;
	const int64_t lvl_max = 2 * split_cap;
	const int64_t fl_cap = lvl_max + 1;
	Buf dlev, dlog;
	DTCHK(dflags.alloc(size_t(fl_cap + 1) * 4));
	DTCHK(drank.alloc(size_t(fl_cap + 1) * 4));
	DTCHK(dinfo.alloc(size_t(fl_cap) * sizeof(SplitInfo)));
	DTCHK(dlev.alloc(2 * sizeof(int32_t)));
	const int log_cap = MAX_SPLIT_LEVEL + 2 + DT_BATCH;
	DTCHK(dlog.alloc(size_t(3 * log_cap) * sizeof(int32_t)));
	{
		const int32_t lev0[2] = {0, 1};
		DTCHK(hipMemcpyAsync(dlev.p, lev0, sizeof(lev0), hipMemcpyHostToDevice, s));
	}
	// the scans' temporary storage: the particle scan and the largest level scan
	size_t scan_bytes = 0, level_scan_bytes = 0;
	DTCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, dlow.as<int32_t>(), dS.as<int32_t>(), int(n + 1), s));
	DTCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, level_scan_bytes, dflags.as<int32_t>(), drank.as<int32_t>(), int(fl_cap), s));
	DTCHK(scan_tmp.alloc(std::max(scan_bytes, level_scan_bytes)));
	std::vector<int> level_begin;
	int64_t le_h = 1;      // nodes made, as of the host's last look
	int64_t nl_bound = 1;  // the current level's nodes, at most
	int level = 0;
	int le = 1;
	bool finished = false;
	while (!finished) {
		if (level + DT_BATCH > log_cap) {
			err = "device tree build: more levels than the split limit allows";
			return hipErrorInvalidValue;
		}
		// node capacity for this batch
		int64_t add = 0;
		for (int64_t i = 0, b = nl_bound; i < DT_BATCH; i++, b = std::min(2 * b, lvl_max)) add += 2 * std::min(b, split_cap);
		if (le_h + add > INT32_MAX) {
			err = "device tree build: node count overflow";
			return hipErrorInvalidValue;
		}
		if (le_h + add > node_cap) {
			// grow the node array, keeping the nodes made so far
			int64_t cap = node_cap;
			while (cap < le_h + add) cap *= 2;
			Buf bigger;
			DTCHK(bigger.alloc(size_t(cap) * sizeof(DNode)));
			DTCHK(hipMemcpyAsync(bigger.p, dnodes.p, size_t(le_h) * sizeof(DNode), hipMemcpyDeviceToDevice, s));
			DTCHK(hipStreamSynchronize(s));
			std::swap(dnodes.p, bigger.p);
			std::swap(dnodes.bytes, bigger.bytes);
			node_cap = cap;
		}
		const int first = level;
		for (int i = 0; i < DT_BATCH; i++, level++) {
			const int bound = int(nl_bound);
			const int32_t* lev = dlev.as<const int32_t>();
			hipLaunchKernelGGL(k_split_flags, dim3(blocks(bound + 1, 256)), dim3(256), 0, s, dnodes.as<const DNode>(), lev, p.parts_per_node,
					MAX_SPLIT_LEVEL, dflags.as<int32_t>(), bound);
			DTLAUNCH("k_split_flags");
			size_t have = scan_tmp.bytes;
			DTCHK(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, have, dflags.as<int32_t>(), drank.as<int32_t>(), bound + 1, s));
			hipLaunchKernelGGL(k_split_setup, dim3(blocks(bound, 256)), dim3(256), 0, s, dnodes.as<const DNode>(), lev,
					dflags.as<const int32_t>(), drank.as<const int32_t>(), dinfo.as<SplitInfo>());
			hipLaunchKernelGGL(k_low_flags, dim3(blocks(n + 1, 256)), dim3(256), 0, s, dpos.as<const int32_t>(), dpnode.as<const int32_t>(), n,
					lev, dinfo.as<const SplitInfo>(), dlow.as<int32_t>());
			DTLAUNCH("k_split_setup/k_low_flags");
			have = scan_tmp.bytes;
			DTCHK(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, have, dlow.as<int32_t>(), dS.as<int32_t>(), int(n + 1), s));
			hipLaunchKernelGGL(k_split_nodes, dim3(blocks(bound, 256)), dim3(256), 0, s, dnodes.as<DNode>(), lev, dinfo.as<SplitInfo>(),
					dS.as<const int32_t>());
			if (n > 0) {
				hipLaunchKernelGGL(k_high_side_lows, dim3(blocks(n, 256)), dim3(256), 0, s, dpnode.as<const int32_t>(), n, lev,
						dinfo.as<const SplitInfo>(), dnodes.as<const DNode>(), dlow.as<const int32_t>(), dS.as<const int32_t>(),
						dmr.as<int32_t>());
				hipLaunchKernelGGL(k_exchange, dim3(blocks(n, 256)), dim3(256), 0, s, dpos.as<int32_t>(), dmass.as<float>(), dorder.as<int64_t>(),
						dpnode.as<int32_t>(), n, lev, dinfo.as<const SplitInfo>(), dnodes.as<const DNode>(), dlow.as<const int32_t>(),
						dS.as<const int32_t>(), dmr.as<const int32_t>());
			}
			hipLaunchKernelGGL(k_next_level, dim3(1), dim3(64), 0, s, dlev.as<int32_t>(), drank.as<const int32_t>(),
					dlog.as<int32_t>() + 3 * level);
			DTLAUNCH("k_split_nodes/k_high_side_lows/k_exchange/k_next_level");
			nl_bound = std::min(2 * nl_bound, lvl_max);
		}
		int32_t lg[3 * DT_BATCH];
		DTCHK(hipMemcpyAsync(lg, dlog.as<int32_t>() + 3 * first, sizeof(lg), hipMemcpyDeviceToHost, s));
		DTCHK(hipStreamSynchronize(s));
		for (int i = 0; i < DT_BATCH; i++) {
			const int32_t lb_i = lg[3 * i], le_i = lg[3 * i + 1], ns = lg[3 * i + 2];
			level_begin.push_back(lb_i);
			le = le_i;
			if (ns == 0) {
				finished = true;
				break;
			}
		}
		le_h = int64_t(lg[3 * DT_BATCH - 2]) + 2 * int64_t(lg[3 * DT_BATCH - 1]);
		nl_bound = std::max<int64_t>(2 * int64_t(lg[3 * DT_BATCH - 1]), 1);  // the next level's nodes, exactly
	}
	level_begin.push_back(le);
	const int nn = le;
	const double t1 = wall();
	// multipoles, levels from the deepest up
	DTCHK(dmulti.alloc(size_t(nn) * sizeof(DMulti)));
	for (size_t l = level_begin.size() - 1; l-- > 0;) {
		const int b = level_begin[l], nl = level_begin[l + 1] - b;
		hipLaunchKernelGGL(k_leaf_moments, dim3(blocks(nl, 4)), dim3(64), 0, s, dnodes.as<const DNode>(), b, nl, dpos.as<const int32_t>(),
				dmass.as<const float>(), dmulti.as<DMulti>());
		hipLaunchKernelGGL(k_interior_moments, dim3(blocks(nl, 64)), dim3(64), 0, s, dnodes.as<const DNode>(), b, nl, dmulti.as<DMulti>());
		DTLAUNCH("k_leaf_moments/k_interior_moments");
	}
	if (keep) {
		// everything stays on the device: the caller's DeviceTree takes the arrays
		DTCHK(hipStreamSynchronize(s));
		keep->release();
		keep->n = n;
		keep->nn = nn;
		keep->level_begin = level_begin;
		keep->pos = static_cast<int32_t*>(dpos.p);
		keep->mass = static_cast<float*>(dmass.p);
		keep->order = static_cast<int64_t*>(dorder.p);
		keep->nodes = static_cast<DNode*>(dnodes.p);
		keep->multi = static_cast<DMulti*>(dmulti.p);
		dpos.p = dmass.p = dorder.p = dnodes.p = dmulti.p = nullptr;
		T.pos.clear();
		T.mass.clear();
		T.order.clear();
		T.nodes.clear();
		T.leaves.clear();
		T.max_level = int(level_begin.size()) - 2;
		if (t_out) {
			t_out[0] = t1 - t0;
			t_out[1] = wall() - t1;
		}
		return hipSuccess;
	}
#undef DTLAUNCH
	{
		DeviceTree view;  // the build's arrays, still owned by the buffers above
		view.n = n;
		view.nn = nn;
		view.pos = static_cast<int32_t*>(dpos.p);
		view.mass = static_cast<float*>(dmass.p);
		view.order = static_cast<int64_t*>(dorder.p);
		view.nodes = static_cast<DNode*>(dnodes.p);
		view.multi = static_cast<DMulti*>(dmulti.p);
		e = device_tree_to_host(view, T, s, err);
		view = DeviceTree();  // not released: the buffers free them
		if (e != hipSuccess) return e;
	}
	if (t_out) {
		t_out[0] = t1 - t0;
		t_out[1] = wall() - t1;
	}
	return hipSuccess;
#undef DTCHK
}

hipError_t device_tree_to_host(const DeviceTree& D, Tree& T, hipStream_t s, std::string& err) {
	hipError_t e = hipSuccess;
#define DTCHK(call)                                                    \
	do {                                                               \
		e = (call);                                                    \
		if (e != hipSuccess) {                                         \
			err = std::string(#call) + ": " + hipGetErrorString(e);    \
			return e;                                                  \
		}                                                              \
	} while (0)
	const int64_t n = D.n;
	const int nn = int(D.nn);
	// through pinned staging
	std::lock_guard<std::mutex> lock(pinned_mtx);
	const size_t b_pos = size_t(3 * n) * 4, b_mass = size_t(n) * 4, b_order = size_t(n) * 8;
	const size_t b_nodes = size_t(nn) * sizeof(DNode), b_multi = size_t(nn) * sizeof(DMulti);
	const size_t o_mass = b_pos, o_order = o_mass + b_mass, o_nodes = (o_order + b_order + 255) / 256 * 256,
			o_multi = (o_nodes + b_nodes + 255) / 256 * 256;
	void* stage = nullptr;
	DTCHK(pinned_staging(o_multi + b_multi, &stage));
	char* st = static_cast<char*>(stage);
	if (n > 0) {
		DTCHK(hipMemcpyAsync(st, D.pos, b_pos, hipMemcpyDeviceToHost, s));
		DTCHK(hipMemcpyAsync(st + o_mass, D.mass, b_mass, hipMemcpyDeviceToHost, s));
		DTCHK(hipMemcpyAsync(st + o_order, D.order, b_order, hipMemcpyDeviceToHost, s));
	}
	DTCHK(hipMemcpyAsync(st + o_nodes, D.nodes, b_nodes, hipMemcpyDeviceToHost, s));
	DTCHK(hipMemcpyAsync(st + o_multi, D.multi, b_multi, hipMemcpyDeviceToHost, s));
	DTCHK(hipStreamSynchronize(s));
	T.pos.resize(size_t(3 * n));
	T.mass.resize(size_t(n));
	T.order.resize(size_t(n));
	const DNode* hn = reinterpret_cast<const DNode*>(st + o_nodes);
	const DMulti* hm = reinterpret_cast<const DMulti*>(st + o_multi);
	{
		// the particle arrays in parallel slices
		const int64_t chunk = 1 << 16;
		const int64_t nch = (n + chunk - 1) / chunk;
#pragma omp parallel for schedule(static)
		for (int64_t c = 0; c < nch; c++) {
			const int64_t b = c * chunk, e2 = std::min(n, b + chunk);
			std::memcpy(T.pos.data() + 3 * b, st + 12 * b, size_t(e2 - b) * 12);
			std::memcpy(T.mass.data() + b, st + o_mass + 4 * b, size_t(e2 - b) * 4);
			std::memcpy(T.order.data() + b, st + o_order + 8 * b, size_t(e2 - b) * 8);
		}
	}
	T.nodes.resize(size_t(nn));
	T.leaves.clear();
	T.max_level = 0;
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nn; i++) {
		Node& o = T.nodes[size_t(i)];
		const DNode& a = hn[size_t(i)];
		const DMulti& m = hm[size_t(i)];
		o.pbegin = a.pbegin;
		o.pend = a.pend;
		o.child[0] = a.c0;
		o.child[1] = a.c1;
		o.parent = a.parent;
		o.level = a.level;
		o.boxid = boxid_t(a.boxid);
		std::memcpy(o.m, m.m, sizeof(o.m));
		o.r = m.r;
		for (int d = 0; d < 3; d++) {
			o.x[d] = m.x[d];
			o.rmin[d] = m.rmin[d];
			o.rmax[d] = m.rmax[d];
		}
		o.N = m.N;
	}
	for (int i = 0; i < nn; i++) {
		if (T.nodes[size_t(i)].leaf()) T.leaves.push_back(i);
		T.max_level = std::max(T.max_level, T.nodes[size_t(i)].level);
	}
	return hipSuccess;
#undef DTCHK
}

}  // namespace tg
