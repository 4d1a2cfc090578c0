// tg_solver.hip -- the batched force solve: every interaction of one solve on the device.
//
// Stream order (all kernels on the solver's stream):
//   1. CC/CP (direct, Ewald) node items -> 35-coefficient partials per item
//   2. reduce partials per node -> dL[node]                  (the reference's L += Lacc.sum())
//   3. per leaf: L = sum over the root-to-leaf path of shift(dL[a]) (src/tree.cpp:400-401,481-491)
//   4. L2P at every leaf particle, minus phi_self            (src/tree.cpp:603-611, gravity.cpp:295)
//   5. PC direct, PP direct, PC Ewald, PP Ewald leaf items    (src/tree.cpp:616-626)
// which is the reference's per-node accumulation order.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "tg_gravity.h"
#include "tg_local.hpp"
#include "tg_solver.hpp"
#include "tg_devmem.hpp"
#include "tg_devtree.hpp"
#include "tg_fof.hpp"
#include "tg_upload.hpp"
#include "tg_walk.hpp"
#include "tg_solver.h"

namespace tg {

hipError_t DeviceArray::alloc(size_t n) {
	if (n <= bytes && p && owned) return hipSuccess;
	release();
	if (n == 0) return hipSuccess;
	hipError_t e = dev_malloc(&p, n);
	if (e == hipSuccess) bytes = n;
	return e;
}

void DeviceArray::release() {
	if (p && owned) dev_free(p);
	p = nullptr;
	bytes = 0;
	owned = true;
}

// ---- kernels ---------------------------------------------------------------------------------
// Each node's own CC/CP contribution: the sum of its partial slots in slot order (the
// reference adds each list's accumulator into the node's L, src/tree.cpp:400-401,481-491).
// One thread per (node, coefficient); consecutive threads read consecutive doubles.
__global__ __launch_bounds__(256) void k_node_sum(const int64_t* __restrict__ slot_ptr,
		const double* __restrict__ partial, double* __restrict__ dL, int64_t nn) {
	const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (i >= nn * 35) return;
	const int64_t a = i / 35;
	const int q = int(i - 35 * a);
	double acc = 0.0;
	for (int64_t z = slot_ptr[a]; z < slot_ptr[a + 1]; z++) acc += partial[35 * z + q];
	dL[i] = acc;
}

// The local-expansion pass, LPG leaves per block of 64 (64 / LPG lanes each).  The shift of a
// degree-4 Taylor polynomial is exact, so shifts compose and a leaf's expansion is the sum over its
// root-to-leaf path of every node's own contribution dL_a re-expanded about the leaf centre:
//   L_leaf = sum_a shift(dL_a, X_leaf - X_a)
// (the reference shifts the running sum node by node, src/tree.cpp:400-401; the two agree up to
// rounding).  Lane k of a leaf shifts the k-th node of its path (and k + 64 / LPG, ...) with the
// unrolled 210-FMA re-expansion, so no lane waits on a table; the leaf's lanes then add the path's
// terms coefficient by coefficient, root first, and evaluate at its particles.  Two leaves per
// block put the lanes a single path leaves idle to work.  leaf item: (node, pbegin, pend, path
// begin); the path has level(node) + 1 <= TG_MAXDEPTH entries.
#define TG_MAXDEPTH 64
template <int LPG>
__global__ __launch_bounds__(64) void k_local(const int4* __restrict__ leaves, int nleaves, const int* __restrict__ path,
		const int* __restrict__ level, const int4* __restrict__ pos, const int4* __restrict__ parts,
		const double* __restrict__ dL, double* __restrict__ out, double phi_self, int maxdepth, int wrap) {
	constexpr int W = 64 / LPG;  // lanes per leaf
	const int t = threadIdx.x;
	const int q = t / W, tl = t % W;
	const int li = int(blockIdx.x) * LPG + q;
	const bool active = li < nleaves;
	const int4 lf = active ? leaves[li] : make_int4(0, 0, 0, 0);
	const int depth = active ? level[lf.x] + 1 : 0;
	// dynamic LDS per leaf: [35][maxdepth] shifted contributions, then [35] the leaf's expansion
	extern __shared__ double lds[];
	double* mine = lds + q * (35 * maxdepth + 35);
	double* Ls = mine + 35 * maxdepth;
	const int4 c = pos[lf.x];
	for (int k = tl; k < depth; k += W) {
		const int a = path[lf.w + k];
		const int4 A = pos[a];
		const double d[3] = {double(int64_t(c.x) - int64_t(A.x)) * POS_INV, double(int64_t(c.y) - int64_t(A.y)) * POS_INV,
				double(int64_t(c.z) - int64_t(A.z)) * POS_INV};
		const double* src = dL + 35 * int64_t(a);
		double La[35], Lt[35];
#pragma unroll
		for (int r = 0; r < 35; r++) La[r] = src[r];
		l2l(Lt, La, d);
#pragma unroll
		for (int r = 0; r < 35; r++) mine[r * maxdepth + k] = Lt[r];
	}
	__syncthreads();
	for (int cc = tl; cc < 35 && active; cc += W) {
		// four partial sums over the path (root first within each), short dependency chains
		const double* row = mine + cc * maxdepth;
		double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
		int k = 0;
		for (; k + 4 <= depth; k += 4) {
			a0 += row[k];
			a1 += row[k + 1];
			a2 += row[k + 2];
			a3 += row[k + 3];
		}
		for (; k < depth; k++) a0 += row[k];
		Ls[cc] = (a0 + a1) + (a2 + a3);
	}
	__syncthreads();
	if (!active) return;
	// evaluate at the leaf's particles
	for (int i = lf.y + tl; i < lf.z; i += W) {
		const int4 p = parts[i];
		// the reference rounds the offset from the leaf centre to float (src/tree.cpp:607)
		const double d[3] = {double(float(double(centre_offset(p.x, c.x, wrap)) * POS_INV)),
				double(float(double(centre_offset(p.y, c.y, wrap)) * POS_INV)),
				double(float(double(centre_offset(p.z, c.z, wrap)) * POS_INV))};
		double phi, g[3];
		l2p(phi, g, Ls, d);
		// added last, onto the leaf kinds' sums
		double* o = out + 4 * size_t(i);
		o[0] += phi - phi_self;
		o[1] += g[0];
		o[2] += g[1];
		o[3] += g[2];
	}
}
constexpr int LOCAL_LPG = 2;

// The local-expansion pass as the reference runs it (src/tree.cpp:400-401): a node's expansion is
// its own contribution plus its parent's shifted to it, L_n = dL_n + shift(L_parent, X_n - X_parent),
// level by level from the root, one thread per node of the level (nodes are numbered level by
// level).  Every node is shifted once, where summing over each leaf's path shifts every ancestor
// once per leaf below it.
__global__ __launch_bounds__(64) void k_l2l_level(int begin, int end, const int* __restrict__ parent, const int4* __restrict__ pos,
		const double* __restrict__ dL, double* __restrict__ L) {
	const int n = begin + int(blockIdx.x) * 64 + int(threadIdx.x);
	if (n >= end) return;
	const int p = parent[n];
	const double* dn = dL + 35 * int64_t(n);
	double* Ln = L + 35 * int64_t(n);
	if (p < 0) {
#pragma unroll
		for (int r = 0; r < 35; r++) Ln[r] = dn[r];
		return;
	}
	const int4 C = pos[n], A = pos[p];
	const double d[3] = {double(int64_t(C.x) - int64_t(A.x)) * POS_INV, double(int64_t(C.y) - int64_t(A.y)) * POS_INV,
			double(int64_t(C.z) - int64_t(A.z)) * POS_INV};
	double Lp[35], Lt[35];
	const double* lp = L + 35 * int64_t(p);
#pragma unroll
	for (int r = 0; r < 35; r++) Lp[r] = lp[r];
	l2l(Lt, Lp, d);
#pragma unroll
	for (int r = 0; r < 35; r++) Ln[r] = Lt[r] + dn[r];
}

// L2P at every particle from its leaf's expansion (one thread per particle), with the reference's
// float rounding of the offset from the leaf centre (src/tree.cpp:607) and phi_self (quirk Q2)
__global__ __launch_bounds__(64) void k_l2p_parts(int64_t n, const int* __restrict__ pleaf, const int4* __restrict__ pos,
		const int4* __restrict__ parts, const double* __restrict__ L, double* __restrict__ out, double phi_self, int wrap) {
	const int64_t i = int64_t(blockIdx.x) * 64 + threadIdx.x;
	if (i >= n) return;
	const int lf = pleaf[i];
	const int4 c = pos[lf], p = parts[i];
	const double d[3] = {double(float(double(centre_offset(p.x, c.x, wrap)) * POS_INV)),
			double(float(double(centre_offset(p.y, c.y, wrap)) * POS_INV)),
			double(float(double(centre_offset(p.z, c.z, wrap)) * POS_INV))};
	double Ls[35];
	const double* src = L + 35 * int64_t(lf);
#pragma unroll
	for (int r = 0; r < 35; r++) Ls[r] = src[r];
	double phi, g[3];
	l2p(phi, g, Ls, d);
	double* o = out + 4 * i;  // added last, onto the leaf kinds' sums
	o[0] += phi - phi_self;
	o[1] += g[0];
	o[2] += g[1];
	o[3] += g[2];
}

// PP-direct items: the near prefix extended to the whole list (softened pair on every tile)
__global__ __launch_bounds__(256) void k_pp_near_all(int4* __restrict__ items, int n) {
	const int i = int(blockIdx.x) * 256 + threadIdx.x;
	if (i < n) items[i].y = pp_item_y(items[i].y & PP_COUNT_MASK, PP_NEAR_ALL);
}

__global__ __launch_bounds__(256) void k_zero_f64(double* __restrict__ p, int64_t n) {
	const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
	if (i < n) p[i] = 0.0;
}

static size_t local_lds_bytes(int maxdepth) { return size_t(LOCAL_LPG) * size_t(35 * maxdepth + 35) * sizeof(double); }

// ---- host side ---------------------------------------------------------------------------------
#define SCHK(call)                                           \
	do {                                                     \
		hipError_t e_ = (call);                              \
		if (e_ != hipSuccess) {                              \
			err = std::string(#call) + ": " + hipGetErrorString(e_); \
			return e_ == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;   \
		}                                                    \
	} while (0)

static double now() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int Solver::init(int dev) {
	device = dev;
	if (dev < 0) return TG_OK;  // host-only: tree, multipoles and lists, no device work
	int nd = 0;
	if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) {
		err = "no HIP device";
		return TG_ENODEV;
	}
	if (dev < 0 || dev >= nd) {
		err = "device ordinal out of range";
		return TG_EINVAL;
	}
	SCHK(hipSetDevice(dev));
	gpu_walk = true;  // the lists on the device by default (bit-identical to the host walk)
	gpu_tree = true;  // the tree and multipoles too (bit-identical to the host build)
	SCHK(init_device_tables(dev));
	SCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
	SCHK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
	SCHK(hipStreamCreateWithFlags(&stream3, hipStreamNonBlocking));
	SCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
	SCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
	SCHK(hipEventCreateWithFlags(&ev_join3, hipEventDisableTiming));
	SCHK(d_queue.alloc(256));
	int cus = 0;
	SCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
	resident_groups = cus * 4 * 6;  // six 64-lane work groups per SIMD
	ev.resize(size_t(2) * TIMER_EVSPANS * TIMER_RING);
	for (auto& e : ev) SCHK(hipEventCreate(&e));
	return TG_OK;
}

void Solver::drop_graph() {
	if (graph_exec) {
		(void) hipGraphExecDestroy(graph_exec);
		graph_exec = nullptr;
	}
}

// The solve's launches captured once into a HIP graph (no timing events inside) and replayed:
// one submission per solve instead of nine.  The capture holds no memset nodes: every buffer a
// solve reads is written earlier in the same solve.
int Solver::solve_graph() {
	SCHK(hipSetDevice(device));
	if (n == 0) return TG_OK;  // no particles: nothing to launch, and no graph to capture
	if (!graph_exec) {
		SCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
		const int rc = enqueue_solve(false);
		hipGraph_t g = nullptr;
		const hipError_t e = hipStreamEndCapture(stream, &g);
		if (rc != TG_OK) {
			if (g) (void) hipGraphDestroy(g);
			return rc;
		}
		SCHK(e);
		const hipError_t ei = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
		(void) hipGraphDestroy(g);
		SCHK(ei);
	}
	SCHK(hipGraphLaunch(graph_exec, stream));
	return TG_OK;
}

Solver::~Solver() {
	if (stream) {
		hipSetDevice(device);
		hipStreamSynchronize(stream);
		if (stream2) hipStreamSynchronize(stream2);
		if (stream3) hipStreamSynchronize(stream3);
		drop_graph();
		for (auto& e : ev)
			if (e) hipEventDestroy(e);
		for (auto& e : adj_ev)
			if (e) hipEventDestroy(e);
		d_parts.release();
		d_multis.release();
		d_nodepos.release();
		d_out.release();
		d_partial.release();
		d_node_slot_ptr.release();
		d_dL.release();
		d_queue.release();
		d_adj_w.release();
		d_adj_grad.release();
		d_adj_lam.release();
		d_adj_mu.release();
		d_adj_up.release();
		for (int e = 0; e < 2; e++) {
			d_adj_cpt_items[e].release();
			d_adj_cpt_sinks[e].release();
			d_adj_pct_items[e].release();
			d_adj_pct_leaves[e].release();
		}
		for (int e = 0; e < 2; e++) {
			d_adj_cct_src[e].release();
			d_adj_cct_snk[e].release();
			d_adj_cct_waves[e].release();
		}
		d_leaf_items.release();
		d_slot_node.release();
		d_path.release();
		d_level.release();
		d_parent.release();
		d_pleaf.release();
		d_Lnode.release();
		for (auto& w : work) {
			w.items.release();
			w.list.release();
			w.owner.release();
		}
		dlists.release();
		dtree.release();
		if (stream2) {
			hipStreamSynchronize(stream2);
			hipStreamDestroy(stream2);
		}
		if (stream3) {
			hipStreamSynchronize(stream3);
			hipStreamDestroy(stream3);
		}
		if (ev_fork) hipEventDestroy(ev_fork);
		if (ev_join) hipEventDestroy(ev_join);
		if (ev_join3) hipEventDestroy(ev_join3);
		hipStreamDestroy(stream);
		dev_trim(device);  // the cached blocks this solver's buffers went back to
	}
}

int Solver::set_particles(const int32_t* x, const float* mass, int64_t count, bool on_device) {
	if (on_device && !(gpu_tree && device >= 0)) {
		err = "particles in device memory need the device tree build (\"gpu_tree\" 1)";
		return TG_EINVAL;
	}
	drop_graph();  // the captured launches point at buffers that may be reallocated
	if (device >= 0) {
		// the last solves may still read the buffers this build hands back to the device cache
		// (tg_devmem.hpp), where the build's own stream would take them again
		SCHK(hipSetDevice(device));
		SCHK(hipStreamSynchronize(stream));
		if (stream2) SCHK(hipStreamSynchronize(stream2));
		if (stream3) SCHK(hipStreamSynchronize(stream3));
	}
	order_on_device = false;
	fof_group.clear();
	n = count;
	TreeParams P;
	P.ewald = ewald;
	P.theta = theta;
	P.soft_len = soft_len;
	P.parts_per_node = parts_per_node;
	double t0 = now();
	uniform_mass = (mass == nullptr);
	double t1, t2;
	// with the walk on the device too, the tree stays there: the walk and the solve's arrays are
	// made from it on the device, and the host copy only on demand (host_tree)
	const bool keep_tree = gpu_tree && gpu_walk && device >= 0;
	tree_on_device = false;
	if (gpu_tree && device >= 0) {
		double tt[2] = {0.0, 0.0};
		const hipError_t e = device_build_tree(tree, x, mass, count, P, m, device, stream, err, tt, on_device, keep_tree ? &dtree : nullptr);
		if (e != hipSuccess) return e == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;
		tree_on_device = keep_tree;
		t1 = t0 + tt[0];
		t2 = now();
	} else {
		tree.build(x, mass, count, P, m);
		t1 = now();
		tree.compute_multipoles();
		t2 = now();
	}
	if (!tree_on_device) dtree.release();
	n_nodes = tree_on_device ? dtree.nn : int64_t(tree.nodes.size());
	n_leaves = tree_on_device ? 0 : int64_t(tree.leaves.size());  // counted by upload_tree_device
	max_level = tree_on_device ? int(dtree.level_begin.size()) - 2 : tree.max_level;
	// the work lists of the last build may be views of the walk's lists, released below
	for (auto& w : work)
		if (!w.list.owned) w.list.release();
	lists_on_device = false;
	if (gpu_walk) {
		// the lists stay on the device; the work is built from them there (upload_lists_device)
		WalkNode* wn = nullptr;
		if (tree_on_device) {
			const hipError_t e = walk_nodes_device(dtree, &wn, stream, err);
			if (e != hipSuccess) return e == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;
		}
		const hipError_t e = device_walk(tree, device, stream, err, t_detail, &dlists, wn, n_nodes, &walk_hint);
		if (e != hipSuccess) return e == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP;
		lists_on_device = true;
	} else {
		dlists.release();
		tree.walk();
	}
	double t3 = now();
	t_tree = t1 - t0;
	t_multipoles = t2 - t1;
	t_walk = t3 - t2;
	int rc = upload();
	t_upload = now() - t3;
	return rc;
}

template <class T, class A>
static hipError_t put(DeviceArray& d, const std::vector<T, A>& h, hipStream_t s);

int Solver::update_particles(const int32_t* x, const float* mass) {
	{
		const int rc = host_tree();  // the moments are recomputed on the host tree
		if (rc != TG_OK) return rc;
	}
	if (tree.nodes.empty() || n == 0) {
		err = "update_particles before set_particles";
		return TG_EINVAL;
	}
	drop_graph();
	std::vector<int32_t> pt(3 * size_t(n));
	std::vector<float> mt;
	for (int64_t i = 0; i < n; i++)
		for (int d = 0; d < 3; d++) pt[3 * i + d] = x[3 * tree.order[i] + d];
	if (mass) {
		mt.resize(n);
		for (int64_t i = 0; i < n; i++) mt[i] = mass[tree.order[i]];
		uniform_mass = false;
	}
	tree.update_moments(pt.data(), mass ? mt.data() : nullptr);
	// The near prefix of each PP-direct item was classified from the leaf boxes of the build; moved
	// particles may now come within h of a source range classed far.  Every PP-direct item of an
	// updated tree therefore takes the softened pair on every tile (the adjoint reads the same
	// flag), so the solve stays the spline-softened sum over every listed pair.
	for (int4& it : host_pp_items) it.y = pp_item_y(it.y & PP_COUNT_MASK, PP_NEAR_ALL);
	if (device < 0) return TG_OK;
	SCHK(hipSetDevice(device));
	const int rc = upload_sources();
	if (rc) return rc;
	KindWork& w = work[PP_DIRECT];
	if (w.nitems > 0) {
		hipLaunchKernelGGL(k_pp_near_all, dim3(unsigned((w.nitems + 255) / 256)), dim3(256), 0, stream, w.items.as<int4>(),
				w.nitems);
		SCHK(hipGetLastError());
	}
	return TG_OK;
}

// particle records and multipoles of the current tree to the device
// the adjoint gathers each particle's PP source-side terms from its own list, which needs leaf B in
// leaf A's PP lists exactly when A is in B's (true of the walk's lists, checked here per tree)
bool Solver::pp_lists_symmetric() const {
	const Tree& T = tree;
	std::vector<int32_t> leaf_at(size_t(n) + 1, -1);
	for (int32_t lf : T.leaves)
		if (!T.nodes[lf].empty()) leaf_at[T.nodes[lf].pbegin] = lf;
	for (int kind = 0; kind < 2; kind++) {
		const hvec<int64_t>& ptr = kind ? T.lists.pp_e_ptr : T.lists.pp_d_ptr;
		const hvec<int32_t>& v = kind ? T.lists.pp_e : T.lists.pp_d;
		std::vector<std::pair<int32_t, int32_t>> pairs;
		for (int32_t lf : T.leaves) {
			if (size_t(lf) + 1 >= ptr.size()) continue;
			for (int64_t k = ptr[lf]; k < ptr[lf + 1]; k += 2) {
				const int32_t src = leaf_at[v[k]];
				if (src < 0 || T.nodes[src].pend != v[k + 1]) return false;  // not a whole leaf
				pairs.emplace_back(lf, src);
			}
		}
		std::vector<std::pair<int32_t, int32_t>> rev(pairs.size());
		for (size_t i = 0; i < pairs.size(); i++) rev[i] = {pairs[i].second, pairs[i].first};
		std::sort(pairs.begin(), pairs.end());
		std::sort(rev.begin(), rev.end());
		if (pairs != rev) return false;
	}
	return true;
}

int Solver::upload_sources() {
	const Tree& T = tree;
	const int64_t nn = int64_t(T.nodes.size());
	std::vector<int4> parts(n);
	for (int64_t i = 0; i < n; i++) {
		int w;
		std::memcpy(&w, &T.mass[i], 4);
		parts[i] = make_int4(T.pos[3 * i], T.pos[3 * i + 1], T.pos[3 * i + 2], w);
	}
	std::vector<MultiSrc> multis(nn);
	for (int64_t i = 0; i < nn; i++) {
		std::memcpy(multis[i].m, T.nodes[i].m, sizeof(T.nodes[i].m));
		std::memcpy(multis[i].x, T.nodes[i].x, sizeof(T.nodes[i].x));
	}
	SCHK(put(d_parts, parts, stream));
	SCHK(put(d_multis, multis, stream));
	SCHK(hipStreamSynchronize(stream));
	return TG_OK;
}

// list expansion: work group g writes runs [64 g, 64 g + 64) one after another, the lanes covering
// a run's entries: out[dst + l] = value + step * l
__global__ __launch_bounds__(64) void k_expand_runs(const ListRun* __restrict__ runs, int64_t nruns, int step, int* __restrict__ out) {
	const int64_t r0 = int64_t(blockIdx.x) * 64;
	const int64_t r1 = r0 + 64 < nruns ? r0 + 64 : nruns;
	for (int64_t r = r0; r < r1; r++) {
		const ListRun g = runs[r];
		for (int l = threadIdx.x; l < g.len; l += 64) out[g.dst + l] = g.value + step * l;
	}
}

template <class T, class A>
static hipError_t put(DeviceArray& d, const std::vector<T, A>& h, hipStream_t s) {
	hipError_t e = d.alloc(std::max<size_t>(h.size() * sizeof(T), 16));
	if (e != hipSuccess) return e;
	if (h.empty()) return hipSuccess;
	return hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

int Solver::host_tree() {
	if (!tree_on_device) return TG_OK;
	SCHK(hipSetDevice(device));
	SCHK(device_tree_to_host(dtree, tree, stream, err));
	tree_on_device = false;
	return TG_OK;
}

int Solver::host_lists() {
	if (!lists_on_device) return TG_OK;
	SCHK(hipSetDevice(device));
	Lists& Ls = tree.lists;
	hvec<int64_t>* ptrs[8] = {&Ls.cc_d_ptr, &Ls.cp_d_ptr, &Ls.cc_e_ptr, &Ls.cp_e_ptr, &Ls.pc_d_ptr, &Ls.pp_d_ptr, &Ls.pc_e_ptr, &Ls.pp_e_ptr};
	hvec<int32_t>* vals[8] = {&Ls.cc_d, &Ls.cp_d, &Ls.cc_e, &Ls.cp_e, &Ls.pc_d, &Ls.pp_d, &Ls.pc_e, &Ls.pp_e};
	const int64_t np = dlists.nn + 1;
	for (int k = 0; k < 8; k++) {
		ptrs[k]->resize(size_t(np));
		vals[k]->resize(size_t(dlists.len[k]));
		SCHK(hipMemcpyAsync(ptrs[k]->data(), dlists.ptr(k), size_t(np) * 8, hipMemcpyDeviceToHost, stream));
		if (dlists.len[k] > 0)
			SCHK(hipMemcpyAsync(vals[k]->data(), dlists.val[k], size_t(dlists.len[k]) * 4, hipMemcpyDeviceToHost, stream));
	}
	SCHK(hipStreamSynchronize(stream));
	lists_on_device = false;
	return TG_OK;
}

int Solver::upload() {
	if (device >= 0) SCHK(hipSetDevice(device));
	const Tree& T = tree;
	const int64_t nn = int64_t(T.nodes.size());
	HostWork W;
	const double tp0 = now();
	if (tree_on_device) {
		// the tree's arrays and the list work, all made on the device
		adj_checked = false;
		adj_lists_ready = false;
		int rc = upload_tree_device();
		const double tp1 = now();
		if (rc == TG_OK) rc = upload_lists_device();
		if (rc != TG_OK) return rc;
		t_detail[4] = now() - tp0;  // preparation: the tree's arrays and the list work
		t_detail[5] = 0.0;          // no transfers
		(void) tp1;
		SCHK(d_partial.alloc(std::max<size_t>(n_slots * 35 * sizeof(double), 16)));
		SCHK(d_dL.alloc(std::max<size_t>(size_t(n_nodes) * 35 * sizeof(double), 16)));
		SCHK(d_out.alloc(std::max<size_t>(n * 4 * sizeof(double), 16)));
		SCHK(hipStreamSynchronize(stream));
		return TG_OK;
	}
	const bool device_lists = lists_on_device && device >= 0;
	{
		// with the lists on the device only the tree's arrays are made here
		const int rc = device_lists ? prepare_tree_work(T, n, W, err) : prepare_work(T, n, ewald, soft_len, node_chunk, W, err);
		if (rc != TG_OK) return rc;
	}
	t_detail[4] = now() - tp0;
	if (T.max_level + 1 > TG_MAXDEPTH) {
		err = "tree deeper than the local-expansion pass supports (64 levels)";
		return TG_EINVAL;
	}
	n_slots = W.n_slots;
	n_leaf_items = int(W.leaf_items.size());
	auto* items = W.items;
	auto* lists = W.lists;
	auto* runs = W.runs;
	auto* owner_runs = W.owner_runs;
	auto* waves = W.waves;
	const int64_t* list_len = W.list_len;
	const hvec<int32_t>* walk_list[NKINDS] = {};  // CC kinds: the walk's flat list, node order
	walk_list[CC_DIRECT] = &T.lists.cc_d;
	walk_list[CC_EWALD] = &T.lists.cc_e;
	auto& parts = W.parts;
	auto& multis = W.multis;
	auto& npos = W.npos;
	auto& slot_node = W.slot_node;
	auto& slot_ptr = W.slot_ptr;
	auto& leaf_items = W.leaf_items;
	auto& path = W.path;
	auto& level = W.level;
	if (device < 0) {  // host-only: keep the PP-direct work for inspection, its list expanded here
		host_pp_items = std::move(items[PP_DIRECT]);
		host_pp_list.assign(size_t(list_len[PP_DIRECT]), 0);
		for (const ListRun& g : runs[PP_DIRECT])
			for (int32_t l = 0; l < g.len; l++) host_pp_list[size_t(g.dst + l)] = g.value + l;
		return TG_OK;
	}
	hipStream_t s = stream;
	SCHK(put(d_parts, parts, s));
	SCHK(put(d_multis, multis, s));
	SCHK(put(d_nodepos, npos, s));
	SCHK(put(d_leaf_items, leaf_items, s));
	adj_checked = false;
	adj_lists_ready = false;
	SCHK(put(d_path, path, s));
	path_len = int64_t(path.size());
	SCHK(put(d_level, level, s));
	{
		// the level-by-level local pass: node parents, level ranges, each particle's leaf
		hvec<int> parent(static_cast<size_t>(nn)), pleaf(static_cast<size_t>(n));
		level_begin.clear();
		for (int64_t i = 0; i < nn; i++) {
			parent[size_t(i)] = T.nodes[size_t(i)].parent;
			if (i == 0 || T.nodes[size_t(i)].level != T.nodes[size_t(i) - 1].level) level_begin.push_back(int(i));
			if (i > 0 && T.nodes[size_t(i)].level < T.nodes[size_t(i) - 1].level) {
				err = "nodes not numbered level by level";
				return TG_EINVAL;
			}
		}
		level_begin.push_back(int(nn));
		for (const int4& li : leaf_items)
			std::fill(pleaf.begin() + li.y, pleaf.begin() + li.z, li.x);
		SCHK(put(d_parent, parent, s));
		SCHK(put(d_pleaf, pleaf, s));
		SCHK(d_Lnode.alloc(std::max<size_t>(size_t(nn) * 35 * sizeof(double), 16)));
	}
	if (device_lists) {
		SCHK(hipStreamSynchronize(s));  // the host arrays above are freed on return
		const double tp1 = now();
		const int rc = upload_lists_device();
		if (rc != TG_OK) return rc;
		t_detail[5] = tp1 - tp0 - t_detail[4];  // transfers: the tree's arrays
		t_detail[4] += now() - tp1;            // preparation: the tree's arrays (host), the list work (device)
		SCHK(d_partial.alloc(std::max<size_t>(n_slots * 35 * sizeof(double), 16)));
		SCHK(d_dL.alloc(std::max<size_t>(T.nodes.size() * 35 * sizeof(double), 16)));
		SCHK(d_out.alloc(std::max<size_t>(n * 4 * sizeof(double), 16)));
		SCHK(hipStreamSynchronize(s));
		return TG_OK;
	}
	SCHK(put(d_node_slot_ptr, slot_ptr, s));
	SCHK(put(d_slot_node, slot_node, s));
	list_bytes = 0;
	// run tables stay allocated until the stream has drained below
	std::vector<DeviceArray> run_tables;
	auto expand_runs = [&](const hvec<ListRun>& r, int step, DeviceArray& out) -> hipError_t {
		if (r.empty()) return hipSuccess;
		run_tables.emplace_back();
		hipError_t e = put(run_tables.back(), r, s);
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(k_expand_runs, dim3(unsigned((r.size() + 63) / 64)), dim3(64), 0, s, run_tables.back().as<ListRun>(),
				int64_t(r.size()), step, out.as<int>());
		return hipGetLastError();
	};
	int rc = TG_OK;
	for (int k = 0; k < NKINDS && rc == TG_OK; k++) {
		const bool node_kind = (k == CC_DIRECT || k == CP_DIRECT || k == CC_EWALD || k == CP_EWALD);
		const size_t lbytes = size_t(list_len[k]) * sizeof(int);
		hipError_t e = hipSuccess;
		if (node_kind) {
			e = put(work[k].items, waves[k], s);
			if (e == hipSuccess) e = work[k].owner.alloc(std::max<size_t>(lbytes, 16));
			if (e == hipSuccess) e = expand_runs(owner_runs[k], 0, work[k].owner);
			work[k].nitems = int(waves[k].size());
			list_bytes += int64_t(waves[k].size() * sizeof(int2) + lbytes);
		} else {
			e = put(work[k].items, items[k], s);
			work[k].nitems = int(items[k].size());
			list_bytes += int64_t(items[k].size() * sizeof(int4));
		}
		if (e == hipSuccess) e = work[k].list.alloc(std::max<size_t>(lbytes, 16));
		if (e == hipSuccess && lbytes > 0) {
			if (walk_list[k])
				e = hipMemcpyAsync(work[k].list.p, walk_list[k]->data(), lbytes, hipMemcpyHostToDevice, s);
			else if (!lists[k].empty())
				e = hipMemcpyAsync(work[k].list.p, lists[k].data(), lbytes, hipMemcpyHostToDevice, s);
			else
				e = expand_runs(runs[k], 1, work[k].list);
		}
		work[k].nlist = list_len[k];
		list_bytes += int64_t(lbytes);
		if (e != hipSuccess) {
			err = hipGetErrorString(e);
			rc = TG_EHIP;
		}
	}
	const hipError_t es = hipStreamSynchronize(s);
	for (auto& t : run_tables) t.release();
	t_detail[5] = now() - tp0 - t_detail[4];
	if (rc != TG_OK) return rc;
	SCHK(es);
	SCHK(d_partial.alloc(std::max<size_t>(n_slots * 35 * sizeof(double), 16)));
	SCHK(d_dL.alloc(std::max<size_t>(T.nodes.size() * 35 * sizeof(double), 16)));
	SCHK(d_out.alloc(std::max<size_t>(n * 4 * sizeof(double), 16)));
	SCHK(hipStreamSynchronize(s));
	return TG_OK;
}


// One solve.  The leaf kinds (PC, PP) accumulate into the zeroed outputs; the node kinds (CC, CP),
// their sums per node and the level-by-level L2L touch none of them, so in an untimed solve (a
// graph capture) they run on a second stream beside the leaf kinds.  Each node kind writes partial
// slots of its own, so with overlap 2 the Ewald node kinds run on a third stream beside the direct
// ones (the table-bound CC_ewald leaves issue slots that the VALU-bound kernels fill); the node
// sums wait for both.  L2P adds last, after everything.  A timed solve runs everything on one
// stream in the same order of additions, so all forms give the same bits.
int Solver::enqueue_solve(bool timed) {
	SCHK(hipSetDevice(device));
	hipStream_t s = stream;
	const int64_t nn = n_nodes;
	int e = 0;
	if (timed && solves_recorded - solves_folded >= TIMER_RING) {
		// the set about to be reused still holds an unfolded solve: add its times first
		const int rc = fold_set(solves_folded);
		if (rc != TG_OK) return rc;
		solves_folded++;
	}
	hipEvent_t* evs = ev.data() + size_t(2) * TIMER_EVSPANS * size_t(solves_recorded % TIMER_RING);
	if (timed) solves_recorded++;
	auto mark = [&](int k) -> hipError_t { return timed ? hipEventRecord(evs[k], s) : hipSuccess; };
	const bool fork = overlap > 0 && !timed;
	const bool fork3 = overlap > 1 && !timed;
	hipStream_t sn = fork ? stream2 : s;     // the node side
	hipStream_t se = fork3 ? stream3 : sn;   // its Ewald kinds
	if (fork) {
		SCHK(hipEventRecord(ev_fork, s));
		SCHK(hipStreamWaitEvent(stream2, ev_fork, 0));
		if (fork3) SCHK(hipStreamWaitEvent(stream3, ev_fork, 0));
	}
	NodeArgs na{};
	na.sinkpos = d_nodepos.as<int4>();
	na.srcs = d_parts.as<int4>();
	na.multis = d_multis.as<MultiSrc>();
	na.partial = d_partial.as<double>();
	na.slot_node = d_slot_node.as<int>();
	na.wrap = ewald;
	const int node_kinds[4] = {CC_DIRECT, CP_DIRECT, CC_EWALD, CP_EWALD};
	for (int k : node_kinds) {
		na.waves = work[k].items.as<int2>();
		na.list = work[k].list.as<int>();
		na.owner = work[k].owner.as<int>();
		na.nwaves = work[k].nitems;
		SCHK(mark(2 * k));
		const bool ew = k == CC_EWALD || k == CP_EWALD;
		if (k == CC_DIRECT || k == CC_EWALD)
			SCHK(launch_cc(na, ew, ew ? se : sn));
		else
			SCHK(launch_cp(na, ew, ew ? se : sn));
		SCHK(mark(2 * k + 1));
	}
	if (fork3) {
		SCHK(hipEventRecord(ev_join3, stream3));
		SCHK(hipStreamWaitEvent(stream2, ev_join3, 0));
	}
	e = 2 * NKINDS;
	SCHK(mark(e));
	const bool local_pass = nn > 0 && n_leaf_items > 0;
	if (local_pass) {
		hipLaunchKernelGGL(k_node_sum, dim3(unsigned((nn * 35 + 255) / 256)), dim3(256), 0, sn,
				d_node_slot_ptr.as<int64_t>(), d_partial.as<double>(), d_dL.as<double>(), nn);
		SCHK(hipGetLastError());
		if (local_sweep)
			for (size_t l = 0; l + 1 < level_begin.size(); l++) {
				const int b = level_begin[l], e2 = level_begin[l + 1];
				hipLaunchKernelGGL(k_l2l_level, dim3(unsigned((e2 - b + 63) / 64)), dim3(64), 0, sn, b, e2, d_parent.as<int>(), d_nodepos.as<int4>(),
						d_dL.as<double>(), d_Lnode.as<double>());
				SCHK(hipGetLastError());
			}
	}
	if (n > 0) {
		// a kernel, not a memset, so the launch sequence stays graph-safe
		hipLaunchKernelGGL(k_zero_f64, dim3(unsigned((4 * n + 255) / 256)), dim3(256), 0, s, d_out.as<double>(), int64_t(4 * n));
		SCHK(hipGetLastError());
	}
	SCHK(mark(e + 1));
	LeafArgs la{};
	la.sinks = d_parts.as<int4>();
	la.srcs = d_parts.as<int4>();
	la.multis = d_multis.as<MultiSrc>();
	la.out = d_out.as<double>();
	la.wrap = ewald;
	la.uniform = uniform_mass ? 1 : 0;
	la.m = float(m);
	const double hu = soft_len * 4294967296.0;
	la.h2u = float(hu * hu);
	la.hinvu = float(1.0 / hu);
	la.hinv3u = float(1.0 / (hu * hu * hu));
	const int leaf_kinds[4] = {PC_DIRECT, PP_DIRECT, PC_EWALD, PP_EWALD};
	for (int k : leaf_kinds) {
		la.queue = (k == PP_DIRECT && pp_persistent) ? d_queue.as<int>() : nullptr;
		la.resident = resident_groups;
		la.items = work[k].items.as<int4>();
		la.list = work[k].list.as<int>();
		la.nitems = work[k].nitems;
		SCHK(mark(2 * k));
		if (k == PC_DIRECT || k == PC_EWALD)
			SCHK(launch_pc(la, k == PC_EWALD, s));
		else if (k == PP_DIRECT)
			SCHK(launch_pp_direct(la, true, s));
		else
			SCHK(launch_pp_ewald(la, s));
		SCHK(mark(2 * k + 1));
	}
	if (fork) {
		SCHK(hipEventRecord(ev_join, stream2));
		SCHK(hipStreamWaitEvent(s, ev_join, 0));
	}
	SCHK(mark(e + 2));
	if (local_pass) {
		const double phi_self = (-315.0 / 128.0) * m / soft_len;
		const int maxdepth = max_level + 1;
		if (local_sweep)
			hipLaunchKernelGGL(k_l2p_parts, dim3(unsigned((n + 63) / 64)), dim3(64), 0, s, n, d_pleaf.as<int>(), d_nodepos.as<int4>(),
					d_parts.as<int4>(), d_Lnode.as<double>(), d_out.as<double>(), phi_self, ewald);
		else
			hipLaunchKernelGGL(k_local<LOCAL_LPG>, dim3(unsigned((n_leaf_items + LOCAL_LPG - 1) / LOCAL_LPG)), dim3(64), local_lds_bytes(maxdepth),
					s, d_leaf_items.as<int4>(), n_leaf_items, d_path.as<int>(), d_level.as<int>(), d_nodepos.as<int4>(), d_parts.as<int4>(),
					d_dL.as<double>(), d_out.as<double>(), phi_self, maxdepth, ewald);
		SCHK(hipGetLastError());
	}
	SCHK(mark(e + 3));
	return TG_OK;
}

// per-kind kernel times (ms) of the last timed solve: 8 kinds then the local-expansion pass
int Solver::kernel_times(std::vector<float>& ms) {
	ms.assign(TIMER_SPANS, 0.0f);
	if (solves_recorded == 0) return TG_OK;
	const hipEvent_t* evs = ev.data() + size_t(2) * TIMER_EVSPANS * size_t((solves_recorded - 1) % TIMER_RING);
	SCHK(hipEventSynchronize(evs[2 * TIMER_EVSPANS - 1]));
	for (int k = 0; k < TIMER_EVSPANS; k++) {
		float t = 0.0f;
		SCHK(hipEventElapsedTime(&t, evs[2 * k], evs[2 * k + 1]));
		ms[std::min(k, TIMER_SPANS - 1)] += t;  // the two local spans report as one
	}
	return TG_OK;
}

int Solver::fold_set(int64_t solve) {
	const hipEvent_t* evs = ev.data() + size_t(2) * TIMER_EVSPANS * size_t(solve % TIMER_RING);
	SCHK(hipEventSynchronize(evs[2 * TIMER_EVSPANS - 1]));
	for (int k = 0; k < TIMER_EVSPANS; k++) {
		float t = 0.0f;
		SCHK(hipEventElapsedTime(&t, evs[2 * k], evs[2 * k + 1]));
		acc_ms[size_t(std::min(k, TIMER_SPANS - 1))] += t;  // the two local spans report as one
	}
	return TG_OK;
}

int Solver::collect_times(double* ms_sum, int64_t* count) {
	while (solves_folded < solves_recorded) {
		const int rc = fold_set(solves_folded);
		if (rc != TG_OK) return rc;
		solves_folded++;
	}
	for (int k = 0; k < TIMER_SPANS; k++) ms_sum[k] = acc_ms[k];
	*count = solves_recorded;
	acc_ms.assign(TIMER_SPANS, 0.0);
	solves_recorded = solves_folded = 0;
	return TG_OK;
}

__global__ void k_forces_input_order(const double* __restrict__ out, const int64_t* __restrict__ order, int64_t n,
		double* __restrict__ f) {
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int64_t o = order[i];
	for (int c = 0; c < 4; c++) f[4 * o + c] = out[4 * i + c];
}

int Solver::forces_device_order() {
	// a tree kept on the device left its order in d_order already
	if (!order_on_device && n > 0) {
		SCHK(hipSetDevice(device));
		SCHK(d_order.alloc(size_t(n) * sizeof(int64_t)));
		SCHK(hipMemcpyAsync(d_order.p, tree.order.data(), size_t(n) * sizeof(int64_t), hipMemcpyHostToDevice, stream));
		SCHK(hipStreamSynchronize(stream));
		order_on_device = true;
	}
	return TG_OK;
}

int Solver::forces_device(double* out4_dev) {
	SCHK(hipSetDevice(device));
	if (n == 0) return TG_OK;
	{
		const int rc = forces_device_order();
		if (rc != TG_OK) return rc;
	}
	hipLaunchKernelGGL(k_forces_input_order, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream, d_out.as<const double>(),
			d_order.as<const int64_t>(), n, out4_dev);
	SCHK(hipGetLastError());
	SCHK(hipStreamSynchronize(stream));
	return TG_OK;
}

int Solver::forces(double* out4) {
	SCHK(hipSetDevice(device));
	if (n == 0) return TG_OK;
	// put into input order on the device, then one copy
	SCHK(d_fin.alloc(size_t(n) * 4 * sizeof(double)));
	const int rc = forces_device(d_fin.as<double>());
	if (rc != TG_OK) return rc;
	SCHK(hipMemcpyAsync(out4, d_fin.p, size_t(n) * 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
	SCHK(hipStreamSynchronize(stream));
	return TG_OK;
}

}  // namespace tg

// ---- C ABI -------------------------------------------------------------------------------------
struct tg_solver {
	tg::Solver s;
};

namespace {
thread_local std::string g_solver_err;
int fail(tg_solver* h, int rc) {
	if (rc != TG_OK && h) g_solver_err = h->s.err;
	return rc;
}
}  // namespace

extern "C" {

int tg_solver_create(const tg_params* p, double theta, int parts_per_node, tg_solver** out) {
	if (!p || !out || !(theta > 0.0) || parts_per_node < 1 || !(p->soft_len > 0.0)) return TG_EINVAL;
	*out = nullptr;
	tg_solver* h = new tg_solver();
	h->s.ewald = p->ewald ? 1 : 0;
	h->s.soft_len = p->soft_len;
	h->s.m = p->m;
	h->s.theta = theta;
	h->s.parts_per_node = parts_per_node;
	int rc = h->s.init(p->device);
	if (rc != TG_OK) {
		g_solver_err = h->s.err;
		delete h;
		return rc;
	}
	*out = h;
	return TG_OK;
}

int tg_solver_destroy(tg_solver* h) {
	delete h;
	return TG_OK;
}

int tg_solver_set_particles(tg_solver* h, const int32_t* x, const float* mass, int64_t n) {
	if (!h || n < 0 || (n > 0 && !x) || n > (int64_t(1) << 31) - 64) return TG_EINVAL;
	return fail(h, h->s.set_particles(x, mass, n));
}

int tg_solver_set_particles_device(tg_solver* h, const int32_t* x, const float* mass, int64_t n) {
	if (!h || n < 0 || (n > 0 && !x) || n > (int64_t(1) << 31) - 64) return TG_EINVAL;
	return fail(h, h->s.set_particles(x, mass, n, true));
}

int tg_solver_get_forces_device(tg_solver* h, tg_force* f) {
	if (!h || h->s.device < 0 || (h->s.n > 0 && !f)) return TG_EINVAL;
	return fail(h, h->s.forces_device(reinterpret_cast<double*>(f)));
}

int tg_solver_find_groups(tg_solver* h, double link_len, uint64_t* group, int64_t* n_groups) {
	if (!h || h->s.device < 0 || (h->s.n > 0 && !group) || !(link_len > 0.0)) return TG_EINVAL;
	tg::Solver& S = h->s;
	const int64_t n = S.n;
	{
		const int rc = S.host_tree();  // the candidate search walks the host tree
		if (rc != TG_OK) return fail(h, rc);
	}
	S.fof_group.assign(size_t(n), TG_FOF_NONE);
	int64_t cand = 0;
	const hipError_t e = tg::fof_device(S.tree, S.d_parts.as<const int4>(), link_len, S.device, S.stream, S.fof_group.data(), &cand, S.err);
	if (e != hipSuccess) return fail(h, e == hipErrorOutOfMemory ? TG_ENOMEM : TG_EHIP);
	int64_t ng = 0;
	for (int64_t i = 0; i < n; i++) {
		group[S.tree.order[size_t(i)]] = S.fof_group[size_t(i)];
		ng += S.fof_group[size_t(i)] == uint64_t(i);
	}
	if (n_groups) *n_groups = ng;
	return TG_OK;
}

int tg_solver_group_catalog(tg_solver* h, const float* v, double a, int32_t min_n, tg_group* out, int64_t capacity, int64_t* n_out) {
	if (!h || !n_out || !(a > 0.0) || (h->s.n > 0 && !v) || capacity < 0) return TG_EINVAL;
	if (h->s.fof_group.size() != size_t(h->s.n)) {
		g_solver_err = "tg_solver_group_catalog before tg_solver_find_groups";
		return TG_EINVAL;
	}
	return tg::group_catalog(h->s.tree, h->s.fof_group, v, a, min_n, out, capacity, n_out);
}

int tg_solver_solve(tg_solver* h) {
	if (!h || h->s.device < 0) return TG_EINVAL;
	return fail(h, h->s.enqueue_solve(true));
}

int tg_solver_tune(tg_solver* h, const char* knob, int value) {
	if (!h || !knob) return TG_EINVAL;
	const std::string k(knob);
	tg::Solver& S = h->s;
	if (k == "pp_persistent") {
		S.pp_persistent = value != 0;
	} else if (k == "node_chunk") {
		if (value < 1 || value > tg::NODE_CHUNK_MAX) return TG_EINVAL;
		S.node_chunk = value;
	} else if (k == "resident") {
		if (value < 1) return TG_EINVAL;
		S.resident_groups = value;
	} else if (k == "overlap") {
		if (value < 0 || value > 2) return TG_EINVAL;
		S.overlap = value;
	} else if (k == "local_sweep") {
		S.local_sweep = value != 0;
	} else if (k == "gpu_walk") {
		if (S.device < 0 && value != 0) return TG_EINVAL;
		S.gpu_walk = value != 0;
	} else if (k == "gpu_tree") {
		if (S.device < 0 && value != 0) return TG_EINVAL;
		S.gpu_tree = value != 0;
	} else {
		g_solver_err = "unknown tuning knob " + k;
		return TG_EINVAL;
	}
	S.drop_graph();
	return TG_OK;
}

int tg_solver_adjoint(tg_solver* h, const float* w, double* grad, double* ms) {
	if (!h || h->s.device < 0 || (h->s.n > 0 && (!w || !grad))) return TG_EINVAL;
	tg::Solver& S = h->s;
	if (!S.adj_checked) {
		int rl = S.host_tree();  // the adjoint's set-up reads the tree and the lists on the host
		if (rl == TG_OK) rl = S.host_lists();
		if (rl != TG_OK) return fail(h, rl);
		S.adj_pp_symmetric = S.pp_lists_symmetric();
		S.adj_checked = true;
	}
	float t = 0.0f;
	const int rc = S.adjoint(w, grad, ms ? &t : nullptr);
	if (ms) *ms = t;
	return fail(h, rc);
}

int tg_solver_update_particles(tg_solver* h, const int32_t* x, const float* mass) {
	if (!h || (h->s.n > 0 && !x)) return TG_EINVAL;
	return fail(h, h->s.update_particles(x, mass));
}

int tg_solver_solve_graph(tg_solver* h) {
	if (!h || h->s.device < 0) return TG_EINVAL;
	return fail(h, h->s.solve_graph());
}

int tg_solver_synchronize(tg_solver* h) {
	if (!h || h->s.device < 0) return TG_EINVAL;
	hipError_t e = hipStreamSynchronize(h->s.stream);
	if (e != hipSuccess) {
		g_solver_err = hipGetErrorString(e);
		return TG_EHIP;
	}
	return TG_OK;
}

int tg_solver_get_forces(tg_solver* h, tg_force* f) {
	if (!h || h->s.device < 0 || (h->s.n > 0 && !f)) return TG_EINVAL;
	return fail(h, h->s.forces(reinterpret_cast<double*>(f)));
}

int tg_solver_kernel_ms(tg_solver* h, double* ms) {
	if (!h || h->s.device < 0 || !ms) return TG_EINVAL;
	std::vector<float> v;
	int rc = h->s.kernel_times(v);
	if (rc != TG_OK) return fail(h, rc);
	for (size_t k = 0; k < v.size(); k++) ms[k] = v[k];
	return TG_OK;
}

int tg_solver_collect_ms(tg_solver* h, double* ms_sum, int64_t* solves) {
	if (!h || h->s.device < 0 || !ms_sum || !solves) return TG_EINVAL;
	return fail(h, h->s.collect_times(ms_sum, solves));
}

int tg_solver_info(tg_solver* h, tg_solve_info* info) {
	if (!h || !info) return TG_EINVAL;
	const tg::Tree& T = h->s.tree;
	std::memset(info, 0, sizeof(*info));
	info->n = h->s.n;
	info->n_nodes = h->s.n_nodes;
	info->n_leaves = h->s.n_leaves;
	info->max_level = h->s.max_level;
	const tg::Census& C = T.census;
	const uint64_t cnt[8] = {C.PP_direct, C.PC_direct, C.CC_direct, C.CP_direct, C.PP_ewald, C.PC_ewald, C.CC_ewald, C.CP_ewald};
	for (int k = 0; k < 8; k++) info->count[k] = cnt[k];
	for (int k = 0; k < 8; k++) {
		info->items[k] = h->s.work[k].nitems;
		info->list_len[k] = h->s.work[k].nlist;
	}
	info->list_bytes = h->s.list_bytes;
	info->t_tree = h->s.t_tree;
	info->t_multipoles = h->s.t_multipoles;
	info->t_walk = h->s.t_walk;
	info->t_upload = h->s.t_upload;
	for (int k = 0; k < 6; k++) info->t_detail[k] = h->s.t_detail[k];
	return TG_OK;
}

// copy a named host-side array of the last build into dst (capacity in bytes); returns the
// array's size in bytes, or a negative status
int64_t tg_solver_export(tg_solver* h, const char* name, void* dst, int64_t capacity) {
	if (!h || !name) return TG_EINVAL;
	if (h->s.device >= 0) {
		int rc = h->s.host_tree();
		if (rc == TG_OK) rc = h->s.host_lists();
		if (rc != TG_OK) return fail(h, rc);
	}
	const tg::Tree& T = h->s.tree;
	std::vector<char> buf;
	auto put = [&](const void* p, size_t bytes) {
		buf.resize(bytes);
		if (bytes) std::memcpy(buf.data(), p, bytes);
	};
	const std::string nm(name);
	const size_t nn = T.nodes.size();
	if (nm == "pos") put(T.pos.data(), T.pos.size() * 4);
	else if (nm == "order") put(T.order.data(), T.order.size() * 8);
	else if (nm == "mass") put(T.mass.data(), T.mass.size() * 4);
	else if (nm == "leaves") put(T.leaves.data(), T.leaves.size() * 4);
	else if (nm == "node_range" || nm == "node_child" || nm == "node_parent" || nm == "node_level" || nm == "node_x" ||
	         nm == "node_m" || nm == "node_r" || nm == "node_boxid" || nm == "node_box" || nm == "node_n") {
		std::vector<char> v;
		for (size_t i = 0; i < nn; i++) {
			const tg::Node& d = T.nodes[i];
			const char* p = nullptr;
			size_t b = 0;
			int32_t tmp[2];
			uint64_t bid[2];
			double box[6];
			if (nm == "node_range") { tmp[0] = d.pbegin; tmp[1] = d.pend; p = (const char*) tmp; b = 8; }
			else if (nm == "node_child") { tmp[0] = d.child[0]; tmp[1] = d.child[1]; p = (const char*) tmp; b = 8; }
			else if (nm == "node_parent") { p = (const char*) &d.parent; b = 4; }
			else if (nm == "node_level") { p = (const char*) &d.level; b = 4; }
			else if (nm == "node_x") { p = (const char*) d.x; b = 12; }
			else if (nm == "node_m") { p = (const char*) d.m; b = 68; }
			else if (nm == "node_r") { p = (const char*) &d.r; b = 4; }
			else if (nm == "node_n") { p = (const char*) &d.N; b = 8; }
			else if (nm == "node_box") {
				std::memcpy(box, d.rmin, 24);
				std::memcpy(box + 3, d.rmax, 24);
				p = (const char*) box;
				b = 48;
			}
			else { bid[0] = uint64_t(d.boxid); bid[1] = uint64_t(d.boxid >> 64); p = (const char*) bid; b = 16; }
			v.insert(v.end(), p, p + b);
		}
		buf.swap(v);
	}
#define TG_EXPORT_LIST(NAME)                                                   \
	else if (nm == #NAME) put(T.lists.NAME.data(), T.lists.NAME.size() * 4);   \
	else if (nm == #NAME "_ptr") put(T.lists.NAME##_ptr.data(), T.lists.NAME##_ptr.size() * 8);
	TG_EXPORT_LIST(cc_d)
	TG_EXPORT_LIST(cp_d)
	TG_EXPORT_LIST(cc_e)
	TG_EXPORT_LIST(cp_e)
	TG_EXPORT_LIST(pc_d)
	TG_EXPORT_LIST(pp_d)
	TG_EXPORT_LIST(pc_e)
	TG_EXPORT_LIST(pp_e)
	else if (nm.rfind("solve_", 0) == 0) {
		// the tree arrays of the last build as the solve reads them on the device
		tg::Solver& S = h->s;
		if (S.device < 0) return TG_EINVAL;
		const size_t n = size_t(S.n), nnd = size_t(S.n_nodes);
		const tg::DeviceArray* a = nullptr;
		size_t bytes = 0;
		if (nm == "solve_parts") { a = &S.d_parts; bytes = n * 16; }
		else if (nm == "solve_multis") { a = &S.d_multis; bytes = nnd * sizeof(tg::MultiSrc); }
		else if (nm == "solve_nodepos") { a = &S.d_nodepos; bytes = nnd * 16; }
		else if (nm == "solve_level") { a = &S.d_level; bytes = nnd * 4; }
		else if (nm == "solve_parent") { a = &S.d_parent; bytes = nnd * 4; }
		else if (nm == "solve_leaf_items") { a = &S.d_leaf_items; bytes = size_t(S.n_leaf_items) * 16; }
		else if (nm == "solve_path") { a = &S.d_path; bytes = size_t(S.path_len) * 4; }
		else if (nm == "solve_pleaf") { a = &S.d_pleaf; bytes = n * 4; }
		else if (nm == "solve_order") {
			const int rc = S.forces_device_order();
			if (rc != TG_OK) return fail(h, rc);
			a = &S.d_order;
			bytes = n * 8;
		}
		else return TG_EINVAL;
		buf.resize(bytes);
		if (bytes) {
			if (hipSetDevice(S.device) != hipSuccess || !a->p || hipMemcpy(buf.data(), a->p, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
				g_solver_err = "tg_solver_export: device copy failed";
				return TG_EHIP;
			}
		}
	}
	else if (nm.rfind("work_", 0) == 0 || nm == "slot_node" || nm == "slot_ptr") {
		// the device work of the last build: work_items_K, work_list_K, work_owner_K (K a kind index)
		tg::Solver& S = h->s;
		const tg::DeviceArray* a = nullptr;
		size_t bytes = 0;
		if (S.device < 0) return TG_EINVAL;
		if (nm == "slot_node") {
			a = &S.d_slot_node;
			bytes = size_t(S.n_slots) * 4;
		} else if (nm == "slot_ptr") {
			a = &S.d_node_slot_ptr;
			bytes = (nn + 1) * 8;
		} else {
			const size_t us = nm.rfind('_');
			const int k = us == std::string::npos ? -1 : std::atoi(nm.c_str() + us + 1);
			if (k < 0 || k >= tg::NKINDS) return TG_EINVAL;
			const std::string what = nm.substr(0, us);
			const bool node_kind = k == tg::CC_DIRECT || k == tg::CP_DIRECT || k == tg::CC_EWALD || k == tg::CP_EWALD;
			const tg::KindWork& w = S.work[k];
			if (what == "work_items") {
				a = &w.items;
				bytes = size_t(w.nitems) * (node_kind ? sizeof(int2) : sizeof(int4));
			} else if (what == "work_list") {
				a = &w.list;
				bytes = size_t(w.nlist) * 4;
			} else if (what == "work_owner" && node_kind) {
				a = &w.owner;
				bytes = size_t(w.nlist) * 4;
			} else {
				return TG_EINVAL;
			}
		}
		buf.resize(bytes);
		if (bytes) {
			if (hipSetDevice(S.device) != hipSuccess || !a->p ||
					hipMemcpy(buf.data(), a->p, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
				g_solver_err = "tg_solver_export: device copy failed";
				return TG_EHIP;
			}
		}
	}
	else if (nm == "pp_d_items") put(h->s.host_pp_items.data(), h->s.host_pp_items.size() * sizeof(int4));
	else if (nm == "pp_d_list") put(h->s.host_pp_list.data(), h->s.host_pp_list.size() * 4);
#undef TG_EXPORT_LIST
	else return TG_EINVAL;
	if (dst) {
		if (capacity < int64_t(buf.size())) return TG_EINVAL;
		std::memcpy(dst, buf.data(), buf.size());
	}
	return int64_t(buf.size());
}

const char* tg_solver_last_error(void) { return g_solver_err.c_str(); }

}  // extern "C"
