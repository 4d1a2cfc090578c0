// tg_solver.hpp -- batched, device-resident force solve over the tree's interaction lists.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "tg_devtree.hpp"
#include "tg_kernels.hpp"
#include "tg_tree.hpp"
#include "tg_upload.hpp"
#include "tg_walk.hpp"

namespace tg {

struct DeviceArray {
	void* p = nullptr;
	size_t bytes = 0;
	bool owned = true;  // false: a view of memory another object frees (release only forgets it)
	hipError_t alloc(size_t n);
	void release();
	template <class T>
	T* as() const {
		return static_cast<T*>(p);
	}
};

// one kind's work on the device
struct KindWork {
	DeviceArray items;  // int4 leaf items (PP/PC) or int2 waves (CC/CP)
	DeviceArray list;   // int32 source indices
	DeviceArray owner;  // int32 partial slot per list entry (CC/CP)
	int nitems = 0;
	int64_t nlist = 0;
};

struct Solver {
	int device = 0;
	hipStream_t stream = nullptr;
	int ewald = 1;
	double soft_len = 0.0, m = 0.0, theta = 0.7;
	int parts_per_node = 64;
	bool uniform_mass = true;
	int node_chunk = 4096;  // sources per CC/CP partial slot (<= NODE_CHUNK_MAX); applies from the next set_particles
	bool pp_persistent = false;  // PP direct as a persistent queue-driven kernel
	// graph replays: 0 one stream; 1 the node kinds on a second stream beside the leaf kinds; 2 as 1,
	// with the Ewald node kinds on a third stream beside the direct ones
	int overlap = 1;
	hipStream_t stream2 = nullptr, stream3 = nullptr;
	hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join3 = nullptr;
	bool local_sweep = false;    // local pass level by level (k_l2l_level) instead of per-leaf path sums (k_local)
	bool gpu_walk = false;       // build the interaction lists on the device (device_walk); on for device solvers
	WalkHint walk_hint;          // the buffer sizes the last device walk reached: the next one starts there
	bool gpu_tree = false;       // build the tree and multipoles on the device (device_build_tree); on for device solvers
	int resident_groups = 1024;  // work groups of the persistent form (set from the CU count)
	Tree tree;
	int64_t n = 0;
	// device state
	DeviceArray d_parts, d_multis, d_nodepos, d_out, d_partial, d_node_slot_ptr;
	DeviceArray d_leaf_items, d_slot_node, d_path, d_level, d_dL, d_queue;
	DeviceArray d_parent, d_pleaf, d_Lnode;  // level-by-level local pass: parents, particle -> leaf, node expansions
	std::vector<int> level_begin;          // first node of each level, then the node count
	// adjoint buffers: loss weights, gradients, expansion and multipole adjoints
	DeviceArray d_adj_w, d_adj_grad, d_adj_lam, d_adj_mu;
	bool adj_pp_symmetric = false;  // checked once per tree at the first adjoint call
	bool adj_checked = false;
	hipEvent_t adj_ev[2] = {nullptr, nullptr};
	// CC entries by source for the adjoint (direct, Ewald): built at the first adjoint call per tree
	DeviceArray d_adj_cct_src[2], d_adj_cct_snk[2], d_adj_cct_waves[2];
	int adj_cct_nwaves[2] = {0, 0};
	// CP entries by source leaf (direct, Ewald): items (first particle, count, sink range) and sink nodes
	DeviceArray d_adj_cpt_items[2], d_adj_cpt_sinks[2];
	int adj_cpt_nitems[2] = {0, 0};
	// PC entries by source node (direct, Ewald): items (node, sink-leaf run) and sink leaves
	DeviceArray d_adj_pct_items[2], d_adj_pct_leaves[2];
	int adj_pct_nitems[2] = {0, 0};
	int adj_pc_lgL = 6;  // lanes per sink leaf in k_adj_pc_t: 2^adj_pc_lgL >= the largest leaf
	// (node, child) pairs for the Lambda up-sweep, grouped by level deepest first: batch l is
	// [adj_up_ptr[l], adj_up_ptr[l + 1])
	DeviceArray d_adj_up;
	std::vector<int> adj_up_ptr;
	bool adj_lists_ready = false;
	KindWork work[NKINDS];
	// host-only solvers (device < 0) keep the PP-direct items and flattened source list
	hvec<int4> host_pp_items;
	std::vector<int> host_pp_list;
	std::vector<uint64_t> fof_group;  // friends-of-friends group per particle, tree order (find_groups)
	int n_leaf_items = 0;
	int64_t n_slots = 0;
	int64_t list_bytes = 0;
	// the device build's tree (gpu_tree with gpu_walk): the walk's node table and the solve's tree
	// arrays are made from it on the device (upload_tree_device), and the host copy in `tree` only
	// on demand (host_tree)
	DeviceTree dtree;
	bool tree_on_device = false;  // tree holds no copy of dtree yet
	int64_t n_nodes = 0;          // nodes of the current tree
	int64_t n_leaves = 0;         // its leaves, empty ones included
	int max_level = 0;            // its deepest level
	int64_t path_len = 0;         // entries of the root-to-leaf paths (d_path)
	// the device walk's lists (gpu_walk): the work of every kind is built from them on the device
	// (upload_lists_device), the CC and PC work lists are views of them, and the host copy in
	// tree.lists is made only on demand (host_lists)
	DeviceLists dlists;
	bool lists_on_device = false;  // tree.lists holds no copy of dlists yet
	// timings (seconds) of the last build and solve
	double t_tree = 0, t_multipoles = 0, t_walk = 0, t_upload = 0;
	double t_detail[6] = {};  // walk: set-up, steps, assembly, census; upload: preparation, transfers
	// per-solve event sets: a ring of TIMER_RING solves, each 2 events per timed span (8 kinds,
	// then the local-expansion pass in two spans: node sums and L2L before the leaf kinds, L2P
	// after them); sets are folded into acc_ms (8 kinds, local) before they are reused
	static constexpr int TIMER_RING = 32, TIMER_SPANS = NKINDS + 1, TIMER_EVSPANS = NKINDS + 2;
	std::vector<hipEvent_t> ev;
	int64_t solves_recorded = 0, solves_folded = 0;
	std::vector<double> acc_ms = std::vector<double>(TIMER_SPANS, 0.0);
	std::string err;

	int init(int dev);
	int set_particles(const int32_t* x, const float* mass, int64_t count, bool on_device = false);
	int upload();
	int upload_lists_device();  // every kind's work from dlists, on the device (tg_prep.hip)
	int upload_tree_device();   // the tree's arrays from dtree, on the device (tg_prep.hip)
	int host_tree();            // copies the device build's tree into tree
	int host_lists();  // copies the device walk's lists into tree.lists
	int upload_sources();
	int update_particles(const int32_t* x, const float* mass);
	// gradients of sum_i w_i . g_i (w: 3 per particle, input order) with respect to positions and
	// masses: grad 4 per particle (d/dx, d/dy, d/dz, d/dm), input order; ms: device time, or NULL
	int adjoint(const float* w, double* grad, float* ms);
	bool pp_lists_symmetric() const;
	int enqueue_solve(bool timed);
	hipGraphExec_t graph_exec = nullptr;
	int solve_graph();
	void drop_graph();
	int kernel_times(std::vector<float>& ms);       // last timed solve
	int collect_times(double* ms_sum, int64_t* n);  // sums over the timed solves since the last collect
	int fold_set(int64_t solve);
	int forces(double* out4);  // N x (phi, gx, gy, gz), input order
	int forces_device(double* out4_dev);  // the same into device memory, on the solver's stream (waits for it)
	int forces_device_order();            // d_order: the input index of each tree-order particle
	DeviceArray d_order;  // input index of each tree-order particle, uploaded by forces_device
	DeviceArray d_fin;    // forces in input order, for the copy to the host
	bool order_on_device = false;
	~Solver();
};

}  // namespace tg
