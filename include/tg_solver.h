/*
 * tg_solver.h -- C ABI of the batched, device-resident force solve.
 *
 * The reference evaluates one force solve as a recursive walk that calls the eight kernels once
 * per tree node / leaf (tree::kick_fmm, src/tree.cpp:378-635, driven by solve_gravity,
 * src/main.cpp:31-73).  This interface runs the same solve in one piece: the host restatement of
 * the tree builds the kd-tree, multipoles and interaction lists (src/tree.cpp:122-296,378-635),
 * uploads them, and every interaction of the solve runs on the device.  Types and status codes
 * are those of tg_gravity.h.
 */
#ifndef TG_SOLVER_H
#define TG_SOLVER_H
#include <stdint.h>

#include "tg_gravity.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tg_solver tg_solver;

typedef struct tg_solve_info {
	int64_t n, n_nodes, n_leaves, max_level;
	/* interaction census as the reference counts it (src/tree.cpp:483-494,618-628), kinds in the
	 * order PP_direct, PC_direct, CC_direct, CP_direct, PP_ewald, PC_ewald, CC_ewald, CP_ewald */
	uint64_t count[8];
	int64_t items[8];      /* device work items per kind */
	int64_t list_len[8];   /* source indices per kind on the device */
	int64_t list_bytes;    /* device bytes of lists and items */
	double t_tree, t_multipoles, t_walk, t_upload;  /* seconds, last tg_solver_set_particles */
	/* finer split of t_walk and t_upload (seconds): device walk set-up, its steps, list assembly
	   and copy to the host, census; work-array preparation on the host, transfers */
	double t_detail[6];
} tg_solve_info;

/* replaces the tree root construction + options set-up (src/main.cpp:133-174): kernel
 * parameters, opening angle theta and leaf size parts_per_node for subsequent solves.
 * p->device = -1 gives a host-only solver: tree, multipoles, lists and export, no solve. */
int tg_solver_create(const tg_params* p, double theta, int parts_per_node, tg_solver** out);
int tg_solver_destroy(tg_solver* s);
/* particles (int32 x3 fixed point, n of them; mass NULL = all p->m): builds the tree and
 * multipoles (tree::refine, compute_multipoles) and the lists (kick_fmm's list construction) on
 * the device or, with "gpu_tree" / "gpu_walk" 0 or a host-only solver, on the host; makes them
 * resident on the device */
int tg_solver_set_particles(tg_solver* s, const int32_t* x, const float* mass, int64_t n);
/* tg_solver_set_particles with x (and mass, if not NULL) in device memory on the solver's device:
 * the tree is built there (needs "gpu_tree" 1, the default), so a time step needs no host copy of the particles
 * (the reference keeps them in its own memory, src/part_vect.cpp) */
int tg_solver_set_particles_device(tg_solver* s, const int32_t* x, const float* mass, int64_t n);
/* new positions (and masses, or NULL to keep them) for the same particles, in input order, with
 * the tree, the interaction lists and the expansion centres of the last tg_solver_set_particles
 * kept: the moments are recomputed about the kept centres.  The next solve evaluates that frozen
 * tree at the new positions -- the function tg_solver_adjoint differentiates (no reference
 * counterpart; the reference rebuilds the tree every step). */
int tg_solver_update_particles(tg_solver* s, const int32_t* x, const float* mass);
/* Adjoint of the last solve's particle set (no reference counterpart): for the loss
 * L = sum_i w_i . g_i (w: 3 floats per particle, input order), grad receives per particle
 * (dL/dx, dL/dy, dL/dz, dL/dm) in input order, positions in box units -- the derivatives of the
 * solve with its tree, lists and expansion centres held fixed (see tg_solver_update_particles).
 * Interactions are recomputed from the stored lists.  Needs leaves of at most 64 particles.
 * *ms, if not NULL, receives the device time of the adjoint kernels. */
int tg_solver_adjoint(tg_solver* s, const float* w, double* grad, double* ms);
/* enqueue one force solve on the solver's stream (replaces tree::kick_fmm's kernel calls) */
int tg_solver_solve(tg_solver* s);
/* tuning knobs (no reference counterpart): "pp_persistent" 0/1 runs PP direct as a persistent
 * queue-driven kernel; "node_chunk" 1..4096 (4096, the default: a whole list) sources per CC/CP partial slot (applies from the next
 * tg_solver_set_particles); "resident" work groups of the persistent form; "local_sweep" 1 runs the
 * local-expansion pass level by level (the reference's order, src/tree.cpp:400-401) instead of
 * per-leaf sums over the root path (0, the default); "gpu_walk" 1 (the default
 * for device solvers) builds the interaction lists on the device, 0 on the host (the same lists
 * either way: the walk of kick_fmm, src/tree.cpp:378-635; from the next tg_solver_set_particles);
 * with the lists on the device, the work of every kind is built from them there and the host
 * receives a copy only when something asks for it (tg_solver_export, the adjoint);
 * "gpu_tree" 1 (the default for device solvers) builds the tree and multipoles on the device, 0 on
 * the host (the same tree bit for bit; from the next tg_solver_set_particles);
 * "overlap" 1 (the default) lets graph replays run the node kinds on a second stream beside the
 * leaf kinds, 2 also runs the Ewald node kinds on a third stream beside the direct ones (the same
 * results bit for bit), 0 keeps one stream */
int tg_solver_tune(tg_solver* s, const char* knob, int value);
/* the same solve replayed from a HIP graph captured on its first call after set_particles (one
 * submission instead of 11 or more launches; records no per-kind timings; with "overlap" the node and
 * leaf kinds run concurrently) */
int tg_solver_solve_graph(tg_solver* s);
int tg_solver_synchronize(tg_solver* s);
/* forces of the last solve, in input particle order (the reference's `force` per particle) */
int tg_solver_get_forces(tg_solver* s, tg_force* f);
/* the same forces into device memory f (n tg_force, input order), for kick (tg_step.h) */
int tg_solver_get_forces_device(tg_solver* s, tg_force* f);
/* device time (ms) of each kind's kernels in the last solve: 8 kinds in tg_solve_info order, then
 * the local-expansion pass (reduce + L2L + L2P) */
int tg_solver_kernel_ms(tg_solver* s, double* ms /* [9] */);
/* sums of the same 9 device times over every solve enqueued since the previous call, and the
 * number of those solves (waits for them); the sums restart at zero afterwards */
int tg_solver_collect_ms(tg_solver* s, double* ms_sum /* [9] */, int64_t* solves);
int tg_solver_info(tg_solver* s, tg_solve_info* info);
/* copy a named host array of the last build (tree order positions "pos", "order", "mass",
 * "leaves", per node "node_range" "node_child" "node_parent" "node_level" "node_x" "node_m"
 * "node_r" "node_boxid" "node_box" "node_n", and the lists "cc_d" "cp_d" "cc_e" "cp_e" "pc_d"
 * "pp_d" "pc_e" "pp_e" with their "_ptr" CSR offsets; for a host-only solver also "pp_d_items",
 * the PP-direct work items (first sink, count | near-prefix length << 7, list begin, list end),
 * and "pp_d_list", their flattened source indices; for a device solver the device work of kind K
 * (tg_solve_info order) "work_items_K" (leaf kinds: int4 items as above; node kinds: int2 waves of
 * list entries), "work_list_K", "work_owner_K" (node kinds: the partial slot of each entry),
 * "slot_node", "slot_ptr" (each slot's node; each node's first slot), and the tree arrays the
 * solve reads: "solve_parts" (int4 x, y, z, mass bits), "solve_multis", "solve_nodepos",
 * "solve_level", "solve_parent", "solve_leaf_items" (node, first particle, end, path offset),
 * "solve_path", "solve_pleaf" (each particle's leaf), "solve_order").  Returns the size in bytes;
 * dst may be NULL to query it. */
int64_t tg_solver_export(tg_solver* s, const char* name, void* dst, int64_t capacity);
const char* tg_solver_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
