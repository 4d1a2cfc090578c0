/*
 * tg_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of tigergrav's gravity-interaction kernels, used as the parity checker for
 * the HIP kernels in tigergrav_amd.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library; the product never links or calls it.
 *
 * Three families live here:
 *   tgo_*    : float per-pair arithmetic with FP64 lane accumulation, following the reference
 *              kernels step by step (src/gravity.cpp:171-785, tigergrav/green.hpp:62-286,
 *              tigergrav/simd.hpp:556-692) including the reference's own erfc/exp/sin
 *              approximations and its 16-lane summation order.  Pinned bit-for-bit against the
 *              compiled reference (oracle/_ref, tests/golden/).
 *   tgo64_*  : the same algorithm (same lattice truncation, same contractions, same quirks) in
 *              FP64 with libm erfc/exp/sin/cos -- the exact value of the formula the reference
 *              approximates.  Pinned by closed forms (Newton, Madelung constant, symmetry).
 *   tgo_direct64_* : FP64 physics oracle (spline-softened direct sum, optional Ewald periodic
 *              sum with converged lattice sums), for RMS force-error reporting.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tgo_params {
	int ewald;        /* options::ewald: wrap dX as periodic minimum image */
	double soft_len;  /* options::soft_len (h) */
	double m;         /* particle mass m_tot / problem_size */
} tgo_params;

typedef struct tgo_force {
	double phi;
	double g[3];
} tgo_force;          /* layout of the reference `force` (tigergrav/expansion.hpp:30-38) */

typedef struct tgo_multi_src {
	float m[17];
	int32_t x[3];
} tgo_multi_src;      /* layout of the reference `multi_src` (tigergrav/gravity.hpp:12-21) */

/* ---- reference-faithful float kernels (return the reference's weighted flop count) ---- */
uint64_t tgo_PP_direct(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj, int do_phi);
uint64_t tgo_PC_direct(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc);
uint64_t tgo_CC_direct(const tgo_params*, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc);
uint64_t tgo_CP_direct(const tgo_params*, double* L, const int32_t* x, const int32_t* y, int64_t nj);
uint64_t tgo_PP_ewald(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj);
uint64_t tgo_PC_ewald(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc);
uint64_t tgo_CC_ewald(const tgo_params*, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc);
uint64_t tgo_CP_ewald(const tgo_params*, double* L, const int32_t* x, const int32_t* y, int64_t nj);

/* building blocks, exposed for closed-form tests */
void tgo_green_direct(const float dx[3], float D[35]);
void tgo_green_ewald(const float dx[3], float D[35]);
float tgo_erfcexp(float x, float* e);
float tgo_exp(float x);
float tgo_sin(float x);

/* ---- same formulas in FP64 (libm math) ---- */
void tgo64_green_direct(const double dx[3], double D[35]);
void tgo64_green_ewald(const double dx[3], double D[35]);
void tgo64_PP_direct(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj, int do_phi);
void tgo64_PC_direct(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc);
void tgo64_CC_direct(const tgo_params*, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc);
void tgo64_CP_direct(const tgo_params*, double* L, const int32_t* x, const int32_t* y, int64_t nj);
void tgo64_PP_ewald(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj);
void tgo64_PC_ewald(const tgo_params*, tgo_force* f, const int32_t* x, int64_t ns, const tgo_multi_src* y, int64_t nc);
void tgo64_CC_ewald(const tgo_params*, double* L, const int32_t* x, const tgo_multi_src* y, int64_t nc);
void tgo64_CP_ewald(const tgo_params*, double* L, const int32_t* x, const int32_t* y, int64_t nj);

/* ---- FP64 physics oracle ----
 * Accelerations g and potentials phi on sinks `xs` (int32 fixed point, n_s of them) from all
 * sources `ys` with masses `m` (n_j), spline-softened with length h (GADGET-2 cubic spline, as the
 * reference, src/gravity.cpp:231-287).  periodic=0: plain open-boundary sum, dX = exact
 * difference.  periodic=1: minimum-image spline part + converged Ewald correction (alpha=2,
 * |n|^2<=25, |h|^2<=16), correction(0) = its r->0 limit.  phi excludes no self term: a sink that
 * coincides with a source gets the spline value at r=0 (and the Ewald self constant).  OpenMP
 * parallel over sinks. */
void tgo_direct64(int periodic, double h, const int32_t* xs, int64_t ns, const int32_t* ys, const double* m, int64_t nj,
		double* phi, double* g);
/* periodic correction (potential, acceleration) of a unit mass at separation dx, FP64 */
void tgo_ewald_corr64(const double dx[3], double* phi, double g[3]);

/* FP64 frozen-tree solve (no float rounding anywhere): moments recomputed in FP64 about the kept
 * centres node_x from the positions pos and masses (tree order; mass NULL = P->m for all), then
 * the solve over the given lists with the tgo64_ formulas and per-particle masses; forces in tree
 * order */
void tgo64_frozen_solve(const tgo_params* P, int64_t N, const int32_t* pos, const double* mass, int64_t nn,
		const int32_t* node_range, const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x,
		const int64_t* const* ptr, const int32_t* const* lst, tgo_force* out);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* the reference's L2L (in place) and L2P, FP64 */
void tgo_l2l(double* L, const double* dX);
void tgo_l2p(const double* L, const double* dX, tgo_force* f);
/* a whole force solve over given lists (see tg_oracle.c) */
uint64_t tgo_solve(const tgo_params* P, int64_t N, const int32_t* pos, int64_t nn, const int32_t* node_range,
		const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x, const float* node_m,
		const int64_t* const* ptr, const int32_t* const* lst, int stride, tgo_force* out, uint64_t* counts);
/* the same solve restricted to the nodes with want[n] != 0 (a set of leaves and all their
 * ancestors gives those leaves' exact results; the other particles' forces are left as they are) */
uint64_t tgo_solve_want(const tgo_params* P, int64_t N, const int32_t* pos, int64_t nn, const int32_t* node_range,
		const int32_t* node_child, const int32_t* node_parent, const int32_t* node_x, const float* node_m,
		const int64_t* const* ptr, const int32_t* const* lst, const uint8_t* want, tgo_force* out);

/* ---- time step: part_vect_kick / part_vect_drift (src/part_vect.cpp:341-436,493-540) ---- */
typedef struct tgo_step_params {
	double G, eta, soft_len, dt_max, t_max, m, problem_size;
	int glass, cosmic, ewald;
	double scale;
	double kick_dt1[31], kick_dt2[31];
} tgo_step_params;
/* stats: g[3], p[3], pot, kin, max rung; returns -1 if a step falls below rung 30 */
int tgo_kick(const tgo_step_params* P, int64_t n, const tgo_force* f, float* v, int8_t* rung, int min_rung, int do_out,
		double* stats);
/* returns the kinetic energy before the move */
double tgo_drift(const tgo_step_params* P, int64_t n, int32_t* x, const float* v, int max_rung, double drift_dt, double a1);


/* ---- friends-of-friends (src/part_vect.cpp:84-120), all pairs, to convergence ---- */
void tgo_fof(int64_t n, const int32_t* pos_tree, uint64_t L2, uint64_t* group);

#ifdef __cplusplus
}
#endif
