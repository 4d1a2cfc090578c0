/*
 * tg_gravity.h -- C ABI of the MI355X (gfx950) gravity-interaction kernels.
 *
 * Drop-in boundary for tigergrav's eight interaction kernels (tigergrav/gravity.hpp:39-46),
 * called from the tree walk (src/tree.cpp:481-491, 616-626).  Plain pointers and sizes only.
 *
 * Conventions shared by every entry point
 *   positions   int32 triples, fixed point: x = p * 2^-32 + 0.5      (tigergrav/position.hpp:27-37)
 *   tg_force    {phi, g[3]} FP64, the reference `force`              (tigergrav/expansion.hpp:30-38)
 *   tg_multi_src{float m[17]; int32 x[3]}, the reference `multi_src` (tigergrav/gravity.hpp:12-21)
 *   expansion   35 FP64 local-expansion coefficients, the reference `expansion<double>` layout
 *   Outputs are ACCUMULATED (+=) into caller memory, as the reference does.
 *   Return value: 0 on success, a negative TG_E* code otherwise (never aborts).  The reference's
 *   weighted flop count (its return value) is written to *flop when flop != NULL.
 *   Kernel parameters come from tg_params on every call instead of the reference's options
 *   singleton frozen at first use (src/gravity.cpp:176-187).
 *   Thread safety: every host thread gets its own HIP stream and staging buffers; calls from
 *   different threads may overlap.
 */
#ifndef TG_GRAVITY_H
#define TG_GRAVITY_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TG_OK 0
#define TG_EINVAL (-1)   /* bad argument (NULL pointer with non-zero size, negative size) */
#define TG_EHIP (-2)     /* HIP runtime error; see tg_last_error() */
#define TG_ENODEV (-3)   /* no HIP device */
#define TG_ENOMEM (-4)   /* device allocation failed */

typedef struct tg_params {
	int ewald;        /* options::ewald -- periodic minimum-image separations in the direct kernels */
	double soft_len;  /* options::soft_len: spline softening length h */
	double m;         /* particle mass, options::m_tot / options::problem_size */
	int device;       /* HIP device ordinal */
} tg_params;

typedef struct tg_force {
	double phi;
	double g[3];
} tg_force;

typedef struct tg_multi_src {
	float m[17];
	int32_t x[3];
} tg_multi_src;

/* ---- the eight reference kernels (per-call drop-in; host pointers) ----------------------- */
/* replaces gravity_PP_direct, tigergrav/gravity.hpp:39 / src/gravity.cpp:171-300 */
int tg_gravity_PP_direct(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj,
		int do_phi, uint64_t* flop);
/* replaces gravity_PC_direct, tigergrav/gravity.hpp:40 / src/gravity.cpp:302-366 */
int tg_gravity_PC_direct(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const tg_multi_src* y, int64_t nc,
		uint64_t* flop);
/* replaces gravity_CC_direct, tigergrav/gravity.hpp:41 / src/gravity.cpp:368-426 */
int tg_gravity_CC_direct(const tg_params* p, double* L, const int32_t* x, const tg_multi_src* y, int64_t nc, uint64_t* flop);
/* replaces gravity_CP_direct, tigergrav/gravity.hpp:42 / src/gravity.cpp:428-486 */
int tg_gravity_CP_direct(const tg_params* p, double* L, const int32_t* x, const int32_t* y, int64_t nj, uint64_t* flop);
/* replaces gravity_PP_ewald, tigergrav/gravity.hpp:43 / src/gravity.cpp:491-614 */
int tg_gravity_PP_ewald(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj,
		uint64_t* flop);
/* replaces gravity_PC_ewald, tigergrav/gravity.hpp:44 / src/gravity.cpp:616-678 */
int tg_gravity_PC_ewald(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const tg_multi_src* y, int64_t nc,
		uint64_t* flop);
/* replaces gravity_CC_ewald, tigergrav/gravity.hpp:45 / src/gravity.cpp:680-731 */
int tg_gravity_CC_ewald(const tg_params* p, double* L, const int32_t* x, const tg_multi_src* y, int64_t nc, uint64_t* flop);
/* replaces gravity_CP_ewald, tigergrav/gravity.hpp:46 / src/gravity.cpp:733-785 */
int tg_gravity_CP_ewald(const tg_params* p, double* L, const int32_t* x, const int32_t* y, int64_t nj, uint64_t* flop);

/* replaces ewald_near_separation / ewald_far_separation (tigergrav/gravity.hpp:49-50,
 * src/gravity.cpp:10-23): minimum-image distance, and max(0.25, that) */
double tg_ewald_near_separation(const double x[3]);
double tg_ewald_far_separation(const double x[3], double r);

/* green_ewald (tigergrav/green.hpp:189-286) in FP64, host code: the periodic correction with the
 * reference's truncation (real space |n|^2 <= 17 within r < 3.6, Fourier |h|^2 <= 9) and its
 * derivatives to seventh order, at n separations x (3 doubles each, |x_i| <= 1/2).  120 doubles per
 * point: orders 0..7 in turn, each order's sorted index tuples in lexicographic order (orders 0..4
 * are the 35 of the expansion layout).  At x = 0 it gives the analytic limit, not the reference's
 * r -> 0 constant.  The device kernels interpolate a table of these values. */
int tg_ewald_green64(const double* x, int64_t n, double* D);

/* ---- no allocation ------------------------------------------------------------------------ */
/* tg_gravity_PP_direct without any allocation (no counterpart in the reference): for callers that
 * cannot allocate -- code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, results (bit for bit), flop count and errors as tg_gravity_PP_direct;
 * its staging lives in a fixed device array of 8.4 MiB that every device holds from the
 * library's load until the process ends, so sizes are limited: ns <= TG_NOALLOC_MAX_SINKS and
 * nj <= TG_NOALLOC_MAX_SOURCES, larger calls return TG_EINVAL before any device work.  The array
 * is shared by every call on one device and keeps the last call's contents: calls that overlap on
 * one device race (serialise them, e.g. on one host thread); calls on different devices do not.
 * Work runs on the calling thread's per-thread default stream. */
#define TG_NOALLOC_MAX_SINKS 16384
#define TG_NOALLOC_MAX_SOURCES 262144
int tg_gravity_PP_direct_noalloc(const tg_params* p, tg_force* f, const int32_t* x, int64_t ns, const int32_t* y, int64_t nj,
		int do_phi, uint64_t* flop);

/* ---- library status ----------------------------------------------------------------------- */
const char* tg_last_error(void);   /* message for the calling thread's last TG_EHIP / TG_ENOMEM */
int tg_device_count(void);
int tg_version(void);

#ifdef __cplusplus
}
#endif
#endif
