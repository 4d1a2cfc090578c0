/*
 * tg_fof.h -- C ABI of the friends-of-friends groups (an analysis consumer of the tree).
 *
 * Replaces the group finding of the reference (tree::find_groups, src/tree.cpp:637-760;
 * part_vect_find_groups / part_vect_init_groups, src/part_vect.cpp:84-133) and its group
 * statistics (src/groups.cpp:129-228) for the particles of a solver's last
 * tg_solver_set_particles.  Status codes are those of tg_gravity.h.
 */
#ifndef TG_FOF_H
#define TG_FOF_H
#include <stdint.h>

#include "tg_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TG_FOF_NONE 0xFFFFFFFFFFFFFFull /* DEFAULT_GROUP (tigergrav/particle.hpp:16): linked to no other particle */

/* groups on the device: particles whose squared int32 separation, as uint64, is below L^2 and not
 * zero, L = floor(N^(-1/3) link_len (2^32 - 1)), are friends -- except across the box edge, which
 * the reference's unwrapped search ranges (src/range.cpp:80-116) leave to the kd boxes' reach and
 * one side only, and which is never a link here; group[i] (input order)
 * receives the smallest tree-order index of particle i's connected component -- the id the
 * reference's relabelling converges to -- or TG_FOF_NONE.  *n_groups, if not NULL, receives the
 * number of groups (components of two or more particles). */
int tg_solver_find_groups(tg_solver* s, double link_len, uint64_t* group, int64_t* n_groups);

/* statistics of one group (struct group, tigergrav/groups.hpp:6-21) */
typedef struct tg_group {
	uint64_t id;
	int32_t N;
	float rmax;      /* largest member distance from the centre, times a */
	float rc;        /* median member distance, times a */
	double x[3];     /* centre (periodic mean) */
	float v[3];      /* mean velocity / a */
	float dv[3];     /* velocity dispersion per axis */
} tg_group;

/* the statistics of the groups of the last tg_solver_find_groups with at least min_n members, in
 * increasing id order, from velocities v (input order, float x3) at scale factor a: the passes of
 * groups_add_particle1 / groups_finish1 / groups_add_particle2 / groups_finish2 over the particles
 * in tree order (src/groups.cpp:129-228).  *n_out receives the number of such groups; out may be
 * NULL to query it. */
int tg_solver_group_catalog(tg_solver* s, const float* v, double a, int32_t min_n, tg_group* out, int64_t capacity, int64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif
