/*
 * tg_step.h -- C ABI of the time step on the device: kick and drift with hierarchical rungs.
 *
 * The reference advances particles between force solves with part_vect_kick
 * (src/part_vect.cpp:341-436: half kicks around a rung re-assignment from |g|, and the energy,
 * momentum and force statistics) and part_vect_drift (src/part_vect.cpp:493-540: every particle
 * moved by the smallest active time step and wrapped into the unit box), on the time grid of
 * tigergrav/time.hpp.  These entry points do the same on particle arrays resident on the device,
 * so a step (solve, kick, drift, rebuild) needs no host round trip of the particle state.
 *
 * Arrays are DEVICE pointers on p->device, one entry per particle in one common particle order:
 *   x     int32 x3 fixed-point positions (tigergrav/position.hpp)
 *   v     float x3 velocities (particle::v, tigergrav/particle.hpp)
 *   rung  int8 rung per particle (particle::flags.rung)
 *   f     tg_force per particle (the solve's potential and acceleration)
 * Per-particle results (v, rung, x) are those of the reference's expressions, operation for
 * operation; the statistics are FP64 sums whose order differs from the reference's.
 * Return value: 0 or a negative TG_E* code (tg_gravity.h); work is complete on return.
 */
#ifndef TG_STEP_H
#define TG_STEP_H
#include <stdint.h>

#include "tg_gravity.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TG_RUNG_MAX 30 /* RUNG_MAX, tigergrav/time.hpp:9 */

typedef struct tg_step_params {
	double G;          /* options::G */
	double eta;        /* options::eta */
	double soft_len;   /* options::soft_len */
	double dt_max;     /* options::dt_max */
	double t_max;      /* options::t_max: rung r has time step t_max / 2^r */
	double m;          /* options::m_tot / options::problem_size */
	double problem_size; /* options::problem_size (the statistics' normalisation) */
	int glass;         /* options::glass: kicks reversed and damped by 1 / (1 + dt) */
	int cosmic;        /* options::cosmic: the first half kick is kick_dt1[rung] */
	int ewald;         /* options::ewald: the second half kick is kick_dt2[rung] (src/part_vect.cpp:389 tests
	                      this flag, not cosmic; without cosmology the table holds zeros) */
	double scale;      /* cosmo_scale().first, 1 without cosmology */
	double kick_dt1[TG_RUNG_MAX + 1]; /* cosmo_kick_dt1 per rung */
	double kick_dt2[TG_RUNG_MAX + 1]; /* cosmo_kick_dt2 per rung */
	int device;
} tg_step_params;

typedef struct tg_kick_stats {
	double g[3];       /* statistics::g: sum of G g / N / a^3 over output particles */
	double p[3];       /* statistics::p: sum of v / N / a^2 (velocities before the kick) */
	double pot;        /* statistics::pot: sum of m phi G / (2 a) over output particles */
	double kin;        /* statistics::kin: sum of m v.v / (2 a^2) (before the kick) */
	int32_t max_rung;  /* kick_return::rung: the largest rung assigned */
} tg_kick_stats;

/* replaces part_vect_kick (src/part_vect.cpp:341-436): particles with rung >= min_rung take the
 * closing half kick of their old rung, a new rung from |G g| (eta sqrt(soft_len / 2.8 / |a|),
 * capped by dt_max, at least min_rung, at most one rung below the old, at least 1) and the opening
 * half kick of it; with do_out every particle's force enters the statistics. */
int tg_kick(const tg_step_params* p, int64_t n, const tg_force* f, float* v, int8_t* rung, int min_rung, int do_out,
		tg_kick_stats* stats);

/* replaces part_vect_drift (src/part_vect.cpp:493-540): every particle moves by v dt with
 * dt = t_max / 2^max_rung (drift_dt with cosmology), wrapped into [0, 1); *ekin receives
 * sum m v.v / (2 a1^2) (a1 = cosmo_scale().second, 1 without cosmology) before the move. */
int tg_drift(const tg_step_params* p, int64_t n, int32_t* x, const float* v, int max_rung, double drift_dt, double a1,
		double* ekin);

/* the time grid (tigergrav/time.hpp:11-48), host functions */
double tg_rung_to_dt(double t_max, int rung);             /* rung_to_dt */
int tg_dt_to_rung(double t_max, double dt);               /* dt_to_rung (-1 where the reference aborts) */
int tg_min_rung(uint32_t t);                              /* min_rung */
uint32_t tg_time_inc(uint32_t t, int max_rung);           /* inc */
double tg_time_to_double(double t_max, uint32_t t);       /* time_to_double */

#ifdef __cplusplus
}
#endif
#endif
