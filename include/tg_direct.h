/* tg_direct.h -- FP64 direct-summation forces, the accuracy reference of a force solve.
 *
 * Replaces the direct sum of the reference's solver test (src/main.cpp:137-164), whose result
 * src/output.cpp:17-51 compares with the tree solve; here in FP64 on the GPU, for a sample of
 * sinks against all particles. */
#ifndef TG_DIRECT_H
#define TG_DIRECT_H

#include <stdint.h>

#include "tg_gravity.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Potential and acceleration at particles sinks[0..ns) from all n particles x (int32 fixed point,
 * 3 per particle, the reference's vect<pos_type>), masses mass[] (NULL: every mass p->m):
 * spline-softened Newton with softening p->soft_len, plus the converged periodic (Ewald)
 * correction when p->ewald.  The self pair contributes the spline's r = 0 value.  out[k] is the
 * result for sinks[k]; FP64 throughout.  Returns TG_OK or a negative status. */
int tg_direct_forces(const tg_params* p, const int32_t* x, const float* mass, int64_t n, const int64_t* sinks, int64_t ns,
		tg_force* out);
/* Gradient of the direct-sum loss L = sum_i w_i . g_i (all particles, w: 3 floats each, the same
 * potential as tg_direct_forces) with respect to the positions and masses of particles sinks[k]:
 * out[4k..4k+3] = (dL/dx, dL/dy, dL/dz, dL/dm), positions in box units; FP64.  No reference
 * counterpart: the reference gradient the adjoint of a solve (tg_solver_adjoint) is compared with. */
int tg_direct_gradient(const tg_params* p, const int32_t* x, const float* mass, int64_t n, const float* w, const int64_t* sinks,
		int64_t ns, double* out);
const char* tg_direct_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
