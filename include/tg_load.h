/*
 * tg_load.h -- C ABI of the initial-condition reader: N-GenIC / Gadget-1 files.
 *
 * Replaces load_header / load_particles (src/load.cpp:13-96, tigergrav/load.hpp:16-46).  Host
 * code; the arrays it fills are host memory.  Status codes are those of tg_gravity.h; the
 * message of a failure is tg_last_error()'s.
 */
#ifndef TG_LOAD_H
#define TG_LOAD_H
#include <stdint.h>

#include "tg_gravity.h"

#ifdef __cplusplus
extern "C" {
#endif

/* io_header_1 of N-GenIC (tigergrav/load.hpp:16-43), 256 bytes */
typedef struct tg_ngenic_header {
	uint32_t npart[6];     /* npart[1]: the particles of the file (other types are ignored) */
	double mass[6];
	double time;           /* scale factor */
	double redshift;
	int32_t flag_sfr, flag_feedback;
	uint32_t npartTotal[6];
	int32_t flag_cooling, num_files;
	double BoxSize, Omega0, OmegaLambda, HubbleParam;
	int32_t flag_stellarage, flag_metals, hashtabsize;
	char fill[84];
} tg_ngenic_header;

/* replaces load_header (src/load.cpp:13-32) */
int tg_ngenic_read_header(const char* path, tg_ngenic_header* h);
/* replaces load_particles (src/load.cpp:34-96): npart[1] particles into x (int32 x3 fixed point,
 * shifted by 0.5 N^(-1/3) and wrapped into the unit box) and v (float x3, times (1 + z)^(-3/2)).
 * *n receives npart[1]; with x and v both NULL only that is done (a size query).  The
 * reference's random choice of output particles (flags.out) is not made. */
int tg_ngenic_load(const char* path, int64_t capacity, int32_t* x, float* v, int64_t* n);

#ifdef __cplusplus
}
#endif
#endif
