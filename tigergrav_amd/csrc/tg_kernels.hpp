// tg_kernels.hpp -- launch interface of the gfx950 interaction kernels (host side).
//
// Work is described by int4 items.  Leaf kinds (PP/PC: sinks are particles):
//   item = (first sink, sink count <= 64, list begin, list end)
// The count field of every leaf item is read through PP_COUNT_MASK.  For PP direct the upper bits
// carry the length of the list's near prefix (sources that may lie within the softening length):
// y = count | near_len << PP_NEAR_SHIFT, with near_len == PP_NEAR_ALL meaning the whole list.
// Results are added to out[4 * sink + {0:phi, 1..3:g}] (FP64; the item owns its sinks, so no
// atomics are needed while kernels of one stream run in order).
// Node kinds (CC/CP: the sink is an expansion centre): a wave takes a run of consecutive list
// entries, waves[w] = [begin, end): at most 64 of one or more partial slots, or any number of one
// slot (lane l then takes entries begin + l + 64 k); entry j belongs to partial slot owner[j],
// whose expansion centre is sinkpos[slot_node[slot]].  Every slot's entries lie inside one wave,
// so each slot's 35 FP64 coefficients are written exactly once (no atomics, deterministic).
// `list` holds source indices (particles for PP/CP, multipoles for PC/CC); NULL means the
// identity list.  For node kinds owner == NULL means one slot per wave (slot = wave index) and
// slot_node == NULL means every slot's centre is sinkpos[0].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace tg {

constexpr int PP_COUNT_MASK = 127;
constexpr int PP_NEAR_SHIFT = 7;
constexpr int PP_NEAR_ALL = (1 << 25) - 1;  // all 25 upper bits set
__host__ __device__ constexpr int pp_item_y(int count, int64_t near_len) {
	return int(unsigned(count) | (unsigned(near_len >= PP_NEAR_ALL ? PP_NEAR_ALL : near_len) << PP_NEAR_SHIFT));
}

// The dispatch order of leaf items: by cost class, heaviest class first, items of one class in leaf
// (tree) order, so that items running together share most of their sources in the L2.  The cost
// is the list length times the sink count rounded up to a power of two; its class is its bit
// length.
__host__ __device__ constexpr int item_class(int64_t list_len, int cnt) {
	int sh = 0;
	while (cnt > 1 && (1 << sh) < cnt) sh++;
	uint64_t c = uint64_t(list_len) << sh;
	int b = 0;
	while (c) {
		b++;
		c >>= 1;
	}
	return b;
}

struct MultiSrc {
	float m[17];
	int32_t x[3];
};

struct LeafArgs {
	const int4* sinks;       // sink particles (x, y, z, mass bits)
	const int4* srcs;        // source particles (PP)
	const MultiSrc* multis;  // source multipoles (PC)
	const int4* items;
	const int* list;
	double* out;             // 4 doubles per sink
	int nitems;
	int wrap;                // periodic minimum-image separations (direct kinds)
	int uniform;             // all source masses equal to m (PP direct sums once, scales at the end)
	float m;                 // the common mass when uniform
	float h2u, hinvu, hinv3u;  // softening in position units (2^32 per box length)
	int* queue;              // PP direct: work counter of the persistent form (NULL: one group per item)
	int resident;            // PP direct persistent form: work groups to launch
};

struct NodeArgs {
	const int4* sinkpos;     // expansion centres (x, y, z, unused)
	const int4* srcs;        // source particles (CP)
	const MultiSrc* multis;  // source multipoles (CC)
	const int2* waves;       // [begin, end) runs of list entries, <= 64 each
	const int* list;         // source index per entry
	const int* owner;        // partial slot per entry
	const int* slot_node;    // expansion-centre index per slot
	double* partial;         // 35 doubles per slot
	int nwaves;
	int wrap;
};

enum Kind { PP_DIRECT = 0, PC_DIRECT, CC_DIRECT, CP_DIRECT, PP_EWALD, PC_EWALD, CC_EWALD, CP_EWALD, NKINDS };

// one-time per-device setup of the tabulated Ewald correction; returns hipSuccess or the error
hipError_t init_device_tables(int device);
// the tabulated correction on a device (after init_device_tables), for other translation units
const float* ewald_table_on(int device);

hipError_t launch_pp_direct(const LeafArgs& a, bool do_phi, hipStream_t s);
hipError_t launch_pc(const LeafArgs& a, bool ewald, hipStream_t s);
hipError_t launch_pp_ewald(const LeafArgs& a, hipStream_t s);
hipError_t launch_cc(const NodeArgs& a, bool ewald, hipStream_t s);
hipError_t launch_cp(const NodeArgs& a, bool ewald, hipStream_t s);

}  // namespace tg
