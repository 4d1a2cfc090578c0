// tg_walk.hpp -- the dual-tree walk (interaction lists) on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "tg_tree.hpp"

namespace tg {

// a node as the device walk reads it
struct WalkNode {
	int32_t pbegin, pend, c0, c1;
	int32_t x[3];
	float r;
	float rmax;  // span of the node's particle box (the sink rmax of the Ewald self test)
};

// fills out[i] for every node of T (host, OpenMP; tg_upload.cpp)
void walk_nodes(const Tree& T, WalkNode* out);

// Tree::walk on the device: fills T.lists and T.census with exactly the lists the host walk
// builds (same entries, same order).  Returns hipSuccess or the first HIP error; err names it.
// t, if given, receives the seconds spent in set-up, the walk steps, list assembly with the copy
// to the host, and the census.  Lists k (in the order cc_d, cp_d, cc_e, cp_e, pc_d, pp_d, pc_e,
// pp_e) with bit k of keep_mask set stay on the device: kept[k] receives the device array of
// their values (the caller frees it with hipFree, whatever the return status) and the host
// vector of values is left empty; their pointer arrays are copied to the host as for the others.
hipError_t device_walk(Tree& T, int device, hipStream_t s, std::string& err, double* t = nullptr, unsigned keep_mask = 0,
		int32_t** kept = nullptr);

}  // namespace tg
