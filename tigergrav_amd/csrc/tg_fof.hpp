// tg_fof.hpp -- friends-of-friends groups on the device (tg_fof.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tg_fof.h"
#include "tg_tree.hpp"

namespace tg {

// group of every particle of T (tree order) into group_tree (host, n entries): the smallest
// tree-order index of its friends-of-friends component, TG_FOF_NONE for a particle linked to no
// other; d_parts are T's particles on the device (x, y, z, mass bits, tree order)
hipError_t fof_device(const Tree& T, const int4* d_parts, double link_len, int device, hipStream_t s, uint64_t* group_tree,
		int64_t* n_candidates, std::string& err);

// the catalog of the groups in group_tree (tg_groups.cpp): groups of at least min_n members, by id
int group_catalog(const Tree& T, const std::vector<uint64_t>& group_tree, const float* v_input, double a, int32_t min_n, tg_group* out,
		int64_t capacity, int64_t* n_out);

}  // namespace tg
