// tg_walk.hpp -- the dual-tree walk (interaction lists) on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "tg_tree.hpp"

namespace tg {

// Tree::walk on the device: fills T.lists and T.census with exactly the lists the host walk
// builds (same entries, same order).  Returns hipSuccess or the first HIP error; err names it.
hipError_t device_walk(Tree& T, int device, hipStream_t s, std::string& err);

}  // namespace tg
