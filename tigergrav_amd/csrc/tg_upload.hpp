// tg_upload.hpp -- host-side preparation of a solver's device work (tg_upload.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "tg_kernels.hpp"
#include "tg_tree.hpp"

namespace tg {

// a run of interaction-list entries expanded on the device: list[dst + l] = value + step * l, l < len
struct ListRun {
	int32_t value;
	int32_t len;
	int64_t dst;
	ListRun() = default;
	ListRun(int32_t v, int32_t l, int64_t d) : value(v), len(l), dst(d) {}
};

// everything a solver uploads, in host memory
struct HostWork {
	hvec<int4> parts;      // (x, y, z, mass bits), tree order
	hvec<MultiSrc> multis;
	hvec<int4> npos;       // node centres
	hvec<int4> items[NKINDS];        // leaf kinds: (first sink, count | extra, list begin, end)
	hvec<int> lists[NKINDS];         // lists built on the host (PC kinds)
	hvec<ListRun> runs[NKINDS];      // range kinds (CP, PP): runs the device expands
	hvec<ListRun> owner_runs[NKINDS];  // node kinds: partial-slot owner of every entry, as runs
	hvec<int2> waves[NKINDS];        // node kinds: [begin, end) entry runs of <= 64
	int64_t list_len[NKINDS] = {};
	hvec<int> slot_node;
	hvec<int64_t> slot_ptr;
	hvec<int4> leaf_items;  // (node, first particle, end, root-path offset)
	hvec<int> path, level;
	int64_t n_slots = 0;
};

// node kinds: list entries per partial slot at most (a chunk longer than 64 is one wave whose lanes
// take its entries 64 apart)
constexpr int NODE_CHUNK_MAX = 4096;

// node kinds' chunks are packed into waves greedily, the packing restarted every WAVE_SEGMENT
// chunks (host and device pack alike)
constexpr int WAVE_SEGMENT = 256;

// the tree's part of W: parts, multis, npos, leaf_items, path, level; TG_OK or a negative status
int prepare_tree_work(const Tree& T, int64_t n, HostWork& W, std::string& err);

// builds W from the tree and its lists (n particles; node_chunk entries per CC/CP partial slot);
// TG_OK or a negative status with err set
int prepare_work(const Tree& T, int64_t n, int ewald, double soft_len, int node_chunk, HostWork& W, std::string& err);

}  // namespace tg
