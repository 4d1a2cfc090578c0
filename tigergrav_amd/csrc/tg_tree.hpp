// tg_tree.hpp -- host-side restatement of tigergrav's tree (the caller of the kernels).
//
// Builds the kd-tree (src/tree.cpp:122-201, src/part_vect.cpp:895-933, src/range.cpp:134-169),
// its multipoles (src/tree.cpp:203-296, src/part_vect.cpp:542-635, tigergrav/multipole.hpp:
// 117-135) and the interaction lists of the dual-tree walk (src/tree.cpp:378-635), with the same
// floating-point steps so the lists -- membership and order -- are those the reference builds.
// Single process, no HPX; subtrees of the walk run as OpenMP tasks.
#pragma once
#include <stdint.h>

#include <memory>
#include <new>
#include <utility>
#include <vector>

namespace tg {

// An allocator whose value-less construct default-initialises: resizing a vector of a trivial type
// leaves the new elements unwritten, so a large host array is first touched (page-faulted) by the
// parallel loop that fills it rather than zeroed on one thread.
template <class T>
struct NoInit : std::allocator<T> {
	template <class U>
	struct rebind {
		using other = NoInit<U>;
	};
	NoInit() = default;
	template <class U>
	NoInit(const NoInit<U>&) {}
	template <class U>
	void construct(U* p) {
		::new (static_cast<void*>(p)) U;
	}
	template <class U, class... A>
	void construct(U* p, A&&... a) {
		::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
	}
};
template <class T>
using hvec = std::vector<T, NoInit<T>>;

using boxid_t = unsigned __int128;

struct TreeParams {
	int ewald = 1;
	double theta = 0.7;
	double soft_len = 0.0;
	int parts_per_node = 64;
};

struct Node {
	int32_t pbegin = 0, pend = 0;
	int32_t child[2] = {-1, -1};
	int32_t parent = -1;
	int32_t level = 0;
	boxid_t boxid = 1;
	// multipole_info (tigergrav/gravity.hpp:24-36)
	float m[17];
	float r = 0.0f;
	int32_t x[3] = {0, 0, 0};
	// particle range bound handed to the parent (src/tree.cpp:211-266)
	double rmin[3], rmax[3];
	uint64_t N = 0;
	bool leaf() const { return child[0] < 0; }
	bool empty() const { return pend == pbegin; }
};

// interaction lists in CSR form; node lists are indexed by node, leaf lists by node too (only
// leaves have entries).  Particle sources are [begin, end) ranges into the tree-ordered particles.
struct Lists {
	// per node (first MAC pass, src/tree.cpp:424-494)
	hvec<int64_t> cc_d_ptr, cp_d_ptr, cc_e_ptr, cp_e_ptr;
	hvec<int32_t> cc_d, cc_e;      // source node ids
	hvec<int32_t> cp_d, cp_e;      // ranges, 2 ints each
	// per leaf (drain loops, src/tree.cpp:530-602)
	hvec<int64_t> pc_d_ptr, pp_d_ptr, pc_e_ptr, pp_e_ptr;
	hvec<int32_t> pc_d, pc_e;      // source node ids
	hvec<int32_t> pp_d, pp_e;      // ranges, 2 ints each
};

// interaction census as the reference counts it (src/tree.cpp:483-494,618-628)
struct Census {
	uint64_t PP_direct = 0, PC_direct = 0, CC_direct = 0, CP_direct = 0;
	uint64_t PP_ewald = 0, PC_ewald = 0, CC_ewald = 0, CP_ewald = 0;
};

// Nodes at this level are not split further.  A box there is 2^-20 of the unit cell per side, so only
// particles that (nearly) coincide reach it: more than parts_per_node of them at one position would
// otherwise be split for ever (every split leaves them all on one side), in the reference too.  They
// stay one leaf; the kernels take its sinks 64 at a time.
constexpr int MAX_SPLIT_LEVEL = 60;

class Tree {
public:
	TreeParams P;
	std::vector<int32_t> pos;   // 3 per particle, tree order
	std::vector<float> mass;    // per particle, tree order
	std::vector<int64_t> order; // order[i] = input index of tree-order particle i
	std::vector<Node> nodes;    // nodes[0] = root; children after parents
	std::vector<int32_t> leaves;
	double m = 0.0;             // particle mass the reference assumes (m_tot / problem_size)
	Lists lists;
	Census census;
	int max_level = 0;

	// build the tree from positions (3 per particle) -- refine until every leaf holds at most
	// parts_per_node particles
	void build(const int32_t* x, const float* masses, int64_t n, const TreeParams& p, double particle_mass);
	// P2M / M2M / radii, bottom up
	void compute_multipoles();
	// new positions / masses (tree order) with the tree, lists and expansion centres kept: the
	// moments are recomputed about the existing centres (leaf P2M, then M2M upward) exactly as
	// compute_multipoles forms them; radii, boxes and lists are left as they are
	void update_moments(const int32_t* pos_tree, const float* mass_tree);
	// dual-tree walk from the root with the reference's MAC; fills `lists` and `census`
	void walk();
	// per node, the span of its particles' bounding box (the walk's sink rmax)
	std::vector<float> sink_spans() const;
	// the interaction census of `lists`
	void count_census();

private:
	void refine(int node, int first_child);
	void multipoles(int node);
	struct Scratch;
	void walk_node(int node, std::vector<int32_t>& dlist, std::vector<int32_t>& elist, std::vector<std::vector<int32_t>>* per_node);
};

// reference conversions (tigergrav/position.hpp:32-51)
inline double pos_to_double(int32_t x) { return double(x) * (1.0 / 4294967296.0) + 0.5; }
// a - b in position units between a particle and a node centre: the int32 wrap-around difference
// when periodic, else the exact difference (equal for a particle inside the node)
inline int64_t centre_offset_host(int32_t a, int32_t b, bool wrap) {
	return wrap ? int64_t(int32_t(uint32_t(a) - uint32_t(b))) : int64_t(a) - int64_t(b);
}
int32_t double_to_pos(double x);
// box_id_to_range (src/range.cpp:134-169)
void box_id_to_range(boxid_t id, double* mn, double* mx);

}  // namespace tg
