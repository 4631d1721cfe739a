// Static bounding-volume tree over mesh triangles (the b3StaticTree layout).
#ifndef B3AMD_STATIC_TREE_H
#define B3AMD_STATIC_TREE_H

#include <cstdint>
#include <vector>

namespace b3amd {

constexpr uint32_t kNullNode = 0xFFFFFFFFu;  // B3_NULL_NODE_S (static_tree.h:25)

// b3StaticTree::b3Node (static_tree.h:62-77): box, child1, child2 or the triangle index
// of a leaf (child1 == kNullNode).  32 bytes, uploaded as is.
struct TreeNode {
    float lower[3];
    float upper[3];
    uint32_t child1;
    uint32_t child2;
};
static_assert(sizeof(TreeNode) == 32, "tree node layout");

// Builds 2 * triangleCount - 1 nodes; false on an out-of-range vertex index.
bool build_static_tree(const float* vertices, uint32_t vertexCount, const uint32_t* triangles,
                       uint32_t triangleCount, std::vector<TreeNode>& nodes);

}  // namespace b3amd

#endif
