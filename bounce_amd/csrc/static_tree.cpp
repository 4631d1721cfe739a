// Host construction of the mesh midphase tree with the node numbering and triangle order
// of b3StaticTree::Build (static_tree.cpp:42-180): one triangle per leaf, 2n - 1 nodes,
// split on the longest axis of the node box at the first centroid past the box centre
// after sorting the node's triangles by centroid (Lomuto partition, last element as
// pivot), falling back to the middle index when one side would be empty.  The device
// query walks these nodes depth first, so equal numbering gives equal triangle order.
#include "static_tree.h"

#include <vector>

namespace b3amd {
namespace {

struct Box {
    float lo[3], hi[3];
};

inline float centroid_axis(const Box& b, int axis) {
    // b3AABB3::Centroid = 0.5f * (lower + upper) per component (aabb3.h:100-103).
    return 0.5f * (b.lo[axis] + b.hi[axis]);
}

inline int longest_axis(const Box& b) {  // aabb3.h:136-154
    float r[3];
    for (int i = 0; i < 3; ++i) r[i] = b.hi[i] - centroid_axis(b, i);
    int axis = 0;
    float mx = r[0];
    if (r[1] > mx) {
        mx = r[1];
        axis = 1;
    }
    if (r[2] > mx) axis = 2;
    return axis;
}

// In-place Lomuto quicksort by centroid; sub-ranges go to an explicit stack (the order in
// which the two disjoint halves are finished does not affect the result).
void sort_by_centroid(const std::vector<Box>& boxes, int axis, uint32_t* ids, uint32_t count,
                      std::vector<uint32_t>& stack) {
    stack.clear();
    stack.push_back(0);
    stack.push_back(count);
    while (!stack.empty()) {
        uint32_t n = stack.back();
        stack.pop_back();
        uint32_t first = stack.back();
        stack.pop_back();
        if (n <= 1) continue;
        uint32_t* p = ids + first;
        const uint32_t pivot = p[n - 1];
        const float pc = centroid_axis(boxes[pivot], axis);
        uint32_t low = 0;
        for (uint32_t i = 0; i + 1 < n; ++i) {
            if (centroid_axis(boxes[p[i]], axis) < pc) {
                uint32_t t = p[i];
                p[i] = p[low];
                p[low] = t;
                ++low;
            }
        }
        p[n - 1] = p[low];
        p[low] = pivot;
        stack.push_back(first);
        stack.push_back(low);
        stack.push_back(first + low + 1);
        stack.push_back(n - 1 - low);
    }
}

}  // namespace

bool build_static_tree(const float* vertices, uint32_t vertexCount, const uint32_t* triangles,
                       uint32_t triangleCount, std::vector<TreeNode>& nodes) {
    nodes.clear();
    if (triangleCount == 0) return true;
    std::vector<Box> boxes(triangleCount);
    for (uint32_t t = 0; t < triangleCount; ++t) {
        const uint32_t* tri = triangles + 3 * (size_t)t;
        if (tri[0] >= vertexCount || tri[1] >= vertexCount || tri[2] >= vertexCount) return false;
        const float* a = vertices + 3 * (size_t)tri[0];
        const float* b = vertices + 3 * (size_t)tri[1];
        const float* c = vertices + 3 * (size_t)tri[2];
        for (int k = 0; k < 3; ++k) {
            // b3AABB3::Set(v1, v2, v3): min(v1, min(v2, v3)) (aabb3.h:64-68), b3Min = a < b ? a : b.
            float m23 = b[k] < c[k] ? b[k] : c[k];
            float M23 = b[k] > c[k] ? b[k] : c[k];
            boxes[t].lo[k] = a[k] < m23 ? a[k] : m23;
            boxes[t].hi[k] = a[k] > M23 ? a[k] : M23;
        }
    }
    nodes.assign(2 * (size_t)triangleCount - 1, TreeNode{});
    std::vector<uint32_t> ids(triangleCount), sortStack;
    for (uint32_t i = 0; i < triangleCount; ++i) ids[i] = i;
    struct Work {
        uint32_t node, first, count;
    };
    std::vector<Work> work;
    work.push_back(Work{0, 0, triangleCount});
    uint32_t nodeCount = 1;
    while (!work.empty()) {
        Work w = work.back();
        work.pop_back();
        uint32_t* p = ids.data() + w.first;
        Box box = boxes[p[0]];
        for (uint32_t i = 1; i < w.count; ++i) {  // b3Combine: component min / max
            const Box& o = boxes[p[i]];
            for (int k = 0; k < 3; ++k) {
                box.lo[k] = box.lo[k] < o.lo[k] ? box.lo[k] : o.lo[k];
                box.hi[k] = box.hi[k] > o.hi[k] ? box.hi[k] : o.hi[k];
            }
        }
        TreeNode& nd = nodes[w.node];
        for (int k = 0; k < 3; ++k) {
            nd.lower[k] = box.lo[k];
            nd.upper[k] = box.hi[k];
        }
        if (w.count <= 1) {
            nd.child1 = kNullNode;
            nd.child2 = p[0];
            continue;
        }
        const int axis = longest_axis(box);
        const float split = centroid_axis(box, axis);
        sort_by_centroid(boxes, axis, p, w.count, sortStack);
        uint32_t middle = 0;
        const uint32_t right = w.count - 1;
        while (middle < right) {
            if (centroid_axis(boxes[p[middle]], axis) > split) break;
            ++middle;
        }
        if (middle == 0 || middle == w.count) middle = right / 2;
        const uint32_t c1 = nodeCount++, c2 = nodeCount++;
        nd.child1 = c1;
        nd.child2 = c2;
        // child1's subtree is numbered before child2's: finish it first.
        work.push_back(Work{c2, w.first + middle, w.count - middle});
        work.push_back(Work{c1, w.first, middle});
    }
    return nodeCount == nodes.size();
}

}  // namespace b3amd
