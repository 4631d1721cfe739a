// The step before the narrowphase on the device: b3BroadPhase (broad_phase.h:36-119,
// broad_phase.cpp:78-183) over fat AABBs kept in HBM.
//
// What the reference's pair order depends on, and how it is kept:
//
//  * Proxy ids are b3DynamicTree node ids (broad_phase.cpp:78-90, dynamic_tree.cpp:41-114).
//    Moving a proxy frees its parent node and allocates a new parent at once
//    (UpdateNode = RemoveLeaf + InsertLeaf, dynamic_tree.cpp:116-126, 199-308), so moves
//    never change the node pool's free list; creates and destroys do.  A destroy pushes
//    the leaf's parent (when the tree had two or more leaves) and then the leaf; a create
//    pops the leaf and then, when the tree is not empty, a parent.  Counting proxies, the
//    top group of the free list was always pushed by the destroy that left the current
//    count, so it holds exactly the node(s) the next create pops: a create's leaf id is
//    the last destroyed leaf still on the list, or the next never-used node.  The host
//    keeps that list of groups (leaf id, parent present) and the never-used counter, which
//    reproduces every proxy id without the tree.
//  * FindPairs (broad_phase.h:160-208) reports every pair (min, max) of a moved proxy and a
//    different proxy whose fat AABBs overlap (b3TestOverlap, aabb3.h:247-251), sorted by
//    (proxy1, proxy2) with duplicates removed.  The tree only decides the order in which
//    the query meets them, so any exact overlap search gives the same list.  Here a
//    linear BVH (Morton order of the fat AABB centres, Karras' radix-tree build, bottom-up
//    refit with exact min/max, stackless traversal over escape links) is rebuilt at each
//    call; each moved proxy walks it and emits a pair once (from the lower id when both
//    moved); one radix sort orders the keys (proxy1 << B | proxy2).
//  * MoveProxy (broad_phase.cpp:99-149): a proxy whose fat AABB does not contain the new
//    AABB gets the AABB extended by b3_aabbExtension and by b3_aabbMultiplier times the
//    displacement on the side it points to, and is marked moved.  Each proxy depends only
//    on its own inputs, so one lane per proxy is exact.
#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "broadphase.h"

namespace b3amd {

__global__ __launch_bounds__(256) void bp_edit_kernel(uint32_t n, const BpEdit* edits, BpState st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const BpEdit e = edits[i];
    if (e.mode == kEditCreate) {
        st.lower[e.id] = make_float4(e.lower[0], e.lower[1], e.lower[2], 0.0f);
        st.upper[e.id] = make_float4(e.upper[0], e.upper[1], e.upper[2], 0.0f);
        st.user[e.id] = e.user;
        st.flags[e.id] = kLive | kMoved;  // BufferMove (broad_phase.cpp:87)
    } else if (e.mode == kEditDestroy) {
        st.flags[e.id] = 0u;  // UnbufferMove + RemoveNode (broad_phase.cpp:92-97)
    } else {
        st.flags[e.id] |= kMoved;  // TouchProxy (broad_phase.cpp:151-154)
    }
}

// CreateProxy (broad_phase.cpp:78-90) of proxy records[2i] with user data records[2i+1]
// and the AABB aabbs[i] ([n][6]), fattened by b3_aabbExtension (aabb3.h:85-90).
__global__ __launch_bounds__(256) void bp_create_kernel(uint32_t n, const uint32_t* records, const float* aabbs,
                                                        BpState st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = records[2 * i];
    const float* a = aabbs + 6ull * i;
    st.lower[id] = make_float4(a[0] - kAabbExtension, a[1] - kAabbExtension, a[2] - kAabbExtension, 0.0f);
    st.upper[id] = make_float4(a[3] + kAabbExtension, a[4] + kAabbExtension, a[5] + kAabbExtension, 0.0f);
    st.user[id] = records[2 * i + 1];
    st.flags[id] = kLive | kMoved;  // BufferMove (broad_phase.cpp:87)
}

// MoveProxy (broad_phase.cpp:99-149) for proxies ids[i].
__global__ __launch_bounds__(256) void bp_move_kernel(uint32_t n, const uint32_t* ids, const float* aabbs,
                                                      const float* disp, BpState st, uint32_t idCap,
                                                      uint32_t* moved, uint32_t* status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = ids[i];
    if (id >= idCap || !(st.flags[id] & kLive)) {  // not a live proxy: nothing moves
        if (moved) moved[i] = 0u;
        atomicMax(status, i + 1u);
        return;
    }
    const float* a = aabbs + 6 * (size_t)i;
    const float* dp = disp + 3 * (size_t)i;
    const bool m = move_proxy(st, id, V3{a[0], a[1], a[2]}, V3{a[3], a[4], a[5]}, V3{dp[0], dp[1], dp[2]});
    if (moved) moved[i] = m ? 1u : 0u;
}

// Float bits mapped so that unsigned order is float order (for atomic min/max).
__device__ __forceinline__ uint32_t order_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_order_bits(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Wave-aggregated append: returns this lane's slot for one item (if want) in *counter.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
    const uint64_t m = __ballot(want);
    if (!m) return 0u;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// Live proxies into `leaves`, moved live proxies into `queries`; bounds of the live fat AABB
// centres into bounds[0..5] (order bits: min x y z, max x y z).
__global__ __launch_bounds__(256) void bp_lists_kernel(uint32_t idCap, BpState st, uint32_t* leaves,
                                                       uint32_t* queries, uint32_t leafCap, uint32_t* counts,
                                                       uint32_t* bounds) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t f = id < idCap ? st.flags[id] : 0u;
    const bool live = (f & kLive) != 0u;
    const bool moved = live && (f & kMoved) != 0u;
    const uint32_t li = wave_append(counts + 0, live);
    const uint32_t qi = wave_append(counts + 1, moved);
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (live && li < leafCap) {
        leaves[li] = id;
        const float4 L = st.lower[id], U = st.upper[id];
        cx = 0.5f * (L.x + U.x), cy = 0.5f * (L.y + U.y), cz = 0.5f * (L.z + U.z);
    }
    if (moved && qi < leafCap) queries[qi] = id;
    uint32_t mn[3] = {live ? order_bits(cx) : 0xFFFFFFFFu, live ? order_bits(cy) : 0xFFFFFFFFu,
                      live ? order_bits(cz) : 0xFFFFFFFFu};
    uint32_t mx[3] = {live ? order_bits(cx) : 0u, live ? order_bits(cy) : 0u, live ? order_bits(cz) : 0u};
    for (int o = 32; o > 0; o >>= 1)
        for (int k = 0; k < 3; ++k) {
            mn[k] = min(mn[k], (uint32_t)__shfl_xor((int)mn[k], o));
            mx[k] = max(mx[k], (uint32_t)__shfl_xor((int)mx[k], o));
        }
    if (__lane_id() == 0u && __ballot(live)) {
        for (int k = 0; k < 3; ++k) {
            atomicMin(bounds + k, mn[k]);
            atomicMax(bounds + 3 + k, mx[k]);
        }
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t quant10(float c, float lo, float hi) {
    const float ext = hi - lo;
    float t = ext > 0.0f ? (c - lo) / ext : 0.0f;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return min((uint32_t)(t * 1024.0f), 1023u);
}

// Key of leaf i: 30-bit Morton code of its centre << 32 | proxy id (unique keys).
__global__ __launch_bounds__(256) void bp_morton_kernel(uint32_t n, const uint32_t* leaves, BpState st,
                                                        const uint32_t* bounds, uint64_t* keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = leaves[i];
    const float4 L = st.lower[id], U = st.upper[id];
    const float cx = 0.5f * (L.x + U.x), cy = 0.5f * (L.y + U.y), cz = 0.5f * (L.z + U.z);
    const float lo[3] = {from_order_bits(bounds[0]), from_order_bits(bounds[1]), from_order_bits(bounds[2])};
    const float hi[3] = {from_order_bits(bounds[3]), from_order_bits(bounds[4]), from_order_bits(bounds[5])};
    const uint32_t m = (spread10(quant10(cx, lo[0], hi[0])) << 2) | (spread10(quant10(cy, lo[1], hi[1])) << 1) |
                       spread10(quant10(cz, lo[2], hi[2]));
    keys[i] = ((uint64_t)m << 32) | id;
}

// Node numbering: internal nodes 0 .. n-2 (root 0), leaf j at n-1+j.
struct Lbvh {
    uint32_t n;
    const uint64_t* keys;  // sorted
    uint32_t* left;        // [n-1] child node numbers
    uint32_t* right;       // [n-1]
    uint32_t* parent;      // [2n-1]
    uint32_t* escape;      // [2n-1]
    float4* lo;            // [n-1] internal node boxes
    float4* hi;
    uint32_t* first;       // [n-1] first and last sorted leaf under each internal node
    uint32_t* last;
    // Refit scratch: per sorted leaf, the box of the leaves from its refit block's start to
    // it (pre) and from it to the block's end (suf); a sparse table over the blocks' boxes
    // (level k, block j: blocks j .. j + 2^k - 1) at blkLo/blkHi[k * blkStride + j].
    float4 *preLo, *preHi, *sufLo, *sufHi;
    float4 *blkLo, *blkHi;
    uint32_t blkStride;
};

__device__ __forceinline__ int common_prefix(const uint64_t* k, uint32_t n, int i, int j) {
    if (j < 0 || j >= (int)n) return -1;
    return __clzll((long long)(k[i] ^ k[j]));
}

// Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees, and k-d trees".
__global__ __launch_bounds__(256) void bp_build_kernel(Lbvh t) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = (int)t.n;
    if (i >= n - 1) return;
    const uint64_t* k = t.keys;
    const int d = common_prefix(k, n, i, i + 1) - common_prefix(k, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = common_prefix(k, n, i, i - d);
    int lmax = 2;
    while (common_prefix(k, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int s = lmax / 2; s >= 1; s /= 2)
        if (common_prefix(k, n, i, i + (l + s) * d) > dmin) l += s;
    const int j = i + l * d;
    const int dnode = common_prefix(k, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int step = (l + div - 1) / div;
        if (common_prefix(k, n, i, i + (s + step) * d) > dnode) s += step;
        if (step <= 1) break;
    }
    const int gamma = i + s * d + min(d, 0);
    const uint32_t first = (uint32_t)min(i, j), last = (uint32_t)max(i, j);
    const uint32_t lc = first == (uint32_t)gamma ? (uint32_t)(n - 1 + gamma) : (uint32_t)gamma;
    const uint32_t rc = last == (uint32_t)gamma + 1u ? (uint32_t)(n - 1 + gamma + 1) : (uint32_t)(gamma + 1);
    t.left[i] = lc;
    t.right[i] = rc;
    t.parent[lc] = (uint32_t)i;
    t.parent[rc] = (uint32_t)i;
    t.first[i] = first;
    t.last[i] = last;
    if (i == 0) t.parent[0] = kNone;
}

__device__ __forceinline__ void node_box(const Lbvh& t, const BpState& st, uint32_t node, float4& l, float4& u) {
    if (node >= t.n - 1u) {
        const uint32_t id = (uint32_t)t.keys[node - (t.n - 1u)];
        l = st.lower[id];
        u = st.upper[id];
    } else {
        l = t.lo[node];
        u = t.hi[node];
    }
}

// Refit: an internal node's box is the exact min/max over the leaves of its range, which
// is order-free, so it is read off range-minimum tables instead of climbing the tree (a
// climb passes every level through device-scope atomics and fences).  Leaves are taken
// in blocks of kRefitBlock in sorted order:
//   * bp_refit_block_kernel: a sparse table of the block's leaf boxes in LDS gives the box
//     of every internal node whose range lies in the block, the box of each leaf's prefix
//     and suffix within the block, and the block's box;
//   * bp_refit_table_kernel: a sparse table over the blocks' boxes (one workgroup);
//   * bp_refit_cross_kernel: a node whose range crosses blocks combines the suffix in its
//     first block, the whole blocks between (two table entries) and the prefix in its last.
constexpr uint32_t kRefitBlock = 256;
constexpr uint32_t kRefitLevels = 8;  // in-block table levels: spans 1 .. 128

struct Box6 {
    float4 l, u;
};

__device__ __forceinline__ Box6 join(const Box6& a, const Box6& b) {
    return Box6{make_float4(fminf(a.l.x, b.l.x), fminf(a.l.y, b.l.y), fminf(a.l.z, b.l.z), 0.0f),
                make_float4(fmaxf(a.u.x, b.u.x), fmaxf(a.u.y, b.u.y), fmaxf(a.u.z, b.u.z), 0.0f)};
}

__device__ __forceinline__ uint32_t floor_log2(uint32_t v) { return 31u - (uint32_t)__clz(v); }

__global__ __launch_bounds__(kRefitBlock) void bp_refit_block_kernel(Lbvh t, BpState st) {
    __shared__ float tab[kRefitLevels][6][kRefitBlock];
    const uint32_t tid = threadIdx.x, b0 = blockIdx.x * kRefitBlock;
    const uint32_t cnt = t.n - b0 < kRefitBlock ? t.n - b0 : kRefitBlock;
    if (tid < cnt) {
        const uint32_t id = (uint32_t)t.keys[b0 + tid];
        const float4 l = st.lower[id], u = st.upper[id];
        tab[0][0][tid] = l.x, tab[0][1][tid] = l.y, tab[0][2][tid] = l.z;
        tab[0][3][tid] = u.x, tab[0][4][tid] = u.y, tab[0][5][tid] = u.z;
    }
    for (uint32_t k = 1; k < kRefitLevels; ++k) {
        __syncthreads();
        const uint32_t half = 1u << (k - 1u);
        if (tid + 2u * half <= cnt) {
#pragma unroll
            for (int c = 0; c < 3; ++c) tab[k][c][tid] = fminf(tab[k - 1][c][tid], tab[k - 1][c][tid + half]);
#pragma unroll
            for (int c = 3; c < 6; ++c) tab[k][c][tid] = fmaxf(tab[k - 1][c][tid], tab[k - 1][c][tid + half]);
        }
    }
    __syncthreads();
    // Box of block-relative leaves f .. l (f <= l < cnt): two windows of the largest span
    // that fits cover the range.
    auto range = [&](uint32_t f, uint32_t l) {
        uint32_t k = floor_log2(l - f + 1u);
        if (k >= kRefitLevels) k = kRefitLevels - 1u;
        const uint32_t g = l + 1u - (1u << k);
        Box6 r;
        r.l = make_float4(fminf(tab[k][0][f], tab[k][0][g]), fminf(tab[k][1][f], tab[k][1][g]),
                          fminf(tab[k][2][f], tab[k][2][g]), 0.0f);
        r.u = make_float4(fmaxf(tab[k][3][f], tab[k][3][g]), fmaxf(tab[k][4][f], tab[k][4][g]),
                          fmaxf(tab[k][5][f], tab[k][5][g]), 0.0f);
        return r;
    };
    if (tid < cnt) {
        const Box6 p = range(0u, tid), q = range(tid, cnt - 1u);
        t.preLo[b0 + tid] = p.l, t.preHi[b0 + tid] = p.u;
        t.sufLo[b0 + tid] = q.l, t.sufHi[b0 + tid] = q.u;
    }
    if (tid == 0) {
        const Box6 a = range(0u, cnt - 1u);
        t.blkLo[blockIdx.x] = a.l, t.blkHi[blockIdx.x] = a.u;
    }
    // Internal node i has i at one end of its range, so a range inside this block has its
    // node here.
    const uint32_t i = b0 + tid;
    if (i + 1u < t.n) {
        const uint32_t f = t.first[i], l = t.last[i];
        if (f >= b0 && l < b0 + cnt) {
            const Box6 r = range(f - b0, l - b0);
            t.lo[i] = r.l, t.hi[i] = r.u;
        }
    }
}

__global__ __launch_bounds__(1024) void bp_refit_table_kernel(Lbvh t, uint32_t nb) {
    for (uint32_t k = 1; (1u << k) <= nb; ++k) {
        __syncthreads();
        const uint32_t half = 1u << (k - 1u);
        const float4* pl = t.blkLo + (k - 1u) * t.blkStride;
        const float4* pu = t.blkHi + (k - 1u) * t.blkStride;
        for (uint32_t j = threadIdx.x; j + 2u * half <= nb; j += blockDim.x) {
            const Box6 r = join(Box6{pl[j], pu[j]}, Box6{pl[j + half], pu[j + half]});
            t.blkLo[k * t.blkStride + j] = r.l;
            t.blkHi[k * t.blkStride + j] = r.u;
        }
    }
}

__global__ __launch_bounds__(256) void bp_refit_cross_kernel(Lbvh t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1u >= t.n) return;
    const uint32_t f = t.first[i], l = t.last[i];
    const uint32_t bf = f / kRefitBlock, bl = l / kRefitBlock;
    if (bf == bl) return;
    Box6 r = join(Box6{t.sufLo[f], t.sufHi[f]}, Box6{t.preLo[l], t.preHi[l]});
    if (bl > bf + 1u) {
        const uint32_t a = bf + 1u, c = bl - 1u, k = floor_log2(c - a + 1u), g = c + 1u - (1u << k);
        const uint32_t o = k * t.blkStride;
        r = join(r, join(Box6{t.blkLo[o + a], t.blkHi[o + a]}, Box6{t.blkLo[o + g], t.blkHi[o + g]}));
    }
    t.lo[i] = r.l;
    t.hi[i] = r.u;
}

// escape[x]: the node after x's subtree in depth-first order (kNone after the last).
__global__ __launch_bounds__(256) void bp_escape_kernel(Lbvh t) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= 2u * t.n - 1u) return;
    uint32_t c = x, e = kNone;
    for (uint32_t up = 0; c != 0u && up < t.n; ++up) {
        const uint32_t p = t.parent[c];
        if (t.left[p] == c) {
            e = t.right[p];
            break;
        }
        c = p;
    }
    t.escape[x] = e;
}

// A pair (p, q) of the moved proxy q and the leaf's proxy p, kept once per unordered pair:
// from the lower id when both moved.
__device__ __forceinline__ bool query_pair(const BpState& st, uint32_t p, uint32_t q, uint32_t bits, uint64_t& key) {
    if (p == q || ((st.flags[p] & kMoved) && p < q)) return false;
    key = ((uint64_t)min(p, q) << bits) | max(p, q);
    return true;
}

// The subtree of `root` walked by one lane over the escape links (the walk leaves the
// subtree at escape[root]), pairs appended one atomic each: the way out for a group whose
// stack of pending nodes is full.
__device__ void query_subtree(const Lbvh& t, const BpState& st, uint32_t root, uint32_t q, float4 ql, float4 qu,
                              uint32_t bits, uint64_t* out, uint32_t cap, uint32_t* count, uint32_t* status) {
    const uint32_t firstLeaf = t.n - 1u, end = t.escape[root];
    uint32_t node = root;
    for (uint32_t steps = 0; node != end; ++steps) {
        if (node == kNone || steps >= 2u * t.n) {  // a malformed tree: stop and report
            atomicMax(status, 1u);
            return;
        }
        float4 l, u;
        node_box(t, st, node, l, u);
        const bool hit = overlap(l, u, ql, qu);
        if (node >= firstLeaf) {
            uint64_t key;
            if (hit && query_pair(st, (uint32_t)t.keys[node - firstLeaf], q, bits, key)) {
                const uint32_t slot = atomicAdd(count, 1u);
                if (slot < cap) out[slot] = key;
            }
            node = t.escape[node];
        } else {
            node = hit ? t.left[node] : t.escape[node];
        }
    }
}

// kQueryLanes lanes per moved proxy q: every overlapping leaf p != q, emitted once per
// unordered pair.  The group keeps the nodes still to test on a stack in LDS; each round
// its lanes pop up to kQueryLanes of them, test their boxes, and push the children of the
// internal nodes hit, so a walk takes about as many rounds as the tree is deep instead of
// one round per node visited.  Children that do not fit on the stack are walked by their
// lane alone.  The pairs found do not depend on the order of the walk.
constexpr uint32_t kQueryLanes = 16;
constexpr uint32_t kQueryGroups = 256 / kQueryLanes;  // moved proxies per block
constexpr uint32_t kQueryStack = 128;                 // pending nodes per moved proxy
static_assert(kQueryLanes < 32 && 64 % kQueryLanes == 0 && kQueryStack >= 3 * kQueryLanes, "query groups");

__device__ __forceinline__ void query_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__global__ __launch_bounds__(256) void bp_query_kernel(Lbvh t, BpState st, uint32_t nq, const uint32_t* queries,
                                                       uint32_t bits, uint64_t* out, uint32_t cap,
                                                       uint32_t* count, uint32_t* status) {
    __shared__ uint32_t stacks[kQueryGroups][kQueryStack];
    const uint32_t g = threadIdx.x / kQueryLanes, r = threadIdx.x % kQueryLanes;
    const uint32_t gbase = (threadIdx.x % 64u) - r;  // the group's first lane in its wave
    const uint32_t qi = blockIdx.x * kQueryGroups + g;
    const bool active = qi < nq;
    const uint32_t q = active ? queries[qi] : 0u;
    float4 ql = make_float4(0, 0, 0, 0), qu = ql;
    if (active) ql = st.lower[q], qu = st.upper[q];
    uint32_t* stk = stacks[g];
    if (r == 0) stk[0] = 0u;  // node 0: the root (the single leaf when n == 1)
    uint32_t top = active ? 1u : 0u;  // the same on every lane of the group
    const uint32_t firstLeaf = t.n - 1u;
    query_lds_sync();
    // Every node is pushed at most once per proxy, so 2n - 1 rounds bound the walk.
    for (uint32_t round = 0; __ballot(top != 0u); ++round) {
        if (round >= 2u * t.n) {  // a malformed tree: stop and report
            if (top != 0u && r == 0) atomicMax(status, 1u);
            break;
        }
        const uint32_t k = top < kQueryLanes ? top : kQueryLanes;
        const uint32_t node = r < k ? stk[top - 1u - r] : kNone;
        top -= k;
        bool emit = false, push = false;
        uint64_t key = 0;
        uint32_t lc = 0, rc = 0;
        if (node != kNone) {
            float4 l, u;
            node_box(t, st, node, l, u);
            const bool hit = overlap(l, u, ql, qu);
            if (node >= firstLeaf) {
                emit = hit && query_pair(st, (uint32_t)t.keys[node - firstLeaf], q, bits, key);
            } else if (hit) {
                push = true;
                lc = t.left[node];
                rc = t.right[node];
            }
        }
        const uint32_t gm = (uint32_t)(__ballot(push) >> gbase) & ((1u << kQueryLanes) - 1u);
        const uint32_t before = (uint32_t)__popc(gm & ((1u << r) - 1u)), total = (uint32_t)__popc(gm);
        const uint32_t room = (kQueryStack - top) / 2u;
        const uint32_t fit = total < room ? total : room;
        query_lds_sync();  // this round's pops are read before its pushes overwrite them
        if (push && before < fit) {
            stk[top + 2u * before] = rc;
            stk[top + 2u * before + 1u] = lc;
        }
        top += 2u * fit;
        const uint32_t slot = wave_append(count, emit);
        if (emit && slot < cap) out[slot] = key;
        if (push && before >= fit) {
            query_subtree(t, st, lc, q, ql, qu, bits, out, cap, count, status);
            query_subtree(t, st, rc, q, ql, qu, bits, out, cap, count, status);
        }
        query_lds_sync();
    }
}

__global__ __launch_bounds__(256) void bp_pairs_kernel(uint32_t n, const uint64_t* keys, uint32_t bits,
                                                       const uint32_t* user, uint32_t* pairs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const uint32_t a = (uint32_t)(k >> bits), b = (uint32_t)(k & ((1ull << bits) - 1ull));
    reinterpret_cast<uint4*>(pairs)[i] = make_uint4(a, b, user[a], user[b]);
}

__global__ __launch_bounds__(256) void bp_clear_moved_kernel(uint32_t n, const uint32_t* queries, uint32_t* flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[queries[i]] &= ~kMoved;
}

__global__ __launch_bounds__(256) void bp_test_overlap_kernel(uint32_t n, const uint32_t* pairs, BpState st,
                                                              uint32_t idCap, uint32_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
    if (a >= idCap || b >= idCap || !(st.flags[a] & kLive) || !(st.flags[b] & kLive)) {
        out[i] = B3AMD_OVERLAP_INVALID;
        return;
    }
    out[i] = overlap(st.lower[a], st.upper[a], st.lower[b], st.upper[b]) ? 1u : 0u;
}

// The find-pairs count words: leaf, query and pair counts 0, the Morton bounds' minimum
// all ones and maximum 0 (word 9, the move status, is kept).
__global__ void bp_init_counts_kernel(uint32_t* counts) {
    const uint32_t i = threadIdx.x;
    if (i < 9u) counts[i] = (i >= 3u && i < 6u) ? 0xFFFFFFFFu : 0u;
}

__global__ void bp_fill_kernel(uint32_t n, uint32_t* p, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct ShapeAabbArgs {
    uint32_t n;
    uint32_t bodyCount;
    const uint32_t* shapes;
    const uint32_t* bodies;
    const float* xfs;     // [B][12]
    const float* sweeps;  // [B][18] b3Sweep, or null
    const ShapeRec* table;
    uint32_t shapeCount;
    const float4* pool;
    float* aabbs;         // [n][6]
    float* disp;          // [n][3] or null
    uint32_t* status;     // last invalid index + 1 (0: all valid)
};

__global__ __launch_bounds__(256) void shape_aabb_kernel(ShapeAabbArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t si = a.shapes[i];
    if (si >= a.shapeCount || a.table[si].type == B3AMD_MESH) {
        atomicMax(a.status, i + 1u);
        return;
    }
    const ShapeRec s = a.table[si];
    const uint32_t b = a.bodies[i];
    if (b >= a.bodyCount) {
        atomicMax(a.status, i + 1u);
        return;
    }
    const float* x = a.xfs + 12 * (size_t)b;
    Xf xf2;
    xf2.R = M33{V3{x[0], x[1], x[2]}, V3{x[3], x[4], x[5]}, V3{x[6], x[7], x[8]}};
    xf2.p = V3{x[9], x[10], x[11]};
    V3 lo, hi;
    shape_aabb(s, a.pool, xf2, lo, hi);
    V3 d{0.0f, 0.0f, 0.0f};
    if (a.sweeps) {
        const Xf xf1 = sweep_transform0(a.sweeps + 18 * (size_t)b);
        V3 lo1, hi1;
        shape_aabb(s, a.pool, xf1, lo1, hi1);
        lo = vmin(lo1, lo);  // b3Combine(aabb1, aabb2) (aabb3.h:238-244)
        hi = vmax(hi1, hi);
        d = xf2.p - xf1.p;
    }
    float* o = a.aabbs + 6 * (size_t)i;
    o[0] = lo.x, o[1] = lo.y, o[2] = lo.z, o[3] = hi.x, o[4] = hi.y, o[5] = hi.z;
    if (a.disp) {
        float* q = a.disp + 3 * (size_t)i;
        q[0] = d.x, q[1] = d.y, q[2] = d.z;
    }
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + 255) / 256); }
// Refit blocks of `cap` leaves (the row length of the block table, whose 32 levels cover
// any count below 2^32).
inline uint32_t refit_blocks(uint32_t cap) { return (cap + kRefitBlock - 1u) / kRefitBlock; }

}  // namespace b3amd

namespace {

int bp_grow_ids(b3amd_broadphase* bp, uint32_t need) {
    if (need <= bp->idCap) return B3AMD_OK;
    b3amd_context* ctx = bp->ctx;
    uint32_t cap = bp->idCap ? bp->idCap : 1024u;
    while (cap < need) cap *= 2u;
    BpState n{nullptr, nullptr, nullptr, nullptr};
    int rc;
    if ((rc = ctx->alloc(16ull * cap, (void**)&n.lower)) || (rc = ctx->alloc(16ull * cap, (void**)&n.upper)) ||
        (rc = ctx->alloc(4ull * cap, (void**)&n.user)) || (rc = ctx->alloc(4ull * cap, (void**)&n.flags)))
        return rc;
    hipStream_t s = ctx->stream;
    bp_fill_kernel<<<blocks(cap), 256, 0, s>>>(cap, n.flags, 0u);
    if (bp->idCap) {
        HIP_TRY(hipMemcpyAsync(n.lower, bp->st.lower, 16ull * bp->idCap, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(n.upper, bp->st.upper, 16ull * bp->idCap, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(n.user, bp->st.user, 4ull * bp->idCap, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(n.flags, bp->st.flags, 4ull * bp->idCap, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        ctx->release(bp->st.lower), ctx->release(bp->st.upper), ctx->release(bp->st.user),
            ctx->release(bp->st.flags);
    }
    bp->st = n;
    bp->idCap = cap;
    bp->live.resize(cap, 0u);
    return B3AMD_OK;
}

// The proxy id of the next create (b3DynamicTree::InsertNode: AllocateNode for the leaf,
// then InsertLeaf allocates a parent unless the tree is empty; dynamic_tree.cpp:102-114,
// 203-214), marked live.
int bp_take_id(b3amd_broadphase* bp, uint32_t* out) {
    uint32_t id;
    if (!bp->freeGroups.empty()) {
        const std::pair<uint32_t, bool> g = bp->freeGroups.back();
        if (g.second != (bp->count >= 1u)) return fail(B3AMD_ERR_ARGUMENT, "node pool out of step (internal)");
        bp->freeGroups.pop_back();
        id = g.first;
    } else {
        id = bp->fresh++;
        if (bp->count >= 1u) bp->fresh++;  // the new parent node
    }
    int rc = bp_grow_ids(bp, id + 1u);
    if (rc) return rc;
    bp->live[id] = 1u;
    bp->count++;
    *out = id;
    return B3AMD_OK;
}

int bp_apply_edits(b3amd_broadphase* bp, const std::vector<BpEdit>& edits) {
    if (edits.empty()) return B3AMD_OK;
    b3amd_context* ctx = bp->ctx;
    int rc;
    if (edits.size() > bp->editCap) {
        if ((rc = ctx->release(bp->dEdits))) return rc;
        bp->dEdits = nullptr;
        bp->editCap = 0;
        const uint32_t cap = std::max<uint32_t>(1024u, (uint32_t)edits.size());
        if ((rc = ctx->alloc(sizeof(BpEdit) * (uint64_t)cap, (void**)&bp->dEdits))) return rc;
        bp->editCap = cap;
    }
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(bp->dEdits, edits.data(), sizeof(BpEdit) * edits.size(), hipMemcpyHostToDevice, s));
    bp_edit_kernel<<<blocks(edits.size()), 256, 0, s>>>((uint32_t)edits.size(), bp->dEdits, bp->st);
    HIP_TRY(hipGetLastError());
    // The edit records are host memory that the copy reads before return.
    HIP_TRY(hipStreamSynchronize(s));
    return B3AMD_OK;
}

int bp_grow_leaves(b3amd_broadphase* bp, uint32_t need) {
    if (need <= bp->leafCap) return B3AMD_OK;
    b3amd_context* ctx = bp->ctx;
    uint32_t cap = std::max<uint32_t>(1024u, bp->leafCap);
    while (cap < need) cap *= 2u;
    void* ptrs[] = {bp->dLeaves, bp->dKeys, bp->dKeys2, bp->dNodes, bp->dBoxes};
    for (void* p : ptrs) ctx->release(p);
    int rc;
    if ((rc = ctx->alloc(8ull * cap, (void**)&bp->dLeaves)) || (rc = ctx->alloc(8ull * cap, (void**)&bp->dKeys)) ||
        (rc = ctx->alloc(8ull * cap, (void**)&bp->dKeys2)) || (rc = ctx->alloc(4ull * 8 * cap, (void**)&bp->dNodes)) ||
        (rc = ctx->alloc(16ull * (6ull * cap + 64ull * refit_blocks(cap)), (void**)&bp->dBoxes)))
        return rc;
    bp->dQueries = bp->dLeaves + cap;
    bp->leafCap = cap;
    return B3AMD_OK;
}

int bp_grow_pairs(b3amd_broadphase* bp, uint32_t need) {
    if (need <= bp->pairCap) return B3AMD_OK;
    b3amd_context* ctx = bp->ctx;
    uint32_t cap = std::max<uint32_t>(4096u, bp->pairCap);
    while (cap < need) cap *= 2u;
    ctx->release(bp->dPairKeys), ctx->release(bp->dPairKeys2), ctx->release(bp->dPairs);
    bp->dPairKeys = bp->dPairKeys2 = nullptr;
    bp->dPairs = nullptr;
    bp->pairCap = 0;
    int rc;
    if ((rc = ctx->alloc(8ull * cap, (void**)&bp->dPairKeys)) || (rc = ctx->alloc(8ull * cap, (void**)&bp->dPairKeys2)) ||
        (rc = ctx->alloc(16ull * cap, (void**)&bp->dPairs)))
        return rc;
    bp->pairCap = cap;
    return B3AMD_OK;
}

int bp_sort(b3amd_broadphase* bp, uint64_t* in, uint64_t* out, uint32_t n, uint32_t endBit) {
    size_t need = 0;
    if (rocprim::radix_sort_keys(nullptr, need, in, out, n, 0u, endBit, bp->ctx->stream) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "radix sort sizing failed");
    if (need > bp->sortTmpBytes) {
        bp->ctx->release(bp->dSortTmp);
        bp->dSortTmp = nullptr;
        bp->sortTmpBytes = 0;
        int rc = bp->ctx->alloc(need, &bp->dSortTmp);
        if (rc) return rc;
        bp->sortTmpBytes = need;
    }
    size_t have = bp->sortTmpBytes;
    if (rocprim::radix_sort_keys(bp->dSortTmp, have, in, out, n, 0u, endBit, bp->ctx->stream) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "radix sort failed");
    return B3AMD_OK;
}

uint32_t id_bits(uint32_t idCap) {
    uint32_t b = 1;
    while ((1ull << b) < idCap) ++b;
    return b;
}

}  // namespace

namespace b3amd {

int bp_create_proxies_device(b3amd_broadphase* bp, uint32_t n, const float* d_aabbs, const uint32_t* users,
                             uint32_t* ids, uint32_t* records, uint32_t* d_records) {
    for (uint32_t i = 0; i < n; ++i) {
        int rc = bp_take_id(bp, ids + i);
        if (rc) return rc;
        records[2 * i] = ids[i];
        records[2 * i + 1] = users[i];
    }
    hipStream_t s = bp->ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_records, records, 8ull * n, hipMemcpyHostToDevice, s));
    bp_create_kernel<<<blocks(n), 256, 0, s>>>(n, d_records, d_aabbs, bp->st);
    HIP_TRY(hipGetLastError());
    return B3AMD_OK;
}

int bp_find_pairs(b3amd_broadphase* bp, uint32_t* pair_count, const uint32_t** d_pairs, bool sync) {
    if (!bp || !pair_count) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_context* ctx = bp->ctx;
    hipStream_t s = ctx->stream;
    int rc;
    *pair_count = 0;
    if (d_pairs) *d_pairs = bp->dPairs;
    bp->lastPairs = 0;
    if (bp->count == 0) return B3AMD_OK;
    if ((rc = bp_grow_leaves(bp, bp->count))) return rc;
    bp_init_counts_kernel<<<1, 64, 0, s>>>(bp->dCounts);
    HIP_TRY(hipGetLastError());
    rc = timed(ctx, s, "bp_lists", [&] {
        bp_lists_kernel<<<blocks(bp->idCap), 256, 0, s>>>(bp->idCap, bp->st, bp->dLeaves, bp->dQueries,
                                                          bp->leafCap, bp->dCounts, bp->dCounts + 3);
        return hipGetLastError();
    });
    if (rc) return rc;
    uint32_t counts[10];
    if ((rc = ctx->read_back(bp->dCounts, counts, sizeof(counts)))) return rc;
    const uint32_t n = counts[0], nq = counts[1];
    if (counts[9]) {
        const uint32_t zero = 0u;
        HIP_TRY(hipMemcpyAsync(bp->dCounts + 9, &zero, 4, hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        return fail(B3AMD_ERR_ARGUMENT, "move_proxies entry %u named a proxy that is not live", counts[9] - 1u);
    }
    if (n != bp->count || nq > n)
        return fail(B3AMD_ERR_DEVICE, "broadphase lists %u live proxies (%u moved), expected %u", n, nq, bp->count);
    if (nq == 0) return B3AMD_OK;
    Lbvh t;
    t.n = n;
    t.keys = bp->dKeys2;
    t.left = bp->dNodes;
    t.right = bp->dNodes + bp->leafCap;
    t.parent = bp->dNodes + 2ull * bp->leafCap;
    t.escape = bp->dNodes + 4ull * bp->leafCap;
    t.first = bp->dNodes + 6ull * bp->leafCap;
    t.last = bp->dNodes + 7ull * bp->leafCap;
    t.lo = bp->dBoxes;
    t.hi = bp->dBoxes + bp->leafCap;
    t.preLo = bp->dBoxes + 2ull * bp->leafCap;
    t.preHi = bp->dBoxes + 3ull * bp->leafCap;
    t.sufLo = bp->dBoxes + 4ull * bp->leafCap;
    t.sufHi = bp->dBoxes + 5ull * bp->leafCap;
    t.blkStride = refit_blocks(bp->leafCap);
    t.blkLo = bp->dBoxes + 6ull * bp->leafCap;
    t.blkHi = t.blkLo + 32ull * t.blkStride;
    bp_morton_kernel<<<blocks(n), 256, 0, s>>>(n, bp->dLeaves, bp->st, bp->dCounts + 3, bp->dKeys);
    HIP_TRY(hipGetLastError());
    if ((rc = bp_sort(bp, bp->dKeys, bp->dKeys2, n, 64u))) return rc;
    if (n > 1u) {
        bp_build_kernel<<<blocks(n - 1u), 256, 0, s>>>(t);
        const uint32_t nb = (n + kRefitBlock - 1u) / kRefitBlock;
        bp_refit_block_kernel<<<nb, kRefitBlock, 0, s>>>(t, bp->st);
        if (nb > 1u) {
            bp_refit_table_kernel<<<1, 1024, 0, s>>>(t, nb);
            bp_refit_cross_kernel<<<blocks(n - 1u), 256, 0, s>>>(t);
        }
    }
    bp_escape_kernel<<<blocks(2ull * n - 1u), 256, 0, s>>>(t);
    HIP_TRY(hipGetLastError());
    const uint32_t bits = id_bits(bp->idCap);
    uint32_t np = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        if ((rc = bp_grow_pairs(bp, std::max<uint32_t>(np, 4u * nq)))) return rc;
        HIP_TRY(launch_zero_words(bp->dCounts + 2, 1, s));
        rc = timed(ctx, s, "bp_query", [&] {
            const uint32_t qb = (nq + kQueryGroups - 1u) / kQueryGroups;
            bp_query_kernel<<<qb, 256, 0, s>>>(t, bp->st, nq, bp->dQueries, bits, bp->dPairKeys, bp->pairCap,
                                               bp->dCounts + 2, bp->dCounts + 10);
            return hipGetLastError();
        });
        if (rc) return rc;
        uint32_t qs[9];
        if ((rc = ctx->read_back(bp->dCounts + 2, qs, sizeof(qs)))) return rc;
        np = qs[0];
        if (qs[8]) return fail(B3AMD_ERR_DEVICE, "broadphase tree walk did not end (internal)");
        if (np <= bp->pairCap) break;
    }
    if (np > bp->pairCap) return fail(B3AMD_ERR_DEVICE, "pair buffer did not grow");
    if (np) {
        if ((rc = bp_sort(bp, bp->dPairKeys, bp->dPairKeys2, np, 2u * bits))) return rc;
        bp_pairs_kernel<<<blocks(np), 256, 0, s>>>(np, bp->dPairKeys2, bits, bp->st.user, bp->dPairs);
    }
    // The move buffer is reset for the next step (broad_phase.h:185-186).
    bp_clear_moved_kernel<<<blocks(nq), 256, 0, s>>>(nq, bp->dQueries, bp->st.flags);
    HIP_TRY(hipGetLastError());
    if (sync) HIP_TRY(hipStreamSynchronize(s));
    bp->lastPairs = np;
    *pair_count = np;
    if (d_pairs) *d_pairs = bp->dPairs;
    return B3AMD_OK;
}

int launch_shape_aabbs(b3amd_context* ctx, uint32_t n, const uint32_t* d_shapes, const uint32_t* d_bodies,
                       uint32_t body_count, const b3amd_transform* d_body_xfs, float* d_aabbs) {
    hipStream_t s = ctx->stream;
    int rc;
    if (!ctx->dStatus && (rc = ctx->alloc(64, (void**)&ctx->dStatus))) return rc;
    ShapeAabbArgs a;
    a.n = n;
    a.bodyCount = body_count;
    a.shapes = d_shapes;
    a.bodies = d_bodies;
    a.xfs = reinterpret_cast<const float*>(d_body_xfs);
    a.sweeps = nullptr;
    a.table = ctx->dShapes;
    a.shapeCount = ctx->shapeCount;
    a.pool = ctx->dPool;
    a.aabbs = d_aabbs;
    a.disp = nullptr;
    a.status = ctx->dStatus;
    return timed(ctx, s, "shape_aabb", [&] {
        shape_aabb_kernel<<<blocks(n), 256, 0, s>>>(a);
        return hipGetLastError();
    });
}

}  // namespace b3amd

extern "C" {

int b3amd_broadphase_create(b3amd_context* ctx, b3amd_broadphase** out) {
    if (!ctx || !out) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_broadphase* bp = new b3amd_broadphase();
    bp->ctx = ctx;
    int rc;
    if ((rc = ctx->alloc(16 * 4, (void**)&bp->dCounts)) || (rc = bp_grow_ids(bp, 1024u))) {
        b3amd_broadphase_destroy(bp);
        return rc;
    }
    bp_fill_kernel<<<1, 64, 0, ctx->stream>>>(16u, bp->dCounts, 0u);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        b3amd_broadphase_destroy(bp);
        return fail(B3AMD_ERR_DEVICE, "broadphase set-up failed");
    }
    *out = bp;
    return B3AMD_OK;
}

int b3amd_broadphase_destroy(b3amd_broadphase* bp) {
    if (!bp) return B3AMD_OK;
    b3amd_context* ctx = bp->ctx;
    hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {bp->st.lower, bp->st.upper, bp->st.user, bp->st.flags, bp->dLeaves, bp->dKeys, bp->dKeys2,
                    bp->dNodes, bp->dBoxes, bp->dCounts, bp->dPairKeys, bp->dPairKeys2, bp->dPairs,
                    bp->dSortTmp, bp->dEdits};
    for (void* p : ptrs) ctx->release(p);
    delete bp;
    return B3AMD_OK;
}

int b3amd_broadphase_create_proxies(b3amd_broadphase* bp, uint32_t n, const b3amd_aabb* aabbs,
                                    const uint32_t* user_data, uint32_t* proxy_ids) {
    if (!bp) return fail(B3AMD_ERR_ARGUMENT, "null broadphase");
    if (n == 0) return B3AMD_OK;
    if (!aabbs || !proxy_ids) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    std::vector<BpEdit> edits(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t id;
        int rc = bp_take_id(bp, &id);
        if (rc) return rc;
        proxy_ids[i] = id;
        BpEdit& e = edits[i];
        e.id = id;
        e.mode = kEditCreate;
        e.user = user_data ? user_data[i] : i;
        e.pad = 0u;
        // fatAABB = aabb; fatAABB.Extend(b3_aabbExtension) (broad_phase.cpp:80-81, aabb3.h:85-90)
        const b3amd_aabb& a = aabbs[i];
        e.lower[0] = a.lower.x - kAabbExtension, e.lower[1] = a.lower.y - kAabbExtension,
        e.lower[2] = a.lower.z - kAabbExtension;
        e.upper[0] = a.upper.x + kAabbExtension, e.upper[1] = a.upper.y + kAabbExtension,
        e.upper[2] = a.upper.z + kAabbExtension;
    }
    return bp_apply_edits(bp, edits);
}

int b3amd_broadphase_destroy_proxies(b3amd_broadphase* bp, uint32_t n, const uint32_t* proxy_ids) {
    if (!bp) return fail(B3AMD_ERR_ARGUMENT, "null broadphase");
    if (n == 0) return B3AMD_OK;
    if (!proxy_ids) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    std::vector<BpEdit> edits(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t id = proxy_ids[i];
        if (id >= bp->idCap || !bp->live[id]) return fail(B3AMD_ERR_ARGUMENT, "proxy %u is not live", id);
        // RemoveLeaf frees the parent when the leaf is not the root, then FreeNode the leaf
        // (dynamic_tree.cpp:72-79, 251-292).
        bp->freeGroups.emplace_back(id, bp->count >= 2u);
        bp->count--;
        bp->live[id] = 0u;
        edits[i] = BpEdit{id, kEditDestroy, 0u, 0u, {0, 0, 0}, {0, 0, 0}};
    }
    return bp_apply_edits(bp, edits);
}

int b3amd_broadphase_touch_proxies(b3amd_broadphase* bp, uint32_t n, const uint32_t* proxy_ids) {
    if (!bp) return fail(B3AMD_ERR_ARGUMENT, "null broadphase");
    if (n == 0) return B3AMD_OK;
    if (!proxy_ids) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    std::vector<BpEdit> edits(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t id = proxy_ids[i];
        if (id >= bp->idCap || !bp->live[id]) return fail(B3AMD_ERR_ARGUMENT, "proxy %u is not live", id);
        edits[i] = BpEdit{id, kEditTouch, 0u, 0u, {0, 0, 0}, {0, 0, 0}};
    }
    return bp_apply_edits(bp, edits);
}

int b3amd_broadphase_move_proxies(b3amd_broadphase* bp, uint32_t n, const uint32_t* d_proxy_ids,
                                  const b3amd_aabb* d_aabbs, const b3amd_vec3* d_displacements, uint32_t* d_moved) {
    if (!bp) return fail(B3AMD_ERR_ARGUMENT, "null broadphase");
    if (n == 0) return B3AMD_OK;
    if (!d_proxy_ids || !d_aabbs || !d_displacements) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    hipStream_t s = bp->ctx->stream;
    // A bad id is counted in the status word and reported by the next find_pairs.
    return timed(bp->ctx, s, "bp_move", [&] {
        bp_move_kernel<<<blocks(n), 256, 0, s>>>(n, d_proxy_ids, reinterpret_cast<const float*>(d_aabbs),
                                                 reinterpret_cast<const float*>(d_displacements), bp->st, bp->idCap,
                                                 d_moved, bp->dCounts + 9);
        return hipGetLastError();
    });
}

int b3amd_broadphase_find_pairs(b3amd_broadphase* bp, uint32_t* pair_count, const uint32_t** d_pairs) {
    return bp_find_pairs(bp, pair_count, d_pairs, true);
}

int b3amd_broadphase_test_overlap(b3amd_broadphase* bp, uint32_t n, const uint32_t* d_proxy_pairs,
                                  uint32_t* d_overlap) {
    if (!bp) return fail(B3AMD_ERR_ARGUMENT, "null broadphase");
    if (n == 0) return B3AMD_OK;
    if (!d_proxy_pairs || !d_overlap) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    hipStream_t s = bp->ctx->stream;
    bp_test_overlap_kernel<<<blocks(n), 256, 0, s>>>(n, d_proxy_pairs, bp->st, bp->idCap, d_overlap);
    HIP_TRY(hipGetLastError());
    return B3AMD_OK;
}

int b3amd_broadphase_info(const b3amd_broadphase* bp, uint32_t* proxy_count, uint32_t* id_capacity) {
    if (!bp) return fail(B3AMD_ERR_ARGUMENT, "null broadphase");
    if (proxy_count) *proxy_count = bp->count;
    if (id_capacity) *id_capacity = bp->idCap;
    return B3AMD_OK;
}

int b3amd_broadphase_read(b3amd_broadphase* bp, uint32_t first, uint32_t count, b3amd_aabb* aabbs,
                          uint32_t* user_data, uint32_t* flags) {
    if (!bp) return fail(B3AMD_ERR_ARGUMENT, "null broadphase");
    if ((uint64_t)first + count > bp->idCap) return fail(B3AMD_ERR_ARGUMENT, "range beyond the id capacity");
    if (count == 0) return B3AMD_OK;
    hipStream_t s = bp->ctx->stream;
    std::vector<float4> lo(count), hi(count);
    HIP_TRY(hipMemcpyAsync(lo.data(), bp->st.lower + first, 16ull * count, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hi.data(), bp->st.upper + first, 16ull * count, hipMemcpyDeviceToHost, s));
    if (user_data) HIP_TRY(hipMemcpyAsync(user_data, bp->st.user + first, 4ull * count, hipMemcpyDeviceToHost, s));
    if (flags) HIP_TRY(hipMemcpyAsync(flags, bp->st.flags + first, 4ull * count, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (aabbs)
        for (uint32_t i = 0; i < count; ++i)
            aabbs[i] = b3amd_aabb{{lo[i].x, lo[i].y, lo[i].z}, {hi[i].x, hi[i].y, hi[i].z}};
    return B3AMD_OK;
}

int b3amd_shape_aabbs(b3amd_context* ctx, uint32_t n, const uint32_t* d_shapes, const uint32_t* d_bodies,
                      uint32_t body_count, const b3amd_transform* d_body_xfs, const b3amd_sweep* d_body_sweeps,
                      b3amd_aabb* d_aabbs, b3amd_vec3* d_displacements) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    if (n == 0) return B3AMD_OK;
    if (!d_shapes || !d_bodies || !d_body_xfs || !d_aabbs) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (!ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no shape table uploaded");
    hipStream_t s = ctx->stream;
    ShapeAabbArgs a;
    a.n = n;
    a.bodyCount = body_count;
    a.shapes = d_shapes;
    a.bodies = d_bodies;
    a.xfs = reinterpret_cast<const float*>(d_body_xfs);
    a.sweeps = reinterpret_cast<const float*>(d_body_sweeps);
    a.table = ctx->dShapes;
    a.shapeCount = ctx->shapeCount;
    a.pool = ctx->dPool;
    a.aabbs = reinterpret_cast<float*>(d_aabbs);
    a.disp = reinterpret_cast<float*>(d_displacements);
    int rc;
    if (!ctx->dStatus && (rc = ctx->alloc(64, (void**)&ctx->dStatus))) return rc;
    a.status = ctx->dStatus;
    HIP_TRY(launch_zero_words(a.status, 1, s));
    rc = timed(ctx, s, "shape_aabb", [&] {
        shape_aabb_kernel<<<blocks(n), 256, 0, s>>>(a);
        return hipGetLastError();
    });
    if (rc) return rc;
    uint32_t status = 0;
    if ((rc = ctx->read_back(a.status, &status, 4))) return rc;
    if (status)
        return fail(B3AMD_ERR_ARGUMENT, "shape entry %u: not a convex shape of the table or body beyond %u",
                    status - 1u, body_count);
    return B3AMD_OK;
}

}  // extern "C"
