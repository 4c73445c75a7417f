// kscr_lbvh.h -- linear BVH over axis-aligned boxes on gfx950.
//
// Same tree as the reference's Kitten GPU LBVH
// (KE/KittenGpuLBVH/lbvh.cu:53-426): 30-bit Morton codes of the box centres
// in the root box, a stable sort of the codes, Karras's internal-node build
// with keys (code << 32 | index), a bottom-up refit, and a self query that
// reports each overlapping pair once.  The sort is this library's own stable
// LSD radix sort.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace kscr {

// 64-byte node: parent/left/right carry the child side / leaf bit in the MSB,
// fence is the other end of the node's leaf range; two child boxes.
struct alignas(64) LbvhNode {
    uint32_t parent, left, right, fence;
    float b[2][6];  // (min xyz, max xyz) of the left and right child
};
// the query reads a node as four float4 words
static_assert(sizeof(LbvhNode) == 64 && offsetof(LbvhNode, fence) == 12 && offsetof(LbvhNode, b) == 16,
              "LbvhNode layout");

// Karras trees over 62-bit keys (30-bit Morton code, 32-bit index) are at
// most 63 levels tall.
constexpr int kMaxLevels = 64;

// Trees over at most 131072 objects are fit by rounds (no height pass); the
// nodes that cross a 512-leaf tile boundary, about 2k of them at 100k objects,
// are fit from LDS up to this many and from global memory beyond.
constexpr int kLbvhFitSlots = 4096;

struct LbvhDev {
    int n = 0;
    int cap = 0;
    int tiles = 0;
    float* aabb = nullptr;  // 6 floats per object
    // boxes the build, refit and query read in place of aabb when set (device
    // memory of the caller's, kept as the reference's LBVH keeps the pointer
    // compute was given)
    const float* src = nullptr;
    const float* boxes() const { return src ? src : aabb; }
    uint32_t* codes = nullptr;
    uint32_t* ids = nullptr;
    uint32_t* codes_tmp = nullptr;
    uint32_t* ids_tmp = nullptr;
    uint32_t* hist = nullptr;      // 256 * tiles
    uint32_t* hist_sums = nullptr;  // scan partials
    float* partial = nullptr;       // root reduction partials (6 per block)
    float* root = nullptr;          // 6 floats
    LbvhNode* nodes = nullptr;
    uint32_t* leaf_parent = nullptr;
    int* flags = nullptr;
    int* max_depth = nullptr;
    // query counters, one block: [0] pairs found, [1] queries to redo,
    // [2] the error word of the library's own LBVH handle
    int* count = nullptr;
    int* redo_count = nullptr;  // count + 1
    int2* redo = nullptr;       // queries deeper than the first pass's stack: (sorted index, pairs reported)
    // internal nodes grouped by height (leaves 0): refit runs one level per
    // launch, from (offset, count) per height kept on the device
    uint32_t* node_height = nullptr;
    uint32_t* level_nodes = nullptr;
    uint32_t* level_count = nullptr;  // kMaxLevels counts, then kMaxLevels scatter cursors
    int2* level_range = nullptr;      // kMaxLevels (offset, count)
    int max_depth_host = -1;          // DFS stack bound of the built tree once read back (lbvh_read_depth)
    // Trees over at most 131072 objects ("small"): the sort runs one launch per
    // digit pass and the fit runs by rounds; then node_height holds each
    // node's DFS stack depth, level_nodes the nodes that cross a tile
    // boundary, redo their child words, flags each such node's index in that
    // list, level_count[0] their number, and codes_tmp the done words of the
    // global-memory fit.
    bool small = false;
    int fit_slots = kLbvhFitSlots;  // crossing nodes fit from LDS (0 .. kLbvhFitSlots)
    bool built = false;
};

hipError_t lbvh_reserve(LbvhDev& d, int n);
// Lays out the buffers of a tree over at most n objects in [base, base +
// bytes) (256-byte aligned pieces) instead of allocating them; returns the
// bytes the layout needs, and sets d's pointers only when they fit (base may
// be null to ask for the size).  Such a view is never passed to lbvh_release.
size_t lbvh_carve(LbvhDev& d, int n, char* base, size_t bytes);
void lbvh_release(LbvhDev& d);
// Builds the tree over d.aabb[0 .. n) (all device work on `stream`).
hipError_t lbvh_build(LbvhDev& d, int n, hipStream_t stream);
// Recomputes the node boxes from d.aabb for the existing structure.
hipError_t lbvh_refit(LbvhDev& d, hipStream_t stream);
// Copies the built tree's DFS stack bound to d.max_depth_host (synchronises
// the stream).
hipError_t lbvh_read_depth(LbvhDev& d, hipStream_t stream);
// Self query: writes pairs (found object, query object), the found leaf lying
// at a larger sorted position than the query, into pairs[0 .. cap) and the
// number found (which may exceed cap) into *d.count.  Pairs whose object ids
// differ by at most min_gap are skipped (min_gap < 0 keeps all).  A traversal
// deeper than the stack limit ends early and ORs 4 into *error (if given).
hipError_t lbvh_query(LbvhDev& d, int2* pairs, int cap, int min_gap, int* error, hipStream_t stream);

// Fills n ints with a value (a kernel, so it is safe under graph capture).
hipError_t fill_i32(int* p, int value, size_t n, hipStream_t stream);

}  // namespace kscr
