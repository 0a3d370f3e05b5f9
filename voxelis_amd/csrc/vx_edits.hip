// vx_edits.hip -- edit lists to Batch buffers on the device.
//
// Input: per tree a contiguous range of edits (x, y, z, value) in submission
// order.  Edits are radix-sorted by (tree, Morton path) with the submission index
// as payload (a stable sort, so the last write to a voxel ends each run), and the
// run ends are scattered into zeroed dense Batch buffers exactly as Batch::just_set
// records them (core/batch.rs:145-175): a non-default value sets the slot's set
// bit, the default value sets its clear bit, the value is stored either way.
// After the apply, the touched bricks are zeroed again so the buffers stay clear.
#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "vx_device.h"

namespace vx {

constexpr int kPathBits = 3 * kMaxDepth;  // 21: Morton path of a depth-7 tree

__device__ inline uint32_t spread3(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x30000FFu;
    v = (v | (v << 8)) & 0x300F00Fu;
    v = (v | (v << 4)) & 0x30C30C3u;
    v = (v | (v << 2)) & 0x9249249u;
    return v;
}

struct EditTree {
    uint64_t first;   // first edit of the tree
    uint64_t mbase;   // byte offset of its masks in the workspace
    uint64_t vbase;   // element offset of its values
    uint32_t depth;
    int32_t status;
};

__device__ inline uint32_t tree_of_edit(const uint64_t* first, uint32_t ntrees, uint64_t e) {
    uint32_t lo = 0, hi = ntrees;  // last tree with first <= e
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Sort keys: tree << 21 | path; out-of-range coordinates (a panic in the
// reference's Vec indexing) sort after every tree and fail their tree.
__global__ void k_edit_keys(uint64_t n, const uint64_t* first, uint32_t ntrees, EditTree* trees,
                            const int32_t* xyz, uint64_t* keys, uint32_t* order) {
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t t = tree_of_edit(first, ntrees, e);
        int32_t x = xyz[3 * e], y = xyz[3 * e + 1], z = xyz[3 * e + 2];
        uint64_t key;
        if (trees[t].status < 0) {
            key = (uint64_t)ntrees << kPathBits;  // tree rejected on the host
        } else if (x < 0 || y < 0 || z < 0 || x >= (1 << trees[t].depth) || y >= (1 << trees[t].depth) ||
                   z >= (1 << trees[t].depth)) {
            trees[t].status = -4;
            key = (uint64_t)ntrees << kPathBits;
        } else {
            uint32_t path = spread3((uint32_t)x) | (spread3((uint32_t)y) << 1) | (spread3((uint32_t)z) << 2);
            key = ((uint64_t)t << kPathBits) | path;
        }
        keys[e] = key;
        order[e] = (uint32_t)e;
    }
}

// Run ends (the last submission of each voxel) record the voxel like just_set.
template <typename T>
__global__ void k_edit_scatter(uint64_t n, uint32_t ntrees, const EditTree* trees, const uint64_t* keys,
                               const uint32_t* order, const T* values, uint8_t* masks_ws, T* values_ws,
                               bool zero) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t key = keys[i];
        uint32_t t = (uint32_t)(key >> kPathBits);
        if (t >= ntrees) continue;
        if (i + 1 < n && keys[i + 1] == key) continue;  // a later write to the same voxel wins
        const EditTree& tr = trees[t];
        uint32_t path = (uint32_t)key & ((1u << kPathBits) - 1);
        uint64_t p = path >> 3;
        uint32_t slot = path & 7u;
        if (zero) {
            masks_ws[tr.mbase + 2 * p] = 0;
            masks_ws[tr.mbase + 2 * p + 1] = 0;
            values_ws[tr.vbase + 8 * p + slot] = 0;
            continue;
        }
        T v = values[order[i]];
        values_ws[tr.vbase + 8 * p + slot] = v;
        uint64_t byte = tr.mbase + 2 * p + (v != 0 ? 0 : 1);
        uint32_t* word = (uint32_t*)(masks_ws + (byte & ~3ull));
        atomicOr(word, (1u << slot) << (8 * (byte & 3)));
    }
}

__global__ void k_edit_status(uint32_t ntrees, const EditTree* trees, int32_t* pre) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntrees; t += gridDim.x * blockDim.x)
        pre[t] = trees[t].status;
}

}  // namespace vx
