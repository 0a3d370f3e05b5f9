// vx_edits.hip -- edit lists to Batch buffers on the device.
//
// Input: per tree a contiguous range of edits (x, y, z, value) in submission
// order.  Edits are radix-sorted by (tree, Morton path) with the submission index
// as payload (the project's stable LSD sort below, so the last write to a voxel ends
// each run), and the
// run ends are scattered into zeroed dense Batch buffers exactly as Batch::just_set
// records them (core/batch.rs:145-175): a non-default value sets the slot's set
// bit, the default value sets its clear bit, the value is stored either way.
// After the apply, the touched bricks are zeroed again so the buffers stay clear.
#include <hip/hip_runtime.h>

#include <cstring>

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

// ---- stable LSD radix sort of (u64 key, u32 value) pairs, 8 key bits per pass
//
// A pass over G blocks, each owning one contiguous range of the input: (1) a digit
// histogram per block; (2) one exclusive scan of the histograms in digit-major order,
// so that every block's keys of a digit land after those of the blocks before it;
// (3) a scatter in which each block walks its range in order, 256 keys per round:
// equal digits are ranked within a wavefront by ballots over the digit's bits, across
// the block's four wavefronts and across rounds through LDS counters.  Order among equal
// digits is the input order, so the sort is stable (core/batch.rs:145-175: the last
// write to a voxel wins).
constexpr unsigned kSortBlock = 256;
constexpr unsigned kSortScanBlock = 1024;

__global__ __launch_bounds__(kSortBlock) void k_sort_hist(const uint64_t* keys, uint64_t n, uint64_t per,
                                                          unsigned shift, uint32_t* hist) {
    __shared__ uint32_t c[256];
    c[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t b0 = (uint64_t)blockIdx.x * per;
    const uint64_t b1 = min(n, b0 + per);
    for (uint64_t i = b0 + threadIdx.x; i < b1; i += kSortBlock) atomicAdd(&c[(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    hist[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = c[threadIdx.x];
}

// Exclusive scan of the digit counts in tiles of kScanTileN: each tile in place (its own
// prefix only), the tile's total in tot[tile].  The tile totals are scanned next by
// k_sort_scan (one block: at most 128 tiles for 2048 sort blocks), and the scatter adds a
// count's tile offset when it reads it.
constexpr unsigned kSortScanPer = 4;
constexpr unsigned kScanTileN = kSortScanBlock * kSortScanPer;
__global__ __launch_bounds__(kSortScanBlock) void k_sort_scan_tiles(uint32_t* hist, uint32_t m, uint32_t* tot) {
    __shared__ uint32_t s_w[kSortScanBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t base = blockIdx.x * kScanTileN + threadIdx.x * kSortScanPer;
    uint32_t v[kSortScanPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kSortScanPer; i++) {
        v[i] = base + i < m ? hist[base + i] : 0u;
        sum += v[i];
    }
    uint32_t x = sum;  // inclusive scan of the lanes' sums
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t before = 0;  // the earlier wavefronts of the block
    for (uint32_t w = 0; w < wave; w++) before += s_w[w];
    uint32_t run = before + x - sum;
#pragma unroll
    for (uint32_t i = 0; i < kSortScanPer; i++) {
        if (base + i < m) hist[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == kSortScanBlock - 1) tot[blockIdx.x] = run;
}

// Exclusive scan of m counts in place, one block: each thread scans a contiguous segment.
__global__ __launch_bounds__(kSortScanBlock) void k_sort_scan(uint32_t* hist, uint32_t m) {
    __shared__ uint32_t part[kSortScanBlock];
    const uint32_t seg = (m + kSortScanBlock - 1) / kSortScanBlock;
    const uint32_t a = min(m, threadIdx.x * seg), b = min(m, a + seg);
    uint32_t sum = 0;
    for (uint32_t i = a; i < b; i++) sum += hist[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kSortScanBlock; off <<= 1) {
        const uint32_t t = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t i = a; i < b; i++) {
        const uint32_t v = hist[i];
        hist[i] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kSortBlock) void k_sort_scatter(const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
                                                             uint32_t* vout, uint64_t n, uint64_t per, unsigned shift,
                                                             const uint32_t* offs, const uint32_t* tile_off) {
    __shared__ uint32_t base[256];             // next output position of each digit (this block)
    __shared__ uint32_t wcnt[kSortBlock / 64][256];  // per wavefront: its digit counts, then offsets
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    const uint64_t at = (uint64_t)tid * gridDim.x + blockIdx.x;  // digit-major count of this block
    base[tid] = offs[at] + tile_off[at / kScanTileN];
    const uint64_t b0 = (uint64_t)blockIdx.x * per;
    const uint64_t b1 = min(n, b0 + per);
    const uint64_t lt = (1ull << lane) - 1ull;
    for (uint64_t r0 = b0; r0 < b1; r0 += kSortBlock) {  // uniform over the block
        const uint64_t i = r0 + tid;
        const bool valid = i < b1;
        const uint64_t key = valid ? kin[i] : 0ull;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
#pragma unroll
        for (uint32_t w = 0; w < kSortBlock / 64; w++) wcnt[w][tid] = 0;
        // lanes of this wavefront holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const bool on = (d >> bit) & 1u;
            const uint64_t bal = __ballot(valid && on);
            peers &= on ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        __syncthreads();  // counters cleared (and the previous round's reads done)
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {  // thread = digit: offsets of the wavefronts in order, then the block's next base
            uint32_t run = base[tid];
#pragma unroll
            for (uint32_t w = 0; w < kSortBlock / 64; w++) {
                const uint32_t t = wcnt[w][tid];
                wcnt[w][tid] = run;
                run += t;
            }
            base[tid] = run;
        }
        __syncthreads();
        if (valid) {
            const uint32_t dst = wcnt[wave][d] + rank;
            kout[dst] = key;
            vout[dst] = val;
        }
        __syncthreads();
    }
}

__global__ void k_edit_status(uint32_t ntrees, const EditTree* trees, int32_t* pre) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntrees; t += gridDim.x * blockDim.x)
        pre[t] = trees[t].status;
}

}  // namespace vx
