// vx_kernels.hip -- gfx950 kernels of the batched SVO-DAG edit path.
//
// apply_batch (spatial/voxtree.rs:724-1118) is rebuilt as a height-synchronous
// sweep over dense per-level slot arrays: every tree contributes 8^(D-h) slots at
// height h, a slot holds the node that replaces the old one at that position
// (INVALID = untouched), and each height is built from the 8 contiguous slots of
// the height below.  Interning (interner/mod.rs:627-829) is a two-kernel step per
// height: build writes candidate keys, probe inserts/looks them up in an HBM
// open-addressing table with atomicCAS; keys of the same launch resolve to one
// owner without any waiting, and commit writes the owners' nodes.
//
// Reference counting follows the reference's net effect: a new node adds one
// reference to each non-empty child slot, a tree owns one reference on its root,
// and the reference's counting quirks are applied as explicit deltas (see
// apply_drift_* below).  Frees run as a level-parallel worklist after the sweep
// (dec_ref_recursive, interner/mod.rs:419-534).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "vx_device.h"

namespace vx {

#define VX_LOOP(i, n)                                                                        \
    for (uint64_t i##_b = (uint64_t)blockIdx.x * blockDim.x; i##_b < (uint64_t)(n);          \
         i##_b += (uint64_t)gridDim.x * blockDim.x)                                          \
        for (uint64_t i = i##_b + threadIdx.x, i##_k = 0; i##_k == 0; i##_k = 1)

enum Stat {
    ST_HIT = 0, ST_MISS, ST_DUP, ST_LEAF_MISS, ST_COLLAPSE, ST_RECYCLED, ST_H2, ST_H3, ST_H4, ST_H5,
    ST_CAND, ST_PROBES, ST_CLAIMS, ST_CLAIMS_LOST, ST_MERGED,
    ST_TOMB,  // tombstones in the pattern table (signed adds; a counter, not a statistic)
    // the brick level's share of ST_CAND, ST_MERGED, ST_HIT, ST_MISS, ST_DUP
    ST_CAND1, ST_MERGED1, ST_HIT1, ST_MISS1, ST_DUP1,
    ST_LTOMB  // tombstones in the hashed leaf table of four-byte voxels
};

__device__ inline uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t b = __ballot(pred);
    if (b == 0) return 0;
    uint32_t lane = __lane_id();
    int leader = __ffsll((unsigned long long)b) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(b));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
}

__device__ inline void wave_count(uint64_t* st, bool pred) {
    uint64_t b = __ballot(pred);
    if (b == 0) return;
    int leader = __ffsll((unsigned long long)b) - 1;
    if ((int)__lane_id() == leader) atomicAdd((unsigned long long*)st, (unsigned long long)__popcll(b));
}

// this block's stripe of a statistics counter
__device__ inline uint64_t* stat(Ctl* ctl, int id) { return &ctl->st[blockIdx.x & (kStatStripes - 1)][id]; }

__device__ inline uint64_t shfl64(uint64_t v, int src) {
    uint32_t lo = __shfl((uint32_t)v, src);
    uint32_t hi = __shfl((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint64_t readlane64(uint64_t v, int lane) {
    uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
    uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline void wave_add64(uint64_t* st, uint32_t v) {
    uint32_t t = wave_sum(v);
    if (__lane_id() == 0 && t) atomicAdd((unsigned long long*)st, (unsigned long long)t);
}

constexpr uint32_t kNoIdx = 0xFFFFFFFFu;
constexpr int kMatchRounds = 4;

// refs[idx] += 1 for every lane with idx != kNoIdx; lanes with equal idx share one
// atomic (hot leaves are referenced from most new nodes).
__device__ inline void wave_ref_inc(uint32_t* refs, uint32_t idx) {
    bool pending = idx != kNoIdx;
    for (int r = 0; r < kMatchRounds; r++) {
        uint64_t act = __ballot(pending);
        if (!act) return;
        int leader = __ffsll((unsigned long long)act) - 1;
        uint32_t li = __shfl(idx, leader);
        bool same = pending && idx == li;
        uint64_t sm = __ballot(same);
        if ((int)__lane_id() == leader) atomicAdd(&refs[li], (uint32_t)__popcll(sm));
        if (same) pending = false;
    }
    if (pending) atomicAdd(&refs[idx], 1u);
}

// refs[idx] -= 1 for every lane with idx != kNoIdx (aggregated like wave_ref_inc).
// Returns true on exactly one lane per index whose count reached zero.  Every round's
// atomic is issued before any result is read, so they are in flight together.
__device__ inline bool wave_ref_dec_zero(uint32_t* refs, uint32_t idx) {
    bool pending = idx != kNoIdx;
    uint32_t old[kMatchRounds], want[kMatchRounds];
#pragma unroll
    for (int r = 0; r < kMatchRounds; r++) {
        want[r] = 0;  // 0: this lane did not lead round r
        old[r] = 0;
        const uint64_t act = __ballot(pending);
        if (act) {  // uniform over the wavefront
            const int leader = __ffsll((unsigned long long)act) - 1;
            const uint32_t li = __shfl(idx, leader);
            const bool same = pending && idx == li;
            const uint64_t sm = __ballot(same);
            if ((int)__lane_id() == leader) {
                want[r] = (uint32_t)__popcll(sm);
                old[r] = atomicSub(&refs[li], want[r]);
            }
            if (same) pending = false;
        }
    }
    const uint32_t rest = pending ? atomicSub(&refs[idx], 1u) : 0u;
    bool owner = pending && rest == 1u;
#pragma unroll
    for (int r = 0; r < kMatchRounds; r++) owner = owner || (want[r] != 0 && old[r] == want[r]);
    return owner;
}

// Block-local combining of reference-count updates: hot nodes (the few leaves
// most new nodes point at) get one global atomic per block instead of one per
// reference.  agg_add returns false when the table has no room.
constexpr int kAggSlots = 512;

__device__ inline void agg_init(uint32_t* akey, uint32_t* acnt) {
    for (int i = threadIdx.x; i < kAggSlots; i += blockDim.x) {
        akey[i] = kNoIdx;
        acnt[i] = 0;
    }
}

__device__ inline bool agg_add(uint32_t* akey, uint32_t* acnt, uint32_t idx, uint32_t n) {
    uint32_t h = (idx * 2654435761u) >> 23;
    // two slots only: a full table (many distinct children) falls back quickly
#pragma unroll 1
    for (int p = 0; p < 2; p++) {
        uint32_t s = (h + p) & (kAggSlots - 1);
        uint32_t k = atomicCAS(&akey[s], kNoIdx, idx);
        if (k == kNoIdx || k == idx) {
            atomicAdd(&acnt[s], n);
            return true;
        }
    }
    return false;
}

struct Vox8x4 {  // 8 four-byte voxels
    uint4 lo, hi;
};

template <typename T>
__device__ inline void load8(const T* v, uint32_t out[8]) {
    if constexpr (sizeof(T) == 1) {
        uint64_t w = *(const uint64_t*)v;
#pragma unroll
        for (int i = 0; i < 8; i++) out[i] = (uint32_t)(w >> (8 * i)) & 0xFFu;
    } else if constexpr (sizeof(T) == 2) {
        uint4 w = *(const uint4*)v;
        uint32_t a[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            out[2 * i] = a[i] & 0xFFFFu;
            out[2 * i + 1] = a[i] >> 16;
        }
    } else {
        const uint4 a = *(const uint4*)v, b = *(const uint4*)(v + 4);
        out[0] = a.x; out[1] = a.y; out[2] = a.z; out[3] = a.w;
        out[4] = b.x; out[5] = b.y; out[6] = b.z; out[7] = b.w;
    }
}

// core/voxel.rs:96-141 calc_average: mode; ties keep the first seen unless the
// leader is the default value and the challenger is not.  Distinct values are taken
// in order of first occurrence, each with its count; every index is static, so the
// routine stays in registers.
__device__ inline uint32_t calc_average(const uint32_t v[8]) {
    uint32_t best = v[0], bc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        bool first = true;
#pragma unroll
        for (int j = 0; j < k; j++) first = first && v[j] != v[k];
        uint32_t c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) c += v[j] == v[k] ? 1u : 0u;
        if (k == 0) {
            bc = c;
        } else if (first && (c > bc || (c == bc && best == 0 && v[k] != 0))) {
            best = v[k];
            bc = c;
        }
    }
    return best;
}

// ------------------------------------------------------------------ leaf table

// value -> leaf node index (get_or_create_leaf's map, interner/mod.rs:627-710).  u8/u16:
// direct-mapped arrays (leaf_map; leaf_need marks values to create).  i32: 2^32 values
// cannot be mapped directly, so an open-addressing table (ltab) of 8-byte slots
// value << 32 | index: 0 empty, kLeafTomb a removed entry (value 0 never has a leaf), and
// index kLeafNeed a value requested by this apply and not created yet; the claim that
// inserts a request lists the value once (need_list) for k_leaf_alloc.  A value is the
// whole key, so a slot compare is the full-key compare.
constexpr uint64_t kLeafTomb = 1ull;
constexpr uint32_t kLeafNeed = 0xFFFFFFFFu;

__device__ inline uint64_t leaf_hash(uint32_t v) { return mix64((uint64_t)v * 0x9E3779B97F4A7C15ull + 0x4C454146ull); }

// i32: the slot holding value v, or ~0 when absent.
template <typename T>
__device__ inline uint64_t ltab_find(const Pool<T>& P, uint32_t v) {
    uint64_t s = leaf_hash(v) & P.lmask;
    for (uint64_t g = 0; g <= P.lmask; g++) {
        const uint64_t e = P.ltab[s];
        if (e == 0) return ~0ull;
        if ((uint32_t)(e >> 32) == v && e != kLeafTomb) return s;
        s = (s + 1) & P.lmask;
    }
    return ~0ull;
}

// The live leaf index of value v (0: none, or requested and not created yet).
template <typename T>
__device__ inline uint32_t leaf_index(const Pool<T>& P, uint32_t v) {
    if constexpr (sizeof(T) <= 2) {
        return P.leaf_map[v];
    } else {
        const uint64_t s = ltab_find(P, v);
        if (s == ~0ull) return 0;
        const uint32_t i = (uint32_t)P.ltab[s];
        return i == kLeafNeed ? 0u : i;
    }
}

// Ask k_leaf_alloc to create the leaf of value v; returns the key of its dry-run
// placeholder id (capacity + key): the value itself for u8/u16, the request's slot for
// i32 (one per value, the same for every requester).
template <typename T>
__device__ inline uint32_t leaf_request(const Pool<T>& P, uint32_t v) {
    if constexpr (sizeof(T) <= 2) {
        P.leaf_need[v] = 1;
        return v;
    } else {
        uint64_t s = leaf_hash(v) & P.lmask;
        uint64_t first_tomb = ~0ull;
        for (uint64_t g = 0; g <= 4 * (P.lmask + 1); g++) {
            const uint64_t e = P.ltab[s];
            if (e != 0 && e != kLeafTomb && (uint32_t)(e >> 32) == v) return (uint32_t)s;
            if (e == kLeafTomb && first_tomb == ~0ull) first_tomb = s;
            if (e == 0 || (first_tomb != ~0ull && g > P.lmask)) {
                const uint64_t target = first_tomb != ~0ull ? first_tomb : s;
                const uint64_t expected = first_tomb != ~0ull ? kLeafTomb : 0ull;
                const uint64_t want = ((uint64_t)v << 32) | kLeafNeed;
                const uint64_t seen = atomicCAS((unsigned long long*)&P.ltab[target], (unsigned long long)expected,
                                                (unsigned long long)want);
                if (seen == expected) {
                    const uint32_t pos = atomicAdd(&P.ctl->n_need, 1u);
                    if (pos < P.need_cap) P.need_list[pos] = v;
                    else atomicOr(&P.ctl->err, kErrTableFull | kErrPoison);
                    return (uint32_t)target;
                }
                if ((uint32_t)(seen >> 32) == v && seen != kLeafTomb) return (uint32_t)target;
                s = (target + 1) & P.lmask;  // another value took the slot: go on behind it
                first_tomb = ~0ull;
                continue;
            }
            s = (s + 1) & P.lmask;
        }
        atomicOr(&P.ctl->err, kErrTableFull | kErrPoison);
        return 0;
    }
}

// Publish the leaf index of value v (k_leaf_alloc, snapshot import).
template <typename T>
__device__ inline void leaf_publish(const Pool<T>& P, uint32_t v, uint32_t idx) {
    if constexpr (sizeof(T) <= 2) {
        P.leaf_map[v] = idx;
    } else {
        const uint64_t want = ((uint64_t)v << 32) | idx;
        uint64_t s = leaf_hash(v) & P.lmask;
        for (uint64_t g = 0; g <= P.lmask; g++) {
            const uint64_t e = P.ltab[s];
            if (e != 0 && e != kLeafTomb && (uint32_t)(e >> 32) == v) {  // its request
                P.ltab[s] = want;
                return;
            }
            if (e == 0 && atomicCAS((unsigned long long*)&P.ltab[s], 0ull, (unsigned long long)want) == 0ull) return;
            s = (s + 1) & P.lmask;
        }
        atomicOr(&P.ctl->err, kErrTableFull | kErrPoison);
    }
}

// Remove value v's leaf (freed, or its creation rolled back).
template <typename T>
__device__ inline void leaf_remove(const Pool<T>& P, uint32_t v) {
    if constexpr (sizeof(T) <= 2) {
        P.leaf_map[v] = 0;
    } else {
        const uint64_t s = ltab_find(P, v);
        if (s != ~0ull) {
            P.ltab[s] = kLeafTomb;
            atomicAdd((unsigned long long*)stat(P.ctl, ST_LTOMB), 1ull);
        }
    }
}

template <typename T>
__device__ inline uint64_t leaf_id_of(const Pool<T>& P, uint32_t v) {
    uint32_t idx = leaf_index(P, v);
    return make_leaf(idx, tmg_gen(P.tmg[idx]));
}

template <typename T>
__device__ inline uint32_t alloc_at(const Pool<T>& P, uint32_t k, bool* ok, uint32_t ft0, uint32_t next0);

template <typename T>
__device__ inline uint32_t alloc_node(const Pool<T>& P, bool* ok) {
    return alloc_at(P, atomicAdd(&P.ctl->alloc_count, 1u), ok, P.ctl->free_top0, P.ctl->next0);
}

// The index of the launch sequence's k-th allocation: the free stack first, then
// fresh indices.  ft0 / next0: the allocation window's snapshot (free_top0, next0), which
// no kernel of the sequence changes before k_end_alloc.
template <typename T>
__device__ inline uint32_t alloc_at(const Pool<T>& P, uint32_t k, bool* ok, uint32_t ft0, uint32_t next0) {
    if (k >= P.alloc_limit) {
        atomicOr(&P.ctl->abort_, 2u);  // beyond the caller's reservation
        *ok = false;
        return 0;
    }
    if (k < ft0) {
        *ok = true;
        return P.free_stack[ft0 - 1 - k];
    }
    uint64_t b = (uint64_t)next0 + (uint64_t)(k - ft0);
    if (b >= P.capacity) {
        atomicOr(&P.ctl->abort_, 1u);  // interner capacity (interner/macros.rs:37-38)
        *ok = false;
        return 0;
    }
    *ok = true;
    return (uint32_t)b;
}

// A reference-count delta applied outside a commit, logged so an aborted apply
// can take it back.
template <typename T>
__device__ inline void ref_delta_logged(const Pool<T>& P, uint32_t chunk, uint32_t idx, int32_t delta,
                                        uint32_t st) {
    atomicAdd(&P.refs[idx], (uint32_t)delta);
    uint32_t pos = atomicAdd(&P.ctl->n_undo, 1u);
    if (pos < P.undo_cap) P.undo[pos] = Drift{chunk, idx, delta, st};
    else atomicOr(&P.ctl->err, kErrPoison);
}

// The group of a level slot (the last group whose first slot is <= slot): its
// first slot, tree depth and first chunk.  Every read of the kernel argument has a
// uniform index, so the selection costs scalar loads and per-lane selects only.
struct GroupPos {
    uint64_t base;
    uint64_t prev;  // the group's first slot one height below
    uint64_t next;  // ... one height above (~0 when the group's depth is this height)
    uint32_t depth;
    uint32_t chunk_base;
};

__device__ inline GroupPos group_pos(const LevelInfo& L, uint64_t slot) {
    GroupPos r{L.slot_base[0], L.prev_base[0], L.next_base[0], L.depth[0], L.chunk_base[0]};
    for (uint32_t k = 1; k < L.ngroups; k++) {
        bool in = slot >= L.slot_base[k];
        r.base = in ? L.slot_base[k] : r.base;
        r.prev = in ? L.prev_base[k] : r.prev;
        r.next = in ? L.next_base[k] : r.next;
        r.depth = in ? L.depth[k] : r.depth;
        r.chunk_base = in ? L.chunk_base[k] : r.chunk_base;
    }
    return r;
}

// group_pos for the slots of one block: when the block's first and last slots lie in
// one group (every block but those that straddle two), that group, found once per block
// from block-uniform values (scalar instructions); else each slot's own.  first and last
// must be the same for every lane.
struct BlockGroup {
    GroupPos G;
    bool one;
};
__device__ inline BlockGroup block_group(const LevelInfo& L, uint64_t first, uint64_t last) {
    const GroupPos a = group_pos(L, first), b = group_pos(L, last);
    return BlockGroup{a, a.base == b.base};
}
__device__ inline GroupPos group_of(const LevelInfo& L, const BlockGroup& B, uint64_t slot) {
    return B.one ? B.G : group_pos(L, slot);
}

// The fields of a chunk record that the build passes read first.  A wavefront whose
// lanes all lie in one chunk (every wavefront of a tree of depth >= 3 at the brick
// level) reads them with scalar loads: vector-memory address processing is what
// bounds these passes, and a wave-wide load of one address costs as much of it as
// any other.
struct ChunkHead {
    const uint8_t* masks;
    const void* values;
    uint64_t root_in;
    uint64_t init;
    int32_t status;
    bool has_fill;
    bool has_patches;
};

typedef const __attribute__((address_space(4))) ChunkDev ConstChunk;  // read-only view: scalar loads

// (A record's status may be set negative by another wavefront of the same pass; a
// reader that misses it is covered by the probe's own status check.)
__device__ inline ChunkHead chunk_head(const ChunkDev* chunks, uint32_t c) {
    const uint32_t c0 = __builtin_amdgcn_readfirstlane(c);
    if (__ballot(c != c0) == 0) {
        ConstChunk& r = *(ConstChunk*)(chunks + c0);
        // the four flag bytes (depth, has_fill, has_patches, skip) as one dword
        const uint32_t flags = *(const __attribute__((address_space(4))) uint32_t*)&r.depth;
        return ChunkHead{r.masks, r.values, r.root_in, r.init, r.status, ((flags >> 8) & 0xFFu) != 0,
                         ((flags >> 16) & 0xFFu) != 0};
    }
    const ChunkDev& r = chunks[c];
    return ChunkHead{r.masks, r.values, r.root_in, r.init, r.status, r.has_fill != 0, r.has_patches != 0};
}

// ------------------------------------------------------------------ launch setup

// Start of a launch sequence: allocator snapshot, per-sequence counters cleared.
__device__ inline void ctl_begin(Ctl* ctl) {
    {
        ctl->free_top0 = ctl->free_top;
        ctl->next0 = ctl->next_index;
        ctl->alloc_count = 0;
        ctl->n_new_leaves = 0;
        ctl->n_deferred = 0;
        ctl->n_drift = 0;
        ctl->dry_owners = 0;
        ctl->dry_new_leaves = 0;
        ctl->abort_ = 0;
        ctl->n_undo = 0;
        ctl->h1_fail = 0;
        ctl->n_owners = 0;
        for (int i = 0; i <= kLevels; i++) ctl->ncand[i] = 0;
        for (int i = 0; i < kLevels + 2; i++) ctl->gc_count[i] = 0;
    }
}

__global__ void k_begin(Ctl* ctl) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_begin(ctl);
}

// i32: requests still open at the end of an apply (a dry run's, whose placeholders were
// counted; a leaf whose creation failed) leave the table.  One block.
template <typename T>
__device__ inline void drop_requests(const Pool<T>& P) {
    if constexpr (sizeof(T) == 4) {
        const uint32_t nn = min(P.ctl->n_need, P.need_cap);
        for (uint32_t i = threadIdx.x; i < nn; i += blockDim.x) {
            const uint64_t sl = ltab_find(P, P.need_list[i]);
            if (sl != ~0ull && (uint32_t)P.ltab[sl] == kLeafNeed) P.ltab[sl] = kLeafTomb;
        }
    }
}

// Close the allocation window of a launch sequence.  An aborted apply
// (rollback = true) gives every allocation back instead: the allocator cursors
// stay, reference counts of the taken indices go back to zero and new leaves
// leave the value map.  counts (optional): new branches, new leaves.
template <typename T>
__global__ void k_end_alloc(Pool<T> P, bool rollback, uint32_t* counts) {
    Ctl* ctl = P.ctl;
    const bool abort = rollback && !P.dry && ctl->abort_ != 0;
    if (abort) {
        uint32_t n = ctl->alloc_count;
        uint32_t ft0 = ctl->free_top0;
        uint64_t room = (uint64_t)ft0 + (P.capacity > ctl->next0 ? P.capacity - ctl->next0 : 0u);
        uint64_t nok = n;
        if (nok > P.alloc_limit) nok = P.alloc_limit;
        if (nok > room) nok = room;
        for (uint64_t i = threadIdx.x; i < nok; i += blockDim.x) {
            uint32_t idx = i < ft0 ? P.free_stack[ft0 - 1 - i] : (uint32_t)(ctl->next0 + (i - ft0));
            P.refs[idx] = 0;
        }
        uint32_t nl = ctl->n_new_leaves;
        for (uint32_t i = threadIdx.x; i < nl; i += blockDim.x)
            leaf_remove(P, (uint32_t)P.values[P.new_leaves[i]]);
        drop_requests(P);
        __syncthreads();
        if (threadIdx.x == 0) {
            ctl->live_leaves -= nl;
            ctl->n_new_leaves = 0;
            ctl->alloc_count = 0;
            if (counts) counts[0] = counts[1] = 0xFFFFFFFFu;
        }
        return;
    }
    drop_requests(P);
    if (threadIdx.x == 0) {
        uint32_t n = ctl->alloc_count;
        ctl->n_owners = P.dry || n < ctl->own_base ? 0u : n - ctl->own_base;
        if (counts) {
            counts[0] = P.dry ? ctl->dry_owners : n - ctl->n_new_leaves;
            counts[1] = P.dry ? ctl->dry_new_leaves : ctl->n_new_leaves;
        }
        uint32_t used_free = n < ctl->free_top0 ? n : ctl->free_top0;
        ctl->free_top = ctl->free_top0 - used_free;
        uint64_t nx = (uint64_t)ctl->next0 + (n - used_free);
        ctl->next_index = (uint32_t)(nx > 0xFFFFFFFFull ? 0xFFFFFFFFull : nx);
        ctl->free_top0 = ctl->free_top;
        ctl->next0 = ctl->next_index;
        ctl->alloc_count = 0;
    }
}

// The previous tree's nodes on the path to one height-3 position (level-3 slot
// `slot`) and the 8 height-2 nodes below it, for the equivalent-node lookups of the
// brick and level passes (voxtree.rs:930-952): the node at height h >= 3 is
// recorded by the lane that owns the first height-3 position under it.  Below a
// leaf every height records that leaf (it covers the position); below an empty
// child, empty.  Trees with a fill start from the fill leaf instead and are
// skipped; a depth-2 tree's height-2 node is its root (recorded by k_prepare).
template <typename T>
__device__ inline void descend_body(const Pool<T>& P, const ChunkDev* chunks, const LevelInfo& L3,
                                    const uint64_t* doff, uint32_t n, const OldNodes& O,
                                    const uint64_t* roots, uint64_t slot) {
    const GroupPos G = group_pos(L3, slot);
    uint64_t local = slot - G.base;
    uint32_t D = G.depth;
    uint32_t sh = 3 * (D - 3);
    uint32_t c = G.chunk_base + (uint32_t)(local >> sh);
    uint32_t q = (uint32_t)(local & ((1ull << sh) - 1));
    const ChunkDev& C = chunks[c];
    if (C.has_fill) return;
    uint64_t cur = roots[C.orig];
    for (int hh = (int)D; hh >= 3; hh--) {
        if (hh < (int)D && cur != kEmpty && !bid_is_leaf(cur)) {
            if (bid_index(cur) >= P.capacity) {
                atomicOr(&P.ctl->err, kErrPoison);
                cur = kEmpty;
            } else {
                cur = P.children[8ull * bid_index(cur) + ((q >> (3 * (hh - 3))) & 7u)];
            }
        }
        uint32_t sub = 3 * (uint32_t)(hh - 3);
        if ((q & ((1u << sub) - 1u)) == 0) O.at[hh][doff[(size_t)hh * n + c] + (q >> sub)] = cur;
    }
    uint64_t below[8];
    const bool branch = cur != kEmpty && !bid_is_leaf(cur);
    if (branch && bid_index(cur) >= P.capacity) atomicOr(&P.ctl->err, kErrPoison);
    const bool ok = branch && bid_index(cur) < P.capacity;
    const uint64_t* src = P.children + 8ull * (ok ? bid_index(cur) : 0u);
#pragma unroll
    for (int i = 0; i < 8; i++) below[i] = ok ? src[i] : (branch ? kEmpty : cur);
    uint64_t* dst = O.at[2] + doff[(size_t)2 * n + c] + 8ull * q;
#pragma unroll
    for (int i = 0; i < 8; i++) dst[i] = below[i];
}

// Read roots, mark the fill leaf as needed (Phase 0, voxtree.rs:742-754).  Blocks
// nprep..nprep+ndesc run the old-tree descent over the level-3 slots instead, and the
// blocks after them (u16) record the leaf id of every value for the brick pass.
template <typename T>
__global__ void k_prepare(Pool<T> P, ChunkDev* chunks, uint32_t n, const uint64_t* roots,
                          const int32_t* pre_status, uint32_t nprep, uint32_t ndesc, LevelInfo L3,
                          const uint64_t* doff, OldNodes O) {
    if (blockIdx.x >= nprep + ndesc) {
        if (P.leaf_ids)
            for (uint32_t v = (blockIdx.x - nprep - ndesc) * blockDim.x + threadIdx.x; v < 65536u;
                 v += (gridDim.x - nprep - ndesc) * blockDim.x) {
                const uint32_t li = P.leaf_map[v];
                P.leaf_ids[v] = li ? make_leaf(li, tmg_gen(P.tmg[li])) : kEmpty;
            }
        return;
    }
    if (blockIdx.x >= nprep) {
        uint64_t stride = (uint64_t)ndesc * blockDim.x;
        for (uint64_t s = (uint64_t)(blockIdx.x - nprep) * blockDim.x + threadIdx.x; s < L3.nslots; s += stride)
            descend_body(P, chunks, L3, doff, n, O, roots, s);
        return;
    }
    // the sequence's control-block reset rides in block 0 (nothing else here reads it)
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_begin(P.ctl);
    if (sizeof(T) == 1 && blockIdx.x == 0)
        for (uint32_t v = threadIdx.x; v < 256; v += blockDim.x) {
            uint32_t li = P.leaf_map[v];
            P.ctl->leaf_ids8[v] = li ? make_leaf(li, tmg_gen(P.tmg[li])) : kEmpty;
        }
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (uint64_t)nprep * blockDim.x) {
        {
            ChunkDev& C = chunks[c];
            C.root_in = roots[C.orig];
            C.init = C.root_in;
            C.status = C.status0;
            // a depth-2 tree's height-2 node is its root (deeper trees: the descent)
            if (C.depth == 2 && C.has_patches && !C.has_fill) O.at[2][doff[(size_t)2 * n + c]] = C.root_in;
            if (P.per_tree) P.per_tree[C.orig] = 0;
            if (pre_status && pre_status[C.orig] < 0 && C.status >= 0) C.status = pre_status[C.orig];
            if (C.status >= 0 && C.has_fill) {
                uint32_t v = (uint32_t)C.fill;
                if (leaf_index(P, v) == 0) leaf_request(P, v);
            }
        }
    }
}

// ------------------------------------------------------------------ Phase 1

// One block: create the leaves marked as needed (interner/mod.rs:669-708), then
// fix each tree's Phase-1 start node.  The fill leaf of a batch that also has
// patches keeps its reference (H3, voxtree.rs:742-754).
// Create the leaves requested so far (one block; a dry run only counts them).  Returns
// the number created (or counted).
template <typename T>
__device__ inline uint32_t create_requested_leaves(const Pool<T>& P, uint32_t nvalues) {
    __shared__ uint32_t created;
    if (threadIdx.x == 0) created = 0;
    __syncthreads();
    const uint32_t nreq = sizeof(T) == 4 ? min(P.ctl->n_need, P.need_cap) : nvalues;
    for (uint32_t r = threadIdx.x; r < nreq; r += blockDim.x) {
        uint32_t v = r;
        uint64_t sl = 0;
        if constexpr (sizeof(T) == 4) {
            v = P.need_list[r];
            sl = ltab_find(P, v);
            if (sl == ~0ull || (uint32_t)P.ltab[sl] != kLeafNeed) continue;
        } else {
            if (!P.leaf_need[v]) continue;
            P.leaf_need[v] = 0;
            if (P.leaf_map[v] != 0) continue;
        }
        if (P.dry) {  // counted; an i32 request stays until the end of the apply (its placeholder)
            atomicAdd(&created, 1u);
            continue;
        }
        bool ok;
        uint32_t idx = alloc_node(P, &ok);
        if (ok) {
            // a recycled node keeps its old children (the free does not clear them): a
            // leaf has none
            uint64_t* ch = P.children + 8ull * idx;
#pragma unroll
            for (int j = 0; j < 8; j++) ch[j] = kEmpty;
            P.values[idx] = (T)v;
            P.tmg[idx] &= 0xFFFF0000u;  // a leaf: no types/mask, the generation kept
            P.refs[idx] = 0;
            if constexpr (sizeof(T) == 4) P.ltab[sl] = ((uint64_t)v << 32) | idx;
            else P.leaf_map[v] = idx;
            uint32_t k = atomicAdd(&P.ctl->n_new_leaves, 1u);
            P.new_leaves[k] = idx;
            atomicAdd(&created, 1u);
        } else if constexpr (sizeof(T) == 4) {
            P.ltab[sl] = kLeafTomb;
        }
    }
    __syncthreads();
    return created;
}

// One block: create the leaves marked as needed (interner/mod.rs:669-708), then
// fix each tree's Phase-1 start node.  The fill leaf of a batch that also has
// patches keeps its reference (H3, voxtree.rs:742-754).
template <typename T>
__global__ void k_leaf_alloc(Pool<T> P, ChunkDev* chunks, uint32_t n, uint32_t nvalues) {
    const uint32_t created = create_requested_leaves(P, nvalues);
    if (threadIdx.x == 0) {
        if (P.dry) {
            P.ctl->dry_new_leaves += created;
        } else {
            atomicAdd((unsigned long long*)&P.ctl->live_leaves, (unsigned long long)created);
            atomicAdd((unsigned long long*)stat(P.ctl, ST_LEAF_MISS), (unsigned long long)created);
        }
    }
    for (uint32_t c = threadIdx.x; c < n; c += blockDim.x) {
        ChunkDev& C = chunks[c];
        if (C.status < 0 || !C.has_fill) continue;
        uint32_t idx = leaf_index(P, (uint32_t)C.fill);
        if (idx == 0 && P.dry) {
            C.init = make_leaf(P.capacity + leaf_request(P, (uint32_t)C.fill), 0);
            continue;
        }
        if (idx == 0) continue;  // allocation failed (OOM flagged)
        C.init = make_leaf(idx, tmg_gen(P.tmg[idx]));
        if (C.has_patches && !P.dry) {
            ref_delta_logged(P, c, idx, 1, ST_H3);
            atomicAdd((unsigned long long*)stat(P.ctl, ST_H3), 1ull);
        }
    }
}

// ------------------------------------------------------------------ Phase 1 build

// Append a branch key to this tile's region of the candidate arrays.
// Compaction of a tile's candidate keys into the tile's region of K (order of
// discovery).  Called by every thread of the block.
__device__ inline void write_candidate(const Cands& K, uint64_t k, const uint64_t ch[8], uint32_t tm, uint64_t h,
                                       uint64_t slot) {
#pragma unroll
    for (int i = 0; i < 8; i++) key_word(K, k, i) = ch[i];
    key_word(K, k, 8) = tm;
    K.hash[k] = h;
    K.slot[k] = slot;
}

// The candidate was written by an earlier launch (every launch starts with coherent
// caches) or by this block before a barrier.
// Every word is loaded before any is compared (no early exit), so the nine loads are in
// flight together: one memory round trip instead of up to nine dependent ones.
__device__ inline bool key_matches_cand(const Cands& K, uint32_t k2, const uint64_t key[9]) {
    uint64_t w[9];
#pragma unroll
    for (int i = 0; i < 9; i++) w[i] = key_word(K, k2, i);
    bool eq = true;
#pragma unroll
    for (int i = 0; i < 9; i++) eq &= w[i] == key[i];
    return eq;
}

template <int TILE>
__device__ inline uint32_t emit_candidate(const Cands& K, uint32_t* s_cnt, bool pred, const uint64_t ch[8],
                                          uint32_t tm, uint64_t slot, uint32_t blk, uint64_t h) {
    uint32_t pos = wave_append(s_cnt, pred);
    if (pred) write_candidate(K, (uint64_t)blk * TILE + pos, ch, tm, h, slot);
    return pos;
}

template <int TILE>
__device__ inline uint32_t emit_candidate(const Cands& K, uint32_t* s_cnt, bool pred, const uint64_t ch[8],
                                          uint32_t tm, uint64_t slot, uint32_t blk) {
    return emit_candidate<TILE>(K, s_cnt, pred, ch, tm, slot, blk, pred ? branch_hash(ch, tm) : 0ull);
}

// Block-wide merge of equal branch keys before they become candidates: a set in LDS
// maps a key's full 64-bit hash to the candidate of the first lane that inserted it
// (the key's leader).  A lane that finds its hash present is a follower once a full-key
// compare against the leader's candidate (written by this block before a barrier)
// confirms it; the merge crosses trees, since a key's node is the same for every tree.
// Returns the set slot (*lead when this lane inserted it), or -1 when no slot was found
// within a bounded scan (the key is then emitted on its own).  Hash 0 marks a free slot.
template <uint32_t N>
__device__ inline int keyset_insert(uint64_t* s_h, uint64_t h, bool* lead) {
    const uint64_t hk = h ? h : 1ull;
    uint32_t s = (uint32_t)hk & (N - 1);
    *lead = false;
#pragma unroll 1
    for (uint32_t g = 0; g < 32; g++) {
        uint64_t cur = s_h[s];
        if (cur == 0) {
            cur = atomicCAS((unsigned long long*)&s_h[s], 0ull, (unsigned long long)hk);
            if (cur == 0) {
                *lead = true;
                return (int)s;
            }
        }
        if (cur == hk) return (int)s;
        s = (s + 1) & (N - 1);
    }
    return -1;
}

// Equality of two bricks' eight values (any voxel width).
template <typename V>
__device__ inline bool vox8_equal(const V& a, const V& b) {
    static_assert(sizeof(V) % 4 == 0, "whole dwords");
    const uint32_t* x = (const uint32_t*)&a;
    const uint32_t* y = (const uint32_t*)&b;
    bool eq = true;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(V) / 4); i++) eq &= x[i] == y[i];
    return eq;
}

// Tile statistics in LDS, flushed with one striped atomic per counter per block.
struct TileStats {
    uint32_t v[8];
};
__device__ inline void tile_count(uint32_t* s, int id, bool pred) {
    uint64_t b = __ballot(pred);
    if (b && (int)__lane_id() == __ffsll((unsigned long long)b) - 1) atomicAdd(&s[id], (uint32_t)__popcll(b));
}

// Per brick (voxtree.rs:778-896): descend the old tree to the brick (792-822), then
// the all-same brick becomes a leaf; otherwise the brick's children are built from
// the old node, a split leaf, or nothing, and the changed slots get leaves.
// Returns: 0 = not pushed, 1 = leaf result in *out, 2 = branch key in ch/tm,
// 3 = a needed leaf does not exist yet (marked; the brick is retried later).
// Reference-count corrections of the reference's split-leaf quirks are queued in
// *drift_idx / *drift_delta (applied once the tree's status is final).
// The leaf id of a value, or kEmpty when no leaf holds it.  leaf_ids (u8: LDS, u16: the
// table k_prepare writes) caches it for every value; otherwise the direct-mapped table is
// read.
template <typename T>
__device__ inline uint64_t leaf_lookup(const Pool<T>& P, const uint64_t* leaf_ids, uint32_t v) {
    if (leaf_ids) return leaf_ids[v];
    uint32_t li = leaf_index(P, v);
    return li ? make_leaf(li, tmg_gen(P.tmg[li])) : kEmpty;
}

template <typename T>
__device__ inline int brick_work(const Pool<T>& P, const ChunkDev& C, uint32_t m, const uint32_t v[8],
                                 uint64_t base, uint64_t* out, uint64_t ch[8], uint32_t* tm,
                                 uint32_t* drift_idx, int32_t* drift_delta, int* flags,
                                 const uint64_t lids[8], bool request = true) {
    *drift_idx = kNoIdx;
    bool all_same = m == 0xFFu;
#pragma unroll
    for (int i = 1; i < 8; i++) all_same = all_same && v[i] == v[0];
    if (all_same) {
        uint64_t lid = lids[0];
        if (lid == kEmpty) {
            const uint32_t key = request ? leaf_request(P, v[0]) : 0u;
            if (!P.dry) return 3;
            *out = make_leaf(P.capacity + key, 0);  // dry run: a placeholder for the new leaf
        } else {
            *out = lid;
        }
        *flags |= 1;  // collapsed
        return 1;
    }
    // the split leaf's value is known without its id (the fill leaf may not exist yet)
    bool missing = false;
    bool fill_split = base == kFillMark;
    if (fill_split) {
        uint32_t fi = leaf_index(P, (uint32_t)C.fill);
        if (fi == 0) {
            const uint32_t key = request ? leaf_request(P, (uint32_t)C.fill) : 0u;
            if (P.dry) {
                base = make_leaf(P.capacity + key, 0);
            } else {
                missing = true;
                base = make_leaf(0, 0);
            }
        } else {
            base = make_leaf(fi, tmg_gen(P.tmg[fi]));
        }
    }
    bool split = bid_is_leaf(base);
    uint32_t split_value = 0;
    if (split) split_value = fill_split ? (uint32_t)C.fill : (uint32_t)P.values[bid_index(base)];
    // A completely written brick replaces every child of an old brick-level node: its key
    // is (the leaves of its values, types = mask = 0xFF) whatever the old children were, so
    // they are not loaded.  The old children only tell whether the brick is unchanged (every
    // new leaf equal to the old child); such a brick is built anyway, its probe finds the
    // old node itself, and the level above counts a child equal to the old one as
    // untouched, as the reference leaves an unchanged brick unpushed (voxtree.rs:865-868).
    if (!split && m == 0xFFu) base = kEmpty;
    uint32_t types = 0, mask = 0;
    if (split) {
        for (int i = 0; i < 8; i++) ch[i] = base;
        types = mask = 0xFFu;
    } else if (base != kEmpty) {
        for (int i = 0; i < 8; i++) ch[i] = P.children[8ull * bid_index(base) + i];
        types = bid_types(base);
        mask = bid_mask(base);
    } else {
        for (int i = 0; i < 8; i++) ch[i] = kEmpty;
    }
    uint32_t modified = 0;
    for (int i = 0; i < 8; i++) {
        if (!(m & (1u << i))) continue;
        // a live leaf is the one leaf of its value: the child holds v[i] exactly when it
        // is that leaf (the children of a brick are leaves)
        uint64_t lid = lids[i];
        if (ch[i] != kEmpty) {
            bool same = split ? split_value == v[i] : (lid != kEmpty && ch[i] == lid);
            if (same) continue;
        }
        if (lid == kEmpty) {
            const uint32_t key = request ? leaf_request(P, v[i]) : 0u;
            if (!P.dry) {
                missing = true;
                continue;
            }
            ch[i] = make_leaf(P.capacity + key, 0);
        } else {
            ch[i] = lid;
        }
        types |= 1u << i;
        mask |= 1u << i;
        modified |= 1u << i;
    }
    if (missing) return 3;
    uint32_t pcm = (uint32_t)__popc(m), pcmod = (uint32_t)__popc(modified);
    if (modified == 0) {
        // no change: not pushed; the split leaf keeps its 8-pc(s) refs (H3)
        if (split) {
            *drift_idx = bid_index(base);
            *drift_delta = (int32_t)(8u - pcm);
            *flags |= 8;
        }
        return 0;
    }
    if (split && pcmod != pcm) {
        // slots equal to the split leaf stay without a ref (H4)
        *drift_idx = bid_index(base);
        *drift_delta = (int32_t)pcmod - (int32_t)pcm;
        *flags |= 16;
    }
    bool uniform = types == 0xFFu;
    for (int i = 1; i < 8; i++) uniform = uniform && ch[i] == ch[0];
    if (uniform) *flags |= 4;  // H2
    *tm = (types << 8) | mask;
    return 2;
}

// Brick slots per lane of the brick pass's several-slot form: 4 (2 for 512-lane
// tiles), so a block takes 512 to 1024 slots.
template <int TILE>
constexpr int brick_rounds() { return TILE >= 512 ? 2 : 4; }

// Two phases per block of R*TILE consecutive brick slots (R per lane, slot =
// block start + j*TILE + lane in round j).  Phase 1 issues every round's loads at
// once (edit mask, the 8 values and the old height-2 node above the brick) and
// settles the bricks that need no old tree: untouched ones and all-same ones whose
// leaf exists.  The rest (mixed bricks; all-same ones whose leaf is still missing)
// are queued in LDS with their loaded data and take phase 2, one per lane, in
// rounds of TILE: the old-tree descent and brick_work.  Round r of phase 2 emits
// into candidate tile blk*R + r; a deferred brick keeps that tile (its round has a
// free place for it) for the retry.  The pass is bound by memory latency, so it runs
// as many waves per SIMD as fit without spilling: 5 for one slot per lane with u8 and
// u16 values (their eight leaf ids are loaded together), 6 with 4-byte values; the
// four-slot form's phase-1 loads and phase-2 keys take more registers and its queue more
// LDS (5 waves for u8 and u16, 3 for 4-byte values).
template <typename T, int TILE, int R>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(R == 1 ? (sizeof(T) == 4 ? 6 : 5) : (sizeof(T) == 4 ? 3 : 5), 8))) void k_brick_build(Pool<T> P, ChunkDev* chunks, LevelInfo L,
                                                       uint64_t* res1, Cands K, uint32_t* tile_cnt,
                                                       uint64_t* deferred, Drift* drift,
                                                       const uint64_t* old2) {
    // a brick's 8 values
    typedef typename std::conditional<sizeof(T) == 1, uint2, typename std::conditional<sizeof(T) == 2, uint4,
                                                                                          Vox8x4>::type>::type Vox8;
    static_assert(sizeof(Vox8) == 8 * sizeof(T), "8 voxels");
    static_assert(R * TILE <= 32768, "queued brick index is 15 bits (bit 15: all eight values equal)");
    __shared__ uint32_t s_nq;
    __shared__ uint32_t s_cnt[R];
    __shared__ uint32_t s_st[8];
    // u8: the leaf id of every value, from LDS (the value map does not change here)
    __shared__ uint64_t s_leaf[sizeof(T) == 1 ? 256 : 1];
    __shared__ uint64_t s_up[R * TILE];
    __shared__ Vox8 s_v[R * TILE];
    __shared__ uint16_t s_idx[R * TILE];
    __shared__ uint8_t s_m[R * TILE];
    // the block's key set (keyset_insert): a quarter as many slots as the block has
    // bricks, enough for the few distinct keys of a dense region (a block of all-distinct
    // keys fills it and emits the rest unmerged), and small enough that five blocks of
    // 256 lanes with u16 values share a CU's LDS
    constexpr uint32_t kSetB = R * TILE / 4;
    __shared__ uint64_t s_kh[kSetB];
    __shared__ uint32_t s_kc[kSetB];
    __shared__ uint16_t s_kq[kSetB];  // the leader's queue entry (its mask and values in LDS)
    for (uint32_t i = threadIdx.x; i < kSetB; i += TILE) s_kh[i] = 0;
    uint64_t lm[sizeof(T) == 1 ? (256 + TILE - 1) / TILE : 1];
    if (sizeof(T) == 1)
        for (uint32_t i = 0, v = threadIdx.x; v < 256; i++, v += TILE) lm[i] = P.ctl->leaf_ids8[v];
    const uint64_t b0 = (uint64_t)blockIdx.x * (R * TILE);
    const BlockGroup BG = block_group(L, min(b0, L.nslots - 1), min(b0 + R * TILE, L.nslots) - 1);
    // phase 1 loads, without branches (lanes past the end load the last slot), in two
    // rounds for all R slots of a lane: the chunk records, then the bricks' masks and
    // values and the height-2 node above each brick.  (Taken one slot after the other,
    // every wait for a record also waits for the previous slot's data: 2R round trips.)
    uint32_t m[R];
    Vox8 vw[R];
    uint64_t up[R];
    {
        uint32_t cj[R], pj[R];
        uint64_t uj[R];
        bool lj[R];
        bool uni = true;  // every slot's chunk is one per wavefront: scalar loads
#pragma unroll
        for (int j = 0; j < R; j++) {
            const uint64_t slot = b0 + (uint64_t)j * TILE + threadIdx.x;
            lj[j] = slot < L.nslots;
            const uint64_t sl = lj[j] ? slot : L.nslots - 1;
            const GroupPos G = group_of(L, BG, sl);
            const uint64_t local = sl - G.base;
            const uint32_t D = G.depth;
            const uint32_t sh = 3 * (D - 1);
            cj[j] = G.chunk_base + (uint32_t)(local >> sh);
            pj[j] = (uint32_t)(local & ((1ull << sh) - 1));
            // the height-2 slot above this brick follows from the slot (a depth-1 tree
            // has none; a tree with a fill does not use it)
            uj[j] = D >= 2 ? G.next + (local >> 3) : 0;
            uni = uni && __ballot(cj[j] != __builtin_amdgcn_readfirstlane(cj[j])) == 0;
        }
        const uint8_t* hm[R];
        const void* hv[R];
        int32_t hs[R];
        if (uni) {
#pragma unroll
            for (int j = 0; j < R; j++) {
                ConstChunk& r = *(ConstChunk*)(chunks + __builtin_amdgcn_readfirstlane(cj[j]));
                hm[j] = r.masks;
                hv[j] = r.values;
                hs[j] = r.status;
            }
        } else {
#pragma unroll
            for (int j = 0; j < R; j++) {
                const ChunkDev& r = chunks[cj[j]];
                hm[j] = r.masks;
                hv[j] = r.values;
                hs[j] = r.status;
            }
        }
        // (global-memory loads: a generic load may return out of order with the others,
        // and a wait for any earlier load then waits for all of them)
        typedef const __attribute__((address_space(1))) uint8_t GByte;
        typedef unsigned int VoxWords __attribute__((ext_vector_type(sizeof(Vox8) / 4)));
        typedef const __attribute__((address_space(1))) VoxWords GVox;
#pragma unroll
        for (int j = 0; j < R; j++) {
            m[j] = ((GByte*)hm[j])[2ull * pj[j]];
            const VoxWords w = *(GVox*)((const T*)hv[j] + 8ull * pj[j]);
            __builtin_memcpy(&vw[j], &w, sizeof(Vox8));
            up[j] = old2[uj[j]];
        }
        // (used here, the masks are loaded with the rest: the compiler would otherwise
        // load a mask only for a live slot, and wait for it before the next slot's loads)
#pragma unroll
        for (int j = 0; j < R; j++) asm volatile("" ::"v"(m[j]));
#pragma unroll
        for (int j = 0; j < R; j++)
            if (!lj[j] || hs[j] < 0) m[j] = 0;
    }
    if (sizeof(T) == 1)
        for (uint32_t i = 0, vv = threadIdx.x; vv < 256; i++, vv += TILE) s_leaf[vv] = lm[i];
    const uint64_t* leaf_ids = sizeof(T) == 1 ? s_leaf : P.leaf_ids;
    if (threadIdx.x < 8) s_st[threadIdx.x] = 0;
    if (threadIdx.x < R) s_cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_nq = 0;
    __syncthreads();
    uint32_t ncollapse = 0;
    // the leaf of every all-same brick's value, for all R slots before any is used (u8
    // and u16: one load from the value's table entry, made for every slot)
    uint64_t lidj[R];
    bool asj[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
        uint32_t v[8];
        load8<T>((const T*)&vw[j], v);
        bool all_same = m[j] == 0xFFu;
#pragma unroll
        for (int i = 1; i < 8; i++) all_same = all_same && v[i] == v[0];
        asj[j] = all_same;
        if constexpr (sizeof(T) <= 2) {
            const uint64_t l = leaf_ids[v[0]];
            lidj[j] = all_same ? l : kEmpty;
        } else {
            lidj[j] = all_same ? leaf_lookup(P, leaf_ids, v[0]) : kEmpty;
        }
    }
#pragma unroll
    for (int j = 0; j < R; j++) {
        const uint64_t slot = b0 + (uint64_t)j * TILE + threadIdx.x;
        const uint64_t lid = lidj[j];
        const bool queue = m[j] != 0 && lid == kEmpty;
        const uint32_t q = wave_append(&s_nq, queue);
        if (queue) {
            s_idx[q] = (uint16_t)((j * TILE + threadIdx.x) | (asj[j] ? 0x8000u : 0u));
            s_m[q] = (uint8_t)m[j];
            s_v[q] = vw[j];
            s_up[q] = up[j];
        } else if (slot < L.nslots) {
            res1[slot] = m[j] != 0 ? lid : kInvalid;
            ncollapse += m[j] != 0;
        }
    }
    {
        const uint32_t t = wave_sum(ncollapse);
        if (__lane_id() == 0 && t) atomicAdd(&s_st[1], t);
    }
    __syncthreads();
    const uint32_t nq = s_nq;
    for (int r = 0; r < R; r++) {
        if ((uint32_t)r * TILE >= nq) break;  // uniform
        const uint32_t q = (uint32_t)r * TILE + threadIdx.x;
        uint64_t out = kInvalid, slot = 0;
        bool cand = false, defer = false;
        int flags = 0;
        uint64_t ch[8];
        uint32_t tm = 0, c = 0;
        if (q < nq) {
            const uint32_t qi = s_idx[q];
            slot = b0 + (qi & 0x7FFFu);
            const uint32_t mm = s_m[q];
            const Vox8 w = s_v[q];
            const uint64_t upq = s_up[q];
            uint32_t v[8];
            load8<T>((const T*)&w, v);
            const GroupPos G = group_of(L, BG, slot);
            const uint64_t local = slot - G.base;
            const uint32_t D = G.depth;
            const uint32_t sh = 3 * (D - 1);
            c = G.chunk_base + (uint32_t)(local >> sh);
            const uint32_t p = (uint32_t)(local & ((1ull << sh) - 1));
            // one round of independent loads: the tree's record, the leaf ids of the values
            // and the node one hop below the recorded height-2 node (used unless the tree
            // has a fill)
            const ChunkDev& C = chunks[c];
            const bool c_fill = C.has_fill != 0;
            const uint64_t c_root = C.root_in;
            // (u8 and u16: a table load for every voxel, edited or not, so that the eight
            // are issued together; brick_work reads the leaf ids of edited voxels only)
            uint64_t lids[8];
            if constexpr (sizeof(T) <= 2) {
#pragma unroll
                for (int i = 0; i < 8; i++) lids[i] = leaf_ids[v[i]];
            } else {
#pragma unroll
                for (int i = 0; i < 8; i++) lids[i] = (mm >> i) & 1u ? leaf_lookup(P, leaf_ids, v[i]) : kEmpty;
            }
            const bool up_branch = D >= 2 && upq != kEmpty && !bid_is_leaf(upq);
            const bool up_ok = up_branch && bid_index(upq) < P.capacity;
            const uint64_t below = P.children[8ull * (up_ok ? bid_index(upq) : 0u) + (p & 7u)];
            // (used here, the record's fields and the node below are loaded with the leaf
            // ids: the compiler would otherwise load each in the branch that needs it,
            // one round trip after the other)
            asm volatile("" ::"v"(below), "v"(c_root), "v"((uint32_t)c_fill));
            const bool all_same = (qi >> 15) != 0;  // taken in phase 1
            uint64_t base = kEmpty;
            bool h1 = false;
            if (!all_same) {
                if (c_fill) {
                    base = kFillMark;
                } else {
                    // the equivalent node at brick level: a leaf above it splits;
                    // a leaf at brick level is the reference's panic
                    uint64_t cur = D >= 2 ? upq : c_root;
                    if (up_branch) {
                        if (!up_ok) {
                            atomicOr(&P.ctl->err, kErrPoison);
                            cur = kEmpty;
                        } else {
                            cur = below;
                        }
                        // BlockId::types() on a leaf brick panics (block_id.rs:273-276)
                        h1 = cur != kEmpty && bid_is_leaf(cur);
                    } else if (D == 1) {
                        h1 = cur != kEmpty && bid_is_leaf(cur);
                    }
                    base = cur;
                }
            }
            if (h1) {
                chunks[c].status = -5;
                P.ctl->h1_fail = 1u;
            } else {
                uint32_t didx = kNoIdx;
                int32_t dd = 0;
                int rc;
                bool lids_all = true;
#pragma unroll
                for (int i = 0; i < 8; i++) lids_all = lids_all && lids[i] != kEmpty;
                if (mm == 0xFFu && !all_same && base != kFillMark && !bid_is_leaf(base) && lids_all) {
                    // brick_work's result for a completely written brick of distinct values
                    // over a branch or nothing, whose leaves all exist (the dense case): the
                    // key is the eight leaves, types = mask = 0xFF.  (Never 8 equal leaves,
                    // brick_work's H2 case: a live leaf is the one leaf of its value, and
                    // the values are not all the same.)
#pragma unroll
                    for (int i = 0; i < 8; i++) ch[i] = lids[i];
                    tm = 0xFFFFu;
                    rc = 2;
                } else {
                    rc = brick_work(P, C, mm, v, base, &out, ch, &tm, &didx, &dd, &flags, lids);
                }
                if (didx != kNoIdx && !P.dry) {
                    uint32_t pos = atomicAdd(&P.ctl->n_drift, 1u);
                    drift[pos] = Drift{c, didx, dd, (uint32_t)(flags & 8 ? ST_H3 : ST_H4)};
                }
                if (rc == 2) {
                    cand = true;
                    out = res_pending(0xFFFFFFFFu);  // replaced by the probe
                } else if (rc == 3) {
                    defer = true;
                    out = base;  // kept for the retry
                }
            }
        }
        // Equal keys of the block (all rounds, any tree) share one candidate: the leader
        // emits it, followers point at it (a pending result, resolved through the
        // candidate's id once the probe has run).
        uint64_t key[9];
#pragma unroll
        for (int i = 0; i < 8; i++) key[i] = ch[i];
        key[8] = tm;
        const uint64_t hk = cand ? branch_hash(ch, tm) : 0ull;
        bool lead = false;
        const int ks = cand && P.brick_merge_rounds ? keyset_insert<kSetB>(s_kh, hk, &lead) : -1;
        const uint32_t tile = blockIdx.x * R + (uint32_t)r;
        const uint32_t pos = emit_candidate<TILE>(K, &s_cnt[r], cand && (lead || ks < 0), ch, tm, slot, tile, hk);
        if (lead) {
            s_kc[ks] = (uint32_t)((uint64_t)tile * TILE) + pos;
            s_kq[ks] = (uint16_t)q;
        }
        __syncthreads();  // the leaders' candidates and their indices are visible to the block
        bool follow = false;
        if (cand && ks >= 0 && !lead) {
            const uint32_t k2 = s_kc[ks];
            // Two completely written bricks have equal keys exactly when their values are
            // equal: every child is then the leaf of its value (a live leaf is the one leaf
            // of its value; a dry run's placeholder is a function of the value too) and
            // types = mask = 0xFF, whatever the old brick was.  Such a pair is compared in
            // LDS; any other pair compares the leader's candidate key.
            const uint32_t lq = s_kq[ks];
            const bool both_full = s_m[q] == 0xFFu && s_m[lq] == 0xFFu;
            if (both_full ? vox8_equal(s_v[q], s_v[lq]) : key_matches_cand(K, k2, key)) {
                out = res_pending(k2);
                follow = true;
            } else {  // another key with the same 64-bit hash: a candidate of its own
                const uint32_t p2 = atomicAdd(&s_cnt[r], 1u);
                write_candidate(K, (uint64_t)tile * TILE + p2, ch, tm, hk, slot);
            }
        }
        if (q < nq) res1[slot] = out;
        const uint32_t dpos = wave_append(&P.ctl->n_deferred, defer);
        if (defer) deferred[dpos] = ((uint64_t)tile << 32) | (uint32_t)slot;
        tile_count(s_st, 0, cand);
        tile_count(s_st, 1, (flags & 1) != 0);
        tile_count(s_st, 2, (flags & 4) != 0);
        tile_count(s_st, 3, follow);
    }
    __syncthreads();
    if (threadIdx.x < R) tile_cnt[blockIdx.x * R + threadIdx.x] = s_cnt[threadIdx.x];
    const int ids[6] = {ST_CAND, ST_COLLAPSE, ST_H2, ST_MERGED, ST_CAND1, ST_MERGED1};
    const int src[6] = {0, 1, 2, 3, 0, 3};
    if (threadIdx.x < 6 && s_st[src[threadIdx.x]])
        atomicAdd((unsigned long long*)stat(P.ctl, ids[threadIdx.x]), (unsigned long long)s_st[src[threadIdx.x]]);
}

// Bricks that waited for new leaves, and the queued reference-count corrections
// of trees that did not fail (their status is final now).  A deferred entry is
// (candidate tile << 32) | slot: the tile of the brick pass's round that queued it.
template <typename T, int TILE>
__global__ void k_brick_retry(Pool<T> P, const ChunkDev* chunks, LevelInfo L, uint64_t* res1, Cands K,
                              uint32_t* tile_cnt, const uint64_t* deferred, const Drift* drift) {
    mark_owner_base(P);
    uint32_t nd = P.ctl->n_deferred;
    VX_LOOP(i, nd) {
        if (i < nd) {
            const uint64_t e = deferred[i];
            const uint64_t slot = (uint32_t)e;
            const uint64_t tile = e >> 32;
            const GroupPos G = group_pos(L, slot);
            uint64_t local = slot - G.base;
            uint32_t D = G.depth;
            uint32_t sh = 3 * (D - 1);
            uint32_t c = G.chunk_base + (uint32_t)(local >> sh);
            uint32_t p = (uint32_t)(local & ((1ull << sh) - 1));
            const ChunkDev& C = chunks[c];
            uint64_t out = kInvalid;
            if (C.status >= 0) {
                uint32_t m = C.masks[2ull * p];
                uint32_t v[8];
                load8<T>((const T*)C.values + 8ull * p, v);
                uint64_t ch[8];
                uint32_t tm = 0, didx;
                int32_t dd;
                int flags = 0;
                // no new requests here: k_leaf_alloc has run, a leaf still missing could
                // not be allocated and the apply aborts
                uint64_t lids[8];
#pragma unroll
                for (int i = 0; i < 8; i++)
                    lids[i] = (m >> i) & 1u ? leaf_lookup(P, (const uint64_t*)nullptr, v[i]) : kEmpty;
                int r = brick_work(P, C, m, v, res1[slot], &out, ch, &tm, &didx, &dd, &flags, lids, false);
                if (didx != kNoIdx) {
                    ref_delta_logged(P, c, didx, dd, flags & 8 ? ST_H3 : ST_H4);
                    atomicAdd((unsigned long long*)stat(P.ctl, flags & 8 ? ST_H3 : ST_H4), 1ull);
                }
                if (flags & 1) atomicAdd((unsigned long long*)stat(P.ctl, ST_COLLAPSE), 1ull);
                if (r == 2) {
                    uint32_t pos = atomicAdd(&tile_cnt[tile], 1u);
                    uint64_t k = tile * TILE + pos;
                    for (int j = 0; j < 8; j++) key_word(K, k, j) = ch[j];
                    key_word(K, k, 8) = tm;
                    K.hash[k] = branch_hash(ch, tm);
                    K.slot[k] = slot;
                    atomicAdd((unsigned long long*)stat(P.ctl, ST_CAND), 1ull);
                    atomicAdd((unsigned long long*)stat(P.ctl, ST_CAND1), 1ull);
                    out = res_pending(0xFFFFFFFFu);
                } else if (r == 3) {
                    atomicOr(&P.ctl->abort_, 1u);  // the leaf could not be created
                    out = kInvalid;
                }
            }
            res1[slot] = out;
        }
    }
    uint32_t ndr = P.ctl->n_drift;
    VX_LOOP(j, ndr) {
        if (j < ndr) {
            Drift d = drift[j];
            if (chunks[d.chunk].status >= 0) {
                atomicAdd(&P.refs[d.idx], (uint32_t)d.delta);
                atomicAdd((unsigned long long*)stat(P.ctl, (int)d.stat), 1ull);
            }
        }
    }
}

// ------------------------------------------------------------------ interning

template <typename T>
__device__ inline bool key_matches_node(const Pool<T>& P, uint32_t idx, const uint64_t key[9], uint32_t* gen) {
    // the node's 64-byte children row and its types/mask + generation word in one round
    const uint32_t tg = P.tmg[idx];
    const uint32_t tm = tmg_tm(tg);
    *gen = tmg_gen(tg);
    const ulonglong2* c = (const ulonglong2*)(P.children + 8ull * idx);
    ulonglong2 w[4];
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = c[i];
    bool eq = tm == (uint32_t)key[8];
#pragma unroll
    for (int i = 0; i < 4; i++) eq &= w[i].x == key[2 * i] && w[i].y == key[2 * i + 1];
    return eq;
}


// get_or_create_branch (interner/mod.rs:716-829) for one launch's candidates:
// committed keys resolve to the existing node, the first inserter of a new key
// owns it (allocates its index), later equal keys point at the owner.  Equal keys
// inside one wavefront are merged first, so only one lane per key probes.

template <typename T>
__device__ inline int match_slot(const Pool<T>& P, uint32_t level, uint64_t e, uint32_t tag,
                                 const uint64_t key[9], uint64_t* result) {
    if (e == kSlotEmpty || e == kSlotTomb || slot_tag_of(e) != tag) return -1;
    uint32_t tm = (uint32_t)key[8];
    if (slot_pending(e)) {
        uint32_t lv = slot_level(e);
        if (!key_matches_cand(P.levels[lv], slot_payload(e), key)) return -1;
        // same level: point at the owner; an earlier level (dry runs keep those pending):
        // its owner's id is final already
        *result = lv == level ? res_pending(slot_payload(e)) : P.levels[lv].id[slot_payload(e)];
        return lv == level ? 1 : 0;
    }
    uint32_t gen;
    if (!key_matches_node(P, slot_payload(e), key, &gen)) return -1;
    *result = make_branch(slot_payload(e), gen, tm);
    return 0;
}

// The scan reads the table with plain loads.  A slot changes at most once during
// one probe launch (empty or tombstone -> this launch's pending key; commits and GC
// run in other launches, and a launch starts with coherent caches), so a plain load
// returns either the slot's value at launch start or its final value.  Every claim of
// a slot that looked free is an atomicCAS, which returns the slot's true value; after
// a failed claim the scan continues behind that slot, whose value is now known.
// (A tombstone is claimed on the same terms: it cannot turn back into a tombstone.)
template <typename T>
__device__ inline int probe_table(const Pool<T>& P, uint32_t level, uint64_t k, const uint64_t key[9],
                                  uint64_t h, uint64_t* result, uint64_t* tslot,
                                  uint32_t* probes, uint32_t* claims) {
    uint32_t tag = slot_tag(h);
    uint64_t s = h & P.tmask;
    uint64_t first_free = ~0ull;  // the first tombstone since the scan (re)started
    uint64_t scanned = 0;         // slots scanned since then
    for (uint64_t guard = 0; guard <= 4 * (P.tmask + 1); guard++) {
        uint64_t e = P.tab[s];
        (*probes)++;
        const bool tomb = e == kSlotTomb;
        if (tomb && first_free == ~0ull) first_free = s;
        // A key that is not present is claimed at the first empty slot, or at the first
        // tombstone before it; when a whole table's worth of slots held no empty one
        // (tombstones pile up between rebuilds), at the first tombstone seen.
        if (e == kSlotEmpty || (first_free != ~0ull && scanned > P.tmask)) {
            const uint64_t target = first_free != ~0ull ? first_free : s;
            const uint64_t expected = first_free != ~0ull ? kSlotTomb : kSlotEmpty;
            uint64_t seen = expected;
            bool won = __hip_atomic_compare_exchange_strong(
                &P.tab[target], &seen, slot_make(tag, true, level, (uint32_t)k), __ATOMIC_RELAXED,
                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            *claims += won ? 1u : 0x10000u;  // low half: won, high half: lost
            if (won) {
                // the caller allocates the owner's node; the slot is kept when the
                // allocation fails too (the abort path restores it), and a dry run
                // restores it after the run
                *tslot = target | (expected == kSlotTomb ? kTslotTomb : 0ull);
                return 2;
            }
            int m = match_slot(P, level, seen, tag, key, result);
            if (m >= 0) return m;
            s = (target + 1) & P.tmask;  // another key took the slot: go on behind it
            first_free = ~0ull;
            scanned = 0;
            continue;
        }
        if (!tomb) {
            int m = match_slot(P, level, e, tag, key, result);
            if (m >= 0) return m;
        }
        s = (s + 1) & P.tmask;
        scanned++;
    }
    atomicOr(&P.ctl->err, kErrTableFull | kErrPoison);
    *result = kInvalid;
    return -1;
}

// u8/u16 interners keep one-leaf keys in a direct-mapped index, single[(slot << 8s) |
// value of the leaf], instead of the hashed table: the pair is the whole key (a live
// leaf is the one leaf of its value), so a lookup is one read of a small, cached array.
// A dry run's placeholder leaf (index capacity + value) carries its value in its index.
template <typename T>
__device__ inline uint64_t single_index(const Pool<T>& P, uint32_t tm, const uint64_t* ch) {
    const uint32_t slot = (uint32_t)__ffs(tm & 0xFFu) - 1u;
    const uint32_t li = bid_index(ch[slot]);
    const uint32_t v = li >= P.capacity ? li - P.capacity : (uint32_t)P.values[li];
    return ((uint64_t)slot << (8 * sizeof(T))) | v;
}

// get_or_create_branch for a one-leaf key: a committed entry is a hit, a claim of this
// launch at the same level a duplicate (at an earlier level -- a dry run keeps those --
// its owner's id is final), an empty entry is claimed with atomicCAS.
template <typename T>
__device__ inline int probe_single(const Pool<T>& P, uint32_t level, uint64_t k, const uint64_t key[9],
                                   uint64_t* result, uint64_t* tslot, uint32_t* probes, uint32_t* claims) {
    const uint64_t ix = single_index(P, (uint32_t)key[8], key);
    uint64_t e = P.single[ix];
    (*probes)++;
    if (e == 0) {
        uint64_t seen = 0;
        const uint64_t want = kSinglePend | ((uint64_t)level << 32) | (uint32_t)k;
        const bool won = __hip_atomic_compare_exchange_strong(&P.single[ix], &seen, want, __ATOMIC_RELAXED,
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *claims += won ? 1u : 0x10000u;
        if (won) {
            *tslot = kTslotSingle | ix;
            return 2;
        }
        e = seen;
    }
    if (e & kSinglePend) {
        const uint32_t lv = (uint32_t)(e >> 32) & 7u, c = (uint32_t)e;
        *result = lv == level ? res_pending(c) : P.levels[lv].id[c];
        return lv == level ? 1 : 0;
    }
    *result = e;
    return 0;
}

// The probe's statistics: each wavefront adds its sums to the block's LDS counters as
// it ends; the last one to end adds them to the global counters.
constexpr uint32_t kProbeStats = 7;

template <typename T>
__device__ inline void probe_stats_flush(const Pool<T>& P, uint32_t h, uint32_t* s_stat, uint32_t* s_done,
                                         uint32_t probes, uint32_t claims, uint32_t lost, uint32_t tomb,
                                         uint32_t hit, uint32_t dup, uint32_t miss) {
    const uint32_t v[kProbeStats] = {wave_sum(probes), wave_sum(claims), wave_sum(lost), wave_sum(tomb),
                                     wave_sum(hit), wave_sum(dup), wave_sum(miss)};
    uint32_t last = 0;
    if (__lane_id() == 0) {
#pragma unroll
        for (uint32_t i = 0; i < kProbeStats; i++)
            if (v[i]) atomicAdd(&s_stat[i], v[i]);
        __threadfence_block();
        last = atomicAdd(s_done, 1u) == blockDim.x / 64 - 1 ? 1u : 0u;
    }
    last = __shfl(last, 0);
    if (!last) return;
    __threadfence_block();
    const uint32_t lane = __lane_id();
    if (lane < kProbeStats) {
        const uint32_t x = atomicAdd(&s_stat[lane], 0u);  // read through the LDS atomic unit
        const int ids[kProbeStats] = {ST_PROBES, ST_CLAIMS, ST_CLAIMS_LOST, ST_TOMB, ST_HIT, ST_DUP, ST_MISS};
        const int ids1[kProbeStats] = {-1, -1, -1, -1, ST_HIT1, ST_DUP1, ST_MISS1};
        if (x) {
            // a claimed tombstone is subtracted
            const unsigned long long add = lane == 3 ? (unsigned long long)(-(long long)x) : (unsigned long long)x;
            atomicAdd((unsigned long long*)stat(P.ctl, ids[lane]), add);
            if (h == 1 && ids1[lane] >= 0) atomicAdd((unsigned long long*)stat(P.ctl, ids1[lane]), add);
        }
    }
}

// Probe of the candidates of tiles [t0, t0+tn) of one level (tn <= 64), packed
// densely over the block's lanes: sparse tiles fill whole wavefronts, and equal keys
// meet in the same pass more often.  Per pass over TILE of them: merge equal keys of
// the pass through an LDS hash set (dedup_rounds 0 turns the merge off), look the
// leaders' keys up / insert them (get_or_create_branch, interner/mod.rs:716-829),
// then allocate the owners' nodes
// with one counter update for the whole block (a counter every block of a launch
// updates is a serial point), and publish the results.  Owners are listed by
// allocation rank for the commit, which runs after the last level's probe.  Every
// candidate was emitted by an earlier launch: a key that another block claimed in
// this launch is compared against that block's candidate, which is visible.  (On a
// multi-die part, publishing a block's keys to the other dies inside one launch
// would take a write-back of its L2.)  A wavefront with no candidate returns.
template <typename T, int TILE>
__device__ inline void probe_tiles(const Pool<T>& P, const Cands& K, const uint32_t* tile_cnt, uint64_t* res,
                                   const ChunkDev* chunks, const LevelInfo& L, uint32_t t0, uint32_t tn) {
    constexpr uint32_t kSlots = 2 * TILE;  // the block's key set of one pass (load <= 1/2)
    __shared__ uint32_t s_pre[65];         // s_pre[t]: candidates of the group's tiles before tile t
    __shared__ uint32_t s_own, s_base;
    __shared__ uint32_t s_set[kSlots];     // hash slot -> lane + 1 of the key's leader (0: free)
    __shared__ uint64_t s_hash[TILE];      // by lane: the pass's key hash
    __shared__ uint32_t s_cand[TILE];      // ... its candidate index
    __shared__ uint64_t s_res[TILE];       // ... a leader's result
    __shared__ int32_t s_kind[TILE];       // ... and kind
    // statistics of the block, flushed once by the last wavefront to end (a global atomic
    // per wavefront and pass would be waited for with every later load: stores and
    // atomics share the load counter on this part)
    __shared__ uint32_t s_stat[kProbeStats];
    __shared__ uint32_t s_done;
    for (uint32_t i = threadIdx.x; i < kSlots; i += TILE) s_set[i] = 0;
    if (threadIdx.x < kProbeStats) s_stat[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_done = 0;
    // the levels' candidate arrays from LDS: a key met as another level's pending entry
    // is compared against (and resolved through) that level's candidates without first
    // loading the level's record
    __shared__ Cands s_lv[kLevels + 1];
    if (threadIdx.x < kLevels + 1) s_lv[threadIdx.x] = P.levels[threadIdx.x];
    Pool<T> Q = P;
    Q.levels = s_lv;
    uint32_t n_probes = 0, n_claims = 0, n_lost = 0, n_tomb = 0, n_hit = 0, n_dup = 0, n_miss = 0;
    // the allocation window of the sequence (constant during this launch)
    const uint32_t ft0 = P.ctl->free_top0, next0 = P.ctl->next0, own_base = P.ctl->own_base;
    if (threadIdx.x < 64) {
        uint32_t x = threadIdx.x < tn ? tile_cnt[t0 + threadIdx.x] : 0u;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t y = __shfl_up(x, d);
            if ((int)threadIdx.x >= d) x += y;
        }
        s_pre[threadIdx.x + 1] = x;
        if (threadIdx.x == 0) {
            s_pre[0] = 0;
            s_own = 0;
        }
    }
    __syncthreads();
    const uint32_t N = s_pre[tn];
    const uint32_t wbase = threadIdx.x & ~63u;
    const int lane = (int)__lane_id();
    const uint32_t me = threadIdx.x;
    // A wavefront with no candidate in any pass ends here (whole wavefronts only: a
    // barrier waits for the wavefronts of the block that have not ended).
    if (wbase >= N) {
        probe_stats_flush(P, L.h, s_stat, &s_done, 0, 0, 0, 0, 0, 0, 0);
        return;
    }
    for (uint32_t j0 = 0; j0 < N; j0 += TILE) {  // uniform over the remaining wavefronts
        const bool wlive = j0 + wbase < N;  // the wavefront has candidates
        const uint32_t j = j0 + threadIdx.x;
        bool valid = j < N;
        uint64_t k = 0, slot = 0, result = kInvalid, tslot = ~0ull, h = 0;
        uint64_t key[9];
        int kind = -1;
        uint32_t leader = me, mine = kNoIdx, tree = kNoIdx;
        if (wlive) {
            const uint32_t jj = valid ? j : j0 + wbase;  // the wavefront's first candidate exists
            uint32_t lo = 0, hi = tn;  // s_pre[lo] <= jj < s_pre[hi]: jj lies in tile lo
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (s_pre[mid] <= jj) lo = mid;
                else hi = mid;
            }
            k = (uint64_t)(t0 + lo) * TILE + (jj - s_pre[lo]);
#pragma unroll
            for (int i = 0; i < 9; i++) key[i] = key_word(K, k, i);
            h = K.hash[k];
            slot = K.slot[k];
            // (a key of a tree that hit the leaf-brick panic is interned all the same:
            // keys merge across trees, so other trees may point at its candidate; the
            // GC releases the nodes nothing references afterwards)
            if (P.per_tree) {
                const GroupPos G = group_pos(L, slot);
                uint32_t c = G.chunk_base + (uint32_t)((slot - G.base) >> (3 * (G.depth - L.h)));
                tree = chunks[c].orig;
            }
            if (valid) {
                s_hash[me] = h;
                s_cand[me] = (uint32_t)k;
            }
        }
        __syncthreads();  // the pass's hashes and candidate indices are published
        // Equal keys of the pass share one leader: the first lane to take the key's free
        // slot of the block's LDS set (linear probing; a full compare against the slot
        // holder's candidate, which an earlier launch wrote).
        if (valid && P.dedup_rounds) {
            uint32_t s = (uint32_t)h & (kSlots - 1);
            for (uint32_t g = 0; g < kSlots; g++) {
                const uint32_t prev = atomicCAS(&s_set[s], 0u, me + 1);
                if (prev == 0) {
                    mine = s;
                    break;
                }
                const uint32_t o = prev - 1;
                if (s_hash[o] == h && key_matches_cand(K, s_cand[o], key)) {
                    leader = o;
                    break;
                }
                s = (s + 1) & (kSlots - 1);
            }
        }
        const bool lead = valid && leader == me;
        uint32_t probes = 0, claims = 0;
        if (lead)
            kind = P.single && tm_is_single((uint32_t)key[8])
                       ? probe_single(Q, L.h, k, key, &result, &tslot, &probes, &claims)
                       : probe_table(Q, L.h, k, key, h, &result, &tslot, &probes, &claims);
        // An owner's LOD value (core/voxel.rs:96-141) from its children's values, which
        // earlier launches wrote (a node gets its value at its own level): its loads are
        // issued now and complete while the block allocates.
        uint32_t lod = 0;
        if (kind == 2 && !P.dry) {
            uint32_t vals[8];
#pragma unroll
            for (int i = 0; i < 8; i++) vals[i] = (uint32_t)P.values[bid_index(key[i])];
            lod = calc_average(vals);
        }
        // a claimed tombstone is a tombstone less (a dry or aborted apply restores it)
        n_tomb += kind == 2 && (tslot & kTslotTomb) ? 1u : 0u;
        n_probes += probes;
        n_claims += (claims & 0xFFFFu) + (claims >> 16);
        n_lost += claims >> 16;
        // owners' allocation ranks: wavefront offsets in LDS, one counter update per block
        const uint64_t ob = __ballot(kind == 2);
        uint32_t woff = 0;
        if (ob && lane == 0) woff = atomicAdd(&s_own, (uint32_t)__popcll(ob));
        woff = __shfl(woff, 0);
        __syncthreads();
        if (threadIdx.x == 0) {
            // s_own is only read here; cleared for the next pass before the barrier
            // that pass's wavefront offsets come after
            const uint32_t own = s_own;
            if (own) s_base = atomicAdd(P.dry ? &P.ctl->dry_owners : &P.ctl->alloc_count, own);
            s_own = 0;
        }
        __syncthreads();  // s_base is next written after the next pass's first barrier
        if (kind == 2) {
            const uint32_t ka = s_base + woff + (uint32_t)__popcll(ob & ((1ull << lane) - 1ull));
            if (P.dry) {
                // a placeholder id (the table slot's previous value is restored after the run)
                result = make_branch(P.capacity + 65536u + ka, 0, (uint32_t)key[8]);
            } else {
                bool ok;
                uint32_t idx = alloc_at(P, ka, &ok, ft0, next0);
                result = ok ? make_branch(idx, tmg_gen(P.tmg[idx]), (uint32_t)key[8]) : kInvalid;
                if (ok) P.values[idx] = (T)lod;
                P.owners[ka - own_base] = ((uint64_t)L.h << 32) | k;  // the commit's list
            }
            if (P.per_tree && tree != kNoIdx) atomicAdd(&P.per_tree[tree], 1u);
        }
        if (lead) {
            s_res[me] = result;
            s_kind[me] = kind;
        }
        __syncthreads();  // the leaders' results are published
        if (valid && !lead) {
            result = s_res[leader];
            kind = s_kind[leader] == 0 ? 0 : 1;  // a follower of an owner or of a duplicate is a duplicate
        }
        if (valid) {
            K.id[k] = result;
            K.tslot[k] = tslot;
            res[slot] = result;
        }
        if (mine != kNoIdx) s_set[mine] = 0;  // every lane's probing ended before the barrier above
        n_hit += valid && kind == 0 ? 1u : 0u;
        n_dup += valid && kind == 1 ? 1u : 0u;
        n_miss += valid && kind == 2 ? 1u : 0u;
    }
    probe_stats_flush(P, L.h, s_stat, &s_done, n_probes, n_claims, n_lost, n_tomb, n_hit, n_dup, n_miss);
}

// The probe kernels hold a key, its hash and the scan state per lane: 5 waves per SIMD
// is the most that fits in registers without spilling to scratch.
template <typename T, int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_probe(Pool<T> P, Cands K, const uint32_t* tile_cnt,
                                                 uint64_t* res, const ChunkDev* chunks, LevelInfo L,
                                                 uint32_t ntiles, uint32_t group) {
    const uint32_t t0 = blockIdx.x * group;
    probe_tiles<T, TILE>(P, K, tile_cnt, res, chunks, L, t0, min(group, ntiles - t0));
}

// The same pass for the brick level, under its own name so that kernel traces tell the
// two apart.
template <typename T, int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_probe_bricks(
    Pool<T> P, Cands K, const uint32_t* tile_cnt, uint64_t* res, const ChunkDev* chunks, LevelInfo L,
    uint32_t ntiles, uint32_t group) {
    const uint32_t t0 = blockIdx.x * group;
    probe_tiles<T, TILE>(P, K, tile_cnt, res, chunks, L, t0, min(group, ntiles - t0));
}

// Owners write their node (children, types/mask; the probe wrote its LOD value),
// publish the table slot, and take one reference on each non-empty child.  The owners of an apply are
// its allocation ranks [own_base, alloc_count) after the last probe (leaves are
// allocated before own_base), listed by rank as (height << 32) | candidate.  No
// probe runs during the commit: until it, every new key is a pending table entry,
// matched against its candidate.  Grid-stride over the owners.
template <typename T, int TILE>
__global__ __launch_bounds__(TILE) void k_commit(Pool<T> P) {
    __shared__ uint32_t akey[kAggSlots], acnt[kAggSlots];
    __shared__ uint32_t s_owned;
    // the levels' candidate arrays, from LDS (read once per owner, ahead of its loads)
    __shared__ Cands s_lv[kLevels + 1];
    agg_init(akey, acnt);
    if (threadIdx.x == 0) s_owned = 0;
    if (threadIdx.x < kLevels + 1) s_lv[threadIdx.x] = P.levels[threadIdx.x];
    __syncthreads();
    uint32_t owned = 0;
    const uint32_t n = P.ctl->alloc_count - P.ctl->own_base;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t e = P.owners[i];
        const Cands& K = s_lv[e >> 32];
        const uint64_t k = (uint32_t)e;
        // the owner's record in one round of loads: the slot and id decide whether it is
        // written, the key and hash are what it writes
        const uint64_t ts = K.tslot[k];
        const uint64_t id = K.id[k];
        uint64_t key[9];
#pragma unroll
        for (int j = 0; j < 9; j++) key[j] = key_word(K, k, j);
        const uint64_t hk = K.hash[k];
        asm volatile("" ::"v"(hk));  // (loaded with the rest, not after the slot test)
        if (ts == ~0ull || id == kInvalid) continue;  // its allocation failed
        owned++;
        uint32_t idx = bid_index(id);
        uint64_t* dst = P.children + 8ull * idx;
#pragma unroll
        for (int j = 0; j < 8; j++) dst[j] = key[j];
        P.tmg[idx] = tmg_make((uint32_t)key[8], bid_gen(id));
        if (ts & kTslotSingle) P.single[tslot_index(ts)] = id;
        else P.tab[tslot_index(ts)] = slot_make(slot_tag(hk), false, 0, idx);
        for (int j = 0; j < 8; j++) {
            if (key[j] == kEmpty) continue;
            uint32_t ci = bid_index(key[j]);
            if (!agg_add(akey, acnt, ci, 1u)) atomicAdd(&P.refs[ci], 1u);
        }
    }
    // the block's owners go to the live count with one update (a counter every block
    // of the launch updates is a serial point)
    const uint32_t wo = wave_sum(owned);
    if (__lane_id() == 0 && wo) atomicAdd(&s_owned, wo);
    __syncthreads();
    for (int i = threadIdx.x; i < kAggSlots; i += blockDim.x)
        if (akey[i] != kNoIdx && acnt[i]) atomicAdd(&P.refs[akey[i]], acnt[i]);
    if (threadIdx.x == 0 && s_owned)
        atomicAdd((unsigned long long*)&P.ctl->live_branches, (unsigned long long)s_owned);
}

// The first allocation rank of the apply's probes (the owners' list base), set by a
// pass that runs after the leaf allocation and before the first probe.
template <typename T>
__device__ inline void mark_owner_base(const Pool<T>& P) {
    if (blockIdx.x == 0 && threadIdx.x == 0) P.ctl->own_base = P.ctl->alloc_count;
}

// A pending result names a candidate of the level; that candidate's id is final, or
// (a brick that shared an equal key's candidate) pending on the owner's, which is.
__device__ inline uint64_t resolve(uint64_t r, const uint64_t* cand_id_prev) {
    if (res_is_pending(r)) r = cand_id_prev[(uint32_t)r];
    if (res_is_pending(r)) r = cand_id_prev[(uint32_t)r];
    return r;
}

// One node at height h >= 2 (voxtree.rs:913-1106): gather the touched children
// from height h-1, clone untouched ones from the old node at this position (or
// fill them with the split leaf), collapse 8 equal leaves, else emit a key.
template <typename T, int TILE>
__device__ inline void level_build_body(const Pool<T>& P, const ChunkDev* chunks, const LevelInfo& L,
                                        const uint64_t* res_prev,
                                        const uint64_t* cand_id_prev, uint64_t* res, const Cands& K,
                                        uint32_t* tile_cnt, uint32_t blk, const uint64_t* old_h) {
    __shared__ uint32_t s_cnt;
    __shared__ uint32_t s_st[8];
    __shared__ uint64_t s_kh[TILE];  // the block's key set (keyset_insert)
    __shared__ uint32_t s_kc[TILE];
    if (threadIdx.x < 8) s_st[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_cnt = 0;
    s_kh[threadIdx.x] = 0;
    __syncthreads();
    {
        // Loads in three rounds without branches between them (lanes past the end
        // load the last slot): the chunk's fields, its level h-1 offset and the
        // recorded old node; the 8 results of height h-1 and the children of the
        // equivalent node; the owners of results still pending.
        uint64_t slot = (uint64_t)blk * TILE + threadIdx.x;
        const bool valid = slot < L.nslots;
        const uint64_t sl = valid ? slot : L.nslots - 1;
        const uint64_t s0 = min((uint64_t)blk * TILE, L.nslots - 1);
        const GroupPos G = group_of(L, block_group(L, s0, min((uint64_t)blk * TILE + TILE, L.nslots) - 1), sl);
        const uint64_t local = sl - G.base;
        const uint32_t D = G.depth;
        const uint32_t sh = 3 * (D - L.h);
        const uint32_t c = G.chunk_base + (uint32_t)(local >> sh);
        const uint32_t q = (uint32_t)(local & ((1ull << sh) - 1));
        const ChunkDev* Cp = chunks + c;
        const int32_t cstatus = Cp->status;
        const bool cpatch = Cp->has_patches != 0;
        const bool cfill = Cp->has_fill != 0;
        const uint64_t cinit = Cp->init;
        const uint64_t recorded = old_h[sl];
        // equivalent node (voxtree.rs:930-952): the descent pass's node at this slot;
        // a tree with a fill descends from its fill leaf
        uint32_t ell = D - L.h;
        uint64_t eq = cinit;
        if (ell > 0 && !cfill) {
            eq = recorded;
            ell = 0;
        }
        uint64_t ch[8], ec[8];
        {
            const uint64_t* src = res_prev + G.prev + (local << 3);  // the 8 child slots
            const bool eqb = eq != kEmpty && !bid_is_leaf(eq) && bid_index(eq) < P.capacity;
            const uint64_t* ecp = P.children + 8ull * (eqb ? bid_index(eq) : 0u);
#pragma unroll
            for (int i = 0; i < 8; i++) {
                ch[i] = src[i];
                ec[i] = ecp[i];
            }
        }
        // A completely written brick is rebuilt even when no value changed (its probe then
        // finds the old node); the reference does not push such a brick, so at height 2 a
        // branch result equal to the old child is untouched.  A leaf result (an all-same
        // brick) is pushed by the reference whether or not it changed (voxtree.rs:888-894),
        // and above height 2 an equal result only follows from such pushes: both stay
        // touched.
        const bool same_old =
            L.h == 2 && eq != kEmpty && !bid_is_leaf(eq) && bid_index(eq) < P.capacity && ell == 0;
        uint32_t given = 0;  // children with a result (kInvalid: untouched)
#pragma unroll
        for (int i = 0; i < 8; i++) given |= (ch[i] != kInvalid ? 1u : 0u) << i;
        // pending results resolve through the previous level's candidate ids (two hops at
        // most: a follower's leader may follow another), each hop one load per child for
        // all eight together (entry 0 for a child already resolved); taken child by child,
        // the compiler waited for every load before the next
#pragma unroll
        for (int hop = 0; hop < 2; hop++) {
            uint64_t v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = cand_id_prev[res_is_pending(ch[i]) ? (uint32_t)ch[i] : 0u];
#pragma unroll
            for (int i = 0; i < 8; i++) ch[i] = res_is_pending(ch[i]) ? v[i] : ch[i];
        }
        uint32_t touched = 0;
#pragma unroll
        for (int i = 0; i < 8; i++)
            touched |= (((given >> i) & 1u) && !(same_old && !bid_is_leaf(ch[i]) && ch[i] == ec[i]) ? 1u : 0u) << i;
        uint64_t out = kInvalid;
        bool cand = false, collapsed = false, h5 = false;
        uint32_t tm = 0;
        if (valid && cstatus >= 0 && cpatch) {
            uint32_t types = 0, mask = 0;
            bool failed = false;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                uint64_t r = ch[i];
                if (touched & (1u << i)) {
                    failed = failed || r == kInvalid;  // its owner could not allocate
                    mask |= 1u << i;
                    types |= (bid_is_leaf(r) ? 1u : 0u) << i;
                } else {
                    ch[i] = kEmpty;
                }
            }
            if (failed) mask = 0;  // the apply aborts; nothing is built above
            if (mask != 0) {
                uint64_t lq = kEmpty;
                bool walked = ell > 0;
                for (uint32_t cd = 0; cd < ell; cd++) {
                    if (bid_is_leaf(eq)) {
                        lq = eq;
                        eq = kEmpty;
                        break;
                    }
                    uint32_t digit = (q >> (3 * (ell - 1 - cd))) & 7u;
                    if (bid_index(eq) >= P.capacity) {
                        atomicOr(&P.ctl->err, kErrPoison);
                        eq = kEmpty;
                        break;
                    }
                    eq = P.children[8ull * bid_index(eq) + digit];
                    if (eq == kEmpty) break;
                }
                if (bid_is_leaf(eq)) {
                    lq = eq;
                    eq = kEmpty;
                }
                uint32_t cloned = bid_mask(eq) & ~mask & 0xFFu;
                if (mask != 0xFFu) {
                    if (cloned) {
                        for (int i = 0; i < 8; i++) {
                            if (!(cloned & (1u << i))) continue;
                            ch[i] = walked ? P.children[8ull * bid_index(eq) + i] : ec[i];
                            types |= (bid_is_leaf(ch[i]) ? 1u : 0u) << i;
                            mask |= 1u << i;
                        }
                    } else if (lq != kEmpty) {
                        uint32_t inv = ~mask & 0xFFu;
                        types |= inv;
                        mask |= inv;
                        for (int i = 0; i < 8; i++)
                            if (inv & (1u << i)) ch[i] = lq;
                    }
                }
                bool all_same = types == 0xFFu;
                for (int i = 1; i < 8; i++) all_same = all_same && ch[i] == ch[0];
                if (all_same) {
                    out = ch[0];
                    collapsed = true;
                    if (cloned) {
                        // dec_ref_by(children[0], min(pc(cloned),7)) instead of
                        // 7 - pc(cloned) (voxtree.rs:1066-1072, H5)
                        uint32_t pc = (uint32_t)__popc(cloned);
                        if (!P.dry) ref_delta_logged(P, c, bid_index(ch[0]), (int32_t)(7u - 2u * pc), ST_H5);
                        h5 = true;
                    }
                } else {
                    tm = (types << 8) | mask;
                    cand = true;
                    out = res_pending(0xFFFFFFFFu);
                }
            }
        }
        // equal keys of the block (any tree) share one candidate, as in the brick pass
        uint64_t key[9];
#pragma unroll
        for (int i = 0; i < 8; i++) key[i] = ch[i];
        key[8] = tm;
        const uint64_t hk = cand ? branch_hash(ch, tm) : 0ull;
        bool lead = false;
        const int ks = cand && P.brick_merge_rounds ? keyset_insert<TILE>(s_kh, hk, &lead) : -1;
        const uint32_t pos = emit_candidate<TILE>(K, &s_cnt, cand && (lead || ks < 0), ch, tm, slot, blk, hk);
        if (lead) s_kc[ks] = (uint32_t)((uint64_t)blk * TILE) + pos;
        __syncthreads();  // the leaders' candidates and their indices are visible to the block
        bool follow = false;
        if (cand && ks >= 0 && !lead) {
            const uint32_t k2 = s_kc[ks];
            if (key_matches_cand(K, k2, key)) {
                out = res_pending(k2);
                follow = true;
            } else {  // another key with the same 64-bit hash: a candidate of its own
                const uint32_t p2 = atomicAdd(&s_cnt, 1u);
                write_candidate(K, (uint64_t)blk * TILE + p2, ch, tm, hk, slot);
            }
        }
        if (valid) res[slot] = out;
        tile_count(s_st, 0, cand);
        tile_count(s_st, 1, collapsed);
        tile_count(s_st, 2, h5);
        tile_count(s_st, 3, follow);
    }
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blk] = s_cnt;
    const int ids[4] = {ST_CAND, ST_COLLAPSE, ST_H5, ST_MERGED};
    if (threadIdx.x < 4 && s_st[threadIdx.x])
        atomicAdd((unsigned long long*)stat(P.ctl, ids[threadIdx.x]), (unsigned long long)s_st[threadIdx.x]);
}

// Height h >= 2 (its probe is the next launch).  6 waves per SIMD: the most that fit
// without spilling.
template <typename T, int TILE>
__global__ __launch_bounds__(TILE) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_level_build(
    Pool<T> P, const ChunkDev* chunks, LevelInfo L,
    const uint64_t* res_prev, const uint64_t* cand_id_prev, uint64_t* res, Cands K,
    uint32_t* tile_cnt, const uint64_t* old_h) {
    level_build_body<T, TILE>(P, chunks, L, res_prev, cand_id_prev, res, K, tile_cnt, blockIdx.x, old_h);
}

// Tiles of every level of one apply, flattened (levels 1..dmax).
struct LevelTiles {
    const uint32_t* cnt[kLevels + 1];
    uint64_t start[kLevels + 2];  // first flattened tile of each level; start[dmax + 1] = total
    uint32_t dmax;
};

// After the last commit: a dry run restores the table slots its owners claimed;
// an apply that could not allocate (abort_) also takes back every reference its
// commits and logged corrections added, so the interner is left as before.
// Allocations themselves are undone by k_end_alloc.
template <typename T, int TILE>
__device__ inline void rollback_body(const Pool<T>& P, const ChunkDev* chunks, const LevelTiles& LT,
                                     const Drift* drift) {
    const bool abort = !P.dry && P.ctl->abort_ != 0;
    if (!P.dry && !abort) return;
    uint32_t owned = 0;
    uint64_t total = LT.start[LT.dmax + 1];
    for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
        uint32_t h = 1;
        while (h < LT.dmax && t >= LT.start[h + 1]) h++;
        uint64_t lt = t - LT.start[h];
        uint32_t n = LT.cnt[h][lt];
        if (threadIdx.x >= n) continue;
        const Cands& K = P.levels[h];
        uint64_t k = lt * TILE + threadIdx.x;
        uint64_t ts = K.tslot[k];
        if (ts == ~0ull) continue;
        if (ts & kTslotSingle) P.single[tslot_index(ts)] = 0;
        else P.tab[tslot_index(ts)] = (ts & kTslotTomb) ? kSlotTomb : kSlotEmpty;
        wave_count(stat(P.ctl, ST_TOMB), (ts & kTslotTomb) != 0);
        uint64_t id = K.id[k];
        if (P.dry || id == kInvalid) continue;
        owned++;
        for (int i = 0; i < 8; i++) {
            uint64_t c = key_word(K, k, i);
            if (c != kEmpty) atomicSub(&P.refs[bid_index(c)], 1u);
        }
    }
    if (!abort) return;
    uint32_t nu = P.ctl->n_undo < P.undo_cap ? P.ctl->n_undo : P.undo_cap;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nu; j += gridDim.x * blockDim.x)
        atomicSub(&P.refs[P.undo[j].idx], (uint32_t)P.undo[j].delta);
    // queued corrections applied by k_brick_retry (trees whose status was final)
    uint32_t ndr = P.ctl->n_drift;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < ndr; j += gridDim.x * blockDim.x) {
        Drift d = drift[j];
        if (chunks[d.chunk].status >= 0) atomicSub(&P.refs[d.idx], (uint32_t)d.delta);
    }
    uint32_t tot = wave_sum(owned);
    if (__lane_id() == 0 && tot)
        atomicAdd((unsigned long long*)&P.ctl->live_branches, (unsigned long long)(-(long long)tot));
}

// ------------------------------------------------------------------ finalize + GC

// voxtree.rs:303-328: adopt the new root, take the tree's reference on it, and
// queue the old root for dec_ref_recursive.
template <typename T>
__device__ inline void finalize_body(const Pool<T>& P, const ChunkDev* chunks, uint32_t n,
                                     const uint64_t* slot_off_all, uint32_t n_stride,
                                     uint64_t* const* res_levels, uint64_t* const* cand_id_levels,
                                     uint64_t* roots, int32_t* status, uint32_t* gc_front) {
    const uint32_t ab = P.ctl->abort_;
    VX_LOOP(c, n) {
        bool push = false;
        uint64_t old = kEmpty;
        if (c < n) {
            const ChunkDev& C = chunks[c];
            int32_t st = C.status;
            if (C.skip) {
                st = 0;
            } else if (st >= 0 && ab) {
                st = (ab & 2u) ? -8 : -2;  // VX_E_BUDGET / VX_E_OOM: the apply was rolled back
            } else if (st >= 0) {
                uint64_t nw;
                if (!C.has_patches) {
                    nw = C.init;
                } else {
                    uint32_t D = C.depth;
                    uint64_t r = res_levels[D][slot_off_all[(uint64_t)D * n_stride + c]];
                    nw = r != kInvalid ? resolve(r, cand_id_levels[D]) : r;
                }
                old = C.root_in;
                if (nw == kInvalid) {
                    st = 0;
                } else if (old != kEmpty && nw == old) {
                    st = -6;  // assert_ne!(new_root_id, self.root_id) panics
                } else if (P.dry) {
                    st = 1;
                } else {
                    if (nw != kEmpty) atomicAdd(&P.refs[bid_index(nw)], 1u);
                    push = old != kEmpty;
                    roots[C.orig] = nw;
                    st = 1;
                }
            }
            status[C.orig] = st;  // the record keeps its status (the rollback may read it)
        }
        uint32_t k = wave_append(&P.ctl->gc_count[0], push);
        if (push) gc_front[k] = bid_index(old);
    }
}

// After the last probe: the rollback (dry runs: restore the table; aborted applies:
// take references back) and the root swap, in one launch.  The swap does not touch
// the table or reference counts of an aborted apply, and leaves the per-tree records
// (read by the rollback) unchanged.
template <typename T, int TILE>
__global__ __launch_bounds__(TILE) void k_finish(Pool<T> P, const ChunkDev* chunks, LevelTiles LT,
                                                  const Drift* drift, uint32_t n, const uint64_t* slot_off_all,
                                                  uint32_t n_stride, uint64_t* const* res_levels,
                                                  uint64_t* const* cand_id_levels, uint64_t* roots,
                                                  int32_t* status, uint32_t* gc_front) {
    rollback_body<T, TILE>(P, chunks, LT, drift);
    finalize_body<T>(P, chunks, n, slot_off_all, n_stride, res_levels, cand_id_levels, roots, status, gc_front);
}

// Leaves created in this sequence that nothing references (their trees hit an
// error) are released again so the interner is left as before for those trees.
template <typename T>
__device__ inline void leaf_cleanup_body(const Pool<T>& P) {
    uint32_t n = P.ctl->n_new_leaves;
    VX_LOOP(i, n) {
        bool freed = false;
        if (i < n) {
            uint32_t idx = P.new_leaves[i];
            if (P.refs[idx] == 0) {
                freed = true;
                leaf_remove(P, (uint32_t)P.values[idx]);
                P.values[idx] = 0;
                const uint32_t g = tmg_gen(P.tmg[idx]) + 1u;
                P.tmg[idx] = tmg_make(0u, g >= kMaxGeneration ? 0u : g);
                uint32_t pos = atomicAdd(&P.ctl->free_top, 1u);
                if (pos < P.capacity) P.free_stack[pos] = idx;
                else atomicOr(&P.ctl->err, kErrPoison);
            }
        }
        uint64_t b = __ballot(freed);
        if (b && __lane_id() == (uint32_t)(__ffsll((unsigned long long)b) - 1))
            atomicAdd((unsigned long long*)&P.ctl->live_leaves, (unsigned long long)(-(long long)__popcll(b)));
    }
}

// One round of dec_ref_recursive over all queued nodes (interner/mod.rs:449-527):
// decrement; a node that reaches zero drops its pattern, is recycled
// (mod.rs:566-625), and decrements its children, queueing those that reach zero.
// Round 0 holds the replaced roots (decremented here); later rounds hold nodes
// whose count already reached zero.  Each pass of a block stages its pushes (freed
// indices onto the free stack, children that reached zero onto the next queue) in
// LDS and reserves room for them with one update of each shared counter: those
// counters are serial points that every block of the launch updates.
//
// A small front (at most kGcLocalShare nodes per block of the launch) is split into
// contiguous shares of at least kGcShareMin nodes, and each block follows its share's
// subtrees itself: the children that reach zero in one pass are the block's next pass,
// in the same launch, while they fit one pass (kGcLocal); more go to the next launch.
// Only the block whose decrement took a count to zero holds that node, so no other
// block has to see it, and a chain of small rounds (a few thousand replaced paths) costs
// one launch instead of one per height.  A launch frees nodes of any depth at or below
// its round's, so the D + 1 launches of a tree of depth D still reach every node.
constexpr uint32_t kGcStage = 2048;      // children staged per pass (8 per lane of a 256-lane block)
constexpr uint32_t kGcLocal = 256;       // the most nodes of one pass a block follows itself
constexpr uint32_t kGcLocalShare = 64;   // a front up to this many nodes per block is followed
constexpr uint32_t kGcShareMin = 16;     // ... in shares of at least this many nodes
constexpr uint32_t kGcFreeStage = 768;   // freed nodes staged before one free-stack update

// Frees node idx, whose count is zero: drops its pattern, marks the slot free and
// releases its children.  Children decremented here that reach zero are staged in
// s_next; children combined in the block's table (akey/acnt) are decremented when the
// table is flushed.
template <typename T>
__device__ inline void gc_free(const Pool<T>& P, uint32_t idx, uint32_t* akey, uint32_t* acnt, uint32_t* s_next,
                               uint32_t* s_nnext, bool* is_leaf, bool* tomb) {
    // the node's columns in one round of loads (a leaf's children are empty).  A free
    // writes only what a later reader needs: the generation, the types/mask word (0
    // marks a free or leaf slot for the table rebuild), the pattern's table slot and the
    // free stack.  The count is already 0, the value and the children are rewritten
    // when the slot is allocated again.
    const uint32_t w = P.tmg[idx];
    const uint32_t tm = tmg_tm(w);
    const uint32_t g = tmg_gen(w) + 1u;
    uint64_t ch[8];
    const uint64_t* src = P.children + 8ull * idx;
#pragma unroll
    for (int j = 0; j < 8; j++) ch[j] = src[j];
    // Both loads complete before the leaf test: without this use of the children the
    // compiler moves their load into the branch path, after the wait for the types/mask
    // word (two round trips instead of one).
    asm volatile("" ::"v"(w), "v"(ch[0]), "v"(ch[1]), "v"(ch[2]), "v"(ch[3]), "v"(ch[4]), "v"(ch[5]), "v"(ch[6]),
                 "v"(ch[7]));
    *is_leaf = tm == 0;
    if (*is_leaf) {
        leaf_remove(P, (uint32_t)P.values[idx]);
    } else {
        // children first: combined in LDS where the block's table has room, the rest
        // decremented with independent atomics (issued under the lane mask, so they are
        // in flight together while the pattern is removed below)
        uint32_t gmask = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (ch[j] != kEmpty && !agg_add(akey, acnt, bid_index(ch[j]), 1u)) gmask |= 1u << j;
        uint32_t old[8];
#pragma unroll
        for (int j = 0; j < 8; j++) old[j] = ((gmask >> j) & 1u) ? atomicSub(&P.refs[bid_index(ch[j])], 1u) : 0u;
        // remove the pattern (mod.rs:518): a one-leaf key from its index (the leaf is
        // still live: this node's reference is still being dropped), any other from the
        // table, leaving a tombstone
        if (P.single && tm_is_single(tm)) {
            P.single[single_index(P, tm, ch)] = 0;
        } else {
            uint64_t h = branch_hash(ch, tm);
            uint64_t want = slot_make(slot_tag(h), false, 0, idx);
            uint64_t s = h & P.tmask;
            for (uint64_t guard = 0; guard <= P.tmask; guard++) {
                uint64_t e = P.tab[s];
                if (e == want) {
                    P.tab[s] = kSlotTomb;
                    *tomb = true;
                    break;
                }
                if (e == kSlotEmpty) break;
                s = (s + 1) & P.tmask;
            }
        }
        // A child reached zero when its old count was 1.  The 1 is hidden from the
        // compiler here: a compare it can place next to each atomic makes it wait for
        // each result before issuing the next (eight round trips in a row); compared
        // after the pattern's removal, the eight atomics are in flight together.
        uint32_t one = 1u;
        asm volatile("" : "+v"(one));
#pragma unroll
        for (int j = 0; j < 8; j++)
            if (((gmask >> j) & 1u) && old[j] == one) s_next[atomicAdd(s_nnext, 1u)] = bid_index(ch[j]);
    }
    P.tmg[idx] = tmg_make(0u, g >= kMaxGeneration ? 0u : g);  // no types/mask: free
}

// Applies the block's combined decrements (a 256-lane block: every lane takes
// kAggSlots / 256 slots, whose atomics are in flight together); slots that reach zero
// are staged in s_next.  reset: clear the slots for further use.
template <typename T>
__device__ inline void gc_agg_flush(const Pool<T>& P, uint32_t* akey, uint32_t* acnt, uint32_t* s_next,
                                    uint32_t* s_nnext, bool reset) {
    constexpr int kPer = kAggSlots / 256;
    uint32_t key[kPer], cnt[kPer], old[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const uint32_t k = threadIdx.x + 256u * u;
        key[u] = akey[k];
        cnt[u] = acnt[k];
        old[u] = key[u] != kNoIdx && cnt[u] ? atomicSub(&P.refs[key[u]], cnt[u]) : 0u;
        if (reset && key[u] != kNoIdx) {
            akey[k] = kNoIdx;
            acnt[k] = 0;
        }
    }
    // (the counts are hidden from the compiler so that it compares after every atomic
    // was issued, not next to each one)
    asm volatile("" : "+v"(cnt[0]));
#pragma unroll
    for (int u = 0; u < kPer; u++)
        if (key[u] != kNoIdx && cnt[u] && old[u] == cnt[u]) s_next[atomicAdd(s_nnext, 1u)] = key[u];
}

template <typename T>
__global__ __launch_bounds__(256) void k_gc_round(Pool<T> P, uint32_t round, const uint32_t* front, uint32_t* next,
                                                  bool decrement_self, uint64_t front_cap, bool leaf_cleanup) {
    __shared__ uint32_t akey[kAggSlots], acnt[kAggSlots];
    __shared__ uint32_t s_next[kGcStage];
    __shared__ uint32_t s_cur[kGcLocal];
    __shared__ uint32_t s_free[kGcFreeStage];
    __shared__ uint32_t s_nnext, s_nfree, s_bnext, s_bfree;
    __shared__ unsigned long long s_leaves, s_branches, s_tombs, s_recycled;
    uint64_t n = P.ctl->gc_count[round];
    if (n > front_cap) n = front_cap;
    // a small front: the block's share [lo, hi), followed locally
    const bool local = n <= (uint64_t)gridDim.x * kGcLocalShare;
    uint64_t lo = 0, hi = 0;
    if (local) {
        const uint64_t share = max((uint64_t)kGcShareMin, (n + gridDim.x - 1) / gridDim.x);
        lo = min(n, (uint64_t)blockIdx.x * share);
        hi = min(n, lo + share);
    }
    // a block past the frontier has nothing to do (the rounds near the leaves hold few
    // nodes); the first round also releases new leaves and orphans with every block
    if (!leaf_cleanup && (local ? lo >= hi : (uint64_t)blockIdx.x * blockDim.x >= n)) return;
    agg_init(akey, acnt);
    if (threadIdx.x == 0) {
        s_nnext = s_nfree = 0;
        s_leaves = s_branches = s_tombs = s_recycled = 0;
    }
    __syncthreads();
    // new leaves nothing references (their trees failed) are released in the first
    // round; no node that this GC visits references them
    if (leaf_cleanup) leaf_cleanup_body(P);
    uint32_t* next_count = &P.ctl->gc_count[round + 1];
    // New branch nodes that nothing references go to the next round: keys of a tree that
    // failed in the brick pass (the leaf-brick panic) are interned like every key, and
    // the ones no other tree shares end up without a reference.  Every other new node
    // holds one from its parent or its tree (the commit and the root swap ran).
    if (leaf_cleanup && P.ctl->h1_fail) {
        const uint32_t no = P.ctl->n_owners;
        VX_LOOP(i, no) {
            uint32_t idx = kNoIdx;
            if (i < no) {
                const uint64_t e = P.owners[i];
                const Cands& K = P.levels[e >> 32];
                const uint64_t id = K.id[(uint32_t)e];
                if (K.tslot[(uint32_t)e] != ~0ull && id != kInvalid && P.refs[bid_index(id)] == 0u)
                    idx = bid_index(id);
            }
            const uint32_t pos = wave_append(next_count, idx != kNoIdx);
            if (idx != kNoIdx) {
                if (pos < front_cap) next[pos] = idx;
                else atomicOr(&P.ctl->err, kErrPoison);
            }
        }
    }
    const int lane = (int)__lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    if (local) {
        // passes over the block's share, then over the children that reached zero in the
        // block, while they fit one pass (the loop is uniform over the block)
        uint32_t cnt = (uint32_t)(hi - lo);
        bool first = true;
        for (;;) {
            uint32_t idx = kNoIdx;
            if (threadIdx.x < cnt) {
                const uint32_t v = first ? front[lo + threadIdx.x] : s_cur[threadIdx.x];
                if (v != 0 && v < P.capacity) idx = v;
            }
            const bool freed = first && decrement_self ? wave_ref_dec_zero(P.refs, idx) : idx != kNoIdx;
            bool is_leaf = false, tomb = false;
            if (freed) gc_free(P, idx, akey, acnt, s_next, &s_nnext, &is_leaf, &tomb);
            const uint64_t fb = __ballot(freed);
            const uint64_t bl = __ballot(freed && is_leaf), bb = __ballot(freed && !is_leaf), bt = __ballot(tomb);
            uint32_t woff = 0;
            if (lane == 0) {
                if (fb) woff = atomicAdd(&s_nfree, (uint32_t)__popcll(fb));
                if (bl) atomicAdd(&s_leaves, (unsigned long long)__popcll(bl));
                if (bb) atomicAdd(&s_branches, (unsigned long long)__popcll(bb));
                if (bt) atomicAdd(&s_tombs, (unsigned long long)__popcll(bt));
                if (fb) atomicAdd(&s_recycled, (unsigned long long)__popcll(fb));
            }
            woff = __shfl(woff, 0);
            if (freed) s_free[woff + (uint32_t)__popcll(fb & below)] = idx;
            __syncthreads();
            // the pass's combined decrements (children shared by several of its nodes):
            // those that reach zero join the children of the next pass
            gc_agg_flush(P, akey, acnt, s_next, &s_nnext, true);
            __syncthreads();
            const uint32_t nn = s_nnext, nf = s_nfree;
            const bool more = nn != 0 && nn <= kGcLocal;
            // the staged frees go to the free stack when the next pass might not fit
            if (nf && (!more || nf + kGcLocal > kGcFreeStage)) {
                if (threadIdx.x == 0) s_bfree = atomicAdd(&P.ctl->free_top, nf);
                __syncthreads();
                for (uint32_t k = threadIdx.x; k < nf; k += blockDim.x) {
                    const uint32_t pos = s_bfree + k;
                    if (pos < P.capacity) P.free_stack[pos] = s_free[k];
                    else atomicOr(&P.ctl->err, kErrPoison);
                }
                __syncthreads();
                if (threadIdx.x == 0) s_nfree = 0;
            }
            if (!more) {
                // more children than one pass: the next launch takes them
                if (nn) {
                    if (threadIdx.x == 0) s_bnext = atomicAdd(next_count, nn);
                    __syncthreads();
                    for (uint32_t k = threadIdx.x; k < nn; k += blockDim.x) {
                        const uint32_t pos = s_bnext + k;
                        if (pos < front_cap) next[pos] = s_next[k];
                        else atomicOr(&P.ctl->err, kErrPoison);
                    }
                }
                __syncthreads();
                if (threadIdx.x == 0) s_nnext = 0;
                break;
            }
            for (uint32_t k = threadIdx.x; k < nn; k += blockDim.x) s_cur[k] = s_next[k];
            __syncthreads();
            if (threadIdx.x == 0) s_nnext = 0;
            __syncthreads();
            cnt = nn;
            first = false;
        }
    } else {
        for (uint64_t b0 = (uint64_t)blockIdx.x * blockDim.x; b0 < n; b0 += (uint64_t)gridDim.x * blockDim.x) {
            const uint64_t i = b0 + threadIdx.x;  // the loop is uniform over the block
            uint32_t idx = kNoIdx;
            if (i < n) {
                uint32_t v = front[i];
                if (v != 0 && v < P.capacity) idx = v;
            }
            bool freed = decrement_self ? wave_ref_dec_zero(P.refs, idx) : idx != kNoIdx;
            bool is_leaf = false, tomb = false;
            if (freed) gc_free(P, idx, akey, acnt, s_next, &s_nnext, &is_leaf, &tomb);
            const uint64_t fb = __ballot(freed);
            const uint64_t bl = __ballot(freed && is_leaf), bb = __ballot(freed && !is_leaf), bt = __ballot(tomb);
            uint32_t woff = 0;
            if (lane == 0) {
                if (fb) woff = atomicAdd(&s_nfree, (uint32_t)__popcll(fb));
                if (bl) atomicAdd(&s_leaves, (unsigned long long)__popcll(bl));
                if (bb) atomicAdd(&s_branches, (unsigned long long)__popcll(bb));
                if (bt) atomicAdd(&s_tombs, (unsigned long long)__popcll(bt));
                if (fb) atomicAdd(&s_recycled, (unsigned long long)__popcll(fb));
            }
            woff = __shfl(woff, 0);
            __syncthreads();
            if (threadIdx.x == 0) {
                if (s_nfree) s_bfree = atomicAdd(&P.ctl->free_top, s_nfree);
                if (s_nnext) s_bnext = atomicAdd(next_count, s_nnext);
            }
            __syncthreads();
            if (freed) {
                uint32_t pos = s_bfree + woff + (uint32_t)__popcll(fb & below);
                if (pos < P.capacity) P.free_stack[pos] = idx;
                else atomicOr(&P.ctl->err, kErrPoison);
            }
            for (uint32_t k = threadIdx.x; k < s_nnext; k += blockDim.x) {
                uint32_t pos = s_bnext + k;
                if (pos < front_cap) next[pos] = s_next[k];
                else atomicOr(&P.ctl->err, kErrPoison);
            }
            __syncthreads();
            if (threadIdx.x == 0) s_nfree = s_nnext = 0;
            __syncthreads();
        }
    }
    // the block's combined decrements; children reaching zero are queued the same way
    __syncthreads();
    gc_agg_flush(P, akey, acnt, s_next, &s_nnext, false);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (s_nnext) s_bnext = atomicAdd(next_count, s_nnext);
        if (s_leaves) atomicAdd((unsigned long long*)&P.ctl->live_leaves, (unsigned long long)(-(long long)s_leaves));
        if (s_branches) {
            atomicAdd((unsigned long long*)&P.ctl->live_branches, (unsigned long long)(-(long long)s_branches));
            if (s_tombs) atomicAdd((unsigned long long*)stat(P.ctl, ST_TOMB), (unsigned long long)s_tombs);
        }
        if (s_recycled) {
            atomicAdd((unsigned long long*)stat(P.ctl, ST_RECYCLED), s_recycled);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < s_nnext; k += blockDim.x) {
        uint32_t pos = s_bnext + k;
        if (pos < front_cap) next[pos] = s_next[k];
        else atomicOr(&P.ctl->err, kErrPoison);
    }
}

// ------------------------------------------------------------------ table rebuild

// Frees leave tombstones in the pattern table.  A later key reuses one only when its
// scan meets it before an empty slot, so between rebuilds the non-empty share of the
// table grows and scans lengthen.  A rebuild clears the table and re-inserts every live
// branch (a node with a types/mask word: the GC clears it on free, leaves have none).
//
// Every write call starts with the decision, made on the device in stream order
// (so it stays current however many calls the host has queued): rebuild when the
// tombstones take a sixteenth of the slots and live keys plus tombstones half of them
// in the pattern table, or three quarters in the i32 leaf table (force: always).  Half
// keeps a miss's scan to the next empty slot at about two probes; the rebuild (a clear
// of the table and one insert per live branch) costs less than the longer scans of a
// fuller table (measured on the 4096-tree uneven apply: 27.1 against 28.3 ms per step).
// Every block of the clear kernel takes the decision (block 0 records it for the kernels
// that follow); the clear and re-insert kernels return at once when no rebuild is due.
__global__ void k_tab_clear(uint64_t* tab, uint64_t nslots, Ctl* ctl, uint64_t lslots, uint32_t force) {
    static_assert(kStatStripes == 64, "one statistics stripe per lane");
    __shared__ uint32_t s_due;
    if (threadIdx.x < 64) {  // every block takes the same decision from the same counters
        const int lane = (int)threadIdx.x;
        uint64_t t = ctl->st[lane][ST_TOMB], lt = ctl->st[lane][ST_LTOMB];
        for (int off = 32; off > 0; off >>= 1) {  // signed adds: the sums wrap to the count
            t += shfl64(t, lane ^ off);
            lt += shfl64(lt, lane ^ off);
        }
        if (lane == 0) {
            const uint64_t live = ctl->live_branches, ll = ctl->live_leaves;
            bool due = force != 0 || (t >= nslots / 16 && live + t > nslots / 2);
            if (lslots) due = due || (lt >= lslots / 16 && ll + lt > lslots / 4 * 3);
            s_due = due ? (lslots ? 3u : 1u) : 0u;
            if (blockIdx.x == 0) {  // for the re-insert and the leaf-table kernels
                ctl->rebuild = s_due;
                if (due) ctl->rebuilds++;
            }
        }
    }
    __syncthreads();
    if (!(s_due & 1u)) return;
    uint4* t4 = (uint4*)tab;
    const uint64_t n4 = nslots / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x)
        t4[i] = make_uint4(0u, 0u, 0u, 0u);
}

template <typename T>
__global__ void k_tab_reinsert(Pool<T> P) {
    if (!(P.ctl->rebuild & 1u)) return;
    // the cleared table holds no tombstone (no other block reads the counter here)
    if (blockIdx.x == 0 && threadIdx.x < kStatStripes) P.ctl->st[threadIdx.x][ST_TOMB] = 0;
    const uint32_t top = P.ctl->next_index;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < top; i += gridDim.x * blockDim.x) {
        const uint32_t tm = tmg_tm(P.tmg[i]);
        if (i == 0 || tm == 0) continue;
        uint64_t ch[8];
        const uint64_t* src = P.children + 8ull * i;
#pragma unroll
        for (int j = 0; j < 8; j++) ch[j] = src[j];
        if (P.single && tm_is_single(tm)) continue;  // in the one-leaf index, which has no tombstones
        const uint64_t h = branch_hash(ch, tm);
        const uint64_t want = slot_make(slot_tag(h), false, 0, i);
        uint64_t s = h & P.tmask;
        for (uint64_t guard = 0; guard <= P.tmask; guard++) {
            if (atomicCAS((unsigned long long*)&P.tab[s], 0ull, (unsigned long long)want) == 0ull) break;
            s = (s + 1) & P.tmask;
        }
    }
}

// The hashed leaf table of four-byte voxels, rebuilt the same way: every live leaf (a
// node without a types/mask word that holds a reference) is inserted again.
__global__ void k_ltab_clear(uint64_t* ltab, uint64_t nslots, Ctl* ctl) {
    if (!(ctl->rebuild & 2u)) return;
    uint4* t4 = (uint4*)ltab;
    const uint64_t n4 = nslots / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x)
        t4[i] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && threadIdx.x < kStatStripes) ctl->st[threadIdx.x][ST_LTOMB] = 0;
}

template <typename T>
__global__ void k_ltab_reinsert(Pool<T> P) {
    if (!(P.ctl->rebuild & 2u)) return;
    const uint32_t top = P.ctl->next_index;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < top; i += gridDim.x * blockDim.x)
        if (i != 0 && tmg_tm(P.tmg[i]) == 0 && P.refs[i] != 0) leaf_publish(P, (uint32_t)P.values[i], i);
}

// ------------------------------------------------------------------ single-voxel set

// VoxOpsWrite::set (voxtree.rs:162-260) on many trees at once, one voxel per tree:
// set_at_depth_iterative (voxtree.rs:393-561) for a non-default value,
// remove_at_depth (voxtree.rs:563-705) for the default one.  The descent records
// the path exactly as those functions see it; the new nodes are then built one
// height per launch (k_set_level) and interned by the shared probe and commit, so
// trees of one call share nodes like any hash-consed apply.  The reference-count
// result equals the reference's: its increments and decrements along one path
// cancel except for +1 per non-empty child of each new node and +1 on the new root
// (no count passes through zero inside a set: a hit never contains a node created
// lower in the same path).

__device__ inline uint32_t child_slot(int32_t x, int32_t y, int32_t z, uint32_t d, uint32_t D) {
    const uint32_t sh = D - d - 1;
    return (((uint32_t)x >> sh) & 1u) | ((((uint32_t)y >> sh) & 1u) << 1) | ((((uint32_t)z >> sh) & 1u) << 2);
}

// Per tree: rec[c][0][d] = the branch at depth d on the path (empty when there is
// none), rec[c][1][d] = the leaf being split above depth d; meta[c] = depth at which
// a remove meets a leaf (the depth itself when the voxel's own leaf is removed).
constexpr uint32_t kSetRec = 2 * kLevels;

template <typename T>
__global__ void k_set_descend(Pool<T> P, ChunkDev* chunks, uint32_t n, const uint64_t* roots, const int32_t* xyz,
                              const T* vals, uint64_t* rec, uint32_t* meta) {
    // the sequence's control-block reset rides in block 0 (nothing else here reads it)
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl_begin(P.ctl);
    VX_LOOP(c, n) {
        if (c < n) {
            ChunkDev& C = chunks[c];
            C.root_in = roots[C.orig];
            C.init = C.root_in;
            C.status = C.status0;
            if (C.status < 0) continue;
            const uint32_t D = C.depth;
            const int32_t x = xyz[3ull * C.orig], y = xyz[3ull * C.orig + 1], z = xyz[3ull * C.orig + 2];
            const uint32_t v = (uint32_t)vals[C.orig];
            const int32_t lim = 1 << D;
            // assert!(position in 0..2^D) (voxtree.rs:164-166)
            if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) {
                C.status = -4;
                continue;
            }
            uint64_t* rn = rec + (uint64_t)c * kSetRec;
            uint64_t* rl = rn + kLevels;
            uint64_t cur = C.root_in;
            bool unchanged = cur == kEmpty && v == 0;  // an empty tree and the default value
            if (!unchanged && v != 0) {
                // voxtree.rs:411-470: branches (empty included) are followed, a leaf is
                // split from there down; the same value already present changes nothing
                uint64_t leaf = kEmpty;
                for (uint32_t d = 0; d < D && !unchanged; d++) {
                    if (bid_index(cur) >= P.capacity) {  // a corrupt id: poisoned, read as empty
                        atomicOr(&P.ctl->err, kErrPoison);
                        cur = kEmpty;
                    }
                    if (!bid_is_leaf(cur)) {
                        rn[d] = cur;
                        rl[d] = leaf;
                        if (cur != kEmpty) cur = P.children[8ull * bid_index(cur) + child_slot(x, y, z, d, D)];
                    } else {
                        unchanged = (uint32_t)P.values[bid_index(cur)] == v;
                        leaf = cur;
                        cur = kEmpty;
                        rn[d] = kEmpty;
                        rl[d] = leaf;
                    }
                }
                if (bid_index(cur) >= P.capacity) {
                    atomicOr(&P.ctl->err, kErrPoison);
                    cur = kEmpty;
                }
                if (!unchanged && bid_is_leaf(cur) && (uint32_t)P.values[bid_index(cur)] == v) unchanged = true;
                if (!unchanged && leaf_index(P, v) == 0) leaf_request(P, v);
            } else if (!unchanged) {
                // voxtree.rs:581-660: present children are followed down to a leaf (split
                // from there, or the voxel itself at depth D); a missing child changes nothing
                uint32_t dl = D;
                for (uint32_t d = 0; d < D; d++) {
                    if (bid_is_leaf(cur)) {
                        dl = d;
                        break;
                    }
                    const uint32_t ci = child_slot(x, y, z, d, D);
                    if (!((bid_mask(cur) >> ci) & 1u)) {
                        unchanged = true;
                        break;
                    }
                    if (bid_index(cur) >= P.capacity) {
                        atomicOr(&P.ctl->err, kErrPoison);
                        unchanged = true;
                        break;
                    }
                    rn[d] = cur;
                    cur = P.children[8ull * bid_index(cur) + ci];
                }
                for (uint32_t d = dl; d < D; d++) rl[d] = cur;  // the leaf being split
                meta[c] = dl;
            }
            if (unchanged) C.skip = 1;
        }
    }
}

// One height of the set / remove paths (one slot per tree): the node at depth
// D - h from the result one height below, as voxtree.rs:476-555 (set) and
// voxtree.rs:597-705 (remove) build it.  Blocks nbuild.. commit height h-1.
template <typename T, int TILE>
__global__ __launch_bounds__(TILE) void k_set_level(
    Pool<T> P, const ChunkDev* chunks, LevelInfo L, const uint64_t* doff, uint32_t n, const int32_t* xyz,
    const T* vals, const uint64_t* rec, const uint32_t* meta, const uint64_t* res_prev,
    const uint64_t* cand_id_prev, uint64_t* res, Cands K, uint32_t* tile_cnt) {
    if (L.h == 1) mark_owner_base(P);
    __shared__ uint32_t s_cnt;
    __shared__ uint32_t s_st[8];
    if (threadIdx.x < 8) s_st[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint64_t slot = (uint64_t)blockIdx.x * TILE + threadIdx.x;
    const bool valid = slot < L.nslots;
    uint64_t out = kInvalid;
    bool cand = false, collapsed = false;
    uint64_t ch[8];
    uint32_t tm = 0;
    if (valid) {
        const GroupPos G = group_pos(L, slot);
        const uint32_t c = G.chunk_base + (uint32_t)(slot - G.base);  // one slot per tree
        const ChunkDev& C = chunks[c];
        if (C.status >= 0 && !C.skip) {
            const uint32_t D = C.depth;
            const uint32_t d = D - L.h;
            const int32_t x = xyz[3ull * C.orig], y = xyz[3ull * C.orig + 1], z = xyz[3ull * C.orig + 2];
            const uint32_t v = (uint32_t)vals[C.orig];
            const uint32_t ci = child_slot(x, y, z, d, D), bit = 1u << ci;
            const uint64_t* rn = rec + (uint64_t)c * kSetRec;
            const uint64_t* rl = rn + kLevels;
            // the result one height below: the voxel's leaf (or empty) at the bottom
            uint64_t cur;
            if (L.h == 1) {
                if (v != 0) {
                    // creation failed: the apply aborts; a dry run counts a placeholder
                    uint32_t li = leaf_index(P, v);
                    cur = li ? make_leaf(li, tmg_gen(P.tmg[li])) : (P.dry ? make_leaf(P.capacity + v, 0) : kInvalid);
                } else {
                    cur = kEmpty;
                }
            } else {
                uint64_t r = res_prev[doff[(uint64_t)(L.h - 1) * n + c]];
                cur = res_is_pending(r) ? cand_id_prev[(uint32_t)r] : r;
            }
            if (cur != kInvalid && v != 0) {
                const uint64_t parent = rn[d], leaf = rl[d];
                if (parent != kEmpty) {
                    // the parent's children with one replaced; 8 equal leaves collapse
                    uint32_t types = bid_types(parent) | ((bid_is_leaf(cur) ? 1u : 0u) << ci);
                    bool all = true;
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        ch[i] = i == (int)ci ? cur : P.children[8ull * bid_index(parent) + i];
                        all = all && ch[i] == cur;
                    }
                    if (types == 0xFFu && all) {
                        out = cur;
                        collapsed = true;
                    } else {
                        tm = (types << 8) | ((bid_mask(parent) | bit) & 0xFFu);
                        cand = true;
                    }
                } else if (leaf == kEmpty) {
#pragma unroll
                    for (int i = 0; i < 8; i++) ch[i] = i == (int)ci ? cur : kEmpty;
                    tm = (((bid_is_leaf(cur) ? 1u : 0u) << ci) << 8) | bit;
                    cand = true;
                } else {
                    // a split leaf: its 7 other slots keep it (inc_ref_by(leaf, 7))
#pragma unroll
                    for (int i = 0; i < 8; i++) ch[i] = i == (int)ci ? cur : leaf;
                    const uint32_t types = (~bit | ((bid_is_leaf(cur) ? 1u : 0u) << ci)) & 0xFFu;
                    tm = (types << 8) | 0xFFu;
                    cand = true;
                }
            } else if (cur != kInvalid) {
                if (d >= meta[c]) {
                    // remove_at_depth_leaf: the split leaf in 7 slots, the result in one
                    const uint64_t N = rl[d];
#pragma unroll
                    for (int i = 0; i < 8; i++) ch[i] = i == (int)ci ? cur : N;
                    const uint32_t types = (~bit | ((bid_is_leaf(cur) ? 1u : 0u) << ci)) & 0xFFu;
                    const uint32_t mask = (~bit | ((cur != kEmpty ? 1u : 0u) << ci)) & 0xFFu;
                    tm = (types << 8) | mask;
                    cand = true;
                } else {
                    // remove_at_depth_branch: an emptied child leaves the mask; no
                    // child left removes the branch
                    const uint64_t node = rn[d];
                    const uint32_t mask = cur == kEmpty ? (bid_mask(node) & ~bit) : bid_mask(node);
                    if (mask == 0) {
                        out = kEmpty;
                    } else {
#pragma unroll
                        for (int i = 0; i < 8; i++) ch[i] = i == (int)ci ? cur : P.children[8ull * bid_index(node) + i];
                        tm = ((bid_types(node) & ~bit) << 8) | mask;
                        cand = true;
                    }
                }
            }
            if (cand) out = res_pending(0xFFFFFFFFu);  // replaced by the probe
        }
    }
    if (valid) res[slot] = out;
    emit_candidate<TILE>(K, &s_cnt, cand, ch, tm, slot, blockIdx.x);
    tile_count(s_st, 0, cand);
    tile_count(s_st, 1, collapsed);
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = s_cnt;
    const int ids[2] = {ST_CAND, ST_COLLAPSE};
    if (threadIdx.x < 2 && s_st[threadIdx.x])
        atomicAdd((unsigned long long*)stat(P.ctl, ids[threadIdx.x]), (unsigned long long)s_st[threadIdx.x]);
}

// ------------------------------------------------------------------ fill / clear

// VoxOpsBulkWrite::fill/clear (voxtree.rs:263-292): queue the old root for
// dec_ref_recursive first, then (after GC and leaf creation) install leaf(v).
template <typename T>
__global__ void k_fill_prepare(Pool<T> P, uint32_t n, const uint64_t* roots, const uint64_t* vals,
                               uint32_t* gc_front) {
    const uint32_t vmask = sizeof(T) == 1 ? 0xFFu : sizeof(T) == 2 ? 0xFFFFu : 0xFFFFFFFFu;
    VX_LOOP(c, n) {
        bool push = false;
        uint64_t r = kEmpty;
        if (c < n) {
            r = roots[c];
            push = r != kEmpty;
            uint32_t v = (uint32_t)vals[c] & vmask;
            if (v != 0 && leaf_index(P, v) == 0) leaf_request(P, v);
        }
        uint32_t k = wave_append(&P.ctl->gc_count[0], push);
        if (push) gc_front[k] = bid_index(r);
    }
}

template <typename T>
__global__ void k_fill_finish(Pool<T> P, uint32_t n, uint64_t* roots, const uint64_t* vals) {
    const uint32_t vmask = sizeof(T) == 1 ? 0xFFu : sizeof(T) == 2 ? 0xFFFFu : 0xFFFFFFFFu;
    VX_LOOP(c, n) {
        if (c < n) {
            uint32_t v = (uint32_t)vals[c] & vmask;
            if (v == 0) {
                roots[c] = kEmpty;
            } else {
                uint32_t idx = leaf_index(P, v);
                if (idx != 0) {
                    atomicAdd(&P.refs[idx], 1u);
                    roots[c] = make_leaf(idx, tmg_gen(P.tmg[idx]));
                }
            }
        }
    }
}

template <typename T>
__global__ void k_leaf_alloc_only(Pool<T> P, uint32_t nvalues) {
    const uint32_t created = create_requested_leaves(P, nvalues);
    if (threadIdx.x == 0) {
        // fill released the old roots first: there is no state to return to
        if (P.ctl->abort_) atomicOr(&P.ctl->err, kErrOom | kErrPoison);
        atomicAdd((unsigned long long*)&P.ctl->live_leaves, (unsigned long long)created);
        atomicAdd((unsigned long long*)stat(P.ctl, ST_LEAF_MISS), (unsigned long long)created);
    }
}

// ------------------------------------------------------------------ reads

// The rest of get_at_depth (utils/common.rs:122-156) from `node` at depth d of a tree
// of depth D: child digits from bit D-d-1 of x, y, z down to a leaf or an empty child.
template <typename T>
__device__ inline void get_walk(const Pool<T>& P, uint64_t node, uint32_t d, uint32_t D, uint32_t x, uint32_t y,
                                uint32_t z, uint32_t* v_out, uint8_t* pr_out, uint8_t* stop_out) {
    uint32_t v = 0;
    uint8_t pr = 0;
    bool at_leaf = false;
    while (node != kEmpty) {
        const uint32_t idx = bid_index(node);
        if (idx >= P.capacity) {
            atomicOr(&P.ctl->err, kErrPoison);
            break;
        }
        if (d >= D) {
            v = (uint32_t)P.values[idx];
            pr = v != 0;
            break;
        }
        if (!bid_is_leaf(node)) {
            const uint32_t sh = D - d - 1;
            const uint32_t ci = ((x >> sh) & 1u) | (((y >> sh) & 1u) << 1) | (((z >> sh) & 1u) << 2);
            node = P.children[8ull * idx + ci];
            d++;
        } else {
            v = (uint32_t)P.values[idx];
            pr = 1;
            at_leaf = true;
            break;
        }
    }
    *v_out = pr ? v : 0u;
    *pr_out = pr;
    *stop_out = (uint8_t)(d | (at_leaf ? 0x80u : 0u));
}

// VoxOpsRead::get (voxtree.rs:145-159) -> get_at_depth (utils/common.rs:122-156),
// one lane per query.
template <typename T>
__global__ void k_get(Pool<T> P, uint64_t nq, uint64_t ntrees, const uint64_t* roots,
                      const uint8_t* depths, const uint32_t* chunk_of, const int32_t* xyz, T* out,
                      uint8_t* present, uint8_t* stop = nullptr) {
    VX_LOOP(i, nq) {
        if (i < nq) {
            uint32_t c = chunk_of[i];
            bool tree_ok = c < ntrees;
            uint64_t node = tree_ok ? roots[c] : kEmpty;
            uint32_t D = tree_ok ? depths[c] : 0u;
            if (D > (uint32_t)kMaxDepth) tree_ok = false;
            int32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
            int32_t lim = 1 << D;
            uint32_t v = 0;
            uint8_t pr = 0, sd = 0;
            if (!tree_ok || x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) {
                atomicOr(&P.ctl->err, kErrRange);
                pr = 0xFF;
            } else {
                get_walk(P, node, 0, D, (uint32_t)x, (uint32_t)y, (uint32_t)z, &v, &pr, &sd);
            }
            out[i] = (T)v;
            if (present) present[i] = pr;
            if (stop) stop[i] = sd;
        }
    }
}

// get for queries grouped by tree: tree t's queries are [qoff[t], qoff[t+1]).  A block
// takes `per` consecutive queries.  For each tree its range meets, the block first walks
// every path of the tree's top k levels (k = 3 for a segment of at least 256 queries,
// 2 from 32, else none) and keeps the node each path reaches in LDS; a query then takes
// the node of its first k child digits from there and walks the rest in HBM
// (utils/common.rs:122-156): a depth-5 read makes three dependent loads instead of six.
// A query outside its tree, or beyond qoff[ntrees], is VX_E_RANGE (present 0xFF).
constexpr uint32_t kGetTop = 3;
template <typename T, int TILE>
__global__ __launch_bounds__(TILE) void k_get_grouped(Pool<T> P, uint64_t nq, uint32_t ntrees,
                                                      const uint64_t* roots, const uint8_t* depths,
                                                      const uint64_t* qoff, const int32_t* xyz, T* out,
                                                      uint8_t* present, uint8_t* stop, uint64_t per) {
    __shared__ uint64_t s_node[1u << (3 * kGetTop)];
    __shared__ uint8_t s_dep[1u << (3 * kGetTop)];
    __shared__ uint32_t s_lo;
    const uint64_t q1 = min(nq, (uint64_t)(blockIdx.x + 1) * per);
    uint64_t q0 = (uint64_t)blockIdx.x * per;
    if (q0 >= q1) return;
    // the tree holding q0, the last t with qoff[t] <= q0: a TILE-way search per round
    {
        uint32_t lo = 0, span = ntrees;
        while (span > 1) {  // uniform
            const uint32_t step = (span + TILE - 1) / TILE;
            const uint32_t t = lo + threadIdx.x * step;
            const int below = t < lo + span && t < ntrees && qoff[t] <= q0;
            const uint32_t cnt = (uint32_t)__syncthreads_count(below);
            lo += (cnt ? cnt - 1 : 0) * step;
            span = min(step, ntrees - lo);
        }
        if (threadIdx.x == 0) s_lo = lo;
        __syncthreads();
    }
    uint32_t t = s_lo;
    while (q0 < q1) {  // uniform: one segment (tree) per pass
        while (t + 1 < ntrees && qoff[t + 1] <= q0) t++;  // skip trees with no query here
        const bool t_ok = t < ntrees;
        const uint64_t a = t_ok ? qoff[t] : 0ull, b = t_ok ? qoff[t + 1] : 0ull;
        const bool seg_ok = t_ok && a <= q0 && b > q0;
        const uint64_t e = seg_ok ? min(q1, b) : q1;
        const uint64_t root = seg_ok ? roots[t] : kEmpty;
        const uint32_t D = seg_ok ? depths[t] : 0xFFu;
        const bool tree_ok = seg_ok && D <= (uint32_t)kMaxDepth;
        const uint64_t n_seg = e - q0;
        const uint32_t k = !tree_ok ? 0u : min(D, n_seg >= 256 ? kGetTop : n_seg >= 32 ? 2u : 0u);
        if (k) {
            const uint32_t np = 1u << (3 * k);
            for (uint32_t p = threadIdx.x; p < np; p += TILE) {
                uint64_t node = root;
                uint32_t d = 0;
                for (uint32_t l = 0; l < k; l++) {
                    if (node == kEmpty || bid_is_leaf(node)) break;
                    const uint32_t idx = bid_index(node);
                    if (idx >= P.capacity) {
                        atomicOr(&P.ctl->err, kErrPoison);
                        node = kEmpty;
                        break;
                    }
                    node = P.children[8ull * idx + ((p >> (3 * (k - 1 - l))) & 7u)];
                    d++;
                }
                s_node[p] = node;
                s_dep[p] = (uint8_t)d;
            }
            __syncthreads();
        }
        for (uint64_t q = q0 + threadIdx.x; q < e; q += TILE) {
            const int32_t x = xyz[3 * q], y = xyz[3 * q + 1], z = xyz[3 * q + 2];
            const int32_t lim = tree_ok ? 1 << D : 0;
            uint32_t v = 0;
            uint8_t pr = 0xFF, sd = 0;
            if (!tree_ok || x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) {
                atomicOr(&P.ctl->err, kErrRange);
            } else {
                uint64_t node = root;
                uint32_t d = 0;
                if (k) {
                    uint32_t pth = 0;
                    for (uint32_t l = 0; l < k; l++) {
                        const uint32_t sh = D - l - 1;
                        pth = (pth << 3) | (((uint32_t)x >> sh) & 1u) | ((((uint32_t)y >> sh) & 1u) << 1) |
                              ((((uint32_t)z >> sh) & 1u) << 2);
                    }
                    node = s_node[pth];
                    d = s_dep[pth];
                }
                get_walk(P, node, d, D, (uint32_t)x, (uint32_t)y, (uint32_t)z, &v, &pr, &sd);
            }
            out[q] = (T)v;
            if (present) present[q] = pr;
            if (stop) stop[q] = sd;
        }
        __syncthreads();  // the staged paths are read before the next segment restages them
        q0 = e;
        t++;
    }
}

// ------------------------------------------------------------------ canonical hash

__device__ inline uint64_t canon_leaf(uint32_t v, uint64_t seed) {
    return mix64(seed ^ mix64((uint64_t)v + 0x4C454146ull));
}
__device__ inline uint64_t canon_branch_init(uint32_t v, uint32_t tm, uint64_t seed) {
    return mix64(seed ^ 0x4252414Eull ^ mix64((uint64_t)v) ^ ((uint64_t)tm << 32));
}
__device__ inline uint64_t canon_combine(uint64_t h, uint64_t child, uint32_t i) {
    return mix64(h ^ (child + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1)));
}

// Merkle hash of a tree, seeds S0/S1; one lane per tree, iterative DFS.
// to_vec (utils/common.rs:158-246) at a reduced depth (MaxDepth::for_lod,
// core/max_depth.rs:137-140): one lane per output voxel, output in y, z, x order.
// A branch reached at the reduced depth gives its LOD value; an empty child gives
// the default value; a leaf covers its whole cube.
template <typename T>
__global__ void k_to_vec(Pool<T> P, uint64_t n, const uint64_t* roots, const uint8_t* depths, uint32_t lod,
                         uint64_t stride, T* out) {
    const uint64_t total = n * stride;
    VX_LOOP(i, total) {
        if (i < total) {
            uint64_t t = i / stride, v = i - t * stride;
            uint32_t D = depths[t];
            if (D > (uint32_t)kMaxDepth) {
                atomicOr(&P.ctl->err, kErrRange);
                continue;
            }
            uint32_t Dl = D > lod ? D - lod : 0u;
            uint64_t N = 1ull << Dl;
            if (N * N * N > stride) {
                if (v == 0) atomicOr(&P.ctl->err, kErrRange);
                continue;
            }
            if (v >= N * N * N) continue;
            uint32_t x = (uint32_t)(v % N), z = (uint32_t)((v / N) % N), y = (uint32_t)(v / (N * N));
            uint64_t node = roots[t];
            for (uint32_t d = 0; d < Dl && node != kEmpty && !bid_is_leaf(node); d++) {
                if (bid_index(node) >= P.capacity) {
                    atomicOr(&P.ctl->err, kErrPoison);
                    node = kEmpty;
                    break;
                }
                uint32_t s = Dl - 1 - d;
                uint32_t ci = ((x >> s) & 1u) | (((y >> s) & 1u) << 1) | (((z >> s) & 1u) << 2);
                node = P.children[8ull * bid_index(node) + ci];
            }
            T val = 0;
            if (node != kEmpty && bid_index(node) < P.capacity) val = P.values[bid_index(node)];
            out[i] = val;
        }
    }
}

template <typename T>
__global__ void k_canon(Pool<T> P, uint32_t n, const uint64_t* roots, uint64_t* out) {
    const uint64_t S0 = 0x736F6D6570736575ull, S1 = 0x646F72616E646F6Dull;
    VX_LOOP(c, n) {
        if (c < n) {
            uint64_t root = roots[c];
            uint64_t r0 = 0, r1 = 0;
            if (bid_index(root) >= P.capacity) root = kEmpty;
            if (root != kEmpty) {
                uint32_t ri = bid_index(root);
                if (bid_is_leaf(root)) {
                    r0 = canon_leaf((uint32_t)P.values[ri], S0);
                    r1 = canon_leaf((uint32_t)P.values[ri], S1);
                } else {
                    uint64_t sid[kLevels + 2], sh0[kLevels + 2], sh1[kLevels + 2];
                    uint32_t snext[kLevels + 2];
                    int top = 0;
                    sid[0] = root;
                    snext[0] = 0;
                    sh0[0] = canon_branch_init((uint32_t)P.values[ri], bid_tm(root), S0);
                    sh1[0] = canon_branch_init((uint32_t)P.values[ri], bid_tm(root), S1);
                    while (top >= 0) {
                        if (snext[top] == 8) {
                            uint64_t h0 = sh0[top], h1 = sh1[top];
                            top--;
                            if (top < 0) {
                                r0 = h0;
                                r1 = h1;
                                break;
                            }
                            sh0[top] = canon_combine(sh0[top], h0, snext[top]);
                            sh1[top] = canon_combine(sh1[top], h1, snext[top]);
                            snext[top]++;
                            continue;
                        }
                        uint64_t cid = P.children[8ull * bid_index(sid[top]) + snext[top]];
                        if (bid_index(cid) >= P.capacity) cid = kEmpty;
                        if (cid == kEmpty) {
                            sh0[top] = canon_combine(sh0[top], 0, snext[top]);
                            sh1[top] = canon_combine(sh1[top], 0, snext[top]);
                            snext[top]++;
                        } else if (bid_is_leaf(cid)) {
                            uint32_t lv = (uint32_t)P.values[bid_index(cid)];
                            sh0[top] = canon_combine(sh0[top], canon_leaf(lv, S0), snext[top]);
                            sh1[top] = canon_combine(sh1[top], canon_leaf(lv, S1), snext[top]);
                            snext[top]++;
                        } else if (top + 1 < kLevels + 2) {
                            top++;
                            uint32_t vi = (uint32_t)P.values[bid_index(cid)];
                            sid[top] = cid;
                            snext[top] = 0;
                            sh0[top] = canon_branch_init(vi, bid_tm(cid), S0);
                            sh1[top] = canon_branch_init(vi, bid_tm(cid), S1);
                        } else {
                            snext[top]++;  // deeper than any valid tree: malformed, skip
                        }
                    }
                }
            }
            out[2 * c] = r0;
            out[2 * c + 1] = r1;
        }
    }
}

// explicit instantiations are produced by the host dispatch in vx_api.hip

}  // namespace vx
