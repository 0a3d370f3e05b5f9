// vx_device.h -- device-side types and helpers of the MI355X SVO-DAG edit path.
//
// Node ids keep the reference's 64-bit BlockId layout (core/block_id.rs:18-30):
//   bit 63 leaf | 62..55 types | 54..47 mask | 46..32 generation | 31..0 index.
// The interner is a set of node columns in HBM indexed by the id's index field,
// a branch pattern table (open addressing, 8-byte slots, atomicCAS insertion) and
// a direct-mapped leaf table indexed by the voxel value.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vx {

constexpr uint64_t kEmpty = 0ull;          // BlockId::EMPTY (block_id.rs:123)
constexpr uint64_t kInvalid = ~0ull;       // BlockId::INVALID (block_id.rs:120)
constexpr uint32_t kMaxGeneration = 0x7FFEu;  // BlockId::MAX_GENERATION (block_id.rs:130)
constexpr int kMaxDepth = 7;               // 6 is the reference's limit; 7 is the extension
constexpr int kLevels = kMaxDepth + 1;

// A level result that still points at the owner candidate of a key inserted in
// the same launch: generation field 0x7FFF (never a valid generation), payload =
// candidate index in that level's candidate list.
constexpr uint64_t kPendingHi = 0x7FFFull;
__host__ __device__ inline bool res_is_pending(uint64_t r) { return (r >> 32) == kPendingHi; }
__host__ __device__ inline uint64_t res_pending(uint32_t cand) { return (kPendingHi << 32) | cand; }
// Phase-1 marker for "the fill leaf of this chunk" before its id is known.
constexpr uint64_t kFillMark = (1ull << 63) | (0x7FFFull << 32);

// Table slots: 0 = empty, 1 = tombstone, else tag(28) | level(3) | pending(1) | payload(32).
// A pending slot's payload is a candidate index of level `level` (1..7).
constexpr uint64_t kSlotEmpty = 0ull;
constexpr uint64_t kSlotTomb = 1ull;

__host__ __device__ inline bool bid_is_leaf(uint64_t id) { return (id >> 63) != 0; }
__host__ __device__ inline uint32_t bid_index(uint64_t id) { return (uint32_t)id; }
// A node's types/mask word and its generation share one 32-bit column entry, so a free or
// a commit touches one sector for both.
__host__ __device__ inline uint32_t tmg_tm(uint32_t w) { return w & 0xFFFFu; }
__host__ __device__ inline uint32_t tmg_gen(uint32_t w) { return w >> 16; }
__host__ __device__ inline uint32_t tmg_make(uint32_t tm, uint32_t gen) { return (gen << 16) | (tm & 0xFFFFu); }
__host__ __device__ inline uint32_t bid_gen(uint64_t id) { return (uint32_t)(id >> 32) & 0x7FFFu; }
__host__ __device__ inline uint32_t bid_tm(uint64_t id) { return (uint32_t)(id >> 47) & 0xFFFFu; }
__host__ __device__ inline uint32_t bid_types(uint64_t id) { return (uint32_t)(id >> 55) & 0xFFu; }
__host__ __device__ inline uint32_t bid_mask(uint64_t id) { return (uint32_t)(id >> 47) & 0xFFu; }
__host__ __device__ inline uint64_t make_leaf(uint32_t idx, uint32_t gen) {
    return (1ull << 63) | ((uint64_t)gen << 32) | idx;
}
// tm = (types << 8) | mask lands on bits 62..47
__host__ __device__ inline uint64_t make_branch(uint32_t idx, uint32_t gen, uint32_t tm) {
    return ((uint64_t)tm << 47) | ((uint64_t)gen << 32) | idx;
}

// splitmix64 finalizer (Steele, Lea & Flood 2014; constants from Vigna's splitmix64.c)
__host__ __device__ inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// Hash of a branch key (children ids, types, mask).  Identity is always decided by
// a full-key compare; the hash only places the key in the table.
//
// NH-style (the UMAC/VHASH inner hash): the sum, mod 2^64, of products of pairs of
// 32-bit halves, each offset (mod 2^32) by its own odd constant, then the splitmix64
// finalizer.  Pairs join halves of different children (low with low, high with high),
// so a product of two data terms keeps structured keys (consecutive leaf ids, one new
// child among empty ones) from cancelling as a linear combination could.  Eight 32x32
// products where mixing every word took eighteen 64-bit ones: the brick and level
// passes hash every key they emit, and their issue time is mostly vector ALU.
__host__ __device__ inline uint64_t branch_hash(const uint64_t ch[8], uint32_t tm) {
    const auto lo = [&](int i) { return (uint32_t)ch[i]; };
    const auto hi = [&](int i) { return (uint32_t)(ch[i] >> 32); };
    uint64_t s = (uint64_t)(tm + 0x50AB4B57u) * 0x80FE62C5u;
    s += (uint64_t)(lo(0) + 0x4ABEA221u) * (lo(1) + 0x23148989u);
    s += (uint64_t)(lo(2) + 0x609DFE03u) * (lo(3) + 0xDEB12801u);
    s += (uint64_t)(lo(4) + 0x6C442CB7u) * (lo(5) + 0xC7E4F8C5u);
    s += (uint64_t)(lo(6) + 0xF8FBA7E5u) * (lo(7) + 0x3CDB9EEBu);
    s += (uint64_t)(hi(0) + 0xF9520069u) * (hi(2) + 0xE116868Bu);
    s += (uint64_t)(hi(1) + 0xA2023CBDu) * (hi(3) + 0xA37B51B9u);
    s += (uint64_t)(hi(4) + 0x524F3905u) * (hi(6) + 0x6CA3B277u);
    s += (uint64_t)(hi(5) + 0x5104E85Bu) * (hi(7) + 0x9FD533B3u);
    return mix64(s);
}

__host__ __device__ inline uint32_t slot_tag(uint64_t h) {
    uint32_t t = (uint32_t)(h >> 36);
    return t ? t : 1u;
}
__host__ __device__ inline uint64_t slot_make(uint32_t tag, bool pending, uint32_t level, uint32_t payload) {
    return ((uint64_t)tag << 36) | ((uint64_t)(level & 7u) << 33) | ((uint64_t)(pending ? 1 : 0) << 32) |
           payload;
}
__host__ __device__ inline uint32_t slot_tag_of(uint64_t s) { return (uint32_t)(s >> 36); }
__host__ __device__ inline uint32_t slot_level(uint64_t s) { return (uint32_t)(s >> 33) & 7u; }
__host__ __device__ inline bool slot_pending(uint64_t s) { return ((s >> 32) & 1ull) != 0; }
__host__ __device__ inline uint32_t slot_payload(uint64_t s) { return (uint32_t)s; }

// Device control block: allocator cursors, live counters, errors, statistics.
enum : uint32_t { kErrOom = 1u, kErrTableFull = 2u, kErrRange = 4u, kErrPoison = 8u };
constexpr int kStatStripes = 64;
// Build/probe/commit kernels work on tiles of TILE level slots (one block each,
// TILE = 128, 256 or 512, a tuning choice); a tile's candidates are compacted into
// its own region of the candidate arrays.

struct Ctl {
    uint32_t free_top;    // entries on the free stack
    uint32_t next_index;  // bump cursor (interner/macros.rs:16-34)
    uint32_t free_top0;   // snapshot at the start of a launch sequence
    uint32_t next0;
    uint32_t alloc_count;  // allocations in the current sequence
    uint32_t err;
    uint32_t n_new_leaves;
    uint32_t n_deferred;  // bricks waiting for leaves created later in the sequence
    uint64_t live_leaves;
    uint64_t live_branches;
    uint32_t rebuild;   // this call rebuilds the pattern table (bit 0) and the i32 leaf table (bit 1)
    uint32_t rebuilds;  // table rebuilds so far
    uint32_t n_drift;  // queued reference-count corrections of the reference's quirks
    uint32_t dry_owners;      // dry run: branch keys that would be new
    uint32_t dry_new_leaves;  // dry run: leaves that would be new
    uint32_t abort_;          // this apply could not allocate: 1 capacity, 2 reservation
    uint32_t n_undo;          // entries of the undo log (reference-count deltas applied)
    uint32_t h1_fail;         // a tree of this apply failed in the brick pass (leaf-brick panic)
    uint32_t n_need;          // i32: values requested by this apply (need_list entries)
    uint32_t n_owners;        // new branch nodes of the last apply (its owner list), for the GC
    uint32_t pad1;
    uint32_t ncand[kLevels + 1];
    uint32_t gc_count[kLevels + 2];
    uint32_t own_base;  // alloc_count when the apply's first probe starts
    // statistics, striped over 64 cache lines by block index (summed when read)
    uint64_t st[kStatStripes][32];
    // u8 interners: the leaf id of every value at the start of the current apply
    // (written by k_prepare, read by the brick pass; leaves created later in the
    // apply are not in it)
    uint64_t leaf_ids8[256];
};

// Per-tree record of one apply launch sequence (trees sorted by depth on the host).
struct ChunkDev {  // 8-byte aligned; depth .. skip form one dword (read as such)
    const uint8_t* masks;
    const void* values;
    uint64_t fill;
    uint64_t root_in;
    uint64_t init;  // the node Phase 1/2 descend from (voxtree.rs:742-754)
    uint32_t orig;  // index in the caller's arrays
    int32_t status;
    uint8_t depth;
    uint8_t has_fill;
    uint8_t has_patches;
    uint8_t skip;  // not applied at all (an edit-list tree without edits)
    int32_t status0;  // status before the launch (a replayed launch starts from it)
};
static_assert(offsetof(ChunkDev, depth) % 4 == 0 && offsetof(ChunkDev, has_fill) == offsetof(ChunkDev, depth) + 1 &&
                  offsetof(ChunkDev, has_patches) == offsetof(ChunkDev, depth) + 2,
              "the flag bytes are read as one dword");

// Slot space of one level: groups of trees of equal depth, each tree owning
// 8^(depth-h) consecutive slots at height h (h = 1: bricks, h = depth: root).
struct LevelInfo {
    uint64_t slot_base[kLevels + 1];
    uint64_t prev_base[kLevels];  // the same group's first slot at height h-1
    uint64_t next_base[kLevels];  // ... at height h+1 (~0 when its depth is h)
    uint32_t chunk_base[kLevels];
    uint32_t depth[kLevels];
    uint32_t ngroups;
    uint32_t h;
    uint64_t nslots;
};

// Per height h >= 2: the previous tree's node at every level-h slot (a branch,
// the leaf that covers the position from above, or empty), written by one
// descent pass at the start of an apply for trees without a fill.
struct OldNodes {
    uint64_t* at[kLevels + 1];
};

// Candidate list of one level (compacted, order of discovery).
struct Cands {
    uint64_t* key;    // 9 words per candidate: children[8], tm
    uint64_t* hash;
    uint64_t* id;     // owner: the new node's id
    uint64_t* tslot;  // owner: its table slot; ~0 otherwise
    uint64_t* slot;   // the level slot this candidate resolves
};

__host__ __device__ inline uint64_t& key_word(const Cands& K, uint64_t k, int i) { return K.key[9ull * k + i]; }

// A queued reference-count correction (applied once the tree's status is final).
struct Drift {
    uint32_t chunk;
    uint32_t idx;
    int32_t delta;
    uint32_t stat;
};

template <typename T>
struct Pool {
    uint64_t* children;  // capacity * 8
    uint32_t* tmg;       // (generation << 16) | (types << 8) | mask; types/mask 0 for leaves and free slots
    T* values;
    uint32_t* refs;
    uint32_t* free_stack;
    uint64_t* tab;
    uint64_t tmask;
    uint64_t* single;  // u8/u16: the direct-mapped index of one-leaf keys (nullptr: none)
    uint32_t* leaf_map;  // u8/u16: value -> node index (0 = none)
    uint64_t* ltab;      // i32: open-addressing leaf table (value << 32 | index)
    uint64_t lmask;      // its slot count - 1
    uint32_t* need_list; // i32: values requested by this apply (k_leaf_alloc creates them)
    uint32_t need_cap;
    uint64_t* leaf_ids;  // u16: the leaf id of every value at the start of the apply (k_prepare)
    uint8_t* leaf_need;
    uint32_t* new_leaves;
    Ctl* ctl;
    uint32_t capacity;
    uint32_t dry;          // 1: count what an apply would create, change nothing
    uint32_t alloc_limit;  // allocations this apply may make (its reservation)
    uint32_t dedup_rounds;  // 0: every key probes the table; else equal keys of a pass merge first
    uint32_t brick_merge_rounds;  // ... of the brick pass, before its keys become candidates
    uint32_t* per_tree;    // optional: new branch nodes owned by each tree (caller's index)
    const Cands* levels;   // candidate arrays of every level (pending keys of earlier levels)
    Drift* undo;           // reference-count deltas applied outside commits (rolled back on abort)
    uint64_t* owners;      // by allocation rank - Ctl::own_base: (height << 32) | candidate
    uint32_t undo_cap;
};

// An owner's table slot with bit 63 set when the slot held a tombstone before, or bit 62
// set when the key lives in the one-leaf index instead (the entry's index).
constexpr uint64_t kTslotTomb = 1ull << 63;
constexpr uint64_t kTslotSingle = 1ull << 62;
__host__ __device__ inline uint64_t tslot_index(uint64_t ts) { return ts & ~(kTslotTomb | kTslotSingle); }

// One-leaf keys: a branch whose only child is a leaf (mask and types one and the same
// bit) -- a brick holding one voxel, the most common key of sparse content.
__host__ __device__ inline bool tm_is_single(uint32_t tm) {
    const uint32_t m = tm & 0xFFu, t = (tm >> 8) & 0xFFu;
    return m != 0 && (m & (m - 1u)) == 0 && t == m;
}
// Index entries: 0 none, the node's BlockId (bit 63 clear), or a claim of the current
// launch (bit 63 set, level << 32 | candidate).
constexpr uint64_t kSinglePend = 1ull << 63;


}  // namespace vx
