/*
 * voxelis_oracle.h -- CPU oracle for the batched SVO-DAG edit path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a literal, single-threaded C restatement
 * of the reference Rust code (WildPixelGames/voxelis) for the path
 * VoxTree::apply_batch / VoxInterner hash-consing / VoxOpsRead::get.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it, and only as the checker.  The product library never links it.
 *
 * Every value is carried as uint64_t and masked to the configured voxel width
 * (1, 2, 4 or 8 bytes); the reference's VoxelTrait only needs equality and
 * T::default() == 0 on this path, so one code path covers u8..i64.
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *  - BlockId bit layout .......... golden raw 0x800001C80000007B (block_id.rs:405-412)
 *                                    and the block_id.rs unit tests
 *  - PATH_MASKS formula .......... the reference table voxtree.rs:41-105 (golden fixture)
 *  - Morton path encoder ......... brute force against bit interleaving
 *  - allocator order ............. pool_allocator_lite.rs tests (LIFO reuse, OOM)
 *  - apply_batch / get / set ..... the 32 voxtree.rs unit tests (ported), the
 *                                    config-1 closed form (2 leaves, 5 branches,
 *                                    root id 0x007F800000000007, refcounts 4/8/1),
 *                                    dense set_sum closed form 3*2^D-2 leaves /
 *                                    3(2^D-1)-2D branches, and dense-shadow reads
 *  - calc_average ................ brute-force mode over all small inputs
 *  - canonical hash .............. invariance under node-id permutation (two
 *                                    insertion orders give equal hashes),
 *                                    and sensitivity to single-voxel changes
 * Parity unpinned: the exact rustc-hash 2.1.1 FxHash values (identity here is
 * a full-key compare, so only hash collisions could differ).
 */
#ifndef VOXELIS_ORACLE_H
#define VOXELIS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Panic codes: the reference panics; the oracle reports the panic instead. */
enum {
    VXO_OK = 0,
    VXO_PANIC_OOM = -1,          /* interner/macros.rs:37-38 "Out of memory" */
    VXO_PANIC_DEPTH = -2,        /* max_depth.rs:77-83, batch underflow for D<2 */
    VXO_PANIC_RANGE = -3,        /* voxtree.rs:146-148, 164-166 */
    VXO_PANIC_LEAF_TYPES = -4,   /* block_id.rs:273-276 types()/mask() on a leaf */
    VXO_PANIC_SAME_ROOT = -5,    /* voxtree.rs:311 assert_ne!(new_root, root) */
    VXO_PANIC_EXPIRED = -6,      /* interner/mod.rs:662-667 */
    VXO_PANIC_BUDGET = -7,       /* interner/mod.rs:63-68 */
    VXO_PANIC_GENERATION = -8,   /* block_id.rs:217 */
    VXO_PANIC_INVALID = -9,      /* voxtree.rs:316-318, 380 */
    VXO_PANIC_REMOVE = -10,      /* voxtree.rs:591, 624 */
};

typedef struct vxo_interner vxo_interner;
typedef struct vxo_batch vxo_batch;

/* Hazard and statistics counters, index into vxo_stats() output. */
enum {
    VXO_ST_LEAF_HITS = 0,
    VXO_ST_LEAF_MISSES,
    VXO_ST_BRANCH_HITS,
    VXO_ST_BRANCH_MISSES,
    VXO_ST_COLLAPSES,
    VXO_ST_RECYCLED,
    VXO_ST_H2_UNCOLLAPSED,   /* Phase-1 brick of 8 equal leaves kept as a branch */
    VXO_ST_H3_LEAKS,         /* Phase-0 fill leaf kept with patches; split with no change */
    VXO_ST_H4_UNDERCOUNT,    /* split brick slot equal to the split leaf, skipped w/o ref */
    VXO_ST_H5_COLLAPSE_CLONED,
    VXO_ST_TRANSIENT_FREE,   /* node freed by a dec inside apply_batch */
    VXO_ST_COUNT
};

/* ---- BlockId helpers (core/block_id.rs) ---- */
uint64_t vxo_bid_new_leaf(uint32_t index, uint16_t generation);
uint64_t vxo_bid_new_branch(uint32_t index, uint16_t generation, uint8_t types, uint8_t mask);
uint32_t vxo_bid_path_mask(uint32_t max_depth, uint32_t level);
uint32_t vxo_encode_child_index_path(int32_t x, int32_t y, int32_t z);
uint64_t vxo_calc_average(const uint64_t* values, size_t n);

/* ---- interner (interner/mod.rs) ---- */
vxo_interner* vxo_interner_new(uint32_t value_size, uint64_t budget_bytes, int* err);
void vxo_interner_free(vxo_interner* o);
void vxo_set_extended_depth(vxo_interner* o, int enable); /* allow D == 7 */
/* Snapshot / restore of the whole mutable interner state (test infrastructure: the
 * device path leaves everything unchanged when the reference would panic). */
typedef struct vxo_snapshot vxo_snapshot;
vxo_snapshot* vxo_snapshot_take(const vxo_interner* o);
void vxo_snapshot_restore(vxo_interner* o, const vxo_snapshot* s);
void vxo_snapshot_free(vxo_snapshot* s);
uint64_t vxo_capacity(const vxo_interner* o);
uint64_t vxo_node_size(uint32_t value_size);
void vxo_counts(const vxo_interner* o, uint64_t* leaf_patterns, uint64_t* branch_patterns);
uint32_t vxo_get_ref(const vxo_interner* o, uint64_t id);
uint64_t vxo_get_value(const vxo_interner* o, uint64_t id);
void vxo_get_children(const vxo_interner* o, uint64_t id, uint64_t out[8]);
void vxo_stats(const vxo_interner* o, uint64_t out[VXO_ST_COUNT]);
int vxo_get_or_create_leaf(vxo_interner* o, uint64_t value, uint64_t* out);
int vxo_get_or_create_branch(vxo_interner* o, const uint64_t children[8], uint8_t types,
                             uint8_t mask, uint64_t* out);
int vxo_dec_ref_recursive(vxo_interner* o, uint64_t id);
uint64_t vxo_alive_nodes(const vxo_interner* o);

/* ---- batch (core/batch.rs) ---- */
vxo_batch* vxo_batch_new(uint8_t max_depth);
void vxo_batch_free(vxo_batch* b);
void vxo_batch_just_set(vxo_batch* b, int32_t x, int32_t y, int32_t z, uint64_t v);
void vxo_batch_set_many(vxo_batch* b, size_t n, const int32_t* xyz, const uint64_t* values);
void vxo_batch_just_fill(vxo_batch* b, uint64_t v);
void vxo_batch_just_clear(vxo_batch* b);
size_t vxo_batch_size(const vxo_batch* b);
int vxo_batch_has_patches(const vxo_batch* b);
void vxo_batch_load(vxo_batch* b, const uint8_t* masks, const void* values, uint32_t vsize,
                    int has_patches);
size_t vxo_batch_bricks(const vxo_batch* b);
const uint8_t* vxo_batch_masks(const vxo_batch* b);   /* 2*P bytes, (set, clear) */
const uint64_t* vxo_batch_values(const vxo_batch* b); /* 8*P values */

/* ---- tree (spatial/voxtree.rs); the caller owns the root id ---- */
int vxo_tree_get(const vxo_interner* o, uint64_t root, uint8_t depth, int32_t x, int32_t y,
                 int32_t z, uint64_t* value, int* present);
int vxo_tree_get_many(const vxo_interner* o, uint64_t root, uint8_t depth, size_t n,
                      const int32_t* xyz, uint64_t* values, uint8_t* present);
int vxo_tree_set(vxo_interner* o, uint64_t* root, uint8_t depth, int32_t x, int32_t y,
                 int32_t z, uint64_t v, int* changed);
int vxo_tree_fill(vxo_interner* o, uint64_t* root, uint8_t depth, uint64_t v);
int vxo_trees_set_round(vxo_interner* o, size_t n, uint64_t* roots, const uint8_t* depths, const int32_t* xyz,
                        const uint64_t* values, int* status);
int vxo_tree_clear(vxo_interner* o, uint64_t* root);
int vxo_tree_apply_batch(vxo_interner* o, uint64_t* root, uint8_t depth, const vxo_batch* b,
                         int* changed);
int vxo_tree_to_vec(const vxo_interner* o, uint64_t root, uint8_t depth, uint64_t* out);
int vxo_tree_canonical_hash(vxo_interner* o, uint64_t root, uint8_t depth, uint64_t out[2]);
/* live patterns: leaf values and branch canonical hash pairs (test infrastructure) */
int vxo_live_patterns(vxo_interner* o, uint64_t* leaf_values, uint64_t* branch_canon, size_t cap,
                      size_t* n_leaves, size_t* n_branches);

#ifdef __cplusplus
}
#endif
#endif
