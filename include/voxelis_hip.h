/*
 * voxelis_hip.h -- C ABI of the MI355X-native batched SVO-DAG edit path.
 *
 * Drop-in boundary for the reference's VoxTree / VoxInterner / Batch / VoxOps
 * API (WildPixelGames/voxelis, Rust).  Plain pointers and sizes only; every
 * entry point returns a vx_status (0 = VX_OK, negative = error).  Reference
 * panics map to error codes and leave the affected tree unchanged.
 *
 * Memory: "device" pointers are HIP device allocations (HBM); "host" pointers
 * are ordinary CPU memory.  One interner owns one HIP stream; entry points are
 * asynchronous on that stream unless documented otherwise, and writers must be
 * serialized per interner (the reference's &mut VoxInterner).
 */
#ifndef VOXELIS_HIP_H
#define VOXELIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    VX_OK = 0,
    VX_E_INVALID_ARG = -1,
    VX_E_OOM = -2,         /* interner/macros.rs:37-38 panic "Out of memory" */
    VX_E_DEPTH = -3,       /* core/max_depth.rs:77-83; batch apply with depth < 2 */
    VX_E_RANGE = -4,       /* spatial/voxtree.rs:146-148 coordinate assert */
    VX_E_LEAF_BRICK = -5,  /* voxtree.rs:829-834 + block_id.rs:273-276: types() on a leaf */
    VX_E_SAME_ROOT = -6,   /* voxtree.rs:311 assert_ne!(new_root, root) */
    VX_E_HIP = -7,         /* a HIP runtime call failed */
    VX_E_BUDGET = -8,      /* interner/mod.rs:63-68 budget too small / too large */
    VX_E_UNSUPPORTED = -9, /* voxel width not built into this library */
    VX_E_TABLE_FULL = -10, /* pattern table exhausted (load factor too high) */
    VX_E_POISONED = -11,   /* a previous device-side failure left the interner unusable */
    VX_E_SCRATCH = -12     /* larger than the fixed scratch of a no-allocation entry point */
} vx_status;

typedef struct vx_interner vx_interner;

/* Dense per-chunk edit record: the buffers of core/batch.rs:39-45 Batch<T>.
 * masks: 2*P bytes, (set_mask, clear_mask) per 2x2x2 brick, P = 8^(depth-1);
 * values: 8*P voxels of the interner's value size, slot i = x | y<<1 | z<<2.
 * Brick p covers Morton path p<<3 (utils/common.rs:24-55). */
typedef struct {
    const uint8_t* masks;  /* device */
    const void* values;    /* device */
    uint64_t fill_value;   /* Batch::to_fill (batch.rs:106-111) when has_fill */
    uint8_t depth;         /* MaxDepth of the target tree */
    uint8_t has_fill;
    uint8_t has_patches;   /* Batch::has_patches (batch.rs:127-132) */
    uint8_t reserved;      /* must be 0 */
} vx_batch_view;

/* ---- interner: interner/mod.rs ---- */

/* VoxInterner::<T>::with_memory_budget (interner/mod.rs:45-155).
 * value_size is sizeof(T): 1 (u8/i8), 2 (u16/i16) or 4 (i32/u32, core/voxel.rs:93; the
 * type of the reference's tests and benchmarks).  Values cross the boundary as their bit
 * patterns (the path compares them for equality only; 0 is T::default()).  capacity =
 * budget_bytes / (78 + value_size), exactly the reference's node_size.  load_factor
 * (0 = default 0.5) sizes the branch pattern table; with value_size 4 the leaf table is
 * hashed as well (2^32 values are not mapped directly) and has as many slots. */
int vx_interner_create(uint32_t value_size, uint64_t budget_bytes, double load_factor,
                       vx_interner** out);
int vx_interner_destroy(vx_interner* h);
/* VoxInterner::capacity (mod.rs:1021-1024) */
int vx_interner_capacity(const vx_interner* h, uint64_t* capacity);
/* leaf_patterns().len(), branch_patterns().len() (mod.rs:1031-1037; the EMPTY
 * branch pattern counts) and live node slots.  Synchronizes the stream. */
int vx_interner_counts(vx_interner* h, uint64_t* leaf_patterns, uint64_t* branch_patterns,
                       uint64_t* alive_nodes);
/* VoxInterner::get_ref (mod.rs:222-232).  Synchronizes. */
int vx_interner_get_ref(vx_interner* h, uint64_t block_id, uint32_t* ref_count);
/* VoxInterner::get_value (mod.rs:167-177).  Synchronizes. */
int vx_interner_get_value(vx_interner* h, uint64_t block_id, uint64_t* value);
/* VoxInterner::get_children (mod.rs:180-191).  Synchronizes. */
int vx_interner_get_children(vx_interner* h, uint64_t block_id, uint64_t children[8]);
/* Counters (hits, misses, collapses, hazards ...); see VX_STAT_* below.  Synchronizes. */
int vx_interner_stats(vx_interner* h, uint64_t* out, uint32_t n);
/* Use the caller's HIP stream (hipStream_t) instead of the interner's own. */
int vx_interner_set_stream(vx_interner* h, void* hip_stream);
/* Launch shape of the apply (no reference counterpart; picked by the autotuner in
 * voxelis_amd/autotune.py).  tile: lanes per block of the build / probe / commit
 * passes (128, 256 or 512; a block owns one tile of level slots and its candidate
 * region); dedup_rounds: 0 probes every key, 1-16 merges the equal keys of each
 * probe pass (block-wide) first, so one leader per key reads the table;
 * commit_grid / gc_grid: blocks of the commit pass and of each GC round after
 * the first (the first takes one block per 16 replaced roots, up to gc_grid);
 * probe_blocks: a probe launch over more tiles than this packs the candidates of
 * up to 16 tiles into one block, keeping at least this many blocks (0: one tile
 * per block), separately for the brick level (probe_blocks_bricks) and the levels
 * above; brick_merge_rounds: wavefront key-merge rounds of the brick pass, before
 * its keys become candidates (0-16; equal keys of one tree share a candidate);
 * brick_rounds: brick slots per lane of the brick pass (1, 2 or 4; 0: 4 when the
 * pass has at least 512 blocks of them, else 1).
 * Defaults {256, 8, 1024, 1024, 32768, 512, 4, 0}.  Synchronizes the interner's
 * stream. */
typedef struct {
    uint32_t tile;
    uint32_t dedup_rounds;
    uint32_t commit_grid;
    uint32_t gc_grid;
    uint32_t probe_blocks_bricks;
    uint32_t probe_blocks;
    uint32_t brick_merge_rounds;
    uint32_t brick_rounds;
} vx_tuning;
int vx_interner_set_tuning(vx_interner* h, const vx_tuning* tuning);
int vx_interner_get_tuning(const vx_interner* h, vx_tuning* tuning);

/* Capture the launch sequence of vx_trees_apply_batches[_ex] as a HIP graph and
 * replay it when a later call has the same arguments (same roots/status arrays,
 * same batch views, same options): repeated batched applies skip planning and
 * per-kernel launches.  Up to 256 sequences are kept (the least recently used is
 * dropped); any workspace growth rebuilds them.  Off by default; profiling (vx_interner_profile) bypasses it. */
int vx_interner_set_graphs(vx_interner* h, int enable);
/* Pattern-table maintenance (no reference counterpart: the reference's HashMap
 * removes entries outright, interner/mod.rs:518).  Frees here leave tombstones; a
 * write call rebuilds the table by itself (clear, re-insert every live branch) once
 * tombstones take 1/16 of the slots and live entries plus tombstones 1/2 (3/4 in the
 * i32 leaf table; decided on the device at the start of the call, in stream order, so
 * queued calls see it).
 * rebuild_table forces a rebuild now and synchronizes; table_stats reports the slot
 * count, the current tombstones and the rebuilds so far (synchronizes).  Either
 * output pointer may be NULL. */
int vx_interner_rebuild_table(vx_interner* h, uint64_t* rebuilds);
int vx_interner_table_stats(vx_interner* h, uint64_t* slots, uint64_t* tombstones, uint64_t* rebuilds);
/* Wait for all work on the interner's stream; returns the first device error. */
int vx_interner_sync(vx_interner* h);

enum {
    VX_STAT_BRANCH_HITS = 0,
    VX_STAT_BRANCH_MISSES,
    VX_STAT_BRANCH_DUPS,       /* same key twice in one launch: resolved to one node */
    VX_STAT_LEAF_MISSES,
    VX_STAT_COLLAPSES,
    VX_STAT_RECYCLED,
    VX_STAT_H2_UNCOLLAPSED,
    VX_STAT_H3_LEAKS,
    VX_STAT_H4_UNDERCOUNT,
    VX_STAT_H5_COLLAPSE_CLONED,
    VX_STAT_CANDIDATES,        /* branch keys built */
    VX_STAT_PROBES,            /* table slots inspected */
    VX_STAT_CLAIMS,            /* atomic claims of a free table slot */
    VX_STAT_CLAIMS_LOST,       /* claims that found the slot taken meanwhile */
    VX_STAT_MERGED,            /* keys that shared an equal key's candidate before the probe */
    VX_STAT_BRICK_KEYS,        /* the brick level's share of CANDIDATES ... */
    VX_STAT_BRICK_MERGED,      /* ... of MERGED */
    VX_STAT_BRICK_HITS,        /* ... of BRANCH_HITS */
    VX_STAT_BRICK_MISSES,      /* ... of BRANCH_MISSES */
    VX_STAT_BRICK_DUPS,        /* ... of BRANCH_DUPS */
    VX_STAT_COUNT
};

/* ---- batched apply: VoxTree::apply_batch (spatial/voxtree.rs:303-328) ---- */

/* Apply batches[i] to tree i for i in [0, n), all in one launch sequence.
 * roots: device, n BlockIds, read as the current roots and overwritten with
 * the new ones.  status: device, n int32: 1 = changed (apply_batch returned
 * true), 0 = unchanged (returned false), < 0 = vx_status of the reference's
 * panic for that tree (its root is left as it was).  Asynchronous. */
int vx_trees_apply_batches(vx_interner* h, size_t n, uint64_t* roots,
                           const vx_batch_view* batches, int32_t* status);

/* Single-tree drop-in for VoxTree::apply_batch with the Batch buffers in host
 * memory (batch.masks / batch.values are host pointers here).  root is a host
 * BlockId, updated in place; *changed receives apply_batch's bool.  Synchronous. */
int vx_tree_apply_batch_host(vx_interner* h, uint64_t* root, const vx_batch_view* batch,
                             int* changed);

/* Edit-list form of vx_trees_apply_batches: tree t's edits are
 * xyz[3e..3e+3) / values[e] for e in [edit_offsets[t], edit_offsets[t+1]) in
 * submission order, as if Batch::just_set (core/batch.rs:145-175) had been called
 * for each on a fresh Batch of that tree; a tree with no edits is not applied
 * (status 0).  The edits are radix-sorted by (tree, Morton path) on the device;
 * the last write to a voxel wins.  depths and edit_offsets (n+1 entries) are
 * host arrays; roots, xyz, values and status are device arrays.  A coordinate
 * outside the tree fails that tree with VX_E_RANGE (the reference's Vec index
 * panics).  Asynchronous. */
int vx_trees_apply_edits(vx_interner* h, size_t n, uint64_t* roots, const uint8_t* depths,
                         const uint64_t* edit_offsets, const int32_t* xyz, const void* values,
                         int32_t* status);

/* Options of the _ex forms below (the plain forms behave as zeroed options).
 *
 * Allocation: an apply whose new nodes (leaves + branches) do not fit the
 * interner's capacity, or exceed node_limit when that is nonzero, is rolled back
 * as a whole: every tree of the call keeps its root and gets status VX_E_OOM
 * (capacity; the reference panics, interner/macros.rs:37-38) or VX_E_BUDGET
 * (node_limit), and the interner is left as before the call.  node_limit lets a
 * caller reserve nodes per apply (admission by memory budget).
 *
 * VX_APPLY_DRY_RUN: run the apply's passes without changing anything (roots,
 * nodes, reference counts and the pattern table are left as they were); status
 * receives what the apply would return.  With counts, this is the exact count of
 * the nodes the apply would create.
 *
 * counts (nullable, device, 2 + n uint32): [0] new branch nodes, [1] new leaf
 * nodes, [2 + t] new branch nodes first created for tree t (a node shared by
 * several trees of the call counts once, for one of them).  0xFFFFFFFF in [0]
 * and [1] when the apply was rolled back. */
enum { VX_APPLY_DRY_RUN = 1u };
typedef struct {
    uint32_t flags;
    uint32_t reserved;   /* must be 0 */
    uint64_t node_limit; /* 0 = none */
    uint32_t* counts;
} vx_apply_opts;

/* vx_trees_apply_batches with options (no counterpart in the reference: the
 * dry run and node_limit serve admission by memory budget). */
int vx_trees_apply_batches_ex(vx_interner* h, size_t n, uint64_t* roots,
                              const vx_batch_view* batches, int32_t* status,
                              const vx_apply_opts* opts);
/* vx_trees_apply_edits with options. */
int vx_trees_apply_edits_ex(vx_interner* h, size_t n, uint64_t* roots, const uint8_t* depths,
                            const uint64_t* edit_offsets, const int32_t* xyz, const void* values,
                            int32_t* status, const vx_apply_opts* opts);

/* ---- batched reads: VoxOpsRead::get (voxtree.rs:145-159, utils/common.rs:122-156) ---- */

/* For q in [0, nq): tree chunk_of_query[q] (< n_trees; roots/depths indexed by
 * it), voxel xyz[3q..3q+3).  out_values receives the value or 0 (None);
 * out_present (nullable) receives 1 for Some.  Coordinates outside
 * [0, 2^depth) or a tree index >= n_trees set out_present = 0xFF and make the
 * next sync report VX_E_RANGE (the reference panics).  All pointers device.
 * Asynchronous. */
int vx_trees_get(vx_interner* h, size_t n_trees, const uint64_t* roots, const uint8_t* depths,
                 size_t nq, const uint32_t* chunk_of_query, const int32_t* xyz, void* out_values,
                 uint8_t* out_present);

/* vx_trees_get on the caller's HIP stream instead of the interner's.  Reads may
 * run concurrently with an apply on the interner's stream as long as that apply
 * does not edit the trees read (nodes reachable from a root that is not replaced
 * are neither rewritten nor freed); the caller orders reads after the applies
 * they depend on (e.g. with hipStreamWaitEvent).  out_stop (nullable, device, nq
 * bytes): the depth where the traversal stopped, + 0x80 when it stopped at a leaf
 * (a leaf at depth d covers a cube of side 2^(depth - d)). */
int vx_trees_get_on(vx_interner* h, void* hip_stream, size_t n_trees, const uint64_t* roots,
                    const uint8_t* depths, size_t nq, const uint32_t* chunk_of_query,
                    const int32_t* xyz, void* out_values, uint8_t* out_present, uint8_t* out_stop);

/* Batched VoxOpsRead::get with the queries grouped by tree, as a caller reading many
 * voxels of each chunk has them: tree t's queries are [query_offsets[t],
 * query_offsets[t+1]) (n_trees + 1 non-decreasing device offsets, the last one nq),
 * voxel xyz[3q..3q+3).  For every run of queries of one tree the top three levels of
 * the tree are walked once into LDS and each query continues from there
 * (utils/common.rs:122-156).  Results, out_present, out_stop and VX_E_RANGE (a voxel
 * outside its tree, a query outside the offsets) as vx_trees_get_on.  hip_stream NULL =
 * the interner's stream.  Asynchronous. */
int vx_trees_get_grouped(vx_interner* h, void* hip_stream, size_t n_trees, const uint64_t* roots,
                         const uint8_t* depths, const uint64_t* query_offsets, size_t nq, const int32_t* xyz,
                         void* out_values, uint8_t* out_present, uint8_t* out_stop);

/* ---- single-voxel write: VoxOpsWrite::set (voxtree.rs:162-260) ---- */

/* set(xyz[i], values[i]) on tree i for i in [0, n), all trees in one launch
 * sequence (a tree appears once per call; several sets on one tree take one call
 * each, in order).  A non-default value follows set_at_depth_iterative
 * (voxtree.rs:393-561), the default value remove_at_depth (voxtree.rs:563-705):
 * leaves are split along the path, 8 equal leaves collapse, emptied branches are
 * removed.  depths: host, n (1..7); roots, xyz (3n int32), values (n of the value
 * type) and status: device.  status[i]: 1 = changed (set returned true), 0 =
 * unchanged (false), < 0 = the reference's panic as a vx_status (VX_E_RANGE for a
 * coordinate outside the tree; VX_E_OOM rolls the whole call back).
 * Asynchronous. */
int vx_trees_set(vx_interner* h, size_t n, uint64_t* roots, const uint8_t* depths, const int32_t* xyz,
                 const void* values, int32_t* status);

/* ---- whole-tree ops: VoxOpsBulkWrite (voxtree.rs:263-292) ---- */

/* fill(value) for value != 0, clear() for value == 0, on trees[i] (device
 * roots).  Asynchronous. */
int vx_trees_fill(vx_interner* h, size_t n, uint64_t* roots, const uint64_t* values);

/* ---- dense extraction: to_vec (utils/common.rs:158-246) ---- */

/* Tree t written densely at out + t * out_stride voxels (device), N = 2^(depth[t] -
 * lod) per axis (MaxDepth::for_lod, core/max_depth.rs:137-140), index y*N*N + z*N +
 * x.  A branch reached at the reduced depth contributes its LOD value
 * (core/voxel.rs:96-141); lod 0 is the full-resolution to_vec.  depths: device.  A
 * tree whose N^3 exceeds out_stride is skipped and the next sync reports VX_E_RANGE.
 * Asynchronous. */
int vx_trees_to_vec(vx_interner* h, size_t n, const uint64_t* roots, const uint8_t* depths, uint32_t lod,
                    uint64_t out_stride, void* out);

/* ---- VTM snapshot of a resident world: VoxModel::serialize / deserialize
 * (world/voxmodel.rs:177-408, world/voxchunk.rs:381-440, io/varint.rs) ---- */

/* The data section of a .vtm file (what export_model_to_vtm hashes and writes after
 * its header, io/export.rs:90-151) for the trees roots[0..n) (device) at the chunk
 * positions positions[3n] (host): the live patterns renumbered 1.. (leaves, then
 * branches, each in node-index order) as varint records, then one chunk record per
 * tree.  *size receives the byte count; out (host, cap bytes) receives the data when
 * it is large enough (out = NULL: the size only; cap too small: VX_E_RANGE).
 * Synchronizes. */
int vx_model_serialize(vx_interner* h, size_t n, const uint64_t* roots, const int32_t* positions, void* out,
                       size_t cap, size_t* size);
/* vx_model_serialize without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate (a serving loop that checkpoints its world, code
 * that must not call hipMalloc).  The same bytes as vx_model_serialize, with every
 * scratch array (ranks, record sizes and offsets, the staged section) carved from one
 * fixed 64 MiB device array per GPU that the library holds until the process ends;
 * a world whose export needs more (about 70 bytes per live node and 36 per chunk, at
 * most) returns VX_E_SCRATCH and changes nothing.  Calls on one GPU share that array:
 * calls that overlap on one GPU (two interners on two threads) race.  Synchronizes. */
int vx_model_serialize_noalloc(vx_interner* h, size_t n, const uint64_t* roots, const int32_t* positions,
                               void* out, size_t cap, size_t* size);
/* A data section rebuilt into a fresh interner (interner/mod.rs:930-993: nodes at
 * their file ids; reference counts = parent references + roots): the chunks' roots
 * (device, ncap) and positions (host, 3 * ncap); *n = the chunk count (VX_E_RANGE
 * when above ncap).  VX_E_INVALID_ARG for a malformed section or an interner that
 * holds nodes already.  Synchronizes. */
int vx_model_deserialize(vx_interner* h, const void* data, size_t size, size_t ncap, uint64_t* roots,
                         int32_t* positions, size_t* n);

/* ---- verification: canonical structural hash (no reference counterpart) ---- */

/* 128-bit Merkle hash of each tree's DAG, independent of node ids: two trees
 * with the same shape, values and LOD values hash equal in any interner.
 * out: device, 2*n uint64.  Asynchronous. */
int vx_trees_canonical_hash(vx_interner* h, size_t n, const uint64_t* roots, uint64_t* out);

/* ---- measurement (no reference counterpart; the reference uses tracy spans) ---- */

/* Time every kernel launch with HIP events on the interner's stream (0 = off). */
int vx_interner_profile(vx_interner* h, int enable);
/* Synchronizes; fills, for kernel id k < n, the milliseconds and launch count
 * accumulated since the previous read, then resets them.  Returns the number of
 * kernel ids (name of id k: vx_kernel_name(k)). */
int vx_interner_profile_read(vx_interner* h, double* ms, uint64_t* launches, uint32_t n);
const char* vx_kernel_name(uint32_t id);

/* ---- node-growth predictor (no reference counterpart: admission by memory budget) ----
 * FP32 MLP 16 -> 64 -> 64 -> 1 with ReLU.  params (device, vx_mlp_param_count()
 * floats): W1[64][16] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3.  x: n rows of 16
 * features.  All pointers device; asynchronous on hip_stream (NULL = default). */
int vx_mlp_param_count(void);
int vx_mlp_forward(void* hip_stream, const float* params, const float* x, uint32_t n, float* y);
/* Gradient of L = (1/n) sum (y - target)^2 with respect to params (grad, same
 * layout), the outputs y, and L (loss, nullable).  workspace: n * 257 floats. */
int vx_mlp_gradient(void* hip_stream, const float* params, const float* x, const float* target, uint32_t n,
                    float* y, float* workspace, float* grad, float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8; step counts from 1). */
int vx_mlp_adam(void* hip_stream, float* params, const float* grad, float* m, float* v, uint32_t step,
                float lr);

/* Atomic-throughput ceiling (no reference counterpart): `ops` operations on
 * uniformly random addresses of buf (device, `slots` entries of 8 bytes for kind 0
 * and 4 bytes otherwise, slots a power of two >= 2048; kind 0: 64-bit CAS, 1: 32-bit
 * add, 2: 32-bit add on 64 hot counters, 3: one add per wavefront on a single counter,
 * 4: one add per wavefront on 64 counters 128 bytes apart; kinds 3-4 count one
 * operation per lane, so their rate is per wavefront x 64).  Kernel
 * time in *ms.  Synchronizes the stream; buf contents are overwritten. */
int vx_bench_atomics(void* hip_stream, int kind, void* buf, uint64_t slots, uint64_t ops, float* ms);
/* Plain copy of `bytes` (a multiple of 16, 16-byte aligned buffers) with the kernel
 * time in *ms: the reference kernel for scaling memory-traffic counters.  Synchronizes. */
int vx_bench_copy(void* hip_stream, void* dst, const void* src, uint64_t bytes, float* ms);

/* ---- device memory helpers for callers without a HIP allocator ---- */
int vx_device_alloc(size_t bytes, void** ptr);
int vx_device_free(void* ptr);
int vx_memcpy_h2d(void* dst, const void* src, size_t bytes);
int vx_memcpy_d2h(void* dst, const void* src, size_t bytes);

const char* vx_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
