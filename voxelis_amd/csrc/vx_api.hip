// vx_api.hip -- host side of the C ABI declared in include/voxelis_hip.h.
//
// One vx_interner = the device node columns + pattern tables + a HIP stream + a
// grow-only workspace for launch sequences.  All work is enqueued on the
// interner's stream; synchronizing calls are documented in the header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <mutex>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/voxelis_hip.h"
#include "vx_device.h"
#include "vx_edits.hip"
#include "vx_kernels.hip"
#include "vx_mlp.hip"
#include "vx_microbench.hip"

using namespace vx;

// One call per voxel width: u8 (1), u16 (2) and four-byte voxels (i32/u32, kept as their
// bit pattern: the path only compares values for equality).
#define VX_WIDTH_DISPATCH(h, fn, ...)                                                 \
    ((h)->vsize == 1 ? fn<uint8_t>(__VA_ARGS__)                                      \
                     : (h)->vsize == 2 ? fn<uint16_t>(__VA_ARGS__) : fn<uint32_t>(__VA_ARGS__))

namespace {

constexpr int kBlock = 256;
constexpr int kGetTile = 256;  // lanes of a grouped-get block
constexpr uint64_t kMaxGrid = 8192;

// Tiles per probe block: candidates of sparse tiles are packed into full
// wavefronts, and the owners' allocation takes one update of the shared counter per
// block; at least `blocks` blocks stay so the launch still spreads over the chip.
inline uint32_t probe_group(uint32_t ntiles, uint32_t blocks) {
    if (blocks == 0) return 1;
    return std::max<uint32_t>(1u, std::min<uint32_t>(16u, ntiles / blocks));
}

inline unsigned grid_for(uint64_t work) {
    uint64_t g = (work + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > kMaxGrid) g = kMaxGrid;
    return (unsigned)g;
}

template <typename P>
bool dmalloc(P** p, size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) return false;
    *p = (P*)q;
    return true;
}

// Bumped whenever any workspace buffer moves: captured launch graphs hold its
// addresses and are rebuilt after a move.
static std::atomic<uint64_t> g_ws_gen{1};

struct Grow {
    void* ptr = nullptr;
    size_t cap = 0;
    bool ensure(size_t bytes, bool tracked = true) {
        if (bytes <= cap) return true;
        if (tracked) g_ws_gen.fetch_add(1);
        // queued kernels may still use the old buffer
        if (ptr) {
            (void)hipDeviceSynchronize();
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 4096);
        if (hipMalloc(&ptr, want) != hipSuccess) return false;
        cap = want;
        return true;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

struct LevelBufs {
    Grow res, key, hash, id, tslot, slot, tiles, old;
};

// Kernel ids of the per-kernel timer (vx_interner_profile_*).
enum KernelId : uint32_t {
    KID_BEGIN = 0, KID_PREPARE, KID_BRICK_SCAN, KID_LEAF_ALLOC, KID_BRICK_BUILD, KID_PROBE_BRICKS,
    KID_UNUSED, KID_LEVEL_BUILD, KID_PROBE_UPPER, KID_COMMIT, KID_FINALIZE,
    KID_END_ALLOC, KID_LEAF_CLEANUP, KID_GC, KID_GET, KID_CANON, KID_FILL, KID_ROLLBACK, KID_TO_VEC,
    KID_SET_DESCEND, KID_SET_LEVEL, KID_GET_GROUPED, KID_EDIT_KEYS, KID_SORT_HIST, KID_SORT_SCAN,
    KID_SORT_SCATTER, KID_EDIT_SCATTER, KID_COUNT
};
const char* const kKernelNames[KID_COUNT] = {
    "k_begin", "k_prepare", "k_brick_retry", "k_leaf_alloc", "k_brick_build", "k_probe[bricks]",
    "", "k_level_build", "k_probe[upper]", "k_commit", "k_finish",
    "k_end_alloc", "k_leaf_cleanup", "k_gc_round", "k_get", "k_canon", "k_fill", "k_rollback", "k_to_vec",
    "k_set_descend", "k_set_level", "k_get_grouped", "k_edit_keys", "k_sort_hist", "k_sort_scan",
    "k_sort_scatter", "k_edit_scatter"};

struct Prof {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<uint32_t, size_t>> pending;  // kernel id, index of its start event
    double ms[KID_COUNT] = {0};
    uint64_t launches[KID_COUNT] = {0};
};

// A launch sequence captured as a HIP graph, replayed while the call's arguments
// (and the workspace addresses it captured) are unchanged.
struct Plan {
    uint64_t key = 0;
    uint64_t ws_gen = 0;
    uint64_t last_use = 0;
    hipGraphExec_t exec = nullptr;
    Grow chunks, slot_off, tables;  // the call's records, owned by the plan
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        chunks.release();
        slot_off.release();
        tables.release();
    }
};

}  // namespace

struct vx_interner {
    uint32_t vsize = 0;
    uint64_t capacity = 0;
    uint64_t tslots = 0;
    uint64_t* children = nullptr;
    uint32_t* tmg = nullptr;  // (generation << 16) | types/mask
    void* values = nullptr;
    uint32_t* refs = nullptr;
    uint32_t* free_stack = nullptr;
    uint64_t* tab = nullptr;
    uint64_t* single = nullptr;  // u8/u16: the one-leaf key index (8 slots x every value)
    uint32_t* leaf_map = nullptr;
    uint64_t* leaf_ids = nullptr;  // u16: the brick pass's leaf id of every value
    uint64_t* ltab = nullptr;      // four-byte voxels: the hashed leaf table
    uint64_t lslots = 0;
    uint32_t* need_list = nullptr;
    uint8_t* leaf_need = nullptr;
    uint32_t* new_leaves = nullptr;
    Ctl* ctl = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int poisoned = 0;
    // workspace
    Grow chunks, slot_off, ptr_tables, gc_front[2], stage, scratch, deferred, drift, undo, owners;
    Grow set_rec, set_meta;  // single-voxel set: the recorded path of every tree
    // edit-list path: zero-kept dense Batch buffers, sort buffers, tree records
    Grow edit_ws, edit_keys[2], edit_order[2], edit_tmp, edit_trees, edit_first, edit_status;
    LevelBufs lv[kLevels + 1];
    // pinned staging ring: a slot is reused only after its copies completed
    static constexpr int kRing = 4;
    void* h_pinned_ring[kRing] = {nullptr};
    size_t h_pinned_cap[kRing] = {0};
    hipEvent_t ring_ev[kRing] = {nullptr};
    bool ring_used[kRing] = {false};
    int ring_next = 0;
    int ring_cur = 0;
    void* h_pinned = nullptr;  // the slot of the current call
    Prof prof;
    vx_tuning tune = {256, 8, 1024, 1024, 32768, 512, 4, 0};
    // captured launch sequences of repeated calls (vx_interner_set_graphs)
    bool graphs = false;
    hipStream_t cap_stream = nullptr;
    static constexpr int kPlans = 256;
    Plan plans[kPlans];
    uint64_t plan_clock = 0;
};

namespace {

template <typename T>
Pool<T> pool_of(vx_interner* h) {
    Pool<T> P;
    P.children = h->children;
    P.tmg = h->tmg;
    P.values = (T*)h->values;
    P.refs = h->refs;
    P.free_stack = h->free_stack;
    P.tab = h->tab;
    P.single = h->single;
    P.tmask = h->tslots - 1;
    P.leaf_map = h->leaf_map;
    P.leaf_ids = h->leaf_ids;
    P.ltab = h->ltab;
    P.lmask = h->lslots ? h->lslots - 1 : 0;
    P.need_list = h->need_list;
    P.need_cap = (uint32_t)std::min<uint64_t>(h->lslots, 0xFFFFFFFFull);
    P.leaf_need = h->leaf_need;
    P.new_leaves = h->new_leaves;
    P.ctl = h->ctl;
    P.capacity = (uint32_t)h->capacity;
    P.dry = 0;
    P.alloc_limit = 0xFFFFFFFFu;
    P.per_tree = nullptr;
    P.levels = nullptr;
    P.undo = nullptr;
    P.undo_cap = 0;
    P.dedup_rounds = h->tune.dedup_rounds;
    P.brick_merge_rounds = h->tune.brick_merge_rounds;
    return P;
}

// Take the next staging slot of the ring, waiting for its previous copies.
bool pinned_ensure(vx_interner* h, size_t bytes) {
    int i = h->ring_next;
    h->ring_next = (i + 1) % vx_interner::kRing;
    if (!h->ring_ev[i] && hipEventCreateWithFlags(&h->ring_ev[i], hipEventDisableTiming) != hipSuccess)
        return false;
    if (h->ring_used[i] && hipEventSynchronize(h->ring_ev[i]) != hipSuccess) return false;
    h->ring_used[i] = false;
    if (bytes > h->h_pinned_cap[i]) {
        if (h->h_pinned_ring[i]) (void)hipHostFree(h->h_pinned_ring[i]);
        h->h_pinned_ring[i] = nullptr;
        h->h_pinned_cap[i] = 0;
        size_t want = std::max<size_t>(bytes, 1 << 16);
        if (hipHostMalloc(&h->h_pinned_ring[i], want, hipHostMallocDefault) != hipSuccess) return false;
        h->h_pinned_cap[i] = want;
    }
    h->h_pinned = h->h_pinned_ring[i];
    h->ring_cur = i;
    return true;
}

// Mark the current staging slot busy until the stream passes this point.
void pinned_release(vx_interner* h) {
    int i = h->ring_cur;
    if (hipEventRecord(h->ring_ev[i], h->stream) == hipSuccess) h->ring_used[i] = true;
}

int err_to_status(uint32_t err) {
    if (err & kErrOom) return VX_E_OOM;
    if (err & kErrTableFull) return VX_E_TABLE_FULL;
    if (err & kErrRange) return VX_E_RANGE;
    if (err & kErrPoison) return VX_E_POISONED;
    return VX_OK;
}

#define HIPCHK(x)                        \
    do {                                 \
        if ((x) != hipSuccess) return VX_E_HIP; \
    } while (0)

uint64_t slots_at(uint32_t d, uint32_t h) { return 1ull << (3 * (d - h)); }

// vx_batch_view.reserved value used internally for trees that are not applied
constexpr uint8_t kViewSkip = 0x5A;

hipEvent_t prof_event(vx_interner* h) {
    Prof& p = h->prof;
    if (p.used == p.pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        p.pool.push_back(e);
    }
    return p.pool[p.used++];
}

#define VX_COMMA ,
// Launch a kernel on the interner's stream, bracketed by timing events when on.
#define VX_LAUNCH(h, kid, kern, grid, block, ...)                                   \
    do {                                                                            \
        size_t ev0__ = 0;                                                           \
        bool prof__ = (h)->prof.on;                                                 \
        if (prof__) {                                                               \
            ev0__ = (h)->prof.used;                                                 \
            hipEvent_t e0__ = prof_event(h);                                        \
            hipEvent_t e1__ = prof_event(h);                                        \
            prof__ = e0__ && e1__ && hipEventRecord(e0__, (h)->stream) == hipSuccess; \
        }                                                                           \
        kern<<<(grid), (block), 0, (h)->stream>>>(__VA_ARGS__);                     \
        if (prof__) {                                                               \
            (void)hipEventRecord((h)->prof.pool[ev0__ + 1], (h)->stream);           \
            (h)->prof.pending.push_back({(uint32_t)(kid), ev0__});                  \
        }                                                                           \
    } while (0)

// Options of one apply (vx_apply_opts, validated by the caller).
struct ApplyOpts {
    bool dry = false;
    uint32_t alloc_limit = 0xFFFFFFFFu;
    uint32_t* counts = nullptr;  // device: new branches, new leaves, then per tree
    // single-voxel set mode (vx_trees_set): one voxel per tree, device xyz[3n] and values[n]
    const int32_t* set_xyz = nullptr;
    const void* set_values = nullptr;
};

template <typename T>
int rebuild_step(vx_interner* h, bool force);

template <typename T, int TILE>
int apply_tiled(vx_interner* h, size_t n, uint64_t* roots, const vx_batch_view* batches,
                int32_t* status, const int32_t* pre_status, const ApplyOpts& opt) {
    constexpr int kBR = brick_rounds<TILE>();
    hipStream_t s = h->stream;
    Pool<T> P = pool_of<T>(h);
    P.dry = opt.dry ? 1u : 0u;
    P.alloc_limit = opt.alloc_limit;
    P.per_tree = opt.counts ? opt.counts + 2 : nullptr;
    // single-voxel set mode: one slot per tree at every height (the voxel's path)
    const bool setm = opt.set_xyz != nullptr;
    // ---- a repeated call replays its captured launch sequence
    Plan* plan = nullptr;
    uint64_t key = 0;
    if (h->graphs && !h->prof.on && !pre_status && !setm) {
        key = mix64(0x5650u ^ (uint64_t)n ^ ((uint64_t)sizeof(T) << 40));
        const uint64_t head[] = {(uint64_t)roots, (uint64_t)status, (uint64_t)opt.dry,
                                 (uint64_t)opt.alloc_limit, (uint64_t)opt.counts, (uint64_t)TILE,
                                 (uint64_t)h->tune.dedup_rounds, (uint64_t)h->tune.commit_grid,
                                 (uint64_t)h->tune.gc_grid, (uint64_t)h->tune.probe_blocks_bricks,
                                 (uint64_t)h->tune.probe_blocks, (uint64_t)h->tune.brick_merge_rounds,
                                 (uint64_t)h->tune.brick_rounds};
        for (uint64_t w : head) key = mix64(key ^ w);
        for (size_t i = 0; i < n; i++) {
            const vx_batch_view& b = batches[i];
            key = mix64(key ^ (uint64_t)b.masks);
            key = mix64(key ^ (uint64_t)b.values);
            key = mix64(key ^ b.fill_value);
            key = mix64(key ^ ((uint64_t)b.depth | (uint64_t)b.has_fill << 8 | (uint64_t)b.has_patches << 16 |
                               (uint64_t)b.reserved << 24));
        }
        uint64_t gen = g_ws_gen.load();
        for (Plan& pl : h->plans) {
            if (pl.exec && pl.key == key && pl.ws_gen == gen) {
                pl.last_use = ++h->plan_clock;
                return hipGraphLaunch(pl.exec, s) == hipSuccess ? VX_OK : VX_E_HIP;
            }
        }
        plan = &h->plans[0];
        for (Plan& pl : h->plans)
            if (pl.last_use < plan->last_use) plan = &pl;
        if (plan->exec && hipStreamSynchronize(s) != hipSuccess) return VX_E_HIP;  // it may be running
        plan->release();
    }
    // ---- plan: trees with patches grouped by depth (others first), dense slot
    // spaces per height
    auto status_of = [&](const vx_batch_view& b) -> int32_t {
        if (b.depth > kMaxDepth) return VX_E_DEPTH;
        if (setm) return b.depth >= 1 ? 0 : VX_E_DEPTH;
        // voxtree.rs:792/913: max_depth - 1 underflows for a patched batch of depth < 2
        if (b.has_patches && b.depth < 2) return VX_E_DEPTH;
        if (b.has_patches && (!b.masks || !b.values)) return VX_E_INVALID_ARG;
        return 0;
    };
    auto sort_key = [&](uint32_t i) -> uint32_t {
        const vx_batch_view& b = batches[i];
        return (b.has_patches && status_of(b) == 0 && b.reserved != kViewSkip) ? b.depth : 0u;
    };
    std::vector<uint32_t> perm(n);
    for (size_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return sort_key(a) < sort_key(b); });
    size_t chunk_bytes = n * sizeof(ChunkDev);
    size_t off_count = (size_t)(kLevels + 1) * n;
    size_t tab_bytes = 2 * (kLevels + 1) * sizeof(void*) + (kLevels + 1) * sizeof(Cands);
    size_t stage_bytes = chunk_bytes + off_count * sizeof(uint64_t) + tab_bytes;
    if (!pinned_ensure(h, stage_bytes)) return VX_E_HIP;
    ChunkDev* hc = (ChunkDev*)h->h_pinned;
    uint64_t* hoff = (uint64_t*)((char*)h->h_pinned + chunk_bytes);
    uint32_t dmax = 0, dall = 0;  // deepest patched tree, deepest tree
    struct Run { uint32_t start, count, depth; };
    std::vector<Run> runs;
    for (size_t c = 0; c < n; c++) {
        const vx_batch_view& b = batches[perm[c]];
        ChunkDev& C = hc[c];
        std::memset(&C, 0, sizeof(C));
        C.masks = b.masks;
        C.values = b.values;
        C.fill = b.fill_value & (sizeof(T) == 1 ? 0xFFull : sizeof(T) == 2 ? 0xFFFFull : 0xFFFFFFFFull);
        C.orig = perm[c];
        C.depth = b.depth;
        C.has_fill = b.has_fill ? 1 : 0;
        C.skip = b.reserved == kViewSkip ? 1 : 0;
        C.status = C.status0 = status_of(b);
        C.has_patches = (b.has_patches && C.status == 0 && !C.skip) ? 1 : 0;
        if (C.status == 0 && !C.skip) dall = std::max<uint32_t>(dall, b.depth);
        if (!C.has_patches) continue;
        uint32_t d = b.depth;
        dmax = std::max<uint32_t>(dmax, d);
        if (!runs.empty() && runs.back().depth == d && runs.back().start + runs.back().count == c)
            runs.back().count++;
        else
            runs.push_back({(uint32_t)c, 1, d});
    }
    for (size_t i = 0; i < off_count; i++) hoff[i] = ~0ull;
    LevelInfo L[kLevels + 1];
    std::memset(L, 0, sizeof(L));
    std::vector<Run> lruns[kLevels + 1];
    for (uint32_t hh = 1; hh <= dmax; hh++) {
        uint64_t base = 0;
        for (const Run& r : runs) {
            if (r.depth < hh) continue;
            lruns[hh].push_back(r);
        }
        L[hh].h = hh;
        L[hh].ngroups = (uint32_t)lruns[hh].size();
        for (size_t g = 0; g < lruns[hh].size(); g++) {
            const Run& r = lruns[hh][g];
            L[hh].slot_base[g] = base;
            L[hh].chunk_base[g] = r.start;
            L[hh].depth[g] = setm ? hh : r.depth;  // set mode: slot = tree, shift 0
            uint64_t per = setm ? 1 : slots_at(r.depth, hh);
            for (uint32_t k = 0; k < r.count; k++) hoff[(size_t)hh * n + r.start + k] = base + k * per;
            base += per * r.count;
        }
        L[hh].slot_base[L[hh].ngroups] = base;
        L[hh].nslots = base;
    }
    // each group's first slot one height below and one above (a group of height hh is
    // a run of trees of depth >= hh; runs keep their order at every height)
    for (uint32_t hh = 1; hh <= dmax; hh++) {
        for (size_t g = 0; g < lruns[hh].size(); g++) {
            const Run& r = lruns[hh][g];
            L[hh].prev_base[g] = ~0ull;
            L[hh].next_base[g] = ~0ull;
            if (hh >= 2)
                for (size_t g2 = 0; g2 < lruns[hh - 1].size(); g2++)
                    if (lruns[hh - 1][g2].start == r.start) L[hh].prev_base[g] = L[hh - 1].slot_base[g2];
            if (hh + 1 <= dmax)
                for (size_t g2 = 0; g2 < lruns[hh + 1].size(); g2++)
                    if (lruns[hh + 1][g2].start == r.start) L[hh].next_base[g] = L[hh + 1].slot_base[g2];
        }
    }
    // ---- workspace (a captured call owns its records)
    Grow& g_chunks = plan ? plan->chunks : h->chunks;
    Grow& g_off = plan ? plan->slot_off : h->slot_off;
    Grow& g_tables = plan ? plan->tables : h->ptr_tables;
    if (!g_chunks.ensure(chunk_bytes, !plan)) return VX_E_HIP;
    if (!g_off.ensure(off_count * sizeof(uint64_t), !plan)) return VX_E_HIP;
    uint64_t* res_ptrs_h[kLevels + 1] = {nullptr};
    uint64_t* cid_ptrs_h[kLevels + 1] = {nullptr};
    Cands K[kLevels + 1];
    std::memset(K, 0, sizeof(K));
    uint32_t* tiles_h[kLevels + 1] = {nullptr};
    uint64_t ntiles[kLevels + 1] = {0};
    int br = 1;  // bricks per lane of the brick pass
    OldNodes O;
    std::memset(&O, 0, sizeof(O));
    for (uint32_t hh = 1; hh <= dmax; hh++) {
        ntiles[hh] = std::max<uint64_t>((L[hh].nslots + TILE - 1) / TILE, 1);
        // the brick pass takes R tiles per block and writes every tile's count
        if (hh == 1 && !setm) {
            // several bricks per lane once the pass has blocks to spare for every
            // compute unit; a small apply keeps one per lane (more blocks, shorter chains)
            br = h->tune.brick_rounds == 1 ? 1
                 : h->tune.brick_rounds == 2 ? 2
                 : h->tune.brick_rounds != 0 ? kBR
                 : L[1].nslots >= (uint64_t)kBR * TILE * 512 ? kBR : 1;
            ntiles[hh] = (ntiles[hh] + br - 1) / br * br;
        }
        uint64_t ns = ntiles[hh] * TILE;
        // slot and candidate indices travel as 32-bit values (pending results, the
        // owner list, deferred bricks)
        if (ns > 0xFFFFFFFFull) return VX_E_INVALID_ARG;
        LevelBufs& B = h->lv[hh];
        if (!B.res.ensure(ns * 8) || !B.key.ensure(ns * 72) || !B.hash.ensure(ns * 8) ||
            !B.id.ensure(ns * 8) || !B.tslot.ensure(ns * 8) || !B.slot.ensure(ns * 8) ||
            !B.tiles.ensure(ntiles[hh] * 4) || (hh >= 2 && !B.old.ensure(ns * 8)))
            return VX_E_HIP;
        if (hh >= 2) O.at[hh] = (uint64_t*)B.old.ptr;
        tiles_h[hh] = (uint32_t*)B.tiles.ptr;
        K[hh].key = (uint64_t*)B.key.ptr;
        K[hh].hash = (uint64_t*)B.hash.ptr;
        K[hh].id = (uint64_t*)B.id.ptr;
        K[hh].tslot = (uint64_t*)B.tslot.ptr;
        K[hh].slot = (uint64_t*)B.slot.ptr;
        res_ptrs_h[hh] = (uint64_t*)B.res.ptr;
        cid_ptrs_h[hh] = (uint64_t*)B.id.ptr;
    }
    // the owners of every level, by allocation rank (at most one per candidate)
    uint64_t nown = 1;
    for (uint32_t hh = 1; hh <= dmax; hh++) nown += ntiles[hh] * TILE;
    if (!h->owners.ensure(nown * 8)) return VX_E_HIP;
    P.owners = (uint64_t*)h->owners.ptr;
    void** hptrs = (void**)((char*)h->h_pinned + chunk_bytes + off_count * sizeof(uint64_t));
    for (int i = 0; i <= kLevels; i++) {
        hptrs[i] = res_ptrs_h[i];
        hptrs[kLevels + 1 + i] = cid_ptrs_h[i];
    }
    std::memcpy(hptrs + 2 * (kLevels + 1), K, (kLevels + 1) * sizeof(Cands));
    if (!g_tables.ensure(tab_bytes, !plan)) return VX_E_HIP;
    size_t gcap = (size_t)h->capacity + n + 16;
    if (!h->gc_front[0].ensure(gcap * 4) || !h->gc_front[1].ensure(gcap * 4)) return VX_E_HIP;
    ChunkDev* dchunks = (ChunkDev*)g_chunks.ptr;
    uint64_t* doff = (uint64_t*)g_off.ptr;
    uint64_t* const* dres = (uint64_t* const*)g_tables.ptr;
    uint64_t* const* dcid = dres + (kLevels + 1);
    HIPCHK(hipMemcpyAsync(dchunks, hc, chunk_bytes, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(doff, hoff, off_count * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_tables.ptr, hptrs, tab_bytes, hipMemcpyHostToDevice, s));
    P.levels = (const Cands*)((void**)g_tables.ptr + 2 * (kLevels + 1));
    pinned_release(h);
    uint32_t nvalues = sizeof(T) <= 2 ? 1u << (8 * sizeof(T)) : 0u;  // four-byte voxels: the request list
    uint64_t nb = std::max<uint64_t>(dmax >= 2 ? L[1].nslots : 1, 1);
    if (!h->deferred.ensure(nb * 8) || !h->drift.ensure(nb * sizeof(Drift))) return VX_E_HIP;
    uint64_t* ddef = (uint64_t*)h->deferred.ptr;
    Drift* ddrift = (Drift*)h->drift.ptr;
    // undo log: one entry per deferred brick, collapsed upper node or fill leaf at most
    uint64_t undo_cap = n;
    LevelTiles LT;
    std::memset(&LT, 0, sizeof(LT));
    LT.dmax = (dmax >= 2 || (setm && dmax >= 1)) ? dmax : 0;
    for (uint32_t hh = 1; hh <= LT.dmax; hh++) {
        undo_cap += L[hh].nslots;
        LT.cnt[hh] = tiles_h[hh];
        LT.start[hh + 1] = LT.start[hh] + ntiles[hh];
    }
    for (uint32_t hh = LT.dmax + 1; hh <= (uint32_t)kLevels; hh++) LT.start[hh + 1] = LT.start[hh];
    if (undo_cap > 0xFFFFFFFFull) return VX_E_INVALID_ARG;
    if (opt.dry && (uint64_t)h->capacity + 65536 + (undo_cap - n) > 0xFFFFFFFFull) return VX_E_INVALID_ARG;
    // a dry run's missing leaf is capacity + its key (the value, or the request's slot)
    if (opt.dry && (uint64_t)h->capacity + (sizeof(T) <= 2 ? 65536ull : h->lslots) > 0xFFFFFFFFull)
        return VX_E_INVALID_ARG;
    if (!h->undo.ensure(std::max<uint64_t>(undo_cap, 1) * sizeof(Drift))) return VX_E_HIP;
    P.undo = (Drift*)h->undo.ptr;
    P.undo_cap = (uint32_t)undo_cap;
    // ---- launch sequence
    // set mode: the recorded paths
    uint64_t* drec = nullptr;
    uint32_t* dmeta = nullptr;
    if (setm) {
        if (!h->set_rec.ensure(std::max<size_t>(n, 1) * kSetRec * 8) || !h->set_meta.ensure(std::max<size_t>(n, 1) * 4))
            return VX_E_HIP;
        drec = (uint64_t*)h->set_rec.ptr;
        dmeta = (uint32_t*)h->set_meta.ptr;
    }
    auto launch_set = [&]() -> int {
    const int32_t* sx = opt.set_xyz;
    const T* sv = (const T*)opt.set_values;
    VX_LAUNCH(h, KID_SET_DESCEND, k_set_descend<T>, grid_for(n), kBlock, P, dchunks, (uint32_t)n, roots, sx, sv,
              drec, dmeta);
    VX_LAUNCH(h, KID_LEAF_ALLOC, k_leaf_alloc<T>, 1, 1024, P, dchunks, (uint32_t)n, nvalues);
    const unsigned wpb = TILE / 64;
    for (uint32_t hh = 1; hh <= dmax; hh++) {
        unsigned th = (unsigned)ntiles[hh];
        VX_LAUNCH(h, KID_SET_LEVEL, k_set_level<T VX_COMMA TILE>, th, TILE, P, dchunks, L[hh], doff,
                  (uint32_t)n, sx, sv, drec, dmeta, res_ptrs_h[hh - 1], cid_ptrs_h[hh - 1], res_ptrs_h[hh], K[hh],
                  tiles_h[hh]);
        uint32_t gh = probe_group(th, h->tune.probe_blocks);
        VX_LAUNCH(h, KID_PROBE_UPPER, k_probe<T VX_COMMA TILE>, (th + gh - 1) / gh, TILE, P, K[hh], tiles_h[hh],
                  res_ptrs_h[hh], dchunks, L[hh], th, gh);
    }
    (void)wpb;
    if (!opt.dry && dmax >= 1)
        VX_LAUNCH(h, KID_COMMIT, k_commit<T VX_COMMA TILE>, h->tune.commit_grid, TILE, P);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
    };
    auto launch_all = [&]() -> int {
    if (!opt.dry) {
        int rc = rebuild_step<T>(h, false);
        if (rc) return rc;
    }
    if (setm) {
        int rc = launch_set();
        if (rc) return rc;
    } else {
    // chunk set-up, plus the old-tree descent over the level-3 slots in extra blocks
    unsigned np = grid_for(n);
    unsigned nd = L[3].nslots ? grid_for(L[3].nslots) : 0u;
    unsigned nl = P.leaf_ids && dmax >= 2 ? 65536u / kBlock : 0u;
    VX_LAUNCH(h, KID_PREPARE, k_prepare<T>, np + nd + nl, kBlock, P, dchunks, (uint32_t)n, roots, pre_status, np,
              nd, L[3], doff, O);
    unsigned t1 = (unsigned)ntiles[1];
    if (dmax >= 2 && br == kBR)
        VX_LAUNCH(h, KID_BRICK_BUILD, k_brick_build<T VX_COMMA TILE VX_COMMA kBR>, t1 / kBR, TILE, P, dchunks, L[1],
                  res_ptrs_h[1], K[1], tiles_h[1], ddef, ddrift, O.at[2]);
    else if (dmax >= 2 && br == 2)
        VX_LAUNCH(h, KID_BRICK_BUILD, k_brick_build<T VX_COMMA TILE VX_COMMA 2>, t1 / 2, TILE, P, dchunks, L[1],
                  res_ptrs_h[1], K[1], tiles_h[1], ddef, ddrift, O.at[2]);
    else if (dmax >= 2)
        VX_LAUNCH(h, KID_BRICK_BUILD, k_brick_build<T VX_COMMA TILE VX_COMMA 1>, t1, TILE, P, dchunks, L[1],
                  res_ptrs_h[1], K[1], tiles_h[1], ddef, ddrift, O.at[2]);
    VX_LAUNCH(h, KID_LEAF_ALLOC, k_leaf_alloc<T>, 1, 1024, P, dchunks, (uint32_t)n, nvalues);
    if (dmax >= 2) {
        VX_LAUNCH(h, KID_BRICK_SCAN, k_brick_retry<T VX_COMMA TILE>, 1024, kBlock, P, dchunks, L[1], res_ptrs_h[1], K[1],
                  tiles_h[1], ddef, ddrift);
        uint32_t g1 = probe_group(t1, h->tune.probe_blocks_bricks);
        VX_LAUNCH(h, KID_PROBE_BRICKS, k_probe_bricks<T VX_COMMA TILE>, (t1 + g1 - 1) / g1, TILE, P, K[1], tiles_h[1],
                  res_ptrs_h[1], dchunks, L[1], t1, g1);
        for (uint32_t hh = 2; hh <= dmax; hh++) {
            unsigned th = (unsigned)ntiles[hh];
            VX_LAUNCH(h, KID_LEVEL_BUILD, k_level_build<T VX_COMMA TILE>, th, TILE, P, dchunks, L[hh],
                      res_ptrs_h[hh - 1], cid_ptrs_h[hh - 1], res_ptrs_h[hh], K[hh], tiles_h[hh], O.at[hh]);
            uint32_t gh = probe_group(th, h->tune.probe_blocks);
            VX_LAUNCH(h, KID_PROBE_UPPER, k_probe<T VX_COMMA TILE>, (th + gh - 1) / gh, TILE, P, K[hh], tiles_h[hh],
                      res_ptrs_h[hh], dchunks, L[hh], th, gh);
        }
        // every level's new nodes, written once no probe runs any more
        if (!opt.dry) VX_LAUNCH(h, KID_COMMIT, k_commit<T VX_COMMA TILE>, h->tune.commit_grid, TILE, P);
    }
    }
    uint32_t* gf0 = (uint32_t*)h->gc_front[0].ptr;
    uint32_t* gf1 = (uint32_t*)h->gc_front[1].ptr;
    // the rollback (dry runs restore the table; an aborted apply takes its references
    // back) and the root swap
    unsigned rg = (unsigned)std::min<uint64_t>(std::max<uint64_t>(LT.start[LT.dmax + 1], 1), 1024);
    unsigned fg = std::max(rg, (unsigned)((n + TILE - 1) / TILE));
    VX_LAUNCH(h, KID_FINALIZE, k_finish<T VX_COMMA TILE>, std::min(fg, 65535u), TILE, P, dchunks, LT, ddrift,
              (uint32_t)n, doff, (uint32_t)n, dres, dcid, roots, status, gf0);
    VX_LAUNCH(h, KID_END_ALLOC, k_end_alloc<T>, 1, 256, P, true, opt.counts);
    if (!opt.dry) {
        // dec_ref_recursive of the replaced roots: a tree of depth D has nodes at
        // heights D..0, one round each; the first round also releases unused new leaves.
        // A small front is split over the blocks in shares of kGcShareMin nodes and
        // followed locally: the first round takes one block per share of the replaced
        // roots (every block of that round also runs the leaf release), the later ones
        // the GC grid
        static_assert(kBlock == 256, "k_gc_round follows up to 256 nodes per pass");
        const unsigned g0 = std::max(grid_for(n), std::min<unsigned>(h->tune.gc_grid,
                                                                    (unsigned)((n + kGcShareMin - 1) / kGcShareMin)));
        for (uint32_t r = 0; r <= dall; r++) {
            uint32_t* fr = (r & 1) ? gf1 : gf0;
            uint32_t* nx = (r & 1) ? gf0 : gf1;
            unsigned grid = r == 0 ? g0 : h->tune.gc_grid;
            VX_LAUNCH(h, KID_GC, k_gc_round<T>, grid, kBlock, P, r, fr, nx, r == 0, (uint64_t)gcap, r == 0);
        }
    }
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
    };
    if (!plan) return launch_all();
    // capture on a private stream, then replay on the interner's stream
    if (!h->cap_stream && hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking) != hipSuccess)
        return VX_E_HIP;
    uint64_t gen = g_ws_gen.load();
    hipStream_t real = h->stream;
    h->stream = h->cap_stream;
    if (hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        h->stream = real;
        return VX_E_HIP;
    }
    int rc = launch_all();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(h->cap_stream, &graph);
    h->stream = real;
    if (rc != VX_OK || e != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return VX_E_HIP;
    }
    e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        plan->exec = nullptr;
        return VX_E_HIP;
    }
    plan->key = key;
    plan->ws_gen = gen;
    plan->last_use = ++h->plan_clock;
    return hipGraphLaunch(plan->exec, s) == hipSuccess ? VX_OK : VX_E_HIP;
}

// The table-rebuild step that starts every write call (enqueued on the interner's
// stream, captured with the call's launch sequence): the device decides whether the
// pattern table (and the i32 leaf table) is rebuilt now, and the clear and re-insert
// kernels do it or return at once.  force: rebuild regardless.
constexpr unsigned kRebuildGrid = 512;
template <typename T>
int rebuild_step(vx_interner* h, bool force) {
    Pool<T> P = pool_of<T>(h);
    hipStream_t s = h->stream;
    k_tab_clear<<<kRebuildGrid, kBlock, 0, s>>>(h->tab, h->tslots, h->ctl, h->ltab ? h->lslots : 0ull,
                                                force ? 1u : 0u);
    k_tab_reinsert<T><<<kRebuildGrid, kBlock, 0, s>>>(P);
    if (h->ltab) {
        k_ltab_clear<<<kRebuildGrid, kBlock, 0, s>>>(h->ltab, h->lslots, h->ctl);
        k_ltab_reinsert<T><<<kRebuildGrid, kBlock, 0, s>>>(P);
    }
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

// Tombstones in the pattern table: a counter striped over the statistics lines.
uint64_t tombstones_of(const Ctl& c, int which = ST_TOMB) {
    uint64_t t = 0;
    for (int i = 0; i < kStatStripes; i++) t += c.st[i][which];
    return t;
}

template <typename T>
int apply_impl(vx_interner* h, size_t n, uint64_t* roots, const vx_batch_view* batches,
               int32_t* status, const int32_t* pre_status = nullptr, const ApplyOpts& opt = ApplyOpts()) {
    int rc = VX_OK;
    switch (h->tune.tile) {
        case 128: rc = apply_tiled<T, 128>(h, n, roots, batches, status, pre_status, opt); break;
        case 512: rc = apply_tiled<T, 512>(h, n, roots, batches, status, pre_status, opt); break;
        default: rc = apply_tiled<T, 256>(h, n, roots, batches, status, pre_status, opt); break;
    }
    return rc;
}

template <typename T>
int apply_edits_impl(vx_interner* h, size_t n, uint64_t* roots, const uint8_t* depths,
                     const uint64_t* offsets, const int32_t* xyz, const void* values, int32_t* status,
                     const ApplyOpts& opt = ApplyOpts()) {
    hipStream_t s = h->stream;
    uint64_t E = offsets[n];
    std::vector<EditTree> trees(n);
    std::vector<uint64_t> first(n);
    uint64_t mbytes = 0, vcount = 0;
    for (size_t t = 0; t < n; t++) {
        uint32_t d = depths[t];
        if (offsets[t + 1] < offsets[t]) return VX_E_INVALID_ARG;
        uint64_t P = d > 0 ? (1ull << (3 * (d - 1))) : 1ull;
        trees[t].first = offsets[t];
        trees[t].depth = d;
        trees[t].status = 0;
        bool has = offsets[t + 1] > offsets[t];
        if (d > (uint32_t)kMaxDepth || (has && d < 2)) trees[t].status = VX_E_DEPTH;
        trees[t].mbase = mbytes;
        trees[t].vbase = vcount;
        first[t] = offsets[t];
        if (d <= (uint32_t)kMaxDepth) {
            mbytes += ((2 * P + 255) / 256) * 256;
            vcount += ((8 * P + 255) / 256) * 256;
        }
    }
    size_t ws_bytes = mbytes + vcount * sizeof(T);
    if (ws_bytes > h->edit_ws.cap) {
        if (!h->edit_ws.ensure(ws_bytes)) return VX_E_HIP;
        HIPCHK(hipMemsetAsync(h->edit_ws.ptr, 0, h->edit_ws.cap, s));  // kept zero from here on
    }
    uint8_t* mws = (uint8_t*)h->edit_ws.ptr;
    T* vws = (T*)(mws + mbytes);
    uint64_t En = std::max<uint64_t>(E, 1);
    if (!h->edit_keys[0].ensure(En * 8) || !h->edit_keys[1].ensure(En * 8) ||
        !h->edit_order[0].ensure(En * 4) || !h->edit_order[1].ensure(En * 4) ||
        !h->edit_trees.ensure(n * sizeof(EditTree)) || !h->edit_first.ensure(n * 8) ||
        !h->edit_status.ensure(n * 4))
        return VX_E_HIP;
    EditTree* dtrees = (EditTree*)h->edit_trees.ptr;
    uint64_t* dfirst = (uint64_t*)h->edit_first.ptr;
    int32_t* dpre = (int32_t*)h->edit_status.ptr;
    // the tree records travel through a pinned staging slot, reused only once these
    // copies completed (no wait for earlier work on the stream)
    if (!pinned_ensure(h, n * (sizeof(EditTree) + 8))) return VX_E_HIP;
    std::memcpy(h->h_pinned, trees.data(), n * sizeof(EditTree));
    std::memcpy((char*)h->h_pinned + n * sizeof(EditTree), first.data(), n * 8);
    HIPCHK(hipMemcpyAsync(dtrees, h->h_pinned, n * sizeof(EditTree), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dfirst, (char*)h->h_pinned + n * sizeof(EditTree), n * 8, hipMemcpyHostToDevice, s));
    pinned_release(h);
    uint64_t* k0 = (uint64_t*)h->edit_keys[0].ptr;
    uint64_t* k1 = (uint64_t*)h->edit_keys[1].ptr;
    uint32_t* o0 = (uint32_t*)h->edit_order[0].ptr;
    uint32_t* o1 = (uint32_t*)h->edit_order[1].ptr;
    unsigned tree_bits = 1;
    while ((1ull << tree_bits) <= n) tree_bits++;
    unsigned end_bit = kPathBits + tree_bits;
    if (E > 0) {
        VX_LAUNCH(h, KID_EDIT_KEYS, k_edit_keys, grid_for(E), kBlock, E, dfirst, (uint32_t)n, dtrees, xyz, k0, o0);
        // stable LSD radix sort, 8 bits per pass, ping-ponging between the two buffers
        const unsigned G = (unsigned)std::min<uint64_t>(2048, (E + kSortBlock - 1) / kSortBlock);
        const uint64_t per = (E + G - 1) / G;
        const uint32_t M = 256u * G;  // digit counts of a pass, digit-major
        const uint32_t ntile = (M + kScanTileN - 1) / kScanTileN;
        if (!h->edit_tmp.ensure((size_t)(M + ntile + 1) * 4)) return VX_E_HIP;
        uint32_t* hist = (uint32_t*)h->edit_tmp.ptr;
        uint32_t* tile_off = hist + M;
        for (unsigned shift = 0; shift < end_bit; shift += 8) {
            VX_LAUNCH(h, KID_SORT_HIST, k_sort_hist, G, kSortBlock, k0, E, per, shift, hist);
            VX_LAUNCH(h, KID_SORT_SCAN, k_sort_scan_tiles, ntile, kSortScanBlock, hist, M, tile_off);
            VX_LAUNCH(h, KID_SORT_SCAN, k_sort_scan, 1, kSortScanBlock, tile_off, ntile);
            VX_LAUNCH(h, KID_SORT_SCATTER, k_sort_scatter, G, kSortBlock, k0, o0, k1, o1, E, per, shift, hist,
                      tile_off);
            std::swap(k0, k1);
            std::swap(o0, o1);
        }
        std::swap(k0, k1);  // the sorted pairs are in (k1, o1)
        std::swap(o0, o1);
        VX_LAUNCH(h, KID_EDIT_SCATTER, k_edit_scatter<T>, grid_for(E), kBlock, E, (uint32_t)n, dtrees, k1, o1,
                  (const T*)values, mws, vws, false);
    }
    // per-tree failures recorded by k_edit_keys go in as pre-statuses
    k_edit_status<<<grid_for(n), kBlock, 0, s>>>((uint32_t)n, dtrees, dpre);
    std::vector<vx_batch_view> views(n);
    for (size_t t = 0; t < n; t++) {
        vx_batch_view& v = views[t];
        std::memset(&v, 0, sizeof(v));
        v.depth = depths[t];
        v.has_patches = offsets[t + 1] > offsets[t] ? 1 : 0;
        if (!v.has_patches) v.reserved = kViewSkip;
        v.masks = mws + trees[t].mbase;
        v.values = vws + trees[t].vbase;
    }
    int rc = apply_impl<T>(h, n, roots, views.data(), status, dpre, opt);
    // the touched bricks are cleared again on every path, a failed apply included, so
    // the workspace stays zero for the next call
    if (E > 0)
        VX_LAUNCH(h, KID_EDIT_SCATTER, k_edit_scatter<T>, grid_for(E), kBlock, E, (uint32_t)n, dtrees, k1, o1,
                  (const T*)values, mws, vws, true);
    if (rc) return rc;
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

template <typename T>
int fill_impl(vx_interner* h, size_t n, uint64_t* roots, const uint64_t* vals) {
    int rc = rebuild_step<T>(h, false);
    if (rc) return rc;
    Pool<T> P = pool_of<T>(h);
    size_t gcap = (size_t)h->capacity + n + 16;
    if (!h->gc_front[0].ensure(gcap * 4) || !h->gc_front[1].ensure(gcap * 4)) return VX_E_HIP;
    uint32_t* gf0 = (uint32_t*)h->gc_front[0].ptr;
    uint32_t* gf1 = (uint32_t*)h->gc_front[1].ptr;
    uint32_t nvalues = sizeof(T) <= 2 ? 1u << (8 * sizeof(T)) : 0u;
    VX_LAUNCH(h, KID_BEGIN, k_begin, 1, 1, h->ctl);
    VX_LAUNCH(h, KID_FILL, k_fill_prepare<T>, grid_for(n), kBlock, P, (uint32_t)n, roots, vals, gf0);
    for (uint32_t r = 0; r < (uint32_t)kLevels + 1; r++) {
        uint32_t* fr = (r & 1) ? gf1 : gf0;
        uint32_t* nx = (r & 1) ? gf0 : gf1;
        unsigned grid = r == 0 ? std::max(grid_for(n), std::min<unsigned>(1024u, (unsigned)((n + kGcShareMin - 1) / kGcShareMin)))
                               : 1024u;
        VX_LAUNCH(h, KID_GC, k_gc_round<T>, grid, kBlock, P, r, fr, nx, r == 0, (uint64_t)gcap, false);
    }
    VX_LAUNCH(h, KID_BEGIN, k_begin, 1, 1, h->ctl);
    VX_LAUNCH(h, KID_LEAF_ALLOC, k_leaf_alloc_only<T>, 1, 1024, P, nvalues);
    VX_LAUNCH(h, KID_FILL, k_fill_finish<T>, grid_for(n), kBlock, P, (uint32_t)n, roots, vals);
    VX_LAUNCH(h, KID_END_ALLOC, k_end_alloc<T>, 1, 256, P, false, (uint32_t*)nullptr);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

template <typename T>
int get_impl(vx_interner* h, size_t nq, size_t ntrees, const uint64_t* roots,
             const uint8_t* depths, const uint32_t* chunk_of, const int32_t* xyz, void* out,
             uint8_t* present) {
    Pool<T> P = pool_of<T>(h);
    VX_LAUNCH(h, KID_GET, k_get<T>, grid_for(nq), kBlock, P, nq, ntrees, roots, depths, chunk_of, xyz,
              (T*)out, present);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

template <typename T>
int get_on_impl(vx_interner* h, hipStream_t st, size_t nq, size_t ntrees, const uint64_t* roots,
                const uint8_t* depths, const uint32_t* chunk_of, const int32_t* xyz, void* out,
                uint8_t* present, uint8_t* stop) {
    Pool<T> P = pool_of<T>(h);
    k_get<T><<<grid_for(nq), kBlock, 0, st>>>(P, nq, ntrees, roots, depths, chunk_of, xyz, (T*)out, present,
                                              stop);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

template <typename T>
int get_grouped_impl(vx_interner* h, hipStream_t st, size_t ntrees, const uint64_t* roots, const uint8_t* depths,
                     const uint64_t* qoff, size_t nq, const int32_t* xyz, void* out, uint8_t* present,
                     uint8_t* stop) {
    Pool<T> P = pool_of<T>(h);
    // about 4096 blocks for large batches; at least 512 queries per block
    uint64_t per = std::min<uint64_t>(4096, std::max<uint64_t>(512, (nq + 4095) / 4096));
    per = (per + kGetTile - 1) / kGetTile * kGetTile;
    const uint64_t blocks = (nq + per - 1) / per;
    if (blocks > 0x7FFFFFFFull) return VX_E_INVALID_ARG;
    if (st) {
        k_get_grouped<T, kGetTile><<<(unsigned)blocks, kGetTile, 0, st>>>(P, nq, (uint32_t)ntrees, roots, depths,
                                                                          qoff, xyz, (T*)out, present, stop, per);
    } else {
        VX_LAUNCH(h, KID_GET_GROUPED, k_get_grouped<T VX_COMMA kGetTile>, (unsigned)blocks, kGetTile, P, nq, (uint32_t)ntrees,
                  roots, depths, qoff, xyz, (T*)out, present, stop, per);
    }
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

template <typename T>
int to_vec_impl(vx_interner* h, size_t n, const uint64_t* roots, const uint8_t* depths, uint32_t lod,
                uint64_t stride, void* out) {
    Pool<T> P = pool_of<T>(h);
    uint64_t total = (uint64_t)n * stride;
    VX_LAUNCH(h, KID_TO_VEC, k_to_vec<T>, grid_for(total), kBlock, P, (uint64_t)n, roots, depths, lod, stride,
              (T*)out);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

template <typename T>
int canon_impl(vx_interner* h, size_t n, const uint64_t* roots, uint64_t* out) {
    Pool<T> P = pool_of<T>(h);
    VX_LAUNCH(h, KID_CANON, k_canon<T>, grid_for(n), 64, P, (uint32_t)n, roots, out);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

int sync_read_ctl(vx_interner* h, Ctl* out) {
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(out, h->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    return VX_OK;
}

}  // namespace

extern "C" {

const char* vx_status_string(int st) {
    switch (st) {
        case VX_OK: return "ok";
        case VX_E_INVALID_ARG: return "invalid argument";
        case VX_E_OOM: return "out of memory (interner capacity)";
        case VX_E_DEPTH: return "max depth out of range";
        case VX_E_RANGE: return "coordinate out of range";
        case VX_E_LEAF_BRICK: return "partial edit of a collapsed brick leaf";
        case VX_E_SAME_ROOT: return "apply produced the current root";
        case VX_E_HIP: return "HIP runtime error";
        case VX_E_BUDGET: return "memory budget invalid";
        case VX_E_UNSUPPORTED: return "unsupported voxel width";
        case VX_E_TABLE_FULL: return "pattern table full";
        case VX_E_POISONED: return "interner poisoned by an earlier device failure";
        case VX_E_SCRATCH: return "exceeds the fixed scratch of a no-allocation entry point";
        default: return "unknown";
    }
}

int vx_interner_create(uint32_t value_size, uint64_t budget_bytes, double load_factor,
                       vx_interner** out) {
    if (!out) return VX_E_INVALID_ARG;
    *out = nullptr;
    if (value_size != 1 && value_size != 2 && value_size != 4) return VX_E_UNSUPPORTED;
    uint64_t node_size = 4 + 2 + 64 + value_size + 8;  // interner/mod.rs:158-164
    uint64_t cap = budget_bytes / node_size;
    if (cap == 0 || cap > 0xFFFFFFFEull) return VX_E_BUDGET;
    if (!(load_factor > 0.0 && load_factor < 1.0)) load_factor = 0.5;
    vx_interner* h = new (std::nothrow) vx_interner();
    if (!h) return VX_E_HIP;
    h->vsize = value_size;
    h->capacity = cap;
    uint64_t want = (uint64_t)((double)cap / load_factor) + 1;
    uint64_t ts = 1024;
    while (ts < want) ts <<= 1;
    h->tslots = ts;
    // u8/u16: a direct map of every value; four-byte voxels: a hashed leaf table as large as
    // the branch table, and every leaf an apply creates listed once
    const bool direct = value_size <= 2;
    uint64_t nv = direct ? 1ull << (8 * value_size) : 1;
    h->lslots = direct ? 0 : ts;
    uint64_t nnew = direct ? nv : std::min<uint64_t>(cap, ts);
    bool ok = dmalloc(&h->children, cap * 64) && dmalloc(&h->tmg, cap * 4) &&
              dmalloc(&h->values, cap * value_size) && dmalloc(&h->refs, cap * 4) &&
              dmalloc(&h->free_stack, cap * 4) &&
              dmalloc(&h->tab, ts * 8) && dmalloc(&h->leaf_map, nv * 4) &&
              dmalloc(&h->leaf_need, nv) && dmalloc(&h->new_leaves, nnew * 4) &&
              (direct || (dmalloc(&h->ltab, h->lslots * 8) && dmalloc(&h->need_list, h->lslots * 4))) &&
              (!direct || dmalloc(&h->single, 8 * nv * 8)) &&
              (value_size != 2 || dmalloc(&h->leaf_ids, nv * 8)) &&
              dmalloc(&h->ctl, sizeof(Ctl));
    if (!ok || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        vx_interner_destroy(h);
        return VX_E_HIP;
    }
    h->own_stream = true;
    // PoolAllocatorLite::new zeroes every column (pool_allocator_lite.rs:29-81);
    // the EMPTY branch is index 0 with everything zero (interner/mod.rs:83-101).
    bool z = hipMemset(h->children, 0, cap * 64) == hipSuccess &&
             hipMemset(h->tmg, 0, cap * 4) == hipSuccess &&
             hipMemset(h->values, 0, cap * value_size) == hipSuccess &&
             hipMemset(h->refs, 0, cap * 4) == hipSuccess &&
             hipMemset(h->tab, 0, ts * 8) == hipSuccess && hipMemset(h->leaf_map, 0, nv * 4) == hipSuccess &&
             hipMemset(h->leaf_need, 0, nv) == hipSuccess &&
             (direct || hipMemset(h->ltab, 0, h->lslots * 8) == hipSuccess) &&
             (!direct || hipMemset(h->single, 0, 8 * nv * 8) == hipSuccess);
    Ctl c;
    std::memset(&c, 0, sizeof(c));
    c.next_index = 1;
    c.next0 = 1;
    z = z && hipMemcpy(h->ctl, &c, sizeof(c), hipMemcpyHostToDevice) == hipSuccess;
    z = z && hipDeviceSynchronize() == hipSuccess;
    if (!z) {
        vx_interner_destroy(h);
        return VX_E_HIP;
    }
    *out = h;
    return VX_OK;
}

int vx_interner_destroy(vx_interner* h) {
    if (!h) return VX_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void* ptrs[] = {h->children, h->tmg, h->values, h->refs, h->free_stack,
                    h->tab, h->leaf_map, h->leaf_need, h->new_leaves, h->ctl, h->leaf_ids,
                    h->ltab, h->need_list, h->single};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    h->chunks.release();
    h->slot_off.release();
    h->ptr_tables.release();
    h->gc_front[0].release();
    h->gc_front[1].release();
    h->stage.release();
    h->scratch.release();
    h->deferred.release();
    h->drift.release();
    h->undo.release();
    h->owners.release();
    for (Plan& pl : h->plans) pl.release();
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    for (Grow* g : {&h->edit_ws, &h->edit_keys[0], &h->edit_keys[1], &h->edit_order[0], &h->edit_order[1],
                    &h->edit_tmp, &h->edit_trees, &h->edit_first, &h->edit_status})
        g->release();
    for (auto& b : h->lv) {
        b.res.release();
        b.key.release();
        b.hash.release();
        b.id.release();
        b.tslot.release();
        b.slot.release();
        b.tiles.release();
        b.old.release();
    }
    for (hipEvent_t e : h->prof.pool) (void)hipEventDestroy(e);
    for (int i = 0; i < vx_interner::kRing; i++) {
        if (h->h_pinned_ring[i]) (void)hipHostFree(h->h_pinned_ring[i]);
        if (h->ring_ev[i]) (void)hipEventDestroy(h->ring_ev[i]);
    }
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return VX_OK;
}

int vx_interner_capacity(const vx_interner* h, uint64_t* capacity) {
    if (!h || !capacity) return VX_E_INVALID_ARG;
    *capacity = h->capacity;
    return VX_OK;
}

int vx_interner_set_stream(vx_interner* h, void* stream) {
    if (!h) return VX_E_INVALID_ARG;
    (void)hipStreamSynchronize(h->stream);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    h->own_stream = false;
    h->stream = (hipStream_t)stream;
    return VX_OK;
}

int vx_interner_sync(vx_interner* h) {
    if (!h) return VX_E_INVALID_ARG;
    Ctl c;
    int rc = sync_read_ctl(h, &c);
    if (rc) return rc;
    if (c.err & kErrPoison) h->poisoned = 1;
    return err_to_status(c.err);
}

int vx_interner_counts(vx_interner* h, uint64_t* leaves, uint64_t* branches, uint64_t* alive) {
    if (!h) return VX_E_INVALID_ARG;
    Ctl c;
    int rc = sync_read_ctl(h, &c);
    if (rc) return rc;
    if (leaves) *leaves = c.live_leaves;
    if (branches) *branches = c.live_branches + 1;  // + the EMPTY pattern (mod.rs:97)
    if (alive) *alive = (uint64_t)c.next_index - c.free_top;
    return VX_OK;
}

int vx_interner_stats(vx_interner* h, uint64_t* out, uint32_t n) {
    if (!h || !out) return VX_E_INVALID_ARG;
    Ctl c;
    int rc = sync_read_ctl(h, &c);
    if (rc) return rc;
    static const int map[VX_STAT_COUNT] = {ST_HIT, ST_MISS, ST_DUP, ST_LEAF_MISS, ST_COLLAPSE,
                                           ST_RECYCLED, ST_H2, ST_H3, ST_H4, ST_H5, ST_CAND,
                                           ST_PROBES, ST_CLAIMS, ST_CLAIMS_LOST, ST_MERGED, ST_CAND1,
                                           ST_MERGED1, ST_HIT1, ST_MISS1, ST_DUP1};
    for (uint32_t i = 0; i < n && i < (uint32_t)VX_STAT_COUNT; i++) {
        uint64_t sum = 0;
        for (int sidx = 0; sidx < kStatStripes; sidx++) sum += c.st[sidx][map[i]];
        out[i] = sum;
    }
    return VX_OK;
}

int vx_interner_get_ref(vx_interner* h, uint64_t id, uint32_t* ref) {
    if (!h || !ref || bid_index(id) >= h->capacity) return VX_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(ref, h->refs + bid_index(id), 4, hipMemcpyDeviceToHost));
    return VX_OK;
}

int vx_interner_get_value(vx_interner* h, uint64_t id, uint64_t* value) {
    if (!h || !value || bid_index(id) >= h->capacity) return VX_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    *value = 0;
    HIPCHK(hipMemcpy(value, (char*)h->values + (size_t)bid_index(id) * h->vsize, h->vsize,
                     hipMemcpyDeviceToHost));
    return VX_OK;
}

int vx_interner_get_children(vx_interner* h, uint64_t id, uint64_t children[8]) {
    if (!h || !children || bid_index(id) >= h->capacity) return VX_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(children, h->children + 8ull * bid_index(id), 64, hipMemcpyDeviceToHost));
    return VX_OK;
}

int vx_trees_apply_batches(vx_interner* h, size_t n, uint64_t* roots, const vx_batch_view* batches,
                           int32_t* status) {
    if (!h || (n && (!roots || !batches || !status))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    if (n == 0) return VX_OK;
    if (n > 0x7FFFFFFFull) return VX_E_INVALID_ARG;
    for (size_t i = 0; i < n; i++)
        if (batches[i].reserved != 0) return VX_E_INVALID_ARG;
    return VX_WIDTH_DISPATCH(h, apply_impl, h, n, roots, batches, status);
}

int vx_tree_apply_batch_host(vx_interner* h, uint64_t* root, const vx_batch_view* batch,
                             int* changed) {
    if (!h || !root || !batch || !changed) return VX_E_INVALID_ARG;
    uint32_t d = batch->depth;
    if (d > kMaxDepth) return VX_E_DEPTH;
    uint64_t P = d > 0 ? (1ull << (3 * (d - 1))) : 1ull;
    size_t mbytes = (batch->has_patches && batch->masks) ? 2 * P : 0;
    size_t vbytes = (batch->has_patches && batch->values) ? 8 * P * h->vsize : 0;
    size_t moff = 64, voff = 64 + ((mbytes + 255) / 256) * 256;
    if (!h->stage.ensure(voff + vbytes + 64)) return VX_E_HIP;
    char* base = (char*)h->stage.ptr;
    uint64_t* droot = (uint64_t*)base;
    int32_t* dstatus = (int32_t*)(base + 8);
    uint8_t* dm = (uint8_t*)(base + moff);
    void* dv = base + voff;
    HIPCHK(hipMemcpyAsync(droot, root, 8, hipMemcpyHostToDevice, h->stream));
    if (mbytes) HIPCHK(hipMemcpyAsync(dm, batch->masks, mbytes, hipMemcpyHostToDevice, h->stream));
    if (vbytes) HIPCHK(hipMemcpyAsync(dv, batch->values, vbytes, hipMemcpyHostToDevice, h->stream));
    vx_batch_view v = *batch;
    v.masks = mbytes ? dm : nullptr;
    v.values = vbytes ? dv : nullptr;
    int rc = vx_trees_apply_batches(h, 1, droot, &v, dstatus);
    if (rc) return rc;
    int32_t st = 0;
    HIPCHK(hipMemcpyAsync(root, droot, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(&st, dstatus, 4, hipMemcpyDeviceToHost, h->stream));
    rc = vx_interner_sync(h);
    if (rc) return rc;
    if (st < 0) {
        *changed = 0;
        return st;
    }
    *changed = st;
    return VX_OK;
}

int vx_trees_apply_edits(vx_interner* h, size_t n, uint64_t* roots, const uint8_t* depths,
                         const uint64_t* edit_offsets, const int32_t* xyz, const void* values,
                         int32_t* status) {
    if (!h || (n && (!roots || !depths || !edit_offsets || !status))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    if (n == 0) return VX_OK;
    if (n > 0x7FFFFFFFull) return VX_E_INVALID_ARG;
    if (edit_offsets[n] > 0 && (!xyz || !values)) return VX_E_INVALID_ARG;
    return VX_WIDTH_DISPATCH(h, apply_edits_impl, h, n, roots, depths, edit_offsets, xyz, values, status);
}

namespace {
int vx_trees_set(vx_interner* h, size_t n, uint64_t* roots, const uint8_t* depths, const int32_t* xyz,
                 const void* values, int32_t* status) {
    if (!h || (n && (!roots || !depths || !xyz || !values || !status))) return VX_E_INVALID_ARG;
    if (n == 0) return VX_OK;
    std::vector<vx_batch_view> views(n);
    for (size_t i = 0; i < n; i++) views[i] = vx_batch_view{nullptr, nullptr, 0, depths[i], 0, 1, 0};
    ApplyOpts o;
    o.set_xyz = xyz;
    o.set_values = values;
    return VX_WIDTH_DISPATCH(h, apply_impl, h, n, roots, views.data(), status, nullptr, o);
}

int opts_of(const vx_interner* h, const vx_apply_opts* o, ApplyOpts* out) {
    *out = ApplyOpts();
    if (!o) return VX_OK;
    if (o->reserved != 0 || (o->flags & ~(uint32_t)VX_APPLY_DRY_RUN)) return VX_E_INVALID_ARG;
    out->dry = (o->flags & VX_APPLY_DRY_RUN) != 0;
    if (o->node_limit != 0) out->alloc_limit = (uint32_t)std::min<uint64_t>(o->node_limit, 0xFFFFFFFEull);
    out->counts = o->counts;
    (void)h;
    return VX_OK;
}
}  // namespace

int vx_trees_apply_batches_ex(vx_interner* h, size_t n, uint64_t* roots, const vx_batch_view* batches,
                              int32_t* status, const vx_apply_opts* opts) {
    if (!h || (n && (!roots || !batches || !status))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    if (n == 0) return VX_OK;
    if (n > 0x7FFFFFFFull) return VX_E_INVALID_ARG;
    for (size_t i = 0; i < n; i++)
        if (batches[i].reserved != 0) return VX_E_INVALID_ARG;
    ApplyOpts o;
    int rc = opts_of(h, opts, &o);
    if (rc) return rc;
    return VX_WIDTH_DISPATCH(h, apply_impl, h, n, roots, batches, status, nullptr, o);
}

int vx_trees_apply_edits_ex(vx_interner* h, size_t n, uint64_t* roots, const uint8_t* depths,
                            const uint64_t* edit_offsets, const int32_t* xyz, const void* values,
                            int32_t* status, const vx_apply_opts* opts) {
    if (!h || (n && (!roots || !depths || !edit_offsets || !status))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    if (n == 0) return VX_OK;
    if (n > 0x7FFFFFFFull) return VX_E_INVALID_ARG;
    if (edit_offsets[n] > 0 && (!xyz || !values)) return VX_E_INVALID_ARG;
    ApplyOpts o;
    int rc = opts_of(h, opts, &o);
    if (rc) return rc;
    return VX_WIDTH_DISPATCH(h, apply_edits_impl, h, n, roots, depths, edit_offsets, xyz, values, status, o);
}

int vx_trees_get(vx_interner* h, size_t ntrees, const uint64_t* roots, const uint8_t* depths,
                 size_t nq, const uint32_t* chunk_of, const int32_t* xyz, void* out,
                 uint8_t* present) {
    if (!h || (nq && (!roots || !depths || !chunk_of || !xyz || !out))) return VX_E_INVALID_ARG;
    if (nq == 0) return VX_OK;
    return VX_WIDTH_DISPATCH(h, get_impl, h, nq, ntrees, roots, depths, chunk_of, xyz, out, present);
}

int vx_trees_get_on(vx_interner* h, void* stream, size_t ntrees, const uint64_t* roots,
                    const uint8_t* depths, size_t nq, const uint32_t* chunk_of, const int32_t* xyz,
                    void* out, uint8_t* present, uint8_t* stop) {
    if (!h || (nq && (!roots || !depths || !chunk_of || !xyz || !out))) return VX_E_INVALID_ARG;
    if (nq == 0) return VX_OK;
    hipStream_t st = (hipStream_t)stream;
    return VX_WIDTH_DISPATCH(h, get_on_impl, h, st, nq, ntrees, roots, depths, chunk_of, xyz, out, present, stop);
}

int vx_trees_get_grouped(vx_interner* h, void* stream, size_t ntrees, const uint64_t* roots,
                         const uint8_t* depths, const uint64_t* query_offsets, size_t nq, const int32_t* xyz,
                         void* out, uint8_t* present, uint8_t* stop) {
    if (!h || (nq && (!roots || !depths || !query_offsets || !xyz || !out))) return VX_E_INVALID_ARG;
    if (ntrees > 0xFFFFFFFFull) return VX_E_INVALID_ARG;
    if (nq == 0) return VX_OK;
    return VX_WIDTH_DISPATCH(h, get_grouped_impl, h, (hipStream_t)stream, ntrees, roots, depths, query_offsets,
                             nq, xyz, out, present, stop);
}

int vx_trees_fill(vx_interner* h, size_t n, uint64_t* roots, const uint64_t* values) {
    if (!h || (n && (!roots || !values))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    if (n == 0) return VX_OK;
    return VX_WIDTH_DISPATCH(h, fill_impl, h, n, roots, values);
}

int vx_trees_to_vec(vx_interner* h, size_t n, const uint64_t* roots, const uint8_t* depths, uint32_t lod,
                    uint64_t out_stride, void* out) {
    if (!h || (n && (!roots || !depths || !out || out_stride == 0))) return VX_E_INVALID_ARG;
    if (n == 0) return VX_OK;
    if (lod > (uint32_t)kMaxDepth || (uint64_t)n * out_stride / out_stride != (uint64_t)n) return VX_E_INVALID_ARG;
    return VX_WIDTH_DISPATCH(h, to_vec_impl, h, n, roots, depths, lod, out_stride, out);
}

int vx_trees_canonical_hash(vx_interner* h, size_t n, const uint64_t* roots, uint64_t* out) {
    if (!h || (n && (!roots || !out))) return VX_E_INVALID_ARG;
    if (n == 0) return VX_OK;
    return VX_WIDTH_DISPATCH(h, canon_impl, h, n, roots, out);
}

int vx_interner_set_tuning(vx_interner* h, const vx_tuning* t) {
    if (!h || !t) return VX_E_INVALID_ARG;
    if (t->tile != 128 && t->tile != 256 && t->tile != 512) return VX_E_INVALID_ARG;
    if (t->dedup_rounds > 16 || t->brick_merge_rounds > 16 ||
        (t->brick_rounds != 0 && t->brick_rounds != 1 && t->brick_rounds != 2 && t->brick_rounds != 4) ||
        t->commit_grid == 0 || t->commit_grid > (1u << 20) || t->gc_grid == 0 ||
        t->gc_grid > (1u << 20))
        return VX_E_INVALID_ARG;
    (void)hipStreamSynchronize(h->stream);
    h->tune = *t;
    return VX_OK;
}

int vx_interner_get_tuning(const vx_interner* h, vx_tuning* t) {
    if (!h || !t) return VX_E_INVALID_ARG;
    *t = h->tune;
    return VX_OK;
}

int vx_interner_rebuild_table(vx_interner* h, uint64_t* rebuilds) {
    if (!h) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    int rc = VX_WIDTH_DISPATCH(h, rebuild_step, h, true);
    if (rc) return rc;
    Ctl c;
    rc = sync_read_ctl(h, &c);
    if (rc) return rc;
    if (rebuilds) *rebuilds = c.rebuilds;
    return VX_OK;
}

int vx_interner_table_stats(vx_interner* h, uint64_t* slots, uint64_t* tombstones, uint64_t* rebuilds) {
    if (!h) return VX_E_INVALID_ARG;
    Ctl c;
    int rc = sync_read_ctl(h, &c);
    if (rc) return rc;
    if (slots) *slots = h->tslots;
    if (tombstones) *tombstones = tombstones_of(c);
    if (rebuilds) *rebuilds = c.rebuilds;
    return VX_OK;
}

int vx_interner_set_graphs(vx_interner* h, int enable) {
    if (!h) return VX_E_INVALID_ARG;
    (void)hipStreamSynchronize(h->stream);
    h->graphs = enable != 0;
    if (!h->graphs)
        for (Plan& pl : h->plans) pl.release();
    return VX_OK;
}

int vx_interner_profile(vx_interner* h, int enable) {
    if (!h) return VX_E_INVALID_ARG;
    h->prof.on = enable != 0;
    return VX_OK;
}

int vx_interner_profile_read(vx_interner* h, double* ms, uint64_t* launches, uint32_t n) {
    if (!h) return VX_E_INVALID_ARG;
    HIPCHK(hipStreamSynchronize(h->stream));
    Prof& p = h->prof;
    for (auto& pr : p.pending) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, p.pool[pr.second], p.pool[pr.second + 1]) == hipSuccess) {
            p.ms[pr.first] += t;
            p.launches[pr.first] += 1;
        }
    }
    p.pending.clear();
    p.used = 0;
    for (uint32_t i = 0; i < n && i < (uint32_t)KID_COUNT; i++) {
        if (ms) ms[i] = p.ms[i];
        if (launches) launches[i] = p.launches[i];
        p.ms[i] = 0;
        p.launches[i] = 0;
    }
    return (int)KID_COUNT;
}

const char* vx_kernel_name(uint32_t id) { return id < (uint32_t)KID_COUNT ? kKernelNames[id] : ""; }

int vx_bench_atomics(void* stream, int kind, void* buf, uint64_t slots, uint64_t ops, float* ms) {
    if (!buf || !ms || kind < 0 || kind > 4 || slots < 2048 || (slots & (slots - 1)) || ops == 0)
        return VX_E_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t per_lane = 64;
    uint64_t lanes = (ops + per_lane - 1) / per_lane;
    unsigned blocks = (unsigned)std::min<uint64_t>((lanes + 255) / 256, 1u << 20);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    k_atomic_bench<<<blocks, 256, 0, st>>>(buf, slots, per_lane, kind, (uint64_t*)buf);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

int vx_bench_copy(void* stream, void* dst, const void* src, uint64_t bytes, float* ms) {
    if (!dst || !src || !ms || bytes == 0 || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        return VX_E_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, st));
    k_copy_bench<<<2048, 256, 0, st>>>((const uint4*)src, (uint4*)dst, bytes / 16);
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

int vx_mlp_param_count(void) { return kMlpParams; }

int vx_mlp_forward(void* stream, const float* params, const float* x, uint32_t n, float* y) {
    if (!params || (n && (!x || !y))) return VX_E_INVALID_ARG;
    if (n == 0) return VX_OK;
    unsigned grid = (unsigned)std::min<uint64_t>((n + kMlpWaves - 1) / kMlpWaves, 1024);
    k_mlp_fwd<<<grid, 64 * kMlpWaves, 0, (hipStream_t)stream>>>(params, x, n, y, nullptr, nullptr);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

int vx_mlp_gradient(void* stream, const float* params, const float* x, const float* target, uint32_t n,
                    float* y, float* workspace, float* grad, float* loss) {
    if (!params || !grad || (n && (!x || !target || !y || !workspace))) return VX_E_INVALID_ARG;
    if (n == 0) return VX_E_INVALID_ARG;
    hipStream_t st = (hipStream_t)stream;
    unsigned grid = (unsigned)std::min<uint64_t>((n + kMlpWaves - 1) / kMlpWaves, 1024);
    k_mlp_fwd<<<grid, 64 * kMlpWaves, 0, st>>>(params, x, n, y, target, workspace);
    k_mlp_grad<<<(kMlpParams + 255) / 256, 256, 0, st>>>(x, target, y, n, workspace, grad, loss);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

int vx_mlp_adam(void* stream, float* params, const float* grad, float* m, float* v, uint32_t step, float lr) {
    if (!params || !grad || !m || !v || step == 0) return VX_E_INVALID_ARG;
    k_mlp_adam<<<(kMlpParams + 255) / 256, 256, 0, (hipStream_t)stream>>>(params, grad, m, v, step, lr, 0.9f,
                                                                         0.999f, 1e-8f);
    return hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
}

int vx_device_alloc(size_t bytes, void** ptr) {
    if (!ptr) return VX_E_INVALID_ARG;
    return hipMalloc(ptr, bytes ? bytes : 1) == hipSuccess ? VX_OK : VX_E_HIP;
}
int vx_device_free(void* ptr) { return hipFree(ptr) == hipSuccess ? VX_OK : VX_E_HIP; }
int vx_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? VX_OK : VX_E_HIP;
}
int vx_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? VX_OK : VX_E_HIP;
}

}  // extern "C"

#include "vx_vtm.hip"
