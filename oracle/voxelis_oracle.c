/*
 * voxelis_oracle.c -- CPU oracle for the batched SVO-DAG edit path.
 *
 * TEST INFRASTRUCTURE ONLY: a literal single-threaded restatement of the
 * reference Rust code, used as the checker by the test suite, the smoke check
 * and the CPU-baseline leg of the benchmark.  The product never links it.
 * Citations are file:line in the reference repository (voxelis/src/...).
 *
 * Deliberate differences from the reference, none observable on the path:
 *  - pattern maps use full-key identity (hash + compare of (children, types,
 *    mask) / value) instead of the 64-bit FxHash alone (interner/hash.rs:15-81);
 *  - the dec_ref_recursive worklist grows instead of writing past its 32768
 *    preallocated entries (interner/mod.rs:431-513);
 *  - panics are reported as negative return codes (see voxelis_oracle.h).
 */
#include "voxelis_oracle.h"

#include <setjmp.h>
#include <stdlib.h>
#include <string.h>

#define MAX_ALLOWED_DEPTH 7u
#define MAX_GENERATION 0x7FFEu
#define EMPTY_ID 0ull
#define INVALID_ID UINT64_MAX

/* ---------------------------------------------------------------- panics */

static __thread jmp_buf* g_jb = NULL;
static __thread int g_panic = 0;

static void panic_with(int code) {
    g_panic = code;
    if (g_jb) longjmp(*g_jb, 1);
    abort();
}

#define API_TRY                                \
    jmp_buf jb__;                              \
    jmp_buf* saved__ = g_jb;                   \
    if (setjmp(jb__)) {                        \
        g_jb = saved__;                        \
        return g_panic;                        \
    }                                          \
    g_jb = &jb__;
#define API_DONE g_jb = saved__;

/* ---------------------------------------------------------------- BlockId */
/* core/block_id.rs:18-30, 216-226: leaf(1) | types(8) | mask(8) | gen(15) | index(32) */

static inline uint64_t bid_new(uint32_t index, uint16_t gen, uint8_t types, uint8_t mask,
                               int leaf) {
    if (gen > MAX_GENERATION) panic_with(VXO_PANIC_GENERATION);
    return ((uint64_t)(leaf ? 1 : 0) << 63) | ((uint64_t)types << 55) | ((uint64_t)mask << 47) |
           (((uint64_t)gen & 0x7FFFu) << 32) | (uint64_t)index;
}
static inline int bid_is_leaf(uint64_t id) { return (id >> 63) == 1; }
static inline int bid_is_branch(uint64_t id) { return (id >> 63) == 0; }
static inline uint32_t bid_index(uint64_t id) { return (uint32_t)(id & 0xFFFFFFFFull); }
static inline uint16_t bid_gen(uint64_t id) { return (uint16_t)((id >> 32) & 0x7FFFu); }
/* block_id.rs:273-294: types()/mask() assert!(is_branch) in release builds too */
static inline uint8_t bid_types(uint64_t id) {
    if (!bid_is_branch(id)) panic_with(VXO_PANIC_LEAF_TYPES);
    return (uint8_t)((id >> 55) & 0xFF);
}
static inline uint8_t bid_mask(uint64_t id) {
    if (!bid_is_branch(id)) panic_with(VXO_PANIC_LEAF_TYPES);
    return (uint8_t)((id >> 47) & 0xFF);
}

uint64_t vxo_bid_new_leaf(uint32_t index, uint16_t generation) {
    jmp_buf* saved = g_jb;
    g_jb = NULL;
    uint64_t r = generation > MAX_GENERATION ? INVALID_ID : bid_new(index, generation, 0, 0, 1);
    g_jb = saved;
    return r;
}
uint64_t vxo_bid_new_branch(uint32_t index, uint16_t generation, uint8_t types, uint8_t mask) {
    if (generation > MAX_GENERATION) return INVALID_ID;
    return bid_new(index, generation, types, mask, 0);
}

/* ---------------------------------------------------------------- paths */

/* voxtree.rs:41-105: PATH_MASKS[max_depth][k] keeps the top k+1 octal digits of a
 * 3*max_depth-bit Morton path.  Row 7 (beyond the reference) follows the formula. */
uint32_t vxo_bid_path_mask(uint32_t max_depth, uint32_t level) {
    if (max_depth == 0 || level >= max_depth) return 0;
    uint32_t digits = level + 1;
    uint32_t ones = (digits * 3 >= 32) ? 0xFFFFFFFFu : ((1u << (digits * 3)) - 1u);
    return ones << (3 * (max_depth - 1 - level));
}

/* utils/common.rs:24-55: x -> bit 3n, y -> 3n+1, z -> 3n+2 over the low 10 bits */
static uint32_t spread3(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 10; b++) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}
uint32_t vxo_encode_child_index_path(int32_t x, int32_t y, int32_t z) {
    return spread3((uint32_t)x & 0x3FFu) | (spread3((uint32_t)y & 0x3FFu) << 1) |
           (spread3((uint32_t)z & 0x3FFu) << 2);
}

/* utils/common.rs:6-21 child_index: shift = max - current - 1 */
static inline uint32_t child_index(int32_t x, int32_t y, int32_t z, uint32_t cur, uint32_t max) {
    uint32_t sh = max - cur - 1;
    return (((uint32_t)x >> sh) & 1u) | ((((uint32_t)y >> sh) & 1u) << 1) |
           ((((uint32_t)z >> sh) & 1u) << 2);
}

/* ---------------------------------------------------------------- calc_average */
/* core/voxel.rs:96-141: mode of the children; ties keep the first seen, except that
 * a non-default value displaces a default maximum of equal count. */
uint64_t vxo_calc_average(const uint64_t* children, size_t n) {
    uint64_t values[8];
    size_t counts[8];
    size_t unique = 0;
    for (size_t k = 0; k < n && k < 8; k++) {
        uint64_t c = children[k];
        size_t i = 0;
        while (i < unique) {
            if (values[i] == c) {
                counts[i]++;
                break;
            }
            i++;
        }
        if (i == unique) {
            values[unique] = c;
            counts[unique] = 1;
            unique++;
        }
    }
    if (unique == 0) return 0;
    size_t max_i = 0, max_cnt = counts[0];
    for (size_t i = 1; i < unique; i++) {
        size_t cnt = counts[i];
        int is_default_i = values[i] == 0;
        int is_default_max = values[max_i] == 0;
        if (cnt > max_cnt || (cnt == max_cnt && is_default_max && !is_default_i)) {
            max_cnt = cnt;
            max_i = i;
        }
    }
    return values[max_i];
}

/* ---------------------------------------------------------------- hashing */
/* splitmix64 finalizer (Steele, Lea, Flood 2014; constants from Vigna's public
 * splitmix64.c).  Used for the pattern maps and the canonical structural hash. */
static inline uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

static uint64_t branch_key_hash(const uint64_t ch[8], uint8_t types, uint8_t mask) {
    uint64_t h = mix64(0x4252414E43480000ull ^ ((uint64_t)types << 8) ^ mask);
    for (int i = 0; i < 8; i++) h = mix64(h ^ (ch[i] + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1)));
    return h;
}
static uint64_t leaf_key_hash(uint64_t v) { return mix64(v ^ 0x4C45414600000000ull); }

/* ---------------------------------------------------------------- pattern map */
/* Open addressing, linear probing, tombstones.  Slot holds (hash, id); identity
 * is resolved by a full-key compare against the node columns. */

typedef struct {
    uint64_t* hash;
    uint64_t* id; /* 0 = empty, INVALID = tombstone */
    size_t cap;   /* power of two */
    size_t live;
    size_t used;  /* live + tombstones */
} pmap;

struct vxo_interner {
    uint32_t vsize;
    uint64_t vmask;
    size_t capacity;
    uint32_t next_index;
    uint32_t* free_indices;
    size_t free_len;
    uint32_t* ref_counts;
    uint16_t* generations;
    uint64_t (*children)[8];
    uint64_t* values;
    uint64_t* hashes;
    pmap maps[2]; /* [0] branch, [1] leaf (interner/consts.rs:9-10) */
    uint64_t* gc_stack;
    size_t gc_cap;
    int extended_depth;
    int in_apply;
    uint64_t stats[VXO_ST_COUNT];
    /* canonical-hash memo: per node index, stamped */
    uint64_t (*memo)[2];
    uint32_t* memo_stamp;
    uint32_t stamp;
};

static void pmap_init(pmap* m, size_t cap) {
    m->cap = cap;
    m->hash = (uint64_t*)calloc(cap, sizeof(uint64_t));
    m->id = (uint64_t*)calloc(cap, sizeof(uint64_t));
    m->live = m->used = 0;
}
static void pmap_free(pmap* m) {
    free(m->hash);
    free(m->id);
}

static int key_equal(const vxo_interner* o, int leaf, uint64_t id, const uint64_t* ch,
                     uint8_t types, uint8_t mask, uint64_t value) {
    uint32_t idx = bid_index(id);
    if (leaf) return o->values[idx] == value;
    if ((uint8_t)((id >> 55) & 0xFF) != types || (uint8_t)((id >> 47) & 0xFF) != mask) return 0;
    return memcmp(o->children[idx], ch, 8 * sizeof(uint64_t)) == 0;
}

/* returns slot of the key or (size_t)-1 */
static size_t pmap_find(const vxo_interner* o, const pmap* m, int leaf, uint64_t h,
                        const uint64_t* ch, uint8_t types, uint8_t mask, uint64_t value) {
    size_t s = (size_t)h & (m->cap - 1);
    for (;;) {
        uint64_t id = m->id[s];
        if (id == 0) return (size_t)-1;
        if (id != INVALID_ID && m->hash[s] == h && key_equal(o, leaf, id, ch, types, mask, value))
            return s;
        s = (s + 1) & (m->cap - 1);
    }
}

static void pmap_insert_raw(pmap* m, uint64_t h, uint64_t id) {
    size_t s = (size_t)h & (m->cap - 1);
    while (m->id[s] != 0 && m->id[s] != INVALID_ID) s = (s + 1) & (m->cap - 1);
    if (m->id[s] == 0) m->used++;
    m->hash[s] = h;
    m->id[s] = id;
    m->live++;
}

static void pmap_grow(pmap* m) {
    pmap n;
    size_t cap = m->cap;
    if (m->live * 2 >= cap / 2) cap *= 2;
    pmap_init(&n, cap);
    for (size_t s = 0; s < m->cap; s++)
        if (m->id[s] != 0 && m->id[s] != INVALID_ID) pmap_insert_raw(&n, m->hash[s], m->id[s]);
    pmap_free(m);
    *m = n;
}

static void pmap_insert(pmap* m, uint64_t h, uint64_t id) {
    if ((m->used + 1) * 2 > m->cap) pmap_grow(m);
    pmap_insert_raw(m, h, id);
}

static void pmap_remove_slot(pmap* m, size_t s) {
    m->id[s] = INVALID_ID;
    m->live--;
}

/* ---------------------------------------------------------------- interner */

uint64_t vxo_node_size(uint32_t vsize) {
    /* interner/mod.rs:158-164: u32 ref + u16 generation + [BlockId;8] + T + u64 hash */
    return 4 + 2 + 64 + vsize + 8;
}

vxo_interner* vxo_interner_new(uint32_t vsize, uint64_t budget, int* err) {
    if (err) *err = 0;
    if (!(vsize == 1 || vsize == 2 || vsize == 4 || vsize == 8)) {
        if (err) *err = VXO_PANIC_BUDGET;
        return NULL;
    }
    uint64_t node = vxo_node_size(vsize);
    uint64_t cap = budget / node; /* interner/mod.rs:49-53 */
    if (cap == 0 || cap > UINT32_MAX) {
        if (err) *err = VXO_PANIC_BUDGET;
        return NULL;
    }
    vxo_interner* o = (vxo_interner*)calloc(1, sizeof(vxo_interner));
    o->vsize = vsize;
    o->vmask = vsize == 8 ? UINT64_MAX : ((1ull << (8 * vsize)) - 1);
    o->capacity = (size_t)cap;
    o->free_indices = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)cap);
    o->ref_counts = (uint32_t*)calloc((size_t)cap, sizeof(uint32_t));
    o->generations = (uint16_t*)calloc((size_t)cap, sizeof(uint16_t));
    o->children = (uint64_t(*)[8])calloc((size_t)cap, sizeof(uint64_t[8]));
    o->values = (uint64_t*)calloc((size_t)cap, sizeof(uint64_t));
    o->hashes = (uint64_t*)calloc((size_t)cap, sizeof(uint64_t));
    if (!o->free_indices || !o->ref_counts || !o->generations || !o->children || !o->values ||
        !o->hashes) {
        vxo_interner_free(o);
        if (err) *err = VXO_PANIC_BUDGET;
        return NULL;
    }
    pmap_init(&o->maps[0], 1024);
    pmap_init(&o->maps[1], 1024);
    o->gc_cap = 32768; /* interner/consts.rs:15 PREALLOCATED_STACK_SIZE */
    o->gc_stack = (uint64_t*)malloc(sizeof(uint64_t) * o->gc_cap);
    /* interner/mod.rs:83-101: EMPTY branch pre-installed at index 0, ref 0 */
    o->hashes[0] = branch_key_hash(o->children[0], 0, 0);
    o->next_index = 1;
    return o;
}

void vxo_interner_free(vxo_interner* o) {
    if (!o) return;
    free(o->free_indices);
    free(o->ref_counts);
    free(o->generations);
    free(o->children);
    free(o->values);
    free(o->hashes);
    pmap_free(&o->maps[0]);
    pmap_free(&o->maps[1]);
    free(o->gc_stack);
    free(o->memo);
    free(o->memo_stamp);
    free(o);
}

/* Snapshot of the mutable interner state, for comparing against a device path whose
 * policy for a reference panic (the process aborts there) is "the call changed
 * nothing": take before an apply, restore if it panicked.  Test infrastructure. */
struct vxo_snapshot {
    uint32_t next_index;
    size_t free_len;
    uint32_t* free_indices;
    uint32_t* ref_counts;
    uint16_t* generations;
    uint64_t (*children)[8];
    uint64_t* values;
    uint64_t* hashes;
    pmap maps[2];
    uint64_t stats[VXO_ST_COUNT];
};

vxo_snapshot* vxo_snapshot_take(const vxo_interner* o) {
    vxo_snapshot* s = (vxo_snapshot*)calloc(1, sizeof(vxo_snapshot));
    size_t n = o->next_index;
    s->next_index = o->next_index;
    s->free_len = o->free_len;
    s->free_indices = (uint32_t*)malloc(sizeof(uint32_t) * (o->free_len + 1));
    memcpy(s->free_indices, o->free_indices, sizeof(uint32_t) * o->free_len);
    s->ref_counts = (uint32_t*)malloc(sizeof(uint32_t) * n);
    memcpy(s->ref_counts, o->ref_counts, sizeof(uint32_t) * n);
    s->generations = (uint16_t*)malloc(sizeof(uint16_t) * n);
    memcpy(s->generations, o->generations, sizeof(uint16_t) * n);
    s->children = (uint64_t(*)[8])malloc(sizeof(uint64_t[8]) * n);
    memcpy(s->children, o->children, sizeof(uint64_t[8]) * n);
    s->values = (uint64_t*)malloc(sizeof(uint64_t) * n);
    memcpy(s->values, o->values, sizeof(uint64_t) * n);
    s->hashes = (uint64_t*)malloc(sizeof(uint64_t) * n);
    memcpy(s->hashes, o->hashes, sizeof(uint64_t) * n);
    for (int i = 0; i < 2; i++) {
        const pmap* m = &o->maps[i];
        s->maps[i] = *m;
        s->maps[i].hash = (uint64_t*)malloc(sizeof(uint64_t) * m->cap);
        s->maps[i].id = (uint64_t*)malloc(sizeof(uint64_t) * m->cap);
        memcpy(s->maps[i].hash, m->hash, sizeof(uint64_t) * m->cap);
        memcpy(s->maps[i].id, m->id, sizeof(uint64_t) * m->cap);
    }
    memcpy(s->stats, o->stats, sizeof(s->stats));
    return s;
}

void vxo_snapshot_restore(vxo_interner* o, const vxo_snapshot* s) {
    size_t n = s->next_index, cur = o->next_index;
    memcpy(o->ref_counts, s->ref_counts, sizeof(uint32_t) * n);
    memcpy(o->generations, s->generations, sizeof(uint16_t) * n);
    memcpy(o->children, s->children, sizeof(uint64_t[8]) * n);
    memcpy(o->values, s->values, sizeof(uint64_t) * n);
    memcpy(o->hashes, s->hashes, sizeof(uint64_t) * n);
    if (cur > n) { /* slots first taken by the undone call go back to never-used */
        memset(o->ref_counts + n, 0, sizeof(uint32_t) * (cur - n));
        memset(o->generations + n, 0, sizeof(uint16_t) * (cur - n));
        memset(o->children + n, 0, sizeof(uint64_t[8]) * (cur - n));
        memset(o->values + n, 0, sizeof(uint64_t) * (cur - n));
        memset(o->hashes + n, 0, sizeof(uint64_t) * (cur - n));
    }
    o->next_index = s->next_index;
    o->free_len = s->free_len;
    memcpy(o->free_indices, s->free_indices, sizeof(uint32_t) * s->free_len);
    for (int i = 0; i < 2; i++) {
        pmap_free(&o->maps[i]);
        o->maps[i] = s->maps[i];
        o->maps[i].hash = (uint64_t*)malloc(sizeof(uint64_t) * s->maps[i].cap);
        o->maps[i].id = (uint64_t*)malloc(sizeof(uint64_t) * s->maps[i].cap);
        memcpy(o->maps[i].hash, s->maps[i].hash, sizeof(uint64_t) * s->maps[i].cap);
        memcpy(o->maps[i].id, s->maps[i].id, sizeof(uint64_t) * s->maps[i].cap);
    }
    memcpy(o->stats, s->stats, sizeof(s->stats));
}

void vxo_snapshot_free(vxo_snapshot* s) {
    if (!s) return;
    free(s->free_indices);
    free(s->ref_counts);
    free(s->generations);
    free(s->children);
    free(s->values);
    free(s->hashes);
    pmap_free(&s->maps[0]);
    pmap_free(&s->maps[1]);
    free(s);
}

void vxo_set_extended_depth(vxo_interner* o, int enable) { o->extended_depth = enable; }
uint64_t vxo_capacity(const vxo_interner* o) { return o->capacity; }

void vxo_counts(const vxo_interner* o, uint64_t* leaves, uint64_t* branches) {
    /* leaf_patterns().len(), branch_patterns().len() (mod.rs:1031-1037); the EMPTY
     * branch pattern inserted at mod.rs:97 is counted. */
    if (leaves) *leaves = o->maps[1].live;
    if (branches) *branches = o->maps[0].live + 1;
}
uint32_t vxo_get_ref(const vxo_interner* o, uint64_t id) { return o->ref_counts[bid_index(id)]; }
uint64_t vxo_get_value(const vxo_interner* o, uint64_t id) { return o->values[bid_index(id)]; }
void vxo_get_children(const vxo_interner* o, uint64_t id, uint64_t out[8]) {
    memcpy(out, o->children[bid_index(id)], sizeof(uint64_t[8]));
}
void vxo_stats(const vxo_interner* o, uint64_t out[VXO_ST_COUNT]) {
    memcpy(out, o->stats, sizeof(o->stats));
}
uint64_t vxo_alive_nodes(const vxo_interner* o) {
    return (uint64_t)(o->next_index - 1) - (uint64_t)o->free_len + 1;
}

/* interner/macros.rs:1-41: LIFO free list, then bump, then "Out of memory" */
static uint32_t next_index(vxo_interner* o) {
    if (o->free_len > 0) return o->free_indices[--o->free_len];
    if (o->next_index < (uint32_t)o->capacity) return o->next_index++;
    panic_with(VXO_PANIC_OOM);
    return 0;
}

/* interner/mod.rs:566-625 */
static void recycle(vxo_interner* o, uint64_t id) {
    uint32_t idx = bid_index(id);
    o->values[idx] = 0;
    memset(o->children[idx], 0, sizeof(uint64_t[8]));
    o->hashes[idx] = 0;
    o->ref_counts[idx] = 0;
    uint16_t g = (uint16_t)(o->generations[idx] + 1);
    if (g >= MAX_GENERATION) g = 0;
    o->generations[idx] = g;
    o->free_indices[o->free_len++] = idx;
    o->stats[VXO_ST_RECYCLED]++;
}

/* remove the pattern of a node that is being freed, by its key */
static void remove_pattern(vxo_interner* o, uint64_t id) {
    uint32_t idx = bid_index(id);
    int leaf = bid_is_leaf(id);
    pmap* m = &o->maps[leaf];
    size_t s;
    if (leaf)
        s = pmap_find(o, m, 1, o->hashes[idx], NULL, 0, 0, o->values[idx]);
    else
        s = pmap_find(o, m, 0, o->hashes[idx], o->children[idx], (uint8_t)((id >> 55) & 0xFF),
                      (uint8_t)((id >> 47) & 0xFF), 0);
    if (s != (size_t)-1 && m->id[s] == id) pmap_remove_slot(m, s);
}

static inline void inc_ref(vxo_interner* o, uint64_t id) { o->ref_counts[bid_index(id)] += 1; }
static inline void inc_ref_by(vxo_interner* o, uint64_t id, uint32_t n) {
    o->ref_counts[bid_index(id)] += n;
}

/* interner/mod.rs:262-296 (u32 arithmetic wraps: overflow-checks=false) */
static void dec_ref(vxo_interner* o, uint64_t id) {
    uint32_t idx = bid_index(id);
    o->ref_counts[idx] -= 1;
    if (o->ref_counts[idx] == 0) {
        if (o->in_apply) o->stats[VXO_ST_TRANSIENT_FREE]++;
        remove_pattern(o, id);
        recycle(o, id);
    }
}

/* interner/mod.rs:389-417 */
static void dec_ref_by(vxo_interner* o, uint64_t id, uint32_t n) {
    uint32_t idx = bid_index(id);
    o->ref_counts[idx] -= n;
    if (o->ref_counts[idx] == 0) {
        if (o->in_apply) o->stats[VXO_ST_TRANSIENT_FREE]++;
        remove_pattern(o, id);
        recycle(o, id);
    }
}

/* interner/mod.rs:298-326 */
static void inc_child_refs(vxo_interner* o, const uint64_t ch[8], size_t skip) {
    for (size_t i = 0; i < 8; i++) {
        if (i == skip) continue;
        if (ch[i] != EMPTY_ID) inc_ref(o, ch[i]);
    }
}

/* interner/mod.rs:536-564 */
static void dec_child_refs(vxo_interner* o, const uint64_t ch[8]) {
    for (size_t i = 0; i < 8; i++)
        if (ch[i] != EMPTY_ID) dec_ref(o, ch[i]);
}

/* interner/mod.rs:419-534 */
static void dec_ref_recursive(vxo_interner* o, uint64_t id) {
    size_t read = 0, write = 1;
    o->gc_stack[0] = id;
    while (read < write) {
        uint64_t cur = o->gc_stack[read++];
        uint32_t ci = bid_index(cur);
        o->ref_counts[ci] -= 1;
        if (o->ref_counts[ci] == 0) {
            for (int k = 0; k < 8; k++) {
                uint64_t child = o->children[ci][k];
                if (child == EMPTY_ID) continue;
                uint32_t* rc = &o->ref_counts[bid_index(child)];
                if (*rc > 1) {
                    *rc -= 1;
                } else {
                    if (write == o->gc_cap) {
                        o->gc_cap *= 2;
                        o->gc_stack = (uint64_t*)realloc(o->gc_stack, sizeof(uint64_t) * o->gc_cap);
                    }
                    o->gc_stack[write++] = child;
                }
            }
            remove_pattern(o, cur);
            recycle(o, cur);
        }
    }
}

/* interner/mod.rs:627-710 */
static uint64_t get_or_create_leaf(vxo_interner* o, uint64_t value) {
    value &= o->vmask;
    uint64_t h = leaf_key_hash(value);
    pmap* m = &o->maps[1];
    size_t s = pmap_find(o, m, 1, h, NULL, 0, 0, value);
    if (s != (size_t)-1) {
        uint64_t existing = m->id[s];
        if (o->generations[bid_index(existing)] != bid_gen(existing)) panic_with(VXO_PANIC_EXPIRED);
        inc_ref(o, existing);
        o->stats[VXO_ST_LEAF_HITS]++;
        return existing;
    }
    uint32_t index = next_index(o);
    uint16_t gen = o->generations[index];
    uint64_t id = bid_new(index, gen, 0, 0, 1);
    o->values[index] = value;
    o->hashes[index] = h;
    pmap_insert(m, h, id);
    inc_ref(o, id);
    o->stats[VXO_ST_LEAF_MISSES]++;
    return id;
}

/* interner/mod.rs:716-829 */
static uint64_t get_or_create_branch(vxo_interner* o, const uint64_t ch[8], uint8_t types,
                                     uint8_t mask) {
    uint64_t h = branch_key_hash(ch, types, mask);
    pmap* m = &o->maps[0];
    size_t s = pmap_find(o, m, 0, h, ch, types, mask, 0);
    if (s != (size_t)-1) {
        uint64_t existing = m->id[s];
        dec_child_refs(o, ch);
        inc_ref(o, existing);
        o->stats[VXO_ST_BRANCH_HITS]++;
        return existing;
    }
    uint32_t index = next_index(o);
    uint16_t gen = o->generations[index];
    uint64_t id = bid_new(index, gen, types, mask, 0);
    uint64_t vals[8];
    for (int i = 0; i < 8; i++) vals[i] = o->values[bid_index(ch[i])];
    memcpy(o->children[index], ch, sizeof(uint64_t[8]));
    o->values[index] = vxo_calc_average(vals, 8);
    o->hashes[index] = h;
    pmap_insert(m, h, id);
    inc_ref(o, id);
    o->stats[VXO_ST_BRANCH_MISSES]++;
    return id;
}

int vxo_get_or_create_leaf(vxo_interner* o, uint64_t value, uint64_t* out) {
    API_TRY
    *out = get_or_create_leaf(o, value);
    API_DONE
    return 0;
}
int vxo_get_or_create_branch(vxo_interner* o, const uint64_t children[8], uint8_t types,
                             uint8_t mask, uint64_t* out) {
    API_TRY
    *out = get_or_create_branch(o, children, types, mask);
    API_DONE
    return 0;
}
int vxo_dec_ref_recursive(vxo_interner* o, uint64_t id) {
    API_TRY
    dec_ref_recursive(o, id);
    API_DONE
    return 0;
}

static inline uint64_t get_child_id(const vxo_interner* o, uint64_t id, uint32_t index) {
    return o->children[bid_index(id)][index];
}

static void check_depth(const vxo_interner* o, uint32_t depth) {
    if (depth >= MAX_ALLOWED_DEPTH + (o->extended_depth ? 1u : 0u)) panic_with(VXO_PANIC_DEPTH);
}

/* ---------------------------------------------------------------- batch */

struct vxo_batch {
    uint8_t max_depth;
    size_t bricks;
    uint8_t* masks;   /* (set, clear) pairs */
    uint64_t* values; /* 8 per brick */
    int has_fill;
    uint64_t fill;
    int has_patches;
};

/* core/batch.rs:63-81 */
vxo_batch* vxo_batch_new(uint8_t max_depth) {
    vxo_batch* b = (vxo_batch*)calloc(1, sizeof(vxo_batch));
    b->max_depth = max_depth;
    uint32_t lower = max_depth > 0 ? max_depth - 1u : 0u;
    b->bricks = (size_t)1 << (3 * lower);
    b->masks = (uint8_t*)calloc(b->bricks * 2, 1);
    b->values = (uint64_t*)calloc(b->bricks * 8, sizeof(uint64_t));
    return b;
}
void vxo_batch_free(vxo_batch* b) {
    if (!b) return;
    free(b->masks);
    free(b->values);
    free(b);
}

/* core/batch.rs:145-175 */
void vxo_batch_just_set(vxo_batch* b, int32_t x, int32_t y, int32_t z, uint64_t v) {
    uint32_t full = vxo_encode_child_index_path(x, y, z);
    size_t p = (size_t)(full >> 3);
    uint32_t index = full & 7u;
    uint8_t bit = (uint8_t)(1u << index);
    if (p >= b->bricks) return; /* debug_assert range in the reference */
    if (v != 0) {
        b->masks[2 * p] |= bit;
        b->masks[2 * p + 1] &= (uint8_t)~bit;
    } else {
        b->masks[2 * p] &= (uint8_t)~bit;
        b->masks[2 * p + 1] |= bit;
    }
    b->values[8 * p + index] = v;
    b->has_patches = 1;
}
void vxo_batch_set_many(vxo_batch* b, size_t n, const int32_t* xyz, const uint64_t* values) {
    for (size_t i = 0; i < n; i++)
        vxo_batch_just_set(b, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], values[i]);
}
/* core/batch.rs:178-195 */
void vxo_batch_just_clear(vxo_batch* b) {
    memset(b->masks, 0, b->bricks * 2);
    memset(b->values, 0, b->bricks * 8 * sizeof(uint64_t));
    b->has_fill = 0;
    b->fill = 0;
    b->has_patches = 0;
}
void vxo_batch_just_fill(vxo_batch* b, uint64_t v) {
    vxo_batch_just_clear(b);
    b->has_fill = 1;
    b->fill = v;
}
/* core/batch.rs:115-123 */
size_t vxo_batch_size(const vxo_batch* b) {
    size_t n = 0;
    for (size_t p = 0; p < b->bricks; p++)
        if (b->masks[2 * p] != 0 || b->masks[2 * p + 1] != 0) n++;
    return n;
}
int vxo_batch_has_patches(const vxo_batch* b) { return b->has_patches; }

/* Load dense buffers recorded elsewhere (masks: 2*P bytes, values: 8*P voxels of
 * vsize bytes) as if just_set had been called for them; has_patches follows. */
void vxo_batch_load(vxo_batch* b, const uint8_t* masks, const void* values, uint32_t vsize,
                    int has_patches) {
    memcpy(b->masks, masks, b->bricks * 2);
    for (size_t i = 0; i < b->bricks * 8; i++) {
        uint64_t v = 0;
        memcpy(&v, (const uint8_t*)values + i * vsize, vsize);
        b->values[i] = v;
    }
    b->has_patches = has_patches;
}
size_t vxo_batch_bricks(const vxo_batch* b) { return b->bricks; }
const uint8_t* vxo_batch_masks(const vxo_batch* b) { return b->masks; }
const uint64_t* vxo_batch_values(const vxo_batch* b) { return b->values; }

/* ---------------------------------------------------------------- get */

/* utils/common.rs:122-156 */
static int get_at_depth(const vxo_interner* o, uint64_t node, int32_t x, int32_t y, int32_t z,
                        uint32_t max_depth, uint64_t* out) {
    uint32_t depth = 0;
    while (node != EMPTY_ID) {
        if (depth >= max_depth) {
            uint64_t v = o->values[bid_index(node)];
            *out = v;
            return v != 0;
        }
        if (bid_is_branch(node)) {
            node = get_child_id(o, node, child_index(x, y, z, depth, max_depth));
            depth++;
        } else {
            *out = o->values[bid_index(node)];
            return 1;
        }
    }
    *out = 0;
    return 0;
}

/* voxtree.rs:145-159 */
int vxo_tree_get(const vxo_interner* o, uint64_t root, uint8_t depth, int32_t x, int32_t y,
                 int32_t z, uint64_t* value, int* present) {
    int32_t n = 1 << depth;
    if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) return VXO_PANIC_RANGE;
    uint64_t v = 0;
    int p = get_at_depth(o, root, x, y, z, depth, &v);
    *value = p ? v : 0;
    *present = p;
    return 0;
}

int vxo_tree_get_many(const vxo_interner* o, uint64_t root, uint8_t depth, size_t n,
                      const int32_t* xyz, uint64_t* values, uint8_t* present) {
    for (size_t i = 0; i < n; i++) {
        int p = 0;
        uint64_t v = 0;
        int rc = vxo_tree_get(o, root, depth, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], &v, &p);
        if (rc) return rc;
        values[i] = v;
        if (present) present[i] = (uint8_t)p;
    }
    return 0;
}

/* ---------------------------------------------------------------- single set */

/* voxtree.rs:393-561 */
static uint64_t set_at_depth_iterative(vxo_interner* o, uint64_t initial, int32_t x, int32_t y,
                                       int32_t z, uint32_t max_depth, uint64_t voxel) {
    uint64_t st_node[MAX_ALLOWED_DEPTH + 1], st_leaf[MAX_ALLOWED_DEPTH + 1];
    uint8_t st_idx[MAX_ALLOWED_DEPTH + 1];
    uint64_t current = initial;
    uint32_t d = 0;
    uint64_t leaf_node = EMPTY_ID;
    while (d < max_depth) {
        uint32_t index = child_index(x, y, z, d, max_depth);
        if (bid_is_branch(current)) {
            st_node[d] = current;
            st_leaf[d] = leaf_node;
            st_idx[d] = (uint8_t)index;
            current = get_child_id(o, current, index);
        } else {
            if (o->values[bid_index(current)] == voxel) return INVALID_ID;
            leaf_node = current;
            current = EMPTY_ID;
            st_node[d] = current;
            st_leaf[d] = leaf_node;
            st_idx[d] = (uint8_t)index;
        }
        d++;
    }
    if (bid_is_leaf(current) && o->values[bid_index(current)] == voxel) return INVALID_ID;
    current = get_or_create_leaf(o, voxel);
    while (d > 0) {
        d--;
        uint64_t parent = st_node[d], leaf = st_leaf[d];
        uint32_t pidx = st_idx[d];
        if (parent != EMPTY_ID) {
            uint8_t types = (uint8_t)(bid_types(parent) | ((bid_is_leaf(current) ? 1u : 0u) << pidx));
            uint64_t branch[8];
            memcpy(branch, o->children[bid_index(parent)], sizeof(branch));
            branch[pidx] = current;
            int all = 1;
            for (int i = 0; i < 8; i++)
                if (branch[i] != current) all = 0;
            if (!(types == 0xFF && all)) {
                uint8_t mask = (uint8_t)(bid_mask(parent) | (1u << pidx));
                inc_child_refs(o, branch, pidx);
                current = get_or_create_branch(o, branch, types, mask);
            }
        } else if (leaf == EMPTY_ID) {
            uint64_t ch[8] = {0};
            ch[pidx] = current;
            uint8_t types = (uint8_t)((bid_is_leaf(current) ? 1u : 0u) << pidx);
            uint8_t mask = (uint8_t)(1u << pidx);
            current = get_or_create_branch(o, ch, types, mask);
        } else {
            uint64_t ch[8];
            for (int i = 0; i < 8; i++) ch[i] = leaf;
            inc_ref_by(o, leaf, 7);
            ch[pidx] = current;
            uint8_t types =
                (uint8_t)(~(1u << pidx) | ((bid_is_leaf(current) ? 1u : 0u) << pidx));
            current = get_or_create_branch(o, ch, types, 0xFF);
        }
    }
    return current;
}

static uint64_t remove_at_depth(vxo_interner* o, uint64_t node, int32_t x, int32_t y, int32_t z,
                                uint32_t cur, uint32_t max);

/* voxtree.rs:662-705 */
static uint64_t remove_at_depth_leaf(vxo_interner* o, uint64_t node, int32_t x, int32_t y,
                                     int32_t z, uint32_t cur, uint32_t max) {
    if (cur == max) return EMPTY_ID;
    uint64_t nn = remove_at_depth_leaf(o, node, x, y, z, cur + 1, max);
    uint32_t index = child_index(x, y, z, cur, max);
    uint64_t ch[8];
    for (int i = 0; i < 8; i++) ch[i] = node;
    ch[index] = nn;
    int is_leaf = bid_is_leaf(nn), is_empty = nn == EMPTY_ID;
    uint8_t types = (uint8_t)(~(1u << index) | ((is_leaf ? 1u : 0u) << index));
    uint8_t mask = (uint8_t)(~(1u << index) | ((!is_empty ? 1u : 0u) << index));
    inc_ref_by(o, node, 7);
    return get_or_create_branch(o, ch, types, mask);
}

/* voxtree.rs:581-660 */
static uint64_t remove_at_depth_branch(vxo_interner* o, uint64_t node, int32_t x, int32_t y,
                                       int32_t z, uint32_t cur, uint32_t max) {
    if (!(cur < max)) panic_with(VXO_PANIC_REMOVE);
    uint32_t index = child_index(x, y, z, cur, max);
    if (bid_mask(node) & (1u << index)) {
        uint64_t branch[8];
        memcpy(branch, o->children[bid_index(node)], sizeof(branch));
        uint64_t child = branch[index];
        uint64_t nc = remove_at_depth(o, child, x, y, z, cur + 1, max);
        if (nc != INVALID_ID) {
            int is_empty = nc == EMPTY_ID;
            uint8_t mask = is_empty ? (uint8_t)(bid_mask(node) & ~(1u << index)) : bid_mask(node);
            if (mask == 0) return EMPTY_ID;
            if (!bid_is_branch(nc)) panic_with(VXO_PANIC_REMOVE);
            uint8_t types = (uint8_t)(bid_types(node) & ~(1u << index));
            branch[index] = nc;
            inc_child_refs(o, branch, index);
            return get_or_create_branch(o, branch, types, mask);
        }
        return INVALID_ID;
    }
    return INVALID_ID;
}

/* voxtree.rs:563-579 */
static uint64_t remove_at_depth(vxo_interner* o, uint64_t node, int32_t x, int32_t y, int32_t z,
                                uint32_t cur, uint32_t max) {
    if (node == INVALID_ID) panic_with(VXO_PANIC_INVALID);
    if (bid_is_branch(node)) return remove_at_depth_branch(o, node, x, y, z, cur, max);
    return remove_at_depth_leaf(o, node, x, y, z, cur, max);
}

/* voxtree.rs:162-260 */
int vxo_tree_set(vxo_interner* o, uint64_t* root, uint8_t depth, int32_t x, int32_t y, int32_t z,
                 uint64_t v_in, int* changed) {
    const uint64_t v = v_in & o->vmask;
    API_TRY
    check_depth(o, depth);
    int32_t n = 1 << depth;
    if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) panic_with(VXO_PANIC_RANGE);
    uint64_t new_root;
    *changed = 0;
    if (*root != EMPTY_ID || v != 0) {
        if (*root == INVALID_ID) panic_with(VXO_PANIC_INVALID);
        new_root = v != 0 ? set_at_depth_iterative(o, *root, x, y, z, depth, v)
                          : remove_at_depth(o, *root, x, y, z, 0, depth);
    } else {
        API_DONE
        return 0;
    }
    if (new_root != INVALID_ID) {
        if (*root != EMPTY_ID) {
            if (new_root == *root) panic_with(VXO_PANIC_SAME_ROOT);
            dec_ref_recursive(o, *root);
        }
        *root = new_root;
        *changed = 1;
    }
    API_DONE
    return 0;
}

/* One set on each of n trees (tree i: roots[i], depths[i], voxel xyz[3i..3i+2], value
 * values[i]), in tree order: the per-voxel CPU baseline without a call per set.
 * status[i]: 1 changed, 0 unchanged, < 0 the panic's code (that tree is left as it was
 * at the panic; the others go on). */
int vxo_trees_set_round(vxo_interner* o, size_t n, uint64_t* roots, const uint8_t* depths, const int32_t* xyz,
                        const uint64_t* values, int* status) {
    for (size_t i = 0; i < n; i++) {
        int ch = 0;
        int rc = vxo_tree_set(o, &roots[i], depths[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], values[i], &ch);
        status[i] = rc < 0 ? rc : ch;
    }
    return 0;
}

/* voxtree.rs:263-292 */
int vxo_tree_clear(vxo_interner* o, uint64_t* root) {
    API_TRY
    if (*root != EMPTY_ID) {
        dec_ref_recursive(o, *root);
        *root = EMPTY_ID;
    }
    API_DONE
    return 0;
}
int vxo_tree_fill(vxo_interner* o, uint64_t* root, uint8_t depth, uint64_t v) {
    (void)depth;
    v &= o->vmask;
    if (v == 0) return vxo_tree_clear(o, root);
    API_TRY
    if (*root != EMPTY_ID) dec_ref_recursive(o, *root);
    *root = get_or_create_leaf(o, v);
    API_DONE
    return 0;
}

/* ---------------------------------------------------------------- apply_batch */

typedef struct {
    size_t* v;
    size_t n, cap;
} pathvec;
static void pv_push(pathvec* p, size_t x) {
    if (p->n == p->cap) {
        p->cap = p->cap ? p->cap * 2 : 64;
        p->v = (size_t*)realloc(p->v, p->cap * sizeof(size_t));
    }
    p->v[p->n++] = x;
}

/* voxtree.rs:724-1118 set_batch_at_depth_iterative; buffers are freed by the caller */
static uint64_t set_batch(vxo_interner* o, uint64_t initial, uint32_t max_depth,
                          const vxo_batch* b, uint64_t** bufs, pathvec* pv) {
    /* Phase 0 (742-758) */
    if (b->has_fill) {
        initial = get_or_create_leaf(o, b->fill);
        if (b->has_patches) o->stats[VXO_ST_H3_LEAKS]++;
    }
    if (!b->has_patches) return initial;
    /* H9: max_depth - 1 underflows for D < 2 and Phase 2 indexes PATH_MASKS out of range */
    if (max_depth < 2) panic_with(VXO_PANIC_DEPTH);

    size_t data_len = b->bricks;
    uint64_t* cur = bufs[0] = (uint64_t*)malloc(data_len * sizeof(uint64_t));
    uint64_t* nxt = bufs[1] = (uint64_t*)malloc(data_len * sizeof(uint64_t));
    for (size_t i = 0; i < data_len; i++) cur[i] = nxt[i] = INVALID_ID;
    pathvec* paths = &pv[0];
    pathvec* next_paths = &pv[1];

    /* Phase 1 (778-897) */
    for (size_t p = 0; p < data_len; p++) {
        uint8_t set_mask = b->masks[2 * p];
        if (set_mask == 0) continue;
        size_t path = p << 3;
        uint64_t current = initial;
        uint64_t leaf_node = EMPTY_ID;
        for (uint32_t cd = 0; cd < max_depth - 1; cd++) {
            if (bid_is_branch(current)) {
                uint32_t index = (uint32_t)((path >> ((max_depth - cd - 1) * 3)) & 7u);
                current = get_child_id(o, current, index);
            } else {
                leaf_node = current;
                current = EMPTY_ID;
            }
            if (current == EMPTY_ID) break;
        }
        const uint64_t* values = &b->values[8 * p];
        int all_same = set_mask == 0xFF;
        for (int i = 1; i < 8 && all_same; i++)
            if (values[i] != values[0]) all_same = 0;
        if (!all_same) {
            uint64_t children[8];
            uint8_t types, mask;
            if (current != EMPTY_ID) {
                memcpy(children, o->children[bid_index(current)], sizeof(children));
                types = bid_types(current); /* H1: panics when current is a leaf */
                mask = bid_mask(current);
            } else if (bid_is_leaf(leaf_node)) {
                uint32_t dl = (uint32_t)__builtin_popcount(set_mask);
                inc_ref_by(o, leaf_node, 8 - dl);
                for (int i = 0; i < 8; i++) children[i] = leaf_node;
                types = mask = 0xFF;
            } else {
                memset(children, 0, sizeof(children));
                types = mask = 0;
            }
            uint8_t modified = 0;
            uint8_t bits = set_mask;
            int skipped_equal_split = 0;
            while (bits) {
                int idx = __builtin_ctz(bits);
                bits &= (uint8_t)~(1u << idx);
                uint64_t value = values[idx] & o->vmask;
                if (children[idx] != EMPTY_ID && o->values[bid_index(children[idx])] == value) {
                    if (leaf_node != EMPTY_ID && children[idx] == leaf_node) skipped_equal_split++;
                    continue;
                }
                children[idx] = get_or_create_leaf(o, value);
                types |= (uint8_t)(1u << idx);
                mask |= (uint8_t)(1u << idx);
                modified |= (uint8_t)(1u << idx);
            }
            if (modified == 0) {
                if (leaf_node != EMPTY_ID) o->stats[VXO_ST_H3_LEAKS]++;
                continue;
            }
            if (skipped_equal_split) o->stats[VXO_ST_H4_UNDERCOUNT]++;
            if (leaf_node == EMPTY_ID) {
                uint8_t nm = (uint8_t)~modified;
                while (nm) {
                    int idx = __builtin_ctz(nm);
                    nm &= (uint8_t)~(1u << idx);
                    if (children[idx] != EMPTY_ID) inc_ref_by(o, children[idx], 1);
                }
            }
            int uniform = types == 0xFF;
            for (int i = 1; i < 8 && uniform; i++)
                if (children[i] != children[0]) uniform = 0;
            if (uniform) o->stats[VXO_ST_H2_UNCOLLAPSED]++;
            uint64_t branch_id = get_or_create_branch(o, children, types, mask);
            cur[p] = branch_id;
            pv_push(paths, path);
        } else {
            o->stats[VXO_ST_COLLAPSES]++;
            uint64_t first = values[__builtin_ctz(set_mask)];
            cur[p] = get_or_create_leaf(o, first);
            pv_push(paths, path);
        }
    }

    /* Phase 2 (905-1117) */
    if (paths->n == 0) return INVALID_ID;
    size_t target_depth = max_depth - 1;
    while (paths->n > 0) {
        size_t path = paths->v[--paths->n];
        size_t pmd = target_depth > 1 ? target_depth - 2 : 0;
        size_t path_mask = vxo_bid_path_mask(max_depth, (uint32_t)pmd);
        uint64_t eq = initial, leaf_id = EMPTY_ID;
        for (size_t cd = 0; cd < target_depth - 1; cd++) {
            if (bid_is_leaf(eq)) {
                leaf_id = eq;
                eq = EMPTY_ID;
                break;
            } else {
                uint32_t index = (uint32_t)((path >> ((max_depth - cd - 1) * 3)) & 7u);
                eq = get_child_id(o, eq, index);
            }
            if (eq == EMPTY_ID) break;
        }
        if (bid_is_leaf(eq)) {
            leaf_id = eq;
            eq = EMPTY_ID;
        }
        uint64_t children[8] = {0};
        uint8_t types = 0, mask = 0;
        int has_next = 1;
        while (has_next) {
            uint32_t ti = (uint32_t)((path >> ((max_depth - target_depth) * 3)) & 7u);
            if (target_depth == 1)
                has_next = paths->n > 0;
            else
                has_next = paths->n > 0 && (path & path_mask) == (paths->v[paths->n - 1] & path_mask);
            size_t cpi = path >> 3;
            uint64_t lid = cur[cpi];
            children[ti] = lid;
            cur[cpi] = INVALID_ID;
            types |= (uint8_t)((bid_is_leaf(lid) ? 1u : 0u) << ti);
            mask |= (uint8_t)(1u << ti);
            if (has_next && paths->n > 0) path = paths->v[--paths->n];
        }
        uint8_t existing_mask = bid_mask(eq);
        uint8_t inv = (uint8_t)~mask;
        uint8_t cloned = existing_mask & inv;
        if (mask != 0xFF) {
            if (cloned != 0) {
                const uint64_t* ex = o->children[bid_index(eq)];
                uint8_t cb = cloned;
                while (cb) {
                    int idx = __builtin_ctz(cb);
                    cb &= (uint8_t)~(1u << idx);
                    children[idx] = ex[idx];
                    types |= (uint8_t)((bid_is_leaf(children[idx]) ? 1u : 0u) << idx);
                    mask |= (uint8_t)(1u << idx);
                }
            } else if (leaf_id != EMPTY_ID) {
                uint32_t n = (uint32_t)__builtin_popcount(inv);
                types |= inv;
                mask |= inv;
                uint8_t lb = inv;
                while (lb) {
                    int idx = __builtin_ctz(lb);
                    children[idx] = leaf_id;
                    lb &= (uint8_t)~(1u << idx);
                }
                inc_ref_by(o, leaf_id, n);
            }
        }
        int all_same = types == 0xFF;
        for (int i = 1; i < 8 && all_same; i++)
            if (children[i] != children[0]) all_same = 0;
        uint64_t new_node;
        if (!all_same) {
            uint8_t cb = cloned;
            while (cb) {
                int idx = __builtin_ctz(cb);
                cb &= (uint8_t)~(1u << idx);
                inc_ref(o, children[idx]);
            }
            new_node = get_or_create_branch(o, children, types, mask);
        } else {
            o->stats[VXO_ST_COLLAPSES]++;
            uint32_t dec = 7;
            if (cloned != 0) {
                uint32_t pc = (uint32_t)__builtin_popcount(cloned);
                dec = pc < 7 ? pc : 7;
                o->stats[VXO_ST_H5_COLLAPSE_CLONED]++;
            }
            dec_ref_by(o, children[0], dec);
            new_node = children[0];
        }
        size_t np = path & path_mask;
        pv_push(next_paths, np);
        nxt[np >> 3] = new_node;
        if (paths->n == 0) {
            uint64_t* t = cur;
            cur = nxt;
            nxt = t;
            pathvec tp = *paths;
            *paths = *next_paths;
            *next_paths = tp;
            next_paths->n = 0;
            target_depth -= 1;
            if (target_depth == 0) break;
        }
    }
    bufs[0] = cur;
    bufs[1] = nxt;
    return cur[paths->v[0] >> 3];
}

/* voxtree.rs:303-328 */
int vxo_tree_apply_batch(vxo_interner* o, uint64_t* root, uint8_t depth, const vxo_batch* b,
                         int* changed) {
    uint64_t* bufs[2] = {NULL, NULL};
    pathvec pv[2] = {{NULL, 0, 0}, {NULL, 0, 0}};
    int rc = 0;
    *changed = 0;
    {
        jmp_buf jb;
        jmp_buf* saved = g_jb;
        if (setjmp(jb)) {
            g_jb = saved;
            rc = g_panic;
            goto out;
        }
        g_jb = &jb;
        check_depth(o, depth);
        if (*root == INVALID_ID) panic_with(VXO_PANIC_INVALID);
        o->in_apply = 1;
        uint64_t new_root = set_batch(o, *root, depth, b, bufs, pv);
        o->in_apply = 0;
        if (new_root != INVALID_ID) {
            if (*root != EMPTY_ID) {
                if (new_root == *root) panic_with(VXO_PANIC_SAME_ROOT);
                dec_ref_recursive(o, *root);
            }
            *root = new_root;
            *changed = 1;
        }
        g_jb = saved;
    }
out:
    o->in_apply = 0;
    free(bufs[0]);
    free(bufs[1]);
    free(pv[0].v);
    free(pv[1].v);
    return rc;
}

/* ---------------------------------------------------------------- to_vec */

static void fill_sub(uint64_t* data, int32_t px, int32_t py, int32_t pz, size_t side, size_t n,
                     uint64_t v) {
    for (size_t y = (size_t)py; y < (size_t)py + side; y++)
        for (size_t z = (size_t)pz; z < (size_t)pz + side; z++)
            for (size_t x = (size_t)px; x < (size_t)px + side; x++) data[y * n * n + z * n + x] = v;
}

static void to_vec_rec(const vxo_interner* o, uint64_t node, int32_t px, int32_t py, int32_t pz,
                       uint32_t depth, uint32_t max_depth, size_t n, uint64_t* data) {
    if (bid_is_branch(node) && depth < max_depth) {
        int32_t half = 1 << (max_depth - depth - 1);
        for (int i = 7; i >= 0; i--) {
            uint64_t c = o->children[bid_index(node)][i];
            if (c == EMPTY_ID) continue;
            to_vec_rec(o, c, px + (i & 1) * half, py + ((i & 2) >> 1) * half,
                       pz + ((i & 4) >> 2) * half, depth + 1, max_depth, n, data);
        }
    } else {
        uint64_t v = o->values[bid_index(node)];
        if (v != 0) fill_sub(data, px, py, pz, (size_t)1 << (max_depth - depth), n, v);
    }
}

/* utils/common.rs:158-246 (layout y-major: index = y*N*N + z*N + x) */
int vxo_tree_to_vec(const vxo_interner* o, uint64_t root, uint8_t depth, uint64_t* out) {
    size_t n = (size_t)1 << depth;
    size_t size = n * n * n;
    if (!bid_is_branch(root)) {
        uint64_t v = o->values[bid_index(root)];
        for (size_t i = 0; i < size; i++) out[i] = v;
        return 0;
    }
    memset(out, 0, size * sizeof(uint64_t));
    if (root == EMPTY_ID) return 0;
    to_vec_rec(o, root, 0, 0, 0, 0, depth, n, out);
    return 0;
}

/* ---------------------------------------------------------------- canonical hash */
/* Merkle hash of the DAG below a root, independent of node ids:
 *   EMPTY          -> (0, 0)
 *   leaf(v)        -> lane s: mix64(S_s ^ mix64(v + 0x4C454146))
 *   branch         -> lane s: h = mix64(S_s ^ 0x4252414E ^ mix64(value) ^ ((types<<8|mask) << 32));
 *                     for i in 0..8: h = mix64(h ^ (child_s[i] + (i+1) * 0x9E3779B97F4A7C15))
 * with S_0 = 0x736F6D6570736575, S_1 = 0x646F72616E646F6D. */
static const uint64_t CANON_SEED[2] = {0x736F6D6570736575ull, 0x646F72616E646F6Dull};

static void canon(vxo_interner* o, uint64_t node, uint64_t out[2]) {
    if (node == EMPTY_ID) {
        out[0] = out[1] = 0;
        return;
    }
    uint32_t idx = bid_index(node);
    if (o->memo_stamp[idx] == o->stamp) {
        out[0] = o->memo[idx][0];
        out[1] = o->memo[idx][1];
        return;
    }
    uint64_t v = o->values[idx];
    if (bid_is_leaf(node)) {
        for (int s = 0; s < 2; s++) out[s] = mix64(CANON_SEED[s] ^ mix64(v + 0x4C454146ull));
    } else {
        uint64_t tm = ((node >> 47) & 0xFFFFull); /* (types << 8) | mask */
        uint64_t ch[8][2];
        for (int i = 0; i < 8; i++) canon(o, o->children[idx][i], ch[i]);
        for (int s = 0; s < 2; s++) {
            uint64_t h = mix64(CANON_SEED[s] ^ 0x4252414Eull ^ mix64(v) ^ (tm << 32));
            for (int i = 0; i < 8; i++) h = mix64(h ^ (ch[i][s] + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1)));
            out[s] = h;
        }
    }
    o->memo[idx][0] = out[0];
    o->memo[idx][1] = out[1];
    o->memo_stamp[idx] = o->stamp;
}

int vxo_tree_canonical_hash(vxo_interner* o, uint64_t root, uint8_t depth, uint64_t out[2]) {
    (void)depth;
    if (!o->memo) {
        o->memo = (uint64_t(*)[2])calloc(o->capacity, sizeof(uint64_t[2]));
        o->memo_stamp = (uint32_t*)calloc(o->capacity, sizeof(uint32_t));
    }
    o->stamp++;
    if (o->stamp == 0) {
        memset(o->memo_stamp, 0, o->capacity * sizeof(uint32_t));
        o->stamp = 1;
    }
    canon(o, root, out);
    return 0;
}

/* Every live pattern of the interner, independent of node ids: leaf values, and the
 * canonical hash pair of every branch (the EMPTY branch excluded).  Test infrastructure:
 * oracles run in separate processes on disjoint trees give the pattern counts of one
 * shared interner as the union of these sets (equal structure <=> equal pattern). */
int vxo_live_patterns(vxo_interner* o, uint64_t* leaf_values, uint64_t* branch_canon, size_t cap,
                      size_t* n_leaves, size_t* n_branches) {
    if (!o->memo) {
        o->memo = (uint64_t(*)[2])calloc(o->capacity, sizeof(uint64_t[2]));
        o->memo_stamp = (uint32_t*)calloc(o->capacity, sizeof(uint32_t));
    }
    o->stamp++;
    if (o->stamp == 0) {
        memset(o->memo_stamp, 0, o->capacity * sizeof(uint32_t));
        o->stamp = 1;
    }
    size_t nl = 0, nb = 0;
    for (int leaf = 0; leaf < 2; leaf++) {
        const pmap* m = &o->maps[leaf];
        for (size_t k = 0; k < m->cap; k++) {
            uint64_t id = m->id[k];
            if (id == 0 || id == INVALID_ID) continue;
            if (leaf) {
                if (nl < cap) leaf_values[nl] = o->values[bid_index(id)];
                nl++;
            } else {
                uint64_t h[2];
                canon(o, id, h);
                if (nb < cap) {
                    branch_canon[2 * nb] = h[0];
                    branch_canon[2 * nb + 1] = h[1];
                }
                nb++;
            }
        }
    }
    *n_leaves = nl;
    *n_branches = nb;
    return (nl > cap || nb > cap) ? -1 : 0;
}
