// vx_vtm.hip -- VTM snapshot of a resident world: VoxModel::serialize /
// VoxModel::deserialize (world/voxmodel.rs:177-408) and the chunk records
// (world/voxchunk.rs:381-440), with the varint of io/varint.rs.
//
// The data section lists the live patterns renumbered 1.. (leaves first, then
// branches, each in node-index order): a leaf as varint id + value (big-endian),
// a branch as varint id + mask + the varint ids of its non-empty children in slot
// order + its LOD value; then the chunk records.  Export runs on the device: live
// nodes are ranked with prefix sums, every record's size is computed, a second
// scan gives its offset, and one pass writes all records.  Import parses on the
// host (the format is sequential) and links the nodes on the device: child
// reference counts, the branch pattern table and the leaf table.  Included by
// vx_api.hip.

namespace vx {

// unsigned LEB128 (io/varint.rs:5-15)
__host__ __device__ inline uint32_t varint_len(uint64_t v) {
    uint32_t k = 1;
    while (v >= 0x80) {
        v >>= 7;
        k++;
    }
    return k;
}

__host__ __device__ inline uint32_t put_varint(uint8_t* dst, uint64_t v) {
    uint32_t k = 0;
    while (v >= 0x80) {
        dst[k++] = (uint8_t)((v & 0x7F) | 0x80);
        v >>= 7;
    }
    dst[k++] = (uint8_t)v;
    return k;
}

// ---- exclusive scan of uint64 (2048 elements per block, then the block sums)

constexpr uint32_t kScanPer = 8;
constexpr uint32_t kScanBlock = 256;
constexpr uint64_t kScanTile = (uint64_t)kScanPer * kScanBlock;

__global__ __launch_bounds__(kScanBlock) void k_scan_tile(uint64_t* data, uint64_t n, uint64_t* tile_sums) {
    __shared__ uint64_t s[kScanBlock];
    const uint64_t base = blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    uint64_t v[kScanPer];
    uint64_t sum = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanPer; i++) {
        v[i] = base + i < n ? data[base + i] : 0;
        sum += v[i];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kScanBlock; off <<= 1) {
        uint64_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint64_t run = s[threadIdx.x] - sum;
    if (threadIdx.x == kScanBlock - 1) tile_sums[blockIdx.x] = s[kScanBlock - 1];
#pragma unroll
    for (uint32_t i = 0; i < kScanPer; i++) {
        if (base + i < n) data[base + i] = run;
        run += v[i];
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint64_t* data, uint64_t n, const uint64_t* tile_offsets) {
    const uint64_t base = blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
    const uint64_t add = tile_offsets[blockIdx.x];
#pragma unroll
    for (uint32_t i = 0; i < kScanPer; i++)
        if (base + i < n) data[base + i] += add;
}

// Device scratch of one export: carved from a fixed region (the no-allocation entry
// point) or, without a region, hipMalloc'd and freed at the end of the call.
// An export takes 7 arrays plus up to 3 tile-sum arrays for each of its 5 scans (a scan
// recurses once per factor of 2048 elements): 22 at most below 2^33 elements.
struct Arena {
    static constexpr int kHeld = 64;
    uint8_t* base = nullptr;  // nullptr: allocate
    size_t cap = 0, used = 0;
    void* held[kHeld] = {nullptr};
    int nheld = 0;
    bool take(size_t bytes, void** out) {
        *out = nullptr;
        if (!base) {
            if (nheld == kHeld || hipMalloc(out, bytes ? bytes : 1) != hipSuccess) return false;
            held[nheld++] = *out;
            return true;
        }
        size_t at = (used + 255) & ~(size_t)255;
        if (at > cap || bytes > cap - at) return false;
        *out = base + at;
        used = at + bytes;
        return true;
    }
    bool fits(size_t bytes) const {
        const size_t at = (used + 255) & ~(size_t)255;
        return !base || (at <= cap && bytes <= cap - at);
    }
    void release() {
        for (int i = 0; i < nheld; i++) (void)hipFree(held[i]);
        nheld = 0;
    }
};

// In place; data holds n + 1 elements, the last one zero, so data[n] ends as the total.
// The tile sums come from the arena.
inline int scan_u64(uint64_t* data, uint64_t n1, hipStream_t s, Arena* a) {
    const uint64_t tiles = (n1 + kScanTile - 1) / kScanTile;
    uint64_t* sums = nullptr;
    if (!a->take((tiles + 1) * 8, (void**)&sums)) return a->base ? VX_E_SCRATCH : VX_E_HIP;
    hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, data, n1, sums);
    int rc = VX_OK;
    if (tiles > 1) {
        rc = scan_u64(sums, tiles, s, a);
        if (rc == VX_OK) hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, data, n1, sums);
    }
    if (hipGetLastError() != hipSuccess) return VX_E_HIP;
    return rc;
}

// ---- export

// Live patterns: index >= 1 with a reference (the EMPTY branch at index 0 is never written).
template <typename T>
__global__ void k_vtm_count(Pool<T> P, uint64_t N, uint64_t* leaf, uint64_t* branch) {
    VX_LOOP(i, N + 1) {
        if (i <= N) {
            const bool live = i > 0 && i < N && P.refs[i] > 0;
            const bool is_leaf = live && tmg_tm(P.tmg[i]) == 0;
            leaf[i] = is_leaf ? 1 : 0;
            branch[i] = live && !is_leaf ? 1 : 0;
        }
    }
}

// New ids (leaves 1..L, branches L+1..), then every record's size.
template <typename T>
__global__ void k_vtm_ids(Pool<T> P, uint64_t N, const uint64_t* rank_leaf, const uint64_t* rank_branch,
                          uint32_t* newid) {
    const uint64_t L = rank_leaf[N];
    VX_LOOP(i, N) {
        if (i < N) {
            uint32_t id = 0;
            if (rank_leaf[i + 1] != rank_leaf[i]) id = (uint32_t)(1 + rank_leaf[i]);
            else if (rank_branch[i + 1] != rank_branch[i]) id = (uint32_t)(1 + L + rank_branch[i]);
            newid[i] = id;
        }
    }
}

template <typename T>
__global__ void k_vtm_sizes(Pool<T> P, uint64_t N, const uint32_t* newid, uint64_t* size_leaf, uint64_t* size_branch) {
    VX_LOOP(i, N + 1) {
        if (i <= N) {
            uint64_t sl = 0, sb = 0;
            if (i < N && newid[i]) {
                if (tmg_tm(P.tmg[i]) == 0) {
                    sl = varint_len(newid[i]) + sizeof(T);
                } else {
                    sb = varint_len(newid[i]) + 1 + sizeof(T);
                    const uint64_t* ch = P.children + 8ull * i;
                    for (int j = 0; j < 8; j++)
                        if (ch[j] != kEmpty) sb += varint_len(newid[bid_index(ch[j])]);
                }
            }
            size_leaf[i] = sl;
            size_branch[i] = sb;
        }
    }
}

template <typename T>
__device__ inline void put_be(uint8_t* dst, uint32_t v) {
#pragma unroll
    for (int k = 0; k < (int)sizeof(T); k++) dst[k] = (uint8_t)(v >> (8 * ((int)sizeof(T) - 1 - k)));
}

// Leaves at 4 + off_leaf[i], branches at 8 + (leaf bytes) + off_branch[i].
template <typename T>
__global__ void k_vtm_write(Pool<T> P, uint64_t N, const uint32_t* newid, const uint64_t* off_leaf,
                            const uint64_t* off_branch, uint8_t* out) {
    const uint64_t LB = off_leaf[N];
    VX_LOOP(i, N) {
        if (i < N && newid[i]) {
            if (tmg_tm(P.tmg[i]) == 0) {
                uint8_t* d = out + 4 + off_leaf[i];
                uint32_t k = put_varint(d, newid[i]);
                put_be<T>(d + k, (uint32_t)P.values[i]);
            } else {
                uint8_t* d = out + 8 + LB + off_branch[i];
                uint32_t k = put_varint(d, newid[i]);
                d[k++] = (uint8_t)(P.tmg[i] & 0xFFu);
                const uint64_t* ch = P.children + 8ull * i;
                for (int j = 0; j < 8; j++)
                    if (ch[j] != kEmpty) k += put_varint(d + k, newid[bid_index(ch[j])]);
                put_be<T>(d + k, (uint32_t)P.values[i]);
            }
        }
    }
}

__global__ void k_vtm_root_ids(const uint64_t* roots, uint64_t n, const uint32_t* newid, uint64_t N,
                               uint32_t* out) {
    VX_LOOP(t, n) {
        if (t < n) {
            uint64_t r = roots[t];
            out[t] = (r == kEmpty || bid_index(r) >= N) ? 0u : newid[bid_index(r)];
        }
    }
}

// ---- import

// Children references, the branch pattern table and the leaf table of imported
// nodes [1, N) (a fresh interner: no other node exists).
template <typename T>
__global__ void k_vtm_link(Pool<T> P, uint64_t N) {
    VX_LOOP(i, N) {
        if (i >= 1 && i < N) {
            const uint32_t tm = tmg_tm(P.tmg[i]);
            if (tm == 0) {
                leaf_publish(P, (uint32_t)P.values[i], (uint32_t)i);
            } else {
                uint64_t ch[8];
                for (int j = 0; j < 8; j++) {
                    ch[j] = P.children[8ull * i + j];
                    if (ch[j] != kEmpty) atomicAdd(&P.refs[bid_index(ch[j])], 1u);
                }
                if (P.single && tm_is_single(tm)) {  // one-leaf keys live in their index
                    P.single[single_index(P, tm, ch)] = make_branch((uint32_t)i, 0, tm);
                    continue;
                }
                const uint64_t h = branch_hash(ch, tm);
                uint64_t s = h & P.tmask;
                for (uint64_t guard = 0; guard <= P.tmask; guard++) {
                    unsigned long long seen = atomicCAS((unsigned long long*)&P.tab[s], (unsigned long long)kSlotEmpty,
                                                        (unsigned long long)slot_make(slot_tag(h), false, 0, (uint32_t)i));
                    if (seen == kSlotEmpty) break;
                    s = (s + 1) & P.tmask;
                }
            }
        }
    }
}

// The chunk records (voxchunk.rs:381-440) on the device: the chunk count, then per
// chunk the magic, its position (3 x i32 big-endian) and the varint id of its root.
__global__ void k_vtm_tail_sizes(const uint32_t* rid, uint64_t n, uint64_t* tsz) {
    VX_LOOP(t, n + 1) {
        if (t <= n) tsz[t] = t < n ? 24 + varint_len(rid[t]) : 0;
    }
}

__device__ inline void put_be32(uint8_t* d, uint32_t x) {
    d[0] = (uint8_t)(x >> 24);
    d[1] = (uint8_t)(x >> 16);
    d[2] = (uint8_t)(x >> 8);
    d[3] = (uint8_t)x;
}

__global__ void k_vtm_tail_write(const uint32_t* rid, const int32_t* pos, uint64_t n, const uint64_t* toff,
                                 uint8_t* tail) {
    const uint8_t magic[12] = {'V', 'o', 'x', 'T', 'r', 'e', 'e', 'C', 'h', 'u', 'n', 'k'};
    VX_LOOP(t, n + 1) {
        if (t == n) put_be32(tail, (uint32_t)n);
        if (t < n) {
            uint8_t* d = tail + 4 + toff[t];
#pragma unroll
            for (int k = 0; k < 12; k++) d[k] = magic[k];
            for (int k = 0; k < 3; k++) put_be32(d + 12 + 4 * k, (uint32_t)pos[3 * t + k]);
            put_varint(d + 24, rid[t]);
        }
    }
}

__global__ void k_vtm_roots(uint32_t* refs, const uint64_t* roots, uint64_t n) {
    VX_LOOP(t, n) {
        if (t < n && roots[t] != kEmpty) atomicAdd(&refs[bid_index(roots[t])], 1u);
    }
}

__global__ void k_vtm_ctl(Ctl* ctl, uint32_t N, uint64_t leaves, uint64_t branches) {
    ctl->next_index = N;
    ctl->next0 = N;
    ctl->live_leaves = leaves;
    ctl->live_branches = branches;
}

}  // namespace vx

namespace {

using namespace vx;

struct Reader {
    const uint8_t* p;
    size_t n, at = 0;
    bool ok = true;
    uint32_t u32() {
        if (at + 4 > n) {
            ok = false;
            return 0;
        }
        uint32_t x = ((uint32_t)p[at] << 24) | ((uint32_t)p[at + 1] << 16) | ((uint32_t)p[at + 2] << 8) | p[at + 3];
        at += 4;
        return x;
    }
    uint32_t u8() {
        if (at + 1 > n) {
            ok = false;
            return 0;
        }
        return p[at++];
    }
    uint32_t be(uint32_t size) {
        uint32_t x = 0;
        for (uint32_t k = 0; k < size; k++) x = (x << 8) | u8();
        return x;
    }
    // io/varint.rs:63-84: at most 32 bits
    uint32_t varint() {
        uint32_t r = 0, shift = 0;
        for (;;) {
            uint32_t b = u8();
            if (!ok) return 0;
            r |= (b & 0x7Fu) << shift;
            if (!(b & 0x80u)) return r;
            shift += 7;
            if (shift >= 32) {
                ok = false;
                return 0;
            }
        }
    }
};

const uint8_t kVtcMagic[12] = {'V', 'o', 'x', 'T', 'r', 'e', 'e', 'C', 'h', 'u', 'n', 'k'};

// The export (VoxModel::serialize, voxmodel.rs:177-283): live nodes ranked by two scans,
// record sizes, two more scans for their offsets, one pass writing every record; the
// chunk records likewise.  All of it is written on the device into one buffer and copied
// to out once; the two counts are filled in on the host.  Scratch comes from `a`.
template <typename T>
int serialize_core(vx_interner* h, size_t n, const uint64_t* roots, const int32_t* positions, void* out,
                   size_t cap, size_t* size, Arena* a) {
    hipStream_t s = h->stream;
    Ctl c;
    int rc = sync_read_ctl(h, &c);
    if (rc) return rc;
    const uint64_t N = c.next_index;
    Pool<T> P = pool_of<T>(h);
    const int short_of = a->base ? VX_E_SCRATCH : VX_E_HIP;
    uint64_t *rl = nullptr, *rb = nullptr, *tsz = nullptr;
    uint32_t *newid = nullptr, *rid = nullptr;
    int32_t* dpos = nullptr;
    uint8_t* dout = nullptr;
    const size_t words = (size_t)N + 1;
    if (!a->take(words * 8, (void**)&rl) || !a->take(words * 8, (void**)&rb) ||
        !a->take(words * 4, (void**)&newid) || !a->take(std::max<size_t>(n, 1) * 4, (void**)&rid) ||
        !a->take(std::max<size_t>(n, 1) * 12, (void**)&dpos) || !a->take((n + 1) * 8, (void**)&tsz))
        return short_of;
    const unsigned g = grid_for(words);
    hipLaunchKernelGGL(k_vtm_count<T>, dim3(g), dim3(kBlock), 0, s, P, N, rl, rb);
    if ((rc = scan_u64(rl, words, s, a)) || (rc = scan_u64(rb, words, s, a))) return rc;
    hipLaunchKernelGGL(k_vtm_ids<T>, dim3(g), dim3(kBlock), 0, s, P, N, rl, rb, newid);
    uint64_t L = 0, B = 0;
    HIPCHK(hipMemcpyAsync(&L, rl + N, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&B, rb + N, 8, hipMemcpyDeviceToHost, s));
    // the rank arrays become the record sizes, then their offsets
    hipLaunchKernelGGL(k_vtm_sizes<T>, dim3(g), dim3(kBlock), 0, s, P, N, newid, rl, rb);
    if ((rc = scan_u64(rl, words, s, a)) || (rc = scan_u64(rb, words, s, a))) return rc;
    uint64_t LB = 0, BB = 0, TB = 0;
    HIPCHK(hipMemcpyAsync(&LB, rl + N, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&BB, rb + N, 8, hipMemcpyDeviceToHost, s));
    if (n) {
        HIPCHK(hipMemcpyAsync(dpos, positions, n * 12, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_vtm_root_ids, dim3(grid_for(n)), dim3(kBlock), 0, s, roots, (uint64_t)n, newid, N, rid);
    }
    hipLaunchKernelGGL(k_vtm_tail_sizes, dim3(grid_for(n + 1)), dim3(kBlock), 0, s, rid, (uint64_t)n, tsz);
    if ((rc = scan_u64(tsz, n + 1, s, a))) return rc;
    HIPCHK(hipMemcpyAsync(&TB, tsz + n, 8, hipMemcpyDeviceToHost, s));
    if (hipStreamSynchronize(s) != hipSuccess) return VX_E_HIP;
    if (L + B >= 0xFFFFFFFFull) return VX_E_RANGE;
    const size_t total = 8 + LB + BB + 4 + TB;
    *size = total;
    if (!a->fits(total)) return VX_E_SCRATCH;  // a size query tells it as well
    if (!out || cap < total) return out ? VX_E_RANGE : VX_OK;
    if (!a->take(total, (void**)&dout)) return short_of;
    hipLaunchKernelGGL(k_vtm_write<T>, dim3(g), dim3(kBlock), 0, s, P, N, newid, rl, rb, dout);
    hipLaunchKernelGGL(k_vtm_tail_write, dim3(grid_for(n + 1)), dim3(kBlock), 0, s, rid, dpos, (uint64_t)n, tsz,
                       dout + 8 + LB + BB);
    uint8_t* o = (uint8_t*)out;
    HIPCHK(hipMemcpyAsync(o, dout, total, hipMemcpyDeviceToHost, s));
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) return VX_E_HIP;
    // the two counts (the branch count leaves out the EMPTY branch, voxmodel.rs:237)
    for (int k = 0; k < 4; k++) {
        o[k] = (uint8_t)(L >> (24 - 8 * k));
        o[4 + LB + k] = (uint8_t)(B >> (24 - 8 * k));
    }
    return VX_OK;
}

template <typename T>
int serialize_impl(vx_interner* h, size_t n, const uint64_t* roots, const int32_t* positions, void* out,
                   size_t cap, size_t* size) {
    Arena a;  // hipMalloc'd scratch
    int rc = serialize_core<T>(h, n, roots, positions, out, cap, size, &a);
    (void)hipStreamSynchronize(h->stream);
    a.release();
    return rc;
}

// ---- export without allocation

// The scratch of vx_model_serialize_noalloc: one file-scope array per GPU, held until the
// process ends.  Host code never names it directly: its address on each device comes
// from hipGetSymbolAddress, once per device.
constexpr size_t kExportScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kExportScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {nullptr};

// This device's address of `scratch` (the current device); nullptr on failure.
void* scratch_of_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

template <typename T>
int serialize_noalloc_impl(vx_interner* h, size_t n, const uint64_t* roots, const int32_t* positions, void* out,
                           size_t cap, size_t* size) {
    void* base = scratch_of_current_device();
    if (!base) return VX_E_HIP;
    Arena a;
    a.base = (uint8_t*)base;
    a.cap = kExportScratchBytes;
    return serialize_core<T>(h, n, roots, positions, out, cap, size, &a);
}

template <typename T>
int deserialize_impl(vx_interner* h, const uint8_t* data, size_t size, size_t ncap, uint64_t* roots,
                     int32_t* positions, size_t* n_out) {
    hipStream_t s = h->stream;
    Ctl c;
    int rc = sync_read_ctl(h, &c);
    if (rc) return rc;
    if (c.next_index != 1 || c.free_top != 0 || c.live_branches != 0 || c.live_leaves != 0)
        return VX_E_INVALID_ARG;  // only into a fresh interner (interner/mod.rs:930-963 assert the ids)
    Reader r{data, size};
    const uint32_t L = r.u32();
    std::vector<T> val(1);
    std::vector<uint32_t> tm(1, 0);  // generation 0
    val.reserve((size_t)L + 1);
    for (uint32_t i = 0; i < L && r.ok; i++) {
        if (r.varint() != i + 1) return VX_E_INVALID_ARG;
        val.push_back((T)r.be(sizeof(T)));
        tm.push_back(0);
    }
    const uint32_t B = r.u32();
    if (!r.ok || (uint64_t)L + B + 1 > h->capacity) return r.ok ? VX_E_OOM : VX_E_INVALID_ARG;
    const uint64_t N = (uint64_t)L + B + 1;
    std::vector<uint32_t> kids;  // file ids, 8 per branch
    kids.reserve((size_t)B * 8);
    for (uint32_t j = 0; j < B && r.ok; j++) {
        if (r.varint() != L + 1 + j) return VX_E_INVALID_ARG;
        const uint32_t mask = r.u8();
        if (mask == 0) return VX_E_INVALID_ARG;
        uint32_t types = 0;
        for (int k = 0; k < 8; k++) {
            uint32_t id = 0;
            if (mask & (1u << k)) {
                id = r.varint();
                if (id == 0 || id >= N) return VX_E_INVALID_ARG;
                if (id <= L) types |= 1u << k;
            }
            kids.push_back(id);
        }
        val.push_back((T)r.be(sizeof(T)));
        tm.push_back((uint32_t)((types << 8) | mask));
    }
    const uint32_t nch = r.u32();
    if (!r.ok) return VX_E_INVALID_ARG;
    *n_out = nch;
    if (nch > ncap) return VX_E_RANGE;
    auto id_of = [&](uint32_t fid) -> uint64_t {
        if (fid == 0) return kEmpty;
        return fid <= L ? make_leaf(fid, 0) : make_branch(fid, 0, tm[fid]);
    };
    std::vector<uint64_t> root_ids(nch);
    for (uint32_t t = 0; t < nch; t++) {
        for (int k = 0; k < 12; k++)
            if (r.u8() != kVtcMagic[k]) return VX_E_INVALID_ARG;
        for (int k = 0; k < 3; k++) positions[3 * t + k] = (int32_t)r.u32();
        const uint32_t fid = r.varint();
        if (!r.ok || fid >= N) return VX_E_INVALID_ARG;
        root_ids[t] = id_of(fid);
    }
    if (!r.ok) return VX_E_INVALID_ARG;
    std::vector<uint64_t> ch((size_t)N * 8, kEmpty);
    for (uint32_t j = 0; j < B; j++)
        for (int k = 0; k < 8; k++) ch[8ull * (L + 1 + j) + k] = id_of(kids[8ull * j + k]);
    // node columns, then the links on the device
    HIPCHK(hipMemcpyAsync(h->values, val.data(), N * sizeof(T), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(h->tmg, tm.data(), N * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(h->children, ch.data(), N * 64, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemsetAsync(h->refs, 0, N * 4, s));
    Pool<T> P = pool_of<T>(h);
    hipLaunchKernelGGL(k_vtm_link<T>, dim3(grid_for(N)), dim3(kBlock), 0, s, P, N);
    if (nch) {
        HIPCHK(hipMemcpyAsync(roots, root_ids.data(), nch * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_vtm_roots, dim3(grid_for(nch)), dim3(kBlock), 0, s, h->refs, roots, (uint64_t)nch);
    }
    hipLaunchKernelGGL(k_vtm_ctl, dim3(1), dim3(1), 0, s, h->ctl, (uint32_t)N, (uint64_t)L, (uint64_t)B);
    rc = hipStreamSynchronize(s) == hipSuccess && hipGetLastError() == hipSuccess ? VX_OK : VX_E_HIP;
    return rc;
}

}  // namespace

extern "C" {

int vx_model_serialize(vx_interner* h, size_t n, const uint64_t* roots, const int32_t* positions, void* out,
                       size_t cap, size_t* size) {
    if (!h || !size || (n && (!roots || !positions))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    return VX_WIDTH_DISPATCH(h, serialize_impl, h, n, roots, positions, out, cap, size);
}

int vx_model_serialize_noalloc(vx_interner* h, size_t n, const uint64_t* roots, const int32_t* positions,
                               void* out, size_t cap, size_t* size) {
    if (!h || !size || (n && (!roots || !positions))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    return VX_WIDTH_DISPATCH(h, serialize_noalloc_impl, h, n, roots, positions, out, cap, size);
}

int vx_model_deserialize(vx_interner* h, const void* data, size_t size, size_t ncap, uint64_t* roots,
                         int32_t* positions, size_t* n) {
    if (!h || !data || !n || (ncap && (!roots || !positions))) return VX_E_INVALID_ARG;
    if (h->poisoned) return VX_E_POISONED;
    return VX_WIDTH_DISPATCH(h, deserialize_impl, h, (const uint8_t*)data, size, ncap, roots, positions, n);
}

}  // extern "C"
