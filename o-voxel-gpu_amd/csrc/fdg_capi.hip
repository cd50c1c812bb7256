// C ABI of libovx_fdg.so: host orchestration of the flexible-dual-grid
// pipeline on one HIP device.  See include/ovx_fdg.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <cstddef>
#include <mutex>
#include <string>
#include <vector>

#include "ovx_fdg.h"
#include "fdg_kernels.h"
#include "dual_mesh.h"
#include "serialize.h"
#include "svo.h"

using namespace ovx;


namespace
{

thread_local std::string g_last_error;

ovx_status fail(ovx_status code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                              \
    do                                                                                                             \
    {                                                                                                              \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess)                                                                                      \
            return fail(OVX_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

struct DevBuf
{
    void *p = nullptr;
    size_t bytes = 0;
    bool pooled = false;     // stream-ordered allocation (freed stream-ordered too)
    bool in_scratch = false; // carved from the device's static scratch (never freed)
};

// restores the caller's current device on scope exit
struct DeviceGuard
{
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        if (prev != dev)
            (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev)
            (void)hipSetDevice(prev);
    }
};

} // namespace

// Scanline launch classes: (index into MARK_TILES, LDS merging of the
// occupancy bits).  A mesh's class comes from its faces per grid
// cross-section (fpc): meshes of few, large triangles take small tiles and
// merge their dense crossings in LDS; with many more faces than voxel rows
// per plane the crossings are sparse, the merge table only costs occupancy,
// and larger tiles pay.  Each class present is one launch over its meshes.
// The class thresholds are per context (ovx_fdg_set_tile_classes).
constexpr int N_MARK_CLASSES = 4;
constexpr int MARK_CLASS_TILE[N_MARK_CLASSES] = {0, 1, 1, 2};
constexpr bool MARK_CLASS_MERGE[N_MARK_CLASSES] = {true, true, false, false};
static int mark_class(const double t[3], double fpc)
{
    return fpc < t[0] ? 0 : (fpc <= t[1] ? 1 : (fpc < t[2] ? 2 : 3));
}
// face-test class from voxels per face: index into FACE_TILES (32 or 64 faces
// per block walked cooperatively; below t[0], one lane per face)
static int face_class(const double t[2], double vpf)
{
    return vpf >= t[1] ? 0 : (vpf >= t[0] ? 1 : 2);
}

struct ovx_fdg_ctx
{
    int device = 0;
    std::mutex mu;
    // workspace
    DevBuf planes, tris, bits, woff, tiles, counters;
    DevBuf dm_table, dm_cnt, dm_off, dm_mask, dm_tiles; // dual grid -> mesh
    DevBuf dm_direct, dm_touched; // direct-mapped voxel table (all zero between calls) and the cells a call set
    int64_t dm_dirty = 0;         // rows in dm_touched whose cells are still set
    // sparse voxel octrees (.vxz): per-code state of the last encode, neighbour-filter table
    DevBuf svo_insf, svo_k, svo_pos, svo_tiles, svo_rows, svo_erow, nb_table, nb_own, nb_nb, nb_acc, nb_jmp;
    SvoCodes svo{};
    uint64_t svo_total = 0;
    uint8_t *svo_bytes = nullptr; // the octree written by the last encode finish
    int svo_state = 0;            // 1 begun, 2 written
    DevBuf vcnt, voff, nbr, erank, bcnt, boff, bedge_v, bedges, ebins, ehe;
    bool edge_csr = false; // the last conversion's boundary edges came from the CSR lists (a vertex overflowed its slots)
    DevBuf events, recs, chunks, heads, choff, bchunks, tiny, active, medium, large, huge, shard_ctr, rowq, faceq;
    // face and boundary hits, and, once the records are made from them, the
    // per-voxel QEFs and the box-search queue in the same bytes
    DevBuf hitqef;
    size_t off_bhits = 0, off_box = 0, off_qbits = 0; // byte offsets inside hitqef
    uint64_t per[SB_N] = {};       // per-shard capacities of the sharded buffers (grow-only)
    uint64_t force_per[SB_N] = {}; // one-shot overrides (tests, tuning)
    double face_hit_ratio = 1.0;          // face hits of the last conversion over the prep's estimate (capacity estimate)
    double bhit_per_edge = 16.0;          // boundary hits per boundary edge seen last time (capacity estimate)
    double chunk_per_rec = 0.1;           // record chunks per record (decaying maximum of the conversions seen; capacity estimate)
    double rec_per_event = 1.5;           // crossing records (one per crossing and brick) per crossing (same; 1..4)
    uint64_t face_geo = 0;                // face-test hits expected from the faces' sizes (prep of the last begin)
    uint64_t n_events = 0;                // crossings recorded by the last mark pass
    int64_t E = 0;           // boundary edges of the last begin
    uint64_t queue_floor = 1ull << 20;    // least items of the crossing and row-chunk queues
    unsigned char *scratch = nullptr;     // this device's static scratch (ovx_fdg_convert_noalloc)
    unsigned long long *h_ctr = nullptr; // pinned mirror of the counters
    // state between begin and finish
    int pending = 0; // 0 none, 1 fdg, 2 occ, 3 dual mesh
    MeshDev mesh{};
    int64_t V = 0;
    uint64_t n_words = 0;  // occupancy words of all meshes
    uint64_t n_bricks = 0; // bricks of all meshes
    // the batch (a single conversion is a batch of one): host tables, their
    // device copy (segbuf) and the pinned staging of the upload / readback
    int M = 0;
    std::vector<int64_t> face_off, vert_off;
    std::vector<uint64_t> word_off, brick_off;
    std::vector<uint64_t> mtile_off[N_MARK_CLASSES]; // scanline tiles per mesh, per launch class
    const uint64_t *d_mtile_off[N_MARK_CLASSES] = {};
    size_t stage_ftile = 0; // byte offset of the face-tile arrays in the staging and device tables
    size_t stage_gd = 0;    // byte offset of the first guide (after the face-tile arrays), and bytes per guide
    size_t gd_bytes = 0;
    std::vector<uint64_t> ftile_off[N_FACE_TILES]; // face-test tiles per mesh, per tile size
    const uint64_t *d_ftile_off[N_FACE_TILES] = {};
    std::vector<MeshSeg> segs;
    DevBuf segbuf;
    unsigned char *h_stage = nullptr;
    size_t h_stage_bytes = 0;
    uint64_t *h_vox_off = nullptr; // inside h_stage: per-mesh first voxel ids, read back by begin
    BatchDev bdev{};
    double mark_fpc[3] = {0.5, 2.0, 8.0}; // scanline class thresholds (faces per grid cross-section)
    double face_vpf[2] = {4.0, 16.0};     // face-test class thresholds (voxels per face)
    double brick_fpv = 3.0;               // 4x4x4 bricks for meshes of at least this many faces per voxel
    float fw = 0, bw = 0, rw = 0;
    int prec = 0;           // 0 reference float path, 1 compensated voxel-frame accumulation
    float cond_max = 1.0e6f; // precision variant: FP64 solve above this condition estimate
    int64_t N = 0;
    // statistics of the last conversion
    int64_t stats[OVX_FDG_N_STATS] = {};
    hipStream_t last_stream = nullptr; // stream of the last call (the null stream is valid)
    bool ran = false;                 // a call has enqueued work on this context
    hipEvent_t ev_done = nullptr;     // end of the last call's work on last_stream
    bool done_recorded = false;
    // per-kernel timing (events around each launch on the caller's stream)
    bool timing = false;
    hipEvent_t kev[2 * 24] = {};
    hipStream_t side = nullptr; // second stream (default priority): boundary edges beside the scanline stage, crossing and boundary records beside the face test
    // graphs of finish's device work (ovx_fdg_set_graph): the plan key of
    // each (every host-side input of the captured launches), its executable
    // graph and last use; a stream of the context to capture on (the null
    // stream cannot be captured)
    struct GraphEntry
    {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    bool graph_mode = false;
    std::vector<GraphEntry> graphs;
    hipStream_t cap = nullptr;
    uint64_t graph_clock = 0;
    int64_t graph_replays = 0, graph_captures = 0;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool kran[24] = {};
    DualMeshArgs dm{}; // dual grid -> mesh between begin and finish

    // every workspace buffer except the counters
    template <class Fn>
    void for_each_buf(Fn &&fn)
    {
        DevBuf *all[] = {&planes, &tris, &bits,   &woff,  &tiles,   &vcnt,   &voff,  &nbr, &erank, &bcnt, &boff, &ebins, &ehe,
                         &bedge_v, &bedges, &events, &recs,   &chunks, &heads, &choff, &bchunks, &tiny, &active, &medium, &large, &huge,
                         &hitqef,  &shard_ctr, &rowq, &faceq, &segbuf,
                         &dm_table, &dm_cnt, &dm_off, &dm_mask, &dm_tiles, &dm_direct, &dm_touched,
                         &svo_insf, &svo_k, &svo_pos, &svo_tiles, &svo_rows, &svo_erow, &nb_table, &nb_own,
                         &nb_nb, &nb_acc, &nb_jmp};
        for (DevBuf *b : all)
            fn(*b);
    }
};

// Scratch of ovx_fdg_convert_noalloc: one copy of the array per device,
// held until the process ends.  In host code the name stands for no device
// address: the address of the current device's copy comes from
// hipGetSymbolAddress (scratch_for_device), and only that is handed to
// kernels.  This file holds no kernel, so its code object (with the array)
// is loaded only by that lookup.
static __device__ __align__(256) unsigned char scratch[OVX_FDG_NOALLOC_SCRATCH_BYTES];

namespace
{

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The address of `device`'s scratch (that device current), looked up once
// per device.
ovx_status scratch_for_device(int device, unsigned char **out);

// Bump allocator over the device's static scratch, active on this thread
// for the duration of an allocation-free call: workspace buffers are carved
// from it instead of allocated.
struct ScratchArena
{
    unsigned char *base;
    size_t size, used;
};
thread_local ScratchArena *tl_arena = nullptr;

// Stream of the API call in progress on this thread (set by StreamScope):
// workspace growth inside a call is stream-ordered on it, so a buffer is
// freed after the work already queued on the context and no other stream
// of the device waits.
thread_local hipStream_t tl_ws_stream = nullptr;
thread_local bool tl_ws_stream_set = false;

// Every call on a context shares its workspace, whatever stream it runs on:
// a call's first enqueue waits for the end of the previous call's work when
// the streams differ, and the call's end is recorded for the next one.  Work
// of one stream is ordered by the stream itself.
struct StreamScope
{
    ovx_fdg_ctx *c;
    hipStream_t s;
    bool active = false;
    StreamScope(ovx_fdg_ctx *c_, hipStream_t s_) : c(c_), s(s_) {}
    hipError_t enter()
    {
        active = true;
        tl_ws_stream = s;
        tl_ws_stream_set = true;
        if (c->done_recorded && s != c->last_stream)
            return hipStreamWaitEvent(s, c->ev_done, 0);
        return hipSuccess;
    }
    ~StreamScope()
    {
        if (!active)
            return;
        tl_ws_stream_set = false;
        tl_ws_stream = nullptr;
        if (hipEventRecord(c->ev_done, s) == hipSuccess)
            c->done_recorded = true;
        c->last_stream = s;
        c->ran = true;
    }
};

// generation of the workspaces: advances whenever a buffer is allocated or
// released, so a graph captured over the old buffers is never replayed
std::atomic<uint64_t> g_ws_gen{1};

// Grows b to at least `bytes` (grow-only).  `slack` adds 1/8 for buffers
// whose sizes creep up from call to call; buffers sized with their own
// margin take none.
ovx_status ensure(DevBuf &b, size_t bytes, bool slack = true)
{
    if (bytes == 0)
        bytes = 256;
    if (b.bytes >= bytes)
        return OVX_OK;
    if (tl_arena)
    {
        // allocation-free call: a fresh piece of the scratch (a piece carved
        // earlier in the call for this buffer is left unused)
        const size_t off = (tl_arena->used + 255) / 256 * 256;
        if (off > tl_arena->size || bytes > tl_arena->size - off)
            return fail(OVX_ERR_CAPACITY,
                        "the allocation-free conversion needs more than its %zu-byte scratch (%zu bytes in use, "
                        "%zu more requested)",
                        tl_arena->size, off, bytes);
        b.p = tl_arena->base + off;
        b.bytes = bytes;
        b.pooled = false;
        b.in_scratch = true;
        tl_arena->used = off + bytes;
        return OVX_OK;
    }
    g_ws_gen.fetch_add(1);
    if (b.p)
    {
        if (b.pooled)
            HIP_TRY(hipFreeAsync(b.p, tl_ws_stream));
        else
            HIP_TRY(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    const size_t want = (slack ? bytes + bytes / 8 : bytes) + 256;
    if (tl_ws_stream_set)
    {
        HIP_TRY(hipMallocAsync(&b.p, want, tl_ws_stream));
        b.pooled = true;
    }
    else
    {
        HIP_TRY(hipMalloc(&b.p, want));
        b.pooled = false;
    }
    b.bytes = want;
    return OVX_OK;
}

// Frees b; callers have waited for the context's work (stream-ordered
// buffers are freed on the null stream, which has nothing of it left to
// wait for).
void release(DevBuf &b)
{
    if (b.in_scratch)
    {
        b = DevBuf{};
        return;
    }
    g_ws_gen.fetch_add(1);
    if (b.p)
        (void)(b.pooled ? hipFreeAsync(b.p, nullptr) : hipFree(b.p));
    b = DevBuf{};
}

int64_t held_bytes(ovx_fdg_ctx *c)
{
    int64_t s = (int64_t)c->counters.bytes;
    c->for_each_buf([&](DevBuf &b) { s += (int64_t)b.bytes; });
    return s;
}

ovx_status check_grid(const ovx_grid *grid, GridDev &g, uint64_t &n_words)
{
    if (!grid)
        return fail(OVX_ERR_INVALID_ARGUMENT, "grid is null");
    for (int a = 0; a < 3; ++a)
    {
        const float vs = grid->voxel_size[a];
        if (!(vs > 0.0f) || !std::isfinite(vs))
            return fail(OVX_ERR_INVALID_ARGUMENT, "voxel_size[%d] = %g must be positive and finite", a, (double)vs);
        const int64_t lo = grid->grid_range[a], hi = grid->grid_range[3 + a];
        if (hi <= lo)
            return fail(OVX_ERR_INVALID_ARGUMENT, "grid_range max[%d] = %lld must exceed min = %lld", a, (long long)hi,
                        (long long)lo);
        if (hi - lo > MAX_GRID_EXTENT)
            return fail(OVX_ERR_CAPACITY, "grid extent %lld along axis %d exceeds the supported maximum %d",
                        (long long)(hi - lo), a, MAX_GRID_EXTENT);
        if (lo < -(1ll << 29) || hi > (1ll << 29))
            return fail(OVX_ERR_CAPACITY, "grid_range along axis %d is outside [-2^29, 2^29]", a);
        g.vs[a] = vs;
        int ex = 0;
        g.ivs[a] = (std::frexp(vs, &ex) == 0.5f && std::isnormal(std::ldexp(1.0f, 1 - ex))) ? std::ldexp(1.0f, 1 - ex)
                                                                                            : 0.0f;
        g.gmin[a] = (int)lo;
        g.gmax[a] = (int)hi;
        g.G[a] = (int)(hi - lo);
    }
    g.wx = (g.G[0] + 63) / 64;
    n_words = (uint64_t)g.G[2] * (uint64_t)g.G[1] * (uint64_t)g.wx;
    if (n_words >= (1ull << 32))
        return fail(OVX_ERR_CAPACITY, "occupancy bitset of %llu words exceeds 2^32 - 1", (unsigned long long)n_words);
    return OVX_OK;
}

ovx_status read_counters(ovx_fdg_ctx *c, hipStream_t s)
{
    HIP_TRY(hipMemcpyAsync(c->h_ctr, c->counters.p, CTR_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return OVX_OK;
}

ovx_status input_error(unsigned long long flags)
{
    if (flags & ERR_FACE_INDEX)
        return fail(OVX_ERR_INPUT, "faces contain a vertex index outside [0, num_vertices)");
    if (flags & ERR_NONFINITE)
        return fail(OVX_ERR_INPUT, "vertices contain a non-finite coordinate");
    if (flags & ERR_RANGE)
        return fail(OVX_ERR_INPUT, "a vertex lies more than 2^30 voxels from the grid origin");
    if (flags & ERR_DDA)
        return fail(OVX_ERR_INPUT, "boundary traversal exceeded its step bound (non-finite geometry?)");
    return fail(OVX_ERR_INPUT, "invalid input (flags %llx)", flags);
}

// timing slots, one per kernel (or kernel group) of the pipeline
enum KSlot
{
    KS_PREP,
    KS_MARK,
    KS_WORDSCAN,
    KS_EDGES,
    KS_COORDS,
    KS_GEN_EVENTS,
    KS_FACE,
    KS_GEN_FHITS,
    KS_BDRY,
    KS_GEN_BHITS,
    KS_ACTIVE,
    KS_FOLD,
    KS_FOLD_LARGE,
    KS_SOLVE,
    KS_SOLVE_BOX,
    KS_DM_INSERT, // dual grid -> mesh: voxel table
    KS_DM_COUNT,  // vertices and quad counts
    KS_DM_SCAN,
    KS_DM_EMIT,   // quads and triangles
    KS_DM_RESTORE, // direct table: clear the cells of the call
    KS_N
};
static_assert(2 * KS_N <= 48, "timing events");
const char *const kslot_names[KS_N] = {"prep",      "mark",     "word_scan", "edges",   "coords",  "gen_events",
                                       "face",      "gen_fhits", "bdry",     "gen_bhits", "active", "fold",
                                       "fold_large", "solve",   "solve_box", "dm_insert", "dm_count", "dm_scan",
                                       "dm_emit",   "dm_restore"};

void tic(ovx_fdg_ctx *c, int k, hipStream_t s)
{
    if (c->timing)
        (void)hipEventRecord(c->kev[2 * k], s);
}
void toc(ovx_fdg_ctx *c, int k, hipStream_t s)
{
    if (c->timing)
    {
        (void)hipEventRecord(c->kev[2 * k + 1], s);
        c->kran[k] = true;
    }
}

// the boundary-edge bins' arrays (n_bins + 1 entries each, in c->ebins) and
// the half-edges binned by smaller vertex (c->ehe)
struct EdgeBins
{
    uint32_t *cnt, *off, *bcnt, *boff;
    uint2 *he;
};
EdgeBins edge_bins(ovx_fdg_ctx *c, uint64_t n_bins)
{
    auto *p = (uint32_t *)c->ebins.p;
    return EdgeBins{p, p + (n_bins + 1), p + 2 * (n_bins + 1), p + 3 * (n_bins + 1), (uint2 *)c->ehe.p};
}

// ShardBuf view of a buffer of `per` items per shard
template <typename T>
ShardBuf<T> shard_view(ovx_fdg_ctx *c, DevBuf &b, int sb, size_t off = 0)
{
    return ShardBuf<T>{(T *)((unsigned char *)b.p + off), c->per[sb],
                       (unsigned long long *)c->shard_ctr.p + (uint64_t)sb * NSHARD * SHARD_STRIDE};
}

// per-shard capacity for `total` items spread over the shards, with headroom
// (1/margin above the shard mean: 4 -> 25%)
uint64_t per_for(uint64_t total, int margin = 4) { return (total + total / margin) / NSHARD + 256; }

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// the side stream (default priority) and its fork / join events, made once per context
ovx_status ensure_side(ovx_fdg_ctx *c)
{
    if (c->side)
        return OVX_OK;
    HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    return OVX_OK;
}

// Sets the per-shard capacity of sharded buffer sb (grow-only, or a one-shot
// override) and returns the bytes of its regions.
size_t size_shard(ovx_fdg_ctx *c, int sb, uint64_t want_per, size_t item)
{
    if (c->force_per[sb])
    {
        want_per = c->force_per[sb];
        c->force_per[sb] = 0;
        c->per[sb] = want_per;
    }
    else if (want_per > c->per[sb])
        c->per[sb] = want_per;
    return (size_t)c->per[sb] * NSHARD * item;
}

// a sharded buffer of its own (its capacity carries its own margin)
ovx_status ensure_shard(ovx_fdg_ctx *c, DevBuf &b, int sb, uint64_t want_per, size_t item)
{
    return ensure(b, size_shard(c, sb, want_per, item), false);
}

ovx_status clear_shards(ovx_fdg_ctx *c, int sb0, int nsb, hipStream_t s)
{
    HIP_TRY(hipMemsetAsync((unsigned long long *)c->shard_ctr.p + (uint64_t)sb0 * NSHARD * SHARD_STRIDE, 0,
                           (size_t)nsb * NSHARD * SHARD_STRIDE * sizeof(unsigned long long), s));
    return OVX_OK;
}

// Guide of a prefix array of M meshes (BatchDev::gd): the smallest shift
// that makes ((total - 1) >> shift) + 2 entries fit guide_len(M), and entry
// j = the owner of index j << shift (largest m < M with off[m] <= j << shift).
template <typename T>
int build_guide(const T *off, int M, int32_t *gd)
{
    const uint64_t total = (uint64_t)off[M];
    int sh = 0;
    while (total > 0 && (int64_t)((total - 1) >> sh) + 2 > guide_len(M))
        ++sh;
    const int64_t n = total ? (int64_t)((total - 1) >> sh) + 2 : 2;
    int m = 0;
    for (int64_t j = 0; j < n; ++j)
    {
        const uint64_t x = (uint64_t)j << sh;
        while (m + 1 < M && (uint64_t)off[m + 1] <= x)
            ++m;
        gd[j] = m;
    }
    return sh;
}

// Batch tables: per-mesh grids, prefix offsets of faces, vertices, bitset
// words, bricks and block tiles; the device copy is uploaded on stream s.
// Pinned staging bytes of a batch of M meshes (layout in setup_batch): the
// tables above and the read-back voxel offsets.
size_t stage_bytes(int M)
{
    const size_t seg_bytes = ((sizeof(MeshSeg) * (size_t)M + 255) / 256) * 256;
    const size_t arr = (size_t)(M + 1) * 8;
    const size_t garr = (((size_t)guide_len(M) * 4 + 7) / 8) * 8;
    return seg_bytes + (5 + N_MARK_CLASSES + N_FACE_TILES) * arr + N_GUIDES * garr + arr;
}

ovx_status setup_batch(ovx_fdg_ctx *c, const MeshDev &m, int M, const int64_t *face_off, const int64_t *vert_off,
                       const ovx_grid *grids, hipStream_t s)
{
    if (M < 1)
        return fail(OVX_ERR_INVALID_ARGUMENT, "a batch needs at least one mesh");
    if (!face_off || !vert_off || !grids)
        return fail(OVX_ERR_INVALID_ARGUMENT, "batch tables are null");
    if (face_off[0] != 0 || vert_off[0] != 0 || face_off[M] != m.F || vert_off[M] != m.V)
        return fail(OVX_ERR_INVALID_ARGUMENT, "batch offsets must start at 0 and end at the face / vertex counts");
    c->M = M;
    c->face_off.assign(face_off, face_off + M + 1);
    c->vert_off.assign(vert_off, vert_off + M + 1);
    c->word_off.assign(M + 1, 0);
    c->brick_off.assign(M + 1, 0);
    for (auto &v : c->mtile_off)
        v.assign(M + 1, 0);
    for (auto &v : c->ftile_off)
        v.assign(M + 1, 0);
    c->segs.assign(M, MeshSeg{});
    for (int i = 0; i < M; ++i)
    {
        if (face_off[i + 1] < face_off[i] || vert_off[i + 1] < vert_off[i])
            return fail(OVX_ERR_INVALID_ARGUMENT, "batch offsets must not decrease (mesh %d)", i);
        GridDev g{};
        uint64_t nw = 0;
        ovx_status st = check_grid(grids + i, g, nw);
        if (st)
            return st;
        MeshSeg &sg = c->segs[i];
        sg.g = g;
        for (int a = 0; a < 3; ++a)
            sg.nb[a] = (g.G[a] + 7) / 8;
        sg.bs = 3; // 8x8x8 until finish chooses (choose_bricks)
        sg.brick0 = c->brick_off[i];
        sg.face0 = face_off[i];
        const int64_t nf = face_off[i + 1] - face_off[i];
        c->word_off[i + 1] = c->word_off[i] + nw;
        c->brick_off[i + 1] = c->brick_off[i] + (uint64_t)sg.nb[0] * (uint64_t)sg.nb[1] * (uint64_t)sg.nb[2];
        // scanline launch class of the mesh from its faces per grid cross-section
        const double cross = std::max({(double)g.G[0] * g.G[1], (double)g.G[1] * g.G[2], (double)g.G[0] * g.G[2]});
        const int mc = mark_class(c->mark_fpc, (double)nf / cross);
        for (int t = 0; t < N_MARK_CLASSES; ++t)
        {
            const uint64_t tile = (uint64_t)MARK_TILES[MARK_CLASS_TILE[t]];
            c->mtile_off[t][i + 1] = c->mtile_off[t][i] + (t == mc ? (3 * (uint64_t)nf + tile - 1) / tile : 0ull);
        }
        for (int t = 0; t < N_FACE_TILES; ++t)
            c->ftile_off[t][i + 1] = c->ftile_off[t][i] + (uint64_t)(nf + FACE_TILES[t] - 1) / FACE_TILES[t];
    }
    c->n_words = c->word_off[M];
    c->n_bricks = c->brick_off[M];
    if (c->n_words >= (1ull << 32))
        return fail(OVX_ERR_CAPACITY, "occupancy bitset of %llu words exceeds 2^32 - 1", (unsigned long long)c->n_words);
    if (c->n_bricks >= (1ull << 32))
        return fail(OVX_ERR_CAPACITY, "brick grids of %llu bricks exceed 2^32", (unsigned long long)c->n_bricks);
    ovx_status st;
    if ((st = ensure(c->bits, c->n_words * sizeof(uint64_t))) || (st = ensure(c->woff, c->n_words * sizeof(uint32_t))))
        return st;
    for (int i = 0; i < M; ++i)
    {
        c->segs[i].g.bits = (uint64_t *)c->bits.p + c->word_off[i];
        c->segs[i].g.woff = (const uint32_t *)c->woff.p + c->word_off[i];
    }
    // device layout: segments, then 5 + N_MARK_CLASSES + N_FACE_TILES offset
    // arrays of M + 1 entries, then N_GUIDES guides of guide_len(M) entries
    const size_t seg_bytes = ((sizeof(MeshSeg) * (size_t)M + 255) / 256) * 256;
    const size_t arr = (size_t)(M + 1) * 8;
    const size_t garr = (((size_t)guide_len(M) * 4 + 7) / 8) * 8;
    constexpr int FT0 = 5 + N_MARK_CLASSES; // first face-tile array
    const size_t gd0 = seg_bytes + (FT0 + N_FACE_TILES) * arr;
    const size_t total = gd0 + N_GUIDES * garr;
    if (total + arr != stage_bytes(M))
        return fail(OVX_ERR_STATE, "staging layout mismatch (internal error)");
    if (c->h_stage_bytes < total + arr)
    {
        if (tl_arena)
            return fail(OVX_ERR_STATE, "the allocation-free conversion needs the staging its context set up");
        if (c->done_recorded)
            HIP_TRY(hipEventSynchronize(c->ev_done)); // the staging may still be read by a copy
        if (c->h_stage)
            HIP_TRY(hipHostFree(c->h_stage));
        c->h_stage = nullptr;
        c->h_stage_bytes = 0;
        HIP_TRY(hipHostMalloc((void **)&c->h_stage, total + arr));
        c->h_stage_bytes = total + arr;
    }
    if ((st = ensure(c->segbuf, total)))
        return st;
    unsigned char *h = c->h_stage;
    std::memcpy(h, c->segs.data(), sizeof(MeshSeg) * (size_t)M);
    std::memcpy(h + seg_bytes + 0 * arr, c->face_off.data(), arr);
    std::memcpy(h + seg_bytes + 1 * arr, c->vert_off.data(), arr);
    std::memcpy(h + seg_bytes + 2 * arr, c->word_off.data(), arr);
    std::memcpy(h + seg_bytes + 3 * arr, c->brick_off.data(), arr);
    std::memset(h + seg_bytes + 4 * arr, 0, arr); // voxel offsets: written on the device
    for (int t = 0; t < N_MARK_CLASSES; ++t)
        std::memcpy(h + seg_bytes + (5 + t) * arr, c->mtile_off[t].data(), arr);
    for (int t = 0; t < N_FACE_TILES; ++t)
        std::memcpy(h + seg_bytes + (FT0 + t) * arr, c->ftile_off[t].data(), arr);
    int gsh[N_GUIDES] = {};
    gsh[GD_FACE] = build_guide(c->face_off.data(), M, (int32_t *)(h + gd0 + GD_FACE * garr));
    gsh[GD_VERT] = build_guide(c->vert_off.data(), M, (int32_t *)(h + gd0 + GD_VERT * garr));
    gsh[GD_WORD] = build_guide(c->word_off.data(), M, (int32_t *)(h + gd0 + GD_WORD * garr));
    gsh[GD_BRICK] = build_guide(c->brick_off.data(), M, (int32_t *)(h + gd0 + GD_BRICK * garr));
    std::memset(h + gd0 + GD_VOX * garr, 0, garr); // voxel guide: built once the voxel offsets are read back
    c->h_vox_off = (uint64_t *)(h + total);
    HIP_TRY(hipMemcpyAsync(c->segbuf.p, h, total, hipMemcpyHostToDevice, s));
    unsigned char *d = (unsigned char *)c->segbuf.p;
    BatchDev &b = c->bdev;
    b.seg = (const MeshSeg *)d;
    b.face_off = (const int64_t *)(d + seg_bytes + 0 * arr);
    b.vert_off = (const int64_t *)(d + seg_bytes + 1 * arr);
    b.word_off = (const uint64_t *)(d + seg_bytes + 2 * arr);
    b.brick_off = (const uint64_t *)(d + seg_bytes + 3 * arr);
    b.vox_off = (const uint64_t *)(d + seg_bytes + 4 * arr);
    for (int t = 0; t < N_MARK_CLASSES; ++t)
        c->d_mtile_off[t] = (const uint64_t *)(d + seg_bytes + (5 + t) * arr);
    for (int t = 0; t < N_FACE_TILES; ++t)
        c->d_ftile_off[t] = (const uint64_t *)(d + seg_bytes + (FT0 + t) * arr);
    c->stage_ftile = seg_bytes + FT0 * arr;
    c->stage_gd = gd0;
    c->gd_bytes = garr;
    b.mtile_off = c->d_mtile_off[0];
    b.ftile_off = c->d_ftile_off[0];
    for (int k = 0; k < N_GUIDES; ++k)
    {
        b.gd[k] = (const int32_t *)(d + gd0 + k * garr);
        b.gsh[k] = gsh[k];
    }
    b.M = M;
    return OVX_OK;
}

// the scanline pass: one launch per launch class present in the batch
ovx_status launch_mark(ovx_fdg_ctx *c, const MeshDev &m, unsigned long long *ctr, const ShardBuf<Event> &eb,
                       const ShardBuf<RowChunk> &rq, hipStream_t s)
{
    for (int t = 0; t < N_MARK_CLASSES; ++t)
    {
        const uint64_t nt = c->mtile_off[t][c->M];
        if (!nt)
            continue;
        BatchDev b = c->bdev;
        b.mtile_off = c->d_mtile_off[t];
        HIP_TRY(launch_scan_events(m, b, MARK_CLASS_TILE[t], nt, ctr, eb, rq, MARK_CLASS_MERGE[t], s));
    }
    return OVX_OK;
}

// the face test: the face-tile size of each mesh from its voxels per face
// (faces covering many voxels take small tiles, sub-voxel faces larger
// ones); per-size tile offsets are uploaded, then one launch per size
// the face-test class of each mesh from its voxels per face; per-class tile
// offsets staged and uploaded for a batch (the kernels of one mesh read only
// entry 0 of its class's offsets, which is 0 in every class)
ovx_status stage_face_tiles(ovx_fdg_ctx *c, hipStream_t s)
{
    const int M = c->M;
    for (auto &v : c->ftile_off)
        v[0] = 0;
    for (int i = 0; i < M; ++i)
    {
        const int64_t nf = c->face_off[i + 1] - c->face_off[i];
        const double nv = (double)(c->h_vox_off[i + 1] - c->h_vox_off[i]);
        const double vpf = nf > 0 ? nv / (double)nf : 0.0;
        const int ft = face_class(c->face_vpf, vpf);
        for (int t = 0; t < N_FACE_TILES; ++t)
            c->ftile_off[t][i + 1] =
                c->ftile_off[t][i] + (t == ft ? (uint64_t)(nf + FACE_TILES[t] - 1) / FACE_TILES[t] : 0ull);
    }
    // the staging region was last read by the set-up copy, complete since
    // begin's readback; one mesh needs no upload (the kernels read only entry
    // 0 of its class's offsets, which is 0 in every class)
    const size_t arr = (size_t)(M + 1) * 8;
    if (M > 1)
    {
        for (int t = 0; t < N_FACE_TILES; ++t)
            std::memcpy(c->h_stage + c->stage_ftile + t * arr, c->ftile_off[t].data(), arr);
        HIP_TRY(hipMemcpyAsync((unsigned char *)c->segbuf.p + c->stage_ftile, c->h_stage + c->stage_ftile,
                               N_FACE_TILES * arr, hipMemcpyHostToDevice, s));
    }
    return OVX_OK;
}

// the face-test launches of the classes present (offsets staged by stage_face_tiles)
ovx_status launch_face_tiles(ovx_fdg_ctx *c, const MeshDev &m, const ShardBuf<FHit> &fb, unsigned long long *ctr,
                             const ShardBuf<FaceChunk> &fq, hipStream_t s)
{
    const int M = c->M;
    for (int t = 0; t < N_FACE_TILES; ++t)
    {
        const uint64_t nt = c->ftile_off[t][M];
        if (!nt)
            continue;
        BatchDev b = c->bdev;
        b.ftile_off = c->d_ftile_off[t];
        // faces of many voxels put their rows in chunks on a queue that a
        // second launch shares out over the grid (row chunks carry their mesh)
        HIP_TRY(launch_face(m, b, t, nt, fb, ctr, fq, s));
    }
    return OVX_OK;
}

// common begin of a batch: validation, (planes), mark, compaction count;
// n_voxels[0..M) receives each mesh's voxel count
ovx_status begin_common(ovx_fdg_ctx *c, const MeshDev &m_in, int M, const int64_t *face_off, const int64_t *vert_off,
                        const ovx_grid *grids, bool planes, hipStream_t s, int64_t *n_voxels)
{
    MeshDev m = m_in; // gains the gathered triangles once the prep kernel is queued
    c->pending = 0;
    c->N = 0;
    for (int i = 0; i < OVX_FDG_N_STATS; ++i)
        c->stats[i] = 0;
    ovx_status st;
    const int64_t V = m.V;
    if (m.F < 0 || V < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "negative size");
    if (m.F > 0x7fffffffll)
        return fail(OVX_ERR_CAPACITY, "more than 2^31-1 faces");
    if (m.F > 0 && !m.v)
        return fail(OVX_ERR_INVALID_ARGUMENT, "vertices pointer is null");
    if (m.F > 0 && m.f && V == 0)
        return fail(OVX_ERR_INPUT, "faces contain a vertex index outside [0, num_vertices)");
    if ((st = ensure(c->counters, CTR_COUNT * sizeof(unsigned long long))))
        return st;
    HIP_TRY(hipMemsetAsync(c->counters.p, 0, CTR_COUNT * sizeof(unsigned long long), s));
    if ((st = setup_batch(c, m, M, face_off, vert_off, grids, s)))
        return st;
    c->mesh = m;
    c->V = V;
    c->last_stream = s;
    c->ran = true;
    const uint64_t n_words = c->n_words;
    if (m.F == 0)
    {
        c->pending = planes ? 1 : 2;
        for (int i = 0; i < M; ++i)
            n_voxels[i] = 0;
        return OVX_OK;
    }
    if ((st = ensure(c->tiles, (scan_tile_count(n_words) + scan_tile_count((uint64_t)m.F * 3 + 1) + 64) * 8)))
        return st;
    if (planes && (st = ensure(c->planes, (size_t)m.F * sizeof(float4))))
        return st;
    // indexed meshes: the prep kernel gathers each face's vertices once, later
    // stages read them in face order (a soup is in that layout already)
    float *tris = nullptr;
    if (planes && m.f)
    {
        if ((st = ensure(c->tris, (size_t)m.F * 9 * sizeof(float))))
            return st;
        tris = (float *)c->tris.p;
    }
    auto *ctr = (unsigned long long *)c->counters.p;

    for (bool &r : c->kran)
        r = false;
    // boundary edges (edge multiset of FDG.cpp:728-739, kept in std::map order)
    // depend on the faces only: they run on the side stream, beside the
    // scanline and compaction stages, and join before the readback.
    // Half-edges are binned by their smaller vertex (bins of EBIN_VB
    // vertices), each bin's half-edges bucketed by vertex in LDS, each
    // vertex's list sorted in registers, singletons kept; a vertex with more
    // than EDGE_SLOTS half-edges re-runs the CSR lists after the readback.
    c->E = 0;
    bool edges_forked = false;
    const uint64_t n_ebins = ((uint64_t)V + EBIN_VB - 1) / EBIN_VB;
    if (planes && c->bw > 0.0f)
    {
        if ((uint64_t)m.F * 3ull >= (1ull << 32))
            return fail(OVX_ERR_CAPACITY, "more than 2^32 half-edges");
        if ((st = ensure(c->ebins, (size_t)(n_ebins + 1) * 4 * 4)) ||
            (st = ensure(c->ehe, (size_t)m.F * 3 * sizeof(uint2) + 16)) ||
            (st = ensure(c->tiles, (scan_tile_count(n_words) + 2 * scan_tile_count((uint64_t)V + 1) +
                                    scan_tile_count((uint64_t)m.F * 3 + 1) + 64) * 8)))
            return st;
        if ((st = ensure_side(c)))
            return st;
        // the fork point; the edge kernels themselves are queued after the
        // scanline launch, so their host-side queueing does not hold it back
        HIP_TRY(hipEventRecord(c->ev_fork, s)); // counters cleared, inputs in place
        edges_forked = true;
        c->edge_csr = false;
    }
    if ((st = ensure(c->shard_ctr, (size_t)SB_N * NSHARD * SHARD_STRIDE * sizeof(unsigned long long))) ||
        (st = clear_shards(c, SB_FACE_EST, 1, s)))
        return st;
    tic(c, KS_PREP, s);
    HIP_TRY(hipMemsetAsync(c->bits.p, 0, n_words * sizeof(uint64_t), s));
    HIP_TRY(launch_prep(m, c->bdev, planes ? (float4 *)c->planes.p : nullptr, tris, ctr,
                        (unsigned long long *)c->shard_ctr.p + (uint64_t)SB_FACE_EST * NSHARD * SHARD_STRIDE, s));
    toc(c, KS_PREP, s);
    m.tri = tris;
    m.pl = planes ? (const float4 *)c->planes.p : nullptr;
    c->mesh = m;
    // crossings are recorded once, here, when the buffer can hold them
    ShardBuf<Event> eb{};
    if (planes)
    {
        if ((st = ensure_shard(c, c->events, SB_EVENTS, per_for(std::max<uint64_t>(c->queue_floor, 2ull * (uint64_t)m.F)),
                               sizeof(Event))) ||
            (st = clear_shards(c, SB_EVENTS, 1, s)))
            return st;
        eb = shard_view<Event>(c, c->events, SB_EVENTS);
    }
    // row chunks of large triangles, shared out over the grid by a second launch
    if ((st = ensure_shard(c, c->rowq, SB_ROWQ, per_for(std::max<uint64_t>(c->queue_floor, (uint64_t)m.F / 4)),
                           sizeof(RowChunk))))
        return st;
    ShardBuf<RowChunk> rq = shard_view<RowChunk>(c, c->rowq, SB_ROWQ);
    if (M > 1)
        rq.items = nullptr; // row chunks are for single-mesh conversions (a batch has other work to balance)
    tic(c, KS_MARK, s);
    if ((st = launch_mark(c, m, ctr, eb, rq, s)))
        return st;
    toc(c, KS_MARK, s);
    if (edges_forked)
    {
        const hipStream_t es = c->side;
        HIP_TRY(hipStreamWaitEvent(es, c->ev_fork, 0));
        EdgeBins eb = edge_bins(c, n_ebins);
        // its scan tiles sit after the word scan's
        uint64_t *etiles = (uint64_t *)c->tiles.p + scan_tile_count(n_words) + 16;
        tic(c, KS_EDGES, es);
        HIP_TRY(hipMemsetAsync(c->ebins.p, 0, (size_t)(n_ebins + 1) * 4 * 4, es));
        HIP_TRY(hipMemsetAsync(ctr + CTR_EDGE_OVER, 0, sizeof(unsigned long long), es));
        HIP_TRY(launch_ebin(m, false, eb.cnt, eb.off, eb.he, es));
        HIP_TRY(launch_scan_u32(eb.cnt, n_ebins + 1, eb.off, etiles, ctr + CTR_EDGE_TOTAL, es));
        HIP_TRY(launch_ebin(m, true, eb.cnt, eb.off, eb.he, es));
        HIP_TRY(launch_ebin_fold(n_ebins, eb.off, eb.he, eb.bcnt, ctr + CTR_EDGE_OVER, es));
        HIP_TRY(launch_scan_u32(eb.bcnt, n_ebins + 1, eb.boff, etiles, ctr + CTR_BEDGE_TOTAL, es));
        toc(c, KS_EDGES, es);
        HIP_TRY(hipEventRecord(c->ev_join, es));
    }
    tic(c, KS_WORDSCAN, s);
    HIP_TRY(launch_scan_popc((const uint64_t *)c->bits.p, n_words, (uint32_t *)c->woff.p, (uint64_t *)c->tiles.p,
                             ctr + CTR_WORDS_TOTAL, s));
    // first voxel id of each mesh, for the solve and for the caller
    uint64_t *d_vox_off = (uint64_t *)c->bdev.vox_off;
    HIP_TRY(launch_seg_voxels(c->bdev, (const uint32_t *)c->woff.p, n_words, ctr + CTR_WORDS_TOTAL, d_vox_off, s));
    HIP_TRY(hipMemcpyAsync(c->h_vox_off, d_vox_off, (size_t)(M + 1) * 8, hipMemcpyDeviceToHost, s));
    toc(c, KS_WORDSCAN, s);
    if (edges_forked)
        HIP_TRY(hipStreamWaitEvent(s, c->ev_join, 0)); // boundary edges done
    HIP_TRY(launch_shard_reduce((unsigned long long *)c->shard_ctr.p, ctr, s));
    if ((st = read_counters(c, s)))
        return st;
    if (c->h_ctr[CTR_ERRORS])
        return input_error(c->h_ctr[CTR_ERRORS]);
    const unsigned long long N = c->h_ctr[CTR_WORDS_TOTAL];
    if (N >= (1ull << 32) - 1)
        return fail(OVX_ERR_CAPACITY, "%llu occupied voxels exceed the 32-bit voxel index", N);
    c->n_events = planes ? c->h_ctr[CTR_SUM + SB_EVENTS] : 0;
    c->face_geo = planes ? c->h_ctr[CTR_SUM + SB_FACE_EST] : 0;
    if (planes)
    {
        if (c->h_ctr[CTR_MAX + SB_EVENTS] > c->per[SB_EVENTS])
        {
            // a shard was too small: grow and record the crossings again
            // (re-marking the bitset is idempotent)
            const uint64_t want = c->h_ctr[CTR_MAX + SB_EVENTS] + c->h_ctr[CTR_MAX + SB_EVENTS] / 4 + 256;
            if ((st = ensure_shard(c, c->events, SB_EVENTS, want, sizeof(Event))) ||
                (st = clear_shards(c, SB_EVENTS, 1, s)))
                return st;
            ShardBuf<Event> eb2 = shard_view<Event>(c, c->events, SB_EVENTS);
            tic(c, KS_MARK, s);
            if ((st = launch_mark(c, m, nullptr, eb2, rq, s)))
                return st;
            toc(c, KS_MARK, s);
        }
        if (c->bw > 0.0f && c->h_ctr[CTR_EDGE_OVER])
        {
            // a vertex has more half-edges than slots: per-vertex lists of any
            // length (count, scan, place, sort in place), read back again
            if ((st = ensure(c->vcnt, (size_t)(V + 1) * 4)) || (st = ensure(c->bcnt, (size_t)(V + 1) * 4)) ||
                (st = ensure(c->boff, (size_t)(V + 1) * 4)) || (st = ensure(c->voff, (size_t)(V + 1) * 4)) ||
                (st = ensure(c->nbr, (size_t)m.F * 3 * 4)) || (st = ensure(c->erank, (size_t)m.F * 3 * 4)))
                return st;
            c->edge_csr = true;
            auto *vcnt = (uint32_t *)c->vcnt.p;
            auto *voff = (uint32_t *)c->voff.p;
            auto *erank = (uint32_t *)c->erank.p;
            tic(c, KS_EDGES, s);
            HIP_TRY(hipMemsetAsync(vcnt, 0, (size_t)(V + 1) * 4, s));
            HIP_TRY(launch_edge_count(m, vcnt, erank, voff, (uint32_t *)c->nbr.p, 0, s));
            HIP_TRY(launch_scan_u32(vcnt, (uint64_t)V + 1, voff, (uint64_t *)c->tiles.p, ctr + CTR_EDGE_TOTAL, s));
            HIP_TRY(launch_edge_count(m, vcnt, erank, voff, (uint32_t *)c->nbr.p, 1, s));
            HIP_TRY(hipMemsetAsync(c->bcnt.p, 0, (size_t)(V + 1) * 4, s));
            HIP_TRY(
                launch_edge_singletons(V, voff, (uint32_t *)c->nbr.p, (uint32_t *)c->bcnt.p, nullptr, nullptr, 0, s));
            HIP_TRY(launch_scan_u32((uint32_t *)c->bcnt.p, (uint64_t)V + 1, (uint32_t *)c->boff.p,
                                    (uint64_t *)c->tiles.p + scan_tile_count(c->n_words) + 16, ctr + CTR_BEDGE_TOTAL,
                                    s));
            toc(c, KS_EDGES, s);
            if ((st = read_counters(c, s)))
                return st;
        }
        c->E = (int64_t)c->h_ctr[CTR_BEDGE_TOTAL];
        if (c->E > 0)
        {
            if ((st = ensure(c->bedge_v, (size_t)c->E * sizeof(int2))) ||
                (st = ensure(c->bedges, (size_t)c->E * sizeof(BEdge))))
                return st;
            if (c->edge_csr)
                HIP_TRY(launch_edge_singletons(V, (const uint32_t *)c->voff.p, (uint32_t *)c->nbr.p,
                                               (uint32_t *)c->bcnt.p, (const uint32_t *)c->boff.p,
                                               (int2 *)c->bedge_v.p, 1, s));
            else
            {
                const EdgeBins eb = edge_bins(c, n_ebins);
                HIP_TRY(launch_ebin_gather(n_ebins, eb.off, eb.boff, eb.he, (int2 *)c->bedge_v.p, s));
            }
            HIP_TRY(launch_bedge_prep(m.v, c->E, (const int2 *)c->bedge_v.p, (BEdge *)c->bedges.p, s));
        }
    }
    c->N = (int64_t)N;
    c->stats[0] = (int64_t)N;
    c->stats[1] = (int64_t)c->n_events;
    c->stats[4] = c->E;
    c->pending = planes ? 1 : 2;
    for (int i = 0; i < M; ++i)
        n_voxels[i] = (int64_t)(c->h_vox_off[i + 1] - c->h_vox_off[i]);
    return OVX_OK;
}

// Brick size of each mesh for grouping and folding its records: 4x4x4 voxels
// where the faces per voxel reach brick_fpv (records are dense there: an
// 8x8x8 brick would hold more than the fold's tiles), else 8x8x8.  Set-up
// uploaded 8x8x8 for every mesh; the segments, brick offsets and brick guide
// are uploaded again when a mesh changes.
ovx_status choose_bricks(ovx_fdg_ctx *c, hipStream_t s)
{
    const int M = c->M;
    bool changed = false;
    for (int i = 0; i < M; ++i)
    {
        MeshSeg &sg = c->segs[i];
        const double nv = (double)(c->h_vox_off[i + 1] - c->h_vox_off[i]);
        const double nf = (double)(c->face_off[i + 1] - c->face_off[i]);
        const int bs = (nv > 0.0 && nf >= c->brick_fpv * nv) ? 2 : 3;
        changed |= bs != sg.bs;
        sg.bs = bs;
        for (int a = 0; a < 3; ++a)
            sg.nb[a] = (sg.g.G[a] + (1 << bs) - 1) >> bs;
        sg.brick0 = c->brick_off[i];
        c->brick_off[i + 1] = c->brick_off[i] + (uint64_t)sg.nb[0] * (uint64_t)sg.nb[1] * (uint64_t)sg.nb[2];
    }
    c->n_bricks = c->brick_off[M];
    if (c->n_bricks >= (1ull << 32))
        return fail(OVX_ERR_CAPACITY, "brick grids of %llu bricks exceed 2^32", (unsigned long long)c->n_bricks);
    if (!changed)
        return OVX_OK;
    // the staging regions were last read by begin's set-up copy, complete since its readback
    const size_t seg_bytes = ((sizeof(MeshSeg) * (size_t)M + 255) / 256) * 256;
    const size_t arr = (size_t)(M + 1) * 8;
    const size_t gb = c->stage_gd + GD_BRICK * c->gd_bytes;
    unsigned char *h = c->h_stage, *d = (unsigned char *)c->segbuf.p;
    std::memcpy(h, c->segs.data(), sizeof(MeshSeg) * (size_t)M);
    std::memcpy(h + seg_bytes + 3 * arr, c->brick_off.data(), arr);
    c->bdev.gsh[GD_BRICK] = build_guide(c->brick_off.data(), M, (int32_t *)(h + gb));
    HIP_TRY(hipMemcpyAsync(d, h, sizeof(MeshSeg) * (size_t)M, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d + seg_bytes + 3 * arr, h + seg_bytes + 3 * arr, arr, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d + gb, h + gb, c->gd_bytes, hipMemcpyHostToDevice, s));
    return OVX_OK;
}

// The device work of a conversion after its voxel set: coordinates, face
// test, records, grouping, fold and solve.  Every buffer is sized before (no
// allocation here) and nothing waits on the host, so the sequence can be
// captured in a graph; buffers are cleared by fill kernels.  A record-stage
// buffer that overflows sets CTR_OVER, and the grouping, fold and solve
// kernels then return at once (the host sizes the buffers and runs it again).
ovx_status enqueue_finish(ovx_fdg_ctx *c, int32_t *coords, float *dual, uint8_t *flags, uint64_t grouped,
                          hipStream_t s)
{
    ovx_status st;
    const MeshDev &m = c->mesh;
    auto *ctr = (unsigned long long *)c->counters.p;
    const uint64_t n_bricks = c->n_bricks;
    const hipStream_t es = c->side;

    tic(c, KS_COORDS, s);
    HIP_TRY(launch_emit_coords(c->bdev, c->n_words, coords, s));
    toc(c, KS_COORDS, s);
    HIP_TRY(launch_fill_u64((unsigned long long *)c->shard_ctr.p + (uint64_t)SB_FHITS * NSHARD * SHARD_STRIDE,
                            (uint64_t)(SB_N - SB_FHITS) * NSHARD * SHARD_STRIDE, 0ull, s));
    HIP_TRY(launch_fill_u64(c->heads.p, n_bricks + 1, 0ull, s));
    BucketArgs ba{};
    ba.recs = shard_view<Rec>(c, c->recs, SB_RECS);
    ba.chunks = shard_view<Chunk>(c, c->chunks, SB_CHUNKS);
    ba.brick_cnt = (unsigned long long *)c->heads.p;
    ba.b = c->bdev;
    ba.bedge_v = (const int2 *)c->bedge_v.p;
    ba.ctr = ctr;
    // records of the crossings and the boundary DDA on the side stream,
    // beside the face test (which alone needs the main stream's order);
    // the face test is queued first, so the side stream's host-side
    // queueing does not hold it back
    HIP_TRY(hipEventRecord(c->ev_fork, s)); // buffers cleared
    const ShardBuf<FHit> fb = shard_view<FHit>(c, c->hitqef, SB_FHITS);
    if (c->fw > 0.0f)
    {
        tic(c, KS_FACE, s);
        ShardBuf<FaceChunk> fq = shard_view<FaceChunk>(c, c->faceq, SB_FACEQ);
        if ((st = launch_face_tiles(c, m, fb, ctr, fq, s)))
            return st;
        toc(c, KS_FACE, s);
    }
    HIP_TRY(hipStreamWaitEvent(es, c->ev_fork, 0));
    tic(c, KS_GEN_EVENTS, es);
    HIP_TRY(launch_gen_events(shard_view<Event>(c, c->events, SB_EVENTS), ba, es));
    toc(c, KS_GEN_EVENTS, es);
    if (c->E > 0)
    {
        // the boundary DDA makes its records directly; its hits are only counted
        const ShardBuf<BHit> hb = shard_view<BHit>(c, c->hitqef, SB_BHITS, c->off_bhits);
        tic(c, KS_BDRY, es);
        HIP_TRY(launch_bdry(c->E, (const BEdge *)c->bedges.p, (const int2 *)c->bedge_v.p, c->bdev, ba, hb, ctr, es));
        toc(c, KS_BDRY, es);
    }
    HIP_TRY(hipEventRecord(c->ev_join, es));
    if (c->fw > 0.0f)
    {
        tic(c, KS_GEN_FHITS, s);
        HIP_TRY(launch_gen_fhits(fb, ba, s));
        toc(c, KS_GEN_FHITS, s);
    }
    HIP_TRY(hipStreamWaitEvent(s, c->ev_join, 0));
    HIP_TRY(launch_shard_reduce((unsigned long long *)c->shard_ctr.p, ctr, s));
    const OverCaps caps{c->per[SB_FHITS], ~0ull, c->per[SB_RECS], c->per[SB_CHUNKS], grouped}; // boundary hits: counts only
    HIP_TRY(launch_over_check(ctr, caps, s));

    // ---- chunks grouped by brick, bricks with records, then the per-brick ordered fold + solve
    tic(c, KS_ACTIVE, s);
    HIP_TRY(launch_scan_lo32((const unsigned long long *)c->heads.p, n_bricks + 1, (uint32_t *)c->choff.p,
                             (uint64_t *)c->tiles.p, ctr + CTR_CHUNK_SCAN, s));
    HIP_TRY(launch_place_chunks(shard_view<Chunk>(c, c->chunks, SB_CHUNKS), (const uint32_t *)c->choff.p,
                                (uint2 *)c->bchunks.p, ctr + CTR_OVER, s));
    TierLists tl;
    tl.list[0] = (uint32_t *)c->tiny.p;
    tl.list[1] = (uint32_t *)c->active.p;
    tl.list[2] = (uint32_t *)c->medium.p;
    tl.list[3] = (uint32_t *)c->large.p;
    tl.list[4] = (uint32_t *)c->huge.p;
    tl.count[0] = ctr + CTR_TINY;
    tl.count[1] = ctr + CTR_ACTIVE;
    tl.count[2] = ctr + CTR_MEDIUM;
    tl.count[3] = ctr + CTR_LARGE;
    tl.count[4] = ctr + CTR_HUGE;
    HIP_TRY(launch_active_bricks((const unsigned long long *)c->heads.p, n_bricks, tl, ctr + CTR_OVER, s));
    toc(c, KS_ACTIVE, s);
    FoldArgs a{};
    a.choff = (const uint32_t *)c->choff.p;
    a.gchunks = (const uint2 *)c->bchunks.p;
    a.recs = (const Rec *)c->recs.p;
    a.tiny = (const uint32_t *)c->tiny.p;
    a.n_tiny = ctr + CTR_TINY;
    a.active = (uint32_t *)c->active.p;
    a.n_active = ctr + CTR_ACTIVE;
    a.medium = (uint32_t *)c->medium.p;
    a.n_medium = ctr + CTR_MEDIUM;
    a.large = (uint32_t *)c->large.p;
    a.n_large = ctr + CTR_LARGE;
    a.huge = (uint32_t *)c->huge.p;
    a.n_huge = ctr + CTR_HUGE;
    a.b = c->bdev; // grid and brick grid are set per brick from its mesh
    a.planes = (const float4 *)c->planes.p;
    a.bedges = (const BEdge *)c->bedges.p;
    a.dual = dual;
    a.flags = flags;
    a.bw = c->bw;
    a.rw = c->rw;
    // the hits are all made into records by now (the side stream is joined):
    // their bytes hold the QEFs and the box queue
    a.qef = (float *)c->hitqef.p;
    a.solve_bits = c->prec ? (uint32_t *)((unsigned char *)c->hitqef.p + c->off_qbits) : nullptr;
    a.box = shard_view<BoxItem>(c, c->hitqef, SB_BOX, c->off_box);
    a.prec = c->prec;
    a.cond_max = c->cond_max;
    a.n64 = (unsigned long long *)c->shard_ctr.p + (uint64_t)SB_FP64 * NSHARD * SHARD_STRIDE;
    a.ctr = ctr;
    // CTR_TINY, CTR_ACTIVE, CTR_MEDIUM, CTR_LARGE, CTR_LARGE_RECS and CTR_HUGE are zeroed with all counters by begin (an
    // attempt whose record buffers overflow leaves them at zero)
    if (c->prec)
        HIP_TRY(launch_fill_u32(a.solve_bits, ((uint64_t)c->N + 31) / 32, 0u, s)); // the hits' bytes before
    tic(c, KS_FOLD, s);
    HIP_TRY(launch_fold_bricks(a, n_bricks, s));
    toc(c, KS_FOLD, s);
    tic(c, KS_FOLD_LARGE, s);
    HIP_TRY(launch_fold_mid(a, n_bricks, s));
    HIP_TRY(launch_fold_large(a, n_bricks, s));
    toc(c, KS_FOLD_LARGE, s);
    // the reference path's folds solved every voxel; the precision variant solves from its stored QEFs
    if (c->prec)
    {
        tic(c, KS_SOLVE, s);
        HIP_TRY(launch_solve_voxels(a, coords, (uint64_t)c->N, s));
        toc(c, KS_SOLVE, s);
    }
    tic(c, KS_SOLVE_BOX, s);
    HIP_TRY(launch_solve_box(a, s));
    toc(c, KS_SOLVE_BOX, s);
    // totals of the queue regions, for the statistics
    HIP_TRY(launch_shard_reduce((unsigned long long *)c->shard_ctr.p, ctr, s));
    return OVX_OK;
}

// Plan key of finish's device work: every host-side value its launches
// depend on (sizes, buffer and output pointers, weights, capacities, the
// staged per-mesh tables) and the workspace generation.  Equal keys enqueue
// identical launches, so a graph captured for one is replayed for the other.
std::vector<uint64_t> finish_key(ovx_fdg_ctx *c, const int32_t *coords, const float *dual, const uint8_t *flags,
                                 uint64_t grouped)
{
    std::vector<uint64_t> k;
    auto fbits = [](double d) {
        uint64_t u;
        std::memcpy(&u, &d, sizeof u);
        return u;
    };
    const MeshDev &m = c->mesh;
    k.insert(k.end(), {g_ws_gen.load(), (uint64_t)m.v, (uint64_t)m.f, (uint64_t)m.F, (uint64_t)m.V, (uint64_t)c->M,
                       (uint64_t)c->N, c->n_words, c->n_bricks, c->n_events, (uint64_t)c->E, (uint64_t)c->edge_csr,
                       fbits(c->fw), fbits(c->bw), fbits(c->rw), (uint64_t)c->prec, fbits(c->cond_max),
                       (uint64_t)coords, (uint64_t)dual, (uint64_t)flags, grouped, fbits(c->face_vpf[0]),
                       fbits(c->face_vpf[1]), (uint64_t)c->off_bhits, (uint64_t)c->off_box});
    for (int i = 0; i < SB_N; ++i)
        k.push_back(c->per[i]);
    // the staging tables (segments, offsets, guides, face-test tiles, voxel offsets), byte for byte
    const size_t n = c->h_stage_bytes;
    k.push_back((uint64_t)n);
    const size_t w = k.size();
    k.resize(w + (n + 7) / 8, 0ull);
    std::memcpy(k.data() + w, c->h_stage, n);
    return k;
}

// Enqueues finish's device work on s through a graph: replayed when a graph
// of the same plan key exists, otherwise captured on the context's capture
// stream and instantiated first (at most GRAPHS_KEPT graphs are kept, the
// least recently used one is dropped).
constexpr size_t GRAPHS_KEPT = 4;
ovx_status graph_finish(ovx_fdg_ctx *c, int32_t *coords, float *dual, uint8_t *flags, uint64_t grouped,
                        hipStream_t s)
{
    std::vector<uint64_t> key = finish_key(c, coords, dual, flags, grouped);
    ++c->graph_clock;
    for (auto &e : c->graphs)
        if (e.key == key)
        {
            e.used = c->graph_clock;
            ++c->graph_replays;
            HIP_TRY(hipGraphLaunch(e.exec, s));
            return OVX_OK;
        }
    if (!c->cap)
        HIP_TRY(hipStreamCreateWithFlags(&c->cap, hipStreamNonBlocking));
    // nothing runs during the capture; the graph is launched on s, after the work queued there
    HIP_TRY(hipStreamBeginCapture(c->cap, hipStreamCaptureModeThreadLocal));
    ovx_status st = enqueue_finish(c, coords, dual, flags, grouped, c->cap);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(c->cap, &g);
    if (st)
    {
        if (g)
            (void)hipGraphDestroy(g);
        return st;
    }
    if (ee != hipSuccess)
        return fail(OVX_ERR_HIP, "graph capture of the conversion failed: %s", hipGetErrorString(ee));
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess)
        return fail(OVX_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
    if (c->graphs.size() >= GRAPHS_KEPT)
    {
        size_t old = 0;
        for (size_t i = 1; i < c->graphs.size(); ++i)
            if (c->graphs[i].used < c->graphs[old].used)
                old = i;
        (void)hipGraphExecDestroy(c->graphs[old].exec);
        c->graphs.erase(c->graphs.begin() + (long)old);
    }
    c->graphs.push_back(ovx_fdg_ctx::GraphEntry{std::move(key), ex, c->graph_clock});
    ++c->graph_captures;
    HIP_TRY(hipGraphLaunch(ex, s));
    return OVX_OK;
}

void drop_graphs(ovx_fdg_ctx *c)
{
    for (auto &e : c->graphs)
        (void)hipGraphExecDestroy(e.exec);
    c->graphs.clear();
}

ovx_status finish_fdg(ovx_fdg_ctx *c, int32_t *coords, float *dual, uint8_t *flags, hipStream_t s)
{
    ovx_status st;
    const MeshDev &m = c->mesh;
    if ((st = choose_bricks(c, s)))
        return st;
    if (c->M > 1)
    {
        // the voxel guide, from the voxel offsets begin read back
        const size_t g = c->stage_gd + GD_VOX * c->gd_bytes;
        c->bdev.gsh[GD_VOX] = build_guide(c->h_vox_off, c->M, (int32_t *)(c->h_stage + g));
        HIP_TRY(hipMemcpyAsync((unsigned char *)c->segbuf.p + g, c->h_stage + g, c->gd_bytes, hipMemcpyHostToDevice, s));
    }
    if ((st = ensure_side(c)))
        return st;
    if (c->fw > 0.0f && (st = stage_face_tiles(c, s)))
        return st;

    // ---- records grouped by brick (bricks of all meshes, numbered mesh after mesh)
    const uint64_t n_bricks = c->n_bricks;
    // phase-0 records are 1 to 4 per crossing (one per brick of its voxels),
    // estimated from the records per crossing seen before; face hits are
    // estimated from the faces' sizes (prep) scaled by the last conversion's
    // hits over that estimate, boundary hits from the last conversion's hits
    // per boundary edge, the chunks from its chunks per record.  Hit regions
    // take 15% above the shard mean for the uneven spread of the claims; an
    // overflow is detected on the device and the record stages run again
    // with exact sizes.
    const uint64_t face_est = (uint64_t)(c->face_hit_ratio * (double)c->face_geo) + 4096;
    const uint64_t bhit_est = (uint64_t)(c->bhit_per_edge * (double)c->E) + 4096;
    const uint64_t ev_est = std::min<uint64_t>(4ull * c->n_events, (uint64_t)(c->rec_per_event * (double)c->n_events) +
                                                                       c->n_events / 16 + 4096);
    const uint64_t rec_est = ev_est + face_est + bhit_est;
    uint64_t want_rec = per_for(rec_est, 8);
    uint64_t want_fhit = per_for(face_est + face_est * 15 / 100);
    // chunks: the last conversion's chunks per record with a margin (they vary with the order of the claims)
    const double cpr = std::min(1.0, std::max(0.02, c->chunk_per_rec) * 1.5);
    uint64_t want_chunk = per_for((uint64_t)(cpr * (double)rec_est));
    uint64_t grouped = (uint64_t)(cpr * (double)rec_est) + 65536;
    const uint64_t lim = 0xffffffffull / NSHARD; // record offsets and chunk indices are 32-bit
    // at most 3 attempts: hits that overflowed are dropped, so the record
    // counts of the attempt that first holds every hit are the exact ones
    for (int attempt = 0; attempt < 3; ++attempt)
    {
        want_rec = std::min(want_rec, lim);
        if ((st = ensure_shard(c, c->recs, SB_RECS, want_rec, sizeof(Rec))) ||
            (st = ensure_shard(c, c->chunks, SB_CHUNKS, std::min(want_chunk, lim), sizeof(Chunk))) ||
            (st = ensure_shard(c, c->faceq, SB_FACEQ, per_for(std::max<uint64_t>(c->queue_floor, (uint64_t)m.F / 4)),
                               sizeof(FaceChunk))) ||
            (st = ensure(c->heads, (n_bricks + 1) * 8)))
            return st;
        grouped = std::min<uint64_t>(grouped, c->per[SB_CHUNKS] * NSHARD);
        if ((st = ensure(c->bchunks, (grouped + 1) * 8)))
            return st;
        // the whole grouped-chunk buffer is usable: its capacity (stable until it grows) keeps the plan of
        // repeated conversions the same while the chunk count varies with the order of the block claims
        grouped = std::min<uint64_t>(c->bchunks.bytes / 8 - 1, c->per[SB_CHUNKS] * NSHARD);
        if ((st = ensure(c->choff, (n_bricks + 1) * 4)) ||
            (st = ensure(c->tiny, n_bricks * 4)) || (st = ensure(c->active, n_bricks * 4)) ||
            (st = ensure(c->medium, n_bricks * 4)) ||
            (st = ensure(c->large, n_bricks * 4)) ||
            (st = ensure(c->huge, n_bricks * 4)) ||
            // box-search queue: a few percent of the voxels typically; the solve
            // kernel solves in place whatever does not fit
            (st = ensure(c->tiles, (scan_tile_count(n_bricks + 1) + 64) * 8)))
            return st;
        {
            // hits first, then (after the records are made) the box-search queue, which is a few percent of
            // the voxels typically (whatever does not fit is solved in place), after the QEFs of the precision
            // variant (the reference path solves each voxel where it is folded and stores no QEF)
            const size_t fh = align_up(size_shard(c, SB_FHITS, c->fw > 0.0f ? want_fhit : 1, sizeof(FHit)), 256);
            const size_t bh = size_shard(c, SB_BHITS, 1, sizeof(BHit)); // counters only
            // precision variant: quadrics of the voxels its folds leave unsolved, and their flags
            const size_t q = c->prec ? align_up((size_t)c->N * 80, 256) : 0;
            const size_t qb = c->prec ? align_up(((size_t)c->N + 31) / 32 * 4, 256) : 0;
            const size_t bx = size_shard(c, SB_BOX, per_for((uint64_t)c->N / 8 + 4096), sizeof(BoxItem));
            if ((st = ensure(c->hitqef, std::max(fh + bh, q + qb + bx), false)))
                return st;
            c->off_bhits = fh;
            c->off_qbits = q;
            c->off_box = q + qb;
        }
        // first attempt through a graph when enabled (timing events stay outside graphs)
        if (c->graph_mode && !c->timing && attempt == 0)
            st = graph_finish(c, coords, dual, flags, grouped, s);
        else
            st = enqueue_finish(c, coords, dual, flags, grouped, s);
        if (st)
            return st;
        // the conversion's one read-back of finish: errors, overflow, statistics
        if ((st = read_counters(c, s)))
            return st;
        c->stats[14] = attempt + 1;
        if (c->h_ctr[CTR_ERRORS])
            return input_error(c->h_ctr[CTR_ERRORS]);
        const unsigned long long *mx = c->h_ctr + CTR_MAX, *sum = c->h_ctr + CTR_SUM;
        if (!c->h_ctr[CTR_OVER])
        {
            if (c->fw > 0.0f && c->face_geo)
                c->face_hit_ratio = std::min(4.0, std::max(0.05, (double)sum[SB_FHITS] / (double)c->face_geo));
            if (c->E > 0)
                c->bhit_per_edge = std::min(1.0e4, std::max(1.0, (double)sum[SB_BHITS] / (double)c->E));
            c->stats[2] = (int64_t)sum[SB_RECS];
            c->stats[3] = (int64_t)sum[SB_FHITS];
            c->stats[5] = (int64_t)sum[SB_BHITS];
            c->stats[9] = (int64_t)sum[SB_CHUNKS];
            // a slowly decaying maximum: conversions of other meshes on the context do not overflow at once
            if (sum[SB_RECS])
                c->chunk_per_rec = std::max((double)sum[SB_CHUNKS] / (double)sum[SB_RECS], 0.9 * c->chunk_per_rec);
            // every hit made one record: the rest are the crossings'
            if (c->n_events && sum[SB_RECS] >= sum[SB_FHITS] + sum[SB_BHITS])
                c->rec_per_event = std::min(
                    4.0, std::max({1.0, (double)(sum[SB_RECS] - sum[SB_FHITS] - sum[SB_BHITS]) / (double)c->n_events,
                                   0.9 * c->rec_per_event}));
            return OVX_OK;
        }
        const bool hits_over = mx[SB_FHITS] > c->per[SB_FHITS];
        if (sum[SB_RECS] >= 0xffffffffull || sum[SB_CHUNKS] >= 0xffffffffull)
            return fail(OVX_ERR_CAPACITY, "%llu QEF records / %llu chunks exceed the 32-bit record index",
                        (unsigned long long)sum[SB_RECS], (unsigned long long)sum[SB_CHUNKS]);
        if (attempt == 2)
            return fail(OVX_ERR_STATE, "record buffers overflowed three times (internal error)");
        // sizes are known now: grow and generate again
        want_fhit = mx[SB_FHITS] + mx[SB_FHITS] / 8 + 256;
        want_rec = mx[SB_RECS] + mx[SB_RECS] / 8 + 256;
        if (hits_over)
        {
            // records of dropped hits were never requested: size records and
            // chunks from all the hits (every hit was counted)
            const uint64_t all = 4ull * c->n_events + sum[SB_FHITS] + sum[SB_BHITS];
            want_rec = std::max(want_rec, per_for(all, 8));
            want_chunk = std::max(want_chunk, per_for((uint64_t)(cpr * (double)all)));
            grouped = std::max<uint64_t>(grouped, (uint64_t)(cpr * (double)all) + 65536);
        }
        else
        {
            // every chunk of this attempt was counted (dropped ones included)
            want_chunk = std::max<uint64_t>(want_chunk, mx[SB_CHUNKS] + mx[SB_CHUNKS] / 4 + 256);
            grouped = std::max<uint64_t>(grouped, sum[SB_CHUNKS] + sum[SB_CHUNKS] / 8 + 256);
        }
    }
    return OVX_OK;
}

ovx_status scratch_for_device(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= SCRATCH_MAX_DEVICES)
        return fail(OVX_ERR_INVALID_ARGUMENT, "device %d beyond the %d devices with a scratch", device,
                    SCRATCH_MAX_DEVICES);
    if (!g_scratch_ptr[device])
    {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OVX_ERR_HIP, "address of the device scratch: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return OVX_OK;
}

// Every workspace buffer of a context, its sharded capacities and its graph
// mode, parked while an allocation-free call runs on the scratch.
struct ParkedWorkspace
{
    static constexpr int MAX_BUFS = 64;
    DevBuf bufs[MAX_BUFS];
    DevBuf counters;
    uint64_t per[SB_N], force_per[SB_N];
    bool graph_mode;
    uint64_t queue_floor;
    int n = 0;

    void park(ovx_fdg_ctx *c)
    {
        n = 0;
        c->for_each_buf([&](DevBuf &b) {
            bufs[n++] = b;
            b = DevBuf{};
        });
        counters = c->counters;
        c->counters = DevBuf{};
        for (int i = 0; i < SB_N; ++i)
        {
            per[i] = c->per[i];
            force_per[i] = c->force_per[i];
            c->per[i] = 0;
            c->force_per[i] = 0;
        }
        graph_mode = c->graph_mode;
        c->graph_mode = false; // graphs are for the context's own buffers
        queue_floor = c->queue_floor;
        c->queue_floor = 1ull << 14; // small queues: a full queue only moves work back to its producer
    }
    void restore(ovx_fdg_ctx *c)
    {
        int i = 0;
        c->for_each_buf([&](DevBuf &b) { b = bufs[i++]; });
        c->counters = counters;
        for (int k = 0; k < SB_N; ++k)
        {
            c->per[k] = per[k];
            c->force_per[k] = force_per[k];
        }
        c->graph_mode = graph_mode;
        c->queue_floor = queue_floor;
    }
};

} // namespace

extern "C" {

const char *ovx_version(void) { return "ovx_fdg 0.1.0 gfx950"; }

const char *ovx_last_error(void) { return g_last_error.c_str(); }

ovx_status ovx_fdg_create(int device, ovx_fdg_ctx **out)
{
    if (!out)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx out-pointer is null");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(OVX_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, n);
    DeviceGuard dg(device);
    auto *c = new ovx_fdg_ctx();
    c->device = device;
    hipError_t e = hipHostMalloc((void **)&c->h_ctr, CTR_COUNT * sizeof(unsigned long long), hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming);
    // the staging of a single-mesh conversion and the side stream, so that
    // ovx_fdg_convert_noalloc finds them in place
    if (e == hipSuccess && (e = hipHostMalloc((void **)&c->h_stage, stage_bytes(1))) == hipSuccess)
        c->h_stage_bytes = stage_bytes(1);
    if (e != hipSuccess)
    {
        if (c->h_ctr)
            (void)hipHostFree(c->h_ctr);
        if (c->h_stage)
            (void)hipHostFree(c->h_stage);
        if (c->ev_done)
            (void)hipEventDestroy(c->ev_done);
        delete c;
        return fail(OVX_ERR_HIP, "context set-up failed: %s", hipGetErrorString(e));
    }
    ovx_status st = ensure_side(c);
    if (st)
    {
        const std::string msg = g_last_error;
        (void)ovx_fdg_destroy(c);
        return fail(st, "%s", msg.c_str());
    }
    *out = c;
    return OVX_OK;
}

ovx_status ovx_fdg_destroy(ovx_fdg_ctx *c)
{
    if (!c)
        return OVX_OK;
    {
        DeviceGuard dg(c->device);
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->done_recorded)
            (void)hipEventSynchronize(c->ev_done);
        c->for_each_buf([](DevBuf &b) { release(b); });
        release(c->counters);
        drop_graphs(c);
        if (c->cap)
            (void)hipStreamDestroy(c->cap);
        if (c->ev_done)
            (void)hipEventDestroy(c->ev_done);
        if (c->h_ctr)
            (void)hipHostFree(c->h_ctr);
        if (c->h_stage)
            (void)hipHostFree(c->h_stage);
        for (auto &e : c->kev)
            if (e)
                (void)hipEventDestroy(e);
        if (c->side)
        {
            (void)hipStreamSynchronize(c->side);
            (void)hipEventDestroy(c->ev_fork);
            (void)hipEventDestroy(c->ev_join);
            (void)hipStreamDestroy(c->side);
        }
    }
    delete c;
    return OVX_OK;
}

ovx_status ovx_fdg_prepare_mesh(const float *d_vertices, const float *d_origin, int64_t n_vertices,
                                const int64_t *d_faces, int64_t n_faces, float *d_out_vertices, int32_t *d_out_faces,
                                void *stream)
{
    if (n_vertices < 0 || n_faces < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "negative vertex or face count");
    if (n_vertices > (1ll << 40) || n_faces > (1ll << 40))
        return fail(OVX_ERR_CAPACITY, "vertex or face count beyond 2^40");
    if (n_vertices > 0 && (!d_vertices || !d_origin || !d_out_vertices))
        return fail(OVX_ERR_INVALID_ARGUMENT, "vertex, origin or output vertex pointer is null");
    if (d_faces && n_faces > 0 && !d_out_faces)
        return fail(OVX_ERR_INVALID_ARGUMENT, "output face pointer is null");
    const uint64_t nf3 = d_faces ? 3ull * (uint64_t)n_faces : 0ull;
    if (n_vertices == 0 && nf3 == 0)
        return OVX_OK;
    HIP_TRY(launch_prepare_mesh(d_vertices, d_origin, (uint64_t)n_vertices, d_out_vertices, d_faces, nf3, d_out_faces,
                                (hipStream_t)stream));
    return OVX_OK;
}

static_assert(sizeof(MeshSrc) == sizeof(ovx_mesh_src) && offsetof(MeshSrc, origin) == offsetof(ovx_mesh_src, origin) &&
                  offsetof(MeshSrc, n_vertices) == offsetof(ovx_mesh_src, n_vertices) &&
                  offsetof(MeshSrc, face_offset) == offsetof(ovx_mesh_src, face_offset) &&
                  offsetof(MeshSrc, faces_int64) == offsetof(ovx_mesh_src, faces_int64),
              "ovx_mesh_src layout");

ovx_status ovx_fdg_batch_prepare(const ovx_mesh_src *d_meshes, int n_meshes, int64_t total_vertices,
                                 int64_t total_faces, float *d_out_vertices, int32_t *d_out_faces, void *stream)
{
    if (n_meshes < 0 || total_vertices < 0 || total_faces < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "negative mesh, vertex or face count");
    if (n_meshes == 0 || (total_vertices == 0 && total_faces == 0))
        return OVX_OK;
    if (!d_meshes)
        return fail(OVX_ERR_INVALID_ARGUMENT, "mesh table pointer is null");
    if (total_vertices > (int64_t)INT32_MAX)
        return fail(OVX_ERR_CAPACITY, "%lld vertices: packed face indices are 32-bit", (long long)total_vertices);
    if (total_faces > (1ll << 40))
        return fail(OVX_ERR_CAPACITY, "face count beyond 2^40");
    if ((total_vertices > 0 && !d_out_vertices) || (total_faces > 0 && !d_out_faces))
        return fail(OVX_ERR_INVALID_ARGUMENT, "output pointer is null");
    HIP_TRY(launch_batch_prepare((const MeshSrc *)d_meshes, n_meshes, (uint64_t)total_vertices, (uint64_t)total_faces,
                                 d_out_vertices, d_out_faces, (hipStream_t)stream));
    return OVX_OK;
}

int64_t ovx_fdg_workspace_bytes(const ovx_fdg_ctx *c) { return c ? held_bytes(const_cast<ovx_fdg_ctx *>(c)) : 0; }

ovx_status ovx_fdg_trim(ovx_fdg_ctx *c)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->done_recorded)
        HIP_TRY(hipEventSynchronize(c->ev_done));
    drop_graphs(c);
    c->for_each_buf([](DevBuf &b) { release(b); });
    for (int i = 0; i < SB_N; ++i)
        c->per[i] = 0;
    c->pending = 0;
    c->dm_dirty = 0; // the direct table is re-zeroed when it is allocated again
    return OVX_OK;
}

ovx_status ovx_fdg_begin(ovx_fdg_ctx *c, const float *d_vertices, int64_t n_vertices, const int32_t *d_faces,
                         int64_t n_faces, const ovx_grid *grid, float face_weight, float boundary_weight,
                         float regularization_weight, void *stream, int64_t *n_voxels)
{
    if (!c || !n_voxels)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx / n_voxels is null");
    if (n_faces > 0 && !d_faces)
        return fail(OVX_ERR_INVALID_ARGUMENT, "faces pointer is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    MeshDev m{d_vertices, d_faces, n_faces, n_vertices};
    c->fw = face_weight;
    c->bw = boundary_weight;
    c->rw = regularization_weight;
    const int64_t fo[2] = {0, n_faces}, vo[2] = {0, n_vertices};
    StreamScope ss(c, (hipStream_t)stream);
    HIP_TRY(ss.enter());
    return begin_common(c, m, 1, fo, vo, grid, true, (hipStream_t)stream, n_voxels);
}

ovx_status ovx_fdg_batch_begin(ovx_fdg_ctx *c, const float *d_vertices, int64_t n_vertices, const int32_t *d_faces,
                               int64_t n_faces, int n_meshes, const int64_t *face_offsets,
                               const int64_t *vertex_offsets, const ovx_grid *grids, float face_weight,
                               float boundary_weight, float regularization_weight, void *stream, int64_t *n_voxels)
{
    if (!c || !n_voxels)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx / n_voxels is null");
    if (n_faces > 0 && !d_faces)
        return fail(OVX_ERR_INVALID_ARGUMENT, "faces pointer is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    MeshDev m{d_vertices, d_faces, n_faces, n_vertices};
    c->fw = face_weight;
    c->bw = boundary_weight;
    c->rw = regularization_weight;
    StreamScope ss(c, (hipStream_t)stream);
    HIP_TRY(ss.enter());
    return begin_common(c, m, n_meshes, face_offsets, vertex_offsets, grids, true, (hipStream_t)stream, n_voxels);
}

ovx_status ovx_fdg_finish(ovx_fdg_ctx *c, int32_t *d_coords, float *d_dual, uint8_t *d_intersected, void *stream)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->pending != 1)
        return fail(OVX_ERR_STATE, "ovx_fdg_finish without a successful ovx_fdg_begin");
    c->pending = 0;
    if (c->N == 0)
        return OVX_OK;
    if (!d_coords || !d_dual || !d_intersected)
        return fail(OVX_ERR_INVALID_ARGUMENT, "output pointer is null");
    StreamScope ss(c, (hipStream_t)stream);
    HIP_TRY(ss.enter());
    return finish_fdg(c, d_coords, d_dual, d_intersected, (hipStream_t)stream);
}

ovx_status ovx_fdg_convert_noalloc(ovx_fdg_ctx *c, const float *d_vertices, int64_t n_vertices,
                                   const int32_t *d_faces, int64_t n_faces, const ovx_grid *grid, float face_weight,
                                   float boundary_weight, float regularization_weight, int32_t *d_coords,
                                   float *d_dual, uint8_t *d_intersected, int64_t max_voxels, void *stream,
                                   int64_t *n_voxels)
{
    if (!c || !n_voxels)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx / n_voxels is null");
    if (n_faces > 0 && !d_faces)
        return fail(OVX_ERR_INVALID_ARGUMENT, "faces pointer is null");
    if (max_voxels < 0 || (max_voxels > 0 && (!d_coords || !d_dual || !d_intersected)))
        return fail(OVX_ERR_INVALID_ARGUMENT, "output pointers / max_voxels invalid");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    // the device's scratch, looked up by the first allocation-free call on
    // the context (its device current): the code object holding the array
    // is loaded, and the array placed, then, so that a process that makes no
    // such call does not pay for it (about 25 ms for 64 MiB)
    if (!c->scratch)
    {
        const ovx_status sst = scratch_for_device(c->device, &c->scratch);
        if (sst)
            return sst;
    }
    *n_voxels = 0;
    MeshDev m{d_vertices, d_faces, n_faces, n_vertices};
    c->fw = face_weight;
    c->bw = boundary_weight;
    c->rw = regularization_weight;
    const int64_t fo[2] = {0, n_faces}, vo[2] = {0, n_vertices};
    ParkedWorkspace parked;
    parked.park(c);
    ScratchArena arena{c->scratch, (size_t)OVX_FDG_NOALLOC_SCRATCH_BYTES, 0};
    tl_arena = &arena;
    ovx_status st = OVX_OK;
    int64_t N = 0;
    {
        StreamScope ss(c, (hipStream_t)stream);
        const hipError_t e = ss.enter();
        if (e != hipSuccess)
            st = fail(OVX_ERR_HIP, "stream ordering: %s", hipGetErrorString(e));
        if (!st)
            st = begin_common(c, m, 1, fo, vo, grid, true, (hipStream_t)stream, &N);
        if (!st && N > max_voxels)
            st = fail(OVX_ERR_CAPACITY, "%lld occupied voxels exceed the output capacity of %lld", (long long)N,
                      (long long)max_voxels);
        if (!st && N > 0)
            st = finish_fdg(c, d_coords, d_dual, d_intersected, (hipStream_t)stream);
    }
    tl_arena = nullptr;
    c->pending = 0;
    parked.restore(c);
    if (st)
        return st;
    *n_voxels = N;
    return OVX_OK;
}

ovx_status ovx_occ_begin(ovx_fdg_ctx *c, const float *d_triangles, int64_t n_triangles, const ovx_grid *grid,
                         void *stream, int64_t *n_voxels)
{
    if (!c || !n_voxels)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx / n_voxels is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    MeshDev m{d_triangles, nullptr, n_triangles, 3 * n_triangles};
    const int64_t fo[2] = {0, n_triangles}, vo[2] = {0, 3 * n_triangles};
    StreamScope ss(c, (hipStream_t)stream);
    HIP_TRY(ss.enter());
    return begin_common(c, m, 1, fo, vo, grid, false, (hipStream_t)stream, n_voxels);
}

ovx_status ovx_occ_finish(ovx_fdg_ctx *c, int32_t *d_coords, void *stream)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->pending != 2)
        return fail(OVX_ERR_STATE, "ovx_occ_finish without a successful ovx_occ_begin");
    c->pending = 0;
    if (c->N == 0)
        return OVX_OK;
    if (!d_coords)
        return fail(OVX_ERR_INVALID_ARGUMENT, "output pointer is null");
    StreamScope ss(c, (hipStream_t)stream);
    HIP_TRY(ss.enter());
    HIP_TRY(launch_emit_coords(c->bdev, c->n_words, d_coords, (hipStream_t)stream));
    return OVX_OK;
}

ovx_status ovx_fdg_stats(ovx_fdg_ctx *c, int64_t *out, int n)
{
    if (!c || !out || n < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid stats arguments");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->ran && c->counters.p)
    {
        ovx_status st = read_counters(c, c->last_stream);
        if (st)
            return st;
        c->stats[7] = (int64_t)c->h_ctr[CTR_GIANT];
        c->stats[8] = (int64_t)(c->h_ctr[CTR_TINY] + c->h_ctr[CTR_ACTIVE] + c->h_ctr[CTR_MEDIUM] + c->h_ctr[CTR_LARGE] +
                                c->h_ctr[CTR_HUGE]);
        c->stats[10] = (int64_t)(c->h_ctr[CTR_LARGE] + c->h_ctr[CTR_HUGE]);
        c->stats[11] = (int64_t)c->h_ctr[CTR_LARGE_RECS];
        c->stats[12] = (int64_t)c->h_ctr[CTR_SUM + SB_BOX];
        c->stats[13] = c->prec ? (int64_t)c->h_ctr[CTR_SUM + SB_FP64] : 0;
    }
    c->stats[6] = held_bytes(c);
    for (int i = 0; i < n && i < OVX_FDG_N_STATS; ++i)
        out[i] = c->stats[i];
    return OVX_OK;
}

ovx_status ovx_fdg_set_timing(ovx_fdg_ctx *c, int enable)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    c->timing = enable != 0;
    if (c->timing)
        for (auto &e : c->kev)
            if (!e)
                HIP_TRY(hipEventCreate(&e));
    return OVX_OK;
}

ovx_status ovx_fdg_set_capacity(ovx_fdg_ctx *c, int64_t events, int64_t hits)
{
    if (!c || events < 0 || hits < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid capacity arguments");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->ran)
        HIP_TRY(hipStreamSynchronize(c->last_stream));
    // capacities are per shard; at least one item per shard
    if (events)
    {
        release(c->events);
        c->per[SB_EVENTS] = 0;
        c->force_per[SB_EVENTS] = std::max<uint64_t>(1, (uint64_t)events / NSHARD);
    }
    if (hits)
    {
        for (int sb : {SB_RECS, SB_CHUNKS, SB_FHITS})
        {
            c->per[sb] = 0;
            c->force_per[sb] = std::max<uint64_t>(1, (uint64_t)hits / NSHARD);
        }
        release(c->recs);
        release(c->chunks);
        release(c->hitqef);
    }
    return OVX_OK;
}

ovx_status ovx_fdg_set_precision(ovx_fdg_ctx *c, int mode, float cond_max)
{
    if (!c || (mode != OVX_PREC_REFERENCE && mode != OVX_PREC_COMPENSATED) || !(cond_max >= 1.0f))
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid precision arguments");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->pending)
        return fail(OVX_ERR_STATE, "precision changed between begin and finish");
    c->prec = mode;
    c->cond_max = cond_max;
    return OVX_OK;
}

ovx_status ovx_fdg_set_tile_classes(ovx_fdg_ctx *c, const double *mark_fpc3, const double *face_vpf2)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null context");
    auto ok = [](const double *t, int n) {
        for (int i = 0; i < n; ++i)
            if (!(t[i] >= 0.0 && t[i] <= 1e30) || (i && t[i] < t[i - 1]))
                return false;
        return true;
    };
    if ((mark_fpc3 && !ok(mark_fpc3, 3)) || (face_vpf2 && !ok(face_vpf2, 2)))
        return fail(OVX_ERR_INVALID_ARGUMENT, "tile class thresholds must be finite, >= 0 and non-decreasing");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->pending)
        return fail(OVX_ERR_STATE, "tile classes changed between begin and finish");
    if (mark_fpc3)
        for (int i = 0; i < 3; ++i)
            c->mark_fpc[i] = mark_fpc3[i];
    if (face_vpf2)
        for (int i = 0; i < 2; ++i)
            c->face_vpf[i] = face_vpf2[i];
    return OVX_OK;
}

ovx_status ovx_fdg_get_tile_classes(ovx_fdg_ctx *c, double *mark_fpc3, double *face_vpf2)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null context");
    std::lock_guard<std::mutex> lk(c->mu);
    if (mark_fpc3)
        for (int i = 0; i < 3; ++i)
            mark_fpc3[i] = c->mark_fpc[i];
    if (face_vpf2)
        for (int i = 0; i < 2; ++i)
            face_vpf2[i] = c->face_vpf[i];
    return OVX_OK;
}

ovx_status ovx_fdg_set_brick_class(ovx_fdg_ctx *c, double faces_per_voxel)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null context");
    if (!(faces_per_voxel >= 0.0 && faces_per_voxel <= 1e30))
        return fail(OVX_ERR_INVALID_ARGUMENT, "brick class threshold must be finite and >= 0");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->pending)
        return fail(OVX_ERR_STATE, "brick class changed between begin and finish");
    c->brick_fpv = faces_per_voxel;
    return OVX_OK;
}

ovx_status ovx_fdg_get_brick_class(ovx_fdg_ctx *c, double *faces_per_voxel)
{
    if (!c || !faces_per_voxel)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    *faces_per_voxel = c->brick_fpv;
    return OVX_OK;
}

ovx_status ovx_fdg_set_graph(ovx_fdg_ctx *c, int enable)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null context");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->pending)
        return fail(OVX_ERR_STATE, "graph mode changed between begin and finish");
    c->graph_mode = enable != 0;
    if (!c->graph_mode)
    {
        if (c->done_recorded)
            HIP_TRY(hipEventSynchronize(c->ev_done));
        drop_graphs(c);
    }
    return OVX_OK;
}

ovx_status ovx_fdg_graph_stats(ovx_fdg_ctx *c, int64_t *replays, int64_t *captures)
{
    if (!c || !replays || !captures)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    *replays = c->graph_replays;
    *captures = c->graph_captures;
    return OVX_OK;
}

int ovx_fdg_kernel_count(void) { return KS_N; }

const char *ovx_fdg_kernel_name(int slot) { return (slot >= 0 && slot < KS_N) ? kslot_names[slot] : ""; }

ovx_status ovx_fdg_kernel_times(ovx_fdg_ctx *c, float *ms, int n)
{
    if (!c || !ms)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->timing)
        return fail(OVX_ERR_STATE, "timing is not enabled");
    for (int k = 0; k < n && k < KS_N; ++k)
    {
        ms[k] = -1.0f;
        if (!c->kran[k])
            continue;
        HIP_TRY(hipEventSynchronize(c->kev[2 * k + 1]));
        HIP_TRY(hipEventElapsedTime(&ms[k], c->kev[2 * k], c->kev[2 * k + 1]));
    }
    return OVX_OK;
}

ovx_status ovx_fdg_stage_times(ovx_fdg_ctx *c, float *ms4)
{
    if (!ms4)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null argument");
    float k[KS_N];
    ovx_status st = ovx_fdg_kernel_times(c, k, KS_N);
    if (st)
        return st;
    auto v = [&](int i) { return k[i] > 0.0f ? k[i] : 0.0f; };
    ms4[0] = v(KS_PREP) + v(KS_MARK) + v(KS_WORDSCAN) + v(KS_COORDS) + v(KS_GEN_EVENTS);
    ms4[1] = v(KS_FACE) + v(KS_GEN_FHITS);
    ms4[2] = v(KS_EDGES) + v(KS_BDRY) + v(KS_GEN_BHITS);
    ms4[3] = v(KS_ACTIVE) + v(KS_FOLD) + v(KS_FOLD_LARGE) + v(KS_SOLVE) + v(KS_SOLVE_BOX);
    return OVX_OK;
}

// ---------------------------------------------------------------------------
// flexible_dual_grid_to_mesh
// ---------------------------------------------------------------------------
ovx_status ovx_dual_mesh_begin(ovx_fdg_ctx *c, const int32_t *d_coords, const float *d_dual, const uint8_t *d_flags,
                               int64_t n_voxels, const int32_t *grid_size, const float *voxel_size,
                               const float *aabb_min, float *d_vertices, void *stream, int64_t *n_quads)
{
    if (!c || !grid_size || !voxel_size || !aabb_min || !n_quads)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null argument");
    if (n_voxels < 0 || n_voxels > (1ll << 30))
        return fail(OVX_ERR_CAPACITY, "n_voxels = %lld outside [0, 2^30]", (long long)n_voxels);
    if (n_voxels > 0 && (!d_coords || !d_dual || !d_flags || !d_vertices))
        return fail(OVX_ERR_INVALID_ARGUMENT, "null device pointer");
    for (int a = 0; a < 3; ++a)
        if (grid_size[a] < 1)
            return fail(OVX_ERR_INVALID_ARGUMENT, "grid_size[%d] = %d must be positive", a, grid_size[a]);
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    hipStream_t s = (hipStream_t)stream;
    c->pending = 0;
    StreamScope ss(c, s);
    HIP_TRY(ss.enter());
    DualMeshArgs &a = c->dm;
    a = DualMeshArgs{};
    a.coords = d_coords;
    a.dual = d_dual;
    a.flags = d_flags;
    a.N = n_voxels;
    a.W = grid_size[0];
    a.H = grid_size[1];
    a.D = grid_size[2];
    for (int k = 0; k < 3; ++k)
    {
        a.vs[k] = voxel_size[k];
        a.a0[k] = aabb_min[k];
    }
    a.verts = d_vertices;
    const uint64_t cells = (uint64_t)a.W * (uint64_t)a.H * (uint64_t)a.D;
    a.mode = cells <= DM_DIRECT_MAX_CELLS ? DM_DIRECT : (cells >= (1ull << 32) ? DM_WIDE : DM_NARROW);
    uint64_t cap = 64;
    while (cap < 2 * (uint64_t)n_voxels)
        cap <<= 1;
    a.cap = a.mode == DM_DIRECT ? cells : cap;
    *n_quads = 0;
    if (n_voxels == 0)
    {
        c->pending = 3;
        return OVX_OK;
    }
    ovx_status st;
    if (c->dm_dirty) // a begin whose finish never ran left cells set
    {
        HIP_TRY(launch_dm_restore((uint32_t *)c->dm_direct.p, (const uint32_t *)c->dm_touched.p, c->dm_dirty, s));
        c->dm_dirty = 0;
    }
    if (a.mode == DM_DIRECT)
    {
        void *old = c->dm_direct.p;
        const size_t old_bytes = c->dm_direct.bytes;
        if ((st = ensure(c->dm_direct, cells * 4)) || (st = ensure(c->dm_touched, (size_t)n_voxels * 4)))
            return st;
        if (c->dm_direct.p != old || c->dm_direct.bytes != old_bytes)
            HIP_TRY(hipMemsetAsync(c->dm_direct.p, 0, c->dm_direct.bytes, s));
    }
    else if ((st = ensure(c->dm_table, cap * (a.mode == DM_WIDE ? 16 : 8))))
        return st;
    if ((st = ensure(c->dm_cnt, (size_t)n_voxels * 4)) ||
        (st = ensure(c->dm_off, (size_t)n_voxels * 4)) || (st = ensure(c->dm_mask, (size_t)n_voxels)) ||
        (st = ensure(c->dm_tiles, scan_tile_count((uint64_t)n_voxels) * 8)) ||
        (st = ensure(c->counters, CTR_COUNT * sizeof(unsigned long long))))
        return st;
    a.table = a.mode == DM_DIRECT ? c->dm_direct.p : c->dm_table.p;
    a.touched = a.mode == DM_DIRECT ? (uint32_t *)c->dm_touched.p : nullptr;
    a.cnt = (uint32_t *)c->dm_cnt.p;
    a.off = (const uint32_t *)c->dm_off.p;
    a.mask = (uint8_t *)c->dm_mask.p;
    for (int k = KS_DM_INSERT; k <= KS_DM_RESTORE; ++k)
        c->kran[k] = false;
    unsigned long long *ctr = (unsigned long long *)c->counters.p;
    tic(c, KS_DM_INSERT, s);
    if (a.mode != DM_DIRECT)
        HIP_TRY(launch_dm_clear(a, s));
    HIP_TRY(launch_dm_insert(a, s));
    toc(c, KS_DM_INSERT, s);
    if (a.mode == DM_DIRECT)
        c->dm_dirty = n_voxels;
    tic(c, KS_DM_COUNT, s);
    HIP_TRY(launch_dm_count(a, s));
    toc(c, KS_DM_COUNT, s);
    tic(c, KS_DM_SCAN, s);
    HIP_TRY(launch_scan_u32(a.cnt, (uint64_t)n_voxels, (uint32_t *)c->dm_off.p, (uint64_t *)c->dm_tiles.p,
                            ctr + CTR_DM_QUADS, s));
    toc(c, KS_DM_SCAN, s);
    HIP_TRY(hipMemcpyAsync(c->h_ctr + CTR_DM_QUADS, ctr + CTR_DM_QUADS, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->last_stream = s;
    c->ran = true;
    *n_quads = (int64_t)c->h_ctr[CTR_DM_QUADS];
    c->pending = 3;
    return OVX_OK;
}

ovx_status ovx_dual_mesh_finish(ovx_fdg_ctx *c, const float *d_split_weight, int split_weight_type, int32_t *d_faces,
                                int32_t *d_quads, void *stream)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx is null");
    if (split_weight_type < OVX_SPLIT_F32 || split_weight_type > OVX_SPLIT_BF16)
        return fail(OVX_ERR_INVALID_ARGUMENT, "split_weight_type %d is not one of OVX_SPLIT_*", split_weight_type);
    if (((uintptr_t)d_quads & 15) != 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "d_quads must be 16-byte aligned");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->pending != 3)
        return fail(OVX_ERR_STATE, "ovx_dual_mesh_finish without a successful ovx_dual_mesh_begin");
    c->pending = 0;
    hipStream_t s = (hipStream_t)stream;
    StreamScope ss(c, s);
    HIP_TRY(ss.enter());
    DualMeshArgs a = c->dm;
    if (a.N > 0 && c->h_ctr[CTR_DM_QUADS] > 0 && (d_faces || d_quads))
    {
        a.split_w = d_split_weight;
        a.faces = d_faces;
        a.quads = d_quads;
        tic(c, KS_DM_EMIT, s);
        HIP_TRY(launch_dm_emit(a, split_weight_type, s));
        toc(c, KS_DM_EMIT, s);
    }
    if (c->dm_dirty)
    {
        tic(c, KS_DM_RESTORE, s);
        HIP_TRY(launch_dm_restore((uint32_t *)c->dm_direct.p, (const uint32_t *)c->dm_touched.p, c->dm_dirty, s));
        toc(c, KS_DM_RESTORE, s);
        c->dm_dirty = 0;
    }
    c->last_stream = s;
    return OVX_OK;
}

// ---------------------------------------------------------------------------
// generic hashmap (reference table format)
// ---------------------------------------------------------------------------
static ovx_status hash_args(const void *tk, int kb, const void *tv, int vb, int64_t cap, const void *in, int64_t n,
                            bool coords)
{
    if ((kb != 4 && kb != 8) || (vb != 4 && vb != 8))
        return fail(OVX_ERR_INVALID_ARGUMENT, "key_bytes / value_bytes must be 4 or 8 (got %d / %d)", kb, vb);
    if (cap <= 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "capacity must be positive");
    if (n < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "n must be non-negative");
    if (n > 0 && (!tk || !tv || !in))
        return fail(OVX_ERR_INVALID_ARGUMENT, "null device pointer");
    if (coords && ((uintptr_t)in & 15) != 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "coords must be 16-byte aligned int32 [n, 4]");
    return OVX_OK;
}

ovx_status ovx_hashmap_insert(void *tk, int kb, void *tv, int vb, int64_t cap, const void *d_keys, const void *d_values,
                              int64_t n, void *stream)
{
    if (ovx_status st = hash_args(tk, kb, tv, vb, cap, d_keys, n, false))
        return st;
    if (n > 0 && !d_values)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null values");
    HIP_TRY(launch_hash_insert(tk, kb, tv, vb, (uint64_t)cap, HASH_KEYS, d_keys, d_values, (uint64_t)n, 0, 0, 0,
                               (hipStream_t)stream));
    return OVX_OK;
}

ovx_status ovx_hashmap_lookup(const void *tk, int kb, const void *tv, int vb, int64_t cap, const void *d_keys,
                              void *d_out, int64_t n, void *stream)
{
    if (ovx_status st = hash_args(tk, kb, tv, vb, cap, d_keys, n, false))
        return st;
    if (n > 0 && !d_out)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null output");
    HIP_TRY(launch_hash_lookup(tk, kb, tv, vb, (uint64_t)cap, HASH_KEYS, d_keys, d_out, (uint64_t)n, 0, 0, 0,
                               (hipStream_t)stream));
    return OVX_OK;
}

ovx_status ovx_hashmap_insert_3d(void *tk, int kb, void *tv, int vb, int64_t cap, const int32_t *d_coords,
                                 const void *d_values, int64_t n, int W, int H, int D, void *stream)
{
    if (ovx_status st = hash_args(tk, kb, tv, vb, cap, d_coords, n, true))
        return st;
    if (n > 0 && !d_values)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null values");
    HIP_TRY(launch_hash_insert(tk, kb, tv, vb, (uint64_t)cap, HASH_COORDS, d_coords, d_values, (uint64_t)n, W, H, D,
                               (hipStream_t)stream));
    return OVX_OK;
}

ovx_status ovx_hashmap_lookup_3d(const void *tk, int kb, const void *tv, int vb, int64_t cap, const int32_t *d_coords,
                                 void *d_out, int64_t n, int W, int H, int D, void *stream)
{
    if (ovx_status st = hash_args(tk, kb, tv, vb, cap, d_coords, n, true))
        return st;
    if (n > 0 && !d_out)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null output");
    HIP_TRY(launch_hash_lookup(tk, kb, tv, vb, (uint64_t)cap, HASH_COORDS, d_coords, d_out, (uint64_t)n, W, H, D,
                               (hipStream_t)stream));
    return OVX_OK;
}

ovx_status ovx_hashmap_insert_3d_idx_as_val(void *tk, int kb, void *tv, int vb, int64_t cap, const int32_t *d_coords,
                                            int64_t n, int W, int H, int D, void *stream)
{
    if (ovx_status st = hash_args(tk, kb, tv, vb, cap, d_coords, n, true))
        return st;
    HIP_TRY(launch_hash_insert(tk, kb, tv, vb, (uint64_t)cap, HASH_COORDS, d_coords, nullptr, (uint64_t)n, W, H, D,
                               (hipStream_t)stream));
    return OVX_OK;
}

// ---------------------------------------------------------------------------
// Z-order / Hilbert codes
// ---------------------------------------------------------------------------
ovx_status ovx_seq_encode(const void *const *d_cols, const int64_t *strides, int elem_bytes, int64_t n, int mode,
                          int32_t *d_codes, void *stream)
{
    if (!d_cols || !strides || n < 0 || (elem_bytes != 4 && elem_bytes != 8) ||
        (mode != OVX_SEQ_Z_ORDER && mode != OVX_SEQ_HILBERT))
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid ovx_seq_encode arguments");
    if (n == 0)
        return OVX_OK;
    SeqCols in{};
    for (int k = 0; k < 3; ++k)
    {
        if (!d_cols[k])
            return fail(OVX_ERR_INVALID_ARGUMENT, "null coordinate column %d", k);
        in.p[k] = d_cols[k];
        in.stride[k] = strides[k];
    }
    if (!d_codes)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null codes");
    HIP_TRY(launch_seq_encode(in, elem_bytes, n, mode, d_codes, (hipStream_t)stream));
    return OVX_OK;
}

ovx_status ovx_seq_decode(const int32_t *d_codes, int64_t n, int mode, int32_t *const *d_cols, const int64_t *strides,
                          void *stream)
{
    if (!d_cols || !strides || n < 0 || (mode != OVX_SEQ_Z_ORDER && mode != OVX_SEQ_HILBERT))
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid ovx_seq_decode arguments");
    if (n == 0)
        return OVX_OK;
    if (!d_codes)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null codes");
    SeqCols out{};
    for (int k = 0; k < 3; ++k)
    {
        if (!d_cols[k])
            return fail(OVX_ERR_INVALID_ARGUMENT, "null coordinate column %d", k);
        out.p[k] = d_cols[k];
        out.stride[k] = strides[k];
    }
    HIP_TRY(launch_seq_decode(d_codes, n, mode, out, (hipStream_t)stream));
    return OVX_OK;
}

// ---------------------------------------------------------------------------
// sparse voxel octrees and attribute filters of the .vxz format
// ---------------------------------------------------------------------------
ovx_status ovx_svo_encode_begin(ovx_fdg_ctx *c, const int32_t *d_codes, const int32_t *d_chunk,
                                const int64_t *d_chunk_off, int64_t n, int n_chunks, int depth, void *stream,
                                int64_t *n_bytes)
{
    if (!c || !n_bytes || n < 0 || depth < 1 || depth > 10 || n_chunks < 0 || n > (1ll << 31))
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid ovx_svo_encode_begin arguments");
    if (n > 0 && (!d_codes || !d_chunk || !d_chunk_off || n_chunks < 1))
        return fail(OVX_ERR_INVALID_ARGUMENT, "null device pointer or no chunk");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    hipStream_t s = (hipStream_t)stream;
    c->svo_state = 0;
    StreamScope ss(c, s);
    HIP_TRY(ss.enter());
    c->svo = SvoCodes{d_codes, d_chunk, d_chunk_off, n, n_chunks, depth};
    *n_bytes = 0;
    c->svo_total = 0;
    if (n > 0)
    {
        ovx_status st;
        if ((st = ensure(c->svo_insf, (size_t)n)) || (st = ensure(c->svo_k, (size_t)n * 4)) ||
            (st = ensure(c->svo_pos, (size_t)n * 4)) || (st = ensure(c->svo_tiles, scan_tile_count((uint64_t)n) * 8)) ||
            (st = ensure(c->counters, CTR_COUNT * sizeof(unsigned long long))))
            return st;
        unsigned long long *ctr = (unsigned long long *)c->counters.p;
        HIP_TRY(launch_svo_prep(c->svo, (uint8_t *)c->svo_insf.p, (uint32_t *)c->svo_k.p, s));
        HIP_TRY(launch_scan_u32((const uint32_t *)c->svo_k.p, (uint64_t)n, (uint32_t *)c->svo_pos.p,
                                (uint64_t *)c->svo_tiles.p, ctr + CTR_SVO_BYTES, s));
        HIP_TRY(hipMemcpyAsync(c->h_ctr + CTR_SVO_BYTES, ctr + CTR_SVO_BYTES, sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        c->svo_total = c->h_ctr[CTR_SVO_BYTES];
        if (c->svo_total >= (1ull << 32) - 4)
            return fail(OVX_ERR_CAPACITY, "octree of %llu bytes exceeds 2^32 - 4", (unsigned long long)c->svo_total);
        c->last_stream = s;
        c->ran = true;
    }
    *n_bytes = (int64_t)c->svo_total;
    c->svo_state = 1;
    return OVX_OK;
}

ovx_status ovx_svo_encode_finish(ovx_fdg_ctx *c, uint8_t *d_svo, int64_t *d_chunk_svo_off, void *stream)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx is null");
    if (((uintptr_t)d_svo & 3) != 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "d_svo must be 4-byte aligned");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->svo_state != 1)
        return fail(OVX_ERR_STATE, "ovx_svo_encode_finish without a successful ovx_svo_encode_begin");
    hipStream_t s = (hipStream_t)stream;
    StreamScope ss(c, s);
    HIP_TRY(ss.enter());
    if (c->svo_total > 0)
    {
        if (!d_svo)
            return fail(OVX_ERR_INVALID_ARGUMENT, "null octree buffer");
        HIP_TRY(hipMemsetAsync(d_svo, 0, (c->svo_total + 3) & ~3ull, s));
        HIP_TRY(launch_svo_bits(c->svo, (const uint8_t *)c->svo_insf.p, (const uint32_t *)c->svo_pos.p, d_svo, s));
    }
    if (d_chunk_svo_off && c->svo.n_chunks > 0)
        HIP_TRY(launch_svo_chunk_offsets(c->svo, (const uint32_t *)c->svo_pos.p, c->svo_total, d_chunk_svo_off, s));
    c->svo_bytes = d_svo;
    c->svo_state = 2;
    c->last_stream = s;
    return OVX_OK;
}

ovx_status ovx_svo_parent_encode(ovx_fdg_ctx *c, const uint8_t *d_attr, int channels, uint8_t *d_delta, void *stream)
{
    if (!c || channels < 1)
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid ovx_svo_parent_encode arguments");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->svo_state != 2)
        return fail(OVX_ERR_STATE, "ovx_svo_parent_encode needs the octree of a finished encode on this context");
    const int64_t n = c->svo.n;
    if (n == 0)
        return OVX_OK;
    if (!d_attr || !d_delta)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null attribute buffer");
    hipStream_t s = (hipStream_t)stream;
    StreamScope ss(c, s);
    HIP_TRY(ss.enter());
    ovx_status st;
    if ((st = ensure(c->svo_rows, c->svo_total * 4)) || (st = ensure(c->svo_erow, c->svo_total * 4)) ||
        (st = ensure(c->svo_tiles, scan_tile_count(std::max<uint64_t>(c->svo_total, (uint64_t)n)) * 8)))
        return st;
    unsigned long long *ctr = (unsigned long long *)c->counters.p;
    HIP_TRY(launch_svo_rows(c->svo_bytes, c->svo_total, (uint32_t *)c->svo_rows.p, s));
    HIP_TRY(launch_scan_u32((const uint32_t *)c->svo_rows.p, c->svo_total, (uint32_t *)c->svo_erow.p,
                            (uint64_t *)c->svo_tiles.p, ctr + CTR_SVO_BYTES + 1, s));
    HIP_TRY(hipMemsetAsync(d_delta, 0, (size_t)n * channels, s));
    HIP_TRY(launch_svo_parent_encode(c->svo, (const uint8_t *)c->svo_insf.p, (const uint32_t *)c->svo_pos.p,
                                     c->svo_bytes, (const uint32_t *)c->svo_erow.p, d_attr, channels, d_delta, s));
    c->last_stream = s;
    return OVX_OK;
}

ovx_status ovx_svo_neighbor_filter(ovx_fdg_ctx *c, const int32_t *d_coords, const int32_t *d_chunk, int64_t n,
                                   int n_chunks, int res, const uint8_t *d_in, int channels, int decode,
                                   uint8_t *d_out, void *stream)
{
    if (!c || n < 0 || n_chunks < 0 || res < 1 || channels < 1 || n > (1ll << 31) - 2)
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid ovx_svo_neighbor_filter arguments");
    const uint64_t cells = (uint64_t)n_chunks * (uint64_t)res * (uint64_t)res * (uint64_t)res;
    if (cells > (1ull << 31))
        return fail(OVX_ERR_CAPACITY, "neighbour filter over %llu cells exceeds 2^31", (unsigned long long)cells);
    if (n == 0)
        return OVX_OK;
    if (!d_coords || !d_chunk || !d_in || !d_out)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null device pointer");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    hipStream_t s = (hipStream_t)stream;
    StreamScope ss(c, s);
    HIP_TRY(ss.enter());
    ovx_status st;
    if ((st = ensure(c->nb_table, cells * 4)) || (st = ensure(c->nb_own, (size_t)n * 4)) ||
        (st = ensure(c->nb_nb, (size_t)n * 4)) ||
        (decode && ((st = ensure(c->nb_acc, (size_t)n * channels * 2)) || (st = ensure(c->nb_jmp, (size_t)n * 8)))))
        return st;
    HIP_TRY(hipMemsetAsync(c->nb_table.p, 0, cells * 4, s));
    NbArgs a{d_coords, d_chunk, n, n_chunks, res, (uint32_t *)c->nb_table.p, (int32_t *)c->nb_own.p,
             (int32_t *)c->nb_nb.p};
    HIP_TRY(launch_nb_insert(a, s));
    HIP_TRY(launch_nb_link(a, s));
    if (decode)
    {
        uint8_t *acc = (uint8_t *)c->nb_acc.p;
        int32_t *jmp = (int32_t *)c->nb_jmp.p;
        HIP_TRY(launch_nb_decode(a, d_in, channels, acc, acc + (size_t)n * channels, jmp, jmp + n, d_out, s));
    }
    else
        HIP_TRY(launch_nb_encode(a, d_in, channels, d_out, s));
    c->last_stream = s;
    c->ran = true;
    return OVX_OK;
}

ovx_status ovx_svo_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, int32_t *codes, int64_t capacity,
                               int64_t *n_codes)
{
    if (!svo || !codes || !n_codes || n_bytes < 1 || depth < 1 || depth > 10 || capacity < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid ovx_svo_decode_host arguments");
    const int64_t r = svo_decode_host(svo, n_bytes, depth, codes, capacity);
    if (r == -2)
        return fail(OVX_ERR_CAPACITY, "octree holds more than %lld leaves", (long long)capacity);
    if (r < 0)
        return fail(OVX_ERR_INPUT, "truncated or malformed octree stream");
    *n_codes = r;
    return OVX_OK;
}

ovx_status ovx_svo_parent_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, const uint8_t *delta, int64_t n,
                                      int channels, uint8_t *attr)
{
    if (!svo || !delta || !attr || n_bytes < 1 || depth < 1 || depth > 10 || n < 1 || channels < 1)
        return fail(OVX_ERR_INVALID_ARGUMENT, "invalid ovx_svo_parent_decode_host arguments");
    const int64_t r = svo_parent_decode_host(svo, n_bytes, depth, delta, n, channels, attr);
    if (r != n)
        return fail(OVX_ERR_INPUT, "octree and delta rows do not match (%lld leaves for %lld rows)", (long long)r,
                    (long long)n);
    return OVX_OK;
}

} // extern "C"
