// C ABI of libovx_fdg.so: host orchestration of the flexible-dual-grid
// pipeline on one HIP device.  See include/ovx_fdg.h for the contract.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "ovx_fdg.h"
#include "fdg_kernels.h"

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

struct ovx_fdg_ctx
{
    int device = 0;
    std::mutex mu;
    // workspace
    DevBuf planes, bits, woff, tiles, counters, cnt, rec_off, keys, zs, gperm;
    DevBuf vcnt, voff, nbr, bcnt, boff, bedge_v, bedges;
    unsigned long long *h_ctr = nullptr; // pinned mirror of the counters
    // state between begin and finish
    int pending = 0; // 0 none, 1 fdg, 2 occ
    MeshDev mesh{};
    int64_t V = 0;
    GridDev grid{};
    uint64_t n_words = 0;
    float fw = 0, bw = 0, rw = 0;
    int64_t N = 0;
    // statistics of the last conversion
    int64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    hipStream_t last_stream = nullptr;
    // per-kernel timing (events around each launch on the caller's stream)
    bool timing = false;
    hipEvent_t kev[2 * 16] = {};
    bool kran[16] = {};
};

namespace
{

ovx_status ensure(DevBuf &b, size_t bytes)
{
    if (bytes == 0)
        bytes = 256;
    if (b.bytes >= bytes)
        return OVX_OK;
    if (b.p)
    {
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    const size_t want = bytes + bytes / 8 + 256;
    HIP_TRY(hipMalloc(&b.p, want));
    b.bytes = want;
    return OVX_OK;
}

void release(DevBuf &b)
{
    if (b.p)
        (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

int64_t held_bytes(const ovx_fdg_ctx *c)
{
    const DevBuf *all[] = {&c->planes, &c->bits, &c->woff, &c->tiles, &c->counters, &c->cnt, &c->rec_off, &c->keys,
                           &c->zs, &c->gperm, &c->vcnt, &c->voff, &c->nbr, &c->bcnt, &c->boff, &c->bedge_v,
                           &c->bedges};
    int64_t s = 0;
    for (const DevBuf *b : all)
        s += (int64_t)b->bytes;
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
        g.gmin[a] = (int)lo;
        g.gmax[a] = (int)hi;
        g.G[a] = (int)(hi - lo);
    }
    g.wx = (g.G[0] + 63) / 64;
    n_words = (uint64_t)g.G[2] * (uint64_t)g.G[1] * (uint64_t)g.wx;
    if (n_words > (1ull << 32))
        return fail(OVX_ERR_CAPACITY, "occupancy bitset of %llu words exceeds 2^32", (unsigned long long)n_words);
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
    if (flags & ERR_RECORDS)
        return fail(OVX_ERR_STATE, "record count and emit passes disagree (internal error)");
    return fail(OVX_ERR_INPUT, "invalid input (flags %llx)", flags);
}

// timing slots, one per kernel (or kernel group) of the pipeline
enum KSlot
{
    KS_PREP,
    KS_MARK,
    KS_WORDSCAN,
    KS_COORDS,
    KS_COUNT_SCAN,
    KS_COUNT_FACE,
    KS_EDGES,
    KS_COUNT_BDRY,
    KS_RECSCAN,
    KS_EMIT_SCAN,
    KS_EMIT_FACE,
    KS_EMIT_BDRY,
    KS_FOLD,
    KS_N
};
const char *const kslot_names[KS_N] = {"prep",       "mark",       "word_scan", "coords",    "count_events",
                                       "count_face", "edges",      "count_bdry", "rec_scan", "emit_events",
                                       "emit_face",  "emit_bdry",  "fold_solve"};

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

// common begin: validation, (planes), mark, compaction count
ovx_status begin_common(ovx_fdg_ctx *c, const MeshDev &m, int64_t V, const ovx_grid *grid, bool planes,
                        hipStream_t s, int64_t *n_voxels)
{
    c->pending = 0;
    c->N = 0;
    for (int i = 0; i < 8; ++i)
        c->stats[i] = 0;
    GridDev g{};
    uint64_t n_words = 0;
    ovx_status st = check_grid(grid, g, n_words);
    if (st)
        return st;
    if (m.F < 0 || V < 0)
        return fail(OVX_ERR_INVALID_ARGUMENT, "negative size");
    if (m.F > 0x7fffffffll)
        return fail(OVX_ERR_CAPACITY, "more than 2^31-1 faces");
    if (m.F > 0 && !m.v)
        return fail(OVX_ERR_INVALID_ARGUMENT, "vertices pointer is null");
    if ((st = ensure(c->counters, CTR_COUNT * sizeof(unsigned long long))))
        return st;
    HIP_TRY(hipMemsetAsync(c->counters.p, 0, CTR_COUNT * sizeof(unsigned long long), s));
    c->mesh = m;
    c->V = V;
    c->last_stream = s;
    if (m.F == 0)
    {
        c->grid = g;
        c->n_words = n_words;
        c->pending = planes ? 1 : 2;
        *n_voxels = 0;
        return OVX_OK;
    }
    if ((st = ensure(c->bits, n_words * sizeof(uint64_t))) || (st = ensure(c->woff, n_words * sizeof(uint32_t))) ||
        (st = ensure(c->tiles, (scan_tile_count(n_words) + scan_tile_count((uint64_t)m.F * 3 + 1) + 64) * 8)))
        return st;
    if (planes && (st = ensure(c->planes, (size_t)m.F * sizeof(float4))))
        return st;
    g.bits = (uint64_t *)c->bits.p;
    g.woff = (const uint32_t *)c->woff.p;
    auto *ctr = (unsigned long long *)c->counters.p;

    for (bool &r : c->kran)
        r = false;
    tic(c, KS_PREP, s);
    HIP_TRY(hipMemsetAsync(c->bits.p, 0, n_words * sizeof(uint64_t), s));
    HIP_TRY(launch_prep(m, V, g, planes ? (float4 *)c->planes.p : nullptr, ctr, s));
    toc(c, KS_PREP, s);
    RecArgs none{};
    tic(c, KS_MARK, s);
    HIP_TRY(launch_scan_events(MODE_MARK, m, g, none, ctr, s));
    toc(c, KS_MARK, s);
    tic(c, KS_WORDSCAN, s);
    HIP_TRY(launch_scan_popc(g.bits, n_words, (uint32_t *)c->woff.p, (uint64_t *)c->tiles.p, ctr + CTR_WORDS_TOTAL,
                             s));
    toc(c, KS_WORDSCAN, s);
    if ((st = read_counters(c, s)))
        return st;
    if (c->h_ctr[CTR_ERRORS])
        return input_error(c->h_ctr[CTR_ERRORS]);
    const unsigned long long N = c->h_ctr[CTR_WORDS_TOTAL];
    if (N >= (1ull << 32) - 1)
        return fail(OVX_ERR_CAPACITY, "%llu occupied voxels exceed the 32-bit voxel index", N);
    c->grid = g;
    c->n_words = n_words;
    c->N = (int64_t)N;
    c->stats[0] = (int64_t)N;
    c->stats[1] = (int64_t)c->h_ctr[CTR_EVENTS];
    c->pending = planes ? 1 : 2;
    *n_voxels = (int64_t)N;
    return OVX_OK;
}

ovx_status finish_fdg(ovx_fdg_ctx *c, int32_t *coords, float *dual, uint8_t *flags, hipStream_t s)
{
    ovx_status st;
    const int64_t N = c->N;
    const MeshDev &m = c->mesh;
    GridDev g = c->grid;
    auto *ctr = (unsigned long long *)c->counters.p;
    tic(c, KS_COORDS, s);
    HIP_TRY(launch_emit_coords(g, c->n_words, coords, s));
    toc(c, KS_COORDS, s);

    // ---- per-voxel record counts
    if ((st = ensure(c->cnt, (size_t)(N + 1) * sizeof(uint32_t))) ||
        (st = ensure(c->rec_off, (size_t)(N + 1) * sizeof(uint32_t))) ||
        (st = ensure(c->tiles, (scan_tile_count((uint64_t)N + 1) + scan_tile_count(c->n_words) +
                                scan_tile_count((uint64_t)c->V + 1) + 64) *
                                   8)))
        return st;
    HIP_TRY(hipMemsetAsync(c->cnt.p, 0, (size_t)(N + 1) * sizeof(uint32_t), s));
    RecArgs r{};
    r.cnt = (uint32_t *)c->cnt.p;
    r.rec_off = (const uint32_t *)c->rec_off.p;
    r.err = ctr + CTR_ERRORS;
    tic(c, KS_COUNT_SCAN, s);
    HIP_TRY(launch_scan_events(MODE_COUNT, m, g, r, ctr, s));
    toc(c, KS_COUNT_SCAN, s);
    if (c->fw > 0.0f)
    {
        tic(c, KS_COUNT_FACE, s);
        HIP_TRY(launch_face(MODE_COUNT, m, g, r, ctr, s));
        toc(c, KS_COUNT_FACE, s);
    }

    // ---- boundary edges (edge multiset of FDG.cpp:728-739, kept in std::map order)
    int64_t E = 0;
    if (c->bw > 0.0f)
    {
        const int64_t V = c->V;
        if ((uint64_t)m.F * 3ull >= (1ull << 32))
            return fail(OVX_ERR_CAPACITY, "more than 2^32 half-edges");
        if ((st = ensure(c->vcnt, (size_t)(V + 1) * 4)) || (st = ensure(c->voff, (size_t)(V + 1) * 4)) ||
            (st = ensure(c->nbr, (size_t)m.F * 3 * 4)) || (st = ensure(c->bcnt, (size_t)(V + 1) * 4)) ||
            (st = ensure(c->boff, (size_t)(V + 1) * 4)))
            return st;
        auto *vcnt = (uint32_t *)c->vcnt.p;
        auto *voff = (uint32_t *)c->voff.p;
        auto *nbr = (uint32_t *)c->nbr.p;
        auto *bcnt = (uint32_t *)c->bcnt.p;
        auto *boff = (uint32_t *)c->boff.p;
        tic(c, KS_EDGES, s);
        HIP_TRY(hipMemsetAsync(vcnt, 0, (size_t)(V + 1) * 4, s));
        HIP_TRY(launch_edge_count(m, vcnt, voff, nbr, 0, s));
        HIP_TRY(launch_scan_u32(vcnt, (uint64_t)V + 1, voff, (uint64_t *)c->tiles.p, ctr + CTR_EDGE_TOTAL, s));
        HIP_TRY(launch_edge_count(m, vcnt, voff, nbr, 1, s));
        HIP_TRY(hipMemsetAsync(bcnt, 0, (size_t)(V + 1) * 4, s));
        HIP_TRY(launch_edge_singletons(V, voff, nbr, bcnt, boff, nullptr, 0, s));
        HIP_TRY(launch_scan_u32(bcnt, (uint64_t)V + 1, boff, (uint64_t *)c->tiles.p, ctr + CTR_BEDGE_TOTAL, s));
        toc(c, KS_EDGES, s);
        if ((st = read_counters(c, s)))
            return st;
        E = (int64_t)c->h_ctr[CTR_BEDGE_TOTAL];
        if (E > 0)
        {
            if ((st = ensure(c->bedge_v, (size_t)E * sizeof(int2))) ||
                (st = ensure(c->bedges, (size_t)E * sizeof(BEdge))))
                return st;
            tic(c, KS_COUNT_BDRY, s);
            HIP_TRY(launch_edge_singletons(V, voff, nbr, bcnt, boff, (int2 *)c->bedge_v.p, 1, s));
            HIP_TRY(launch_bedge_prep(m.v, E, (const int2 *)c->bedge_v.p, (BEdge *)c->bedges.p, s));
            HIP_TRY(launch_bdry(MODE_COUNT, E, (const BEdge *)c->bedges.p, g, r, ctr, s));
            toc(c, KS_COUNT_BDRY, s);
        }
    }
    c->stats[4] = E;

    // ---- record offsets
    tic(c, KS_RECSCAN, s);
    HIP_TRY(launch_scan_u32(r.cnt, (uint64_t)N + 1, (uint32_t *)c->rec_off.p, (uint64_t *)c->tiles.p,
                            ctr + CTR_REC_TOTAL, s));
    toc(c, KS_RECSCAN, s);
    if ((st = read_counters(c, s)))
        return st;
    if (c->h_ctr[CTR_ERRORS])
        return input_error(c->h_ctr[CTR_ERRORS]);
    const unsigned long long R = c->h_ctr[CTR_REC_TOTAL];
    if (R >= (1ull << 32))
        return fail(OVX_ERR_CAPACITY, "%llu QEF records exceed the 32-bit record index", R);
    c->stats[2] = (int64_t)R;
    c->stats[3] = (int64_t)c->h_ctr[CTR_REC_FACE];
    c->stats[5] = (int64_t)c->h_ctr[CTR_REC_BOUNDARY];
    if ((st = ensure(c->keys, (size_t)R * 8)) || (st = ensure(c->zs, (size_t)R * 4)) ||
        (st = ensure(c->gperm, (size_t)R * 4)))
        return st;
    r.keys = (uint64_t *)c->keys.p;
    r.zs = (float *)c->zs.p;

    // ---- emit records (the counters run back down to zero)
    tic(c, KS_EMIT_SCAN, s);
    HIP_TRY(launch_scan_events(MODE_EMIT, m, g, r, nullptr, s));
    toc(c, KS_EMIT_SCAN, s);
    if (c->fw > 0.0f)
    {
        tic(c, KS_EMIT_FACE, s);
        HIP_TRY(launch_face(MODE_EMIT, m, g, r, nullptr, s));
        toc(c, KS_EMIT_FACE, s);
    }
    if (E > 0)
    {
        tic(c, KS_EMIT_BDRY, s);
        HIP_TRY(launch_bdry(MODE_EMIT, E, (const BEdge *)c->bedges.p, g, r, nullptr, s));
        toc(c, KS_EMIT_BDRY, s);
    }

    // ---- ordered fold + solve
    FoldArgs a{};
    a.N = (uint64_t)N;
    a.rec_off = (const uint32_t *)c->rec_off.p;
    a.keys = (const uint64_t *)c->keys.p;
    a.zs = (const float *)c->zs.p;
    a.gperm = (uint32_t *)c->gperm.p;
    a.planes = (const float4 *)c->planes.p;
    a.bedges = (const BEdge *)c->bedges.p;
    a.coords = coords;
    a.dual = dual;
    a.flags = flags;
    for (int i = 0; i < 3; ++i)
    {
        a.vs[i] = g.vs[i];
        a.gmin[i] = g.gmin[i];
    }
    a.bw = c->bw;
    a.rw = c->rw;
    a.counters = ctr;
    tic(c, KS_FOLD, s);
    HIP_TRY(launch_fold_solve(a, s));
    toc(c, KS_FOLD, s);
    return OVX_OK;
}

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
    if (e != hipSuccess)
    {
        delete c;
        return fail(OVX_ERR_HIP, "hipHostMalloc failed: %s", hipGetErrorString(e));
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
        if (c->last_stream)
            (void)hipStreamSynchronize(c->last_stream);
        DevBuf *all[] = {&c->planes, &c->bits, &c->woff, &c->tiles, &c->counters, &c->cnt, &c->rec_off, &c->keys,
                         &c->zs, &c->gperm, &c->vcnt, &c->voff, &c->nbr, &c->bcnt, &c->boff, &c->bedge_v,
                         &c->bedges};
        for (DevBuf *b : all)
            release(*b);
        if (c->h_ctr)
            (void)hipHostFree(c->h_ctr);
        for (auto &e : c->kev)
            if (e)
                (void)hipEventDestroy(e);
    }
    delete c;
    return OVX_OK;
}

int64_t ovx_fdg_workspace_bytes(const ovx_fdg_ctx *c) { return c ? held_bytes(c) : 0; }

ovx_status ovx_fdg_trim(ovx_fdg_ctx *c)
{
    if (!c)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->last_stream)
        HIP_TRY(hipStreamSynchronize(c->last_stream));
    DevBuf *all[] = {&c->planes, &c->bits, &c->woff, &c->tiles, &c->cnt, &c->rec_off, &c->keys, &c->zs, &c->gperm,
                     &c->vcnt, &c->voff, &c->nbr, &c->bcnt, &c->boff, &c->bedge_v, &c->bedges};
    for (DevBuf *b : all)
        release(*b);
    c->pending = 0;
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
    MeshDev m{d_vertices, d_faces, n_faces};
    c->fw = face_weight;
    c->bw = boundary_weight;
    c->rw = regularization_weight;
    return begin_common(c, m, n_vertices, grid, true, (hipStream_t)stream, n_voxels);
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
    c->last_stream = (hipStream_t)stream;
    return finish_fdg(c, d_coords, d_dual, d_intersected, (hipStream_t)stream);
}

ovx_status ovx_occ_begin(ovx_fdg_ctx *c, const float *d_triangles, int64_t n_triangles, const ovx_grid *grid,
                         void *stream, int64_t *n_voxels)
{
    if (!c || !n_voxels)
        return fail(OVX_ERR_INVALID_ARGUMENT, "ctx / n_voxels is null");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    MeshDev m{d_triangles, nullptr, n_triangles};
    return begin_common(c, m, 3 * n_triangles, grid, false, (hipStream_t)stream, n_voxels);
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
    HIP_TRY(launch_emit_coords(c->grid, c->n_words, d_coords, (hipStream_t)stream));
    return OVX_OK;
}

ovx_status ovx_fdg_stats(ovx_fdg_ctx *c, int64_t *out8)
{
    if (!c || !out8)
        return fail(OVX_ERR_INVALID_ARGUMENT, "null argument");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->last_stream && c->counters.p)
    {
        ovx_status st = read_counters(c, c->last_stream);
        if (st)
            return st;
        c->stats[7] = (int64_t)c->h_ctr[CTR_GIANT];
    }
    c->stats[6] = held_bytes(c);
    for (int i = 0; i < 8; ++i)
        out8[i] = c->stats[i];
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

ovx_status ovx_fdg_stage_times(ovx_fdg_ctx *c, float *ms5)
{
    float k[KS_N];
    ovx_status st = ovx_fdg_kernel_times(c, k, KS_N);
    if (st)
        return st;
    auto v = [&](int i) { return k[i] > 0.0f ? k[i] : 0.0f; };
    ms5[0] = v(KS_PREP) + v(KS_MARK);
    ms5[1] = v(KS_WORDSCAN) + v(KS_COORDS);
    ms5[2] = v(KS_COUNT_SCAN) + v(KS_COUNT_FACE) + v(KS_EDGES) + v(KS_COUNT_BDRY) + v(KS_RECSCAN);
    ms5[3] = v(KS_EMIT_SCAN) + v(KS_EMIT_FACE) + v(KS_EMIT_BDRY);
    ms5[4] = v(KS_FOLD);
    return OVX_OK;
}

} // extern "C"
