// Flexible dual grid on MI355X (gfx950): kernels.
//
// Pipeline for one mesh (all on the caller's stream):
//   prep      validate faces / vertices, per-face plane  (FDG.cpp:180-185)
//   mark      scanline events -> atomicOr into a dense occupancy bitset
//   compact   popcount prefix scan over bitset words -> voxel ids, coords
//   count     per-voxel record counts (scan events, face test, boundary DDA)
//   offsets   prefix scan of the counts -> record segments
//   emit      the same traversals write one 64-bit order key per record
//   fold      per voxel: sort its keys in LDS, replay the contributions in the
//             reference's serial order, regularize and solve (CPQR)
// The scan, face test and DDA traversals are expanded block-cooperatively
// (256 lanes share the rows / columns / voxels of 256 work items) so that
// large and small triangles balance across the wave64 lanes.
#include "fdg_device.h"
#include "fdg_kernels.h"

#include <type_traits>
#include "fdg_solve.h"

namespace ovx
{

// ---------------------------------------------------------------------------
// block primitives (256 threads = 4 waves of 64)
// ---------------------------------------------------------------------------
constexpr int BT = 256;
constexpr int NWAVE = BT / 64;

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T x)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        const T y = __shfl_up(x, d, 64);
        if (lane >= d)
            x += y;
    }
    return x;
}

// Exclusive scan over a block of NW waves; `sh` holds NW elements of LDS.
// Contains barriers: every thread of the block must call it.
template <int NW, typename T>
__device__ __forceinline__ T block_excl_scan_w(T v, T *sh, T &total)
{
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const T inc = wave_incl_scan(v);
    if (lane == 63)
        sh[wid] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w)
    {
        const T t = sh[w];
        pre += (w < wid) ? t : T(0);
        tot += t;
    }
    __syncthreads();
    total = tot;
    return pre + inc - v;
}

// the same for blocks of BT threads
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *sh, T &total)
{
    return block_excl_scan_w<NWAVE>(v, sh, total);
}

template <typename T>
__device__ __forceinline__ T block_reduce_sum(T v, T *sh)
{
    T tot;
    block_excl_scan(v, sh, tot);
    return tot;
}

// smallest index i in [0, n) with incl[i] > g (incl is non-decreasing)
template <typename T>
__device__ __forceinline__ int upper_find(const T *incl, int n, T g)
{
    int lo = 0, hi = n - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (incl[mid] > g)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ int sel3i(int i, int a, int b, int c) { return i == 0 ? a : (i == 1 ? b : c); }
__device__ __forceinline__ float sel3f(int i, float a, float b, float c) { return i == 0 ? a : (i == 1 ? b : c); }

// bitset word / bit of an absolute voxel coordinate
__device__ __forceinline__ uint64_t word_of(const GridDev &g, int c0, int c1, int c2, int &bit)
{
    const int lx = c0 - g.gmin[0];
    const int ly = c1 - g.gmin[1];
    const int lz = c2 - g.gmin[2];
    bit = lx & 63;
    return ((uint64_t)lz * (uint64_t)g.G[1] + (uint64_t)ly) * (uint64_t)g.wx + (uint64_t)(lx >> 6);
}

__device__ __forceinline__ uint32_t voxel_id(const GridDev &g, uint64_t w, int bit)
{
    const uint64_t m = g.bits[w];
    return g.woff[w] + (uint32_t)__popcll(m & ((1ull << bit) - 1ull));
}

// brick of an absolute voxel coordinate: id in the dense brick grid of the
// mesh (bricks of 2^bs voxels along each axis) and the voxel's index inside
// its brick, laid out as in an 8x8x8 brick (x | y << 3 | z << 6) for either size
__device__ __forceinline__ void brick_of(const MeshSeg &sg, int c0, int c1, int c2, uint32_t &brick, uint32_t &lvox)
{
    const GridDev &g = sg.g;
    const int lx = c0 - g.gmin[0], ly = c1 - g.gmin[1], lz = c2 - g.gmin[2];
    const int bs = sg.bs, m = (1 << bs) - 1;
    brick = (uint32_t)(lx >> bs) + (uint32_t)sg.nb[0] * ((uint32_t)(ly >> bs) + (uint32_t)sg.nb[1] * (uint32_t)(lz >> bs));
    lvox = (uint32_t)((lx & m) | ((ly & m) << 3) | ((lz & m) << 6));
}

// Wave-aggregated append: one atomic per wave.  All lanes of the wave must call it.
__device__ __forceinline__ uint64_t wave_append(bool pred, unsigned long long *counter)
{
    const uint64_t m = __ballot(pred);
    const int lane = threadIdx.x & 63;
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    unsigned long long base = 0;
    if (m)
    {
        const int leader = __ffsll((long long)m) - 1;
        if (lane == leader)
            base = atomicAdd(counter, (unsigned long long)__popcll(m));
        base = __shfl(base, leader, 64);
    }
    return (uint64_t)base + below;
}

// Claims n items of shard `sh` of b; returns the global index of the first.
// ok = false when the region is full (the counter still advances, so the
// totals stay exact for re-sizing).
template <typename T>
__device__ __forceinline__ uint64_t shard_claim(const ShardBuf<T> &b, uint32_t sh, uint64_t n, bool &ok)
{
    const unsigned long long off = atomicAdd(b.cnt + (uint64_t)sh * SHARD_STRIDE, (unsigned long long)n);
    ok = off + n <= b.per;
    return (uint64_t)sh * b.per + off;
}

// Claims n items of shard `sh` only if they fit: the counter never passes
// the capacity, so every slot below it holds an item (for queues that are
// consumed without a re-run; a failed claim is handled by the caller).
template <typename T>
__device__ __forceinline__ bool shard_claim_fit(const ShardBuf<T> &b, uint32_t sh, uint64_t n, uint64_t &first)
{
    unsigned long long *cnt = b.cnt + (uint64_t)sh * SHARD_STRIDE;
    unsigned long long cur = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (cur + n <= b.per)
    {
        const unsigned long long prev = atomicCAS(cnt, cur, cur + n);
        if (prev == cur)
        {
            first = (uint64_t)sh * b.per + cur;
            return true;
        }
        cur = prev;
    }
    return false;
}

// Dense numbering of the items of a sharded buffer: the exclusive prefix
// of the regions' fill counts, one region per thread, so that consumers walk
// the items rather than the capacity.  A region whose counter passed its
// capacity had a claim dropped and counts as empty (the host re-runs the
// producer with larger regions then; all-or-nothing queues never overflow).
// Contains barriers.
template <typename T>
__device__ __forceinline__ uint64_t queue_prefix(const ShardBuf<T> &b, uint64_t *pre, uint64_t *sh)
{
    static_assert(NSHARD == BT, "one region per thread");
    const unsigned long long c = b.cnt[(uint64_t)threadIdx.x * SHARD_STRIDE];
    const uint64_t n = c <= b.per ? (uint64_t)c : 0ull;
    uint64_t tot;
    const uint64_t ex = block_excl_scan(n, sh, tot);
    pre[threadIdx.x] = ex + n; // inclusive
    __syncthreads();
    return tot;
}

// slot of the v-th item (v < total)
template <typename T>
__device__ __forceinline__ uint64_t queue_slot(const ShardBuf<T> &b, const uint64_t *pre, uint64_t v)
{
    const int s = upper_find(pre, NSHARD, v);
    return (uint64_t)s * b.per + (v - (s ? pre[s - 1] : 0ull));
}

// dense numbering in LDS (one block's copy)
struct QueueMap
{
    uint64_t pre[NSHARD];
    uint64_t sh[NWAVE];
};

// LDS table of one block's records grouped by brick
template <int LOG2>
struct BucketLds
{
    static constexpr int N = 1 << LOG2;
    uint32_t key[N]; // brick of the slot, then the first record of its chunk
    uint32_t cnt[N];
    unsigned long long sh[NWAVE];
    unsigned long long rbase, cbase;
    bool fits;
};

// Block-level bucketed emission.  Each thread passes K records with their
// brick ids, validity flags and voxel entries (the voxels of the brick a
// record stands for: 1 to 4 for a crossing, 1 otherwise).  Records of one
// brick are counted in LDS (records in the low 16 bits of a table slot's
// count, entries in the high 16); the block reserves space for all of its
// records and chunks with one atomic each (on the counters of its shard),
// gives every brick it touched one contiguous chunk, adds the chunk and its
// entries to the brick's counts, and writes the records contiguously.
// Capacity misses are counted (the totals stay exact) and nothing is written
// out of bounds.  Contains barriers: every thread of the block must call it.
template <int K, int LOG2>
__device__ __forceinline__ void bucket_emit(const uint32_t (&bk)[K], const Rec (&rc)[K], const bool (&ok)[K],
                                            const uint32_t (&ent)[K], const BucketArgs &a, BucketLds<LOG2> &L)
{
    static_assert(BT * K < 65536 && 4 * BT * K < 65536, "records and entries of one emission in 16 bits each");
    constexpr int N = BucketLds<LOG2>::N;
    constexpr int PER = N / BT; // table entries per thread
    static_assert(N > BT * K, "bucket table too small");
    static_assert(N % BT == 0, "table size must be a multiple of the block size");
    for (int s = threadIdx.x; s < N; s += BT)
    {
        L.key[s] = 0xffffffffu;
        L.cnt[s] = 0u;
    }
    __syncthreads();
    int slot[K];
    uint32_t rank[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        slot[k] = 0;
        rank[k] = 0;
    }
    // a thread's records of one brick (the 4 voxels around a crossing mostly
    // share one) take one table lookup and one count update together
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        if (!ok[k])
            continue;
        bool led = false; // an earlier record of this thread has the brick
#pragma unroll
        for (int j = 0; j < k; ++j)
            led = led || (ok[j] && bk[j] == bk[k]);
        if (led)
            continue;
        uint32_t n = 1, ne = ent[k];
#pragma unroll
        for (int j = k + 1; j < K; ++j)
            if (ok[j] && bk[j] == bk[k])
            {
                ++n;
                ne += ent[j];
            }
        uint32_t h = (bk[k] * 2654435761u) >> (32 - LOG2);
        while (true)
        {
            const uint32_t prev = atomicCAS(&L.key[h], 0xffffffffu, bk[k]);
            if (prev == 0xffffffffu || prev == bk[k])
                break;
            h = (h + 1u) & (uint32_t)(N - 1);
        }
        const uint32_t r0 = atomicAdd(&L.cnt[h], n | (ne << 16)) & 0xffffu;
        slot[k] = (int)h;
        rank[k] = r0;
        uint32_t i = 1;
#pragma unroll
        for (int j = k + 1; j < K; ++j)
            if (ok[j] && bk[j] == bk[k])
            {
                slot[j] = (int)h;
                rank[j] = r0 + i;
                ++i;
            }
    }
    __syncthreads();
    // this thread's entries: PER consecutive table slots
    const int s0 = threadIdx.x * PER;
    unsigned long long my_n = 0, my_c = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e)
    {
        const uint32_t n = L.cnt[s0 + e] & 0xffffu;
        my_n += n;
        my_c += n ? 1ull : 0ull;
    }
    unsigned long long tot_n, tot_c;
    const unsigned long long ex_n = block_excl_scan(my_n, L.sh, tot_n);
    const unsigned long long ex_c = block_excl_scan(my_c, L.sh, tot_c);
    if (threadIdx.x == 0)
    {
        const uint32_t sh = blockIdx.x % NSHARD;
        bool ok_r = true, ok_c = true;
        L.rbase = tot_n ? shard_claim(a.recs, sh, tot_n, ok_r) : 0ull;
        L.cbase = tot_c ? shard_claim(a.chunks, sh, tot_c, ok_c) : 0ull;
        L.fits = ok_r && ok_c;
    }
    __syncthreads();
    const bool fits = L.fits;
    unsigned long long rpos = L.rbase + ex_n, cpos = L.cbase + ex_c;
#pragma unroll
    for (int e = 0; e < PER; ++e)
    {
        const uint32_t n = L.cnt[s0 + e] & 0xffffu, ne = L.cnt[s0 + e] >> 16;
        if (n && fits)
        {
            // slots s0 .. s0+PER-1 are this thread's alone here
            const uint32_t b = L.key[s0 + e];
            // one chunk and ne voxel entries more for the brick; the old chunk count is this chunk's rank
            const uint32_t r = (uint32_t)atomicAdd(a.brick_cnt + b, 1ull | ((unsigned long long)ne << 32));
            a.chunks.items[cpos] = Chunk{(uint32_t)rpos, n, b, r};
            L.key[s0 + e] = (uint32_t)rpos;
        }
        rpos += n;
        cpos += n ? 1ull : 0ull;
    }
    __syncthreads();
    if (fits)
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (ok[k])
                a.recs.items[(uint64_t)L.key[slot[k]] + rank[k]] = rc[k];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// prep: validate inputs, compute per-face planes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BT) void k_prep(MeshDev m, BatchDev bt, float4 *planes, float *tri,
                                             unsigned long long *counters, unsigned long long *est_ctr)
{
    const int64_t stride = (int64_t)gridDim.x * BT;
    unsigned flags = 0;
    unsigned long long est = 0; // face-test hits expected of this thread's faces
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < m.F; i += stride)
    {
        const int ms = mesh_of_face(bt, i);
        if (m.f)
        {
            // indices are batch-global: each must name a vertex of the face's own mesh
            const int64_t v0 = bt.vert_off[ms], v1 = bt.vert_off[ms + 1];
            const int32_t *fi = m.f + 3 * i;
            const int64_t a = fi[0], b = fi[1], c = fi[2];
            if (a < v0 || b < v0 || c < v0 || a >= v1 || b >= v1 || c >= v1)
            {
                flags |= ERR_FACE_INDEX;
                continue;
            }
        }
        if (planes)
        {
            float v0[3], v1[3], v2[3], pl[4];
            load_tri(m, i, v0, v1, v2);
            tri_plane(v0, v1, v2, pl);
            planes[i] = make_float4(pl[0], pl[1], pl[2], pl[3]);
            {
                // voxels a triangle overlaps: about the sum of its areas
                // projected on the three axis planes, half its perimeter
                // measured along the axes, and one (all in voxel units); sizes
                // the face-hit buffer, whose overflow is detected anyway
                const GridDev &g = bt.seg[ms].g;
                float e0[3], e1[3], e2[3];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                {
                    e0[a] = div_vs(v1[a] - v0[a], g, a);
                    e1[a] = div_vs(v2[a] - v1[a], g, a);
                    e2[a] = div_vs(v0[a] - v2[a], g, a);
                }
                const float area = fabsf(e0[1] * e2[2] - e0[2] * e2[1]) + fabsf(e0[2] * e2[0] - e0[0] * e2[2]) +
                                   fabsf(e0[0] * e2[1] - e0[1] * e2[0]);
                float per = 0.0f;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    per += fabsf(e0[a]) + fabsf(e1[a]) + fabsf(e2[a]);
                const float x = 0.5f * (area + per) + 1.0f;
                est += x < 1.0e9f ? (unsigned long long)x + 1ull : 1000000000ull;
            }
            if (tri)
            {
                // the gathered triangle, read coalesced by the later stages
                float *t = tri + 9 * i;
#pragma unroll
                for (int k = 0; k < 3; ++k)
                {
                    t[k] = v0[k];
                    t[3 + k] = v1[k];
                    t[6 + k] = v2[k];
                }
            }
        }
    }
    // vertices: finite and |v / vs| < 2^30 (so every int conversion of the
    // scanline and DDA is in range)
    const int64_t nv = m.f ? m.V : 3 * m.F;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < nv; i += stride)
    {
        const GridDev &g = bt.seg[mesh_of_vert(bt, i)].g;
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            const float x = m.v[3 * i + a];
            if (!(fabsf(x) <= 3.0e38f))
                flags |= ERR_NONFINITE;
            else if (!(fabs(div_vs((double)x, g, a)) < 1073741824.0))
                flags |= ERR_RANGE;
        }
    }
    if (flags)
        atomicOr(counters + CTR_ERRORS, (unsigned long long)flags);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1)
        est += __shfl_xor(est, d, 64);
    if ((threadIdx.x & 63) == 0 && est)
        atomicAdd(est_ctr + (uint64_t)((blockIdx.x * NWAVE + (threadIdx.x >> 6)) % NSHARD) * SHARD_STRIDE, est);
}

// ---------------------------------------------------------------------------
// scanline events (FDG.cpp:195-283)
// ---------------------------------------------------------------------------
struct ScanItem
{
    float t[3][3]; // y-sorted vertices in (ax0, ax1, ax2) coordinates
    int start, mid, end; // rows [start, end) handled here (a part of a large triangle's rows); mid splits the halves
    int ax2;
    uint32_t face;
};

// A scanline item whose rows x columns exceed MARK_VOX_CAP is split into
// row chunks of about that many cells; all but the last go to the row queue
constexpr int MARK_VOX_CAP = 512;
constexpr uint32_t QUEUE_GROUP = 32; // row chunks claimed together (from one shard)

// Sorts the triangle's vertices by the scan axis and returns its row range
// [s, e) and middle row (FDG.cpp:195-230).
__device__ __forceinline__ void scan_item_setup(const MeshDev &m, const GridDev &g, int64_t face, int ax2,
                                                ScanItem &it)
{
    const int ax0 = axis_next(ax2);
    const int ax1 = axis_prev(ax2);
    float v[3][3];
    load_tri(m, face, v[0], v[1], v[2]);
    float t[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        t[k][0] = sel3f(ax0, v[k][0], v[k][1], v[k][2]);
        t[k][1] = sel3f(ax1, v[k][0], v[k][1], v[k][2]);
        t[k][2] = sel3f(ax2, v[k][0], v[k][1], v[k][2]);
    }
    // stable sort of 3 by y (std::sort on 3 elements is insertion sort)
    auto swp = [&](int a, int b) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            const float tmp = t[a][i];
            t[a][i] = t[b][i];
            t[b][i] = tmp;
        }
    };
    if (t[1][1] < t[0][1])
        swp(0, 1);
    if (t[2][1] < t[1][1])
    {
        swp(1, 2);
        if (t[1][1] < t[0][1])
            swp(0, 1);
    }
    const int lo = g.gmin[ax1], hi = g.gmax[ax1] - 1;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            it.t[k][i] = t[k][i];
    it.start = iclamp((int)div_vs((double)t[0][1], g, ax1), lo, hi);
    it.mid = iclamp((int)div_vs((double)t[1][1], g, ax1), lo, hi);
    it.end = iclamp((int)div_vs((double)t[2][1], g, ax1), lo, hi);
    it.ax2 = ax2;
    it.face = (uint32_t)face;
}

struct ScanRow
{
    double t3x, t3z, t4x, t4z;
    // row and first column as offsets from the grid's minimum (rows and
    // columns are clamped to the grid, whose extent is at most 8192), and
    // the row's item in the tile: 40 B a row
    uint16_t y_off;
    uint16_t x_off;
    uint16_t item;
};

// the 4 voxels around a crossing, in the reference's (dx, dy) order
__device__ __forceinline__ void event_voxels(int ax0, int ax1, int x_idx, int y_idx, int z_idx, int (&c)[4][3])
{
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
        {
            const int k = dx * 2 + dy;
            const int cx = x_idx + dx, cy = y_idx + dy;
            c[k][0] = ax0 == 0 ? cx : (ax1 == 0 ? cy : z_idx);
            c[k][1] = ax0 == 1 ? cx : (ax1 == 1 ? cy : z_idx);
            c[k][2] = ax0 == 2 ? cx : (ax1 == 2 ? cy : z_idx);
        }
}

// Block-local OR-combining table for occupancy words: the crossings of a
// block's faces fall on few bitset words, so their bits are merged in LDS and
// each distinct word is written with one global atomicOr when the table is
// flushed (device-scope atomics are far costlier than LDS ones).
// merge table entries: 2048 for the smallest tiles (meshes of large
// triangles, whose crossings are densest), 1024 otherwise (a smaller table
// leaves LDS for a fourth block per CU)
constexpr int mark_log2(int tile)
{
    return tile <= 64 ? 11 : 10;
}
constexpr uint32_t MARK_EMPTY = 0xffffffffu; // word indices are < 2^32 - 1

template <int LOG2>
struct MarkTable
{
    static constexpr int LOG = LOG2;
    static constexpr int N = 1 << LOG2;
    uint32_t key[N];
    unsigned long long mask[N];
};
struct NoMarkTable // kernel variant without merging: no LDS
{
    static constexpr int LOG = 0;
    static constexpr int N = 1;
    uint32_t key[1];
    unsigned long long mask[1];
};

template <class Tab>
__device__ __forceinline__ void mark_word(const GridDev &g, Tab &T, uint64_t w, uint64_t msk)
{
    const uint32_t key = (uint32_t)w;
    uint32_t h = (key * 2654435761u) >> (32 - Tab::LOG);
#pragma unroll 1
    for (int probe = 0; probe < 4; ++probe)
    {
        const uint32_t prev = atomicCAS(&T.key[h], MARK_EMPTY, key);
        if (prev == MARK_EMPTY || prev == key)
        {
            atomicOr(&T.mask[h], (unsigned long long)msk);
            return;
        }
        h = (h + 1u) & (uint32_t)(Tab::N - 1);
    }
    atomicOr((unsigned long long *)(g.bits + w), (unsigned long long)msk); // table crowded
}

// writes the table's words to the bitset and empties it; contains barriers
template <class Tab>
__device__ __forceinline__ void mark_flush(const GridDev &g, Tab &T)
{
    __syncthreads();
    for (int i = threadIdx.x; i < Tab::N; i += BT)
    {
        const uint32_t key = T.key[i];
        if (key != MARK_EMPTY)
        {
            atomicOr((unsigned long long *)(g.bits + key), T.mask[i]);
            T.key[i] = MARK_EMPTY;
            T.mask[i] = 0ull;
        }
    }
    __syncthreads();
}

// set the occupancy bits of the 4 voxels of a crossing (one update per word,
// in the table or directly)
template <bool MERGE, class Tab>
__device__ __forceinline__ void mark_event(const GridDev &g, Tab &T, int ax0, int ax1, int x_idx, int y_idx, int z_idx)
{
    int c[4][3];
    event_voxels(ax0, ax1, x_idx, y_idx, z_idx, c);
    uint64_t w[4], msk[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        int b;
        w[k] = word_of(g, c[k][0], c[k][1], c[k][2], b);
        msk[k] = 1ull << b;
    }
#pragma unroll
    for (int k = 1; k < 4; ++k)
    {
        bool merged = false;
#pragma unroll
        for (int j = 0; j < k; ++j)
            if (!merged && msk[j] != 0 && w[j] == w[k])
            {
                msk[j] |= msk[k];
                merged = true;
            }
        if (merged)
            msk[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (msk[k])
        {
            if constexpr (MERGE)
                mark_word(g, T, w[k], msk[k]);
            else
                atomicOr((unsigned long long *)(g.bits + w[k]), (unsigned long long)msk[k]);
        }
}

// MERGE: the occupancy bits of a block iteration's crossings are merged in an
// LDS table first (pays when crossings are dense; the table costs occupancy)
// QUEUED = false: one item per (face, axis); QUEUED = true: the queued row
// chunks of large triangles (items index the queue's slots)
template <bool MERGE, bool QUEUED, int MARK_TILE>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4))) void k_scan_events(MeshDev m, BatchDev bt, uint64_t n_tiles,
                                                    unsigned long long *counters, ShardBuf<Event> eb,
                                                    ShardBuf<RowChunk> rq)
{
    static_assert(MARK_TILE <= BT, "a tile's items are set up one per thread");
    __shared__ unsigned long long sbase;
    __shared__ bool sfit;
    __shared__ ScanItem items[MARK_TILE];
    __shared__ ScanRow rows[BT];
    __shared__ uint32_t incl_items[MARK_TILE];
    __shared__ uint32_t incl_rows[BT];
    __shared__ uint32_t sh[NWAVE];
    __shared__ typename std::conditional<MERGE, MarkTable<mark_log2(MARK_TILE)>, NoMarkTable>::type T;
    __shared__ QueueMap qm;
    const int tid = threadIdx.x;
    int64_t n_items = 0;
    if constexpr (QUEUED)
    {
        n_items = (int64_t)queue_prefix(rq, qm.pre, qm.sh);
        n_tiles = (uint64_t)(n_items + MARK_TILE - 1) / MARK_TILE;
    }
    uint32_t n_ev = 0;
    if constexpr (MERGE)
    {
        for (int i = tid; i < T.N; i += BT)
        {
            T.key[i] = MARK_EMPTY;
            T.mask[i] = 0ull;
        }
        __syncthreads();
    }

    // a block takes tiles of BT items of one mesh (queued row chunks exist
    // for single-mesh conversions only)
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        // the tile's mesh; its grid is read in place (a local copy of the
        // struct would go to scratch memory under runtime indexing)
        const int ms = QUEUED ? 0 : seg_find(bt.mtile_off, bt.M, tile);
        const GridDev &g = bt.seg[ms].g;
        int64_t item;
        bool live;
        if constexpr (!QUEUED)
        {
            const int64_t f0 = bt.face_off[ms];
            const int64_t loc = (int64_t)(tile - bt.mtile_off[ms]) * MARK_TILE + tid;
            item = 3 * f0 + loc;
            live = tid < MARK_TILE && loc < 3 * (bt.face_off[ms + 1] - f0);
        }
        else
        {
            item = (int64_t)tile * MARK_TILE + tid;
            live = tid < MARK_TILE && item < n_items;
        }
        uint32_t R = 0;
        if constexpr (!QUEUED)
        {
            if (live)
            {
                const int64_t face = item / 3;
                ScanItem &it = items[tid];
                scan_item_setup(m, g, face, (int)(item - face * 3), it);
                // a non-finite vertex (reported as an input error after this
                // kernel) can break the sort: an empty range then, never a
                // negative one
                R = it.end > it.start ? (uint32_t)(it.end - it.start) : 0u;
                // columns of a row are bounded by the x extent of the triangle
                const int a0 = axis_next(it.ax2);
                const float xlo = div_vs(sminf(sminf(it.t[0][0], it.t[1][0]), it.t[2][0]), g, a0);
                const float xhi = div_vs(smaxf(smaxf(it.t[0][0], it.t[1][0]), it.t[2][0]), g, a0);
                const float ext = xhi - xlo + 2.0f;
                const uint32_t rpc =
                    ext < (float)MARK_VOX_CAP ? (uint32_t)((float)MARK_VOX_CAP / (ext > 1.0f ? ext : 1.0f)) : 1u;
                if (R > rpc && rq.items)
                {
                    // all but the last rpc rows go to the row queue, in chunks
                    // of rpc rows, claimed QUEUE_GROUP chunks at a time from
                    // successive shards; the rows of the first group that does
                    // not fit stay with this block
                    const uint32_t nq = (R - 1) / rpc;
                    uint32_t done = 0;
                    for (uint32_t c0 = 0; c0 < nq; c0 += QUEUE_GROUP)
                    {
                        const uint32_t n = nq - c0 < QUEUE_GROUP ? nq - c0 : QUEUE_GROUP;
                        uint64_t q0;
                        if (!shard_claim_fit(rq, (uint32_t)((item + c0 / QUEUE_GROUP) % NSHARD), n, q0))
                            break;
                        for (uint32_t c = 0; c < n; ++c)
                        {
                            const int r0 = it.start + (int)(rpc * (c0 + c));
                            rq.items[q0 + c] = RowChunk{it.face, it.ax2, r0, r0 + (int)rpc};
                        }
                        done = c0 + n;
                    }
                    it.start += (int)(rpc * done);
                    R = (uint32_t)(it.end - it.start);
                }
            }
        }
        else
        {
            if (live)
            {
                const RowChunk ch = rq.items[queue_slot(rq, qm.pre, (uint64_t)item)];
                ScanItem &it = items[tid];
                scan_item_setup(m, g, (int64_t)ch.face, ch.ax2, it);
                it.start = ch.r0;
                it.end = ch.r1;
                R = (uint32_t)(ch.r1 - ch.r0);
            }
        }
        uint32_t tot_r;
        const uint32_t ex_r = block_excl_scan(R, sh, tot_r);
        if (tid < MARK_TILE)
            incl_items[tid] = ex_r + R;
        __syncthreads();

        for (uint32_t rb = 0; rb < tot_r; rb += BT)
        {
            const uint32_t gi = rb + tid;
            uint32_t C = 0;
            if (gi < tot_r)
            {
                const int L = upper_find(incl_items, MARK_TILE, gi);
                const ScanItem &it = items[L];
                const uint32_t R_L = (uint32_t)(it.end - it.start);
                const int y_idx = it.start + (int)(gi - (incl_items[L] - R_L));
                const int ax2 = it.ax2;
                const int ax0 = axis_next(ax2);
                const int ax1 = axis_prev(ax2);
                const bool first = y_idx < it.mid;
                const int i0 = first ? 0 : 2, i2 = first ? 2 : 0;
                const double t0x = it.t[i0][0], t0y = it.t[i0][1], t0z = it.t[i0][2];
                const double t1x = it.t[1][0], t1y = it.t[1][1], t1z = it.t[1][2];
                const double t2x = it.t[i2][0], t2y = it.t[i2][1], t2z = it.t[i2][2];
                const float vs1 = sel3f(ax1, g.vs[0], g.vs[1], g.vs[2]);
                const double y = (double)((float)(y_idx + 1) * vs1);
                double t3x = lerp1(t0y, t1y, y, t0x, t1x);
                double t3z = lerp1(t0y, t1y, y, t0z, t1z);
                double t4x = lerp1(t0y, t2y, y, t0x, t2x);
                double t4z = lerp1(t0y, t2y, y, t0z, t2z);
                if (t3x > t4x)
                {
                    double tmp = t3x;
                    t3x = t4x;
                    t4x = tmp;
                    tmp = t3z;
                    t3z = t4z;
                    t4z = tmp;
                }
                const int lo = sel3i(ax0, g.gmin[0], g.gmin[1], g.gmin[2]);
                const int hi = sel3i(ax0, g.gmax[0], g.gmax[1], g.gmax[2]) - 1;
                const int ls = iclamp((int)div_vs(t3x, g, ax0), lo, hi);
                const int le = iclamp((int)div_vs(t4x, g, ax0), lo, hi);
                ScanRow &rw = rows[tid];
                rw.t3x = t3x;
                rw.t3z = t3z;
                rw.t4x = t4x;
                rw.t4z = t4z;
                rw.y_off = (uint16_t)(y_idx - sel3i(ax1, g.gmin[0], g.gmin[1], g.gmin[2]));
                rw.x_off = (uint16_t)(ls - lo);
                rw.item = (uint16_t)L;
                C = le > ls ? (uint32_t)(le - ls) : 0u; // le < ls only for non-finite input
            }
            uint32_t tot_c;
            const uint32_t ex_c = block_excl_scan(C, sh, tot_c);
            incl_rows[tid] = ex_c + C;
            __syncthreads();
            for (uint32_t cb = 0; cb < tot_c; cb += BT)
            {
                const uint32_t h = cb + tid;
                Event evr;
                uint32_t has = 0;
                if (h < tot_c)
                {
                    const int rr = upper_find(incl_rows, BT, h);
                    const ScanRow &rw = rows[rr];
                    const ScanItem &it = items[rw.item];
                    const uint32_t ex = (rr == 0) ? 0u : incl_rows[rr - 1];
                    const int ax2 = it.ax2;
                    const int ax0 = axis_next(ax2);
                    const int ax1 = axis_prev(ax2);
                    const int ylo = sel3i(ax1, g.gmin[0], g.gmin[1], g.gmin[2]);
                    const int xlo = sel3i(ax0, g.gmin[0], g.gmin[1], g.gmin[2]);
                    const int y_idx = ylo + (int)rw.y_off;
                    const int x_idx = xlo + (int)rw.x_off + (int)(h - ex);
                    const float vs0 = sel3f(ax0, g.vs[0], g.vs[1], g.vs[2]);
                    const double x = (double)((float)(x_idx + 1) * vs0);
                    const double z = lerp1(rw.t3x, rw.t4x, x, rw.t3z, rw.t4z);
                    const int z_idx = (int)div_vs(z, g, ax2);
                    const int zlo = sel3i(ax2, g.gmin[0], g.gmin[1], g.gmin[2]);
                    const int zhi = sel3i(ax2, g.gmax[0], g.gmax[1], g.gmax[2]);
                    if (z_idx >= zlo && z_idx < zhi)
                    {
                        ++n_ev;
                        const uint64_t keybase = ((uint64_t)it.face << KEY_FACE_SHIFT) |
                                                 ((uint64_t)ax2 << KEY_AX_SHIFT) |
                                                 ((uint64_t)rw.y_off << KEY_Y_SHIFT) |
                                                 ((uint64_t)(x_idx - xlo) << KEY_X_SHIFT);
                        mark_event<MERGE>(g, T, ax0, ax1, x_idx, y_idx, z_idx);
                        evr.key = keybase;
                        evr.z = (float)z;
                        evr.zi = z_idx;
                        has = 1;
                    }
                }
                if (eb.items)
                {
                    // block-aggregated append: one atomic per block iteration, consecutive slots
                    uint32_t tot_e;
                    const uint32_t off = block_excl_scan(has, sh, tot_e);
                    if (tot_e)
                    {
                        if (tid == 0)
                        {
                            bool okk;
                            sbase = shard_claim(eb, blockIdx.x % NSHARD, tot_e, okk);
                            sfit = okk;
                        }
                        __syncthreads();
                        if (has && sfit)
                            eb.items[sbase + off] = evr;
                    }
                }
            }
            __syncthreads();
        }
        if constexpr (MERGE)
            mark_flush(g, T);
    }
    (void)counters;
    (void)n_ev;
}

// ---------------------------------------------------------------------------
// face term: triangle / voxel overlap test (FDG.cpp:304-413), active voxels only
// ---------------------------------------------------------------------------
struct FaceItem
{
    float n[3], d1, d2;
    float nxy[3][2], dxy[3], nyz[3][2], dyz[3], nzx[3][2], dzx[3];
    int bx0, bx1, by0, bz0, ny;
    uint32_t face;
    uint32_t r0; // first bounding-box row (linear (y, z) index) handled here
    uint32_t ms; // the face's mesh (queued row chunks of a batch mix meshes)
};
// A face whose bounding box holds more than FACE_VOX_CAP voxels is split into
// chunks of bounding-box rows of about that many voxels; all but the last are queued
constexpr uint32_t FACE_VOX_CAP = 1024;

struct FaceRow
{
    uint64_t word; // the row's active candidates when they lie in one bitset word (else 0)
    int item, y, z;
    int x0, x1; // candidate voxels [x0, x1) of the row
    int one;    // 1 when the candidates lie in one bitset word (`word`)
};

// Voxels of bounding-box row (y, z) that can pass the slab test of
// face_pass: with p the voxel's min corner, n.p must lie between -d1 and
// -d2, an interval of x for |n_x| large enough.  The interval is computed in
// double and widened by a bound on the FP32 test's rounding (a few ulp of
// the terms of n.p + d, over |n_x|) plus 2 voxels; for |n_x| < 0.05, and for
// rows of at most 4 voxels, the whole bounding-box row is kept.  The terms
// that do not depend on the row are formed once per face (face_row_pre).
struct RowNarrow
{
    double k;     // 1 / (n_x vs_x): voxels of x per unit of -(d + n_y p_y + n_z p_z)
    double base;  // |n_x| max(|bx0|, |bx1|) vs_x + |d1| + |d2|: the row-independent rounding terms
    double scale; // 1e-6 / (|n_x| vs_x): rounding terms -> voxels
    bool on;      // narrowing applies to this face
};

__device__ __forceinline__ RowNarrow face_row_pre(const FaceItem &it, const GridDev &g)
{
    RowNarrow r;
    const double nx = (double)it.n[0];
    r.on = fabs(nx) >= 0.05 && it.bx1 - it.bx0 > 4;
    const double vs0 = (double)g.vs[0];
    const double anv = fabs(nx) * vs0;
    r.k = r.on ? 1.0 / (nx * vs0) : 0.0;
    r.base = anv * fmax(fabs((double)it.bx0), fabs((double)it.bx1)) + fabs((double)it.d1) + fabs((double)it.d2);
    r.scale = r.on ? 1.0e-6 / anv : 0.0;
    return r;
}

__device__ __forceinline__ void face_row_range(const FaceItem &it, const GridDev &g, const RowNarrow &r, int y, int z,
                                               int &x0, int &x1)
{
    x0 = it.bx0;
    x1 = it.bx1;
    if (!r.on)
        return;
    const double py = (double)g.vs[1] * (double)y, pz = (double)g.vs[2] * (double)z;
    const double ky = (double)it.n[1] * py, kz = (double)it.n[2] * pz;
    const double K = ky + kz;
    const double a = (-(double)it.d1 - K) * r.k, b = (-(double)it.d2 - K) * r.k;
    const double margin = 2.0 + (r.base + fabs(ky) + fabs(kz)) * r.scale;
    const double lo = fmax(fmin(a, b) - margin, (double)it.bx0 - 1.0);
    const double hi = fmin(fmax(a, b) + margin, (double)it.bx1 + 1.0);
    if (!(lo <= hi))
    {
        x1 = x0; // empty (also for non-finite planes)
        return;
    }
    const int l = (int)floor(lo), h = (int)floor(hi) + 1;
    x0 = l > x0 ? l : x0;
    x1 = h < x1 ? h : x1;
    if (x1 < x0)
        x1 = x0;
}

// Candidate range of bounding-box row (y, z) for the block-cooperative
// face test, whose faces have long rows: besides the slab, the row is cut to
// where the 6 projected-edge half-planes that depend on x can hold, and
// dropped when one of the 3 that do not (the yz ones, evaluated exactly as
// face_pass does) fails.  Each x bound comes from the exact line a vs_x x =
// -(c + d) moved by 1e-6 of the test's terms (the FP32 evaluation errs by
// less than 4 ulp of them) and by 2 more voxels, so every voxel that passes
// face_pass stays in the range; face_pass re-tests them all.
struct RowNarrowFull
{
    double ks;    // 1 / (n_x vs_x), 0 when n_x = 0
    double sbase; // |n_x| pxm + |d1| + |d2|
    double pxm;   // vs_x max(|bx0|, |bx1|): bound on |p_x| over the box
    double ke[6]; // 1 / (a vs_x) of the xy edges' and the zx edges' x coefficient a, 0 when a = 0
};

__device__ __forceinline__ RowNarrowFull face_row_pre_full(const FaceItem &it, const GridDev &g)
{
    RowNarrowFull r;
    const double vs0 = (double)g.vs[0];
    const double nx = (double)it.n[0];
    r.pxm = vs0 * fmax(fabs((double)it.bx0), fabs((double)it.bx1));
    r.ks = nx != 0.0 ? 1.0 / (nx * vs0) : 0.0;
    r.sbase = fabs(nx) * r.pxm + fabs((double)it.d1) + fabs((double)it.d2);
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const double a = (double)it.nxy[k][0], b = (double)it.nzx[k][1];
        r.ke[k] = a != 0.0 ? 1.0 / (a * vs0) : 0.0;
        r.ke[3 + k] = b != 0.0 ? 1.0 / (b * vs0) : 0.0;
    }
    return r;
}

// one x-dependent half-plane a p_x + c + d >= 0 (c the FP32 product of the
// other coordinate's term, as face_pass forms it); false when the row fails
// it for every x (a = 0 and the exact FP32 test fails)
__device__ __forceinline__ bool row_edge_bound(float a, float c, float d, double ke, double pxm, float p0, double &lo,
                                               double &hi)
{
    if (ke == 0.0)
        return !((a * p0 + c) + d < 0.0f);
    const double m = 1.0e-6 * (fabs((double)a) * pxm + fabs((double)c) + fabs((double)d));
    const double t = (-((double)c + (double)d) - m) * ke;
    if (a > 0.0f)
        lo = fmax(lo, t - 2.0);
    else
        hi = fmin(hi, t + 2.0);
    return true;
}

__device__ __forceinline__ void face_row_range_full(const FaceItem &it, const GridDev &g, const RowNarrowFull &r, int y,
                                                    int z, int &x0, int &x1)
{
    x0 = it.bx0;
    x1 = it.bx1;
    const float p1 = g.vs[1] * (float)y, p2 = g.vs[2] * (float)z;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if ((it.nyz[k][0] * p1 + it.nyz[k][1] * p2) + it.dyz[k] < 0.0f)
        {
            x1 = x0;
            return;
        }
    if (it.bx1 - it.bx0 <= 4)
        return; // short rows: the box range costs less than narrowing it
    double lo = (double)it.bx0, hi = (double)(it.bx1 - 1); // inclusive
    {
        // slab: n_x vs_x x within [-d1 - K, -d2 - K] (either order), K = n_y p_y + n_z p_z
        const double ky = (double)it.n[1] * ((double)g.vs[1] * (double)y);
        const double kz = (double)it.n[2] * ((double)g.vs[2] * (double)z);
        const double K = ky + kz;
        const double m = 1.0e-6 * (r.sbase + fabs(ky) + fabs(kz));
        const double u = -(double)it.d1 - K, v = -(double)it.d2 - K;
        const double tlo = fmin(u, v) - m, thi = fmax(u, v) + m;
        if (r.ks != 0.0)
        {
            const double a = tlo * r.ks, b = thi * r.ks;
            lo = fmax(lo, fmin(a, b) - 2.0);
            hi = fmin(hi, fmax(a, b) + 2.0);
        }
        else if (!(tlo <= 0.0 && 0.0 <= thi))
        {
            x1 = x0;
            return;
        }
    }
    const float p0 = g.vs[0] * (float)it.bx0; // only for half-planes that do not depend on x
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        ok = ok && row_edge_bound(it.nxy[k][0], it.nxy[k][1] * p1, it.dxy[k], r.ke[k], r.pxm, p0, lo, hi);
        ok = ok && row_edge_bound(it.nzx[k][1], it.nzx[k][0] * p2, it.dzx[k], r.ke[3 + k], r.pxm, p0, lo, hi);
    }
    if (!ok || !(lo <= hi))
    {
        x1 = x0; // empty (also for non-finite planes)
        return;
    }
    const int l = (int)floor(lo), h = (int)floor(hi) + 1;
    x0 = l > x0 ? l : x0;
    x1 = h < x1 ? h : x1;
    if (x1 < x0)
        x1 = x0;
}

__device__ __forceinline__ bool face_pass(const FaceItem &it, const float vs[3], int x, int y, int z)
{
    const float p0 = vs[0] * (float)x, p1 = vs[1] * (float)y, p2 = vs[2] * (float)z;
    const float ndp = dot3f(it.n[0], it.n[1], it.n[2], p0, p1, p2);
    if (((ndp + it.d1) * (ndp + it.d2)) > 0.0f)
        return false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if ((it.nxy[k][0] * p0 + it.nxy[k][1] * p1) + it.dxy[k] < 0.0f)
            return false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if ((it.nyz[k][0] * p1 + it.nyz[k][1] * p2) + it.dyz[k] < 0.0f)
            return false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if ((it.nzx[k][0] * p2 + it.nzx[k][1] * p0) + it.dzx[k] < 0.0f)
            return false;
    return true;
}

// position of the k-th (0-based) set bit of m; k < popcount(m)
__device__ __forceinline__ int select_bit(uint64_t m, uint32_t k)
{
    int pos = 0;
#pragma unroll
    for (int half = 32; half >= 1; half >>= 1)
    {
        const uint64_t low = m & ((1ull << half) - 1ull);
        const uint32_t c = (uint32_t)__popcll(low);
        if (k >= c)
        {
            k -= c;
            m >>= half;
            pos += half;
        }
        else
            m = low;
    }
    return pos;
}

// Face-test set-up of one face (FDG.cpp:304-360): plane, slab offsets, the
// 9 projected edge half-planes and the clamped bounding box; returns the
// number of bounding-box rows (y, z).
__device__ __forceinline__ uint64_t face_item_setup(const MeshDev &m, const GridDev &g, int64_t face, FaceItem &it)
{
    float v0[3], v1[3], v2[3], pl[4];
    load_tri(m, face, v0, v1, v2);
    if (m.pl)
    {
        // the plane the prep kernel computed with the same operations
        const float4 p4 = m.pl[face];
        pl[0] = p4.x;
        pl[1] = p4.y;
        pl[2] = p4.z;
    }
    else
        tri_plane(v0, v1, v2, pl);
    const float n0 = pl[0], n1 = pl[1], n2 = pl[2];
    int bmin[3], bmax[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const float mn = div_vs(sminf(sminf(v0[i], v1[i]), v2[i]), g, i);
        const float mx = div_vs(smaxf(smaxf(v0[i], v1[i]), v2[i]), g, i);
        const int a = (int)mn;
        const int b = (int)(mx + 1.0f);
        bmin[i] = a > g.gmin[i] ? a : g.gmin[i];
        bmax[i] = b < g.gmax[i] ? b : g.gmax[i];
    }
    const float c0 = n0 > 0.0f ? g.vs[0] : 0.0f;
    const float c1 = n1 > 0.0f ? g.vs[1] : 0.0f;
    const float c2 = n2 > 0.0f ? g.vs[2] : 0.0f;
    it.n[0] = n0;
    it.n[1] = n1;
    it.n[2] = n2;
    it.d1 = dot3f(n0, n1, n2, c0 - v0[0], c1 - v0[1], c2 - v0[2]);
    it.d2 = dot3f(n0, n1, n2, (g.vs[0] - c0) - v0[0], (g.vs[1] - c1) - v0[1], (g.vs[2] - c2) - v0[2]);
    const float mxy = (n2 < 0.0f) ? -1.0f : 1.0f;
    const float myz = (n0 < 0.0f) ? -1.0f : 1.0f;
    const float mzx = (n1 < 0.0f) ? -1.0f : 1.0f;
    const float *P[3] = {v0, v1, v2};
    float E[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        E[0][i] = v1[i] - v0[i];
        E[1][i] = v2[i] - v1[i];
        E[2][i] = v0[i] - v2[i];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const float *p = P[k];
        float a0 = (-mxy) * E[k][1], a1 = mxy * E[k][0];
        it.nxy[k][0] = a0;
        it.nxy[k][1] = a1;
        it.dxy[k] = -(a0 * p[0] + a1 * p[1]) + (smaxf(a0, 0.0f) * g.vs[0] + smaxf(a1, 0.0f) * g.vs[1]);
        a0 = (-myz) * E[k][2];
        a1 = myz * E[k][1];
        it.nyz[k][0] = a0;
        it.nyz[k][1] = a1;
        it.dyz[k] = -(a0 * p[1] + a1 * p[2]) + (smaxf(a0, 0.0f) * g.vs[1] + smaxf(a1, 0.0f) * g.vs[2]);
        a0 = (-mzx) * E[k][0];
        a1 = mzx * E[k][2];
        it.nzx[k][0] = a0;
        it.nzx[k][1] = a1;
        it.dzx[k] = -(a0 * p[2] + a1 * p[0]) + (smaxf(a0, 0.0f) * g.vs[2] + smaxf(a1, 0.0f) * g.vs[0]);
    }
    it.bx0 = bmin[0];
    it.bx1 = bmax[0];
    it.by0 = bmin[1];
    it.bz0 = bmin[2];
    it.ny = bmax[1] - bmin[1];
    it.face = (uint32_t)face;
    if (bmax[0] > bmin[0] && bmax[1] > bmin[1] && bmax[2] > bmin[2])
        return (uint64_t)(bmax[1] - bmin[1]) * (uint64_t)(bmax[2] - bmin[2]);
    return 0;
}

__device__ __forceinline__ uint64_t row_mask(int lo_bit, int hi_bit)
{
    // bits [lo_bit, hi_bit) of a 64-bit word, 0 <= lo_bit < hi_bit <= 64
    const uint64_t hi = (hi_bit >= 64) ? ~0ull : ((1ull << hi_bit) - 1ull);
    const uint64_t lo = (1ull << lo_bit) - 1ull;
    return hi & ~lo;
}

// QUEUED = false: one item per face; QUEUED = true: the queued row chunks of
// large faces (items index the queue's slots)
template <bool QUEUED, int FACE_TILE>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(7))) void k_face(MeshDev m, BatchDev bt, uint64_t n_tiles, ShardBuf<FHit> hb,
                                             unsigned long long *ctr, ShardBuf<FaceChunk> fq)
{
    static_assert(FACE_TILE <= BT, "a tile's faces are set up one per thread");
    __shared__ FaceItem items[FACE_TILE];
    __shared__ RowNarrowFull narrows[FACE_TILE];
    __shared__ FaceRow rows[BT];
    __shared__ uint64_t incl_items[FACE_TILE];
    __shared__ uint32_t incl_rows[BT];
    __shared__ uint64_t sh64[NWAVE];
    __shared__ uint32_t sh32[NWAVE];
    __shared__ QueueMap qm;
    const int tid = threadIdx.x;
    uint32_t n_hit = 0;
    int64_t n_items = 0;
    if constexpr (QUEUED)
    {
        n_items = (int64_t)queue_prefix(fq, qm.pre, qm.sh);
        n_tiles = (uint64_t)(n_items + FACE_TILE - 1) / FACE_TILE;
    }

    // a block takes tiles of FACE_TILE faces of one mesh (queued row chunks exist
    // for single-mesh conversions only)
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        // the tile's mesh and grid (a queued item names its own)
        const int ms = QUEUED ? 0 : seg_find(bt.ftile_off, bt.M, tile);
        const GridDev &gt = bt.seg[ms].g;
        int64_t item;
        bool live;
        if constexpr (!QUEUED)
        {
            const int64_t f0 = bt.face_off[ms];
            const int64_t loc = (int64_t)(tile - bt.ftile_off[ms]) * FACE_TILE + tid;
            item = f0 + loc;
            live = tid < FACE_TILE && loc < bt.face_off[ms + 1] - f0;
        }
        else
        {
            item = (int64_t)tile * FACE_TILE + tid;
            live = tid < FACE_TILE && item < n_items;
        }
        uint64_t R = 0;
        if constexpr (!QUEUED)
        {
            if (live)
            {
                FaceItem &it = items[tid];
                R = face_item_setup(m, gt, item, it);
                narrows[tid] = face_row_pre_full(it, gt);
                it.ms = (uint32_t)ms;
                it.r0 = 0;
                const uint32_t ext = (uint32_t)(it.bx1 > it.bx0 ? it.bx1 - it.bx0 : 1);
                const uint64_t rpc = ext < FACE_VOX_CAP ? FACE_VOX_CAP / ext : 1u;
                if (R > rpc && fq.items)
                {
                    // all but the last rpc rows go to the row queue, in chunks
                    // of rpc rows, claimed QUEUE_GROUP chunks at a time from
                    // successive shards; the rows of the first group that does
                    // not fit stay with this block
                    const uint64_t nq = (R - 1) / rpc;
                    uint64_t done = 0;
                    for (uint64_t c0 = 0; c0 < nq; c0 += QUEUE_GROUP)
                    {
                        const uint64_t n = nq - c0 < QUEUE_GROUP ? nq - c0 : QUEUE_GROUP;
                        uint64_t q0;
                        if (!shard_claim_fit(fq, (uint32_t)((item + c0 / QUEUE_GROUP) % NSHARD), n, q0))
                            break;
                        for (uint64_t c = 0; c < n; ++c)
                        {
                            const uint64_t r0 = rpc * (c0 + c);
                            fq.items[q0 + c] = FaceChunk{it.face, (uint32_t)r0, (uint32_t)(r0 + rpc), (uint32_t)ms};
                        }
                        done = c0 + n;
                    }
                    it.r0 = (uint32_t)(rpc * done);
                    R -= rpc * done;
                }
            }
        }
        else
        {
            if (live)
            {
                const FaceChunk ch = fq.items[queue_slot(fq, qm.pre, (uint64_t)item)];
                FaceItem &it = items[tid];
                face_item_setup(m, bt.seg[ch.ms].g, (int64_t)ch.face, it);
                narrows[tid] = face_row_pre_full(it, bt.seg[ch.ms].g);
                it.ms = ch.ms;
                it.r0 = ch.r0;
                R = (uint64_t)(ch.r1 - ch.r0);
            }
        }
        uint64_t tot_r;
        const uint64_t ex_r = block_excl_scan(R, sh64, tot_r);
        if (tid < FACE_TILE)
            incl_items[tid] = ex_r + R;
        __syncthreads();

        for (uint64_t rb = 0; rb < tot_r; rb += BT)
        {
            const uint64_t gi = rb + tid;
            uint32_t A = 0;
            if (gi < tot_r)
            {
                const int L = upper_find(incl_items, FACE_TILE, gi);
                const FaceItem &it = items[L];
                const GridDev &g = QUEUED ? bt.seg[it.ms].g : gt;
                const uint64_t R_L = incl_items[L] - (L ? incl_items[L - 1] : 0ull);
                // a bounding box has fewer than 2^26 rows: 32-bit division
                const uint32_t rr = it.r0 + (uint32_t)(gi - (incl_items[L] - R_L));
                const int y = it.by0 + (int)(rr % (uint32_t)it.ny);
                const int z = it.bz0 + (int)(rr / (uint32_t)it.ny);
                FaceRow &row = rows[tid];
                row.item = L;
                row.y = y;
                row.z = z;
                face_row_range_full(it, g, narrows[L], y, z, row.x0, row.x1);
                const int lx0 = row.x0 - g.gmin[0], lx1 = row.x1 - g.gmin[0];
                const uint64_t rowbase = ((uint64_t)(z - g.gmin[2]) * g.G[1] + (uint64_t)(y - g.gmin[1])) * g.wx;
                uint64_t last = 0ull;
                for (int w = lx0 >> 6; lx1 > lx0 && w <= ((lx1 - 1) >> 6); ++w)
                {
                    const int blo = (w == (lx0 >> 6)) ? (lx0 & 63) : 0;
                    const int bhi = (w == ((lx1 - 1) >> 6)) ? (((lx1 - 1) & 63) + 1) : 64;
                    last = g.bits[rowbase + w] & row_mask(blo, bhi);
                    A += (uint32_t)__popcll(last);
                }
                // a row inside one word keeps its candidates for the test below
                row.one = (lx1 > lx0 && (lx0 >> 6) == ((lx1 - 1) >> 6)) ? 1 : 0;
                row.word = row.one ? last : 0ull;
            }
            uint32_t tot_a;
            const uint32_t ex_a = block_excl_scan(A, sh32, tot_a);
            incl_rows[tid] = ex_a + A;
            __syncthreads();
            for (uint32_t ab = 0; ab < tot_a; ab += BT)
            {
                const uint32_t h = ab + tid;
                FHit hit;
                bool has = false;
                if (h < tot_a)
                {
                    const int rr = upper_find(incl_rows, BT, h);
                    const FaceRow &row = rows[rr];
                    const FaceItem &it = items[row.item];
                    const GridDev &g = QUEUED ? bt.seg[it.ms].g : gt;
                    uint32_t k = h - ((rr == 0) ? 0u : incl_rows[rr - 1]);
                    const int lx0 = row.x0 - g.gmin[0], lx1 = row.x1 - g.gmin[0];
                    int x = 0;
                    bool found = row.one != 0; // grid coordinates may be negative: no sentinel x
                    if (found)
                        x = g.gmin[0] + (lx0 & ~63) + select_bit(row.word, k);
                    else
                    {
                        const uint64_t rowbase =
                            ((uint64_t)(row.z - g.gmin[2]) * g.G[1] + (uint64_t)(row.y - g.gmin[1])) * g.wx;
                        for (int w = lx0 >> 6; w <= ((lx1 - 1) >> 6); ++w)
                        {
                            const int blo = (w == (lx0 >> 6)) ? (lx0 & 63) : 0;
                            const int bhi = (w == ((lx1 - 1) >> 6)) ? (((lx1 - 1) & 63) + 1) : 64;
                            const uint64_t mm = g.bits[rowbase + w] & row_mask(blo, bhi);
                            const uint32_t pc = (uint32_t)__popcll(mm);
                            if (k >= pc)
                            {
                                k -= pc;
                                continue;
                            }
                            x = g.gmin[0] + w * 64 + select_bit(mm, k);
                            found = true;
                            break;
                        }
                    }
                    if (found && face_pass(it, g.vs, x, row.y, row.z))
                    {
                        ++n_hit;
                        const int hm = QUEUED ? (int)it.ms : ms;
                        hit = FHit{it.face, (uint16_t)(x - g.gmin[0]), (uint16_t)(row.y - g.gmin[1]),
                                   (uint16_t)(row.z - g.gmin[2]), bt.M <= 65535 ? (uint16_t)hm : FHIT_NO_MESH};
                        has = true;
                    }
                }
                // one claim per wave for the wave's hits (no block barrier):
                // a wave's hits come from consecutive candidates of the tile's
                // rows, so the records made of them fall in few bricks per
                // generator block
                const uint64_t m = __ballot(has);
                if (m)
                {
                    const int lane = tid & 63;
                    unsigned long long base = 0;
                    if (lane == 0)
                    {
                        bool okk = false;
                        const uint32_t sh = (blockIdx.x * NWAVE + (uint32_t)(tid >> 6)) % NSHARD;
                        const unsigned long long b0 = shard_claim(hb, sh, (uint64_t)__popcll(m), okk);
                        base = okk ? b0 : ~0ull;
                    }
                    base = __shfl(base, 0, 64);
                    if (has && base != ~0ull)
                        hb.items[base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull))] = hit;
                }
            }
            __syncthreads();
        }
        __syncthreads();
    }
    (void)ctr;
    (void)n_hit;
}

// Walks the candidate voxels of bounding-box rows [0, n_rows) of one face
// (the active voxels of each row's slab range) and returns the number of
// faces-test hits; with WRITE the hits are also stored from `out` on.
template <bool WRITE>
__device__ __forceinline__ uint32_t face_walk(const FaceItem &it, const GridDev &g, uint64_t n_rows, FHit *out,
                                              uint16_t hm = FHIT_NO_MESH)
{
    uint32_t n = 0;
    const RowNarrow nr = face_row_pre(it, g);
    int ry = 0, rz = 0; // row rr = rz * ny + ry, stepped instead of divided
    for (uint64_t rr = 0; rr < n_rows; ++rr)
    {
        const int y = it.by0 + ry;
        const int z = it.bz0 + rz;
        if (++ry == it.ny)
        {
            ry = 0;
            ++rz;
        }
        int x0, x1;
        face_row_range(it, g, nr, y, z, x0, x1);
        const int lx0 = x0 - g.gmin[0], lx1 = x1 - g.gmin[0];
        if (lx1 <= lx0)
            continue;
        const uint64_t rowbase = ((uint64_t)(z - g.gmin[2]) * g.G[1] + (uint64_t)(y - g.gmin[1])) * g.wx;
        for (int w = lx0 >> 6; w <= ((lx1 - 1) >> 6); ++w)
        {
            const int blo = (w == (lx0 >> 6)) ? (lx0 & 63) : 0;
            const int bhi = (w == ((lx1 - 1) >> 6)) ? (((lx1 - 1) & 63) + 1) : 64;
            uint64_t mm = g.bits[rowbase + w] & row_mask(blo, bhi);
            while (mm)
            {
                const int b = __builtin_ctzll(mm);
                mm &= mm - 1ull;
                const int x = g.gmin[0] + w * 64 + b;
                if (face_pass(it, g.vs, x, y, z))
                {
                    if constexpr (WRITE)
                        out[n] = FHit{it.face, (uint16_t)(x - g.gmin[0]), (uint16_t)(y - g.gmin[1]),
                                      (uint16_t)(z - g.gmin[2]), hm};
                    ++n;
                }
            }
        }
    }
    return n;
}

// Hits of a face whose bounding box is at most 8 voxels long along x, with
// at most FACE_MASK_ROWS (y, z) rows, each row inside one bitset word, as a
// bit mask over the box (bit 8 r + (x - bx0) for row r).  The rows' words
// are loaded together and their active voxels gathered into one candidate
// mask, so a wave's lanes test their candidates in one loop (as many steps
// as the lane with the most candidates, not that maximum per row summed).
constexpr int FACE_MASK_ROWS = 8;
__device__ __forceinline__ bool face_mask_fits(const FaceItem &it, const GridDev &g, uint64_t R, int ext)
{
    const int sh = (it.bx0 - g.gmin[0]) & 63;
    return R > 0 && R <= (uint64_t)FACE_MASK_ROWS && ext <= 8 && sh + ext <= 64;
}

// (y, z) offsets of the box's rows r < 8 from its min corner, 4 bits each,
// row r in byte r (rows run y fastest)
__device__ __forceinline__ uint64_t face_mask_rows(const FaceItem &it)
{
    uint64_t t = 0ull;
    int dy = 0, dz = 0;
#pragma unroll
    for (int r = 0; r < FACE_MASK_ROWS; ++r)
    {
        t |= (uint64_t)(dy | (dz << 4)) << (8 * r);
        if (++dy == it.ny)
        {
            dy = 0;
            ++dz;
        }
    }
    return t;
}

__device__ __forceinline__ uint64_t face_mask_hits(const FaceItem &it, const GridDev &g, int R, int ext,
                                                   uint64_t ryz)
{
    const int lx0 = it.bx0 - g.gmin[0];
    const int w0 = lx0 >> 6, sh = lx0 & 63;
    uint64_t wv[FACE_MASK_ROWS];
#pragma unroll
    for (int r = 0; r < FACE_MASK_ROWS; ++r)
    {
        wv[r] = 0ull;
        if (r < R)
        {
            const int y = it.by0 + (int)((ryz >> (8 * r)) & 15u), z = it.bz0 + (int)((ryz >> (8 * r + 4)) & 15u);
            wv[r] = g.bits[((uint64_t)(z - g.gmin[2]) * g.G[1] + (uint64_t)(y - g.gmin[1])) * g.wx + w0];
        }
    }
    const uint64_t emask = (1ull << ext) - 1ull;
    // long rows are narrowed to the plane's slab (face_row_range); at most 8
    // voxels here, so only rows of 5 .. 8 voxels
    const RowNarrow nr = face_row_pre(it, g);
    const bool narrow = nr.on;
    uint64_t cand = 0ull;
#pragma unroll
    for (int r = 0; r < FACE_MASK_ROWS; ++r)
    {
        uint64_t act = (wv[r] >> sh) & emask;
        if (narrow && r < R && act)
        {
            int x0, x1;
            face_row_range(it, g, nr, it.by0 + (int)((ryz >> (8 * r)) & 15u),
                           it.bz0 + (int)((ryz >> (8 * r + 4)) & 15u), x0, x1);
            const int j0 = x0 - it.bx0, j1 = x1 - it.bx0;
            act &= ((1ull << j1) - 1ull) & ~((1ull << j0) - 1ull);
        }
        cand |= act << (8 * r);
    }
    uint64_t hits = 0ull;
    while (cand)
    {
        const int b = __builtin_ctzll(cand);
        cand &= cand - 1ull;
        const uint32_t yz = (uint32_t)(ryz >> (b & ~7)) & 0xffu;
        if (face_pass(it, g.vs, it.bx0 + (b & 7), it.by0 + (int)(yz & 15u), it.bz0 + (int)(yz >> 4)))
            hits |= 1ull << b;
    }
    return hits;
}

// The face test with one lane per face, for meshes whose faces cover few
// voxels (block-cooperative walking costs more than it balances there).  A
// block takes BT consecutive faces of one mesh.  A face of a small box (at
// most 8 x 8 voxels) is tested once into a hit mask; a larger one is walked twice, to
// count its hits and, after one claim per wave, to store them.  A face whose
// bounding box holds more than FACE_LANE_CAP voxels goes to the row-chunk
// queue, which the block-cooperative kernel consumes; if the queue is full
// the lane walks it.
constexpr uint32_t FACE_LANE_CAP = 192;
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4))) void k_face_lane(MeshDev m, BatchDev bt, uint64_t n_tiles, ShardBuf<FHit> hb,
                                                  ShardBuf<FaceChunk> fq)
{
    const int tid = threadIdx.x, lane = tid & 63;
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x)
    {
        const int ms = seg_find(bt.ftile_off, bt.M, tile);
        const GridDev &g = bt.seg[ms].g;
        const int64_t f0 = bt.face_off[ms];
        const int64_t loc = (int64_t)(tile - bt.ftile_off[ms]) * BT + tid;
        FaceItem it;
        uint64_t R = 0, hits = 0, ryz = 0;
        int ext = 1;
        bool masked = false;
        if (loc < bt.face_off[ms + 1] - f0)
        {
            R = face_item_setup(m, g, f0 + loc, it);
            ext = it.bx1 > it.bx0 ? it.bx1 - it.bx0 : 1;
            if (face_mask_fits(it, g, R, ext))
            {
                ryz = face_mask_rows(it);
                hits = face_mask_hits(it, g, (int)R, ext, ryz);
                masked = true;
            }
            else if (R * (uint64_t)ext > FACE_LANE_CAP && fq.items)
            {
                // chunks of rpc rows of about FACE_VOX_CAP voxels, all queued or none
                const uint64_t rpc = (uint64_t)ext < FACE_VOX_CAP ? FACE_VOX_CAP / (uint64_t)ext : 1u;
                const uint64_t nq = (R + rpc - 1) / rpc;
                uint64_t q0;
                if (shard_claim_fit(fq, it.face % NSHARD, nq, q0))
                {
                    for (uint64_t c = 0; c < nq; ++c)
                    {
                        const uint64_t r0 = rpc * c, r1 = r0 + rpc < R ? r0 + rpc : R;
                        fq.items[q0 + c] = FaceChunk{it.face, (uint32_t)r0, (uint32_t)r1, (uint32_t)ms};
                    }
                    R = 0;
                }
            }
        }
        const uint32_t cnt = masked ? (uint32_t)__popcll(hits) : (R ? face_walk<false>(it, g, R, nullptr) : 0u);
        // one claim per wave for the wave's hits (consecutive faces stay together)
        const uint32_t incl = wave_incl_scan(cnt);
        const uint32_t total = __shfl(incl, 63, 64);
        if (total == 0)
            continue;
        unsigned long long base = 0;
        if (lane == 63)
        {
            bool ok = false;
            const uint64_t b0 = shard_claim(hb, (uint32_t)((blockIdx.x * NWAVE + (tid >> 6)) % NSHARD), total, ok);
            base = ok ? (unsigned long long)b0 : ~0ull;
        }
        base = __shfl(base, 63, 64);
        if (!cnt || base == ~0ull)
            continue;
        FHit *out = hb.items + base + (incl - cnt);
        const uint16_t hm = bt.M <= 65535 ? (uint16_t)ms : FHIT_NO_MESH;
        if (masked)
        {
            uint32_t k = 0;
            while (hits)
            {
                const int b = __builtin_ctzll(hits);
                hits &= hits - 1ull;
                const uint32_t yz = (uint32_t)(ryz >> (b & ~7)) & 0xffu;
                const int x = it.bx0 + (b & 7), y = it.by0 + (int)(yz & 15u), z = it.bz0 + (int)(yz >> 4);
                out[k++] = FHit{it.face, (uint16_t)(x - g.gmin[0]), (uint16_t)(y - g.gmin[1]), (uint16_t)(z - g.gmin[2]),
                                hm};
            }
        }
        else
            face_walk<true>(it, g, R, out, hm);
    }
}

// ---------------------------------------------------------------------------
// boundary edges (FDG.cpp:724-756): per-min-vertex adjacency lists, sorted,
// singletons are boundary edges; ranks follow (min, max) ascending order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BT) void k_edge_count(MeshDev m, uint32_t *vcnt, uint32_t *rank, const uint32_t *voff,
                                                   uint32_t *nbr, int pass)
{
    const int64_t stride = (int64_t)gridDim.x * BT;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < 3 * m.F; i += stride)
    {
        const int64_t f = i / 3;
        const int k = (int)(i - 3 * f);
        const int32_t *fi = m.f + 3 * f;
        if (!face_ok(m, fi[0], fi[1], fi[2]))
            continue; // flagged by the prep kernel
        int a = fi[k];
        int b = fi[(k + 1) % 3];
        if (a > b)
        {
            const int t = a;
            a = b;
            b = t;
        }
        // pass 0: count the half-edges of each smaller vertex, keeping each
        // one's rank; pass 1: place the larger vertex at offset + rank
        if (pass == 0)
            rank[i] = atomicAdd(vcnt + a, 1u);
        else
            nbr[voff[a] + rank[i]] = (uint32_t)b;
    }
}

// Boundary edges by vertex bins: half-edges are binned by their smaller
// vertex, EBIN_VB consecutive vertices per bin, so that no half-edge takes a
// returning global atomic of its own.
//   k_ebin<false>  a block of faces counts its half-edges per bin in an LDS
//                  table and adds each bin's count once;
//   (scan of the bin counts: each bin's range of the half-edge array)
//   k_ebin<true>   the same blocks claim one range per bin and store their
//                  half-edges (smaller, larger vertex) there;
//   k_ebin_fold    one block per bin buckets the bin's half-edges by vertex
//                  in LDS (at most EDGE_SLOTS per vertex), sorts each vertex's
//                  list in registers and stores the singletons, vertex by
//                  vertex, over the bin's own half-edges;
//   (scan of the bins' singleton counts: the boundary-edge total)
//   k_ebin_gather  concatenates the bins' singletons: the boundary edges in
//                  (min, max) order.
constexpr int EBIN_LOG_VB = 9;
static_assert((1 << EBIN_LOG_VB) == EBIN_VB, "bin width");
static_assert(EBIN_VB == 2 * BT, "a bin's vertices are sorted two per thread");
constexpr int EBIN_FPT = 2;       // faces per thread of the binning kernels
constexpr int EBIN_LOG_HASH = 11; // LDS table slots of a binning block
constexpr int EBIN_HASH = 1 << EBIN_LOG_HASH;
static_assert(EBIN_HASH >= 3 * BT * EBIN_FPT, "every bin of a block's half-edges finds a table slot");

template <bool PLACE>
__global__ __launch_bounds__(BT) void k_ebin(MeshDev m, uint32_t *bin_cnt, const uint32_t *bin_off, uint2 *he)
{
    constexpr int K = 3 * EBIN_FPT;
    __shared__ uint32_t key[EBIN_HASH];
    __shared__ uint32_t cnt[EBIN_HASH]; // half-edges per bin; once claimed (PLACE), the block's first position
    for (int i = threadIdx.x; i < EBIN_HASH; i += BT)
    {
        key[i] = 0xffffffffu;
        cnt[i] = 0;
    }
    __syncthreads();
    const int64_t fbase = (int64_t)blockIdx.x * (BT * EBIN_FPT) + threadIdx.x;
    int32_t ea[K], eb[K];
    bool ok[K];
#pragma unroll
    for (int j = 0; j < EBIN_FPT; ++j)
    {
        const int64_t f = fbase + (int64_t)j * BT;
        int32_t v[3] = {0, 0, 0};
        bool live = f < m.F;
        if (live)
        {
            const int32_t *fi = m.f + 3 * f;
            v[0] = fi[0];
            v[1] = fi[1];
            v[2] = fi[2];
            live = face_ok(m, v[0], v[1], v[2]); // else flagged by the prep kernel
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
        {
            const int32_t p = v[k], q = v[k == 2 ? 0 : k + 1];
            ea[3 * j + k] = p < q ? p : q;
            eb[3 * j + k] = p < q ? q : p;
            ok[3 * j + k] = live;
        }
    }
    // a thread's half-edges of one bin (a face's mostly share one) take one
    // table lookup and one count update together
    uint32_t slot[K], rank[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        slot[k] = 0;
        rank[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
    {
        if (!ok[k])
            continue;
        const uint32_t bin = (uint32_t)ea[k] >> EBIN_LOG_VB;
        bool led = false;
#pragma unroll
        for (int j = 0; j < k; ++j)
            led = led || (ok[j] && ((uint32_t)ea[j] >> EBIN_LOG_VB) == bin);
        if (led)
            continue;
        uint32_t n = 1;
#pragma unroll
        for (int j = k + 1; j < K; ++j)
            n += (ok[j] && ((uint32_t)ea[j] >> EBIN_LOG_VB) == bin) ? 1u : 0u;
        uint32_t h = (bin * 2654435761u) >> (32 - EBIN_LOG_HASH);
        while (true)
        {
            const uint32_t prev = atomicCAS(&key[h], 0xffffffffu, bin);
            if (prev == 0xffffffffu || prev == bin)
                break;
            h = (h + 1u) & (uint32_t)(EBIN_HASH - 1);
        }
        const uint32_t r0 = atomicAdd(&cnt[h], n);
        slot[k] = h;
        rank[k] = r0;
        uint32_t i = 1;
#pragma unroll
        for (int j = k + 1; j < K; ++j)
            if (ok[j] && ((uint32_t)ea[j] >> EBIN_LOG_VB) == bin)
            {
                slot[j] = h;
                rank[j] = r0 + i;
                ++i;
            }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < EBIN_HASH; i += BT)
    {
        const uint32_t n = cnt[i];
        if (!n)
            continue;
        const uint32_t bin = key[i];
        if constexpr (!PLACE)
            atomicAdd(bin_cnt + bin, n);
        else
        {
            // each block lowers the bin's count by its share: its range ends
            // where the count stood (the count pass saw the same half-edges)
            const uint32_t old = atomicSub(bin_cnt + bin, n);
            cnt[i] = old >= n ? bin_off[bin] + (old - n) : 0xffffffffu;
        }
    }
    if constexpr (PLACE)
    {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            if (!ok[k])
                continue;
            const uint32_t b0 = cnt[slot[k]];
            if (b0 != 0xffffffffu)
                he[(uint64_t)b0 + rank[k]] = make_uint2((uint32_t)ea[k], (uint32_t)eb[k]);
        }
    }
}

// ascending sort of 16 values in registers (bitonic network)
__device__ __forceinline__ void sort16(uint32_t (&x)[16])
{
#pragma unroll
    for (int size = 2; size <= 16; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int i = 0; i < 16; ++i)
            {
                const int j = i ^ stride;
                if (j > i)
                {
                    const bool up = (i & size) == 0;
                    const uint32_t lo = x[i] < x[j] ? x[i] : x[j];
                    const uint32_t hi = x[i] < x[j] ? x[j] : x[i];
                    x[i] = up ? lo : hi;
                    x[j] = up ? hi : lo;
                }
            }
}

// one block per bin: the bin's half-edges bucketed by their smaller vertex
// in LDS, each vertex's list sorted in registers; the singletons (boundary
// edges) are stored in vertex order from the bin's first half-edge on, and
// their count in bin_bcnt.  A vertex with more than EDGE_SLOTS half-edges
// sets `over` (the conversion then re-runs the stage on per-vertex lists of
// any length).
__global__ __launch_bounds__(BT) void k_ebin_fold(const uint32_t *bin_off, uint2 *he, uint32_t *bin_bcnt,
                                                  unsigned long long *over)
{
    __shared__ uint32_t slots[EDGE_SLOTS * EBIN_VB]; // slot r of vertex l at r * EBIN_VB + l
    __shared__ uint32_t vc[EBIN_VB];
    __shared__ uint32_t sh[NWAVE];
    const int tid = threadIdx.x;
    const uint32_t bin = blockIdx.x;
    const uint32_t s = bin_off[bin], e = bin_off[bin + 1];
    const uint32_t v0 = bin << EBIN_LOG_VB;
    vc[tid] = 0;
    vc[tid + BT] = 0;
    __syncthreads();
    for (uint32_t i = s + tid; i < e; i += BT)
    {
        const uint2 h = he[i];
        const uint32_t l = h.x - v0; // < EBIN_VB: the half-edge was binned by h.x
        if (l < (uint32_t)EBIN_VB)
        {
            const uint32_t r = atomicAdd(&vc[l], 1u);
            if (r < (uint32_t)EDGE_SLOTS)
                slots[r * EBIN_VB + l] = h.y;
        }
    }
    __syncthreads(); // every half-edge of the bin is read: its range may be overwritten
    uint32_t base = 0;
#pragma unroll 1
    for (int half = 0; half < 2; ++half)
    {
        const int l = half * BT + tid;
        uint32_t c = vc[l];
        if (c > (uint32_t)EDGE_SLOTS)
        {
            atomicOr(over, 1ull);
            c = 0;
        }
        uint32_t x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i)
            x[i] = (uint32_t)i < c ? slots[i * EBIN_VB + l] : 0xffffffffu; // unused slots sort last
        if (c > 1)
            sort16(x);
        bool single[16];
        uint32_t n = 0;
#pragma unroll
        for (int i = 0; i < 16; ++i)
        {
            single[i] = (uint32_t)i < c && (i == 0 || x[i] != x[i - 1]) && ((uint32_t)(i + 1) >= c || x[i] != x[i + 1]);
            n += single[i] ? 1u : 0u;
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan(n, sh, tot);
        uint32_t o = s + base + ex;
#pragma unroll
        for (int i = 0; i < 16; ++i)
            if (single[i])
                he[o++] = make_uint2(v0 + (uint32_t)l, x[i]);
        base += tot;
    }
    if (tid == 0)
        bin_bcnt[bin] = base;
}

// one wave per bin: the bin's singletons to their place in the boundary-edge list
__global__ __launch_bounds__(BT) void k_ebin_gather(uint64_t n_bins, const uint32_t *bin_off, const uint32_t *bin_boff,
                                                    const uint2 *he, int2 *bedge_v)
{
    const int lane = threadIdx.x & 63;
    for (uint64_t bin = (uint64_t)blockIdx.x * NWAVE + (threadIdx.x >> 6); bin < n_bins; bin += (uint64_t)gridDim.x * NWAVE)
    {
        const uint32_t s = bin_off[bin], o = bin_boff[bin], n = bin_boff[bin + 1] - o;
        for (uint32_t i = lane; i < n; i += 64)
        {
            const uint2 h = he[s + i];
            bedge_v[o + i] = make_int2((int)h.x, (int)h.y);
        }
    }
}

// sort each vertex's neighbour list, count singletons (pass 0) or write the
// boundary edges in rank order (pass 1)
__global__ __launch_bounds__(BT) void k_edge_singletons(int64_t V, const uint32_t *voff, uint32_t *nbr, uint32_t *bcnt,
                                                         const uint32_t *boff, int2 *bedge_v, int pass)
{
    const int64_t stride = (int64_t)gridDim.x * BT;
    for (int64_t v = (int64_t)blockIdx.x * BT + threadIdx.x; v < V; v += stride)
    {
        const uint32_t s = voff[v], e = voff[v + 1];
        if (pass == 0)
        {
            for (uint32_t i = s + 1; i < e; ++i)
            {
                const uint32_t x = nbr[i];
                uint32_t j = i;
                while (j > s && nbr[j - 1] > x)
                {
                    nbr[j] = nbr[j - 1];
                    --j;
                }
                nbr[j] = x;
            }
        }
        uint32_t n = 0;
        uint32_t i = s;
        while (i < e)
        {
            uint32_t j = i + 1;
            while (j < e && nbr[j] == nbr[i])
                ++j;
            if (j - i == 1)
            {
                if (pass == 1)
                    bedge_v[boff[v] + n] = make_int2((int)v, (int)nbr[i]);
                ++n;
            }
            i = j;
        }
        if (pass == 0)
            bcnt[v] = n;
    }
}

// boundary quadric (FDG.cpp:433-453) and DDA set-up
__global__ __launch_bounds__(BT) void k_bedge_prep(const float *verts, int64_t E, const int2 *bedge_v, BEdge *be)
{
    const int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    if (i >= E)
        return;
    const int2 ab = bedge_v[i];
    const float *v0 = verts + 3 * (int64_t)ab.x;
    const float *v1 = verts + 3 * (int64_t)ab.y;
    BEdge out;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        out.v0[a] = v0[a];
    double dir[3] = {(double)(v1[0] - v0[0]), (double)(v1[1] - v0[1]), (double)(v1[2] - v0[2])};
    // Vector3d squaredNorm with SSE2 packets: (d0^2 + d1^2) + d2^2
    const double sq = (dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2];
    const double len = sqrt(sq);
    out.valid = !(len < 1e-6);
    if (sq > 0.0)
    {
        const double rn = sqrt(sq);
        dir[0] = dir[0] / rn;
        dir[1] = dir[1] / rn;
        dir[2] = dir[2] / rn;
    }
    float A[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            A[a][b] = (a == b ? 1.0f : 0.0f) - (float)(dir[a] * dir[b]);
    const float p0 = v0[0], p1 = v0[1], p2 = v0[2];
    float bb[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        bb[a] = -(A[a][0] * p0 + (A[a][1] * p1 + A[a][2] * p2));
    float t[3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
        t[b] = p0 * A[0][b] + (p1 * A[1][b] + p2 * A[2][b]);
    const float c = t[0] * p0 + (t[1] * p1 + t[2] * p2);
    out.q[qidx(0, 0)] = A[0][0];
    out.q[qidx(0, 1)] = A[0][1];
    out.q[qidx(0, 2)] = A[0][2];
    out.q[qidx(1, 1)] = A[1][1];
    out.q[qidx(1, 2)] = A[1][2];
    out.q[qidx(2, 2)] = A[2][2];
    out.q[qidx(0, 3)] = bb[0];
    out.q[qidx(1, 3)] = bb[1];
    out.q[qidx(2, 3)] = bb[2];
    out.q[qidx(3, 3)] = c;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        out.dir[a] = dir[a];
    out.len = len;
    be[i] = out;
}

// Boundary QEF terms (FDG.cpp:455-522): one thread per boundary segment walks
// the voxels the segment passes with the reference's DDA (FP64 steps), and
// every occupied voxel on the way gets a phase-2 record (key = the edge's
// rank).  The block emits its threads' records in rounds of up to 4 each
// (bucketed by brick, like the other record generators) until every segment
// of the block is walked; the hits are only counted (SB_BHITS counters:
// statistics and record-buffer estimates).
__global__ __launch_bounds__(BT) void k_bdry(int64_t E, const BEdge *be, const int2 *bedge_v, BatchDev bt,
                                             BucketArgs ba, ShardBuf<BHit> hcount, unsigned long long *counters)
{
    __shared__ BucketLds<11> L;
    const int tid = threadIdx.x;
    unsigned err = 0;
    uint32_t n_hit = 0;
    for (int64_t base = (int64_t)blockIdx.x * BT; base < E; base += (int64_t)gridDim.x * BT)
    {
        const int64_t i = base + tid;
        bool alive = false;
        int ms = 0;
        int cur[3] = {0, 0, 0}, step[3] = {0, 0, 0};
        double tMax[3] = {0.0, 0.0, 0.0}, tDelta[3] = {0.0, 0.0, 0.0};
        double len = 0.0;
        int64_t steps = 0, max_steps = 0;
        if (i < E && be[i].valid)
        {
            const BEdge &e = be[i];
            ms = mesh_of_vert(bt, (int64_t)bedge_v[i].x); // the edge's mesh
            const GridDev &g = bt.seg[ms].g;
            double cap = 16.0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
            {
                cur[a] = (int)floorf(div_vs(e.v0[a], g, a));
                step[a] = e.dir[a] > 0.0 ? 1 : -1;
                if (e.dir[a] == 0.0)
                {
                    tMax[a] = __builtin_huge_val();
                    tDelta[a] = __builtin_huge_val();
                }
                else
                {
                    const float border = g.vs[a] * (float)(cur[a] + (step[a] > 0 ? 1 : 0));
                    tMax[a] = (double)(border - e.v0[a]) / e.dir[a];
                    tDelta[a] = (double)g.vs[a] / fabs(e.dir[a]);
                    cap += ceil(e.len * fabs(e.dir[a]) / (double)g.vs[a]) + 2.0;
                }
            }
            len = e.len;
            max_steps = (int64_t)(cap < 4.0e9 ? cap : 4.0e9);
            alive = true;
        }
        while (__syncthreads_or(alive ? 1 : 0))
        {
            uint32_t bk[4] = {0, 0, 0, 0};
            Rec rc[4];
            bool ok[4] = {false, false, false, false};
            int nh = 0;
            if (alive)
            {
                const MeshSeg &sg = bt.seg[ms];
                const GridDev &g = sg.g;
                // one voxel per iteration: its record when it is occupied, then the step to the next
                while (alive && nh < 4)
                {
                    if (cur[0] >= g.gmin[0] && cur[0] < g.gmax[0] && cur[1] >= g.gmin[1] && cur[1] < g.gmax[1] &&
                        cur[2] >= g.gmin[2] && cur[2] < g.gmax[2])
                    {
                        int b;
                        const uint64_t w = word_of(g, cur[0], cur[1], cur[2], b);
                        if ((g.bits[w] >> b) & 1ull)
                        {
                            uint32_t lvox, brick;
                            brick_of(sg, cur[0], cur[1], cur[2], brick, lvox);
#pragma unroll
                            for (int k = 0; k < 4; ++k)
                                if (k == nh)
                                {
                                    bk[k] = brick + (uint32_t)sg.brick0;
                                    rc[k] = Rec{(2ull << KEY_PHASE_SHIFT) | (uint64_t)i, 0.0f, lvox};
                                    ok[k] = true;
                                }
                            ++nh;
                        }
                    }
                    int axis;
                    if (tMax[0] < tMax[1])
                        axis = (tMax[0] < tMax[2]) ? 0 : 2;
                    else
                        axis = (tMax[1] < tMax[2]) ? 1 : 2;
                    const double tm = axis == 0 ? tMax[0] : (axis == 1 ? tMax[1] : tMax[2]);
                    if (tm > len)
                        alive = false;
                    else if (++steps > max_steps)
                    {
                        err |= ERR_DDA;
                        alive = false;
                    }
                    else if (axis == 0)
                    {
                        cur[0] += step[0];
                        tMax[0] += tDelta[0];
                    }
                    else if (axis == 1)
                    {
                        cur[1] += step[1];
                        tMax[1] += tDelta[1];
                    }
                    else
                    {
                        cur[2] += step[2];
                        tMax[2] += tDelta[2];
                    }
                }
            }
            n_hit += (uint32_t)nh;
            const uint32_t ent[4] = {1u, 1u, 1u, 1u};
            bucket_emit<4, 11>(bk, rc, ok, ent, ba, L);
        }
    }
    // the block's hits onto its shard's counter (no items are stored)
    uint32_t tot;
    __shared__ uint32_t sh[NWAVE];
    block_excl_scan(n_hit, sh, tot);
    if (tid == 0 && tot)
        atomicAdd(hcount.cnt + (uint64_t)(blockIdx.x % NSHARD) * SHARD_STRIDE, (unsigned long long)tot);
    if (counters && err)
        atomicOr(counters + CTR_ERRORS, (unsigned long long)err);
}

// ---------------------------------------------------------------------------
// record generation from the recorded crossings and boundary hits
// ---------------------------------------------------------------------------
// phase-0 records: one per (crossing, brick of its 4 voxels), the crossing's
// key without (dx, dy), its z and its voxel index zi along the axis (in the
// lvox field); the fold expands it into its voxels inside the brick
// (crossing_voxels), whose keys (face*3+axis, y, x, dx, dy) order them
__global__ __launch_bounds__(BT) void k_gen_events(ShardBuf<Event> ev, BucketArgs ba)
{
    __shared__ BucketLds<11> L;
    __shared__ QueueMap qm;
    const uint64_t n = queue_prefix(ev, qm.pre, qm.sh);
    for (uint64_t base = (uint64_t)blockIdx.x * BT; base < n; base += (uint64_t)gridDim.x * BT)
    {
        const uint64_t i = base + threadIdx.x;
        uint32_t bk[4] = {0, 0, 0, 0};
        uint32_t ent[4] = {0, 0, 0, 0};
        Rec rc[4];
        bool ok[4] = {false, false, false, false};
        if (i < n)
        {
            const Event e = ev.items[queue_slot(ev, qm.pre, i)];
            const uint64_t face = (e.key >> KEY_FACE_SHIFT) & KEY_FACE_MASK;
            const MeshSeg &sg = ba.b.seg[mesh_of_face(ba.b, (int64_t)face)];
            const GridDev &g = sg.g;
            const int ax2 = (int)((e.key >> KEY_AX_SHIFT) & 3ull);
            const int ax0 = axis_next(ax2);
            const int ax1 = axis_prev(ax2);
            const int y_idx =
                (int)((e.key >> KEY_Y_SHIFT) & KEY_IDX_MASK) + sel3i(ax1, g.gmin[0], g.gmin[1], g.gmin[2]);
            const int x_idx =
                (int)((e.key >> KEY_X_SHIFT) & KEY_IDX_MASK) + sel3i(ax0, g.gmin[0], g.gmin[1], g.gmin[2]);
            int c[4][3];
            event_voxels(ax0, ax1, x_idx, y_idx, e.zi, c);
#pragma unroll
            for (int k = 0; k < 4; ++k)
            {
                uint32_t lvox;
                brick_of(sg, c[k][0], c[k][1], c[k][2], bk[k], lvox);
                bk[k] += (uint32_t)sg.brick0;
                rc[k] = Rec{e.key, e.z, (uint32_t)e.zi};
            }
            // the first of the 4 voxels in each distinct brick stands for the brick's voxels
#pragma unroll
            for (int k = 0; k < 4; ++k)
            {
                bool first = true;
#pragma unroll
                for (int j = 0; j < k; ++j)
                    first = first && bk[j] != bk[k];
                ok[k] = first;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    ent[k] += bk[j] == bk[k] ? 1u : 0u;
            }
        }
        bucket_emit<4, 11>(bk, rc, ok, ent, ba, L);
    }
}

// phase-1 records from the face-test hits, key = face index
__global__ __launch_bounds__(BT) void k_gen_fhits(ShardBuf<FHit> h, BucketArgs ba)
{
    __shared__ BucketLds<11> L;
    __shared__ QueueMap qm;
    const uint64_t n = queue_prefix(h, qm.pre, qm.sh);
    for (uint64_t base = (uint64_t)blockIdx.x * (4 * BT); base < n; base += (uint64_t)gridDim.x * (4 * BT))
    {
        uint32_t bk[4] = {0, 0, 0, 0};
        Rec rc[4];
        bool ok[4] = {false, false, false, false};
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const uint64_t i = base + 4ull * threadIdx.x + (uint64_t)k; // 4 consecutive hits: mostly one face's
            if (i < n)
            {
                const FHit x = h.items[queue_slot(h, qm.pre, i)];
                const MeshSeg &sg = ba.b.seg[x.ms != FHIT_NO_MESH ? (int)x.ms : mesh_of_face(ba.b, (int64_t)x.face)];
                const GridDev &g = sg.g;
                uint32_t lvox;
                brick_of(sg, g.gmin[0] + (int)x.lx, g.gmin[1] + (int)x.ly, g.gmin[2] + (int)x.lz, bk[k], lvox);
                bk[k] += (uint32_t)sg.brick0;
                rc[k] = Rec{(1ull << KEY_PHASE_SHIFT) | (uint64_t)x.face, 0.0f, lvox};
                ok[k] = true;
            }
        }
        const uint32_t ent[4] = {1u, 1u, 1u, 1u};
        bucket_emit<4, 11>(bk, rc, ok, ent, ba, L);
    }
}

// sum and largest shard of each sharded buffer's requests
__global__ __launch_bounds__(BT) void k_shard_reduce(const unsigned long long *shard_ctr, unsigned long long *ctr)
{
    __shared__ unsigned long long sh[NWAVE];
    const int b = blockIdx.x; // one block per buffer
    const unsigned long long v = shard_ctr[((uint64_t)b * NSHARD + threadIdx.x) * SHARD_STRIDE];
    unsigned long long tot;
    block_excl_scan(v, sh, tot);
    // max via the wave / LDS reduction of the scan's building blocks
    unsigned long long m = v;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
    {
        const unsigned long long o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long mx = sh[0];
        for (int w = 1; w < NWAVE; ++w)
            mx = sh[w] > mx ? sh[w] : mx;
        ctr[CTR_SUM + b] = tot;
        ctr[CTR_MAX + b] = mx;
    }
}

// chunks grouped by brick as (first record, count): position = first chunk of
// the brick + rank.  PLACE_SUB blocks walk each shard's chunks (no lookup of
// a dense index's shard), PLACE_UNROLL chunks in flight per thread.
constexpr int PLACE_SUB = 32;
constexpr int PLACE_UNROLL = 8;
__global__ __launch_bounds__(BT) void k_place_chunks(ShardBuf<Chunk> chunks, const uint32_t *choff, uint2 *gchunks,
                                                     const unsigned long long *over)
{
    if (*over)
        return; // dropped chunks: the host sizes the buffers again and re-runs
    const uint32_t sh = blockIdx.x % NSHARD, sub = blockIdx.x / NSHARD;
    const unsigned long long c = chunks.cnt[(uint64_t)sh * SHARD_STRIDE];
    const uint64_t n = c <= chunks.per ? (uint64_t)c : 0ull;
    const Chunk *src = chunks.items + (uint64_t)sh * chunks.per;
    const uint64_t step = (uint64_t)(gridDim.x / NSHARD) * BT;
    for (uint64_t i0 = (uint64_t)sub * BT + threadIdx.x; i0 < n; i0 += step * PLACE_UNROLL)
    {
        Chunk ch[PLACE_UNROLL];
        uint32_t off[PLACE_UNROLL];
#pragma unroll
        for (int u = 0; u < PLACE_UNROLL; ++u)
            if (i0 + u * step < n)
                ch[u] = src[i0 + u * step];
#pragma unroll
        for (int u = 0; u < PLACE_UNROLL; ++u)
            if (i0 + u * step < n)
                off[u] = choff[ch[u].brick];
#pragma unroll
        for (int u = 0; u < PLACE_UNROLL; ++u)
            if (i0 + u * step < n)
                gchunks[off[u] + ch[u].rank] = make_uint2(ch[u].start, ch[u].count);
    }
}

// fold tiers by voxel entries per brick (held in LDS by the tier's kernel)
constexpr int FOLD_TINY_CAP = 384;  // first tier: one wave per brick (10 KB of LDS, 16 per CU)
constexpr int FOLD_TINY_NT = 64;
constexpr int FOLD_TINY_WAVES = 4;  // waves per SIMD the first tier is compiled for
constexpr int FOLD_SMALL_CAP = 768; // second tier: blocks of FOLD_SMALL_NT threads (15 KB of LDS)
constexpr int FOLD_SMALL_NT = 128;
constexpr int FOLD_SMALL = 1536; // third tier (25 KB of LDS, 5 blocks per CU)
constexpr int FOLD_MID = 3584;   // fourth tier (3 blocks per CU); larger bricks: k_fold_large

// bricks that received records, each appended to the list of the fold tier
// its record count selects: each block takes ACT_PER consecutive bricks per
// thread and appends them with one atomic per tier
constexpr int ACT_PER = 8;
__global__ __launch_bounds__(BT) void k_active_bricks(const unsigned long long *cnt, uint64_t n_bricks, TierLists t,
                                                      const unsigned long long *over)
{
    if (*over)
        return;
    __shared__ uint32_t sh[NWAVE];
    __shared__ unsigned long long sbase[N_FOLD_TIERS];
    const uint64_t span = (uint64_t)BT * ACT_PER;
    for (uint64_t b0 = (uint64_t)blockIdx.x * span; b0 < n_bricks; b0 += (uint64_t)gridDim.x * span)
    {
        const uint64_t first = b0 + (uint64_t)threadIdx.x * ACT_PER;
        uint32_t mine[N_FOLD_TIERS] = {0u, 0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < ACT_PER; ++e)
        {
            const unsigned long long c = first + e < n_bricks ? cnt[first + e] : 0ull;
            if ((uint32_t)c == 0u)
                continue;
            const uint32_t nr = (uint32_t)(c >> 32);
            const int tier = nr <= (uint32_t)FOLD_TINY_CAP    ? 0
                             : nr <= (uint32_t)FOLD_SMALL_CAP ? 1
                             : nr <= (uint32_t)FOLD_SMALL     ? 2
                             : nr <= (uint32_t)FOLD_MID       ? 3
                                                              : 4;
            mine[tier] |= 1u << e;
        }
#pragma unroll
        for (int tier = 0; tier < N_FOLD_TIERS; ++tier)
        {
            uint32_t tot;
            const uint32_t ex = block_excl_scan((uint32_t)__popc(mine[tier]), sh, tot);
            if (threadIdx.x == 0)
                sbase[tier] = tot ? atomicAdd(t.count[tier], (unsigned long long)tot) : 0ull;
            __syncthreads();
            uint64_t pos = sbase[tier] + ex;
#pragma unroll
            for (int e = 0; e < ACT_PER; ++e)
                if (mine[tier] & (1u << e))
                    t.list[tier][pos++] = (uint32_t)(first + e);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// device-wide exclusive scan (3 kernels): reduce tiles, scan tile sums, down-sweep
// ---------------------------------------------------------------------------
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BT * SCAN_ITEMS;

template <class Load>
__global__ __launch_bounds__(BT) void k_scan_reduce(Load ld, uint64_t n, uint64_t *tile_sums)
{
    __shared__ uint64_t sh[NWAVE];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
    {
        const uint64_t i = base + (uint64_t)k * BT + threadIdx.x;
        if (i < n)
            s += ld(i);
    }
    const uint64_t tot = block_reduce_sum(s, sh);
    if (threadIdx.x == 0)
        tile_sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(BT) void k_scan_tiles(uint64_t *tile_sums, uint64_t nt, unsigned long long *total_out)
{
    __shared__ uint64_t sh[NWAVE];
    uint64_t carry = 0;
    for (uint64_t b = 0; b < nt; b += BT)
    {
        const uint64_t i = b + threadIdx.x;
        const uint64_t v = (i < nt) ? tile_sums[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_excl_scan(v, sh, tot);
        if (i < nt)
            tile_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0)
        *total_out = carry;
}

template <class Load>
__global__ __launch_bounds__(BT) void k_scan_down(Load ld, uint64_t n, const uint64_t *tile_sums, uint32_t *out)
{
    __shared__ uint64_t sh[NWAVE];
    // each thread owns SCAN_ITEMS consecutive elements
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
    {
        const uint64_t i = base + k;
        v[k] = (i < n) ? ld(i) : 0u;
        s += v[k];
    }
    uint64_t tot;
    uint64_t run = tile_sums[blockIdx.x] + block_excl_scan(s, sh, tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
    {
        const uint64_t i = base + k;
        if (i < n)
            out[i] = (uint32_t)run;
        run += v[k];
    }
}

// ---------------------------------------------------------------------------
// voxel coordinates from the bitset (ascending x + Gx*(y + Gy*z))
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BT) void k_emit_coords(BatchDev bt, uint64_t n_words, int32_t *coords)
{
    // Each wave takes 64 consecutive words of the shared numbering (their
    // voxel ids are consecutive too, across meshes) and its lanes write the
    // words' voxels at consecutive ids: contiguous stores across the wave.
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * BT + (uint64_t)threadIdx.x) >> 6;
    const uint64_t n_waves = ((uint64_t)gridDim.x * BT) >> 6;
    for (uint64_t gw0 = wave * 64ull; gw0 < n_words; gw0 += n_waves * 64ull)
    {
        const uint64_t gw = gw0 + (uint64_t)lane;
        uint64_t m = 0;
        uint32_t id = 0;
        int xb = 0, y = 0, z = 0;
        if (gw < n_words)
        {
            const int ms = mesh_of_word(bt, gw);
            const GridDev &g = bt.seg[ms].g;
            const uint64_t w = gw - bt.word_off[ms]; // word of the mesh
            m = g.bits[w];
            if (m)
            {
                id = g.woff[w];
                // a mesh has fewer than 2^32 bitset words: 32-bit division
                const uint32_t w32 = (uint32_t)w, row = w32 / (uint32_t)g.wx;
                xb = g.gmin[0] + 64 * (int)(w32 - row * (uint32_t)g.wx);
                y = (int)(row % (uint32_t)g.G[1]) + g.gmin[1];
                z = (int)(row / (uint32_t)g.G[1]) + g.gmin[2];
            }
        }
        const uint32_t pc = (uint32_t)__popcll(m);
        const uint32_t incl = wave_incl_scan(pc);
        const uint32_t total = __shfl(incl, 63, 64);
        if (total == 0)
            continue; // wave-uniform
        const uint64_t nz = __ballot(pc != 0u);
        const uint32_t base = __shfl(id, __ffsll((long long)nz) - 1, 64); // id of the wave's first voxel
        const uint32_t excl = incl - pc;
        // every lane takes part in every round (the shuffles read all lanes)
        for (uint32_t k0 = 0; k0 < total; k0 += 64u)
        {
            const uint32_t k = k0 + (uint32_t)lane;
            // the word of voxel k: lane j = number of lanes with incl <= k
            int j = 0;
#pragma unroll
            for (int step = 32; step >= 1; step >>= 1)
            {
                const uint32_t v = __shfl(incl, j + step - 1, 64);
                j += (v <= k) ? step : 0;
            }
            j = j < 63 ? j : 63;
            const uint64_t mj = __shfl(m, j, 64);
            const uint32_t ej = __shfl(excl, j, 64);
            const int xj = __shfl(xb, j, 64), yj = __shfl(y, j, 64), zj = __shfl(z, j, 64);
            if (k < total)
            {
                const int bit = select_bit(mj, k - ej);
                int32_t *o = coords + 3 * (uint64_t)(base + k);
                o[0] = xj + bit;
                o[1] = yj;
                o[2] = zj;
            }
        }
    }
}

// first voxel id of each mesh (the word offset at its first word)
__global__ __launch_bounds__(BT) void k_seg_voxels(BatchDev bt, const uint32_t *woff, uint64_t n_words,
                                                   const unsigned long long *n_voxels, uint64_t *vox_off)
{
    const int m = (int)(blockIdx.x * BT + threadIdx.x);
    if (m > bt.M)
        return;
    const uint64_t w = bt.word_off[m];
    vox_off[m] = w < n_words ? (uint64_t)woff[w] : (uint64_t)*n_voxels;
}

// Running replay of one voxel's contributions in the reference's serial order
// (FDG.cpp:254-271 intersection, 409 face, 521 boundary), then regularization
// and solve (FDG.cpp:780-982).
struct FoldState
{
    float q[10];
    float m0, m1, m2, cnt;
    int fl;

    __device__ __forceinline__ void init(int, int, int, const FoldArgs &)
    {
#pragma unroll
        for (int k = 0; k < 10; ++k)
            q[k] = 0.0f;
        m0 = m1 = m2 = cnt = 0.0f;
        fl = 0;
    }

    // face whose plane a record uses (phases 0 and 1), NO_PLANE for boundary records
    static constexpr uint32_t NO_PLANE = 0xffffffffu;
    __device__ __forceinline__ static uint32_t plane_of(uint64_t key)
    {
        const int phase = (int)(key >> KEY_PHASE_SHIFT);
        if (phase == 0)
            return (uint32_t)((key >> KEY_FACE_SHIFT) & KEY_FACE_MASK);
        if (phase == 1)
            return (uint32_t)(key & 0xffffffffull);
        return NO_PLANE;
    }

    __device__ __forceinline__ void add(uint64_t key, float z, const FoldArgs &a)
    {
        const uint32_t f = plane_of(key);
        add_pl(key, z, f != NO_PLANE ? a.planes[f] : make_float4(0.0f, 0.0f, 0.0f, 0.0f), a);
    }

    // Start values from the voxel's first record (records are replayed in
    // key order, so a voxel's intersections come first): the reference stores
    // the first intersection's point and quadric as they are, which is -0 + x
    // (exact for every x, signed zeros included); a voxel without
    // intersections adds its face and boundary terms to +0.
    __device__ __forceinline__ void begin(uint64_t first_key)
    {
        const float q0 = (first_key >> KEY_PHASE_SHIFT) == 0 ? -0.0f : 0.0f;
#pragma unroll
        for (int k = 0; k < 10; ++k)
            q[k] = q0;
        m0 = m1 = m2 = -0.0f;
    }

    // one record; pl4 = the plane of its face (unused for boundary records).
    // The quadric term and its addition are shared by all phases, so a wave
    // whose lanes are at records of different phases runs them once.
    __device__ __forceinline__ void add_pl(uint64_t key, float z, const float4 pl4, const FoldArgs &a)
    {
        const int phase = (int)(key >> KEY_PHASE_SHIFT);
        float qf[10];
        if (phase == 2)
        {
            const uint64_t e = key & ((1ull << KEY_PHASE_SHIFT) - 1ull);
            const float *qe = a.bedges[e].q;
#pragma unroll
            for (int k = 0; k < 10; ++k)
                qf[k] = a.bw * qe[k];
        }
        else
        {
            const float pl[4] = {pl4.x, pl4.y, pl4.z, pl4.w};
            plane_quadric(pl, qf);
        }
#pragma unroll
        for (int k = 0; k < 10; ++k)
            q[k] = q[k] + qf[k];
        if (phase == 0)
        {
            const int ax2 = (int)((key >> KEY_AX_SHIFT) & 3ull);
            const int ax0 = axis_next(ax2);
            const int ax1 = axis_prev(ax2);
            const int y_idx =
                (int)((key >> KEY_Y_SHIFT) & KEY_IDX_MASK) + sel3i(ax1, a.g.gmin[0], a.g.gmin[1], a.g.gmin[2]);
            const int x_idx =
                (int)((key >> KEY_X_SHIFT) & KEY_IDX_MASK) + sel3i(ax0, a.g.gmin[0], a.g.gmin[1], a.g.gmin[2]);
            const float px = (float)(x_idx + 1) * sel3f(ax0, a.g.vs[0], a.g.vs[1], a.g.vs[2]);
            const float py = (float)(y_idx + 1) * sel3f(ax1, a.g.vs[0], a.g.vs[1], a.g.vs[2]);
            m0 = m0 + (ax0 == 0 ? px : (ax1 == 0 ? py : z));
            m1 = m1 + (ax0 == 1 ? px : (ax1 == 1 ? py : z));
            m2 = m2 + (ax0 == 2 ? px : (ax1 == 2 ? py : z));
            cnt = cnt + 1.0f;
            if ((key & 3ull) == 0ull)
                fl |= 1 << ax2;
        }
    }

    // regularised quadric and flags of the voxel in output row vid
    __device__ __forceinline__ void finish_at(uint64_t vid, const FoldArgs &a)
    {
        if (a.rw > 0.0f && cnt > 0.0f)
        {
            const float p0 = m0 / cnt, p1 = m1 / cnt, p2 = m2 / cnt;
            const float s = a.rw * cnt;
            const float R[10] = {1.0f, 0.0f, 0.0f, -p0, 1.0f, 0.0f, -p1, 1.0f, -p2, dot3f(p0, p1, p2, p0, p1, p2)};
#pragma unroll
            for (int k = 0; k < 10; ++k)
                q[k] = q[k] + s * R[k];
        }
        float *qo = a.qef + 10 * vid;
#pragma unroll
        for (int i = 0; i < 10; ++i)
            qo[i] = q[i];
        uint8_t *fo = a.flags + 3 * vid;
        fo[0] = (uint8_t)(fl & 1);
        fo[1] = (uint8_t)((fl >> 1) & 1);
        fo[2] = (uint8_t)((fl >> 2) & 1);
    }

    __device__ __forceinline__ void finish(int x, int y, int z, const FoldArgs &a)
    {
        int b;
        const uint64_t w = word_of(a.g, x, y, z, b);
        finish_at((uint64_t)voxel_id(a.g, w, b), a);
    }

    // Regularised quadric and flags of voxel (x, y, z) in output row vid,
    // and its dual vertex solved here from the registers (FDG.cpp:763-983):
    // the unconstrained minimiser when it lies in the voxel, else the voxel
    // is queued for the box-constrained search (solved in place when its
    // queue region is full).  The active lanes of the wave claim queue space
    // together (any subset of the wave's lanes may call it).
    __device__ __forceinline__ void finish_out(uint64_t vid, int x, int y, int z, const FoldArgs &a)
    {
        if (a.rw > 0.0f && cnt > 0.0f)
        {
            const float p0 = m0 / cnt, p1 = m1 / cnt, p2 = m2 / cnt;
            const float s = a.rw * cnt;
            const float R[10] = {1.0f, 0.0f, 0.0f, -p0, 1.0f, 0.0f, -p1, 1.0f, -p2, dot3f(p0, p1, p2, p0, p1, p2)};
#pragma unroll
            for (int k = 0; k < 10; ++k)
                q[k] = q[k] + s * R[k];
        }
        uint8_t *fo = a.flags + 3 * vid;
        fo[0] = (uint8_t)(fl & 1);
        fo[1] = (uint8_t)((fl >> 1) & 1);
        fo[2] = (uint8_t)((fl >> 2) & 1);
        const int c[3] = {x, y, z};
        const float vs[3] = {a.g.vs[0], a.g.vs[1], a.g.vs[2]};
        float d[3];
        const bool inside = solve_free(q, c, vs, d);
        const uint64_t m = __ballot(!inside);
        if (m)
        {
            const int leader = __ffsll((long long)m) - 1;
            const int lane = (int)(threadIdx.x & 63);
            unsigned long long base = ~0ull;
            if (lane == leader)
            {
                uint64_t first;
                const uint32_t shard = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % NSHARD;
                base = shard_claim_fit(a.box, shard, (uint64_t)__popcll(m), first) ? first : ~0ull;
            }
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, leader);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(base >> 32), leader);
            base = ((unsigned long long)hi << 32) | lo;
            if (!inside)
            {
                if (base != ~0ull)
                {
                    BoxItem &e = a.box.items[base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull))];
#pragma unroll
                    for (int i = 0; i < 10; ++i)
                        e.q[i] = q[i];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                    {
                        e.v[i] = d[i];
                        e.c[i] = c[i];
                    }
                    e.vid = vid;
                    return;
                }
                solve_box(q, c, vs, d);
            }
        }
        float *o = a.dual + 3 * vid;
        o[0] = d[0];
        o[1] = d[1];
        o[2] = d[2];
    }
};

// ---------------------------------------------------------------------------
// precision variant: the quadric of each voxel is accumulated in the voxel's
// own frame u = (p - c) / vs (c = the voxel's min corner), where its entries
// are O(vs^2) instead of the O(1) - O(1) cancellations of world coordinates:
// FP64 sums of the exact (FP64) products of FP32 plane coefficients, whose
// offsets n.c + d, like the intersection means, are compensated FP32 sums
// (TwoSum) of error-free products (TwoProduct); the world problem
// min_p [p,1]^T Q [p,1] over the voxel is the same problem over u in [0,1]^3.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void two_prod(float a, float b, float &p, float &e)
{
    p = a * b;
    e = __fmaf_rn(a, b, -p);
}

// compensated sum: value = h + l
struct CSum
{
    float h, l;
    __device__ __forceinline__ void zero() { h = l = 0.0f; }
    __device__ __forceinline__ void add(float x)
    {
        float s, e;
        two_sum(h, x, s, e);
        h = s;
        l = l + e;
    }
    __device__ __forceinline__ void add_prod(float a, float b)
    {
        float p, e;
        two_prod(a, b, p, e);
        add(p);
        l = l + e;
    }
    __device__ __forceinline__ void add_double(double x)
    {
        const float xh = (float)x;
        add(xh);
        add((float)(x - (double)xh));
    }
};

// FP64 accumulator of a quadric entry: products of two FP32 values are exact
// in FP64 (48 significant bits); read back as an unevaluated FP32 pair (h, l).
// Half the instructions of a compensated FP32 sum of TwoProducts.
struct DSum
{
    double v;
    __device__ __forceinline__ void zero() { v = 0.0; }
    __device__ __forceinline__ void add(float x) { v = v + (double)x; }
    __device__ __forceinline__ void add_prod(float a, float b) { v = v + (double)a * (double)b; }
    __device__ __forceinline__ void add_double(double x) { v = v + x; }
    __device__ __forceinline__ float hi() const { return (float)v; }
    __device__ __forceinline__ float lo() const { return (float)(v - (double)(float)v); }
};

struct FoldStateC
{
    DSum q[10];
    CSum m[3]; // sum of the intersection points in the voxel frame
    float cnt;
    int fl;
    float c[3]; // voxel min corner (world)

    __device__ __forceinline__ void init(int x, int y, int z, const FoldArgs &a)
    {
#pragma unroll
        for (int k = 0; k < 10; ++k)
            q[k].zero();
#pragma unroll
        for (int k = 0; k < 3; ++k)
            m[k].zero();
        cnt = 0.0f;
        fl = 0;
        c[0] = (float)x * a.g.vs[0];
        c[1] = (float)y * a.g.vs[1];
        c[2] = (float)z * a.g.vs[2];
    }

    // the compensated sums start at zero whatever the first record is
    __device__ __forceinline__ void begin(uint64_t)
    {
    }

    // plane n.p + d = 0 in the voxel frame: (n * vs).u + (n.c + d), the
    // offset formed from exact products and compensated sums
    __device__ __forceinline__ void add_plane(const float4 pl4, const FoldArgs &a)
    {
        float w[4];
        w[0] = pl4.x * a.g.vs[0];
        w[1] = pl4.y * a.g.vs[1];
        w[2] = pl4.z * a.g.vs[2];
        CSum d;
        d.h = pl4.w;
        d.l = 0.0f;
        d.add_prod(pl4.x, c[0]);
        d.add_prod(pl4.y, c[1]);
        d.add_prod(pl4.z, c[2]);
        w[3] = d.h + d.l;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = i; j < 4; ++j)
                q[qidx(i, j)].add_prod(w[i], w[j]);
    }

    __device__ __forceinline__ void add_pl(uint64_t key, float z, const float4 pl4, const FoldArgs &a)
    {
        const int phase = (int)(key >> KEY_PHASE_SHIFT);
        if (phase == 0)
        {
            const int ax2 = (int)((key >> KEY_AX_SHIFT) & 3ull);
            const int ax0 = axis_next(ax2);
            const int ax1 = axis_prev(ax2);
            const int y_idx =
                (int)((key >> KEY_Y_SHIFT) & KEY_IDX_MASK) + sel3i(ax1, a.g.gmin[0], a.g.gmin[1], a.g.gmin[2]);
            const int x_idx =
                (int)((key >> KEY_X_SHIFT) & KEY_IDX_MASK) + sel3i(ax0, a.g.gmin[0], a.g.gmin[1], a.g.gmin[2]);
            const float px = (float)(x_idx + 1) * sel3f(ax0, a.g.vs[0], a.g.vs[1], a.g.vs[2]);
            const float py = (float)(y_idx + 1) * sel3f(ax1, a.g.vs[0], a.g.vs[1], a.g.vs[2]);
            const float p[3] = {ax0 == 0 ? px : (ax1 == 0 ? py : z), ax0 == 1 ? px : (ax1 == 1 ? py : z),
                                ax0 == 2 ? px : (ax1 == 2 ? py : z)};
#pragma unroll
            for (int i = 0; i < 3; ++i)
                m[i].add(div_vs(p[i] - c[i], a.g, i));
            cnt = cnt + 1.0f;
            if ((key & 3ull) == 0ull)
                fl |= 1 << ax2;
        }
        if (phase != 2)
            add_plane(pl4, a);
        else
        {
            // bw |A (p - p0)|^2 with A = I - d d^T: in the voxel frame
            // Q_ij = vs_i vs_j A_ij, Q_i3 = -vs_i (A r0)_i, Q_33 = r0^T A r0, r0 = p0 - c
            // entries formed one row of A at a time (few doubles live at once)
            const uint64_t e = key & ((1ull << KEY_PHASE_SHIFT) - 1ull);
            const BEdge &be = a.bedges[e];
            const double d[3] = {be.dir[0], be.dir[1], be.dir[2]};
            const double r0[3] = {(double)be.v0[0] - (double)c[0], (double)be.v0[1] - (double)c[1],
                                  (double)be.v0[2] - (double)c[2]};
            const double bw = (double)a.bw;
            double rAr = 0.0;
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                double Ai[3];
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    Ai[j] = (i == j ? 1.0 : 0.0) - d[i] * d[j];
                const double Ar = Ai[0] * r0[0] + Ai[1] * r0[1] + Ai[2] * r0[2];
#pragma unroll
                for (int j = i; j < 3; ++j)
                    q[qidx(i, j)].add_double(bw * ((double)a.g.vs[i] * (double)a.g.vs[j] * Ai[j]));
                q[qidx(i, 3)].add_double(bw * (-(double)a.g.vs[i] * Ar));
                rAr = i == 0 ? r0[0] * Ar : rAr + r0[i] * Ar;
            }
            q[qidx(3, 3)].add_double(bw * rAr);
        }
    }

    __device__ __forceinline__ void add(uint64_t key, float z, const FoldArgs &a)
    {
        const uint32_t f = FoldState::plane_of(key);
        add_pl(key, z, f != FoldState::NO_PLANE ? a.planes[f] : make_float4(0.0f, 0.0f, 0.0f, 0.0f), a);
    }

    // regularisation rw cnt |p - mean|^2 in the voxel frame, scaling by a
    // power of two (largest diagonal entry to [1, 2)), quadric (h and l
    // parts) and flags of the voxel in output row vid
    __device__ __forceinline__ void finish_at(uint64_t vid, const FoldArgs &a)
    {
        if (a.rw > 0.0f && cnt > 0.0f)
        {
            const float s = a.rw * cnt;
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                const float mu = (m[i].h + m[i].l) / cnt;
                const float t = s * (a.g.vs[i] * a.g.vs[i]);
                q[qidx(i, i)].add(t);
                q[qidx(i, 3)].add_prod(-t, mu);
                float tm, te;
                two_prod(t, mu, tm, te);
                q[qidx(3, 3)].add_prod(tm, mu);
                q[qidx(3, 3)].add_prod(te, mu);
            }
        }
        const float dmax = fmaxf(fmaxf(q[qidx(0, 0)].hi(), q[qidx(1, 1)].hi()), q[qidx(2, 2)].hi());
        int ex = 0;
        if (dmax > 0.0f && dmax <= 3.0e38f)
            frexpf(dmax, &ex);
        float qh[10], ql[10];
#pragma unroll
        for (int i = 0; i < 10; ++i)
        {
            qh[i] = ldexpf(q[i].hi(), 1 - ex);
            ql[i] = ldexpf(q[i].lo(), 1 - ex);
        }
        uint8_t *fo = a.flags + 3 * vid;
        fo[0] = (uint8_t)(fl & 1);
        fo[1] = (uint8_t)((fl >> 1) & 1);
        fo[2] = (uint8_t)((fl >> 2) & 1);
        // the FP32 solve of k_solve_voxels_c, done here when the system is
        // well conditioned and its minimiser lies in the voxel (the common
        // case); otherwise the scaled quadric is stored and the voxel
        // flagged for that kernel (FP64 fallback, box search)
        float Q[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                Q[i][j] = qh[qidx(i, j)] + ql[qidx(i, j)];
        const float A[3][3] = {{Q[0][0], Q[0][1], Q[0][2]}, {Q[1][0], Q[1][1], Q[1][2]}, {Q[2][0], Q[2][1], Q[2][2]}};
        const float b[3] = {-Q[0][3], -Q[1][3], -Q[2][3]};
        Cpqr<3> f;
        f.factor(A);
        const float r0 = fabsf(f.qr[0][0]), rl = fabsf(f.qr[2][2]);
        bool solved = false;
        if (!(f.rk < 3 || !(r0 <= a.cond_max * rl)))
        {
            float u[3];
            f.solve(b, u);
            if (u[0] >= 0.0f && u[0] <= 1.0f && u[1] >= 0.0f && u[1] <= 1.0f && u[2] >= 0.0f && u[2] <= 1.0f)
            {
                float *o = a.dual + 3 * vid;
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    o[i] = c[i] + a.g.vs[i] * u[i];
                solved = true;
            }
        }
        if (!solved)
        {
            float *qo = a.qef + 20 * vid;
#pragma unroll
            for (int i = 0; i < 10; ++i)
            {
                qo[i] = qh[i];
                qo[10 + i] = ql[i];
            }
            atomicOr(a.solve_bits + (vid >> 5), 1u << (uint32_t)(vid & 31ull));
        }
    }

    // flags of voxel (x, y, z) in output row vid and its dual vertex, or its
    // quadric for k_solve_voxels_c
    __device__ __forceinline__ void finish_out(uint64_t vid, int, int, int, const FoldArgs &a)
    {
        finish_at(vid, a);
    }
};

template <int P>
struct FoldPick
{
    using type = FoldState;
};
template <>
struct FoldPick<1>
{
    using type = FoldStateC;
};

constexpr int FOLD_WAVES = 6;    // occupancy of the reference-path 256-thread fold (waves per SIMD: at most 85 VGPRs; 6 blocks per CU, its LDS limit)
constexpr int FOLD_CAP = 4096;   // records held in LDS at a time by the kernel for large bricks (2 blocks per CU)

// LDS view of a batch of up to NT chunks of one brick (NT = threads of the block)
template <int NT>
struct ChunkBatchT
{
    uint32_t start[NT];
    uint32_t incl[NT];
    uint32_t sh[NT / 64];
};
using ChunkBatch = ChunkBatchT<BT>;

// one brick's records in LDS and the per-voxel grouping built over them
template <int CAP>
struct FoldLds
{
    static_assert(CAP <= 65536, "16-bit record indices");
    uint64_t key[CAP];
    float z[CAP];
    uint16_t lv[CAP];
    uint16_t perm[CAP]; // record indices grouped by voxel, then key-ordered inside each voxel
    uint32_t off[513];   // first grouped slot of each voxel
    uint32_t cur[512];   // per-voxel record counts
    uint16_t vlist[512]; // voxels with records
    uint64_t rbits[64];  // occupancy word of each (y, z) row of the brick
    uint32_t roff[64];   // ... and its first voxel id
    uint32_t sh[NWAVE];
    uint32_t n, m;
    ChunkBatch cb;
};

// Voxel entries of a record in the brick with min corner (x0, y0, z0) of the
// current mesh: returns a mask of entries k (bit k) and their keys' low bits
// and brick-local voxel indices.  A crossing record (phase 0) stands for the
// crossing's 4 voxels (k = dx * 2 + dy, FDG.cpp:239-261) that lie in the
// brick, with keys key | k; any other record is one entry, its own voxel.
__device__ __forceinline__ uint32_t rec_entries(const FoldArgs &a, uint64_t key, uint32_t lvox, int x0, int y0,
                                                int z0, uint32_t (&lv)[4])
{
    if ((key >> KEY_PHASE_SHIFT) != 0)
    {
        lv[0] = lvox;
        lv[1] = lv[2] = lv[3] = 0u;
        return 1u;
    }
    const int ax2 = (int)((key >> KEY_AX_SHIFT) & 3ull);
    const int ax0 = axis_next(ax2);
    const int ax1 = axis_prev(ax2);
    const int y_idx = (int)((key >> KEY_Y_SHIFT) & KEY_IDX_MASK) + sel3i(ax1, a.g.gmin[0], a.g.gmin[1], a.g.gmin[2]);
    const int x_idx = (int)((key >> KEY_X_SHIFT) & KEY_IDX_MASK) + sel3i(ax0, a.g.gmin[0], a.g.gmin[1], a.g.gmin[2]);
    int c[4][3];
    event_voxels(ax0, ax1, x_idx, y_idx, (int)lvox, c);
    const uint32_t e = 1u << a.bs;
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const uint32_t lx = (uint32_t)(c[k][0] - x0), ly = (uint32_t)(c[k][1] - y0), lz = (uint32_t)(c[k][2] - z0);
        lv[k] = lx | (ly << 3) | (lz << 6);
        m |= (lx < e && ly < e && lz < e) ? (1u << k) : 0u;
    }
    return m;
}

// Calls f(rec) for every record of `brick`, all threads cooperating.
// Contains barriers: every thread of the block must call it.
template <class Fn>
__device__ __forceinline__ void for_each_rec(const FoldArgs &a, uint32_t brick, ChunkBatch &cb, Fn &&f)
{
    const uint32_t c0 = a.choff[brick], c1 = a.choff[brick + 1];
    for (uint32_t cbeg = c0; cbeg < c1; cbeg += BT)
    {
        const uint32_t ci = cbeg + threadIdx.x;
        const int nch = (int)(c1 - cbeg < (uint32_t)BT ? c1 - cbeg : (uint32_t)BT);
        uint32_t cnt = 0;
        if (ci < c1)
        {
            const uint2 ch = a.gchunks[ci];
            cb.start[threadIdx.x] = ch.x;
            cnt = ch.y;
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan(cnt, cb.sh, tot);
        cb.incl[threadIdx.x] = ex + cnt;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < tot; j += BT)
        {
            const int k = upper_find(cb.incl, nch, j);
            const uint32_t before = k ? cb.incl[k - 1] : 0u;
            f(a.recs[(uint64_t)cb.start[k] + (j - before)]);
        }
        __syncthreads();
    }
}

// Calls f(key, z, lvox) for every voxel entry of the records of `brick`
// (min corner x0, y0, z0), all threads cooperating.  Contains barriers.
template <class Fn>
__device__ __forceinline__ void for_each_entry(const FoldArgs &a, uint32_t brick, ChunkBatch &cb, int x0, int y0,
                                               int z0, Fn &&f)
{
    for_each_rec(a, brick, cb, [&](const Rec &r) {
        uint32_t lv[4];
        const uint32_t m = rec_entries(a, r.key, r.lvox, x0, y0, z0, lv);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (m & (1u << k))
                f(r.key | (uint64_t)k, r.z, lv[k]);
    });
}

// key, crossing coordinate and face plane of the grouped records i .. i+FETCH-1 below s1
// (zeros past s1; boundary records have no plane)
template <int CAP, int FETCH>
__device__ __forceinline__ void fetch_recs(const FoldArgs &a, const FoldLds<CAP> &L, uint32_t i, uint32_t s1,
                                           uint64_t kk[FETCH], float zz[FETCH], float4 pp[FETCH])
{
#pragma unroll
    for (int u = 0; u < FETCH; ++u)
    {
        kk[u] = 0;
        zz[u] = 0.0f;
        pp[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (i + u < s1)
        {
            const uint32_t j = L.perm[i + u];
            kk[u] = L.key[j];
            zz[u] = L.z[j];
            const uint32_t f = FoldState::plane_of(kk[u]);
            if (f != FoldState::NO_PLANE)
                pp[u] = a.planes[f];
        }
    }
}

// The brick's n records are in L.key/z/lv and L.cur holds their per-voxel
// histogram: group them by voxel (counting sort), order each voxel's records
// by key (rank sort, keys are unique), replay them in that order, one thread
// per voxel with records, and solve.  Contains barriers.
template <int CAP, int P>
__device__ __forceinline__ void fold_loaded(const FoldArgs &a, FoldLds<CAP> &L, uint32_t n, int x0, int y0, int z0)
{
    const int tid = threadIdx.x;
    uint32_t nv;
    {
        const uint32_t c0 = L.cur[2 * tid], c1 = L.cur[2 * tid + 1];
        uint32_t tot;
        const uint32_t ex = block_excl_scan(c0 + c1, L.sh, tot);
        L.off[2 * tid] = ex;
        L.off[2 * tid + 1] = ex + c0;
        if (tid == 0)
            L.off[512] = tot;
        const uint32_t nz = (c0 ? 1u : 0u) + (c1 ? 1u : 0u);
        const uint32_t exv = block_excl_scan(nz, L.sh, nv);
        if (c0)
            L.vlist[exv] = (uint16_t)(2 * tid);
        if (c1)
            L.vlist[exv + (c0 ? 1u : 0u)] = (uint16_t)(2 * tid + 1);
    }
    __syncthreads();
    for (uint32_t i = tid; i < n; i += BT)
    {
        const uint32_t l = L.lv[i];
        L.perm[L.off[l] + atomicSub(&L.cur[l], 1u) - 1u] = (uint16_t)i;
    }
    __syncthreads();
    // rank of each record among its voxel's keys; the permutation is
    // rewritten in place once every rank is known
    constexpr int PER = (CAP + BT - 1) / BT;
    uint32_t rk[PER];
#pragma unroll
    for (int it = 0; it < PER; ++it)
    {
        const uint32_t p = (uint32_t)(tid + it * BT);
        rk[it] = 0xffffffffu;
        if (p < n)
        {
            const uint32_t me = L.perm[p];
            const uint32_t l = L.lv[me];
            const uint32_t s0 = L.off[l], s1 = L.off[l + 1];
            const uint64_t k = L.key[me];
            uint32_t rank = 0;
            for (uint32_t t = s0; t < s1; ++t)
                rank += (L.key[L.perm[t]] < k) ? 1u : 0u;
            rk[it] = ((s0 + rank) << 16) | me;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < PER; ++it)
        if (rk[it] != 0xffffffffu)
            L.perm[rk[it] >> 16] = (uint16_t)(rk[it] & 0xffffu);
    __syncthreads();
    for (uint32_t vi = tid; vi < nv; vi += BT)
    {
        const int l = L.vlist[vi];
        const uint32_t s0 = L.off[l], s1 = L.off[l + 1];
        typename FoldPick<P>::type st;
        st.init(x0 + (l & 7), y0 + ((l >> 3) & 7), z0 + (l >> 6), a);
        st.begin(L.key[L.perm[s0]]);
        // the planes of up to FETCH records are fetched together, then folded in order
        constexpr int FETCH = P ? 1 : 4;
        for (uint32_t i = s0; i < s1; i += FETCH)
        {
            uint64_t kk[FETCH];
            float zz[FETCH];
            float4 pp[FETCH];
            fetch_recs<CAP, FETCH>(a, L, i, s1, kk, zz, pp);
#pragma unroll
            for (int u = 0; u < FETCH; ++u)
                if (i + u < s1)
                    st.add_pl(kk[u], zz[u], pp[u], a);
        }
        st.finish_out(row_vid(a, L, x0, l), x0 + (l & 7), y0 + ((l >> 3) & 7), z0 + (l >> 6), a);
    }
    __syncthreads();
}

// Output rows of the brick's voxels: the 8 voxels of a (y, z) row of a brick
// share one 64-bit occupancy word (bricks start at multiples of 8 along x).
// Loaded by the first wave; read after the block's next barrier.
template <class LDS>
__device__ __forceinline__ void load_rows(const FoldArgs &a, LDS &L, int x0, int y0, int z0)
{
    const int r = threadIdx.x;
    if (r < 64)
    {
        const int y = y0 + (r & 7), z = z0 + (r >> 3);
        uint64_t m = 0;
        uint32_t o = 0;
        if (y < a.g.gmax[1] && z < a.g.gmax[2])
        {
            int bit;
            const uint64_t w = word_of(a.g, x0, y, z, bit);
            m = a.g.bits[w];
            o = a.g.woff[w];
        }
        L.rbits[r] = m;
        L.roff[r] = o;
    }
}

template <class LDS>
__device__ __forceinline__ uint64_t row_vid(const FoldArgs &a, const LDS &L, int x0, int l)
{
    const int row = l >> 3;
    const int bit = (x0 - a.g.gmin[0] + (l & 7)) & 63;
    return (uint64_t)L.roff[row] + (uint64_t)__popcll(L.rbits[row] & ((1ull << bit) - 1ull));
}

// Sets the grid and brick grid of the mesh owning `brick` (a brick of the
// batch's shared numbering) and returns the brick's index inside its mesh.
__device__ __forceinline__ uint32_t enter_brick(FoldArgs &a, uint32_t brick)
{
    const MeshSeg &sg = a.b.seg[mesh_of_brick(a.b, (uint64_t)brick)];
    a.g = sg.g;
    a.nb[0] = sg.nb[0];
    a.nb[1] = sg.nb[1];
    a.nb[2] = sg.nb[2];
    a.bs = sg.bs;
    return brick - (uint32_t)sg.brick0;
}

// min corner of a brick of the current mesh (lb = index inside the mesh)
__device__ __forceinline__ void brick_origin(const FoldArgs &a, uint32_t brick, int &x0, int &y0, int &z0)
{
    const int bx = (int)(brick % (uint32_t)a.nb[0]);
    const int byz = (int)(brick / (uint32_t)a.nb[0]);
    const int by = byz % a.nb[1];
    const int bz = byz / a.nb[1];
    x0 = a.g.gmin[0] + (bx << a.bs);
    y0 = a.g.gmin[1] + (by << a.bs);
    z0 = a.g.gmin[2] + (bz << a.bs);
}

// LDS of the register-grouped brick fold (blocks of NT threads): the brick's
// records grouped by voxel and, inside each voxel, in key order
template <int CAP, int NT = BT>
struct FoldLdsG
{
    // 12 B per record and 16-bit offsets and counters: 24 KB at 1536
    // records (6 blocks of 256 threads per CU), 14 KB at 768 (11 blocks of
    // 128 threads)
    static_assert(CAP <= 65535, "16-bit record offsets");
    uint64_t key[CAP];
    float z[CAP];
    uint16_t off[513];   // first record of each voxel
    // per-voxel record counts, then fill cursors: in blocks of at most 256
    // threads 16 bits each (voxel l in half l & 1 of word l >> 1), 1 KB less
    // LDS; 32 bits in larger blocks, whose threads own one voxel each
    static constexpr bool PACK = NT <= 256;
    static constexpr int CUR_WORDS = PACK ? 256 : 512;
    uint32_t cur[CUR_WORDS];
    uint16_t vlist[512]; // voxels with records
    __device__ __forceinline__ uint32_t cnt_get(int l) const
    {
        return PACK ? (cur[l >> 1] >> ((l & 1) * 16)) & 0xffffu : cur[l];
    }
    // adds one to voxel l's counter and returns its previous value (the
    // counters stay below CAP < 2^16, so no carry reaches the other half)
    __device__ __forceinline__ uint32_t cnt_inc(uint32_t l)
    {
        if constexpr (PACK)
            return (atomicAdd(&cur[l >> 1], 1u << ((l & 1u) * 16u)) >> ((l & 1u) * 16u)) & 0xffffu;
        else
            return atomicAdd(&cur[l], 1u);
    }
    uint64_t rbits[64];  // occupancy word of each (y, z) row of the brick
    uint32_t roff[64];   // ... and its first voxel id
    uint32_t sh[NT / 64];
    ChunkBatchT<NT> cb;
};

// The brick's n records are in registers (record g = tid + k*NT in slot k)
// and L.cur holds the per-voxel histogram of their voxel entries
// (rec_entries): each entry goes to its voxel's segment, is ranked there
// among the voxel's keys (keys are unique) and moved to its rank; then one
// thread per voxel with records replays them in key order and solves.
// Returns false, before anything is folded, when the entries exceed CAP.
// Contains barriers.
template <int CAP, int P, int NT = BT>
__device__ __forceinline__ bool fold_grouped(const FoldArgs &a, FoldLdsG<CAP, NT> &L, uint32_t n,
                                             const uint64_t (&rk)[CAP / NT], const float (&rz)[CAP / NT],
                                             const uint32_t (&rl)[CAP / NT], int x0, int y0, int z0)
{
    constexpr int PER = CAP / NT;
    constexpr int NW = NT / 64;
    constexpr int VPT = 512 / NT; // voxel slots of the brick per thread
    const int tid = threadIdx.x;
    uint32_t nv, tot;
    {
        uint32_t c[VPT], sum = 0, nz = 0;
#pragma unroll
        for (int j = 0; j < VPT; ++j)
        {
            c[j] = L.cnt_get(VPT * tid + j);
            sum += c[j];
            nz += c[j] ? 1u : 0u;
        }
        uint32_t ex = block_excl_scan_w<NW>(sum, L.sh, tot);
        if (tot > (uint32_t)CAP)
            return false; // block-uniform
        uint32_t exj[VPT];
#pragma unroll
        for (int j = 0; j < VPT; ++j)
        {
            L.off[VPT * tid + j] = (uint16_t)ex;
            exj[j] = ex;
            ex += c[j];
        }
        // the fill cursors start at the segment offsets
        if constexpr (FoldLdsG<CAP, NT>::PACK)
        {
            static_assert(VPT % 2 == 0, "whole counter words per thread");
#pragma unroll
            for (int j = 0; j < VPT; j += 2)
                L.cur[(VPT * tid + j) >> 1] = exj[j] | (exj[j + 1] << 16);
        }
        else
        {
#pragma unroll
            for (int j = 0; j < VPT; ++j)
                L.cur[VPT * tid + j] = exj[j];
        }
        if (tid == 0)
            L.off[512] = (uint16_t)tot;
        uint32_t exv = block_excl_scan_w<NW>(nz, L.sh, nv);
#pragma unroll
        for (int j = 0; j < VPT; ++j)
            if (c[j])
                L.vlist[exv++] = (uint16_t)(VPT * tid + j);
    }
    __syncthreads();
    // entries into their voxels' segments, in any order (the slot loops stop
    // at the block's last slot in use)
#pragma unroll
    for (int k = 0; k < PER; ++k)
    {
        if ((uint32_t)(k * NT) >= n)
            break;
        if ((uint32_t)(tid + k * NT) < n)
        {
            uint32_t lv[4];
            const uint32_t em = rec_entries(a, rk[k], rl[k], x0, y0, z0, lv);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (em & (1u << q))
                {
                    const uint32_t u = L.cnt_inc(lv[q]);
                    L.key[u] = rk[k] | (uint64_t)q;
                    L.z[u] = rz[k];
                }
        }
    }
    __syncthreads();
    // rank of the entry at each position among its voxel's keys
    uint64_t ek[PER];
    float ez[PER];
    uint32_t pos[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i)
    {
        pos[i] = 0;
        ek[i] = 0;
        ez[i] = 0.0f;
        if ((uint32_t)(i * NT) >= tot)
            break;
        const uint32_t u = (uint32_t)(tid + i * NT);
        if (u < tot)
        {
            const uint64_t key = L.key[u];
            ek[i] = key;
            ez[i] = L.z[u];
            const int v = upper_find(L.off + 1, 512, (uint16_t)u); // the voxel whose segment holds u
            const uint32_t s0 = L.off[v], s1 = L.off[v + 1];
            uint32_t rank = 0;
            for (uint32_t t = s0; t < s1; ++t)
                rank += (L.key[t] < key) ? 1u : 0u;
            pos[i] = s0 + rank;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i)
    {
        if ((uint32_t)(i * NT) >= tot)
            break;
        if ((uint32_t)(tid + i * NT) < tot)
        {
            L.key[pos[i]] = ek[i];
            L.z[pos[i]] = ez[i];
        }
    }
    __syncthreads();
    for (uint32_t vi = tid; vi < nv; vi += NT)
    {
        const int l = L.vlist[vi];
        const uint32_t s0 = L.off[l], s1 = L.off[l + 1];
        typename FoldPick<P>::type st;
        st.init(x0 + (l & 7), y0 + ((l >> 3) & 7), z0 + (l >> 6), a);
        st.begin(L.key[s0]);
        // the planes of up to FETCH records are fetched together, then folded
        // in order (one at a time for the precision variant, whose larger
        // state leaves no registers for more)
        constexpr int FETCH = P ? 1 : 4;
        for (uint32_t i = s0; i < s1; i += FETCH)
        {
            uint64_t kk[FETCH];
            float zz[FETCH];
            float4 pp[FETCH];
#pragma unroll
            for (int u = 0; u < FETCH; ++u)
            {
                kk[u] = 0;
                zz[u] = 0.0f;
                pp[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (i + u < s1)
                {
                    kk[u] = L.key[i + u];
                    zz[u] = L.z[i + u];
                    const uint32_t f = FoldState::plane_of(kk[u]);
                    if (f != FoldState::NO_PLANE)
                        pp[u] = a.planes[f];
                }
            }
#pragma unroll
            for (int u = 0; u < FETCH; ++u)
                if (i + u < s1)
                    st.add_pl(kk[u], zz[u], pp[u], a);
        }
        st.finish_out(row_vid(a, L, x0, l), x0 + (l & 7), y0 + ((l >> 3) & 7), z0 + (l >> 6), a);
    }
    __syncthreads();
    return true;
}

// One workgroup of NT threads per brick of `list`, for bricks of at most CAP
// records: the records are read once into registers, grouped and ordered
// in LDS, replayed and solved there.  Larger bricks are appended to `next`
// for the next tier; with count_recs the records folded here are counted.
template <int P, int CAP, int NT>
__device__ __forceinline__ void fold_bricks_body(const FoldArgs &ain, const uint32_t *list, uint64_t n_list,
                                                 uint32_t *next, unsigned long long *n_next, bool count_recs)
{
    __shared__ FoldLdsG<CAP, NT> L;
    constexpr int PER = CAP / NT;
    constexpr int NW = NT / 64;
    static_assert(CAP % NT == 0, "record slots per thread");
    static_assert(NT >= 64 && NT <= 512 && 512 % NT == 0, "block size");
    const int tid = threadIdx.x;
    for (uint64_t bi = blockIdx.x; bi < n_list; bi += gridDim.x)
    {
        const uint32_t brick = list[bi];
        FoldArgs a = ain;
        const uint32_t lb = enter_brick(a, brick);
        const uint32_t c0 = a.choff[brick], c1 = a.choff[brick + 1];
        int x0, y0, z0;
        brick_origin(a, lb, x0, y0, z0);
        load_rows(a, L, x0, y0, z0);
        for (int i = tid; i < FoldLdsG<CAP, NT>::CUR_WORDS; i += NT)
            L.cur[i] = 0;
        // the brick's chunks, NT at a time: record g of the brick goes to
        // slot g / NT of thread g % NT
        uint64_t rk[PER];
        float rz[PER];
        uint32_t rl[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k)
        {
            rk[k] = 0;
            rz[k] = 0.0f;
            rl[k] = 0;
        }
        uint32_t n = 0;
        bool over = false;
        for (uint32_t cbeg = c0; cbeg < c1; cbeg += NT)
        {
            const int nb = (int)(c1 - cbeg < (uint32_t)NT ? c1 - cbeg : (uint32_t)NT);
            uint32_t cnt = 0;
            if (tid < nb)
            {
                const uint2 ch = a.gchunks[cbeg + tid];
                L.cb.start[tid] = ch.x;
                cnt = ch.y;
            }
            uint32_t tot;
            const uint32_t ex = block_excl_scan_w<NW>(cnt, L.cb.sh, tot);
            if (n + tot > (uint32_t)CAP)
            {
                over = true;
                break;
            }
            L.cb.incl[tid] = ex + cnt;
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PER; ++k)
            {
                // slots of this batch's records only
                if ((uint32_t)(k * NT) >= n + tot)
                    break;
                if ((uint32_t)((k + 1) * NT) <= n)
                    continue;
                const uint32_t g = (uint32_t)(tid + k * NT);
                if (g >= n && g < n + tot)
                {
                    const uint32_t j = g - n;
                    const int c = upper_find(L.cb.incl, nb, j);
                    const uint32_t before = c ? L.cb.incl[c - 1] : 0u;
                    const Rec r = a.recs[(uint64_t)L.cb.start[c] + (j - before)];
                    rk[k] = r.key;
                    rz[k] = r.z;
                    rl[k] = r.lvox;
                    uint32_t lv[4];
                    const uint32_t em = rec_entries(a, r.key, r.lvox, x0, y0, z0, lv);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (em & (1u << q))
                            L.cnt_inc(lv[q]);
                }
            }
            __syncthreads();
            n += tot;
        }
        if (over)
        {
            if (tid == 0)
                next[atomicAdd(n_next, 1ull)] = brick;
            __syncthreads();
            continue;
        }
        if (!fold_grouped<CAP, P, NT>(a, L, n, rk, rz, rl, x0, y0, z0))
        {
            // more voxel entries than the tile holds (not expected: bricks are
            // listed by their exact entry counts): the next tier folds it
            if (tid == 0)
                next[atomicAdd(n_next, 1ull)] = brick;
            __syncthreads();
            continue;
        }
        if (count_recs && tid == 0)
            atomicAdd(a.ctr + CTR_LARGE_RECS, (unsigned long long)n);
    }
}

// The kernels, one per tier of bricks by voxel entries (a conversion whose
// record buffers overflowed skips the fold: the host sizes the buffers
// again and re-runs its record stages):
//  - k_fold_tiny: bricks of up to FOLD_TINY_CAP entries, one wave each
//    (10 KB of LDS: 16 bricks per CU, and no barrier waits on another wave);
//  - k_fold_small: bricks of up to FOLD_SMALL_CAP entries, in blocks of 128
//    threads (5 waves per SIMD and 15 KB of LDS: 10 blocks per CU), so that
//    a CU folds 10 bricks at a time;
//  - k_fold_bricks(_ref): bricks of up to FOLD_SMALL records (25 KB of LDS:
//    6 blocks of 256 threads per CU on the reference path, 5 otherwise);
//  - k_fold_mid(_ref): bricks of up to FOLD_MID records (3 blocks per CU).
// Each brick is listed for its tier by k_active_bricks; a brick too large
// for a tier's tile (not expected: the counts are exact) goes to the next.
template <int P>
__global__ __launch_bounds__(FOLD_TINY_NT) __attribute__((amdgpu_waves_per_eu(FOLD_TINY_WAVES))) void k_fold_tiny(FoldArgs a)
{
    if (a.ctr[CTR_OVER])
        return;
    fold_bricks_body<P, FOLD_TINY_CAP, FOLD_TINY_NT>(a, a.tiny, *a.n_tiny, a.active, a.n_active, false);
}
template <int P>
__global__ __launch_bounds__(FOLD_SMALL_NT) __attribute__((amdgpu_waves_per_eu(5))) void k_fold_small(FoldArgs a)
{
    if (a.ctr[CTR_OVER])
        return;
    fold_bricks_body<P, FOLD_SMALL_CAP, FOLD_SMALL_NT>(a, a.active, *a.n_active, a.medium, a.n_medium, false);
}
template <int P>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(5))) void k_fold_bricks(FoldArgs a)
{
    if (a.ctr[CTR_OVER])
        return;
    fold_bricks_body<P, FOLD_SMALL, BT>(a, a.medium, *a.n_medium, a.large, a.n_large, false);
}
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(FOLD_WAVES))) void k_fold_bricks_ref(FoldArgs a)
{
    if (a.ctr[CTR_OVER])
        return;
    fold_bricks_body<0, FOLD_SMALL, BT>(a, a.medium, *a.n_medium, a.large, a.n_large, false);
}
template <int P>
__global__ __launch_bounds__(BT) void k_fold_mid(FoldArgs a)
{
    if (a.ctr[CTR_OVER])
        return;
    fold_bricks_body<P, FOLD_MID, BT>(a, a.large, *a.n_large, a.huge, a.n_huge, true);
}
// the reference path's mid tier in blocks of FOLD_MID_REF_NT threads: the
// same 3 bricks per CU (LDS), twice the waves to hide their loads, and every
// voxel of a brick replayed in one round
constexpr int FOLD_MID_REF_NT = 512;
__global__ __launch_bounds__(FOLD_MID_REF_NT) __attribute__((amdgpu_waves_per_eu(6))) void k_fold_mid_ref(FoldArgs a)
{
    if (a.ctr[CTR_OVER])
        return;
    fold_bricks_body<0, FOLD_MID, FOLD_MID_REF_NT>(a, a.large, *a.n_large, a.huge, a.n_huge, true);
}

// Bricks with more records than k_fold_bricks holds.  A brick
// whose records fit FOLD_CAP is folded like there; otherwise voxel ranges
// are sized to the LDS tile, and a voxel with more records than the tile is
// replayed in key-ordered batches selected by a binary search over the key.
template <int P>
__global__ __launch_bounds__(BT) void k_fold_large(FoldArgs ain)
{
    __shared__ FoldLds<FOLD_CAP> L;
    __shared__ uint64_t next_bi;
    const int tid = threadIdx.x;
    if (ain.ctr[CTR_OVER])
        return;
    const uint64_t n_huge = *ain.n_huge;
    if (blockIdx.x >= n_huge)
        return; // no brick would be left for this block

    // bricks are taken one at a time from a shared counter (their sizes vary
    // widely: a fixed share per block leaves a long tail)
    for (;;)
    {
        __syncthreads(); // every thread has read the last brick's index
        if (tid == 0)
            next_bi = (uint64_t)atomicAdd(ain.ctr + CTR_HUGE_NEXT, 1ull);
        __syncthreads();
        const uint64_t bi = next_bi;
        if (bi >= n_huge)
            break;
        const uint32_t brick = ain.huge[bi];
        FoldArgs a = ain;
        const uint32_t lb = enter_brick(a, brick);
        int x0, y0, z0;
        brick_origin(a, lb, x0, y0, z0);
        load_rows(a, L, x0, y0, z0);

        // one pass over the brick's records: per-voxel histogram, and the
        // records themselves into LDS while they fit
        for (int i = tid; i < 512; i += BT)
            L.cur[i] = 0;
        if (tid == 0)
            L.n = 0;
        __syncthreads();
        for_each_entry(a, brick, L.cb, x0, y0, z0, [&](uint64_t key, float z, uint32_t lvox) {
            atomicAdd(&L.cur[lvox], 1u);
            const uint32_t i = atomicAdd(&L.n, 1u);
            if (i < (uint32_t)FOLD_CAP)
            {
                L.key[i] = key;
                L.z[i] = z;
                L.lv[i] = (uint16_t)lvox;
            }
        });
        const uint32_t n_all = L.n;
        if (tid == 0)
            atomicAdd(a.ctr + CTR_LARGE_RECS, (unsigned long long)n_all);
        if (n_all <= (uint32_t)FOLD_CAP)
        {
            fold_loaded<FOLD_CAP, P>(a, L, n_all, x0, y0, z0);
            continue;
        }
        {
            // exclusive scan of the 512 counts (2 per thread)
            const uint32_t c0 = L.cur[2 * tid], c1 = L.cur[2 * tid + 1];
            uint32_t tot;
            const uint32_t ex = block_excl_scan(c0 + c1, L.sh, tot);
            L.off[2 * tid] = ex;
            L.off[2 * tid + 1] = ex + c0;
            if (tid == 0)
                L.off[512] = tot;
        }
        __syncthreads();

        int lo = 0;
        while (lo < 512)
        {
            const uint32_t nlo = L.off[lo + 1] - L.off[lo];
            if (nlo == 0)
            {
                ++lo;
                continue;
            }
            if (nlo > (uint32_t)FOLD_CAP)
            {
                // ---- one voxel with more records than the tile: key-ordered batches
                const uint32_t lv = (uint32_t)lo;
                typename FoldPick<P>::type st;
                st.init(x0 + (int)(lv & 7), y0 + (int)((lv >> 3) & 7), z0 + (int)(lv >> 6), a);
                uint64_t last = 0;
                bool have_last = false;
                uint32_t done = 0;
                while (done < nlo)
                {
                    uint64_t tlo = have_last ? last + 1ull : 0ull, thi = ~0ull;
                    // largest T with count(last < key <= T) <= FOLD_CAP
                    while (tlo < thi)
                    {
                        const uint64_t d = thi - tlo;
                        const uint64_t mid = tlo + d / 2 + (d & 1ull);
                        uint32_t cnt = 0;
                        for_each_entry(a, brick, L.cb, x0, y0, z0, [&](uint64_t key, float, uint32_t lvox) {
                            cnt += (lvox == lv && (!have_last || key > last) && key <= mid) ? 1u : 0u;
                        });
                        uint32_t tot;
                        block_excl_scan(cnt, L.sh, tot);
                        if (tot <= (uint32_t)FOLD_CAP)
                            tlo = mid;
                        else
                            thi = mid - 1ull;
                    }
                    const uint64_t T = tlo;
                    if (tid == 0)
                        L.m = 0;
                    __syncthreads();
                    for_each_entry(a, brick, L.cb, x0, y0, z0, [&](uint64_t key, float z, uint32_t lvox) {
                        if (lvox == lv && (!have_last || key > last) && key <= T)
                        {
                            const uint32_t p = atomicAdd(&L.m, 1u);
                            L.key[p] = key;
                            L.z[p] = z;
                        }
                    });
                    const uint32_t m = L.m;
                    for (uint32_t i = tid; i < m; i += BT)
                    {
                        const uint64_t k = L.key[i];
                        uint32_t rank = 0;
                        for (uint32_t t = 0; t < m; ++t)
                            rank += (L.key[t] < k) ? 1u : 0u;
                        L.perm[rank] = (uint16_t)i;
                    }
                    __syncthreads();
                    if (tid == 0)
                        for (uint32_t i = 0; i < m; ++i)
                        {
                            if (done == 0 && i == 0)
                                st.begin(L.key[L.perm[0]]);
                            st.add(L.key[L.perm[i]], L.z[L.perm[i]], a);
                        }
                    __syncthreads();
                    last = T;
                    have_last = true;
                    done += m;
                    if (m == 0)
                        break; // cannot happen with unique keys; keeps the loop finite
                }
                if (tid == 0)
                {
                    st.finish_out(row_vid(a, L, x0, (int)lv), x0 + (int)(lv & 7), y0 + (int)((lv >> 3) & 7),
                                  z0 + (int)(lv >> 6), a);
                    atomicAdd(a.ctr + CTR_GIANT, 1ull);
                }
                __syncthreads();
                ++lo;
                continue;
            }
            // largest hi with off[hi] - off[lo] <= FOLD_CAP
            int l = lo + 1, h = 512;
            while (l < h)
            {
                const int mid = (l + h + 1) >> 1;
                if (L.off[mid] - L.off[lo] <= (uint32_t)FOLD_CAP)
                    l = mid;
                else
                    h = mid - 1;
            }
            const int hi = l;
            const uint32_t base = L.off[lo];
            const uint32_t nrec = L.off[hi] - base;
            for (int i = tid; i < 512; i += BT)
                L.cur[i] = 0;
            __syncthreads();
            // counting sort of the range's records by voxel
            for_each_entry(a, brick, L.cb, x0, y0, z0, [&](uint64_t key, float z, uint32_t lvox) {
                if (lvox >= (uint32_t)lo && lvox < (uint32_t)hi)
                {
                    const uint32_t p = L.off[lvox] - base + atomicAdd(&L.cur[lvox], 1u);
                    L.key[p] = key;
                    L.z[p] = z;
                }
            });
            // rank of each record among its voxel's keys (keys are unique)
            for (uint32_t i = tid; i < nrec; i += BT)
            {
                int vl = lo, vh = hi - 1;
                while (vl < vh)
                {
                    const int mid = (vl + vh + 1) >> 1;
                    if (L.off[mid] - base <= i)
                        vl = mid;
                    else
                        vh = mid - 1;
                }
                const uint32_t s0 = L.off[vl] - base, s1 = L.off[vl + 1] - base;
                const uint64_t k = L.key[i];
                uint32_t rank = 0;
                for (uint32_t t = s0; t < s1; ++t)
                    rank += (L.key[t] < k) ? 1u : 0u;
                L.perm[s0 + rank] = (uint16_t)i;
            }
            __syncthreads();
            for (int lv = lo + tid; lv < hi; lv += BT)
            {
                const uint32_t s0 = L.off[lv] - base, s1 = L.off[lv + 1] - base;
                if (s1 == s0)
                    continue;
                typename FoldPick<P>::type st;
                st.init(x0 + (lv & 7), y0 + ((lv >> 3) & 7), z0 + (lv >> 6), a);
                st.begin(L.key[L.perm[s0]]);
                for (uint32_t i = s0; i < s1; ++i)
                {
                    const uint32_t j = L.perm[i];
                    st.add(L.key[j], L.z[j], a);
                }
                st.finish_out(row_vid(a, L, x0, lv), x0 + (lv & 7), y0 + ((lv >> 3) & 7), z0 + (lv >> 6), a);
            }
            __syncthreads();
            lo = hi;
        }
        __syncthreads();
    }
}

// Dual vertex of every voxel from its voxel-frame quadric (precision
// variant): FP32 column-pivoted QR; when it finds the system rank-deficient
// or its condition estimate |R00 / R22| exceeds cond_max, the pivoting /
// rank decision and the back substitution are redone in FP64 on the
// compensated (h + l) quadric, in world coordinates like the reference.  A
// minimiser outside [0,1]^3 goes through the box search (FP32, voxel frame).
// The voxels that took the FP64 path are counted.
__global__ __launch_bounds__(BT) void k_solve_voxels_c(const float *qef, const uint32_t *solve_bits,
                                                       const int32_t *coords, uint64_t n, BatchDev bt, float *dual,
                                                       float cond_max, unsigned long long *n64,
                                                       const unsigned long long *over)
{
    if (*over)
        return;
    const int lane = threadIdx.x & 63;
    unsigned long long *cnt64 = n64 + (uint64_t)((blockIdx.x * NWAVE + (threadIdx.x >> 6)) % NSHARD) * SHARD_STRIDE;
    // each wave takes 64 consecutive words of the flags (2048 voxels) and its
    // lanes take the flagged voxels among them one each, in rounds of 64
    const uint64_t n_words = (n + 31ull) / 32ull;
    const uint64_t wave = ((uint64_t)blockIdx.x * BT + (uint64_t)threadIdx.x) >> 6;
    const uint64_t n_waves = ((uint64_t)gridDim.x * BT) >> 6;
    for (uint64_t w0 = wave * 64ull; w0 < n_words; w0 += n_waves * 64ull)
    for (uint32_t k0 = 0, total = 1; k0 < total; k0 += 64u)
    {
        const uint64_t wi = w0 + (uint64_t)lane;
        const uint32_t word = wi < n_words ? solve_bits[wi] : 0u;
        const uint32_t pc = (uint32_t)__popc(word);
        const uint32_t incl = wave_incl_scan(pc);
        total = __shfl(incl, 63, 64);
        const uint32_t k = k0 + (uint32_t)lane;
        // the word of flagged voxel k: lane j = number of lanes with incl <= k
        int jw = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1)
        {
            const uint32_t v = __shfl(incl, jw + step - 1, 64);
            jw += (v <= k) ? step : 0;
        }
        jw = jw < 63 ? jw : 63;
        const uint32_t mj = __shfl(word, jw, 64);
        const uint32_t ej = __shfl(incl - pc, jw, 64);
        const bool live = k < total;
        const uint64_t vid = live ? (w0 + (uint64_t)jw) * 32ull + (uint64_t)select_bit((uint64_t)mj, k - ej) : 0ull;
        bool used64 = false;
        if (live)
        {
            const GridDev &g = bt.seg[mesh_of_voxel(bt, vid)].g; // the voxel's mesh
            float qh[10], ql[10];
#pragma unroll
            for (int i = 0; i < 10; ++i)
            {
                qh[i] = qef[20 * vid + i];
                ql[i] = qef[20 * vid + 10 + i];
            }
            float Q[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    Q[i][j] = qh[qidx(i, j)] + ql[qidx(i, j)];
            const float A[3][3] = {{Q[0][0], Q[0][1], Q[0][2]}, {Q[1][0], Q[1][1], Q[1][2]}, {Q[2][0], Q[2][1], Q[2][2]}};
            const float b[3] = {-Q[0][3], -Q[1][3], -Q[2][3]};
            float u[3];
            Cpqr<3> f;
            f.factor(A);
            const float r0 = fabsf(f.qr[0][0]), rl = fabsf(f.qr[2][2]);
            used64 = f.rk < 3 || !(r0 <= cond_max * rl);
            if (used64)
            {
                // FP64 solve in world coordinates x = c + vs u, so that a
                // rank-deficient system gets the reference's basic solution
                // (free world coordinates zero): A_w = S^-1 A_u S^-1,
                // b_w_i = -(Q_u[i][3] - sum_j A_u[i][j] c_j / s_j) / s_i
                double Qd[4][4], cd[3], sd[3], Ad[3][3], bd[3], xd[3];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        Qd[i][j] = (double)qh[qidx(i, j)] + (double)ql[qidx(i, j)];
#pragma unroll
                for (int i = 0; i < 3; ++i)
                {
                    sd[i] = (double)g.vs[i];
                    cd[i] = (double)((float)coords[3 * vid + i] * g.vs[i]);
                }
#pragma unroll
                for (int i = 0; i < 3; ++i)
                {
                    double gi = Qd[i][3];
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                    {
                        Ad[i][j] = Qd[i][j] / (sd[i] * sd[j]);
                        gi = gi - Qd[i][j] * (cd[j] / sd[j]);
                    }
                    bd[i] = -(gi / sd[i]);
                }
                Cpqr<3, double> fd;
                fd.factor(Ad);
                fd.solve(bd, xd);
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    u[i] = (float)((xd[i] - cd[i]) / sd[i]);
            }
            else
                f.solve(b, u);
            const float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {1.0f, 1.0f, 1.0f};
            if (!(u[0] >= 0.0f && u[0] <= 1.0f && u[1] >= 0.0f && u[1] <= 1.0f && u[2] >= 0.0f && u[2] <= 1.0f))
                box_search(Q, mn, mx, u);
            float *o = dual + 3 * vid;
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                const float c = (float)coords[3 * vid + i] * g.vs[i];
                o[i] = c + g.vs[i] * u[i];
            }
        }
        const uint64_t m = __ballot(used64);
        if (m && lane == __ffsll((long long)m) - 1)
            atomicAdd(cnt64, (unsigned long long)__popcll(m));
    }
}

// box-constrained search for the queued voxels, one per lane
__global__ __launch_bounds__(BT) void k_solve_box(ShardBuf<BoxItem> box, BatchDev bt, float *dual,
                                                  const unsigned long long *over)
{
    __shared__ QueueMap qm;
    if (*over)
        return;
    const uint64_t n = queue_prefix(box, qm.pre, qm.sh);
    const uint64_t stride = (uint64_t)gridDim.x * BT;
    for (uint64_t qi = (uint64_t)blockIdx.x * BT + threadIdx.x; qi < n; qi += stride)
    {
        const BoxItem e = box.items[queue_slot(box, qm.pre, qi)];
        const GridDev &g = bt.seg[mesh_of_voxel(bt, e.vid)].g; // the voxel's mesh
        float v[3] = {e.v[0], e.v[1], e.v[2]};
        solve_box(e.q, e.c, g.vs, v);
        float *o = dual + 3 * e.vid;
        o[0] = v[0];
        o[1] = v[1];
        o[2] = v[2];
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t n, int64_t cap = 1 << 20)
{
    int64_t b = (n + BT - 1) / BT;
    if (b < 1)
        b = 1;
    if (b > cap)
        b = cap;
    return (unsigned)b;
}

hipError_t launch_prep(const MeshDev &m, const BatchDev &b, float4 *planes, float *tri, unsigned long long *ctr,
                       unsigned long long *est_ctr, hipStream_t s)
{
    const int64_t n = m.F > (m.f ? m.V : 3 * m.F) ? m.F : (m.f ? m.V : 3 * m.F);
    MeshDev mi = m;
    mi.tri = nullptr; // the prep kernel writes the gathered triangles
    hipLaunchKernelGGL(k_prep, dim3(grid_for(n)), dim3(BT), 0, s, mi, b, planes, tri, ctr, est_ctr);
    return hipGetLastError();
}

template <bool MERGE, int TILE>
static void scan_pair(const MeshDev &m, const BatchDev &b, unsigned gm, uint64_t n_tiles, unsigned long long *ctr,
                      const ShardBuf<Event> &eb, const ShardBuf<RowChunk> &rq, hipStream_t s)
{
    hipLaunchKernelGGL((k_scan_events<MERGE, false, TILE>), dim3(gm), dim3(BT), 0, s, m, b, n_tiles, ctr, eb, rq);
    // consumers loop over the filled slots of the row-chunk queue
    hipLaunchKernelGGL((k_scan_events<MERGE, true, TILE>), dim3(2048), dim3(BT), 0, s, m, b, 0ull, ctr, eb, rq);
}

hipError_t launch_scan_events(const MeshDev &m, const BatchDev &b, int tile, uint64_t n_tiles, unsigned long long *ctr,
                              const ShardBuf<Event> &eb, const ShardBuf<RowChunk> &rq, bool merge, hipStream_t s)
{
    const unsigned gm = (unsigned)(n_tiles < (1ull << 20) ? (n_tiles ? n_tiles : 1) : (1ull << 20));
    // row chunks of large triangles: queue emptied, filled by the first
    // launch, consumed by the second
    hipError_t e = hipMemsetAsync(rq.cnt, 0, (size_t)NSHARD * SHARD_STRIDE * sizeof(unsigned long long), s);
    if (e != hipSuccess)
        return e;
    static_assert(MARK_TILES[0] == 64 && MARK_TILES[1] == 128 && MARK_TILES[2] == 256, "scanline tile instantiations");
    if (merge)
    {
        if (tile == 0)
            scan_pair<true, 64>(m, b, gm, n_tiles, ctr, eb, rq, s);
        else if (tile == 1)
            scan_pair<true, 128>(m, b, gm, n_tiles, ctr, eb, rq, s);
        else
            scan_pair<true, 256>(m, b, gm, n_tiles, ctr, eb, rq, s);
    }
    else
    {
        if (tile == 0)
            scan_pair<false, 64>(m, b, gm, n_tiles, ctr, eb, rq, s);
        else if (tile == 1)
            scan_pair<false, 128>(m, b, gm, n_tiles, ctr, eb, rq, s);
        else
            scan_pair<false, 256>(m, b, gm, n_tiles, ctr, eb, rq, s);
    }
    return hipGetLastError();
}

template <int TILE>
static void face_pair(const MeshDev &m, const BatchDev &b, unsigned gf, uint64_t n_tiles, const ShardBuf<FHit> &hb,
                      unsigned long long *ctr, const ShardBuf<FaceChunk> &fq, hipStream_t s)
{
    hipLaunchKernelGGL((k_face<false, TILE>), dim3(gf), dim3(BT), 0, s, m, b, n_tiles, hb, ctr, fq);
    // loops over the filled slots of the row-chunk queue
    hipLaunchKernelGGL((k_face<true, TILE>), dim3(2048), dim3(BT), 0, s, m, b, 0ull, hb, ctr, fq);
}

hipError_t launch_face(const MeshDev &m, const BatchDev &b, int tile, uint64_t n_tiles, const ShardBuf<FHit> &hb,
                       unsigned long long *ctr, const ShardBuf<FaceChunk> &fq, hipStream_t s)
{
    const unsigned gf = (unsigned)(n_tiles < (1ull << 20) ? (n_tiles ? n_tiles : 1) : (1ull << 20));
    // row chunks of large faces: queue emptied, filled by the first launch,
    // consumed by the second
    hipError_t e = launch_fill_u64(fq.cnt, (uint64_t)NSHARD * SHARD_STRIDE, 0ull, s);
    if (e != hipSuccess)
        return e;
    static_assert(FACE_TILES[0] == 32 && FACE_TILES[1] == 64 && FACE_TILES[FACE_LANE_CLASS] == BT,
                  "face tile instantiations");
    if (tile == 0)
        face_pair<32>(m, b, gf, n_tiles, hb, ctr, fq, s);
    else if (tile == 1)
        face_pair<64>(m, b, gf, n_tiles, hb, ctr, fq, s);
    else
    {
        hipLaunchKernelGGL(k_face_lane, dim3(gf), dim3(BT), 0, s, m, b, n_tiles, hb, fq);
        // the queued faces of many voxels, block-cooperatively
        hipLaunchKernelGGL((k_face<true, 64>), dim3(2048), dim3(BT), 0, s, m, b, 0ull, hb, ctr, fq);
    }
    return hipGetLastError();
}

hipError_t launch_gen_fhits(const ShardBuf<FHit> &h, const BucketArgs &ba, hipStream_t s)
{
    hipLaunchKernelGGL(k_gen_fhits, dim3(4096), dim3(BT), 0, s, h, ba); // loops over the hits
    return hipGetLastError();
}

hipError_t launch_gen_events(const ShardBuf<Event> &ev, const BucketArgs &ba, hipStream_t s)
{
    hipLaunchKernelGGL(k_gen_events, dim3(4096), dim3(BT), 0, s, ev, ba); // loops over the crossings
    return hipGetLastError();
}

__global__ __launch_bounds__(BT) void k_fill64(uint64_t *p, uint64_t n, uint64_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * BT;
    for (uint64_t i = (uint64_t)blockIdx.x * BT + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

__global__ __launch_bounds__(BT) void k_fill32(uint32_t *p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * BT;
    for (uint64_t i = (uint64_t)blockIdx.x * BT + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

// A conversion's inputs from the caller's arrays, in one pass: grid-local
// vertices (each coordinate minus the aabb's min corner, one FP32
// subtraction) and, when the faces are 64-bit, their indices as 32-bit ones;
// an index outside [0, 2^31) becomes -1, which the conversion rejects as a
// face index out of range.  nvf = 3 x vertices floats, nf3 = 3 x faces
// indices; with vec, 16-byte loads (4 floats, 2 indices) and the few
// elements past the last whole vector one at a time.
__device__ __forceinline__ int32_t face_index_i32(int64_t x)
{
    return (x >= 0 && x <= (int64_t)INT32_MAX) ? (int32_t)x : -1;
}

template <bool VEC>
__global__ __launch_bounds__(BT) void k_prepare_mesh(const float *v, const float *origin, uint64_t nvf, float *vout,
                                                     const int64_t *f64, uint64_t nf3, int32_t *f32)
{
    // float k of the vertex array is coordinate k % 3 of its vertex
    const float o0 = origin[0], o1 = origin[1], o2 = origin[2];
    const uint64_t stride = (uint64_t)gridDim.x * BT;
    const uint64_t t0 = (uint64_t)blockIdx.x * BT + threadIdx.x;
    uint64_t vdone = 0, fdone = 0;
    if (VEC)
    {
        const uint64_t n4 = nvf / 4, n2 = nf3 / 2;
        for (uint64_t i = t0; i < n4; i += stride)
        {
            const float4 x = reinterpret_cast<const float4 *>(v)[i];
            const int b = (int)((4 * i) % 3); // coordinate of the vector's first float
            const float p0 = sel3f(b, o0, o1, o2);
            const float p1 = sel3f(b == 2 ? 0 : b + 1, o0, o1, o2);
            const float p2 = sel3f(b == 0 ? 2 : b - 1, o0, o1, o2);
            float4 y;
            y.x = x.x - p0;
            y.y = x.y - p1;
            y.z = x.z - p2;
            y.w = x.w - p0;
            reinterpret_cast<float4 *>(vout)[i] = y;
        }
        for (uint64_t i = t0; i < n2; i += stride)
        {
            const longlong2 x = reinterpret_cast<const longlong2 *>(f64)[i];
            reinterpret_cast<int2 *>(f32)[i] = make_int2(face_index_i32(x.x), face_index_i32(x.y));
        }
        vdone = 4 * n4;
        fdone = 2 * n2;
    }
    for (uint64_t k = vdone + t0; k < nvf; k += stride)
        vout[k] = v[k] - sel3f((int)(k % 3), o0, o1, o2);
    for (uint64_t j = fdone + t0; j < nf3; j += stride)
        f32[j] = face_index_i32(f64[j]);
}

hipError_t launch_prepare_mesh(const float *v, const float *origin, uint64_t nv, float *vout, const int64_t *f64,
                               uint64_t nf3, int32_t *f32, hipStream_t s)
{
    const uint64_t nvf = 3 * nv;
    const uint64_t n = nvf > nf3 ? nvf : nf3;
    const bool vec = ((uintptr_t)v % 16 == 0) && ((uintptr_t)vout % 16 == 0) && ((uintptr_t)f64 % 16 == 0) &&
                     ((uintptr_t)f32 % 8 == 0);
    const unsigned grid = grid_for((int64_t)(vec ? n / 2 + 1 : n), 8192);
    if (vec)
        hipLaunchKernelGGL(k_prepare_mesh<true>, dim3(grid), dim3(BT), 0, s, v, origin, nvf, vout, f64, nf3, f32);
    else
        hipLaunchKernelGGL(k_prepare_mesh<false>, dim3(grid), dim3(BT), 0, s, v, origin, nvf, vout, f64, nf3, f32);
    return hipGetLastError();
}

// Packed inputs of a batched conversion from each mesh's own arrays (the
// concatenation, the shift by each mesh's origin and the rebase of the face
// indices the batch wrapper needs) in one pass.  Block b writes the packed
// vertex floats and the packed face indices [b * PACK_CH, (b + 1) * PACK_CH):
// it finds the mesh of its first element by a binary search over the meshes'
// first rows and walks on across mesh boundaries.  A face index outside
// [0, the mesh's vertex count) becomes -1, which the conversion rejects as
// out of range.  Writes stay inside the packed outputs whatever the table
// holds; reads stay inside each mesh's arrays given its counts.
constexpr int PACK_CH = 4096;

template <int FACES>
__device__ __forceinline__ int pack_first_mesh(const MeshSrc *t, int M, uint64_t e0)
{
    int lo = 0, hi = M - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (3ull * (uint64_t)(FACES ? t[mid].face_offset : t[mid].vertex_offset) <= e0)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <typename I>
__device__ __forceinline__ void pack_faces(const I *f, uint64_t mb, uint64_t lo, uint64_t hi, int64_t nv, int64_t base,
                                           int32_t *fout)
{
    for (uint64_t k = lo + threadIdx.x; k < hi; k += BT)
    {
        const int64_t x = (int64_t)f[k - mb];
        fout[k] = (x >= 0 && x < nv && x + base <= (int64_t)INT32_MAX) ? (int32_t)(x + base) : -1;
    }
}

__global__ __launch_bounds__(BT) void k_batch_prepare(const MeshSrc *t, int M, uint64_t nvf, uint64_t nf3, float *vout,
                                                      int32_t *fout)
{
    const uint64_t e0 = (uint64_t)blockIdx.x * PACK_CH;
    if (e0 < nvf)
    {
        const uint64_t e1 = e0 + PACK_CH < nvf ? e0 + PACK_CH : nvf;
        for (int m = pack_first_mesh<0>(t, M, e0); m < M; ++m)
        {
            const MeshSrc s = t[m];
            const uint64_t mb = 3ull * (uint64_t)s.vertex_offset, me = mb + 3ull * (uint64_t)s.n_vertices;
            if (mb >= e1)
                break;
            const uint64_t lo = e0 > mb ? e0 : mb, hi = e1 < me ? e1 : me;
            if (hi <= lo)
                continue;
            const float o0 = s.origin[0], o1 = s.origin[1], o2 = s.origin[2];
            for (uint64_t k = lo + threadIdx.x; k < hi; k += BT)
            {
                const uint64_t i = k - mb;
                vout[k] = s.vertices[i] - sel3f((int)(i % 3), o0, o1, o2);
            }
        }
    }
    if (e0 < nf3)
    {
        const uint64_t e1 = e0 + PACK_CH < nf3 ? e0 + PACK_CH : nf3;
        for (int m = pack_first_mesh<1>(t, M, e0); m < M; ++m)
        {
            const MeshSrc s = t[m];
            const uint64_t mb = 3ull * (uint64_t)s.face_offset, me = mb + 3ull * (uint64_t)s.n_faces;
            if (mb >= e1)
                break;
            const uint64_t lo = e0 > mb ? e0 : mb, hi = e1 < me ? e1 : me;
            if (hi <= lo)
                continue;
            if (s.faces_int64)
                pack_faces((const int64_t *)s.faces, mb, lo, hi, s.n_vertices, s.vertex_offset, fout);
            else
                pack_faces((const int32_t *)s.faces, mb, lo, hi, s.n_vertices, s.vertex_offset, fout);
        }
    }
}

hipError_t launch_batch_prepare(const MeshSrc *t, int M, uint64_t nv, uint64_t nf, float *vout, int32_t *fout,
                                hipStream_t s)
{
    const uint64_t nvf = 3 * nv, nf3 = 3 * nf;
    const uint64_t n = nvf > nf3 ? nvf : nf3;
    if (n == 0 || M <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_batch_prepare, dim3((unsigned)((n + PACK_CH - 1) / PACK_CH)), dim3(BT), 0, s, t, M, nvf, nf3,
                       vout, fout);
    return hipGetLastError();
}

hipError_t launch_fill_u64(void *p, uint64_t n, uint64_t v, hipStream_t s)
{
    if (n)
        hipLaunchKernelGGL(k_fill64, dim3(grid_for((int64_t)n, 4096)), dim3(BT), 0, s, (uint64_t *)p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u32(void *p, uint64_t n, uint32_t v, hipStream_t s)
{
    if (n)
        hipLaunchKernelGGL(k_fill32, dim3(grid_for((int64_t)n, 4096)), dim3(BT), 0, s, (uint32_t *)p, n, v);
    return hipGetLastError();
}

__global__ void k_over_check(unsigned long long *ctr, OverCaps caps)
{
    const unsigned long long *mx = ctr + CTR_MAX;
    const bool over = mx[SB_FHITS] > caps.fhits || mx[SB_BHITS] > caps.bhits || mx[SB_RECS] > caps.recs ||
                      mx[SB_CHUNKS] > caps.chunks || ctr[CTR_SUM + SB_CHUNKS] > caps.grouped;
    ctr[CTR_OVER] = over ? 1ull : 0ull;
}

hipError_t launch_over_check(unsigned long long *ctr, const OverCaps &caps, hipStream_t s)
{
    hipLaunchKernelGGL(k_over_check, dim3(1), dim3(1), 0, s, ctr, caps);
    return hipGetLastError();
}

hipError_t launch_shard_reduce(unsigned long long *shard_ctr, unsigned long long *ctr, hipStream_t s)
{
    static_assert(NSHARD == BT, "one thread per shard");
    hipLaunchKernelGGL(k_shard_reduce, dim3(SB_N), dim3(BT), 0, s, shard_ctr, ctr);
    return hipGetLastError();
}

hipError_t launch_place_chunks(const ShardBuf<Chunk> &chunks, const uint32_t *choff, uint2 *gchunks,
                               const unsigned long long *over, hipStream_t s)
{
    hipLaunchKernelGGL(k_place_chunks, dim3(NSHARD * PLACE_SUB), dim3(BT), 0, s, chunks, choff, gchunks,
                       (const unsigned long long *)over); // loops over the chunks
    return hipGetLastError();
}

hipError_t launch_active_bricks(const unsigned long long *cnt, uint64_t n_bricks, TierLists t,
                                const unsigned long long *over, hipStream_t s)
{
    hipLaunchKernelGGL(k_active_bricks, dim3(grid_for((int64_t)((n_bricks + ACT_PER - 1) / ACT_PER), 65536)), dim3(BT),
                       0, s, cnt, n_bricks, t, over);
    return hipGetLastError();
}

hipError_t launch_fold_bricks(const FoldArgs &a, uint64_t max_bricks, hipStream_t s)
{
    uint64_t nt = max_bricks < 262144ull ? max_bricks : 262144ull;
    if (nt == 0)
        nt = 1;
    if (a.prec)
        hipLaunchKernelGGL(k_fold_tiny<1>, dim3((unsigned)nt), dim3(FOLD_TINY_NT), 0, s, a);
    else
        hipLaunchKernelGGL(k_fold_tiny<0>, dim3((unsigned)nt), dim3(FOLD_TINY_NT), 0, s, a);
    uint64_t ns = max_bricks < 131072ull ? max_bricks : 131072ull;
    if (ns == 0)
        ns = 1;
    if (a.prec)
        hipLaunchKernelGGL(k_fold_small<1>, dim3((unsigned)ns), dim3(FOLD_SMALL_NT), 0, s, a);
    else
        hipLaunchKernelGGL(k_fold_small<0>, dim3((unsigned)ns), dim3(FOLD_SMALL_NT), 0, s, a);
    uint64_t nb = max_bricks < 65536ull ? max_bricks : 65536ull;
    if (nb == 0)
        nb = 1;
    if (a.prec)
        hipLaunchKernelGGL(k_fold_bricks<1>, dim3((unsigned)nb), dim3(BT), 0, s, a);
    else
        hipLaunchKernelGGL(k_fold_bricks_ref, dim3((unsigned)nb), dim3(BT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fold_mid(const FoldArgs &a, uint64_t max_bricks, hipStream_t s)
{
    uint64_t nb = max_bricks < 4096ull ? max_bricks : 4096ull;
    if (nb == 0)
        nb = 1;
    if (a.prec)
        hipLaunchKernelGGL(k_fold_mid<1>, dim3((unsigned)nb), dim3(BT), 0, s, a);
    else
        hipLaunchKernelGGL(k_fold_mid_ref, dim3((unsigned)nb), dim3(FOLD_MID_REF_NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_solve_voxels(const FoldArgs &a, const int32_t *coords, uint64_t n, hipStream_t s)
{
    // precision variant only: the reference path's folds solve each voxel where it is folded
    hipLaunchKernelGGL(k_solve_voxels_c, dim3(grid_for((int64_t)((n + 31) / 32), 4096)), dim3(BT), 0, s,
                       (const float *)a.qef, (const uint32_t *)a.solve_bits, coords, n, a.b, a.dual, a.cond_max, a.n64,
                       (const unsigned long long *)(a.ctr + CTR_OVER));
    return hipGetLastError();
}

hipError_t launch_solve_box(const FoldArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_solve_box, dim3(2048), dim3(BT), 0, s, a.box, a.b, a.dual,
                       (const unsigned long long *)(a.ctr + CTR_OVER)); // loops over the queue
    return hipGetLastError();
}

hipError_t launch_fold_large(const FoldArgs &a, uint64_t max_bricks, hipStream_t s)
{
    // the resident blocks (2 per CU) take the bricks one by one
    uint64_t nb = max_bricks < 512ull ? max_bricks : 512ull;
    if (nb == 0)
        nb = 1;
    if (a.prec)
        hipLaunchKernelGGL(k_fold_large<1>, dim3((unsigned)nb), dim3(BT), 0, s, a);
    else
        hipLaunchKernelGGL(k_fold_large<0>, dim3((unsigned)nb), dim3(BT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_edge_count(const MeshDev &m, uint32_t *vcnt, uint32_t *rank, const uint32_t *voff, uint32_t *nbr,
                             int pass, hipStream_t s)
{
    hipLaunchKernelGGL(k_edge_count, dim3(grid_for(3 * m.F, 65536)), dim3(BT), 0, s, m, vcnt, rank, voff, nbr, pass);
    return hipGetLastError();
}

hipError_t launch_ebin(const MeshDev &m, bool place, uint32_t *bin_cnt, const uint32_t *bin_off, uint2 *he,
                       hipStream_t s)
{
    if (m.F <= 0)
        return hipSuccess;
    const int64_t per = (int64_t)BT * EBIN_FPT;
    const unsigned grid = (unsigned)((m.F + per - 1) / per);
    if (place)
        hipLaunchKernelGGL(k_ebin<true>, dim3(grid), dim3(BT), 0, s, m, bin_cnt, bin_off, he);
    else
        hipLaunchKernelGGL(k_ebin<false>, dim3(grid), dim3(BT), 0, s, m, bin_cnt, bin_off, he);
    return hipGetLastError();
}

hipError_t launch_ebin_fold(uint64_t n_bins, const uint32_t *bin_off, uint2 *he, uint32_t *bin_bcnt,
                            unsigned long long *over, hipStream_t s)
{
    if (n_bins == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_ebin_fold, dim3((unsigned)n_bins), dim3(BT), 0, s, bin_off, he, bin_bcnt, over);
    return hipGetLastError();
}

hipError_t launch_ebin_gather(uint64_t n_bins, const uint32_t *bin_off, const uint32_t *bin_boff, const uint2 *he,
                              int2 *bedge_v, hipStream_t s)
{
    if (n_bins == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_ebin_gather, dim3(grid_for((int64_t)((n_bins + NWAVE - 1) / NWAVE) * BT, 65536)), dim3(BT), 0,
                       s, n_bins, bin_off, bin_boff, he, bedge_v);
    return hipGetLastError();
}

hipError_t launch_edge_singletons(int64_t V, const uint32_t *voff, uint32_t *nbr, uint32_t *bcnt, const uint32_t *boff,
                                  int2 *bedge_v, int pass, hipStream_t s)
{
    hipLaunchKernelGGL(k_edge_singletons, dim3(grid_for(V, 65536)), dim3(BT), 0, s, V, voff, nbr, bcnt, boff, bedge_v,
                       pass);
    return hipGetLastError();
}

hipError_t launch_bedge_prep(const float *verts, int64_t E, const int2 *bedge_v, BEdge *be, hipStream_t s)
{
    hipLaunchKernelGGL(k_bedge_prep, dim3(grid_for(E)), dim3(BT), 0, s, verts, E, bedge_v, be);
    return hipGetLastError();
}

hipError_t launch_bdry(int64_t E, const BEdge *be, const int2 *bedge_v, const BatchDev &b, const BucketArgs &ba,
                       const ShardBuf<BHit> &hcount, unsigned long long *ctr, hipStream_t s)
{
    if (E == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_bdry, dim3(grid_for(E, 65536)), dim3(BT), 0, s, E, be, bedge_v, b, ba, hcount, ctr);
    return hipGetLastError();
}

struct LoadPopc
{
    const uint64_t *bits;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return (uint32_t)__popcll(bits[i]); }
};
struct LoadU32
{
    const uint32_t *a;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return a[i]; }
};
struct LoadLo32
{
    const unsigned long long *a;
    __device__ __forceinline__ uint32_t operator()(uint64_t i) const { return (uint32_t)a[i]; }
};

template <class Load>
static hipError_t scan_impl(Load ld, uint64_t n, uint32_t *out, uint64_t *tile_sums, unsigned long long *total,
                            hipStream_t s)
{
    const uint64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nt == 0)
        return hipMemsetAsync(total, 0, sizeof(unsigned long long), s);
    hipLaunchKernelGGL(k_scan_reduce<Load>, dim3((unsigned)nt), dim3(BT), 0, s, ld, n, tile_sums);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(BT), 0, s, tile_sums, nt, total);
    hipLaunchKernelGGL(k_scan_down<Load>, dim3((unsigned)nt), dim3(BT), 0, s, ld, n, (const uint64_t *)tile_sums, out);
    return hipGetLastError();
}

uint64_t scan_tile_count(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

hipError_t launch_scan_popc(const uint64_t *bits, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                            unsigned long long *total, hipStream_t s)
{
    return scan_impl(LoadPopc{bits}, n, out, tile_sums, total, s);
}

hipError_t launch_scan_lo32(const unsigned long long *in, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                            unsigned long long *total, hipStream_t s)
{
    return scan_impl(LoadLo32{in}, n, out, tile_sums, total, s);
}

hipError_t launch_scan_u32(const uint32_t *in, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                           unsigned long long *total, hipStream_t s)
{
    return scan_impl(LoadU32{in}, n, out, tile_sums, total, s);
}

hipError_t launch_emit_coords(const BatchDev &b, uint64_t n_words, int32_t *coords, hipStream_t s)
{
    hipLaunchKernelGGL(k_emit_coords, dim3(grid_for((int64_t)n_words, 65536)), dim3(BT), 0, s, b, n_words, coords);
    return hipGetLastError();
}

hipError_t launch_seg_voxels(const BatchDev &b, const uint32_t *woff, uint64_t n_words,
                             const unsigned long long *n_voxels, uint64_t *vox_off, hipStream_t s)
{
    hipLaunchKernelGGL(k_seg_voxels, dim3(grid_for((int64_t)b.M + 1)), dim3(BT), 0, s, b, woff, n_words, n_voxels,
                       vox_off);
    return hipGetLastError();
}

} // namespace ovx
