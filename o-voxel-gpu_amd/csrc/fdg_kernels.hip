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

// Exclusive block scan; `sh` holds NWAVE elements of LDS.  Contains barriers:
// every thread of the block must call it.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *sh, T &total)
{
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const T inc = wave_incl_scan(v);
    if (lane == 63)
        sh[wid] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w)
    {
        const T t = sh[w];
        pre += (w < wid) ? t : T(0);
        tot += t;
    }
    __syncthreads();
    total = tot;
    return pre + inc - v;
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

// count / emit one record for voxel `vid`
template <int MODE>
__device__ __forceinline__ void record(uint32_t vid, uint64_t key, float z, const RecArgs &r)
{
    if (MODE == MODE_COUNT)
    {
        atomicAdd(r.cnt + vid, 1u);
    }
    else
    {
        // cnt[vid] runs back down from the COUNT pass; a zero here would mean
        // the two passes disagree -- never write outside the segment then.
        const uint32_t old = atomicSub(r.cnt + vid, 1u);
        if (old == 0u)
        {
            atomicAdd(r.cnt + vid, 1u);
            atomicOr(r.err, (unsigned long long)ERR_RECORDS);
            return;
        }
        const uint32_t slot = r.rec_off[vid] + old - 1u;
        r.keys[slot] = key;
        r.zs[slot] = z;
    }
}

// ---------------------------------------------------------------------------
// prep: validate inputs, compute per-face planes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BT) void k_prep(MeshDev m, int64_t V, GridDev g, float4 *planes,
                                             unsigned long long *counters)
{
    const int64_t stride = (int64_t)gridDim.x * BT;
    unsigned flags = 0;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < m.F; i += stride)
    {
        if (m.f)
        {
            const int32_t *fi = m.f + 3 * i;
            const int32_t a = fi[0], b = fi[1], c = fi[2];
            if (a < 0 || b < 0 || c < 0 || a >= V || b >= V || c >= V)
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
        }
    }
    // vertices: finite and |v / vs| < 2^30 (so every int conversion of the
    // scanline and DDA is in range)
    const int64_t nv = m.f ? V : 3 * m.F;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < nv; i += stride)
    {
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            const float x = m.v[3 * i + a];
            if (!(fabsf(x) <= 3.0e38f))
                flags |= ERR_NONFINITE;
            else if (!(fabs((double)x / (double)g.vs[a]) < 1073741824.0))
                flags |= ERR_RANGE;
        }
    }
    if (flags)
        atomicOr(counters + CTR_ERRORS, (unsigned long long)flags);
}

// ---------------------------------------------------------------------------
// scanline events (FDG.cpp:195-283)
// ---------------------------------------------------------------------------
struct ScanItem
{
    float t[3][3]; // y-sorted vertices in (ax0, ax1, ax2) coordinates
    int start, mid, end;
    int ax2;
    uint32_t face;
};

struct ScanRow
{
    double t3x, t3z, t4x, t4z;
    int y_idx;
    int line_start;
    int item;
};

template <int MODE>
__device__ __forceinline__ void event_action(const GridDev &g, int ax0, int ax1, int ax2, int x_idx, int y_idx,
                                             int z_idx, float zf, uint64_t keybase, const RecArgs &r)
{
    int c[4][3];
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
    if (MODE == MODE_MARK)
    {
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
                atomicOr((unsigned long long *)(g.bits + w[k]), (unsigned long long)msk[k]);
    }
    else
    {
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            int b;
            const uint64_t w = word_of(g, c[k][0], c[k][1], c[k][2], b);
            const uint32_t vid = voxel_id(g, w, b);
            record<MODE>(vid, keybase | (uint64_t)k, zf, r);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(BT) void k_scan_events(MeshDev m, GridDev g, RecArgs r, unsigned long long *counters)
{
    __shared__ ScanItem items[BT];
    __shared__ ScanRow rows[BT];
    __shared__ uint32_t incl_items[BT];
    __shared__ uint32_t incl_rows[BT];
    __shared__ uint32_t sh[NWAVE];
    const int tid = threadIdx.x;
    const int64_t n_items = 3 * m.F;
    uint32_t n_ev = 0;

    for (int64_t base = (int64_t)blockIdx.x * BT; base < n_items; base += (int64_t)gridDim.x * BT)
    {
        const int64_t item = base + tid;
        uint32_t R = 0;
        if (item < n_items)
        {
            const int64_t face = item / 3;
            const int ax2 = (int)(item - face * 3);
            const int ax0 = (ax2 + 1) % 3;
            const int ax1 = (ax2 + 2) % 3;
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
            const double vs1 = (double)g.vs[ax1];
            const int lo = g.gmin[ax1], hi = g.gmax[ax1] - 1;
            const int s = iclamp((int)((double)t[0][1] / vs1), lo, hi);
            const int mi = iclamp((int)((double)t[1][1] / vs1), lo, hi);
            const int e = iclamp((int)((double)t[2][1] / vs1), lo, hi);
            ScanItem &it = items[tid];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    it.t[k][i] = t[k][i];
            it.start = s;
            it.mid = mi;
            it.end = e;
            it.ax2 = ax2;
            it.face = (uint32_t)face;
            R = (uint32_t)(e - s);
        }
        uint32_t tot_r;
        const uint32_t ex_r = block_excl_scan(R, sh, tot_r);
        incl_items[tid] = ex_r + R;
        __syncthreads();

        for (uint32_t rb = 0; rb < tot_r; rb += BT)
        {
            const uint32_t gi = rb + tid;
            uint32_t C = 0;
            if (gi < tot_r)
            {
                const int L = upper_find(incl_items, BT, gi);
                const ScanItem &it = items[L];
                const uint32_t R_L = (uint32_t)(it.end - it.start);
                const int y_idx = it.start + (int)(gi - (incl_items[L] - R_L));
                const int ax2 = it.ax2;
                const int ax0 = (ax2 + 1) % 3;
                const int ax1 = (ax2 + 2) % 3;
                const bool first = y_idx < it.mid;
                const int i0 = first ? 0 : 2, i2 = first ? 2 : 0;
                const double t0x = it.t[i0][0], t0y = it.t[i0][1], t0z = it.t[i0][2];
                const double t1x = it.t[1][0], t1y = it.t[1][1], t1z = it.t[1][2];
                const double t2x = it.t[i2][0], t2y = it.t[i2][1], t2z = it.t[i2][2];
                const float vs1 = sel3f(ax1, g.vs[0], g.vs[1], g.vs[2]);
                const float vs0 = sel3f(ax0, g.vs[0], g.vs[1], g.vs[2]);
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
                const int ls = iclamp((int)(t3x / (double)vs0), lo, hi);
                const int le = iclamp((int)(t4x / (double)vs0), lo, hi);
                ScanRow &rw = rows[tid];
                rw.t3x = t3x;
                rw.t3z = t3z;
                rw.t4x = t4x;
                rw.t4z = t4z;
                rw.y_idx = y_idx;
                rw.line_start = ls;
                rw.item = L;
                C = (uint32_t)(le - ls);
            }
            uint32_t tot_c;
            const uint32_t ex_c = block_excl_scan(C, sh, tot_c);
            incl_rows[tid] = ex_c + C;
            __syncthreads();
            for (uint32_t cb = 0; cb < tot_c; cb += BT)
            {
                const uint32_t h = cb + tid;
                if (h < tot_c)
                {
                    const int rr = upper_find(incl_rows, BT, h);
                    const ScanRow &rw = rows[rr];
                    const ScanItem &it = items[rw.item];
                    const uint32_t ex = (rr == 0) ? 0u : incl_rows[rr - 1];
                    const int x_idx = rw.line_start + (int)(h - ex);
                    const int ax2 = it.ax2;
                    const int ax0 = (ax2 + 1) % 3;
                    const int ax1 = (ax2 + 2) % 3;
                    const float vs0 = sel3f(ax0, g.vs[0], g.vs[1], g.vs[2]);
                    const float vs2 = sel3f(ax2, g.vs[0], g.vs[1], g.vs[2]);
                    const double x = (double)((float)(x_idx + 1) * vs0);
                    const double z = lerp1(rw.t3x, rw.t4x, x, rw.t3z, rw.t4z);
                    const int z_idx = (int)(z / (double)vs2);
                    const int zlo = sel3i(ax2, g.gmin[0], g.gmin[1], g.gmin[2]);
                    const int zhi = sel3i(ax2, g.gmax[0], g.gmax[1], g.gmax[2]);
                    if (z_idx >= zlo && z_idx < zhi)
                    {
                        ++n_ev;
                        const int ylo = sel3i(ax1, g.gmin[0], g.gmin[1], g.gmin[2]);
                        const int xlo = sel3i(ax0, g.gmin[0], g.gmin[1], g.gmin[2]);
                        const uint64_t fa = (uint64_t)it.face * 3u + (uint64_t)ax2;
                        const uint64_t keybase = (fa << KEY_FA_SHIFT) | ((uint64_t)(rw.y_idx - ylo) << KEY_Y_SHIFT) |
                                                 ((uint64_t)(x_idx - xlo) << KEY_X_SHIFT);
                        event_action<MODE>(g, ax0, ax1, ax2, x_idx, rw.y_idx, z_idx, (float)z, keybase, r);
                    }
                }
            }
            __syncthreads();
        }
        __syncthreads();
    }
    if (counters)
    {
        __shared__ uint32_t shr[NWAVE];
        const uint32_t tot = block_reduce_sum(n_ev, shr);
        if (tid == 0 && tot)
            atomicAdd(counters + (MODE == MODE_MARK ? CTR_EVENTS : CTR_REC_INTERSECT),
                      (unsigned long long)tot * (MODE == MODE_MARK ? 1ull : 4ull));
    }
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
};

struct FaceRow
{
    int item, y, z;
};

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

__device__ __forceinline__ uint64_t row_mask(int lo_bit, int hi_bit)
{
    // bits [lo_bit, hi_bit) of a 64-bit word, 0 <= lo_bit < hi_bit <= 64
    const uint64_t hi = (hi_bit >= 64) ? ~0ull : ((1ull << hi_bit) - 1ull);
    const uint64_t lo = (1ull << lo_bit) - 1ull;
    return hi & ~lo;
}

template <int MODE>
__global__ __launch_bounds__(BT) void k_face(MeshDev m, GridDev g, RecArgs r, unsigned long long *counters)
{
    __shared__ FaceItem items[BT];
    __shared__ FaceRow rows[BT];
    __shared__ uint64_t incl_items[BT];
    __shared__ uint32_t incl_rows[BT];
    __shared__ uint64_t sh64[NWAVE];
    __shared__ uint32_t sh32[NWAVE];
    const int tid = threadIdx.x;
    uint32_t n_hit = 0;

    for (int64_t base = (int64_t)blockIdx.x * BT; base < m.F; base += (int64_t)gridDim.x * BT)
    {
        const int64_t face = base + tid;
        uint64_t R = 0;
        if (face < m.F)
        {
            float v0[3], v1[3], v2[3], pl[4];
            load_tri(m, face, v0, v1, v2);
            tri_plane(v0, v1, v2, pl);
            FaceItem &it = items[tid];
            const float n0 = pl[0], n1 = pl[1], n2 = pl[2];
            int bmin[3], bmax[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                const float mn = sminf(sminf(v0[i], v1[i]), v2[i]) / g.vs[i];
                const float mx = smaxf(smaxf(v0[i], v1[i]), v2[i]) / g.vs[i];
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
                R = (uint64_t)(bmax[1] - bmin[1]) * (uint64_t)(bmax[2] - bmin[2]);
        }
        uint64_t tot_r;
        const uint64_t ex_r = block_excl_scan(R, sh64, tot_r);
        incl_items[tid] = ex_r + R;
        __syncthreads();

        for (uint64_t rb = 0; rb < tot_r; rb += BT)
        {
            const uint64_t gi = rb + tid;
            uint32_t A = 0;
            if (gi < tot_r)
            {
                const int L = upper_find(incl_items, BT, gi);
                const FaceItem &it = items[L];
                const uint64_t R_L = incl_items[L] - (L ? incl_items[L - 1] : 0ull);
                const uint64_t rr = gi - (incl_items[L] - R_L);
                const int y = it.by0 + (int)(rr % (uint64_t)it.ny);
                const int z = it.bz0 + (int)(rr / (uint64_t)it.ny);
                FaceRow &row = rows[tid];
                row.item = L;
                row.y = y;
                row.z = z;
                const int lx0 = it.bx0 - g.gmin[0], lx1 = it.bx1 - g.gmin[0];
                const uint64_t rowbase = ((uint64_t)(z - g.gmin[2]) * g.G[1] + (uint64_t)(y - g.gmin[1])) * g.wx;
                for (int w = lx0 >> 6; w <= ((lx1 - 1) >> 6); ++w)
                {
                    const int blo = (w == (lx0 >> 6)) ? (lx0 & 63) : 0;
                    const int bhi = (w == ((lx1 - 1) >> 6)) ? (((lx1 - 1) & 63) + 1) : 64;
                    A += (uint32_t)__popcll(g.bits[rowbase + w] & row_mask(blo, bhi));
                }
            }
            uint32_t tot_a;
            const uint32_t ex_a = block_excl_scan(A, sh32, tot_a);
            incl_rows[tid] = ex_a + A;
            __syncthreads();
            for (uint32_t ab = 0; ab < tot_a; ab += BT)
            {
                const uint32_t h = ab + tid;
                if (h < tot_a)
                {
                    const int rr = upper_find(incl_rows, BT, h);
                    const FaceRow &row = rows[rr];
                    const FaceItem &it = items[row.item];
                    uint32_t k = h - ((rr == 0) ? 0u : incl_rows[rr - 1]);
                    const int lx0 = it.bx0 - g.gmin[0], lx1 = it.bx1 - g.gmin[0];
                    const uint64_t rowbase =
                        ((uint64_t)(row.z - g.gmin[2]) * g.G[1] + (uint64_t)(row.y - g.gmin[1])) * g.wx;
                    for (int w = lx0 >> 6; w <= ((lx1 - 1) >> 6); ++w)
                    {
                        const int blo = (w == (lx0 >> 6)) ? (lx0 & 63) : 0;
                        const int bhi = (w == ((lx1 - 1) >> 6)) ? (((lx1 - 1) & 63) + 1) : 64;
                        const uint64_t word = g.bits[rowbase + w];
                        uint64_t mm = word & row_mask(blo, bhi);
                        const uint32_t pc = (uint32_t)__popcll(mm);
                        if (k >= pc)
                        {
                            k -= pc;
                            continue;
                        }
                        for (uint32_t j = 0; j < k; ++j)
                            mm &= mm - 1ull;
                        const int b = __builtin_ctzll(mm);
                        const int x = g.gmin[0] + w * 64 + b;
                        if (face_pass(it, g.vs, x, row.y, row.z))
                        {
                            ++n_hit;
                            if (MODE != MODE_MARK)
                            {
                                const uint32_t vid =
                                    g.woff[rowbase + w] + (uint32_t)__popcll(word & ((1ull << b) - 1ull));
                                record<MODE>(vid, (1ull << KEY_PHASE_SHIFT) | (uint64_t)it.face, 0.0f, r);
                            }
                        }
                        break;
                    }
                }
            }
            __syncthreads();
        }
        __syncthreads();
    }
    if (counters && MODE == MODE_COUNT)
    {
        __shared__ uint32_t shr[NWAVE];
        const uint32_t tot = block_reduce_sum(n_hit, shr);
        if (tid == 0 && tot)
            atomicAdd(counters + CTR_REC_FACE, (unsigned long long)tot);
    }
}

// ---------------------------------------------------------------------------
// boundary edges (FDG.cpp:724-756): per-min-vertex adjacency lists, sorted,
// singletons are boundary edges; ranks follow (min, max) ascending order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BT) void k_edge_count(MeshDev m, uint32_t *vcnt, uint32_t *voff, uint32_t *nbr, int pass)
{
    const int64_t stride = (int64_t)gridDim.x * BT;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < 3 * m.F; i += stride)
    {
        const int64_t f = i / 3;
        const int k = (int)(i - 3 * f);
        int a = m.f[3 * f + k];
        int b = m.f[3 * f + (k + 1) % 3];
        if (a > b)
        {
            const int t = a;
            a = b;
            b = t;
        }
        if (pass == 0)
            atomicAdd(vcnt + a, 1u);
        else
        {
            const uint32_t pos = voff[a] + atomicSub(vcnt + a, 1u) - 1u;
            nbr[pos] = (uint32_t)b;
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

// DDA over a boundary segment (FDG.cpp:455-522); one lane per segment.
template <int MODE>
__global__ __launch_bounds__(BT) void k_bdry(int64_t E, const BEdge *be, GridDev g, RecArgs r,
                                             unsigned long long *counters)
{
    const int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x;
    uint32_t n_hit = 0;
    unsigned err = 0;
    if (i < E && be[i].valid)
    {
        const BEdge &e = be[i];
        int cur[3], step[3];
        double tMax[3], tDelta[3];
        double cap = 16.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            cur[a] = (int)floorf(e.v0[a] / g.vs[a]);
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
        const uint64_t key = (2ull << KEY_PHASE_SHIFT) | (uint64_t)i;
        int64_t steps = 0;
        const int64_t max_steps = (int64_t)(cap < 4.0e9 ? cap : 4.0e9);
        while (true)
        {
            if (cur[0] >= g.gmin[0] && cur[0] < g.gmax[0] && cur[1] >= g.gmin[1] && cur[1] < g.gmax[1] &&
                cur[2] >= g.gmin[2] && cur[2] < g.gmax[2])
            {
                int b;
                const uint64_t w = word_of(g, cur[0], cur[1], cur[2], b);
                const uint64_t word = g.bits[w];
                if ((word >> b) & 1ull)
                {
                    ++n_hit;
                    if (MODE != MODE_MARK)
                        record<MODE>(g.woff[w] + (uint32_t)__popcll(word & ((1ull << b) - 1ull)), key, 0.0f, r);
                }
            }
            int axis;
            if (tMax[0] < tMax[1])
                axis = (tMax[0] < tMax[2]) ? 0 : 2;
            else
                axis = (tMax[1] < tMax[2]) ? 1 : 2;
            const double tm = axis == 0 ? tMax[0] : (axis == 1 ? tMax[1] : tMax[2]);
            if (tm > e.len)
                break;
            if (++steps > max_steps)
            {
                err |= ERR_DDA;
                break;
            }
            if (axis == 0)
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
    if (counters)
    {
        __shared__ uint32_t shr[NWAVE];
        const uint32_t tot = block_reduce_sum(n_hit, shr);
        if (threadIdx.x == 0 && tot && MODE == MODE_COUNT)
            atomicAdd(counters + CTR_REC_BOUNDARY, (unsigned long long)tot);
        if (err)
            atomicOr(counters + CTR_ERRORS, (unsigned long long)err);
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
__global__ __launch_bounds__(BT) void k_emit_coords(GridDev g, uint64_t n_words, int32_t *coords)
{
    const uint64_t stride = (uint64_t)gridDim.x * BT;
    for (uint64_t w = (uint64_t)blockIdx.x * BT + threadIdx.x; w < n_words; w += stride)
    {
        uint64_t m = g.bits[w];
        if (!m)
            continue;
        uint32_t id = g.woff[w];
        const uint64_t row = w / (uint64_t)g.wx;
        const int xw = (int)(w - row * (uint64_t)g.wx);
        const int y = (int)(row % (uint64_t)g.G[1]) + g.gmin[1];
        const int z = (int)(row / (uint64_t)g.G[1]) + g.gmin[2];
        while (m)
        {
            const int b = __builtin_ctzll(m);
            m &= m - 1ull;
            int32_t *o = coords + 3 * (uint64_t)id;
            o[0] = g.gmin[0] + xw * 64 + b;
            o[1] = y;
            o[2] = z;
            ++id;
        }
    }
}

// ---------------------------------------------------------------------------
// QEF solve (FDG.cpp:763-983): Eigen ColPivHouseholderQR restated for n = 2, 3
// ---------------------------------------------------------------------------
template <int n>
__device__ __forceinline__ void cpqr_solve(const float (&A)[n][n], const float (&b)[n], float (&x)[n])
{
    // one spare row/column keeps the unrolled (dead) tail indices in bounds
    float qr[n + 1][n + 1];
#pragma unroll
    for (int i = 0; i <= n; ++i)
#pragma unroll
        for (int j = 0; j <= n; ++j)
            qr[i][j] = (i < n && j < n) ? A[i][j] : 0.0f;
    float hco[n];
    int transp[n];
    float nu[n], nd[n];
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
        float s;
        if constexpr (n == 3)
            s = qr[0][k] * qr[0][k] + (qr[1][k] * qr[1][k] + qr[2][k] * qr[2][k]);
        else
            s = qr[0][k] * qr[0][k] + qr[1][k] * qr[1][k];
        nd[k] = sqrtf(s);
        nu[k] = nd[k];
    }
    float maxn = nu[0];
#pragma unroll
    for (int k = 1; k < n; ++k)
        if (nu[k] > maxn)
            maxn = nu[k];
    const float me = maxn * FLT_EPSILON;
    const float th = (me * me) / (float)n;
    const float ndt = sqrtf(FLT_EPSILON);
    int nzp = n;
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
        int big = k;
        float bigv = nu[k];
#pragma unroll
        for (int j = k + 1; j < n; ++j)
            if (nu[j] > bigv)
            {
                bigv = nu[j];
                big = j;
            }
        if (nzp == n && bigv * bigv < th * (float)(n - k))
            nzp = k;
        transp[k] = big;
#pragma unroll
        for (int j = k + 1; j < n; ++j)
            if (j == big)
            {
#pragma unroll
                for (int rr = 0; rr < n; ++rr)
                {
                    const float t = qr[rr][k];
                    qr[rr][k] = qr[rr][j];
                    qr[rr][j] = t;
                }
                float t = nu[k];
                nu[k] = nu[j];
                nu[j] = t;
                t = nd[k];
                nd[k] = nd[j];
                nd[j] = t;
            }
        const int tl = n - k - 1;
        float tail_sq = 0.0f;
        if (tl > 0)
        {
            tail_sq = qr[k + 1][k] * qr[k + 1][k];
#pragma unroll
            for (int i = 1; i < n - k - 1; ++i)
                tail_sq = tail_sq + qr[k + 1 + i][k] * qr[k + 1 + i][k];
        }
        const float c0 = qr[k][k];
        float beta, tau;
        if (tail_sq <= FLT_MIN)
        {
            tau = 0.0f;
            beta = c0;
#pragma unroll
            for (int i = 0; i < n - k - 1; ++i)
                qr[k + 1 + i][k] = 0.0f;
        }
        else
        {
            beta = sqrtf(c0 * c0 + tail_sq);
            if (c0 >= 0.0f)
                beta = -beta;
            const float den = c0 - beta;
#pragma unroll
            for (int i = 0; i < n - k - 1; ++i)
                qr[k + 1 + i][k] = qr[k + 1 + i][k] / den;
            tau = (beta - c0) / beta;
        }
        hco[k] = tau;
        qr[k][k] = beta;
        if (n - k > 1 && tau != 0.0f)
        {
#pragma unroll
            for (int j = k + 1; j < n; ++j)
            {
                float tmp = qr[k + 1][k] * qr[k + 1][j];
#pragma unroll
                for (int i = 1; i < n - k - 1; ++i)
                    tmp = tmp + qr[k + 1 + i][k] * qr[k + 1 + i][j];
                tmp = tmp + qr[k][j];
                qr[k][j] = qr[k][j] - tau * tmp;
#pragma unroll
                for (int i = 0; i < n - k - 1; ++i)
                {
                    const float te = tau * qr[k + 1 + i][k];
                    qr[k + 1 + i][j] = qr[k + 1 + i][j] - tmp * te;
                }
            }
        }
#pragma unroll
        for (int j = k + 1; j < n; ++j)
        {
            if (nu[j] != 0.0f)
            {
                const float a = fabsf(qr[k][j]);
                float temp = a / nu[j];
                temp = (1.0f + temp) * (1.0f - temp);
                temp = temp < 0.0f ? 0.0f : temp;
                const float ratio = nu[j] / nd[j];
                const float temp2 = temp * (ratio * ratio);
                if (temp2 <= ndt)
                {
                    float s = 0.0f;
                    if (k + 1 < n)
                    {
                        s = qr[k + 1][j] * qr[k + 1][j];
#pragma unroll
                        for (int i = k + 2; i < n; ++i)
                            s = s + qr[i][j] * qr[i][j];
                        s = sqrtf(s);
                    }
                    nd[j] = s;
                    nu[j] = s;
                }
                else
                {
                    nu[j] = nu[j] * sqrtf(temp);
                }
            }
        }
    }
    int perm[n];
#pragma unroll
    for (int i = 0; i < n; ++i)
        perm[i] = i;
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
#pragma unroll
        for (int j = k + 1; j < n; ++j)
            if (transp[k] == j)
            {
                const int t = perm[k];
                perm[k] = perm[j];
                perm[j] = t;
            }
    }
    const int rk = nzp;
    float c[n + 1];
#pragma unroll
    for (int i = 0; i < n; ++i)
        c[i] = b[i];
    c[n] = 0.0f;
    if (rk == 0)
    {
#pragma unroll
        for (int i = 0; i < n; ++i)
            x[i] = 0.0f;
        return;
    }
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
        if (k < rk)
        {
            const float tau = hco[k];
            if (n - k == 1)
            {
                c[k] = c[k] * (1.0f - tau);
            }
            else if (tau != 0.0f)
            {
                float tmp = qr[k + 1][k] * c[k + 1];
#pragma unroll
                for (int i = 1; i < n - k - 1; ++i)
                    tmp = tmp + qr[k + 1 + i][k] * c[k + 1 + i];
                tmp = tmp + c[k];
                c[k] = c[k] - tau * tmp;
#pragma unroll
                for (int i = 0; i < n - k - 1; ++i)
                {
                    const float te = tau * qr[k + 1 + i][k];
                    c[k + 1 + i] = c[k + 1 + i] - tmp * te;
                }
            }
        }
    }
#pragma unroll
    for (int i = n - 1; i >= 0; --i)
    {
        if (i < rk && c[i] != 0.0f)
        {
            c[i] = c[i] / qr[i][i];
#pragma unroll
            for (int s = 0; s < i; ++s)
                c[s] = c[s] - c[i] * qr[s][i];
        }
    }
#pragma unroll
    for (int j = 0; j < n; ++j)
    {
        float v = 0.0f;
#pragma unroll
        for (int i = 0; i < n; ++i)
            if (perm[i] == j && i < rk)
                v = c[i];
        x[j] = v;
    }
}

// p^T Q p in the Packet4f predux order (x0 + x2) + (x1 + x3)
__device__ __forceinline__ float quad_form(const float p[4], const float Q[4][4])
{
    float t[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        t[j] = (p[0] * Q[0][j] + p[2] * Q[2][j]) + (p[1] * Q[1][j] + p[3] * Q[3][j]);
    return (t[0] * p[0] + t[2] * p[2]) + (t[1] * p[1] + t[3] * p[3]);
}

__device__ __forceinline__ void solve_voxel(const float q[10], const int coord[3], const float vs[3], float out[3])
{
    float Q[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            Q[i][j] = q[qidx(i, j)];
    float mn[3], mx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        mn[a] = (float)coord[a] * vs[a];
        mx[a] = (float)(coord[a] + 1) * vs[a];
    }
    const float A[3][3] = {{Q[0][0], Q[0][1], Q[0][2]}, {Q[1][0], Q[1][1], Q[1][2]}, {Q[2][0], Q[2][1], Q[2][2]}};
    const float bv[3] = {-Q[0][3], -Q[1][3], -Q[2][3]};
    float v[3];
    cpqr_solve<3>(A, bv, v);
    if (!(v[0] >= mn[0] && v[0] <= mx[0] && v[1] >= mn[1] && v[1] <= mx[1] && v[2] >= mn[2] && v[2] <= mx[2]))
    {
        float best = __builtin_huge_valf();
        auto consider = [&](const float p[4]) {
            const float err = quad_form(p, Q);
            if (err < best)
            {
                best = err;
                v[0] = p[0];
                v[1] = p[1];
                v[2] = p[2];
            }
        };
#pragma unroll
        for (int fixed = 0; fixed < 3; ++fixed)
        {
            const int a1 = (fixed + 1) % 3, a2 = (fixed + 2) % 3;
            const float A2[2][2] = {{Q[a1][a1], Q[a1][a2]}, {Q[a2][a1], Q[a2][a2]}};
            const float B[2][2] = {{Q[a1][fixed], Q[a1][3]}, {Q[a2][fixed], Q[a2][3]}};
#pragma unroll
            for (int side = 0; side < 2; ++side)
            {
                const float q0 = side ? mx[fixed] : mn[fixed];
                const float rhs[2] = {(-B[0][0]) * q0 + (-B[0][1]) * 1.0f, (-B[1][0]) * q0 + (-B[1][1]) * 1.0f};
                float x2[2];
                cpqr_solve<2>(A2, rhs, x2);
                if (x2[0] >= mn[a1] && x2[0] <= mx[a1] && x2[1] >= mn[a2] && x2[1] <= mx[a2])
                {
                    float p[4];
                    p[fixed] = q0;
                    p[a1] = x2[0];
                    p[a2] = x2[1];
                    p[3] = 1.0f;
                    consider(p);
                }
            }
        }
#pragma unroll
        for (int fr = 0; fr < 3; ++fr)
        {
            const int a1 = (fr + 1) % 3, a2 = (fr + 2) % 3;
            const float a = Q[fr][fr];
            const float b0 = Q[fr][a1], b1 = Q[fr][a2], b2 = Q[fr][3];
#pragma unroll
            for (int s1 = 0; s1 < 2; ++s1)
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2)
                {
                    const float q1 = s1 ? mx[a1] : mn[a1];
                    const float q2 = s2 ? mx[a2] : mn[a2];
                    const float xv = -(dot3f(b0, b1, b2, q1, q2, 1.0f)) / a;
                    if (xv >= mn[fr] && xv <= mx[fr])
                    {
                        float p[4];
                        p[fr] = xv;
                        p[a1] = q1;
                        p[a2] = q2;
                        p[3] = 1.0f;
                        consider(p);
                    }
                }
        }
#pragma unroll
        for (int cx = 0; cx < 2; ++cx)
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cz = 0; cz < 2; ++cz)
                {
                    const float p[4] = {cx ? mn[0] : mx[0], cy ? mn[1] : mx[1], cz ? mn[2] : mx[2], 1.0f};
                    consider(p);
                }
    }
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
}

// Replay one voxel's contributions in key order and solve.
template <class Get>
__device__ __forceinline__ void fold_solve(uint32_t vid, int n, Get get, const FoldArgs &a)
{
    float q[10];
#pragma unroll
    for (int k = 0; k < 10; ++k)
        q[k] = 0.0f;
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f, cnt = 0.0f;
    int fl = 0;
    for (int i = 0; i < n; ++i)
    {
        uint64_t key;
        float z;
        get(i, key, z);
        const int phase = (int)(key >> KEY_PHASE_SHIFT);
        if (phase == 0)
        {
            const uint64_t fa = (key >> KEY_FA_SHIFT) & KEY_FA_MASK;
            const uint32_t f = (uint32_t)(fa / 3u);
            const int ax2 = (int)(fa - 3ull * f);
            const int ax0 = (ax2 + 1) % 3;
            const int ax1 = (ax2 + 2) % 3;
            const int y_idx = (int)((key >> KEY_Y_SHIFT) & KEY_IDX_MASK) + sel3i(ax1, a.gmin[0], a.gmin[1], a.gmin[2]);
            const int x_idx = (int)((key >> KEY_X_SHIFT) & KEY_IDX_MASK) + sel3i(ax0, a.gmin[0], a.gmin[1], a.gmin[2]);
            const float px = (float)(x_idx + 1) * sel3f(ax0, a.vs[0], a.vs[1], a.vs[2]);
            const float py = (float)(y_idx + 1) * sel3f(ax1, a.vs[0], a.vs[1], a.vs[2]);
            const float p0 = ax0 == 0 ? px : (ax1 == 0 ? py : z);
            const float p1 = ax0 == 1 ? px : (ax1 == 1 ? py : z);
            const float p2 = ax0 == 2 ? px : (ax1 == 2 ? py : z);
            const float4 pl4 = a.planes[f];
            const float pl[4] = {pl4.x, pl4.y, pl4.z, pl4.w};
            float qf[10];
            plane_quadric(pl, qf);
            if (i == 0)
            {
                m0 = p0;
                m1 = p1;
                m2 = p2;
                cnt = 1.0f;
#pragma unroll
                for (int k = 0; k < 10; ++k)
                    q[k] = qf[k];
            }
            else
            {
                m0 = m0 + p0;
                m1 = m1 + p1;
                m2 = m2 + p2;
                cnt = cnt + 1.0f;
#pragma unroll
                for (int k = 0; k < 10; ++k)
                    q[k] = q[k] + qf[k];
            }
            if ((key & 3ull) == 0ull)
                fl |= 1 << ax2;
        }
        else if (phase == 1)
        {
            const uint32_t f = (uint32_t)(key & 0xffffffffull);
            const float4 pl4 = a.planes[f];
            const float pl[4] = {pl4.x, pl4.y, pl4.z, pl4.w};
            float qf[10];
            plane_quadric(pl, qf);
#pragma unroll
            for (int k = 0; k < 10; ++k)
                q[k] = q[k] + qf[k];
        }
        else
        {
            const uint64_t e = key & ((1ull << KEY_PHASE_SHIFT) - 1ull);
            const float *qe = a.bedges[e].q;
#pragma unroll
            for (int k = 0; k < 10; ++k)
                q[k] = q[k] + a.bw * qe[k];
        }
    }
    const int coord[3] = {a.coords[3 * (uint64_t)vid + 0], a.coords[3 * (uint64_t)vid + 1],
                          a.coords[3 * (uint64_t)vid + 2]};
    if (a.rw > 0.0f && n > 0)
    {
        const float p0 = m0 / cnt, p1 = m1 / cnt, p2 = m2 / cnt;
        const float s = a.rw * cnt;
        const float R[10] = {1.0f, 0.0f, 0.0f, -p0, 1.0f, 0.0f, -p1, 1.0f, -p2, dot3f(p0, p1, p2, p0, p1, p2)};
#pragma unroll
        for (int k = 0; k < 10; ++k)
            q[k] = q[k] + s * R[k];
    }
    float d[3];
    solve_voxel(q, coord, a.vs, d);
    float *o = a.dual + 3 * (uint64_t)vid;
    o[0] = d[0];
    o[1] = d[1];
    o[2] = d[2];
    uint8_t *fo = a.flags + 3 * (uint64_t)vid;
    fo[0] = (uint8_t)(fl & 1);
    fo[1] = (uint8_t)((fl >> 1) & 1);
    fo[2] = (uint8_t)((fl >> 2) & 1);
}

constexpr int FOLD_CAP = 3072;

__global__ __launch_bounds__(BT) void k_fold_solve(FoldArgs a)
{
    __shared__ uint64_t skey[FOLD_CAP];
    __shared__ float sz[FOLD_CAP];
    __shared__ uint16_t sperm[FOLD_CAP];
    __shared__ uint32_t soff[BT + 1];
    const int tid = threadIdx.x;
    const uint64_t v0 = (uint64_t)blockIdx.x * BT;
    if (v0 >= a.N)
        return;
    const int nv = (int)((a.N - v0) < (uint64_t)BT ? (a.N - v0) : (uint64_t)BT);
    for (int i = tid; i <= nv; i += BT)
        soff[i] = a.rec_off[v0 + i];
    __syncthreads();

    int cv = 0;
    while (cv < nv)
    {
        const uint32_t base = soff[cv];
        if (soff[cv + 1] - base > (uint32_t)FOLD_CAP)
        {
            // one voxel with more records than the LDS tile: rank-sort it in
            // global memory, then fold it on one lane
            const uint32_t nrec = soff[cv + 1] - base;
            for (uint32_t i = tid; i < nrec; i += BT)
            {
                const uint64_t k = a.keys[base + i];
                uint32_t rank = 0;
                for (uint32_t t = 0; t < nrec; ++t)
                    rank += (a.keys[base + t] < k) ? 1u : 0u;
                a.gperm[base + rank] = i;
            }
            __syncthreads();
            if (tid == 0)
            {
                const uint64_t *K = a.keys + base;
                const float *Z = a.zs + base;
                const uint32_t *P = a.gperm + base;
                fold_solve((uint32_t)(v0 + cv), (int)nrec,
                           [&](int i, uint64_t &key, float &z) {
                               const uint32_t j = P[i];
                               key = K[j];
                               z = Z[j];
                           },
                           a);
                atomicAdd(a.counters + CTR_GIANT, 1ull);
            }
            __syncthreads();
            cv += 1;
            continue;
        }
        // largest ce with soff[ce] - base <= FOLD_CAP
        int lo = cv + 1, hi = nv;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (soff[mid] - base <= (uint32_t)FOLD_CAP)
                lo = mid;
            else
                hi = mid - 1;
        }
        const int ce = lo;
        const uint32_t nrec = soff[ce] - base;
        for (uint32_t i = tid; i < nrec; i += BT)
        {
            skey[i] = a.keys[base + i];
            sz[i] = a.zs[base + i];
        }
        __syncthreads();
        for (uint32_t i = tid; i < nrec; i += BT)
        {
            // voxel of record i: largest j in [cv, ce) with soff[j] - base <= i
            int l = cv, h = ce - 1;
            while (l < h)
            {
                const int mid = (l + h + 1) >> 1;
                if (soff[mid] - base <= i)
                    l = mid;
                else
                    h = mid - 1;
            }
            const uint32_t s0 = soff[l] - base, s1 = soff[l + 1] - base;
            const uint64_t k = skey[i];
            uint32_t rank = 0;
            for (uint32_t t = s0; t < s1; ++t)
                rank += (skey[t] < k) ? 1u : 0u;
            sperm[s0 + rank] = (uint16_t)i;
        }
        __syncthreads();
        if (tid >= cv && tid < ce)
        {
            const uint32_t s0 = soff[tid] - base;
            const int n = (int)(soff[tid + 1] - soff[tid]);
            fold_solve((uint32_t)(v0 + tid), n,
                       [&](int i, uint64_t &key, float &z) {
                           const uint32_t j = sperm[s0 + i];
                           key = skey[j];
                           z = sz[j];
                       },
                       a);
        }
        __syncthreads();
        cv = ce;
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

hipError_t launch_prep(const MeshDev &m, int64_t V, const GridDev &g, float4 *planes, unsigned long long *ctr,
                       hipStream_t s)
{
    const int64_t n = m.F > V ? m.F : V;
    hipLaunchKernelGGL(k_prep, dim3(grid_for(n, 8192)), dim3(BT), 0, s, m, V, g, planes, ctr);
    return hipGetLastError();
}

hipError_t launch_scan_events(int mode, const MeshDev &m, const GridDev &g, const RecArgs &r,
                              unsigned long long *ctr, hipStream_t s)
{
    const unsigned nb = grid_for(3 * m.F);
    if (mode == MODE_MARK)
        hipLaunchKernelGGL(k_scan_events<MODE_MARK>, dim3(nb), dim3(BT), 0, s, m, g, r, ctr);
    else if (mode == MODE_COUNT)
        hipLaunchKernelGGL(k_scan_events<MODE_COUNT>, dim3(nb), dim3(BT), 0, s, m, g, r, ctr);
    else
        hipLaunchKernelGGL(k_scan_events<MODE_EMIT>, dim3(nb), dim3(BT), 0, s, m, g, r, ctr);
    return hipGetLastError();
}

hipError_t launch_face(int mode, const MeshDev &m, const GridDev &g, const RecArgs &r, unsigned long long *ctr,
                       hipStream_t s)
{
    const unsigned nb = grid_for(m.F);
    if (mode == MODE_COUNT)
        hipLaunchKernelGGL(k_face<MODE_COUNT>, dim3(nb), dim3(BT), 0, s, m, g, r, ctr);
    else
        hipLaunchKernelGGL(k_face<MODE_EMIT>, dim3(nb), dim3(BT), 0, s, m, g, r, ctr);
    return hipGetLastError();
}

hipError_t launch_edge_count(const MeshDev &m, uint32_t *vcnt, uint32_t *voff, uint32_t *nbr, int pass, hipStream_t s)
{
    hipLaunchKernelGGL(k_edge_count, dim3(grid_for(3 * m.F, 65536)), dim3(BT), 0, s, m, vcnt, voff, nbr, pass);
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

hipError_t launch_bdry(int mode, int64_t E, const BEdge *be, const GridDev &g, const RecArgs &r,
                       unsigned long long *ctr, hipStream_t s)
{
    const unsigned nb = grid_for(E);
    if (mode == MODE_COUNT)
        hipLaunchKernelGGL(k_bdry<MODE_COUNT>, dim3(nb), dim3(BT), 0, s, E, be, g, r, ctr);
    else
        hipLaunchKernelGGL(k_bdry<MODE_EMIT>, dim3(nb), dim3(BT), 0, s, E, be, g, r, ctr);
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

hipError_t launch_scan_u32(const uint32_t *in, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                           unsigned long long *total, hipStream_t s)
{
    return scan_impl(LoadU32{in}, n, out, tile_sums, total, s);
}

hipError_t launch_emit_coords(const GridDev &g, uint64_t n_words, int32_t *coords, hipStream_t s)
{
    hipLaunchKernelGGL(k_emit_coords, dim3(grid_for((int64_t)n_words, 65536)), dim3(BT), 0, s, g, n_words, coords);
    return hipGetLastError();
}

hipError_t launch_fold_solve(const FoldArgs &a, hipStream_t s)
{
    const uint64_t nb = (a.N + BT - 1) / BT;
    if (nb == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fold_solve, dim3((unsigned)nb), dim3(BT), 0, s, a);
    return hipGetLastError();
}

} // namespace ovx
