// Dual grid -> triangle mesh (flexible_dual_grid_to_mesh) and the generic
// hashmap of the extension boundary, for MI355X (gfx950).
//
// flexible_dual_grid_to_mesh (reference o_voxel/convert/flexible_dual_grid.py:331-477)
// builds a voxel table, looks up the four voxels around every intersected
// grid edge, keeps the quads whose four voxels all exist (in (voxel, axis)
// order) and splits each into two triangles.  Here that is four launches:
// clear + insert the private voxel table, a count pass (one lane per voxel:
// mesh vertex, quad mask, quad count), a scan of the counts (fdg_kernels),
// and an emit pass that writes each voxel's quads and triangles at its scan
// offset, so the output order is the reference's without any compaction of
// a [N, 3, 4] candidate array.
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "dual_mesh.h"

namespace ovx
{
namespace
{

constexpr int HB = 256;
constexpr uint32_t NONE = 0xffffffffu;

unsigned grid_of(uint64_t n, uint64_t cap_blocks)
{
    uint64_t b = (n + HB - 1) / HB;
    if (b > cap_blocks)
        b = cap_blocks;
    return (unsigned)(b ? b : 1);
}

// Murmur3 64/32-bit finalizers: the hash of the reference's table
// (src/hash/hash.cuh:1-20); tables stay interchangeable with it.
__device__ __forceinline__ uint32_t fmix(uint32_t k)
{
    k ^= k >> 16;
    k *= 0x85ebca6bu;
    k ^= k >> 13;
    k *= 0xc2b2ae35u;
    k ^= k >> 16;
    return k;
}

__device__ __forceinline__ uint64_t fmix(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ uint32_t cas(uint32_t *p, uint32_t cmp, uint32_t val)
{
    return atomicCAS((unsigned int *)p, cmp, val);
}

__device__ __forceinline__ uint64_t cas(uint64_t *p, uint64_t cmp, uint64_t val)
{
    return (uint64_t)atomicCAS((unsigned long long *)p, (unsigned long long)cmp, (unsigned long long)val);
}

// b*W*H*D + x*H*D + y*D + z with every int widened to a 64-bit unsigned
// (sign-extended, products modulo 2^64) as the reference's size_t arithmetic
// does (src/hash/hash.cu, insert_3d / lookup_3d kernels)
__device__ __forceinline__ uint64_t flat_key(int b, int x, int y, int z, int W, int H, int D)
{
    const uint64_t w = (uint64_t)(int64_t)W, h = (uint64_t)(int64_t)H, d = (uint64_t)(int64_t)D;
    return (uint64_t)(int64_t)b * w * h * d + (uint64_t)(int64_t)x * h * d + (uint64_t)(int64_t)y * d +
           (uint64_t)(int64_t)z;
}

// ---------------------------------------------------------------------------
// generic table (reference layout): separate key / value arrays, all-ones
// empty key, hash modulo capacity, linear probing.  Probing stops after one
// full sweep: a full table drops the key (insert) or reports it missing
// (lookup) instead of probing forever.
// ---------------------------------------------------------------------------
template <class K, class V>
__global__ __launch_bounds__(HB) void k_hash_insert(K *tk, V *tv, uint64_t cap, int src, const void *keys,
                                                    const V *vals, uint64_t n, int W, int H, int D)
{
    const K EMPTY = (K)~(K)0;
    const uint64_t stride = (uint64_t)gridDim.x * HB;
    for (uint64_t i = (uint64_t)blockIdx.x * HB + threadIdx.x; i < n; i += stride)
    {
        K key;
        if (src == HASH_KEYS)
            key = ((const K *)keys)[i];
        else
        {
            const int4 c = ((const int4 *)keys)[i];
            key = (K)flat_key(c.x, c.y, c.z, c.w, W, H, D);
        }
        const V val = vals ? vals[i] : (V)i;
        uint64_t slot = (uint64_t)fmix(key) % cap;
        for (uint64_t p = 0; p < cap; ++p)
        {
            const K prev = cas(&tk[slot], EMPTY, key);
            if (prev == EMPTY || prev == key)
            {
                tv[slot] = val;
                break;
            }
            if (++slot == cap)
                slot = 0;
        }
    }
}

template <class K, class V>
__global__ __launch_bounds__(HB) void k_hash_lookup(const K *tk, const V *tv, uint64_t cap, int src, const void *keys,
                                                    V *out, uint64_t n, int W, int H, int D)
{
    const K EMPTY = (K)~(K)0;
    const V MISSING = (V)~(V)0;
    const uint64_t stride = (uint64_t)gridDim.x * HB;
    for (uint64_t i = (uint64_t)blockIdx.x * HB + threadIdx.x; i < n; i += stride)
    {
        K key;
        if (src == HASH_KEYS)
            key = ((const K *)keys)[i];
        else
        {
            const int4 c = ((const int4 *)keys)[i];
            if (c.y < 0 || c.y >= W || c.z < 0 || c.z >= H || c.w < 0 || c.w >= D)
            {
                out[i] = MISSING;
                continue;
            }
            key = (K)flat_key(c.x, c.y, c.z, c.w, W, H, D);
        }
        V r = MISSING;
        uint64_t slot = (uint64_t)fmix(key) % cap;
        for (uint64_t p = 0; p < cap; ++p)
        {
            const K k = tk[slot];
            if (k == EMPTY)
                break;
            if (k == key)
            {
                r = tv[slot];
                break;
            }
            if (++slot == cap)
                slot = 0;
        }
        out[i] = r;
    }
}

template <class K, class V>
hipError_t hash_insert_t(void *tk, void *tv, uint64_t cap, int src, const void *keys, const void *vals, uint64_t n,
                         int W, int H, int D, hipStream_t s)
{
    hipLaunchKernelGGL((k_hash_insert<K, V>), dim3(grid_of(n, 1u << 16)), dim3(HB), 0, s, (K *)tk, (V *)tv, cap, src,
                       keys, (const V *)vals, n, W, H, D);
    return hipGetLastError();
}

template <class K, class V>
hipError_t hash_lookup_t(const void *tk, const void *tv, uint64_t cap, int src, const void *keys, void *out,
                         uint64_t n, int W, int H, int D, hipStream_t s)
{
    hipLaunchKernelGGL((k_hash_lookup<K, V>), dim3(grid_of(n, 1u << 16)), dim3(HB), 0, s, (const K *)tk,
                       (const V *)tv, cap, src, keys, (V *)out, n, W, H, D);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// private voxel table of the dual-grid -> mesh pass.  One probe is one load:
// narrow entries pack (key << 32 | row) into 8 bytes; wide entries are
// 16 bytes (u64 key, u64 row).  Duplicate coordinates keep the largest row
// (the last one a sequential insertion would store).
// ---------------------------------------------------------------------------
constexpr uint64_t EMPTY8 = ~0ull;

__device__ __forceinline__ void put_narrow(uint64_t *t, uint64_t mask, uint32_t key, uint32_t row)
{
    const uint64_t ent = ((uint64_t)key << 32) | row;
    uint64_t slot = fmix(key) & mask;
    for (uint64_t p = 0; p <= mask; ++p)
    {
        const uint64_t prev = cas(&t[slot], EMPTY8, ent);
        if (prev == EMPTY8)
            return;
        if ((uint32_t)(prev >> 32) == key)
        {
            atomicMax((unsigned long long *)&t[slot], (unsigned long long)ent);
            return;
        }
        slot = (slot + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t get_narrow(const uint64_t *t, uint64_t mask, uint32_t key)
{
    uint64_t slot = fmix(key) & mask;
    for (uint64_t p = 0; p <= mask; ++p)
    {
        const uint64_t e = t[slot];
        if (e == EMPTY8)
            return NONE;
        if ((uint32_t)(e >> 32) == key)
            return (uint32_t)e;
        slot = (slot + 1) & mask;
    }
    return NONE;
}

__device__ __forceinline__ void put_wide(ulonglong2 *t, uint64_t mask, uint64_t key, uint32_t row)
{
    uint64_t slot = fmix(key) & mask;
    for (uint64_t p = 0; p <= mask; ++p)
    {
        const uint64_t prev = cas((uint64_t *)&t[slot].x, EMPTY8, key);
        if (prev == EMPTY8 || prev == key)
        {
            atomicMax((unsigned long long *)&t[slot].y, (unsigned long long)row);
            return;
        }
        slot = (slot + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t get_wide(const ulonglong2 *t, uint64_t mask, uint64_t key)
{
    uint64_t slot = fmix(key) & mask;
    for (uint64_t p = 0; p <= mask; ++p)
    {
        const ulonglong2 e = t[slot];
        if (e.x == EMPTY8)
            return NONE;
        if (e.x == key)
            return (uint32_t)e.y;
        slot = (slot + 1) & mask;
    }
    return NONE;
}

// The four voxels around the grid edge of axis a at the (+,+) corner of a
// voxel, in quad order (reference flexible_dual_grid.py:367-374).
__constant__ int c_edge_nbr[3][4][3] = {
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}},
    {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}},
    {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}},
};

// cell of the direct table: x fastest, so lanes holding consecutive x (the
// canonical voxel order) read neighbouring words
__device__ __forceinline__ uint64_t direct_cell(uint32_t x, uint32_t y, uint32_t z, const DualMeshArgs &a)
{
    return ((uint64_t)z * (uint32_t)a.H + y) * (uint32_t)a.W + x;
}

template <int MODE>
__device__ __forceinline__ uint32_t find_voxel(const DualMeshArgs &a, int x, int y, int z, int ax, int k)
{
    // neighbour coordinates wrap like the reference's int32 tensor sum; the
    // unsigned compare rejects everything outside [0, W) x [0, H) x [0, D)
    const uint32_t xx = (uint32_t)x + (uint32_t)c_edge_nbr[ax][k][0];
    const uint32_t yy = (uint32_t)y + (uint32_t)c_edge_nbr[ax][k][1];
    const uint32_t zz = (uint32_t)z + (uint32_t)c_edge_nbr[ax][k][2];
    if (xx >= (uint32_t)a.W || yy >= (uint32_t)a.H || zz >= (uint32_t)a.D)
        return NONE;
    if (MODE == DM_DIRECT)
    {
        // 0 (empty) -> NONE; a row beyond this call's voxels cannot come from
        // this call's insertions and is treated as empty as well
        const uint32_t r = ((const uint32_t *)a.table)[direct_cell(xx, yy, zz, a)] - 1u;
        return r < (uint64_t)a.N ? r : NONE;
    }
    const uint64_t key = ((uint64_t)xx * (uint32_t)a.H + yy) * (uint32_t)a.D + zz;
    const uint64_t mask = a.cap - 1;
    if (MODE == DM_WIDE)
        return get_wide((const ulonglong2 *)a.table, mask, key);
    return get_narrow((const uint64_t *)a.table, mask, (uint32_t)key);
}

__global__ __launch_bounds__(HB) void k_dm_restore(uint32_t *table, const uint32_t *touched, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * HB + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t t = touched[i];
    if (t != NONE)
        table[t] = 0u;
}

__global__ __launch_bounds__(HB) void k_dm_clear(void *table, uint64_t n, int wide)
{
    const uint64_t stride = (uint64_t)gridDim.x * HB;
    for (uint64_t i = (uint64_t)blockIdx.x * HB + threadIdx.x; i < n; i += stride)
    {
        if (wide)
            ((ulonglong2 *)table)[i] = make_ulonglong2(EMPTY8, 0ull);
        else
            ((uint64_t *)table)[i] = EMPTY8;
    }
}

template <int MODE>
__global__ __launch_bounds__(HB) void k_dm_insert(DualMeshArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * HB + threadIdx.x;
    if (i >= a.N)
        return;
    const int x = a.coords[3 * i], y = a.coords[3 * i + 1], z = a.coords[3 * i + 2];
    // the key the reference inserts for (0, x, y, z), truncated to the key width
    const uint64_t key = flat_key(0, x, y, z, a.W, a.H, a.D);
    if (MODE == DM_DIRECT)
    {
        // a coordinate outside the grid lands where its truncated key names
        // an in-grid cell (as in the reference's u32-keyed table), else nowhere
        uint32_t cx = (uint32_t)x, cy = (uint32_t)y, cz = (uint32_t)z;
        bool in = cx < (uint32_t)a.W && cy < (uint32_t)a.H && cz < (uint32_t)a.D;
        if (!in)
        {
            const uint32_t k32 = (uint32_t)key;
            const uint32_t hd = (uint32_t)a.H * (uint32_t)a.D;
            in = k32 < (uint32_t)a.W * hd;
            cx = k32 / hd;
            cy = (k32 - cx * hd) / (uint32_t)a.D;
            cz = k32 - cx * hd - cy * (uint32_t)a.D;
        }
        uint32_t t = NONE;
        if (in)
        {
            t = (uint32_t)direct_cell(cx, cy, cz, a);
            atomicMax((unsigned int *)a.table + t, (unsigned int)i + 1u); // duplicates: the last row
        }
        a.touched[i] = t;
        return;
    }
    // an all-ones 64-bit key (only from coordinates outside the grid) would
    // read as an empty slot; no in-grid lookup can match it
    if (MODE == DM_WIDE && key == EMPTY8)
        return;
    if (MODE == DM_WIDE)
        put_wide((ulonglong2 *)a.table, a.cap - 1, key, (uint32_t)i);
    else
        put_narrow((uint64_t *)a.table, a.cap - 1, (uint32_t)key, (uint32_t)i);
}

// mesh vertex (coords + dual) * voxel_size + aabb_min (one rounding per
// operation, reference flexible_dual_grid.py:437); quads whose four voxels
// all exist (reference :422-431)
template <int MODE>
__global__ __launch_bounds__(HB) void k_dm_count(DualMeshArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * HB + threadIdx.x;
    if (i >= a.N)
        return;
    const int c[3] = {a.coords[3 * i], a.coords[3 * i + 1], a.coords[3 * i + 2]};
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        float t = (float)c[k] + a.dual[3 * i + k];
        t = t * a.vs[k];
        a.verts[3 * i + k] = t + a.a0[k];
    }
    uint32_t m = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
    {
        if (!a.flags[3 * i + ax])
            continue;
        bool ok = true;
        for (int k = 0; k < 4 && ok; ++k)
            ok = find_voxel<MODE>(a, c[0], c[1], c[2], ax, k) != NONE;
        if (ok)
            m |= 1u << ax;
    }
    a.mask[i] = (uint8_t)m;
    a.cnt[i] = __popc(m);
}

struct V3
{
    float x, y, z;
};

__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ __forceinline__ float abs_dot(V3 a, V3 b)
{
    const float s = (a.x * b.x + a.y * b.y) + a.z * b.z;
    return fabsf(s);
}

// The reference's split choice without split weights
// (flexible_dual_grid.py:440-450): for each candidate split it takes the
// triangle columns (0,1,2) and (1,2,3) of the 6-index split pattern.  For
// split 1 = (0,1,2,0,2,3) those are triangles (q0,q1,q2) and (q1,q2,q0), the
// same triangle; for split 2 = (0,1,3,3,1,2) the second is (q1,q3,q3), whose
// normal is exactly zero for finite vertices.  Split 1 wins when its
// alignment is strictly larger.
__device__ __forceinline__ bool split_first_by_normals(const V3 v[4])
{
    const V3 n0 = cross(sub(v[1], v[0]), sub(v[2], v[0]));
    const V3 n1 = cross(sub(v[2], v[1]), sub(v[0], v[1]));
    const float align0 = abs_dot(n0, n1);
    const V3 m0 = cross(sub(v[1], v[0]), sub(v[3], v[0]));
    const V3 m1 = cross(sub(v[3], v[1]), sub(v[3], v[1]));
    const float align1 = abs_dot(m0, m1);
    return align0 > align1;
}

// product of two split weights rounded to the weights' own dtype (products
// of two float16 / bfloat16 values are exact in float32, so one rounding
// reproduces the dtype's multiply)
__device__ __forceinline__ float weight_product(float a, float b, int wtype)
{
    const float p = a * b;
    if (wtype == 1)
        return __half2float(__float2half_rn(p));
    if (wtype == 2)
        return __bfloat162float(__float2bfloat16(p));
    return p;
}

template <int MODE>
__global__ __launch_bounds__(HB) void k_dm_emit(DualMeshArgs a, int wtype)
{
    const int64_t i = (int64_t)blockIdx.x * HB + threadIdx.x;
    if (i >= a.N)
        return;
    const uint32_t m = a.mask[i];
    if (!m)
        return;
    const int x = a.coords[3 * i], y = a.coords[3 * i + 1], z = a.coords[3 * i + 2];
    uint64_t j = a.off[i];
    for (int ax = 0; ax < 3; ++ax)
    {
        if (!(m >> ax & 1u))
            continue;
        uint32_t q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            q[k] = find_voxel<MODE>(a, x, y, z, ax, k);
        if (a.quads)
            *(int4 *)(a.quads + 4 * j) = make_int4((int)q[0], (int)q[1], (int)q[2], (int)q[3]);
        if (a.faces)
        {
            bool first;
            if (a.split_w)
            {
                const float w02 = weight_product(a.split_w[q[0]], a.split_w[q[2]], wtype);
                const float w13 = weight_product(a.split_w[q[1]], a.split_w[q[3]], wtype);
                first = w02 > w13;
            }
            else
            {
                V3 v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    v[k] = {a.verts[3 * (uint64_t)q[k]], a.verts[3 * (uint64_t)q[k] + 1],
                            a.verts[3 * (uint64_t)q[k] + 2]};
                first = split_first_by_normals(v);
            }
            int32_t *f = a.faces + 6 * j;
            if (first)
            {
                f[0] = (int)q[0], f[1] = (int)q[1], f[2] = (int)q[2];
                f[3] = (int)q[0], f[4] = (int)q[2], f[5] = (int)q[3];
            }
            else
            {
                f[0] = (int)q[0], f[1] = (int)q[1], f[2] = (int)q[3];
                f[3] = (int)q[3], f[4] = (int)q[1], f[5] = (int)q[2];
            }
        }
        ++j;
    }
}

} // namespace

hipError_t launch_hash_insert(void *tkeys, int key_bytes, void *tvals, int val_bytes, uint64_t cap, int src,
                              const void *keys, const void *vals, uint64_t n, int W, int H, int D, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    if (key_bytes == 4)
        return val_bytes == 4 ? hash_insert_t<uint32_t, uint32_t>(tkeys, tvals, cap, src, keys, vals, n, W, H, D, s)
                              : hash_insert_t<uint32_t, uint64_t>(tkeys, tvals, cap, src, keys, vals, n, W, H, D, s);
    return val_bytes == 4 ? hash_insert_t<uint64_t, uint32_t>(tkeys, tvals, cap, src, keys, vals, n, W, H, D, s)
                          : hash_insert_t<uint64_t, uint64_t>(tkeys, tvals, cap, src, keys, vals, n, W, H, D, s);
}

hipError_t launch_hash_lookup(const void *tkeys, int key_bytes, const void *tvals, int val_bytes, uint64_t cap, int src,
                              const void *keys, void *out, uint64_t n, int W, int H, int D, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    if (key_bytes == 4)
        return val_bytes == 4 ? hash_lookup_t<uint32_t, uint32_t>(tkeys, tvals, cap, src, keys, out, n, W, H, D, s)
                              : hash_lookup_t<uint32_t, uint64_t>(tkeys, tvals, cap, src, keys, out, n, W, H, D, s);
    return val_bytes == 4 ? hash_lookup_t<uint64_t, uint32_t>(tkeys, tvals, cap, src, keys, out, n, W, H, D, s)
                          : hash_lookup_t<uint64_t, uint64_t>(tkeys, tvals, cap, src, keys, out, n, W, H, D, s);
}

#define DM_DISPATCH(kern, g, ...)                                                                                   \
    do                                                                                                             \
    {                                                                                                              \
        if (a.mode == DM_DIRECT)                                                                                   \
            hipLaunchKernelGGL(kern<DM_DIRECT>, g, dim3(HB), 0, s, __VA_ARGS__);                                    \
        else if (a.mode == DM_WIDE)                                                                                \
            hipLaunchKernelGGL(kern<DM_WIDE>, g, dim3(HB), 0, s, __VA_ARGS__);                                      \
        else                                                                                                       \
            hipLaunchKernelGGL(kern<DM_NARROW>, g, dim3(HB), 0, s, __VA_ARGS__);                                    \
    } while (0)

hipError_t launch_dm_clear(const DualMeshArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_dm_clear, dim3(grid_of(a.cap, 1u << 16)), dim3(HB), 0, s, a.table, a.cap,
                       a.mode == DM_WIDE ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_dm_restore(uint32_t *table, const uint32_t *touched, int64_t n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_dm_restore, dim3((unsigned)((n + HB - 1) / HB)), dim3(HB), 0, s, table, touched, n);
    return hipGetLastError();
}

hipError_t launch_dm_insert(const DualMeshArgs &a, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    const dim3 g((unsigned)((a.N + HB - 1) / HB));
    DM_DISPATCH(k_dm_insert, g, a);
    return hipGetLastError();
}

hipError_t launch_dm_count(const DualMeshArgs &a, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    const dim3 g((unsigned)((a.N + HB - 1) / HB));
    DM_DISPATCH(k_dm_count, g, a);
    return hipGetLastError();
}

hipError_t launch_dm_emit(const DualMeshArgs &a, int wtype, hipStream_t s)
{
    if (a.N == 0)
        return hipSuccess;
    const dim3 g((unsigned)((a.N + HB - 1) / HB));
    DM_DISPATCH(k_dm_emit, g, a, wtype);
    return hipGetLastError();
}

} // namespace ovx
