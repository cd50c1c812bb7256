// Sparse voxel octrees of the .vxz format and its attribute filters, for
// MI355X (gfx950).  Reference: src/io/svo.cpp, src/io/filter_parent.cpp,
// src/io/filter_neighbor.cpp (sequential CPU code, one chunk per call).
//
// Encoding is parallel over the voxels of all chunks at once.  An octree is
// written in pre-order; with the codes ascending, the bytes a code opens are
// the nodes below the first level where it leaves the previous code's path,
// so an exclusive scan of "nodes opened per code" places every byte, and a
// code's child bit in a node it did not open goes to the node opened by the
// first code with the same prefix (a binary search over the chunk's codes).
// The parent filter's rows follow from a scan of (children - 1) over the
// bytes.  The neighbour filter uses a direct-mapped cell table per chunk;
// its inverse (each value = delta + its lower neighbour's value) is solved
// by pointer jumping.  Parsing a pre-order stream (reading) is sequential
// and stays on the host, beside the decompression that bounds it.
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "svo.h"

namespace ovx
{
namespace
{

constexpr int VB = 256;

unsigned blocks(int64_t n)
{
    const int64_t b = (n + VB - 1) / VB;
    return (unsigned)(b < 1 ? 1 : b);
}

__device__ __forceinline__ uint32_t digit_of(uint32_t code, int D, int j) { return (code >> (3 * (D - 1 - j))) & 7u; }

// first index a in [lo, hi) with codes[a] >= key (codes ascending)
__device__ __forceinline__ int64_t lower_bound(const int32_t *codes, int64_t lo, int64_t hi, uint32_t key)
{
    while (lo < hi)
    {
        const int64_t mid = (lo + hi) >> 1;
        if ((uint32_t)codes[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// pre-order position of the level-j node holding code i (j >= 1): opened by
// the first code with the same j-digit prefix
__device__ __forceinline__ uint32_t node_pos(const SvoCodes &c, const uint8_t *insf, const uint32_t *pos, int64_t lo,
                                             int64_t i, uint32_t code, int j)
{
    if (j == 0)
        return pos[lo];
    const int sh = 3 * (c.depth - j);
    const uint32_t key = (code >> sh) << sh;
    const int64_t a = lower_bound(c.codes, lo, i + 1, key);
    // codes that are not ascending could name a code that did not open this
    // level; stay inside that code's own bytes
    if (a != lo && insf[a] >= j)
        return pos[a];
    return pos[a] + (a == lo ? 1u : 0u) + (uint32_t)(j - (a == lo ? 0 : insf[a]) - 1);
}

__device__ __forceinline__ bool same_as_prev(const SvoCodes &c, int64_t i, int64_t lo, uint32_t mask)
{
    return i > lo && (((uint32_t)c.codes[i] ^ (uint32_t)c.codes[i - 1]) & mask) == 0;
}

__global__ __launch_bounds__(VB) void k_svo_prep(SvoCodes c, uint8_t *insf, uint32_t *k)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= c.n)
        return;
    const int D = c.depth;
    const int64_t lo = c.off[c.chunk[i]];
    if (i == lo)
    {
        insf[i] = 0;
        k[i] = (uint32_t)D; // the chunk's root and one node per level below it
        return;
    }
    const uint32_t x = ((uint32_t)c.codes[i] ^ (uint32_t)c.codes[i - 1]) & ((1u << (3 * D)) - 1u);
    int f = D - 1; // equal digits (a repeated code): only the leaf bit, already set
    if (x)
    {
        const int b = 31 - __clz(x);
        f = D - 1 - b / 3;
    }
    insf[i] = (uint8_t)f;
    k[i] = (uint32_t)(D - 1 - f);
}

__global__ __launch_bounds__(VB) void k_svo_bits(SvoCodes c, const uint8_t *insf, const uint32_t *pos, uint8_t *svo)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= c.n)
        return;
    const int D = c.depth;
    const int64_t lo = c.off[c.chunk[i]];
    if (same_as_prev(c, i, lo, (1u << (3 * D)) - 1u))
        return;
    const uint32_t code = (uint32_t)c.codes[i];
    const int f = insf[i];
    const uint32_t base = pos[i] + (i == lo ? 1u : 0u);
    for (int j = f; j < D; ++j)
    {
        const uint32_t p = j > f ? base + (uint32_t)(j - f - 1) : node_pos(c, insf, pos, lo, i, code, j);
        const uint32_t bit = (1u << digit_of(code, D, j)) << (8 * (p & 3u));
        atomicOr((unsigned int *)(svo + (p & ~3u)), bit);
    }
}

__global__ __launch_bounds__(VB) void k_svo_chunk_offsets(SvoCodes c, const uint32_t *pos, uint64_t total,
                                                          int64_t *chunk_off)
{
    const int64_t k = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (k > c.n_chunks)
        return;
    const int64_t first = c.off[k];
    chunk_off[k] = (k == c.n_chunks || first >= c.n) ? (int64_t)total : (int64_t)pos[first];
}

__global__ __launch_bounds__(VB) void k_svo_rows(const uint8_t *svo, uint64_t n, uint32_t *rows)
{
    const uint64_t i = (uint64_t)blockIdx.x * VB + threadIdx.x;
    if (i < n)
        rows[i] = (uint32_t)__popc(svo[i]) - 1u;
}

__global__ __launch_bounds__(VB) void k_svo_parent_encode(SvoCodes c, const uint8_t *insf, const uint32_t *pos,
                                                          const uint8_t *svo, const uint32_t *erow,
                                                          const uint8_t *attr, int C, uint8_t *delta)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= c.n)
        return;
    const int D = c.depth;
    const int64_t lo = c.off[c.chunk[i]];
    if (i == lo)
    {
        // the chunk's first row: the root's value, its first leaf's attributes
        for (int ch = 0; ch < C; ++ch)
            delta[lo * C + ch] = attr[i * C + ch];
        return;
    }
    if (same_as_prev(c, i, lo, (1u << (3 * D)) - 1u))
        return;
    // code i opens a non-first child at level j of an existing node: that
    // child's value minus its previous sibling's (their first leaves)
    const uint32_t code = (uint32_t)c.codes[i];
    const int j = insf[i];
    const uint32_t d = digit_of(code, D, j);
    const uint32_t q = node_pos(c, insf, pos, lo, i, code, j);
    const uint32_t below = (uint32_t)svo[q] & ((1u << d) - 1u);
    if (!below) // only for codes that are not ascending
        return;
    const uint32_t rank = (uint32_t)__popc(below);
    const uint32_t dprev = 31u - __clz(below);
    const int sh = 3 * (D - 1 - j);
    const uint32_t key = ((((code >> sh) >> 3) << 3) | dprev) << sh;
    const int64_t ip = lower_bound(c.codes, lo, i, key);
    const int64_t row = lo + 1 + (int64_t)(erow[q] - erow[pos[lo]]) + (int64_t)rank - 1;
    if (row <= lo || row >= c.off[c.chunk[i] + 1])
        return;
    for (int ch = 0; ch < C; ++ch)
        delta[row * C + ch] = (uint8_t)(attr[i * C + ch] - attr[ip * C + ch]);
}

// ---------------------------------------------------------------------------
// neighbour filter
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool nb_valid(const NbArgs &a, int64_t chunk, int x, int y, int z)
{
    return chunk >= 0 && chunk < a.n_chunks && (unsigned)x < (unsigned)a.res && (unsigned)y < (unsigned)a.res &&
           (unsigned)z < (unsigned)a.res;
}

__device__ __forceinline__ uint64_t nb_cell(const NbArgs &a, int64_t chunk, int x, int y, int z)
{
    const uint64_t r = (uint64_t)a.res;
    return (uint64_t)chunk * r * r * r + ((uint64_t)z * r + (uint64_t)y) * r + (uint64_t)x;
}

__global__ __launch_bounds__(VB) void k_nb_insert(NbArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= a.n)
        return;
    const int x = a.coords[3 * i], y = a.coords[3 * i + 1], z = a.coords[3 * i + 2];
    if (nb_valid(a, a.chunk[i], x, y, z))
        atomicMax((unsigned int *)a.table + nb_cell(a, a.chunk[i], x, y, z), (unsigned int)i + 1u);
}

// lower neighbours, highest priority first (the reference takes the last
// present of x, y, z, xy, xz, yz, xyz)
__constant__ int c_nb_off[7][3] = {{-1, -1, -1}, {0, -1, -1}, {-1, 0, -1}, {-1, -1, 0},
                                   {0, 0, -1},   {0, -1, 0},  {-1, 0, 0}};

__global__ __launch_bounds__(VB) void k_nb_link(NbArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= a.n)
        return;
    const int x = a.coords[3 * i], y = a.coords[3 * i + 1], z = a.coords[3 * i + 2];
    const int64_t ck = a.chunk[i];
    if (!nb_valid(a, ck, x, y, z))
    {
        // outside its chunk: kept as is, with no neighbour
        a.own[i] = (int32_t)i;
        a.nb[i] = -1;
        return;
    }
    a.own[i] = (int32_t)a.table[nb_cell(a, ck, x, y, z)] - 1;
    int32_t r = -1;
    for (int k = 0; k < 7 && r < 0; ++k)
    {
        const int xx = x + c_nb_off[k][0], yy = y + c_nb_off[k][1], zz = z + c_nb_off[k][2];
        if (xx < 0 || yy < 0 || zz < 0)
            continue;
        r = (int32_t)a.table[nb_cell(a, ck, xx, yy, zz)] - 1;
    }
    a.nb[i] = r;
}

__global__ __launch_bounds__(VB) void k_nb_encode(NbArgs a, const uint8_t *attr, int C, uint8_t *delta)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= a.n)
        return;
    const int64_t o = a.own[i], r = a.nb[i];
    for (int ch = 0; ch < C; ++ch)
        delta[i * C + ch] = (uint8_t)(attr[o * C + ch] - (r >= 0 ? attr[r * C + ch] : 0));
}

__global__ __launch_bounds__(VB) void k_nb_decode_init(NbArgs a, const uint8_t *delta, int C, uint8_t *acc,
                                                       int32_t *jmp)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= a.n)
        return;
    const int64_t o = a.own[i];
    for (int ch = 0; ch < C; ++ch)
        acc[i * C + ch] = delta[o * C + ch];
    jmp[i] = a.nb[i];
}

// value(i) = delta(own(i)) + value(nb(i)): each round adds the value at the
// current jump target and doubles the jump
__global__ __launch_bounds__(VB) void k_nb_jump(int64_t n, int C, const uint8_t *acc_in, const int32_t *jmp_in,
                                                uint8_t *acc_out, int32_t *jmp_out)
{
    const int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x;
    if (i >= n)
        return;
    const int32_t j = jmp_in[i];
    for (int ch = 0; ch < C; ++ch)
        acc_out[i * C + ch] = (uint8_t)(acc_in[i * C + ch] + (j >= 0 ? acc_in[(int64_t)j * C + ch] : 0));
    jmp_out[i] = j >= 0 ? jmp_in[j] : -1;
}

} // namespace

hipError_t launch_svo_prep(const SvoCodes &c, uint8_t *insf, uint32_t *k, hipStream_t s)
{
    if (c.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_svo_prep, dim3(blocks(c.n)), dim3(VB), 0, s, c, insf, k);
    return hipGetLastError();
}

hipError_t launch_svo_bits(const SvoCodes &c, const uint8_t *insf, const uint32_t *pos, uint8_t *svo, hipStream_t s)
{
    if (c.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_svo_bits, dim3(blocks(c.n)), dim3(VB), 0, s, c, insf, pos, svo);
    return hipGetLastError();
}

hipError_t launch_svo_chunk_offsets(const SvoCodes &c, const uint32_t *pos, uint64_t total, int64_t *chunk_off,
                                    hipStream_t s)
{
    hipLaunchKernelGGL(k_svo_chunk_offsets, dim3(blocks(c.n_chunks + 1)), dim3(VB), 0, s, c, pos, total, chunk_off);
    return hipGetLastError();
}

hipError_t launch_svo_rows(const uint8_t *svo, uint64_t n_bytes, uint32_t *rows, hipStream_t s)
{
    if (n_bytes == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_svo_rows, dim3(blocks((int64_t)n_bytes)), dim3(VB), 0, s, svo, n_bytes, rows);
    return hipGetLastError();
}

hipError_t launch_svo_parent_encode(const SvoCodes &c, const uint8_t *insf, const uint32_t *pos, const uint8_t *svo,
                                    const uint32_t *erow, const uint8_t *attr, int C, uint8_t *delta, hipStream_t s)
{
    if (c.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_svo_parent_encode, dim3(blocks(c.n)), dim3(VB), 0, s, c, insf, pos, svo, erow, attr, C,
                       delta);
    return hipGetLastError();
}

hipError_t launch_nb_insert(const NbArgs &a, hipStream_t s)
{
    if (a.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_nb_insert, dim3(blocks(a.n)), dim3(VB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_nb_link(const NbArgs &a, hipStream_t s)
{
    if (a.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_nb_link, dim3(blocks(a.n)), dim3(VB), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_nb_encode(const NbArgs &a, const uint8_t *attr, int C, uint8_t *delta, hipStream_t s)
{
    if (a.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_nb_encode, dim3(blocks(a.n)), dim3(VB), 0, s, a, attr, C, delta);
    return hipGetLastError();
}

hipError_t launch_nb_decode(const NbArgs &a, const uint8_t *delta, int C, uint8_t *acc0, uint8_t *acc1, int32_t *jmp0,
                            int32_t *jmp1, uint8_t *attr, hipStream_t s)
{
    if (a.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_nb_decode_init, dim3(blocks(a.n)), dim3(VB), 0, s, a, delta, C, acc0, jmp0);
    // a chain of lower neighbours has at most 3 (res - 1) + 1 links
    int rounds = 1;
    while ((1ll << (rounds - 1)) < 3ll * a.res)
        ++rounds;
    uint8_t *ai = acc0, *ao = acc1;
    int32_t *ji = jmp0, *jo = jmp1;
    for (int r = 0; r < rounds; ++r)
    {
        hipLaunchKernelGGL(k_nb_jump, dim3(blocks(a.n)), dim3(VB), 0, s, a.n, C, ai, ji, ao, jo);
        std::swap(ai, ao);
        std::swap(ji, jo);
    }
    return hipMemcpyAsync(attr, ai, (size_t)a.n * C, hipMemcpyDeviceToDevice, s);
}

// ---------------------------------------------------------------------------
// host parsers
// ---------------------------------------------------------------------------
int64_t svo_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, int32_t *codes, int64_t capacity)
{
    if (n_bytes <= 0 || depth < 1 || depth > 10)
        return -1;
    uint8_t node[10];
    int bit[10];
    uint32_t prefix[10];
    int64_t pos = 0, out = 0;
    int level = 0;
    node[0] = svo[pos++];
    bit[0] = 0;
    prefix[0] = 0;
    while (level >= 0)
    {
        int b = bit[level];
        while (b < 8 && !(node[level] >> b & 1))
            ++b;
        if (b == 8)
        {
            --level;
            continue;
        }
        bit[level] = b + 1;
        const uint32_t child = (prefix[level] << 3) | (uint32_t)b;
        if (level == depth - 1)
        {
            if (out >= capacity)
                return -2;
            codes[out++] = (int32_t)child;
        }
        else
        {
            if (pos >= n_bytes)
                return -3; // truncated stream
            ++level;
            node[level] = svo[pos++];
            bit[level] = 0;
            prefix[level] = child;
        }
    }
    return out;
}

int64_t svo_parent_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, const uint8_t *delta, int64_t n, int C,
                               uint8_t *attr)
{
    if (n_bytes <= 0 || depth < 1 || depth > 10 || C <= 0 || n <= 0)
        return -1;
    std::vector<uint8_t> val((size_t)(depth + 1) * C); // running child value per level
    uint8_t node[10];
    int bit[10], idx[10];
    int64_t base[10];
    int64_t pos = 0, row = 1, leaf = 0;
    int level = 0;
    node[0] = svo[pos++];
    bit[0] = idx[0] = 0;
    base[0] = row;
    row += __builtin_popcount(node[0]) - 1;
    for (int ch = 0; ch < C; ++ch)
        val[ch] = delta[ch];
    while (level >= 0)
    {
        int b = bit[level];
        while (b < 8 && !(node[level] >> b & 1))
            ++b;
        if (b == 8)
        {
            --level;
            continue;
        }
        bit[level] = b + 1;
        uint8_t *v = &val[(size_t)level * C];
        if (idx[level] > 0)
        {
            const int64_t r = base[level] + idx[level] - 1;
            if (r >= n)
                return -2;
            for (int ch = 0; ch < C; ++ch)
                v[ch] = (uint8_t)(v[ch] + delta[r * C + ch]);
        }
        ++idx[level];
        if (level == depth - 1)
        {
            if (leaf >= n)
                return -2;
            for (int ch = 0; ch < C; ++ch)
                attr[leaf * C + ch] = v[ch];
            ++leaf;
        }
        else
        {
            if (pos >= n_bytes)
                return -3;
            ++level;
            node[level] = svo[pos++];
            bit[level] = idx[level] = 0;
            base[level] = row;
            row += __builtin_popcount(node[level]) - 1;
            for (int ch = 0; ch < C; ++ch)
                val[(size_t)level * C + ch] = v[ch];
        }
    }
    return leaf;
}

} // namespace ovx
