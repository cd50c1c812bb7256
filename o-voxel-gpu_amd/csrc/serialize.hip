// 30-bit Z-order (Morton) and Hilbert codes of voxel coordinates, for MI355X
// (gfx950).  Reference: src/serialize/z_order.cu, src/serialize/hilbert.cu
// (bound src/ext.cpp:24-31, called by o_voxel/serialize.py).
//
// Codes are a storage format (they order the voxels of a .vxz file), so the
// bit manipulation follows the reference's operation for operation,
// including what it does with coordinates outside [0, 1024).  The kernels
// read and write the coordinates through element strides: the columns of an
// [N, 3] int32 / int64 tensor are used in place, in any permutation, with no
// per-column copies.
#include <hip/hip_runtime.h>

#include "serialize.h"

namespace ovx
{
namespace
{

constexpr int SB = 256;

// 10-bit value -> 30 bits with two zeros after each bit (multiply-and-mask
// spreading, the reference's expandBits)
__device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// inverse of spread3 on bits 0, 3, 6, ... (the reference's extractBits)
__device__ __forceinline__ uint32_t gather3(uint32_t v)
{
    v &= 0x49249249u;
    v = (v ^ (v >> 2)) & 0x030C30C3u;
    v = (v ^ (v >> 4)) & 0x0300F00Fu;
    v = (v ^ (v >> 8)) & 0x030000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}

__device__ __forceinline__ uint32_t interleave(uint32_t a, uint32_t b, uint32_t c)
{
    return spread3(a) * 4u + spread3(b) * 2u + spread3(c);
}

// Skilling's Hilbert transform (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 2004), axes -> transposed index, 10 bits,
// in the reference's loop order (hilbert.cu encode)
__device__ __forceinline__ uint32_t hilbert_code(uint32_t p0, uint32_t p1, uint32_t p2)
{
    uint32_t p[3] = {p0, p1, p2};
    for (uint32_t q = 1u << 9; q > 1; q >>= 1)
    {
        const uint32_t m = q - 1;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            if (p[i] & q)
                p[0] ^= m;
            else
            {
                const uint32_t t = (p[0] ^ p[i]) & m;
                p[0] ^= t;
                p[i] ^= t;
            }
        }
    }
    p[1] ^= p[0];
    p[2] ^= p[1];
    uint32_t t = 0;
    for (uint32_t q = 1u << 9; q > 1; q >>= 1)
        if (p[2] & q)
            t ^= q - 1;
    return interleave(p[0] ^ t, p[1] ^ t, p[2] ^ t);
}

// transposed index -> axes (hilbert.cu decode)
__device__ __forceinline__ void hilbert_point(uint32_t code, uint32_t p[3])
{
    p[0] = gather3(code >> 2);
    p[1] = gather3(code >> 1);
    p[2] = gather3(code);
    const uint32_t t = p[2] >> 1;
    p[2] ^= p[1];
    p[1] ^= p[0];
    p[0] ^= t;
    for (uint32_t q = 2; q != (2u << 9); q <<= 1)
    {
        const uint32_t m = q - 1;
#pragma unroll
        for (int i = 2; i >= 0; --i)
        {
            if (p[i] & q)
                p[0] ^= m;
            else
            {
                const uint32_t u = (p[0] ^ p[i]) & m;
                p[0] ^= u;
                p[i] ^= u;
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ uint32_t load_coord(const void *base, int64_t stride, int64_t i)
{
    return (uint32_t)((const T *)base)[i * stride]; // int64 -> low 32 bits, as torch's .int()
}

template <typename T, int MODE>
__global__ __launch_bounds__(SB) void k_seq_encode(SeqCols in, int64_t n, int32_t *codes)
{
    const int64_t stride = (int64_t)gridDim.x * SB;
    for (int64_t i = (int64_t)blockIdx.x * SB + threadIdx.x; i < n; i += stride)
    {
        const uint32_t a = load_coord<T>(in.p[0], in.stride[0], i);
        const uint32_t b = load_coord<T>(in.p[1], in.stride[1], i);
        const uint32_t c = load_coord<T>(in.p[2], in.stride[2], i);
        codes[i] = (int32_t)(MODE == SEQ_Z_ORDER ? interleave(a, b, c) : hilbert_code(a, b, c));
    }
}

template <int MODE>
__global__ __launch_bounds__(SB) void k_seq_decode(const int32_t *codes, int64_t n, SeqCols out)
{
    const int64_t stride = (int64_t)gridDim.x * SB;
    for (int64_t i = (int64_t)blockIdx.x * SB + threadIdx.x; i < n; i += stride)
    {
        const uint32_t c = (uint32_t)codes[i];
        uint32_t p[3];
        if (MODE == SEQ_Z_ORDER)
        {
            p[0] = gather3(c >> 2);
            p[1] = gather3(c >> 1);
            p[2] = gather3(c);
        }
        else
            hilbert_point(c, p);
#pragma unroll
        for (int k = 0; k < 3; ++k)
            ((int32_t *)out.p[k])[i * out.stride[k]] = (int32_t)p[k];
    }
}

unsigned blocks_for(int64_t n)
{
    int64_t b = (n + SB - 1) / SB;
    if (b > 65536)
        b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

} // namespace

hipError_t launch_seq_encode(const SeqCols &in, int elem_bytes, int64_t n, int mode, int32_t *codes, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    const dim3 g(blocks_for(n)), b(SB);
    if (elem_bytes == 8)
    {
        if (mode == SEQ_Z_ORDER)
            hipLaunchKernelGGL((k_seq_encode<int64_t, SEQ_Z_ORDER>), g, b, 0, s, in, n, codes);
        else
            hipLaunchKernelGGL((k_seq_encode<int64_t, SEQ_HILBERT>), g, b, 0, s, in, n, codes);
    }
    else
    {
        if (mode == SEQ_Z_ORDER)
            hipLaunchKernelGGL((k_seq_encode<int32_t, SEQ_Z_ORDER>), g, b, 0, s, in, n, codes);
        else
            hipLaunchKernelGGL((k_seq_encode<int32_t, SEQ_HILBERT>), g, b, 0, s, in, n, codes);
    }
    return hipGetLastError();
}

hipError_t launch_seq_decode(const int32_t *codes, int64_t n, int mode, const SeqCols &out, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    const dim3 g(blocks_for(n)), b(SB);
    if (mode == SEQ_Z_ORDER)
        hipLaunchKernelGGL((k_seq_decode<SEQ_Z_ORDER>), g, b, 0, s, codes, n, out);
    else
        hipLaunchKernelGGL((k_seq_decode<SEQ_HILBERT>), g, b, 0, s, codes, n, out);
    return hipGetLastError();
}

} // namespace ovx
