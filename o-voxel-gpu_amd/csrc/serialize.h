// Launchers of the Z-order / Hilbert code kernels (serialize.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ovx
{

enum SeqMode
{
    SEQ_Z_ORDER = 0,
    SEQ_HILBERT = 1
};

// three coordinate columns, each a base pointer and an element stride
struct SeqCols
{
    const void *p[3];
    int64_t stride[3];
};

// codes[i] from the columns (int32 or int64 elements, elem_bytes 4 / 8)
hipError_t launch_seq_encode(const SeqCols &in, int elem_bytes, int64_t n, int mode, int32_t *codes, hipStream_t s);
// columns (int32) from codes[i]
hipError_t launch_seq_decode(const int32_t *codes, int64_t n, int mode, const SeqCols &out, hipStream_t s);

} // namespace ovx
