// Launchers of the sparse-voxel-octree kernels (svo.hip) and the host-side
// pre-order parsers of the .vxz reader.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ovx
{

// Chunked code list: codes[n] ascending inside each chunk, chunk[n] the chunk
// of each code, off[n_chunks + 1] the first code of each chunk (device).
struct SvoCodes
{
    const int32_t *codes;
    const int32_t *chunk;
    const int64_t *off;
    int64_t n;
    int n_chunks;
    int depth; // octree levels above the leaves (log2 of the chunk size)
};

// per code: first level where it leaves the previous code's path (insf) and
// the number of octree bytes it opens (k, the chunk's root counted on its
// first code)
hipError_t launch_svo_prep(const SvoCodes &c, uint8_t *insf, uint32_t *k, hipStream_t s);
// child-mask bytes (OR of the child bits) at the pre-order positions given
// by pos = exclusive scan of k; svo must be zeroed and 4-byte aligned
hipError_t launch_svo_bits(const SvoCodes &c, const uint8_t *insf, const uint32_t *pos, uint8_t *svo, hipStream_t s);
// byte offset of each chunk's octree: chunk_off[c] = pos[off[c]], chunk_off[n_chunks] = total
hipError_t launch_svo_chunk_offsets(const SvoCodes &c, const uint32_t *pos, uint64_t total, int64_t *chunk_off,
                                    hipStream_t s);
// popcount(byte) - 1 per octree byte (the rows a node owns in the parent filter)
hipError_t launch_svo_rows(const uint8_t *svo, uint64_t n_bytes, uint32_t *rows, hipStream_t s);
// parent filter: delta rows from sibling differences (erow = exclusive scan of rows)
hipError_t launch_svo_parent_encode(const SvoCodes &c, const uint8_t *insf, const uint32_t *pos, const uint8_t *svo,
                                    const uint32_t *erow, const uint8_t *attr, int C, uint8_t *delta, hipStream_t s);

// neighbour filter over chunk-local coordinates; table: n_chunks * res^3
// zeroed u32 cells
struct NbArgs
{
    const int32_t *coords; // [n][3] in [0, res)
    const int32_t *chunk;  // [n]
    int64_t n;
    int n_chunks;
    int res;
    uint32_t *table;
    int32_t *own; // [n] row holding each voxel's cell
    int32_t *nb;  // [n] chosen lower neighbour's row or -1
};
hipError_t launch_nb_insert(const NbArgs &a, hipStream_t s);
hipError_t launch_nb_link(const NbArgs &a, hipStream_t s);
hipError_t launch_nb_encode(const NbArgs &a, const uint8_t *attr, int C, uint8_t *delta, hipStream_t s);
// decode by pointer jumping: rounds enough for chains of 3 * res links
hipError_t launch_nb_decode(const NbArgs &a, const uint8_t *delta, int C, uint8_t *acc0, uint8_t *acc1, int32_t *jmp0,
                            int32_t *jmp1, uint8_t *attr, hipStream_t s);

// host parsers (pre-order streams are parsed sequentially)
int64_t svo_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, int32_t *codes, int64_t capacity);
int64_t svo_parent_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, const uint8_t *delta, int64_t n, int C,
                               uint8_t *attr);

} // namespace ovx
