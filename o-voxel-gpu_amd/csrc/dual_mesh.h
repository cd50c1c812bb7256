// Launchers of the dual-grid -> mesh kernels and of the generic hashmap
// kernels (dual_mesh.hip).  Host orchestration lives in fdg_capi.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ovx
{

// Generic open-addressing table in the layout of the reference's hashmap
// (src/hash/hash.cuh): keys[cap] (u32 or u64, all ones = empty) and
// values[cap] (u32 or u64) as separate arrays, Murmur3 finalizer modulo cap,
// linear probing.  `src` selects the keys: HASH_KEYS reads keys[n];
// HASH_COORDS reads int32 coords[n][4] = (b, x, y, z) and forms
// b*W*H*D + x*H*D + y*D + z.  Values come from vals[n], or are the row index
// when vals is null.
enum HashSrc
{
    HASH_KEYS = 0,
    HASH_COORDS = 1
};

hipError_t launch_hash_insert(void *tkeys, int key_bytes, void *tvals, int val_bytes, uint64_t cap, int src,
                              const void *keys, const void *vals, uint64_t n, int W, int H, int D, hipStream_t s);
hipError_t launch_hash_lookup(const void *tkeys, int key_bytes, const void *tvals, int val_bytes, uint64_t cap, int src,
                              const void *keys, void *out, uint64_t n, int W, int H, int D, hipStream_t s);

// Dual grid -> mesh.  The voxel table is private to the library:
//   DM_DIRECT   grids of at most DM_DIRECT_MAX_CELLS cells: one u32 per grid
//               cell (row + 1, 0 = empty), x fastest, kept all-zero between
//               calls by clearing the cells a call set (touched list)
//   DM_NARROW   open addressing with packed 8-byte entries (u32 key << 32 |
//               row), grids of fewer than 2^32 cells
//   DM_WIDE     open addressing with 16-byte entries (u64 key, u64 row)
// Hash capacities are a power of two >= 2N.
enum DmMode
{
    DM_NARROW = 0,
    DM_WIDE = 1,
    DM_DIRECT = 2
};
constexpr uint64_t DM_DIRECT_MAX_CELLS = 1ull << 30;

struct DualMeshArgs
{
    const int32_t *coords; // [N][3]
    const float *dual;     // [N][3], voxel units
    const uint8_t *flags;  // [N][3], nonzero = intersected grid edge
    const float *split_w;  // [N] or null
    int64_t N;
    int W, H, D;   // grid size (key dimensions and lookup bounds)
    float vs[3];   // voxel size
    float a0[3];   // aabb minimum
    int mode;      // DmMode
    uint64_t cap;  // hash table entries (power of two); DM_DIRECT: cells
    void *table;
    uint32_t *touched;    // DM_DIRECT: [N] cell set by each row (all ones: none)
    uint32_t *cnt;        // [N] quads of each voxel
    uint8_t *mask;        // [N] bit a: the quad around axis a exists
    const uint32_t *off;  // [N] exclusive scan of cnt
    float *verts;         // [N][3] mesh vertices
    int32_t *faces;       // [2L][3] or null
    int32_t *quads;       // [L][4] or null
};

hipError_t launch_dm_clear(const DualMeshArgs &a, hipStream_t s);
// DM_DIRECT: zero the cells listed in touched[0, n)
hipError_t launch_dm_restore(uint32_t *table, const uint32_t *touched, int64_t n, hipStream_t s);
hipError_t launch_dm_insert(const DualMeshArgs &a, hipStream_t s);
hipError_t launch_dm_count(const DualMeshArgs &a, hipStream_t s);
// wtype: rounding of split-weight products (0 float32, 1 float16, 2 bfloat16)
hipError_t launch_dm_emit(const DualMeshArgs &a, int wtype, hipStream_t s);

} // namespace ovx
