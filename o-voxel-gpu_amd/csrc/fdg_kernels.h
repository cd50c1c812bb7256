// Kernel launch interface of the flexible-dual-grid library (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fdg_device.h"

namespace ovx
{

// faces per tile of the face-test kernel (a block of BT threads takes one
// tile): small tiles when faces cover many voxels (more, shorter blocks),
// larger ones for sub-voxel faces (less per-block overhead)
// (face, axis) items per tile of the scanline kernel (a block of BT threads
// takes one tile): small tiles spread meshes with large triangles over more
// blocks, large ones keep the per-tile merging of occupancy bits for meshes
// of many small triangles
constexpr int N_MARK_TILES = 3;
constexpr int MARK_TILES[N_MARK_TILES] = {64, 128, 256};
// face-test launch classes: blocks of 32 or 64 faces walked cooperatively,
// or 256 faces walked one per lane (k_face_lane)
constexpr int N_FACE_TILES = 3;
constexpr int FACE_TILES[N_FACE_TILES] = {32, 64, 256};
constexpr int FACE_LANE_CLASS = 2;

// device counters (unsigned long long each)
enum
{
    CTR_TINY = 0,        // bricks listed for the first fold tier (one wave per brick)
    CTR_ERRORS = 1,
    CTR_WORDS_TOTAL = 2, // total of the bitset popcount scan (= N)
    CTR_OVER = 3,        // a record-stage buffer of the conversion overflowed (the fold and solve stages skip it)
    CTR_MEDIUM = 4,      // bricks left to the third fold tier
    CTR_GIANT = 5,
    CTR_EDGE_TOTAL = 6,  // half-edge count scan total
    CTR_BEDGE_TOTAL = 7, // boundary edges
    CTR_ACTIVE = 8,      // bricks with records
    CTR_CHUNK_SCAN = 9,  // total of the per-brick chunk-count scan
    CTR_LARGE = 10,      // bricks left to the fold kernel for large bricks
    CTR_LARGE_RECS = 11, // their records
    CTR_DM_QUADS = 12,   // dual grid -> mesh: quads (scan total)
    CTR_SVO_BYTES = 13,  // octree bytes of an encode (scan total)
    CTR_HUGE = 14,       // bricks left to the voxel-range fold (beyond the mid-size tile)
    CTR_EDGE_OVER = 15,  // a vertex has more half-edges than its slots (boundary edges re-run on the CSR lists)
    // sum and largest shard of each sharded buffer (filled by k_shard_reduce)
    CTR_SUM = 16,        // CTR_SUM + SB_*  : total items requested
    CTR_MAX = 32,        // CTR_MAX + SB_*  : largest per-shard request
    CTR_HUGE_NEXT = 42,  // next entry of the huge-brick list to fold (blocks of k_fold_large take bricks one by one)
    CTR_COUNT = 48
};

// sharded buffers
enum
{
    SB_EVENTS = 0,
    SB_FHITS = 1,
    SB_BHITS = 2,
    SB_RECS = 3,
    SB_CHUNKS = 4,
    SB_BOX = 5, // voxels queued for the box-constrained search
    SB_FP64 = 6, // counters only: voxels solved on the FP64 path (precision variant)
    SB_ROWQ = 7, // row chunks of large triangles (scanline)
    SB_FACEQ = 8, // row chunks of large triangles (face test)
    SB_FACE_EST = 9, // counters only: face-test hits expected from the faces' sizes (prep; sizes the face hits)
    SB_N = 10
};

// Append / allocation counters are split over NSHARD counters 128 bytes
// apart: a single global counter serialises every wave of the GPU on one
// address.  A buffer is NSHARD equal regions, region i filled through counter i.
constexpr int NSHARD = 256;
constexpr int SHARD_STRIDE = 16;

template <typename T>
struct ShardBuf
{
    T *items;
    uint64_t per;            // capacity of one region
    unsigned long long *cnt; // [NSHARD * SHARD_STRIDE] region fill counters
};

// error flags (CTR_ERRORS)
enum
{
    ERR_FACE_INDEX = 1,
    ERR_NONFINITE = 2,
    ERR_RANGE = 4,
    ERR_DDA = 8
};

// One scanline crossing, recorded once by the mark pass:
//   key  = (face*3 + axis) << 28 | ly << 15 | lx << 2  (the phase-0 key without dx, dy)
//   z    = FP32 crossing coordinate along the axis, zi = its voxel index
struct Event
{
    uint64_t key;
    float z;
    int zi;
};

static_assert(SB_N <= 16, "counter layout: CTR_SUM and CTR_MAX hold 16 entries each");

// Rows [r0, r1) of one (face, axis) scanline item, queued when a triangle has
// more rows than one block iteration should take (balances large triangles
// over the grid)
struct RowChunk
{
    uint32_t face;
    int ax2, r0, r1;
};

// Bounding-box rows [r0, r1) (linear (y, z) index) of one face for the face test
struct FaceChunk
{
    uint32_t face;
    uint32_t r0, r1;
    uint32_t ms; // the face's mesh
};

// one QEF contribution: order key, crossing z (phase 0), voxel index in its brick
struct Rec
{
    uint64_t key;
    float z;
    uint32_t lvox;
};

// contiguous run of one brick's records written by one block; `rank` is the
// chunk's position among its brick's chunks
struct Chunk
{
    uint32_t start;
    uint32_t count;
    uint32_t brick;
    uint32_t rank;
};

// face-test hit: face index and voxel (grid-local coordinates)
struct FHit
{
    uint32_t face;
    uint16_t lx, ly, lz;
    uint16_t ms; // the face's mesh (FHIT_NO_MESH in batches of more meshes: looked up from the face)
};
constexpr uint16_t FHIT_NO_MESH = 0xffff;

// boundary DDA hit: boundary-edge rank and voxel
struct BHit
{
    uint32_t edge;
    int x, y, z;
};

// One mesh of a batch (a single conversion is a batch of one).  Its grid's
// bits / woff point at the mesh's first word of the shared occupancy bitset
// and word offsets, so word indices inside a mesh are local and voxel ids
// (word offsets) are global: the meshes' voxels follow each other in mesh order.
struct MeshSeg
{
    GridDev g;
    int nb[3];      // brick grid
    int bs;         // log2 of the brick edge: 3 (8x8x8 voxels) or 2 (4x4x4, meshes of dense records)
    uint64_t brick0; // first brick of the mesh in the shared brick numbering
    int64_t face0;   // first face
};

// the batch on the device: segments and prefix offsets (M + 1 entries each)
struct BatchDev
{
    const MeshSeg *seg;
    const int64_t *face_off;   // faces
    const int64_t *vert_off;   // vertices
    const uint64_t *word_off;  // bitset words
    const uint64_t *brick_off; // bricks
    const uint64_t *vox_off;   // voxels (written on the device after the word scan)
    const uint64_t *mtile_off; // scanline tiles (ceil(3 F_m / tile)) for the tile size of the launch
    const uint64_t *ftile_off; // face-test tiles (ceil(F_m / tile)) for the tile size of the launch
    // guides of the face / vertex / word / brick / voxel prefixes:
    // gd[k][j] = owner of index j << gsh[k], for j up to ((total - 1) >> gsh[k]) + 1
    const int32_t *gd[5];
    int gsh[5];
    int M;
};

enum SegGuide
{
    GD_FACE = 0,
    GD_VERT = 1,
    GD_WORD = 2,
    GD_BRICK = 3,
    GD_VOX = 4,
    N_GUIDES = 5
};

// guide entries per prefix array for M meshes (the shift makes the guide
// cover the array's total in at most this many steps)
__host__ __device__ constexpr int64_t guide_len(int M)
{
    return 8 * (int64_t)M + 2;
}

// the mesh owning index x of a prefix array (largest m with off[m] <= x;
// empty meshes are skipped over), searched inside [lo, hi]
template <typename T>
__device__ __forceinline__ int seg_search(const T *off, int lo, int hi, T x)
{
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

template <typename T>
__device__ __forceinline__ int seg_find(const T *off, int M, T x)
{
    return seg_search(off, 0, M - 1, x);
}

// the same through the prefix's guide: the owner of x lies between the
// owners of the guide points below and above x, usually one mesh
template <typename T>
__device__ __forceinline__ int seg_find(const BatchDev &b, int k, const T *off, T x)
{
    if (b.M == 1)
        return 0;
    const uint64_t j = (uint64_t)x >> b.gsh[k];
    return seg_search(off, b.gd[k][j], b.gd[k][j + 1], x);
}

__device__ __forceinline__ int mesh_of_face(const BatchDev &b, int64_t f)
{
    return seg_find(b, GD_FACE, b.face_off, f);
}
__device__ __forceinline__ int mesh_of_vert(const BatchDev &b, int64_t v)
{
    return seg_find(b, GD_VERT, b.vert_off, v);
}
__device__ __forceinline__ int mesh_of_word(const BatchDev &b, uint64_t w)
{
    return seg_find(b, GD_WORD, b.word_off, w);
}
__device__ __forceinline__ int mesh_of_brick(const BatchDev &b, uint64_t k)
{
    return seg_find(b, GD_BRICK, b.brick_off, k);
}
__device__ __forceinline__ int mesh_of_voxel(const BatchDev &b, uint64_t v)
{
    return seg_find(b, GD_VOX, b.vox_off, v);
}

struct BucketArgs
{
    ShardBuf<Rec> recs;
    ShardBuf<Chunk> chunks;
    unsigned long long *brick_cnt; // [bricks] chunks (low 32 bits) and records (high 32 bits) per brick
    BatchDev b;              // meshes (grid, brick grid and first brick of each)
    const int2 *bedge_v;     // boundary edges (vertex pairs), to find a boundary hit's mesh
    unsigned long long *ctr; // device counters
};

// boundary segment: quadric (FDG.cpp:433-453) and DDA set-up
struct BEdge
{
    double dir[3];
    double len;
    float q[10];
    float v0[3];
    int valid;
};

// voxel queued for the box-constrained search: regularised quadric,
// unconstrained minimiser, voxel coordinates, output row
struct BoxItem
{
    float q[10];
    float v[3];
    int c[3];
    uint64_t vid;
};

struct FoldArgs
{
    const uint32_t *choff;              // [bricks+1] first chunk of each brick in `bchunks`
    const uint2 *gchunks;               // (first record, count) of each chunk, grouped by brick
    const Rec *recs;
    const uint32_t *tiny;               // bricks with the fewest voxel entries (first fold tier)
    const unsigned long long *n_tiny;   // their count (device)
    uint32_t *active;                   // bricks of the second fold tier
    unsigned long long *n_active;       // their count (device, zeroed by the host)
    uint32_t *medium;                   // bricks too large for the second fold tier
    unsigned long long *n_medium;       // their count (device, zeroed by the host)
    uint32_t *large;                    // bricks too large for the main fold tile
    unsigned long long *n_large;        // their count (device, zeroed by the host)
    uint32_t *huge;                     // bricks too large for the mid-size tile
    unsigned long long *n_huge;         // their count (device, zeroed by the host)
    BatchDev b;                         // meshes
    int nb[3];  // brick grid of the brick being folded (set per brick from its mesh)
    int bs;     // log2 of its brick edge
    GridDev g;  // grid of the brick's mesh (set per brick)
    const float4 *planes;
    const BEdge *bedges;
    float *dual;
    uint8_t *flags;
    float *qef; // precision variant: [N][20] scaled quadric (h then l) of the voxels its folds leave unsolved
    uint32_t *solve_bits; // precision variant: [ceil(N / 32)] those voxels (cleared before the folds)
    float bw, rw;
    ShardBuf<BoxItem> box; // box-search queue
    int prec;              // 0 reference float path, 1 compensated voxel-frame accumulation
    float cond_max;        // precision variant: FP64 solve above this condition estimate
    unsigned long long *n64; // precision variant: sharded counters of FP64 solves
    unsigned long long *ctr;
};

// est_ctr: the NSHARD counters of SB_FACE_EST (cleared by the caller)
hipError_t launch_prep(const MeshDev &m, const BatchDev &b, float4 *planes, float *tri, unsigned long long *ctr,
                       unsigned long long *est_ctr, hipStream_t s);
// tile: index into MARK_TILES; b.mtile_off must hold that tile size's offsets
hipError_t launch_scan_events(const MeshDev &m, const BatchDev &b, int tile, uint64_t n_tiles, unsigned long long *ctr,
                              const ShardBuf<Event> &eb, const ShardBuf<RowChunk> &rq, bool merge, hipStream_t s);
// tile: index into FACE_TILES; b.ftile_off must hold that tile size's offsets
hipError_t launch_face(const MeshDev &m, const BatchDev &b, int tile, uint64_t n_tiles, const ShardBuf<FHit> &hb,
                       unsigned long long *ctr, const ShardBuf<FaceChunk> &fq, hipStream_t s);
hipError_t launch_gen_events(const ShardBuf<Event> &ev, const BucketArgs &ba, hipStream_t s);
hipError_t launch_gen_fhits(const ShardBuf<FHit> &h, const BucketArgs &ba, hipStream_t s);
hipError_t launch_shard_reduce(unsigned long long *shard_ctr, unsigned long long *ctr, hipStream_t s);
// One mesh of a batch to pack (the layout of the C ABI's ovx_mesh_src)
struct MeshSrc
{
    const float *vertices; // [n_vertices, 3]
    const void *faces;     // [n_faces, 3], int32 or int64
    const float *origin;   // 3 floats subtracted from the vertices
    int64_t n_vertices, n_faces;
    int64_t vertex_offset, face_offset; // first packed rows
    int32_t faces_int64;
    int32_t reserved;
};
// packed vertices (nv rows) and faces (nf rows) of the M meshes of table t (device)
hipError_t launch_batch_prepare(const MeshSrc *t, int M, uint64_t nv, uint64_t nf, float *vout, int32_t *fout,
                                hipStream_t s);
// vout = v - origin per vertex (nv vertices); f32 = f64 (nf3 indices, -1 where
// outside [0, 2^31)); either part may be empty
hipError_t launch_prepare_mesh(const float *v, const float *origin, uint64_t nv, float *vout, const int64_t *f64,
                               uint64_t nf3, int32_t *f32, hipStream_t s);
// fills n 64-bit / 32-bit words with v (a kernel: capturable in a graph)
hipError_t launch_fill_u64(void *p, uint64_t n, uint64_t v, hipStream_t s);
hipError_t launch_fill_u32(void *p, uint64_t n, uint32_t v, hipStream_t s);
// Sets ctr[CTR_OVER] when a record-stage buffer overflowed: the largest
// region request of the face hits, boundary hits, records or chunks above its
// per-region capacity, or the chunks above the grouped-chunk capacity.  Runs
// after k_shard_reduce.
struct OverCaps
{
    uint64_t fhits, bhits, recs, chunks; // per-region capacities
    uint64_t grouped;                    // chunks the grouped list holds
};
hipError_t launch_over_check(unsigned long long *ctr, const OverCaps &caps, hipStream_t s);
hipError_t launch_place_chunks(const ShardBuf<Chunk> &chunks, const uint32_t *choff, uint2 *gchunks,
                               const unsigned long long *over, hipStream_t s);
// Bricks with records, each listed for the fold tier its voxel-entry count
// selects: lists[t] / counts[t] for t = 0 (at most FOLD_TINY_CAP entries),
// 1 (FOLD_SMALL_CAP), 2 (FOLD_SMALL), 3 (FOLD_MID), 4 (more).
constexpr int N_FOLD_TIERS = 5;
struct TierLists
{
    uint32_t *list[N_FOLD_TIERS];
    unsigned long long *count[N_FOLD_TIERS];
};
hipError_t launch_active_bricks(const unsigned long long *cnt, uint64_t n_bricks, TierLists t,
                                const unsigned long long *over, hipStream_t s);
hipError_t launch_fold_bricks(const FoldArgs &a, uint64_t max_bricks, hipStream_t s);
// bricks of the `large` list up to the mid-size tile; the rest go to `huge`
hipError_t launch_fold_mid(const FoldArgs &a, uint64_t max_bricks, hipStream_t s);
// bricks of the `huge` list: whole, by voxel ranges, or a giant voxel in key-ordered batches
hipError_t launch_fold_large(const FoldArgs &a, uint64_t max_bricks, hipStream_t s);
// dual vertices of the precision variant's voxels that its folds left unsolved (flagged in a.solve_bits, their
// quadrics in a.qef): the FP64 fallback and the box search in the voxel frame
hipError_t launch_solve_voxels(const FoldArgs &a, const int32_t *coords, uint64_t n, hipStream_t s);
hipError_t launch_solve_box(const FoldArgs &a, hipStream_t s);
hipError_t launch_edge_count(const MeshDev &m, uint32_t *vcnt, uint32_t *rank, const uint32_t *voff, uint32_t *nbr,
                             int pass, hipStream_t s);
// Boundary edges by bins of EBIN_VB consecutive vertices (a half-edge goes
// to the bin of its smaller vertex).  bin_cnt, bin_off, bin_bcnt and bin_boff
// hold n_bins + 1 entries each, bin_cnt and bin_bcnt zeroed first; he holds
// every half-edge (3 per face).  Order on one stream:
//   launch_ebin(place = false)          half-edges per bin into bin_cnt
//   scan bin_cnt -> bin_off              each bin's range of he
//   launch_ebin(place = true)           the half-edges into their ranges (bin_cnt drops to 0)
//   launch_ebin_fold                     each bin's singletons at the start of its range, counts in bin_bcnt;
//                                        a vertex with more than EDGE_SLOTS half-edges sets `over` (the
//                                        caller then re-runs the CSR lists below)
//   scan bin_bcnt -> bin_boff            the boundary-edge total
//   launch_ebin_gather                   the boundary edges in (min, max) order into bedge_v
constexpr int EDGE_SLOTS = 16;
constexpr int EBIN_VB = 512;
hipError_t launch_ebin(const MeshDev &m, bool place, uint32_t *bin_cnt, const uint32_t *bin_off, uint2 *he,
                       hipStream_t s);
hipError_t launch_ebin_fold(uint64_t n_bins, const uint32_t *bin_off, uint2 *he, uint32_t *bin_bcnt,
                            unsigned long long *over, hipStream_t s);
hipError_t launch_ebin_gather(uint64_t n_bins, const uint32_t *bin_off, const uint32_t *bin_boff, const uint2 *he,
                              int2 *bedge_v, hipStream_t s);
hipError_t launch_edge_singletons(int64_t V, const uint32_t *voff, uint32_t *nbr, uint32_t *bcnt, const uint32_t *boff,
                                  int2 *bedge_v, int pass, hipStream_t s);
hipError_t launch_bedge_prep(const float *verts, int64_t E, const int2 *bedge_v, BEdge *be, hipStream_t s);
// boundary DDA and its phase-2 records (hits counted on hcount's counters only)
hipError_t launch_bdry(int64_t E, const BEdge *be, const int2 *bedge_v, const BatchDev &b, const BucketArgs &ba,
                       const ShardBuf<BHit> &hcount, unsigned long long *ctr, hipStream_t s);
uint64_t scan_tile_count(uint64_t n);
hipError_t launch_scan_popc(const uint64_t *bits, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                            unsigned long long *total, hipStream_t s);
// exclusive scan of the low 32 bits of 64-bit values (the chunk counts of `brick_cnt`)
hipError_t launch_scan_lo32(const unsigned long long *in, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                            unsigned long long *total, hipStream_t s);
hipError_t launch_scan_u32(const uint32_t *in, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                           unsigned long long *total, hipStream_t s);
hipError_t launch_seg_voxels(const BatchDev &b, const uint32_t *woff, uint64_t n_words,
                             const unsigned long long *n_voxels, uint64_t *vox_off, hipStream_t s);
hipError_t launch_emit_coords(const BatchDev &b, uint64_t n_words, int32_t *coords, hipStream_t s);

} // namespace ovx
