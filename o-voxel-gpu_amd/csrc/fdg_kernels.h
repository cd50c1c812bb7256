// Kernel launch interface of the flexible-dual-grid library (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fdg_device.h"

namespace ovx
{

enum
{
    MODE_MARK = 0,
    MODE_COUNT = 1,
    MODE_EMIT = 2
};

// device counters (unsigned long long each)
enum
{
    CTR_EVENTS = 0,
    CTR_ERRORS = 1,
    CTR_WORDS_TOTAL = 2, // total of the bitset popcount scan (= N)
    CTR_REC_TOTAL = 3,   // total of the record-count scan (= R)
    CTR_REC_INTERSECT = 4,
    CTR_REC_FACE = 5,
    CTR_REC_BOUNDARY = 6,
    CTR_GIANT = 7,
    CTR_EDGE_TOTAL = 8, // half-edge count scan total
    CTR_BEDGE_TOTAL = 9, // boundary edges
    CTR_COUNT = 16
};

// error flags (CTR_ERRORS)
enum
{
    ERR_FACE_INDEX = 1,
    ERR_NONFINITE = 2,
    ERR_RANGE = 4,
    ERR_DDA = 8,
    ERR_RECORDS = 16
};

struct RecArgs
{
    uint32_t *cnt;           // [N] per-voxel record counters
    const uint32_t *rec_off; // [N+1] record segment offsets
    uint64_t *keys;          // [R] order keys
    float *zs;               // [R] intersection z (phase 0 records)
    unsigned long long *err; // error flags word
};

struct BEdge
{
    double dir[3];
    double len;
    float q[10];
    float v0[3];
    int valid;
};

struct FoldArgs
{
    uint64_t N;
    const uint32_t *rec_off;
    const uint64_t *keys;
    const float *zs;
    uint32_t *gperm;
    const float4 *planes;
    const BEdge *bedges;
    const int32_t *coords;
    float *dual;
    uint8_t *flags;
    float vs[3];
    int gmin[3];
    float bw, rw;
    unsigned long long *counters;
};

hipError_t launch_prep(const MeshDev &m, int64_t V, const GridDev &g, float4 *planes, unsigned long long *ctr,
                       hipStream_t s);
hipError_t launch_scan_events(int mode, const MeshDev &m, const GridDev &g, const RecArgs &r,
                              unsigned long long *ctr, hipStream_t s);
hipError_t launch_face(int mode, const MeshDev &m, const GridDev &g, const RecArgs &r, unsigned long long *ctr,
                       hipStream_t s);
hipError_t launch_edge_count(const MeshDev &m, uint32_t *vcnt, uint32_t *voff, uint32_t *nbr, int pass,
                             hipStream_t s);
hipError_t launch_edge_singletons(int64_t V, const uint32_t *voff, uint32_t *nbr, uint32_t *bcnt, const uint32_t *boff,
                                  int2 *bedge_v, int pass, hipStream_t s);
hipError_t launch_bedge_prep(const float *verts, int64_t E, const int2 *bedge_v, BEdge *be, hipStream_t s);
hipError_t launch_bdry(int mode, int64_t E, const BEdge *be, const GridDev &g, const RecArgs &r,
                       unsigned long long *ctr, hipStream_t s);
uint64_t scan_tile_count(uint64_t n);
hipError_t launch_scan_popc(const uint64_t *bits, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                            unsigned long long *total, hipStream_t s);
hipError_t launch_scan_u32(const uint32_t *in, uint64_t n, uint32_t *out, uint64_t *tile_sums,
                           unsigned long long *total, hipStream_t s);
hipError_t launch_emit_coords(const GridDev &g, uint64_t n_words, int32_t *coords, hipStream_t s);
hipError_t launch_fold_solve(const FoldArgs &a, hipStream_t s);

} // namespace ovx
