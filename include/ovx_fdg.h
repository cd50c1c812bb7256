/*
 * ovx_fdg.h -- C ABI of the MI355X (gfx950) flexible-dual-grid library
 * (libovx_fdg.so).  Plain pointers, sizes and a hipStream_t; no torch types.
 *
 * Conventions
 *   - Every function returns an ovx_status; 0 is success.  On failure
 *     ovx_last_error() returns a message for the calling thread.
 *   - Device pointers (d_*) must be on the context's device.  Input buffers
 *     are borrowed read-only; output buffers are owned by the caller.
 *   - A context owns a grow-only device workspace.  Calls on one context are
 *     serialised by an internal lock; use one context per concurrent caller
 *     to run conversions in parallel.
 *   - A conversion is two calls: *_begin computes the occupied voxel set and
 *     returns its size N (one blocking read-back); the caller allocates the
 *     [N,3] outputs and calls *_finish, which enqueues the rest on `stream`
 *     and blocks once, at its end, on a read-back of the error and overflow
 *     flags (when a record buffer overflowed it sizes the buffers exactly and
 *     runs the record stages again, blocking once more).  The outputs are
 *     complete when *_finish returns OVX_OK; the work of ovx_occ_finish,
 *     ovx_dual_mesh_finish and the svo calls is only enqueued.
 *   - Calls on one context may use different streams: a call's work waits
 *     for the previous call's work on the context (an event recorded at the
 *     end of every call), since they share the workspace.  Work on other
 *     streams of the caller is never waited for.
 *
 * Drop-in boundary: these entry points are what the reference's extension
 * module binds for this path (src/ext.cpp:18-21), split into begin/finish so
 * that the caller's allocator owns the outputs.
 */
#ifndef OVX_FDG_H
#define OVX_FDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ovx_status;
#define OVX_OK 0
#define OVX_ERR_INVALID_ARGUMENT 1 /* bad shape / size / parameter          */
#define OVX_ERR_INPUT 2            /* face index out of range, non-finite   */
#define OVX_ERR_CAPACITY 3         /* a count exceeds a supported limit     */
#define OVX_ERR_HIP 4              /* a HIP runtime call failed             */
#define OVX_ERR_STATE 5            /* finish without a matching begin, ...  */

typedef struct ovx_fdg_ctx ovx_fdg_ctx;

/* Grid description shared by every entry point.  Mirrors the reference's
 * `voxel_size` (3 floats) and `grid_range` ([min_x, min_y, min_z, max_x,
 * max_y, max_z]) arguments (src/convert/api.h:31-39). */
typedef struct ovx_grid
{
    float voxel_size[3];
    int64_t grid_range[6];
} ovx_grid;

/* Library version string. */
const char *ovx_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char *ovx_last_error(void);

/* Create / destroy a context bound to HIP device `device`. */
ovx_status ovx_fdg_create(int device, ovx_fdg_ctx **ctx);
ovx_status ovx_fdg_destroy(ovx_fdg_ctx *ctx);

/* Bytes of device workspace currently held by the context. */
int64_t ovx_fdg_workspace_bytes(const ovx_fdg_ctx *ctx);

/* Release the context's workspace (it is re-grown on the next call). */
ovx_status ovx_fdg_trim(ovx_fdg_ctx *ctx);

/*
 * A conversion's inputs from the caller's arrays, as the public wrapper
 * prepares them before the extension call (o_voxel/convert/flexible_dual_grid.py:89-90
 * faces cast to int32, :181 vertices - aabb[0]), in one kernel on `stream`
 * (the current device's), so that a conversion queues no framework kernel:
 *   d_out_vertices[i][k] = d_vertices[i][k] - d_origin[k]   (FP32, n_vertices rows)
 *   d_out_faces[j][k]    = d_faces[j][k] as int32           (n_faces rows; an index
 *                          outside [0, 2^31) becomes -1, which ovx_fdg_begin
 *                          rejects as out of range)
 * d_faces may be NULL (faces already int32: only the vertices are written).
 * Needs no context; returns without waiting.
 */
ovx_status ovx_fdg_prepare_mesh(const float *d_vertices, const float *d_origin, int64_t n_vertices,
                                const int64_t *d_faces, int64_t n_faces, float *d_out_vertices, int32_t *d_out_faces,
                                void *stream);

/*
 * Packed inputs of a batched conversion (ovx_fdg_batch_begin) from each
 * mesh's own arrays, in one kernel on `stream`: what the batch wrapper
 * needs (no counterpart in the reference, which converts one mesh per call,
 * FDG.py:89-90,181 per mesh).  d_meshes is a DEVICE array of n_meshes
 * records; mesh m's vertices minus its origin go to rows
 * [vertex_offset, vertex_offset + n_vertices) of d_out_vertices
 * [total_vertices, 3], its faces plus vertex_offset, as int32, to rows
 * [face_offset, face_offset + n_faces) of d_out_faces [total_faces, 3].
 * Offsets ascend mesh after mesh.  A face index outside [0, n_vertices of
 * its mesh) becomes -1, which ovx_fdg_batch_begin rejects as out of range.
 * Writes stay inside the outputs whatever the table holds.  total_vertices
 * must be < 2^31 (OVX_ERR_CAPACITY).  Needs no context; returns without
 * waiting.
 */
typedef struct ovx_mesh_src
{
    const float *vertices; /* device, [n_vertices, 3] float32 */
    const void *faces;     /* device, [n_faces, 3] int32, or int64 when faces_int64 */
    const float *origin;   /* device, 3 floats subtracted from the mesh's vertices */
    int64_t n_vertices, n_faces;
    int64_t vertex_offset, face_offset; /* first packed rows of the mesh */
    int32_t faces_int64;
    int32_t reserved;
} ovx_mesh_src;
ovx_status ovx_fdg_batch_prepare(const ovx_mesh_src *d_meshes, int n_meshes, int64_t total_vertices,
                                 int64_t total_faces, float *d_out_vertices, int32_t *d_out_faces, void *stream);

/*
 * mesh_to_flexible_dual_grid, device path.
 * Replaces o_voxel::fdg::mesh_to_flexible_dual_grid_cuda
 * (src/convert/mesh_to_flexible_dual_grid_gpu/mesh_to_flexible_dual_grid.cu:797-895;
 *  declared src/convert/api.h:164-172, bound src/ext.cpp:19) and, run on host
 * copies, the semantics of mesh_to_flexible_dual_grid_cpu
 * (src/convert/flexible_dual_grid.cpp:660-992, bound src/ext.cpp:18).
 *
 *   d_vertices  float32 [n_vertices, 3], grid-local (already shifted by aabb[0])
 *   d_faces     int32   [n_faces, 3]
 *   weights     face / boundary / regularization weights (reference semantics:
 *               face_weight > 0 enables the unscaled face term, boundary
 *               terms are scaled by boundary_weight, regularization by
 *               regularization_weight * count)
 * begin returns N (number of occupied voxels) in *n_voxels.
 * finish writes
 *   d_coords       int32   [N,3]  voxel coordinates, ascending x + Gx*(y + Gy*z)
 *   d_dual         float32 [N,3]  dual vertices in grid-local units
 *   d_intersected  uint8   [N,3]  0/1 intersected-edge flags (torch.bool layout)
 */
ovx_status ovx_fdg_begin(ovx_fdg_ctx *ctx, const float *d_vertices, int64_t n_vertices, const int32_t *d_faces,
                         int64_t n_faces, const ovx_grid *grid, float face_weight, float boundary_weight,
                         float regularization_weight, void *stream, int64_t *n_voxels);
/*
 * Batched conversion: M meshes of any sizes and grids in one pass, every
 * stage one launch over all of them.  No counterpart in the reference (its
 * API converts one mesh per call, FDG.py:38-205); each mesh's results are the
 * ones ovx_fdg_begin / ovx_fdg_finish give for it alone.
 *   d_vertices  float32 [n_vertices, 3]: the meshes' grid-local vertices, mesh after mesh
 *   d_faces     int32 [n_faces, 3]: faces, mesh after mesh, indices into d_vertices
 *               (batch-global: mesh m's indices lie in [vertex_offsets[m], vertex_offsets[m+1]))
 *   face_offsets, vertex_offsets  host int64 [n_meshes + 1] prefix offsets (first 0, last n_faces / n_vertices)
 *   grids       host [n_meshes]: each mesh's grid
 *   n_voxels    host int64 [n_meshes]: each mesh's voxel count
 * Then ovx_fdg_finish writes all meshes' outputs, mesh after mesh (mesh m's
 * rows start at the sum of the voxel counts before it).  Blocks once, like
 * ovx_fdg_begin.
 */
ovx_status ovx_fdg_batch_begin(ovx_fdg_ctx *ctx, const float *d_vertices, int64_t n_vertices, const int32_t *d_faces,
                               int64_t n_faces, int n_meshes, const int64_t *face_offsets,
                               const int64_t *vertex_offsets, const ovx_grid *grids, float face_weight,
                               float boundary_weight, float regularization_weight, void *stream, int64_t *n_voxels);

ovx_status ovx_fdg_finish(ovx_fdg_ctx *ctx, int32_t *d_coords, float *d_dual, uint8_t *d_intersected, void *stream);

/*
 * Allocation-free conversion: ovx_fdg_begin + ovx_fdg_finish in one call
 * that allocates nothing.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under HIP graph capture
 * of their own, tight serving loops, harnesses that forbid hipMalloc).
 * Every workspace buffer is carved from a static per-device scratch of
 * OVX_FDG_NOALLOC_SCRATCH_BYTES (held by each device the library runs on
 * until the process ends; its address is looked up by the context's first
 * allocation-free call, which also pays for placing it on the device);
 * the context's own workspace is left as it was.  The caller passes the
 * outputs with room for max_voxels rows (layout as ovx_fdg_finish) and gets
 * the voxel count in *n_voxels.
 * Errors: OVX_ERR_CAPACITY when the conversion needs more than the scratch
 * (checked before the stage that would need it is queued) or has more than
 * max_voxels voxels (checked before anything is written to the outputs).
 * All calls on one device share its scratch and leave their contents in
 * it: calls on one context are ordered, but calls that overlap on one GPU
 * from different contexts race.  Blocks twice, like ovx_fdg_begin and
 * ovx_fdg_finish; precision follows ovx_fdg_set_precision; graph mode does
 * not apply.
 */
#define OVX_FDG_NOALLOC_SCRATCH_BYTES (64ll << 20)
ovx_status ovx_fdg_convert_noalloc(ovx_fdg_ctx *ctx, const float *d_vertices, int64_t n_vertices,
                                   const int32_t *d_faces, int64_t n_faces, const ovx_grid *grid, float face_weight,
                                   float boundary_weight, float regularization_weight, int32_t *d_coords,
                                   float *d_dual, uint8_t *d_intersected, int64_t max_voxels, void *stream,
                                   int64_t *n_voxels);

/*
 * intersect_occ, device path: occupancy of a triangle soup.
 * Replaces o_voxel::fdg::intersect_occ_cuda
 * (src/convert/mesh_to_flexible_dual_grid_gpu/intersect_qef.cu:896-920;
 *  declared src/convert/api.h:63-66, bound src/ext.cpp:21).
 *   d_triangles float32 [n_triangles, 3, 3], grid-local.
 * finish writes d_coords int32 [N,3] in ascending x + Gx*(y + Gy*z) order.
 */
ovx_status ovx_occ_begin(ovx_fdg_ctx *ctx, const float *d_triangles, int64_t n_triangles, const ovx_grid *grid,
                         void *stream, int64_t *n_voxels);
ovx_status ovx_occ_finish(ovx_fdg_ctx *ctx, int32_t *d_coords, void *stream);

/*
 * Arithmetic of the QEF accumulation and solve for the next conversions on
 * this context.  No counterpart in the reference: it is the precision
 * variant beside the reference-faithful path.
 *   OVX_PREC_REFERENCE (default): the reference CPU float path
 *     (FDG.cpp:158-983) operation for operation; duals bit-identical to it.
 *   OVX_PREC_COMPENSATED: each voxel's quadric is accumulated in the voxel's
 *     own frame u = (p - min corner) / voxel_size with compensated FP32 sums
 *     of error-free products; the 3x3 solve runs in FP32 and falls back to
 *     FP64 (pivoting, rank decision, back substitution) when the system is
 *     rank-deficient or |R00 / R22| > cond_max.  Occupancy and flags are the
 *     same as the reference path; duals approach the FP64 solution.
 * cond_max >= 1.  Must not be called between begin and finish.
 */
#define OVX_PREC_REFERENCE 0
#define OVX_PREC_COMPENSATED 1
ovx_status ovx_fdg_set_precision(ovx_fdg_ctx *ctx, int mode, float cond_max);

/*
 * Launch-class thresholds of the scanline and face-test kernels for the next
 * conversions on this context.  No counterpart in the reference (its launch
 * shapes are fixed); the autotuner o_voxel.convert.autotune_tile_classes sets
 * them from timed trial conversions.  A mesh's scanline class comes from its
 * faces per grid cross-section fpc (F / the largest of Gx*Gy, Gy*Gz, Gx*Gz):
 *   fpc <  mark_fpc3[0]: 64-item tiles with a 2048-word LDS merge table
 *   fpc <= mark_fpc3[1]: 128-item tiles with a 1024-word merge table
 *   fpc <  mark_fpc3[2]: 128-item tiles, no table
 *   otherwise:           256-item tiles, no table
 * and its face-test class from its voxels per face vpf:
 *   vpf >= face_vpf2[1]: 32 faces per block; vpf >= face_vpf2[0]: 64; else 128.
 * Defaults {0.5, 2, 8} and {4, 16} (the autotuner's choice on a mixed batch
 * of 256 meshes, 500..500k faces, grids 64..512).  Thresholds are finite, >= 0 and
 * non-decreasing; a null array leaves its set unchanged.  Results do not
 * depend on them, only the launch shapes do.  Must not be called between
 * begin and finish.
 */
ovx_status ovx_fdg_set_tile_classes(ovx_fdg_ctx *ctx, const double *mark_fpc3, const double *face_vpf2);
ovx_status ovx_fdg_get_tile_classes(ovx_fdg_ctx *ctx, double *mark_fpc3, double *face_vpf2);

/*
 * Brick size of the record grouping and fold for the next conversions on
 * this context.  No counterpart in the reference (its brick hash uses 8x8x8
 * bricks, src/convert/mesh_to_flexible_dual_grid_gpu/types.cuh:47-55).  A
 * mesh with at least `faces_per_voxel` faces per occupied voxel is grouped in
 * 4x4x4 bricks (its records are dense: an 8x8x8 brick would exceed the fold's
 * tiles), any other in 8x8x8 bricks.  Default 4; the autotuner
 * o_voxel.convert.autotune_tile_classes sets it from timed trial
 * conversions.  Results do not depend on it.  Must not be called between
 * begin and finish.
 */
ovx_status ovx_fdg_set_brick_class(ovx_fdg_ctx *ctx, double faces_per_voxel);
ovx_status ovx_fdg_get_brick_class(ovx_fdg_ctx *ctx, double *faces_per_voxel);

/*
 * Graph mode for the next conversions on this context.  No counterpart in the
 * reference (it launches every kernel from the host).  With enable != 0,
 * ovx_fdg_finish enqueues its device work (every launch between the voxel
 * set and the final read-back) as a HIP graph: captured the first time a plan
 * occurs, replayed when a conversion has the same plan (the same mesh sizes,
 * grids, weights, buffers and output pointers, whatever the vertex values).
 * Up to 4 graphs are kept per context.  Outputs are identical to the ones
 * without graphs.  Disabling releases the graphs.  Timing
 * (ovx_fdg_set_timing) bypasses graphs.  Must not be called between begin and
 * finish.  ovx_fdg_graph_stats: graph launches that replayed a kept graph,
 * and graphs captured, since the context was created.
 */
ovx_status ovx_fdg_set_graph(ovx_fdg_ctx *ctx, int enable);
ovx_status ovx_fdg_graph_stats(ovx_fdg_ctx *ctx, int64_t *replays, int64_t *captures);

/*
 * Per-stage statistics of the last conversion on this context; the first
 * min(n, OVX_FDG_N_STATS) are written:
 *   out[0] N voxels          out[1] scanline events     out[2] QEF records
 *   out[3] face records      out[4] boundary edges      out[5] boundary records
 *   out[6] workspace bytes   out[7] voxels solved on the slow (giant) path
 *   out[8] bricks with records                         out[9] record chunks
 *   out[10] bricks folded by the large-brick kernel    out[11] their records
 *   out[12] voxels queued for the box-constrained search (the rest of the voxels
 *           whose minimiser left the box are searched in place)
 *   out[13] voxels solved on the FP64 path (precision variant; 0 otherwise)
 *   out[14] times finish generated the records (1, or more when a record
 *           buffer overflowed and was sized again)
 * Reading them synchronises the context's last stream.
 */
#define OVX_FDG_N_STATS 15
ovx_status ovx_fdg_stats(ovx_fdg_ctx *ctx, int64_t *out, int n);

/*
 * Capacities (in items) for the next conversion on this context: the
 * scanline-crossing buffer and the QEF-record buffer (0 keeps the current
 * sizing).  A conversion that needs more re-runs the affected pass with exact
 * sizes, so any value is correct; it only trades memory against a second pass.
 */
ovx_status ovx_fdg_set_capacity(ovx_fdg_ctx *ctx, int64_t events, int64_t hits);

/*
 * Device timing of the last conversion, recorded with HIP events around each
 * kernel launch on the caller's stream once enabled with
 * ovx_fdg_set_timing(ctx, 1).
 *   ovx_fdg_kernel_times: per timing slot (ovx_fdg_kernel_count() slots, named
 *     by ovx_fdg_kernel_name), -1 for kernels that did not run.
 *   ovx_fdg_stage_times: the reference's four timed stages (FDG.cpp:711-986):
 *     [intersection QEF, face QEF, boundary QEF, dual vertices], each the sum
 *     of the kernels that build or consume that stage's records.
 */
ovx_status ovx_fdg_set_timing(ovx_fdg_ctx *ctx, int enable);
int ovx_fdg_kernel_count(void);
const char *ovx_fdg_kernel_name(int slot);
ovx_status ovx_fdg_kernel_times(ovx_fdg_ctx *ctx, float *ms, int n);
ovx_status ovx_fdg_stage_times(ovx_fdg_ctx *ctx, float *ms4);

/*
 * flexible_dual_grid_to_mesh, device path: the quads around intersected grid
 * edges whose four voxels all exist, split into triangles.
 * Replaces the body of o_voxel.convert.flexible_dual_grid_to_mesh
 * (o_voxel/convert/flexible_dual_grid.py:422-454: its voxel hashmap built by
 * _C.hashmap_insert_3d_idx_as_val_cuda, the neighbour lookups by
 * _C.hashmap_lookup_3d_cuda and the torch glue that forms the quads and
 * picks each split).
 *   d_coords      int32   [n_voxels, 3] voxel coordinates
 *   d_dual        float32 [n_voxels, 3] dual vertices in voxel units
 *   d_flags       uint8   [n_voxels, 3] intersected-edge flags (nonzero = set)
 *   grid_size     host [3] (W, H, D): lookups outside [0,W)x[0,H)x[0,D) fail
 *   voxel_size, aabb_min   host float [3]
 *   d_vertices    float32 [n_voxels, 3] out: (coords + dual) * voxel_size + aabb_min
 * begin returns the quad count L in *n_quads (one blocking readback).
 * finish writes, in (voxel, axis) order,
 *   d_faces  int32 [2L, 3] (may be null): two triangles per quad, split by the
 *            larger product of opposite split weights when d_split_weight is
 *            given (float32 [n_voxels] values; split_weight_type says which
 *            dtype's multiply rounds the products: OVX_SPLIT_F32 / F16 /
 *            BF16), otherwise by the reference's normal-alignment rule
 *   d_quads  int32 [L, 4] (may be null): the quads' voxel rows
 * Duplicate coordinates resolve to the largest row.  Uses the context's
 * workspace; not between another begin / finish pair on the same context.
 */
#define OVX_SPLIT_F32 0
#define OVX_SPLIT_F16 1
#define OVX_SPLIT_BF16 2
ovx_status ovx_dual_mesh_begin(ovx_fdg_ctx *ctx, const int32_t *d_coords, const float *d_dual, const uint8_t *d_flags,
                               int64_t n_voxels, const int32_t *grid_size, const float *voxel_size,
                               const float *aabb_min, float *d_vertices, void *stream, int64_t *n_quads);
ovx_status ovx_dual_mesh_finish(ovx_fdg_ctx *ctx, const float *d_split_weight, int split_weight_type,
                                int32_t *d_faces, int32_t *d_quads, void *stream);

/*
 * Generic device hashmap in the reference's table format
 * (src/hash/hash.cuh, src/hash/hash.cu; declared src/hash/api.h:27-110,
 * bound src/ext.cpp:12-16): caller-owned key and value arrays of `capacity`
 * entries (key_bytes / value_bytes 4 or 8; empty keys all ones), Murmur3
 * finalizer modulo capacity, linear probing.  Enqueued on `stream` of the
 * current device; no context.
 *   ovx_hashmap_insert / _lookup: keys [n] of the table's key type.
 *   *_3d: coords int32 [n, 4] (b, x, y, z), 16-byte aligned, key
 *     b*W*H*D + x*H*D + y*D + z; lookups of (x, y, z) outside the grid miss.
 *   ovx_hashmap_insert_3d_idx_as_val: the value of row i is i.
 * Lookups write all ones for a missing key.  Inserting an existing key
 * overwrites its value.  A full table drops the keys it cannot place (the
 * reference probes forever).
 */
ovx_status ovx_hashmap_insert(void *table_keys, int key_bytes, void *table_values, int value_bytes, int64_t capacity,
                              const void *d_keys, const void *d_values, int64_t n, void *stream);
ovx_status ovx_hashmap_lookup(const void *table_keys, int key_bytes, const void *table_values, int value_bytes,
                              int64_t capacity, const void *d_keys, void *d_out, int64_t n, void *stream);
ovx_status ovx_hashmap_insert_3d(void *table_keys, int key_bytes, void *table_values, int value_bytes,
                                 int64_t capacity, const int32_t *d_coords, const void *d_values, int64_t n, int W,
                                 int H, int D, void *stream);
ovx_status ovx_hashmap_lookup_3d(const void *table_keys, int key_bytes, const void *table_values, int value_bytes,
                                 int64_t capacity, const int32_t *d_coords, void *d_out, int64_t n, int W, int H,
                                 int D, void *stream);
ovx_status ovx_hashmap_insert_3d_idx_as_val(void *table_keys, int key_bytes, void *table_values, int value_bytes,
                                            int64_t capacity, const int32_t *d_coords, int64_t n, int W, int H,
                                            int D, void *stream);

/*
 * 30-bit voxel serialization codes (10 bits per axis): Z-order (Morton,
 * x's bit the most significant of each triple) and Hilbert (Skilling's
 * transform, then the same interleave).
 * Replaces z_order_encode/decode_cuda and hilbert_encode/decode_cuda
 * (src/serialize/api.cu, z_order.cu, hilbert.cu; declared
 * src/serialize/api.h; bound src/ext.cpp:28-31) and, on host copies, their
 * _cpu twins (src/ext.cpp:24-27).  Codes are bit-identical to the
 * reference's for every int32 input (coordinates outside [0, 1024) included).
 *   d_cols[3], strides[3]  the three coordinate columns (element strides, so
 *                          the columns of an [N, 3] tensor are read in place)
 *   elem_bytes             4 (int32) or 8 (int64: the low 32 bits are used)
 *   d_codes                int32 [n]
 * Decode writes int32 columns through the same column/stride description.
 */
#define OVX_SEQ_Z_ORDER 0
#define OVX_SEQ_HILBERT 1
ovx_status ovx_seq_encode(const void *const *d_cols, const int64_t *strides, int elem_bytes, int64_t n, int mode,
                          int32_t *d_codes, void *stream);
ovx_status ovx_seq_decode(const int32_t *d_codes, int64_t n, int mode, int32_t *const *d_cols,
                          const int64_t *strides, void *stream);

/*
 * Sparse voxel octrees and attribute filters of the .vxz file format.
 * Replace encode/decode_sparse_voxel_octree_cpu (src/io/svo.cpp:17-138) and
 * encode/decode_sparse_voxel_octree_attr_{parent,neighbor}_cpu
 * (src/io/filter_parent.cpp, src/io/filter_neighbor.cpp; declared
 * src/io/api.h), which the reference binds from its extension and calls per
 * chunk from o_voxel/io/vxz.py.  Octree bytes and filter outputs are
 * identical to the reference's.
 *
 * Encoding runs on the device over all chunks at once:
 *   d_codes      int32 [n] Morton codes of chunk-local coordinates (< 2^(3 depth)),
 *                ascending inside each chunk; d_chunk int32 [n] chunk of each code;
 *                d_chunk_off int64 [n_chunks + 1] first code of each chunk (device)
 *   begin returns the total octree bytes (one blocking readback); finish
 *   writes the chunks' octrees one after another into d_svo (4-byte aligned,
 *   room for n_bytes rounded up to 4) and, if d_chunk_svo_off is given, each
 *   chunk's byte offset (int64 [n_chunks + 1], device).
 *   ovx_svo_parent_encode: after finish, the parent-filter rows of the
 *   chunks (d_attr, d_delta uint8 [n, channels], in code order).
 * ovx_svo_neighbor_filter: the neighbour filter (decode = 0) or its inverse
 *   (decode = 1) over chunk-local int32 [n, 3] coordinates in [0, res),
 *   n_chunks * res^3 <= 2^31.
 * Reading parses pre-order streams on the host (one chunk per call):
 *   ovx_svo_decode_host writes the ascending codes (at most capacity);
 *   ovx_svo_parent_decode_host inverts the parent filter into attr [n, channels].
 * Codes must be distinct within a chunk (repeated codes add no octree byte;
 * the parent filter leaves their rows zero).
 */
ovx_status ovx_svo_encode_begin(ovx_fdg_ctx *ctx, const int32_t *d_codes, const int32_t *d_chunk,
                                const int64_t *d_chunk_off, int64_t n, int n_chunks, int depth, void *stream,
                                int64_t *n_bytes);
ovx_status ovx_svo_encode_finish(ovx_fdg_ctx *ctx, uint8_t *d_svo, int64_t *d_chunk_svo_off, void *stream);
ovx_status ovx_svo_parent_encode(ovx_fdg_ctx *ctx, const uint8_t *d_attr, int channels, uint8_t *d_delta,
                                 void *stream);
ovx_status ovx_svo_neighbor_filter(ovx_fdg_ctx *ctx, const int32_t *d_coords, const int32_t *d_chunk, int64_t n,
                                   int n_chunks, int res, const uint8_t *d_in, int channels, int decode,
                                   uint8_t *d_out, void *stream);
ovx_status ovx_svo_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, int32_t *codes, int64_t capacity,
                               int64_t *n_codes);
ovx_status ovx_svo_parent_decode_host(const uint8_t *svo, int64_t n_bytes, int depth, const uint8_t *delta, int64_t n,
                                      int channels, uint8_t *attr);

#ifdef __cplusplus
}
#endif
#endif /* OVX_FDG_H */
