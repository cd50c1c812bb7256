// Device-side arithmetic of the flexible-dual-grid path (gfx950, wave64).
//
// Every floating-point expression here reproduces, operation by operation,
// the reference CPU path src/convert/flexible_dual_grid.cpp (FDG.cpp below):
// same operand precision, same association order, no fused multiply-add
// (the library is compiled with -ffp-contract=off), IEEE division and square
// root.  That is what makes occupancy / flags bit-exact and lets the ordered
// QEF fold reproduce the CPU float path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

namespace ovx
{

// Record keys: the 2 top bits hold the accumulation phase, the rest the
// position of the contribution in the CPU's serial order.
//   phase 0 (intersection): face << 30 | axis << 28 | ly << 15 | lx << 2 | dx << 1 | dy
//                           (the order of face*3 + axis, with both as plain fields)
//   phase 1 (face term):    face
//   phase 2 (boundary):     rank of the (min, max) vertex pair
constexpr int KEY_PHASE_SHIFT = 62;
constexpr int KEY_FACE_SHIFT = 30;
constexpr int KEY_AX_SHIFT = 28;
constexpr int KEY_Y_SHIFT = 15;
constexpr int KEY_X_SHIFT = 2;
constexpr uint64_t KEY_IDX_MASK = 0x1fff;  // 13 bits: grid extent <= 8192
constexpr uint64_t KEY_FACE_MASK = 0xffffffffull;

// the axes after ax in cyclic order: (ax + 1) % 3 and (ax + 2) % 3
__host__ __device__ __forceinline__ int axis_next(int ax)
{
    return ax == 2 ? 0 : ax + 1;
}
__host__ __device__ __forceinline__ int axis_prev(int ax)
{
    return ax == 0 ? 2 : ax - 1;
}
constexpr int MAX_GRID_EXTENT = 8192;

struct GridDev
{
    float vs[3];
    float ivs[3];        // 1 / vs where vs is a power of two (x / vs == x * ivs exactly), else 0
    int gmin[3];
    int gmax[3];
    int G[3];
    int wx;              // 64-bit words per x row of the occupancy bitset
    uint64_t *bits;      // [G2 * G1 * wx] occupancy bitset, x fastest
    const uint32_t *woff; // [words] exclusive prefix popcount (voxel id base)
};

struct MeshDev
{
    const float *v;     // [V,3]
    const int32_t *f;   // [F,3], or nullptr for a triangle soup (v is [F,3,3])
    int64_t F;
    int64_t V;          // vertices (>= 1 when f is set and F > 0)
    const float *tri = nullptr; // [F,9] the faces' vertices gathered once (by the prep kernel), or nullptr
    const float4 *pl = nullptr; // [F] the faces' planes (prep kernel), or nullptr
};

// A face with a vertex index outside [0, V) is read as vertex 0 three times
// (a point, which no stage turns into a contribution); the prep kernel
// flags it and the conversion reports the input error, so kernels that run
// before that report never read outside the vertex array.
__device__ __forceinline__ bool face_ok(const MeshDev &m, int32_t a, int32_t b, int32_t c)
{
    return a >= 0 && b >= 0 && c >= 0 && a < m.V && b < m.V && c < m.V;
}

__device__ __forceinline__ void load_tri(const MeshDev &m, int64_t face, float v0[3], float v1[3], float v2[3])
{
    const float *p0, *p1, *p2;
    if (m.tri)
    {
        p0 = m.tri + 9 * face;
        p1 = p0 + 3;
        p2 = p0 + 6;
    }
    else if (m.f)
    {
        const int32_t *fi = m.f + 3 * face;
        int32_t a = fi[0], b = fi[1], c = fi[2];
        if (!face_ok(m, a, b, c))
            a = b = c = 0;
        p0 = m.v + 3 * (int64_t)a;
        p1 = m.v + 3 * (int64_t)b;
        p2 = m.v + 3 * (int64_t)c;
    }
    else
    {
        p0 = m.v + 9 * face;
        p1 = p0 + 3;
        p2 = p0 + 6;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        v0[i] = p0[i];
        v1[i] = p1[i];
        v2[i] = p2[i];
    }
}

__device__ __forceinline__ int iclamp(int v, int lo, int hi) { return v < lo ? lo : (hi < v ? hi : v); }

// FDG.cpp:32-39
// x / vs along axis a: a product by the exact reciprocal where vs is a power
// of two (the same value, without the division sequence)
__device__ __forceinline__ double div_vs(double x, const GridDev &g, int a)
{
    return g.ivs[a] != 0.0f ? x * (double)g.ivs[a] : x / (double)g.vs[a];
}
__device__ __forceinline__ float div_vs(float x, const GridDev &g, int a)
{
    return g.ivs[a] != 0.0f ? x * g.ivs[a] : x / g.vs[a];
}

__device__ __forceinline__ double lerp1(double a, double b, double t, double va, double vb)
{
    if (a == b)
        return va;
    const double alpha = (t - a) / (b - a);
    return (1.0 - alpha) * va + alpha * vb;
}

// Eigen Vector3f dot: a0 + (a1 + a2)
__device__ __forceinline__ float dot3f(float a0, float a1, float a2, float b0, float b1, float b2)
{
    return a0 * b0 + (a1 * b1 + a2 * b2);
}

__device__ __forceinline__ float smaxf(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float sminf(float a, float b) { return (b < a) ? b : a; }

// Unit normal and plane offset of a triangle (FDG.cpp:180-184, 305-310).
__device__ __forceinline__ void tri_plane(const float v0[3], const float v1[3], const float v2[3], float pl[4])
{
    const float e00 = v1[0] - v0[0], e01 = v1[1] - v0[1], e02 = v1[2] - v0[2];
    const float e10 = v2[0] - v1[0], e11 = v2[1] - v1[1], e12 = v2[2] - v1[2];
    const float c0 = e01 * e12 - e02 * e11;
    const float c1 = e02 * e10 - e00 * e12;
    const float c2 = e00 * e11 - e01 * e10;
    const float s = dot3f(c0, c1, c2, c0, c1, c2);
    float n0 = c0, n1 = c1, n2 = c2;
    if (s > 0.0f)
    {
        const float r = sqrtf(s);
        n0 = c0 / r;
        n1 = c1 / r;
        n2 = c2 / r;
    }
    pl[0] = n0;
    pl[1] = n1;
    pl[2] = n2;
    pl[3] = -dot3f(n0, n1, n2, v0[0], v0[1], v0[2]);
}

// Symmetric 4x4 quadric, upper triangle: q00 q01 q02 q03 q11 q12 q13 q22 q23 q33
struct Q10
{
    float q[10];
};

__device__ __forceinline__ void plane_quadric(const float p[4], float q[10])
{
    q[0] = p[0] * p[0];
    q[1] = p[0] * p[1];
    q[2] = p[0] * p[2];
    q[3] = p[0] * p[3];
    q[4] = p[1] * p[1];
    q[5] = p[1] * p[2];
    q[6] = p[1] * p[3];
    q[7] = p[2] * p[2];
    q[8] = p[2] * p[3];
    q[9] = p[3] * p[3];
}

// index of (i,j) in the Q10 layout.  Not recursive: a recursive form is
// turned into a loop before the unrolled callers' constant indices reach it,
// which leaves the index dynamic and puts the indexed quadric in scratch
// memory.
__host__ __device__ constexpr int qidx(int i, int j)
{
    const int a = i < j ? i : j, b = i < j ? j : i;
    return a == 0 ? b : (a == 1 ? 3 + b : (a == 2 ? 5 + b : 9));
}

} // namespace ovx
