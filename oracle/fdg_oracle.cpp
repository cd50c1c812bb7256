// TEST INFRASTRUCTURE -- the parity oracle for mesh_to_flexible_dual_grid.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, and only as the checker (or as the timed CPU baseline).  The
// product path (o-voxel-gpu_amd/) never links, loads or calls it.
//
// What it is: a CPU restatement, without Eigen, of the reference CPU
// algorithm in src/convert/flexible_dual_grid.cpp (cited below as FDG.cpp).
// Two variants share one code path:
//   * the float path (prec = 0) mirrors FDG.cpp's arithmetic operation by
//     operation: FP64 scanline, FP32 QEF accumulation in first-insertion
//     order, FP32 column-pivoted Householder QR solve.  Built with
//     -O3 -std=c++17 -ffp-contract=off and no -march, so no FMA contraction
//     and no AVX (the reference compiles with -O3 -std=c++17, setup.py:65-68).
//   * the FP64 path (prec = 1) keeps every membership decision of the float
//     path (which voxels exist, which face / boundary terms reach which voxel)
//     and evaluates every QEF term, the accumulation and the solve in double.
//     It is the "exact" reference the float path's rounding is measured
//     against.
//
// Eigen evaluation orders that no reference test pins are restated from our
// reading of Eigen 3.4 (SSE2 baseline, no AVX):
//   * Vector3f dot / squaredNorm: a0 + (a1 + a2)   (unrolled redux halves)
//   * Vector3d squaredNorm:       (a0 + a1) + a2   (one Packet2d + tail)
//   * Vector4f/Matrix4f p^T Q p:  (x0 + x2) + (x1 + x3) per 4-term sum
//     (Packet4f predux), first p^T Q into a row temp, then the inner product
//   * ColPivHouseholderQR: LAPACK xGEQPF pivoting with the LAWN-176 norm
//     downdate, nonzero-pivot threshold (max_norm*eps)^2/rows*(rows-k),
//     Householder apply tmp = (sum_i e_i r_i) + r_0, column-oriented
//     back substitution that skips zero right-hand-side entries.
// These orders are "parity unpinned" against the Eigen build (Eigen is not
// available here); the GPU kernels use the same orders, so GPU-vs-oracle
// parity does not depend on them.

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// small helpers with the reference's evaluation orders
// ---------------------------------------------------------------------------

// Eigen Vector3f dot / squaredNorm (FDG.cpp:182,184,308,310,329-330,789)
template <typename T>
inline T dot3(const T *a, const T *b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }

// Eigen Vector3d squaredNorm with SSE2 packets (FDG.cpp:434,437)
inline double sqnorm3d_packet(const double *a) { return (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]; }

// Eigen Vector2f dot (FDG.cpp:338-401)
inline float dot2(float a0, float a1, float b0, float b1) { return a0 * b0 + a1 * b1; }

// std::min / std::max semantics (FDG.cpp:314-315, 338-360: cwiseMin, cwiseMax)
inline float smin(float a, float b) { return (b < a) ? b : a; }
inline float smax(float a, float b) { return (a < b) ? b : a; }

inline int iclamp(int v, int lo, int hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

// FDG.cpp:32-39 -- lerp with the exact a == b guard, in double, no FMA.
inline double lerp1(double a, double b, double t, double va, double vb)
{
    if (a == b)
        return va;
    double alpha = (t - a) / (b - a);
    return (1.0 - alpha) * va + alpha * vb;
}

// 4-term sum in Packet4f predux order.
template <typename T>
inline T sum4_predux(T x0, T x1, T x2, T x3) { return (x0 + x2) + (x1 + x3); }

// p^T Q p for Vector4f p, Matrix4f Q (FDG.cpp:836,857,891,909,927,945,970).
template <typename T>
inline T quad_form4(const T *p, const T *Q /* row-major 4x4, symmetric */)
{
    T t[4];
    for (int j = 0; j < 4; ++j)
        t[j] = sum4_predux(p[0] * Q[0 * 4 + j], p[1] * Q[1 * 4 + j], p[2] * Q[2 * 4 + j], p[3] * Q[3 * 4 + j]);
    return sum4_predux(t[0] * p[0], t[1] * p[1], t[2] * p[2], t[3] * p[3]);
}

// ---------------------------------------------------------------------------
// Column-pivoted Householder QR solve of an n x n system (n = 2 or 3),
// restating Eigen's ColPivHouseholderQR::computeInPlace / _solve_impl as used
// at FDG.cpp:797 and FDG.cpp:821-847.
// A is row-major n x n.
// ---------------------------------------------------------------------------
template <typename T>
struct Eps;
template <>
struct Eps<float>
{
    static float eps() { return FLT_EPSILON; }
    static float tiny() { return FLT_MIN; }
};
template <>
struct Eps<double>
{
    static double eps() { return DBL_EPSILON; }
    static double tiny() { return DBL_MIN; }
};

template <typename T>
inline T col_norm(const T *qr, int n, int col, int row0)
{
    // squaredNorm of column `col` from row0 on.  For the full 3-column
    // (fixed size) case Eigen's unrolled redux gives a0 + (a1 + a2); for the
    // dynamic-size tails used in the downdate it is a sequential sum.
    int len = n - row0;
    if (row0 == 0 && n == 3)
    {
        T a0 = qr[0 * 3 + col], a1 = qr[1 * 3 + col], a2 = qr[2 * 3 + col];
        return std::sqrt(a0 * a0 + (a1 * a1 + a2 * a2));
    }
    T s = qr[row0 * n + col] * qr[row0 * n + col];
    for (int i = 1; i < len; ++i)
        s = s + qr[(row0 + i) * n + col] * qr[(row0 + i) * n + col];
    return std::sqrt(s);
}

template <typename T, int n>
void cpqr_solve(const T *A, const T *b, T *x)
{
    T qr[n * n];
    for (int i = 0; i < n * n; ++i)
        qr[i] = A[i];
    T hcoeff[n];
    int transp[n];
    T norms_upd[n], norms_dir[n];
    for (int k = 0; k < n; ++k)
    {
        norms_dir[k] = col_norm<T>(qr, n, k, 0);
        norms_upd[k] = norms_dir[k];
    }
    // maxCoeff: first index of the maximum
    T maxn = norms_upd[0];
    for (int k = 1; k < n; ++k)
        if (norms_upd[k] > maxn)
            maxn = norms_upd[k];
    const T eps = Eps<T>::eps();
    const T me = maxn * eps;
    const T threshold_helper = (me * me) / T(n);
    const T norm_downdate_threshold = std::sqrt(eps);
    int nonzero_pivots = n;

    for (int k = 0; k < n; ++k)
    {
        int big = k;
        T bigv = norms_upd[k];
        for (int j = k + 1; j < n; ++j)
            if (norms_upd[j] > bigv)
            {
                bigv = norms_upd[j];
                big = j;
            }
        const T big_sq = bigv * bigv;
        if (nonzero_pivots == n && big_sq < threshold_helper * T(n - k))
            nonzero_pivots = k;
        transp[k] = big;
        if (k != big)
        {
            for (int r = 0; r < n; ++r)
                std::swap(qr[r * n + k], qr[r * n + big]);
            std::swap(norms_upd[k], norms_upd[big]);
            std::swap(norms_dir[k], norms_dir[big]);
        }
        // makeHouseholderInPlace on column k rows k..n-1
        const int tl = n - k - 1;
        T tail_sq = T(0);
        if (tl > 0)
        {
            tail_sq = qr[(k + 1) * n + k] * qr[(k + 1) * n + k];
            for (int i = 1; i < tl; ++i)
                tail_sq = tail_sq + qr[(k + 1 + i) * n + k] * qr[(k + 1 + i) * n + k];
        }
        const T c0 = qr[k * n + k];
        T beta, tau;
        if (tail_sq <= Eps<T>::tiny())
        {
            tau = T(0);
            beta = c0;
            for (int i = 0; i < tl; ++i)
                qr[(k + 1 + i) * n + k] = T(0);
        }
        else
        {
            beta = std::sqrt(c0 * c0 + tail_sq);
            if (c0 >= T(0))
                beta = -beta;
            const T den = c0 - beta;
            for (int i = 0; i < tl; ++i)
                qr[(k + 1 + i) * n + k] = qr[(k + 1 + i) * n + k] / den;
            tau = (beta - c0) / beta;
        }
        hcoeff[k] = tau;
        qr[k * n + k] = beta;
        // applyHouseholderOnTheLeft to the bottom-right corner (rows k.., cols k+1..)
        if (n - k > 1 && tau != T(0))
        {
            for (int j = k + 1; j < n; ++j)
            {
                T tmp = qr[(k + 1) * n + k] * qr[(k + 1) * n + j];
                for (int i = 1; i < tl; ++i)
                    tmp = tmp + qr[(k + 1 + i) * n + k] * qr[(k + 1 + i) * n + j];
                tmp = tmp + qr[k * n + j];
                qr[k * n + j] = qr[k * n + j] - tau * tmp;
                for (int i = 0; i < tl; ++i)
                {
                    const T te = tau * qr[(k + 1 + i) * n + k];
                    qr[(k + 1 + i) * n + j] = qr[(k + 1 + i) * n + j] - tmp * te;
                }
            }
        }
        // LAWN-176 norm downdate
        for (int j = k + 1; j < n; ++j)
        {
            if (norms_upd[j] != T(0))
            {
                T a = qr[k * n + j];
                a = a < T(0) ? -a : a;
                T temp = a / norms_upd[j];
                temp = (T(1) + temp) * (T(1) - temp);
                temp = temp < T(0) ? T(0) : temp;
                const T ratio = norms_upd[j] / norms_dir[j];
                const T temp2 = temp * (ratio * ratio);
                if (temp2 <= norm_downdate_threshold)
                {
                    norms_dir[j] = (k + 1 < n) ? col_norm<T>(qr, n, j, k + 1) : T(0);
                    norms_upd[j] = norms_dir[j];
                }
                else
                {
                    norms_upd[j] = norms_upd[j] * std::sqrt(temp);
                }
            }
        }
    }
    // column permutation from the transpositions
    int perm[n];
    for (int i = 0; i < n; ++i)
        perm[i] = i;
    for (int k = 0; k < n; ++k)
        std::swap(perm[k], perm[transp[k]]);

    const int r = nonzero_pivots;
    if (r == 0)
    {
        for (int i = 0; i < n; ++i)
            x[i] = T(0);
        return;
    }
    T c[n];
    for (int i = 0; i < n; ++i)
        c[i] = b[i];
    // c = Q^T c: reflectors 0..r-1 in order
    for (int k = 0; k < r; ++k)
    {
        const T tau = hcoeff[k];
        const int rows = n - k;
        if (rows == 1)
        {
            c[k] = c[k] * (T(1) - tau);
        }
        else if (tau != T(0))
        {
            const int tl = rows - 1;
            T tmp = qr[(k + 1) * n + k] * c[k + 1];
            for (int i = 1; i < tl; ++i)
                tmp = tmp + qr[(k + 1 + i) * n + k] * c[k + 1 + i];
            tmp = tmp + c[k];
            c[k] = c[k] - tau * tmp;
            for (int i = 0; i < tl; ++i)
            {
                const T te = tau * qr[(k + 1 + i) * n + k];
                c[k + 1 + i] = c[k + 1 + i] - tmp * te;
            }
        }
    }
    // upper-triangular solve on the leading r x r block, column oriented
    for (int i = r - 1; i >= 0; --i)
    {
        if (c[i] != T(0))
        {
            c[i] = c[i] / qr[i * n + i];
            for (int s = 0; s < i; ++s)
                c[s] = c[s] - c[i] * qr[s * n + i];
        }
    }
    for (int i = 0; i < r; ++i)
        x[perm[i]] = c[i];
    for (int i = r; i < n; ++i)
        x[perm[i]] = T(0);
}

// ---------------------------------------------------------------------------
// voxel hash table of the reference (FDG.cpp:48-78)
// ---------------------------------------------------------------------------
struct VoxelCoord
{
    int x, y, z;
    bool operator==(const VoxelCoord &o) const { return x == o.x && y == o.y && z == o.z; }
};
struct VoxelHash
{
    size_t operator()(const VoxelCoord &v) const
    {
        const size_t p1 = 73856093, p2 = 19349663, p3 = 83492791;
        return ((size_t)(v.x) * p1) ^ ((size_t)(v.y) * p2) ^ ((size_t)(v.z) * p3);
    }
};
using VoxelMap = std::unordered_map<VoxelCoord, size_t, VoxelHash>;

struct Grid
{
    float vs[3];
    int gmin[3], gmax[3];
};

// ---------------------------------------------------------------------------
// scanline event generator (FDG.cpp:195-283, identical in intersect_occ
// FDG.cpp:96-151).  Calls emit(ax0, ax1, ax2, x_idx, y_idx, z_idx, x, y, z)
// once per in-range grid-line crossing, in the reference's traversal order.
// ---------------------------------------------------------------------------
struct V3d
{
    double x, y, z;
};

template <typename Emit>
inline void scan_axis(const float *v0, const float *v1, const float *v2, const Grid &g, int ax2, Emit &&emit)
{
    const int ax0 = (ax2 + 1) % 3;
    const int ax1 = (ax2 + 2) % 3;
    V3d t[3] = {{(double)v0[ax0], (double)v0[ax1], (double)v0[ax2]},
                {(double)v1[ax0], (double)v1[ax1], (double)v1[ax2]},
                {(double)v2[ax0], (double)v2[ax1], (double)v2[ax2]}};
    // std::sort of 3 elements by .y: libstdc++ uses insertion sort below
    // 16 elements, which is stable; a stable 3-sort gives the same order.
    if (t[1].y < t[0].y)
        std::swap(t[0], t[1]);
    if (t[2].y < t[1].y)
    {
        std::swap(t[1], t[2]);
        if (t[1].y < t[0].y)
            std::swap(t[0], t[1]);
    }
    const double vs0 = (double)g.vs[ax0], vs1 = (double)g.vs[ax1], vs2 = (double)g.vs[ax2];
    const int start = iclamp((int)(t[0].y / vs1), g.gmin[ax1], g.gmax[ax1] - 1);
    const int mid = iclamp((int)(t[1].y / vs1), g.gmin[ax1], g.gmax[ax1] - 1);
    const int end = iclamp((int)(t[2].y / vs1), g.gmin[ax1], g.gmax[ax1] - 1);

    auto half = [&](int row_start, int row_end, const V3d &t0, const V3d &t1, const V3d &t2) {
        for (int y_idx = row_start; y_idx < row_end; ++y_idx)
        {
            const double y = (double)((float)(y_idx + 1) * g.vs[ax1]);
            double t3x = lerp1(t0.y, t1.y, y, t0.x, t1.x);
            double t3z = lerp1(t0.y, t1.y, y, t0.z, t1.z);
            double t4x = lerp1(t0.y, t2.y, y, t0.x, t2.x);
            double t4z = lerp1(t0.y, t2.y, y, t0.z, t2.z);
            if (t3x > t4x)
            {
                std::swap(t3x, t4x);
                std::swap(t3z, t4z);
            }
            const int line_start = iclamp((int)(t3x / vs0), g.gmin[ax0], g.gmax[ax0] - 1);
            const int line_end = iclamp((int)(t4x / vs0), g.gmin[ax0], g.gmax[ax0] - 1);
            for (int x_idx = line_start; x_idx < line_end; ++x_idx)
            {
                const double x = (double)((float)(x_idx + 1) * g.vs[ax0]);
                const double z = lerp1(t3x, t4x, x, t3z, t4z);
                const int z_idx = (int)(z / vs2);
                if (z_idx >= g.gmin[ax2] && z_idx < g.gmax[ax2])
                    emit(ax0, ax1, ax2, x_idx, y_idx, z_idx, x, y, z);
            }
        }
    };
    half(start, mid, t[0], t[1], t[2]);
    half(mid, end, t[2], t[1], t[0]);
}

// ---------------------------------------------------------------------------
// triangle plane and its quadric (FDG.cpp:180-185 and 305-311)
// ---------------------------------------------------------------------------
template <typename T>
inline void triangle_plane(const float *v0f, const float *v1f, const float *v2f, T plane[4], T n_out[3])
{
    T v0[3] = {(T)v0f[0], (T)v0f[1], (T)v0f[2]};
    T e0[3], e1[3];
    for (int i = 0; i < 3; ++i)
    {
        e0[i] = (T)v1f[i] - (T)v0f[i];
        e1[i] = (T)v2f[i] - (T)v1f[i];
    }
    T c[3] = {e0[1] * e1[2] - e0[2] * e1[1],
              e0[2] * e1[0] - e0[0] * e1[2],
              e0[0] * e1[1] - e0[1] * e1[0]};
    const T s = dot3(c, c);
    T n[3];
    if (s > T(0))
    {
        const T r = std::sqrt(s);
        n[0] = c[0] / r;
        n[1] = c[1] / r;
        n[2] = c[2] / r;
    }
    else
    {
        n[0] = c[0];
        n[1] = c[1];
        n[2] = c[2];
    }
    plane[0] = n[0];
    plane[1] = n[1];
    plane[2] = n[2];
    plane[3] = -dot3(n, v0);
    if (n_out)
    {
        n_out[0] = n[0];
        n_out[1] = n[1];
        n_out[2] = n[2];
    }
}

// ---------------------------------------------------------------------------
// triangle / voxel-box overlap test of the face stage (FDG.cpp:304-413).
// Calls hit(x, y, z) for every bbox voxel that passes all ten tests, in the
// reference's z, y, x loop order.  Always evaluated in float.
// ---------------------------------------------------------------------------
template <typename Hit>
inline void face_test(const float *v0, const float *v1, const float *v2, const Grid &g, Hit &&hit)
{
    const float *vs = g.vs;
    float e0[3], e1[3], e2[3];
    for (int i = 0; i < 3; ++i)
    {
        e0[i] = v1[i] - v0[i];
        e1[i] = v2[i] - v1[i];
        e2[i] = v0[i] - v2[i];
    }
    float plane[4], n[3];
    triangle_plane<float>(v0, v1, v2, plane, n);

    int bb_min[3], bb_max[3];
    for (int i = 0; i < 3; ++i)
    {
        const float mn = smin(smin(v0[i], v1[i]), v2[i]) / vs[i];
        const float mx = smax(smax(v0[i], v1[i]), v2[i]) / vs[i];
        bb_min[i] = std::max((int)mn, g.gmin[i]);
        bb_max[i] = std::min((int)(mx + 1.0f), g.gmax[i]);
    }
    const float c[3] = {n[0] > 0.0f ? vs[0] : 0.0f, n[1] > 0.0f ? vs[1] : 0.0f, n[2] > 0.0f ? vs[2] : 0.0f};
    const float cm[3] = {c[0] - v0[0], c[1] - v0[1], c[2] - v0[2]};
    const float d1 = dot3(n, cm);
    const float vc[3] = {(vs[0] - c[0]) - v0[0], (vs[1] - c[1]) - v0[1], (vs[2] - c[2]) - v0[2]};
    const float d2 = dot3(n, vc);

    // Edge normals of the three axis projections.  (-mul) * e is exact, so
    // the float products equal the reference's int-times-float products.
    const float *E[3] = {e0, e1, e2};
    const float *P[3] = {v0, v1, v2};
    float nxy[3][2], dxy[3], nyz[3][2], dyz[3], nzx[3][2], dzx[3];
    const float mxy = (n[2] < 0.0f) ? -1.0f : 1.0f;
    const float myz = (n[0] < 0.0f) ? -1.0f : 1.0f;
    const float mzx = (n[1] < 0.0f) ? -1.0f : 1.0f;
    for (int k = 0; k < 3; ++k)
    {
        const float *e = E[k];
        const float *p = P[k];
        nxy[k][0] = (-mxy) * e[1];
        nxy[k][1] = mxy * e[0];
        dxy[k] = -dot2(nxy[k][0], nxy[k][1], p[0], p[1]) + dot2(smax(nxy[k][0], 0.0f), smax(nxy[k][1], 0.0f), vs[0], vs[1]);
        nyz[k][0] = (-myz) * e[2];
        nyz[k][1] = myz * e[1];
        dyz[k] = -dot2(nyz[k][0], nyz[k][1], p[1], p[2]) + dot2(smax(nyz[k][0], 0.0f), smax(nyz[k][1], 0.0f), vs[1], vs[2]);
        nzx[k][0] = (-mzx) * e[0];
        nzx[k][1] = mzx * e[2];
        dzx[k] = -dot2(nzx[k][0], nzx[k][1], p[2], p[0]) + dot2(smax(nzx[k][0], 0.0f), smax(nzx[k][1], 0.0f), vs[2], vs[0]);
    }

    for (int z = bb_min[2]; z < bb_max[2]; ++z)
        for (int y = bb_min[1]; y < bb_max[1]; ++y)
            for (int x = bb_min[0]; x < bb_max[0]; ++x)
            {
                const float p[3] = {vs[0] * (float)x, vs[1] * (float)y, vs[2] * (float)z};
                const float ndp = dot3(n, p);
                if (((ndp + d1) * (ndp + d2)) > 0.0f)
                    continue;
                bool ok = true;
                for (int k = 0; k < 3 && ok; ++k)
                    if (dot2(nxy[k][0], nxy[k][1], p[0], p[1]) + dxy[k] < 0.0f)
                        ok = false;
                for (int k = 0; k < 3 && ok; ++k)
                    if (dot2(nyz[k][0], nyz[k][1], p[1], p[2]) + dyz[k] < 0.0f)
                        ok = false;
                for (int k = 0; k < 3 && ok; ++k)
                    if (dot2(nzx[k][0], nzx[k][1], p[2], p[0]) + dzx[k] < 0.0f)
                        ok = false;
                if (ok)
                    hit(x, y, z);
            }
}

// ---------------------------------------------------------------------------
// boundary segment quadric and DDA (FDG.cpp:427-507).
// Returns false for the skipped (degenerate) segments.  Q is row-major 4x4.
// ---------------------------------------------------------------------------
template <typename T>
inline bool boundary_quadric(const float *v0, const float *v1, T Q[16], double dir_out[3], double *len_out)
{
    double dir[3] = {(double)(v1[0] - v0[0]), (double)(v1[1] - v0[1]), (double)(v1[2] - v0[2])};
    const double len = std::sqrt(sqnorm3d_packet(dir));
    if (len < 1e-6)
        return false;
    const double z2 = sqnorm3d_packet(dir);
    if (z2 > 0.0)
    {
        const double r = std::sqrt(z2);
        dir[0] /= r;
        dir[1] /= r;
        dir[2] /= r;
    }
    T A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            A[i][j] = (i == j ? T(1) : T(0)) - (T)(dir[i] * dir[j]);
    const T p[3] = {(T)v0[0], (T)v0[1], (T)v0[2]};
    T b[3];
    for (int i = 0; i < 3; ++i)
        b[i] = -(A[i][0] * p[0] + (A[i][1] * p[1] + A[i][2] * p[2]));
    T t[3];
    for (int j = 0; j < 3; ++j)
        t[j] = p[0] * A[0][j] + (p[1] * A[1][j] + p[2] * A[2][j]);
    const T c = t[0] * p[0] + (t[1] * p[1] + t[2] * p[2]);
    for (int i = 0; i < 3; ++i)
    {
        for (int j = 0; j < 3; ++j)
            Q[i * 4 + j] = A[i][j];
        Q[i * 4 + 3] = b[i];
        Q[3 * 4 + i] = b[i];
    }
    Q[15] = c;
    for (int i = 0; i < 3; ++i)
        dir_out[i] = dir[i];
    *len_out = len;
    return true;
}

template <typename Visit>
inline void boundary_dda(const float *v0, const double *dir, double len, const Grid &g, Visit &&visit)
{
    int cur[3], step[3];
    double tMax[3], tDelta[3];
    for (int a = 0; a < 3; ++a)
    {
        cur[a] = (int)std::floor(v0[a] / g.vs[a]);
        step[a] = dir[a] > 0.0 ? 1 : -1;
    }
    for (int a = 0; a < 3; ++a)
    {
        if (dir[a] == 0.0)
        {
            tMax[a] = std::numeric_limits<double>::infinity();
            tDelta[a] = std::numeric_limits<double>::infinity();
        }
        else
        {
            const float border = g.vs[a] * (float)(cur[a] + (step[a] > 0 ? 1 : 0));
            tMax[a] = (double)(border - v0[a]) / dir[a];
            tDelta[a] = (double)g.vs[a] / std::fabs(dir[a]);
        }
    }
    auto in_grid = [&](const int *c) {
        return !((c[0] < g.gmin[0] || c[0] >= g.gmax[0]) || (c[1] < g.gmin[1] || c[1] >= g.gmax[1]) ||
                 (c[2] < g.gmin[2] || c[2] >= g.gmax[2]));
    };
    if (in_grid(cur))
        visit(cur[0], cur[1], cur[2]);
    while (true)
    {
        int axis;
        if (tMax[0] < tMax[1])
            axis = (tMax[0] < tMax[2]) ? 0 : 2;
        else
            axis = (tMax[1] < tMax[2]) ? 1 : 2;
        if (tMax[axis] > len)
            break;
        cur[axis] += step[axis];
        tMax[axis] += tDelta[axis];
        if (in_grid(cur))
            visit(cur[0], cur[1], cur[2]);
    }
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
struct oracle_fdg_result
{
    std::vector<int32_t> coords;   // [N,3]
    std::vector<uint8_t> flags;    // [N,3]
    std::vector<float> dual_f32;   // [N,3]  (prec 0)
    std::vector<double> dual_f64;  // [N,3]  (prec 1)
    std::vector<double> qef;       // [N,16] accumulated quadric before regularization
    std::vector<double> mean;      // [N,3]  mean of intersection points
    std::vector<double> cnt;       // [N]
    int64_t n_boundary_edges = 0;
    int64_t n_events = 0;
    int64_t n_face_hits = 0;
    int64_t n_boundary_hits = 0;
    double t_intersect = 0, t_face = 0, t_boundary = 0, t_solve = 0;
};

namespace {

template <typename T>
void run_fdg(const float *verts, int64_t V, const int32_t *faces, int64_t F, const Grid &g, float face_weight,
             float boundary_weight, float reg_weight, oracle_fdg_result &res, bool keep_qef)
{
    (void)V;
    VoxelMap table;
    std::vector<int32_t> &voxels = res.coords;
    std::vector<T> means, cnt, qefs;
    std::vector<uint8_t> &inter = res.flags;

    // ---- intersection stage (FDG.cpp:158-285, driver 695-709)
    for (int64_t f = 0; f < F; ++f)
    {
        const float *v0 = verts + 3 * (int64_t)faces[3 * f + 0];
        const float *v1 = verts + 3 * (int64_t)faces[3 * f + 1];
        const float *v2 = verts + 3 * (int64_t)faces[3 * f + 2];
        T plane[4];
        triangle_plane<T>(v0, v1, v2, plane, nullptr);
        T Q[16];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                Q[i * 4 + j] = plane[i] * plane[j];
        for (int ax2 = 0; ax2 < 3; ++ax2)
        {
            scan_axis(v0, v1, v2, g, ax2,
                      [&](int ax0, int ax1, int axz, int x_idx, int y_idx, int z_idx, double x, double y, double z) {
                          res.n_events++;
                          T pt[3];
                          pt[ax0] = (T)x;
                          pt[ax1] = (T)y;
                          pt[axz] = (T)z;
                          for (int dx = 0; dx < 2; ++dx)
                              for (int dy = 0; dy < 2; ++dy)
                              {
                                  VoxelCoord vc;
                                  int *cc = &vc.x;
                                  cc[ax0] = x_idx + dx;
                                  cc[ax1] = y_idx + dy;
                                  cc[axz] = z_idx;
                                  auto it = table.find(vc);
                                  if (it == table.end())
                                  {
                                      const size_t id = voxels.size() / 3;
                                      table.emplace(vc, id);
                                      voxels.push_back(vc.x);
                                      voxels.push_back(vc.y);
                                      voxels.push_back(vc.z);
                                      means.push_back(pt[0]);
                                      means.push_back(pt[1]);
                                      means.push_back(pt[2]);
                                      cnt.push_back(T(1));
                                      inter.push_back(0);
                                      inter.push_back(0);
                                      inter.push_back(0);
                                      for (int k = 0; k < 16; ++k)
                                          qefs.push_back(Q[k]);
                                      if (dx == 0 && dy == 0)
                                          inter[3 * id + axz] = 1;
                                  }
                                  else
                                  {
                                      const size_t id = it->second;
                                      means[3 * id + 0] += pt[0];
                                      means[3 * id + 1] += pt[1];
                                      means[3 * id + 2] += pt[2];
                                      cnt[id] += T(1);
                                      if (dx == 0 && dy == 0)
                                          inter[3 * id + axz] = 1;
                                      T *q = &qefs[16 * id];
                                      for (int k = 0; k < 16; ++k)
                                          q[k] += Q[k];
                                  }
                              }
                      });
        }
    }
    const size_t N = voxels.size() / 3;

    // ---- face stage (FDG.cpp:287-415, gated at FDG.cpp:715; Q added unscaled at 409)
    if (face_weight > 0.0f)
    {
        for (int64_t f = 0; f < F; ++f)
        {
            const float *v0 = verts + 3 * (int64_t)faces[3 * f + 0];
            const float *v1 = verts + 3 * (int64_t)faces[3 * f + 1];
            const float *v2 = verts + 3 * (int64_t)faces[3 * f + 2];
            T plane[4];
            triangle_plane<T>(v0, v1, v2, plane, nullptr);
            T Q[16];
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    Q[i * 4 + j] = plane[i] * plane[j];
            face_test(v0, v1, v2, g, [&](int x, int y, int z) {
                auto it = table.find(VoxelCoord{x, y, z});
                if (it != table.end())
                {
                    res.n_face_hits++;
                    T *q = &qefs[16 * it->second];
                    for (int k = 0; k < 16; ++k)
                        q[k] += Q[k];
                }
            });
        }
    }

    // ---- boundary stage (FDG.cpp:724-761, boundry_qef 417-524)
    if (boundary_weight > 0.0f)
    {
        std::map<std::pair<int, int>, int> edge_count;
        for (int64_t f = 0; f < F; ++f)
            for (int k = 0; k < 3; ++k)
            {
                int a = faces[3 * f + k];
                int b = faces[3 * f + (k + 1) % 3];
                if (a > b)
                    std::swap(a, b);
                edge_count[std::make_pair(a, b)]++;
            }
        const T bw = (T)boundary_weight;
        for (const auto &e : edge_count)
        {
            if (e.second != 1)
                continue;
            res.n_boundary_edges++;
            const float *v0 = verts + 3 * (int64_t)e.first.first;
            const float *v1 = verts + 3 * (int64_t)e.first.second;
            T Q[16];
            double dir[3], len;
            if (!boundary_quadric<T>(v0, v1, Q, dir, &len))
                continue;
            boundary_dda(v0, dir, len, g, [&](int x, int y, int z) {
                auto it = table.find(VoxelCoord{x, y, z});
                if (it == table.end())
                    return;
                res.n_boundary_hits++;
                T *q = &qefs[16 * it->second];
                for (int k = 0; k < 16; ++k)
                    q[k] += bw * Q[k];
            });
        }
    }

    if (keep_qef)
    {
        res.qef.assign(qefs.begin(), qefs.end());
        res.mean.assign(means.begin(), means.end());
        res.cnt.assign(cnt.begin(), cnt.end());
    }

    // ---- solve (FDG.cpp:763-983)
    std::vector<T> dual(3 * N);
    const T rw = (T)reg_weight;
    for (size_t i = 0; i < N; ++i)
    {
        const int *coord = &voxels[3 * i];
        T Q[16];
        std::memcpy(Q, &qefs[16 * i], sizeof(Q));
        T mn[3], mx[3];
        for (int a = 0; a < 3; ++a)
        {
            // int * float in the value precision (exact for the float path)
            mn[a] = (T)coord[a] * (T)g.vs[a];
            mx[a] = (T)(coord[a] + 1) * (T)g.vs[a];
        }
        if (reg_weight > 0.0f)
        {
            const T p[3] = {means[3 * i + 0] / cnt[i], means[3 * i + 1] / cnt[i], means[3 * i + 2] / cnt[i]};
            T R[16] = {1, 0, 0, -p[0], 0, 1, 0, -p[1], 0, 0, 1, -p[2], -p[0], -p[1], -p[2], dot3(p, p)};
            const T s = rw * cnt[i];
            for (int k = 0; k < 16; ++k)
                Q[k] = Q[k] + s * R[k];
        }
        T A[9] = {Q[0], Q[1], Q[2], Q[4], Q[5], Q[6], Q[8], Q[9], Q[10]};
        T b[3] = {-Q[3], -Q[7], -Q[11]};
        T v[3];
        cpqr_solve<T, 3>(A, b, v);
        if (!(v[0] >= mn[0] && v[0] <= mx[0] && v[1] >= mn[1] && v[1] <= mx[1] && v[2] >= mn[2] && v[2] <= mx[2]))
        {
            T best = std::numeric_limits<T>::infinity();
            auto consider = [&](const T p[4]) {
                const T err = quad_form4(p, Q);
                if (err < best)
                {
                    best = err;
                    v[0] = p[0];
                    v[1] = p[1];
                    v[2] = p[2];
                }
            };
            for (int fixed = 0; fixed < 3; ++fixed)
            {
                const int a1 = (fixed + 1) % 3, a2 = (fixed + 2) % 3;
                const T A2[4] = {Q[a1 * 4 + a1], Q[a1 * 4 + a2], Q[a2 * 4 + a1], Q[a2 * 4 + a2]};
                const T B[4] = {Q[a1 * 4 + fixed], Q[a1 * 4 + 3], Q[a2 * 4 + fixed], Q[a2 * 4 + 3]};
                for (int side = 0; side < 2; ++side)
                {
                    const T q0 = side ? mx[fixed] : mn[fixed];
                    const T rhs[2] = {(-B[0]) * q0 + (-B[1]) * T(1), (-B[2]) * q0 + (-B[3]) * T(1)};
                    T x2[2];
                    cpqr_solve<T, 2>(A2, rhs, x2);
                    if (x2[0] >= mn[a1] && x2[0] <= mx[a1] && x2[1] >= mn[a2] && x2[1] <= mx[a2])
                    {
                        T p[4];
                        p[fixed] = q0;
                        p[a1] = x2[0];
                        p[a2] = x2[1];
                        p[3] = T(1);
                        consider(p);
                    }
                }
            }
            for (int fr = 0; fr < 3; ++fr)
            {
                const int a1 = (fr + 1) % 3, a2 = (fr + 2) % 3;
                const T a = Q[fr * 4 + fr];
                const T bb[3] = {Q[fr * 4 + a1], Q[fr * 4 + a2], Q[fr * 4 + 3]};
                for (int s1 = 0; s1 < 2; ++s1)
                    for (int s2 = 0; s2 < 2; ++s2)
                    {
                        const T q[3] = {s1 ? mx[a1] : mn[a1], s2 ? mx[a2] : mn[a2], T(1)};
                        const T xv = -(dot3(bb, q)) / a;
                        if (xv >= mn[fr] && xv <= mx[fr])
                        {
                            T p[4];
                            p[fr] = xv;
                            p[a1] = q[0];
                            p[a2] = q[1];
                            p[3] = T(1);
                            consider(p);
                        }
                    }
            }
            for (int cx = 0; cx < 2; ++cx)
                for (int cy = 0; cy < 2; ++cy)
                    for (int cz = 0; cz < 2; ++cz)
                    {
                        // FDG.cpp:965-967: a set flag selects the MIN corner
                        const T p[4] = {cx ? mn[0] : mx[0], cy ? mn[1] : mx[1], cz ? mn[2] : mx[2], T(1)};
                        consider(p);
                    }
        }
        dual[3 * i + 0] = v[0];
        dual[3 * i + 1] = v[1];
        dual[3 * i + 2] = v[2];
    }
    if (sizeof(T) == sizeof(float))
        res.dual_f32.assign(dual.begin(), dual.end());
    else
    {
        res.dual_f64.assign(dual.begin(), dual.end());
        res.dual_f32.resize(dual.size());
        for (size_t k = 0; k < dual.size(); ++k)
            res.dual_f32[k] = (float)dual[k];
    }
}

} // namespace

extern "C" {

// Full conversion.  prec 0 = float path, 1 = FP64 path.  keep_qef != 0 keeps
// the accumulated per-voxel quadric, mean and count for inspection.
// Returns N, or -1 on invalid arguments.
int64_t oracle_fdg_run(const float *verts, int64_t V, const int32_t *faces, int64_t F, const float *voxel_size,
                       const int64_t *grid_range, float face_weight, float boundary_weight, float reg_weight, int prec,
                       int keep_qef, oracle_fdg_result **out)
{
    if (!out || (F > 0 && (!verts || !faces)))
        return -1;
    Grid g;
    for (int i = 0; i < 3; ++i)
    {
        g.vs[i] = voxel_size[i];
        g.gmin[i] = (int)grid_range[i];
        g.gmax[i] = (int)grid_range[3 + i];
    }
    for (int64_t k = 0; k < 3 * F; ++k)
        if (faces[k] < 0 || faces[k] >= V)
            return -1;
    auto *r = new oracle_fdg_result();
    if (prec == 0)
        run_fdg<float>(verts, V, faces, F, g, face_weight, boundary_weight, reg_weight, *r, keep_qef != 0);
    else
        run_fdg<double>(verts, V, faces, F, g, face_weight, boundary_weight, reg_weight, *r, keep_qef != 0);
    *out = r;
    return (int64_t)(r->coords.size() / 3);
}

// Occupancy only, triangle soup [T,3,3] (FDG.cpp:80-156, 605-644).
int64_t oracle_intersect_occ(const float *tris, int64_t T, const float *voxel_size, const int64_t *grid_range,
                             oracle_fdg_result **out)
{
    if (!out || (T > 0 && !tris))
        return -1;
    Grid g;
    for (int i = 0; i < 3; ++i)
    {
        g.vs[i] = voxel_size[i];
        g.gmin[i] = (int)grid_range[i];
        g.gmax[i] = (int)grid_range[3 + i];
    }
    auto *r = new oracle_fdg_result();
    VoxelMap table;
    for (int64_t t = 0; t < T; ++t)
    {
        const float *v0 = tris + 9 * t, *v1 = v0 + 3, *v2 = v0 + 6;
        for (int ax2 = 0; ax2 < 3; ++ax2)
            scan_axis(v0, v1, v2, g, ax2, [&](int ax0, int ax1, int axz, int x_idx, int y_idx, int z_idx, double, double, double) {
                for (int dx = 0; dx < 2; ++dx)
                    for (int dy = 0; dy < 2; ++dy)
                    {
                        VoxelCoord vc;
                        int *cc = &vc.x;
                        cc[ax0] = x_idx + dx;
                        cc[ax1] = y_idx + dy;
                        cc[axz] = z_idx;
                        if (table.find(vc) == table.end())
                        {
                            table.emplace(vc, r->coords.size() / 3);
                            r->coords.push_back(vc.x);
                            r->coords.push_back(vc.y);
                            r->coords.push_back(vc.z);
                        }
                    }
            });
    }
    *out = r;
    return (int64_t)(r->coords.size() / 3);
}

void oracle_fdg_copy(const oracle_fdg_result *r, int32_t *coords, float *dual_f32, double *dual_f64, uint8_t *flags)
{
    if (coords && !r->coords.empty())
        std::memcpy(coords, r->coords.data(), r->coords.size() * sizeof(int32_t));
    if (dual_f32 && !r->dual_f32.empty())
        std::memcpy(dual_f32, r->dual_f32.data(), r->dual_f32.size() * sizeof(float));
    if (dual_f64 && !r->dual_f64.empty())
        std::memcpy(dual_f64, r->dual_f64.data(), r->dual_f64.size() * sizeof(double));
    if (flags && !r->flags.empty())
        std::memcpy(flags, r->flags.data(), r->flags.size());
}

// Accumulated quadric (row-major 4x4, before regularization), mean sum and
// count per voxel; only filled when the run kept them.
int oracle_fdg_copy_qef(const oracle_fdg_result *r, double *qef16, double *mean3, double *cnt)
{
    if (r->qef.empty() && !r->coords.empty())
        return -1;
    if (qef16 && !r->qef.empty())
        std::memcpy(qef16, r->qef.data(), r->qef.size() * sizeof(double));
    if (mean3 && !r->mean.empty())
        std::memcpy(mean3, r->mean.data(), r->mean.size() * sizeof(double));
    if (cnt && !r->cnt.empty())
        std::memcpy(cnt, r->cnt.data(), r->cnt.size() * sizeof(double));
    return 0;
}

// Stage counters: events, face hits, boundary edges, boundary hits.
void oracle_fdg_stats(const oracle_fdg_result *r, int64_t *out4)
{
    out4[0] = r->n_events;
    out4[1] = r->n_face_hits;
    out4[2] = r->n_boundary_edges;
    out4[3] = r->n_boundary_hits;
}

void oracle_fdg_free(oracle_fdg_result *r) { delete r; }

// Building blocks exposed for unit tests.
void oracle_cpqr_solve3_f32(const float *A, const float *b, float *x) { cpqr_solve<float, 3>(A, b, x); }
void oracle_cpqr_solve2_f32(const float *A, const float *b, float *x) { cpqr_solve<float, 2>(A, b, x); }
void oracle_cpqr_solve3_f64(const double *A, const double *b, double *x) { cpqr_solve<double, 3>(A, b, x); }
void oracle_triangle_plane_f32(const float *v0, const float *v1, const float *v2, float *plane4)
{
    triangle_plane<float>(v0, v1, v2, plane4, nullptr);
}

// Scanline events of one triangle: writes up to cap rows of
// (ax2, x_idx, y_idx, z_idx) and returns the total count.
int64_t oracle_scan_events(const float *tri9, const float *voxel_size, const int64_t *grid_range, int32_t *out4,
                           int64_t cap)
{
    Grid g;
    for (int i = 0; i < 3; ++i)
    {
        g.vs[i] = voxel_size[i];
        g.gmin[i] = (int)grid_range[i];
        g.gmax[i] = (int)grid_range[3 + i];
    }
    int64_t n = 0;
    for (int ax2 = 0; ax2 < 3; ++ax2)
        scan_axis(tri9, tri9 + 3, tri9 + 6, g, ax2, [&](int, int, int axz, int x_idx, int y_idx, int z_idx, double, double, double) {
            if (n < cap)
            {
                out4[4 * n + 0] = axz;
                out4[4 * n + 1] = x_idx;
                out4[4 * n + 2] = y_idx;
                out4[4 * n + 3] = z_idx;
            }
            ++n;
        });
    return n;
}

// Face-stage hits of one triangle (all bbox voxels passing the ten tests,
// without the active-voxel filter): writes up to cap (x,y,z), returns count.
int64_t oracle_face_hits(const float *tri9, const float *voxel_size, const int64_t *grid_range, int32_t *out3,
                         int64_t cap)
{
    Grid g;
    for (int i = 0; i < 3; ++i)
    {
        g.vs[i] = voxel_size[i];
        g.gmin[i] = (int)grid_range[i];
        g.gmax[i] = (int)grid_range[3 + i];
    }
    int64_t n = 0;
    face_test(tri9, tri9 + 3, tri9 + 6, g, [&](int x, int y, int z) {
        if (n < cap)
        {
            out3[3 * n + 0] = x;
            out3[3 * n + 1] = y;
            out3[3 * n + 2] = z;
        }
        ++n;
    });
    return n;
}

// Boundary-segment DDA voxels (in-grid ones, in visit order) and quadric.
int64_t oracle_boundary_voxels(const float *seg6, const float *voxel_size, const int64_t *grid_range, int32_t *out3,
                               int64_t cap, float *Q16)
{
    Grid g;
    for (int i = 0; i < 3; ++i)
    {
        g.vs[i] = voxel_size[i];
        g.gmin[i] = (int)grid_range[i];
        g.gmax[i] = (int)grid_range[3 + i];
    }
    float Q[16];
    double dir[3], len;
    if (!boundary_quadric<float>(seg6, seg6 + 3, Q, dir, &len))
        return -1;
    if (Q16)
        std::memcpy(Q16, Q, sizeof(Q));
    int64_t n = 0;
    boundary_dda(seg6, dir, len, g, [&](int x, int y, int z) {
        if (n < cap)
        {
            out3[3 * n + 0] = x;
            out3[3 * n + 1] = y;
            out3[3 * n + 2] = z;
        }
        ++n;
    });
    return n;
}

} // extern "C"
