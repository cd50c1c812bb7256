// QEF solve of one voxel (FDG.cpp:763-983), device side.
// cpqr_solve restates Eigen's ColPivHouseholderQR (pivoting, LAWN-176 norm
// downdate, nonzero-pivot threshold, Householder application, column-oriented
// back substitution) for n = 2, 3; solve_free / solve_box are the unconstrained
// solve and the box-constrained search of the dual vertex.
#pragma once

#include "fdg_device.h"

namespace ovx
{

// ---------------------------------------------------------------------------
// float / double constants and functions of the factorisation
template <typename T>
struct Num;
template <>
struct Num<float>
{
    __device__ static constexpr float eps() { return FLT_EPSILON; }
    __device__ static constexpr float tiny() { return FLT_MIN; }
    __device__ static float sqrt_(float x) { return sqrtf(x); }
    __device__ static float abs_(float x) { return fabsf(x); }
};
template <>
struct Num<double>
{
    __device__ static constexpr double eps() { return DBL_EPSILON; }
    __device__ static constexpr double tiny() { return DBL_MIN; }
    __device__ static double sqrt_(double x) { return sqrt(x); }
    __device__ static double abs_(double x) { return fabs(x); }
};

template <int n, typename T = float>
struct Cpqr
{
    T qr[n + 1][n + 1]; // one spare row/column keeps the unrolled (dead) tail indices in bounds
    T hco[n];
    int perm[n];
    int rk;

    __device__ __forceinline__ void factor(const T (&A)[n][n]);
    __device__ __forceinline__ void solve(const T (&b)[n], T (&x)[n]) const;
};

// Householder QR with column pivoting of A (the factorisation does not depend
// on the right-hand side, so one factorisation serves several solves)
template <int n, typename T>
__device__ __forceinline__ void Cpqr<n, T>::factor(const T (&A)[n][n])
{
#pragma unroll
    for (int i = 0; i <= n; ++i)
#pragma unroll
        for (int j = 0; j <= n; ++j)
            qr[i][j] = (i < n && j < n) ? A[i][j] : T(0);
    int transp[n];
    T nu[n], nd[n];
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
        T s;
        if constexpr (n == 3)
            s = qr[0][k] * qr[0][k] + (qr[1][k] * qr[1][k] + qr[2][k] * qr[2][k]);
        else
            s = qr[0][k] * qr[0][k] + qr[1][k] * qr[1][k];
        nd[k] = Num<T>::sqrt_(s);
        nu[k] = nd[k];
    }
    T maxn = nu[0];
#pragma unroll
    for (int k = 1; k < n; ++k)
        if (nu[k] > maxn)
            maxn = nu[k];
    const T me = maxn * Num<T>::eps();
    const T th = (me * me) / (T)n;
    const T ndt = Num<T>::sqrt_(Num<T>::eps());
    int nzp = n;
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
        int big = k;
        T bigv = nu[k];
#pragma unroll
        for (int j = k + 1; j < n; ++j)
            if (nu[j] > bigv)
            {
                bigv = nu[j];
                big = j;
            }
        if (nzp == n && bigv * bigv < th * (T)(n - k))
            nzp = k;
        transp[k] = big;
#pragma unroll
        for (int j = k + 1; j < n; ++j)
            if (j == big)
            {
#pragma unroll
                for (int rr = 0; rr < n; ++rr)
                {
                    const T t = qr[rr][k];
                    qr[rr][k] = qr[rr][j];
                    qr[rr][j] = t;
                }
                T t = nu[k];
                nu[k] = nu[j];
                nu[j] = t;
                t = nd[k];
                nd[k] = nd[j];
                nd[j] = t;
            }
        const int tl = n - k - 1;
        T tail_sq = T(0);
        if (tl > 0)
        {
            tail_sq = qr[k + 1][k] * qr[k + 1][k];
#pragma unroll
            for (int i = 1; i < n - k - 1; ++i)
                tail_sq = tail_sq + qr[k + 1 + i][k] * qr[k + 1 + i][k];
        }
        const T c0 = qr[k][k];
        T beta, tau;
        if (tail_sq <= Num<T>::tiny())
        {
            tau = T(0);
            beta = c0;
#pragma unroll
            for (int i = 0; i < n - k - 1; ++i)
                qr[k + 1 + i][k] = T(0);
        }
        else
        {
            beta = Num<T>::sqrt_(c0 * c0 + tail_sq);
            if (c0 >= T(0))
                beta = -beta;
            const T den = c0 - beta;
#pragma unroll
            for (int i = 0; i < n - k - 1; ++i)
                qr[k + 1 + i][k] = qr[k + 1 + i][k] / den;
            tau = (beta - c0) / beta;
        }
        hco[k] = tau;
        qr[k][k] = beta;
        if (n - k > 1 && tau != T(0))
        {
#pragma unroll
            for (int j = k + 1; j < n; ++j)
            {
                T tmp = qr[k + 1][k] * qr[k + 1][j];
#pragma unroll
                for (int i = 1; i < n - k - 1; ++i)
                    tmp = tmp + qr[k + 1 + i][k] * qr[k + 1 + i][j];
                tmp = tmp + qr[k][j];
                qr[k][j] = qr[k][j] - tau * tmp;
#pragma unroll
                for (int i = 0; i < n - k - 1; ++i)
                {
                    const T te = tau * qr[k + 1 + i][k];
                    qr[k + 1 + i][j] = qr[k + 1 + i][j] - tmp * te;
                }
            }
        }
#pragma unroll
        for (int j = k + 1; j < n; ++j)
        {
            if (nu[j] != T(0))
            {
                const T a = Num<T>::abs_(qr[k][j]);
                T temp = a / nu[j];
                temp = (T(1) + temp) * (T(1) - temp);
                temp = temp < T(0) ? T(0) : temp;
                const T ratio = nu[j] / nd[j];
                const T temp2 = temp * (ratio * ratio);
                if (temp2 <= ndt)
                {
                    T s = T(0);
                    if (k + 1 < n)
                    {
                        s = qr[k + 1][j] * qr[k + 1][j];
#pragma unroll
                        for (int i = k + 2; i < n; ++i)
                            s = s + qr[i][j] * qr[i][j];
                        s = Num<T>::sqrt_(s);
                    }
                    nd[j] = s;
                    nu[j] = s;
                }
                else
                {
                    nu[j] = nu[j] * Num<T>::sqrt_(temp);
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < n; ++i)
        perm[i] = i;
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
#pragma unroll
        for (int j = k + 1; j < n; ++j)
            if (transp[k] == j)
            {
                const int t = perm[k];
                perm[k] = perm[j];
                perm[j] = t;
            }
    }
    rk = nzp;
}

template <int n, typename T>
__device__ __forceinline__ void Cpqr<n, T>::solve(const T (&b)[n], T (&x)[n]) const
{
    T c[n + 1];
#pragma unroll
    for (int i = 0; i < n; ++i)
        c[i] = b[i];
    c[n] = T(0);
    if (rk == 0)
    {
#pragma unroll
        for (int i = 0; i < n; ++i)
            x[i] = T(0);
        return;
    }
#pragma unroll
    for (int k = 0; k < n; ++k)
    {
        if (k < rk)
        {
            const T tau = hco[k];
            if (n - k == 1)
            {
                c[k] = c[k] * (T(1) - tau);
            }
            else if (tau != T(0))
            {
                T tmp = qr[k + 1][k] * c[k + 1];
#pragma unroll
                for (int i = 1; i < n - k - 1; ++i)
                    tmp = tmp + qr[k + 1 + i][k] * c[k + 1 + i];
                tmp = tmp + c[k];
                c[k] = c[k] - tau * tmp;
#pragma unroll
                for (int i = 0; i < n - k - 1; ++i)
                {
                    const T te = tau * qr[k + 1 + i][k];
                    c[k + 1 + i] = c[k + 1 + i] - tmp * te;
                }
            }
        }
    }
#pragma unroll
    for (int i = n - 1; i >= 0; --i)
    {
        if (i < rk && c[i] != T(0))
        {
            c[i] = c[i] / qr[i][i];
#pragma unroll
            for (int s = 0; s < i; ++s)
                c[s] = c[s] - c[i] * qr[s][i];
        }
    }
#pragma unroll
    for (int j = 0; j < n; ++j)
    {
        T v = T(0);
#pragma unroll
        for (int i = 0; i < n; ++i)
            if (perm[i] == j && i < rk)
                v = c[i];
        x[j] = v;
    }
}

template <int n>
__device__ __forceinline__ void cpqr_solve(const float (&A)[n][n], const float (&b)[n], float (&x)[n])
{
    Cpqr<n> f;
    f.factor(A);
    f.solve(b, x);
}

// p^T Q p in the Packet4f predux order (x0 + x2) + (x1 + x3)
__device__ __forceinline__ float quad_form(const float p[4], const float Q[4][4])
{
    float t[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        t[j] = (p[0] * Q[0][j] + p[2] * Q[2][j]) + (p[1] * Q[1][j] + p[3] * Q[3][j]);
    return (t[0] * p[0] + t[2] * p[2]) + (t[1] * p[1] + t[3] * p[3]);
}

__device__ __forceinline__ void load_q44(const float q[10], float (&Q)[4][4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            Q[i][j] = q[qidx(i, j)];
}

__device__ __forceinline__ void voxel_box(const int coord[3], const float vs[3], float (&mn)[3], float (&mx)[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        mn[a] = (float)coord[a] * vs[a];
        mx[a] = (float)(coord[a] + 1) * vs[a];
    }
}

// Unconstrained minimiser of the (regularised) quadric; true when it lies in
// the voxel's box, which makes it the dual vertex (FDG.cpp:795-812).
__device__ __forceinline__ bool solve_free(const float q[10], const int coord[3], const float vs[3], float v[3])
{
    float Q[4][4];
    load_q44(q, Q);
    float mn[3], mx[3];
    voxel_box(coord, vs, mn, mx);
    const float A[3][3] = {{Q[0][0], Q[0][1], Q[0][2]}, {Q[1][0], Q[1][1], Q[1][2]}, {Q[2][0], Q[2][1], Q[2][2]}};
    const float bv[3] = {-Q[0][3], -Q[1][3], -Q[2][3]};
    float x[3];
    cpqr_solve<3>(A, bv, x);
    v[0] = x[0];
    v[1] = x[1];
    v[2] = x[2];
    return v[0] >= mn[0] && v[0] <= mx[0] && v[1] >= mn[1] && v[1] <= mx[1] && v[2] >= mn[2] && v[2] <= mx[2];
}

// Box-constrained search of FDG.cpp:813-981 for a minimiser outside the box:
// the 6 faces (2x2 solves), 12 edges (1x1) and 8 corners, keeping the
// smallest p^T Q p with a strict <.  v enters as the unconstrained minimiser
// (kept if no candidate compares smaller, e.g. non-finite quadrics).
__device__ __forceinline__ void box_search(const float (&Q)[4][4], const float (&mn)[3], const float (&mx)[3], float v[3])
{
    float best = __builtin_huge_valf();
    auto consider = [&](const float p[4]) {
        const float err = quad_form(p, Q);
        if (err < best)
        {
            best = err;
            v[0] = p[0];
            v[1] = p[1];
            v[2] = p[2];
        }
    };
#pragma unroll
    for (int fixed = 0; fixed < 3; ++fixed)
    {
        const int a1 = (fixed + 1) % 3, a2 = (fixed + 2) % 3;
        const float A2[2][2] = {{Q[a1][a1], Q[a1][a2]}, {Q[a2][a1], Q[a2][a2]}};
        const float B[2][2] = {{Q[a1][fixed], Q[a1][3]}, {Q[a2][fixed], Q[a2][3]}};
        Cpqr<2> f2;
        f2.factor(A2); // both faces normal to `fixed` share the matrix
#pragma unroll
        for (int side = 0; side < 2; ++side)
        {
            const float q0 = side ? mx[fixed] : mn[fixed];
            const float rhs[2] = {(-B[0][0]) * q0 + (-B[0][1]) * 1.0f, (-B[1][0]) * q0 + (-B[1][1]) * 1.0f};
            float x2[2];
            f2.solve(rhs, x2);
            if (x2[0] >= mn[a1] && x2[0] <= mx[a1] && x2[1] >= mn[a2] && x2[1] <= mx[a2])
            {
                float p[4];
                p[fixed] = q0;
                p[a1] = x2[0];
                p[a2] = x2[1];
                p[3] = 1.0f;
                consider(p);
            }
        }
    }
#pragma unroll
    for (int fr = 0; fr < 3; ++fr)
    {
        const int a1 = (fr + 1) % 3, a2 = (fr + 2) % 3;
        const float a = Q[fr][fr];
        const float b0 = Q[fr][a1], b1 = Q[fr][a2], b2 = Q[fr][3];
#pragma unroll
        for (int s1 = 0; s1 < 2; ++s1)
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2)
            {
                const float q1 = s1 ? mx[a1] : mn[a1];
                const float q2 = s2 ? mx[a2] : mn[a2];
                const float xv = -(dot3f(b0, b1, b2, q1, q2, 1.0f)) / a;
                if (xv >= mn[fr] && xv <= mx[fr])
                {
                    float p[4];
                    p[fr] = xv;
                    p[a1] = q1;
                    p[a2] = q2;
                    p[3] = 1.0f;
                    consider(p);
                }
            }
    }
#pragma unroll
    for (int cx = 0; cx < 2; ++cx)
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int cz = 0; cz < 2; ++cz)
            {
                const float p[4] = {cx ? mn[0] : mx[0], cy ? mn[1] : mx[1], cz ? mn[2] : mx[2], 1.0f};
                consider(p);
            }
}

__device__ __forceinline__ void solve_box(const float q[10], const int coord[3], const float vs[3], float v[3])
{
    float Q[4][4];
    load_q44(q, Q);
    float mn[3], mx[3];
    voxel_box(coord, vs, mn, mx);
    box_search(Q, mn, mx, v);
}

// dual vertex of one voxel: the unconstrained minimiser, else the box search
__device__ __forceinline__ void solve_voxel(const float q[10], const int coord[3], const float vs[3], float out[3])
{
    if (!solve_free(q, coord, vs, out))
        solve_box(q, coord, vs, out);
}

} // namespace ovx
