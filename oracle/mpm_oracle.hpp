// TEST INFRASTRUCTURE -- not part of the product.
//
// CPU restatement of the reference MPM substep (yafei-ou/CRESSim-MPM v2.2.0),
// templated on the scalar type: T = float reproduces the reference CPU solvers
// bit for bit (built with -O2 -ffp-contract=off), T = double is the FP64 host
// replay used to report rounding error.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it, and only as the checker.
//
// Every function names the reference lines it restates.  The restatement keeps
// the reference's evaluation order (including its non-textbook details):
//   * computeStress falls through from co-rotational into linear elastic
//     (src/solver/constitutive_models.h:104-117);
//   * Vec3::outerProduct(a, b) builds the matrix M(r, c) = a_c * b_r
//     (include/math/mat33.h:287-306: the 9-float constructor is column-major);
//   * Mat3x3f::inverse() returns the transpose of the inverse, and the zero
//     matrix for |det| < 1e-6 (include/math/mat33.h:256-283);
//   * svd3 passes A^T to the row-major Sifakis kernel (include/math/svd3.h:1265-1275);
//   * node mass cutoffs 1e-4 (MLS, standard) and 1e-8 (PB) leave light nodes
//     holding raw momentum (src/solver/mls_mpm_algorithm.h:264-266).
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc
{
    // ------------------------------------------------------------------
    // Scalar helpers.  rsq follows the host fallback of rsqrtf
    // (include/utils/cuda_compat.h:55): 1 / sqrt(x), two roundings.
    // ------------------------------------------------------------------
    template <class T> inline T rsq(T x) { return T(1) / std::sqrt(x); }
    template <class T> inline T fmn(T a, T b) { return std::fmin(a, b); }
    template <class T> inline T fmx(T a, T b) { return std::fmax(a, b); }
    template <class T> inline T sq(T a) { return a * a; } // powf(a, 2): g++ folds it to a*a

    template <class T>
    struct V3
    {
        T x, y, z;
        V3() : x(0), y(0), z(0) {}
        V3(T a, T b, T c) : x(a), y(b), z(c) {}
        T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
        const T &operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    };
    template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return V3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
    template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return V3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
    template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return V3<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
    template <class T> inline V3<T> operator*(V3<T> a, T s) { return V3<T>(a.x * s, a.y * s, a.z * s); }
    template <class T> inline V3<T> operator/(V3<T> a, T s) { return V3<T>(a.x / s, a.y / s, a.z / s); }
    template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    template <class T> inline T sqn(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
    template <class T> inline T norm(V3<T> a) { return std::sqrt(sqn(a)); }
    template <class T> inline V3<T> cross(V3<T> a, V3<T> b)
    {
        return V3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
    }
    // Vec3::normalize (include/math/vec3.h:344-351)
    template <class T> inline void normalize(V3<T> &a)
    {
        T s = sqn(a);
        if (s > T(0))
            a = a * rsq(s);
    }

    // Column-major 3x3 (include/math/mat33.h:46-70); c[j] is column j.
    template <class T>
    struct M3
    {
        V3<T> c[3];
        T &operator()(int r, int col) { return c[col][r]; }
        T operator()(int r, int col) const { return c[col][r]; }
        static M3 zero() { return M3(); }
        static M3 identity()
        {
            M3 m;
            m.c[0].x = T(1);
            m.c[1].y = T(1);
            m.c[2].z = T(1);
            return m;
        }
        static M3 scalar(T v)
        {
            M3 m;
            m.c[0].x = v;
            m.c[1].y = v;
            m.c[2].z = v;
            return m;
        }
    };
    // Mat * vec = (col0*x + col1*y) + col2*z  (include/math/mat33.h:137-148)
    template <class T> inline V3<T> mul(const M3<T> &m, V3<T> v)
    {
        return m.c[0] * v.x + m.c[1] * v.y + m.c[2] * v.z;
    }
    template <class T> inline M3<T> mul(const M3<T> &a, const M3<T> &b)
    {
        M3<T> r;
        for (int j = 0; j < 3; ++j)
            r.c[j] = mul(a, b.c[j]);
        return r;
    }
    template <class T> inline M3<T> operator+(const M3<T> &a, const M3<T> &b)
    {
        M3<T> r;
        for (int j = 0; j < 3; ++j)
            r.c[j] = a.c[j] + b.c[j];
        return r;
    }
    template <class T> inline M3<T> operator-(const M3<T> &a, const M3<T> &b)
    {
        M3<T> r;
        for (int j = 0; j < 3; ++j)
            r.c[j] = a.c[j] - b.c[j];
        return r;
    }
    template <class T> inline M3<T> operator*(const M3<T> &a, T s)
    {
        M3<T> r;
        for (int j = 0; j < 3; ++j)
            r.c[j] = a.c[j] * s;
        return r;
    }
    template <class T> inline M3<T> transpose(const M3<T> &a)
    {
        M3<T> r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                r(i, j) = a(j, i);
        return r;
    }
    template <class T> inline T trace(const M3<T> &a) { return a(0, 0) + a(1, 1) + a(2, 2); }
    // include/math/mat33.h:209-222
    template <class T> inline T det(const M3<T> &m)
    {
        T a = m(0, 0), b = m(0, 1), c = m(0, 2), d = m(1, 0), e = m(1, 1), f = m(1, 2), g = m(2, 0), h = m(2, 1), i = m(2, 2);
        return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    }
    // include/math/mat33.h:256-283.  Columns are built from the cofactor rows,
    // so the result is inverse(m)^T; |det| < 1e-6 gives the zero matrix.
    template <class T> inline M3<T> inverse_t(const M3<T> &m)
    {
        T a = m(0, 0), b = m(0, 1), c = m(0, 2), d = m(1, 0), e = m(1, 1), f = m(1, 2), g = m(2, 0), h = m(2, 1), i = m(2, 2);
        T dt = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
        if (std::fabs(dt) < T(1e-6f))
            return M3<T>();
        T inv = T(1) / dt;
        M3<T> adj;
        adj.c[0] = V3<T>((e * i - f * h), -(b * i - c * h), (b * f - c * e));
        adj.c[1] = V3<T>(-(d * i - f * g), (a * i - c * g), -(a * f - c * d));
        adj.c[2] = V3<T>((d * h - e * g), -(a * h - b * g), (a * e - b * d));
        return adj * inv;
    }
    // Vec3::outerProduct (include/math/mat33.h:287-306): M(r, c) = a_c * b_r.
    template <class T> inline M3<T> outer_ref(V3<T> a, V3<T> b)
    {
        M3<T> m;
        m.c[0] = V3<T>(a.x * b.x, a.x * b.y, a.x * b.z);
        m.c[1] = V3<T>(a.y * b.x, a.y * b.y, a.y * b.z);
        m.c[2] = V3<T>(a.z * b.x, a.z * b.y, a.z * b.z);
        return m;
    }

    // Quaternion (x, y, z, w) rotations (include/math/quat.h:117-152).
    template <class T> inline V3<T> qrot(const T q[4], V3<T> v)
    {
        const T x = q[0], y = q[1], z = q[2], w = q[3];
        const T vx = T(2) * v.x, vy = T(2) * v.y, vz = T(2) * v.z;
        const T w2 = w * w - T(0.5f);
        const T dot2 = (x * vx + y * vy + z * vz);
        return V3<T>((vx * w2 + (y * vz - z * vy) * w + x * dot2),
                     (vy * w2 + (z * vx - x * vz) * w + y * dot2),
                     (vz * w2 + (x * vy - y * vx) * w + z * dot2));
    }
    template <class T> inline V3<T> qrot_inv(const T q[4], V3<T> v)
    {
        const T x = q[0], y = q[1], z = q[2], w = q[3];
        const T vx = T(2) * v.x, vy = T(2) * v.y, vz = T(2) * v.z;
        const T w2 = w * w - T(0.5f);
        const T dot2 = (x * vx + y * vy + z * vz);
        return V3<T>((vx * w2 - (y * vz - z * vy) * w + x * dot2),
                     (vy * w2 - (z * vx - x * vz) * w + y * dot2),
                     (vz * w2 - (x * vy - y * vx) * w + z * dot2));
    }

    // ------------------------------------------------------------------
    // 3x3 SVD: McAdams, Selle, Tamstorf, Teran, Sifakis (2011), "Computing
    // the singular value decomposition of 3x3 matrices with minimal
    // branching and elementary floating point operations", as the reference
    // evaluates it (include/math/svd3.h:74-1260): 4 Jacobi sweeps on A^T A
    // with quaternion accumulation, sorting, Givens QR.  Bit masks become
    // selects; the arithmetic sequence is the same.
    // ------------------------------------------------------------------
    template <class T>
    inline void svd_jacobi_conjugation(T &sPP, T &sQQ, T &sPQ, T &sRR, T &sRP, T &sRQ,
                                       T &qs, T &qR, T &qP, T &qQ)
    {
        const T tiny = T(1.e-20f), four_gamma_sq = T(5.82842712474619f);
        const T sin_pi8 = T(0.3826834323650897f), cos_pi8 = T(0.9238795325112867f);
        T sh = sPQ * T(0.5f);
        T t5 = sPP - sQQ;
        T t2 = sh * sh;
        bool m = t2 >= tiny;
        sh = m ? sh : T(0);
        T ch = m ? t5 : T(1);
        T t1 = sh * sh;
        t2 = ch * ch;
        T t3 = t1 + t2;
        T t4 = rsq(t3);
        sh = t4 * sh;
        ch = t4 * ch;
        t1 = four_gamma_sq * t1;
        m = t2 <= t1;
        sh = m ? sin_pi8 : sh;
        ch = m ? cos_pi8 : ch;
        t1 = sh * sh;
        t2 = ch * ch;
        T c = t2 - t1;
        T s = ch * sh;
        s = s + s;
        t3 = t1 + t2;
        sRR = sRR * t3;
        sRP = sRP * t3;
        sRQ = sRQ * t3;
        sRR = sRR * t3; // applied twice, as in the reference kernel
        t1 = s * sRP;
        t2 = s * sRQ;
        sRP = c * sRP;
        sRQ = c * sRQ;
        sRP = t2 + sRP;
        sRQ = sRQ - t1;
        t2 = s * s;
        t1 = sQQ * t2;
        t3 = sPP * t2;
        t4 = c * c;
        sPP = sPP * t4;
        sQQ = sQQ * t4;
        sPP = sPP + t1;
        sQQ = sQQ + t3;
        t4 = t4 - t2;
        t2 = sPQ + sPQ;
        sPQ = sPQ * t4;
        t4 = c * s;
        t2 = t2 * t4;
        t5 = t5 * t4;
        sPP = sPP + t2;
        sPQ = sPQ - t5;
        sQQ = sQQ - t2;
        const T tP = sh * qP, tQ = sh * qQ, tR = sh * qR;
        const T shq = sh * qs;
        qs = ch * qs;
        qP = ch * qP;
        qQ = ch * qQ;
        qR = ch * qR;
        qR = qR + shq;
        qs = qs - tR;
        qP = qP + tQ;
        qQ = qQ - tP;
    }

    // Newton-refined reciprocal square root used by the kernel's QR steps.
    template <class T> inline T svd_refined_rsq(T t2)
    {
        T t1 = rsq(t2);
        T t4 = t1 * T(0.5f);
        T t3 = t1 * t4;
        t3 = t1 * t3;
        t3 = t2 * t3;
        t1 = t1 + t4;
        t1 = t1 - t3;
        return t1;
    }

    // One Givens step of the QR stage: eliminates b[q][p] using b[p][p];
    // rotates rows p, q of b and columns p, q of u.
    template <class T> inline void svd_qr_givens(T b[3][3], T u[3][3], int p, int q)
    {
        const T small = T(1.e-12f);
        T sh = b[q][p] * b[q][p];
        sh = (sh >= small) ? b[q][p] : T(0);
        T ch = T(0) - b[p][p];
        ch = fmx(ch, b[p][p]);
        ch = fmx(ch, small);
        const bool m5 = b[p][p] >= T(0);
        T t1 = ch * ch;
        T t2 = sh * sh;
        t2 = t1 + t2;
        t1 = svd_refined_rsq(t2);
        t1 = t1 * t2;
        ch = ch + t1;
        if (!m5)
        {
            T tmp = ch;
            ch = sh;
            sh = tmp;
        }
        t1 = ch * ch;
        t2 = sh * sh;
        t2 = t1 + t2;
        t1 = svd_refined_rsq(t2);
        ch = ch * t1;
        sh = sh * t1;
        T c = ch * ch;
        T s = sh * sh;
        c = c - s;
        s = sh * ch;
        s = s + s;
        for (int j = 0; j < 3; ++j)
        {
            t1 = s * b[p][j];
            t2 = s * b[q][j];
            b[p][j] = c * b[p][j];
            b[q][j] = c * b[q][j];
            b[p][j] = b[p][j] + t2;
            b[q][j] = b[q][j] - t1;
        }
        for (int r = 0; r < 3; ++r)
        {
            t1 = s * u[r][p];
            t2 = s * u[r][q];
            u[r][p] = c * u[r][p];
            u[r][q] = c * u[r][q];
            u[r][p] = u[r][p] + t2;
            u[r][q] = u[r][q] - t1;
        }
    }

    // Conditional column swap of b and v (with rho swap) followed by the
    // sign flip of column `neg` when swapped (include/math/svd3.h:1000-1120).
    template <class T> inline void svd_sort_swap(T b[3][3], T v[3][3], T rho[3], int i, int j, int neg)
    {
        const bool m = rho[i] < rho[j];
        if (m)
        {
            for (int r = 0; r < 3; ++r)
            {
                T t = b[r][i];
                b[r][i] = b[r][j];
                b[r][j] = t;
                t = v[r][i];
                v[r][i] = v[r][j];
                v[r][j] = t;
            }
            T t = rho[i];
            rho[i] = rho[j];
            rho[j] = t;
        }
        const T sgn = m ? T(-1) : T(1);
        for (int r = 0; r < 3; ++r)
        {
            b[r][neg] = b[r][neg] * sgn;
            v[r][neg] = v[r][neg] * sgn;
        }
    }

    // Kernel-level SVD: a = u diag(sig) v^T with a, u, v row-major [row][col].
    template <class T> inline void sifakis_svd(const T a_in[3][3], T u[3][3], T v[3][3], T sig[3])
    {
        T a[3][3];
        std::memcpy(a, a_in, sizeof(a));
        // S = A^T A (lower triangle), include/math/svd3.h:426-455
        T s11 = a[0][0] * a[0][0];
        s11 = a[1][0] * a[1][0] + s11;
        s11 = a[2][0] * a[2][0] + s11;
        T s21 = a[0][1] * a[0][0];
        s21 = a[1][1] * a[1][0] + s21;
        s21 = a[2][1] * a[2][0] + s21;
        T s31 = a[0][2] * a[0][0];
        s31 = a[1][2] * a[1][0] + s31;
        s31 = a[2][2] * a[2][0] + s31;
        T s22 = a[0][1] * a[0][1];
        s22 = a[1][1] * a[1][1] + s22;
        s22 = a[2][1] * a[2][1] + s22;
        T s32 = a[0][2] * a[0][1];
        s32 = a[1][2] * a[1][1] + s32;
        s32 = a[2][2] * a[2][1] + s32;
        T s33 = a[0][2] * a[0][2];
        s33 = a[1][2] * a[1][2] + s33;
        s33 = a[2][2] * a[2][2] + s33;

        T qs = T(1), qx = T(0), qy = T(0), qz = T(0);
        for (int sweep = 0; sweep < 4; ++sweep)
        {
            svd_jacobi_conjugation(s11, s22, s21, s33, s31, s32, qs, qz, qx, qy);
            svd_jacobi_conjugation(s22, s33, s32, s11, s21, s31, qs, qx, qy, qz);
            svd_jacobi_conjugation(s33, s11, s31, s22, s32, s21, qs, qy, qz, qx);
        }

        // Normalise the quaternion (include/math/svd3.h:845-866)
        T t2 = qs * qs;
        T t1 = qx * qx;
        t2 = t1 + t2;
        t1 = qy * qy;
        t2 = t1 + t2;
        t1 = qz * qz;
        t2 = t1 + t2;
        t1 = svd_refined_rsq(t2);
        qs = qs * t1;
        qx = qx * t1;
        qy = qy * t1;
        qz = qz * t1;

        // Quaternion to V (include/math/svd3.h:868-893)
        t1 = qx * qx;
        t2 = qy * qy;
        T t3 = qz * qz;
        T v11 = qs * qs;
        T v22 = v11 - t1;
        T v33 = v22 - t2;
        v33 = v33 + t3;
        v22 = v22 + t2;
        v22 = v22 - t3;
        v11 = v11 + t1;
        v11 = v11 - t2;
        v11 = v11 - t3;
        t1 = qx + qx;
        t2 = qy + qy;
        t3 = qz + qz;
        T v32 = qs * t1;
        T v13 = qs * t2;
        T v21 = qs * t3;
        t1 = qy * t1;
        t2 = qz * t2;
        t3 = qx * t3;
        T v12 = t1 - v21;
        T v23 = t2 - v32;
        T v31 = t3 - v13;
        v21 = t1 + v21;
        v32 = t2 + v32;
        v13 = t3 + v13;
        v[0][0] = v11; v[0][1] = v12; v[0][2] = v13;
        v[1][0] = v21; v[1][1] = v22; v[1][2] = v23;
        v[2][0] = v31; v[2][1] = v32; v[2][2] = v33;

        // B = A V, row by row (include/math/svd3.h:894-945)
        for (int r = 0; r < 3; ++r)
        {
            T c2 = a[r][1], c3 = a[r][2];
            a[r][1] = v12 * a[r][0];
            a[r][2] = v13 * a[r][0];
            a[r][0] = v11 * a[r][0];
            a[r][0] = a[r][0] + v21 * c2;
            a[r][0] = a[r][0] + v31 * c3;
            a[r][1] = a[r][1] + v22 * c2;
            a[r][1] = a[r][1] + v32 * c3;
            a[r][2] = a[r][2] + v23 * c2;
            a[r][2] = a[r][2] + v33 * c3;
        }

        // Sort columns by decreasing norm (include/math/svd3.h:947-1120)
        T rho[3];
        for (int j = 0; j < 3; ++j)
        {
            T r0 = a[0][j] * a[0][j];
            r0 = r0 + a[1][j] * a[1][j];
            r0 = r0 + a[2][j] * a[2][j];
            rho[j] = r0;
        }
        svd_sort_swap(a, v, rho, 0, 1, 1);
        svd_sort_swap(a, v, rho, 0, 2, 0);
        svd_sort_swap(a, v, rho, 1, 2, 2);

        // QR by Givens rotations (include/math/svd3.h:1121-1238)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                u[i][j] = (i == j) ? T(1) : T(0);
        svd_qr_givens(a, u, 0, 1);
        svd_qr_givens(a, u, 0, 2);
        svd_qr_givens(a, u, 1, 2);
        sig[0] = a[0][0];
        sig[1] = a[1][1];
        sig[2] = a[2][2];
    }

    // svd3(A, u, vt, sigma) (include/math/svd3.h:1265-1275): the kernel sees
    // a_ij = A(j, i); u(i, j) = U_ij; vt(i, j) = V_ji.  So A^T = u sigma vt.
    template <class T> inline void svd3(const M3<T> &A, M3<T> &u, M3<T> &vt, V3<T> &sigma)
    {
        T a[3][3], uu[3][3], vv[3][3], s[3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                a[i][j] = A(j, i);
        sifakis_svd(a, uu, vv, s);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                u(i, j) = uu[i][j];
                vt(i, j) = vv[j][i];
            }
        sigma = V3<T>(s[0], s[1], s[2]);
    }

    // ------------------------------------------------------------------
    // Enumerations, values as in include/engine/geometry.h:42-55,
    // include/engine/shape.h:43-48, include/solver/material_data.h:41-46.
    // ------------------------------------------------------------------
    enum GeomType { eBox = 0, eSphere, ePlane, eCapsule, eTriangleMesh, eQuadSlicer, eTriangleMeshSlicer, eConnectedLineSegments, eArc };
    enum ShapeTypeE { eStatic = 0, eKinematic, eDynamic };
    enum MatType { eCoRotational = 0, eNeoHookean, eLinearElastic };

    // Shape/geometry/SDF tables; inputs stay FP32 (they are FP32 in the
    // reference) and are widened where used.
    struct Scene
    {
        int numShapes = 0;
        std::vector<int> shapeIds;
        std::vector<int> shapeType, shapeGeom;
        std::vector<float> pos, rot, invScale, linVel, angVel, comInvMass, inertia0, inertia1, params0; // float4 each
        std::vector<double> force, torque;                                                              // float4 each
        std::vector<int> geomType;
        std::vector<float> geomParams; // float4 each
        std::vector<int> sdfDims;      // int4 each
        std::vector<float> sdfLbcs;    // float4 each
        std::vector<float> sdfData;    // float4 each
    };

    template <class T> inline V3<T> f4v(const float *p) { return V3<T>(T(p[0]), T(p[1]), T(p[2])); }
    inline int f2i(float f)
    {
        int i;
        std::memcpy(&i, &f, 4);
        return i;
    }

    template <class T>
    struct Sdf4
    {
        V3<T> g;
        T d;
    };

    // Trilinear distance of an SDF grid (include/geometry_utils/sdf.h:204-268)
    template <class T> inline T sdf_trilinear(const Scene &sc, const int dim[3], int offset, V3<int> base, V3<T> frac)
    {
        T gv[8];
        for (int dx = 0; dx <= 1; ++dx)
            for (int dy = 0; dy <= 1; ++dy)
                for (int dz = 0; dz <= 1; ++dz)
                {
                    int idx = (dx << 2) | (dy << 1) | dz;
                    int gi = offset + ((base.z + dz) * dim[0] * dim[1]) + ((base.y + dy) * dim[0]) + (base.x + dx);
                    gv[idx] = T(sc.sdfData[4 * gi + 3]);
                }
        T fx = frac.x, fy = frac.y, fz = frac.z;
        T ox = T(1) - fx, oy = T(1) - fy, oz = T(1) - fz;
        T w[8] = {ox * oy * oz, fx * oy * oz, ox * fy * oz, fx * fy * oz, ox * oy * fz, fx * oy * fz, ox * fy * fz, fx * fy * fz};
        T acc = 0;
        for (int i = 0; i < 8; ++i)
            acc += gv[i] * w[i];
        return acc;
    }

    // getGeometrySdf (include/geometry_utils/sdf.h:67-496).  Returns "in spine".
    template <class T>
    inline bool geometry_sdf(const Scene &sc, V3<T> lp, int type, const float *p0, Sdf4<T> &out, V3<T> *tangent)
    {
        bool inSpine = false;
        switch (type)
        {
        case eBox:
        {
            V3<T> pen(std::fabs(lp.x) - T(p0[0]), std::fabs(lp.y) - T(p0[1]), std::fabs(lp.z) - T(p0[2]));
            V3<T> s(lp.x < 0.0 ? T(-1.0) : T(1.0), lp.y < 0.0 ? T(-1.0) : T(1.0), lp.z < 0.0 ? T(-1.0) : T(1.0));
            T g = fmx(pen.x, fmx(pen.y, pen.z));
            V3<T> q(fmx(pen.x, T(0)), fmx(pen.y, T(0)), fmx(pen.z, T(0)));
            T l = norm(q);
            V3<T> dir;
            if (g > 0.0)
                dir = q / l;
            else if (pen.x > pen.y)
                dir = (pen.x > pen.z) ? V3<T>(1, 0, 0) : V3<T>(0, 0, 1);
            else
                dir = (pen.y > pen.z) ? V3<T>(0, 1, 0) : V3<T>(0, 0, 1);
            out.g = s * dir;
            out.d = l + fmn(g, T(0));
            break;
        }
        case eSphere:
        {
            T len = norm(lp);
            out.g = lp / fmx(len, T(FLT_EPSILON));
            out.d = len - T(p0[0]);
            break;
        }
        case ePlane:
            out.g = V3<T>(0, 1, 0);
            out.d = lp.y;
            break;
        case eCapsule:
        {
            const T hcyl = T(p0[1]) * T(2);
            V3<T> pa = lp;
            pa.y -= T(p0[1]);
            T h = -pa.y / hcyl;
            h = fmn(fmx(h, T(0.0)), T(1.0));
            pa.y += hcyl * h;
            T d = norm(pa);
            out.g = pa / fmx(d, T(FLT_EPSILON));
            out.d = d - T(p0[0]);
            break;
        }
        case eArc:
        {
            T lo = T(p0[0]), hi = T(p0[1]);
            T ang = std::atan2(lp.z, lp.x);
            if (ang < lo || ang > hi)
            {
                out.g = V3<T>(0, 1, 0);
                out.d = T(NAN);
                break;
            }
            V3<T> pp(std::cos(ang), T(0), std::sin(ang));
            V3<T> d = lp - pp;
            T dist = norm(d);
            out.g = d / fmx(dist, T(FLT_EPSILON));
            out.d = dist;
            if (tangent)
                *tangent = V3<T>(-pp.z, T(0), pp.x);
            break;
        }
        case eTriangleMesh:
        {
            const int off = f2i(p0[0]);
            const int id = f2i(p0[1]);
            const float *lb = &sc.sdfLbcs[4 * id];
            const int dim[3] = {sc.sdfDims[4 * id], sc.sdfDims[4 * id + 1], sc.sdfDims[4 * id + 2]};
            V3<T> sp = (lp - f4v<T>(lb)) / T(lb[3]);
            V3<int> mq(int(sp.x + T(0.5f)), int(sp.y + T(0.5f)), int(sp.z + T(0.5f)));
            if (mq.x < 0 || mq.x >= dim[0] || mq.y < 0 || mq.y >= dim[1] || mq.z < 0 || mq.z >= dim[2])
            {
                out.g = V3<T>(0, 1, 0);
                out.d = T(3.4028234663852885981170418348452e+38F);
                break;
            }
            int qi = off + mq.z * dim[0] * dim[1] + mq.y * dim[0] + mq.x;
            out.g = f4v<T>(&sc.sdfData[4 * qi]);
            V3<int> base(int(sp.x), int(sp.y), int(sp.z));
            V3<T> fr(sp.x - T(base.x), sp.y - T(base.y), sp.z - T(base.z));
            if (base.x < 0 || base.x >= dim[0] - 1 || base.y < 0 || base.y >= dim[1] - 1 || base.z < 0 || base.z >= dim[2] - 1)
            {
                out.d = T(sc.sdfData[4 * qi + 3]);
                break;
            }
            out.d = sdf_trilinear<T>(sc, dim, off, base, fr);
            break;
        }
        case eQuadSlicer:
        {
            T hx = T(p0[0]), hz = T(p0[1]);
            if (lp.x < -hx || lp.x > hx || lp.z < -hz || lp.z > hz)
            {
                out.g = V3<T>(0, 0, 0);
                out.d = T(3.4028234663852885981170418348452e+38F);
            }
            else
            {
                out.g = V3<T>(0, std::copysign(T(1), lp.y), 0);
                out.d = std::fabs(lp.y);
            }
            break;
        }
        case eTriangleMeshSlicer:
        {
            out.g = V3<T>(0, 1, 0);
            out.d = T(NAN);
            const int off = f2i(p0[0]) + 1;
            const int id = f2i(p0[1]);
            const float *lb = &sc.sdfLbcs[4 * id];
            const int dim[3] = {sc.sdfDims[4 * id], sc.sdfDims[4 * id + 1], sc.sdfDims[4 * id + 2]};
            V3<T> sp = (lp - f4v<T>(lb)) / T(lb[3]);
            V3<int> mq(int(sp.x + T(0.5f)), int(sp.y + T(0.5f)), int(sp.z + T(0.5f)));
            if (mq.x < 1 || mq.x >= dim[0] - 1 || mq.y < 1 || mq.y >= dim[1] - 1 || mq.z < 1 || mq.z >= dim[2] - 1)
                break;
            int qi = off + mq.z * dim[0] * dim[1] + mq.y * dim[0] + mq.x;
            T qd = T(sc.sdfData[4 * qi + 3]);
            if (std::isnan(qd))
                break;
            out.g = f4v<T>(&sc.sdfData[4 * qi]);
            V3<int> base(int(sp.x), int(sp.y), int(sp.z));
            V3<T> fr(sp.x - T(base.x), sp.y - T(base.y), sp.z - T(base.z));
            T iv = sdf_trilinear<T>(sc, dim, off, base, fr);
            if (std::isnan(iv))
                break;
            out.d = std::fabs(iv);
            if (qd * iv < 0)
                out.g = V3<T>(-out.g.x, -out.g.y, -out.g.z);
            const float *sp4 = &sc.sdfData[4 * (off - 1)];
            if (std::isnan(sp4[0]))
                break;
            V3<T> sn = f4v<T>(sp4);
            T len = norm(sn);
            T side = dot(lp, sn / len) + T(sp4[3]);
            if (side > 0)
            {
                out.d -= len - T(1);
                inSpine = true;
            }
            break;
        }
        case eConnectedLineSegments:
        {
            out.g = V3<T>(0, 1, 0);
            out.d = T(NAN);
            const int off = f2i(p0[0]);
            const int id = f2i(p0[1]);
            const int size = f2i(p0[2]);
            V3<T> bmin = f4v<T>(&sc.sdfLbcs[4 * id]);
            V3<T> bmax = f4v<T>(&sc.sdfData[4 * off]);
            if (lp.x < bmin.x || lp.x > bmax.x || lp.y < bmin.y || lp.y > bmax.y || lp.z < bmin.z || lp.z > bmax.z)
                break;
            int closestIdx = 0;
            V3<T> closest;
            T closestD = T(3.4028234663852885981170418348452e+38F);
            V3<T> p0v, p1v;
            for (int p = off + 1; p < off + size - 1; ++p)
            {
                p0v = f4v<T>(&sc.sdfData[4 * p]);
                p1v = f4v<T>(&sc.sdfData[4 * (p + 1)]);
                V3<T> lv = p1v - p0v;
                V3<T> pv = lp - p0v;
                T t = dot(pv, lv) / sqn(lv);
                t = fmx(T(0.0f), fmn(T(1.0f), t));
                V3<T> pr = p0v + lv * t;
                T d = norm(lp - pr);
                if (d < closestD)
                {
                    closestD = d;
                    closest = pr;
                    closestIdx = p;
                }
            }
            p0v = f4v<T>(&sc.sdfData[4 * (off + 1)]);
            if (sqn(closest - p0v) < T(FLT_EPSILON) || sqn(closest - p1v) < T(FLT_EPSILON))
                break;
            V3<T> g = lp - closest;
            normalize(g);
            if (tangent)
            {
                *tangent = f4v<T>(&sc.sdfData[4 * closestIdx]) - f4v<T>(&sc.sdfData[4 * (closestIdx + 1)]);
                normalize(*tangent);
            }
            out.g = g;
            out.d = closestD;
            break;
        }
        default:
            break;
        }
        return inSpine;
    }

    // getGeometryModSdf (include/geometry_utils/sdf.h:504-662). Returns
    // whether the point projects onto the slicer.
    template <class T>
    inline bool geometry_mod_sdf(const Scene &sc, V3<T> lp, int type, const float *p0, Sdf4<T> &out)
    {
        if (type == eQuadSlicer)
        {
            T hx = T(p0[0]), hz = T(p0[1]);
            if (lp.x < -hx || lp.x > hx || lp.z < -hz || lp.z > hz)
                return false;
            out.g = V3<T>(0, std::copysign(T(1), lp.y), 0);
            out.d = lp.y;
            return true;
        }
        if (type == eTriangleMeshSlicer)
        {
            out.g = V3<T>(0, 1, 0);
            out.d = T(NAN);
            int off = f2i(p0[0]);
            const float *sp4 = &sc.sdfData[4 * off];
            if (std::isnan(sp4[0]))
                return false;
            T side = dot(lp, f4v<T>(sp4)) + T(sp4[3]);
            if (side > 0)
                return false;
            off++;
            const int id = f2i(p0[1]);
            const float *lb = &sc.sdfLbcs[4 * id];
            const int dim[3] = {sc.sdfDims[4 * id], sc.sdfDims[4 * id + 1], sc.sdfDims[4 * id + 2]};
            V3<T> sp = (lp - f4v<T>(lb)) / T(lb[3]);
            V3<int> mq(int(sp.x + T(0.5f)), int(sp.y + T(0.5f)), int(sp.z + T(0.5f)));
            if (mq.x < 1 || mq.x >= dim[0] - 1 || mq.y < 1 || mq.y >= dim[1] - 1 || mq.z < 1 || mq.z >= dim[2] - 1)
                return false;
            int qi = off + mq.z * dim[0] * dim[1] + mq.y * dim[0] + mq.x;
            T qd = T(sc.sdfData[4 * qi + 3]);
            if (std::isnan(qd))
                return false;
            out.g = f4v<T>(&sc.sdfData[4 * qi]);
            V3<int> base(int(sp.x), int(sp.y), int(sp.z));
            V3<T> fr(sp.x - T(base.x), sp.y - T(base.y), sp.z - T(base.z));
            T iv = sdf_trilinear<T>(sc, dim, off, base, fr);
            if (std::isnan(iv))
                return false;
            out.d = iv;
            if (qd * iv < 0)
                out.g = V3<T>(-out.g.x, -out.g.y, -out.g.z);
            return true;
        }
        return false;
    }

    // resolveRigidCollision (src/solver/mpm_rigid_contact.h:101-399).
    // vel holds (x, y, z, mass) for grid nodes.  Returns nothing; modifies
    // vel (nodes) or pos (particles); dynamic shapes accumulate force/torque.
    template <class T, bool IsParticle>
    inline void resolve_rigid(Scene &sc, T vel[4], V3<T> &pos, bool effectiveMass)
    {
        for (int si = 0; si < sc.numShapes; ++si)
        {
            const int id = sc.shapeIds[si];
            const int stype = sc.shapeType[id];
            const V3<T> invScale = f4v<T>(&sc.invScale[4 * id]);
            const int gid = sc.shapeGeom[id];
            const int gtype = sc.geomType[gid];
            if (IsParticle && (gtype == eConnectedLineSegments || gtype == eArc))
                continue;
            const V3<T> spos = f4v<T>(&sc.pos[4 * id]);
            const T q[4] = {T(sc.rot[4 * id]), T(sc.rot[4 * id + 1]), T(sc.rot[4 * id + 2]), T(sc.rot[4 * id + 3])};
            V3<T> worldRel = pos - spos;
            V3<T> rel = qrot_inv(q, worldRel);
            V3<T> local = rel * invScale;
            Sdf4<T> sd;
            V3<T> tangent;
            bool onSpine = geometry_sdf(sc, local, gtype, &sc.geomParams[4 * gid], sd, &tangent);
            V3<T> gradient = sd.g;
            T distance = sd.d;
            if (std::isnan(distance))
                continue;
            gradient = gradient * invScale;
            const T is = rsq(sqn(gradient));
            distance = distance * is;
            gradient = gradient * is;
            const float *sp0 = &sc.params0[4 * id];
            distance -= T(sp0[3]);
            T smooth = IsParticle ? T(0) : T(sp0[0]);
            if (!(distance < smooth))
                continue;
            V3<T> normal = qrot(q, gradient);
            V3<T> rigidVel = (stype == eStatic) ? V3<T>() : (f4v<T>(&sc.linVel[4 * id]) + cross(f4v<T>(&sc.angVel[4 * id]), worldRel));
            V3<T> relVel = V3<T>(vel[0], vel[1], vel[2]) - rigidVel;
            T vn = dot(relVel, normal);
            V3<T> nComp = normal * vn;
            if (IsParticle)
            {
                T inside = fmn(distance, T(0));
                V3<T> corr = normal * inside;
                pos.x -= corr.x;
                pos.y -= corr.y;
                pos.z -= corr.z;
                continue;
            }
            if (vn >= T(0))
                continue;
            T friction = onSpine ? T(1e2f) : T(sp0[2]);
            const T sticky = T(sp0[1]);
            V3<T> response;
            if (effectiveMass && stype == eDynamic)
            {
                // Two-body impulse (src/solver/mpm_rigid_contact.h:234-336)
                const T gridMass = vel[3];
                const float *cim = &sc.comInvMass[4 * id];
                const T invMassRigid = T(cim[3]);
                const T rigidMass = invMassRigid > T(0) ? T(1) / invMassRigid : T(0);
                // collisionResponse is zero-initialised in the reference: a
                // dynamic shape without mass zeroes the node velocity.
                response = V3<T>();
                if (!(rigidMass > T(0)))
                {
                    vel[0] = response.x;
                    vel[1] = response.y;
                    vel[2] = response.z;
                    continue;
                }
                const V3<T> r = worldRel - qrot(q, f4v<T>(cim));
                V3<T> rxn = cross(r, normal);
                V3<T> rxnl = qrot_inv(q, rxn);
                const float *i0 = &sc.inertia0[4 * id];
                const float *i1 = &sc.inertia1[4 * id];
                V3<T> wl(T(i0[0]) * rxnl.x + T(i1[0]) * rxnl.y + T(i1[1]) * rxnl.z,
                         T(i1[0]) * rxnl.x + T(i0[1]) * rxnl.y + T(i1[2]) * rxnl.z,
                         T(i1[1]) * rxnl.x + T(i1[2]) * rxnl.y + T(i0[2]) * rxnl.z);
                V3<T> ww = qrot(q, wl);
                V3<T> t = cross(ww, r);
                T invMassRigidN = invMassRigid + dot(normal, t);
                T invMassGrid = T(1) / gridMass;
                T invMassComb = invMassGrid + invMassRigidN;
                if (!(invMassComb > T(0)))
                {
                    vel[0] = response.x;
                    vel[1] = response.y;
                    vel[2] = response.z;
                    continue;
                }
                T Jn = T(-1) * vn / invMassComb;
                V3<T> impN = normal * Jn;
                V3<T> vto = relVel - (normal * vn);
                T vtn = norm(vto);
                V3<T> impT(0, 0, 0);
                if (vtn > T(FLT_EPSILON))
                {
                    T stopJt = vtn / invMassComb;
                    T maxJt = friction * Jn;
                    T fJt = fmn(maxJt, stopJt);
                    T left = vtn - (fJt * invMassComb);
                    T fin = left * sticky;
                    T totJt = (vtn - fin) / invMassComb;
                    impT = (vto * T(-1) / vtn) * totJt;
                }
                V3<T> totImp = impN + impT;
                response = V3<T>(vel[0], vel[1], vel[2]) + (totImp * invMassGrid);
                V3<T> mom = totImp * T(-1);
                V3<T> tq = cross(r, mom);
                for (int k = 0; k < 3; ++k)
                {
                    sc.force[4 * id + k] = double(T(sc.force[4 * id + k]) + mom[k]);
                    sc.torque[4 * id + k] = double(T(sc.torque[4 * id + k]) + tq[k]);
                }
            }
            else
            {
                V3<T> tang = relVel - nComp;
                if (gtype == eConnectedLineSegments || gtype == eArc)
                {
                    tangent = tangent * invScale;
                    normalize(tangent);
                    tangent = qrot(q, tangent);
                    tang = tangent * dot(tang, tangent);
                }
                T tn = norm(tang);
                T fc = fmx(friction * vn / (tn + T(FLT_EPSILON)), T(-1));
                response = tang + tang * fc;
                response = response * sticky;
                response = response + rigidVel;
                if (!effectiveMass && stype == eDynamic)
                {
                    V3<T> mom = V3<T>(vel[0], vel[1], vel[2]) - response;
                    mom = mom * vel[3];
                    V3<T> tq = cross(pos - spos, mom);
                    for (int k = 0; k < 3; ++k)
                    {
                        sc.force[4 * id + k] = double(T(sc.force[4 * id + k]) + mom[k]);
                        sc.torque[4 * id + k] = double(T(sc.torque[4 * id + k]) + tq[k]);
                    }
                }
            }
            vel[0] = response.x;
            vel[1] = response.y;
            vel[2] = response.z;
        }
    }

    // checkOnDifferentSides (src/solver/mpm_rigid_contact.h:401-478).  The
    // returned normal stays in the shape-local frame.
    template <class T>
    inline bool different_sides(const Scene &sc, V3<T> p1, V3<T> p2, V3<T> *normal)
    {
        for (int si = 0; si < sc.numShapes; ++si)
        {
            const int id = sc.shapeIds[si];
            const int gid = sc.shapeGeom[id];
            const int gtype = sc.geomType[gid];
            if (gtype != eQuadSlicer && gtype != eTriangleMeshSlicer)
                continue;
            const V3<T> invScale = f4v<T>(&sc.invScale[4 * id]);
            const V3<T> spos = f4v<T>(&sc.pos[4 * id]);
            const T q[4] = {T(sc.rot[4 * id]), T(sc.rot[4 * id + 1]), T(sc.rot[4 * id + 2]), T(sc.rot[4 * id + 3])};
            V3<T> r1 = qrot_inv(q, p1 - spos) * invScale;
            V3<T> r2 = qrot_inv(q, p2 - spos) * invScale;
            Sdf4<T> s1, s2;
            if (!geometry_mod_sdf(sc, r1, gtype, &sc.geomParams[4 * gid], s1))
                continue;
            bool on2 = geometry_mod_sdf(sc, r2, gtype, &sc.geomParams[4 * gid], s2);
            if (on2 && !(std::signbit(s1.d) == std::signbit(s2.d)))
            {
                if (normal)
                {
                    V3<T> g = s2.g * invScale;
                    const T is = rsq(sqn(g));
                    *normal = g * is;
                }
                return true;
            }
        }
        return false;
    }

    // Quadratic B-spline weights and grid-unit gradients (src/solver/shape_functions.h:44-75)
    template <class T> inline void bspline_w(V3<T> fx, T w[3][3])
    {
        for (int a = 0; a < 3; ++a)
        {
            T f = fx[a];
            w[0][a] = T(0.5f) * sq(T(1.5f) - f);
            w[1][a] = T(0.75f) - sq(f - T(1.0f));
            w[2][a] = T(0.5f) * sq(f - T(0.5f));
        }
    }
    template <class T> inline void bspline_dw(V3<T> fx, T dw[3][3])
    {
        for (int a = 0; a < 3; ++a)
        {
            T f = fx[a];
            dw[0][a] = T(double(f) - 1.5); // the reference subtracts a double literal
            dw[1][a] = T(2.0f) * (T(1.0f) - f);
            dw[2][a] = f - T(0.5f);
        }
    }

    // computeStress (src/solver/constitutive_models.h:95-118).  eCoRotational
    // falls through to the linear-elastic (StVK) branch.
    template <class T> inline M3<T> compute_stress(const float *p0, int type, const M3<T> &F)
    {
        const T lambda = T(p0[0]), mu = T(p0[1]);
        if (type == eNeoHookean)
        {
            T J = det(F);
            if (J < T(1e-6))
                J = T(1e-6f);
            M3<T> b = mul(F, transpose(F));
            T lj = T(std::log(J));
            return (b - M3<T>::identity()) * mu + M3<T>::scalar(lambda * lj);
        }
        if (type == eCoRotational || type == eLinearElastic)
        {
            M3<T> strain = (mul(transpose(F), F) - M3<T>::identity()) * T(0.5f);
            M3<T> S = M3<T>::scalar(trace(strain)) * lambda + strain * (T(2) * mu);
            return mul(mul(F, S), transpose(F));
        }
        return M3<T>();
    }

    // solveMaterialConstraints (src/solver/constitutive_models.h:123-156)
    template <class T> inline void pb_constraint(T dt, const float *p0, int type, const M3<T> &F, M3<T> &C)
    {
        if (type != eNeoHookean && type != eCoRotational)
            return;
        M3<T> targetF = mul(M3<T>::identity() + C * dt, F);
        M3<T> u, vt;
        V3<T> sig;
        svd3(targetF, u, vt, sig);
        T detF = det(targetF);
        T cl = fmn(fmx(std::fabs(detF), T(0.1f)), T(1000.0f));
        M3<T> Q = targetF * (T(1.0f) / (std::copysign(T(1.0f), detF) * std::sqrt(cl)));
        const T r = T(p0[1]);
        targetF = mul(u, vt) * r + Q * (T(1.0f) - r);
        const T alpha = T(p0[0]);
        C = C + ((mul(targetF, inverse_t(F)) - M3<T>::identity()) - C * dt) * alpha;
    }

    // ------------------------------------------------------------------
    // Solver state and drivers (src/solver/cpu_{mls,pb,standard}_mpm_solver.cpp)
    // ------------------------------------------------------------------
    enum Kind { kMls = 0, kPb = 1, kStd = 2 };

    template <class T>
    struct Solver
    {
        int kind = kMls;
        int maxParticles = 0;
        T dx, invdx, cellVol;
        V3<T> bmin, bmax;
        int nx, ny, nz, numNodes;
        V3<T> gravity;
        T dt = T(0.002f);
        int iterations = 10;
        bool effectiveMass = false;
        unsigned numActive = 0;
        // particle state
        std::vector<T> xm;     // 4 per particle (x, y, z, mass)
        std::vector<T> v;      // 3
        std::vector<T> F, C;   // 9 each, column-major
        std::vector<T> V0;     // 1
        std::vector<float> params; // 4
        std::vector<int> type;
        std::vector<uint8_t> mask;
        // grid
        std::vector<T> grid;  // 4 per node
        std::vector<T> force; // 3 per node (standard MPM)
        Scene scene;

        Solver(int k, int maxP, float cellSize, const float lo[3], const float hi[3])
        {
            kind = k;
            maxParticles = maxP;
            // Ctor arithmetic is FP32 in the reference (src/solver/cpu_mls_mpm_solver.cpp:39-50)
            float fdx = cellSize;
            float finv = 1 / cellSize;
            int n[3];
            for (int a = 0; a < 3; ++a)
                n[a] = int((hi[a] - lo[a]) / cellSize + 1.0f);
            nx = n[0];
            ny = n[1];
            nz = n[2];
            numNodes = nx * ny * nz;
            dx = T(fdx);
            invdx = T(finv);
            cellVol = T(fdx * fdx * fdx);
            bmin = V3<T>(T(lo[0]), T(lo[1]), T(lo[2]));
            bmax = V3<T>(T(hi[0]), T(hi[1]), T(hi[2]));
            xm.assign(4 * maxP, T(0));
            v.assign(3 * maxP, T(0));
            F.assign(9 * maxP, T(0));
            C.assign(9 * maxP, T(0));
            V0.assign(maxP, T(0));
            params.assign(4 * maxP, 0.f);
            type.assign(maxP, 0);
            mask.assign(maxP, 0);
            grid.assign(4 * size_t(numNodes), T(0));
            force.assign(3 * size_t(numNodes), T(0));
        }

        M3<T> getM(const std::vector<T> &a, int p) const
        {
            M3<T> m;
            for (int j = 0; j < 3; ++j)
                m.c[j] = V3<T>(a[9 * p + 3 * j], a[9 * p + 3 * j + 1], a[9 * p + 3 * j + 2]);
            return m;
        }
        void setM(std::vector<T> &a, int p, const M3<T> &m)
        {
            for (int j = 0; j < 3; ++j)
            {
                a[9 * p + 3 * j] = m.c[j].x;
                a[9 * p + 3 * j + 1] = m.c[j].y;
                a[9 * p + 3 * j + 2] = m.c[j].z;
            }
        }
        V3<T> getX(int p) const { return V3<T>(xm[4 * p], xm[4 * p + 1], xm[4 * p + 2]); }
        void setX(int p, V3<T> x)
        {
            xm[4 * p] = x.x;
            xm[4 * p + 1] = x.y;
            xm[4 * p + 2] = x.z;
        }
        V3<T> getV(int p) const { return V3<T>(v[3 * p], v[3 * p + 1], v[3 * p + 2]); }
        void setV(int p, V3<T> a)
        {
            v[3 * p] = a.x;
            v[3 * p + 1] = a.y;
            v[3 * p + 2] = a.z;
        }

        // Base cell (truncation toward zero) and fractional position
        // (src/solver/mls_mpm_algorithm.h:172-174)
        void cellOf(V3<T> x, V3<int> &base, V3<T> &fx) const
        {
            V3<T> s = (x - bmin) * invdx;
            base = V3<int>(int(s.x - T(0.5f)), int(s.y - T(0.5f)), int(s.z - T(0.5f)));
            fx = V3<T>(s.x - T(base.x), s.y - T(base.y), s.z - T(base.z));
        }
        bool nodeInRange(V3<int> n) const
        {
            return !(n.x < 0 || n.x >= nx || n.y < 0 || n.y >= ny || n.z < 0 || n.z >= nz);
        }
        int nodeIndex(V3<int> n) const { return n.x + n.y * nx + n.z * nx * ny; }

        // computeInitialData (src/solver/cpu_mls_mpm_solver.cpp:62-95; the
        // algorithm is src/solver/mls_mpm_algorithm.h:46-145).  The grid is not
        // cleared first.
        void computeInitialData(unsigned n, const unsigned *idx)
        {
            for (unsigned i = 0; i < n; ++i)
            {
                int p = int(idx[i]);
                setM(F, p, M3<T>::identity());
                V3<int> base;
                V3<T> fx;
                cellOf(getX(p), base, fx);
                T w[3][3];
                bspline_w(fx, w);
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int c = 0; c < 3; ++c)
                        {
                            V3<int> nc(base.x + a, base.y + b, base.z + c);
                            if (!nodeInRange(nc))
                                continue;
                            T wt = w[a][0] * w[b][1] * w[c][2];
                            grid[4 * size_t(nodeIndex(nc)) + 3] += wt * xm[4 * p + 3];
                        }
            }
            for (unsigned i = 0; i < n; ++i)
            {
                int p = int(idx[i]);
                V3<int> base;
                V3<T> fx;
                cellOf(getX(p), base, fx);
                T w[3][3];
                bspline_w(fx, w);
                T density = 0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int c = 0; c < 3; ++c)
                        {
                            V3<int> nc(base.x + a, base.y + b, base.z + c);
                            if (!nodeInRange(nc))
                                continue;
                            T wt = w[a][0] * w[b][1] * w[c][2];
                            T nd = grid[4 * size_t(nodeIndex(nc)) + 3] / cellVol;
                            density += nd * wt;
                        }
                V0[p] = (density > T(1e-8f)) ? xm[4 * p + 3] / density : T(1e-8f);
            }
        }

        void resetGrid()
        {
            std::fill(grid.begin(), grid.end(), T(0));
            if (kind == kStd)
                std::fill(force.begin(), force.end(), T(0));
        }

        // mlsMpmParticleToGrid (src/solver/mls_mpm_algorithm.h:147-241)
        void mlsP2G(int p)
        {
            V3<int> base;
            V3<T> fx;
            const V3<T> x = getX(p);
            cellOf(x, base, fx);
            T w[3][3];
            bspline_w(fx, w);
            V3<T> vel = getV(p) + gravity * dt;
            setV(p, vel);
            const M3<T> Fp = getM(F, p);
            const M3<T> Cp = getM(C, p);
            const M3<T> stress = compute_stress(&params[4 * p], type[p], Fp);
            const T s = -dt * T(4) * invdx * invdx * V0[p];
            const T m = xm[4 * p + 3];
            const M3<T> affine = stress * s + Cp * m;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                    {
                        V3<int> nc(base.x + a, base.y + b, base.z + c);
                        if (!nodeInRange(nc))
                            continue;
                        V3<T> dist = (V3<T>(T(a), T(b), T(c)) - fx) * dx;
                        if (different_sides<T>(scene, x, x + dist, nullptr))
                            continue;
                        const size_t ni = size_t(nodeIndex(nc));
                        const T wt = w[a][0] * w[b][1] * w[c][2];
                        grid[4 * ni + 3] += wt * m;
                        V3<T> contrib = (vel * m + mul(affine, dist)) * wt;
                        grid[4 * ni] += contrib.x;
                        grid[4 * ni + 1] += contrib.y;
                        grid[4 * ni + 2] += contrib.z;
                    }
        }

        // pbMpmParticleToGrid (src/solver/pb_mpm_algorithm.h:150-231); the
        // CPU driver stores the constrained C (src/solver/cpu_pb_mpm_solver.cpp:123-140).
        void pbP2G(int p)
        {
            M3<T> Cp = getM(C, p);
            const M3<T> Fp = getM(F, p);
            pb_constraint(dt, &params[4 * p], type[p], Fp, Cp);
            setM(C, p, Cp);
            V3<int> base;
            V3<T> fx;
            const V3<T> x = getX(p);
            cellOf(x, base, fx);
            T w[3][3];
            bspline_w(fx, w);
            const T m = xm[4 * p + 3];
            const V3<T> vel = getV(p);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                    {
                        V3<int> nc(base.x + a, base.y + b, base.z + c);
                        if (!nodeInRange(nc))
                            continue;
                        V3<T> dist = (V3<T>(T(a), T(b), T(c)) - fx) * dx;
                        if (different_sides<T>(scene, x, x + dist, nullptr))
                            continue;
                        const size_t ni = size_t(nodeIndex(nc));
                        const T wt = w[a][0] * w[b][1] * w[c][2];
                        const T wm = wt * m;
                        grid[4 * ni + 3] += wm;
                        V3<T> contrib = (vel + mul(Cp, dist)) * wm;
                        grid[4 * ni] += contrib.x;
                        grid[4 * ni + 1] += contrib.y;
                        grid[4 * ni + 2] += contrib.z;
                    }
        }

        // standardMpmParticleToGrid (src/solver/standard_mpm_algorithm.h:146-215)
        void stdP2G(int p)
        {
            const T m = xm[4 * p + 3];
            V3<int> base;
            V3<T> fx;
            cellOf(getX(p), base, fx);
            T w[3][3], dw[3][3];
            bspline_w(fx, w);
            bspline_dw(fx, dw);
            const M3<T> stress = compute_stress(&params[4 * p], type[p], getM(F, p));
            const V3<T> vel = getV(p);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                    {
                        V3<int> nc(base.x + a, base.y + b, base.z + c);
                        if (!nodeInRange(nc))
                            continue;
                        const size_t ni = size_t(nodeIndex(nc));
                        const T wt = w[a][0] * w[b][1] * w[c][2];
                        const V3<T> wg(dw[a][0] * w[b][1] * w[c][2], w[a][0] * dw[b][1] * w[c][2], w[a][0] * w[b][1] * dw[c][2]);
                        const V3<T> ext = gravity * (wt * m);
                        const V3<T> inr = mul(stress, wg) * (-V0[p]);
                        grid[4 * ni + 3] += wt * m;
                        V3<T> mom = vel * (wt * m);
                        grid[4 * ni] += mom.x;
                        grid[4 * ni + 1] += mom.y;
                        grid[4 * ni + 2] += mom.z;
                        V3<T> f = ext + inr;
                        force[3 * ni] += f.x;
                        force[3 * ni + 1] += f.y;
                        force[3 * ni + 2] += f.z;
                    }
        }

        // mlsMpmUpdateGrid / pbMpmUpdateGrid / standardMpmUpdateGrid
        // (src/solver/mls_mpm_algorithm.h:243-294, pb 233-284, standard 217-272)
        void updateNode(int idx)
        {
            const int i = idx % nx;
            const int j = (idx / nx) % ny;
            const int k = idx / (nx * ny);
            T *g = &grid[4 * size_t(idx)];
            const T cutoff = (kind == kPb) ? T(1e-8f) : T(1e-4f);
            if (g[3] < cutoff)
                return;
            if (kind == kStd)
            {
                const T *f = &force[3 * size_t(idx)];
                g[0] += f[0] * dt;
                g[1] += f[1] * dt;
                g[2] += f[2] * dt;
            }
            g[0] /= g[3];
            g[1] /= g[3];
            g[2] /= g[3];
            V3<T> np = V3<T>(T(i) * dx, T(j) * dx, T(k) * dx) + bmin;
            np = np + V3<T>(g[0], g[1], g[2]) * dt;
            resolve_rigid<T, false>(scene, g, np, effectiveMass);
            if (i < 2 || i > nx - 3)
                g[0] = 0;
            if (j < 2 || j > ny - 3)
                g[1] = 0;
            if (k < 2 || k > nz - 3)
                g[2] = 0;
        }

        // G2P gather shared by MLS and PB (src/solver/mls_mpm_algorithm.h:318-371,
        // src/solver/pb_mpm_algorithm.h:307-360): returns PIC v and C.
        void gatherVC(int p, V3<T> &vOut, M3<T> &cOut) const
        {
            V3<int> base;
            V3<T> fx;
            const V3<T> x = getX(p);
            cellOf(x, base, fx);
            T w[3][3];
            bspline_w(fx, w);
            const V3<T> vp = getV(p);
            V3<T> vacc;
            M3<T> aacc;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                    {
                        V3<int> nc(base.x + a, base.y + b, base.z + c);
                        if (!nodeInRange(nc))
                            continue;
                        const size_t ni = size_t(nodeIndex(nc));
                        V3<T> nv(grid[4 * ni], grid[4 * ni + 1], grid[4 * ni + 2]);
                        V3<T> dist = (V3<T>(T(a), T(b), T(c)) - fx) * dx;
                        V3<T> n;
                        if (different_sides<T>(scene, x, x + dist, &n))
                            nv = vp - n * dot(vp, n) * T(2.0f);
                        const T wt = w[a][0] * w[b][1] * w[c][2];
                        vacc = vacc + nv * wt;
                        aacc = aacc + outer_ref(nv * wt, dist);
                    }
            vOut = vacc;
            cOut = aacc * (T(4) * invdx);
        }

        // Clamp to grid bounds (include/math/math_helpers.h:246-251)
        V3<T> clampBounds(V3<T> x) const
        {
            return V3<T>(fmn(fmx(x.x, bmin.x), bmax.x), fmn(fmx(x.y, bmin.y), bmax.y), fmn(fmx(x.z, bmin.z), bmax.z));
        }

        // mlsMpmGridToParticle (src/solver/mls_mpm_algorithm.h:296-390)
        void mlsG2P(int p)
        {
            V3<T> vel;
            M3<T> Cp;
            gatherVC(p, vel, Cp);
            setV(p, vel);
            setM(C, p, Cp);
            V3<T> x = getX(p) + vel * dt;
            x = clampBounds(x);
            T vv[4] = {vel.x, vel.y, vel.z, T(0)};
            resolve_rigid<T, true>(scene, vv, x, false);
            setX(p, x);
            M3<T> Fp = getM(F, p);
            setM(F, p, mul(M3<T>::identity() + Cp * dt, Fp));
        }

        // pbMpmGridToParticle (src/solver/pb_mpm_algorithm.h:286-361)
        void pbG2P(int p)
        {
            V3<T> vel;
            M3<T> Cp;
            gatherVC(p, vel, Cp);
            setV(p, vel);
            setM(C, p, Cp);
        }

        // pbMpmIntegrateParticle (src/solver/pb_mpm_algorithm.h:363-408)
        void pbIntegrate(int p)
        {
            V3<T> vel = getV(p);
            V3<T> x = getX(p) + vel * dt;
            x = clampBounds(x);
            vel = vel + gravity * dt;
            T vv[4] = {vel.x, vel.y, vel.z, T(0)};
            resolve_rigid<T, true>(scene, vv, x, false);
            setX(p, x);
            setV(p, vel);
            setM(F, p, mul(M3<T>::identity() + getM(C, p) * dt, getM(F, p)));
        }

        // standardMpmGridToParticle (src/solver/standard_mpm_algorithm.h:274-352)
        void stdG2P(int p)
        {
            V3<int> base;
            V3<T> fx;
            cellOf(getX(p), base, fx);
            T w[3][3], dw[3][3];
            bspline_w(fx, w);
            bspline_dw(fx, dw);
            V3<T> vacc;
            M3<T> gv;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    for (int c = 0; c < 3; ++c)
                    {
                        V3<int> nc(base.x + a, base.y + b, base.z + c);
                        if (!nodeInRange(nc))
                            continue;
                        const size_t ni = size_t(nodeIndex(nc));
                        const T wt = w[a][0] * w[b][1] * w[c][2];
                        V3<T> nv(grid[4 * ni], grid[4 * ni + 1], grid[4 * ni + 2]);
                        vacc = vacc + nv * wt;
                        const V3<T> wg(dw[a][0] * w[b][1] * w[c][2], w[a][0] * dw[b][1] * w[c][2], w[a][0] * w[b][1] * dw[c][2]);
                        gv = gv + outer_ref(wg, nv);
                    }
            setV(p, vacc);
            V3<T> x = getX(p) + vacc * dt;
            x = clampBounds(x);
            T vv[4] = {vacc.x, vacc.y, vacc.z, T(0)};
            resolve_rigid<T, true>(scene, vv, x, false);
            setX(p, x);
            setM(F, p, mul(M3<T>::identity() + gv * dt, getM(F, p)));
        }

        void updateGridAll()
        {
            for (int n = 0; n < numNodes; ++n)
                updateNode(n);
        }

        // step() of the three CPU solvers.  The standard P2G ignores the mask
        // (src/solver/cpu_standard_mpm_solver.cpp:101-120).
        void step()
        {
            const int na = int(numActive);
            if (kind == kMls)
            {
                resetGrid();
                for (int p = 0; p < na; ++p)
                    if (mask[p])
                        mlsP2G(p);
                updateGridAll();
                for (int p = 0; p < na; ++p)
                    if (mask[p])
                        mlsG2P(p);
            }
            else if (kind == kPb)
            {
                for (int it = 0; it < iterations; ++it)
                {
                    resetGrid();
                    for (int p = 0; p < na; ++p)
                        if (mask[p])
                            pbP2G(p);
                    updateGridAll();
                    for (int p = 0; p < na; ++p)
                        if (mask[p])
                            pbG2P(p);
                }
                for (int p = 0; p < na; ++p)
                    if (mask[p])
                        pbIntegrate(p);
            }
            else
            {
                resetGrid();
                for (int p = 0; p < na; ++p)
                    stdP2G(p);
                updateGridAll();
                for (int p = 0; p < na; ++p)
                    if (mask[p])
                        stdG2P(p);
            }
        }
    };
    // ------------------------------------------------------------------
    // ParticleObjectImpl::initialize (src/engine/particle_object_impl.cpp:36-88):
    // lattice over the geometry bounds fattened by two spacings
    // (Bounds3::fattenFast, include/math/bounds3.h:85-94), the points whose
    // SDF is <= 0 (GeometryImpl::queryPointSdf, src/engine/geometry_impl.cpp:37-45),
    // mapped to the world, in loop order.  FP32 only, as in the engine.
    // ------------------------------------------------------------------
    inline unsigned particle_lattice(const Scene &sc, int type, const float *p0, const float *bmin, const float *bmax,
                                     float spacing, const float *pos, const float *rot, const float *invScale, float mass,
                                     std::vector<float> &out)
    {
        const float fat = 2 * spacing;
        float lo[3], hi[3];
        int dim[3];
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = bmin[k] - fat;
            hi[k] = bmax[k] + fat;
            // ((maximum - minimum) / particleSpacing / invScale).cast<int>()
            dim[k] = static_cast<int>((hi[k] - lo[k]) / spacing / invScale[k]);
        }
        unsigned n = 0;
        for (int ix = 0; ix < dim[0]; ++ix)
            for (int iy = 0; iy < dim[1]; ++iy)
                for (int iz = 0; iz < dim[2]; ++iz)
                {
                    // Vec3f(ix, iy, iz) * particleSpacing * invScale + minimum
                    const V3<float> q(float(ix) * spacing * invScale[0] + lo[0], float(iy) * spacing * invScale[1] + lo[1],
                                      float(iz) * spacing * invScale[2] + lo[2]);
                    Sdf4<float> sd;
                    geometry_sdf(sc, q, type, p0, sd, static_cast<V3<float> *>(nullptr));
                    if (sd.d <= 0)
                    {
                        V3<float> w(q.x / invScale[0], q.y / invScale[1], q.z / invScale[2]);
                        w = qrot(rot, w);
                        w = V3<float>(w.x + pos[0], w.y + pos[1], w.z + pos[2]);
                        out.push_back(w.x);
                        out.push_back(w.y);
                        out.push_back(w.z);
                        out.push_back(mass);
                        ++n;
                    }
                }
        return n;
    }

    // ------------------------------------------------------------------
    // Triangle mesh -> SDF table, as the engine builds it for a triangle-mesh
    // geometry (src/engine/simulation_factory_impl.cpp:628-635).  FP32 only.
    // ------------------------------------------------------------------

    // Closest point of triangle abc to p, region by region (Ericson, Real-Time
    // Collision Detection 5.1.5), as src/geometry_utils/aabb_tree.cpp:336-439
    // evaluates it; returns the squared distance.
    // Regions of the closest point: vertices a, b, c, edges ab, bc, ac, face.
    enum : int { kRegA = 0, kRegB, kRegC, kRegAB, kRegBC, kRegAC, kRegFace };

    inline float tri_closest(V3<float> p, V3<float> a, V3<float> b, V3<float> c, V3<float> &cp, int *region = nullptr)
    {
        const V3<float> ab = b - a, ac = c - a, bc = c - b, ap = p - a, bp = p - b, cpv = p - c;
        const float d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp), d5 = dot(ab, cpv),
                    d6 = dot(ac, cpv);
        const float unom = d4 - d3, udenom = d5 - d6;
        int r;
        if (d1 <= 0.0f && d2 <= 0.0f)
            cp = a, r = kRegA;
        else if (d3 >= 0.0f && d3 >= d4)
            cp = b, r = kRegB;
        else if (d6 >= 0.0f && d6 >= d5)
            cp = c, r = kRegC;
        else
        {
            const float vc = d1 * d4 - d3 * d2;
            const float va = d3 * d6 - d5 * d4;
            const float vb = d5 * d2 - d1 * d6;
            if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f)
                cp = a + ab * (d1 / (d1 - d3)), r = kRegAB;
            else if (va <= 0.0f && d4 >= d3 && d5 >= d6)
                cp = b + bc * (unom / (unom + udenom)), r = kRegBC;
            else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f)
                cp = a + ac * (d2 / (d2 - d6)), r = kRegAC;
            else
            {
                const float den = 1.0f / (va + vb + vc);
                cp = a + ab * (vb * den) + ac * (vc * den), r = kRegFace;
            }
        }
        if (region)
            *region = r;
        const V3<float> v = p - cp;
        return dot(v, v);
    }

    // The same in a unit box: everything shifted by the componentwise
    // min(max(a, b), min(c, q)) and scaled by 1 / max(1e-12, largest extent)
    // (aabb_tree.cpp:443-466; the shift is not the box minimum, kept as written).
    inline float tri_closest_unit(V3<float> q, V3<float> a, V3<float> b, V3<float> c, V3<float> &cp,
                                  int *region = nullptr)
    {
        auto vmin = [](V3<float> u, V3<float> v) {
            return V3<float>(std::fmin(u.x, v.x), std::fmin(u.y, v.y), std::fmin(u.z, v.z));
        };
        auto vmax = [](V3<float> u, V3<float> v) {
            return V3<float>(std::fmax(u.x, v.x), std::fmax(u.y, v.y), std::fmax(u.z, v.z));
        };
        const V3<float> lo = vmin(vmax(a, b), vmin(c, q));
        const V3<float> size = vmax(vmax(a, b), vmax(c, q)) - lo;
        const float inv = std::fmax(1e-12f, std::fmax(size.x, std::fmax(size.y, size.z)));
        const float s = 1.0f / inv;
        V3<float> u;
        const float r = tri_closest((q - lo) * s, (a - lo) * s, (b - lo) * s, (c - lo) * s, u, region);
        cp = u * inv + lo;
        return r * inv * inv;
    }

    // One triangle's term of the exact winding number: its solid angle over
    // 4 pi (evaluateExact, include/geometry_utils/winding_number.h:79-93,
    // the Van Oosterom-Strackee formula), expression order kept.
    inline float winding_term(V3<float> a, V3<float> b, V3<float> c, V3<float> p)
    {
        const float k = float(0.5 / 3.141592653589793238462643383);
        a = a - p;
        b = b - p;
        c = c - p;
        const float la = norm(a), lb = norm(b), lc = norm(c);
        const float y = a.x * b.y * c.z - a.x * b.z * c.y - a.y * b.x * c.z + a.y * b.z * c.x + a.z * b.x * c.y -
                        a.z * b.y * c.x;
        const float x = (la * lb * lc + (a.x * b.x + a.y * b.y + a.z * b.z) * lc + (b.x * c.x + b.y * c.y + b.z * c.z) * la +
                         (c.x * a.x + c.y * a.y + c.z * a.z) * lb);
        return k * std::atan2(y, x);
    }

    // SdfBuilder::precomputeSdfDesc (src/geometry_utils/sdf_builder.cpp:84-101):
    // the AABB tree's root box -- the union of the triangles' boxes, each grown
    // by 1e-4 (buildTree, winding_number.h:148-161) -- grown by fatten;
    // dims = int(size / cellSize + 1).
    inline int mesh_sdf_desc(const float *pts, const unsigned *tris, int numTris, float cell, float fatten, int *dims,
                             float *lower)
    {
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int t = 0; t < 3 * numTris; ++t)
            for (int k = 0; k < 3; ++k)
            {
                lo[k] = std::fmin(lo[k], pts[3 * tris[t] + k] - 1e-4f);
                hi[k] = std::fmax(hi[k], pts[3 * tris[t] + k] + 1e-4f);
            }
        long count = 1;
        for (int k = 0; k < 3; ++k)
        {
            lower[k] = lo[k] - fatten;
            const float size = (hi[k] + fatten) - lower[k];
            dims[k] = static_cast<int>(size / cell + 1.0f);
            count *= dims[k];
        }
        return static_cast<int>(count);
    }

    // SdfBuilder::computeSdf + computeSdfGradient (sdf_builder.cpp:103-160,
    // 242-330).  Node (x, y, z) sits at lower + (x, y, z) * cellSize, row
    // z * dx * dy + y * dx + x.  The closest triangle is found by exhaustive
    // search in index order with a strict <, where the reference walks its AABB
    // tree: the minimum is the same, and where several triangles share the
    // closest point (an edge or vertex) the one picked may differ, with the
    // distance equal up to rounding.  The sign is -1 where the exact winding
    // number (computeWindingNumberExact, winding_number.h:136-146) exceeds 0.5;
    // the reference thresholds its far-field approximation of the same number
    // (beta = 2, sdf_builder.cpp:150), which differs from it only by the
    // approximation error, i.e. at nodes on the surface itself.
    // computeSdfGradient (sdf_builder.cpp:242-330) over the .w column of grid4.
    inline void sdf_gradient(const int *dims, float cell, float *grid4);

    inline void mesh_sdf(const float *pts, const unsigned *tris, int numTris, float cell, const int *dims,
                         const float *lower, float *grid4)
    {
        const int dx = dims[0], dy = dims[1], dz = dims[2];
        auto P = [&](unsigned i) { return V3<float>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); };
        for (int z = 0; z < dz; ++z)
            for (int y = 0; y < dy; ++y)
                for (int x = 0; x < dx; ++x)
                {
                    const V3<float> q(lower[0] + float(x) * cell, lower[1] + float(y) * cell, lower[2] + float(z) * cell);
                    float best = FLT_MAX, wn = 0.0f;
                    V3<float> bestP;
                    for (int t = 0; t < numTris; ++t)
                    {
                        const V3<float> a = P(tris[3 * t]), b = P(tris[3 * t + 1]), c = P(tris[3 * t + 2]);
                        V3<float> cp;
                        const float d2 = tri_closest_unit(q, a, b, c, cp);
                        if (d2 < best)
                        {
                            best = d2;
                            bestP = cp;
                        }
                        wn += winding_term(a, b, c, q);
                    }
                    const float dist = norm(bestP - q);
                    grid4[4 * ((z * dy + y) * dx + x) + 3] = wn > 0.5f ? -dist : dist;
                }
        sdf_gradient(dims, cell, grid4);
    }

    // The mesh-slicer table (src/engine/simulation_factory_impl.cpp:670-690):
    // row 0 holds the spine area (a plane: normal xyz, offset w; x NaN for
    // none), rows 1.. the grid of SdfBuilder::computeModSdf +
    // computeSdfGradient + postProcessModSdfGradient (sdf_builder.cpp:162-240,
    // 242-330, 332-357): the unsigned distance to the closest triangle and its
    // normalised gradient, the whole row NaN where the closest point is not
    // on the surface -- where it falls on a boundary edge (one adjacent
    // triangle) or a vertex of one (distanceProjectionPointTriangleSquared,
    // src/geometry_utils/aabb_tree.cpp:519-688) -- unless the node lies on
    // the positive side of the spine plane (q . normalize(n) + w > 0).
    // The closest triangle is the first in index order, as in mesh_sdf.  The
    // reference marks the node off the surface when any candidate its tree
    // walk accepts on the way (d <= the best so far, aabb_tree.cpp:724-760)
    // falls on the boundary, so it may drop a node whose final closest point
    // is on the surface near the boundary; here only the final one counts.
    inline void mesh_slicer_sdf(const float *pts, const unsigned *tris, int numTris, float cell, const int *dims,
                                const float *lower, const float *spine, float *rows4)
    {
        const int dx = dims[0], dy = dims[1], dz = dims[2];
        auto P = [&](unsigned i) { return V3<float>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); };
        // boundary edges: one adjacent triangle (buildEdgeTriangleMap /
        // buildEdgeVerticesList, sdf_builder.cpp:45-82)
        std::vector<std::pair<uint64_t, int>> edges;
        edges.reserve(3 * numTris);
        auto key = [](unsigned u, unsigned v) { return (uint64_t)std::min(u, v) << 32 | std::max(u, v); };
        for (int t = 0; t < numTris; ++t)
            for (int j = 0; j < 3; ++j)
                edges.push_back({key(tris[3 * t + j], tris[3 * t + (j + 1) % 3]), t});
        std::sort(edges.begin(), edges.end());
        std::vector<uint64_t> boundaryEdges;
        std::vector<unsigned> boundaryVerts;
        for (size_t i = 0; i < edges.size();)
        {
            size_t j = i;
            while (j < edges.size() && edges[j].first == edges[i].first)
                ++j;
            if (j - i < 2)
            {
                boundaryEdges.push_back(edges[i].first);
                boundaryVerts.push_back((unsigned)(edges[i].first >> 32));
                boundaryVerts.push_back((unsigned)(edges[i].first & 0xffffffffu));
            }
            i = j;
        }
        std::sort(boundaryVerts.begin(), boundaryVerts.end());
        auto isBE = [&](unsigned u, unsigned v) {
            return std::binary_search(boundaryEdges.begin(), boundaryEdges.end(), key(u, v));
        };
        auto isBV = [&](unsigned u) { return std::binary_search(boundaryVerts.begin(), boundaryVerts.end(), u); };
        const bool hasSpine = !std::isnan(spine[0]);
        V3<float> sn(spine[0], spine[1], spine[2]);
        {
            const float s2 = sqn(sn);
            sn = s2 > 0.0f ? sn * rsq(s2) : V3<float>();
        }
        for (int k = 0; k < 4; ++k)
            rows4[k] = spine[k];
        float *grid4 = rows4 + 4;
        std::vector<char> onMesh((size_t)dx * dy * dz);
        for (int z = 0; z < dz; ++z)
            for (int y = 0; y < dy; ++y)
                for (int x = 0; x < dx; ++x)
                {
                    const V3<float> q(lower[0] + float(x) * cell, lower[1] + float(y) * cell, lower[2] + float(z) * cell);
                    float best = FLT_MAX;
                    int bestT = 0;
                    V3<float> bestP;
                    for (int t = 0; t < numTris; ++t)
                    {
                        V3<float> cp;
                        const float d2 = tri_closest_unit(q, P(tris[3 * t]), P(tris[3 * t + 1]), P(tris[3 * t + 2]), cp);
                        if (d2 < best)
                        {
                            best = d2;
                            bestT = t;
                            bestP = cp;
                        }
                    }
                    const size_t node = ((size_t)z * dy + y) * dx + x;
                    grid4[4 * node + 3] = norm(bestP - q);
                    V3<float> cp;
                    int region;
                    const unsigned *tv = tris + 3 * bestT;
                    tri_closest_unit(q, P(tv[0]), P(tv[1]), P(tv[2]), cp, &region);
                    bool off = false;
                    switch (region)
                    {
                    case kRegA: off = isBV(tv[0]); break;
                    case kRegB: off = isBV(tv[1]); break;
                    case kRegC: off = isBV(tv[2]); break;
                    case kRegAB: off = isBE(tv[0], tv[1]); break;
                    case kRegBC: off = isBE(tv[1], tv[2]); break;
                    case kRegAC: off = isBE(tv[0], tv[2]); break;
                    default: break;
                    }
                    if (hasSpine && dot(q, sn) + spine[3] > 0.0f)
                        off = false;
                    onMesh[node] = !off;
                }
        sdf_gradient(dims, cell, grid4);
        for (size_t i = 0; i < onMesh.size(); ++i)
            if (!onMesh[i])
                for (int k = 0; k < 4; ++k)
                    grid4[4 * i + k] = std::nanf("0");
    }

    inline void sdf_gradient(const int *dims, float cell, float *grid4)
    {
        const int dx = dims[0], dy = dims[1], dz = dims[2];
        const float inv = 1.0f / cell;
        auto W = [&](int x, int y, int z) { return grid4[4 * ((z * dy + y) * dx + x) + 3]; };
        auto diff = [&](int i, int n, float lo, float mid, float hi) {
            if (i > 0 && i < n - 1)
                return (hi - lo) * 0.5f * inv;
            return i == 0 ? (hi - mid) * inv : (mid - lo) * inv;
        };
        for (int z = 0; z < dz; ++z)
            for (int y = 0; y < dy; ++y)
                for (int x = 0; x < dx; ++x)
                {
                    const float m = W(x, y, z);
                    // a one-node axis has no neighbour: the reference then reads
                    // past the row; such tables are refused by the builders here
                    V3<float> g(diff(x, dx, x > 0 ? W(x - 1, y, z) : m, m, x < dx - 1 ? W(x + 1, y, z) : m),
                                diff(y, dy, y > 0 ? W(x, y - 1, z) : m, m, y < dy - 1 ? W(x, y + 1, z) : m),
                                diff(z, dz, z > 0 ? W(x, y, z - 1) : m, m, z < dz - 1 ? W(x, y, z + 1) : m));
                    normalize(g);
                    float *o = grid4 + 4 * ((z * dy + y) * dx + x);
                    o[0] = g.x;
                    o[1] = g.y;
                    o[2] = g.z;
                }
    }

} // namespace orc
