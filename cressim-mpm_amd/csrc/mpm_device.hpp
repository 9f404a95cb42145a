// Device math for the MPM substep kernels (gfx950).
//
// Scalar FP32 helpers, column-major 3x3 matrices, quaternion rotation, the
// 3x3 SVD, analytic/grid SDFs, rigid contact and slicer connectivity.  The
// library is compiled with -ffp-contract=off and IEEE division/sqrt, so the
// per-particle arithmetic follows the reference CPU path's evaluation order
// term for term (rsqrtf is 1/sqrt there, include/utils/cuda_compat.h:55);
// the summation order of the tile reductions differs, and the SVD (used by
// PB-MPM once per particle per iteration) is contracted to FMAs with the
// hardware rsqrt, as in the reference's CUDA build.
// Reference semantics reproduced (file:line of the reference):
//   Vec3::outerProduct(a, b)(r, c) = a_c * b_r     include/math/mat33.h:287-306
//   Mat3x3f::inverse() = inverse^T, 0 if |det|<1e-6 include/math/mat33.h:256-283
//   svd3 factors A^T (u * vt = R(A)^T)              include/math/svd3.h:1265-1275
//   rsqrtf on the host is 1/sqrt                    include/utils/cuda_compat.h:55
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CR_DEV __device__ __forceinline__
#define CR_DEV_NOINLINE __device__ __noinline__

namespace crhip
{
    constexpr float kFltEps = 1.19209290e-07f;
    constexpr float kFltMax = 3.4028234663852885981170418348452e+38F;

    // Geometry, shape and material enums (include/engine/geometry.h:42-55,
    // include/engine/shape.h:43-48, include/solver/material_data.h:41-46).
    enum : int { gBox = 0, gSphere, gPlane, gCapsule, gTriangleMesh, gQuadSlicer, gTriangleMeshSlicer, gLines, gArc };
    enum : int { sStatic = 0, sKinematic, sDynamic };
    enum : int { mCoRotational = 0, mNeoHookean, mLinearElastic };

    CR_DEV float rsq(float x) { return 1.0f / sqrtf(x); }
    // Hardware reciprocal square root (v_rsq_f32, about 1 ulp).  Used only in
    // the SVD's Jacobi and Givens steps, whose McAdams formulation is built for
    // an approximate rsqrt and refines it with a Newton step where it matters;
    // the reference's CUDA build takes the fast rsqrtf there as well.
    CR_DEV float rsq_svd(float x) { return __builtin_amdgcn_rsqf(x); }

    struct V3
    {
        float x, y, z;
    };
    CR_DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
    CR_DEV V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }
    CR_DEV V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
    CR_DEV V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
    CR_DEV V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
    CR_DEV V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
    CR_DEV V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
    CR_DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
    CR_DEV float sqn(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
    CR_DEV float norm(V3 a) { return sqrtf(sqn(a)); }
    CR_DEV V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
    CR_DEV void normalize(V3 &a)
    {
        float s = sqn(a);
        if (s > 0.0f)
            a = a * rsq(s);
    }

    // Column-major 3x3: c0, c1, c2 are columns.
    struct M3
    {
        V3 c0, c1, c2;
    };
    CR_DEV M3 m3_identity() { return M3{V3{1, 0, 0}, V3{0, 1, 0}, V3{0, 0, 1}}; }
    CR_DEV M3 m3_scalar(float s) { return M3{V3{s, 0, 0}, V3{0, s, 0}, V3{0, 0, s}}; }
    CR_DEV V3 mul(const M3 &m, V3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }
    CR_DEV M3 mul(const M3 &a, const M3 &b) { return M3{mul(a, b.c0), mul(a, b.c1), mul(a, b.c2)}; }
    CR_DEV M3 operator+(const M3 &a, const M3 &b) { return M3{a.c0 + b.c0, a.c1 + b.c1, a.c2 + b.c2}; }
    CR_DEV M3 operator-(const M3 &a, const M3 &b) { return M3{a.c0 - b.c0, a.c1 - b.c1, a.c2 - b.c2}; }
    CR_DEV M3 operator*(const M3 &a, float s) { return M3{a.c0 * s, a.c1 * s, a.c2 * s}; }
    CR_DEV M3 transpose(const M3 &a)
    {
        return M3{V3{a.c0.x, a.c1.x, a.c2.x}, V3{a.c0.y, a.c1.y, a.c2.y}, V3{a.c0.z, a.c1.z, a.c2.z}};
    }
    CR_DEV float trace(const M3 &a) { return a.c0.x + a.c1.y + a.c2.z; }
    // a = m(0,0) b = m(0,1) c = m(0,2); d..f row 1; g..i row 2
    CR_DEV float det(const M3 &m)
    {
        const float a = m.c0.x, b = m.c1.x, c = m.c2.x, d = m.c0.y, e = m.c1.y, f = m.c2.y, g = m.c0.z, h = m.c1.z, i = m.c2.z;
        return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
    }
    // Transposed inverse, zero for |det| < 1e-6 (reference Mat3x3f::inverse).
    CR_DEV M3 inverse_t(const M3 &m)
    {
        const float a = m.c0.x, b = m.c1.x, c = m.c2.x, d = m.c0.y, e = m.c1.y, f = m.c2.y, g = m.c0.z, h = m.c1.z, i = m.c2.z;
        const float dt = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
        if (fabsf(dt) < 1e-6f)
            return M3{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}};
        const float inv = 1.0f / dt;
        M3 adj{V3{(e * i - f * h), -(b * i - c * h), (b * f - c * e)},
               V3{-(d * i - f * g), (a * i - c * g), -(a * f - c * d)},
               V3{(d * h - e * g), -(a * h - b * g), (a * e - b * d)}};
        return adj * inv;
    }
    // Reference outer product: M(r, c) = a_c * b_r.
    CR_DEV M3 outer_ref(V3 a, V3 b)
    {
        return M3{V3{a.x * b.x, a.x * b.y, a.x * b.z}, V3{a.y * b.x, a.y * b.y, a.y * b.z}, V3{a.z * b.x, a.z * b.y, a.z * b.z}};
    }

    // Quaternion rotation, q = (x, y, z, w) (include/math/quat.h:117-152)
    CR_DEV V3 qrot(float4 q, V3 v)
    {
        const float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
        const float w2 = q.w * q.w - 0.5f;
        const float dot2 = (q.x * vx + q.y * vy + q.z * vz);
        return V3{(vx * w2 + (q.y * vz - q.z * vy) * q.w + q.x * dot2),
                  (vy * w2 + (q.z * vx - q.x * vz) * q.w + q.y * dot2),
                  (vz * w2 + (q.x * vy - q.y * vx) * q.w + q.z * dot2)};
    }
    CR_DEV V3 qrot_inv(float4 q, V3 v)
    {
        const float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
        const float w2 = q.w * q.w - 0.5f;
        const float dot2 = (q.x * vx + q.y * vy + q.z * vz);
        return V3{(vx * w2 - (q.y * vz - q.z * vy) * q.w + q.x * dot2),
                  (vy * w2 - (q.z * vx - q.x * vz) * q.w + q.y * dot2),
                  (vz * w2 - (q.x * vy - q.y * vx) * q.w + q.z * dot2)};
    }

    // ------------------------------------------------------------------
    // 3x3 SVD, McAdams et al. 2011 (branch-free Jacobi on A^T A with
    // quaternion accumulation, 4 sweeps; sort; Givens QR), all selects,
    // no divergent branches.  The sequence of operations is that of the
    // reference's svd3.h (include/math/svd3.h:74-1260), which derives from
    // "Computing the Singular Value Decomposition of 3x3 matrices with minimal
    // branching and elementary floating point operations" by A. McAdams,
    // A. Selle, R. Tamstorf, J. Teran and E. Sifakis (University of
    // Wisconsin-Madison technical report TR1690, 2011) and its BSD-licensed
    // reference implementation (Copyright (c) 2011 the authors): kept so that
    // PB-MPM results match the reference's bit for bit.
    // ------------------------------------------------------------------
    CR_DEV void svd_conj(float &sPP, float &sQQ, float &sPQ, float &sRR, float &sRP, float &sRQ,
                         float &qs, float &qR, float &qP, float &qQ)
    {
#pragma clang fp contract(fast)
        float sh = sPQ * 0.5f;
        float t5 = sPP - sQQ;
        float t2 = sh * sh;
        bool m = t2 >= 1.e-20f;
        sh = m ? sh : 0.0f;
        float ch = m ? t5 : 1.0f;
        float t1 = sh * sh;
        t2 = ch * ch;
        float t3 = t1 + t2;
        float t4 = rsq_svd(t3);
        sh = t4 * sh;
        ch = t4 * ch;
        t1 = 5.82842712474619f * t1;
        m = t2 <= t1;
        sh = m ? 0.3826834323650897f : sh;
        ch = m ? 0.9238795325112867f : ch;
        t1 = sh * sh;
        t2 = ch * ch;
        const float c = t2 - t1;
        float s = ch * sh;
        s = s + s;
        t3 = t1 + t2;
        sRR = sRR * t3;
        sRP = sRP * t3;
        sRQ = sRQ * t3;
        sRR = sRR * t3;
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
        const float tP = sh * qP, tQ = sh * qQ, tR = sh * qR;
        const float shq = sh * qs;
        qs = ch * qs;
        qP = ch * qP;
        qQ = ch * qQ;
        qR = ch * qR;
        qR = qR + shq;
        qs = qs - tR;
        qP = qP + tQ;
        qQ = qQ - tP;
    }

    CR_DEV float svd_refined_rsq(float t2)
    {
#pragma clang fp contract(fast)
        float t1 = rsq_svd(t2);
        float t4 = t1 * 0.5f;
        float t3 = t1 * t4;
        t3 = t1 * t3;
        t3 = t2 * t3;
        t1 = t1 + t4;
        return t1 - t3;
    }

    // Givens rotation eliminating b(q,p) with pivot b(p,p).  Rows of b and
    // columns of u are passed explicitly so everything stays in registers.
    CR_DEV void svd_givens(float app, float aqp, float &c, float &s)
    {
#pragma clang fp contract(fast)
        float sh = aqp * aqp;
        sh = (sh >= 1.e-12f) ? aqp : 0.0f;
        float ch = 0.0f - app;
        ch = fmaxf(ch, app);
        ch = fmaxf(ch, 1.e-12f);
        const bool m5 = app >= 0.0f;
        float t1 = ch * ch;
        float t2 = sh * sh;
        t2 = t1 + t2;
        t1 = svd_refined_rsq(t2);
        t1 = t1 * t2;
        ch = ch + t1;
        const float chs = m5 ? ch : sh;
        const float shs = m5 ? sh : ch;
        ch = chs;
        sh = shs;
        t1 = ch * ch;
        t2 = sh * sh;
        t2 = t1 + t2;
        t1 = svd_refined_rsq(t2);
        ch = ch * t1;
        sh = sh * t1;
        c = ch * ch;
        float ss = sh * sh;
        c = c - ss;
        s = sh * ch;
        s = s + s;
    }
    CR_DEV void svd_rot(float &a, float &b, float c, float s)
    {
#pragma clang fp contract(fast)
        const float t1 = s * a;
        const float t2 = s * b;
        a = c * a;
        b = c * b;
        a = a + t2;
        b = b - t1;
    }
    CR_DEV void cswap(bool m, float &a, float &b)
    {
        const float ta = m ? b : a;
        const float tb = m ? a : b;
        a = ta;
        b = tb;
    }

    // Kernel-level SVD of the row-major a (a_ij): a = U diag(s) V^T.
    CR_DEV void svd_kernel(float a11, float a12, float a13, float a21, float a22, float a23, float a31, float a32, float a33,
                           float (&U)[3][3], float (&V)[3][3], float (&sig)[3])
    {
#pragma clang fp contract(fast)
        float s11 = a11 * a11;
        s11 = a21 * a21 + s11;
        s11 = a31 * a31 + s11;
        float s21 = a12 * a11;
        s21 = a22 * a21 + s21;
        s21 = a32 * a31 + s21;
        float s31 = a13 * a11;
        s31 = a23 * a21 + s31;
        s31 = a33 * a31 + s31;
        float s22 = a12 * a12;
        s22 = a22 * a22 + s22;
        s22 = a32 * a32 + s22;
        float s32 = a13 * a12;
        s32 = a23 * a22 + s32;
        s32 = a33 * a32 + s32;
        float s33 = a13 * a13;
        s33 = a23 * a23 + s33;
        s33 = a33 * a33 + s33;
        float qs = 1.0f, qx = 0.0f, qy = 0.0f, qz = 0.0f;
#pragma unroll
        for (int sweep = 0; sweep < 4; ++sweep)
        {
            svd_conj(s11, s22, s21, s33, s31, s32, qs, qz, qx, qy);
            svd_conj(s22, s33, s32, s11, s21, s31, qs, qx, qy, qz);
            svd_conj(s33, s11, s31, s22, s32, s21, qs, qy, qz, qx);
        }
        float t2 = qs * qs;
        float t1 = qx * qx;
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
        t1 = qx * qx;
        t2 = qy * qy;
        float t3 = qz * qz;
        float v11 = qs * qs;
        float v22 = v11 - t1;
        float v33 = v22 - t2;
        v33 = v33 + t3;
        v22 = v22 + t2;
        v22 = v22 - t3;
        v11 = v11 + t1;
        v11 = v11 - t2;
        v11 = v11 - t3;
        t1 = qx + qx;
        t2 = qy + qy;
        t3 = qz + qz;
        float v32 = qs * t1;
        float v13 = qs * t2;
        float v21 = qs * t3;
        t1 = qy * t1;
        t2 = qz * t2;
        t3 = qx * t3;
        float v12 = t1 - v21;
        float v23 = t2 - v32;
        float v31 = t3 - v13;
        v21 = t1 + v21;
        v32 = t2 + v32;
        v13 = t3 + v13;
        // B = A V
        float b[3][3];
        const float ar[3][3] = {{a11, a12, a13}, {a21, a22, a23}, {a31, a32, a33}};
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
            float x0 = v11 * ar[r][0];
            x0 = x0 + v21 * ar[r][1];
            x0 = x0 + v31 * ar[r][2];
            float x1 = v12 * ar[r][0];
            x1 = x1 + v22 * ar[r][1];
            x1 = x1 + v32 * ar[r][2];
            float x2 = v13 * ar[r][0];
            x2 = x2 + v23 * ar[r][1];
            x2 = x2 + v33 * ar[r][2];
            b[r][0] = x0;
            b[r][1] = x1;
            b[r][2] = x2;
        }
        float v[3][3] = {{v11, v12, v13}, {v21, v22, v23}, {v31, v32, v33}};
        float rho[3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
        {
            float r0 = b[0][j] * b[0][j];
            r0 = r0 + b[1][j] * b[1][j];
            r0 = r0 + b[2][j] * b[2][j];
            rho[j] = r0;
        }
        // sort: (1,2) negate 2; (1,3) negate 1; (2,3) negate 3
        {
            const bool m = rho[0] < rho[1];
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                cswap(m, b[r][0], b[r][1]);
                cswap(m, v[r][0], v[r][1]);
            }
            cswap(m, rho[0], rho[1]);
            const float sg = m ? -1.0f : 1.0f;
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                b[r][1] = b[r][1] * sg;
                v[r][1] = v[r][1] * sg;
            }
        }
        {
            const bool m = rho[0] < rho[2];
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                cswap(m, b[r][0], b[r][2]);
                cswap(m, v[r][0], v[r][2]);
            }
            cswap(m, rho[0], rho[2]);
            const float sg = m ? -1.0f : 1.0f;
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                b[r][0] = b[r][0] * sg;
                v[r][0] = v[r][0] * sg;
            }
        }
        {
            const bool m = rho[1] < rho[2];
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                cswap(m, b[r][1], b[r][2]);
                cswap(m, v[r][1], v[r][2]);
            }
            cswap(m, rho[1], rho[2]);
            const float sg = m ? -1.0f : 1.0f;
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                b[r][2] = b[r][2] * sg;
                v[r][2] = v[r][2] * sg;
            }
        }
        float u[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        float c, s;
        // eliminate b21 (rows 0,1; U cols 0,1)
        svd_givens(b[0][0], b[1][0], c, s);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            svd_rot(b[0][j], b[1][j], c, s);
#pragma unroll
        for (int r = 0; r < 3; ++r)
            svd_rot(u[r][0], u[r][1], c, s);
        // eliminate b31 (rows 0,2)
        svd_givens(b[0][0], b[2][0], c, s);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            svd_rot(b[0][j], b[2][j], c, s);
#pragma unroll
        for (int r = 0; r < 3; ++r)
            svd_rot(u[r][0], u[r][2], c, s);
        // eliminate b32 (rows 1,2)
        svd_givens(b[1][1], b[2][1], c, s);
#pragma unroll
        for (int j = 0; j < 3; ++j)
            svd_rot(b[1][j], b[2][j], c, s);
#pragma unroll
        for (int r = 0; r < 3; ++r)
            svd_rot(u[r][1], u[r][2], c, s);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
            {
                U[i][j] = u[i][j];
                V[i][j] = v[i][j];
            }
        sig[0] = b[0][0];
        sig[1] = b[1][1];
        sig[2] = b[2][2];
    }

    // svd3(A): the kernel factors A^T; returns u and vt with A^T = u * S * vt.
    CR_DEV void svd3(const M3 &A, M3 &u, M3 &vt, V3 &sigma)
    {
        float U[3][3], V[3][3], s[3];
        // a_ij = A(j, i): row i of the kernel input is column i of A
        svd_kernel(A.c0.x, A.c0.y, A.c0.z, A.c1.x, A.c1.y, A.c1.z, A.c2.x, A.c2.y, A.c2.z, U, V, s);
        u = M3{V3{U[0][0], U[1][0], U[2][0]}, V3{U[0][1], U[1][1], U[2][1]}, V3{U[0][2], U[1][2], U[2][2]}};
        // vt(i, j) = V(j, i): column j of vt is row j of V
        vt = M3{V3{V[0][0], V[0][1], V[0][2]}, V3{V[1][0], V[1][1], V[1][2]}, V3{V[2][0], V[2][1], V[2][2]}};
        sigma = V3{s[0], s[1], s[2]};
    }

    // ------------------------------------------------------------------
    // Constitutive models (src/solver/constitutive_models.h:49-156)
    // ------------------------------------------------------------------
    CR_DEV M3 stress_kirchhoff(float lambda, float mu, int type, const M3 &F)
    {
        if (type == mNeoHookean)
        {
            float J = det(F);
            if (J < 1e-6f)
                J = 1e-6f;
            const M3 b = mul(F, transpose(F));
            return (b - m3_identity()) * mu + m3_scalar(lambda * logf(J));
        }
        if (type == mCoRotational || type == mLinearElastic)
        {
            // co-rotational falls through to linear elastic (StVK)
            const M3 strain = (mul(transpose(F), F) - m3_identity()) * 0.5f;
            const M3 S = m3_scalar(trace(strain)) * lambda + strain * (2.0f * mu);
            return mul(mul(F, S), transpose(F));
        }
        return M3{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}};
    }

    // PB-MPM material constraint on C (params: alpha = relaxation, r = ratio)
    CR_DEV void pb_constraint(float dt, float alpha, float ratio, int type, const M3 &F, M3 &C)
    {
        if (type != mNeoHookean && type != mCoRotational)
            return;
        M3 targetF = mul(m3_identity() + C * dt, F);
        M3 u, vt;
        V3 sig;
        svd3(targetF, u, vt, sig);
        const float detF = det(targetF);
        const float cl = fminf(fmaxf(fabsf(detF), 0.1f), 1000.0f);
        const M3 Q = targetF * (1.0f / (copysignf(1.0f, detF) * sqrtf(cl)));
        targetF = mul(u, vt) * ratio + Q * (1.0f - ratio);
        C = C + ((mul(targetF, inverse_t(F)) - m3_identity()) - C * dt) * alpha;
    }

    // Quadratic B-spline weights (src/solver/shape_functions.h:44-57)
    CR_DEV void bspline_w(float f, float &w0, float &w1, float &w2)
    {
        const float a = 1.5f - f, b = f - 1.0f, c = f - 0.5f;
        w0 = 0.5f * (a * a);
        w1 = 0.75f - (b * b);
        w2 = 0.5f * (c * c);
    }
    // Grid-unit gradients (src/solver/shape_functions.h:62-75); 1.5 is a double literal there
    CR_DEV void bspline_dw(float f, float &d0, float &d1, float &d2)
    {
        d0 = (float)((double)f - 1.5);
        d1 = 2.0f * (1.0f - f);
        d2 = f - 0.5f;
    }

    // ------------------------------------------------------------------
    // Shapes: a per-step copy of the bound shape tables in shapeIds order.
    // ------------------------------------------------------------------
    struct ShapeRec
    {
        int type, geomType, shapeId, pad;
        float4 pos, rot, invScale, linVel, angVel, params0, gparams, comInvMass, inertia0, inertia1;
    };

    struct ShapeView
    {
        const ShapeRec *recs;
        const int *counts;   // [0] shapes, [1] slicers
        const int *slicers;  // indices into recs
        float4 *force;       // engine ShapeData.force (dynamic coupling)
        float4 *torque;
        const int4 *sdfDims;
        const float4 *sdfLbcs;
        const float4 *sdfData;
        int effectiveMass;
    };

    // Shape records, slicer lists, SDF table headers and polyline points are
    // written by the host or by an earlier launch, never by the kernel that
    // reads them: they are read through the constant address space, so a
    // wavefront-uniform read is a scalar load, which neither occupies the
    // vector memory pipe nor joins its in-order completion counter (a
    // particle loop's loads would otherwise wait behind the previous chunk's
    // stores for every shape field it reads).  (The attribute only exists in
    // the device compilation pass.)
#if defined(__HIP_DEVICE_COMPILE__)
    template <class T>
    using ConstMem = const __attribute__((address_space(4))) T;
#else
    template <class T>
    using ConstMem = const T;
#endif
    template <class T>
    CR_DEV T const_load(const T *p, int i)
    {
        return ((ConstMem<T> *)(size_t)p)[i];
    }
    CR_DEV int shape_count(const ShapeView &sv) { return const_load(sv.counts, 0); }
    CR_DEV int slicer_count(const ShapeView &sv) { return const_load(sv.counts, 1); }

    CR_DEV int f2i(float f) { return __float_as_int(f); }

    CR_DEV float sdf_trilinear(const ShapeView &sv, int4 dim, int offset, int bx, int by, int bz, V3 fr)
    {
        float gv[8];
#pragma unroll
        for (int dx = 0; dx <= 1; ++dx)
#pragma unroll
            for (int dy = 0; dy <= 1; ++dy)
#pragma unroll
                for (int dz = 0; dz <= 1; ++dz)
                {
                    const int gi = offset + ((bz + dz) * dim.x * dim.y) + ((by + dy) * dim.x) + (bx + dx);
                    gv[(dx << 2) | (dy << 1) | dz] = sv.sdfData[gi].w;
                }
        const float fx = fr.x, fy = fr.y, fz = fr.z;
        const float ox = 1.0f - fx, oy = 1.0f - fy, oz = 1.0f - fz;
        const float w[8] = {ox * oy * oz, fx * oy * oz, ox * fy * oz, fx * fy * oz, ox * oy * fz, fx * oy * fz, ox * fy * fz, fx * fy * fz};
        float acc = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            acc += gv[i] * w[i];
        return acc;
    }

    // getGeometrySdf (include/geometry_utils/sdf.h:67-496); returns "in spine".
    CR_DEV bool geometry_sdf(const ShapeView &sv, V3 lp, int type, float4 p0, V3 &g, float &d, V3 &tangent)
    {
        bool inSpine = false;
        switch (type)
        {
        case gBox:
        {
            const V3 pen{fabsf(lp.x) - p0.x, fabsf(lp.y) - p0.y, fabsf(lp.z) - p0.z};
            const V3 s{lp.x < 0.0f ? -1.0f : 1.0f, lp.y < 0.0f ? -1.0f : 1.0f, lp.z < 0.0f ? -1.0f : 1.0f};
            const float gg = fmaxf(pen.x, fmaxf(pen.y, pen.z));
            const V3 q{fmaxf(pen.x, 0.0f), fmaxf(pen.y, 0.0f), fmaxf(pen.z, 0.0f)};
            const float l = norm(q);
            V3 dir;
            if (gg > 0.0f)
                dir = q / l;
            else if (pen.x > pen.y)
                dir = (pen.x > pen.z) ? V3{1, 0, 0} : V3{0, 0, 1};
            else
                dir = (pen.y > pen.z) ? V3{0, 1, 0} : V3{0, 0, 1};
            g = s * dir;
            d = l + fminf(gg, 0.0f);
            break;
        }
        case gSphere:
        {
            const float len = norm(lp);
            g = lp / fmaxf(len, kFltEps);
            d = len - p0.x;
            break;
        }
        case gPlane:
            g = V3{0, 1, 0};
            d = lp.y;
            break;
        case gCapsule:
        {
            const float hcyl = p0.y * 2.0f;
            V3 pa = lp;
            pa.y -= p0.y;
            float h = -pa.y / hcyl;
            h = fminf(fmaxf(h, 0.0f), 1.0f);
            pa.y += hcyl * h;
            const float dd = norm(pa);
            g = pa / fmaxf(dd, kFltEps);
            d = dd - p0.x;
            break;
        }
        case gArc:
        {
            const float ang = atan2f(lp.z, lp.x);
            if (ang < p0.x || ang > p0.y)
            {
                g = V3{0, 1, 0};
                d = __int_as_float(0x7fc00000);
                break;
            }
            const V3 pp{cosf(ang), 0.0f, sinf(ang)};
            const V3 dv = lp - pp;
            const float dist = norm(dv);
            g = dv / fmaxf(dist, kFltEps);
            d = dist;
            tangent = V3{-pp.z, 0.0f, pp.x};
            break;
        }
        case gTriangleMesh:
        {
            const int off = f2i(p0.x), id = f2i(p0.y);
            const float4 lb = const_load(sv.sdfLbcs, id);
            const int4 dim = const_load(sv.sdfDims, id);
            const V3 sp = (lp - v3(lb)) / lb.w;
            const int mx = (int)(sp.x + 0.5f), my = (int)(sp.y + 0.5f), mz = (int)(sp.z + 0.5f);
            if (mx < 0 || mx >= dim.x || my < 0 || my >= dim.y || mz < 0 || mz >= dim.z)
            {
                g = V3{0, 1, 0};
                d = kFltMax;
                break;
            }
            const int qi = off + mz * dim.x * dim.y + my * dim.x + mx;
            const float4 qv = sv.sdfData[qi];
            g = v3(qv);
            const int bx = (int)sp.x, by = (int)sp.y, bz = (int)sp.z;
            const V3 fr{sp.x - (float)bx, sp.y - (float)by, sp.z - (float)bz};
            if (bx < 0 || bx >= dim.x - 1 || by < 0 || by >= dim.y - 1 || bz < 0 || bz >= dim.z - 1)
            {
                d = qv.w;
                break;
            }
            d = sdf_trilinear(sv, dim, off, bx, by, bz, fr);
            break;
        }
        case gQuadSlicer:
        {
            if (lp.x < -p0.x || lp.x > p0.x || lp.z < -p0.y || lp.z > p0.y)
            {
                g = V3{0, 0, 0};
                d = kFltMax;
            }
            else
            {
                g = V3{0.0f, copysignf(1.0f, lp.y), 0.0f};
                d = fabsf(lp.y);
            }
            break;
        }
        case gTriangleMeshSlicer:
        {
            g = V3{0, 1, 0};
            d = __int_as_float(0x7fc00000);
            const int off = f2i(p0.x) + 1, id = f2i(p0.y);
            const float4 lb = const_load(sv.sdfLbcs, id);
            const int4 dim = const_load(sv.sdfDims, id);
            const V3 sp = (lp - v3(lb)) / lb.w;
            const int mx = (int)(sp.x + 0.5f), my = (int)(sp.y + 0.5f), mz = (int)(sp.z + 0.5f);
            if (mx < 1 || mx >= dim.x - 1 || my < 1 || my >= dim.y - 1 || mz < 1 || mz >= dim.z - 1)
                break;
            const int qi = off + mz * dim.x * dim.y + my * dim.x + mx;
            const float4 qv = sv.sdfData[qi];
            if (isnan(qv.w))
                break;
            g = v3(qv);
            const int bx = (int)sp.x, by = (int)sp.y, bz = (int)sp.z;
            const V3 fr{sp.x - (float)bx, sp.y - (float)by, sp.z - (float)bz};
            const float iv = sdf_trilinear(sv, dim, off, bx, by, bz, fr);
            if (isnan(iv))
                break;
            d = fabsf(iv);
            if (qv.w * iv < 0.0f)
                g = V3{-g.x, -g.y, -g.z};
            const float4 spn = const_load(sv.sdfData, off - 1);
            if (isnan(spn.x))
                break;
            const V3 sn = v3(spn);
            const float len = norm(sn);
            const float side = dot(lp, sn / len) + spn.w;
            if (side > 0.0f)
            {
                d -= len - 1.0f;
                inSpine = true;
            }
            break;
        }
        case gLines:
        {
            g = V3{0, 1, 0};
            d = __int_as_float(0x7fc00000);
            const int off = f2i(p0.x), id = f2i(p0.y), size = f2i(p0.z);
            const V3 bmin = v3(const_load(sv.sdfLbcs, id));
            const V3 bmax = v3(const_load(sv.sdfData, off));
            if (lp.x < bmin.x || lp.x > bmax.x || lp.y < bmin.y || lp.y > bmax.y || lp.z < bmin.z || lp.z > bmax.z)
                break;
            int closestIdx = 0;
            V3 closest{0, 0, 0};
            float closestD = kFltMax;
            V3 pa{0, 0, 0}, pb{0, 0, 0};
            for (int p = off + 1; p < off + size - 1; ++p)
            {
                pa = v3(const_load(sv.sdfData, p));
                pb = v3(const_load(sv.sdfData, p + 1));
                const V3 lv = pb - pa;
                const V3 pv = lp - pa;
                float t = dot(pv, lv) / sqn(lv);
                t = fmaxf(0.0f, fminf(1.0f, t));
                const V3 pr = pa + lv * t;
                const float dd = norm(lp - pr);
                if (dd < closestD)
                {
                    closestD = dd;
                    closest = pr;
                    closestIdx = p;
                }
            }
            pa = v3(const_load(sv.sdfData, off + 1));
            if (sqn(closest - pa) < kFltEps || sqn(closest - pb) < kFltEps)
                break;
            V3 gg = lp - closest;
            normalize(gg);
            tangent = v3(sv.sdfData[closestIdx]) - v3(sv.sdfData[closestIdx + 1]);
            normalize(tangent);
            g = gg;
            d = closestD;
            break;
        }
        default:
            break;
        }
        return inSpine;
    }

    // getGeometryModSdf for slicers (include/geometry_utils/sdf.h:504-662)
    CR_DEV bool geometry_mod_sdf(const ShapeView &sv, V3 lp, int type, float4 p0, V3 &g, float &d)
    {
        if (type == gQuadSlicer)
        {
            if (lp.x < -p0.x || lp.x > p0.x || lp.z < -p0.y || lp.z > p0.y)
                return false;
            g = V3{0.0f, copysignf(1.0f, lp.y), 0.0f};
            d = lp.y;
            return true;
        }
        if (type == gTriangleMeshSlicer)
        {
            g = V3{0, 1, 0};
            d = __int_as_float(0x7fc00000);
            int off = f2i(p0.x);
            const float4 spn = const_load(sv.sdfData, off);
            if (isnan(spn.x))
                return false;
            const float side = dot(lp, v3(spn)) + spn.w;
            if (side > 0.0f)
                return false;
            off++;
            const int id = f2i(p0.y);
            const float4 lb = const_load(sv.sdfLbcs, id);
            const int4 dim = const_load(sv.sdfDims, id);
            const V3 sp = (lp - v3(lb)) / lb.w;
            const int mx = (int)(sp.x + 0.5f), my = (int)(sp.y + 0.5f), mz = (int)(sp.z + 0.5f);
            if (mx < 1 || mx >= dim.x - 1 || my < 1 || my >= dim.y - 1 || mz < 1 || mz >= dim.z - 1)
                return false;
            const int qi = off + mz * dim.x * dim.y + my * dim.x + mx;
            const float4 qv = sv.sdfData[qi];
            if (isnan(qv.w))
                return false;
            g = v3(qv);
            const int bx = (int)sp.x, by = (int)sp.y, bz = (int)sp.z;
            const V3 fr{sp.x - (float)bx, sp.y - (float)by, sp.z - (float)bz};
            const float iv = sdf_trilinear(sv, dim, off, bx, by, bz, fr);
            if (isnan(iv))
                return false;
            d = iv;
            if (qv.w * iv < 0.0f)
                g = V3{-g.x, -g.y, -g.z};
            return true;
        }
        return false;
    }

    // checkOnDifferentSides (src/solver/mpm_rigid_contact.h:401-478); the
    // normal stays in the slicer's local frame.
    template <bool ComputeNormal>
    CR_DEV bool different_sides(const ShapeView &sv, int nSlicers, V3 p1, V3 p2, V3 &normal)
    {
        for (int k = 0; k < nSlicers; ++k)
        {
            const ShapeRec r = const_load(sv.recs, const_load(sv.slicers, k));
            const V3 is = v3(r.invScale);
            const V3 sp = v3(r.pos);
            const V3 r1 = qrot_inv(r.rot, p1 - sp) * is;
            const V3 r2 = qrot_inv(r.rot, p2 - sp) * is;
            V3 g1, g2;
            float d1, d2;
            if (!geometry_mod_sdf(sv, r1, r.geomType, r.gparams, g1, d1))
                continue;
            const bool on2 = geometry_mod_sdf(sv, r2, r.geomType, r.gparams, g2, d2);
            if (on2 && !(signbit(d1) == signbit(d2)))
            {
                if (ComputeNormal)
                {
                    const V3 g = g2 * is;
                    const float inv = rsq(sqn(g));
                    normal = g * inv;
                }
                return true;
            }
        }
        return false;
    }

    // True when no particle-node pair of the particle at x can straddle a
    // slicer, so different_sides is false for all 27 nodes: nodes lie within
    // 1.5 sqrt(3) dx of x, and a quad slicer only cuts a pair whose particle
    // is inside its extent with the node on the other side of its plane.
    // Triangle-mesh slicers are never ruled out here.
    CR_DEV bool slicer_free(const ShapeView &sv, int nSlicers, V3 x, float dx)
    {
        const float reach = 2.5980762f * dx; // 1.5 * sqrt(3) * dx
        for (int k = 0; k < nSlicers; ++k)
        {
            const ShapeRec r = const_load(sv.recs, const_load(sv.slicers, k));
            if (r.geomType != gQuadSlicer)
                return false;
            const V3 is = v3(r.invScale);
            const V3 l = qrot_inv(r.rot, x - v3(r.pos)) * is;
            const bool inside = !(l.x < -r.gparams.x || l.x > r.gparams.x || l.z < -r.gparams.y || l.z > r.gparams.y);
            if (inside && fabsf(l.y) <= reach * fabsf(is.y) * 1.0001f)
                return false;
        }
        return true;
    }

    // Momentum a dynamic shape receives from one wavefront's nodes: the lanes'
    // (momentum, arm) pairs are summed over the wavefront in a fixed shuffle
    // tree and lane 0 adds the sums with one set of atomics, instead of one
    // set per node (addMomentumToRigidBody, src/solver/mpm_rigid_contact.h:59-81).
    // All 64 lanes must call it.
    CR_DEV void wave_add_momentum(float4 *force, float4 *torque, int id, V3 mom, V3 arm)
    {
        const V3 tq = cross(arm, mom);
        float v[6] = {mom.x, mom.y, mom.z, tq.x, tq.y, tq.z};
        const bool any = (v[0] != 0.0f) || (v[1] != 0.0f) || (v[2] != 0.0f) || (v[3] != 0.0f) || (v[4] != 0.0f) ||
                         (v[5] != 0.0f);
        if (__ballot(any) == 0ull)
            return;
#pragma unroll
        for (int k = 0; k < 6; ++k)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
                v[k] += __shfl_xor(v[k], o, 64);
        if (__lane_id() == 0)
        {
            atomicAdd(&force[id].x, v[0]);
            atomicAdd(&force[id].y, v[1]);
            atomicAdd(&force[id].z, v[2]);
            atomicAdd(&torque[id].x, v[3]);
            atomicAdd(&torque[id].y, v[4]);
            atomicAdd(&torque[id].z, v[5]);
        }
    }

    // resolveRigidCollision (src/solver/mpm_rigid_contact.h:101-399).
    // Nodes: vel = (v, mass), pos = predicted node position; every lane of
    // the wavefront calls it (active false: a node without contact work), so
    // a dynamic shape's coupling momentum is summed per wavefront.
    // Particles: pushes pos out of shapes; vel is untouched.
    template <bool IsParticle>
    CR_DEV void resolve_rigid(const ShapeView &sv, int nShapes, float4 &vel, V3 &pos, bool active = true)
    {
        for (int si = 0; si < nShapes; ++si)
        {
            const ShapeRec r = const_load(sv.recs, si);
            const int gtype = r.geomType;
            if (IsParticle && (gtype == gLines || gtype == gArc))
                continue;
            V3 mom{0, 0, 0}, arm{0, 0, 0}; // this node's coupling momentum for a dynamic shape
            [&]() {
                if (!active)
                    return;
                const V3 is = v3(r.invScale);
                const V3 spos = v3(r.pos);
                const V3 worldRel = pos - spos;
                const V3 local = qrot_inv(r.rot, worldRel) * is;
                V3 gradient, tangent{0, 0, 0};
                float distance;
                const bool onSpine = geometry_sdf(sv, local, gtype, r.gparams, gradient, distance, tangent);
                if (isnan(distance))
                    return;
                gradient = gradient * is;
                const float inv = rsq(sqn(gradient));
                distance = distance * inv;
                gradient = gradient * inv;
                distance -= r.params0.w;
                const float smooth = IsParticle ? 0.0f : r.params0.x;
                if (!(distance < smooth))
                    return;
                const V3 normal = qrot(r.rot, gradient);
                const V3 rigidVel = (r.type == sStatic) ? V3{0, 0, 0} : (v3(r.linVel) + cross(v3(r.angVel), worldRel));
                const V3 relVel = v3(vel) - rigidVel;
                const float vn = dot(relVel, normal);
                const V3 nComp = normal * vn;
                if (IsParticle)
                {
                    const float inside = fminf(distance, 0.0f);
                    const V3 corr = normal * inside;
                    pos.x -= corr.x;
                    pos.y -= corr.y;
                    pos.z -= corr.z;
                    return;
                }
                if (vn >= 0.0f)
                    return;
                const float friction = onSpine ? 1e2f : r.params0.z;
                const float sticky = r.params0.y;
                V3 response{0, 0, 0};
                if (sv.effectiveMass && r.type == sDynamic)
                {
                    const float gridMass = vel.w;
                    const float invMassRigid = r.comInvMass.w;
                    const float rigidMass = invMassRigid > 0.0f ? 1.0f / invMassRigid : 0.0f;
                    if (rigidMass > 0.0f)
                    {
                        const V3 rr = worldRel - qrot(r.rot, v3(r.comInvMass));
                        const V3 rxnl = qrot_inv(r.rot, cross(rr, normal));
                        const float4 i0 = r.inertia0, i1 = r.inertia1;
                        const V3 wl{i0.x * rxnl.x + i1.x * rxnl.y + i1.y * rxnl.z,
                                    i1.x * rxnl.x + i0.y * rxnl.y + i1.z * rxnl.z,
                                    i1.y * rxnl.x + i1.z * rxnl.y + i0.z * rxnl.z};
                        const V3 t = cross(qrot(r.rot, wl), rr);
                        const float invMassRigidN = invMassRigid + dot(normal, t);
                        const float invMassGrid = 1.0f / gridMass;
                        const float invMassComb = invMassGrid + invMassRigidN;
                        if (invMassComb > 0.0f)
                        {
                            const float Jn = -1.0f * vn / invMassComb;
                            const V3 impN = normal * Jn;
                            const V3 vto = relVel - (normal * vn);
                            const float vtn = norm(vto);
                            V3 impT{0, 0, 0};
                            if (vtn > kFltEps)
                            {
                                const float stopJt = vtn / invMassComb;
                                const float maxJt = friction * Jn;
                                const float fJt = fminf(maxJt, stopJt);
                                const float left = vtn - (fJt * invMassComb);
                                const float fin = left * sticky;
                                const float totJt = (vtn - fin) / invMassComb;
                                impT = (vto * -1.0f / vtn) * totJt;
                            }
                            const V3 totImp = impN + impT;
                            response = v3(vel) + (totImp * invMassGrid);
                            mom = totImp * -1.0f;
                            arm = rr;
                        }
                    }
                }
                else
                {
                    V3 tang = relVel - nComp;
                    if (gtype == gLines || gtype == gArc)
                    {
                        tangent = tangent * is;
                        normalize(tangent);
                        tangent = qrot(r.rot, tangent);
                        tang = tangent * dot(tang, tangent);
                    }
                    const float tn = norm(tang);
                    const float fc = fmaxf(friction * vn / (tn + kFltEps), -1.0f);
                    response = tang + tang * fc;
                    response = response * sticky;
                    response = response + rigidVel;
                    if (!sv.effectiveMass && r.type == sDynamic)
                    {
                        mom = (v3(vel) - response) * vel.w;
                        arm = pos - spos;
                    }
                }
                vel.x = response.x;
                vel.y = response.y;
                vel.z = response.z;
            }();
            if (!IsParticle && r.type == sDynamic)
                wave_add_momentum(sv.force, sv.torque, r.shapeId, mom, arm);
        }
    }
} // namespace crhip
