// Triangle mesh -> SDF table on the device (gfx950).
//
// Builds the table the engine computes on the host for a triangle-mesh
// geometry (SdfBuilder::precomputeSdfDesc / computeSdf / computeSdfGradient,
// src/geometry_utils/sdf_builder.cpp:84-160 and 242-330, called from
// src/engine/simulation_factory_impl.cpp:628-635): at node (x, y, z), placed
// at lowerBound + (x, y, z) * cellSize and stored at row z*dx*dy + y*dx + x,
// the signed distance to the mesh in .w and its normalised finite-difference
// gradient in .xyz.
//
// Work: every node against every triangle.  The reference walks an AABB tree
// per node for the closest triangle and a clustered far-field approximation
// of the winding number for the sign; here a wavefront owns a 4x4x4 brick of
// nodes and streams the triangles through LDS, each lane evaluating the
// point-triangle distance and the triangle's solid angle for its node.  The
// bricks keep the lanes of a wavefront in the same Voronoi region of most
// triangles, so the region branches rarely diverge.  Small tables split the
// triangle list into slices over more workgroups; a combine pass takes the
// slices in order.
//
// Results against the reference's arithmetic (the library is compiled with
// -ffp-contract=off and IEEE division and square root):
//   * distance: the closest triangle in index order with a strict <, its
//     closest point from the same unit-box evaluation as the reference, so
//     |distance| is the reference's bit for bit wherever the minimum is
//     unique; where triangles share the closest point the reference's tree
//     order may pick another one, equal up to rounding;
//   * sign: the exact winding number (the sum the reference's
//     computeWindingNumberExact forms, include/geometry_utils/winding_number.h:136-146)
//     against 0.5; the reference thresholds its beta = 2 approximation of the
//     same number (sdf_builder.cpp:150), so the two differ only where the
//     winding number is within that approximation's error of 0.5: on the
//     surface of a closed mesh, near the rim of an open one;
//   * gradient: the reference's differences and normalisation, so equal to it
//     wherever the neighbouring distances are.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "crmpm_hip.h"
#include "mpm_device.hpp"

namespace crhip
{
    namespace
    {
        constexpr int kSdfWG = 256;          // threads: 4 wavefronts, one 4x4x4 node brick each
        constexpr int kSdfChunk = 256;       // triangles staged in LDS per round
        constexpr int kSdfTargetWGs = 2048;  // slice the triangles until about this many workgroups

        struct MeshSdfArgs
        {
            const float4 *tri;  // 3 float4 per triangle: a, b, c (w unused)
            float4 *part;       // per slice and node: closest point, squared distance
            float *partWn;      // per slice and node: winding number of the slice's triangles
            float3 lower;
            float cell;
            int dx, dy, dz;
            int bx, by;         // bricks per axis (x, y)
            int nBricks;
            int numTris;
            int sliceTris;      // triangles per slice
            int groups;         // workgroups per slice
            long long count;    // nodes
        };

        // Closest point of triangle abc to p, region by region (Ericson,
        // Real-Time Collision Detection 5.1.5), with the reference's tests
        // and expressions (src/geometry_utils/aabb_tree.cpp:336-439).
        CR_DEV float tri_closest(V3 p, V3 a, V3 b, V3 c, V3 &cp)
        {
            const V3 ab = b - a, ac = c - a, bc = c - b, ap = p - a, bp = p - b, cq = p - c;
            const float d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp), d5 = dot(ab, cq),
                        d6 = dot(ac, cq);
            if (d1 <= 0.0f && d2 <= 0.0f)
                cp = a;
            else if (d3 >= 0.0f && d3 >= d4)
                cp = b;
            else if (d6 >= 0.0f && d6 >= d5)
                cp = c;
            else
            {
                const float vc = d1 * d4 - d3 * d2;
                const float va = d3 * d6 - d5 * d4;
                const float vb = d5 * d2 - d1 * d6;
                if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f)
                    cp = a + ab * (d1 / (d1 - d3));
                else if (va <= 0.0f && d4 >= d3 && d5 >= d6)
                {
                    const float unom = d4 - d3, udenom = d5 - d6;
                    cp = b + bc * (unom / (unom + udenom));
                }
                else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f)
                    cp = a + ac * (d2 / (d2 - d6));
                else
                {
                    const float den = 1.0f / (va + vb + vc);
                    cp = a + ab * (vb * den) + ac * (vc * den);
                }
            }
            const V3 v = p - cp;
            return dot(v, v);
        }

        CR_DEV V3 vmin(V3 u, V3 v) { return v3(fminf(u.x, v.x), fminf(u.y, v.y), fminf(u.z, v.z)); }
        CR_DEV V3 vmax(V3 u, V3 v) { return v3(fmaxf(u.x, v.x), fmaxf(u.y, v.y), fmaxf(u.z, v.z)); }

        // The reference evaluates the distance with point and triangle mapped
        // into a unit box: shifted by the componentwise min(max(a, b), min(c, q))
        // -- not the box minimum, kept as written -- and scaled by
        // 1 / max(1e-12, largest extent) (aabb_tree.cpp:443-466).
        CR_DEV float tri_closest_unit(V3 q, V3 a, V3 b, V3 c, V3 &cp)
        {
            const V3 lo = vmin(vmax(a, b), vmin(c, q));
            const V3 size = vmax(vmax(a, b), vmax(c, q)) - lo;
            const float inv = fmaxf(1e-12f, fmaxf(size.x, fmaxf(size.y, size.z)));
            const float s = 1.0f / inv;
            V3 u;
            const float r = tri_closest((q - lo) * s, (a - lo) * s, (b - lo) * s, (c - lo) * s, u);
            cp = u * inv + lo;
            return r * inv * inv;
        }

        // Solid angle of triangle abc seen from p over 4 pi (Van Oosterom and
        // Strackee), in the reference's expression order (evaluateExact,
        // include/geometry_utils/winding_number.h:79-93).
        CR_DEV float winding_term(V3 a, V3 b, V3 c, V3 p)
        {
            const float k = float(0.5 / 3.141592653589793238462643383);
            a = a - p;
            b = b - p;
            c = c - p;
            const float la = norm(a), lb = norm(b), lc = norm(c);
            const float y = a.x * b.y * c.z - a.x * b.z * c.y - a.y * b.x * c.z + a.y * b.z * c.x + a.z * b.x * c.y -
                            a.z * b.y * c.x;
            const float x = (la * lb * lc + (a.x * b.x + a.y * b.y + a.z * b.z) * lc +
                             (b.x * c.x + b.y * c.y + b.z * c.z) * la + (c.x * a.x + c.y * a.y + c.z * a.z) * lb);
            return k * atan2f(y, x);
        }

        CR_DEV V3 node_point(const MeshSdfArgs &a, int x, int y, int z)
        {
            return v3(a.lower.x + float(x) * a.cell, a.lower.y + float(y) * a.cell, a.lower.z + float(z) * a.cell);
        }

        // One workgroup: 4 bricks (one per wavefront) against one slice of the
        // triangles.  Writes the slice's closest point, squared distance and
        // winding-number sum per node.
        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_slices(MeshSdfArgs a)
        {
            __shared__ float4 tri[kSdfChunk * 3];
            const int slice = blockIdx.x / a.groups;
            const int brick = (blockIdx.x - slice * a.groups) * (kSdfWG / 64) + (int)(threadIdx.x >> 6);
            const int lane = threadIdx.x & 63;
            const int bz = brick / (a.bx * a.by), rem = brick - bz * a.bx * a.by;
            const int by = rem / a.bx, bx = rem - by * a.bx;
            const int x = bx * 4 + (lane & 3), y = by * 4 + ((lane >> 2) & 3), z = bz * 4 + (lane >> 4);
            const bool valid = brick < a.nBricks && x < a.dx && y < a.dy && z < a.dz;
            const V3 q = node_point(a, x, y, z);

            const int t0 = slice * a.sliceTris;
            const int t1 = min(a.numTris, t0 + a.sliceTris);
            float best = kFltMax, wn = 0.0f;
            V3 bestP = q;
            for (int base = t0; base < t1; base += kSdfChunk)
            {
                const int n = min(kSdfChunk, t1 - base);
                __syncthreads();
                for (int i = threadIdx.x; i < 3 * n; i += kSdfWG)
                    tri[i] = a.tri[3 * (long long)base + i];
                __syncthreads();
                for (int t = 0; t < n; ++t)
                {
                    const V3 ta = v3(tri[3 * t]), tb = v3(tri[3 * t + 1]), tc = v3(tri[3 * t + 2]);
                    V3 cp;
                    const float d2 = tri_closest_unit(q, ta, tb, tc, cp);
                    if (d2 < best)
                    {
                        best = d2;
                        bestP = cp;
                    }
                    wn += winding_term(ta, tb, tc, q);
                }
            }
            if (valid)
            {
                const long long node = ((long long)z * a.dy + y) * a.dx + x;
                const long long o = (long long)slice * a.count + node;
                a.part[o] = make_float4(bestP.x, bestP.y, bestP.z, best);
                a.partWn[o] = wn;
            }
        }

        // Slices in order (strict <, so the first closest triangle in index
        // order wins, as in one sequential pass); signed distance
        // |closest - q|, negative where the winding number exceeds 0.5
        // (sdf_builder.cpp:141-157).
        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_combine(MeshSdfArgs a, int slices, float *dist)
        {
            const long long node = (long long)blockIdx.x * kSdfWG + threadIdx.x;
            if (node >= a.count)
                return;
            const int x = (int)(node % a.dx), y = (int)((node / a.dx) % a.dy), z = (int)(node / ((long long)a.dx * a.dy));
            const V3 q = node_point(a, x, y, z);
            float best = kFltMax, wn = 0.0f;
            V3 bestP = q;
            for (int s = 0; s < slices; ++s)
            {
                const long long o = (long long)s * a.count + node;
                const float4 p = a.part[o];
                if (p.w < best)
                {
                    best = p.w;
                    bestP = v3(p);
                }
                wn += a.partWn[o];
            }
            const float d = norm(bestP - q);
            dist[node] = wn > 0.5f ? -d : d;
        }

        // computeSdfGradient (sdf_builder.cpp:242-330): central differences
        // inside, one-sided on the faces, normalised when non-zero.
        CR_DEV float sdf_diff(int i, int n, const float *w, long long at, long long step, float inv)
        {
            if (i > 0 && i < n - 1)
                return (w[at + step] - w[at - step]) * 0.5f * inv;
            if (i == 0)
                return (w[at + step] - w[at]) * inv;
            return (w[at] - w[at - step]) * inv;
        }

        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_gradient(int dx, int dy, int dz, long long count,
                                                                     float cell, const float *dist, float4 *out)
        {
            const long long node = (long long)blockIdx.x * kSdfWG + threadIdx.x;
            if (node >= count)
                return;
            const int x = (int)(node % dx), y = (int)((node / dx) % dy), z = (int)(node / ((long long)dx * dy));
            const float inv = 1.0f / cell;
            V3 g = v3(sdf_diff(x, dx, dist, node, 1, inv), sdf_diff(y, dy, dist, node, dx, inv),
                      sdf_diff(z, dz, dist, node, (long long)dx * dy, inv));
            normalize(g);
            out[node] = make_float4(g.x, g.y, g.z, dist[node]);
        }
    } // namespace
} // namespace crhip

using namespace crhip;

extern "C" {

int crHipMeshSdfDesc(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                     float cellSize, float fatten, int dims[3], float lowerBound[3])
{
    if (!points || !triangles || !dims || !lowerBound || numPoints < 1 || numTriangles < 1 || !(cellSize > 0.0f) ||
        !std::isfinite(cellSize) || !(fatten >= 0.0f) || !std::isfinite(fatten))
        return CRMPM_HIP_ERR_ARG;
    // The AABB tree's root box: the union of the triangles' boxes, each grown
    // by 1e-4 (buildTree, include/geometry_utils/winding_number.h:148-161),
    // then grown by fatten; dims = int(size / cellSize + 1)
    // (SdfBuilder::precomputeSdfDesc, src/geometry_utils/sdf_builder.cpp:84-101).
    float lo[3] = {kFltMax, kFltMax, kFltMax}, hi[3] = {-kFltMax, -kFltMax, -kFltMax};
    for (long long t = 0; t < 3ll * numTriangles; ++t)
    {
        if (triangles[t] >= (unsigned)numPoints)
            return CRMPM_HIP_ERR_ARG;
        for (int k = 0; k < 3; ++k)
        {
            const float v = points[3ll * triangles[t] + k];
            if (!std::isfinite(v))
                return CRMPM_HIP_ERR_ARG;
            lo[k] = std::fmin(lo[k], v - 1e-4f);
            hi[k] = std::fmax(hi[k], v + 1e-4f);
        }
    }
    long long count = 1;
    for (int k = 0; k < 3; ++k)
    {
        lowerBound[k] = lo[k] - fatten;
        const float n = ((hi[k] + fatten) - lowerBound[k]) / cellSize + 1.0f;
        if (!(n < 2147483647.0f))
            return CRMPM_HIP_ERR_CAPACITY;
        dims[k] = (int)n;
        count *= dims[k];
    }
    return count >= (1ll << 31) ? CRMPM_HIP_ERR_CAPACITY : CRMPM_HIP_OK;
}

int crHipMeshSdf(const float *points, int numPoints, const unsigned int *triangles, int numTriangles, float cellSize,
                 float fatten, int dims[3], float lowerBound[3], float *gradientSignedDistance,
                 unsigned long long capacity, void *stream)
{
    int rc = crHipMeshSdfDesc(points, numPoints, triangles, numTriangles, cellSize, fatten, dims, lowerBound);
    if (rc || !gradientSignedDistance)
        return rc;
    // one node along an axis leaves the reference's one-sided difference
    // reading outside the table; such tables are refused
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2)
        return CRMPM_HIP_ERR_ARG;
    const long long count = (long long)dims[0] * dims[1] * dims[2];
    if (capacity < (unsigned long long)count)
        return CRMPM_HIP_ERR_CAPACITY;

    MeshSdfArgs a{};
    a.lower = make_float3(lowerBound[0], lowerBound[1], lowerBound[2]);
    a.cell = cellSize;
    a.dx = dims[0];
    a.dy = dims[1];
    a.dz = dims[2];
    a.bx = (a.dx + 3) / 4;
    a.by = (a.dy + 3) / 4;
    a.nBricks = a.bx * a.by * ((a.dz + 3) / 4);
    a.numTris = numTriangles;
    a.count = count;
    a.groups = (a.nBricks + kSdfWG / 64 - 1) / (kSdfWG / 64);
    const int chunks = (numTriangles + kSdfChunk - 1) / kSdfChunk;
    const int slices = std::max(1, std::min(chunks, (kSdfTargetWGs + a.groups - 1) / a.groups));
    a.sliceTris = ((numTriangles + slices - 1) / slices + 63) / 64 * 64;
    const int usedSlices = (numTriangles + a.sliceTris - 1) / a.sliceTris;

    // triangle corners gathered on the host: 3 float4 per triangle
    std::vector<float4> tri(3 * (size_t)numTriangles);
    for (size_t t = 0; t < 3 * (size_t)numTriangles; ++t)
    {
        const float *p = points + 3 * (size_t)triangles[t];
        tri[t] = make_float4(p[0], p[1], p[2], 0.0f);
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    float4 *dTri = nullptr, *dPart = nullptr;
    float *dWn = nullptr, *dDist = nullptr;
    rc = CRMPM_HIP_ERR_ALLOC;
    if (hipMalloc(&dTri, tri.size() * sizeof(float4)) == hipSuccess &&
        hipMalloc(&dPart, (size_t)usedSlices * count * sizeof(float4)) == hipSuccess &&
        hipMalloc(&dWn, (size_t)usedSlices * count * sizeof(float)) == hipSuccess &&
        hipMalloc(&dDist, (size_t)count * sizeof(float)) == hipSuccess)
    {
        rc = CRMPM_HIP_ERR_LAUNCH;
        a.tri = dTri;
        a.part = dPart;
        a.partWn = dWn;
        const int nodeWGs = (int)((count + kSdfWG - 1) / kSdfWG);
        if (hipMemcpyAsync(dTri, tri.data(), tri.size() * sizeof(float4), hipMemcpyHostToDevice, st) == hipSuccess)
        {
            k_mesh_sdf_slices<<<a.groups * usedSlices, kSdfWG, 0, st>>>(a);
            k_mesh_sdf_combine<<<nodeWGs, kSdfWG, 0, st>>>(a, usedSlices, dDist);
            k_mesh_sdf_gradient<<<nodeWGs, kSdfWG, 0, st>>>(a.dx, a.dy, a.dz, count, cellSize, dDist,
                                                            reinterpret_cast<float4 *>(gradientSignedDistance));
            const bool launched = hipGetLastError() == hipSuccess;
            // the copy reads the host vector: wait for it even after a failed launch
            if (hipStreamSynchronize(st) == hipSuccess && launched)
                rc = CRMPM_HIP_OK;
        }
    }
    (void)hipFree(dTri);
    (void)hipFree(dPart);
    (void)hipFree(dWn);
    (void)hipFree(dDist);
    return rc;
}

} // extern "C"
