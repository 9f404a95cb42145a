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
// nodes and walks the triangles in clusters of 16: a cluster far from the
// brick adds its far-field winding term (with an error bound), a near one
// each triangle's solid angle, and the triangles' distances are evaluated
// unless their boxes are provably too far from the brick for any lane to
// take them; triangles are staged through LDS only for clusters some
// wavefront of the workgroup needs.  The bricks keep the lanes of a wavefront in the
// same Voronoi region of most triangles, so the region branches rarely
// diverge.  Small tables split the triangle list into slices over more
// workgroups; a combine pass takes the slices in order, and the nodes whose
// sign the clustered far-field winding sum leaves undecided get the exact sum
// in one more pass.
//
// Results against the reference's arithmetic (the distances are evaluated
// under -ffp-contract=off with IEEE division and square root):
//   * distance: the closest triangle in index order with a strict <, its
//     closest point from the same unit-box evaluation as the reference, so
//     |distance| is the reference's bit for bit wherever the minimum is
//     unique; where triangles share the closest point the reference's tree
//     order may pick another one, equal up to rounding;
//   * sign: the exact winding number (the sum the reference's
//     computeWindingNumberExact forms, include/geometry_utils/winding_number.h:136-146,
//     here with FMAs and a polynomial atan2, about 1e-7 per triangle from its
//     FP32 evaluation) against 0.5; the reference thresholds its beta = 2
//     approximation of the same number (sdf_builder.cpp:150), so the two
//     differ only where the winding number is within that approximation's
//     error of 0.5: on the surface of a closed mesh, near the rim of an open one;
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
        constexpr int kClu = 16;             // triangles per cluster (consecutive in index order)
        constexpr int kCluPerChunk = kSdfChunk / kClu;
        // A cluster of radius rho is summed by its far-field term for a brick
        // at least kFar * rho from its centre.
        constexpr float kFar = 6.0f;
        // Winding numbers within this of 0.5 plus the far-field bound are
        // summed again exactly.
        constexpr float kWnSlack = 1e-3f;

        struct MeshSdfArgs
        {
            const float4 *tri;  // 4 float4 per triangle: (a, lo.x), (b, lo.y), (c, lo.z), (hi, cull slack)
            const float4 *cen;  // per triangle: centroid
            const float4 *clu;  // 4 float4 per cluster: (centre, radius), (vector area / 4 pi,
                                // 2 radius area / 4 pi), (box lo, cull slack), (box hi, 0)
            float *partErr;     // per slice and node: bound on the far-field error of the winding number
            float4 *part;       // per slice and node: closest point, squared distance
            float *partWn;      // per slice and node: winding number of the slice's triangles
                                // (mesh-slicer table: the closest triangle's index)
            const int *flags;   // mesh-slicer table: per triangle, boundary edges ab, bc, ac (bits 0-2)
                                // and boundary vertices a, b, c (bits 3-5)
            float4 spine;       // mesh-slicer table: spine area (x NaN: none)
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
        // Regions of the closest point: vertices a, b, c, edges ab, bc, ac, face.
        enum : int { kRegA = 0, kRegB, kRegC, kRegAB, kRegBC, kRegAC, kRegFace };

        CR_DEV float tri_closest(V3 p, V3 a, V3 b, V3 c, V3 &cp, int &region)
        {
            const V3 ab = b - a, ac = c - a, bc = c - b, ap = p - a, bp = p - b, cq = p - c;
            const float d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp), d5 = dot(ab, cq),
                        d6 = dot(ac, cq);
            if (d1 <= 0.0f && d2 <= 0.0f)
                cp = a, region = kRegA;
            else if (d3 >= 0.0f && d3 >= d4)
                cp = b, region = kRegB;
            else if (d6 >= 0.0f && d6 >= d5)
                cp = c, region = kRegC;
            else
            {
                const float vc = d1 * d4 - d3 * d2;
                const float va = d3 * d6 - d5 * d4;
                const float vb = d5 * d2 - d1 * d6;
                if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f)
                    cp = a + ab * (d1 / (d1 - d3)), region = kRegAB;
                else if (va <= 0.0f && d4 >= d3 && d5 >= d6)
                {
                    const float unom = d4 - d3, udenom = d5 - d6;
                    cp = b + bc * (unom / (unom + udenom)), region = kRegBC;
                }
                else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f)
                    cp = a + ac * (d2 / (d2 - d6)), region = kRegAC;
                else
                {
                    const float den = 1.0f / (va + vb + vc);
                    cp = a + ab * (vb * den) + ac * (vc * den), region = kRegFace;
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
        CR_DEV float tri_closest_unit(V3 q, V3 a, V3 b, V3 c, V3 &cp, int &region)
        {
            const V3 lo = vmin(vmax(a, b), vmin(c, q));
            const V3 size = vmax(vmax(a, b), vmax(c, q)) - lo;
            const float inv = fmaxf(1e-12f, fmaxf(size.x, fmaxf(size.y, size.z)));
            const float s = 1.0f / inv;
            V3 u;
            const float r = tri_closest((q - lo) * s, (a - lo) * s, (b - lo) * s, (c - lo) * s, u, region);
            cp = u * inv + lo;
            return r * inv * inv;
        }

        CR_DEV float tri_closest_unit(V3 q, V3 a, V3 b, V3 c, V3 &cp)
        {
            int region;
            return tri_closest_unit(q, a, b, c, cp, region);
        }

        // atan2 from an odd polynomial for atan on [0, 1] and the octant
        // identities atan(t) = pi/2 - atan(1/t), atan2(y, -x) = pi - atan2(y, x).
        // Coefficients fitted here (Lawson-reweighted least squares on [0, 1]);
        // absolute error 1.4e-7 rad evaluated in FP32.  atan2(0, 0) = 0.
        CR_DEV float atan2_poly(float y, float x)
        {
            const float ax = fabsf(x), ay = fabsf(y);
            const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
            const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
            const float s = t * t;
            float p = -0.004054564516991377f;
            p = fmaf(p, s, 0.021862953901290894f);
            p = fmaf(p, s, -0.055912334471940994f);
            p = fmaf(p, s, 0.0964219942688942f);
            p = fmaf(p, s, -0.1390863060951233f);
            p = fmaf(p, s, 0.19946566224098206f);
            p = fmaf(p, s, -0.33329862356185913f);
            p = fmaf(p, s, 0.9999993443489075f);
            float r = p * t;
            r = ay > ax ? 1.5707963267948966f - r : r;
            r = x < 0.0f ? 3.141592653589793f - r : r;
            return copysignf(r, y);
        }

        // Solid angle of triangle abc seen from p over 4 pi (Van Oosterom and
        // Strackee): the term of the exact winding number
        // (computeWindingNumberExact / evaluateExact,
        // include/geometry_utils/winding_number.h:79-93, 136-146), evaluated
        // with FMAs, the hardware square root and atan2_poly.  Only its sum's
        // side of 0.5 is used, and the sum differs from the reference's FP32
        // evaluation by about 1e-7 per term.
        CR_DEV float winding_term(V3 a, V3 b, V3 c, V3 p)
        {
            const float k = float(0.5 / 3.141592653589793238462643383);
            a = a - p;
            b = b - p;
            c = c - p;
            const float la = __builtin_amdgcn_sqrtf(fmaf(a.x, a.x, fmaf(a.y, a.y, a.z * a.z)));
            const float lb = __builtin_amdgcn_sqrtf(fmaf(b.x, b.x, fmaf(b.y, b.y, b.z * b.z)));
            const float lc = __builtin_amdgcn_sqrtf(fmaf(c.x, c.x, fmaf(c.y, c.y, c.z * c.z)));
            const float bcx = fmaf(b.y, c.z, -b.z * c.y), bcy = fmaf(b.z, c.x, -b.x * c.z),
                        bcz = fmaf(b.x, c.y, -b.y * c.x);
            const float y = fmaf(a.x, bcx, fmaf(a.y, bcy, a.z * bcz));
            const float ab = fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z));
            const float bc = fmaf(b.x, c.x, fmaf(b.y, c.y, b.z * c.z));
            const float ca = fmaf(c.x, a.x, fmaf(c.y, a.y, c.z * a.z));
            const float x = fmaf(ca, lb, fmaf(bc, la, fmaf(ab, lc, la * lb * lc)));
            return k * atan2_poly(y, x);
        }

        // Squared gap between two boxes (0 where they overlap): a lower bound
        // of the squared distance between any points of the two.
        CR_DEV float box_gap2(V3 lo1, V3 hi1, V3 lo2, V3 hi2)
        {
            const float gx = fmaxf(0.0f, fmaxf(lo2.x - hi1.x, lo1.x - hi2.x));
            const float gy = fmaxf(0.0f, fmaxf(lo2.y - hi1.y, lo1.y - hi2.y));
            const float gz = fmaxf(0.0f, fmaxf(lo2.z - hi1.z, lo1.z - hi2.z));
            return gx * gx + gy * gy + gz * gz;
        }

        // Squared distance from a point to a box (0 inside).
        CR_DEV float point_box_gap2(V3 p, V3 lo, V3 hi)
        {
            const float gx = fmaxf(0.0f, fmaxf(lo.x - p.x, p.x - hi.x));
            const float gy = fmaxf(0.0f, fmaxf(lo.y - p.y, p.y - hi.y));
            const float gz = fmaxf(0.0f, fmaxf(lo.z - p.z, p.z - hi.z));
            return gx * gx + gy * gy + gz * gz;
        }

        CR_DEV float wave_max(float v)
        {
            for (int o = 32; o > 0; o >>= 1)
                v = fmaxf(v, __shfl_xor(v, o));
            return v;
        }

        // Smallest d over the wavefront, ties to the smaller index; every lane
        // ends with the same pair.
        CR_DEV void wave_argmin(float &d, int &i)
        {
            for (int o = 32; o > 0; o >>= 1)
            {
                const float od = __shfl_xor(d, o);
                const int oi = __shfl_xor(i, o);
                if (od < d || (od == d && oi < i))
                {
                    d = od;
                    i = oi;
                }
            }
        }

        CR_DEV V3 node_point(const MeshSdfArgs &a, int x, int y, int z)
        {
            return v3(a.lower.x + float(x) * a.cell, a.lower.y + float(y) * a.cell, a.lower.z + float(z) * a.cell);
        }

        // One workgroup: 4 bricks (one per wavefront) against one slice of the
        // triangles.  Writes the slice's closest point, squared distance and
        // winding-number sum per node.
        //
        // The distance evaluation is skipped, for the whole wavefront, for a
        // triangle whose box lies farther from the brick's box than the
        // largest squared distance any lane still has to beat (plus a relative
        // 1e-5 and an absolute 1e-5 of the triangle's squared extent, far
        // above the evaluation's rounding): no lane could take it.  The first
        // bound comes from the triangle whose centroid is nearest the brick's
        // centre, then from the lanes' running minima after each chunk.  The
        // triangles are still visited in index order with a strict <, so the
        // result is the one of the full evaluation.
        // Slicer: the mesh-slicer table, unsigned, so no winding number; the
        // closest triangle's index is kept instead.
        template <bool Slicer>
        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_slices(MeshSdfArgs a)
        {
            __shared__ float4 lds[kSdfChunk * 4 + kCluPerChunk * 4];
            const int slice = blockIdx.x / a.groups;
            const int brick = (blockIdx.x - slice * a.groups) * (kSdfWG / 64) + (int)(threadIdx.x >> 6);
            const int lane = threadIdx.x & 63;
            const int bz = brick / (a.bx * a.by), rem = brick - bz * a.bx * a.by;
            const int by = rem / a.bx, bx = rem - by * a.bx;
            const int x = bx * 4 + (lane & 3), y = by * 4 + ((lane >> 2) & 3), z = bz * 4 + (lane >> 4);
            const bool valid = brick < a.nBricks && x < a.dx && y < a.dy && z < a.dz;
            const V3 q = node_point(a, x, y, z);
            // the box of all 64 lanes' points, valid or not
            const V3 blo = node_point(a, bx * 4, by * 4, bz * 4), bhi = node_point(a, bx * 4 + 3, by * 4 + 3, bz * 4 + 3);

            const int t0 = slice * a.sliceTris;
            const int t1 = min(a.numTris, t0 + a.sliceTris);

            // seed: the slice's triangle with the centroid nearest the brick's centre
            const V3 mid = (blo + bhi) * 0.5f;
            float cd = kFltMax;
            int ci = t0;
            for (int base = t0; base < t1; base += 4 * kSdfChunk)
            {
                const int n = min(4 * kSdfChunk, t1 - base);
                __syncthreads();
                for (int i = threadIdx.x; i < n; i += kSdfWG)
                    lds[i] = a.cen[base + i];
                __syncthreads();
                for (int t = lane; t < n; t += 64)
                {
                    const float d = sqn(v3(lds[t]) - mid);
                    if (d < cd)
                    {
                        cd = d;
                        ci = base + t;
                    }
                }
            }
            wave_argmin(cd, ci);
            ci = __builtin_amdgcn_readfirstlane(ci);
            V3 cp;
            float bound = wave_max(tri_closest_unit(q, v3(a.tri[4 * ci]), v3(a.tri[4 * ci + 1]),
                                                    v3(a.tri[4 * ci + 2]), cp));

            float best = kFltMax, wn = 0.0f, err = 0.0f;
            int bestT = t0;
            V3 bestP = q;
            float4 *cl = lds + 4 * kSdfChunk;
            for (int base = t0; base < t1; base += kSdfChunk)
            {
                const int n = min(kSdfChunk, t1 - base);
                const int nc = (n + kClu - 1) / kClu;
                __syncthreads();
                for (int i = threadIdx.x; i < 4 * nc; i += kSdfWG)
                    cl[i] = a.clu[4 * (long long)(base / kClu) + i];
                __syncthreads();
                // per cluster: is any distance still to be evaluated, and is it
                // near enough to need the exact winding terms?  Far clusters
                // add their far-field term and its error bound here.
                unsigned need = 0, nearMask = 0;
                for (int k = 0; k < nc; ++k)
                {
                    const float4 c0 = cl[4 * k], c1 = cl[4 * k + 1], c2 = cl[4 * k + 2], c3 = cl[4 * k + 3];
                    const bool dist = !(box_gap2(blo, bhi, v3(c2), v3(c3)) > fmaf(bound, 1.00001f, c2.w));
                    bool near = false;
                    if (!Slicer)
                    {
                        near = !(point_box_gap2(v3(c0), blo, bhi) > kFar * kFar * c0.w * c0.w);
                        if (!near)
                        {
                            // The cluster's solid angle from q is the sum over its
                            // triangles of int_T n.(x - q) / |x - q|^3 dA; with x
                            // replaced by the centre c it is N.(c - q) / R^3 (N the
                            // vector area), and the integrand's gradient in x is at
                            // most 2 / r^3, so the error is at most
                            // 2 rho A / (R - rho)^3 for R = |c - q| > rho.
                            const V3 v = v3(c0) - q;
                            const float r2 = fmaf(v.x, v.x, fmaf(v.y, v.y, v.z * v.z));
                            const float r = __builtin_amdgcn_sqrtf(r2);
                            const float term = fmaf(c1.x, v.x, fmaf(c1.y, v.y, c1.z * v.z)) *
                                               __builtin_amdgcn_rcpf(r2 * r);
                            const float g = r - c0.w;
                            wn += term;
                            err += fmaf(c1.w, __builtin_amdgcn_rcpf(g * g * g), 1e-5f * fabsf(term));
                        }
                    }
                    need |= (unsigned)(dist || near) << k;
                    nearMask |= (unsigned)near << k;
                }
                need = __builtin_amdgcn_readfirstlane(need);
                nearMask = __builtin_amdgcn_readfirstlane(nearMask);
                if (__syncthreads_or(need != 0))
                {
                    for (int i = threadIdx.x; i < 4 * n; i += kSdfWG)
                        lds[i] = a.tri[4 * (long long)base + i];
                    __syncthreads();
                    while (need)
                    {
                        const int k = __builtin_ctz(need);
                        need &= need - 1;
                        const bool near = (nearMask >> k) & 1;
                        for (int t = k * kClu; t < min(n, (k + 1) * kClu); ++t)
                        {
                            const float4 r0 = lds[4 * t], r1 = lds[4 * t + 1], r2 = lds[4 * t + 2],
                                         r3 = lds[4 * t + 3];
                            const V3 ta = v3(r0), tb = v3(r1), tc = v3(r2);
                            const float gap = box_gap2(blo, bhi, v3(r0.w, r1.w, r2.w), v3(r3));
                            if (!__builtin_amdgcn_readfirstlane((int)(gap > fmaf(bound, 1.00001f, r3.w))))
                            {
                                const float d2 = tri_closest_unit(q, ta, tb, tc, cp);
                                if (d2 < best)
                                {
                                    best = d2;
                                    bestP = cp;
                                    bestT = base + t;
                                }
                            }
                            if (!Slicer && near)
                                wn += winding_term(ta, tb, tc, q);
                        }
                    }
                }
                bound = fminf(bound, wave_max(best));
            }
            if (valid)
            {
                const long long node = ((long long)z * a.dy + y) * a.dx + x;
                const long long o = (long long)slice * a.count + node;
                a.part[o] = make_float4(bestP.x, bestP.y, bestP.z, best);
                a.partWn[o] = Slicer ? __int_as_float(bestT) : wn;
                if (!Slicer)
                    a.partErr[o] = err;
            }
        }

        // Slices in order (strict <, so the first closest triangle in index
        // order wins, as in one sequential pass); signed distance
        // |closest - q|, negative where the winding number exceeds 0.5
        // (sdf_builder.cpp:141-157).
        //
        // Mesh-slicer table: the unsigned distance, and whether the closest
        // point lies on the surface: not on a boundary edge (one adjacent
        // triangle) or a vertex of one, as distanceProjectionPointTriangleSquared
        // classifies it (src/geometry_utils/aabb_tree.cpp:519-688), or the node
        // is on the positive side of the spine plane (sdf_builder.cpp:226-236).
        // The region is re-evaluated for the closest triangle with the same
        // arithmetic.  The reference's tree walk drops the node when any
        // candidate it accepts on the way falls on the boundary
        // (aabb_tree.cpp:724-760); here only the closest one counts, so the
        // nodes kept are a superset of the reference's.
        // Unsigned: the distance only (the reference rule signs it with
        // k_mesh_sdf_sign_tree afterwards); the slices then carry the closest
        // triangle's index, as for the slicer.
        template <bool Slicer, bool Unsigned = false>
        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_combine(MeshSdfArgs a, int slices, float *dist,
                                                                    unsigned char *onMesh)
        {
            const long long node = (long long)blockIdx.x * kSdfWG + threadIdx.x;
            if (node >= a.count)
                return;
            const int x = (int)(node % a.dx), y = (int)((node / a.dx) % a.dy), z = (int)(node / ((long long)a.dx * a.dy));
            const V3 q = node_point(a, x, y, z);
            float best = kFltMax, wn = 0.0f, err = 0.0f;
            int bestT = 0;
            V3 bestP = q;
            for (int s = 0; s < slices; ++s)
            {
                const long long o = (long long)s * a.count + node;
                const float4 p = a.part[o];
                if (p.w < best)
                {
                    best = p.w;
                    bestP = v3(p);
                    bestT = __float_as_int(a.partWn[o]);
                }
                if (!Slicer && !Unsigned)
                {
                    wn += a.partWn[o];
                    err += a.partErr[o];
                }
            }
            const float d = norm(bestP - q);
            if (Unsigned)
            {
                dist[node] = d;
                return;
            }
            if (!Slicer)
            {
                // where the far-field sum cannot decide the side of 0.5 the
                // node is flagged (onMesh carries the flag here) and
                // k_mesh_sdf_exact sums every triangle's term
                const bool undecided = !(fabsf(wn - 0.5f) > err + kWnSlack);
                onMesh[node] = undecided ? 1 : 0;
                dist[node] = (!undecided && wn > 0.5f) ? -d : d;
                return;
            }
            dist[node] = d;
            V3 cp;
            int region;
            tri_closest_unit(q, v3(a.tri[4 * bestT]), v3(a.tri[4 * bestT + 1]), v3(a.tri[4 * bestT + 2]), cp, region);
            const int f = a.flags[bestT];
            bool off = region != kRegFace && ((f >> (region < 3 ? region + 3 : region - 3)) & 1);
            if (!isnan(a.spine.x))
            {
                V3 n = v3(a.spine.x, a.spine.y, a.spine.z);
                const float s2 = sqn(n);
                n = s2 > 0.0f ? n * rsq(s2) : v3(0.0f, 0.0f, 0.0f);
                if (dot(q, n) + a.spine.w > 0.0f)
                    off = false;
            }
            onMesh[node] = off ? 0 : 1;
        }

        // The exact winding number of the nodes the combine pass left
        // undecided, bricks as in k_mesh_sdf_slices, every triangle staged
        // through LDS; workgroups whose four bricks hold no undecided node
        // leave at once.
        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_exact(MeshSdfArgs a, const unsigned char *undecided,
                                                                  float *dist)
        {
            __shared__ float4 lds[kSdfChunk * 3];
            const int brick = (int)blockIdx.x * (kSdfWG / 64) + (int)(threadIdx.x >> 6);
            const int lane = threadIdx.x & 63;
            const int bz = brick / (a.bx * a.by), rem = brick - bz * a.bx * a.by;
            const int by = rem / a.bx, bx = rem - by * a.bx;
            const int x = bx * 4 + (lane & 3), y = by * 4 + ((lane >> 2) & 3), z = bz * 4 + (lane >> 4);
            const bool valid = brick < a.nBricks && x < a.dx && y < a.dy && z < a.dz;
            const long long node = ((long long)z * a.dy + y) * a.dx + x;
            const bool need = valid && undecided[node];
            if (!__syncthreads_or(need))
                return;
            const V3 q = node_point(a, x, y, z);
            float wn = 0.0f;
            for (int base = 0; base < a.numTris; base += kSdfChunk)
            {
                const int n = min(kSdfChunk, a.numTris - base);
                __syncthreads();
                for (int i = threadIdx.x; i < 3 * n; i += kSdfWG)
                    lds[i] = a.tri[4 * (long long)base + 4 * (i / 3) + i % 3];
                __syncthreads();
                if (need)
                    for (int t = 0; t < n; ++t)
                        wn += winding_term(v3(lds[3 * t]), v3(lds[3 * t + 1]), v3(lds[3 * t + 2]), q);
            }
            if (need && wn > 0.5f)
                dist[node] = -dist[node];
        }

        // ------------------------------------------------------------------
        // The reference's rules (the default), over its AABB tree built on the
        // host (build_ref_tree below).  Node i of the tree: lo = (box minimum,
        // data bits), hi = (box maximum, bits of the node after i's subtree in
        // a positive-child-first walk, -1 after the last); data is the
        // reference's flattened node word (leaf: triangle << 5 | 2 | 1,
        // internal: positive child << 1, negative child next to it).
        // ------------------------------------------------------------------
        struct RefTreeArgs
        {
            const float4 *lo, *hi;
            const float4 *cluC; // per node: area-weighted centroid, radius
            const float4 *cluN; // per node: area-weighted normal sum
        };

        CR_DEV unsigned tree_data(const RefTreeArgs &t, int i) { return __float_as_uint(t.lo[i].w); }

        // evaluateExact (include/geometry_utils/winding_number.h:79-93) in the
        // reference's own FP32 operations: IEEE square roots, no
        // contraction, the math library's atan2f.
        CR_DEV float winding_term_ref(V3 a, V3 b, V3 c, V3 p)
        {
            const float k = float(0.5 / 3.141592653589793238462643383);
            a = a - p;
            b = b - p;
            c = c - p;
            const float la = sqrtf(dot(a, a)), lb = sqrtf(dot(b, b)), lc = sqrtf(dot(c, c));
            const float y = a.x * b.y * c.z - a.x * b.z * c.y - a.y * b.x * c.z + a.y * b.z * c.x + a.z * b.x * c.y -
                            a.z * b.y * c.x;
            const float x = (la * lb * lc + (a.x * b.x + a.y * b.y + a.z * b.z) * lc +
                             (b.x * c.x + b.y * c.y + b.z * c.z) * la + (c.x * a.x + c.y * a.y + c.z * a.z) * lb);
            return k * atan2f(y, x);
        }

        // SdfBuilder::computeSdf's sign (sdf_builder.cpp:148-155):
        // computeWindingNumberApprox with beta = 2 (winding_number.h:128-133,
        // winding_number.cpp:149-171) -- exact terms at the leaves, a
        // cluster's first-order term once the node is farther than two radii
        // from its centroid, summed in the walk's order (positive child
        // first, so the escape links replace the walk's stack) -- and the
        // distance negated where the sum exceeds 0.5.  One thread per node.
        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_sign_tree(MeshSdfArgs a, RefTreeArgs t, float *dist)
        {
            const long long node = (long long)blockIdx.x * kSdfWG + threadIdx.x;
            if (node >= a.count)
                return;
            const int x = (int)(node % a.dx), y = (int)((node / a.dx) % a.dy), z = (int)(node / ((long long)a.dx * a.dy));
            const V3 q = node_point(a, x, y, z);
            float w = 0.0f;
            int i = 0;
            while (i >= 0)
            {
                const unsigned data = tree_data(t, i);
                const int next = __float_as_int(t.hi[i].w);
                if (data & 1u)
                {
                    const int tri = (int)(data >> 5);
                    w += winding_term_ref(v3(a.tri[4 * tri]), v3(a.tri[4 * tri + 1]), v3(a.tri[4 * tri + 2]), q);
                    i = next;
                    continue;
                }
                const float4 cc = t.cluC[i];
                const V3 cen = v3(cc);
                const V3 dq = q - cen;
                const float d2 = dot(dq, dq);
                const float th = 2.0f * cc.w;
                if (d2 > th * th)
                {
                    // ((0.25 / pi) / (l l l)) n . dir: the double constant makes it a double expression
                    const V3 dir = cen - q;
                    const float l = sqrtf(dot(dir, dir));
                    const float ll = l * l * l;
                    w += (float)(((0.25 / 3.141592653589793238462643383) / (double)ll) * (double)dot(v3(t.cluN[i]), dir));
                    i = next;
                    continue;
                }
                i = (int)(data >> 1);
            }
            if (w > 0.5f)
                dist[node] = -dist[node];
        }

        constexpr int kWalkStack = 64;
        // passes of the slicer walk before its chain of warm starts must have
        // settled (a node's walk depends on its predecessor's only where
        // triangles tie, so chains are short)
        constexpr int kWalkPasses = 64;

        CR_DEV float box_dist2(const RefTreeArgs &t, int i, V3 q)
        {
            const V3 cpt = vmax(v3(t.lo[i]), vmin(v3(t.hi[i]), q));
            const V3 d = cpt - q;
            return dot(d, d);
        }

        // SdfBuilder::computeModSdf's walk (sdf_builder.cpp:187-217;
        // ClosestDistanceProjectionToTrimeshTraversalController, aabb_tree.cpp:724-779)
        // for node `node`, warm-started from triangle warm[node - 1] (the
        // previous node's closest triangle in row order; none for node 0):
        // the nearer child first, a subtree pruned when its box is no nearer
        // than the best so far, a leaf accepted when its distance is <= the
        // best.  The on-surface flag falls for good when an accepted
        // candidate's closest point lies on a boundary edge or boundary vertex.
        // Writes the walk's closest triangle, the distance |closest point - q|
        // (0 when the walk accepts nothing past the warm start: the
        // reference's side stays 0) and the flag with the spine override.
        __global__ __launch_bounds__(kSdfWG) void k_mesh_slicer_walk(MeshSdfArgs a, RefTreeArgs t, const int *warm,
                                                                     int *closest, float *dist, unsigned char *onMesh)
        {
            const long long node = (long long)blockIdx.x * kSdfWG + threadIdx.x;
            if (node >= a.count)
                return;
            const int x = (int)(node % a.dx), y = (int)((node / a.dx) % a.dy), z = (int)(node / ((long long)a.dx * a.dy));
            const V3 q = node_point(a, x, y, z);
            float best = kFltMax;
            int bestT = -1;
            V3 bestP = v3(0.0f, 0.0f, 0.0f);
            bool on = true, accepted = false;
            const int w0 = node > 0 ? warm[node - 1] : -1;
            if (w0 >= 0)
            {
                best = tri_closest_unit(q, v3(a.tri[4 * w0]), v3(a.tri[4 * w0 + 1]), v3(a.tri[4 * w0 + 2]), bestP);
                bestT = w0;
            }
            int stack[kWalkStack];
            int sp = 0, i = 0;
            while (true)
            {
                const unsigned data = tree_data(t, i);
                int go = 0; // 0: no deeper, 1: positive child first, 2: negative child first
                if (box_dist2(t, i, q) < best)
                {
                    if (data & 1u)
                    {
                        const int tri = (int)(data >> 5);
                        V3 cp;
                        int region;
                        const float d = tri_closest_unit(q, v3(a.tri[4 * tri]), v3(a.tri[4 * tri + 1]),
                                                         v3(a.tri[4 * tri + 2]), cp, region);
                        if (d <= best)
                        {
                            best = d;
                            bestT = tri;
                            bestP = cp;
                            accepted = true;
                            const int f = a.flags[tri];
                            if (region != kRegFace && ((f >> (region < 3 ? region + 3 : region - 3)) & 1))
                                on = false;
                        }
                    }
                    else
                    {
                        const int pos = (int)(data >> 1);
                        const float dp = box_dist2(t, pos, q), dn = box_dist2(t, pos + 1, q);
                        go = dp < dn ? (dp < best ? 1 : 0) : (dn < best ? 2 : 0);
                    }
                }
                if (go != 0)
                {
                    const int pos = (int)(data >> 1);
                    stack[sp++] = go == 1 ? pos + 1 : pos; // depth checked on the host
                    i = go == 1 ? pos : pos + 1;
                    continue;
                }
                if (sp == 0)
                    break;
                i = stack[--sp];
            }
            closest[node] = bestT;
            dist[node] = accepted ? norm(bestP - q) : 0.0f;
            if (!isnan(a.spine.x))
            {
                V3 n = v3(a.spine.x, a.spine.y, a.spine.z);
                const float s2 = dot(n, n);
                n = s2 > 0.0f ? n * rsq(s2) : v3(0.0f, 0.0f, 0.0f);
                if (dot(q, n) + a.spine.w > 0.0f)
                    on = true;
            }
            onMesh[node] = on ? 1 : 0;
        }

        // *changed = 1 when some node's closest triangle differs from the warm
        // start the walk gave the next node.
        __global__ __launch_bounds__(kSdfWG) void k_mesh_slicer_changed(long long count, const int *warm,
                                                                        const int *closest, int *changed)
        {
            const long long node = (long long)blockIdx.x * kSdfWG + threadIdx.x;
            if (node + 1 < count && warm[node] != closest[node])
                *changed = 1;
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

        // onMesh (mesh-slicer table): rows off the surface become NaN after the
        // gradient is taken from all distances (postProcessModSdfGradient,
        // sdf_builder.cpp:332-357).
        __global__ __launch_bounds__(kSdfWG) void k_mesh_sdf_gradient(int dx, int dy, int dz, long long count,
                                                                     float cell, const float *dist,
                                                                     const unsigned char *onMesh, float4 *out)
        {
            const long long node = (long long)blockIdx.x * kSdfWG + threadIdx.x;
            if (node >= count)
                return;
            if (onMesh && !onMesh[node])
            {
                const float nan = __int_as_float(0x7fc00000);
                out[node] = make_float4(nan, nan, nan, nan);
                return;
            }
            const int x = (int)(node % dx), y = (int)((node / dx) % dy), z = (int)(node / ((long long)dx * dy));
            const float inv = 1.0f / cell;
            V3 g = v3(sdf_diff(x, dx, dist, node, 1, inv), sdf_diff(y, dy, dist, node, dx, inv),
                      sdf_diff(z, dz, dist, node, (long long)dx * dy, inv));
            normalize(g);
            out[node] = make_float4(g.x, g.y, g.z, dist[node]);
        }

        // Per triangle: boundary edges ab, bc, ac (bits 0-2) and boundary
        // vertices a, b, c (bits 3-5); a boundary edge has one adjacent
        // triangle, a boundary vertex ends one (buildEdgeTriangleMap /
        // buildEdgeVerticesList, sdf_builder.cpp:45-82).
        std::vector<int> boundary_flags(const unsigned *tris, int numTris)
        {
            auto key = [](unsigned u, unsigned v) {
                return (unsigned long long)std::min(u, v) << 32 | std::max(u, v);
            };
            std::vector<unsigned long long> edges(3 * (size_t)numTris);
            for (size_t t = 0; t < (size_t)numTris; ++t)
                for (int j = 0; j < 3; ++j)
                    edges[3 * t + j] = key(tris[3 * t + j], tris[3 * t + (j + 1) % 3]);
            std::vector<unsigned long long> sorted(edges);
            std::sort(sorted.begin(), sorted.end());
            auto single = [&](unsigned long long e) {
                auto r = std::equal_range(sorted.begin(), sorted.end(), e);
                return r.second - r.first < 2;
            };
            std::vector<unsigned> bverts;
            for (size_t i = 0; i < sorted.size(); ++i)
                if ((i == 0 || sorted[i] != sorted[i - 1]) && single(sorted[i]))
                {
                    bverts.push_back((unsigned)(sorted[i] >> 32));
                    bverts.push_back((unsigned)(sorted[i] & 0xffffffffu));
                }
            std::sort(bverts.begin(), bverts.end());
            std::vector<int> flags(numTris, 0);
            for (size_t t = 0; t < (size_t)numTris; ++t)
            {
                const unsigned *v = tris + 3 * t;
                int f = 0;
                // edges in the order ab, bc, ca: bit 2 is edge ac
                for (int j = 0; j < 3; ++j)
                    if (single(edges[3 * t + j]))
                        f |= 1 << j;
                for (int j = 0; j < 3; ++j)
                    if (std::binary_search(bverts.begin(), bverts.end(), v[j]))
                        f |= 8 << j;
                flags[t] = f;
            }
            return flags;
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
    // A triangle's box starts as the zero-initialised Bounds3 and grows to the
    // corners, so the union always holds the origin.
    float lo[3] = {kFltMax, kFltMax, kFltMax}, hi[3] = {-kFltMax, -kFltMax, -kFltMax};
    for (long long t = 0; t < numTriangles; ++t)
    {
        float bl[3] = {0.0f, 0.0f, 0.0f}, bh[3] = {0.0f, 0.0f, 0.0f};
        for (int j = 0; j < 3; ++j)
        {
            const unsigned i = triangles[3 * t + j];
            if (i >= (unsigned)numPoints)
                return CRMPM_HIP_ERR_ARG;
            for (int k = 0; k < 3; ++k)
            {
                const float v = points[3ll * i + k];
                if (!std::isfinite(v))
                    return CRMPM_HIP_ERR_ARG;
                bl[k] = std::fmin(bl[k], v);
                bh[k] = std::fmax(bh[k], v);
            }
        }
        for (int k = 0; k < 3; ++k)
        {
            lo[k] = std::fmin(lo[k], bl[k] - 1e-4f);
            hi[k] = std::fmax(hi[k], bh[k] + 1e-4f);
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

} // extern "C"

namespace
{
    // The reference's AABB tree over the triangles and the cluster data of its
    // winding-number approximation, as SdfBuilder::precomputeSdfDesc builds
    // them (buildTree, include/geometry_utils/winding_number.h:148-161;
    // buildAABBTree / flattenTree, src/geometry_utils/aabb_tree.cpp:97-333;
    // precomputeClusterInformation, src/geometry_utils/winding_number.cpp:44-124),
    // in the device layout of RefTreeArgs.  FP32 host arithmetic in the
    // reference's order of operations, so the tree, its boxes and the
    // clusters are the reference's.
    struct RefTree
    {
        std::vector<float4> lo, hi, cluC, cluN;
        int height = 0; // longest root-to-leaf path, in edges
    };

    struct H3
    {
        float x, y, z;
        float operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
    };
    H3 operator+(H3 a, H3 b) { return H3{a.x + b.x, a.y + b.y, a.z + b.z}; }
    H3 operator-(H3 a, H3 b) { return H3{a.x - b.x, a.y - b.y, a.z - b.z}; }
    H3 operator*(H3 a, H3 b) { return H3{a.x * b.x, a.y * b.y, a.z * b.z}; }
    H3 operator*(H3 a, float s) { return H3{a.x * s, a.y * s, a.z * s}; }
    H3 operator/(H3 a, float s) { return H3{a.x / s, a.y / s, a.z / s}; }
    H3 hmin(H3 a, H3 b) { return H3{std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
    H3 hmax(H3 a, H3 b) { return H3{std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
    float hsqn(H3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

    RefTree build_ref_tree(const float *points, const unsigned *tris, int numTris)
    {
        auto P = [&](unsigned i) { return H3{points[3 * (size_t)i], points[3 * (size_t)i + 1], points[3 * (size_t)i + 2]}; };
        // triangle boxes: the zero box grown to the corners, fattened by 1e-4; centres
        std::vector<H3> blo(numTris), bhi(numTris), centre(numTris);
        for (int t = 0; t < numTris; ++t)
        {
            H3 l{0.0f, 0.0f, 0.0f}, h{0.0f, 0.0f, 0.0f};
            for (int j = 0; j < 3; ++j)
            {
                const H3 p = P(tris[3 * (size_t)t + j]);
                l = hmin(l, p);
                h = hmax(h, p);
            }
            blo[t] = H3{l.x - 1e-4f, l.y - 1e-4f, l.z - 1e-4f};
            bhi[t] = H3{h.x + 1e-4f, h.y + 1e-4f, h.z + 1e-4f};
            centre[t] = (bhi[t] + blo[t]) * 0.5f;
        }
        // depth-first build, one triangle per leaf: split at the box centre
        // along the axis of largest centre variance, positive side first in
        // the index list, 50-50 when one side is empty; children in pairs
        struct BN
        {
            H3 lo, hi;
            int pos = -1, depth = 0;
            unsigned first = 0, count = 0;
        };
        std::vector<unsigned> idx(numTris);
        for (int i = 0; i < numTris; ++i)
            idx[i] = (unsigned)i;
        std::vector<BN> pool;
        pool.reserve(2 * (size_t)numTris);
        pool.push_back(BN{});
        pool[0].count = (unsigned)numTris;
        std::vector<int> stack{0};
        RefTree tree;
        while (!stack.empty())
        {
            const int ni = stack.back();
            stack.pop_back();
            unsigned *prims = idx.data() + pool[ni].first;
            const unsigned nb = pool[ni].count;
            H3 lo = blo[prims[0]], hi = bhi[prims[0]], mean = centre[prims[0]];
            for (unsigned i = 1; i < nb; ++i)
            {
                mean = mean + centre[prims[i]];
                lo = hmin(lo, blo[prims[i]]);
                hi = hmax(hi, bhi[prims[i]]);
            }
            pool[ni].lo = lo;
            pool[ni].hi = hi;
            tree.height = std::max(tree.height, pool[ni].depth);
            if (nb <= 1)
                continue;
            mean = mean * (1.0f / float(nb));
            H3 var{0.0f, 0.0f, 0.0f};
            for (unsigned i = 0; i < nb; ++i)
            {
                H3 c = centre[prims[i]] - mean;
                c = c * c;
                var = var + c;
            }
            var = var * (1.0f / float(nb - 1));
            const float ax = std::fabs(var.x), ay = std::fabs(var.y), az = std::fabs(var.z);
            const int axis = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
            const float split = (lo[axis] + hi[axis]) * 0.5f;
            unsigned nPos = 0;
            for (unsigned i = 0; i < nb; ++i)
            {
                const unsigned p = prims[i];
                if (centre[p][axis] > split)
                {
                    prims[i] = prims[nPos];
                    prims[nPos] = p;
                    ++nPos;
                }
            }
            if (nPos == 0 || nPos == nb)
                nPos = nb >> 1;
            const int c0 = (int)pool.size();
            pool.push_back(BN{});
            pool.push_back(BN{});
            pool[ni].pos = c0;
            for (int k = 0; k < 2; ++k)
                pool[c0 + k].depth = pool[ni].depth + 1;
            pool[c0].first = pool[ni].first;
            pool[c0].count = nPos;
            pool[c0 + 1].first = pool[ni].first + nPos;
            pool[c0 + 1].count = nb - nPos;
            stack.push_back(c0 + 1);
            stack.push_back(c0);
        }
        const size_t n = pool.size();
        std::vector<unsigned> data(n);
        std::vector<int> next(n, -1);
        for (size_t i = 0; i < n; ++i)
            data[i] = pool[i].pos < 0 ? (idx[pool[i].first] << 5) | ((pool[i].count & 15u) << 1) | 1u
                                      : (unsigned)pool[i].pos << 1;
        // escape links of the positive-child-first walk
        for (size_t i = 0; i < n; ++i)
            if (pool[i].pos >= 0)
            {
                next[pool[i].pos] = pool[i].pos + 1;
                next[pool[i].pos + 1] = next[i];
            }
        // children come after their parent in the pool: the loop above sees a
        // parent's link before setting its children's
        tree.lo.resize(n);
        tree.hi.resize(n);
        for (size_t i = 0; i < n; ++i)
        {
            tree.lo[i] = make_float4(pool[i].lo.x, pool[i].lo.y, pool[i].lo.z, __builtin_bit_cast(float, data[i]));
            tree.hi[i] = make_float4(pool[i].hi.x, pool[i].hi.y, pool[i].hi.z, __builtin_bit_cast(float, next[i]));
        }
        // clusters: per triangle area-weighted normal, area and centroid; each
        // internal node's triangles in the order of a negative-child-first walk
        std::vector<float> area(numTris);
        std::vector<H3> nrmA(numTris), cen(numTris);
        for (int t = 0; t < numTris; ++t)
        {
            const H3 a = P(tris[3 * (size_t)t]), b = P(tris[3 * (size_t)t + 1]), c = P(tris[3 * (size_t)t + 2]);
            const H3 u = b - a, v = c - a;
            nrmA[t] = H3{u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x} * 0.5f;
            area[t] = std::sqrt(hsqn(nrmA[t]));
            cen[t] = (a + b + c) * float(1.0 / 3.0);
        }
        tree.cluC.assign(n, make_float4(0.f, 0.f, 0.f, 0.f));
        tree.cluN.assign(n, make_float4(0.f, 0.f, 0.f, 0.f));
        std::vector<int> todo{0}, order;
        std::vector<std::pair<int, int>> sections;
        while (!todo.empty())
        {
            int id = todo.back();
            todo.pop_back();
            if (id >= 0)
            {
                if (data[id] & 1u)
                {
                    order.push_back((int)(data[id] >> 5));
                    sections.push_back({(int)order.size() - 1, (int)order.size()});
                    continue;
                }
                const int pos = (int)(data[id] >> 1);
                todo.push_back(-id - 1);
                todo.push_back(pos);
                todo.push_back(pos + 1);
                continue;
            }
            const std::pair<int, int> sa = sections.back();
            sections.pop_back();
            const std::pair<int, int> sb = sections.back();
            sections.pop_back();
            const std::pair<int, int> sum{sb.first, sa.second};
            id = -id - 1;
            H3 wc{0.0f, 0.0f, 0.0f}, wn{0.0f, 0.0f, 0.0f};
            float as = 0.0f;
            for (int i = sum.first; i < sum.second; ++i)
            {
                const int t = order[i];
                as += area[t];
                wc = wc + cen[t] * area[t];
                wn = wn + nrmA[t];
            }
            wc = wc / as;
            float r2 = 0.0f;
            for (int i = sum.first; i < sum.second; ++i)
                for (int j = 0; j < 3; ++j)
                {
                    const float d2 = hsqn(wc - P(tris[3 * (size_t)order[i] + j]));
                    if (d2 > r2)
                        r2 = d2;
                }
            tree.cluC[id] = make_float4(wc.x, wc.y, wc.z, std::sqrt(r2));
            tree.cluN[id] = make_float4(wn.x, wn.y, wn.z, as);
            sections.push_back(sum);
        }
        return tree;
    }

    // Both tables: the triangle-mesh SDF (slicer false: signed distance,
    // rows at out) and the mesh-slicer table (spine row at out, then the
    // grid).  Synchronous on stream.
    int build_mesh_table(bool slicer, int rule, const float *points, int numPoints, const unsigned *triangles,
                         int numTriangles, float cellSize, float fatten, const float *spine, int dims[3],
                         float lowerBound[3], float *out, unsigned long long capacity, void *stream)
    {
        if (rule != CRMPM_HIP_MESH_RULE_REFERENCE && rule != CRMPM_HIP_MESH_RULE_EXACT)
            return CRMPM_HIP_ERR_ARG;
        int rc = crHipMeshSdfDesc(points, numPoints, triangles, numTriangles, cellSize, fatten, dims, lowerBound);
        if (rc || !out)
            return rc;
        const bool refRule = rule == CRMPM_HIP_MESH_RULE_REFERENCE;
        // one node along an axis leaves the reference's one-sided difference
        // reading outside the table; such tables are refused
        if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2)
            return CRMPM_HIP_ERR_ARG;
        const long long count = (long long)dims[0] * dims[1] * dims[2];
        if (capacity < (unsigned long long)count + (slicer ? 1 : 0))
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
        if (slicer)
            a.spine = make_float4(spine[0], spine[1], spine[2], spine[3]);
        const int chunks = (numTriangles + kSdfChunk - 1) / kSdfChunk;
        const int slices = std::max(1, std::min(chunks, (kSdfTargetWGs + a.groups - 1) / a.groups));
        a.sliceTris = ((numTriangles + slices - 1) / slices + 63) / 64 * 64;
        const int usedSlices = (numTriangles + a.sliceTris - 1) / a.sliceTris;

        // triangle records gathered on the host: corners with the box's lower
        // corner in .w, then the upper corner and the culling slack (1e-5 of
        // the squared largest extent); centroids apart
        std::vector<float4> tri(4 * (size_t)numTriangles), cen(numTriangles);
        for (size_t t = 0; t < (size_t)numTriangles; ++t)
        {
            const float *v[3];
            for (int k = 0; k < 3; ++k)
                v[k] = points + 3 * (size_t)triangles[3 * t + k];
            float lo[3], hi[3], c[3], ext = 0.0f;
            for (int k = 0; k < 3; ++k)
            {
                lo[k] = std::min(v[0][k], std::min(v[1][k], v[2][k]));
                hi[k] = std::max(v[0][k], std::max(v[1][k], v[2][k]));
                c[k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0f;
                ext = std::max(ext, hi[k] - lo[k]);
            }
            for (int k = 0; k < 3; ++k)
                tri[4 * t + k] = make_float4(v[k][0], v[k][1], v[k][2], lo[k]);
            tri[4 * t + 3] = make_float4(hi[0], hi[1], hi[2], 1e-5f * ext * ext);
            cen[t] = make_float4(c[0], c[1], c[2], 0.0f);
        }
        // clusters of kClu consecutive triangles: centre and radius of their
        // box, vector area and area (for the far-field winding term and its
        // error bound, both pre-divided by 4 pi), the box and its cull slack
        const int numClu = (numTriangles + kClu - 1) / kClu;
        std::vector<float4> clu(4 * (size_t)numClu);
        const double k4pi = 1.0 / (4.0 * 3.141592653589793238462643383);
        for (int k = 0; k < numClu; ++k)
        {
            double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, N[3] = {0, 0, 0}, A = 0;
            for (int t = k * kClu; t < std::min(numTriangles, (k + 1) * kClu); ++t)
            {
                double v[3][3];
                for (int j = 0; j < 3; ++j)
                    for (int c = 0; c < 3; ++c)
                    {
                        v[j][c] = points[3 * (size_t)triangles[3 * (size_t)t + j] + c];
                        lo[c] = std::min(lo[c], v[j][c]);
                        hi[c] = std::max(hi[c], v[j][c]);
                    }
                const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
                const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
                const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                      e1[0] * e2[1] - e1[1] * e2[0]};
                for (int c = 0; c < 3; ++c)
                    N[c] += 0.5 * cr[c];
                A += 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            }
            const double rho = 0.5 * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                               (hi[2] - lo[2]) * (hi[2] - lo[2])) * (1.0 + 1e-6) + 1e-12;
            const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
            clu[4 * k] = make_float4((float)(0.5 * (lo[0] + hi[0])), (float)(0.5 * (lo[1] + hi[1])),
                                     (float)(0.5 * (lo[2] + hi[2])), (float)rho);
            clu[4 * k + 1] = make_float4((float)(N[0] * k4pi), (float)(N[1] * k4pi), (float)(N[2] * k4pi),
                                         (float)(2.0 * rho * A * k4pi * (1.0 + 1e-5)));
            clu[4 * k + 2] = make_float4((float)lo[0], (float)lo[1], (float)lo[2], (float)(1e-5 * ext * ext));
            clu[4 * k + 3] = make_float4((float)hi[0], (float)hi[1], (float)hi[2], 0.0f);
        }
        std::vector<int> flags;
        if (slicer)
            flags = boundary_flags(triangles, numTriangles);
        RefTree tree;
        if (refRule)
        {
            tree = build_ref_tree(points, triangles, numTriangles);
            // the slicer walk keeps the deferred subtrees of one root-to-leaf path
            if (slicer && tree.height > kWalkStack)
                return CRMPM_HIP_ERR_CAPACITY;
        }
        const float spineRow[4] = {slicer ? spine[0] : 0.0f, slicer ? spine[1] : 0.0f, slicer ? spine[2] : 0.0f,
                                   slicer ? spine[3] : 0.0f};

        hipStream_t st = static_cast<hipStream_t>(stream);
        float4 *dTri = nullptr, *dCen = nullptr, *dClu = nullptr, *dPart = nullptr;
        float *dWn = nullptr, *dErr = nullptr, *dDist = nullptr;
        int *dFlags = nullptr;
        unsigned char *dMask = nullptr; // per node: slicer on-surface flag, or SDF sign left undecided
        float4 *dTree = nullptr;        // reference rule: node lo, node hi, cluster centroid, cluster normal
        int *dWalk = nullptr;           // reference slicer: two closest-triangle arrays and a flag
        const size_t tn = tree.lo.size();
        int walkRc = CRMPM_HIP_OK;
        rc = CRMPM_HIP_ERR_ALLOC;
        if (hipMalloc(&dTri, tri.size() * sizeof(float4)) == hipSuccess &&
            hipMalloc(&dCen, cen.size() * sizeof(float4)) == hipSuccess &&
            hipMalloc(&dClu, clu.size() * sizeof(float4)) == hipSuccess &&
            hipMalloc(&dErr, (size_t)usedSlices * count * sizeof(float)) == hipSuccess &&
            hipMalloc(&dPart, (size_t)usedSlices * count * sizeof(float4)) == hipSuccess &&
            hipMalloc(&dWn, (size_t)usedSlices * count * sizeof(float)) == hipSuccess &&
            hipMalloc(&dDist, (size_t)count * sizeof(float)) == hipSuccess &&
            hipMalloc(&dMask, (size_t)count) == hipSuccess &&
            (!slicer || hipMalloc(&dFlags, flags.size() * sizeof(int)) == hipSuccess) &&
            (!refRule || hipMalloc(&dTree, 4 * tn * sizeof(float4)) == hipSuccess) &&
            (!(refRule && slicer) || hipMalloc(&dWalk, (2 * (size_t)count + 1) * sizeof(int)) == hipSuccess))
        {
            rc = CRMPM_HIP_ERR_LAUNCH;
            a.tri = dTri;
            a.cen = dCen;
            a.clu = dClu;
            a.partErr = dErr;
            a.part = dPart;
            a.partWn = dWn;
            a.flags = dFlags;
            const int nodeWGs = (int)((count + kSdfWG - 1) / kSdfWG);
            float4 *rows = reinterpret_cast<float4 *>(out) + (slicer ? 1 : 0);
            bool ok = hipMemcpyAsync(dTri, tri.data(), tri.size() * sizeof(float4), hipMemcpyHostToDevice, st) ==
                          hipSuccess &&
                      hipMemcpyAsync(dCen, cen.data(), cen.size() * sizeof(float4), hipMemcpyHostToDevice, st) ==
                          hipSuccess &&
                      hipMemcpyAsync(dClu, clu.data(), clu.size() * sizeof(float4), hipMemcpyHostToDevice, st) ==
                          hipSuccess &&
                      (!slicer ||
                       (hipMemcpyAsync(dFlags, flags.data(), flags.size() * sizeof(int), hipMemcpyHostToDevice, st) ==
                            hipSuccess &&
                        hipMemcpyAsync(out, spineRow, sizeof(spineRow), hipMemcpyHostToDevice, st) == hipSuccess));
            RefTreeArgs ta{};
            if (ok && refRule)
            {
                ta = RefTreeArgs{dTree, dTree + tn, dTree + 2 * tn, dTree + 3 * tn};
                const std::vector<float4> *parts[4] = {&tree.lo, &tree.hi, &tree.cluC, &tree.cluN};
                for (int k = 0; k < 4 && ok; ++k)
                    ok = hipMemcpyAsync(dTree + k * tn, parts[k]->data(), tn * sizeof(float4), hipMemcpyHostToDevice,
                                        st) == hipSuccess;
            }
            if (ok && refRule && slicer)
            {
                // the reference's warm-started walk, node after node: run it for
                // all nodes from the previous pass's closest triangles (none at
                // first) until every node's warm start is its predecessor's result
                int *warm = dWalk, *closest = dWalk + count, *changed = dWalk + 2 * count;
                ok = hipMemsetAsync(warm, 0xff, (size_t)count * sizeof(int), st) == hipSuccess;
                int pass = 0;
                for (; ok && pass < kWalkPasses; ++pass)
                {
                    int flag = 0;
                    k_mesh_slicer_walk<<<nodeWGs, kSdfWG, 0, st>>>(a, ta, warm, closest, dDist, dMask);
                    ok = hipMemsetAsync(changed, 0, sizeof(int), st) == hipSuccess;
                    k_mesh_slicer_changed<<<nodeWGs, kSdfWG, 0, st>>>(count, warm, closest, changed);
                    ok = ok && hipGetLastError() == hipSuccess &&
                         hipMemcpyAsync(&flag, changed, sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
                         hipStreamSynchronize(st) == hipSuccess;
                    if (!ok || !flag)
                        break;
                    std::swap(warm, closest);
                }
                if (ok && pass == kWalkPasses)
                    walkRc = CRMPM_HIP_ERR_STATE; // the chain of warm starts did not settle
            }
            if (ok)
            {
                if (slicer && refRule)
                {
                    // distances and flags come from the walk
                }
                else if (slicer)
                {
                    k_mesh_sdf_slices<true><<<a.groups * usedSlices, kSdfWG, 0, st>>>(a);
                    k_mesh_sdf_combine<true><<<nodeWGs, kSdfWG, 0, st>>>(a, usedSlices, dDist, dMask);
                }
                else if (refRule)
                {
                    // closest distances from the slices, sign from the reference's winding walk
                    k_mesh_sdf_slices<true><<<a.groups * usedSlices, kSdfWG, 0, st>>>(a);
                    k_mesh_sdf_combine<false, true><<<nodeWGs, kSdfWG, 0, st>>>(a, usedSlices, dDist, dMask);
                    k_mesh_sdf_sign_tree<<<nodeWGs, kSdfWG, 0, st>>>(a, ta, dDist);
                }
                else
                {
                    k_mesh_sdf_slices<false><<<a.groups * usedSlices, kSdfWG, 0, st>>>(a);
                    k_mesh_sdf_combine<false><<<nodeWGs, kSdfWG, 0, st>>>(a, usedSlices, dDist, dMask);
                    k_mesh_sdf_exact<<<a.groups, kSdfWG, 0, st>>>(a, dMask, dDist);
                }
                // dMask: the slicer's on-surface flags; for the SDF, the undecided nodes (done with)
                k_mesh_sdf_gradient<<<nodeWGs, kSdfWG, 0, st>>>(a.dx, a.dy, a.dz, count, cellSize, dDist,
                                                                slicer ? dMask : nullptr, rows);
                ok = hipGetLastError() == hipSuccess;
            }
            // the copies read host memory: wait for them whatever failed
            if (hipStreamSynchronize(st) == hipSuccess && ok)
                rc = walkRc;
        }
        (void)hipFree(dTri);
        (void)hipFree(dCen);
        (void)hipFree(dClu);
        (void)hipFree(dErr);
        (void)hipFree(dPart);
        (void)hipFree(dWn);
        (void)hipFree(dDist);
        (void)hipFree(dFlags);
        (void)hipFree(dMask);
        (void)hipFree(dTree);
        (void)hipFree(dWalk);
        return rc;
    }
} // namespace

extern "C" {

int crHipMeshSdf(const float *points, int numPoints, const unsigned int *triangles, int numTriangles, float cellSize,
                 float fatten, int dims[3], float lowerBound[3], float *gradientSignedDistance,
                 unsigned long long capacity, void *stream)
{
    return build_mesh_table(false, CRMPM_HIP_MESH_RULE_REFERENCE, points, numPoints, triangles, numTriangles, cellSize,
                            fatten, nullptr, dims, lowerBound, gradientSignedDistance, capacity, stream);
}

int crHipMeshSdfWithRule(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                         float cellSize, float fatten, int rule, int dims[3], float lowerBound[3],
                         float *gradientSignedDistance, unsigned long long capacity, void *stream)
{
    return build_mesh_table(false, rule, points, numPoints, triangles, numTriangles, cellSize, fatten, nullptr, dims,
                            lowerBound, gradientSignedDistance, capacity, stream);
}

int crHipMeshSlicerSdf(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                       float cellSize, float fatten, const float spineArea[4], int dims[3], float lowerBound[3],
                       float *rows, unsigned long long capacity, void *stream)
{
    if (!spineArea)
        return CRMPM_HIP_ERR_ARG;
    return build_mesh_table(true, CRMPM_HIP_MESH_RULE_REFERENCE, points, numPoints, triangles, numTriangles, cellSize,
                            fatten, spineArea, dims, lowerBound, rows, capacity, stream);
}

int crHipMeshSlicerSdfWithRule(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                               float cellSize, float fatten, const float spineArea[4], int rule, int dims[3],
                               float lowerBound[3], float *rows, unsigned long long capacity, void *stream)
{
    if (!spineArea)
        return CRMPM_HIP_ERR_ARG;
    return build_mesh_table(true, rule, points, numPoints, triangles, numTriangles, cellSize, fatten, spineArea, dims,
                            lowerBound, rows, capacity, stream);
}

} // extern "C"
