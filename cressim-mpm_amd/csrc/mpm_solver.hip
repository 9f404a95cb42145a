// MPM substep solver for MI355X (gfx950): kernels and the C-ABI of include/crmpm_hip.h.
//
// Data layout in HBM
//   engine-owned (bound):  positionMass float4[N], velocity float4[N], params0 float4[N],
//                          type int[N], mask u8[N], shape/geometry/SDF tables
//   solver-owned:          F, C  9 x float planes [9][stride] (column-major element k),
//                          V0 float[N]
//   grid:                  4x4x4-node blocks, 1 KiB each (float4 (momentum|velocity, mass) per
//                          node), block b = bx + nbx * (by + nby * bz), node (lx,ly,lz) at
//                          b * 64 + lz * 16 + ly * 4 + lx
//   per step:              pblock key per particle, stable LSD radix sort (key, index),
//                          one "tile" per non-empty pblock: a 4^3-cell block whose particles
//                          touch the 6^3 nodes [4b, 4b + 6)
//
// Substep (MLS):  prepare -> shape cache -> bin -> radix sort -> segments -> mark active node
//                 blocks -> P2G (one wavefront per tile, LDS accumulation, tile written to a
//                 partial buffer) -> update (one wavefront per active node block: sums the
//                 <= 8 overlapping tiles in a fixed order, then cutoff / contact / walls)
//                 -> G2P (one wavefront per tile, 6^3 nodes staged in LDS).
// P2G never issues global atomics and every sum has a fixed order, so a substep is bitwise
// reproducible run to run.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "crmpm_hip.h"
#include "mpm_device.hpp"

using namespace crhip;

namespace
{
    constexpr int kTB = 4;            // cells per tile edge
    constexpr int kTN = kTB + 2;      // nodes per tile edge
    constexpr int kTileNodes = kTN * kTN * kTN;
    constexpr int kWG = 256;          // threads per workgroup
    constexpr int kWaves = kWG / 64;  // tiles per workgroup
    constexpr int kRadixItems = 8;
    constexpr int kRadixTile = kWG * kRadixItems;

    struct GridGeom
    {
        float3 bmin, bmax;
        float dx, invdx, cellVol;
        int nx, ny, nz;
        int nbx, nby, nbz, numBlocks;
    };

    struct ParticleBind
    {
        float4 *xm;
        float4 *v;
        const float4 *params;
        const int *type;
        const unsigned char *mask;
    };

    struct SolverState
    {
        float *F;   // [9][stride]
        float *C;   // [9][stride]
        float *V0;
        int stride;
    };

    CR_DEV int block_key(const GridGeom &g, int bx, int by, int bz) { return bx + g.nbx * (by + g.nby * bz); }

    CR_DEV int grid_index(const GridGeom &g, int x, int y, int z)
    {
        const int b = block_key(g, x >> 2, y >> 2, z >> 2);
        return b * 64 + ((z & 3) << 4) + ((y & 3) << 2) + (x & 3);
    }

    CR_DEV bool in_grid(const GridGeom &g, int x, int y, int z)
    {
        return !(x < 0 || x >= g.nx || y < 0 || y >= g.ny || z < 0 || z >= g.nz);
    }

    // Base cell (truncation toward zero) and fraction (src/solver/mls_mpm_algorithm.h:172-174)
    CR_DEV void cell_of(const GridGeom &g, V3 x, int &bx, int &by, int &bz, V3 &fx)
    {
        const V3 s = (x - v3(g.bmin.x, g.bmin.y, g.bmin.z)) * g.invdx;
        bx = (int)(s.x - 0.5f);
        by = (int)(s.y - 0.5f);
        bz = (int)(s.z - 0.5f);
        fx = V3{s.x - (float)bx, s.y - (float)by, s.z - (float)bz};
    }

    // pblock of a base cell; clamped so every in-range node of the particle
    // lies inside the tile's node box [4b, 4b + 6).
    CR_DEV int pblock_key(const GridGeom &g, int bx, int by, int bz)
    {
        bx = min(max(bx, 0), g.nx - 1) >> 2;
        by = min(max(by, 0), g.ny - 1) >> 2;
        bz = min(max(bz, 0), g.nz - 1) >> 2;
        return block_key(g, bx, by, bz);
    }

    CR_DEV void decode_block(const GridGeom &g, int key, int &bx, int &by, int &bz)
    {
        bx = key % g.nbx;
        by = (key / g.nbx) % g.nby;
        bz = key / (g.nbx * g.nby);
    }

    CR_DEV M3 load_m3(const float *plane, int stride, int p)
    {
        return M3{V3{plane[0 * stride + p], plane[1 * stride + p], plane[2 * stride + p]},
                  V3{plane[3 * stride + p], plane[4 * stride + p], plane[5 * stride + p]},
                  V3{plane[6 * stride + p], plane[7 * stride + p], plane[8 * stride + p]}};
    }
    CR_DEV void store_m3(float *plane, int stride, int p, const M3 &m)
    {
        plane[0 * stride + p] = m.c0.x;
        plane[1 * stride + p] = m.c0.y;
        plane[2 * stride + p] = m.c0.z;
        plane[3 * stride + p] = m.c1.x;
        plane[4 * stride + p] = m.c1.y;
        plane[5 * stride + p] = m.c1.z;
        plane[6 * stride + p] = m.c2.x;
        plane[7 * stride + p] = m.c2.y;
        plane[8 * stride + p] = m.c2.z;
    }

    CR_DEV void wave_sync()
    {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }

    struct Weights
    {
        float w[3][3]; // w[i][axis]
    };
    CR_DEV void weights_of(V3 fx, Weights &W)
    {
        bspline_w(fx.x, W.w[0][0], W.w[1][0], W.w[2][0]);
        bspline_w(fx.y, W.w[0][1], W.w[1][1], W.w[2][1]);
        bspline_w(fx.z, W.w[0][2], W.w[1][2], W.w[2][2]);
    }

    // ------------------------------------------------------------------
    // Per-step bookkeeping
    // ------------------------------------------------------------------

    // Copy one bound shape into a compact record (shapeIds order).
    CR_DEV void cache_shape(int s, const int *shapeIds, const CrHipShapeData &sd, const CrHipGeometryData &gd, ShapeRec *recs)
    {
        const int id = shapeIds[s];
        const int gi = sd.geometryIdx[id];
        ShapeRec r;
        r.type = sd.type[id];
        r.geomType = gd.type[gi];
        r.shapeId = id;
        r.pad = 0;
        const float4 *f4 = nullptr;
        r.pos = reinterpret_cast<const float4 *>(sd.position)[id];
        r.rot = reinterpret_cast<const float4 *>(sd.rotation)[id];
        r.invScale = reinterpret_cast<const float4 *>(sd.invScale)[id];
        r.linVel = reinterpret_cast<const float4 *>(sd.linearVelocity)[id];
        r.angVel = reinterpret_cast<const float4 *>(sd.angularVelocity)[id];
        r.params0 = reinterpret_cast<const float4 *>(sd.params0)[id];
        r.gparams = reinterpret_cast<const float4 *>(gd.params0)[gi];
        f4 = reinterpret_cast<const float4 *>(sd.comInvMass);
        r.comInvMass = f4 ? f4[id] : make_float4(0.f, 0.f, 0.f, 0.f);
        f4 = reinterpret_cast<const float4 *>(sd.inertiaInv0);
        r.inertia0 = f4 ? f4[id] : make_float4(0.f, 0.f, 0.f, 0.f);
        f4 = reinterpret_cast<const float4 *>(sd.inertiaInv1);
        r.inertia1 = f4 ? f4[id] : make_float4(0.f, 0.f, 0.f, 0.f);
        recs[s] = r;
    }

    struct ShapeTables
    {
        int numShapes;
        const int *shapeIds;
        CrHipShapeData sd;
        CrHipGeometryData gd;
        ShapeRec *recs;
        int *counts;
        int *slicers;
    };

    // Start of a step.  Workgroups [0, blockWGs): zero the grid blocks written
    // since the last step and mark every pblock empty (one wavefront per
    // block).  The workgroups after them copy the bound shape tables into a
    // compact record array in shapeIds order; the first of those also lists
    // the slicers (shapes whose geometry is a slicer).
    __global__ __launch_bounds__(kWG) void k_prepare(GridGeom g, int blockWGs, float4 *grid, float4 *gridForce,
                                                     unsigned char *dirty, int *segStart, int *segEnd, ShapeTables st)
    {
        const int lane = threadIdx.x & 63;
        if ((int)blockIdx.x >= blockWGs)
        {
            const int sw = blockIdx.x - blockWGs;
            const int s = sw * kWG + threadIdx.x;
            if (s < st.numShapes)
                cache_shape(s, st.shapeIds, st.sd, st.gd, st.recs);
            if (sw == 0 && threadIdx.x == 0)
            {
                int ns = 0;
                for (int k = 0; k < st.numShapes; ++k)
                {
                    const int gt = st.gd.type[st.sd.geometryIdx[st.shapeIds[k]]];
                    if (gt == gQuadSlicer || gt == gTriangleMeshSlicer)
                        st.slicers[ns++] = k;
                }
                st.counts[0] = st.numShapes;
                st.counts[1] = ns;
            }
            return;
        }
        const int b = blockIdx.x * kWaves + (threadIdx.x >> 6);
        if (b >= g.numBlocks)
            return;
        if (dirty[b])
        {
            grid[b * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gridForce)
                gridForce[b * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        wave_sync();
        if (lane == 0)
        {
            dirty[b] = 0;
            segStart[b] = 0;
            segEnd[b] = 0;
        }
    }

    // ---------------- stable LSD radix sort, 8-bit digits ----------------
    CR_DEV unsigned wave_inclusive_scan(unsigned v, int lane)
    {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const unsigned t = __shfl_up(v, o, 64);
            if (lane >= o)
                v += t;
        }
        return v;
    }

    // Exclusive scan of one value per thread over a 256-thread workgroup.
    CR_DEV unsigned block_exclusive_scan(unsigned v, unsigned &total, unsigned *sh /* >= 4 */)
    {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const unsigned inc = wave_inclusive_scan(v, lane);
        if (lane == 63)
            sh[wave] = inc;
        __syncthreads();
        unsigned off = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
        {
            const unsigned s = sh[w];
            off += (w < wave) ? s : 0u;
            tot += s;
        }
        __syncthreads();
        total = tot;
        return off + inc - v;
    }

    __global__ __launch_bounds__(kWG) void k_radix_hist(const unsigned *keys, int n, int shift, unsigned *counts, int numTiles)
    {
        __shared__ unsigned h[256];
        h[threadIdx.x] = 0;
        __syncthreads();
        const int base = blockIdx.x * kRadixTile;
#pragma unroll
        for (int r = 0; r < kRadixItems; ++r)
        {
            const int idx = base + r * kWG + threadIdx.x;
            if (idx < n)
                atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
        }
        __syncthreads();
        counts[threadIdx.x * numTiles + blockIdx.x] = h[threadIdx.x];
    }

    // Key per particle (its pblock, or the sentinel numBlocks when it does
    // not take part), the identity permutation, and the histogram of the
    // first radix digit: k_radix_hist fused with the binning pass.
    __global__ __launch_bounds__(kWG) void k_bin_hist(GridGeom g, int n, ParticleBind P, int useMask, unsigned *keys,
                                                      unsigned *vals, unsigned *counts, int numTiles)
    {
        __shared__ unsigned h[256];
        h[threadIdx.x] = 0;
        __syncthreads();
        const int base = blockIdx.x * kRadixTile;
#pragma unroll
        for (int r = 0; r < kRadixItems; ++r)
        {
            const int i = base + r * kWG + threadIdx.x;
            if (i < n)
            {
                unsigned key = (unsigned)g.numBlocks;
                if (!useMask || P.mask[i])
                {
                    const float4 x = P.xm[i];
                    int bx, by, bz;
                    V3 fx;
                    cell_of(g, v3(x), bx, by, bz, fx);
                    key = (unsigned)pblock_key(g, bx, by, bz);
                }
                keys[i] = key;
                vals[i] = (unsigned)i;
                atomicAdd(&h[key & 255u], 1u);
            }
        }
        __syncthreads();
        counts[threadIdx.x * numTiles + blockIdx.x] = h[threadIdx.x];
    }

    // One workgroup per digit: exclusive scan over the tiles, digit total out.
    __global__ __launch_bounds__(kWG) void k_radix_scan(unsigned *counts, int numTiles, unsigned *digitTotals)
    {
        __shared__ unsigned sh[kWaves];
        const int d = blockIdx.x;
        unsigned carry = 0;
        for (int c = 0; c < numTiles; c += kWG)
        {
            const int i = c + threadIdx.x;
            const unsigned v = (i < numTiles) ? counts[d * numTiles + i] : 0u;
            unsigned tot;
            const unsigned ex = block_exclusive_scan(v, tot, sh);
            if (i < numTiles)
                counts[d * numTiles + i] = carry + ex;
            carry += tot;
        }
        if (threadIdx.x == 0)
            digitTotals[d] = carry;
    }

    __global__ __launch_bounds__(kWG) void k_radix_scatter(const unsigned *keysIn, const unsigned *valsIn, int n, int shift,
                                                           const unsigned *counts, int numTiles, const unsigned *digitTotals,
                                                           unsigned *keysOut, unsigned *valsOut)
    {
        __shared__ unsigned digitOff[256];
        __shared__ unsigned waveCnt[kWaves][256];
        __shared__ unsigned roundTot[256];
        __shared__ unsigned sh[kWaves];
        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        unsigned tot;
        const unsigned ex = block_exclusive_scan(digitTotals[tid], tot, sh);
        digitOff[tid] = ex + counts[tid * numTiles + blockIdx.x];
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
            waveCnt[w][tid] = 0;
        __syncthreads();
        const unsigned long long ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        const int base = blockIdx.x * kRadixTile;
        for (int r = 0; r < kRadixItems; ++r)
        {
            const int idx = base + r * kWG + tid;
            const bool valid = idx < n;
            const unsigned key = valid ? keysIn[idx] : 0u;
            const unsigned val = valid ? valsIn[idx] : 0u;
            const unsigned d = (key >> shift) & 255u;
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; ++bit)
            {
                const bool set = (d >> bit) & 1u;
                const unsigned long long bal = __ballot(valid && set);
                peers &= set ? bal : ~bal;
            }
            const unsigned rankInWave = (unsigned)__popcll(peers & ltMask);
            if (valid && rankInWave == 0)
                waveCnt[wave][d] = (unsigned)__popcll(peers);
            __syncthreads();
            unsigned run = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w)
            {
                const unsigned c = waveCnt[w][tid];
                waveCnt[w][tid] = run;
                run += c;
            }
            roundTot[tid] = run;
            __syncthreads();
            if (valid)
            {
                const unsigned pos = digitOff[d] + waveCnt[wave][d] + rankInWave;
                keysOut[pos] = key;
                valsOut[pos] = val;
            }
            __syncthreads();
            digitOff[tid] += roundTot[tid];
#pragma unroll
            for (int w = 0; w < kWaves; ++w)
                waveCnt[w][tid] = 0;
            __syncthreads();
        }
    }

    // Segment boundaries of the sorted keys: pblock k holds the sorted
    // particles [segStart[k], segEnd[k]); both stay 0 for an empty pblock.
    // The tile of pblock k is tile k, so no tile list is built.
    __global__ void k_segments(int n, int numBlocks, const unsigned *keys, int *segStart, int *segEnd)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n)
            return;
        const unsigned key = keys[k];
        if (key >= (unsigned)numBlocks)
            return;
        if (k == 0 || keys[k - 1] != key)
            segStart[key] = k;
        if (k == n - 1 || keys[k + 1] != key)
            segEnd[key] = k + 1;
    }

    // ------------------------------------------------------------------
    // Tile helpers
    // ------------------------------------------------------------------
    struct TileInfo
    {
        int bx, by, bz, start, end;
    };
    // Tile t is pblock t; false when it holds no particle.
    CR_DEV bool tile_info(const GridGeom &g, int t, const int *segStart, const int *segEnd, TileInfo &ti)
    {
        if (t >= g.numBlocks)
            return false;
        ti.start = segStart[t];
        ti.end = segEnd[t];
        if (ti.end <= ti.start)
            return false;
        decode_block(g, t, ti.bx, ti.by, ti.bz);
        return true;
    }

    // Stage the grid nodes of a tile in LDS (xyz velocity, w mass).
    CR_DEV void load_tile_grid(const GridGeom &g, const TileInfo &ti, const float4 *grid, float4 *lds, int lane)
    {
        for (int i = lane; i < kTileNodes; i += 64)
        {
            const int lx = i % kTN, ly = (i / kTN) % kTN, lz = i / (kTN * kTN);
            const int x = ti.bx * kTB + lx, y = ti.by * kTB + ly, z = ti.bz * kTB + lz;
            lds[i] = in_grid(g, x, y, z) ? grid[grid_index(g, x, y, z)] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }

    // Local node index in the tile box, or -1.
    CR_DEV int tile_local(const TileInfo &ti, int x, int y, int z)
    {
        const int lx = x - ti.bx * kTB, ly = y - ti.by * kTB, lz = z - ti.bz * kTB;
        if ((unsigned)lx >= (unsigned)kTN || (unsigned)ly >= (unsigned)kTN || (unsigned)lz >= (unsigned)kTN)
            return -1;
        return lx + kTN * (ly + kTN * lz);
    }

    // Weight i (0, 1, 2) of fraction f and its grid-unit gradient, evaluated
    // directly (same expressions as bspline_w / bspline_dw) so that a rolled
    // node loop needs no indexed weight table, which would live in scratch.
    CR_DEV float w_of(float f, int i)
    {
        const float a = 1.5f - f, b = f - 1.0f, c = f - 0.5f;
        return i == 0 ? 0.5f * (a * a) : (i == 1 ? 0.75f - (b * b) : 0.5f * (c * c));
    }
    CR_DEV float dw_of(float f, int i)
    {
        return i == 0 ? (float)((double)f - 1.5) : (i == 1 ? 2.0f * (1.0f - f) : f - 0.5f);
    }

    // All 27 nodes of base cell (bx, by, bz) inside the grid: the tile box then
    // holds them all and the node loop needs no checks.
    CR_DEV bool nodes_inside(const GridGeom &g, int bx, int by, int bz)
    {
        return bx >= 0 && by >= 0 && bz >= 0 && bx + 2 < g.nx && by + 2 < g.ny && bz + 2 < g.nz;
    }

    // Visit the 27 nodes of a particle in the reference's (i, j, k) order:
    // fn(li, a, b, c, dist, wt) with li the node's index in the tile box,
    // dist = ((a, b, c) - fx) * dx and wt the B-spline weight.  Fast: fully
    // unrolled and unchecked (requires nodes_inside).  Checked: out-of-grid
    // nodes skipped, loops left rolled to keep the code small.
    template <bool Fast, class Fn>
    CR_DEV void for_each_node(const GridGeom &g, const TileInfo &ti, int bx, int by, int bz, V3 fx, const Weights &W, Fn &&fn)
    {
        if constexpr (Fast)
        {
            const int l0 = (bx - ti.bx * kTB) + kTN * ((by - ti.by * kTB) + kTN * (bz - ti.bz * kTB));
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
                    {
                        const V3 dist = (V3{(float)a, (float)b, (float)cc} - fx) * g.dx;
                        const float wt = W.w[a][0] * W.w[b][1] * W.w[cc][2];
                        fn(l0 + a + kTN * (b + kTN * cc), a, b, cc, dist, wt);
                    }
        }
        else
        {
#pragma unroll 1
            for (int a = 0; a < 3; ++a)
#pragma unroll 1
                for (int b = 0; b < 3; ++b)
#pragma unroll 1
                    for (int cc = 0; cc < 3; ++cc)
                    {
                        const int nx = bx + a, ny = by + b, nz = bz + cc;
                        if (!in_grid(g, nx, ny, nz))
                            continue;
                        const int li = tile_local(ti, nx, ny, nz);
                        if (li < 0)
                            continue;
                        const V3 dist = (V3{(float)a, (float)b, (float)cc} - fx) * g.dx;
                        const float wt = w_of(fx.x, a) * w_of(fx.y, b) * w_of(fx.z, cc);
                        fn(li, a, b, cc, dist, wt);
                    }
        }
    }

    // ------------------------------------------------------------------
    // Conflict-free LDS scatter.  Two lanes can only hit the same node in one
    // node-loop instruction if their particles share a base cell, so lanes are
    // grouped by base cell and each pass lets one lane per cell scatter with a
    // plain read-modify-write (ds_read_b128 + ds_write_b128 per node) instead
    // of four ds_add_f32 atomics.  Lanes sharing a cell take turns in lane
    // order, so the summation order is fixed and the result reproducible.
    // ------------------------------------------------------------------
    CR_DEV void compiler_fence() { asm volatile("" ::: "memory"); }

    // Makes values opaque to the optimiser at this point (no instruction), so
    // the per-node work of a scatter pass is not hoisted out of the pass loop.
    CR_DEV void opaque(float &v) { asm volatile("" : "+v"(v)); }
    CR_DEV void opaque(V3 &v)
    {
        opaque(v.x);
        opaque(v.y);
        opaque(v.z);
    }
    CR_DEV void opaque(M3 &m)
    {
        opaque(m.c0);
        opaque(m.c1);
        opaque(m.c2);
    }
    CR_DEV void opaque(Weights &W)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int a = 0; a < 3; ++a)
                opaque(W.w[i][a]);
    }

    // Key of the base cell inside the tile box: 3 bits per axis of the local
    // base offset, clamped to [-2, 5] (a particle whose offset is clamped has
    // no node in the box on that side, so sharing a key only serialises it).
    CR_DEV unsigned cell_key(const TileInfo &ti, int bx, int by, int bz)
    {
        const int lx = min(max(bx - ti.bx * kTB + 2, 0), 7);
        const int ly = min(max(by - ti.by * kTB + 2, 0), 7);
        const int lz = min(max(bz - ti.bz * kTB + 2, 0), 7);
        return (unsigned)(lx | (ly << 3) | (lz << 6));
    }

    // Active lanes whose 9-bit key equals this lane's.
    CR_DEV unsigned long long match_lanes(unsigned key)
    {
        unsigned long long peers = __ballot(1);
#pragma unroll
        for (int b = 0; b < 9; ++b)
        {
            const bool set = (key >> b) & 1u;
            const unsigned long long bal = __ballot(set);
            peers &= set ? bal : ~bal;
        }
        return peers;
    }

    // Runs body() on every active lane, in passes in which no two lanes share a key.
    template <class Body>
    CR_DEV void in_conflict_free_passes(unsigned key, Body &&body)
    {
        const unsigned long long peers = match_lanes(key);
        const int lane = __lane_id();
        bool pending = true;
        for (;;)
        {
            const unsigned long long pend = __ballot(pending);
            if (pend == 0ull)
                break;
            if (pending && lane == __ffsll((unsigned long long)(pend & peers)) - 1)
            {
                body();
                pending = false;
            }
            compiler_fence();
        }
    }

    CR_DEV void lds_add(float4 *A, int li, float x, float y, float z, float w)
    {
        float4 cur = A[li];
        cur.x += x;
        cur.y += y;
        cur.z += z;
        cur.w += w;
        A[li] = cur;
        compiler_fence();
    }

    // ------------------------------------------------------------------
    // Order of the particles inside each tile: by (rank within the base cell,
    // base cell), ranks counted in sorted order.  The tile kernels run 64
    // lanes over consecutive particles of a tile, and lanes that share a base
    // cell take turns in the conflict-free scatter.  In this order each rank
    // level holds one particle per occupied cell, so a dense tile (several
    // particles per cell) scatters in one or two passes per 64-lane chunk
    // instead of one pass per particle of its fullest cell.  One wavefront per
    // tile; tmp holds (rank << 6 | cell) per sorted particle.
    // ------------------------------------------------------------------
    __global__ __launch_bounds__(kWG) void k_cell_rank_order(GridGeom g, const int *segStart, const int *segEnd,
                                                             const float4 *xm, const unsigned *order, unsigned *tmp,
                                                             unsigned *orderOut)
    {
        __shared__ int cnt[kWaves][64];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        TileInfo ti;
        const int t = blockIdx.x * kWaves + wave;
        if (!tile_info(g, t, segStart, segEnd, ti))
            return;
        int *C = cnt[wave];
        C[lane] = 0;
        wave_sync();
        const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        for (int base = ti.start; base < ti.end; base += 64)
        {
            const int k = base + lane;
            const bool valid = k < ti.end;
            unsigned cell = 64u;
            if (valid)
            {
                int bx, by, bz;
                V3 fx;
                cell_of(g, v3(xm[order[k]]), bx, by, bz, fx);
                bx = min(max(bx, 0), g.nx - 1);
                by = min(max(by, 0), g.ny - 1);
                bz = min(max(bz, 0), g.nz - 1);
                cell = (unsigned)((bx & 3) | ((by & 3) << 2) | ((bz & 3) << 4));
            }
            const unsigned long long peers = match_lanes(cell);
            const int below = __popcll(peers & lt);
            const int r = valid ? C[cell] + below : 0;
            wave_sync();
            if (valid && below == 0)
                C[cell] += __popcll(peers);
            wave_sync();
            if (valid)
                tmp[k] = ((unsigned)r << 6) | cell;
        }
        const int mine = C[lane];
        for (int base = ti.start; base < ti.end; base += 64)
        {
            const int k = base + lane;
            const bool valid = k < ti.end;
            const unsigned rc = valid ? tmp[k] : 0u;
            const int r = (int)(rc >> 6), cell = (int)(rc & 63u);
            // particles of lower rank, then rank-r particles of lower cells
            int pos = 0;
#pragma unroll 16
            for (int c = 0; c < 64; ++c)
            {
                const int nc = __builtin_amdgcn_readlane(mine, c);
                pos += min(nc, r) + ((c < cell && nc > r) ? 1 : 0);
            }
            if (valid)
                orderOut[ti.start + pos] = order[k];
        }
    }

    // ------------------------------------------------------------------
    // Separable node sums.  For a particle whose 27 nodes all lie in the tile
    // box and need no slicer test, the node offset dist = ((a, b, c) - fx) dx
    // has per-axis components, so the per-node outer products of G2P and the
    // per-node matrix-vector products of P2G factor into per-axis terms: the
    // same sums as the reference's node loops, associated differently
    // (about 13 instead of 25 VALU operations per node).
    // ------------------------------------------------------------------
    struct AxisOffsets
    {
        float d[3][3]; // d[axis][i] = (i - fx[axis]) * dx
    };
    CR_DEV AxisOffsets axis_offsets(V3 fx, float dx)
    {
        AxisOffsets o;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            o.d[0][i] = ((float)i - fx.x) * dx;
            o.d[1][i] = ((float)i - fx.y) * dx;
            o.d[2][i] = ((float)i - fx.z) * dx;
        }
        return o;
    }
    // Local index of node (a, b, c) of base cell local offset l0.
    CR_DEV int node_li(int l0, int a, int b, int c) { return l0 + a + kTN * (b + kTN * c); }
    CR_DEV int base_li(const TileInfo &ti, int bx, int by, int bz)
    {
        return (bx - ti.bx * kTB) + kTN * ((by - ti.by * kTB) + kTN * (bz - ti.bz * kTB));
    }

    // G2P: v = sum w v_i and B = sum outer_ref(w v_i, dist); column k of B is
    // sum (w v_i)_k dist, and dist's r-component depends only on the node's
    // r-axis index, so B(r, k) = sum_i d[r][i] * S_r[i]_k with S_r[i] the sum
    // of w v over the nodes whose r-axis index is i.
    CR_DEV void gather_separable(const float4 *G, int l0, const Weights &W, const AxisOffsets &o, V3 &v, M3 &B)
    {
        V3 sx[3], sy[3], sz[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            sx[i] = sy[i] = sz[i] = V3{0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
            {
                const float wab = W.w[a][0] * W.w[b][1];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                {
                    const V3 nw = v3(G[node_li(l0, a, b, c)]) * (wab * W.w[c][2]);
                    sx[a] = sx[a] + nw;
                    sy[b] = sy[b] + nw;
                    sz[c] = sz[c] + nw;
                }
            }
        v = (sx[0] + sx[1]) + sx[2];
        auto col = [&](int k) {
            auto comp = [&](const V3 *s3, int r) {
                const float *q0 = &s3[0].x, *q1 = &s3[1].x, *q2 = &s3[2].x;
                return o.d[r][0] * q0[k] + o.d[r][1] * q1[k] + o.d[r][2] * q2[k];
            };
            return V3{comp(sx, 0), comp(sy, 1), comp(sz, 2)};
        };
        B = M3{col(0), col(1), col(2)};
    }

    // P2G terms (v + C dist) * w * m: C dist = c0 d_x + c1 d_y + c2 d_z, so
    // va[a] = v + c0 d[0][a], py[b] = c1 d[1][b], pz[c] = c2 d[2][c].
    struct ScatterTerms
    {
        V3 va[3], py[3], pz[3];
        float wxm[3], wy[3], wz[3]; // x weights times the particle mass
    };
    CR_DEV ScatterTerms scatter_terms(V3 v, const M3 &Cm, float m, const Weights &W, const AxisOffsets &o)
    {
        ScatterTerms t;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            t.va[i] = v + Cm.c0 * o.d[0][i];
            t.py[i] = Cm.c1 * o.d[1][i];
            t.pz[i] = Cm.c2 * o.d[2][i];
            t.wxm[i] = W.w[i][0] * m;
            t.wy[i] = W.w[i][1];
            t.wz[i] = W.w[i][2];
        }
        return t;
    }
    CR_DEV void opaque(ScatterTerms &t)
    {
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            opaque(t.va[i]);
            opaque(t.py[i]);
            opaque(t.pz[i]);
            opaque(t.wxm[i]);
            opaque(t.wy[i]);
            opaque(t.wz[i]);
        }
    }
    // Scatter (term * w m, w m) to the 27 nodes.
    CR_DEV void scatter_separable(float4 *A, int l0, const ScatterTerms &t)
    {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int b = 0; b < 3; ++b)
            {
                const V3 q = t.py[b] + t.pz[c];
                const float wbc = t.wy[b] * t.wz[c];
#pragma unroll
                for (int a = 0; a < 3; ++a)
                {
                    const float wm = t.wxm[a] * wbc;
                    const V3 contrib = (t.va[a] + q) * wm;
                    lds_add(A, node_li(l0, a, b, c), contrib.x, contrib.y, contrib.z, wm);
                }
            }
    }

    // ------------------------------------------------------------------
    // MLS-MPM P2G (src/solver/mls_mpm_algorithm.h:147-241)
    // ------------------------------------------------------------------
    struct StepCommon
    {
        GridGeom g;
        float3 gravity;
        float dt;
        ParticleBind P;
        SolverState S;
        ShapeView sv;
        const int *segStart;
        const int *segEnd;
        const unsigned *order; // sorted particle indices
        float4 *tilebuf;
        float4 *grid;
    };

    __global__ __launch_bounds__(kWG) void k_mls_p2g(StepCommon c)
    {
        __shared__ float4 acc[kWaves][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        TileInfo ti;
        const int t = blockIdx.x * kWaves + wave;
        if (!tile_info(c.g, t, c.segStart, c.segEnd, ti))
            return;
        float4 *A = acc[wave];
        for (int i = lane; i < kTileNodes; i += 64)
            A[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        wave_sync();
        const int nSlicers = c.sv.counts[1];
        const GridGeom &g = c.g;
        for (int k = ti.start + lane; k < ti.end; k += 64)
        {
            const int p = (int)c.order[k];
            const float4 xm = c.P.xm[p];
            const V3 x = v3(xm);
            const float m = xm.w;
            int bx, by, bz;
            V3 fx;
            cell_of(g, x, bx, by, bz, fx);
            Weights W;
            weights_of(fx, W);
            float4 vv = c.P.v[p];
            const V3 vel = v3(vv) + v3(c.gravity.x, c.gravity.y, c.gravity.z) * c.dt;
            c.P.v[p] = make_float4(vel.x, vel.y, vel.z, vv.w);
            const M3 F = load_m3(c.S.F, c.S.stride, p);
            const M3 Cm = load_m3(c.S.C, c.S.stride, p);
            const float4 prm = c.P.params[p];
            const M3 stress = stress_kirchhoff(prm.x, prm.y, c.P.type[p], F);
            const float sc = -c.dt * 4.0f * g.invdx * g.invdx * c.S.V0[p];
            const M3 affine = stress * sc + Cm * m;
            const V3 vm = vel * m;
            const bool fast = nSlicers == 0 && nodes_inside(g, bx, by, bz);
            in_conflict_free_passes(cell_key(ti, bx, by, bz), [&]() {
                M3 af = affine;
                V3 v_m = vm, f = fx;
                Weights w = W;
                opaque(af);
                opaque(v_m);
                opaque(f);
                opaque(w);
                auto sc = [&](int li, int, int, int, V3 dist, float wt) {
                    const V3 contrib = (v_m + mul(af, dist)) * wt;
                    lds_add(A, li, contrib.x, contrib.y, contrib.z, wt * m);
                };
                if (fast)
                    for_each_node<true>(g, ti, bx, by, bz, f, w, sc);
                else
                    for_each_node<false>(g, ti, bx, by, bz, f, w, [&](int li, int a, int b, int cc, V3 dist, float wt) {
                        V3 nrm;
                        if (nSlicers > 0 && different_sides<false>(c.sv, nSlicers, x, x + dist, nrm))
                            return;
                        sc(li, a, b, cc, dist, wt);
                    });
            });
        }
        wave_sync();
        float4 *out = c.tilebuf + (size_t)t * kTileNodes;
        for (int i = lane; i < kTileNodes; i += 64)
            out[i] = A[i];
    }

    // ------------------------------------------------------------------
    // Grid update over active node blocks: tile sums in a fixed order, then
    // mlsMpmUpdateGrid / pbMpmUpdateGrid / standardMpmUpdateGrid
    // (src/solver/mls_mpm_algorithm.h:243-294, pb 233-284, standard 217-272).
    // ------------------------------------------------------------------
    template <int Kind> // 0 MLS, 1 PB, 2 standard
    __global__ __launch_bounds__(kWG) void k_update(StepCommon c, const float4 *tileForce, unsigned char *dirty)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nb = blockIdx.x * kWaves + wave;
        const GridGeom &g = c.g;
        if (nb >= g.numBlocks)
            return;
        int bx, by, bz;
        decode_block(g, nb, bx, by, bz);
        // Tiles whose 6^3 box covers this block: pblocks nb - (dx, dy, dz).
        // All occupancy loads are issued before any partial is read.
        bool src[8];
        bool any = false;
#pragma unroll
        for (int d = 0; d < 8; ++d)
        {
            const int px = bx - (d & 1), py = by - ((d >> 1) & 1), pz = bz - (d >> 2);
            const int key = (px >= 0 && py >= 0 && pz >= 0) ? block_key(g, px, py, pz) : -1;
            src[d] = key >= 0 && c.segEnd[key] > c.segStart[key];
            any = any || src[d];
        }
        if (!any)
            return;
        const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
        float4 q[8], qf[8];
#pragma unroll
        for (int d = 0; d < 8; ++d)
        {
            const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
            const int ux = lx + 4 * dx, uy = ly + 4 * dy, uz = lz + 4 * dz;
            const bool use = src[d] && ux < kTN && uy < kTN && uz < kTN;
            const size_t at = (size_t)(nb - dx - g.nbx * (dy + g.nby * dz)) * kTileNodes + (ux + kTN * (uy + kTN * uz));
            q[d] = use ? c.tilebuf[at] : make_float4(0.f, 0.f, 0.f, 0.f);
            if (Kind == 2)
                qf[d] = use ? tileForce[at] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        // fixed summation order d = 0..7 (bitwise reproducible)
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int d = 0; d < 8; ++d)
        {
            s.x += q[d].x;
            s.y += q[d].y;
            s.z += q[d].z;
            s.w += q[d].w;
            if (Kind == 2)
            {
                f.x += qf[d].x;
                f.y += qf[d].y;
                f.z += qf[d].z;
            }
        }
        if (lane == 0)
            dirty[nb] = 1;
        const int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
        if (!in_grid(g, i, j, k))
            return;
        const int gi = nb * 64 + lane;
        const float cutoff = (Kind == 1) ? 1e-8f : 1e-4f;
        if (!(s.w < cutoff))
        {
            if (Kind == 2)
            {
                s.x += f.x * c.dt;
                s.y += f.y * c.dt;
                s.z += f.z * c.dt;
            }
            s.x /= s.w;
            s.y /= s.w;
            s.z /= s.w;
            V3 np = V3{(float)i * g.dx, (float)j * g.dx, (float)k * g.dx} + v3(g.bmin.x, g.bmin.y, g.bmin.z);
            np = np + v3(s) * c.dt;
            resolve_rigid<false>(c.sv, c.sv.counts[0], s, np);
            if (i < 2 || i > g.nx - 3)
                s.x = 0.0f;
            if (j < 2 || j > g.ny - 3)
                s.y = 0.0f;
            if (k < 2 || k > g.nz - 3)
                s.z = 0.0f;
        }
        c.grid[gi] = s;
    }

    // Gather of v and C from the LDS tile (src/solver/mls_mpm_algorithm.h:318-371)
    CR_DEV void gather_vc(const StepCommon &c, const TileInfo &ti, const float4 *G, int nSlicers, V3 x, V3 vp, V3 &vOut,
                          M3 &cOut)
    {
        const GridGeom &g = c.g;
        int bx, by, bz;
        V3 fx;
        cell_of(g, x, bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        V3 vacc{0, 0, 0};
        M3 aacc{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}};
        auto gather = [&](V3 nv, V3 dist, float wt) {
            const V3 nw = nv * wt;
            vacc = vacc + nw;
            aacc = aacc + outer_ref(nw, dist);
        };
        if (nSlicers == 0 && nodes_inside(g, bx, by, bz))
        {
            gather_separable(G, base_li(ti, bx, by, bz), W, axis_offsets(fx, g.dx), vacc, aacc);
        }
        else
            for_each_node<false>(g, ti, bx, by, bz, fx, W, [&](int li, int, int, int, V3 dist, float wt) {
                V3 nv = v3(G[li]);
                V3 n;
                if (nSlicers > 0 && different_sides<true>(c.sv, nSlicers, x, x + dist, n))
                    nv = vp - n * dot(vp, n) * 2.0f;
                gather(nv, dist, wt);
            });
        vOut = vacc;
        cOut = aacc * (4.0f * g.invdx);
    }

    CR_DEV V3 clamp_bounds(const GridGeom &g, V3 x)
    {
        return V3{fminf(fmaxf(x.x, g.bmin.x), g.bmax.x), fminf(fmaxf(x.y, g.bmin.y), g.bmax.y),
                  fminf(fmaxf(x.z, g.bmin.z), g.bmax.z)};
    }

    // MLS G2P (src/solver/mls_mpm_algorithm.h:296-390)
    __global__ __launch_bounds__(kWG) void k_mls_g2p(StepCommon c)
    {
        __shared__ float4 lg[kWaves][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        TileInfo ti;
        const int t = blockIdx.x * kWaves + wave;
        if (!tile_info(c.g, t, c.segStart, c.segEnd, ti))
            return;
        float4 *G = lg[wave];
        load_tile_grid(c.g, ti, c.grid, G, lane);
        wave_sync();
        const int nSlicers = c.sv.counts[1];
        const int nShapes = c.sv.counts[0];
        for (int k = ti.start + lane; k < ti.end; k += 64)
        {
            const int p = (int)c.order[k];
            const float4 xm = c.P.xm[p];
            V3 x = v3(xm);
            V3 vel;
            M3 Cm;
            gather_vc(c, ti, G, nSlicers, x, v3(c.P.v[p]), vel, Cm);
            x = clamp_bounds(c.g, x + vel * c.dt);
            float4 vv = make_float4(vel.x, vel.y, vel.z, 0.0f);
            resolve_rigid<true>(c.sv, nShapes, vv, x);
            c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
            c.P.v[p] = make_float4(vel.x, vel.y, vel.z, 0.0f);
            store_m3(c.S.C, c.S.stride, p, Cm);
            const M3 F = load_m3(c.S.F, c.S.stride, p);
            store_m3(c.S.F, c.S.stride, p, mul(m3_identity() + Cm * c.dt, F));
        }
    }

    // ------------------------------------------------------------------
    // PB-MPM (src/solver/pb_mpm_algorithm.h:150-408).  Particle state of the
    // substep is staged in sorted order; x, F, params and type are fixed over
    // the iterations, v and C change.
    // ------------------------------------------------------------------
    struct PbStage
    {
        float4 *xs;   // sorted x|m
        float *vs;    // [3][stride]
        float *Fs;    // [9][stride]
        float *Cs;    // [9][stride]
        float2 *ps;   // (alpha, ratio)
        int *ts;      // material type
        int stride;
    };

    __global__ void k_pb_stage(int n, int numBlocks, const unsigned *keys, const unsigned *order, ParticleBind P,
                               SolverState S, PbStage st)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n || keys[k] >= (unsigned)numBlocks)
            return;
        const int p = (int)order[k];
        st.xs[k] = P.xm[p];
        const float4 v = P.v[p];
        st.vs[k] = v.x;
        st.vs[st.stride + k] = v.y;
        st.vs[2 * st.stride + k] = v.z;
#pragma unroll
        for (int e = 0; e < 9; ++e)
        {
            st.Fs[e * st.stride + k] = S.F[e * S.stride + p];
            st.Cs[e * st.stride + k] = S.C[e * S.stride + p];
        }
        const float4 prm = P.params[p];
        st.ps[k] = make_float2(prm.x, prm.y);
        st.ts[k] = P.type[p];
    }

    // Scatter of one PB particle into the LDS tile (pbMpmParticleToGrid loop)
    CR_DEV void pb_scatter(const StepCommon &c, const TileInfo &ti, float4 *A, int nSlicers, V3 x, float m, V3 vel,
                           const M3 &Cm)
    {
        const GridGeom &g = c.g;
        int bx, by, bz;
        V3 fx;
        cell_of(g, x, bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        const bool fast = nSlicers == 0 && nodes_inside(g, bx, by, bz);
        const ScatterTerms terms = scatter_terms(vel, Cm, m, W, axis_offsets(fx, g.dx));
        const int l0 = fast ? base_li(ti, bx, by, bz) : 0;
        in_conflict_free_passes(cell_key(ti, bx, by, bz), [&]() {
            if (fast)
            {
                ScatterTerms tt = terms;
                opaque(tt);
                scatter_separable(A, l0, tt);
                return;
            }
            M3 cm = Cm;
            V3 v = vel, f = fx;
            Weights w = W;
            opaque(cm);
            opaque(v);
            opaque(f);
            opaque(w);
            auto scatter = [&](int li, int, int, int, V3 dist, float wt) {
                const float wm = wt * m;
                const V3 contrib = (v + mul(cm, dist)) * wm;
                lds_add(A, li, contrib.x, contrib.y, contrib.z, wm);
            };
            for_each_node<false>(g, ti, bx, by, bz, f, w, [&](int li, int a, int b, int cc, V3 dist, float wt) {
                V3 nrm;
                if (nSlicers > 0 && different_sides<false>(c.sv, nSlicers, x, x + dist, nrm))
                    return;
                scatter(li, a, b, cc, dist, wt);
            });
        });
    }

    // Mode 0: P2G of the first iteration.  Mode 1: G2P of the previous
    // iteration fused with P2G of this one.  Mode 2: last G2P fused with
    // pbMpmIntegrateParticle, results written back in particle order.
    template <int Mode>
    __global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(3))) void k_pb_tile(StepCommon c, PbStage st, float3 gravity)
    {
        __shared__ float4 acc[(Mode < 2) ? kWaves : 1][kTileNodes];
        __shared__ float4 lg[(Mode > 0) ? kWaves : 1][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        TileInfo ti;
        const int t = blockIdx.x * kWaves + wave;
        if (!tile_info(c.g, t, c.segStart, c.segEnd, ti))
            return;
        float4 *A = acc[(Mode < 2) ? wave : 0];
        float4 *G = lg[(Mode > 0) ? wave : 0];
        if (Mode < 2)
            for (int i = lane; i < kTileNodes; i += 64)
                A[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (Mode > 0)
            load_tile_grid(c.g, ti, c.grid, G, lane);
        wave_sync();
        const int nSlicers = c.sv.counts[1];
        const int nShapes = c.sv.counts[0];
        const int S = st.stride;
        for (int k = ti.start + lane; k < ti.end; k += 64)
        {
            const float4 xm = st.xs[k];
            V3 x = v3(xm);
            V3 vel{st.vs[k], st.vs[S + k], st.vs[2 * S + k]};
            M3 Cm;
            if (Mode == 0)
                Cm = load_m3(st.Cs, S, k);
            else
            {
                V3 vnew;
                gather_vc(c, ti, G, nSlicers, x, vel, vnew, Cm);
                vel = vnew;
            }
            if (Mode < 2)
            {
                const M3 F = load_m3(st.Fs, S, k);
                const float2 pr = st.ps[k];
                pb_constraint(c.dt, pr.x, pr.y, st.ts[k], F, Cm);
                if (Mode == 1 && nSlicers > 0)
                {
                    st.vs[k] = vel.x;
                    st.vs[S + k] = vel.y;
                    st.vs[2 * S + k] = vel.z;
                }
                pb_scatter(c, ti, A, nSlicers, x, xm.w, vel, Cm);
            }
            else
            {
                // pbMpmIntegrateParticle (src/solver/pb_mpm_algorithm.h:363-408)
                const int p = (int)c.order[k];
                x = clamp_bounds(c.g, x + vel * c.dt);
                vel = vel + v3(gravity.x, gravity.y, gravity.z) * c.dt;
                float4 vv = make_float4(vel.x, vel.y, vel.z, 0.0f);
                resolve_rigid<true>(c.sv, nShapes, vv, x);
                c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
                c.P.v[p] = make_float4(vel.x, vel.y, vel.z, 0.0f);
                store_m3(c.S.C, c.S.stride, p, Cm);
                const M3 F = load_m3(st.Fs, S, k);
                store_m3(c.S.F, c.S.stride, p, mul(m3_identity() + Cm * c.dt, F));
            }
        }
        if (Mode < 2)
        {
            wave_sync();
            float4 *out = c.tilebuf + (size_t)t * kTileNodes;
            for (int i = lane; i < kTileNodes; i += 64)
                out[i] = A[i];
        }
    }

    // Zero iterations: integrate with the current v and C.
    __global__ void k_pb_integrate_only(StepCommon c, int n, const unsigned *keys, PbStage st, float3 gravity)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n || keys[k] >= (unsigned)c.g.numBlocks)
            return;
        const int S = st.stride;
        const int p = (int)c.order[k];
        const float4 xm = st.xs[k];
        V3 vel{st.vs[k], st.vs[S + k], st.vs[2 * S + k]};
        const M3 Cm = load_m3(st.Cs, S, k);
        V3 x = clamp_bounds(c.g, v3(xm) + vel * c.dt);
        vel = vel + v3(gravity.x, gravity.y, gravity.z) * c.dt;
        float4 vv = make_float4(vel.x, vel.y, vel.z, 0.0f);
        resolve_rigid<true>(c.sv, c.sv.counts[0], vv, x);
        c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
        c.P.v[p] = make_float4(vel.x, vel.y, vel.z, 0.0f);
        const M3 F = load_m3(st.Fs, S, k);
        store_m3(c.S.F, c.S.stride, p, mul(m3_identity() + Cm * c.dt, F));
    }

    // ------------------------------------------------------------------
    // Standard MPM (src/solver/standard_mpm_algorithm.h:146-352)
    // ------------------------------------------------------------------
    __global__ __launch_bounds__(kWG) void k_std_p2g(StepCommon c, float4 *tileForce)
    {
        __shared__ float4 acc[kWaves][2][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        TileInfo ti;
        const int t = blockIdx.x * kWaves + wave;
        if (!tile_info(c.g, t, c.segStart, c.segEnd, ti))
            return;
        float4 *A = acc[wave][0];
        float4 *AF = acc[wave][1];
        for (int i = lane; i < kTileNodes; i += 64)
        {
            A[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            AF[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        wave_sync();
        const GridGeom &g = c.g;
        const V3 grav = v3(c.gravity.x, c.gravity.y, c.gravity.z);
        for (int k = ti.start + lane; k < ti.end; k += 64)
        {
            const int p = (int)c.order[k];
            const float4 xm = c.P.xm[p];
            const float m = xm.w;
            int bx, by, bz;
            V3 fx;
            cell_of(g, v3(xm), bx, by, bz, fx);
            Weights W;
            weights_of(fx, W);
            const float4 prm = c.P.params[p];
            const M3 stress = stress_kirchhoff(prm.x, prm.y, c.P.type[p], load_m3(c.S.F, c.S.stride, p));
            const V3 vel = v3(c.P.v[p]);
            const float nv0 = -c.S.V0[p];
            const bool fast = nodes_inside(g, bx, by, bz);
            in_conflict_free_passes(cell_key(ti, bx, by, bz), [&]() {
                M3 st = stress;
                V3 v = vel, f = fx;
                Weights w = W;
                opaque(st);
                opaque(v);
                opaque(f);
                opaque(w);
                auto scatter = [&](int li, int a, int b, int cc, V3, float wt) {
                    const float wa = w_of(f.x, a), wb = w_of(f.y, b), wc = w_of(f.z, cc);
                    const V3 wg{dw_of(f.x, a) * wb * wc, wa * dw_of(f.y, b) * wc, wa * wb * dw_of(f.z, cc)};
                    const float wm = wt * m;
                    const V3 ext = grav * wm;
                    const V3 inr = mul(st, wg) * nv0;
                    const V3 mom = v * wm;
                    const V3 fo = ext + inr;
                    lds_add(A, li, mom.x, mom.y, mom.z, wm);
                    lds_add(AF, li, fo.x, fo.y, fo.z, 0.0f);
                };
                if (fast)
                    for_each_node<true>(g, ti, bx, by, bz, f, w, scatter);
                else
                    for_each_node<false>(g, ti, bx, by, bz, f, w, scatter);
            });
        }
        wave_sync();
        float4 *out = c.tilebuf + (size_t)t * kTileNodes;
        float4 *outF = tileForce + (size_t)t * kTileNodes;
        for (int i = lane; i < kTileNodes; i += 64)
        {
            out[i] = A[i];
            outF[i] = AF[i];
        }
    }

    __global__ __launch_bounds__(kWG) void k_std_g2p(StepCommon c)
    {
        __shared__ float4 lg[kWaves][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        TileInfo ti;
        const int t = blockIdx.x * kWaves + wave;
        if (!tile_info(c.g, t, c.segStart, c.segEnd, ti))
            return;
        float4 *G = lg[wave];
        load_tile_grid(c.g, ti, c.grid, G, lane);
        wave_sync();
        const GridGeom &g = c.g;
        const int nShapes = c.sv.counts[0];
        for (int k = ti.start + lane; k < ti.end; k += 64)
        {
            const int p = (int)c.order[k];
            if (!c.P.mask[p])
                continue;
            const float4 xm = c.P.xm[p];
            int bx, by, bz;
            V3 fx;
            cell_of(g, v3(xm), bx, by, bz, fx);
            Weights W;
            weights_of(fx, W);
            V3 vacc{0, 0, 0};
            M3 gv{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}};
            auto gather = [&](int li, int a, int b, int cc, V3, float wt) {
                const V3 nv = v3(G[li]);
                vacc = vacc + nv * wt;
                const float wa = w_of(fx.x, a), wb = w_of(fx.y, b), wc = w_of(fx.z, cc);
                const V3 wg{dw_of(fx.x, a) * wb * wc, wa * dw_of(fx.y, b) * wc, wa * wb * dw_of(fx.z, cc)};
                gv = gv + outer_ref(wg, nv);
            };
            if (nodes_inside(g, bx, by, bz))
                for_each_node<true>(g, ti, bx, by, bz, fx, W, gather);
            else
                for_each_node<false>(g, ti, bx, by, bz, fx, W, gather);
            V3 x = clamp_bounds(g, v3(xm) + vacc * c.dt);
            float4 vv = make_float4(vacc.x, vacc.y, vacc.z, 0.0f);
            resolve_rigid<true>(c.sv, nShapes, vv, x);
            c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
            c.P.v[p] = make_float4(vacc.x, vacc.y, vacc.z, 0.0f);
            const M3 F = load_m3(c.S.F, c.S.stride, p);
            store_m3(c.S.F, c.S.stride, p, mul(m3_identity() + gv * c.dt, F));
        }
    }

    // ------------------------------------------------------------------
    // computeInitialData (src/solver/mls_mpm_algorithm.h:46-145): mass onto the
    // grid as it stands (global atomics, rarely called), then V0.
    // ------------------------------------------------------------------
    __global__ void k_init_mass(GridGeom g, unsigned n, const unsigned *idx, const float4 *xm, SolverState S, float4 *grid,
                                unsigned char *dirty)
    {
        const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const int p = (int)idx[i];
        store_m3(S.F, S.stride, p, m3_identity());
        const float4 x = xm[p];
        int bx, by, bz;
        V3 fx;
        cell_of(g, v3(x), bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c)
                {
                    const int nx = bx + a, ny = by + b, nz = bz + c;
                    if (!in_grid(g, nx, ny, nz))
                        continue;
                    const float wt = W.w[a][0] * W.w[b][1] * W.w[c][2];
                    atomicAdd(&grid[grid_index(g, nx, ny, nz)].w, wt * x.w);
                    dirty[block_key(g, nx >> 2, ny >> 2, nz >> 2)] = 1;
                }
    }

    __global__ void k_init_volume(GridGeom g, unsigned n, const unsigned *idx, const float4 *xm, SolverState S,
                                  const float4 *grid)
    {
        const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const int p = (int)idx[i];
        store_m3(S.F, S.stride, p, m3_identity());
        const float4 x = xm[p];
        int bx, by, bz;
        V3 fx;
        cell_of(g, v3(x), bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        float density = 0.0f;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c)
                {
                    const int nx = bx + a, ny = by + b, nz = bz + c;
                    if (!in_grid(g, nx, ny, nz))
                        continue;
                    const float wt = W.w[a][0] * W.w[b][1] * W.w[c][2];
                    const float nd = grid[grid_index(g, nx, ny, nz)].w / g.cellVol;
                    density += nd * wt;
                }
        S.V0[p] = (density > 1e-8f) ? x.w / density : 1e-8f;
    }

    __global__ void k_copy_grid(GridGeom g, const float4 *grid, float4 *dst)
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        const int total = g.nx * g.ny * g.nz;
        if (i >= total)
            return;
        const int x = i % g.nx, y = (i / g.nx) % g.ny, z = i / (g.nx * g.ny);
        dst[i] = grid[grid_index(g, x, y, z)];
    }

    int grid1d(long long n, int bs) { return (int)((n + bs - 1) / bs); }
} // namespace

// ----------------------------------------------------------------------
// Solver object behind the C-ABI
// ----------------------------------------------------------------------
struct CrHipSolver
{
    int kind = CRMPM_HIP_MLS;
    int maxParticles = 0;
    GridGeom g{};
    float3 gravity = make_float3(0.f, 0.f, 0.f);
    float dt = 0.002f; // mpm_solver_base.h default
    int iterations = 10;
    int contactModel = CRMPM_HIP_CONTACT_KINEMATIC;
    hipStream_t stream = nullptr;
    bool initialized = false;

    ParticleBind P{};
    unsigned numActive = 0;
    int numShapes = 0;
    int *shapeIds = nullptr;
    CrHipShapeData shapes{};
    CrHipGeometryData geoms{};
    CrHipGeometrySdfData sdf{};

    // owned
    float *F = nullptr, *C = nullptr, *V0 = nullptr;
    float4 *grid = nullptr, *gridForce = nullptr;
    unsigned char *dirty = nullptr;
    int *segStart = nullptr, *segEnd = nullptr;
    unsigned *keysA = nullptr, *keysB = nullptr, *valsA = nullptr, *valsB = nullptr;
    unsigned *radixCounts = nullptr, *digitTotals = nullptr;
    float4 *tilebuf = nullptr, *tileForce = nullptr;
    long long maxTiles = 0;
    ShapeRec *shapeRecs = nullptr;
    int *shapeCounts = nullptr, *slicerList = nullptr;
    int shapeCap = 0;
    PbStage st{};
    float *stageMem = nullptr;

    // optional per-launch event timing (crHipSolverProfileBegin/End)
    int profMax = 0, profUsed = 0;
    unsigned profMask = ~0u;
    hipEvent_t *profEvents = nullptr;
    int *profCat = nullptr;
};

namespace
{
    template <class T>
    int dalloc(T **p, size_t count)
    {
        if (count == 0)
            count = 1;
        if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess)
        {
            *p = nullptr;
            return CRMPM_HIP_ERR_ALLOC;
        }
        return CRMPM_HIP_OK;
    }
    template <class T>
    void dfree(T *&p)
    {
        if (p)
            (void)hipFree((void *)p);
        p = nullptr;
    }

    int check_launch()
    {
        return hipGetLastError() == hipSuccess ? CRMPM_HIP_OK : CRMPM_HIP_ERR_LAUNCH;
    }

    int ensure_shape_capacity(CrHipSolver *s, int n)
    {
        if (n <= s->shapeCap && s->shapeRecs)
            return CRMPM_HIP_OK;
        const int cap = std::max(n, 16);
        dfree(s->shapeRecs);
        dfree(s->slicerList);
        if (dalloc(&s->shapeRecs, cap) || dalloc(&s->slicerList, cap))
            return CRMPM_HIP_ERR_ALLOC;
        s->shapeCap = cap;
        return CRMPM_HIP_OK;
    }

    ShapeView shape_view(const CrHipSolver *s)
    {
        ShapeView v;
        v.recs = s->shapeRecs;
        v.counts = s->shapeCounts;
        v.slicers = s->slicerList;
        v.force = reinterpret_cast<float4 *>(s->shapes.force);
        v.torque = reinterpret_cast<float4 *>(s->shapes.torque);
        v.sdfDims = reinterpret_cast<const int4 *>(s->sdf.dimension);
        v.sdfLbcs = reinterpret_cast<const float4 *>(s->sdf.lowerBoundCellSize);
        v.sdfData = reinterpret_cast<const float4 *>(s->sdf.gradientSignedDistance);
        v.effectiveMass = s->contactModel == CRMPM_HIP_CONTACT_EFFECTIVE_MASS;
        return v;
    }

    // Kernel launch on the solver stream.  While profiling, the dispatch itself
    // records a start and a stop event (hipExtLaunchKernelGGL), so the timing
    // covers the kernel's execution only.
    template <typename... Params, typename... Args>
    void launch(CrHipSolver *s, int cat, void (*kernel)(Params...), dim3 grid, dim3 block, Args... args)
    {
        if (s->profEvents && ((s->profMask >> cat) & 1u) && s->profUsed < s->profMax)
        {
            const int slot = s->profUsed++;
            s->profCat[slot] = cat;
            hipExtLaunchKernelGGL(kernel, grid, block, 0, s->stream, s->profEvents[2 * slot], s->profEvents[2 * slot + 1], 0,
                                  static_cast<Params>(args)...);
        }
        else
            hipLaunchKernelGGL(kernel, grid, block, 0, s->stream, static_cast<Params>(args)...);
    }

    // Keys per particle and a stable LSD radix sort of (key, index).  Returns
    // the sorted keys and particle order (one of the two ping-pong buffers).
    int sort_particles(CrHipSolver *s, int n, int useMask, const unsigned **orderOut, const unsigned **keysOut)
    {
        const GridGeom &g = s->g;
        int bits = 0;
        while ((1ll << bits) <= (long long)g.numBlocks)
            ++bits;
        const int passes = (bits + 7) / 8;
        const int numTiles = grid1d(n, kRadixTile);
        unsigned *kin = s->keysA, *vin = s->valsA, *kout = s->keysB, *vout = s->valsB;
        for (int pass = 0; pass < passes; ++pass)
        {
            const int shift = 8 * pass;
            if (pass == 0)
                launch(s, CRMPM_HIP_PROF_BINNING, k_bin_hist, numTiles, kWG, g, n, s->P, useMask, kin, vin, s->radixCounts,
                       numTiles);
            else
                launch(s, CRMPM_HIP_PROF_BINNING, k_radix_hist, numTiles, kWG, kin, n, shift, s->radixCounts, numTiles);
            launch(s, CRMPM_HIP_PROF_BINNING, k_radix_scan, 256, kWG, s->radixCounts, numTiles, s->digitTotals);
            launch(s, CRMPM_HIP_PROF_BINNING, k_radix_scatter, numTiles, kWG, kin, vin, n, shift, s->radixCounts, numTiles,
                   s->digitTotals, kout, vout);
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        *orderOut = vin;
        *keysOut = kin;
        return check_launch();
    }

    // Bin, sort and segment the particles taking part in this step.
    int build_tiles(CrHipSolver *s, int n, int useMask, const unsigned **orderOut, const unsigned **keysOut)
    {
        int rc = sort_particles(s, n, useMask, orderOut, keysOut);
        if (rc)
            return rc;
        launch(s, CRMPM_HIP_PROF_BINNING, k_segments, grid1d(n, 256), 256, n, s->g.numBlocks, *keysOut, s->segStart,
               s->segEnd);
        // the sort's spare ping-pong buffers take the cell ranks and the tile-ordered permutation
        unsigned *tmp = (*keysOut == s->keysA) ? s->keysB : s->keysA;
        unsigned *order2 = (*orderOut == s->valsA) ? s->valsB : s->valsA;
        launch(s, CRMPM_HIP_PROF_BINNING, k_cell_rank_order, grid1d(s->g.numBlocks, kWaves), kWG, s->g, s->segStart,
               s->segEnd, s->P.xm, *orderOut, tmp, order2);
        *orderOut = order2;
        return check_launch();
    }

    StepCommon common(CrHipSolver *s, const unsigned *order)
    {
        StepCommon c;
        c.g = s->g;
        c.gravity = s->gravity;
        c.dt = s->dt;
        c.P = s->P;
        c.S = SolverState{s->F, s->C, s->V0, s->maxParticles};
        c.sv = shape_view(s);
        c.segStart = s->segStart;
        c.segEnd = s->segEnd;
        c.order = order;
        c.tilebuf = s->tilebuf;
        c.grid = s->grid;
        return c;
    }
} // namespace

extern "C" {

int crHipSolverCreate(int kind, int numMaxParticles, float cellSize, const float boundMin[3], const float boundMax[3],
                      CrHipSolver **out)
{
    if (!out || !boundMin || !boundMax || numMaxParticles < 0 || !(cellSize > 0.0f))
        return CRMPM_HIP_ERR_ARG;
    if (kind != CRMPM_HIP_MLS && kind != CRMPM_HIP_PB && kind != CRMPM_HIP_STANDARD)
        return CRMPM_HIP_ERR_ARG;
    CrHipSolver *s = new (std::nothrow) CrHipSolver();
    if (!s)
        return CRMPM_HIP_ERR_ALLOC;
    s->kind = kind;
    s->maxParticles = numMaxParticles;
    GridGeom &g = s->g;
    g.bmin = make_float3(boundMin[0], boundMin[1], boundMin[2]);
    g.bmax = make_float3(boundMax[0], boundMax[1], boundMax[2]);
    g.dx = cellSize;
    g.invdx = 1 / cellSize;
    g.cellVol = cellSize * cellSize * cellSize;
    int n[3];
    for (int a = 0; a < 3; ++a)
        n[a] = (int)((boundMax[a] - boundMin[a]) / cellSize + 1.0f);
    g.nx = n[0];
    g.ny = n[1];
    g.nz = n[2];
    if (g.nx < 1 || g.ny < 1 || g.nz < 1)
    {
        delete s;
        return CRMPM_HIP_ERR_ARG;
    }
    g.nbx = (g.nx + 3) / 4;
    g.nby = (g.ny + 3) / 4;
    g.nbz = (g.nz + 3) / 4;
    g.numBlocks = g.nbx * g.nby * g.nbz;
    *out = s;
    return CRMPM_HIP_OK;
}

int crHipSolverDestroy(CrHipSolver *s)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    if (s->initialized)
        (void)hipStreamSynchronize(s->stream);
    dfree(s->F);
    dfree(s->C);
    dfree(s->V0);
    dfree(s->grid);
    dfree(s->gridForce);
    dfree(s->dirty);
    dfree(s->segStart);
    dfree(s->segEnd);
    dfree(s->keysA);
    dfree(s->keysB);
    dfree(s->valsA);
    dfree(s->valsB);
    dfree(s->radixCounts);
    dfree(s->digitTotals);
    dfree(s->tilebuf);
    dfree(s->tileForce);
    dfree(s->shapeRecs);
    dfree(s->shapeCounts);
    dfree(s->slicerList);
    dfree(s->stageMem);
    delete s;
    return CRMPM_HIP_OK;
}

int crHipSolverGetGridDims(const CrHipSolver *s, int dims[3])
{
    if (!s || !dims)
        return CRMPM_HIP_ERR_ARG;
    dims[0] = s->g.nx;
    dims[1] = s->g.ny;
    dims[2] = s->g.nz;
    return CRMPM_HIP_OK;
}

int crHipSolverSetStream(CrHipSolver *s, void *stream)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    s->stream = (hipStream_t)stream;
    return CRMPM_HIP_OK;
}

int crHipSolverSetGravity(CrHipSolver *s, float gx, float gy, float gz)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    s->gravity = make_float3(gx, gy, gz);
    return CRMPM_HIP_OK;
}

int crHipSolverSetIntegrationStepSize(CrHipSolver *s, float dt)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    s->dt = dt;
    return CRMPM_HIP_OK;
}

int crHipSolverSetSolverIterations(CrHipSolver *s, int n)
{
    if (!s || n < 0)
        return CRMPM_HIP_ERR_ARG;
    s->iterations = n;
    return CRMPM_HIP_OK;
}

int crHipSolverSetShapeContactModel(CrHipSolver *s, int model)
{
    if (!s || (model != CRMPM_HIP_CONTACT_KINEMATIC && model != CRMPM_HIP_CONTACT_EFFECTIVE_MASS))
        return CRMPM_HIP_ERR_ARG;
    s->contactModel = model;
    return CRMPM_HIP_OK;
}

int crHipSolverBindParticleData(CrHipSolver *s, float *positionMass, float *velocity, int size)
{
    if (!s || !positionMass || !velocity || size < s->maxParticles)
        return CRMPM_HIP_ERR_ARG;
    s->P.xm = reinterpret_cast<float4 *>(positionMass);
    s->P.v = reinterpret_cast<float4 *>(velocity);
    return CRMPM_HIP_OK;
}

int crHipSolverBindParticleMaterials(CrHipSolver *s, float *params0, int *type, int size)
{
    if (!s || !params0 || !type || size < s->maxParticles)
        return CRMPM_HIP_ERR_ARG;
    s->P.params = reinterpret_cast<const float4 *>(params0);
    s->P.type = type;
    return CRMPM_HIP_OK;
}

int crHipSolverBindActiveMask(CrHipSolver *s, unsigned char *activeMask)
{
    if (!s || !activeMask)
        return CRMPM_HIP_ERR_ARG;
    s->P.mask = activeMask;
    return CRMPM_HIP_OK;
}

int crHipSolverSetNumActiveParticles(CrHipSolver *s, unsigned int n)
{
    if (!s || n > (unsigned)s->maxParticles)
        return CRMPM_HIP_ERR_ARG;
    s->numActive = n;
    return CRMPM_HIP_OK;
}

int crHipSolverSetNumShapes(CrHipSolver *s, int n)
{
    if (!s || n < 0)
        return CRMPM_HIP_ERR_ARG;
    s->numShapes = n;
    return ensure_shape_capacity(s, n);
}

int crHipSolverBindShapeIds(CrHipSolver *s, int *shapeIds)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    s->shapeIds = shapeIds;
    return CRMPM_HIP_OK;
}

int crHipSolverBindShapeData(CrHipSolver *s, const CrHipShapeData *d)
{
    if (!s || !d)
        return CRMPM_HIP_ERR_ARG;
    s->shapes = *d;
    return CRMPM_HIP_OK;
}

int crHipSolverBindGeometryData(CrHipSolver *s, const CrHipGeometryData *d)
{
    if (!s || !d)
        return CRMPM_HIP_ERR_ARG;
    s->geoms = *d;
    return CRMPM_HIP_OK;
}

int crHipSolverBindGeometrySdfData(CrHipSolver *s, const CrHipGeometrySdfData *d)
{
    if (!s || !d)
        return CRMPM_HIP_ERR_ARG;
    s->sdf = *d;
    return CRMPM_HIP_OK;
}

int crHipSolverInitialize(CrHipSolver *s)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    if (s->initialized)
        return CRMPM_HIP_ERR_STATE;
    const size_t N = (size_t)std::max(s->maxParticles, 1);
    const GridGeom &g = s->g;
    const size_t nodes = (size_t)g.numBlocks * 64;
    // tile k is pblock k, so the partial buffers hold one 6^3 box per pblock
    s->maxTiles = g.numBlocks;
    const int radixTiles = grid1d((long long)N, kRadixTile);
    int rc = 0;
    rc |= dalloc(&s->F, 9 * N);
    rc |= dalloc(&s->C, 9 * N);
    rc |= dalloc(&s->V0, N);
    rc |= dalloc(&s->grid, nodes);
    if (s->kind == CRMPM_HIP_STANDARD)
    {
        rc |= dalloc(&s->gridForce, nodes);
        rc |= dalloc(&s->tileForce, (size_t)s->maxTiles * kTileNodes);
    }
    rc |= dalloc(&s->dirty, g.numBlocks);
    rc |= dalloc(&s->segStart, g.numBlocks);
    rc |= dalloc(&s->segEnd, g.numBlocks);
    rc |= dalloc(&s->keysA, N);
    rc |= dalloc(&s->keysB, N);
    rc |= dalloc(&s->valsA, N);
    rc |= dalloc(&s->valsB, N);
    rc |= dalloc(&s->radixCounts, (size_t)radixTiles * 256);
    rc |= dalloc(&s->digitTotals, 256);
    rc |= dalloc(&s->tilebuf, (size_t)s->maxTiles * kTileNodes);
    rc |= dalloc(&s->shapeCounts, 4);
    rc |= ensure_shape_capacity(s, std::max(s->numShapes, 1));
    if (s->kind == CRMPM_HIP_PB)
    {
        // xs (4) + vs (3) + Fs (9) + Cs (9) + ps (2) + ts (1) floats per particle
        rc |= dalloc(&s->stageMem, 28 * N);
        if (!rc)
        {
            float *m = s->stageMem;
            s->st.stride = (int)N;
            s->st.xs = reinterpret_cast<float4 *>(m);
            m += 4 * N;
            s->st.vs = m;
            m += 3 * N;
            s->st.Fs = m;
            m += 9 * N;
            s->st.Cs = m;
            m += 9 * N;
            s->st.ps = reinterpret_cast<float2 *>(m);
            m += 2 * N;
            s->st.ts = reinterpret_cast<int *>(m);
        }
    }
    if (rc)
        return CRMPM_HIP_ERR_ALLOC;
    hipStream_t st = s->stream;
    rc |= hipMemsetAsync(s->F, 0, 9 * N * sizeof(float), st);
    rc |= hipMemsetAsync(s->C, 0, 9 * N * sizeof(float), st);
    rc |= hipMemsetAsync(s->V0, 0, N * sizeof(float), st);
    rc |= hipMemsetAsync(s->grid, 0, nodes * sizeof(float4), st);
    if (s->gridForce)
        rc |= hipMemsetAsync(s->gridForce, 0, nodes * sizeof(float4), st);
    rc |= hipMemsetAsync(s->dirty, 0, g.numBlocks, st);
    rc |= hipMemsetAsync(s->segStart, 0, g.numBlocks * sizeof(int), st);
    rc |= hipMemsetAsync(s->segEnd, 0, g.numBlocks * sizeof(int), st);
    rc |= hipMemsetAsync(s->shapeCounts, 0, 4 * sizeof(int), st);
    if (rc)
        return CRMPM_HIP_ERR_LAUNCH;
    s->initialized = true;
    return CRMPM_HIP_OK;
}

int crHipSolverComputeInitialData(CrHipSolver *s, unsigned int n, const unsigned int *indices)
{
    if (!s || (n > 0 && !indices))
        return CRMPM_HIP_ERR_ARG;
    if (!s->initialized || !s->P.xm)
        return CRMPM_HIP_ERR_STATE;
    if (n == 0)
        return CRMPM_HIP_OK;
    SolverState S{s->F, s->C, s->V0, s->maxParticles};
    k_init_mass<<<grid1d(n, 256), 256, 0, s->stream>>>(s->g, n, indices, s->P.xm, S, s->grid, s->dirty);
    k_init_volume<<<grid1d(n, 256), 256, 0, s->stream>>>(s->g, n, indices, s->P.xm, S, s->grid);
    return check_launch();
}

int crHipSolverStep(CrHipSolver *s, float *dtOut)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    if (!s->initialized || !s->P.xm || !s->P.v || !s->P.params || !s->P.type || !s->P.mask)
        return CRMPM_HIP_ERR_STATE;
    if (s->numShapes > 0 && (!s->shapeIds || !s->shapes.type || !s->geoms.type))
        return CRMPM_HIP_ERR_STATE;
    if (dtOut)
        *dtOut = s->dt;
    const GridGeom &g = s->g;
    const int blockWGs = grid1d(g.numBlocks, kWaves);
    const ShapeTables tables{s->numShapes, s->shapeIds, s->shapes, s->geoms, s->shapeRecs, s->shapeCounts, s->slicerList};
    launch(s, CRMPM_HIP_PROF_BINNING, k_prepare, blockWGs + std::max(1, grid1d(s->numShapes, kWG)), kWG, g, blockWGs,
           s->grid, s->gridForce, s->dirty, s->segStart, s->segEnd, tables);
    const int n = (int)s->numActive;
    if (n == 0)
        return check_launch();
    const unsigned *order = nullptr, *keys = nullptr;
    // The standard solver's P2G ignores the active mask (src/solver/cpu_standard_mpm_solver.cpp:101-120).
    int rc = build_tiles(s, n, s->kind != CRMPM_HIP_STANDARD, &order, &keys);
    if (rc)
        return rc;
    StepCommon c = common(s, order);
    const int tileWGs = grid1d(g.numBlocks, kWaves);
    const int updWGs = grid1d(g.numBlocks, kWaves);
    const float4 *noForce = nullptr;
    if (s->kind == CRMPM_HIP_MLS)
    {
        launch(s, CRMPM_HIP_PROF_MLS_P2G, k_mls_p2g, tileWGs, kWG, c);
        launch(s, CRMPM_HIP_PROF_UPDATE, k_update<0>, updWGs, kWG, c, noForce, s->dirty);
        launch(s, CRMPM_HIP_PROF_MLS_G2P, k_mls_g2p, tileWGs, kWG, c);
    }
    else if (s->kind == CRMPM_HIP_STANDARD)
    {
        launch(s, CRMPM_HIP_PROF_STD_P2G, k_std_p2g, tileWGs, kWG, c, s->tileForce);
        launch(s, CRMPM_HIP_PROF_UPDATE, k_update<2>, updWGs, kWG, c, s->tileForce, s->dirty);
        launch(s, CRMPM_HIP_PROF_STD_G2P, k_std_g2p, tileWGs, kWG, c);
    }
    else
    {
        SolverState S{s->F, s->C, s->V0, s->maxParticles};
        launch(s, CRMPM_HIP_PROF_BINNING, k_pb_stage, grid1d(n, 256), 256, n, g.numBlocks, keys, order, s->P, S, s->st);
        const int K = s->iterations;
        if (K == 0)
        {
            launch(s, CRMPM_HIP_PROF_PB_G2P, k_pb_integrate_only, grid1d(n, 256), 256, c, n, keys, s->st, s->gravity);
        }
        else
        {
            launch(s, CRMPM_HIP_PROF_PB_P2G, k_pb_tile<0>, tileWGs, kWG, c, s->st, s->gravity);
            launch(s, CRMPM_HIP_PROF_UPDATE, k_update<1>, updWGs, kWG, c, noForce, s->dirty);
            for (int it = 1; it < K; ++it)
            {
                launch(s, CRMPM_HIP_PROF_PB_FUSED, k_pb_tile<1>, tileWGs, kWG, c, s->st, s->gravity);
                launch(s, CRMPM_HIP_PROF_UPDATE, k_update<1>, updWGs, kWG, c, noForce, s->dirty);
            }
            launch(s, CRMPM_HIP_PROF_PB_G2P, k_pb_tile<2>, tileWGs, kWG, c, s->st, s->gravity);
        }
    }
    return check_launch();
}

int crHipSolverBinParticles(CrHipSolver *s, unsigned int *keysOut, unsigned int *orderOut)
{
    if (!s || !keysOut || !orderOut)
        return CRMPM_HIP_ERR_ARG;
    if (!s->initialized || !s->P.xm || !s->P.mask)
        return CRMPM_HIP_ERR_STATE;
    const int n = (int)s->numActive;
    if (n == 0)
        return CRMPM_HIP_OK;
    const unsigned *order = nullptr, *keys = nullptr;
    int rc = sort_particles(s, n, s->kind != CRMPM_HIP_STANDARD, &order, &keys);
    if (rc)
        return rc;
    if (hipMemcpyAsync(keysOut, keys, n * sizeof(unsigned), hipMemcpyDeviceToDevice, s->stream) != hipSuccess ||
        hipMemcpyAsync(orderOut, order, n * sizeof(unsigned), hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
        return CRMPM_HIP_ERR_LAUNCH;
    return CRMPM_HIP_OK;
}

int crHipSolverProfileBegin(CrHipSolver *s, int maxLaunches)
{
    if (!s || maxLaunches <= 0)
        return CRMPM_HIP_ERR_ARG;
    if (s->profEvents)
        return CRMPM_HIP_ERR_STATE;
    s->profEvents = new (std::nothrow) hipEvent_t[2 * (size_t)maxLaunches];
    s->profCat = new (std::nothrow) int[maxLaunches];
    if (!s->profEvents || !s->profCat)
        return CRMPM_HIP_ERR_ALLOC;
    for (int i = 0; i < 2 * maxLaunches; ++i)
        if (hipEventCreate(&s->profEvents[i]) != hipSuccess)
            return CRMPM_HIP_ERR_ALLOC;
    s->profMax = maxLaunches;
    s->profUsed = 0;
    s->profMask = ~0u;
    return CRMPM_HIP_OK;
}

int crHipSolverProfileMask(CrHipSolver *s, unsigned int mask)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    if (!s->profEvents)
        return CRMPM_HIP_ERR_STATE;
    s->profMask = mask;
    return CRMPM_HIP_OK;
}

int crHipSolverProfileEnd(CrHipSolver *s, float *msPerCategory, int *launchesPerCategory, int numCategories)
{
    if (!s || !s->profEvents)
        return CRMPM_HIP_ERR_STATE;
    int rc = CRMPM_HIP_OK;
    if (hipStreamSynchronize(s->stream) != hipSuccess)
        rc = CRMPM_HIP_ERR_LAUNCH;
    for (int c = 0; c < numCategories; ++c)
    {
        if (msPerCategory)
            msPerCategory[c] = 0.0f;
        if (launchesPerCategory)
            launchesPerCategory[c] = 0;
    }
    for (int i = 0; i < s->profUsed && rc == CRMPM_HIP_OK; ++i)
    {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, s->profEvents[2 * i], s->profEvents[2 * i + 1]) != hipSuccess)
        {
            rc = CRMPM_HIP_ERR_LAUNCH;
            break;
        }
        const int c = s->profCat[i];
        if (c >= 0 && c < numCategories)
        {
            if (msPerCategory)
                msPerCategory[c] += ms;
            if (launchesPerCategory)
                launchesPerCategory[c] += 1;
        }
    }
    for (int i = 0; i < 2 * s->profMax; ++i)
        (void)hipEventDestroy(s->profEvents[i]);
    delete[] s->profEvents;
    delete[] s->profCat;
    s->profEvents = nullptr;
    s->profCat = nullptr;
    s->profMax = s->profUsed = 0;
    return rc;
}

int crHipSolverFetchResults(CrHipSolver *s)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    return hipStreamSynchronize(s->stream) == hipSuccess ? CRMPM_HIP_OK : CRMPM_HIP_ERR_LAUNCH;
}

int crHipSolverGetStateArrays(CrHipSolver *s, float **F, float **C, float **V0, int *stride)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    if (!s->initialized)
        return CRMPM_HIP_ERR_STATE;
    if (F)
        *F = s->F;
    if (C)
        *C = s->C;
    if (V0)
        *V0 = s->V0;
    if (stride)
        *stride = s->maxParticles;
    return CRMPM_HIP_OK;
}

int crHipSolverCopyGrid(CrHipSolver *s, float *dst)
{
    if (!s || !dst)
        return CRMPM_HIP_ERR_ARG;
    if (!s->initialized)
        return CRMPM_HIP_ERR_STATE;
    const long long total = (long long)s->g.nx * s->g.ny * s->g.nz;
    k_copy_grid<<<grid1d(total, 256), 256, 0, s->stream>>>(s->g, s->grid, reinterpret_cast<float4 *>(dst));
    return check_launch();
}

} // extern "C"
