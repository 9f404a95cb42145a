// MPM substep solver for MI355X (gfx950): kernels and the C-ABI of include/crmpm_hip.h.
//
// Data layout in HBM
//   engine-owned (bound):  positionMass float4[N], velocity float4[N], params0 float4[N],
//                          type int[N], mask u8[N], shape/geometry/SDF tables
//   solver-owned:          F, C  9 floats per particle, column-major (F[9p + 3c + r]),
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
// Substep (PB):   the same binning, then stage the particles in sorted order and run K
//                 iterations, each a per-particle kernel (G2P of the previous iteration and
//                 the material constraint, 64 sorted particles per wavefront) and a
//                 per-tile scatter (P2G), then the update; the last G2P integrates.
// P2G never issues global atomics and every sum has a fixed order, so a substep is bitwise
// reproducible run to run.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
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
    // Radix digits are 8 to 10 bits wide: the fewest passes that cover the
    // key range (pblock index and sentinel), each as narrow as that allows.
    constexpr int kMaxDigitBits = 10;
    constexpr int kMaxDigits = 1 << kMaxDigitBits;

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

    // Solver-owned particle state, particle order.  F and C are three float4
    // columns per particle (the reference's Mat3x3f layout), so a particle's
    // matrix is 48 contiguous bytes: the tile kernels visit particles in
    // sorted order, and per-particle records keep those scattered accesses
    // to few cache lines.
    struct SolverState
    {
        float *F;  // [9 * stride]: element (r, c) of particle p at 9 p + 3 c + r
        float *C;
        float *V0;
        int stride;
    };

    // One scene as the kernels see it: its grid geometry and parameters, and
    // the batch-wide arrays with the scene's pblock-indexed ones (segments,
    // tile partials, grid) offset to the scene's first pblock, so tile and
    // block indices below are scene-local.
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
        const unsigned *order;       // particle indices by tile, each tile in (cell rank, cell) order (global)
        const unsigned *gatherOrder; // the same tiles, each in particle index order (kernels that do not scatter)
        float4 *tilebuf;
        float4 *tileForce;
        float4 *grid;
        int subSize, subMax, subStride; // sub-tile split (see BatchArgs)
    };

    // Per-scene record in device memory.  A batch is a list of scenes whose
    // pblocks are numbered consecutively (scene s owns global pblocks
    // [blockOffset, blockOffset + g.numBlocks)) and whose particles are
    // slices of the bound particle arrays ([particleOffset, + numActive)).
    struct SceneRec
    {
        GridGeom g;
        float3 gravity;
        float dt;
        ShapeView sv;
        int blockOffset;
        int particleOffset;
        int numActive;
        int activeOffset; // sort item of the scene's first active particle
        // A resting scene sits out the substep: no particle is processed and
        // its grid blocks are neither cleared nor written, so its grid keeps
        // the last update's state (an engine scene that has taken its
        // substeps while the others of its batch finish).
        int resting;
    };

    // Shape tables bound to a scene (engine-owned), read when a step starts.
    struct SceneShapes
    {
        int numShapes;
        const int *shapeIds;
        CrHipShapeData sd;
        CrHipGeometryData gd;
    };

    // Kernel argument of the batched launches: every scene in one launch.
    struct BatchArgs
    {
        const SceneRec *scenes;
        const int *blockScene; // global pblock -> scene
        int numScenes;
        int totalBlocks;
        ParticleBind P;
        SolverState S;
        const int *segStart; // global pblock
        const int *segEnd;
        const unsigned *order;
        const unsigned *gatherOrder;
        // Node blocks that receive tile partials this substep, compacted:
        // entry (global block << 8 | mask of its 8 source pblocks holding
        // particles), activeCount entries.
        const unsigned *activeList;
        const int *activeCount;
        // Occupied pblocks (global keys), compacted, tileCount entries: the
        // items of the persistent tile loops.
        const unsigned *tileList;
        const int *tileCount;
        // When set, the grid update also writes each updated node here (the
        // batch's grid layout): the adjoint's tape of a PB iteration.
        float4 *tapeGrid;
        float4 *tilebuf;
        float4 *tileForce;
        float4 *grid;
        // A pblock's particles are split into sub-tiles of subSize particles
        // (a multiple of 64; the last of at most subMax sub-tiles takes the
        // rest), each run by its own wavefront with its own partial box at
        // slot pblock * subStride + j.
        int subSize, subMax, subStride;
        // Split of the gather kernels (G2P), which write no partials: a
        // pblock's particles in sub-tiles of gSubSize, at most gSubMax.
        int gSubSize, gSubMax;
        // Tiles a scatter launch takes (tile_class_ok): 0 all, 1 those of at
        // most 64 particles, 2 the others.
        int tileClass;
    };

    // Scene records are never written by a kernel: reading them through the
    // constant address space makes the loads invariant, so the compiler may
    // reload a field instead of holding it in registers for the whole kernel.
    // (The attribute only exists in the device compilation pass.)
#if defined(__HIP_DEVICE_COMPILE__)
    using ConstSceneRec = const __attribute__((address_space(4))) SceneRec;
#else
    using ConstSceneRec = const SceneRec;
#endif

    CR_DEV StepCommon scene_view_of(const BatchArgs &b, int s)
    {
        ConstSceneRec &r = ((ConstSceneRec *)(size_t)b.scenes)[s];
        StepCommon c;
        c.g = r.g;
        c.gravity = r.gravity;
        c.dt = r.dt;
        c.P = b.P;
        c.S = b.S;
        c.sv = r.sv;
        c.segStart = b.segStart + r.blockOffset;
        c.segEnd = b.segEnd + r.blockOffset;
        c.order = b.order;
        c.gatherOrder = b.gatherOrder;
        c.tilebuf = b.tilebuf + (size_t)r.blockOffset * b.subStride * kTileNodes;
        c.tileForce = b.tileForce ? b.tileForce + (size_t)r.blockOffset * b.subStride * kTileNodes : nullptr;
        c.grid = b.grid + (size_t)r.blockOffset * 64;
        c.subSize = b.subSize;
        c.subMax = b.subMax;
        c.subStride = b.subStride;
        return c;
    }

    // View of the scene owning global pblock tg; t receives the scene-local pblock.
    CR_DEV StepCommon scene_view(const BatchArgs &b, int tg, int &t)
    {
        // one scene: no map load in front of the dependent loads that follow
        const int s = b.numScenes == 1 ? 0 : __builtin_amdgcn_readfirstlane(b.blockScene[tg]);
        t = tg - ((ConstSceneRec *)(size_t)b.scenes)[s].blockOffset;
        return scene_view_of(b, s);
    }

    // Persistent tile loop: wavefront w of the launch visits global pblocks
    // w, w + W, w + 2W, ... (W = wavefronts in the launch), so a launch of any
    // size covers the batch and the autotuner picks the size.
    CR_DEV int first_tile(int wave) { return __builtin_amdgcn_readfirstlane((int)blockIdx.x * kWaves + wave); }
    // The same for a launch of Waves wavefronts per workgroup.  Launches with
    // one tile per wavefront use one-wavefront workgroups: a workgroup's LDS
    // is returned when all of its wavefronts are done, so in a four-wavefront
    // workgroup one tile with many particles would hold the other three
    // wavefronts' LDS (PB scatter on the slab: 23.3 -> 19.8 us).
    template <int Waves>
    CR_DEV int first_tile_of(int wave)
    {
        return __builtin_amdgcn_readfirstlane((int)blockIdx.x * Waves + wave);
    }
    CR_DEV int tile_stride() { return (int)gridDim.x * kWaves; }
    template <int Waves>
    CR_DEV int tile_stride_of()
    {
        return (int)gridDim.x * Waves;
    }
    CR_DEV int num_subtiles(const BatchArgs &b) { return b.totalBlocks * b.subMax; }
    // Whether sub-tile tg (global pblock tg / subMax, index tg % subMax) has
    // particles, from the global segment arrays alone (an empty sub-tile
    // skips the scene record loads).
    CR_DEV bool split_used(const BatchArgs &b, int tg, int subSize, int subMax)
    {
        const int pb = tg / subMax, j = tg - pb * subMax;
        const int count = b.segEnd[pb] - b.segStart[pb];
        return count > 0 && (j == 0 || count > j * subSize);
    }
    CR_DEV bool subtile_used(const BatchArgs &b, int tg) { return split_used(b, tg, b.subSize, b.subMax); }
    // Items of a tile loop.  Non-persistent (a launch of one wavefront per
    // slot): every (sub-)tile slot of the batch, each item its slot.
    // Persistent (a launch of any size whose wavefronts stride over the
    // items): the occupied pblocks' sub-tiles, from the compacted list
    // k_active_blocks builds, so no wavefront visits an empty pblock.
    template <bool Persistent>
    CR_DEV int loop_items(const BatchArgs &b, int slots, int subMax)
    {
        return Persistent ? __builtin_amdgcn_readfirstlane(*b.tileCount) * subMax : slots;
    }
    // The slot (global pblock * subMax + sub-tile) of item it.
    template <bool Persistent>
    CR_DEV int loop_tile(const BatchArgs &b, int it, int subMax)
    {
        if (!Persistent)
            return it;
        const int i = it / subMax;
        return __builtin_amdgcn_readfirstlane((int)b.tileList[i]) * subMax + (it - i * subMax);
    }
    // Next pblock of a tile loop; a non-persistent launch (one wavefront per
    // pblock) compiles to straight-line code, which needs fewer registers.
    template <bool Persistent, int Waves = kWaves>
    CR_DEV int next_tile(int tg, int totalBlocks)
    {
        return Persistent ? tg + tile_stride_of<Waves>() : totalBlocks;
    }

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

    // Solver-owned particle matrices (F, C): 9 floats per particle, element
    // (r, c) of particle p at 9 p + 3 c + r (36 bytes, the columns packed).
    CR_DEV M3 load_pm(const float *m, int p)
    {
        const float *a = m + 9 * (size_t)p;
        return M3{V3{a[0], a[1], a[2]}, V3{a[3], a[4], a[5]}, V3{a[6], a[7], a[8]}};
    }
    CR_DEV void store_pm(float *m, int p, const M3 &v)
    {
        float *a = m + 9 * (size_t)p;
        a[0] = v.c0.x;
        a[1] = v.c0.y;
        a[2] = v.c0.z;
        a[3] = v.c1.x;
        a[4] = v.c1.y;
        a[5] = v.c1.z;
        a[6] = v.c2.x;
        a[7] = v.c2.y;
        a[8] = v.c2.z;
    }
    // Three float4 columns per particle (the adjoints' own scratch arrays).
    CR_DEV M3 load_cols(const float4 *m, int p)
    {
        const float4 a = m[3 * p], b = m[3 * p + 1], c = m[3 * p + 2];
        return M3{V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}};
    }
    CR_DEV void store_cols(float4 *m, int p, const M3 &v)
    {
        m[3 * p] = make_float4(v.c0.x, v.c0.y, v.c0.z, 0.0f);
        m[3 * p + 1] = make_float4(v.c1.x, v.c1.y, v.c1.z, 0.0f);
        m[3 * p + 2] = make_float4(v.c2.x, v.c2.y, v.c2.z, 0.0f);
    }
    // Planes [9][stride] (element 3 column + row): the PB staging in sorted order.
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

    // A chunk's solver-owned matrices, written through LDS.  A store of one
    // 36-byte record per lane touches a cache line per lane and instruction;
    // so each lane puts its particle's C and F into the wavefront's stage in
    // chunk order, and lane j stores the chunk's j-th float: where the
    // chunk's particles are consecutive (index order), each store
    // instruction writes 256 contiguous bytes.  All 64 lanes call it;
    // p < 0: no particle.
    struct ChunkStage
    {
        float m[2][9 * 64]; // C, F: element e of the chunk's particle q at 9 q + e
        int p[64];
    };
    CR_DEV void store_chunk_mats(ChunkStage &S, int lane, int p, const M3 &Cm, const M3 &Fm, float *C, float *F)
    {
        S.p[lane] = p;
        const float cv[9] = {Cm.c0.x, Cm.c0.y, Cm.c0.z, Cm.c1.x, Cm.c1.y, Cm.c1.z, Cm.c2.x, Cm.c2.y, Cm.c2.z};
        const float fv[9] = {Fm.c0.x, Fm.c0.y, Fm.c0.z, Fm.c1.x, Fm.c1.y, Fm.c1.z, Fm.c2.x, Fm.c2.y, Fm.c2.z};
#pragma unroll
        for (int e = 0; e < 9; ++e)
        {
            S.m[0][9 * lane + e] = cv[e];
            S.m[1][9 * lane + e] = fv[e];
        }
        wave_sync();
#pragma unroll
        for (int i = 0; i < 9; ++i)
        {
            const int j = 64 * i + lane, q = j / 9, e = j - 9 * q;
            const int pq = S.p[q];
            if (pq >= 0)
            {
                C[9 * (size_t)pq + e] = S.m[0][j];
                F[9 * (size_t)pq + e] = S.m[1][j];
            }
        }
        wave_sync(); // the next chunk reuses the stage
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

    // Start of a step.  Workgroups [0, blockWGs): zero the grid blocks written
    // since the last step and mark every pblock empty (64 blocks per
    // wavefront).  Workgroup blockWGs + s copies scene s's bound shape tables into
    // its compact record array in shapeIds order and lists its slicers
    // (shapes whose geometry is a slicer).
    __global__ __launch_bounds__(kWG) void k_prepare(int totalBlocks, int blockWGs, float4 *grid, float4 *gridForce,
                                                     unsigned char *dirty, int *segStart, int *segEnd,
                                                     const SceneShapes *shapes, const SceneRec *scenes,
                                                     const int *blockScene, int numScenes, int *activeCount)
    {
        const int lane = threadIdx.x & 63;
        if ((int)blockIdx.x >= blockWGs)
        {
            const int sc = blockIdx.x - blockWGs;
            const SceneShapes st = shapes[sc];
            ShapeRec *recs = const_cast<ShapeRec *>(scenes[sc].sv.recs);
            int *counts = const_cast<int *>(scenes[sc].sv.counts);
            int *slicers = const_cast<int *>(scenes[sc].sv.slicers);
            for (int s = threadIdx.x; s < st.numShapes; s += kWG)
                cache_shape(s, st.shapeIds, st.sd, st.gd, recs);
            if (threadIdx.x == 0)
            {
                int ns = 0;
                for (int k = 0; k < st.numShapes; ++k)
                {
                    const int gt = st.gd.type[st.sd.geometryIdx[st.shapeIds[k]]];
                    if (gt == gQuadSlicer || gt == gTriangleMeshSlicer)
                        slicers[ns++] = k;
                }
                counts[0] = st.numShapes;
                counts[1] = ns;
            }
            return;
        }
        // a wavefront takes 64 node blocks: one lane per block resets its flags
        // and segment, then the whole wavefront clears each dirty block
        const int b0 = (blockIdx.x * kWaves + (threadIdx.x >> 6)) * 64;
        const int b = b0 + lane;
        if (b == 0)
        {
            activeCount[0] = 0; // active node blocks
            activeCount[1] = 0; // occupied pblocks
        }
        bool wasDirty = false;
        if (b < totalBlocks)
        {
            // a resting scene's blocks keep their contents and their flags
            const int sc = numScenes == 1 ? 0 : blockScene[b];
            if (!scenes[sc].resting)
            {
                wasDirty = dirty[b] != 0;
                dirty[b] = 0;
            }
            segStart[b] = 0;
            segEnd[b] = 0;
        }
        unsigned long long pending = __ballot(wasDirty);
        while (pending)
        {
            const int i = __ffsll(pending) - 1;
            pending &= pending - 1ull;
            const size_t node = (size_t)(b0 + i) * 64 + lane;
            grid[node] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (gridForce)
                gridForce[node] = make_float4(0.f, 0.f, 0.f, 0.f);
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

    // Histogram of one tile's digits (digit width `bits`), digit-major.
    template <int Bits>
    __global__ __launch_bounds__(kWG) void k_radix_hist(const unsigned *keys, int n, int shift, unsigned *counts,
                                                        int numTiles)
    {
        constexpr int digits = 1 << Bits;
        __shared__ unsigned h[digits];
        for (int d = threadIdx.x; d < digits; d += kWG)
            h[d] = 0;
        __syncthreads();
        const int base = blockIdx.x * kRadixTile;
#pragma unroll
        for (int r = 0; r < kRadixItems; ++r)
        {
            const int idx = base + r * kWG + threadIdx.x;
            if (idx < n)
                atomicAdd(&h[(keys[idx] >> shift) & (unsigned)(digits - 1)], 1u);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < digits; d += kWG)
            counts[d * numTiles + blockIdx.x] = h[d];
    }

    // Sort item i of a step is active particle i - activeOffset of the scene
    // whose [activeOffset, activeOffset + numActive) holds i.
    CR_DEV int scene_of_item(const int *activePrefix, int numScenes, int i)
    {
        int lo = 0, hi = numScenes - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (activePrefix[mid] <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        return lo;
    }

    // Key per sort item (the global pblock of its particle, or the sentinel
    // totalBlocks when the particle does not take part), the particle index,
    // and the histogram of the first radix digit: k_radix_hist fused with the
    // binning pass.
    template <int Bits>
    __global__ __launch_bounds__(kWG) void k_bin_hist(BatchArgs b, const int *activePrefix, int n, int useMask,
                                                      unsigned *keys, unsigned *vals, unsigned *counts, int numTiles)
    {
        constexpr int digits = 1 << Bits;
        __shared__ unsigned h[digits];
        for (int d = threadIdx.x; d < digits; d += kWG)
            h[d] = 0;
        __syncthreads();
        const int base = blockIdx.x * kRadixTile;
        auto bin = [&](int i, int sc) {
            ConstSceneRec &rec = ((ConstSceneRec *)(size_t)b.scenes)[sc];
            const int p = rec.particleOffset + (i - const_load(activePrefix, sc));
            unsigned key = (unsigned)b.totalBlocks;
            if (!useMask || b.P.mask[p])
            {
                const float4 x = b.P.xm[p];
                const GridGeom g = rec.g;
                int bx, by, bz;
                V3 fx;
                cell_of(g, v3(x), bx, by, bz, fx);
                key = (unsigned)(rec.blockOffset + pblock_key(g, bx, by, bz));
            }
            keys[i] = key;
            vals[i] = (unsigned)p;
            atomicAdd(&h[key & (unsigned)(digits - 1)], 1u);
        };
        // one scene: its record's fields are wavefront-uniform scalar loads
        if (b.numScenes == 1)
        {
#pragma unroll
            for (int r = 0; r < kRadixItems; ++r)
            {
                const int i = base + r * kWG + threadIdx.x;
                if (i < n)
                    bin(i, 0);
            }
        }
        else
        {
#pragma unroll
            for (int r = 0; r < kRadixItems; ++r)
            {
                const int i = base + r * kWG + threadIdx.x;
                if (i < n)
                    bin(i, scene_of_item(activePrefix, b.numScenes, i));
            }
        }
        __syncthreads();
        for (int d = threadIdx.x; d < digits; d += kWG)
            counts[d * numTiles + blockIdx.x] = h[d];
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

    // Stable scatter of one tile by the digit (key >> shift) & (2^bits - 1):
    // the tile's items go in rounds of 256; within a round, items of equal
    // digit keep their order (rank within the wavefront by ballots, then
    // wavefronts in order).  Thread t keeps the digits [t R, t R + R).
    template <int Bits>
    __global__ __launch_bounds__(kWG) void k_radix_scatter(const unsigned *keysIn, const unsigned *valsIn, int n, int shift,
                                                           const unsigned *counts, int numTiles,
                                                           const unsigned *digitTotals, unsigned *keysOut,
                                                           unsigned *valsOut)
    {
        constexpr int digits = 1 << Bits, R = digits / kWG;
        __shared__ unsigned digitOff[digits];
        __shared__ unsigned waveCnt[kWaves][digits];
        __shared__ unsigned roundTot[digits];
        __shared__ unsigned sh[kWaves];
        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        const int d0 = tid * R;
        unsigned mine = 0;
#pragma unroll
        for (int r = 0; r < R; ++r)
            mine += digitTotals[d0 + r];
        unsigned tot;
        unsigned ex = block_exclusive_scan(mine, tot, sh);
#pragma unroll
        for (int r = 0; r < R; ++r)
        {
            digitOff[d0 + r] = ex + counts[(d0 + r) * numTiles + blockIdx.x];
            ex += digitTotals[d0 + r];
#pragma unroll
            for (int w = 0; w < kWaves; ++w)
                waveCnt[w][d0 + r] = 0;
        }
        __syncthreads();
        const unsigned long long ltMask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        const int base = blockIdx.x * kRadixTile;
        for (int rd = 0; rd < kRadixItems; ++rd)
        {
            const int idx = base + rd * kWG + tid;
            const bool valid = idx < n;
            const unsigned key = valid ? keysIn[idx] : 0u;
            const unsigned val = valid ? valsIn[idx] : 0u;
            const unsigned d = (key >> shift) & (unsigned)(digits - 1);
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < Bits; ++bit)
            {
                const bool set = (d >> bit) & 1u;
                const unsigned long long bal = __ballot(valid && set);
                peers &= set ? bal : ~bal;
            }
            const unsigned rankInWave = (unsigned)__popcll(peers & ltMask);
            if (valid && rankInWave == 0)
                waveCnt[wave][d] = (unsigned)__popcll(peers);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                unsigned run = 0;
#pragma unroll
                for (int w = 0; w < kWaves; ++w)
                {
                    const unsigned c = waveCnt[w][d0 + r];
                    waveCnt[w][d0 + r] = run;
                    run += c;
                }
                roundTot[d0 + r] = run;
            }
            __syncthreads();
            if (valid)
            {
                const unsigned pos = digitOff[d] + waveCnt[wave][d] + rankInWave;
                keysOut[pos] = key;
                valsOut[pos] = val;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
            {
                digitOff[d0 + r] += roundTot[d0 + r];
#pragma unroll
                for (int w = 0; w < kWaves; ++w)
                    waveCnt[w][d0 + r] = 0;
            }
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
        int slot; // partial box of this (sub-)tile, scene-local
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
        ti.slot = t;
        return true;
    }

    // Number of sub-tiles of a pblock holding `count` particles.
    CR_DEV int subtiles_of(int count, int subSize, int subMax)
    {
        return count <= 0 ? 0 : min((count + subSize - 1) / subSize, subMax);
    }

    // Sub-tile j of scene-local pblock t: false when the pblock is empty or
    // has fewer sub-tiles.
    CR_DEV bool split_info(const StepCommon &c, int t, int j, int subSize, int subMax, TileInfo &ti)
    {
        const int start = c.segStart[t], end = c.segEnd[t];
        const int ns = subtiles_of(end - start, subSize, subMax);
        if (j >= ns)
            return false;
        ti.start = start + j * subSize;
        ti.end = (j == ns - 1) ? end : ti.start + subSize;
        decode_block(c.g, t, ti.bx, ti.by, ti.bz);
        ti.slot = t * c.subStride + j;
        return true;
    }
    CR_DEV bool subtile_info(const StepCommon &c, int t, int j, TileInfo &ti)
    {
        return split_info(c, t, j, c.subSize, c.subMax, ti);
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
    // grouped by base cell and each pass lets one lane per cell and
    // accumulation tile scatter with a plain read-modify-write (ds_read_b128 +
    // ds_write_b128 per node) instead of four ds_add_f32 atomics (which
    // measure 1.5x slower even when only the second and later lanes of a cell
    // use them, 4.5x when all lanes do).  The r-th lane of a cell adds into
    // tile r % Tiles in pass r / Tiles: with several tiles per wavefront, a
    // sparse tile (cells holding 0-3 particles) scatters in one or two passes
    // instead of one per particle of its fullest cell, and the tiles are
    // summed in a fixed order when written out.  Lanes of a cell take turns
    // in lane order, so the summation order is fixed and the result
    // reproducible.
    // ------------------------------------------------------------------
    CR_DEV void compiler_fence() { asm volatile("" ::: "memory"); }
    // Makes the wavefront wait for the load of v here.  A particle loop that
    // prefetches the next chunk's index settles the first one before the
    // loop: otherwise the wait at the loop header, which serves the entry and
    // the back edge alike, is sized for the entry (nothing issued after the
    // load) and waits for every outstanding access, the previous chunk's
    // stores included, on each iteration.
    CR_DEV void settle(int v) { asm volatile("" ::"v"(v)); }

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

    // Runs body(tile) on every active lane, in passes in which no two lanes
    // with the same tile share a key: the lane of rank r among the active
    // lanes of its key (in lane order) runs in pass r / Tiles with tile
    // r % Tiles.
    template <int Tiles, class Body>
    CR_DEV void in_conflict_free_passes(unsigned key, Body &&body)
    {
        const unsigned long long peers = match_lanes(key);
        const int lane = __lane_id();
        const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        const int rank = __popcll(peers & below);
        const int pass = rank / Tiles, tile = rank - pass * Tiles;
        for (int p = 0; __ballot(pass >= p) != 0ull; ++p)
        {
            if (pass == p)
                body(tile);
            compiler_fence();
        }
    }

    // Accumulation tiles per wavefront of MLS P2G.
    constexpr int kScatterTiles = 2;

    // Momentum-only accumulation node (the PB scatter of iterations after
    // the first: a pblock's node masses do not change over a substep's
    // iterations).
    struct F3
    {
        float x, y, z;
    };

    template <int Tiles>
    CR_DEV void zero_tiles(float4 *A, int lane)
    {
        for (int i = lane; i < Tiles * kTileNodes; i += 64)
            A[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    template <int Tiles>
    CR_DEV void zero_tiles(F3 *A, int lane)
    {
        for (int i = lane; i < Tiles * kTileNodes; i += 64)
            A[i] = F3{0.f, 0.f, 0.f};
    }
    // out = sum of the Tiles accumulation tiles, in tile order (Threads
    // threads share the nodes; tid is the thread's index among them).
    template <int Tiles, int Threads = 64>
    CR_DEV void store_tile_sum(const float4 *A, float4 *out, int tid)
    {
        for (int i = tid; i < kTileNodes; i += Threads)
        {
            float4 v = A[i];
#pragma unroll
            for (int j = 1; j < Tiles; ++j)
            {
                const float4 u = A[j * kTileNodes + i];
                v.x += u.x;
                v.y += u.y;
                v.z += u.z;
                v.w += u.w;
            }
            out[i] = v;
        }
    }
    // Momentum tiles: out's xyz = the sum in tile order; out's w (the node
    // mass the substep's first iteration stored there) is kept.
    template <int Tiles, int Threads = 64>
    CR_DEV void store_tile_sum(const F3 *A, float4 *out, int tid)
    {
        for (int i = tid; i < kTileNodes; i += Threads)
        {
            F3 v = A[i];
#pragma unroll
            for (int j = 1; j < Tiles; ++j)
            {
                const F3 u = A[j * kTileNodes + i];
                v.x += u.x;
                v.y += u.y;
                v.z += u.z;
            }
            *reinterpret_cast<F3 *>(out + i) = v;
        }
    }

    // The scatter of one (sub-)tile by the W wavefronts of a workgroup, for
    // dense pblocks (which otherwise run all their chunks on one wavefront):
    // wavefront w takes chunks w, w + W, ... into its own KAcc accumulation
    // tiles, and the workgroup writes the sum over wavefronts, then tiles, in
    // that fixed order to the tile's partial box.  A tile still writes one
    // partial box, so the grid update's work does not grow with W.
    // scatterOne(A, k) scatters sorted particle k into the tiles at A.
    template <int W, int KAcc, typename Acc, typename Fn>
    CR_DEV void wg_tile_scatter(Acc *acc, const TileInfo &ti, float4 *out, Fn &&scatterOne)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        Acc *A = acc + wave * KAcc * kTileNodes;
        zero_tiles<KAcc>(A, lane);
        wave_sync();
        for (int k = ti.start + 64 * wave + lane; k < ti.end; k += 64 * W)
            scatterOne(A, k);
        __syncthreads();
        store_tile_sum<W * KAcc, 64 * W>(acc, out, (int)threadIdx.x);
    }

    // Whether a scatter launch of tile class cls takes a (sub-)tile of count
    // particles: with adaptive wavefronts per tile, one launch takes the
    // tiles of one chunk on one wavefront each and another the larger ones
    // on four; a tile's class depends on its count alone, so its summation
    // order is fixed.
    CR_DEV bool tile_class_ok(int cls, int count)
    {
        return cls == 0 || ((count <= 64) == (cls == 1));
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
    CR_DEV void lds_add(F3 *A, int li, float x, float y, float z, float)
    {
        F3 cur = A[li];
        cur.x += x;
        cur.y += y;
        cur.z += z;
        A[li] = cur;
        compiler_fence();
    }

    // ------------------------------------------------------------------
    // Order of the particles inside each tile: by (rank pair, base cell,
    // rank), ranks counted in sorted order.  The tile kernels run 64 lanes
    // over consecutive particles of a tile, and lanes that share a base cell
    // take turns in the conflict-free scatter.  In this order each level
    // holds up to two particles per occupied cell (ranks 2p and 2p + 1, which
    // the two accumulation tiles take in one pass), so a dense tile scatters
    // in one or two passes per 64-lane chunk instead of one pass per particle
    // of its fullest cell, and a chunk reads both members of each pair of
    // neighbouring particles.  One wavefront per tile; tmp holds
    // (rank << 6 | cell) per sorted particle.
    // ------------------------------------------------------------------
    __global__ __launch_bounds__(kWG) void k_cell_rank_order(BatchArgs b, const unsigned *order, unsigned *tmp,
                                                             unsigned *orderOut)
    {
        // 128 rank levels cover cells of up to 127 particles (a body packed at
        // under a quarter of the cell spacing) without the per-cell fallback
        constexpr int kLevels = 128, kAhead = 4;
        __shared__ int cnt[kWaves][64];
        __shared__ unsigned long long levelMask[kWaves][kLevels];
        __shared__ int levelStart[kWaves][kLevels];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const float4 *xm = b.P.xm;
        // The lanes test the wavefront's next 64 pblocks (tg0, tg0 + W, ...)
        // for particles at once, and the wavefront visits only the occupied
        // ones: most of a batch of sparse grids is empty.
        const int W = tile_stride();
        for (int tg0 = first_tile(wave); tg0 < b.totalBlocks; tg0 += 64 * W)
        {
            const long long mineTile = (long long)tg0 + (long long)lane * W;
            const bool used = mineTile < b.totalBlocks && b.segEnd[mineTile] > b.segStart[mineTile];
            unsigned long long todo = __ballot(used);
            while (todo != 0ull)
            {
                const int tg = __builtin_amdgcn_readfirstlane(tg0 + (__ffsll((long long)todo) - 1) * W);
                todo &= todo - 1ull;
                int t;
                const StepCommon sv = scene_view(b, tg, t);
                const GridGeom &g = sv.g;
                TileInfo ti;
                if (!tile_info(g, t, sv.segStart, sv.segEnd, ti))
                    continue;
                wave_sync();
                int *C = cnt[wave];
                C[lane] = 0;
                wave_sync();
                const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
                // kAhead chunks' cells are loaded before their ranks are counted:
                // the order[k] -> xm gather chains of a dense tile overlap instead
                // of running one chunk at a time.
                for (int base = ti.start; base < ti.end; base += 64 * kAhead)
                {
                    unsigned cells[kAhead];
    #pragma unroll
                    for (int u = 0; u < kAhead; ++u)
                    {
                        const int k = base + 64 * u + lane;
                        cells[u] = 64u;
                        if (k < ti.end)
                        {
                            int bx, by, bz;
                            V3 fx;
                            cell_of(g, v3(xm[order[k]]), bx, by, bz, fx);
                            bx = min(max(bx, 0), g.nx - 1);
                            by = min(max(by, 0), g.ny - 1);
                            bz = min(max(bz, 0), g.nz - 1);
                            cells[u] = (unsigned)((bx & 3) | ((by & 3) << 2) | ((bz & 3) << 4));
                        }
                    }
    #pragma unroll
                    for (int u = 0; u < kAhead; ++u)
                    {
                        if (base + 64 * u >= ti.end)
                            break;
                        const int k = base + 64 * u + lane;
                        const bool valid = k < ti.end;
                        const unsigned cell = cells[u];
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
                }
                const int mine = C[lane];
                // Position of (rank r, cell) = particles of lower rank + rank-r
                // particles of lower cells.  Per rank level up to kLevels: the mask
                // of cells holding a rank-r particle and the level's start; deeper
                // levels (very dense cells) take the per-cell sum.
                int top = mine;
    #pragma unroll
                for (int o = 32; o > 0; o >>= 1)
                    top = max(top, __shfl_xor(top, o, 64));
                const int levels = min(top, kLevels);
                unsigned long long *LM = levelMask[wave];
                int *LS = levelStart[wave];
                int start = 0;
                for (int r = 0; r < levels; ++r)
                {
                    const unsigned long long m = __ballot(mine > r);
                    if (lane == 0)
                    {
                        LM[r] = m;
                        LS[r] = start;
                    }
                    start += __popcll(m);
                }
                wave_sync();
                for (int base = ti.start; base < ti.end; base += 64 * kAhead)
                {
                    unsigned rcs[kAhead], src[kAhead];
    #pragma unroll
                    for (int u = 0; u < kAhead; ++u)
                    {
                        const int k = base + 64 * u + lane;
                        rcs[u] = 0u;
                        src[u] = 0u;
                        if (k < ti.end)
                        {
                            rcs[u] = tmp[k];
                            src[u] = order[k];
                        }
                    }
    #pragma unroll
                    for (int u = 0; u < kAhead; ++u)
                    {
                        const int k = base + 64 * u + lane;
                        const bool valid = k < ti.end;
                        const unsigned rc = rcs[u];
                        const int r = (int)(rc >> 6), cell = (int)(rc & 63u);
                        const unsigned long long below = (cell == 0) ? 0ull : ((~0ull) >> (64 - cell));
                        int pos;
                        if (r < (levels & ~1))
                        {
                            // ranks in pairs: level p holds ranks 2p and 2p + 1 of each
                            // cell side by side, so a 64-lane chunk reads whole runs of
                            // neighbouring particles and still scatters in one pass of
                            // the two accumulation tiles
                            const int r2 = r & ~1;
                            pos = LS[r2] + __popcll(LM[r2] & below) + __popcll(LM[r2 + 1] & below) + (r & 1);
                        }
                        else if (r < levels)
                            pos = LS[r] + __popcll(LM[r] & below);
                        else
                        {
                            pos = 0;
                            for (int c = 0; c < 64; ++c)
                            {
                                const int nc = __builtin_amdgcn_readlane(mine, c);
                                pos += min(nc, r) + ((c < cell && nc > r) ? 1 : 0);
                            }
                        }
                        if (valid)
                            orderOut[ti.start + pos] = src[u];
                    }
                }
            }
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
    // node(a, b, c) returns node (a, b, c) of the particle's 3^3 stencil.
    template <class Node>
    CR_DEV void gather_separable_by(Node &&node, const Weights &W, const AxisOffsets &o, V3 &v, M3 &B)
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
                    const V3 nw = v3(node(a, b, c)) * (wab * W.w[c][2]);
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
    // The same sums with the stencil read from global memory one x-slice (9
    // nodes) at a time, at most two slices in flight: issued all at once the
    // 27 loads hold 81 registers, which costs the per-particle kernel a
    // wavefront per SIMD.  Same nodes, products and summation order as
    // gather_separable_by.
    template <class Node>
    CR_DEV void gather_separable_slices(Node &&node, const Weights &W, const AxisOffsets &o, V3 &v, M3 &B)
    {
        V3 sx[3], sy[3], sz[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            sx[i] = sy[i] = sz[i] = V3{0.f, 0.f, 0.f};
        float4 q0[9], q1[9];
        auto load = [&](int a, float4 *q) {
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    q[3 * b + c] = node(a, b, c);
        };
        auto add = [&](int a, const float4 *q) {
#pragma unroll
            for (int b = 0; b < 3; ++b)
            {
                const float wab = W.w[a][0] * W.w[b][1];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                {
                    const V3 nw = v3(q[3 * b + c]) * (wab * W.w[c][2]);
                    sx[a] = sx[a] + nw;
                    sy[b] = sy[b] + nw;
                    sz[c] = sz[c] + nw;
                }
            }
        };
        load(0, q0);
        load(1, q1);
        __builtin_amdgcn_sched_barrier(0);
        add(0, q0);
        __builtin_amdgcn_sched_barrier(0);
        load(2, q0);
        __builtin_amdgcn_sched_barrier(0);
        add(1, q1);
        add(2, q0);
        v = (sx[0] + sx[1]) + sx[2];
        auto col = [&](int k) {
            auto comp = [&](const V3 *s3, int r) {
                const float *p0 = &s3[0].x, *p1 = &s3[1].x, *p2 = &s3[2].x;
                return o.d[r][0] * p0[k] + o.d[r][1] * p1[k] + o.d[r][2] * p2[k];
            };
            return V3{comp(sx, 0), comp(sy, 1), comp(sz, 2)};
        };
        B = M3{col(0), col(1), col(2)};
    }
    // The stencil from the tile box staged in LDS (G), base cell at local index l0.
    CR_DEV void gather_separable(const float4 *G, int l0, const Weights &W, const AxisOffsets &o, V3 &v, M3 &B)
    {
        gather_separable_by([&](int a, int b, int c) { return G[node_li(l0, a, b, c)]; }, W, o, v, B);
    }

    // P2G terms (v + C dist) * w * m: C dist = c0 d_x + c1 d_y + c2 d_z, so
    // va[a] = v + c0 d[0][a], py[b] = c1 d[1][b], pz[c] = c2 d[2][c].
    struct ScatterTerms
    {
        V3 va[3], py[3], pz[3];
        float wxm[3], wy[3], wz[3]; // x weights times m (the caller passes m = 1 to keep them plain)
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
    // Scatter (term * w, w * mfac) to the 27 nodes, w = wxm[a] wy[b] wz[c].
    template <typename Acc>
    CR_DEV void scatter_separable(Acc *A, int l0, const ScatterTerms &t, float mfac)
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
                    lds_add(A, node_li(l0, a, b, c), contrib.x, contrib.y, contrib.z, wm * mfac);
                }
            }
    }

    // ------------------------------------------------------------------
    // MLS-MPM P2G (src/solver/mls_mpm_algorithm.h:147-241)
    // ------------------------------------------------------------------

    // One particle of MLS P2G (src/solver/mls_mpm_algorithm.h:147-241):
    // its momentum, mass and affine terms scattered into the tile's LDS
    // accumulation tiles A0 in conflict-free passes.
    CR_DEV void mls_p2g_particle(const StepCommon &c, const TileInfo &ti, float4 *A0, int nSlicers, int p)
    {
        const GridGeom &g = c.g;
        const float4 xm = c.P.xm[p];
        const float4 vv = c.P.v[p];
        const M3 F = load_pm(c.S.F, p);
        const M3 Cm = load_pm(c.S.C, p);
        const float4 prm = c.P.params[p];
        const int type = c.P.type[p];
        const float v0 = c.S.V0[p];
        const V3 x = v3(xm);
        const float m = xm.w;
        int bx, by, bz;
        V3 fx;
        cell_of(g, x, bx, by, bz, fx);
        // v += g dt before the scatter (Q9).  The reference writes it
        // back; only G2P's slicer reflection reads it before G2P
        // overwrites v, so G2P forms the same sum instead and P2G
        // issues no global store until its tile flush.
        const V3 vel = v3(vv) + v3(c.gravity.x, c.gravity.y, c.gravity.z) * c.dt;
        const M3 stress = stress_kirchhoff(prm.x, prm.y, type, F);
        const float sc = -c.dt * 4.0f * g.invdx * g.invdx * v0;
        const M3 affine = stress * sc + Cm * m;
        const V3 vm = vel * m;
        const bool fast = nodes_inside(g, bx, by, bz) && (nSlicers == 0 || slicer_free(c.sv, nSlicers, x, g.dx));
        const int l0 = fast ? base_li(ti, bx, by, bz) : 0;
        // The weights and the separable terms are formed inside each
        // pass from inputs made opaque there (a chunk takes one pass
        // in the rank-pair order), so 15 inputs stay live across the
        // passes instead of the 36 terms.
        M3 af = affine;
        V3 v_m = vm, f = fx;
        in_conflict_free_passes<kScatterTiles>(cell_key(ti, bx, by, bz), [&](int tile) {
            float4 *A = A0 + tile * kTileNodes;
            opaque(af);
            opaque(v_m);
            opaque(f);
            Weights w;
            weights_of(f, w);
            if (fast)
            {
                // (v m + A d) w and w m: the separable form with terms v m + A d, mass factor m
                scatter_separable(A, l0, scatter_terms(v_m, af, 1.0f, w, axis_offsets(f, g.dx)), m);
                return;
            }
            auto sc = [&](int li, int, int, int, V3 dist, float wt) {
                const V3 contrib = (v_m + mul(af, dist)) * wt;
                lds_add(A, li, contrib.x, contrib.y, contrib.z, wt * m);
            };
            for_each_node<false>(g, ti, bx, by, bz, f, w, [&](int li, int a, int b, int cc, V3 dist, float wt) {
                V3 nrm;
                if (nSlicers > 0 && different_sides<false>(c.sv, nSlicers, x, x + dist, nrm))
                    return;
                sc(li, a, b, cc, dist, wt);
            });
        });
    }

    template <bool Persistent, int Waves = kWaves>
    __global__ __launch_bounds__(64 * Waves) void k_mls_p2g(BatchArgs b)
    {
        static_assert(Persistent ? Waves == kWaves : true, "persistent tile loops stride by kWaves per workgroup");
        __shared__ float4 acc[Waves * kScatterTiles][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nIt = loop_items<Persistent>(b, num_subtiles(b), b.subMax);
        for (int it = first_tile_of<Waves>(wave); it < nIt; it = next_tile<Persistent, Waves>(it, nIt))
        {
            const int tg = loop_tile<Persistent>(b, it, b.subMax);
            if (!subtile_used(b, tg))
                continue;
            int t;
            const StepCommon c = scene_view(b, tg / b.subMax, t);
            TileInfo ti;
            if (!subtile_info(c, t, tg % b.subMax, ti))
                continue;
            if (!tile_class_ok(b.tileClass, ti.end - ti.start))
                continue;
            wave_sync(); // the previous tile's LDS reads are done
            float4 *A0 = acc[wave * kScatterTiles];
            zero_tiles<kScatterTiles>(A0, lane);
            wave_sync();
            const int nSlicers = slicer_count(c.sv);
            // the next chunk's particle index is loaded one chunk ahead
            int pNext = (int)c.order[min(ti.start + lane, ti.end - 1)];
            settle(pNext);
            for (int k = ti.start + lane; k < ti.end; k += 64)
            {
                const int p = pNext;
                pNext = (int)c.order[min(k + 64, ti.end - 1)];
                mls_p2g_particle(c, ti, A0, nSlicers, p);
            }
            wave_sync();
            store_tile_sum<kScatterTiles>(A0, c.tilebuf + (size_t)ti.slot * kTileNodes, lane);
        }
    }

    // k_mls_p2g with W wavefronts per (sub-)tile (wg_tile_scatter), one
    // workgroup per tile slot: for dense pblocks, whose chunks otherwise run
    // on one wavefront.
    template <int W>
    __global__ __launch_bounds__(64 * W) void k_mls_p2g_wg(BatchArgs b)
    {
        __shared__ float4 acc[W * kScatterTiles][kTileNodes];
        const int tg = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
        if (tg >= num_subtiles(b) || !subtile_used(b, tg))
            return;
        int t;
        const StepCommon c = scene_view(b, tg / b.subMax, t);
        TileInfo ti;
        if (!subtile_info(c, t, tg % b.subMax, ti))
            return;
        if (!tile_class_ok(b.tileClass, ti.end - ti.start))
            return;
        const int nSlicers = slicer_count(c.sv);
        wg_tile_scatter<W, kScatterTiles>(acc[0], ti, c.tilebuf + (size_t)ti.slot * kTileNodes, [&](float4 *A, int k) {
            mls_p2g_particle(c, ti, A, nSlicers, (int)c.order[k]);
        });
    }

    // ------------------------------------------------------------------
    // Grid update over active node blocks: tile sums in a fixed order, then
    // mlsMpmUpdateGrid / pbMpmUpdateGrid / standardMpmUpdateGrid
    // (src/solver/mls_mpm_algorithm.h:243-294, pb 233-284, standard 217-272).
    // ------------------------------------------------------------------
    // Node blocks whose 6^3-box sources (the pblocks nb - (dx, dy, dz), d in
    // {0, 1}^3) hold particles this substep, listed once per substep for the
    // grid updates that follow (20 per substep in PB-MPM).  One thread per
    // global block; each wavefront appends its blocks with one atomic, so the
    // list's order varies between runs while each block's result does not.
    // Also appends each occupied pblock (bit 0 of its own node block's mask)
    // to tiles (count[1] entries): the items of the persistent tile loops.
    __global__ __launch_bounds__(256) void k_active_blocks(BatchArgs b, unsigned *list, int *count, unsigned *tiles)
    {
        const int nbg = blockIdx.x * blockDim.x + threadIdx.x;
        const int lane = threadIdx.x & 63;
        unsigned mask = 0u;
        if (nbg < b.totalBlocks)
        {
            const int s = (b.numScenes == 1) ? 0 : b.blockScene[nbg];
            ConstSceneRec &r = ((ConstSceneRec *)(size_t)b.scenes)[s];
            const GridGeom g = r.g;
            const int off = r.blockOffset;
            int bx, by, bz;
            decode_block(g, nbg - off, bx, by, bz);
#pragma unroll
            for (int d = 0; d < 8; ++d)
            {
                const int px = bx - (d & 1), py = by - ((d >> 1) & 1), pz = bz - (d >> 2);
                if (px >= 0 && py >= 0 && pz >= 0)
                {
                    const int key = off + block_key(g, px, py, pz);
                    if (b.segEnd[key] > b.segStart[key])
                        mask |= 1u << d;
                }
            }
        }
        const unsigned long long act = __ballot(mask != 0u);
        int base = 0;
        if (lane == 0 && act != 0ull)
            base = atomicAdd(count, __popcll(act));
        base = __shfl(base, 0, 64);
        const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
        if (mask != 0u)
            list[base + __popcll(act & below)] = ((unsigned)nbg << 8) | mask;
        const unsigned long long occ = __ballot((mask & 1u) != 0u);
        int tbase = 0;
        if (lane == 0 && occ != 0ull)
            tbase = atomicAdd(count + 1, __popcll(occ));
        tbase = __shfl(tbase, 0, 64);
        if (mask & 1u)
            tiles[tbase + __popcll(occ & below)] = (unsigned)nbg;
    }

    // Multi: pblocks may have more than one sub-tile (subMax > 1); the
    // single-sub-tile variant needs fewer registers.  The kernel is bound by
    // the latency of its dependent loads (list entry, then partials), so it
    // is capped at 80 VGPRs for 6 wavefronts per SIMD.  It visits the listed
    // blocks only; unlisted blocks receive no partial and keep their cleared
    // state.
    template <int Kind, bool Multi, bool Persistent> // Kind: 0 MLS, 1 PB, 2 standard
    __global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(6))) void k_update(BatchArgs b,
                                                                                            unsigned char *dirty)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nActive = __builtin_amdgcn_readfirstlane(*b.activeCount);
        for (int item = first_tile(wave); item < nActive; item = next_tile<Persistent>(item, nActive))
        {
            const unsigned entry = __builtin_amdgcn_readfirstlane(b.activeList[item]);
            const int nbg = (int)(entry >> 8);
            int nb;
            const StepCommon c = scene_view(b, nbg, nb);
            const GridGeom &g = c.g;
            int bx, by, bz;
            decode_block(g, nb, bx, by, bz);
            // Tiles whose 6^3 box covers this block: pblocks nb - (dx, dy, dz),
            // those holding particles flagged in the entry.
            int nsub[8];
#pragma unroll
            for (int d = 0; d < 8; ++d)
            {
                const bool src = (entry >> d) & 1u;
                if (Multi)
                {
                    const int key = src ? block_key(g, bx - (d & 1), by - ((d >> 1) & 1), bz - (d >> 2)) : 0;
                    nsub[d] = src ? subtiles_of(c.segEnd[key] - c.segStart[key], c.subSize, c.subMax) : 0;
                }
                else
                    nsub[d] = src ? 1 : 0;
            }
            const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
            // First sub-tile of every source pblock: all loads issued before any
            // is summed.  Box pointers and sub-tile counts are wave-uniform; only
            // the node offset inside a box is per lane.
            float4 q[8], qf[8];
            bool more = false;
#pragma unroll
            for (int d = 0; d < 8; ++d)
            {
                const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
                const int ux = lx + 4 * dx, uy = ly + 4 * dy, uz = lz + 4 * dz;
                const bool use = nsub[d] > 0 && ux < kTN && uy < kTN && uz < kTN;
                const size_t box = (size_t)(nb - dx - g.nbx * (dy + g.nby * dz)) * c.subStride * kTileNodes;
                const int li = ux + kTN * (uy + kTN * uz);
                more = more || nsub[d] > 1;
                q[d] = use ? (c.tilebuf + box)[li] : make_float4(0.f, 0.f, 0.f, 0.f);
                if (Kind == 2)
                    qf[d] = use ? (c.tileForce + box)[li] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            // fixed summation order: d = 0..7, each followed by its further sub-tiles (bitwise reproducible)
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
                const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
                const int ux = lx + 4 * dx, uy = ly + 4 * dy, uz = lz + 4 * dz;
                if (Multi && more && ux < kTN && uy < kTN && uz < kTN)
                    for (int j = 1; j < nsub[d]; ++j)
                    {
                        const size_t at = ((size_t)(nb - dx - g.nbx * (dy + g.nby * dz)) * c.subStride + j) * kTileNodes +
                                          (ux + kTN * (uy + kTN * uz));
                        const float4 r = c.tilebuf[at];
                        s.x += r.x;
                        s.y += r.y;
                        s.z += r.z;
                        s.w += r.w;
                        if (Kind == 2)
                        {
                            const float4 rf = c.tileForce[at];
                            f.x += rf.x;
                            f.y += rf.y;
                            f.z += rf.z;
                        }
                    }
            }
            if (lane == 0)
                dirty[nbg] = 1;
            const int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
            const bool inGrid = in_grid(g, i, j, k);
            const float cutoff = (Kind == 1) ? 1e-8f : 1e-4f;
            const bool heavy = inGrid && !(s.w < cutoff);
            V3 np{0, 0, 0};
            if (heavy)
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
                np = V3{(float)i * g.dx, (float)j * g.dx, (float)k * g.dx} + v3(g.bmin.x, g.bmin.y, g.bmin.z);
                np = np + v3(s) * c.dt;
            }
            // all lanes take part (coupling momentum is summed per wavefront); light nodes are left as they are
            resolve_rigid<false>(c.sv, shape_count(c.sv), s, np, heavy);
            if (heavy)
            {
                if (i < 2 || i > g.nx - 3)
                    s.x = 0.0f;
                if (j < 2 || j > g.ny - 3)
                    s.y = 0.0f;
                if (k < 2 || k > g.nz - 3)
                    s.z = 0.0f;
            }
            if (inGrid)
            {
                c.grid[nb * 64 + lane] = s;
                if (b.tapeGrid)
                    b.tapeGrid[(size_t)nbg * 64 + lane] = s;
            }
        }
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
        if (nodes_inside(g, bx, by, bz) && (nSlicers == 0 || slicer_free(c.sv, nSlicers, x, g.dx)))
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

    // The same gather read straight from the scene's grid in global memory
    // (the per-particle PB kernels, which stage no tile): the same nodes,
    // sums and order as gather_vc.  Every in-grid node of a particle lies in
    // its tile's box, so skipping out-of-grid nodes is the tile's test.
    CR_DEV void gather_vc_grid(const StepCommon &c, int nSlicers, V3 x, V3 vp, V3 &vOut, M3 &cOut)
    {
        const GridGeom &g = c.g;
        int bx, by, bz;
        V3 fx;
        cell_of(g, x, bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        V3 vacc{0, 0, 0};
        M3 aacc{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}};
        if (nodes_inside(g, bx, by, bz) && (nSlicers == 0 || slicer_free(c.sv, nSlicers, x, g.dx)))
        {
            // node (x, y, z) is at block (x>>2, y>>2, z>>2) * 64 + (z&3) * 16 + (y&3) * 4 + (x&3):
            // a sum of per-axis parts
            int ox[3], oy[3], oz[3];
            const int rowY = g.nbx * 64, rowZ = g.nbx * g.nby * 64;
#pragma unroll
            for (int i = 0; i < 3; ++i)
            {
                const int X = bx + i, Y = by + i, Z = bz + i;
                ox[i] = ((X >> 2) << 6) + (X & 3);
                oy[i] = (Y >> 2) * rowY + ((Y & 3) << 2);
                oz[i] = (Z >> 2) * rowZ + ((Z & 3) << 4);
            }
            const float4 *grid = c.grid;
            gather_separable_slices([&](int a, int b, int cc) { return grid[ox[a] + oy[b] + oz[cc]]; }, W,
                                    axis_offsets(fx, g.dx), vacc, aacc);
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
                        const V3 dist = (V3{(float)a, (float)b, (float)cc} - fx) * g.dx;
                        const float wt = w_of(fx.x, a) * w_of(fx.y, b) * w_of(fx.z, cc);
                        V3 nv = v3(c.grid[grid_index(g, nx, ny, nz)]);
                        V3 n;
                        if (nSlicers > 0 && different_sides<true>(c.sv, nSlicers, x, x + dist, n))
                            nv = vp - n * dot(vp, n) * 2.0f;
                        const V3 nw = nv * wt;
                        vacc = vacc + nw;
                        aacc = aacc + outer_ref(nw, dist);
                    }
        }
        vOut = vacc;
        cOut = aacc * (4.0f * g.invdx);
    }

    // Standard MPM's gather (v and the velocity gradient) read straight from
    // the scene's grid in global memory: the same nodes, terms and order as
    // k_std_g2p's LDS gather.
    CR_DEV void gather_vgrad_grid(const StepCommon &c, V3 x, V3 &vOut, M3 &gvOut)
    {
        const GridGeom &g = c.g;
        int bx, by, bz;
        V3 fx;
        cell_of(g, x, bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        V3 vacc{0, 0, 0};
        M3 gv{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}};
        auto gather = [&](V3 nv, int a, int b, int cc, float wt) {
            vacc = vacc + nv * wt;
            const float wa = w_of(fx.x, a), wb = w_of(fx.y, b), wc = w_of(fx.z, cc);
            const V3 wg{dw_of(fx.x, a) * wb * wc, wa * dw_of(fx.y, b) * wc, wa * wb * dw_of(fx.z, cc)};
            gv = gv + outer_ref(wg, nv);
        };
        if (nodes_inside(g, bx, by, bz))
        {
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
                        gather(v3(c.grid[grid_index(g, bx + a, by + b, bz + cc)]), a, b, cc,
                               W.w[a][0] * W.w[b][1] * W.w[cc][2]);
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
                        gather(v3(c.grid[grid_index(g, nx, ny, nz)]), a, b, cc,
                               w_of(fx.x, a) * w_of(fx.y, b) * w_of(fx.z, cc));
                    }
        }
        vOut = vacc;
        gvOut = gv;
    }

    CR_DEV V3 clamp_bounds(const GridGeom &g, V3 x)
    {
        return V3{fminf(fmaxf(x.x, g.bmin.x), g.bmax.x), fminf(fmaxf(x.y, g.bmin.y), g.bmax.y),
                  fminf(fmaxf(x.z, g.bmin.z), g.bmax.z)};
    }

    // MLS G2P (src/solver/mls_mpm_algorithm.h:296-390)
    template <bool Persistent, int Waves = kWaves>
    __global__ __launch_bounds__(64 * Waves) void k_mls_g2p(BatchArgs b)
    {
        static_assert(Persistent ? Waves == kWaves : true, "persistent tile loops stride by kWaves per workgroup");
        __shared__ float4 lg[Waves][kTileNodes];
        __shared__ ChunkStage stage[Waves];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nIt = loop_items<Persistent>(b, b.totalBlocks * b.gSubMax, b.gSubMax);
        for (int it = first_tile_of<Waves>(wave); it < nIt; it = next_tile<Persistent, Waves>(it, nIt))
        {
            const int tg = loop_tile<Persistent>(b, it, b.gSubMax);
            if (!split_used(b, tg, b.gSubSize, b.gSubMax))
                continue;
            int t;
            const StepCommon c = scene_view(b, tg / b.gSubMax, t);
            TileInfo ti;
            if (!split_info(c, t, tg % b.gSubMax, b.gSubSize, b.gSubMax, ti))
                continue;
            wave_sync(); // the previous tile's LDS reads are done
            float4 *G = lg[wave];
            load_tile_grid(c.g, ti, c.grid, G, lane);
            wave_sync();
            const int nSlicers = slicer_count(c.sv);
            const int nShapes = shape_count(c.sv);
            // the next chunk's particle index is loaded one chunk ahead, and a
            // particle's x, v and F are loaded together before any store; all
            // lanes run every chunk (the staged C and F stores need them)
            int pNext = (int)c.gatherOrder[min(ti.start + lane, ti.end - 1)];
            settle(pNext);
            for (int k0 = ti.start; k0 < ti.end; k0 += 64)
            {
                const int k = k0 + lane;
                const bool valid = k < ti.end;
                const int p = pNext;
                pNext = (int)c.gatherOrder[min(k + 64, ti.end - 1)];
                M3 Cm = m3_identity(), Fn = m3_identity();
                if (valid)
                {
                    const float4 xm = c.P.xm[p];
                    const float4 vp = c.P.v[p];
                    const M3 F = load_pm(c.S.F, p);
                    V3 x = v3(xm);
                    V3 vel;
                    // the particle velocity after P2G's gravity update (P2G does not store it)
                    const V3 vpg = v3(vp) + v3(c.gravity.x, c.gravity.y, c.gravity.z) * c.dt;
                    gather_vc(c, ti, G, nSlicers, x, vpg, vel, Cm);
                    x = clamp_bounds(c.g, x + vel * c.dt);
                    float4 vv = make_float4(vel.x, vel.y, vel.z, 0.0f);
                    resolve_rigid<true>(c.sv, nShapes, vv, x);
                    c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
                    c.P.v[p] = make_float4(vel.x, vel.y, vel.z, 0.0f);
                    Fn = mul(m3_identity() + Cm * c.dt, F);
                }
                store_chunk_mats(stage[wave], lane, valid ? p : -1, Cm, Fn, c.S.C, c.S.F);
            }
        }
    }

    // G2P per particle in particle index order, for a batch of one scene
    // (the gather split "per particle"): 64 consecutive particles per
    // wavefront, the stencil read from the grid in global memory
    // (gather_vc_grid: the same nodes, sums and order as the tile kernels'
    // LDS gathers), so every particle array is read and written in whole
    // cache lines instead of the tile's rows of 8 particles.  The same
    // particles take part as in the tile kernels (the binning's test:
    // MLS: in the active mask; standard: P2G takes every particle, G2P the
    // masked-in ones), and the results are bitwise those of the tile kernels.
    // i counts the scene's particles; n is 0 for a resting scene.
    template <int Kind> // 0 MLS, 2 standard
    __global__ __launch_bounds__(kWG) void k_g2p_dense(BatchArgs b, int n)
    {
        __shared__ ChunkStage stage[kWaves];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int i = blockIdx.x * kWG + threadIdx.x;
        if (__builtin_amdgcn_readfirstlane(i - lane) >= n)
            return; // the whole wavefront is past the end
        const StepCommon c = scene_view_of(b, 0);
        const GridGeom &g = c.g;
        const int p = ((ConstSceneRec *)(size_t)b.scenes)[0].particleOffset + i;
        const bool valid = i < n && c.P.mask[p];
        const int nSlicers = slicer_count(c.sv);
        const int nShapes = shape_count(c.sv);
        M3 Cm = m3_identity(), Fn = m3_identity();
        if (valid)
        {
            const float4 xm = c.P.xm[p];
            // the incoming velocity only enters a slicer's reflection
            const float4 vp = (Kind == 0 && nSlicers > 0) ? c.P.v[p] : make_float4(0.f, 0.f, 0.f, 0.f);
            const M3 F = load_pm(c.S.F, p);
            V3 x = v3(xm);
            V3 vel;
            if (Kind == 0)
            {
                // the particle velocity after P2G's gravity update (P2G does not store it)
                const V3 vpg = v3(vp) + v3(c.gravity.x, c.gravity.y, c.gravity.z) * c.dt;
                gather_vc_grid(c, nSlicers, x, vpg, vel, Cm);
                Fn = mul(m3_identity() + Cm * c.dt, F);
            }
            else
            {
                M3 gv;
                gather_vgrad_grid(c, x, vel, gv);
                Fn = mul(m3_identity() + gv * c.dt, F);
            }
            x = clamp_bounds(g, x + vel * c.dt);
            float4 vv = make_float4(vel.x, vel.y, vel.z, 0.0f);
            resolve_rigid<true>(c.sv, nShapes, vv, x);
            c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
            c.P.v[p] = make_float4(vel.x, vel.y, vel.z, 0.0f);
        }
        if (Kind == 0)
            store_chunk_mats(stage[wave], lane, valid ? p : -1, Cm, Fn, c.S.C, c.S.F);
        else if (valid)
            store_pm(c.S.F, p, Fn);
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
        store_m3(st.Fs, st.stride, k, load_pm(S.F, p));
        store_m3(st.Cs, st.stride, k, load_pm(S.C, p));
        const float4 prm = P.params[p];
        st.ps[k] = make_float2(prm.x, prm.y);
        st.ts[k] = P.type[p];
    }

    // Scatter of one PB particle into the LDS tiles A0[0 .. Tiles)
    // (pbMpmParticleToGrid loop)
    template <int Tiles, typename Acc>
    CR_DEV void pb_scatter(const StepCommon &c, const TileInfo &ti, Acc *A0, int nSlicers, V3 x, float m, V3 vel,
                           const M3 &Cm)
    {
        const GridGeom &g = c.g;
        int bx, by, bz;
        V3 fx;
        cell_of(g, x, bx, by, bz, fx);
        const bool fast = nodes_inside(g, bx, by, bz) && (nSlicers == 0 || slicer_free(c.sv, nSlicers, x, g.dx));
        const int l0 = fast ? base_li(ti, bx, by, bz) : 0;
        // The weights and per-node terms are formed inside each pass (a
        // pblock takes one or two) from inputs made opaque there, so only the
        // 16 inputs stay live across passes, not the 36 separable terms.
        M3 cm = Cm;
        V3 v = vel, f = fx;
        in_conflict_free_passes<Tiles>(cell_key(ti, bx, by, bz), [&](int tile) {
            Acc *A = A0 + tile * kTileNodes;
            opaque(cm);
            opaque(v);
            opaque(f);
            Weights W;
            weights_of(f, W);
            if (fast)
            {
                scatter_separable(A, l0, scatter_terms(v, cm, m, W, axis_offsets(f, g.dx)), 1.0f);
                return;
            }
            auto scatter = [&](int li, int, int, int, V3 dist, float wt) {
                const float wm = wt * m;
                const V3 contrib = (v + mul(cm, dist)) * wm;
                lds_add(A, li, contrib.x, contrib.y, contrib.z, wm);
            };
            for_each_node<false>(g, ti, bx, by, bz, f, W, [&](int li, int a, int b, int cc, V3 dist, float wt) {
                V3 nrm;
                if (nSlicers > 0 && different_sides<false>(c.sv, nSlicers, x, x + dist, nrm))
                    return;
                scatter(li, a, b, cc, dist, wt);
            });
        });
    }

    // ------------------------------------------------------------------
    // PB iterations split by where the lanes are full.  A slab-like body
    // holds about 37 particles per pblock, so a wavefront that takes one
    // pblock keeps 57% of its lanes busy, and the per-particle work (the
    // gather and above all the SVD-based constraint) dominates.  So each
    // iteration runs the per-particle half over 64 consecutive sorted
    // particles per wavefront (k_pb_dense: every lane busy, the stencil read
    // from the grid in global memory, no LDS) and the scatter half per pblock
    // (k_pb_scatter: LDS accumulation tiles only, few registers).  The staged
    // v and C carry the result from one to the other.  Every sum keeps the
    // order it has when one wavefront runs a pblock's whole iteration (the
    // stencil in the same separable order, the scatter in the same
    // conflict-free passes), and the split measures bitwise equal to that
    // form: 42.2 us per iteration there, 16.5 + 19.5 us here on the slab.
    //   Mode 0: the first iteration's constraint (pbMpmParticleToGrid's
    //           solveMaterialConstraints on the staged C).
    //   Mode 1: G2P of iteration k-1 (pbMpmGridToParticle) and the
    //           constraint of iteration k.
    //   Mode 2: the last G2P and pbMpmIntegrateParticle, written back in
    //           particle order (src/solver/pb_mpm_algorithm.h:150-408).
    // ------------------------------------------------------------------
    template <int Mode>
    CR_DEV void pb_dense_particle(const StepCommon &c, const PbStage &st, int k)
    {
        const int S = st.stride;
        if (Mode == 0)
        {
            // the substep's staging (k_pb_stage's gather into sorted order),
            // then the first iteration's constraint on the staged C
            const int p = (int)c.order[k];
            st.xs[k] = c.P.xm[p];
            const float4 v = c.P.v[p];
            st.vs[k] = v.x;
            st.vs[S + k] = v.y;
            st.vs[2 * S + k] = v.z;
            const M3 F = load_pm(c.S.F, p);
            M3 Cm = load_pm(c.S.C, p);
            const float4 prm = c.P.params[p];
            const int type = c.P.type[p];
            store_m3(st.Fs, S, k, F);
            st.ps[k] = make_float2(prm.x, prm.y);
            st.ts[k] = type;
            pb_constraint(c.dt, prm.x, prm.y, type, F, Cm);
            store_m3(st.Cs, S, k, Cm);
            return;
        }
        const float4 xm = st.xs[k];
        V3 x = v3(xm);
        const V3 vp{st.vs[k], st.vs[S + k], st.vs[2 * S + k]};
        V3 vel;
        M3 Cm;
        gather_vc_grid(c, slicer_count(c.sv), x, vp, vel, Cm);
        // F and the material are loaded after the gather: held across it
        // (its slicer path is the kernel's register peak) they cost a
        // wavefront per SIMD (155 against 127 registers; 16.4 against 14.4 us
        // on the slab)
        const M3 F = load_m3(st.Fs, S, k);
        if (Mode == 1)
        {
            const float2 pr = st.ps[k];
            const int type = st.ts[k];
            pb_constraint(c.dt, pr.x, pr.y, type, F, Cm);
            st.vs[k] = vel.x;
            st.vs[S + k] = vel.y;
            st.vs[2 * S + k] = vel.z;
            store_m3(st.Cs, S, k, Cm);
        }
        else
        {
            // pbMpmIntegrateParticle (src/solver/pb_mpm_algorithm.h:363-408)
            const int p = (int)c.order[k];
            x = clamp_bounds(c.g, x + vel * c.dt);
            vel = vel + v3(c.gravity.x, c.gravity.y, c.gravity.z) * c.dt;
            float4 vv = make_float4(vel.x, vel.y, vel.z, 0.0f);
            resolve_rigid<true>(c.sv, shape_count(c.sv), vv, x);
            c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
            c.P.v[p] = make_float4(vel.x, vel.y, vel.z, 0.0f);
            store_pm(c.S.C, p, Cm);
            store_pm(c.S.F, p, mul(m3_identity() + Cm * c.dt, F));
        }
    }

    // One thread per sorted particle; n sort items, keys sorted (masked and
    // inactive particles, keyed past the last pblock, come last).  A
    // wavefront whose particles belong to several scenes of a batch takes
    // them one scene at a time, so scene records stay wavefront-uniform.
    template <int Mode>
    __global__ __launch_bounds__(kWG) void k_pb_dense(BatchArgs b, PbStage st, int n, const unsigned *keys)
    {
        const int k = blockIdx.x * kWG + threadIdx.x;
        const unsigned key = k < n ? keys[k] : 0xffffffffu;
        if (key >= (unsigned)b.totalBlocks)
            return;
        const int s = b.numScenes == 1 ? 0 : b.blockScene[key];
        for (;;)
        {
            const int s0 = __builtin_amdgcn_readfirstlane(s);
            if (s == s0)
            {
                pb_dense_particle<Mode>(scene_view_of(b, s0), st, k);
                return;
            }
        }
    }

    // The scatter half of an iteration: one wavefront per pblock tile (or
    // sub-tile), the staged x, v and C (the constraint's output) scattered
    // into 3 LDS accumulation tiles in conflict-free passes.  Mass: the first
    // iteration, which also accumulates the node masses (16-byte nodes);
    // later iterations accumulate momentum only (12-byte nodes) and leave the
    // first iteration's masses in the partial boxes.
    template <bool Mass, bool Persistent, int Waves = kWaves>
    __global__ __launch_bounds__(64 * Waves) void k_pb_scatter(BatchArgs b, PbStage st)
    {
        constexpr int kAcc = 3;
        using Acc = typename std::conditional<Mass, float4, F3>::type;
        __shared__ Acc acc[Waves * kAcc][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nIt = loop_items<Persistent>(b, num_subtiles(b), b.subMax);
        for (int it = first_tile_of<Waves>(wave); it < nIt; it = next_tile<Persistent, Waves>(it, nIt))
        {
            const int tg = loop_tile<Persistent>(b, it, b.subMax);
            if (!subtile_used(b, tg))
                continue;
            int t;
            const StepCommon c = scene_view(b, tg / b.subMax, t);
            TileInfo ti;
            if (!subtile_info(c, t, tg % b.subMax, ti))
                continue;
            if (!tile_class_ok(b.tileClass, ti.end - ti.start))
                continue;
            wave_sync(); // the previous tile's LDS reads are done
            Acc *A = acc[wave * kAcc];
            zero_tiles<kAcc>(A, lane);
            wave_sync();
            const int nSlicers = slicer_count(c.sv);
            const int S = st.stride;
            for (int k = ti.start + lane; k < ti.end; k += 64)
            {
                const float4 xm = st.xs[k];
                const V3 vel{st.vs[k], st.vs[S + k], st.vs[2 * S + k]};
                const M3 Cm = load_m3(st.Cs, S, k);
                pb_scatter<kAcc>(c, ti, A, nSlicers, v3(xm), xm.w, vel, Cm);
            }
            wave_sync();
            store_tile_sum<kAcc>(A, c.tilebuf + (size_t)ti.slot * kTileNodes, lane);
        }
    }

    // k_pb_scatter with W wavefronts per (sub-)tile (wg_tile_scatter), one
    // workgroup per tile slot.
    template <bool Mass, int W>
    __global__ __launch_bounds__(64 * W) void k_pb_scatter_wg(BatchArgs b, PbStage st)
    {
        constexpr int kAcc = 3;
        using Acc = typename std::conditional<Mass, float4, F3>::type;
        __shared__ Acc acc[W * kAcc][kTileNodes];
        const int tg = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
        if (tg >= num_subtiles(b) || !subtile_used(b, tg))
            return;
        int t;
        const StepCommon c = scene_view(b, tg / b.subMax, t);
        TileInfo ti;
        if (!subtile_info(c, t, tg % b.subMax, ti))
            return;
        if (!tile_class_ok(b.tileClass, ti.end - ti.start))
            return;
        const int nSlicers = slicer_count(c.sv);
        const int S = st.stride;
        wg_tile_scatter<W, kAcc>(acc[0], ti, c.tilebuf + (size_t)ti.slot * kTileNodes, [&](Acc *A, int k) {
            const float4 xm = st.xs[k];
            const V3 vel{st.vs[k], st.vs[S + k], st.vs[2 * S + k]};
            const M3 Cm = load_m3(st.Cs, S, k);
            pb_scatter<kAcc>(c, ti, A, nSlicers, v3(xm), xm.w, vel, Cm);
        });
    }

    // Zero iterations: integrate with the current v and C.
    __global__ void k_pb_integrate_only(BatchArgs b, int n, const unsigned *keys, PbStage st)
    {
        const int k = blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n || keys[k] >= (unsigned)b.totalBlocks)
            return;
        const StepCommon c = scene_view_of(b, b.blockScene[keys[k]]);
        const int S = st.stride;
        const int p = (int)c.order[k];
        const float4 xm = st.xs[k];
        V3 vel{st.vs[k], st.vs[S + k], st.vs[2 * S + k]};
        const M3 Cm = load_m3(st.Cs, S, k);
        V3 x = clamp_bounds(c.g, v3(xm) + vel * c.dt);
        vel = vel + v3(c.gravity.x, c.gravity.y, c.gravity.z) * c.dt;
        float4 vv = make_float4(vel.x, vel.y, vel.z, 0.0f);
        resolve_rigid<true>(c.sv, shape_count(c.sv), vv, x);
        c.P.xm[p] = make_float4(x.x, x.y, x.z, xm.w);
        c.P.v[p] = make_float4(vel.x, vel.y, vel.z, 0.0f);
        const M3 F = load_m3(st.Fs, S, k);
        store_pm(c.S.F, p, mul(m3_identity() + Cm * c.dt, F));
    }

    // ------------------------------------------------------------------
    // Standard MPM (src/solver/standard_mpm_algorithm.h:146-352)
    // ------------------------------------------------------------------
    template <bool Persistent, int Waves = kWaves>
    __global__ __launch_bounds__(64 * Waves) void k_std_p2g(BatchArgs b)
    {
        static_assert(Persistent ? Waves == kWaves : true, "persistent tile loops stride by kWaves per workgroup");
        __shared__ float4 acc[Waves][2][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nIt = loop_items<Persistent>(b, num_subtiles(b), b.subMax);
        for (int it = first_tile_of<Waves>(wave); it < nIt; it = next_tile<Persistent, Waves>(it, nIt))
        {
            const int tg = loop_tile<Persistent>(b, it, b.subMax);
            if (!subtile_used(b, tg))
                continue;
            int t;
            const StepCommon c = scene_view(b, tg / b.subMax, t);
            TileInfo ti;
            if (!subtile_info(c, t, tg % b.subMax, ti))
                continue;
            wave_sync(); // the previous tile's LDS reads are done
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
                const M3 stress = stress_kirchhoff(prm.x, prm.y, c.P.type[p], load_pm(c.S.F, p));
                const V3 vel = v3(c.P.v[p]);
                const float nv0 = -c.S.V0[p];
                const bool fast = nodes_inside(g, bx, by, bz);
                in_conflict_free_passes<1>(cell_key(ti, bx, by, bz), [&](int) {
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
            float4 *out = c.tilebuf + (size_t)ti.slot * kTileNodes;
            float4 *outF = c.tileForce + (size_t)ti.slot * kTileNodes;
            for (int i = lane; i < kTileNodes; i += 64)
            {
                out[i] = A[i];
                outF[i] = AF[i];
            }
        }
    }

    template <bool Persistent, int Waves = kWaves>
    __global__ __launch_bounds__(64 * Waves) void k_std_g2p(BatchArgs b)
    {
        static_assert(Persistent ? Waves == kWaves : true, "persistent tile loops stride by kWaves per workgroup");
        __shared__ float4 lg[Waves][kTileNodes];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nIt = loop_items<Persistent>(b, b.totalBlocks * b.gSubMax, b.gSubMax);
        for (int it = first_tile_of<Waves>(wave); it < nIt; it = next_tile<Persistent, Waves>(it, nIt))
        {
            const int tg = loop_tile<Persistent>(b, it, b.gSubMax);
            if (!split_used(b, tg, b.gSubSize, b.gSubMax))
                continue;
            int t;
            const StepCommon c = scene_view(b, tg / b.gSubMax, t);
            TileInfo ti;
            if (!split_info(c, t, tg % b.gSubMax, b.gSubSize, b.gSubMax, ti))
                continue;
            wave_sync(); // the previous tile's LDS reads are done
            float4 *G = lg[wave];
            load_tile_grid(c.g, ti, c.grid, G, lane);
            wave_sync();
            const GridGeom &g = c.g;
            const int nShapes = shape_count(c.sv);
            for (int k = ti.start + lane; k < ti.end; k += 64)
            {
                const int p = (int)c.gatherOrder[k];
                if (!c.P.mask[p])
                    continue;
                const float4 xm = c.P.xm[p];
                const M3 F = load_pm(c.S.F, p); // loaded with x, before the stores
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
                store_pm(c.S.F, p, mul(m3_identity() + gv * c.dt, F));
            }
        }
    }

    // ------------------------------------------------------------------
    // computeInitialData (src/solver/mls_mpm_algorithm.h:46-145): mass onto the
    // grid as it stands (global atomics, rarely called), then V0.
    // ------------------------------------------------------------------
    // computeInitialData's mass scatter (src/solver/mls_mpm_algorithm.h:46-101)
    // in 2^-40 kg fixed point: integer atomics add in any order to the same
    // sum, so the initial volumes are bitwise reproducible; the sum is added
    // to the grid's existing mass once per node by k_init_apply.
    constexpr double kMassFixed = 1099511627776.0; // 2^40

    // idx: scene-local particle indices of scene s; acc: one counter per node
    // of the scene (scene-local node index), zeroed by the caller.
    __global__ void k_init_mass(BatchArgs b, int s, unsigned n, const unsigned *idx, unsigned long long *acc)
    {
        const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const StepCommon c = scene_view_of(b, s);
        const GridGeom &g = c.g;
        const int p = b.scenes[s].particleOffset + (int)idx[i];
        store_pm(c.S.F, p, m3_identity());
        const float4 x = c.P.xm[p];
        int bx, by, bz;
        V3 fx;
        cell_of(g, v3(x), bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb)
                for (int cc = 0; cc < 3; ++cc)
                {
                    const int nx = bx + a, ny = by + bb, nz = bz + cc;
                    if (!in_grid(g, nx, ny, nz))
                        continue;
                    const float wt = W.w[a][0] * W.w[bb][1] * W.w[cc][2];
                    const long long q = llrint((double)(wt * x.w) * kMassFixed);
                    atomicAdd(&acc[grid_index(g, nx, ny, nz)], (unsigned long long)q);
                }
    }

    // Add the fixed-point mass sums to the grid (one wavefront per node block).
    __global__ __launch_bounds__(kWG) void k_init_apply(BatchArgs b, int s, const unsigned long long *acc,
                                                        unsigned char *dirty)
    {
        const StepCommon c = scene_view_of(b, s);
        const int blk = blockIdx.x * kWaves + (threadIdx.x >> 6);
        if (blk >= c.g.numBlocks)
            return;
        const int gi = blk * 64 + (threadIdx.x & 63);
        const long long q = (long long)acc[gi];
        const bool touched = q != 0;
        if (touched)
            c.grid[gi].w += (float)((double)q / kMassFixed);
        if (__ballot(touched) != 0ull && (threadIdx.x & 63) == 0)
            dirty[b.scenes[s].blockOffset + blk] = 1;
    }

    __global__ void k_init_volume(BatchArgs b, int s, unsigned n, const unsigned *idx)
    {
        const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n)
            return;
        const StepCommon c = scene_view_of(b, s);
        const GridGeom &g = c.g;
        const int p = b.scenes[s].particleOffset + (int)idx[i];
        store_pm(c.S.F, p, m3_identity());
        const float4 x = c.P.xm[p];
        int bx, by, bz;
        V3 fx;
        cell_of(g, v3(x), bx, by, bz, fx);
        Weights W;
        weights_of(fx, W);
        float density = 0.0f;
        for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb)
                for (int cc = 0; cc < 3; ++cc)
                {
                    const int nx = bx + a, ny = by + bb, nz = bz + cc;
                    if (!in_grid(g, nx, ny, nz))
                        continue;
                    const float wt = W.w[a][0] * W.w[bb][1] * W.w[cc][2];
                    const float nd = c.grid[grid_index(g, nx, ny, nz)].w / g.cellVol;
                    density += nd * wt;
                }
        c.S.V0[p] = (density > 1e-8f) ? x.w / density : 1e-8f;
    }

    __global__ void k_copy_grid(BatchArgs b, int s, float4 *dst)
    {
        const StepCommon c = scene_view_of(b, s);
        const GridGeom &g = c.g;
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        const int total = g.nx * g.ny * g.nz;
        if (i >= total)
            return;
        const int x = i % g.nx, y = (i / g.nx) % g.ny, z = i / (g.nx * g.ny);
        dst[i] = c.grid[grid_index(g, x, y, z)];
    }

    int grid1d(long long n, int bs) { return (int)((n + bs - 1) / bs); }
} // namespace


// ----------------------------------------------------------------------
// Adjoint of the MLS-MPM substep (reverse mode), for one scene without
// rigid shapes.  Forward per substep (src/solver/mls_mpm_algorithm.h):
//   P2G   v' = v + g dt, A = s tau(F) + m C (s = -4 dt V0 / dx^2),
//         p_i = sum_p w (m v' + A d),  m_i = sum_p w m          (d = x_i - x_p)
//   grid  u_i = p_i / m_i with the wall band zeroed, or u_i = p_i when m_i < 1e-4
//   G2P   v+ = sum w u_i, C+ = (4 / dx) sum w d u_i^T, x+ = clamp(x + dt v+),
//         F+ = (I + dt C+) F
// The kernels below apply the transposed Jacobians in the opposite order:
// k_adj_g2p (per particle, scatters the grid velocity adjoint through the
// LDS tile), k_adj_grid (per node), k_adj_p2g (per particle, gathers the
// momentum and mass adjoints, then the stress adjoint).  Parameter adjoints
// (lambda, mu) are accumulated per particle in FP64.
// ----------------------------------------------------------------------
namespace
{
    CR_DEV float mdot(const M3 &a, const M3 &b) { return dot(a.c0, b.c0) + dot(a.c1, b.c1) + dot(a.c2, b.c2); }

    // Particle state (x|m, v, F, C) in the solver's layouts.
    struct StateRef
    {
        const float4 *x;
        const float4 *v;
        const float *F; // packed, as SolverState
        const float *C;
        int stride;
    };

    // Adjoint of the particle state, indexed by particle.
    struct AdjRef
    {
        float4 *x;
        float4 *v;
        float4 *F; // columns, as SolverState
        float4 *C;
        float4 *xTmp; // G2P-adjoint part of the x adjoint
        float4 *FTmp; // G2P-adjoint part of the F adjoint
        double *lambdaBar;
        double *muBar;
        int stride;
    };

    struct NodeW
    {
        float w[3][3], dw[3][3]; // [i][axis]: weight and d weight / d fraction
    };
    CR_DEV void node_weights(V3 fx, NodeW &n)
    {
        const float f[3] = {fx.x, fx.y, fx.z};
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            bspline_w(f[a], n.w[0][a], n.w[1][a], n.w[2][a]);
            bspline_dw(f[a], n.dw[0][a], n.dw[1][a], n.dw[2][a]);
        }
    }

    // Visit the 27 nodes with weight, position gradient of the weight
    // (d w / d x_p) and d = x_i - x_p; out-of-grid nodes are skipped.
    template <class Fn>
    CR_DEV void for_each_node_grad(const GridGeom &g, const TileInfo &ti, int bx, int by, int bz, V3 fx, const NodeW &n,
                                   Fn &&fn)
    {
#pragma unroll 1
        for (int a = 0; a < 3; ++a)
#pragma unroll 1
            for (int b = 0; b < 3; ++b)
#pragma unroll 3
                for (int c = 0; c < 3; ++c)
                {
                    const int nx = bx + a, ny = by + b, nz = bz + c;
                    if (!in_grid(g, nx, ny, nz))
                        continue;
                    const int li = tile_local(ti, nx, ny, nz);
                    if (li < 0)
                        continue;
                    const float wa = w_of(fx.x, a), wb = w_of(fx.y, b), wc = w_of(fx.z, c);
                    const float da = dw_of(fx.x, a), db = dw_of(fx.y, b), dc = dw_of(fx.z, c);
                    const float wt = wa * wb * wc;
                    const V3 gw = V3{da * wb * wc, wa * db * wc, wa * wb * dc} * g.invdx;
                    const V3 dist = (V3{(float)a, (float)b, (float)c} - fx) * g.dx;
                    fn(li, dist, wt, gw);
                }
    }

    // The same visit for a particle whose stencil is read from a grid in
    // global memory: fn(node index in the scene's grid, d, w, grad w).
    template <class Fn>
    CR_DEV void for_each_node_grad_global(const GridGeom &g, int bx, int by, int bz, V3 fx, const NodeW &n, Fn &&fn)
    {
#pragma unroll 1
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll 3
                for (int c = 0; c < 3; ++c)
                {
                    const int nx = bx + a, ny = by + b, nz = bz + c;
                    if (!in_grid(g, nx, ny, nz))
                        continue;
                    const float wa = w_of(fx.x, a), wb = w_of(fx.y, b), wc = w_of(fx.z, c);
                    const float da = dw_of(fx.x, a), db = dw_of(fx.y, b), dc = dw_of(fx.z, c);
                    const float wt = wa * wb * wc;
                    const V3 gw = V3{da * wb * wc, wa * db * wc, wa * wb * dc} * g.invdx;
                    const V3 dist = (V3{(float)a, (float)b, (float)c} - fx) * g.dx;
                    fn(grid_index(g, nx, ny, nz), dist, wt, gw);
                }
    }

    // The weights' x adjoint of a G2P (MLS, or PB iteration k) for one particle:
    // v_k = sum_i w U_i, C_k = (4 / dx) sum_i w d U_i^T, so with B_bar =
    // (4 / dx) C_bar it is sum_i grad w (v_bar . U_i + d . B_bar U_i) -
    // w B_bar U_i.  U (the taped grid of the particle's scene) is read from
    // global memory: every in-grid node of a particle lies in its pblock's
    // box, and out-of-grid nodes are skipped, as in the forward's gather.
    CR_DEV V3 pbadj_g2p_x(const GridGeom &g, const float4 *U, V3 x, V3 vb, const M3 &Bb)
    {
        int bx, by, bz;
        V3 fx;
        cell_of(g, x, bx, by, bz, fx);
        NodeW nw;
        node_weights(fx, nw);
        V3 xacc{0.f, 0.f, 0.f};
#pragma unroll 1
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll 3
                for (int c = 0; c < 3; ++c)
                {
                    const int nx = bx + a, ny = by + b, nz = bz + c;
                    if (!in_grid(g, nx, ny, nz))
                        continue;
                    const float wa = w_of(fx.x, a), wb = w_of(fx.y, b), wc = w_of(fx.z, c);
                    const float da = dw_of(fx.x, a), db = dw_of(fx.y, b), dc = dw_of(fx.z, c);
                    const float wt = wa * wb * wc;
                    const V3 gw = V3{da * wb * wc, wa * db * wc, wa * wb * dc} * g.invdx;
                    const V3 dist = (V3{(float)a, (float)b, (float)c} - fx) * g.dx;
                    const V3 u = v3(U[grid_index(g, nx, ny, nz)]);
                    const V3 Bu = mul(Bb, u);
                    xacc = xacc + gw * (dot(vb, u) + dot(dist, Bu)) - Bu * wt;
                }
        return xacc;
    }

    // What the MLS G2P adjoint's scatter needs per particle (tile order):
    // the adjoint of the gathered velocity and of the gathered affine matrix.
    struct MlsScatterIn
    {
        float4 *v; // v_bar (xyz)
        float4 *B; // B_bar = (4 / dx) (C_bar + dt F_bar F^T), 3 columns
    };

    // G2P adjoint of one substep, per particle (one thread per sorted
    // particle).  s0: state at the substep's start, s1: at its end; grid: the
    // substep's updated grid (read from global memory: every in-grid node of
    // a particle's stencil lies in a block that received partials).  The
    // adjoint of s1 comes in through ad; the G2P parts of the x and F
    // adjoints go to ad.xTmp / ad.FTmp, and the inputs of the velocity
    // adjoint's scatter to si.
    __global__ __launch_bounds__(kWG) void k_adj_g2p(BatchArgs b, int n, const unsigned *keys, StateRef s0, StateRef s1,
                                                     const float4 *grid, AdjRef ad, MlsScatterIn si)
    {
        const int q = blockIdx.x * kWG + threadIdx.x;
        const unsigned key = q < n ? keys[q] : 0xffffffffu;
        if (key >= (unsigned)b.totalBlocks)
            return;
        const StepCommon c = scene_view_of(b, 0);
        const GridGeom &g = c.g;
        const int p = (int)c.order[q];
        const V3 x = v3(s0.x[p]);
        const V3 vn = v3(s1.v[p]);
        const M3 Cn = load_pm(s1.C, p);
        const M3 F = load_pm(s0.F, p);
        const V3 xb = v3(ad.x[p]);
        const V3 vb = v3(ad.v[p]);
        const M3 Fb = load_cols(ad.F, p);
        const M3 Cb = load_cols(ad.C, p);
        // x+ = clamp(x + dt v+): the adjoint passes where the clamp is inactive
        const V3 xp = x + vn * c.dt;
        const V3 keep{(xp.x > g.bmin.x && xp.x < g.bmax.x) ? 1.f : 0.f, (xp.y > g.bmin.y && xp.y < g.bmax.y) ? 1.f : 0.f,
                      (xp.z > g.bmin.z && xp.z < g.bmax.z) ? 1.f : 0.f};
        const V3 xbk = xb * keep;
        const V3 vaccb = vb + xbk * c.dt;
        // F+ = (I + dt C+) F
        const M3 Bb = (Cb + mul(Fb, transpose(F)) * c.dt) * (4.0f * g.invdx);
        store_cols(ad.FTmp, p, mul(transpose(m3_identity() + Cn * c.dt), Fb));
        // v+ = sum w u_i, C+ = (4 / dx) sum w d u_i^T: the weights' x adjoint
        const V3 xacc = xbk + pbadj_g2p_x(g, grid + (c.grid - b.grid), x, vaccb, Bb);
        ad.xTmp[p] = make_float4(xacc.x, xacc.y, xacc.z, 0.0f);
        si.v[q] = make_float4(vaccb.x, vaccb.y, vaccb.z, 0.0f);
        store_cols(si.B, q, Bb);
    }

    // The grid velocity adjoint of the G2P, w (v_bar + B_bar^T d) per node,
    // through the forward's separable scatter (pb_scatter with m = 1) into the
    // tile partials; one pblock per one-wavefront workgroup.
    // W wavefronts per tile as in the forward's scatter (wg_tile_scatter).
    template <int W>
    __global__ __launch_bounds__(64 * W) void k_adj_uscatter(BatchArgs b, StateRef s0, MlsScatterIn si)
    {
        constexpr int kAcc = 3;
        __shared__ F3 acc[W * kAcc][kTileNodes];
        const int tg = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
        if (tg >= num_subtiles(b) || !subtile_used(b, tg))
            return;
        int t;
        const StepCommon c = scene_view(b, tg / b.subMax, t);
        TileInfo ti;
        if (!subtile_info(c, t, tg % b.subMax, ti))
            return;
        if (!tile_class_ok(b.tileClass, ti.end - ti.start))
            return;
        wg_tile_scatter<W, kAcc>(acc[0], ti, c.tilebuf + (size_t)ti.slot * kTileNodes, [&](F3 *A, int q) {
            pb_scatter<kAcc>(c, ti, A, 0, v3(s0.x[c.order[q]]), 1.0f, v3(si.v[q]), transpose(load_cols(si.B, q)));
        });
    }

    // Grid adjoint: sum the tile partials of the velocity adjoint (fixed
    // order), then through u = p / m (or u = p below the mass cutoff: 1e-4
    // for MLS, Kind 0; 1e-8 for PB, Kind 1) to the momentum and mass
    // adjoints (p_bar xyz, m_bar w).  One wavefront per node block of the
    // substep's active list (the blocks that received tile partials, with
    // the mask of their source pblocks holding particles).
    template <int Kind>
    __global__ __launch_bounds__(kWG) void k_adj_grid(BatchArgs b, const float4 *grid, float4 *gridAdj)
    {
        const float cutoff = Kind == 1 ? 1e-8f : 1e-4f;
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nActive = __builtin_amdgcn_readfirstlane(*b.activeCount);
        for (int item = first_tile(wave); item < nActive; item += tile_stride())
        {
            const unsigned entry = __builtin_amdgcn_readfirstlane(b.activeList[item]);
            const int nbg = (int)(entry >> 8);
            int nb;
            const StepCommon c = scene_view(b, nbg, nb);
            const GridGeom &g = c.g;
            int bx, by, bz;
            decode_block(g, nb, bx, by, bz);
            int nsub[8];
#pragma unroll
            for (int d = 0; d < 8; ++d)
            {
                const int key = ((entry >> d) & 1u) ? block_key(g, bx - (d & 1), by - ((d >> 1) & 1), bz - (d >> 2)) : -1;
                nsub[d] = key >= 0 ? subtiles_of(c.segEnd[key] - c.segStart[key], c.subSize, c.subMax) : 0;
            }
            const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
            V3 ub{0.f, 0.f, 0.f};
#pragma unroll
            for (int d = 0; d < 8; ++d)
            {
                const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
                const int ux = lx + 4 * dx, uy = ly + 4 * dy, uz = lz + 4 * dz;
                if (!(ux < kTN && uy < kTN && uz < kTN))
                    continue;
                const size_t box = (size_t)(nb - dx - g.nbx * (dy + g.nby * dz)) * c.subStride;
                const int li = ux + kTN * (uy + kTN * uz);
                for (int j = 0; j < nsub[d]; ++j)
                    ub = ub + v3(c.tilebuf[(box + j) * kTileNodes + li]);
            }
            const int i = bx * 4 + lx, j = by * 4 + ly, k = bz * 4 + lz;
            if (!in_grid(g, i, j, k))
                continue;
            const size_t gi = (size_t)(c.grid - b.grid) + (size_t)nb * 64 + lane;
            const float4 u = grid[gi];
            float4 r;
            if (u.w < cutoff)
                r = make_float4(ub.x, ub.y, ub.z, 0.0f);
            else
            {
                const bool kx = !(i < 2 || i > g.nx - 3), ky = !(j < 2 || j > g.ny - 3), kz = !(k < 2 || k > g.nz - 3);
                r.x = kx ? ub.x / u.w : 0.0f;
                r.y = ky ? ub.y / u.w : 0.0f;
                r.z = kz ? ub.z / u.w : 0.0f;
                r.w = -(ub.x * u.x + ub.y * u.y + ub.z * u.z) / u.w;
            }
            gridAdj[gi] = r;
        }
    }

    // FP64 3x3 matrices for the affine and stress adjoints (row-major d[r][c]).
    struct D3
    {
        double d[3][3];
    };
    CR_DEV D3 d3_zero()
    {
        D3 a;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                a.d[r][c] = 0.0;
        return a;
    }
    CR_DEV D3 d3_of(const M3 &m)
    {
        const V3 col[3] = {m.c0, m.c1, m.c2};
        D3 a;
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
            a.d[0][c] = col[c].x;
            a.d[1][c] = col[c].y;
            a.d[2][c] = col[c].z;
        }
        return a;
    }
    CR_DEV M3 m3_of(const D3 &a)
    {
        return M3{V3{(float)a.d[0][0], (float)a.d[1][0], (float)a.d[2][0]},
                  V3{(float)a.d[0][1], (float)a.d[1][1], (float)a.d[2][1]},
                  V3{(float)a.d[0][2], (float)a.d[1][2], (float)a.d[2][2]}};
    }
    CR_DEV D3 d3_mul(const D3 &a, const D3 &b)
    {
        D3 o;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                o.d[r][c] = a.d[r][0] * b.d[0][c] + a.d[r][1] * b.d[1][c] + a.d[r][2] * b.d[2][c];
        return o;
    }
    CR_DEV D3 d3_t(const D3 &a)
    {
        D3 o;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                o.d[r][c] = a.d[c][r];
        return o;
    }
    CR_DEV double d3_tr(const D3 &a) { return a.d[0][0] + a.d[1][1] + a.d[2][2]; }
    CR_DEV double d3_dot(const D3 &a, const D3 &b)
    {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                s += a.d[r][c] * b.d[r][c];
        return s;
    }
    // a x + b y + z I
    CR_DEV D3 d3_comb(const D3 &a, double x, const D3 &b, double y, double z)
    {
        D3 o;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                o.d[r][c] = a.d[r][c] * x + b.d[r][c] * y + (r == c ? z : 0.0);
        return o;
    }

    // P2G adjoint of one substep, one thread per sorted particle: gathers
    // (p_bar, m_bar) from the grid adjoint in global memory, then the affine,
    // stress and parameter adjoints; writes the adjoint of the substep's
    // start state.
    __global__ __launch_bounds__(kWG) void k_adj_p2g(BatchArgs b, int n, const unsigned *keys, StateRef s0,
                                                     const float4 *gridAdj, AdjRef ad)
    {
        const int q = blockIdx.x * kWG + threadIdx.x;
        const unsigned key = q < n ? keys[q] : 0xffffffffu;
        if (key >= (unsigned)b.totalBlocks)
            return;
        {
            const StepCommon c = scene_view_of(b, 0);
            const GridGeom &g = c.g;
            const float4 *G = gridAdj + (c.grid - b.grid);
            {
                const int p = (int)c.order[q];
                const float4 xm = s0.x[p];
                const V3 x = v3(xm);
                const float m = xm.w;
                const V3 vel = v3(s0.v[p]) + v3(c.gravity.x, c.gravity.y, c.gravity.z) * c.dt;
                const M3 F = load_pm(s0.F, p);
                const M3 Cm = load_pm(s0.C, p);
                const float4 prm = c.P.params[p];
                const int type = c.P.type[p];
                const M3 tau = stress_kirchhoff(prm.x, prm.y, type, F);
                const float sc = -c.dt * 4.0f * g.invdx * g.invdx * c.S.V0[p];
                const M3 A = tau * sc + Cm * m;
                const M3 AT = transpose(A);
                const V3 vm = vel * m;
                V3 vpb{0.f, 0.f, 0.f}, xacc{0.f, 0.f, 0.f};
                D3 Ab = d3_zero(); // affine adjoint in FP64: its trace is a first moment with cancellation
                int bx, by, bz;
                V3 fx;
                cell_of(g, x, bx, by, bz, fx);
                NodeW nw;
                node_weights(fx, nw);
                for_each_node_grad_global(g, bx, by, bz, fx, nw, [&](int gi, V3 dist, float wt, V3 gw) {
                    const float4 q = G[gi];
                    const V3 pb = v3(q);
                    vpb = vpb + pb * (wt * m);
                    const double pr[3] = {pb.x, pb.y, pb.z}, dc[3] = {dist.x, dist.y, dist.z};
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int cc = 0; cc < 3; ++cc)
                            Ab.d[r][cc] += (double)wt * pr[r] * dc[cc];
                    const float wbar = dot(pb, vm + mul(A, dist)) + q.w * m;
                    xacc = xacc + gw * wbar - mul(AT, pb) * wt;
                });
                // A = s tau + m C; the stress and parameter adjoints in FP64
                const D3 taub = d3_comb(Ab, (double)sc, Ab, 0.0, 0.0);
                const D3 Fd = d3_of(F), FdT = d3_t(Fd);
                const double lam = prm.x, muv = prm.y;
                D3 Fb = d3_of(load_cols(ad.FTmp, p));
                double lb = 0.0, mb = 0.0;
                if (type == mNeoHookean)
                {
                    // tau = mu (F F^T - I) + lambda ln(J) I, J = max(det F, 1e-6)
                    const float Jr = det(F);
                    const double J = Jr < 1e-6f ? 1e-6 : (double)Jr;
                    Fb = d3_comb(Fb, 1.0, d3_mul(d3_comb(taub, 1.0, d3_t(taub), 1.0, 0.0), Fd), muv, 0.0);
                    if (Jr >= 1e-6f) // d ln(J) / dF = F^-T (inverse_t)
                        Fb = d3_comb(Fb, 1.0, d3_of(inverse_t(F)), lam * d3_tr(taub), 0.0);
                    mb = d3_dot(taub, d3_comb(d3_mul(Fd, FdT), 1.0, Fd, 0.0, -1.0));
                    lb = d3_tr(taub) * log(J);
                }
                else if (type == mCoRotational || type == mLinearElastic)
                {
                    // tau = F S F^T, S = lambda tr(E) I + 2 mu E, E = (F^T F - I) / 2
                    const D3 E = d3_comb(d3_mul(FdT, Fd), 0.5, Fd, 0.0, -0.5);
                    const D3 S = d3_comb(E, 2.0 * muv, E, 0.0, lam * d3_tr(E));
                    const D3 Sb = d3_mul(d3_mul(FdT, taub), Fd);
                    const D3 Eb = d3_comb(Sb, 2.0 * muv, Sb, 0.0, lam * d3_tr(Sb));
                    const D3 t1 = d3_mul(d3_mul(taub, Fd), d3_t(S));
                    const D3 t2 = d3_mul(d3_mul(d3_t(taub), Fd), S);
                    const D3 t3 = d3_mul(Fd, d3_comb(Eb, 0.5, d3_t(Eb), 0.5, 0.0));
                    Fb = d3_comb(d3_comb(Fb, 1.0, t1, 1.0, 0.0), 1.0, d3_comb(t2, 1.0, t3, 1.0, 0.0), 1.0, 0.0);
                    lb = d3_tr(E) * d3_tr(Sb);
                    mb = 2.0 * d3_dot(Sb, E);
                }
                const V3 xt = v3(ad.xTmp[p]) + xacc;
                ad.x[p] = make_float4(xt.x, xt.y, xt.z, 0.0f);
                ad.v[p] = make_float4(vpb.x, vpb.y, vpb.z, 0.0f);
                store_cols(ad.F, p, m3_of(Fb));
                store_cols(ad.C, p, m3_of(d3_comb(Ab, (double)m, Ab, 0.0, 0.0)));
                ad.lambdaBar[p] += lb;
                ad.muBar[p] += mb;
            }
        }
    }

    // ------------------------------------------------------------------
    // Adjoint of the PB-MPM substep (reverse mode), one scene without rigid
    // shapes.  Forward (src/solver/pb_mpm_algorithm.h:150-408; positions,
    // and so weights and node masses M, are fixed over the K iterations):
    //   iteration k = 1..K   Cc = constraint(C_{k-1}, F)                (per particle)
    //                        P_i = sum_p w m (v_{k-1} + Cc d),  U_i = P_i / M_i with the
    //                        wall band zeroed, or P_i where M_i < 1e-8
    //                        v_k = sum_i w U_i,  C_k = (4 / dx) sum_i w d U_i^T
    //   integration          x+ = clamp(x + dt v_K), v+ = v_K + g dt,
    //                        F+ = (I + dt C_K) F, C+ = C_K
    // The constraint (constitutive_models.h:123-154): T = (I + dt C) F,
    // R = u vt of svd3(T) (the polar rotation of T^T), Q = T / s with
    // s = sign(det T) sqrt(clamp(|det T|, 0.1, 1000)), B = r R + (1 - r) Q,
    // Cc = C + alpha ((B F^-T - I) - dt C).
    // The driver re-runs each substep with every iteration's grid U_k kept
    // (the tape), so v_{k-1} and C_{k-1} are re-gathered exactly as the
    // forward formed them.  The reverse kernels work in the substep's tile
    // order on its staged particle arrays (x|m, F, (alpha, r), type):
    //   k_pbadj_integrate    the integration's adjoint, per particle
    //   k_pbadj_g2p_x        the weights' x adjoint of iteration K's G2P
    //   k = K..1             k_pbadj_uscatter (per pblock: the velocity
    //                        adjoint scattered into tile partials),
    //                        k_adj_grid<1> (momentum and mass adjoints),
    //                        k_pbadj_particle (per particle: the scatter's
    //                        inputs re-formed, the momentum and mass adjoints
    //                        gathered, the constraint's adjoint through the
    //                        SVD's polar rotation, iteration k-1's x adjoint
    //                        of the weights)
    //   k_pbadj_finalize     the adjoint of F^-T; back to particle order
    // As in the forward, the per-particle work (gathers, SVD) runs 64 sorted
    // particles per wavefront, reading the grids from global memory, and
    // only the scatter is per pblock, through LDS.
    // ------------------------------------------------------------------

    // Adjoint state of the substep being reversed, in tile (sorted) order.
    struct PbAdjRef
    {
        float4 *x;  // x adjoint (xyz), accumulated over the substep
        float4 *v;  // adjoint of v_k
        float4 *C;  // adjoint of C_k, 3 columns
        float4 *F;  // adjoint of F, 3 columns, accumulated
        float4 *Fi; // adjoint of F^-T, 3 columns, accumulated
    };

    // Cofactor matrix of m (d det(m) / dm): the adjugate's transpose.
    CR_DEV M3 cofactor(const M3 &m)
    {
        const float a = m.c0.x, b = m.c1.x, c = m.c2.x, d = m.c0.y, e = m.c1.y, f = m.c2.y, g = m.c0.z, h = m.c1.z, i = m.c2.z;
        return M3{V3{(e * i - f * h), -(b * i - c * h), (b * f - c * e)},
                  V3{-(d * i - f * g), (a * i - c * g), -(a * f - c * d)},
                  V3{(d * h - e * g), -(a * h - b * g), (a * e - b * d)}};
    }
    CR_DEV float m3_at(const M3 &m, int r, int c)
    {
        const V3 col = c == 0 ? m.c0 : (c == 1 ? m.c1 : m.c2);
        return r == 0 ? col.x : (r == 1 ? col.y : col.z);
    }

    // Adjoint of pb_constraint for an active material: Ccb is the adjoint of
    // the constraint's output; returns the adjoint of its input C and adds
    // to those of F, of F^-T (Fi) and of alpha and the ratio.  The forward's
    // intermediates are recomputed in FP32 exactly as pb_constraint forms
    // them (the same SVD).  FP64 where it matters: the parameter reductions
    // (the residual (B F^-T - I) - dt C is a small difference of O(1)
    // terms) and the polar rotation's derivative, whose terms divide by
    // sigma_i + sigma_j.
    CR_DEV M3 pb_constraint_adjoint(float dt, float alpha, float ratio, const M3 &F, const M3 &C, const M3 &Ccb,
                                    M3 &Fb, M3 &Fib, double &ab, double &rb)
    {
        const M3 T = mul(m3_identity() + C * dt, F);
        M3 u, vt;
        V3 sig;
        svd3(T, u, vt, sig);
        const float detT = det(T);
        const float adet = fabsf(detT);
        const float cl = fminf(fmaxf(adet, 0.1f), 1000.0f);
        const float s = copysignf(1.0f, detT) * sqrtf(cl);
        const M3 Q = T * (1.0f / s);
        const M3 R = mul(u, vt);
        const M3 B = R * ratio + Q * (1.0f - ratio);
        const M3 Fi = inverse_t(F);
        // Cc = C + alpha ((B Fi - I) - dt C)
        ab += d3_dot(d3_of(Ccb), d3_comb(d3_mul(d3_of(B), d3_of(Fi)), 1.0, d3_of(C), -(double)dt, -1.0));
        const M3 Bb = mul(Ccb, transpose(Fi)) * alpha;
        Fib = Fib + mul(transpose(B), Ccb) * alpha;
        // B = r R + (1 - r) Q
        rb += d3_dot(d3_of(Bb), d3_comb(d3_of(R), 1.0, d3_of(Q), -1.0, 0.0));
        const M3 Rb = Bb * ratio;
        // Q = T / s, s = sign(d) sqrt(clamp(|d|, 0.1, 1000)), d = det T
        M3 Tb = Bb * ((1.0f - ratio) / s);
        if (adet > 0.1f && adet < 1000.0f)
        {
            const float sb = -(1.0f - ratio) * mdot(Bb, T) / (s * s);
            Tb = Tb + cofactor(T) * (sb * 0.5f / sqrtf(adet));
        }
        // R = U V^T for T^T = U diag(sigma) V^T (U = u, V^T = vt):
        // dL/d(T^T) = U K V^T, K_ij = (G_ij - G_ji) / (sigma_i + sigma_j), G = U^T Rb V
        const M3 G = mul(mul(transpose(u), Rb), transpose(vt));
        const double sg[3] = {sig.x, sig.y, sig.z};
        float k[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)
            {
                const double den = sg[i] + sg[j];
                k[i][j] = (i != j && fabs(den) > 1e-12)
                              ? (float)(((double)m3_at(G, i, j) - (double)m3_at(G, j, i)) / den)
                              : 0.0f;
            }
        const M3 K{V3{k[0][0], k[1][0], k[2][0]}, V3{k[0][1], k[1][1], k[2][1]}, V3{k[0][2], k[1][2], k[2][2]}};
        Tb = Tb + transpose(mul(mul(u, K), vt));
        // T = (I + dt C) F
        Fb = Fb + mul(transpose(m3_identity() + C * dt), Tb);
        return Ccb * (1.0f - alpha * dt) + mul(Tb, transpose(F)) * dt;
    }

    // Adjoint of the integration (pbMpmIntegrateParticle without shapes),
    // one thread per sorted particle: the adjoint of the substep's end state
    // (particle order, ad) to that of (v_K, C_K) and the direct x and F
    // terms (tile order, pa).  s1 is the substep's end state; F is the
    // staged start F.
    __global__ void k_pbadj_integrate(BatchArgs b, PbStage st, int n, const unsigned *keys, StateRef s1, AdjRef ad,
                                      PbAdjRef pa)
    {
        const int q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n || keys[q] >= (unsigned)b.totalBlocks)
            return;
        const StepCommon c = scene_view_of(b, b.numScenes == 1 ? 0 : b.blockScene[keys[q]]);
        const GridGeom &g = c.g;
        const int p = (int)c.order[q];
        const V3 xe = v3(s1.x[p]);
        const M3 CK = load_pm(s1.C, p);
        const M3 F = load_m3(st.Fs, st.stride, q);
        const V3 xb = v3(ad.x[p]), vb = v3(ad.v[p]);
        const M3 Fb = load_cols(ad.F, p), Cb = load_cols(ad.C, p);
        // x+ = clamp(x + dt v_K): the adjoint passes where the clamp is inactive
        const V3 keep{(xe.x > g.bmin.x && xe.x < g.bmax.x) ? 1.f : 0.f, (xe.y > g.bmin.y && xe.y < g.bmax.y) ? 1.f : 0.f,
                      (xe.z > g.bmin.z && xe.z < g.bmax.z) ? 1.f : 0.f};
        const V3 xbk = xb * keep;
        const V3 vK = vb + xbk * c.dt;
        pa.x[q] = make_float4(xbk.x, xbk.y, xbk.z, 0.0f);
        pa.v[q] = make_float4(vK.x, vK.y, vK.z, 0.0f);
        // F+ = (I + dt C_K) F, C+ = C_K
        store_cols(pa.C, q, Cb + mul(Fb, transpose(F)) * c.dt);
        store_cols(pa.F, q, mul(transpose(m3_identity() + CK * c.dt), Fb));
        const M3 z{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}};
        store_cols(pa.Fi, q, z);
    }

    // The weights' x adjoint of the last iteration's G2P (the earlier
    // iterations' is added by k_pbadj_particle), one thread per sorted
    // particle; U is iteration K's taped grid (the batch's layout).
    __global__ __launch_bounds__(kWG) void k_pbadj_g2p_x(BatchArgs b, PbStage st, int n, const unsigned *keys,
                                                         const float4 *U, PbAdjRef pa)
    {
        const int q = blockIdx.x * kWG + threadIdx.x;
        const unsigned key = q < n ? keys[q] : 0xffffffffu;
        if (key >= (unsigned)b.totalBlocks)
            return;
        const StepCommon c = scene_view_of(b, 0);
        const M3 Bb = load_cols(pa.C, q) * (4.0f * c.g.invdx);
        const V3 xs = v3(pa.x[q]) + pbadj_g2p_x(c.g, U + (c.grid - b.grid), v3(st.xs[q]), v3(pa.v[q]), Bb);
        pa.x[q] = make_float4(xs.x, xs.y, xs.z, 0.0f);
    }

    // The grid velocity adjoint of iteration k's G2P, w (v_bar + B_bar^T d)
    // per node: the form of the forward scatter's momentum with m = 1, so it
    // goes through pb_scatter (separable node sums, three accumulation tiles
    // in conflict-free passes) into the tile partials; one pblock per
    // one-wavefront workgroup, so the grid adjoint sums at most 8 partials
    // per node.
    template <int W>
    __global__ __launch_bounds__(64 * W) void k_pbadj_uscatter(BatchArgs b, PbStage st, PbAdjRef pa)
    {
        constexpr int kAcc = 3;
        __shared__ F3 acc[W * kAcc][kTileNodes];
        const int tg = __builtin_amdgcn_readfirstlane((int)blockIdx.x);
        if (tg >= num_subtiles(b) || !subtile_used(b, tg))
            return;
        int t;
        const StepCommon c = scene_view(b, tg / b.subMax, t);
        TileInfo ti;
        if (!subtile_info(c, t, tg % b.subMax, ti))
            return;
        if (!tile_class_ok(b.tileClass, ti.end - ti.start))
            return;
        wg_tile_scatter<W, kAcc>(acc[0], ti, c.tilebuf + (size_t)ti.slot * kTileNodes, [&](F3 *A, int q) {
            const M3 Bb = load_cols(pa.C, q) * (4.0f * c.g.invdx);
            pb_scatter<kAcc>(c, ti, A, 0, v3(st.xs[q]), 1.0f, v3(pa.v[q]), transpose(Bb));
        });
    }

    // The per-particle part of reverse iteration k, one thread per sorted
    // particle (the phases run in one kernel: at the sizes of a
    // differentiable scene a launch holds about one wavefront per SIMD, so
    // each launch costs its latency):
    //  1. the forward's inputs of iteration k's scatter re-formed
    //     (k_pb_dense's arithmetic): (v_{k-1}, C_{k-1}) gathered from the
    //     previous iteration's taped grid Uprev -- or, for k = 1, read from
    //     the substep's start state s0 -- and the constraint applied;
    //  2. P2G adjoint: the momentum and mass adjoints (gridAdj) gathered into
    //     the adjoint of v_{k-1} and of the constraint's output (the latter
    //     accumulated in FP64, a first moment with cancellation), and the
    //     weights' x adjoint;
    //  3. the constraint's adjoint (active materials; others pass the
    //     adjoint through): the adjoint of C_{k-1}, the F and F^-T adjoints
    //     accumulated in pa, those of alpha and the ratio in FP64 per particle
    //     (ad.lambdaBar, ad.muBar);
    //  4. for k > 1, the weights' x adjoint of iteration k-1's G2P.
    // Grids are read from global memory (every in-grid node of a particle's
    // stencil lies in a block that received partials).
    __global__ __launch_bounds__(kWG) void k_pbadj_particle(BatchArgs b, PbStage st, int n, const unsigned *keys,
                                                            StateRef s0, const float4 *gridAdj, const float4 *Uprev,
                                                            PbAdjRef pa, AdjRef ad)
    {
        const int q = blockIdx.x * kWG + threadIdx.x;
        const unsigned key = q < n ? keys[q] : 0xffffffffu;
        if (key >= (unsigned)b.totalBlocks)
            return;
        StepCommon c = scene_view_of(b, 0);
        const GridGeom &g = c.g;
        const size_t off = (size_t)(c.grid - b.grid); // the scene's first node in the batch's layout
        const float4 xm = st.xs[q];
        const V3 x = v3(xm);
        const float m = xm.w;
        // 1. the scatter's inputs
        V3 vel;
        M3 Cm;
        if (!Uprev)
        {
            const int p = (int)c.order[q];
            vel = v3(s0.v[p]);
            Cm = load_pm(s0.C, p);
        }
        else
        {
            c.grid = const_cast<float4 *>(Uprev) + off;
            gather_vc_grid(c, 0, x, V3{0.f, 0.f, 0.f}, vel, Cm);
        }
        const M3 F = load_m3(st.Fs, st.stride, q);
        const float2 pr = st.ps[q];
        const int type = st.ts[q];
        M3 Cc = Cm;
        pb_constraint(c.dt, pr.x, pr.y, type, F, Cc);
        // 2. P2G adjoint
        const float4 *G = gridAdj + off;
        const M3 CcT = transpose(Cc);
        V3 vpb{0.f, 0.f, 0.f}, xacc{0.f, 0.f, 0.f};
        D3 Ccb = d3_zero();
        {
            int bx, by, bz;
            V3 fx;
            cell_of(g, x, bx, by, bz, fx);
            NodeW nw;
            node_weights(fx, nw);
#pragma unroll 1
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int bb = 0; bb < 3; ++bb)
#pragma unroll 3
                    for (int cc = 0; cc < 3; ++cc)
                    {
                        const int nx = bx + a, ny = by + bb, nz = bz + cc;
                        if (!in_grid(g, nx, ny, nz))
                            continue;
                        const float wa = w_of(fx.x, a), wb = w_of(fx.y, bb), wc = w_of(fx.z, cc);
                        const float da = dw_of(fx.x, a), db = dw_of(fx.y, bb), dc = dw_of(fx.z, cc);
                        const float wt = wa * wb * wc;
                        const V3 gw = V3{da * wb * wc, wa * db * wc, wa * wb * dc} * g.invdx;
                        const V3 dist = (V3{(float)a, (float)bb, (float)cc} - fx) * g.dx;
                        const float4 gq = G[grid_index(g, nx, ny, nz)];
                        const V3 pb = v3(gq);
                        const float wm = wt * m;
                        vpb = vpb + pb * wm;
                        const double pr3[3] = {pb.x, pb.y, pb.z}, dcs[3] = {dist.x, dist.y, dist.z};
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int k2 = 0; k2 < 3; ++k2)
                                Ccb.d[r][k2] += (double)wm * pr3[r] * dcs[k2];
                        const float wbar = m * (dot(pb, vel + mul(Cc, dist)) + gq.w);
                        xacc = xacc + gw * wbar - mul(CcT, pb) * wm;
                    }
        }
        // 3. the constraint's adjoint
        M3 Cb = m3_of(Ccb);
        if (type == mNeoHookean || type == mCoRotational)
        {
            M3 Fb = load_cols(pa.F, q), Fib = load_cols(pa.Fi, q);
            double ab = 0.0, rb = 0.0;
            Cb = pb_constraint_adjoint(c.dt, pr.x, pr.y, F, Cm, Cb, Fb, Fib, ab, rb);
            store_cols(pa.F, q, Fb);
            store_cols(pa.Fi, q, Fib);
            const int p = (int)c.order[q];
            ad.lambdaBar[p] += ab;
            ad.muBar[p] += rb;
        }
        // 4. iteration k-1's weights' x adjoint
        if (Uprev)
            xacc = xacc + pbadj_g2p_x(g, Uprev + off, x, vpb, Cb * (4.0f * g.invdx));
        const V3 xs = v3(pa.x[q]) + xacc;
        pa.x[q] = make_float4(xs.x, xs.y, xs.z, 0.0f);
        pa.v[q] = make_float4(vpb.x, vpb.y, vpb.z, 0.0f);
        store_cols(pa.C, q, Cb);
    }

    // End of a substep's reverse: the adjoint of F^-T added to F's
    // (d F^-T = -F^-T dF^T F^-T), and the substep's start-state adjoint
    // written back in particle order.
    __global__ void k_pbadj_finalize(BatchArgs b, PbStage st, int n, const unsigned *keys, PbAdjRef pa, AdjRef ad)
    {
        const int q = blockIdx.x * blockDim.x + threadIdx.x;
        if (q >= n || keys[q] >= (unsigned)b.totalBlocks)
            return;
        const int p = (int)b.order[q];
        const M3 Fi = inverse_t(load_m3(st.Fs, st.stride, q));
        const M3 Fb = load_cols(pa.F, q) - mul(mul(Fi, transpose(load_cols(pa.Fi, q))), Fi);
        ad.x[p] = pa.x[q];
        ad.v[p] = pa.v[q];
        store_cols(ad.F, p, Fb);
        store_cols(ad.C, p, load_cols(pa.C, q));
    }
} // namespace

// ----------------------------------------------------------------------
// Batch object behind the C-ABI.  A batch is a list of scenes of one solver
// kind advanced together: every phase of a substep is one launch over all
// scenes.  The single-scene solver (CrHipSolver) is a batch of one.
// ----------------------------------------------------------------------
namespace
{
    struct HostScene
    {
        GridGeom g{};
        float3 gravity = make_float3(0.f, 0.f, 0.f);
        float dt = 0.002f; // mpm_solver_base.h default
        int contactModel = CRMPM_HIP_CONTACT_KINEMATIC;
        int maxParticles = 0;
        int particleOffset = 0;
        unsigned numActive = 0;
        bool resting = false;
        int numShapes = 0;
        int *shapeIds = nullptr;
        CrHipShapeData shapes{};
        CrHipGeometryData geoms{};
        CrHipGeometrySdfData sdf{};
        int blockOffset = 0;
        int shapeOffset = 0, shapeCap = 0;
    };

    constexpr int kUploadSlots = 4;

    bool grid_geom(float cellSize, const float lo[3], const float hi[3], GridGeom &g)
    {
        g.bmin = make_float3(lo[0], lo[1], lo[2]);
        g.bmax = make_float3(hi[0], hi[1], hi[2]);
        g.dx = cellSize;
        g.invdx = 1 / cellSize;
        g.cellVol = cellSize * cellSize * cellSize;
        int n[3];
        for (int a = 0; a < 3; ++a)
            n[a] = (int)((hi[a] - lo[a]) / cellSize + 1.0f);
        g.nx = n[0];
        g.ny = n[1];
        g.nz = n[2];
        if (g.nx < 1 || g.ny < 1 || g.nz < 1)
            return false;
        g.nbx = (g.nx + 3) / 4;
        g.nby = (g.ny + 3) / 4;
        g.nbz = (g.nz + 3) / 4;
        g.numBlocks = g.nbx * g.nby * g.nbz;
        return true;
    }
} // namespace

struct CrHipBatch
{
    int kind = CRMPM_HIP_MLS;
    int iterations = 10;
    HostScene *scenes = nullptr;
    int numScenes = 0;
    long long totalBlocks = 0;
    int capacity = 0; // particles in the bound arrays
    hipStream_t stream = nullptr;
    bool initialized = false;
    bool scenesDirty = true;
    ParticleBind P{};

    // owned
    float *F = nullptr, *C = nullptr; // 9 floats per particle (SolverState)
    float *V0 = nullptr;
    float4 *grid = nullptr, *gridForce = nullptr;
    unsigned char *dirty = nullptr;
    int *segStart = nullptr, *segEnd = nullptr;
    unsigned *activeList = nullptr; // node blocks with partials this substep
    int *activeCount = nullptr;     // [0] active node blocks, [1] occupied pblocks
    unsigned *tileList = nullptr;   // occupied pblocks this substep
    unsigned *keysA = nullptr, *keysB = nullptr, *valsA = nullptr, *valsB = nullptr;
    unsigned *radixCounts = nullptr, *digitTotals = nullptr;
    float4 *tilebuf = nullptr, *tileForce = nullptr;
    ShapeRec *shapeRecs = nullptr;
    int *shapeCounts = nullptr, *slicerList = nullptr;
    int shapeCapTotal = 0;
    SceneRec *dScenes = nullptr;
    SceneShapes *dShapes = nullptr;
    int *dActivePrefix = nullptr;
    int *dBlockScene = nullptr;
    PbStage st{};
    float *stageMem = nullptr;

    // pinned staging of the per-scene tables (a ring, so an upload never
    // overwrites one the stream has not consumed yet)
    char *hStage = nullptr;
    size_t stageBytes = 0;
    hipEvent_t stageDone[kUploadSlots] = {};
    int stageSlot = 0;
    int totalActive = 0;

    // launch sizes in wavefronts; 0 = one wavefront per (sub-)tile
    int tileWaves = 0, updateWaves = 0;
    // sub-tile split: particles per sub-tile, sub-tiles per pblock (<= subStride,
    // the allocated partial boxes per pblock)
    int subSize = 64, subMax = 1, subStride = 1;
    int gSubSize = 1 << 20, gSubMax = 1; // gather kernels' split
    // wavefronts per tile of the PB and MLS scatters (P2G) and the
    // adjoints' velocity scatters (1, 2 or 4; wg_tile_scatter), or 0:
    // four for tiles of more than 64 particles, one for the others
    int scatterWaves = 1;
    // tile-ordered particle permutation of the last enqueued substep, and its sorted keys
    const unsigned *lastOrder = nullptr;
    const unsigned *lastKeys = nullptr;
    // When set, the grid update also writes its nodes (the active node
    // blocks, in the grid's layout) to tape (MLS) or tape + iteration * nodes
    // (PB): the adjoint's re-run of a substep
    float4 *tape = nullptr;
    const unsigned *gatherOrder = nullptr; // this substep's tiles in particle index order

    // hipGraph replay of the substep
    bool graphMode = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    long long graphKey[10] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    long long bindEpoch = 0;

    // optional per-launch event timing (ProfileBegin/End)
    int profMax = 0, profUsed = 0;
    unsigned profMask = ~0u;
    hipEvent_t *profEvents = nullptr;
    int *profCat = nullptr;
};

struct CrHipSolver
{
    CrHipBatch *b = nullptr;
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

    bool valid_scene(const CrHipBatch *b, int s) { return b && s >= 0 && s < b->numScenes; }

    // Shape record regions: scene s owns [shapeOffset, shapeOffset + shapeCap).
    int ensure_shape_capacity(CrHipBatch *b)
    {
        bool grow = !b->shapeRecs;
        for (int s = 0; s < b->numScenes; ++s)
            grow = grow || b->scenes[s].numShapes > b->scenes[s].shapeCap;
        if (!grow)
            return CRMPM_HIP_OK;
        if (b->initialized)
            (void)hipStreamSynchronize(b->stream);
        int total = 0;
        for (int s = 0; s < b->numScenes; ++s)
        {
            HostScene &h = b->scenes[s];
            h.shapeCap = std::max(h.shapeCap, std::max(h.numShapes, 16));
            h.shapeOffset = total;
            total += h.shapeCap;
        }
        dfree(b->shapeRecs);
        dfree(b->slicerList);
        if (dalloc(&b->shapeRecs, total) || dalloc(&b->slicerList, total))
            return CRMPM_HIP_ERR_ALLOC;
        b->shapeCapTotal = total;
        b->scenesDirty = true;
        ++b->bindEpoch;
        return CRMPM_HIP_OK;
    }

    ShapeView shape_view(const CrHipBatch *b, int s)
    {
        const HostScene &h = b->scenes[s];
        ShapeView v;
        v.recs = b->shapeRecs + h.shapeOffset;
        v.counts = b->shapeCounts + 4 * s;
        v.slicers = b->slicerList + h.shapeOffset;
        v.force = reinterpret_cast<float4 *>(h.shapes.force);
        v.torque = reinterpret_cast<float4 *>(h.shapes.torque);
        v.sdfDims = reinterpret_cast<const int4 *>(h.sdf.dimension);
        v.sdfLbcs = reinterpret_cast<const float4 *>(h.sdf.lowerBoundCellSize);
        v.sdfData = reinterpret_cast<const float4 *>(h.sdf.gradientSignedDistance);
        v.effectiveMass = h.contactModel == CRMPM_HIP_CONTACT_EFFECTIVE_MASS;
        return v;
    }

    // Upload the per-scene records, shape bindings and active-particle prefix
    // when any of them changed since the last step.
    int upload_scenes(CrHipBatch *b)
    {
        if (!b->scenesDirty)
            return CRMPM_HIP_OK;
        int rc = ensure_shape_capacity(b);
        if (rc)
            return rc;
        const int ns = b->numScenes;
        char *slot = b->hStage + (size_t)b->stageSlot * b->stageBytes;
        if (hipEventSynchronize(b->stageDone[b->stageSlot]) != hipSuccess)
            return CRMPM_HIP_ERR_LAUNCH;
        SceneRec *rec = reinterpret_cast<SceneRec *>(slot);
        SceneShapes *sh = reinterpret_cast<SceneShapes *>(slot + ns * sizeof(SceneRec));
        int *prefix = reinterpret_cast<int *>(slot + ns * (sizeof(SceneRec) + sizeof(SceneShapes)));
        int active = 0;
        for (int s = 0; s < ns; ++s)
        {
            const HostScene &h = b->scenes[s];
            rec[s].g = h.g;
            rec[s].gravity = h.gravity;
            rec[s].dt = h.dt;
            rec[s].sv = shape_view(b, s);
            rec[s].blockOffset = h.blockOffset;
            rec[s].particleOffset = h.particleOffset;
            const unsigned na = h.resting ? 0u : h.numActive;
            rec[s].numActive = (int)na;
            rec[s].activeOffset = active;
            rec[s].resting = h.resting ? 1 : 0;
            sh[s].numShapes = h.numShapes;
            sh[s].shapeIds = h.shapeIds;
            sh[s].sd = h.shapes;
            sh[s].gd = h.geoms;
            prefix[s] = active;
            active += (int)na;
        }
        prefix[ns] = active;
        b->totalActive = active;
        if (hipMemcpyAsync(b->dScenes, rec, ns * sizeof(SceneRec), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
            hipMemcpyAsync(b->dShapes, sh, ns * sizeof(SceneShapes), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
            hipMemcpyAsync(b->dActivePrefix, prefix, (ns + 1) * sizeof(int), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
            hipEventRecord(b->stageDone[b->stageSlot], b->stream) != hipSuccess)
            return CRMPM_HIP_ERR_LAUNCH;
        b->stageSlot = (b->stageSlot + 1) % kUploadSlots;
        b->scenesDirty = false;
        return CRMPM_HIP_OK;
    }

    // Kernel launch on the batch stream.  While profiling, the dispatch itself
    // records a start and a stop event (hipExtLaunchKernelGGL), so the timing
    // covers the kernel's execution only.
    template <typename... Params, typename... Args>
    void launch(CrHipBatch *b, int cat, void (*kernel)(Params...), dim3 grid, dim3 block, Args... args)
    {
        if (b->profEvents && ((b->profMask >> cat) & 1u) && b->profUsed < b->profMax)
        {
            const int slot = b->profUsed++;
            b->profCat[slot] = cat;
            hipExtLaunchKernelGGL(kernel, grid, block, 0, b->stream, b->profEvents[2 * slot], b->profEvents[2 * slot + 1], 0,
                                  static_cast<Params>(args)...);
        }
        else
            hipLaunchKernelGGL(kernel, grid, block, 0, b->stream, static_cast<Params>(args)...);
    }

    BatchArgs batch_args(const CrHipBatch *b, const unsigned *order)
    {
        BatchArgs a;
        a.scenes = b->dScenes;
        a.blockScene = b->dBlockScene;
        a.numScenes = b->numScenes;
        a.totalBlocks = (int)b->totalBlocks;
        a.P = b->P;
        a.S = SolverState{b->F, b->C, b->V0, b->capacity};
        a.segStart = b->segStart;
        a.segEnd = b->segEnd;
        a.order = order;
        a.gatherOrder = b->gatherOrder;
        a.activeList = b->activeList;
        a.activeCount = b->activeCount;
        a.tileList = b->tileList;
        a.tileCount = b->activeCount + 1;
        a.tapeGrid = nullptr;
        a.tilebuf = b->tilebuf;
        a.tileForce = b->tileForce;
        a.grid = b->grid;
        a.subSize = b->subSize;
        a.subMax = b->subMax;
        a.subStride = b->subStride;
        a.gSubSize = b->gSubSize;
        a.gSubMax = b->gSubMax;
        a.tileClass = 0;
        return a;
    }

    // Workgroups of a launch over `items` (pblocks or sub-tiles); waves = 0:
    // one wavefront per item, otherwise a persistent launch of that many.
    int pblock_wgs(long long items, int waves)
    {
        const long long all = grid1d(items, kWaves);
        return (int)std::max(1ll, waves > 0 ? std::min<long long>(all, grid1d(waves, kWaves)) : all);
    }

    // Keys per particle and a stable LSD radix sort of (key, particle index).
    // Returns the sorted keys and particle order (one of the ping-pong buffers).
    int sort_particles(CrHipBatch *b, int n, int useMask, const unsigned **orderOut, const unsigned **keysOut)
    {
        int bits = 0;
        while ((1ll << bits) <= b->totalBlocks)
            ++bits;
        const int passes = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
        const int digitBits = std::max(8, (bits + passes - 1) / passes);
        const int numTiles = grid1d(n, kRadixTile);
        const BatchArgs a = batch_args(b, nullptr);
        unsigned *kin = b->keysA, *vin = b->valsA, *kout = b->keysB, *vout = b->valsB;
        for (int pass = 0; pass < passes; ++pass)
        {
            const int shift = digitBits * pass;
            auto passKernels = [&](auto bitsTag) {
                constexpr int B = decltype(bitsTag)::value;
                if (pass == 0)
                    launch(b, CRMPM_HIP_PROF_BINNING, k_bin_hist<B>, numTiles, kWG, a, b->dActivePrefix, n, useMask, kin,
                           vin, b->radixCounts, numTiles);
                else
                    launch(b, CRMPM_HIP_PROF_BINNING, k_radix_hist<B>, numTiles, kWG, kin, n, shift, b->radixCounts,
                           numTiles);
                launch(b, CRMPM_HIP_PROF_BINNING, k_radix_scan, 1 << B, kWG, b->radixCounts, numTiles, b->digitTotals);
                launch(b, CRMPM_HIP_PROF_BINNING, k_radix_scatter<B>, numTiles, kWG, kin, vin, n, shift, b->radixCounts,
                       numTiles, b->digitTotals, kout, vout);
            };
            if (digitBits == 8)
                passKernels(std::integral_constant<int, 8>());
            else if (digitBits == 9)
                passKernels(std::integral_constant<int, 9>());
            else
                passKernels(std::integral_constant<int, 10>());
            std::swap(kin, kout);
            std::swap(vin, vout);
        }
        *orderOut = vin;
        *keysOut = kin;
        return check_launch();
    }

    // Bin, sort and segment the particles taking part in this step, then
    // order each tile by (cell rank, cell).
    int build_tiles(CrHipBatch *b, int n, int useMask, const unsigned **orderOut, const unsigned **keysOut)
    {
        int rc = sort_particles(b, n, useMask, orderOut, keysOut);
        if (rc)
            return rc;
        launch(b, CRMPM_HIP_PROF_BINNING, k_segments, grid1d(n, 256), 256, n, (int)b->totalBlocks, *keysOut, b->segStart,
               b->segEnd);
        launch(b, CRMPM_HIP_PROF_BINNING, k_active_blocks, grid1d(b->totalBlocks, 256), 256, batch_args(b, nullptr),
               b->activeList, b->activeCount, b->tileList);
        // the sort's spare ping-pong buffers take the cell ranks and the tile-ordered permutation
        unsigned *tmp = (*keysOut == b->keysA) ? b->keysB : b->keysA;
        unsigned *order2 = (*orderOut == b->valsA) ? b->valsB : b->valsA;
        launch(b, CRMPM_HIP_PROF_BINNING, k_cell_rank_order, pblock_wgs(b->totalBlocks, b->tileWaves), kWG,
               batch_args(b, nullptr), *orderOut, tmp, order2);
        // G2P visits each tile in particle index order, the sort's own order:
        // neighbouring lanes then write neighbouring particles' records.
        b->gatherOrder = *orderOut;
        *orderOut = order2;
        return check_launch();
    }

    int validate_for_step(const CrHipBatch *b)
    {
        if (!b->initialized || !b->P.xm || !b->P.v || !b->P.params || !b->P.type || !b->P.mask)
            return CRMPM_HIP_ERR_STATE;
        for (int s = 0; s < b->numScenes; ++s)
        {
            const HostScene &h = b->scenes[s];
            if (h.numShapes > 0 && (!h.shapeIds || !h.shapes.type || !h.geoms.type))
                return CRMPM_HIP_ERR_STATE;
        }
        return CRMPM_HIP_OK;
    }

    template <int Kind>
    auto update_kernel(bool multi, bool persistent)
    {
        return multi ? (persistent ? k_update<Kind, true, true> : k_update<Kind, true, false>)
                     : (persistent ? k_update<Kind, false, true> : k_update<Kind, false, false>);
    }

    // The launches of one substep (all scenes).
    int enqueue_step(CrHipBatch *b)
    {
        const int blockWGs = grid1d(b->totalBlocks, 64 * kWaves); // 64 node blocks per wavefront
        launch(b, CRMPM_HIP_PROF_BINNING, k_prepare, blockWGs + b->numScenes, kWG, (int)b->totalBlocks, blockWGs, b->grid,
               b->gridForce, b->dirty, b->segStart, b->segEnd, (const SceneShapes *)b->dShapes, (const SceneRec *)b->dScenes,
               (const int *)b->dBlockScene, b->numScenes, b->activeCount);
        const int n = b->totalActive;
        if (n == 0)
            return check_launch();
        const unsigned *order = nullptr, *keys = nullptr;
        // The standard solver's P2G ignores the active mask (src/solver/cpu_standard_mpm_solver.cpp:101-120).
        int rc = build_tiles(b, n, b->kind != CRMPM_HIP_STANDARD, &order, &keys);
        if (rc)
            return rc;
        b->lastOrder = order;
        b->lastKeys = keys;
        const BatchArgs a = batch_args(b, order);
        const int tileWGs = pblock_wgs(b->totalBlocks * b->subMax, b->tileWaves);
        const int gatherWGs = pblock_wgs(b->totalBlocks * b->gSubMax, b->tileWaves);
        // gather split 0: G2P per particle in particle index order (one scene)
        const bool gDense = b->gSubMax == 0;
        const int updWGs = pblock_wgs(b->totalBlocks, b->updateWaves);
        const bool pers = b->tileWaves > 0;
        const bool multi = b->subMax > 1, persU = b->updateWaves > 0;
        // PB scatter: one-wavefront workgroups, one per item or, persistent,
        // one per launched wavefront
        const long long tileItems = (long long)b->totalBlocks * b->subMax;
        const dim3 tileGrid((unsigned)std::max(1ll, pers ? std::min<long long>(b->tileWaves, tileItems) : tileItems));
        const dim3 tileBlock(64);
        auto upd1 = update_kernel<1>(multi, persU);
        if (b->kind == CRMPM_HIP_MLS)
        {
            // MLS and standard MPM keep four-wavefront workgroups: their sparse
            // grids have many empty items, whose workgroups cost dispatch time
            // (one-wavefront groups: MLS cube substep 0.109 -> 0.117 ms)
            const dim3 wgGrid((unsigned)std::max(1ll, tileItems));
            if (b->scatterWaves == 4)
                launch(b, CRMPM_HIP_PROF_MLS_P2G, k_mls_p2g_wg<4>, wgGrid, dim3(256), a);
            else if (b->scatterWaves == 2)
                launch(b, CRMPM_HIP_PROF_MLS_P2G, k_mls_p2g_wg<2>, wgGrid, dim3(128), a);
            else if (b->scatterWaves == 1)
                launch(b, CRMPM_HIP_PROF_MLS_P2G, pers ? k_mls_p2g<true> : k_mls_p2g<false>, tileWGs, kWG, a);
            else
            {
                // adaptive: tiles of one chunk on one wavefront, the others on four
                BatchArgs small = a, large = a;
                small.tileClass = 1;
                large.tileClass = 2;
                launch(b, CRMPM_HIP_PROF_MLS_P2G, pers ? k_mls_p2g<true> : k_mls_p2g<false>, tileWGs, kWG, small);
                launch(b, CRMPM_HIP_PROF_MLS_P2G, k_mls_p2g_wg<4>, wgGrid, dim3(256), large);
            }
            BatchArgs au = a;
            au.tapeGrid = b->tape;
            launch(b, CRMPM_HIP_PROF_UPDATE, update_kernel<0>(multi, persU), updWGs, kWG, au, b->dirty);
            if (gDense)
                launch(b, CRMPM_HIP_PROF_MLS_G2P, k_g2p_dense<0>, grid1d(n, kWG), kWG, a, n);
            else
                launch(b, CRMPM_HIP_PROF_MLS_G2P, pers ? k_mls_g2p<true> : k_mls_g2p<false>, gatherWGs, kWG, a);
        }
        else if (b->kind == CRMPM_HIP_STANDARD)
        {
            launch(b, CRMPM_HIP_PROF_STD_P2G, pers ? k_std_p2g<true> : k_std_p2g<false>, tileWGs, kWG, a);
            launch(b, CRMPM_HIP_PROF_UPDATE, update_kernel<2>(multi, persU), updWGs, kWG, a, b->dirty);
            if (gDense)
                launch(b, CRMPM_HIP_PROF_STD_G2P, k_g2p_dense<2>, grid1d(n, kWG), kWG, a, n);
            else
                launch(b, CRMPM_HIP_PROF_STD_G2P, pers ? k_std_g2p<true> : k_std_g2p<false>, gatherWGs, kWG, a);
        }
        else
        {
            const int K = b->iterations;
            if (K == 0)
            {
                launch(b, CRMPM_HIP_PROF_BINNING, k_pb_stage, grid1d(n, 256), 256, n, (int)b->totalBlocks, keys, order, b->P,
                       a.S, b->st);
                launch(b, CRMPM_HIP_PROF_PB_G2P, k_pb_integrate_only, grid1d(n, 256), 256, a, n, keys, b->st);
            }
            else
            {
                const int denseWGs = grid1d(n, kWG);
                const size_t nodes = (size_t)b->totalBlocks * 64;
                // the adjoint's re-run of a substep: each iteration's update
                // also writes its nodes to the tape
                auto taped = [&](int it) {
                    BatchArgs t = a;
                    t.tapeGrid = b->tape ? b->tape + (size_t)it * nodes : nullptr;
                    return t;
                };
                // the scatter: one wavefront per tile (persistent or not), or
                // scatterWaves wavefronts per tile, one workgroup per slot
                const dim3 wgGrid((unsigned)std::max(1ll, tileItems));
                auto scatterW = [&](bool mass, int prof, int w, const BatchArgs &ar) {
                    switch (w)
                    {
                    case 2:
                        launch(b, prof, mass ? k_pb_scatter_wg<true, 2> : k_pb_scatter_wg<false, 2>, wgGrid, dim3(128), ar,
                               b->st);
                        break;
                    case 4:
                        launch(b, prof, mass ? k_pb_scatter_wg<true, 4> : k_pb_scatter_wg<false, 4>, wgGrid, dim3(256), ar,
                               b->st);
                        break;
                    default:
                        if (mass)
                            launch(b, prof, pers ? k_pb_scatter<true, true, 1> : k_pb_scatter<true, false, 1>, tileGrid,
                                   tileBlock, ar, b->st);
                        else
                            launch(b, prof, pers ? k_pb_scatter<false, true, 1> : k_pb_scatter<false, false, 1>, tileGrid,
                                   tileBlock, ar, b->st);
                    }
                };
                // adaptive (0): tiles of one chunk on one wavefront, the others on four
                auto scatter = [&](bool mass, int prof) {
                    if (b->scatterWaves != 0)
                        return scatterW(mass, prof, b->scatterWaves, a);
                    BatchArgs small = a, large = a;
                    small.tileClass = 1;
                    large.tileClass = 2;
                    scatterW(mass, prof, 1, small);
                    scatterW(mass, prof, 4, large);
                };
                launch(b, CRMPM_HIP_PROF_PB_P2G, k_pb_dense<0>, denseWGs, kWG, a, b->st, n, keys);
                scatter(true, CRMPM_HIP_PROF_PB_P2G);
                launch(b, CRMPM_HIP_PROF_UPDATE, upd1, updWGs, kWG, taped(0), b->dirty);
                for (int it = 1; it < K; ++it)
                {
                    launch(b, CRMPM_HIP_PROF_PB_FUSED, k_pb_dense<1>, denseWGs, kWG, a, b->st, n, keys);
                    scatter(false, CRMPM_HIP_PROF_PB_SCATTER);
                    launch(b, CRMPM_HIP_PROF_UPDATE, upd1, updWGs, kWG, taped(it), b->dirty);
                }

                launch(b, CRMPM_HIP_PROF_PB_G2P, k_pb_dense<2>, denseWGs, kWG, a, b->st, n, keys);
            }
        }
        return check_launch();
    }

    void drop_graph(CrHipBatch *b)
    {
        if (b->graphExec)
            (void)hipGraphExecDestroy(b->graphExec);
        if (b->graph)
            (void)hipGraphDestroy(b->graph);
        b->graphExec = nullptr;
        b->graph = nullptr;
        b->graphKey[0] = -1;
    }

    // One substep: uploads scene records if needed, then either enqueues the
    // launches or replays the captured graph of them.  The graph holds only
    // kernel nodes (no memset or copy nodes); it is recaptured whenever the
    // launch configuration changes.
    int batch_step(CrHipBatch *b)
    {
        int rc = validate_for_step(b);
        if (rc)
            return rc;
        rc = upload_scenes(b);
        if (rc)
            return rc;
        // substeps whose launches record profiling events run eagerly
        const bool useGraph = b->graphMode && (!b->profEvents || b->profMask == 0u) && b->stream != nullptr;
        if (!useGraph)
            return enqueue_step(b);
        const long long key[10] = {b->totalActive, b->iterations, b->tileWaves, b->updateWaves, b->bindEpoch,
                                   b->subSize,       b->subMax,     b->gSubSize,  b->gSubMax,    b->scatterWaves};
        if (!b->graphExec || !std::equal(key, key + 10, b->graphKey))
        {
            drop_graph(b);
            if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return CRMPM_HIP_ERR_LAUNCH;
            const int erc = enqueue_step(b);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(b->stream, &g);
            if (erc || ec != hipSuccess || !g)
            {
                if (g)
                    (void)hipGraphDestroy(g);
                return erc ? erc : CRMPM_HIP_ERR_LAUNCH;
            }
            b->graph = g;
            if (hipGraphInstantiate(&b->graphExec, g, nullptr, nullptr, 0) != hipSuccess)
            {
                drop_graph(b);
                return CRMPM_HIP_ERR_LAUNCH;
            }
            std::copy(key, key + 10, b->graphKey);
        }
        return hipGraphLaunch(b->graphExec, b->stream) == hipSuccess ? CRMPM_HIP_OK : CRMPM_HIP_ERR_LAUNCH;
    }

    // One substep, launched eagerly, with each grid update also writing its
    // active node blocks to tape (MLS: one grid of the batch's node count;
    // PB: K, one per iteration).
    int batch_step_taped(CrHipBatch *b, float4 *tape)
    {
        int rc = validate_for_step(b);
        if (!rc)
            rc = upload_scenes(b);
        if (rc)
            return rc;
        b->tape = tape;
        rc = enqueue_step(b);
        b->tape = nullptr;
        return rc;
    }

    int batch_destroy(CrHipBatch *b)
    {
        if (b->initialized)
            (void)hipStreamSynchronize(b->stream);
        drop_graph(b);
        dfree(b->F);
        dfree(b->C);
        dfree(b->V0);
        dfree(b->grid);
        dfree(b->gridForce);
        dfree(b->dirty);
        dfree(b->activeList);
        dfree(b->activeCount);
        dfree(b->tileList);
        dfree(b->segStart);
        dfree(b->segEnd);
        dfree(b->keysA);
        dfree(b->keysB);
        dfree(b->valsA);
        dfree(b->valsB);
        dfree(b->radixCounts);
        dfree(b->digitTotals);
        dfree(b->tilebuf);
        dfree(b->tileForce);
        dfree(b->shapeRecs);
        dfree(b->shapeCounts);
        dfree(b->slicerList);
        dfree(b->dScenes);
        dfree(b->dShapes);
        dfree(b->dActivePrefix);
        dfree(b->dBlockScene);
        dfree(b->stageMem);
        for (int i = 0; i < kUploadSlots; ++i)
            if (b->stageDone[i])
                (void)hipEventDestroy(b->stageDone[i]);
        if (b->hStage)
            (void)hipHostFree(b->hStage);
        if (b->profEvents)
        {
            for (int i = 0; i < 2 * b->profMax; ++i)
                (void)hipEventDestroy(b->profEvents[i]);
            delete[] b->profEvents;
            delete[] b->profCat;
        }
        delete[] b->scenes;
        delete b;
        return CRMPM_HIP_OK;
    }

    CrHipBatch *batch_new(int kind, int numScenes)
    {
        CrHipBatch *b = new (std::nothrow) CrHipBatch();
        if (!b)
            return nullptr;
        b->scenes = new (std::nothrow) HostScene[numScenes];
        if (!b->scenes)
        {
            delete b;
            return nullptr;
        }
        b->kind = kind;
        b->numScenes = numScenes;
        return b;
    }

    int batch_initialize(CrHipBatch *b)
    {
        if (b->initialized)
            return CRMPM_HIP_ERR_STATE;
        const size_t N = (size_t)std::max(b->capacity, 1);
        const size_t blocks = (size_t)b->totalBlocks;
        const size_t nodes = blocks * 64;
        const int radixTiles = grid1d((long long)N, kRadixTile);
        const int ns = b->numScenes;
        int rc = 0;
        rc |= dalloc(&b->F, 9 * N);
        rc |= dalloc(&b->C, 9 * N);
        rc |= dalloc(&b->V0, N);
        rc |= dalloc(&b->grid, nodes);
        if (b->kind == CRMPM_HIP_STANDARD)
            rc |= dalloc(&b->gridForce, nodes);
        rc |= dalloc(&b->dirty, blocks);
        rc |= dalloc(&b->activeList, blocks);
        rc |= dalloc(&b->activeCount, 2);
        rc |= dalloc(&b->tileList, blocks);
        rc |= dalloc(&b->segStart, blocks);
        rc |= dalloc(&b->segEnd, blocks);
        rc |= dalloc(&b->keysA, N);
        rc |= dalloc(&b->keysB, N);
        rc |= dalloc(&b->valsA, N);
        rc |= dalloc(&b->valsB, N);
        rc |= dalloc(&b->radixCounts, (size_t)radixTiles * kMaxDigits);
        rc |= dalloc(&b->digitTotals, kMaxDigits);
        // partial boxes: subStride per pblock (up to 8, fewer for very large
        // batches so the boxes stay within 8 GiB)
        const size_t box = (size_t)kTileNodes * sizeof(float4);
        b->subStride = (int)std::max<size_t>(1, std::min<size_t>(8, (size_t(8) << 30) / std::max<size_t>(blocks * box, 1)));
        b->subMax = std::min(b->subMax, b->subStride);
        rc |= dalloc(&b->tilebuf, blocks * b->subStride * kTileNodes);
        if (b->kind == CRMPM_HIP_STANDARD)
            rc |= dalloc(&b->tileForce, blocks * b->subStride * kTileNodes);
        rc |= dalloc(&b->shapeCounts, 4 * (size_t)ns);
        rc |= dalloc(&b->dScenes, ns);
        rc |= dalloc(&b->dShapes, ns);
        rc |= dalloc(&b->dActivePrefix, ns + 1);
        rc |= dalloc(&b->dBlockScene, blocks);
        if (b->kind == CRMPM_HIP_PB)
        {
            // xs (4) + vs (3) + Fs (9) + Cs (9) + ps (2) + ts (1) floats per particle
            rc |= dalloc(&b->stageMem, 28 * N);
            if (!rc)
            {
                float *m = b->stageMem;
                b->st.stride = (int)N;
                b->st.xs = reinterpret_cast<float4 *>(m);
                m += 4 * N;
                b->st.vs = m;
                m += 3 * N;
                b->st.Fs = m;
                m += 9 * N;
                b->st.Cs = m;
                m += 9 * N;
                b->st.ps = reinterpret_cast<float2 *>(m);
                m += 2 * N;
                b->st.ts = reinterpret_cast<int *>(m);
            }
        }
        b->stageBytes = ((ns * (sizeof(SceneRec) + sizeof(SceneShapes)) + (ns + 1) * sizeof(int)) + 255) & ~size_t(255);
        if (!rc && hipHostMalloc((void **)&b->hStage, b->stageBytes * kUploadSlots, hipHostMallocDefault) != hipSuccess)
        {
            b->hStage = nullptr;
            rc = CRMPM_HIP_ERR_ALLOC;
        }
        for (int i = 0; i < kUploadSlots && !rc; ++i)
            if (hipEventCreateWithFlags(&b->stageDone[i], hipEventDisableTiming) != hipSuccess)
                rc = CRMPM_HIP_ERR_ALLOC;
        if (!rc)
            rc = ensure_shape_capacity(b);
        if (rc)
            return CRMPM_HIP_ERR_ALLOC;
        // pblock -> scene map
        int *map = new (std::nothrow) int[blocks];
        if (!map)
            return CRMPM_HIP_ERR_ALLOC;
        for (int s = 0; s < ns; ++s)
            for (int k = 0; k < b->scenes[s].g.numBlocks; ++k)
                map[b->scenes[s].blockOffset + k] = s;
        const hipError_t mc = hipMemcpy(b->dBlockScene, map, blocks * sizeof(int), hipMemcpyHostToDevice);
        delete[] map;
        if (mc != hipSuccess)
            return CRMPM_HIP_ERR_LAUNCH;
        hipStream_t st = b->stream;
        rc |= hipMemsetAsync(b->F, 0, 9 * N * sizeof(float), st);
        rc |= hipMemsetAsync(b->C, 0, 9 * N * sizeof(float), st);
        rc |= hipMemsetAsync(b->V0, 0, N * sizeof(float), st);
        rc |= hipMemsetAsync(b->grid, 0, nodes * sizeof(float4), st);
        if (b->gridForce)
            rc |= hipMemsetAsync(b->gridForce, 0, nodes * sizeof(float4), st);
        rc |= hipMemsetAsync(b->dirty, 0, blocks, st);
        rc |= hipMemsetAsync(b->segStart, 0, blocks * sizeof(int), st);
        rc |= hipMemsetAsync(b->segEnd, 0, blocks * sizeof(int), st);
        rc |= hipMemsetAsync(b->shapeCounts, 0, 4 * ns * sizeof(int), st);
        if (rc)
            return CRMPM_HIP_ERR_LAUNCH;
        b->initialized = true;
        b->scenesDirty = true;
        ++b->bindEpoch;
        return CRMPM_HIP_OK;
    }

    int batch_compute_initial_data(CrHipBatch *b, int s, unsigned n, const unsigned *indices)
    {
        if (!valid_scene(b, s) || (n > 0 && !indices))
            return CRMPM_HIP_ERR_ARG;
        if (!b->initialized || !b->P.xm)
            return CRMPM_HIP_ERR_STATE;
        if (n == 0)
            return CRMPM_HIP_OK;
        int rc = upload_scenes(b);
        if (rc)
            return rc;
        const BatchArgs a = batch_args(b, nullptr);
        // the scene's tile partial region holds its fixed-point mass counters
        // (216 float4 per pblock >= 64 counters per node block)
        const HostScene &h = b->scenes[s];
        unsigned long long *acc =
            reinterpret_cast<unsigned long long *>(b->tilebuf + (size_t)h.blockOffset * b->subStride * kTileNodes);
        if (hipMemsetAsync(acc, 0, (size_t)h.g.numBlocks * 64 * sizeof(unsigned long long), b->stream) != hipSuccess)
            return CRMPM_HIP_ERR_LAUNCH;
        k_init_mass<<<grid1d(n, 256), 256, 0, b->stream>>>(a, s, n, indices, acc);
        k_init_apply<<<grid1d(h.g.numBlocks, kWaves), kWG, 0, b->stream>>>(a, s, acc, b->dirty);
        k_init_volume<<<grid1d(n, 256), 256, 0, b->stream>>>(a, s, n, indices);
        return check_launch();
    }

    int batch_copy_grid(CrHipBatch *b, int s, float *dst)
    {
        if (!valid_scene(b, s) || !dst)
            return CRMPM_HIP_ERR_ARG;
        if (!b->initialized)
            return CRMPM_HIP_ERR_STATE;
        int rc = upload_scenes(b);
        if (rc)
            return rc;
        const GridGeom &g = b->scenes[s].g;
        const long long total = (long long)g.nx * g.ny * g.nz;
        k_copy_grid<<<grid1d(total, 256), 256, 0, b->stream>>>(batch_args(b, nullptr), s, reinterpret_cast<float4 *>(dst));
        return check_launch();
    }

    int profile_begin(CrHipBatch *b, int maxLaunches)
    {
        if (!b || maxLaunches <= 0)
            return CRMPM_HIP_ERR_ARG;
        if (b->profEvents)
            return CRMPM_HIP_ERR_STATE;
        b->profEvents = new (std::nothrow) hipEvent_t[2 * (size_t)maxLaunches]();
        b->profCat = new (std::nothrow) int[maxLaunches];
        if (!b->profEvents || !b->profCat)
            return CRMPM_HIP_ERR_ALLOC;
        b->profMax = maxLaunches;
        for (int i = 0; i < 2 * maxLaunches; ++i)
            if (hipEventCreate(&b->profEvents[i]) != hipSuccess)
                return CRMPM_HIP_ERR_ALLOC;
        b->profUsed = 0;
        b->profMask = ~0u;
        return CRMPM_HIP_OK;
    }

    int profile_end(CrHipBatch *b, float *msPerCategory, int *launchesPerCategory, int numCategories)
    {
        if (!b || !b->profEvents)
            return CRMPM_HIP_ERR_STATE;
        int rc = CRMPM_HIP_OK;
        if (hipStreamSynchronize(b->stream) != hipSuccess)
            rc = CRMPM_HIP_ERR_LAUNCH;
        for (int c = 0; c < numCategories; ++c)
        {
            if (msPerCategory)
                msPerCategory[c] = 0.0f;
            if (launchesPerCategory)
                launchesPerCategory[c] = 0;
        }
        for (int i = 0; i < b->profUsed && rc == CRMPM_HIP_OK; ++i)
        {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, b->profEvents[2 * i], b->profEvents[2 * i + 1]) != hipSuccess)
            {
                rc = CRMPM_HIP_ERR_LAUNCH;
                break;
            }
            const int c = b->profCat[i];
            if (c >= 0 && c < numCategories)
            {
                if (msPerCategory)
                    msPerCategory[c] += ms;
                if (launchesPerCategory)
                    launchesPerCategory[c] += 1;
            }
        }
        for (int i = 0; i < 2 * b->profMax; ++i)
            (void)hipEventDestroy(b->profEvents[i]);
        delete[] b->profEvents;
        delete[] b->profCat;
        b->profEvents = nullptr;
        b->profCat = nullptr;
        b->profMax = b->profUsed = 0;
        return rc;
    }

    // Autotuner state copy: everything a substep changes.
    struct TuneBackup
    {
        void *mem = nullptr;
        size_t bytes = 0;
    };

    struct Span
    {
        void *p;
        size_t bytes;
    };

    int tune_spans(CrHipBatch *b, Span *sp, int maxSpans)
    {
        const size_t N = (size_t)std::max(b->capacity, 1);
        int k = 0;
        auto add = [&](void *p, size_t bytes) {
            if (p && bytes && k < maxSpans)
                sp[k++] = Span{p, bytes};
        };
        add(b->P.xm, N * sizeof(float4));
        add(b->P.v, N * sizeof(float4));
        add(b->F, 9 * N * sizeof(float));
        add(b->C, 9 * N * sizeof(float));
        add(b->V0, N * sizeof(float));
        add(b->grid, (size_t)b->totalBlocks * 64 * sizeof(float4));
        add(b->gridForce, (size_t)b->totalBlocks * 64 * sizeof(float4));
        add(b->dirty, (size_t)b->totalBlocks);
        for (int s = 0; s < b->numScenes; ++s)
        {
            const HostScene &h = b->scenes[s];
            if (h.numShapes > 0)
            {
                add(h.shapes.force, (size_t)h.shapes.size * sizeof(float4));
                add(h.shapes.torque, (size_t)h.shapes.size * sizeof(float4));
            }
        }
        return k;
    }

    int tune_copy(CrHipBatch *b, char *backup, bool save)
    {
        Span sp[64];
        const int k = tune_spans(b, sp, 64);
        size_t off = 0;
        for (int i = 0; i < k; ++i)
        {
            void *dst = save ? (void *)(backup + off) : sp[i].p;
            const void *src = save ? sp[i].p : (const void *)(backup + off);
            if (hipMemcpyAsync(dst, src, sp[i].bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
                return CRMPM_HIP_ERR_LAUNCH;
            off += (sp[i].bytes + 255) & ~size_t(255);
        }
        return CRMPM_HIP_OK;
    }

    // Time `steps` substeps from the saved state with the given launch sizes
    // and sub-tile split.
    int tune_trial(CrHipBatch *b, char *backup, int steps, int tileWaves, int updateWaves, int subSize, int subMax,
                   int gSubSize, int gSubMax, float *ms)
    {
        b->tileWaves = tileWaves;
        b->updateWaves = updateWaves;
        b->subSize = subSize;
        b->subMax = subMax;
        b->gSubSize = gSubSize;
        b->gSubMax = gSubMax;
        int rc = tune_copy(b, backup, false);
        if (rc)
            return rc;
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess)
            return CRMPM_HIP_ERR_ALLOC;
        if (hipEventCreate(&e1) != hipSuccess)
        {
            (void)hipEventDestroy(e0);
            return CRMPM_HIP_ERR_ALLOC;
        }
        rc = enqueue_step(b); // warm-up substep
        if (!rc && hipEventRecord(e0, b->stream) != hipSuccess)
            rc = CRMPM_HIP_ERR_LAUNCH;
        for (int i = 0; i < steps && !rc; ++i)
            rc = enqueue_step(b);
        if (!rc && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(ms, e0, e1) != hipSuccess))
            rc = CRMPM_HIP_ERR_LAUNCH;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    }

    int batch_autotune(CrHipBatch *b, int steps, int *tileWavesOut, int *updateWavesOut, int *subSizeOut,
                       int *subMaxOut)
    {
        if (!b || steps <= 0)
            return CRMPM_HIP_ERR_ARG;
        int rc = validate_for_step(b);
        if (rc)
            return rc;
        if (b->profEvents)
            return CRMPM_HIP_ERR_STATE;
        rc = upload_scenes(b);
        if (rc)
            return rc;
        Span sp[64];
        const int k = tune_spans(b, sp, 64);
        size_t bytes = 0;
        for (int i = 0; i < k; ++i)
            bytes += (sp[i].bytes + 255) & ~size_t(255);
        char *backup = nullptr;
        if (dalloc(&backup, bytes))
            return CRMPM_HIP_ERR_ALLOC;
        // the launch configuration in force, restored if a trial fails
        const int prev[7] = {b->tileWaves, b->updateWaves, b->subSize,     b->subMax,
                             b->gSubSize,  b->gSubMax,    b->scatterWaves};
        b->scatterWaves = 1; // steps 1-4 time one wavefront per scatter tile
        rc = tune_copy(b, backup, true);
        const bool saved = rc == CRMPM_HIP_OK;
        // 1. sub-tile split: 64, 128 or 256 particles per sub-tile, or whole
        //    pblocks (one sub-tile each)
        const int splits[4][2] = {{64, b->subStride}, {128, b->subStride}, {256, b->subStride}, {1 << 20, 1}};
        int bestS = b->subSize, bestM = b->subMax;
        float best = 3.4e38f;
        for (const auto &sp : splits)
        {
            float ms = 0.0f;
            if (rc)
                break;
            rc = tune_trial(b, backup, steps, 0, 0, sp[0], sp[1], 1 << 20, 1, &ms);
            if (!rc && ms < best)
            {
                best = ms;
                bestS = sp[0];
                bestM = sp[1];
            }
        }
        // 2. gather (G2P) split: whole pblocks, sub-tiles of 256 or 128
        //    particles (no partials to sum, only more wavefronts), or, for a
        //    batch of one scene, per particle in particle index order
        const int gsplits[4][2] = {{1 << 20, 1}, {256, 4}, {128, 8}, {1 << 20, 0}};
        int bestGS = 1 << 20, bestGM = 1;
        best = 3.4e38f;
        for (const auto &gs : gsplits)
        {
            float ms = 0.0f;
            if (rc || b->kind == CRMPM_HIP_PB) // PB has no separate gather kernel
                break;
            if (gs[1] == 0 && b->numScenes != 1)
                continue;
            rc = tune_trial(b, backup, steps, 0, 0, bestS, bestM, gs[0], gs[1], &ms);
            if (!rc && ms < best)
            {
                best = ms;
                bestGS = gs[0];
                bestGM = gs[1];
            }
        }
        // 3. launch sizes: one wavefront per item, or a persistent launch of
        //    2, 4, 8 or 16 thousand wavefronts looping over the items
        const int cand[] = {0, 2048, 4096, 8192, 16384};
        int bestT = 0, bestU = 0;
        best = 3.4e38f;
        for (int c : cand)
        {
            float ms = 0.0f;
            if (rc || (c > 0 && c >= b->totalBlocks * bestM))
                continue;
            rc = tune_trial(b, backup, steps, c, 0, bestS, bestM, bestGS, bestGM, &ms);
            if (!rc && ms < best)
            {
                best = ms;
                bestT = c;
            }
        }
        best = 3.4e38f;
        for (int c : cand)
        {
            float ms = 0.0f;
            if (rc || (c > 0 && c >= b->totalBlocks))
                continue;
            rc = tune_trial(b, backup, steps, bestT, c, bestS, bestM, bestGS, bestGM, &ms);
            if (!rc && ms < best)
            {
                best = ms;
                bestU = c;
            }
        }
        // 5. PB and MLS: 2 or 4 wavefronts per scatter tile, or 4 for the
        //    tiles of more than one chunk only (dense pblocks)
        int bestW = 1;
        for (int w : {2, 4, 0})
        {
            float ms = 0.0f;
            if (rc || b->kind == CRMPM_HIP_STANDARD)
                break;
            b->scatterWaves = w;
            rc = tune_trial(b, backup, steps, bestT, bestU, bestS, bestM, bestGS, bestGM, &ms);
            if (!rc && ms < best)
            {
                best = ms;
                bestW = w;
            }
        }
        // the state goes back to where it was whether or not every trial ran
        if (saved)
        {
            const int rr = tune_copy(b, backup, false);
            rc = rc ? rc : rr;
        }
        (void)hipStreamSynchronize(b->stream);
        dfree(backup);
        if (rc)
        {
            b->tileWaves = prev[0];
            b->updateWaves = prev[1];
            b->subSize = prev[2];
            b->subMax = prev[3];
            b->gSubSize = prev[4];
            b->gSubMax = prev[5];
            b->scatterWaves = prev[6];
            return rc;
        }
        b->tileWaves = bestT;
        b->updateWaves = bestU;
        b->subSize = bestS;
        b->subMax = bestM;
        b->gSubSize = bestGS;
        b->gSubMax = bestGM;
        b->scatterWaves = bestW;
        if (tileWavesOut)
            *tileWavesOut = bestT;
        if (updateWavesOut)
            *updateWavesOut = bestU;
        if (subSizeOut)
            *subSizeOut = bestS;
        if (subMaxOut)
            *subMaxOut = bestM;
        return rc;
    }
} // namespace

namespace
{
    // ------------------------------------------------------------------
    // Particle lattice of a particle object: ParticleObjectImpl::initialize
    // (src/engine/particle_object_impl.cpp:36-88).  Lattice point (ix, iy, iz)
    // of the fattened geometry bounds is q = (ix, iy, iz) spacing invScale +
    // min in the geometry's frame; it becomes a particle where the geometry's
    // SDF is <= 0, at rotate(q / invScale) + position.  Particles come out in
    // the reference's loop order (ix outer, iz inner): a count pass per
    // 2048-point block, a scan of the block counts, and a write pass that
    // ranks the inside points of each block.
    // ------------------------------------------------------------------
    struct LatticeArgs
    {
        ShapeView sv; // SDF tables only
        float4 p0;    // geometry params0
        float4 rot;   // quaternion (x, y, z, w)
        float3 lo, invScale, pos;
        float spacing, mass;
        int gtype, ny, nz;
        long long total; // nx * ny * nz
    };

    CR_DEV bool lattice_point(const LatticeArgs &a, long long L, V3 &q)
    {
        const long long yz = (long long)a.ny * a.nz;
        const int ix = (int)(L / yz);
        const int rem = (int)(L - (long long)ix * yz);
        const int iy = rem / a.nz, iz = rem - iy * a.nz;
        q = V3{((float)ix * a.spacing) * a.invScale.x + a.lo.x, ((float)iy * a.spacing) * a.invScale.y + a.lo.y,
               ((float)iz * a.spacing) * a.invScale.z + a.lo.z};
        V3 g, tangent{0, 0, 0};
        float d;
        geometry_sdf(a.sv, q, a.gtype, a.p0, g, d, tangent);
        return d <= 0.0f;
    }

    // Each thread takes kLatticePts consecutive lattice points, a workgroup
    // kLatticeBlock of them.
    constexpr int kLatticePts = 8;
    constexpr int kLatticeBlock = kWG * kLatticePts;

    // Exclusive prefix of v over the workgroup's threads (in thread order)
    // and the workgroup total.
    CR_DEV unsigned wg_exclusive_scan(unsigned v, unsigned &total)
    {
        __shared__ unsigned waveSum[kWaves];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        unsigned inc = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
        {
            const unsigned t = __shfl_up(inc, o, 64);
            if (lane >= o)
                inc += t;
        }
        if (lane == 63)
            waveSum[wave] = inc;
        __syncthreads();
        unsigned before = 0;
        total = 0;
        for (int w = 0; w < kWaves; ++w)
        {
            before += (w < wave) ? waveSum[w] : 0u;
            total += waveSum[w];
        }
        __syncthreads(); // waveSum is reused by the next call
        return before + inc - v;
    }

    // Inside flags of this thread's points (bit j: point first + j).
    CR_DEV unsigned lattice_flags(const LatticeArgs &a, long long first)
    {
        unsigned f = 0;
        for (int j = 0; j < kLatticePts; ++j)
        {
            V3 q;
            if (first + j < a.total && lattice_point(a, first + j, q))
                f |= 1u << j;
        }
        return f;
    }

    // svd3 on n column-major matrices, one per thread (crHipSvd3).
    __global__ __launch_bounds__(kWG) void k_svd3(const float *A9, int n, float *U9, float *Vt9, float *S3)
    {
        const int i = blockIdx.x * kWG + threadIdx.x;
        if (i >= n)
            return;
        const float *a = A9 + 9 * (size_t)i;
        const M3 A{V3{a[0], a[1], a[2]}, V3{a[3], a[4], a[5]}, V3{a[6], a[7], a[8]}};
        M3 u, vt;
        V3 sig;
        svd3(A, u, vt, sig);
        float *uo = U9 + 9 * (size_t)i, *vo = Vt9 + 9 * (size_t)i;
        const V3 *uc = &u.c0, *vc = &vt.c0;
        for (int c = 0; c < 3; ++c)
        {
            uo[3 * c] = uc[c].x;
            uo[3 * c + 1] = uc[c].y;
            uo[3 * c + 2] = uc[c].z;
            vo[3 * c] = vc[c].x;
            vo[3 * c + 1] = vc[c].y;
            vo[3 * c + 2] = vc[c].z;
        }
        S3[3 * (size_t)i] = sig.x;
        S3[3 * (size_t)i + 1] = sig.y;
        S3[3 * (size_t)i + 2] = sig.z;
    }

    __global__ __launch_bounds__(kWG) void k_lattice_count(LatticeArgs a, unsigned *blockCounts)
    {
        const long long first = (long long)blockIdx.x * kLatticeBlock + (long long)threadIdx.x * kLatticePts;
        unsigned total;
        wg_exclusive_scan((unsigned)__popc(lattice_flags(a, first)), total);
        if (threadIdx.x == 0)
            blockCounts[blockIdx.x] = total;
    }

    // Exclusive scan of n block counts in place (one workgroup); total to counts[n].
    __global__ __launch_bounds__(kWG) void k_lattice_scan(int n, unsigned *counts)
    {
        unsigned carry = 0;
        for (int base = 0; base < n; base += kWG)
        {
            const int i = base + (int)threadIdx.x;
            const unsigned v = (i < n) ? counts[i] : 0u;
            unsigned total;
            const unsigned ex = wg_exclusive_scan(v, total);
            if (i < n)
                counts[i] = carry + ex;
            carry += total;
        }
        if (threadIdx.x == 0)
            counts[n] = carry;
    }

    __global__ __launch_bounds__(kWG) void k_lattice_write(LatticeArgs a, const unsigned *blockOffsets, float4 *out,
                                                           unsigned capacity)
    {
        const long long first = (long long)blockIdx.x * kLatticeBlock + (long long)threadIdx.x * kLatticePts;
        const unsigned f = lattice_flags(a, first);
        unsigned total;
        unsigned idx = blockOffsets[blockIdx.x] + wg_exclusive_scan((unsigned)__popc(f), total);
        for (int j = 0; j < kLatticePts; ++j)
        {
            if (!((f >> j) & 1u))
                continue;
            if (idx < capacity)
            {
                V3 q;
                lattice_point(a, first + j, q);
                const V3 pl{q.x / a.invScale.x, q.y / a.invScale.y, q.z / a.invScale.z};
                const V3 pw = qrot(a.rot, pl) + v3(a.pos.x, a.pos.y, a.pos.z);
                out[idx] = make_float4(pw.x, pw.y, pw.z, a.mass);
            }
            ++idx;
        }
    }

    // Argument checks and launch arguments shared by the lattice entry points:
    // bounds fattened by 2 spacings (Bounds3::fattenFast, include/math/bounds3.h:85-94);
    // lattice dimensions ((max - min) / spacing / invScale).cast<int>().  blocks = 0
    // for an empty lattice (*numParticles is then 0).
    int lattice_setup(int geomType, const float params0[4], const CrHipGeometrySdfData *sdf, const float boundsMin[3],
                      const float boundsMax[3], float particleSpacing, const float position[3], const float rotation[4],
                      const float invScale[3], float particleMass, const float *positionMassOut, unsigned capacity,
                      unsigned *numParticles, LatticeArgs &a, int &blocks)
    {
        blocks = 0;
        if (!params0 || !boundsMin || !boundsMax || !position || !rotation || !invScale || !numParticles ||
            (capacity > 0 && !positionMassOut) || geomType < gBox || geomType > gArc || !(particleSpacing > 0.0f))
            return CRMPM_HIP_ERR_ARG;
        const bool table = geomType == gTriangleMesh || geomType == gTriangleMeshSlicer || geomType == gLines;
        if (table && (!sdf || !sdf->dimension || !sdf->lowerBoundCellSize || !sdf->gradientSignedDistance))
            return CRMPM_HIP_ERR_ARG;
        float lo[3], dimf[3];
        int dims[3];
        for (int k = 0; k < 3; ++k)
        {
            if (!(invScale[k] > 0.0f) || !std::isfinite(boundsMin[k]) || !std::isfinite(boundsMax[k]))
                return CRMPM_HIP_ERR_ARG;
            const float d = 2.0f * particleSpacing;
            lo[k] = boundsMin[k] - d;
            const float hi = boundsMax[k] + d;
            dimf[k] = (hi - lo[k]) / particleSpacing / invScale[k];
            if (!(dimf[k] < 2147483647.0f))
                return CRMPM_HIP_ERR_CAPACITY;
            dims[k] = std::max(0, (int)dimf[k]);
        }
        const long long total = (long long)dims[0] * dims[1] * dims[2];
        if (total >= (1ll << 31))
            return CRMPM_HIP_ERR_CAPACITY;
        *numParticles = 0;
        if (total == 0)
            return CRMPM_HIP_OK;
        if (table)
        {
            a.sv.sdfDims = reinterpret_cast<const int4 *>(sdf->dimension);
            a.sv.sdfLbcs = reinterpret_cast<const float4 *>(sdf->lowerBoundCellSize);
            a.sv.sdfData = reinterpret_cast<const float4 *>(sdf->gradientSignedDistance);
        }
        a.p0 = make_float4(params0[0], params0[1], params0[2], params0[3]);
        a.rot = make_float4(rotation[0], rotation[1], rotation[2], rotation[3]);
        a.lo = make_float3(lo[0], lo[1], lo[2]);
        a.invScale = make_float3(invScale[0], invScale[1], invScale[2]);
        a.pos = make_float3(position[0], position[1], position[2]);
        a.spacing = particleSpacing;
        a.mass = particleMass;
        a.gtype = geomType;
        a.ny = dims[1];
        a.nz = dims[2];
        a.total = total;
        blocks = (int)((total + kLatticeBlock - 1) / kLatticeBlock);
        return CRMPM_HIP_OK;
    }

    // Count, scan and write passes with the per-block counts in `counts`
    // (blocks + 1 entries, each written before it is read), then the total back
    // to the host.
    int lattice_run(const LatticeArgs &a, int blocks, unsigned *counts, float *positionMassOut, unsigned capacity,
                    unsigned *numParticles, hipStream_t st)
    {
        k_lattice_count<<<blocks, kWG, 0, st>>>(a, counts);
        k_lattice_scan<<<1, kWG, 0, st>>>(blocks, counts);
        k_lattice_write<<<blocks, kWG, 0, st>>>(a, counts, reinterpret_cast<float4 *>(positionMassOut), capacity);
        unsigned n = 0;
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&n, counts + blocks, sizeof(unsigned), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return CRMPM_HIP_ERR_LAUNCH;
        *numParticles = n;
        return CRMPM_HIP_OK;
    }

    // Scratch of crHipParticleLatticeNoAlloc: the per-block counts and their
    // total for lattices of up to kLatticeScratchBlocks blocks of kLatticeBlock
    // points (2^28 points), 512 KiB per GPU, held until the process ends.  In
    // host code the name stands for no device address: kernels and copies take
    // the per-device address from lattice_scratch().
    constexpr int kLatticeScratchBlocks = 1 << 17;
    static __device__ __align__(256) unsigned scratch[kLatticeScratchBlocks + 1];

    constexpr int kMaxScratchDevices = 64;
    std::mutex gScratchLock;
    void *gScratchAddr[kMaxScratchDevices] = {};

    // The current device's copy of `scratch`: looked up once per device (with
    // that device current) by hipGetSymbolAddress, then kept.
    int lattice_scratch(unsigned **out)
    {
        int dev = -1;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxScratchDevices)
            return CRMPM_HIP_ERR_LAUNCH;
        std::lock_guard<std::mutex> lock(gScratchLock);
        if (!gScratchAddr[dev])
        {
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
                return CRMPM_HIP_ERR_LAUNCH;
            gScratchAddr[dev] = p;
        }
        *out = static_cast<unsigned *>(gScratchAddr[dev]);
        return CRMPM_HIP_OK;
    }
} // namespace

extern "C" {

// ---------------------------- batch ----------------------------

int crHipBatchCreate(int kind, int numScenes, const CrHipSceneDesc *scenes, CrHipBatch **out)
{
    if (!out || numScenes <= 0 || !scenes)
        return CRMPM_HIP_ERR_ARG;
    if (kind != CRMPM_HIP_MLS && kind != CRMPM_HIP_PB && kind != CRMPM_HIP_STANDARD)
        return CRMPM_HIP_ERR_ARG;
    CrHipBatch *b = batch_new(kind, numScenes);
    if (!b)
        return CRMPM_HIP_ERR_ALLOC;
    long long blocks = 0;
    long long cap = 0;
    for (int s = 0; s < numScenes; ++s)
    {
        const CrHipSceneDesc &d = scenes[s];
        HostScene &h = b->scenes[s];
        if (d.numMaxParticles < 0 || d.particleOffset < 0 || !(d.cellSize > 0.0f) ||
            !grid_geom(d.cellSize, d.boundMin, d.boundMax, h.g))
        {
            batch_destroy(b);
            return CRMPM_HIP_ERR_ARG;
        }
        h.maxParticles = d.numMaxParticles;
        h.particleOffset = d.particleOffset;
        h.blockOffset = (int)blocks;
        blocks += h.g.numBlocks;
        cap = std::max(cap, (long long)d.particleOffset + d.numMaxParticles);
    }
    // the grid update's block list packs a block index in 24 bits
    if (blocks >= (1ll << 24) || cap >= (1ll << 31))
    {
        batch_destroy(b);
        return CRMPM_HIP_ERR_CAPACITY;
    }
    b->totalBlocks = blocks;
    b->capacity = (int)cap;
    *out = b;
    return CRMPM_HIP_OK;
}

int crHipBatchDestroy(CrHipBatch *b)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    return batch_destroy(b);
}

int crHipBatchSetStream(CrHipBatch *b, void *stream)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    b->stream = (hipStream_t)stream;
    drop_graph(b);
    return CRMPM_HIP_OK;
}

int crHipBatchSetSolverIterations(CrHipBatch *b, int n)
{
    if (!b || n < 0)
        return CRMPM_HIP_ERR_ARG;
    b->iterations = n;
    return CRMPM_HIP_OK;
}

int crHipBatchBindParticleData(CrHipBatch *b, float *positionMass, float *velocity, int size)
{
    if (!b || !positionMass || !velocity || size < b->capacity)
        return CRMPM_HIP_ERR_ARG;
    b->P.xm = reinterpret_cast<float4 *>(positionMass);
    b->P.v = reinterpret_cast<float4 *>(velocity);
    ++b->bindEpoch;
    return CRMPM_HIP_OK;
}

int crHipBatchBindParticleMaterials(CrHipBatch *b, float *params0, int *type, int size)
{
    if (!b || !params0 || !type || size < b->capacity)
        return CRMPM_HIP_ERR_ARG;
    b->P.params = reinterpret_cast<const float4 *>(params0);
    b->P.type = type;
    ++b->bindEpoch;
    return CRMPM_HIP_OK;
}

int crHipBatchBindActiveMask(CrHipBatch *b, unsigned char *activeMask)
{
    if (!b || !activeMask)
        return CRMPM_HIP_ERR_ARG;
    b->P.mask = activeMask;
    ++b->bindEpoch;
    return CRMPM_HIP_OK;
}

int crHipBatchInitialize(CrHipBatch *b)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    return batch_initialize(b);
}

int crHipBatchGetGridDims(const CrHipBatch *b, int scene, int dims[3])
{
    if (!valid_scene(b, scene) || !dims)
        return CRMPM_HIP_ERR_ARG;
    dims[0] = b->scenes[scene].g.nx;
    dims[1] = b->scenes[scene].g.ny;
    dims[2] = b->scenes[scene].g.nz;
    return CRMPM_HIP_OK;
}

int crHipBatchSetSceneGravity(CrHipBatch *b, int scene, float gx, float gy, float gz)
{
    if (!valid_scene(b, scene))
        return CRMPM_HIP_ERR_ARG;
    b->scenes[scene].gravity = make_float3(gx, gy, gz);
    b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipBatchSetSceneStepSize(CrHipBatch *b, int scene, float dt)
{
    if (!valid_scene(b, scene))
        return CRMPM_HIP_ERR_ARG;
    b->scenes[scene].dt = dt;
    b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipBatchSetSceneContactModel(CrHipBatch *b, int scene, int model)
{
    if (!valid_scene(b, scene) || (model != CRMPM_HIP_CONTACT_KINEMATIC && model != CRMPM_HIP_CONTACT_EFFECTIVE_MASS))
        return CRMPM_HIP_ERR_ARG;
    b->scenes[scene].contactModel = model;
    b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipBatchSetSceneNumActiveParticles(CrHipBatch *b, int scene, unsigned int n)
{
    if (!valid_scene(b, scene) || n > (unsigned)b->scenes[scene].maxParticles)
        return CRMPM_HIP_ERR_ARG;
    b->scenes[scene].numActive = n;
    b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipBatchSetSceneResting(CrHipBatch *b, int scene, int resting)
{
    if (!valid_scene(b, scene))
        return CRMPM_HIP_ERR_ARG;
    b->scenes[scene].resting = resting != 0;
    b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipBatchBindSceneShapes(CrHipBatch *b, int scene, int numShapes, int *shapeIds, const CrHipShapeData *shapes,
                              const CrHipGeometryData *geometries, const CrHipGeometrySdfData *sdf)
{
    if (!valid_scene(b, scene) || numShapes < 0)
        return CRMPM_HIP_ERR_ARG;
    HostScene &h = b->scenes[scene];
    h.numShapes = numShapes;
    h.shapeIds = shapeIds;
    if (shapes)
        h.shapes = *shapes;
    if (geometries)
        h.geoms = *geometries;
    if (sdf)
        h.sdf = *sdf;
    b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipBatchComputeInitialData(CrHipBatch *b, int scene, unsigned int n, const unsigned int *indices)
{
    return batch_compute_initial_data(b, scene, n, indices);
}

int crHipBatchStep(CrHipBatch *b)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    return batch_step(b);
}

int crHipBatchSetGraphMode(CrHipBatch *b, int enabled)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    b->graphMode = enabled != 0;
    if (!b->graphMode)
        drop_graph(b);
    return CRMPM_HIP_OK;
}

int crHipBatchAutotune(CrHipBatch *b, int stepsPerTrial, int *tileWaves, int *updateWaves, int *subSize, int *subMax)
{
    return batch_autotune(b, stepsPerTrial, tileWaves, updateWaves, subSize, subMax);
}

int crHipBatchSetTileSplit(CrHipBatch *b, int subSize, int subMax)
{
    if (!b || subSize < 64 || subSize % 64 != 0 || subMax < 1 || subMax > 8)
        return CRMPM_HIP_ERR_ARG;
    if (b->initialized && subMax > b->subStride)
        return CRMPM_HIP_ERR_CAPACITY;
    b->subSize = subSize;
    b->subMax = subMax;
    return CRMPM_HIP_OK;
}

int crHipBatchSetGatherSplit(CrHipBatch *b, int subSize, int subMax)
{
    if (!b || subSize < 64 || subSize % 64 != 0 || subMax < 0 || subMax > 64)
        return CRMPM_HIP_ERR_ARG;
    if (subMax == 0 && b->numScenes != 1) // per particle: batches of one scene
        return CRMPM_HIP_ERR_ARG;
    b->gSubSize = subSize;
    b->gSubMax = subMax;
    return CRMPM_HIP_OK;
}

int crHipBatchSetScatterWaves(CrHipBatch *b, int waves)
{
    if (!b || (waves != 0 && waves != 1 && waves != 2 && waves != 4))
        return CRMPM_HIP_ERR_ARG;
    b->scatterWaves = waves;
    return CRMPM_HIP_OK;
}

int crHipBatchGetScatterWaves(const CrHipBatch *b, int *waves)
{
    if (!b || !waves)
        return CRMPM_HIP_ERR_ARG;
    *waves = b->scatterWaves;
    return CRMPM_HIP_OK;
}

int crHipBatchGetGatherSplit(const CrHipBatch *b, int *subSize, int *subMax)
{
    if (!b || !subSize || !subMax)
        return CRMPM_HIP_ERR_ARG;
    *subSize = b->gSubSize;
    *subMax = b->gSubMax;
    return CRMPM_HIP_OK;
}

int crHipBatchSetLaunchSizes(CrHipBatch *b, int tileWaves, int updateWaves)
{
    if (!b || tileWaves < 0 || updateWaves < 0)
        return CRMPM_HIP_ERR_ARG;
    b->tileWaves = tileWaves;
    b->updateWaves = updateWaves;
    return CRMPM_HIP_OK;
}

int crHipBatchFetchResults(CrHipBatch *b)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    return hipStreamSynchronize(b->stream) == hipSuccess ? CRMPM_HIP_OK : CRMPM_HIP_ERR_LAUNCH;
}

int crHipBatchGetStateArrays(CrHipBatch *b, float **F, float **C, float **V0, int *stride)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    if (!b->initialized)
        return CRMPM_HIP_ERR_STATE;
    if (F)
        *F = b->F;
    if (C)
        *C = b->C;
    if (V0)
        *V0 = b->V0;
    if (stride)
        *stride = b->capacity;
    return CRMPM_HIP_OK;
}

int crHipBatchCopyGrid(CrHipBatch *b, int scene, float *dst)
{
    return batch_copy_grid(b, scene, dst);
}

int crHipBatchProfileBegin(CrHipBatch *b, int maxLaunches)
{
    return profile_begin(b, maxLaunches);
}

int crHipBatchProfileMask(CrHipBatch *b, unsigned int mask)
{
    if (!b)
        return CRMPM_HIP_ERR_ARG;
    if (!b->profEvents)
        return CRMPM_HIP_ERR_STATE;
    b->profMask = mask;
    return CRMPM_HIP_OK;
}

int crHipBatchProfileEnd(CrHipBatch *b, float *msPerCategory, int *launchesPerCategory, int numCategories)
{
    return profile_end(b, msPerCategory, launchesPerCategory, numCategories);
}

// ------------------------- single scene -------------------------

int crHipSolverCreate(int kind, int numMaxParticles, float cellSize, const float boundMin[3], const float boundMax[3],
                      CrHipSolver **out)
{
    if (!out || !boundMin || !boundMax)
        return CRMPM_HIP_ERR_ARG;
    CrHipSceneDesc d{};
    d.numMaxParticles = numMaxParticles;
    d.particleOffset = 0;
    d.cellSize = cellSize;
    for (int a = 0; a < 3; ++a)
    {
        d.boundMin[a] = boundMin[a];
        d.boundMax[a] = boundMax[a];
    }
    CrHipSolver *s = new (std::nothrow) CrHipSolver();
    if (!s)
        return CRMPM_HIP_ERR_ALLOC;
    const int rc = crHipBatchCreate(kind, 1, &d, &s->b);
    if (rc)
    {
        delete s;
        return rc;
    }
    *out = s;
    return CRMPM_HIP_OK;
}

int crHipSolverDestroy(CrHipSolver *s)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    const int rc = batch_destroy(s->b);
    delete s;
    return rc;
}

int crHipSolverGetGridDims(const CrHipSolver *s, int dims[3])
{
    return s ? crHipBatchGetGridDims(s->b, 0, dims) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetStream(CrHipSolver *s, void *stream)
{
    return s ? crHipBatchSetStream(s->b, stream) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetGravity(CrHipSolver *s, float gx, float gy, float gz)
{
    return s ? crHipBatchSetSceneGravity(s->b, 0, gx, gy, gz) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetIntegrationStepSize(CrHipSolver *s, float dt)
{
    return s ? crHipBatchSetSceneStepSize(s->b, 0, dt) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetSolverIterations(CrHipSolver *s, int n)
{
    return s ? crHipBatchSetSolverIterations(s->b, n) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetShapeContactModel(CrHipSolver *s, int model)
{
    return s ? crHipBatchSetSceneContactModel(s->b, 0, model) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverBindParticleData(CrHipSolver *s, float *positionMass, float *velocity, int size)
{
    return s ? crHipBatchBindParticleData(s->b, positionMass, velocity, size) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverBindParticleMaterials(CrHipSolver *s, float *params0, int *type, int size)
{
    return s ? crHipBatchBindParticleMaterials(s->b, params0, type, size) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverBindActiveMask(CrHipSolver *s, unsigned char *activeMask)
{
    return s ? crHipBatchBindActiveMask(s->b, activeMask) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetNumActiveParticles(CrHipSolver *s, unsigned int n)
{
    return s ? crHipBatchSetSceneNumActiveParticles(s->b, 0, n) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetNumShapes(CrHipSolver *s, int n)
{
    if (!s || n < 0)
        return CRMPM_HIP_ERR_ARG;
    s->b->scenes[0].numShapes = n;
    s->b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipSolverBindShapeIds(CrHipSolver *s, int *shapeIds)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    s->b->scenes[0].shapeIds = shapeIds;
    s->b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipSolverBindShapeData(CrHipSolver *s, const CrHipShapeData *d)
{
    if (!s || !d)
        return CRMPM_HIP_ERR_ARG;
    s->b->scenes[0].shapes = *d;
    s->b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipSolverBindGeometryData(CrHipSolver *s, const CrHipGeometryData *d)
{
    if (!s || !d)
        return CRMPM_HIP_ERR_ARG;
    s->b->scenes[0].geoms = *d;
    s->b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipSolverBindGeometrySdfData(CrHipSolver *s, const CrHipGeometrySdfData *d)
{
    if (!s || !d)
        return CRMPM_HIP_ERR_ARG;
    s->b->scenes[0].sdf = *d;
    s->b->scenesDirty = true;
    return CRMPM_HIP_OK;
}

int crHipSolverInitialize(CrHipSolver *s)
{
    return s ? batch_initialize(s->b) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverComputeInitialData(CrHipSolver *s, unsigned int n, const unsigned int *indices)
{
    return s ? batch_compute_initial_data(s->b, 0, n, indices) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverStep(CrHipSolver *s, float *dtOut)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    if (dtOut)
        *dtOut = s->b->scenes[0].dt;
    return batch_step(s->b);
}

int crHipSolverAutotune(CrHipSolver *s, int stepsPerTrial, int *tileWaves, int *updateWaves, int *subSize, int *subMax)
{
    return s ? batch_autotune(s->b, stepsPerTrial, tileWaves, updateWaves, subSize, subMax) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetLaunchConfig(CrHipSolver *s, int tileWaves, int updateWaves, int subSize, int subMax,
                               int gatherSubSize, int gatherSubMax)
{
    if (!s)
        return CRMPM_HIP_ERR_ARG;
    int rc = crHipBatchSetTileSplit(s->b, subSize, subMax);
    rc = rc ? rc : crHipBatchSetGatherSplit(s->b, gatherSubSize, gatherSubMax);
    return rc ? rc : crHipBatchSetLaunchSizes(s->b, tileWaves, updateWaves);
}

int crHipSolverSetGraphMode(CrHipSolver *s, int enabled)
{
    return s ? crHipBatchSetGraphMode(s->b, enabled) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverSetScatterWaves(CrHipSolver *s, int waves)
{
    return s ? crHipBatchSetScatterWaves(s->b, waves) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverGetScatterWaves(const CrHipSolver *s, int *waves)
{
    return s ? crHipBatchGetScatterWaves(s->b, waves) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverGetGatherSplit(const CrHipSolver *s, int *subSize, int *subMax)
{
    return s ? crHipBatchGetGatherSplit(s->b, subSize, subMax) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverBinParticles(CrHipSolver *s, unsigned int *keysOut, unsigned int *orderOut)
{
    if (!s || !keysOut || !orderOut)
        return CRMPM_HIP_ERR_ARG;
    CrHipBatch *b = s->b;
    if (!b->initialized || !b->P.xm || !b->P.mask)
        return CRMPM_HIP_ERR_STATE;
    int rc = upload_scenes(b);
    if (rc)
        return rc;
    const int n = b->totalActive;
    if (n == 0)
        return CRMPM_HIP_OK;
    const unsigned *order = nullptr, *keys = nullptr;
    rc = sort_particles(b, n, b->kind != CRMPM_HIP_STANDARD, &order, &keys);
    if (rc)
        return rc;
    if (hipMemcpyAsync(keysOut, keys, n * sizeof(unsigned), hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(orderOut, order, n * sizeof(unsigned), hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
        return CRMPM_HIP_ERR_LAUNCH;
    return CRMPM_HIP_OK;
}

int crHipSolverProfileBegin(CrHipSolver *s, int maxLaunches)
{
    return s ? profile_begin(s->b, maxLaunches) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverProfileMask(CrHipSolver *s, unsigned int mask)
{
    return s ? crHipBatchProfileMask(s->b, mask) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverProfileEnd(CrHipSolver *s, float *msPerCategory, int *launchesPerCategory, int numCategories)
{
    return s ? profile_end(s->b, msPerCategory, launchesPerCategory, numCategories) : CRMPM_HIP_ERR_STATE;
}

int crHipSolverFetchResults(CrHipSolver *s)
{
    return s ? crHipBatchFetchResults(s->b) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverGetStateArrays(CrHipSolver *s, float **F, float **C, float **V0, int *stride)
{
    return s ? crHipBatchGetStateArrays(s->b, F, C, V0, stride) : CRMPM_HIP_ERR_ARG;
}

int crHipSolverCopyGrid(CrHipSolver *s, float *dst)
{
    return s ? batch_copy_grid(s->b, 0, dst) : CRMPM_HIP_ERR_ARG;
}

} // extern "C"

// ----------------------------------------------------------------------
// Adjoint driver: checkpointed forward, segment-wise recomputation and
// reverse sweep (one scene, MLS-MPM, no rigid shapes).
// ----------------------------------------------------------------------
namespace
{
    struct StateBuf
    {
        float4 *x = nullptr, *v = nullptr;
        float *F = nullptr, *C = nullptr; // as SolverState
    };
} // namespace

struct CrHipAdjoint
{
    CrHipSolver *solver = nullptr;
    int steps = 0, every = 0, segments = 0;
    bool forwardDone = false;
    StateBuf *ckpt = nullptr;  // state at substeps 0, K, 2K, ... and the final state
    StateBuf *seg = nullptr;   // states of the substeps of one segment (K + 1)
    float4 **segGrid = nullptr;
    unsigned **segOrder = nullptr;
    int **segStart = nullptr, **segEnd = nullptr;
    unsigned **segActive = nullptr; // active node block list and its count ([totalBlocks] then the count)
    unsigned **segKeys = nullptr;   // sorted pblock keys
    float4 *gridAdj = nullptr;
    float4 *xTmp = nullptr;
    float4 *FTmp = nullptr;
    // PB: the tape (a segment's substeps, or the one substep being reversed;
    // see pb_reverse), its size and budget, and the adjoint state in tile order
    char *tape = nullptr; // the PB tape's bytes (see pb_tape_slot)
    size_t tapeBytes = 0;
    size_t tapeBudget = size_t(8) << 30;
    float4 *paX = nullptr, *paV = nullptr, *paC = nullptr, *paF = nullptr, *paFi = nullptr;
    // PB, two segment tapes (pb_reverse): the second set of segment states,
    // the reverse's own partial boxes and stream, and the events that order
    // the re-runs (batch stream) against the reverses (reverse stream)
    StateBuf *seg2 = nullptr;
    float4 *revTiles = nullptr;
    size_t revTilesCount = 0;
    hipStream_t revStream = nullptr;
    hipEvent_t ran[2] = {nullptr, nullptr}, rev[2] = {nullptr, nullptr}, start = nullptr;
};

namespace
{
    int state_alloc(StateBuf &s, size_t N)
    {
        int rc = dalloc(&s.x, N);
        rc |= dalloc(&s.v, N);
        rc |= dalloc(&s.F, 9 * N);
        rc |= dalloc(&s.C, 9 * N);
        return rc;
    }
    void state_free(StateBuf &s)
    {
        dfree(s.x);
        dfree(s.v);
        dfree(s.F);
        dfree(s.C);
    }
    // live particle state <-> buffer (whole arrays: x|m, v, F, C)
    int state_copy(CrHipBatch *b, StateBuf &s, bool save)
    {
        const size_t N = (size_t)std::max(b->capacity, 1);
        auto cp = [&](void *live, void *buf, size_t bytes) {
            return hipMemcpyAsync(save ? buf : live, save ? live : buf, bytes, hipMemcpyDeviceToDevice, b->stream) ==
                   hipSuccess;
        };
        const bool ok = cp(b->P.xm, s.x, N * sizeof(float4)) && cp(b->P.v, s.v, N * sizeof(float4)) &&
                        cp(b->F, s.F, 9 * N * sizeof(float)) && cp(b->C, s.C, 9 * N * sizeof(float));
        return ok ? CRMPM_HIP_OK : CRMPM_HIP_ERR_LAUNCH;
    }
    StateRef state_ref(const StateBuf &s, int stride) { return StateRef{s.x, s.v, s.F, s.C, stride}; }

    // What the reverse kernels of one PB substep read of its forward run:
    // the iterations' grids, the sort products and the staged particle
    // arrays (x|m, F, (alpha, ratio), type in tile order).
    struct PbSubstepTape
    {
        const float4 *U; // K grids of the batch's node count
        const unsigned *order, *keys;
        const int *segStart, *segEnd;
        const unsigned *activeList;
        const int *activeCount;
        PbStage st;
    };

    // The reverse kernels of one PB substep; s0 and s1 are its start and
    // end states.
    // st: the stream of the reverse kernels; tiles: the partial boxes of the
    // velocity-adjoint scatter (null: the batch's own)
    int pb_reverse_substep(CrHipAdjoint *a, const AdjRef &ad, const PbSubstepTape &tp, StateRef s0, StateRef s1,
                           hipStream_t st, float4 *tiles)
    {
        CrHipBatch *b = a->solver->b;
        const size_t nodes = (size_t)b->totalBlocks * 64;
        const int K = b->iterations, n = b->totalActive;
        if (n == 0)
            return CRMPM_HIP_OK;
        const PbAdjRef pa{a->paX, a->paV, a->paC, a->paF, a->paFi};
        BatchArgs args = batch_args(b, tp.order);
        args.gatherOrder = tp.order;
        args.segStart = tp.segStart;
        args.segEnd = tp.segEnd;
        args.activeList = tp.activeList;
        args.activeCount = tp.activeCount;
        // the velocity-adjoint scatter and the grid adjoint take whole pblocks
        // (the grid adjoint then sums at most 8 partials per node)
        BatchArgs whole = args;
        whole.subSize = 1 << 30;
        whole.subMax = 1;
        if (tiles)
            whole.tilebuf = tiles;
        BatchArgs wholeSmall = whole, wholeLarge = whole; // adaptive wavefronts per tile
        wholeSmall.tileClass = 1;
        wholeLarge.tileClass = 2;
        const int wholeItems = std::max(1, (int)b->totalBlocks); // one workgroup per pblock
        const int blockWgs = grid1d(b->totalBlocks, kWaves);
        const int denseWgs = grid1d(n, kWG);
        const unsigned *keys = tp.keys;
        k_pbadj_integrate<<<grid1d(n, 256), 256, 0, st>>>(args, tp.st, n, keys, s1, ad, pa);
        if (K > 0)
            k_pbadj_g2p_x<<<denseWgs, kWG, 0, st>>>(args, tp.st, n, keys, tp.U + (size_t)(K - 1) * nodes, pa);
        for (int k = K; k >= 1; --k)
        {
            const float4 *U = tp.U + (size_t)(k - 1) * nodes;
            const float4 *Uprev = k > 1 ? U - nodes : nullptr;
            if (b->scatterWaves == 4)
                k_pbadj_uscatter<4><<<wholeItems, 256, 0, st>>>(whole, tp.st, pa);
            else if (b->scatterWaves == 2)
                k_pbadj_uscatter<2><<<wholeItems, 128, 0, st>>>(whole, tp.st, pa);
            else if (b->scatterWaves == 1)
                k_pbadj_uscatter<1><<<wholeItems, 64, 0, st>>>(whole, tp.st, pa);
            else
            {
                k_pbadj_uscatter<1><<<wholeItems, 64, 0, st>>>(wholeSmall, tp.st, pa);
                k_pbadj_uscatter<4><<<wholeItems, 256, 0, st>>>(wholeLarge, tp.st, pa);
            }
            k_adj_grid<1><<<blockWgs, kWG, 0, st>>>(whole, U, a->gridAdj);
            k_pbadj_particle<<<denseWgs, kWG, 0, st>>>(args, tp.st, n, keys, s0, a->gridAdj, Uprev, pa, ad);
        }
        k_pbadj_finalize<<<grid1d(n, 256), 256, 0, st>>>(args, tp.st, n, keys, pa, ad);
        return check_launch();
    }

    // Bytes per substep of a whole segment's tape: the iterations' grids,
    // the sort products (order, keys, segments, active list and count) and
    // the staged x|m, F, (alpha, ratio) and type.
    size_t pb_substep_tape_bytes(const CrHipBatch *b, int K)
    {
        const size_t N = (size_t)std::max(b->capacity, 1), B = (size_t)b->totalBlocks;
        return (size_t)K * B * 64 * sizeof(float4) + N * (2 * sizeof(unsigned) + sizeof(float4) + 9 * sizeof(float) +
                                                          sizeof(float2) + sizeof(int)) +
               B * (2 * sizeof(int) + sizeof(unsigned)) + 2 * sizeof(int);
    }

    int pb_ensure_tape(CrHipAdjoint *a, int K, int substeps)
    {
        CrHipBatch *b = a->solver->b;
        const size_t need = pb_substep_tape_bytes(b, K) * (size_t)substeps;
        if (a->tape && a->tapeBytes >= need)
            return CRMPM_HIP_OK;
        (void)hipStreamSynchronize(b->stream);
        dfree(a->tape);
        a->tapeBytes = 0;
        if (dalloc(&a->tape, need))
            return CRMPM_HIP_ERR_ALLOC;
        a->tapeBytes = need;
        return CRMPM_HIP_OK;
    }

    // The tape of substep j of a segment, carved from a->tape.
    PbSubstepTape pb_tape_slot(CrHipAdjoint *a, int K, int j)
    {
        CrHipBatch *b = a->solver->b;
        const size_t N = (size_t)std::max(b->capacity, 1), B = (size_t)b->totalBlocks;
        char *m = a->tape + pb_substep_tape_bytes(b, K) * (size_t)j;
        PbSubstepTape t{};
        auto take = [&](size_t bytes) {
            char *r = m;
            m += bytes;
            return r;
        };
        t.U = reinterpret_cast<const float4 *>(take((size_t)K * B * 64 * sizeof(float4)));
        t.st.xs = reinterpret_cast<float4 *>(take(N * sizeof(float4)));
        t.st.Fs = reinterpret_cast<float *>(take(9 * N * sizeof(float)));
        t.st.ps = reinterpret_cast<float2 *>(take(N * sizeof(float2)));
        t.order = reinterpret_cast<const unsigned *>(take(N * sizeof(unsigned)));
        t.keys = reinterpret_cast<const unsigned *>(take(N * sizeof(unsigned)));
        t.st.ts = reinterpret_cast<int *>(take(N * sizeof(int)));
        t.segStart = reinterpret_cast<const int *>(take(B * sizeof(int)));
        t.segEnd = reinterpret_cast<const int *>(take(B * sizeof(int)));
        t.activeList = reinterpret_cast<const unsigned *>(take(B * sizeof(unsigned)));
        t.activeCount = reinterpret_cast<const int *>(take(2 * sizeof(int)));
        t.st.vs = nullptr;
        t.st.Cs = nullptr;
        t.st.stride = (int)N;
        return t;
    }

    // Copy of what the reverse kernels read of the substep just run (its
    // grids were taped by batch_step_taped) into a segment tape slot.
    int pb_save_substep(CrHipBatch *b, const PbSubstepTape &t)
    {
        const size_t N = (size_t)std::max(b->capacity, 1), B = (size_t)b->totalBlocks;
        const int n = b->totalActive;
        hipStream_t st = b->stream;
        auto cp = [&](const void *dst, const void *src, size_t bytes) {
            return hipMemcpyAsync(const_cast<void *>(dst), src, bytes, hipMemcpyDeviceToDevice, st) == hipSuccess;
        };
        const bool ok = n == 0 ||
                        (cp(t.st.xs, b->st.xs, N * sizeof(float4)) && cp(t.st.Fs, b->st.Fs, 9 * N * sizeof(float)) &&
                         cp(t.st.ps, b->st.ps, N * sizeof(float2)) && cp(t.st.ts, b->st.ts, N * sizeof(int)) &&
                         cp(t.order, b->lastOrder, N * sizeof(unsigned)) && cp(t.keys, b->lastKeys, N * sizeof(unsigned)) &&
                         cp(t.segStart, b->segStart, B * sizeof(int)) && cp(t.segEnd, b->segEnd, B * sizeof(int)) &&
                         cp(t.activeList, b->activeList, B * sizeof(unsigned)) &&
                         cp(t.activeCount, b->activeCount, 2 * sizeof(int)));
        return ok ? CRMPM_HIP_OK : CRMPM_HIP_ERR_LAUNCH;
    }

    // Reverse sweep of a PB rollout.  Per segment (checkpoint interval):
    // when a whole segment's tape fits the adjoint's tape budget, the
    // segment is re-run once from its checkpoint with every substep's
    // iteration grids, sort products and staged particle arrays kept, and
    // then reversed; otherwise the re-run keeps only every substep's start
    // state and each substep is re-run once more, last first, with its own
    // tape before its reverse.  Both give bitwise the same gradients.
    // What two segment tapes need beside the tape: a second set of segment
    // states, partial boxes of the reverse's own, a stream (default
    // priority) and the events ordering the two streams.
    int pb_ensure_overlap(CrHipAdjoint *a)
    {
        CrHipBatch *b = a->solver->b;
        // the partial boxes follow the batch's layout, which a later tile
        // split may widen
        const size_t tiles = (size_t)b->totalBlocks * b->subStride * kTileNodes;
        if (a->revTilesCount < tiles)
        {
            if (a->revStream)
                (void)hipStreamSynchronize(a->revStream);
            dfree(a->revTiles);
            a->revTilesCount = 0;
            if (dalloc(&a->revTiles, tiles))
                return CRMPM_HIP_ERR_ALLOC;
            a->revTilesCount = tiles;
        }
        if (a->revStream)
            return CRMPM_HIP_OK;
        const size_t N = (size_t)std::max(b->capacity, 1);
        if (!a->seg2)
        {
            a->seg2 = new (std::nothrow) StateBuf[a->every + 1]();
            if (!a->seg2)
                return CRMPM_HIP_ERR_ALLOC;
            int rc = 0;
            for (int i = 0; i < a->every + 1; ++i)
                rc |= state_alloc(a->seg2[i], N);
            if (rc)
                return CRMPM_HIP_ERR_ALLOC; // freed with the adjoint
        }
        for (hipEvent_t *e : {&a->ran[0], &a->ran[1], &a->rev[0], &a->rev[1], &a->start})
            if (!*e && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
            {
                *e = nullptr;
                return CRMPM_HIP_ERR_ALLOC;
            }
        if (hipStreamCreateWithFlags(&a->revStream, hipStreamNonBlocking) != hipSuccess)
        {
            a->revStream = nullptr;
            return CRMPM_HIP_ERR_ALLOC;
        }
        return CRMPM_HIP_OK;
    }

    // MLS: the bytes of one set of per-segment buffers (grids, sort products,
    // segments and active lists of a segment's substeps).
    size_t mls_set_bytes(const CrHipAdjoint *a)
    {
        const CrHipBatch *b = a->solver->b;
        const size_t N = (size_t)std::max(b->capacity, 1), B = (size_t)b->totalBlocks;
        return (size_t)a->every * (B * 64 * sizeof(float4) + 2 * N * sizeof(unsigned) + (3 * B + 1) * sizeof(int));
    }

    int pb_ensure_overlap(CrHipAdjoint *a);

    // MLS: the second set of per-segment buffers, beside what
    // pb_ensure_overlap provides (second segment states, partial boxes,
    // stream, events).
    int mls_ensure_overlap(CrHipAdjoint *a)
    {
        CrHipBatch *b = a->solver->b;
        const size_t N = (size_t)std::max(b->capacity, 1), B = (size_t)b->totalBlocks;
        int rc = 0;
        for (int q = a->every; q < 2 * a->every; ++q)
        {
            if (!a->segGrid[q])
                rc |= dalloc(&a->segGrid[q], B * 64);
            if (!a->segOrder[q])
                rc |= dalloc(&a->segOrder[q], N);
            if (!a->segStart[q])
                rc |= dalloc(&a->segStart[q], B);
            if (!a->segEnd[q])
                rc |= dalloc(&a->segEnd[q], B);
            if (!a->segActive[q])
                rc |= dalloc(&a->segActive[q], B + 1);
            if (!a->segKeys[q])
                rc |= dalloc(&a->segKeys[q], N);
        }
        return rc ? CRMPM_HIP_ERR_ALLOC : pb_ensure_overlap(a);
    }

    // The overlapped backward schedule: segment s is reversed on the
    // adjoint's stream while segment s - 1 is re-run on the batch's stream,
    // the two sets of per-segment buffers alternating.  rerun(sgm, set)
    // enqueues a segment's re-run on the batch's stream; reverse(sgm, set,
    // stream) its reverse.  Events order each re-run after the reverse that
    // last read its set and each reverse after its re-run; the batch's
    // stream goes on after the last reverse.
    template <class Rerun, class Reverse>
    int overlapped_reverse(CrHipAdjoint *a, Rerun &&rerun, Reverse &&reverse)
    {
        CrHipBatch *b = a->solver->b;
        auto ok = [](hipError_t e) { return e == hipSuccess; };
        hipStream_t rs = a->revStream;
        // the reverse stream starts after the work enqueued so far
        if (!ok(hipEventRecord(a->start, b->stream)) || !ok(hipStreamWaitEvent(rs, a->start, 0)))
            return CRMPM_HIP_ERR_LAUNCH;
        const int last = a->segments - 1;
        int rc = rerun(last, last & 1);
        if (!rc && !ok(hipEventRecord(a->ran[last & 1], b->stream)))
            rc = CRMPM_HIP_ERR_LAUNCH;
        for (int sgm = last; sgm >= 0 && !rc; --sgm)
        {
            const int set = sgm & 1, other = set ^ 1;
            if (!ok(hipStreamWaitEvent(rs, a->ran[set], 0)))
                rc = CRMPM_HIP_ERR_LAUNCH;
            if (!rc && sgm > 0)
            {
                // set `other` was read by the reverse of segment sgm + 1
                if (sgm < last && !ok(hipStreamWaitEvent(b->stream, a->rev[other], 0)))
                    rc = CRMPM_HIP_ERR_LAUNCH;
                if (!rc)
                    rc = rerun(sgm - 1, other);
                if (!rc && !ok(hipEventRecord(a->ran[other], b->stream)))
                    rc = CRMPM_HIP_ERR_LAUNCH;
            }
            if (!rc)
                rc = reverse(sgm, set, rs);
            if (!rc && !ok(hipEventRecord(a->rev[set], rs)))
                rc = CRMPM_HIP_ERR_LAUNCH;
        }
        // the batch stream goes on after the last reverse (or, on an error,
        // after whatever the reverse stream holds)
        if (ok(hipEventRecord(a->start, rs)))
            (void)hipStreamWaitEvent(b->stream, a->start, 0);
        return rc;
    }

    // Reverse sweep of a PB rollout, three ways by the tape budget:
    //  * two segment tapes fit: segment s is reversed on the adjoint's stream
    //    while segment s - 1 is re-run (taped) on the batch's stream, the
    //    two sets of tapes and segment states alternating.  The reverse
    //    kernels hold about one wavefront per SIMD at the sizes of a
    //    differentiable scene, so the re-run mostly fills idle SIMDs;
    //  * one fits: each segment is re-run, then reversed;
    //  * none: each substep is re-run once more before it is reversed.
    // The gradients are bitwise the same in all three (each reversed
    // substep reads the same taped values).
    int pb_reverse(CrHipAdjoint *a, const AdjRef &ad)
    {
        CrHipBatch *b = a->solver->b;
        const int K = b->iterations;
        const int stride = b->capacity;
        const size_t segBytes = pb_substep_tape_bytes(b, K) * (size_t)a->every;
        const bool overlap = a->segments > 1 && 2 * segBytes <= a->tapeBudget;
        const bool whole = overlap || segBytes <= a->tapeBudget;
        int rc = pb_ensure_tape(a, K, whole ? a->every * (overlap ? 2 : 1) : 1);
        if (!rc && overlap)
            rc = pb_ensure_overlap(a);
        if (rc)
            return rc;
        const StateRef live{b->P.xm, b->P.v, b->F, b->C, stride};
        auto states = [&](int set) { return set ? a->seg2 : a->seg; };
        // re-run segment sgm from its checkpoint into tape set `set`
        auto rerun = [&](int sgm, int set) {
            const int t0 = sgm * a->every, t1 = std::min(t0 + a->every, a->steps);
            StateBuf *S = states(set);
            int r = state_copy(b, a->ckpt[sgm], false);
            for (int t = t0; t < t1 && !r; ++t)
            {
                const int j = t - t0;
                r = state_copy(b, S[j], true);
                if (r)
                    break;
                const PbSubstepTape tp = pb_tape_slot(a, K, set * a->every + j);
                r = batch_step_taped(b, const_cast<float4 *>(tp.U));
                if (!r)
                    r = pb_save_substep(b, tp);
            }
            return r ? r : state_copy(b, S[t1 - t0], true);
        };
        auto reverse = [&](int sgm, int set, hipStream_t st, float4 *tiles) {
            const int t0 = sgm * a->every, t1 = std::min(t0 + a->every, a->steps);
            StateBuf *S = states(set);
            int r = CRMPM_HIP_OK;
            for (int t = t1 - 1; t >= t0 && !r; --t)
            {
                const int j = t - t0;
                r = pb_reverse_substep(a, ad, pb_tape_slot(a, K, set * a->every + j), state_ref(S[j], stride),
                                       state_ref(S[j + 1], stride), st, tiles);
            }
            return r;
        };
        if (overlap)
            return overlapped_reverse(a, rerun, [&](int sgm, int set, hipStream_t rs) {
                return reverse(sgm, set, rs, a->revTiles);
            });
        for (int sgm = a->segments - 1; sgm >= 0 && !rc; --sgm)
        {
            if (whole)
            {
                rc = rerun(sgm, 0);
                if (!rc)
                    rc = reverse(sgm, 0, b->stream, nullptr);
                continue;
            }
            const int t0 = sgm * a->every, t1 = std::min(t0 + a->every, a->steps);
            rc = state_copy(b, a->ckpt[sgm], false);
            for (int t = t0; t < t1 && !rc; ++t)
            {
                rc = state_copy(b, a->seg[t - t0], true);
                if (!rc && t + 1 < t1)
                    rc = batch_step(b);
            }
            for (int t = t1 - 1; t >= t0 && !rc; --t)
            {
                const int j = t - t0;
                rc = state_copy(b, a->seg[j], false);
                PbSubstepTape tp{};
                tp.U = reinterpret_cast<const float4 *>(a->tape);
                if (!rc)
                    rc = batch_step_taped(b, reinterpret_cast<float4 *>(a->tape));
                if (rc)
                    break;
                tp.order = b->lastOrder;
                tp.keys = b->lastKeys;
                tp.segStart = b->segStart;
                tp.segEnd = b->segEnd;
                tp.activeList = b->activeList;
                tp.activeCount = b->activeCount;
                tp.st = b->st;
                rc = pb_reverse_substep(a, ad, tp, state_ref(a->seg[j], stride), live, b->stream, nullptr);
            }
        }
        return rc;
    }

    int adjoint_destroy(CrHipAdjoint *a)
    {
        if (a->solver && a->solver->b->initialized)
            (void)hipStreamSynchronize(a->solver->b->stream);
        for (int i = 0; a->ckpt && i < a->segments + 1; ++i)
            state_free(a->ckpt[i]);
        for (int i = 0; a->seg && i < a->every + 1; ++i)
            state_free(a->seg[i]);
        for (int i = 0; i < 2 * a->every; ++i)
        {
            if (a->segGrid)
                dfree(a->segGrid[i]);
            if (a->segOrder)
                dfree(a->segOrder[i]);
            if (a->segStart)
                dfree(a->segStart[i]);
            if (a->segEnd)
                dfree(a->segEnd[i]);
            if (a->segActive)
                dfree(a->segActive[i]);
            if (a->segKeys)
                dfree(a->segKeys[i]);
        }
        delete[] a->ckpt;
        delete[] a->seg;
        delete[] a->segGrid;
        delete[] a->segOrder;
        delete[] a->segStart;
        delete[] a->segEnd;
        delete[] a->segActive;
        delete[] a->segKeys;
        dfree(a->gridAdj);
        dfree(a->xTmp);
        dfree(a->FTmp);
        dfree(a->tape);
        dfree(a->paX);
        dfree(a->paV);
        dfree(a->paC);
        dfree(a->paF);
        dfree(a->paFi);
        if (a->revStream)
            (void)hipStreamSynchronize(a->revStream);
        for (int i = 0; a->seg2 && i < a->every + 1; ++i)
            state_free(a->seg2[i]);
        delete[] a->seg2;
        dfree(a->revTiles);
        for (hipEvent_t e : {a->ran[0], a->ran[1], a->rev[0], a->rev[1], a->start})
            if (e)
                (void)hipEventDestroy(e);
        if (a->revStream)
            (void)hipStreamDestroy(a->revStream);
        delete a;
        return CRMPM_HIP_OK;
    }
} // namespace

extern "C" {

int crHipAdjointCreate(CrHipSolver *s, int numSteps, int checkpointEvery, CrHipAdjoint **out)
{
    if (!s || !out || numSteps <= 0 || checkpointEvery <= 0)
        return CRMPM_HIP_ERR_ARG;
    CrHipBatch *b = s->b;
    if ((b->kind != CRMPM_HIP_MLS && b->kind != CRMPM_HIP_PB) || !b->initialized)
        return CRMPM_HIP_ERR_STATE;
    const bool mls = b->kind == CRMPM_HIP_MLS;
    CrHipAdjoint *a = new (std::nothrow) CrHipAdjoint();
    if (!a)
        return CRMPM_HIP_ERR_ALLOC;
    a->solver = s;
    a->steps = numSteps;
    a->every = std::min(checkpointEvery, numSteps);
    a->segments = (numSteps + a->every - 1) / a->every;
    const size_t N = (size_t)std::max(b->capacity, 1);
    const size_t nodes = (size_t)b->totalBlocks * 64;
    a->ckpt = new (std::nothrow) StateBuf[a->segments + 1];
    a->seg = new (std::nothrow) StateBuf[a->every + 1];
    a->segGrid = new (std::nothrow) float4 *[2 * a->every](); // a second set for overlapped reverses
    a->segOrder = new (std::nothrow) unsigned *[2 * a->every]();
    a->segStart = new (std::nothrow) int *[2 * a->every]();
    a->segEnd = new (std::nothrow) int *[2 * a->every]();
    a->segActive = new (std::nothrow) unsigned *[2 * a->every]();
    a->segKeys = new (std::nothrow) unsigned *[2 * a->every]();
    int rc = (a->ckpt && a->seg && a->segGrid && a->segOrder && a->segStart && a->segEnd && a->segActive && a->segKeys)
                 ? 0
                 : CRMPM_HIP_ERR_ALLOC;
    for (int i = 0; !rc && i < a->segments + 1; ++i)
        rc |= state_alloc(a->ckpt[i], N);
    for (int i = 0; !rc && i < a->every + 1; ++i)
        rc |= state_alloc(a->seg[i], N);
    // the MLS reverse keeps each substep's grid and tile order from the
    // segment's re-run; the PB reverse re-runs each substep with its tape
    for (int i = 0; !rc && mls && i < a->every; ++i)
    {
        rc |= dalloc(&a->segGrid[i], nodes);
        rc |= dalloc(&a->segOrder[i], N);
        rc |= dalloc(&a->segStart[i], (size_t)b->totalBlocks);
        rc |= dalloc(&a->segEnd[i], (size_t)b->totalBlocks);
        rc |= dalloc(&a->segActive[i], (size_t)b->totalBlocks + 1);
        rc |= dalloc(&a->segKeys[i], N);
    }
    if (!rc)
    {
        rc |= dalloc(&a->gridAdj, nodes);
        rc |= dalloc(&a->xTmp, N);
        rc |= dalloc(&a->FTmp, 3 * N);
    }
    // PB: the reverse state in tile order; MLS: paV and paC hold the G2P
    // adjoint's scatter inputs
    if (!rc)
    {
        rc |= dalloc(&a->paV, N);
        rc |= dalloc(&a->paC, 3 * N);
    }
    if (!rc && !mls)
    {
        rc |= dalloc(&a->paX, N);
        rc |= dalloc(&a->paF, 3 * N);
        rc |= dalloc(&a->paFi, 3 * N);
    }
    if (rc)
    {
        adjoint_destroy(a);
        return CRMPM_HIP_ERR_ALLOC;
    }
    *out = a;
    return CRMPM_HIP_OK;
}

int crHipAdjointSetTapeBudget(CrHipAdjoint *a, unsigned long long bytes)
{
    if (!a)
        return CRMPM_HIP_ERR_ARG;
    a->tapeBudget = (size_t)bytes;
    return CRMPM_HIP_OK;
}

int crHipAdjointGetSegmentTapeBytes(const CrHipAdjoint *a, unsigned long long *bytes)
{
    if (!a || !bytes)
        return CRMPM_HIP_ERR_ARG;
    const CrHipBatch *b = a->solver->b;
    *bytes = b->kind == CRMPM_HIP_PB ? (unsigned long long)(pb_substep_tape_bytes(b, b->iterations) * (size_t)a->every)
                                     : (unsigned long long)mls_set_bytes(a);
    return CRMPM_HIP_OK;
}

int crHipAdjointDestroy(CrHipAdjoint *a)
{
    if (!a)
        return CRMPM_HIP_ERR_ARG;
    return adjoint_destroy(a);
}

int crHipAdjointForward(CrHipAdjoint *a)
{
    if (!a)
        return CRMPM_HIP_ERR_ARG;
    CrHipBatch *b = a->solver->b;
    int rc = validate_for_step(b);
    if (rc)
        return rc;
    if (b->scenes[0].numShapes > 0)
        return CRMPM_HIP_ERR_STATE;
    for (int t = 0; t < a->steps && !rc; ++t)
    {
        if (t % a->every == 0)
            rc = state_copy(b, a->ckpt[t / a->every], true);
        if (!rc)
            rc = batch_step(b);
    }
    if (!rc)
        rc = state_copy(b, a->ckpt[a->segments], true);
    a->forwardDone = rc == CRMPM_HIP_OK;
    return rc;
}

int crHipAdjointBackward(CrHipAdjoint *a, const float *dLdx, const float *dLdv, double *dLdLambda, double *dLdMu,
                         float *dLdx0, float *dLdv0)
{
    if (!a || !dLdx || !dLdLambda || !dLdMu || !dLdv0)
        return CRMPM_HIP_ERR_ARG;
    if (!a->forwardDone)
        return CRMPM_HIP_ERR_STATE;
    CrHipBatch *b = a->solver->b;
    if (b->profEvents)
        return CRMPM_HIP_ERR_STATE;
    const size_t N = (size_t)std::max(b->capacity, 1);
    hipStream_t st = b->stream;
    // Adjoint of the state: x and v in the caller's output arrays (x in a
    // scratch array when dL/dx0 is not wanted), F and C in scratch arrays.
    // The final state's adjoint is (dL/dx, dL/dv, 0, 0).
    float4 *xOwn = nullptr;
    float4 *Fb = nullptr, *Cb = nullptr;
    int rc = CRMPM_HIP_OK;
    if (dalloc(&Fb, 3 * N) || dalloc(&Cb, 3 * N) || (!dLdx0 && dalloc(&xOwn, N)))
        rc = CRMPM_HIP_ERR_ALLOC;
    float4 *xb = dLdx0 ? reinterpret_cast<float4 *>(dLdx0) : xOwn;
    float4 *vb = reinterpret_cast<float4 *>(dLdv0);
    if (!rc && (hipMemcpyAsync(xb, dLdx, N * sizeof(float4), hipMemcpyDeviceToDevice, st) != hipSuccess ||
                (dLdv ? hipMemcpyAsync(vb, dLdv, N * sizeof(float4), hipMemcpyDeviceToDevice, st)
                      : hipMemsetAsync(vb, 0, N * sizeof(float4), st)) != hipSuccess ||
                hipMemsetAsync(Fb, 0, 3 * N * sizeof(float4), st) != hipSuccess ||
                hipMemsetAsync(Cb, 0, 3 * N * sizeof(float4), st) != hipSuccess ||
                hipMemsetAsync(dLdLambda, 0, N * sizeof(double), st) != hipSuccess ||
                hipMemsetAsync(dLdMu, 0, N * sizeof(double), st) != hipSuccess))
        rc = CRMPM_HIP_ERR_LAUNCH;
    const AdjRef ad{xb, vb, Fb, Cb, a->xTmp, a->FTmp, dLdLambda, dLdMu, b->capacity};
    const int stride = b->capacity;
    if (!rc && b->kind == CRMPM_HIP_PB)
        rc = pb_reverse(a, ad);
    // MLS: re-run a segment from its checkpoint into buffer set `set`,
    // keeping every substep's start state, updated grid (its active blocks,
    // written by the grid update) and tiles
    auto states = [&](int set) { return set ? a->seg2 : a->seg; };
    auto mrerun = [&](int sgm, int set) {
        const int t0 = sgm * a->every, t1 = std::min(t0 + a->every, a->steps);
        StateBuf *S = states(set);
        int r = state_copy(b, a->ckpt[sgm], false);
        for (int t = t0; t < t1 && !r; ++t)
        {
            const int j = t - t0, q = set * a->every + j;
            r = state_copy(b, S[j], true);
            if (!r)
                r = batch_step_taped(b, a->segGrid[q]);
            if (!r &&
                (hipMemcpyAsync(a->segOrder[q], b->lastOrder, N * sizeof(unsigned), hipMemcpyDeviceToDevice, st) != hipSuccess ||
                 hipMemcpyAsync(a->segKeys[q], b->lastKeys, N * sizeof(unsigned), hipMemcpyDeviceToDevice, st) != hipSuccess ||
                 hipMemcpyAsync(a->segStart[q], b->segStart, (size_t)b->totalBlocks * sizeof(int), hipMemcpyDeviceToDevice,
                                st) != hipSuccess ||
                 hipMemcpyAsync(a->segEnd[q], b->segEnd, (size_t)b->totalBlocks * sizeof(int), hipMemcpyDeviceToDevice,
                                st) != hipSuccess ||
                 hipMemcpyAsync(a->segActive[q], b->activeList, (size_t)b->totalBlocks * sizeof(unsigned),
                                hipMemcpyDeviceToDevice, st) != hipSuccess ||
                 hipMemcpyAsync(a->segActive[q] + b->totalBlocks, b->activeCount, sizeof(int), hipMemcpyDeviceToDevice,
                                st) != hipSuccess))
                r = CRMPM_HIP_ERR_LAUNCH;
        }
        return r ? r : state_copy(b, S[t1 - t0], true);
    };
    // the reverse sweep over a segment, on stream rs; tiles: the velocity
    // adjoint's partial boxes (null: the batch's own)
    auto mreverse = [&](int sgm, int set, hipStream_t rs, float4 *tiles) {
        const int t0 = sgm * a->every, t1 = std::min(t0 + a->every, a->steps);
        StateBuf *S = states(set);
        const int n = b->totalActive;
        int r = CRMPM_HIP_OK;
        for (int t = t1 - 1; t >= t0 && !r && n > 0; --t)
        {
            const int j = t - t0, q = set * a->every + j;
            BatchArgs args = batch_args(b, a->segOrder[q]);
            args.gatherOrder = args.order;
            args.segStart = a->segStart[q];
            args.segEnd = a->segEnd[q];
            args.activeList = a->segActive[q];
            args.activeCount = reinterpret_cast<const int *>(a->segActive[q] + b->totalBlocks);
            // the velocity adjoint's scatter and the grid adjoint take whole
            // pblocks (at most 8 partials per node); the per-particle kernels
            // run 64 sorted particles per wavefront
            BatchArgs whole = args;
            whole.subSize = 1 << 30;
            whole.subMax = 1;
            if (tiles)
                whole.tilebuf = tiles;
            const int wholeItems = std::max(1, (int)b->totalBlocks); // one workgroup per pblock
            const int blockWgs = grid1d(b->totalBlocks, kWaves);
            const int denseWgs = grid1d(n, kWG);
            const MlsScatterIn si{a->paV, a->paC};
            const StateRef s0 = state_ref(S[j], stride);
            k_adj_g2p<<<denseWgs, kWG, 0, rs>>>(args, n, a->segKeys[q], s0, state_ref(S[j + 1], stride), a->segGrid[q],
                                                ad, si);
            if (b->scatterWaves == 4)
                k_adj_uscatter<4><<<wholeItems, 256, 0, rs>>>(whole, s0, si);
            else if (b->scatterWaves == 2)
                k_adj_uscatter<2><<<wholeItems, 128, 0, rs>>>(whole, s0, si);
            else if (b->scatterWaves == 1)
                k_adj_uscatter<1><<<wholeItems, 64, 0, rs>>>(whole, s0, si);
            else
            {
                BatchArgs wholeSmall = whole, wholeLarge = whole;
                wholeSmall.tileClass = 1;
                wholeLarge.tileClass = 2;
                k_adj_uscatter<1><<<wholeItems, 64, 0, rs>>>(wholeSmall, s0, si);
                k_adj_uscatter<4><<<wholeItems, 256, 0, rs>>>(wholeLarge, s0, si);
            }
            k_adj_grid<0><<<blockWgs, kWG, 0, rs>>>(whole, a->segGrid[q], a->gridAdj);
            k_adj_p2g<<<denseWgs, kWG, 0, rs>>>(args, n, a->segKeys[q], s0, a->gridAdj, ad);
            r = check_launch();
        }
        return r;
    };
    if (!rc && b->kind == CRMPM_HIP_MLS)
    {
        // two sets of segment buffers when the second fits the tape budget
        const bool overlap = a->segments > 1 && mls_set_bytes(a) <= a->tapeBudget && mls_ensure_overlap(a) == CRMPM_HIP_OK;
        if (overlap)
            rc = overlapped_reverse(a, mrerun, [&](int sgm, int set, hipStream_t rs) {
                return mreverse(sgm, set, rs, a->revTiles);
            });
        for (int sgm = a->segments - 1; !overlap && sgm >= 0 && !rc; --sgm)
        {
            rc = mrerun(sgm, 0);
            if (!rc)
                rc = mreverse(sgm, 0, st, nullptr);
        }
    }
    // leave the solver in the state the forward pass ended in
    if (!rc)
        rc = state_copy(b, a->ckpt[a->segments], false);
    (void)hipStreamSynchronize(st);
    dfree(Fb);
    dfree(Cb);
    dfree(xOwn);
    return rc;
}


int crHipSvd3(const float *A9, int n, float *U9, float *Vt9, float *S3, void *stream)
{
    if (n < 0 || (n > 0 && (!A9 || !U9 || !Vt9 || !S3)))
        return CRMPM_HIP_ERR_ARG;
    if (n == 0)
        return CRMPM_HIP_OK;
    k_svd3<<<grid1d(n, kWG), kWG, 0, (hipStream_t)stream>>>(A9, n, U9, Vt9, S3);
    return check_launch();
}

int crHipParticleLattice(int geomType, const float params0[4], const CrHipGeometrySdfData *sdf, const float boundsMin[3],
                         const float boundsMax[3], float particleSpacing, const float position[3],
                         const float rotation[4], const float invScale[3], float particleMass, float *positionMassOut,
                         unsigned int capacity, unsigned int *numParticles, void *stream)
{
    LatticeArgs a{};
    int blocks = 0;
    int rc = lattice_setup(geomType, params0, sdf, boundsMin, boundsMax, particleSpacing, position, rotation, invScale,
                           particleMass, positionMassOut, capacity, numParticles, a, blocks);
    if (rc || blocks == 0)
        return rc;
    unsigned *counts = nullptr;
    if (dalloc(&counts, (size_t)blocks + 1))
        return CRMPM_HIP_ERR_ALLOC;
    rc = lattice_run(a, blocks, counts, positionMassOut, capacity, numParticles, static_cast<hipStream_t>(stream));
    dfree(counts);
    return rc;
}

int crHipParticleLatticeNoAlloc(int geomType, const float params0[4], const CrHipGeometrySdfData *sdf,
                                const float boundsMin[3], const float boundsMax[3], float particleSpacing,
                                const float position[3], const float rotation[4], const float invScale[3],
                                float particleMass, float *positionMassOut, unsigned int capacity,
                                unsigned int *numParticles, void *stream)
{
    LatticeArgs a{};
    int blocks = 0;
    int rc = lattice_setup(geomType, params0, sdf, boundsMin, boundsMax, particleSpacing, position, rotation, invScale,
                           particleMass, positionMassOut, capacity, numParticles, a, blocks);
    if (rc || blocks == 0)
        return rc;
    if (blocks > kLatticeScratchBlocks)
        return CRMPM_HIP_ERR_CAPACITY; // more than 2^28 lattice points: the scratch array holds no more
    unsigned *counts = nullptr;
    if ((rc = lattice_scratch(&counts)))
        return rc;
    return lattice_run(a, blocks, counts, positionMassOut, capacity, numParticles, static_cast<hipStream_t>(stream));
}

} // extern "C"

