/*
 * crmpm_hip.h -- C-ABI of the MI355X (gfx950) MPM substep solver.
 *
 * One opaque handle per solver instance replaces one GPU solver object of the
 * reference (GpuMlsMpmSolver / GpuPbMpmSolver / GpuStandardMpmSolver behind
 * crmpm::MpmSolver, include/solver/mpm_solver.h:63-190 and
 * src/solver/mpm_solver_gpu.h:43-57).  Each entry point below names the
 * reference member it stands in for.  Device pointers are plain pointers into
 * HBM with the reference's element layouts:
 *   positionMass  float4[N]  (x, y, z, mass)        include/solver/particle_data.h:48
 *   velocity      float4[N]  (Vec3f, 16 B)           include/solver/particle_data.h:49
 *   params0       float4[N], type int32[N]           include/solver/material_data.h:53-54
 *   activeMask    uint8[N]                           include/solver/mpm_solver.h:133
 * All calls return 0 on success or a CRMPM_HIP_* error code; none throws.
 * Handles are independent: no global mutable state, so different handles may
 * be driven from different host threads.  Calls on one handle must not overlap.
 */
#ifndef CRMPM_HIP_H
#define CRMPM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Solver kinds: the values of crmpm::MpmSolverType for the GPU solvers
 * (include/solver/mpm_solver.h:51-61). */
#define CRMPM_HIP_MLS 1
#define CRMPM_HIP_PB 3
#define CRMPM_HIP_STANDARD 5

/* crmpm::ShapeContactModel (include/solver/mpm_solver.h:45-49) */
#define CRMPM_HIP_CONTACT_KINEMATIC 0
#define CRMPM_HIP_CONTACT_EFFECTIVE_MASS 1

#define CRMPM_HIP_OK 0
#define CRMPM_HIP_ERR_ARG 1
#define CRMPM_HIP_ERR_STATE 2
#define CRMPM_HIP_ERR_ALLOC 3
#define CRMPM_HIP_ERR_LAUNCH 4
#define CRMPM_HIP_ERR_CAPACITY 5

typedef struct CrHipSolver CrHipSolver;

/* ShapeData SoA of device pointers (include/engine/shape.h:58-89).  Vec3f and
 * Quat entries are 16-byte float4 records. */
typedef struct CrHipShapeData
{
    int *type;           /* ShapeType: 0 static, 1 kinematic, 2 dynamic */
    int *geometryIdx;
    float *position;     /* float4 */
    float *rotation;     /* float4 quaternion (x, y, z, w) */
    float *invScale;     /* float4 */
    float *linearVelocity;
    float *angularVelocity;
    float *comInvMass;   /* float4: centre of mass, inverse mass */
    float *inertiaInv0;  /* float4: Ixx, Iyy, Izz */
    float *inertiaInv1;  /* float4: Ixy, Ixz, Iyz */
    float *force;        /* float4, accumulated coupling force */
    float *torque;       /* float4, accumulated coupling torque */
    float *params0;      /* float4: sdf smooth distance, drag, friction, fatten */
    int size;
} CrHipShapeData;

/* GeometryData SoA (include/engine/geometry.h:65-108) */
typedef struct CrHipGeometryData
{
    int *type;      /* GeometryType 0..8 */
    float *params0; /* float4 */
    int size;
} CrHipGeometryData;

/* GeometrySdfData SoA (include/engine/geometry.h:113-118) */
typedef struct CrHipGeometrySdfData
{
    int *dimension;               /* int4 (Vec3i) per SDF */
    float *lowerBoundCellSize;    /* float4 per SDF */
    float *gradientSignedDistance;/* float4 per SDF sample */
} CrHipGeometrySdfData;

/* Constructor GpuMlsMpmSolver(int numParticles, float cellSize, Bounds3 gridBound)
 * (src/solver/gpu_mls_mpm_solver.cu:40-52; PB :40-52; standard :37-49).
 * kind is CRMPM_HIP_MLS, CRMPM_HIP_PB or CRMPM_HIP_STANDARD. */
int crHipSolverCreate(int kind, int numMaxParticles, float cellSize, const float boundMin[3],
                      const float boundMax[3], CrHipSolver **out);
/* Releasable::release -> _release (src/solver/gpu_mls_mpm_solver.cu:111-124) */
int crHipSolverDestroy(CrHipSolver *s);

/* Grid node counts per axis, int(extent / cellSize + 1) (src/solver/gpu_mls_mpm_solver.cu:48-50) */
int crHipSolverGetGridDims(const CrHipSolver *s, int dims[3]);

/* MpmSolverGpu::setCudaStream (src/solver/mpm_solver_gpu.cpp:47-50); stream is a hipStream_t */
int crHipSolverSetStream(CrHipSolver *s, void *stream);
/* MpmSolver::setGravity (include/solver/mpm_solver.h:84) */
int crHipSolverSetGravity(CrHipSolver *s, float gx, float gy, float gz);
/* MpmSolver::setIntegrationStepSize (include/solver/mpm_solver.h:89) */
int crHipSolverSetIntegrationStepSize(CrHipSolver *s, float dt);
/* MpmSolver::setSolverIterations (include/solver/mpm_solver.h:96; PB only) */
int crHipSolverSetSolverIterations(CrHipSolver *s, int n);
/* MpmSolver::setShapeContactModel (include/solver/mpm_solver.h:182) */
int crHipSolverSetShapeContactModel(CrHipSolver *s, int model);

/* MpmSolver::bindParticleData (include/solver/mpm_solver.h:112) */
int crHipSolverBindParticleData(CrHipSolver *s, float *positionMass, float *velocity, int size);
/* MpmSolver::bindParticleMaterials (include/solver/mpm_solver.h:119) */
int crHipSolverBindParticleMaterials(CrHipSolver *s, float *params0, int *type, int size);
/* MpmSolver::bindActiveMask (include/solver/mpm_solver.h:136) */
int crHipSolverBindActiveMask(CrHipSolver *s, unsigned char *activeMask);
/* MpmSolver::setNumActiveParticles (include/solver/mpm_solver.h:127) */
int crHipSolverSetNumActiveParticles(CrHipSolver *s, unsigned int n);
/* MpmSolver::setNumShapes / bindShapeIds (include/solver/mpm_solver.h:141,147) */
int crHipSolverSetNumShapes(CrHipSolver *s, int n);
int crHipSolverBindShapeIds(CrHipSolver *s, int *shapeIds);
/* MpmSolver::bindShapeData / bindGeometryData / bindGeometrySdfData
 * (include/solver/mpm_solver.h:153,159,165); the structs are copied. */
int crHipSolverBindShapeData(CrHipSolver *s, const CrHipShapeData *d);
int crHipSolverBindGeometryData(CrHipSolver *s, const CrHipGeometryData *d);
int crHipSolverBindGeometrySdfData(CrHipSolver *s, const CrHipGeometrySdfData *d);

/* MpmSolver::initialize (include/solver/mpm_solver.h:70): allocates F, C, V0
 * and the grid, zeroed (src/solver/gpu_mls_mpm_solver.cu:54-83). */
int crHipSolverInitialize(CrHipSolver *s);
/* MpmSolver::computeInitialData (include/solver/mpm_solver.h:104): F := I and
 * V0 from the mass already on the grid plus these particles' mass; indices is
 * a device array of n particle indices. */
int crHipSolverComputeInitialData(CrHipSolver *s, unsigned int n, const unsigned int *indices);
/* MpmSolver::step (include/solver/mpm_solver.h:75): enqueue one substep on the
 * solver stream; *dtOut (may be NULL) receives the step size. */
int crHipSolverStep(CrHipSolver *s, float *dtOut);
/* MpmSolver::fetchResults (src/solver/mpm_solver_gpu.cpp:37-40): stream sync */
int crHipSolverFetchResults(CrHipSolver *s);

/* Solver-owned per-particle state, which the reference keeps private
 * (src/solver/gpu_mls_mpm_solver.cuh:61-77).  Device pointers into the
 * solver's arrays.  F and C hold 9 floats per particle, the reference
 * Mat3x3f's element order with its column padding dropped: element (row r,
 * column c) of particle p is F[9 * p + 3 * c + r]; same for C; V0[p].  stride is the number of particle slots.  Used for checkpoints,
 * the adjoint and parity checks. */
int crHipSolverGetStateArrays(CrHipSolver *s, float **F, float **C, float **V0, int *stride);
/* Binning stage of a substep, exposed for verification; it has no counterpart
 * in the reference, which scatters in particle order.  For the current
 * positions of the numActive particles: keysOut[k] is the k-th smallest
 * 4x4x4-cell block key (bx + nbx * (by + nby * bz) of the base cell
 * trunc((x - min) / dx - 0.5) >> 2, clamped to the grid; inactive particles
 * get the key nbx * nby * nbz) and orderOut[k] the particle index, ties in
 * index order (a stable sort).  Both are device arrays of numActive entries. */
int crHipSolverBinParticles(CrHipSolver *s, unsigned int *keysOut, unsigned int *orderOut);
/* crHipBatchAutotune for the solver's one-scene batch (no counterpart in the
 * reference): picks the tile split and launch sizes on the current state. */
int crHipSolverAutotune(CrHipSolver *s, int stepsPerTrial, int *tileWaves, int *updateWaves, int *subSize, int *subMax);
/* A launch configuration as crHipSolverAutotune reports it (tile split,
 * gather split, launch sizes), set without timing trials: crHipBatchSetTileSplit,
 * crHipBatchSetGatherSplit and crHipBatchSetLaunchSizes for the solver's
 * one-scene batch (e.g. to profile the configuration a timed run used).  No
 * counterpart in the reference. */
int crHipSolverSetLaunchConfig(CrHipSolver *s, int tileWaves, int updateWaves, int subSize, int subMax,
                               int gatherSubSize, int gatherSubMax);
/* crHipBatchSetGraphMode for the solver's one-scene batch: replay the
 * substep as a hipGraph (no counterpart in the reference). */
int crHipSolverSetGraphMode(CrHipSolver *s, int enabled);
/* crHipBatchGetGatherSplit for the solver's one-scene batch (the split the
 * autotuner picked for the G2P kernels). */
int crHipSolverGetGatherSplit(const CrHipSolver *s, int *subSize, int *subMax);
/* crHipBatchSetScatterWaves / crHipBatchGetScatterWaves for the solver's
 * one-scene batch (no counterpart in the reference). */
int crHipSolverSetScatterWaves(CrHipSolver *s, int waves);
int crHipSolverGetScatterWaves(const CrHipSolver *s, int *waves);
/* Launch timing, an addition of this solver (the reference times nothing on
 * the device).  Between Begin and End every particle-transfer, grid-update and
 * binning launch of crHipSolverStep is bracketed by a pair of hipEvents on the
 * solver stream (at most maxLaunches launches are recorded).  End synchronizes
 * the stream and returns the summed milliseconds and launch counts per
 * category (CRMPM_HIP_PROF_*). */
#define CRMPM_HIP_PROF_PB_FUSED 0  /* PB: G2P of iteration k-1 with the constraint of iteration k, per particle */
#define CRMPM_HIP_PROF_PB_P2G 1    /* PB: the first iteration's constraint and P2G */
#define CRMPM_HIP_PROF_PB_G2P 2    /* PB: last G2P fused with the particle integration */
#define CRMPM_HIP_PROF_UPDATE 3    /* grid update, all solvers */
#define CRMPM_HIP_PROF_MLS_P2G 4
#define CRMPM_HIP_PROF_MLS_G2P 5
#define CRMPM_HIP_PROF_STD_P2G 6
#define CRMPM_HIP_PROF_STD_G2P 7
#define CRMPM_HIP_PROF_BINNING 8   /* binning, radix sort, tiling, PB staging */
#define CRMPM_HIP_PROF_PB_SCATTER 9 /* PB: P2G scatter of iterations 1 .. K-1, per pblock */
#define CRMPM_HIP_PROF_CATEGORIES 10
int crHipSolverProfileBegin(CrHipSolver *s, int maxLaunches);
int crHipSolverProfileEnd(CrHipSolver *s, float *msPerCategory, int *launchesPerCategory, int numCategories);
/* Between ProfileBegin and ProfileEnd, record events only for launches whose
 * category bit (1 << CRMPM_HIP_PROF_*) is set in mask; 0 pauses recording.
 * ProfileBegin sets the mask to all categories.  Each recorded launch costs a
 * few microseconds of dispatch time, so timing runs sample a subset. */
int crHipSolverProfileMask(CrHipSolver *s, unsigned int mask);
/* Copies the grid (mass-weighted momentum or velocity in xyz, mass in w) to a
 * device float4 array in node order x + y * nx + z * nx * ny. */
int crHipSolverCopyGrid(CrHipSolver *s, float *dst);

/* ------------------------------------------------------------------------
 * Batched scenes.  A batch advances many scenes of one solver kind together:
 * each phase of a substep (binning, sort, P2G, grid update, G2P; PB: its
 * iteration kernels) is one launch over the pblocks of every scene, so small
 * scenes share the GPU instead of each paying its own launch sequence.  The
 * scenes' particles are slices of one set of bound particle arrays, as the
 * engine's per-scene particle ranges are slices of the factory's particle
 * block (src/engine/scene_impl.h:143-160); scenes may differ in particle
 * count, grid and parameters.  Replaces SimulationFactory::advanceAll's one
 * solver step per scene (src/engine/simulation_factory_impl.cpp:1106-1150).
 * A CrHipSolver is a batch of one scene. */
typedef struct CrHipBatch CrHipBatch;

typedef struct CrHipSceneDesc
{
    int numMaxParticles; /* particles of the scene (its slice of the bound arrays) */
    int particleOffset;  /* index of the scene's first particle in the bound arrays */
    float cellSize;      /* SceneDesc::gridCellSize */
    float boundMin[3];   /* SceneDesc::gridBounds */
    float boundMax[3];
} CrHipSceneDesc;

/* CRMPM_HIP_ERR_CAPACITY when the scenes' grids total 2^24 node blocks of
 * 4x4x4 nodes (about 1.07e9 nodes) or more, or their particles 2^31. */
int crHipBatchCreate(int kind, int numScenes, const CrHipSceneDesc *scenes, CrHipBatch **out);
int crHipBatchDestroy(CrHipBatch *b);
int crHipBatchSetStream(CrHipBatch *b, void *stream);
/* PB-MPM iterations, one count for the whole batch */
int crHipBatchSetSolverIterations(CrHipBatch *b, int n);
/* Bound arrays hold every scene's slice: size >= max(particleOffset + numMaxParticles). */
int crHipBatchBindParticleData(CrHipBatch *b, float *positionMass, float *velocity, int size);
int crHipBatchBindParticleMaterials(CrHipBatch *b, float *params0, int *type, int size);
int crHipBatchBindActiveMask(CrHipBatch *b, unsigned char *activeMask);
int crHipBatchInitialize(CrHipBatch *b);
int crHipBatchGetGridDims(const CrHipBatch *b, int scene, int dims[3]);
/* Per-scene settings (MpmSolverBase setters of each scene's solver) */
int crHipBatchSetSceneGravity(CrHipBatch *b, int scene, float gx, float gy, float gz);
int crHipBatchSetSceneStepSize(CrHipBatch *b, int scene, float dt);
int crHipBatchSetSceneContactModel(CrHipBatch *b, int scene, int model);
/* Particles [particleOffset, particleOffset + n) of the scene take part. */
int crHipBatchSetSceneNumActiveParticles(CrHipBatch *b, int scene, unsigned int n);
/* The scene sits out the following substeps while resting != 0: none of its
 * particles is processed and its grid is neither cleared nor written, so the
 * grid keeps its last update (what a solver that is not stepped holds, e.g.
 * one that has taken its substeps in advanceAll,
 * src/engine/simulation_factory_impl.cpp:1123-1136; computeInitialData then
 * adds mass onto it, src/solver/gpu_mls_mpm_solver.cu:85-100).  No
 * counterpart in the reference, whose solvers are stepped one by one. */
int crHipBatchSetSceneResting(CrHipBatch *b, int scene, int resting);
/* A scene's shapes: setNumShapes + bindShapeIds + bindShapeData + bindGeometryData
 * + bindGeometrySdfData of its solver (null table pointers keep the previous binding). */
int crHipBatchBindSceneShapes(CrHipBatch *b, int scene, int numShapes, int *shapeIds, const CrHipShapeData *shapes,
                              const CrHipGeometryData *geometries, const CrHipGeometrySdfData *sdf);
/* computeInitialData of one scene; indices are scene-local particle indices (device). */
int crHipBatchComputeInitialData(CrHipBatch *b, int scene, unsigned int n, const unsigned int *indices);
/* One substep of every scene, enqueued on the batch stream. */
int crHipBatchStep(CrHipBatch *b);
/* With graph mode on (and a non-null stream), crHipBatchStep captures its
 * launches into a hipGraph once and replays it while the launch configuration
 * (active particle count, iterations, launch sizes, bindings) is unchanged. */
int crHipBatchSetGraphMode(CrHipBatch *b, int enabled);
/* Tile split: each pblock's particles run as up to subMax sub-tiles of subSize
 * particles (a multiple of 64; the last sub-tile takes the rest), one
 * wavefront and one partial box each, so dense pblocks spread over several
 * wavefronts.  subMax <= 8 (fewer when the partial boxes of a very large
 * batch would exceed 8 GiB: CRMPM_HIP_ERR_CAPACITY).  Default: whole pblocks
 * (subMax 1). */
int crHipBatchSetTileSplit(CrHipBatch *b, int subSize, int subMax);
/* Autotuner: times substeps of the batch's current state (saved and restored
 * around the trials, so the state is unchanged afterwards) for several tile
 * splits, gather splits and launch sizes and keeps the fastest.  Launch sizes are in
 * wavefronts; 0 means one wavefront per (sub-)tile or node block.  Any
 * output pointer may be null. */
int crHipBatchAutotune(CrHipBatch *b, int stepsPerTrial, int *tileWaves, int *updateWaves, int *subSize, int *subMax);
int crHipBatchSetLaunchSizes(CrHipBatch *b, int tileWaves, int updateWaves);
/* Split of the gather kernels (G2P), which write no partial boxes: a
 * pblock's particles in sub-tiles of subSize (a multiple of 64), at most
 * subMax (1..64) of them, the last taking the rest.  subMax 0, for a batch of
 * one scene only (else CRMPM_HIP_ERR_ARG): G2P per particle in particle index
 * order, 64 consecutive particles per wavefront, with the same results.
 * Default whole pblocks; crHipBatchAutotune also picks it.  No counterpart in
 * the reference. */
int crHipBatchSetGatherSplit(CrHipBatch *b, int subSize, int subMax);
int crHipBatchGetGatherSplit(const CrHipBatch *b, int *subSize, int *subMax);
/* Wavefronts per tile (1, 2 or 4, or 0: four for tiles of more than 64
 * particles and one for the others; else CRMPM_HIP_ERR_ARG) of the PB-MPM
 * and MLS scatters (P2G) and of the adjoints' velocity-adjoint scatters:
 * with more than one, a tile's chunks are shared by the wavefronts of one
 * workgroup, each into its own LDS accumulation tiles, summed in a fixed
 * order into the tile's one partial box (for dense pblocks, whose chunks
 * otherwise run on one wavefront).  Deterministic for a given setting; the summation order, and so
 * the rounding, depends on it.  Default 1; crHipBatchAutotune picks it for PB
 * and MLS batches.  No counterpart in the reference. */
int crHipBatchSetScatterWaves(CrHipBatch *b, int waves);
int crHipBatchGetScatterWaves(const CrHipBatch *b, int *waves);
int crHipBatchFetchResults(CrHipBatch *b);
int crHipBatchGetStateArrays(CrHipBatch *b, float **F, float **C, float **V0, int *stride);
int crHipBatchCopyGrid(CrHipBatch *b, int scene, float *dst);
int crHipBatchProfileBegin(CrHipBatch *b, int maxLaunches);
int crHipBatchProfileMask(CrHipBatch *b, unsigned int mask);
int crHipBatchProfileEnd(CrHipBatch *b, float *msPerCategory, int *launchesPerCategory, int numCategories);

/* ------------------------------------------------------------------------
 * Engine runtime: the GPU side of SimulationFactoryImpl and SceneImpl
 * (src/engine/simulation_factory_impl.cpp:897-955, 1106-1150;
 * src/engine/scene_impl.cpp:180-263).  An engine owns the particle block and
 * the shape / geometry / SDF tables, each as a pinned host copy and a device
 * copy; scenes are slices of the particle block.  The caller writes the host
 * copies and marks what changed; crHipEngineAdvanceAll uploads the marked
 * arrays, runs the queued computeInitialData, resets the shapes' coupling
 * force and torque, advances every scene by dt (scenes of one solver kind and
 * iteration count as one CrHipBatch), and downloads positions, velocities
 * and the coupling forces into the host copies; it runs on a worker thread
 * (SceneImpl::advance submits to a thread manager) and
 * crHipEngineFetchResultsAll waits for it.  While an advance is in flight
 * every other call returns CRMPM_HIP_ERR_STATE. */
typedef struct CrHipEngine CrHipEngine;

/* SceneDataDirtyFlags (include/engine/gpu_data_dirty_flags.h:86-99), plus the
 * active mask, which the reference writes to the device at once. */
#define CRMPM_HIP_SCENE_POSITION_MASS (1u << 0)
#define CRMPM_HIP_SCENE_VELOCITY (1u << 1)
#define CRMPM_HIP_SCENE_PARAMS0 (1u << 2)
#define CRMPM_HIP_SCENE_TYPE (1u << 3)
#define CRMPM_HIP_SCENE_NUM_PARTICLES (1u << 4)
#define CRMPM_HIP_SCENE_COMPUTE_INITIAL_DATA (1u << 5)
#define CRMPM_HIP_SCENE_ACTIVE_MASK (1u << 6)
#define CRMPM_HIP_SCENE_ALL 0x7fu
/* Table groups of SimulationFactoryGpuDataDirtyFlags
 * (include/engine/gpu_data_dirty_flags.h): every shape array but force and
 * torque; geometry type and params; SDF dims and lower bound|cell size; SDF rows. */
#define CRMPM_HIP_ENGINE_SHAPES (1u << 0)
#define CRMPM_HIP_ENGINE_GEOMETRIES (1u << 1)
#define CRMPM_HIP_ENGINE_SDF (1u << 2)
#define CRMPM_HIP_ENGINE_SDF_ROWS (1u << 3)

/* createFactory's capacities (src/engine/simulation_factory_impl.cpp): the
 * particle block, shapes, geometries, SDF tables and SDF rows. */
int crHipEngineCreate(int particleCapacity, int shapeCapacity, int geometryCapacity, int sdfTableCapacity,
                      int sdfRowCapacity, CrHipEngine **out);
int crHipEngineDestroy(CrHipEngine *e);
/* SimulationFactory::createScene (simulation_factory_impl.cpp:278-340): a
 * slice of numMaxParticles particles and a grid; kind CRMPM_HIP_MLS / PB /
 * STANDARD.  The set of scenes is fixed at the first advance. */
int crHipEngineCreateScene(CrHipEngine *e, int kind, int numMaxParticles, float cellSize, const float boundMin[3],
                           const float boundMax[3], int *sceneId);
/* SceneDesc's solverIntegrationStepSize, gravity, contactModel and
 * solverIterations (include/engine/simulation_factory.h:49-59; defaults
 * 0.002, (0, -9.81, 0), kinematic, 10).  After the first advance the step
 * size, gravity and contact model may still change (the solver setters); a
 * PB scene's iteration count may not (CRMPM_HIP_ERR_STATE). */
int crHipEngineSetSceneParams(CrHipEngine *e, int scene, float dt, const float gravity[3], int contactModel,
                              int iterations);
/* Scene::getParticleData / getParticleMaterialData / getActiveMask: the
 * scene's slice of the pinned host copies (position|mass and velocity as
 * float4, params0 float4, type int, mask u8).  Any output may be NULL. */
int crHipEngineGetSceneHostData(CrHipEngine *e, int scene, float **positionMass, float **velocity, float **params0,
                                int **type, unsigned char **activeMask);
/* Scene::markDirty: CRMPM_HIP_SCENE_* flags */
int crHipEngineMarkSceneDirty(CrHipEngine *e, int scene, unsigned int flags);
/* The scene's allocated particle count (setNumActiveParticles at the next advance). */
int crHipEngineSetSceneNumParticles(CrHipEngine *e, int scene, unsigned int n);
/* Scene::getComputeInitialDataIndices().pushBack: scene-local indices whose
 * F, V0 are set up at the next advance. */
int crHipEngineAddComputeInitialData(CrHipEngine *e, int scene, unsigned int n, const unsigned int *indices);
/* Scene::setActiveMaskRange (scene_impl.cpp:130-139) */
int crHipEngineSetActiveMaskRange(CrHipEngine *e, int scene, unsigned int start, unsigned int length,
                                  unsigned char value);
/* Scene::addShape: the shape ids (rows of the engine's shape table) the scene collides with. */
int crHipEngineSetSceneShapes(CrHipEngine *e, int scene, int numShapes, const int *shapeIds);
/* The engine's tables: device != 0 for the device copies (e.g. to build SDF
 * rows there with crHipMeshSdf), 0 for the pinned host copies. */
int crHipEngineGetTables(CrHipEngine *e, int device, CrHipShapeData *shapes, CrHipGeometryData *geometries,
                         CrHipGeometrySdfData *sdf);
/* markDirty of the factory's GPU data: CRMPM_HIP_ENGINE_* groups. */
int crHipEngineMarkTablesDirty(CrHipEngine *e, unsigned int flags);
/* SimulationFactory::advanceAll(dt) / fetchResultsAll(): each scene takes
 * the reference's substep count (dt accumulated in FP32 until it reaches
 * dt); within a batch a scene that has taken its count rests
 * (crHipBatchSetSceneResting) while the others finish, so its grid keeps its
 * last update, as the reference's solver that is not stepped does, and
 * computeInitialData queued for it at the next advance adds mass onto that
 * grid.  FetchResultsAll returns the advance's status; CRMPM_HIP_ERR_STATE
 * if no advance is in flight. */
int crHipEngineAdvanceAll(CrHipEngine *e, float dt);
int crHipEngineFetchResultsAll(CrHipEngine *e);
/* crHipBatchSetGraphMode / crHipBatchAutotune on every batch of the engine
 * (autotune after the first advance: it times the current state and leaves
 * it unchanged).  No counterpart in the reference. */
int crHipEngineSetGraphMode(CrHipEngine *e, int enabled);
int crHipEngineAutotune(CrHipEngine *e, int stepsPerTrial);
/* Substeps the scene took in the last advance (no counterpart in the reference). */
int crHipEngineGetSceneSubsteps(const CrHipEngine *e, int scene, int *substeps);
/* The scene's solver state arrays (as crHipBatchGetStateArrays; the scene's
 * particles start at index *first).  After the first advance.  No
 * counterpart in the reference. */
int crHipEngineGetSceneStateArrays(CrHipEngine *e, int scene, float **F, float **C, float **V0, int *stride,
                                   int *first);

/* ------------------------------------------------------------------------
 * Triangle mesh -> SDF table on the device: the table the engine builds on
 * the host for a triangle-mesh geometry (SimulationFactoryImpl::createGeometry,
 * src/engine/simulation_factory_impl.cpp:628-635).
 *
 * crHipMeshSdfDesc replaces SdfBuilder::precomputeSdfDesc
 * (src/geometry_utils/sdf_builder.cpp:84-101): from host arrays points
 * (float xyz per vertex) and triangles (3 vertex indices each), the table's
 * dims (int(size / cellSize + 1) per axis) and lower bound (the union of
 * the triangles' boxes -- each the reference's zero-initialised box grown to
 * its corners, so holding the origin -- grown by 1e-4 and by fatten).
 *
 * crHipMeshSdf replaces SdfBuilder::computeSdf + computeSdfGradient
 * (sdf_builder.cpp:103-160, 242-330): writes dims and lowerBound as above
 * and, when gradientSignedDistance (device float4[count], the geometry's rows
 * of CrHipGeometrySdfData::gradientSignedDistance) is not NULL, fills row
 * z*dx*dy + y*dx + x with the normalised gradient (xyz) and signed distance
 * (w) at lowerBound + (x, y, z) * cellSize; negative inside: where the
 * reference's approximate winding number (its AABB tree's clusters, beta = 2,
 * src/geometry_utils/winding_number.cpp:136-171) exceeds 0.5.  Synchronous on
 * stream.  Errors: CRMPM_HIP_ERR_ARG for a missing pointer, no triangles,
 * cellSize <= 0, fatten < 0, a vertex index beyond numPoints, a non-finite
 * vertex, or a table with one node along an axis; CRMPM_HIP_ERR_CAPACITY for
 * capacity < dx*dy*dz rows or 2^31 rows or more.
 *
 * crHipMeshSdfWithRule: the same with rule CRMPM_HIP_MESH_RULE_REFERENCE
 * (crHipMeshSdf) or CRMPM_HIP_MESH_RULE_EXACT: the sign from the exact
 * winding number (computeWindingNumberExact, winding_number.h:136-146), which
 * differs from the reference's approximation near 0.5 (nodes on a closed
 * surface, near the rim of an open one).  No counterpart in the reference. */
#define CRMPM_HIP_MESH_RULE_REFERENCE 0
#define CRMPM_HIP_MESH_RULE_EXACT 1
int crHipMeshSdfDesc(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                     float cellSize, float fatten, int dims[3], float lowerBound[3]);
int crHipMeshSdf(const float *points, int numPoints, const unsigned int *triangles, int numTriangles, float cellSize,
                 float fatten, int dims[3], float lowerBound[3], float *gradientSignedDistance,
                 unsigned long long capacity, void *stream);
int crHipMeshSdfWithRule(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                         float cellSize, float fatten, int rule, int dims[3], float lowerBound[3],
                         float *gradientSignedDistance, unsigned long long capacity, void *stream);

/* The mesh-slicer table the engine builds for a triangle-mesh slicer geometry
 * (src/engine/simulation_factory_impl.cpp:653-690): SdfBuilder::computeModSdf
 * + computeSdfGradient + postProcessModSdfGradient (sdf_builder.cpp:162-240,
 * 242-357).  rows (device float4[dx*dy*dz + 1]) receives the spine area
 * (normal xyz, offset w; x NaN for none) in row 0, then per node the unsigned
 * distance (w) and its normalised gradient (xyz), all four NaN where the
 * reference's warm-started tree walk (each node starting from the previous
 * node's closest triangle, aabb_tree.cpp:724-779) accepted a candidate whose
 * closest point lies on a boundary edge or boundary vertex of the mesh (a
 * boundary edge has one adjacent triangle), unless the node is on the
 * positive side of the spine plane; a node whose walk accepts nothing past
 * its warm start gets distance 0, as in the reference.  Errors as
 * crHipMeshSdf; capacity counts the spine row; CRMPM_HIP_ERR_CAPACITY also
 * for a tree deeper than 64 levels; CRMPM_HIP_ERR_STATE if the chain of warm
 * starts has not settled after 16 passes.
 *
 * crHipMeshSlicerSdfWithRule with CRMPM_HIP_MESH_RULE_EXACT drops a row only
 * where the final closest point (the first closest triangle in index order)
 * lies on the boundary: a superset of the reference's rows.  No counterpart
 * in the reference. */
int crHipMeshSlicerSdf(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                       float cellSize, float fatten, const float spineArea[4], int dims[3], float lowerBound[3],
                       float *rows, unsigned long long capacity, void *stream);
int crHipMeshSlicerSdfWithRule(const float *points, int numPoints, const unsigned int *triangles, int numTriangles,
                               float cellSize, float fatten, const float spineArea[4], int rule, int dims[3],
                               float lowerBound[3], float *rows, unsigned long long capacity, void *stream);

/* ------------------------------------------------------------------------
 * Particle lattice of a particle object on the device:
 * ParticleObjectImpl::initialize (src/engine/particle_object_impl.cpp:36-88),
 * which SimulationFactoryImpl::createParticleObject runs on the host
 * (src/engine/simulation_factory_impl.cpp:776-783).  boundsMin/boundsMax are
 * the geometry's Geometry::getBounds(); the lattice covers them fattened by
 * two spacings, with ((max - min) / spacing / invScale) points per axis
 * (truncated).  Point q = (ix, iy, iz) * spacing * invScale + min becomes a
 * particle where the geometry's SDF at q is <= 0, at rotation.rotate(q /
 * invScale) + position with mass particleMass.  Particles are written as
 * float4 position|mass in the reference's order (ix outer, iz inner), at most
 * capacity of them, to the device array positionMassOut (e.g. the bound
 * positionMass array at the object's offset); *numParticles receives the
 * full count (call with capacity 0 to size an allocation).  Synchronous on
 * stream.  sdf is required for the table geometries (triangle mesh, mesh
 * slicer, line segments) and may be NULL otherwise.  Errors:
 * CRMPM_HIP_ERR_ARG for a missing pointer, a geometry type outside 0..8,
 * spacing <= 0, invScale <= 0 or unbounded bounds (plane, arc);
 * CRMPM_HIP_ERR_CAPACITY for 2^31 lattice points or more. */
int crHipParticleLattice(int geomType, const float params0[4], const CrHipGeometrySdfData *sdf, const float boundsMin[3],
                         const float boundsMax[3], float particleSpacing, const float position[3],
                         const float rotation[4], const float invScale[3], float particleMass, float *positionMassOut,
                         unsigned int capacity, unsigned int *numParticles, void *stream);
/* The device 3x3 SVD that PB-MPM's material constraint runs per particle
 * and iteration, on n matrices: svd() of include/math/svd3.h:1265-1275
 * (McAdams et al., 4 Jacobi sweeps), which factors the transpose:
 * u diag(sigma) vt = A^T.  Matrices are column-major, element (r, c) of
 * matrix i at [9 i + 3 c + r], the reference Mat3x3f's element order; A9,
 * U9, Vt9 and S3 (n x 3) are device arrays.  The reference runs this only
 * inside its solvers; the entry point exists so the device SVD can be
 * checked on its own.  Asynchronous on stream.  Errors: CRMPM_HIP_ERR_ARG
 * for n < 0 or a missing array. */
int crHipSvd3(const float *A9, int n, float *U9, float *Vt9, float *S3, void *stream);
/* crHipParticleLattice without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * stream capture, tight loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors, except that the per-block counts live in a
 * fixed device array of 512 KiB per GPU (taken on the first call on a device
 * and held until the process ends): lattices of more than 2^28 points return
 * CRMPM_HIP_ERR_CAPACITY.  Calls that overlap on one GPU (other threads,
 * other streams) share that array and race: serialise them. */
int crHipParticleLatticeNoAlloc(int geomType, const float params0[4], const CrHipGeometrySdfData *sdf,
                                const float boundsMin[3], const float boundsMax[3], float particleSpacing,
                                const float position[3], const float rotation[4], const float invScale[3],
                                float particleMass, float *positionMassOut, unsigned int capacity,
                                unsigned int *numParticles, void *stream);

/* ------------------------------------------------------------------------
 * Differentiable MLS-MPM and PB-MPM (one scene, no rigid shapes).  The
 * reference has no adjoint; this is the differentiable form of its MLS
 * substep (src/solver/mls_mpm_algorithm.h:147-390) and of its PB substep
 * (src/solver/pb_mpm_algorithm.h:150-408 with the SVD-based constraint of
 * src/solver/constitutive_models.h:123-154), same arithmetic and quirks.
 * Forward runs numSteps substeps from the solver's current state and keeps
 * the state every checkpointEvery substeps; Backward takes dL/dx (and
 * optionally dL/dv) of the final state, recomputes each segment from its
 * checkpoint and sweeps the adjoint kernels back to substep 0 (PB: each
 * substep is re-run once more with its iterations' grids kept).  Outputs
 * (device arrays, particle order): the gradients w.r.t. params0.x and
 * params0.y per particle (FP64, summed over the substeps) -- dL/dlambda and
 * dL/dmu for MLS, dL/dalpha and dL/dratio (the constraint's relaxation and
 * rotation ratio) for PB -- and dL/dv and dL/dx of the initial state
 * (float4, xyz used; dLdx0 may be null).  Backward leaves the solver in the
 * state Forward ended in.  Other solver kinds: CRMPM_HIP_ERR_STATE. */
typedef struct CrHipAdjoint CrHipAdjoint;
int crHipAdjointCreate(CrHipSolver *s, int numSteps, int checkpointEvery, CrHipAdjoint **out);
int crHipAdjointForward(CrHipAdjoint *a);
int crHipAdjointBackward(CrHipAdjoint *a, const float *dLdx, const float *dLdv, double *dLdLambda, double *dLdMu,
                         float *dLdx0, float *dLdv0);
/* No counterpart in the reference; for callers that trade device memory for
 * time: the device memory the backward pass may take for checkpoint
 * segments' tapes.  PB: a segment's tape is every substep's iteration grids,
 * sort products and staged particle arrays, about
 * iterations x nodes x 16 B + 72 B x particles per substep.  When two
 * segments' tapes fit, segment s is reversed on a stream of the adjoint's
 * own while segment s - 1 is re-run on the solver's stream; when one fits,
 * each segment is re-run once, then reversed; otherwise each substep is
 * re-run once more before it is reversed.  MLS: when a second set of
 * per-segment buffers (grids and sort products, nodes x 16 B + 8 B x
 * particles per substep) fits, the re-runs overlap the reverses the same way.
 * With overlap the backward's work is complete when the solver's stream is.
 * Default 8 GiB.  The gradients are bitwise the same in every mode. */
int crHipAdjointSetTapeBudget(CrHipAdjoint *a, unsigned long long bytes);
/* The bytes of one segment's tape (PB) or of one set of per-segment buffers
 * (MLS); no counterpart in the reference. */
int crHipAdjointGetSegmentTapeBytes(const CrHipAdjoint *a, unsigned long long *bytes);
int crHipAdjointDestroy(CrHipAdjoint *a);

#ifdef __cplusplus
}
#endif

#endif /* CRMPM_HIP_H */
