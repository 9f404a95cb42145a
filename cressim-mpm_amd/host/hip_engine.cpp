// Engine-side GPU runtime over the batched solver.
//
// The GPU half of the reference's SimulationFactoryImpl and SceneImpl
// (src/engine/simulation_factory_impl.cpp:897-955, 1106-1150,
// src/engine/scene_impl.cpp:180-263): host copies of the particle block and
// of the shape / geometry / SDF tables in pinned memory, dirty flags that say
// what to upload before the next advance, the per-scene initial-data queue,
// the coupling-force reset, the substep loop of advanceAll, and the download
// of positions, velocities and coupling forces afterwards.  What differs is
// the substep loop: the reference steps each scene's solver in turn
// (`while (t < dt) t += solver->step()`); here the scenes that share a solver
// kind and iteration count are one CrHipBatch, advanced by one launch per
// phase, and a scene whose float-accumulated time has reached dt rests for
// the remaining substeps of the batch: its particles are not processed and
// its grid keeps the last update, as a solver that is not stepped does.
//
// advanceAll runs on a worker thread, as the reference submits it to its
// thread manager; fetchResultsAll joins it.  Calls other than fetch are
// refused while an advance is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <thread>
#include <utility>
#include <vector>

#include "crmpm_hip.h"

namespace
{
    template <class T> int host_alloc(T **p, size_t n)
    {
        *p = nullptr;
        if (n == 0)
            n = 1;
        if (hipHostMalloc(reinterpret_cast<void **>(p), n * sizeof(T), hipHostMallocDefault) != hipSuccess)
            return CRMPM_HIP_ERR_ALLOC;
        std::memset(*p, 0, n * sizeof(T));
        return CRMPM_HIP_OK;
    }

    template <class T> int dev_alloc(T **p, size_t n)
    {
        *p = nullptr;
        if (n == 0)
            n = 1;
        if (hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T)) != hipSuccess)
            return CRMPM_HIP_ERR_ALLOC;
        return hipMemset(*p, 0, n * sizeof(T)) == hipSuccess ? CRMPM_HIP_OK : CRMPM_HIP_ERR_ALLOC;
    }

    // One set of shape / geometry / SDF tables (host or device).  Field order
    // and element layouts are CrHipShapeData's, CrHipGeometryData's and
    // CrHipGeometrySdfData's.
    struct Tables
    {
        int *type = nullptr, *geom = nullptr;
        float4 *f4[11] = {};  // position .. params0 in CrHipShapeData order (force = 8, torque = 9)
        int *gtype = nullptr;
        float4 *gparams = nullptr;
        int4 *sdfDims = nullptr;
        float4 *sdfLbcs = nullptr, *sdfRows = nullptr;
    };
    constexpr int kForce = 8, kTorque = 9;

    CrHipShapeData shape_view(const Tables &t, int cap)
    {
        CrHipShapeData s;
        s.type = t.type;
        s.geometryIdx = t.geom;
        float **f[11] = {&s.position, &s.rotation, &s.invScale, &s.linearVelocity, &s.angularVelocity, &s.comInvMass,
                         &s.inertiaInv0, &s.inertiaInv1, &s.force, &s.torque, &s.params0};
        for (int k = 0; k < 11; ++k)
            *f[k] = reinterpret_cast<float *>(t.f4[k]);
        s.size = cap;
        return s;
    }
} // namespace

struct CrHipEngine
{
    int particleCap = 0, shapeCap = 0, geomCap = 0, sdfTableCap = 0, sdfRowCap = 0;
    hipStream_t stream = nullptr;
    // particle block: pinned host copy and device arrays
    float4 *hXm = nullptr, *hV = nullptr, *hParams = nullptr, *dXm = nullptr, *dV = nullptr, *dParams = nullptr;
    int *hType = nullptr, *dType = nullptr;
    unsigned char *hMask = nullptr, *dMask = nullptr;
    Tables h, d;
    unsigned tablesDirty = 0;

    struct Scene
    {
        int kind;
        CrHipSceneDesc desc;
        float dt = 0.002f;
        float gravity[3] = {0.0f, -9.81f, 0.0f};
        int contactModel = CRMPM_HIP_CONTACT_KINEMATIC;
        int iterations = 10;
        unsigned numAllocated = 0;
        unsigned dirty = 0;
        std::vector<unsigned> initIdx;
        std::vector<int> shapeIds;
        bool shapesChanged = true;
        int *dShapeIds = nullptr;
        int batch = -1, slot = -1;  // batch and scene index in it
        int substeps = 0;           // taken by the last advance
    };
    std::vector<Scene> scenes;
    int nextOffset = 0;

    struct Batch
    {
        CrHipBatch *b = nullptr;
        std::vector<int> scenes;  // engine scene ids, in batch order
    };
    std::vector<Batch> batches;
    bool built = false;
    bool graphMode = false;

    unsigned *dIdx = nullptr;
    size_t dIdxCap = 0;

    std::thread worker;
    bool busy = false;
    int workerRc = CRMPM_HIP_OK;
};

namespace
{
    int build_batches(CrHipEngine *e)
    {
        // scenes grouped by solver kind and iteration count, one batch each
        std::map<std::pair<int, int>, std::vector<int>> groups;
        for (int i = 0; i < (int)e->scenes.size(); ++i)
        {
            const auto &s = e->scenes[i];
            groups[{s.kind, s.kind == CRMPM_HIP_PB ? s.iterations : 0}].push_back(i);
        }
        for (auto &g : groups)
        {
            CrHipEngine::Batch bt;
            bt.scenes = g.second;
            std::vector<CrHipSceneDesc> descs;
            for (int i : bt.scenes)
                descs.push_back(e->scenes[i].desc);
            int rc = crHipBatchCreate(g.first.first, (int)descs.size(), descs.data(), &bt.b);
            if (rc)
                return rc;
            e->batches.push_back(bt);
            CrHipBatch *b = bt.b;
            const int bi = (int)e->batches.size() - 1;
            if ((rc = crHipBatchSetStream(b, e->stream)) ||
                (rc = crHipBatchSetSolverIterations(b, std::max(1, g.first.second))) ||
                (rc = crHipBatchBindParticleData(b, reinterpret_cast<float *>(e->dXm), reinterpret_cast<float *>(e->dV),
                                                 e->particleCap)) ||
                (rc = crHipBatchBindParticleMaterials(b, reinterpret_cast<float *>(e->dParams), e->dType,
                                                      e->particleCap)) ||
                (rc = crHipBatchBindActiveMask(b, e->dMask)))
                return rc;
            for (int k = 0; k < (int)bt.scenes.size(); ++k)
            {
                auto &s = e->scenes[bt.scenes[k]];
                s.batch = bi;
                s.slot = k;
                if ((rc = crHipBatchSetSceneGravity(b, k, s.gravity[0], s.gravity[1], s.gravity[2])) ||
                    (rc = crHipBatchSetSceneStepSize(b, k, s.dt)) ||
                    (rc = crHipBatchSetSceneContactModel(b, k, s.contactModel)))
                    return rc;
            }
            if ((rc = crHipBatchInitialize(b)) || (rc = crHipBatchSetGraphMode(b, e->graphMode ? 1 : 0)))
                return rc;
            for (int k = 0; k < (int)bt.scenes.size(); ++k)
                if ((rc = crHipBatchSetSceneNumActiveParticles(b, k, e->scenes[bt.scenes[k]].numAllocated)))
                    return rc;
        }
        e->built = true;
        return CRMPM_HIP_OK;
    }

    int copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t st)
    {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess ? CRMPM_HIP_OK
                                                                                         : CRMPM_HIP_ERR_LAUNCH;
    }

    // SimulationFactoryImpl::syncCpuToGpuIfNeeded (simulation_factory_impl.cpp:897-940)
    int sync_tables(CrHipEngine *e)
    {
        int rc = CRMPM_HIP_OK;
        const unsigned f = e->tablesDirty;
        const size_t sc = e->shapeCap, gc = e->geomCap;
        if (f & CRMPM_HIP_ENGINE_SHAPES)
        {
            rc |= copy_h2d(e->d.type, e->h.type, sc * sizeof(int), e->stream);
            rc |= copy_h2d(e->d.geom, e->h.geom, sc * sizeof(int), e->stream);
            for (int k = 0; k < 11; ++k)
                if (k != kForce && k != kTorque)
                    rc |= copy_h2d(e->d.f4[k], e->h.f4[k], sc * sizeof(float4), e->stream);
        }
        if (f & CRMPM_HIP_ENGINE_GEOMETRIES)
        {
            rc |= copy_h2d(e->d.gtype, e->h.gtype, gc * sizeof(int), e->stream);
            rc |= copy_h2d(e->d.gparams, e->h.gparams, gc * sizeof(float4), e->stream);
        }
        if (f & CRMPM_HIP_ENGINE_SDF)
        {
            rc |= copy_h2d(e->d.sdfDims, e->h.sdfDims, (size_t)e->sdfTableCap * sizeof(int4), e->stream);
            rc |= copy_h2d(e->d.sdfLbcs, e->h.sdfLbcs, (size_t)e->sdfTableCap * sizeof(float4), e->stream);
        }
        if (f & CRMPM_HIP_ENGINE_SDF_ROWS)
            rc |= copy_h2d(e->d.sdfRows, e->h.sdfRows, (size_t)e->sdfRowCap * sizeof(float4), e->stream);
        e->tablesDirty = 0;
        return rc ? CRMPM_HIP_ERR_LAUNCH : CRMPM_HIP_OK;
    }

    // SceneImpl::beforeAdvance + syncDataIfNeeded (scene_impl.cpp:180-248)
    int sync_scene(CrHipEngine *e, CrHipEngine::Scene &s)
    {
        CrHipBatch *b = e->batches[s.batch].b;
        const size_t off = s.desc.particleOffset, n = s.desc.numMaxParticles;
        int rc = CRMPM_HIP_OK;
        if (s.shapesChanged)
        {
            if (!s.shapeIds.empty() &&
                (rc = copy_h2d(s.dShapeIds, s.shapeIds.data(), s.shapeIds.size() * sizeof(int), e->stream)))
                return rc;
            const CrHipShapeData sd = shape_view(e->d, e->shapeCap);
            const CrHipGeometryData gd{e->d.gtype, reinterpret_cast<float *>(e->d.gparams), e->geomCap};
            const CrHipGeometrySdfData sdf{reinterpret_cast<int *>(e->d.sdfDims), reinterpret_cast<float *>(e->d.sdfLbcs),
                                           reinterpret_cast<float *>(e->d.sdfRows)};
            if ((rc = crHipBatchBindSceneShapes(b, s.slot, (int)s.shapeIds.size(), s.dShapeIds, &sd, &gd, &sdf)))
                return rc;
            s.shapesChanged = false;
        }
        if (s.dirty & CRMPM_HIP_SCENE_NUM_PARTICLES)
            if ((rc = crHipBatchSetSceneNumActiveParticles(b, s.slot, s.numAllocated)))
                return rc;
        if (s.dirty & CRMPM_HIP_SCENE_POSITION_MASS)
            rc |= copy_h2d(e->dXm + off, e->hXm + off, n * sizeof(float4), e->stream);
        if (s.dirty & CRMPM_HIP_SCENE_VELOCITY)
            rc |= copy_h2d(e->dV + off, e->hV + off, n * sizeof(float4), e->stream);
        if (s.dirty & CRMPM_HIP_SCENE_PARAMS0)
            rc |= copy_h2d(e->dParams + off, e->hParams + off, n * sizeof(float4), e->stream);
        if (s.dirty & CRMPM_HIP_SCENE_TYPE)
            rc |= copy_h2d(e->dType + off, e->hType + off, n * sizeof(int), e->stream);
        if (s.dirty & CRMPM_HIP_SCENE_ACTIVE_MASK)
            rc |= copy_h2d(e->dMask + off, e->hMask + off, n, e->stream);
        if (rc)
            return CRMPM_HIP_ERR_LAUNCH;
        if ((s.dirty & CRMPM_HIP_SCENE_COMPUTE_INITIAL_DATA) && !s.initIdx.empty())
        {
            const size_t m = s.initIdx.size();
            if (m > e->dIdxCap)
            {
                if (hipStreamSynchronize(e->stream) != hipSuccess)
                    return CRMPM_HIP_ERR_LAUNCH;
                (void)hipFree(e->dIdx);
                e->dIdxCap = 0;
                if ((rc = dev_alloc(&e->dIdx, m)))
                    return rc;
                e->dIdxCap = m;
            }
            // the pageable index list is read by the copy: wait for it below
            if ((rc = copy_h2d(e->dIdx, s.initIdx.data(), m * sizeof(unsigned), e->stream)) ||
                (rc = crHipBatchComputeInitialData(b, s.slot, (unsigned)m, e->dIdx)))
                return rc;
            if (hipStreamSynchronize(e->stream) != hipSuccess)
                return CRMPM_HIP_ERR_LAUNCH;
            s.initIdx.clear();
        }
        s.dirty = 0;
        return CRMPM_HIP_OK;
    }

    // SimulationFactoryImpl::_advanceAllTaskFn (simulation_factory_impl.cpp:1106-1150)
    int advance_all(CrHipEngine *e, float dt)
    {
        int rc = CRMPM_HIP_OK;
        if (!e->built && (rc = build_batches(e)))
            return rc;
        if ((rc = sync_tables(e)))
            return rc;
        // ShapeImpl::resetCouplingForce for every shape (scene_impl.cpp:235-244)
        if (hipMemsetAsync(e->d.f4[kForce], 0, (size_t)e->shapeCap * sizeof(float4), e->stream) != hipSuccess ||
            hipMemsetAsync(e->d.f4[kTorque], 0, (size_t)e->shapeCap * sizeof(float4), e->stream) != hipSuccess)
            return CRMPM_HIP_ERR_LAUNCH;
        for (auto &s : e->scenes)
            if ((rc = sync_scene(e, s)))
                return rc;
        for (auto &bt : e->batches)
        {
            // substeps per scene: `while (timeAdvanced < dt) timeAdvanced += step()`
            int most = 0;
            for (int i : bt.scenes)
            {
                auto &s = e->scenes[i];
                float t = 0.0f;
                int n = 0;
                while (t < dt && s.dt > 0.0f)
                {
                    t += s.dt;
                    ++n;
                }
                s.substeps = n;
                most = std::max(most, n);
            }
            std::vector<int> resting;
            for (int k = 0; k < most && !rc; ++k)
            {
                for (int i : bt.scenes)
                {
                    auto &s = e->scenes[i];
                    if (!rc && s.substeps == k)
                    {
                        rc = crHipBatchSetSceneResting(bt.b, s.slot, 1);
                        resting.push_back(i);
                    }
                }
                if (!rc)
                    rc = crHipBatchStep(bt.b);
            }
            // resting scenes take part again at the next advance, even after a failed step
            for (int i : resting)
            {
                const int r = crHipBatchSetSceneResting(bt.b, e->scenes[i].slot, 0);
                rc = rc ? rc : r;
            }
            if (rc)
                return rc;
        }
        // SceneImpl::afterAdvance and afterSceneAdvance: positions,
        // velocities and coupling forces back to the host. Scenes tile
        // [0, nextOffset) back to back, so one copy per array moves every
        // scene's range (per-scene copies cost a launch each).
        const size_t all = (size_t)e->nextOffset;
        if (hipMemcpyAsync(e->hXm, e->dXm, all * sizeof(float4), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->hV, e->dV, all * sizeof(float4), hipMemcpyDeviceToHost, e->stream) != hipSuccess)
            return CRMPM_HIP_ERR_LAUNCH;
        for (int k : {kForce, kTorque})
            if (hipMemcpyAsync(e->h.f4[k], e->d.f4[k], (size_t)e->shapeCap * sizeof(float4), hipMemcpyDeviceToHost,
                               e->stream) != hipSuccess)
                return CRMPM_HIP_ERR_LAUNCH;
        return hipStreamSynchronize(e->stream) == hipSuccess ? CRMPM_HIP_OK : CRMPM_HIP_ERR_LAUNCH;
    }

    bool valid_scene(const CrHipEngine *e, int scene)
    {
        return e && scene >= 0 && scene < (int)e->scenes.size();
    }

    void free_tables(Tables &t, bool host)
    {
        auto f = [host](void *p) {
            if (p)
                (void)(host ? hipHostFree(p) : hipFree(p));
        };
        f(t.type);
        f(t.geom);
        for (auto *p : t.f4)
            f(p);
        f(t.gtype);
        f(t.gparams);
        f(t.sdfDims);
        f(t.sdfLbcs);
        f(t.sdfRows);
    }
} // namespace

extern "C" {

int crHipEngineCreate(int particleCapacity, int shapeCapacity, int geometryCapacity, int sdfTableCapacity,
                      int sdfRowCapacity, CrHipEngine **out)
{
    if (!out || particleCapacity < 1 || shapeCapacity < 1 || geometryCapacity < 1 || sdfTableCapacity < 1 ||
        sdfRowCapacity < 1)
        return CRMPM_HIP_ERR_ARG;
    *out = nullptr;
    CrHipEngine *e = new (std::nothrow) CrHipEngine();
    if (!e)
        return CRMPM_HIP_ERR_ALLOC;
    e->particleCap = particleCapacity;
    e->shapeCap = shapeCapacity;
    e->geomCap = geometryCapacity;
    e->sdfTableCap = sdfTableCapacity;
    e->sdfRowCap = sdfRowCapacity;
    const size_t n = particleCapacity;
    int rc = CRMPM_HIP_OK;
    if (hipStreamCreate(&e->stream) != hipSuccess)
        rc = CRMPM_HIP_ERR_ALLOC;
    rc = rc ? rc : host_alloc(&e->hXm, n);
    rc = rc ? rc : host_alloc(&e->hV, n);
    rc = rc ? rc : host_alloc(&e->hParams, n);
    rc = rc ? rc : host_alloc(&e->hType, n);
    rc = rc ? rc : host_alloc(&e->hMask, n);
    rc = rc ? rc : dev_alloc(&e->dXm, n);
    rc = rc ? rc : dev_alloc(&e->dV, n);
    rc = rc ? rc : dev_alloc(&e->dParams, n);
    rc = rc ? rc : dev_alloc(&e->dType, n);
    rc = rc ? rc : dev_alloc(&e->dMask, n);
    for (int hostSide = 0; hostSide < 2 && !rc; ++hostSide)
    {
        Tables &t = hostSide ? e->h : e->d;
        auto al = [&](auto **p, size_t m) { return hostSide ? host_alloc(p, m) : dev_alloc(p, m); };
        rc = rc ? rc : al(&t.type, shapeCapacity);
        rc = rc ? rc : al(&t.geom, shapeCapacity);
        for (int k = 0; k < 11 && !rc; ++k)
            rc = al(&t.f4[k], shapeCapacity);
        rc = rc ? rc : al(&t.gtype, geometryCapacity);
        rc = rc ? rc : al(&t.gparams, geometryCapacity);
        rc = rc ? rc : al(&t.sdfDims, sdfTableCapacity);
        rc = rc ? rc : al(&t.sdfLbcs, sdfTableCapacity);
        rc = rc ? rc : al(&t.sdfRows, sdfRowCapacity);
    }
    if (!rc)
    {
        // identity rotations and unit inverse scales, as a fresh shape has
        for (int i = 0; i < shapeCapacity; ++i)
        {
            e->h.f4[1][i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            e->h.f4[2][i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        }
        e->tablesDirty = CRMPM_HIP_ENGINE_SHAPES | CRMPM_HIP_ENGINE_GEOMETRIES | CRMPM_HIP_ENGINE_SDF |
                         CRMPM_HIP_ENGINE_SDF_ROWS;
        *out = e;
        return CRMPM_HIP_OK;
    }
    crHipEngineDestroy(e);
    return rc;
}

int crHipEngineDestroy(CrHipEngine *e)
{
    if (!e)
        return CRMPM_HIP_ERR_ARG;
    if (e->worker.joinable())
        e->worker.join();
    for (auto &bt : e->batches)
        crHipBatchDestroy(bt.b);
    for (auto &s : e->scenes)
        (void)hipFree(s.dShapeIds);
    for (void *p : {(void *)e->hXm, (void *)e->hV, (void *)e->hParams, (void *)e->hType, (void *)e->hMask})
        if (p)
            (void)hipHostFree(p);
    for (void *p : {(void *)e->dXm, (void *)e->dV, (void *)e->dParams, (void *)e->dType, (void *)e->dMask,
                    (void *)e->dIdx})
        if (p)
            (void)hipFree(p);
    free_tables(e->h, true);
    free_tables(e->d, false);
    if (e->stream)
        (void)hipStreamDestroy(e->stream);
    delete e;
    return CRMPM_HIP_OK;
}

int crHipEngineCreateScene(CrHipEngine *e, int kind, int numMaxParticles, float cellSize, const float boundMin[3],
                           const float boundMax[3], int *sceneId)
{
    if (!e || !boundMin || !boundMax || !sceneId || numMaxParticles < 1 || !(cellSize > 0.0f) ||
        (kind != CRMPM_HIP_MLS && kind != CRMPM_HIP_PB && kind != CRMPM_HIP_STANDARD))
        return CRMPM_HIP_ERR_ARG;
    if (e->busy || e->built)
        return CRMPM_HIP_ERR_STATE;
    if ((long long)e->nextOffset + numMaxParticles > e->particleCap)
        return CRMPM_HIP_ERR_CAPACITY;
    CrHipEngine::Scene s;
    s.kind = kind;
    s.desc.numMaxParticles = numMaxParticles;
    s.desc.particleOffset = e->nextOffset;
    s.desc.cellSize = cellSize;
    for (int k = 0; k < 3; ++k)
    {
        s.desc.boundMin[k] = boundMin[k];
        s.desc.boundMax[k] = boundMax[k];
    }
    if (dev_alloc(&s.dShapeIds, e->shapeCap))
        return CRMPM_HIP_ERR_ALLOC;
    e->nextOffset += numMaxParticles;
    e->scenes.push_back(std::move(s));
    *sceneId = (int)e->scenes.size() - 1;
    return CRMPM_HIP_OK;
}

int crHipEngineSetSceneParams(CrHipEngine *e, int scene, float dt, const float gravity[3], int contactModel,
                              int iterations)
{
    if (!valid_scene(e, scene) || !gravity || !(dt > 0.0f) || iterations < 1 ||
        (contactModel != CRMPM_HIP_CONTACT_KINEMATIC && contactModel != CRMPM_HIP_CONTACT_EFFECTIVE_MASS))
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    auto &s = e->scenes[scene];
    if (e->built)
    {
        // the batches are formed: step size, gravity and contact model go to
        // the scene's batch; the iteration count chose the batch and stays
        if (iterations != s.iterations && s.kind == CRMPM_HIP_PB)
            return CRMPM_HIP_ERR_STATE;
        CrHipBatch *b = e->batches[s.batch].b;
        int rc;
        if ((rc = crHipBatchSetSceneStepSize(b, s.slot, dt)) ||
            (rc = crHipBatchSetSceneGravity(b, s.slot, gravity[0], gravity[1], gravity[2])) ||
            (rc = crHipBatchSetSceneContactModel(b, s.slot, contactModel)))
            return rc;
    }
    s.dt = dt;
    std::copy(gravity, gravity + 3, s.gravity);
    s.contactModel = contactModel;
    s.iterations = iterations;
    return CRMPM_HIP_OK;
}

int crHipEngineGetSceneHostData(CrHipEngine *e, int scene, float **positionMass, float **velocity, float **params0,
                                int **type, unsigned char **activeMask)
{
    if (!valid_scene(e, scene))
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE; // the advance writes these arrays
    const size_t off = e->scenes[scene].desc.particleOffset;
    if (positionMass)
        *positionMass = reinterpret_cast<float *>(e->hXm + off);
    if (velocity)
        *velocity = reinterpret_cast<float *>(e->hV + off);
    if (params0)
        *params0 = reinterpret_cast<float *>(e->hParams + off);
    if (type)
        *type = e->hType + off;
    if (activeMask)
        *activeMask = e->hMask + off;
    return CRMPM_HIP_OK;
}

int crHipEngineMarkSceneDirty(CrHipEngine *e, int scene, unsigned int flags)
{
    if (!valid_scene(e, scene) || (flags & ~CRMPM_HIP_SCENE_ALL))
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    e->scenes[scene].dirty |= flags;
    return CRMPM_HIP_OK;
}

int crHipEngineSetSceneNumParticles(CrHipEngine *e, int scene, unsigned int n)
{
    if (!valid_scene(e, scene) || n > (unsigned)e->scenes[scene].desc.numMaxParticles)
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    e->scenes[scene].numAllocated = n;
    e->scenes[scene].dirty |= CRMPM_HIP_SCENE_NUM_PARTICLES;
    return CRMPM_HIP_OK;
}

int crHipEngineAddComputeInitialData(CrHipEngine *e, int scene, unsigned int n, const unsigned int *indices)
{
    if (!valid_scene(e, scene) || (n && !indices))
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    auto &s = e->scenes[scene];
    for (unsigned i = 0; i < n; ++i)
        if (indices[i] >= (unsigned)s.desc.numMaxParticles)
            return CRMPM_HIP_ERR_ARG;
    s.initIdx.insert(s.initIdx.end(), indices, indices + n);
    s.dirty |= CRMPM_HIP_SCENE_COMPUTE_INITIAL_DATA;
    return CRMPM_HIP_OK;
}

int crHipEngineSetActiveMaskRange(CrHipEngine *e, int scene, unsigned int start, unsigned int length,
                                  unsigned char value)
{
    if (!valid_scene(e, scene) || (unsigned long long)start + length > (unsigned)e->scenes[scene].desc.numMaxParticles)
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    const size_t off = e->scenes[scene].desc.particleOffset + (size_t)start;
    std::memset(e->hMask + off, value, length);
    e->scenes[scene].dirty |= CRMPM_HIP_SCENE_ACTIVE_MASK;
    return CRMPM_HIP_OK;
}

int crHipEngineSetSceneShapes(CrHipEngine *e, int scene, int numShapes, const int *shapeIds)
{
    if (!valid_scene(e, scene) || numShapes < 0 || numShapes > e->shapeCap || (numShapes && !shapeIds))
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    for (int i = 0; i < numShapes; ++i)
        if (shapeIds[i] < 0 || shapeIds[i] >= e->shapeCap)
            return CRMPM_HIP_ERR_ARG;
    auto &s = e->scenes[scene];
    s.shapeIds.assign(shapeIds, shapeIds + numShapes);
    s.shapesChanged = true;
    return CRMPM_HIP_OK;
}

int crHipEngineGetTables(CrHipEngine *e, int device, CrHipShapeData *shapes, CrHipGeometryData *geometries,
                         CrHipGeometrySdfData *sdf)
{
    if (!e)
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    const Tables &t = device ? e->d : e->h;
    if (shapes)
        *shapes = shape_view(t, e->shapeCap);
    if (geometries)
        *geometries = CrHipGeometryData{t.gtype, reinterpret_cast<float *>(t.gparams), e->geomCap};
    if (sdf)
        *sdf = CrHipGeometrySdfData{reinterpret_cast<int *>(t.sdfDims), reinterpret_cast<float *>(t.sdfLbcs),
                                    reinterpret_cast<float *>(t.sdfRows)};
    return CRMPM_HIP_OK;
}

int crHipEngineMarkTablesDirty(CrHipEngine *e, unsigned int flags)
{
    if (!e || (flags & ~(CRMPM_HIP_ENGINE_SHAPES | CRMPM_HIP_ENGINE_GEOMETRIES | CRMPM_HIP_ENGINE_SDF |
                         CRMPM_HIP_ENGINE_SDF_ROWS)))
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    e->tablesDirty |= flags;
    return CRMPM_HIP_OK;
}

int crHipEngineAdvanceAll(CrHipEngine *e, float dt)
{
    if (!e || !(dt >= 0.0f) || e->scenes.empty())
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    if (e->worker.joinable())
        e->worker.join();
    e->busy = true;
    e->workerRc = CRMPM_HIP_OK;
    try
    {
        e->worker = std::thread([e, dt] { e->workerRc = advance_all(e, dt); });
    }
    catch (...)
    {
        e->busy = false;
        return CRMPM_HIP_ERR_STATE;
    }
    return CRMPM_HIP_OK;
}

int crHipEngineFetchResultsAll(CrHipEngine *e)
{
    if (!e)
        return CRMPM_HIP_ERR_ARG;
    if (!e->busy)
        return CRMPM_HIP_ERR_STATE;
    e->worker.join();
    e->busy = false;
    return e->workerRc;
}

int crHipEngineGetSceneSubsteps(const CrHipEngine *e, int scene, int *substeps)
{
    if (!valid_scene(e, scene) || !substeps)
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    *substeps = e->scenes[scene].substeps;
    return CRMPM_HIP_OK;
}

int crHipEngineSetGraphMode(CrHipEngine *e, int enabled)
{
    if (!e)
        return CRMPM_HIP_ERR_ARG;
    if (e->busy)
        return CRMPM_HIP_ERR_STATE;
    e->graphMode = enabled != 0;
    for (auto &bt : e->batches)
    {
        const int rc = crHipBatchSetGraphMode(bt.b, enabled);
        if (rc)
            return rc;
    }
    return CRMPM_HIP_OK;
}

int crHipEngineAutotune(CrHipEngine *e, int stepsPerTrial)
{
    if (!e || stepsPerTrial < 1)
        return CRMPM_HIP_ERR_ARG;
    if (e->busy || !e->built)
        return CRMPM_HIP_ERR_STATE;
    for (auto &bt : e->batches)
    {
        const int rc = crHipBatchAutotune(bt.b, stepsPerTrial, nullptr, nullptr, nullptr, nullptr);
        if (rc)
            return rc;
    }
    return CRMPM_HIP_OK;
}

int crHipEngineGetSceneStateArrays(CrHipEngine *e, int scene, float **F, float **C, float **V0, int *stride,
                                   int *first)
{
    if (!valid_scene(e, scene) || !F || !C || !V0 || !stride || !first)
        return CRMPM_HIP_ERR_ARG;
    if (e->busy || !e->built)
        return CRMPM_HIP_ERR_STATE;
    const auto &s = e->scenes[scene];
    *first = s.desc.particleOffset;
    return crHipBatchGetStateArrays(e->batches[s.batch].b, F, C, V0, stride);
}

} // extern "C"
