// MpmSolverHip: the reference solver contract over the C-ABI of libcrmpm_hip.
#include "hip_mpm_solver.h"

#include <cstdio>

namespace crmpm_hip
{
    namespace
    {
        int kindOf(MpmSolverType type)
        {
            switch (type)
            {
            case MpmSolverType::eGpuMlsMpmSolver:
                return CRMPM_HIP_MLS;
            case MpmSolverType::eGpuPbMpmSolver:
                return CRMPM_HIP_PB;
            case MpmSolverType::eGpuStandardMpmSolver:
                return CRMPM_HIP_STANDARD;
            default:
                return -1;
            }
        }
    } // namespace

    MpmSolverHip::MpmSolverHip(MpmSolverType type, int numMaxParticles, float cellSize, const Bounds3 &gridBound)
    {
        const int kind = kindOf(type);
        const float lo[3] = {gridBound.minimum.x, gridBound.minimum.y, gridBound.minimum.z};
        const float hi[3] = {gridBound.maximum.x, gridBound.maximum.y, gridBound.maximum.z};
        if (kind < 0)
        {
            report(CRMPM_HIP_ERR_ARG, "MpmSolverHip: not a GPU solver type");
            return;
        }
        report(crHipSolverCreate(kind, numMaxParticles, cellSize, lo, hi, &mHandle), "crHipSolverCreate");
    }

    MpmSolverHip::~MpmSolverHip() { release(); }

    void MpmSolverHip::release()
    {
        if (mHandle)
            report(crHipSolverDestroy(mHandle), "crHipSolverDestroy");
        mHandle = nullptr;
    }

    void MpmSolverHip::report(int rc, const char *what) const
    {
        if (rc != CRMPM_HIP_OK)
            std::fprintf(stderr, "[crmpm_hip] %s failed with status %d\n", what, rc);
    }

    void MpmSolverHip::initialize() { report(crHipSolverInitialize(mHandle), "initialize"); }

    float MpmSolverHip::step()
    {
        float dt = mDt;
        report(crHipSolverStep(mHandle, &dt), "step");
        return dt;
    }

    void MpmSolverHip::fetchResults() { report(crHipSolverFetchResults(mHandle), "fetchResults"); }

    void MpmSolverHip::setGravity(const Float4 &g) { report(crHipSolverSetGravity(mHandle, g.x, g.y, g.z), "setGravity"); }

    void MpmSolverHip::setIntegrationStepSize(float dt)
    {
        mDt = dt;
        report(crHipSolverSetIntegrationStepSize(mHandle, dt), "setIntegrationStepSize");
    }

    void MpmSolverHip::setSolverIterations(int n) { report(crHipSolverSetSolverIterations(mHandle, n), "setSolverIterations"); }

    void MpmSolverHip::computeInitialData(unsigned int n, const unsigned int *indices)
    {
        report(crHipSolverComputeInitialData(mHandle, n, indices), "computeInitialData");
    }

    void MpmSolverHip::bindParticleData(ParticleData &d)
    {
        mParticleData = &d;
        report(crHipSolverBindParticleData(mHandle, &d.positionMass->x, &d.velocity->x, d.size), "bindParticleData");
    }

    void MpmSolverHip::bindParticleMaterials(ParticleMaterialData &d)
    {
        report(crHipSolverBindParticleMaterials(mHandle, &d.params0->x, reinterpret_cast<int *>(d.type), d.size),
               "bindParticleMaterials");
    }

    void MpmSolverHip::setNumActiveParticles(unsigned int n)
    {
        report(crHipSolverSetNumActiveParticles(mHandle, n), "setNumActiveParticles");
    }

    void MpmSolverHip::bindActiveMask(unsigned char *mask) { report(crHipSolverBindActiveMask(mHandle, mask), "bindActiveMask"); }

    void MpmSolverHip::setNumShapes(int n) { report(crHipSolverSetNumShapes(mHandle, n), "setNumShapes"); }

    void MpmSolverHip::bindShapeIds(int *ids) { report(crHipSolverBindShapeIds(mHandle, ids), "bindShapeIds"); }

    void MpmSolverHip::bindShapeData(const CrHipShapeData &d) { report(crHipSolverBindShapeData(mHandle, &d), "bindShapeData"); }

    void MpmSolverHip::bindGeometryData(const CrHipGeometryData &d)
    {
        report(crHipSolverBindGeometryData(mHandle, &d), "bindGeometryData");
    }

    void MpmSolverHip::bindGeometrySdfData(const CrHipGeometrySdfData &d)
    {
        report(crHipSolverBindGeometrySdfData(mHandle, &d), "bindGeometrySdfData");
    }

    void MpmSolverHip::setShapeContactModel(ShapeContactModel m)
    {
        report(crHipSolverSetShapeContactModel(mHandle, m == ShapeContactModel::eEffectiveMass ? CRMPM_HIP_CONTACT_EFFECTIVE_MASS
                                                                                                : CRMPM_HIP_CONTACT_KINEMATIC),
               "setShapeContactModel");
    }

    void MpmSolverHip::setStream(void *stream)
    {
        mStream = stream;
        report(crHipSolverSetStream(mHandle, stream), "setStream");
    }

    MpmSolverHip *createMpmSolverHip(MpmSolverType type, int numMaxParticles, float cellSize, const Bounds3 &gridBound)
    {
        MpmSolverHip *s = new MpmSolverHip(type, numMaxParticles, cellSize, gridBound);
        if (!s->valid())
        {
            delete s;
            return nullptr;
        }
        return s;
    }
} // namespace crmpm_hip
