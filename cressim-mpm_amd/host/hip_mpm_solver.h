// C++ host interface of the MI355X MPM solver.
//
// MpmSolverHip mirrors the reference's solver contract member for member
// (crmpm::MpmSolver, include/solver/mpm_solver.h:63-190; MpmSolverBase,
// src/solver/mpm_solver_base.h:41-172; MpmSolverGpu, src/solver/mpm_solver_gpu.h:43-57):
// same member names, argument meanings and call order (bind*, initialize,
// computeInitialData, step, fetchResults).  It holds a C-ABI handle
// (include/crmpm_hip.h) and forwards every call; errors are logged to stderr
// and execution continues, like the reference's CR_CHECK_CUDA
// (src/utils/check_cuda.cuh:35-44).  The POD records below have the
// reference's sizes and alignments (16-byte float4 / Vec3f / Quat records).
#pragma once

#include <cstdint>

#include "crmpm_hip.h"

namespace crmpm_hip
{
    struct alignas(16) Float4
    {
        float x, y, z, w;
    };
    struct alignas(16) Int4
    {
        int x, y, z, w;
    };
    static_assert(sizeof(Float4) == 16 && alignof(Float4) == 16, "float4 record");

    struct Bounds3
    {
        Float4 minimum; // xyz used
        Float4 maximum;
    };

    // Enumerations with the reference's values (include/solver/mpm_solver.h:45-61,
    // include/solver/material_data.h:41-46, include/engine/shape.h:43-48,
    // include/engine/geometry.h:42-55).
    enum class MpmSolverType
    {
        eCpuMlsMpmSolver,
        eGpuMlsMpmSolver,
        eCpuPbMpmSolver,
        eGpuPbMpmSolver,
        eCpuStandardMpmSolver,
        eGpuStandardMpmSolver,
    };
    enum class ShapeContactModel
    {
        eKinematic,
        eEffectiveMass,
    };
    enum class ParticleMaterialType : int
    {
        eCoRotational,
        eNeoHookean,
        eLinearElastic,
    };

    // Particle SoA (include/solver/particle_data.h:44-58)
    struct ParticleData
    {
        Float4 *positionMass = nullptr;
        Float4 *velocity = nullptr;
        int size = 0;
    };
    // Material SoA (include/solver/material_data.h:50-64)
    struct ParticleMaterialData
    {
        Float4 *params0 = nullptr;
        ParticleMaterialType *type = nullptr;
        int size = 0;
    };

    class MpmSolverHip
    {
    public:
        // GpuMlsMpmSolver / GpuPbMpmSolver / GpuStandardMpmSolver(int numParticles,
        // float cellSize, Bounds3 gridBound); type must be one of the eGpu* values.
        MpmSolverHip(MpmSolverType type, int numMaxParticles, float cellSize, const Bounds3 &gridBound);
        ~MpmSolverHip();
        MpmSolverHip(const MpmSolverHip &) = delete;
        MpmSolverHip &operator=(const MpmSolverHip &) = delete;

        bool valid() const { return mHandle != nullptr; }
        CrHipSolver *handle() { return mHandle; }

        // crmpm::MpmSolver
        void initialize();
        float step();
        void fetchResults();
        void setGravity(const Float4 &gravity);
        void setIntegrationStepSize(float dt);
        void setSolverIterations(int n);
        void computeInitialData(unsigned int numParticlesToCompute, const unsigned int *indices);
        ParticleData &getParticleData() { return *mParticleData; }
        void bindParticleData(ParticleData &particleData);
        void bindParticleMaterials(ParticleMaterialData &particleMaterialData);
        void setNumActiveParticles(unsigned int numActiveParticles);
        void bindActiveMask(unsigned char *activeMask);
        void setNumShapes(int numShapes);
        void bindShapeIds(int *shapeIds);
        void bindShapeData(const CrHipShapeData &shapeData);
        void bindGeometryData(const CrHipGeometryData &geometryData);
        void bindGeometrySdfData(const CrHipGeometrySdfData &geometrySdfData);
        void setShapeContactModel(ShapeContactModel model);

        // MpmSolverGpu (the stream is a hipStream_t)
        void *getStream() const { return mStream; }
        void setStream(void *stream);

        // release() of crmpm::Releasable
        void release();

    private:
        void report(int rc, const char *what) const;

        CrHipSolver *mHandle = nullptr;
        ParticleData *mParticleData = nullptr;
        void *mStream = nullptr;
        float mDt = 0.002f;
    };

    // The createScene switch (src/engine/simulation_factory_impl.cpp:296-349)
    // for the GPU solver types; returns nullptr for the CPU types or on failure.
    MpmSolverHip *createMpmSolverHip(MpmSolverType type, int numMaxParticles, float cellSize, const Bounds3 &gridBound);
} // namespace crmpm_hip
