// TEST INFRASTRUCTURE -- not part of the product; built only where the
// reference sources are present (oracle/Makefile, target `ref`).
//
// Drives the reference's own CPU solvers (src/solver/cpu_{mls,pb,standard}_mpm_solver.cpp,
// compiled from the reference tree) through their MpmSolverBase API
// (src/solver/mpm_solver_base.h:41-172) and exposes the same C entry points as
// oracle_capi.cpp under the prefix ref_, so tests can run one scenario through
// both and compare bit for bit.  The solvers keep F, C and V0 in private
// members; the harness reads them by compiling the solver headers with
// `private` mapped to `public` (standard headers are included first).
#include <algorithm>
#include <cassert>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <functional>
#include <mutex>
#include <new>
#include <unordered_map>
#include <utility>
#include <vector>

#define private public
#include "cpu_mls_mpm_solver.h"
#include "cpu_pb_mpm_solver.h"
#include "cpu_standard_mpm_solver.h"
#undef private
#include "svd3.h"
#include "constitutive_models.h"

using namespace crmpm;

namespace
{
    struct RefHandle
    {
        int kind;
        MpmSolverBase *solver = nullptr;
        int maxParticles;
        Vec3i dims;
        // engine-owned particle arrays
        std::vector<float4> xm;
        std::vector<Vec3f> v;
        std::vector<float4> params;
        std::vector<ParticleMaterialType> type;
        std::vector<unsigned char> mask;
        ParticleData *pdata = nullptr;
        ParticleMaterialData *mdata = nullptr;
        // shapes
        std::vector<int> shapeIds;
        std::vector<ShapeType> shapeType;
        std::vector<int> shapeGeom;
        std::vector<Vec3f> pos, invScale, linVel, angVel;
        std::vector<Quat> rot;
        std::vector<float4> comInvMass, inertia0, inertia1, force, torque, sparams;
        std::vector<GeometryType> geomType;
        std::vector<float4> geomParams;
        std::vector<Vec3i> sdfDims;
        std::vector<float4> sdfLbcs, sdfData;
        ShapeData shapeData;
        GeometryData geomData;
        GeometrySdfData sdfDataS;
    };

    Vec3f f4v(const float *p) { return Vec3f(p[0], p[1], p[2]); }
    float4 f4(const float *p) { return make_float4(p[0], p[1], p[2], p[3]); }

    void rebind(RefHandle *h)
    {
        h->shapeData.type = h->shapeType.data();
        h->shapeData.geometryIdx = h->shapeGeom.data();
        h->shapeData.position = h->pos.data();
        h->shapeData.rotation = h->rot.data();
        h->shapeData.invScale = h->invScale.data();
        h->shapeData.linearVelocity = h->linVel.data();
        h->shapeData.angularVelocity = h->angVel.data();
        h->shapeData.comInvMass = h->comInvMass.data();
        h->shapeData.inertiaInv0 = h->inertia0.data();
        h->shapeData.inertiaInv1 = h->inertia1.data();
        h->shapeData.force = h->force.data();
        h->shapeData.torque = h->torque.data();
        h->shapeData.params0 = h->sparams.data();
        h->shapeData.size = int(h->shapeType.size());
        h->geomData.type = h->geomType.data();
        h->geomData.params0 = h->geomParams.data();
        h->geomData.size = int(h->geomType.size());
        h->sdfDataS.dimemsion = h->sdfDims.data();
        h->sdfDataS.lowerBoundCellSize = h->sdfLbcs.data();
        h->sdfDataS.gradientSignedDistance = h->sdfData.data();
        h->solver->bindShapeData(h->shapeData);
        h->solver->bindGeometryData(h->geomData);
        h->solver->bindGeometrySdfData(h->sdfDataS);
        h->solver->bindShapeIds(h->shapeIds.empty() ? nullptr : h->shapeIds.data());
        h->solver->setNumShapes(int(h->shapeIds.size()));
    }

    template <class P>
    void readParticles(P *parts, int n, double *F9, double *C9, double *V0)
    {
        for (int p = 0; p < n; ++p)
        {
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                {
                    if (F9)
                        F9[9 * p + 3 * c + r] = parts->gradientDeformationTensor[p](r, c);
                    if (C9)
                        C9[9 * p + 3 * c + r] = parts->affineMomentum[p](r, c);
                }
            if (V0)
                V0[p] = parts->initialVolume[p];
        }
    }
} // namespace

extern "C"
{
    void *ref_create(int kind, int fp64, int maxParticles, float cellSize, const float *lo, const float *hi)
    {
        if (fp64)
            return nullptr;
        RefHandle *h = new RefHandle();
        h->kind = kind;
        h->maxParticles = maxParticles;
        Bounds3 b(Vec3f(lo[0], lo[1], lo[2]), Vec3f(hi[0], hi[1], hi[2]));
        if (kind == 0)
            h->solver = new CpuMlsMpmSolver(float(maxParticles), cellSize, b);
        else if (kind == 1)
            h->solver = new CpuPbMpmSolver(float(maxParticles), cellSize, b);
        else
            h->solver = new CpuStandardMpmSolver(float(maxParticles), cellSize, b);
        h->xm.assign(maxParticles, make_float4(0, 0, 0, 0));
        h->v.assign(maxParticles, Vec3f());
        h->params.assign(maxParticles, make_float4(0, 0, 0, 0));
        h->type.assign(maxParticles, ParticleMaterialType::eCoRotational);
        h->mask.assign(maxParticles, 0);
        h->pdata = new ParticleData(maxParticles);
        h->pdata->positionMass = h->xm.data();
        h->pdata->velocity = h->v.data();
        h->mdata = new ParticleMaterialData(maxParticles);
        h->mdata->params0 = h->params.data();
        h->mdata->type = h->type.data();
        h->solver->bindParticleData(*h->pdata);
        h->solver->bindParticleMaterials(*h->mdata);
        h->solver->bindActiveMask(h->mask.data());
        h->geomType.assign(1, GeometryType::eBox);
        h->geomParams.assign(1, make_float4(0, 0, 0, 0));
        h->sdfDims.assign(1, Vec3i());
        h->sdfLbcs.assign(1, make_float4(0, 0, 0, 0));
        h->sdfData.assign(1, make_float4(0, 0, 0, 0));
        rebind(h);
        h->solver->initialize();
        Vec3f gridSize = b.maximum - b.minimum;
        h->dims = (gridSize / cellSize + 1.0f).cast<int>();
        return h;
    }

    void ref_destroy(void *hp)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        // The reference solvers do not free their buffers in their destructors; leaked deliberately.
        delete h->pdata;
        delete h->mdata;
        delete h;
    }

    void ref_grid_dims(void *hp, int *dims)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        dims[0] = h->dims.x();
        dims[1] = h->dims.y();
        dims[2] = h->dims.z();
    }

    void ref_set_gravity(void *hp, float gx, float gy, float gz) { static_cast<RefHandle *>(hp)->solver->setGravity(Vec3f(gx, gy, gz)); }
    void ref_set_dt(void *hp, float dt) { static_cast<RefHandle *>(hp)->solver->setIntegrationStepSize(dt); }
    void ref_set_iterations(void *hp, int n) { static_cast<RefHandle *>(hp)->solver->setSolverIterations(n); }
    void ref_set_contact_model(void *hp, int m)
    {
        static_cast<RefHandle *>(hp)->solver->setShapeContactModel(m == 1 ? ShapeContactModel::eEffectiveMass : ShapeContactModel::eKinematic);
    }

    void ref_set_particles(void *hp, int n, const float *xm4, const float *v4, const float *params4, const int *type,
                           const unsigned char *mask)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        for (int p = 0; p < n; ++p)
        {
            if (xm4)
                h->xm[p] = f4(&xm4[4 * p]);
            if (v4)
                h->v[p] = f4v(&v4[4 * p]);
            if (params4)
                h->params[p] = f4(&params4[4 * p]);
            if (type)
                h->type[p] = ParticleMaterialType(type[p]);
            if (mask)
                h->mask[p] = mask[p];
        }
    }

    void ref_set_num_active(void *hp, unsigned n) { static_cast<RefHandle *>(hp)->solver->setNumActiveParticles(n); }

    void ref_set_shapes(void *hp, int numShapes, const int *shapeIds, int cap, const int *type, const int *geom,
                        const float *pos4, const float *rot4, const float *invScale4, const float *linVel4,
                        const float *angVel4, const float *comInvMass4, const float *inertia0, const float *inertia1,
                        const float *params0)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        h->shapeIds.assign(shapeIds, shapeIds + numShapes);
        h->shapeType.resize(cap);
        h->shapeGeom.resize(cap);
        h->pos.resize(cap);
        h->rot.resize(cap);
        h->invScale.resize(cap);
        h->linVel.resize(cap);
        h->angVel.resize(cap);
        h->comInvMass.resize(cap);
        h->inertia0.resize(cap);
        h->inertia1.resize(cap);
        h->sparams.resize(cap);
        if (int(h->force.size()) != cap)
        {
            h->force.assign(cap, make_float4(0, 0, 0, 0));
            h->torque.assign(cap, make_float4(0, 0, 0, 0));
        }
        const float zero4[4] = {0, 0, 0, 0};
        auto at = [&](const float *a, int i) { return a ? &a[4 * i] : zero4; };
        for (int i = 0; i < cap; ++i)
        {
            h->shapeType[i] = ShapeType(type[i]);
            h->shapeGeom[i] = geom[i];
            h->pos[i] = f4v(at(pos4, i));
            const float *q = at(rot4, i);
            h->rot[i] = Quat(q[0], q[1], q[2], q[3]);
            h->invScale[i] = f4v(at(invScale4, i));
            h->linVel[i] = f4v(at(linVel4, i));
            h->angVel[i] = f4v(at(angVel4, i));
            h->comInvMass[i] = f4(at(comInvMass4, i));
            h->inertia0[i] = f4(at(inertia0, i));
            h->inertia1[i] = f4(at(inertia1, i));
            h->sparams[i] = f4(at(params0, i));
        }
        rebind(h);
    }

    void ref_set_geometries(void *hp, int n, const int *type, const float *params4)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        h->geomType.resize(n);
        h->geomParams.resize(n);
        for (int i = 0; i < n; ++i)
        {
            h->geomType[i] = GeometryType(type[i]);
            h->geomParams[i] = f4(&params4[4 * i]);
        }
        rebind(h);
    }

    void ref_set_sdf(void *hp, int nSdf, const int *dims4, const float *lbcs4, int nData, const float *data4)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        h->sdfDims.resize(nSdf);
        h->sdfLbcs.resize(nSdf);
        for (int i = 0; i < nSdf; ++i)
        {
            h->sdfDims[i] = Vec3i(dims4[4 * i], dims4[4 * i + 1], dims4[4 * i + 2]);
            h->sdfLbcs[i] = f4(&lbcs4[4 * i]);
        }
        h->sdfData.resize(nData);
        for (int i = 0; i < nData; ++i)
            h->sdfData[i] = f4(&data4[4 * i]);
        rebind(h);
    }

    void ref_reset_shape_force(void *hp)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        std::fill(h->force.begin(), h->force.end(), make_float4(0, 0, 0, 0));
        std::fill(h->torque.begin(), h->torque.end(), make_float4(0, 0, 0, 0));
    }

    void ref_get_shape_force(void *hp, int cap, double *f4o, double *t4o)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        for (int i = 0; i < cap && i < int(h->force.size()); ++i)
        {
            const float *f = &h->force[i].x;
            const float *t = &h->torque[i].x;
            for (int k = 0; k < 4; ++k)
            {
                f4o[4 * i + k] = f[k];
                t4o[4 * i + k] = t[k];
            }
        }
    }

    void ref_compute_initial_data(void *hp, unsigned n, const unsigned *idx)
    {
        static_cast<RefHandle *>(hp)->solver->computeInitialData(n, idx);
    }

    void ref_step(void *hp, int n)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        for (int i = 0; i < n; ++i)
            h->solver->step();
    }

    void ref_get_particles(void *hp, int n, double *xm4, double *v3, double *F9, double *C9, double *V0)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        for (int p = 0; p < n; ++p)
        {
            if (xm4)
            {
                const float *x = &h->xm[p].x;
                for (int k = 0; k < 4; ++k)
                    xm4[4 * p + k] = x[k];
            }
            if (v3)
                for (int k = 0; k < 3; ++k)
                    v3[3 * p + k] = h->v[p][k];
        }
        if (h->kind == 0)
            readParticles(static_cast<CpuMlsMpmSolver *>(h->solver)->mParticles, n, F9, C9, V0);
        else if (h->kind == 1)
            readParticles(static_cast<CpuPbMpmSolver *>(h->solver)->mParticles, n, F9, C9, V0);
        else
        {
            auto *parts = static_cast<CpuStandardMpmSolver *>(h->solver)->mParticles;
            for (int p = 0; p < n; ++p)
            {
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r)
                    {
                        if (F9)
                            F9[9 * p + 3 * c + r] = parts->gradientDeformationTensor[p](r, c);
                        if (C9)
                            C9[9 * p + 3 * c + r] = 0.0;
                    }
                if (V0)
                    V0[p] = parts->initialVolume[p];
            }
        }
    }

    void ref_set_deformation(void *hp, int n, const float *F9, const float *C9)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        auto put = [&](auto *parts, bool hasC) {
            for (int p = 0; p < n; ++p)
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r)
                    {
                        if (F9)
                            parts->gradientDeformationTensor[p](r, c) = F9[9 * p + 3 * c + r];
                        if (C9 && hasC)
                            parts->affineMomentum[p](r, c) = C9[9 * p + 3 * c + r];
                    }
        };
        if (h->kind == 0)
            put(static_cast<CpuMlsMpmSolver *>(h->solver)->mParticles, true);
        else if (h->kind == 1)
            put(static_cast<CpuPbMpmSolver *>(h->solver)->mParticles, true);
    }

    void ref_get_grid(void *hp, double *g4)
    {
        RefHandle *h = static_cast<RefHandle *>(hp);
        int nn = h->dims.x() * h->dims.y() * h->dims.z();
        float4 *g = nullptr;
        if (h->kind == 0)
            g = static_cast<CpuMlsMpmSolver *>(h->solver)->mGrid->mMomentumVelocityMass;
        else if (h->kind == 1)
            g = static_cast<CpuPbMpmSolver *>(h->solver)->mGrid->mMomentumVelocityMass;
        else
            g = static_cast<CpuStandardMpmSolver *>(h->solver)->mGrid->mMomentumVelocityMass;
        for (int i = 0; i < nn; ++i)
        {
            g4[4 * i] = g[i].x;
            g4[4 * i + 1] = g[i].y;
            g4[4 * i + 2] = g[i].z;
            g4[4 * i + 3] = g[i].w;
        }
    }

    void ref_svd3(int fp64, int n, const float *A9, double *U9, double *Vt9, double *S3)
    {
        (void)fp64;
        for (int i = 0; i < n; ++i)
        {
            Mat3x3f A, u, vt, s;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    A(r, c) = A9[9 * i + 3 * c + r];
            svd3(A, u, vt, s);
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                {
                    U9[9 * i + 3 * c + r] = u(r, c);
                    Vt9[9 * i + 3 * c + r] = vt(r, c);
                }
            S3[3 * i] = s(0, 0);
            S3[3 * i + 1] = s(1, 1);
            S3[3 * i + 2] = s(2, 2);
        }
    }

    void ref_stress(int n, const float *params4, const int *type, const float *F9, double *out9)
    {
        for (int i = 0; i < n; ++i)
        {
            Mat3x3f F, s;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    F(r, c) = F9[9 * i + 3 * c + r];
            computeStress(f4(&params4[4 * i]), ParticleMaterialType(type[i]), F, Mat3x3f(), 0.f, 0.f, s);
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    out9[9 * i + 3 * c + r] = s(r, c);
        }
    }

    void ref_pb_constraint(int n, float dt, const float *params4, const int *type, const float *F9, const float *C9, double *out9)
    {
        for (int i = 0; i < n; ++i)
        {
            Mat3x3f F, C;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                {
                    F(r, c) = F9[9 * i + 3 * c + r];
                    C(r, c) = C9[9 * i + 3 * c + r];
                }
            solveMaterialConstraints(dt, f4(&params4[4 * i]), ParticleMaterialType(type[i]), F, C);
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    out9[9 * i + 3 * c + r] = C(r, c);
        }
    }
}
