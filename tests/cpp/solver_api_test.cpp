// GPU test of the C++ solver interface (MpmSolverHip) -- the call sequence of
// the reference engine for one scene (src/engine/scene_impl.cpp:42-93,180-263):
// bind particle/material/mask/shape data, initialize, computeInitialData,
// step, fetchResults.  A falling elastic cube (MLS, co-rotational) hits a
// static ground plane; the result is checked against the CPU oracle
// (oracle/liboracle.so) fed the same inputs.  Exit status 0 on success.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hip_mpm_solver.h"

extern "C"
{
    void *orc_create(int kind, int fp64, int maxParticles, float cellSize, const float *lo, const float *hi);
    void orc_destroy(void *h);
    void orc_set_gravity(void *h, float gx, float gy, float gz);
    void orc_set_dt(void *h, float dt);
    void orc_set_particles(void *h, int n, const float *xm4, const float *v4, const float *params4, const int *type,
                           const unsigned char *mask);
    void orc_set_num_active(void *h, unsigned n);
    void orc_set_shapes(void *h, int numShapes, const int *shapeIds, int cap, const int *type, const int *geom,
                        const float *pos4, const float *rot4, const float *invScale4, const float *linVel4,
                        const float *angVel4, const float *comInvMass4, const float *inertia0, const float *inertia1,
                        const float *params0);
    void orc_set_geometries(void *h, int n, const int *type, const float *params4);
    void orc_compute_initial_data(void *h, unsigned n, const unsigned *idx);
    void orc_step(void *h, int n);
    void orc_get_particles(void *h, int n, double *xm4, double *v3, double *F9, double *C9, double *V0);
}

#define HIPCHECK(x)                                                              \
    do                                                                           \
    {                                                                            \
        if ((x) != hipSuccess)                                                   \
        {                                                                        \
            std::fprintf(stderr, "HIP error at %s:%d\n", __FILE__, __LINE__);   \
            return 2;                                                            \
        }                                                                        \
    } while (0)

template <class T>
static T *toDevice(const std::vector<T> &h)
{
    T *d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess)
        return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return nullptr;
    return d;
}

int main()
{
    using namespace crmpm_hip;
    const int side = 10, N = side * side * side, steps = 50;
    const float dx = 1.0f / 32, h = dx / 2, E = 1e4f, nu = 0.3f;
    const float lambda = (E * nu) / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    std::vector<Float4> xm(N), v(N), params(N);
    std::vector<int> type(N, 0); // co-rotational
    std::vector<unsigned char> mask(N, 1);
    std::vector<unsigned> idx(N);
    for (int k = 0, p = 0; k < side; ++k)
        for (int j = 0; j < side; ++j)
            for (int i = 0; i < side; ++i, ++p)
            {
                xm[p] = Float4{0.4f + (i + 0.5f) * h, 0.15f + (j + 0.5f) * h, 0.4f + (k + 0.5f) * h, 1000.0f * h * h * h};
                v[p] = Float4{0.1f, -0.5f, 0.0f, 0.0f};
                params[p] = Float4{lambda, mu, 0.0f, 0.0f};
                idx[p] = p;
            }
    // one static plane at y = 0.1 (shape 0, geometry 0)
    std::vector<int> shapeType{0}, shapeGeom{0}, shapeIds{0}, geomType{2};
    std::vector<Float4> pos{{0, 0.1f, 0, 0}}, rot{{0, 0, 0, 1}}, invScale{{1, 1, 1, 0}}, zero{{0, 0, 0, 0}},
        sparams{{0.0f, 1.0f, 0.2f, 0.0f}}, gparams{{0, 0, 0, 0}};

    Float4 *dXm = toDevice(xm), *dV = toDevice(v), *dParams = toDevice(params);
    int *dType = toDevice(type), *dShapeType = toDevice(shapeType), *dShapeGeom = toDevice(shapeGeom),
        *dShapeIds = toDevice(shapeIds), *dGeomType = toDevice(geomType);
    unsigned char *dMask = toDevice(mask);
    unsigned *dIdx = toDevice(idx);
    Float4 *dPos = toDevice(pos), *dRot = toDevice(rot), *dInvScale = toDevice(invScale), *dZero = toDevice(zero),
           *dSparams = toDevice(sparams), *dGparams = toDevice(gparams), *dForce = toDevice(zero), *dTorque = toDevice(zero);
    Int4 *dSdfDims = toDevice(std::vector<Int4>{Int4{0, 0, 0, 0}});
    Float4 *dSdfLb = toDevice(zero), *dSdf = toDevice(zero);
    if (!dXm || !dV || !dParams || !dType || !dMask || !dIdx || !dSdf)
        return 2;
    hipStream_t stream;
    HIPCHECK(hipStreamCreate(&stream));

    Bounds3 bounds{{0, 0, 0, 0}, {1, 1, 1, 0}};
    MpmSolverHip *solver = createMpmSolverHip(MpmSolverType::eGpuMlsMpmSolver, N, dx, bounds);
    if (!solver)
        return 3;
    ParticleData pd;
    pd.positionMass = dXm;
    pd.velocity = dV;
    pd.size = N;
    ParticleMaterialData md;
    md.params0 = dParams;
    md.type = reinterpret_cast<ParticleMaterialType *>(dType);
    md.size = N;
    CrHipShapeData sd{};
    sd.type = dShapeType;
    sd.geometryIdx = dShapeGeom;
    sd.position = &dPos->x;
    sd.rotation = &dRot->x;
    sd.invScale = &dInvScale->x;
    sd.linearVelocity = &dZero->x;
    sd.angularVelocity = &dZero->x;
    sd.comInvMass = &dZero->x;
    sd.inertiaInv0 = &dZero->x;
    sd.inertiaInv1 = &dZero->x;
    sd.force = &dForce->x;
    sd.torque = &dTorque->x;
    sd.params0 = &dSparams->x;
    sd.size = 1;
    CrHipGeometryData gd{dGeomType, &dGparams->x, 1};
    CrHipGeometrySdfData sdf{&dSdfDims->x, &dSdfLb->x, &dSdf->x};

    solver->setStream(stream);
    solver->bindShapeData(sd);
    solver->bindGeometryData(gd);
    solver->bindGeometrySdfData(sdf);
    solver->setGravity(Float4{0, -9.81f, 0, 0});
    solver->setIntegrationStepSize(1e-4f);
    solver->setShapeContactModel(ShapeContactModel::eKinematic);
    solver->bindParticleData(pd);
    solver->bindParticleMaterials(md);
    solver->bindActiveMask(dMask);
    solver->setNumShapes(1);
    solver->bindShapeIds(dShapeIds);
    solver->initialize();
    solver->setNumActiveParticles(N);
    solver->computeInitialData(N, dIdx);
    float t = 0.0f;
    for (int s = 0; s < steps; ++s)
        t += solver->step();
    solver->fetchResults();
    std::vector<Float4> xOut(N), vOut(N);
    HIPCHECK(hipMemcpy(xOut.data(), dXm, N * sizeof(Float4), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(vOut.data(), dV, N * sizeof(Float4), hipMemcpyDeviceToHost));
    solver->release();
    delete solver;

    // the same scene on the CPU oracle
    const float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    void *o = orc_create(0, 0, N, dx, lo, hi);
    orc_set_gravity(o, 0, -9.81f, 0);
    orc_set_dt(o, 1e-4f);
    orc_set_particles(o, N, &xm[0].x, &v[0].x, &params[0].x, type.data(), mask.data());
    orc_set_num_active(o, N);
    orc_set_geometries(o, 1, geomType.data(), &gparams[0].x);
    orc_set_shapes(o, 1, shapeIds.data(), 1, shapeType.data(), shapeGeom.data(), &pos[0].x, &rot[0].x, &invScale[0].x,
                   &zero[0].x, &zero[0].x, &zero[0].x, &zero[0].x, &zero[0].x, &sparams[0].x);
    orc_compute_initial_data(o, N, idx.data());
    orc_step(o, steps);
    std::vector<double> ox(4 * N), ov(3 * N);
    orc_get_particles(o, N, ox.data(), ov.data(), nullptr, nullptr, nullptr);
    orc_destroy(o);

    double ex = 0, ev = 0, sx = 0, sv = 0;
    for (int p = 0; p < N; ++p)
    {
        const float *gx = &xOut[p].x, *gv = &vOut[p].x;
        for (int k = 0; k < 3; ++k)
        {
            ex = std::fmax(ex, std::fabs(gx[k] - ox[4 * p + k]));
            ev = std::fmax(ev, std::fabs(gv[k] - ov[3 * p + k]));
            sx = std::fmax(sx, std::fabs(ox[4 * p + k]));
            sv = std::fmax(sv, std::fabs(ov[3 * p + k]));
        }
    }
    const double rx = ex / sx, rv = ev / sv;
    std::printf("solver_api_test: %d substeps (t = %.4f s), relative error x %.3e v %.3e\n", steps, t, rx, rv);
    return (rx <= 1e-4 && rv <= 1e-4 && t > 0.0f) ? 0 : 1;
}
