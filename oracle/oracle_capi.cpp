// TEST INFRASTRUCTURE -- not part of the product.
//
// C entry points over the CPU restatement in mpm_oracle.hpp, loaded with
// ctypes by tests/ and bench.py (cpu_baseline).  The entry points take the
// same inputs the reference CPU solvers bind (include/solver/mpm_solver.h:63-190):
// float4 position|mass, velocity, params0, material type, active mask, and the
// ShapeData/GeometryData/GeometrySdfData tables (include/engine/shape.h:58-89,
// include/engine/geometry.h:65-118), as flat host arrays.
#include "mpm_oracle.hpp"
#include "mesh_sdf_ref.hpp"

#include <algorithm>

namespace
{
    struct Handle
    {
        int fp64;
        orc::Solver<float> *sf = nullptr;
        orc::Solver<double> *sd = nullptr;
    };

    template <class F>
    void with(void *h, F &&f)
    {
        Handle *H = static_cast<Handle *>(h);
        if (H->fp64)
            f(*H->sd);
        else
            f(*H->sf);
    }

    template <class T>
    void setParticles(orc::Solver<T> &s, int n, const float *xm4, const float *v4, const float *params4,
                      const int *type, const unsigned char *mask)
    {
        for (int p = 0; p < n; ++p)
        {
            if (xm4)
                for (int k = 0; k < 4; ++k)
                    s.xm[4 * p + k] = T(xm4[4 * p + k]);
            if (v4)
                for (int k = 0; k < 3; ++k)
                    s.v[3 * p + k] = T(v4[4 * p + k]);
            if (params4)
                for (int k = 0; k < 4; ++k)
                    s.params[4 * p + k] = params4[4 * p + k];
            if (type)
                s.type[p] = type[p];
            if (mask)
                s.mask[p] = mask[p];
        }
    }
} // namespace

extern "C"
{
    void *orc_create(int kind, int fp64, int maxParticles, float cellSize, const float *lo, const float *hi)
    {
        Handle *h = new Handle();
        h->fp64 = fp64;
        if (fp64)
            h->sd = new orc::Solver<double>(kind, maxParticles, cellSize, lo, hi);
        else
            h->sf = new orc::Solver<float>(kind, maxParticles, cellSize, lo, hi);
        return h;
    }

    void orc_destroy(void *h)
    {
        Handle *H = static_cast<Handle *>(h);
        delete H->sf;
        delete H->sd;
        delete H;
    }

    void orc_grid_dims(void *h, int *dims)
    {
        with(h, [&](auto &s) { dims[0] = s.nx; dims[1] = s.ny; dims[2] = s.nz; });
    }

    void orc_set_gravity(void *h, float gx, float gy, float gz)
    {
        with(h, [&](auto &s) { using T = std::decay_t<decltype(s.dt)>; s.gravity = orc::V3<T>(T(gx), T(gy), T(gz)); });
    }

    void orc_set_dt(void *h, float dt)
    {
        with(h, [&](auto &s) { using T = std::decay_t<decltype(s.dt)>; s.dt = T(dt); });
    }

    void orc_set_iterations(void *h, int n)
    {
        with(h, [&](auto &s) { s.iterations = n; });
    }

    void orc_set_contact_model(void *h, int m)
    {
        with(h, [&](auto &s) { s.effectiveMass = (m == 1); });
    }

    void orc_set_particles(void *h, int n, const float *xm4, const float *v4, const float *params4,
                           const int *type, const unsigned char *mask)
    {
        with(h, [&](auto &s) { setParticles(s, n, xm4, v4, params4, type, mask); });
    }

    void orc_set_num_active(void *h, unsigned n)
    {
        with(h, [&](auto &s) { s.numActive = n; });
    }

    void orc_set_shapes(void *h, int numShapes, const int *shapeIds, int cap, const int *type, const int *geom,
                        const float *pos4, const float *rot4, const float *invScale4, const float *linVel4,
                        const float *angVel4, const float *comInvMass4, const float *inertia0, const float *inertia1,
                        const float *params0)
    {
        with(h, [&](auto &s) {
            orc::Scene &sc = s.scene;
            sc.numShapes = numShapes;
            sc.shapeIds.assign(shapeIds, shapeIds + numShapes);
            sc.shapeType.assign(type, type + cap);
            sc.shapeGeom.assign(geom, geom + cap);
            auto cp = [&](std::vector<float> &dst, const float *src) {
                if (src)
                    dst.assign(src, src + 4 * cap);
                else
                    dst.assign(4 * cap, 0.f);
            };
            cp(sc.pos, pos4);
            cp(sc.rot, rot4);
            cp(sc.invScale, invScale4);
            cp(sc.linVel, linVel4);
            cp(sc.angVel, angVel4);
            cp(sc.comInvMass, comInvMass4);
            cp(sc.inertia0, inertia0);
            cp(sc.inertia1, inertia1);
            cp(sc.params0, params0);
            if (int(sc.force.size()) != 4 * cap)
            {
                sc.force.assign(4 * cap, 0.0);
                sc.torque.assign(4 * cap, 0.0);
            }
        });
    }

    void orc_set_geometries(void *h, int n, const int *type, const float *params4)
    {
        with(h, [&](auto &s) {
            s.scene.geomType.assign(type, type + n);
            s.scene.geomParams.assign(params4, params4 + 4 * n);
        });
    }

    void orc_set_sdf(void *h, int nSdf, const int *dims4, const float *lbcs4, int nData, const float *data4)
    {
        with(h, [&](auto &s) {
            s.scene.sdfDims.assign(dims4, dims4 + 4 * nSdf);
            s.scene.sdfLbcs.assign(lbcs4, lbcs4 + 4 * nSdf);
            s.scene.sdfData.assign(data4, data4 + 4 * nData);
        });
    }

    void orc_reset_shape_force(void *h)
    {
        with(h, [&](auto &s) {
            std::fill(s.scene.force.begin(), s.scene.force.end(), 0.0);
            std::fill(s.scene.torque.begin(), s.scene.torque.end(), 0.0);
        });
    }

    void orc_get_shape_force(void *h, int cap, double *f4, double *t4)
    {
        with(h, [&](auto &s) {
            for (int i = 0; i < 4 * cap && i < int(s.scene.force.size()); ++i)
            {
                f4[i] = s.scene.force[i];
                t4[i] = s.scene.torque[i];
            }
        });
    }

    void orc_compute_initial_data(void *h, unsigned n, const unsigned *idx)
    {
        with(h, [&](auto &s) { s.computeInitialData(n, idx); });
    }

    void orc_step(void *h, int n)
    {
        with(h, [&](auto &s) {
            for (int i = 0; i < n; ++i)
                s.step();
        });
    }

    void orc_get_particles(void *h, int n, double *xm4, double *v3, double *F9, double *C9, double *V0)
    {
        with(h, [&](auto &s) {
            for (int p = 0; p < n; ++p)
            {
                if (xm4)
                    for (int k = 0; k < 4; ++k)
                        xm4[4 * p + k] = double(s.xm[4 * p + k]);
                if (v3)
                    for (int k = 0; k < 3; ++k)
                        v3[3 * p + k] = double(s.v[3 * p + k]);
                if (F9)
                    for (int k = 0; k < 9; ++k)
                        F9[9 * p + k] = double(s.F[9 * p + k]);
                if (C9)
                    for (int k = 0; k < 9; ++k)
                        C9[9 * p + k] = double(s.C[9 * p + k]);
                if (V0)
                    V0[p] = double(s.V0[p]);
            }
        });
    }

    void orc_set_deformation(void *h, int n, const float *F9, const float *C9)
    {
        with(h, [&](auto &s) {
            using T = std::decay_t<decltype(s.dt)>;
            for (int p = 0; p < n; ++p)
                for (int k = 0; k < 9; ++k)
                {
                    if (F9)
                        s.F[9 * p + k] = T(F9[9 * p + k]);
                    if (C9)
                        s.C[9 * p + k] = T(C9[9 * p + k]);
                }
        });
    }

    // Base cell of each particle, trunc((x - min) * invdx - 0.5) per axis
    // (src/solver/mls_mpm_algorithm.h:172-174).
    void orc_base_cells(void *h, int n, int *base3)
    {
        with(h, [&](auto &s) {
            for (int p = 0; p < n; ++p)
            {
                orc::V3<int> b;
                std::decay_t<decltype(s.getX(0))> fx;
                s.cellOf(s.getX(p), b, fx);
                base3[3 * p] = b.x;
                base3[3 * p + 1] = b.y;
                base3[3 * p + 2] = b.z;
            }
        });
    }

    void orc_get_grid(void *h, double *g4)
    {
        with(h, [&](auto &s) {
            for (size_t i = 0; i < s.grid.size(); ++i)
                g4[i] = double(s.grid[i]);
        });
    }

    // ParticleObjectImpl::initialize on a geometry (params0, SDF tables) with
    // bounds [bmin, bmax]; writes up to cap float4 particles, returns the count.
    unsigned orc_particle_lattice(int type, const float *p0, int numSdf, const int *sdfDims4, const float *sdfLbcs4,
                                  int sdfRows, const float *sdfData4, const float *bmin, const float *bmax,
                                  float spacing, const float *pos, const float *rot4, const float *invScale,
                                  float mass, float *out4, unsigned cap)
    {
        orc::Scene sc;
        sc.sdfDims.assign(sdfDims4, sdfDims4 + 4 * numSdf);
        sc.sdfLbcs.assign(sdfLbcs4, sdfLbcs4 + 4 * numSdf);
        sc.sdfData.assign(sdfData4, sdfData4 + 4 * sdfRows);
        std::vector<float> out;
        const unsigned n = orc::particle_lattice(sc, type, p0, bmin, bmax, spacing, pos, rot4, invScale, mass, out);
        std::copy(out.begin(), out.begin() + 4 * std::min(n, cap), out4);
        return n;
    }

    // Mesh -> SDF table (src/engine/simulation_factory_impl.cpp:628-635):
    // writes dims and lower; fills grid4 when the row count fits capacity.
    // -1 for a vertex index beyond numPoints.
    int orc_mesh_sdf(const float *pts, int numPoints, const unsigned *tris, int numTris, float cell, float fatten,
                     int *dims3, float *lower3, float *grid4, int capacity)
    {
        for (int t = 0; t < 3 * numTris; ++t)
            if (tris[t] >= static_cast<unsigned>(numPoints))
                return -1;
        const int n = orc::mesh_sdf_desc_ref(pts, tris, numTris, cell, fatten, dims3, lower3);
        if (grid4 && n <= capacity)
            orc::mesh_sdf_ref(pts, tris, numTris, cell, dims3, lower3, grid4);
        return n;
    }

    // The same table with the exact winding number for the sign and the
    // closest triangle in index order for the distance (the builders'
    // optional exact rule).
    int orc_mesh_sdf_exact(const float *pts, int numPoints, const unsigned *tris, int numTris, float cell,
                           float fatten, int *dims3, float *lower3, float *grid4, int capacity)
    {
        for (int t = 0; t < 3 * numTris; ++t)
            if (tris[t] >= static_cast<unsigned>(numPoints))
                return -1;
        const int n = orc::mesh_sdf_desc_ref(pts, tris, numTris, cell, fatten, dims3, lower3);
        if (grid4 && n <= capacity)
            orc::mesh_sdf(pts, tris, numTris, cell, dims3, lower3, grid4);
        return n;
    }

    // Mesh-slicer table (src/engine/simulation_factory_impl.cpp:670-690):
    // dims and lower as above; count + 1 rows (spine area, then the grid)
    // when they fit capacity.  Returns the grid's node count, -1 for a
    // vertex index beyond numPoints.
    int orc_mesh_slicer_sdf(const float *pts, int numPoints, const unsigned *tris, int numTris, float cell,
                            float fatten, const float *spine4, int *dims3, float *lower3, float *rows4, int capacity)
    {
        for (int t = 0; t < 3 * numTris; ++t)
            if (tris[t] >= static_cast<unsigned>(numPoints))
                return -1;
        const int n = orc::mesh_sdf_desc_ref(pts, tris, numTris, cell, fatten, dims3, lower3);
        if (rows4 && n + 1 <= capacity)
            orc::mesh_slicer_sdf_ref(pts, tris, numTris, cell, dims3, lower3, spine4, rows4);
        return n;
    }

    // The mesh-slicer table with the exact rule: a node is off the surface
    // when its closest triangle (first in index order) puts the closest point
    // on a boundary edge or vertex.
    int orc_mesh_slicer_sdf_exact(const float *pts, int numPoints, const unsigned *tris, int numTris, float cell,
                                  float fatten, const float *spine4, int *dims3, float *lower3, float *rows4,
                                  int capacity)
    {
        for (int t = 0; t < 3 * numTris; ++t)
            if (tris[t] >= static_cast<unsigned>(numPoints))
                return -1;
        const int n = orc::mesh_sdf_desc_ref(pts, tris, numTris, cell, fatten, dims3, lower3);
        if (rows4 && n + 1 <= capacity)
            orc::mesh_slicer_sdf(pts, tris, numTris, cell, dims3, lower3, spine4, rows4);
        return n;
    }

    // svd3 on n column-major 3x3 matrices (include/math/svd3.h:1265-1275)
    void orc_svd3(int fp64, int n, const float *A9, double *U9, double *Vt9, double *S3)
    {
        auto run = [&](auto tag) {
            using T = decltype(tag);
            for (int i = 0; i < n; ++i)
            {
                orc::M3<T> A, u, vt;
                orc::V3<T> s;
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r)
                        A(r, c) = T(A9[9 * i + 3 * c + r]);
                orc::svd3(A, u, vt, s);
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r)
                    {
                        U9[9 * i + 3 * c + r] = double(u(r, c));
                        Vt9[9 * i + 3 * c + r] = double(vt(r, c));
                    }
                for (int k = 0; k < 3; ++k)
                    S3[3 * i + k] = double(s[k]);
            }
        };
        if (fp64)
            run(double());
        else
            run(float());
    }

    // computeStress / solveMaterialConstraints on n particles, column-major
    void orc_stress(int n, const float *params4, const int *type, const float *F9, double *out9)
    {
        for (int i = 0; i < n; ++i)
        {
            orc::M3<float> F;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    F(r, c) = F9[9 * i + 3 * c + r];
            orc::M3<float> s = orc::compute_stress<float>(&params4[4 * i], type[i], F);
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    out9[9 * i + 3 * c + r] = s(r, c);
        }
    }

    void orc_pb_constraint(int n, float dt, const float *params4, const int *type, const float *F9, const float *C9, double *out9)
    {
        for (int i = 0; i < n; ++i)
        {
            orc::M3<float> F, C;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                {
                    F(r, c) = F9[9 * i + 3 * c + r];
                    C(r, c) = C9[9 * i + 3 * c + r];
                }
            orc::pb_constraint<float>(dt, &params4[4 * i], type[i], F, C);
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r)
                    out9[9 * i + 3 * c + r] = C(r, c);
        }
    }
}
