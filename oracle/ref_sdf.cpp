// TEST INFRASTRUCTURE -- not part of the product.
//
// C entry point over the reference's own host mesh-to-SDF builder
// (SdfBuilder::precomputeSdfDesc / computeSdf / computeSdfGradient,
// src/geometry_utils/sdf_builder.cpp:84-340), compiled from the reference
// sources by oracle/Makefile target `ref`.  tests/golden/make_golden.py uses
// it to write the SDF-table fixtures; it never travels as source.
#include "sdf_builder.h"

#include <vector>

extern "C" {

// Triangle mesh (points xyz, triangles as vertex-index triples) -> SDF table
// as the engine builds it for a triangle-mesh geometry
// (src/engine/simulation_factory_impl.cpp:602-649).  Writes dims and the
// lower bound; when dims' product fits capacity, the float4 rows (gradient,
// signed distance).  Returns the row count.
int ref_mesh_sdf(const float *points3, int numPoints, const unsigned *tris, int numTris, float cellSize, float fatten,
                 int *dims3, float *lower3, float *grid4, int capacity)
{
    using namespace crmpm;
    std::vector<Vec3f3> pts(numPoints);
    for (int i = 0; i < numPoints; ++i)
        pts[i] = Vec3f3(points3[3 * i], points3[3 * i + 1], points3[3 * i + 2]);
    std::vector<unsigned> t(tris, tris + 3 * numTris);
    TriangleMesh mesh{pts.data(), t.data(), numPoints, numTris};
    SdfDesc d;
    d.cellSize = cellSize;
    d.fatten = fatten;
    SdfBuilder::precomputeSdfDesc(d, mesh);
    dims3[0] = d.sdfDimension.x();
    dims3[1] = d.sdfDimension.y();
    dims3[2] = d.sdfDimension.z();
    lower3[0] = d.lowerBound.x();
    lower3[1] = d.lowerBound.y();
    lower3[2] = d.lowerBound.z();
    const int count = dims3[0] * dims3[1] * dims3[2];
    if (count <= capacity && grid4)
    {
        float4 *g = reinterpret_cast<float4 *>(grid4);
        SdfBuilder::computeSdf(d, mesh, g);
        SdfBuilder::computeSdfGradient(d, g);
    }
    delete d.tree;
    delete d.clusters;
    return count;
}

// The mesh-slicer table as the engine builds it
// (src/engine/simulation_factory_impl.cpp:670-690): row 0 the spine area,
// rows 1.. computeModSdf + computeSdfGradient + postProcessModSdfGradient.
// Returns the grid's node count; fills count + 1 rows when they fit capacity.
int ref_mesh_slicer_sdf(const float *points3, int numPoints, const unsigned *tris, int numTris, float cellSize,
                        float fatten, const float *spine4, int *dims3, float *lower3, float *rows4, int capacity)
{
    using namespace crmpm;
    std::vector<Vec3f3> pts(numPoints);
    for (int i = 0; i < numPoints; ++i)
        pts[i] = Vec3f3(points3[3 * i], points3[3 * i + 1], points3[3 * i + 2]);
    std::vector<unsigned> t(tris, tris + 3 * numTris);
    TriangleMesh mesh{pts.data(), t.data(), numPoints, numTris};
    SdfDesc d;
    d.cellSize = cellSize;
    d.fatten = fatten;
    d.slicerSpineArea = make_float4(spine4[0], spine4[1], spine4[2], spine4[3]);
    SdfBuilder::precomputeSdfDesc(d, mesh);
    dims3[0] = d.sdfDimension.x();
    dims3[1] = d.sdfDimension.y();
    dims3[2] = d.sdfDimension.z();
    lower3[0] = d.lowerBound.x();
    lower3[1] = d.lowerBound.y();
    lower3[2] = d.lowerBound.z();
    const int count = dims3[0] * dims3[1] * dims3[2];
    if (count + 1 <= capacity && rows4)
    {
        float4 *g = reinterpret_cast<float4 *>(rows4);
        g[0] = d.slicerSpineArea;
        Array<bool> onMesh;
        SdfBuilder::computeModSdf(d, mesh, g + 1, onMesh);
        SdfBuilder::computeSdfGradient(d, g + 1);
        SdfBuilder::postProcessModSdfGradient(d, g + 1, onMesh);
    }
    delete d.tree;
    delete d.clusters;
    return count;
}
}
