// TEST INFRASTRUCTURE (the CPU oracle): restatement of the reference's host
// SDF builder for triangle meshes, statement for statement, so that the
// oracle's tables equal the reference's bit for bit:
//   * buildTree / buildAABBTree / flattenTree (include/geometry_utils/winding_number.h:148-161,
//     src/geometry_utils/aabb_tree.cpp:97-333): one primitive per leaf, nodes split at
//     the centre of the box along the axis of largest centre variance, a 50-50 split
//     when all centres fall on one side, children allocated in pairs in the order the
//     depth-first build (positive child first) subdivides them;
//   * precomputeClusterInformation / approximateCluster (src/geometry_utils/winding_number.cpp:44-124)
//     and computeWindingNumberApprox with beta = 2 (winding_number.h:128-133,
//     winding_number.cpp:136-171): the sign of a node;
//   * the closest-point tree walks with the warm start from the previous node's
//     triangle (SdfBuilder::computeSdf / computeModSdf, src/geometry_utils/sdf_builder.cpp:105-240;
//     ClosestDistance(Projection)ToTrimeshTraversalController, aabb_tree.cpp:473-515, 724-779):
//     the distance, and for the mesh slicer the on-surface flag, which a walk
//     clears for good once it accepts (d <= best so far) a candidate whose closest
//     point lies on a boundary edge or boundary vertex.
// A triangle's box starts as Bounds3's zero-initialised box and grows to its
// corners, so every box, and the table's bounds, include the origin.
// Nothing in the product includes this file.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace orc
{
    namespace ref_sdf
    {
        using F3 = V3<float>;

        inline F3 vmin(F3 a, F3 b) { return F3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
        inline F3 vmax(F3 a, F3 b) { return F3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
        inline F3 cross(F3 a, F3 b)
        {
            return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
        }

        struct Box
        {
            F3 lo, hi;
        };

        struct Node
        {
            Box bv;
            unsigned data; // leaf: triangle << 5 | count << 1 | 1; internal: positive child << 1
            bool leaf() const { return data & 1u; }
            unsigned prim() const { return data >> 5; }
            unsigned pos() const { return data >> 1; }
            unsigned neg() const { return (data >> 1) + 1; }
        };

        struct Cluster
        {
            float radius = 0.0f, area = 0.0f;
            F3 centroid, normalSum;
        };

        struct Tree
        {
            std::vector<Node> nodes;
            std::vector<Cluster> clusters; // per node (internal nodes only)
        };

        // buildTree: the triangles' boxes (zero-initialised, grown to the three
        // corners, fattened by 1e-4), then buildAABBTree with a limit of one
        // primitive per leaf.
        inline Tree build(const float *pts, const unsigned *tris, int numTris)
        {
            auto P = [&](unsigned i) { return F3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); };
            std::vector<Box> boxes(numTris);
            std::vector<F3> centre(numTris);
            for (int t = 0; t < numTris; ++t)
            {
                Box b{F3(0.0f, 0.0f, 0.0f), F3(0.0f, 0.0f, 0.0f)};
                for (int j = 0; j < 3; ++j)
                {
                    const F3 p = P(tris[3 * t + j]);
                    b.lo = vmin(b.lo, p);
                    b.hi = vmax(b.hi, p);
                }
                const float e = 1e-4f;
                b.lo = F3(b.lo.x - e, b.lo.y - e, b.lo.z - e);
                b.hi = F3(b.hi.x + e, b.hi.y + e, b.hi.z + e);
                boxes[t] = b;
                centre[t] = (b.hi + b.lo) * 0.5f;
            }
            struct BuildNode
            {
                Box bv;
                int pos = -1; // pool index of the positive child, -1 for a leaf
                unsigned first = 0, count = 0;
            };
            std::vector<unsigned> idx(numTris);
            for (int i = 0; i < numTris; ++i)
                idx[i] = i;
            std::vector<BuildNode> pool;
            pool.reserve(2 * (size_t)numTris);
            pool.push_back(BuildNode{});
            pool[0].count = numTris;
            std::vector<int> stack{0};
            while (!stack.empty())
            {
                const int ni = stack.back();
                stack.pop_back();
                // AABBTreeBuildNode::subdivide
                unsigned *prims = idx.data() + pool[ni].first;
                const unsigned nb = pool[ni].count;
                Box bv = boxes[prims[0]];
                F3 mean = centre[prims[0]];
                for (unsigned i = 1; i < nb; ++i)
                {
                    mean = mean + centre[prims[i]];
                    bv.lo = vmin(bv.lo, boxes[prims[i]].lo);
                    bv.hi = vmax(bv.hi, boxes[prims[i]].hi);
                }
                pool[ni].bv = bv;
                mean = mean * (1.0f / float(nb));
                if (nb > 1)
                {
                    F3 var(0.0f, 0.0f, 0.0f);
                    for (unsigned i = 0; i < nb; ++i)
                    {
                        F3 c = centre[prims[i]] - mean;
                        c = c * c;
                        var = var + c;
                    }
                    var = var * (1.0f / float(nb - 1));
                    const float ax = std::fabs(var.x), ay = std::fabs(var.y), az = std::fabs(var.z);
                    const int axis = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
                    const float split = (bv.lo[axis] + bv.hi[axis]) * 0.5f;
                    unsigned nPos = 0;
                    for (unsigned i = 0; i < nb; ++i)
                    {
                        const unsigned p = prims[i];
                        if (centre[p][axis] > split)
                        {
                            prims[i] = prims[nPos];
                            prims[nPos] = p;
                            ++nPos;
                        }
                    }
                    if (nPos == 0 || nPos == nb)
                        nPos = nb >> 1;
                    const int c0 = (int)pool.size();
                    pool.push_back(BuildNode{});
                    pool.push_back(BuildNode{});
                    pool[ni].pos = c0;
                    pool[c0].first = pool[ni].first;
                    pool[c0].count = nPos;
                    pool[c0 + 1].first = pool[ni].first + nPos;
                    pool[c0 + 1].count = nb - nPos;
                    stack.push_back(c0 + 1);
                    stack.push_back(c0);
                }
            }
            Tree tree;
            tree.nodes.resize(pool.size());
            for (size_t i = 0; i < pool.size(); ++i)
            {
                tree.nodes[i].bv = pool[i].bv;
                tree.nodes[i].data = pool[i].pos < 0 ? (idx[pool[i].first] << 5) | ((pool[i].count & 15u) << 1) | 1u
                                                     : (unsigned)pool[i].pos << 1;
            }
            // precomputeClusterInformation: per-triangle area, area-weighted
            // normal and centroid; each internal node's triangles in the order
            // of a walk that takes the negative child first
            std::vector<float> area(numTris);
            std::vector<F3> nrmA(numTris), cen(numTris);
            for (int t = 0; t < numTris; ++t)
            {
                const F3 a = P(tris[3 * t]), b = P(tris[3 * t + 1]), c = P(tris[3 * t + 2]);
                nrmA[t] = cross(b - a, c - a) * 0.5f;
                area[t] = norm(nrmA[t]);
                cen[t] = (a + b + c) * float(1.0 / 3.0);
            }
            tree.clusters.resize(pool.size());
            std::vector<int> todo{0};
            std::vector<int> order;
            std::vector<std::pair<int, int>> sections;
            while (!todo.empty())
            {
                int id = todo.back();
                todo.pop_back();
                if (id >= 0)
                {
                    const Node &n = tree.nodes[id];
                    if (n.leaf())
                    {
                        order.push_back((int)n.prim());
                        sections.push_back({(int)order.size() - 1, (int)order.size()});
                        continue;
                    }
                    todo.push_back(-id - 1);
                    todo.push_back((int)n.pos());
                    todo.push_back((int)n.pos() + 1);
                }
                else
                {
                    const std::pair<int, int> a = sections.back();
                    sections.pop_back();
                    const std::pair<int, int> b = sections.back();
                    sections.pop_back();
                    const std::pair<int, int> sum{b.first, a.second};
                    id = -id - 1;
                    // approximateCluster
                    F3 wc(0.0f, 0.0f, 0.0f), wn(0.0f, 0.0f, 0.0f);
                    float as = 0.0f;
                    for (int i = sum.first; i < sum.second; ++i)
                    {
                        const int t = order[i];
                        as += area[t];
                        wc = wc + cen[t] * area[t];
                        wn = wn + nrmA[t];
                    }
                    wc = wc / as;
                    float r2 = 0.0f;
                    for (int i = sum.first; i < sum.second; ++i)
                    {
                        const unsigned *tv = tris + 3 * order[i];
                        for (int j = 0; j < 3; ++j)
                        {
                            const float d2 = sqn(wc - P(tv[j]));
                            if (d2 > r2)
                                r2 = d2;
                        }
                    }
                    Cluster &c = tree.clusters[id];
                    c.radius = std::sqrt(r2);
                    c.area = as;
                    c.centroid = wc;
                    c.normalSum = wn;
                    sections.push_back(sum);
                }
            }
            return tree;
        }

        // traverseBVH (include/geometry_utils/aabb_tree.h:245-274): visit(node,
        // index) returns 0 (do not go deeper), 1 (positive child first) or 2
        // (negative child first).
        template <class Visit> inline void traverse(const Tree &tree, Visit &&visit)
        {
            std::vector<unsigned> todo;
            unsigned index = 0;
            while (true)
            {
                const Node &n = tree.nodes[index];
                const int c = visit(n, index);
                if (!n.leaf() && c != 0)
                {
                    if (c == 2)
                    {
                        todo.push_back(n.pos());
                        index = n.neg();
                    }
                    else
                    {
                        todo.push_back(n.neg());
                        index = n.pos();
                    }
                    continue;
                }
                if (todo.empty())
                    break;
                index = todo.back();
                todo.pop_back();
            }
        }

        // computeWindingNumberApprox with beta = 2: exact terms at the
        // leaves, a cluster's first-order term (evaluated in double, as its
        // double constant promotes it) once the query point is farther than
        // beta radii from its centroid
        inline float winding_approx(const Tree &tree, const float *pts, const unsigned *tris, F3 q, float beta = 2.0f)
        {
            auto P = [&](unsigned i) { return F3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); };
            float w = 0.0f;
            traverse(tree, [&](const Node &n, unsigned index) {
                if (n.leaf())
                {
                    const unsigned *tv = tris + 3 * n.prim();
                    w += winding_term(P(tv[0]), P(tv[1]), P(tv[2]), q);
                    return 0;
                }
                const Cluster &c = tree.clusters[index];
                const float d2 = sqn(q - c.centroid);
                const float th = beta * c.radius;
                if (d2 > th * th)
                {
                    const F3 dir = c.centroid - q;
                    const float l = norm(dir);
                    w += (float)(((0.25 / 3.141592653589793238462643383) / (l * l * l)) * dot(c.normalSum, dir));
                    return 0;
                }
                return 1;
            });
            return w;
        }

        inline float box_dist2(const Box &b, F3 q) { return sqn(vmax(b.lo, vmin(b.hi, q)) - q); }

        // Per-triangle boundary data of the mesh slicer (buildEdgeTriangleMap
        // / buildEdgeVerticesList, src/geometry_utils/sdf_builder.cpp:45-82):
        // an edge is a boundary edge when one triangle holds it, a vertex a
        // boundary vertex when it ends a boundary edge.
        struct Boundary
        {
            std::vector<unsigned char> edge;   // per triangle, bits ab, bc, ca
            std::vector<unsigned char> vertex; // per triangle, bits a, b, c
        };
        inline Boundary boundary(const unsigned *tris, int numTris)
        {
            auto key = [](unsigned u, unsigned v) { return (uint64_t)std::min(u, v) << 32 | std::max(u, v); };
            std::unordered_map<uint64_t, int> count;
            for (int t = 0; t < numTris; ++t)
                for (int j = 0; j < 3; ++j)
                    ++count[key(tris[3 * t + j], tris[3 * t + (j + 1) % 3])];
            std::unordered_map<unsigned, bool> bverts;
            for (const auto &e : count)
                if (e.second < 2)
                {
                    bverts[(unsigned)(e.first >> 32)] = true;
                    bverts[(unsigned)(e.first & 0xffffffffu)] = true;
                }
            Boundary b;
            b.edge.assign(numTris, 0);
            b.vertex.assign(numTris, 0);
            for (int t = 0; t < numTris; ++t)
                for (int j = 0; j < 3; ++j)
                {
                    if (count[key(tris[3 * t + j], tris[3 * t + (j + 1) % 3])] < 2)
                        b.edge[t] |= (unsigned char)(1u << j);
                    if (bverts.count(tris[3 * t + j]))
                        b.vertex[t] |= (unsigned char)(1u << j);
                }
            return b;
        }

        // The closest triangle of q by the reference's walk: start from the
        // warm triangle's distance (none when warm < 0), accept a leaf when
        // d < best (project = false) or d <= best (project = true).  With
        // project, onSurface turns false for good when an accepted candidate's
        // closest point lies on a boundary edge or vertex, and side becomes +-1
        // once a candidate is accepted (distanceProjectionPointTriangleSquared,
        // aabb_tree.cpp:519-556; the warm start sets no side, so a node whose
        // walk accepts nothing keeps side 0 and distance 0).
        struct Walk
        {
            float best = FLT_MAX;
            int tri = -1;
            F3 cp;
            bool onSurface = true;
            int side = 0;
        };
        inline Walk closest(const Tree &tree, const float *pts, const unsigned *tris, F3 q, int warm, bool project,
                            const Boundary *bnd)
        {
            auto P = [&](unsigned i) { return F3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); };
            Walk w;
            if (warm >= 0)
            {
                const unsigned *tv = tris + 3 * warm;
                w.best = tri_closest_unit(q, P(tv[0]), P(tv[1]), P(tv[2]), w.cp);
                w.tri = warm;
            }
            traverse(tree, [&](const Node &n, unsigned) {
                if (box_dist2(n.bv, q) >= w.best)
                    return 0;
                if (n.leaf())
                {
                    const int t = (int)n.prim();
                    const unsigned *tv = tris + 3 * t;
                    F3 cp;
                    int region;
                    const F3 a = P(tv[0]), b = P(tv[1]), c = P(tv[2]);
                    const float d = tri_closest_unit(q, a, b, c, cp, &region);
                    if (project ? d <= w.best : d < w.best)
                    {
                        w.best = d;
                        w.tri = t;
                        w.cp = cp;
                        if (project)
                        {
                            bool outside = false;
                            switch (region)
                            {
                            case kRegA: outside = bnd->vertex[t] & 1; break;
                            case kRegB: outside = bnd->vertex[t] & 2; break;
                            case kRegC: outside = bnd->vertex[t] & 4; break;
                            case kRegAB: outside = bnd->edge[t] & 1; break;
                            case kRegBC: outside = bnd->edge[t] & 2; break;
                            case kRegAC: outside = bnd->edge[t] & 4; break;
                            default: break;
                            }
                            if (outside)
                                w.onSurface = false;
                            // the candidate's side of its plane, +1 or -1; only
                            // its magnitude reaches the table (|distance * side|)
                            w.side = 1;
                        }
                    }
                    return 0;
                }
                const float dp = box_dist2(tree.nodes[n.pos()].bv, q), dn = box_dist2(tree.nodes[n.neg()].bv, q);
                if (dp < dn)
                    return dp < w.best ? 1 : 0;
                return dn < w.best ? 2 : 0;
            });
            return w;
        }
    } // namespace ref_sdf

    // SdfBuilder::precomputeSdfDesc with the reference's tree bounds: the root
    // box (the union of the triangles' origin-including boxes) grown by fatten.
    inline int mesh_sdf_desc_ref(const float *pts, const unsigned *tris, int numTris, float cell, float fatten,
                                 int *dims, float *lower)
    {
        using ref_sdf::F3;
        F3 lo(FLT_MAX, FLT_MAX, FLT_MAX), hi(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        for (int t = 0; t < numTris; ++t)
        {
            F3 bl(0.0f, 0.0f, 0.0f), bh(0.0f, 0.0f, 0.0f);
            for (int j = 0; j < 3; ++j)
            {
                const unsigned i = tris[3 * t + j];
                const F3 p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
                bl = ref_sdf::vmin(bl, p);
                bh = ref_sdf::vmax(bh, p);
            }
            lo = ref_sdf::vmin(lo, F3(bl.x - 1e-4f, bl.y - 1e-4f, bl.z - 1e-4f));
            hi = ref_sdf::vmax(hi, F3(bh.x + 1e-4f, bh.y + 1e-4f, bh.z + 1e-4f));
        }
        long count = 1;
        for (int k = 0; k < 3; ++k)
        {
            lower[k] = lo[k] - fatten;
            const float size = (hi[k] + fatten) - lower[k];
            dims[k] = static_cast<int>(size / cell + 1.0f);
            count *= dims[k];
        }
        return static_cast<int>(count);
    }

    // SdfBuilder::computeSdf + computeSdfGradient: distance from the warm-started
    // walk, sign from the beta = 2 winding approximation.
    inline void mesh_sdf_ref(const float *pts, const unsigned *tris, int numTris, float cell, const int *dims,
                             const float *lower, float *grid4)
    {
        const ref_sdf::Tree tree = ref_sdf::build(pts, tris, numTris);
        const int dx = dims[0], dy = dims[1], dz = dims[2];
        int last = -1;
        for (int z = 0; z < dz; ++z)
            for (int y = 0; y < dy; ++y)
                for (int x = 0; x < dx; ++x)
                {
                    const V3<float> q(lower[0] + float(x) * cell, lower[1] + float(y) * cell, lower[2] + float(z) * cell);
                    const ref_sdf::Walk w = ref_sdf::closest(tree, pts, tris, q, last, false, nullptr);
                    last = w.tri;
                    const float dist = norm(w.cp - q);
                    const float wn = ref_sdf::winding_approx(tree, pts, tris, q);
                    grid4[4 * ((z * dy + y) * dx + x) + 3] = dist * (wn > 0.5f ? -1.0f : 1.0f);
                }
        sdf_gradient(dims, cell, grid4);
    }

    // SdfBuilder::computeModSdf + computeSdfGradient + postProcessModSdfGradient:
    // |distance * side| from the warm-started projection walk (0 when no
    // candidate past the warm start was accepted), rows off the surface NaN
    // unless on the positive side of the spine plane.
    inline void mesh_slicer_sdf_ref(const float *pts, const unsigned *tris, int numTris, float cell, const int *dims,
                                    const float *lower, const float *spine, float *rows4)
    {
        const ref_sdf::Tree tree = ref_sdf::build(pts, tris, numTris);
        const ref_sdf::Boundary bnd = ref_sdf::boundary(tris, numTris);
        const int dx = dims[0], dy = dims[1], dz = dims[2];
        const bool hasSpine = !std::isnan(spine[0]);
        V3<float> sn(spine[0], spine[1], spine[2]);
        {
            const float s2 = sqn(sn);
            sn = s2 > 0.0f ? sn * rsq(s2) : V3<float>();
        }
        for (int k = 0; k < 4; ++k)
            rows4[k] = spine[k];
        float *grid4 = rows4 + 4;
        std::vector<char> onMesh((size_t)dx * dy * dz);
        int last = -1;
        for (int z = 0; z < dz; ++z)
            for (int y = 0; y < dy; ++y)
                for (int x = 0; x < dx; ++x)
                {
                    const V3<float> q(lower[0] + float(x) * cell, lower[1] + float(y) * cell, lower[2] + float(z) * cell);
                    const ref_sdf::Walk w = ref_sdf::closest(tree, pts, tris, q, last, true, &bnd);
                    last = w.tri;
                    const size_t node = ((size_t)z * dy + y) * dx + x;
                    grid4[4 * node + 3] = std::fabs(norm(w.cp - q) * float(w.side));
                    bool on = w.onSurface;
                    if (hasSpine && dot(q, sn) + spine[3] > 0.0f)
                        on = true;
                    onMesh[node] = on;
                }
        sdf_gradient(dims, cell, grid4);
        for (size_t i = 0; i < onMesh.size(); ++i)
            if (!onMesh[i])
                for (int k = 0; k < 4; ++k)
                    grid4[4 * i + k] = std::nanf("0");
    }
} // namespace orc
