// orf_bvh.h -- binned-SAH BVH2 builder (host) producing the 64 B wide-box node layout.
//
// Replaces the OptiX "Trbvh"/"Sbvh" builders the reference selects by name
// (ORF/OptixSceneLoader.cpp:134-137, ORF/ObjLoader.cpp:387-402).  Degenerate
// triangles (zero or non-finite area) get no box and are never intersected,
// as mesh_bounds invalidates them (CU/triangle_mesh.cu:88-105).
#pragma once

#include <cstdint>
#include <vector>

#include "orf_vec.h"

namespace orf {

struct BvhStats {
    int nodes = 0;
    int leaves = 0;
    int max_depth = 0;
    int max_leaf = 0;
    int64_t tris = 0;        // triangles referenced by leaves
    int64_t degenerate = 0;  // skipped triangles
    double sah_cost = 0.0;   // SAH cost relative to the root area
};

struct BvhResult {
    std::vector<float4> nodes;  // 4 float4 per node
    // The same tree collapsed to 4 children per node (each BVH2 node absorbs
    // the children of its largest internal children): 8 float4 per node,
    // lo.x, hi.x, lo.y, hi.y, lo.z, hi.z of the 4 children, their links, pad.
    // The first top4 nodes are the top levels breadth-first; the rest are
    // depth-first subtrees.
    std::vector<float4> nodes4;
    int depth4 = 0;             // deepest BVH4 node level
    int top4 = 0;               // leading BVH4 nodes in breadth-first order (kBvhTopOrdered at most)
    int stack4 = 0;             // most traversal-stack entries a BVH4 ray can hold
    std::vector<float4> tris;   // 3 float4 per leaf-ordered triangle (p0|id, e1, e2)
    unsigned root_link = 0;     // leaf link when the whole scene is one leaf
    bool root_is_leaf = false;
    bool empty = true;
    BvhStats stats;
};

// SAH cost of a node visit relative to one triangle test, largest leaf, bins.
#ifndef ORF_BVH_TRAV_COST
#define ORF_BVH_TRAV_COST 1.5
#endif
#ifndef ORF_BVH_MAX_LEAF
#define ORF_BVH_MAX_LEAF 4
#endif
#ifndef ORF_BVH_BINS
#define ORF_BVH_BINS 16
#endif
constexpr double kBvhTravCost = ORF_BVH_TRAV_COST;

// verts: 3 floats per vertex; idx: 3 ints per triangle (object = world space).
BvhResult build_bvh(const float* verts, int64_t nverts, const int32_t* idx, int64_t ntris,
                    int max_leaf = ORF_BVH_MAX_LEAF, int bins = ORF_BVH_BINS);

}  // namespace orf
