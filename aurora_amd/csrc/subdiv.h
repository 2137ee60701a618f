// subdiv.h -- Loop subdivision and displacement on the GPU (preprocessing).
#pragma once

#include <stdint.h>

#include <vector>

#include "context.h"

namespace aur {

struct SubdivResult {
    std::vector<float> verts;           // camera space, 3 per vertex
    std::vector<int> tris;              // 3 per triangle, local indices
    std::vector<float> corner_normals;  // camera space, 9 per triangle
    std::vector<float> corner_uvs;      // 6 per triangle (zero without `vt`)
};

// Object-space indexed mesh after subdivision (parity entry point).
struct LoopMesh {
    std::vector<float> P;  // 3 per vertex
    std::vector<int> F;    // 3 per face
    std::vector<float> N;  // 3 per vertex (limit normals, or area-weighted after displacement)
    std::vector<float> UV; // 6 per face: corner uvs, when the input had them
};

// Welds the de-indexed corners by exact position, runs `levels` Loop steps,
// pushes to the limit surface, computes normals, optionally displaces.
// UV (optional): 6 per input face, the uvs of its corners.
LoopMesh loop_subdivide(Ctx *c, const std::vector<float> &P_indexed, const std::vector<int> &F, int levels,
                        const Displace &disp, const std::vector<float> *UV = nullptr);
SubdivResult subdivide_object(Ctx *c, const HostObject &o);

}  // namespace aur
