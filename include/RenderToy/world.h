// RenderToy/world.h -- World (include/RenderToy/world.h in RenderToy).
#ifndef RENDERTOY_AMD_WORLD_H
#define RENDERTOY_AMD_WORLD_H

#include <vector>

#include "material.h"
#include "object.h"
#include "rtmath.h"

namespace RenderToy {

struct World {
    std::vector<Mesh *> meshes;
    std::vector<Triangle *> triangles;
    std::vector<Camera> cameras;
    std::vector<PrincipledBSDF *> materials;
    std::vector<Light *> lights;

    Vector3f sky_emission;
    Vector3f ground_reflection;

    // Collects the triangles of every mesh whose emission is not zero, meshes in
    // order, each mesh's triangles in Mesh::tris order.  Call before rendering.
    void PrepareDirectLightSampling();
    int CountEmitters() const { return (int)emissive_triangles.size(); }
    Vector3f GetDefaultEmission(const Vector3f &back_dir) const {
        return (back_dir[2] < 0.0f) ? sky_emission : ground_reflection;
    }
    const std::vector<const Triangle *> &EmissiveTriangles() const { return emissive_triangles; }

  private:
    std::vector<const Triangle *> emissive_triangles;
};

}  // namespace RenderToy

#endif
