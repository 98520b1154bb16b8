// RenderToy/ray.h -- Ray and RayState (include/RenderToy/ray.h in RenderToy).
#ifndef RENDERTOY_AMD_RAY_H
#define RENDERTOY_AMD_RAY_H

#include "rtmath.h"

namespace RenderToy {

struct Ray {
    Ray() = delete;
    Ray(Vector3f src_, Vector3f normalized_direction_) : src(src_), direction(normalized_direction_) {}
    Vector3f src;
    Vector3f direction;
    Ray operator-(const Ray &) const { return Ray(src, -direction); }
};

struct RayState {
    float eta = 1.0f;
    Vector3f absorption;
    Vector3f ffnormal;
};

}  // namespace RenderToy

#endif
