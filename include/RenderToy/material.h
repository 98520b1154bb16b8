// RenderToy/material.h -- PrincipledBSDF parameters (include/RenderToy/material.h in
// RenderToy).  Evaluation and sampling run on the GPU (rendertoy_amd/csrc/dev_bsdf.h);
// this host type carries the parameters with the same defaults and constructor.
#ifndef RENDERTOY_AMD_MATERIAL_H
#define RENDERTOY_AMD_MATERIAL_H

#include <string>

#include "rtmath.h"

namespace RenderToy {

struct PrincipledBSDF {
    std::string name = "unnamed";
    Vector3f base_color = Vector3f::White;
    Vector3f emission = Vector3f::O;
    float roughness = 0.5f;
    float metallic = 0.0f;
    float anisotropic = 0.0f;
    float subsurface = 0.0f;
    float specular_tint = 0.0f;
    float sheen = 0.0f;
    float sheen_tint = 0.5f;
    float clearcoat = 0.0f;
    float clearcoat_roughness = 0.03f;
    float spec_trans = 0.0f;
    float ior = 1.45f;
    float at_distance = 1.0f;
    Vector3f extinction = Vector3f::White;

    PrincipledBSDF() = default;
    // Argument order and defaults as RenderToy's; roughness is clamped to >= 0.001.
    PrincipledBSDF(const Vector3f &base_color_, const Vector3f &emission_ = Vector3f::O, const float roughness_ = 0.5f,
                   const float metallic_ = 0.0f, const float spec_trans_ = 0.0f, const float ior_ = 1.45f,
                   const float anisotropic_ = 0.0f, const float subsurface_ = 0.0f, const float specular_tint_ = 0.0f,
                   const float sheen_ = 0.0f, const float sheen_tint_ = 0.0f, const float clearcoat_ = 0.0f,
                   const float clearcoat_roughness_ = 0.0f, const float at_distance_ = 1.0f,
                   const Vector3f &extinction_ = Vector3f::White);
};

}  // namespace RenderToy

#endif
