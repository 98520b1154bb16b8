// RenderToy/rendertoy.h -- umbrella header of the RenderToy API on MI355X.
#ifndef RENDERTOY_AMD_RENDERTOY_H
#define RENDERTOY_AMD_RENDERTOY_H

#include "compositor.h"
#include "exception.h"
#include "exporter.h"
#include "importer.h"
#include "material.h"
#include "object.h"
#include "ray.h"
#include "renderer.h"
#include "rtmath.h"
#include "world.h"

#endif
