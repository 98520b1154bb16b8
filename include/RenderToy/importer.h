// RenderToy/importer.h -- OBJ import (include/RenderToy/importer.h in RenderToy).
#ifndef RENDERTOY_AMD_IMPORTER_H
#define RENDERTOY_AMD_IMPORTER_H

#include <string>

#include "world.h"

namespace RenderToy {

struct OBJModelImporter {
    // o / v / vn / vt / s / f records; polygons are fan-triangulated, appended
    // to World::triangles in file order and prepended to their Mesh::tris.
    static void Import(World &world, const std::string &path);
};

}  // namespace RenderToy

#endif
