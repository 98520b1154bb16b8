// RenderToy/exporter.h -- image writers (include/RenderToy/exporter.h in RenderToy).
#ifndef RENDERTOY_AMD_EXPORTER_H
#define RENDERTOY_AMD_EXPORTER_H

#include <ostream>

#include "compositor.h"

namespace RenderToy {

// 24-bit BMP, rows bottom-up (buffer row 0 first), each channel int(255.99*clamp(v,0,1)).
class BMPExporter {
    const Image &img;

  public:
    explicit BMPExporter(const Image &img_) : img(img_) {}
    void Export(std::ostream &os) const;
};

// Binary PPM (P6), rows top-down.
class PPMExporter {
    const Image &img;

  public:
    explicit PPMExporter(const Image &img_) : img(img_) {}
    void Export(std::ostream &os) const;
};

}  // namespace RenderToy

#endif
