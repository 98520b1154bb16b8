// RenderToy/exception.h -- exception types (include/RenderToy/exception.h in RenderToy).
#ifndef RENDERTOY_AMD_EXCEPTION_H
#define RENDERTOY_AMD_EXCEPTION_H

#include <stdexcept>
#include <string>

namespace RenderToy::Exception {

class IRenderToyException : public std::runtime_error {
  public:
    explicit IRenderToyException(const std::string &what) : std::runtime_error(what) {}
};
class ImageSizeNotMatchException : public IRenderToyException {
  public:
    ImageSizeNotMatchException() : IRenderToyException("image sizes do not match") {}
    explicit ImageSizeNotMatchException(const std::string &what) : IRenderToyException(what) {}
};
// A failure reported by the device library (rt_last_error()).
class DeviceException : public IRenderToyException {
  public:
    explicit DeviceException(const std::string &what) : IRenderToyException(what) {}
};

}  // namespace RenderToy::Exception

#endif
