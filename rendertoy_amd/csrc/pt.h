// pt.h -- wavefront path tracer (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

struct rt_scene;

namespace rtamd {
struct PtState;
void pt_destroy(PtState *p);
int pt_device_output(rt_scene *s, void **dev_ptr, size_t *bytes);
uint32_t persistent_grid(rt_scene *s, int blocks_per_cu);
void ensure_overflow(rt_scene *s, uint32_t threads);
}  // namespace rtamd
