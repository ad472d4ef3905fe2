// engine_internal.h — what other parts of the library (the MD step) need
// from an engine handle beyond the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct dyn_engine;

namespace dyn {
struct Profiler;
Profiler* engine_profiler(dyn_engine* eng);
hipStream_t engine_stream(const dyn_engine* eng);
int engine_device(const dyn_engine* eng);
void engine_sizes(const dyn_engine* eng, uint32_t* n_dyn, uint32_t* n_water);
void engine_box(const dyn_engine* eng, float box[3]);
}  // namespace dyn
