// engine_internal.h — what other parts of the library (the MD step and the
// batch of systems) need from engine handles beyond the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "batch.h"
#include "common.h"
#include "forces.h"

struct dyn_engine;

namespace dyn {

// One evaluation pass over engines that share a stream: every kernel is one
// launch for all of them.  A single engine is a set of one.  The SPME chain
// (and the caller's side work, such as bonded forces) runs on a second
// stream of the same priority, forked after the preparation launch and
// joined before the finalize launch, so that it overlaps the list build,
// the pair kernel and the corrections.
struct EngineSet {
  std::vector<dyn_engine*> e;
  BatchRun R;          // main stream
  BatchRun side;       // second stream (side.stream null: everything on R)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  Profiler* prof = nullptr;
};

// Enqueue one evaluation of every engine in the set (DYN_COMPUTE_* flags,
// the same for all).  Each engine decides on the device whether it rebuilds
// its pair list (DYN_COMPUTE_REBUILD, or a list never built, forces it).
// side_work, when given, enqueues further work on the side stream (or the
// main one) after the SPME chain; it must not read the nonbonded forces.
// Internal compute flag (beside the public DYN_COMPUTE_* bits): list builds
// the device decides on go to the staging list on the second stream, beside
// the pair kernel, and become current at the next evaluation; an evaluation
// whose current list has gone stale is void (status) and its call repeats.
constexpr uint32_t kComputeStagedBuild = 1u << 30;
// The caller's next launch sums the site forces from their parts and ends
// the evaluation's list bookkeeping (engine_force_parts); no finalize launch,
// and the engine's force array stays stale until a later evaluation
// finalizes.
constexpr uint32_t kComputeFusedFinalize = 1u << 29;
void engine_set_enqueue(EngineSet& S, uint32_t flags,
                        const std::function<void(const BatchRun&)>& side_work = nullptr);
// Wait for the set's work and check each engine's device status (throws the
// first engine's error; buffers that overflowed have been enlarged).
void engine_set_check(EngineSet& S);
// The set of this engine alone (its stream, argument packs and readings).
EngineSet engine_self_set(dyn_engine* eng);

Profiler* engine_profiler(dyn_engine* eng);
hipStream_t engine_stream(const dyn_engine* eng);
// The engine's device status word (non-zero after a failed evaluation).
const unsigned int* engine_status_device(const dyn_engine* eng);
// Whether the engine's evaluations may be recorded into a graph (no test
// hook pending, no profiling).
bool engine_graph_safe(const dyn_engine* eng);
// Whether the engine asks for every kernel on one stream (no second stream).
bool engine_single_stream(const dyn_engine* eng);
void engine_set_stream(dyn_engine* eng, hipStream_t s);  // null: its own stream again
int engine_device(const dyn_engine* eng);
void engine_sizes(const dyn_engine* eng, uint32_t* n_dyn, uint32_t* n_water);
void engine_box(const dyn_engine* eng, float box[3]);
int engine_mode(const dyn_engine* eng);  // DYN_MODE_*
const int64_t* engine_forces_device(const dyn_engine* eng);
// The parts of the last evaluation's forces, and whether the engine's force
// array is stale (the evaluation was enqueued with kComputeFusedFinalize).
ForceParts engine_force_parts(const dyn_engine* eng);
bool engine_forces_stale(const dyn_engine* eng);
// Keep / put back the last evaluation's forces, SPME forces and energies
// (enqueued on the engine's stream); a restore also invalidates the list.
void engine_checkpoint(dyn_engine* eng);
void engine_restore(dyn_engine* eng);
}  // namespace dyn
