// engine_internal.h — what other parts of the library (the MD step and the
// batch of systems) need from engine handles beyond the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "batch.h"
#include "common.h"

struct dyn_engine;

namespace dyn {

// Displacement readings since the last list build of a set of engines, one
// per system (squared Å as the bit pattern of a non-negative float): a
// device array the displacement kernel maximises into, a ring of pinned
// host copies of asynchronous evaluations (read once their event has
// completed) and a copy for synchronous ones.
struct DispRing {
  static constexpr int kRing = 4;
  // an asynchronous evaluation decides with every reading at least kLag
  // evaluations old in hand (the host waits for them)
  static constexpr int kLag = 2;
  // rebuild when a site has moved factor × buffer (a repeat after a
  // flagged evaluation lowers it)
  float factor = 0.25f;
  uint32_t n = 0;
  DevBuf<unsigned int> bits;
  PinnedBuf<unsigned int> h;       // kRing × n
  PinnedBuf<unsigned int> h_sync;  // n
  hipEvent_t ev[kRing] = {};
  uint64_t gen[kRing] = {};
  uint64_t eval_of[kRing] = {};    // the asynchronous evaluation of each reading
  uint64_t issued = 0;             // asynchronous readings recorded
  bool pending[kRing] = {};
  int next = 0;
  uint64_t builds = 0;             // list builds of the set (reading generation)
  std::vector<float> seen;         // largest reading per system of the current lists
  void init(uint32_t systems);
  DispRing() = default;
  DispRing(const DispRing&) = delete;
  DispRing& operator=(const DispRing&) = delete;
  ~DispRing();
};

// One evaluation pass over engines that share a stream: every kernel is one
// launch for all of them.  A single engine is a set of one.
struct EngineSet {
  std::vector<dyn_engine*> e;
  BatchRun R;
  Profiler* prof = nullptr;
  DispRing* disp = nullptr;
};

// Enqueue one evaluation of every engine in the set (DYN_COMPUTE_* flags,
// the same for all); a list rebuild, when one engine needs it, rebuilds all.
void engine_set_enqueue(EngineSet& S, uint32_t flags);
// Wait for the set's work and check each engine's device status (throws the
// first engine's error; buffers that overflowed have been enlarged).
void engine_set_check(EngineSet& S);
// The set of this engine alone (its stream, argument packs and readings).
EngineSet engine_self_set(dyn_engine* eng);

Profiler* engine_profiler(dyn_engine* eng);
hipStream_t engine_stream(const dyn_engine* eng);
void engine_set_stream(dyn_engine* eng, hipStream_t s);  // null: its own stream again
int engine_device(const dyn_engine* eng);
void engine_sizes(const dyn_engine* eng, uint32_t* n_dyn, uint32_t* n_water);
void engine_box(const dyn_engine* eng, float box[3]);
int engine_mode(const dyn_engine* eng);  // DYN_MODE_*
const int64_t* engine_forces_device(const dyn_engine* eng);
// Keep / put back the last evaluation's forces, SPME forces and energies
// (enqueued on the engine's stream); a restore also invalidates the list.
void engine_checkpoint(dyn_engine* eng);
void engine_restore(dyn_engine* eng);
}  // namespace dyn
