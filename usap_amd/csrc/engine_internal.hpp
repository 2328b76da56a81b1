// engine_internal.hpp -- engine accessors used by the host driver inside the
// library (not part of the public C-ABI).
#pragma once
#include <string>
#include <vector>

#include "usap.h"

namespace usap {
// Host copy of the engine's graph (refreshed from the device after edge
// appends); null if that copy failed.
const usap_graph* engine_host_graph(usap_engine* e);
// Per-phase records of the running / last partition (usap_partition_log).
std::vector<usap_phase>& engine_phase_log(usap_engine* e);
// usap_nodal_sweep; with S null the sweep starts from the entropy of the
// current state, computed on the device (no host wait for it).
int nodal_sweep(usap_engine* e, int32_t batch, uint64_t stream0, const double* S, double mean,
                double thr, int32_t check, usap_sweep* out);
const usap_params* engine_params(const usap_engine* e);
int32_t engine_rebuilds(const usap_engine* e);
// Rows moved to a larger slab between batches since the engine was created.
int64_t engine_grows(usap_engine* e);
// Cumulative growth counters: rows grown or compacted, compactions, failed
// insertions, pool exhaustions.
void engine_grow_debug(usap_engine* e, long long out[4]);
void engine_set_error(usap_engine* e, const char* msg);
// Error of a call that has no engine handle to hold it (batch, serve): the
// first one recorded since the last clear; usap_last_error(NULL) returns it.
void set_detached_error(const std::string& msg);
// A host worker's device context: one HIP stream and one stream-ordered
// memory pool that the engines the worker creates one after another share
// (batched graphs), so that no stream or pool is created or destroyed per
// graph and no freed block moves between workers' streams.  Contexts live as
// long as the process: worker_ctx_destroy synchronises the stream, trims the
// pool and returns the context to a free list that worker_ctx_create takes
// from first.
struct WorkerCtx;
int worker_ctx_create(int32_t device, WorkerCtx** out);
void worker_ctx_destroy(WorkerCtx* w);
// usap_create on a worker context (null: own stream, the device's pool).
int create_in(WorkerCtx* ctx, const usap_graph* g, const usap_params* p, usap_engine** out);
void clear_detached_error();
const char* detached_error();
}  // namespace usap
