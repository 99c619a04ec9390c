// Batched execution: many queries evaluated by one launch sequence on one of a table's lanes.
//
// A lane is a HIP stream plus device buffers sized for `max_queries` queries (filter words,
// cardinalities, Mutations scratch, packed output) and pinned host buffers for the staged inputs
// and the results.  A batch is cut into lane jobs; a job is staged with one host-to-device copy,
// run as one fixed kernel sequence (captured once per lane and shape into a HIP graph and replayed
// afterwards), and brought back with one device-to-host copy.  Lanes run concurrently on their own
// streams, so jobs submitted while others are on the GPU overlap with them.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace rdb {

struct BatchResult;

// Batches of at least this many queries compile, plan and emit their filters on host threads.
constexpr uint32_t kParallelCompileQueries = 256;

struct BatchLane {
   Table* table = nullptr;
   hipStream_t stream = nullptr;
   hipEvent_t done = nullptr;
   uint32_t max_queries = 0;
   uint32_t row_capacity = 0;
   uint32_t instr_capacity = 0;
   uint32_t range_capacity = 0;
   uint32_t hist_capacity = 0;
   size_t params_bytes = 0;
   size_t out_bytes = 0;
   DeviceBuffer partial_c, octet_masks;  // octet shared scan only, allocated on first use
   DeviceBuffer params, filters, cards, lists, list_counts, partial, partial_b, heavy_sums, sparse_counts, sparse_bnd, totals, rows,
      n_rows, hist, out;
   void* host_params = nullptr;
   void* host_out = nullptr;
   // by (depth bucket, Mutations column, sparse switch, kernel knobs): what the captured launches bake in
   std::map<std::tuple<uint32_t, int, long long, uint32_t>, hipGraphExec_t> graphs;

   // the job on the lane (busy until retired)
   bool busy = false;
   uint64_t job_seq = 0;  // submission order of the job
   BatchResult* owner = nullptr;
   std::vector<uint32_t> job_queries;  // batch query index of each lane slot
   uint32_t job_total_bins = 0;
   size_t job_first_bytes = 0;
   // timing of the job's dense counting kernel (profiling only)
   hipEvent_t scan_events[2] = {nullptr, nullptr};
   bool job_timed = false;

   BatchLane() = default;
   BatchLane(const BatchLane&) = delete;
   BatchLane& operator=(const BatchLane&) = delete;
   ~BatchLane();
};

// Results of one query of a batch.
struct BatchQueryResult {
   int status = RDB_OK;
   std::string message;
   int action = 0;
   std::string column_name;
   int column = -1;
   double min_proportion = 0.0;
   std::string filter;
   std::string dims;
   uint64_t cardinality = 0;
   std::vector<MutationRowDev> rows;
   std::vector<uint32_t> agg_ids;
   std::vector<uint64_t> agg_counts;
};

struct BatchResult {
   Table* table = nullptr;
   std::vector<BatchQueryResult> queries;
   std::vector<BatchLane*> pending;  // lanes holding this batch's unretired jobs
};

// Compiles and enqueues the queries; results appear in `out` once every pending job is retired.
void batchSubmit(Table& t, BatchResult& out);
// Waits for and retires every pending job of `b`.
void batchWait(BatchResult& b);
// Waits for the pending jobs of `b` and frees their lanes without collecting results (after an
// error, or before `b` goes away with jobs still pending).
void batchDetach(BatchResult& b);
// Retires the jobs of `b` that have finished, without waiting; true when none is left.
bool batchPoll(BatchResult& b);
// Device bytes of one lane (its Mutations scratch serves every sequence column).
size_t laneBytes(const Table& t, uint32_t max_queries, uint32_t row_capacity);

namespace query {
class Emitter;
}
// Cardinalities of compiled filter programs, evaluated as one batch.
std::vector<uint64_t> countPrograms(Table& t, const std::vector<query::Emitter>& programs);

}  // namespace rdb
