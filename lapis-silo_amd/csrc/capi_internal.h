// Helpers shared by the C-ABI translation units: exceptions become status codes at the boundary.
#pragma once

#include <string>

#include "engine.h"

namespace rdb::capi {

inline thread_local std::string g_error;

template <typename F>
int guarded(Context* ctx, F&& f) {
   try {
      f();
      return RDB_OK;
   } catch (const Error& e) {
      g_error = e.what();
      if (ctx) ctx->error = e.what();
      return e.code;
   } catch (const std::bad_alloc&) {
      g_error = "host allocation failed";
      if (ctx) ctx->error = g_error;
      return RDB_ENOMEM;
   } catch (const std::exception& e) {
      g_error = e.what();
      if (ctx) ctx->error = e.what();
      return RDB_EINVAL;
   }
}

inline Table& tableOf(rdb_table* t) {
   if (!t) fail(RDB_EINVAL, "null table");
   return t->t;
}

inline void requireFinalized(const Table& t) {
   if (!t.finalized) fail(RDB_EINVAL, "table is not finalized (call rdb_table_finalize)");
}

// Grouping dimensions of the Aggregated action: dictionary columns or `column@position` of a
// sequence column.
struct AggDim {
   bool is_seq = false;
   const DictColumn* dict = nullptr;
   const SeqColumn* seq = nullptr;
   int seq_index = -1;
   uint32_t position = 0;
   uint32_t n_labels = 0;  // including the null label
};

// Parses "dim,dim,..." (at most four); RDB_EQUERY for unknown columns or bad positions.
std::vector<AggDim> parseDims(const Table& t, const char* dims);

// Strides of the combined label (last dimension fastest); the number of bins.
uint64_t dimStrides(const std::vector<AggDim>& ds, uint32_t* stride);

// Non-empty bins of a histogram as (ids, count) in BitmapAggregationNode's order
// (bitmap_aggregation_node.cpp:91-130, 226-253): per dimension by value string (symbol order for a
// sequence position) with the null group (UINT32_MAX) last, depth first.
void decodeHistogram(const std::vector<AggDim>& ds, const uint32_t* stride, const uint32_t* hist, uint64_t bins,
                     std::vector<uint32_t>& ids, std::vector<uint64_t>& counts);
// The same from (bin, count) pairs in any order.
void decodeBins(const std::vector<AggDim>& ds, const uint32_t* stride, const std::vector<std::pair<uint32_t, uint32_t>>& bins,
                std::vector<uint32_t>& ids, std::vector<uint64_t>& counts);
// The Aggregated action under a device filter (single-query path).
void aggregateRows(Table& t, const DeviceBitmap& f, const char* dims, std::vector<uint32_t>& ids,
                   std::vector<uint64_t>& counts);

// Mutations of column `idx` under a device filter, rows sorted in the reference's output order.
std::vector<MutationRowDev> mutationRows(Table& t, const DeviceBitmap& f, int idx, double min_proportion);
void sortMutationRows(std::vector<MutationRowDev>& rows);
// Rows into the caller's arrays (position 1-based, from = original reference, proportion in double).
void emitMutationRows(const SeqColumn& col, const std::vector<MutationRowDev>& rows, uint32_t* position, uint8_t* from,
                      uint8_t* to, uint32_t* count, uint32_t* coverage, double* proportion, uint64_t capacity,
                      uint64_t* n);
// Column index of a Mutations query; RDB_EQUERY for an unknown column or a bad minProportion.
int mutationColumn(const Table& t, const char* column, double min_proportion);

// Mutations rows (1-based positions, symbols of the column's alphabet) for the result sinks.
struct MutationRowsView {
   const Alphabet* alphabet = nullptr;
   const char* sequence_name = "";
   uint32_t fields = 0;  // rdb_mutation_field bits
   const uint32_t* position = nullptr;
   const uint8_t* from = nullptr;
   const uint8_t* to = nullptr;
   const uint32_t* count = nullptr;
   const uint32_t* coverage = nullptr;
   const double* proportion = nullptr;
   uint64_t n = 0;
};
uint32_t validFields(uint32_t fields);
// RDB_EINVAL when a row's from / to symbol lies outside the alphabet.
void checkSymbols(const MutationRowsView& rows);
// NDJSON lines as the reference's NdjsonSink writes them (exec_node/ndjson_sink.cpp).
std::string mutationRowsNdjson(const MutationRowsView& rows);
// An Arrow IPC stream (schema, one record batch, end of stream) as arrow::ipc::MakeStreamWriter
// writes the reference's ArrowIpcSink output (exec_node/arrow_ipc_sink.cpp).
std::vector<uint8_t> mutationRowsArrowIpc(const MutationRowsView& rows);

}  // namespace rdb::capi
