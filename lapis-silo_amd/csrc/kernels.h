// Host-side launch wrappers for the HIP kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_abi.h"

namespace rdb {

struct GroupDims {
   uint32_t n;
   const uint32_t* labels[4];  // per-row labels, padded row indexing
   uint32_t stride[4];
};

hipError_t launchFill(uint32_t* p, uint32_t value, size_t n, hipStream_t stream);

size_t filterLdsBytes(uint32_t depth);

// Evaluates n_queries programs over every (chunk, 4096-row slice) of the table.  Output words of
// query q start at out + q * n_chunks * 1024; cards[q] and chunk_cards[q * n_chunks + c] must be
// zero on entry and receive the popcounts.
hipError_t launchFilter(const TableView& tv, const ProgramDesc* progs, const Instr* instrs, const uint32_t* ranges,
                        uint32_t n_queries, uint32_t depth, uint64_t* out, unsigned long long* cards,
                        uint32_t* chunk_cards, uint32_t max_blocks, hipStream_t stream);

// Mutations for one filter: per-slot AND-popcounts (one partial sum per chunk group, n_groups <=
// kMaxGroups, `partial` holds n_groups * n_slots u32), boundary prefix sums, local-reference
// accumulation and thresholding.  counts_out ([n_symbols][L]) and rows may be null.
hipError_t launchMutations(const SeqColumnView* col_dev, const SeqColumnView& col_host, uint32_t n_chunks,
                           uint32_t n_groups, const uint64_t* filter, const uint32_t* filter_chunk_cards,
                           const unsigned long long* filter_card, uint32_t* partial, int64_t* totals,
                           double min_proportion, uint32_t* counts_out, MutationRowDev* rows, uint32_t row_capacity,
                           uint32_t* n_rows, hipStream_t stream, hipEvent_t count_start = nullptr,
                           hipEvent_t count_stop = nullptr);

hipError_t launchSymbolLabels(const SeqColumnView* col_dev, const uint32_t* chunk_sizes, uint32_t position,
                              uint32_t n_chunks, uint32_t block, uint32_t slot_lo, uint32_t slot_hi, uint32_t* labels,
                              hipStream_t stream);

hipError_t launchGroupHistogram(uint32_t n_chunks, const uint64_t* filter, const uint32_t* filter_chunk_cards,
                                const GroupDims& dims, uint32_t n_bins, uint32_t* hist, hipStream_t stream);

}  // namespace rdb
