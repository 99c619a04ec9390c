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
// Streaming read of `bytes` (a multiple of 16) with `blocks` work-groups of 256; `out` holds
// blocks * 256 u32 (written only for a data value zero-filled memory never gives).
hipError_t launchStreamRead(const void* src, size_t bytes, uint32_t* out, uint32_t blocks, hipStream_t stream);

// The allocation-free Mutations entry point's device scratch: a file-scope device array of this
// many bytes per GPU; noallocScratch resolves its device address on `device` (once, then cached).
constexpr size_t kNoAllocScratchBytes = size_t{128} << 20;
hipError_t noallocScratch(int device, void** ptr);

size_t filterLdsBytes(uint32_t depth);
// Lets the filter kernel allocate up to max_lds_bytes of dynamic LDS (deep operator trees).
hipError_t configureFilterKernel(size_t max_lds_bytes);

// Evaluates n_queries programs over every (chunk, 4096-row slice) of the table (when n_queries_dev
// is set, the number of programs is read from it and n_queries only sizes the grid).  Output words of
// query q start at out + q * n_chunks * 1024; cards[q] and chunk_cards[q * n_chunks + c] must be
// zero on entry and receive the popcounts.
hipError_t launchFilter(const TableView& tv, const ProgramDesc* progs, const Instr* instrs, const uint32_t* ranges,
                        uint32_t n_queries, uint32_t depth, uint64_t* out, unsigned long long* cards,
                        uint32_t* chunk_cards, uint32_t max_blocks, hipStream_t stream,
                        const uint32_t* n_queries_dev = nullptr);

// Mutations for one filter.  Two methods, both launched; the filter's device-side cardinality
// picks one:
//  * dense (cardinality > sparse_max, or sparse_max < 0): per-slot AND-popcounts of the vertical
//    index against the filter (one partial sum per chunk group, n_groups <= kMaxGroups, `partial`
//    holds n_groups * n_slots u32);
//  * sparse (cardinality <= sparse_max): each filter row adds its own mutations and boundaries;
//    sparse_counts ([n_symbols][L] u32) and sparse_bnd ([3][L] i32) must be zero on entry and are
//    zero again on exit.
// Then boundary prefix sums, the implicit (reference) count and thresholding.  counts_out
// ([n_symbols][L]) and rows may be null.
struct MutationsLaunch {
   const SeqColumnView* col_dev = nullptr;
   const SeqColumnView* col_host = nullptr;
   const uint32_t* chunk_sizes = nullptr;  // device
   uint32_t n_chunks = 0;
   uint32_t n_groups = 1;
   uint32_t unroll = 8;                    // 16-byte loads in flight per lane in the dense kernel (8 or 16)
   bool coop = false;                      // a work-group's wavefronts count each chunk tile together
   const uint64_t* filter = nullptr;
   const uint32_t* filter_chunk_cards = nullptr;
   const unsigned long long* filter_card = nullptr;
   long long sparse_max = -1;
   uint32_t* partial = nullptr;
   uint32_t* heavy_sums = nullptr;         // [n_slots], zero; set: 16-bit packed LDS counters (more CU occupancy)
   uint32_t* sparse_counts = nullptr;
   int32_t* sparse_bnd = nullptr;
   int64_t* totals = nullptr;
   double min_proportion = 0.0;
   uint32_t* counts_out = nullptr;
   MutationRowDev* rows = nullptr;
   uint32_t row_capacity = 0;
   uint32_t* n_rows = nullptr;
   hipEvent_t count_start = nullptr;  // recorded around the dense counting kernel when set
   hipEvent_t count_stop = nullptr;
};

hipError_t launchMutations(const MutationsLaunch& m, hipStream_t stream);

// Mutations for a batch of queries whose filters were evaluated together (query q's words at
// filters + q * n_chunks * 1024).  A planning kernel sorts the Mutations queries by their
// device-side cardinality into a dense list and a sparse list (empty filters in neither); the
// counting kernels then work through the lists, each list entry owning its scratch (dense: n_slots
// partial sums; sparse: the zeroed counts and boundary arrays), and the finalize kernels append
// query q's rows at rows + q * row_capacity with their number in n_rows[q] (which may exceed the
// capacity).  Every grid is sized for max_queries and reads the actual count from the device, so
// the launch sequence can be captured once and replayed for any batch of up to max_queries.
constexpr uint32_t kOctetParts = 4;  // chunk parts of the octet shared scan (slot-sum arrays)
constexpr uint32_t kOctetMaxPartChunks = 1024;   // chunks of one part of the octet scan
constexpr uint32_t kSharedMaxPartChunks = 4096;  // chunks of one half of the four-query scan

// The dense scan a batch job runs: the one asked for (0 one pass per query, 1 four queries per
// pass, 2 eight per pass), or the next smaller one where the table is too large for it.  The lane
// allocates the eight-query buffers, and its footprint counts them, only when this returns 2.
inline uint32_t effectiveSharedScan(uint32_t requested, uint32_t max_queries, uint32_t n_chunks) {
   uint32_t mode = requested > 2 ? 2 : requested;
   const uint64_t max_octets = (static_cast<uint64_t>(max_queries) + 7) / 8;
   if (mode == 2 && (max_octets * n_chunks > 65535 || (n_chunks + kOctetParts - 1) / kOctetParts > kOctetMaxPartChunks)) {
      mode = 1;
   }
   if (mode == 1 && (n_chunks + 1) / 2 > kSharedMaxPartChunks) mode = 0;
   return mode;
}

struct MutationsBatch {
   const SeqColumnView* col_dev = nullptr;
   const SeqColumnView* col_host = nullptr;
   const uint32_t* chunk_sizes = nullptr;
   uint32_t n_chunks = 0;
   uint32_t max_queries = 0;
   // query slots the grids cover (the job's size rounded up; 0: max_queries); the buffers' strides
   // stay max_queries
   uint32_t grid_queries = 0;
   const uint32_t* n_queries = nullptr;        // device
   const uint64_t* filters = nullptr;
   const uint32_t* chunk_cards = nullptr;      // [q * n_chunks + c]
   const unsigned long long* cards = nullptr;  // [q]
   const uint32_t* is_mutations = nullptr;     // [q] 1 for Mutations queries on this column
   const double* min_proportion = nullptr;     // [q]
   long long sparse_max = -1;
   uint32_t unroll = 8;
   bool coop = false;                          // per-query dense kernel: wavefronts share each tile
   uint32_t* lists = nullptr;                  // [2 * max_queries]: dense ids, then sparse ids
   uint32_t* list_counts = nullptr;            // [2]
   uint32_t* partial = nullptr;                // [max_queries * n_slots]
   uint32_t* heavy_sums = nullptr;             // [max_queries * n_slots], zero: set = shared-scan kernel
   uint32_t* partial_b = nullptr;              // [max_queries * n_slots]: the shared scan's second part
   // octet shared scan (shared_mode 2): parts 2 and 3 of the slot sums, [2 * max_queries * n_slots],
   // and the row bytes of every octet, [ceil(max_queries / 8) * n_chunks * 65536]
   uint32_t shared_mode = 0;                   // 0: one pass per query, 1: four queries, 2: eight
   uint32_t* partial_c = nullptr;
   uint8_t* octet_masks = nullptr;
   uint32_t persistent_groups = 256;           // work-groups of the persistent octet kernel (one per CU)
   uint32_t octet_threads = 1024;              // their size: 512 or 1024
   uint32_t* sparse_counts = nullptr;          // [max_queries * n_symbols * L], zero
   int32_t* sparse_bnd = nullptr;              // [max_queries * 3 * L], zero
   int64_t* totals = nullptr;                  // [2 * max_queries * position blocks * 3]
   MutationRowDev* rows = nullptr;
   uint32_t row_capacity = 0;
   uint32_t* n_rows = nullptr;                 // [max_queries]
   hipEvent_t scan_start = nullptr;            // recorded around the dense counting kernel when set
   hipEvent_t scan_stop = nullptr;
};

hipError_t launchMutationsBatch(const MutationsBatch& m, hipStream_t stream);

// Aggregated group counts for a batch: entry e (e < *n_agg) counts the rows of query agg_query[e]
// into hist[hist_offset[e] .. + bins[e]), which the launch zeroes first (*total_bins words), then
// compacts the non-empty bins into the query's row slots.
struct AggregateBatch {
   uint32_t n_chunks = 0;
   uint32_t max_entries = 0;
   const uint32_t* n_agg = nullptr;         // device
   const uint32_t* total_bins = nullptr;    // device
   uint32_t max_total_bins = 0;
   const uint32_t* agg_query = nullptr;
   const uint32_t* bins = nullptr;
   const uint32_t* hist_offset = nullptr;
   const GroupDims* dims = nullptr;
   const uint64_t* filters = nullptr;
   const uint32_t* chunk_cards = nullptr;
   uint32_t* hist = nullptr;
   // non-empty bins of entry e go to rows + agg_query[e] * row_capacity as {bin, count, 0, 0}, their
   // number to n_rows[agg_query[e]] (which must be zero; it may exceed the capacity)
   MutationRowDev* rows = nullptr;
   uint32_t row_capacity = 0;
   uint32_t* n_rows = nullptr;
};

hipError_t launchAggregateBatch(const AggregateBatch& a, hipStream_t stream);

// Packs the Mutations rows of a batch: header[q] = {n_rows[q], offset, card low, card high} and
// query q's first min(n_rows[q], row_capacity) rows at packed[offset ..].
hipError_t launchPackRows(const uint32_t* n_queries, uint32_t max_queries, const uint32_t* n_rows,
                          const unsigned long long* cards, const MutationRowDev* rows, uint32_t row_capacity,
                          uint32_t* header, MutationRowDev* packed, hipStream_t stream);

hipError_t launchSymbolLabels(const SeqColumnView* col_dev, const uint32_t* chunk_sizes, uint32_t position,
                              uint32_t n_chunks, uint32_t block, uint32_t slot_lo, uint32_t slot_hi, uint32_t* labels,
                              hipStream_t stream);

hipError_t launchGroupHistogram(uint32_t n_chunks, const uint64_t* filter, const uint32_t* filter_chunk_cards,
                                const GroupDims& dims, uint32_t n_bins, uint32_t* hist, hipStream_t stream);

// ---- ingestion (ingest.hip) -------------------------------------------------------------------------
// Diff of aligned sequences against the reference.  Row i's characters are data[row_ptr[i] ..
// row_ptr[i + 1]), placed from position offsets[i] (0 when offsets is null).  The count pass writes
// per row the covered range [start, end) ([0, 0) when nothing is covered), the number of mutations
// and of internal missing runs, and the first illegal character's position (0xFFFFFFFF: none); the
// write pass writes the mutations at mut_off[i] and the runs at run_off[i].
struct DiffBatch {
   const char* data = nullptr;
   const uint64_t* row_ptr = nullptr;
   const uint32_t* offsets = nullptr;
   const uint8_t* is_null = nullptr;
   const uint8_t* lut = nullptr;  // 256 entries: the symbol of a character, 0xFF for an illegal one
   const uint8_t* ref = nullptr;  // reference symbols
   uint32_t n_rows = 0;
   uint32_t length = 0;
   uint32_t missing = 0;  // the missing symbol
   uint32_t* start = nullptr;
   uint32_t* end = nullptr;
   uint32_t* n_mut = nullptr;
   uint32_t* n_run = nullptr;
   uint32_t* err = nullptr;
   const uint64_t* mut_off = nullptr;
   const uint64_t* run_off = nullptr;
   uint32_t* mut_pos = nullptr;
   uint8_t* mut_sym = nullptr;
   uint32_t* run_start = nullptr;
   uint32_t* run_end = nullptr;
};
hipError_t launchDiff(const DiffBatch& d, bool write, hipStream_t stream);

// A sequence column's rows in diff form on the device: internal missing runs normalized (inside the
// covered range, sorted, disjoint), packed start | end << 16.
struct DiffRowsDev {
   uint64_t n_rows = 0;
   uint32_t length = 0;
   uint32_t n_symbols = 0;
   const uint8_t* is_null = nullptr;
   const uint32_t* cov_start = nullptr;
   const uint32_t* cov_end = nullptr;
   const uint64_t* mut_ptr = nullptr;
   const uint32_t* mut_pos = nullptr;
   const uint8_t* mut_sym = nullptr;
   const uint64_t* nrun_ptr = nullptr;
   const uint32_t* runs = nullptr;
   const uint8_t* reference = nullptr;
};

// Coverage and symbol counts per position, then the local reference of every position, without
// atomics: every row writes one key `position << 5 | code` per mutation (code = symbol) and per
// coverage event (code 28: coverage +1 at a range start or run end, 29: -1 at a range end or run
// start; a row without a covered range writes two keys past the end), the keys are radix-sorted,
// and a run of equal keys [start, end) is its count.  keys / keys_alt: key_count u32 each (row i's
// keys at mut_ptr[i] + 2 nrun_ptr[i] + 2 i); run_start / run_end: (length + 1) * 32 u32, zero.
constexpr uint32_t kAdaptCodeUp = 28, kAdaptCodeDown = 29;
struct AdaptCount {
   DiffRowsDev rows;
   uint32_t* keys = nullptr;
   uint32_t* keys_alt = nullptr;
   uint64_t key_count = 0;
   uint32_t* run_start = nullptr;
   uint32_t* run_end = nullptr;
   uint32_t* coverage = nullptr;
   uint8_t* local_ref = nullptr;
};
// Key writing, sort (temp == null: sets *temp_bytes only), run boundaries and the rule.
hipError_t launchAdapt(const AdaptCount& a, void* temp, size_t* temp_bytes, hipStream_t stream);

// The container entries of every row (chunk << 40 | position << 24 | symbol << 16 | row in chunk):
// the count pass writes their number per row, the write pass writes them at offsets[i].
struct EntryBatch {
   DiffRowsDev rows;
   const uint64_t* chunk_offset = nullptr;  // [n_chunks + 1]: first row of each chunk
   uint32_t n_chunks = 0;
   const uint8_t* flip_to = nullptr;  // per position: the local reference symbol, 0xFF if unflipped
   const uint32_t* flipped = nullptr;  // flipped positions, ascending
   uint32_t n_flipped = 0;
   uint32_t* counts = nullptr;
   const uint64_t* offsets = nullptr;
   uint64_t* keys = nullptr;
};
hipError_t launchEntries(const EntryBatch& e, bool write, hipStream_t stream);
// Exclusive prefix sum of n u32 counts into u64 offsets (temp == null: sets *temp_bytes).
hipError_t scanEntryCounts(const uint32_t* counts, uint64_t* offsets, uint64_t n, void* temp, size_t* temp_bytes,
                           hipStream_t stream);
// Radix sort of n (< 2^31) keys over bits [0, 56); the sorted keys end in *sorted (keys or
// keys_alt).  temp == null: sets *temp_bytes.
hipError_t sortEntryKeys(uint64_t* keys, uint64_t* keys_alt, uint64_t n, void* temp, size_t* temp_bytes,
                         hipStream_t stream, uint64_t** sorted);

}  // namespace rdb
