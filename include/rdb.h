/*
 * rdb.h — C ABI of the MI355X query engine for SILO/RhyDB's filter, Mutations and Aggregated path.
 *
 * The engine keeps sequence and metadata columns resident on one GPU (gfx950) and evaluates
 * filter expressions, Mutations counts and group-by counts with hand-written HIP kernels.  All
 * arguments are plain pointers and sizes; host arrays are borrowed for the duration of a call.
 * Every entry point returns an `int` status (RDB_OK on success); the text of the last error of a
 * context is available from rdb_last_error().  No C++ exception crosses this boundary.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   - storage upload: SequenceColumn::appendChunk/finalize (src/rhydb/storage/column/
 *     sequence_column.cpp:99-212), Date32Column / Int32Column / DictionaryEncodedColumn appendChunk,
 *     Table::finalize (src/rhydb/storage/table.cpp:87-94);
 *   - filter evaluation: computeFilter (src/rhydb/query_engine/operators/compute_filter.cpp:14-21)
 *     = ScalarExpression::rewrite + compile + Operator::evaluate
 *     (src/rhydb/query_engine/filter/operators/operator.h:24-37);
 *   - Mutations action: calculateMutationsPerPosition + addMutationsToOutput
 *     (src/rhydb/query_engine/operators/mutations_node.cpp:269-366);
 *   - Aggregated action: BitmapAggregationNode::addToExecPlan
 *     (src/rhydb/query_engine/operators/bitmap_aggregation_node.cpp:304-356);
 *   - Count action: CountFilterNode::addToExecPlan (src/rhydb/query_engine/operators/
 *     count_filter_node.cpp:35-71).
 */
#ifndef RDB_H
#define RDB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum rdb_status {
   RDB_OK = 0,
   RDB_EINVAL = 1,   /* malformed argument or ingest data */
   RDB_ENOMEM = 2,   /* device or host allocation failed */
   RDB_EDEVICE = 3,  /* HIP runtime error */
   RDB_EQUERY = 4,   /* illegal query (reference: IllegalQueryException, HTTP 400) */
   RDB_ECOMPILE = 5  /* operator construction rule violated (reference: QueryCompilationException) */
};

enum rdb_alphabet { RDB_NUCLEOTIDE = 0, RDB_AMINO_ACID = 1 };

typedef struct rdb_ctx rdb_ctx;       /* one GPU, one stream, device memory pool */
typedef struct rdb_table rdb_table;   /* row layout + device-resident columns */
typedef struct rdb_bitmap rdb_bitmap; /* device-resident row set (result of a filter) */

/* ---- context ---------------------------------------------------------------------------------- */
/* Replaces: nothing in the reference (the reference has no device); owns the HIP stream. */
int rdb_open(int device, rdb_ctx** out);
void rdb_close(rdb_ctx* ctx);
const char* rdb_last_error(const rdb_ctx* ctx);

/* ---- tables ----------------------------------------------------------------------------------- */
/* Row layout of 2^16-row chunks (storage/column/row_layout.h:27-47, row_id.h:16-37).  chunk_sizes
 * holds the rows of each chunk (<= 65536); n_chunks < 65535. */
int rdb_table_create(rdb_ctx* ctx, const uint32_t* chunk_sizes, uint32_t n_chunks, rdb_table** out);
void rdb_table_free(rdb_table* table);

/* A sequence column in post-diff form (storage/column/sequence_column.h:66-71 BufferedSequence,
 * common/aligned_sequence.cpp:18-122): per row the covered range [cov_start, cov_end), internal
 * missing runs [nrun_start, nrun_end) (CSR by nrun_ptr) and mutations against `reference`
 * (CSR by mut_ptr; positions strictly increasing; symbols in the alphabet's order, never the
 * missing symbol).  genome_length < 65536. */
int rdb_table_add_sequence_column(rdb_table* table, const char* name, int alphabet, const uint8_t* reference,
                                  uint32_t genome_length, const uint8_t* is_null, const uint32_t* cov_start,
                                  const uint32_t* cov_end, const uint64_t* mut_ptr, const uint32_t* mut_pos,
                                  const uint8_t* mut_sym, const uint64_t* nrun_ptr, const uint32_t* nrun_start,
                                  const uint32_t* nrun_end);
/* A sequence column in the state the reference's SequenceColumn holds after finalize
 * (storage/column/sequence_column.h:86-97), as a maintainer reads it out of a finalized or loaded
 * column:
 *   - reference: SequenceColumnMetadata::reference_sequence; local_reference: the symbols of
 *     local_reference_sequence_string (sequence_column.h:118 getLocalReference);
 *   - the vertical index (vertical_sequence_index.h:22-44): n_containers entries of
 *     vertical_bitmaps, keyed (position, chunk = v_index, symbol), each container as
 *     RoaringContainer serializes it (roaring_container.h:105-116): its cardinality, its CRoaring
 *     typecode (1 bitset, 2 array, 3 run) and the bytes container_write writes for it, at
 *     payload_offset in `payload` (bitset 8,192 bytes; array 2 per value; run a u16 run count and
 *     (start, length - 1) u16 pairs);
 *   - the horizontal coverage index (horizontal_coverage_index.h:21-35): every row's [start, end)
 *     from start_end, and the missing positions horizontal_bitmaps holds for it (CSR by row:
 *     missing_ptr[n_rows + 1], missing_pos ascending per row);
 *   - null_bitmap, as a flag per row.
 * Rows are in row-layout order (row r of chunk c at index chunk offset + r).  The engine keeps the
 * given local reference (the reference adapts it again after each append, relative to the one
 * before) and answers every query as for the same column uploaded in diff form; RDB_EINVAL for
 * containers that do not fit the row layout, hold the local reference symbol, or rows whose
 * mutations lie outside their covered range. */
int rdb_table_add_sequence_column_finalized(rdb_table* table, const char* name, int alphabet, const uint8_t* reference,
                                            const uint8_t* local_reference, uint32_t genome_length, uint64_t n_containers,
                                            const uint32_t* key_position, const uint16_t* key_chunk,
                                            const uint8_t* key_symbol, const uint8_t* typecode,
                                            const uint32_t* cardinality, const uint64_t* payload_offset,
                                            const uint8_t* payload, uint64_t payload_bytes, const uint8_t* is_null,
                                            const uint32_t* cov_start, const uint32_t* cov_end,
                                            const uint64_t* missing_ptr, const uint32_t* missing_pos);
/* Scanners for saved sequence columns (the reference's Boost binary archives; no counterpart in the
 * reference, which reads them through Boost.Serialization).  rdb_archive_scan_containers reads n
 * vertical-index records starting at buf[pos] (after the first record's class information: u32
 * position, u16 chunk, i32 symbol, u32 cardinality, u8 typecode, u64 length, the container_write
 * bytes; vertical_sequence_index.h:31-44, roaring_container.h:105-116), fills the keys and each
 * container's offset in the packed payload, copies the bytes into payload when it is non-null
 * (payload_capacity bytes), and returns the payload's size and the position after the records.
 * rdb_archive_scan_row_bitmaps reads n horizontal-index records (u32 row id, u64 length, a
 * portable roaring bitmap; horizontal_coverage_index.h:21-35) and returns where each bitmap lies. */
int rdb_archive_scan_containers(const uint8_t* buf, uint64_t len, uint64_t pos, uint64_t n, uint32_t* position,
                                uint16_t* chunk, int32_t* symbol, uint32_t* cardinality, uint8_t* typecode,
                                uint64_t* payload_offset, uint8_t* payload, uint64_t payload_capacity,
                                uint64_t* payload_bytes, uint64_t* end);
/* The maximal runs [start, end) of n portable roaring bitmaps (bitmap i at buf[start[i]], length[i]
 * bytes; CRoaring's portable format, roaring_serialize.h:16-48): the missing positions of a saved
 * horizontal coverage index as runs (horizontal_coverage_index.cpp:35-45).  Called first with
 * run_start / run_end null it fills run_ptr[n + 1] (CSR offsets); called again with arrays of
 * run_ptr[n] entries (capacity) it writes the runs.  RDB_EINVAL for a malformed bitmap.  No
 * counterpart in the reference, which deserializes through CRoaring. */
int rdb_archive_bitmap_runs(const uint8_t* buf, uint64_t len, const uint64_t* start, const uint64_t* length, uint64_t n,
                            uint64_t* run_ptr, uint32_t* run_start, uint32_t* run_end, uint64_t capacity);
int rdb_archive_scan_row_bitmaps(const uint8_t* buf, uint64_t len, uint64_t pos, uint64_t n, uint32_t* row,
                                 uint64_t* bitmap_start, uint64_t* bitmap_length, uint64_t* end);
/* rdb_table_add_sequence_column_finalized with each row's missing positions given as the runs
 * [start, end) of its horizontal bitmap (CSR: missing_run_ptr[n_rows + 1], runs ascending and
 * disjoint, ends at most the genome length): the form a saved state's run-optimized roaring
 * bitmaps hold (horizontal_coverage_index.cpp:35-45), without expanding every position. */
int rdb_table_add_sequence_column_finalized_runs(
   rdb_table* table, const char* name, int alphabet, const uint8_t* reference, const uint8_t* local_reference,
   uint32_t genome_length, uint64_t n_containers, const uint32_t* key_position, const uint16_t* key_chunk,
   const uint8_t* key_symbol, const uint8_t* typecode, const uint32_t* cardinality, const uint64_t* payload_offset,
   const uint8_t* payload, uint64_t payload_bytes, const uint8_t* is_null, const uint32_t* cov_start,
   const uint32_t* cov_end, const uint64_t* missing_run_ptr, const uint32_t* missing_run_start,
   const uint32_t* missing_run_end);
/* A sequence column built from aligned sequences, as the reference's SequenceColumnBuilder
 * (storage/column/sequence_column.h:178-233) buffers rows and SequenceColumn::appendChunk /
 * finalize (sequence_column.cpp:99-212) index them.  Each append diffs its rows on the GPU against
 * `reference` (symbols; extractCoverageAndMutationsFromSequence, common/aligned_sequence.cpp:18-122):
 * row i's characters are data[row_ptr[i] .. row_ptr[i + 1]), placed from position offsets[i]
 * (SequenceColumnBuilder::insert's offset, sequence_column.cpp:312-341; offsets may be null: 0),
 * is_null may be null.  Errors as the reference's insert: RDB_EINVAL "illegal character 'c' at
 * position p in the input sequence", or "the sequence '...' which was inserted with an offset o is
 * larger than the length of the reference genome: L"; a failed append adds none of its rows.
 * finish (every row of the table appended, in row order) builds the column into the table: the
 * same layout rdb_table_add_sequence_column builds from the same rows in diff form, the local-
 * reference adaptation and the sorted container entries computed on the GPU.  Free a builder
 * before its table. */
typedef struct rdb_seq_builder rdb_seq_builder;
int rdb_seq_builder_create(rdb_table* table, const char* name, int alphabet, const uint8_t* reference,
                           uint32_t genome_length, rdb_seq_builder** out);
int rdb_seq_builder_append(rdb_seq_builder* builder, const char* data, const uint64_t* row_ptr, const uint32_t* offsets,
                           const uint8_t* is_null, uint64_t n_rows);
int rdb_seq_builder_finish(rdb_seq_builder* builder);
void rdb_seq_builder_free(rdb_seq_builder* builder);
/* No reference counterpart: a 64-bit digest of a sequence column's device buffers and host-side
 * index state, equal for equal layouts (build checks). */
int rdb_table_column_digest(rdb_table* table, const char* column, uint64_t* digest);
/* int32 or date32 column (storage/column/int_column.h:140, date32_column.h:20-80). */
int rdb_table_add_int_column(rdb_table* table, const char* name, int is_date, const int32_t* values,
                             const uint8_t* is_null);
/* Indexed dictionary column (storage/column/dictionary_encoded_column.h:67-140); names are the
 * dictionary strings (n_values of them, concatenated, name_offsets[n_values + 1]).  A non-null
 * lineage_parent (-1 = root) also builds the lineage index (storage/column/lineage_index.h:17-42). */
int rdb_table_add_dict_column(rdb_table* table, const char* name, const uint32_t* ids, const uint8_t* is_null,
                              uint32_t n_values, const char* names, const uint32_t* name_offsets,
                              const int32_t* lineage_parent);
/* rdb_table_add_dict_column with lineage aliases: lineage_alias[v] is the value v resolves to (-1 =
 * none; LineageTree::resolveAlias, common/lineage_tree.cpp:164-169).  Rows are indexed under their
 * value's resolved lineage and a lineage filter on an alias gives its lineage's rows
 * (storage/column/lineage_index.cpp:17-45); string equality keeps the stored values. */
int rdb_table_add_lineage_column(rdb_table* table, const char* name, const uint32_t* ids, const uint8_t* is_null,
                                 uint32_t n_values, const char* names, const uint32_t* name_offsets,
                                 const int32_t* lineage_parent, const int32_t* lineage_alias);
/* A literal row set usable in filters as (bitmap ID) — the IndexScan over a literal bitmap that the
 * reference's operator tests use (filter/operators/index_scan.h).  Row ids must lie in existing
 * chunks. */
int rdb_table_add_bitmap(rdb_table* table, const uint32_t* rows, uint64_t n, uint32_t* id_out);
/* Local-reference adaptation, container build and upload (sequence_column.cpp:157-212). */
int rdb_table_finalize(rdb_table* table);
/* The local reference after adaptation (SequenceColumn::getLocalReference, sequence_column.h:118). */
int rdb_table_local_reference(const rdb_table* table, const char* column, uint8_t* out);
/* Device bytes held by the table and the vertical-index payload statistics of a sequence column. */
int rdb_table_stats(const rdb_table* table, const char* column, uint64_t* device_bytes, uint64_t* n_containers,
                    uint64_t* payload_bytes);

/* No reference counterpart: wall time (ms) of the phases of a sequence column's index build, in
 * order: checks of the rows, local-reference adaptation, containers (entries generated, sorted and
 * packed), slots / position blocks / tiles, row-level arrays, layout checks, upload.  Phases past
 * the seventh read 0. */
int rdb_table_build_profile(const rdb_table* table, const char* column, double* phase_ms, uint32_t n_phases);

/* ---- filters ---------------------------------------------------------------------------------- */
/* computeFilter (operators/compute_filter.cpp:14-21).  `filter` is the expression tree in the
 * engine's S-expression form, e.g.
 *   (and (date_between date 18262 18351) (str_eq country "Switzerland"))
 *   (nof 2 0 (sym_eq main 122 T) (has_mut main 2000) (not (lineage lineage "B.1" 1)))
 * The result stays on the device; *cardinality receives its row count. */
int rdb_filter(rdb_table* table, const char* filter, rdb_bitmap** out, uint64_t* cardinality);
int rdb_bitmap_cardinality(const rdb_bitmap* bitmap, uint64_t* cardinality);
/* Global row ids (chunk << 16 | row), ascending (CopyOnWriteBitmap::toRoaring, copy_on_write_bitmap.cpp:352-364). */
int rdb_bitmap_rows(rdb_bitmap* bitmap, uint32_t* rows, uint64_t capacity, uint64_t* n);
void rdb_bitmap_free(rdb_bitmap* bitmap);
/* Operator::toString (filter/operators/operator.h:33) of the operator tree the filter compiles to
 * (rewrite + compile, compute_filter.cpp:14-20, without evaluating it), e.g. "RangeSelection(...)"
 * for a range on a sorted date column.  *length receives the text's length; the text (with a
 * terminating zero) is written when capacity > length, else RDB_ENOMEM. */
int rdb_filter_explain(rdb_table* table, const char* filter, char* out, uint64_t capacity, uint64_t* length);
/* CountFilterNode: |filter| (count_filter_node.cpp:35-71). */
int rdb_count(rdb_table* table, const char* filter, uint64_t* count);

/* ---- actions ---------------------------------------------------------------------------------- */
/* Per-symbol counts counts[symbol][position] (u32, n_symbols x genome_length, symbols of the
 * column's alphabet) after local-reference accumulation (mutations_node.cpp:206-288). */
int rdb_mutation_counts(rdb_table* table, const rdb_bitmap* filter, const char* column, uint32_t* counts);
/* Mutations rows (mutations_node.cpp:292-366) in output order: position ascending (1-based), then
 * symbol in VALID_MUTATION_SYMBOLS order.  proportion = count / coverage in double. */
int rdb_mutations(rdb_table* table, const rdb_bitmap* filter, const char* column, double min_proportion,
                  uint32_t* position, uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage,
                  double* proportion, uint64_t capacity, uint64_t* n);
/* MutationsNode::addToExecPlan as one call (mutations_node.cpp:372-428): computeFilter on the
 * device, then the Mutations rows as rdb_mutations returns them, with a single wait. */
int rdb_mutations_query(rdb_table* table, const char* filter, const char* column, double min_proportion,
                        uint32_t* position, uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage,
                        double* proportion, uint64_t capacity, uint64_t* n);
/* No counterpart in the reference.  rdb_mutations_query for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same results, but every device buffer the call needs beyond the table (staged filter program,
 * filter words, partial sums, accumulators, rows) is carved from one file-scope device array of
 * 128 MiB per GPU, held until the process ends; the call makes no hipMalloc, hipFree or pinned
 * allocation.  RDB_EINVAL when the table needs more scratch than that (about 8 KiB per chunk for
 * the filter plus 4 bytes per (position, symbol) slot per chunk group).  All calls on one GPU share
 * the array: calls that overlap on one GPU (from different contexts or threads) race. */
int rdb_mutations_query_noalloc(rdb_table* table, const char* filter, const char* column, double min_proportion,
                                uint32_t* position, uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage,
                                double* proportion, uint64_t capacity, uint64_t* n);
/* ---- result sinks ------------------------------------------------------------------------------ */
/* Output columns of the Mutations action (MutationsNode::getOutputSchema, operators/
 * mutations_node.h:60-84), written in that order whatever the order of the bits. */
enum rdb_mutation_field {
   RDB_FIELD_MUTATION_FROM = 1,  /* utf8: the reference symbol */
   RDB_FIELD_MUTATION_TO = 2,    /* utf8 */
   RDB_FIELD_SEQUENCE_NAME = 4,  /* utf8 */
   RDB_FIELD_POSITION = 8,       /* int32, 1-based */
   RDB_FIELD_PROPORTION = 16,    /* float64 */
   RDB_FIELD_COVERAGE = 32,      /* int32 */
   RDB_FIELD_COUNT = 64,         /* int32 */
   RDB_FIELDS_ALL = 127
};
/* Mutations rows (as rdb_mutations returns them, symbols of `alphabet`) written as the reference's
 * NdjsonSink writes a result batch (query_engine/exec_node/ndjson_sink.cpp: one JSON object per
 * line, fields in schema order, doubles in shortest round-trip form).  *length receives the byte
 * count; RDB_ENOMEM (nothing written) when capacity is smaller.  Host only: no device work. */
int rdb_mutation_rows_ndjson(int alphabet, const char* sequence_name, uint32_t fields, const uint32_t* position,
                             const uint8_t* from, const uint8_t* to, const uint32_t* count, const uint32_t* coverage,
                             const double* proportion, uint64_t n, char* out, uint64_t capacity, uint64_t* length);
/* The same rows as an Arrow IPC stream (schema message, one record batch, end-of-stream marker),
 * the format the reference's ArrowIpcSink writes through arrow::ipc::MakeStreamWriter
 * (exec_node/arrow_ipc_sink.cpp:49-60); string columns utf8, position / coverage / count int32,
 * proportion float64, all nullable.  Host only. */
int rdb_mutation_rows_arrow_ipc(int alphabet, const char* sequence_name, uint32_t fields, const uint32_t* position,
                                const uint8_t* from, const uint8_t* to, const uint32_t* count, const uint32_t* coverage,
                                const double* proportion, uint64_t n, uint8_t* out, uint64_t capacity, uint64_t* length);
/* The Mutations action under `filter` on `column` (as rdb_mutations), its rows written to NDJSON
 * (addMutationsToOutput -> NdjsonSink, mutations_node.cpp:292-366) or to an Arrow IPC stream
 * (-> ArrowIpcSink).  *length and RDB_ENOMEM as above. */
int rdb_mutations_ndjson(rdb_table* table, const rdb_bitmap* filter, const char* column, double min_proportion,
                         uint32_t fields, char* out, uint64_t capacity, uint64_t* length);
int rdb_mutations_arrow_ipc(rdb_table* table, const rdb_bitmap* filter, const char* column, double min_proportion,
                            uint32_t fields, uint8_t* out, uint64_t capacity, uint64_t* length);

/* Aggregated count group-by (bitmap_aggregation_node.cpp:91-130, 189-253).  dims: comma-separated
 * dimensions; "name" = indexed dictionary column, "name@pos" = symbol at 0-based position of a
 * sequence column.  Output: combinations in the reference's depth-first order, n_dims ids each
 * (dictionary value id or symbol; UINT32_MAX = null group) and their counts. */
int rdb_aggregate(rdb_table* table, const rdb_bitmap* filter, const char* dims, uint32_t* ids, uint64_t* counts,
                  uint64_t capacity, uint64_t* n);

/* ---- measurement ------------------------------------------------------------------------------ */
/* No reference counterpart (the reference's evobench probes time CPU scopes, src/evobench/
 * evobench.hpp:144-230).  Enables (and resets) HIP-event timing of the Mutations counting kernel
 * and the filter kernel on the context's stream; timing adds a wait after each timed launch. */
/* No reference counterpart: the bandwidth a plain streaming read of `bytes` (a multiple of 16)
 * reaches on this device, averaged over `reps` back-to-back launches timed with HIP events (the
 * device-side reference point for the counting kernels' roofline fractions). */
int rdb_measure_read_bandwidth(rdb_ctx* ctx, uint64_t bytes, int reps, double* gb_per_s);
int rdb_profile(rdb_ctx* ctx, int enable);
int rdb_profile_read(rdb_ctx* ctx, double* mutations_ms, uint64_t* mutations_launches, double* filter_ms,
                     uint64_t* filter_launches);
/* The batched path's dense counting kernel while profiling: its total time, launches (one per lane
 * job with Mutations queries) and the dense Mutations queries those launches counted.  Profiled
 * jobs run without graph replay, with HIP events on their lane's stream around the kernel. */
int rdb_profile_read_scan(rdb_ctx* ctx, double* scan_ms, uint64_t* scan_launches, uint64_t* dense_queries);
/* Algorithmic bytes of one Mutations evaluation under `filter` by the method the context's tuning
 * selects.  Dense: stored container bytes (array 2*card, bitset 8192, run 4*runs, plus an 8-byte
 * descriptor each) of the chunks the filter touches, 8192 per such chunk (filter tile), 8 per row
 * and 4 per internal missing run of those chunks (coverage), and 4 * n_symbols * genome_length
 * (counts).  Sparse: the filter words, 20 bytes per filter row plus 4 per mutation and per missing
 * run of those rows, and the counts and boundary arrays read and cleared. */
int rdb_mutations_traffic(rdb_table* table, const rdb_bitmap* filter, const char* column, uint64_t* bytes);

/* ---- batched queries ------------------------------------------------------------------------- */
/* No single reference counterpart: the reference runs one query per worker thread
 * (src/rhydb/api/query_handler.cpp:38-99), each as computeFilter + the action's node.  A batch runs
 * many such queries with one launch sequence per lane job; each query's results are those of the
 * single-query entry points (rdb_count, rdb_mutations_query, rdb_aggregate) on the same filter. */
enum rdb_action { RDB_ACTION_COUNT = 0, RDB_ACTION_MUTATIONS = 1, RDB_ACTION_AGGREGATE = 2 };

typedef struct rdb_query {
   const char* filter;     /* filter expression, as for rdb_filter */
   int action;             /* rdb_action */
   const char* column;     /* Mutations: sequence column */
   double min_proportion;  /* Mutations: minProportion */
   const char* dims;       /* Aggregated: comma-separated dictionary columns */
} rdb_query;

typedef struct rdb_batch rdb_batch;

/* Compiles the n queries and enqueues them; returns without waiting for the GPU.  A query whose
 * filter or action is invalid does not fail the batch: its getters return its own status. */
int rdb_batch_submit(rdb_table* table, const rdb_query* queries, uint32_t n, rdb_batch** out);
/* Waits for every job of the batch (the getters wait too). */
int rdb_batch_wait(rdb_batch* batch);
/* Collects the batch's finished jobs without waiting; *done = 1 once all have finished. */
int rdb_batch_poll(rdb_batch* batch, int* done);
/* Device memory one lane holds for jobs of up to max_queries queries keeping up to row_capacity
 * records each: the basis of admission by free device memory. */
int rdb_batch_lane_bytes(rdb_table* table, uint32_t max_queries, uint32_t row_capacity, uint64_t* bytes);
/* Free and total device memory of the context's GPU (hipMemGetInfo). */
int rdb_device_memory(rdb_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* Cardinality of query q's filter (any action): CountFilterNode (count_filter_node.cpp:35-71). */
int rdb_batch_count(rdb_batch* batch, uint32_t q, uint64_t* count);
/* Rows of Mutations query q, as rdb_mutations. */
int rdb_batch_mutations(rdb_batch* batch, uint32_t q, uint32_t* position, uint8_t* from, uint8_t* to,
                        uint32_t* count, uint32_t* coverage, double* proportion, uint64_t capacity, uint64_t* n);
/* Every query's status (RDB_OK or the error it failed with), filter cardinality and result size
 * (Mutations rows, Aggregated groups, 0 for Count) in one call, for the first n_queries queries:
 * no counterpart in the reference (a batch is this engine's own); for callers that collect many
 * results at once. */
int rdb_batch_summary(rdb_batch* batch, uint32_t n_queries, int* status, uint64_t* cardinality, uint64_t* size);
/* Rows of every Mutations query of the batch, concatenated in query order: query q's rows are
 * [offsets[q], offsets[q + 1]) (empty for other actions and for failed queries); offsets has
 * n_queries + 1 entries and the row arrays are as rdb_mutations'.  RDB_ENOMEM (nothing written but
 * *n, the total) when capacity is too small.  No counterpart in the reference. */
int rdb_batch_mutations_all(rdb_batch* batch, uint32_t n_queries, uint64_t* offsets, uint32_t* position,
                            uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage, double* proportion,
                            uint64_t capacity, uint64_t* n);
/* Groups of Aggregated query q, as rdb_aggregate. */
int rdb_batch_aggregate(rdb_batch* batch, uint32_t q, uint32_t* ids, uint64_t* counts, uint64_t capacity,
                        uint64_t* n);
void rdb_batch_free(rdb_batch* batch);

/* ---- learned planner -------------------------------------------------------------------------- */
/* No reference counterpart: the reference plans with fixed rules (Intersection sorts evaluated
 * children by cardinality, intersection.cpp:74-86; Selection orders predicates by a fixed
 * selectivity guess, selection.cpp:37, and switches to per-row matching for small children,
 * selection.cpp:105).  Here a multilayer perceptron (ReLU hidden layers, BF16 parameters, FP32
 * accumulation; forward, backward and Adam as HIP kernels) estimates every operator node's
 * cardinality from per-node features, and the batched path uses the estimates to order
 * conjuncts / disjuncts and to skip the rest of an And (Or) in every 4096-row slice where it is
 * already empty (full).  No plan changes a result. */
typedef struct rdb_mlp rdb_mlp;

/* Parameters: per layer W (out x in, row-major) then b; inputs -> n_hidden_layers x n_hidden -> 1.
 * He-uniform initial weights from `seed`, zero biases. */
int rdb_mlp_create(rdb_ctx* ctx, uint32_t n_in, uint32_t n_hidden, uint32_t n_hidden_layers, uint64_t seed,
                   rdb_mlp** out);
void rdb_mlp_free(rdb_mlp* mlp);
int rdb_mlp_parameter_count(rdb_mlp* mlp, uint64_t* n);
/* The BF16 parameter values the kernels use, as float. */
int rdb_mlp_get_parameters(rdb_mlp* mlp, float* out, uint64_t n);
/* Sets the FP32 master parameters (rounded to BF16 for the products) and resets Adam. */
int rdb_mlp_set_parameters(rdb_mlp* mlp, const float* p, uint64_t n);
/* y[i] = network(x[i * n_in ..]) for n rows. */
int rdb_mlp_forward(rdb_mlp* mlp, const float* x, uint32_t n, float* y);
/* Mean squared error over n rows and its gradient with respect to every parameter. */
int rdb_mlp_gradients(rdb_mlp* mlp, const float* x, const float* y, uint32_t n, float* loss, float* grads);
/* Adam (lr, 0.9, 0.999, 1e-8) over `epochs` shuffled passes of minibatches; the last pass's MSE. */
int rdb_mlp_train(rdb_mlp* mlp, const float* x, const float* y, uint64_t n, uint32_t epochs, uint32_t batch, float lr,
                  uint64_t seed, float* final_loss);
/* Training rows of the planner: one per operator node of every filter's compiled tree (post-order),
 * 32 features with the children's exact values in the child-summary features, and the target:
 * log(1 + rows) / log(1 + table rows) of the node's exact cardinality (evaluated on the device).
 * tree[i] = the filter the row belongs to.  Any output may be null. */
int rdb_plan_samples(rdb_table* table, const char* const* filters, uint32_t n_filters, float* features, float* targets,
                     uint32_t* tree, uint64_t capacity, uint64_t* n);
/* Per node of every filter: the learned estimate (rows; 0 when mlp is null), the independence
 * estimate (rows) and the exact cardinality (exact may be null to skip evaluating). */
int rdb_plan_estimates(rdb_table* table, rdb_mlp* mlp, const char* const* filters, uint32_t n_filters, float* estimate,
                       float* independence, uint64_t* exact, uint32_t* tree, uint64_t capacity, uint64_t* n);
/* Planner of the batched path: 0 = the reference's order, nothing skipped (default); 1 = order by
 * independence estimates, short-circuit on; 2 = order by the estimator `mlp`, short-circuit on. */
int rdb_set_planner(rdb_ctx* ctx, int mode, rdb_mlp* mlp);

/* ---- tuning ----------------------------------------------------------------------------------- */
/* No reference counterpart: knobs of the engine's own kernels, set by hand or by the autotuner.
 *   "sparse_max_rows"  Mutations under a filter of at most this many rows walk the rows' own
 *                      mutations instead of streaming the index (-1, the default: never);
 *   "mutation_groups"  chunk groups (partial sums) of the dense Mutations kernel, 0 = automatic
 *                      (at most 32);
 *   "mutation_packed"  1: the dense Mutations kernel keeps two 16-bit slot counters per LDS word
 *                      (used when a group holds at most 15 chunks), 0: one 32-bit (default);
 *   "mutation_unroll"  16-byte loads per lane and round in the dense Mutations kernel (two rounds
 *                      in flight): 2, 4 (default), 8 or 16;
 *   "mutation_coop"    1: the four wavefronts of a dense-kernel work-group count each chunk's tile
 *                      together (default), 0: each wavefront takes its own chunks;
 *   "batch_max_queries" queries per lane job of the batched path (1..4096, default 256);
 *   "batch_lanes"      lanes (HIP streams with their own buffers) per table (1..16, default 2);
 *   "batch_graphs"     1: replay each lane's launch sequence from a captured HIP graph (default);
 *   "batch_shared_scan" how a job's dense Mutations queries pass over the index: 0 one pass per
 *                      query (default), 1 four queries per pass, 2 eight per pass (row bytes of
 *                      eight filters; the lane then also holds 1 byte per row per 8 queries);
 *   "batch_octet_threads" work-group size of the eight-per-pass scan, 512 or 1024 (default);
 *   "batch_octet_groups" its persistent work-groups, 0 (default: one per CU) .. 4096; fewer leave
 *                      CUs to concurrent batches;
 *   "batch_row_capacity" Mutations rows per query kept by a lane job (more: the query runs again
 *                      alone), default 4096;
 *   "index_build_device" 1: the local-reference adaptation and container entries of a sequence
 *                      column's index build run on the GPU (default), 0: on the host;
 *   "batch_fail_job"   fault injection for tests: the n-th lane job enqueued from now on fails with
 *                      RDB_ENOMEM after its inputs were staged, failing its rdb_batch_submit (0, the
 *                      default: never).
 * Results never depend on them.  RDB_EINVAL for an unknown key or an out-of-range value. */
int rdb_set_tuning(rdb_ctx* ctx, const char* key, int64_t value);
int rdb_get_tuning(rdb_ctx* ctx, const char* key, int64_t* value);

/* ---- ingest helper ---------------------------------------------------------------------------- */
/* extractCoverageAndMutationsFromSequence (common/aligned_sequence.cpp:18-122) with offset 0.
 * Returns RDB_EINVAL on an illegal character and RDB_ENOMEM when a capacity is too small. */
int rdb_diff_sequence(int alphabet, const char* sequence, uint32_t length, const char* reference,
                      uint32_t* start, uint32_t* end, uint32_t* missing, uint64_t missing_capacity,
                      uint64_t* n_missing, uint32_t* mut_pos, uint8_t* mut_sym, uint64_t mut_capacity,
                      uint64_t* n_mut);

#ifdef __cplusplus
}
#endif

#endif /* RDB_H */
