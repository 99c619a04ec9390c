// The C ABI (include/rdb.h): status codes at the boundary, exceptions inside.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "capi_internal.h"
#include "engine.h"
#include "ingest.h"
#include "query.h"

using namespace rdb;

namespace rdb {

void Context::init(int dev) {
   device = dev;
   hipCheck(hipSetDevice(dev), "hipSetDevice");
   hipCheck(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
   hipDeviceProp_t prop{};
   if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0) cu_count = prop.multiProcessorCount;
   int lds = 0;
   if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds >= 64 * 1024) {
      max_lds_bytes = static_cast<size_t>(lds);
   }
   if (configureFilterKernel(max_lds_bytes) != hipSuccess) {
      (void)hipGetLastError();
      max_lds_bytes = 64 * 1024;
   }
   hipCheck(hipHostMalloc(&stage_host, kStageSlots * kStageSlotBytes, hipHostMallocDefault), "pinned staging");
   stage_dev = DeviceBuffer(kStageSlots * kStageSlotBytes);
   for (auto& e : stage_events) hipCheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
   for (auto& e : prof_events) hipCheck(hipEventCreate(&e), "hipEventCreate");
   hipCheck(noallocScratch(dev, &noalloc_scratch), "allocation-free scratch address");
}

void Context::profileAccumulate(int which) {
   hipCheck(hipEventSynchronize(prof_events[2 * which + 1]), "profile wait");
   float ms = 0;
   hipCheck(hipEventElapsedTime(&ms, prof_events[2 * which], prof_events[2 * which + 1]), "profile time");
   prof_ms[which] += ms;
   prof_launches[which] += 1;
}

Context::~Context() {
   if (stream) (void)hipStreamSynchronize(stream);
   for (auto& e : stage_events) {
      if (e) (void)hipEventDestroy(e);
   }
   for (auto& e : prof_events) {
      if (e) (void)hipEventDestroy(e);
   }
   if (stage_host) (void)hipHostFree(stage_host);
   if (pinned) (void)hipHostFree(pinned);
   if (stream) (void)hipStreamDestroy(stream);
}

void* Context::stage(const void* host, size_t bytes) {
   if (bytes > kStageSlotBytes) {
      hipCheck(hipStreamSynchronize(stream), "synchronize");
      if (stage_big.bytes() < bytes) stage_big = DeviceBuffer(bytes);
      hipCheck(hipMemcpy(stage_big.as<void>(), host, bytes, hipMemcpyHostToDevice), "upload");
      return stage_big.as<void>();
   }
   const size_t k = stage_next;
   stage_next = (stage_next + 1) % kStageSlots;
   if (stage_used[k]) hipCheck(hipEventSynchronize(stage_events[k]), "staging slot wait");
   auto* h = static_cast<uint8_t*>(stage_host) + k * kStageSlotBytes;
   auto* d = stage_dev.as<uint8_t>() + k * kStageSlotBytes;
   std::memcpy(h, host, bytes);
   hipCheck(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream), "upload");
   hipCheck(hipEventRecord(stage_events[k], stream), "hipEventRecord");
   stage_used[k] = true;
   return d;
}

uint32_t maxFilterDepth(const Context& ctx) {
   uint32_t d = 1;
   while (filterLdsBytes(d + 1) <= ctx.max_lds_bytes) ++d;
   return d;
}

}  // namespace rdb

namespace {

using capi::g_error;
using capi::guarded;
using capi::requireFinalized;
using capi::tableOf;

void requireOpen(const Table& t) {
   if (t.finalized) fail(RDB_EINVAL, "table is already finalized");
}

bool nameTaken(const Table& t, const std::string& name) {
   return t.seqIndex(name) >= 0 || t.intIndex(name) >= 0 || t.dictColumn(name) != nullptr;
}

const DeviceBitmap& filterOf(const Table& t, const rdb_bitmap* f) {
   if (!f || !f->b.words) fail(RDB_EINVAL, "null filter bitmap");
   if (f->b.table != &t) fail(RDB_EINVAL, "filter bitmap belongs to another table");
   return f->b;
}

// Chunk groups of the Mutations kernel: at most kMaxGroups partial sums, and at least enough groups
// for the packed counters (at most 15 chunks per group).  When each wavefront takes its own chunks,
// a group per four chunks (one per wavefront); when the work-group shares each tile, a group per
// three and a half chunks (measured fastest on 62-chunk tables: 18 groups, every work-group
// resident at once at six wavefronts per SIMD).
uint32_t mutationGroups(const Table& t, const SeqColumn& /*col*/) {
   if (t.ctx->mutation_groups) return std::clamp<uint32_t>(t.ctx->mutation_groups, 1u, kMaxGroups);
   const uint32_t at_least = std::max<uint32_t>(1, (t.n_chunks + 14) / 15);
   const uint32_t groups = t.ctx->mutation_coop ? (2 * t.n_chunks + 6) / 7 : (t.n_chunks + 3) / 4;
   return std::clamp<uint32_t>(groups, at_least, kMaxGroups);
}

// Launch description of the Mutations kernels for column `idx` under filter `f`, with the
// context's dense/sparse switch and the table's scratch buffers.
MutationsLaunch mutationsLaunch(Table& t, int idx, const DeviceBitmap& f) {
   SeqColumn& col = *t.seq[idx];
   MutationsLaunch m;
   m.col_dev = t.seq_views_dev.as<SeqColumnView>() + idx;
   m.col_host = &col.view;
   m.chunk_sizes = t.chunk_sizes_dev.as<uint32_t>();
   m.n_chunks = t.n_chunks;
   m.n_groups = mutationGroups(t, col);
   m.unroll = t.ctx->mutation_unroll;
   m.coop = t.ctx->mutation_coop != 0;
   m.filter = f.words;
   m.filter_chunk_cards = f.chunk_cards;
   m.filter_card = f.card_dev;
   m.sparse_max = t.ctx->sparse_max_rows;
   m.partial = t.slot_counts.as<uint32_t>();
   m.heavy_sums = t.ctx->mutation_packed ? t.heavy_sums.as<uint32_t>() : nullptr;
   m.sparse_counts = col.sparse_counts.as<uint32_t>();
   m.sparse_bnd = col.sparse_bnd.as<int32_t>();
   m.totals = t.totals.as<int64_t>();
   return m;
}

}  // namespace

extern "C" {

int rdb_open(int device, rdb_ctx** out) {
   return guarded(nullptr, [&] {
      if (!out) fail(RDB_EINVAL, "null output");
      int n = 0;
      hipCheck(hipGetDeviceCount(&n), "hipGetDeviceCount");
      if (device < 0 || device >= n) fail(RDB_EINVAL, "no such device");
      auto* c = new rdb_ctx;
      try {
         c->c.init(device);
      } catch (...) {
         delete c;
         throw;
      }
      *out = c;
   });
}

void rdb_close(rdb_ctx* ctx) { delete ctx; }

const char* rdb_last_error(const rdb_ctx* ctx) { return ctx ? ctx->c.error.c_str() : g_error.c_str(); }

int rdb_table_create(rdb_ctx* ctx, const uint32_t* chunk_sizes, uint32_t n_chunks, rdb_table** out) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx || !out || (n_chunks && !chunk_sizes)) fail(RDB_EINVAL, "null argument");
      if (n_chunks >= 0xFFFF) fail(RDB_EINVAL, "too many chunks (row_layout.h:44 allows fewer than 65535)");
      auto* t = new rdb_table;
      Table& tb = t->t;
      tb.ctx = &ctx->c;
      uint64_t off = 0;
      for (uint32_t c = 0; c < n_chunks; ++c) {
         if (chunk_sizes[c] > kChunkRows) {
            delete t;
            fail(RDB_EINVAL, "a chunk holds at most 65536 rows");
         }
         tb.chunk_sizes.push_back(chunk_sizes[c]);
         tb.chunk_offset.push_back(off);
         off += chunk_sizes[c];
      }
      tb.chunk_offset.push_back(off);
      tb.n_chunks = n_chunks;
      tb.num_rows = off;
      *out = t;
   });
}

int rdb_device_memory(rdb_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx || !free_bytes || !total_bytes) fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      size_t f = 0, tot = 0;
      hipCheck(hipSetDevice(ctx->c.device), "hipSetDevice");
      hipCheck(hipMemGetInfo(&f, &tot), "hipMemGetInfo");
      *free_bytes = f;
      *total_bytes = tot;
   });
}

int rdb_measure_read_bandwidth(rdb_ctx* ctx, uint64_t bytes, int reps, double* gb_per_s) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx || !gb_per_s) fail(RDB_EINVAL, "null argument");
      if (bytes < 16 || bytes % 16 || reps < 1) fail(RDB_EINVAL, "bytes must be a positive multiple of 16, reps >= 1");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      hipCheck(hipSetDevice(ctx->c.device), "hipSetDevice");
      // enough work-groups to fill the device, none without work
      const uint64_t words = bytes / 16;
      const uint32_t blocks = static_cast<uint32_t>(std::max<uint64_t>(
         1, std::min<uint64_t>(static_cast<uint64_t>(ctx->c.cu_count) * 8, (words + 1023) / 1024)));
      DeviceBuffer src(bytes), out(static_cast<size_t>(blocks) * 256 * 4);
      hipStream_t s = ctx->c.stream;
      hipCheck(hipMemsetAsync(src.as<void>(), 0, bytes, s), "hipMemsetAsync");
      hipEvent_t a, b;
      hipCheck(hipEventCreate(&a), "hipEventCreate");
      hipCheck(hipEventCreate(&b), "hipEventCreate");
      hipCheck(launchStreamRead(src.as<void>(), bytes, out.as<uint32_t>(), blocks, s), "stream read");
      hipCheck(hipEventRecord(a, s), "hipEventRecord");
      for (int r = 0; r < reps; ++r) {
         hipCheck(launchStreamRead(src.as<void>(), bytes, out.as<uint32_t>(), blocks, s), "stream read");
      }
      hipCheck(hipEventRecord(b, s), "hipEventRecord");
      hipCheck(hipEventSynchronize(b), "hipEventSynchronize");
      float ms = 0;
      hipCheck(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      *gb_per_s = static_cast<double>(bytes) * reps / (ms * 1e-3) / 1e9;
   });
}

int rdb_set_tuning(rdb_ctx* ctx, const char* key, int64_t value) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx || !key) fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      const std::string k(key);
      if (k == "sparse_max_rows") {
         ctx->c.sparse_max_rows = value < 0 ? -1 : static_cast<long long>(value);
      } else if (k == "mutation_groups") {
         if (value < 0 || value > kMaxGroups) fail(RDB_EINVAL, "mutation_groups must be in [0, 32]");
         ctx->c.mutation_groups = static_cast<uint32_t>(value);
      } else if (k == "mutation_unroll") {
         if (value != 2 && value != 4 && value != 8 && value != 16) {
            fail(RDB_EINVAL, "mutation_unroll must be 2, 4, 8 or 16");
         }
         ctx->c.mutation_unroll = static_cast<uint32_t>(value);
      } else if (k == "mutation_packed") {
         ctx->c.mutation_packed = value != 0;
      } else if (k == "mutation_coop") {
         ctx->c.mutation_coop = value != 0;
      } else if (k == "batch_max_queries") {
         if (value < 1 || value > 4096) fail(RDB_EINVAL, "batch_max_queries must be in [1, 4096]");
         ctx->c.batch_max_queries = static_cast<uint32_t>(value);
      } else if (k == "batch_lanes") {
         if (value < 1 || value > 16) fail(RDB_EINVAL, "batch_lanes must be in [1, 16]");
         ctx->c.batch_lanes = static_cast<uint32_t>(value);
      } else if (k == "batch_graphs") {
         ctx->c.batch_graphs = value != 0;
      } else if (k == "batch_shared_scan") {
         if (value < 0 || value > 2) fail(RDB_EINVAL, "batch_shared_scan must be 0, 1 or 2");
         ctx->c.batch_shared_scan = static_cast<uint32_t>(value);
      } else if (k == "batch_octet_threads") {
         if (value != 512 && value != 1024) fail(RDB_EINVAL, "batch_octet_threads must be 512 or 1024");
         ctx->c.batch_octet_threads = static_cast<uint32_t>(value);
      } else if (k == "batch_octet_groups") {
         if (value < 0 || value > 4096) fail(RDB_EINVAL, "batch_octet_groups must be in [0, 4096]");
         ctx->c.batch_octet_groups = static_cast<uint32_t>(value);
      } else if (k == "batch_row_capacity") {
         if (value < 16 || value > (1 << 20)) fail(RDB_EINVAL, "batch_row_capacity must be in [16, 2^20]");
         ctx->c.batch_row_capacity = static_cast<uint32_t>(value);
      } else if (k == "batch_fail_job") {
         if (value < 0 || value > 1000000) fail(RDB_EINVAL, "batch_fail_job must be in [0, 10^6]");
         ctx->c.batch_fail_job = static_cast<uint32_t>(value);
      } else if (k == "index_build_device") {
         if (value != 0 && value != 1) fail(RDB_EINVAL, "index_build_device must be 0 or 1");
         ctx->c.index_build_device = static_cast<uint32_t>(value);
      } else {
         fail(RDB_EINVAL, "unknown tuning key '" + k + "'");
      }
   });
}

int rdb_get_tuning(rdb_ctx* ctx, const char* key, int64_t* value) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx || !key || !value) fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      const std::string k(key);
      if (k == "sparse_max_rows") {
         *value = ctx->c.sparse_max_rows;
      } else if (k == "mutation_groups") {
         *value = ctx->c.mutation_groups;
      } else if (k == "mutation_unroll") {
         *value = ctx->c.mutation_unroll;
      } else if (k == "mutation_packed") {
         *value = ctx->c.mutation_packed;
      } else if (k == "mutation_coop") {
         *value = ctx->c.mutation_coop;
      } else if (k == "batch_max_queries") {
         *value = ctx->c.batch_max_queries;
      } else if (k == "batch_lanes") {
         *value = ctx->c.batch_lanes;
      } else if (k == "batch_graphs") {
         *value = ctx->c.batch_graphs;
      } else if (k == "batch_shared_scan") {
         *value = ctx->c.batch_shared_scan;
      } else if (k == "batch_octet_threads") {
         *value = ctx->c.batch_octet_threads;
      } else if (k == "batch_octet_groups") {
         *value = ctx->c.batch_octet_groups;
      } else if (k == "batch_row_capacity") {
         *value = ctx->c.batch_row_capacity;
      } else if (k == "batch_fail_job") {
         *value = ctx->c.batch_fail_job;
      } else if (k == "index_build_device") {
         *value = ctx->c.index_build_device;
      } else {
         fail(RDB_EINVAL, "unknown tuning key '" + k + "'");
      }
   });
}

void rdb_table_free(rdb_table* table) {
   if (!table) return;
   Context* ctx = table->t.ctx;
   std::lock_guard<std::recursive_mutex> lock(ctx->mu);
   (void)hipStreamSynchronize(ctx->stream);
   delete table;
}

int rdb_table_add_sequence_column(rdb_table* table, const char* name, int alphabet, const uint8_t* reference,
                                  uint32_t genome_length, const uint8_t* is_null, const uint32_t* cov_start,
                                  const uint32_t* cov_end, const uint64_t* mut_ptr, const uint32_t* mut_pos,
                                  const uint8_t* mut_sym, const uint64_t* nrun_ptr, const uint32_t* nrun_start,
                                  const uint32_t* nrun_end) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireOpen(t);
      if (!name || !reference || !is_null || !cov_start || !cov_end || !mut_ptr || !nrun_ptr) fail(RDB_EINVAL, "null argument");
      if (nameTaken(t, name)) fail(RDB_EINVAL, std::string("duplicate column '") + name + "'");
      if (alphabet != RDB_NUCLEOTIDE && alphabet != RDB_AMINO_ACID) fail(RDB_EINVAL, "unknown alphabet");
      auto col = std::make_unique<SeqColumn>();
      col->name = name;
      col->alphabet = alphabet;
      col->abc = &alphabetFor(alphabet);
      col->genome_length = genome_length;
      col->reference.assign(reference, reference + genome_length);
      buildSequenceColumn(t, *col, is_null, cov_start, cov_end, mut_ptr, mut_pos, mut_sym, nrun_ptr, nrun_start, nrun_end);
      t.seq.push_back(std::move(col));
   });
}

extern "C++" {
namespace {
int addFinalized(rdb_table* table, const char* name, int alphabet, const uint8_t* reference, uint32_t genome_length,
                 const FinalizedSequence& f) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireOpen(t);
      if (!name || !reference) fail(RDB_EINVAL, "null argument");
      if (nameTaken(t, name)) fail(RDB_EINVAL, std::string("duplicate column '") + name + "'");
      if (alphabet != RDB_NUCLEOTIDE && alphabet != RDB_AMINO_ACID) fail(RDB_EINVAL, "unknown alphabet");
      auto col = std::make_unique<SeqColumn>();
      col->name = name;
      col->alphabet = alphabet;
      col->abc = &alphabetFor(alphabet);
      col->genome_length = genome_length;
      col->reference.assign(reference, reference + genome_length);
      buildSequenceColumnFinalized(t, *col, f);
      t.seq.push_back(std::move(col));
   });
}

FinalizedSequence finalizedIndex(const uint8_t* local_reference, uint64_t n_containers, const uint32_t* key_position,
                                 const uint16_t* key_chunk, const uint8_t* key_symbol, const uint8_t* typecode,
                                 const uint32_t* cardinality, const uint64_t* payload_offset, const uint8_t* payload,
                                 uint64_t payload_bytes, const uint8_t* is_null, const uint32_t* cov_start,
                                 const uint32_t* cov_end) {
   FinalizedSequence f;
   f.local_reference = local_reference;
   f.n_containers = n_containers;
   f.key_position = key_position;
   f.key_chunk = key_chunk;
   f.key_symbol = key_symbol;
   f.typecode = typecode;
   f.cardinality = cardinality;
   f.payload_offset = payload_offset;
   f.payload = payload;
   f.payload_bytes = payload_bytes;
   f.is_null = is_null;
   f.cov_start = cov_start;
   f.cov_end = cov_end;
   return f;
}
}  // namespace
}  // extern "C++"

int rdb_table_add_sequence_column_finalized(rdb_table* table, const char* name, int alphabet, const uint8_t* reference,
                                            const uint8_t* local_reference, uint32_t genome_length, uint64_t n_containers,
                                            const uint32_t* key_position, const uint16_t* key_chunk,
                                            const uint8_t* key_symbol, const uint8_t* typecode,
                                            const uint32_t* cardinality, const uint64_t* payload_offset,
                                            const uint8_t* payload, uint64_t payload_bytes, const uint8_t* is_null,
                                            const uint32_t* cov_start, const uint32_t* cov_end,
                                            const uint64_t* missing_ptr, const uint32_t* missing_pos) {
   FinalizedSequence f = finalizedIndex(local_reference, n_containers, key_position, key_chunk, key_symbol, typecode,
                                        cardinality, payload_offset, payload, payload_bytes, is_null, cov_start, cov_end);
   f.missing_ptr = missing_ptr;
   f.missing_pos = missing_pos;
   return addFinalized(table, name, alphabet, reference, genome_length, f);
}

int rdb_table_add_sequence_column_finalized_runs(
   rdb_table* table, const char* name, int alphabet, const uint8_t* reference, const uint8_t* local_reference,
   uint32_t genome_length, uint64_t n_containers, const uint32_t* key_position, const uint16_t* key_chunk,
   const uint8_t* key_symbol, const uint8_t* typecode, const uint32_t* cardinality, const uint64_t* payload_offset,
   const uint8_t* payload, uint64_t payload_bytes, const uint8_t* is_null, const uint32_t* cov_start,
   const uint32_t* cov_end, const uint64_t* missing_run_ptr, const uint32_t* missing_run_start,
   const uint32_t* missing_run_end) {
   FinalizedSequence f = finalizedIndex(local_reference, n_containers, key_position, key_chunk, key_symbol, typecode,
                                        cardinality, payload_offset, payload, payload_bytes, is_null, cov_start, cov_end);
   if (!missing_run_ptr) {
      Context* ctx = table ? table->t.ctx : nullptr;
      return guarded(ctx, [] { fail(RDB_EINVAL, "null argument"); });
   }
   f.run_ptr = missing_run_ptr;
   f.run_start = missing_run_start;
   f.run_end = missing_run_end;
   return addFinalized(table, name, alphabet, reference, genome_length, f);
}

struct rdb_seq_builder {
   rdb::SeqBuilder b;
};

int rdb_seq_builder_create(rdb_table* table, const char* name, int alphabet, const uint8_t* reference,
                           uint32_t genome_length, rdb_seq_builder** out) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireOpen(t);
      if (!name || !reference || !out) fail(RDB_EINVAL, "null argument");
      *out = nullptr;
      if (nameTaken(t, name)) fail(RDB_EINVAL, std::string("duplicate column '") + name + "'");
      if (alphabet != RDB_NUCLEOTIDE && alphabet != RDB_AMINO_ACID) fail(RDB_EINVAL, "unknown alphabet");
      if (genome_length == 0 || genome_length > 65535) fail(RDB_EINVAL, "genome length must be in [1, 65535]");
      auto h = std::make_unique<rdb_seq_builder>();
      SeqBuilder& b = h->b;
      b.table = &t;
      b.col = std::make_unique<SeqColumn>();
      b.col->name = name;
      b.col->alphabet = alphabet;
      b.col->abc = &alphabetFor(alphabet);
      b.col->genome_length = genome_length;
      b.col->reference.assign(reference, reference + genome_length);
      for (uint32_t p = 0; p < genome_length; ++p) {
         if (reference[p] >= b.col->abc->count) fail(RDB_EINVAL, "invalid reference symbol");
      }
      uint8_t lut[256];
      for (int c = 0; c < 256; ++c) {
         const int sym = b.col->abc->fromChar(static_cast<char>(c));
         lut[c] = sym < 0 ? 0xFF : static_cast<uint8_t>(sym);
      }
      b.lut = DeviceBuffer(lut, sizeof(lut));
      b.ref = DeviceBuffer(reference, genome_length);
      *out = h.release();
   });
}

int rdb_seq_builder_append(rdb_seq_builder* builder, const char* data, const uint64_t* row_ptr, const uint32_t* offsets,
                           const uint8_t* is_null, uint64_t n_rows) {
   Context* ctx = builder ? builder->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!builder) fail(RDB_EINVAL, "null builder");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      seqBuilderAppend(builder->b, data, row_ptr, offsets, is_null, n_rows);
   });
}

int rdb_seq_builder_finish(rdb_seq_builder* builder) {
   Context* ctx = builder ? builder->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!builder) fail(RDB_EINVAL, "null builder");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      Table& t = *builder->b.table;
      requireOpen(t);
      if (!builder->b.finished && nameTaken(t, builder->b.col->name)) {
         fail(RDB_EINVAL, std::string("duplicate column '") + builder->b.col->name + "'");
      }
      seqBuilderFinish(builder->b);
   });
}

void rdb_seq_builder_free(rdb_seq_builder* builder) { delete builder; }

int rdb_table_column_digest(rdb_table* table, const char* column, uint64_t* digest) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      if (!column || !digest) fail(RDB_EINVAL, "null argument");
      *digest = seqColumnDigest(t, t.seqColumn(column));
   });
}

int rdb_table_add_int_column(rdb_table* table, const char* name, int is_date, const int32_t* values, const uint8_t* is_null) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireOpen(t);
      if (!name || (t.num_rows && (!values || !is_null))) fail(RDB_EINVAL, "null argument");
      if (nameTaken(t, name)) fail(RDB_EINVAL, std::string("duplicate column '") + name + "'");
      auto col = std::make_unique<IntColumn>();
      col->name = name;
      col->is_date = is_date != 0;
      buildIntColumn(t, *col, values, is_null);
      t.ints.push_back(std::move(col));
   });
}

int rdb_table_add_dict_column(rdb_table* table, const char* name, const uint32_t* ids, const uint8_t* is_null,
                              uint32_t n_values, const char* names, const uint32_t* name_offsets,
                              const int32_t* lineage_parent) {
   return rdb_table_add_lineage_column(table, name, ids, is_null, n_values, names, name_offsets, lineage_parent,
                                       nullptr);
}

int rdb_table_add_lineage_column(rdb_table* table, const char* name, const uint32_t* ids, const uint8_t* is_null,
                                 uint32_t n_values, const char* names, const uint32_t* name_offsets,
                                 const int32_t* lineage_parent, const int32_t* lineage_alias) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireOpen(t);
      if (!name || (t.num_rows && (!ids || !is_null)) || (n_values && (!names || !name_offsets))) fail(RDB_EINVAL, "null argument");
      if (nameTaken(t, name)) fail(RDB_EINVAL, std::string("duplicate column '") + name + "'");
      auto col = std::make_unique<DictColumn>();
      col->name = name;
      for (uint32_t v = 0; v < n_values; ++v) {
         if (name_offsets[v + 1] < name_offsets[v]) fail(RDB_EINVAL, "bad dictionary offsets");
         col->dictionary.emplace_back(names + name_offsets[v], names + name_offsets[v + 1]);
         if (!col->ids.emplace(col->dictionary.back(), v).second) fail(RDB_EINVAL, "duplicate dictionary value");
      }
      if (lineage_alias && !lineage_parent) fail(RDB_EINVAL, "lineage aliases need the lineage parents");
      buildDictColumn(t, *col, ids, is_null, lineage_parent, lineage_alias);
      t.dicts.push_back(std::move(col));
   });
}

int rdb_table_add_bitmap(rdb_table* table, const uint32_t* rows, uint64_t n, uint32_t* id_out) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireOpen(t);
      if ((n && !rows) || !id_out) fail(RDB_EINVAL, "null argument");
      std::vector<uint32_t> v(rows, rows + n);
      t.literal_bitmaps.push_back(t.addStoreBitmap(v));
      *id_out = static_cast<uint32_t>(t.literal_bitmaps.size() - 1);
   });
}

int rdb_table_finalize(rdb_table* table) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireOpen(t);
      finalizeTable(t);
   });
}

int rdb_table_local_reference(const rdb_table* table, const char* column, uint8_t* out) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      if (!table || !column || !out) fail(RDB_EINVAL, "null argument");
      const SeqColumn& col = table->t.seqColumn(column);
      std::copy(col.local_ref.begin(), col.local_ref.end(), out);
   });
}

int rdb_table_build_profile(const rdb_table* table, const char* column, double* phase_ms, uint32_t n_phases) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      if (!table || !column || (!phase_ms && n_phases)) fail(RDB_EINVAL, "null argument");
      const SeqColumn& col = table->t.seqColumn(column);
      for (uint32_t k = 0; k < n_phases; ++k) phase_ms[k] = k < kBuildPhases ? col.build_ms[k] : 0.0;
   });
}

int rdb_table_stats(const rdb_table* table, const char* column, uint64_t* device_bytes, uint64_t* n_containers,
                    uint64_t* payload_bytes) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      if (!table) fail(RDB_EINVAL, "null table");
      const Table& t = table->t;
      uint64_t bytes = 0;
      for (const auto& c : t.seq) {
         for (const auto& b : c->buffers) bytes += b.bytes();
      }
      for (const auto& c : t.ints) {
         for (const auto& b : c->buffers) bytes += b.bytes();
      }
      for (const auto& c : t.dicts) bytes += c->labels.bytes();
      for (const auto& b : t.store_buffers) bytes += b.bytes();
      if (device_bytes) *device_bytes = bytes;
      if (column) {
         const SeqColumn& col = t.seqColumn(column);
         if (n_containers) *n_containers = col.n_containers;
         if (payload_bytes) *payload_bytes = col.payload_bytes;
      }
   });
}

int rdb_filter(rdb_table* table, const char* filter, rdb_bitmap** out, uint64_t* cardinality) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!filter || !out) fail(RDB_EINVAL, "null argument");
      auto op = query::compileFilter(t, filter);
      auto* bm = new rdb_bitmap;
      bm->b = op->evaluate();
      if (cardinality) {
         try {
            *cardinality = bm->b.cardinality();
         } catch (...) {
            delete bm;
            throw;
         }
      }
      *out = bm;
   });
}

int rdb_bitmap_cardinality(const rdb_bitmap* bitmap, uint64_t* cardinality) {
   return guarded(nullptr, [&] {
      if (!bitmap || !cardinality || !bitmap->b.table) fail(RDB_EINVAL, "null argument");
      auto& b = const_cast<DeviceBitmap&>(bitmap->b);
      std::lock_guard<std::recursive_mutex> lock(b.table->ctx->mu);
      *cardinality = b.cardinality();
   });
}

int rdb_bitmap_rows(rdb_bitmap* bitmap, uint32_t* rows, uint64_t capacity, uint64_t* n) {
   return guarded(nullptr, [&] {
      if (!bitmap || !n || !bitmap->b.table) fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(bitmap->b.table->ctx->mu);
      const auto v = bitmap->b.rows();
      *n = v.size();
      if (v.size() > capacity) fail(RDB_ENOMEM, "output capacity too small");
      if (!v.empty()) std::copy(v.begin(), v.end(), rows);
   });
}

void rdb_bitmap_free(rdb_bitmap* bitmap) {
   if (!bitmap) return;
   if (bitmap->b.table) {
      std::lock_guard<std::recursive_mutex> lock(bitmap->b.table->ctx->mu);
      delete bitmap;
   } else {
      delete bitmap;
   }
}

int rdb_filter_explain(rdb_table* table, const char* filter, char* out, uint64_t capacity, uint64_t* length) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!filter || !length) fail(RDB_EINVAL, "null argument");
      const std::string text = query::compileFilter(t, filter)->toString();
      *length = text.size();
      if (out == nullptr || capacity <= text.size()) fail(RDB_ENOMEM, "explain: output capacity too small");
      std::memcpy(out, text.c_str(), text.size() + 1);
   });
}

int rdb_count(rdb_table* table, const char* filter, uint64_t* count) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!filter || !count) fail(RDB_EINVAL, "null argument");
      auto op = query::compileFilter(t, filter);
      DeviceBitmap b = op->evaluate();
      *count = b.cardinality();
   });
}

int rdb_mutation_counts(rdb_table* table, const rdb_bitmap* filter, const char* column, uint32_t* counts) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!column || !counts) fail(RDB_EINVAL, "null argument");
      const DeviceBitmap& f = filterOf(t, filter);
      const int idx = t.seqIndex(column);
      if (idx < 0) fail(RDB_EQUERY, std::string("Database does not contain the sequence '") + column + "'");
      const SeqColumn& col = *t.seq[idx];
      const size_t bytes = static_cast<size_t>(col.abc->count) * col.genome_length * 4;
      void* dev = t.ctx->pool.get(bytes);
      try {
         MutationsLaunch m = mutationsLaunch(t, idx, f);
         m.counts_out = static_cast<uint32_t*>(dev);
         m.n_rows = t.mut_rows.as<uint32_t>();
         hipCheck(launchMutations(m, t.ctx->stream), "mutations kernels");
         hipCheck(hipMemcpyAsync(counts, dev, bytes, hipMemcpyDeviceToHost, t.ctx->stream), "copy counts");
         hipCheck(hipStreamSynchronize(t.ctx->stream), "synchronize");
      } catch (...) {
         t.ctx->pool.put(dev, bytes);
         throw;
      }
      t.ctx->pool.put(dev, bytes);
   });
}

}  // extern "C"

namespace rdb::capi {

// Mutations rows under a device filter: kernels, one readback, reference output order
// (mutations_node.cpp:292-366).
std::vector<MutationRowDev> mutationRows(Table& t, const DeviceBitmap& f, int idx, double min_proportion) {
   // device buffer: a 16-byte header holding the row count, then the rows
   auto* base = t.mut_rows.as<uint8_t>();
   auto* n_dev = reinterpret_cast<uint32_t*>(base);
   auto* rows_dev = reinterpret_cast<MutationRowDev*>(base + 16);
   const uint32_t row_cap = static_cast<uint32_t>((t.mut_rows.bytes() - 16) / sizeof(MutationRowDev));
   MutationsLaunch m = mutationsLaunch(t, idx, f);
   m.min_proportion = min_proportion;
   m.rows = rows_dev;
   m.row_capacity = row_cap;
   m.n_rows = n_dev;
   if (t.ctx->profiling) {
      m.count_start = t.ctx->prof_events[0];
      m.count_stop = t.ctx->prof_events[1];
   }
   hipCheck(launchMutations(m, t.ctx->stream), "mutations kernels");
   if (t.ctx->profiling) t.ctx->profileAccumulate(0);
   // one copy brings the count and the first rows; a second one only for long outputs
   constexpr size_t kFirst = 2048;
   const size_t first_bytes = std::min(16 + kFirst * sizeof(MutationRowDev), t.mut_rows.bytes());
   auto* host = static_cast<uint8_t*>(t.ctx->pinnedBuffer(first_bytes));
   hipCheck(hipMemcpyAsync(host, base, first_bytes, hipMemcpyDeviceToHost, t.ctx->stream), "copy rows");
   hipCheck(hipStreamSynchronize(t.ctx->stream), "synchronize");
   uint32_t rows_n = 0;
   std::memcpy(&rows_n, host, 4);
   if (rows_n > row_cap) fail(RDB_EDEVICE, "internal: mutation row buffer overflow");
   std::vector<MutationRowDev> rows(rows_n);
   const size_t in_first = std::min<size_t>(rows_n, (first_bytes - 16) / sizeof(MutationRowDev));
   if (in_first) std::memcpy(rows.data(), host + 16, in_first * sizeof(MutationRowDev));
   if (rows_n > in_first) {
      hipCheck(hipMemcpyAsync(rows.data() + in_first, rows_dev + in_first, (rows_n - in_first) * sizeof(MutationRowDev),
                              hipMemcpyDeviceToHost, t.ctx->stream),
               "copy rows");
      hipCheck(hipStreamSynchronize(t.ctx->stream), "synchronize");
   }
   sortMutationRows(rows);
   return rows;
}

void sortMutationRows(std::vector<MutationRowDev>& rows) {
   // mutations_node.cpp:302-364 emits rows by position, then in VALID_MUTATION_SYMBOLS order
   std::sort(rows.begin(), rows.end(), [](const MutationRowDev& a, const MutationRowDev& b) {
      return a.position != b.position ? a.position < b.position : a.symbol < b.symbol;
   });
}

void emitMutationRows(const SeqColumn& col, const std::vector<MutationRowDev>& rows, uint32_t* position, uint8_t* from,
                      uint8_t* to, uint32_t* count, uint32_t* coverage, double* proportion, uint64_t capacity,
                      uint64_t* n) {
   *n = rows.size();
   if (rows.size() > capacity) fail(RDB_ENOMEM, "output capacity too small");
   for (size_t i = 0; i < rows.size(); ++i) {
      if (position) position[i] = rows[i].position + 1;
      if (from) from[i] = col.reference[rows[i].position];
      if (to) to[i] = static_cast<uint8_t>(rows[i].symbol);
      if (count) count[i] = rows[i].count;
      if (coverage) coverage[i] = rows[i].coverage;
      if (proportion) proportion[i] = static_cast<double>(rows[i].count) / static_cast<double>(rows[i].coverage);
   }
}

int mutationColumn(const Table& t, const char* column, double min_proportion) {
   if (!(min_proportion >= 0.0 && min_proportion <= 1.0)) fail(RDB_EQUERY, "minProportion must be in interval [0.0,1.0]");
   const int idx = t.seqIndex(column);
   if (idx < 0) fail(RDB_EQUERY, std::string("Database does not contain the sequence '") + column + "'");
   return idx;
}

}  // namespace rdb::capi

namespace {

void runMutations(Table& t, const DeviceBitmap& f, const char* column, double min_proportion, uint32_t* position,
                  uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage, double* proportion,
                  uint64_t capacity, uint64_t* n) {
   const int idx = capi::mutationColumn(t, column, min_proportion);
   const auto rows = capi::mutationRows(t, f, idx, min_proportion);
   capi::emitMutationRows(*t.seq[idx], rows, position, from, to, count, coverage, proportion, capacity, n);
}

}  // namespace

extern "C" {

int rdb_mutations(rdb_table* table, const rdb_bitmap* filter, const char* column, double min_proportion,
                  uint32_t* position, uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage,
                  double* proportion, uint64_t capacity, uint64_t* n) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!column || !n) fail(RDB_EINVAL, "null argument");
      runMutations(t, filterOf(t, filter), column, min_proportion, position, from, to, count, coverage, proportion,
                   capacity, n);
   });
}

int rdb_mutations_query(rdb_table* table, const char* filter, const char* column, double min_proportion,
                        uint32_t* position, uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage,
                        double* proportion, uint64_t capacity, uint64_t* n) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!filter || !column || !n) fail(RDB_EINVAL, "null argument");
      DeviceBitmap f = query::compileFilter(t, filter)->evaluate();
      runMutations(t, f, column, min_proportion, position, from, to, count, coverage, proportion, capacity, n);
   });
}

// Mutations query whose kernels use only the device's file-scope scratch array: the staged filter
// program, the filter words, partial sums, heavy sums, totals, sparse accumulators and rows are
// carved from it (256-byte aligned); nothing is allocated on the device or pinned on the host.
int rdb_mutations_query_noalloc(rdb_table* table, const char* filter, const char* column, double min_proportion,
                                uint32_t* position, uint8_t* from, uint8_t* to, uint32_t* count, uint32_t* coverage,
                                double* proportion, uint64_t capacity, uint64_t* n) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!filter || !column || !n) fail(RDB_EINVAL, "null argument");
      if (t.ctx->noalloc_scratch == nullptr) fail(RDB_EDEVICE, "allocation-free scratch not resolved");
      const int idx = capi::mutationColumn(t, column, min_proportion);
      SeqColumn& col = *t.seq[idx];
      query::Emitter program;
      query::compileFilter(t, filter)->emit(program);
      query::programDepth(*t.ctx, program);
      const std::vector<uint8_t> blob = query::programBlob(program);
      const uint32_t L = col.genome_length;
      const uint32_t n_slots = col.view.n_slots;
      const uint32_t n_pblocks = (L + kMutPosBlock - 1) / kMutPosBlock;
      const size_t row_cap = std::max<size_t>(static_cast<size_t>(L) * 32, 2048);
      // carve the scratch; fewer chunk groups (partial sums) when the table needs it
      struct Carve {
         size_t blob, words, card, partial, heavy, totals, sparse_counts, sparse_bnd, rows, end;
      } c{};
      auto layout = [&](uint32_t groups) {
         size_t o = 0;
         auto take = [&](size_t bytes) {
            const size_t at = o;
            o = (o + bytes + 255) / 256 * 256;
            return at;
         };
         c.blob = take(blob.size());
         c.words = take(static_cast<size_t>(t.n_chunks) * kChunkWords * 8);
         c.card = take(64 + 4ULL * t.n_chunks);  // cardinality, then the per-chunk cardinalities
         c.partial = take(4ULL * groups * n_slots);
         c.heavy = take(4ULL * n_slots);
         c.totals = take(static_cast<size_t>(n_pblocks) * 3 * sizeof(int64_t));
         c.sparse_counts = take(4ULL * L * col.abc->count);
         c.sparse_bnd = take(4ULL * 3 * L);
         c.rows = take(16 + row_cap * sizeof(MutationRowDev));
         c.end = o;
      };
      uint32_t groups = mutationGroups(t, col);
      layout(groups);
      while (c.end > kNoAllocScratchBytes && groups > 1) {
         groups = (groups + 1) / 2;
         layout(groups);
      }
      if (c.end > kNoAllocScratchBytes) {
         fail(RDB_EINVAL, "table too large for rdb_mutations_query_noalloc: needs " + std::to_string(c.end) +
                             " bytes of scratch, the limit is " + std::to_string(kNoAllocScratchBytes));
      }
      auto* base = static_cast<uint8_t*>(t.ctx->noalloc_scratch);
      hipStream_t stream = t.ctx->stream;
      hipCheck(hipMemcpyAsync(base + c.blob, blob.data(), blob.size(), hipMemcpyHostToDevice, stream), "stage program");
      auto* card = reinterpret_cast<unsigned long long*>(base + c.card);
      auto* chunk_cards = reinterpret_cast<uint32_t*>(base + c.card + 64);
      query::launchProgram(t, program, base + c.blob, reinterpret_cast<uint64_t*>(base + c.words), card, chunk_cards);
      // the accumulators that must start at zero: heavy sums and the sparse arrays (earlier calls
      // leave whatever they left)
      hipCheck(launchFill(reinterpret_cast<uint32_t*>(base + c.heavy), 0, n_slots, stream), "clear heavy sums");
      hipCheck(launchFill(reinterpret_cast<uint32_t*>(base + c.sparse_counts), 0, static_cast<size_t>(L) * col.abc->count,
                          stream),
               "clear sparse counts");
      hipCheck(launchFill(reinterpret_cast<uint32_t*>(base + c.sparse_bnd), 0, 3ULL * L, stream), "clear sparse boundaries");
      MutationsLaunch m;
      m.col_dev = t.seq_views_dev.as<SeqColumnView>() + idx;
      m.col_host = &col.view;
      m.chunk_sizes = t.chunk_sizes_dev.as<uint32_t>();
      m.n_chunks = t.n_chunks;
      m.n_groups = groups;
      m.unroll = t.ctx->mutation_unroll;
      m.coop = t.ctx->mutation_coop != 0;
      m.filter = reinterpret_cast<const uint64_t*>(base + c.words);
      m.filter_chunk_cards = chunk_cards;
      m.filter_card = card;
      m.sparse_max = t.ctx->sparse_max_rows;
      m.partial = reinterpret_cast<uint32_t*>(base + c.partial);
      m.heavy_sums = t.ctx->mutation_packed ? reinterpret_cast<uint32_t*>(base + c.heavy) : nullptr;
      m.sparse_counts = reinterpret_cast<uint32_t*>(base + c.sparse_counts);
      m.sparse_bnd = reinterpret_cast<int32_t*>(base + c.sparse_bnd);
      m.totals = reinterpret_cast<int64_t*>(base + c.totals);
      m.min_proportion = min_proportion;
      m.n_rows = reinterpret_cast<uint32_t*>(base + c.rows);
      m.rows = reinterpret_cast<MutationRowDev*>(base + c.rows + 16);
      m.row_capacity = static_cast<uint32_t>(row_cap);
      hipCheck(launchMutations(m, stream), "mutations kernels");
      uint32_t rows_n = 0;
      hipCheck(hipMemcpyAsync(&rows_n, base + c.rows, 4, hipMemcpyDeviceToHost, stream), "copy row count");
      hipCheck(hipStreamSynchronize(stream), "synchronize");
      if (rows_n > row_cap) fail(RDB_EDEVICE, "internal: mutation row buffer overflow");
      std::vector<MutationRowDev> rows(rows_n);
      if (rows_n) {
         hipCheck(hipMemcpyAsync(rows.data(), base + c.rows + 16, rows_n * sizeof(MutationRowDev), hipMemcpyDeviceToHost,
                                 stream),
                  "copy rows");
         hipCheck(hipStreamSynchronize(stream), "synchronize");
      }
      capi::sortMutationRows(rows);
      capi::emitMutationRows(col, rows, position, from, to, count, coverage, proportion, capacity, n);
   });
}

int rdb_mutation_rows_ndjson(int alphabet, const char* sequence_name, uint32_t fields, const uint32_t* position,
                             const uint8_t* from, const uint8_t* to, const uint32_t* count, const uint32_t* coverage,
                             const double* proportion, uint64_t n, char* out, uint64_t capacity, uint64_t* length) {
   return guarded(nullptr, [&] {
      if (!length || !sequence_name) fail(RDB_EINVAL, "null argument");
      if (n && (!position || !from || !to || !count || !coverage || !proportion)) fail(RDB_EINVAL, "null row array");
      if (alphabet != RDB_NUCLEOTIDE && alphabet != RDB_AMINO_ACID) fail(RDB_EINVAL, "unknown alphabet");
      capi::MutationRowsView v{&alphabetFor(alphabet), sequence_name, capi::validFields(fields), position, from, to, count,
                               coverage, proportion, n};
      capi::checkSymbols(v);
      const std::string text = capi::mutationRowsNdjson(v);
      *length = text.size();
      if (text.size() > capacity || (!out && !text.empty())) fail(RDB_ENOMEM, "output capacity too small");
      if (!text.empty()) std::memcpy(out, text.data(), text.size());
   });
}

int rdb_mutation_rows_arrow_ipc(int alphabet, const char* sequence_name, uint32_t fields, const uint32_t* position,
                                const uint8_t* from, const uint8_t* to, const uint32_t* count, const uint32_t* coverage,
                                const double* proportion, uint64_t n, uint8_t* out, uint64_t capacity, uint64_t* length) {
   return guarded(nullptr, [&] {
      if (!length || !sequence_name) fail(RDB_EINVAL, "null argument");
      if (n && (!position || !from || !to || !count || !coverage || !proportion)) fail(RDB_EINVAL, "null row array");
      if (alphabet != RDB_NUCLEOTIDE && alphabet != RDB_AMINO_ACID) fail(RDB_EINVAL, "unknown alphabet");
      capi::MutationRowsView v{&alphabetFor(alphabet), sequence_name, capi::validFields(fields), position, from, to, count,
                               coverage, proportion, n};
      capi::checkSymbols(v);
      const std::vector<uint8_t> bytes = capi::mutationRowsArrowIpc(v);
      *length = bytes.size();
      if (bytes.size() > capacity || !out) fail(RDB_ENOMEM, "output capacity too small");
      std::memcpy(out, bytes.data(), bytes.size());
   });
}

extern "C++" {
namespace {

// The Mutations rows of `filter` on `column` as row arrays for the sinks.
struct RowArrays {
   std::vector<uint32_t> position, count, coverage;
   std::vector<uint8_t> from, to;
   std::vector<double> proportion;
   capi::MutationRowsView view(const SeqColumn& col, uint32_t fields) const {
      return {col.abc, col.name.c_str(), capi::validFields(fields), position.data(), from.data(), to.data(),
              count.data(), coverage.data(), proportion.data(), position.size()};
   }
};

RowArrays mutationRowArrays(Table& t, const DeviceBitmap& f, int idx, double min_proportion) {
   const auto rows = capi::mutationRows(t, f, idx, min_proportion);
   RowArrays a;
   const size_t n = rows.size();
   a.position.resize(n);
   a.count.resize(n);
   a.coverage.resize(n);
   a.from.resize(n);
   a.to.resize(n);
   a.proportion.resize(n);
   uint64_t k = 0;
   capi::emitMutationRows(*t.seq[idx], rows, a.position.data(), a.from.data(), a.to.data(), a.count.data(),
                          a.coverage.data(), a.proportion.data(), n, &k);
   return a;
}

}  // namespace
}  // extern "C++"

int rdb_mutations_ndjson(rdb_table* table, const rdb_bitmap* filter, const char* column, double min_proportion,
                         uint32_t fields, char* out, uint64_t capacity, uint64_t* length) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!column || !length) fail(RDB_EINVAL, "null argument");
      capi::validFields(fields);
      const int idx = capi::mutationColumn(t, column, min_proportion);
      const RowArrays a = mutationRowArrays(t, filterOf(t, filter), idx, min_proportion);
      const std::string text = capi::mutationRowsNdjson(a.view(*t.seq[idx], fields));
      *length = text.size();
      if (text.size() > capacity || (!out && !text.empty())) fail(RDB_ENOMEM, "output capacity too small");
      if (!text.empty()) std::memcpy(out, text.data(), text.size());
   });
}

int rdb_mutations_arrow_ipc(rdb_table* table, const rdb_bitmap* filter, const char* column, double min_proportion,
                            uint32_t fields, uint8_t* out, uint64_t capacity, uint64_t* length) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!column || !length) fail(RDB_EINVAL, "null argument");
      capi::validFields(fields);
      const int idx = capi::mutationColumn(t, column, min_proportion);
      const RowArrays a = mutationRowArrays(t, filterOf(t, filter), idx, min_proportion);
      const std::vector<uint8_t> bytes = capi::mutationRowsArrowIpc(a.view(*t.seq[idx], fields));
      *length = bytes.size();
      if (bytes.size() > capacity || !out) fail(RDB_ENOMEM, "output capacity too small");
      std::memcpy(out, bytes.data(), bytes.size());
   });
}

}  // extern "C"

namespace rdb::capi {

void aggregateRows(Table& t, const DeviceBitmap& f, const char* dims, std::vector<uint32_t>& ids,
                   std::vector<uint64_t>& counts) {
   const std::vector<capi::AggDim> ds = capi::parseDims(t, dims);
   GroupDims gd{};
   for (size_t k = 0; k < ds.size(); ++k) {
      if (ds[k].is_seq) {
         const SeqColumn& col = *ds[k].seq;
         const uint32_t p = ds[k].position;
         const size_t padded = static_cast<size_t>(t.n_chunks) << 16;
         if (t.labels_scratch[k].bytes() < padded * 4) t.labels_scratch[k] = DeviceBuffer(padded * 4);
         const uint32_t blk = col.block_of_pos[p];
         const uint32_t base = col.block_slot_start[blk];
         hipCheck(launchSymbolLabels(t.seq_views_dev.as<SeqColumnView>() + ds[k].seq_index, t.chunk_sizes_dev.as<uint32_t>(),
                                     p, t.n_chunks, blk, col.pos_slot_ptr[p] - base, col.pos_slot_ptr[p + 1] - base,
                                     t.labels_scratch[k].as<uint32_t>(), t.ctx->stream),
                  "symbol labels");
         gd.labels[k] = t.labels_scratch[k].as<uint32_t>();
      } else {
         gd.labels[k] = ds[k].dict->labels.as<uint32_t>();
      }
   }
   gd.n = static_cast<uint32_t>(ds.size());
   if (ds.empty()) {
      // partition() at depth 0 == #dimensions emits the single empty combination with |F|
      // (bitmap_aggregation_node.cpp:91-101), also when |F| is 0
      ids.clear();
      counts.assign(1, const_cast<DeviceBitmap&>(f).cardinality());
      return;
   }
   const uint64_t bins = capi::dimStrides(ds, gd.stride);
   const size_t hb = bins * 4;
   void* hist = t.ctx->pool.get(hb);
   std::vector<uint32_t> h(bins);
   try {
      hipCheck(launchGroupHistogram(t.n_chunks, f.words, f.chunk_cards, gd, static_cast<uint32_t>(bins),
                                    static_cast<uint32_t*>(hist), t.ctx->stream),
               "group histogram");
      hipCheck(hipMemcpyAsync(h.data(), hist, hb, hipMemcpyDeviceToHost, t.ctx->stream), "copy histogram");
      hipCheck(hipStreamSynchronize(t.ctx->stream), "synchronize");
   } catch (...) {
      t.ctx->pool.put(hist, hb);
      throw;
   }
   t.ctx->pool.put(hist, hb);
   decodeHistogram(ds, gd.stride, h.data(), bins, ids, counts);
}

}  // namespace rdb::capi

extern "C" {

int rdb_aggregate(rdb_table* table, const rdb_bitmap* filter, const char* dims, uint32_t* ids, uint64_t* counts,
                  uint64_t capacity, uint64_t* n) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!dims || !n) fail(RDB_EINVAL, "null argument");
      const DeviceBitmap& f = filterOf(t, filter);
      std::vector<uint32_t> out_ids;
      std::vector<uint64_t> out_counts;
      capi::aggregateRows(t, f, dims, out_ids, out_counts);
      *n = out_counts.size();
      if (out_counts.size() > capacity) fail(RDB_ENOMEM, "output capacity too small");
      std::copy(out_ids.begin(), out_ids.end(), ids);
      std::copy(out_counts.begin(), out_counts.end(), counts);
   });
}

int rdb_profile(rdb_ctx* ctx, int enable) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx) fail(RDB_EINVAL, "null context");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      ctx->c.profiling = enable != 0;
      for (int k = 0; k < 3; ++k) {
         ctx->c.prof_ms[k] = 0;
         ctx->c.prof_launches[k] = 0;
      }
      ctx->c.prof_dense_queries = 0;
   });
}

int rdb_profile_read_scan(rdb_ctx* ctx, double* scan_ms, uint64_t* scan_launches, uint64_t* dense_queries) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx) fail(RDB_EINVAL, "null context");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      if (scan_ms) *scan_ms = ctx->c.prof_ms[2];
      if (scan_launches) *scan_launches = ctx->c.prof_launches[2];
      if (dense_queries) *dense_queries = ctx->c.prof_dense_queries;
   });
}

int rdb_profile_read(rdb_ctx* ctx, double* mutations_ms, uint64_t* mutations_launches, double* filter_ms,
                     uint64_t* filter_launches) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx) fail(RDB_EINVAL, "null context");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      if (mutations_ms) *mutations_ms = ctx->c.prof_ms[0];
      if (mutations_launches) *mutations_launches = ctx->c.prof_launches[0];
      if (filter_ms) *filter_ms = ctx->c.prof_ms[1];
      if (filter_launches) *filter_launches = ctx->c.prof_launches[1];
   });
}

int rdb_mutations_traffic(rdb_table* table, const rdb_bitmap* filter, const char* column, uint64_t* bytes) {
   Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!column || !bytes) fail(RDB_EINVAL, "null argument");
      const DeviceBitmap& f = filterOf(t, filter);
      const SeqColumn& col = t.seqColumn(column);
      std::vector<uint32_t> cards(t.n_chunks);
      hipCheck(hipMemcpyAsync(cards.data(), f.chunk_cards, cards.size() * 4, hipMemcpyDeviceToHost, t.ctx->stream), "copy");
      hipCheck(hipStreamSynchronize(t.ctx->stream), "synchronize");
      uint64_t card = 0;
      for (uint32_t c : cards) card += c;
      const uint64_t counts_bytes = 4ULL * col.abc->count * col.genome_length;
      uint64_t b = 0;
      if (t.ctx->sparse_max_rows >= 0 && card <= static_cast<uint64_t>(t.ctx->sparse_max_rows)) {
         // sparse method: the filter words, then per filter row its coverage word, two CSR offset
         // pairs, its mutations and its missing runs; the counts are read and cleared
         DeviceBitmap& fm = const_cast<DeviceBitmap&>(f);
         b = 8ULL * t.n_chunks * kChunkWords + 2 * counts_bytes + 2ULL * 12 * col.genome_length;
         for (uint32_t r : fm.rows()) {
            b += 4 + 16 + 4ULL * (col.row_mut_ptr[r + 1] - col.row_mut_ptr[r]) +
                 4ULL * (col.row_nrun_ptr[r + 1] - col.row_nrun_ptr[r]);
         }
      } else {
         b = counts_bytes;
         for (uint32_t c = 0; c < t.n_chunks; ++c) {
            if (!cards[c]) continue;
            b += col.chunk_index_bytes[c] + 8192ULL + 8ULL * t.chunk_sizes[c] + 4ULL * col.chunk_nruns[c];
         }
      }
      *bytes = b;
   });
}

int rdb_diff_sequence(int alphabet, const char* sequence, uint32_t length, const char* reference, uint32_t* start,
                      uint32_t* end, uint32_t* missing, uint64_t missing_capacity, uint64_t* n_missing,
                      uint32_t* mut_pos, uint8_t* mut_sym, uint64_t mut_capacity, uint64_t* n_mut) {
   return guarded(nullptr, [&] {
      if (!sequence || !reference || !start || !end || !n_missing || !n_mut) fail(RDB_EINVAL, "null argument");
      const Alphabet& a = alphabetFor(alphabet);
      // aligned_sequence.cpp:18-122: N is missing, a symbol equal to the reference's is no mutation,
      // leading/trailing missing positions define the covered range
      std::vector<uint32_t> miss;
      std::vector<std::pair<uint32_t, uint8_t>> muts;
      for (uint32_t i = 0; i < length; ++i) {
         const int s = a.fromChar(sequence[i]);
         if (s < 0) {
            fail(RDB_EINVAL, std::string("illegal character '") + sequence[i] + "' at position " + std::to_string(i) +
                                " in the input sequence");
         }
         if (s == a.missing) {
            miss.push_back(i);
         } else if (s != a.fromChar(reference[i])) {
            muts.emplace_back(i, static_cast<uint8_t>(s));
         }
      }
      size_t lead = 0;
      while (lead < miss.size() && miss[lead] == lead) ++lead;
      size_t trail = 0;
      while (trail < miss.size() - lead && miss[miss.size() - 1 - trail] == length - 1 - trail) ++trail;
      if (lead + trail == length) {
         *start = 0;
         *end = 0;
         miss.clear();
      } else {
         *start = static_cast<uint32_t>(lead);
         *end = static_cast<uint32_t>(length - trail);
         miss.assign(miss.begin() + static_cast<ptrdiff_t>(lead), miss.end() - static_cast<ptrdiff_t>(trail));
      }
      *n_missing = miss.size();
      *n_mut = muts.size();
      if (miss.size() > missing_capacity || muts.size() > mut_capacity) fail(RDB_ENOMEM, "output capacity too small");
      std::copy(miss.begin(), miss.end(), missing);
      for (size_t i = 0; i < muts.size(); ++i) {
         mut_pos[i] = muts[i].first;
         mut_sym[i] = muts[i].second;
      }
   });
}

}  // extern "C"
