// Batched execution of filter + Count / Mutations / Aggregated queries (see batch.h).
#include "batch.h"

#include <algorithm>
#include <cstring>

#include "capi_internal.h"
#include "planner.h"
#include "query.h"

namespace rdb {

namespace {

constexpr size_t kAlign = 256;
size_t alignUp(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Layout of a lane's staged inputs (device and pinned host alike).
struct ParamsLayout {
   size_t header = 0;        // u32: n_queries, n_agg, total_bins, pad
   size_t is_mutations = 0;  // u32 [max_queries]
   size_t min_prop = 0;      // f64 [max_queries]
   size_t agg_query = 0;     // u32 [max_queries]
   size_t agg_bins = 0;      // u32 [max_queries]
   size_t agg_offset = 0;    // u32 [max_queries]
   size_t dims = 0;          // GroupDims [max_queries]
   size_t progs = 0;         // ProgramDesc [max_queries]
   size_t instrs = 0;        // Instr [instr_capacity]
   size_t ranges = 0;        // u32 [range_capacity]
   size_t bytes = 0;

   ParamsLayout(uint32_t q, uint32_t instr_cap, uint32_t range_cap) {
      size_t o = 0;
      auto take = [&](size_t n) {
         const size_t at = o;
         o = alignUp(o + n);
         return at;
      };
      header = take(16);
      is_mutations = take(4ULL * q);
      min_prop = take(8ULL * q);
      agg_query = take(4ULL * q);
      agg_bins = take(4ULL * q);
      agg_offset = take(4ULL * q);
      dims = take(sizeof(GroupDims) * q);
      progs = take(sizeof(ProgramDesc) * q);
      instrs = take(sizeof(Instr) * instr_cap);
      ranges = take(4ULL * range_cap);
      bytes = o;
   }
};

// Layout of a lane's results: per-query header {n_rows, offset, card lo, card hi}, then the packed
// 16-byte records: Mutations rows, or {bin, count} of the non-empty groups of Aggregated queries.
struct OutLayout {
   size_t header = 0;
   size_t rows = 0;
   size_t bytes = 0;
   OutLayout(uint32_t q, uint32_t /*hist_cap*/, uint32_t row_cap) {
      header = 0;
      rows = alignUp(16ULL * q);
      bytes = rows + sizeof(MutationRowDev) * static_cast<size_t>(q) * row_cap;
   }
};

struct CompiledQuery {
   query::Emitter program;
   uint32_t depth = 1;
};

// Stack depth a lane's filter launch is sized for: rounded up to a multiple of 8 so that few graph
// shapes cover all programs.
// Histogram words of a lane: room for every query of a job to group by up to 2^17 combinations.
uint32_t histCapacity(uint32_t max_queries) {
   return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(uint64_t{max_queries} << 17, 1u << 20), 1u << 30));
}

uint32_t depthBucket(uint32_t depth, uint32_t max_depth) { return std::min((depth + 7) / 8 * 8, max_depth); }

BatchLane& acquireLane(Table& t);
void retire(BatchLane& lane);

// Mutations scratch of one lane slot, sized for the largest sequence column so that a lane serves
// any column.
struct MutScratch {
   size_t slots = 0, counts = 0, bnd = 0, pblocks = 0;
};
MutScratch mutScratch(const Table& t) {
   MutScratch m;
   for (const auto& c : t.seq) {
      const size_t L = c->genome_length;
      m.slots = std::max<size_t>(m.slots, c->view.n_slots);
      m.counts = std::max<size_t>(m.counts, L * c->abc->count);
      m.bnd = std::max<size_t>(m.bnd, 3 * L);
      m.pblocks = std::max<size_t>(m.pblocks, (L + kMutPosBlock - 1) / kMutPosBlock);
   }
   return m;
}

void setupLane(Table& t, BatchLane& lane) {
   Context& ctx = *t.ctx;
   lane.table = &t;
   lane.max_queries = ctx.batch_max_queries;
   lane.row_capacity = ctx.batch_row_capacity;
   lane.instr_capacity = lane.max_queries * 256;
   lane.range_capacity = lane.max_queries * 256;
   lane.hist_capacity = histCapacity(lane.max_queries);
   hipCheck(hipStreamCreateWithFlags(&lane.stream, hipStreamNonBlocking), "hipStreamCreate");
   hipCheck(hipEventCreateWithFlags(&lane.done, hipEventDisableTiming), "hipEventCreate");
   const uint32_t Q = lane.max_queries;
   const ParamsLayout pl(Q, lane.instr_capacity, lane.range_capacity);
   const OutLayout ol(Q, lane.hist_capacity, lane.row_capacity);
   lane.params_bytes = pl.bytes;
   lane.out_bytes = ol.bytes;
   lane.params = DeviceBuffer(pl.bytes);
   lane.out = DeviceBuffer(ol.bytes);
   hipCheck(hipHostMalloc(&lane.host_params, pl.bytes, hipHostMallocDefault), "pinned batch inputs");
   hipCheck(hipHostMalloc(&lane.host_out, ol.bytes, hipHostMallocDefault), "pinned batch results");
   const size_t words = static_cast<size_t>(t.n_chunks) * kChunkWords;
   lane.filters = DeviceBuffer(words * 8 * Q);
   lane.cards = DeviceBuffer(8ULL * Q + 4ULL * Q * t.n_chunks);
   lane.lists = DeviceBuffer(8ULL * Q);
   lane.list_counts = DeviceBuffer(64);  // dense / sparse list sizes, work counters
   lane.n_rows = DeviceBuffer(4ULL * Q);
   lane.hist = DeviceBuffer(4ULL * lane.hist_capacity);
   lane.rows = DeviceBuffer(sizeof(MutationRowDev) * static_cast<size_t>(Q) * lane.row_capacity);
   if (!t.seq.empty()) {
      const MutScratch ms = mutScratch(t);
      lane.partial = DeviceBuffer(4ULL * Q * ms.slots);
      lane.heavy_sums = DeviceBuffer(4ULL * Q * ms.slots);
      lane.partial_b = DeviceBuffer(4ULL * Q * ms.slots);
      hipCheck(hipMemset(lane.heavy_sums.as<void>(), 0, lane.heavy_sums.bytes()), "clear heavy sums");
      lane.sparse_counts = DeviceBuffer(4ULL * Q * ms.counts);
      lane.sparse_bnd = DeviceBuffer(4ULL * Q * ms.bnd);
      lane.totals = DeviceBuffer(2ULL * Q * ms.pblocks * 3 * sizeof(int64_t));
      hipCheck(hipMemset(lane.sparse_counts.as<void>(), 0, lane.sparse_counts.bytes()), "clear sparse counts");
      hipCheck(hipMemset(lane.sparse_bnd.as<void>(), 0, lane.sparse_bnd.bytes()), "clear sparse boundaries");
   }
}

// The lane's fixed kernel sequence for programs of at most `depth` stack entries whose Mutations
// queries read sequence column `column` (-1: none); the histogram kernels only when the job has
// Aggregated queries.
void enqueueKernels(Table& t, BatchLane& lane, uint32_t depth, int column, bool aggregates, uint32_t grid_q) {
   Context& ctx = *t.ctx;
   const uint32_t Q = lane.max_queries;
   const ParamsLayout pl(Q, lane.instr_capacity, lane.range_capacity);
   const OutLayout ol(Q, lane.hist_capacity, lane.row_capacity);
   auto* pd = lane.params.as<uint8_t>();
   auto* od = lane.out.as<uint8_t>();
   const auto* n_queries = reinterpret_cast<const uint32_t*>(pd + pl.header);
   auto* cards = lane.cards.as<unsigned long long>();
   auto* chunk_cards = reinterpret_cast<uint32_t*>(lane.cards.as<uint8_t>() + 8ULL * Q);
   hipCheck(launchFill(reinterpret_cast<uint32_t*>(cards), 0, lane.cards.bytes() / 4, lane.stream), "clear cardinalities");
   hipCheck(launchFilter(t.view, reinterpret_cast<const ProgramDesc*>(pd + pl.progs), reinterpret_cast<const Instr*>(pd + pl.instrs),
                         reinterpret_cast<const uint32_t*>(pd + pl.ranges), grid_q, depth, lane.filters.as<uint64_t>(), cards,
                         chunk_cards, static_cast<uint32_t>(ctx.cu_count) * 8, lane.stream, n_queries),
            "filter kernel");
   if (column >= 0) {
      const SeqColumn& col = *t.seq[column];
      MutationsBatch m;
      m.col_dev = t.seq_views_dev.as<SeqColumnView>() + column;
      m.col_host = &col.view;
      m.chunk_sizes = t.chunk_sizes_dev.as<uint32_t>();
      m.n_chunks = t.n_chunks;
      m.max_queries = Q;
      m.grid_queries = grid_q;
      m.n_queries = n_queries;
      m.filters = lane.filters.as<uint64_t>();
      m.chunk_cards = chunk_cards;
      m.cards = cards;
      m.is_mutations = reinterpret_cast<const uint32_t*>(pd + pl.is_mutations);
      m.min_proportion = reinterpret_cast<const double*>(pd + pl.min_prop);
      m.sparse_max = ctx.sparse_max_rows;
      m.unroll = ctx.mutation_unroll;
      m.coop = ctx.mutation_coop != 0;
      m.lists = lane.lists.as<uint32_t>();
      m.list_counts = lane.list_counts.as<uint32_t>();
      m.partial = lane.partial.as<uint32_t>();
      m.heavy_sums = ctx.batch_shared_scan ? lane.heavy_sums.as<uint32_t>() : nullptr;
      m.partial_b = lane.partial_b.as<uint32_t>();
      m.shared_mode = ctx.batch_shared_scan;
      m.partial_c = lane.partial_c.bytes() ? lane.partial_c.as<uint32_t>() : nullptr;
      m.octet_masks = lane.octet_masks.bytes() ? lane.octet_masks.as<uint8_t>() : nullptr;
      m.persistent_groups = ctx.batch_octet_groups ? ctx.batch_octet_groups : static_cast<uint32_t>(ctx.cu_count);
      m.octet_threads = ctx.batch_octet_threads;
      m.sparse_counts = lane.sparse_counts.as<uint32_t>();
      m.sparse_bnd = lane.sparse_bnd.as<int32_t>();
      m.totals = lane.totals.as<int64_t>();
      m.rows = lane.rows.as<MutationRowDev>();
      m.row_capacity = lane.row_capacity;
      m.n_rows = lane.n_rows.as<uint32_t>();
      if (ctx.profiling) {
         for (auto& e : lane.scan_events) {
            if (e == nullptr) hipCheck(hipEventCreate(&e), "hipEventCreate");
         }
         m.scan_start = lane.scan_events[0];
         m.scan_stop = lane.scan_events[1];
         lane.job_timed = true;
      }
      hipCheck(launchMutationsBatch(m, lane.stream), "batched Mutations kernels");
   } else {
      hipCheck(launchFill(lane.n_rows.as<uint32_t>(), 0, Q, lane.stream), "clear row counts");
   }
   if (aggregates) {
      AggregateBatch a;
      a.n_chunks = t.n_chunks;
      a.max_entries = grid_q;
      a.n_agg = n_queries + 1;
      a.total_bins = n_queries + 2;
      a.max_total_bins = lane.hist_capacity;
      a.agg_query = reinterpret_cast<const uint32_t*>(pd + pl.agg_query);
      a.bins = reinterpret_cast<const uint32_t*>(pd + pl.agg_bins);
      a.hist_offset = reinterpret_cast<const uint32_t*>(pd + pl.agg_offset);
      a.dims = reinterpret_cast<const GroupDims*>(pd + pl.dims);
      a.filters = lane.filters.as<uint64_t>();
      a.chunk_cards = chunk_cards;
      a.hist = lane.hist.as<uint32_t>();
      a.rows = lane.rows.as<MutationRowDev>();
      a.row_capacity = lane.row_capacity;
      a.n_rows = lane.n_rows.as<uint32_t>();
      hipCheck(launchAggregateBatch(a, lane.stream), "batched group histogram");
   }
   hipCheck(launchPackRows(n_queries, grid_q, lane.n_rows.as<uint32_t>(), cards, lane.rows.as<MutationRowDev>(), lane.row_capacity,
                           reinterpret_cast<uint32_t*>(od + ol.header), reinterpret_cast<MutationRowDev*>(od + ol.rows),
                           lane.stream),
            "pack rows");
}

// Buffers of the octet shared scan (outside any graph capture): the slot sums of chunk parts 2 and 3
// and one byte per row per eight queries.
void ensureOctetBuffers(const Table& t, BatchLane& lane) {
   if (t.seq.empty() || lane.octet_masks.bytes() != 0 ||
       effectiveSharedScan(t.ctx->batch_shared_scan, lane.max_queries, t.n_chunks) != 2) {
      return;
   }
   const uint32_t Q = lane.max_queries;
   const MutScratch ms = mutScratch(t);
   lane.partial_c = DeviceBuffer(4ULL * (kOctetParts - 2) * Q * ms.slots);
   lane.octet_masks = DeviceBuffer(static_cast<size_t>((Q + 7) / 8) * t.n_chunks * kChunkRows);
}

void launchLane(Table& t, BatchLane& lane, uint32_t depth, int column, bool aggregates, uint32_t n_queries) {
   Context& ctx = *t.ctx;
   ensureOctetBuffers(t, lane);
   const uint32_t bucket = depthBucket(depth, maxFilterDepth(ctx));
   // the per-query grids cover the job's size rounded up to a power of two, not the lane's
   // capacity: a small job (a serving loop's few queries) then does not dispatch the work-groups of
   // the capacity's empty query slots
   uint32_t grid_q = 1;
   while (grid_q < n_queries) grid_q <<= 1;
   grid_q = std::min(grid_q, lane.max_queries);
   lane.job_timed = false;
   if (!ctx.batch_graphs || ctx.profiling) {  // timed jobs run eagerly: events around their kernels
      enqueueKernels(t, lane, bucket, column, aggregates, grid_q);
      return;
   }
   const uint32_t knobs = ctx.mutation_unroll | (ctx.batch_shared_scan << 8) | (ctx.mutation_coop ? 1u << 12 : 0u) |
                          (ctx.batch_octet_threads == 512 ? 1u << 13 : 0u) | (aggregates ? 1u << 14 : 0u) |
                          (ctx.batch_octet_groups << 16);
   const auto key = std::make_tuple(bucket | (grid_q << 16), column, static_cast<long long>(ctx.sparse_max_rows), knobs);
   auto it = lane.graphs.find(key);
   if (it == lane.graphs.end()) {
      hipGraph_t graph = nullptr;
      hipCheck(hipStreamBeginCapture(lane.stream, hipStreamCaptureModeThreadLocal), "begin capture");
      try {
         enqueueKernels(t, lane, bucket, column, aggregates, grid_q);
      } catch (...) {
         (void)hipStreamEndCapture(lane.stream, &graph);
         if (graph) (void)hipGraphDestroy(graph);
         throw;
      }
      hipCheck(hipStreamEndCapture(lane.stream, &graph), "end capture");
      hipGraphExec_t exec = nullptr;
      const hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      hipCheck(err, "graph instantiate");
      it = lane.graphs.emplace(key, exec).first;
   }
   hipCheck(hipGraphLaunch(it->second, lane.stream), "graph launch");
}

struct JobPlan {
   std::vector<uint32_t> queries;  // batch indices
   uint32_t instrs = 0, ranges = 0, bins = 0;
   int column = -1;
};

void enqueueJob(Table& t, BatchResult& b, const std::vector<CompiledQuery>& cq, const JobPlan& job) {
   BatchLane& lane = acquireLane(t);
   const uint32_t Q = lane.max_queries;
   const ParamsLayout pl(Q, lane.instr_capacity, lane.range_capacity);
   const OutLayout ol(Q, lane.hist_capacity, lane.row_capacity);
   auto* hp = static_cast<uint8_t*>(lane.host_params);
   auto* is_mut = reinterpret_cast<uint32_t*>(hp + pl.is_mutations);
   auto* min_prop = reinterpret_cast<double*>(hp + pl.min_prop);
   auto* agg_query = reinterpret_cast<uint32_t*>(hp + pl.agg_query);
   auto* agg_bins = reinterpret_cast<uint32_t*>(hp + pl.agg_bins);
   auto* agg_offset = reinterpret_cast<uint32_t*>(hp + pl.agg_offset);
   auto* dims = reinterpret_cast<GroupDims*>(hp + pl.dims);
   auto* progs = reinterpret_cast<ProgramDesc*>(hp + pl.progs);
   auto* instrs = reinterpret_cast<Instr*>(hp + pl.instrs);
   auto* ranges = reinterpret_cast<uint32_t*>(hp + pl.ranges);
   uint32_t n_instr = 0, n_range = 0, n_agg = 0, total_bins = 0, depth = 1;
   const uint32_t n = static_cast<uint32_t>(job.queries.size());
   for (uint32_t slot = 0; slot < n; ++slot) {
      const uint32_t qi = job.queries[slot];
      const BatchQueryResult& r = b.queries[qi];
      const query::Emitter& e = cq[qi].program;
      progs[slot] = ProgramDesc{n_instr, static_cast<uint32_t>(e.code.size()), n_range, slot,
                                n_range + static_cast<uint32_t>(e.ranges.size()), static_cast<uint32_t>(e.prefetch.size())};
      // (empty vectors may hold null data(): memcpy is only called for non-empty ones)
      if (!e.code.empty()) std::memcpy(instrs + n_instr, e.code.data(), e.code.size() * sizeof(Instr));
      if (!e.ranges.empty()) std::memcpy(ranges + n_range, e.ranges.data(), e.ranges.size() * 4);
      if (!e.prefetch.empty()) std::memcpy(ranges + n_range + e.ranges.size(), e.prefetch.data(), e.prefetch.size() * 4);
      n_instr += static_cast<uint32_t>(e.code.size());
      n_range += static_cast<uint32_t>(e.ranges.size() + e.prefetch.size());
      depth = std::max(depth, cq[qi].depth);
      is_mut[slot] = r.status == RDB_OK && r.action == RDB_ACTION_MUTATIONS ? 1u : 0u;
      min_prop[slot] = r.min_proportion;
      if (r.status == RDB_OK && r.action == RDB_ACTION_AGGREGATE) {
         const auto ds = capi::parseDims(t, r.dims.c_str());
         GroupDims gd{};
         gd.n = static_cast<uint32_t>(ds.size());
         const uint64_t bins = capi::dimStrides(ds, gd.stride);
         for (size_t k = 0; k < ds.size(); ++k) gd.labels[k] = ds[k].dict->labels.as<uint32_t>();
         agg_query[n_agg] = slot;
         agg_bins[n_agg] = static_cast<uint32_t>(bins);
         agg_offset[n_agg] = total_bins;
         dims[n_agg] = gd;
         total_bins += static_cast<uint32_t>(bins);
         ++n_agg;
      }
   }
   auto* header = reinterpret_cast<uint32_t*>(hp + pl.header);
   header[0] = n;
   header[1] = n_agg;
   header[2] = total_bins;
   header[3] = 0;
   // the header and up to 1 MiB of packed records (the rest, if any, on retirement)
   lane.job_first_bytes = std::min(lane.out_bytes, ol.rows + (size_t{1} << 20));
   try {
      hipCheck(hipMemcpyAsync(lane.params.as<void>(), hp, pl.bytes, hipMemcpyHostToDevice, lane.stream), "stage batch");
      if (t.ctx->batch_fail_job > 0 && --t.ctx->batch_fail_job == 0) fail(RDB_ENOMEM, "injected lane-job failure");
      launchLane(t, lane, depth, job.column, n_agg > 0, n);
      hipCheck(hipMemcpyAsync(lane.host_out, lane.out.as<void>(), lane.job_first_bytes, hipMemcpyDeviceToHost, lane.stream),
               "copy batch results");
      hipCheck(hipEventRecord(lane.done, lane.stream), "hipEventRecord");
   } catch (...) {
      // whatever was enqueued finishes before the lane takes another job; the lane stays free
      (void)hipStreamSynchronize(lane.stream);
      throw;
   }
   // the lane holds the job only once all of its work is enqueued
   lane.busy = true;
   lane.owner = &b;
   lane.job_seq = ++t.job_counter;
   lane.job_queries = job.queries;
   lane.job_total_bins = total_bins;
   b.pending.push_back(&lane);
}

BatchLane& acquireLane(Table& t) {
   Context& ctx = *t.ctx;
   auto shaped = [&](const BatchLane& l) {
      return l.max_queries == ctx.batch_max_queries && l.row_capacity == ctx.batch_row_capacity;
   };
   for (auto& l : t.lanes) {
      if (!l->busy && shaped(*l)) return *l;
   }
   // drop idle lanes of another shape, then create one if the budget allows
   t.lanes.erase(std::remove_if(t.lanes.begin(), t.lanes.end(),
                                [&](const std::shared_ptr<BatchLane>& l) { return !l->busy && !shaped(*l); }),
                 t.lanes.end());
   if (t.lanes.size() < std::max<uint32_t>(ctx.batch_lanes, 1)) {
      auto lane = std::make_shared<BatchLane>();
      setupLane(t, *lane);
      t.lanes.push_back(lane);
      return *lane;
   }
   // every lane is busy: retire the one whose job was enqueued first
   BatchLane* oldest = t.lanes.front().get();
   for (auto& l : t.lanes) {
      if (l->busy && l->job_seq < oldest->job_seq) oldest = l.get();
   }
   hipCheck(hipEventSynchronize(oldest->done), "batch wait");
   retire(*oldest);
   return acquireLane(t);
}

// Copies a finished job's results into its batch and frees the lane.  A job detached from its
// batch (owner null) is dropped.
void retire(BatchLane& lane) {
   if (!lane.busy) return;
   if (lane.owner == nullptr) {
      lane.busy = false;
      lane.job_queries.clear();
      return;
   }
   Table& t = *lane.table;
   BatchResult& b = *lane.owner;
   const uint32_t Q = lane.max_queries;
   const OutLayout ol(Q, lane.hist_capacity, lane.row_capacity);
   auto* ho = static_cast<uint8_t*>(lane.host_out);
   const auto* header = reinterpret_cast<const uint32_t*>(ho + ol.header);
   const uint32_t n = static_cast<uint32_t>(lane.job_queries.size());
   if (lane.job_timed) {
      // the dense counting kernel's time and the number of dense queries it counted (Mutations
      // queries with a non-empty filter above the dense/sparse switch)
      float ms = 0;
      hipCheck(hipEventElapsedTime(&ms, lane.scan_events[0], lane.scan_events[1]), "profile time");
      Context& ctx = *t.ctx;
      ctx.prof_ms[2] += ms;
      ctx.prof_launches[2] += 1;
      for (uint32_t s = 0; s < n; ++s) {
         const BatchQueryResult& r = b.queries[lane.job_queries[s]];
         const uint64_t card = static_cast<uint64_t>(header[4 * s + 2]) | (static_cast<uint64_t>(header[4 * s + 3]) << 32);
         if (r.status == RDB_OK && r.action == RDB_ACTION_MUTATIONS && card > 0 &&
             (ctx.sparse_max_rows < 0 || card > static_cast<uint64_t>(ctx.sparse_max_rows))) {
            ++ctx.prof_dense_queries;
         }
      }
      lane.job_timed = false;
   }
   // rows beyond the first copy
   size_t kept = 0;
   for (uint32_t s = 0; s < n; ++s) kept += std::min(header[4 * s], lane.row_capacity);
   const size_t rows_bytes = kept * sizeof(MutationRowDev);
   if (ol.rows + rows_bytes > lane.job_first_bytes) {
      hipCheck(hipMemcpyAsync(ho + lane.job_first_bytes, lane.out.as<uint8_t>() + lane.job_first_bytes,
                              ol.rows + rows_bytes - lane.job_first_bytes, hipMemcpyDeviceToHost, lane.stream),
               "copy batch rows");
      hipCheck(hipStreamSynchronize(lane.stream), "synchronize");
   }
   const auto* packed = reinterpret_cast<const MutationRowDev*>(ho + ol.rows);
   std::vector<uint32_t> overflow;
   for (uint32_t s = 0; s < n; ++s) {
      BatchQueryResult& r = b.queries[lane.job_queries[s]];
      if (r.status != RDB_OK) continue;
      r.cardinality = static_cast<uint64_t>(header[4 * s + 2]) | (static_cast<uint64_t>(header[4 * s + 3]) << 32);
      if (r.action == RDB_ACTION_MUTATIONS) {
         const uint32_t nr = header[4 * s];
         if (nr > lane.row_capacity) {
            overflow.push_back(lane.job_queries[s]);
            continue;
         }
         r.rows.assign(packed + header[4 * s + 1], packed + header[4 * s + 1] + nr);
         capi::sortMutationRows(r.rows);
      } else if (r.action == RDB_ACTION_AGGREGATE) {
         const auto ds = capi::parseDims(t, r.dims.c_str());
         if (ds.empty()) {
            // a single empty combination with |F| (bitmap_aggregation_node.cpp:91-101)
            r.agg_counts.assign(1, r.cardinality);
            continue;
         }
         const uint32_t nb = header[4 * s];
         if (nb > lane.row_capacity) {
            overflow.push_back(lane.job_queries[s]);
            continue;
         }
         uint32_t stride[4] = {};
         capi::dimStrides(ds, stride);
         std::vector<std::pair<uint32_t, uint32_t>> bins;
         bins.reserve(nb);
         for (uint32_t k = 0; k < nb; ++k) {
            const MutationRowDev& rec = packed[header[4 * s + 1] + k];
            bins.emplace_back(rec.position, rec.symbol);
         }
         capi::decodeBins(ds, stride, bins, r.agg_ids, r.agg_counts);
      }
   }
   lane.busy = false;
   lane.owner = nullptr;
   lane.job_queries.clear();
   b.pending.erase(std::remove(b.pending.begin(), b.pending.end(), &lane), b.pending.end());
   // queries whose records did not fit the lane's capacity run again on the single-query path
   for (uint32_t qi : overflow) {
      BatchQueryResult& r = b.queries[qi];
      DeviceBitmap f = query::compileFilter(t, r.filter)->evaluate();
      if (r.action == RDB_ACTION_MUTATIONS) {
         r.rows = capi::mutationRows(t, f, r.column, r.min_proportion);
      } else {
         capi::aggregateRows(t, f, r.dims.c_str(), r.agg_ids, r.agg_counts);
      }
   }
}

}  // namespace

BatchLane::~BatchLane() {
   if (stream) (void)hipStreamSynchronize(stream);
   for (auto& [k, g] : graphs) (void)hipGraphExecDestroy(g);
   for (auto& e : scan_events) {
      if (e) (void)hipEventDestroy(e);
   }
   if (done) (void)hipEventDestroy(done);
   if (host_params) (void)hipHostFree(host_params);
   if (host_out) (void)hipHostFree(host_out);
   if (stream) (void)hipStreamDestroy(stream);
}

void submitCompiled(Table& t, BatchResult& b, std::vector<CompiledQuery>& cq);

void batchSubmit(Table& t, BatchResult& b) {
   Context& ctx = *t.ctx;
   const uint32_t n = static_cast<uint32_t>(b.queries.size());
   std::vector<CompiledQuery> cq(n);
   std::vector<query::OperatorPtr> ops(n);
   // host threads for the per-query work of large batches (a small batch's work is shorter than
   // starting the threads)
   const unsigned threads = n >= kParallelCompileQueries ? 0u : 1u;
   // host side: parse, rewrite and compile every filter; a query that fails keeps its status and
   // runs as the empty program
   parallelFor(n, [&](size_t i) {
      BatchQueryResult& r = b.queries[i];
      try {
         if (r.action == RDB_ACTION_MUTATIONS) {
            r.column = capi::mutationColumn(t, r.column_name.c_str(), r.min_proportion);
         } else if (r.action == RDB_ACTION_AGGREGATE) {
            for (const auto& d : capi::parseDims(t, r.dims.c_str())) {
               if (d.is_seq) fail(RDB_EQUERY, "batched Aggregated queries group by dictionary columns only");
            }
         } else if (r.action != RDB_ACTION_COUNT) {
            fail(RDB_EINVAL, "unknown action " + std::to_string(r.action));
         }
         ops[i] = query::compileFilter(t, r.filter);
      } catch (const Error& e) {
         r.status = e.code;
         r.message = e.what();
      }
   }, threads);
   // the planner orders conjuncts and disjuncts and turns on short-circuit (results unchanged); the
   // compiled order's stack depth is kept to see whether a plan needs a deeper stack
   const bool planned = ctx.plan_mode != plan::REFERENCE;
   std::vector<uint32_t> ref_depth(planned ? n : 0, 1);
   if (planned) {
      parallelFor(n, [&](size_t i) {
         if (!ops[i]) return;
         query::Emitter e;
         ops[i]->emit(e);
         ref_depth[i] = std::max<uint32_t>(e.max_sp, 1);
      }, threads);
   }
   if (planned) {
      std::vector<query::Operator*> roots;
      for (auto& op : ops) {
         if (op) roots.push_back(op.get());
      }
      plan::planAll(t, roots, static_cast<plan::Mode>(ctx.plan_mode), ctx.planner, threads);
   }
   parallelFor(n, [&](size_t i) {
      BatchQueryResult& r = b.queries[i];
      if (r.status == RDB_OK) {
         try {
            ops[i]->emit(cq[i].program);
            const uint32_t max_depth = maxFilterDepth(ctx);
            if (planned && depthBucket(std::max<uint32_t>(cq[i].program.max_sp, 1), max_depth) >
                              depthBucket(ref_depth[i], max_depth)) {
               // the plan's order needs a deeper stack, i.e. more LDS per filter work-group and so
               // fewer of them per CU, than the compiled order: keep that order, with short-circuit
               auto op = query::compileFilter(t, r.filter);
               plan::shortCircuitOnly(*op);
               cq[i].program = query::Emitter();
               op->emit(cq[i].program);
            }
            if (cq[i].program.sp != 1) fail(RDB_ECOMPILE, "internal: filter program leaves a wrong stack");
            cq[i].depth = std::max<uint32_t>(cq[i].program.max_sp, 1);
            if (cq[i].depth > maxFilterDepth(ctx)) fail(RDB_EQUERY, "filter expression nests too deeply");
         } catch (const Error& e) {
            r.status = e.code;
            r.message = e.what();
         }
      }
      if (r.status != RDB_OK) {
         cq[i].program = query::Emitter();
         cq[i].program.leaf(Instr{OP_EMPTY, 0, 0, 0, 0, 0, 0, 0});
         cq[i].depth = 1;
      }
   }, threads);
   try {
      submitCompiled(t, b, cq);
   } catch (...) {
      batchDetach(b);
      throw;
   }
}

void batchDetach(BatchResult& b) {
   for (BatchLane* l : b.pending) {
      // the lane's job finishes before the lane takes another one; its results are dropped
      (void)hipEventSynchronize(l->done);
      l->owner = nullptr;
      retire(*l);
   }
   b.pending.clear();
}

void submitCompiled(Table& t, BatchResult& b, std::vector<CompiledQuery>& cq) {
   Context& ctx = *t.ctx;
   const uint32_t n = static_cast<uint32_t>(b.queries.size());
   // cut into lane jobs: one Mutations column per job, capacities of the lane
   const uint32_t max_q = ctx.batch_max_queries;
   const uint32_t instr_cap = max_q * 256, range_cap = max_q * 256;
   const uint32_t hist_cap = histCapacity(max_q);
   std::vector<JobPlan> jobs;
   JobPlan cur;
   auto flush = [&] {
      if (!cur.queries.empty()) jobs.push_back(std::move(cur));
      cur = JobPlan();
   };
   for (uint32_t i = 0; i < n; ++i) {
      const BatchQueryResult& r = b.queries[i];
      const uint32_t ni = static_cast<uint32_t>(cq[i].program.code.size());
      const uint32_t nr = static_cast<uint32_t>(cq[i].program.ranges.size() + cq[i].program.prefetch.size());
      uint32_t bins = 0;
      if (r.status == RDB_OK && r.action == RDB_ACTION_AGGREGATE) {
         uint32_t stride[4];
         bins = static_cast<uint32_t>(capi::dimStrides(capi::parseDims(t, r.dims.c_str()), stride));
      }
      if (ni > instr_cap || nr > range_cap || bins > hist_cap) {
         b.queries[i].status = RDB_EQUERY;
         b.queries[i].message = "query too large for the batched path";
         continue;
      }
      const int col = r.status == RDB_OK && r.action == RDB_ACTION_MUTATIONS ? r.column : -1;
      if (cur.queries.size() == max_q || cur.instrs + ni > instr_cap || cur.ranges + nr > range_cap ||
          cur.bins + bins > hist_cap || (col >= 0 && cur.column >= 0 && col != cur.column)) {
         flush();
      }
      if (col >= 0) cur.column = col;
      cur.queries.push_back(i);
      cur.instrs += ni;
      cur.ranges += nr;
      cur.bins += bins;
   }
   flush();
   for (const auto& job : jobs) enqueueJob(t, b, cq, job);
}

std::vector<uint64_t> countPrograms(Table& t, const std::vector<query::Emitter>& programs) {
   BatchResult b;
   b.table = &t;
   b.queries.resize(programs.size());
   std::vector<CompiledQuery> cq(programs.size());
   for (size_t i = 0; i < programs.size(); ++i) {
      b.queries[i].action = RDB_ACTION_COUNT;
      cq[i].program = programs[i];
      cq[i].depth = std::max<uint32_t>(programs[i].max_sp, 1);
      if (cq[i].depth > maxFilterDepth(*t.ctx)) fail(RDB_EQUERY, "filter expression nests too deeply");
   }
   try {
      submitCompiled(t, b, cq);
      batchWait(b);
   } catch (...) {
      batchDetach(b);
      throw;
   }
   std::vector<uint64_t> out(programs.size());
   for (size_t i = 0; i < programs.size(); ++i) {
      if (b.queries[i].status != RDB_OK) fail(b.queries[i].status, b.queries[i].message);
      out[i] = b.queries[i].cardinality;
   }
   return out;
}

bool batchPoll(BatchResult& b) {
   std::vector<BatchLane*> lanes = b.pending;
   for (BatchLane* l : lanes) {
      const hipError_t err = hipEventQuery(l->done);
      if (err == hipErrorNotReady) return false;
      hipCheck(err, "batch poll");
      if (l->busy && l->owner == &b) retire(*l);
   }
   return b.pending.empty();
}

size_t laneBytes(const Table& t, uint32_t max_queries, uint32_t row_capacity) {
   const uint32_t Q = max_queries;
   const uint32_t instr_cap = Q * 256, range_cap = Q * 256;
   const uint32_t hist_cap = histCapacity(Q);
   const ParamsLayout pl(Q, instr_cap, range_cap);
   const OutLayout ol(Q, hist_cap, row_capacity);
   size_t b = pl.bytes + ol.bytes;                                            // staged inputs, results
   b += static_cast<size_t>(t.n_chunks) * kChunkWords * 8 * Q;                // filter words
   b += 8ULL * Q + 4ULL * Q * t.n_chunks + 8ULL * Q + 16 + 4ULL * Q;         // cardinalities, lists
   b += sizeof(MutationRowDev) * static_cast<size_t>(Q) * row_capacity;      // rows before packing
   b += 4ULL * hist_cap;
   if (!t.seq.empty()) {
      const MutScratch ms = mutScratch(t);
      b += 4ULL * Q * (3 * ms.slots + ms.counts + ms.bnd) + 2ULL * Q * ms.pblocks * 3 * sizeof(int64_t);
      if (effectiveSharedScan(t.ctx->batch_shared_scan, Q, t.n_chunks) == 2) {
         b += 4ULL * (kOctetParts - 2) * Q * ms.slots + static_cast<size_t>((Q + 7) / 8) * t.n_chunks * kChunkRows;
      }
   }
   return b;
}

void batchWait(BatchResult& b) {
   Context& ctx = *b.table->ctx;
   while (!b.pending.empty()) {
      BatchLane* lane = b.pending.front();
      hipEvent_t ev = lane->done;
      ctx.mu.unlock();
      const hipError_t err = hipEventSynchronize(ev);
      ctx.mu.lock();
      hipCheck(err, "batch wait");
      if (lane->busy && lane->owner == &b) {
         retire(*lane);
      } else {
         b.pending.erase(std::remove(b.pending.begin(), b.pending.end(), lane), b.pending.end());
      }
   }
}

}  // namespace rdb

struct rdb_batch {
   rdb::BatchResult b;
};

using rdb::capi::guarded;
using rdb::capi::requireFinalized;
using rdb::capi::tableOf;

namespace {

// The results of query q after every job of the batch has finished; a query that failed returns
// its own status and message.
rdb::BatchQueryResult& batchQuery(rdb_batch* batch, uint32_t q) {
   if (!batch) rdb::fail(RDB_EINVAL, "null batch");
   rdb::batchWait(batch->b);
   if (q >= batch->b.queries.size()) rdb::fail(RDB_EINVAL, "query index out of range");
   rdb::BatchQueryResult& r = batch->b.queries[q];
   if (r.status != RDB_OK) rdb::fail(r.status, r.message);
   return r;
}

}  // namespace

extern "C" {

int rdb_batch_submit(rdb_table* table, const rdb_query* queries, uint32_t n, rdb_batch** out) {
   rdb::Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      rdb::Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!out || (n && !queries)) rdb::fail(RDB_EINVAL, "null argument");
      auto batch = std::make_unique<rdb_batch>();
      batch->b.table = &t;
      batch->b.queries.resize(n);
      for (uint32_t i = 0; i < n; ++i) {
         rdb::BatchQueryResult& r = batch->b.queries[i];
         const rdb_query& q = queries[i];
         r.action = q.action;
         r.filter = q.filter ? q.filter : "";
         r.column_name = q.column ? q.column : "";
         r.min_proportion = q.min_proportion;
         r.dims = q.dims ? q.dims : "";
         if (!q.filter) {
            r.status = RDB_EINVAL;
            r.message = "null filter";
         }
      }
      rdb::batchSubmit(t, batch->b);
      *out = batch.release();
   });
}

int rdb_batch_wait(rdb_batch* batch) {
   rdb::Context* ctx = batch && batch->b.table ? batch->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!batch) rdb::fail(RDB_EINVAL, "null batch");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      rdb::batchWait(batch->b);
   });
}

int rdb_batch_poll(rdb_batch* batch, int* done) {
   rdb::Context* ctx = batch && batch->b.table ? batch->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!batch || !done) rdb::fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      *done = rdb::batchPoll(batch->b) ? 1 : 0;
   });
}

int rdb_batch_lane_bytes(rdb_table* table, uint32_t max_queries, uint32_t row_capacity, uint64_t* bytes) {
   rdb::Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      rdb::Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!bytes || max_queries == 0) rdb::fail(RDB_EINVAL, "bad argument");
      *bytes = rdb::laneBytes(t, max_queries, row_capacity);
   });
}

int rdb_batch_count(rdb_batch* batch, uint32_t q, uint64_t* count) {
   rdb::Context* ctx = batch && batch->b.table ? batch->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!batch || !count) rdb::fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      *count = batchQuery(batch, q).cardinality;
   });
}

int rdb_batch_mutations(rdb_batch* batch, uint32_t q, uint32_t* position, uint8_t* from, uint8_t* to, uint32_t* count,
                        uint32_t* coverage, double* proportion, uint64_t capacity, uint64_t* n) {
   rdb::Context* ctx = batch && batch->b.table ? batch->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!batch || !n) rdb::fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      rdb::BatchQueryResult& r = batchQuery(batch, q);
      if (r.action != RDB_ACTION_MUTATIONS) rdb::fail(RDB_EINVAL, "query is not a Mutations query");
      rdb::capi::emitMutationRows(*batch->b.table->seq[r.column], r.rows, position, from, to, count, coverage,
                                  proportion, capacity, n);
   });
}

int rdb_batch_summary(rdb_batch* batch, uint32_t n_queries, int* status, uint64_t* cardinality, uint64_t* size) {
   rdb::Context* ctx = batch && batch->b.table ? batch->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!batch || (n_queries && (!status || !cardinality || !size))) rdb::fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      rdb::batchWait(batch->b);
      if (n_queries > batch->b.queries.size()) rdb::fail(RDB_EINVAL, "more queries than the batch holds");
      for (uint32_t q = 0; q < n_queries; ++q) {
         const rdb::BatchQueryResult& r = batch->b.queries[q];
         status[q] = r.status;
         cardinality[q] = r.status == RDB_OK ? r.cardinality : 0;
         size[q] = r.status != RDB_OK                    ? 0
                   : r.action == RDB_ACTION_MUTATIONS    ? r.rows.size()
                   : r.action == RDB_ACTION_AGGREGATE    ? r.agg_counts.size()
                                                         : 0;
      }
   });
}

int rdb_batch_mutations_all(rdb_batch* batch, uint32_t n_queries, uint64_t* offsets, uint32_t* position, uint8_t* from,
                            uint8_t* to, uint32_t* count, uint32_t* coverage, double* proportion, uint64_t capacity,
                            uint64_t* n) {
   rdb::Context* ctx = batch && batch->b.table ? batch->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!batch || !offsets || !n) rdb::fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      rdb::batchWait(batch->b);
      if (n_queries > batch->b.queries.size()) rdb::fail(RDB_EINVAL, "more queries than the batch holds");
      auto rowsOf = [&](uint32_t q) -> const rdb::BatchQueryResult* {
         const rdb::BatchQueryResult& r = batch->b.queries[q];
         return r.status == RDB_OK && r.action == RDB_ACTION_MUTATIONS ? &r : nullptr;
      };
      uint64_t total = 0;
      offsets[0] = 0;
      for (uint32_t q = 0; q < n_queries; ++q) {
         const rdb::BatchQueryResult* r = rowsOf(q);
         total += r ? r->rows.size() : 0;
         offsets[q + 1] = total;
      }
      *n = total;
      if (total > capacity) rdb::fail(RDB_ENOMEM, "output capacity too small");
      for (uint32_t q = 0; q < n_queries; ++q) {
         const rdb::BatchQueryResult* r = rowsOf(q);
         if (!r || r->rows.empty()) continue;
         const uint64_t o = offsets[q];
         uint64_t k = 0;
         rdb::capi::emitMutationRows(*batch->b.table->seq[r->column], r->rows, position ? position + o : nullptr,
                                     from ? from + o : nullptr, to ? to + o : nullptr, count ? count + o : nullptr,
                                     coverage ? coverage + o : nullptr, proportion ? proportion + o : nullptr,
                                     r->rows.size(), &k);
      }
      *n = total;
   });
}

int rdb_batch_aggregate(rdb_batch* batch, uint32_t q, uint32_t* ids, uint64_t* counts, uint64_t capacity, uint64_t* n) {
   rdb::Context* ctx = batch && batch->b.table ? batch->b.table->ctx : nullptr;
   return guarded(ctx, [&] {
      if (!batch || !n) rdb::fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      rdb::BatchQueryResult& r = batchQuery(batch, q);
      if (r.action != RDB_ACTION_AGGREGATE) rdb::fail(RDB_EINVAL, "query is not an Aggregated query");
      *n = r.agg_counts.size();
      if (r.agg_counts.size() > capacity) rdb::fail(RDB_ENOMEM, "output capacity too small");
      if (!r.agg_ids.empty()) std::copy(r.agg_ids.begin(), r.agg_ids.end(), ids);
      std::copy(r.agg_counts.begin(), r.agg_counts.end(), counts);
   });
}

void rdb_batch_free(rdb_batch* batch) {
   if (!batch) return;
   if (batch->b.table) {
      rdb::Context* ctx = batch->b.table->ctx;
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      try {
         rdb::batchWait(batch->b);
      } catch (...) {
         // the lanes stay valid; their jobs finish and their results are dropped with the batch
         rdb::batchDetach(batch->b);
      }
   }
   delete batch;
}

}  // extern "C"
