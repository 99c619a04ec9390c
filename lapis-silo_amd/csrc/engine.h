// Host side of the engine: context, device-resident tables, and the device row-set type that the
// filter operators return.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "device_abi.h"
#include "kernels.h"
#include "rdb.h"

namespace rdb {

// ---- errors --------------------------------------------------------------------------------------
struct Error : std::runtime_error {
   int code;
   Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

inline void hipCheck(hipError_t e, const char* what) {
   if (e != hipSuccess) fail(RDB_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Runs f(i) for i in [0, n) on up to `threads` host threads.
void parallelFor(size_t n, const std::function<void(size_t)>& f, unsigned threads = 0);

// An allocator whose value-less construct leaves elements uninitialised: the build's large arrays
// are written in full by parallel passes, and zeroing them first (on one thread, taking every page
// fault there) cost more than the passes.
template <typename T>
struct NoInit : std::allocator<T> {
   template <typename U>
   struct rebind {
      using other = NoInit<U>;
   };
   NoInit() = default;
   template <typename U>
   NoInit(const NoInit<U>&) {}
   template <typename U, typename... A>
   void construct(U* p, A&&... a) {
      if constexpr (sizeof...(A) == 0) {
         ::new (static_cast<void*>(p)) U;
      } else {
         ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
      }
   }
};
template <typename T>
using RawVector = std::vector<T, NoInit<T>>;

// ---- alphabets (common/nucleotide_symbols.h:23-106, common/aa_symbols.h:20-118) ----------------
struct Alphabet {
   uint32_t count;
   uint8_t missing;
   std::string chars;
   std::vector<uint8_t> valid;                    // VALID_MUTATION_SYMBOLS
   std::vector<std::vector<uint8_t>> codes_for;   // CODES_FOR
   std::vector<std::vector<uint8_t>> ambiguity;   // AMBIGUITY_SYMBOLS (derived)
   int fromChar(char c) const;
};
const Alphabet& alphabetFor(int alphabet);

// ---- device memory ---------------------------------------------------------------------------------
// Long-lived allocation owned by a table.
class DeviceBuffer {
  public:
   DeviceBuffer() = default;
   explicit DeviceBuffer(size_t bytes);
   DeviceBuffer(const void* host, size_t bytes);
   ~DeviceBuffer();
   DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
      o.ptr_ = nullptr;
      o.bytes_ = 0;
   }
   DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
   DeviceBuffer(const DeviceBuffer&) = delete;
   DeviceBuffer& operator=(const DeviceBuffer&) = delete;
   template <typename T>
   T* as() const {
      return static_cast<T*>(ptr_);
   }
   size_t bytes() const { return bytes_; }

  private:
   void* ptr_ = nullptr;
   size_t bytes_ = 0;
};

// Size-class cache of device allocations for per-query results and scratch.
class DevicePool {
  public:
   ~DevicePool();
   void* get(size_t bytes);
   void put(void* p, size_t bytes);

  private:
   std::mutex mu_;
   std::map<size_t, std::vector<void*>> free_;
};

struct Context {
   static constexpr size_t kStageSlots = 64;
   static constexpr size_t kStageSlotBytes = 1 << 20;

   int device = 0;
   hipStream_t stream = nullptr;
   DevicePool pool;
   std::recursive_mutex mu;
   std::string error;
   int cu_count = 256;
   // device address of this GPU's allocation-free scratch (resolved by init)
   void* noalloc_scratch = nullptr;
   size_t max_lds_bytes = 64 * 1024;  // LDS a work-group may allocate
   // pinned host staging
   void* pinned = nullptr;
   size_t pinned_bytes = 0;
   void* pinnedBuffer(size_t bytes);
   // ring of (pinned host slot, device slot, event) for small asynchronous uploads
   void* stage_host = nullptr;
   DeviceBuffer stage_dev;
   hipEvent_t stage_events[kStageSlots] = {};
   bool stage_used[kStageSlots] = {};
   size_t stage_next = 0;
   DeviceBuffer stage_big;
   // Copies `bytes` of host data to device memory that stays valid for work enqueued afterwards on
   // `stream` until the ring wraps (64 further uploads); large uploads synchronize the stream.
   void* stage(const void* host, size_t bytes);
   // kernel timing with HIP events on `stream` (enabled by rdb_profile)
   bool profiling = false;
   hipEvent_t prof_events[4] = {};
   // [0] Mutations counting kernel, [1] filter kernel, [2] dense counting kernel of batch jobs (the
   // jobs then run without graphs, with events on their lane's stream)
   double prof_ms[3] = {0, 0, 0};
   uint64_t prof_launches[3] = {0, 0, 0};
   uint64_t prof_dense_queries = 0;  // dense Mutations queries of the timed batch jobs
   void profileAccumulate(int which);  // waits for the pair's stop event and adds its time
   // tuning knobs (rdb_set_tuning)
   long long sparse_max_rows = -1;     // filters with at most this many rows take the sparse path
   uint32_t mutation_groups = 0;       // chunk groups of the dense Mutations kernel (0 = automatic)
   uint32_t mutation_unroll = 4;       // 16-byte loads in flight per lane in the dense Mutations kernel
   uint32_t mutation_packed = 0;       // dense Mutations kernel with 16-bit packed LDS counters
   uint32_t mutation_coop = 1;         // dense Mutations kernel: a work-group's wavefronts share each tile
   uint32_t batch_max_queries = 256;   // queries per lane job of the batched path
   uint32_t batch_lanes = 2;           // lanes (streams with their own buffers) per table
   uint32_t batch_graphs = 1;          // capture each lane's launch sequence into a HIP graph
   uint32_t batch_row_capacity = 4096; // Mutations rows per query a lane job keeps (more: rerun alone)
   uint32_t batch_shared_scan = 0;     // dense Mutations of a job: 1 four, 2 eight queries per pass over the index
   uint32_t batch_octet_threads = 1024;  // work-group size of the eight-query scan (512 or 1024)
   uint32_t batch_octet_groups = 0;      // its persistent work-groups (0: one per CU)
   // fault injection for tests: the n-th lane job enqueued from now on fails with RDB_ENOMEM after
   // its inputs were staged (0: never)
   uint32_t batch_fail_job = 0;
   // sequence index build: 1 local-reference adaptation and container entries on the device, 0 host
   uint32_t index_build_device = 1;
   // planner of the batched path: 0 reference order, 1 independence estimates, 2 learned estimator
   int plan_mode = 0;
   struct Mlp* planner = nullptr;
   void init(int dev);
   ~Context();
};

// ---- containers (host build) -------------------------------------------------------------------
struct ArenaBuilder {
   std::vector<Desc> desc;
   std::vector<uint16_t> arr;
   std::vector<uint32_t> runs;
   std::vector<uint64_t> bits;
   // Appends the container of the sorted, distinct `rows` of one 2^16 block and returns its descriptor.
   Desc add(const uint16_t* rows, uint32_t n, uint32_t slot);
};

// Bytes of a container as the reference stores it: array 2*card, bitset 8192, run 4*runs.
uint64_t containerPayloadBytes(const Desc& d);

// ---- columns --------------------------------------------------------------------------------------
constexpr uint32_t kBuildPhases = 7;

struct SeqColumn {
   std::string name;
   int alphabet = 0;
   const Alphabet* abc = nullptr;
   uint32_t genome_length = 0;
   std::vector<uint8_t> reference;
   std::vector<uint8_t> local_ref;
   // a local reference given with the column (a finalized upload): used as is instead of adapting
   std::vector<uint8_t> local_ref_given;
   // host copies used when compiling filter leaves
   std::vector<uint32_t> block_pos_start, block_slot_start, pos_slot_ptr, slot_info;
   std::vector<uint32_t> block_of_pos;
   uint32_t null_bitmap_id = kInvalidId;
   bool has_nulls = false;
   SeqColumnView view{};
   std::vector<DeviceBuffer> buffers;
   uint64_t n_containers = 0, payload_bytes = 0, desc_bytes = 0;
   std::vector<uint64_t> chunk_index_bytes;  // per chunk: descriptors + payloads of stored symbols
   std::vector<uint64_t> chunk_nruns;        // per chunk: internal missing runs
   uint32_t n_flipped = 0;
   std::vector<uint32_t> slot_rows;          // rows per slot over all chunks (planner statistics)
   std::vector<uint32_t> coverage;           // covered rows per position (planner statistics)
   uint64_t row_major_bytes = 0;             // row-major mutation CSR (sparse path)
   RawVector<uint32_t> row_mut_ptr, row_nrun_ptr;  // host copies over padded rows (traffic accounting)
   DeviceBuffer sparse_counts, sparse_bnd;   // zero between calls; cleared by the sparse finalize
   // wall time (ms) of the index build's phases: checks, adaptation, containers, slots and tiles,
   // row-level arrays, layout checks, upload
   double build_ms[kBuildPhases] = {};
};

struct IntColumn {
   std::string name;
   bool is_date = false;
   bool is_sorted = true;
   std::vector<int32_t> values;  // dense row order
   std::vector<int32_t> sorted_values;  // non-null values, sorted (planner statistics)
   uint64_t null_count = 0;
   uint32_t null_bitmap_id = kInvalidId;
   IntColumnView view{};
   std::vector<DeviceBuffer> buffers;
};

struct DictColumn {
   std::string name;
   std::vector<std::string> dictionary;
   std::unordered_map<std::string, uint32_t> ids;
   std::vector<uint32_t> value_bitmap;      // value id -> store id (kInvalidId when no row has it)
   std::vector<uint32_t> sublineage_bitmap; // value id -> store id (lineage index)
   std::vector<uint32_t> lineage_bitmap;    // value id -> store id of the rows whose value resolves to
                                            // the same lineage (aliases resolved; lineage index)
   uint32_t null_bitmap_id = kInvalidId;
   bool has_nulls = false;
   bool lineage = false;
   std::vector<uint32_t> rank;              // value id -> rank by string order
   DeviceBuffer labels;                     // per padded row: id, n_values for null, kInvalidId outside
};

struct BatchLane;

// Deepest filter program (operand stack entries) the filter kernel's LDS holds on this device.
uint32_t maxFilterDepth(const Context& ctx);

struct Table {
   Context* ctx = nullptr;
   std::vector<uint32_t> chunk_sizes;
   std::vector<uint64_t> chunk_offset;
   uint32_t n_chunks = 0;
   uint64_t num_rows = 0;
   bool finalized = false;

   std::vector<std::unique_ptr<SeqColumn>> seq;
   std::vector<std::unique_ptr<IntColumn>> ints;
   std::vector<std::unique_ptr<DictColumn>> dicts;
   std::vector<uint32_t> literal_bitmaps;  // literal id -> store id

   // bitmap store: rows per (bitmap, chunk), uploaded at finalize
   std::vector<std::vector<std::vector<uint16_t>>> store_rows;
   std::vector<uint64_t> store_card;  // rows of each store bitmap (planner statistics)
   ArenaBuilder store_builder;
   std::vector<DeviceBuffer> store_buffers;

   DeviceBuffer chunk_sizes_dev, seq_views_dev, int_views_dev;
   TableView view{};

   // per-call scratch (sized at finalize)
   DeviceBuffer slot_counts, heavy_sums, totals, mut_rows, mut_n, labels_scratch[4], hist;
   size_t hist_bins_capacity = 0;
   // lanes of the batched path (created on demand)
   std::vector<std::shared_ptr<BatchLane>> lanes;
   uint64_t job_counter = 0;

   uint32_t addStoreBitmap(const std::vector<uint32_t>& sorted_global_rows);
   // a store bitmap's rows split by chunk (sorted, unique; thread-safe), and its registration
   std::vector<std::vector<uint16_t>> storeChunks(const std::vector<uint32_t>& global_rows) const;
   uint32_t addStoreChunks(std::vector<std::vector<uint16_t>>&& per_chunk);
   uint32_t globalRow(uint64_t dense) const;

   const SeqColumn& seqColumn(const std::string& name) const;
   const IntColumn* intColumn(const std::string& name) const;
   const DictColumn* dictColumn(const std::string& name) const;
   int seqIndex(const std::string& name) const;
   int intIndex(const std::string& name) const;
};

// Builds a sequence column's device layout from its rows in diff form.  `trusted`: the rows come
// from the device diff (mutations ascending and valid, runs internal, sorted and disjoint), so they
// are not checked or normalized again.
void buildSequenceColumn(Table& t, SeqColumn& col, const uint8_t* is_null, const uint32_t* cov_start,
                         const uint32_t* cov_end, const uint64_t* mut_ptr, const uint32_t* mut_pos,
                         const uint8_t* mut_sym, const uint64_t* nrun_ptr, const uint32_t* nrun_start,
                         const uint32_t* nrun_end, bool trusted = false);
void buildIntColumn(Table& t, IntColumn& col, const int32_t* values, const uint8_t* is_null);

// A sequence column in the state the reference's SequenceColumn holds after finalize, brought to
// the diff form buildSequenceColumn takes (mutations against the column's reference, missing runs,
// covered ranges).  Containers are CRoaring containers as container_write writes them.
struct FinalizedSequence {
   const uint8_t* local_reference = nullptr;
   uint64_t n_containers = 0;
   const uint32_t* key_position = nullptr;
   const uint16_t* key_chunk = nullptr;
   const uint8_t* key_symbol = nullptr;
   const uint8_t* typecode = nullptr;  // 1 bitset, 2 array, 3 run (CRoaring's container types)
   const uint32_t* cardinality = nullptr;
   const uint64_t* payload_offset = nullptr;
   const uint8_t* payload = nullptr;
   uint64_t payload_bytes = 0;
   const uint8_t* is_null = nullptr;
   const uint32_t* cov_start = nullptr;
   const uint32_t* cov_end = nullptr;
   const uint64_t* missing_ptr = nullptr;  // [n_rows + 1]
   const uint32_t* missing_pos = nullptr;
   // or the missing positions as runs [start, end) per row (CSR, run_ptr[n_rows + 1])
   const uint64_t* run_ptr = nullptr;
   const uint32_t* run_start = nullptr;
   const uint32_t* run_end = nullptr;
};
void buildSequenceColumnFinalized(Table& t, SeqColumn& col, const FinalizedSequence& f);
// lineage_parent (-1 = root) builds the lineage index; lineage_alias (-1 = none) names the value
// an alias resolves to (lineage_tree.cpp:164-169), only with lineage_parent
void buildDictColumn(Table& t, DictColumn& col, const uint32_t* ids, const uint8_t* is_null,
                     const int32_t* lineage_parent, const int32_t* lineage_alias = nullptr);
void finalizeTable(Table& t);

// ---- device row sets ---------------------------------------------------------------------------
// The engine's counterpart of CopyOnWriteBitmap (query_engine/copy_on_write_bitmap.h:28-147):
// dense 1024 x u64 words per chunk on the device, with the total and per-chunk cardinalities.
struct DeviceBitmap {
   Table* table = nullptr;
   uint64_t* words = nullptr;
   uint32_t* chunk_cards = nullptr;
   unsigned long long* card_dev = nullptr;
   size_t bytes = 0;
   uint64_t card = 0;
   bool card_known = false;

   DeviceBitmap() = default;
   explicit DeviceBitmap(Table* t);
   ~DeviceBitmap();
   DeviceBitmap(DeviceBitmap&& o) noexcept;
   DeviceBitmap& operator=(DeviceBitmap&& o) noexcept;
   DeviceBitmap(const DeviceBitmap&) = delete;
   DeviceBitmap& operator=(const DeviceBitmap&) = delete;

   uint64_t cardinality();
   std::vector<uint32_t> rows();
};

}  // namespace rdb

struct rdb_ctx {
   rdb::Context c;
};
struct rdb_table {
   rdb::Table t;
};
struct rdb_bitmap {
   rdb::DeviceBitmap b;
};
