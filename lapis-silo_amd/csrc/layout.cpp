// Building the device-resident layout of a table from post-diff ingest arrays.
//
// The sequence-column build restates SequenceColumn::appendChunk / finalize
// (storage/column/sequence_column.cpp:99-212): containers per (position, chunk, symbol) for every
// symbol that differs from the local reference, with the per-position local reference adapted to
// the majority symbol of the covered rows (vertical_sequence_index.cpp:56-164, strict '>' in symbol
// order, coverage excluding missing rows).  The containers are then cut into (position block,
// chunk) tiles for the Mutations kernel and uploaded once.
#include <algorithm>
#include <bit>
#include <chrono>
#include <cstring>
#include <numeric>
#include <atomic>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>

#include "engine.h"
#include "ingest.h"

namespace rdb {

namespace {

// Host worker threads started once and reused by every parallelFor that asks for all of them, so
// that the short parallel sections of a batch submission (compiling, planning and emitting a
// thousand filters) do not pay for starting threads each time.
thread_local bool t_pool_worker = false;

class WorkerPool {
  public:
   explicit WorkerPool(unsigned workers) {
      for (unsigned k = 0; k < workers; ++k) threads_.emplace_back([this] { loop(); });
   }
   // Runs f(0 .. n-1) on the workers and the calling thread; false (nothing run) when another job
   // holds the pool.
   bool run(size_t n, const std::function<void(size_t)>& f) {
      std::unique_lock<std::mutex> job_lock(job_mu_, std::try_to_lock);
      if (!job_lock.owns_lock()) return false;
      // the calling thread works as one of the pool's threads: a parallelFor inside f runs serially
      t_pool_worker = true;
      struct Reset {
         ~Reset() { t_pool_worker = false; }
      } reset;
      auto job = std::make_shared<Job>();
      job->f = &f;
      job->n = n;
      {
         std::lock_guard<std::mutex> lock(mu_);
         current_ = job;
      }
      cv_.notify_all();
      work(*job);
      {
         // every item has finished (a worker may still be inside f until then)
         std::unique_lock<std::mutex> lock(mu_);
         done_cv_.wait(lock, [&] { return job->done.load() == n; });
         current_.reset();
      }
      if (job->error) std::rethrow_exception(job->error);
      return true;
   }

  private:
   struct Job {
      const std::function<void(size_t)>* f = nullptr;
      size_t n = 0;
      std::atomic<size_t> next{0};
      std::atomic<size_t> done{0};
      std::exception_ptr error;
      std::mutex error_mu;
   };

   void work(Job& j) {
      for (;;) {
         const size_t i = j.next.fetch_add(1);
         if (i >= j.n) return;
         try {
            (*j.f)(i);
         } catch (...) {
            std::lock_guard<std::mutex> lock(j.error_mu);
            if (!j.error) j.error = std::current_exception();
         }
         if (j.done.fetch_add(1) + 1 == j.n) {
            std::lock_guard<std::mutex> lock(mu_);
            done_cv_.notify_all();
         }
      }
   }

   void loop() {
      t_pool_worker = true;
      std::shared_ptr<Job> last;
      for (;;) {
         std::shared_ptr<Job> job;
         {
            std::unique_lock<std::mutex> lock(mu_);
            cv_.wait(lock, [&] { return current_ != nullptr && current_ != last; });
            job = current_;
         }
         last = job;
         work(*job);
      }
   }

   std::vector<std::thread> threads_;
   std::mutex job_mu_, mu_;
   std::condition_variable cv_, done_cv_;
   std::shared_ptr<Job> current_;
};

unsigned defaultThreads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

}  // namespace

void parallelFor(size_t n, const std::function<void(size_t)>& f, unsigned threads) {
   const unsigned all = defaultThreads();
   if (threads == 0) threads = all;
   if (n <= 1 || threads <= 1 || t_pool_worker) {
      for (size_t i = 0; i < n; ++i) f(i);
      return;
   }
   if (threads >= all) {
      // created on first use and kept for the life of the process (its workers wait on a condition)
      static WorkerPool* pool = new WorkerPool(all - 1);
      if (pool->run(n, f)) return;
   }
   std::atomic<size_t> next{0};
   std::vector<std::thread> pool;
   std::exception_ptr error;
   std::mutex error_mu;
   for (unsigned k = 0; k < std::min<size_t>(threads, n); ++k) {
      pool.emplace_back([&] {
         for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n) break;
            try {
               f(i);
            } catch (...) {
               std::lock_guard<std::mutex> lock(error_mu);
               if (!error) error = std::current_exception();
            }
         }
      });
   }
   for (auto& th : pool) th.join();
   if (error) std::rethrow_exception(error);
}

// ---- alphabets ---------------------------------------------------------------------------------
int Alphabet::fromChar(char c) const {
   char u = c;
   if (u >= 'a' && u <= 'z') u = static_cast<char>(u - 'a' + 'A');
   if (count == 16 && u == 'U') u = 'T';
   const auto p = chars.find(u);
   return p == std::string::npos ? -1 : static_cast<int>(p);
}

namespace {
void deriveAmbiguity(Alphabet& a) {
   a.ambiguity.assign(a.count, {});
   for (uint8_t s = 0; s < a.count; ++s) {
      for (uint8_t cand = 0; cand < a.count; ++cand) {
         const auto& cs = a.codes_for[s];
         const auto& cc = a.codes_for[cand];
         const bool superset =
            std::all_of(cs.begin(), cs.end(), [&](uint8_t x) { return std::find(cc.begin(), cc.end(), x) != cc.end(); });
         if (superset) a.ambiguity[s].push_back(cand);
      }
   }
}
}  // namespace

const Alphabet& alphabetFor(int alphabet) {
   static const Alphabet nuc = [] {
      Alphabet a;
      a.count = 16;
      a.chars = "-ACGTRYSWKMBDHVN";
      a.missing = 15;
      a.valid = {0, 1, 2, 3, 4};
      a.codes_for = {{0}, {1}, {2}, {3}, {4}, {1, 3}, {2, 4}, {2, 3}, {1, 4}, {3, 4}, {1, 2}, {2, 3, 4}, {1, 3, 4}, {1, 2, 4}, {1, 2, 3}, {}};
      for (uint8_t s = 0; s < 16; ++s) a.codes_for[15].push_back(s);
      deriveAmbiguity(a);
      return a;
   }();
   static const Alphabet aa = [] {
      Alphabet a;
      a.count = 28;
      a.chars = "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X";
      a.missing = 27;
      for (uint8_t s = 0; s < 23; ++s) a.valid.push_back(s);
      a.valid.push_back(26);
      a.codes_for.resize(28);
      for (uint8_t s = 0; s < 23; ++s) a.codes_for[s] = {s};
      a.codes_for[23] = {3, 12};
      a.codes_for[24] = {8, 10};
      a.codes_for[25] = {4, 15};
      a.codes_for[26] = {26};
      for (uint8_t s = 0; s < 28; ++s) a.codes_for[27].push_back(s);
      deriveAmbiguity(a);
      return a;
   }();
   return alphabet == RDB_AMINO_ACID ? aa : nuc;
}

// ---- device memory -----------------------------------------------------------------------------
DeviceBuffer::DeviceBuffer(size_t bytes) : bytes_(bytes) {
   if (bytes_ == 0) return;
   if (hipMalloc(&ptr_, bytes_) != hipSuccess) {
      ptr_ = nullptr;
      fail(RDB_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
   }
}

DeviceBuffer::DeviceBuffer(const void* host, size_t bytes) : DeviceBuffer(std::max<size_t>(bytes, 16)) {
   // empty arrays still get a (small) allocation so that every device pointer is valid; the bytes
   // past a short array are zero, so equal arrays make equal buffers
   if (bytes < 16) hipCheck(hipMemset(ptr_, 0, 16), "clear");
   if (bytes > 0 && host != nullptr) hipCheck(hipMemcpy(ptr_, host, bytes, hipMemcpyHostToDevice), "upload");
}

DeviceBuffer::~DeviceBuffer() {
   if (ptr_) (void)hipFree(ptr_);
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
   if (this != &o) {
      if (ptr_) (void)hipFree(ptr_);
      ptr_ = o.ptr_;
      bytes_ = o.bytes_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
   }
   return *this;
}

DevicePool::~DevicePool() {
   for (auto& [size, list] : free_) {
      for (void* p : list) (void)hipFree(p);
   }
}

void* DevicePool::get(size_t bytes) {
   {
      std::lock_guard<std::mutex> lock(mu_);
      auto it = free_.find(bytes);
      if (it != free_.end() && !it->second.empty()) {
         void* p = it->second.back();
         it->second.pop_back();
         return p;
      }
   }
   void* p = nullptr;
   if (hipMalloc(&p, bytes) != hipSuccess) fail(RDB_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
   return p;
}

void DevicePool::put(void* p, size_t bytes) {
   if (!p) return;
   std::lock_guard<std::mutex> lock(mu_);
   free_[bytes].push_back(p);
}

void* Context::pinnedBuffer(size_t bytes) {
   if (bytes > pinned_bytes) {
      if (pinned) (void)hipHostFree(pinned);
      pinned = nullptr;
      pinned_bytes = 0;
      hipCheck(hipHostMalloc(&pinned, bytes, hipHostMallocDefault), "pinned allocation");
      pinned_bytes = bytes;
   }
   return pinned;
}

// ---- containers --------------------------------------------------------------------------------
Desc ArenaBuilder::add(const uint16_t* rows, uint32_t n, uint32_t slot) {
   if (n <= 2) {
      uint32_t w1 = 0;
      for (uint32_t k = 0; k < n; ++k) w1 |= static_cast<uint32_t>(rows[k]) << (16 * k);
      Desc d{makeW0(slot, CT_INLINE, n), w1};
      desc.push_back(d);
      return d;
   }
   uint32_t n_runs = 1;
   for (uint32_t k = 1; k < n; ++k) n_runs += rows[k] != rows[k - 1] + 1;
   const uint64_t run_bytes = 4ULL * n_runs;
   const uint64_t plain_bytes = n <= 4096 ? 2ULL * n : 8192ULL;
   Desc d{};
   if (run_bytes < plain_bytes) {
      d = Desc{makeW0(slot, CT_RUN, n_runs), static_cast<uint32_t>(runs.size())};
      uint32_t k = 0;
      while (k < n) {
         uint32_t j = k;
         while (j + 1 < n && rows[j + 1] == rows[j] + 1) ++j;
         runs.push_back(static_cast<uint32_t>(rows[k]) | (static_cast<uint32_t>(rows[j] - rows[k]) << 16));
         k = j + 1;
      }
   } else if (n <= 4096) {
      if (arr.size() + n > 0xFFFFFFFFULL) fail(RDB_ENOMEM, "array arena exceeds 2^32 entries");
      d = Desc{makeW0(slot, CT_ARRAY, n), static_cast<uint32_t>(arr.size())};
      arr.insert(arr.end(), rows, rows + n);
   } else {
      d = Desc{makeW0(slot, CT_BITSET, 0), static_cast<uint32_t>(bits.size() / kChunkWords)};
      const size_t base = bits.size();
      bits.resize(base + kChunkWords, 0);
      for (uint32_t k = 0; k < n; ++k) bits[base + (rows[k] >> 6)] |= 1ULL << (rows[k] & 63);
   }
   desc.push_back(d);
   return d;
}

uint64_t containerPayloadBytes(const Desc& d) {
   switch (descType(d.w0)) {
      case CT_INLINE:
      case CT_ARRAY:
         return 2ULL * descCard(d.w0);
      case CT_RUN:
         return 4ULL * descCard(d.w0);
      default:
         return 8192;
   }
}

// Every descriptor must point inside its arenas: checked on the host before upload so that a layout
// bug surfaces as an error and never as an out-of-bounds device access.
void validateDescriptors(const Desc* desc, size_t n, size_t arr_size, size_t run_size, size_t bit_blocks,
                         uint32_t slot_limit) {
   for (size_t k = 0; k < n; ++k) {
      const Desc& d = desc[k];
      const uint32_t type = descType(d.w0), card = descCard(d.w0);
      bool ok = descSlot(d.w0) < slot_limit;
      switch (type) {
         case CT_INLINE:
            ok = ok && card <= 2;
            break;
         case CT_ARRAY:
            ok = ok && card <= 4096 && static_cast<uint64_t>(d.w1) + card <= arr_size;
            break;
         case CT_RUN:
            ok = ok && card > 0 && static_cast<uint64_t>(d.w1) + card <= run_size;
            break;
         default:
            ok = ok && d.w1 < bit_blocks;
            break;
      }
      if (!ok) fail(RDB_EDEVICE, "internal: invalid container descriptor " + std::to_string(k));
   }
}

// ---- table helpers -----------------------------------------------------------------------------
uint32_t Table::globalRow(uint64_t dense) const {
   const auto it = std::upper_bound(chunk_offset.begin(), chunk_offset.end(), dense) - 1;
   const uint64_t c = static_cast<uint64_t>(it - chunk_offset.begin());
   return static_cast<uint32_t>((c << 16) | (dense - *it));
}

std::vector<std::vector<uint16_t>> Table::storeChunks(const std::vector<uint32_t>& rows) const {
   std::vector<std::vector<uint16_t>> per_chunk(n_chunks);
   for (uint32_t g : rows) {
      const uint32_t c = g >> 16;
      if (c >= n_chunks) fail(RDB_EINVAL, "row id " + std::to_string(g) + " lies outside the table's chunks");
      per_chunk[c].push_back(static_cast<uint16_t>(g & 0xFFFF));
   }
   for (auto& v : per_chunk) {
      if (!std::is_sorted(v.begin(), v.end())) std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
   }
   return per_chunk;
}

uint32_t Table::addStoreChunks(std::vector<std::vector<uint16_t>>&& per_chunk) {
   store_rows.push_back(std::move(per_chunk));
   return static_cast<uint32_t>(store_rows.size() - 1);
}

uint32_t Table::addStoreBitmap(const std::vector<uint32_t>& rows) { return addStoreChunks(storeChunks(rows)); }

const SeqColumn& Table::seqColumn(const std::string& name) const {
   for (const auto& c : seq) {
      if (c->name == name) return *c;
   }
   fail(RDB_EQUERY, "Database does not contain the sequence '" + name + "'");
}

int Table::seqIndex(const std::string& name) const {
   for (size_t i = 0; i < seq.size(); ++i) {
      if (seq[i]->name == name) return static_cast<int>(i);
   }
   return -1;
}

const IntColumn* Table::intColumn(const std::string& name) const {
   for (const auto& c : ints) {
      if (c->name == name) return c.get();
   }
   return nullptr;
}

int Table::intIndex(const std::string& name) const {
   for (size_t i = 0; i < ints.size(); ++i) {
      if (ints[i]->name == name) return static_cast<int>(i);
   }
   return -1;
}

const DictColumn* Table::dictColumn(const std::string& name) const {
   for (const auto& c : dicts) {
      if (c->name == name) return c.get();
   }
   return nullptr;
}

// ---- sequence column ---------------------------------------------------------------------------
namespace {


struct ChunkContainers {
   std::vector<uint32_t> key;  // position << 8 | symbol, sorted
   std::vector<uint32_t> rows; // rows of each container
   ArenaBuilder arenas;
};

constexpr uint32_t kBoundaryCount = 4;

bool inRuns(const uint32_t* runs, uint64_t n, uint32_t p) {
   // runs sorted, disjoint: start | end << 16
   uint64_t lo = 0, hi = n;
   while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if ((runs[mid] & 0xFFFFu) <= p) {
         lo = mid + 1;
      } else {
         hi = mid;
      }
   }
   return lo > 0 && p < (runs[lo - 1] >> 16);
}

}  // namespace

void buildSequenceColumn(Table& t, SeqColumn& col, const uint8_t* is_null, const uint32_t* cov_start,
                         const uint32_t* cov_end, const uint64_t* mut_ptr, const uint32_t* mut_pos,
                         const uint8_t* mut_sym, const uint64_t* nrun_ptr, const uint32_t* nrun_start,
                         const uint32_t* nrun_end, bool trusted) {
   // wall time of the build's phases (SeqColumn::build_ms)
   auto clock0 = std::chrono::steady_clock::now();
   auto phase = [&](int k) {
      const auto now = std::chrono::steady_clock::now();
      col.build_ms[k] = std::chrono::duration<double, std::milli>(now - clock0).count();
      clock0 = now;
   };
   const uint32_t L = col.genome_length;
   const Alphabet& A = *col.abc;
   const uint32_t nsym = A.count;
   const uint64_t n = t.num_rows;
   if (L == 0 || L > 65535) fail(RDB_EINVAL, "genome length must be in [1, 65535]");
   for (uint32_t p = 0; p < L; ++p) {
      if (col.reference[p] >= nsym) fail(RDB_EINVAL, "invalid reference symbol");
   }

   // normalized internal missing runs: clipped to [start, end) as insertCoverage does
   // (horizontal_coverage_index.cpp:34-37), sorted and merged
   RawVector<uint64_t> nptr(n + 1);  // every entry written below
   nptr[0] = 0;
   std::vector<std::vector<uint32_t>> row_runs_tmp(trusted ? 0 : t.n_chunks);
   if (!trusted) parallelFor(t.n_chunks, [&](size_t c) {
      auto& out = row_runs_tmp[c];
      std::vector<std::pair<uint32_t, uint32_t>> runs;
      for (uint64_t i = t.chunk_offset[c]; i < t.chunk_offset[c + 1]; ++i) {
         runs.clear();
         if (!is_null[i]) {
            if (cov_start[i] > cov_end[i] || cov_end[i] > L) fail(RDB_EINVAL, "invalid coverage range");
            for (uint64_t k = mut_ptr[i]; k < mut_ptr[i + 1]; ++k) {
               if (mut_pos[k] >= L) fail(RDB_EINVAL, "mutation position out of range");
               if (mut_sym[k] >= nsym || mut_sym[k] == A.missing) fail(RDB_EINVAL, "invalid mutation symbol");
               if (k > mut_ptr[i] && mut_pos[k] <= mut_pos[k - 1]) {
                  fail(RDB_EINVAL, "mutation positions of a row must be strictly increasing");
               }
            }
            for (uint64_t k = nrun_ptr[i]; k < nrun_ptr[i + 1]; ++k) {
               const uint32_t a = std::max(nrun_start[k], cov_start[i]);
               const uint32_t b = std::min(nrun_end[k], cov_end[i]);
               if (a < b) runs.emplace_back(a, b);
            }
            std::sort(runs.begin(), runs.end());
         }
         uint32_t count = 0;
         for (size_t k = 0; k < runs.size(); ++k) {
            if (count > 0) {
               uint32_t& last = out.back();
               const uint32_t le = last >> 16;
               if (runs[k].first <= le) {
                  const uint32_t ne = std::max(le, runs[k].second);
                  last = (last & 0xFFFFu) | (ne << 16);
                  continue;
               }
            }
            out.push_back(runs[k].first | (runs[k].second << 16));
            ++count;
         }
         nptr[i + 1] = count;
         if (!is_null[i]) {
            // the diff form of aligned_sequence.cpp:18-122: a mutation lies inside the covered
            // range, outside every missing run, and differs from the reference symbol
            const uint32_t* rr = out.data() + (out.size() - count);
            uint32_t j = 0;
            for (uint64_t k = mut_ptr[i]; k < mut_ptr[i + 1]; ++k) {
               const uint32_t p = mut_pos[k];
               if (p < cov_start[i] || p >= cov_end[i]) fail(RDB_EINVAL, "mutation outside the covered range");
               if (mut_sym[k] == col.reference[p]) fail(RDB_EINVAL, "mutation equal to the reference symbol");
               while (j < count && (rr[j] >> 16) <= p) ++j;
               if (j < count && (rr[j] & 0xFFFFu) <= p) fail(RDB_EINVAL, "mutation inside a missing run");
            }
         }
      }
   });
   RawVector<uint32_t> runs_all;
   if (trusted) {
      // rows of the device diff: runs already internal, sorted and disjoint
      std::copy(nrun_ptr, nrun_ptr + n + 1, nptr.begin());
      runs_all.resize(nptr[n]);
      for (uint64_t k = 0; k < nptr[n]; ++k) runs_all[k] = nrun_start[k] | (nrun_end[k] << 16);
   } else {
      // nptr[i + 1] holds row i's run count: offsets per chunk from the chunks' totals, then each
      // chunk's running sums and its runs in parallel
      std::vector<uint64_t> chunk_base(t.n_chunks + 1, 0);
      for (uint32_t c = 0; c < t.n_chunks; ++c) chunk_base[c + 1] = chunk_base[c] + row_runs_tmp[c].size();
      runs_all.resize(chunk_base[t.n_chunks]);
      parallelFor(t.n_chunks, [&](size_t c) {
         uint64_t acc = chunk_base[c];
         for (uint64_t i = t.chunk_offset[c]; i < t.chunk_offset[c + 1]; ++i) {
            acc += nptr[i + 1];
            nptr[i + 1] = acc;
         }
         auto& v = row_runs_tmp[c];
         std::copy(v.begin(), v.end(), runs_all.begin() + static_cast<ptrdiff_t>(chunk_base[c]));
         std::vector<uint32_t>().swap(v);
      });
      if (nptr[n] != chunk_base[t.n_chunks]) fail(RDB_EDEVICE, "internal: missing-run offsets");
   }

   phase(0);
   // --- local reference adaptation (sequence_column.cpp:157-212) ---------------------------------
   // On the device (the default) the coverage and symbol counts and the rule below run as kernels
   // over the uploaded rows, which then also generate and sort the container entries.
   std::unique_ptr<DeviceIndexBuild> dev;
   std::vector<uint32_t> dev_coverage;
   std::vector<uint8_t> dev_best;
   std::vector<uint32_t> sym_counts;
   std::vector<int64_t> changes;
   if (t.ctx->index_build_device) {
      dev = std::make_unique<DeviceIndexBuild>(t, col, is_null, cov_start, cov_end, mut_ptr, mut_pos, mut_sym,
                                               nptr.data(), runs_all.data());
      dev->adapt(dev_coverage, dev_best);
   } else {
      sym_counts.assign(static_cast<size_t>(L) * nsym, 0);
      changes.assign(L + 1, 0);
      for (uint64_t i = 0; i < n; ++i) {
         if (is_null[i]) continue;
         changes[cov_start[i]] += 1;
         changes[cov_end[i]] -= 1;
         for (uint64_t k = nptr[i]; k < nptr[i + 1]; ++k) {
            changes[runs_all[k] & 0xFFFFu] -= 1;
            changes[runs_all[k] >> 16] += 1;
         }
         for (uint64_t k = mut_ptr[i]; k < mut_ptr[i + 1]; ++k) sym_counts[static_cast<size_t>(mut_pos[k]) * nsym + mut_sym[k]] += 1;
      }
   }
   col.local_ref = col.reference;
   std::vector<uint8_t> flip_to(L, 0xFF);
   std::vector<uint32_t> flipped;
   int64_t coverage = 0;
   std::vector<uint32_t> counts(nsym);
   col.coverage.assign(L, 0);
   for (uint32_t p = 0; p < L; ++p) {
      const uint8_t current = col.reference[p];
      uint8_t best = current;
      if (dev) {
         col.coverage[p] = dev_coverage[p];
         best = dev_best[p];
      } else {
         coverage += changes[p];
         col.coverage[p] = static_cast<uint32_t>(coverage);
         for (uint32_t s = 0; s < nsym; ++s) counts[s] = sym_counts[static_cast<size_t>(p) * nsym + s];
         uint32_t ref_count = static_cast<uint32_t>(coverage);
         for (uint32_t s = 0; s < nsym; ++s) {
            if (s != current) ref_count -= counts[s];
         }
         counts[current] = ref_count;
         uint32_t best_count = counts[current];
         for (uint32_t s = 0; s < nsym; ++s) {
            if (s == current) continue;
            if (counts[s] > best_count) {
               best = static_cast<uint8_t>(s);
               best_count = counts[s];
            }
         }
      }
      // a column uploaded finalized keeps its own local reference: the reference adapts it again
      // after every append relative to the previous one, which ties can leave elsewhere
      if (!col.local_ref_given.empty()) best = col.local_ref_given[p];
      if (best != current) {
         flip_to[p] = best;
         flipped.push_back(p);
         col.local_ref[p] = best;
      }
   }
   col.n_flipped = static_cast<uint32_t>(flipped.size());
   std::vector<int64_t>().swap(changes);
   std::vector<uint32_t>().swap(sym_counts);

   phase(1);
   // --- per-chunk containers ----------------------------------------------------------------------
   // entries position << 24 | symbol << 16 | row, sorted per chunk (from the device: with the chunk
   // in bits 40 and up, all chunks in one array)
   RawVector<uint64_t> dev_keys;
   std::vector<uint64_t> dev_chunk_ptr;
   if (dev) {
      dev->entries(flip_to, flipped, dev_keys, dev_chunk_ptr);
      dev.reset();
   }
   std::vector<ChunkContainers> chunks(t.n_chunks);
   parallelFor(t.n_chunks, [&](size_t c) {
      std::vector<uint64_t> entries;
      const uint64_t* ebegin = nullptr;
      size_t ecount = 0;
      if (!dev_chunk_ptr.empty()) {
         ebegin = dev_keys.data() + dev_chunk_ptr[c];
         ecount = dev_chunk_ptr[c + 1] - dev_chunk_ptr[c];
      }
      const uint64_t off = t.chunk_offset[c];
      const uint32_t size = t.chunk_sizes[c];
      auto push = [&](uint32_t p, uint32_t sym, uint32_t r) {
         entries.push_back((static_cast<uint64_t>(p) << 24) | (static_cast<uint64_t>(sym) << 16) | r);
      };
      for (uint32_t r = 0; r < (ebegin ? 0 : size); ++r) {
         const uint64_t i = off + r;
         if (is_null[i]) {
            push(0, nsym + BND_COV_END, r);  // no covered range: uncovered everywhere
            continue;
         }
         // an empty covered range is stored as [0, 0): uncovered from position 0 on
         const bool has_cov = cov_start[i] < cov_end[i];
         const uint32_t s0 = has_cov ? cov_start[i] : 0, e0 = has_cov ? cov_end[i] : 0;
         for (uint64_t k = mut_ptr[i]; k < mut_ptr[i + 1]; ++k) {
            const uint32_t p = mut_pos[k];
            if (flip_to[p] != 0xFF && mut_sym[k] == flip_to[p]) continue;
            push(p, mut_sym[k], r);
         }
         for (uint32_t p : flipped) {
            if (p < s0 || p >= e0) continue;
            if (inRuns(runs_all.data() + nptr[i], nptr[i + 1] - nptr[i], p)) continue;
            const uint32_t* mb = mut_pos + mut_ptr[i];
            const uint32_t* me = mut_pos + mut_ptr[i + 1];
            const uint32_t* it = std::lower_bound(mb, me, p);
            if (it != me && *it == p) continue;
            push(p, col.reference[p], r);
         }
         if (s0 > 0) push(s0, nsym + BND_COV_START, r);
         if (e0 < L) push(e0, nsym + BND_COV_END, r);
         for (uint64_t k = nptr[i]; k < nptr[i + 1]; ++k) {
            const uint32_t a = runs_all[k] & 0xFFFFu, b = runs_all[k] >> 16;
            push(a, nsym + BND_N_START, r);
            if (b < L) push(b, nsym + BND_N_END, r);
         }
      }
      if (!ebegin) {
         std::sort(entries.begin(), entries.end());
         ebegin = entries.data();
         ecount = entries.size();
      }
      auto& out = chunks[c];
      std::vector<uint16_t> rows;
      size_t k = 0;
      constexpr uint64_t kInChunk = (uint64_t{1} << 40) - 1;
      while (k < ecount) {
         const uint64_t key = (ebegin[k] & kInChunk) >> 16;
         rows.clear();
         while (k < ecount && ((ebegin[k] & kInChunk) >> 16) == key) {
            rows.push_back(static_cast<uint16_t>(ebegin[k] & 0xFFFF));
            ++k;
         }
         out.key.push_back(static_cast<uint32_t>(key));
         out.rows.push_back(static_cast<uint32_t>(rows.size()));
         out.arenas.add(rows.data(), static_cast<uint32_t>(rows.size()), 0);
      }
   });

   phase(2);
   // --- slots: distinct (position, symbol) pairs, sorted ------------------------------------------
   const uint32_t total_syms = nsym + kBoundaryCount;
   if (total_syms > 32) fail(RDB_EINVAL, "alphabet too large");
   std::vector<uint32_t> sym_mask(L, 0);
   std::vector<uint64_t> pos_bytes(L, 0);
   col.chunk_index_bytes.assign(t.n_chunks, 0);
   // Per-position statistics in parallel over position ranges: every chunk's keys are sorted by
   // position, so a range's keys of a chunk are found by binary search and each position is written
   // by one range only.  Per-chunk sums are kept per range and added after.
   const uint32_t n_ranges = std::min<uint32_t>(64, L);
   auto rangeKeys = [&](const ChunkContainers& ch, uint32_t r) {
      const uint32_t p0 = static_cast<uint32_t>(uint64_t{L} * r / n_ranges), p1 = static_cast<uint32_t>(uint64_t{L} * (r + 1) / n_ranges);
      const auto a = std::lower_bound(ch.key.begin(), ch.key.end(), p0 << 8);
      const auto e = std::lower_bound(a, ch.key.end(), p1 << 8);
      return std::pair<size_t, size_t>(static_cast<size_t>(a - ch.key.begin()), static_cast<size_t>(e - ch.key.begin()));
   };
   std::vector<uint64_t> range_chunk_bytes(static_cast<size_t>(n_ranges) * t.n_chunks, 0), range_containers(n_ranges, 0),
      range_payload(n_ranges, 0);
   parallelFor(n_ranges, [&](size_t r) {
      for (uint32_t c = 0; c < t.n_chunks; ++c) {
         const auto& ch = chunks[c];
         const auto [k0, k1] = rangeKeys(ch, static_cast<uint32_t>(r));
         for (size_t k = k0; k < k1; ++k) {
            const uint32_t p = ch.key[k] >> 8, s = ch.key[k] & 0xFF;
            sym_mask[p] |= 1u << s;
            const uint64_t b = containerPayloadBytes(ch.arenas.desc[k]) + sizeof(Desc);
            pos_bytes[p] += b;
            if (s < nsym) {
               range_chunk_bytes[r * t.n_chunks + c] += b;
               range_containers[r] += 1;
               range_payload[r] += containerPayloadBytes(ch.arenas.desc[k]);
            }
         }
      }
   });
   for (uint32_t r = 0; r < n_ranges; ++r) {
      for (uint32_t c = 0; c < t.n_chunks; ++c) col.chunk_index_bytes[c] += range_chunk_bytes[static_cast<size_t>(r) * t.n_chunks + c];
      col.n_containers += range_containers[r];
      col.payload_bytes += range_payload[r];
   }
   col.pos_slot_ptr.assign(L + 1, 0);
   for (uint32_t p = 0; p < L; ++p) col.pos_slot_ptr[p + 1] = col.pos_slot_ptr[p] + std::popcount(sym_mask[p]);
   const uint32_t n_slots = col.pos_slot_ptr[L];
   col.slot_info.resize(n_slots);
   for (uint32_t p = 0; p < L; ++p) {
      uint32_t m = sym_mask[p], s = col.pos_slot_ptr[p];
      while (m) {
         const uint32_t sym = static_cast<uint32_t>(std::countr_zero(m));
         m &= m - 1;
         col.slot_info[s++] = p | (sym << 16);
      }
   }

   // rows per slot over all chunks (planner statistics)
   col.slot_rows.assign(n_slots, 0);
   parallelFor(n_ranges, [&](size_t r) {
      for (uint32_t c = 0; c < t.n_chunks; ++c) {
         const auto& ch = chunks[c];
         const auto [k0, k1] = rangeKeys(ch, static_cast<uint32_t>(r));
         for (size_t k = k0; k < k1; ++k) {
            const uint32_t p = ch.key[k] >> 8, sym = ch.key[k] & 0xFF;
            const uint32_t slot = col.pos_slot_ptr[p] + static_cast<uint32_t>(std::popcount(sym_mask[p] & ((1u << sym) - 1u)));
            col.slot_rows[slot] += ch.rows[k];
         }
      }
   });

   // --- position blocks: <= 4096 slots and about equal bytes --------------------------------------
   const uint64_t total_bytes = std::accumulate(pos_bytes.begin(), pos_bytes.end(), uint64_t{0});
   const uint64_t target = std::max<uint64_t>(total_bytes / 64, 1);
   col.block_pos_start.assign(1, 0);
   col.block_of_pos.assign(L, 0);
   {
      uint64_t bytes = 0;
      uint32_t slots = 0;
      for (uint32_t p = 0; p < L; ++p) {
         const uint32_t ps = col.pos_slot_ptr[p + 1] - col.pos_slot_ptr[p];
         if (p > col.block_pos_start.back() && (slots + ps > kMaxBlockSlots || bytes + pos_bytes[p] > target)) {
            col.block_pos_start.push_back(p);
            bytes = 0;
            slots = 0;
         }
         bytes += pos_bytes[p];
         slots += ps;
         col.block_of_pos[p] = static_cast<uint32_t>(col.block_pos_start.size() - 1);
      }
      col.block_pos_start.push_back(L);
   }
   const uint32_t n_blocks = static_cast<uint32_t>(col.block_pos_start.size() - 1);
   col.block_slot_start.resize(n_blocks + 1);
   for (uint32_t b = 0; b <= n_blocks; ++b) col.block_slot_start[b] = col.pos_slot_ptr[col.block_pos_start[b]];

   // --- tiles -------------------------------------------------------------------------------------
   std::vector<size_t> arr_base(t.n_chunks + 1, 0), run_base(t.n_chunks + 1, 0), bit_base(t.n_chunks + 1, 0);
   for (uint32_t c = 0; c < t.n_chunks; ++c) {
      arr_base[c + 1] = arr_base[c] + chunks[c].arenas.arr.size();
      run_base[c + 1] = run_base[c] + chunks[c].arenas.runs.size();
      bit_base[c + 1] = bit_base[c] + chunks[c].arenas.bits.size() / kChunkWords;
   }
   if (arr_base[t.n_chunks] > 0xFFFFFFFFULL || run_base[t.n_chunks] > 0xFFFFFFFFULL || bit_base[t.n_chunks] > 0xFFFFFFFFULL) {
      fail(RDB_ENOMEM, "container arena exceeds 32-bit offsets");
   }
   // Light containers (at most kLightMaxRows rows) become u32 entries slot << 16 | row; the rest stay
   // descriptors into the concatenated arenas.
   const size_t n_tiles = static_cast<size_t>(n_blocks) * t.n_chunks;
   std::vector<uint32_t> light_ptr(n_tiles + 1, 0), heavy_ptr(n_tiles + 1, 0), group_ptr(n_tiles + 1, 0);
   // per (chunk, position): [lo, hi) of the position's light entries, groups and descriptors, and
   // the mask of 4096-row slices that hold any of its rows
   RawVector<uint32_t> pos_ranges(static_cast<size_t>(t.n_chunks) * L * kPosRecord);  // zeroed per chunk below
   auto sliceMask = [&](const ChunkContainers& ch, const Desc& d) -> uint32_t {
      const uint32_t type = descType(d.w0), card = descCard(d.w0);
      uint32_t m = 0;
      if (type == CT_INLINE) {
         for (uint32_t j = 0; j < card; ++j) m |= 1u << (((d.w1 >> (16 * j)) & 0xFFFFu) >> 12);
      } else if (type == CT_ARRAY) {
         for (uint32_t j = 0; j < card; ++j) m |= 1u << (ch.arenas.arr[d.w1 + j] >> 12);
      } else if (type == CT_RUN) {
         for (uint32_t j = 0; j < card; ++j) {
            const uint32_t x = ch.arenas.runs[d.w1 + j];
            const uint32_t a = (x & 0xFFFFu) >> 12, e = ((x & 0xFFFFu) + (x >> 16)) >> 12;
            for (uint32_t sl = a; sl <= e; ++sl) m |= 1u << sl;
         }
      } else {
         const uint64_t* w = ch.arenas.bits.data() + static_cast<size_t>(d.w1) * kChunkWords;
         for (uint32_t sl = 0; sl < kSlicesPerChunk; ++sl) {
            for (uint32_t k = 0; k < kSliceWords; ++k) {
               if (w[sl * kSliceWords + k]) {
                  m |= 1u << sl;
                  break;
               }
            }
         }
      }
      return m;
   };
   // rows a container holds (a bitset counts as a full chunk): at most kLightMaxRows make light entries
   auto containerRows = [&](const ChunkContainers& ch, const Desc& d) -> uint32_t {
      const uint32_t type = descType(d.w0), card = descCard(d.w0);
      if (type == CT_BITSET) return kChunkRows;
      if (type != CT_RUN) return card;
      uint32_t rows = 0;
      for (uint32_t j = 0; j < card; ++j) rows += (ch.arenas.runs[d.w1 + j] >> 16) + 1;
      return rows;
   };
   // Tiles are laid out block-major ((block, chunk) order) but built per chunk in parallel: a first
   // pass counts each tile's light entries (padded to a multiple of 4), groups and descriptors, a
   // prefix sum over the tiles places them, and a second pass writes them.
   std::vector<uint64_t> n_light(n_tiles), n_group(n_tiles), n_desc(n_tiles);
   parallelFor(t.n_chunks, [&](size_t c) {
      const auto& ch = chunks[c];
      size_t k = 0;
      for (uint32_t b = 0; b < n_blocks; ++b) {
         const uint32_t pend = col.block_pos_start[b + 1];
         uint64_t nl = 0, ng = 0, nd = 0;
         for (; k < ch.key.size() && (ch.key[k] >> 8) < pend; ++k) {
            const Desc& d = ch.arenas.desc[k];
            const uint32_t rows = containerRows(ch, d);
            if (rows <= kLightMaxRows) {
               nl += rows;
            } else if (descType(d.w0) == CT_ARRAY) {
               ng += (descCard(d.w0) + 7) / 8;
            } else {
               nd += 1;
            }
         }
         const size_t tile = static_cast<size_t>(b) * t.n_chunks + c;
         n_light[tile] = (nl + 3) & ~uint64_t{3};
         n_group[tile] = ng;
         n_desc[tile] = nd;
      }
   });
   {
      uint64_t l = 0, g = 0, d = 0;
      for (size_t tile = 0; tile < n_tiles; ++tile) {
         light_ptr[tile] = static_cast<uint32_t>(l);
         group_ptr[tile] = static_cast<uint32_t>(g);
         heavy_ptr[tile] = static_cast<uint32_t>(d);
         l += n_light[tile];
         g += n_group[tile];
         d += n_desc[tile];
         if (l > 0xFFFFFFF0ULL || d > 0xFFFFFFF0ULL || g > 0xFFFFFFF0ULL) fail(RDB_ENOMEM, "too many containers");
      }
      light_ptr[n_tiles] = static_cast<uint32_t>(l);
      group_ptr[n_tiles] = static_cast<uint32_t>(g);
      heavy_ptr[n_tiles] = static_cast<uint32_t>(d);
   }
   RawVector<uint32_t> light(light_ptr[n_tiles]);
   RawVector<uint16_t> group_rows(static_cast<size_t>(group_ptr[n_tiles]) * 8);  // 8 per group
   RawVector<uint16_t> group_hdr(group_ptr[n_tiles]);
   RawVector<Desc> desc(heavy_ptr[n_tiles]);
   parallelFor(t.n_chunks, [&](size_t c) {
      const auto& ch = chunks[c];
      std::fill_n(pos_ranges.data() + c * L * kPosRecord, static_cast<size_t>(L) * kPosRecord, 0u);
      size_t k = 0;
      for (uint32_t b = 0; b < n_blocks; ++b) {
         const uint32_t pend = col.block_pos_start[b + 1];
         const uint32_t sbase = col.block_slot_start[b];
         const size_t tile = static_cast<size_t>(b) * t.n_chunks + c;
         uint32_t li = light_ptr[tile], gi = group_ptr[tile], di = heavy_ptr[tile];
         uint32_t range_pos = kInvalidId;
         for (; k < ch.key.size() && (ch.key[k] >> 8) < pend; ++k) {
            const uint32_t p = ch.key[k] >> 8, s = ch.key[k] & 0xFF;
            uint32_t* pr = pos_ranges.data() + (static_cast<size_t>(c) * L + p) * kPosRecord;
            if (p != range_pos) {
               range_pos = p;
               pr[0] = li;
               pr[2] = gi;
               pr[4] = di;
            }
            const uint32_t slot = col.pos_slot_ptr[p] + std::popcount(sym_mask[p] & ((1u << s) - 1u)) - sbase;
            Desc d = ch.arenas.desc[k];
            pr[6] |= sliceMask(ch, d);
            const uint32_t type = descType(d.w0), card = descCard(d.w0);
            if (containerRows(ch, d) <= kLightMaxRows) {
               const uint32_t hi = slot << 16;
               if (type == CT_INLINE) {
                  for (uint32_t j = 0; j < card; ++j) light[li++] = hi | ((d.w1 >> (16 * j)) & 0xFFFFu);
               } else if (type == CT_ARRAY) {
                  for (uint32_t j = 0; j < card; ++j) light[li++] = hi | ch.arenas.arr[d.w1 + j];
               } else {
                  for (uint32_t j = 0; j < card; ++j) {
                     const uint32_t x = ch.arenas.runs[d.w1 + j];
                     for (uint32_t r = (x & 0xFFFFu); r <= (x & 0xFFFFu) + (x >> 16); ++r) light[li++] = hi | r;
                  }
               }
            } else if (type == CT_ARRAY) {
               const uint16_t* a = ch.arenas.arr.data() + d.w1;
               for (uint32_t j = 0; j < card; j += 8) {
                  const uint32_t m = std::min<uint32_t>(8, card - j);
                  for (uint32_t u = 0; u < 8; ++u) group_rows[static_cast<size_t>(gi) * 8 + u] = u < m ? a[j + u] : 0;
                  group_hdr[gi++] = static_cast<uint16_t>(slot | ((m - 1) << 12));
               }
            } else {
               d.w0 = makeW0(slot, type, card);
               if (type == CT_RUN) d.w1 += static_cast<uint32_t>(run_base[c]);
               if (type == CT_BITSET) d.w1 += static_cast<uint32_t>(bit_base[c]);
               desc[di++] = d;
            }
            pr[1] = li;
            pr[3] = gi;
            pr[5] = di;
         }
         while (li % 4 != 0) light[li++] = kPadSlot << 16;
         if (li != light_ptr[tile + 1] || gi != group_ptr[tile + 1] || di != heavy_ptr[tile + 1]) {
            fail(RDB_EDEVICE, "internal: tile sizes");
         }
      }
   });

   RawVector<uint16_t> arr_all(arr_base[t.n_chunks]);
   RawVector<uint32_t> run_all(run_base[t.n_chunks]);
   RawVector<uint64_t> bit_all(bit_base[t.n_chunks] * kChunkWords);
   parallelFor(t.n_chunks, [&](size_t c) {
      auto& a = chunks[c].arenas;
      std::copy(a.arr.begin(), a.arr.end(), arr_all.begin() + static_cast<ptrdiff_t>(arr_base[c]));
      std::copy(a.runs.begin(), a.runs.end(), run_all.begin() + static_cast<ptrdiff_t>(run_base[c]));
      std::copy(a.bits.begin(), a.bits.end(), bit_all.begin() + static_cast<ptrdiff_t>(bit_base[c] * kChunkWords));
      ArenaBuilder().desc.swap(a.desc);
      std::vector<uint16_t>().swap(a.arr);
      std::vector<uint32_t>().swap(a.runs);
      std::vector<uint64_t>().swap(a.bits);
   });
   chunks.clear();
   col.desc_bytes = desc.size() * sizeof(Desc) + light.size() * 4 + group_hdr.size() * 18;

   phase(3);
   // --- row-level arrays (padded to 2^16 rows per chunk) ------------------------------------------
   const size_t padded = static_cast<size_t>(t.n_chunks) << 16;
   RawVector<uint32_t> cov(padded), pptr(padded + 1);  // every entry written per chunk below
   std::vector<uint64_t> null_bits(static_cast<size_t>(t.n_chunks) * kChunkWords, 0);
   col.chunk_nruns.assign(t.n_chunks, 0);
   if (nptr[n] > 0xFFFFFFFFULL) fail(RDB_ENOMEM, "too many missing runs");
   // Every array below is chunk-local except for running offsets, which come from per-chunk counts:
   // the chunks are filled in parallel.  Null rows per chunk, concatenated in chunk order after.
   std::vector<std::vector<uint32_t>> chunk_null_rows(t.n_chunks);
   std::vector<uint64_t> chunk_muts(t.n_chunks + 1, 0);
   parallelFor(t.n_chunks, [&](size_t c) {
      const uint64_t off = t.chunk_offset[c];
      const uint32_t size = t.chunk_sizes[c];
      uint64_t muts = 0;
      std::fill_n(cov.data() + (c << 16), kChunkRows, 0u);
      for (uint32_t r = 0; r < size; ++r) {
         const uint64_t i = off + r;
         const size_t g = (c << 16) | r;
         if (is_null[i]) {
            null_bits[g >> 6] |= 1ULL << (g & 63);
            chunk_null_rows[c].push_back(static_cast<uint32_t>(g));
         } else {
            if (cov_start[i] < cov_end[i]) cov[g] = cov_start[i] | (cov_end[i] << 16);  // empty range: [0, 0)
            muts += mut_ptr[i + 1] - mut_ptr[i];
         }
      }
      col.chunk_nruns[c] = nptr[off + size] - nptr[off];
      chunk_muts[c + 1] = muts;
      // pptr over padded rows: the runs of every row before, null rows included
      for (uint32_t r = 0; r < kChunkRows; ++r) pptr[(c << 16) | r] = static_cast<uint32_t>(nptr[off + std::min(r, size)] - nptr[0]);
   });
   pptr[padded] = static_cast<uint32_t>(nptr[n] - nptr[0]);
   std::vector<uint32_t> null_rows;
   for (auto& v : chunk_null_rows) null_rows.insert(null_rows.end(), v.begin(), v.end());
   // missing-run index by (chunk, start bucket or "long", slice), rows in order: each chunk owns its
   // lists, so counting and filling run per chunk
   const uint32_t n_rb = (L + kRunBucket - 1) / kRunBucket;
   const size_t n_lists = static_cast<size_t>(t.n_chunks) * (n_rb + 1) * kSlicesPerChunk;
   std::vector<uint32_t> ri_ptr(n_lists + 1, 0);
   RawVector<uint64_t> ri;
   {
      auto listOf = [&](uint32_t c, uint32_t r, uint32_t run) {
         const uint32_t a = run & 0xFFFFu, b = run >> 16;
         const uint32_t k = b - a > kShortRun ? n_rb : a / kRunBucket;
         return (static_cast<size_t>(c) * (n_rb + 1) + k) * kSlicesPerChunk + (r >> 12);
      };
      parallelFor(t.n_chunks, [&](size_t c) {
         for (uint32_t r = 0; r < t.chunk_sizes[c]; ++r) {
            const uint64_t i = t.chunk_offset[c] + r;
            for (uint64_t k = nptr[i]; k < nptr[i + 1]; ++k) ri_ptr[listOf(static_cast<uint32_t>(c), r, runs_all[k]) + 1] += 1;
         }
      });
      for (size_t l = 0; l < n_lists; ++l) ri_ptr[l + 1] += ri_ptr[l];
      if (ri_ptr[n_lists] != nptr[n]) fail(RDB_EDEVICE, "internal: missing-run index size");
      ri.resize(ri_ptr[n_lists]);
      std::vector<uint32_t> fill(ri_ptr.begin(), ri_ptr.end() - 1);
      parallelFor(t.n_chunks, [&](size_t c) {
         for (uint32_t r = 0; r < t.chunk_sizes[c]; ++r) {
            const uint64_t i = t.chunk_offset[c] + r;
            for (uint64_t k = nptr[i]; k < nptr[i + 1]; ++k) {
               const uint32_t run = runs_all[k];
               ri[fill[listOf(static_cast<uint32_t>(c), r, run)]++] = r | (static_cast<uint64_t>(run & 0xFFFFu) << 16) |
                                                                     (static_cast<uint64_t>(run >> 16) << 32);
            }
         }
      });
   }
   col.has_nulls = !null_rows.empty();
   col.null_bitmap_id = t.addStoreBitmap(null_rows);

   // row-major mutations over padded rows (null rows have none)
   for (uint32_t c = 0; c < t.n_chunks; ++c) chunk_muts[c + 1] += chunk_muts[c];
   if (chunk_muts[t.n_chunks] > 0xFFFFFFFFULL) fail(RDB_ENOMEM, "too many mutations for 32-bit offsets");
   RawVector<uint32_t> rmptr(padded + 1), rmuts(chunk_muts[t.n_chunks]);
   parallelFor(t.n_chunks, [&](size_t c) {
      uint32_t acc = static_cast<uint32_t>(chunk_muts[c]);
      for (uint32_t r = 0; r < kChunkRows; ++r) {
         const size_t g = (c << 16) | r;
         rmptr[g] = acc;
         if (r >= t.chunk_sizes[c]) continue;
         const uint64_t i = t.chunk_offset[c] + r;
         if (is_null[i]) continue;
         for (uint64_t k = mut_ptr[i]; k < mut_ptr[i + 1]; ++k) rmuts[acc++] = mut_pos[k] | (static_cast<uint32_t>(mut_sym[k]) << 16);
      }
   });
   rmptr[padded] = static_cast<uint32_t>(chunk_muts[t.n_chunks]);

   phase(4);
   // --- checks ------------------------------------------------------------------------------------
   validateDescriptors(desc.data(), desc.size(), arr_all.size(), run_all.size(), bit_all.size() / kChunkWords, kMaxBlockSlots);
   for (uint32_t b = 0; b < n_blocks; ++b) {
      const uint32_t ns = col.block_slot_start[b + 1] - col.block_slot_start[b];
      if (ns > kMaxBlockSlots) fail(RDB_EDEVICE, "internal: block with too many slots");
      for (uint32_t c = 0; c < t.n_chunks; ++c) {
         const size_t tile = static_cast<size_t>(b) * t.n_chunks + c;
         if (light_ptr[tile] > light_ptr[tile + 1] || heavy_ptr[tile] > heavy_ptr[tile + 1] || light_ptr[tile] % 4 != 0) {
            fail(RDB_EDEVICE, "internal: tile pointers out of order");
         }
         for (uint32_t k = heavy_ptr[tile]; k < heavy_ptr[tile + 1]; ++k) {
            if (descSlot(desc[k].w0) >= ns) fail(RDB_EDEVICE, "internal: descriptor slot outside its block");
         }
         if (group_ptr[tile] > group_ptr[tile + 1]) fail(RDB_EDEVICE, "internal: group pointers out of order");
         for (uint32_t k = group_ptr[tile]; k < group_ptr[tile + 1]; ++k) {
            if ((group_hdr[k] & 0xFFFu) >= ns) fail(RDB_EDEVICE, "internal: group slot outside its block");
            if (k > group_ptr[tile] && (group_hdr[k] & 0xFFFu) < (group_hdr[k - 1] & 0xFFFu)) {
               fail(RDB_EDEVICE, "internal: groups not sorted");
            }
         }
         for (uint32_t k = light_ptr[tile]; k < light_ptr[tile + 1]; ++k) {
            const uint32_t slot = light[k] >> 16;
            if (slot >= ns && slot != kPadSlot) fail(RDB_EDEVICE, "internal: light entry slot outside its block");
            if (k > light_ptr[tile] && light[k] < light[k - 1]) fail(RDB_EDEVICE, "internal: light entries not sorted");
         }
      }
   }
   for (size_t k = 0; k < runs_all.size(); ++k) {
      if ((runs_all[k] & 0xFFFFu) >= (runs_all[k] >> 16) || (runs_all[k] >> 16) > L) fail(RDB_EDEVICE, "internal: bad missing run");
   }

   phase(5);
   // --- upload ------------------------------------------------------------------------------------
   auto up = [&](const void* p, size_t bytes) -> const void* {
      col.buffers.emplace_back(p, bytes);
      return col.buffers.back().as<void>();
   };
   SeqColumnView& v = col.view;
   v.genome_length = L;
   v.n_symbols = nsym;
   v.n_blocks = n_blocks;
   v.n_slots = n_slots;
   v.n_valid = static_cast<uint32_t>(A.valid.size());
   v.valid_mask = 0;
   for (uint8_t s : A.valid) v.valid_mask |= 1u << s;
   v.vert.desc = static_cast<const Desc*>(up(desc.data(), desc.size() * sizeof(Desc)));
   v.vert.arr = static_cast<const uint16_t*>(up(arr_all.data(), arr_all.size() * 2));
   v.vert.runs = static_cast<const uint32_t*>(up(run_all.data(), run_all.size() * 4));
   v.vert.bits = static_cast<const uint64_t*>(up(bit_all.data(), bit_all.size() * 8));
   v.block_pos_start = static_cast<const uint32_t*>(up(col.block_pos_start.data(), col.block_pos_start.size() * 4));
   v.block_slot_start = static_cast<const uint32_t*>(up(col.block_slot_start.data(), col.block_slot_start.size() * 4));
   v.light_ptr = static_cast<const uint32_t*>(up(light_ptr.data(), light_ptr.size() * 4));
   v.light = static_cast<const uint32_t*>(up(light.data(), light.size() * 4));
   v.heavy_ptr = static_cast<const uint32_t*>(up(heavy_ptr.data(), heavy_ptr.size() * 4));
   v.group_ptr = static_cast<const uint32_t*>(up(group_ptr.data(), group_ptr.size() * 4));
   v.groups = static_cast<const uint16_t*>(up(group_rows.data(), group_rows.size() * 2));
   v.group_hdr = static_cast<const uint16_t*>(up(group_hdr.data(), group_hdr.size() * 2));
   v.slot_info = static_cast<const uint32_t*>(up(col.slot_info.data(), col.slot_info.size() * 4));
   v.pos_slot_ptr = static_cast<const uint32_t*>(up(col.pos_slot_ptr.data(), col.pos_slot_ptr.size() * 4));
   v.local_ref = static_cast<const uint8_t*>(up(col.local_ref.data(), L));
   v.reference = static_cast<const uint8_t*>(up(col.reference.data(), L));
   v.cov = static_cast<const uint32_t*>(up(cov.data(), cov.size() * 4));
   {
      std::vector<uint32_t> bounds(static_cast<size_t>(t.n_chunks) * kSlicesPerChunk, 0xFFFFu);
      std::vector<uint64_t> never(static_cast<size_t>(t.n_chunks) * kChunkWords, 0);
      for (uint32_t c = 0; c < t.n_chunks; ++c) {
         for (uint32_t sl = 0; sl < kSlicesPerChunk; ++sl) {
            uint32_t lo = 0, hi = 0xFFFFu;
            bool any = false;
            for (uint32_t r = sl * 4096u; r < std::min<uint32_t>((sl + 1) * 4096u, t.chunk_sizes[c]); ++r) {
               const size_t g = (static_cast<size_t>(c) << 16) | r;
               const uint32_t x = cov[g];
               if (x == 0) {
                  never[g >> 6] |= 1ULL << (g & 63);  // null row or empty range
               } else {
                  lo = std::max(lo, x & 0xFFFFu);
                  hi = std::min(hi, x >> 16);
                  any = true;
               }
            }
            bounds[static_cast<size_t>(c) * kSlicesPerChunk + sl] = any ? (lo | (hi << 16)) : 0xFFFFu;
         }
      }
      v.cov_bounds = static_cast<const uint32_t*>(up(bounds.data(), bounds.size() * 4));
      v.cov_never = static_cast<const uint64_t*>(up(never.data(), never.size() * 8));
   }
   v.nrun_ptr = static_cast<const uint32_t*>(up(pptr.data(), pptr.size() * 4));
   v.nruns = static_cast<const uint32_t*>(up(runs_all.data(), runs_all.size() * 4));
   v.n_run_buckets = n_rb;
   v.run_index_ptr = static_cast<const uint32_t*>(up(ri_ptr.data(), ri_ptr.size() * 4));
   v.run_index = static_cast<const uint64_t*>(up(ri.data(), ri.size() * 8));
   v.null_bits = static_cast<const uint64_t*>(up(null_bits.data(), null_bits.size() * 8));
   v.pos_ranges = static_cast<const uint32_t*>(up(pos_ranges.data(), pos_ranges.size() * 4));
   v.row_mut_ptr = static_cast<const uint32_t*>(up(rmptr.data(), rmptr.size() * 4));
   v.row_muts = static_cast<const uint32_t*>(up(rmuts.data(), rmuts.size() * 4));
   col.row_major_bytes = (rmptr.size() + rmuts.size()) * 4;
   col.row_mut_ptr = std::move(rmptr);
   col.row_nrun_ptr = std::move(pptr);
   phase(6);
}

// ---- metadata columns --------------------------------------------------------------------------
void buildIntColumn(Table& t, IntColumn& col, const int32_t* values, const uint8_t* is_null) {
   const size_t padded = static_cast<size_t>(t.n_chunks) << 16;
   std::vector<int32_t> vals(padded, 0);
   std::vector<uint64_t> null_bits(static_cast<size_t>(t.n_chunks) * kChunkWords, 0);
   std::vector<uint32_t> null_rows;
   col.values.assign(values, values + t.num_rows);
   bool have_last = false;
   int32_t last = 0;
   for (uint32_t c = 0; c < t.n_chunks; ++c) {
      for (uint32_t r = 0; r < t.chunk_sizes[c]; ++r) {
         const uint64_t i = t.chunk_offset[c] + r;
         const size_t g = (static_cast<size_t>(c) << 16) | r;
         if (is_null[i]) {
            null_bits[g >> 6] |= 1ULL << (g & 63);
            null_rows.push_back(static_cast<uint32_t>(g));
            col.values[i] = 0;
            col.is_sorted = false;  // date32_column.cpp:27-32
         } else {
            vals[g] = values[i];
            if (have_last && values[i] < last) col.is_sorted = false;
            last = values[i];
            have_last = true;
         }
      }
   }
   col.null_bitmap_id = t.addStoreBitmap(null_rows);
   // sorted non-null values: the planner's exact counts of value predicates
   col.null_count = null_rows.size();
   col.sorted_values.clear();
   col.sorted_values.reserve(t.num_rows - null_rows.size());
   int64_t vmin = INT64_MAX, vmax = INT64_MIN;
   for (uint64_t i = 0; i < t.num_rows; ++i) {
      if (is_null[i]) continue;
      vmin = std::min<int64_t>(vmin, values[i]);
      vmax = std::max<int64_t>(vmax, values[i]);
   }
   if (vmin <= vmax && static_cast<uint64_t>(vmax - vmin) < (uint64_t{1} << 24)) {
      // counting sort over a narrow range (dates span a few thousand days)
      std::vector<uint64_t> count(static_cast<size_t>(vmax - vmin) + 1, 0);
      for (uint64_t i = 0; i < t.num_rows; ++i) {
         if (!is_null[i]) ++count[static_cast<size_t>(values[i] - vmin)];
      }
      for (size_t k = 0; k < count.size(); ++k) col.sorted_values.insert(col.sorted_values.end(), count[k], static_cast<int32_t>(vmin + static_cast<int64_t>(k)));
   } else {
      for (uint64_t i = 0; i < t.num_rows; ++i) {
         if (!is_null[i]) col.sorted_values.push_back(values[i]);
      }
      std::sort(col.sorted_values.begin(), col.sorted_values.end());
   }
   col.buffers.emplace_back(vals.data(), vals.size() * 4);
   col.view.values = col.buffers.back().as<int32_t>();
   col.buffers.emplace_back(null_bits.data(), null_bits.size() * 8);
   col.view.null_bits = col.buffers.back().as<uint64_t>();
   // bit-sliced copy for comparisons: a date range then reads ceil(log2(span)) bits per row instead
   // of 32
   col.view.planes = nullptr;
   col.view.plane_min = 0;
   col.view.n_planes = 0;
   if (!col.sorted_values.empty()) {
      const int64_t lo = col.sorted_values.front(), hi = col.sorted_values.back();
      const uint64_t span = static_cast<uint64_t>(hi - lo);
      if (span < (uint64_t{1} << kMaxIntPlanes)) {
         const uint32_t np = std::max<uint32_t>(1, static_cast<uint32_t>(std::bit_width(span)));
         const size_t plane_words = static_cast<size_t>(t.n_chunks) * kChunkWords;
         std::vector<uint64_t> planes(np * plane_words, 0);
         // each chunk's plane words are its own
         parallelFor(t.n_chunks, [&](size_t c) {
            for (uint32_t r = 0; r < t.chunk_sizes[c]; ++r) {
               const uint64_t i = t.chunk_offset[c] + r;
               if (is_null[i]) continue;
               const size_t g = (static_cast<size_t>(c) << 16) | r;
               const uint64_t x = static_cast<uint64_t>(static_cast<int64_t>(values[i]) - lo);
               for (uint32_t p = 0; p < np; ++p) {
                  if ((x >> p) & 1u) planes[p * plane_words + (g >> 6)] |= 1ULL << (g & 63);
               }
            }
         });
         col.buffers.emplace_back(planes.data(), planes.size() * 8);
         col.view.planes = col.buffers.back().as<uint64_t>();
         col.view.plane_min = static_cast<int32_t>(lo);
         col.view.n_planes = np;
      }
   }
}

void buildDictColumn(Table& t, DictColumn& col, const uint32_t* ids, const uint8_t* is_null,
                     const int32_t* lineage_parent, const int32_t* lineage_alias) {
   const uint32_t nv = static_cast<uint32_t>(col.dictionary.size());
   std::vector<std::vector<uint32_t>> rows(nv);
   std::vector<uint32_t> null_rows;
   const size_t padded = static_cast<size_t>(t.n_chunks) << 16;
   std::vector<uint32_t> labels(padded, kInvalidId);
   for (uint32_t c = 0; c < t.n_chunks; ++c) {
      for (uint32_t r = 0; r < t.chunk_sizes[c]; ++r) {
         const uint64_t i = t.chunk_offset[c] + r;
         const uint32_t g = (c << 16) | r;
         if (is_null[i]) {
            null_rows.push_back(g);
            labels[g] = nv;
         } else {
            if (ids[i] >= nv) fail(RDB_EINVAL, "dictionary id out of range");
            rows[ids[i]].push_back(g);
            labels[g] = ids[i];
         }
      }
   }
   // Store bitmaps are split into chunks in parallel, then registered in value order (the ids a
   // serial build gives).
   auto addAll = [&](const std::vector<std::vector<uint32_t>>& lists, const std::vector<uint8_t>& use,
                     std::vector<uint32_t>& ids) {
      std::vector<std::vector<std::vector<uint16_t>>> split(lists.size());
      parallelFor(lists.size(), [&](size_t v) {
         if (use[v]) split[v] = t.storeChunks(lists[v]);
      });
      for (size_t v = 0; v < lists.size(); ++v) {
         if (use[v]) ids[v] = t.addStoreChunks(std::move(split[v]));
      }
   };
   col.value_bitmap.assign(nv, kInvalidId);
   {
      std::vector<uint8_t> use(nv);
      for (uint32_t v = 0; v < nv; ++v) use[v] = !rows[v].empty();
      addAll(rows, use, col.value_bitmap);
   }
   col.has_nulls = !null_rows.empty();
   col.null_bitmap_id = t.addStoreBitmap(null_rows);
   if (lineage_parent) {
      // lineage_index.cpp:17-45: rows are indexed under their value's resolved alias, a value's
      // sublineage rows are its own and those of all its descendants, and a query value resolves
      // its alias before the lookup
      col.lineage = true;
      std::vector<uint32_t> resolved(nv);
      for (uint32_t v = 0; v < nv; ++v) {
         const int32_t a = lineage_alias ? lineage_alias[v] : -1;
         if (a >= 0 && static_cast<uint32_t>(a) >= nv) fail(RDB_EINVAL, "lineage alias out of range");
         resolved[v] = a >= 0 ? static_cast<uint32_t>(a) : v;
      }
      std::vector<std::vector<uint32_t>> own(nv);
      for (uint32_t v = 0; v < nv; ++v) own[resolved[v]].insert(own[resolved[v]].end(), rows[v].begin(), rows[v].end());
      col.lineage_bitmap.assign(nv, kInvalidId);
      {
         std::vector<uint8_t> use(nv);
         for (uint32_t v = 0; v < nv; ++v) use[v] = resolved[v] == v && !own[v].empty();
         parallelFor(nv, [&](size_t v) {
            if (use[v]) std::sort(own[v].begin(), own[v].end());
         });
         addAll(own, use, col.lineage_bitmap);
      }
      col.sublineage_bitmap.assign(nv, kInvalidId);
      std::vector<std::vector<uint32_t>> children(nv);
      for (uint32_t v = 0; v < nv; ++v) {
         const int32_t p = lineage_parent[v];
         if (resolved[v] == v && p >= 0 && static_cast<uint32_t>(p) < nv && static_cast<uint32_t>(p) != v) {
            children[p].push_back(v);
         }
      }
      // the lineages under each lineage (a walk over lineages, each visited once per start even
      // if the parent links form a cycle), then their rows gathered and sorted in parallel
      std::vector<std::vector<uint32_t>> under(nv);
      std::vector<uint32_t> seen(nv, kInvalidId);
      for (uint32_t v = 0; v < nv; ++v) {
         if (resolved[v] != v) continue;
         std::vector<uint32_t> stack{v};
         while (!stack.empty()) {
            const uint32_t w = stack.back();
            stack.pop_back();
            if (seen[w] == v) continue;
            seen[w] = v;
            under[v].push_back(w);
            for (uint32_t ch : children[w]) stack.push_back(ch);
         }
      }
      std::vector<std::vector<std::vector<uint16_t>>> split(nv);
      std::vector<uint8_t> use(nv, 0);
      parallelFor(nv, [&](size_t v) {
         uint64_t total = 0;
         for (uint32_t w : under[v]) total += own[w].size();
         if (total == 0) return;
         use[v] = 1;
         if (total < (uint64_t{1} << 16)) {
            std::vector<uint32_t> acc;
            acc.reserve(total);
            for (uint32_t w : under[v]) acc.insert(acc.end(), own[w].begin(), own[w].end());
            std::sort(acc.begin(), acc.end());
            split[v] = t.storeChunks(acc);
            return;
         }
         // a large subtree (a root's can hold every row): its rows set bits of per-chunk bitsets,
         // read back in order, instead of sorting them
         std::vector<uint64_t> bits(static_cast<size_t>(t.n_chunks) * kChunkWords, 0);
         for (uint32_t w : under[v]) {
            for (uint32_t g : own[w]) bits[g >> 6] |= 1ULL << (g & 63);  // own rows lie in the table
         }
         auto& out = split[v];
         out.resize(t.n_chunks);
         for (uint32_t c = 0; c < t.n_chunks; ++c) {
            const uint64_t* cw = bits.data() + static_cast<size_t>(c) * kChunkWords;
            for (uint32_t k = 0; k < kChunkWords; ++k) {
               for (uint64_t x = cw[k]; x; x &= x - 1) out[c].push_back(static_cast<uint16_t>(64 * k + std::countr_zero(x)));
            }
         }
      });
      for (uint32_t v = 0; v < nv; ++v) {
         if (use[v]) col.sublineage_bitmap[v] = t.addStoreChunks(std::move(split[v]));
      }
      for (uint32_t v = 0; v < nv; ++v) {
         if (resolved[v] == v) continue;
         col.lineage_bitmap[v] = col.lineage_bitmap[resolved[v]];
         col.sublineage_bitmap[v] = col.sublineage_bitmap[resolved[v]];
      }
   }
   std::vector<uint32_t> order(nv);
   std::iota(order.begin(), order.end(), 0);
   std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return col.dictionary[a] < col.dictionary[b]; });
   col.rank.assign(nv, 0);
   for (uint32_t k = 0; k < nv; ++k) col.rank[order[k]] = k;
   col.labels = DeviceBuffer(labels.data(), labels.size() * 4);
}

// ---- finalize: bitmap store + views + scratch ----------------------------------------------------
void finalizeTable(Table& t) {
   // bitmap store
   ArenaBuilder& sb = t.store_builder;
   sb = ArenaBuilder();
   for (const auto& bm : t.store_rows) {
      for (uint32_t c = 0; c < t.n_chunks; ++c) {
         const auto& r = bm[c];
         sb.add(r.data(), static_cast<uint32_t>(r.size()), 0);
      }
   }
   if (sb.desc.size() != t.store_rows.size() * t.n_chunks) fail(RDB_EDEVICE, "internal: bitmap store size mismatch");
   t.store_card.assign(t.store_rows.size(), 0);
   for (size_t b = 0; b < t.store_rows.size(); ++b) {
      for (const auto& r : t.store_rows[b]) t.store_card[b] += r.size();
   }

   validateDescriptors(sb.desc.data(), sb.desc.size(), sb.arr.size(), sb.runs.size(), sb.bits.size() / kChunkWords, 1);
   t.store_buffers.clear();
   auto up = [&](const void* p, size_t bytes) -> void* {
      t.store_buffers.emplace_back(p, bytes);
      return t.store_buffers.back().as<void>();
   };
   t.view.store.desc = static_cast<const Desc*>(up(sb.desc.data(), sb.desc.size() * sizeof(Desc)));
   t.view.store.arr = static_cast<const uint16_t*>(up(sb.arr.data(), sb.arr.size() * 2));
   t.view.store.runs = static_cast<const uint32_t*>(up(sb.runs.data(), sb.runs.size() * 4));
   t.view.store.bits = static_cast<const uint64_t*>(up(sb.bits.data(), sb.bits.size() * 8));
   sb = ArenaBuilder();
   std::vector<SeqColumnView> sv;
   for (const auto& c : t.seq) sv.push_back(c->view);
   std::vector<IntColumnView> iv;
   for (const auto& c : t.ints) iv.push_back(c->view);
   t.chunk_sizes_dev = DeviceBuffer(t.chunk_sizes.data(), t.chunk_sizes.size() * 4);
   t.seq_views_dev = DeviceBuffer(sv.data(), sv.size() * sizeof(SeqColumnView));
   t.int_views_dev = DeviceBuffer(iv.data(), iv.size() * sizeof(IntColumnView));
   t.view.n_chunks = t.n_chunks;
   t.view.chunk_sizes = t.chunk_sizes_dev.as<uint32_t>();
   t.view.seq = t.seq_views_dev.as<SeqColumnView>();
   t.view.ints = t.int_views_dev.as<IntColumnView>();
   uint32_t max_slots = 1, max_L = 1;
   for (const auto& c : t.seq) {
      max_slots = std::max(max_slots, c->view.n_slots);
      max_L = std::max(max_L, c->genome_length);
   }
   t.slot_counts = DeviceBuffer(static_cast<size_t>(max_slots) * 4 * kMaxGroups);
   t.heavy_sums = DeviceBuffer(static_cast<size_t>(max_slots) * 4);
   hipCheck(hipMemset(t.heavy_sums.as<void>(), 0, t.heavy_sums.bytes()), "clear heavy sums");
   t.totals = DeviceBuffer(static_cast<size_t>((max_L + kMutPosBlock - 1) / kMutPosBlock) * 3 * sizeof(int64_t));
   // at least the 2048 rows the Mutations path copies back in its first transfer
   t.mut_rows = DeviceBuffer(16 + std::max<size_t>(static_cast<size_t>(max_L) * 32, 2048) * sizeof(MutationRowDev));
   t.mut_n = DeviceBuffer(64);
   for (auto& c : t.seq) {
      // sparse-path accumulators start (and are left) zero
      const size_t L = c->genome_length;
      c->sparse_counts = DeviceBuffer(L * c->abc->count * 4);
      c->sparse_bnd = DeviceBuffer(L * 3 * 4);
      hipCheck(hipMemset(c->sparse_counts.as<void>(), 0, c->sparse_counts.bytes()), "clear sparse counts");
      hipCheck(hipMemset(c->sparse_bnd.as<void>(), 0, c->sparse_bnd.bytes()), "clear sparse boundaries");
   }
   t.finalized = true;
}

// ---- device bitmaps ----------------------------------------------------------------------------
DeviceBitmap::DeviceBitmap(Table* t) : table(t) {
   const size_t n_words = static_cast<size_t>(t->n_chunks) * kChunkWords;
   bytes = n_words * 8 + static_cast<size_t>(t->n_chunks) * 4 + 64;
   auto* base = static_cast<uint8_t*>(t->ctx->pool.get(bytes));
   words = reinterpret_cast<uint64_t*>(base);
   card_dev = reinterpret_cast<unsigned long long*>(base + n_words * 8);
   chunk_cards = reinterpret_cast<uint32_t*>(base + n_words * 8 + 64);
}

DeviceBitmap::~DeviceBitmap() {
   if (words && table) table->ctx->pool.put(words, bytes);
}

DeviceBitmap::DeviceBitmap(DeviceBitmap&& o) noexcept { *this = std::move(o); }

DeviceBitmap& DeviceBitmap::operator=(DeviceBitmap&& o) noexcept {
   if (this != &o) {
      if (words && table) table->ctx->pool.put(words, bytes);
      table = o.table;
      words = o.words;
      chunk_cards = o.chunk_cards;
      card_dev = o.card_dev;
      bytes = o.bytes;
      card = o.card;
      card_known = o.card_known;
      o.words = nullptr;
      o.table = nullptr;
   }
   return *this;
}

uint64_t DeviceBitmap::cardinality() {
   if (!card_known) {
      unsigned long long v = 0;
      hipCheck(hipMemcpyAsync(&v, card_dev, sizeof(v), hipMemcpyDeviceToHost, table->ctx->stream), "copy cardinality");
      hipCheck(hipStreamSynchronize(table->ctx->stream), "synchronize");
      card = v;
      card_known = true;
   }
   return card;
}

std::vector<uint32_t> DeviceBitmap::rows() {
   const size_t n_words = static_cast<size_t>(table->n_chunks) * kChunkWords;
   std::vector<uint64_t> host(n_words);
   hipCheck(hipMemcpyAsync(host.data(), words, n_words * 8, hipMemcpyDeviceToHost, table->ctx->stream), "copy bitmap");
   hipCheck(hipStreamSynchronize(table->ctx->stream), "synchronize");
   std::vector<uint32_t> out;
   for (size_t w = 0; w < n_words; ++w) {
      uint64_t x = host[w];
      while (x) {
         out.push_back(static_cast<uint32_t>(w * 64 + static_cast<size_t>(std::countr_zero(x))));
         x &= x - 1;
      }
   }
   return out;
}

}  // namespace rdb

namespace rdb {

// The reference's finalized SequenceColumn (sequence_column.h:86-97) back in diff form: a stored
// container (position p, chunk, symbol s) gives its rows the mutation (p, s) unless s is the
// column's reference symbol (a position whose local reference was adapted stores the reference
// symbol's rows, vertical_sequence_index.cpp:118-164); at such a position the rows left implicit
// (covered at p, p not missing, in no stored container) hold the local reference symbol, which is
// a mutation against the reference (symbol_in_set.cpp:185-212 reads them the same way).  Missing
// positions inside a row's covered range (horizontal_coverage_index.h:25-27) become runs.
// Sorts keys row << 24 | position << 8 | symbol (row < 2^16, position < 2^16) in ascending order
// with three counting passes (symbol, position, row): linear in the keys, where a comparison sort
// of a chunk's millions of entries dominated the finalized upload.
static void sortRowKeys(std::vector<uint64_t>& keys) {
   if (keys.size() < 2) return;
   std::vector<uint64_t> tmp(keys.size());
   std::vector<uint32_t> count(1u << 16);
   auto pass = [&](const std::vector<uint64_t>& src, std::vector<uint64_t>& dst, uint32_t shift, uint32_t bits) {
      const uint64_t mask = (1ull << bits) - 1;
      std::fill(count.begin(), count.begin() + (1u << bits), 0u);
      for (uint64_t k : src) ++count[(k >> shift) & mask];
      uint32_t sum = 0;
      for (uint32_t b = 0; b < (1u << bits); ++b) {
         const uint32_t c = count[b];
         count[b] = sum;
         sum += c;
      }
      for (uint64_t k : src) dst[count[(k >> shift) & mask]++] = k;
   };
   pass(keys, tmp, 0, 8);
   pass(tmp, keys, 8, 16);
   pass(keys, tmp, 24, 16);
   keys.swap(tmp);
}

void buildSequenceColumnFinalized(Table& t, SeqColumn& col, const FinalizedSequence& f) {
   const uint32_t L = col.genome_length;
   const uint32_t nsym = col.abc->count;
   const uint64_t n = t.num_rows;
   if (L == 0 || L > 65535) fail(RDB_EINVAL, "genome length must be in [1, 65535]");
   const bool given_runs = f.run_ptr != nullptr;
   if (!f.local_reference || !f.is_null || !f.cov_start || !f.cov_end || (!f.missing_ptr && !given_runs)) {
      fail(RDB_EINVAL, "null argument");
   }
   if (f.n_containers && (!f.key_position || !f.key_chunk || !f.key_symbol || !f.typecode || !f.cardinality ||
                          !f.payload_offset || !f.payload)) {
      fail(RDB_EINVAL, "null container argument");
   }
   for (uint32_t p = 0; p < L; ++p) {
      if (f.local_reference[p] >= nsym || col.reference[p] >= nsym) fail(RDB_EINVAL, "invalid reference symbol");
   }
   std::vector<uint64_t> nrun_ptr(n + 1, 0);
   std::vector<uint32_t> nrun_start, nrun_end;
   if (given_runs) {
      if (!f.run_start || !f.run_end) fail(RDB_EINVAL, "null argument");
      if (f.run_ptr[0] != 0) fail(RDB_EINVAL, "missing run offsets must start at 0");
      for (uint64_t i = 0; i < n; ++i) {
         if (f.run_ptr[i + 1] < f.run_ptr[i]) fail(RDB_EINVAL, "missing run offsets must be non-decreasing");
         for (uint64_t k = f.run_ptr[i]; k < f.run_ptr[i + 1]; ++k) {
            if (f.run_start[k] >= f.run_end[k] || f.run_end[k] > L) fail(RDB_EINVAL, "empty missing run or one past the genome");
            if (k > f.run_ptr[i] && f.run_start[k] < f.run_end[k - 1]) fail(RDB_EINVAL, "missing runs of a row must ascend");
         }
         nrun_ptr[i + 1] = f.run_ptr[i + 1];
      }
      nrun_start.assign(f.run_start, f.run_start + f.run_ptr[n]);
      nrun_end.assign(f.run_end, f.run_end + f.run_ptr[n]);
   } else {
      if (f.missing_ptr[0] != 0) fail(RDB_EINVAL, "missing_ptr must start at 0");
      for (uint64_t i = 0; i < n; ++i) {
         if (f.missing_ptr[i + 1] < f.missing_ptr[i]) fail(RDB_EINVAL, "missing_ptr must be non-decreasing");
      }
      // missing positions -> runs [start, end)
      for (uint64_t i = 0; i < n; ++i) {
         const uint32_t* a = f.missing_pos + f.missing_ptr[i];
         const uint32_t* e = f.missing_pos + f.missing_ptr[i + 1];
         for (const uint32_t* q = a; q < e;) {
            const uint32_t s = *q;
            uint32_t last = s;
            ++q;
            while (q < e && *q == last + 1) last = *q++;
            if (q < e && *q <= last) fail(RDB_EINVAL, "missing positions of a row must ascend");
            nrun_start.push_back(s);
            nrun_end.push_back(last + 1);
         }
         nrun_ptr[i + 1] = nrun_start.size();
      }
   }
   // p inside one of row i's missing runs
   auto isMissing = [&](uint64_t i, uint32_t p) {
      const uint32_t* a = nrun_start.data() + nrun_ptr[i];
      const uint32_t* e = nrun_start.data() + nrun_ptr[i + 1];
      const uint32_t* q = std::upper_bound(a, e, p);
      return q != a && p < nrun_end[static_cast<size_t>(q - nrun_start.data()) - 1];
   };
   // containers by chunk
   std::vector<std::vector<uint64_t>> by_chunk(t.n_chunks);
   for (uint64_t k = 0; k < f.n_containers; ++k) {
      const uint32_t c = f.key_chunk[k];
      if (c >= t.n_chunks) fail(RDB_EINVAL, "container chunk outside the row layout");
      const uint32_t p = f.key_position[k], s = f.key_symbol[k];
      if (p >= L) fail(RDB_EINVAL, "container position outside the genome");
      if (s >= nsym || s == col.abc->missing) fail(RDB_EINVAL, "invalid container symbol");
      if (s == f.local_reference[p]) fail(RDB_EINVAL, "a stored container holds the local reference symbol");
      by_chunk[c].push_back(k);
   }
   std::vector<uint32_t> adapted;
   for (uint32_t p = 0; p < L; ++p) {
      if (f.local_reference[p] != col.reference[p]) adapted.push_back(p);
   }
   std::vector<std::vector<uint64_t>> muts(t.n_chunks);  // row << 24 | position << 8 | symbol
   std::vector<std::string> errors(t.n_chunks);
   parallelFor(t.n_chunks, [&](size_t c) {
      const uint32_t size = t.chunk_sizes[c];
      const uint64_t off = t.chunk_offset[c];
      auto& out = muts[c];
      // rows stored at each adapted position of this chunk
      std::vector<uint64_t> stored(adapted.size() * kChunkWords, 0);
      std::vector<uint32_t> rows;
      auto decode = [&](uint64_t k) -> bool {
         rows.clear();
         const uint8_t type = f.typecode[k];
         const uint64_t o = f.payload_offset[k];
         const uint32_t card = f.cardinality[k];
         // [o, o + n) inside the payload, without wrapping for offsets near 2^64
         auto fits = [&](uint64_t n) { return o <= f.payload_bytes && n <= f.payload_bytes - o; };
         if (!fits(0)) return false;
         const uint8_t* b = f.payload + o;
         if (type == 1) {  // bitset: 1024 little-endian u64 words
            if (!fits(8192)) return false;
            for (uint32_t w = 0; w < kChunkWords; ++w) {
               uint64_t x;
               std::memcpy(&x, b + 8 * w, 8);
               while (x) {
                  rows.push_back(w * 64 + static_cast<uint32_t>(std::countr_zero(x)));
                  x &= x - 1;
               }
            }
         } else if (type == 2) {  // array: card sorted u16
            if (!fits(2ULL * card)) return false;
            for (uint32_t j = 0; j < card; ++j) {
               uint16_t v;
               std::memcpy(&v, b + 2 * j, 2);
               if (j && v <= rows.back()) return false;
               rows.push_back(v);
            }
         } else if (type == 3) {  // run: u16 run count, then (start, length - 1) pairs
            if (!fits(2)) return false;
            uint16_t nr;
            std::memcpy(&nr, b, 2);
            if (!fits(2 + 4ULL * nr)) return false;
            for (uint32_t j = 0; j < nr; ++j) {
               uint16_t v, len;
               std::memcpy(&v, b + 2 + 4 * j, 2);
               std::memcpy(&len, b + 4 + 4 * j, 2);
               if (static_cast<uint32_t>(v) + len > 0xFFFFu) return false;
               if (!rows.empty() && v <= rows.back()) return false;
               for (uint32_t r = v; r <= static_cast<uint32_t>(v) + len; ++r) rows.push_back(r);
            }
         } else {
            return false;
         }
         return rows.size() == card;
      };
      for (uint64_t k : by_chunk[c]) {
         if (!decode(k)) {
            errors[c] = "malformed container payload";
            return;
         }
         const uint32_t p = f.key_position[k], s = f.key_symbol[k];
         const auto ai = std::lower_bound(adapted.begin(), adapted.end(), p);
         const bool is_adapted = ai != adapted.end() && *ai == p;
         for (uint32_t r : rows) {
            if (r >= size) {
               errors[c] = "container row outside its chunk";
               return;
            }
            if (is_adapted) stored[static_cast<size_t>(ai - adapted.begin()) * kChunkWords + r / 64] |= 1ULL << (r % 64);
            if (s != col.reference[p]) out.push_back((static_cast<uint64_t>(r) << 24) | (static_cast<uint64_t>(p) << 8) | s);
         }
      }
      for (size_t a = 0; a < adapted.size(); ++a) {
         const uint32_t p = adapted[a];
         for (uint32_t r = 0; r < size; ++r) {
            const uint64_t i = off + r;
            if (f.is_null[i] || p < f.cov_start[i] || p >= f.cov_end[i]) continue;
            if ((stored[a * kChunkWords + r / 64] >> (r % 64)) & 1ULL) continue;
            if (isMissing(i, p)) continue;
            out.push_back((static_cast<uint64_t>(r) << 24) | (static_cast<uint64_t>(p) << 8) | f.local_reference[p]);
         }
      }
      sortRowKeys(out);
   });
   for (const auto& e : errors) {
      if (!e.empty()) fail(RDB_EINVAL, e);
   }
   // each chunk counts and writes its own rows' range of the diff form
   std::vector<uint64_t> mut_ptr(n + 1, 0);
   parallelFor(t.n_chunks, [&](size_t c) {
      for (uint64_t m : muts[c]) mut_ptr[t.chunk_offset[c] + (m >> 24) + 1] += 1;
   });
   for (uint64_t i = 0; i < n; ++i) mut_ptr[i + 1] += mut_ptr[i];
   std::vector<uint32_t> mut_pos(mut_ptr[n]);
   std::vector<uint8_t> mut_sym(mut_ptr[n]);
   parallelFor(t.n_chunks, [&](size_t c) {
      uint64_t w = mut_ptr[t.chunk_offset[c]];
      for (uint64_t m : muts[c]) {  // sorted by row, then position
         mut_pos[w] = static_cast<uint32_t>((m >> 8) & 0xFFFF);
         mut_sym[w] = static_cast<uint8_t>(m & 0xFF);
         ++w;
      }
      std::vector<uint64_t>().swap(muts[c]);
   });
   col.local_ref_given.assign(f.local_reference, f.local_reference + L);
   buildSequenceColumn(t, col, f.is_null, f.cov_start, f.cov_end, mut_ptr.data(), mut_pos.data(), mut_sym.data(),
                       nrun_ptr.data(), nrun_start.data(), nrun_end.data());
}

}  // namespace rdb
