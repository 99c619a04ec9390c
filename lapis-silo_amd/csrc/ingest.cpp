// Host side of device ingestion: staging of aligned sequences for the diff kernel, the builder's
// diff-form rows, and the device stages of a sequence column's index build.
#include "ingest.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <string>

namespace rdb {

namespace {

template <typename T>
T* alloc(std::vector<DeviceBuffer>& bufs, size_t count) {
   bufs.emplace_back(std::max<size_t>(count * sizeof(T), 16));
   return bufs.back().as<T>();
}

template <typename T>
const T* upload(std::vector<DeviceBuffer>& bufs, const T* host, size_t count) {
   bufs.emplace_back(host, count * sizeof(T));
   return bufs.back().as<T>();
}

template <typename T>
void download(std::vector<T>& out, const T* dev, size_t count, hipStream_t stream) {
   out.resize(count);
   if (count == 0) return;
   hipCheck(hipMemcpyAsync(out.data(), dev, count * sizeof(T), hipMemcpyDeviceToHost, stream), "download");
}

// Appends at most this many characters, or rows, per diff launch pair.
constexpr uint64_t kPieceBytes = uint64_t{256} << 20;
constexpr uint64_t kPieceRows = uint64_t{1} << 20;

}  // namespace

// ---- diff of aligned sequences (aligned_sequence.cpp:18-122, sequence_column.cpp:312-341) -----------
void seqBuilderAppend(SeqBuilder& b, const char* data, const uint64_t* row_ptr, const uint32_t* offsets,
                      const uint8_t* is_null, uint64_t n_rows) {
   if (b.finished) fail(RDB_EINVAL, "the sequence column builder has finished");
   if (n_rows == 0) return;
   if (!row_ptr) fail(RDB_EINVAL, "null argument");
   if (!data && row_ptr[n_rows] > row_ptr[0]) fail(RDB_EINVAL, "null argument");
   Table& t = *b.table;
   const SeqColumn& col = *b.col;
   const uint32_t L = col.genome_length;
   if (b.is_null.size() + n_rows > t.num_rows) fail(RDB_EINVAL, "more rows appended than the table holds");
   for (uint64_t i = 0; i < n_rows; ++i) {
      if (row_ptr[i + 1] < row_ptr[i]) fail(RDB_EINVAL, "row offsets must not decrease");
      if (is_null && is_null[i]) continue;
      const uint64_t len = row_ptr[i + 1] - row_ptr[i];
      const uint64_t off = offsets ? offsets[i] : 0;
      if (len + off > L) {
         fail(RDB_EINVAL, "the sequence '" + std::string(data + row_ptr[i], len) + "' which was inserted with an offset " +
                             std::to_string(off) + " is larger than the length of the reference genome: " +
                             std::to_string(L));
      }
   }
   const size_t rows0 = b.is_null.size(), muts0 = b.mut_pos.size(), runs0 = b.nrun_start.size();
   auto rollback = [&] {
      b.is_null.resize(rows0);
      b.cov_start.resize(rows0);
      b.cov_end.resize(rows0);
      b.mut_ptr.resize(rows0 + 1);
      b.nrun_ptr.resize(rows0 + 1);
      b.mut_pos.resize(muts0);
      b.mut_sym.resize(muts0);
      b.nrun_start.resize(runs0);
      b.nrun_end.resize(runs0);
   };
   const hipStream_t stream = t.ctx->stream;
   try {
      uint64_t a = 0;
      while (a < n_rows) {
         uint64_t e = a + 1;
         while (e < n_rows && e - a < kPieceRows && row_ptr[e + 1] - row_ptr[a] <= kPieceBytes) ++e;
         const uint32_t n = static_cast<uint32_t>(e - a);
         std::vector<DeviceBuffer> bufs;
         std::vector<uint64_t> ptr(n + 1);
         for (uint32_t i = 0; i <= n; ++i) ptr[i] = row_ptr[a + i] - row_ptr[a];
         std::vector<uint8_t> nulls(n, 0);
         if (is_null) std::copy(is_null + a, is_null + e, nulls.begin());
         DiffBatch d;
         d.data = upload(bufs, data ? data + row_ptr[a] : nullptr, ptr[n]);
         d.row_ptr = upload(bufs, ptr.data(), ptr.size());
         d.offsets = offsets ? upload(bufs, offsets + a, n) : nullptr;
         d.is_null = upload(bufs, nulls.data(), n);
         d.lut = b.lut.as<uint8_t>();
         d.ref = b.ref.as<uint8_t>();
         d.n_rows = n;
         d.length = L;
         d.missing = col.abc->missing;
         uint32_t* stats = alloc<uint32_t>(bufs, 5ULL * n);
         d.start = stats;
         d.end = stats + n;
         d.n_mut = stats + 2ULL * n;
         d.n_run = stats + 3ULL * n;
         d.err = stats + 4ULL * n;
         hipCheck(launchDiff(d, false, stream), "diff count launch");
         std::vector<uint32_t> h;
         download(h, stats, 5ULL * n, stream);
         hipCheck(hipStreamSynchronize(stream), "diff count");
         for (uint32_t i = 0; i < n; ++i) {
            const uint32_t p = h[4ULL * n + i];
            if (p == 0xFFFFFFFFu) continue;
            const uint32_t off = offsets ? offsets[a + i] : 0;
            const char c = data[row_ptr[a + i] + (p - off)];
            fail(RDB_EINVAL, std::string("illegal character '") + c + "' at position " + std::to_string(p) +
                                " in the input sequence");
         }
         std::vector<uint64_t> mut_off(n), run_off(n);
         uint64_t nm = 0, nr = 0;
         for (uint32_t i = 0; i < n; ++i) {
            mut_off[i] = nm;
            run_off[i] = nr;
            nm += h[2ULL * n + i];
            nr += h[3ULL * n + i];
         }
         d.mut_off = upload(bufs, mut_off.data(), n);
         d.run_off = upload(bufs, run_off.data(), n);
         d.mut_pos = alloc<uint32_t>(bufs, nm);
         d.mut_sym = alloc<uint8_t>(bufs, nm);
         d.run_start = alloc<uint32_t>(bufs, nr);
         d.run_end = alloc<uint32_t>(bufs, nr);
         hipCheck(launchDiff(d, true, stream), "diff write launch");
         b.mut_pos.resize(b.mut_pos.size() + nm);
         b.mut_sym.resize(b.mut_pos.size());
         b.nrun_start.resize(b.nrun_start.size() + nr);
         b.nrun_end.resize(b.nrun_start.size());
         if (nm) {
            hipCheck(hipMemcpyAsync(b.mut_pos.data() + (b.mut_pos.size() - nm), d.mut_pos, nm * 4, hipMemcpyDeviceToHost, stream), "download");
            hipCheck(hipMemcpyAsync(b.mut_sym.data() + (b.mut_sym.size() - nm), d.mut_sym, nm, hipMemcpyDeviceToHost, stream), "download");
         }
         if (nr) {
            hipCheck(hipMemcpyAsync(b.nrun_start.data() + (b.nrun_start.size() - nr), d.run_start, nr * 4, hipMemcpyDeviceToHost, stream), "download");
            hipCheck(hipMemcpyAsync(b.nrun_end.data() + (b.nrun_end.size() - nr), d.run_end, nr * 4, hipMemcpyDeviceToHost, stream), "download");
         }
         hipCheck(hipStreamSynchronize(stream), "diff write");
         for (uint32_t i = 0; i < n; ++i) {
            b.is_null.push_back(nulls[i]);
            b.cov_start.push_back(h[i]);
            b.cov_end.push_back(h[static_cast<size_t>(n) + i]);
            b.mut_ptr.push_back(b.mut_ptr.back() + h[2ULL * n + i]);
            b.nrun_ptr.push_back(b.nrun_ptr.back() + h[3ULL * n + i]);
         }
         a = e;
      }
   } catch (...) {
      rollback();
      throw;
   }
}

void seqBuilderFinish(SeqBuilder& b) {
   if (b.finished) fail(RDB_EINVAL, "the sequence column builder has finished");
   Table& t = *b.table;
   if (b.is_null.size() != t.num_rows) {
      fail(RDB_EINVAL, "the builder holds " + std::to_string(b.is_null.size()) + " rows, the table " +
                          std::to_string(t.num_rows));
   }
   buildSequenceColumn(t, *b.col, b.is_null.data(), b.cov_start.data(), b.cov_end.data(), b.mut_ptr.data(),
                       b.mut_pos.data(), b.mut_sym.data(), b.nrun_ptr.data(), b.nrun_start.data(), b.nrun_end.data(),
                       true);
   t.seq.push_back(std::move(b.col));
   b.finished = true;
   b.lut = DeviceBuffer();
   b.ref = DeviceBuffer();
}

// ---- device stages of the index build ------------------------------------------------------------
DeviceIndexBuild::DeviceIndexBuild(Table& t, const SeqColumn& col, const uint8_t* is_null, const uint32_t* cov_start,
                                   const uint32_t* cov_end, const uint64_t* mut_ptr, const uint32_t* mut_pos,
                                   const uint8_t* mut_sym, const uint64_t* nrun_ptr, const uint32_t* runs)
    : t_(t) {
   const uint64_t n = t.num_rows;
   if (n > static_cast<uint64_t>(INT_MAX)) fail(RDB_EINVAL, "too many rows for the device index build");
   rows_.n_rows = n;
   rows_.length = col.genome_length;
   rows_.n_symbols = col.abc->count;
   rows_.is_null = upload(bufs_, is_null, n);
   rows_.cov_start = upload(bufs_, cov_start, n);
   rows_.cov_end = upload(bufs_, cov_end, n);
   rows_.mut_ptr = upload(bufs_, mut_ptr, n + 1);
   rows_.mut_pos = upload(bufs_, mut_pos, mut_ptr[n]);
   rows_.mut_sym = upload(bufs_, mut_sym, mut_ptr[n]);
   rows_.nrun_ptr = upload(bufs_, nrun_ptr, n + 1);
   rows_.runs = upload(bufs_, runs, nrun_ptr[n]);
   mut_total_ = mut_ptr[n];
   run_total_ = nrun_ptr[n];
   rows_.reference = upload(bufs_, col.reference.data(), col.genome_length);
}

void DeviceIndexBuild::adapt(std::vector<uint32_t>& coverage, std::vector<uint8_t>& local_ref) {
   const hipStream_t stream = t_.ctx->stream;
   const uint32_t L = rows_.length;
   const uint64_t n = rows_.n_rows;
   std::vector<DeviceBuffer> tmp;
   AdaptCount a;
   a.rows = rows_;
   a.key_count = mut_total_ + 2 * run_total_ + 2 * n;
   if (a.key_count > static_cast<uint64_t>(INT_MAX)) fail(RDB_ENOMEM, "too many rows for the device adaptation sort");
   a.keys = alloc<uint32_t>(tmp, a.key_count);
   a.keys_alt = alloc<uint32_t>(tmp, a.key_count);
   const size_t n_keys = (static_cast<size_t>(L) + 1) * 32;
   a.run_start = alloc<uint32_t>(tmp, n_keys);
   a.run_end = alloc<uint32_t>(tmp, n_keys);
   a.coverage = alloc<uint32_t>(tmp, L);
   a.local_ref = alloc<uint8_t>(tmp, L);
   hipCheck(hipMemsetAsync(a.run_start, 0, n_keys * 4, stream), "clear");
   hipCheck(hipMemsetAsync(a.run_end, 0, n_keys * 4, stream), "clear");
   size_t temp_bytes = 0;
   hipCheck(launchAdapt(a, nullptr, &temp_bytes, stream), "adaptation sort size");
   DeviceBuffer temp(std::max<size_t>(temp_bytes, 16));
   hipCheck(launchAdapt(a, temp.as<void>(), &temp_bytes, stream), "adaptation launch");
   download(coverage, a.coverage, L, stream);
   download(local_ref, a.local_ref, L, stream);
   hipCheck(hipStreamSynchronize(stream), "adaptation");
}

void DeviceIndexBuild::entries(const std::vector<uint8_t>& flip_to, const std::vector<uint32_t>& flipped,
                               RawVector<uint64_t>& keys, std::vector<uint64_t>& chunk_ptr) {
   const hipStream_t stream = t_.ctx->stream;
   const uint64_t n = rows_.n_rows;
   std::vector<DeviceBuffer> tmp;
   EntryBatch e;
   e.rows = rows_;
   e.chunk_offset = upload(tmp, t_.chunk_offset.data(), t_.chunk_offset.size());
   e.n_chunks = t_.n_chunks;
   e.flip_to = upload(tmp, flip_to.data(), flip_to.size());
   e.flipped = upload(tmp, flipped.data(), flipped.size());
   e.n_flipped = static_cast<uint32_t>(flipped.size());
   e.counts = alloc<uint32_t>(tmp, n);
   uint64_t* offsets = alloc<uint64_t>(tmp, n + 1);
   e.offsets = offsets;
   hipCheck(launchEntries(e, false, stream), "entry count launch");
   size_t scan_bytes = 0;
   hipCheck(scanEntryCounts(e.counts, offsets, n, nullptr, &scan_bytes, stream), "scan size");
   {
      DeviceBuffer scan_tmp(std::max<size_t>(scan_bytes, 16));
      hipCheck(scanEntryCounts(e.counts, offsets, n, scan_tmp.as<void>(), &scan_bytes, stream), "entry scan");
      hipCheck(hipStreamSynchronize(stream), "entry scan");
   }
   // chunk c's entries start at the offset of its first row
   std::vector<uint64_t> row_off;
   download(row_off, offsets, n, stream);
   uint32_t last_count = 0;
   if (n > 0) hipCheck(hipMemcpyAsync(&last_count, e.counts + (n - 1), 4, hipMemcpyDeviceToHost, stream), "download");
   hipCheck(hipStreamSynchronize(stream), "entry offsets");
   const uint64_t total = n > 0 ? row_off[n - 1] + last_count : 0;
   chunk_ptr.assign(t_.n_chunks + 1, total);
   for (uint32_t c = 0; c < t_.n_chunks; ++c) {
      if (t_.chunk_offset[c] < n) chunk_ptr[c] = row_off[t_.chunk_offset[c]];
   }
   uint64_t* k0 = alloc<uint64_t>(tmp, total);
   uint64_t* k1 = alloc<uint64_t>(tmp, total);
   e.keys = k0;
   hipCheck(launchEntries(e, true, stream), "entry write launch");
   // sort groups of whole chunks, each under 2^30 keys (chunk ids order the groups)
   // the host copy of the sorted keys: allocated without zeroing and first touched by the worker
   // pool, so that the page faults of a gigabyte are not taken by the copy's one thread
   keys.resize(total);
   {
      constexpr uint64_t kPiece = uint64_t{1} << 21;
      parallelFor((total + kPiece - 1) / kPiece, [&](size_t k) {
         const uint64_t a = k * kPiece, b = std::min(total, a + kPiece);
         std::memset(keys.data() + a, 0, (b - a) * 8);
      });
   }
   constexpr uint64_t kGroupKeys = uint64_t{1} << 30;
   std::vector<std::pair<uint64_t, uint64_t>> groups;
   {
      uint32_t c = 0;
      while (c < t_.n_chunks) {
         uint32_t d = c + 1;
         if (chunk_ptr[d] - chunk_ptr[c] > static_cast<uint64_t>(INT_MAX)) {
            fail(RDB_ENOMEM, "a chunk holds too many container entries for the device sort");
         }
         while (d < t_.n_chunks && chunk_ptr[d + 1] - chunk_ptr[c] <= kGroupKeys) ++d;
         groups.emplace_back(chunk_ptr[c], chunk_ptr[d]);
         c = d;
      }
   }
   uint64_t max_group = 0;
   for (auto& g : groups) max_group = std::max(max_group, g.second - g.first);
   size_t sort_bytes = 0;
   hipCheck(sortEntryKeys(k0, k1, max_group, nullptr, &sort_bytes, stream, nullptr), "sort size");
   DeviceBuffer sort_tmp(std::max<size_t>(sort_bytes, 16));
   for (auto& g : groups) {
      const uint64_t m = g.second - g.first;
      if (m == 0) continue;
      uint64_t* sorted = nullptr;
      size_t bytes = sort_bytes;
      hipCheck(sortEntryKeys(k0 + g.first, k1 + g.first, m, sort_tmp.as<void>(), &bytes, stream, &sorted), "entry sort");
      hipCheck(hipMemcpyAsync(keys.data() + g.first, sorted, m * 8, hipMemcpyDeviceToHost, stream), "download");
   }
   hipCheck(hipStreamSynchronize(stream), "entry sort");
   bufs_.clear();
}

// ---- digest -------------------------------------------------------------------------------------------
namespace {

struct Hasher {
   uint64_t h = 0x9E3779B97F4A7C15ULL;
   void words(const void* p, size_t bytes) {
      const auto* c = static_cast<const unsigned char*>(p);
      size_t i = 0;
      for (; i + 8 <= bytes; i += 8) {
         uint64_t w;
         std::memcpy(&w, c + i, 8);
         mix(w);
      }
      uint64_t tail = 0;
      if (bytes > i) std::memcpy(&tail, c + i, bytes - i);  // an empty vector's data() may be null
      mix(tail ^ (static_cast<uint64_t>(bytes) << 56));
   }
   void mix(uint64_t w) {
      h ^= w;
      h *= 0xFF51AFD7ED558CCDULL;
      h ^= h >> 29;
   }
   template <typename T, typename A>
   void vec(const std::vector<T, A>& v) {
      mix(v.size());
      words(v.data(), v.size() * sizeof(T));
   }
};

}  // namespace

uint64_t seqColumnDigest(const Table& t, const SeqColumn& col) {
   Hasher hs;
   std::vector<unsigned char> host;
   for (const DeviceBuffer& b : col.buffers) {
      host.resize(b.bytes());
      if (b.bytes()) hipCheck(hipMemcpy(host.data(), b.as<void>(), b.bytes(), hipMemcpyDeviceToHost), "download");
      hs.vec(host);
   }
   hs.vec(col.local_ref);
   hs.vec(col.coverage);
   hs.vec(col.block_pos_start);
   hs.vec(col.block_slot_start);
   hs.vec(col.pos_slot_ptr);
   hs.vec(col.slot_info);
   hs.vec(col.block_of_pos);
   hs.vec(col.slot_rows);
   hs.vec(col.chunk_index_bytes);
   hs.vec(col.chunk_nruns);
   hs.vec(col.row_mut_ptr);
   hs.vec(col.row_nrun_ptr);
   hs.mix(col.n_containers);
   hs.mix(col.payload_bytes);
   hs.mix(col.desc_bytes);
   hs.mix(col.n_flipped);
   hs.mix(col.row_major_bytes);
   hs.mix(col.has_nulls);
   hs.mix(t.num_rows);
   hs.mix(col.null_bitmap_id);
   if (col.null_bitmap_id < t.store_rows.size()) {
      for (const auto& rows : t.store_rows[col.null_bitmap_id]) hs.vec(rows);
   }
   return hs.h;
}

}  // namespace rdb
