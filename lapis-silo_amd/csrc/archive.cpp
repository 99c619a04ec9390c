// Scanners for the two long runs of records in a saved sequence column (the reference's Boost
// binary archives, storage/column/sequence_column.h:86-97): the vertical index's container
// entries and the horizontal index's per-row bitmaps.  After the first entry of each (which carries
// the class information) the records have a fixed layout, so they are scanned here in one pass
// instead of record by record in the host language.
#include <cstring>

#include "capi_internal.h"

using namespace rdb;
using namespace rdb::capi;

namespace {
template <typename T>
T load(const uint8_t* p) {
   T v;
   std::memcpy(&v, p, sizeof(T));
   return v;
}
// The maximal runs [a, b) of one portable roaring bitmap (CRoaring's portable format, the one
// Roaring::write produces: roaring_serialize.h:16-48), as `emit(a, b)` calls in ascending order.
// Malformed input (a short buffer, a bad cookie, keys out of order, container values out of order
// or not matching their cardinality, a run past the end of its 2^16 block) fails with RDB_EINVAL.
template <typename Emit>
void bitmapRuns(const uint8_t* b, uint64_t n, Emit&& emit) {
   constexpr uint32_t kCookieNoRun = 12346, kCookie = 12347, kNoOffsetThreshold = 4;
   auto need = [&](uint64_t pos, uint64_t bytes) {
      if (pos > n || bytes > n - pos) fail(RDB_EINVAL, "truncated roaring bitmap");
   };
   need(0, 4);
   const uint32_t cookie = load<uint32_t>(b);
   uint64_t pos = 4, size = 0;
   const uint8_t* run_flags = nullptr;
   if ((cookie & 0xFFFFu) == kCookie) {
      size = (cookie >> 16) + 1;
      need(pos, (size + 7) / 8);
      run_flags = b + pos;
      pos += (size + 7) / 8;
   } else if (cookie == kCookieNoRun) {
      need(pos, 4);
      size = load<uint32_t>(b + pos);
      pos += 4;
   } else {
      fail(RDB_EINVAL, "bad roaring cookie");
   }
   if (size > 65536) fail(RDB_EINVAL, "truncated roaring bitmap");
   need(pos, 4 * size);
   const uint8_t* hdr = b + pos;
   pos += 4 * size;
   if (!run_flags || size >= kNoOffsetThreshold) pos += 4 * size;  // container offsets
   bool open = false;
   uint64_t ra = 0, rb = 0;  // the run being extended
   auto add = [&](uint64_t a, uint64_t e) {  // values [a, e), after every earlier value
      if (open && a == rb) {
         rb = e;
         return;
      }
      if (open) emit(ra, rb);
      open = true;
      ra = a;
      rb = e;
   };
   for (uint64_t k = 0; k < size; ++k) {
      const uint64_t key = load<uint16_t>(hdr + 4 * k);
      const uint64_t card = uint64_t{load<uint16_t>(hdr + 4 * k + 2)} + 1;
      if (k > 0 && key <= load<uint16_t>(hdr + 4 * (k - 1))) fail(RDB_EINVAL, "container keys out of order");
      const uint64_t base = key << 16;
      uint64_t got = 0;
      int64_t last = -1;
      if (run_flags && ((run_flags[k / 8] >> (k % 8)) & 1u)) {
         need(pos, 2);
         const uint32_t n_runs = load<uint16_t>(b + pos);
         need(pos + 2, 4ULL * n_runs);
         for (uint32_t j = 0; j < n_runs; ++j) {
            const uint32_t s = load<uint16_t>(b + pos + 2 + 4 * j), l = load<uint16_t>(b + pos + 4 + 4 * j);
            if (s + l > 0xFFFFu) fail(RDB_EINVAL, "run past the end of its block");
            if (static_cast<int64_t>(s) <= last) fail(RDB_EINVAL, "container values do not match its cardinality");
            add(base + s, base + s + l + 1);
            got += l + 1;
            last = s + l;
         }
         pos += 2 + 4ULL * n_runs;
      } else if (card > 4096) {
         need(pos, 8192);
         for (uint32_t w = 0; w < 1024; ++w) {
            uint64_t x = load<uint64_t>(b + pos + 8 * w);
            while (x) {
               const uint32_t bit = static_cast<uint32_t>(__builtin_ctzll(x));
               const uint64_t ones = ~x >> bit;  // the run of set bits from `bit` up
               const uint32_t len = ones ? static_cast<uint32_t>(__builtin_ctzll(ones)) : 64 - bit;
               add(base + 64 * w + bit, base + 64 * w + bit + len);
               got += len;
               x = len + bit >= 64 ? 0 : x & (~uint64_t{0} << (bit + len));
            }
         }
         pos += 8192;
      } else {
         need(pos, 2 * card);
         for (uint64_t j = 0; j < card; ++j) {
            const int64_t v = load<uint16_t>(b + pos + 2 * j);
            if (v <= last) fail(RDB_EINVAL, "container values do not match its cardinality");
            add(base + v, base + v + 1);
            last = v;
         }
         got = card;
         pos += 2 * card;
      }
      if (got != card) fail(RDB_EINVAL, "container values do not match its cardinality");
   }
   if (open) emit(ra, rb);
}

}  // namespace

extern "C" {

int rdb_archive_bitmap_runs(const uint8_t* buf, uint64_t len, const uint64_t* start, const uint64_t* length, uint64_t n,
                            uint64_t* run_ptr, uint32_t* run_start, uint32_t* run_end, uint64_t capacity) {
   return guarded(nullptr, [&] {
      if (!buf || !run_ptr || (n && (!start || !length))) fail(RDB_EINVAL, "null argument");
      for (uint64_t i = 0; i < n; ++i) {
         if (start[i] > len || length[i] > len - start[i]) fail(RDB_EINVAL, "roaring bitmap outside the buffer");
      }
      // counts per bitmap (into run_ptr[i + 1]), then offsets; with output arrays, the runs
      const bool fill = run_start && run_end;
      if (!fill) {
         run_ptr[0] = 0;
         parallelFor(n, [&](size_t i) {
            uint64_t c = 0;
            bitmapRuns(buf + start[i], length[i], [&](uint64_t, uint64_t) { ++c; });
            run_ptr[i + 1] = c;
         });
         for (uint64_t i = 0; i < n; ++i) run_ptr[i + 1] += run_ptr[i];
         return;
      }
      if (n && run_ptr[n] > capacity) fail(RDB_EINVAL, "run arrays too small");
      parallelFor(n, [&](size_t i) {
         uint64_t o = run_ptr[i];
         const uint64_t e = run_ptr[i + 1];
         bitmapRuns(buf + start[i], length[i], [&](uint64_t a, uint64_t b) {
            if (o >= e) fail(RDB_EINVAL, "run counts changed between passes");
            run_start[o] = static_cast<uint32_t>(a);
            run_end[o] = static_cast<uint32_t>(b);
            ++o;
         });
         if (o != e) fail(RDB_EINVAL, "run counts changed between passes");
      });
   });
}

int rdb_archive_scan_containers(const uint8_t* buf, uint64_t len, uint64_t pos, uint64_t n, uint32_t* position,
                                uint16_t* chunk, int32_t* symbol, uint32_t* cardinality, uint8_t* typecode,
                                uint64_t* payload_offset, uint8_t* payload, uint64_t payload_capacity,
                                uint64_t* payload_bytes, uint64_t* end) {
   return guarded(nullptr, [&] {
      if (!buf || !payload_bytes || !end || (n && (!position || !chunk || !symbol || !cardinality || !typecode ||
                                                   !payload_offset))) {
         fail(RDB_EINVAL, "null argument");
      }
      // record: u32 position, u16 chunk, i32 symbol (vertical_sequence_index.h:31-38), u32
      // cardinality, u8 typecode, u64 length + the container_write bytes (roaring_container.h:105-116)
      constexpr uint64_t kHeader = 4 + 2 + 4 + 4 + 1 + 8;
      uint64_t total = 0;
      for (uint64_t i = 0; i < n; ++i) {
         if (pos > len || len - pos < kHeader) fail(RDB_EINVAL, "truncated vertical index");
         const uint8_t* p = buf + pos;
         position[i] = load<uint32_t>(p);
         chunk[i] = load<uint16_t>(p + 4);
         symbol[i] = load<int32_t>(p + 6);
         cardinality[i] = load<uint32_t>(p + 10);
         typecode[i] = p[14];
         const uint64_t bytes = load<uint64_t>(p + 15);
         pos += kHeader;
         if (bytes > len - pos) fail(RDB_EINVAL, "truncated container");
         payload_offset[i] = total;
         if (payload) {
            if (total + bytes > payload_capacity) fail(RDB_EINVAL, "payload buffer too small");
            std::memcpy(payload + total, buf + pos, bytes);
         }
         total += bytes;
         pos += bytes;
      }
      *payload_bytes = total;
      *end = pos;
   });
}

int rdb_archive_scan_row_bitmaps(const uint8_t* buf, uint64_t len, uint64_t pos, uint64_t n, uint32_t* row,
                                 uint64_t* bitmap_start, uint64_t* bitmap_length, uint64_t* end) {
   return guarded(nullptr, [&] {
      if (!buf || !end || (n && (!row || !bitmap_start || !bitmap_length))) fail(RDB_EINVAL, "null argument");
      // record: u32 row id, u64 length + a portable roaring bitmap (roaring_serialize.h:16-48)
      for (uint64_t i = 0; i < n; ++i) {
         if (pos > len || len - pos < 12) fail(RDB_EINVAL, "truncated coverage index");
         row[i] = load<uint32_t>(buf + pos);
         const uint64_t bytes = load<uint64_t>(buf + pos + 4);
         pos += 12;
         if (bytes > len - pos) fail(RDB_EINVAL, "truncated bitmap");
         bitmap_start[i] = pos;
         bitmap_length[i] = bytes;
         pos += bytes;
      }
      *end = pos;
   });
}

}  // extern "C"
