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
}  // namespace

extern "C" {

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
