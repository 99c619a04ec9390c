// Mutations rows written the way the reference's result sinks write them: NDJSON
// (query_engine/exec_node/ndjson_sink.cpp) and an Arrow IPC stream of one record batch
// (exec_node/arrow_ipc_sink.cpp, arrow::ipc::MakeStreamWriter), with the columns of
// MutationsNode::getOutputSchema (operators/mutations_node.h:60-84) in its order, restricted to the
// requested fields: mutationFrom, mutationTo, sequenceName (utf8), position (int32), proportion
// (float64), coverage, count (int32).
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "engine.h"

namespace rdb {

namespace {

enum FieldBit : uint32_t {
   F_FROM = RDB_FIELD_MUTATION_FROM,
   F_TO = RDB_FIELD_MUTATION_TO,
   F_SEQ = RDB_FIELD_SEQUENCE_NAME,
   F_POS = RDB_FIELD_POSITION,
   F_PROP = RDB_FIELD_PROPORTION,
   F_COV = RDB_FIELD_COVERAGE,
   F_COUNT = RDB_FIELD_COUNT,
};
// getOutputSchema's order
constexpr FieldBit kOrder[] = {F_FROM, F_TO, F_SEQ, F_POS, F_PROP, F_COV, F_COUNT};
const char* fieldName(FieldBit f) {
   switch (f) {
      case F_FROM: return "mutationFrom";
      case F_TO: return "mutationTo";
      case F_SEQ: return "sequenceName";
      case F_POS: return "position";
      case F_PROP: return "proportion";
      case F_COV: return "coverage";
      case F_COUNT: return "count";
   }
   return "";
}

// ---- JSON values ---------------------------------------------------------------------------------
void jsonString(std::string& out, const char* s, size_t n) {
   static const char* hex = "0123456789abcdef";
   out.push_back('"');
   for (size_t i = 0; i < n; ++i) {
      const unsigned char c = static_cast<unsigned char>(s[i]);
      switch (c) {
         case '"': out += "\\\""; break;
         case '\\': out += "\\\\"; break;
         case '\b': out += "\\b"; break;
         case '\f': out += "\\f"; break;
         case '\n': out += "\\n"; break;
         case '\r': out += "\\r"; break;
         case '\t': out += "\\t"; break;
         default:
            if (c < 0x20) {
               out += "\\u00";
               out.push_back(hex[c >> 4]);
               out.push_back(hex[c & 15]);
            } else {
               out.push_back(static_cast<char>(c));
            }
      }
   }
   out.push_back('"');
}

// A finite double as its shortest round-trip digits, in fixed notation when the decimal point
// falls within (-4, 15] digits of the first digit, else d.ddde±XX; integral values keep ".0".
void jsonDouble(std::string& out, double v) {
   if (v == 0.0) {
      out += std::signbit(v) ? "-0.0" : "0.0";
      return;
   }
   char sci[64];
   const auto r = std::to_chars(sci, sci + sizeof(sci), v, std::chars_format::scientific);
   std::string s(sci, r.ptr);
   std::string sign;
   if (s[0] == '-') {
      sign = "-";
      s.erase(0, 1);
   }
   const size_t e = s.find('e');
   const int exp10 = std::stoi(s.substr(e + 1));
   std::string digits = s.substr(0, e);
   digits.erase(std::remove(digits.begin(), digits.end(), '.'), digits.end());
   const int k = static_cast<int>(digits.size());
   const int n = exp10 + 1;  // position of the decimal point from the first digit
   out += sign;
   if (k <= n && n <= 15) {
      out += digits;
      out.append(static_cast<size_t>(n - k), '0');
      out += ".0";
   } else if (0 < n && n <= 15) {
      out += digits.substr(0, static_cast<size_t>(n));
      out.push_back('.');
      out += digits.substr(static_cast<size_t>(n));
   } else if (-4 < n && n <= 0) {
      out += "0.";
      out.append(static_cast<size_t>(-n), '0');
      out += digits;
   } else {
      out.push_back(digits[0]);
      if (k > 1) {
         out.push_back('.');
         out += digits.substr(1);
      }
      out.push_back('e');
      const int x = n - 1;
      out.push_back(x < 0 ? '-' : '+');
      const int a = x < 0 ? -x : x;
      if (a < 10) out.push_back('0');
      out += std::to_string(a);
   }
}

// ---- flatbuffers (front to back: a table's vtable before it, its children after it) --------------
class Flat {
  public:
   std::vector<uint8_t> b;
   size_t pos() const { return b.size(); }
   void align(size_t a) {
      while (b.size() % a) b.push_back(0);
   }
   template <class T>
   void put(T v) {
      uint8_t raw[sizeof(T)];
      std::memcpy(raw, &v, sizeof(T));
      b.insert(b.end(), raw, raw + sizeof(T));
   }
   template <class T>
   void patch(size_t at, T v) {
      std::memcpy(b.data() + at, &v, sizeof(T));
   }
   // an offset slot at `at` pointing forward to `target`
   void link(size_t at, size_t target) { patch<uint32_t>(at, static_cast<uint32_t>(target - at)); }

   // A table: fields given as (slot, size, alignment) in slot order of the layout below; returns
   // the table's position and each field's position.
   struct Field {
      int slot;
      size_t size;
   };
   std::pair<size_t, std::vector<size_t>> table(const std::vector<Field>& fields, int n_slots) {
      // vtable: u16 vtable size, u16 table size, u16 per slot
      align(2);
      // the table (soffset, then fields each aligned to its size) must start 4-aligned; its fields
      // are laid out after the soffset in the given order
      size_t table_size = 4;
      std::vector<size_t> field_off(fields.size());
      for (size_t i = 0; i < fields.size(); ++i) {
         while (table_size % fields[i].size) ++table_size;
         field_off[i] = table_size;
         table_size += fields[i].size;
      }
      const size_t vt_size = 4 + 2 * static_cast<size_t>(n_slots);
      // place the vtable so that the table right after it (padded) is aligned to 8
      size_t vt = pos();
      size_t tbl = vt + vt_size;
      while (tbl % 8) ++tbl;
      // pad before the vtable so that the vtable ends exactly where the table (8-aligned) begins
      const size_t pad = (tbl - (vt + vt_size));
      b.insert(b.end(), pad, 0);
      vt += pad;
      tbl = vt + vt_size;
      put<uint16_t>(static_cast<uint16_t>(vt_size));
      put<uint16_t>(static_cast<uint16_t>(table_size));
      std::vector<uint16_t> slots(static_cast<size_t>(n_slots), 0);
      for (size_t i = 0; i < fields.size(); ++i) slots[static_cast<size_t>(fields[i].slot)] = static_cast<uint16_t>(field_off[i]);
      for (uint16_t s : slots) put<uint16_t>(s);
      put<int32_t>(static_cast<int32_t>(tbl - vt));  // soffset: vtable = table - soffset
      b.resize(tbl + table_size, 0);
      std::vector<size_t> at(fields.size());
      for (size_t i = 0; i < fields.size(); ++i) at[i] = tbl + field_off[i];
      return {tbl, at};
   }
   size_t string(const std::string& s) {
      align(4);
      const size_t at = pos();
      put<uint32_t>(static_cast<uint32_t>(s.size()));
      b.insert(b.end(), s.begin(), s.end());
      b.push_back(0);
      return at;
   }
   // a vector of n offsets to fill; returns (vector position, first element position)
   std::pair<size_t, size_t> offsetVector(size_t n) {
      align(4);
      const size_t at = pos();
      put<uint32_t>(static_cast<uint32_t>(n));
      const size_t first = pos();
      b.resize(b.size() + 4 * n, 0);
      return {at, first};
   }
   // a vector of 16-byte structs of two i64
   size_t structVector(const std::vector<std::pair<int64_t, int64_t>>& v) {
      // the elements must be 8-aligned: the u32 length right before them
      align(4);
      if ((pos() + 4) % 8) put<uint32_t>(0);
      const size_t at = pos();
      put<uint32_t>(static_cast<uint32_t>(v.size()));
      for (const auto& [x, y] : v) {
         put<int64_t>(x);
         put<int64_t>(y);
      }
      return at;
   }
};

// Arrow Type union members and Message header members (Schema.fbs, Message.fbs)
constexpr uint8_t kTypeInt = 2, kTypeFloatingPoint = 3, kTypeUtf8 = 5;
constexpr uint8_t kHeaderSchema = 1, kHeaderRecordBatch = 3;
constexpr int16_t kMetadataV5 = 4;

struct ColumnOut {
   FieldBit field;
   uint8_t type;
   std::vector<std::vector<uint8_t>> buffers;  // validity (empty), then offsets/data or values
};

// A Message flatbuffer with the given header; the header's table is written by `header` after the
// message table and returns its position.
template <class H>
std::vector<uint8_t> message(uint8_t header_type, int64_t body_length, H&& header) {
   Flat f;
   f.put<uint32_t>(0);  // root offset
   // Message: 0 version (short), 1 header_type (ubyte), 2 header (offset), 3 bodyLength (long)
   auto [tbl, at] = f.table({{3, 8}, {2, 4}, {0, 2}, {1, 1}}, 4);
   f.link(0, tbl);
   f.patch<int64_t>(at[0], body_length);
   f.patch<int16_t>(at[2], kMetadataV5);
   f.patch<uint8_t>(at[3], header_type);
   const size_t h = header(f);
   f.link(at[1], h);
   f.align(8);
   return f.b;
}

void appendMessage(std::vector<uint8_t>& out, const std::vector<uint8_t>& meta, const std::vector<uint8_t>& body) {
   // encapsulated message: continuation marker, metadata length (padded to 8), metadata, body
   const uint32_t cont = 0xFFFFFFFFu;
   size_t len = meta.size();
   while ((8 + len) % 8) ++len;
   const int32_t len32 = static_cast<int32_t>(len);
   const auto* c = reinterpret_cast<const uint8_t*>(&cont);
   out.insert(out.end(), c, c + 4);
   const auto* l = reinterpret_cast<const uint8_t*>(&len32);
   out.insert(out.end(), l, l + 4);
   out.insert(out.end(), meta.begin(), meta.end());
   out.insert(out.end(), len - meta.size(), 0);
   out.insert(out.end(), body.begin(), body.end());
}

}  // namespace

namespace capi {

uint32_t validFields(uint32_t fields) {
   if (fields == 0 || (fields & ~static_cast<uint32_t>(RDB_FIELDS_ALL)) != 0) fail(RDB_EINVAL, "unknown output fields");
   return fields;
}

void checkSymbols(const MutationRowsView& r) {
   for (uint64_t i = 0; i < r.n; ++i) {
      if (r.from[i] >= r.alphabet->count || r.to[i] >= r.alphabet->count) fail(RDB_EINVAL, "symbol outside the alphabet");
   }
}

std::string mutationRowsNdjson(const MutationRowsView& r) {
   const std::string& chars = r.alphabet->chars;
   std::string out;
   out.reserve(r.n * 128);
   for (uint64_t i = 0; i < r.n; ++i) {
      out.push_back('{');
      bool first = true;
      for (FieldBit f : kOrder) {
         if (!(r.fields & f)) continue;
         if (!first) out.push_back(',');
         first = false;
         out.push_back('"');
         out += fieldName(f);
         out += "\":";
         switch (f) {
            case F_FROM: jsonString(out, &chars[r.from[i]], 1); break;
            case F_TO: jsonString(out, &chars[r.to[i]], 1); break;
            case F_SEQ: jsonString(out, r.sequence_name, std::strlen(r.sequence_name)); break;
            case F_POS: out += std::to_string(static_cast<int32_t>(r.position[i])); break;
            case F_PROP: jsonDouble(out, r.proportion[i]); break;
            case F_COV: out += std::to_string(static_cast<int32_t>(r.coverage[i])); break;
            case F_COUNT: out += std::to_string(static_cast<int32_t>(r.count[i])); break;
         }
      }
      out += "}\n";
   }
   return out;
}

std::vector<uint8_t> mutationRowsArrowIpc(const MutationRowsView& r) {
   const std::string& chars = r.alphabet->chars;
   const uint64_t n = r.n;
   std::vector<ColumnOut> cols;
   auto i32col = [&](FieldBit f, auto get) {
      ColumnOut c{f, kTypeInt, {{}, std::vector<uint8_t>(4 * n)}};
      for (uint64_t i = 0; i < n; ++i) {
         const int32_t v = static_cast<int32_t>(get(i));
         std::memcpy(c.buffers[1].data() + 4 * i, &v, 4);
      }
      cols.push_back(std::move(c));
   };
   auto strcol = [&](FieldBit f, auto get) {
      ColumnOut c{f, kTypeUtf8, {{}, std::vector<uint8_t>(4 * (n + 1)), {}}};
      int32_t off = 0;
      std::memcpy(c.buffers[1].data(), &off, 4);
      for (uint64_t i = 0; i < n; ++i) {
         const std::string s = get(i);
         c.buffers[2].insert(c.buffers[2].end(), s.begin(), s.end());
         off = static_cast<int32_t>(c.buffers[2].size());
         std::memcpy(c.buffers[1].data() + 4 * (i + 1), &off, 4);
      }
      cols.push_back(std::move(c));
   };
   for (FieldBit f : kOrder) {
      if (!(r.fields & f)) continue;
      switch (f) {
         case F_FROM: strcol(f, [&](uint64_t i) { return std::string(1, chars[r.from[i]]); }); break;
         case F_TO: strcol(f, [&](uint64_t i) { return std::string(1, chars[r.to[i]]); }); break;
         case F_SEQ: strcol(f, [&](uint64_t) { return std::string(r.sequence_name); }); break;
         case F_POS: i32col(f, [&](uint64_t i) { return r.position[i]; }); break;
         case F_COV: i32col(f, [&](uint64_t i) { return r.coverage[i]; }); break;
         case F_COUNT: i32col(f, [&](uint64_t i) { return r.count[i]; }); break;
         case F_PROP: {
            ColumnOut c{f, kTypeFloatingPoint, {{}, std::vector<uint8_t>(8 * n)}};
            if (n) std::memcpy(c.buffers[1].data(), r.proportion, 8 * n);
            cols.push_back(std::move(c));
            break;
         }
      }
   }
   std::vector<uint8_t> out;
   // ---- schema message
   const auto schema = message(kHeaderSchema, 0, [&](Flat& f) {
      // Schema: 0 endianness (short, Little = 0), 1 fields
      auto [tbl, at] = f.table({{1, 4}, {0, 2}}, 2);
      auto [vec, first] = f.offsetVector(cols.size());
      f.link(at[0], vec);
      for (size_t k = 0; k < cols.size(); ++k) {
         // Field: 0 name, 1 nullable, 2 type_type, 3 type, 5 children
         auto [ft, fa] = f.table({{0, 4}, {3, 4}, {5, 4}, {1, 1}, {2, 1}}, 6);
         f.link(first + 4 * k, ft);
         f.patch<uint8_t>(fa[3], 1);
         f.patch<uint8_t>(fa[4], cols[k].type);
         f.link(fa[0], f.string(fieldName(cols[k].field)));
         size_t type_tbl;
         if (cols[k].type == kTypeInt) {
            auto [it, ia] = f.table({{0, 4}, {1, 1}}, 2);  // Int: bitWidth, is_signed
            f.patch<int32_t>(ia[0], 32);
            f.patch<uint8_t>(ia[1], 1);
            type_tbl = it;
         } else if (cols[k].type == kTypeFloatingPoint) {
            auto [pt, pa] = f.table({{0, 2}}, 1);  // FloatingPoint: precision DOUBLE = 2
            f.patch<int16_t>(pa[0], 2);
            type_tbl = pt;
         } else {
            type_tbl = f.table({}, 0).first;  // Utf8 {}
         }
         f.link(fa[1], type_tbl);
         f.link(fa[2], f.offsetVector(0).first);
      }
      return tbl;
   });
   appendMessage(out, schema, {});
   // ---- record batch message
   std::vector<uint8_t> body;
   std::vector<std::pair<int64_t, int64_t>> nodes, buffers;
   for (const auto& c : cols) {
      nodes.emplace_back(static_cast<int64_t>(n), 0);
      for (const auto& buf : c.buffers) {
         buffers.emplace_back(static_cast<int64_t>(body.size()), static_cast<int64_t>(buf.size()));
         body.insert(body.end(), buf.begin(), buf.end());
         while (body.size() % 8) body.push_back(0);
      }
   }
   const auto batch = message(kHeaderRecordBatch, static_cast<int64_t>(body.size()), [&](Flat& f) {
      // RecordBatch: 0 length (long), 1 nodes, 2 buffers
      auto [tbl, at] = f.table({{0, 8}, {1, 4}, {2, 4}}, 3);
      f.patch<int64_t>(at[0], static_cast<int64_t>(n));
      f.link(at[1], f.structVector(nodes));
      f.link(at[2], f.structVector(buffers));
      return tbl;
   });
   appendMessage(out, batch, body);
   // end of stream
   const uint32_t eos[2] = {0xFFFFFFFFu, 0};
   const auto* e = reinterpret_cast<const uint8_t*>(eos);
   out.insert(out.end(), e, e + 8);
   return out;
}

}  // namespace capi

}  // namespace rdb
