// Memory-safety check of the engine's host code through its C ABI.  The Makefile's `sanitize`
// target compiles every host source of the engine with AddressSanitizer and UndefinedBehavior-
// Sanitizer (device code is built as usual) and links them with this driver, the counterpart of
// the reference's Debug build with ASAN.  Both sanitizers stop the process at their first finding.
//
//   sanitize_check host    entry points without device work: the diff of aligned sequences, the
//                          saved-state scanners and bitmap decoder on whole, truncated and
//                          corrupted buffers, the
//                          NDJSON and Arrow IPC writers with every capacity below the needed one;
//   sanitize_check device  on GPU 0: tables built from the same rows by the device builder and by
//                          the diff-form upload (equal digests), sequence, date, integer,
//                          dictionary and lineage columns, every filter leaf kind, Count /
//                          Mutations / Aggregated / the sinks / the allocation-free entry point /
//                          batches under each scan mode against the single-query results, a failed
//                          batch submission, malformed uploads and queries.
//
// Exit status 0 when every check passes; a failed check prints its line and exits with 1.
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "rdb.h"

namespace {

int failures = 0;

#define CHECK(cond)                                                             \
   do {                                                                         \
      if (!(cond)) {                                                            \
         std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
         ++failures;                                                            \
      }                                                                         \
   } while (0)

struct Rng {
   uint64_t s;
   uint64_t next() {  // splitmix64
      uint64_t z = (s += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
   }
   uint32_t below(uint32_t n) { return static_cast<uint32_t>(next() % n); }
   bool chance(double p) { return (next() >> 11) * (1.0 / 9007199254740992.0) < p; }
};

template <typename T>
void put(std::vector<uint8_t>& b, T v) {
   const auto* p = reinterpret_cast<const uint8_t*>(&v);
   b.insert(b.end(), p, p + sizeof(T));
}

// ---- host mode ------------------------------------------------------------------------------------

void checkDiff(Rng& rng) {
   const char* nuc = "ACGT-RYSWKMBDHVN";
   const uint32_t L = 200;
   std::string ref(L, 'A');
   for (auto& c : ref) c = "ACGT"[rng.below(4)];
   std::vector<uint32_t> missing(L), mpos(L);
   std::vector<uint8_t> msym(L);
   for (int i = 0; i < 200; ++i) {
      std::string s(L, 'N');
      const uint32_t a = rng.below(L), b = a + rng.below(L - a + 1);
      for (uint32_t p = a; p < b; ++p) s[p] = rng.chance(0.8) ? ref[p] : nuc[rng.below(16)];
      uint32_t start = 0, end = 0;
      uint64_t nm = 0, nmut = 0;
      CHECK(rdb_diff_sequence(RDB_NUCLEOTIDE, s.data(), L, ref.data(), &start, &end, missing.data(), L, &nm,
                              mpos.data(), msym.data(), L, &nmut) == RDB_OK);
      CHECK(start <= end && end <= L && nm <= L && nmut <= L);
      if (nmut > 0) {  // too small a capacity is an error, not an overrun
         CHECK(rdb_diff_sequence(RDB_NUCLEOTIDE, s.data(), L, ref.data(), &start, &end, missing.data(), L, &nm,
                                 mpos.data(), msym.data(), nmut - 1, &nmut) == RDB_ENOMEM);
      }
   }
   std::string bad(L, 'A');
   bad[17] = 'J';
   uint32_t start, end;
   uint64_t nm, nmut;
   CHECK(rdb_diff_sequence(RDB_NUCLEOTIDE, bad.data(), L, ref.data(), &start, &end, missing.data(), L, &nm,
                           mpos.data(), msym.data(), L, &nmut) == RDB_EINVAL);
}

// Vertical-index records followed by horizontal-index records, as a saved sequence column holds them.
void checkArchiveScanners(Rng& rng) {
   std::vector<uint8_t> buf;
   const uint64_t n = 40;
   uint64_t want_payload = 0;
   for (uint64_t i = 0; i < n; ++i) {
      const uint64_t bytes = 2 * rng.below(300);
      put<uint32_t>(buf, static_cast<uint32_t>(i));
      put<uint16_t>(buf, static_cast<uint16_t>(rng.below(3)));
      put<int32_t>(buf, static_cast<int32_t>(rng.below(16)));
      put<uint32_t>(buf, static_cast<uint32_t>(bytes / 2));
      put<uint8_t>(buf, 2);
      put<uint64_t>(buf, bytes);
      for (uint64_t k = 0; k < bytes; ++k) buf.push_back(static_cast<uint8_t>(rng.next()));
      want_payload += bytes;
   }
   const uint64_t rows_at = buf.size();
   for (uint64_t i = 0; i < n; ++i) {
      const uint64_t bytes = rng.below(64);
      put<uint32_t>(buf, static_cast<uint32_t>(i));
      put<uint64_t>(buf, bytes);
      for (uint64_t k = 0; k < bytes; ++k) buf.push_back(static_cast<uint8_t>(rng.next()));
   }
   std::vector<uint32_t> position(n), card(n), row(n);
   std::vector<uint16_t> chunk(n);
   std::vector<int32_t> symbol(n);
   std::vector<uint8_t> tc(n), payload(want_payload);
   std::vector<uint64_t> off(n), bstart(n), blen(n);
   uint64_t pb = 0, end = 0;
   CHECK(rdb_archive_scan_containers(buf.data(), buf.size(), 0, n, position.data(), chunk.data(), symbol.data(),
                                     card.data(), tc.data(), off.data(), payload.data(), payload.size(), &pb,
                                     &end) == RDB_OK);
   CHECK(pb == want_payload && end == rows_at);
   if (want_payload > 0) {
      CHECK(rdb_archive_scan_containers(buf.data(), buf.size(), 0, n, position.data(), chunk.data(), symbol.data(),
                                        card.data(), tc.data(), off.data(), payload.data(), want_payload - 1, &pb,
                                        &end) == RDB_EINVAL);
   }
   CHECK(rdb_archive_scan_row_bitmaps(buf.data(), buf.size(), rows_at, n, row.data(), bstart.data(), blen.data(),
                                      &end) == RDB_OK);
   CHECK(end == buf.size());
   // every truncation is reported (the scanners read exactly the buffer they are given: a copy of
   // the prefix lets the sanitizer see any read past it)
   for (uint64_t cut = 0; cut < buf.size(); cut += 1 + rng.below(7)) {
      std::vector<uint8_t> head(buf.begin(), buf.begin() + cut);
      CHECK(rdb_archive_scan_containers(head.data(), cut, 0, n, position.data(), chunk.data(), symbol.data(),
                                        card.data(), tc.data(), off.data(), nullptr, 0, &pb, &end) ==
            (cut >= rows_at ? RDB_OK : RDB_EINVAL));
      if (cut >= rows_at) {
         CHECK(rdb_archive_scan_row_bitmaps(head.data(), cut, rows_at, n, row.data(), bstart.data(), blen.data(),
                                            &end) == (cut == buf.size() ? RDB_OK : RDB_EINVAL));
      }
   }
   // lengths near 2^64 must not wrap a bounds check
   for (int k = 0; k < 200; ++k) {
      std::vector<uint8_t> bad(buf);
      const uint64_t at = 15 + 8 * rng.below(4);
      const uint64_t huge = ~0ull - rng.below(1000);
      if (at + 8 <= bad.size()) std::memcpy(bad.data() + at, &huge, 8);
      rdb_archive_scan_containers(bad.data(), bad.size(), 0, n, position.data(), chunk.data(), symbol.data(),
                                  card.data(), tc.data(), off.data(), nullptr, 0, &pb, &end);
      rdb_archive_scan_row_bitmaps(bad.data(), bad.size(), rows_at, n, row.data(), bstart.data(), blen.data(), &end);
   }
   CHECK(rdb_archive_scan_containers(buf.data(), buf.size(), buf.size() + 5, 1, position.data(), chunk.data(),
                                     symbol.data(), card.data(), tc.data(), off.data(), nullptr, 0, &pb,
                                     &end) == RDB_EINVAL);
}

// A portable roaring bitmap with a run container (key 0: [5, 10) and [20, 25)), an array container
// (key 1: 7, 8, 9) and a bitset container (key 2: its first 5,000 rows), decoded to runs whole and
// after random damage (every outcome is a result or RDB_EINVAL).
void checkBitmapRuns(Rng& rng) {
   std::vector<uint8_t> b;
   put<uint32_t>(b, 12347u | (2u << 16));  // three containers, with the run cookie
   put<uint8_t>(b, 0x1);                    // container 0 holds runs
   for (uint16_t v : {0, 9, 1, 2, 2, 4999}) put<uint16_t>(b, v);  // (key, cardinality - 1)
   for (uint16_t v : {2, 5, 4, 20, 4}) put<uint16_t>(b, v);        // run count, (start, length - 1)
   for (uint16_t v : {7, 8, 9}) put<uint16_t>(b, v);
   for (uint32_t w = 0; w < 1024; ++w) {
      const uint32_t lo = 64 * w;
      const uint64_t word = lo + 64 <= 5000 ? ~0ull : (lo >= 5000 ? 0ull : (1ull << (5000 - lo)) - 1);
      put<uint64_t>(b, word);
   }
   const uint64_t start = 0, len = b.size();
   uint64_t ptr[2] = {};
   CHECK(rdb_archive_bitmap_runs(b.data(), b.size(), &start, &len, 1, ptr, nullptr, nullptr, 0) == RDB_OK);
   CHECK(ptr[1] == 4);
   uint32_t rs[4], re[4];
   CHECK(rdb_archive_bitmap_runs(b.data(), b.size(), &start, &len, 1, ptr, rs, re, 4) == RDB_OK);
   CHECK(rs[0] == 5 && re[0] == 10 && rs[1] == 20 && re[1] == 25 && rs[2] == 65543 && re[2] == 65546 &&
         rs[3] == 131072 && re[3] == 131072 + 5000);
   for (int trial = 0; trial < 400; ++trial) {
      std::vector<uint8_t> bad(b.begin(), b.begin() + (trial % 3 == 0 ? rng.below(static_cast<uint32_t>(b.size())) : b.size()));
      if (trial % 3 != 0) {
         for (uint32_t k = 1 + rng.below(3); k > 0; --k) bad[rng.below(static_cast<uint32_t>(bad.size()))] = static_cast<uint8_t>(rng.next());
      }
      const uint64_t blen = bad.size();
      uint64_t bp[2] = {};
      int rc = rdb_archive_bitmap_runs(bad.data(), blen, &start, &blen, 1, bp, nullptr, nullptr, 0);
      CHECK(rc == RDB_OK || rc == RDB_EINVAL);
      if (rc == RDB_OK) {
         std::vector<uint32_t> a(bp[1] + 1), e(bp[1] + 1);
         CHECK(rdb_archive_bitmap_runs(bad.data(), blen, &start, &blen, 1, bp, a.data(), e.data(), bp[1]) == RDB_OK);
      }
   }
   const uint64_t past = b.size() + 1;
   CHECK(rdb_archive_bitmap_runs(b.data(), b.size(), &start, &past, 1, ptr, nullptr, nullptr, 0) == RDB_EINVAL);
}

void checkSinks(Rng& rng) {
   for (int alphabet : {RDB_NUCLEOTIDE, RDB_AMINO_ACID}) {
      const uint32_t n_sym = alphabet == RDB_NUCLEOTIDE ? 16 : 28;
      const uint64_t n = 1 + rng.below(50);
      std::vector<uint32_t> pos(n), count(n), cov(n);
      std::vector<uint8_t> from(n), to(n);
      std::vector<double> prop(n);
      for (uint64_t i = 0; i < n; ++i) {
         pos[i] = 1 + static_cast<uint32_t>(i) * 3;
         from[i] = static_cast<uint8_t>(rng.below(n_sym));
         to[i] = static_cast<uint8_t>(rng.below(n_sym));
         cov[i] = 1 + rng.below(100000);
         count[i] = rng.below(cov[i] + 1);
         prop[i] = static_cast<double>(count[i]) / cov[i];
      }
      for (uint32_t fields : {127u, 8u, 1u | 64u, 4u | 16u}) {
         uint64_t need = 0;
         CHECK(rdb_mutation_rows_ndjson(alphabet, "seq", fields, pos.data(), from.data(), to.data(), count.data(),
                                        cov.data(), prop.data(), n, nullptr, 0, &need) == RDB_ENOMEM);
         for (uint64_t cap = 0; cap <= need; cap += 1 + need / 9) {
            std::vector<char> out(cap);
            uint64_t len = 0;
            const int rc = rdb_mutation_rows_ndjson(alphabet, "seq", fields, pos.data(), from.data(), to.data(),
                                                    count.data(), cov.data(), prop.data(), n, out.data(), cap, &len);
            CHECK(rc == (cap >= need ? RDB_OK : RDB_ENOMEM) && len == need);
         }
         std::vector<char> text(need);
         uint64_t len = 0;
         CHECK(rdb_mutation_rows_ndjson(alphabet, "seq", fields, pos.data(), from.data(), to.data(), count.data(),
                                        cov.data(), prop.data(), n, text.data(), need, &len) == RDB_OK);
         CHECK(len == need && need > 0 && text[need - 1] == '\n');
         CHECK(rdb_mutation_rows_arrow_ipc(alphabet, "seq", fields, pos.data(), from.data(), to.data(), count.data(),
                                           cov.data(), prop.data(), n, nullptr, 0, &need) == RDB_ENOMEM);
         for (uint64_t cap = 0; cap <= need; cap += 1 + need / 9) {
            std::vector<uint8_t> out(cap);
            const int rc = rdb_mutation_rows_arrow_ipc(alphabet, "seq", fields, pos.data(), from.data(), to.data(),
                                                       count.data(), cov.data(), prop.data(), n, out.data(), cap,
                                                       &len);
            CHECK(rc == (cap >= need ? RDB_OK : RDB_ENOMEM) && len == need);
         }
      }
   }
}

// ---- device mode ----------------------------------------------------------------------------------

struct Rows {
   std::string reference;
   std::vector<std::string> seq;  // aligned strings
   std::vector<uint8_t> is_null;
};

// Rows with scattered mutations, internal N runs, partial coverage and nulls, plus three planted
// positions: one symbol on a tenth of the rows at random (bitset containers in a full chunk), one
// on a contiguous block of rows (run containers), and one on most rows (the local reference flips).
Rows makeRows(Rng& rng, int alphabet, uint32_t L, uint64_t n) {
   const char* base = alphabet == RDB_NUCLEOTIDE ? "ACGT" : "ACDEFGHIKLMNPQRSTVWY";
   const char* any = alphabet == RDB_NUCLEOTIDE ? "ACGT-RYSWKMBDHVN" : "ACDEFGHIKLMNPQRSTVWY*X-";
   const uint32_t n_base = static_cast<uint32_t>(std::strlen(base)), n_any = static_cast<uint32_t>(std::strlen(any));
   const char missing = alphabet == RDB_NUCLEOTIDE ? 'N' : 'X';
   Rows r;
   r.reference.resize(L);
   for (auto& c : r.reference) c = base[rng.below(n_base)];
   r.seq.resize(n);
   r.is_null.assign(n, 0);
   for (uint64_t i = 0; i < n; ++i) {
      if (rng.chance(0.02)) {
         r.is_null[i] = 1;
         continue;
      }
      std::string s = r.reference;
      for (uint32_t p = 0; p < L; ++p)
         if (rng.chance(0.01)) s[p] = any[rng.below(n_any)];
      if (rng.chance(0.1)) s[5] = base[(std::strchr(base, r.reference[5]) - base + 1) % n_base];
      if (i % 65536 >= 1000 && i % 65536 < 9000) s[7] = base[(std::strchr(base, r.reference[7]) - base + 1) % n_base];
      if (rng.chance(0.7)) s[9] = base[(std::strchr(base, r.reference[9]) - base + 2) % n_base];
      for (int k = rng.below(3); k > 0; --k) {
         const uint32_t a = rng.below(L), b = std::min(L, a + 1 + rng.below(12));
         for (uint32_t p = a; p < b; ++p) s[p] = missing;
      }
      const uint32_t head = rng.below(15), tail = rng.below(15);
      for (uint32_t p = 0; p < head; ++p) s[p] = missing;
      for (uint32_t p = L - tail; p < L; ++p) s[p] = missing;
      r.seq[i] = std::move(s);
   }
   return r;
}

std::vector<uint8_t> symbols(int alphabet, const std::string& s) {
   const char* order = alphabet == RDB_NUCLEOTIDE ? "-ACGTRYSWKMBDHVN" : "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X";
   std::vector<uint8_t> out(s.size());
   for (size_t i = 0; i < s.size(); ++i) out[i] = static_cast<uint8_t>(std::strchr(order, s[i]) - order);
   return out;
}

bool addByBuilder(rdb_table* t, const char* name, int alphabet, const Rows& r) {
   const auto ref = symbols(alphabet, r.reference);
   rdb_seq_builder* b = nullptr;
   if (rdb_seq_builder_create(t, name, alphabet, ref.data(), static_cast<uint32_t>(ref.size()), &b) != RDB_OK) return false;
   const uint64_t n = r.seq.size(), half = n / 3;
   bool ok = true;
   for (auto [lo, hi] : {std::pair<uint64_t, uint64_t>{0, half}, {half, n}}) {
      std::string data;
      std::vector<uint64_t> ptr{0};
      for (uint64_t i = lo; i < hi; ++i) {
         data += r.seq[i];
         ptr.push_back(data.size());
      }
      ok = ok && rdb_seq_builder_append(b, data.data(), ptr.data(), nullptr, r.is_null.data() + lo, hi - lo) == RDB_OK;
   }
   ok = ok && rdb_seq_builder_finish(b) == RDB_OK;
   rdb_seq_builder_free(b);
   return ok;
}

bool addByDiff(rdb_table* t, const char* name, int alphabet, const Rows& r) {
   const auto ref = symbols(alphabet, r.reference);
   const uint32_t L = static_cast<uint32_t>(ref.size());
   const uint64_t n = r.seq.size();
   std::vector<uint32_t> cs(n, 0), ce(n, 0), mpos, ns, ne, miss(L), mp(L);
   std::vector<uint8_t> msym, ms(L);
   std::vector<uint64_t> mptr{0}, nptr{0};
   for (uint64_t i = 0; i < n; ++i) {
      if (!r.is_null[i]) {
         uint64_t nm = 0, nmut = 0;
         if (rdb_diff_sequence(alphabet, r.seq[i].data(), L, r.reference.data(), &cs[i], &ce[i], miss.data(), L, &nm,
                               mp.data(), ms.data(), L, &nmut) != RDB_OK) {
            return false;
         }
         mpos.insert(mpos.end(), mp.begin(), mp.begin() + nmut);
         msym.insert(msym.end(), ms.begin(), ms.begin() + nmut);
         for (uint64_t k = 0; k < nm; ++k) {
            if (k > 0 && miss[k] == ne.back()) {
               ne.back() = miss[k] + 1;
            } else {
               ns.push_back(miss[k]);
               ne.push_back(miss[k] + 1);
            }
         }
      }
      mptr.push_back(mpos.size());
      nptr.push_back(ns.size());
   }
   mpos.push_back(0), msym.push_back(0), ns.push_back(0), ne.push_back(0);  // never empty arrays
   return rdb_table_add_sequence_column(t, name, alphabet, ref.data(), L, r.is_null.data(), cs.data(), ce.data(),
                                        mptr.data(), mpos.data(), msym.data(), nptr.data(), ns.data(),
                                        ne.data()) == RDB_OK;
}

struct Meta {
   std::vector<int32_t> date, sorted_date, age;
   std::vector<uint8_t> date_null, no_null, age_null;
   std::vector<uint32_t> country, lineage;
   std::vector<uint8_t> country_null;
};

Meta makeMeta(Rng& rng, uint64_t n) {
   Meta m;
   int32_t d = 18000;
   for (uint64_t i = 0; i < n; ++i) {
      m.date.push_back(18000 + static_cast<int32_t>(rng.below(900)));
      m.date_null.push_back(rng.chance(0.05));
      d += rng.chance(0.01) ? 1 : 0;
      m.sorted_date.push_back(d);
      m.no_null.push_back(0);
      m.age.push_back(static_cast<int32_t>(rng.below(100)));
      m.age_null.push_back(rng.chance(0.1));
      m.country.push_back(rng.below(5));
      m.country_null.push_back(rng.chance(0.03));
      m.lineage.push_back(rng.below(6));
   }
   return m;
}

bool addMeta(rdb_table* t, const Meta& m) {
   const char* countries = "ChinaGermanyIndiaPeruSwitzerland";
   const uint32_t country_off[] = {0, 5, 12, 17, 21, 32};
   const char* lineages = "BB.1B.1.1XB.1.1.7Q";  // Q is an alias of B.1.1.7
   const uint32_t lineage_off[] = {0, 1, 4, 9, 10, 17, 18};
   const int32_t parent[] = {-1, 0, 1, -1, 2, -1};
   const int32_t alias[] = {-1, -1, -1, -1, -1, 4};
   return rdb_table_add_int_column(t, "date", 1, m.date.data(), m.date_null.data()) == RDB_OK &&
          rdb_table_add_int_column(t, "sorted_date", 1, m.sorted_date.data(), m.no_null.data()) == RDB_OK &&
          rdb_table_add_int_column(t, "age", 0, m.age.data(), m.age_null.data()) == RDB_OK &&
          rdb_table_add_dict_column(t, "country", m.country.data(), m.country_null.data(), 5, countries, country_off,
                                    nullptr) == RDB_OK &&
          rdb_table_add_lineage_column(t, "lineage", m.lineage.data(), m.no_null.data(), 6, lineages, lineage_off,
                                       parent, alias) == RDB_OK;
}

const char* kFilters[] = {
   "(true)",
   "(false)",
   "(sym_eq main 5 C)",
   "(has_mut main 7)",
   "(has_mut main 9)",
   "(covered main 3)",
   "(sym_eq main 0 N)",
   "(sym_in gene 4 \"AX\")",
   "(and (date_between date 18100 18400) (not (str_eq country \"Peru\")))",
   "(date_between sorted_date 18100 _)",
   "(date_between sorted_date _ 18050)",
   "(int_between age 30 60)",
   "(lineage lineage \"B\" 1)",
   "(lineage lineage \"Q\" 0)",
   "(or (str_eq country \"China\") (has_mut gene 11))",
   "(nof 2 0 (has_mut main 5) (has_mut main 7) (lineage lineage \"B.1\" 1) (covered gene 2))",
   "(nof 1 1 (sym_eq main 9 A) (sym_eq main 9 C))",
   "(bitmap 0)",
};

struct Result {
   std::vector<uint32_t> pos, count, cov;
   std::vector<uint8_t> from, to;
   std::vector<double> prop;
   uint64_t n = 0;
   bool operator==(const Result& o) const {
      return n == o.n && pos == o.pos && count == o.count && cov == o.cov && from == o.from && to == o.to &&
             prop == o.prop;
   }
   void resize(uint64_t k) {
      pos.resize(k), count.resize(k), cov.resize(k), from.resize(k), to.resize(k), prop.resize(k);
   }
};

Result mutationsQuery(rdb_table* t, const char* f, const char* col, double p, bool noalloc = false) {
   Result r;
   r.resize(1 << 14);
   const auto call = noalloc ? rdb_mutations_query_noalloc : rdb_mutations_query;
   CHECK(call(t, f, col, p, r.pos.data(), r.from.data(), r.to.data(), r.count.data(), r.cov.data(), r.prop.data(),
              r.pos.size(), &r.n) == RDB_OK);
   r.resize(r.n);
   return r;
}

std::vector<uint32_t> filterRows(rdb_table* t, const char* f) {
   rdb_bitmap* bm = nullptr;
   uint64_t card = 0, n = 0;
   if (rdb_filter(t, f, &bm, &card) != RDB_OK) return {UINT32_MAX};
   std::vector<uint32_t> rows(card + 1);
   CHECK(rdb_bitmap_rows(bm, rows.data(), rows.size(), &n) == RDB_OK && n == card);
   rows.resize(n);
   uint64_t count = 0;
   CHECK(rdb_count(t, f, &count) == RDB_OK && count == card);
   char text[4096];
   uint64_t len = 0;
   CHECK(rdb_filter_explain(t, f, text, sizeof text, &len) == RDB_OK && len < sizeof text);
   CHECK(rdb_filter_explain(t, f, text, len, &len) == RDB_ENOMEM);
   for (const char* col : {"main", "gene"}) {
      Result a;
      a.resize(1 << 14);
      CHECK(rdb_mutations(t, bm, col, 0.05, a.pos.data(), a.from.data(), a.to.data(), a.count.data(), a.cov.data(),
                          a.prop.data(), a.pos.size(), &a.n) == RDB_OK);
      a.resize(a.n);
      CHECK(a == mutationsQuery(t, f, col, 0.05));
      CHECK(a == mutationsQuery(t, f, col, 0.05, true));
      if (a.n > 1) {  // a too-small output is an error, not an overrun
         Result s;
         s.resize(a.n - 1);
         CHECK(rdb_mutations(t, bm, col, 0.05, s.pos.data(), s.from.data(), s.to.data(), s.count.data(),
                             s.cov.data(), s.prop.data(), s.pos.size(), &s.n) == RDB_ENOMEM);
      }
      uint64_t need = 0, got = 0;
      CHECK(rdb_mutations_ndjson(t, bm, col, 0.05, RDB_FIELDS_ALL, nullptr, 0, &need) ==
            (a.n ? RDB_ENOMEM : RDB_OK));
      std::vector<char> text2(need + 1);
      CHECK(rdb_mutations_ndjson(t, bm, col, 0.05, RDB_FIELDS_ALL, text2.data(), text2.size(), &got) == RDB_OK &&
            got == need);
      CHECK(rdb_mutations_arrow_ipc(t, bm, col, 0.05, RDB_FIELDS_ALL, nullptr, 0, &need) == RDB_ENOMEM);
      std::vector<uint8_t> ipc(need);
      CHECK(rdb_mutations_arrow_ipc(t, bm, col, 0.05, RDB_FIELDS_ALL, ipc.data(), ipc.size(), &got) == RDB_OK &&
            got == need);
      std::vector<uint32_t> counts((std::string(col) == "main" ? 16 : 28) * 300);
      CHECK(rdb_mutation_counts(t, bm, col, counts.data()) == RDB_OK);
   }
   std::vector<uint32_t> ids(3 * 4096);
   std::vector<uint64_t> counts(4096);
   uint64_t groups = 0;
   CHECK(rdb_aggregate(t, bm, "country,lineage,main@9", ids.data(), counts.data(), counts.size(), &groups) == RDB_OK);
   uint64_t total = 0;
   for (uint64_t g = 0; g < groups; ++g) total += counts[g];
   CHECK(total == card);
   if (groups > 1) {
      CHECK(rdb_aggregate(t, bm, "country,lineage,main@9", ids.data(), counts.data(), groups - 1, &groups) ==
            RDB_ENOMEM);
   }
   rdb_bitmap_free(bm);
   return rows;
}

void checkBatch(rdb_ctx* ctx, rdb_table* t) {
   std::vector<rdb_query> qs;
   for (const char* f : kFilters) {
      qs.push_back({f, RDB_ACTION_MUTATIONS, "main", 0.02, nullptr});
      qs.push_back({f, RDB_ACTION_MUTATIONS, "gene", 0.0, nullptr});
      qs.push_back({f, RDB_ACTION_COUNT, nullptr, 0.0, nullptr});
      qs.push_back({f, RDB_ACTION_AGGREGATE, nullptr, 0.0, "lineage,country"});
   }
   qs.push_back({"(has_mut nosuch 1)", RDB_ACTION_MUTATIONS, "main", 0.02, nullptr});
   qs.push_back({"(and", RDB_ACTION_COUNT, nullptr, 0.0, nullptr});
   const uint32_t nq = static_cast<uint32_t>(qs.size());
   for (int64_t scan : {0, 1, 2}) {
      for (int64_t max_q : {256, 5}) {
         CHECK(rdb_set_tuning(ctx, "batch_shared_scan", scan) == RDB_OK);
         CHECK(rdb_set_tuning(ctx, "batch_max_queries", max_q) == RDB_OK);
         rdb_batch* b = nullptr;
         CHECK(rdb_batch_submit(t, qs.data(), nq, &b) == RDB_OK);
         CHECK(rdb_batch_wait(b) == RDB_OK);
         std::vector<int> status(nq);
         std::vector<uint64_t> card(nq), size(nq), offsets(nq + 1);
         CHECK(rdb_batch_summary(b, nq, status.data(), card.data(), size.data()) == RDB_OK);
         CHECK(status[nq - 1] != RDB_OK && status[nq - 2] != RDB_OK);
         uint64_t total = 0;
         const int rc = rdb_batch_mutations_all(b, nq, offsets.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                                nullptr, 0, &total);
         CHECK(rc == (total ? RDB_ENOMEM : RDB_OK));
         Result all;
         all.resize(total);
         CHECK(rdb_batch_mutations_all(b, nq, offsets.data(), all.pos.data(), all.from.data(), all.to.data(),
                                       all.count.data(), all.cov.data(), all.prop.data(), total, &all.n) == RDB_OK);
         CHECK(all.n == total && offsets[nq] == total);
         for (uint32_t q = 0; q + 2 < nq; ++q) {
            CHECK(status[q] == RDB_OK);
            uint64_t count = 0;
            CHECK(rdb_batch_count(b, q, &count) == RDB_OK);
            CHECK(rdb_count(t, qs[q].filter, &count) == RDB_OK && count == card[q]);
            if (qs[q].action == RDB_ACTION_MUTATIONS) {
               Result r;
               r.resize(size[q]);
               CHECK(rdb_batch_mutations(b, q, r.pos.data(), r.from.data(), r.to.data(), r.count.data(), r.cov.data(),
                                         r.prop.data(), size[q], &r.n) == RDB_OK);
               CHECK(r == mutationsQuery(t, qs[q].filter, qs[q].column, qs[q].min_proportion));
            } else if (qs[q].action == RDB_ACTION_AGGREGATE) {
               std::vector<uint32_t> ids(2 * size[q] + 2);
               std::vector<uint64_t> counts(size[q] + 1);
               uint64_t n = 0;
               CHECK(rdb_batch_aggregate(b, q, ids.data(), counts.data(), counts.size(), &n) == RDB_OK &&
                     n == size[q]);
            }
         }
         rdb_batch_free(b);
      }
   }
   // a submission whose second job fails leaves the lanes usable
   CHECK(rdb_set_tuning(ctx, "batch_max_queries", 5) == RDB_OK);
   CHECK(rdb_set_tuning(ctx, "batch_fail_job", 2) == RDB_OK);
   rdb_batch* b = nullptr;
   CHECK(rdb_batch_submit(t, qs.data(), nq, &b) == RDB_ENOMEM);
   CHECK(rdb_batch_submit(t, qs.data(), nq, &b) == RDB_OK);
   CHECK(rdb_batch_wait(b) == RDB_OK);
   rdb_batch_free(b);
   CHECK(rdb_set_tuning(ctx, "batch_max_queries", 256) == RDB_OK);
   CHECK(rdb_set_tuning(ctx, "batch_shared_scan", 0) == RDB_OK);
}

void checkMalformed(rdb_ctx* ctx) {
   const uint32_t sizes[] = {10};
   rdb_table* t = nullptr;
   CHECK(rdb_table_create(ctx, sizes, 1, &t) == RDB_OK);
   const uint8_t ref[4] = {1, 2, 3, 4};
   const uint8_t nulls[10] = {};
   uint32_t cs[10] = {}, ce[10], mpos[1] = {7}, ns[1] = {0}, ne[1] = {0};
   for (auto& e : ce) e = 4;
   const uint8_t msym[1] = {2};
   uint64_t mptr[11] = {0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1}, nptr[11] = {};
   // a mutation past the genome
   CHECK(rdb_table_add_sequence_column(t, "s", RDB_NUCLEOTIDE, ref, 4, nulls, cs, ce, mptr, mpos, msym, nptr, ns,
                                       ne) == RDB_EINVAL);
   // a finalized container whose payload lies past the buffer, and one holding a row past the chunk
   const uint32_t kp[1] = {1}, card[1] = {3};
   const uint16_t kc[1] = {0};
   const uint8_t ks[1] = {3}, tc[1] = {2};
   const uint64_t off[1] = {2}, miss_ptr[11] = {};
   const uint16_t rows[3] = {1, 2, 11};
   CHECK(rdb_table_add_sequence_column_finalized(t, "f", RDB_NUCLEOTIDE, ref, ref, 4, 1, kp, kc, ks, tc, card, off,
                                                 reinterpret_cast<const uint8_t*>(rows), 6, nulls, cs, ce, miss_ptr,
                                                 ns) == RDB_EINVAL);
   const uint64_t off0[1] = {0};
   CHECK(rdb_table_add_sequence_column_finalized(t, "f", RDB_NUCLEOTIDE, ref, ref, 4, 1, kp, kc, ks, tc, card, off0,
                                                 reinterpret_cast<const uint8_t*>(rows), 6, nulls, cs, ce, miss_ptr,
                                                 ns) == RDB_EINVAL);
   // offsets whose end wraps past 2^64, for each container type
   for (uint8_t type : {1, 2, 3}) {
      const uint8_t tcw[1] = {type};
      const uint64_t wrap[1] = {~0ull - 3};
      CHECK(rdb_table_add_sequence_column_finalized(t, "f", RDB_NUCLEOTIDE, ref, ref, 4, 1, kp, kc, ks, tcw, card,
                                                    wrap, reinterpret_cast<const uint8_t*>(rows), 6, nulls, cs, ce,
                                                    miss_ptr, ns) == RDB_EINVAL);
   }
   // a builder row with an illegal character adds nothing
   rdb_seq_builder* b = nullptr;
   CHECK(rdb_seq_builder_create(t, "g", RDB_NUCLEOTIDE, ref, 4, &b) == RDB_OK);
   const uint64_t ptr[2] = {0, 4};
   CHECK(rdb_seq_builder_append(b, "ACJT", ptr, nullptr, nullptr, 1) == RDB_EINVAL);
   const uint32_t offsets[1] = {3};
   CHECK(rdb_seq_builder_append(b, "ACGT", ptr, offsets, nullptr, 1) == RDB_EINVAL);
   rdb_seq_builder_free(b);
   const uint32_t ids[10] = {0, 1, 2, 0, 0, 0, 0, 0, 0, 9};
   CHECK(rdb_table_add_dict_column(t, "d", ids, nulls, 3, "abc", std::array<uint32_t, 4>{0, 1, 2, 3}.data(),
                                   nullptr) == RDB_EINVAL);
   const uint32_t far[1] = {70000};
   uint32_t id = 0;
   CHECK(rdb_table_add_bitmap(t, far, 1, &id) == RDB_EINVAL);
   CHECK(rdb_table_finalize(t) == RDB_OK);
   uint64_t count = 0;
   for (const char* f : {"(and", "(has_mut s 1)", "(sym_eq nosuch 1 A)", "(str_eq d \"x\")", "(date_between date 1 2)",
                         "(bitmap 5)", ")", "", "(lineage x \"B\" 0)"}) {
      const int rc = rdb_count(t, f, &count);
      CHECK(rc != RDB_OK);
      CHECK(std::strlen(rdb_last_error(ctx)) > 0);
      if (rc == RDB_OK) std::fprintf(stderr, "  accepted: %s\n", f);
   }
   // more matchers than children is not an error: nothing matches (nof.cpp's rewrite to False)
   CHECK(rdb_count(t, "(nof 3 0 (true))", &count) == RDB_OK && count == 0);
   rdb_table_free(t);
}

void checkPlanner(rdb_ctx* ctx, rdb_table* t) {
   rdb_mlp* mlp = nullptr;
   const uint32_t n_in = 32;
   CHECK(rdb_mlp_create(ctx, n_in, 32, 2, 7, &mlp) == RDB_OK);
   uint64_t np = 0;
   CHECK(rdb_mlp_parameter_count(mlp, &np) == RDB_OK && np > 0);
   std::vector<float> p(np);
   CHECK(rdb_mlp_get_parameters(mlp, p.data(), np) == RDB_OK);
   CHECK(rdb_mlp_set_parameters(mlp, p.data(), np) == RDB_OK);
   const uint32_t nf = static_cast<uint32_t>(std::size(kFilters));
   uint64_t ns = 0;
   CHECK(rdb_plan_samples(t, kFilters, nf, nullptr, nullptr, nullptr, 0, &ns) != RDB_OK || ns > 0);
   std::vector<float> x(ns * n_in), y(ns), est(ns), ind(ns), g(np);
   std::vector<uint64_t> exact(ns);
   std::vector<uint32_t> tree(ns);
   uint64_t got = 0;
   CHECK(rdb_plan_samples(t, kFilters, nf, x.data(), y.data(), tree.data(), ns, &got) == RDB_OK && got == ns);
   float loss = 0;
   CHECK(rdb_mlp_gradients(mlp, x.data(), y.data(), static_cast<uint32_t>(ns), &loss, g.data()) == RDB_OK);
   CHECK(rdb_mlp_train(mlp, x.data(), y.data(), ns, 3, 16, 1e-3f, 5, &loss) == RDB_OK);
   CHECK(rdb_mlp_forward(mlp, x.data(), static_cast<uint32_t>(ns), y.data()) == RDB_OK);
   CHECK(rdb_plan_estimates(t, mlp, kFilters, nf, est.data(), ind.data(), exact.data(), tree.data(), ns, &got) ==
         RDB_OK && got == ns);
   std::vector<rdb_query> qs;
   for (const char* f : kFilters) qs.push_back({f, RDB_ACTION_COUNT, nullptr, 0.0, nullptr});
   for (int mode : {1, 2}) {
      CHECK(rdb_set_planner(ctx, mode, mode == 2 ? mlp : nullptr) == RDB_OK);
      rdb_batch* b = nullptr;
      CHECK(rdb_batch_submit(t, qs.data(), nf, &b) == RDB_OK);
      for (uint32_t q = 0; q < nf; ++q) {
         uint64_t a = 0, c = 1;
         CHECK(rdb_batch_count(b, q, &a) == RDB_OK && rdb_count(t, kFilters[q], &c) == RDB_OK && a == c);
      }
      rdb_batch_free(b);
   }
   CHECK(rdb_set_planner(ctx, 0, nullptr) == RDB_OK);
   rdb_mlp_free(mlp);
}

int deviceMode() {
   rdb_ctx* ctx = nullptr;
   if (rdb_open(0, &ctx) != RDB_OK) {
      std::fprintf(stderr, "rdb_open failed: %s\n", rdb_last_error(ctx));
      return 1;
   }
   Rng rng{31};
   const std::vector<uint32_t> chunks = {65536, 3000, 1, 4097};
   uint64_t n = 0;
   for (uint32_t c : chunks) n += c;
   const Rows main = makeRows(rng, RDB_NUCLEOTIDE, 300, n);
   const Rows gene = makeRows(rng, RDB_AMINO_ACID, 120, n);
   const Meta meta = makeMeta(rng, n);
   std::vector<uint32_t> literal;
   for (uint32_t c = 0, base = 0; c < chunks.size(); base += 65536, ++c)
      for (uint32_t r = 0; r < chunks[c]; r += 1 + rng.below(40)) literal.push_back(base + r);
   rdb_table* tables[2] = {};
   for (int k = 0; k < 2; ++k) {
      CHECK(rdb_table_create(ctx, chunks.data(), static_cast<uint32_t>(chunks.size()), &tables[k]) == RDB_OK);
      const bool ok = k == 0 ? addByBuilder(tables[k], "main", RDB_NUCLEOTIDE, main) &&
                                   addByBuilder(tables[k], "gene", RDB_AMINO_ACID, gene)
                             : addByDiff(tables[k], "main", RDB_NUCLEOTIDE, main) &&
                                   addByDiff(tables[k], "gene", RDB_AMINO_ACID, gene);
      CHECK(ok);
      CHECK(addMeta(tables[k], meta));
      uint32_t id = 9;
      CHECK(rdb_table_add_bitmap(tables[k], literal.data(), literal.size(), &id) == RDB_OK && id == 0);
      CHECK(rdb_table_finalize(tables[k]) == RDB_OK);
   }
   for (const char* col : {"main", "gene"}) {
      uint64_t da = 0, db = 1;
      CHECK(rdb_table_column_digest(tables[0], col, &da) == RDB_OK);
      CHECK(rdb_table_column_digest(tables[1], col, &db) == RDB_OK);
      CHECK(da == db);
      std::vector<uint8_t> la(300), lb(300);
      CHECK(rdb_table_local_reference(tables[0], col, la.data()) == RDB_OK);
      CHECK(rdb_table_local_reference(tables[1], col, lb.data()) == RDB_OK);
      CHECK(la == lb);
      uint64_t bytes = 0, containers = 0, payload = 0;
      CHECK(rdb_table_stats(tables[0], col, &bytes, &containers, &payload) == RDB_OK && containers > 0);
   }
   for (const char* f : kFilters) {
      const auto a = filterRows(tables[0], f), b = filterRows(tables[1], f);
      CHECK(a == b);
      if (a != b) std::fprintf(stderr, "  filter %s\n", f);
   }
   CHECK(filterRows(tables[0], "(true)").size() == n);
   checkBatch(ctx, tables[0]);
   checkPlanner(ctx, tables[0]);
   checkMalformed(ctx);
   rdb_table_free(tables[0]);
   rdb_table_free(tables[1]);
   rdb_close(ctx);
   return 0;
}

}  // namespace

int main(int argc, char** argv) {
   const std::string mode = argc > 1 ? argv[1] : "host";
   Rng rng{7};
   if (mode == "host") {
      checkDiff(rng);
      checkArchiveScanners(rng);
      checkBitmapRuns(rng);
      checkSinks(rng);
   } else if (mode == "device") {
      if (deviceMode() != 0) return 1;
   } else {
      std::fprintf(stderr, "usage: %s host|device\n", argv[0]);
      return 2;
   }
   std::printf("%s checks: %d failed\n", mode.c_str(), failures);
   return failures == 0 ? 0 : 1;
}
