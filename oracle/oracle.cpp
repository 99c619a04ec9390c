// ORACLE — test infrastructure only (the checker and the CPU baseline).  The product never links,
// loads or calls anything in this directory.
//
// A CPU restatement of the reference's query-engine hot path (GenSpectrum/LAPIS-SILO, "RhyDB"
// 0.13.3): reference-diffed sequence storage with local-reference adaptation, the filter-expression
// rewrite/compile rules, the `Operator::evaluate` algebra over roaring-style bitmaps, and the
// Mutations, Aggregated (bitmap group-by) and Count actions.  Each function names the reference
// file:line it follows.  The reference itself cannot be compiled in this image (it needs CRoaring,
// Boost, spdlog, ... via Conan), so this restatement is pinned instead by the reference's own
// known-answer tests (ported as fixtures under tests/golden/) and by brute force over materialized
// sequences on small inputs (oracle/brute.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "roaring_lite.h"

namespace oracle {

struct QueryError : std::runtime_error {
   using std::runtime_error::runtime_error;
};

// ------------------------------------------------------------------------------------------------
// Alphabets: common/nucleotide_symbols.h:23-106, nucleotide_symbols.cpp:11-46 (CODES_FOR and the
// derived AMBIGUITY_SYMBOLS), common/aa_symbols.h:20-118, aa_symbols.cpp:13-72.
// ------------------------------------------------------------------------------------------------
struct Alphabet {
   uint32_t count;
   uint8_t missing;
   std::string chars;
   std::vector<uint8_t> valid;  // VALID_MUTATION_SYMBOLS
   std::vector<std::vector<uint8_t>> codes_for;
   std::vector<std::vector<uint8_t>> ambiguity;  // {Y : CODES_FOR[S] subset of CODES_FOR[Y]}

   int fromChar(char c) const {
      char u = c;
      if (u >= 'a' && u <= 'z') u = static_cast<char>(u - 'a' + 'A');
      if (count == 16 && u == 'U') u = 'T';
      const auto p = chars.find(u);
      if (p == std::string::npos) return -1;
      return static_cast<int>(p);
   }

   void deriveAmbiguity() {
      ambiguity.assign(count, {});
      for (uint8_t s = 0; s < count; ++s) {
         for (uint8_t cand = 0; cand < count; ++cand) {
            bool sup = true;
            for (uint8_t x : codes_for[s]) {
               if (std::find(codes_for[cand].begin(), codes_for[cand].end(), x) == codes_for[cand].end()) {
                  sup = false;
               }
            }
            if (sup) ambiguity[s].push_back(cand);
         }
      }
   }
};

const Alphabet& nucleotide() {
   static const Alphabet a = [] {
      Alphabet n;
      n.count = 16;
      n.chars = "-ACGTRYSWKMBDHVN";
      n.missing = 15;
      n.valid = {0, 1, 2, 3, 4};
      // GAP A C G T R Y S W K M B D H V N
      n.codes_for = {{0},       {1},       {2},       {3},       {4},    {1, 3}, {2, 4}, {3, 2},
                     {1, 4},    {3, 4},    {1, 2},    {2, 3, 4}, {1, 3, 4}, {1, 2, 4}, {1, 2, 3}, {}};
      for (uint8_t s = 0; s < 16; ++s) n.codes_for[15].push_back(s);
      n.deriveAmbiguity();
      return n;
   }();
   return a;
}

const Alphabet& aminoAcid() {
   static const Alphabet a = [] {
      Alphabet n;
      n.count = 28;
      n.chars = "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X";
      n.missing = 27;
      for (uint8_t s = 0; s <= 22; ++s) n.valid.push_back(s);
      n.valid.push_back(26);  // STOP
      n.codes_for.resize(28);
      for (uint8_t s = 0; s < 23; ++s) n.codes_for[s] = {s};
      n.codes_for[23] = {3, 12};   // B = D or N
      n.codes_for[24] = {10, 8};   // J = L or I
      n.codes_for[25] = {15, 4};   // Z = Q or E
      n.codes_for[26] = {26};      // STOP
      for (uint8_t s = 0; s < 28; ++s) n.codes_for[27].push_back(s);
      n.deriveAmbiguity();
      return n;
   }();
   return a;
}

// ------------------------------------------------------------------------------------------------
// Row layout: storage/column/row_layout.{h,cpp} (fullBitmap :9-16, complementInPlace :18-24) and
// row_id.h:16-37 (global id = chunk << 16 | row_in_chunk).
// ------------------------------------------------------------------------------------------------
struct RowLayout {
   std::vector<uint32_t> sizes;
   uint64_t num_rows = 0;

   Bitmap full() const {
      Bitmap b;
      for (size_t c = 0; c < sizes.size(); ++c) {
         if (sizes[c] > 0) b.push(static_cast<uint16_t>(c), rangeC(0, sizes[c]));
      }
      return b;
   }

   Bitmap complement(const Bitmap& in) const {
      Bitmap r = in;
      for (size_t c = 0; c < sizes.size(); ++c) {
         const uint64_t start = static_cast<uint64_t>(c) << 16;
         r = flipB(r, start, start + sizes[c]);
      }
      return r;
   }
};

// ------------------------------------------------------------------------------------------------
// Sequence column: storage/column/sequence_column.{h,cpp}, vertical_sequence_index.{h,cpp},
// horizontal_coverage_index.{h,cpp}.
// ------------------------------------------------------------------------------------------------
struct SeqKey {
   uint32_t pos;
   uint16_t chunk;
   uint8_t sym;
   bool operator<(const SeqKey& o) const {
      if (pos != o.pos) return pos < o.pos;
      if (chunk != o.chunk) return chunk < o.chunk;
      return sym < o.sym;
   }
};

Bitmap bitmapFromRuns(const std::vector<std::pair<uint32_t, uint32_t>>& runs) {
   // runs: sorted disjoint [start, end)
   Bitmap b;
   size_t i = 0;
   while (i < runs.size()) {
      const uint16_t key = static_cast<uint16_t>(runs[i].first >> 16);
      Container c;
      c.type = Container::RUN;
      while (i < runs.size() && (runs[i].first >> 16) == key) {
         const uint32_t s = runs[i].first;
         const uint32_t e = std::min<uint32_t>(runs[i].second, (static_cast<uint32_t>(key) + 1) << 16);
         c.runs.push_back(static_cast<uint16_t>(s & 0xFFFF));
         c.runs.push_back(static_cast<uint16_t>(e - s - 1));
         c.card += e - s;
         if (runs[i].second > e) {
            // a run crossing a 2^16 boundary continues in the next container
            auto rest = runs;
            rest[i].first = e;
            std::vector<std::pair<uint32_t, uint32_t>> tail(rest.begin() + static_cast<ptrdiff_t>(i), rest.end());
            Bitmap head;
            head.push(key, std::move(c));
            Bitmap t = bitmapFromRuns(tail);
            Bitmap out = b;
            for (size_t k = 0; k < head.keys.size(); ++k) out.pushShared(head.keys[k], head.cs[k]);
            for (size_t k = 0; k < t.keys.size(); ++k) out.pushShared(t.keys[k], t.cs[k]);
            return out;
         }
         ++i;
      }
      // keep the smaller of run and array forms
      if (c.card <= kArrayMax && 2 + 2 * static_cast<size_t>(c.card) < 2 + 2 * c.runs.size()) {
         c = c.normalized();
      }
      b.push(key, std::move(c));
   }
   return b;
}

struct SequenceColumn {
   std::string name;
   const Alphabet* alphabet = nullptr;
   uint32_t genome_length = 0;
   std::vector<uint8_t> reference;   // metadata->reference_sequence
   std::vector<uint8_t> local_ref;   // local_reference_sequence_string
   std::map<SeqKey, Container> vertical;
   std::vector<std::vector<std::pair<uint32_t, uint32_t>>> start_end;  // per chunk
   std::vector<std::pair<uint32_t, uint32_t>> batch_start_ends;
   std::map<uint32_t, Bitmap> horizontal;  // row -> internal missing positions
   Bitmap null_bitmap;
   uint32_t sequence_count = 0;

   std::pair<uint32_t, uint32_t> coverageRange(uint32_t g) const {
      return start_end.at(g >> 16).at(g & 0xFFFF);
   }

   // horizontal_coverage_index.h:56-98 (BatchSize = 1)
   Bitmap coverageBitmapAt(uint32_t position) const {
      std::vector<uint32_t> rows;
      for (size_t c = 0; c < start_end.size(); ++c) {
         const auto [bs, be] = batch_start_ends[c];
         if (be <= position || bs >= position + 1) continue;
         const uint32_t base = static_cast<uint32_t>(c) << 16;
         for (size_t r = 0; r < start_end[c].size(); ++r) {
            const auto [s, e] = start_end[c][r];
            if (s <= position && position < e) rows.push_back(base | static_cast<uint32_t>(r));
         }
      }
      Bitmap b = Bitmap::fromSorted(rows.data(), rows.size());
      std::vector<uint32_t> remove;
      for (const auto& [row, bm] : horizontal) {
         if (bm.contains(position)) remove.push_back(row);
      }
      if (!remove.empty()) b = andnotB(b, Bitmap::fromSorted(remove.data(), remove.size()));
      return b;
   }

   // vertical_sequence_index.cpp:205-227 (getMatchingContainerViews) + copy_on_write_bitmap.cpp
   // fromContainerViews :299-350 (OR-merge of containers sharing a key)
   Bitmap matching(uint32_t position, const std::vector<uint8_t>& symbols) const {
      Bitmap b;
      auto it = vertical.lower_bound(SeqKey{position, 0, 0});
      auto end = vertical.lower_bound(SeqKey{position + 1, 0, 0});
      for (; it != end; ++it) {
         if (std::find(symbols.begin(), symbols.end(), it->first.sym) == symbols.end()) continue;
         const uint16_t key = it->first.chunk;
         if (!b.keys.empty() && b.keys.back() == key) {
            b.cs.back() = std::make_shared<const Container>(orC(*b.cs.back(), it->second));
         } else {
            b.keys.push_back(key);
            b.cs.push_back(std::make_shared<const Container>(it->second));
         }
      }
      return b;
   }
};

struct IntColumn {
   std::string name;
   bool is_date = false;
   bool is_sorted = true;  // date32_column.cpp:17-33: non-decreasing and no nulls
   std::vector<int32_t> values;  // dense row order
   Bitmap null_bitmap;
};

struct DictColumn {
   std::string name;
   std::vector<std::string> dictionary;
   std::unordered_map<std::string, uint32_t> ids;
   std::vector<uint32_t> values;  // dense row order
   Bitmap null_bitmap;
   std::map<uint32_t, Bitmap> indexed_values;          // value id -> rows (non-null only)
   std::map<uint32_t, Bitmap> including_sublineages;   // lineage index, value id -> rows
   bool has_lineage_index = false;
};

struct Table {
   RowLayout layout;
   std::vector<uint64_t> chunk_offset;
   std::map<std::string, SequenceColumn> seq;
   std::map<std::string, IntColumn> ints;
   std::map<std::string, DictColumn> dicts;
   std::vector<Bitmap> literals;
   std::string error;

   size_t dense(uint32_t g) const { return chunk_offset[g >> 16] + (g & 0xFFFF); }
   uint32_t global(uint64_t i) const {
      const auto it = std::upper_bound(chunk_offset.begin(), chunk_offset.end(), i) - 1;
      const uint64_t c = static_cast<uint64_t>(it - chunk_offset.begin());
      return static_cast<uint32_t>((c << 16) | (i - *it));
   }
   template <typename F>
   void forEachRow(F&& f) const {
      for (size_t c = 0; c < layout.sizes.size(); ++c) {
         for (uint32_t r = 0; r < layout.sizes[c]; ++r) f((static_cast<uint32_t>(c) << 16) | r);
      }
   }
};

// sequence_column.cpp:99-155 (appendChunk/appendValue/appendNullValue),
// horizontal_coverage_index.cpp:17-46 (insertCoverage), vertical_sequence_index.cpp:16-38
// (addSymbolsToPositions) and sequence_column.cpp:157-212 (finalize: local-reference adaptation
// then runOptimize).
void buildSequenceColumn(
   SequenceColumn& col,
   const Table& t,
   const uint8_t* is_null,
   const uint32_t* cov_start,
   const uint32_t* cov_end,
   const uint64_t* mut_ptr,
   const uint32_t* mut_pos,
   const uint8_t* mut_sym,
   const uint64_t* nrun_ptr,
   const uint32_t* nrun_start,
   const uint32_t* nrun_end
) {
   const uint32_t L = col.genome_length;
   col.local_ref = col.reference;
   std::vector<uint64_t> entries;
   std::vector<uint32_t> null_rows;
   for (size_t c = 0; c < t.layout.sizes.size(); ++c) {
      col.start_end.emplace_back();
      col.batch_start_ends.emplace_back(UINT32_MAX, 0);
      entries.clear();
      const uint64_t off = t.chunk_offset[c];
      for (uint32_t r = 0; r < t.layout.sizes[c]; ++r) {
         const uint64_t i = off + r;
         const uint32_t g = (static_cast<uint32_t>(c) << 16) | r;
         col.sequence_count++;
         uint32_t s = 0, e = 0;
         if (is_null[i]) {
            null_rows.push_back(g);
         } else {
            s = cov_start[i];
            e = cov_end[i];
            for (uint64_t k = mut_ptr[i]; k < mut_ptr[i + 1]; ++k) {
               if (mut_pos[k] >= L || mut_sym[k] >= col.alphabet->count || mut_sym[k] == col.alphabet->missing) {
                  throw QueryError("invalid mutation in ingest data");
               }
               entries.push_back(
                  (static_cast<uint64_t>(mut_pos[k]) << 24) | (static_cast<uint64_t>(mut_sym[k]) << 16) | r
               );
            }
            std::vector<std::pair<uint32_t, uint32_t>> runs;
            for (uint64_t k = nrun_ptr[i]; k < nrun_ptr[i + 1]; ++k) {
               // insertCoverage trims missing positions to [start, end)
               const uint32_t a = std::max(nrun_start[k], s);
               const uint32_t b = std::min(nrun_end[k], e);
               if (a < b) runs.emplace_back(a, b);
            }
            if (!runs.empty()) col.horizontal.emplace(g, bitmapFromRuns(runs));
         }
         col.start_end.back().emplace_back(s, e);
         auto& [bs, be] = col.batch_start_ends.back();
         bs = std::min(bs, s);
         be = std::max(be, e);
      }
      std::sort(entries.begin(), entries.end());
      size_t k = 0;
      std::vector<uint16_t> rows;
      while (k < entries.size()) {
         const uint64_t key = entries[k] >> 16;
         rows.clear();
         while (k < entries.size() && (entries[k] >> 16) == key) {
            rows.push_back(static_cast<uint16_t>(entries[k] & 0xFFFF));
            ++k;
         }
         col.vertical.emplace(
            SeqKey{static_cast<uint32_t>(key >> 8), static_cast<uint16_t>(c), static_cast<uint8_t>(key & 0xFF)},
            Container::fromSorted(rows.data(), rows.size())
         );
      }
   }
   col.null_bitmap = Bitmap::fromSorted(null_rows.data(), null_rows.size());

   // horizontal_coverage_index.cpp:53-79 computeCoverageCardinalities
   std::vector<int64_t> changes(L + 1, 0);
   for (const auto& chunk : col.start_end) {
      for (const auto& [s, e] : chunk) {
         changes[s] += 1;
         changes[e] -= 1;
      }
   }
   for (const auto& [row, bm] : col.horizontal) {
      bm.forEach([&](uint32_t p) {
         changes[p] -= 1;
         changes[p + 1] += 1;
      });
   }
   std::vector<uint64_t> coverage(L);
   int64_t running = 0;
   for (uint32_t p = 0; p < L; ++p) {
      running += changes[p];
      coverage[p] = static_cast<uint32_t>(running);
   }

   const uint32_t nsym = col.alphabet->count;
   for (uint32_t p = 0; p < L; ++p) {
      // vertical_sequence_index.cpp:56-115 computeSymbolCountsForPosition /
      // getSymbolWithHighestCount / findBetterLocalReferenceSymbol
      const uint8_t current = col.local_ref[p];
      std::vector<uint32_t> counts(nsym, 0);
      counts[current] = static_cast<uint32_t>(coverage[p]);
      auto it = col.vertical.lower_bound(SeqKey{p, 0, 0});
      auto end = col.vertical.lower_bound(SeqKey{p + 1, 0, 0});
      for (auto j = it; j != end; ++j) {
         counts[j->first.sym] += j->second.card;
         counts[current] -= j->second.card;
      }
      uint8_t best = current;
      uint32_t best_count = counts[current];
      for (uint8_t s = 0; s < nsym; ++s) {
         if (s == current) continue;
         if (counts[s] > best_count) {
            best = s;
            best_count = counts[s];
         }
      }
      if (best == current) continue;
      // vertical_sequence_index.cpp:118-164 adaptLocalReference
      Bitmap old_ref = col.coverageBitmapAt(p);
      std::vector<uint8_t> all(nsym);
      for (uint8_t s = 0; s < nsym; ++s) all[s] = s;
      old_ref = andnotB(old_ref, col.matching(p, all));
      for (size_t k = 0; k < old_ref.keys.size(); ++k) {
         col.vertical.emplace(SeqKey{p, old_ref.keys[k], current}, *old_ref.cs[k]);
      }
      std::vector<SeqKey> drop;
      for (auto j = col.vertical.lower_bound(SeqKey{p, 0, 0}); j != col.vertical.lower_bound(SeqKey{p + 1, 0, 0}); ++j) {
         if (j->first.sym == best) drop.push_back(j->first);
      }
      for (const auto& key : drop) col.vertical.erase(key);
      col.local_ref[p] = best;
   }
   for (auto& [key, c] : col.vertical) c.runOptimize();
}

// ------------------------------------------------------------------------------------------------
// Predicates: filter/operators/selection.h:24-166 (Predicate, CompareToValueSelection),
// is_in_covered_region.cpp:31-98.
// ------------------------------------------------------------------------------------------------
enum class Comparator { EQUALS, LESS, HIGHER, LESS_OR_EQUALS, HIGHER_OR_EQUALS, NOT_EQUALS };

struct Predicate {
   virtual ~Predicate() = default;
   virtual bool match(uint32_t g) const = 0;
   virtual Bitmap makeBitmap(const Table& t) const {
      std::vector<uint32_t> rows;
      t.forEachRow([&](uint32_t g) {
         if (match(g)) rows.push_back(g);
      });
      return Bitmap::fromSorted(rows.data(), rows.size());
   }
   virtual double estimateSelectivity() const { return 0.5; }
   virtual std::unique_ptr<Predicate> negate() const = 0;
};
using PredicateVector = std::vector<std::unique_ptr<Predicate>>;

struct CompareToValue : Predicate {
   const Table* table;
   const IntColumn* column;
   Comparator cmp;
   int32_t value;
   bool with_nulls;
   CompareToValue(const Table* t, const IntColumn* c, Comparator k, int32_t v, bool n)
       : table(t), column(c), cmp(k), value(v), with_nulls(n) {}
   bool match(uint32_t g) const override {
      if (column->null_bitmap.contains(g)) return with_nulls;
      const int32_t x = column->values[table->dense(g)];
      switch (cmp) {
         case Comparator::EQUALS: return x == value;
         case Comparator::NOT_EQUALS: return x != value;
         case Comparator::LESS: return x < value;
         case Comparator::HIGHER_OR_EQUALS: return x >= value;
         case Comparator::HIGHER: return x > value;
         case Comparator::LESS_OR_EQUALS: return x <= value;
      }
      return false;
   }
   std::unique_ptr<Predicate> negate() const override {
      Comparator n = cmp;
      switch (cmp) {
         case Comparator::EQUALS: n = Comparator::NOT_EQUALS; break;
         case Comparator::NOT_EQUALS: n = Comparator::EQUALS; break;
         case Comparator::LESS: n = Comparator::HIGHER_OR_EQUALS; break;
         case Comparator::HIGHER_OR_EQUALS: n = Comparator::LESS; break;
         case Comparator::HIGHER: n = Comparator::LESS_OR_EQUALS; break;
         case Comparator::LESS_OR_EQUALS: n = Comparator::HIGHER; break;
      }
      return std::make_unique<CompareToValue>(table, column, n, value, !with_nulls);
   }
};

struct IsInCoveredRegion : Predicate {
   const Table* table;
   const SequenceColumn* column;
   uint32_t position;
   bool covered;
   IsInCoveredRegion(const Table* t, const SequenceColumn* c, uint32_t p, bool cov)
       : table(t), column(c), position(p), covered(cov) {}
   bool isCovered(uint32_t g) const {
      const auto [s, e] = column->coverageRange(g);
      if (position < s || position >= e) return false;
      auto it = column->horizontal.find(g);
      if (it != column->horizontal.end()) return !it->second.contains(position);
      return true;
   }
   bool match(uint32_t g) const override { return isCovered(g) == covered; }
   Bitmap makeBitmap(const Table& t) const override {
      Bitmap b = column->coverageBitmapAt(position);
      if (!covered) return t.layout.complement(b);
      return b;
   }
   double estimateSelectivity() const override { return 0.1; }
   std::unique_ptr<Predicate> negate() const override {
      return std::make_unique<IsInCoveredRegion>(table, column, position, !covered);
   }
};

// ------------------------------------------------------------------------------------------------
// Operators: filter/operators/*.cpp.
// ------------------------------------------------------------------------------------------------
enum class OpType { EMPTY, FULL, INDEX_SCAN, INTERSECTION, COMPLEMENT, RANGE_SELECTION, SELECTION, THRESHOLD, UNION };

struct Operator;
using OperatorPtr = std::unique_ptr<Operator>;
using OperatorVector = std::vector<OperatorPtr>;

struct Operator {
   const Table* table = nullptr;
   explicit Operator(const Table* t) : table(t) {}
   virtual ~Operator() = default;
   virtual OpType type() const = 0;
   virtual Bitmap evaluate() const = 0;
   static OperatorPtr negate(OperatorPtr op);
};

struct Empty : Operator {
   using Operator::Operator;
   OpType type() const override { return OpType::EMPTY; }
   Bitmap evaluate() const override { return {}; }  // empty.cpp:25-27
};

struct Full : Operator {
   using Operator::Operator;
   OpType type() const override { return OpType::FULL; }
   Bitmap evaluate() const override { return table->layout.full(); }  // full.cpp:26-29
};

struct IndexScan : Operator {
   Bitmap bitmap;
   IndexScan(const Table* t, Bitmap b) : Operator(t), bitmap(std::move(b)) {}
   OpType type() const override { return OpType::INDEX_SCAN; }
   Bitmap evaluate() const override { return bitmap; }  // index_scan.cpp:43-46
};

struct Complement : Operator {
   OperatorPtr child;
   Complement(const Table* t, OperatorPtr c) : Operator(t), child(std::move(c)) {}
   OpType type() const override { return OpType::COMPLEMENT; }
   // complement.cpp:51-56
   Bitmap evaluate() const override { return table->layout.complement(child->evaluate()); }
};

struct Intersection : Operator {
   OperatorVector children, negated;
   Intersection(const Table* t, OperatorVector c, OperatorVector n)
       : Operator(t), children(std::move(c)), negated(std::move(n)) {
      // intersection.cpp:18-41
      if (children.empty()) {
         throw QueryError("Compilation bug: Intersection without non-negated children is not allowed.");
      }
      if (children.size() + negated.size() < 2) {
         throw QueryError("Compilation bug: Intersection needs at least two children.");
      }
   }
   OpType type() const override { return OpType::INTERSECTION; }
   // intersection.cpp:59-106
   Bitmap evaluate() const override {
      std::vector<Bitmap> pos, neg;
      for (const auto& c : children) pos.push_back(c->evaluate());
      for (const auto& c : negated) neg.push_back(c->evaluate());
      std::sort(pos.begin(), pos.end(), [](const Bitmap& a, const Bitmap& b) {
         return a.cardinality() < b.cardinality();
      });
      std::sort(neg.begin(), neg.end(), [](const Bitmap& a, const Bitmap& b) {
         return a.cardinality() > b.cardinality();
      });
      Bitmap result = pos[0];
      for (size_t i = 1; i < pos.size(); ++i) result = andB(result, pos[i]);
      for (const auto& n : neg) result = andnotB(result, n);
      return result;
   }
};

struct Union : Operator {
   OperatorVector children;
   Union(const Table* t, OperatorVector c) : Operator(t), children(std::move(c)) {}
   OpType type() const override { return OpType::UNION; }
   // union.cpp:34-42 + copy_on_write_bitmap.cpp:286-297 (fastUnion: sequential OR fold)
   Bitmap evaluate() const override {
      Bitmap result;
      for (const auto& c : children) result = orB(result, c->evaluate());
      return result;
   }
};

struct RangeSelection : Operator {
   std::vector<std::pair<uint32_t, uint32_t>> ranges;  // global [start, end) as (chunk<<16|row)
   RangeSelection(const Table* t, std::vector<std::pair<uint32_t, uint32_t>> r) : Operator(t), ranges(std::move(r)) {}
   OpType type() const override { return OpType::RANGE_SELECTION; }
   // range_selection.cpp:54-87
   Bitmap evaluate() const override {
      std::vector<uint32_t> rows;
      const auto& sizes = table->layout.sizes;
      for (const auto& [start, end] : ranges) {
         const uint32_t sc = start >> 16, ec = end >> 16;
         auto add = [&](uint64_t a, uint64_t b) {
            for (uint64_t x = a; x < b; ++x) rows.push_back(static_cast<uint32_t>(x));
         };
         if (sc == ec) {
            add(start, end);
         } else {
            const uint64_t end_of_start = (static_cast<uint64_t>(sc) << 16) + sizes[sc];
            if (start != end_of_start) add(start, end_of_start);
            for (uint32_t c = sc + 1; c < ec; ++c) add(static_cast<uint64_t>(c) << 16, (static_cast<uint64_t>(c) << 16) + sizes[c]);
            const uint64_t start_of_end = static_cast<uint64_t>(ec) << 16;
            if (end != start_of_end) add(start_of_end, end);
         }
      }
      std::sort(rows.begin(), rows.end());
      rows.erase(std::unique(rows.begin(), rows.end()), rows.end());  // addRange is idempotent
      return Bitmap::fromSorted(rows.data(), rows.size());
   }
};

struct Selection : Operator {
   std::optional<OperatorPtr> child;
   PredicateVector predicates;
   Selection(const Table* t, std::optional<OperatorPtr> c, PredicateVector p)
       : Operator(t), child(std::move(c)), predicates(std::move(p)) {
      // selection.cpp:17-29: most selective first
      std::stable_sort(predicates.begin(), predicates.end(), [](const auto& a, const auto& b) {
         return a->estimateSelectivity() < b->estimateSelectivity();
      });
   }
   OpType type() const override { return OpType::SELECTION; }
   // selection.cpp:94-141
   Bitmap evaluate() const override {
      Bitmap candidates;
      if (child.has_value()) {
         Bitmap cb = (*child)->evaluate();
         if (cb.cardinality() <= table->layout.num_rows / 10) {
            std::vector<uint32_t> rows;
            cb.forEach([&](uint32_t g) {
               for (const auto& p : predicates) {
                  if (!p->match(g)) return;
               }
               rows.push_back(g);
            });
            return Bitmap::fromSorted(rows.data(), rows.size());
         }
         candidates = andB(cb, predicates.front()->makeBitmap(*table));
      } else {
         candidates = predicates.front()->makeBitmap(*table);
      }
      if (predicates.size() == 1) return candidates;
      std::vector<uint32_t> rows;
      candidates.forEach([&](uint32_t g) {
         for (size_t i = 1; i < predicates.size(); ++i) {
            if (!predicates[i]->match(g)) return;
         }
         rows.push_back(g);
      });
      return Bitmap::fromSorted(rows.data(), rows.size());
   }
};

struct Threshold : Operator {
   OperatorVector non_negated, negated;
   uint32_t n;
   bool exact;
   Threshold(const Table* t, OperatorVector a, OperatorVector b, uint32_t num, bool ex)
       : Operator(t), non_negated(std::move(a)), negated(std::move(b)), n(num), exact(ex) {
      // threshold.cpp:18-41
      if (n >= non_negated.size() + negated.size()) {
         throw QueryError("Compilation Error: number_of_matchers must be less than the number of children of a threshold expression");
      }
      if (n == 0) throw QueryError("Compilation Error: number_of_matchers must be greater than zero");
   }
   OpType type() const override { return OpType::THRESHOLD; }
   // threshold.cpp:64-138: DP over bitmaps with pruning bounds
   Bitmap evaluate() const override {
      const uint32_t dp_size = exact ? n + 1 : n;
      std::vector<Bitmap> bm(dp_size);
      bm[0] = non_negated.empty() ? negated[0]->evaluate() : non_negated[0]->evaluate();
      if (non_negated.empty()) bm[0] = table->layout.complement(bm[0]);
      const int max_idx = static_cast<int>(dp_size - 1);
      const int nn = static_cast<int>(non_negated.size());
      const int ng = static_cast<int>(negated.size());
      const int num = static_cast<int>(n);
      const int k = nn + ng;
      for (int i = 1; i < nn; ++i) {
         const Bitmap b = non_negated[i]->evaluate();
         for (int j = std::min(max_idx, i); j > std::max(0, num - k + i - 1); --j) {
            bm[j] = orB(bm[j], andB(bm[j - 1], b));
         }
         if (k - i > num - 1) bm[0] = orB(bm[0], b);
      }
      const int took_first = non_negated.empty() ? 1 : 0;
      for (int li = took_first; li < ng; ++li) {
         Bitmap b = negated[li]->evaluate();
         const int i = li + nn;
         for (int j = std::min(max_idx, i); j > std::max(0, num - k + i - 1); --j) {
            bm[j] = orB(bm[j], andnotB(bm[j - 1], b));
         }
         if (k - i > num - 1) {
            b = table->layout.complement(b);
            bm[0] = orB(bm[0], b);
         }
      }
      if (exact) return andnotB(bm[n - 1], bm[n]);
      return bm.back();
   }
};

// operator.cpp:23-67 and the per-type negate functions
OperatorPtr Operator::negate(OperatorPtr op) {
   const Table* t = op->table;
   switch (op->type()) {
      case OpType::EMPTY:
         return std::make_unique<Full>(t);
      case OpType::FULL:
         return std::make_unique<Empty>(t);
      case OpType::COMPLEMENT:
         return std::move(static_cast<Complement*>(op.get())->child);
      case OpType::RANGE_SELECTION: {
         // range_selection.cpp:89-115
         auto* rs = static_cast<RangeSelection*>(op.get());
         std::vector<std::pair<uint32_t, uint32_t>> nr;
         if (t->layout.sizes.empty()) return std::make_unique<RangeSelection>(t, nr);
         uint32_t last_end = 0;
         for (const auto& [s, e] : rs->ranges) {
            if (last_end != s) nr.emplace_back(last_end, s);
            last_end = e;
         }
         const uint32_t ranges_end = static_cast<uint32_t>(t->layout.sizes.size()) << 16;
         if (last_end != ranges_end) nr.emplace_back(last_end, ranges_end);
         return std::make_unique<RangeSelection>(t, nr);
      }
      case OpType::SELECTION: {
         // selection.cpp:143-151
         auto* sel = static_cast<Selection*>(op.get());
         if (!sel->child.has_value() && sel->predicates.size() == 1) {
            PredicateVector p;
            p.push_back(sel->predicates[0]->negate());
            return std::make_unique<Selection>(t, std::nullopt, std::move(p));
         }
         return std::make_unique<Complement>(t, std::move(op));
      }
      default:
         return std::make_unique<Complement>(t, std::move(op));
   }
}

// complement.cpp:23-49
OperatorPtr fromDeMorgan(const Table* t, OperatorVector disjunction) {
   OperatorVector non_negated, negated;
   for (auto& c : disjunction) {
      if (c->type() == OpType::COMPLEMENT) {
         negated.push_back(Operator::negate(std::move(c)));
      } else {
         non_negated.push_back(std::move(c));
      }
   }
   auto inter = std::make_unique<Intersection>(t, std::move(negated), std::move(non_negated));
   return std::make_unique<Complement>(t, std::move(inter));
}

// ------------------------------------------------------------------------------------------------
// Filter expressions (logical AST) with rewrite(NONE) and compile: scalar_expressions/*.cpp.
// ------------------------------------------------------------------------------------------------
struct Expr;
using ExprPtr = std::unique_ptr<Expr>;

struct Expr {
   enum Kind {
      TRUE_, FALSE_, AND, OR, NOT, NOF, SYM_EQ, HAS_MUT, SYM_IN, STR_EQ, LINEAGE,
      DATE_BETWEEN, INT_BETWEEN, INT_EQ, BITMAP, COVERED,
      // operator-level forms used by the reference's operator tests: they build the operator
      // directly with its constructor rules instead of through an expression's compile()
      OP_THRESHOLD, OP_INTERSECTION, OP_UNION, OP_COMPLEMENT, OP_RANGES, OP_SELECT
   } kind;
   struct ValuePredicate {  // OP_SELECT: CompareToValueSelection (selection.h:105-131)
      std::string column;
      int comparator;
      int64_t value;
   };
   std::vector<ExprPtr> children;
   std::string column;
   uint32_t position = 0;
   char symbol_char = 0;
   std::vector<uint8_t> symbols;
   std::string symbol_chars;
   int number = 0;
   bool exact = false;
   std::optional<std::string> str;
   bool sublineage = false;
   std::optional<int64_t> from, to;
   int64_t value = 0;
   uint32_t bitmap_id = 0;
   std::vector<uint32_t> row_ranges;  // OP_RANGES: global [start, end) pairs
   std::vector<ValuePredicate> predicates;

   ExprPtr clone() const {
      auto e = std::make_unique<Expr>(*this, 0);
      for (const auto& c : children) e->children.push_back(c->clone());
      return e;
   }
   Expr(Kind k) : kind(k) {}
   Expr(const Expr& o, int)
       : kind(o.kind), column(o.column), position(o.position), symbol_char(o.symbol_char),
         symbols(o.symbols), symbol_chars(o.symbol_chars), number(o.number), exact(o.exact), str(o.str),
         sublineage(o.sublineage), from(o.from), to(o.to), value(o.value), bitmap_id(o.bitmap_id),
         row_ranges(o.row_ranges), predicates(o.predicates) {}
};

// --- S-expression reader -------------------------------------------------------------------------
struct Reader {
   const std::string& s;
   size_t i = 0;
   explicit Reader(const std::string& src) : s(src) {}
   void ws() {
      while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
   }
   bool peekClose() {
      ws();
      return i < s.size() && s[i] == ')';
   }
   bool peekOpen() {
      ws();
      return i < s.size() && s[i] == '(';
   }
   void expect(char c) {
      ws();
      if (i >= s.size() || s[i] != c) throw QueryError(std::string("malformed filter: expected '") + c + "'");
      ++i;
   }
   // returns {is_string, text}; "_" means absent
   std::pair<bool, std::string> atom() {
      ws();
      if (i >= s.size()) throw QueryError("malformed filter: unexpected end");
      if (s[i] == '"') {
         ++i;
         std::string out;
         while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) ++i;
            out.push_back(s[i++]);
         }
         if (i >= s.size()) throw QueryError("malformed filter: unterminated string");
         ++i;
         return {true, out};
      }
      size_t b = i;
      while (i < s.size() && s[i] != ' ' && s[i] != ')' && s[i] != '(' && s[i] != '\n' && s[i] != '\t') ++i;
      if (b == i) throw QueryError("malformed filter: empty atom");
      return {false, s.substr(b, i - b)};
   }
   int64_t integer() {
      auto [q, t] = atom();
      if (q) throw QueryError("malformed filter: expected integer");
      try {
         size_t used = 0;
         const long long v = std::stoll(t, &used);
         if (used != t.size()) throw QueryError("malformed filter: bad integer " + t);
         return v;
      } catch (const std::logic_error&) {
         throw QueryError("malformed filter: bad integer " + t);
      }
   }
   std::optional<int64_t> optInteger() {
      ws();
      size_t save = i;
      auto [q, t] = atom();
      if (!q && t == "_") return std::nullopt;
      i = save;
      return integer();
   }
   std::string name() {
      auto [q, t] = atom();
      return t;
   }
   std::optional<std::string> optString() {
      auto [q, t] = atom();
      if (!q && t == "null") return std::nullopt;
      return t;
   }

   ExprPtr expr() {
      expect('(');
      const std::string op = name();
      ExprPtr e;
      auto children = [&](Expr& x) {
         while (!peekClose()) x.children.push_back(expr());
      };
      if (op == "true") {
         e = std::make_unique<Expr>(Expr::TRUE_);
      } else if (op == "false") {
         e = std::make_unique<Expr>(Expr::FALSE_);
      } else if (op == "and") {
         e = std::make_unique<Expr>(Expr::AND);
         children(*e);
      } else if (op == "or") {
         e = std::make_unique<Expr>(Expr::OR);
         children(*e);
      } else if (op == "not") {
         e = std::make_unique<Expr>(Expr::NOT);
         e->children.push_back(expr());
      } else if (op == "nof") {
         e = std::make_unique<Expr>(Expr::NOF);
         e->number = static_cast<int>(integer());
         e->exact = integer() != 0;
         children(*e);
      } else if (op == "sym_eq") {
         e = std::make_unique<Expr>(Expr::SYM_EQ);
         e->column = name();
         e->position = static_cast<uint32_t>(integer());
         auto t = atom().second;
         if (t.size() != 1) throw QueryError("malformed filter: symbol must be one character");
         e->symbol_char = t[0];
      } else if (op == "has_mut") {
         e = std::make_unique<Expr>(Expr::HAS_MUT);
         e->column = name();
         e->position = static_cast<uint32_t>(integer());
      } else if (op == "sym_in") {
         e = std::make_unique<Expr>(Expr::SYM_IN);
         e->column = name();
         e->position = static_cast<uint32_t>(integer());
         e->symbol_chars = atom().second;
      } else if (op == "str_eq") {
         e = std::make_unique<Expr>(Expr::STR_EQ);
         e->column = name();
         e->str = optString();
      } else if (op == "lineage") {
         e = std::make_unique<Expr>(Expr::LINEAGE);
         e->column = name();
         e->str = optString();
         e->sublineage = integer() != 0;
      } else if (op == "date_between" || op == "int_between") {
         e = std::make_unique<Expr>(op == "date_between" ? Expr::DATE_BETWEEN : Expr::INT_BETWEEN);
         e->column = name();
         e->from = optInteger();
         e->to = optInteger();
      } else if (op == "int_eq") {
         e = std::make_unique<Expr>(Expr::INT_EQ);
         e->column = name();
         e->value = integer();
      } else if (op == "bitmap") {
         e = std::make_unique<Expr>(Expr::BITMAP);
         e->bitmap_id = static_cast<uint32_t>(integer());
      } else if (op == "covered") {
         e = std::make_unique<Expr>(Expr::COVERED);
         e->column = name();
         e->position = static_cast<uint32_t>(integer());
      } else if (op == "op_threshold") {
         e = std::make_unique<Expr>(Expr::OP_THRESHOLD);
         e->number = static_cast<int>(integer());
         e->exact = integer() != 0;
         children(*e);
      } else if (op == "op_intersection" || op == "op_union") {
         e = std::make_unique<Expr>(op == "op_union" ? Expr::OP_UNION : Expr::OP_INTERSECTION);
         children(*e);
      } else if (op == "op_complement") {
         e = std::make_unique<Expr>(Expr::OP_COMPLEMENT);
         e->children.push_back(expr());
      } else if (op == "op_range") {
         // RangeSelection(ranges, row_layout) (range_selection.h) over global row-id pairs
         e = std::make_unique<Expr>(Expr::OP_RANGES);
         while (!peekClose()) {
            const int64_t v = integer();
            if (v < 0 || v > UINT32_MAX) throw QueryError("malformed filter: row id out of range");
            e->row_ranges.push_back(static_cast<uint32_t>(v));
         }
         if (e->row_ranges.size() % 2 != 0) throw QueryError("malformed filter: op_range takes start/end pairs");
      } else if (op == "op_select") {
         // Selection(child?, predicates) (selection.h) with (cmp COLUMN OP VALUE) predicates
         e = std::make_unique<Expr>(Expr::OP_SELECT);
         while (!peekClose()) {
            const size_t save = i;
            expect('(');
            auto [q, head] = atom();
            if (!q && head == "cmp") {
               Expr::ValuePredicate p;
               p.column = name();
               const std::string c = name();
               const char* const names[] = {"eq", "lt", "gt", "le", "ge", "ne"};
               p.comparator = -1;
               for (int k = 0; k < 6; ++k) {
                  if (c == names[k]) p.comparator = k;
               }
               if (p.comparator < 0) throw QueryError("malformed filter: unknown comparator " + c);
               p.value = integer();
               expect(')');
               e->predicates.push_back(std::move(p));
            } else {
               i = save;
               if (!e->children.empty()) throw QueryError("malformed filter: op_select takes at most one child");
               e->children.push_back(expr());
            }
         }
      } else {
         throw QueryError("malformed filter: unknown expression '" + op + "'");
      }
      expect(')');
      return e;
   }
};

ExprPtr parseFilter(const std::string& text) {
   Reader r(text);
   ExprPtr e = r.expr();
   r.ws();
   if (r.i != text.size()) throw QueryError("malformed filter: trailing characters");
   return e;
}

const SequenceColumn& seqColumn(const Table& t, const std::string& name) {
   auto it = t.seq.find(name);
   if (it == t.seq.end()) throw QueryError("Database does not contain the sequence '" + name + "'");
   return it->second;
}

// Rewrite with AmbiguityMode NONE (compute_filter.cpp:14-21).
ExprPtr rewrite(const Table& t, const Expr& e) {
   switch (e.kind) {
      case Expr::AND:
      case Expr::NOF:
      case Expr::NOT:
      case Expr::OP_THRESHOLD:
      case Expr::OP_INTERSECTION:
      case Expr::OP_UNION:
      case Expr::OP_COMPLEMENT:
      case Expr::OP_RANGES:
      case Expr::OP_SELECT: {
         // and.cpp:139-147, nof.cpp:249-257 (NONE keeps exact NOf), negation.cpp:18-21
         auto r = std::make_unique<Expr>(e, 0);
         for (const auto& c : e.children) r->children.push_back(rewrite(t, *c));
         return r;
      }
      case Expr::OR: {
         // or.cpp:173-192: collect nested Or children, rewrite, drop false / short-circuit true,
         // merge SymbolInSet children of the same (column, position)
         std::vector<const Expr*> queue, collected;
         for (const auto& c : e.children) queue.push_back(c.get());
         while (!queue.empty()) {
            const Expr* cur = queue.back();
            queue.pop_back();
            if (cur->kind == Expr::OR) {
               for (const auto& c : cur->children) queue.push_back(c.get());
            } else {
               collected.push_back(cur);
            }
         }
         std::vector<ExprPtr> unprocessed;
         for (const Expr* c : collected) unprocessed.push_back(rewrite(t, *c));
         std::vector<ExprPtr> non_trivial;
         while (!unprocessed.empty()) {
            auto c = std::move(unprocessed.back());
            unprocessed.pop_back();
            if (c->kind == Expr::TRUE_) {
               std::vector<ExprPtr> single;
               single.push_back(std::make_unique<Expr>(Expr::TRUE_));
               non_trivial = std::move(single);
               break;
            }
            if (c->kind == Expr::FALSE_) continue;
            if (c->kind == Expr::OR) {
               for (auto& cc : c->children) unprocessed.push_back(std::move(cc));
            } else {
               non_trivial.push_back(std::move(c));
            }
         }
         std::vector<ExprPtr> merged;
         std::map<std::pair<std::string, uint32_t>, std::vector<uint8_t>> sym_in;
         for (auto& c : non_trivial) {
            if (c->kind == Expr::SYM_IN) {
               auto& v = sym_in[{c->column, c->position}];
               v.insert(v.end(), c->symbols.begin(), c->symbols.end());
            } else {
               merged.push_back(std::move(c));
            }
         }
         for (auto& [key, syms] : sym_in) {
            auto x = std::make_unique<Expr>(Expr::SYM_IN);
            x->column = key.first;
            x->position = key.second;
            x->symbols = syms;
            merged.push_back(std::move(x));
         }
         if (merged.size() == 1) return std::move(merged[0]);
         auto r = std::make_unique<Expr>(Expr::OR);
         r->children = std::move(merged);
         return r;
      }
      case Expr::SYM_EQ: {
         // symbol_equals.cpp:65-93 ('.' means the reference symbol at that position)
         const auto& col = seqColumn(t, e.column);
         if (e.position >= col.genome_length) {
            throw QueryError("SymbolEquals position is out of bounds");
         }
         int sym;
         if (e.symbol_char == '.') {
            sym = col.reference[e.position];
         } else {
            sym = col.alphabet->fromChar(e.symbol_char);
            if (sym < 0) throw QueryError(std::string("invalid symbol '") + e.symbol_char + "'");
         }
         auto r = std::make_unique<Expr>(Expr::SYM_IN);
         r->column = e.column;
         r->position = e.position;
         r->symbols = {static_cast<uint8_t>(sym)};
         return r;
      }
      case Expr::HAS_MUT: {
         // has_mutation.cpp:34-67 (mode NONE: drop every symbol that could be the reference base)
         const auto& col = seqColumn(t, e.column);
         if (e.position >= col.genome_length) {
            throw QueryError("HasMutation position is out of bounds");
         }
         const uint8_t ref = col.reference[e.position];
         auto r = std::make_unique<Expr>(Expr::SYM_IN);
         r->column = e.column;
         r->position = e.position;
         const auto& amb = col.alphabet->ambiguity[ref];
         for (uint8_t s = 0; s < col.alphabet->count; ++s) {
            if (std::find(amb.begin(), amb.end(), s) == amb.end()) r->symbols.push_back(s);
         }
         return r;
      }
      case Expr::SYM_IN: {
         auto r = std::make_unique<Expr>(e, 0);
         if (!e.symbol_chars.empty()) {
            const auto& col = seqColumn(t, e.column);
            r->symbols.clear();
            for (char c : e.symbol_chars) {
               const int s = col.alphabet->fromChar(c);
               if (s < 0) throw QueryError(std::string("invalid symbol '") + c + "'");
               r->symbols.push_back(static_cast<uint8_t>(s));
            }
            r->symbol_chars.clear();
         }
         return r;
      }
      default:
         return e.clone();
   }
}

OperatorPtr compile(const Table& t, const Expr& e);

// symbol_in_set.cpp:95-226 (the four compilation branches) and :231-264
OperatorPtr makeDifference(const Table* t, OperatorPtr left, OperatorPtr right) {
   OperatorVector a, b;
   a.push_back(std::move(left));
   b.push_back(std::move(right));
   return std::make_unique<Intersection>(t, std::move(a), std::move(b));
}

OperatorPtr excludeNull(const Table* t, OperatorPtr op, const SequenceColumn& col) {
   if (col.null_bitmap.empty()) return op;
   return makeDifference(t, std::move(op), std::make_unique<IndexScan>(t, col.null_bitmap));
}

OperatorPtr compileSymbolInSet(const Table& t, const SequenceColumn& col, uint32_t pos, const std::vector<uint8_t>& symbols) {
   if (pos >= col.genome_length) throw QueryError("SymbolInSet position is out of bounds");
   const uint8_t local = col.local_ref[pos];
   const uint8_t missing = col.alphabet->missing;
   const bool inc_ref = std::find(symbols.begin(), symbols.end(), local) != symbols.end();
   const bool inc_missing = std::find(symbols.begin(), symbols.end(), missing) != symbols.end();
   const Table* tp = &t;
   auto negated = [&](bool exclude_missing) {
      std::vector<uint8_t> out;
      for (uint8_t s = 0; s < col.alphabet->count; ++s) {
         if (exclude_missing && s == missing) continue;
         if (std::find(symbols.begin(), symbols.end(), s) == symbols.end()) out.push_back(s);
      }
      return out;
   };
   if (inc_ref && inc_missing) {
      auto bm = col.matching(pos, negated(false));
      return excludeNull(tp, std::make_unique<Complement>(tp, std::make_unique<IndexScan>(tp, std::move(bm))), col);
   }
   if (inc_missing) {
      auto bm = col.matching(pos, symbols);
      OperatorVector u;
      PredicateVector p;
      p.push_back(std::make_unique<IsInCoveredRegion>(tp, &col, pos, false));
      u.push_back(std::make_unique<Selection>(tp, std::nullopt, std::move(p)));
      u.push_back(std::make_unique<IndexScan>(tp, std::move(bm)));
      return excludeNull(tp, std::make_unique<Union>(tp, std::move(u)), col);
   }
   if (inc_ref) {
      auto bm = col.matching(pos, negated(true));
      PredicateVector p;
      p.push_back(std::make_unique<IsInCoveredRegion>(tp, &col, pos, true));
      return makeDifference(
         tp, std::make_unique<Selection>(tp, std::nullopt, std::move(p)), std::make_unique<IndexScan>(tp, std::move(bm))
      );
   }
   return std::make_unique<IndexScan>(tp, col.matching(pos, symbols));
}

// nof.cpp:16-139 (handleTrivialCases, handleAndCase, handleOrCase, toOperator)
OperatorPtr nofToOperator(const Table* t, int n, OperatorVector a, OperatorVector b, bool exact) {
   const int count = static_cast<int>(a.size() + b.size());
   if (n > count) return std::make_unique<Empty>(t);
   if (n < 0) {
      if (exact) return std::make_unique<Empty>(t);
      return std::make_unique<Full>(t);
   }
   if (n == 0) {
      if (!exact) return std::make_unique<Full>(t);
      if (count == 0) return std::make_unique<Full>(t);
      if (count == 1) {
         if (a.empty()) return std::move(b[0]);
         return std::make_unique<Complement>(t, std::move(a[0]));
      }
      if (b.empty()) return std::make_unique<Complement>(t, std::make_unique<Union>(t, std::move(a)));
      return std::make_unique<Intersection>(t, std::move(b), std::move(a));
   }
   if (n == 1 && count == 1) {
      if (b.empty()) return std::move(a[0]);
      return std::make_unique<Complement>(t, std::move(b[0]));
   }
   if (n == count) {
      if (a.empty()) return std::make_unique<Complement>(t, std::make_unique<Union>(t, std::move(b)));
      return std::make_unique<Intersection>(t, std::move(a), std::move(b));
   }
   if (n == 1 && !exact) {
      if (b.empty()) return std::make_unique<Union>(t, std::move(a));
      return std::make_unique<Complement>(t, std::make_unique<Intersection>(t, std::move(b), std::move(a)));
   }
   return std::make_unique<Threshold>(t, std::move(a), std::move(b), static_cast<uint32_t>(n), exact);
}

OperatorPtr compile(const Table& t, const Expr& e) {
   const Table* tp = &t;
   switch (e.kind) {
      case Expr::TRUE_:
         return std::make_unique<Full>(tp);  // literal.cpp:115-121
      case Expr::FALSE_:
         return std::make_unique<Empty>(tp);
      case Expr::AND: {
         // and.cpp:91-219
         OperatorVector unprocessed;
         for (const auto& c : e.children) unprocessed.push_back(compile(t, *c));
         OperatorVector pos, neg;
         PredicateVector preds;
         while (!unprocessed.empty()) {
            auto c = std::move(unprocessed.back());
            unprocessed.pop_back();
            if (c->type() == OpType::FULL) continue;
            if (c->type() == OpType::EMPTY) return std::make_unique<Empty>(tp);
            if (c->type() == OpType::INTERSECTION) {
               auto* x = static_cast<Intersection*>(c.get());
               for (auto& cc : x->children) pos.push_back(std::move(cc));
               for (auto& cc : x->negated) neg.push_back(std::move(cc));
            } else if (c->type() == OpType::COMPLEMENT) {
               neg.push_back(Operator::negate(std::move(c)));
            } else if (c->type() == OpType::SELECTION) {
               auto* s = static_cast<Selection*>(c.get());
               for (auto& p : s->predicates) preds.push_back(std::move(p));
               if (s->child.has_value()) unprocessed.push_back(std::move(*s->child));
            } else {
               pos.push_back(std::move(c));
            }
         }
         if (pos.empty() && neg.empty()) {
            if (preds.empty()) return std::make_unique<Full>(tp);
            return std::make_unique<Selection>(tp, std::nullopt, std::move(preds));
         }
         OperatorPtr idx;
         if (pos.size() == 1 && neg.empty()) {
            idx = std::move(pos[0]);
         } else if (neg.size() == 1 && pos.empty()) {
            idx = std::make_unique<Complement>(tp, std::move(neg[0]));
         } else if (pos.empty()) {
            idx = std::make_unique<Complement>(tp, std::make_unique<Union>(tp, std::move(neg)));
         } else {
            idx = std::make_unique<Intersection>(tp, std::move(pos), std::move(neg));
         }
         if (preds.empty()) return idx;
         return std::make_unique<Selection>(tp, std::optional<OperatorPtr>(std::move(idx)), std::move(preds));
      }
      case Expr::OR: {
         // or.cpp:194-241
         OperatorVector all;
         for (const auto& c : e.children) all.push_back(compile(t, *c));
         OperatorVector filtered;
         for (auto& c : all) {
            if (c->type() == OpType::EMPTY) continue;
            if (c->type() == OpType::FULL) return std::make_unique<Full>(tp);
            if (c->type() == OpType::UNION) {
               for (auto& cc : static_cast<Union*>(c.get())->children) filtered.push_back(std::move(cc));
            } else {
               filtered.push_back(std::move(c));
            }
         }
         if (filtered.empty()) return std::make_unique<Empty>(tp);
         if (filtered.size() == 1) return std::move(filtered[0]);
         const bool any_complement = std::any_of(filtered.begin(), filtered.end(), [](const auto& c) {
            return c->type() == OpType::COMPLEMENT;
         });
         if (any_complement) return fromDeMorgan(tp, std::move(filtered));
         return std::make_unique<Union>(tp, std::move(filtered));
      }
      case Expr::NOT:
         return Operator::negate(compile(t, *e.children[0]));  // negation.cpp:32-34
      case Expr::NOF: {
         // nof.cpp:181-214 mapChildExpressions, :259-277 compile
         OperatorVector a, b;
         int n = e.number;
         for (const auto& c : e.children) {
            auto op = compile(t, *c);
            if (op->type() == OpType::EMPTY) continue;
            if (op->type() == OpType::FULL) {
               --n;
               continue;
            }
            if (op->type() == OpType::COMPLEMENT) {
               b.push_back(Operator::negate(std::move(op)));
               continue;
            }
            a.push_back(std::move(op));
         }
         if (n < 0) {
            if (e.exact) return std::make_unique<Empty>(tp);
            return std::make_unique<Full>(tp);
         }
         return nofToOperator(tp, n, std::move(a), std::move(b), e.exact);
      }
      case Expr::SYM_IN: {
         const auto& col = seqColumn(t, e.column);
         return compileSymbolInSet(t, col, e.position, e.symbols);
      }
      case Expr::STR_EQ: {
         // equals.cpp:158-187 (indexed dictionary column) and dictionary_encoded_column.cpp filter
         auto it = t.dicts.find(e.column);
         if (it == t.dicts.end()) throw QueryError("The database does not contain the column '" + e.column + "'");
         const auto& col = it->second;
         const Bitmap* bm = nullptr;
         if (!e.str.has_value()) {
            bm = &col.null_bitmap;
         } else {
            auto id = col.ids.find(*e.str);
            if (id != col.ids.end()) {
               auto iv = col.indexed_values.find(id->second);
               if (iv != col.indexed_values.end()) bm = &iv->second;
            }
         }
         if (bm == nullptr || bm->empty()) return std::make_unique<Empty>(tp);
         return std::make_unique<IndexScan>(tp, *bm);
      }
      case Expr::LINEAGE: {
         // lineage_filter.cpp:30-100
         auto it = t.dicts.find(e.column);
         if (it == t.dicts.end() || !it->second.has_lineage_index) {
            throw QueryError("The database does not contain a lineage index for the column '" + e.column + "'");
         }
         const auto& col = it->second;
         const Bitmap* bm = nullptr;
         if (!e.str.has_value()) {
            bm = &col.null_bitmap;
         } else {
            auto id = col.ids.find(*e.str);
            if (id == col.ids.end()) {
               throw QueryError("The lineage '" + *e.str + "' is not a valid lineage for column '" + e.column + "'.");
            }
            const auto& index = e.sublineage ? col.including_sublineages : col.indexed_values;
            auto iv = index.find(id->second);
            if (iv != index.end()) bm = &iv->second;
         }
         if (bm == nullptr) return std::make_unique<Empty>(tp);
         return std::make_unique<IndexScan>(tp, *bm);
      }
      case Expr::DATE_BETWEEN: {
         // date_between.cpp:61-135
         auto it = t.ints.find(e.column);
         if (it == t.ints.end() || !it->second.is_date) throw QueryError("The column '" + e.column + "' is not of type date");
         const auto& col = it->second;
         const int32_t from = e.from.has_value() ? static_cast<int32_t>(*e.from) : std::numeric_limits<int32_t>::min();
         const int32_t to = e.to.has_value() ? static_cast<int32_t>(*e.to) : std::numeric_limits<int32_t>::max();
         if (col.is_sorted) {
            std::vector<std::pair<uint32_t, uint32_t>> ranges;
            for (size_t c = 0; c < t.layout.sizes.size(); ++c) {
               const int32_t* b = col.values.data() + t.chunk_offset[c];
               const int32_t* en = b + t.layout.sizes[c];
               const auto lo = std::lower_bound(b, en, from) - b;
               const auto hi = (e.to.has_value() ? std::upper_bound(b, en, to) : en) - b;
               const uint32_t size = t.layout.sizes[c];
               auto rid = [&](int64_t idx) {
                  if (idx == static_cast<int64_t>(size)) return static_cast<uint32_t>((c + 1) << 16);
                  return static_cast<uint32_t>((c << 16) | static_cast<uint32_t>(idx));
               };
               ranges.emplace_back(rid(lo), rid(hi));
            }
            return std::make_unique<RangeSelection>(tp, std::move(ranges));
         }
         PredicateVector p;
         p.push_back(std::make_unique<CompareToValue>(tp, &col, Comparator::HIGHER_OR_EQUALS, from, false));
         p.push_back(std::make_unique<CompareToValue>(tp, &col, Comparator::LESS_OR_EQUALS, to, false));
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case Expr::INT_BETWEEN: {
         // int_between.cpp:71-131
         auto it = t.ints.find(e.column);
         if (it == t.ints.end() || it->second.is_date) throw QueryError("The column '" + e.column + "' is not of type int32 or int64");
         const auto& col = it->second;
         PredicateVector p;
         if (e.from.has_value()) {
            p.push_back(std::make_unique<CompareToValue>(tp, &col, Comparator::HIGHER_OR_EQUALS, static_cast<int32_t>(*e.from), false));
         }
         if (e.to.has_value()) {
            p.push_back(std::make_unique<CompareToValue>(tp, &col, Comparator::LESS_OR_EQUALS, static_cast<int32_t>(*e.to), false));
         }
         if (p.empty()) {
            return std::make_unique<Complement>(tp, std::make_unique<IndexScan>(tp, col.null_bitmap));
         }
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case Expr::INT_EQ: {
         // equals.cpp:62-99 compileEquals / compileIntEquals
         auto it = t.ints.find(e.column);
         if (it == t.ints.end()) throw QueryError("The column '" + e.column + "' is not of type int");
         if (e.value < std::numeric_limits<int32_t>::min() || e.value > std::numeric_limits<int32_t>::max()) {
            throw QueryError("Cannot cast value to int32. Value out of range");
         }
         PredicateVector p;
         p.push_back(std::make_unique<CompareToValue>(tp, &it->second, Comparator::EQUALS, static_cast<int32_t>(e.value), false));
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case Expr::BITMAP: {
         if (e.bitmap_id >= t.literals.size()) throw QueryError("unknown bitmap literal");
         return std::make_unique<IndexScan>(tp, t.literals[e.bitmap_id]);
      }
      case Expr::COVERED: {
         const auto& col = seqColumn(t, e.column);
         PredicateVector p;
         p.push_back(std::make_unique<IsInCoveredRegion>(tp, &col, e.position, true));
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case Expr::OP_THRESHOLD:
      case Expr::OP_INTERSECTION: {
         OperatorVector a, b;
         for (const auto& c : e.children) {
            auto op = compile(t, *c);
            if (op->type() == OpType::COMPLEMENT) {
               b.push_back(Operator::negate(std::move(op)));
            } else {
               a.push_back(std::move(op));
            }
         }
         if (e.kind == Expr::OP_INTERSECTION) return std::make_unique<Intersection>(tp, std::move(a), std::move(b));
         if (e.number < 0) throw QueryError("Compilation Error: number_of_matchers must be greater than zero");
         return std::make_unique<Threshold>(tp, std::move(a), std::move(b), static_cast<uint32_t>(e.number), e.exact);
      }
      case Expr::OP_UNION: {
         OperatorVector a;
         for (const auto& c : e.children) a.push_back(compile(t, *c));
         return std::make_unique<Union>(tp, std::move(a));
      }
      case Expr::OP_COMPLEMENT:
         return std::make_unique<Complement>(tp, compile(t, *e.children[0]));
      case Expr::OP_RANGES: {
         // range_selection.cpp:54-60 asserts every end is a row of the layout or the layout's end
         const auto& sizes = t.layout.sizes;
         const uint64_t layout_end = static_cast<uint64_t>(sizes.size()) << 16;
         auto valid = [&](uint32_t id) {
            const uint32_t c = id >> 16;
            return (c < sizes.size() && (id & 0xFFFFu) < sizes[c]) || id == layout_end;
         };
         std::vector<std::pair<uint32_t, uint32_t>> ranges;
         for (size_t k = 0; k < e.row_ranges.size(); k += 2) {
            const uint32_t a = e.row_ranges[k], b = e.row_ranges[k + 1];
            if (!valid(a) || !valid(b) || a > b) throw QueryError("Compilation Error: RangeSelection range outside the row layout");
            ranges.emplace_back(a, b);
         }
         return std::make_unique<RangeSelection>(tp, std::move(ranges));
      }
      case Expr::OP_SELECT: {
         if (e.predicates.empty()) throw QueryError("Compilation Error: Selection needs at least one predicate");
         PredicateVector p;
         for (const auto& vp : e.predicates) {
            auto it = t.ints.find(vp.column);
            if (it == t.ints.end()) throw QueryError("The column '" + vp.column + "' is not of type int or date");
            if (vp.value < INT32_MIN || vp.value > INT32_MAX) throw QueryError("Value out of int32 range");
            p.push_back(std::make_unique<CompareToValue>(tp, &it->second, static_cast<Comparator>(vp.comparator),
                                                         static_cast<int32_t>(vp.value), false));
         }
         std::optional<OperatorPtr> child;
         if (!e.children.empty()) child = compile(t, *e.children[0]);
         return std::make_unique<Selection>(tp, std::move(child), std::move(p));
      }
      default:
         throw QueryError("expression must be rewritten before compilation");
   }
}

Bitmap computeFilter(const Table& t, const std::string& text) {
   auto e = parseFilter(text);
   auto r = rewrite(t, *e);
   auto op = compile(t, *r);
   return op->evaluate();
}

// ------------------------------------------------------------------------------------------------
// Mutations: operators/mutations_node.cpp:39-288 (counting) and :292-366 (output rows).
// ------------------------------------------------------------------------------------------------
void subtractCumulative(std::vector<uint32_t>& count, uint32_t L, const std::vector<size_t>& cs, const std::vector<size_t>& ce) {
   // mutations_node.cpp:63-90
   size_t run_s = cs.at(L);
   size_t p = L - 1;
   while (true) {
      count.at(p) -= static_cast<uint32_t>(run_s);
      run_s += cs.at(p);
      if (p == 0) break;
      p -= 1;
   }
   size_t run_e = ce.at(0);
   p = 0;
   while (true) {
      count.at(p) -= static_cast<uint32_t>(run_e);
      run_e += ce.at(p + 1);
      if (p == L - 1) break;
      p += 1;
   }
}

std::vector<std::vector<uint32_t>> mutationCounts(const Table& t, const SequenceColumn& col, const Bitmap& filter) {
   const uint32_t L = col.genome_length;
   std::vector<std::vector<uint32_t>> counts(col.alphabet->count, std::vector<uint32_t>(L, 0));
   const uint64_t card = filter.cardinality();
   std::vector<uint32_t> ref(L, 0);
   if (card == t.layout.num_rows) {
      // addMutationCountsForFullBitmaps :240-266
      for (auto& x : ref) x += col.sequence_count;
      for (const auto& [row, bm] : col.horizontal) bm.forEach([&](uint32_t p) { ref[p] -= 1; });
      std::vector<size_t> cs(L + 1, 0), ce(L + 1, 0);
      for (const auto& chunk : col.start_end) {
         for (const auto& [s, e] : chunk) {
            cs.at(s) += 1;
            ce.at(e) += 1;
         }
      }
      subtractCumulative(ref, L, cs, ce);
      for (const auto& [key, c] : col.vertical) {
         counts[key.sym][key.pos] += c.card;
         ref[key.pos] -= c.card;
      }
   } else if (card > 0) {
      // addMutationCountsForMixedBitmaps :206-237
      for (auto& x : ref) x += static_cast<uint32_t>(card);
      std::vector<size_t> cs(L + 1, 0), ce(L + 1, 0);
      filter.forEach([&](uint32_t g) {
         auto it = col.horizontal.find(g);
         if (it != col.horizontal.end()) it->second.forEach([&](uint32_t p) { ref[p] -= 1; });
         const auto [s, e] = col.coverageRange(g);
         cs.at(s) += 1;
         ce.at(e) += 1;
      });
      subtractCumulative(ref, L, cs, ce);
      // countActualFilteredMutations :154-189
      std::unordered_map<uint16_t, const Container*> fc;
      for (size_t i = 0; i < filter.keys.size(); ++i) fc[filter.keys[i]] = filter.cs[i].get();
      for (const auto& [key, c] : col.vertical) {
         auto it = fc.find(key.chunk);
         if (it == fc.end()) continue;
         const uint32_t n = andCardinality(*it->second, c);
         counts[key.sym][key.pos] += n;
         ref[key.pos] -= n;
      }
   } else {
      return counts;
   }
   for (uint32_t p = 0; p < L; ++p) counts[col.local_ref[p]][p] += ref[p];  // accumulateFinalCounts
   return counts;
}

struct MutationRow {
   uint32_t position;  // 1-based
   uint8_t from, to;
   uint32_t count, coverage;
   double proportion;
};

std::vector<MutationRow> mutationRows(const SequenceColumn& col, const std::vector<std::vector<uint32_t>>& counts, double min_proportion) {
   std::vector<MutationRow> out;
   const auto& valid = col.alphabet->valid;
   for (uint32_t p = 0; p < col.genome_length; ++p) {
      uint32_t total = 0;
      for (uint8_t s : valid) total += counts[s][p];
      if (total == 0) continue;
      const uint32_t thr = min_proportion == 0
                              ? 0
                              : static_cast<uint32_t>(std::ceil(static_cast<double>(total) * min_proportion) - 1);
      const uint8_t orig = col.reference[p];
      for (uint8_t s : valid) {
         if (s == orig) continue;
         const uint32_t c = counts[s][p];
         if (c > thr) {
            out.push_back({p + 1, orig, s, c, total, static_cast<double>(c) / static_cast<double>(total)});
         }
      }
   }
   return out;
}

// ------------------------------------------------------------------------------------------------
// Aggregated: operators/bitmap_aggregation_node.cpp:53-130, 226-253.
// ------------------------------------------------------------------------------------------------
struct Group {
   uint32_t id;  // dictionary id / symbol; UINT32_MAX = null group
   Bitmap rows;
};

std::vector<Group> buildDictGroups(const DictColumn& col, const Bitmap& filter) {
   std::vector<std::pair<std::string, Group>> groups;
   for (const auto& [id, bm] : col.indexed_values) {
      Bitmap g = andB(filter, bm);
      if (g.empty()) continue;
      groups.push_back({col.dictionary[id], Group{id, std::move(g)}});
   }
   std::stable_sort(groups.begin(), groups.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
   std::vector<Group> out;
   for (auto& g : groups) out.push_back(std::move(g.second));
   Bitmap nulls = andB(col.null_bitmap, filter);
   if (!nulls.empty()) out.push_back(Group{UINT32_MAX, std::move(nulls)});
   return out;
}

std::vector<Group> buildSymbolGroups(const Table& t, const SequenceColumn& col, uint32_t pos, const Bitmap& filter) {
   std::vector<Group> out;
   for (uint8_t s = 0; s < col.alphabet->count; ++s) {
      auto op = compileSymbolInSet(t, col, pos, {s});
      Bitmap b = andB(op->evaluate(), filter);
      if (!b.empty()) out.push_back(Group{s, std::move(b)});
   }
   Bitmap nulls = andB(filter, col.null_bitmap);
   if (!nulls.empty()) out.push_back(Group{UINT32_MAX, std::move(nulls)});
   return out;
}

void partition(const Bitmap& current, size_t depth, const std::vector<std::vector<Group>>& dims,
               std::vector<uint32_t>& acc, std::vector<std::pair<std::vector<uint32_t>, uint64_t>>& out) {
   if (depth == dims.size()) {
      out.emplace_back(acc, current.cardinality());
      return;
   }
   for (const auto& g : dims[depth]) {
      Bitmap inter = andB(current, g.rows);
      if (inter.empty()) continue;
      acc.push_back(g.id);
      partition(inter, depth + 1, dims, acc, out);
      acc.pop_back();
   }
}

}  // namespace oracle

// ------------------------------------------------------------------------------------------------
// C API used by the tests and by the CPU-baseline leg of the benchmark.
// ------------------------------------------------------------------------------------------------
using namespace oracle;

namespace {
thread_local std::string g_error;
int fail(Table* t, const std::string& msg) {
   if (t) t->error = msg;
   g_error = msg;
   return -1;
}
}  // namespace

extern "C" {

void* or_table_create(const uint32_t* chunk_sizes, uint32_t n_chunks) {
   if (n_chunks >= 0xFFFF) return nullptr;  // row_layout.h:44
   auto* t = new Table;
   uint64_t off = 0;
   for (uint32_t c = 0; c < n_chunks; ++c) {
      if (chunk_sizes[c] > 65536) {
         delete t;
         return nullptr;
      }
      t->layout.sizes.push_back(chunk_sizes[c]);
      t->chunk_offset.push_back(off);
      off += chunk_sizes[c];
   }
   t->chunk_offset.push_back(off);
   t->layout.num_rows = off;
   return t;
}

void or_table_free(void* t) { delete static_cast<Table*>(t); }

const char* or_last_error(void* t) { return t ? static_cast<Table*>(t)->error.c_str() : g_error.c_str(); }

int or_add_sequence_column(
   void* tv, const char* name, int alphabet, const uint8_t* reference, uint32_t L,
   const uint8_t* is_null, const uint32_t* cov_start, const uint32_t* cov_end,
   const uint64_t* mut_ptr, const uint32_t* mut_pos, const uint8_t* mut_sym,
   const uint64_t* nrun_ptr, const uint32_t* nrun_start, const uint32_t* nrun_end
) {
   auto* t = static_cast<Table*>(tv);
   try {
      if (L == 0) return fail(t, "empty reference");
      SequenceColumn col;
      col.name = name;
      col.alphabet = alphabet == 0 ? &nucleotide() : &aminoAcid();
      col.genome_length = L;
      col.reference.assign(reference, reference + L);
      const uint64_t n = t->layout.num_rows;
      for (uint64_t i = 0; i < n; ++i) {
         if (!is_null[i] && (cov_start[i] > cov_end[i] || cov_end[i] > L)) return fail(t, "invalid coverage");
         for (uint64_t k = mut_ptr[i] + 1; k < mut_ptr[i + 1]; ++k) {
            if (mut_pos[k] <= mut_pos[k - 1]) return fail(t, "mutation positions of a row must be strictly increasing");
         }
      }
      buildSequenceColumn(col, *t, is_null, cov_start, cov_end, mut_ptr, mut_pos, mut_sym, nrun_ptr, nrun_start, nrun_end);
      t->seq[name] = std::move(col);
      return 0;
   } catch (const std::exception& e) {
      return fail(t, e.what());
   }
}

int or_add_int_column(void* tv, const char* name, int is_date, const int32_t* values, const uint8_t* is_null) {
   auto* t = static_cast<Table*>(tv);
   IntColumn col;
   col.name = name;
   col.is_date = is_date != 0;
   const uint64_t n = t->layout.num_rows;
   col.values.assign(values, values + n);
   std::vector<uint32_t> nulls;
   std::optional<int32_t> last;
   for (uint64_t i = 0; i < n; ++i) {
      if (is_null[i]) {
         nulls.push_back(t->global(i));
         col.values[i] = 0;
         col.is_sorted = false;
      } else {
         if (last.has_value() && values[i] < *last) col.is_sorted = false;
         last = values[i];
      }
   }
   col.null_bitmap = Bitmap::fromSorted(nulls.data(), nulls.size());
   t->ints[name] = std::move(col);
   return 0;
}

int or_add_dict_column(
   void* tv, const char* name, const uint32_t* ids, const uint8_t* is_null, uint32_t n_values,
   const char* names, const uint32_t* name_offsets, const int32_t* lineage_parent
) {
   auto* t = static_cast<Table*>(tv);
   DictColumn col;
   col.name = name;
   for (uint32_t v = 0; v < n_values; ++v) {
      col.dictionary.emplace_back(names + name_offsets[v], names + name_offsets[v + 1]);
      col.ids[col.dictionary.back()] = v;
   }
   const uint64_t n = t->layout.num_rows;
   col.values.assign(ids, ids + n);
   std::vector<std::vector<uint32_t>> rows(n_values);
   std::vector<uint32_t> nulls;
   for (uint64_t i = 0; i < n; ++i) {
      const uint32_t g = t->global(i);
      if (is_null[i]) {
         nulls.push_back(g);
      } else {
         if (ids[i] >= n_values) return fail(t, "dictionary id out of range");
         rows[ids[i]].push_back(g);
      }
   }
   for (uint32_t v = 0; v < n_values; ++v) {
      if (!rows[v].empty()) col.indexed_values[v] = Bitmap::fromSorted(rows[v].data(), rows[v].size());
   }
   col.null_bitmap = Bitmap::fromSorted(nulls.data(), nulls.size());
   if (lineage_parent != nullptr) {
      // lineage_index.h:17-42: per value, the rows of the value and of all its descendants
      col.has_lineage_index = true;
      for (uint32_t v = 0; v < n_values; ++v) {
         std::vector<uint32_t> acc;
         for (uint32_t w = 0; w < n_values; ++w) {
            int32_t x = static_cast<int32_t>(w);
            bool desc = false;
            int guard = 0;
            while (x >= 0 && guard++ <= static_cast<int32_t>(n_values)) {
               if (static_cast<uint32_t>(x) == v) {
                  desc = true;
                  break;
               }
               x = lineage_parent[x];
            }
            if (desc) acc.insert(acc.end(), rows[w].begin(), rows[w].end());
         }
         std::sort(acc.begin(), acc.end());
         if (!acc.empty()) col.including_sublineages[v] = Bitmap::fromSorted(acc.data(), acc.size());
      }
   }
   t->dicts[name] = std::move(col);
   return 0;
}

int or_add_bitmap(void* tv, const uint32_t* rows, uint64_t n) {
   auto* t = static_cast<Table*>(tv);
   std::vector<uint32_t> v(rows, rows + n);
   std::sort(v.begin(), v.end());
   t->literals.push_back(Bitmap::fromSorted(v.data(), v.size()));
   return static_cast<int>(t->literals.size() - 1);
}

int or_local_reference(void* tv, const char* col, uint8_t* out) {
   auto* t = static_cast<Table*>(tv);
   auto it = t->seq.find(col);
   if (it == t->seq.end()) return fail(t, "unknown sequence column");
   std::copy(it->second.local_ref.begin(), it->second.local_ref.end(), out);
   return 0;
}

/// The column's finalized storage in the form a finalized reference SequenceColumn holds
/// (sequence_column.h:86-97): vertical containers keyed (position, chunk, symbol) in map order with
/// their cardinality, CRoaring typecode (1 bitset, 2 array, 3 run) and container_write bytes
/// (roaring_container.h:105-116); per row (row-layout order) the null flag, [start, end) and the
/// internal missing positions (CSR).  With null outputs it reports the sizes only.
int or_export_sequence_column(void* tv, const char* col, uint64_t* n_containers, uint64_t* payload_bytes,
                              uint64_t* n_missing, uint8_t* local_ref, uint32_t* key_position, uint16_t* key_chunk,
                              uint8_t* key_symbol, uint8_t* typecode, uint32_t* cardinality, uint64_t* payload_offset,
                              uint8_t* payload, uint8_t* is_null, uint32_t* cov_start, uint32_t* cov_end,
                              uint64_t* missing_ptr, uint32_t* missing_pos) {
   auto* t = static_cast<Table*>(tv);
   auto it = t->seq.find(col);
   if (it == t->seq.end()) return fail(t, "unknown sequence column");
   const SequenceColumn& c = it->second;
   uint64_t k = 0, bytes = 0;
   for (const auto& [key, cont] : c.vertical) {
      const uint64_t b = cont.type == Container::BITSET ? 8192 : cont.type == Container::ARRAY ? 2ULL * cont.arr.size()
                                                                                             : 2 + 2ULL * cont.runs.size();
      if (key_position) {
         key_position[k] = key.pos;
         key_chunk[k] = key.chunk;
         key_symbol[k] = key.sym;
         typecode[k] = cont.type == Container::BITSET ? 1 : cont.type == Container::ARRAY ? 2 : 3;
         cardinality[k] = cont.card;
         payload_offset[k] = bytes;
         uint8_t* out = payload + bytes;
         if (cont.type == Container::BITSET) {
            std::memcpy(out, cont.bits.data(), 8192);
         } else if (cont.type == Container::ARRAY) {
            std::memcpy(out, cont.arr.data(), 2 * cont.arr.size());
         } else {
            const uint16_t nr = static_cast<uint16_t>(cont.runs.size() / 2);
            std::memcpy(out, &nr, 2);
            std::memcpy(out + 2, cont.runs.data(), 2 * cont.runs.size());
         }
      }
      bytes += b;
      ++k;
   }
   *n_containers = k;
   *payload_bytes = bytes;
   uint64_t miss = 0;
   for (const auto& [row, bm] : c.horizontal) miss += bm.cardinality();
   *n_missing = miss;
   if (!key_position && !is_null) return 0;
   std::copy(c.local_ref.begin(), c.local_ref.end(), local_ref);
   uint64_t i = 0, m = 0;
   missing_ptr[0] = 0;
   for (size_t ch = 0; ch < c.start_end.size(); ++ch) {
      for (size_t r = 0; r < c.start_end[ch].size(); ++r) {
         const uint32_t g = (static_cast<uint32_t>(ch) << 16) | static_cast<uint32_t>(r);
         is_null[i] = c.null_bitmap.contains(g) ? 1 : 0;
         cov_start[i] = c.start_end[ch][r].first;
         cov_end[i] = c.start_end[ch][r].second;
         auto h = c.horizontal.find(g);
         if (h != c.horizontal.end()) h->second.forEach([&](uint32_t p) { missing_pos[m++] = p; });
         missing_ptr[++i] = m;
      }
   }
   return 0;
}

/// Number of stored vertical containers and their total payload bytes (array 2*card, bitset 8192,
/// run 4*runs) for the sequence column.
int or_index_stats(void* tv, const char* col, uint64_t* n_containers, uint64_t* payload_bytes) {
   auto* t = static_cast<Table*>(tv);
   auto it = t->seq.find(col);
   if (it == t->seq.end()) return fail(t, "unknown sequence column");
   uint64_t n = 0, b = 0;
   for (const auto& [k, c] : it->second.vertical) {
      ++n;
      b += c.type == Container::ARRAY ? 2ULL * c.card : c.type == Container::BITSET ? 8192ULL : 2ULL * c.runs.size();
   }
   *n_containers = n;
   *payload_bytes = b;
   return 0;
}

int or_filter(void* tv, const char* expr, uint32_t* out_rows, uint64_t cap, uint64_t* card) {
   auto* t = static_cast<Table*>(tv);
   try {
      Bitmap b = computeFilter(*t, expr);
      *card = b.cardinality();
      if (out_rows != nullptr) {
         if (*card > cap) return fail(t, "output capacity too small");
         uint64_t k = 0;
         b.forEach([&](uint32_t g) { out_rows[k++] = g; });
      }
      return 0;
   } catch (const std::exception& e) {
      return fail(t, e.what());
   }
}

// The type of the compiled filter's root operator (OpType order of filter/operators/operator.h:11-22):
// tests use it to check which operator a filter compiles to.
int or_filter_root_type(void* tv, const char* expr, int* type) {
   auto* t = static_cast<Table*>(tv);
   try {
      auto e = parseFilter(expr);
      auto r = rewrite(*t, *e);
      *type = static_cast<int>(compile(*t, *r)->type());
      return 0;
   } catch (const std::exception& e) {
      return fail(t, e.what());
   }
}

int or_mutation_counts(void* tv, const char* expr, const char* column, uint32_t* out) {
   auto* t = static_cast<Table*>(tv);
   try {
      const auto& col = seqColumn(*t, column);
      Bitmap b = computeFilter(*t, expr);
      auto counts = mutationCounts(*t, col, b);
      for (size_t s = 0; s < counts.size(); ++s) std::copy(counts[s].begin(), counts[s].end(), out + s * col.genome_length);
      return 0;
   } catch (const std::exception& e) {
      return fail(t, e.what());
   }
}

int or_mutations(
   void* tv, const char* expr, const char* column, double min_proportion, uint32_t* pos, uint8_t* from,
   uint8_t* to, uint32_t* count, uint32_t* coverage, double* proportion, uint64_t cap, uint64_t* n_out
) {
   auto* t = static_cast<Table*>(tv);
   try {
      if (!(min_proportion >= 0.0 && min_proportion <= 1.0)) return fail(t, "minProportion must be in [0, 1]");
      const auto& col = seqColumn(*t, column);
      Bitmap b = computeFilter(*t, expr);
      auto rows = mutationRows(col, mutationCounts(*t, col, b), min_proportion);
      *n_out = rows.size();
      if (rows.size() > cap) return fail(t, "output capacity too small");
      for (size_t i = 0; i < rows.size(); ++i) {
         pos[i] = rows[i].position;
         from[i] = rows[i].from;
         to[i] = rows[i].to;
         count[i] = rows[i].count;
         coverage[i] = rows[i].coverage;
         proportion[i] = rows[i].proportion;
      }
      return 0;
   } catch (const std::exception& e) {
      return fail(t, e.what());
   }
}

/// dims: comma-separated list; "name" = indexed dictionary column, "name@pos" = symbol at 0-based
/// position of a sequence column.  Output: n_out combinations in emission order; out_ids holds
/// n_dims ids per combination (UINT32_MAX = null group).
int or_aggregate(void* tv, const char* expr, const char* dims_spec, uint32_t* out_ids, uint64_t* out_counts,
                 uint64_t cap, uint64_t* n_out) {
   auto* t = static_cast<Table*>(tv);
   try {
      Bitmap filter = computeFilter(*t, expr);
      std::vector<std::vector<Group>> dims;
      std::string spec(dims_spec);
      size_t b = 0;
      while (b <= spec.size()) {
         size_t e = spec.find(',', b);
         if (e == std::string::npos) e = spec.size();
         const std::string d = spec.substr(b, e - b);
         b = e + 1;
         if (d.empty()) continue;
         const auto at = d.find('@');
         if (at != std::string::npos) {
            const auto& col = seqColumn(*t, d.substr(0, at));
            dims.push_back(buildSymbolGroups(*t, col, static_cast<uint32_t>(std::stoul(d.substr(at + 1))), filter));
         } else {
            auto it = t->dicts.find(d);
            if (it == t->dicts.end()) return fail(t, "unknown dictionary column " + d);
            dims.push_back(buildDictGroups(it->second, filter));
         }
      }
      std::vector<std::pair<std::vector<uint32_t>, uint64_t>> combos;
      std::vector<uint32_t> acc;
      partition(filter, 0, dims, acc, combos);
      *n_out = combos.size();
      if (combos.size() > cap) return fail(t, "output capacity too small");
      for (size_t i = 0; i < combos.size(); ++i) {
         for (size_t d = 0; d < dims.size(); ++d) out_ids[i * dims.size() + d] = combos[i].first[d];
         out_counts[i] = combos[i].second;
      }
      return 0;
   } catch (const std::exception& e) {
      return fail(t, e.what());
   }
}

/// aligned_sequence.cpp:18-122 extractCoverageAndMutationsFromSequence (offset 0).
/// Returns 0 and fills start/end, the missing positions inside [start, end) and the mutations;
/// -2 when an output capacity is too small; -1 on an illegal character.
int or_diff_sequence(int alphabet, const char* seq, uint32_t len, const char* reference, uint32_t* start,
                     uint32_t* end, uint32_t* missing, uint64_t missing_cap, uint64_t* n_missing,
                     uint32_t* mut_pos, uint8_t* mut_sym, uint64_t mut_cap, uint64_t* n_mut) {
   const Alphabet& a = alphabet == 0 ? nucleotide() : aminoAcid();
   std::vector<uint32_t> miss;
   std::vector<std::pair<uint32_t, uint8_t>> muts;
   for (uint32_t i = 0; i < len; ++i) {
      const int s = a.fromChar(seq[i]);
      if (s < 0) {
         g_error = std::string("illegal character '") + seq[i] + "' at position " + std::to_string(i) + " in the input sequence";
         return -1;
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
   while (trail < miss.size() - lead && miss[miss.size() - 1 - trail] == len - 1 - trail) ++trail;
   if (lead + trail == len) {
      *start = 0;
      *end = 0;
      miss.clear();
   } else {
      *start = static_cast<uint32_t>(lead);
      *end = static_cast<uint32_t>(len - trail);
      miss.erase(miss.end() - static_cast<ptrdiff_t>(trail), miss.end());
      miss.erase(miss.begin(), miss.begin() + static_cast<ptrdiff_t>(lead));
   }
   *n_missing = miss.size();
   *n_mut = muts.size();
   if (miss.size() > missing_cap || muts.size() > mut_cap) return -2;
   std::copy(miss.begin(), miss.end(), missing);
   for (size_t i = 0; i < muts.size(); ++i) {
      mut_pos[i] = muts[i].first;
      mut_sym[i] = muts[i].second;
   }
   return 0;
}

}  // extern "C"
