// Filter expressions -> filter operators -> filter-kernel program.
#include "query.h"

#include <algorithm>
#include <bit>
#include <cstring>
#include <limits>
#include <map>

namespace rdb::query {

namespace {
std::string joinChildren(const OperatorVector& v, const char* sep = ", ") {
   std::string s;
   for (size_t i = 0; i < v.size(); ++i) {
      if (i) s += sep;
      s += v[i]->toString();
   }
   return s;
}

uint32_t cmpCode(Comparator c) {
   switch (c) {
      case Comparator::EQUALS: return CMP_EQ;
      case Comparator::LESS: return CMP_LT;
      case Comparator::HIGHER: return CMP_GT;
      case Comparator::LESS_OR_EQUALS: return CMP_LE;
      case Comparator::HIGHER_OR_EQUALS: return CMP_GE;
      default: return CMP_NE;
   }
}

const char* cmpText(Comparator c) {
   switch (c) {
      case Comparator::EQUALS: return "=";
      case Comparator::LESS: return "<";
      case Comparator::HIGHER: return ">";
      case Comparator::LESS_OR_EQUALS: return "<=";
      case Comparator::HIGHER_OR_EQUALS: return ">=";
      default: return "!=";
   }
}
}  // namespace

// ---- program execution -------------------------------------------------------------------------
std::vector<uint8_t> programBlob(const Emitter& program) {
   std::vector<uint8_t> blob(sizeof(ProgramDesc) + program.code.size() * sizeof(Instr) +
                             (program.ranges.size() + program.prefetch.size()) * 4 + 16);
   ProgramDesc pd{0, static_cast<uint32_t>(program.code.size()), 0, 0, static_cast<uint32_t>(program.ranges.size()),
                  static_cast<uint32_t>(program.prefetch.size())};
   std::memcpy(blob.data(), &pd, sizeof(pd));
   uint8_t* at = blob.data() + sizeof(pd);
   // (empty vectors may hold null data(): memcpy is only called for non-empty ones)
   auto append = [&](const void* src, size_t bytes) {
      if (bytes) std::memcpy(at, src, bytes);
      at += bytes;
   };
   append(program.code.data(), program.code.size() * sizeof(Instr));
   append(program.ranges.data(), program.ranges.size() * 4);
   append(program.prefetch.data(), program.prefetch.size() * 4);
   return blob;
}

uint32_t programDepth(const Context& ctx, const Emitter& program) {
   if (program.sp != 1) fail(RDB_ECOMPILE, "internal: filter program leaves " + std::to_string(program.sp) + " results");
   const uint32_t depth = std::max<uint32_t>(program.max_sp, 1);
   if (depth > maxFilterDepth(ctx)) fail(RDB_EQUERY, "filter expression nests too deeply");
   return depth;
}

void launchProgram(Table& t, const Emitter& program, const uint8_t* dev_blob, uint64_t* words, unsigned long long* card,
                   uint32_t* chunk_cards) {
   Context& ctx = *t.ctx;
   const uint32_t depth = programDepth(ctx, program);
   if (reinterpret_cast<uint8_t*>(chunk_cards) == reinterpret_cast<uint8_t*>(card) + 64) {
      // the cardinality (64 bytes) directly followed by the per-chunk cardinalities: one clear
      hipCheck(launchFill(reinterpret_cast<uint32_t*>(card), 0, (64 + 4 * static_cast<size_t>(t.n_chunks)) / 4, ctx.stream),
               "clear cardinalities");
   } else {
      hipCheck(launchFill(reinterpret_cast<uint32_t*>(card), 0, 64 / 4, ctx.stream), "clear cardinality");
      hipCheck(launchFill(chunk_cards, 0, t.n_chunks, ctx.stream), "clear chunk cardinalities");
   }
   if (ctx.profiling) hipCheck(hipEventRecord(ctx.prof_events[2], ctx.stream), "hipEventRecord");
   hipCheck(launchFilter(t.view, reinterpret_cast<const ProgramDesc*>(dev_blob),
                         reinterpret_cast<const Instr*>(dev_blob + sizeof(ProgramDesc)),
                         reinterpret_cast<const uint32_t*>(dev_blob + sizeof(ProgramDesc) + program.code.size() * sizeof(Instr)),
                         1, depth, words, card, chunk_cards, 0, ctx.stream),
            "filter kernel");
   if (ctx.profiling) {
      hipCheck(hipEventRecord(ctx.prof_events[3], ctx.stream), "hipEventRecord");
      ctx.profileAccumulate(1);
   }
}

DeviceBitmap runProgram(Table& t, const Emitter& program) {
   Context& ctx = *t.ctx;
   programDepth(ctx, program);
   DeviceBitmap out(&t);
   const std::vector<uint8_t> blob = programBlob(program);
   auto* dev = static_cast<uint8_t*>(ctx.stage(blob.data(), blob.size()));
   launchProgram(t, program, dev, out.words, out.card_dev, out.chunk_cards);
   return out;
}

DeviceBitmap Operator::evaluate() const {
   Emitter e;
   emit(e);
   return runProgram(*table, e);
}

// ---- predicates --------------------------------------------------------------------------------
std::string CompareToValueSelection::toString() const {
   return "$" + table_->ints[column_]->name + " " + cmpText(cmp_) + " " + std::to_string(value_);
}

// selection.h:133-160
std::unique_ptr<Predicate> CompareToValueSelection::negate() const {
   Comparator n = cmp_;
   switch (cmp_) {
      case Comparator::EQUALS: n = Comparator::NOT_EQUALS; break;
      case Comparator::NOT_EQUALS: n = Comparator::EQUALS; break;
      case Comparator::LESS: n = Comparator::HIGHER_OR_EQUALS; break;
      case Comparator::HIGHER_OR_EQUALS: n = Comparator::LESS; break;
      case Comparator::HIGHER: n = Comparator::LESS_OR_EQUALS; break;
      case Comparator::LESS_OR_EQUALS: n = Comparator::HIGHER; break;
   }
   return std::make_unique<CompareToValueSelection>(table_, column_, n, value_, !with_nulls_);
}

void CompareToValueSelection::emit(Emitter& e) const {
   e.leaf(Instr{OP_CMP, static_cast<uint32_t>(column_), cmpCode(cmp_), static_cast<uint32_t>(value_), with_nulls_ ? 1u : 0u, 0, 0, 0});
}

bool CompareToValueSelection::emitPair(Emitter& e, const Predicate* other) const {
   const auto* o = dynamic_cast<const CompareToValueSelection*>(other);
   if (o == nullptr || o->column_ != column_) return false;
   e.leaf(Instr{OP_CMP, static_cast<uint32_t>(column_), cmpCode(cmp_), static_cast<uint32_t>(value_), with_nulls_ ? 1u : 0u,
                cmpCode(o->cmp_), static_cast<uint32_t>(o->value_), 1u | (o->with_nulls_ ? 2u : 0u)});
   return true;
}

uint64_t CompareToValueSelection::tableCount(const Table& t) const {
   const IntColumn& c = *t.ints[column_];
   const auto& v = c.sorted_values;
   const auto lo = [&](int32_t x) { return static_cast<uint64_t>(std::lower_bound(v.begin(), v.end(), x) - v.begin()); };
   const auto hi = [&](int32_t x) { return static_cast<uint64_t>(std::upper_bound(v.begin(), v.end(), x) - v.begin()); };
   uint64_t n = 0;
   switch (cmp_) {
      case Comparator::EQUALS: n = hi(value_) - lo(value_); break;
      case Comparator::NOT_EQUALS: n = v.size() - (hi(value_) - lo(value_)); break;
      case Comparator::LESS: n = lo(value_); break;
      case Comparator::LESS_OR_EQUALS: n = hi(value_); break;
      case Comparator::HIGHER: n = v.size() - hi(value_); break;
      case Comparator::HIGHER_OR_EQUALS: n = v.size() - lo(value_); break;
   }
   return n + (with_nulls_ ? c.null_count : 0);
}

uint64_t IsInCoveredRegion::tableCount(const Table& t) const {
   const uint64_t covered = t.seq[column_]->coverage[position_];
   return cmp_ == Comparator::IS_COVERED ? covered : t.num_rows - covered;
}

std::string IsInCoveredRegion::toString() const {
   return std::string(cmp_ == Comparator::IS_COVERED ? "" : "!") + "IsInCoveredRegion(" + std::to_string(position_) + ")";
}

std::unique_ptr<Predicate> IsInCoveredRegion::negate() const {
   return std::make_unique<IsInCoveredRegion>(
      table_, column_, position_, cmp_ == Comparator::IS_COVERED ? Comparator::IS_NOT_COVERED : Comparator::IS_COVERED
   );
}

// is_in_covered_region.cpp:53-62: the coverage bitmap, complemented within the layout when negated
void IsInCoveredRegion::emit(Emitter& e) const {
   e.leaf(Instr{OP_COVERED, static_cast<uint32_t>(column_), position_, 0, 0, 0, 0, 0});
   if (cmp_ == Comparator::IS_NOT_COVERED) e.unary(OP_NOT);
}

// ---- operators ---------------------------------------------------------------------------------
std::string IndexScan::toString() const {
   if (source.kind == BitmapSource::STORE) return "IndexScan(store " + std::to_string(source.store_id) + ")";
   return "IndexScan(" + table->seq[source.column]->name + ":" + std::to_string(source.position + 1) + " mask " +
          std::to_string(source.symbol_mask) + ")";
}

void IndexScan::emit(Emitter& e) const {
   if (source.kind == BitmapSource::STORE) {
      e.leaf(Instr{OP_BITMAP, source.store_id, e.prefetchSlot((1u << 31) | source.store_id), 0, 0, 0, 0, 0});
      return;
   }
   const SeqColumn& col = *table->seq[source.column];
   const uint32_t p = source.position;
   const uint32_t b = col.block_of_pos[p];
   const uint32_t base = col.block_slot_start[b];
   // the position's slots (at most 32: one per symbol and boundary kind) whose symbol is in the
   // mask, as bits relative to the position's first slot: the kernel tests an entry's slot
   // without looking its symbol up
   uint32_t slot_mask = 0;
   for (uint32_t s = col.pos_slot_ptr[p]; s < col.pos_slot_ptr[p + 1]; ++s) {
      const uint32_t sym = col.slot_info[s] >> 16;
      if (sym < 32 && ((source.symbol_mask >> sym) & 1u)) slot_mask |= 1u << (s - col.pos_slot_ptr[p]);
   }
   e.leaf(Instr{OP_SYMBOLS, static_cast<uint32_t>(source.column), b, col.pos_slot_ptr[p] - base,
                col.pos_slot_ptr[p + 1] - base, slot_mask, p,
                e.prefetchSlot((static_cast<uint32_t>(source.column) << 16) | p)});
}

// intersection.cpp:18-41
Intersection::Intersection(Table* t, OperatorVector&& c, OperatorVector&& n)
    : Operator(t), children(std::move(c)), negated_children(std::move(n)) {
   if (children.empty()) {
      fail(RDB_ECOMPILE, "Compilation bug: Intersection without non-negated children is not allowed. Should be compiled as a union.");
   }
   if (children.size() + negated_children.size() < 2) {
      fail(RDB_ECOMPILE, "Compilation bug: Intersection needs at least two children.");
   }
}

std::string Intersection::toString() const {
   return "Intersection(non_negated: (" + joinChildren(children) + ") negated: (" + joinChildren(negated_children) + ") )";
}

// intersection.cpp:59-106: AND of the children, then AND NOT of the negated children
void Intersection::emit(Emitter& e) const {
   std::vector<size_t> skips;
   children[0]->emit(e);
   for (size_t i = 1; i < children.size(); ++i) {
      if (short_circuit) skips.push_back(e.skip(OP_SKIP_IF_EMPTY));
      children[i]->emit(e);
      e.binary(OP_AND);
   }
   for (const auto& n : negated_children) {
      if (short_circuit) skips.push_back(e.skip(OP_SKIP_IF_EMPTY));
      n->emit(e);
      e.binary(OP_ANDNOT);
   }
   for (size_t at : skips) e.patch(at, e.code.size());
}

// complement.cpp:51-56
void Complement::emit(Emitter& e) const {
   child->emit(e);
   e.unary(OP_NOT);
}

// complement.cpp:23-49
OperatorPtr Complement::fromDeMorgan(Table* t, OperatorVector disjunction) {
   OperatorVector non_negated, negated;
   for (auto& c : disjunction) {
      if (c->type() == COMPLEMENT) {
         negated.push_back(Operator::negate(std::move(c)));
      } else {
         non_negated.push_back(std::move(c));
      }
   }
   auto inter = std::make_unique<Intersection>(t, std::move(negated), std::move(non_negated));
   return std::make_unique<Complement>(t, std::move(inter));
}

std::string RangeSelection::toString() const {
   std::string s = "RangeSelection(";
   for (size_t i = 0; i < ranges.size(); ++i) {
      if (i) s += ", ";
      s += std::to_string(ranges[i].start) + "-" + std::to_string(ranges[i].end);
   }
   return s + ")";
}

// range_selection.cpp:54-87: a range spanning chunks covers the populated part of each chunk
void RangeSelection::emit(Emitter& e) const {
   const uint32_t first = static_cast<uint32_t>(e.ranges.size() / 2);
   auto add = [&](uint64_t lo, uint64_t hi) {
      if (lo < hi) {
         e.ranges.push_back(static_cast<uint32_t>(lo));
         e.ranges.push_back(static_cast<uint32_t>(hi));
      }
   };
   for (const auto& r : ranges) {
      const uint32_t sc = r.start >> 16, ec = r.end >> 16;
      if (sc == ec) {
         add(r.start, r.end);
         continue;
      }
      const uint64_t end_of_start = (static_cast<uint64_t>(sc) << 16) + (sc < table->n_chunks ? table->chunk_sizes[sc] : 0);
      if (r.start != end_of_start) add(r.start, end_of_start);
      for (uint32_t c = sc + 1; c < ec; ++c) add(static_cast<uint64_t>(c) << 16, (static_cast<uint64_t>(c) << 16) + table->chunk_sizes[c]);
      add(static_cast<uint64_t>(ec) << 16, r.end);
   }
   const uint32_t count = static_cast<uint32_t>(e.ranges.size() / 2) - first;
   e.leaf(Instr{OP_RANGES, first, count, 0, 0, 0, 0, 0});
}

// selection.cpp:17-29 (most selective predicate first)
Selection::Selection(Table* t, std::optional<OperatorPtr> child, PredicateVector&& p)
    : Operator(t), child_operator(std::move(child)), predicates(std::move(p)) {
   std::stable_sort(predicates.begin(), predicates.end(), [](const auto& a, const auto& b) {
      return a->estimateSelectivity() < b->estimateSelectivity();
   });
}

std::string Selection::toString() const {
   std::string s = "Select[";
   for (size_t i = 0; i < predicates.size(); ++i) {
      if (i) s += ",";
      s += predicates[i]->toString();
   }
   s += "](";
   if (child_operator.has_value()) s += (*child_operator)->toString();
   return s + ")";
}

// selection.cpp:94-141: rows of the child (when present) that satisfy every predicate
void Selection::emit(Emitter& e) const {
   if (predicates.empty()) fail(RDB_ECOMPILE, "Selection without predicates");
   // predicates are conjoined; two adjacent comparisons of one column become one instruction
   auto emitFrom = [&](size_t i) -> size_t {
      const auto* c = dynamic_cast<const CompareToValueSelection*>(predicates[i].get());
      if (c != nullptr && i + 1 < predicates.size() && c->emitPair(e, predicates[i + 1].get())) return i + 2;
      predicates[i]->emit(e);
      return i + 1;
   };
   size_t i = 0;
   if (child_operator.has_value()) {
      (*child_operator)->emit(e);
   } else {
      i = emitFrom(0);
   }
   while (i < predicates.size()) {
      i = emitFrom(i);
      e.binary(OP_AND);
   }
}

// threshold.cpp:18-41
Threshold::Threshold(Table* t, OperatorVector&& a, OperatorVector&& b, uint32_t n, bool exact)
    : Operator(t), non_negated_children(std::move(a)), negated_children(std::move(b)), number_of_matchers(n),
      match_exactly(exact) {
   if (number_of_matchers >= non_negated_children.size() + negated_children.size()) {
      fail(RDB_ECOMPILE, "Compilation Error: number_of_matchers must be less than the number of children of a threshold expression");
   }
   if (number_of_matchers == 0) fail(RDB_ECOMPILE, "Compilation Error: number_of_matchers must be greater than zero");
}

std::string Threshold::toString() const {
   return std::string("Threshold(") + (match_exactly ? "=" : ">=") + std::to_string(number_of_matchers) +
          "-of non_negated: (" + joinChildren(non_negated_children) + ") negated: (" + joinChildren(negated_children) + ") )";
}

// Bit-sliced counter over the children in the DP's order (non-negated first), see OP_TH_ADD.
void Threshold::emit(Emitter& e) const {
   const uint32_t m = static_cast<uint32_t>(non_negated_children.size() + negated_children.size());
   const uint32_t planes = 32u - static_cast<uint32_t>(std::countl_zero(m));
   e.thresholdBegin(planes);
   // with short-circuit, before each later child: skip the rest once every row's count c is such
   // that c + (children left) cannot change the comparison (each child adds at most one)
   std::vector<size_t> skips;
   uint32_t done = 0;
   auto skipPoint = [&] {
      if (!short_circuit || done == 0) return;
      e.code.push_back(Instr{OP_TH_SKIP, 0, planes, number_of_matchers, match_exactly ? 1u : 0u, m - done, 0, 0});
      skips.push_back(e.code.size() - 1);
   };
   for (const auto& c : non_negated_children) {
      skipPoint();
      c->emit(e);
      e.thresholdAdd(planes, false);
      ++done;
   }
   for (const auto& c : negated_children) {
      skipPoint();
      c->emit(e);
      e.thresholdAdd(planes, true);
      ++done;
   }
   for (size_t at : skips) e.patch(at, e.code.size());
   e.thresholdEnd(planes, number_of_matchers, match_exactly);
}

std::string Union::toString() const { return "(" + joinChildren(children, " | ") + ")"; }

// union.cpp:34-42
void Union::emit(Emitter& e) const {
   if (children.empty()) {
      e.leaf(Instr{OP_EMPTY, 0, 0, 0, 0, 0, 0, 0});
      return;
   }
   std::vector<size_t> skips;
   children[0]->emit(e);
   for (size_t i = 1; i < children.size(); ++i) {
      if (short_circuit) skips.push_back(e.skip(OP_SKIP_IF_FULL));
      children[i]->emit(e);
      e.binary(OP_OR);
   }
   for (size_t at : skips) e.patch(at, e.code.size());
}

// operator.cpp:23-67 with the per-type negate rules
OperatorPtr Operator::negate(OperatorPtr&& op) {
   Table* t = op->table;
   switch (op->type()) {
      case EMPTY:
         return std::make_unique<Full>(t);
      case FULL:
         return std::make_unique<Empty>(t);
      case COMPLEMENT:
         return std::move(static_cast<Complement*>(op.get())->child);
      case RANGE_SELECTION: {
         // range_selection.cpp:89-115
         auto* rs = static_cast<RangeSelection*>(op.get());
         std::vector<RangeSelection::Range> nr;
         if (t->n_chunks == 0) return std::make_unique<RangeSelection>(t, nr);
         uint32_t last_end = 0;
         for (const auto& r : rs->ranges) {
            if (last_end != r.start) nr.push_back({last_end, r.start});
            last_end = r.end;
         }
         const uint32_t ranges_end = t->n_chunks << 16;
         if (last_end != ranges_end) nr.push_back({last_end, ranges_end});
         return std::make_unique<RangeSelection>(t, std::move(nr));
      }
      case SELECTION: {
         // selection.cpp:143-151
         auto* sel = static_cast<Selection*>(op.get());
         if (!sel->child_operator.has_value() && sel->predicates.size() == 1) {
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

// ---- expressions: parsing ----------------------------------------------------------------------
ExpressionPtr Expression::shallowCopy() const {
   auto e = std::make_unique<Expression>(kind);
   e->column = column;
   e->position = position;
   e->symbol_char = symbol_char;
   e->symbol_chars = symbol_chars;
   e->symbols = symbols;
   e->number = number;
   e->exact = exact;
   e->value = value;
   e->sublineage = sublineage;
   e->from = from;
   e->to = to;
   e->int_value = int_value;
   e->bitmap_id = bitmap_id;
   e->row_ranges = row_ranges;
   e->predicates = predicates;
   return e;
}

ExpressionPtr Expression::clone() const {
   auto e = shallowCopy();
   for (const auto& c : children) e->children.push_back(c->clone());
   return e;
}

namespace {

class Parser {
  public:
   explicit Parser(const std::string& s) : s_(s) {}

   ExpressionPtr parse() {
      auto e = expr();
      skip();
      if (i_ != s_.size()) bad("trailing characters");
      return e;
   }

  private:
   const std::string& s_;
   size_t i_ = 0;

   [[noreturn]] void bad(const std::string& what) { fail(RDB_EQUERY, "malformed filter: " + what); }

   void skip() {
      while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
   }
   bool at(char c) {
      skip();
      return i_ < s_.size() && s_[i_] == c;
   }
   void expect(char c) {
      if (!at(c)) bad(std::string("expected '") + c + "'");
      ++i_;
   }
   // {quoted, text}
   std::pair<bool, std::string> atom() {
      skip();
      if (i_ >= s_.size()) bad("unexpected end");
      if (s_[i_] == '"') {
         ++i_;
         std::string out;
         while (i_ < s_.size() && s_[i_] != '"') {
            if (s_[i_] == '\\' && i_ + 1 < s_.size()) ++i_;
            out.push_back(s_[i_++]);
         }
         if (i_ >= s_.size()) bad("unterminated string");
         ++i_;
         return {true, out};
      }
      const size_t b = i_;
      while (i_ < s_.size() && !std::isspace(static_cast<unsigned char>(s_[i_])) && s_[i_] != '(' && s_[i_] != ')') ++i_;
      if (b == i_) bad("empty atom");
      return {false, s_.substr(b, i_ - b)};
   }
   int64_t integer() {
      const auto [q, t] = atom();
      if (q) bad("expected an integer");
      char* end = nullptr;
      errno = 0;
      const long long v = std::strtoll(t.c_str(), &end, 10);
      if (errno != 0 || end == t.c_str() || *end != '\0') bad("bad integer '" + t + "'");
      return v;
   }
   std::optional<int64_t> optionalInteger() {
      skip();
      const size_t save = i_;
      const auto [q, t] = atom();
      if (!q && t == "_") return std::nullopt;
      i_ = save;
      return integer();
   }
   std::optional<std::string> optionalString() {
      const auto [q, t] = atom();
      if (!q && t == "null") return std::nullopt;
      return t;
   }
   uint32_t position() {
      const int64_t v = integer();
      if (v < 0 || v > std::numeric_limits<uint32_t>::max()) bad("position out of range");
      return static_cast<uint32_t>(v);
   }

   ExpressionPtr expr() {
      expect('(');
      const std::string op = atom().second;
      ExpressionPtr e;
      auto rest = [&](Expression& x) {
         while (!at(')')) x.children.push_back(expr());
      };
      using K = Expression::Kind;
      if (op == "true") {
         e = std::make_unique<Expression>(K::TRUE_);
      } else if (op == "false") {
         e = std::make_unique<Expression>(K::FALSE_);
      } else if (op == "and" || op == "or") {
         e = std::make_unique<Expression>(op == "and" ? K::AND : K::OR);
         rest(*e);
      } else if (op == "not") {
         e = std::make_unique<Expression>(K::NOT);
         e->children.push_back(expr());
      } else if (op == "nof") {
         e = std::make_unique<Expression>(K::NOF);
         e->number = static_cast<int>(integer());
         e->exact = integer() != 0;
         rest(*e);
      } else if (op == "sym_eq") {
         e = std::make_unique<Expression>(K::SYMBOL_EQUALS);
         e->column = atom().second;
         e->position = position();
         const auto t = atom().second;
         if (t.size() != 1) bad("a symbol is one character");
         e->symbol_char = t[0];
      } else if (op == "has_mut") {
         e = std::make_unique<Expression>(K::HAS_MUTATION);
         e->column = atom().second;
         e->position = position();
      } else if (op == "sym_in") {
         e = std::make_unique<Expression>(K::SYMBOL_IN_SET);
         e->column = atom().second;
         e->position = position();
         e->symbol_chars = atom().second;
      } else if (op == "str_eq") {
         e = std::make_unique<Expression>(K::STRING_EQUALS);
         e->column = atom().second;
         e->value = optionalString();
      } else if (op == "lineage") {
         e = std::make_unique<Expression>(K::LINEAGE);
         e->column = atom().second;
         e->value = optionalString();
         e->sublineage = integer() != 0;
      } else if (op == "date_between" || op == "int_between") {
         e = std::make_unique<Expression>(op == "date_between" ? K::DATE_BETWEEN : K::INT_BETWEEN);
         e->column = atom().second;
         e->from = optionalInteger();
         e->to = optionalInteger();
      } else if (op == "int_eq") {
         e = std::make_unique<Expression>(K::INT_EQUALS);
         e->column = atom().second;
         e->int_value = integer();
      } else if (op == "bitmap") {
         e = std::make_unique<Expression>(K::BITMAP);
         e->bitmap_id = position();
      } else if (op == "covered") {
         e = std::make_unique<Expression>(K::COVERED);
         e->column = atom().second;
         e->position = position();
      } else if (op == "op_threshold") {
         e = std::make_unique<Expression>(K::OP_THRESHOLD);
         e->number = static_cast<int>(integer());
         e->exact = integer() != 0;
         rest(*e);
      } else if (op == "op_intersection" || op == "op_union") {
         e = std::make_unique<Expression>(op == "op_union" ? K::OP_UNION : K::OP_INTERSECTION);
         rest(*e);
      } else if (op == "op_complement") {
         e = std::make_unique<Expression>(K::OP_COMPLEMENT);
         e->children.push_back(expr());
      } else if (op == "op_range") {
         // RangeSelection over global row-id ranges: (op_range start end start end ...)
         e = std::make_unique<Expression>(K::OP_RANGES);
         while (!at(')')) e->row_ranges.push_back(position());
         if (e->row_ranges.size() % 2 != 0) bad("op_range takes start / end pairs");
      } else if (op == "op_select") {
         // Selection with an optional child and (cmp COLUMN OP VALUE) predicates, OP one of
         // eq lt gt le ge ne
         e = std::make_unique<Expression>(K::OP_SELECT);
         while (!at(')')) {
            const size_t save = i_;
            expect('(');
            const auto [quoted, head] = atom();
            if (!quoted && head == "cmp") {
               Expression::ValuePredicate p;
               p.column = atom().second;
               const std::string c = atom().second;
               static const char* const kCmp[] = {"eq", "lt", "gt", "le", "ge", "ne"};
               p.comparator = -1;
               for (int k = 0; k < 6; ++k) {
                  if (c == kCmp[k]) p.comparator = k;
               }
               if (p.comparator < 0) bad("unknown comparator '" + c + "'");
               p.value = integer();
               expect(')');
               e->predicates.push_back(std::move(p));
            } else {
               i_ = save;
               if (!e->children.empty()) bad("op_select takes at most one child");
               e->children.push_back(expr());
            }
         }
      } else {
         bad("unknown expression '" + op + "'");
      }
      expect(')');
      return e;
   }
};

}  // namespace

ExpressionPtr parseFilter(const std::string& text) { return Parser(text).parse(); }

// ---- expressions: rewrite ----------------------------------------------------------------------
ExpressionPtr rewrite(const Table& t, const Expression& e) {
   using K = Expression::Kind;
   switch (e.kind) {
      case K::AND:
      case K::NOF:
      case K::NOT:
      case K::OP_THRESHOLD:
      case K::OP_INTERSECTION:
      case K::OP_UNION:
      case K::OP_COMPLEMENT:
      case K::OP_RANGES:
      case K::OP_SELECT: {
         // and.cpp:139-147, nof.cpp:249-257 (mode NONE keeps an exact NOf), negation.cpp
         auto r = e.shallowCopy();
         for (const auto& c : e.children) r->children.push_back(rewrite(t, *c));
         return r;
      }
      case K::OR: {
         // or.cpp:173-192
         std::vector<const Expression*> queue, collected;
         for (const auto& c : e.children) queue.push_back(c.get());
         while (!queue.empty()) {
            const Expression* cur = queue.back();
            queue.pop_back();
            if (cur->kind == K::OR) {
               for (const auto& c : cur->children) queue.push_back(c.get());
            } else {
               collected.push_back(cur);
            }
         }
         std::vector<ExpressionPtr> unprocessed;
         for (const Expression* c : collected) unprocessed.push_back(rewrite(t, *c));
         std::vector<ExpressionPtr> non_trivial;
         while (!unprocessed.empty()) {
            auto c = std::move(unprocessed.back());
            unprocessed.pop_back();
            if (c->kind == K::TRUE_) {
               non_trivial.clear();
               non_trivial.push_back(std::make_unique<Expression>(K::TRUE_));
               break;
            }
            if (c->kind == K::FALSE_) continue;
            if (c->kind == K::OR) {
               for (auto& cc : c->children) unprocessed.push_back(std::move(cc));
            } else {
               non_trivial.push_back(std::move(c));
            }
         }
         std::vector<ExpressionPtr> merged;
         std::map<std::pair<std::string, uint32_t>, std::vector<uint8_t>> symbol_sets;
         for (auto& c : non_trivial) {
            if (c->kind == K::SYMBOL_IN_SET) {
               auto& v = symbol_sets[{c->column, c->position}];
               v.insert(v.end(), c->symbols.begin(), c->symbols.end());
            } else {
               merged.push_back(std::move(c));
            }
         }
         for (auto& [key, syms] : symbol_sets) {
            auto x = std::make_unique<Expression>(K::SYMBOL_IN_SET);
            x->column = key.first;
            x->position = key.second;
            x->symbols = syms;
            merged.push_back(std::move(x));
         }
         if (merged.size() == 1) return std::move(merged[0]);
         auto r = std::make_unique<Expression>(K::OR);
         r->children = std::move(merged);
         return r;
      }
      case K::SYMBOL_EQUALS: {
         // symbol_equals.cpp:65-93
         const SeqColumn& col = t.seqColumn(e.column);
         if (e.position >= col.genome_length) {
            fail(RDB_EQUERY, "SymbolEquals position is out of bounds " + std::to_string(e.position + 1) + " > " +
                                std::to_string(col.genome_length));
         }
         int sym;
         if (e.symbol_char == '.') {
            sym = col.reference[e.position];
         } else {
            sym = col.abc->fromChar(e.symbol_char);
            if (sym < 0) fail(RDB_EQUERY, std::string("illegal symbol '") + e.symbol_char + "'");
         }
         auto r = std::make_unique<Expression>(K::SYMBOL_IN_SET);
         r->column = e.column;
         r->position = e.position;
         r->symbols = {static_cast<uint8_t>(sym)};
         return r;
      }
      case K::HAS_MUTATION: {
         // has_mutation.cpp:34-67
         const SeqColumn& col = t.seqColumn(e.column);
         if (e.position >= col.genome_length) {
            fail(RDB_EQUERY, "HasMutation position is out of bounds " + std::to_string(e.position + 1) + " > " +
                                std::to_string(col.genome_length));
         }
         const auto& amb = col.abc->ambiguity[col.reference[e.position]];
         auto r = std::make_unique<Expression>(K::SYMBOL_IN_SET);
         r->column = e.column;
         r->position = e.position;
         for (uint8_t s = 0; s < col.abc->count; ++s) {
            if (std::find(amb.begin(), amb.end(), s) == amb.end()) r->symbols.push_back(s);
         }
         return r;
      }
      case K::SYMBOL_IN_SET: {
         auto r = e.shallowCopy();
         if (!e.symbol_chars.empty()) {
            const SeqColumn& col = t.seqColumn(e.column);
            r->symbols.clear();
            for (char ch : e.symbol_chars) {
               const int s = col.abc->fromChar(ch);
               if (s < 0) fail(RDB_EQUERY, std::string("illegal symbol '") + ch + "'");
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

// ---- expressions: compile ------------------------------------------------------------------------
namespace {

OperatorPtr makeDifference(Table* t, OperatorPtr left, OperatorPtr right) {
   OperatorVector a, b;
   a.push_back(std::move(left));
   b.push_back(std::move(right));
   return std::make_unique<Intersection>(t, std::move(a), std::move(b));
}

uint32_t maskOf(const std::vector<uint8_t>& symbols) {
   uint32_t m = 0;
   for (uint8_t s : symbols) m |= 1u << s;
   return m;
}

// symbol_in_set.cpp:95-226 and :231-264
OperatorPtr compileSymbolInSet(Table& t, int column, uint32_t position, const std::vector<uint8_t>& symbols) {
   const SeqColumn& col = *t.seq[column];
   if (position >= col.genome_length) {
      fail(RDB_EQUERY, "SymbolInSet position is out of bounds " + std::to_string(position + 1) + " > " +
                          std::to_string(col.genome_length));
   }
   const uint32_t all = (col.abc->count >= 32) ? 0xFFFFFFFFu : ((1u << col.abc->count) - 1u);
   const uint32_t mask = maskOf(symbols);
   const uint8_t local = col.local_ref[position];
   const uint32_t missing_bit = 1u << col.abc->missing;
   const bool includes_reference = (mask >> local) & 1u;
   const bool includes_missing = (mask & missing_bit) != 0;
   auto scan = [&](uint32_t m) {
      return std::make_unique<IndexScan>(&t, BitmapSource{BitmapSource::SYMBOLS, 0, column, position, m});
   };
   auto exclude_nulls = [&](OperatorPtr op) -> OperatorPtr {
      if (!col.has_nulls) return op;
      return makeDifference(&t, std::move(op),
                            std::make_unique<IndexScan>(&t, BitmapSource{BitmapSource::STORE, col.null_bitmap_id}));
   };
   if (includes_reference && includes_missing) {
      return exclude_nulls(std::make_unique<Complement>(&t, scan(all & ~mask)));
   }
   if (includes_missing) {
      OperatorVector u;
      PredicateVector p;
      p.push_back(std::make_unique<IsInCoveredRegion>(&t, column, position, IsInCoveredRegion::Comparator::IS_NOT_COVERED));
      u.push_back(std::make_unique<Selection>(&t, std::nullopt, std::move(p)));
      u.push_back(scan(mask));
      return exclude_nulls(std::make_unique<Union>(&t, std::move(u)));
   }
   if (includes_reference) {
      PredicateVector p;
      p.push_back(std::make_unique<IsInCoveredRegion>(&t, column, position, IsInCoveredRegion::Comparator::IS_COVERED));
      return makeDifference(&t, std::make_unique<Selection>(&t, std::nullopt, std::move(p)), scan(all & ~mask & ~missing_bit));
   }
   return scan(mask);
}

// nof.cpp:16-139
OperatorPtr nofToOperator(Table* t, int n, OperatorVector a, OperatorVector b, bool exact) {
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

}  // namespace

OperatorPtr compile(Table& t, const Expression& e) {
   using K = Expression::Kind;
   Table* tp = &t;
   switch (e.kind) {
      case K::TRUE_:
         return std::make_unique<Full>(tp);
      case K::FALSE_:
         return std::make_unique<Empty>(tp);
      case K::AND: {
         // and.cpp:91-219
         OperatorVector unprocessed;
         for (const auto& c : e.children) unprocessed.push_back(compile(t, *c));
         OperatorVector pos, neg;
         PredicateVector preds;
         while (!unprocessed.empty()) {
            auto c = std::move(unprocessed.back());
            unprocessed.pop_back();
            if (c->type() == FULL) continue;
            if (c->type() == EMPTY) return std::make_unique<Empty>(tp);
            if (c->type() == INTERSECTION) {
               auto* x = static_cast<Intersection*>(c.get());
               for (auto& cc : x->children) pos.push_back(std::move(cc));
               for (auto& cc : x->negated_children) neg.push_back(std::move(cc));
            } else if (c->type() == COMPLEMENT) {
               neg.push_back(Operator::negate(std::move(c)));
            } else if (c->type() == SELECTION) {
               auto* s = static_cast<Selection*>(c.get());
               for (auto& p : s->predicates) preds.push_back(std::move(p));
               if (s->child_operator.has_value()) unprocessed.push_back(std::move(*s->child_operator));
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
      case K::OR: {
         // or.cpp:194-241
         OperatorVector filtered;
         for (const auto& ce : e.children) {
            auto c = compile(t, *ce);
            if (c->type() == EMPTY) continue;
            if (c->type() == FULL) return std::make_unique<Full>(tp);
            if (c->type() == UNION) {
               for (auto& cc : static_cast<Union*>(c.get())->children) filtered.push_back(std::move(cc));
            } else {
               filtered.push_back(std::move(c));
            }
         }
         if (filtered.empty()) return std::make_unique<Empty>(tp);
         if (filtered.size() == 1) return std::move(filtered[0]);
         if (std::any_of(filtered.begin(), filtered.end(), [](const auto& c) { return c->type() == COMPLEMENT; })) {
            return Complement::fromDeMorgan(tp, std::move(filtered));
         }
         return std::make_unique<Union>(tp, std::move(filtered));
      }
      case K::NOT:
         return Operator::negate(compile(t, *e.children[0]));  // negation.cpp:32-34
      case K::NOF: {
         // nof.cpp:181-214, 259-277
         OperatorVector a, b;
         int n = e.number;
         for (const auto& ce : e.children) {
            auto op = compile(t, *ce);
            if (op->type() == EMPTY) continue;
            if (op->type() == FULL) {
               --n;
               continue;
            }
            if (op->type() == COMPLEMENT) {
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
      case K::SYMBOL_IN_SET: {
         const int idx = t.seqIndex(e.column);
         if (idx < 0) fail(RDB_EQUERY, "Database does not contain the sequence '" + e.column + "'");
         return compileSymbolInSet(t, idx, e.position, e.symbols);
      }
      case K::STRING_EQUALS: {
         // equals.cpp:158-187
         const DictColumn* col = t.dictColumn(e.column);
         if (!col) fail(RDB_EQUERY, "The database does not contain the column '" + e.column + "'");
         uint32_t id = kInvalidId;
         if (!e.value.has_value()) {
            if (col->has_nulls) id = col->null_bitmap_id;
         } else {
            auto it = col->ids.find(*e.value);
            if (it != col->ids.end()) id = col->value_bitmap[it->second];
         }
         if (id == kInvalidId) return std::make_unique<Empty>(tp);
         return std::make_unique<IndexScan>(tp, BitmapSource{BitmapSource::STORE, id});
      }
      case K::LINEAGE: {
         // lineage_filter.cpp:30-100
         const DictColumn* col = t.dictColumn(e.column);
         if (!col || !col->lineage) {
            fail(RDB_EQUERY, "The database does not contain a lineage index for the column '" + e.column + "'");
         }
         uint32_t id;
         if (!e.value.has_value()) {
            id = col->null_bitmap_id;
         } else {
            auto it = col->ids.find(*e.value);
            if (it == col->ids.end()) {
               fail(RDB_EQUERY, "The lineage '" + *e.value + "' is not a valid lineage for column '" + e.column + "'.");
            }
            id = e.sublineage ? col->sublineage_bitmap[it->second] : col->lineage_bitmap[it->second];
         }
         if (id == kInvalidId) return std::make_unique<Empty>(tp);
         return std::make_unique<IndexScan>(tp, BitmapSource{BitmapSource::STORE, id});
      }
      case K::DATE_BETWEEN: {
         // date_between.cpp:61-135
         const int idx = t.intIndex(e.column);
         if (idx < 0 || !t.ints[idx]->is_date) fail(RDB_EQUERY, "The column '" + e.column + "' is not of type date");
         const IntColumn& col = *t.ints[idx];
         const int32_t from = e.from.has_value() ? static_cast<int32_t>(*e.from) : std::numeric_limits<int32_t>::min();
         const int32_t to = e.to.has_value() ? static_cast<int32_t>(*e.to) : std::numeric_limits<int32_t>::max();
         if (col.is_sorted) {
            std::vector<RangeSelection::Range> ranges;
            for (uint32_t c = 0; c < t.n_chunks; ++c) {
               const int32_t* b = col.values.data() + t.chunk_offset[c];
               const int32_t* en = b + t.chunk_sizes[c];
               const auto lo = std::lower_bound(b, en, from) - b;
               const auto hi = (e.to.has_value() ? std::upper_bound(b, en, to) : en) - b;
               auto rid = [&](int64_t i) {
                  if (i == static_cast<int64_t>(t.chunk_sizes[c])) return (c + 1) << 16;
                  return (c << 16) | static_cast<uint32_t>(i);
               };
               ranges.push_back({rid(lo), rid(hi)});
            }
            return std::make_unique<RangeSelection>(tp, std::move(ranges));
         }
         PredicateVector p;
         p.push_back(std::make_unique<CompareToValueSelection>(tp, idx, Comparator::HIGHER_OR_EQUALS, from, false));
         p.push_back(std::make_unique<CompareToValueSelection>(tp, idx, Comparator::LESS_OR_EQUALS, to, false));
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case K::INT_BETWEEN: {
         // int_between.cpp:71-131
         const int idx = t.intIndex(e.column);
         if (idx < 0 || t.ints[idx]->is_date) fail(RDB_EQUERY, "The column '" + e.column + "' is not of type int32 or int64");
         auto check = [&](const std::optional<int64_t>& v) {
            if (v.has_value() && (*v < std::numeric_limits<int32_t>::min() || *v > std::numeric_limits<int32_t>::max())) {
               fail(RDB_EQUERY, "Value out of int32 range");
            }
         };
         check(e.from);
         check(e.to);
         PredicateVector p;
         if (e.from.has_value()) {
            p.push_back(std::make_unique<CompareToValueSelection>(tp, idx, Comparator::HIGHER_OR_EQUALS, static_cast<int32_t>(*e.from), false));
         }
         if (e.to.has_value()) {
            p.push_back(std::make_unique<CompareToValueSelection>(tp, idx, Comparator::LESS_OR_EQUALS, static_cast<int32_t>(*e.to), false));
         }
         if (p.empty()) {
            return std::make_unique<Complement>(tp, std::make_unique<IndexScan>(tp, BitmapSource{BitmapSource::STORE, t.ints[idx]->null_bitmap_id}));
         }
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case K::INT_EQUALS: {
         // equals.cpp:62-99
         const int idx = t.intIndex(e.column);
         if (idx < 0) fail(RDB_EQUERY, "The column '" + e.column + "' is not of type int");
         if (e.int_value < std::numeric_limits<int32_t>::min() || e.int_value > std::numeric_limits<int32_t>::max()) {
            fail(RDB_EQUERY, "Cannot cast " + std::to_string(e.int_value) + " to int32. Value out of range");
         }
         PredicateVector p;
         p.push_back(std::make_unique<CompareToValueSelection>(tp, idx, Comparator::EQUALS, static_cast<int32_t>(e.int_value), false));
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case K::BITMAP: {
         if (e.bitmap_id >= t.literal_bitmaps.size()) fail(RDB_EQUERY, "unknown bitmap literal " + std::to_string(e.bitmap_id));
         return std::make_unique<IndexScan>(tp, BitmapSource{BitmapSource::STORE, t.literal_bitmaps[e.bitmap_id]});
      }
      case K::COVERED: {
         const int idx = t.seqIndex(e.column);
         if (idx < 0) fail(RDB_EQUERY, "Database does not contain the sequence '" + e.column + "'");
         if (e.position >= t.seq[idx]->genome_length) fail(RDB_EQUERY, "position out of bounds");
         PredicateVector p;
         p.push_back(std::make_unique<IsInCoveredRegion>(tp, idx, e.position, IsInCoveredRegion::Comparator::IS_COVERED));
         return std::make_unique<Selection>(tp, std::nullopt, std::move(p));
      }
      case K::OP_THRESHOLD:
      case K::OP_INTERSECTION: {
         OperatorVector a, b;
         for (const auto& ce : e.children) {
            auto op = compile(t, *ce);
            if (op->type() == COMPLEMENT) {
               b.push_back(Operator::negate(std::move(op)));
            } else {
               a.push_back(std::move(op));
            }
         }
         if (e.kind == K::OP_INTERSECTION) return std::make_unique<Intersection>(tp, std::move(a), std::move(b));
         if (e.number < 0) fail(RDB_ECOMPILE, "Compilation Error: number_of_matchers must be greater than zero");
         return std::make_unique<Threshold>(tp, std::move(a), std::move(b), static_cast<uint32_t>(e.number), e.exact);
      }
      case K::OP_UNION: {
         OperatorVector a;
         for (const auto& ce : e.children) a.push_back(compile(t, *ce));
         return std::make_unique<Union>(tp, std::move(a));
      }
      case K::OP_COMPLEMENT:
         return std::make_unique<Complement>(tp, compile(t, *e.children[0]));
      case K::OP_RANGES: {
         // range_selection.cpp:54-60: every end is a row of the layout or the end of the last chunk
         std::vector<RangeSelection::Range> ranges;
         const uint64_t layout_end = static_cast<uint64_t>(t.n_chunks) << 16;
         auto valid = [&](uint32_t id) {
            const uint32_t c = id >> 16;
            return (c < t.n_chunks && (id & 0xFFFFu) < t.chunk_sizes[c]) || id == layout_end;
         };
         for (size_t k = 0; k < e.row_ranges.size(); k += 2) {
            const uint32_t a = e.row_ranges[k], b = e.row_ranges[k + 1];
            if (!valid(a) || !valid(b) || a > b) fail(RDB_ECOMPILE, "RangeSelection: range outside the row layout");
            ranges.push_back({a, b});
         }
         return std::make_unique<RangeSelection>(tp, std::move(ranges));
      }
      case K::OP_SELECT: {
         // selection.h:24-166: Selection(child?, predicates); at least one predicate
         if (e.predicates.empty()) fail(RDB_ECOMPILE, "Selection needs at least one predicate");
         PredicateVector p;
         for (const auto& vp : e.predicates) {
            const int idx = t.intIndex(vp.column);
            if (idx < 0) fail(RDB_EQUERY, "The column '" + vp.column + "' is not of type int or date");
            if (vp.value < std::numeric_limits<int32_t>::min() || vp.value > std::numeric_limits<int32_t>::max()) {
               fail(RDB_EQUERY, "Value out of int32 range");
            }
            p.push_back(std::make_unique<CompareToValueSelection>(tp, idx, static_cast<Comparator>(vp.comparator),
                                                                  static_cast<int32_t>(vp.value), false));
         }
         std::optional<OperatorPtr> child;
         if (!e.children.empty()) child = compile(t, *e.children[0]);
         return std::make_unique<Selection>(tp, std::move(child), std::move(p));
      }
      default:
         fail(RDB_ECOMPILE, "expression must be rewritten before compilation");
   }
}

OperatorPtr compileFilter(Table& t, const std::string& text) {
   auto parsed = parseFilter(text);
   auto rewritten = rewrite(t, *parsed);
   return compile(t, *rewritten);
}

}  // namespace rdb::query
