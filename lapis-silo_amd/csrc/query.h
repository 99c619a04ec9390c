// Filter expressions and filter operators, mirroring the reference's API surface:
//   * expressions: query_engine/scalar_expressions/*.{h,cpp} (rewrite with AmbiguityMode NONE, then
//     compile to an operator tree);
//   * operators: query_engine/filter/operators/*.{h,cpp} (Operator::type / evaluate / toString /
//     negate, and the same construction rules).
// evaluate() runs on the GPU: the operator tree is emitted as post-order stack code and executed by
// the fused filter kernel over every 4096-row slice of the table; the result is a DeviceBitmap.
#pragma once

#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "engine.h"

namespace rdb::query {

// ---- program emission ----------------------------------------------------------------------------
class Emitter {
  public:
   std::vector<Instr> code;
   std::vector<uint32_t> ranges;  // global [lo, hi) pairs
   std::vector<uint32_t> prefetch;  // leaf record codes loaded before the program runs
   uint32_t sp = 0;
   uint32_t max_sp = 0;

   // prefetch slot of a leaf record (the same record shares one slot), kNoPrefetch when full
   uint32_t prefetchSlot(uint32_t code) {
      for (size_t i = 0; i < prefetch.size(); ++i) {
         if (prefetch[i] == code) return static_cast<uint32_t>(i);
      }
      if (prefetch.size() >= kPrefetchSlots) return kNoPrefetch;
      prefetch.push_back(code);
      return static_cast<uint32_t>(prefetch.size() - 1);
   }

   void leaf(const Instr& ins) {
      code.push_back(ins);
      bump(1);
   }
   void binary(uint32_t op) {
      code.push_back(Instr{op, 0, 0, 0, 0, 0, 0, 0});
      --sp;
   }
   void unary(uint32_t op) { code.push_back(Instr{op, 0, 0, 0, 0, 0, 0, 0}); }
   void thresholdBegin(uint32_t planes) {
      code.push_back(Instr{OP_TH_BEGIN, planes, 0, 0, 0, 0, 0, 0});
      bump(planes + 1);
   }
   void thresholdAdd(uint32_t planes, bool negated) {
      code.push_back(Instr{OP_TH_ADD, planes, negated ? 1u : 0u, 0, 0, 0, 0, 0});
      --sp;
   }
   void thresholdEnd(uint32_t planes, uint32_t n, bool exact) {
      code.push_back(Instr{OP_TH_END, planes, n, exact ? 1u : 0u, 0, 0, 0, 0});
      sp -= planes;
   }
   // a forward jump whose target is set by patch(); leaves the stack as it is
   size_t skip(uint32_t op) {
      code.push_back(Instr{op, 0, 0, 0, 0, 0, 0, 0});
      return code.size() - 1;
   }
   void patch(size_t at, size_t target) { code[at].a = static_cast<uint32_t>(target); }

  private:
   void bump(uint32_t k) {
      sp += k;
      if (sp > max_sp) max_sp = sp;
   }
};

// Runs one emitted program over the table.
DeviceBitmap runProgram(Table& table, const Emitter& program);
// The staged form of one program: its ProgramDesc, instructions, ranges and prefetch codes.
std::vector<uint8_t> programBlob(const Emitter& program);
// The operand-stack depth the program needs (errors when it leaves a wrong stack or nests too deep).
uint32_t programDepth(const Context& ctx, const Emitter& program);
// Enqueues the filter kernel for a program staged on the device in programBlob's layout; clears
// and writes the row words, the cardinality (64 bytes) and the per-chunk cardinalities.
void launchProgram(Table& table, const Emitter& program, const uint8_t* dev_blob, uint64_t* words,
                   unsigned long long* card, uint32_t* chunk_cards);

// ---- predicates (filter/operators/selection.h:24-166, is_in_covered_region.h) -------------------
enum class Comparator { EQUALS, LESS, HIGHER, LESS_OR_EQUALS, HIGHER_OR_EQUALS, NOT_EQUALS };

class Predicate {
  public:
   virtual ~Predicate() = default;
   virtual std::string toString() const = 0;
   virtual double estimateSelectivity() const { return 0.5; }
   // rows of the whole table the predicate matches, from host-side column statistics
   virtual uint64_t tableCount(const Table& t) const = 0;
   virtual std::unique_ptr<Predicate> copy() const = 0;
   virtual std::unique_ptr<Predicate> negate() const = 0;
   virtual void emit(Emitter& e) const = 0;
};
using PredicateVector = std::vector<std::unique_ptr<Predicate>>;

class CompareToValueSelection : public Predicate {
  public:
   CompareToValueSelection(const Table* t, int column, Comparator cmp, int32_t value, bool with_nulls)
       : table_(t), column_(column), cmp_(cmp), value_(value), with_nulls_(with_nulls) {}
   std::string toString() const override;
   std::unique_ptr<Predicate> copy() const override {
      return std::make_unique<CompareToValueSelection>(table_, column_, cmp_, value_, with_nulls_);
   }
   std::unique_ptr<Predicate> negate() const override;
   void emit(Emitter& e) const override;
   // one instruction for this predicate and, when `other` compares the same column, that one too
   // (their conjunction: a value range reads the column once)
   bool emitPair(Emitter& e, const Predicate* other) const;
   uint64_t tableCount(const Table& t) const override;

  private:
   const Table* table_;
   int column_;
   Comparator cmp_;
   int32_t value_;
   bool with_nulls_;
};

class IsInCoveredRegion : public Predicate {
  public:
   enum class Comparator { IS_COVERED, IS_NOT_COVERED };
   IsInCoveredRegion(const Table* t, int column, uint32_t position, Comparator cmp)
       : table_(t), column_(column), position_(position), cmp_(cmp) {}
   std::string toString() const override;
   double estimateSelectivity() const override { return 0.1; }
   std::unique_ptr<Predicate> copy() const override {
      return std::make_unique<IsInCoveredRegion>(table_, column_, position_, cmp_);
   }
   std::unique_ptr<Predicate> negate() const override;
   void emit(Emitter& e) const override;
   uint64_t tableCount(const Table& t) const override;

  private:
   const Table* table_;
   int column_;
   uint32_t position_;
   Comparator cmp_;
};

// ---- operators ------------------------------------------------------------------------------------
enum Type : uint8_t { EMPTY, FULL, INDEX_SCAN, INTERSECTION, COMPLEMENT, RANGE_SELECTION, SELECTION, THRESHOLD, UNION };

class Operator;
using OperatorPtr = std::unique_ptr<Operator>;
using OperatorVector = std::vector<OperatorPtr>;

class Operator {
  public:
   explicit Operator(Table* t) : table(t) {}
   virtual ~Operator() = default;
   virtual Type type() const = 0;
   virtual std::string toString() const = 0;
   virtual void emit(Emitter& e) const = 0;
   // Operator::evaluate (filter/operators/operator.h:32) on the device.
   DeviceBitmap evaluate() const;
   // Operator::negate (operator.cpp:23-67).
   static OperatorPtr negate(OperatorPtr&& op);
   Table* table;
};

class Empty : public Operator {
  public:
   using Operator::Operator;
   Type type() const override { return EMPTY; }
   std::string toString() const override { return "Empty"; }
   void emit(Emitter& e) const override { e.leaf(Instr{OP_EMPTY, 0, 0, 0, 0, 0, 0, 0}); }
};

class Full : public Operator {
  public:
   using Operator::Operator;
   Type type() const override { return FULL; }
   std::string toString() const override { return "Full"; }
   void emit(Emitter& e) const override { e.leaf(Instr{OP_FULL, 0, 0, 0, 0, 0, 0, 0}); }
};

// The row set an IndexScan reads: a bitmap-store entry (dictionary value, null bitmap, lineage,
// literal) or the stored containers of a set of symbols at one position of a sequence column (the
// views that getMatchingContainerViews returns, vertical_sequence_index.cpp:205-227).
struct BitmapSource {
   enum Kind { STORE, SYMBOLS } kind = STORE;
   uint32_t store_id = 0;
   int column = -1;
   uint32_t position = 0;
   uint32_t symbol_mask = 0;
};

class IndexScan : public Operator {
  public:
   IndexScan(Table* t, BitmapSource s) : Operator(t), source(s) {}
   Type type() const override { return INDEX_SCAN; }
   std::string toString() const override;
   void emit(Emitter& e) const override;
   BitmapSource source;
};

class Intersection : public Operator {
  public:
   Intersection(Table* t, OperatorVector&& children, OperatorVector&& negated_children);
   Type type() const override { return INTERSECTION; }
   std::string toString() const override;
   void emit(Emitter& e) const override;
   OperatorVector children;
   OperatorVector negated_children;
   // planner: skip the remaining children of a slice once the partial result is empty
   bool short_circuit = false;
};

class Complement : public Operator {
  public:
   Complement(Table* t, OperatorPtr c) : Operator(t), child(std::move(c)) {}
   Type type() const override { return COMPLEMENT; }
   std::string toString() const override { return "!" + child->toString(); }
   void emit(Emitter& e) const override;
   static OperatorPtr fromDeMorgan(Table* t, OperatorVector disjunction);
   OperatorPtr child;
};

class RangeSelection : public Operator {
  public:
   struct Range {
      uint32_t start;  // global row ids (chunk << 16 | row); end exclusive
      uint32_t end;
   };
   RangeSelection(Table* t, std::vector<Range> r) : Operator(t), ranges(std::move(r)) {}
   Type type() const override { return RANGE_SELECTION; }
   std::string toString() const override;
   void emit(Emitter& e) const override;
   std::vector<Range> ranges;
};

class Selection : public Operator {
  public:
   Selection(Table* t, std::optional<OperatorPtr> child, PredicateVector&& predicates);
   Type type() const override { return SELECTION; }
   std::string toString() const override;
   void emit(Emitter& e) const override;
   std::optional<OperatorPtr> child_operator;
   PredicateVector predicates;
};

class Threshold : public Operator {
  public:
   Threshold(Table* t, OperatorVector&& non_negated, OperatorVector&& negated, uint32_t number_of_matchers,
             bool match_exactly);
   Type type() const override { return THRESHOLD; }
   std::string toString() const override;
   void emit(Emitter& e) const override;
   OperatorVector non_negated_children;
   OperatorVector negated_children;
   uint32_t number_of_matchers;
   bool match_exactly;
   // planner: skip the remaining children of a slice once every row's outcome is decided
   bool short_circuit = false;
};

class Union : public Operator {
  public:
   Union(Table* t, OperatorVector&& c) : Operator(t), children(std::move(c)) {}
   Type type() const override { return UNION; }
   std::string toString() const override;
   void emit(Emitter& e) const override;
   OperatorVector children;
   // planner: skip the remaining children of a slice once the partial result is all ones
   bool short_circuit = false;
};

// ---- expressions ----------------------------------------------------------------------------------
struct Expression;
using ExpressionPtr = std::unique_ptr<Expression>;

struct Expression {
   enum Kind {
      TRUE_, FALSE_, AND, OR, NOT, NOF, SYMBOL_EQUALS, HAS_MUTATION, SYMBOL_IN_SET, STRING_EQUALS, LINEAGE,
      DATE_BETWEEN, INT_BETWEEN, INT_EQUALS, BITMAP, COVERED,
      // operator-level forms (the reference's operator tests build operators directly with their
      // constructor rules)
      OP_THRESHOLD, OP_INTERSECTION, OP_UNION, OP_COMPLEMENT, OP_RANGES, OP_SELECT
   };
   // OP_SELECT: a CompareToValueSelection on an int/date column
   struct ValuePredicate {
      std::string column;
      int comparator;  // Comparator
      int64_t value;
   };
   Kind kind;
   std::vector<ExpressionPtr> children;
   std::string column;
   uint32_t position = 0;
   char symbol_char = 0;
   std::string symbol_chars;
   std::vector<uint8_t> symbols;
   int number = 0;
   bool exact = false;
   std::optional<std::string> value;
   bool sublineage = false;
   std::optional<int64_t> from, to;
   int64_t int_value = 0;
   uint32_t bitmap_id = 0;
   std::vector<uint32_t> row_ranges;  // OP_RANGES: global [start, end) pairs
   std::vector<ValuePredicate> predicates;

   explicit Expression(Kind k) : kind(k) {}
   ExpressionPtr shallowCopy() const;
   ExpressionPtr clone() const;
};

ExpressionPtr parseFilter(const std::string& text);
// ScalarExpression::rewrite(table, AmbiguityMode::NONE)
ExpressionPtr rewrite(const Table& table, const Expression& e);
// ScalarExpression::compile(table)
OperatorPtr compile(Table& table, const Expression& e);
// computeFilter (operators/compute_filter.cpp:14-21)
OperatorPtr compileFilter(Table& table, const std::string& text);

}  // namespace rdb::query
