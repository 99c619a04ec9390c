// Planning of compiled filter operator trees: per-node features for a learned cardinality
// estimator, and the plan it drives.
//
// A plan never changes a result.  It only
//   * orders the children of Intersection (ascending estimated rows) and Union (descending), both
//     commutative in the filter kernel's word arithmetic, and
//   * turns on their short-circuit, with which the filter kernel skips the remaining children of a
//     4096-row slice once the partial conjunction is empty (or the partial disjunction full), or
//     once every row's n-of-m count is decided whatever the remaining children add.
// Without an estimator the compiled order stands and nothing is skipped: the reference's plan,
// which evaluates every child (intersection.cpp:59-106, union.cpp:34-42).
#pragma once

#include <cstdint>
#include <vector>

#include "mlp.h"
#include "query.h"

namespace rdb::plan {

constexpr uint32_t kFeatures = 32;

struct Node {
   query::Operator* op = nullptr;
   int parent = -1;
   std::vector<int> children;       // node indices (post-order: children come first)
   std::vector<uint8_t> negated;    // per child: negated child of an Intersection or Threshold
   float features[kFeatures] = {};  // static part; the child summaries are filled by compose()
   uint32_t depth = 0;
   bool leaf_hint = false;
};

// Normalised log-cardinality: log(1 + rows) / log(1 + table rows), in [0, 1].
float normRows(double rows, double table_rows);
double denormRows(float v, double table_rows);

// Post-order node list of an operator tree with the static features of every node.
std::vector<Node> collect(const Table& t, query::Operator& root);

// Fills node k's child-summary features from per-node values (normalised log-cardinalities:
// exact ones for training, the estimator's own for planning).
void compose(const Table& t, std::vector<Node>& nodes, size_t k, const float* values);

// An estimate without the network: leaf statistics combined under independence (product for
// And, complement rule for Or, binomial tail for n-of-m).  Also one of the network's features.
float independenceEstimate(const Table& t, const std::vector<Node>& nodes, size_t k, const float* values);

// Orders the children of every Intersection, Union and Threshold (non-negated ones) by the
// estimates and turns on their short-circuit.
void apply(std::vector<Node>& nodes, const float* estimates);

enum Mode { REFERENCE = 0, INDEPENDENCE = 1, LEARNED = 2 };

// Estimates (normalised log-cardinalities) of every node of every tree, bottom up: the learned
// estimator composes each node's features from its children's estimates and runs level by level
// (all trees' nodes of one height per launch); INDEPENDENCE uses independenceEstimate alone.
std::vector<std::vector<float>> estimate(const Table& t, std::vector<std::vector<Node>>& trees, Mode mode, Mlp* mlp,
                                         unsigned threads = 1);

// Plans compiled filters in place (no effect in REFERENCE mode).
// Turns on short-circuit everywhere in the tree, keeping the compiled order of every operator's
// children.
void shortCircuitOnly(query::Operator& root);

// `threads`: host threads for the per-tree work (0: as many as useful, 1: this thread only).
void planAll(const Table& t, std::vector<query::Operator*>& roots, Mode mode, Mlp* mlp, unsigned threads = 1);

}  // namespace rdb::plan
