// Planning of compiled filter operator trees (see planner.h).
#include "planner.h"

#include <algorithm>
#include <bit>
#include <cmath>
#include <unordered_map>

namespace rdb::plan {

namespace {

enum Feature : uint32_t {
   F_EMPTY = 0,
   F_FULL,
   F_SCAN_STORE,
   F_SCAN_SYMBOLS,
   F_INTERSECTION,
   F_COMPLEMENT,
   F_RANGE,
   F_SELECTION_CHILD,
   F_SELECTION_LEAF,
   F_THRESHOLD,
   F_UNION,
   F_LEAF_ROWS,     // normalised rows of a leaf, from host statistics
   F_HAS_LEAF_ROWS,
   F_N_POS,         // children / 20
   F_N_NEG,         // negated children / 20
   F_POS_MIN,       // summaries of the children's values
   F_POS_MAX,
   F_POS_MEAN,
   F_INDEPENDENCE,  // independence estimate of the node
   F_NEG_MIN,
   F_NEG_MAX,
   F_NEG_MEAN,
   F_TH_RATIO,      // n / m of a Threshold
   F_TH_EXACT,
   F_DEPTH,         // depth / 8
   F_SYMBOLS,       // symbols of a sequence IndexScan / 16
   F_PREDICATES,    // predicates of a Selection / 4
   F_TABLE_SCALE,   // log2(table rows) / 32
   F_PRED_PRODUCT,  // normalised rows of the Selection's predicates under independence
};

double selectivity(float v, double n) { return n > 0 ? denormRows(v, n) / n : 0.0; }

void visit(const Table& t, query::Operator& op, uint32_t depth, int parent, std::vector<Node>& out, int& index) {
   using namespace query;
   std::vector<std::pair<Operator*, bool>> kids;
   switch (op.type()) {
      case INTERSECTION: {
         auto& o = static_cast<Intersection&>(op);
         for (auto& c : o.children) kids.emplace_back(c.get(), false);
         for (auto& c : o.negated_children) kids.emplace_back(c.get(), true);
         break;
      }
      case UNION:
         for (auto& c : static_cast<Union&>(op).children) kids.emplace_back(c.get(), false);
         break;
      case COMPLEMENT:
         kids.emplace_back(static_cast<Complement&>(op).child.get(), false);
         break;
      case SELECTION: {
         auto& o = static_cast<Selection&>(op);
         if (o.child_operator.has_value()) kids.emplace_back(o.child_operator->get(), false);
         break;
      }
      case THRESHOLD: {
         auto& o = static_cast<Threshold&>(op);
         for (auto& c : o.non_negated_children) kids.emplace_back(c.get(), false);
         for (auto& c : o.negated_children) kids.emplace_back(c.get(), true);
         break;
      }
      default:
         break;
   }
   std::vector<int> child_idx;
   for (auto& [c, neg] : kids) {
      visit(t, *c, depth + 1, -2, out, index);
      child_idx.push_back(index - 1);
   }
   Node n;
   n.op = &op;
   n.parent = parent;
   n.depth = depth;
   n.children = child_idx;
   for (auto& [c, neg] : kids) n.negated.push_back(neg ? 1 : 0);
   const double rows = static_cast<double>(t.num_rows);
   float* f = n.features;
   f[F_DEPTH] = static_cast<float>(depth) / 8.0f;
   f[F_TABLE_SCALE] = static_cast<float>(std::log2(rows + 1.0) / 32.0);
   auto leaf = [&](double r) {
      n.leaf_hint = true;
      f[F_LEAF_ROWS] = normRows(r, rows);
      f[F_HAS_LEAF_ROWS] = 1.0f;
   };
   switch (op.type()) {
      case EMPTY:
         f[F_EMPTY] = 1;
         leaf(0);
         break;
      case FULL:
         f[F_FULL] = 1;
         leaf(rows);
         break;
      case INDEX_SCAN: {
         const auto& s = static_cast<IndexScan&>(op).source;
         if (s.kind == BitmapSource::STORE) {
            f[F_SCAN_STORE] = 1;
            leaf(static_cast<double>(t.store_card[s.store_id]));
         } else {
            f[F_SCAN_SYMBOLS] = 1;
            const SeqColumn& col = *t.seq[s.column];
            double r = 0;
            for (uint32_t slot = col.pos_slot_ptr[s.position]; slot < col.pos_slot_ptr[s.position + 1]; ++slot) {
               const uint32_t sym = col.slot_info[slot] >> 16;
               if (sym < col.abc->count && ((s.symbol_mask >> sym) & 1u)) r += col.slot_rows[slot];
            }
            leaf(r);
            f[F_SYMBOLS] = static_cast<float>(std::popcount(s.symbol_mask)) / 16.0f;
         }
         break;
      }
      case RANGE_SELECTION: {
         f[F_RANGE] = 1;
         double r = 0;
         for (const auto& g : static_cast<RangeSelection&>(op).ranges) r += static_cast<double>(g.end) - g.start;
         leaf(r);
         break;
      }
      case SELECTION: {
         auto& o = static_cast<Selection&>(op);
         double sel = 1.0;
         for (const auto& p : o.predicates) sel *= rows > 0 ? static_cast<double>(p->tableCount(t)) / rows : 0.0;
         f[F_PREDICATES] = static_cast<float>(o.predicates.size()) / 4.0f;
         f[F_PRED_PRODUCT] = normRows(sel * rows, rows);
         if (o.child_operator.has_value()) {
            f[F_SELECTION_CHILD] = 1;
         } else {
            f[F_SELECTION_LEAF] = 1;
            leaf(sel * rows);
         }
         break;
      }
      case INTERSECTION:
         f[F_INTERSECTION] = 1;
         break;
      case UNION:
         f[F_UNION] = 1;
         break;
      case COMPLEMENT:
         f[F_COMPLEMENT] = 1;
         break;
      case THRESHOLD: {
         auto& o = static_cast<Threshold&>(op);
         f[F_THRESHOLD] = 1;
         const double m = static_cast<double>(o.non_negated_children.size() + o.negated_children.size());
         f[F_TH_RATIO] = static_cast<float>(o.number_of_matchers / m);
         f[F_TH_EXACT] = o.match_exactly ? 1.0f : 0.0f;
         break;
      }
   }
   size_t np = 0, nn = 0;
   for (uint8_t g : n.negated) (g ? nn : np) += 1;
   f[F_N_POS] = static_cast<float>(np) / 20.0f;
   f[F_N_NEG] = static_cast<float>(nn) / 20.0f;
   out.push_back(std::move(n));
   for (int c : out.back().children) out[c].parent = index;
   ++index;
}

}  // namespace

float normRows(double rows, double table_rows) {
   if (table_rows <= 0) return 0.0f;
   const double r = std::clamp(rows, 0.0, table_rows);
   return static_cast<float>(std::log1p(r) / std::log1p(table_rows));
}

double denormRows(float v, double table_rows) {
   const double x = std::clamp(static_cast<double>(v), 0.0, 1.0);
   return std::expm1(x * std::log1p(table_rows));
}

std::vector<Node> collect(const Table& t, query::Operator& root) {
   std::vector<Node> out;
   int index = 0;
   visit(t, root, 0, -1, out, index);
   return out;
}

float independenceEstimate(const Table& t, const std::vector<Node>& nodes, size_t k, const float* values) {
   using namespace query;
   const Node& n = nodes[k];
   const double rows = static_cast<double>(t.num_rows);
   if (n.leaf_hint) return n.features[F_LEAF_ROWS];
   std::vector<double> pos, neg;
   for (size_t i = 0; i < n.children.size(); ++i) {
      (n.negated[i] ? neg : pos).push_back(selectivity(values[n.children[i]], rows));
   }
   double s = 0;
   switch (n.op->type()) {
      case INTERSECTION:
         s = 1;
         for (double x : pos) s *= x;
         for (double x : neg) s *= 1 - x;
         break;
      case UNION: {
         double none = 1;
         for (double x : pos) none *= 1 - x;
         s = 1 - none;
         break;
      }
      case COMPLEMENT:
         s = 1 - pos.at(0);
         break;
      case SELECTION:
         s = (pos.empty() ? 1.0 : pos[0]) * selectivity(n.features[F_PRED_PRODUCT], rows);
         break;
      case THRESHOLD: {
         // Poisson-binomial distribution of the number of matching children
         const auto& o = static_cast<const Threshold&>(*n.op);
         std::vector<double> p(1, 1.0);
         auto add = [&](double x) {
            std::vector<double> q(p.size() + 1, 0.0);
            for (size_t j = 0; j < p.size(); ++j) {
               q[j] += p[j] * (1 - x);
               q[j + 1] += p[j] * x;
            }
            p.swap(q);
         };
         for (double x : pos) add(x);
         for (double x : neg) add(1 - x);
         s = 0;
         for (size_t j = o.number_of_matchers; j < p.size(); ++j) {
            if (!o.match_exactly || j == o.number_of_matchers) s += p[j];
         }
         break;
      }
      default:
         s = 0.5;
         break;
   }
   return normRows(std::clamp(s, 0.0, 1.0) * rows, rows);
}

void compose(const Table& t, std::vector<Node>& nodes, size_t k, const float* values) {
   Node& n = nodes[k];
   float* f = n.features;
   float pmin = 1, pmax = 0, psum = 0, nmin = 1, nmax = 0, nsum = 0;
   size_t np = 0, nn = 0;
   for (size_t i = 0; i < n.children.size(); ++i) {
      const float v = values[n.children[i]];
      if (n.negated[i]) {
         nmin = std::min(nmin, v);
         nmax = std::max(nmax, v);
         nsum += v;
         ++nn;
      } else {
         pmin = std::min(pmin, v);
         pmax = std::max(pmax, v);
         psum += v;
         ++np;
      }
   }
   f[F_POS_MIN] = np ? pmin : 0;
   f[F_POS_MAX] = np ? pmax : 0;
   f[F_POS_MEAN] = np ? psum / static_cast<float>(np) : 0;
   f[F_NEG_MIN] = nn ? nmin : 0;
   f[F_NEG_MAX] = nn ? nmax : 0;
   f[F_NEG_MEAN] = nn ? nsum / static_cast<float>(nn) : 0;
   f[F_INDEPENDENCE] = independenceEstimate(t, nodes, k, values);
}

void apply(std::vector<Node>& nodes, const float* estimates) {
   using namespace query;
   std::unordered_map<const Operator*, float> est;
   for (size_t k = 0; k < nodes.size(); ++k) est[nodes[k].op] = estimates[k];
   auto by = [&](bool ascending) {
      return [&, ascending](const OperatorPtr& a, const OperatorPtr& b) {
         return ascending ? est[a.get()] < est[b.get()] : est[a.get()] > est[b.get()];
      };
   };
   for (auto& n : nodes) {
      if (n.op->type() == INTERSECTION) {
         auto& o = static_cast<Intersection&>(*n.op);
         // smallest conjunct first: the partial result empties early; negated children that
         // remove the most rows first
         std::stable_sort(o.children.begin(), o.children.end(), by(true));
         std::stable_sort(o.negated_children.begin(), o.negated_children.end(), by(false));
         o.short_circuit = true;
      } else if (n.op->type() == UNION) {
         auto& o = static_cast<Union&>(*n.op);
         std::stable_sort(o.children.begin(), o.children.end(), by(false));
         o.short_circuit = true;
      } else if (n.op->type() == THRESHOLD) {
         // only the non-negated children are reordered: their counting does not depend on the
         // order (the negated ones' handling of rows outside the layout does)
         auto& o = static_cast<Threshold&>(*n.op);
         const size_t m = o.non_negated_children.size() + o.negated_children.size();
         // few matchers needed: the largest children first reach n soonest; many: the smallest
         // first rule rows out soonest
         const bool largest_first = 2 * o.number_of_matchers <= m;
         std::stable_sort(o.non_negated_children.begin(), o.non_negated_children.end(), by(!largest_first));
         o.short_circuit = true;
      }
   }
}

// A leaf whose statistic is its exact row count: Empty, Full, an index scan (store cardinality or
// the stored symbols' rows), row ranges, a selection of one predicate (its table count); a
// selection of several predicates multiplies their selectivities and is estimated like an
// operator.
bool exactLeaf(const Node& n) {
   if (!n.leaf_hint) return false;
   if (n.op->type() == query::SELECTION) {
      return static_cast<const query::Selection&>(*n.op).predicates.size() <= 1;
   }
   return true;
}

std::vector<std::vector<float>> estimate(const Table& t, std::vector<std::vector<Node>>& trees, Mode mode, Mlp* mlp,
                                         unsigned threads) {
   std::vector<std::vector<float>> est(trees.size());
   for (size_t q = 0; q < trees.size(); ++q) est[q].assign(trees[q].size(), 0.0f);
   if (mode != LEARNED || mlp == nullptr) {
      parallelFor(trees.size(), [&](size_t q) {
         for (size_t k = 0; k < trees[q].size(); ++k) est[q][k] = independenceEstimate(t, trees[q], k, est[q].data());
      }, threads);
      return est;
   }
   // heights: leaves 0; a node's height is one more than its highest child
   std::vector<std::vector<uint32_t>> height(trees.size());
   uint32_t max_h = 0;
   for (size_t q = 0; q < trees.size(); ++q) {
      height[q].assign(trees[q].size(), 0);
      for (size_t k = 0; k < trees[q].size(); ++k) {
         uint32_t h = 0;
         for (int c : trees[q][k].children) h = std::max(h, height[q][c] + 1);
         height[q][k] = h;
         max_h = std::max(max_h, h);
      }
   }
   if (mlp->dims[0] != kFeatures) fail(RDB_EINVAL, "the estimator takes " + std::to_string(mlp->dims[0]) + " features, the planner " + std::to_string(kFeatures));
   std::vector<float> x, y;
   std::vector<std::pair<uint32_t, uint32_t>> who;
   for (uint32_t h = 0; h <= max_h; ++h) {
      x.clear();
      who.clear();
      // the nodes of height h compose their features from their children's estimates (per tree on
      // the host threads), then one estimator launch takes them all
      parallelFor(trees.size(), [&](size_t q) {
         for (size_t k = 0; k < trees[q].size(); ++k) {
            if (height[q][k] == h) compose(t, trees[q], k, est[q].data());
         }
      }, threads);
      for (size_t q = 0; q < trees.size(); ++q) {
         for (size_t k = 0; k < trees[q].size(); ++k) {
            if (height[q][k] != h) continue;
            const Node& n = trees[q][k];
            if (exactLeaf(n)) {
               // the host statistic is the leaf's row count: no estimate can improve on it
               est[q][k] = n.features[F_LEAF_ROWS];
               continue;
            }
            x.insert(x.end(), n.features, n.features + kFeatures);
            who.emplace_back(static_cast<uint32_t>(q), static_cast<uint32_t>(k));
         }
      }
      if (who.empty()) continue;
      y.assign(who.size(), 0.0f);
      mlp->predict(x.data(), who.size(), y.data());
      for (size_t i = 0; i < who.size(); ++i) est[who[i].first][who[i].second] = std::clamp(y[i], 0.0f, 1.0f);
   }
   return est;
}

void shortCircuitOnly(query::Operator& op) {
   using namespace query;
   switch (op.type()) {
      case INTERSECTION: {
         auto& o = static_cast<Intersection&>(op);
         o.short_circuit = true;
         for (auto& c : o.children) shortCircuitOnly(*c);
         for (auto& c : o.negated_children) shortCircuitOnly(*c);
         break;
      }
      case UNION: {
         auto& o = static_cast<Union&>(op);
         o.short_circuit = true;
         for (auto& c : o.children) shortCircuitOnly(*c);
         break;
      }
      case COMPLEMENT:
         shortCircuitOnly(*static_cast<Complement&>(op).child);
         break;
      case SELECTION: {
         auto& o = static_cast<Selection&>(op);
         if (o.child_operator.has_value()) shortCircuitOnly(**o.child_operator);
         break;
      }
      case THRESHOLD: {
         auto& o = static_cast<Threshold&>(op);
         o.short_circuit = true;
         for (auto& c : o.non_negated_children) shortCircuitOnly(*c);
         for (auto& c : o.negated_children) shortCircuitOnly(*c);
         break;
      }
      default:
         break;
   }
}

void planAll(const Table& t, std::vector<query::Operator*>& roots, Mode mode, Mlp* mlp, unsigned threads) {
   if (mode == REFERENCE) return;
   std::vector<std::vector<Node>> trees(roots.size());
   parallelFor(roots.size(), [&](size_t q) { trees[q] = collect(t, *roots[q]); }, threads);
   const auto est = estimate(t, trees, mode, mlp, threads);
   parallelFor(trees.size(), [&](size_t q) { apply(trees[q], est[q].data()); }, threads);
}

}  // namespace rdb::plan
