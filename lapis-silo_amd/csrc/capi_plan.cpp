// C ABI of the learned planner: the estimator network (rdb_mlp_*) and the planner's training rows,
// estimates and switch (include/rdb.h).
#include <cstring>

#include "batch.h"
#include "capi_internal.h"
#include "mlp.h"
#include "planner.h"

struct rdb_mlp {
   std::unique_ptr<rdb::Mlp> m;
};

using rdb::capi::guarded;
using rdb::capi::requireFinalized;
using rdb::capi::tableOf;

namespace {

rdb::Mlp& mlpOf(rdb_mlp* m) {
   if (!m || !m->m) rdb::fail(RDB_EINVAL, "null estimator");
   return *m->m;
}

// Compiles the filters and collects their operator trees' nodes.
struct Trees {
   std::vector<rdb::query::OperatorPtr> ops;
   std::vector<std::vector<rdb::plan::Node>> nodes;
   size_t total = 0;
};

Trees collectTrees(rdb::Table& t, const char* const* filters, uint32_t n) {
   Trees tr;
   for (uint32_t i = 0; i < n; ++i) {
      if (!filters[i]) rdb::fail(RDB_EINVAL, "null filter");
      tr.ops.push_back(rdb::query::compileFilter(t, filters[i]));
      tr.nodes.push_back(rdb::plan::collect(t, *tr.ops.back()));
      tr.total += tr.nodes.back().size();
   }
   return tr;
}

// Exact cardinality of every node: each node's subtree as its own program, one batch.
std::vector<std::vector<uint64_t>> exactCounts(rdb::Table& t, Trees& tr) {
   std::vector<rdb::query::Emitter> programs;
   programs.reserve(tr.total);
   for (auto& nodes : tr.nodes) {
      for (auto& n : nodes) {
         rdb::query::Emitter e;
         n.op->emit(e);
         programs.push_back(std::move(e));
      }
   }
   const auto flat = rdb::countPrograms(t, programs);
   std::vector<std::vector<uint64_t>> out(tr.nodes.size());
   size_t o = 0;
   for (size_t q = 0; q < tr.nodes.size(); ++q) {
      out[q].assign(flat.begin() + static_cast<ptrdiff_t>(o), flat.begin() + static_cast<ptrdiff_t>(o + tr.nodes[q].size()));
      o += tr.nodes[q].size();
   }
   return out;
}

}  // namespace

extern "C" {

int rdb_mlp_create(rdb_ctx* ctx, uint32_t n_in, uint32_t n_hidden, uint32_t n_hidden_layers, uint64_t seed,
                   rdb_mlp** out) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx || !out) rdb::fail(RDB_EINVAL, "null argument");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      auto m = std::make_unique<rdb_mlp>();
      m->m = std::make_unique<rdb::Mlp>(&ctx->c, n_in, n_hidden, n_hidden_layers, seed);
      *out = m.release();
   });
}

void rdb_mlp_free(rdb_mlp* mlp) {
   if (!mlp) return;
   if (mlp->m) {
      rdb::Context* ctx = mlp->m->ctx;
      std::lock_guard<std::recursive_mutex> lock(ctx->mu);
      if (ctx->planner == mlp->m.get()) {
         ctx->planner = nullptr;
         if (ctx->plan_mode == rdb::plan::LEARNED) ctx->plan_mode = rdb::plan::REFERENCE;
      }
      mlp->m.reset();
   }
   delete mlp;
}

int rdb_mlp_parameter_count(rdb_mlp* mlp, uint64_t* n) {
   return guarded(mlp && mlp->m ? mlp->m->ctx : nullptr, [&] {
      if (!n) rdb::fail(RDB_EINVAL, "null argument");
      *n = mlpOf(mlp).n_params;
   });
}

int rdb_mlp_get_parameters(rdb_mlp* mlp, float* out, uint64_t n) {
   return guarded(mlp && mlp->m ? mlp->m->ctx : nullptr, [&] {
      rdb::Mlp& m = mlpOf(mlp);
      std::lock_guard<std::recursive_mutex> lock(m.ctx->mu);
      if (!out || n != m.n_params) rdb::fail(RDB_EINVAL, "parameter count mismatch");
      const auto p = m.parameters();
      std::memcpy(out, p.data(), p.size() * 4);
   });
}

int rdb_mlp_set_parameters(rdb_mlp* mlp, const float* p, uint64_t n) {
   return guarded(mlp && mlp->m ? mlp->m->ctx : nullptr, [&] {
      rdb::Mlp& m = mlpOf(mlp);
      std::lock_guard<std::recursive_mutex> lock(m.ctx->mu);
      if (!p || n != m.n_params) rdb::fail(RDB_EINVAL, "parameter count mismatch");
      m.setParameters(p);
   });
}

int rdb_mlp_forward(rdb_mlp* mlp, const float* x, uint32_t n, float* y) {
   return guarded(mlp && mlp->m ? mlp->m->ctx : nullptr, [&] {
      rdb::Mlp& m = mlpOf(mlp);
      std::lock_guard<std::recursive_mutex> lock(m.ctx->mu);
      if ((!x || !y) && n) rdb::fail(RDB_EINVAL, "null argument");
      m.predict(x, n, y);
   });
}

int rdb_mlp_gradients(rdb_mlp* mlp, const float* x, const float* y, uint32_t n, float* loss, float* grads) {
   return guarded(mlp && mlp->m ? mlp->m->ctx : nullptr, [&] {
      rdb::Mlp& m = mlpOf(mlp);
      std::lock_guard<std::recursive_mutex> lock(m.ctx->mu);
      if (!x || !y || !loss || !grads) rdb::fail(RDB_EINVAL, "null argument");
      *loss = m.gradients(x, y, n, grads);
   });
}

int rdb_mlp_train(rdb_mlp* mlp, const float* x, const float* y, uint64_t n, uint32_t epochs, uint32_t batch, float lr,
                  uint64_t seed, float* final_loss) {
   return guarded(mlp && mlp->m ? mlp->m->ctx : nullptr, [&] {
      rdb::Mlp& m = mlpOf(mlp);
      std::lock_guard<std::recursive_mutex> lock(m.ctx->mu);
      if (!x || !y) rdb::fail(RDB_EINVAL, "null argument");
      const float l = m.train(x, y, n, epochs, batch, lr, seed);
      if (final_loss) *final_loss = l;
   });
}

int rdb_plan_samples(rdb_table* table, const char* const* filters, uint32_t n_filters, float* features, float* targets,
                     uint32_t* tree, uint64_t capacity, uint64_t* n) {
   rdb::Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      rdb::Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!n || (n_filters && !filters)) rdb::fail(RDB_EINVAL, "null argument");
      Trees tr = collectTrees(t, filters, n_filters);
      *n = tr.total;
      if (tr.total > capacity) rdb::fail(RDB_ENOMEM, "output capacity too small");
      const auto exact = exactCounts(t, tr);
      const double rows = static_cast<double>(t.num_rows);
      size_t o = 0;
      for (size_t q = 0; q < tr.nodes.size(); ++q) {
         std::vector<float> v(tr.nodes[q].size());
         for (size_t k = 0; k < v.size(); ++k) v[k] = rdb::plan::normRows(static_cast<double>(exact[q][k]), rows);
         for (size_t k = 0; k < v.size(); ++k) {
            // children's exact values stand in for estimates (teacher forcing)
            rdb::plan::compose(t, tr.nodes[q], k, v.data());
            if (features) std::memcpy(features + o * rdb::plan::kFeatures, tr.nodes[q][k].features, rdb::plan::kFeatures * 4);
            if (targets) targets[o] = v[k];
            if (tree) tree[o] = static_cast<uint32_t>(q);
            ++o;
         }
      }
   });
}

int rdb_plan_estimates(rdb_table* table, rdb_mlp* mlp, const char* const* filters, uint32_t n_filters, float* estimate,
                       float* independence, uint64_t* exact, uint32_t* tree, uint64_t capacity, uint64_t* n) {
   rdb::Context* ctx = table ? table->t.ctx : nullptr;
   return guarded(ctx, [&] {
      rdb::Table& t = tableOf(table);
      std::lock_guard<std::recursive_mutex> lock(t.ctx->mu);
      requireFinalized(t);
      if (!n || (n_filters && !filters)) rdb::fail(RDB_EINVAL, "null argument");
      Trees tr = collectTrees(t, filters, n_filters);
      *n = tr.total;
      if (tr.total > capacity) rdb::fail(RDB_ENOMEM, "output capacity too small");
      const double rows = static_cast<double>(t.num_rows);
      const auto ind = rdb::plan::estimate(t, tr.nodes, rdb::plan::INDEPENDENCE, nullptr);
      std::vector<std::vector<float>> learned;
      if (mlp) learned = rdb::plan::estimate(t, tr.nodes, rdb::plan::LEARNED, &mlpOf(mlp));
      std::vector<std::vector<uint64_t>> ex;
      if (exact) ex = exactCounts(t, tr);
      size_t o = 0;
      for (size_t q = 0; q < tr.nodes.size(); ++q) {
         for (size_t k = 0; k < tr.nodes[q].size(); ++k) {
            if (estimate) estimate[o] = mlp ? static_cast<float>(rdb::plan::denormRows(learned[q][k], rows)) : 0.0f;
            if (independence) independence[o] = static_cast<float>(rdb::plan::denormRows(ind[q][k], rows));
            if (exact) exact[o] = ex[q][k];
            if (tree) tree[o] = static_cast<uint32_t>(q);
            ++o;
         }
      }
   });
}

int rdb_set_planner(rdb_ctx* ctx, int mode, rdb_mlp* mlp) {
   return guarded(ctx ? &ctx->c : nullptr, [&] {
      if (!ctx) rdb::fail(RDB_EINVAL, "null context");
      std::lock_guard<std::recursive_mutex> lock(ctx->c.mu);
      if (mode < rdb::plan::REFERENCE || mode > rdb::plan::LEARNED) rdb::fail(RDB_EINVAL, "unknown planner mode");
      if (mode == rdb::plan::LEARNED && !mlp) rdb::fail(RDB_EINVAL, "the learned planner needs an estimator");
      if (mlp && mlpOf(mlp).dims[0] != rdb::plan::kFeatures) rdb::fail(RDB_EINVAL, "estimator input width differs from the planner's features");
      ctx->c.plan_mode = mode;
      ctx->c.planner = mode == rdb::plan::LEARNED ? mlp->m.get() : nullptr;
   });
}

}  // extern "C"
