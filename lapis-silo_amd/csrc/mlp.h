// The planner's cardinality estimator: a multilayer perceptron (ReLU hidden layers, one linear
// output) with BF16 parameters and FP32 accumulation, evaluated and trained on the GPU by the
// kernels of mlp.hip.
#pragma once

#include <cstdint>
#include <vector>

#include "engine.h"

namespace rdb {

struct Mlp {
   Context* ctx = nullptr;
   std::vector<uint32_t> dims;              // inputs, hidden..., 1
   std::vector<size_t> w_off, b_off;        // per layer: W (out x in, row-major) and b offsets
   size_t n_params = 0;
   DeviceBuffer master, bf16, eff, grad, m1, m2;  // f32 master, its BF16 copy and that as f32, Adam moments
   uint64_t adam_step = 0;
   // per-batch scratch
   size_t rows_capacity = 0;
   DeviceBuffer x, y, acts, dz0, dz1, loss;

   Mlp(Context* c, uint32_t n_in, uint32_t n_hidden, uint32_t n_hidden_layers, uint64_t seed);
   void setParameters(const float* p);       // host f32 -> master, BF16 copy
   std::vector<float> parameters() const;    // the BF16 values the products use, as f32
   // y = f(x) for n rows of host features; host output
   void predict(const float* x_host, size_t n, float* y_host);
   // mean squared error and its parameter gradients (host) for n rows
   float gradients(const float* x_host, const float* y_host, size_t n, float* grads_host);
   // Adam over shuffled minibatches; returns the last epoch's mean squared error
   float train(const float* x_host, const float* y_host, size_t n, uint32_t epochs, uint32_t batch, float lr,
               uint64_t seed);

  private:
   void reserve(size_t n);
   float* act(size_t layer) const;  // output of layer `layer` (n x dims[layer + 1])
   void forwardDevice(const float* x_dev, size_t n);
   float backwardDevice(const float* x_dev, const float* y_dev, size_t n);  // fills grad; returns the sum of squares
   void refreshCopies();
};

}  // namespace rdb
