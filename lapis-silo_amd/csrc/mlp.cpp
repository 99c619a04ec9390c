// Host side of the planner's estimator (see mlp.h).
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "mlp_kernels.h"

namespace rdb {

Mlp::Mlp(Context* c, uint32_t n_in, uint32_t n_hidden, uint32_t n_hidden_layers, uint64_t seed) : ctx(c) {
   if (n_in == 0 || n_hidden == 0 || n_hidden_layers == 0 || n_in > 4096 || n_hidden > 4096 || n_hidden_layers > 16) {
      fail(RDB_EINVAL, "estimator shape out of range");
   }
   dims.push_back(n_in);
   for (uint32_t l = 0; l < n_hidden_layers; ++l) dims.push_back(n_hidden);
   dims.push_back(1);
   for (size_t l = 0; l + 1 < dims.size(); ++l) {
      w_off.push_back(n_params);
      n_params += static_cast<size_t>(dims[l]) * dims[l + 1];
      b_off.push_back(n_params);
      n_params += dims[l + 1];
   }
   // He-uniform weights, zero biases
   std::vector<float> p(n_params, 0.0f);
   std::mt19937_64 rng(seed);
   for (size_t l = 0; l + 1 < dims.size(); ++l) {
      const float a = std::sqrt(6.0f / static_cast<float>(dims[l]));
      std::uniform_real_distribution<float> u(-a, a);
      for (size_t k = 0; k < static_cast<size_t>(dims[l]) * dims[l + 1]; ++k) p[w_off[l] + k] = u(rng);
   }
   master = DeviceBuffer(n_params * 4);
   bf16 = DeviceBuffer(n_params * 2);
   eff = DeviceBuffer(n_params * 4);
   grad = DeviceBuffer(n_params * 4);
   m1 = DeviceBuffer(n_params * 4);
   m2 = DeviceBuffer(n_params * 4);
   loss = DeviceBuffer(16);
   setParameters(p.data());
}

void Mlp::refreshCopies() {
   hipCheck(mlpToBf16(static_cast<long>(n_params), master.as<float>(), bf16.as<void>(), ctx->stream), "estimator bf16");
   hipCheck(mlpFromBf16(static_cast<long>(n_params), bf16.as<void>(), eff.as<float>(), ctx->stream), "estimator f32");
}

void Mlp::setParameters(const float* p) {
   hipCheck(hipMemcpyAsync(master.as<void>(), p, n_params * 4, hipMemcpyHostToDevice, ctx->stream), "upload parameters");
   hipCheck(hipMemsetAsync(m1.as<void>(), 0, n_params * 4, ctx->stream), "clear moments");
   hipCheck(hipMemsetAsync(m2.as<void>(), 0, n_params * 4, ctx->stream), "clear moments");
   adam_step = 0;
   refreshCopies();
   hipCheck(hipStreamSynchronize(ctx->stream), "synchronize");
}

std::vector<float> Mlp::parameters() const {
   std::vector<float> p(n_params);
   hipCheck(hipMemcpyAsync(p.data(), eff.as<void>(), n_params * 4, hipMemcpyDeviceToHost, ctx->stream), "copy parameters");
   hipCheck(hipStreamSynchronize(ctx->stream), "synchronize");
   return p;
}

void Mlp::reserve(size_t n) {
   if (n <= rows_capacity) return;
   const size_t cap = std::max<size_t>(n, 1024);
   size_t width = 0;
   for (size_t l = 1; l < dims.size(); ++l) width += dims[l];
   const uint32_t widest = *std::max_element(dims.begin(), dims.end());
   x = DeviceBuffer(cap * dims[0] * 4);
   y = DeviceBuffer(cap * 4);
   acts = DeviceBuffer(cap * width * 4);
   dz0 = DeviceBuffer(cap * widest * 4);
   dz1 = DeviceBuffer(cap * widest * 4);
   rows_capacity = cap;
}

float* Mlp::act(size_t layer) const {
   size_t off = 0;
   for (size_t l = 0; l < layer; ++l) off += dims[l + 1];
   return acts.as<float>() + off * rows_capacity;
}

void Mlp::forwardDevice(const float* x_dev, size_t n) {
   const size_t L = dims.size() - 1;
   const float* in = x_dev;
   for (size_t l = 0; l < L; ++l) {
      const long din = dims[l], dout = dims[l + 1];
      hipCheck(mlpGemmBf16(static_cast<int>(n), static_cast<int>(dout), static_cast<int>(din), in, din, 1,
                           bf16.as<uint16_t>() + w_off[l], 1, din, act(l), dout, eff.as<float>() + b_off[l], l + 1 < L,
                           false, ctx->stream),
               "estimator forward");
      in = act(l);
   }
}

float Mlp::backwardDevice(const float* x_dev, const float* y_dev, size_t n) {
   const size_t L = dims.size() - 1;
   hipCheck(hipMemsetAsync(loss.as<void>(), 0, 4, ctx->stream), "clear loss");
   float* dz = dz0.as<float>();
   float* dz_prev = dz1.as<float>();
   hipCheck(mlpOutputGrad(static_cast<int>(n), act(L - 1), y_dev, dz, loss.as<float>(), ctx->stream), "output gradient");
   for (size_t l = L; l-- > 0;) {
      const long din = dims[l], dout = dims[l + 1];
      const float* in = l == 0 ? x_dev : act(l - 1);
      // dW = dZ^T in
      hipCheck(mlpGemm(static_cast<int>(dout), static_cast<int>(din), static_cast<int>(n), dz, 1, dout, in, din, 1,
                       grad.as<float>() + w_off[l], din, nullptr, false, false, ctx->stream),
               "weight gradient");
      hipCheck(mlpColumnSum(static_cast<int>(n), static_cast<int>(dout), dz, grad.as<float>() + b_off[l], ctx->stream),
               "bias gradient");
      if (l == 0) break;
      // dA = dZ W, then the relu derivative of the layer below
      hipCheck(mlpGemmBf16(static_cast<int>(n), static_cast<int>(din), static_cast<int>(dout), dz, dout, 1,
                           bf16.as<uint16_t>() + w_off[l], din, 1, dz_prev, din, nullptr, false, false, ctx->stream),
               "activation gradient");
      hipCheck(mlpReluGrad(static_cast<long>(n) * din, act(l - 1), dz_prev, ctx->stream), "relu gradient");
      std::swap(dz, dz_prev);
   }
   float sq = 0.0f;
   hipCheck(hipMemcpyAsync(&sq, loss.as<void>(), 4, hipMemcpyDeviceToHost, ctx->stream), "copy loss");
   hipCheck(hipStreamSynchronize(ctx->stream), "synchronize");
   return sq;
}

void Mlp::predict(const float* x_host, size_t n, float* y_host) {
   if (n == 0) return;
   reserve(n);
   hipCheck(hipMemcpyAsync(x.as<void>(), x_host, n * dims[0] * 4, hipMemcpyHostToDevice, ctx->stream), "upload features");
   forwardDevice(x.as<float>(), n);
   hipCheck(hipMemcpyAsync(y_host, act(dims.size() - 2), n * 4, hipMemcpyDeviceToHost, ctx->stream), "copy estimates");
   hipCheck(hipStreamSynchronize(ctx->stream), "synchronize");
}

float Mlp::gradients(const float* x_host, const float* y_host, size_t n, float* grads_host) {
   if (n == 0) fail(RDB_EINVAL, "no rows");
   reserve(n);
   hipCheck(hipMemcpyAsync(x.as<void>(), x_host, n * dims[0] * 4, hipMemcpyHostToDevice, ctx->stream), "upload features");
   hipCheck(hipMemcpyAsync(y.as<void>(), y_host, n * 4, hipMemcpyHostToDevice, ctx->stream), "upload targets");
   forwardDevice(x.as<float>(), n);
   const float sq = backwardDevice(x.as<float>(), y.as<float>(), n);
   hipCheck(hipMemcpyAsync(grads_host, grad.as<void>(), n_params * 4, hipMemcpyDeviceToHost, ctx->stream), "copy gradients");
   hipCheck(hipStreamSynchronize(ctx->stream), "synchronize");
   return sq / static_cast<float>(n);
}

float Mlp::train(const float* x_host, const float* y_host, size_t n, uint32_t epochs, uint32_t batch, float lr,
                 uint64_t seed) {
   if (n == 0 || batch == 0) fail(RDB_EINVAL, "no rows");
   const size_t d = dims[0];
   // the whole (shuffled) set is resident; minibatches are slices of it
   DeviceBuffer xs(n * d * 4), ys(n * 4);
   std::vector<size_t> perm(n);
   std::iota(perm.begin(), perm.end(), size_t{0});
   std::mt19937_64 rng(seed);
   std::vector<float> xb(n * d), yb(n);
   reserve(std::min<size_t>(batch, n));
   constexpr float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
   float last = 0.0f;
   for (uint32_t e = 0; e < epochs; ++e) {
      std::shuffle(perm.begin(), perm.end(), rng);
      for (size_t i = 0; i < n; ++i) {
         std::memcpy(xb.data() + i * d, x_host + perm[i] * d, d * 4);
         yb[i] = y_host[perm[i]];
      }
      hipCheck(hipMemcpyAsync(xs.as<void>(), xb.data(), n * d * 4, hipMemcpyHostToDevice, ctx->stream), "upload features");
      hipCheck(hipMemcpyAsync(ys.as<void>(), yb.data(), n * 4, hipMemcpyHostToDevice, ctx->stream), "upload targets");
      double sum = 0.0;
      for (size_t s = 0; s < n; s += batch) {
         const size_t m = std::min<size_t>(batch, n - s);
         forwardDevice(xs.as<float>() + s * d, m);
         sum += backwardDevice(xs.as<float>() + s * d, ys.as<float>() + s, m);
         ++adam_step;
         const float c1 = 1.0f - std::pow(b1, static_cast<float>(adam_step));
         const float c2 = 1.0f - std::pow(b2, static_cast<float>(adam_step));
         hipCheck(mlpAdam(static_cast<long>(n_params), master.as<float>(), bf16.as<void>(), grad.as<float>(), m1.as<float>(),
                          m2.as<float>(), lr, b1, b2, eps, c1, c2, ctx->stream),
                  "adam");
         hipCheck(mlpFromBf16(static_cast<long>(n_params), bf16.as<void>(), eff.as<float>(), ctx->stream), "estimator f32");
      }
      last = static_cast<float>(sum / static_cast<double>(n));
   }
   hipCheck(hipStreamSynchronize(ctx->stream), "synchronize");
   return last;
}

}  // namespace rdb
