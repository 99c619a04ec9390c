// Launch wrappers of the estimator's kernels (mlp.hip).  Row-major float matrices; `b` operands of
// the *Bf16 variant are BF16 parameter matrices.
#pragma once

#include <hip/hip_runtime.h>

namespace rdb {

// C[m, n] (+)= sum_k A(m, k) B(k, n) (+ bias[n]) (relu); A(m, k) = a[m * sam + k * sak],
// B(k, n) = b[k * sbk + n * sbn].
hipError_t mlpGemm(int M, int N, int K, const float* a, long sam, long sak, const float* b, long sbk, long sbn, float* c,
                   long ldc, const float* bias, bool relu, bool accumulate, hipStream_t s);
hipError_t mlpGemmBf16(int M, int N, int K, const float* a, long sam, long sak, const void* b, long sbk, long sbn,
                       float* c, long ldc, const float* bias, bool relu, bool accumulate, hipStream_t s);
// dz = 2 (z - y) / n; *loss += sum (z - y)^2
hipError_t mlpOutputGrad(int n, const float* z, const float* y, float* dz, float* loss, hipStream_t s);
hipError_t mlpReluGrad(long count, const float* z, float* dz, hipStream_t s);
hipError_t mlpColumnSum(int rows, int cols, const float* dz, float* db, hipStream_t s);
hipError_t mlpAdam(long n, float* master, void* w, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                   float eps, float c1, float c2, hipStream_t s);
hipError_t mlpToBf16(long n, const float* src, void* dst, hipStream_t s);
hipError_t mlpFromBf16(long n, const void* src, float* dst, hipStream_t s);

}  // namespace rdb
