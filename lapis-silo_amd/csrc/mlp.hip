// HIP kernels of the planner's cardinality estimator: a multilayer perceptron with BF16 weights
// and FP32 accumulation, trained on the device.
//
//   forward   Z_l = A_{l-1} W_l^T + b_l,  A_l = relu(Z_l) (hidden layers), output Z_L (one unit)
//   loss      mean over the batch of (Z_L - y)^2
//   backward  dZ_L = 2 (Z_L - y) / N;  dW_l = dZ_l^T A_{l-1};  db_l = column sums of dZ_l;
//             dA_{l-1} = dZ_l W_l;  dZ_{l-1} = dA_{l-1} * [Z_{l-1} > 0]
//   update    Adam on FP32 master parameters, then the BF16 copy the products read
//
// The products are small (batch x 128 x 128), so they use one generic LDS-tiled kernel: 64 x 64
// output tiles, 16-deep slices of both operands staged in LDS, 4 x 4 outputs per thread.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_kernels.h"

namespace rdb {

namespace {

constexpr int kTile = 64;
constexpr int kDepth = 16;

__device__ inline float loadB(const float* p, size_t i) { return p[i]; }
__device__ inline float loadB(const __hip_bfloat16* p, size_t i) { return __bfloat162float(p[i]); }

// C[m, n] (+)= sum_k A(m, k) * B(k, n) with A(m, k) = a[m * sam + k * sak] and
// B(k, n) = b[k * sbk + n * sbn]; then + bias[n] and relu when asked.
template <typename TB>
__global__ __launch_bounds__(256) void gemmKernel(int M, int N, int K, const float* __restrict__ a, long sam, long sak,
                                                  const TB* __restrict__ b, long sbk, long sbn, float* __restrict__ c,
                                                  long ldc, const float* __restrict__ bias, int relu, int accumulate) {
   __shared__ float as[kDepth][kTile + 1];
   __shared__ float bs[kDepth][kTile + 1];
   const int tx = threadIdx.x & 15;  // 16 x 16 threads, 4 x 4 outputs each
   const int ty = threadIdx.x >> 4;
   const int m0 = blockIdx.y * kTile;
   const int n0 = blockIdx.x * kTile;
   float acc[4][4] = {};
   for (int k0 = 0; k0 < K; k0 += kDepth) {
      for (int i = threadIdx.x; i < kDepth * kTile; i += 256) {
         const int kk = i / kTile, mm = i % kTile;
         const int m = m0 + mm, k = k0 + kk;
         as[kk][mm] = (m < M && k < K) ? a[static_cast<long>(m) * sam + static_cast<long>(k) * sak] : 0.0f;
         const int n = n0 + mm;
         bs[kk][mm] = (n < N && k < K) ? loadB(b, static_cast<size_t>(static_cast<long>(k) * sbk + static_cast<long>(n) * sbn))
                                       : 0.0f;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < kDepth; ++kk) {
         float av[4], bv[4];
#pragma unroll
         for (int r = 0; r < 4; ++r) av[r] = as[kk][ty * 4 + r];
#pragma unroll
         for (int r = 0; r < 4; ++r) bv[r] = bs[kk][tx * 4 + r];
#pragma unroll
         for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
         }
      }
      __syncthreads();
   }
#pragma unroll
   for (int i = 0; i < 4; ++i) {
      const int m = m0 + ty * 4 + i;
      if (m >= M) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
         const int n = n0 + tx * 4 + j;
         if (n >= N) continue;
         float v = acc[i][j];
         if (bias) v += bias[n];
         if (relu) v = v > 0.0f ? v : 0.0f;
         float* out = c + static_cast<long>(m) * ldc + n;
         *out = accumulate ? *out + v : v;
      }
   }
}

// dZ = 2 (z - y) / n for the output unit, and the summed squared error.
__global__ void outputGradKernel(int n, const float* __restrict__ z, const float* __restrict__ y, float* __restrict__ dz,
                                 float* __restrict__ loss) {
   __shared__ float part[4];
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   float e2 = 0.0f;
   if (i < n) {
      const float e = z[i] - y[i];
      dz[i] = 2.0f * e / static_cast<float>(n);
      e2 = e * e;
   }
   for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
   if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e2;
   __syncthreads();
   if (threadIdx.x == 0) atomicAdd(loss, part[0] + part[1] + part[2] + part[3]);
}

// dZ *= [Z > 0] (relu derivative), in place.
__global__ void reluGradKernel(long count, const float* __restrict__ z, float* __restrict__ dz) {
   const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
   if (i < count && !(z[i] > 0.0f)) dz[i] = 0.0f;
}

// db[o] = sum over rows of dz[r, o]; one work-group per output unit.
__global__ void columnSumKernel(int rows, int cols, const float* __restrict__ dz, float* __restrict__ db) {
   __shared__ float part[4];
   const int o = blockIdx.x;
   float s = 0.0f;
   for (int r = threadIdx.x; r < rows; r += blockDim.x) s += dz[static_cast<long>(r) * cols + o];
   for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
   if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
   __syncthreads();
   if (threadIdx.x == 0) db[o] = part[0] + part[1] + part[2] + part[3];
}

__global__ void adamKernel(long n, float* __restrict__ master, __hip_bfloat16* __restrict__ w, const float* __restrict__ g,
                           float* __restrict__ m1, float* __restrict__ m2, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
   const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
   if (i >= n) return;
   const float gi = g[i];
   const float a = b1 * m1[i] + (1.0f - b1) * gi;
   const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
   m1[i] = a;
   m2[i] = b;
   const float p = master[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
   master[i] = p;
   w[i] = __float2bfloat16(p);
}

__global__ void toBf16Kernel(long n, const float* __restrict__ src, __hip_bfloat16* __restrict__ dst) {
   const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
   if (i < n) dst[i] = __float2bfloat16(src[i]);
}

__global__ void fromBf16Kernel(long n, const __hip_bfloat16* __restrict__ src, float* __restrict__ dst) {
   const long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
   if (i < n) dst[i] = __bfloat162float(src[i]);
}

unsigned blocksFor(long n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

hipError_t mlpGemm(int M, int N, int K, const float* a, long sam, long sak, const float* b, long sbk, long sbn, float* c,
                   long ldc, const float* bias, bool relu, bool accumulate, hipStream_t s) {
   if (M <= 0 || N <= 0) return hipSuccess;
   const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
   hipLaunchKernelGGL(gemmKernel<float>, grid, dim3(256), 0, s, M, N, K, a, sam, sak, b, sbk, sbn, c, ldc, bias,
                      relu ? 1 : 0, accumulate ? 1 : 0);
   return hipGetLastError();
}

hipError_t mlpGemmBf16(int M, int N, int K, const float* a, long sam, long sak, const void* b, long sbk, long sbn,
                       float* c, long ldc, const float* bias, bool relu, bool accumulate, hipStream_t s) {
   if (M <= 0 || N <= 0) return hipSuccess;
   const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
   hipLaunchKernelGGL(gemmKernel<__hip_bfloat16>, grid, dim3(256), 0, s, M, N, K, a, sam, sak,
                      static_cast<const __hip_bfloat16*>(b), sbk, sbn, c, ldc, bias, relu ? 1 : 0, accumulate ? 1 : 0);
   return hipGetLastError();
}

hipError_t mlpOutputGrad(int n, const float* z, const float* y, float* dz, float* loss, hipStream_t s) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(outputGradKernel, dim3(blocksFor(n)), dim3(256), 0, s, n, z, y, dz, loss);
   return hipGetLastError();
}

hipError_t mlpReluGrad(long count, const float* z, float* dz, hipStream_t s) {
   if (count <= 0) return hipSuccess;
   hipLaunchKernelGGL(reluGradKernel, dim3(blocksFor(count)), dim3(256), 0, s, count, z, dz);
   return hipGetLastError();
}

hipError_t mlpColumnSum(int rows, int cols, const float* dz, float* db, hipStream_t s) {
   if (cols <= 0) return hipSuccess;
   hipLaunchKernelGGL(columnSumKernel, dim3(cols), dim3(256), 0, s, rows, cols, dz, db);
   return hipGetLastError();
}

hipError_t mlpAdam(long n, float* master, void* w, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                   float eps, float c1, float c2, hipStream_t s) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(adamKernel, dim3(blocksFor(n)), dim3(256), 0, s, n, master, static_cast<__hip_bfloat16*>(w), g, m1,
                      m2, lr, b1, b2, eps, c1, c2);
   return hipGetLastError();
}

hipError_t mlpToBf16(long n, const float* src, void* dst, hipStream_t s) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(toBf16Kernel, dim3(blocksFor(n)), dim3(256), 0, s, n, src, static_cast<__hip_bfloat16*>(dst));
   return hipGetLastError();
}

hipError_t mlpFromBf16(long n, const void* src, float* dst, hipStream_t s) {
   if (n <= 0) return hipSuccess;
   hipLaunchKernelGGL(fromBf16Kernel, dim3(blocksFor(n)), dim3(256), 0, s, n, static_cast<const __hip_bfloat16*>(src), dst);
   return hipGetLastError();
}

}  // namespace rdb
