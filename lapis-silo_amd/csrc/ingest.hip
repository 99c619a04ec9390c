// Device side of ingestion: the diff of aligned sequences against the reference
// (extractCoverageAndMutationsFromSequence, common/aligned_sequence.cpp:18-122), the per-position
// coverage and symbol counts and the local-reference rule of SequenceColumn::finalize
// (sequence_column.cpp:157-212, vertical_sequence_index.cpp:86-95), and the (chunk, position,
// symbol, row) entries of the vertical index's containers (VerticalSequenceIndex::
// addSymbolsToPositions, vertical_sequence_index.cpp:16-38), radix-sorted.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "kernels.h"

namespace rdb {

namespace {

constexpr uint32_t kDiffWaves = 4;  // rows (one per wavefront) per work-group

__device__ inline uint64_t laneMaskLt(uint32_t lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

__device__ inline uint32_t firstBit(uint64_t m) { return static_cast<uint32_t>(__ffsll(static_cast<long long>(m)) - 1); }

// One wavefront per row, 64 consecutive characters per step, classified by ballots.  `j` counts
// characters of the row, its position is offset + j.  Pass 1 (kWrite false) records the covered
// range, the number of mutations and internal missing runs and the first illegal character; pass 2
// writes mutations and runs at the row's offsets.
template <bool kWrite>
__global__ __launch_bounds__(64 * kDiffWaves) void diffKernel(DiffBatch d) {
   __shared__ uint8_t lut[256];
   for (uint32_t t = threadIdx.x; t < 256; t += blockDim.x) lut[t] = d.lut[t];
   __syncthreads();
   const uint32_t lane = threadIdx.x & 63;
   const uint32_t r = blockIdx.x * kDiffWaves + (threadIdx.x >> 6);
   if (r >= d.n_rows) return;
   if (d.is_null[r]) {
      if (!kWrite && lane == 0) {
         d.start[r] = 0;
         d.end[r] = 0;
         d.n_mut[r] = 0;
         d.n_run[r] = 0;
         d.err[r] = 0xFFFFFFFFu;
      }
      return;
   }
   const char* s = d.data + d.row_ptr[r];
   const uint32_t len = static_cast<uint32_t>(d.row_ptr[r + 1] - d.row_ptr[r]);
   const uint32_t off = d.offsets ? d.offsets[r] : 0;
   const uint8_t* ref = d.ref + off;
   const uint32_t row_start = kWrite ? d.start[r] : 0, row_end = kWrite ? d.end[r] : 0;
   uint64_t mut_w = kWrite ? d.mut_off[r] : 0, run_sw = kWrite ? d.run_off[r] : 0, run_ew = run_sw;
   uint32_t first = 0xFFFFFFFFu, last = 0, n_mut = 0, n_starts = 0, err = 0xFFFFFFFFu;
   uint64_t prev_miss = 0;
   for (uint32_t base = 0; base < len; base += 64) {
      const uint32_t j = base + lane;
      const bool valid = j < len;
      const uint8_t sym = valid ? lut[static_cast<uint8_t>(s[j])] : 0;
      const bool illegal = valid && sym == 0xFF;
      const bool miss = valid && sym == d.missing;
      const bool mut = valid && !illegal && !miss && sym != ref[j];
      const uint64_t bm = __ballot(miss), bmut = __ballot(mut), bill = __ballot(illegal), bvalid = __ballot(valid);
      const uint64_t before = (bm << 1) | prev_miss;  // bit k: character k - 1 is missing
      const uint64_t starts = bm & ~before;
      const uint64_t nonmiss = bvalid & ~bm & ~bill;
      if constexpr (!kWrite) {
         n_mut += static_cast<uint32_t>(__popcll(bmut));
         n_starts += static_cast<uint32_t>(__popcll(starts));
         if (nonmiss) {
            if (first == 0xFFFFFFFFu) first = base + firstBit(nonmiss);
            last = base + 64 - static_cast<uint32_t>(__clzll(static_cast<long long>(nonmiss)));
         }
         if (bill && err == 0xFFFFFFFFu) err = off + base + firstBit(bill);
      } else {
         const uint32_t p = off + j;
         if (mut) {
            const uint64_t o = mut_w + __popcll(bmut & laneMaskLt(lane));
            d.mut_pos[o] = p;
            d.mut_sym[o] = sym;
         }
         mut_w += __popcll(bmut);
         // internal runs: starts, and ends (the first non-missing character after a run), strictly
         // inside the covered range
         const uint64_t bin = __ballot(valid && p > row_start && p < row_end);
         const uint64_t is = starts & bin;
         const uint64_t ie = nonmiss & before & bin;
         if ((is >> lane) & 1ull) d.run_start[run_sw + __popcll(is & laneMaskLt(lane))] = p;
         if ((ie >> lane) & 1ull) d.run_end[run_ew + __popcll(ie & laneMaskLt(lane))] = p;
         run_sw += __popcll(is);
         run_ew += __popcll(ie);
      }
      prev_miss = (bm >> 63) & 1ull;
   }
   if constexpr (!kWrite) {
      if (lane == 0) {
         d.err[r] = err;
         if (first == 0xFFFFFFFFu) {  // nothing covered
            d.start[r] = 0;
            d.end[r] = 0;
            d.n_mut[r] = 0;
            d.n_run[r] = 0;
         } else {
            d.start[r] = off + first;
            d.end[r] = off + last;
            d.n_mut[r] = n_mut;
            d.n_run[r] = n_starts - (first > 0 ? 1u : 0u) - (last < len ? 1u : 0u);
         }
      }
   }
}

// Row i's adaptation keys (see AdaptCount) at mut_ptr[i] + 2 nrun_ptr[i] + 2 i.
__global__ void adaptKeysKernel(AdaptCount a) {
   const DiffRowsDev& R = a.rows;
   const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
   if (i >= R.n_rows) return;
   const uint32_t past = (R.length + 1) << 5;  // sorts after every real key
   uint32_t* k = a.keys + R.mut_ptr[i] + 2 * R.nrun_ptr[i] + 2 * i;
   const bool live = !R.is_null[i];
   const uint32_t s = R.cov_start[i], e = R.cov_end[i];
   const bool cov = live && s < e;
   k[0] = cov ? (s << 5 | kAdaptCodeUp) : past;
   k[1] = cov ? (e << 5 | kAdaptCodeDown) : past;
   k += 2;
   // a null row counts nothing: whatever it lists fills its keys with `past`
   for (uint64_t r = R.nrun_ptr[i]; r < R.nrun_ptr[i + 1]; ++r) {
      *k++ = live ? (R.runs[r] & 0xFFFFu) << 5 | kAdaptCodeDown : past;
      *k++ = live ? (R.runs[r] >> 16) << 5 | kAdaptCodeUp : past;
   }
   for (uint64_t m = R.mut_ptr[i]; m < R.mut_ptr[i + 1]; ++m) *k++ = live ? R.mut_pos[m] << 5 | R.mut_sym[m] : past;
}

// Boundaries of the runs of equal sorted keys: run_start[key] and run_end[key] (count = end - start).
__global__ void adaptRunsKernel(const uint32_t* __restrict__ keys, uint64_t n, uint32_t past, uint32_t* run_start,
                                uint32_t* run_end) {
   const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
   if (i >= n) return;
   const uint32_t k = keys[i];
   if (k >= past) return;
   if (i == 0 || keys[i - 1] != k) run_start[k] = static_cast<uint32_t>(i);
   if (i + 1 == n || keys[i + 1] != k) run_end[k] = static_cast<uint32_t>(i + 1);
}

__device__ inline uint32_t adaptCount(const AdaptCount& a, uint32_t key) { return a.run_end[key] - a.run_start[key]; }

// One work-group: coverage[p] = prefix sum of the coverage events; per position the reference
// symbol's count is the coverage minus the other symbols' counts, and the local reference is the
// reference symbol unless another symbol's count is strictly larger (the first such symbol with
// the largest count).
constexpr uint32_t kAdaptThreads = 1024;
__global__ __launch_bounds__(kAdaptThreads) void adaptRuleKernel(AdaptCount a) {
   __shared__ long long part[kAdaptThreads];
   const DiffRowsDev& R = a.rows;
   const uint32_t L = R.length, nsym = R.n_symbols;
   const uint32_t per = (L + kAdaptThreads - 1) / kAdaptThreads;
   const uint32_t lo = min(L, threadIdx.x * per), hi = min(L, lo + per);
   auto change = [&](uint32_t p) -> long long {
      return static_cast<long long>(adaptCount(a, p << 5 | kAdaptCodeUp)) -
             static_cast<long long>(adaptCount(a, p << 5 | kAdaptCodeDown));
   };
   long long sum = 0;
   for (uint32_t p = lo; p < hi; ++p) sum += change(p);
   part[threadIdx.x] = sum;
   __syncthreads();
   for (uint32_t step = 1; step < kAdaptThreads; step <<= 1) {
      const long long v = threadIdx.x >= step ? part[threadIdx.x - step] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
   }
   long long coverage = part[threadIdx.x] - sum;
   for (uint32_t p = lo; p < hi; ++p) {
      coverage += change(p);
      a.coverage[p] = static_cast<uint32_t>(coverage);
      const uint32_t current = R.reference[p];
      uint32_t ref_count = static_cast<uint32_t>(coverage);
      for (uint32_t s = 0; s < nsym; ++s) {
         if (s != current) ref_count -= adaptCount(a, p << 5 | s);
      }
      uint32_t best = current, best_count = ref_count;
      for (uint32_t s = 0; s < nsym; ++s) {
         const uint32_t c = adaptCount(a, p << 5 | s);
         if (s != current && c > best_count) {
            best = s;
            best_count = c;
         }
      }
      a.local_ref[p] = static_cast<uint8_t>(best);
   }
}

// The container entries of one row: its mutations except those to a flipped position's local
// reference symbol; the reference symbol at each flipped position it covers with neither a mutation
// nor a missing run there; and the boundary pseudo-symbols of its covered range and runs (a null
// row, or one that covers nothing: uncovered from position 0).
template <bool kWrite>
__global__ void entriesKernel(EntryBatch e) {
   const DiffRowsDev& R = e.rows;
   const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
   if (i >= R.n_rows) return;
   uint32_t lo = 0, hi = e.n_chunks;  // the chunk c with chunk_offset[c] <= i < chunk_offset[c + 1]
   while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (e.chunk_offset[mid] <= i) {
         lo = mid;
      } else {
         hi = mid;
      }
   }
   const uint64_t key_hi = (static_cast<uint64_t>(lo) << 40) | (i - e.chunk_offset[lo]);
   const uint32_t L = R.length, nsym = R.n_symbols;
   const uint64_t w = kWrite ? e.offsets[i] : 0;
   uint32_t n = 0;
   auto push = [&](uint32_t p, uint32_t sym) {
      if constexpr (kWrite) e.keys[w + n] = key_hi | (static_cast<uint64_t>(p) << 24) | (static_cast<uint64_t>(sym) << 16);
      ++n;
   };
   if (R.is_null[i]) {
      push(0, nsym + BND_COV_END);
   } else {
      const bool has_cov = R.cov_start[i] < R.cov_end[i];
      const uint32_t s0 = has_cov ? R.cov_start[i] : 0, e0 = has_cov ? R.cov_end[i] : 0;
      const uint64_t m0 = R.mut_ptr[i], m1 = R.mut_ptr[i + 1];
      const uint64_t r0 = R.nrun_ptr[i], r1 = R.nrun_ptr[i + 1];
      for (uint64_t k = m0; k < m1; ++k) {
         const uint32_t p = R.mut_pos[k];
         if (e.flip_to[p] != 0xFF && R.mut_sym[k] == e.flip_to[p]) continue;
         push(p, R.mut_sym[k]);
      }
      uint64_t mk = m0, rk = r0;  // merge walk: flipped positions, mutations and runs all ascend
      for (uint32_t f = 0; f < e.n_flipped; ++f) {
         const uint32_t p = e.flipped[f];
         if (p < s0) continue;
         if (p >= e0) break;
         while (mk < m1 && R.mut_pos[mk] < p) ++mk;
         while (rk < r1 && (R.runs[rk] >> 16) <= p) ++rk;
         if (mk < m1 && R.mut_pos[mk] == p) continue;
         if (rk < r1 && (R.runs[rk] & 0xFFFFu) <= p) continue;
         push(p, R.reference[p]);
      }
      if (s0 > 0) push(s0, nsym + BND_COV_START);
      if (e0 < L) push(e0, nsym + BND_COV_END);
      for (uint64_t k = r0; k < r1; ++k) {
         push(R.runs[k] & 0xFFFFu, nsym + BND_N_START);
         if ((R.runs[k] >> 16) < L) push(R.runs[k] >> 16, nsym + BND_N_END);
      }
   }
   if constexpr (!kWrite) e.counts[i] = n;
}

}  // namespace

hipError_t launchDiff(const DiffBatch& d, bool write, hipStream_t stream) {
   if (d.n_rows == 0) return hipSuccess;
   const dim3 grid((d.n_rows + kDiffWaves - 1) / kDiffWaves);
   if (write) {
      hipLaunchKernelGGL(diffKernel<true>, grid, dim3(64 * kDiffWaves), 0, stream, d);
   } else {
      hipLaunchKernelGGL(diffKernel<false>, grid, dim3(64 * kDiffWaves), 0, stream, d);
   }
   return hipGetLastError();
}

hipError_t launchAdapt(const AdaptCount& a, void* temp, size_t* temp_bytes, hipStream_t stream) {
   const uint32_t past = (a.rows.length + 1) << 5;
   int end_bit = 1;
   while ((1u << end_bit) <= past) ++end_bit;
   hipcub::DoubleBuffer<uint32_t> buf(a.keys, a.keys_alt);
   if (temp == nullptr) {
      return hipcub::DeviceRadixSort::SortKeys(nullptr, *temp_bytes, buf, static_cast<int>(a.key_count), 0, end_bit,
                                               stream);
   }
   hipError_t err = hipSuccess;
   if (a.rows.n_rows > 0) {
      hipLaunchKernelGGL(adaptKeysKernel, dim3(static_cast<uint32_t>((a.rows.n_rows + 255) / 256)), dim3(256), 0, stream, a);
      if ((err = hipGetLastError()) != hipSuccess) return err;
   }
   if (a.key_count > 0) {
      err = hipcub::DeviceRadixSort::SortKeys(temp, *temp_bytes, buf, static_cast<int>(a.key_count), 0, end_bit, stream);
      if (err != hipSuccess) return err;
      hipLaunchKernelGGL(adaptRunsKernel, dim3(static_cast<uint32_t>((a.key_count + 255) / 256)), dim3(256), 0, stream,
                         buf.Current(), a.key_count, past, a.run_start, a.run_end);
      if ((err = hipGetLastError()) != hipSuccess) return err;
   }
   hipLaunchKernelGGL(adaptRuleKernel, dim3(1), dim3(kAdaptThreads), 0, stream, a);
   return hipGetLastError();
}

hipError_t launchEntries(const EntryBatch& e, bool write, hipStream_t stream) {
   if (e.rows.n_rows == 0) return hipSuccess;
   const dim3 grid(static_cast<uint32_t>((e.rows.n_rows + 255) / 256));
   if (write) {
      hipLaunchKernelGGL(entriesKernel<true>, grid, dim3(256), 0, stream, e);
   } else {
      hipLaunchKernelGGL(entriesKernel<false>, grid, dim3(256), 0, stream, e);
   }
   return hipGetLastError();
}

hipError_t scanEntryCounts(const uint32_t* counts, uint64_t* offsets, uint64_t n, void* temp, size_t* temp_bytes,
                           hipStream_t stream) {
   return hipcub::DeviceScan::ExclusiveScan(temp, *temp_bytes, counts, offsets, hipcub::Sum(), uint64_t{0},
                                            static_cast<int>(n), stream);
}

hipError_t sortEntryKeys(uint64_t* keys, uint64_t* keys_alt, uint64_t n, void* temp, size_t* temp_bytes,
                         hipStream_t stream, uint64_t** sorted) {
   hipcub::DoubleBuffer<uint64_t> buf(keys, keys_alt);
   const hipError_t err = hipcub::DeviceRadixSort::SortKeys(temp, *temp_bytes, buf, static_cast<int>(n), 0, 56, stream);
   if (sorted) *sorted = buf.Current();
   return err;
}

}  // namespace rdb
