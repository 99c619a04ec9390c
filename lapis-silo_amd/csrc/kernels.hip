// HIP kernels for gfx950 (MI355X): filter evaluation, Mutations counting, group-by counting.
//
// Everything here is integer, HBM-bandwidth-bound work: no MFMA.  Row sets are dense 1024 x u64
// words per 2^16-row chunk; a wavefront (64 lanes) owns a 64-word slice (4096 rows), one word per
// lane, so ballots over 64 consecutive rows map directly onto a word.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_abi.h"
#include "kernels.h"

namespace rdb {

namespace {

constexpr int kVmWaves = 4;
constexpr int kVmThreads = kVmWaves * 64;
constexpr uint32_t kVmUnitSlices = 4;      // most slices of one (program, chunk) a wavefront takes in a row
constexpr uint32_t kVmUnitMinUnits = 8192;  // units of a launch for more than one slice a unit (two per
                                            // wavefront resident on 256 CUs)
constexpr uint32_t kProgCache = 64;    // instructions of a program a wavefront keeps in LDS
constexpr int kMutWaves = 4;
constexpr int kMutThreads = kMutWaves * 64;
constexpr int kPosBlock = static_cast<int>(kMutPosBlock);
constexpr int kPosWaves = kPosBlock / 64;

__device__ inline void waveSync() {
   __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
   __builtin_amdgcn_wave_barrier();
}

__device__ inline uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Index arrays are read through global-address-space pointers: the column view holds them as
// generic pointers, which would compile to flat loads (counted against the LDS counter too, so every
// LDS wait would also wait for them).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef const u32x4 __attribute__((address_space(1))) GlobalU32x4;
typedef const uint16_t __attribute__((address_space(1))) GlobalU16;
typedef const uint32_t __attribute__((address_space(1))) GlobalU32;
typedef const uint64_t __attribute__((address_space(1))) GlobalU64;
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
typedef const u64x2 __attribute__((address_space(1))) GlobalU64x2;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef const u32x2 __attribute__((address_space(1))) GlobalU32x2;
template <typename G>
__device__ inline const G* toGlobal(const void* p) {
   return (const G*)(p);
}


__device__ inline uint64_t layoutMask(uint32_t chunk_size, uint32_t word) {
   const uint32_t lo = word * 64;
   if (chunk_size >= lo + 64) return ~0ULL;
   if (chunk_size <= lo) return 0ULL;
   return (1ULL << (chunk_size - lo)) - 1ULL;
}

// Lower bound by the whole wavefront: the first index i in [lo, hi) with !before(i), for a predicate
// that holds on a prefix.  Each round samples 64 evenly spaced indices with one load per lane and
// keeps the interval between the last sample before the bound and the first one after it, so a
// search takes about log64(n) dependent loads instead of log2(n).  Must be called by all 64 lanes
// with the same arguments; returns the same value in every lane.
template <typename Before>
__device__ inline uint32_t waveLowerBound(uint32_t lo, uint32_t hi, Before before) {
   const uint32_t lane = threadIdx.x & 63;
   while (hi - lo > 64) {
      const uint32_t step = (hi - lo + 63) / 64;
      const uint32_t idx = lo + lane * step;
      const bool b = idx < hi && before(idx);
      const uint32_t cnt = static_cast<uint32_t>(__popcll(__ballot(b)));
      if (cnt == 0) return lo;
      const uint32_t last_before = lo + (cnt - 1) * step;
      const uint32_t next = lo + cnt * step;
      lo = last_before + 1;
      if (next < hi) hi = next;
   }
   const uint32_t idx = lo + lane;
   const bool b = idx < hi && before(idx);
   return lo + static_cast<uint32_t>(__popcll(__ballot(b)));
}

__device__ inline uint64_t rangeMask(uint32_t a, uint32_t b) {
   // bits a..b inclusive of one word (0 <= a <= b <= 63)
   const uint64_t hi = (b == 63) ? ~0ULL : ((1ULL << (b + 1)) - 1ULL);
   return hi & (~0ULL << a);
}

// ORs the part of one container inside slice `slice` into the lane's word: bitsets and inline
// containers directly into `acc`, arrays and runs through the wave's LDS scratch slice (LDS
// atomics), which the caller zeroes before and reads after.
__device__ inline void accumulateContainer(const Desc d, const ContainerArenas& ar, uint32_t slice, uint32_t lane,
                                           uint64_t& acc, uint64_t* scratch) {
   const uint32_t type = descType(d.w0);
   const uint32_t card = descCard(d.w0);
   const uint32_t lo = slice * 4096u;
   const uint32_t hi = lo + 4096u;
   const uint32_t my_word = slice * 64u + lane;
   if (type == CT_INLINE) {
      for (uint32_t k = 0; k < card; ++k) {
         const uint32_t v = (d.w1 >> (16 * k)) & 0xFFFFu;
         if ((v >> 6) == my_word) acc |= 1ULL << (v & 63);
      }
   } else if (type == CT_BITSET) {
      acc |= toGlobal<GlobalU64>(ar.bits)[static_cast<size_t>(d.w1) * kChunkWords + my_word];
   } else if (type == CT_ARRAY) {
      const GlobalU16* a = toGlobal<GlobalU16>(ar.arr + d.w1);
      const uint32_t b = waveLowerBound(0, card, [&](uint32_t i) { return static_cast<uint32_t>(a[i]) < lo; });
      const uint32_t e = hi >= kChunkRows ? card
                                          : waveLowerBound(b, card, [&](uint32_t i) { return static_cast<uint32_t>(a[i]) < hi; });
      for (uint32_t k = b + lane; k < e; k += 64) {
         const uint32_t v = a[k];
         atomicOr(reinterpret_cast<unsigned long long*>(&scratch[(v >> 6) & 63]), 1ULL << (v & 63));
      }
   } else {
      const GlobalU32* r = toGlobal<GlobalU32>(ar.runs + d.w1);
      // the first run whose inclusive end reaches the slice (runs are sorted and disjoint, so their
      // ends are sorted too): a search by the whole wavefront, 64 probes per round
      const uint32_t first = waveLowerBound(0, card, [&](uint32_t i) {
         const uint32_t x = r[i];
         return (x & 0xFFFFu) + (x >> 16) < lo;
      });
      for (uint32_t k = first + lane; k < card; k += 64) {
         const uint32_t s = r[k] & 0xFFFFu;
         const uint32_t e = s + (r[k] >> 16);  // inclusive
         if (s >= hi) continue;
         const uint32_t a = s > lo ? s : lo;
         const uint32_t b = e < hi - 1 ? e : hi - 1;
         if (a > b) continue;
         for (uint32_t w = a >> 6; w <= (b >> 6); ++w) {
            const uint32_t x = w * 64 > a ? 0 : a - w * 64;
            const uint32_t y = w * 64 + 63 < b ? 63 : b - w * 64;
            atomicOr(reinterpret_cast<unsigned long long*>(&scratch[w & 63]), rangeMask(x, y));
         }
      }
   }
}

__device__ inline bool compareInt(uint32_t cmp, int32_t x, int32_t v) {
   switch (cmp) {
      case CMP_EQ: return x == v;
      case CMP_LT: return x < v;
      case CMP_GT: return x > v;
      case CMP_LE: return x <= v;
      case CMP_GE: return x >= v;
      default: return x != v;
   }
}

__device__ inline uint32_t waveSum(uint32_t v) {
   for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
   return v;
}

__device__ inline uint64_t waveSum64(uint64_t v) {
   for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
   return v;
}

// ------------------------------------------------------------------------------------------------
// Filter VM: one wavefront evaluates one program over one (chunk, slice) item; items of several
// programs (queries) are spread over the grid.  The operand stack lives in LDS, one 64-word row
// slice per entry, so any operator tree depth fits that the host sized the launch for.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kVmThreads) void filterVmKernel(TableView tv, const ProgramDesc* __restrict__ progs,
                                                             const Instr* __restrict__ instrs,
                                                             const uint32_t* __restrict__ ranges, uint32_t n_items,
                                                             uint32_t depth, uint64_t* __restrict__ out,
                                                             unsigned long long* __restrict__ cards,
                                                             uint32_t* __restrict__ chunk_cards,
                                                             const uint32_t* __restrict__ n_queries_dev,
                                                             uint32_t unit_slices) {
   extern __shared__ uint64_t lds[];
   const uint32_t wave = threadIdx.x >> 6;
   const uint32_t lane = threadIdx.x & 63;
   uint64_t* stack = lds + static_cast<size_t>(wave) * (depth + 1) * 64;
   uint64_t* scratch = stack + static_cast<size_t>(depth) * 64;
   uint32_t* pf = reinterpret_cast<uint32_t*>(lds + static_cast<size_t>(kVmWaves) * (depth + 1) * 64) +
                  static_cast<size_t>(wave) * kPrefetchSlots * 8;
   u32x4* prog = reinterpret_cast<u32x4*>(lds + static_cast<size_t>(kVmWaves) * (depth + 1) * 64 +
                                          static_cast<size_t>(kVmWaves) * kPrefetchSlots * 4) +
                 static_cast<size_t>(wave) * kProgCache * 2;
   // the unit's result words, stored together after its last slice (a store before a later load
   // would make that load's wait include the store)
   uint64_t* res = lds + static_cast<size_t>(kVmWaves) * ((depth + 1) * 64 + kPrefetchSlots * 4 + kProgCache * 4) +
                   static_cast<size_t>(wave) * kVmUnitSlices * 64;
   const uint32_t n_chunks = tv.n_chunks;
   const uint32_t per_query = n_chunks * kSlicesPerChunk;
   if (n_queries_dev) n_items = uniform(*n_queries_dev * per_query);
   // A wavefront takes units of unit_slices (1, 2 or 4) consecutive slices of one (program, chunk):
   // the program (its first kProgCache instructions) and the leaf records are loaded into LDS once
   // per unit, so that an instruction costs an LDS read instead of a memory round trip.
   const uint32_t n_units = n_items / unit_slices;
   for (uint32_t unit = uniform(blockIdx.x * kVmWaves + wave); unit < n_units; unit += gridDim.x * kVmWaves) {
   const uint32_t unit_item = unit * unit_slices;
   const uint32_t qi = unit_item / per_query;
   const uint32_t chunk = (unit_item - qi * per_query) / kSlicesPerChunk;
   const ProgramDesc pd = progs[qi];
   uint32_t prog_base = 0;  // the LDS copy holds instructions prog_base .. prog_base + kProgCache - 1
   {
      waveSync();  // the previous unit's reads of the records and the program are done
      // the program's first instructions, one per lane
      const uint32_t n_cached = pd.n_instr < kProgCache ? pd.n_instr : kProgCache;
      if (lane < n_cached) {
         const GlobalU32x4* src = toGlobal<GlobalU32x4>(instrs + pd.instr_begin + lane);
         const u32x4 a = src[0], b = src[1];
         prog[2 * lane] = a;
         prog[2 * lane + 1] = b;
      }
      // the program's leaf records for this chunk, one per lane, in one round of loads
      {
         const uint32_t n_pf = pd.n_pf < kPrefetchSlots ? pd.n_pf : kPrefetchSlots;
         if (lane < n_pf) {
            const uint32_t code = ranges[pd.pf_begin + lane];
            uint32_t* dst = pf + lane * 8;
            if (code >> 31) {
               const u32x2 d = toGlobal<GlobalU32x2>(tv.store.desc)[static_cast<size_t>(code & 0x7FFFFFFFu) * n_chunks + chunk];
               dst[0] = d[0];
               dst[1] = d[1];
            } else {
               const SeqColumnView& c = tv.seq[code >> 16];
               const GlobalU32* r = toGlobal<GlobalU32>(c.pos_ranges) + (static_cast<size_t>(chunk) * c.genome_length + (code & 0xFFFFu)) * kPosRecord;
               uint32_t v[7];
#pragma unroll
               for (uint32_t k = 0; k < 7; ++k) v[k] = r[k];
#pragma unroll
               for (uint32_t k = 0; k < 7; ++k) dst[k] = v[k];
            }
         }
      }
      waveSync();
   }
   uint32_t unit_rows = 0;  // this lane's rows of the unit's slices
   const uint32_t slice0 = (unit_item - qi * per_query) % kSlicesPerChunk;
   for (uint32_t slice = slice0; slice < slice0 + unit_slices; ++slice) {
      const uint32_t chunk_size = tv.chunk_sizes[chunk];
      const uint32_t my_word = slice * 64 + lane;
      const uint64_t lmask = layoutMask(chunk_size, my_word);
      const size_t row_base = (static_cast<size_t>(chunk) << 16) + slice * 4096u;
      uint32_t sp = 0;
      auto push = [&](uint64_t v) {
         stack[sp * 64 + lane] = v;
         ++sp;
      };
      auto pop = [&]() -> uint64_t {
         --sp;
         return stack[sp * 64 + lane];
      };
      // instruction i: from the LDS copy, or (past kProgCache) from memory
      // Instruction i from the LDS copy, a window of kProgCache instructions: a longer program
      // moves the window forward (jumps only go forward) and back to its start for the next slice,
      // one round of loads per window instead of one load per instruction.
      auto fetch = [&](uint32_t i) -> Instr {
         if (i - prog_base >= kProgCache) {
            waveSync();  // this window's readers are done
            prog_base = i;
            if (i + lane < pd.n_instr) {
               const GlobalU32x4* src = toGlobal<GlobalU32x4>(instrs + pd.instr_begin + i + lane);
               const u32x4 a = src[0], b = src[1];
               prog[2 * lane] = a;
               prog[2 * lane + 1] = b;
            }
            waveSync();
         }
         const uint32_t k = i - prog_base;
         const u32x4 a = prog[2 * k], b = prog[2 * k + 1];
         return Instr{uniform(a[0]), uniform(a[1]), uniform(a[2]), uniform(a[3]),
                      uniform(b[0]), uniform(b[1]), uniform(b[2]), uniform(b[3])};
      };
      for (uint32_t i = 0; i < pd.n_instr; ++i) {
         const Instr ins = fetch(i);
         switch (ins.op) {
            case OP_EMPTY:
               push(0);
               break;
            case OP_FULL:
               push(lmask);
               break;
            case OP_BITMAP: {
               Desc d;
               if (ins.b < kPrefetchSlots) {
                  d = Desc{pf[ins.b * 8], pf[ins.b * 8 + 1]};
               } else {
                  const u32x2 w = toGlobal<GlobalU32x2>(tv.store.desc)[static_cast<size_t>(ins.a) * n_chunks + chunk];
                  d = Desc{w[0], w[1]};
               }
               uint64_t acc = 0;
               scratch[lane] = 0;
               waveSync();
               accumulateContainer(d, tv.store, slice, lane, acc, scratch);
               waveSync();
               push(acc | scratch[lane]);
               break;
            }
            case OP_SYMBOLS: {
               // the position's entries in this chunk's tile, located by one record of six offsets
               const SeqColumnView& col = tv.seq[ins.a];
               uint32_t prv[7];
               if (ins.g < kPrefetchSlots) {
#pragma unroll
                  for (uint32_t k = 0; k < 7; ++k) prv[k] = pf[ins.g * 8 + k];
               } else {
                  const GlobalU32* pr = toGlobal<GlobalU32>(col.pos_ranges) + (static_cast<size_t>(chunk) * col.genome_length + ins.f) * kPosRecord;
#pragma unroll
                  for (uint32_t k = 0; k < 7; ++k) prv[k] = pr[k];
               }
               if (!((uniform(prv[6]) >> slice) & 1u)) {
                  // no row of the position in this slice
                  push(0);
                  break;
               }
               const uint32_t l_lo = uniform(prv[0]), l_hi = uniform(prv[1]);
               const uint32_t g_lo = uniform(prv[2]), g_hi = uniform(prv[3]);
               const uint32_t h_lo = uniform(prv[4]), h_hi = uniform(prv[5]);
               const GlobalU32* light = toGlobal<GlobalU32>(col.light);
               const GlobalU16* group_hdr = toGlobal<GlobalU16>(col.group_hdr);
               const GlobalU32x4* groups = toGlobal<GlobalU32x4>(col.groups);
               const GlobalU32x2* descs = toGlobal<GlobalU32x2>(col.vert.desc);
               uint64_t acc = 0;
               scratch[lane] = 0;
               // the first 64 light entries, the first 64 groups and the first container are loaded
               // together (most positions have no more in a chunk); an entry's slot is taken when its
               // bit (relative to the position's first slot) is set in the mask
               const bool has_l = l_hi > l_lo, has_g = g_hi > g_lo, has_h = h_hi > h_lo;
               uint32_t x0 = 0, hdr0 = 0;
               u32x4 rows0 = {0, 0, 0, 0};
               u32x2 dw0 = {0, 0};
               if (has_l) x0 = light[l_lo + lane < l_hi ? l_lo + lane : l_hi - 1];
               if (has_g) {
                  const uint32_t k = g_lo + lane < g_hi ? g_lo + lane : g_hi - 1;
                  hdr0 = group_hdr[k];
                  rows0 = groups[k];
               }
               if (has_h) dw0 = descs[h_lo];
               waveSync();
               auto light1 = [&](uint32_t x) {
                  const uint32_t row = x & 0xFFFFu;
                  if (((ins.e >> ((x >> 16) - ins.c)) & 1u) && (row >> 12) == slice) {
                     atomicOr(reinterpret_cast<unsigned long long*>(&scratch[(row >> 6) & 63]), 1ULL << (row & 63));
                  }
               };
               auto group1 = [&](uint32_t hdr, const u32x4 rows) {
                  if (!((ins.e >> ((hdr & 0xFFFu) - ins.c)) & 1u)) return;
                  const uint32_t r8[8] = {rows[0] & 0xFFFFu, rows[0] >> 16, rows[1] & 0xFFFFu, rows[1] >> 16,
                                          rows[2] & 0xFFFFu, rows[2] >> 16, rows[3] & 0xFFFFu, rows[3] >> 16};
                  const uint32_t m = (hdr >> 12) + 1;
#pragma unroll
                  for (uint32_t q = 0; q < 8; ++q) {
                     if (q < m && (r8[q] >> 12) == slice) {
                        atomicOr(reinterpret_cast<unsigned long long*>(&scratch[(r8[q] >> 6) & 63]), 1ULL << (r8[q] & 63));
                     }
                  }
               };
               if (has_l && l_lo + lane < l_hi) light1(x0);
               for (uint32_t k = l_lo + 64 + lane; k < l_hi; k += 64) light1(light[k]);
               if (has_g && g_lo + lane < g_hi) group1(hdr0, rows0);
               for (uint32_t k = g_lo + 64 + lane; k < g_hi; k += 64) group1(group_hdr[k], groups[k]);
               for (uint32_t k = h_lo; k < h_hi; ++k) {
                  const u32x2 dw = k == h_lo ? dw0 : descs[k];
                  const Desc d{dw[0], dw[1]};
                  if ((ins.e >> (descSlot(d.w0) - ins.c)) & 1u) accumulateContainer(d, col.vert, slice, lane, acc, scratch);
               }
               waveSync();
               push(acc | scratch[lane]);
               break;
            }
            case OP_COVERED: {
               // rows whose covered range holds p (each lane its 64 consecutive rows), minus the rows
               // with a missing run over p, found in the missing-run index: the runs of this slice's
               // rows that start in the buckets of [p - kShortRun + 1, p], and the long ones
               const SeqColumnView& col = tv.seq[ins.a];
               const uint32_t p = ins.b;
               const GlobalU32* cv = toGlobal<GlobalU32>(col.cov) + row_base;
               // candidate lists: lane j < n_lists loads list j's bounds
               const uint32_t nb = col.n_run_buckets;
               const uint32_t k_hi = p / kRunBucket;
               const uint32_t k_lo = p + 1 >= kShortRun ? (p + 1 - kShortRun) / kRunBucket : 0;
               const uint32_t n_lists = k_hi - k_lo + 2;  // the buckets and the long runs
               uint32_t my_lo = 0, my_len = 0;
               if (lane < n_lists) {
                  const uint32_t k = lane + 1 < n_lists ? k_lo + lane : nb;
                  const GlobalU32* rp = toGlobal<GlobalU32>(col.run_index_ptr) + (static_cast<size_t>(chunk) * (nb + 1) + k) * kSlicesPerChunk + slice;
                  my_lo = rp[0];
                  my_len = rp[1] - rp[0];
               }
               scratch[lane] = 0;
               // inside the slice's coverage bounds every row with a covered range covers p; else
               // rows are read 64 at a time, consecutive lanes on consecutive rows, and the test's
               // ballot is the word of those 64 rows, kept by the lane that owns it
               const uint32_t cb = uniform(toGlobal<GlobalU32>(col.cov_bounds)[static_cast<size_t>(chunk) * kSlicesPerChunk + slice]);
               const uint64_t never = toGlobal<GlobalU64>(col.cov_never)[static_cast<size_t>(chunk) * kChunkWords + my_word];
               uint64_t word = 0;
               if ((cb & 0xFFFFu) <= p && p < (cb >> 16)) {
                  word = ~never;
               } else {
               for (uint32_t u0 = 0; u0 < 64; u0 += 32) {
                  uint32_t cs[32];
#pragma unroll
                  for (uint32_t u = 0; u < 32; ++u) cs[u] = cv[(u0 + u) * 64 + lane];
#pragma unroll
                  for (uint32_t u = 0; u < 32; ++u) {
                     const uint64_t bw = __ballot((cs[u] & 0xFFFFu) <= p && p < (cs[u] >> 16));
                     if (lane == u0 + u) word = bw;
                  }
               }
               }
               // flat index over the lists' entries
               uint32_t lo[8], start[9];
               start[0] = 0;
#pragma unroll
               for (uint32_t j = 0; j < 8; ++j) {
                  lo[j] = __shfl(my_lo, j, 64);
                  const uint32_t len = j < n_lists ? __shfl(my_len, j, 64) : 0u;
                  start[j + 1] = start[j] + len;
               }
               const uint32_t total = start[8];
               waveSync();
               for (uint32_t base = 0; base < total; base += 64 * 4) {
                  uint64_t e[4];
#pragma unroll
                  for (uint32_t u = 0; u < 4; ++u) {
                     const uint32_t idx = base + u * 64 + lane;
                     e[u] = 0;  // row 0, an empty run [0, 0)
                     if (idx < total) {
                        uint32_t j = 0;
#pragma unroll
                        for (uint32_t t = 1; t < 8; ++t) j += idx >= start[t] ? 1u : 0u;
                        e[u] = toGlobal<GlobalU64>(col.run_index)[lo[j] + (idx - start[j])];
                     }
                  }
#pragma unroll
                  for (uint32_t u = 0; u < 4; ++u) {
                     const uint32_t a = static_cast<uint32_t>(e[u] >> 16) & 0xFFFFu;
                     const uint32_t b = static_cast<uint32_t>(e[u] >> 32) & 0xFFFFu;
                     if (a <= p && p < b) {
                        const uint32_t row = static_cast<uint32_t>(e[u]) & 0xFFFu;  // within the slice
                        atomicOr(reinterpret_cast<unsigned long long*>(&scratch[row >> 6]), 1ULL << (row & 63));
                     }
                  }
               }
               waveSync();
               push(word & ~scratch[lane] & lmask);
               break;
            }
            case OP_CMP: {
               // one or two comparisons (g & 1) of the column, conjoined; a null row matches a
               // comparison iff its with_nulls flag is set
               const IntColumnView& col = tv.ints[ins.a];
               const int32_t v = static_cast<int32_t>(ins.c);
               const int32_t v2 = static_cast<int32_t>(ins.f);
               const bool two = (ins.g & 1u) != 0;
               const uint64_t nulls = toGlobal<GlobalU64>(col.null_bits)[static_cast<size_t>(chunk) * kChunkWords + my_word];
               uint64_t word = 0;
               const uint32_t np = col.n_planes;
               if (np > 0) {
                  // bit-sliced: the lane's word of every plane, then a most-significant-first
                  // comparison of (value - plane_min) with (v - plane_min), rows below and equal
                  const GlobalU64* pl = toGlobal<GlobalU64>(col.planes) + static_cast<size_t>(chunk) * kChunkWords + my_word;
                  const size_t plane_stride = static_cast<size_t>(n_chunks) * kChunkWords;
                  uint64_t pw[kMaxIntPlanes];
#pragma unroll
                  for (uint32_t p = 0; p < kMaxIntPlanes; ++p) pw[p] = p < np ? pl[p * plane_stride] : 0ull;
                  auto compareWord = [&](uint32_t cmp, int32_t value) -> uint64_t {
                     const int64_t t = static_cast<int64_t>(value) - col.plane_min;
                     uint64_t lt = 0, eq = 0;
                     if (t >= (int64_t{1} << np)) {
                        lt = ~0ull;  // every value is below
                     } else if (t >= 0) {
                        eq = ~0ull;
#pragma unroll
                        for (int p = static_cast<int>(kMaxIntPlanes) - 1; p >= 0; --p) {
                           if (static_cast<uint32_t>(p) < np) {
                              if ((t >> p) & 1) {
                                 lt |= eq & ~pw[p];
                                 eq &= pw[p];
                              } else {
                                 eq &= ~pw[p];
                              }
                           }
                        }
                     }  // t < 0: every value is above
                     switch (cmp) {
                        case CMP_EQ: return eq;
                        case CMP_LT: return lt;
                        case CMP_GT: return ~(lt | eq);
                        case CMP_LE: return lt | eq;
                        case CMP_GE: return ~lt;
                        default: return ~eq;
                     }
                  };
                  word = compareWord(ins.b, v) & (two ? compareWord(ins.e, v2) : ~0ull);
               } else {
               // rows read 64 at a time, consecutive lanes on consecutive rows; the comparison's
               // ballot is the word of those 64 rows, kept by the lane that owns it
               const GlobalU32* vals = toGlobal<GlobalU32>(col.values) + row_base;
               for (uint32_t u0 = 0; u0 < 64; u0 += 32) {
                  uint32_t x[32];
#pragma unroll
                  for (uint32_t u = 0; u < 32; ++u) x[u] = vals[(u0 + u) * 64 + lane];
#pragma unroll
                  for (uint32_t u = 0; u < 32; ++u) {
                     const int32_t xv = static_cast<int32_t>(x[u]);
                     const uint64_t bw = __ballot(compareInt(ins.b, xv, v) && (!two || compareInt(ins.e, xv, v2)));
                     if (lane == u0 + u) word = bw;
                  }
               }
               }
               // null rows: the stored value is ignored
               const uint64_t null_match = (ins.d != 0) && (!two || (ins.g & 2u) != 0) ? nulls : 0ULL;
               word = (word & ~nulls) | null_match;
               push(word & lmask);
               break;
            }
            case OP_RANGES: {
               const uint32_t g0 = (chunk << 16) + my_word * 64u;
               uint64_t word = 0;
               for (uint32_t j = 0; j < ins.b; ++j) {
                  const uint32_t lo = ranges[pd.range_begin + 2 * (ins.a + j)];
                  const uint32_t hi = ranges[pd.range_begin + 2 * (ins.a + j) + 1];
                  const uint32_t a = lo > g0 ? lo : g0;
                  const uint32_t b = hi < g0 + 64 ? hi : g0 + 64;
                  if (a < b) word |= rangeMask(a - g0, b - 1 - g0);
               }
               push(word);
               break;
            }
            case OP_AND: {
               const uint64_t b = pop();
               const uint64_t a = pop();
               push(a & b);
               break;
            }
            case OP_OR: {
               const uint64_t b = pop();
               const uint64_t a = pop();
               push(a | b);
               break;
            }
            case OP_ANDNOT: {
               const uint64_t b = pop();
               const uint64_t a = pop();
               push(a & ~b);
               break;
            }
            case OP_NOT:
               push(pop() ^ lmask);
               break;
            case OP_TH_BEGIN:
               for (uint32_t j = 0; j <= ins.a; ++j) push(0);
               break;
            case OP_TH_ADD: {
               // Counting that reproduces Threshold's DP (threshold.cpp:64-138) bit for bit: a row
               // enters the count through `e` (the child, or its complement within the layout
               // for a negated child) and, once entered, gains one per child through `p` (the
               // child, or the plain set difference for a negated child).  Inside the layout
               // e == p and this is a population count of matching children.
               const uint64_t x = pop();
               const uint32_t base = sp - (ins.a + 1);
               const uint64_t started = stack[base * 64 + lane];
               const uint64_t p = ins.b ? ~x : x;
               const uint64_t e = ins.b ? (x ^ lmask) : x;
               uint64_t carry = (started & p) | (~started & e);
               stack[base * 64 + lane] = started | e;
               for (uint32_t j = 0; j < ins.a; ++j) {
                  const uint64_t plane = stack[(base + 1 + j) * 64 + lane];
                  stack[(base + 1 + j) * 64 + lane] = plane ^ carry;
                  carry &= plane;
               }
               break;
            }
            case OP_TH_END: {
               const uint32_t base = sp - (ins.a + 1);
               uint64_t gt = 0, eq = ~0ULL;
               for (int j = static_cast<int>(ins.a) - 1; j >= 0; --j) {
                  const uint64_t plane = stack[(base + 1 + j) * 64 + lane];
                  if ((ins.b >> j) & 1u) {
                     eq &= plane;
                  } else {
                     gt |= eq & plane;
                     eq &= ~plane;
                  }
               }
               sp = base;
               push(ins.c ? eq : (gt | eq));
               break;
            }
            case OP_TH_SKIP: {
               // counts of the rows (bit-sliced over ins.b planes): decided when count >= n (> n
               // for an exact match) or when count + left < n
               const uint32_t base = sp - (ins.b + 1);
               auto compare = [&](uint32_t k, uint64_t& gt, uint64_t& eq) {
                  gt = 0;
                  eq = ~0ULL;
                  for (int j = static_cast<int>(ins.b) - 1; j >= 0; --j) {
                     const uint64_t plane = stack[(base + 1 + j) * 64 + lane];
                     if ((k >> j) & 1u) {
                        eq &= plane;
                     } else {
                        gt |= eq & plane;
                        eq &= ~plane;
                     }
                  }
               };
               uint64_t gt_n, eq_n;
               compare(ins.c, gt_n, eq_n);
               uint64_t decided = ins.d ? gt_n : (gt_n | eq_n);
               if (ins.c > ins.e) {
                  uint64_t gt_k, eq_k;
                  compare(ins.c - ins.e, gt_k, eq_k);
                  decided |= ~(gt_k | eq_k);  // count < n - left
               }
               if (__ballot(decided != ~0ULL) == 0ULL) {
                  i = ins.a - 1;
               }
               break;
            }
            case OP_SKIP_IF_EMPTY:
               // the slice's partial conjunction is empty: the remaining conjuncts cannot change it
               if (__ballot(stack[(sp - 1) * 64 + lane] != 0ULL) == 0ULL) {
                  i = ins.a - 1;
               }
               break;
            case OP_SKIP_IF_FULL:
               if (__ballot(stack[(sp - 1) * 64 + lane] != ~0ULL) == 0ULL) {
                  i = ins.a - 1;
               }
               break;
            default:
               break;
         }
      }
      const uint64_t result = stack[lane];
      res[(slice - slice0) * 64 + lane] = result;
      unit_rows += static_cast<uint32_t>(__popcll(result));
   }
   uint64_t* dst = out + static_cast<size_t>(pd.query) * n_chunks * kChunkWords + static_cast<size_t>(chunk) * kChunkWords +
                   slice0 * 64 + lane;
   for (uint32_t k = 0; k < unit_slices; ++k) dst[k * 64] = res[k * 64 + lane];
   const uint32_t n = waveSum(unit_rows);
   if (lane == 0) {
      // A chunk's counter gathers its units' rows in the low 20 bits and their slices above; the
      // unit that completes the chunk leaves the chunk's rows alone in it and adds them to the
      // filter's cardinality, so that the cardinality takes one atomic per chunk (atomics on one
      // address serialise), and a unit waits for one atomic, not one per slice.
      uint32_t* cc = &chunk_cards[static_cast<size_t>(pd.query) * n_chunks + chunk];
      const uint32_t old = atomicAdd(cc, (unit_slices << 20) | n);
      if ((old >> 20) + unit_slices == kSlicesPerChunk) {
         const uint32_t total = (old & 0xFFFFFu) + n;
         *cc = total;
         if (total) atomicAdd(&cards[pd.query], static_cast<unsigned long long>(total));
      }
   }
   }
}

// ------------------------------------------------------------------------------------------------
// Mutations counting.  Work-group (position block b, chunk group g): for every chunk c of the group
// that the filter touches, the filter's 8 KiB tile of c is staged in LDS and every container of
// tile (b, c) is AND-popcounted against it; the per-slot sums stay in LDS and are added to the
// global per-slot counts once at the end.
// ------------------------------------------------------------------------------------------------
// The dense (index-streaming) and sparse (row-walking) Mutations kernels are both launched; the
// filter's cardinality, known on the device, activates exactly one of them.  sparse_max < 0
// disables the sparse path.
__device__ inline bool sparseActive(const unsigned long long* card, long long sparse_max) {
   return sparse_max >= 0 && *card <= static_cast<unsigned long long>(sparse_max);
}
__device__ inline bool denseActive(const unsigned long long* card, long long sparse_max) {
   return !sparseActive(card, sparse_max);
}

// bit v of a filter tile: one 32-bit LDS read (a 64-bit read spans two banks)
__device__ inline uint32_t bitAt(const uint64_t* f, uint32_t v) {
   return (reinterpret_cast<const uint32_t*>(f)[v >> 5] >> (v & 31)) & 1u;
}

__device__ inline uint32_t rangePopcount(const uint64_t* f, uint32_t a, uint32_t b) {
   // bits a..b inclusive
   uint32_t n = 0;
   const uint32_t wa = a >> 6, wb = b >> 6;
   for (uint32_t w = wa; w <= wb; ++w) {
      const uint32_t x = w == wa ? (a & 63) : 0;
      const uint32_t y = w == wb ? (b & 63) : 63;
      n += __popcll(f[w] & rangeMask(x, y));
   }
   return n;
}

// Per-slot AND-popcounts of position block b over the chunks {g + k * n_groups} against one filter;
// writes the block's slot sums to out[0 .. slots of b).  Called by a whole work-group.
//
// kPacked: two slots share one 32-bit LDS counter (16 bits each), halving the counter array so
// that four work-groups fit a CU.  A field gathers light entries and eight-row groups only (at
// most 4096 rows per slot and chunk), over at most 15 chunks per work-group (the caller's chunk
// grouping guarantees it), so it stays below 2^16; bitset and run containers, which may hold all
// 65,536 rows of a chunk, add to the global `heavy` sums instead.
constexpr uint32_t kPackedMaxChunks = 15;

template <uint32_t kUnroll, bool kPacked, bool kCoop>
__device__ __forceinline__ void countBlock(const SeqColumnView& col, uint32_t b, uint32_t g, uint32_t n_chunks,
                                           uint32_t n_groups, const uint64_t* __restrict__ filter,
                                           const uint32_t* __restrict__ filter_chunk_cards, uint32_t* __restrict__ out,
                                           uint64_t (*fs)[kChunkWords], uint32_t* cnt, uint32_t* __restrict__ heavy) {
   const uint32_t tid = threadIdx.x;
   const uint32_t lane = tid & 63;
   const uint32_t wave = tid >> 6;
   // kCoop: the work-group's four wavefronts count each chunk's tile together (one filter tile in
   // LDS, entries split over 256 threads); otherwise each wavefront takes its own chunks
   constexpr uint32_t kThreads = kCoop ? kMutThreads : 64;
   const uint32_t me = kCoop ? tid : lane;
   const uint32_t s0 = col.block_slot_start[b];
   const uint32_t ns = col.block_slot_start[b + 1] - s0;
   constexpr uint32_t kWords = kPacked ? (kMaxBlockSlots + 1) / 2 : kMaxBlockSlots + 1;
   for (uint32_t i = tid; i < kWords; i += kMutThreads) cnt[i] = 0;
   __syncthreads();
   auto addCount = [&](uint32_t slot, uint32_t v) {
      if (kPacked) {
         atomicAdd(&cnt[slot >> 1], v << ((slot & 1u) << 4));
      } else {
         atomicAdd(&cnt[slot], v);
      }
   };
   uint64_t* f = fs[kCoop ? 0 : wave];
   const uint32_t* fw = reinterpret_cast<const uint32_t*>(f);
   const Desc* __restrict__ desc = col.vert.desc;
   const GlobalU32* light_ptr = toGlobal<GlobalU32>(col.light_ptr);
   const GlobalU32* group_ptr = toGlobal<GlobalU32>(col.group_ptr);
   const GlobalU32* heavy_ptr = toGlobal<GlobalU32>(col.heavy_ptr);
   auto countTile = [&](uint32_t l0, uint32_t l1, uint32_t g0, uint32_t g1, uint32_t h0, uint32_t h1) {
         // light entries: each lane walks a contiguous span of kUnroll / 2 16-byte groups per round;
         // the next round's groups are loaded before this round's are used (kUnroll in flight), all filter lookups of a round
         // are issued before any result is used, and consecutive entries of a slot are summed before
         // one LDS atomic.  The round count is uniform over the wavefront; places past the end of the
         // tile re-read its last group and count nothing.
         {
            const GlobalU32x4* e4 = toGlobal<GlobalU32x4>(col.light + l0);
            const uint32_t n4 = (l1 - l0) >> 2;
            constexpr uint32_t kRound = kUnroll / 2;
            constexpr uint32_t kStride = kThreads * kRound;
            const uint32_t rounds = (n4 + kStride - 1) / kStride;
            uint32_t cur = kPadSlot, acc = 0;
            u32x4 bufA[kRound], bufB[kRound];
            auto load = [&](uint32_t round, u32x4* buf) {
               const uint32_t base = round * kStride + me * kRound;
   #pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
                  const uint32_t j = base + u;
                  buf[u] = e4[j < n4 ? j : n4 - 1];
               }
            };
            auto use = [&](uint32_t round, const u32x4* buf) {
               const uint32_t base = round * kStride + me * kRound;
               uint32_t word[4 * kRound];
   #pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
   #pragma unroll
                  for (uint32_t k = 0; k < 4; ++k) word[4 * u + k] = fw[(buf[u][k] & 0xFFFFu) >> 5];
               }
   #pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
                  const uint32_t valid = base + u < n4 ? 1u : 0u;
   #pragma unroll
                  for (uint32_t k = 0; k < 4; ++k) {
                     const uint32_t x = buf[u][k];
                     const uint32_t slot = x >> 16;
                     if (slot != cur) {
                        if (acc) addCount(cur, acc);
                        cur = slot;
                        acc = 0;
                     }
                     acc += (word[4 * u + k] >> (x & 31u)) & valid;
                  }
               }
            };
            if (rounds > 0) {
               load(0, bufA);
               for (uint32_t round = 0; round < rounds; round += 2) {
                  if (round + 1 < rounds) load(round + 1, bufB);
                  use(round, bufA);
                  if (round + 1 < rounds) {
                     if (round + 2 < rounds) load(round + 2, bufA);
                     use(round + 1, bufB);
                  }
               }
            }
            if (acc) addCount(cur, acc);
         }
         // larger arrays: groups of eight rows sharing one slot, walked like the light entries
         {
            const GlobalU32x4* gr = toGlobal<GlobalU32x4>(col.groups) + g0;
            const GlobalU16* gh = toGlobal<GlobalU16>(col.group_hdr) + g0;
            const uint32_t ng = g1 - g0;
            constexpr uint32_t kRound = kUnroll / 2;
            constexpr uint32_t kStride = kThreads * kRound;
            const uint32_t rounds = (ng + kStride - 1) / kStride;
            uint32_t cur = kPadSlot, acc = 0;
            u32x4 rowsA[kRound], rowsB[kRound];
            uint32_t hdrA[kRound], hdrB[kRound];
            auto load = [&](uint32_t round, u32x4* rows, uint32_t* hdr) {
               const uint32_t base = round * kStride + me * kRound;
   #pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
                  const uint32_t j = base + u < ng ? base + u : ng - 1;
                  rows[u] = gr[j];
                  hdr[u] = gh[j];
               }
            };
            auto use = [&](uint32_t round, const u32x4* rows, const uint32_t* hdr) {
               const uint32_t base = round * kStride + me * kRound;
               // lookups four groups (32 rows) at a time, to bound the registers they hold
   #pragma unroll
               for (uint32_t u0 = 0; u0 < kRound; u0 += 4) {
               constexpr uint32_t kQuad = kRound < 4 ? kRound : 4;
               uint32_t word[8 * kQuad];
   #pragma unroll
               for (uint32_t u = u0; u < u0 + kQuad; ++u) {
   #pragma unroll
                  for (uint32_t q = 0; q < 8; ++q) {
                     const uint32_t r = (q & 1u) ? rows[u][q >> 1] >> 16 : rows[u][q >> 1] & 0xFFFFu;
                     word[8 * (u - u0) + q] = fw[r >> 5];
                  }
               }
   #pragma unroll
               for (uint32_t u = u0; u < u0 + kQuad; ++u) {
                  const uint32_t h = base + u < ng ? hdr[u] : 0xFFFFu;
                  const uint32_t slot = h & 0xFFFu;
                  const uint32_t m = h == 0xFFFFu ? 0u : (h >> 12) + 1;
                  uint32_t n = 0;
   #pragma unroll
                  for (uint32_t q = 0; q < 8; ++q) {
                     const uint32_t r = (q & 1u) ? rows[u][q >> 1] >> 16 : rows[u][q >> 1];
                     n += q < m ? (word[8 * (u - u0) + q] >> (r & 31u)) & 1u : 0u;
                  }
                  if (slot != cur) {
                     if (acc) addCount(cur, acc);
                     cur = slot;
                     acc = 0;
                  }
                  acc += n;
               }
               }
            };
            if (rounds > 0) {
               load(0, rowsA, hdrA);
               for (uint32_t round = 0; round < rounds; round += 2) {
                  if (round + 1 < rounds) load(round + 1, rowsB, hdrB);
                  use(round, rowsA, hdrA);
                  if (round + 1 < rounds) {
                     if (round + 2 < rounds) load(round + 2, rowsA, hdrA);
                     use(round + 1, rowsB, hdrB);
                  }
               }
            }
            if (acc) addCount(cur, acc);
         }
         // bitset and run containers: one per wavefront
         for (uint32_t d = h0 + (kCoop ? wave : 0); d < h1; d += kCoop ? kMutWaves : 1) {
            const u32x2 dw = toGlobal<GlobalU32x2>(desc)[d];
            const Desc dd{dw[0], dw[1]};
            const uint32_t type = descType(dd.w0);
            const uint32_t card = descCard(dd.w0);
            uint32_t n = 0;
            if (type == CT_BITSET) {
               const GlobalU64x2* bits = toGlobal<GlobalU64x2>(col.vert.bits + static_cast<size_t>(dd.w1) * kChunkWords);
               u64x2 w[8];
   #pragma unroll
               for (uint32_t u = 0; u < 8; ++u) w[u] = bits[lane + 64 * u];
   #pragma unroll
               for (uint32_t u = 0; u < 8; ++u) {
                  const uint32_t j = lane + 64 * u;
                  n += __popcll(w[u][0] & f[2 * j]) + __popcll(w[u][1] & f[2 * j + 1]);
               }
            } else if (type == CT_ARRAY) {
               const GlobalU16* a = toGlobal<GlobalU16>(col.vert.arr + dd.w1);
               for (uint32_t j0 = lane; j0 < card; j0 += 256) {
                  uint32_t v[4];
   #pragma unroll
                  for (uint32_t u = 0; u < 4; ++u) v[u] = j0 + 64 * u < card ? a[j0 + 64 * u] : 0xFFFFFFFFu;
   #pragma unroll
                  for (uint32_t u = 0; u < 4; ++u) {
                     if (v[u] != 0xFFFFFFFFu) n += bitAt(f, v[u]);
                  }
               }
            } else if (type == CT_RUN) {
               const GlobalU32* r = toGlobal<GlobalU32>(col.vert.runs + dd.w1);
               for (uint32_t j = lane; j < card; j += 64) {
                  const uint32_t st = r[j] & 0xFFFFu;
                  n += rangePopcount(f, st, st + (r[j] >> 16));
               }
            } else if (lane == 0) {
               for (uint32_t j = 0; j < card; ++j) n += bitAt(f, (dd.w1 >> (16 * j)) & 0xFFFFu);
            }
            n = waveSum(n);
            if (lane == 0 && n) {
               if (kPacked) {
                  atomicAdd(&heavy[s0 + descSlot(dd.w0)], n);
               } else {
                  atomicAdd(&cnt[descSlot(dd.w0)], n);
               }
            }
         }
   };
   if constexpr (kCoop) {
      // chunks g, g + n_groups, ... with rows of the filter and a non-empty tile, in order
      auto next = [&](uint32_t c) -> uint32_t {
         for (; c < n_chunks; c += n_groups) {
            if (filter_chunk_cards[c] == 0) continue;
            const uint32_t tile = b * n_chunks + c;
            if (light_ptr[tile] != light_ptr[tile + 1] || group_ptr[tile] != group_ptr[tile + 1] ||
                heavy_ptr[tile] != heavy_ptr[tile + 1]) {
               break;
            }
         }
         return c;
      };
      // the next chunk's filter tile is loaded (two 16-byte loads per thread) while the current
      // one is counted
      u64x2 pf[2];
      auto fetch = [&](uint32_t c) {
         const GlobalU64x2* src = toGlobal<GlobalU64x2>(filter + static_cast<size_t>(c) * kChunkWords);
         pf[0] = src[tid];
         pf[1] = src[tid + kMutThreads];
      };
      uint32_t c = next(g);
      if (c < n_chunks) fetch(c);
      while (c < n_chunks) {
         __syncthreads();  // the previous tile's readers are done
         reinterpret_cast<u64x2*>(f)[tid] = pf[0];
         reinterpret_cast<u64x2*>(f)[tid + kMutThreads] = pf[1];
         const uint32_t tile = b * n_chunks + c;
         const uint32_t l0 = light_ptr[tile], l1 = light_ptr[tile + 1];
         const uint32_t h0 = heavy_ptr[tile], h1 = heavy_ptr[tile + 1];
         const uint32_t g0 = group_ptr[tile], g1 = group_ptr[tile + 1];
         const uint32_t cn = next(c + n_groups);
         if (cn < n_chunks) fetch(cn);
         __syncthreads();
         countTile(l0, l1, g0, g1, h0, h1);
         c = cn;
      }
   } else {
      // each wavefront takes its own chunks of the group; no block-wide barrier until the end
      for (uint32_t k = wave;; k += kMutWaves) {
         const uint32_t c = g + k * n_groups;
         if (c >= n_chunks) break;
         // the chunk's cardinality and tile offsets in one round of loads
         const uint32_t card = toGlobal<GlobalU32>(filter_chunk_cards)[c];
         const uint32_t tile = b * n_chunks + c;
         const uint32_t l0 = light_ptr[tile], l1 = light_ptr[tile + 1];
         const uint32_t h0 = heavy_ptr[tile], h1 = heavy_ptr[tile + 1];
         const uint32_t g0 = group_ptr[tile], g1 = group_ptr[tile + 1];
         if (card == 0 || (l0 == l1 && h0 == h1 && g0 == g1)) continue;
         {
            // stage the 8 KiB filter tile: eight independent 16-byte loads per lane
            const GlobalU64x2* src = toGlobal<GlobalU64x2>(filter + static_cast<size_t>(c) * kChunkWords);
            u64x2 w[8];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) w[u] = src[lane + 64 * u];
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) reinterpret_cast<u64x2*>(f)[lane + 64 * u] = w[u];
         }
         waveSync();
         countTile(l0, l1, g0, g1, h0, h1);
         waveSync();
      }
   }
   __syncthreads();
   for (uint32_t i = tid; i < ns; i += kMutThreads) {
      out[s0 + i] = kPacked ? (cnt[i >> 1] >> ((i & 1u) << 4)) & 0xFFFFu : cnt[i];
   }
}


template <uint32_t kUnroll, bool kPacked, bool kCoop>
// Occupancy target: with the work-group sharing each tile and two 16-byte loads in flight per lane,
// the kernel fits 73 VGPRs and six wavefronts per SIMD (the 24 KiB of LDS a work-group takes allow
// six per CU); larger unrolls keep four, where they do not spill.
__global__ __launch_bounds__(kMutThreads) __attribute__((amdgpu_waves_per_eu(kCoop && kUnroll <= 4 ? 6 : (kPacked || kCoop ? 4 : 3)))) void mutationsCountKernel(const SeqColumnView* __restrict__ colp,
                                                                    uint32_t n_chunks, uint32_t n_groups,
                                                                    const uint64_t* __restrict__ filter,
                                                                    const uint32_t* __restrict__ filter_chunk_cards,
                                                                    const unsigned long long* __restrict__ filter_card,
                                                                    long long sparse_max,
                                                                    uint32_t* __restrict__ partial,
                                                                    uint32_t* __restrict__ heavy) {
   __shared__ uint64_t fs[kCoop ? 1 : kMutWaves][kChunkWords];
   __shared__ uint32_t cnt[kPacked ? (kMaxBlockSlots + 1) / 2 : kMaxBlockSlots + 1];
   if (!denseActive(filter_card, sparse_max)) return;
   const SeqColumnView& col = *colp;
   countBlock<kUnroll, kPacked, kCoop>(col, blockIdx.x, blockIdx.y, n_chunks, n_groups, filter, filter_chunk_cards,
                                partial + static_cast<size_t>(blockIdx.y) * col.n_slots, fs, cnt, heavy);
}


// ------------------------------------------------------------------------------------------------
// Sparse Mutations counting: each filter row adds its own stored mutations (row-major CSR in the
// original-reference frame), its covered-range boundaries and its internal missing runs with
// global atomics.  Work is proportional to the filter's rows, not to the index.
//   counts[s * L + p]  rows of F with symbol s at p (s != reference[p])
//   bnd[p], bnd[L + p], bnd[2L + p]  A(p), B(p), C(p) as in the dense path
// Both arrays must be zero on entry; the sparse finalize kernel clears what it reads.
// ------------------------------------------------------------------------------------------------
// Sparse counts go through a per-work-group LDS table before global memory: rows of one filter share
// most of their mutations (a lineage's) and their covered-range ends, so global atomics would pile
// onto a few addresses.  Keys: a counts index (symbol * L + position), or kSparseBndKey | a bnd index.
constexpr uint32_t kSparseTable = 4096;
constexpr uint32_t kSparseProbes = 16;
constexpr uint32_t kSparseEmpty = 0xFFFFFFFFu;
constexpr uint32_t kSparseBndKey = 1u << 30;

struct SparseTable {
   uint32_t* keys;
   int32_t* vals;
   uint32_t* counts;
   int32_t* bnd;

   __device__ void clear() const {
      for (uint32_t i = threadIdx.x; i < kSparseTable; i += blockDim.x) {
         keys[i] = kSparseEmpty;
         vals[i] = 0;
      }
   }
   // adds v under key: in the table when the key has a place within kSparseProbes probes, else
   // straight to global memory
   __device__ void add(uint32_t key, int32_t v) const {
      uint32_t h = (key * 2654435761u) >> 20;  // 12 bits
      for (uint32_t probe = 0; probe < kSparseProbes; ++probe) {
         uint32_t k = keys[h];
         if (k == kSparseEmpty) k = atomicCAS(&keys[h], kSparseEmpty, key) == kSparseEmpty ? key : keys[h];
         if (k == key) {
            atomicAdd(&vals[h], v);
            return;
         }
         h = (h + 1) & (kSparseTable - 1);
      }
      flushOne(key, v);
   }
   __device__ void flushOne(uint32_t key, int32_t v) const {
      if (key & kSparseBndKey) {
         atomicAdd(&bnd[key & ~kSparseBndKey], v);
      } else {
         atomicAdd(&counts[key], static_cast<uint32_t>(v));
      }
   }
   __device__ void flush() const {
      for (uint32_t i = threadIdx.x; i < kSparseTable; i += blockDim.x) {
         if (keys[i] != kSparseEmpty && vals[i] != 0) flushOne(keys[i], vals[i]);
      }
   }
};

// Rows of a filter: a wavefront takes 64 filter words (one per lane) and visits their rows
// kRowsPerStep at a time, 64 / kRowsPerStep lanes per row; lane l of a row's group adds its
// mutations l, l + 64 / kRowsPerStep, ... and missing runs likewise, the group's first lane its
// covered-range boundaries, all through the work-group's table.  A step costs about two dependent
// loads.
constexpr uint32_t kRowsPerStep = 8;
__device__ inline void sparseCountRows(const SeqColumnView& col, const uint32_t* chunk_sizes, const uint64_t* filter,
                                       uint32_t n_words, const SparseTable& tb, uint32_t wave_id, uint32_t n_waves) {
   const uint32_t lane = threadIdx.x & 63;
   const uint32_t L = col.genome_length;
   const GlobalU32* cov = toGlobal<GlobalU32>(col.cov);
   const GlobalU32* nrun_ptr = toGlobal<GlobalU32>(col.nrun_ptr);
   const GlobalU32* nruns = toGlobal<GlobalU32>(col.nruns);
   const GlobalU32* mut_ptr = toGlobal<GlobalU32>(col.row_mut_ptr);
   const GlobalU32* muts = toGlobal<GlobalU32>(col.row_muts);
   const GlobalU64* fw = toGlobal<GlobalU64>(filter);
   // the next 64 words are loaded while a step's rows are counted
   auto words = [&](uint32_t base) -> uint64_t {
      const uint32_t wd = base + lane;
      const uint32_t wc = wd < n_words ? wd : n_words - 1;
      // bits outside the row layout (left by Threshold's complement of negated children) have no
      // entries in the index: the dense path counts them covered with the local reference
      const uint64_t x = fw[wc] & layoutMask(toGlobal<GlobalU32>(chunk_sizes)[wc >> 10], wc & (kChunkWords - 1));
      return wd < n_words ? x : 0ull;
   };
   uint64_t next = wave_id * 64 < n_words ? words(wave_id * 64) : 0ull;
   for (uint32_t base = wave_id * 64; base < n_words; base += n_waves * 64) {
      uint64_t x = next;
      if (base + n_waves * 64 < n_words) next = words(base + n_waves * 64);
      uint64_t pending = __ballot(x != 0);
      while (pending) {
         // up to kRowsPerStep rows at a time, 64 / kRowsPerStep lanes each
         uint32_t rs[kRowsPerStep] = {};
         uint32_t nr = 0;
#pragma unroll
         for (uint32_t j = 0; j < kRowsPerStep; ++j) {
            if (pending) {
               const uint32_t src = static_cast<uint32_t>(__builtin_ctzll(pending));
               const uint64_t xs = __shfl(x, static_cast<int>(src), 64);
               rs[j] = ((base + src) << 6) | static_cast<uint32_t>(__builtin_ctzll(xs));
               if (lane == src) x &= x - 1;
               pending = __ballot(x != 0);
               nr = j + 1;
            }
         }
         constexpr uint32_t kLanesPerRow = 64 / kRowsPerStep;
         const uint32_t g = lane / kLanesPerRow, sub = lane % kLanesPerRow;
         const bool have = g < nr;
         uint32_t r = rs[0];
#pragma unroll
         for (uint32_t j = 1; j < kRowsPerStep; ++j) r = g == j ? rs[j] : r;
         const uint32_t cv = cov[r];
         const uint32_t n0 = nrun_ptr[r], n1 = nrun_ptr[r + 1];
         const uint32_t m0 = mut_ptr[r], m1 = mut_ptr[r + 1];
         if (have) {
            if (sub == 0) {
               const uint32_t st = cv & 0xFFFFu, en = cv >> 16;
               if (st > 0) tb.add(kSparseBndKey | st, 1);
               if (en < L) tb.add(kSparseBndKey | (L + en), 1);
            }
            for (uint32_t k = n0 + sub; k < n1; k += kLanesPerRow) {
               const uint32_t run = nruns[k];
               tb.add(kSparseBndKey | (2 * L + (run & 0xFFFFu)), 1);
               if ((run >> 16) < L) tb.add(kSparseBndKey | (2 * L + (run >> 16)), -1);
            }
            for (uint32_t k = m0 + sub; k < m1; k += kLanesPerRow) {
               const uint32_t mu = muts[k];
               tb.add((mu >> 16) * L + (mu & 0xFFFFu), 1);
            }
         }
      }
   }
}

__global__ __launch_bounds__(256) void sparseCountKernel(const SeqColumnView* __restrict__ colp,
                                                         const uint32_t* __restrict__ chunk_sizes, uint32_t n_words,
                                                         const uint64_t* __restrict__ filter,
                                                         const unsigned long long* __restrict__ filter_card,
                                                         long long sparse_max, uint32_t* __restrict__ counts,
                                                         int32_t* __restrict__ bnd) {
   __shared__ uint32_t keys[kSparseTable];
   __shared__ int32_t vals[kSparseTable];
   if (!sparseActive(filter_card, sparse_max)) return;
   const SparseTable tb{keys, vals, counts, bnd};
   tb.clear();
   __syncthreads();
   sparseCountRows(*colp, chunk_sizes, filter, n_words, tb, blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6),
                   gridDim.x * (blockDim.x / 64));
   __syncthreads();
   tb.flush();
}

// Sums the chunk groups' partial slot counts into group 0, in place (one thread per slot).
__global__ __launch_bounds__(256) void reducePartialsKernel(uint32_t* __restrict__ partial, uint32_t n_groups,
                                                            uint32_t n_slots,
                                                            const unsigned long long* __restrict__ filter_card,
                                                            long long sparse_max, uint32_t* __restrict__ heavy) {
   if (!denseActive(filter_card, sparse_max)) return;
   const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
   if (s >= n_slots) return;
   uint32_t v = partial[s];
   for (uint32_t g = 1; g < n_groups; ++g) v += partial[static_cast<size_t>(g) * n_slots + s];
   if (heavy) {
      v += heavy[s];
      heavy[s] = 0;
   }
   partial[s] = v;
}

// Values of one block of kPosBlock positions, staged in LDS by the totals and finalize kernels:
// stored counts per symbol and the boundary quantities
//   a: rows of the filter whose covered range starts at p (p > 0),
//   b: rows whose covered range ends at p,
//   c: internal missing runs starting at p minus those ending at p.
constexpr uint32_t kMaxSymbols = 28;
struct BlockValues {
   uint32_t v[kMaxSymbols][kPosBlock];
   uint32_t a[kPosBlock];
   uint32_t b[kPosBlock];
   int32_t c[kPosBlock];
};

// Sources of the per-position quantities for the totals and finalize kernels; stage() fills a
// BlockValues for positions [p0, p0 + kPosBlock) and is called by the whole work-group.
//  * SlotSource: per-slot sums of the dense kernel (one thread per slot of the block); the implicit
//    symbol (rows covered at p with no stored entry) is the local reference.
//  * SparseSource: per-(symbol, position) counts of the sparse kernel in the original-reference
//    frame (one thread per position); the implicit symbol is the original reference.  With
//    `clear`, the counts read are zeroed for the next query.
struct SlotSource {
   const SeqColumnView* col;
   const uint32_t* partial;
   // further parts of the batched shared scans, summed here: part two, parts three and four
   // (`n_more` planes `plane` words apart) and the heavy-container sums (zeroed when staged with
   // `clear`, for the next batch)
   const uint32_t* partial_b = nullptr;
   const uint32_t* partial_c = nullptr;
   size_t plane = 0;
   uint32_t n_more = 0;
   uint32_t* heavy = nullptr;

   __device__ uint32_t value(uint32_t s, bool clear) const {
      uint32_t v = partial[s];
      if (partial_b) v += partial_b[s];
      for (uint32_t k = 0; k < n_more; ++k) v += partial_c[k * plane + s];
      if (heavy) {
         const uint32_t h = heavy[s];
         v += h;
         if (clear && h) heavy[s] = 0;
      }
      return v;
   }

   __device__ void stage(uint32_t p0, BlockValues& bv, bool clear) const {
      const uint32_t t = threadIdx.x;
      const uint32_t ns = col->n_symbols;
      for (uint32_t k = 0; k < ns; ++k) bv.v[k][t] = 0;
      bv.a[t] = 0;
      bv.b[t] = 0;
      bv.c[t] = 0;
      __syncthreads();
      const uint32_t L = col->genome_length;
      const GlobalU32* pos_ptr = toGlobal<GlobalU32>(col->pos_slot_ptr);
      const GlobalU32* info = toGlobal<GlobalU32>(col->slot_info);
      const uint32_t s0 = pos_ptr[p0];
      const uint32_t s1 = pos_ptr[p0 + kPosBlock < L ? p0 + kPosBlock : L];
      for (uint32_t s = s0 + t; s < s1; s += kPosBlock) {
         const uint32_t x = info[s];
         const uint32_t v = value(s, clear);
         const uint32_t pl = (x & 0xFFFFu) - p0;
         const uint32_t sym = x >> 16;
         if (sym < ns) {
            bv.v[sym][pl] = v;
         } else {
            switch (sym - ns) {
               case BND_COV_START: bv.a[pl] = v; break;
               case BND_COV_END: bv.b[pl] = v; break;
               case BND_N_START: atomicAdd(&bv.c[pl], static_cast<int32_t>(v)); break;
               default: atomicSub(&bv.c[pl], static_cast<int32_t>(v)); break;
            }
         }
      }
      __syncthreads();
   }
   __device__ uint32_t implicitSymbol(uint32_t p) const { return col->local_ref[p]; }
   // this thread's share of the block sums of a, b and c (boundary slots are the last of a position)
   __device__ void blockSums(uint32_t p0, int64_t& a, int64_t& b, int64_t& c) const {
      const uint32_t L = col->genome_length;
      const uint32_t ns = col->n_symbols;
      const GlobalU32* pos_ptr = toGlobal<GlobalU32>(col->pos_slot_ptr);
      const GlobalU32* info = toGlobal<GlobalU32>(col->slot_info);
      const uint32_t s0 = pos_ptr[p0];
      const uint32_t s1 = pos_ptr[p0 + kPosBlock < L ? p0 + kPosBlock : L];
      for (uint32_t s = s0 + threadIdx.x; s < s1; s += kPosBlock) {
         const uint32_t sym = info[s] >> 16;
         if (sym < ns) continue;
         const int64_t v = value(s, false);
         switch (sym - ns) {
            case BND_COV_START: a += v; break;
            case BND_COV_END: b += v; break;
            case BND_N_START: c += v; break;
            default: c -= v; break;
         }
      }
   }
};

struct SparseSource {
   const SeqColumnView* col;
   uint32_t* counts;
   int32_t* bnd;

   __device__ void stage(uint32_t p0, BlockValues& bv, bool clear) const {
      const uint32_t t = threadIdx.x;
      const uint32_t L = col->genome_length;
      const uint32_t p = p0 + t;
      const uint32_t ns = col->n_symbols;
      if (p < L) {
         for (uint32_t k = 0; k < ns; ++k) {
            bv.v[k][t] = counts[static_cast<size_t>(k) * L + p];
            if (clear) counts[static_cast<size_t>(k) * L + p] = 0;
         }
         bv.a[t] = static_cast<uint32_t>(bnd[p]);
         bv.b[t] = static_cast<uint32_t>(bnd[L + p]);
         bv.c[t] = bnd[2 * L + p];
         if (clear) {
            bnd[p] = 0;
            bnd[L + p] = 0;
            bnd[2 * L + p] = 0;
         }
      } else {
         for (uint32_t k = 0; k < ns; ++k) bv.v[k][t] = 0;
         bv.a[t] = 0;
         bv.b[t] = 0;
         bv.c[t] = 0;
      }
      __syncthreads();
   }
   __device__ uint32_t implicitSymbol(uint32_t p) const { return col->reference[p]; }
   __device__ void blockSums(uint32_t p0, int64_t& a, int64_t& b, int64_t& c) const {
      const uint32_t L = col->genome_length;
      const uint32_t p = p0 + threadIdx.x;
      if (p < L) {
         a += static_cast<uint32_t>(bnd[p]);
         b += static_cast<uint32_t>(bnd[L + p]);
         c += bnd[2 * L + p];
      }
   }
};

__device__ inline bool sourceActive(const SlotSource&, const unsigned long long* card, long long sparse_max) {
   return denseActive(card, sparse_max);
}
__device__ inline bool sourceActive(const SparseSource&, const unsigned long long* card, long long sparse_max) {
   return sparseActive(card, sparse_max);
}

template <typename T>
__device__ inline T waveInclusiveScan(T v) {
   const uint32_t lane = threadIdx.x & 63;
   for (int off = 1; off < 64; off <<= 1) {
      const T y = __shfl_up(v, off, 64);
      if (lane >= static_cast<uint32_t>(off)) v += y;
   }
   return v;
}

template <typename T>
__device__ inline T waveTotal(T v) {
   for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
   return v;
}

// Inclusive scan over the kPosBlock threads of a work-group; tmp holds kPosWaves values.
template <typename T>
__device__ inline T blockInclusiveScan(T v, T* tmp) {
   const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   v = waveInclusiveScan(v);
   if (lane == 63) tmp[wave] = v;
   __syncthreads();
   T prefix = 0;
   for (uint32_t w = 0; w < wave; ++w) prefix += tmp[w];
   __syncthreads();
   return v + prefix;
}

// Per kPosBlock-position block: totals of the three boundary quantities a, b and c.
template <typename Src>
__device__ __forceinline__ void totalsBody(const Src& src, uint32_t pblock, int64_t* __restrict__ totals) {
   __shared__ int64_t tmp[3][kPosWaves];
   int64_t v[3] = {0, 0, 0};
   src.blockSums(pblock * kPosBlock, v[0], v[1], v[2]);
   const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   for (int q = 0; q < 3; ++q) {
      const int64_t x = waveTotal(v[q]);
      if (lane == 0) tmp[q][wave] = x;
   }
   __syncthreads();
   if (threadIdx.x < 3) {
      int64_t sum = 0;
      for (int w = 0; w < kPosWaves; ++w) sum += tmp[threadIdx.x][w];
      totals[pblock * 3 + threadIdx.x] = sum;
   }
}

template <typename Src>
__global__ __launch_bounds__(kPosBlock) void mutationsTotalsKernel(Src src, const unsigned long long* __restrict__ filter_card,
                                                                   long long sparse_max, int64_t* __restrict__ totals,
                                                                   uint32_t* __restrict__ n_rows) {
   if (!sourceActive(src, filter_card, sparse_max)) return;
   if (blockIdx.x == 0 && threadIdx.x == 0) *n_rows = 0;  // the finalize kernel appends rows after this one
   totalsBody(src, blockIdx.x, totals);
}

// LDS of the finalize step, one instance per kernel whichever sources it finalizes
struct FinalizeShared {
   BlockValues bv;
   int64_t scan_tmp[kPosWaves];
   int64_t part[4][kPosWaves];
   int64_t prefix[4];  // previous-block sums of a, b, c and the total of a
};
__device__ inline FinalizeShared& finalizeShared() {
   __shared__ FinalizeShared sh;
   return sh;
}

template <typename Src>
__device__ __forceinline__ void finalizeBody(const Src& src, uint32_t pblock, const int64_t* __restrict__ totals,
                                             uint32_t n_pblocks, uint64_t filter_card, double min_proportion,
                                             uint32_t* __restrict__ counts_out, MutationRowDev* __restrict__ rows,
                                             uint32_t row_capacity, uint32_t* __restrict__ n_rows) {
   FinalizeShared& sh = finalizeShared();
   BlockValues& bv = sh.bv;
   int64_t* scan_tmp = sh.scan_tmp;
   int64_t (*part)[kPosWaves] = sh.part;
   int64_t* prefix = sh.prefix;
   const SeqColumnView& col = *src.col;
   {
      // block totals: thread k < n_pblocks (<= kPosBlock for genome lengths below 2^16) takes block k
      const uint32_t k = threadIdx.x;
      int64_t a = 0, b = 0, c = 0, ta = 0;
      if (k < n_pblocks) {
         ta = totals[3 * k];
         if (k < pblock) {
            a = ta;
            b = totals[3 * k + 1];
            c = totals[3 * k + 2];
         }
      }
      const int64_t v[4] = {a, b, c, ta};
      const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      for (int q = 0; q < 4; ++q) {
         const int64_t x = waveTotal(v[q]);
         if (lane == 0) part[q][wave] = x;
      }
      __syncthreads();
      if (threadIdx.x < 4) {
         int64_t sum = 0;
         for (int w = 0; w < kPosWaves; ++w) sum += part[threadIdx.x][w];
         prefix[threadIdx.x] = sum;
      }
   }
   src.stage(pblock * kPosBlock, bv, true);  // ends with a barrier, which also publishes prefix
   const uint32_t L = col.genome_length;
   const uint32_t t = threadIdx.x;
   const uint32_t p = pblock * kPosBlock + t;
   const int64_t pa = blockInclusiveScan<int64_t>(bv.a[t], scan_tmp) + prefix[0];
   const int64_t pb = blockInclusiveScan<int64_t>(bv.b[t], scan_tmp) + prefix[1];
   const int64_t pc = blockInclusiveScan<int64_t>(bv.c[t], scan_tmp) + prefix[2];
   if (p >= L) return;
   const uint32_t ns = col.n_symbols;
   const int64_t card = static_cast<int64_t>(filter_card);
   // rows covered at p: not starting after p, not ended at or before p, not inside a missing run
   const int64_t covered = card - (prefix[3] - pa) - pb - pc;
   int64_t stored = 0;
   for (uint32_t sym = 0; sym < ns; ++sym) stored += bv.v[sym][t];
   const uint32_t implicit_count = card == 0 ? 0u : static_cast<uint32_t>(covered - stored);
   const uint32_t implicit = src.implicitSymbol(p);
   // per-symbol counts in symbol order
   uint32_t total = 0;
   for (uint32_t sym = 0; sym < ns; ++sym) {
      uint32_t v = bv.v[sym][t];
      if (sym == implicit) v += implicit_count;
      if (counts_out) counts_out[static_cast<size_t>(sym) * L + p] = v;
      if ((col.valid_mask >> sym) & 1u) total += v;
   }
   if (total != 0 && rows != nullptr) {
      const uint32_t thr = min_proportion == 0.0
                              ? 0u
                              : static_cast<uint32_t>(ceil(static_cast<double>(total) * min_proportion) - 1.0);
      const uint32_t orig = col.reference[p];
      for (uint32_t sym = 0; sym < ns; ++sym) {
         uint32_t v = bv.v[sym][t];
         if (sym == implicit) v += implicit_count;
         if (!((col.valid_mask >> sym) & 1u) || sym == orig) continue;
         if (v > thr) {
            const uint32_t idx = atomicAdd(n_rows, 1u);
            if (idx < row_capacity) rows[idx] = MutationRowDev{p, sym, v, total};
         }
      }
   }
}

template <typename Src>
__global__ __launch_bounds__(kPosBlock) void mutationsFinalizeKernel(
   Src src, const int64_t* __restrict__ totals, uint32_t n_pblocks, const unsigned long long* __restrict__ filter_card,
   long long sparse_max, double min_proportion, uint32_t* __restrict__ counts_out, MutationRowDev* __restrict__ rows,
   uint32_t row_capacity, uint32_t* __restrict__ n_rows) {
   if (!sourceActive(src, filter_card, sparse_max)) return;
   finalizeBody(src, blockIdx.x, totals, n_pblocks, *filter_card, min_proportion, counts_out, rows, row_capacity, n_rows);
}

// ------------------------------------------------------------------------------------------------
// Batched Mutations (see MutationsBatch).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void batchPlanKernel(MutationsBatch m) {
   __shared__ uint32_t n_dense, n_sparse;
   if (threadIdx.x == 0) {
      n_dense = 0;
      n_sparse = 0;
   }
   __syncthreads();
   const uint32_t nq = *m.n_queries;
   for (uint32_t q = threadIdx.x; q < nq && q < m.max_queries; q += blockDim.x) {
      m.n_rows[q] = 0;
      if (!m.is_mutations[q] || m.cards[q] == 0) continue;
      if (m.sparse_max >= 0 && m.cards[q] <= static_cast<unsigned long long>(m.sparse_max)) {
         m.lists[m.max_queries + atomicAdd(&n_sparse, 1u)] = q;
      } else {
         m.lists[atomicAdd(&n_dense, 1u)] = q;
      }
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      m.list_counts[0] = n_dense;
      m.list_counts[1] = n_sparse;
      m.list_counts[2] = 0;  // work counter of the persistent octet kernel
   }
}

// Work-group w handles position block b of dense list entry i.  Blocks are dealt to the eight XCDs
// (work-group w runs on XCD w % 8) so that the entries sharing a block run on one XCD, close in
// time, and its L2 serves the block's tiles to all but the first of them.
template <uint32_t kUnroll, bool kCoop>
__global__ __launch_bounds__(kMutThreads) __attribute__((amdgpu_waves_per_eu(kCoop ? 4 : 3))) void mutationsCountBatchKernel(MutationsBatch m) {
   __shared__ uint64_t fs[kCoop ? 1 : kMutWaves][kChunkWords];
   __shared__ uint32_t cnt[kMaxBlockSlots + 1];
   const uint32_t n_dense = m.list_counts[0];
   if (n_dense == 0) return;
   const SeqColumnView& col = *m.col_dev;
   const uint32_t w = blockIdx.x;
   const uint32_t xcd = w & 7u;
   const uint32_t i = w >> 3;
   const uint32_t b = xcd + 8u * (i / n_dense);
   if (b >= col.n_blocks) return;
   const uint32_t e = i % n_dense;
   const uint32_t q = m.lists[e];
   countBlock<kUnroll, false, kCoop>(col, b, 0, m.n_chunks, 1, m.filters + static_cast<size_t>(q) * m.n_chunks * kChunkWords,
                       m.chunk_cards + static_cast<size_t>(q) * m.n_chunks, m.partial + static_cast<size_t>(e) * col.n_slots,
                              fs, cnt, nullptr);
}

// ---- shared scan: four dense queries per pass over the index --------------------------------------
// Work-group (block b, quad j, part h) streams block b's tiles of the chunks of part h (first or
// second half of the chunks) once for the dense list entries 4j .. 4j+3.  Per chunk, the four
// filters are staged in LDS interleaved by 32-bit word (word i of entry 4j+k at 4 i + k), so that
// one 16-byte LDS read gives a row's word of all four filters.  Light entries and group rows are
// counted as in the one-query kernel, into four 16-bit counters per slot held in one u64 (one LDS
// atomic per run of a slot).  A counter gathers at most kSharedFlushChunks chunks x 4096 rows before
// the work-group adds it to its part's slot sums (which it owns), so it cannot overflow.  Bitset and
// run containers are counted per query (two wavefronts per bitset, a half each) and added with
// global atomics to heavy_sums; the totals and finalize kernels add the second part and heavy_sums
// to partial as they read it.  The next chunk's filter words are loaded while the current tile is
// counted.
constexpr uint32_t kSharedThreads = 512;
constexpr uint32_t kSharedFlushChunks = 15;

// popcount of bits a..b (inclusive) of query k's words in an interleaved tile
__device__ inline uint32_t rangePopcountQ(const uint32_t* fq, uint32_t k, uint32_t a, uint32_t b) {
   uint32_t n = 0;
   const uint32_t wa = a >> 5, wb = b >> 5;
   for (uint32_t w = wa; w <= wb; ++w) {
      const uint32_t x = w == wa ? (a & 31) : 0;
      const uint32_t y = w == wb ? (b & 31) : 31;
      const uint32_t mask = (y == 31 ? 0xFFFFFFFFu : ((1u << (y + 1)) - 1u)) & ~((1u << x) - 1u);
      n += __popc(fq[4 * w + k] & mask);
   }
   return n;
}

// the four filter bits of `row` packed as 16-bit lanes: (bit 0 | bit 1 << 16, bit 2 | bit 3 << 16)
__device__ inline void quadBits(const u32x4 w, uint32_t row, uint32_t& a01, uint32_t& a23) {
   const uint32_t sh = row & 31u;
   a01 += ((w[0] >> sh) & 1u) | (((w[1] >> sh) & 1u) << 16);
   a23 += ((w[2] >> sh) & 1u) | (((w[3] >> sh) & 1u) << 16);
}

template <uint32_t kRound>
__global__ __launch_bounds__(kSharedThreads) __attribute__((amdgpu_waves_per_eu(4))) void mutationsCountSharedKernel(MutationsBatch m) {
   __shared__ __align__(16) uint32_t fq[kChunkWords * 2 * 4];
   __shared__ unsigned long long cnt[kMaxBlockSlots + 1];
   __shared__ uint8_t work[kSharedMaxPartChunks];
   const u32x4* fq4 = reinterpret_cast<const u32x4*>(fq);
   const uint32_t n_dense = m.list_counts[0];
   const uint32_t n_quads = (n_dense + 3) / 4;
   if (n_quads == 0) return;
   const SeqColumnView& col = *m.col_dev;
   // work-group w runs on XCD w % 8: the work-groups of one block share an XCD and its L2
   const uint32_t xcd = blockIdx.x & 7u;
   const uint32_t i = blockIdx.x >> 3;
   const uint32_t b = xcd + 8u * (i / (2 * n_quads));
   if (b >= col.n_blocks) return;
   const uint32_t r = i % (2 * n_quads);
   const uint32_t j = r >> 1;
   const uint32_t part = r & 1u;
   const uint32_t n_chunks = m.n_chunks;
   const uint32_t n_slots = col.n_slots;
   const uint32_t half = (n_chunks + 1) / 2;
   const uint32_t c_end = part ? n_chunks : half;
   uint32_t* out = part ? m.partial_b : m.partial;
   const uint32_t tid = threadIdx.x;
   const uint32_t lane = tid & 63;
   const uint32_t wave = tid >> 6;
   const GlobalU32* light_ptr = toGlobal<GlobalU32>(col.light_ptr);
   const GlobalU32* group_ptr = toGlobal<GlobalU32>(col.group_ptr);
   const GlobalU32* heavy_ptr = toGlobal<GlobalU32>(col.heavy_ptr);
   const uint32_t s0 = col.block_slot_start[b];
   const uint32_t ns = col.block_slot_start[b + 1] - s0;
   // entries past the end take the last entry's query id (their loads stay unconditional, which
   // keeps them from being serialised) and are masked out
   uint32_t q[4];
   bool valid[4];
#pragma unroll
   for (uint32_t k = 0; k < 4; ++k) {
      valid[k] = 4 * j + k < n_dense;
      q[k] = m.lists[valid[k] ? 4 * j + k : n_dense - 1];
   }
   for (uint32_t t = tid; t <= kMaxBlockSlots; t += kSharedThreads) cnt[t] = 0;
   uint32_t since_flush = 0;
   bool first_flush = true;
   auto flush = [&]() {
      __syncthreads();
      for (uint32_t t = tid; t < ns; t += kSharedThreads) {
         const unsigned long long v = cnt[t];
         cnt[t] = 0;
#pragma unroll
         for (uint32_t k = 0; k < 4; ++k) {
            if (!valid[k]) continue;
            uint32_t* o = out + static_cast<size_t>(4 * j + k) * n_slots + s0 + t;
            const uint32_t add = static_cast<uint32_t>((v >> (16 * k)) & 0xFFFFu);
            *o = first_flush ? add : *o + add;
         }
      }
      first_flush = false;
      since_flush = 0;
      __syncthreads();
   };
   // which chunks of the part have work: one of the four filters has rows there and the tile is not
   // empty (one thread per chunk, all loads issued together)
   const uint32_t c_begin = part ? half : 0;
   for (uint32_t t = tid; c_begin + t < c_end; t += kSharedThreads) {
      const uint32_t c = c_begin + t;
      uint32_t cards[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) cards[k] = m.chunk_cards[static_cast<size_t>(q[k]) * n_chunks + c];
      const uint32_t tile = b * n_chunks + c;
      const bool nonempty = light_ptr[tile] != light_ptr[tile + 1] || group_ptr[tile] != group_ptr[tile + 1] ||
                            heavy_ptr[tile] != heavy_ptr[tile + 1];
      bool any = false;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) any |= valid[k] & (cards[k] != 0);
      work[t] = any && nonempty ? 1 : 0;
   }
   __syncthreads();
   auto next = [&](uint32_t c) -> uint32_t {
      while (c < c_end && !work[c - c_begin]) ++c;
      return c;
   };
   // filter words of chunk c: this thread's words tid and tid + 512 of the four entries
   uint64_t pf[2][4];
   auto fetch = [&](uint32_t c) {
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h) {
#pragma unroll
         for (uint32_t k = 0; k < 4; ++k) pf[h][k] = m.filters[(static_cast<size_t>(q[k]) * n_chunks + c) * kChunkWords + tid + h * 512];
#pragma unroll
         for (uint32_t k = 0; k < 4; ++k) pf[h][k] = valid[k] ? pf[h][k] : 0ull;
      }
   };
   uint32_t c = next(c_begin);
   if (c < c_end) fetch(c);
   while (c < c_end) {
      if (since_flush == kSharedFlushChunks) flush();
      ++since_flush;
      __syncthreads();  // the previous tile's readers of fq are done
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h) {
         const uint32_t w = tid + h * 512;
         u32x4 lo, hi;
#pragma unroll
         for (uint32_t k = 0; k < 4; ++k) {
            lo[k] = static_cast<uint32_t>(pf[h][k]);
            hi[k] = static_cast<uint32_t>(pf[h][k] >> 32);
         }
         reinterpret_cast<u32x4*>(fq)[2 * w] = lo;
         reinterpret_cast<u32x4*>(fq)[2 * w + 1] = hi;
      }
      const uint32_t tile = b * n_chunks + c;
      const uint32_t l0 = light_ptr[tile], l1 = light_ptr[tile + 1];
      const uint32_t h0 = heavy_ptr[tile], h1 = heavy_ptr[tile + 1];
      const uint32_t g0 = group_ptr[tile], g1 = group_ptr[tile + 1];
      const uint32_t cn = next(c + 1);
      if (cn < c_end) fetch(cn);
      __syncthreads();
      // light entries
      {
         const GlobalU32x4* e4 = toGlobal<GlobalU32x4>(col.light + l0);
         const uint32_t n4 = (l1 - l0) >> 2;
         constexpr uint32_t kStride = kSharedThreads * kRound;
         const uint32_t rounds = (n4 + kStride - 1) / kStride;
         uint32_t cur = kPadSlot, a01 = 0, a23 = 0;
         u32x4 bufA[kRound], bufB[kRound];
         auto load = [&](uint32_t round, u32x4* buf) {
            const uint32_t base = round * kStride + tid * kRound;
#pragma unroll
            for (uint32_t u = 0; u < kRound; ++u) {
               const uint32_t jj = base + u;
               buf[u] = e4[jj < n4 ? jj : n4 - 1];
            }
         };
         auto use = [&](uint32_t round, const u32x4* buf) {
            const uint32_t base = round * kStride + tid * kRound;
#pragma unroll
            for (uint32_t u = 0; u < kRound; ++u) {
               u32x4 word[4];
#pragma unroll
               for (uint32_t k = 0; k < 4; ++k) word[k] = fq4[(buf[u][k] & 0xFFFFu) >> 5];
               const bool ok = base + u < n4;
#pragma unroll
               for (uint32_t k = 0; k < 4; ++k) {
                  const uint32_t x = buf[u][k];
                  const uint32_t slot = x >> 16;
                  if (slot != cur) {
                     if (a01 | a23) {
                        atomicAdd(&cnt[cur], static_cast<unsigned long long>(a01) | (static_cast<unsigned long long>(a23) << 32));
                     }
                     cur = slot;
                     a01 = 0;
                     a23 = 0;
                  }
                  if (ok) quadBits(word[k], x, a01, a23);
               }
            }
         };
         if (rounds > 0) {
            load(0, bufA);
            for (uint32_t round = 0; round < rounds; round += 2) {
               if (round + 1 < rounds) load(round + 1, bufB);
               use(round, bufA);
               if (round + 1 < rounds) {
                  if (round + 2 < rounds) load(round + 2, bufA);
                  use(round + 1, bufB);
               }
            }
         }
         // slot 4095 is the padding slot: its counter is never read
         if (a01 | a23) atomicAdd(&cnt[cur], static_cast<unsigned long long>(a01) | (static_cast<unsigned long long>(a23) << 32));
      }
      // groups of eight rows of one slot
      {
         const GlobalU32x4* gr = toGlobal<GlobalU32x4>(col.groups) + g0;
         const GlobalU16* gh = toGlobal<GlobalU16>(col.group_hdr) + g0;
         const uint32_t ng = g1 - g0;
         constexpr uint32_t kStride = kSharedThreads * kRound;
         const uint32_t rounds = (ng + kStride - 1) / kStride;
         uint32_t cur = kPadSlot, a01 = 0, a23 = 0;
         u32x4 rowsA[kRound], rowsB[kRound];
         uint32_t hdrA[kRound], hdrB[kRound];
         auto load = [&](uint32_t round, u32x4* rows, uint32_t* hdr) {
            const uint32_t base = round * kStride + tid * kRound;
#pragma unroll
            for (uint32_t u = 0; u < kRound; ++u) {
               const uint32_t jj = base + u < ng ? base + u : ng - 1;
               rows[u] = gr[jj];
               hdr[u] = gh[jj];
            }
         };
         auto use = [&](uint32_t round, const u32x4* rows, const uint32_t* hdr) {
            const uint32_t base = round * kStride + tid * kRound;
#pragma unroll
            for (uint32_t u = 0; u < kRound; ++u) {
               u32x4 word[8];
#pragma unroll
               for (uint32_t t = 0; t < 8; ++t) {
                  const uint32_t row = (t & 1u) ? rows[u][t >> 1] >> 16 : rows[u][t >> 1] & 0xFFFFu;
                  word[t] = fq4[row >> 5];
               }
               const uint32_t h = base + u < ng ? hdr[u] : 0xFFFFu;
               const uint32_t slot = h & 0xFFFu;
               const uint32_t mm = h == 0xFFFFu ? 0u : (h >> 12) + 1;
               uint32_t n01 = 0, n23 = 0;
#pragma unroll
               for (uint32_t t = 0; t < 8; ++t) {
                  const uint32_t row = (t & 1u) ? rows[u][t >> 1] >> 16 : rows[u][t >> 1];
                  if (t < mm) quadBits(word[t], row, n01, n23);
               }
               if (slot != cur) {
                  if (a01 | a23) {
                     atomicAdd(&cnt[cur], static_cast<unsigned long long>(a01) | (static_cast<unsigned long long>(a23) << 32));
                  }
                  cur = slot;
                  a01 = 0;
                  a23 = 0;
               }
               a01 += n01;
               a23 += n23;
            }
         };
         if (rounds > 0) {
            load(0, rowsA, hdrA);
            for (uint32_t round = 0; round < rounds; round += 2) {
               if (round + 1 < rounds) load(round + 1, rowsB, hdrB);
               use(round, rowsA, hdrA);
               if (round + 1 < rounds) {
                  if (round + 2 < rounds) load(round + 2, rowsA, hdrA);
                  use(round + 1, rowsB, hdrB);
               }
            }
         }
         if (a01 | a23) atomicAdd(&cnt[cur], static_cast<unsigned long long>(a01) | (static_cast<unsigned long long>(a23) << 32));
      }
      // bitset and run containers: two wavefronts per container, each a half.  A bitset half's
      // 1,024 u32 words are read lane-consecutive (16 loads of 256 bytes per wavefront, all in
      // flight at once), so the matching filter words come out of LDS without bank conflicts; sums
      // in 32-bit fields (a container and a filter may share all 65,536 rows).
      for (uint32_t t = 2 * h0 + wave; t < 2 * h1; t += kSharedThreads / 64) {
         const uint32_t d = t >> 1, half = t & 1u;
         const u32x2 dw = toGlobal<GlobalU32x2>(col.vert.desc)[d];
         const uint32_t type = descType(dw[0]);
         const uint32_t slot = descSlot(dw[0]);
         uint32_t n[4] = {0, 0, 0, 0};
         if (type == CT_BITSET) {
            constexpr uint32_t kHalfWords = kChunkWords / 64;  // u32 words of a half per lane
            const GlobalU32* b32 = toGlobal<GlobalU32>(reinterpret_cast<const uint32_t*>(
                                      col.vert.bits + static_cast<size_t>(dw[1]) * kChunkWords)) +
                                   half * kChunkWords;
            const u32x4* f4 = fq4 + half * kChunkWords;
            uint32_t x[kHalfWords];
#pragma unroll
            for (uint32_t u = 0; u < kHalfWords; ++u) x[u] = b32[64 * u + lane];
#pragma unroll
            for (uint32_t u = 0; u < kHalfWords; ++u) {
               const u32x4 f = f4[64 * u + lane];
#pragma unroll
               for (uint32_t q4 = 0; q4 < 4; ++q4) n[q4] += __popc(x[u] & f[q4]);
            }
         } else if (type == CT_RUN && half == 0) {
            const uint32_t card = descCard(dw[0]);
            const GlobalU32* rr = toGlobal<GlobalU32>(col.vert.runs + dw[1]);
            for (uint32_t r = lane; r < card; r += 64) {
               const uint32_t st = rr[r] & 0xFFFFu;
               const uint32_t en = st + (rr[r] >> 16);
#pragma unroll
               for (uint32_t q4 = 0; q4 < 4; ++q4) n[q4] += rangePopcountQ(fq, q4, st, en);
            }
         }
         // per lane at most 1,024 a query: two queries per 64-bit sum in 32-bit fields
         const uint64_t s01 = waveSum64(static_cast<uint64_t>(n[0]) | (static_cast<uint64_t>(n[1]) << 32));
         const uint64_t s23 = waveSum64(static_cast<uint64_t>(n[2]) | (static_cast<uint64_t>(n[3]) << 32));
         if (lane == 0) {
            const uint32_t v[4] = {static_cast<uint32_t>(s01), static_cast<uint32_t>(s01 >> 32),
                                   static_cast<uint32_t>(s23), static_cast<uint32_t>(s23 >> 32)};
#pragma unroll
            for (uint32_t q4 = 0; q4 < 4; ++q4) {
               if (valid[q4] && v[q4]) atomicAdd(&m.heavy_sums[static_cast<size_t>(4 * j + q4) * n_slots + s0 + slot], v[q4]);
            }
         }
      }
      c = cn;
   }
   flush();
}


// ---- octet shared scan: eight dense queries per pass over the index -----------------------------
// A pre-pass turns the filters of dense list entries 8o .. 8o+7 into one byte per row (bit k: the row
// is in entry 8o+k's filter), 64 KiB per chunk.  The counting kernel is persistent: each work-group
// (1024 threads, one per CU) takes (position block, chunk part, octet) items from a counter.  Per
// chunk it stages the chunk's row bytes in LDS; a light entry or group row then costs one LDS byte
// read, spread into eight 8-bit fields by two 24-bit multiplies (spread8) and added to the lane's
// run sum with one 64-bit add, for all eight queries at once.  Runs end in two 64-bit LDS
// atomics on counters of four 16-bit fields (flushed to the part's slot sums every 15 chunks, so
// they cannot overflow).  Bitset and run containers AND against the eight filters' words in global
// memory and add to heavy_sums.  The finalize step adds the parts' sums as it reads them.
// While a tile is counted, the next tile's row bytes and the first round of its light entries and
// groups are loaded into registers; the tile pointers of the item's chunks are read into LDS when
// the item starts.
constexpr uint32_t kOctetMaskThreads = 256;
constexpr uint32_t kOctetDescCap = 128;    // containers whose sums are gathered in LDS at a time

__device__ inline uint64_t transpose8x8(uint64_t x) {
   // bit 8r + c <-> bit 8c + r
   uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
   x = x ^ t ^ (t << 7);
   t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
   x = x ^ t ^ (t << 14);
   t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
   x = x ^ t ^ (t << 28);
   return x;
}

// byte k of the result = bit k of b (b < 256): each nibble times 1 + 2^7 + 2^14 + 2^21 puts its bit k
// at bit 8k with no overlapping partial products (so no carries), then the mask keeps those bits
__device__ inline uint64_t spread8(uint32_t b) {
   const uint32_t lo = ((b & 0xFu) * 0x00204081u) & 0x01010101u;
   const uint32_t hi = (((b >> 4) & 0xFu) * 0x00204081u) & 0x01010101u;
   return (static_cast<uint64_t>(hi) << 32) | lo;
}

// four 8-bit fields -> four 16-bit fields
__device__ inline uint64_t widen4(uint32_t a) {
   return static_cast<uint64_t>((a & 0xFFu) | ((a & 0xFF00u) << 8)) |
          (static_cast<uint64_t>(((a >> 16) & 0xFFu) | ((a >> 8) & 0xFF0000u)) << 32);
}

// Grid (kChunkWords / kOctetMaskThreads, max octets * n_chunks): thread = (octet, chunk, word).
__global__ __launch_bounds__(kOctetMaskThreads) void octetMasksKernel(MutationsBatch m) {
   const uint32_t n_dense = m.list_counts[0];
   const uint32_t n_chunks = m.n_chunks;
   const uint32_t o = blockIdx.y / n_chunks;
   const uint32_t c = blockIdx.y % n_chunks;
   if (8 * o >= n_dense) return;
   const uint32_t w = blockIdx.x * kOctetMaskThreads + threadIdx.x;
   const GlobalU64* filters = toGlobal<GlobalU64>(m.filters);
   // loads without conditions (a condition would serialise them): entries past the end read the
   // last entry's words and are masked out
   uint32_t qs[8];
#pragma unroll
   for (uint32_t k = 0; k < 8; ++k) qs[k] = m.lists[8 * o + k < n_dense ? 8 * o + k : n_dense - 1];
   uint64_t x[8];
#pragma unroll
   for (uint32_t k = 0; k < 8; ++k) x[k] = filters[(static_cast<size_t>(qs[k]) * n_chunks + c) * kChunkWords + w];
#pragma unroll
   for (uint32_t k = 0; k < 8; ++k) x[k] = 8 * o + k < n_dense ? x[k] : 0ull;
   uint64_t y[8];
#pragma unroll
   for (uint32_t j = 0; j < 8; ++j) {
      uint64_t v = 0;
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k) v |= ((x[k] >> (8 * j)) & 0xFFull) << (8 * k);
      y[j] = transpose8x8(v);  // byte i: row 8j + i, bit k: entry 8o + k
   }
   u64x2* dst = reinterpret_cast<u64x2*>(m.octet_masks + ((static_cast<size_t>(o) * n_chunks + c) << 16) + 64 * w);
#pragma unroll
   for (uint32_t h = 0; h < 4; ++h) dst[h] = u64x2{y[2 * h], y[2 * h + 1]};
}

template <uint32_t kRound, uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void mutationsCountOctetKernel(MutationsBatch m) {
   static_assert(kChunkWords % kThreads == 0 && kThreads % 64 == 0, "whole bitset words per thread");
   constexpr uint32_t kWords = kChunkWords / kThreads;     // bitset words per thread
   constexpr uint32_t kRowVecs = kChunkRows / 16 / kThreads;  // 16-byte pieces of the row bytes per thread
   __shared__ __align__(16) uint8_t mk[kChunkRows];
   __shared__ unsigned long long cnt[2][kMaxBlockSlots + 1];
   __shared__ uint32_t hacc[kOctetDescCap][8];
   __shared__ uint8_t work[kOctetMaxPartChunks];
   // tile pointers of the item's chunks: entry t is chunk c_lo + t's start, entry t + 1 its end
   __shared__ uint32_t tl[kOctetMaxPartChunks + 1], tg[kOctetMaxPartChunks + 1], th[kOctetMaxPartChunks + 1];
   __shared__ uint32_t item_s;
   const uint32_t n_dense = m.list_counts[0];
   const uint32_t n_oct = (n_dense + 7) / 8;
   if (n_oct == 0) return;
   const SeqColumnView& col = *m.col_dev;
   const uint32_t n_chunks = m.n_chunks;
   const uint32_t n_slots = col.n_slots;
   const uint32_t n_items = col.n_blocks * kOctetParts * n_oct;
   const uint32_t per_part = (n_chunks + kOctetParts - 1) / kOctetParts;
   const uint32_t tid = threadIdx.x;
   const uint32_t lane = tid & 63;
   const GlobalU32* light_ptr = toGlobal<GlobalU32>(col.light_ptr);
   const GlobalU32* group_ptr = toGlobal<GlobalU32>(col.group_ptr);
   const GlobalU32* heavy_ptr = toGlobal<GlobalU32>(col.heavy_ptr);
   const GlobalU64* filters = toGlobal<GlobalU64>(m.filters);
   for (;;) {
      __syncthreads();  // the previous item is finished
      if (tid == 0) item_s = atomicAdd(&m.list_counts[2], 1u);
      __syncthreads();
      const uint32_t item = item_s;
      if (item >= n_items) break;
      const uint32_t o = item % n_oct;
      const uint32_t part = (item / n_oct) % kOctetParts;
      const uint32_t b = item / (n_oct * kOctetParts);
      const uint32_t c_lo = part * per_part;
      const uint32_t c_hi = c_lo + per_part < n_chunks ? c_lo + per_part : n_chunks;
      uint32_t* out = part == 0   ? m.partial
                      : part == 1 ? m.partial_b
                                  : m.partial_c + static_cast<size_t>(part - 2) * m.max_queries * n_slots;
      const uint32_t s0 = col.block_slot_start[b];
      const uint32_t ns = col.block_slot_start[b + 1] - s0;
      // entries past the end take the last entry's query id (their loads stay unconditional) and
      // are masked out
      uint32_t q[8];
      bool valid[8];
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k) {
         valid[k] = 8 * o + k < n_dense;
         q[k] = m.lists[valid[k] ? 8 * o + k : n_dense - 1];
      }
      for (uint32_t t = tid; t <= kMaxBlockSlots; t += kThreads) {
         cnt[0][t] = 0;
         cnt[1][t] = 0;
      }
      for (uint32_t i = tid; i < kOctetDescCap * 8; i += kThreads) hacc[i >> 3][i & 7] = 0;
      // which chunks of the part have work: one of the eight filters has rows there and the tile
      // is not empty (one thread per chunk, all loads issued together)
      for (uint32_t t = tid; c_lo + t < c_hi; t += kThreads) {
         const uint32_t c = c_lo + t;
         uint32_t cards[8];
#pragma unroll
         for (uint32_t k = 0; k < 8; ++k) cards[k] = m.chunk_cards[static_cast<size_t>(q[k]) * n_chunks + c];
         const uint32_t tile = b * n_chunks + c;
         const uint32_t la = light_ptr[tile], lb = light_ptr[tile + 1];
         const uint32_t ga = group_ptr[tile], gb = group_ptr[tile + 1];
         const uint32_t ha = heavy_ptr[tile], hb = heavy_ptr[tile + 1];
         // entry t + 1 is also written by the thread of chunk c + 1, with the same value
         tl[t] = la;
         tl[t + 1] = lb;
         tg[t] = ga;
         tg[t + 1] = gb;
         th[t] = ha;
         th[t + 1] = hb;
         const bool nonempty = la != lb || ga != gb || ha != hb;
         bool any = false;
#pragma unroll
         for (uint32_t k = 0; k < 8; ++k) any |= valid[k] & (cards[k] != 0);
         work[t] = any && nonempty ? 1 : 0;
      }
      __syncthreads();
      uint32_t since_flush = 0;
      bool first_flush = true;
      auto flush = [&]() {
         __syncthreads();
         for (uint32_t t = tid; t < ns; t += kThreads) {
            const unsigned long long v[2] = {cnt[0][t], cnt[1][t]};
            cnt[0][t] = 0;
            cnt[1][t] = 0;
#pragma unroll
            for (uint32_t k = 0; k < 8; ++k) {
               if (!valid[k]) continue;
               uint32_t* ot = out + static_cast<size_t>(8 * o + k) * n_slots + s0 + t;
               const uint32_t add = static_cast<uint32_t>((v[k >> 2] >> (16 * (k & 3))) & 0xFFFFu);
               *ot = first_flush ? add : *ot + add;
            }
         }
         first_flush = false;
         since_flush = 0;
         __syncthreads();
      };
      auto addRun = [&](uint32_t slot, uint64_t lo, uint64_t hi) {
         if (lo) atomicAdd(&cnt[0][slot], static_cast<unsigned long long>(lo));
         if (hi) atomicAdd(&cnt[1][slot], static_cast<unsigned long long>(hi));
      };
      auto next = [&](uint32_t c) -> uint32_t {
         while (c < c_hi && !work[c - c_lo]) ++c;
         return c;
      };
      // the next tile: this thread's share of the chunk's row bytes and of the first round of the
      // light entries and of the groups (left unloaded where the tile has none)
      u32x4 pf[kRowVecs];
      u32x4 pl[kRound], pg[kRound];
      uint32_t ph[kRound];
      auto fetchRows = [&](uint32_t c) {
         const GlobalU32x4* src = toGlobal<GlobalU32x4>(m.octet_masks + ((static_cast<size_t>(o) * n_chunks + c) << 16));
#pragma unroll
         for (uint32_t h = 0; h < kRowVecs; ++h) pf[h] = src[kRowVecs * tid + h];
      };
      auto fetchLight = [&](uint32_t c) {
         const uint32_t t = c - c_lo;
         const uint32_t la = tl[t], n4 = (tl[t + 1] - la) >> 2;
         if (n4 > 0) {
            const GlobalU32x4* e4 = toGlobal<GlobalU32x4>(col.light + la);
#pragma unroll
            for (uint32_t u = 0; u < kRound; ++u) {
               const uint32_t jj = tid * kRound + u;
               pl[u] = e4[jj < n4 ? jj : n4 - 1];
            }
         }
      };
      auto fetchGroups = [&](uint32_t c) {
         const uint32_t t = c - c_lo;
         const uint32_t ga = tg[t], ng = tg[t + 1] - ga;
         if (ng > 0) {
            const GlobalU32x4* gr = toGlobal<GlobalU32x4>(col.groups) + ga;
            const GlobalU16* gh = toGlobal<GlobalU16>(col.group_hdr) + ga;
#pragma unroll
            for (uint32_t u = 0; u < kRound; ++u) {
               const uint32_t jj = tid * kRound + u < ng ? tid * kRound + u : ng - 1;
               pg[u] = gr[jj];
               ph[u] = gh[jj];
            }
         }
      };
      uint32_t c = next(c_lo);
      if (c < c_hi) {
         fetchRows(c);
         fetchLight(c);
         fetchGroups(c);
      }
      // the next tile's loads are issued as soon as the registers they go to are free: row bytes
      // once this tile's are in LDS, light entries and groups once this tile's first rounds are used
      while (c < c_hi) {
         if (since_flush == kSharedFlushChunks) flush();
         ++since_flush;
         __syncthreads();  // the previous tile's readers of mk are done
#pragma unroll
         for (uint32_t h = 0; h < kRowVecs; ++h) reinterpret_cast<u32x4*>(mk)[kRowVecs * tid + h] = pf[h];
         const uint32_t ct = c - c_lo;
         const uint32_t l0 = tl[ct], l1 = tl[ct + 1];
         const uint32_t h0 = th[ct], h1 = th[ct + 1];
         const uint32_t g0 = tg[ct], g1 = tg[ct + 1];
         const uint32_t cn = next(c + 1);
         if (cn < c_hi) fetchRows(cn);
         __syncthreads();
         // light entries: runs of a slot summed in eight 8-bit fields (a light slot has at most 32
         // rows in a chunk)
         {
            const GlobalU32x4* e4 = toGlobal<GlobalU32x4>(col.light + l0);
            const uint32_t n4 = (l1 - l0) >> 2;
            constexpr uint32_t kStride = kThreads * kRound;
            const uint32_t rounds = (n4 + kStride - 1) / kStride;
            uint32_t cur = kPadSlot;
            uint64_t acc = 0;
            u32x4 bufA[kRound], bufB[kRound];
            auto load = [&](uint32_t round, u32x4* buf) {
               const uint32_t base = round * kStride + tid * kRound;
#pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
                  const uint32_t jj = base + u;
                  buf[u] = e4[jj < n4 ? jj : n4 - 1];
               }
            };
            auto use = [&](uint32_t round, const u32x4* buf) {
               const uint32_t base = round * kStride + tid * kRound;
               uint32_t mb[4 * kRound];
#pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
#pragma unroll
                  for (uint32_t k = 0; k < 4; ++k) mb[4 * u + k] = mk[buf[u][k] & 0xFFFFu];
               }
               uint64_t sp[4 * kRound];
#pragma unroll
               for (uint32_t i = 0; i < 4 * kRound; ++i) sp[i] = spread8(mb[i]);
#pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
                  const bool ok = base + u < n4;
#pragma unroll
                  for (uint32_t k = 0; k < 4; ++k) {
                     const uint32_t slot = buf[u][k] >> 16;
                     if (slot != cur) {
                        if (acc) addRun(cur, widen4(static_cast<uint32_t>(acc)), widen4(static_cast<uint32_t>(acc >> 32)));
                        cur = slot;
                        acc = 0;
                     }
                     if (ok) acc += sp[4 * u + k];
                  }
               }
            };
            if (rounds > 0) {
               // round 0 is in pl; later rounds alternate between bufB (odd) and bufA (even)
               if (rounds > 1) load(1, bufB);
               use(0, pl);
               for (uint32_t round = 1; round < rounds; round += 2) {
                  if (round + 1 < rounds) load(round + 1, bufA);
                  use(round, bufB);
                  if (round + 1 < rounds) {
                     if (round + 2 < rounds) load(round + 2, bufB);
                     use(round + 1, bufA);
                  }
               }
            }
            // slot 4095 is the padding slot: its counters are never read
            if (acc) addRun(cur, widen4(static_cast<uint32_t>(acc)), widen4(static_cast<uint32_t>(acc >> 32)));
         }
         if (cn < c_hi) fetchLight(cn);
         // groups of eight rows of one slot: a group's eight rows summed in 8-bit fields, runs in
         // 16-bit fields (at most 4096 rows of a slot in a chunk)
         {
            const GlobalU32x4* gr = toGlobal<GlobalU32x4>(col.groups) + g0;
            const GlobalU16* gh = toGlobal<GlobalU16>(col.group_hdr) + g0;
            const uint32_t ng = g1 - g0;
            constexpr uint32_t kStride = kThreads * kRound;
            const uint32_t rounds = (ng + kStride - 1) / kStride;
            uint32_t cur = kPadSlot;
            uint64_t lo = 0, hi = 0;
            u32x4 rowsA[kRound], rowsB[kRound];
            uint32_t hdrA[kRound], hdrB[kRound];
            auto load = [&](uint32_t round, u32x4* rows, uint32_t* hdr) {
               const uint32_t base = round * kStride + tid * kRound;
#pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
                  const uint32_t jj = base + u < ng ? base + u : ng - 1;
                  rows[u] = gr[jj];
                  hdr[u] = gh[jj];
               }
            };
            auto use = [&](uint32_t round, const u32x4* rows, const uint32_t* hdr) {
               const uint32_t base = round * kStride + tid * kRound;
#pragma unroll
               for (uint32_t u = 0; u < kRound; ++u) {
                  uint32_t mb[8];
#pragma unroll
                  for (uint32_t t = 0; t < 8; ++t) {
                     const uint32_t row = (t & 1u) ? rows[u][t >> 1] >> 16 : rows[u][t >> 1] & 0xFFFFu;
                     mb[t] = mk[row];
                  }
                  const uint32_t h = base + u < ng ? hdr[u] : 0xFFFFu;
                  const uint32_t slot = h & 0xFFFu;
                  const uint32_t mm = h == 0xFFFFu ? 0u : (h >> 12) + 1;
                  uint64_t g8 = 0;
#pragma unroll
                  for (uint32_t t = 0; t < 8; ++t) {
                     if (t < mm) g8 += spread8(mb[t]);
                  }
                  if (slot != cur) {
                     addRun(cur, lo, hi);
                     cur = slot;
                     lo = 0;
                     hi = 0;
                  }
                  lo += widen4(static_cast<uint32_t>(g8));
                  hi += widen4(static_cast<uint32_t>(g8 >> 32));
               }
            };
            if (rounds > 0) {
               // round 0 is in pg / ph; later rounds alternate between the B (odd) and A (even) buffers
               if (rounds > 1) load(1, rowsB, hdrB);
               use(0, pg, ph);
               for (uint32_t round = 1; round < rounds; round += 2) {
                  if (round + 1 < rounds) load(round + 1, rowsA, hdrA);
                  use(round, rowsB, hdrB);
                  if (round + 1 < rounds) {
                     if (round + 2 < rounds) load(round + 2, rowsB, hdrB);
                     use(round + 1, rowsA, hdrA);
                  }
               }
            }
            addRun(cur, lo, hi);
         }
         if (cn < c_hi) fetchGroups(cn);
         // containers (a tile holds only run and bitset containers: smaller ones are light entries
         // or groups).  Bitsets: each thread takes kWords words of every bitset of the tile, against
         // the eight filters' words (loaded once per tile); a wavefront's eight sums go through two
         // 64-bit reductions of 16-bit fields (at most 64 lanes x kWords x 64 rows < 2^16) into
         // 32-bit LDS sums per container.  Runs: the container's runs spread over all threads, 32-bit
         // sums (a run container can hold all 65,536 rows of a chunk).  Every container's eight sums
         // are then added to heavy_sums with global atomics.
         // hacc is zero here: cleared when the item starts and by each flush below, and the barriers
         // at a tile's start order the flush's clearing before the next tile's sums
         static_assert(kOctetDescCap == 128, "two descriptors per lane");
         for (uint32_t d0 = h0; d0 < h1; d0 += kOctetDescCap) {
            const uint32_t d1 = h1 - d0 < kOctetDescCap ? h1 : d0 + kOctetDescCap;
            const GlobalU32x2* descs = toGlobal<GlobalU32x2>(col.vert.desc);
            // the group's descriptors, two per lane, loaded together with the filter words
            const u32x2 da = descs[d0 + lane < d1 ? d0 + lane : d1 - 1];
            const u32x2 db = descs[d0 + 64 + lane < d1 ? d0 + 64 + lane : d1 - 1];
            uint64_t fwd[kWords][8];
#pragma unroll
            for (uint32_t w = 0; w < kWords; ++w) {
#pragma unroll
               for (uint32_t k = 0; k < 8; ++k) {
                  fwd[w][k] = filters[(static_cast<size_t>(q[k]) * n_chunks + c) * kChunkWords + w * kThreads + tid];
               }
            }
#pragma unroll
            for (uint32_t w = 0; w < kWords; ++w) {
#pragma unroll
               for (uint32_t k = 0; k < 8; ++k) fwd[w][k] = valid[k] ? fwd[w][k] : 0ull;
            }
            for (uint32_t d = d0; d < d1; ++d) {
               const uint32_t j = d - d0;
               const u32x2 dsel = j < 64 ? da : db;
               const u32x2 dw = {uniform(__shfl(dsel[0], static_cast<int>(j & 63), 64)),
                                 uniform(__shfl(dsel[1], static_cast<int>(j & 63), 64))};
               const uint32_t type = descType(dw[0]);
               if (type == CT_BITSET) {
                  uint64_t p01 = 0, p23 = 0;  // four 16-bit fields each: queries 0-3 and 4-7
                  const GlobalU64* bits = toGlobal<GlobalU64>(col.vert.bits) + static_cast<size_t>(dw[1]) * kChunkWords;
                  uint64_t x[kWords];
#pragma unroll
                  for (uint32_t w = 0; w < kWords; ++w) x[w] = bits[w * kThreads + tid];
#pragma unroll
                  for (uint32_t w = 0; w < kWords; ++w) {
#pragma unroll
                     for (uint32_t k = 0; k < 4; ++k) {
                        p01 += static_cast<uint64_t>(__popcll(x[w] & fwd[w][k])) << (16 * k);
                        p23 += static_cast<uint64_t>(__popcll(x[w] & fwd[w][k + 4])) << (16 * k);
                     }
                  }
                  const uint64_t s01 = waveSum64(p01);
                  const uint64_t s23 = waveSum64(p23);
                  if (lane == 0 && (s01 | s23)) {
#pragma unroll
                     for (uint32_t k = 0; k < 4; ++k) {
                        const uint32_t v0 = static_cast<uint32_t>((s01 >> (16 * k)) & 0xFFFFu);
                        const uint32_t v1 = static_cast<uint32_t>((s23 >> (16 * k)) & 0xFFFFu);
                        if (v0) atomicAdd(&hacc[d - d0][k], v0);
                        if (v1) atomicAdd(&hacc[d - d0][k + 4], v1);
                     }
                  }
               } else if (type == CT_RUN) {
                  // a run container can hold a thousand runs: they are spread over all threads, with
                  // 32-bit sums per query reduced per wavefront
                  const uint32_t card = descCard(dw[0]);
                  const GlobalU32* rr = toGlobal<GlobalU32>(col.vert.runs + dw[1]);
                  uint32_t n[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                  for (uint32_t t = tid; t < card; t += kThreads) {
                     const uint32_t x = rr[t];
                     const uint32_t st = x & 0xFFFFu;
                     const uint32_t en = st + (x >> 16);  // inclusive, inside the chunk
                     for (uint32_t w = st >> 6; w <= (en >> 6); ++w) {
                        const uint64_t rm = rangeMask(w == (st >> 6) ? (st & 63) : 0, w == (en >> 6) ? (en & 63) : 63);
#pragma unroll
                        for (uint32_t k = 0; k < 8; ++k) {
                           n[k] += static_cast<uint32_t>(__popcll(filters[(static_cast<size_t>(q[k]) * n_chunks + c) * kChunkWords + w] & rm));
                        }
                     }
                  }
#pragma unroll
                  for (uint32_t k = 0; k < 8; ++k) {
                     const uint32_t v = waveSum(n[k]);
                     if (lane == 0 && valid[k] && v) atomicAdd(&hacc[d - d0][k], v);
                  }
               }
            }
            __syncthreads();
            for (uint32_t i = tid; i < (d1 - d0) * 8; i += kThreads) {
               const uint32_t k = i & 7;
               const uint32_t v = hacc[i >> 3][k];
               hacc[i >> 3][k] = 0;
               if (valid[k] && v) {
                  const uint32_t slot = descSlot(descs[d0 + (i >> 3)][0]);
                  atomicAdd(&m.heavy_sums[static_cast<size_t>(8 * o + k) * n_slots + s0 + slot], v);
               }
            }
            if (d1 < h1) __syncthreads();  // the next group's sums go to the cleared hacc
         }
         c = cn;
      }
      flush();
   }
}

// Work-groups of eight wavefronts: one LDS table sums a key over the rows of 512 filter words
// before its single global atomic, and a CU holds 32 wavefronts of it.
constexpr uint32_t kSparseBatchThreads = 512;
__global__ __launch_bounds__(kSparseBatchThreads) void sparseCountBatchKernel(MutationsBatch m) {
   __shared__ uint32_t keys[kSparseTable];
   __shared__ int32_t vals[kSparseTable];
   const uint32_t e = blockIdx.y;
   if (e >= m.list_counts[1]) return;
   const uint32_t q = m.lists[m.max_queries + e];
   const SeqColumnView& col = *m.col_dev;
   const size_t L = col.genome_length;
   const SparseTable tb{keys, vals, m.sparse_counts + static_cast<size_t>(e) * col.n_symbols * L,
                        m.sparse_bnd + static_cast<size_t>(e) * 3 * L};
   tb.clear();
   __syncthreads();
   sparseCountRows(col, m.chunk_sizes, m.filters + static_cast<size_t>(q) * m.n_chunks * kChunkWords,
                   m.n_chunks * kChunkWords, tb, blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6),
                   gridDim.x * (blockDim.x / 64));
   __syncthreads();
   tb.flush();
}

// Dense list entry e's slot sums: the parts of the shared scans (`dense_parts`: 1 for the one-pass-
// per-query kernel, 2 for the four-query scan, 4 for the eight-query scan) plus the heavy sums.
__device__ inline SlotSource denseSource(const MutationsBatch& m, uint32_t e, uint32_t dense_parts) {
   const size_t n_slots = m.col_dev->n_slots;
   SlotSource src{m.col_dev, m.partial + e * n_slots};
   if (dense_parts > 1) {
      src.partial_b = m.partial_b + e * n_slots;
      src.heavy = m.heavy_sums + e * n_slots;
   }
   if (dense_parts > 2) {
      src.partial_c = m.partial_c + e * n_slots;
      src.plane = static_cast<size_t>(m.max_queries) * n_slots;
      src.n_more = dense_parts - 2;
   }
   return src;
}
__device__ inline SparseSource sparseSource(const MutationsBatch& m, uint32_t e) {
   const size_t L = m.col_dev->genome_length;
   return SparseSource{m.col_dev, m.sparse_counts + static_cast<size_t>(e) * m.col_dev->n_symbols * L,
                       m.sparse_bnd + static_cast<size_t>(e) * 3 * L};
}

// Both lists in one launch: work-group row y below the dense count is dense entry y, the rows after
// it the sparse entries.
__global__ __launch_bounds__(kPosBlock) void mutationsTotalsBatchKernel(MutationsBatch m, uint32_t n_pblocks,
                                                                        uint32_t dense_parts) {
   const uint32_t n_dense = m.list_counts[0];
   const uint32_t y = blockIdx.y;
   if (y < n_dense) {
      totalsBody(denseSource(m, y, dense_parts), blockIdx.x, m.totals + static_cast<size_t>(y) * n_pblocks * 3);
   } else if (y - n_dense < m.list_counts[1]) {
      const uint32_t e = y - n_dense;
      totalsBody(sparseSource(m, e), blockIdx.x, m.totals + (static_cast<size_t>(m.max_queries) + e) * n_pblocks * 3);
   }
}

__global__ __launch_bounds__(kPosBlock) void mutationsFinalizeBatchKernel(MutationsBatch m, uint32_t n_pblocks,
                                                                          uint32_t dense_parts) {
   const uint32_t n_dense = m.list_counts[0];
   const uint32_t y = blockIdx.y;
   if (y < n_dense) {
      const uint32_t q = m.lists[y];
      finalizeBody(denseSource(m, y, dense_parts), blockIdx.x, m.totals + static_cast<size_t>(y) * n_pblocks * 3,
                   n_pblocks, m.cards[q], m.min_proportion[q], nullptr, m.rows + static_cast<size_t>(q) * m.row_capacity,
                   m.row_capacity, m.n_rows + q);
   } else if (y - n_dense < m.list_counts[1]) {
      const uint32_t e = y - n_dense;
      const uint32_t q = m.lists[m.max_queries + e];
      finalizeBody(sparseSource(m, e), blockIdx.x, m.totals + (static_cast<size_t>(m.max_queries) + e) * n_pblocks * 3,
                   n_pblocks, m.cards[q], m.min_proportion[q], nullptr, m.rows + static_cast<size_t>(q) * m.row_capacity,
                   m.row_capacity, m.n_rows + q);
   }
}

// ------------------------------------------------------------------------------------------------
// Group-by: per-row labels of a sequence position (symbol, missing, or null), then a histogram of
// combined labels over the filter's rows.
// ------------------------------------------------------------------------------------------------
__global__ void symbolLabelsKernel(const SeqColumnView* __restrict__ colp, const uint32_t* __restrict__ chunk_sizes,
                                   uint32_t position, uint32_t n_rows_padded, uint32_t* __restrict__ labels) {
   const SeqColumnView& col = *colp;
   const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= n_rows_padded) return;
   if ((r & 0xFFFFu) >= chunk_sizes[r >> 16]) {
      labels[r] = kInvalidId;  // not a row of the table
      return;
   }
   const bool is_null = (col.null_bits[r >> 6] >> (r & 63)) & 1ULL;
   uint32_t label;
   if (is_null) {
      label = col.n_symbols;  // null group
   } else {
      const uint32_t c = col.cov[r];
      bool covered = (c & 0xFFFFu) <= position && position < (c >> 16);
      if (covered) {
         for (uint32_t j = col.nrun_ptr[r]; j < col.nrun_ptr[r + 1]; ++j) {
            const uint32_t x = col.nruns[j];
            if ((x & 0xFFFFu) <= position && position < (x >> 16)) {
               covered = false;
               break;
            }
         }
      }
      label = covered ? col.local_ref[position] : (col.n_symbols == 16 ? 15u : 27u);
   }
   labels[r] = label;
}

// One work-group per chunk: rows stored under a symbol at `position` get that symbol.
__global__ void symbolScatterKernel(const SeqColumnView* __restrict__ colp, uint32_t n_chunks, uint32_t block,
                                    uint32_t slot_lo, uint32_t slot_hi, uint32_t* __restrict__ labels) {
   const SeqColumnView& col = *colp;
   const uint32_t c = blockIdx.x;
   const uint32_t tile = block * n_chunks + c;
   const uint32_t slot_base = col.block_slot_start[block];
   uint32_t* lab = labels + (static_cast<size_t>(c) << 16);
   for (uint32_t k = col.light_ptr[tile] + threadIdx.x; k < col.light_ptr[tile + 1]; k += blockDim.x) {
      const uint32_t x = col.light[k];
      const uint32_t slot = x >> 16;
      if (slot < slot_lo || slot >= slot_hi) continue;
      const uint32_t sym = col.slot_info[slot_base + slot] >> 16;
      if (sym < col.n_symbols) lab[x & 0xFFFFu] = sym;
   }
   for (uint32_t k = col.group_ptr[tile] + threadIdx.x; k < col.group_ptr[tile + 1]; k += blockDim.x) {
      const uint32_t hdr = col.group_hdr[k];
      const uint32_t slot = hdr & 0xFFFu;
      if (slot < slot_lo || slot >= slot_hi) continue;
      const uint32_t sym = col.slot_info[slot_base + slot] >> 16;
      if (sym >= col.n_symbols) continue;
      for (uint32_t q = 0; q <= (hdr >> 12); ++q) lab[col.groups[static_cast<size_t>(k) * 8 + q]] = sym;
   }
   for (uint32_t k = col.heavy_ptr[tile]; k < col.heavy_ptr[tile + 1]; ++k) {
      const Desc d = col.vert.desc[k];
      const uint32_t slot = descSlot(d.w0);
      if (slot < slot_lo || slot >= slot_hi) continue;
      const uint32_t sym = col.slot_info[slot_base + slot] >> 16;
      if (sym >= col.n_symbols) continue;
      const uint32_t type = descType(d.w0), card = descCard(d.w0);
      if (type == CT_INLINE) {
         if (threadIdx.x < card) lab[(d.w1 >> (16 * threadIdx.x)) & 0xFFFFu] = sym;
      } else if (type == CT_ARRAY) {
         for (uint32_t j = threadIdx.x; j < card; j += blockDim.x) lab[col.vert.arr[d.w1 + j]] = sym;
      } else if (type == CT_RUN) {
         for (uint32_t j = 0; j < card; ++j) {
            const uint32_t x = col.vert.runs[d.w1 + j];
            const uint32_t st = x & 0xFFFFu, e = st + (x >> 16);
            for (uint32_t v = st + threadIdx.x; v <= e; v += blockDim.x) lab[v] = sym;
         }
      } else {
         const uint64_t* bits = col.vert.bits + static_cast<size_t>(d.w1) * kChunkWords;
         for (uint32_t w = threadIdx.x; w < kChunkWords; w += blockDim.x) {
            uint64_t x = bits[w];
            while (x) {
               lab[w * 64 + __builtin_ctzll(x)] = sym;
               x &= x - 1;
            }
         }
      }
   }
}

// LDS bins per pass: 16-bit counters, two per word (a work-group counts one quarter chunk, at most
// 16,384 rows, so a counter cannot overflow)
constexpr uint32_t kHistLdsBins = 32768;
constexpr uint32_t kHistLdsWords = kHistLdsBins / 2;

__device__ inline void histAdd(uint32_t* lh, uint32_t k, uint32_t v) { atomicAdd(&lh[k >> 1], v << ((k & 1u) << 4)); }
__device__ inline uint32_t histGet(const uint32_t* lh, uint32_t k) { return (lh[k >> 1] >> ((k & 1u) << 4)) & 0xFFFFu; }
// Bin ranges counted in LDS by separate work-groups; bins beyond the last range go to global atomics.
constexpr uint32_t kHistRanges = 8;

// Adds one to LDS counter k for every active lane whose k is not ~0u, lanes holding the same k combined
// first: one LDS atomic per distinct k of the wavefront instead of one per lane (group keys of
// skewed columns repeat within 64 rows, and same-address LDS atomics serialise).
__device__ inline void addCombined(uint32_t* lh, uint32_t k) {
   uint64_t todo = __ballot(k != ~0u);
   while (todo) {
      const uint32_t lead = static_cast<uint32_t>(__builtin_ctzll(todo));
      const uint32_t lk = __shfl(k, static_cast<int>(lead), 64);
      const uint64_t same = __ballot(k == lk);
      if ((threadIdx.x & 63) == lead) histAdd(lh, lk, static_cast<uint32_t>(__popcll(same)));
      todo &= ~same;
   }
}

// Histograms of at most this many bins combine equal keys within a wavefront before counting.
constexpr uint32_t kHistCombineBins = 4096;

// Rows of one quarter chunk (16,384 rows) of a filter, each visited once: for a sparse quarter by
// iterating the set bits of its words, for a dense one by walking the rows in order (consecutive
// threads on consecutive rows, so the label loads coalesce).
// Bin keys of U rows: the labels of every dimension are loaded together (a load behind a
// per-row condition would wait for the loads before it); ~0u for a row with a null label.
constexpr uint32_t kHistBatch = 8;
template <uint32_t U>
__device__ inline void binKeys(const GroupDims& dims, const size_t (&r)[U], uint32_t (&key)[U]) {
   const uint32_t n = dims.n;
   uint32_t lab[4][U];
#pragma unroll
   for (uint32_t d = 0; d < 4; ++d) {
      if (d < n) {
         const GlobalU32* l = toGlobal<GlobalU32>(dims.labels[d]);
#pragma unroll
         for (uint32_t u = 0; u < U; ++u) lab[d][u] = l[r[u]];
      }
   }
#pragma unroll
   for (uint32_t u = 0; u < U; ++u) {
      uint32_t k = 0;
      bool ok = true;
#pragma unroll
      for (uint32_t d = 0; d < 4; ++d) {
         if (d < n) {
            ok &= lab[d][u] != kInvalidId;
            k += lab[d][u] * dims.stride[d];
         }
      }
      key[u] = ok ? k : ~0u;
   }
}

// Calls count(key) for the rows of a quarter chunk (16,384 rows from row_base) in the filter words f,
// kHistBatch rows per thread at a time, every lane of a wavefront in every call (key ~0u: no row).
// Dense quarters (chunk_card > 8192) read consecutive rows on consecutive lanes; sparse ones take
// the set bits of one filter word per thread.  blockDim.x is 256.
template <typename Count>
__device__ inline void histQuarter(const uint64_t* __restrict__ f, uint32_t chunk_card, size_t row_base,
                                   uint64_t* words_lds, const GroupDims& dims, Count count) {
   const uint32_t tid = threadIdx.x;
   if (chunk_card > 8192) {
      words_lds[tid] = f[tid];
      __syncthreads();
      for (uint32_t i0 = 0; i0 < 16384; i0 += 256 * kHistBatch) {
         size_t r[kHistBatch];
         uint32_t key[kHistBatch];
#pragma unroll
         for (uint32_t u = 0; u < kHistBatch; ++u) r[u] = row_base + i0 + u * 256 + tid;
         binKeys(dims, r, key);
#pragma unroll
         for (uint32_t u = 0; u < kHistBatch; ++u) {
            const uint32_t i = i0 + u * 256 + tid;
            count(((words_lds[i >> 6] >> (i & 63)) & 1ULL) ? key[u] : ~0u);
         }
      }
   } else {
      uint64_t x = f[tid];
      while (__ballot(x != 0)) {
         size_t r[kHistBatch];
         bool have[kHistBatch];
         uint32_t key[kHistBatch];
#pragma unroll
         for (uint32_t u = 0; u < kHistBatch; ++u) {
            have[u] = x != 0;
            const uint32_t bit = have[u] ? static_cast<uint32_t>(__builtin_ctzll(x)) : 0u;
            x &= x - 1;
            r[u] = row_base + tid * 64 + bit;
         }
         binKeys(dims, r, key);
#pragma unroll
         for (uint32_t u = 0; u < kHistBatch; ++u) count(have[u] ? key[u] : ~0u);
      }
   }
}

// Work-group (chunk quarter, bin range y): labels of up to four dimensions combined with strides
// into one bin; the rows whose bin falls in [y * kHistLdsBins, (y + 1) * kHistLdsBins) are counted
// in LDS, and the last range also takes every bin above it with global atomics.
__global__ __launch_bounds__(256) void groupHistogramKernel(uint32_t n_chunks, const uint64_t* __restrict__ filter,
                                                            const uint32_t* __restrict__ filter_chunk_cards,
                                                            GroupDims dims, uint32_t n_bins,
                                                            uint32_t* __restrict__ hist) {
   __shared__ uint32_t lh[kHistLdsWords];
   __shared__ uint64_t words[256];
   const uint32_t c = blockIdx.x >> 2;
   const uint32_t quarter = blockIdx.x & 3;
   const uint32_t lo = blockIdx.y * kHistLdsBins;
   const uint32_t card = filter_chunk_cards[c];
   if (card == 0 || lo >= n_bins) return;
   const uint32_t span = n_bins - lo < kHistLdsBins ? n_bins - lo : kHistLdsBins;
   const bool last = blockIdx.y + 1 == gridDim.y;
   for (uint32_t i = threadIdx.x; i < (span + 1) / 2; i += blockDim.x) lh[i] = 0;
   __syncthreads();
   const uint64_t* f = filter + static_cast<size_t>(c) * kChunkWords + quarter * 256;
   const bool combine = n_bins <= kHistCombineBins;
   histQuarter(f, card, (static_cast<size_t>(c) << 16) + quarter * 16384u, words, dims, [&](uint32_t key) {
      const bool in_table = key != ~0u;
      const bool here = in_table && key >= lo && key - lo < kHistLdsBins;
      if (combine) {
         addCombined(lh, here ? key - lo : ~0u);
      } else if (here) {
         histAdd(lh, key - lo, 1u);
      }
      if (in_table && last && key >= lo + kHistLdsBins) atomicAdd(&hist[key], 1u);
   });
   __syncthreads();
   for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) {
      const uint32_t v = histGet(lh, i);
      if (v) atomicAdd(&hist[lo + i], v);
   }
}

// Work-group (quarter chunk, entry e): the rows are counted in passes over bin ranges
// [z * kHistLdsBins, (z + 1) * kHistLdsBins), each in LDS; the last pass also takes every bin above
// its range, with global atomics.  The passes run inside the work-group so that the grid (fixed for
// graph replay) does not grow with the bin count.
__global__ __launch_bounds__(256) void groupHistogramBatchKernel(AggregateBatch a) {
   __shared__ uint32_t lh[kHistLdsWords];
   __shared__ uint64_t words[256];
   const uint32_t e = blockIdx.y;
   if (e >= *a.n_agg) return;
   const uint32_t n_bins = a.bins[e];
   const uint32_t q = a.agg_query[e];
   const uint32_t c = blockIdx.x >> 2;
   const uint32_t quarter = blockIdx.x & 3;
   const uint32_t card = a.chunk_cards[static_cast<size_t>(q) * a.n_chunks + c];
   if (card == 0 || n_bins == 0) return;
   const GroupDims dims = a.dims[e];
   uint32_t* hist = a.hist + a.hist_offset[e];
   const uint64_t* f = a.filters + static_cast<size_t>(q) * a.n_chunks * kChunkWords + static_cast<size_t>(c) * kChunkWords +
                       quarter * 256;
   uint32_t passes = (n_bins + kHistLdsBins - 1) / kHistLdsBins;
   passes = passes > kHistRanges ? kHistRanges : passes;
   const bool combine = n_bins <= kHistCombineBins;
   for (uint32_t z = 0; z < passes; ++z) {
      const uint32_t lo = z * kHistLdsBins;
      const uint32_t span = n_bins - lo < kHistLdsBins ? n_bins - lo : kHistLdsBins;
      const bool last = z + 1 == passes;
      for (uint32_t i = threadIdx.x; i < (span + 1) / 2; i += blockDim.x) lh[i] = 0;
      __syncthreads();
      histQuarter(f, card, (static_cast<size_t>(c) << 16) + quarter * 16384u, words, dims, [&](uint32_t key) {
         const bool in_table = key != ~0u;
         const bool here = in_table && key >= lo && key - lo < kHistLdsBins;
         if (combine) {
            addCombined(lh, here ? key - lo : ~0u);
         } else if (here) {
            histAdd(lh, key - lo, 1u);
         }
         if (in_table && last && key >= lo + kHistLdsBins) atomicAdd(&hist[key], 1u);
      });
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) {
         const uint32_t v = histGet(lh, i);
         if (v) atomicAdd(&hist[lo + i], v);
      }
      __syncthreads();
   }
}

// Non-empty bins of every entry as {bin, count} records in its query's row slots (any order).
__global__ __launch_bounds__(256) void compactHistogramKernel(AggregateBatch a) {
   const uint32_t e = blockIdx.y;
   if (e >= *a.n_agg) return;
   const uint32_t q = a.agg_query[e];
   const uint32_t n_bins = a.bins[e];
   const uint32_t* hist = a.hist + a.hist_offset[e];
   MutationRowDev* out = a.rows + static_cast<size_t>(q) * a.row_capacity;
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_bins; i += gridDim.x * blockDim.x) {
      const uint32_t v = hist[i];
      if (v == 0) continue;
      const uint32_t k = atomicAdd(&a.n_rows[q], 1u);
      if (k < a.row_capacity) out[k] = MutationRowDev{i, v, 0, 0};
   }
}

__global__ void fillCountKernel(uint32_t* __restrict__ p, uint32_t value, const uint32_t* __restrict__ n_dev) {
   const uint32_t n = *n_dev;
   for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = value;
}

__global__ __launch_bounds__(256) void packRowsKernel(const uint32_t* __restrict__ n_queries, const uint32_t* __restrict__ n_rows,
                                                      const unsigned long long* __restrict__ cards,
                                                      const MutationRowDev* __restrict__ rows, uint32_t row_capacity,
                                                      uint32_t* __restrict__ header, MutationRowDev* __restrict__ packed) {
   __shared__ uint32_t part[4];
   const uint32_t q = blockIdx.x;
   if (q >= *n_queries) return;
   // offset of query q: rows kept by the queries before it
   uint32_t acc = 0;
   for (uint32_t j = threadIdx.x; j < q; j += blockDim.x) acc += min(n_rows[j], row_capacity);
   acc = waveTotal(acc);
   if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
   __syncthreads();
   const uint32_t offset = part[0] + part[1] + part[2] + part[3];
   const uint32_t n = min(n_rows[q], row_capacity);
   if (threadIdx.x == 0) {
      const unsigned long long card = cards[q];
      header[4 * q] = n_rows[q];
      header[4 * q + 1] = offset;
      header[4 * q + 2] = static_cast<uint32_t>(card);
      header[4 * q + 3] = static_cast<uint32_t>(card >> 32);
   }
   const MutationRowDev* src = rows + static_cast<size_t>(q) * row_capacity;
   for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) packed[offset + i] = src[i];
}

__global__ void fillKernel(uint32_t* __restrict__ p, uint32_t value, size_t n) {
   for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
        i += static_cast<size_t>(gridDim.x) * blockDim.x) {
      p[i] = value;
   }
}

}  // namespace

// ---- launch wrappers ----------------------------------------------------------------------------
// A streaming read of n16 16-byte words, four loads in flight per lane: the bandwidth a pure read
// of a given size reaches on this device (the reference point for the counting kernels).  A word
// is written only if the XOR of everything read has a value the caller's data never gives.
__global__ __launch_bounds__(256) void streamReadKernel(const uint4* __restrict__ src, size_t n16,
                                                       uint32_t* __restrict__ out) {
   uint32_t acc = 0;
   const size_t stride = static_cast<size_t>(gridDim.x) * 256 * 4;
   for (size_t i = blockIdx.x * 1024ull + threadIdx.x; i < n16; i += stride) {
      uint4 v[4];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) v[u] = i + 256 * u < n16 ? src[i + 256 * u] : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
   }
   if (acc == 0x9E3779B9u) out[blockIdx.x * 256 + threadIdx.x] = acc;
}

hipError_t launchStreamRead(const void* src, size_t bytes, uint32_t* out, uint32_t blocks, hipStream_t stream) {
   if (bytes < 16 || blocks == 0) return hipErrorInvalidValue;
   hipLaunchKernelGGL(streamReadKernel, dim3(blocks), dim3(256), 0, stream, static_cast<const uint4*>(src), bytes / 16,
                      out);
   return hipGetLastError();
}

hipError_t launchFill(uint32_t* p, uint32_t value, size_t n, hipStream_t stream) {
   if (n == 0) return hipSuccess;
   size_t blocks = (n + 255) / 256;
   if (blocks > 4096) blocks = 4096;
   hipLaunchKernelGGL(fillKernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream, p, value, n);
   return hipGetLastError();
}

hipError_t configureFilterKernel(size_t max_lds_bytes) {
   if (max_lds_bytes <= 64 * 1024) return hipSuccess;
   return hipFuncSetAttribute(reinterpret_cast<const void*>(&filterVmKernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(max_lds_bytes));
}

size_t filterLdsBytes(uint32_t depth) {
   // per wave: depth stack entries and a scratch entry of 64 words, the prefetched leaf records, the
   // cached instructions and a unit's result words
   return static_cast<size_t>(kVmWaves) * ((depth + 1) * 64 * sizeof(uint64_t) + kPrefetchSlots * 8 * sizeof(uint32_t) +
                                           kProgCache * sizeof(Instr) + kVmUnitSlices * 64 * sizeof(uint64_t));
}

hipError_t launchFilter(const TableView& tv, const ProgramDesc* progs, const Instr* instrs, const uint32_t* ranges,
                        uint32_t n_queries, uint32_t depth, uint64_t* out, unsigned long long* cards,
                        uint32_t* chunk_cards, uint32_t max_blocks, hipStream_t stream,
                        const uint32_t* n_queries_dev) {
   const uint32_t n_items = n_queries * tv.n_chunks * kSlicesPerChunk;
   if (n_items == 0) return hipSuccess;
   // more slices a unit only where the units still outnumber the wavefronts well (a single query
   // needs every slice in parallel); unit_slices divides kSlicesPerChunk
   const uint32_t unit = n_items >= kVmUnitSlices * kVmUnitMinUnits ? kVmUnitSlices
                         : n_items >= 2 * kVmUnitMinUnits           ? 2u
                                                                    : 1u;
   static_assert(kSlicesPerChunk % kVmUnitSlices == 0, "units stay inside a chunk");
   uint32_t blocks = (n_items / unit + kVmWaves - 1) / kVmWaves;
   if (max_blocks && blocks > max_blocks) blocks = max_blocks;
   hipLaunchKernelGGL(filterVmKernel, dim3(blocks), dim3(kVmThreads), filterLdsBytes(depth), stream, tv, progs,
                      instrs, ranges, n_items, depth, out, cards, chunk_cards, n_queries_dev, unit);
   return hipGetLastError();
}

hipError_t launchMutations(const MutationsLaunch& m, hipStream_t stream) {
   if (m.n_groups == 0 || m.n_groups > kMaxGroups) return hipErrorInvalidValue;
   if (m.sparse_max >= 0 && (m.sparse_counts == nullptr || m.sparse_bnd == nullptr)) return hipErrorInvalidValue;
   hipError_t err = hipSuccess;
   const SeqColumnView& h = *m.col_host;
   if (m.count_start && (err = hipEventRecord(m.count_start, stream)) != hipSuccess) return err;
   // packed counters need at most kPackedMaxChunks chunks per work-group and the heavy sums
   const bool packed = m.heavy_sums != nullptr && (m.n_chunks + m.n_groups - 1) / m.n_groups <= kPackedMaxChunks;
   uint32_t* heavy = packed ? m.heavy_sums : nullptr;
   const dim3 grid(h.n_blocks, m.n_groups);
#define RDB_LAUNCH_K1(U, P, C)                                                                                         \
   hipLaunchKernelGGL((mutationsCountKernel<U, P, C>), grid, dim3(kMutThreads), 0, stream, m.col_dev, m.n_chunks,         \
                      m.n_groups, m.filter, m.filter_chunk_cards, m.filter_card, m.sparse_max, m.partial, heavy)
#define RDB_LAUNCH_K1_C(U, P)       \
   if (m.coop) {                    \
      RDB_LAUNCH_K1(U, P, true);    \
   } else {                         \
      RDB_LAUNCH_K1(U, P, false);   \
   }
   if (m.unroll == 16) {
      if (packed) {
         RDB_LAUNCH_K1_C(16, true);
      } else {
         RDB_LAUNCH_K1_C(16, false);
      }
   } else if (m.unroll == 4) {
      if (packed) {
         RDB_LAUNCH_K1_C(4, true);
      } else {
         RDB_LAUNCH_K1_C(4, false);
      }
   } else if (m.unroll == 2) {
      if (packed) {
         RDB_LAUNCH_K1_C(2, true);
      } else {
         RDB_LAUNCH_K1_C(2, false);
      }
   } else {
      if (packed) {
         RDB_LAUNCH_K1_C(8, true);
      } else {
         RDB_LAUNCH_K1_C(8, false);
      }
   }
#undef RDB_LAUNCH_K1_C
#undef RDB_LAUNCH_K1
   if ((err = hipGetLastError()) != hipSuccess) return err;
   if (m.count_stop && (err = hipEventRecord(m.count_stop, stream)) != hipSuccess) return err;
   if ((m.n_groups > 1 || heavy) && h.n_slots > 0) {
      hipLaunchKernelGGL(reducePartialsKernel, dim3((h.n_slots + 255) / 256), dim3(256), 0, stream, m.partial,
                         m.n_groups, h.n_slots, m.filter_card, m.sparse_max, heavy);
      if ((err = hipGetLastError()) != hipSuccess) return err;
   }
   const uint32_t n_pblocks = (h.genome_length + kPosBlock - 1) / kPosBlock;
   const SlotSource dense{m.col_dev, m.partial};
   hipLaunchKernelGGL(mutationsTotalsKernel<SlotSource>, dim3(n_pblocks), dim3(kPosBlock), 0, stream, dense,
                      m.filter_card, m.sparse_max, m.totals, m.n_rows);
   if ((err = hipGetLastError()) != hipSuccess) return err;
   hipLaunchKernelGGL(mutationsFinalizeKernel<SlotSource>, dim3(n_pblocks), dim3(kPosBlock), 0, stream, dense,
                      m.totals, n_pblocks, m.filter_card, m.sparse_max, m.min_proportion, m.counts_out, m.rows,
                      m.row_capacity, m.n_rows);
   if ((err = hipGetLastError()) != hipSuccess) return err;
   if (m.sparse_max < 0) return hipSuccess;
   const uint32_t n_words = m.n_chunks * kChunkWords;
   if (n_words == 0) return hipSuccess;  // no chunks: no rows
   hipLaunchKernelGGL(sparseCountKernel, dim3((n_words + 255) / 256), dim3(256), 0, stream, m.col_dev, m.chunk_sizes,
                      n_words, m.filter, m.filter_card, m.sparse_max, m.sparse_counts, m.sparse_bnd);
   if ((err = hipGetLastError()) != hipSuccess) return err;
   const SparseSource sparse{m.col_dev, m.sparse_counts, m.sparse_bnd};
   hipLaunchKernelGGL(mutationsTotalsKernel<SparseSource>, dim3(n_pblocks), dim3(kPosBlock), 0, stream, sparse,
                      m.filter_card, m.sparse_max, m.totals, m.n_rows);
   if ((err = hipGetLastError()) != hipSuccess) return err;
   hipLaunchKernelGGL(mutationsFinalizeKernel<SparseSource>, dim3(n_pblocks), dim3(kPosBlock), 0, stream, sparse,
                      m.totals, n_pblocks, m.filter_card, m.sparse_max, m.min_proportion, m.counts_out, m.rows,
                      m.row_capacity, m.n_rows);
   return hipGetLastError();
}

hipError_t launchMutationsBatch(const MutationsBatch& m, hipStream_t stream) {
   if (m.max_queries == 0) return hipSuccess;
   if (m.max_queries > 65535 || m.grid_queries > m.max_queries) return hipErrorInvalidValue;
   const uint32_t gq = m.grid_queries ? m.grid_queries : m.max_queries;  // grid rows of per-query kernels
   const SeqColumnView& h = *m.col_host;
   hipError_t err = hipSuccess;
   hipLaunchKernelGGL(batchPlanKernel, dim3(1), dim3(1024), 0, stream, m);
   if ((err = hipGetLastError()) != hipSuccess) return err;
   // the dense scan asked for, or the next smaller one where a table is too large for it (the
   // knob never changes results); the lane holds the buffers of the scan this chooses
   const uint32_t max_octets = (gq + 7) / 8;
   uint32_t mode = m.heavy_sums == nullptr ? 0u : effectiveSharedScan(m.shared_mode, m.max_queries, m.n_chunks);
   if (mode == 2 && (m.octet_masks == nullptr || m.partial_c == nullptr || m.partial_b == nullptr)) {
      return hipErrorInvalidValue;
   }
   if (mode == 1 && m.partial_b == nullptr) return hipErrorInvalidValue;
   if (m.scan_start && (err = hipEventRecord(m.scan_start, stream)) != hipSuccess) return err;
   if (mode == 2) {
      // shared scan, eight dense queries per pass: row bytes, then the persistent counting kernel
      if (m.n_chunks > 0) {
         hipLaunchKernelGGL(octetMasksKernel, dim3(kChunkWords / kOctetMaskThreads, max_octets * m.n_chunks),
                            dim3(kOctetMaskThreads), 0, stream, m);
         if ((err = hipGetLastError()) != hipSuccess) return err;
      }
      const uint32_t groups = m.persistent_groups == 0 ? 1 : m.persistent_groups;
      if (m.octet_threads == 512) {
         if (m.unroll == 16) {
            hipLaunchKernelGGL((mutationsCountOctetKernel<4, 512>), dim3(groups), dim3(512), 0, stream, m);
         } else {
            hipLaunchKernelGGL((mutationsCountOctetKernel<2, 512>), dim3(groups), dim3(512), 0, stream, m);
         }
      } else {
         if (m.unroll == 16) {
            hipLaunchKernelGGL((mutationsCountOctetKernel<4, 1024>), dim3(groups), dim3(1024), 0, stream, m);
         } else {
            hipLaunchKernelGGL((mutationsCountOctetKernel<2, 1024>), dim3(groups), dim3(1024), 0, stream, m);
         }
      }
      if ((err = hipGetLastError()) != hipSuccess) return err;
      if (m.scan_stop && (err = hipEventRecord(m.scan_stop, stream)) != hipSuccess) return err;
   } else if (mode == 1) {
      // shared scan, four dense queries per pass, chunks in two parts
      const uint32_t n_xblocks = 8 * ((h.n_blocks + 7) / 8);
      const dim3 grid(n_xblocks * 2 * ((gq + 3) / 4));
      if (m.unroll == 16) {
         hipLaunchKernelGGL(mutationsCountSharedKernel<4>, grid, dim3(kSharedThreads), 0, stream, m);
      } else if (m.unroll == 2) {
         hipLaunchKernelGGL(mutationsCountSharedKernel<1>, grid, dim3(kSharedThreads), 0, stream, m);
      } else {
         hipLaunchKernelGGL(mutationsCountSharedKernel<2>, grid, dim3(kSharedThreads), 0, stream, m);
      }
      if ((err = hipGetLastError()) != hipSuccess) return err;
      if (m.scan_stop && (err = hipEventRecord(m.scan_stop, stream)) != hipSuccess) return err;
   } else {
      const uint32_t n_xblocks = 8 * ((h.n_blocks + 7) / 8);
      const dim3 count_grid(n_xblocks * gq);
      if (m.unroll == 16) {
         if (m.coop) {
            hipLaunchKernelGGL((mutationsCountBatchKernel<16, true>), count_grid, dim3(kMutThreads), 0, stream, m);
         } else {
            hipLaunchKernelGGL((mutationsCountBatchKernel<16, false>), count_grid, dim3(kMutThreads), 0, stream, m);
         }
      } else {
         if (m.coop) {
            hipLaunchKernelGGL((mutationsCountBatchKernel<8, true>), count_grid, dim3(kMutThreads), 0, stream, m);
         } else {
            hipLaunchKernelGGL((mutationsCountBatchKernel<8, false>), count_grid, dim3(kMutThreads), 0, stream, m);
         }
      }
      if ((err = hipGetLastError()) != hipSuccess) return err;
      if (m.scan_stop && (err = hipEventRecord(m.scan_stop, stream)) != hipSuccess) return err;
   }
   if ((err = hipGetLastError()) != hipSuccess) return err;
   // the sparse entries' counts, then totals and finalize of both lists (the dense slot sums add the
   // scan's parts and heavy sums as they are read)
   const uint32_t n_words = m.n_chunks * kChunkWords;
   if (m.sparse_max >= 0 && n_words > 0) {
      hipLaunchKernelGGL(sparseCountBatchKernel, dim3((n_words + kSparseBatchThreads - 1) / kSparseBatchThreads, gq),
                         dim3(kSparseBatchThreads), 0, stream, m);
      if ((err = hipGetLastError()) != hipSuccess) return err;
   }
   const uint32_t dense_parts = mode == 2 ? kOctetParts : (mode == 1 ? 2u : 1u);
   const uint32_t n_pblocks = (h.genome_length + kPosBlock - 1) / kPosBlock;
   hipLaunchKernelGGL(mutationsTotalsBatchKernel, dim3(n_pblocks, gq), dim3(kPosBlock), 0, stream, m,
                      n_pblocks, dense_parts);
   if ((err = hipGetLastError()) != hipSuccess) return err;
   hipLaunchKernelGGL(mutationsFinalizeBatchKernel, dim3(n_pblocks, gq), dim3(kPosBlock), 0, stream, m,
                      n_pblocks, dense_parts);
   return hipGetLastError();
}

hipError_t launchAggregateBatch(const AggregateBatch& a, hipStream_t stream) {
   if (a.max_entries == 0) return hipSuccess;
   if (a.max_entries > 65535) return hipErrorInvalidValue;
   uint32_t blocks = (a.max_total_bins + 255) / 256;
   if (blocks > 1024) blocks = 1024;
   if (blocks == 0) blocks = 1;
   hipLaunchKernelGGL(fillCountKernel, dim3(blocks), dim3(256), 0, stream, a.hist, 0u, a.total_bins);
   hipError_t err = hipGetLastError();
   if (err != hipSuccess) return err;
   if (a.n_chunks == 0) return hipSuccess;  // no rows: the histograms stay zero
   hipLaunchKernelGGL(groupHistogramBatchKernel, dim3(a.n_chunks * 4, a.max_entries), dim3(256), 0, stream, a);
   if ((err = hipGetLastError()) != hipSuccess) return err;
   hipLaunchKernelGGL(compactHistogramKernel, dim3(64, a.max_entries), dim3(256), 0, stream, a);
   return hipGetLastError();
}

hipError_t launchPackRows(const uint32_t* n_queries, uint32_t max_queries, const uint32_t* n_rows,
                          const unsigned long long* cards, const MutationRowDev* rows, uint32_t row_capacity,
                          uint32_t* header, MutationRowDev* packed, hipStream_t stream) {
   if (max_queries == 0) return hipSuccess;
   hipLaunchKernelGGL(packRowsKernel, dim3(max_queries), dim3(256), 0, stream, n_queries, n_rows, cards, rows,
                      row_capacity, header, packed);
   return hipGetLastError();
}

hipError_t launchSymbolLabels(const SeqColumnView* col_dev, const uint32_t* chunk_sizes, uint32_t position,
                              uint32_t n_chunks, uint32_t block, uint32_t slot_lo, uint32_t slot_hi, uint32_t* labels,
                              hipStream_t stream) {
   const uint32_t n = n_chunks << 16;
   if (n_chunks == 0) return hipSuccess;
   hipLaunchKernelGGL(symbolLabelsKernel, dim3((n + 255) / 256), dim3(256), 0, stream, col_dev, chunk_sizes, position, n,
                      labels);
   hipError_t err = hipGetLastError();
   if (err != hipSuccess) return err;
   hipLaunchKernelGGL(symbolScatterKernel, dim3(n_chunks), dim3(256), 0, stream, col_dev, n_chunks, block, slot_lo,
                      slot_hi, labels);
   return hipGetLastError();
}

hipError_t launchGroupHistogram(uint32_t n_chunks, const uint64_t* filter, const uint32_t* filter_chunk_cards,
                                const GroupDims& dims, uint32_t n_bins, uint32_t* hist, hipStream_t stream) {
   hipError_t err = launchFill(hist, 0, n_bins, stream);
   if (err != hipSuccess) return err;
   if (n_chunks == 0) return hipSuccess;
   uint32_t ranges = (n_bins + kHistLdsBins - 1) / kHistLdsBins;
   ranges = ranges == 0 ? 1 : (ranges > kHistRanges ? kHistRanges : ranges);
   hipLaunchKernelGGL(groupHistogramKernel, dim3(n_chunks * 4, ranges), dim3(256), 0, stream, n_chunks, filter,
                      filter_chunk_cards, dims, n_bins, hist);
   return hipGetLastError();
}

}  // namespace rdb
