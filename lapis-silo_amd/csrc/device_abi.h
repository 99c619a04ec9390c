// Plain structures shared by the host code and the HIP kernels.
//
// Device layout of one table (all row-indexed arrays are padded to n_chunks * 2^16 rows so that the
// global row id `chunk << 16 | row` indexes them directly; padding rows are zero):
//
//  * Containers — one per (position, chunk, symbol) of a sequence column and one per (bitmap,
//    chunk) of the bitmap store.  An 8-byte descriptor {w0, w1}:
//      w0 = slot (12 bits) | type (2 bits) << 12 | card (13 bits) << 14
//      w1 = payload: INLINE: up to two u16 row ids; ARRAY: offset into the u16 arena;
//           RUN: offset into the u32 run arena (start | (length - 1) << 16 per run);
//           BITSET: index of a 1024 x u64 block in the bitset arena.
//    `card` is the element count of INLINE/ARRAY and the run count of RUN.  An INLINE descriptor
//    with card 0 is the empty container.
//  * Vertical index of a sequence column — positions are cut into blocks (contiguous position
//    ranges with at most 4095 distinct (position, symbol) "slots" and roughly equal payload
//    bytes).  Tile (block b, chunk c) holds the block's containers in chunk c in two parts:
//    light containers (at most 32 rows) flattened into u32 entries `slot << 16 | row`, sorted by
//    slot then row and padded to a multiple of four with entries of the unused slot 4095; larger
//    arrays as groups of eight rows (16 bytes, unused places zero) with a u16 header per group
//    `slot | (rows_in_group - 1) << 12`; and bitset and run containers as descriptors (processed
//    one per wavefront).  Besides the stored (non-local-reference)
//    symbols, four boundary pseudo-symbols hold, per position p, the rows whose covered range
//    starts at p (p > 0), ends at p (p < L), and whose internal missing runs start / end at p —
//    the vertical form of the reference's horizontal coverage index.
//  * Row-level arrays for filter leaves: coverage (start | end << 16), internal missing runs (CSR),
//    int/date values, dictionary ids and null bitsets.
//  * Row-major mutations of a sequence column (CSR over padded rows, `position | symbol << 16`,
//    relative to the original reference): the input of the sparse Mutations kernel.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define RDB_HD __host__ __device__
#else
#define RDB_HD
#endif

namespace rdb {

constexpr uint32_t kChunkRows = 1u << 16;
constexpr uint32_t kChunkWords = 1024;
constexpr uint32_t kSliceWords = 64;
constexpr uint32_t kSlicesPerChunk = kChunkWords / kSliceWords;
// Per-(chunk, position) record of a sequence column: [lo, hi) of the position's light entries,
// groups and descriptors in its tile, then a mask of the 4096-row slices holding any of its rows.
constexpr uint32_t kPosRecord = 8;
constexpr uint32_t kMaxBlockSlots = 4095;   // slot 4095 is the padding slot of light entries
constexpr uint32_t kPadSlot = 4095;
constexpr uint32_t kLightMaxRows = 32;
constexpr uint32_t kMaxGroups = 32;         // chunk groups of the Mutations kernel
constexpr uint32_t kInvalidId = 0xFFFFFFFFu;
constexpr uint32_t kMutPosBlock = 256;      // positions per work-group of the Mutations finalize kernels
constexpr uint32_t kRunBucket = 256;        // positions per start bucket of the missing-run index
constexpr uint32_t kShortRun = 1024;        // longer missing runs go to a per-chunk list of their own

enum ContainerType : uint32_t { CT_INLINE = 0, CT_ARRAY = 1, CT_RUN = 2, CT_BITSET = 3 };

RDB_HD inline uint32_t descSlot(uint32_t w0) { return w0 & 0xFFFu; }
RDB_HD inline uint32_t descType(uint32_t w0) { return (w0 >> 12) & 3u; }
RDB_HD inline uint32_t descCard(uint32_t w0) { return (w0 >> 14) & 0x1FFFu; }
RDB_HD inline uint32_t makeW0(uint32_t slot, uint32_t type, uint32_t card) {
   return (slot & 0xFFFu) | (type << 12) | (card << 14);
}

struct Desc {
   uint32_t w0;
   uint32_t w1;
};

struct ContainerArenas {
   const Desc* desc;
   const uint16_t* arr;
   const uint32_t* runs;
   const uint64_t* bits;
};

// Boundary pseudo-symbols follow the alphabet's real symbols.
enum BoundaryKind : uint32_t { BND_COV_START = 0, BND_COV_END = 1, BND_N_START = 2, BND_N_END = 3 };

struct SeqColumnView {
   uint32_t genome_length;
   uint32_t n_symbols;   // alphabet size (16 or 28); pseudo-symbols are n_symbols + BoundaryKind
   uint32_t n_blocks;
   uint32_t n_slots;
   uint32_t n_valid;     // number of VALID_MUTATION_SYMBOLS
   uint32_t valid_mask;  // bit s set for every valid mutation symbol s
   ContainerArenas vert;
   const uint32_t* block_pos_start;   // [n_blocks + 1]
   const uint32_t* block_slot_start;  // [n_blocks + 1]
   const uint32_t* light_ptr;         // [n_blocks * n_chunks + 1], multiples of 4
   const uint32_t* light;             // slot << 16 | row
   const uint32_t* group_ptr;         // [n_blocks * n_chunks + 1] into groups / group_hdr
   const uint16_t* groups;            // eight u16 rows per group (16-byte aligned)
   const uint16_t* group_hdr;         // slot | (rows - 1) << 12
   const uint32_t* heavy_ptr;         // [n_blocks * n_chunks + 1] into vert.desc
   const uint32_t* slot_info;         // [n_slots] position | symbol << 16
   const uint32_t* pos_slot_ptr;      // [genome_length + 1]
   const uint8_t* local_ref;          // [genome_length]
   const uint8_t* reference;          // [genome_length]
   const uint32_t* cov;               // [n_chunks * 2^16] start | end << 16
   // per (chunk, slice): the largest start and smallest end over the slice's rows that have a
   // covered range (start | end << 16; 0xFFFF | 0 when none), and the rows without one (null or
   // empty range): a position inside the bounds is covered by every other row of the slice
   const uint32_t* cov_bounds;        // [n_chunks * 16]
   const uint64_t* cov_never;         // [n_chunks * 1024]
   const uint32_t* nrun_ptr;          // [n_chunks * 2^16 + 1]
   const uint32_t* nruns;             // start | end << 16, end exclusive
   // missing-run index by position: list (chunk c, bucket k, slice s) holds the runs of slice s's
   // rows (row | start << 16 | end << 32, row within the chunk, in row order) that start in
   // positions [k * kRunBucket, (k + 1) * kRunBucket) and are at most kShortRun long; bucket
   // n_run_buckets holds the longer runs.  A run covering p starts in one of the buckets of
   // [p - kShortRun + 1, p], or is long.
   uint32_t n_run_buckets;
   const uint32_t* run_index_ptr;     // [(c * (n_run_buckets + 1) + k) * 16 + s], and one more
   const uint64_t* run_index;
   const uint64_t* null_bits;         // [n_chunks * 1024]
   const uint32_t* pos_ranges;        // [(chunk * L + position) * 6]: [lo, hi) of the position's light
                                      // entries, groups and descriptors in its tile
   const uint32_t* row_mut_ptr;       // [n_chunks * 2^16 + 1] row-major mutations (sparse path)
   const uint32_t* row_muts;          // position | symbol << 16, original-reference frame
};

// One row of Mutations output before the host computes the proportion.
struct MutationRowDev {
   uint32_t position;  // 0-based
   uint32_t symbol;
   uint32_t count;
   uint32_t coverage;
};

struct IntColumnView {
   const int32_t* values;     // [n_chunks * 2^16]
   const uint64_t* null_bits; // [n_chunks * 1024]
   // Bit-sliced copy when the non-null values span fewer than 2^16: plane p, word w of chunk c at
   // planes[(p * n_chunks + c) * 1024 + w] holds bit p of (value - plane_min) of the word's rows
   // (0 for null rows).  n_planes == 0: no copy (comparisons read `values`).
   const uint64_t* planes;
   int32_t plane_min;
   uint32_t n_planes;
};
constexpr uint32_t kMaxIntPlanes = 16;

struct TableView {
   uint32_t n_chunks;
   const uint32_t* chunk_sizes;      // [n_chunks]
   ContainerArenas store;            // bitmap store, descriptor of (bitmap, chunk) at bitmap * n_chunks + chunk
   const SeqColumnView* seq;         // device array
   const IntColumnView* ints;        // device array
};

// ---- filter program (post-order stack code over 64-word row slices) -------------------------------
enum OpCode : uint32_t {
   OP_EMPTY = 0,
   OP_FULL,
   OP_BITMAP,    // a = bitmap-store id, b = prefetch slot
   OP_SYMBOLS,   // a = seq column, b = block, c = slot lo, d = slot hi (block-relative), e = mask of
                 // the slots c + j taken, f = position, g = prefetch slot
   OP_COVERED,   // a = seq column, b = position
   OP_CMP,       // a = int column, b = comparator, c = value (bits), d = with_nulls
   OP_RANGES,    // a = first range, b = count (global [lo, hi) pairs in the program's range table)
   OP_AND,
   OP_OR,
   OP_ANDNOT,    // second-from-top & ~top
   OP_NOT,       // complement within the row layout
   OP_TH_BEGIN,  // a = planes: push started-mask and `a` counter planes
   OP_TH_ADD,    // a = planes, b = negated child
   OP_TH_END,    // a = planes, b = n, c = exact
   OP_SKIP_IF_EMPTY,  // a = target: jump when the top word is zero in every lane (rest of an And)
   OP_SKIP_IF_FULL,   // a = target: jump when the top word is all ones in every lane (rest of an Or)
   OP_TH_SKIP,        // a = target (the OP_TH_END), b = planes, c = n, d = exact, e = children left:
                      // jump when every row's count is decided whatever the children left add
};

enum CmpCode : uint32_t { CMP_EQ = 0, CMP_LT, CMP_GT, CMP_LE, CMP_GE, CMP_NE };

struct Instr {
   uint32_t op, a, b, c, d, e, f, g;
};

// Leaves whose per-chunk records (a sequence position's six tile offsets, a store bitmap's
// descriptor) a wavefront loads together, one per lane, before running the program: codes
// column << 16 | position, or 1 << 31 | store id.  Leaf instructions name their prefetch slot
// (OP_SYMBOLS in g, OP_BITMAP in b); kNoPrefetch loads the record at the leaf instead.
constexpr uint32_t kPrefetchSlots = 64;
constexpr uint32_t kNoPrefetch = 0xFFFFu;

struct ProgramDesc {
   uint32_t instr_begin;
   uint32_t n_instr;
   uint32_t range_begin;
   uint32_t query;       // index into the per-query outputs
   uint32_t pf_begin;    // prefetch codes at ranges[pf_begin ..]
   uint32_t n_pf;
};

}  // namespace rdb
